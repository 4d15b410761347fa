// Host build of the engine's field arithmetic (omr-ntt_amd/csrc/field.h)
// checked against plain 128-bit remainders: add/sub, Montgomery and
// Goldilocks products, multiplier-form round trips, the butterfly
// difference times a multiplier, Goldilocks x * 2^k,
// (a - b) * 2^k with the sign folded, x * 2^-m.
#include <cstdio>
#include <random>

#include "../../omr-ntt_amd/csrc/field.h"

using namespace omr;
typedef unsigned __int128 u128;
static int fails = 0;
#define CHECK(c, ...)                         \
  do {                                        \
    if (!(c)) {                               \
      if (fails++ < 20) std::printf(__VA_ARGS__); \
    }                                         \
  } while (0)

template <class F>
void check_field(std::mt19937_64 &rng) {
  const uint64_t p = F::P;
  uint64_t edges[] = {0, 1, 2, p - 1, p - 2, p / 2, p / 2 + 1, 0xFFFFFFFFull % p, (1ull << 31) % p};
  for (int it = 0; it < 200000; it++) {
    uint64_t a = it < 729 ? edges[it % 81 / 9] : rng() % p, b = it < 729 ? edges[it % 9] : rng() % p;
    using T = typename F::T;
    T x = (T)a, y = (T)b;
    CHECK((uint64_t)F::add(x, y) == (uint64_t)(((u128)a + b) % p), "add %d\n", F::ID);
    CHECK((uint64_t)F::sub(x, y) == (uint64_t)(((u128)a + p - b) % p), "sub %d\n", F::ID);
    CHECK((uint64_t)F::mul(x, y) == (uint64_t)(((u128)a * b) % p), "mul %d %llu %llu\n", F::ID,
          (unsigned long long)a, (unsigned long long)b);
    CHECK((uint64_t)F::from_w(F::to_w(x)) == a, "w round trip %d\n", F::ID);
    CHECK((uint64_t)F::mulw(x, F::to_w(y)) == (uint64_t)(((u128)a * b) % p), "mulw %d\n", F::ID);
    CHECK((uint64_t)F::from_w(F::wmul(F::to_w(x), F::to_w(y))) == (uint64_t)(((u128)a * b) % p), "wmul %d\n", F::ID);
    // (a - b) w for a multiplier w of c: the unreduced difference a + p - b
    // reaches 2p - 1 (a = p - 1, b = 0), and c runs over the edges too
    const uint64_t c = it < 729 ? edges[(it / 81) % 9] : rng() % p;
    CHECK((uint64_t)F::diff_mulw(x, y, F::to_w((T)c)) == (uint64_t)((((u128)a + p - b) % p * c) % p),
          "diff_mulw %d %llu %llu %llu\n", F::ID, (unsigned long long)a, (unsigned long long)b,
          (unsigned long long)c);
  }
}

int main() {
  std::mt19937_64 rng(5);
  check_field<BabyBear>(rng);
  check_field<KoalaBear>(rng);
  check_field<TeddyBear>(rng);
  check_field<Goldilocks>(rng);
  const uint64_t p = Goldilocks::P;
  for (int it = 0; it < 20000; it++) {
    uint64_t x = it < 3 ? (uint64_t[]){0, 1, p - 1}[it] : rng() % p;
    uint64_t y = it < 9 ? (uint64_t[]){0, 1, p - 1}[it % 3] : rng() % p;
    // arbitrary 64-bit words, including the non-canonical range [p, 2^64)
    uint64_t z = it < 6 ? (uint64_t[]){0, 1, p - 1, p, ~0ull, ~0ull - 1}[it] : (it & 1 ? rng() : p + rng() % (~0ull - p));
    uint64_t pw = 1;
    for (uint32_t k = 0; k < 192; k++) {
      CHECK(Goldilocks::mul_pow2(x, k) == (uint64_t)(((u128)x * pw) % p), "mul_pow2 k=%u\n", k);
      CHECK(Goldilocks::diff_pow2(x, y, k) == (uint64_t)((((u128)x + p - y) % p * pw) % p),
            "diff_pow2 k=%u\n", k);
      pw = (uint64_t)(((u128)pw * 2) % p);
    }
    for (uint32_t m = 1; m <= 32; m++)  // any 64-bit input
      CHECK((uint64_t)(((u128)Goldilocks::div_pow2(z, m) << m) % p) == (uint64_t)(z % p) &&
                Goldilocks::div_pow2(z, m) < p,
            "div_pow2 m=%u\n", m);
    for (uint32_t j = 0; j < 32; j++) {
      uint64_t r = Goldilocks::neg_mul_pow2_hi(z, j);
      uint64_t want = (uint64_t)(((u128)(z % p) * h_pow<Goldilocks>(2, 32 + j)) % p);
      CHECK(r < p && (r + want) % p == 0, "neg_mul_pow2_hi j=%u\n", j);
    }
    for (uint32_t k = 1; k < 32; k++)
      CHECK(Goldilocks::mul_pow2_lo(x, k) == (uint64_t)(((u128)x << k) % p), "mul_pow2_lo k=%u\n", k);
  }
  if (fails) {
    std::printf("%d failures\n", fails);
    return 1;
  }
  std::printf("field arithmetic ok\n");
  return 0;
}
