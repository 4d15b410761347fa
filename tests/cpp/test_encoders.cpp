// C++ mirror of the reference's unit tests (src/lib.rs:9-181), run against
// the gfx950 engine through include/omrntt.hpp, with the CPU oracle
// (oracle/liborc.so) in the role of ArkRadix2 and Dft.  Extended from
// BabyBear (src/lib.rs:14) to all four fields.
#include <cstdio>
#include <cstdlib>
#include <random>
#include <string>
#include <vector>

#include "../../include/omrntt.hpp"

extern "C" {
void orc_ntt(int f, uint64_t *buf, uint64_t n, uint64_t omega);
void orc_dft(int f, uint64_t *buf, uint64_t n, uint64_t omega);
void orc_domain(int f, uint64_t n, uint64_t omega, uint64_t *twiddles, uint64_t *bitrev_twiddles);
}

using namespace omrntt;

static const size_t POWERS_OF_TWO[] = {2, 4, 8, 16, 32, 64, 128, 256, 512, 1024, 2048,
                                       4096, 8192, 16384, 32768, 65536};
static const size_t EVEN_POWERS_OF_TWO[] = {4, 16, 64, 256, 1024, 4096, 16384, 65536};
static const size_t POWERS_OF_TWO_DIV_BY_3[] = {8, 64, 512, 4096, 32768};

static int failures = 0;
#define EXPECT(cond, ...)                  \
  do {                                     \
    if (!(cond)) {                         \
      failures++;                          \
      std::fprintf(stderr, __VA_ARGS__);   \
      std::fprintf(stderr, "\n");          \
    }                                      \
  } while (0)

static uint64_t modulus(int id) {
  uint64_t p;
  omr_field_info(id, &p, nullptr, nullptr);
  return p;
}

template <class F>
std::vector<typename F::T> gen_random(size_t n, std::mt19937_64 &rng) {
  const uint64_t p = modulus(F::id);
  std::vector<typename F::T> v(n);
  for (auto &x : v) {
    uint64_t r;
    do r = rng(); while (r >= (~0ull / p) * p);
    x = (typename F::T)(r % p);
  }
  return v;
}

template <class F>
std::vector<typename F::T> ark(const std::vector<typename F::T> &in, const NttDomain<F> &d) {
  std::vector<uint64_t> t(in.begin(), in.end());
  orc_ntt(F::id, t.data(), t.size(), d.omega);
  return std::vector<typename F::T>(t.begin(), t.end());
}

// assert_agrees_with_ark (src/lib.rs:34-54)
template <class F, size_t K>
void assert_agrees_with_ark(const NttEncoder<F> &enc, const size_t (&sizes)[K]) {
  std::mt19937_64 rng(1);
  for (size_t n : sizes) {
    NttDomain<F> domain(n);
    auto input = gen_random<F>(n, rng);
    auto expected = ark<F>(input, domain);
    auto actual = input;
    enc.ntt(actual, domain);
    EXPECT(expected == actual, "%s: ntt mismatch at N=%zu (field %d)", enc.name(), n, F::id);
  }
}

// assert_prefix_agrees_with_ark (src/lib.rs:56-81)
template <class F>
void assert_prefix_agrees_with_ark(const NttEncoder<F> &enc, size_t s, const std::vector<size_t> &cases) {
  std::mt19937_64 rng(2);
  for (size_t n : cases) {
    size_t m = std::min(2 * s, n);
    NttDomain<F> domain(n);
    auto input = gen_random<F>(n, rng);
    auto expected = ark<F>(input, domain);
    auto actual = input;
    enc.ntt(actual, domain);
    bool ok = std::equal(actual.begin(), actual.begin() + m, expected.begin());
    EXPECT(ok, "%s: first %zu outputs mismatch at N=%zu (field %d)", enc.name(), m, n, F::id);
  }
}

template <class F>
void run_field() {
  // dft_agrees_with_ark (src/lib.rs:85-88)
  {
    std::mt19937_64 rng(3);
    for (size_t n : {64, 128, 256}) {
      NttDomain<F> d(n);
      auto in = gen_random<F>(n, rng);
      std::vector<uint64_t> t(in.begin(), in.end());
      orc_dft(F::id, t.data(), n, d.omega);
      auto got = in;
      Dft<F>().ntt(got, d);
      EXPECT(std::vector<typename F::T>(t.begin(), t.end()) == got, "Dft mismatch N=%zu", n);
      EXPECT(ark<F>(in, d) == got, "Dft vs ark N=%zu", n);
    }
  }
  // full encoders (src/lib.rs:114-139)
  assert_agrees_with_ark(LambdaBowers<F>(), POWERS_OF_TWO);
  assert_agrees_with_ark(WinterfellFourStep<F>(), POWERS_OF_TWO);
  assert_agrees_with_ark(Plonky3Radix2LayerSplit<F>(), POWERS_OF_TWO);
  assert_agrees_with_ark(ArkRadix2<F>(), POWERS_OF_TWO);
  assert_agrees_with_ark(LambdaRadix4<F>(), EVEN_POWERS_OF_TWO);
  assert_agrees_with_ark(TfheStockhamRadix8<F>(), POWERS_OF_TWO_DIV_BY_3);
  // partial encoders (src/lib.rs:141-165)
  std::vector<size_t> pw(std::begin(POWERS_OF_TWO), std::end(POWERS_OF_TWO));
  for (size_t s : {2, 4, 8, 16, 50}) {
    assert_prefix_agrees_with_ark(DftPartial<F>(s), s, {64, 128, 256});
    assert_prefix_agrees_with_ark(TransformDecomposition<F>(s), s, pw);
    assert_prefix_agrees_with_ark(TransformDecompositionV2<F>(s), s, {1024, 4096, 32768, 65536});
  }
  // NttDomain's public tables (src/encoder.rs:14-18) against the oracle's restatement
  for (size_t n : {(size_t)1, (size_t)2, (size_t)8, (size_t)4096, (size_t)65536}) {
    NttDomain<F> d(n);
    std::vector<uint64_t> tw(n), br(n / 2 ? n / 2 : 1);
    orc_domain(F::id, n, d.omega, tw.data(), br.data());
    const auto gt = d.twiddles();
    const auto gb = d.bitrev_twiddles();
    EXPECT(gt.size() == n && gb.size() == n / 2, "domain table sizes n=%zu", n);
    for (size_t k = 0; k < gt.size(); k++) EXPECT(gt[k] == tw[k], "twiddles n=%zu k=%zu", n, k);
    for (size_t i = 0; i < gb.size(); i++) EXPECT(gb[i] == br[i], "bitrev_twiddles n=%zu i=%zu", n, i);
  }
  // registry (src/encoders/mod.rs:32-54)
  EXPECT(encoders::all<F>(12).size() == 6, "all(12)");
  EXPECT(encoders::all<F>(7).size() == 4, "all(7)");
  EXPECT(encoders::all_partial<F>(50).size() == 3, "all_partial");
  // panics of the reference become exceptions (src/encoder.rs:24, lambdaworks_radix4.rs:17)
  bool threw = false;
  try {
    NttDomain<F> bad(12);
  } catch (const Error &e) {
    threw = e.status == OMR_E_SIZE;
  }
  EXPECT(threw, "NttDomain(12) must fail");
  threw = false;
  try {
    std::vector<typename F::T> b(8);
    LambdaRadix4<F>().ntt(b, NttDomain<F>(8));
  } catch (const Error &e) {
    threw = e.status == OMR_E_RADIX;
  }
  EXPECT(threw, "LambdaRadix4 at N=8 must fail");
}

int main() {
  run_field<BabyBear>();
  run_field<KoalaBear>();
  run_field<Goldilocks>();
  run_field<TeddyBear>();
  if (failures) {
    std::fprintf(stderr, "%d failures\n", failures);
    return 1;
  }
  std::printf("all encoder tests passed\n");
  return 0;
}
