// field.h -- prime-field arithmetic for the NTT engine (host and gfx950).
//
// The four fields of src/fields.rs:3-17.  Elements cross the C ABI and live
// in HBM as canonical integers in [0, p).  Multipliers (twiddles, scales) are
// kept in "multiplier form" W so that one product gives a canonical result:
//   * 31/32-bit fields: Montgomery form w * 2^32 mod p, and
//     mulw(x, W) = x * W * 2^-32 = x * w  (Montgomery reduction, R = 2^32).
//   * Goldilocks: W = w, mulw is the 64x64 product with the special
//     reduction 2^64 = 2^32 - 1, 2^96 = -1 (mod p).
// CDNA4 has no 64-bit integer multiply; products are built from 32-bit
// mul_lo / mul_hi (v_mul_lo_u32, v_mul_hi_u32, v_mad_u64_u32).
#pragma once
#include <stdint.h>

#if defined(__HIPCC__)
#include <hip/hip_runtime.h>
#define OMR_HD __host__ __device__ __forceinline__
#else
#define OMR_HD inline
#endif

namespace omr {

enum FieldId : int { kBabyBear = 0, kKoalaBear = 1, kGoldilocks = 2, kTeddyBear = 3 };

constexpr uint32_t inv32(uint32_t p) {  // p^-1 mod 2^32 by Newton iteration
  uint32_t x = p;
  for (int i = 0; i < 5; i++) x *= 2u - p * x;
  return x;
}

// 31/32-bit prime with Montgomery multipliers.  BIG = p > 2^31 (TeddyBear),
// where a + b can carry out of 32 bits.
template <uint32_t P_, uint32_t G_, int TWO_ADICITY_, int ID_>
struct Mont32 {
  using T = uint32_t;
  static constexpr uint32_t P = P_;
  static constexpr uint64_t G = G_;
  static constexpr int TWO_ADICITY = TWO_ADICITY_;
  static constexpr int ID = ID_;
  static constexpr bool BIG = P_ > 0x80000000u;
  static constexpr uint32_t PINV = inv32(P_);
  static constexpr uint32_t R_MOD = (uint32_t)((1ull << 32) % P_);
  static constexpr uint32_t R2 = (uint32_t)(((unsigned __int128)1 << 64) % P_);

  OMR_HD static T add(T a, T b) {
    if constexpr (!BIG) {
      uint32_t s = a + b;
      uint32_t t = s - P;
      return t < s ? t : s;  // s >= P  <=>  s - P does not wrap
    } else {
      uint32_t nb = P - b;  // in (0, P]
      uint32_t r = a - nb;
      return a < nb ? r + P : r;
    }
  }
  // For p < 2^31 a difference d in (-p, p) held mod 2^32 is reduced by
  // min(d, d + p): a wrapped d is >= 2^32 - p > p while d + p < p, and an
  // unwrapped d < p is below d + p < 2^32.  For p > 2^31 the borrow decides.
  OMR_HD static T fix_diff(uint32_t d, bool borrow) {
    if constexpr (!BIG) {
      (void)borrow;
      const uint32_t e = d + P;
      return e < d ? e : d;
    } else {
      return borrow ? d + P : d;
    }
  }
  OMR_HD static T sub(T a, T b) { return fix_diff(a - b, a < b); }
  OMR_HD static T neg(T a) { return a ? P - a : 0; }
  // Montgomery product a * b * 2^-32 mod p for a < p, b < 2^32 (or, for
  // p < 2^31, any a < 2^32 with b < p: either way a b < p 2^32, so the
  // high-half difference lies in (-p, p)).
  OMR_HD static T mont(T a, T b) {
    uint64_t t = (uint64_t)a * b;
    uint32_t tlo = (uint32_t)t, thi = (uint32_t)(t >> 32);
    uint32_t m = tlo * PINV;
    uint32_t mh = (uint32_t)(((uint64_t)m * P) >> 32);
    return fix_diff(thi - mh, thi < mh);
  }
  OMR_HD static T mulw(T a, T w) { return mont(a, w); }
  // (a - b) * w for canonical a, b and multiplier w < p.  For p < 2^31 the
  // unreduced a + (p - b) < 2p < 2^32 is a valid Montgomery operand: with
  // w < p the product stays below p 2^32, so mont's result is canonical.
  // One instruction fewer than mulw(sub(a, b), w).
  OMR_HD static T diff_mulw(T a, T b, T w) {
    if constexpr (!BIG) return mont(a + (P - b), w);
    else return mont(sub(a, b), w);
  }
  OMR_HD static T wmul(T w1, T w2) { return mont(w1, w2); }  // W x W -> W
  OMR_HD static T to_w(T a) { return mont(a, R2); }
  OMR_HD static T from_w(T w) { return mont(w, 1); }
  OMR_HD static T mul(T a, T b) { return mont(mont(a, b), R2); }
  OMR_HD static T one_w() { return R_MOD; }
  static constexpr bool kPow2Roots = false;
};

using BabyBear = Mont32<2013265921u, 31, 27, kBabyBear>;
using KoalaBear = Mont32<2130706433u, 3, 24, kKoalaBear>;
using TeddyBear = Mont32<3221225473u, 5, 30, kTeddyBear>;

struct Goldilocks {
  using T = uint64_t;
  static constexpr uint64_t P = 0xFFFFFFFF00000001ull;
  static constexpr uint64_t EPS = 0xFFFFFFFFull;  // 2^64 mod p
  static constexpr uint64_t G = 7;
  static constexpr int TWO_ADICITY = 32;
  static constexpr int ID = kGoldilocks;

  OMR_HD static T pack(uint32_t lo, uint32_t hi) { return ((uint64_t)hi << 32) | lo; }
  OMR_HD static T add(T a, T b) {  // canonical in and out
    uint64_t s = a + b;
    bool carry = s < a;
    return (carry || s >= P) ? s - P : s;  // on carry s - P == s + EPS mod 2^64
  }
  // 32-bit borrow chain (v_sub_co / v_subb); a borrow out of the top word
  // adds P (mod 2^64).
  OMR_HD static T sub(T a, T b) {
    uint32_t b1, b2;
    const uint32_t lo = __builtin_subc((uint32_t)a, (uint32_t)b, 0u, &b1);
    const uint32_t hi = __builtin_subc((uint32_t)(a >> 32), (uint32_t)(b >> 32), b1, &b2);
    return pack(lo, hi) + (b2 ? P : 0);
  }
  OMR_HD static T neg(T a) { return a ? P - a : 0; }
  // 128-bit product from four 32x32->64 partial products.
  OMR_HD static void mul128(uint64_t a, uint64_t b, uint64_t &lo, uint64_t &hi) {
    uint32_t a0 = (uint32_t)a, a1 = (uint32_t)(a >> 32);
    uint32_t b0 = (uint32_t)b, b1 = (uint32_t)(b >> 32);
    uint64_t p00 = (uint64_t)a0 * b0;
    uint64_t p01 = (uint64_t)a0 * b1 + (p00 >> 32);
    uint64_t p10 = (uint64_t)a1 * b0 + (uint32_t)p01;
    uint64_t p11 = (uint64_t)a1 * b1 + (p01 >> 32) + (p10 >> 32);
    lo = (p10 << 32) | (uint32_t)p00;
    hi = p11;
  }
  // x = lo + hi * 2^64 -> canonical, for any lo, hi < 2^64.  With
  // 2^64 = 2^32 - 1 and 2^96 = -1:  x = lo + hl 2^32 - hl - hh.  Writing
  // lo_hi + hl = s + c 2^32 (c a carry) and 2^64 = 2^32 - 1 once more,
  //   x = (H:lo_lo) - D,  H = s + c < 2^32,  D = hl + hh + c < 2^33,
  // a value in (-2^33, 2^64): a borrow adds P, otherwise one conditional
  // subtraction of P.  32-bit carry chains (v_add_co / v_addc / v_sub_co /
  // v_subb) keep the carries in condition registers.
  OMR_HD static T reduce128(uint64_t lo, uint64_t hi) {
    const uint32_t hh = (uint32_t)(hi >> 32), hl = (uint32_t)hi;
    uint32_t c, c2, b1, b2;
    const uint32_t s = __builtin_addc((uint32_t)(lo >> 32), hl, 0u, &c);
    const uint32_t h = s + c;
    const uint32_t dl = __builtin_addc(hl, hh, c, &c2);
    const uint32_t rl = __builtin_subc((uint32_t)lo, dl, 0u, &b1);
    const uint32_t rh = __builtin_subc(h, c2, b1, &b2);
    const uint64_t r = pack(rl, rh) + (b2 ? P : 0);
    return r >= P ? r - P : r;
  }
  OMR_HD static T mul(T a, T b) {
    uint64_t lo, hi;
    mul128(a, b, lo, hi);
    return reduce128(lo, hi);
  }
  OMR_HD static T mulw(T a, T w) { return mul(a, w); }
  OMR_HD static T diff_mulw(T a, T b, T w) { return mul(sub(a, b), w); }
  OMR_HD static T wmul(T w1, T w2) { return mul(w1, w2); }
  OMR_HD static T to_w(T a) { return a; }
  OMR_HD static T from_w(T w) { return w; }
  OMR_HD static T one_w() { return 1; }

  // 2 has multiplicative order 192 (2^96 = -1), so every root of unity of
  // order dividing 64 is a power of 2 and multiplying by it is a shift plus
  // the reduction -- no 64x64 product.
  static constexpr bool kPow2Roots = true;
  // x * 2^k for 0 < k < 64
  OMR_HD static T shl_reduce(T x, uint32_t k) { return reduce128(x << k, x >> (64 - k)); }
  // x * 2^k for 0 < k < 32: x 2^k = lo + y2 2^64 with y2 < 2^k, and
  // 2^64 = 2^32 - 1, so one 32x32 multiply-add folds y2 in.
  // After a carry t < y2 (2^32 - 1) < 2^63, so t + 2^64 = t + (2^32 - 1) is
  // below P; without one, t >= P subtracts P -- the same addend either way.
  OMR_HD static T mul_pow2_lo(T x, uint32_t k) {
    const uint64_t lo = x << k;
    const uint32_t y2 = (uint32_t)(x >> (64 - k));
    const uint64_t t = (uint64_t)y2 * 0xFFFFFFFFu + lo;  // carry out <=> t < lo
    return t + (((t < lo) | (t >= P)) ? EPS : 0);
  }
  // -(x * 2^(32 + j)) for 0 <= j < 32, any x < 2^64.  With y = x 2^j =
  // y0 + y1 2^32 + y2 2^64 (32-bit words) and 2^64 = 2^32 - 1, 2^96 = -1:
  //   x 2^(32+j) = (y0 + y1) 2^32 - y1 - y2 = H 2^32 - (y1 + y2 + c),
  // where y0 + y1 = A + c 2^32 and H = A + c < 2^32.  The negation
  // L - H 2^32 (L = y1 + y2 + c < 2^34) lies in (-P, 2^34): one borrow fix.
  OMR_HD static T neg_mul_pow2_hi(T x, uint32_t j) {
    const uint32_t xh = (uint32_t)(x >> 32);
    const uint32_t y0 = (uint32_t)x << j;
    const uint32_t y1 = (uint32_t)(x >> (32 - j));
    const uint32_t y2 = j ? xh >> (32 - j) : 0u;
    uint32_t c, c2, br;
    const uint32_t a = __builtin_addc(y0, y1, 0u, &c);
    const uint32_t h = a + c;
    const uint32_t l = __builtin_addc(y1, y2, c, &c2);
    const uint32_t hi = __builtin_subc(c2, h, 0u, &br);
    return pack(l, hi) + (br ? P : 0);
  }
  // x * 2^-m for 0 < m <= 32, any x < 2^64.  With x = q 2^m + r (r < 2^m)
  // and 2^-m = 2^(32-m) - 2^(64-m) (multiply by 2^m: 2^32 - 2^64 = 1 - p):
  //   x 2^-m = q + r 2^(32-m) - r 2^(64-m) = w - u,
  // u = r 2^(64-m) = (x << (64-m)) <= P - 1 has a zero low word, so one
  // borrow fix gives the canonical value.
  OMR_HD static T div_pow2(T x, uint32_t m) {
    const uint32_t uh = (uint32_t)x << (32 - m);
    const uint64_t w = (x >> m) + uh;
    uint32_t br;
    const uint32_t hi = __builtin_subc((uint32_t)(w >> 32), uh, 0u, &br);
    return pack((uint32_t)w, hi) + (br ? P : 0);
  }
  // (a - b) * 2^k for 0 <= k < 192 with k a constant after unrolling: the
  // sign of 2^k (2^96 = -1) goes into the order of the subtraction, and for
  // 64 <= k mod 96 < 96, 2^k = -2^-(96 - k mod 96).  Canonical in and out.
  OMR_HD static T diff_pow2(T a, T b, uint32_t k) {
    const bool neg = k >= 96;
    if (neg) k -= 96;
    if (k == 0) return neg ? sub(b, a) : sub(a, b);
    if (k < 32) return mul_pow2_lo(neg ? sub(b, a) : sub(a, b), k);
    if (k < 64) return neg_mul_pow2_hi(neg ? sub(a, b) : sub(b, a), k - 32);
    return div_pow2(neg ? sub(a, b) : sub(b, a), 96 - k);
  }
  // x * 2^k mod p for 0 <= k < 192, x canonical
  OMR_HD static T mul_pow2(T x, uint32_t k) {
    const bool negate = k >= 96;
    if (negate) k -= 96;
    if (k >= 64) {
      x = shl_reduce(x, 32);
      k -= 32;
    }
    if (k) x = shl_reduce(x, k);
    return negate ? neg(x) : x;
  }
};

// Host helpers on canonical values (uint64_t for every field).
template <class F>
inline uint64_t h_mul(uint64_t a, uint64_t b) {
  return (uint64_t)(((unsigned __int128)a * b) % F::P);
}
template <class F>
inline uint64_t h_pow(uint64_t a, uint64_t e) {
  uint64_t r = 1;
  while (e) {
    if (e & 1) r = h_mul<F>(r, a);
    a = h_mul<F>(a, a);
    e >>= 1;
  }
  return r;
}
template <class F>
inline uint64_t h_inv(uint64_t a) {
  return h_pow<F>(a, (uint64_t)F::P - 2);
}
// ark-ff get_root_of_unity convention: g^((p-1)/n) for n = 2^k.
template <class F>
inline uint64_t h_root_of_unity(uint64_t n) {
  return h_pow<F>(F::G, ((uint64_t)F::P - 1) / n);
}

}  // namespace omr
