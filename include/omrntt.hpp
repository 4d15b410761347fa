// omrntt.hpp -- C++ host mirror of the reference crate's operator API over
// the C ABI of omrntt.h.
//
//   reference (Rust)                                   here (C++)
//   NttDomain<F>::new(N)  src/encoder.rs:23-52          omrntt::NttDomain<F>(N)
//   trait NttEncoder<F>   src/encoder.rs:70-83          omrntt::NttEncoder<F>
//   ArkRadix2 .. TransformDecompositionV2               omrntt::ArkRadix2 .. (same names)
//   encoders::all / all_partial  src/encoders/mod.rs:32-54   omrntt::encoders::all / all_partial
//   BabyBear, KoalaBear, Goldilocks, TeddyBear  src/fields.rs:3-19
//
// Elements are canonical integers (uint32_t, or uint64_t for Goldilocks).
// The reference panics on invalid input; this mirror throws omrntt::Error
// carrying the same condition as an omr_status.
#pragma once
#include <cstdint>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

#include "omrntt.h"

namespace omrntt {

struct Error : std::runtime_error {
  omr_status status;
  Error(omr_status s, const std::string &what)
      : std::runtime_error(what + ": " + omr_status_str(s)), status(s) {}
};

inline void check(omr_status s, const char *what) {
  if (s != OMR_OK) throw Error(s, what);
}

struct BabyBear {
  using T = uint32_t;
  static constexpr int id = OMR_BABYBEAR;
  static constexpr uint32_t TWO_ADICITY = 27;
};
struct KoalaBear {
  using T = uint32_t;
  static constexpr int id = OMR_KOALABEAR;
  static constexpr uint32_t TWO_ADICITY = 24;
};
struct Goldilocks {
  using T = uint64_t;
  static constexpr int id = OMR_GOLDILOCKS;
  static constexpr uint32_t TWO_ADICITY = 32;
};
struct TeddyBear {
  using T = uint32_t;
  static constexpr int id = OMR_TEDDYBEAR;
  static constexpr uint32_t TWO_ADICITY = 30;
};
using DefaultField = BabyBear;

// Process-wide engine context on device 0 (created on first use).
inline omr_ctx *context() {
  struct Holder {
    omr_ctx *c = nullptr;
    Holder() { check(omr_ctx_create(0, &c), "omr_ctx_create"); }
    ~Holder() {
      if (c) omr_ctx_destroy(c);
    }
  };
  static Holder h;
  return h.c;
}

// NttDomain::new (src/encoder.rs:23-52): N, log_N and omega.  The device
// derives its own twiddle tables from omega; the reference's public tables
// twiddles / bitrev_twiddles (src/encoder.rs:14-18) are built on request.
template <class F>
struct NttDomain {
  uint64_t N;
  uint32_t log_N;
  uint64_t omega;
  explicit NttDomain(uint64_t n) : N(n), log_N(0), omega(0) {
    check(omr_root_of_unity(F::id, n, &omega), "NttDomain::new");
    while ((1ull << log_N) < n) log_N++;
  }
  NttDomain(uint64_t n, uint64_t w) : NttDomain(n) { omega = w; }
  // twiddles[k] = omega^k, k < N
  std::vector<typename F::T> twiddles() const {
    std::vector<typename F::T> t(N);
    check(omr_domain_tables(context(), F::id, N, omega, t.data(), nullptr), "NttDomain::twiddles");
    return t;
  }
  // bitrev_twiddles[i] = omega^bitrev(i, log_N - 1), i < N/2
  std::vector<typename F::T> bitrev_twiddles() const {
    std::vector<typename F::T> t(N / 2);
    check(omr_domain_tables(context(), F::id, N, omega, nullptr, t.empty() ? nullptr : t.data()),
          "NttDomain::bitrev_twiddles");
    return t;
  }
};

// trait NttEncoder<F> (src/encoder.rs:70-74).
template <class F>
struct NttEncoder {
  virtual ~NttEncoder() = default;
  virtual void ntt(typename F::T *buf, size_t len, const NttDomain<F> &domain) const = 0;
  void ntt(std::vector<typename F::T> &buf, const NttDomain<F> &domain) const {
    ntt(buf.data(), buf.size(), domain);
  }
  virtual const char *name() const = 0;
};

template <class F>
struct GpuEncoder : NttEncoder<F> {
  int variant;
  uint64_t s;
  GpuEncoder(int v, uint64_t s_ = 0) : variant(v), s(s_) {}
  using NttEncoder<F>::ntt;
  void ntt(typename F::T *buf, size_t len, const NttDomain<F> &d) const override {
    if (len != d.N) throw Error(OMR_E_ARG, "buf.len() != N");
    check(omr_ntt(context(), F::id, variant, buf, d.N, d.omega, s), name());
  }
  const char *name() const override { return omr_variant_name(variant); }
};

#define OMRNTT_FULL(NAME, V) \
  template <class F>         \
  struct NAME : GpuEncoder<F> { NAME() : GpuEncoder<F>(V) {} };
#define OMRNTT_PARTIAL(NAME, V) \
  template <class F>            \
  struct NAME : GpuEncoder<F> { explicit NAME(uint64_t s) : GpuEncoder<F>(V, s) {} };

OMRNTT_FULL(ArkRadix2, OMR_ARK_RADIX2)
OMRNTT_FULL(Dft, OMR_DFT)
OMRNTT_FULL(LambdaBowers, OMR_LAMBDA_BOWERS)
OMRNTT_FULL(LambdaRadix4, OMR_LAMBDA_RADIX4)
OMRNTT_FULL(Plonky3Radix2LayerSplit, OMR_PLONKY3_LAYER_SPLIT)
OMRNTT_FULL(WinterfellFourStep, OMR_WINTERFELL_FOUR_STEP)
OMRNTT_FULL(TfheStockhamRadix8, OMR_TFHE_STOCKHAM_RADIX8)
OMRNTT_PARTIAL(DftPartial, OMR_DFT_PARTIAL)
OMRNTT_PARTIAL(TransformDecomposition, OMR_TRANSFORM_DECOMPOSITION)
OMRNTT_PARTIAL(TransformDecompositionV2, OMR_TRANSFORM_DECOMPOSITION_V2)
#undef OMRNTT_FULL
#undef OMRNTT_PARTIAL

namespace encoders {
// src/encoders/mod.rs:32-46: Radix4 only for even log N, Stockham8 only for
// log N divisible by 3, Dft excluded.
template <class F>
std::vector<std::unique_ptr<NttEncoder<F>>> all(uint32_t log_n) {
  std::vector<std::unique_ptr<NttEncoder<F>>> v;
  v.emplace_back(new ArkRadix2<F>());
  v.emplace_back(new LambdaBowers<F>());
  v.emplace_back(new WinterfellFourStep<F>());
  v.emplace_back(new Plonky3Radix2LayerSplit<F>());
  if (log_n % 2 == 0) v.emplace_back(new LambdaRadix4<F>());
  if (log_n % 3 == 0) v.emplace_back(new TfheStockhamRadix8<F>());
  return v;
}
// src/encoders/mod.rs:48-54
template <class F>
std::vector<std::unique_ptr<NttEncoder<F>>> all_partial(uint64_t s) {
  std::vector<std::unique_ptr<NttEncoder<F>>> v;
  v.emplace_back(new DftPartial<F>(s));
  v.emplace_back(new TransformDecomposition<F>(s));
  v.emplace_back(new TransformDecompositionV2<F>(s));
  return v;
}
}  // namespace encoders

}  // namespace omrntt
