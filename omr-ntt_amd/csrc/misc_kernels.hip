// misc_kernels.hip -- field dispatch for the non-pass kernels of misc.h.
#include "dispatch.h"
#include "internal.h"
#include "misc.h"

#include <algorithm>
#include <map>
#include <mutex>

namespace omr {

#define OMR_FIELD_SWITCH(f, BODY)          \
  switch (f) {                             \
    case kBabyBear: {                      \
      using F = BabyBear;                  \
      BODY;                                \
    } break;                               \
    case kKoalaBear: {                     \
      using F = KoalaBear;                 \
      BODY;                                \
    } break;                               \
    case kGoldilocks: {                    \
      using F = Goldilocks;                \
      BODY;                                \
    } break;                               \
    case kTeddyBear: {                     \
      using F = TeddyBear;                 \
      BODY;                                \
    } break;                               \
    default:                               \
      return OMR_E_FIELD;                  \
  }

static omr_status last_launch() { return hipGetLastError() == hipSuccess ? OMR_OK : OMR_E_DEVICE; }
static unsigned blocks(uint64_t n, unsigned b) { return (unsigned)((n + b - 1) / b); }
static Tw2 tw2_of(const TwTables *t) { return Tw2{t->lo, t->hi, t->lo_bits}; }

omr_status launch_scale(int f, const void *in, void *out, uint64_t n, uint64_t w, int apply,
                        hipStream_t s) {
  OMR_FIELD_SWITCH(f, {
    using T = typename F::T;
    hipLaunchKernelGGL(k_scale<F>, dim3(blocks(n, 256)), dim3(256), 0, s, (const T *)in, (T *)out, n,
                       (T)w, apply);
  });
  return last_launch();
}

omr_status launch_dft_full(int f, const void *in, void *out, uint64_t n, uint32_t logN,
                           const TwTables *t, hipStream_t s) {
  OMR_FIELD_SWITCH(f, {
    using T = typename F::T;
    hipLaunchKernelGGL(k_dft_full<F>, dim3(blocks(n, 256)), dim3(256), 0, s, (const T *)in, (T *)out,
                       n, logN, tw2_of(t));
  });
  return last_launch();
}

omr_status launch_dft_partial(int f, const void *in, void *part, uint64_t n, uint32_t out_len,
                              uint32_t chunk, uint32_t nchunks, const TwTables *t, hipStream_t s) {
  OMR_FIELD_SWITCH(f, {
    using T = typename F::T;
    hipLaunchKernelGGL(k_dft_partial_chunks<F>, dim3(nchunks), dim3(256), 0, s, (const T *)in,
                       (T *)part, n, out_len, chunk, tw2_of(t));
  });
  return last_launch();
}

omr_status launch_reduce_rows(int f, const void *part, void *out, uint32_t nrows, uint64_t ld,
                              uint64_t ncols, hipStream_t s) {
  OMR_FIELD_SWITCH(f, {
    using T = typename F::T;
    hipLaunchKernelGGL(k_reduce_rows<F>, dim3(blocks(ncols, 256)), dim3(256), 0, s,
                       (const T *)part, (T *)out, nrows, ld, ncols);
  });
  return last_launch();
}

omr_status launch_td_recombine(int f, const void *Y, void *part, uint64_t out_len, uint64_t n2,
                               uint32_t logN1, uint32_t splits, uint32_t logN, const TwTables *t,
                               hipStream_t s) {
  const uint64_t rps = (n2 + splits - 1) / splits;
  OMR_FIELD_SWITCH(f, {
    using T = typename F::T;
    if (n2 >= 256) {
      hipLaunchKernelGGL((k_td_recombine<F, 1>), dim3(blocks(out_len, 1), splits), dim3(256), 0, s,
                         (const T *)Y, (T *)part, out_len, n2, logN1, rps, logN, tw2_of(t));
    } else if (n2 >= 32) {
      hipLaunchKernelGGL((k_td_recombine<F, 8>), dim3(blocks(out_len, 8), splits), dim3(256), 0, s,
                         (const T *)Y, (T *)part, out_len, n2, logN1, rps, logN, tw2_of(t));
    } else {
      hipLaunchKernelGGL((k_td_recombine<F, 64>), dim3(blocks(out_len, 64), splits), dim3(256), 0, s,
                         (const T *)Y, (T *)part, out_len, n2, logN1, rps, logN, tw2_of(t));
    }
  });
  return last_launch();
}

omr_status launch_fold(int f, const void *x, void *part, uint64_t n, uint64_t m, uint32_t nchunks,
                       hipStream_t s, uint32_t nvec, uint64_t xs, uint64_t ps) {
  const uint64_t nt = n / m;
  const uint64_t tpc = (nt + nchunks - 1) / nchunks;
  OMR_FIELD_SWITCH(f, {
    using T = typename F::T;
    hipLaunchKernelGGL(k_fold<F>, dim3(blocks(m, 256), nchunks, nvec), dim3(256), 0, s, (const T *)x,
                       (T *)part, n, m, tpc, xs, ps);
  });
  return last_launch();
}

bool stride_fused_plan(int es, uint64_t N, uint64_t m, StrideFusedPlan *plan) {
  if (m < 16 || m > (1ull << kSfMaxLogM) || m >= N || (m & (m - 1)) || (N & (N - 1))) return false;
  const uint64_t tpr = m / (16 / es);  // threads per row (16-byte loads)
  const uint64_t rsub = kSfThreads / tpr;
  const uint64_t rows = N / m;
  if (rows < rsub) return false;
  // powers of two throughout: rows_per_wg is a multiple of rsub, grid of group
  plan->grid = (uint32_t)std::min<uint64_t>((uint64_t)kSfGroup * kSfMaxGroups, rows / rsub);
  plan->group = std::min<uint32_t>(kSfGroup, plan->grid);
  plan->rows_per_wg = (uint32_t)(rows / plan->grid);
  plan->logm = (uint32_t)__builtin_ctzll(m);
  return true;
}

omr_status launch_stride_fused(int f, const void *x, void *part, void *part2, uint32_t *cnt, void *out,
                               const void *root_m, const StrideFusedPlan &plan, hipStream_t s) {
  OMR_FIELD_SWITCH(f, {
    using T = typename F::T;
    hipLaunchKernelGGL(k_stride_fused<F>, dim3(plan.grid), dim3(kSfThreads), 0, s, (const T *)x, (T *)part,
                       (T *)part2, cnt, (T *)out, (const T *)root_m, plan.logm, plan.rows_per_wg, plan.group);
  });
  return last_launch();
}

omr_status launch_sum_contig(int f, const void *part, void *out, uint64_t n, uint64_t m, uint32_t nvec,
                             uint64_t os, hipStream_t s) {
  OMR_FIELD_SWITCH(f, {
    using T = typename F::T;
    if (n <= 32)  // short sums: a thread each
      hipLaunchKernelGGL(k_sum_short<F>, dim3((uint32_t)((m + 255) / 256), nvec), dim3(256), 0, s,
                         (const T *)part, (T *)out, (uint32_t)n, m, os);
    else
      hipLaunchKernelGGL(k_sum_contig<F>, dim3((uint32_t)m, nvec), dim3(256), 0, s, (const T *)part, (T *)out, n,
                         m, os);
  });
  return last_launch();
}

omr_status launch_domain_tables(int f, void *tw, void *brev, uint64_t n, uint32_t logN, const TwTables *t,
                                hipStream_t s) {
  if (n == 0) return OMR_OK;
  OMR_FIELD_SWITCH(f, {
    using T = typename F::T;
    hipLaunchKernelGGL(k_domain_tables<F>, dim3(blocks(n, 256)), dim3(256), 0, s, (T *)tw, (T *)brev, n, logN,
                       tw2_of(t));
  });
  return last_launch();
}

omr_status launch_fill_u32(uint32_t *p, uint64_t n, uint32_t v, hipStream_t s) {
  if (n == 0) return OMR_OK;
  hipLaunchKernelGGL(k_fill_u32, dim3(blocks(n, 256)), dim3(256), 0, s, p, n, v);
  return last_launch();
}

omr_status launch_narrow(const uint64_t *in, uint32_t *out, uint64_t n, hipStream_t s) {
  if (n == 0) return OMR_OK;
  hipLaunchKernelGGL(k_narrow, dim3(blocks(n, 256)), dim3(256), 0, s, in, out, n);
  return last_launch();
}

omr_status launch_widen(const uint32_t *in, uint64_t *out, uint64_t n, hipStream_t s) {
  if (n == 0) return OMR_OK;
  hipLaunchKernelGGL(k_widen, dim3(blocks(n, 256)), dim3(256), 0, s, in, out, n);
  return last_launch();
}

omr_status launch_field_op(int f, int op, const void *a, const void *b, void *out, uint64_t n,
                           hipStream_t s) {
  if (n == 0) return OMR_OK;
  OMR_FIELD_SWITCH(f, {
    using T = typename F::T;
    hipLaunchKernelGGL(k_field_op<F>, dim3(blocks(n, 256)), dim3(256), 0, s, (const T *)a,
                       (const T *)b, (T *)out, n, op);
  });
  return last_launch();
}

hipError_t launch_pass(int field, int logR, int logE, int unit, const PassArgs &a, uint64_t total_tiles,
                       dim3 block, hipStream_t stream) {
  // the pass reads its roots from the pyramid's level logR
  if (!a.root_pyr || logR < 0 || (uint32_t)logR > a.log_rmax) return hipErrorInvalidValue;
  int epi = pass_epilogue(a);
  if (epi == kEpiGeneral && lde4_ok(a, logR, logE)) epi = kEpiLde;
  if (fs4096_ok(field, logR, unit, a.logN, a.logS, epi))
    return launch_pass_fs4096(unit, epi, a, (((uint64_t)1 << a.logN) >> 15) * a.nvec, stream);
  switch (field) {
    case kBabyBear: return launch_pass_babybear(logR, logE, epi, a, total_tiles, block, stream);
    case kKoalaBear: return launch_pass_koalabear(logR, logE, epi, a, total_tiles, block, stream);
    case kGoldilocks:
      if (unit == 39) return launch_pass_goldilocks_u39(logR, logE, epi, a, total_tiles, block, stream);
      if (unit == 153) return launch_pass_goldilocks_u153(logR, logE, epi, a, total_tiles, block, stream);
      return launch_pass_goldilocks_u0(logR, logE, epi, a, total_tiles, block, stream);
    case kTeddyBear: return launch_pass_teddybear(logR, logE, epi, a, total_tiles, block, stream);
  }
  return hipErrorInvalidValue;
}

unsigned pass_grid(const void *, unsigned, uint64_t total_tiles) {
  return (unsigned)std::min<uint64_t>(total_tiles, 0x7FFFFFFFull);
}

}  // namespace omr
