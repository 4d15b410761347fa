// partial.hip -- partial-output transforms (prefix and stride subset).
//
// The reference computes out[0 .. min(2s, N)) by three encoders:
//   DftPartial            direct sums          (src/encoders/dft.rs:40-88)
//   TransformDecomposition(V2)  n2 sub-NTTs of size n1 on x[r::n2] and a
//                         twiddled recombine   (src/encoders/transform_decomposition.rs:26-97)
// Device methods:
//   1  direct: chunked sums of x[i] w^(ik), then a row reduction;
//   2  decomposition: one Stockham pass of radix n1 over stride n2 computes
//      Y[k1 n2 + r] = NTT_n1(x[r::n2])[k1] for the needed rows k1 only, then
//      out[k] = sum_r w^(rk) Y[(k mod n1) n2 + r];
//   3  full multi-pass transform whose last pass stores only k < out_len.
// Stride subset out[j] = X[j N/m]: fold y[i] = sum_t x[i + m t], then an
// m-point transform with w^(N/m).
#include <algorithm>

#include "dispatch.h"
#include "internal.h"

namespace omr {

static int ilog2u(uint64_t n) { return 63 - __builtin_clzll(n); }

// One Stockham pass of radix 2^logR over stride 2^(logN - logR) (the first
// pass of an N-point transform: no inter-pass twiddle), storing outputs
// with index < store_limit.
static omr_status first_pass(omr_ctx *ctx, int f, const void *in, void *out, uint32_t logN,
                             int logR, uint64_t omega, uint64_t store_limit, hipStream_t s) {
  const int es = field_bytes(f);
  const TwTables *tab;
  omr_status st = get_tables(ctx, f, logN, omega, &tab);
  if (st != OMR_OK) return st;
  const int logE = logR <= 2 ? logR : std::min(4, logR);
  const int logC = pass_log_cols((int)logN, logR, logE, es);
  PassArgs a{};
  a.src = in;
  a.dst = out;
  a.tw_lo = tab->lo;
  a.tw_hi = tab->hi;
  a.root = tab->root;
  a.lo_bits = tab->lo_bits;
  a.root_shift = tab->log_rmax - logR;
  a.tw_mid = tab->mid;
  a.log_rmax = tab->log_rmax;
  a.log_mid = tab->log_mid;
  a.logN = logN;
  a.logS = logN - logR;
  a.logC = logC;
  a.vec_stride_in = a.vec_stride_out = 0;
  a.store_limit = store_limit;
  a.nvec = 1;
  const uint64_t N = 1ull << logN;
  dim3 block((1u << logC) << (logR - logE));
  return launch_pass(f, logR, logE, tab->unit(), a, (N >> logR) >> logC, block, s) == hipSuccess ? OMR_OK
                                                                                                  : OMR_E_DEVICE;
}

omr_status partial_impl(omr_ctx *ctx, int f, const void *d_in, void *d_out, uint32_t logN,
                        uint64_t omega, uint64_t out_len, int mode, int method, uint32_t log_n1,
                        hipStream_t stream) {
  const int es = field_bytes(f);
  const uint64_t N = 1ull << logN;
  const int maxlr = max_log_radix(es);
  omr_status st;

  if (mode == 1) {  // stride subset, m = out_len
    const uint64_t m = out_len;
    Schedule sch;
    XformOpts o;
    if (m == N) {
      default_schedule(f, logN, sch);
      return run_transform(ctx, f, d_in, d_out, logN, omega, sch, 1, N, N, o, stream);
    }
    // fold in levels: each level sums ~32 rows of m per thread (x viewed as
    // rows of m elements) until one row is left in scratch[2]
    if ((st = ensure_buf(ctx->scratch[2], (size_t)m * es)) != OMR_OK) return st;
    const void *cur = d_in;
    uint64_t rows = N / m;
    int level = 0;
    while (rows > 1) {
      const uint64_t nch = std::min<uint64_t>((rows + 31) / 32, 65535);
      void *dst;
      if (nch == 1) {
        dst = ctx->scratch[2].p;
      } else {
        DevBuf &b = (level % 2 == 0) ? ctx->part : ctx->part2;
        if ((st = ensure_buf(b, (size_t)nch * m * es)) != OMR_OK) return st;
        dst = b.p;
      }
      if ((st = launch_fold(f, cur, dst, rows * m, m, (uint32_t)nch, stream)) != OMR_OK) return st;
      cur = dst;
      rows = nch;
      level++;
    }
    if (cur != ctx->scratch[2].p &&
        hipMemcpyAsync(ctx->scratch[2].p, cur, (size_t)m * es, hipMemcpyDeviceToDevice, stream) != hipSuccess)
      return OMR_E_DEVICE;
    const uint32_t logm = ilog2u(m);
    if (logm == 0) {
      return hipMemcpyAsync(d_out, ctx->scratch[2].p, es, hipMemcpyDeviceToDevice, stream) == hipSuccess
                 ? OMR_OK
                 : OMR_E_DEVICE;
    }
    default_schedule(f, logm, sch);
    return run_transform(ctx, f, ctx->scratch[2].p, d_out, logm, fpow_host(f, omega, N / m), sch, 1,
                         m, m, o, stream);
  }

  if (method == 0 && (out_len & (out_len - 1)) == 0 &&
      tuned_partial(f, logN, ilog2u(out_len), method, log_n1)) {
    // autotuned choice
  } else if (method == 0) {
    if (out_len <= 4) method = 1;
    else if (out_len <= (1ull << 13)) method = 2;
    else method = 3;
  }
  if (method == 1) {
    const TwTables *tab;
    if ((st = get_tables(ctx, f, logN, omega, &tab)) != OMR_OK) return st;
    const uint32_t chunk = 1u << 14;  // CHUNK of src/encoders/dft.rs:54
    const uint64_t nchunks = (N + chunk - 1) / chunk;
    if (nchunks == 1)
      return launch_dft_partial(f, d_in, d_out, N, (uint32_t)out_len, chunk, 1, tab, stream);
    if ((st = ensure_buf(ctx->part, (size_t)nchunks * out_len * es)) != OMR_OK) return st;
    if ((st = launch_dft_partial(f, d_in, ctx->part.p, N, (uint32_t)out_len, chunk, (uint32_t)nchunks,
                                 tab, stream)) != OMR_OK)
      return st;
    return launch_reduce_rows(f, ctx->part.p, d_out, (uint32_t)nchunks, out_len, out_len, stream);
  }
  if (method == 2) {
    if (log_n1 == 0) log_n1 = std::min<uint32_t>(logN, (uint32_t)maxlr);
    if (log_n1 == 0 || (int)log_n1 > maxlr || log_n1 > logN) return OMR_E_SCHEDULE;
    const uint64_t n1 = 1ull << log_n1, n2 = N >> log_n1;
    const uint64_t rows = std::min<uint64_t>(out_len, n1);
    if ((st = ensure_buf(ctx->scratch[2], (size_t)rows * n2 * es)) != OMR_OK) return st;
    if ((st = first_pass(ctx, f, d_in, ctx->scratch[2].p, logN, (int)log_n1, omega, rows * n2,
                         stream)) != OMR_OK)
      return st;
    const TwTables *tab;
    if ((st = get_tables(ctx, f, logN, omega, &tab)) != OMR_OK) return st;
    const uint64_t kt = n2 >= 256 ? 1 : (n2 >= 32 ? 8 : 64);
    const uint64_t kblocks = (out_len + kt - 1) / kt;
    uint64_t splits = std::max<uint64_t>(1, 2048 / kblocks);
    splits = std::min<uint64_t>(splits, std::max<uint64_t>(1, n2 / 256));
    splits = std::min<uint64_t>(splits, 65535);
    if (splits == 1)
      return launch_td_recombine(f, ctx->scratch[2].p, d_out, out_len, n2, log_n1, 1, logN, tab, stream);
    if ((st = ensure_buf(ctx->part, (size_t)splits * out_len * es)) != OMR_OK) return st;
    if ((st = launch_td_recombine(f, ctx->scratch[2].p, ctx->part.p, out_len, n2, log_n1,
                                  (uint32_t)splits, logN, tab, stream)) != OMR_OK)
      return st;
    return launch_reduce_rows(f, ctx->part.p, d_out, (uint32_t)splits, out_len, out_len, stream);
  }
  // method 3: full transform, last pass stores k < out_len only
  Schedule sch;
  if (!tuned_schedule(f, logN, sch)) default_schedule(f, logN, sch);
  XformOpts o;
  o.store_limit = out_len;
  return run_transform(ctx, f, d_in, d_out, logN, omega, sch, 1, N, out_len, o, stream);
}

}  // namespace omr
