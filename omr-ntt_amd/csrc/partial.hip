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

// Sums `rows` rows of m elements (row r at src + r m) into dst, for nvec
// independent row sets (set v at src + v src_stride, result at
// dst + v dst_stride).  Levels of ~32 rows per thread, ping-ponging between
// ctx->part and ctx->part2, the last level writing dst.
static omr_status fold_rows(omr_ctx *ctx, int f, const void *src, uint64_t rows, uint64_t m, uint32_t nvec,
                            uint64_t src_stride, void *dst, uint64_t dst_stride, hipStream_t stream) {
  const int es = field_bytes(f);
  omr_status st;
  constexpr uint32_t kMaxZ = 65535;  // grid z limit: sets in chunks
  if (nvec > kMaxZ) {
    for (uint32_t v0 = 0; v0 < nvec; v0 += kMaxZ) {
      const uint32_t nv = std::min(kMaxZ, nvec - v0);
      if ((st = fold_rows(ctx, f, (const char *)src + (size_t)v0 * src_stride * es, rows, m, nv, src_stride,
                          (char *)dst + (size_t)v0 * dst_stride * es, dst_stride, stream)) != OMR_OK)
        return st;
    }
    return OMR_OK;
  }
  const void *cur = src;
  uint64_t cur_stride = src_stride;
  int level = 0;
  do {
    const uint64_t nch = std::min<uint64_t>((rows + 31) / 32, 65535);
    void *out;
    uint64_t out_stride;
    if (nch == 1) {
      out = dst;
      out_stride = dst_stride;
    } else {
      DevBuf &b = (level % 2 == 0) ? ctx->part : ctx->part2;
      if ((st = ensure_buf(b, (size_t)nvec * nch * m * es)) != OMR_OK) return st;
      out = b.p;
      out_stride = nch * m;
    }
    if ((st = launch_fold(f, cur, out, rows * m, m, (uint32_t)nch, stream, nvec, cur_stride, out_stride)) != OMR_OK)
      return st;
    cur = out;
    cur_stride = out_stride;
    rows = nch;
    level++;
  } while (rows > 1);
  return OMR_OK;
}

// Prefix out[k] = X[k], k < out_len <= 2^log_n1, of nvec vectors (vector v at
// d_in + v in_stride, its prefix at d_out + v out_stride): one radix-2^log_n1
// pass whose epilogue folds w_N^(q k) Y[q][k] over each tile's columns
// (kEpiPrefix), then each vector's per-tile partial sums added up.  The input
// is read once and nothing N-sized is written.
static omr_status prefix_fused(omr_ctx *ctx, int f, const void *d_in, uint64_t in_stride, void *d_out,
                               uint64_t out_stride, uint32_t logN, uint64_t omega, uint64_t out_len,
                               uint32_t log_n1, uint32_t nvec, hipStream_t stream) {
  const int es = field_bytes(f);
  const TwTables *tab;
  omr_status st = get_tables(ctx, f, logN, omega, &tab);
  if (st != OMR_OK) return st;
  const int logR = (int)log_n1;
  const int logE = logR <= 2 ? logR : std::min(4, logR);
  const int logC = pass_log_cols((int)logN, logR, logE, es);
  const uint64_t tpv = ((1ull << logN) >> logR) >> logC;  // tiles per vector
  if ((st = ensure_buf(ctx->scratch[2], (size_t)nvec * tpv * out_len * es)) != OMR_OK) return st;
  PassArgs a{};
  a.src = d_in;
  a.dst = ctx->scratch[2].p;
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
  a.vec_stride_in = in_stride;
  a.vec_stride_out = 0;
  a.store_limit = out_len;
  a.flags = kPrefixSum;
  a.nvec = nvec;
  dim3 block((1u << logC) << (logR - logE));
  if (launch_pass(f, logR, logE, tab->unit(), a, tpv * nvec, block, stream) != hipSuccess) return OMR_E_DEVICE;
  // partials [vector][k][tile]: one block per (k, vector) sums a vector's tiles
  constexpr uint32_t kMaxY = 65535;
  for (uint32_t v0 = 0; v0 < nvec; v0 += kMaxY) {
    const uint32_t nv = std::min(kMaxY, nvec - v0);
    if ((st = launch_sum_contig(f, (const char *)ctx->scratch[2].p + (size_t)v0 * tpv * out_len * es,
                                (char *)d_out + (size_t)v0 * out_stride * es, tpv, out_len, nv, out_stride,
                                stream)) != OMR_OK)
      return st;
  }
  return OMR_OK;
}

// Sub-NTT size of the fused prefix for out_len: the smallest power of two
// >= out_len, at least 2^6 (fewer, larger tiles), at most 2^10 and N.
static uint32_t prefix_log_n1(uint32_t logN, uint64_t out_len) {
  uint32_t l = 6;
  while ((1ull << l) < out_len) l++;
  return std::min<uint32_t>(std::min<uint32_t>(l, (uint32_t)kPrefixMaxLogR), logN);
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
    // fold x viewed as N/m rows of m into scratch[2]
    if ((st = ensure_buf(ctx->scratch[2], (size_t)m * es)) != OMR_OK) return st;
    if ((st = fold_rows(ctx, f, d_in, N / m, m, 1, 0, ctx->scratch[2].p, 0, stream)) != OMR_OK) return st;
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
    if (out_len <= 4) {
      method = 1;
    } else if (out_len <= (1ull << kPrefixMaxLogR)) {
      method = 2;
      log_n1 = prefix_log_n1(logN, out_len);
    } else if (out_len <= (1ull << 13)) {
      method = 2;
    } else {
      method = 3;
    }
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
    if (out_len <= n1 && (int)log_n1 <= kPrefixMaxLogR)  // recombine fused into the pass
      return prefix_fused(ctx, f, d_in, N, d_out, out_len, logN, omega, out_len, log_n1, 1, stream);
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

// Prefix of every vector of a batch (see omr_partial_batch_device).
omr_status partial_batch_impl(omr_ctx *ctx, int f, const void *d_in, void *d_out, uint32_t logN,
                              uint64_t batch, uint64_t omega, uint64_t out_len, hipStream_t stream) {
  const uint64_t N = 1ull << logN;
  if (out_len <= (1ull << kPrefixMaxLogR))
    return prefix_fused(ctx, f, d_in, N, d_out, out_len, logN, omega, out_len, prefix_log_n1(logN, out_len),
                        (uint32_t)batch, stream);
  // longer prefixes: the full transform of every vector, its last pass storing k < out_len
  Schedule sch;
  if (!tuned_schedule(f, logN, sch)) default_schedule(f, logN, sch);
  XformOpts o;
  o.store_limit = out_len;
  return run_transform(ctx, f, d_in, d_out, logN, omega, sch, batch, N, out_len, o, stream);
}

}  // namespace omr
