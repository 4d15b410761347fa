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
  a.root_pyr = tab->pyr;
  a.lo_bits = tab->lo_bits;
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
    const uint64_t nch = std::min<uint64_t>((rows + 31) / 32, 65535);  // 32 rows a chunk: the fold kernel's unrolled case
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

// Pass radices of the sub-NTTs of size 2^log_n1 in the fused prefix: one
// pass up to 2^10 (the prefix epilogue's largest radix), else the fewest
// passes of at most 2^10, balanced, larger ones first.
static std::vector<int> prefix_split(uint32_t log_n1) {
  const int np = ((int)log_n1 + kPrefixMaxLogR - 1) / kPrefixMaxLogR;
  std::vector<int> lr(np, (int)log_n1 / np);
  for (int i = 0; i < (int)log_n1 % np; i++) lr[i]++;
  return lr;
}

// Whether the fused prefix runs for (N, n1): the last pass's tiles must not
// straddle two values of p (C <= s = N/n1).
static bool prefix_fused_ok(uint32_t logN, uint32_t log_n1) {
  if (log_n1 == 0 || log_n1 > logN) return false;
  const std::vector<int> lr = prefix_split(log_n1);
  if (lr.size() == 1) return (int)log_n1 <= kPrefixMaxLogR;
  return logN - log_n1 >= 2;  // s >= 4 columns per tile
}

// Whether the decomposition with sub-NTT size 2^log_n1 can run a prefix of
// out_len: fused (out_len <= n1, sub-NTTs in passes of up to 2^10), or one
// pass of radix n1 and a separate recombine.
bool decomposition_ok(int es, uint32_t logN, uint32_t log_n1, uint64_t out_len) {
  if (log_n1 == 0 || log_n1 > logN) return false;
  if (out_len <= (1ull << log_n1) && prefix_fused_ok(logN, log_n1)) return true;
  return (int)log_n1 <= max_log_radix(es);
}

// Prefix out[k] = X[k], k < out_len <= n1 = 2^log_n1, of nvec vectors (vector
// v at d_in + v in_stride, its prefix at d_out + v out_stride): the first
// passes of an N-point transform covering log n1 input digits (their
// outputs are the n1-point sub-NTTs of x[j :: N/n1]), the last of them with
// the prefix epilogue folding w_N^(j k) Z_j[k] over each tile's columns
// (kEpiPrefix), then each vector's per-tile partial sums added up.  With
// one pass the input is read once and nothing N-sized is written.
static omr_status prefix_fused(omr_ctx *ctx, int f, const void *d_in, uint64_t in_stride, void *d_out,
                               uint64_t out_stride, uint32_t logN, uint64_t omega, uint64_t out_len,
                               uint32_t log_n1, uint32_t nvec, hipStream_t stream) {
  const int es = field_bytes(f);
  const uint64_t N = 1ull << logN;
  const TwTables *tab;
  omr_status st = get_tables(ctx, f, logN, omega, &tab);
  if (st != OMR_OK) return st;
  if (log_n1 == 0 || log_n1 > logN) return OMR_E_SCHEDULE;
  const std::vector<int> lr = prefix_split(log_n1);
  const int np = (int)lr.size();
  const uint32_t logS_last = logN - log_n1;
  if (np >= 2) {
    if ((st = ensure_buf(ctx->scratch[0], (size_t)nvec * N * es)) != OMR_OK) return st;
    if (np >= 3 && (st = ensure_buf(ctx->scratch[1], (size_t)nvec * N * es)) != OMR_OK) return st;
  }
  const int logR_last = lr.back();
  const int logE_last = logR_last <= 2 ? logR_last : std::min(4, logR_last);
  const int logC_last = std::min<int>(pass_log_cols((int)logN, logR_last, logE_last, es), (int)logS_last);
  const uint64_t tps = (1ull << logS_last) >> logC_last;  // last-pass tiles per value of p
  if ((st = ensure_buf(ctx->scratch[2], (size_t)nvec * tps * out_len * es)) != OMR_OK) return st;
  const void *src = d_in;
  uint32_t consumed = 0;
  for (int i = 0; i < np; i++) {
    const int logR = lr[i];
    consumed += logR;
    const bool last = i == np - 1;
    const int logE = logR <= 2 ? logR : std::min(4, logR);
    const int logC = last ? logC_last : pass_log_cols((int)logN, logR, logE, es);
    PassArgs a{};
    a.src = src;
    a.dst = last ? ctx->scratch[2].p : ctx->scratch[i % 2].p;
    a.tw_lo = tab->lo;
    a.tw_hi = tab->hi;
    a.root_pyr = tab->pyr;
    a.lo_bits = tab->lo_bits;
    a.tw_mid = tab->mid;
    a.log_rmax = tab->log_rmax;
    a.log_mid = tab->log_mid;
    a.logN = logN;
    a.logS = logN - consumed;
    a.logC = logC;
    a.vec_stride_in = i == 0 ? in_stride : N;
    a.vec_stride_out = last ? 0 : N;
    a.store_limit = last ? out_len : ~0ull;
    a.flags = (i > 0 ? kTwiddle : 0u) | (last ? kPrefixSum : 0u);
    a.nvec = nvec;
    dim3 block((1u << logC) << (logR - logE));
    const uint64_t tiles = ((N >> logR) >> logC) * nvec;
    if (launch_pass(f, logR, logE, tab->unit(), a, tiles, block, stream) != hipSuccess) return OMR_E_DEVICE;
    src = a.dst;
  }
  // partials [vector][k][tile]: one block per (k, vector) sums a vector's tiles
  constexpr uint32_t kMaxY = 65535;
  for (uint32_t v0 = 0; v0 < nvec; v0 += kMaxY) {
    const uint32_t nv = std::min(kMaxY, nvec - v0);
    if ((st = launch_sum_contig(f, (const char *)ctx->scratch[2].p + (size_t)v0 * tps * out_len * es,
                                (char *)d_out + (size_t)v0 * out_stride * es, tps, out_len, nv, out_stride,
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

// Sub-NTT size of the multi-pass fused prefix for out_len > 2^10: four times
// out_len, at most N/64 (the last pass keeps 64 columns per value of p), or
// 0 where that is below out_len (the full transform with a pruned last pass
// is then faster: N = 2^22, out_len >= 2^18, measured).
uint32_t multipass_log_n1(uint32_t logN, uint64_t out_len) {
  uint32_t lo = 0;
  while ((1ull << lo) < out_len) lo++;
  if (logN < 6) return 0;
  const uint32_t l = std::min<uint32_t>(lo + 2, logN - 6);
  return (l >= lo && l > (uint32_t)kPrefixMaxLogR) ? l : 0;
}

static omr_status stride_fused(omr_ctx *ctx, int f, const void *d_in, void *d_out, uint32_t logN, uint64_t omega,
                               uint64_t m, const StrideFusedPlan &plan, hipStream_t stream) {
  const int es = field_bytes(f);
  const TwTables *tab;
  omr_status st = get_tables(ctx, f, logN, omega, &tab);
  if (st != OMR_OK) return st;
  if ((st = ensure_buf(ctx->part, (size_t)plan.grid * m * es)) != OMR_OK) return st;
  if ((st = ensure_buf(ctx->part2, (size_t)(plan.grid / plan.group) * m * es)) != OMR_OK) return st;
  if (!ctx->counters.p) {
    if ((st = ensure_buf(ctx->counters, 4096)) != OMR_OK) return st;
    if (hipMemset(ctx->counters.p, 0, 4096) != hipSuccess) return OMR_E_DEVICE;
  }
  const void *root_m = (const char *)tab->pyr + m * es;  // level log m of the root pyramid: w_m^k, k < m
  return launch_stride_fused(f, d_in, ctx->part.p, ctx->part2.p, (uint32_t *)ctx->counters.p, d_out, root_m, plan,
                             stream);
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
    // (16-byte loads: an input aligned to 16 bytes)
    StrideFusedPlan sfp;
    if ((uintptr_t)d_in % 16 == 0 && stride_fused_plan(es, N, m, &sfp))
      return stride_fused(ctx, f, d_in, d_out, logN, omega, m, sfp, stream);
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

  if (logN == 0)  // N = 1: the transform is the identity
    return hipMemcpyAsync(d_out, d_in, es, hipMemcpyDeviceToDevice, stream) == hipSuccess ? OMR_OK : OMR_E_DEVICE;
  if (method == 0 && (out_len & (out_len - 1)) == 0 &&
      tuned_partial(f, logN, ilog2u(out_len), method, log_n1)) {
    // autotuned choice
  } else if (method == 0) {
    if (out_len <= 4) {
      method = 1;
    } else if (out_len <= (1ull << kPrefixMaxLogR)) {
      method = 2;
      log_n1 = prefix_log_n1(logN, out_len);
    } else if ((log_n1 = multipass_log_n1(logN, out_len)) != 0) {
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
    // One chunk writes its sums straight to d_out, unless d_out overlaps the
    // input: every output reads all N inputs, so an in-place call must not
    // store out[k] before the sums of out[k + 1 ..] have read x[k].
    const char *ib = (const char *)d_in, *ob = (const char *)d_out;
    const bool overlap = ob < ib + N * es && ib < ob + out_len * es;
    if (nchunks == 1 && !overlap)
      return launch_dft_partial(f, d_in, d_out, N, (uint32_t)out_len, chunk, 1, tab, stream);
    if ((st = ensure_buf(ctx->part, (size_t)nchunks * out_len * es)) != OMR_OK) return st;
    if ((st = launch_dft_partial(f, d_in, ctx->part.p, N, (uint32_t)out_len, chunk, (uint32_t)nchunks,
                                 tab, stream)) != OMR_OK)
      return st;
    return launch_reduce_rows(f, ctx->part.p, d_out, (uint32_t)nchunks, out_len, out_len, stream);
  }
  if (method == 2) {
    if (log_n1 == 0) log_n1 = std::min<uint32_t>(logN, (uint32_t)maxlr);
    if (log_n1 == 0 || log_n1 > logN) return OMR_E_SCHEDULE;
    const uint64_t n1 = 1ull << log_n1, n2 = N >> log_n1;
    if (out_len <= n1 && prefix_fused_ok(logN, log_n1))  // recombine fused into the last pass
      return prefix_fused(ctx, f, d_in, N, d_out, out_len, logN, omega, out_len, log_n1, 1, stream);
    if ((int)log_n1 > maxlr) return OMR_E_SCHEDULE;  // one pass computes the sub-NTTs
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
  if (logN == 0)  // N = 1: every vector's transform is itself
    return hipMemcpyAsync(d_out, d_in, batch * field_bytes(f), hipMemcpyDeviceToDevice, stream) == hipSuccess
               ? OMR_OK
               : OMR_E_DEVICE;
  // Batches run in chunks of at most kMaxBatch vectors and 2^32 elements:
  // a pass launch has one workgroup per tile of every vector (at most
  // 2^28 work-items here, one per 16 elements), and the pruned full transform
  // takes at most kMaxBatch vectors.
  const uint64_t chunk = std::min<uint64_t>(kMaxBatch, std::max<uint64_t>(1, (1ull << 32) >> logN));
  if (batch > chunk) {
    const int es = field_bytes(f);
    for (uint64_t b0 = 0; b0 < batch; b0 += chunk) {
      const omr_status st = partial_batch_impl(ctx, f, (const char *)d_in + b0 * N * es,
                                               (char *)d_out + b0 * out_len * es, logN,
                                               std::min(chunk, batch - b0), omega, out_len, stream);
      if (st != OMR_OK) return st;
    }
    return OMR_OK;
  }
  if (out_len <= (1ull << kPrefixMaxLogR))
    return prefix_fused(ctx, f, d_in, N, d_out, out_len, logN, omega, out_len, prefix_log_n1(logN, out_len),
                        (uint32_t)batch, stream);
  const uint32_t l = multipass_log_n1(logN, out_len);
  if (l != 0)
    return prefix_fused(ctx, f, d_in, N, d_out, out_len, logN, omega, out_len, l, (uint32_t)batch, stream);
  // long prefixes: the full transform of every vector, its last pass storing k < out_len
  Schedule sch;
  if (!tuned_schedule_for(f, logN, batch, sch)) default_schedule(f, logN, sch);
  XformOpts o;
  o.store_limit = out_len;
  return run_transform(ctx, f, d_in, d_out, logN, omega, sch, batch, N, out_len, o, stream);
}

}  // namespace omr
