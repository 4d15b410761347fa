// lde.hip -- coset low-degree extension and its decode (column batches).
//
// The reference has no encoder beyond the forward NTT (src/decode.rs:1 is a
// stub); these are the closed forms the engine defines:
//   encode: evals[i] = sum_{j<n} c[j] (h w_m^i)^j, i < m = n * blowup,
//           = NTT_m(c[j] h^j zero-padded): h^j is fused into the first pass
//           and the zero tail is never read from HBM;
//   decode: c[j] = h^-j m^-1 sum_i evals[i] w_m^(-ij) for j < n; the last
//           pass fuses h^-j m^-1, stores j < n and counts nonzero j >= n
//           (a codeword has none).
#include <algorithm>
#include <new>

#include "dispatch.h"
#include "internal.h"

namespace omr {
static int ilog2v(uint64_t n) { return 63 - __builtin_clzll(n); }

// Runs a column batch in groups of ~128 MiB of intermediates (group x
// 2^logM elements), so that the passes after the first read them back
// largely from the 256 MB Infinity Cache instead of HBM.  The groups are
// independent: they alternate between the caller's stream and a second one
// (forked from the caller's stream and joined back to it), so that one
// group's launches and pass tails overlap the other's; the second stream
// has its own intermediates.  256 x 2^20 BabyBear columns at blowup 4,
// MI355X: encode 5.99 ms on one stream in 96 MiB groups; on two streams
// 5.46 / 5.29 / 5.26 / 5.45 ms in groups of 48 / 96 / 128 / 192 MiB.
static omr_status run_columns(omr_ctx *ctx, int field, const void *in, void *out, uint32_t logM,
                              uint64_t root, const Schedule &s, uint64_t ncols, uint64_t in_stride,
                              uint64_t out_stride, const XformOpts &o, hipStream_t hs) {
  const int es = field_bytes(field);
  const uint64_t per_col = (uint64_t)es << logM;
  const uint64_t group = std::max<uint64_t>(1, (128ull << 20) / per_col);
  const bool fork = ncols > group;
  omr_status st = OMR_OK;
  if (fork) {
    if (!ctx->side && hipStreamCreateWithFlags(&ctx->side, hipStreamNonBlocking) != hipSuccess) {
      ctx->side = nullptr;
      return OMR_E_DEVICE;
    }
    if (!ctx->side_fork && hipEventCreateWithFlags(&ctx->side_fork, hipEventDisableTiming) != hipSuccess) {
      ctx->side_fork = nullptr;
      return OMR_E_DEVICE;
    }
    if (!ctx->side_join && hipEventCreateWithFlags(&ctx->side_join, hipEventDisableTiming) != hipSuccess) {
      ctx->side_join = nullptr;
      return OMR_E_DEVICE;
    }
    for (auto &b : ctx->side_scratch)
      if ((st = ensure_buf(b, group * per_col)) != OMR_OK) return st;
    // the second stream starts after the work already queued on hs
    if (hipEventRecord(ctx->side_fork, hs) != hipSuccess || hipStreamWaitEvent(ctx->side, ctx->side_fork, 0) != hipSuccess)
      return OMR_E_DEVICE;
  }
  uint64_t i = 0;
  for (uint64_t c0 = 0; c0 < ncols; c0 += group, i++) {
    const uint64_t g = std::min(group, ncols - c0);
    const bool side = fork && (i & 1);
    XformOpts og = o;
    if (side) {
      og.scratch[0] = ctx->side_scratch[0].p;
      og.scratch[1] = ctx->side_scratch[1].p;
    }
    st = run_transform(ctx, field, (const char *)in + c0 * in_stride * es, (char *)out + c0 * out_stride * es, logM,
                       root, s, g, in_stride, out_stride, og, side ? ctx->side : hs);
    if (st != OMR_OK) break;
  }
  if (fork) {  // hs continues after the second stream's groups, also after an error
    if (hipEventRecord(ctx->side_join, ctx->side) != hipSuccess || hipStreamWaitEvent(hs, ctx->side_join, 0) != hipSuccess)
      st = OMR_E_DEVICE;
  }
  return st;
}

omr_status lde_args(int field, uint64_t n, uint64_t ncols, uint32_t blowup, uint64_t &shift, uint32_t &logM) {
  if (!field_ok(field)) return OMR_E_FIELD;
  if (n == 0 || (n & (n - 1)) || blowup == 0 || (blowup & (blowup - 1))) return OMR_E_SIZE;
  if (ncols == 0 || ncols > 65535) return OMR_E_ARG;
  int two_adicity = 0;
  omr_field_info(field, nullptr, nullptr, &two_adicity);
  logM = ilog2v(n) + ilog2v(blowup);
  if ((int)logM > two_adicity) return OMR_E_TWO_ADICITY;
  if (shift == 0) shift = field_generator(field);
  if (shift >= field_modulus(field)) return OMR_E_ARG;
  return OMR_OK;
}
}  // namespace omr

using namespace omr;

extern "C" omr_status omr_coset_lde_device(omr_ctx *ctx, int field, const void *d_coeffs,
                                           void *d_evals, uint64_t n, uint64_t ncols,
                                           uint32_t blowup, uint64_t shift, void *stream) try {
  uint32_t logM;
  omr_status st = lde_args(field, n, ncols, blowup, shift, logM);
  if (st != OMR_OK) return st;
  if (!ctx || !d_coeffs || !d_evals) return OMR_E_NULL;
  if (d_coeffs == d_evals) return OMR_E_ARG;
  std::lock_guard<std::mutex> lk(ctx->mu);
  if (hipSetDevice(ctx->device) != hipSuccess) return OMR_E_DEVICE;
  StreamOrder ord(ctx, (hipStream_t)stream);
  uint64_t w;
  if ((st = omr_root_of_unity(field, 1ull << logM, &w)) != OMR_OK) return st;
  const TwTables *tw, *sh;
  if ((st = get_tables(ctx, field, logM, w, &tw)) != OMR_OK) return st;
  if ((st = get_tables(ctx, field, logM, shift, &sh)) != OMR_OK) return st;
  Schedule s;
  if (!tuned_schedule_for(field, logM, ncols, s)) default_schedule(field, logM, s);
  XformOpts o;
  o.zero_tail = true;
  o.log_valid = ilog2v(n);
  if (shift != 1) {
    o.pre_lo = sh->lo;
    o.pre_hi = sh->hi;
  }
  return run_columns(ctx, field, d_coeffs, d_evals, logM, w, s, ncols, n, 1ull << logM, o,
                     (hipStream_t)stream);
} catch (const std::bad_alloc &) {
  return OMR_E_ALLOC;
} catch (...) {
  return OMR_E_DEVICE;
}

extern "C" omr_status omr_coset_decode_device(omr_ctx *ctx, int field, const void *d_evals,
                                              void *d_coeffs, uint64_t n, uint64_t ncols,
                                              uint32_t blowup, uint64_t shift, uint32_t *d_nonzero,
                                              void *stream) try {
  uint32_t logM;
  omr_status st = lde_args(field, n, ncols, blowup, shift, logM);
  if (st != OMR_OK) return st;
  if (!ctx || !d_evals || !d_coeffs) return OMR_E_NULL;
  if (d_coeffs == d_evals) return OMR_E_ARG;
  std::lock_guard<std::mutex> lk(ctx->mu);
  if (hipSetDevice(ctx->device) != hipSuccess) return OMR_E_DEVICE;
  hipStream_t hs = (hipStream_t)stream;
  StreamOrder ord(ctx, hs);
  uint64_t w;
  if ((st = omr_root_of_unity(field, 1ull << logM, &w)) != OMR_OK) return st;
  const uint64_t M = 1ull << logM;
  const uint64_t winv = finv_host(field, w);
  const uint64_t sinv = finv_host(field, shift);
  const TwTables *tw, *sh;
  if ((st = get_tables(ctx, field, logM, winv, &tw)) != OMR_OK) return st;
  if ((st = get_tables(ctx, field, logM, sinv, &sh)) != OMR_OK) return st;
  if (d_nonzero && (st = launch_fill_u32(d_nonzero, 1, 0, hs)) != OMR_OK) return st;
  Schedule s;
  if (!tuned_schedule_for(field, logM, ncols, s)) default_schedule(field, logM, s);
  XformOpts o;
  o.scale = true;
  o.scale_w = to_w_host(field, finv_host(field, M % field_modulus(field)));
  if (sinv != 1) {
    o.post_lo = sh->lo;
    o.post_hi = sh->hi;
  }
  o.store_limit = n;
  o.nonzero_count = d_nonzero;
  return run_columns(ctx, field, d_evals, d_coeffs, logM, winv, s, ncols, M, n, o, hs);
} catch (const std::bad_alloc &) {
  return OMR_E_ALLOC;
} catch (...) {
  return OMR_E_DEVICE;
}
