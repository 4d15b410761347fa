// internal.h -- shared host-side types of the engine library.
#pragma once
#include <hip/hip_runtime.h>

#include <map>
#include <mutex>
#include <thread>
#include <tuple>
#include <vector>

#include "../../include/omrntt.h"

namespace omr {

// Vectors per call of the dense transform entry points (one pass launch
// covers them all; larger batches are split by the caller).
constexpr uint64_t kMaxBatch = 65535;

struct DevBuf {
  void *p = nullptr;
  size_t bytes = 0;
};

// Device twiddle tables of one root of order 2^logN (multiplier form).
struct TwTables {
  void *lo = nullptr;    // root^e,               e < 2^lo_bits
  void *hi = nullptr;    // root^(e 2^lo_bits),   e < 2^(logN - lo_bits)
  void *root = nullptr;  // root^(m 2^(logN - log_rmax)), m < 2^log_rmax
  // root pyramid: level L < log_rmax + 1 at [2^L, 2^(L+1)) holds w_(2^L)^m,
  // m < 2^L (w_(2^L) = root^(2^(logN - L))), so the roots of every radix are
  // one contiguous block
  void *pyr = nullptr;
  uint32_t lo_bits = 0;
  uint32_t log_rmax = 0;
  void *mid = nullptr;     // root^(j 2^(logN - log_mid)), j < 2^log_mid: single-lookup twiddles
  uint32_t log_mid = 0;
  uint32_t pow2_unit = 0;  // Goldilocks: 2^pow2_unit = root^(2^logN / M), M = 2^log_m = min(2^logN, 64)
  uint32_t log_m = 0;
  uint32_t pins = 0;  // holders beyond one call (batch plans): the cache never frees a pinned entry
  // register-DFT constant unit of the pass kernels (0 = table constants)
  int unit() const { return (log_m == 6 && (pow2_unit == 39 || pow2_unit == 153)) ? (int)pow2_unit : 0; }
};

// A transform schedule: pass radices 2^logR[i] (sum = logN), register radix
// 2^logE (0 = per-pass default).
struct Schedule {
  std::vector<int> logR;
  int logE = 0;
};

// Per-call options of run_transform.
struct XformOpts {
  bool scale = false;           // multiply every output by scale_w (last pass)
  uint64_t scale_w = 0;         // multiplier form
  bool zero_tail = false;       // input positions >= 2^log_valid read as zero (pass 0)
  uint32_t log_valid = 0;
  const void *pre_lo = nullptr;  // input-position multiplier tables (pass 0), or null
  const void *pre_hi = nullptr;
  uint64_t store_limit = ~0ull;  // only outputs with index < store_limit are stored
  const void *post_lo = nullptr;  // output-position multiplier tables (last pass), or null
  const void *post_hi = nullptr;
  uint32_t *nonzero_count = nullptr;  // counts nonzero outputs at index >= store_limit
  const uint64_t *off_in = nullptr;   // device table: vector v's input at element off_in[v] (not in_stride)
  const uint64_t *off_out = nullptr;  // device table: vector v's output at element off_out[v] (not out_stride)
  bool wide_in = false;   // 31/32-bit fields: input vectors are 64-bit words (value in the low half)
  bool wide_out = false;  // 31/32-bit fields: output vectors are written as 64-bit words
  void *scratch[2] = {nullptr, nullptr};  // caller-owned intermediates (batch x N elements each), else the context's
  const struct TwTables *tables = nullptr;  // caller-built twiddle tables of (logN, root), else the context's cache
};

// Pinned staging ring of the host-buffer entry points (hostio.hip): kSlots
// chunks of kChunk bytes, one copy stream and one event per slot.
struct HostStage {
  static constexpr size_t kChunk = 16ull << 20;
  static constexpr int kSlots = 4;
  std::mutex mu;
  void *ring = nullptr;
  hipStream_t stream = nullptr;
  hipEvent_t ev[kSlots] = {};
};

}  // namespace omr

struct omr_ctx {
  int device = 0;
  std::mutex mu;
  std::map<std::tuple<int, uint32_t, uint64_t>, omr::TwTables> tables;
  omr::DevBuf scratch[3];
  omr::DevBuf io[2];
  omr::DevBuf part, part2;
  omr::DevBuf counters;  // k_stride_fused's "last one done" counters, zeroed at allocation
  // second stream of the column batches (coset LDE / decode): created on
  // first use, forked from and joined back to the caller's stream within a
  // call, with its own intermediates
  hipStream_t side = nullptr;
  hipEvent_t side_fork = nullptr, side_join = nullptr;
  omr::DevBuf side_scratch[2];
  omr::DevBuf words[2];  // narrowed input / output of the word-width entry points
  // staging of the host-buffer forms (omr_intt, omr_ntt_batched,
  // omr_coset_lde, omr_coset_decode), under their own lock: they call the
  // device entry points, which take `mu`
  std::mutex host_mu;
  omr::DevBuf hio[3];
  omr::HostStage stage;  // its own lock, taken last
  // the stream of the last call that used the shared buffers above
  // (omr::StreamOrder)
  hipStream_t last_stream = nullptr;
  std::thread::id last_thread;  // for hipStreamPerThread, one stream per thread
  bool has_last = false;
};

namespace omr {

// The device entry points share the context's scratch and staging buffers,
// and the context lock orders only their host side: two calls on different
// streams could otherwise overlap on the GPU and overwrite each other's
// intermediates.  So a call on a stream other than the previous call's first
// waits until the device is idle (calls that stay on one stream cost a
// pointer compare; a stream switch costs one device synchronisation).  A
// stream under graph capture is left alone: a captured call is ordered by
// whoever replays the graph.  Construct under ctx->mu.
struct StreamOrder {
  StreamOrder(omr_ctx *ctx, hipStream_t s) {
    const std::thread::id me = std::this_thread::get_id();
    if (ctx->has_last && ctx->last_stream == s && (s != hipStreamPerThread || ctx->last_thread == me)) return;
    hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
    if (hipStreamIsCapturing(s, &cs) != hipSuccess) {
      (void)hipGetLastError();
      cs = hipStreamCaptureStatusNone;
    }
    if (cs != hipStreamCaptureStatusNone) return;
    if (ctx->has_last) (void)hipDeviceSynchronize();
    ctx->last_stream = s;
    ctx->last_thread = me;
    ctx->has_last = true;
  }
  StreamOrder(const StreamOrder &) = delete;
  StreamOrder &operator=(const StreamOrder &) = delete;
};

int field_bytes(int f);
bool field_ok(int f);
uint64_t fmul_host(int f, uint64_t a, uint64_t b);
uint64_t fpow_host(int f, uint64_t a, uint64_t e);
uint64_t finv_host(int f, uint64_t a);
uint64_t to_w_host(int f, uint64_t a);
uint64_t field_modulus(int f);
uint64_t field_generator(int f);
omr_status validate_domain(int f, uint64_t n, uint64_t omega);

omr_status ensure_buf(DevBuf &b, size_t bytes);
void free_buf(DevBuf &b);
omr_status get_tables(omr_ctx *ctx, int f, uint32_t logN, uint64_t root, const TwTables **out);
void pow2_unit_of(int f, uint32_t logN, uint64_t root, TwTables &t);
// No-allocation entry point (noalloc.hip): looks up this device's static
// scratch array once; called by omr_ctx_create with the device current.
void noalloc_init_device(int device);
// Pins (delta +1) or unpins (-1) a cached entry for a holder that keeps its
// device pointers past the call (batch plans bake them into device tables).
// Call under ctx->mu.
void pin_tables(omr_ctx *ctx, int f, uint32_t logN, uint64_t root, int delta);

void default_schedule(int f, uint32_t logN, Schedule &s);
omr_status variant_schedule(int f, int variant, uint32_t logN, Schedule &s);
// Autotuned schedule for (field, logN) if the tuning table has one.
bool tuned_schedule(int f, uint32_t logN, Schedule &s);
// The tuned schedule for `batch` vectors: batches of 32 MB or more use a slot
// of their own (false when it is empty: take the default schedule).
bool tuned_schedule_for(int f, uint32_t logN, uint64_t batch, Schedule &s);
bool batched_slot(int f, uint32_t logN, uint64_t batch);
void set_tuned_schedule(int f, uint32_t logN, const Schedule &s);
// Autotuned partial-output method for (field, logN, log out_len), if any.
bool tuned_partial(int f, uint32_t logN, uint32_t logOut, int &method, uint32_t &log_n1);

// Schedule-class vocabulary shared by the autotuner and the cost model.
constexpr int kLogMid = 16;  // mid twiddle table: 2^16 entries
constexpr int kFullClasses = 30;
constexpr int kNumClasses = 36;
bool class_schedule(int f, int cls, uint32_t logN, Schedule &s);
bool class_partial(int f, int cls, uint32_t logN, uint32_t logOut, int &method, uint32_t &log_n1);

omr_status run_transform(omr_ctx *ctx, int f, const void *d_in, void *d_out, uint32_t logN,
                         uint64_t root, const Schedule &sch, uint64_t batch, uint64_t in_stride,
                         uint64_t out_stride, const XformOpts &o, hipStream_t stream);
omr_status forward_full(omr_ctx *ctx, int f, int variant, const void *d_in, void *d_out,
                        uint32_t logN, uint64_t batch, uint64_t omega, hipStream_t stream, bool wide = false);
omr_status partial_impl(omr_ctx *ctx, int f, const void *d_in, void *d_out, uint32_t logN,
                        uint64_t omega, uint64_t out_len, int mode, int method, uint32_t log_n1,
                        hipStream_t stream);
// Argument checks of the coset LDE / decode entry points (shift 0 -> g).
// Host <-> device copies of the host-buffer entry points: pinned staging for
// large pageable buffers.  device_to_host first waits for the null stream.
omr_status host_to_device(omr_ctx *ctx, void *dev, const void *host, size_t bytes);
omr_status device_to_host(omr_ctx *ctx, void *host, const void *dev, size_t bytes);
void free_stage(HostStage &s);
uint32_t multipass_log_n1(uint32_t logN, uint64_t out_len);
bool decomposition_ok(int es, uint32_t logN, uint32_t log_n1, uint64_t out_len);
omr_status lde_args(int field, uint64_t n, uint64_t ncols, uint32_t blowup, uint64_t &shift, uint32_t &logM);
omr_status tune_classes_subset(omr_ctx *ctx, int field, uint64_t n, uint64_t out_len, uint64_t batch, int reps,
                               const bool *only, float *times_us);
omr_status partial_batch_impl(omr_ctx *ctx, int f, const void *d_in, void *d_out, uint32_t logN,
                              uint64_t batch, uint64_t omega, uint64_t out_len, hipStream_t stream);

// misc kernel launchers (misc_kernels.hip)
omr_status launch_scale(int f, const void *in, void *out, uint64_t n, uint64_t w, int apply,
                        hipStream_t s);
omr_status launch_dft_full(int f, const void *in, void *out, uint64_t n, uint32_t logN,
                           const TwTables *t, hipStream_t s);
omr_status launch_dft_partial(int f, const void *in, void *part, uint64_t n, uint32_t out_len,
                              uint32_t chunk, uint32_t nchunks, const TwTables *t, hipStream_t s);
omr_status launch_reduce_rows(int f, const void *part, void *out, uint32_t nrows, uint64_t ld,
                              uint64_t ncols, hipStream_t s);
omr_status launch_td_recombine(int f, const void *Y, void *part, uint64_t out_len, uint64_t n2,
                               uint32_t logN1, uint32_t splits, uint32_t logN, const TwTables *t,
                               hipStream_t s);
omr_status launch_fold(int f, const void *x, void *part, uint64_t n, uint64_t m, uint32_t nchunks,
                       hipStream_t s, uint32_t nvec = 1, uint64_t xs = 0, uint64_t ps = 0);
// One-launch stride subset (k_stride_fused) of m outputs of an N-element
// vector of es-byte words: grid workgroups of rows_per_wg rows, first-level
// groups of `group` workgroups.  stride_fused_plan returns false where the
// form does not apply (it needs 16 <= m <= 512, m < N and a row per row
// subset of a workgroup).  part holds grid x m words, part2 grid / group x m,
// cnt grid / group + 1 counters that are zero on entry and on exit.
struct StrideFusedPlan {
  uint32_t grid = 0, group = 0, rows_per_wg = 0, logm = 0;
};
bool stride_fused_plan(int es, uint64_t N, uint64_t m, StrideFusedPlan *plan);
omr_status launch_stride_fused(int f, const void *x, void *part, void *part2, uint32_t *cnt, void *out,
                               const void *root_m, const StrideFusedPlan &plan, hipStream_t s);
// out[v os + k] = sum_{t < n} part[(v m + k) n + t], k < m (<= 2^31), v < nvec (<= 65535)
omr_status launch_sum_contig(int f, const void *part, void *out, uint64_t n, uint64_t m, uint32_t nvec,
                             uint64_t os, hipStream_t s);
omr_status launch_fill_u32(uint32_t *p, uint64_t n, uint32_t v, hipStream_t s);
// NttDomain's twiddles (n words) and bitrev_twiddles (n/2 words) from the
// tables of (log n, w); either output may be null
omr_status launch_domain_tables(int f, void *tw, void *brev, uint64_t n, uint32_t logN, const TwTables *t,
                                hipStream_t s);
// 64-bit words <-> 32-bit values: out32[i] = low half of in64[i] / out64[i] = in32[i]
omr_status launch_narrow(const uint64_t *in, uint32_t *out, uint64_t n, hipStream_t s);
omr_status launch_widen(const uint32_t *in, uint64_t *out, uint64_t n, hipStream_t s);
omr_status launch_field_op(int f, int op, const void *a, const void *b, void *out, uint64_t n,
                           hipStream_t s);

}  // namespace omr
