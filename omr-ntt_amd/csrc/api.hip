// api.hip -- C ABI of the NTT engine: validation, plans, schedules and launches.
//
// Host code only; the kernels live in pass_<field>.hip (Stockham passes) and
// misc_kernels.hip.  See include/omrntt.h for the contract.
#include <hip/hip_runtime.h>
#include <new>

#include <dlfcn.h>

#include <algorithm>
#include <cstdlib>
#include <string>
#include <cstring>
#include <map>
#include <mutex>
#include <tuple>
#include <vector>

#include "../../include/omrntt.h"
#include "dispatch.h"
#include "internal.h"

using namespace omr;

// ---------------------------------------------------------------------------
// status strings, field facts
// ---------------------------------------------------------------------------
extern "C" const char *omr_status_str(omr_status s) {
  switch (s) {
    case OMR_OK: return "ok";
    case OMR_E_NULL: return "null pointer argument";
    case OMR_E_SIZE: return "NTT size N must be a power of 2";
    case OMR_E_TWO_ADICITY: return "N requires log2(N) <= TWO_ADICITY (N must divide p-1)";
    case OMR_E_FIELD: return "unknown field";
    case OMR_E_VARIANT: return "unknown encoder variant";
    case OMR_E_RADIX: return "variant constraint violated (radix-4 needs log2 N even, radix-8 needs log2 N divisible by 3)";
    case OMR_E_OMEGA: return "omega is not a primitive N-th root of unity";
    case OMR_E_OUTLEN: return "invalid partial output length";
    case OMR_E_ALLOC: return "allocation failed";
    case OMR_E_DEVICE: return "HIP device error";
    case OMR_E_SCHEDULE: return "invalid schedule";
    case OMR_E_ARG: return "invalid argument";
    case OMR_E_SCRATCH: return "request exceeds the no-allocation scratch (n * batch * element bytes > OMR_NOALLOC_MAX_BYTES)";
  }
  return "unknown status";
}

extern "C" const char *omr_version(void) { return "omrntt-gfx950 0.1.0"; }

namespace omr {

static const uint64_t kP[4] = {BabyBear::P, KoalaBear::P, Goldilocks::P, TeddyBear::P};
static const uint64_t kG[4] = {31, 3, 7, 5};
static const int kTwoAdicity[4] = {27, 24, 32, 30};

int field_bytes(int f) { return f == kGoldilocks ? 8 : 4; }
bool field_ok(int f) { return f >= 0 && f <= 3; }

uint64_t fmul_host(int f, uint64_t a, uint64_t b) {
  return (uint64_t)(((unsigned __int128)a * b) % kP[f]);
}
uint64_t fpow_host(int f, uint64_t a, uint64_t e) {
  uint64_t r = 1;
  a %= kP[f];
  while (e) {
    if (e & 1) r = fmul_host(f, r, a);
    a = fmul_host(f, a, a);
    e >>= 1;
  }
  return r;
}
uint64_t finv_host(int f, uint64_t a) { return fpow_host(f, a, kP[f] - 2); }
uint64_t to_w_host(int f, uint64_t a) {
  // multiplier form: Montgomery (a * 2^32 mod p) for 32-bit fields, a for Goldilocks
  if (f == kGoldilocks) return a;
  return (uint64_t)((((unsigned __int128)a) << 32) % kP[f]);
}
uint64_t field_modulus(int f) { return kP[f]; }
uint64_t field_generator(int f) { return kG[f]; }

static int ilog2(uint64_t n) { return 63 - __builtin_clzll(n); }

// NttDomain::new checks (src/encoder.rs:24-32) plus a primitive-root check of
// the caller's omega.
omr_status validate_domain(int f, uint64_t n, uint64_t omega) {
  if (!field_ok(f)) return OMR_E_FIELD;
  if (n == 0 || (n & (n - 1))) return OMR_E_SIZE;
  if (ilog2(n) > kTwoAdicity[f]) return OMR_E_TWO_ADICITY;
  if (omega >= kP[f]) return OMR_E_OMEGA;
  // the primitivity check costs two modular powers: the last few domains
  // that passed it are remembered per thread
  struct Seen {
    int f;
    uint64_t n, omega;
  };
  thread_local Seen seen[8] = {};
  thread_local unsigned next = 0;
  for (const Seen &d : seen)
    if (d.n == n && d.omega == omega && d.f == f) return OMR_OK;
  if (fpow_host(f, omega, n) != 1) return OMR_E_OMEGA;
  if (n > 1 && fpow_host(f, omega, n / 2) == 1) return OMR_E_OMEGA;
  seen[next++ % 8] = {f, n, omega};
  return OMR_OK;
}

// ---------------------------------------------------------------------------
// device buffers and the twiddle-table cache
// ---------------------------------------------------------------------------
static bool alloc_fail_injected() {
  const char *e = getenv("OMR_FAIL_ALLOC");
  return e && e[0] == '1';
}

omr_status ensure_buf(DevBuf &b, size_t bytes) {
  if (bytes <= b.bytes) return OMR_OK;
  if (alloc_fail_injected()) return OMR_E_ALLOC;
  if (b.p) {
    if (hipDeviceSynchronize() != hipSuccess) return OMR_E_DEVICE;
    (void)hipFree(b.p);
    b.p = nullptr;
    b.bytes = 0;
  }
  size_t want = bytes < 4096 ? 4096 : bytes;
  if (hipMalloc(&b.p, want) != hipSuccess) {
    (void)hipGetLastError();
    b.p = nullptr;
    return OMR_E_ALLOC;
  }
  b.bytes = want;
  return OMR_OK;
}

void free_buf(DevBuf &b) {
  if (b.p) (void)hipFree(b.p);
  b.p = nullptr;
  b.bytes = 0;
}

// Goldilocks: the discrete log base 2 of the root of order min(N, 64) (the
// register-DFT constants of the pass kernels are then compile-time shifts);
// other fields, and values that are not roots of unity (coset shifts): none.
void pow2_unit_of(int f, uint32_t logN, uint64_t root, TwTables &t) {
  t.log_m = 0;
  t.pow2_unit = 0;
  if (f != kGoldilocks) return;
  const uint32_t log_m = std::min<uint32_t>(logN, 6);
  const uint64_t wm = fpow_host(f, root, 1ull << (logN - log_m));
  uint64_t pw = 1;
  for (uint32_t k = 0; k < 192; k++, pw = fmul_host(f, pw, 2))
    if (pw == wm) {
      t.pow2_unit = k;
      t.log_m = log_m;
      return;
    }
}

// Two-level power table of `root` (order 2^logN) plus the radix root table.
void pin_tables(omr_ctx *ctx, int f, uint32_t logN, uint64_t root, int delta) {
  auto it = ctx->tables.find(std::make_tuple(f, logN, root));
  if (it == ctx->tables.end()) return;
  if (delta > 0) it->second.pins++;
  else if (it->second.pins) it->second.pins--;
}

static void free_tables(TwTables &t) {
  for (void **p : {&t.lo, &t.hi, &t.root, &t.pyr, &t.mid})
    if (*p) (void)hipFree(*p), *p = nullptr;
}

omr_status get_tables(omr_ctx *ctx, int f, uint32_t logN, uint64_t root, const TwTables **out) {
  auto key = std::make_tuple(f, logN, root);
  auto it = ctx->tables.find(key);
  if (it != ctx->tables.end()) {
    *out = &it->second;
    return OMR_OK;
  }
  if (ctx->tables.size() >= 128) {  // bounded cache: drop every entry no plan holds
    if (hipDeviceSynchronize() != hipSuccess) return OMR_E_DEVICE;
    for (auto it2 = ctx->tables.begin(); it2 != ctx->tables.end();) {
      if (it2->second.pins) {
        ++it2;
        continue;
      }
      free_tables(it2->second);
      it2 = ctx->tables.erase(it2);
    }
  }
  if (alloc_fail_injected()) return OMR_E_ALLOC;
  const int es = field_bytes(f);
  TwTables t{};
  t.lo_bits = std::min<uint32_t>(logN, 12);
  const uint64_t nlo = 1ull << t.lo_bits, nhi = 1ull << (logN - t.lo_bits);
  // root table: root^(m 2^(logN - log_rmax)), m < 2^log_rmax (the largest pass
  // radix): in-tile roots of every pass radix, and single-lookup inter-pass
  // twiddles for passes with N/s <= 2^log_rmax.  Kept compact: strided
  // lookups into a larger table miss the L1 cache.
  t.log_rmax = std::min<uint32_t>(logN, (uint32_t)max_log_radix(es));
  const uint64_t nroot = 1ull << t.log_rmax;
  std::vector<uint64_t> lo(nlo), hi(nhi), rt(nroot);
  uint64_t cur = 1;
  for (uint64_t i = 0; i < nlo; i++) {
    lo[i] = to_w_host(f, cur);
    cur = fmul_host(f, cur, root);
  }
  const uint64_t step_hi = cur;  // root^(2^lo_bits)
  cur = 1;
  for (uint64_t i = 0; i < nhi; i++) {
    hi[i] = to_w_host(f, cur);
    cur = fmul_host(f, cur, step_hi);
  }
  const uint64_t step_rt = fpow_host(f, root, 1ull << (logN - t.log_rmax));
  cur = 1;
  for (uint64_t i = 0; i < nroot; i++) {
    rt[i] = to_w_host(f, cur);
    cur = fmul_host(f, cur, step_rt);
  }
  std::vector<uint64_t> pyr(2 * nroot);  // level L: w_(2^L)^m = rt[m 2^(log_rmax - L)]
  pyr[0] = to_w_host(f, 1);
  for (uint32_t L = 0; L <= t.log_rmax; L++)
    for (uint64_t m = 0; m < (1ull << L); m++) pyr[(1ull << L) + m] = rt[m << (t.log_rmax - L)];
  auto upload = [&](std::vector<uint64_t> &v, void **dp) -> omr_status {
    std::vector<uint32_t> v32;
    const void *src = v.data();
    if (es == 4) {
      v32.resize(v.size());
      for (size_t i = 0; i < v.size(); i++) v32[i] = (uint32_t)v[i];
      src = v32.data();
    }
    if (hipMalloc(dp, v.size() * es) != hipSuccess) {
      (void)hipGetLastError();
      return OMR_E_ALLOC;
    }
    if (hipMemcpy(*dp, src, v.size() * es, hipMemcpyHostToDevice) != hipSuccess) return OMR_E_DEVICE;
    return OMR_OK;
  };
  pow2_unit_of(f, logN, root, t);
  omr_status st = OMR_OK;
  t.log_mid = std::min<uint32_t>(logN, kLogMid);
  if (t.log_mid > t.log_rmax) {  // mid table root^(j N/2^log_mid), j < 2^log_mid
    std::vector<uint64_t> mid(1ull << t.log_mid);
    const uint64_t step = fpow_host(f, root, 1ull << (logN - t.log_mid));
    cur = 1;
    for (auto &v : mid) {
      v = to_w_host(f, cur);
      cur = fmul_host(f, cur, step);
    }
    st = upload(mid, &t.mid);
  } else {
    t.log_mid = t.log_rmax;
  }
  if (st == OMR_OK) st = upload(lo, &t.lo);
  if (st == OMR_OK) st = upload(hi, &t.hi);
  if (st == OMR_OK) st = upload(rt, &t.root);
  if (st == OMR_OK) st = upload(pyr, &t.pyr);
  if (st != OMR_OK) {  // nothing of a partial set is kept
    free_tables(t);
    return st;
  }
  auto res = ctx->tables.emplace(key, t);
  *out = &res.first->second;
  return OMR_OK;
}

// ---------------------------------------------------------------------------
// schedules
// ---------------------------------------------------------------------------
// Largest pass radix whose tile still has >= 16 columns (64-128 B runs).
static int coalesced_max_log_radix(int es) {
  return ilog2((uint64_t)(kLdsBudget / es / 16));
}

// Split `total` radix-2 layers into `parts` balanced pieces (larger first).
static void balanced(int total, int parts, std::vector<int> &out) {
  int base = total / parts, rem = total % parts;
  for (int i = 0; i < parts; i++) out.push_back(base + (i < rem ? 1 : 0));
}

void default_schedule(int f, uint32_t logN, Schedule &s) {
  s.logR.clear();
  s.logE = 0;
  if (logN == 0) return;
  const int es = field_bytes(f);
  if ((int)logN <= max_log_radix(es)) {
    s.logR.push_back(logN);
    return;
  }
  const int cm = coalesced_max_log_radix(es);
  const int parts = (logN + cm - 1) / cm;
  balanced(logN, parts, s.logR);
}

// Split `total` layers into pieces no larger than `cap`, each a multiple of `unit`.
static bool split_units(int total, int cap, int unit, std::vector<int> &out) {
  if (total % unit) return false;
  int units = total / unit, capu = std::max(1, cap / unit);
  int parts = (units + capu - 1) / capu;
  int base = units / parts, rem = units % parts;
  for (int i = 0; i < parts; i++) out.push_back((base + (i < rem ? 1 : 0)) * unit);
  return true;
}

omr_status variant_schedule(int f, int variant, uint32_t logN, Schedule &s) {
  const int es = field_bytes(f);
  const int cap = max_log_radix(es);
  const int ccap = coalesced_max_log_radix(es);
  s.logR.clear();
  s.logE = 0;
  if (logN == 0) return OMR_OK;
  switch (variant) {
    case OMR_ARK_RADIX2:
      default_schedule(f, logN, s);
      return OMR_OK;
    case OMR_LAMBDA_BOWERS:  // two fused radix-2 layers per register step
      default_schedule(f, logN, s);
      s.logE = 2;
      for (int &r : s.logR)
        if (r < 2) s.logE = 0;
      return OMR_OK;
    case OMR_LAMBDA_RADIX4:  // radix-4 digits only (lambdaworks_radix4.rs:17-21)
      if (logN % 2) return OMR_E_RADIX;
      split_units(logN, (int)logN <= cap ? cap : ccap, 2, s.logR);
      s.logE = 2;
      return OMR_OK;
    case OMR_TFHE_STOCKHAM_RADIX8:  // radix-8 digits only (tfhe_stockham_radix8.rs:17-21)
      if (logN % 3) return OMR_E_RADIX;
      split_units(logN, (int)logN <= cap ? cap : ccap + 1, 3, s.logR);
      s.logE = 3;
      return OMR_OK;
    case OMR_WINTERFELL_FOUR_STEP: {  // inner = 2^floor(logN/2) first (winterfell_four_step.rs:40-44)
      int inner = logN / 2, outer = logN - inner;
      if (inner) split_units(inner, cap, 1, s.logR);
      split_units(outer, cap, 1, s.logR);
      return OMR_OK;
    }
    case OMR_PLONKY3_LAYER_SPLIT: {  // mid = ceil(logN/2) (plonky3_radix2_layer_split.rs:33)
      int mid = (logN + 1) / 2, rest = logN - mid;
      split_units(mid, cap, 1, s.logR);
      if (rest) split_units(rest, cap, 1, s.logR);
      return OMR_OK;
    }
    default:
      return OMR_E_VARIANT;
  }
}

static int auto_log_e(int logR) {
  if (logR <= 2) return logR;
  return std::min(4, logR);
}

// ---------------------------------------------------------------------------
// the multi-pass transform
// ---------------------------------------------------------------------------
omr_status run_transform(omr_ctx *ctx, int f, const void *d_in, void *d_out, uint32_t logN,
                         uint64_t root, const Schedule &sch, uint64_t batch, uint64_t in_stride,
                         uint64_t out_stride, const XformOpts &o, hipStream_t stream) {
  const int es = field_bytes(f);
  const uint64_t N = 1ull << logN;
  if (batch == 0) return OMR_OK;
  if (batch > kMaxBatch) return OMR_E_ARG;
  if (logN == 0) {  // size-1 transform: identity (times the scale)
    const TwTables *tt;
    (void)tt;
    for (uint64_t b = 0; b < batch; b++) {
      omr_status st = launch_scale(f, (const char *)d_in + b * in_stride * es,
                                   (char *)d_out + b * out_stride * es, 1, o.scale_w,
                                   o.scale ? 1 : 0, stream);
      if (st != OMR_OK) return st;
    }
    return OMR_OK;
  }
  int sum = 0;
  for (int r : sch.logR) {
    if (r < 1 || r > max_log_radix(es)) return OMR_E_SCHEDULE;
    sum += r;
  }
  if (sum != (int)logN || sch.logR.empty() || sch.logR.size() > 8) return OMR_E_SCHEDULE;
  const int np = (int)sch.logR.size();

  const TwTables *tab = o.tables;
  omr_status st = OMR_OK;
  if (!tab && (st = get_tables(ctx, f, logN, root, &tab)) != OMR_OK) return st;

  // ping-pong buffers X_0 = in, X_np = out, X_1..X_np-1 from scratch
  std::vector<const void *> X(np + 1);
  X[0] = d_in;
  X[np] = d_out;
  const size_t vec_bytes = (size_t)N * es * batch;
  // d_out doubles as a ping-pong buffer only when it holds whole dense vectors
  const bool use_out = !o.off_in && !o.off_out && (o.store_limit >= N) && ((d_in != d_out) || (np % 2 == 0));
  void *S[2] = {o.scratch[0], o.scratch[1]};  // caller-owned, else the context's
  auto need = [&](int k) -> omr_status {
    if (S[k]) return OMR_OK;
    const omr_status s2 = ensure_buf(ctx->scratch[k], vec_bytes);
    if (s2 == OMR_OK) S[k] = ctx->scratch[k].p;
    return s2;
  };
  if (np >= 2 && (st = need(0)) != OMR_OK) return st;
  if (np >= 3 && !use_out && (st = need(1)) != OMR_OK) return st;
  for (int i = 1; i < np; i++) {
    if (use_out) X[i] = ((np - i) % 2 == 0) ? d_out : S[0];
    else X[i] = ((np - i) % 2 == 1) ? S[0] : S[1];
  }
  // with use_out and an out-of-place call whose out vectors are strided, the
  // intermediates must be dense N-element vectors
  if ((use_out && np >= 2 && out_stride != N) || (!use_out && np >= 2)) {
    if ((st = need(1)) != OMR_OK) return st;
    for (int i = 1; i < np; i++) X[i] = ((np - i) % 2 == 1) ? S[0] : S[1];
  }

  // post[i]: pass i applies pass i+1's inter-pass twiddle to its outputs
  // (kPostTwiddle) and pass i+1 runs without one.  Goldilocks vectors of
  // 2^22 and more: the later passes are the VALU-bound ones (the first pass
  // has no twiddle and the most memory stalls), so the chain products move
  // one pass earlier, where they are cheaper (A/B on one box, forward /
  // inverse: 2^22 [8,7,7] 67.6 / 66.0 -> 63.8 / 62.0 us, 2^24 [8,8,8]
  // 195.7 / 195.9 -> 193.3 / 193.4 us); at 2^20 and below the extra table
  // lookups cost more than that (2^20 30.4 -> 30.8 us, one launch sequence
  // per vector of the uneven batch 56.7 -> 58.6 ms).  Not on a first pass
  // with an LDE's zero tail or coset multiplier (the general epilogue), nor
  // on radix-4096 passes (the four-step kernel, which has no such
  // epilogue).  An inverse's N^-1 rides in the last such pass's chain.
  // (Moving only the first pass's twiddle into pass 0: 2^22 66.0 us, between
  // the two.)  OMR_POSTTW=0 in the environment turns it off.
  static const bool post_on = [] {
    const char *e = getenv("OMR_POSTTW");
    return !(e && e[0] == '0' && e[1] == 0);
  }();
  std::vector<char> post(np, 0);
  if (post_on && logN >= 22)
    for (int i = 0; i + 1 < np; i++) {
      bool ok = sch.logR[i] < 12 && (i > 0 || (!o.zero_tail && !o.pre_lo && !o.wide_in));
      if (es == 4) {
        int le = sch.logE;
        if (!has_pass_kernel(sch.logR[i], le, es)) le = auto_log_e(sch.logR[i]);
        ok = ok && post4_ok(sch.logR[i], le, pass_log_cols((int)logN, sch.logR[i], le, es));
      }
      post[i] = ok;
    }
  int last_post = -1;
  for (int i = 0; i + 1 < np; i++)
    if (post[i]) last_post = i;
  // the scale rides in the last post pass only when the last pass has no
  // chain of its own to carry it
  const bool scale_in_post = o.scale && last_post >= 0 && last_post == np - 2;

  int consumed = 0;
  for (int i = 0; i < np; i++) {
    const int logR = sch.logR[i];
    consumed += logR;
    int logE = sch.logE;
    if (!has_pass_kernel(logR, logE, es)) logE = auto_log_e(logR);
    if (!has_pass_kernel(logR, logE, es)) return OMR_E_SCHEDULE;
    // Small jobs (batch x N words) take narrower tiles than the coalescing
    // rule gives, so that their few tiles spread over more CUs: a 2^16-word
    // vector has 16 tiles of radix 256 at full width.  Kernel time per
    // transform, MI355X: BabyBear 2^16 [8,8] 10.7 -> 8.2 us at 4 columns;
    // Goldilocks 2^16 13.7 -> 12.3 us at 8, 2^18 16.5 -> 14.8 us at 16;
    // from 2^18 (BabyBear) / 2^20 (Goldilocks) on, full width is faster.
    int logC = pass_log_cols((int)logN, logR, logE, es);
    const uint64_t job = (uint64_t)batch << logN;
    if (job <= (1ull << 16)) logC = std::min(logC, es == 4 ? 2 : 3);
    else if (es == 8 && job <= (1ull << 18)) logC = std::min(logC, 4);
    PassArgs a{};
    a.src = X[i];
    a.dst = (void *)X[i + 1];
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
    a.vec_stride_in = (i == 0) ? in_stride : N;
    a.vec_stride_out = (i == np - 1) ? out_stride : N;
    a.vec_off_in = (i == 0) ? o.off_in : nullptr;
    a.vec_off_out = (i == np - 1) ? o.off_out : nullptr;
    a.store_limit = (i == np - 1) ? o.store_limit : ~0ull;
    a.flags = 0;
    if (i > 0 && !post[i - 1]) a.flags |= kTwiddle;
    if (post[i]) {
      a.flags |= kPostTwiddle;
      a.post_logR = (uint32_t)sch.logR[i + 1];
      if (scale_in_post && i == last_post) {
        a.flags |= kScale;
        a.scale_w = o.scale_w;
      }
    }
    if (i == np - 1) {
      // a last pass with an inter-pass twiddle takes the scale in the start
      // value of its twiddle chains (one product per thread), otherwise
      // every output is scaled
      if (o.scale && !scale_in_post) {
        a.flags |= kScale;
        a.scale_w = o.scale_w;
      }
      a.post_lo = o.post_lo;
      a.post_hi = o.post_hi;
      a.nonzero_count = o.nonzero_count;
    }
    if (i == 0 && o.wide_in) a.flags |= kWideIn;
    if (i == np - 1 && o.wide_out) a.flags |= kWideOut;
    if (i == 0) {
      if (o.zero_tail) {
        a.flags |= kZeroTail;
        a.logValid = o.log_valid;
      }
      a.pre_lo = o.pre_lo;
      a.pre_hi = o.pre_hi;
    }
    a.nvec = (uint32_t)batch;
    dim3 block((1u << logC) << (logR - logE));
    const uint64_t tiles = ((N >> logR) >> logC) * batch;
    if (launch_pass(f, logR, logE, tab->unit(), a, tiles, block, stream) != hipSuccess) return OMR_E_DEVICE;
  }
  return OMR_OK;
}

}  // namespace omr

// ---------------------------------------------------------------------------
// C ABI
// ---------------------------------------------------------------------------
extern "C" omr_status omr_tune_load(const char *path);

// The first context loads the shipped tuning table: $OMR_TUNE_FILE, else
// tuning/mi355x.tune beside the library (written by the exhaustive autotuner,
// omr_ntt.tune_db).  OMR_TUNE_FILE=none disables it.
static void load_default_tuning() {
  static std::once_flag once;
  std::call_once(once, [] {
    const char *env = getenv("OMR_TUNE_FILE");
    if (env) {
      if (std::string(env) != "none") (void)omr_tune_load(env);
      return;
    }
    Dl_info info;
    if (dladdr((void *)&omr_tune_load, &info) && info.dli_fname) {
      std::string p(info.dli_fname);
      const size_t slash = p.rfind('/');
      p = (slash == std::string::npos ? std::string(".") : p.substr(0, slash)) + "/tuning/mi355x.tune";
      (void)omr_tune_load(p.c_str());
    }
  });
}

extern "C" omr_status omr_ctx_create(int device, omr_ctx **out) try {
  if (!out) return OMR_E_NULL;
  load_default_tuning();
  *out = nullptr;
  int count = 0;
  if (hipGetDeviceCount(&count) != hipSuccess || device < 0 || device >= count) {
    (void)hipGetLastError();
    return OMR_E_DEVICE;
  }
  if (hipSetDevice(device) != hipSuccess) return OMR_E_DEVICE;
  noalloc_init_device(device);  // once per device, with the device current
  omr_ctx *c = new (std::nothrow) omr_ctx();
  if (!c) return OMR_E_ALLOC;
  c->device = device;
  *out = c;
  return OMR_OK;
} catch (const std::bad_alloc &) {
  return OMR_E_ALLOC;
} catch (...) {
  return OMR_E_DEVICE;
}

extern "C" omr_status omr_ctx_destroy(omr_ctx *ctx) try {
  if (!ctx) return OMR_E_NULL;
  (void)hipSetDevice(ctx->device);
  (void)hipDeviceSynchronize();
  for (auto &kv : ctx->tables) free_tables(kv.second);
  for (auto &b : ctx->scratch) free_buf(b);
  free_buf(ctx->io[0]);
  free_buf(ctx->io[1]);
  free_buf(ctx->part);
  free_buf(ctx->part2);
  free_buf(ctx->counters);
  for (auto &b : ctx->side_scratch) free_buf(b);
  if (ctx->side_fork) (void)hipEventDestroy(ctx->side_fork);
  if (ctx->side_join) (void)hipEventDestroy(ctx->side_join);
  if (ctx->side) (void)hipStreamDestroy(ctx->side);
  for (auto &b : ctx->hio) free_buf(b);
  for (auto &b : ctx->words) free_buf(b);
  free_stage(ctx->stage);
  delete ctx;
  return OMR_OK;
} catch (const std::bad_alloc &) {
  return OMR_E_ALLOC;
} catch (...) {
  return OMR_E_DEVICE;
}

extern "C" omr_status omr_ctx_synchronize(omr_ctx *ctx, void *stream) try {
  if (!ctx) return OMR_E_NULL;
  return hipStreamSynchronize((hipStream_t)stream) == hipSuccess ? OMR_OK : OMR_E_DEVICE;
} catch (const std::bad_alloc &) {
  return OMR_E_ALLOC;
} catch (...) {
  return OMR_E_DEVICE;
}

extern "C" omr_status omr_field_info(int field, uint64_t *modulus, uint64_t *generator,
                                     int *two_adicity) try {
  if (!field_ok(field)) return OMR_E_FIELD;
  if (modulus) *modulus = kP[field];
  if (generator) *generator = kG[field];
  if (two_adicity) *two_adicity = kTwoAdicity[field];
  return OMR_OK;
} catch (const std::bad_alloc &) {
  return OMR_E_ALLOC;
} catch (...) {
  return OMR_E_DEVICE;
}

extern "C" int omr_field_bytes(int field) { return field_ok(field) ? field_bytes(field) : 0; }

extern "C" omr_status omr_root_of_unity(int field, uint64_t n, uint64_t *omega) try {
  if (!omega) return OMR_E_NULL;
  if (!field_ok(field)) return OMR_E_FIELD;
  if (n == 0 || (n & (n - 1))) return OMR_E_SIZE;
  if (ilog2(n) > kTwoAdicity[field]) return OMR_E_TWO_ADICITY;
  *omega = fpow_host(field, kG[field], (kP[field] - 1) / n);
  return OMR_OK;
} catch (const std::bad_alloc &) {
  return OMR_E_ALLOC;
} catch (...) {
  return OMR_E_DEVICE;
}

static const char *kVariantNames[OMR_VARIANT_COUNT] = {
    "ArkRadix2",          "Dft",
    "LambdaBowers",       "LambdaRadix4",
    "Plonky3Radix2LayerSplit", "WinterfellFourStep",
    "TfheStockhamRadix8", "DftPartial",
    "TransformDecomposition", "TransformDecompositionV2"};

extern "C" const char *omr_variant_name(int variant) {
  if (variant < 0 || variant >= OMR_VARIANT_COUNT) return nullptr;
  return kVariantNames[variant];
}

extern "C" int omr_variant_supports(int variant, uint32_t log_n) {
  if (variant == OMR_LAMBDA_RADIX4) return log_n % 2 == 0;
  if (variant == OMR_TFHE_STOCKHAM_RADIX8) return log_n % 3 == 0;
  return variant >= 0 && variant < OMR_VARIANT_COUNT;
}

extern "C" omr_status omr_default_schedule(int field, uint64_t n, int *log_radix, int *npass) try {
  if (!log_radix || !npass) return OMR_E_NULL;
  if (!field_ok(field)) return OMR_E_FIELD;
  if (n == 0 || (n & (n - 1))) return OMR_E_SIZE;
  if (ilog2(n) > kTwoAdicity[field]) return OMR_E_TWO_ADICITY;
  Schedule s;
  const uint32_t logN = ilog2(n);
  if (!tuned_schedule(field, logN, s)) default_schedule(field, logN, s);
  *npass = (int)s.logR.size();
  for (size_t i = 0; i < s.logR.size(); i++) log_radix[i] = s.logR[i];
  return OMR_OK;
} catch (const std::bad_alloc &) {
  return OMR_E_ALLOC;
} catch (...) {
  return OMR_E_DEVICE;
}

static bool is_partial_variant(int v) {
  return v == OMR_DFT_PARTIAL || v == OMR_TRANSFORM_DECOMPOSITION || v == OMR_TRANSFORM_DECOMPOSITION_V2;
}

static uint64_t next_pow2(uint64_t x) {
  uint64_t r = 1;
  while (r < x) r <<= 1;
  return r;
}

// The partial encoders' output count min(2s, N) (transform_decomposition.rs:30,
// dft.rs:43), without wrapping 2s for s >= 2^63.
static uint64_t partial_out_len(uint64_t s, uint64_t n) { return s > n / 2 ? n : std::min<uint64_t>(2 * s, n); }

// TransformDecomposition(V2)'s sub-NTT size, one rule for the host drop-in
// (omr_ntt) and the device form (omr_ntt_device): the reference's
// n1 = nextpow2(2 out_len) clamped to [2, N] (transform_decomposition.rs:38-39,
// transform_decomposition_v2.rs:36); where the sub-NTTs cannot run as fused
// passes, n1 is capped at the largest pass radix.  Any n1 gives the same
// outputs (w_n1^(jk) = w_n1^(j (k mod n1))); the rule fixes the kernels run.
static uint32_t td_log_n1(int es, uint32_t logN, uint64_t out_len) {
  const uint64_t n = 1ull << logN;
  const uint64_t n1 = std::min<uint64_t>(std::max<uint64_t>(next_pow2(2 * out_len), 2), n);  // out_len <= n <= 2^32
  uint32_t log_n1 = ilog2(n1);
  if (!decomposition_ok(es, logN, log_n1, out_len)) log_n1 = std::min<uint32_t>(log_n1, (uint32_t)max_log_radix(es));
  return log_n1;
}

namespace omr {

// Full forward transform of `batch` dense vectors under `variant`.
omr_status forward_full(omr_ctx *ctx, int f, int variant, const void *d_in, void *d_out,
                        uint32_t logN, uint64_t batch, uint64_t omega, hipStream_t stream, bool wide) {
  const int es = field_bytes(f);
  const uint64_t N = 1ull << logN;
  if (variant == OMR_DFT) {
    const void *src = d_in;
    if (d_in == d_out) {
      omr_status st = ensure_buf(ctx->scratch[0], (size_t)N * batch * es);
      if (st != OMR_OK) return st;
      if (hipMemcpyAsync(ctx->scratch[0].p, d_in, (size_t)N * batch * es, hipMemcpyDeviceToDevice,
                         stream) != hipSuccess)
        return OMR_E_DEVICE;
      src = ctx->scratch[0].p;
    }
    const TwTables *tab;
    omr_status st = get_tables(ctx, f, logN, omega, &tab);
    if (st != OMR_OK) return st;
    for (uint64_t b = 0; b < batch; b++) {
      st = launch_dft_full(f, (const char *)src + b * N * es, (char *)d_out + b * N * es, N, logN,
                           tab, stream);
      if (st != OMR_OK) return st;
    }
    return OMR_OK;
  }
  Schedule s;
  omr_status st = variant_schedule(f, variant, logN, s);
  if (st != OMR_OK) return st;
  if (variant == OMR_ARK_RADIX2 && !tuned_schedule_for(f, logN, batch, s) && batched_slot(f, logN, batch))
    default_schedule(f, logN, s);
  XformOpts o;
  o.wide_in = o.wide_out = wide;
  return run_transform(ctx, f, d_in, d_out, logN, omega, s, batch, N, N, o, stream);
}

}  // namespace omr

extern "C" omr_status omr_ntt_device(omr_ctx *ctx, int field, int variant, const void *d_in,
                                     void *d_out, uint64_t n, uint64_t batch, uint64_t omega,
                                     uint64_t s, void *stream) try {
  if (variant < 0 || variant >= OMR_VARIANT_COUNT) return OMR_E_VARIANT;
  omr_status st = validate_domain(field, n, omega);
  if (st != OMR_OK) return st;
  const uint32_t logN = ilog2(n);
  if (!omr_variant_supports(variant, logN)) return OMR_E_RADIX;
  if (batch > kMaxBatch) return OMR_E_ARG;
  if (!ctx || !d_in || !d_out) return OMR_E_NULL;
  std::lock_guard<std::mutex> lk(ctx->mu);
  if (hipSetDevice(ctx->device) != hipSuccess) return OMR_E_DEVICE;
  hipStream_t hs = (hipStream_t)stream;
  StreamOrder ord(ctx, hs);
  if (!is_partial_variant(variant)) return forward_full(ctx, field, variant, d_in, d_out, logN, batch, omega, hs);
  const uint64_t out_len = partial_out_len(s, n);
  if (out_len == 0) return OMR_OK;
  const int es = field_bytes(field);
  const int method = variant == OMR_DFT_PARTIAL ? 1 : 2;
  const uint32_t log_n1 = method == 2 ? td_log_n1(es, logN, out_len) : 0;
  for (uint64_t b = 0; b < batch; b++) {
    const char *in = (const char *)d_in + b * n * es;
    char *out = (char *)d_out + b * n * es;
    st = partial_impl(ctx, field, in, out, logN, omega, out_len, 0, method, log_n1, hs);
    if (st != OMR_OK) return st;
  }
  return OMR_OK;
} catch (const std::bad_alloc &) {
  return OMR_E_ALLOC;
} catch (...) {
  return OMR_E_DEVICE;
}

// NttEncoder::ntt on buffers of 64-bit words for a 31/32-bit field (the
// one-limb layout of src/fields.rs's define_field! types): full variants of
// the pass engine narrow in their first pass's loads and widen in their last
// pass's stores; Dft and the partial variants run on a narrowed copy.
extern "C" omr_status omr_ntt_device_words(omr_ctx *ctx, int field, int variant, const void *d_in, void *d_out,
                                           uint64_t n, uint64_t batch, uint64_t omega, uint64_t s,
                                           uint32_t word_bytes, void *stream) try {
  if (!field_ok(field)) return OMR_E_FIELD;
  const int es = field_bytes(field);
  if (word_bytes == 0 || word_bytes == (uint32_t)es)
    return omr_ntt_device(ctx, field, variant, d_in, d_out, n, batch, omega, s, stream);
  if (word_bytes != 8) return OMR_E_ARG;  // 32-bit fields in 64-bit words only
  if (variant < 0 || variant >= OMR_VARIANT_COUNT) return OMR_E_VARIANT;
  omr_status st = validate_domain(field, n, omega);
  if (st != OMR_OK) return st;
  const uint32_t logN = ilog2(n);
  if (!omr_variant_supports(variant, logN)) return OMR_E_RADIX;
  if (batch > kMaxBatch) return OMR_E_ARG;
  if (!ctx || !d_in || !d_out) return OMR_E_NULL;
  if (batch == 0) return OMR_OK;
  std::lock_guard<std::mutex> lk(ctx->mu);
  if (hipSetDevice(ctx->device) != hipSuccess) return OMR_E_DEVICE;
  hipStream_t hs = (hipStream_t)stream;
  StreamOrder ord(ctx, hs);
  const uint64_t *in64 = (const uint64_t *)d_in;
  uint64_t *out64 = (uint64_t *)d_out;
  if (n == 1)  // the size-1 transform is the identity
    return d_in == d_out || hipMemcpyAsync(d_out, d_in, batch * 8, hipMemcpyDeviceToDevice, hs) == hipSuccess
               ? OMR_OK
               : OMR_E_DEVICE;
  if (!is_partial_variant(variant) && variant != OMR_DFT)
    return forward_full(ctx, field, variant, d_in, d_out, logN, batch, omega, hs, /*wide=*/true);
  // narrowed copy in, the 32-bit path, widened copy of what it wrote
  if ((st = ensure_buf(ctx->words[0], (size_t)n * batch * 4)) != OMR_OK) return st;
  if ((st = launch_narrow(in64, (uint32_t *)ctx->words[0].p, n * batch, hs)) != OMR_OK) return st;
  const uint32_t *x32 = (const uint32_t *)ctx->words[0].p;
  if (variant == OMR_DFT) {
    if ((st = ensure_buf(ctx->words[1], (size_t)n * batch * 4)) != OMR_OK) return st;
    if ((st = forward_full(ctx, field, variant, x32, ctx->words[1].p, logN, batch, omega, hs)) != OMR_OK) return st;
    return launch_widen((const uint32_t *)ctx->words[1].p, out64, n * batch, hs);
  }
  const uint64_t out_len = partial_out_len(s, n);
  if (out_len == 0) return OMR_OK;
  const int method = variant == OMR_DFT_PARTIAL ? 1 : 2;
  const uint32_t log_n1 = method == 2 ? td_log_n1(es, logN, out_len) : 0;
  if ((st = ensure_buf(ctx->words[1], (size_t)out_len * 4)) != OMR_OK) return st;
  for (uint64_t b = 0; b < batch; b++) {
    if ((st = partial_impl(ctx, field, x32 + b * n, ctx->words[1].p, logN, omega, out_len, 0, method, log_n1, hs)) !=
        OMR_OK)
      return st;
    if ((st = launch_widen((const uint32_t *)ctx->words[1].p, out64 + b * n, out_len, hs)) != OMR_OK) return st;
  }
  return OMR_OK;
} catch (const std::bad_alloc &) {
  return OMR_E_ALLOC;
} catch (...) {
  return OMR_E_DEVICE;
}

extern "C" omr_status omr_intt_device(omr_ctx *ctx, int field, const void *d_in, void *d_out,
                                      uint64_t n, uint64_t batch, uint64_t omega, void *stream) try {
  omr_status st = validate_domain(field, n, omega);
  if (st != OMR_OK) return st;
  if (batch > kMaxBatch) return OMR_E_ARG;
  if (!ctx || !d_in || !d_out) return OMR_E_NULL;
  std::lock_guard<std::mutex> lk(ctx->mu);
  if (hipSetDevice(ctx->device) != hipSuccess) return OMR_E_DEVICE;
  StreamOrder ord(ctx, (hipStream_t)stream);
  const uint32_t logN = ilog2(n);
  Schedule s;
  if (!tuned_schedule_for(field, logN, batch, s)) default_schedule(field, logN, s);
  XformOpts o;
  o.scale = true;
  o.scale_w = to_w_host(field, finv_host(field, n % kP[field]));
  return run_transform(ctx, field, d_in, d_out, logN, finv_host(field, omega), s, batch, n, n, o,
                       (hipStream_t)stream);
} catch (const std::bad_alloc &) {
  return OMR_E_ALLOC;
} catch (...) {
  return OMR_E_DEVICE;
}

extern "C" omr_status omr_ntt_schedule_device(omr_ctx *ctx, int field, const void *d_in,
                                              void *d_out, uint64_t n, uint64_t batch,
                                              uint64_t omega, int inverse, const int *log_radix,
                                              int npass, int log_e, void *stream) try {
  omr_status st = validate_domain(field, n, omega);
  if (st != OMR_OK) return st;
  if (batch > kMaxBatch) return OMR_E_ARG;
  if (!log_radix) return OMR_E_NULL;
  if (npass < 1 || npass > 8) return OMR_E_SCHEDULE;
  {
    int sum = 0;
    for (int i = 0; i < npass; i++) {
      if (log_radix[i] < 1 || log_radix[i] > max_log_radix(field_bytes(field))) return OMR_E_SCHEDULE;
      sum += log_radix[i];
    }
    if (sum != ilog2(n)) return OMR_E_SCHEDULE;
  }
  if (!ctx || !d_in || !d_out) return OMR_E_NULL;
  if (log_e != 0 && (log_e < 1 || log_e > 4)) return OMR_E_SCHEDULE;
  Schedule s;
  for (int i = 0; i < npass; i++) s.logR.push_back(log_radix[i]);
  s.logE = log_e;
  std::lock_guard<std::mutex> lk(ctx->mu);
  if (hipSetDevice(ctx->device) != hipSuccess) return OMR_E_DEVICE;
  StreamOrder ord(ctx, (hipStream_t)stream);
  const uint32_t logN = ilog2(n);
  XformOpts o;
  uint64_t root = omega;
  if (inverse) {
    root = finv_host(field, omega);
    o.scale = true;
    o.scale_w = to_w_host(field, finv_host(field, n % kP[field]));
  }
  return run_transform(ctx, field, d_in, d_out, logN, root, s, batch, n, n, o, (hipStream_t)stream);
} catch (const std::bad_alloc &) {
  return OMR_E_ALLOC;
} catch (...) {
  return OMR_E_DEVICE;
}

extern "C" omr_status omr_partial_device(omr_ctx *ctx, int field, const void *d_in, void *d_out,
                                         uint64_t n, uint64_t omega, uint64_t out_len, int mode,
                                         int method, uint32_t log_n1, void *stream) try {
  omr_status st = validate_domain(field, n, omega);
  if (st != OMR_OK) return st;
  if (out_len == 0 || out_len > n) return OMR_E_OUTLEN;
  if (mode != 0 && mode != 1) return OMR_E_ARG;
  if (mode == 1 && (out_len & (out_len - 1))) return OMR_E_OUTLEN;
  if (method < 0 || method > 3) return OMR_E_ARG;
  const int es = field_bytes(field);
  if (mode == 0 && method == 2 && log_n1 != 0 && !decomposition_ok(es, ilog2(n), log_n1, out_len))
    return OMR_E_SCHEDULE;
  if (!ctx || !d_in || !d_out) return OMR_E_NULL;
  std::lock_guard<std::mutex> lk(ctx->mu);
  if (hipSetDevice(ctx->device) != hipSuccess) return OMR_E_DEVICE;
  StreamOrder ord(ctx, (hipStream_t)stream);
  return partial_impl(ctx, field, d_in, d_out, ilog2(n), omega, out_len, mode, method, log_n1,
                      (hipStream_t)stream);
} catch (const std::bad_alloc &) {
  return OMR_E_ALLOC;
} catch (...) {
  return OMR_E_DEVICE;
}

extern "C" omr_status omr_partial_batch_device(omr_ctx *ctx, int field, const void *d_in, void *d_out,
                                               uint64_t n, uint64_t batch, uint64_t omega, uint64_t out_len,
                                               void *stream) try {
  omr_status st = validate_domain(field, n, omega);
  if (st != OMR_OK) return st;
  if (out_len == 0 || out_len > n) return OMR_E_OUTLEN;
  if (batch == 0 || batch > 0xFFFFFFFFull) return OMR_E_ARG;
  if (!ctx || !d_in || !d_out) return OMR_E_NULL;
  std::lock_guard<std::mutex> lk(ctx->mu);
  if (hipSetDevice(ctx->device) != hipSuccess) return OMR_E_DEVICE;
  StreamOrder ord(ctx, (hipStream_t)stream);
  return partial_batch_impl(ctx, field, d_in, d_out, ilog2(n), batch, omega, out_len, (hipStream_t)stream);
} catch (const std::bad_alloc &) {
  return OMR_E_ALLOC;
} catch (...) {
  return OMR_E_DEVICE;
}

extern "C" omr_status omr_field_op_device(int field, int op, const void *d_a, const void *d_b,
                                          void *d_out, uint64_t n, void *stream) try {
  if (!d_a || !d_b || !d_out) return OMR_E_NULL;
  if (!field_ok(field)) return OMR_E_FIELD;
  if (op < 0 || op > 4) return OMR_E_ARG;
  return launch_field_op(field, op, d_a, d_b, d_out, n, (hipStream_t)stream);
} catch (const std::bad_alloc &) {
  return OMR_E_ALLOC;
} catch (...) {
  return OMR_E_DEVICE;
}

// NttEncoder::ntt on a host buffer (src/encoder.rs:70-74): copy in, run, copy out.
extern "C" omr_status omr_ntt(omr_ctx *ctx, int field, int variant, void *buf, uint64_t n,
                              uint64_t omega, uint64_t s) try {
  if (variant < 0 || variant >= OMR_VARIANT_COUNT) return OMR_E_VARIANT;
  omr_status st = validate_domain(field, n, omega);
  if (st != OMR_OK) return st;
  const uint32_t logN = ilog2(n);
  if (!omr_variant_supports(variant, logN)) return OMR_E_RADIX;
  if (!ctx || !buf) return OMR_E_NULL;
  const int es = field_bytes(field);
  std::lock_guard<std::mutex> lk(ctx->mu);
  if (hipSetDevice(ctx->device) != hipSuccess) return OMR_E_DEVICE;
  StreamOrder ord(ctx, nullptr);  // this form runs on the null stream
  if ((st = ensure_buf(ctx->io[0], (size_t)n * es)) != OMR_OK) return st;
  if ((st = host_to_device(ctx, ctx->io[0].p, buf, (size_t)n * es)) != OMR_OK) return st;
  uint64_t copy_back = n;
  void *result = ctx->io[0].p;
  if (!is_partial_variant(variant)) {
    st = forward_full(ctx, field, variant, ctx->io[0].p, ctx->io[0].p, logN, 1, omega, nullptr);
  } else {
    copy_back = partial_out_len(s, n);
    if (copy_back == 0) return OMR_OK;
    if ((st = ensure_buf(ctx->io[1], (size_t)copy_back * es)) != OMR_OK) return st;
    const int method = variant == OMR_DFT_PARTIAL ? 1 : 2;
    const uint32_t log_n1 = method == 2 ? td_log_n1(es, logN, copy_back) : 0;
    st = partial_impl(ctx, field, ctx->io[0].p, ctx->io[1].p, logN, omega, copy_back, 0, method,
                      log_n1, nullptr);
    result = ctx->io[1].p;
  }
  if (st != OMR_OK) return st;
  return device_to_host(ctx, buf, result, (size_t)copy_back * es);
} catch (const std::bad_alloc &) {
  return OMR_E_ALLOC;
} catch (...) {
  return OMR_E_DEVICE;
}

// NttEncoder::ntt on a host buffer of word_bytes-wide words (64-bit words for
// a 31/32-bit field: the crate's own `[F]` storage, one u64 limb each, whose
// Montgomery form passes through unchanged): copy in, omr_ntt_device_words,
// copy out.  word_bytes 0 or the field's width is omr_ntt.
extern "C" omr_status omr_ntt_words(omr_ctx *ctx, int field, int variant, void *buf, uint64_t n, uint64_t omega,
                                    uint64_t s, uint32_t word_bytes) try {
  if (!field_ok(field)) return OMR_E_FIELD;
  const int es = field_bytes(field);
  if (word_bytes == 0 || word_bytes == (uint32_t)es) return omr_ntt(ctx, field, variant, buf, n, omega, s);
  if (word_bytes != 8) return OMR_E_ARG;
  if (variant < 0 || variant >= OMR_VARIANT_COUNT) return OMR_E_VARIANT;
  omr_status st = validate_domain(field, n, omega);
  if (st != OMR_OK) return st;
  if (!omr_variant_supports(variant, ilog2(n))) return OMR_E_RADIX;
  if (!ctx || !buf) return OMR_E_NULL;
  std::lock_guard<std::mutex> lk(ctx->host_mu);
  if (hipSetDevice(ctx->device) != hipSuccess) return OMR_E_DEVICE;
  const size_t bytes = (size_t)n * 8;
  const uint64_t back = is_partial_variant(variant) ? partial_out_len(s, n) : n;
  if (back == 0) return OMR_OK;
  if ((st = ensure_buf(ctx->hio[0], bytes)) != OMR_OK) return st;
  if ((st = host_to_device(ctx, ctx->hio[0].p, buf, bytes)) != OMR_OK) return st;
  if ((st = omr_ntt_device_words(ctx, field, variant, ctx->hio[0].p, ctx->hio[0].p, n, 1, omega, s, 8, nullptr)) !=
      OMR_OK)
    return st;
  return device_to_host(ctx, buf, ctx->hio[0].p, (size_t)back * 8);
} catch (const std::bad_alloc &) {
  return OMR_E_ALLOC;
} catch (...) {
  return OMR_E_DEVICE;
}

// ---- host-buffer forms of the device entry points ---------------------------
// Synchronous like omr_ntt: copy in, run on the null stream, copy out.  The
// device entry point validates first, so a bad argument never touches the GPU.

static omr_status host_stage(omr_ctx *ctx, int i, size_t bytes, const void *src) {
  omr_status st = ensure_buf(ctx->hio[i], bytes);
  if (st != OMR_OK) return st;
  return src ? host_to_device(ctx, ctx->hio[i].p, src, bytes) : OMR_OK;
}

extern "C" omr_status omr_intt(omr_ctx *ctx, int field, void *buf, uint64_t n, uint64_t omega) try {
  omr_status st = validate_domain(field, n, omega);
  if (st != OMR_OK) return st;
  if (!ctx || !buf) return OMR_E_NULL;
  std::lock_guard<std::mutex> lk(ctx->host_mu);
  if (hipSetDevice(ctx->device) != hipSuccess) return OMR_E_DEVICE;
  const size_t bytes = (size_t)n * field_bytes(field);
  if ((st = host_stage(ctx, 0, bytes, buf)) != OMR_OK) return st;
  if ((st = omr_intt_device(ctx, field, ctx->hio[0].p, ctx->hio[0].p, n, 1, omega, nullptr)) != OMR_OK) return st;
  return device_to_host(ctx, buf, ctx->hio[0].p, bytes);
} catch (const std::bad_alloc &) {
  return OMR_E_ALLOC;
} catch (...) {
  return OMR_E_DEVICE;
}

extern "C" omr_status omr_ntt_batched(omr_ctx *ctx, int field, void *buf, uint64_t n, uint64_t batch,
                                      uint64_t omega) try {
  omr_status st = validate_domain(field, n, omega);
  if (st != OMR_OK) return st;
  if (batch == 0 || batch > kMaxBatch) return OMR_E_ARG;
  if (!ctx || !buf) return OMR_E_NULL;
  std::lock_guard<std::mutex> lk(ctx->host_mu);
  if (hipSetDevice(ctx->device) != hipSuccess) return OMR_E_DEVICE;
  const size_t bytes = (size_t)n * batch * field_bytes(field);
  if ((st = host_stage(ctx, 0, bytes, buf)) != OMR_OK) return st;
  if ((st = omr_ntt_device(ctx, field, OMR_ARK_RADIX2, ctx->hio[0].p, ctx->hio[0].p, n, batch, omega, 0,
                           nullptr)) != OMR_OK)
    return st;
  return device_to_host(ctx, buf, ctx->hio[0].p, bytes);
} catch (const std::bad_alloc &) {
  return OMR_E_ALLOC;
} catch (...) {
  return OMR_E_DEVICE;
}

extern "C" omr_status omr_coset_lde(omr_ctx *ctx, int field, const void *coeffs, void *evals, uint64_t n,
                                    uint64_t ncols, uint32_t blowup, uint64_t shift) try {
  uint64_t sh = shift;
  uint32_t logM;
  omr_status st = lde_args(field, n, ncols, blowup, sh, logM);
  if (st != OMR_OK) return st;
  if (!ctx || !coeffs || !evals) return OMR_E_NULL;
  std::lock_guard<std::mutex> lk(ctx->host_mu);
  if (hipSetDevice(ctx->device) != hipSuccess) return OMR_E_DEVICE;
  const int es = field_bytes(field);
  const size_t in_bytes = (size_t)n * ncols * es, out_bytes = in_bytes * blowup;
  if ((st = host_stage(ctx, 0, in_bytes, coeffs)) != OMR_OK) return st;
  if ((st = host_stage(ctx, 1, out_bytes, nullptr)) != OMR_OK) return st;
  if ((st = omr_coset_lde_device(ctx, field, ctx->hio[0].p, ctx->hio[1].p, n, ncols, blowup, shift, nullptr)) !=
      OMR_OK)
    return st;
  return device_to_host(ctx, evals, ctx->hio[1].p, out_bytes);
} catch (const std::bad_alloc &) {
  return OMR_E_ALLOC;
} catch (...) {
  return OMR_E_DEVICE;
}

extern "C" omr_status omr_coset_decode(omr_ctx *ctx, int field, const void *evals, void *coeffs, uint64_t n,
                                       uint64_t ncols, uint32_t blowup, uint64_t shift, uint32_t *nonzero) try {
  uint64_t sh = shift;
  uint32_t logM;
  omr_status st = lde_args(field, n, ncols, blowup, sh, logM);
  if (st != OMR_OK) return st;
  if (!ctx || !evals || !coeffs) return OMR_E_NULL;
  std::lock_guard<std::mutex> lk(ctx->host_mu);
  if (hipSetDevice(ctx->device) != hipSuccess) return OMR_E_DEVICE;
  const int es = field_bytes(field);
  const size_t out_bytes = (size_t)n * ncols * es, in_bytes = out_bytes * blowup;
  if ((st = host_stage(ctx, 0, in_bytes, evals)) != OMR_OK) return st;
  if ((st = host_stage(ctx, 1, out_bytes, nullptr)) != OMR_OK) return st;
  if (nonzero && (st = host_stage(ctx, 2, sizeof(uint32_t), nullptr)) != OMR_OK) return st;
  if ((st = omr_coset_decode_device(ctx, field, ctx->hio[0].p, ctx->hio[1].p, n, ncols, blowup, shift,
                                    nonzero ? (uint32_t *)ctx->hio[2].p : nullptr, nullptr)) != OMR_OK)
    return st;
  if ((st = device_to_host(ctx, coeffs, ctx->hio[1].p, out_bytes)) != OMR_OK) return st;
  return nonzero ? device_to_host(ctx, nonzero, ctx->hio[2].p, sizeof(uint32_t)) : OMR_OK;
} catch (const std::bad_alloc &) {
  return OMR_E_ALLOC;
} catch (...) {
  return OMR_E_DEVICE;
}

// NttDomain's public twiddles / bitrev_twiddles (src/encoder.rs:14-18).
extern "C" omr_status omr_domain_tables_device(omr_ctx *ctx, int field, uint64_t n, uint64_t omega,
                                               void *d_twiddles, void *d_bitrev_twiddles, void *stream) try {
  omr_status st = validate_domain(field, n, omega);
  if (st != OMR_OK) return st;
  if (!ctx) return OMR_E_NULL;
  std::lock_guard<std::mutex> lk(ctx->mu);
  if (hipSetDevice(ctx->device) != hipSuccess) return OMR_E_DEVICE;
  hipStream_t hs = (hipStream_t)stream;
  StreamOrder ord(ctx, hs);
  uint32_t logN = 0;
  while ((1ull << logN) < n) logN++;
  if (n == 1) {  // twiddles = [1], bitrev_twiddles = []
    if (!d_twiddles) return OMR_OK;
    static const uint64_t one = 1;  // canonical 1, little-endian: the low word serves 4-byte fields
    return hipMemcpyAsync(d_twiddles, &one, field_bytes(field), hipMemcpyHostToDevice, hs) == hipSuccess
               ? OMR_OK
               : OMR_E_DEVICE;
  }
  const TwTables *tab;
  if ((st = get_tables(ctx, field, logN, omega, &tab)) != OMR_OK) return st;
  return launch_domain_tables(field, d_twiddles, d_bitrev_twiddles, n, logN, tab, hs);
} catch (const std::bad_alloc &) {
  return OMR_E_ALLOC;
} catch (...) {
  return OMR_E_DEVICE;
}

extern "C" omr_status omr_domain_tables(omr_ctx *ctx, int field, uint64_t n, uint64_t omega, void *twiddles,
                                        void *bitrev_twiddles) try {
  omr_status st = validate_domain(field, n, omega);
  if (st != OMR_OK) return st;
  if (!ctx) return OMR_E_NULL;
  std::lock_guard<std::mutex> lk(ctx->host_mu);
  if (hipSetDevice(ctx->device) != hipSuccess) return OMR_E_DEVICE;
  const size_t es = field_bytes(field);
  if (twiddles && (st = ensure_buf(ctx->hio[0], n * es)) != OMR_OK) return st;
  if (bitrev_twiddles && n >= 2 && (st = ensure_buf(ctx->hio[1], n / 2 * es)) != OMR_OK) return st;
  if ((st = omr_domain_tables_device(ctx, field, n, omega, twiddles ? ctx->hio[0].p : nullptr,
                                     (bitrev_twiddles && n >= 2) ? ctx->hio[1].p : nullptr, nullptr)) != OMR_OK)
    return st;
  if (twiddles && (st = device_to_host(ctx, twiddles, ctx->hio[0].p, n * es)) != OMR_OK) return st;
  if (bitrev_twiddles && n >= 2 && (st = device_to_host(ctx, bitrev_twiddles, ctx->hio[1].p, n / 2 * es)) != OMR_OK)
    return st;
  return OMR_OK;
} catch (const std::bad_alloc &) {
  return OMR_E_ALLOC;
} catch (...) {
  return OMR_E_DEVICE;
}
