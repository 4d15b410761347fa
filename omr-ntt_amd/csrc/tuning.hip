// tuning.hip -- schedule classes, the timing autotuner and the tuned-schedule
// table (host code).
//
// A shape is (field, N = 2^logN, out_len = 2^logS, batch).  Candidate
// schedules come from a fixed vocabulary of kNumClasses classes so that the
// cost model (mlp.hip) can score all of them with one output vector:
//   full transforms (out_len == N): classes 0..29
//     class = 0..2:   one pass, register radix 2^(2 + class)
//     class = 3..29:  P = 2 + (c-3)/9 passes (2..4), split shape (c-3)/3 % 3
//                     (0 balanced, 1 first pass largest, 2 last pass largest),
//                     register radix 2^(2 + (c-3) % 3)
//   partial prefix (out_len < N): classes 30..35
//     30: direct partial DFT; 31..33: decomposition with n1 = 2^6, 2^8,
//     2^10; 34: decomposition with n1 = 4 out_len (at most N/64, above 2^10),
//     its sub-NTTs in several passes with the recombine fused into the last;
//     35: full transform, pruned last pass.
// The timing autotuner measures every valid class of a shape; its best is
// the reference the cost model's regret is measured against.
#include <hip/hip_runtime.h>
#include <new>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <map>
#include <mutex>
#include <tuple>
#include <utility>
#include <vector>

#include "dispatch.h"
#include "internal.h"

namespace omr {

static std::mutex g_tune_mu;
static std::map<std::pair<int, uint32_t>, Schedule> g_tuned;        // single vectors
static std::map<std::pair<int, uint32_t>, Schedule> g_tuned_batch;  // batches of >= 32 MB
static std::map<std::tuple<int, uint32_t, uint32_t>, std::pair<int, uint32_t>> g_tuned_partial;

bool batched_slot(int f, uint32_t logN, uint64_t batch) {
  return batch >= 2 && ((batch << logN) * (uint64_t)field_bytes(f)) >= (32ull << 20);
}

bool tuned_schedule(int f, uint32_t logN, Schedule &s) {
  std::lock_guard<std::mutex> lk(g_tune_mu);
  auto it = g_tuned.find({f, logN});
  if (it == g_tuned.end()) return false;
  s = it->second;
  return true;
}

// Batches of 32 MB or more have a slot of their own: the single-vector
// timings favour small cache-resident passes, which lose up to 40 % on such
// batches.  An empty batched slot is not filled from the single one: the
// caller's default (coalescing-first) schedule is the better guess there.
bool tuned_schedule_for(int f, uint32_t logN, uint64_t batch, Schedule &s) {
  if (!batched_slot(f, logN, batch)) return tuned_schedule(f, logN, s);
  std::lock_guard<std::mutex> lk(g_tune_mu);
  auto it = g_tuned_batch.find({f, logN});
  if (it == g_tuned_batch.end()) return false;
  s = it->second;
  return true;
}

void set_tuned_schedule(int f, uint32_t logN, const Schedule &s) {
  std::lock_guard<std::mutex> lk(g_tune_mu);
  g_tuned[{f, logN}] = s;
}

static void set_tuned_schedule_batch(int f, uint32_t logN, const Schedule &s) {
  std::lock_guard<std::mutex> lk(g_tune_mu);
  g_tuned_batch[{f, logN}] = s;
}

bool tuned_partial(int f, uint32_t logN, uint32_t logOut, int &method, uint32_t &log_n1) {
  std::lock_guard<std::mutex> lk(g_tune_mu);
  auto it = g_tuned_partial.find(std::make_tuple(f, logN, logOut));
  if (it == g_tuned_partial.end()) return false;
  method = it->second.first;
  log_n1 = it->second.second;
  return true;
}

// Concrete schedule of a full-transform class for 2^logN; false if invalid.
bool class_schedule(int f, int cls, uint32_t logN, Schedule &s) {
  const int es = field_bytes(f);
  const int maxlr = max_log_radix(es);
  s.logR.clear();
  if (cls < 0 || cls >= kFullClasses || logN < 1) return false;
  if (cls < 3) {
    if ((int)logN > maxlr) return false;
    s.logR.push_back((int)logN);
    s.logE = 2 + cls;
    return s.logE <= (int)logN;
  }
  const int c = cls - 3;
  const int P = 2 + c / 9, shape = (c / 3) % 3;
  s.logE = 2 + c % 3;
  const int L = (int)logN;
  if (L < P * s.logE) return false;  // every pass at least one register radix
  if ((L + P - 1) / P > maxlr) return false;
  std::vector<int> parts;
  if (shape == 0) {
    const int base = L / P, rem = L % P;
    for (int i = 0; i < P; i++) parts.push_back(base + (i < rem ? 1 : 0));
  } else {
    // one pass as large as allowed, the others balanced (each >= logE)
    int big = std::min(maxlr, L - (P - 1) * s.logE);
    const int rest = L - big;
    if (big < (rest + P - 2) / (P - 1)) return false;  // not the largest: same as balanced
    std::vector<int> small;
    const int base = rest / (P - 1), rem = rest % (P - 1);
    for (int i = 0; i < P - 1; i++) small.push_back(base + (i < rem ? 1 : 0));
    if (shape == 1) {
      parts.push_back(big);
      parts.insert(parts.end(), small.begin(), small.end());
    } else {
      parts = small;
      parts.push_back(big);
    }
  }
  for (int r : parts)
    if (r < 1 || r > maxlr || r < s.logE) return false;
  s.logR = parts;
  return true;
}

// Partial classes (prefix): method and log n1 (0 = not used).
bool class_partial(int f, int cls, uint32_t logN, uint32_t logOut, int &method, uint32_t &log_n1) {
  const int maxlr = max_log_radix(field_bytes(f));
  if (cls < kFullClasses || cls >= kNumClasses || logOut >= logN) return false;
  const int c = cls - kFullClasses;
  log_n1 = 0;
  if (c == 0) {
    method = 1;
    return logOut <= 6;  // out_len * N products: only tiny prefixes
  }
  if (c == 4) {  // multi-pass fused decomposition
    method = 2;
    log_n1 = multipass_log_n1(logN, 1ull << logOut);
    return log_n1 != 0;
  }
  if (c <= 3) {
    static const int n1s[3] = {6, 8, 10};
    const int ln1 = n1s[c - 1];
    if (ln1 > (int)logN || ln1 > maxlr) return false;
    method = 2;
    log_n1 = (uint32_t)ln1;
    // recombine work out_len * N / n1 must stay below ~4 N
    return (int)logOut - ln1 <= 2;
  }
  method = 3;
  return true;
}

}  // namespace omr

using namespace omr;

extern "C" int omr_num_schedule_classes(void) { return kNumClasses; }

extern "C" omr_status omr_schedule_class(int field, uint64_t n, uint64_t out_len, int cls,
                                         int *log_radix, int *npass, int *log_e, int *method,
                                         uint32_t *log_n1) try {
  if (!field_ok(field)) return OMR_E_FIELD;
  if (n == 0 || (n & (n - 1)) || out_len == 0 || out_len > n || (out_len & (out_len - 1)))
    return OMR_E_SIZE;
  const uint32_t logN = 63 - __builtin_clzll(n), logOut = 63 - __builtin_clzll(out_len);
  if (out_len == n) {
    Schedule s;
    if (!class_schedule(field, cls, logN, s)) return OMR_E_SCHEDULE;
    if (npass) *npass = (int)s.logR.size();
    if (log_radix)
      for (size_t i = 0; i < s.logR.size(); i++) log_radix[i] = s.logR[i];
    if (log_e) *log_e = s.logE;
    if (method) *method = 0;
    if (log_n1) *log_n1 = 0;
    return OMR_OK;
  }
  int m;
  uint32_t l1;
  if (!class_partial(field, cls, logN, logOut, m, l1)) return OMR_E_SCHEDULE;
  if (method) *method = m;
  if (log_n1) *log_n1 = l1;
  if (npass) *npass = 0;
  return OMR_OK;
} catch (const std::bad_alloc &) {
  return OMR_E_ALLOC;
} catch (...) {
  return OMR_E_DEVICE;
}

// Times every valid class of a shape: times_us[cls] = median microseconds,
// or -1 for classes invalid for the shape.  Inputs are pseudo-random field
// elements in a context-owned buffer.
extern "C" omr_status omr_tune_classes(omr_ctx *ctx, int field, uint64_t n, uint64_t out_len,
                                       uint64_t batch, int reps, float *times_us) try {
  return tune_classes_subset(ctx, field, n, out_len, batch, reps, nullptr, times_us);
} catch (const std::bad_alloc &) {
  return OMR_E_ALLOC;
} catch (...) {
  return OMR_E_DEVICE;
}

// Times the valid classes with only[cls] set (every valid class when only is
// null); the others get -1.
omr_status omr::tune_classes_subset(omr_ctx *ctx, int field, uint64_t n, uint64_t out_len, uint64_t batch,
                                    int reps, const bool *only, float *times_us) {
  if (!field_ok(field)) return OMR_E_FIELD;
  if (!times_us) return OMR_E_NULL;
  if (n == 0 || (n & (n - 1)) || out_len == 0 || out_len > n || (out_len & (out_len - 1)))
    return OMR_E_SIZE;
  int two_adicity = 0;
  omr_field_info(field, nullptr, nullptr, &two_adicity);
  const uint32_t logN = 63 - __builtin_clzll(n), logOut = 63 - __builtin_clzll(out_len);
  if ((int)logN > two_adicity) return OMR_E_TWO_ADICITY;
  if (batch == 0 || batch > kMaxBatch || (out_len < n && batch != 1)) return OMR_E_ARG;
  if (reps < 1) reps = 1;
  if (!ctx) return OMR_E_NULL;
  std::lock_guard<std::mutex> lk(ctx->mu);
  if (hipSetDevice(ctx->device) != hipSuccess) return OMR_E_DEVICE;
  const int es = field_bytes(field);
  const size_t bytes = (size_t)n * batch * es;
  omr_status st;
  if ((st = ensure_buf(ctx->io[0], bytes)) != OMR_OK) return st;
  if ((st = ensure_buf(ctx->io[1], bytes)) != OMR_OK) return st;
  // fill the input with reduced pseudo-random words (any canonical data will do)
  {
    std::vector<uint64_t> h(std::min<uint64_t>(n * batch, 1u << 20));
    uint64_t z = 0x9E3779B97F4A7C15ull ^ n;
    for (auto &v : h) {
      z += 0x9E3779B97F4A7C15ull;
      uint64_t r = z;
      r = (r ^ (r >> 30)) * 0xBF58476D1CE4E5B9ull;
      r = (r ^ (r >> 27)) * 0x94D049BB133111EBull;
      v = (r ^ (r >> 31)) % field_modulus(field);
    }
    std::vector<uint32_t> h32;
    const void *src = h.data();
    if (es == 4) {
      h32.assign(h.begin(), h.end());
      src = h32.data();
    }
    for (size_t off = 0; off < bytes; off += h.size() * es) {
      size_t chunk = std::min(bytes - off, h.size() * (size_t)es);
      if (hipMemcpy((char *)ctx->io[0].p + off, src, chunk, hipMemcpyHostToDevice) != hipSuccess)
        return OMR_E_DEVICE;
    }
  }
  uint64_t w;
  omr_root_of_unity(field, n, &w);
  hipStream_t s;
  if (hipStreamCreateWithFlags(&s, hipStreamNonBlocking) != hipSuccess) return OMR_E_DEVICE;
  StreamOrder ord(ctx, s);
  hipEvent_t e0, e1;
  (void)hipEventCreate(&e0);
  (void)hipEventCreate(&e1);
  for (int cls = 0; cls < kNumClasses; cls++) {
    times_us[cls] = -1.f;
    if (only && !only[cls]) continue;
    Schedule sch;
    int method = 0;
    uint32_t ln1 = 0;
    bool full = out_len == n;
    if (full ? !class_schedule(field, cls, logN, sch) : !class_partial(field, cls, logN, logOut, method, ln1))
      continue;
    auto run = [&]() -> omr_status {
      if (full) {
        XformOpts o;
        return run_transform(ctx, field, ctx->io[0].p, ctx->io[1].p, logN, w, sch, batch, n, n, o, s);
      }
      return partial_impl(ctx, field, ctx->io[0].p, ctx->io[1].p, logN, w, out_len, 0, method, ln1, s);
    };
    if ((st = run()) != OMR_OK) break;  // warm-up (tables, scratch)
    std::vector<float> ts;
    for (int r = 0; r < reps; r++) {
      (void)hipEventRecord(e0, s);
      if ((st = run()) != OMR_OK) break;
      (void)hipEventRecord(e1, s);
      if (hipEventSynchronize(e1) != hipSuccess) {
        st = OMR_E_DEVICE;
        break;
      }
      float ms = 0;
      (void)hipEventElapsedTime(&ms, e0, e1);
      ts.push_back(ms * 1000.f);
    }
    if (st != OMR_OK) break;
    std::sort(ts.begin(), ts.end());
    times_us[cls] = ts[ts.size() / 2];
  }
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  (void)hipStreamSynchronize(s);
  (void)hipStreamDestroy(s);
  return st;
}

// Class of the untuned default choice for a shape (the method without tuning
// or the cost model): default_schedule for full transforms, the automatic
// partial rule otherwise.
extern "C" int omr_default_class(int field, uint64_t n, uint64_t out_len) {
  if (!field_ok(field) || n < 2 || (n & (n - 1)) || out_len == 0 || out_len > n) return -1;
  const uint32_t logN = 63 - __builtin_clzll(n);
  const int maxlr = max_log_radix(field_bytes(field));
  if (out_len < n) {  // the rule partial_impl applies without a tuned entry
    if (out_len <= 4) return kFullClasses;
    if (out_len <= (1ull << kPrefixMaxLogR)) {  // one fused pass: the class of the next n1 up
      const uint32_t ln1 = out_len <= 64 ? 6 : (out_len <= 256 ? 8 : 10);
      if (ln1 <= logN) return kFullClasses + (int)(ln1 - 4) / 2;
      return kFullClasses + 5;
    }
    return multipass_log_n1(logN, out_len) ? kFullClasses + 4 : kFullClasses + 5;
  }
  Schedule d;
  default_schedule(field, logN, d);
  // default register radix: 2^min(4, logR) per pass
  const int want_e = std::min(4, *std::min_element(d.logR.begin(), d.logR.end()));
  for (int c = 0; c < kFullClasses; c++) {
    Schedule s;
    if (class_schedule(field, c, logN, s) && s.logR == d.logR && s.logE == want_e) return c;
  }
  (void)maxlr;
  return -1;
}

// Installs class `cls` as the tuned schedule of (field, n, out_len).
extern "C" omr_status omr_tune_set_class(int field, uint64_t n, uint64_t out_len, int cls) try {
  if (!field_ok(field)) return OMR_E_FIELD;
  if (n == 0 || (n & (n - 1)) || out_len == 0 || out_len > n || (out_len & (out_len - 1)))
    return OMR_E_SIZE;
  const uint32_t logN = 63 - __builtin_clzll(n), logOut = 63 - __builtin_clzll(out_len);
  if (out_len == n) {
    Schedule s;
    if (!class_schedule(field, cls, logN, s)) return OMR_E_SCHEDULE;
    set_tuned_schedule(field, logN, s);
    return OMR_OK;
  }
  int m;
  uint32_t l1;
  if (!class_partial(field, cls, logN, logOut, m, l1)) return OMR_E_SCHEDULE;
  std::lock_guard<std::mutex> lk(g_tune_mu);
  g_tuned_partial[std::make_tuple(field, logN, logOut)] = {m, l1};
  return OMR_OK;
} catch (const std::bad_alloc &) {
  return OMR_E_ALLOC;
} catch (...) {
  return OMR_E_DEVICE;
}

// Installs class `cls` for `batch` vectors of (field, n, out_len): the
// batched slot for full transforms of 32 MB or more, else as omr_tune_set_class.
extern "C" omr_status omr_tune_set_class_batch(int field, uint64_t n, uint64_t out_len, uint64_t batch,
                                               int cls) try {
  if (!field_ok(field)) return OMR_E_FIELD;
  if (n == 0 || (n & (n - 1)) || out_len == 0 || out_len > n || (out_len & (out_len - 1)))
    return OMR_E_SIZE;
  if (batch == 0) return OMR_E_ARG;
  const uint32_t logN = 63 - __builtin_clzll(n);
  if (out_len == n && batched_slot(field, logN, batch)) {
    Schedule s;
    if (!class_schedule(field, cls, logN, s)) return OMR_E_SCHEDULE;
    set_tuned_schedule_batch(field, logN, s);
    return OMR_OK;
  }
  return omr_tune_set_class(field, n, out_len, cls);
} catch (const std::bad_alloc &) {
  return OMR_E_ALLOC;
} catch (...) {
  return OMR_E_DEVICE;
}

extern "C" omr_status omr_tune_clear(void) try {
  std::lock_guard<std::mutex> lk(g_tune_mu);
  g_tuned.clear();
  g_tuned_batch.clear();
  g_tuned_partial.clear();
  return OMR_OK;
} catch (const std::bad_alloc &) {
  return OMR_E_ALLOC;
} catch (...) {
  return OMR_E_DEVICE;
}

// Text checkpoint of the tuned tables: one line per entry,
//   F <field> <logN> <logE> <npass> <logR...>   (single vectors),
//   B <field> <logN> <logE> <npass> <logR...>   (batches of 32 MB or more), or
//   P <field> <logN> <logOut> <method> <log_n1>
extern "C" omr_status omr_tune_save(const char *path) try {
  if (!path) return OMR_E_NULL;
  std::lock_guard<std::mutex> lk(g_tune_mu);
  FILE *fp = fopen(path, "w");
  if (!fp) return OMR_E_ARG;
  for (int slot = 0; slot < 2; slot++) {
    for (auto &kv : slot ? g_tuned_batch : g_tuned) {
      fprintf(fp, "%c %d %u %d %zu", slot ? 'B' : 'F', kv.first.first, kv.first.second, kv.second.logE,
              kv.second.logR.size());
      for (int r : kv.second.logR) fprintf(fp, " %d", r);
      fprintf(fp, "\n");
    }
  }
  for (auto &kv : g_tuned_partial)
    fprintf(fp, "P %d %u %u %d %u\n", std::get<0>(kv.first), std::get<1>(kv.first), std::get<2>(kv.first),
            kv.second.first, kv.second.second);
  fclose(fp);
  return OMR_OK;
} catch (const std::bad_alloc &) {
  return OMR_E_ALLOC;
} catch (...) {
  return OMR_E_DEVICE;
}

extern "C" omr_status omr_tune_load(const char *path) try {
  if (!path) return OMR_E_NULL;
  FILE *fp = fopen(path, "r");
  if (!fp) return OMR_E_ARG;
  char kind;
  omr_status st = OMR_OK;
  std::lock_guard<std::mutex> lk(g_tune_mu);
  while (fscanf(fp, " %c", &kind) == 1) {
    if (kind == 'F' || kind == 'B') {
      int f, le;
      unsigned ln;
      size_t np;
      if (fscanf(fp, "%d %u %d %zu", &f, &ln, &le, &np) != 4 || !field_ok(f) || np < 1 || np > 8) {
        st = OMR_E_ARG;
        break;
      }
      Schedule s;
      s.logE = le;
      int sum = 0;
      for (size_t i = 0; i < np; i++) {
        int r;
        if (fscanf(fp, "%d", &r) != 1 || r < 1 || r > max_log_radix(field_bytes(f))) {
          st = OMR_E_ARG;
          break;
        }
        s.logR.push_back(r);
        sum += r;
      }
      if (st != OMR_OK || sum != (int)ln) {
        st = OMR_E_ARG;
        break;
      }
      (kind == 'F' ? g_tuned : g_tuned_batch)[{f, ln}] = s;
    } else if (kind == 'P') {
      int f, m;
      unsigned ln, lo, l1;
      if (fscanf(fp, "%d %u %u %d %u", &f, &ln, &lo, &m, &l1) != 5 || !field_ok(f) || m < 1 || m > 3) {
        st = OMR_E_ARG;
        break;
      }
      g_tuned_partial[std::make_tuple(f, ln, lo)] = {m, l1};
    } else {
      st = OMR_E_ARG;
      break;
    }
  }
  fclose(fp);
  return st;
} catch (const std::bad_alloc &) {
  return OMR_E_ALLOC;
} catch (...) {
  return OMR_E_DEVICE;
}
