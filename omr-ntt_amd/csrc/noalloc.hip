// noalloc.hip -- full transforms that allocate nothing.
//
// omr_ntt_device_noalloc runs the same pass kernels as omr_ntt_device /
// omr_intt_device, but its intermediates (the ping-pong vectors of a
// multi-pass schedule) and its twiddle tables live in one file-scope device
// array, `scratch`, instead of buffers the context allocates on first use.
// It is for callers that must not allocate device memory (code under graph
// capture, tight serving loops, harnesses that forbid hipMalloc).
//
// Layout of `scratch` (byte offsets, each 256-aligned):
//   [0, 1 MiB)                twiddle tables of the last (field, N, root):
//                             lo @ 0, hi @ 64 KiB, root pyramid @ 128 KiB,
//                             mid table @ 256 KiB (2^16 entries)
//   [1 MiB, 1 MiB + 32 MiB)   intermediate vector set A
//   [33 MiB, 65 MiB)          intermediate vector set B
// Every call on a device shares that device's array and leaves its contents
// behind; the tables are rebuilt on the device (a kernel raises the root to
// each entry's exponent) whenever (field, N, root) differs from the previous
// call's, and the intermediates are fully written before they are read, so
// nothing needs clearing.  Calls that overlap in time on one device (other
// streams, other host threads) race on the array.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <mutex>
#include <new>

#include "../../include/omrntt.h"
#include "dispatch.h"
#include "internal.h"

namespace omr {

constexpr size_t kNaTableBytes = 1ull << 20;
constexpr size_t kNaVecBytes = (size_t)OMR_NOALLOC_MAX_BYTES;  // each intermediate set
constexpr size_t kNaBytes = kNaTableBytes + 2 * kNaVecBytes;
constexpr size_t kNaLo = 0, kNaHi = 64ull << 10, kNaPyr = 128ull << 10, kNaMid = 256ull << 10;

static __device__ __align__(256) unsigned char scratch[kNaBytes];

// Host-side view of each device's copy of `scratch`: its device address
// (from hipGetSymbolAddress with that device current) and which tables it
// holds.
struct NaDevice {
  bool looked_up = false;
  void *ptr = nullptr;  // nullptr: the lookup failed
  bool tables_valid = false;
  int field = -1;
  uint32_t logN = 0;
  uint64_t root = 0;
};
constexpr int kNaMaxDevices = 64;
static std::mutex g_na_mu;
static NaDevice g_na[kNaMaxDevices];

void noalloc_init_device(int device) {
  if (device < 0 || device >= kNaMaxDevices) return;
  std::lock_guard<std::mutex> lk(g_na_mu);
  NaDevice &d = g_na[device];
  if (d.looked_up) return;
  d.looked_up = true;
  void *p = nullptr;
  if (hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)) != hipSuccess) {
    (void)hipGetLastError();
    p = nullptr;
  }
  d.ptr = p;
}

// out[i] = root^e(i) in multiplier form, i < count.  kind 0: e = i << shift;
// kind 1 (root pyramid): e(0) = 0, e(2^L + m) = m << (logN - L).
template <class F>
__global__ void __launch_bounds__(256) k_root_powers(typename F::T *__restrict__ out, uint32_t count,
                                                     typename F::T root, uint32_t logN, uint32_t kind,
                                                     uint32_t shift) {
  using T = typename F::T;
  const uint32_t i = blockIdx.x * 256 + threadIdx.x;
  if (i >= count) return;
  uint64_t e;
  if (kind == 0) {
    e = (uint64_t)i << shift;
  } else if (i == 0) {
    e = 0;
  } else {
    const uint32_t L = 31 - __builtin_clz(i);
    e = (uint64_t)(i - (1u << L)) << (logN - L);
  }
  T r = 1, a = root;  // canonical
  while (e) {
    if (e & 1) r = F::mul(r, a);
    a = F::mul(a, a);
    e >>= 1;
  }
  out[i] = F::to_w(r);
}

template <class F>
static omr_status fill_powers(void *out, uint32_t count, uint64_t root, uint32_t logN, uint32_t kind,
                              uint32_t shift, hipStream_t s) {
  using T = typename F::T;
  hipLaunchKernelGGL(k_root_powers<F>, dim3((count + 255) / 256), dim3(256), 0, s, (T *)out, count, (T)root, logN,
                     kind, shift);
  return hipGetLastError() == hipSuccess ? OMR_OK : OMR_E_DEVICE;
}

static omr_status fill_table(int f, void *out, uint32_t count, uint64_t root, uint32_t logN, uint32_t kind,
                             uint32_t shift, hipStream_t s) {
  switch (f) {
    case kBabyBear: return fill_powers<BabyBear>(out, count, root, logN, kind, shift, s);
    case kKoalaBear: return fill_powers<KoalaBear>(out, count, root, logN, kind, shift, s);
    case kGoldilocks: return fill_powers<Goldilocks>(out, count, root, logN, kind, shift, s);
    case kTeddyBear: return fill_powers<TeddyBear>(out, count, root, logN, kind, shift, s);
  }
  return OMR_E_FIELD;
}

// Table layout of get_tables (api.hip) for (f, logN, root), placed in the
// table region at `base`; fills it on stream s unless it already holds it.
static omr_status na_tables(NaDevice &d, int f, uint32_t logN, uint64_t root, hipStream_t s, TwTables &t) {
  const int es = field_bytes(f);
  unsigned char *base = (unsigned char *)d.ptr;
  t = TwTables{};
  t.lo_bits = std::min<uint32_t>(logN, 12);
  t.log_rmax = std::min<uint32_t>(logN, (uint32_t)max_log_radix(es));
  t.log_mid = std::min<uint32_t>(logN, (uint32_t)kLogMid);
  t.lo = base + kNaLo;
  t.hi = base + kNaHi;
  t.pyr = base + kNaPyr;
  t.mid = nullptr;
  if (t.log_mid > t.log_rmax) t.mid = base + kNaMid;
  else t.log_mid = t.log_rmax;
  pow2_unit_of(f, logN, root, t);
  // the regions hold every size this entry point accepts
  const uint64_t nlo = 1ull << t.lo_bits, nhi = 1ull << (logN - t.lo_bits), npyr = 2ull << t.log_rmax;
  if (nlo * es > kNaHi - kNaLo || nhi * es > kNaPyr - kNaHi || npyr * es > kNaMid - kNaPyr ||
      (t.mid && (1ull << t.log_mid) * es > kNaTableBytes - kNaMid))
    return OMR_E_SCRATCH;
  if (d.tables_valid && d.field == f && d.logN == logN && d.root == root) return OMR_OK;
  d.tables_valid = false;
  omr_status st;
  if ((st = fill_table(f, t.lo, (uint32_t)nlo, root, logN, 0, 0, s)) != OMR_OK) return st;
  if ((st = fill_table(f, t.hi, (uint32_t)nhi, root, logN, 0, t.lo_bits, s)) != OMR_OK) return st;
  if ((st = fill_table(f, t.pyr, (uint32_t)npyr, root, logN, 1, 0, s)) != OMR_OK) return st;
  if (t.mid && (st = fill_table(f, t.mid, 1u << t.log_mid, root, logN, 0, logN - t.log_mid, s)) != OMR_OK)
    return st;
  d.tables_valid = true;
  d.field = f;
  d.logN = logN;
  d.root = root;
  return OMR_OK;
}

}  // namespace omr

using namespace omr;

extern "C" omr_status omr_ntt_device_noalloc(omr_ctx *ctx, int field, int variant, const void *d_in, void *d_out,
                                             uint64_t n, uint64_t batch, uint64_t omega, int inverse,
                                             void *stream) try {
  if (variant < 0 || variant >= OMR_VARIANT_COUNT) return OMR_E_VARIANT;
  if (variant == OMR_DFT || variant == OMR_DFT_PARTIAL || variant == OMR_TRANSFORM_DECOMPOSITION ||
      variant == OMR_TRANSFORM_DECOMPOSITION_V2)
    return OMR_E_VARIANT;  // full transforms of the pass engine only
  omr_status st = validate_domain(field, n, omega);
  if (st != OMR_OK) return st;
  const uint32_t logN = 63 - __builtin_clzll(n);
  if (!omr_variant_supports(variant, logN)) return OMR_E_RADIX;
  if (batch > kMaxBatch) return OMR_E_ARG;
  if (!ctx || !d_in || !d_out) return OMR_E_NULL;
  const int es = field_bytes(field);
  if (batch == 0) return OMR_OK;
  if (n * batch * es > kNaVecBytes) return OMR_E_SCRATCH;
  if (ctx->device < 0 || ctx->device >= kNaMaxDevices) return OMR_E_DEVICE;
  if (hipSetDevice(ctx->device) != hipSuccess) return OMR_E_DEVICE;
  std::lock_guard<std::mutex> lk(g_na_mu);
  NaDevice &d = g_na[ctx->device];
  if (!d.ptr) return OMR_E_DEVICE;  // the symbol lookup at omr_ctx_create failed
  Schedule sch;
  if ((st = variant_schedule(field, variant, logN, sch)) != OMR_OK) return st;
  if (variant == OMR_ARK_RADIX2 && !tuned_schedule_for(field, logN, batch, sch)) default_schedule(field, logN, sch);
  XformOpts o;
  uint64_t root = omega;
  if (inverse) {
    root = finv_host(field, omega);
    o.scale = true;
    o.scale_w = to_w_host(field, finv_host(field, n % field_modulus(field)));
  }
  hipStream_t hs = (hipStream_t)stream;
  TwTables t;
  if (logN > 0 && (st = na_tables(d, field, logN, root, hs, t)) != OMR_OK) return st;
  o.tables = &t;
  o.scratch[0] = (unsigned char *)d.ptr + kNaTableBytes;
  o.scratch[1] = (unsigned char *)d.ptr + kNaTableBytes + kNaVecBytes;
  return run_transform(ctx, field, d_in, d_out, logN, root, sch, batch, n, n, o, hs);
} catch (const std::bad_alloc &) {
  return OMR_E_ALLOC;
} catch (...) {
  return OMR_E_DEVICE;
}
