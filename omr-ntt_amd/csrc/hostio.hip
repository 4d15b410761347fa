// hostio.hip -- host <-> device copies of the host-buffer entry points.
//
// A pageable host buffer is copied through a ring of pinned slots: host
// threads memcpy one chunk into a slot while the DMA engine moves the chunk
// before it, so the PCIe link sees pinned transfers.  Buffers that the
// caller allocated with omr_host_alloc are pinned already and go straight to
// hipMemcpy.  Copies below kStageMin bytes also take plain hipMemcpy, and
// so does every copy when OMR_HOST_STAGING=0.
#include <algorithm>
#include <new>
#include <cstdlib>
#include <cstring>
#include <map>
#include <thread>
#include <vector>

#include "internal.h"

namespace {

constexpr size_t kStageMin = 8ull << 20;  // smaller copies: plain hipMemcpy

std::mutex g_pin_mu;
std::map<uintptr_t, size_t> g_pinned;  // omr_host_alloc blocks: start -> bytes

bool is_pinned(const void *p, size_t bytes) {
  std::lock_guard<std::mutex> lk(g_pin_mu);
  const uintptr_t a = (uintptr_t)p;
  auto it = g_pinned.upper_bound(a);
  if (it == g_pinned.begin()) return false;
  --it;
  return a >= it->first && a + bytes <= it->first + it->second;
}

// OMR_HOST_STAGING=0 sends every copy through plain hipMemcpy (read per call)
bool staging_enabled() {
  const char *e = std::getenv("OMR_HOST_STAGING");
  return !(e && e[0] == '0' && e[1] == 0);
}

int copy_threads() {
  static const int n = [] {
    if (const char *e = std::getenv("OMR_HOST_THREADS")) {
      const int v = std::atoi(e);
      if (v >= 1 && v <= 64) return v;
    }
    const int hw = (int)std::thread::hardware_concurrency();
    return std::max(1, std::min(8, hw));
  }();
  return n;
}

// memcpy split over `threads` host threads in page-aligned pieces
void par_memcpy(void *dst, const void *src, size_t bytes, int threads) {
  const size_t piece = ((bytes + threads - 1) / threads + 4095) & ~(size_t)4095;
  if (threads <= 1 || piece >= bytes) {
    std::memcpy(dst, src, bytes);
    return;
  }
  // no exception may leave the C ABI: a piece whose thread cannot be
  // started is copied by the calling thread
  std::vector<std::thread> pool;
  size_t o = piece;
  try {
    pool.reserve((bytes + piece - 1) / piece);
    for (; o < bytes; o += piece) {
      const size_t l = std::min(piece, bytes - o);
      pool.emplace_back([=] { std::memcpy((char *)dst + o, (const char *)src + o, l); });
    }
  } catch (...) {
  }
  for (size_t r = o; r < bytes; r += piece)
    std::memcpy((char *)dst + r, (const char *)src + r, std::min(piece, bytes - r));
  std::memcpy(dst, src, piece);  // the calling thread takes the first piece
  for (auto &t : pool) t.join();
}

// Creates the ring on first use; the ring pointer is set last, so a partial
// failure is retried on the next call.
omr_status ensure_stage(omr::HostStage &s) {
  if (s.ring) return OMR_OK;
  if (!s.stream && hipStreamCreate(&s.stream) != hipSuccess) return s.stream = nullptr, OMR_E_DEVICE;
  for (auto &e : s.ev)
    if (!e && hipEventCreateWithFlags(&e, hipEventDisableTiming) != hipSuccess) return e = nullptr, OMR_E_DEVICE;
  void *ring = nullptr;
  if (hipHostMalloc(&ring, omr::HostStage::kSlots * omr::HostStage::kChunk, hipHostMallocDefault) != hipSuccess) {
    (void)hipGetLastError();
    return OMR_E_ALLOC;
  }
  s.ring = ring;
  return OMR_OK;
}

}  // namespace

namespace omr {

omr_status host_to_device(omr_ctx *ctx, void *dev, const void *host, size_t bytes) {
  if (bytes == 0) return OMR_OK;
  if (bytes < kStageMin || !staging_enabled() || is_pinned(host, bytes))
    return hipMemcpy(dev, host, bytes, hipMemcpyHostToDevice) == hipSuccess ? OMR_OK : OMR_E_DEVICE;
  HostStage &s = ctx->stage;
  std::lock_guard<std::mutex> lk(s.mu);
  omr_status st = ensure_stage(s);
  if (st != OMR_OK) return st;
  const int th = copy_threads();
  const size_t nch = (bytes + HostStage::kChunk - 1) / HostStage::kChunk;
  for (size_t i = 0; i < nch; i++) {
    const int k = (int)(i % HostStage::kSlots);
    char *slot = (char *)s.ring + k * HostStage::kChunk;
    const size_t o = i * HostStage::kChunk, l = std::min(HostStage::kChunk, bytes - o);
    if (i >= HostStage::kSlots && hipEventSynchronize(s.ev[k]) != hipSuccess) return OMR_E_DEVICE;
    par_memcpy(slot, (const char *)host + o, l, th);
    if (hipMemcpyAsync((char *)dev + o, slot, l, hipMemcpyHostToDevice, s.stream) != hipSuccess ||
        hipEventRecord(s.ev[k], s.stream) != hipSuccess)
      return OMR_E_DEVICE;
  }
  return hipStreamSynchronize(s.stream) == hipSuccess ? OMR_OK : OMR_E_DEVICE;
}

omr_status device_to_host(omr_ctx *ctx, void *host, const void *dev, size_t bytes) {
  if (bytes == 0) return OMR_OK;
  // the producer ran on the null stream: it completes before any copy starts
  if (hipStreamSynchronize(nullptr) != hipSuccess) return OMR_E_DEVICE;
  if (bytes < kStageMin || !staging_enabled() || is_pinned(host, bytes))
    return hipMemcpy(host, dev, bytes, hipMemcpyDeviceToHost) == hipSuccess ? OMR_OK : OMR_E_DEVICE;
  HostStage &s = ctx->stage;
  std::lock_guard<std::mutex> lk(s.mu);
  omr_status st = ensure_stage(s);
  if (st != OMR_OK) return st;
  const int th = copy_threads();
  const size_t nch = (bytes + HostStage::kChunk - 1) / HostStage::kChunk;
  auto issue = [&](size_t i) {
    const int k = (int)(i % HostStage::kSlots);
    const size_t o = i * HostStage::kChunk, l = std::min(HostStage::kChunk, bytes - o);
    return hipMemcpyAsync((char *)s.ring + k * HostStage::kChunk, (const char *)dev + o, l,
                          hipMemcpyDeviceToHost, s.stream) == hipSuccess &&
           hipEventRecord(s.ev[k], s.stream) == hipSuccess;
  };
  for (size_t i = 0; i < std::min<size_t>(nch, HostStage::kSlots); i++)
    if (!issue(i)) return OMR_E_DEVICE;
  for (size_t i = 0; i < nch; i++) {
    const int k = (int)(i % HostStage::kSlots);
    const size_t o = i * HostStage::kChunk, l = std::min(HostStage::kChunk, bytes - o);
    if (hipEventSynchronize(s.ev[k]) != hipSuccess) return OMR_E_DEVICE;
    par_memcpy((char *)host + o, (const char *)s.ring + k * HostStage::kChunk, l, th);
    if (i + HostStage::kSlots < nch && !issue(i + HostStage::kSlots)) return OMR_E_DEVICE;
  }
  return OMR_OK;
}

void free_stage(HostStage &s) {
  for (auto &e : s.ev)
    if (e) (void)hipEventDestroy(e), e = nullptr;
  if (s.stream) (void)hipStreamDestroy(s.stream), s.stream = nullptr;
  if (s.ring) (void)hipHostFree(s.ring), s.ring = nullptr;
}

}  // namespace omr

// Pinned host memory for callers that keep their buffers across calls.
extern "C" omr_status omr_host_alloc(size_t bytes, void **out) try {
  if (!out) return OMR_E_NULL;
  *out = nullptr;
  if (bytes == 0) return OMR_E_ARG;
  void *p = nullptr;
  if (hipHostMalloc(&p, bytes, hipHostMallocDefault) != hipSuccess || !p) {
    (void)hipGetLastError();
    return OMR_E_ALLOC;
  }
  try {
    std::lock_guard<std::mutex> lk(g_pin_mu);
    g_pinned[(uintptr_t)p] = bytes;
  } catch (...) {  // the registry could not grow
    (void)hipHostFree(p);
    return OMR_E_ALLOC;
  }
  *out = p;
  return OMR_OK;
} catch (const std::bad_alloc &) {
  return OMR_E_ALLOC;
} catch (...) {
  return OMR_E_DEVICE;
}

extern "C" omr_status omr_host_free(void *p) try {
  if (!p) return OMR_E_NULL;
  {
    std::lock_guard<std::mutex> lk(g_pin_mu);
    auto it = g_pinned.find((uintptr_t)p);
    if (it == g_pinned.end()) return OMR_E_ARG;  // not a block of omr_host_alloc
    g_pinned.erase(it);
  }
  return hipHostFree(p) == hipSuccess ? OMR_OK : OMR_E_DEVICE;
} catch (const std::bad_alloc &) {
  return OMR_E_ALLOC;
} catch (...) {
  return OMR_E_DEVICE;
}
