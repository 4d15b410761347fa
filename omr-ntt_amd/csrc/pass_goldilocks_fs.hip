// Radix-4096 passes of 64-point register DFTs for Goldilocks (fourstep.h):
// forward and inverse ark-convention roots, plain and scaled epilogues,
// with and without the inter-pass twiddle.
#include "dispatch.h"
#include "fourstep.h"

namespace omr {

bool fs4096_ok(int field, int logR, int unit, uint32_t logN, uint32_t logS, int epi) {
  return field == kGoldilocks && logR == 12 && (unit == 39 || unit == 153) && logN >= 15 && logN <= 29 &&
         (logS == 0 || logS >= 3) && (epi == kEpiPlain || epi == kEpiScale);
}

template <int U>
static const void *fs_kernel(int epi, bool twiddle) {
  if (epi == kEpiScale)
    return twiddle ? (const void *)k_fs4096<U, kEpiScale, kFsTwChain> : (const void *)k_fs4096<U, kEpiScale, kFsTwNone>;
  return twiddle ? (const void *)k_fs4096<U, kEpiPlain, kFsTwChain> : (const void *)k_fs4096<U, kEpiPlain, kFsTwNone>;
}

hipError_t launch_pass_fs4096(int unit, int epi, const PassArgs &a, uint64_t total_tiles, hipStream_t stream) {
  const bool tw = (a.flags & kTwiddle) != 0;
  const void *kern = unit == 39 ? fs_kernel<39>(epi, tw) : fs_kernel<153>(epi, tw);
  // one workgroup per 8 columns of every vector
  const uint64_t groups = total_tiles;
  if (groups == 0 || groups > 0x7FFFFFFFull) return hipErrorInvalidValue;
  PassArgs arg = a;
  void *args[] = {&arg};
  return hipLaunchKernel(kern, dim3((unsigned)groups), dim3(kFsThreads), args, 0, stream);
}

}  // namespace omr
