// dispatch.h -- host-side view of the compiled pass kernels.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "engine.h"

namespace omr {

// Largest pass radix compiled per element size (tile data <= kLdsBudget).
constexpr int max_log_radix(int esize) { return esize == 8 ? 12 : 13; }

// Register radices compiled for a pass of radix 2^logR.
inline bool has_pass_kernel(int logR, int logE, int esize) {
  if (logR < 1 || logR > max_log_radix(esize)) return false;
  if (logE < 1 || logE > 4 || logE > logR) return false;
  if (logE == 1) return logR <= 2;
  return logR - logE <= 10;  // one thread per E elements of a column: <= 1024 threads
}

// Columns per tile the kernel for (esize, logR, logE) was compiled for
// (engine.h: PassKernel::CMAX, wide tiles included).
inline int pass_cmax(int logR, int esize, int logE = 0) {
  const bool wide = esize == 4 && (logR == 10 || logR == 11) && logE == 4;
  int c = ((wide ? kLdsBudgetWide : kLdsBudget) / esize) >> logR;
  if (c < 1) c = 1;
  if (c > 32) c = 32;
  return c;
}

// Columns per tile actually launched: the compiled maximum, at most the
// number of columns, and at most 1024 threads per workgroup.
inline int pass_log_cols(int logN, int logR, int logE, int esize) {
  int lc = 0;
  while ((2 << lc) <= pass_cmax(logR, esize, logE)) lc++;
  if (lc > logN - logR) lc = logN - logR;
  if (lc > 10 - (logR - logE)) lc = 10 - (logR - logE);
  return lc < 0 ? 0 : lc;
}

// Grid of a pass kernel: one workgroup per tile.
unsigned pass_grid(const void *kernel, unsigned block, uint64_t total_tiles);

// Launches one Stockham pass over `total_tiles` tiles (a.nvec vectors);
// returns hipSuccess or the launch error.
// unit: 0, or (Goldilocks, N >= 64) 39 / 153 when the 64th root of the
// transform's root is 2^39 / 2^153 -- register-DFT constants become
// compile-time shifts.
hipError_t launch_pass(int field, int logR, int logE, int unit, const PassArgs &a, uint64_t total_tiles,
                       dim3 block, hipStream_t stream);

// The cheapest pass-kernel variant that honours every option set in `a`.
inline int pass_epilogue(const PassArgs &a) {
  if (a.flags & kPrefixSum) return kEpiPrefix;
  if (a.pre_lo || a.post_lo || a.nonzero_count || (a.flags & kZeroTail) || a.store_limit < (1ull << a.logN))
    return kEpiGeneral;
  if (a.flags & kPostTwiddle) return kEpiPost;
  return (a.flags & kScale) ? kEpiScale : kEpiPlain;
}

// Whether a 4-byte-field pass can carry the next pass's twiddle (kEpiPost
// is compiled for those fields at register radix 16, pass radix 2^6..2^11,
// full tiles only).
inline bool post4_ok(int logR, int logE, int logC) {
  return logE == 4 && logR >= 6 && logR <= 11 && (1 << logC) == pass_cmax(logR, 4, logE);
}

// Whether a kEpiGeneral pass is the first pass of a blowup-4 coset
// extension that kEpiLde runs: zero tail from N/4 on, an input multiplier,
// nothing else; register radix 16 with at least two in-tile stages; tiles
// whose rows are strided (s >= C).  The pass kernel's own column count
// check (compile-time C) comes on top.
inline bool lde4_ok(const PassArgs &a, int logR, int logE) {
  return logE == 4 && logR >= 5 && logR <= 12 && (a.flags & kZeroTail) && a.logN == a.logValid + 2 &&
         a.logS + (uint32_t)logR == a.logN && a.logS >= a.logC && a.pre_lo && !a.post_lo && !a.nonzero_count &&
         a.store_limit >= (1ull << a.logN) && !(a.flags & (kScale | kTwiddle | kPrefixSum));
}

// Goldilocks radix-4096 passes as two 64-point register DFTs (fourstep.h):
// whether a pass qualifies (forward / inverse ark roots, 2^15 <= N <= 2^29,
// 8 columns of one p per tile, plain or scaled epilogue), and its launch
// (total_tiles = 8-column tiles of all vectors, 512 threads each).
bool fs4096_ok(int field, int logR, int unit, uint32_t logN, uint32_t logS, int epi);
hipError_t launch_pass_fs4096(int unit, int epi, const PassArgs &a, uint64_t total_tiles, hipStream_t stream);

// Per-field entry points (one translation unit each).
hipError_t launch_pass_babybear(int logR, int logE, int epi, const PassArgs &a, uint64_t t, dim3 b,
                                hipStream_t s);
hipError_t launch_pass_koalabear(int logR, int logE, int epi, const PassArgs &a, uint64_t t, dim3 b,
                                 hipStream_t s);
hipError_t launch_pass_goldilocks_u0(int logR, int logE, int epi, const PassArgs &a, uint64_t t, dim3 b,
                                     hipStream_t s);
hipError_t launch_pass_goldilocks_u39(int logR, int logE, int epi, const PassArgs &a, uint64_t t, dim3 b,
                                      hipStream_t s);
hipError_t launch_pass_goldilocks_u153(int logR, int logE, int epi, const PassArgs &a, uint64_t t, dim3 b,
                                       hipStream_t s);
hipError_t launch_pass_teddybear(int logR, int logE, int epi, const PassArgs &a, uint64_t t, dim3 b,
                                 hipStream_t s);

}  // namespace omr
