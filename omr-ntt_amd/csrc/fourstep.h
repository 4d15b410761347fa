// fourstep.h -- radix-4096 Stockham pass for Goldilocks as two 64-point
// register DFTs (the passes of a four-step 2^12 x 2^12 transform).
//
// Same pass semantics as engine.h (radix R = 4096, stride s, columns
// q = p s + j):
//   dst[k (N/R) + q] = sum_k' w_R^(k k') * w_(N/s)^(k' p) * src[(p R + k') s + j].
// With k' = u + 64 b and k = c1 + 64 d (u, b, c1, d < 64):
//   w_4096^(k k') = w_4096^(c1 u) w_64^(c1 b) w_64^(d u),
// so the pass is, per column: a 64-point DFT over b for each u, the twiddle
// w_4096^(c1 u), a 64-point DFT over u for each c1 (winterfell_four_step.rs:
// 31-81 takes the same split one level up, with transposes in memory).  A
// thread holds 64 elements of one column in registers, and the 64-point
// DFTs are radix-2 networks whose constants w_64^m are compile-time powers
// of 2 (field.h, diff_pow2), so the whole pass has one general product per
// element inside the tile (plus the inter-pass twiddle) and one LDS
// exchange, against three stages, two general products and two exchanges
// for a radix-4096 tile of 16-point register DFTs.
//
// A workgroup owns 8 consecutive columns (512 threads, thread = (c, u) with
// the column c fastest): every load and store instruction of a wave covers
// 8 columns x 8 rows, i.e. 64-byte runs.  Workgroups are mapped to tiles so
// that consecutive tiles (which share 128-byte lines) run on one XCD at the
// same time: the hardware deals workgroups to the 8 XCDs round robin.  The
// exchange between the two DFTs goes through LDS four columns at a time
// (128 KiB): a thread writes its 64 stage-1 outputs and reads its 64
// stage-2 inputs in the same round, into the same registers.
#pragma once
#include "engine.h"

namespace omr {

constexpr int kFsCols = 8;        // columns per tile
constexpr int kFsThreads = 512;   // kFsCols x 64
constexpr int kFsPadU = 65;       // LDS stride between the 64 values of one (column, c1)
constexpr int kFsColStride = 64 * kFsPadU + 4;  // LDS stride between the round's columns
constexpr int kFsLds = 4 * kFsColStride;        // elements (~130 KiB of Goldilocks words)
#ifndef OMR_FS_CHUNK
#define OMR_FS_CHUNK 8
#endif
constexpr int kFsChunk = OMR_FS_CHUNK;  // twiddle lookups in flight per thread

// Tile of workgroup b: consecutive tiles on one XCD (workgroup b runs on XCD
// b mod 8), when the grid divides evenly.
__device__ __forceinline__ uint32_t fs_tile_of(uint32_t b, uint32_t G) {
  return (G & 7u) ? b : (b & 7u) * (G >> 3) + (b >> 3);
}

// Inter-pass twiddle: none (the first pass) or power chains.
enum : int { kFsTwNone = 0, kFsTwChain = 1 };

template <int U, int EPI, int TW>
__device__ __forceinline__ void fs4096_body(const PassArgs &a) {
  using F = Goldilocks;
  using T = uint64_t;
  static_assert(U != 0, "the 64-point register DFTs need compile-time constants");
  __shared__ T lds[kFsLds];

  const uint32_t tid = threadIdx.x;
  const uint32_t c = tid & 7u, u = tid >> 3;
  const uint32_t logS = a.logS;
  const uint32_t logTiles = a.logN - 12 - 3;
  const uint64_t gt = fs_tile_of(blockIdx.x, gridDim.x);
  const uint64_t vec = gt >> logTiles;
  const uint32_t tile = (uint32_t)(gt & ((1ull << logTiles) - 1));
  const T *__restrict__ src = (const T *)a.src + (a.vec_off_in ? a.vec_off_in[vec] : vec * a.vec_stride_in);
  T *__restrict__ dst = (T *)a.dst + (a.vec_off_out ? a.vec_off_out[vec] : vec * a.vec_stride_out);
  const T *__restrict__ pyr = (const T *)a.root_pyr;

  // column q = p s + j; for s = 1 the 8 columns are 8 consecutive p
  const uint32_t q = (tile << 3) + c;
  const uint32_t p = q >> logS, j = q & ((1u << logS) - 1);
  const uint32_t pos0 = ((p << 12) << logS) + j;

  // In-vector byte offsets fit 32 bits (N <= 2^29 for this kernel): a
  // uniform base plus a 32-bit offset lets the loads and stores address
  // through the scalar base (global_* saddr form) with one VGPR offset each.
  const char *sbase = (const char *)src;
  char *dbase = (char *)dst;
  T x[64];
  __builtin_amdgcn_s_setprio(1);
#pragma unroll
  for (int b0 = 0; b0 < 64; b0 += kFsChunk) {
#pragma unroll
    for (int b = b0; b < b0 + kFsChunk; b++) {
      const uint32_t off = (pos0 + ((u + 64u * (uint32_t)b) << logS)) * 8u;
      x[b] = ld_stream((const T *)(sbase + off));
    }
    __builtin_amdgcn_sched_barrier(0);
  }
  __builtin_amdgcn_s_setprio(0);

  if (TW == kFsTwChain) {
    // w_(N/s)^((u + 64 b) p) = A B^b with A = w_(N/s)^(u p), B = w_(N/s)^(64 p),
    // from the two-level table of w_N (w_(N/s)^e = w_N^(e s)): four chains
    // b = r (mod 4) stepped by B^4, two products per element like a
    // two-level lookup, without 128 table loads per thread.  A scale
    // (kEpiScale: inverse transforms) rides in A, so every element gets it
    // once.
    const T *tlo = (const T *)a.tw_lo, *thi = (const T *)a.tw_hi;
    const uint32_t lomask = (1u << a.lo_bits) - 1;
    auto tw = [&](uint32_t e) { return F::wmul(tlo[e & lomask], thi[e >> a.lo_bits]); };  // w_N^e, e < N
    T A = tw((u * p) << logS);
    if (EPI == kEpiScale) A = F::wmul(A, (T)a.scale_w);
    const T B = tw((64u * p) << logS);
    const T B2 = F::wmul(B, B);
    const T B4 = F::wmul(B2, B2);
    T t[4] = {A, F::wmul(A, B), F::wmul(A, B2), F::wmul(F::wmul(A, B2), B)};
#pragma unroll
    for (int b = 0; b < 64; b += 4) {
#pragma unroll
      for (int r = 0; r < 4; r++) {
        x[b + r] = F::mulw(x[b + r], t[r]);
        if (b + 4 < 64) t[r] = F::wmul(t[r], B4);
      }
    }
  }

  T wE[33];
  load_consts<F, 6, U>(wE, nullptr, 0, 0);
  dft_reg<F, 6, 6, 0, U>(x, wE);  // x[c1] = sum_b w_64^(c1 b) in[u + 64 b]

  {  // w_4096^(c1 u), c1 u < 4096: level 12 of the root pyramid
    const T *r12 = pyr + 4096;
#pragma unroll
    for (int c0 = 0; c0 < 64; c0 += kFsChunk) {
#pragma unroll
      for (int c1 = c0; c1 < c0 + kFsChunk; c1++)
        if (c1 > 0) x[c1] = F::mulw(x[c1], r12[u * (uint32_t)c1]);
      __builtin_amdgcn_sched_barrier(0);
    }
  }

  // exchange: thread (c, u) holds Y_u[c1]; thread (c, k2 = u) needs Y_0..63[k2]
  const uint32_t cc = c & 3u;
#pragma unroll
  for (uint32_t r = 0; r < 2; r++) {
    __syncthreads();
    if ((c >> 2) == r) {
#pragma unroll
      for (int c1 = 0; c1 < 64; c1++) lds[cc * kFsColStride + (uint32_t)c1 * kFsPadU + u] = x[c1];
    }
    __syncthreads();
    if ((c >> 2) == r) {
#pragma unroll
      for (int uu = 0; uu < 64; uu++) x[uu] = lds[cc * kFsColStride + u * kFsPadU + (uint32_t)uu];
    }
  }

  dft_reg<F, 6, 6, 0, U>(x, wE);  // x[d] = out[k2 + 64 d]

  const uint32_t logNR = a.logN - 12;
  const uint32_t k2 = u;
  if (EPI == kEpiScale && TW == kFsTwNone) {  // (with a twiddle, the scale rode in A)
    const T sw = (T)a.scale_w;
#pragma unroll
    for (int d = 0; d < 64; d++) x[d] = F::mulw(x[d], sw);
  }
#pragma unroll
  for (int d0 = 0; d0 < 64; d0 += kFsChunk) {
#pragma unroll
    for (int d = d0; d < d0 + kFsChunk; d++) {
      const uint32_t off = (((k2 + 64u * (uint32_t)d) << logNR) + q) * 8u;
      *(T *)(dbase + off) = x[d];
    }
    __builtin_amdgcn_sched_barrier(0);
  }
}

template <int U, int EPI, int TW>
__global__ void __launch_bounds__(kFsThreads) k_fs4096(PassArgs a) {
  fs4096_body<U, EPI, TW>(a);
}

}  // namespace omr
