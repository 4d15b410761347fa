// engine.h -- LDS-fused Stockham pass kernel for gfx950.
//
// An N-point NTT (N = 2^logN, natural order in and out) runs as a short
// sequence of HBM passes.  Pass i has radix R = 2^LOGR, stride s and
// L = N / (R s) groups; it maps
//   dst[k (N/R) + p s + j] = sum_k' w_R^(k k') * w_N^(k' p s) * src[(p R + k') s + j]
// for p < L, j < s, k < R: the Stockham autosort step that the reference's
// radix-8 encoder takes per radix-2^3 layer group
// (src/encoders/tfhe_stockham_radix8.rs:80-139), here with R up to 2^12 so
// that a 2^24 transform needs 2 or 3 passes instead of 8.
//
// One workgroup owns a tile of C consecutive "columns" q = p s + j (all R
// rows k).  Stores are always C-element contiguous runs; loads are C-element
// runs when s >= C, and one contiguous block of C*R elements otherwise (staged
// through LDS).  Inside the tile the R-point DFT is itself a Stockham sequence
// of register radix-E DFTs with LDS exchanges between them.
#pragma once
#include <type_traits>

#include "field.h"

namespace omr {

struct PassArgs {
  const void *src;
  void *dst;
  const void *tw_lo;   // w_N^e, e < 2^lo_bits                     (multiplier form)
  const void *tw_hi;   // w_N^(e 2^lo_bits), e < N / 2^lo_bits     (multiplier form)
  const void *pre_lo;  // optional input-position multiplier h^i:   h^(i mod 2^lo_bits)
  const void *pre_hi;  //                                            h^((i >> lo_bits) 2^lo_bits)
  const void *post_lo;  // optional output-position multiplier, same two-level layout
  const void *post_hi;
  uint32_t *nonzero_count;  // if set: outputs at index >= store_limit that are nonzero are counted
  uint64_t vec_stride_in;   // elements between consecutive input vectors
  uint64_t vec_stride_out;  // elements between consecutive output vectors
  const uint64_t *vec_off_in;   // if set: vector v's input starts at element vec_off_in[v]
  const uint64_t *vec_off_out;  // if set: vector v's output starts at element vec_off_out[v]
  uint64_t scale_w;         // multiplier applied to every output when kScale
  uint64_t store_limit;     // only outputs with index < store_limit are stored
  uint32_t logN, logS, logC, lo_bits;
  uint32_t flags;     // kTwiddle | kScale | kZeroTail
  uint32_t logValid;  // with kZeroTail: input positions >= 2^logValid read as zero
  uint32_t nvec;              // vectors (consecutive, vec_stride_in/out apart)
  const void *tw_mid;         // w_N^(j N/2^log_mid), j < 2^log_mid (single-lookup twiddles)
  uint32_t log_rmax;          // largest level of the root pyramid
  uint32_t log_mid;
  const void *root_pyr;  // root pyramid: w_(2^L)^m at [2^L + m], m < 2^L, L <= log_rmax (multiplier form)
  uint32_t post_logR;    // with kPostTwiddle: log radix of the next pass
};
// kWideIn / kWideOut (31/32-bit fields): the pass reads / writes its vectors
// as 64-bit words, one canonical value in the low half of each (a one-limb
// big-integer layout), narrowing on load and widening on store.
// kPostTwiddle: every output times the inter-pass twiddle that the next
// pass (radix 2^post_logR, stride s / 2^post_logR) would apply to it; that
// pass then runs without kTwiddle.  With kScale as well, the outputs are
// also scaled (the last such pass of an inverse carries its N^-1).
enum : uint32_t {
  kTwiddle = 1u, kScale = 2u, kZeroTail = 4u, kPrefixSum = 8u, kWideIn = 16u, kWideOut = 32u, kPostTwiddle = 64u
};

// Pass-kernel variants by the options they honour (chosen per launch): the
// options' per-element tests, if compiled in, cost VALU and split the store
// loop into small blocks the compiler cannot interleave.
//   kEpiPlain   -- kTwiddle only
//   kEpiScale   -- kTwiddle, every output times scale_w (inverse transforms)
//   kEpiGeneral -- everything: zero tail, pre/post multipliers, scale,
//                  store_limit pruning and the nonzero count
//   kEpiPrefix  -- (kPrefixSum, the last pass run of a prefix-partial
//                  transform, s >= C) output k' < store_limit of column
//                  j = q mod s times w_N^(j k'), summed over the tile's
//                  columns: one partial sum per (k', tile) instead of the
//                  pass output
//   kEpiLde     -- the first pass of a blowup-4 coset extension (zero tail
//                  from N/4 on, input multiplier h^pos, plain store) at
//                  register radix 16: a thread's rows t + m R/16 are zero
//                  for m >= 4, so it loads and shifts 4 rows and its first
//                  register DFT skips the two butterfly levels that only
//                  see zeros
//   kEpiPost    -- kTwiddle and kPostTwiddle (a pass before the last)
enum : int { kEpiPlain = 0, kEpiScale = 1, kEpiGeneral = 2, kEpiPrefix = 3, kEpiLde = 4, kEpiPost = 5 };
constexpr int kPrefixMaxLogR = 10;  // radices compiled with the prefix epilogue

template <int LOGX>
__device__ __forceinline__ constexpr int brev_c(int i) {
  int r = 0;
  for (int b = 0; b < LOGX; b++) r |= ((i >> b) & 1) << (LOGX - 1 - b);
  return r;
}

// Butterfly difference times a register-DFT constant, (u - v) w_E^idx.
// U != 0 (fields with power-of-2 roots, N >= 64): w_64 = 2^U, so w_E^idx =
// 2^((U idx 64/E) mod 192) is a compile-time power of 2 once the DFT loops
// are unrolled, and its sign goes into the order of the subtraction.
// U == 0: c is the multiplier-form constant.
template <class F, int U, int LOGE>
__device__ __forceinline__ typename F::T diff_mul(typename F::T u, typename F::T v, typename F::T c, int idx) {
  if constexpr (U != 0) return F::diff_pow2(u, v, (uint32_t)((U * idx * (64 >> LOGE)) % 192));
  else return F::diff_mulw(u, v, c);
}

// Register-DFT constants w_E^k, k <= E/2: multiplier-form values from the
// root table (unused when the constants are compile-time shifts, U != 0).
template <class F, int LOGE, int U = 0>
__device__ __forceinline__ void load_consts(typename F::T (&wE)[(1 << LOGE) / 2 + 1], const typename F::T *root,
                                            uint32_t RE, uint32_t root_shift) {
#pragma unroll
  for (int k = 0; k <= (1 << LOGE) / 2; k++) {
    if constexpr (U != 0)
      wE[k] = 0;
    else
      wE[k] = root[((uint32_t)k * RE) << root_shift];
  }
}

// In-register DFT of size X = 2^LOGX on x[OFF .. OFF+X), natural order in
// and out.  wE[k] = w_E^k for the register radix E = 2^LOGE; the sub-DFT
// uses w_X = w_E^(E/X).  Radix-2 decimation in frequency followed by a
// compile-time register permutation (free).  LEN0 < X: the butterfly
// levels of length > LEN0 are already done (dft_reg_nz4).
template <class F, int LOGE, int LOGX, int OFF, int U = 0, int LEN0 = (1 << LOGX)>
__device__ __forceinline__ void dft_reg(typename F::T (&x)[1 << LOGE],
                                        const typename F::T (&wE)[(1 << LOGE) / 2 + 1]) {
  using T = typename F::T;
  constexpr int X = 1 << LOGX;
  constexpr int E = 1 << LOGE;
#pragma unroll
  for (int len = LEN0; len >= 2; len >>= 1) {
    const int half = len >> 1;
#pragma unroll
    for (int b = 0; b < X; b += len) {
#pragma unroll
      for (int j = 0; j < half; j++) {
        T u = x[OFF + b + j], v = x[OFF + b + j + half];
        x[OFF + b + j] = F::add(u, v);
        x[OFF + b + j + half] =
            (j == 0) ? F::sub(u, v) : diff_mul<F, U, LOGE>(u, v, wE[j * (E / len)], j * (E / len));
      }
    }
  }
  T y[X];
#pragma unroll
  for (int i = 0; i < X; i++) y[brev_c<LOGX>(i)] = x[OFF + i];
#pragma unroll
  for (int i = 0; i < X; i++) x[OFF + i] = y[i];
}

// dft_reg of the full register radix E = 2^LOGE >= 16 when x[m] = 0 for
// m >= E/4 (canonical x[m] for m < E/4): the first level's differences are
// u w^j (v = 0) and its sums u, and the second level sees a zero v again;
// only the twiddle products of those two levels remain.  Same output as
// dft_reg on the zero-padded input.
template <class F, int LOGE, int U = 0>
__device__ __forceinline__ void dft_reg_nz4(typename F::T (&x)[1 << LOGE],
                                            const typename F::T (&wE)[(1 << LOGE) / 2 + 1]) {
  using T = typename F::T;
  constexpr int E = 1 << LOGE, Q = E / 4;
  static_assert(LOGE >= 2, "register radix of at least 4");
  const T zero = 0;
#pragma unroll
  for (int j = 0; j < Q; j++)  // level len = E: (u, 0) -> (u, u w_E^j)
    x[j + 2 * Q] = (j == 0) ? x[j] : diff_mul<F, U, LOGE>(x[j], zero, wE[j], j);
#pragma unroll
  for (int b = 0; b < E; b += 2 * Q) {  // level len = E/2: (u, 0) -> (u, u w_(E/2)^j)
#pragma unroll
    for (int j = 0; j < Q; j++)
      x[b + j + Q] = (j == 0) ? x[b + j] : diff_mul<F, U, LOGE>(x[b + j], zero, wE[2 * j], 2 * j);
  }
  dft_reg<F, LOGE, LOGE, 0, U, E / 4>(x, wE);
}

// E / X independent sub-DFTs of size X = 2^LOGX on consecutive register groups.
template <class F, int LOGE, int LOGX, int G, int U = 0>
struct GroupDft {
  __device__ __forceinline__ static void run(typename F::T (&x)[1 << LOGE],
                                             const typename F::T (&wE)[(1 << LOGE) / 2 + 1]) {
    GroupDft<F, LOGE, LOGX, G - 1, U>::run(x, wE);
    dft_reg<F, LOGE, LOGX, (G - 1) << LOGX, U>(x, wE);
  }
};
template <class F, int LOGE, int LOGX, int U>
struct GroupDft<F, LOGE, LOGX, 0, U> {
  __device__ __forceinline__ static void run(typename F::T (&)[1 << LOGE],
                                             const typename F::T (&)[(1 << LOGE) / 2 + 1]) {}
};

#ifndef OMR_LDS_BUDGET
#define OMR_LDS_BUDGET 32768
#endif
constexpr int kLdsBudget = OMR_LDS_BUDGET;  // bytes of tile data per workgroup

// Wide tiles: 4-byte fields at radix 2^10 / 2^11 get 64 KiB of tile data
// (16 / 8 columns, 64 / 32-byte runs) instead of 8 / 4 columns, so that a
// 2^22 transform runs in two HBM passes of radix 2^11 with runs the L2
// merges (tiles in XCD order, see xcd_tile); 1024 threads at register radix
// 16.
constexpr int kLdsBudgetWide = 65536;
template <class T, int LOGR, int LOGE>
constexpr bool wide_tile() {
  return sizeof(T) == 4 && (LOGR == 10 || LOGR == 11) && LOGE == 4;
}

// Tile of workgroup b in a grid of G when consecutive tiles share cache
// lines: the hardware deals workgroups to the 8 XCDs round robin (b mod 8),
// so tile (b mod 8) G/8 + b/8 puts consecutive tiles on one XCD at about
// the same time; the grid must be a multiple of 8, else tiles stay in order.
__device__ __forceinline__ uint32_t xcd_tile(uint32_t b, uint32_t G) {
  return (G & 7u) ? b : (b & 7u) * (G >> 3) + (b >> 3);
}

template <class F, int LOGR, int LOGE>
struct PassKernel {
  using T = typename F::T;
  static constexpr int R = 1 << LOGR;
  static constexpr int E = 1 << LOGE;
  static constexpr int RE = R / E;
  static constexpr int NSTAGE = (LOGR + LOGE - 1) / LOGE;
  static constexpr int LOGLAST = LOGR - (NSTAGE - 1) * LOGE;  // radix of the last in-tile stage
  static constexpr bool WIDE = wide_tile<T, LOGR, LOGE>();
  static constexpr int CMAX0 = ((WIDE ? kLdsBudgetWide : kLdsBudget) / (int)sizeof(T)) >> LOGR;
  static constexpr int CMAX = CMAX0 < 1 ? 1 : (CMAX0 > 32 ? 32 : CMAX0);
  static constexpr int MAXT = CMAX * RE > 1024 ? 1024 : CMAX * RE;  // hardware limit
  static constexpr int LDS_N = CMAX * R;
};

// Padded LDS index: one spare slot every 32 elements spreads the power-of-two
// strides of the late in-tile stages over the banks.
#ifndef OMR_LDS_PAD
#define OMR_LDS_PAD 1
#endif
__device__ __forceinline__ uint32_t lpad(uint32_t i) { return OMR_LDS_PAD ? i + (i >> 5) : i; }
constexpr int lds_padded(int n) { return OMR_LDS_PAD ? n + n / 32 + 1 : n; }
// lpad((row << logC) + col) for col < 2^logC.  FOLD (logC known at compile
// time): split so that the row term is linear, and a row step of a
// compile-time amount becomes an instruction offset instead of a
// per-element shift.  With a run-time logC the split form only costs more.
template <bool FOLD>
__device__ __forceinline__ uint32_t lrc(uint32_t row, uint32_t col, uint32_t logC) {
  if (!OMR_LDS_PAD) return (row << logC) + col;
  if constexpr (!FOLD) return lpad((row << logC) + col);
  if (logC > 5) return (row << logC) + (row << (logC - 5)) + col + (col >> 5);
  if (logC == 5) return (row << 5) + row + col;  // col < 32
  return (row << logC) + (row >> (5 - logC)) + col;
}

// In-tile stage ST >= 1 of the R-point Stockham: reads the previous stage's
// outputs from LDS, applies w_R^(k' p s), runs the register DFTs.  The
// previous stage's registers are written by thread (tw, cw) -- its own
// (t, c), or for stage 1 of a row-fast pass the load mapping's.
template <class F, int LOGR, int LOGE, int ST, int U = 0, bool FOLD = false>
__device__ __forceinline__ void tile_stage(typename F::T (&x)[1 << LOGE],
                                           const typename F::T (&wE)[(1 << LOGE) / 2 + 1],
                                           typename F::T *lds, const typename F::T *root,
                                           uint32_t t, uint32_t c, uint32_t logC,
                                           uint32_t root_shift, uint32_t tw, uint32_t cw) {
  using K = PassKernel<F, LOGR, LOGE>;
  using T = typename F::T;
  if constexpr (ST < K::NSTAGE) {
    constexpr bool LAST = (ST == K::NSTAGE - 1);
    constexpr int LOGEI = LAST ? K::LOGLAST : LOGE;
    constexpr int EI = 1 << LOGEI;
    constexpr int LOGS = LOGR - ST * LOGE - LOGEI;  // in-tile stride of this stage
    __syncthreads();
#pragma unroll
    for (int m = 0; m < K::E; m++) lds[lrc<FOLD>(tw + m * K::RE, cw, logC)] = x[m];
    __syncthreads();
#pragma unroll
    for (int g = 0; g < K::E / EI; g++) {
      const uint32_t tt = t + g * K::RE;
      const uint32_t ps = (tt >> LOGS) << LOGS;  // p * s within the tile
      const uint32_t jj = tt & ((1u << LOGS) - 1);
#pragma unroll
      for (int kk = 0; kk < EI; kk++) {
        const uint32_t row = ps * EI + ((uint32_t)kk << LOGS) + jj;
        T v = lds[lrc<FOLD>(row, c, logC)];
        const uint32_t e = (uint32_t)kk * ps;
        if (kk != 0) v = F::mulw(v, root[e << root_shift]);  // e = 0: times root[0] = 1
        x[g * EI + kk] = v;
      }
    }
    GroupDft<F, LOGE, LOGEI, K::E / EI, U>::run(x, wE);
    tile_stage<F, LOGR, LOGE, ST + 1, U, FOLD>(x, wE, lds, root, t, c, logC, root_shift, t, c);
  }
}

// Raw loads of one tile into registers (no arithmetic).  Direct path
// (s >= C): the thread's E rows of its column.  Staged path (s < C): the
// thread's E elements tid + m*nthr of the tile's contiguous C*R block,
// transposed through LDS later.  Offsets inside a vector are 32-bit
// (N <= 2^32); the zero-tail test runs only when the flag is set.
// Tile loads are read-once streams: the nontemporal hint keeps them from
// displacing the twiddle and root tables in the caches.  (Measured on
// MI355X, 2^24 forward with the issue priority below: Goldilocks -4 %,
// BabyBear -5 %; nontemporal stores of the pass output were slower.)
template <class T>
__device__ __forceinline__ T ld_stream(const T *p) {
  return __builtin_nontemporal_load(p);
}

// Raw buffer access to one vector of less than 2^32 bytes (the descriptor's
// record count is 32-bit and an access past it reads 0 and drops the store,
// so a 2^32-byte vector keeps 64-bit addresses): a descriptor over
// the vector, a 32-bit per-thread byte offset, and the uniform part of each
// row's offset in an SGPR (the hardware adds the two).  A thread's E rows
// then cost no per-element 64-bit address arithmetic and their offsets one
// VGPR between them instead of two per row.  (gfx9 descriptor word 3:
// 32-bit data format, no swizzle; raw buffers are byte-addressed.)
typedef unsigned int omr_u32x2 __attribute__((__vector_size__(8)));
__device__ __forceinline__ __amdgpu_buffer_rsrc_t vec_rsrc(const void *base, uint32_t bytes) {
  return __builtin_amdgcn_make_buffer_rsrc(const_cast<void *>(base), (short)0, (int)bytes, 0x00020000);
}
// NT: the nontemporal hint of ld_stream (cache-policy bit 1 on gfx9)
template <class T, bool NT>
__device__ __forceinline__ T buf_load(__amdgpu_buffer_rsrc_t r, uint32_t voff, uint32_t soff) {
  constexpr int aux = NT ? 2 : 0;
  if constexpr (sizeof(T) == 4) {
    return (T)__builtin_amdgcn_raw_buffer_load_b32(r, (int)voff, (int)soff, aux);
  } else {
    const omr_u32x2 v = __builtin_amdgcn_raw_buffer_load_b64(r, (int)voff, (int)soff, aux);
    return ((T)v[1] << 32) | (T)v[0];
  }
}
template <class T>
__device__ __forceinline__ void buf_store(T v, __amdgpu_buffer_rsrc_t r, uint32_t voff, uint32_t soff) {
  if constexpr (sizeof(T) == 4) {
    __builtin_amdgcn_raw_buffer_store_b32((uint32_t)v, r, (int)voff, (int)soff, 0);
  } else {
    const omr_u32x2 w = {(uint32_t)v, (uint32_t)((uint64_t)v >> 32)};
    __builtin_amdgcn_raw_buffer_store_b64(w, r, (int)voff, (int)soff, 0);
  }
}

template <class F, int LOGR, int LOGE, int EPI>
__device__ __forceinline__ void load_tile(typename F::T (&xn)[1 << LOGE], const typename F::T *__restrict__ src,
                                          const uint64_t *__restrict__ srcw, bool wide, const PassArgs &a,
                                          uint32_t tile, uint32_t tid, uint32_t c, uint32_t t, bool staged,
                                          uint32_t logC, bool buf, __amdgpu_buffer_rsrc_t rs) {
  using K = PassKernel<F, LOGR, LOGE>;
  using T = typename F::T;
  const uint32_t logS = a.logS;
  uint32_t pos0, step;
  if (!staged) {
    const uint32_t q = (tile << logC) + c;
    pos0 = (((q >> logS) << LOGR) << logS) + (q & ((1u << logS) - 1)) + (t << logS);
    step = K::RE << logS;
  } else {
    pos0 = (tile << (logC + LOGR)) + tid;
    step = (1u << logC) * K::RE;
  }
  if constexpr (EPI == kEpiLde) {  // rows m >= E/4 lie in the zero tail
#pragma unroll
    for (int m = 0; m < K::E; m++) {
      if (m < K::E / 4) {
        const uint32_t pos = pos0 + (uint32_t)m * step;
        if (buf && !wide) xn[m] = buf_load<T, true>(rs, pos0 * (uint32_t)sizeof(T), (uint32_t)m * step * (uint32_t)sizeof(T));
        else if constexpr (sizeof(T) == 4) xn[m] = wide ? (T)ld_stream(srcw + pos) : ld_stream(src + pos);
        else xn[m] = ld_stream(src + pos);
      } else {
        xn[m] = 0;
      }
    }
    return;
  }
  if constexpr (sizeof(T) == 4) {
    if (wide) {  // 64-bit words: the value is the low half
      if (EPI == kEpiGeneral && (a.flags & kZeroTail)) {
        const uint64_t valid = 1ull << a.logValid;
#pragma unroll
        for (int m = 0; m < K::E; m++) {
          const uint32_t pos = pos0 + (uint32_t)m * step;
          xn[m] = ((uint64_t)pos < valid) ? (T)ld_stream(srcw + pos) : (T)0;
        }
      } else {
#pragma unroll
        for (int m = 0; m < K::E; m++) xn[m] = (T)ld_stream(srcw + pos0 + (uint32_t)m * step);
      }
      return;
    }
  }
  if (EPI == kEpiGeneral && (a.flags & kZeroTail)) {
    const uint64_t valid = 1ull << a.logValid;
#pragma unroll
    for (int m = 0; m < K::E; m++) {
      const uint32_t pos = pos0 + (uint32_t)m * step;
      xn[m] = ((uint64_t)pos < valid) ? ld_stream(src + pos) : (T)0;
    }
  } else if (buf) {
#pragma unroll
    for (int m = 0; m < K::E; m++)
      xn[m] = buf_load<T, true>(rs, pos0 * (uint32_t)sizeof(T), (uint32_t)m * step * (uint32_t)sizeof(T));
  } else {
#pragma unroll
    for (int m = 0; m < K::E; m++) xn[m] = ld_stream(src + pos0 + (uint32_t)m * step);
  }
}

// Pass kernel: workgroup b owns tile b of the a.nvec vectors' tiles.  (A
// grid-stride loop over several tiles per workgroup made the compiler keep
// per-tile invariants live across iterations: 2-3x the VGPRs and a quarter
// of the occupancy, measured slower.)
// FIXC: the tile has the full K::CMAX columns (a.logC == log2 CMAX, the case
// of every pass over a vector of N >= R CMAX), known at compile time, so the
// LDS exchange addresses fold into instruction offsets.
template <class F, int LOGR, int LOGE, int U, int EPI, bool FIXC>
__device__ __forceinline__ void stockham_pass_body(const PassArgs &a) {
  using K = PassKernel<F, LOGR, LOGE>;
  using T = typename F::T;
  constexpr int R = K::R, E = K::E, RE = K::RE;
  __shared__ T lds[lds_padded(K::LDS_N)];

  const uint32_t logC = FIXC ? (uint32_t)__builtin_ctz(K::CMAX) : a.logC;
  // 4-byte fields in full narrow tiles (C >= 16 columns, compile-time logC):
  // LDS indices folded into instruction offsets, buffer loads and stores
  constexpr bool kNarrow4 = sizeof(T) == 4 && FIXC && !K::WIDE;
  const uint32_t C = 1u << logC;
  const uint32_t tid = threadIdx.x;
  const uint32_t c = tid & (C - 1), t = tid >> logC;
  const uint32_t logS = a.logS;
  const uint32_t S = 1u << logS;  // s <= N/R <= 2^31
  const bool staged = S < C;
  // A pass with s = 1 and at least two in-tile stages loads with the row
  // fastest (thread (t0, c0) = (tid mod R/E, tid div R/E)): its columns are
  // contiguous rows of R elements, so those loads are coalesced as they are,
  // and stage 1's exchange through LDS hands the values to the column-fast
  // mapping the stores need -- no staging transpose.  Other passes with
  // s < C stage a contiguous block through LDS.  Goldilocks only: with
  // 4-byte words the row-fast exchange writes conflict on the LDS banks and
  // the pass got slower (BabyBear 2^24: 107 -> 120 us; Goldilocks 207 -> 204).
  constexpr uint32_t LOGRE = LOGR - LOGE;
  const bool rowfast = sizeof(T) == 8 && K::NSTAGE >= 2 && S == 1;
  const uint32_t t0 = rowfast ? (tid & (RE - 1)) : t;
  const uint32_t c0 = rowfast ? (tid >> LOGRE) : c;
  const bool stage_lds = staged && !rowfast;
  // the pass's roots w_R^m, m < R: one contiguous block of the root pyramid
  const T *__restrict__ pyr = (const T *)a.root_pyr;
  const T *__restrict__ rootR = pyr + R;
  const uint32_t logTiles = a.logN - LOGR - logC;
  const uint32_t nthr = C * RE;

  T wE[E / 2 + 1];
  load_consts<F, LOGE, U>(wE, rootR, RE, 0);

  {
    // grid = nvec << logTiles tiles; tiles whose column runs are shorter
    // than a 128-byte line (wide tiles, small jobs' narrow tiles, Goldilocks
    // radix >= 2^9) in XCD order, so that the tiles sharing a line meet in one
    // L2 (A/B on one box: the radix-8 variant at 2^24, [9,9,6] with 64-byte
    // runs at radix 512, 277.5-279.4 -> 241.3-244.6 us; 128-byte runs keep
    // the plain order, which measured faster for them)
    const bool short_runs = K::WIDE || ((uint32_t)sizeof(T) << logC) < 128u;
    const uint64_t gt = short_runs ? xcd_tile(blockIdx.x, gridDim.x) : blockIdx.x;
    const uint64_t vec = gt >> logTiles;
    const uint32_t tile = (uint32_t)(gt & ((1ull << logTiles) - 1));
    const uint64_t voff_in = a.vec_off_in ? a.vec_off_in[vec] : vec * a.vec_stride_in;
    const uint64_t voff_out = a.vec_off_out ? a.vec_off_out[vec] : vec * a.vec_stride_out;
    const T *__restrict__ src = (const T *)a.src + voff_in;
    T *__restrict__ dst = (T *)a.dst + voff_out;
    const bool wide_in = sizeof(T) == 4 && (a.flags & kWideIn) != 0;
    const uint64_t *__restrict__ srcw = (const uint64_t *)a.src + voff_in;
    uint64_t *__restrict__ dstw = (uint64_t *)a.dst + voff_out;
    // Buffer loads and stores: vectors of less than 2^32 bytes in the
    // kNarrow4 kernels.  Measured on MI355X against 64-bit addresses:
    // BabyBear / KoalaBear 2^24 -6 %, the 256 x 2^20 blowup-4 encoder -10 %
    // (VGPRs 66 -> 38 at radix 256); Goldilocks 2^24 +2 %, the wide
    // radix-2^11 tiles +4 %.
    const bool buf = kNarrow4 && a.logN + 2u <= 31u;
    const uint32_t vbytes = buf ? (uint32_t)sizeof(T) << a.logN : 0u;
    const __amdgpu_buffer_rsrc_t rs_in = vec_rsrc(src, vbytes), rs_out = vec_rsrc(dst, vbytes);
    if constexpr (EPI == kEpiPrefix) {  // tiles whose p has no k' = p + P k < out_len
      if (((tile << logC) >> logS) >= (uint32_t)a.store_limit) return;
    }
    T x[E];
    // raised issue priority while the tile's loads go out: a wave starting
    // a tile gets its memory requests in ahead of waves in their arithmetic
    __builtin_amdgcn_s_setprio(1);
    load_tile<F, LOGR, LOGE, EPI>(x, src, srcw, wide_in, a, tile, tid, c0, t0, stage_lds, logC, buf, rs_in);
    __builtin_amdgcn_s_setprio(0);

    // column, p and first input position of the load mapping
    const uint32_t q0 = (tile << logC) + c0;
    const uint32_t p = q0 >> logS, j = q0 & (S - 1);
    const uint32_t pos0 = ((p << LOGR) << logS) + j;

    if (stage_lds) {
      // element i = (pp R + k) S + jj of the block is column pp S + jj, row k
      __syncthreads();
      if (kNarrow4 && logS + LOGE <= logC) {
        // i = tid + m nthr with nthr / S a multiple of R: the thread's row k
        // is the same for every m and its column steps by C / E
        const uint32_t k = (tid >> logS) & (R - 1);
        const uint32_t col = ((tid >> (logS + LOGR)) << logS) + (tid & (S - 1));
#pragma unroll
        for (int m = 0; m < E; m++) lds[lrc<kNarrow4>(k, col + ((uint32_t)m << (logC >= (uint32_t)LOGE ? logC - LOGE : 0)), logC)] = x[m];
      } else {
#pragma unroll
        for (int m = 0; m < E; m++) {
          const uint32_t i = tid + (uint32_t)m * nthr;
          const uint32_t jj = i & (S - 1);
          const uint32_t k = (i >> logS) & (R - 1);
          const uint32_t pp = i >> (logS + LOGR);
          lds[lrc<kNarrow4>(k, (pp << logS) + jj, logC)] = x[m];
        }
      }
      __syncthreads();
#pragma unroll
      for (int m = 0; m < E; m++) x[m] = lds[lrc<kNarrow4>(t + m * RE, c, logC)];
    }

    if ((EPI == kEpiGeneral || EPI == kEpiLde) && a.pre_lo) {  // h^pos on input position pos (coset shift)
      const T *plo = (const T *)a.pre_lo, *phi = (const T *)a.pre_hi;
      const uint32_t lomask = (1u << a.lo_bits) - 1;
#pragma unroll
      for (int m = 0; m < (EPI == kEpiLde ? E / 4 : E); m++) {
        const uint32_t pos = pos0 + ((t0 + m * RE) << logS);
        x[m] = F::mulw(x[m], F::wmul(plo[pos & lomask], phi[pos >> a.lo_bits]));
      }
    }

    bool scaled_in = false;  // kEpiScale folded into the twiddle chains
    if (a.flags & kTwiddle) {  // w_N^(k p s) = w_(N/s)^(k p) on input row k
      const uint32_t logNs = a.logN - logS;
      const bool one = logNs <= a.log_mid;
      // Rows k = t0 + m RE: w^(k p) = A B^m with A = w^(t0 p), B = w^(RE p), in
      // two chains (even / odd m) stepped by B^2: two table lookups per
      // thread instead of one or two per element.  For Goldilocks that costs
      // a second product per element against one lookup per element where a
      // single table covers N/s, and still measured faster (2^24 [8,8,8]:
      // 203.5 -> 201.5 us): the per-element gathers of the middle pass
      // waited longer than the products took.
      {
        const T *tlo = (const T *)a.tw_lo, *thi = (const T *)a.tw_hi;
        const bool small = logNs <= a.log_rmax;
        const T *tab = small ? pyr + (1u << logNs) : (const T *)a.tw_mid;
        const uint32_t sh = small ? 0u : a.log_mid - logNs;
        const uint32_t lomask = (1u << a.lo_bits) - 1;
        auto tw = [&](uint32_t e) -> T {  // w_(N/s)^e, e < N/s
          if (one) return tab[e << sh];
          const uint32_t en = e << logS;  // as a power of w_N
          return F::wmul(tlo[en & lomask], thi[en >> a.lo_bits]);
        };
        T A = tw(t0 * p);
        if (EPI == kEpiScale) {  // the inverse's scale rides in A: once per element
          A = F::wmul(A, (T)a.scale_w);
          scaled_in = true;
        }
        const T B = tw((uint32_t)RE * p);
        const T B2 = F::wmul(B, B);
        T ch[2] = {A, F::wmul(A, B)};
#pragma unroll
        for (int m = 0; m < E; m += 2) {
#pragma unroll
          for (int r = 0; r < 2; r++) {
            x[m + r] = F::mulw(x[m + r], ch[r]);
            if (m + 2 < E) ch[r] = F::wmul(ch[r], B2);
          }
        }
      }
    }

    // in-tile stage 0: stride R/E, one group per thread, no in-tile twiddle
    if constexpr (EPI == kEpiLde && K::NSTAGE >= 2) dft_reg_nz4<F, LOGE, U>(x, wE);
    else dft_reg<F, LOGE, (K::NSTAGE == 1 ? K::LOGLAST : LOGE), 0, U>(x, wE);
    tile_stage<F, LOGR, LOGE, 1, U, kNarrow4>(x, wE, lds, rootR, t, c, logC, 0, t0, c0);
    const uint32_t q = (tile << logC) + c;  // the store mapping's column

    // register slot (g, kk) holds tile output row kk (R/EL) + t + g (R/E)
    constexpr int EL = 1 << K::LOGLAST;
    if constexpr (EPI == kEpiPrefix) {
      // Output row k of column q = p s + j (j < s) holds
      // Z_j[k'] = sum_a x[j + s a] w_(N/s)^(a k'), k' = p + P k, P = N/(R s):
      // the N/s-point transform of x[j :: s], whose last digit this pass
      // added (the first pass, s = N/R, has P = 1, p = 0, Z_j = NTT_R(x[j :: N/R])).
      // The prefix is X[k'] = sum_j w_N^(j k') Z_j[k'] for k' < out_len <= N/s
      // (the decomposition of transform_decomposition.rs:63-94 with n1 = N/s).
      // A tile holds one p and the columns j = j0 + c (C <= s):
      // w_N^(j0 k') sum_c w_N^(c k') Z_(j0 + c)[k'], one partial sum per (k', tile).
      const uint32_t out_len = (uint32_t)a.store_limit;  // <= N/s
      const uint32_t logP = a.logN - LOGR - logS;
      const uint32_t nrows = min((uint32_t)R, (out_len - p + (1u << logP) - 1) >> logP);  // k' < out_len
      const T *tlo = (const T *)a.tw_lo, *thi = (const T *)a.tw_hi;
      const uint32_t lomask = (1u << a.lo_bits) - 1;
      const bool one = ((C << LOGR) << logP) <= (1u << a.lo_bits);  // c k' < C N/s
      auto twN = [&](uint32_t e) -> T {  // w_N^e, e < C N/s <= N
        return one ? tlo[e] : F::wmul(tlo[e & lomask], thi[e >> a.lo_bits]);
      };
      const uint32_t cP = c << logP;
      __syncthreads();  // the last in-tile exchange has been read
      if (one || sizeof(T) == 8) {  // one or two lookups per element
#pragma unroll
        for (int g = 0; g < E / EL; g++) {
#pragma unroll
          for (int kk = 0; kk < EL; kk++) {
            const uint32_t krow = (uint32_t)kk * (R / EL) + t + g * RE;
            if (krow < nrows) lds[lpad((krow << logC) + c)] = F::mulw(x[g * EL + kk], twN(c * p + cP * krow));
          }
        }
      } else {
        // 4-byte fields, two-level table: w_N^(c k') on the thread's rows
        // krow = kk R/EL + t + g RE is A_g B^kk per row group, A_g =
        // w_N^(c (p + P (t + g RE))), B = w_N^(c P R/EL): two power chains
        // (even / odd kk, stepped by B^2), the same two products per element
        // as the two-level lookup, two lookups per group instead of two per
        // element.  N = 2^22 BabyBear, prefix of 2^16: 31.9-32.2 -> 28.1-28.4
        // us (2^8, 2^12, 2^20 unchanged).  Goldilocks, whose products cost 22
        // VALU slots against 6, keeps the lookups.
        const T B = twN(cP * (uint32_t)(R / EL));
        const T B2 = F::wmul(B, B);
#pragma unroll
        for (int g = 0; g < E / EL; g++) {
          const T A = twN(c * p + cP * (t + (uint32_t)g * RE));
          T ch[2] = {A, F::wmul(A, B)};
#pragma unroll
          for (int kk = 0; kk < EL; kk++) {
            const uint32_t krow = (uint32_t)kk * (R / EL) + t + g * RE;
            if (krow < nrows) {  // rows grow with kk: a chain past nrows stops
              lds[lpad((krow << logC) + c)] = F::mulw(x[g * EL + kk], ch[kk & 1]);
              if (kk + 2 < EL) ch[kk & 1] = F::wmul(ch[kk & 1], B2);
            }
          }
        }
      }
      __syncthreads();
      // partial sums laid out [vector][k'][tile of this p]: the reduction over
      // tiles reads contiguous runs
      const uint32_t j0 = (tile << logC) & (S - 1);
      const uint64_t tps = (uint64_t)S >> logC;  // tiles per p
      T *__restrict__ part = (T *)a.dst + (gt >> logTiles) * tps * out_len + (j0 >> logC);
      const bool big = a.logN > a.log_mid;
      const bool small = a.logN <= a.log_rmax;
      const T *tab = small ? pyr + (1u << a.logN) : (const T *)a.tw_mid;
      const uint32_t sh = small ? 0u : a.log_mid - a.logN;
      for (uint32_t r = tid; r < nrows; r += nthr) {
        T acc[4] = {0, 0, 0, 0};  // independent chains over the columns
        uint32_t cc = 0;
        for (; cc + 4 <= C; cc += 4) {
#pragma unroll
          for (int u = 0; u < 4; u++) acc[u] = F::add(acc[u], lds[lpad((r << logC) + cc + u)]);
        }
        for (; cc < C; cc++) acc[0] = F::add(acc[0], lds[lpad((r << logC) + cc)]);
        const T sum = F::add(F::add(acc[0], acc[1]), F::add(acc[2], acc[3]));
        const uint32_t k = p + (r << logP);
        const uint32_t e = k * j0;  // < N
        const T w = big ? F::wmul(tlo[e & lomask], thi[e >> a.lo_bits]) : tab[e << sh];
        part[(uint64_t)k * tps] = F::mulw(sum, w);
      }
      return;
    }
    const uint32_t logNR = a.logN - LOGR;
    const bool scale = (EPI == kEpiScale && !scaled_in) || (EPI == kEpiGeneral && (a.flags & kScale) != 0);
    const bool pruned = EPI == kEpiGeneral && a.store_limit < (1ull << a.logN);
    const T sw = (T)a.scale_w;
    uint32_t nonzero = 0;
    // kEpiPost: the next pass's inter-pass twiddle w_N^(k2 p2 s2) on output
    // row krow of column q, where the next pass (radix R2, stride s2 = s/R2)
    // reads this output as its row k2 = (q / s2) mod R2 of column
    // p2 = krow L + q / s (L = N / (R s)): the exponent is c0 + c1 krow with
    // c1 = k2 L s2, c0 = k2 (q / s) s2 (mod N), a power chain over the
    // thread's rows krow = kk R/EL + t + g RE.
    T pch[E / EL], pstep = 0;
    if constexpr (EPI == kEpiPost) {
      const T *tlo = (const T *)a.tw_lo, *thi = (const T *)a.tw_hi;
      const uint32_t lomask = (1u << a.lo_bits) - 1;
      const uint64_t nmask = (1ull << a.logN) - 1;
      auto twN = [&](uint64_t e) -> T {  // w_N^e
        e &= nmask;
        return F::wmul(tlo[(uint32_t)e & lomask], thi[e >> a.lo_bits]);
      };
      const uint32_t logS2 = logS - a.post_logR;
      const uint32_t logL = a.logN - LOGR - logS;
      const uint64_t k2 = (q >> logS2) & ((1u << a.post_logR) - 1);
      const uint64_t c1 = k2 << (logL + logS2);
      const uint64_t c0 = (k2 * (uint64_t)(q >> logS)) << logS2;
      pstep = twN(c1 * (uint64_t)(R / EL));
#pragma unroll
      for (int g = 0; g < E / EL; g++) {
        pch[g] = twN(c0 + c1 * (uint64_t)(t + g * RE));
        if (a.flags & kScale) pch[g] = F::wmul(pch[g], (T)a.scale_w);  // an inverse's N^-1, once per element
      }
    }
    // Decode's output multiplier (post_lo / post_hi: the two-level table of
    // g^e, e < N) on output oidx = (krow << logNR) + q, 4-byte fields: power
    // chains per row group, och = g^(((t + g RE) << logNR) + q) stepped by
    // g^((2 R/EL) << logNR) over even / odd kk, two lookups per group and
    // the same two products per element as the lookup.
    T och[E / EL][2], ostep2 = 0;
    if constexpr (sizeof(T) == 4 && EPI == kEpiGeneral) {
      if (a.post_lo) {
        const T *plo = (const T *)a.post_lo, *phi = (const T *)a.post_hi;
        const uint32_t lomask = (1u << a.lo_bits) - 1;
        auto pw = [&](uint32_t e) -> T { return F::wmul(plo[e & lomask], phi[e >> a.lo_bits]); };
        const T ostep = pw((uint32_t)(R / EL) << logNR);
        ostep2 = F::wmul(ostep, ostep);
#pragma unroll
        for (int g = 0; g < E / EL; g++) {
          och[g][0] = pw(((t + (uint32_t)g * RE) << logNR) + q);
          och[g][1] = F::wmul(och[g][0], ostep);
        }
      }
    }
    // WIDE: 64-bit output words (a uniform choice: the loop is compiled twice)
    auto store_all = [&](auto wide_tag) {
      constexpr bool WIDE = decltype(wide_tag)::value;
#pragma unroll
      for (int g = 0; g < E / EL; g++) {
#pragma unroll
        for (int kk = 0; kk < EL; kk++) {
          const uint32_t krow = (uint32_t)kk * (R / EL) + t + g * RE;
          T v = x[g * EL + kk];
          if constexpr (EPI == kEpiPost) {
            v = F::mulw(v, pch[g]);
            if (kk + 1 < EL) pch[g] = F::wmul(pch[g], pstep);
          }
          if (scale) v = F::mulw(v, sw);
          const uint32_t oidx = (krow << logNR) + q;
          if (EPI == kEpiGeneral && a.post_lo) {
            if constexpr (sizeof(T) == 4) {
              v = F::mulw(v, och[g][kk & 1]);
              if (kk + 2 < EL) och[g][kk & 1] = F::wmul(och[g][kk & 1], ostep2);
            } else {
              const T *plo = (const T *)a.post_lo, *phi = (const T *)a.post_hi;
              v = F::mulw(v, F::wmul(plo[oidx & ((1u << a.lo_bits) - 1)], phi[oidx >> a.lo_bits]));
            }
          }
          if (!pruned || (uint64_t)oidx < a.store_limit) {
            if constexpr (WIDE) dstw[oidx] = (uint64_t)v;
            else if (buf)  // oidx = (t << logNR) + q + (the row's uniform part << logNR)
              buf_store<T>(v, rs_out, ((t << logNR) + q) * (uint32_t)sizeof(T),
                           (((uint32_t)kk * (R / EL) + (uint32_t)g * RE) << logNR) * (uint32_t)sizeof(T));
            else dst[oidx] = v;
          } else {
            nonzero += (v != 0);
          }
        }
      }
    };
    if (sizeof(T) == 4 && (a.flags & kWideOut) != 0) store_all(std::integral_constant<bool, sizeof(T) == 4>{});
    else store_all(std::false_type{});
    if (EPI == kEpiGeneral && a.nonzero_count && nonzero) atomicAdd(a.nonzero_count, nonzero);
  }
}

template <class F, int LOGR, int LOGE, int U, int EPI, bool FIXC>
__global__ void __attribute__((amdgpu_flat_work_group_size(1, PassKernel<F, LOGR, LOGE>::MAXT)))
k_stockham_pass(PassArgs a) {
  stockham_pass_body<F, LOGR, LOGE, U, EPI, FIXC>(a);
}

}  // namespace omr
