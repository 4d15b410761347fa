// batch_kernel.h -- persistent Stockham tile kernel with a runtime radix,
// for batches of vectors of uneven power-of-two sizes.
//
// Every task is one tile (batch_log_tile: 8192 or 4096 elements) of one
// Stockham pass (the pass semantics of engine.h) of one vector, or of
// several whole small vectors when a vector has fewer than C columns.  All
// tasks use tile / 16 threads and a register radix of 16; the pass radix
// 2^logR (4 <= logR <= 12), the vector
// size and the twiddle tables are per-task data, so one kernel serves every
// size in the batch.  Workgroups walk the task list with a grid stride
// (tasks are sorted by cost, largest first).
#pragma once
#include "engine.h"

namespace omr {

// log2 of the tile of field id f.  2^13 elements (64 KiB of Goldilocks
// words): a radix-2^10 pass still has 8 columns, i.e. 64-byte runs, so
// vectors up to 2^20 take two passes.  TeddyBear keeps 2^12: its kernel
// (99 VGPRs) runs 4 waves per SIMD with either tile, and in 8-wave
// workgroups each barrier holds twice the waves (4096 vectors of
// log n ~ U{10..20}: 10.1 ms with 2^12 tiles, 10.8 ms with 2^13).
constexpr int batch_log_tile(int field) { return field == kTeddyBear ? 12 : 13; }
constexpr int batch_tile_threads(int field) { return (1 << batch_log_tile(field)) / 16; }

struct BatchTask {
  uint32_t vec0;   // first vector of the task (index into the pass's vector table)
  uint32_t tile;   // tile index within the vector (0 when the task holds several vectors)
  uint8_t logN, logR, logS, flags;  // flags: kTwiddle | kScale
  uint16_t tset;   // twiddle-table set
  uint16_t nvec;   // vectors held by the task (columns of vectors >= nvec are idle)
};

struct BatchVec {  // one vector in one pass: buffer ids (0 in, 1 out, 2 A, 3 B) and offsets
  uint64_t src_off, dst_off;
  uint32_t src_buf, dst_buf;
};

struct BatchTables {  // tables of one root of order 2^logN, multiplier form
  const void *lo, *hi, *pyr, *mid;  // pyr: the root pyramid (w_(2^L)^m at [2^L + m], L <= log_rmax)
  uint64_t scale_w;
  uint32_t lo_bits, log_rmax, log_mid;
  uint32_t unit;  // register-DFT constants: 0 = table constants, 39 / 153 = 2^k shifts (Goldilocks)
};

struct BatchBases {
  const void *buf[4];
};

template <class F, int LEI, int U>
__device__ __forceinline__ void gen_stage(typename F::T (&x)[16], const typename F::T (&wE)[9],
                                          const typename F::T *lds, const typename F::T *root,
                                          uint32_t t, uint32_t c, uint32_t logC, uint32_t RE,
                                          uint32_t logs) {
  using T = typename F::T;
  constexpr int EI = 1 << LEI;
#pragma unroll
  for (int g = 0; g < 16 / EI; g++) {
    const uint32_t tt = t + g * RE;
    const uint32_t ps = (tt >> logs) << logs;
    const uint32_t jj = tt & ((1u << logs) - 1);
#pragma unroll
    for (int kk = 0; kk < EI; kk++) {
      const uint32_t row = ps * EI + ((uint32_t)kk << logs) + jj;
      T v = lds[lpad((row << logC) + c)];
      const uint32_t e = (uint32_t)kk * ps;
      if (kk != 0) v = F::mulw(v, root[e]);  // e = 0: times root[0] = 1
      x[g * EI + kk] = v;
    }
  }
  GroupDft<F, 4, LEI, 16 / EI, U>::run(x, wE);
}

// One task (tile) with register-DFT constant unit U (0: table constants).
template <class F, int U>
__device__ __forceinline__ void batch_task(const BatchTask &d, const BatchTables &tb,
                                           const BatchVec *__restrict__ vecs, const BatchBases &bases,
                                           typename F::T *lds, uint32_t tid) {
  using T = typename F::T;
  const T *__restrict__ pyr = (const T *)tb.pyr;
  const uint32_t logN = d.logN, logR = d.logR, logS = d.logS;
  const uint32_t R = 1u << logR, RE = R >> 4;
  const T *__restrict__ root = pyr + R;  // w_R^m, m < R: one contiguous block
  constexpr int kLogTile = batch_log_tile(F::ID);
  constexpr uint32_t kTileElems = 1u << kLogTile, kTileThreads = kTileElems / 16;
  const uint32_t logC = kLogTile - logR;  // C = kTileElems / R columns per task
  const uint32_t c = tid & ((1u << logC) - 1), t = tid >> logC;
  const uint32_t logCols = logN - logR;  // columns per vector
  const uint32_t logCv = logCols < logC ? logCols : logC;
  const uint32_t S = 1u << logS;
  // A pass with s = 1 and at least two in-tile stages (Goldilocks) loads
  // with the row fastest, thread (t0, c0) = (tid mod R/16, tid div R/16): a
  // column is a contiguous run of R elements, so the loads are coalesced as
  // they are and stage 1's LDS exchange hands the values to the column-fast
  // mapping of the stores; no staging transpose (whose 8-byte column writes
  // conflict on the LDS banks).  Other passes with s < C stage a contiguous
  // block through LDS.
  const bool rowfast = sizeof(T) == 8 && logR >= 5 && logS == 0;
  const uint32_t t0 = rowfast ? (tid & (RE - 1)) : t;
  const uint32_t c0 = rowfast ? (tid >> (logR - 4)) : c;
  // offsets inside one vector are 32-bit (N <= 2^32)
  const uint32_t u0 = c0 >> logCv;  // vector of the load mapping's column
  const bool active0 = u0 < d.nvec;
  const uint32_t q0 = (d.tile << logC) + (c0 & ((1u << logCv) - 1));
  const uint32_t p = q0 >> logS, j = q0 & (S - 1);
  const BatchVec bv = vecs[d.vec0 + (active0 ? u0 : 0)];
  const T *__restrict__ src = (const T *)bases.buf[bv.src_buf] + bv.src_off;
  T wE[9];
  load_consts<F, 4, U>(wE, root, RE, 0);

  T x[16];
  if (rowfast || S >= (1u << logCv)) {
    const T *base = src + (((p << logR) << logS) + j);
#pragma unroll
    for (int m = 0; m < 16; m++) x[m] = active0 ? base[(t0 + m * RE) << logS] : (T)0;
  } else {
    // each vector of the task contributes the contiguous block
    // [q0 R, (q0 + Cv) R) of its data, q0 = tile * C
    if (sizeof(T) == 4 && d.nvec == 1) {
      // one vector (u0 = 0, src is its base): the block is one run of
      // kTileElems elements, 16 independent loads per thread.  (Goldilocks
      // stages only passes of stride 2 .. C - 1, rare in batches, and this
      // form would cost its kernel spills.)
      const T *blk = src + (d.tile << (logC + logR));
#pragma unroll
      for (uint32_t m = 0; m < 16; m++) {
        const uint32_t i = tid + m * kTileThreads;
        const uint32_t jj = i & (S - 1), k = (i >> logS) & (R - 1), pp = i >> (logS + logR);
        lds[lpad((k << logC) + (pp << logS) + jj)] = blk[i];
      }
    } else {
      const uint32_t vblk = 1u << (logCv + logR);
      for (uint32_t i = tid; i < kTileElems; i += kTileThreads) {
        const uint32_t uu = i >> (logCv + logR);
        if (uu >= d.nvec) continue;
        const uint32_t ii = i & (vblk - 1);
        const uint32_t jj = ii & (S - 1);
        const uint32_t k = (ii >> logS) & (R - 1);
        const uint32_t pp = ii >> (logS + logR);
        const BatchVec v2 = vecs[d.vec0 + uu];
        const T *s2 = (const T *)bases.buf[v2.src_buf] + v2.src_off;
        const uint32_t col = (uu << logCv) + (pp << logS) + jj;
        lds[lpad((k << logC) + col)] = s2[(d.tile << (logC + logR)) + ii];
      }
    }
    __syncthreads();
#pragma unroll
    for (int m = 0; m < 16; m++) x[m] = lds[lpad(((t + m * RE) << logC) + c)];
  }

  bool scaled = false;  // the inverse's 1/N folded into the twiddle chain
  if (d.flags & kTwiddle) {  // w_(N/s)^(k p) on input row k = t0 + m RE
    const uint32_t logNs = logN - logS;
    const bool one = logNs <= tb.log_mid;
    const bool small = logNs <= tb.log_rmax;
    const T *tab = small ? pyr + (1u << logNs) : (const T *)tb.mid;
    const uint32_t sh = small ? 0u : tb.log_mid - logNs;
    if (sizeof(T) == 4 || !one) {  // power chains A B^m, as in the pass engine
      const T *tlo = (const T *)tb.lo, *thi = (const T *)tb.hi;
      const uint32_t lomask = (1u << tb.lo_bits) - 1;
      auto tw = [&](uint32_t e) -> T {  // w_(N/s)^e, e < N/s
        if (one) return tab[e << sh];
        const uint32_t en = e << logS;
        return F::wmul(tlo[en & lomask], thi[en >> tb.lo_bits]);
      };
      T A = tw(t0 * p);
      const T B = tw(RE * p);
      if (d.flags & kScale) {
        A = F::wmul(A, (T)tb.scale_w);
        scaled = true;
      }
      const T B2 = F::wmul(B, B);
      T ch[2] = {A, F::wmul(A, B)};
#pragma unroll
      for (int m = 0; m < 16; m += 2) {
#pragma unroll
        for (int r = 0; r < 2; r++) {
          x[m + r] = F::mulw(x[m + r], ch[r]);
          if (m + 2 < 16) ch[r] = F::wmul(ch[r], B2);
        }
      }
    } else {
#pragma unroll
      for (int m = 0; m < 16; m++) x[m] = F::mulw(x[m], tab[((t0 + m * RE) * p) << sh]);
    }
  }

  dft_reg<F, 4, 4, 0, U>(x, wE);
  const uint32_t nst = (logR + 3) >> 2;
  const uint32_t loglast = logR - 4 * (nst - 1);
  for (uint32_t st = 1; st < nst; st++) {
    const uint32_t logEi = (st == nst - 1) ? loglast : 4;
    const uint32_t logs = logR - 4 * st - logEi;
    const uint32_t tw = st == 1 ? t0 : t, cw = st == 1 ? c0 : c;  // stage 1: registers of the load mapping
    __syncthreads();
#pragma unroll
    for (int m = 0; m < 16; m++) lds[lpad(((tw + m * RE) << logC) + cw)] = x[m];
    __syncthreads();
    switch (logEi) {
      case 4: gen_stage<F, 4, U>(x, wE, lds, root, t, c, logC, RE, logs); break;
      case 3: gen_stage<F, 3, U>(x, wE, lds, root, t, c, logC, RE, logs); break;
      case 2: gen_stage<F, 2, U>(x, wE, lds, root, t, c, logC, RE, logs); break;
      default: gen_stage<F, 1, U>(x, wE, lds, root, t, c, logC, RE, logs); break;
    }
  }

  if ((d.flags & kScale) && !scaled) {  // a block of its own: the store loop stays branch-free
    const T sw = (T)tb.scale_w;
#pragma unroll
    for (int m = 0; m < 16; m++) x[m] = F::mulw(x[m], sw);
  }
  // stores: column-fast mapping (t, c)
  const uint32_t u = c >> logCv;
  const bool active = u < d.nvec;
  const uint32_t q = (d.tile << logC) + (c & ((1u << logCv) - 1));
  const BatchVec bs = rowfast ? vecs[d.vec0 + (active ? u : 0)] : bv;
  T *__restrict__ dst = (T *)bases.buf[bs.dst_buf] + bs.dst_off;
  const uint32_t logNR = logN - logR;
  const uint32_t ELm = (1u << loglast) - 1;
#pragma unroll
  for (int m = 0; m < 16; m++) {
    const uint32_t g = (uint32_t)m >> loglast, kk = (uint32_t)m & ELm;
    const uint32_t krow = kk * (R >> loglast) + t + g * RE;
    if (active) dst[(krow << logNR) + q] = x[m];
  }
}

// 4 waves per SIMD: what the Goldilocks LDS tile allows (2 workgroups of 8
// waves per CU); Goldilocks fits it in 128 VGPRs without spills.  With
// 4096-element tiles the cap ran 7 % faster than 3 waves per SIMD without it
// (4096 Goldilocks vectors of log n ~ U{10..20}).
template <class F>
__global__ void __attribute__((amdgpu_flat_work_group_size(1, batch_tile_threads(F::ID)), amdgpu_waves_per_eu(4)))
k_batch_tiles(const BatchTask *__restrict__ tasks, uint32_t ntasks, const BatchVec *__restrict__ vecs,
              const BatchTables *__restrict__ tsets, BatchBases bases) {
  using T = typename F::T;
  __shared__ T lds[lds_padded(1 << batch_log_tile(F::ID))];
  const uint32_t tid = threadIdx.x;

  // the next task's descriptor is loaded while this one runs: its tables,
  // vector entry and data loads then start without a chain of dependent
  // descriptor latencies
  // Each round of G tasks gives XCD x (workgroups b = x mod 8, dealt round
  // robin) the x-th contiguous eighth, so consecutive tiles of a vector,
  // whose 64-byte column runs share 128-byte lines, meet in one L2.
  const uint32_t G = gridDim.x;
  const uint32_t slot = (G & 7u) ? blockIdx.x : (blockIdx.x & 7u) * (G >> 3) + (blockIdx.x >> 3);
  BatchTask dn = tasks[slot < ntasks ? slot : 0];
  for (uint32_t ti = slot; ti < ntasks; ti += G) {
    const BatchTask d = dn;
    const BatchTables tb = tsets[d.tset];
    if ((uint64_t)ti + G < ntasks) dn = tasks[ti + G];
    if constexpr (F::kPow2Roots) {  // per-task register-DFT constants: forward / inverse roots
      if (tb.unit == 39) batch_task<F, 39>(d, tb, vecs, bases, lds, tid);
      else if (tb.unit == 153) batch_task<F, 153>(d, tb, vecs, bases, lds, tid);
      else batch_task<F, 0>(d, tb, vecs, bases, lds, tid);
    } else {
      batch_task<F, 0>(d, tb, vecs, bases, lds, tid);
    }
    __syncthreads();  // LDS is reused by the next task
  }
}

}  // namespace omr
