// misc.h -- non-pass kernels: brute-force DFT, partial-output reductions,
// stride-subset fold, elementwise field ops.  Templated over the field.
#pragma once
#include "field.h"

namespace omr {

struct Tw2 {  // two-level table of w^e:  w^e = lo[e & mask] * hi[e >> bits]  (multiplier form)
  const void *lo, *hi;
  uint32_t bits;
};

template <class F>
__device__ __forceinline__ typename F::T tw2(const Tw2 &t, uint64_t e) {
  using T = typename F::T;
  const T *lo = (const T *)t.lo, *hi = (const T *)t.hi;
  return F::wmul(lo[e & ((1ull << t.bits) - 1)], hi[e >> t.bits]);
}

// Full brute-force DFT (the reference's Dft, src/encoders/dft.rs:9-22):
// out[k] = sum_i x[i] w^(ik), one thread per output k, w^(ik) stepped by w^k.
template <class F>
__global__ void __launch_bounds__(256) k_dft_full(const typename F::T *__restrict__ x,
                                                  typename F::T *__restrict__ out, uint64_t n,
                                                  uint32_t logN, Tw2 tw) {
  using T = typename F::T;
  const uint64_t k = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= n) return;
  const T step = tw2<F>(tw, k);
  T t = F::one_w(), acc = 0;
  for (uint64_t i = 0; i < n; i++) {
    acc = F::add(acc, F::mulw(x[i], t));
    t = F::wmul(t, step);
  }
  out[k] = acc;
}

// Direct partial DFT (the reference's DftPartial, src/encoders/dft.rs:40-88):
// each workgroup sums a chunk of inputs for every output k < out_len and
// writes part[blockIdx.x * out_len + k]; k_reduce_rows finishes the sum.
template <class F>
__global__ void __launch_bounds__(256) k_dft_partial_chunks(const typename F::T *__restrict__ x,
                                                            typename F::T *__restrict__ part,
                                                            uint64_t n, uint32_t out_len,
                                                            uint32_t chunk, Tw2 tw) {
  using T = typename F::T;
  __shared__ T red[256];
  const uint64_t i0 = (uint64_t)blockIdx.x * chunk;
  for (uint32_t k = 0; k < out_len; k++) {
    T acc = 0;
    for (uint32_t i = threadIdx.x; i < chunk && i0 + i < n; i += blockDim.x) {
      const uint64_t e = ((i0 + i) * (uint64_t)k) & (n - 1);
      acc = F::add(acc, F::mulw(x[i0 + i], tw2<F>(tw, e)));
    }
    red[threadIdx.x] = acc;
    __syncthreads();
    for (uint32_t s = blockDim.x / 2; s > 0; s >>= 1) {
      if (threadIdx.x < s) red[threadIdx.x] = F::add(red[threadIdx.x], red[threadIdx.x + s]);
      __syncthreads();
    }
    if (threadIdx.x == 0) part[(uint64_t)blockIdx.x * out_len + k] = red[0];
    __syncthreads();
  }
}

// out[k] = sum_{b < nrows} part[b * ld + k] for k < ncols (mod p).
template <class F>
__global__ void __launch_bounds__(256) k_reduce_rows(const typename F::T *__restrict__ part,
                                                     typename F::T *__restrict__ out, uint32_t nrows,
                                                     uint64_t ld, uint64_t ncols) {
  using T = typename F::T;
  const uint64_t k = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= ncols) return;
  T acc = 0;
  for (uint32_t b = 0; b < nrows; b++) acc = F::add(acc, part[(uint64_t)b * ld + k]);
  out[k] = acc;
}

// Transform-decomposition recombine (src/encoders/transform_decomposition.rs
// :63-94, restated for a device layout): Y[k1 * n2 + r] holds the n1-point
// NTT of column r at k1 < rows.  For output k,
//   out[k] = sum_{r < n2} w_N^(r k) Y[(k mod n1) * n2 + r].
// Grid (ceil(out_len / KT), splits); each workgroup covers KT consecutive k
// and an r-range; results go to part[split * out_len + k].
template <class F, int KT>
__global__ void __launch_bounds__(256) k_td_recombine(const typename F::T *__restrict__ Y,
                                                      typename F::T *__restrict__ part,
                                                      uint64_t out_len, uint64_t n2,
                                                      uint32_t logN1, uint64_t r_per_split,
                                                      uint32_t logN, Tw2 tw) {
  using T = typename F::T;
  __shared__ T red[KT][256 / KT + 1];
  const uint32_t lane = threadIdx.x % (256 / KT);
  const uint32_t kl = threadIdx.x / (256 / KT);
  const uint64_t k = (uint64_t)blockIdx.x * KT + kl;
  const uint64_t r0 = (uint64_t)blockIdx.y * r_per_split;
  const uint64_t r1 = r0 + r_per_split < n2 ? r0 + r_per_split : n2;
  T acc = 0;
  if (k < out_len) {
    const T *row = Y + (k & ((1ull << logN1) - 1)) * n2;
    const uint64_t mask = (1ull << logN) - 1;
    const uint64_t stride = 256 / KT;
    if (r0 + lane < r1) {
      T t = tw2<F>(tw, ((r0 + lane) * k) & mask);
      const T step = tw2<F>(tw, (stride * k) & mask);
      for (uint64_t r = r0 + lane; r < r1; r += stride) {
        acc = F::add(acc, F::mulw(row[r], t));
        t = F::wmul(t, step);
      }
    }
  }
  red[kl][lane] = acc;
  __syncthreads();
  for (uint32_t s = (256 / KT) / 2; s > 0; s >>= 1) {
    if (lane < s) red[kl][lane] = F::add(red[kl][lane], red[kl][lane + s]);
    __syncthreads();
  }
  if (lane == 0 && k < out_len) part[(uint64_t)blockIdx.y * out_len + k] = red[kl][0];
}

constexpr int kFoldRows = 32;  // rows per chunk of the row folds

// Row fold: part[c * m + i] = sum_{t in chunk c} x[i + m t] (mod p), grid
// (ceil(m / 256), nchunks, nvec); chunk c covers t in [c tpc, (c+1) tpc);
// set z reads x + z xs and writes part + z ps.
template <class F>
__global__ void __launch_bounds__(256) k_fold(const typename F::T *__restrict__ x,
                                              typename F::T *__restrict__ part, uint64_t n,
                                              uint64_t m, uint64_t tpc, uint64_t xs, uint64_t ps) {
  using T = typename F::T;
  const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= m) return;
  x += (uint64_t)blockIdx.z * xs;
  part += (uint64_t)blockIdx.z * ps;
  const uint64_t nt = n / m;
  const uint64_t t0 = (uint64_t)blockIdx.y * tpc;
  const uint64_t t1 = t0 + tpc < nt ? t0 + tpc : nt;
  const T *__restrict__ col = x + i + m * t0;
  T acc = 0;
  const uint64_t cnt = t1 > t0 ? t1 - t0 : 0;
  if (cnt == kFoldRows) {  // full chunk: every load issued before the adds, four chains
    T a[4] = {0, 0, 0, 0};
#pragma unroll
    for (int r = 0; r < kFoldRows; r++) a[r & 3] = F::add(a[r & 3], col[(uint64_t)r * m]);
    acc = F::add(F::add(a[0], a[1]), F::add(a[2], a[3]));
  } else if (cnt < kFoldRows) {  // short chunk (a last level): the same, loads guarded
    T a[4] = {0, 0, 0, 0};
#pragma unroll
    for (int r = 0; r < kFoldRows; r++)
      if ((uint64_t)r < cnt) a[r & 3] = F::add(a[r & 3], col[(uint64_t)r * m]);
    acc = F::add(F::add(a[0], a[1]), F::add(a[2], a[3]));
  } else {
    for (uint64_t t = 0; t < cnt; t++) acc = F::add(acc, col[m * t]);
  }
  part[(uint64_t)blockIdx.y * m + i] = acc;
}

// Stride subset out[j] = X[j N/m] of one vector in one launch, for
// 16 <= m <= 512 (powers of two).  X[j N/m] = sum_i y[i] w_m^(ij) with the
// fold y[i] = sum_t x[i + m t], so:
// * workgroup b of G adds rows [b R, (b+1) R) of x viewed as N/m rows of m
//   (one contiguous block of R m elements; 16-byte loads, kSfUnroll rows in
//   flight per thread; the 512 threads are m/V column groups of V = 16 bytes
//   x 512 V/m row subsets, added up through LDS) into part[b m + i];
// * the workgroups finish in two levels of "last one done": the last of
//   each group of `group` workgroups (counter cnt[g]) adds its group's rows
//   into part2[g m + i]; the last of those (counter cnt[G / group]) adds
//   the part2 rows, runs the m-point NTT in LDS (radix-2 decimation in
//   frequency over the roots w_m^k, level log m of the root pyramid) and
//   writes out[j] in natural order, then clears the counters for the next
//   launch.  No workgroup waits for another: the ones that are not last
//   simply end.
constexpr int kSfThreads = 512;
constexpr int kSfUnroll = 8;
constexpr int kSfMaxLogM = 9;
constexpr int kSfGroup = 16;      // workgroups per first-level group
constexpr int kSfMaxGroups = 32;  // 512 workgroups at most

// Partial sums that one workgroup writes and another reads within the
// launch go through agent-scope atomic stores and loads (coherent across
// the XCDs' L2s without cache maintenance), ordered around the counters by
// workgroup-scope fences (which wait for the stores to complete).  Plain
// stores with agent-scope release / acquire fences were bit-exact too but
// write back and invalidate an XCD's L2 per workgroup: 83 / 110 us against
// 12.6 / 16.6 us (BabyBear / Goldilocks, N = 2^22, m = 256).
template <class T>
__device__ __forceinline__ void sf_store(T *p, T v) {
  __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
template <class T>
__device__ __forceinline__ T sf_load(const T *p) {
  return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ void sf_release() { __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup"); }
__device__ __forceinline__ void sf_acquire() { __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup"); }

// Column sums of nrows rows of m (<= kSfThreads) partials written by other
// workgroups of this launch: thread (i, sub) = (tid mod m, tid div m) loads
// rows sub, sub + 512/m, ... all at once (at most MAXR), the row subsets
// meet in LDS; returns column tid's sum in threads tid < m.
template <class F, int MAXR>
__device__ __forceinline__ typename F::T sf_sum_rows(const typename F::T *src, uint32_t nrows, uint32_t m,
                                                     uint32_t tid, typename F::T *lds) {
  using T = typename F::T;
  const uint32_t cs = kSfThreads / m;
  const uint32_t i = tid % m, sub = tid / m;
  T v[MAXR];
#pragma unroll
  for (int k = 0; k < MAXR; k++) {
    const uint32_t r = sub + (uint32_t)k * cs;
    v[k] = r < nrows ? sf_load(&src[(uint64_t)r * m + i]) : (T)0;
  }
  T s = 0;
#pragma unroll
  for (int k = 0; k < MAXR; k++) s = F::add(s, v[k]);
  __syncthreads();  // lds is free
  lds[sub * m + i] = s;
  __syncthreads();
  T t = 0;
  if (tid < m)
    for (uint32_t u = 0; u < cs; u++) t = F::add(t, lds[u * m + tid]);
  return t;
}

template <class F>
__global__ void __launch_bounds__(kSfThreads) k_stride_fused(const typename F::T *__restrict__ x,
                                                             typename F::T *__restrict__ part,
                                                             typename F::T *__restrict__ part2, uint32_t *cnt,
                                                             typename F::T *__restrict__ out,
                                                             const typename F::T *__restrict__ root_m,
                                                             uint32_t logm, uint32_t rows_per_wg, uint32_t group) {
  using T = typename F::T;
  constexpr int V = 16 / (int)sizeof(T);
  typedef T VT __attribute__((ext_vector_type(V)));
  __shared__ T lds[kSfThreads * V];
  __shared__ uint32_t ticket;
  const uint32_t m = 1u << logm;
  const uint32_t tpr = m / V;              // threads per row
  const uint32_t rsub = kSfThreads / tpr;  // row subsets
  const uint32_t tid = threadIdx.x;
  const uint32_t cg = tid % tpr, rs = tid / tpr;

  {  // fold of this workgroup's rows: thread (cg, rs) takes rows rs + k rsub
    const T *base = x + (uint64_t)blockIdx.x * rows_per_wg * m + (uint64_t)cg * V;
    const uint32_t K = rows_per_wg / rsub;
    T acc[V];
#pragma unroll
    for (int v = 0; v < V; v++) acc[v] = 0;
    for (uint32_t k0 = 0; k0 < K; k0 += kSfUnroll) {
      VT w[kSfUnroll];
#pragma unroll
      for (int u = 0; u < kSfUnroll; u++)
        if (k0 + u < K)
          w[u] = __builtin_nontemporal_load((const VT *)(base + (uint64_t)(rs + (k0 + u) * rsub) * m));
#pragma unroll
      for (int u = 0; u < kSfUnroll; u++)
        if (k0 + u < K) {
#pragma unroll
          for (int v = 0; v < V; v++) acc[v] = F::add(acc[v], w[u][v]);
        }
    }
#pragma unroll
    for (int v = 0; v < V; v++) lds[rs * m + cg * V + v] = acc[v];
  }
  __syncthreads();
  if (tid < m) {
    T s = 0;
    for (uint32_t r = 0; r < rsub; r++) s = F::add(s, lds[r * m + tid]);
    sf_store(&part[(uint64_t)blockIdx.x * m + tid], s);
  }

  const uint32_t g = blockIdx.x / group, ngroups = gridDim.x / group;
  sf_release();
  __syncthreads();
  if (tid == 0) ticket = __hip_atomic_fetch_add(&cnt[g], 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  __syncthreads();
  if (ticket != group - 1) return;  // not the last of its group
  sf_acquire();
  {
    const T s = sf_sum_rows<F, kSfGroup>(part + (uint64_t)g * group * m, group, m, tid, lds);
    if (tid < m) sf_store(&part2[(uint64_t)g * m + tid], s);
  }
  sf_release();
  __syncthreads();
  if (tid == 0) ticket = __hip_atomic_fetch_add(&cnt[ngroups], 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  __syncthreads();
  if (ticket != ngroups - 1) return;  // not the last group
  sf_acquire();
  {
    const T s = sf_sum_rows<F, kSfMaxGroups>(part2, ngroups, m, tid, lds);
    __syncthreads();  // lds is the NTT's now
    if (tid < m) lds[tid] = s;
  }
  if (tid <= ngroups) cnt[tid] = 0;  // every workgroup has counted: clear for the next launch
  __syncthreads();
  for (uint32_t len = m; len >= 2; len >>= 1) {  // y -> NTT_m(y), bit-reversed order
    const uint32_t half = len >> 1;
    if (tid < m / 2) {
      const uint32_t j = tid & (half - 1);
      const uint32_t i0 = (tid - j) * 2 + j;  // block (tid / half) of length len, offset j
      const T u = lds[i0], v = lds[i0 + half];
      lds[i0] = F::add(u, v);
      lds[i0 + half] = F::mulw(F::sub(u, v), root_m[j * (m / len)]);  // j = 0: times root_m[0] = 1
    }
    __syncthreads();
  }
  if (tid < m) out[tid] = lds[__builtin_bitreverse32(tid) >> (32 - logm)];
}

// Contiguous row sums: out[v os + k] = sum_{t < n} part[(v m + k) n + t],
// grid (m, nvec), one 256-thread block per sum (coalesced loads, LDS tree).
template <class F>
__global__ void __launch_bounds__(256) k_sum_contig(const typename F::T *__restrict__ part,
                                                    typename F::T *__restrict__ out, uint64_t n, uint64_t m,
                                                    uint64_t os) {
  using T = typename F::T;
  __shared__ T red[256];
  const T *row = part + ((uint64_t)blockIdx.y * m + blockIdx.x) * n;
  T acc = 0;
  for (uint64_t t = threadIdx.x; t < n; t += 256) acc = F::add(acc, row[t]);
  red[threadIdx.x] = acc;
  __syncthreads();
  for (uint32_t s = 128; s > 0; s >>= 1) {
    if (threadIdx.x < s) red[threadIdx.x] = F::add(red[threadIdx.x], red[threadIdx.x + s]);
    __syncthreads();
  }
  if (threadIdx.x == 0) out[(uint64_t)blockIdx.y * os + blockIdx.x] = red[0];
}

// The same sums when each is short (n <= 32): one thread per sum, grid
// (ceil(m / 256), nvec); four independent chains over t.
template <class F>
__global__ void __launch_bounds__(256) k_sum_short(const typename F::T *__restrict__ part,
                                                   typename F::T *__restrict__ out, uint32_t n, uint64_t m,
                                                   uint64_t os) {
  using T = typename F::T;
  const uint64_t k = (uint64_t)blockIdx.x * 256 + threadIdx.x;
  if (k >= m) return;
  const T *row = part + ((uint64_t)blockIdx.y * m + k) * n;
  T acc[4] = {0, 0, 0, 0};
  uint32_t t = 0;
  for (; t + 4 <= n; t += 4) {
#pragma unroll
    for (int u = 0; u < 4; u++) acc[u] = F::add(acc[u], row[t + u]);
  }
  for (; t < n; t++) acc[0] = F::add(acc[0], row[t]);
  out[(uint64_t)blockIdx.y * os + k] = F::add(F::add(acc[0], acc[1]), F::add(acc[2], acc[3]));
}

// NttDomain's public tables (src/encoder.rs:14-18, built at :37-44):
// twiddles[k] = w^k for k < n and bitrev_twiddles[i] = w^bitrev(i, log n - 1)
// for i < n/2 (bitrev as src/encoder.rs:54-57), canonical, from the
// engine's two-level table of w; either output may be null.  One thread
// per k.
template <class F>
__global__ void __launch_bounds__(256) k_domain_tables(typename F::T *__restrict__ tw, typename F::T *__restrict__ brev,
                                                       uint64_t n, uint32_t logN, Tw2 t) {
  const uint64_t k = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= n) return;
  if (tw) tw[k] = F::from_w(tw2<F>(t, k));
  if (brev && k < n / 2) {
    const uint32_t lh = logN - 1;  // n >= 2 here
    const uint64_t e = lh ? (__builtin_bitreverse64(k) >> (64 - lh)) : 0;
    brev[k] = F::from_w(tw2<F>(t, e));
  }
}

// Elementwise field ops on canonical operands (device arithmetic unit test):
// op 0 add, 1 sub, 2 mul, 3 to_w then from_w round trip of a, 4 mulw(a, to_w(b)).
template <class F>
__global__ void __launch_bounds__(256) k_field_op(const typename F::T *__restrict__ a,
                                                  const typename F::T *__restrict__ b,
                                                  typename F::T *__restrict__ out, uint64_t n, int op) {
  const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  typename F::T r;
  if (op == 0) r = F::add(a[i], b[i]);
  else if (op == 1) r = F::sub(a[i], b[i]);
  else if (op == 2) r = F::mul(a[i], b[i]);
  else if (op == 3) r = F::from_w(F::to_w(a[i]));
  else r = F::mulw(a[i], F::to_w(b[i]));
  out[i] = r;
}

// out[i] = x[i] * w (multiplier form), used for standalone scaling / copies.
template <class F>
__global__ void __launch_bounds__(256) k_scale(const typename F::T *__restrict__ x,
                                               typename F::T *__restrict__ out, uint64_t n,
                                               typename F::T w, int apply) {
  const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  out[i] = apply ? F::mulw(x[i], w) : x[i];
}

__global__ void __launch_bounds__(256) k_narrow(const uint64_t *__restrict__ in, uint32_t *__restrict__ out,
                                                uint64_t n) {
  const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) out[i] = (uint32_t)in[i];
}

__global__ void __launch_bounds__(256) k_widen(const uint32_t *__restrict__ in, uint64_t *__restrict__ out,
                                               uint64_t n) {
  const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) out[i] = in[i];
}

__global__ void __launch_bounds__(256) k_fill_u32(uint32_t *p, uint64_t n, uint32_t v) {
  const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = v;
}

}  // namespace omr
