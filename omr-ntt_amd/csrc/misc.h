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
