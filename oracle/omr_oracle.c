/*
 * omr_oracle.c -- CPU restatement of the omr-ntt reference algorithms.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the checker: tests/, the smoke
 * entry point and the CPU-baseline leg of bench.py may load it; the product
 * library (libomrntt.so) never links, loads or calls it.
 *
 * Every function names the reference file:line it restates (paths relative
 * to the benbencik/omr-ntt tree).  Field elements are canonical integers in
 * [0, p) held in uint64_t, for all four fields.
 *
 * Fields (src/fields.rs:3-17): 0 BabyBear, 1 KoalaBear, 2 Goldilocks,
 * 3 TeddyBear.  Arithmetic of the reference lives in ark-ff (git
 * arkworks-rs/algebra master#e341a1a, not vendored); its published behaviour
 * is integer arithmetic mod p, restated here as orc_add/orc_sub/orc_mul.
 * orc_mul_textbook is the plain 128-bit remainder used to pin orc_mul.
 *
 * Parity pinning (tests/test_oracle.py): brute-force DFT vs pure-Python
 * big-int golden vectors, closed-form roots of unity, cross-agreement of all
 * restated encoders, inverse round trips, linearity.
 */
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

typedef unsigned __int128 u128;

/* ---- fields ------------------------------------------------------------ */

static const uint64_t ORC_P[4] = {2013265921ull, 2130706433ull,
                                  18446744069414584321ull, 3221225473ull};
static const uint64_t ORC_G[4] = {31, 3, 7, 5};            /* src/fields.rs:4,7,10-14,17 */
static const int ORC_TWO_ADICITY[4] = {27, 24, 32, 30};  /* src/fields.rs:3,6,9,16 */

int orc_field_info(int f, uint64_t *p, uint64_t *g, int *two_adicity) {
  if (f < 0 || f > 3) return -1;
  if (p) *p = ORC_P[f];
  if (g) *g = ORC_G[f];
  if (two_adicity) *two_adicity = ORC_TWO_ADICITY[f];
  return 0;
}

static inline uint64_t fadd(int f, uint64_t a, uint64_t b) {
  uint64_t p = ORC_P[f];
  u128 s = (u128)a + b;
  if (s >= p) s -= p;
  return (uint64_t)s;
}
static inline uint64_t fsub(int f, uint64_t a, uint64_t b) {
  uint64_t p = ORC_P[f];
  return a >= b ? a - b : a + (p - b);
}
static inline uint64_t fneg(int f, uint64_t a) { return a == 0 ? 0 : ORC_P[f] - a; }

/* Goldilocks: 2^64 = 2^32 - 1 (mod p), 2^96 = -1 (mod p). */
static inline uint64_t gl_reduce128(u128 x) {
  const uint64_t P = 18446744069414584321ull, EPS = 0xFFFFFFFFull;
  uint64_t lo = (uint64_t)x, hi = (uint64_t)(x >> 64);
  uint64_t hh = hi >> 32, hl = hi & EPS;
  uint64_t t0 = lo - hh;
  if (lo < hh) t0 -= EPS;
  uint64_t t1 = (hl << 32) - hl;
  uint64_t r = t0 + t1;
  if (r < t1) r += EPS;
  if (r >= P) r -= P;
  return r;
}

/* 31/32-bit fields: the remainder by a constant modulus (the compiler turns
 * it into a multiply-shift), the same value as (a * b) % p. */
static inline uint64_t fmul(int f, uint64_t a, uint64_t b) {
  switch (f) {
    case 0: return (a * b) % 2013265921ull;
    case 1: return (a * b) % 2130706433ull;
    case 2: return gl_reduce128((u128)a * b);
    default: return (a * b) % 3221225473ull;
  }
}

static uint64_t fpow(int f, uint64_t a, uint64_t e) {
  uint64_t r = 1;
  while (e) {
    if (e & 1) r = fmul(f, r, a);
    a = fmul(f, a, a);
    e >>= 1;
  }
  return r;
}
static uint64_t finv(int f, uint64_t a) { return fpow(f, a, ORC_P[f] - 2); }

uint64_t orc_add(int f, uint64_t a, uint64_t b) { return fadd(f, a, b); }
uint64_t orc_sub(int f, uint64_t a, uint64_t b) { return fsub(f, a, b); }
uint64_t orc_mul(int f, uint64_t a, uint64_t b) { return fmul(f, a, b); }
uint64_t orc_mul_textbook(int f, uint64_t a, uint64_t b) {
  return (uint64_t)(((u128)a * b) % ORC_P[f]);
}
uint64_t orc_pow(int f, uint64_t a, uint64_t e) { return fpow(f, a, e); }
uint64_t orc_inv(int f, uint64_t a) { return finv(f, a); }

/* ark-ff FftField::get_root_of_unity(n) for n = 2^k:
 * TWO_ADIC_ROOT_OF_UNITY = g^t (p - 1 = 2^s t), squared s - k times.
 * Called at src/encoder.rs:34-35. Returns 0 if n is not a power of two
 * within the two-adicity. */
uint64_t orc_root_of_unity(int f, uint64_t n) {
  if (n == 0 || (n & (n - 1))) return 0;
  int k = __builtin_ctzll(n);
  int s = ORC_TWO_ADICITY[f];
  if (k > s) return 0;
  uint64_t t = (ORC_P[f] - 1) >> s;
  uint64_t w = fpow(f, ORC_G[f], t);
  for (int i = k; i < s; i++) w = fmul(f, w, w);
  return w;
}

/* src/encoder.rs:54-57 and src/encoders/utils.rs:117-119:
 * a.reverse_bits() >> (64 - log_len).  The 64-bit reversal swaps adjacent
 * bits, pairs and nibbles, then the bytes. */
static inline uint64_t rev64(uint64_t x) {
  x = ((x >> 1) & 0x5555555555555555ull) | ((x & 0x5555555555555555ull) << 1);
  x = ((x >> 2) & 0x3333333333333333ull) | ((x & 0x3333333333333333ull) << 2);
  x = ((x >> 4) & 0x0F0F0F0F0F0F0F0Full) | ((x & 0x0F0F0F0F0F0F0F0Full) << 4);
  return __builtin_bswap64(x);
}
static inline uint64_t bitrev(uint64_t a, int log_len) {
  if (log_len == 0) return 0;
  return rev64(a) >> (64 - log_len);
}
uint64_t orc_bitrev(uint64_t a, int log_len) { return bitrev(a, log_len); }

/* NttDomain::new (src/encoder.rs:23-52): twiddles[k] = w^k, k < N (powers,
 * :60-68); bitrev_twiddles[i] = twiddles[bitrev(i, logN - 1)], i < N/2. */
void orc_domain(int f, uint64_t n, uint64_t omega, uint64_t *twiddles,
                uint64_t *bitrev_twiddles) {
  uint64_t cur = 1;
  for (uint64_t k = 0; k < n; k++) {
    twiddles[k] = cur;
    cur = fmul(f, cur, omega);
  }
  if (bitrev_twiddles && n >= 2) {
    int log_half = __builtin_ctzll(n) - 1;
    for (uint64_t i = 0; i < n / 2; i++) bitrev_twiddles[i] = twiddles[bitrev(i, log_half)];
  }
}

/* ---- permutations and radix-2 kernels (src/encoders/utils.rs) ---------- */

/* derange (utils.rs:85-101): swap each bit-reversed pair once, in parallel
 * over the indices like the reference's par_iter (index idx owns the pair
 * (idx, bitrev(idx)) with idx < bitrev(idx), so no two threads touch one
 * element).  Below 2^16 elements it runs serially, as derange_seq
 * (utils.rs:103-113) does inside the row transforms. */
void orc_derange(uint64_t *xi, uint64_t n, int log_len) {
  if (n < 3) return;
#pragma omp parallel for schedule(static) if (n >= (1u << 16))
  for (int64_t i = 1; i < (int64_t)n - 1; i++) {
    const uint64_t idx = (uint64_t)i, r = bitrev(idx, log_len);
    if (idx < r) {
      uint64_t t = xi[idx];
      xi[idx] = xi[r];
      xi[r] = t;
    }
  }
}

/* layer_twiddles (utils.rs:5-16): per-layer contiguous table of n1 - 1. */
static void layer_twiddles(uint64_t n1, uint64_t n2, const uint64_t *tw, uint64_t *out) {
  uint64_t pos = 0;
  for (uint64_t layer = 1; layer < n1; layer *= 2) {
    uint64_t step = (n1 / (2 * layer)) * n2;
    for (uint64_t j = 0; j < layer; j++) out[pos++] = tw[j * step];
  }
}

/* inplace_radix2_dit (utils.rs:26-55).  Serial: the reference runs it on
 * each row inside a parallel loop. */
static void radix2_dit(int f, uint64_t *xi, uint64_t n, const uint64_t *ltw) {
  if (n < 2) return;
  orc_derange(xi, n, __builtin_ctzll(n));
  for (uint64_t length = 2; length <= n; length *= 2) {
    uint64_t half = length / 2;
    const uint64_t *cur = ltw + (half - 1);
    for (uint64_t b = 0; b < n; b += length) {
      uint64_t *lo = xi + b, *hi = xi + b + half;
      for (uint64_t j = 0; j < half; j++) {
        uint64_t u = lo[j], v = fmul(f, hi[j], cur[j]);
        lo[j] = fadd(f, u, v);
        hi[j] = fsub(f, u, v);
      }
    }
  }
}

/* The same transform with each layer's butterflies spread over the host's
 * threads (the ArkRadix2 role: ark-poly's radix-2 FFT with its "parallel"
 * feature, Cargo.toml:9, splits every layer's butterflies over rayon):
 * over the blocks while a layer has at least 64 of them, else over j
 * inside each block.  Same butterflies, same order within each one. */
static void radix2_dit_par(int f, uint64_t *xi, uint64_t n, const uint64_t *ltw) {
  if (n < (1u << 12)) {
    radix2_dit(f, xi, n, ltw);
    return;
  }
  orc_derange(xi, n, __builtin_ctzll(n));
#pragma omp parallel
  for (uint64_t length = 2; length <= n; length *= 2) {
    const uint64_t half = length / 2, nblk = n / length;
    const uint64_t *cur = ltw + (half - 1);
    if (nblk >= 64) {
#pragma omp for schedule(static)
      for (int64_t bb = 0; bb < (int64_t)nblk; bb++) {
        uint64_t *lo = xi + (uint64_t)bb * length, *hi = lo + half;
        for (uint64_t j = 0; j < half; j++) {
          uint64_t u = lo[j], v = fmul(f, hi[j], cur[j]);
          lo[j] = fadd(f, u, v);
          hi[j] = fsub(f, u, v);
        }
      }
    } else {
      for (uint64_t b = 0; b < n; b += length) {
        uint64_t *lo = xi + b, *hi = xi + b + half;
#pragma omp for schedule(static)
        for (int64_t jj = 0; jj < (int64_t)half; jj++) {
          const uint64_t j = (uint64_t)jj;
          uint64_t u = lo[j], v = fmul(f, hi[j], cur[j]);
          lo[j] = fadd(f, u, v);
          hi[j] = fsub(f, u, v);
        }
      }
    }
  }
}

/* inplace_radix2_dif_no_derange (utils.rs:58-83). */
static void radix2_dif_no_derange(int f, uint64_t *xi, uint64_t n, const uint64_t *ltw) {
  if (n < 2) return;
  for (uint64_t length = n; length >= 2; length /= 2) {
    uint64_t half = length / 2;
    const uint64_t *cur = ltw + (half - 1);
    for (uint64_t b = 0; b < n; b += length) {
      uint64_t *lo = xi + b, *hi = xi + b + half;
      for (uint64_t j = 0; j < half; j++) {
        uint64_t u = lo[j], v = hi[j];
        lo[j] = fadd(f, u, v);
        hi[j] = fmul(f, fsub(f, u, v), cur[j]);
      }
    }
  }
}

/* transpose_par (transpose_out_of_place.rs:19-29, 340-366):
 * out[y + x*H] = in[x + y*W] for x < W, y < H.  Like the reference it works
 * in 16 x 16 tiles and, for large matrices, on all host threads (the
 * reference splits recursively with rayon::join; here the tiles are shared
 * out by OpenMP).  Same index map, same result. */
void orc_transpose(const uint64_t *in, uint64_t *out, uint64_t w, uint64_t h) {
  const uint64_t B = 16, wb = (w + B - 1) / B, hb = (h + B - 1) / B;
#pragma omp parallel for collapse(2) schedule(static) if (w * h >= (1u << 16))
  for (int64_t ty = 0; ty < (int64_t)hb; ty++)
    for (int64_t tx = 0; tx < (int64_t)wb; tx++) {
      const uint64_t y0 = (uint64_t)ty * B, x0 = (uint64_t)tx * B;
      const uint64_t y1 = y0 + B < h ? y0 + B : h, x1 = x0 + B < w ? x0 + B : w;
      for (uint64_t y = y0; y < y1; y++)
        for (uint64_t x = x0; x < x1; x++) out[y + x * h] = in[x + y * w];
    }
}

/* ---- full-transform encoders ------------------------------------------- */

/* Textbook radix-2 DIT on the domain's twiddles; the role of ArkRadix2
 * (src/encoders/ark_radix2.rs:13-19, natural order in and out). */
void orc_radix2(int f, uint64_t *buf, uint64_t n, const uint64_t *twiddles) {
  if (n < 2) return;
  uint64_t *ltw = (uint64_t *)malloc(sizeof(uint64_t) * n);
  layer_twiddles(n, 1, twiddles, ltw);
  radix2_dit_par(f, buf, n, ltw);
  free(ltw);
}

/* Dft (src/encoders/dft.rs:9-22): out[j] = sum_i in[i] w^(ij). */
void orc_dft(int f, uint64_t *buf, uint64_t n, uint64_t omega) {
  uint64_t *in = (uint64_t *)malloc(sizeof(uint64_t) * n);
  memcpy(in, buf, sizeof(uint64_t) * n);
  for (uint64_t j = 0; j < n; j++) {
    uint64_t step = fpow(f, omega, j), t = 1, acc = 0;
    for (uint64_t i = 0; i < n; i++) {
      acc = fadd(f, acc, fmul(f, in[i], t));
      t = fmul(f, t, step);
    }
    buf[j] = acc;
  }
  free(in);
}

/* DftPartial (src/encoders/dft.rs:40-88): buf[j] for j < min(2s, N), chunked
 * partial sums over CHUNK = 2^14 inputs; buf[out_len..] keeps the input. */
void orc_dft_partial(int f, uint64_t *buf, uint64_t n, uint64_t omega, uint64_t s) {
  uint64_t out_len = 2 * s < n ? 2 * s : n;
  if (out_len == 0) return;
  const uint64_t CHUNK = 1ull << 14;
  uint64_t *acc = (uint64_t *)calloc(out_len, sizeof(uint64_t));
  for (uint64_t start = 0; start < n; start += CHUNK) {
    uint64_t end = start + CHUNK < n ? start + CHUNK : n;
    for (uint64_t j = 0; j < out_len; j++) {
      uint64_t step = fpow(f, omega, j);
      uint64_t t = fpow(f, omega, start * j);
      uint64_t sum = 0;
      for (uint64_t i = start; i < end; i++) {
        sum = fadd(f, sum, fmul(f, buf[i], t));
        t = fmul(f, t, step);
      }
      acc[j] = fadd(f, acc[j], sum);
    }
  }
  memcpy(buf, acc, sizeof(uint64_t) * out_len);
  free(acc);
}

/* LambdaRadix4 (src/encoders/lambdaworks_radix4.rs:16-24, 35-74):
 * radix-4 DIF with bit-reversed twiddles, then derange. log N even. */
int orc_lambda_radix4(int f, uint64_t *in, uint64_t n, const uint64_t *btw) {
  int log_n = __builtin_ctzll(n);
  if (log_n % 2) return -1;
  uint64_t group_count = 1, group_size = n;
  while (group_count < n) {
    for (uint64_t g = 0; g < group_count; g++) {
      uint64_t first = g * group_size, next = first + group_size / 4;
      uint64_t w1 = btw[g], w2 = btw[2 * g], w3 = btw[2 * g + 1];
      for (uint64_t i = first; i < next; i++) {
        uint64_t j = i + group_size / 4, k = i + group_size / 2, l = i + 3 * group_size / 4;
        uint64_t zw1 = fmul(f, w1, in[k]);
        uint64_t tw1 = fmul(f, w1, in[l]);
        uint64_t a = fmul(f, w2, fadd(f, in[j], tw1));
        uint64_t b = fmul(f, w3, fsub(f, in[j], tw1));
        uint64_t ip = fadd(f, in[i], zw1), im = fsub(f, in[i], zw1);
        in[i] = fadd(f, ip, a);
        in[j] = fsub(f, ip, a);
        in[k] = fadd(f, im, b);
        in[l] = fsub(f, im, b);
      }
    }
    group_count *= 4;
    group_size /= 4;
  }
  orc_derange(in, n, log_n);
  return 0;
}

/* LambdaBowers (src/encoders/lambdaworks_bowers.rs:15-19, 42-172). */
static void bowers_single(int f, uint64_t *blk, const uint64_t *tw, uint64_t half) {
  if (half > 0) {
    uint64_t s = fadd(f, blk[0], blk[half]), d = fsub(f, blk[0], blk[half]);
    blk[0] = s;
    blk[half] = d;
  }
  for (uint64_t j = 1; j < half; j++) {
    uint64_t s = fadd(f, blk[j], blk[j + half]), d = fsub(f, blk[j], blk[j + half]);
    blk[j] = s;
    blk[j + half] = fmul(f, tw[j], d);
  }
}
static void bowers_fused(int f, uint64_t *blk, uint64_t len, const uint64_t *t0, const uint64_t *t1) {
  uint64_t q = len >> 2;
  {
    uint64_t w1 = t0[q];
    uint64_t s02 = fadd(f, blk[0], blk[2 * q]), d02 = fsub(f, blk[0], blk[2 * q]);
    uint64_t s13 = fadd(f, blk[q], blk[3 * q]), d13 = fsub(f, blk[q], blk[3 * q]);
    uint64_t d13w = fmul(f, w1, d13);
    blk[0] = fadd(f, s02, s13);
    blk[q] = fsub(f, s02, s13);
    blk[2 * q] = fadd(f, d02, d13w);
    blk[3 * q] = fsub(f, d02, d13w);
  }
  for (uint64_t j = 1; j < q; j++) {
    uint64_t w0 = t0[j], w1 = t0[j + q], w2 = t1[j];
    uint64_t s02 = fadd(f, blk[j], blk[j + 2 * q]), d02 = fsub(f, blk[j], blk[j + 2 * q]);
    uint64_t d02w = fmul(f, w0, d02);
    uint64_t s13 = fadd(f, blk[j + q], blk[j + 3 * q]), d13 = fsub(f, blk[j + q], blk[j + 3 * q]);
    uint64_t d13w = fmul(f, w1, d13);
    blk[j] = fadd(f, s02, s13);
    blk[j + q] = fmul(f, w2, fsub(f, s02, s13));
    blk[j + 2 * q] = fadd(f, d02w, d13w);
    blk[j + 3 * q] = fmul(f, w2, fsub(f, d02w, d13w));
  }
}
void orc_lambda_bowers(int f, uint64_t *in, uint64_t n, const uint64_t *twiddles) {
  int log_n = __builtin_ctzll(n);
  if (n <= 1) return;
  /* layer_twiddles (lambdaworks_bowers.rs:28-37): layer k has N >> (k+1)
   * entries tw[i << k]. */
  uint64_t **layers = (uint64_t **)malloc(sizeof(uint64_t *) * log_n);
  for (int k = 0; k < log_n; k++) {
    uint64_t cnt = n >> (k + 1);
    layers[k] = (uint64_t *)malloc(sizeof(uint64_t) * (cnt ? cnt : 1));
    for (uint64_t i = 0; i < cnt; i++) layers[k][i] = twiddles[i << k];
  }
  int layer = 0;
  if (n <= 4) {
    for (; layer < log_n; layer++) {
      uint64_t bs = n >> layer;
      for (uint64_t b = 0; b < n; b += bs) bowers_single(f, in + b, layers[layer], bs >> 1);
    }
  } else {
    while (layer + 1 < log_n) {
      uint64_t bs = n >> layer;
      if (bs < 4) break;
      for (uint64_t b = 0; b < n; b += bs) bowers_fused(f, in + b, bs, layers[layer], layers[layer + 1]);
      layer += 2;
    }
    for (; layer < log_n; layer++) {
      uint64_t bs = n >> layer;
      for (uint64_t b = 0; b < n; b += bs) bowers_single(f, in + b, layers[layer], bs >> 1);
    }
  }
  for (int k = 0; k < log_n; k++) free(layers[k]);
  free(layers);
  orc_derange(in, n, log_n);
}

/* Plonky3Radix2LayerSplit (src/encoders/plonky3_radix2_layer_split.rs:30-97). */
void orc_plonky3_layer_split(int f, uint64_t *a, uint64_t n, const uint64_t *twiddles,
                             const uint64_t *btw) {
  int log_n = __builtin_ctzll(n);
  int mid = (log_n + 1) / 2;
  orc_derange(a, n, log_n);
  uint64_t cs = 1ull << mid;
#pragma omp parallel for schedule(static)
  for (int64_t c = 0; c < (int64_t)(n / cs); c++) {
    uint64_t *chunk = a + (uint64_t)c * cs;
    for (int layer = 0; layer < mid; layer++) {
      uint64_t hb = 1ull << layer, bs = hb * 2, step = n >> (layer + 1);
      for (uint64_t b0 = 0; b0 < cs; b0 += bs)
        for (uint64_t j = 0; j < hb; j++) {
          uint64_t lo = chunk[b0 + j], hi = fmul(f, chunk[b0 + j + hb], twiddles[j * step]);
          chunk[b0 + j] = fadd(f, lo, hi);
          chunk[b0 + j + hb] = fsub(f, lo, hi);
        }
    }
  }
  orc_derange(a, n, log_n);
  uint64_t cs2 = 1ull << (log_n - mid);
#pragma omp parallel for schedule(static)
  for (int64_t th = 0; th < (int64_t)(n / cs2); th++) {
    uint64_t *chunk = a + (uint64_t)th * cs2;
    for (int layer = mid; layer < log_n; layer++) {
      uint64_t first_block = (uint64_t)th << (layer - mid);
      int lr = log_n - 1 - layer;
      uint64_t hb = 1ull << lr, bs = hb * 2;
      for (uint64_t b = 0; b * bs < cs2; b++) {
        uint64_t tw = btw[first_block + b];
        uint64_t *blk = chunk + b * bs;
        for (uint64_t i = 0; i < hb; i++) {
          uint64_t lo = blk[i], hi = fmul(f, blk[i + hb], tw);
          blk[i] = fadd(f, lo, hi);
          blk[i + hb] = fsub(f, lo, hi);
        }
      }
    }
  }
  orc_derange(a, n, log_n);
}

/* fft_in_place (src/encoders/winterfell_four_step.rs:86-130): recursive
 * split producing bit-reversed output, bit-reversed twiddle table. */
static void wf_fft_in_place(int f, uint64_t *v, uint64_t len, const uint64_t *tw, uint64_t count,
                            uint64_t stride, uint64_t offset) {
  const uint64_t MAX_LOOP = 256;
  uint64_t size = len / stride;
  if (size > 2) {
    if (stride == count && count < MAX_LOOP) {
      wf_fft_in_place(f, v, len, tw, 2 * count, 2 * stride, offset);
    } else {
      wf_fft_in_place(f, v, len, tw, count, 2 * stride, offset);
      wf_fft_in_place(f, v, len, tw, count, 2 * stride, offset + stride);
    }
  }
  for (uint64_t i = offset; i < offset + count; i++) {
    uint64_t j = i + stride, t = v[i];
    v[i] = fadd(f, t, v[j]);
    v[j] = fsub(f, t, v[j]);
  }
  uint64_t last = offset + size * stride, idx = 0;
  for (uint64_t off = offset; off < last; off += 2 * stride, idx++) {
    if (idx == 0) continue;
    uint64_t w = tw[idx];
    for (uint64_t j = off; j < off + count; j++) {
      uint64_t t = v[j];
      v[j + stride] = fmul(f, v[j + stride], w);
      uint64_t s = v[j + stride];
      v[j] = fadd(f, t, s);
      v[j + stride] = fsub(f, t, s);
    }
  }
}

/* WinterfellFourStep (winterfell_four_step.rs:21-24, 31-81). */
void orc_winterfell_four_step(int f, uint64_t *values, uint64_t n, const uint64_t *btw) {
  if (n <= 1) return;
  if (n <= 2) {
    wf_fft_in_place(f, values, n, btw, 1, 1, 0);
    orc_derange(values, n, n > 1 ? __builtin_ctzll(n) : 0);
    return;
  }
  int log_n = __builtin_ctzll(n);
  uint64_t inner = 1ull << (log_n / 2), outer = n / inner;
  uint64_t g = btw[(n / 2) / 2];
  int log_inner = __builtin_ctzll(inner);
  uint64_t *scratch = (uint64_t *)calloc(n, sizeof(uint64_t));
  orc_transpose(values, scratch, outer, inner);
#pragma omp parallel for schedule(static)
  for (int64_t r = 0; r < (int64_t)outer; r++) wf_fft_in_place(f, scratch + (uint64_t)r * inner, inner, btw, 1, 1, 0);
  orc_transpose(scratch, values, inner, outer);
#pragma omp parallel for schedule(static)
  for (int64_t i = 0; i < (int64_t)inner; i++) {
    uint64_t *row = values + (uint64_t)i * outer;
    if (i > 0) {
      uint64_t ip = bitrev((uint64_t)i, log_inner);
      uint64_t itw = fpow(f, g, ip), otw = itw;
      for (uint64_t e = 1; e < outer; e++) {
        row[e] = fmul(f, row[e], otw);
        otw = fmul(f, otw, itw);
      }
    }
    wf_fft_in_place(f, row, outer, btw, 1, 1, 0);
  }
  free(scratch);
  orc_derange(values, n, log_n);
}

/* TfheStockhamRadix8 (src/encoders/tfhe_stockham_radix8.rs:16-24, 47-139).
 * log N divisible by 3. */
int orc_tfhe_stockham_radix8(int f, uint64_t *buf, uint64_t n, const uint64_t *twd) {
  int log_n = __builtin_ctzll(n);
  if (log_n % 3) return -1;
  if (n == 1) return 0;
  uint64_t imag = fneg(f, twd[n / 4]), w8 = twd[n / 8], v8 = fneg(f, twd[3 * n / 8]);
  uint64_t *src = (uint64_t *)malloc(sizeof(uint64_t) * n);
  uint64_t *dst = (uint64_t *)calloc(n, sizeof(uint64_t));
  memcpy(src, buf, sizeof(uint64_t) * n);
  for (uint64_t s = n / 8; s >= 1; s /= 8) {
    uint64_t groups = n / (8 * s);
    for (uint64_t p = 0; p < groups; p++) {
      uint64_t tb = p * s, tw[8];
      for (int k = 0; k < 8; k++) tw[k] = twd[k * tb];
      for (uint64_t j = 0; j < s; j++) {
        uint64_t base = 8 * s * p + j, y[8];
        for (int k = 0; k < 8; k++) y[k] = fmul(f, tw[k], src[base + k * s]);
        uint64_t a04 = fadd(f, y[0], y[4]), s04 = fsub(f, y[0], y[4]);
        uint64_t a26 = fadd(f, y[2], y[6]), is26 = fmul(f, imag, fsub(f, y[2], y[6]));
        uint64_t a15 = fadd(f, y[1], y[5]), s15 = fsub(f, y[1], y[5]);
        uint64_t a37 = fadd(f, y[3], y[7]), is37 = fmul(f, imag, fsub(f, y[3], y[7]));
        uint64_t A = fadd(f, a04, a26), B = fadd(f, a15, a37);
        uint64_t Cc = fsub(f, s04, is26), D = fsub(f, s15, is37);
        uint64_t E = fsub(f, a04, a26), Fm = fmul(f, imag, fsub(f, a15, a37));
        uint64_t Gp = fadd(f, s04, is26), H = fadd(f, s15, is37);
        uint64_t o[8];
        o[0] = fadd(f, A, B);
        o[1] = fadd(f, Cc, fmul(f, w8, D));
        o[2] = fsub(f, E, Fm);
        o[3] = fsub(f, Gp, fmul(f, v8, H));
        o[4] = fsub(f, A, B);
        o[5] = fsub(f, Cc, fmul(f, w8, D));
        o[6] = fadd(f, E, Fm);
        o[7] = fadd(f, Gp, fmul(f, v8, H));
        uint64_t ob = s * p + j;
        for (int k = 0; k < 8; k++) dst[k * (n / 8) + ob] = o[k];
      }
    }
    uint64_t *t = src;
    src = dst;
    dst = t;
    if (s == 1) break;
  }
  memcpy(buf, src, sizeof(uint64_t) * n);
  free(src);
  free(dst);
  return 0;
}

/* ---- partial-output encoders ------------------------------------------- */

static uint64_t next_pow2(uint64_t x) {
  uint64_t r = 1;
  while (r < x) r <<= 1;
  return r;
}

/* TransformDecomposition (src/encoders/transform_decomposition.rs:26-97)
 * and V2 (transform_decomposition_v2.rs:24-100, v2 != 0): buf[k] for
 * k < min(2s, N) = sum_r w^(rk) NTT_n1(x[r::n2])[k]. */
void orc_transform_decomposition(int f, uint64_t *buf, uint64_t n, uint64_t omega,
                                 const uint64_t *twiddles, uint64_t s, int v2) {
  uint64_t out_len = 2 * s < n ? 2 * s : n;
  if (out_len == 0) return;
  uint64_t n1 = next_pow2(2 * out_len);
  if (n1 < 2) n1 = 2;
  if (n1 > n) n1 = n;
  uint64_t n2 = n / n1;
  int log1 = __builtin_ctzll(n1);
  uint64_t *ltw = (uint64_t *)malloc(sizeof(uint64_t) * (n1 > 1 ? n1 : 2));
  layer_twiddles(n1, n2, twiddles, ltw);
  uint64_t *tr = (uint64_t *)malloc(sizeof(uint64_t) * n);
  orc_transpose(buf, tr, n2, n1);
#pragma omp parallel for schedule(static)
  for (int64_t r = 0; r < (int64_t)n2; r++) {
    if (v2) radix2_dif_no_derange(f, tr + (uint64_t)r * n1, n1, ltw);
    else radix2_dit(f, tr + (uint64_t)r * n1, n1, ltw);
  }
  uint64_t *acc = (uint64_t *)calloc(out_len, sizeof(uint64_t));
#pragma omp parallel
  {
    uint64_t *local = (uint64_t *)calloc(out_len, sizeof(uint64_t));
#pragma omp for schedule(static)
    for (int64_t r = 0; r < (int64_t)n2; r++) {
      uint64_t texp = fpow(f, omega, (uint64_t)r), tstep = 1;
      const uint64_t *row = tr + (uint64_t)r * n1;
      for (uint64_t j = 0; j < out_len; j++) {
        uint64_t v = v2 ? row[bitrev(j, log1)] : row[j];
        local[j] = fadd(f, local[j], fmul(f, v, tstep));
        tstep = fmul(f, tstep, texp);
      }
    }
#pragma omp critical
    for (uint64_t j = 0; j < out_len; j++) acc[j] = fadd(f, acc[j], local[j]);
    free(local);
  }
  memcpy(buf, acc, sizeof(uint64_t) * out_len);
  free(acc);
  free(tr);
  free(ltw);
}

/* ---- operations the reference does not define (closed forms) ----------- */

/* Inverse: x[j] = N^-1 sum_k X[k] w^(-jk). */
void orc_intt(int f, uint64_t *buf, uint64_t n, uint64_t omega) {
  if (n < 2) return;
  uint64_t winv = finv(f, omega);
  uint64_t *tw = (uint64_t *)malloc(sizeof(uint64_t) * n);
  orc_domain(f, n, winv, tw, NULL);
  orc_radix2(f, buf, n, tw);
  uint64_t ninv = finv(f, n % ORC_P[f]);
  for (uint64_t i = 0; i < n; i++) buf[i] = fmul(f, buf[i], ninv);
  free(tw);
}

/* Forward NTT through the textbook radix-2 path with a given omega. */
void orc_ntt(int f, uint64_t *buf, uint64_t n, uint64_t omega) {
  if (n < 2) return;
  uint64_t *tw = (uint64_t *)malloc(sizeof(uint64_t) * n);
  orc_domain(f, n, omega, tw, NULL);
  orc_radix2(f, buf, n, tw);
  free(tw);
}

/* A batch of vectors of uneven power-of-two sizes: vector i is the
 * 2^log_n[i] elements at buf + off[i], forward (inv[i] == 0) or inverse
 * transformed in place; vectors spread over the host threads. */
void orc_ntt_batch(int f, uint64_t *buf, const uint64_t *off, const uint32_t *log_n, const uint32_t *inv,
                   uint64_t count) {
#pragma omp parallel for schedule(dynamic, 1)
  for (int64_t i = 0; i < (int64_t)count; i++) {
    const uint64_t n = 1ull << log_n[i];
    const uint64_t w = orc_root_of_unity(f, n);
    if (inv[i]) orc_intt(f, buf + off[i], n, w);
    else orc_ntt(f, buf + off[i], n, w);
  }
}

/* Coset low-degree extension: out[i] = sum_j c[j] (shift * w_M^i)^j for
 * i < M = blowup * n, w_M the ark-convention root of unity of order M. */
int orc_coset_lde(int f, const uint64_t *coeffs, uint64_t n, uint64_t blowup, uint64_t shift,
                  uint64_t *out) {
  uint64_t m = n * blowup;
  uint64_t w = orc_root_of_unity(f, m);
  if (!w) return -1;
  memset(out, 0, sizeof(uint64_t) * m);
  uint64_t gp = 1;
  for (uint64_t j = 0; j < n; j++) {
    out[j] = fmul(f, coeffs[j], gp);
    gp = fmul(f, gp, shift);
  }
  orc_ntt(f, out, m, w);
  return 0;
}

/* Decode of a coset LDE: coefficients c[j] = shift^-j INTT_M(evals)[j];
 * returns the number of nonzero coefficients at j >= n (0 for a codeword). */
int64_t orc_coset_decode(int f, const uint64_t *evals, uint64_t n, uint64_t blowup, uint64_t shift,
                         uint64_t *coeffs) {
  uint64_t m = n * blowup;
  uint64_t w = orc_root_of_unity(f, m);
  if (!w) return -1;
  uint64_t *t = (uint64_t *)malloc(sizeof(uint64_t) * m);
  memcpy(t, evals, sizeof(uint64_t) * m);
  orc_intt(f, t, m, w);
  uint64_t sinv = finv(f, shift), gp = 1;
  int64_t bad = 0;
  for (uint64_t j = 0; j < m; j++) {
    uint64_t c = fmul(f, t[j], gp);
    gp = fmul(f, gp, sinv);
    if (j < n) coeffs[j] = c;
    else if (c) bad++;
  }
  free(t);
  return bad;
}

/* Stride subset of the full transform: out[j] = X[j * N/m] for j < m,
 * via the fold identity out = NTT_m(fold_m(x)), w_m = w^(N/m). */
void orc_stride_subset(int f, const uint64_t *x, uint64_t n, uint64_t omega, uint64_t m,
                       uint64_t *out) {
  memset(out, 0, sizeof(uint64_t) * m);
  for (uint64_t i = 0; i < n; i++) out[i % m] = fadd(f, out[i % m], x[i]);
  orc_ntt(f, out, m, fpow(f, omega, n / m));
}

/* Elementwise helpers for tests of the device field arithmetic. */
void orc_vec_op(int f, int op, const uint64_t *a, const uint64_t *b, uint64_t *out, uint64_t n) {
  for (uint64_t i = 0; i < n; i++) {
    if (op == 0) out[i] = fadd(f, a[i], b[i]);
    else if (op == 1) out[i] = fsub(f, a[i], b[i]);
    else out[i] = fmul(f, a[i], b[i]);
  }
}

int orc_num_threads(void) {
#ifdef _OPENMP
  return omp_get_max_threads();
#else
  return 1;
#endif
}
void orc_set_num_threads(int t) {
#ifdef _OPENMP
  if (t > 0) omp_set_num_threads(t);
#else
  (void)t;
#endif
}
