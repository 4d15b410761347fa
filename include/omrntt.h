/*
 * omrntt.h -- C ABI of the MI355X (gfx950) NTT engine.
 *
 * Drop-in boundary for the reference crate's operator API (benbencik/omr-ntt):
 *   trait NttEncoder<F> { fn ntt(&self, buf: &mut [F], domain: &NttDomain<F>);
 *                         fn name(&self) -> &str; }      (src/encoder.rs:70-74)
 * plus the domain checks of NttDomain::new (src/encoder.rs:23-52) and the
 * field set of src/fields.rs:3-17.
 *
 * Conventions
 *  - Field elements cross this boundary as canonical integers in [0, p):
 *    uint32_t for BabyBear/KoalaBear/TeddyBear, uint64_t for Goldilocks.
 *  - `omega` is the caller's primitive N-th root of unity (the Rust shim
 *    passes NttDomain::omega, canonical), so results follow the caller's
 *    domain exactly.  omr_root_of_unity gives the ark-ff convention
 *    g^((p-1)/N).
 *  - Every function validates all arguments before touching the GPU and
 *    returns an omr_status; nothing panics or aborts, and no C++ exception
 *    leaves the library (host allocation failure: OMR_E_ALLOC).
 *  - `_device` entry points take device pointers and a hipStream_t (passed as
 *    void*, NULL = default stream) and are asynchronous on that stream.
 *    Calls sharing one omr_ctx share its scratch buffers, so the context
 *    orders them: a call on a stream other than the previous call's first
 *    waits until the device is idle (calls on one stream are not slowed).
 *    For concurrent streams use one context per stream.
 *  - The host-buffer entry point omr_ntt is synchronous and thread safe
 *    (it holds the context lock), matching `NttEncoder: Send + Sync`.
 */
#ifndef OMRNTT_H
#define OMRNTT_H
#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef enum {
  OMR_OK = 0,
  OMR_E_NULL = 1,         /* null pointer argument */
  OMR_E_SIZE = 2,         /* N not a power of two   (src/encoder.rs:24) */
  OMR_E_TWO_ADICITY = 3,  /* log2 N > two-adicity   (src/encoder.rs:28-32) */
  OMR_E_FIELD = 4,        /* unknown field id */
  OMR_E_VARIANT = 5,      /* unknown encoder variant */
  OMR_E_RADIX = 6,        /* variant constraint (lambdaworks_radix4.rs:17-21,
                             tfhe_stockham_radix8.rs:17-21) */
  OMR_E_OMEGA = 7,        /* omega is not a primitive N-th root of unity */
  OMR_E_OUTLEN = 8,       /* invalid partial-output length / subset */
  OMR_E_ALLOC = 9,        /* device or host allocation failed */
  OMR_E_DEVICE = 10,      /* HIP runtime or launch error */
  OMR_E_SCHEDULE = 11,    /* explicit schedule invalid for this N / field */
  OMR_E_ARG = 12,         /* other invalid argument */
  OMR_E_SCRATCH = 13      /* request exceeds the fixed scratch of omr_ntt_device_noalloc */
} omr_status;

typedef enum {
  OMR_BABYBEAR = 0,   /* p = 2^31 - 2^27 + 1, g = 31 (src/fields.rs:4) */
  OMR_KOALABEAR = 1,  /* p = 2^31 - 2^24 + 1, g = 3  (src/fields.rs:7) */
  OMR_GOLDILOCKS = 2, /* p = 2^64 - 2^32 + 1, g = 7  (src/fields.rs:10-14) */
  OMR_TEDDYBEAR = 3   /* p = 2^32 - 2^30 + 1, g = 5  (src/fields.rs:17) */
} omr_field;

/* Encoder variants: the reference's registry (src/encoders/mod.rs:15-54).
 * Full variants compute out[k] = sum_j in[j] w^(jk) for all k < N; each maps
 * to a schedule of the same fused-LDS Stockham engine.  Partial variants
 * compute out[k] for k < out_len only. */
typedef enum {
  OMR_ARK_RADIX2 = 0,              /* src/encoders/ark_radix2.rs:10-24 (tuned schedule) */
  OMR_DFT = 1,                     /* src/encoders/dft.rs:7-27 (O(N^2) kernel) */
  OMR_LAMBDA_BOWERS = 2,           /* src/encoders/lambdaworks_bowers.rs:11-173 */
  OMR_LAMBDA_RADIX4 = 3,           /* src/encoders/lambdaworks_radix4.rs:12-75 (log N even) */
  OMR_PLONKY3_LAYER_SPLIT = 4,     /* src/encoders/plonky3_radix2_layer_split.rs:16-97 */
  OMR_WINTERFELL_FOUR_STEP = 5,    /* src/encoders/winterfell_four_step.rs:17-130 */
  OMR_TFHE_STOCKHAM_RADIX8 = 6,    /* src/encoders/tfhe_stockham_radix8.rs:13-140 (log N % 3 == 0) */
  OMR_DFT_PARTIAL = 7,             /* src/encoders/dft.rs:29-94 */
  OMR_TRANSFORM_DECOMPOSITION = 8, /* src/encoders/transform_decomposition.rs:10-102 */
  OMR_TRANSFORM_DECOMPOSITION_V2 = 9, /* src/encoders/transform_decomposition_v2.rs:8-105 */
  OMR_VARIANT_COUNT = 10
} omr_variant;

typedef struct omr_ctx omr_ctx;

/* ---- library / context -------------------------------------------------- */
const char *omr_status_str(omr_status s);
const char *omr_version(void);
/* Creates a context bound to HIP device `device`. */
omr_status omr_ctx_create(int device, omr_ctx **out);
omr_status omr_ctx_destroy(omr_ctx *ctx);
/* Waits for all work the context issued on `stream`. */
omr_status omr_ctx_synchronize(omr_ctx *ctx, void *stream);

/* ---- fields and domains (src/fields.rs, src/encoder.rs:23-52) ----------- */
omr_status omr_field_info(int field, uint64_t *modulus, uint64_t *generator, int *two_adicity);
/* Element width in bytes (4 or 8). */
int omr_field_bytes(int field);
/* NttDomain::new checks plus F::get_root_of_unity(N) (src/encoder.rs:23-35). */
omr_status omr_root_of_unity(int field, uint64_t n, uint64_t *omega);
/* NttDomain's public tables (src/encoder.rs:14-18, built at :37-44):
 * twiddles[k] = omega^k for k < n and bitrev_twiddles[i] =
 * omega^bitrev(i, log2 n - 1) for i < n/2, canonical words.  The engine's
 * transforms never need them (it builds its own compact tables from omega);
 * they are for hosts that read the fields.  Either output may be NULL.
 * Device form: d_* are device buffers of n and n/2 words, filled on
 * `stream`; host form: host buffers, synchronous. */
omr_status omr_domain_tables_device(omr_ctx *ctx, int field, uint64_t n, uint64_t omega, void *d_twiddles,
                                    void *d_bitrev_twiddles, void *stream);
omr_status omr_domain_tables(omr_ctx *ctx, int field, uint64_t n, uint64_t omega, void *twiddles,
                             void *bitrev_twiddles);
/* NttEncoder::name (src/encoder.rs:73); NULL for an unknown variant. */
const char *omr_variant_name(int variant);
/* Whether `variant` accepts N = 2^log_n (registry rule, src/encoders/mod.rs:28-46). */
int omr_variant_supports(int variant, uint32_t log_n);

/* ---- the drop-in: NttEncoder::ntt on a host buffer ---------------------- */
/* In place on `buf` (n canonical elements).  Full variants overwrite all n;
 * partial variants (s > 0) write buf[0 .. min(2s, n)) and leave the tail
 * holding the input, as the reference does (transform_decomposition.rs:96,
 * dft.rs:88).  Synchronous. */
omr_status omr_ntt(omr_ctx *ctx, int field, int variant, void *buf, uint64_t n, uint64_t omega,
                   uint64_t s);
/* The same on a buffer of `word_bytes`-wide words: 8 for BabyBear /
 * KoalaBear / TeddyBear held one value per 64-bit word (the one-u64-limb
 * storage of the crate's define_field! types, src/fields.rs:1-17, so a Rust
 * `&mut [F]` is handed over as is; each word < p with a zero upper half;
 * ark-ff's Montgomery words pass through because every variant is linear).
 * 0 or the field's width is omr_ntt.  OMR_E_ARG for other widths. */
omr_status omr_ntt_words(omr_ctx *ctx, int field, int variant, void *buf, uint64_t n, uint64_t omega,
                         uint64_t s, uint32_t word_bytes);

/* ---- device-resident entry points --------------------------------------- */
/* Forward transform of `batch` vectors of n elements (consecutive in memory,
 * stride n) under `variant`.  d_in may equal d_out.  For partial variants
 * out_len = min(2s, n) elements are written per vector.  batch <= 65535
 * here and in the other dense batched entry points (OMR_E_ARG beyond). */
omr_status omr_ntt_device(omr_ctx *ctx, int field, int variant, const void *d_in, void *d_out,
                          uint64_t n, uint64_t batch, uint64_t omega, uint64_t s, void *stream);
/* omr_ntt_device on vectors of `word_bytes`-wide words (see omr_ntt_words;
 * vector b at element b*n of 64-bit words).  Full variants other than Dft
 * read 64-bit words in their first pass and write them in their last pass
 * (no extra pass over memory); Dft and the partial variants run on a
 * narrowed copy and widen the out_len results. */
omr_status omr_ntt_device_words(omr_ctx *ctx, int field, int variant, const void *d_in, void *d_out,
                                uint64_t n, uint64_t batch, uint64_t omega, uint64_t s, uint32_t word_bytes,
                                void *stream);

/* No-allocation form of omr_ntt_device (full-transform variants except
 * OMR_DFT) and omr_intt_device (inverse != 0).  No counterpart in the
 * reference: it is for callers that cannot allocate device memory (code
 * under graph capture, tight serving loops, harnesses that forbid
 * hipMalloc).  It calls no hipMalloc / hipFree: the twiddle tables and the
 * intermediate vectors live in a fixed device array of the library (one per
 * GPU, held until the process ends), the tables rebuilt on the device when
 * (field, n, omega) changes.  n * batch * element bytes must be at most
 * OMR_NOALLOC_MAX_BYTES (32 MiB: Goldilocks 2^22, 32-bit fields 2^23), else
 * OMR_E_SCRATCH before any GPU work.  Calls that overlap in time on one GPU
 * (other streams or host threads) race on that array.  Same results, bit for
 * bit, as omr_ntt_device / omr_intt_device.  d_in may equal d_out. */
#define OMR_NOALLOC_MAX_BYTES (32ull << 20)
omr_status omr_ntt_device_noalloc(omr_ctx *ctx, int field, int variant, const void *d_in, void *d_out,
                                  uint64_t n, uint64_t batch, uint64_t omega, int inverse, void *stream);

/* Inverse: x[j] = n^-1 sum_k X[k] omega^(-jk).  d_in may equal d_out. */
omr_status omr_intt_device(omr_ctx *ctx, int field, const void *d_in, void *d_out, uint64_t n,
                           uint64_t batch, uint64_t omega, void *stream);

/* Transform under an explicit pass schedule: npass passes of radix
 * 2^log_radix[i] (sum = log2 n), register radix 2^log_e (2..4; 0 = auto).
 * inverse != 0 selects the inverse transform.  Used by the autotuner. */
omr_status omr_ntt_schedule_device(omr_ctx *ctx, int field, const void *d_in, void *d_out,
                                   uint64_t n, uint64_t batch, uint64_t omega, int inverse,
                                   const int *log_radix, int npass, int log_e, void *stream);

/* Schedule the engine would use by default for (field, n): writes up to 8
 * log radices, returns the pass count in *npass. */
omr_status omr_default_schedule(int field, uint64_t n, int *log_radix, int *npass);

/* Partial output.  mode 0 (prefix): out[k] = X[k] for k < out_len, the
 * reference's `out[0..min(2s, N)]` with out_len = min(2s, N).  mode 1
 * (stride subset): out[j] = X[j * n / out_len] for j < out_len, out_len a
 * power of two dividing n.  method: 0 auto, 1 direct partial DFT,
 * 2 decomposition with sub-NTT size 2^log_n1 (0 = auto; with out_len <= n1
 * <= n/4 the sub-NTTs run in passes of up to 2^10 with the recombine fused
 * into the last, else n1 is at most the largest pass radix, 2^12 for
 * Goldilocks and 2^13 for the 32-bit fields), 3 full transform then prefix.
 * d_out holds out_len elements. */
omr_status omr_partial_device(omr_ctx *ctx, int field, const void *d_in, void *d_out, uint64_t n,
                              uint64_t omega, uint64_t out_len, int mode, int method,
                              uint32_t log_n1, void *stream);

/* Prefix outputs of a batch of vectors: out[b][k] = X_b[k] for k < out_len,
 * vector b at d_in + b*n, its prefix at d_out + b*out_len.  No counterpart
 * in the reference: the OMR caller's epoch, where every receiver needs
 * k = 2s outputs of its own N-element vector (benches/common/mod.rs:10-12).
 * out_len <= 1024: one fused pass per batch (each input read once);
 * longer prefixes: batched full transforms whose last pass stores only
 * k < out_len. */
omr_status omr_partial_batch_device(omr_ctx *ctx, int field, const void *d_in, void *d_out, uint64_t n,
                                    uint64_t batch, uint64_t omega, uint64_t out_len, void *stream);

/* Coset low-degree extension of `ncols` columns (the encoder path; the
 * reference's decode.rs is a stub, so this is the engine's own closed form):
 *   evals[c][i] = sum_{j<n} coeffs[c][j] (shift * w_m^i)^j,  i < m = n * blowup,
 * w_m the ark-convention root of unity of order m; shift = 0 selects the
 * field generator g (the coset gH).  Column c of coeffs is at d_coeffs + c*n,
 * of evals at d_evals + c*m.  n and blowup powers of two. */
omr_status omr_coset_lde_device(omr_ctx *ctx, int field, const void *d_coeffs, void *d_evals,
                                uint64_t n, uint64_t ncols, uint32_t blowup, uint64_t shift,
                                void *stream);

/* Decode of omr_coset_lde_device: coeffs[c][j] = shift^-j m^-1 sum_i
 * evals[c][i] w_m^(-ij) for j < n.  If d_nonzero (device uint32) is given it
 * is set to the number of nonzero coefficients at j >= n over all columns
 * (0 for codewords). */
omr_status omr_coset_decode_device(omr_ctx *ctx, int field, const void *d_evals, void *d_coeffs,
                                   uint64_t n, uint64_t ncols, uint32_t blowup, uint64_t shift,
                                   uint32_t *d_nonzero, void *stream);

/* ---- batches of vectors of uneven power-of-two sizes --------------------- */
/* One vector of a batch: element offsets into the batch's input and output
 * buffers, size 2^log_n (4 <= log_n <= two-adicity), direction (0 forward,
 * 1 inverse).  Each vector uses the ark-convention root of order 2^log_n.
 * Vectors must not overlap (a vector may be transformed in place). */
typedef struct {
  uint64_t in_offset;
  uint64_t out_offset;
  uint32_t log_n;
  uint32_t inverse;
} omr_batch_item;

typedef struct omr_batch_plan omr_batch_plan;

/* Plans a batch: pass schedules per vector, tile task lists per pass,
 * twiddle tables and intermediate buffers.  Every execution of a plan uses
 * the plan's intermediate buffers: executions queued on one stream run in
 * order, but executions on different streams must not overlap in time (order
 * them with an event), or they overwrite each other's intermediates. */
omr_status omr_batch_plan_create(omr_ctx *ctx, int field, const omr_batch_item *items,
                                 uint32_t count, omr_batch_plan **out);
omr_status omr_batch_plan_destroy(omr_batch_plan *plan);
/* Total elements, kernel launches per execution and tile tasks of a plan. */
omr_status omr_batch_plan_info(const omr_batch_plan *plan, uint64_t *total_elems,
                               uint32_t *nlaunch, uint32_t *ntasks);
/* Runs the batch: one persistent-kernel launch per pass index. */
omr_status omr_batch_execute(omr_batch_plan *plan, const void *d_in, void *d_out, void *stream);
/* Same work, captured once into a hipGraph (per d_in/d_out pair) and replayed. */
omr_status omr_batch_execute_graph(omr_batch_plan *plan, const void *d_in, void *d_out,
                                   void *stream);
/* Baseline form: the single-vector transform launched once per vector. */
omr_status omr_batch_execute_plain(omr_batch_plan *plan, const void *d_in, void *d_out,
                                   void *stream);
/* Grouped form: the vectors of each (size, direction) as one batched
 * transform of the pass engine; the first pass reads and the last pass
 * writes through per-vector offset tables.  A group of 32 MB or more runs its
 * size's default (coalescing-first) schedule, a smaller one the tuned table's
 * (whose entries are single-vector timings); the schedules are fixed when the
 * plan is made.  The groups run on 2 library streams at once (OMR_BATCH_LANES
 * = 1..4 when the plan is made changes that), forked
 * from `stream` and joined back to it (work queued on `stream` afterwards
 * sees every output).
 * The _graph variant captures the launches for a (d_in, d_out) pair and
 * replays them; its capturing call runs them eagerly as well. */
omr_status omr_batch_execute_grouped(omr_batch_plan *plan, const void *d_in, void *d_out, void *stream);
omr_status omr_batch_execute_grouped_graph(omr_batch_plan *plan, const void *d_in, void *d_out,
                                           void *stream);
/* Largest pass radix (2^5 .. 2^12; OMR_E_SCHEDULE otherwise) the batch
 * planner gives multi-pass vectors of later plans.  0 (the default) is
 * automatic: 2^10 for Goldilocks, 2^9 for the 4-byte fields. */
omr_status omr_batch_set_max_log_radix(int log_radix);
/* Pass radices the batch planner gives one vector of 2^log_n elements under
 * a radix cap of 2^max_log_radix (0 = the current setting; when that is
 * automatic, the Goldilocks cap 2^10): one pass up to
 * 2^12, else ceil(log_n / cap) balanced passes, each 2^4 .. 2^12.  Writes up
 * to 9 log radices and the pass count.  Host only (no GPU work). */
omr_status omr_batch_vector_schedule(uint32_t log_n, int max_log_radix, int *log_radix, int *npass);

/* Elementwise device field arithmetic on canonical arrays (test hook):
 * op 0 add, 1 sub, 2 mul, 3 multiplier-form round trip of a, 4 a * b via
 * multiplier form. */
omr_status omr_field_op_device(int field, int op, const void *d_a, const void *d_b, void *d_out,
                               uint64_t n, void *stream);

/* ---- host-buffer forms --------------------------------------------------------
 * Synchronous like omr_ntt (copy in, run, copy out) for callers holding host
 * slices: the boundary companions of NttEncoder::ntt
 * (src/encoder.rs:70-74) -- inverse, batched forward, and the encoder /
 * decoder the reference's src/decode.rs stubs.  omega is the forward root
 * the domain holds (domain.omega); same status codes as the device forms. */
/* buf[0..n) <- inverse NTT (n^-1 sum_k buf[k] omega^-jk), in place. */
omr_status omr_intt(omr_ctx *ctx, int field, void *buf, uint64_t n, uint64_t omega);
/* batch consecutive vectors of n, each forward-transformed in place. */
omr_status omr_ntt_batched(omr_ctx *ctx, int field, void *buf, uint64_t n, uint64_t batch, uint64_t omega);
/* Host forms of omr_coset_lde_device / omr_coset_decode_device (column c of
 * coeffs at c*n, of evals at c*n*blowup); *nonzero (may be NULL) receives the
 * decoder's count of nonzero coefficients at index >= n. */
omr_status omr_coset_lde(omr_ctx *ctx, int field, const void *coeffs, void *evals, uint64_t n, uint64_t ncols,
                         uint32_t blowup, uint64_t shift);
omr_status omr_coset_decode(omr_ctx *ctx, int field, const void *evals, void *coeffs, uint64_t n, uint64_t ncols,
                            uint32_t blowup, uint64_t shift, uint32_t *nonzero);

/* Pinned host memory.  No counterpart in the reference: `NttEncoder::ntt`
 * takes any `&mut [F]` (src/encoder.rs:70-74).  The host-buffer entry points
 * copy a pageable buffer of 8 MiB or more through a pinned staging ring
 * (host threads fill one 16 MiB slot while the DMA engine moves the one
 * before it; OMR_HOST_THREADS sets the thread count, default min(8, cores);
 * OMR_HOST_STAGING=0 turns the ring off).
 * A buffer inside a block from omr_host_alloc is copied directly.  Blocks
 * stay valid until omr_host_free; omr_host_free of any other pointer returns
 * OMR_E_ARG. */
omr_status omr_host_alloc(size_t bytes, void **out);
omr_status omr_host_free(void *p);

/* ---- autotuner ------------------------------------------------------------
 * No counterpart in the reference (its schedules are fixed per encoder).  A
 * shape is (field, n, out_len, batch); candidate schedules come from a fixed
 * vocabulary of classes: 0..29 full transforms (pass count, pass-radix split,
 * register radix), 30..35 prefix methods (direct, decomposition with n1 =
 * 2^6 / 2^8 / 2^10 / 4 out_len in several fused passes, full transform with
 * a pruned last pass).  Tuned choices form a process-wide table consulted by every
 * transform; the first context loads tuning/mi355x.tune beside the library
 * ($OMR_TUNE_FILE overrides, "none" disables). */
int omr_num_schedule_classes(void);
/* Concrete schedule of class cls for a shape: full classes write up to 8 log
 * radices, *npass and *log_e; prefix classes write *method and *log_n1.
 * OMR_E_SCHEDULE if the class is invalid for the shape.  Any output may be
 * NULL. */
omr_status omr_schedule_class(int field, uint64_t n, uint64_t out_len, int cls, int *log_radix, int *npass,
                              int *log_e, int *method, uint32_t *log_n1);
/* The class the engine uses for a shape without tuning or a model (-1 if the
 * shape is invalid). */
int omr_default_class(int field, uint64_t n, uint64_t out_len);
/* Median microseconds of every valid class over `reps` timed runs on
 * context-owned pseudo-random inputs: times_us[omr_num_schedule_classes()],
 * -1 for classes invalid for the shape.  batch > 1 only for full transforms. */
omr_status omr_tune_classes(omr_ctx *ctx, int field, uint64_t n, uint64_t out_len, uint64_t batch, int reps,
                            float *times_us);
/* Installs class cls as the tuned choice for (field, n, out_len). */
omr_status omr_tune_set_class(int field, uint64_t n, uint64_t out_len, int cls);
/* Installs class cls for `batch` vectors of the shape.  Full transforms of
 * two or more vectors totalling 32 MB or more have a slot of their own in the
 * table: the single-vector timings favour small cache-resident passes,
 * which lose up to 40 % on such batches.  Calls of that size read only that
 * slot and take the default schedule when it is empty.  Other shapes behave
 * as with omr_tune_set_class.  omr_autotune_shape installs through this. */
omr_status omr_tune_set_class_batch(int field, uint64_t n, uint64_t out_len, uint64_t batch, int cls);
omr_status omr_tune_clear(void);
/* Text table of the tuned choices (one shape per line). */
omr_status omr_tune_save(const char *path);
omr_status omr_tune_load(const char *path);

/* ---- learned cost model -----------------------------------------------------
 * No counterpart in the reference.  A 4 -> 64 -> 64 -> omr_num_schedule_classes()
 * fp32 MLP (ReLU) whose outputs are predicted centred log-times of the
 * classes for a shape's features (the field as its element width,
 * (bytes - 4) / 4; log n / 32; log out_len / 32; the batch as its total
 * elements, log(n batch) / 40); trained with Adam on autotuner timings by the engine's
 * own HIP forward, backward and update kernels.  Parameters are
 * W1[64][4] b1[64] W2[64][64] b2[64] W3[K][64] b3[K], K = outputs. */
typedef struct omr_cost_model omr_cost_model;
omr_status omr_cost_model_create(omr_ctx *ctx, uint32_t seed, omr_cost_model **out);
omr_status omr_cost_model_destroy(omr_cost_model *cm);
int omr_cost_model_num_params(void);
int omr_cost_model_num_inputs(void);
int omr_cost_model_num_outputs(void);
omr_status omr_cost_model_get_params(omr_cost_model *cm, float *params);
omr_status omr_cost_model_set_params(omr_cost_model *cm, const float *params);
/* Feature vector x[omr_cost_model_num_inputs()] of a shape. */
omr_status omr_cost_model_features(int field, uint64_t n, uint64_t out_len, uint64_t batch, float *x);
/* y[B][K] = MLP(x[B][4]) on the GPU. */
omr_status omr_cost_model_forward(omr_cost_model *cm, const float *x, uint32_t B, float *y);
/* Loss = mean over (b, k) with mask[b][k] != 0 of (y - t)^2, and its gradient
 * with respect to every parameter (grad[num_params]), on the GPU. */
omr_status omr_cost_model_gradients(omr_cost_model *cm, const float *x, const float *t, const float *mask,
                                    uint32_t B, float *grad, float *loss);
/* `steps` full-batch Adam steps at learning rate lr; *loss_out = the loss
 * before the last step. */
omr_status omr_cost_model_train(omr_cost_model *cm, const float *x, const float *t, const float *mask,
                                uint32_t B, uint32_t steps, float lr, float *loss_out);
/* The valid class with the lowest predicted time for a shape. */
omr_status omr_cost_model_choose(omr_cost_model *cm, int field, uint64_t n, uint64_t out_len, uint64_t batch,
                                 int *cls);
omr_status omr_cost_model_save(omr_cost_model *cm, const char *path);
omr_status omr_cost_model_load(omr_cost_model *cm, const char *path);
/* Picks and installs a schedule for a shape: use_model = 1 the model's choice
 * (no timing runs), use_model = k >= 2 the fastest of the model's k
 * best-predicted classes (k timing runs), use_model = 0 (or an untrained or
 * NULL model) exhaustive timing of every class.  *cls = the class installed,
 * *timed = classes timed (may be NULL). */
omr_status omr_autotune_shape(omr_ctx *ctx, omr_cost_model *cm, int use_model, int field, uint64_t n,
                              uint64_t out_len, uint64_t batch, int reps, int *cls, int *timed);

#ifdef __cplusplus
}
#endif
#endif /* OMRNTT_H */
