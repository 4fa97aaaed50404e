/*
 * cufe_ipfe.h -- C ABI of libcufe_ipfe.so, an MI355X (gfx950) implementation of
 * the RLWE inner-product functional-encryption path ("rlwe_sife") and its
 * linear-SVM caller.  Plain C types only; every buffer is caller-owned host
 * memory unless a function says "device".
 *
 * Parameters are the reference's only set (src/params.h:14-28): N = 4096,
 * three RNS primes, L = 785, p = 50241.  Residue arrays are uint32 in
 * [0, q_j], limb-major: poly[slot][limb][coefficient].
 *
 * Randomness: the reference seeds its AES-256-CTR streams from getrandom and
 * leaves two streams uninitialised (src/rlwe_sife.c:40,120).  This library fixes
 * the convention that makes it bit-exact with the reference CPU path:
 *   setup:   s_i from AES-256-CTR(seed2, nonce 0), e_i from AES-256-CTR(seed3, nonce 0);
 *   encrypt: message b of a batch uses nonce b: r and f_0 from (seed2, b),
 *            f_0 .. f_{L-1} from (seed3, b).  Hence repeat <= 65536 per seed pair.
 */
#ifndef CUFE_IPFE_H
#define CUFE_IPFE_H

#include <stdint.h>
#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

#define CUFE_SIFE_N 4096
#define CUFE_SIFE_NMODULI 3
#define CUFE_SIFE_L 785
#define CUFE_SIFE_P 50241

/* ===================================================================== */
/* 1. Reference drop-in entry points                                      */
/*    Same symbols, parameter layouts and call semantics as the reference */
/*    GPU module in its PERF build (src/params.h:9 defines PERF, so the   */
/*    trailing float* accumulates elapsed device milliseconds; NULL is    */
/*    accepted).  Each call allocates, computes and frees its own device  */
/*    memory, as the reference does.  They return void like the           */
/*    reference; failures print to stderr and set cufe_last_error().      */
/* ===================================================================== */

/* Replaces rlwe_sife_setup_gpu (src/rlwe_sife_gpu.cu:19).  On entry mpk[785]
 * holds the uniform polynomial a in the coefficient domain (what
 * sample_polya(seed1) writes, src/rlwe_sife.c:97); on return mpk[785] = NTT(a),
 * mpk[0..784] = a_hat . NTT(s_i) + NTT(e_i) and msk[i] = s_i (coefficients). */
void rlwe_sife_setup_gpu(uint32_t mpk[CUFE_SIFE_L + 1][CUFE_SIFE_NMODULI][CUFE_SIFE_N],
                         uint32_t msk[CUFE_SIFE_L][CUFE_SIFE_NMODULI][CUFE_SIFE_N], unsigned char *seed2,
                         unsigned char *seed3, float *part2_time);

/* Replaces rlwe_sife_encrypt_gpu (src/rlwe_sife_gpu.cu:102).  m is
 * [repeat][785], c is [repeat][786][3][4096] in the coefficient domain. */
void rlwe_sife_encrypt_gpu(uint32_t *m, uint32_t mpk[CUFE_SIFE_L + 1][CUFE_SIFE_NMODULI][CUFE_SIFE_N], uint32_t *c,
                           unsigned char *seed2, unsigned char *seed3, int repeat, float *part2_time);

/* Replaces rlwe_sife_keygen_gpu (src/rlwe_sife_gpu.cu:191).  y is
 * [repeat][785], sk_y is [repeat][3][4096]. */
void rlwe_sife_keygen_gpu(const uint32_t *y, const uint32_t msk[CUFE_SIFE_L][CUFE_SIFE_NMODULI][CUFE_SIFE_N],
                          uint32_t *sk_y, int repeat, float *part2_time);

/* Replaces rlwe_sife_decrypt_gmp_gpu (src/rlwe_sife_gpu.cu:231): one
 * ciphertext against repeat keys; d_y [repeat][3][4096] is the RNS result that
 * round_extract_gmp2 turns into the inner product. */
void rlwe_sife_decrypt_gmp_gpu(uint32_t c[CUFE_SIFE_L + 1][CUFE_SIFE_NMODULI][CUFE_SIFE_N], const uint32_t *y,
                               uint32_t *sk_y, uint32_t *d_y, int repeat, float *part2_time);

/* ===================================================================== */
/* 2. Caller-side helpers the reference runs on the CPU with GMP          */
/* ===================================================================== */

/* sample_polya (src/sample.c:56) on the GPU: a[3][4096]. Returns 0 or -1. */
int cufe_sample_polya(const unsigned char seed[32], uint32_t *a);

/* round_extract_gmp2 (src/rlwe_sife.c:334-393) without GMP: CRT of coefficient
 * 0 of d_y[3][4096] and round(p x / Q), as a double. */
double cufe_round_extract2(const uint32_t *d_y);

/* Plain linear SVM model (the fields of libsvm's struct svm_model that
 * svm_predict_gpu reads, svm/svm.h:52-72); support vectors dense [l][785]. */
typedef struct {
    int nr_class;
    int l;                  /* total support vectors */
    const int *nSV;         /* [nr_class] */
    const int *label;       /* [nr_class] */
    const double *sv_coef;  /* [nr_class-1][l] row-major */
    const double *rho;      /* [nr_class*(nr_class-1)/2] */
    const uint32_t *sv;     /* [l][785] dense, feature index = slot */
} cufe_svm_model;

/* svm_predict_gpu = svm_predict5 (main.c:270-390) for one ciphertext
 * c[786][3][4096]: keygen for every SV, decryption, round_extract, vote.
 * Returns the label; kvalue (optional, [l]) receives the decrypted inner
 * products.  The plain per-call form: keys are recomputed on every call.
 * On failure it returns NaN (and fills kvalue with NaN) with the reason in
 * cufe_last_error(). */
double cufe_svm_predict_gpu(const cufe_svm_model *model,
                            uint32_t c[CUFE_SIFE_L + 1][CUFE_SIFE_NMODULI][CUFE_SIFE_N],
                            const uint32_t msk[CUFE_SIFE_L][CUFE_SIFE_NMODULI][CUFE_SIFE_N], double *kvalue);

/* ===================================================================== */
/* 3. Status-returning API: 0 on success, -1 on failure with the reason   */
/*    in cufe_last_error() (thread-local).                                */
/* ===================================================================== */

const char *cufe_last_error(void);
/* Device name, CU count and HBM size of the current GPU. */
int cufe_device_info(char *name, int name_len, int *cu_count, uint64_t *hbm_bytes);

/* Single operations on host buffers (one call = one batched launch). */
int cufe_ntt_forward(uint32_t *polys, int npoly);  /* [npoly][3][4096] in place: CT_forward */
int cufe_ntt_inverse(uint32_t *polys, int npoly);  /* GS_reverse incl. N^-1 */
/* nstreams independent AES-256-CTR Gaussian streams, stream t keyed by
 * keys[32 t] with nonces[t]; ncalls consecutive 4096-sample calls each.
 * which = 1 (sigma_1, k=266), 2 (k=304214978), 3 (k=608429953).
 * out: [nstreams][ncalls][3][4096]. */
int cufe_gauss_sample(const unsigned char *keys, const uint16_t *nonces, int nstreams, int which, int ncalls,
                      uint32_t *out);
/* The same streams sampled the way the resident row encryption samples its
 * f_i: only sample 0 of every call is produced (the stream is still walked
 * in full).  out: [nstreams][ncalls][3], equal to cufe_gauss_sample's
 * coefficient 0. */
int cufe_gauss_sample_first(const unsigned char *keys, const uint16_t *nonces, int nstreams, int which, int ncalls,
                            uint32_t *out);
/* No counterpart in the reference.  For callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc): cufe_gauss_sample with the output in DEVICE memory the
 * caller owns, d_out [nstreams][ncalls][3][4096], and every scratch buffer of
 * the sampler in one static 64 MiB device array of the library (held per GPU
 * until the process ends).  Nothing is allocated: a single stream whose
 * speculative window does not fit the array (about 600 calls of sampler 3,
 * more for 1 and 2) is rejected with an error; more streams run in groups.
 * Calls on one GPU share the array and are serialised within the process;
 * calls that overlap on one GPU from elsewhere would race.  Runs on the
 * device the library was first used on, on the default stream, and returns
 * after the samples are written.  0 or -1 (reason in cufe_last_error()). */
int cufe_gauss_sample_noalloc(const unsigned char *keys, const uint16_t *nonces, int nstreams, int which, int ncalls,
                              uint32_t *d_out);
int cufe_aes256ctr(const unsigned char key[32], uint16_t nonce, uint32_t ctr0, uint32_t nblocks, unsigned char *out);
/* CRT + rounding of coefficient coef of nvec RNS polys d[nvec][3][4096]. */
int cufe_crt_round(const uint32_t *d, int nvec, int coef, uint32_t *r);


/* The scheme at a vector length l <= 785 (uneven-length batching).  With
 * zero-padding, a length-l encryption equals the first l polynomials plus c_L
 * of the length-785 one (the noise order does not depend on x), so these are
 * the reference functions restricted to l slots.  Layouts: mpk [l+1][3][N]
 * (index l = a_hat), msk [l][3][N], m [repeat][l], c [repeat][l+1][3][N]
 * (index l = c_L), y [repeat][l], sk / d [repeat][3][N].  Encryption of
 * message b uses nonce nonce0 + b. */
int cufe_sife_setup(uint32_t *mpk, uint32_t *msk, const unsigned char seed2[32], const unsigned char seed3[32], int l);
int cufe_sife_encrypt(const uint32_t *m, const uint32_t *mpk, uint32_t *c, const unsigned char seed2[32],
                      const unsigned char seed3[32], int repeat, int l, int nonce0);
int cufe_sife_keygen(const uint32_t *y, const uint32_t *msk, uint32_t *sk, int repeat, int l);
int cufe_sife_decrypt(const uint32_t *c, const uint32_t *y, const uint32_t *sk, uint32_t *d, int repeat, int l);

/* Device-resident timing of the single operations: runs `iters` launches of
 * the kernel over npoly synthetic polys / nstreams streams already in HBM and
 * returns the mean milliseconds per launch measured with HIP events. */
int cufe_time_ntt(int npoly, int inverse, int iters, float *ms_per_launch);
int cufe_time_gauss(int nstreams, int which, int iters, float *ms_per_launch);
/* nstreams independent streams of ncalls consecutive calls each. */
int cufe_time_gauss_calls(int nstreams, int ncalls, int which, int iters, float *ms_per_run);

/* ===================================================================== */
/* 3b. Parameter sets                                                      */
/*    pset 0: N = 2048, 2 limbs (2^31-2^17+1, 2^31-2^24+1), sampler widths  */
/*            k = 266, 262147, 524287;                                      */
/*    pset 1: the reference's set (src/params.h:14-28), the set of every    */
/*            other function in this header;                                */
/*    pset 2: N = 8192, 4 limbs (the reference's primes and 2^26-2^16+1),   */
/*            the reference's sampler widths.                               */
/*    L = 785, p = 50241 in each.  The reference defines only set 1; the    */
/*    others follow its comment on other levels (src/sample.c:16-19).       */
/*    Layouts as above with nmod limbs and N coefficients:                  */
/*    poly[slot][limb][coefficient].  Same seed convention as section 1.    */
/* ===================================================================== */

/* N, limbs, key length, moduli q[4] (unused limbs 0) and p of a set. */
int cufe_pset_info(int pset, int *n, int *nmod, int *l, uint32_t *q, uint32_t *p);
/* CT_forward / GS_reverse (src/ntt.c:38-114) of every limb of polys
 * [npoly][nmod][N], in place, with the shape the autotuner picked for the set
 * and direction on this GPU (tuned on first use). */
int cufe_ps_ntt(int pset, int inverse, uint32_t *polys, int npoly);
/* The same with one given shape (all shapes give identical results). */
int cufe_ps_ntt_variant(int pset, int variant, int inverse, uint32_t *polys, int npoly);
/* Number of NTT shapes for the set's N; names (optional) receives one name per
 * line ("-" for a shape the set cannot use). */
int cufe_ntt_variants(int pset, char *names, int len);
/* The NTT autotuner's timings and choices so far.  Returns the length. */
int cufe_ntt_autotune_report(char *buf, int len);
/* sample_polya (src/sample.c:56-128) at the set's N: a [nmod][N]. */
int cufe_ps_sample_polya(int pset, const unsigned char seed[32], uint32_t *a);
/* gaussian_sampler_S{which} with the set's widths and N samples per call:
 * out [nstreams][ncalls][nmod][N]. */
int cufe_ps_gauss_sample(int pset, const unsigned char *keys, const uint16_t *nonces, int nstreams, int which,
                         int ncalls, uint32_t *out);
/* rlwe_sife_setup / encrypt / keygen / decrypt_gmp at the set's parameters,
 * with the layouts of cufe_sife_* (mpk [l+1][nmod][N] with mpk[l] = a on entry
 * of setup and a_hat on return, msk [l][nmod][N], c [repeat][l+1][nmod][N],
 * sk / d [repeat][nmod][N]). */
int cufe_ps_setup(int pset, uint32_t *mpk, uint32_t *msk, const unsigned char seed2[32],
                  const unsigned char seed3[32], int l);
int cufe_ps_encrypt(int pset, const uint32_t *m, const uint32_t *mpk, uint32_t *c, const unsigned char seed2[32],
                    const unsigned char seed3[32], int repeat, int l, int nonce0);
int cufe_ps_keygen(int pset, const uint32_t *y, const uint32_t *msk, uint32_t *sk, int repeat, int l);
int cufe_ps_decrypt(int pset, const uint32_t *c, const uint32_t *y, const uint32_t *sk, uint32_t *d, int repeat,
                    int l);
/* crt_reverse_gmp + round_extract_gmp (src/crt.c:46-82, src/rlwe_sife.c:288-332)
 * on the GPU: coefficient coef of d [nvec][nmod][N] -> r [nvec], or every
 * coefficient when coef < 0 -> r [nvec][N]. */
int cufe_ps_crt_round(int pset, const uint32_t *d, int nvec, int coef, uint32_t *r);
/* The same for one coefficient's residues res [nmod], on the host. */
int cufe_ps_crt_round_host(int pset, const uint32_t *res, uint32_t *r);
/* Device-resident timing: the tuned transform over npoly polys of the set
 * (ms per launch), and nstreams one-call Gaussian streams (ms per run). */
int cufe_time_ps_ntt(int pset, int npoly, int inverse, int iters, float *ms_per_launch);
int cufe_time_ps_gauss(int pset, int nstreams, int which, int iters, float *ms_per_run);

/* ===================================================================== */
/* 4. Resident-key context and batched prediction                          */
/*    Keys, functional keys and encrypted rows stay in HBM; one predict     */
/*    call classifies many rows against every support vector at once.       */
/*    Outputs equal svm_predict_gpu's (main.c:270-390) row by row.          */
/* ===================================================================== */

typedef struct cufe_ctx cufe_ctx;

int cufe_ctx_create(cufe_ctx **out);
void cufe_ctx_destroy(cufe_ctx *ctx);
/* Master keys of length l (<= 785) generated on the device: a =
 * sample_polya(seedA), s_i from (seed2, nonce 0), e_i from (seed3, nonce 0). */
int cufe_ctx_setup(cufe_ctx *ctx, const unsigned char seedA[32], const unsigned char seed2[32],
                   const unsigned char seed3[32], int l);
/* Master keys from host memory: mpk [l+1][3][4096], msk [l][3][4096]. */
int cufe_ctx_load_keys(cufe_ctx *ctx, const uint32_t *mpk, const uint32_t *msk, int l);
int cufe_ctx_get_keys(cufe_ctx *ctx, uint32_t *mpk, uint32_t *msk);
/* Generates and keeps one functional key per support vector (rlwe_sife_keygen
 * over the key length, src/rlwe_sife.c:215-245). */
int cufe_ctx_set_model(cufe_ctx *ctx, const cufe_svm_model *model);
/* Capacity of the encrypted-row store: cap rows of l slots each; l must exceed
 * every feature index the model uses. */
int cufe_ctx_rows_reserve(cufe_ctx *ctx, int cap, int l);
/* Encrypts n rows x [n][l] and appends them to the store; row b uses nonce
 * nonce0 + b (as rlwe_sife_encrypt_gpu numbers its batch).  ms: device time. */
int cufe_ctx_encrypt_rows(cufe_ctx *ctx, const uint32_t *x, int n, const unsigned char seed2[32],
                          const unsigned char seed3[32], int nonce0, float *ms);
int cufe_ctx_rows_count(const cufe_ctx *ctx);
/* Labels [nrows] of stored rows [row0, row0+nrows); kvalues (optional)
 * [nrows][model l] = the decrypted inner products; ms (optional, 4 floats) =
 * device milliseconds of contraction, key-value epilogue, vote, total. */
int cufe_ctx_predict(cufe_ctx *ctx, int row0, int nrows, double *labels, uint32_t *kvalues, float *ms);
/* Contraction time per limb (ms[3], summed over row chunks) of the last
 * cufe_ctx_predict call (limb 0 may run a cheaper kernel: its residues fit
 * three int8 digits).  Chunks of at most 256 rows (and 2^19 row-key pairs)
 * run the three limbs in one split-K launch, whose whole time counts as
 * limb 0's.  In the fraction form: ms[0] = its contraction (all limbs),
 * ms[1] = its exact kernel, ms[2] = 0. */
int cufe_ctx_contract_limb_ms(const cufe_ctx *ctx, float *ms);
/* Form of the last cufe_ctx_predict call's contraction: 2 = the fraction form
 * ran on at least one row chunk (every limb in one int8 contraction over
 * 72-bit fractions of each column's CRT value, pairs near a rounding change
 * recomputed exactly; keys of two digits that are the same integers in every
 * limb; chunks of at most 256 rows run the split-K per-limb launch instead),
 * 1 = int8 per limb, 0 = fp64.  exact_pairs (optional): the pairs recomputed exactly.  Same key
 * values in every form.  CUFE_CONTRACT=limbs turns the fraction form off. */
int cufe_ctx_contract_form(const cufe_ctx *ctx, unsigned long long *exact_pairs);
/* Number of block shapes of the fraction-form contraction; CUFE_FRAC_VARIANT=v
 * (0 <= v < this) forces shape v instead of the autotuned one. */
int cufe_contract_frac_variants(void);
/* Reduction interval of the fp64 contraction and the largest balanced
 * functional-key coefficient, per limb. */
int cufe_ctx_model_info(const cufe_ctx *ctx, int *kb, unsigned int *maxabs);
/* Functional keys [model l][3][4096] and one stored row (c_L [3][4096] and
 * coefficient 0 of c_0..c_{l-1} as [l][3]), for inspection. */
int cufe_ctx_get_model_keys(const cufe_ctx *ctx, uint32_t *sk);
/* The contraction runs on int8 matrix cores when every functional-key
 * coefficient fits two signed base-256 digits (|b| <= 32512), else on fp64;
 * CUFE_CONTRACT=sgpr|lds forces an fp64 kernel.  The int8 kernel's block shape
 * is tuned on the first launch of each (keys, rows, k steps) shape; this
 * report lists the candidates' times and choices.  Returns its length. */
int cufe_autotune_report(char *buf, int len);
/* Key digits of the loaded model's int8 contraction: 2 (|b| <= 32512), 3
 * (|b| <= 8322943) or 0 (fp64 kernels). */
int cufe_ctx_contract_digits(const cufe_ctx *ctx);
/* Number of int8 contraction variants (block shapes and pipeline depths);
 * valid_mask bit v = variant v runs with nb key digits.  CUFE_I8_VARIANT=v
 * forces variant v instead of the tuned choice. */
int cufe_contract_variants(int nb, unsigned long long *valid_mask);
int cufe_ctx_get_row(const cufe_ctx *ctx, int row, uint32_t *cl, uint32_t *c0);
/* Stored rows [row0, row0+n) in compact wire form: c_L [n][3][4096] and
 * coefficient 0 of c_0..c_{l-1} as [n][l][3] (what a serving request carries). */
int cufe_ctx_export_rows(const cufe_ctx *ctx, int row0, int n, uint32_t *cl, uint32_t *c0);

/* Batched encryption of messages of uneven lengths (rlwe_sife_encrypt,     */
/* src/rlwe_sife.c:112-181, once per message, as rlwe_sife_encrypt_gpu's    */
/* repeat loop, src/rlwe_sife_gpu.cu:102, but with a length per message).  */
/* Message b of length l_b (1 <= l_b <= key length) uses nonce nonce0 + b;  */
/* its ciphertext [l_b+1][3][4096] (index l_b = c_L) equals c_0..c_{l_b-1}  */
/* and c_L of the message zero-padded to the full key length.  Ciphertexts  */
/* stay in HBM, packed in message order.  The context must outlive the     */
/* batch.                                                                  */
typedef struct cufe_encbatch cufe_encbatch;
int cufe_encbatch_create(const cufe_ctx *ctx, const int *lens, int n, cufe_encbatch **out);
/* x: packed messages (sum of lens words).  use_graph: replay the batch as one
 * captured hipGraph (captured on first use).  ms: device time of the run. */
int cufe_encbatch_run(cufe_encbatch *eb, const uint32_t *x, const unsigned char seed2[32],
                      const unsigned char seed3[32], int nonce0, int use_graph, float *ms);
/* Ciphertexts of messages [b0, b0+nb) to host memory, packed. */
int cufe_encbatch_export(const cufe_encbatch *eb, int b0, int nb, uint32_t *c);
/* Device checksum of all ciphertext words: sum of (2i+1) w_i mod 2^64.  info
 * (optional): messages, words, sampler groups, groups redone. */
int cufe_encbatch_checksum(cufe_encbatch *eb, uint64_t *sum, long long *info);
void cufe_encbatch_destroy(cufe_encbatch *eb);

/* ===================================================================== */
/* 5. Serving: continuous batching over resident keys                      */
/*    Replaces the one-row-per-call loop of svm_predict_gpu (main.c:270-390, */
/*    which ships msk and a full ciphertext per row and regenerates every   */
/*    functional key).  Requests carry rows in compact form (c_L and         */
/*    coefficient 0 of c_0..c_{l-1}, as cufe_ctx_export_rows gives them); a */
/*    worker thread packs queued requests of any loaded models into batches */
/*    of up to one slot's capacity.  The HBM budget fixes at start the slot  */
/*    capacity (rows) and the number of slots (2 to 4) that run side by     */
/*    side; a request larger than a slot runs as slot-sized parts.          */
/*    Results equal cufe_ctx_predict's for the same rows, alone, mixed or   */
/*    split.                                                                */
/* ===================================================================== */

typedef struct cufe_server cufe_server;

/* ctx supplies the master keys and must outlive the server.  hbm_budget:
 * device bytes for keys and the batch slots (0 = 16 GiB): a slot holds up to
 * max_batch_rows rows (0 = 4096) if two such slots fit, fewer otherwise, and
 * the budget left over buys up to two more slots; batching 0 runs one request
 * (part) per batch. */
int cufe_server_create(cufe_ctx *ctx, size_t hbm_budget, int max_batch_rows, int batching, cufe_server **out);
/* Generates and keeps the functional keys of one model (before start). */
int cufe_server_add_model(cufe_server *srv, const cufe_svm_model *model, int *model_id);
/* Sizes and allocates the batch slots and starts the worker thread. */
int cufe_server_start(cufe_server *srv);
/* Queues n rows of l slots (model's largest feature < l <= key length); n
 * above the slot capacity runs as parts of at most that many rows:
 * cl [n][3][4096], c0 [n][l][3].  labels [n] and kvalues [n][model l]
 * (optional) receive the results; every buffer stays the caller's and must
 * live until the request completes. */
int cufe_server_submit(cufe_server *srv, int model, int n, int l, const uint32_t *cl, const uint32_t *c0,
                       double *labels, uint32_t *kvalues, uint64_t *ticket);
/* 1 done, 0 pending after timeout_ms (< 0 waits without limit), -1 failed.
 * A completed ticket is released.  latency_ms: submit to completion. */
int cufe_server_wait(cufe_server *srv, uint64_t ticket, double timeout_ms, double *latency_ms);
/* Open-loop replay of nreq requests at the given arrival offsets (seconds);
 * latency_ms[i] is measured from the scheduled arrival.  Returns the number
 * of failed requests, or -1. */
int cufe_server_replay(cufe_server *srv, int nreq, const double *arrival_s, const int *model, const int *nrows,
                       const int *l, const uint32_t *const *cl, const uint32_t *const *c0, double *const *labels,
                       double *latency_ms, double *makespan_s);
/* requests, rows, batches, failed, largest batch rows, slot capacity rows,
 * model key bytes, bytes per slot. */
int cufe_server_stats(cufe_server *srv, long long *out8);
/* Number of batch slots, requests split into parts, peak device bytes held by
 * batches in flight (rows x bytes_per_row) and the device bytes one batch row
 * occupies (row store, staging, key values, vote). */
int cufe_server_info(cufe_server *srv, int *nslots, long long *split_requests, long long *peak_inflight_bytes,
                     long long *bytes_per_row);
/* Zeroes the request, row, batch, failure and largest-batch counters (after
 * an untimed warm-up, so that reported stats cover the timed requests only). */
int cufe_server_reset_stats(cufe_server *srv);
/* Completes queued requests, stops the worker, frees device memory. */
void cufe_server_destroy(cufe_server *srv);
/* Pinned host memory (hipHostMalloc).  Request rows that lie inside such a
 * block are copied to the device by DMA directly; other buffers go through
 * the server's pinned staging first. */
int cufe_host_alloc(size_t bytes, void **out);
void cufe_host_free(void *p);

/* ===================================================================== */
/* 6. Wire format for requests, ciphertexts and keys                       */
/*    No counterpart in the reference, which passes in-memory arrays only.  */
/*    A 64-byte little-endian header (magic "CUFEIPFE", version 1, kind,   */
/*    limbs, N, count, slots l, the three moduli, a caller tag, CRC-32 of   */
/*    the payload, payload bytes) followed by the residue arrays exactly as */
/*    the API takes them; every residue lies in [0, q] of its limb.         */
/* ===================================================================== */

enum {
    CUFE_WIRE_ROWS = 1,            /* c_L [count][3][4096], then c0 [count][l][3] */
    CUFE_WIRE_CIPHERTEXTS = 2,     /* [count][l+1][3][4096], index l = c_L */
    CUFE_WIRE_PUBLIC_KEY = 3,      /* [l+1][3][4096], index l = a_hat, count 1 */
    CUFE_WIRE_FUNCTIONAL_KEYS = 4  /* [count][3][4096] */
};

typedef struct {
    int kind, count, l;
    uint32_t tag;
    const uint32_t *data; /* payload inside the blob */
    const uint32_t *c0;   /* rows: the c0 block inside the blob; else NULL */
} cufe_wire_info;

/* Bytes of a blob of this kind and shape (0 if the shape is invalid). */
size_t cufe_wire_size(int kind, int count, int l);
/* Packs data (and, for rows, data2 = c0) into out. */
int cufe_wire_pack(int kind, const uint32_t *data, const uint32_t *data2, int count, int l, uint32_t tag, void *out,
                   size_t cap);
/* Validates a blob (header, parameter set, length, CRC, residue ranges) and
 * describes it; the pointers point into the blob (4-byte aligned blobs). */
int cufe_wire_parse(const void *blob, size_t len, cufe_wire_info *info);

#ifdef __cplusplus
}
#endif
#endif
