/*
 * ipfe_oracle.h -- CPU restatement of the reference RLWE-IPFE ("rlwe_sife") path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker: it is linked or loaded
 * only by tests/, by __graft_entry__.smoke() and by bench.py's cpu_baseline leg.
 * The product (libcufe_ipfe.so) never calls it and has no CPU fallback.
 *
 * Every function restates one reference function (cited file:line, relative to
 * the reference tree).  Randomness follows a fixed seed convention that replaces
 * the reference's getrandom() seeds and its two uninitialised AES states:
 *   setup(seedA, seedS, seedE): a <- sample_polya(seedA); secret stream
 *     S = AES-256-CTR(seedS, nonce 0); error stream E = AES-256-CTR(seedE, nonce 0).
 *   encrypt(x, seed2, seed3, nonce): r, f0 from AES-256-CTR(seed2, nonce);
 *     f_0 .. f_{l-1} from AES-256-CTR(seed3, nonce).
 *
 * Parity status of each function (how it is pinned) is listed in tests/ and in
 * the design notes; none is "parity unpinned".
 */
#ifndef CUFE_IPFE_ORACLE_H
#define CUFE_IPFE_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define ORC_N 4096
#define ORC_NMODULI 3
#define ORC_L 785

/* AES-256-CTR stream: counter block = BE16(nonce) || 0^10 || BE32(ctr)
 * (src/aes256ctr.c:27-40,92). */
typedef struct {
    uint32_t rk[60];   /* FIPS-197 expanded key, big-endian words */
    uint32_t ctr;      /* next block counter */
    uint16_t nonce;
} orc_stream;

void orc_stream_init(orc_stream *st, const uint8_t key[32], uint16_t nonce);
/* Emits nblocks128 * 128 bytes, i.e. 8 AES blocks per unit (src/aes256ctr.c:145-155). */
void orc_stream_squeeze(orc_stream *st, uint8_t *out, uint64_t nblocks128);
void orc_aes256_encrypt_block(const uint32_t rk[60], const uint8_t in[16], uint8_t out[16]);
void orc_aes256_expand(const uint8_t key[32], uint32_t rk[60]);

/* src/arith_rns.c:22-103 */
uint32_t orc_mod_prime(uint64_t m, uint32_t sel);
uint32_t orc_add_mod(uint32_t a, uint32_t b, uint32_t sel);
uint32_t orc_sub_mod(uint32_t a, uint32_t b, uint32_t sel);
uint32_t orc_mul_mod(uint32_t a, uint32_t b, uint32_t sel);

/* Parameters (src/params.h:14-28, src/consts.h:7-11) */
uint32_t orc_modulus(uint32_t sel);
uint32_t orc_ninv(uint32_t sel);
uint32_t orc_scale_mod(uint32_t sel);           /* floor(Q/p) mod q_sel */
const uint32_t *orc_psi(uint32_t sel);         /* w^bitrev12(k) */
const uint32_t *orc_psi_inv(uint32_t sel);     /* w^-bitrev12(k) */

/* src/sample.c:56-128 (SEC_LEVEL==0 branch) */
void orc_sample_polya(const uint8_t seed[32], uint32_t a[ORC_NMODULI][ORC_N]);

/* src/gauss.c:705-785 (S1_ori), :915-993 (S2_ori), :1086-1163 (S3).  which = 1,2,3 */
void orc_gauss(orc_stream *st, int which, uint32_t out[ORC_NMODULI][ORC_N], uint32_t slen);
/* One 384-byte chunk -> 8 candidates (value z, sign, accept).  Exposed for tests
 * and for the GPU candidate-kernel parity check.  Returns 0, or -1 when the
 * uniform sampler would read past the chunk. */
int orc_gauss_chunk(const uint8_t r[384], int which, uint64_t z[8], uint8_t sign[8], uint8_t acc[8]);
/* The same for any sampler width k (S1/S2/S3 use k = 266, 304214978, 608429953). */
int orc_gauss_chunk_k(const uint8_t r[384], uint64_t k, uint64_t z[8], uint8_t sign[8], uint8_t acc[8]);

/* src/ntt.c:38-67, :70-114, :28-36 */
void orc_ct_forward(uint32_t a[ORC_N], uint32_t sel);
void orc_gs_reverse(uint32_t a[ORC_N], uint32_t sel);
void orc_point_mul(const uint32_t *a, const uint32_t *b, uint32_t *c, uint32_t sel);
void orc_poly_mul_ntt(const uint32_t *a, const uint32_t *b, uint32_t *c, uint32_t sel);

/* Scheme.  Layouts: mpk [l+1][3][N] (index l = a-hat), msk [l][3][N],
 * c [l+1][3][N] (index l = c_L).  l <= 785; the reference always uses l = 785,
 * and a shorter l yields exactly the first l polynomials (prefix property). */
void orc_setup(const uint8_t seedA[32], const uint8_t seedS[32], const uint8_t seedE[32],
               int l, uint32_t *mpk, uint32_t *msk);
void orc_setup_from_a(const uint32_t *a_coeff, const uint8_t seedS[32], const uint8_t seedE[32],
                      int l, uint32_t *mpk, uint32_t *msk);
void orc_encrypt(const uint32_t *x, int l, const uint32_t *mpk, const uint8_t seed2[32],
                 const uint8_t seed3[32], uint16_t nonce, uint32_t *c);
/* Dense keygen over all l slots (src/rlwe_sife.c:215-245). */
void orc_keygen(const uint32_t *y, int l, const uint32_t *msk, uint32_t *sk);
/* RNS decryption d_y[3][N] (src/rlwe_sife.c:248-280, before crt_reverse_gmp). */
void orc_decrypt(const uint32_t *c, const uint32_t *y, int l, const uint32_t *sk, uint32_t *d);
/* src/crt.c:46-82 (Garner, mpz semantics) -> x as (lo, hi) 64-bit halves. */
void orc_crt_reverse(const uint32_t d0, const uint32_t d1, const uint32_t d2, uint64_t *lo, uint64_t *hi);
/* src/rlwe_sife.c:288-332: round(p*x/Q) with no reduction mod p. */
uint32_t orc_round_extract(uint64_t lo, uint64_t hi);
/* round_extract_gmp2 (src/rlwe_sife.c:334-393): coefficient 0 only, as double. */
double orc_round_extract_coef0(const uint32_t *d);

/* Key values round_extract(CRT(d_y[.][0])) of every (row, key) pair from
 * coefficient 0 of the decryption alone (see ipfe_oracle.c).  cl [R][3][N],
 * c0 [R][l][3] (coefficient 0 of c_0..c_{l-1}), y [S][ly], sk [S][3][N],
 * kv [R][S].  Returns the number of pairs whose value depends on the INTT's
 * representative of 0 (their kv is 0xffffffff); 0 when every key ends on a
 * zero slot. */
int orc_kvalues_coef0(const uint32_t *cl, const uint32_t *c0, int R, int l, const uint32_t *y, int ly,
                      const uint32_t *sk, int S, uint32_t *kv);

/* One-vs-one decision (main.c:338-389).  kvalue[l_sv]; sv_coef is
 * [nr_class-1][l_sv] row-major; returns the label.  use_fma selects the
 * contraction gcc applies to "sum += coef * kvalue" at -O3 -march=native. */
double orc_svm_decide(int nr_class, int l_sv, const int *nSV, const int *label,
                      const double *sv_coef, const double *rho, const double *kvalue, int use_fma);

/* Whole svm_predict_cpu (main.c:137-267) for one row: for each SV s
 * keygen(y_s) -> decrypt -> round -> kvalue[s]; then decide. y is [l_sv][l]. */
double orc_svm_predict_cpu(const uint32_t *c, int l, const uint32_t *msk, int nr_class, int l_sv,
                           const int *nSV, const int *label, const double *sv_coef,
                           const double *rho, const uint32_t *y, double *kvalue_out, int use_fma);

/* ------------------------------------------------------------------------- */
/* The path for any of the library's three parameter sets (pset_oracle.c):     */
/* set 0: N = 2048, 2 limbs; set 1: the reference's set; set 2: N = 8192, 4    */
/* limbs.  Layouts are those above with nmod limbs and N coefficients.         */
/* ------------------------------------------------------------------------- */
typedef struct {
    int id, N, logN, nmod, L;
    uint32_t q[4], fa[4], fb[4], root[4]; /* q = 2^fa - 2^fb + 1; root: primitive 2N-th */
    uint64_t k[4];                        /* sampler widths k[1..3] */
    uint32_t p;
} orc_pset;

int orc_pset_get(int id, orc_pset *out);
uint32_t orc_ps_mod_prime(const orc_pset *ps, uint64_t m, int j);
void orc_ps_psi(const orc_pset *ps, int j, uint32_t *psi, uint32_t *psi_inv);
uint32_t orc_ps_ninv(const orc_pset *ps, int j);
uint32_t orc_ps_scale(const orc_pset *ps, int j);
void orc_ps_ct_forward(const orc_pset *ps, uint32_t *a, int j);
void orc_ps_gs_reverse(const orc_pset *ps, uint32_t *a, int j);
void orc_ps_sample_polya(const orc_pset *ps, const uint8_t seed[32], uint32_t *a);
void orc_ps_gauss(const orc_pset *ps, orc_stream *st, int which, uint32_t *out);
void orc_ps_setup_from_a(const orc_pset *ps, const uint32_t *a_coeff, const uint8_t seedS[32],
                         const uint8_t seedE[32], int l, uint32_t *mpk, uint32_t *msk);
void orc_ps_encrypt(const orc_pset *ps, const uint32_t *x, int l, const uint32_t *mpk, const uint8_t seed2[32],
                    const uint8_t seed3[32], uint16_t nonce, uint32_t *c);
void orc_ps_keygen(const orc_pset *ps, const uint32_t *y, int l, const uint32_t *msk, uint32_t *sk);
void orc_ps_decrypt(const orc_pset *ps, const uint32_t *c, const uint32_t *y, int l, const uint32_t *sk,
                    uint32_t *d);
void orc_ps_crt(const orc_pset *ps, const uint32_t *res, uint64_t *lo, uint64_t *hi);
uint32_t orc_ps_round(const orc_pset *ps, uint64_t lo, uint64_t hi);

#ifdef __cplusplus
}
#endif
#endif
