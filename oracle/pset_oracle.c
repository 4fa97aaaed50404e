/*
 * pset_oracle.c -- the CPU restatement of the RLWE-IPFE path for any of the
 * three parameter sets the library supports:
 *
 *   set 0: N = 2048, 2 limbs (q = 2^31-2^17+1, 2^31-2^24+1), sampler widths
 *          k = 266, 262147, 524287;
 *   set 1: the reference's set (src/params.h:14-28): N = 4096, 3 limbs, k = 266,
 *          304214978, 608429953;
 *   set 2: N = 8192, 4 limbs (the reference's three primes and 2^26-2^16+1),
 *          the reference's sampler widths.
 *
 * All sets keep L = 785, B_x = 4, B_y = 16, p = 50241.  Sets 0 and 2 have no
 * reference: they follow the upstream comment on other security levels
 * (src/sample.c:16-19: N = 2048 with 64 bits of Q, N = 8192 with more), and
 * their constants are chosen here -- primes of the form 2^a - 2^b + 1 with
 * q = 1 mod 2N (so the fold of src/arith_rns.c:61-83 applies), sampler widths
 * of set 0 scaled so that 15 standard deviations of the worst-case decryption
 * noise (|y| = B_y sqrt(L)) stay below Q / (2p) (and are not powers of two,
 * whose uniform bound k floor(2^32/k) would not fit 32 bits), and for sets 0 and 2 the
 * primitive 2N-th root w = g^((q-1)/2N) with g the smallest primitive root
 * mod q (set 1 keeps the reference's roots, src/consts.h:7-11).  Q p < 2^128
 * for every set, so CRT and rounding fit unsigned __int128.
 *
 * Every function follows the reference function cited at it, with N, the limb
 * count, the primes and the widths taken from the set; at set 1 each equals
 * the fixed-set oracle in ipfe_oracle.c (checked in tests).
 *
 * TEST INFRASTRUCTURE ONLY, like ipfe_oracle.c.
 */
#include <stdlib.h>
#include <string.h>

#include "ipfe_oracle.h"

static const orc_pset PSETS[3] = {
    {0, 2048, 11, 2, 785, {2147352577u, 2130706433u, 0, 0}, {31, 31, 0, 0}, {17, 24, 0, 0},
     {1543607367u, 1989134074u, 0, 0}, {0, 266ull, 262147ull, 524287ull}, 50241u},
    {1, 4096, 12, 3, 785, {16760833u, 2147352577u, 2130706433u, 0}, {24, 31, 31, 0}, {14, 17, 24, 0},
     {3689u, 1689795u, 456478u, 0}, {0, 266ull, 304214978ull, 608429953ull}, 50241u},
    {2, 8192, 13, 4, 785, {16760833u, 2147352577u, 2130706433u, 67043329u}, {24, 31, 31, 26}, {14, 17, 24, 16},
     {5768458u, 772388167u, 1047035213u, 8157893u}, {0, 266ull, 304214978ull, 608429953ull}, 50241u},
};

int orc_pset_get(int id, orc_pset *out)
{
    if (id < 0 || id > 2) return -1;
    *out = PSETS[id];
    return 0;
}

/* mod_prime (src/arith_rns.c:61-83) for limb j of the set. */
uint32_t orc_ps_mod_prime(const orc_pset *ps, uint64_t m, int j)
{
    const uint64_t q = ps->q[j];
    const uint32_t a = ps->fa[j], b = ps->fb[j];
    const uint64_t lo = ((uint64_t)1 << a) - 1;
    while (m > 2 * q) {
        uint64_t hi = m >> a;
        m = (m & lo) + ((hi << b) - hi);
    }
    if (m >= q) m -= q;
    return (uint32_t)m;
}

static uint32_t ps_add(const orc_pset *ps, uint32_t a, uint32_t b, int j)
{
    uint64_t c = (uint64_t)a + b;
    if (c >= ps->q[j]) c -= ps->q[j];
    return (uint32_t)c;
}

static uint32_t ps_sub(const orc_pset *ps, uint32_t a, uint32_t b, int j)
{
    uint64_t c = (uint64_t)a + ps->q[j] - b;
    if (c >= ps->q[j]) c -= ps->q[j];
    return (uint32_t)c;
}

static uint32_t ps_mul(const orc_pset *ps, uint32_t a, uint32_t b, int j)
{
    return orc_ps_mod_prime(ps, (uint64_t)a * b, j);
}

static uint32_t ps_pow(uint64_t b, uint64_t e, uint64_t q)
{
    uint64_t r = 1;
    b %= q;
    while (e) {
        if (e & 1) r = r * b % q;
        b = b * b % q;
        e >>= 1;
    }
    return (uint32_t)r;
}

static uint32_t bitrev(uint32_t k, int bits)
{
    uint32_t r = 0;
    for (int i = 0; i < bits; ++i) r |= ((k >> i) & 1u) << (bits - 1 - i);
    return r;
}

/* psi[k] = w^bitrev(k), psi_inv[k] = w^-bitrev(k) (the tables of
 * src/consts.h:9-11 at the set's N). */
void orc_ps_psi(const orc_pset *ps, int j, uint32_t *psi, uint32_t *psi_inv)
{
    const uint32_t q = ps->q[j];
    const uint32_t winv = ps_pow(ps->root[j], q - 2, q);
    for (int k = 0; k < ps->N; ++k) {
        if (psi) psi[k] = ps_pow(ps->root[j], bitrev((uint32_t)k, ps->logN), q);
        if (psi_inv) psi_inv[k] = ps_pow(winv, bitrev((uint32_t)k, ps->logN), q);
    }
}

uint32_t orc_ps_ninv(const orc_pset *ps, int j) { return ps_pow((uint64_t)ps->N, ps->q[j] - 2, ps->q[j]); }

/* floor(Q/p) mod q_j (src/params.h:27-28 at the set's Q). */
uint32_t orc_ps_scale(const orc_pset *ps, int j)
{
    unsigned __int128 Q = 1;
    for (int i = 0; i < ps->nmod; ++i) Q *= ps->q[i];
    return (uint32_t)((Q / ps->p) % ps->q[j]);
}

/* CT_forward (src/ntt.c:38-67) at the set's N. */
void orc_ps_ct_forward(const orc_pset *ps, uint32_t *a, int j)
{
    const int N = ps->N;
    uint32_t *psi = malloc(sizeof(uint32_t) * (size_t)N);
    orc_ps_psi(ps, j, psi, NULL);
    uint32_t t = (uint32_t)N;
    for (uint32_t m = 1; m < (uint32_t)N; m <<= 1) {
        t >>= 1;
        for (uint32_t i = 0; i < m; ++i) {
            uint32_t j1 = 2 * i * t, S = psi[m + i];
            for (uint32_t x = j1; x < j1 + t; ++x) {
                uint32_t U = a[x], V = ps_mul(ps, a[x + t], S, j);
                a[x] = ps_add(ps, U, V, j);
                a[x + t] = ps_sub(ps, U, V, j);
            }
        }
    }
    free(psi);
}

/* GS_reverse (src/ntt.c:70-114) at the set's N, including N^-1. */
void orc_ps_gs_reverse(const orc_pset *ps, uint32_t *a, int j)
{
    const int N = ps->N;
    uint32_t *psi_inv = malloc(sizeof(uint32_t) * (size_t)N);
    orc_ps_psi(ps, j, NULL, psi_inv);
    uint32_t t = 1;
    for (uint32_t m = (uint32_t)N; m > 1; m >>= 1) {
        uint32_t j1 = 0, h = m >> 1;
        for (uint32_t i = 0; i < h; ++i) {
            uint32_t S = psi_inv[h + i];
            for (uint32_t x = j1; x < j1 + t; ++x) {
                uint32_t U = a[x], V = a[x + t];
                a[x] = ps_add(ps, U, V, j);
                a[x + t] = ps_mul(ps, ps_sub(ps, U, V, j), S, j);
            }
            j1 += 2 * t;
        }
        t <<= 1;
    }
    const uint32_t ninv = orc_ps_ninv(ps, j);
    for (int i = 0; i < N; ++i) a[i] = ps_mul(ps, a[i], ninv, j);
    free(psi_inv);
}

/* sample_polya (src/sample.c:56-128, SEC_LEVEL 0 branch): coefficient i =
 * LE64(keystream[8i..8i+8)) mod q_j, keystream AES-256-CTR(seed, nonce 0). */
void orc_ps_sample_polya(const orc_pset *ps, const uint8_t seed[32], uint32_t *a)
{
    const int N = ps->N;
    orc_stream st;
    uint8_t *buf = malloc((size_t)8 * N);
    orc_stream_init(&st, seed, 0);
    orc_stream_squeeze(&st, buf, (uint64_t)N / 16);
    for (int i = 0; i < N; ++i) {
        uint64_t v = 0;
        for (int b = 7; b >= 0; --b) v = (v << 8) | buf[8 * i + b];
        for (int j = 0; j < ps->nmod; ++j) a[(size_t)j * N + i] = (uint32_t)(v % ps->q[j]);
    }
    free(buf);
}

/* gaussian_sampler_S{which} (src/gauss.c:705-785, :915-993, :1086-1163) with
 * N samples per call and the set's width: out [nmod][N]. */
void orc_ps_gauss(const orc_pset *ps, orc_stream *st, int which, uint32_t *out)
{
    const int N = ps->N;
    uint8_t r[384];
    uint64_t z[8];
    uint8_t sg[8], ac[8];
    int got = 0;
    while (got < N) {
        orc_stream_squeeze(st, r, 3);
        if (orc_gauss_chunk_k(r, ps->k[which], z, sg, ac) != 0) abort();
        for (int c = 0; c < 8 && got < N; ++c) {
            if (!ac[c]) continue;
            for (int j = 0; j < ps->nmod; ++j) {
                uint32_t m = orc_ps_mod_prime(ps, z[c], j);
                out[(size_t)j * N + got] = sg[c] ? orc_ps_mod_prime(ps, (uint32_t)(ps->q[j] - m), j) : m;
            }
            ++got;
        }
    }
}

#define PSPOLY(ps, base, i, j) ((base) + ((size_t)(i) * (ps)->nmod + (j)) * (ps)->N)

/* rlwe_sife_setup (src/rlwe_sife.c:32-82) given a in the coefficient domain:
 * mpk [l+1][nmod][N] (index l = a-hat), msk [l][nmod][N]. */
void orc_ps_setup_from_a(const orc_pset *ps, const uint32_t *a_coeff, const uint8_t seedS[32],
                         const uint8_t seedE[32], int l, uint32_t *mpk, uint32_t *msk)
{
    const int N = ps->N, nm = ps->nmod;
    orc_stream S, E;
    uint32_t *e = malloc(sizeof(uint32_t) * (size_t)nm * N);
    uint32_t *tmp = malloc(sizeof(uint32_t) * (size_t)N);
    orc_stream_init(&S, seedS, 0);
    orc_stream_init(&E, seedE, 0);
    for (int j = 0; j < nm; ++j) {
        memcpy(PSPOLY(ps, mpk, l, j), a_coeff + (size_t)j * N, sizeof(uint32_t) * N);
        orc_ps_ct_forward(ps, PSPOLY(ps, mpk, l, j), j);
    }
    for (int i = 0; i < l; ++i) {
        orc_ps_gauss(ps, &S, 1, PSPOLY(ps, msk, i, 0));
        orc_ps_gauss(ps, &E, 1, e);
        for (int j = 0; j < nm; ++j) {
            uint32_t *pk = PSPOLY(ps, mpk, i, j);
            memcpy(tmp, PSPOLY(ps, msk, i, j), sizeof(uint32_t) * N);
            orc_ps_ct_forward(ps, tmp, j);
            orc_ps_ct_forward(ps, e + (size_t)j * N, j);
            for (int k = 0; k < N; ++k)
                pk[k] = ps_add(ps, ps_mul(ps, PSPOLY(ps, mpk, l, j)[k], tmp[k], j), e[(size_t)j * N + k], j);
        }
    }
    free(e);
    free(tmp);
}

/* rlwe_sife_encrypt (src/rlwe_sife.c:112-181) of one message of length l:
 * c [l+1][nmod][N] (index l = c_L). */
void orc_ps_encrypt(const orc_pset *ps, const uint32_t *x, int l, const uint32_t *mpk, const uint8_t seed2[32],
                    const uint8_t seed3[32], uint16_t nonce, uint32_t *c)
{
    const int N = ps->N, nm = ps->nmod;
    orc_stream S2, S3;
    uint32_t *r = malloc(sizeof(uint32_t) * (size_t)nm * N);
    uint32_t *f = malloc(sizeof(uint32_t) * (size_t)nm * N);
    uint32_t *mcrt = malloc(sizeof(uint32_t) * (size_t)nm * (size_t)(l > 0 ? l : 1));
    /* crt_convert_generic + scaling by floor(Q/p) (src/crt.c:34-44, src/rlwe_sife.c:129-137) */
    for (int i = 0; i < l; ++i)
        for (int j = 0; j < nm; ++j) {
            uint32_t mj = orc_ps_mod_prime(ps, x[i], j);
            mcrt[(size_t)j * l + i] = orc_ps_mod_prime(ps, (uint64_t)mj * orc_ps_scale(ps, j), j);
        }
    orc_stream_init(&S2, seed2, nonce);
    orc_stream_init(&S3, seed3, nonce);
    orc_ps_gauss(ps, &S2, 2, r);
    orc_ps_gauss(ps, &S2, 2, f);
    for (int j = 0; j < nm; ++j) orc_ps_ct_forward(ps, r + (size_t)j * N, j);
    for (int j = 0; j < nm; ++j) {
        uint32_t *cl = PSPOLY(ps, c, l, j);
        for (int k = 0; k < N; ++k) cl[k] = ps_mul(ps, PSPOLY(ps, mpk, l, j)[k], r[(size_t)j * N + k], j);
        orc_ps_gs_reverse(ps, cl, j);
        for (int k = 0; k < N; ++k) cl[k] = ps_add(ps, cl[k], f[(size_t)j * N + k], j);
    }
    for (int i = 0; i < l; ++i) {
        orc_ps_gauss(ps, &S3, 3, f);
        for (int j = 0; j < nm; ++j) {
            uint32_t *ci = PSPOLY(ps, c, i, j);
            for (int k = 0; k < N; ++k) ci[k] = ps_mul(ps, PSPOLY(ps, mpk, i, j)[k], r[(size_t)j * N + k], j);
            orc_ps_gs_reverse(ps, ci, j);
            for (int k = 0; k < N; ++k) ci[k] = ps_add(ps, ci[k], f[(size_t)j * N + k], j);
            for (int k = 0; k < N; ++k) ci[k] = orc_ps_mod_prime(ps, (uint32_t)(ci[k] + mcrt[(size_t)j * l + i]), j);
        }
    }
    free(r);
    free(f);
    free(mcrt);
}

/* rlwe_sife_keygen (src/rlwe_sife.c:215-245): sk [nmod][N], dense over l slots. */
void orc_ps_keygen(const orc_pset *ps, const uint32_t *y, int l, const uint32_t *msk, uint32_t *sk)
{
    const int N = ps->N, nm = ps->nmod;
    memset(sk, 0, sizeof(uint32_t) * (size_t)nm * N);
    for (int i = 0; i < l; ++i)
        for (int j = 0; j < nm; ++j) {
            uint64_t yj = orc_ps_mod_prime(ps, y[i], j);
            const uint32_t *m = PSPOLY(ps, msk, i, j);
            uint32_t *s = sk + (size_t)j * N;
            for (int k = 0; k < N; ++k) s[k] = orc_ps_mod_prime(ps, yj * m[k] + s[k], j);
        }
}

/* rlwe_sife_decrypt_gmp (src/rlwe_sife.c:248-280) up to the RNS result d [nmod][N]. */
void orc_ps_decrypt(const orc_pset *ps, const uint32_t *c, const uint32_t *y, int l, const uint32_t *sk,
                    uint32_t *d)
{
    const int N = ps->N, nm = ps->nmod;
    uint32_t *a = malloc(sizeof(uint32_t) * (size_t)N);
    uint32_t *b = malloc(sizeof(uint32_t) * (size_t)N);
    memset(d, 0, sizeof(uint32_t) * (size_t)nm * N);
    for (int i = 0; i < l; ++i)
        for (int j = 0; j < nm; ++j) {
            uint64_t yj = orc_ps_mod_prime(ps, y[i], j);
            const uint32_t *ci = PSPOLY(ps, c, i, j);
            uint32_t *dj = d + (size_t)j * N;
            for (int k = 0; k < N; ++k) dj[k] = orc_ps_mod_prime(ps, yj * ci[k] + dj[k], j);
        }
    /* poly_mul_ntt (src/ntt.c:10-26) of c_L and sk_y, then poly_sub_mod */
    for (int j = 0; j < nm; ++j) {
        memcpy(a, PSPOLY(ps, c, l, j), sizeof(uint32_t) * N);
        memcpy(b, sk + (size_t)j * N, sizeof(uint32_t) * N);
        orc_ps_ct_forward(ps, a, j);
        orc_ps_ct_forward(ps, b, j);
        for (int k = 0; k < N; ++k) a[k] = ps_mul(ps, a[k], b[k], j);
        orc_ps_gs_reverse(ps, a, j);
        uint32_t *dj = d + (size_t)j * N;
        for (int k = 0; k < N; ++k) dj[k] = ps_sub(ps, dj[k], a[k], j);
    }
    free(a);
    free(b);
}

/* crt_reverse_gmp (src/crt.c:46-82) for one coefficient: Garner starting at the
 * limb-0 residue as given, u_i = ((v_i - x) C_i) mod q_i taken non-negative as
 * mpz_mod_ui does, with C_i = (q_0 ... q_{i-1})^-1 mod q_i, x += u_i q_0..q_{i-1}. */
void orc_ps_crt(const orc_pset *ps, const uint32_t *res, uint64_t *lo, uint64_t *hi)
{
    unsigned __int128 x = res[0], M = 1;
    for (int i = 1; i < ps->nmod; ++i) {
        const uint64_t q = ps->q[i];
        M *= ps->q[i - 1];
        const uint64_t Mq = (uint64_t)(M % q);
        const uint64_t Ci = ps_pow(Mq, q - 2, q);
        const uint64_t xm = (uint64_t)(x % q);
        const uint64_t diff = ((uint64_t)res[i] % q + q - xm) % q;
        const uint64_t u = (uint64_t)(((unsigned __int128)diff * Ci) % q);
        x += (unsigned __int128)u * M;
    }
    *lo = (uint64_t)x;
    *hi = (uint64_t)(x >> 64);
}

/* round_extract_gmp (src/rlwe_sife.c:288-332): floor(x p / Q), plus one when the
 * remainder exceeds floor(Q/2); no reduction mod p. */
uint32_t orc_ps_round(const orc_pset *ps, uint64_t lo, uint64_t hi)
{
    unsigned __int128 Q = 1;
    for (int i = 0; i < ps->nmod; ++i) Q *= ps->q[i];
    const unsigned __int128 x = ((unsigned __int128)hi << 64) | lo;
    const unsigned __int128 a = x * ps->p;
    unsigned __int128 quo = a / Q;
    const unsigned __int128 rem = a % Q;
    if (rem > Q / 2) quo += 1;
    return (uint32_t)quo;
}
