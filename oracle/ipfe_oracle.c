/*
 * ipfe_oracle.c -- CPU restatement of the reference RLWE-IPFE path.
 *
 * TEST INFRASTRUCTURE ONLY (the checker).  Loaded by tests/, by
 * __graft_entry__.smoke() and by bench.py's cpu_baseline leg; never by the
 * product library.  Compile with -ffp-contract=off: the Bernoulli step must use
 * exactly the IEEE operations the reference's AVX2 intrinsics perform.
 *
 * Written from the reference's behaviour, not from its source text: AES is a
 * plain FIPS-197 implementation whose tables are generated from GF(2^8)
 * arithmetic at start-up, and the NTT twiddles are generated from the closed
 * form w^bitrev12(k).
 */
#include "ipfe_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

/* ------------------------------------------------------------------------- */
/* Parameters: src/params.h:14-28                                             */
/* ------------------------------------------------------------------------- */
static const uint32_t Q[3] = {16760833u, 2147352577u, 2130706433u};
/* q_j = 2^a - 2^b + 1 with (a, b) per limb; used by the fold in mod_prime
 * (src/arith_rns.c:7-16). */
static const uint32_t FOLD_A[3] = {24, 31, 31};
static const uint32_t FOLD_B[3] = {14, 17, 24};
/* primitive 2N-th roots of unity generating the twiddle tables */
static const uint32_t ROOT[3] = {3689u, 1689795u, 456478u};
/* p = B_x * B_y * L + 1 (src/params.h:26) */
#define ORC_P 50241u

/* Q = q0*q1*q2 = 76687145727357674227351553 (src/params.h:14) */
static unsigned __int128 big_Q(void)
{
    return (unsigned __int128)Q[0] * Q[1] * Q[2];
}

uint32_t orc_modulus(uint32_t sel) { return Q[sel]; }

static uint32_t powmod(uint32_t b, uint64_t e, uint32_t q)
{
    uint64_t r = 1, x = b % q;
    while (e) {
        if (e & 1) r = r * x % q;
        x = x * x % q;
        e >>= 1;
    }
    return (uint32_t)r;
}

uint32_t orc_ninv(uint32_t sel) { return powmod(ORC_N, Q[sel] - 2, Q[sel]); }

/* floor(Q/p) mod q_j (src/params.h:27-28 SIFE_SCALE_M_MOD_Q_I) */
uint32_t orc_scale_mod(uint32_t sel)
{
    unsigned __int128 s = big_Q() / ORC_P;
    return (uint32_t)(s % Q[sel]);
}

/* ------------------------------------------------------------------------- */
/* AES-256 (FIPS-197), counter mode as in src/aes256ctr.c:13-155              */
/* ------------------------------------------------------------------------- */
static uint8_t SBOX[256];
static uint32_t TE[4][256];
static int aes_ready;

static uint8_t gf_mul(uint8_t a, uint8_t b)
{
    uint8_t r = 0;
    while (b) {
        if (b & 1) r ^= a;
        a = (uint8_t)((a << 1) ^ ((a & 0x80) ? 0x1b : 0));
        b >>= 1;
    }
    return r;
}

static uint8_t rotl8(uint8_t x, int s) { return (uint8_t)((x << s) | (x >> (8 - s))); }

static void aes_tables(void)
{
    if (aes_ready) return;
    for (int x = 0; x < 256; ++x) {
        uint8_t inv = 0;
        if (x) {
            for (int y = 1; y < 256; ++y)
                if (gf_mul((uint8_t)x, (uint8_t)y) == 1) { inv = (uint8_t)y; break; }
        }
        SBOX[x] = (uint8_t)(inv ^ rotl8(inv, 1) ^ rotl8(inv, 2) ^ rotl8(inv, 3) ^ rotl8(inv, 4) ^ 0x63);
    }
    for (int x = 0; x < 256; ++x) {
        uint8_t s = SBOX[x];
        uint32_t w = ((uint32_t)gf_mul(s, 2) << 24) | ((uint32_t)s << 16) | ((uint32_t)s << 8) | gf_mul(s, 3);
        TE[0][x] = w;
        TE[1][x] = (w >> 8) | (w << 24);
        TE[2][x] = (w >> 16) | (w << 16);
        TE[3][x] = (w >> 24) | (w << 8);
    }
    aes_ready = 1;
}

static uint32_t be32(const uint8_t *p)
{
    return ((uint32_t)p[0] << 24) | ((uint32_t)p[1] << 16) | ((uint32_t)p[2] << 8) | p[3];
}

static void put_be32(uint8_t *p, uint32_t v)
{
    p[0] = (uint8_t)(v >> 24); p[1] = (uint8_t)(v >> 16); p[2] = (uint8_t)(v >> 8); p[3] = (uint8_t)v;
}

static uint32_t sub_word(uint32_t w)
{
    return ((uint32_t)SBOX[w >> 24] << 24) | ((uint32_t)SBOX[(w >> 16) & 255] << 16) |
           ((uint32_t)SBOX[(w >> 8) & 255] << 8) | SBOX[w & 255];
}

/* AES-256 key schedule (src/aes256ctr.c:83-139 computes the same schedule with
 * aeskeygenassist). */
void orc_aes256_expand(const uint8_t key[32], uint32_t rk[60])
{
    static const uint8_t rcon[7] = {0x01, 0x02, 0x04, 0x08, 0x10, 0x20, 0x40};
    aes_tables();
    for (int i = 0; i < 8; ++i) rk[i] = be32(key + 4 * i);
    for (int i = 8; i < 60; ++i) {
        uint32_t t = rk[i - 1];
        if (i % 8 == 0)
            t = sub_word((t << 8) | (t >> 24)) ^ ((uint32_t)rcon[i / 8 - 1] << 24);
        else if (i % 8 == 4)
            t = sub_word(t);
        rk[i] = rk[i - 8] ^ t;
    }
}

void orc_aes256_encrypt_block(const uint32_t rk[60], const uint8_t in[16], uint8_t out[16])
{
    aes_tables();
    uint32_t s0 = be32(in) ^ rk[0], s1 = be32(in + 4) ^ rk[1];
    uint32_t s2 = be32(in + 8) ^ rk[2], s3 = be32(in + 12) ^ rk[3];
    for (int r = 1; r < 14; ++r) {
        uint32_t t0 = TE[0][s0 >> 24] ^ TE[1][(s1 >> 16) & 255] ^ TE[2][(s2 >> 8) & 255] ^ TE[3][s3 & 255] ^ rk[4 * r];
        uint32_t t1 = TE[0][s1 >> 24] ^ TE[1][(s2 >> 16) & 255] ^ TE[2][(s3 >> 8) & 255] ^ TE[3][s0 & 255] ^ rk[4 * r + 1];
        uint32_t t2 = TE[0][s2 >> 24] ^ TE[1][(s3 >> 16) & 255] ^ TE[2][(s0 >> 8) & 255] ^ TE[3][s1 & 255] ^ rk[4 * r + 2];
        uint32_t t3 = TE[0][s3 >> 24] ^ TE[1][(s0 >> 16) & 255] ^ TE[2][(s1 >> 8) & 255] ^ TE[3][s2 & 255] ^ rk[4 * r + 3];
        s0 = t0; s1 = t1; s2 = t2; s3 = t3;
    }
    uint32_t o0 = ((uint32_t)SBOX[s0 >> 24] << 24) | ((uint32_t)SBOX[(s1 >> 16) & 255] << 16) |
                  ((uint32_t)SBOX[(s2 >> 8) & 255] << 8) | SBOX[s3 & 255];
    uint32_t o1 = ((uint32_t)SBOX[s1 >> 24] << 24) | ((uint32_t)SBOX[(s2 >> 16) & 255] << 16) |
                  ((uint32_t)SBOX[(s3 >> 8) & 255] << 8) | SBOX[s0 & 255];
    uint32_t o2 = ((uint32_t)SBOX[s2 >> 24] << 24) | ((uint32_t)SBOX[(s3 >> 16) & 255] << 16) |
                  ((uint32_t)SBOX[(s0 >> 8) & 255] << 8) | SBOX[s1 & 255];
    uint32_t o3 = ((uint32_t)SBOX[s3 >> 24] << 24) | ((uint32_t)SBOX[(s0 >> 16) & 255] << 16) |
                  ((uint32_t)SBOX[(s1 >> 8) & 255] << 8) | SBOX[s2 & 255];
    put_be32(out, o0 ^ rk[56]);
    put_be32(out + 4, o1 ^ rk[57]);
    put_be32(out + 8, o2 ^ rk[58]);
    put_be32(out + 12, o3 ^ rk[59]);
}

void orc_stream_init(orc_stream *st, const uint8_t key[32], uint16_t nonce)
{
    orc_aes256_expand(key, st->rk);
    st->ctr = 0;
    st->nonce = nonce;
}

/* Block t of the stream encrypts BE16(nonce) || 0^10 || BE32(t): the
 * _mm_set_epi64x / add_epi32 / byte-swap sequence of src/aes256ctr.c:27-40 with
 * the state set at src/aes256ctr.c:92. */
void orc_stream_squeeze(orc_stream *st, uint8_t *out, uint64_t nblocks128)
{
    uint8_t blk[16];
    for (uint64_t u = 0; u < nblocks128; ++u) {
        for (int b = 0; b < 8; ++b) {
            memset(blk, 0, sizeof blk);
            blk[0] = (uint8_t)(st->nonce >> 8);
            blk[1] = (uint8_t)(st->nonce & 255);
            put_be32(blk + 12, st->ctr + (uint32_t)b);
            orc_aes256_encrypt_block(st->rk, blk, out + 128 * u + 16 * b);
        }
        st->ctr += 8;
    }
}

/* ------------------------------------------------------------------------- */
/* RNS arithmetic: src/arith_rns.c:22-103                                     */
/* ------------------------------------------------------------------------- */

/* Special-form fold, repeated while m > 2q, then one conditional subtract
 * (src/arith_rns.c:61-83).  Returns q (not 0) when the fold stops exactly at 2q. */
uint32_t orc_mod_prime(uint64_t m, uint32_t sel)
{
    const uint64_t q = Q[sel];
    const uint32_t a = FOLD_A[sel], b = FOLD_B[sel];
    const uint64_t lo = ((uint64_t)1 << a) - 1;
    while (m > 2 * q) {
        uint64_t hi = m >> a;
        m = (m & lo) + ((hi << b) - hi);
    }
    if (m >= q) m -= q;
    return (uint32_t)m;
}

uint32_t orc_add_mod(uint32_t a, uint32_t b, uint32_t sel)
{
    uint64_t c = (uint64_t)a + b;
    if (c >= Q[sel]) c -= Q[sel];
    return (uint32_t)c;
}

uint32_t orc_sub_mod(uint32_t a, uint32_t b, uint32_t sel)
{
    uint64_t c = (uint64_t)a + Q[sel] - b;
    if (c >= Q[sel]) c -= Q[sel];
    return (uint32_t)c;
}

uint32_t orc_mul_mod(uint32_t a, uint32_t b, uint32_t sel)
{
    return orc_mod_prime((uint64_t)a * b, sel);
}

/* ------------------------------------------------------------------------- */
/* Twiddles: psi[k] = w^bitrev12(k), psi_inv[k] = w^-bitrev12(k)              */
/* (the tables of src/consts.h:9-11, generated from their closed form)        */
/* ------------------------------------------------------------------------- */
static uint32_t PSI[3][ORC_N], PSI_INV[3][ORC_N];
static int psi_ready;

static uint32_t bitrev12(uint32_t k)
{
    uint32_t r = 0;
    for (int i = 0; i < 12; ++i) r |= ((k >> i) & 1u) << (11 - i);
    return r;
}

static void psi_tables(void)
{
    if (psi_ready) return;
    for (int s = 0; s < 3; ++s) {
        uint32_t winv = powmod(ROOT[s], Q[s] - 2, Q[s]);
        for (uint32_t k = 0; k < ORC_N; ++k) {
            PSI[s][k] = powmod(ROOT[s], bitrev12(k), Q[s]);
            PSI_INV[s][k] = powmod(winv, bitrev12(k), Q[s]);
        }
    }
    psi_ready = 1;
}

const uint32_t *orc_psi(uint32_t sel) { psi_tables(); return PSI[sel]; }
const uint32_t *orc_psi_inv(uint32_t sel) { psi_tables(); return PSI_INV[sel]; }

/* ------------------------------------------------------------------------- */
/* NTT: src/ntt.c:10-114                                                      */
/* ------------------------------------------------------------------------- */

/* Cooley-Tukey, natural order in, bit-reversed order out (src/ntt.c:38-67). */
void orc_ct_forward(uint32_t a[ORC_N], uint32_t sel)
{
    psi_tables();
    uint32_t t = ORC_N;
    for (uint32_t m = 1; m < ORC_N; m <<= 1) {
        t >>= 1;
        for (uint32_t i = 0; i < m; ++i) {
            uint32_t j1 = 2 * i * t;
            uint32_t S = PSI[sel][m + i];
            for (uint32_t j = j1; j < j1 + t; ++j) {
                uint32_t U = a[j];
                uint32_t V = orc_mul_mod(a[j + t], S, sel);
                a[j] = orc_add_mod(U, V, sel);
                a[j + t] = orc_sub_mod(U, V, sel);
            }
        }
    }
}

/* Gentleman-Sande, bit-reversed in, natural out, times N^-1 (src/ntt.c:70-114). */
void orc_gs_reverse(uint32_t a[ORC_N], uint32_t sel)
{
    psi_tables();
    uint32_t t = 1;
    for (uint32_t m = ORC_N; m > 1; m >>= 1) {
        uint32_t j1 = 0, h = m >> 1;
        for (uint32_t i = 0; i < h; ++i) {
            uint32_t S = PSI_INV[sel][h + i];
            for (uint32_t j = j1; j < j1 + t; ++j) {
                uint32_t U = a[j], V = a[j + t];
                a[j] = orc_add_mod(U, V, sel);
                a[j + t] = orc_mul_mod(orc_sub_mod(U, V, sel), S, sel);
            }
            j1 += 2 * t;
        }
        t <<= 1;
    }
    uint32_t ninv = orc_ninv(sel);
    for (uint32_t i = 0; i < ORC_N; ++i) a[i] = orc_mul_mod(a[i], ninv, sel);
}

void orc_point_mul(const uint32_t *a, const uint32_t *b, uint32_t *c, uint32_t sel)
{
    for (int i = 0; i < ORC_N; ++i) c[i] = orc_mul_mod(a[i], b[i], sel);
}

/* src/ntt.c:10-26 */
void orc_poly_mul_ntt(const uint32_t *a, const uint32_t *b, uint32_t *c, uint32_t sel)
{
    uint32_t at[ORC_N], bt[ORC_N];
    memcpy(at, a, sizeof at);
    memcpy(bt, b, sizeof bt);
    orc_ct_forward(at, sel);
    orc_ct_forward(bt, sel);
    orc_point_mul(at, bt, c, sel);
    orc_gs_reverse(c, sel);
}

/* ------------------------------------------------------------------------- */
/* Uniform polynomial a: src/sample.c:56-128                                  */
/* ------------------------------------------------------------------------- */
static uint64_t le64(const uint8_t *p)
{
    uint64_t v = 0;
    for (int i = 7; i >= 0; --i) v = (v << 8) | p[i];
    return v;
}

static uint32_t le32(const uint8_t *p)
{
    return (uint32_t)p[0] | ((uint32_t)p[1] << 8) | ((uint32_t)p[2] << 16) | ((uint32_t)p[3] << 24);
}

/* a_j[i] = LE64(ks[8i..8i+8)) mod q_j for the AES-256-CTR(seed, nonce 0) keystream.
 * Every draw is kept because 2^64 < Q, so the refill loop (src/sample.c:101-123)
 * never runs.  The reference squeezes 352 x 128 bytes and reads the first 32 KiB. */
void orc_sample_polya(const uint8_t seed[32], uint32_t a[ORC_NMODULI][ORC_N])
{
    orc_stream st;
    uint8_t *buf = (uint8_t *)malloc(352 * 128);
    orc_stream_init(&st, seed, 0);
    orc_stream_squeeze(&st, buf, 352);
    for (int i = 0; i < ORC_N; ++i) {
        uint64_t v = le64(buf + 8 * i);
        for (int j = 0; j < 3; ++j) a[j][i] = (uint32_t)(v % Q[j]);
    }
    free(buf);
}

/* ------------------------------------------------------------------------- */
/* Discrete Gaussian sampler: src/gauss.c:11-1163 (AVX2 "_ori" / S3 path)     */
/* ------------------------------------------------------------------------- */
static const uint64_t SAMPLER_K[4] = {0, 266ull, 304214978ull, 608429953ull};

/* CDT rows (low word, high word): src/gauss.c:62-72 */
static const uint64_t CDT_TAB[9][2] = {
    {2200310400551559144ull, 3327841033070651387ull},
    {7912151619254726620ull, 380075531178589176ull},
    {5167367257772081627ull, 11604843442081400ull},
    {5081592746475748971ull, 90134450315532ull},
    {6522074513864805092ull, 175786317361ull},
    {2579734681240182346ull, 85801740ull},
    {8175784047440310133ull, 10472ull},
    {2947787991558061753ull, 0ull},
    {22489665999543ull, 0ull}};

/* Polynomial for 2^a on [0,1): src/gauss.c:82-103 (EXP_COFF, IEEE-754 bit patterns) */
static const uint64_t EXP_COEF[10] = {
    0x3e833b70ffa2c5d4ull, 0x3eb4a480fda7e6e1ull, 0x3ef01b254493363full, 0x3f242e0e0aa273ccull,
    0x3f55d8a2334ed31bull, 0x3f83b2aa56db0f1aull, 0x3fac6b08e11fc57eull, 0x3fcebfbdff556072ull,
    0x3fe62e42fefa7fe6ull, 0x3ff0000000000000ull};

static double as_double(uint64_t u) { double d; memcpy(&d, &u, 8); return d; }
static uint64_t as_u64(double d) { uint64_t u; memcpy(&u, &d, 8); return u; }

/* Bernoulli(2^(-b/k^2)) decision bit for one candidate, given its 9 random
 * bytes (8 at rb + 8*lane, 1 at rb + 32 + lane).  src/gauss.c:204-264 (S1),
 * :361-416 (S2), :521-576 (S3); the three differ only in k. */
static int bernoulli(uint64_t b, uint64_t k, const uint8_t *rb, int lane)
{
    const double k2inv = -1.0 / (double)(int64_t)(k * k);
    /* int64 -> double by OR-ing in the exponent of 2^52; wrong on purpose for
     * b >= 2^52, exactly as the reference computes it */
    double vx = as_double(b | 0x4330000000000000ull) - 4503599627370496.0;
    vx = vx * k2inv;
    double fl = floor(vx);
    double shifted = fl + 6755399441055744.0; /* 1.5 * 2^52 */
    uint64_t t = (as_u64(shifted) - 0x4338000000000000ull) << 52;
    double a = vx - fl;
    double s = fma(as_double(EXP_COEF[0]), a, as_double(EXP_COEF[1]));
    for (int i = 2; i < 10; ++i) s = fma(s, a, as_double(EXP_COEF[i]));
    uint64_t res = t + as_u64(s);
    uint64_t res_mant = (res & ((1ull << 52) - 1)) | (1ull << 52);
    uint64_t e = (res >> 52) + (uint64_t)(int64_t)-1003; /* R_EXPONENT_L - 1023 + 1 */
    uint64_t res_exp = (e < 64) ? (1ull << e) : 0;        /* vpsllvq: counts >= 64 give 0 */
    uint64_t rm = le64(rb + 8 * lane);
    uint64_t r_exp = (rm >> 53) | ((uint64_t)rb[32 + lane] << 11);
    uint64_t r_mant = rm & ((1ull << 53) - 1);
    uint64_t bits = ((r_mant - res_mant) & (r_exp - res_exp)) |
                    (res == 0x3FF0000000000000ull ? ~0ull : 0ull);
    return (int)(bits >> 63);
}

/* One 384-byte refill -> 8 candidates.  Layout of the chunk (src/gauss.c:725-783):
 *   group g = c>>2 (base 100g): CDT words at base .. base+63, Bernoulli bytes at
 *   base+64 .. base+99; uniform words from byte 200; sign bits in byte 280. */
int orc_gauss_chunk(const uint8_t r[384], int which, uint64_t z[8], uint8_t sign[8], uint8_t acc[8])
{
    return orc_gauss_chunk_k(r, SAMPLER_K[which], z, sign, acc);
}

/* The chunk step for any sampler width k (S1 / S2 / S3 differ only in k). */
int orc_gauss_chunk_k(const uint8_t r[384], uint64_t k, uint64_t z[8], uint8_t sign[8], uint8_t acc[8])
{
    const uint64_t bf = (1ull << 32) / k;
    const uint64_t uq = k * bf;
    uint64_t y[8];
    /* uniform_sampler_S*: keep 32-bit words below k*floor(2^32/k) (src/gauss.c:609-703) */
    int u = 0;
    for (int j = 0; j < 8; ++j) {
        uint64_t w;
        do {
            if (200 + 4 * u + 4 > 384) return -1;
            w = le32(r + 200 + 4 * u);
            ++u;
        } while (!(w < uq));
        uint64_t x = w - ((((w * bf) >> 32) + 1) * k);
        x = x + (x >> 63) * k;
        y[j] = x;
    }
    for (int c = 0; c < 8; ++c) {
        const int g = c >> 2, lane = c & 3;
        const uint8_t *base = r + 100 * g;
        /* cdt_sampler (src/gauss.c:122-154) */
        uint64_t r1 = le64(base + 8 * lane) & 0x7fffffffffffffffull;
        uint64_t r2 = le64(base + 32 + 8 * lane) & 0x7fffffffffffffffull;
        uint64_t x = 0;
        for (int i = 0; i < 9; ++i) {
            uint64_t lt_lo = r1 - CDT_TAB[i][0];
            uint64_t lt_hi = r2 - CDT_TAB[i][1];
            uint64_t eq_hi = (r2 == CDT_TAB[i][1]) ? ~0ull : 0ull;
            x += ((lt_lo & eq_hi) | lt_hi) >> 63;
        }
        /* _mm256_mul_epu32: low 32 bits of each operand */
        uint64_t xk = (uint64_t)(uint32_t)x * (uint64_t)(uint32_t)k;
        uint64_t zz = xk + y[c];
        uint64_t bin = (uint64_t)(uint32_t)(zz + xk) * (uint64_t)(uint32_t)y[c];
        int bit = bernoulli(bin, k, base + 64, lane);
        int s = (r[280] >> c) & 1;
        z[c] = zz;
        sign[c] = (uint8_t)s;
        /* accept iff Bernoulli bit and not (z == 0 and sign == 0) */
        acc[c] = (uint8_t)(bit && (zz != 0 || s != 0));
    }
    return 0;
}

/* gaussian_sampler_S1_ori / S2_ori / S3: each call starts on a fresh chunk, takes
 * accepted candidates in order and drops the rest of its last chunk. */
void orc_gauss(orc_stream *st, int which, uint32_t out[ORC_NMODULI][ORC_N], uint32_t slen)
{
    uint8_t r[384];
    uint64_t z[8];
    uint8_t sg[8], ac[8];
    uint32_t j = 0;
    while (j < slen) {
        orc_stream_squeeze(st, r, 3);
        if (orc_gauss_chunk(r, which, z, sg, ac) != 0) abort();
        for (int c = 0; c < 8 && j < slen; ++c) {
            if (!ac[c]) continue;
            for (uint32_t s = 0; s < 3; ++s) {
                uint32_t m = orc_mod_prime(z[c], s);
                out[s][j] = sg[c] ? orc_mod_prime((uint32_t)(Q[s] - m), s) : m;
            }
            ++j;
        }
    }
}

/* ------------------------------------------------------------------------- */
/* Scheme: src/rlwe_sife.c                                                    */
/* ------------------------------------------------------------------------- */
#define POLY(base, i, j) ((base) + ((size_t)(i) * 3 + (j)) * ORC_N)

/* rlwe_sife_setup (src/rlwe_sife.c:32-82) given a in the coefficient domain. */
void orc_setup_from_a(const uint32_t *a_coeff, const uint8_t seedS[32], const uint8_t seedE[32],
                      int l, uint32_t *mpk, uint32_t *msk)
{
    orc_stream S, E;
    uint32_t (*e)[ORC_N] = malloc(sizeof(uint32_t) * 3 * ORC_N);
    uint32_t *tmp = malloc(sizeof(uint32_t) * ORC_N);
    orc_stream_init(&S, seedS, 0);
    orc_stream_init(&E, seedE, 0);
    for (int j = 0; j < 3; ++j) {
        memcpy(POLY(mpk, l, j), a_coeff + (size_t)j * ORC_N, sizeof(uint32_t) * ORC_N);
        orc_ct_forward(POLY(mpk, l, j), j);
    }
    for (int i = 0; i < l; ++i) {
        orc_gauss(&S, 1, (uint32_t(*)[ORC_N])POLY(msk, i, 0), ORC_N);
        orc_gauss(&E, 1, e, ORC_N);
        for (int j = 0; j < 3; ++j) {
            memcpy(tmp, POLY(msk, i, j), sizeof(uint32_t) * ORC_N);
            orc_ct_forward(tmp, j);
            orc_ct_forward(e[j], j);
            orc_point_mul(POLY(mpk, l, j), tmp, POLY(mpk, i, j), j);
            for (int k = 0; k < ORC_N; ++k) POLY(mpk, i, j)[k] = orc_add_mod(POLY(mpk, i, j)[k], e[j][k], j);
        }
    }
    free(e);
    free(tmp);
}

void orc_setup(const uint8_t seedA[32], const uint8_t seedS[32], const uint8_t seedE[32],
               int l, uint32_t *mpk, uint32_t *msk)
{
    uint32_t (*a)[ORC_N] = malloc(sizeof(uint32_t) * 3 * ORC_N);
    orc_sample_polya(seedA, a);
    orc_setup_from_a(&a[0][0], seedS, seedE, l, mpk, msk);
    free(a);
}

/* rlwe_sife_encrypt (src/rlwe_sife.c:112-181) for one message of length l. */
void orc_encrypt(const uint32_t *x, int l, const uint32_t *mpk, const uint8_t seed2[32],
                 const uint8_t seed3[32], uint16_t nonce, uint32_t *c)
{
    orc_stream S2, S3;
    uint32_t (*r)[ORC_N] = malloc(sizeof(uint32_t) * 3 * ORC_N);
    uint32_t (*f)[ORC_N] = malloc(sizeof(uint32_t) * 3 * ORC_N);
    uint32_t *mcrt = malloc(sizeof(uint32_t) * 3 * (size_t)(l > 0 ? l : 1));
    /* crt_convert_generic + scaling by floor(Q/p) (src/crt.c:34-44, src/rlwe_sife.c:129-137) */
    for (int i = 0; i < l; ++i)
        for (int j = 0; j < 3; ++j) {
            uint32_t mj = orc_mod_prime(x[i], j);
            mcrt[(size_t)j * l + i] = orc_mod_prime((uint64_t)mj * orc_scale_mod(j), j);
        }
    orc_stream_init(&S2, seed2, nonce);
    orc_stream_init(&S3, seed3, nonce);
    orc_gauss(&S2, 2, r, ORC_N);
    orc_gauss(&S2, 2, f, ORC_N);
    for (int j = 0; j < 3; ++j) orc_ct_forward(r[j], j);
    for (int j = 0; j < 3; ++j) {
        uint32_t *cl = POLY(c, l, j);
        orc_point_mul(POLY(mpk, l, j), r[j], cl, j);
        orc_gs_reverse(cl, j);
        for (int k = 0; k < ORC_N; ++k) cl[k] = orc_add_mod(cl[k], f[j][k], j);
    }
    for (int i = 0; i < l; ++i) {
        orc_gauss(&S3, 3, f, ORC_N);
        for (int j = 0; j < 3; ++j) {
            uint32_t *ci = POLY(c, i, j);
            orc_point_mul(POLY(mpk, i, j), r[j], ci, j);
            orc_gs_reverse(ci, j);
            for (int k = 0; k < ORC_N; ++k) ci[k] = orc_add_mod(ci[k], f[j][k], j);
            for (int k = 0; k < ORC_N; ++k) ci[k] = orc_mod_prime((uint32_t)(ci[k] + mcrt[(size_t)j * l + i]), j);
        }
    }
    free(r);
    free(f);
    free(mcrt);
}

/* rlwe_sife_keygen (src/rlwe_sife.c:215-245): dense over all l slots. */
void orc_keygen(const uint32_t *y, int l, const uint32_t *msk, uint32_t *sk)
{
    memset(sk, 0, sizeof(uint32_t) * 3 * ORC_N);
    for (int i = 0; i < l; ++i)
        for (int j = 0; j < 3; ++j) {
            uint64_t yj = orc_mod_prime(y[i], j);
            const uint32_t *m = POLY(msk, i, j);
            uint32_t *s = sk + (size_t)j * ORC_N;
            for (int k = 0; k < ORC_N; ++k) s[k] = orc_mod_prime(yj * m[k] + s[k], j);
        }
}

/* rlwe_sife_decrypt_gmp (src/rlwe_sife.c:248-280) up to the RNS result d_y. */
void orc_decrypt(const uint32_t *c, const uint32_t *y, int l, const uint32_t *sk, uint32_t *d)
{
    uint32_t *prod = malloc(sizeof(uint32_t) * ORC_N);
    memset(d, 0, sizeof(uint32_t) * 3 * ORC_N);
    for (int i = 0; i < l; ++i)
        for (int j = 0; j < 3; ++j) {
            uint64_t yj = orc_mod_prime(y[i], j);
            const uint32_t *ci = POLY(c, i, j);
            uint32_t *dj = d + (size_t)j * ORC_N;
            for (int k = 0; k < ORC_N; ++k) dj[k] = orc_mod_prime(yj * ci[k] + dj[k], j);
        }
    for (int j = 0; j < 3; ++j) {
        orc_poly_mul_ntt(POLY(c, l, j), sk + (size_t)j * ORC_N, prod, j);
        uint32_t *dj = d + (size_t)j * ORC_N;
        for (int k = 0; k < ORC_N; ++k) dj[k] = orc_sub_mod(dj[k], prod[k], j);
    }
    free(prod);
}

/* crt_reverse_gmp (src/crt.c:46-82): Garner with x starting at the limb-0
 * residue as given (q0 stays q0), u_i = ((v_i - x) * C_i) mod q_i taken
 * non-negative as mpz_mod_ui does, x += u_i * prod_{j<i} q_j.  The CRT
 * constants (src/params.h:17) are q0^-1 mod q1 and an unreduced 64-bit
 * representative of (q0 q1)^-1 mod q2; only their residues matter. */
void orc_crt_reverse(const uint32_t d0, const uint32_t d1, const uint32_t d2, uint64_t *lo, uint64_t *hi)
{
    const uint64_t c1 = powmod(Q[0] % Q[1], Q[1] - 2, Q[1]);
    const uint64_t q01m2 = ((uint64_t)Q[0] * Q[1]) % Q[2];
    const uint64_t c2 = powmod((uint32_t)q01m2, Q[2] - 2, Q[2]);
    unsigned __int128 x = d0;
    /* i = 1 */
    {
        uint64_t xm = (uint64_t)(x % Q[1]);
        uint64_t diff = ((uint64_t)d1 % Q[1] + Q[1] - xm) % Q[1];
        uint64_t u = (diff * c1) % Q[1];
        x += (unsigned __int128)u * Q[0];
    }
    /* i = 2 */
    {
        uint64_t xm = (uint64_t)(x % Q[2]);
        uint64_t diff = ((uint64_t)d2 % Q[2] + Q[2] - xm) % Q[2];
        uint64_t u = (diff * c2) % Q[2];
        x += (unsigned __int128)u * Q[0] * Q[1];
    }
    *lo = (uint64_t)x;
    *hi = (uint64_t)(x >> 64);
}

/* round_extract_gmp (src/rlwe_sife.c:288-332): q = floor(x p / Q), plus one when
 * the remainder exceeds floor(Q/2).  No reduction mod p. */
uint32_t orc_round_extract(uint64_t lo, uint64_t hi)
{
    unsigned __int128 x = ((unsigned __int128)hi << 64) | lo;
    unsigned __int128 Qb = big_Q();
    unsigned __int128 a = x * ORC_P;
    unsigned __int128 quo = a / Qb, rem = a % Qb;
    if (rem > Qb / 2) quo += 1;
    return (uint32_t)quo;
}

double orc_round_extract_coef0(const uint32_t *d)
{
    uint64_t lo, hi;
    orc_crt_reverse(d[0], d[ORC_N], d[2 * ORC_N], &lo, &hi);
    return (double)orc_round_extract(lo, hi);
}

/* Key values of many (row, key) pairs from coefficient 0 alone: what
 * svm_predict_cpu (main.c:170-207) keeps of each decryption, kvalue =
 * round_extract(CRT(d_y[.][0])).  Per limb, d_y[0] = sub_mod(acc, prod0)
 * (src/rlwe_sife.c:265-277) with
 *   acc   = the fold acc <- mod_prime(y_i c_i[0] + acc) over i < ly, where slots
 *           at and beyond the row length l hold the encryption of 0, whose
 *           product with y_i = 0 is 0 (the key must be 0 there);
 *   prod0 = coefficient 0 of the negacyclic product c_L * sk_y, i.e.
 *           a_0 b_0 - sum_{j>=1} a_j b_{N-j} mod q (the closed form of
 *           GS_reverse(CT_forward(a) . CT_forward(b))[0]), canonical.
 * The reference's INTT may return q instead of 0 for prod0; that changes d_y
 * only when acc = q (a fold ending on a non-zero slot).  Such pairs are
 * counted in the return value and their key value is set to 0xffffffff.
 * Layouts: cl [R][3][N], c0 [R][l][3], y [S][ly], sk [S][3][N], kv [R][S]. */
int orc_kvalues_coef0(const uint32_t *cl, const uint32_t *c0, int R, int l, const uint32_t *y, int ly,
                      const uint32_t *sk, int S, uint32_t *kv)
{
    int ambiguous = 0;
    for (int r = 0; r < R; ++r)
        for (int v = 0; v < S; ++v) {
            const uint32_t *yv = y + (size_t)v * ly;
            uint32_t d[3];
            int amb = 0;
            for (int j = 0; j < 3; ++j) {
                const uint64_t q = Q[j];
                uint32_t acc = 0;
                for (int i = 0; i < ly; ++i) {
                    const uint64_t ci = i < l ? c0[((size_t)r * l + i) * 3 + j] : 0;
                    acc = orc_mod_prime((uint64_t)orc_mod_prime(yv[i], j) * ci + acc, j);
                }
                const uint32_t *a = cl + ((size_t)r * 3 + j) * ORC_N;
                const uint32_t *b = sk + ((size_t)v * 3 + j) * ORC_N;
                unsigned __int128 minus = 0;
                for (int k = 1; k < ORC_N; ++k) minus += (uint64_t)a[k] * b[ORC_N - k];
                const uint64_t plus = (uint64_t)(((unsigned __int128)a[0] * b[0]) % q);
                const uint64_t prod0 = (plus + q - (uint64_t)(minus % q)) % q;
                if (acc == q) amb = 1;
                d[j] = orc_sub_mod(acc, (uint32_t)prod0, j);
            }
            uint64_t lo, hi;
            orc_crt_reverse(d[0], d[1], d[2], &lo, &hi);
            kv[(size_t)r * S + v] = amb ? 0xffffffffu : orc_round_extract(lo, hi);
            ambiguous += amb;
        }
    return ambiguous;
}

/* One-vs-one vote (main.c:338-389). */
double orc_svm_decide(int nr_class, int l_sv, const int *nSV, const int *label,
                      const double *sv_coef, const double *rho, const double *kvalue, int use_fma)
{
    int *start = malloc(sizeof(int) * (size_t)nr_class);
    int *vote = calloc((size_t)nr_class, sizeof(int));
    start[0] = 0;
    for (int i = 1; i < nr_class; ++i) start[i] = start[i - 1] + nSV[i - 1];
    int p = 0;
    for (int i = 0; i < nr_class; ++i)
        for (int j = i + 1; j < nr_class; ++j) {
            double sum = 0;
            const double *coef1 = sv_coef + (size_t)(j - 1) * l_sv;
            const double *coef2 = sv_coef + (size_t)i * l_sv;
            for (int k = 0; k < nSV[i]; ++k) {
                int s = start[i] + k;
                sum = use_fma ? fma(coef1[s], kvalue[s], sum) : sum + coef1[s] * kvalue[s];
            }
            for (int k = 0; k < nSV[j]; ++k) {
                int s = start[j] + k;
                sum = use_fma ? fma(coef2[s], kvalue[s], sum) : sum + coef2[s] * kvalue[s];
            }
            sum -= rho[p];
            if (sum > 0) ++vote[i]; else ++vote[j];
            ++p;
        }
    int best = 0;
    for (int i = 1; i < nr_class; ++i)
        if (vote[i] > vote[best]) best = i;
    double out = label[best];
    free(start);
    free(vote);
    return out;
}

/* svm_predict_cpu (main.c:137-267): per SV a dense keygen, a full decryption,
 * CRT and rounding of all N coefficients (as crt_reverse_gmp and
 * round_extract_gmp do), keeping coefficient 0. */
double orc_svm_predict_cpu(const uint32_t *c, int l, const uint32_t *msk, int nr_class, int l_sv,
                           const int *nSV, const int *label, const double *sv_coef,
                           const double *rho, const uint32_t *y, double *kvalue_out, int use_fma)
{
    uint32_t *sk = malloc(sizeof(uint32_t) * 3 * ORC_N);
    uint32_t *d = malloc(sizeof(uint32_t) * 3 * ORC_N);
    double *kv = kvalue_out ? kvalue_out : malloc(sizeof(double) * (size_t)l_sv);
    volatile uint32_t sink = 0;
    for (int s = 0; s < l_sv; ++s) {
        const uint32_t *ys = y + (size_t)s * l;
        orc_keygen(ys, l, msk, sk);
        orc_decrypt(c, ys, l, sk, d);
        uint32_t r0 = 0;
        for (int k = 0; k < ORC_N; ++k) {
            uint64_t lo, hi;
            orc_crt_reverse(d[k], d[ORC_N + k], d[2 * ORC_N + k], &lo, &hi);
            uint32_t r = orc_round_extract(lo, hi);
            if (k == 0) r0 = r;
            sink ^= r;
        }
        kv[s] = (double)r0;
    }
    (void)sink;
    double lab = orc_svm_decide(nr_class, l_sv, nSV, label, sv_coef, rho, kv, use_fma);
    free(sk);
    free(d);
    if (!kvalue_out) free(kv);
    return lab;
}
