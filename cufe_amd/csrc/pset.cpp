// pset.cpp -- the three parameter sets (pset.h), their device tables, and the
// single-operation C API for every set: NTT (through the autotuned shapes),
// Gaussian sampling, sample_polya, setup / encrypt / keygen / decrypt and CRT
// + rounding, each following the reference function of the same name with
// the set's N, limbs, primes and sampler widths.
#include <stdio.h>
#include <string.h>

#include <mutex>
#include <string>
#include <vector>

#include "../../include/cufe_ipfe.h"
#include "kernels_pset.h"
#include "ntt_variants.h"
#include "runtime.h"

namespace ipfe {

namespace {

const PsetSpec kSets[PSET_COUNT] = {
    {0, 2048, 11, 2, 785, {2147352577u, 2130706433u, 0, 0}, {31, 31, 0, 0}, {17, 24, 0, 0},
     {1543607367u, 1989134074u, 0, 0}, {0, 266ull, 262147ull, 524287ull}, {0, 1, 4, 5}, 50241u},
    {1, 4096, 12, 3, 785, {16760833u, 2147352577u, 2130706433u, 0}, {24, 31, 31, 0}, {14, 17, 24, 0},
     {3689u, 1689795u, 456478u, 0}, {0, 266ull, 304214978ull, 608429953ull}, {0, 1, 2, 3}, 50241u},
    {2, 8192, 13, 4, 785, {16760833u, 2147352577u, 2130706433u, 67043329u}, {24, 31, 31, 26}, {14, 17, 24, 16},
     {5768458u, 772388167u, 1047035213u, 8157893u}, {0, 266ull, 304214978ull, 608429953ull}, {0, 1, 2, 3}, 50241u},
};

// Twiddle tables of sets 0 and 2 ([4][nmod][N] each); set 1 uses the
// library's reference tables.  Static arrays: no allocation.
__device__ __align__(256) uint32_t d_tw_set0[4 * 2 * 2048];
__device__ __align__(256) uint32_t d_tw_set2[4 * 4 * 8192];

uint32_t powmod(uint64_t b, uint64_t e, uint64_t q)
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

uint32_t bitrev(uint32_t k, int bits)
{
    uint32_t r = 0;
    for (int i = 0; i < bits; ++i) r |= ((k >> i) & 1u) << (bits - 1 - i);
    return r;
}

PsetDev make_dev(const PsetSpec &sp)
{
    PsetDev d;
    memset(&d, 0, sizeof d);
    d.N = sp.N;
    d.logN = sp.logN;
    d.nmod = sp.nmod;
    d.id = sp.id;
    d.p = sp.p;
    unsigned __int128 Q = 1;
    for (int j = 0; j < sp.nmod; ++j) Q *= sp.q[j];
    d.Q_lo = (uint64_t)Q;
    d.Q_hi = (uint64_t)(Q >> 64);
    d.halfQ_lo = (uint64_t)(Q / 2);
    d.halfQ_hi = (uint64_t)((Q / 2) >> 64);
    unsigned __int128 M = 1;
    for (int j = 0; j < sp.nmod; ++j) {
        const uint32_t q = sp.q[j];
        d.q[j] = q;
        d.fa[j] = sp.fa[j];
        d.fb[j] = sp.fb[j];
        d.ninv[j] = powmod((uint64_t)sp.N, q - 2, q);
        d.ninvp[j] = (uint32_t)(((uint64_t)d.ninv[j] << 32) / q);
        d.ninvf[j] = mod_prime_gen((uint64_t)q * d.ninv[j], q, sp.fa[j], sp.fb[j]) == q ? 1u : 0u;
        d.scale[j] = (uint32_t)((Q / sp.p) % q);
        const Barrett b = make_barrett(q);
        d.mu_hi[j] = b.mu_hi;
        d.mu_lo[j] = b.mu_lo;
        if (j > 0) {
            d.garner_c[j] = powmod((uint64_t)(M % q), q - 2, q);
            d.garner_m_lo[j] = (uint64_t)M;
            d.garner_m_hi[j] = (uint64_t)(M >> 64);
        }
        M *= q;
    }
    return d;
}

// [4][nmod][N]: w^bitrev(k) | flag << 31, Shoup companion, the same for w^-1;
// flag = 1 when mod_prime(q w) returns q.
std::vector<uint32_t> make_twiddles(const PsetSpec &sp)
{
    const int N = sp.N, nm = sp.nmod;
    std::vector<uint32_t> t((size_t)4 * nm * N);
    for (int s = 0; s < nm; ++s) {
        const uint32_t q = sp.q[s];
        const uint32_t winv = powmod(sp.root[s], q - 2, q);
        for (int k = 0; k < N; ++k) {
            const uint32_t br = bitrev((uint32_t)k, sp.logN);
            const uint32_t w = powmod(sp.root[s], br, q), wi = powmod(winv, br, q);
            const uint32_t fw = mod_prime_gen((uint64_t)q * w, q, sp.fa[s], sp.fb[s]) == q ? 1u : 0u;
            const uint32_t fwi = mod_prime_gen((uint64_t)q * wi, q, sp.fa[s], sp.fb[s]) == q ? 1u : 0u;
            t[((size_t)0 * nm + s) * N + k] = w | (fw << 31);
            t[((size_t)1 * nm + s) * N + k] = (uint32_t)(((uint64_t)w << 32) / q);
            t[((size_t)2 * nm + s) * N + k] = wi | (fwi << 31);
            t[((size_t)3 * nm + s) * N + k] = (uint32_t)(((uint64_t)wi << 32) / q);
        }
    }
    return t;
}

std::mutex g_ps_mu;
PsetState g_states[PSET_COUNT];
bool g_ready[PSET_COUNT] = {false, false, false};

}  // namespace

const PsetSpec *pset_spec(int id) { return (id >= 0 && id < PSET_COUNT) ? &kSets[id] : nullptr; }

const PsetState *pset_state(int id)
{
    const PsetSpec *sp = pset_spec(id);
    if (!sp) {
        ipfe_set_error("no parameter set %d (0: N=2048/2 limbs, 1: reference N=4096/3 limbs, 2: N=8192/4 limbs)", id);
        return nullptr;
    }
    Device *dev = device();
    if (!dev) return nullptr;
    std::lock_guard<std::mutex> lk(g_ps_mu);
    if (g_ready[id]) return &g_states[id];
    PsetState &st = g_states[id];
    st.dev = make_dev(*sp);
    if (id == 1) {
        st.tw = dev->tw_all;  // the reference tables, same layout
    } else {
        const std::vector<uint32_t> t = make_twiddles(*sp);
        void *p = nullptr;
        const hipError_t e = id == 0 ? hipGetSymbolAddress(&p, HIP_SYMBOL(d_tw_set0))
                                     : hipGetSymbolAddress(&p, HIP_SYMBOL(d_tw_set2));
        if (e != hipSuccess || !p) {
            ipfe_set_error("parameter set %d: twiddle table address unavailable", id);
            return nullptr;
        }
        if (t.size() * 4 != (id == 0 ? sizeof(d_tw_set0) : sizeof(d_tw_set2)) ||
            hipMemcpy(p, t.data(), t.size() * 4, hipMemcpyHostToDevice) != hipSuccess) {
            ipfe_set_error("parameter set %d: twiddle upload failed", id);
            return nullptr;
        }
        st.tw = (uint32_t *)p;
    }
    g_ready[id] = true;
    return &st;
}

NttArgs ntt_args(const PsetState &st)
{
    NttArgs a;
    memset(&a, 0, sizeof a);
    a.nmod = st.dev.nmod;
    for (int s = 0; s < st.dev.nmod; ++s) {
        a.q[s] = st.dev.q[s];
        a.ninv[s] = st.dev.ninv[s];
        a.ninvp[s] = st.dev.ninvp[s];
        a.ninvf[s] = st.dev.ninvf[s];
    }
    return a;
}

// In-place (or out-of-place) transforms of ntrans limb-transforms of set st
// with the tuned shape.
int ps_ntt(const PsetState &st, bool inverse, const uint32_t *in, uint32_t *out, int ntrans, hipStream_t s)
{
    if (ntrans <= 0) return 0;
    const NttArgs a = ntt_args(st);
    const int v = ntt_tuned_variant(st.dev.id, st.dev.logN, inverse, a, st.tw, st.dev.id == 1, s);
    if (v < 0) return -1;
    return ntt_run_variant(st.dev.logN, v, inverse, in, out, ntrans, a, st.tw, st.dev.id == 1, s);
}

}  // namespace ipfe

using namespace ipfe;

namespace {

const PsetState *check_set(int pset, const char *fn)
{
    const PsetState *st = pset_state(pset);
    if (!st) {
        const std::string e = cufe_last_error();
        ipfe_set_error("%s: %s", fn, e.c_str());
    }
    return st;
}

size_t poly_words(const PsetState &st) { return (size_t)st.dev.nmod * st.dev.N; }

}  // namespace

extern "C" int cufe_pset_info(int pset, int *n, int *nmod, int *l, uint32_t *q, uint32_t *p)
{
    const PsetSpec *sp = pset_spec(pset);
    if (!sp) {
        ipfe_set_error("cufe_pset_info: no parameter set %d", pset);
        return -1;
    }
    if (n) *n = sp->N;
    if (nmod) *nmod = sp->nmod;
    if (l) *l = sp->L;
    if (q)
        for (int j = 0; j < 4; ++j) q[j] = j < sp->nmod ? sp->q[j] : 0u;
    if (p) *p = sp->p;
    return 0;
}

static int ps_ntt_host(int pset, int inverse, int variant, uint32_t *polys, int npoly, const char *fn)
{
    const PsetState *st = check_set(pset, fn);
    if (!st) return -1;
    if (npoly < 0 || !polys) {
        ipfe_set_error("%s: bad arguments", fn);
        return -1;
    }
    if (npoly == 0) return 0;
    hipStream_t s = device()->stream;
    const size_t words = (size_t)npoly * poly_words(*st);
    DevBuf<uint32_t> d;
    if (d.alloc(words)) return -1;
    HIP_OK(hipMemcpyAsync(d.get(), polys, words * 4, hipMemcpyHostToDevice, s));
    const int ntrans = npoly * st->dev.nmod;
    if (variant < 0) {
        if (ps_ntt(*st, inverse != 0, d.get(), d.get(), ntrans, s)) return -1;
    } else if (ntt_run_variant(st->dev.logN, variant, inverse != 0, d.get(), d.get(), ntrans, ntt_args(*st), st->tw,
                               pset == 1, s)) {
        return -1;
    }
    HIP_OK(hipMemcpyAsync(polys, d.get(), words * 4, hipMemcpyDeviceToHost, s));
    return sync_stream(s);
}

extern "C" int cufe_ps_ntt(int pset, int inverse, uint32_t *polys, int npoly)
{
    return ps_ntt_host(pset, inverse, -1, polys, npoly, "cufe_ps_ntt");
}

extern "C" int cufe_ps_ntt_variant(int pset, int variant, int inverse, uint32_t *polys, int npoly)
{
    const PsetSpec *sp = pset_spec(pset);
    if (!sp || !ntt_variant_ok(sp->logN, variant, pset == 1)) {
        ipfe_set_error("cufe_ps_ntt_variant: no variant %d for parameter set %d", variant, pset);
        return -1;
    }
    return ps_ntt_host(pset, inverse, variant, polys, npoly, "cufe_ps_ntt_variant");
}

extern "C" int cufe_ntt_variants(int pset, char *names, int len)
{
    const PsetSpec *sp = pset_spec(pset);
    if (!sp) {
        ipfe_set_error("cufe_ntt_variants: no parameter set %d", pset);
        return -1;
    }
    const int n = ntt_variant_count(sp->logN);
    std::string s;
    for (int v = 0; v < n; ++v) {
        s += ntt_variant_ok(sp->logN, v, pset == 1) ? ntt_variant_name(sp->logN, v) : "-";
        s += "\n";
    }
    if (names && len > 0) snprintf(names, (size_t)len, "%s", s.c_str());
    return n;
}

extern "C" int cufe_ntt_autotune_report(char *buf, int len)
{
    const std::string r = ntt_autotune_report();
    if (buf && len > 0) snprintf(buf, (size_t)len, "%s", r.c_str());
    return (int)r.size();
}

extern "C" int cufe_ps_sample_polya(int pset, const unsigned char seed[32], uint32_t *a)
{
    const PsetState *st = check_set(pset, "cufe_ps_sample_polya");
    if (!st) return -1;
    hipStream_t s = device()->stream;
    uint32_t rk[60];
    aes256_expand_key(seed, rk);
    DevBuf<uint32_t> d_rk, d_a;
    if (d_rk.alloc(60) || d_a.alloc(poly_words(*st))) return -1;
    HIP_OK(hipMemcpyAsync(d_rk.get(), rk, sizeof rk, hipMemcpyHostToDevice, s));
    if (launch_sample_polya_ps(d_rk.get(), d_a.get(), st->dev.N, st->dev.nmod, st->dev.q, s)) return -1;
    HIP_OK(hipMemcpyAsync(a, d_a.get(), poly_words(*st) * 4, hipMemcpyDeviceToHost, s));
    return sync_stream(s);
}

// Streams of set pset into device memory d [nstreams][ncalls][nmod][N].
static int ps_gauss_dev(const PsetState &st, const unsigned char *keys, const uint16_t *nonces, int nstreams,
                        int which, int ncalls, uint32_t *d, hipStream_t s)
{
    const PsetSpec &sp = *pset_spec(st.dev.id);
    const size_t per_call = poly_words(st), per_stream = (size_t)ncalls * per_call;
    std::vector<StreamReq> reqs;
    for (int t = 0; t < nstreams; ++t) {
        uint32_t *o = d + (size_t)t * per_stream;
        reqs.push_back(make_stream_ps(sp, keys + 32 * (size_t)t, nonces[t], which, ncalls, o, o + per_call,
                                      (long long)per_call));
    }
    return run_gauss(reqs, s);
}

extern "C" int cufe_ps_gauss_sample(int pset, const unsigned char *keys, const uint16_t *nonces, int nstreams,
                                    int which, int ncalls, uint32_t *out)
{
    const PsetState *st = check_set(pset, "cufe_ps_gauss_sample");
    if (!st) return -1;
    if (which < 1 || which > 3 || ncalls < 0 || nstreams < 0) {
        ipfe_set_error("cufe_ps_gauss_sample: bad arguments (which=%d ncalls=%d nstreams=%d)", which, ncalls,
                       nstreams);
        return -1;
    }
    if (ncalls == 0 || nstreams == 0) return 0;
    hipStream_t s = device()->stream;
    const size_t words = (size_t)nstreams * ncalls * poly_words(*st);
    DevBuf<uint32_t> d;
    if (d.alloc(words)) return -1;
    if (ps_gauss_dev(*st, keys, nonces, nstreams, which, ncalls, d.get(), s)) return -1;
    HIP_OK(hipMemcpyAsync(out, d.get(), words * 4, hipMemcpyDeviceToHost, s));
    return sync_stream(s);
}

// rlwe_sife_setup (src/rlwe_sife.c:32-82): on entry mpk[l] holds a (coefficient
// domain); s_i from AES-256-CTR(seed2, nonce 0), e_i from (seed3, nonce 0).
extern "C" int cufe_ps_setup(int pset, uint32_t *mpk, uint32_t *msk, const unsigned char seed2[32],
                             const unsigned char seed3[32], int l)
{
    const PsetState *st = check_set(pset, "cufe_ps_setup");
    if (!st) return -1;
    if (l < 1 || l > st->dev.N || !mpk || !msk) {
        ipfe_set_error("cufe_ps_setup: bad arguments (l=%d)", l);
        return -1;
    }
    hipStream_t s = device()->stream;
    const size_t pw = poly_words(*st);
    DevBuf<uint32_t> d_mpk, d_msk, d_e, d_ap;
    if (d_mpk.alloc((size_t)(l + 1) * pw) || d_msk.alloc((size_t)l * pw) || d_e.alloc((size_t)l * pw) ||
        d_ap.alloc(pw))
        return -1;
    uint32_t *ahat = d_mpk.get() + (size_t)l * pw;
    HIP_OK(hipMemcpyAsync(ahat, mpk + (size_t)l * pw, pw * 4, hipMemcpyHostToDevice, s));
    if (ps_ntt(*st, false, ahat, ahat, st->dev.nmod, s) || ps_launch_shoup(*st, ahat, d_ap.get(), pw, s)) return -1;
    const uint16_t zero = 0;
    const PsetSpec &sp = *pset_spec(pset);
    std::vector<StreamReq> reqs;
    reqs.push_back(make_stream_ps(sp, seed2, zero, 1, l, d_msk.get(), d_msk.get() + pw, (long long)pw));
    reqs.push_back(make_stream_ps(sp, seed3, zero, 1, l, d_e.get(), d_e.get() + pw, (long long)pw));
    if (run_gauss(reqs, s)) return -1;
    if (ps_launch_setup_pk(*st, d_msk.get(), d_e.get(), ahat, d_ap.get(), d_mpk.get(), l, s)) return -1;
    HIP_OK(hipMemcpyAsync(mpk, d_mpk.get(), (size_t)(l + 1) * pw * 4, hipMemcpyDeviceToHost, s));
    HIP_OK(hipMemcpyAsync(msk, d_msk.get(), (size_t)l * pw * 4, hipMemcpyDeviceToHost, s));
    return sync_stream(s);
}

// rlwe_sife_encrypt (src/rlwe_sife.c:112-181) of repeat messages m [repeat][l];
// message b draws r, f_0 from (seed2, nonce0 + b) and f_i from (seed3, nonce0 + b).
extern "C" int cufe_ps_encrypt(int pset, const uint32_t *m, const uint32_t *mpk, uint32_t *c,
                               const unsigned char seed2[32], const unsigned char seed3[32], int repeat, int l,
                               int nonce0)
{
    const PsetState *st = check_set(pset, "cufe_ps_encrypt");
    if (!st) return -1;
    if (l < 0 || repeat < 0 || nonce0 < 0 || nonce0 + repeat > 65536 || !mpk || !c || (l > 0 && !m)) {
        ipfe_set_error("cufe_ps_encrypt: bad arguments (l=%d repeat=%d nonce0=%d)", l, repeat, nonce0);
        return -1;
    }
    if (repeat == 0) return 0;
    hipStream_t s = device()->stream;
    const size_t pw = poly_words(*st), ct = (size_t)(l + 1) * pw;
    DevBuf<uint32_t> d_mpk, d_c, d_r, d_rp, d_x;
    if (d_mpk.alloc((size_t)(l + 1) * pw) || d_c.alloc((size_t)repeat * ct) || d_r.alloc((size_t)repeat * pw) ||
        d_rp.alloc((size_t)repeat * pw) || d_x.alloc((size_t)repeat * l + 1))
        return -1;
    HIP_OK(hipMemcpyAsync(d_mpk.get(), mpk, (size_t)(l + 1) * pw * 4, hipMemcpyHostToDevice, s));
    if (l > 0) HIP_OK(hipMemcpyAsync(d_x.get(), m, (size_t)repeat * l * 4, hipMemcpyHostToDevice, s));
    const PsetSpec &sp = *pset_spec(pset);
    std::vector<StreamReq> reqs;
    for (int b = 0; b < repeat; ++b) {
        const uint16_t nonce = (uint16_t)(nonce0 + b);
        uint32_t *cb = d_c.get() + (size_t)b * ct;
        // r (call 0) and f_0 into c_L (call 1) on seed2; f_0 .. f_{l-1} into c_i on seed3
        reqs.push_back(make_stream_ps(sp, seed2, nonce, 2, 2, d_r.get() + (size_t)b * pw, cb + (size_t)l * pw, 0));
        if (l > 0) reqs.push_back(make_stream_ps(sp, seed3, nonce, 3, l, cb, cb + pw, (long long)pw));
    }
    if (run_gauss(reqs, s)) return -1;
    if (ps_ntt(*st, false, d_r.get(), d_r.get(), repeat * st->dev.nmod, s) ||
        ps_launch_shoup(*st, d_r.get(), d_rp.get(), (size_t)repeat * pw, s) ||
        ps_launch_enc_poly(*st, d_mpk.get(), d_r.get(), d_rp.get(), d_x.get(), d_c.get(), repeat, l, s))
        return -1;
    HIP_OK(hipMemcpyAsync(c, d_c.get(), (size_t)repeat * ct * 4, hipMemcpyDeviceToHost, s));
    return sync_stream(s);
}

// rlwe_sife_keygen (src/rlwe_sife.c:215-245): sk [repeat][nmod][N] from y [repeat][l].
extern "C" int cufe_ps_keygen(int pset, const uint32_t *y, const uint32_t *msk, uint32_t *sk, int repeat, int l)
{
    const PsetState *st = check_set(pset, "cufe_ps_keygen");
    if (!st) return -1;
    if (l < 1 || repeat < 0 || !y || !msk || !sk) {
        ipfe_set_error("cufe_ps_keygen: bad arguments (l=%d repeat=%d)", l, repeat);
        return -1;
    }
    if (repeat == 0) return 0;
    hipStream_t s = device()->stream;
    const size_t pw = poly_words(*st);
    DevBuf<uint32_t> d_y, d_msk, d_sk;
    if (d_y.alloc((size_t)repeat * l) || d_msk.alloc((size_t)l * pw) || d_sk.alloc((size_t)repeat * pw)) return -1;
    HIP_OK(hipMemcpyAsync(d_y.get(), y, (size_t)repeat * l * 4, hipMemcpyHostToDevice, s));
    HIP_OK(hipMemcpyAsync(d_msk.get(), msk, (size_t)l * pw * 4, hipMemcpyHostToDevice, s));
    if (ps_launch_fold(*st, d_y.get(), repeat, l, d_msk.get(), d_sk.get(), s)) return -1;
    HIP_OK(hipMemcpyAsync(sk, d_sk.get(), (size_t)repeat * pw * 4, hipMemcpyDeviceToHost, s));
    return sync_stream(s);
}

// rlwe_sife_decrypt_gmp (src/rlwe_sife.c:248-280) up to the RNS result: one
// ciphertext c [l+1][nmod][N] against repeat keys -> d [repeat][nmod][N].
extern "C" int cufe_ps_decrypt(int pset, const uint32_t *c, const uint32_t *y, const uint32_t *sk, uint32_t *d,
                               int repeat, int l)
{
    const PsetState *st = check_set(pset, "cufe_ps_decrypt");
    if (!st) return -1;
    if (l < 1 || repeat < 0 || !c || !y || !sk || !d) {
        ipfe_set_error("cufe_ps_decrypt: bad arguments (l=%d repeat=%d)", l, repeat);
        return -1;
    }
    if (repeat == 0) return 0;
    hipStream_t s = device()->stream;
    const size_t pw = poly_words(*st), nk = (size_t)repeat * pw;
    DevBuf<uint32_t> d_c, d_y, d_sk, d_d, d_cl;
    if (d_c.alloc((size_t)(l + 1) * pw) || d_y.alloc((size_t)repeat * l) || d_sk.alloc(nk) || d_d.alloc(nk) ||
        d_cl.alloc(nk))
        return -1;
    HIP_OK(hipMemcpyAsync(d_c.get(), c, (size_t)(l + 1) * pw * 4, hipMemcpyHostToDevice, s));
    HIP_OK(hipMemcpyAsync(d_y.get(), y, (size_t)repeat * l * 4, hipMemcpyHostToDevice, s));
    HIP_OK(hipMemcpyAsync(d_sk.get(), sk, nk * 4, hipMemcpyHostToDevice, s));
    // sum part: fold over the l slots
    if (ps_launch_fold(*st, d_y.get(), repeat, l, d_c.get(), d_d.get(), s)) return -1;
    // poly_mul_ntt(c_L, sk_v) (src/ntt.c:10-26): NTT(c_L) once, copied per key
    uint32_t *cl = d_c.get() + (size_t)l * pw;
    if (ps_ntt(*st, false, cl, cl, st->dev.nmod, s)) return -1;
    for (int v = 0; v < repeat; ++v)
        HIP_OK(hipMemcpyAsync(d_cl.get() + (size_t)v * pw, cl, pw * 4, hipMemcpyDeviceToDevice, s));
    if (ps_ntt(*st, false, d_sk.get(), d_sk.get(), repeat * st->dev.nmod, s) ||
        ps_launch_point_mul(*st, d_cl.get(), d_sk.get(), d_cl.get(), nk, s) ||
        ps_ntt(*st, true, d_cl.get(), d_cl.get(), repeat * st->dev.nmod, s) ||
        ps_launch_sub(*st, d_d.get(), d_cl.get(), nk, s))
        return -1;
    HIP_OK(hipMemcpyAsync(d, d_d.get(), nk * 4, hipMemcpyDeviceToHost, s));
    return sync_stream(s);
}

// CRT + rounding (crt_reverse_gmp + round_extract_gmp) of coefficient coef of
// d [nvec][nmod][N] (coef < 0: every coefficient, r [nvec][N]).
extern "C" int cufe_ps_crt_round(int pset, const uint32_t *d, int nvec, int coef, uint32_t *r)
{
    const PsetState *st = check_set(pset, "cufe_ps_crt_round");
    if (!st) return -1;
    if (nvec < 0 || coef >= st->dev.N || !d || !r) {
        ipfe_set_error("cufe_ps_crt_round: bad arguments");
        return -1;
    }
    if (nvec == 0) return 0;
    hipStream_t s = device()->stream;
    const size_t nin = (size_t)nvec * poly_words(*st), nout = (size_t)nvec * (coef < 0 ? st->dev.N : 1);
    DevBuf<uint32_t> d_d, d_r;
    if (d_d.alloc(nin) || d_r.alloc(nout)) return -1;
    HIP_OK(hipMemcpyAsync(d_d.get(), d, nin * 4, hipMemcpyHostToDevice, s));
    if (ps_launch_crt_round(*st, d_d.get(), nvec, coef, d_r.get(), s)) return -1;
    HIP_OK(hipMemcpyAsync(r, d_r.get(), nout * 4, hipMemcpyDeviceToHost, s));
    return sync_stream(s);
}

// The same CRT + rounding on the host (no device): for callers that hold the
// RNS decryption in host memory, as round_extract_gmp2 does.
extern "C" int cufe_ps_crt_round_host(int pset, const uint32_t *res, uint32_t *r)
{
    const PsetSpec *sp = pset_spec(pset);
    if (!sp || !res || !r) {
        ipfe_set_error("cufe_ps_crt_round_host: bad arguments");
        return -1;
    }
    static std::once_flag once;
    static PsetDev devs[PSET_COUNT];
    std::call_once(once, [] {
        for (int i = 0; i < PSET_COUNT; ++i) devs[i] = make_dev(kSets[i]);
    });
    *r = ps_crt_round_value(devs[pset], res);
    return 0;
}

// Device-resident timing of the tuned transforms (npoly polys of the set) and
// of nstreams one-call Gaussian streams.
extern "C" int cufe_time_ps_ntt(int pset, int npoly, int inverse, int iters, float *ms_per_launch)
{
    const PsetState *st = check_set(pset, "cufe_time_ps_ntt");
    if (!st) return -1;
    if (npoly < 1 || iters < 1) {
        ipfe_set_error("cufe_time_ps_ntt: bad arguments");
        return -1;
    }
    hipStream_t s = device()->stream;
    const size_t words = (size_t)npoly * poly_words(*st);
    DevBuf<uint32_t> d;
    if (d.alloc(words)) return -1;
    HIP_OK(hipMemsetD32Async((hipDeviceptr_t)d.get(), 65537u, words, s));
    const int ntrans = npoly * st->dev.nmod;
    for (int w = 0; w < 2; ++w)
        if (ps_ntt(*st, inverse != 0, d.get(), d.get(), ntrans, s)) return -1;
    hipEvent_t a, b;
    HIP_OK(hipEventCreate(&a));
    HIP_OK(hipEventCreate(&b));
    HIP_OK(hipEventRecord(a, s));
    for (int it = 0; it < iters; ++it)
        if (ps_ntt(*st, inverse != 0, d.get(), d.get(), ntrans, s)) return -1;
    HIP_OK(hipEventRecord(b, s));
    HIP_OK(hipEventSynchronize(b));
    float ms = 0;
    HIP_OK(hipEventElapsedTime(&ms, a, b));
    (void)hipEventDestroy(a);
    (void)hipEventDestroy(b);
    if (ms_per_launch) *ms_per_launch = ms / iters;
    return sync_stream(s);
}

extern "C" int cufe_time_ps_gauss(int pset, int nstreams, int which, int iters, float *ms_per_run)
{
    const PsetState *st = check_set(pset, "cufe_time_ps_gauss");
    if (!st) return -1;
    if (nstreams < 1 || nstreams > 65536 || which < 1 || which > 3 || iters < 1) {
        ipfe_set_error("cufe_time_ps_gauss: bad arguments");
        return -1;
    }
    hipStream_t s = device()->stream;
    DevBuf<uint32_t> d;
    if (d.alloc((size_t)nstreams * poly_words(*st))) return -1;
    std::vector<unsigned char> keys((size_t)nstreams * 32, 0);
    std::vector<uint16_t> nonces(nstreams);
    for (int t = 0; t < nstreams; ++t) {
        keys[(size_t)t * 32] = (unsigned char)(which * 17 + 1);
        nonces[t] = (uint16_t)t;
    }
    if (ps_gauss_dev(*st, keys.data(), nonces.data(), nstreams, which, 1, d.get(), s)) return -1;  // warm-up
    hipEvent_t a, b;
    HIP_OK(hipEventCreate(&a));
    HIP_OK(hipEventCreate(&b));
    HIP_OK(hipEventRecord(a, s));
    for (int it = 0; it < iters; ++it)
        if (ps_gauss_dev(*st, keys.data(), nonces.data(), nstreams, which, 1, d.get(), s)) return -1;
    HIP_OK(hipEventRecord(b, s));
    HIP_OK(hipEventSynchronize(b));
    float ms = 0;
    HIP_OK(hipEventElapsedTime(&ms, a, b));
    (void)hipEventDestroy(a);
    (void)hipEventDestroy(b);
    if (ms_per_run) *ms_per_run = ms / iters;
    return sync_stream(s);
}
