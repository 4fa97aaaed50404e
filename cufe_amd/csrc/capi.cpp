// capi.cpp -- the C ABI of include/cufe_ipfe.h: reference drop-ins, host helpers
// and single-operation entry points.
#include <stdio.h>
#include <string.h>

#include <cmath>
#include <vector>

#include "../../include/cufe_ipfe.h"
#include "crt_dev.h"
#include "runtime.h"
#include "kernels_pset.h"

using namespace ipfe;

namespace {

constexpr size_t kPoly = (size_t)3 * IPFE_N;  // words per [3][N] polynomial

struct EventTimer {
    hipEvent_t a = nullptr, b = nullptr;
    hipStream_t st;
    explicit EventTimer(hipStream_t s) : st(s)
    {
        (void)hipEventCreate(&a);
        (void)hipEventCreate(&b);
        (void)hipEventRecord(a, st);
    }
    float stop()
    {
        float ms = 0;
        (void)hipEventRecord(b, st);
        (void)hipEventSynchronize(b);
        (void)hipEventElapsedTime(&ms, a, b);
        return ms;
    }
    ~EventTimer()
    {
        if (a) (void)hipEventDestroy(a);
        if (b) (void)hipEventDestroy(b);
    }
};

void report(const char *fn)
{
    fprintf(stderr, "[cufe_ipfe] %s failed: %s\n", fn, cufe_last_error());
}

// Sparse form of dense vectors y[n][len] (entries with y != 0, ascending).
struct HostSparse {
    std::vector<int> off, idx;
    std::vector<uint32_t> val;
};
HostSparse to_sparse(const uint32_t *y, int n, int len)
{
    HostSparse h;
    h.off.resize(n + 1);
    for (int v = 0; v < n; ++v) {
        h.off[v] = (int)h.idx.size();
        for (int i = 0; i < len; ++i)
            if (y[(size_t)v * len + i]) {
                h.idx.push_back(i);
                h.val.push_back(y[(size_t)v * len + i]);
            }
    }
    h.off[n] = (int)h.idx.size();
    if (h.idx.empty()) {  // keep device pointers valid
        h.idx.push_back(0);
        h.val.push_back(0);
    }
    return h;
}

struct DevSparse {
    DevBuf<int> off, idx;
    DevBuf<uint32_t> val;
    SparseY view{};
    int upload(const HostSparse &h, int len, hipStream_t st)
    {
        if (off.alloc(h.off.size()) || idx.alloc(h.idx.size()) || val.alloc(h.val.size())) return -1;
        HIP_OK(hipMemcpyAsync(off.get(), h.off.data(), h.off.size() * 4, hipMemcpyHostToDevice, st));
        HIP_OK(hipMemcpyAsync(idx.get(), h.idx.data(), h.idx.size() * 4, hipMemcpyHostToDevice, st));
        HIP_OK(hipMemcpyAsync(val.get(), h.val.data(), h.val.size() * 4, hipMemcpyHostToDevice, st));
        view.off = off.get();
        view.idx = idx.get();
        view.val = val.get();
        view.len = len;
        return 0;
    }
};

// ---------------------------------------------------------------------------
int do_setup(uint32_t *mpk, uint32_t *msk, const unsigned char *seed2, const unsigned char *seed3, int l, float *ms)
{
    Device *dev = device();
    if (!dev) return -1;
    hipStream_t st = dev->stream;
    EventTimer tm(st);
    DevBuf<uint32_t> d_mpk, d_msk, d_e, d_ap;
    if (d_mpk.alloc((l + 1) * kPoly) || d_msk.alloc(l * kPoly) || d_e.alloc(l * kPoly) || d_ap.alloc(kPoly)) return -1;
    uint32_t *d_ahat = d_mpk.get() + (size_t)l * kPoly;
    HIP_OK(hipMemcpyAsync(d_ahat, mpk + (size_t)l * kPoly, kPoly * 4, hipMemcpyHostToDevice, st));
    if (launch_ntt_fwd(d_ahat, d_ahat, 3, dev->tw_all, st)) return -1;
    if (launch_shoup_prep(d_ahat, d_ap.get(), kPoly, st)) return -1;
    std::vector<StreamReq> reqs;
    reqs.push_back(make_stream(seed2, 0, 1, l, d_msk.get(), d_msk.get() + kPoly, (long long)kPoly));
    reqs.push_back(make_stream(seed3, 0, 1, l, d_e.get(), d_e.get() + kPoly, (long long)kPoly));
    if (run_gauss(reqs, st)) return -1;
    if (launch_setup_pk(d_msk.get(), d_e.get(), d_ahat, d_ap.get(), d_mpk.get(), l, dev->tw_all, st)) return -1;
    HIP_OK(hipMemcpyAsync(mpk, d_mpk.get(), (l + 1) * kPoly * 4, hipMemcpyDeviceToHost, st));
    HIP_OK(hipMemcpyAsync(msk, d_msk.get(), l * kPoly * 4, hipMemcpyDeviceToHost, st));
    float e = tm.stop();
    if (ms) *ms += e;
    return sync_stream(st);
}

// Encrypt `repeat` messages of length l against mpk [l+1][3][N]; c is
// [repeat][l+1][3][N].  Message b uses nonce nonce0 + b.
int do_encrypt(const uint32_t *m, const uint32_t *mpk, uint32_t *c, const unsigned char *seed2,
               const unsigned char *seed3, int repeat, int l, int nonce0, float *ms)
{
    Device *dev = device();
    if (!dev) return -1;
    if (repeat < 0 || nonce0 < 0 || nonce0 + repeat > 65536) {
        ipfe_set_error("encrypt: nonce range [%d, %d) exceeds the 16-bit nonce space", nonce0, nonce0 + repeat);
        return -1;
    }
    hipStream_t st = dev->stream;
    EventTimer tm(st);
    DevBuf<uint32_t> d_mpk;
    if (d_mpk.alloc((l + 1) * kPoly)) return -1;
    HIP_OK(hipMemcpyAsync(d_mpk.get(), mpk, (l + 1) * kPoly * 4, hipMemcpyHostToDevice, st));
    const size_t ct_words = (size_t)(l + 1) * kPoly;
    // messages per group: bound ciphertext + sampler scratch to ~16 GiB
    const size_t per_msg = ct_words * 4 + gauss_bytes(3, l, 0) + gauss_bytes(2, 2, 0) + 2 * kPoly * 4;
    int group = (int)std::max<size_t>(1, ((size_t)16 << 30) / per_msg);
    for (int b0 = 0; b0 < repeat; b0 += group) {
        const int g = std::min(group, repeat - b0);
        DevBuf<uint32_t> d_x, d_c, d_r, d_rp;
        DevBuf<EncMsg> d_msgs;
        DevBuf<EncItem> d_items;
        if (d_x.alloc((size_t)g * l + 1) || d_c.alloc((size_t)g * ct_words) || d_r.alloc((size_t)g * kPoly) ||
            d_rp.alloc((size_t)g * kPoly) || d_msgs.alloc(g) || d_items.alloc((size_t)g * (l + 1)))
            return -1;
        if (l > 0)
            HIP_OK(hipMemcpyAsync(d_x.get(), m + (size_t)b0 * l, (size_t)g * l * 4, hipMemcpyHostToDevice, st));
        std::vector<StreamReq> reqs;
        std::vector<EncMsg> msgs(g);
        std::vector<EncItem> items;
        for (int b = 0; b < g; ++b) {
            uint32_t *cb = d_c.get() + (size_t)b * ct_words;
            const uint16_t nonce = (uint16_t)(nonce0 + b0 + b);
            // r, f_0 on (seed2, nonce); f_0 lands in the c_L slot
            reqs.push_back(make_stream(seed2, nonce, 2, 2, d_r.get() + (size_t)b * kPoly, cb + (size_t)l * kPoly, 0));
            if (l > 0) reqs.push_back(make_stream(seed3, nonce, 3, l, cb, cb + kPoly, (long long)kPoly));
            msgs[b].x = d_x.get() + (size_t)b * l;
            msgs[b].rhat = d_r.get() + (size_t)b * kPoly;
            msgs[b].rhatp = d_rp.get() + (size_t)b * kPoly;
            msgs[b].c = cb;
            msgs[b].cl = cb + (size_t)l * kPoly;
            msgs[b].len = l;
            for (int i = 0; i <= l; ++i) items.push_back(EncItem{b, i});
        }
        if (run_gauss(reqs, st)) return -1;
        if (launch_ntt_fwd(d_r.get(), d_r.get(), g * 3, dev->tw_all, st)) return -1;
        if (launch_shoup_prep(d_r.get(), d_rp.get(), (size_t)g * kPoly, st)) return -1;
        HIP_OK(hipMemcpyAsync(d_msgs.get(), msgs.data(), sizeof(EncMsg) * g, hipMemcpyHostToDevice, st));
        HIP_OK(hipMemcpyAsync(d_items.get(), items.data(), sizeof(EncItem) * items.size(), hipMemcpyHostToDevice, st));
        if (launch_enc_poly(d_items.get(), (int)items.size(), d_msgs.get(), d_mpk.get(), l, dev->tw_all, dev->ninv,
                            dev->scale, st))
            return -1;
        HIP_OK(hipMemcpyAsync(c + (size_t)b0 * ct_words, d_c.get(), (size_t)g * ct_words * 4, hipMemcpyDeviceToHost, st));
        if (sync_stream(st)) return -1;
    }
    float e = tm.stop();
    if (ms) *ms += e;
    return 0;
}

int do_keygen(const uint32_t *y, const uint32_t *msk, uint32_t *sk, int repeat, int l, float *ms)
{
    Device *dev = device();
    if (!dev) return -1;
    hipStream_t st = dev->stream;
    EventTimer tm(st);
    HostSparse hsp = to_sparse(y, repeat, l);
    DevSparse dsp;
    DevBuf<uint32_t> d_msk, d_sk;
    if (dsp.upload(hsp, l, st) || d_msk.alloc((size_t)l * kPoly + 1) || d_sk.alloc((size_t)repeat * kPoly)) return -1;
    if (l > 0) HIP_OK(hipMemcpyAsync(d_msk.get(), msk, (size_t)l * kPoly * 4, hipMemcpyHostToDevice, st));
    if (launch_keygen(dsp.view, repeat, d_msk.get(), d_sk.get(), st)) return -1;
    HIP_OK(hipMemcpyAsync(sk, d_sk.get(), (size_t)repeat * kPoly * 4, hipMemcpyDeviceToHost, st));
    float e = tm.stop();
    if (ms) *ms += e;
    return sync_stream(st);
}

// One ciphertext c [l+1][3][N] against `repeat` keys -> d_y [repeat][3][N].
int do_decrypt(const uint32_t *c, const uint32_t *y, const uint32_t *sk, uint32_t *d, int repeat, int l, float *ms)
{
    Device *dev = device();
    if (!dev) return -1;
    hipStream_t st = dev->stream;
    EventTimer tm(st);
    HostSparse hsp = to_sparse(y, repeat, l);
    DevSparse dsp;
    DevBuf<uint32_t> d_c, d_sk, d_d, d_chat, d_prod;
    if (dsp.upload(hsp, l, st) || d_c.alloc((size_t)(l + 1) * kPoly) || d_sk.alloc((size_t)repeat * kPoly) ||
        d_d.alloc((size_t)repeat * kPoly) || d_chat.alloc((size_t)repeat * kPoly) ||
        d_prod.alloc((size_t)repeat * kPoly))
        return -1;
    HIP_OK(hipMemcpyAsync(d_c.get(), c, (size_t)(l + 1) * kPoly * 4, hipMemcpyHostToDevice, st));
    HIP_OK(hipMemcpyAsync(d_sk.get(), sk, (size_t)repeat * kPoly * 4, hipMemcpyHostToDevice, st));
    if (launch_dec_sum(dsp.view, repeat, d_c.get(), d_d.get(), st)) return -1;
    // poly_mul_ntt(c_L, sk) (src/ntt.c:10-26): NTT both, pointwise, INTT
    if (launch_ntt_fwd(d_c.get() + (size_t)l * kPoly, d_chat.get(), 3, dev->tw_all, st)) return -1;
    if (launch_ntt_fwd(d_sk.get(), d_prod.get(), repeat * 3, dev->tw_all, st)) return -1;
    for (int v = 1; v < repeat; ++v)
        HIP_OK(hipMemcpyAsync(d_chat.get() + (size_t)v * kPoly, d_chat.get(), kPoly * 4, hipMemcpyDeviceToDevice, st));
    if (launch_point_mul(d_chat.get(), d_prod.get(), d_prod.get(), (size_t)repeat * kPoly, dev->barrett, st)) return -1;
    if (launch_ntt_inv(d_prod.get(), d_prod.get(), repeat * 3, dev->tw_all, dev->ninv, st)) return -1;
    if (launch_poly_sub(d_d.get(), d_prod.get(), (size_t)repeat * kPoly, st)) return -1;
    HIP_OK(hipMemcpyAsync(d, d_d.get(), (size_t)repeat * kPoly * 4, hipMemcpyDeviceToHost, st));
    float e = tm.stop();
    if (ms) *ms += e;
    return sync_stream(st);
}

}  // namespace

// ===========================================================================
// 1. Reference drop-ins
// ===========================================================================
extern "C" void rlwe_sife_setup_gpu(uint32_t mpk[CUFE_SIFE_L + 1][CUFE_SIFE_NMODULI][CUFE_SIFE_N],
                                    uint32_t msk[CUFE_SIFE_L][CUFE_SIFE_NMODULI][CUFE_SIFE_N], unsigned char *seed2,
                                    unsigned char *seed3, float *part2_time)
{
    if (do_setup(&mpk[0][0][0], &msk[0][0][0], seed2, seed3, IPFE_L, part2_time)) report("rlwe_sife_setup_gpu");
}

extern "C" void rlwe_sife_encrypt_gpu(uint32_t *m, uint32_t mpk[CUFE_SIFE_L + 1][CUFE_SIFE_NMODULI][CUFE_SIFE_N],
                                      uint32_t *c, unsigned char *seed2, unsigned char *seed3, int repeat,
                                      float *part2_time)
{
    if (do_encrypt(m, &mpk[0][0][0], c, seed2, seed3, repeat, IPFE_L, 0, part2_time)) report("rlwe_sife_encrypt_gpu");
}

extern "C" void rlwe_sife_keygen_gpu(const uint32_t *y, const uint32_t msk[CUFE_SIFE_L][CUFE_SIFE_NMODULI][CUFE_SIFE_N],
                                     uint32_t *sk_y, int repeat, float *part2_time)
{
    if (do_keygen(y, &msk[0][0][0], sk_y, repeat, IPFE_L, part2_time)) report("rlwe_sife_keygen_gpu");
}

extern "C" void rlwe_sife_decrypt_gmp_gpu(uint32_t c[CUFE_SIFE_L + 1][CUFE_SIFE_NMODULI][CUFE_SIFE_N], const uint32_t *y,
                                          uint32_t *sk_y, uint32_t *d_y, int repeat, float *part2_time)
{
    if (do_decrypt(&c[0][0][0], y, sk_y, d_y, repeat, IPFE_L, part2_time)) report("rlwe_sife_decrypt_gmp_gpu");
}

// ===========================================================================
// 2. Host helpers
// ===========================================================================
extern "C" int cufe_sample_polya(const unsigned char seed[32], uint32_t *a)
{
    Device *dev = device();
    if (!dev) return -1;
    hipStream_t st = dev->stream;
    uint32_t rk[60];
    aes256_expand_key(seed, rk);
    DevBuf<uint32_t> d_rk, d_a;
    if (d_rk.alloc(60) || d_a.alloc(kPoly)) return -1;
    HIP_OK(hipMemcpyAsync(d_rk.get(), rk, sizeof rk, hipMemcpyHostToDevice, st));
    if (launch_sample_polya(d_rk.get(), d_a.get(), st)) return -1;
    HIP_OK(hipMemcpyAsync(a, d_a.get(), kPoly * 4, hipMemcpyDeviceToHost, st));
    return sync_stream(st);
}

extern "C" double cufe_round_extract2(const uint32_t *d_y)
{
    return (double)crt_decode(d_y[0], d_y[IPFE_N], d_y[2 * IPFE_N]);
}

// svm_predict_gpu (main.c:270-390), plain per-call form.
extern "C" double cufe_svm_predict_gpu(const cufe_svm_model *model,
                                       uint32_t c[CUFE_SIFE_L + 1][CUFE_SIFE_NMODULI][CUFE_SIFE_N],
                                       const uint32_t msk[CUFE_SIFE_L][CUFE_SIFE_NMODULI][CUFE_SIFE_N], double *kvalue)
{
    if (!model || !c || !msk || model->l < 1 || model->nr_class < 2) {
        ipfe_set_error("cufe_svm_predict_gpu: bad arguments");
        return std::nan("");
    }
    const int l = model->l;
    std::vector<uint32_t> sk((size_t)l * kPoly), dy((size_t)l * kPoly);
    // the steps of rlwe_sife_keygen_gpu and rlwe_sife_decrypt_gmp_gpu, with
    // their status checked: a failed device call yields NaN, never a label
    // voted from unwritten key values
    if (do_keygen(model->sv, &msk[0][0][0], sk.data(), l, IPFE_L, nullptr) ||
        do_decrypt(&c[0][0][0], model->sv, sk.data(), dy.data(), l, IPFE_L, nullptr)) {
        report("cufe_svm_predict_gpu");
        if (kvalue)
            for (int i = 0; i < l; ++i) kvalue[i] = std::nan("");
        return std::nan("");
    }
    std::vector<double> kv(l);
    for (int i = 0; i < l; ++i) {
        // round_extract_gmp2 reconstructs every coefficient and keeps only the first
        for (int k = 0; k < IPFE_N; ++k) {
            const uint32_t *dk = dy.data() + (size_t)i * kPoly + k;
            uint32_t r = crt_decode(dk[0], dk[IPFE_N], dk[2 * IPFE_N]);
            if (k == 0) kv[i] = (double)r;
        }
    }
    if (kvalue) memcpy(kvalue, kv.data(), sizeof(double) * l);
    // one-vs-one vote (main.c:338-389)
    const int nr = model->nr_class;
    std::vector<int> start(nr), vote(nr, 0);
    start[0] = 0;
    for (int i = 1; i < nr; ++i) start[i] = start[i - 1] + model->nSV[i - 1];
    int p = 0;
    for (int i = 0; i < nr; ++i)
        for (int j = i + 1; j < nr; ++j) {
            double sum = 0;
            const double *coef1 = model->sv_coef + (size_t)(j - 1) * l;
            const double *coef2 = model->sv_coef + (size_t)i * l;
            for (int k = 0; k < model->nSV[i]; ++k) sum += coef1[start[i] + k] * kv[start[i] + k];
            for (int k = 0; k < model->nSV[j]; ++k) sum += coef2[start[j] + k] * kv[start[j] + k];
            sum -= model->rho[p];
            if (sum > 0) ++vote[i]; else ++vote[j];
            ++p;
        }
    int best = 0;
    for (int i = 1; i < nr; ++i)
        if (vote[i] > vote[best]) best = i;
    return model->label[best];
}

// ===========================================================================
// 3. Status-returning single operations
// ===========================================================================
extern "C" int cufe_device_info(char *name, int name_len, int *cu_count, uint64_t *hbm_bytes)
{
    hipDeviceProp_t prop;
    int d = 0;
    HIP_OK(hipGetDevice(&d));
    HIP_OK(hipGetDeviceProperties(&prop, d));
    if (name && name_len > 0) snprintf(name, (size_t)name_len, "%s (%s)", prop.name, prop.gcnArchName);
    if (cu_count) *cu_count = prop.multiProcessorCount;
    if (hbm_bytes) *hbm_bytes = prop.totalGlobalMem;
    return 0;
}

extern "C" int cufe_ntt_forward(uint32_t *polys, int npoly)
{
    Device *dev = device();
    if (!dev) return -1;
    hipStream_t st = dev->stream;
    DevBuf<uint32_t> d;
    if (d.alloc((size_t)npoly * kPoly)) return -1;
    HIP_OK(hipMemcpyAsync(d.get(), polys, (size_t)npoly * kPoly * 4, hipMemcpyHostToDevice, st));
    const PsetState *ps = pset_state(1);  // the reference set, tuned transform shape
    if (!ps || ps_ntt(*ps, false, d.get(), d.get(), npoly * 3, st)) return -1;
    HIP_OK(hipMemcpyAsync(polys, d.get(), (size_t)npoly * kPoly * 4, hipMemcpyDeviceToHost, st));
    return sync_stream(st);
}

extern "C" int cufe_ntt_inverse(uint32_t *polys, int npoly)
{
    Device *dev = device();
    if (!dev) return -1;
    hipStream_t st = dev->stream;
    DevBuf<uint32_t> d;
    if (d.alloc((size_t)npoly * kPoly)) return -1;
    HIP_OK(hipMemcpyAsync(d.get(), polys, (size_t)npoly * kPoly * 4, hipMemcpyHostToDevice, st));
    const PsetState *ps = pset_state(1);
    if (!ps || ps_ntt(*ps, true, d.get(), d.get(), npoly * 3, st)) return -1;
    HIP_OK(hipMemcpyAsync(polys, d.get(), (size_t)npoly * kPoly * 4, hipMemcpyDeviceToHost, st));
    return sync_stream(st);
}

extern "C" int cufe_gauss_sample(const unsigned char *keys, const uint16_t *nonces, int nstreams, int which, int ncalls,
                                 uint32_t *out)
{
    Device *dev = device();
    if (!dev) return -1;
    if (which < 1 || which > 3 || ncalls < 0 || nstreams < 0) {
        ipfe_set_error("cufe_gauss_sample: bad arguments (which=%d ncalls=%d nstreams=%d)", which, ncalls, nstreams);
        return -1;
    }
    if (ncalls == 0 || nstreams == 0) return 0;
    hipStream_t st = dev->stream;
    DevBuf<uint32_t> d;
    const size_t per_stream = (size_t)ncalls * kPoly;
    if (d.alloc((size_t)nstreams * per_stream)) return -1;
    std::vector<StreamReq> reqs;
    for (int t = 0; t < nstreams; ++t) {
        uint32_t *o = d.get() + (size_t)t * per_stream;
        reqs.push_back(make_stream(keys + 32 * (size_t)t, nonces[t], which, ncalls, o, o + kPoly, (long long)kPoly));
    }
    if (run_gauss(reqs, st)) return -1;
    HIP_OK(hipMemcpyAsync(out, d.get(), (size_t)nstreams * per_stream * 4, hipMemcpyDeviceToHost, st));
    return sync_stream(st);
}

extern "C" int cufe_gauss_sample_first(const unsigned char *keys, const uint16_t *nonces, int nstreams, int which,
                                       int ncalls, uint32_t *out)
{
    Device *dev = device();
    if (!dev) return -1;
    if (which < 1 || which > 3 || ncalls < 0 || nstreams < 0) {
        ipfe_set_error("cufe_gauss_sample_first: bad arguments (which=%d ncalls=%d nstreams=%d)", which, ncalls,
                       nstreams);
        return -1;
    }
    if (ncalls == 0 || nstreams == 0) return 0;
    hipStream_t st = dev->stream;
    DevBuf<uint32_t> d;
    const size_t per_stream = (size_t)ncalls * 3;
    if (d.alloc((size_t)nstreams * per_stream)) return -1;
    std::vector<StreamReq> reqs;
    for (int t = 0; t < nstreams; ++t)
        reqs.push_back(make_stream(keys + 32 * (size_t)t, nonces[t], which, ncalls, d.get() + (size_t)t * per_stream,
                                   nullptr, 0, 1));
    if (run_gauss(reqs, st)) return -1;
    HIP_OK(hipMemcpyAsync(out, d.get(), (size_t)nstreams * per_stream * 4, hipMemcpyDeviceToHost, st));
    return sync_stream(st);
}

extern "C" int cufe_aes256ctr(const unsigned char key[32], uint16_t nonce, uint32_t ctr0, uint32_t nblocks,
                              unsigned char *out)
{
    Device *dev = device();
    if (!dev) return -1;
    hipStream_t st = dev->stream;
    uint32_t rk[60];
    aes256_expand_key(key, rk);
    DevBuf<uint32_t> d_rk;
    DevBuf<uint8_t> d_out;
    if (d_rk.alloc(60) || d_out.alloc((size_t)nblocks * 16 + 16)) return -1;
    HIP_OK(hipMemcpyAsync(d_rk.get(), rk, sizeof rk, hipMemcpyHostToDevice, st));
    if (launch_aes_ctr(d_rk.get(), nonce, ctr0, nblocks, d_out.get(), st)) return -1;
    HIP_OK(hipMemcpyAsync(out, d_out.get(), (size_t)nblocks * 16, hipMemcpyDeviceToHost, st));
    return sync_stream(st);
}

extern "C" int cufe_crt_round(const uint32_t *d, int nvec, int coef, uint32_t *r)
{
    Device *dev = device();
    if (!dev) return -1;
    if (coef < 0 || coef >= IPFE_N) {
        ipfe_set_error("cufe_crt_round: coefficient %d out of range", coef);
        return -1;
    }
    hipStream_t st = dev->stream;
    DevBuf<uint32_t> d_d, d_r;
    if (d_d.alloc((size_t)nvec * kPoly) || d_r.alloc((size_t)nvec + 1)) return -1;
    HIP_OK(hipMemcpyAsync(d_d.get(), d, (size_t)nvec * kPoly * 4, hipMemcpyHostToDevice, st));
    if (launch_crt_round(d_d.get(), nvec, coef, d_r.get(), st)) return -1;
    HIP_OK(hipMemcpyAsync(r, d_r.get(), (size_t)nvec * 4, hipMemcpyDeviceToHost, st));
    return sync_stream(st);
}

extern "C" int cufe_time_ntt(int npoly, int inverse, int iters, float *ms_per_launch)
{
    Device *dev = device();
    if (!dev) return -1;
    hipStream_t st = dev->stream;
    DevBuf<uint32_t> d;
    if (d.alloc((size_t)npoly * kPoly)) return -1;
    // synthetic residues: any fixed pattern in [0, q0)
    HIP_OK(hipMemsetD32Async((hipDeviceptr_t)d.get(), 65537u, (size_t)npoly * kPoly, st));
    // the reference set's tuned transform shape (tuned on first use)
    const PsetState *ps = pset_state(1);
    if (!ps) return -1;
    for (int w = 0; w < 2; ++w)
        if (ps_ntt(*ps, inverse != 0, d.get(), d.get(), npoly * 3, st)) return -1;
    EventTimer tm(st);
    for (int it = 0; it < iters; ++it)
        if (ps_ntt(*ps, inverse != 0, d.get(), d.get(), npoly * 3, st)) return -1;
    float e = tm.stop();
    if (ms_per_launch) *ms_per_launch = e / (iters > 0 ? iters : 1);
    return sync_stream(st);
}

// nstreams streams of ncalls calls each (sampler which), timed per run.
extern "C" int cufe_time_gauss_calls(int nstreams, int ncalls, int which, int iters, float *ms_per_run)
{
    if (nstreams < 1 || nstreams > 65536 || ncalls < 1 || which < 1 || which > 3) {
        ipfe_set_error("cufe_time_gauss_calls: %d streams x %d calls of S%d", nstreams, ncalls, which);
        return -1;
    }
    Device *dev = device();
    if (!dev) return -1;
    hipStream_t st = dev->stream;
    DevBuf<uint32_t> d;
    if (d.alloc((size_t)nstreams * ncalls * kPoly)) return -1;
    std::vector<StreamReq> reqs;
    for (int t = 0; t < nstreams; ++t) {
        uint8_t key[32];
        for (int i = 0; i < 32; ++i) key[i] = (uint8_t)(i * 7 + 3);
        uint32_t *o = d.get() + (size_t)t * ncalls * kPoly;
        reqs.push_back(make_stream(key, (uint16_t)t, which, ncalls, o, o + kPoly, (long long)kPoly));
    }
    if (run_gauss(reqs, st)) return -1;
    EventTimer tm(st);
    for (int it = 0; it < iters; ++it)
        if (run_gauss(reqs, st)) return -1;
    float e = tm.stop();
    if (ms_per_run) *ms_per_run = e / (iters > 0 ? iters : 1);
    return 0;
}

extern "C" int cufe_time_gauss(int nstreams, int which, int iters, float *ms_per_launch)
{
    Device *dev = device();
    if (!dev) return -1;
    hipStream_t st = dev->stream;
    DevBuf<uint32_t> d;
    if (d.alloc((size_t)nstreams * kPoly)) return -1;
    std::vector<StreamReq> reqs;
    for (int t = 0; t < nstreams; ++t) {
        uint8_t key[32];
        for (int i = 0; i < 32; ++i) key[i] = (uint8_t)(i * 7 + 3);
        reqs.push_back(make_stream(key, (uint16_t)t, which, 1, d.get() + (size_t)t * kPoly, nullptr, 0));
    }
    if (run_gauss(reqs, st)) return -1;
    EventTimer tm(st);
    for (int it = 0; it < iters; ++it)
        if (run_gauss(reqs, st)) return -1;
    float e = tm.stop();
    if (ms_per_launch) *ms_per_launch = e / (iters > 0 ? iters : 1);
    return 0;
}

// ===========================================================================
// Variable-length scheme entry points
// ===========================================================================
static int check_l(int l)
{
    if (l < 0 || l > IPFE_L) {
        ipfe_set_error("vector length l=%d outside [0, %d]", l, IPFE_L);
        return -1;
    }
    return 0;
}

extern "C" int cufe_sife_setup(uint32_t *mpk, uint32_t *msk, const unsigned char seed2[32],
                               const unsigned char seed3[32], int l)
{
    if (check_l(l)) return -1;
    return do_setup(mpk, msk, seed2, seed3, l, nullptr);
}

extern "C" int cufe_sife_encrypt(const uint32_t *m, const uint32_t *mpk, uint32_t *c, const unsigned char seed2[32],
                                 const unsigned char seed3[32], int repeat, int l, int nonce0)
{
    if (check_l(l)) return -1;
    return do_encrypt(m, mpk, c, seed2, seed3, repeat, l, nonce0, nullptr);
}

extern "C" int cufe_sife_keygen(const uint32_t *y, const uint32_t *msk, uint32_t *sk, int repeat, int l)
{
    if (check_l(l)) return -1;
    return do_keygen(y, msk, sk, repeat, l, nullptr);
}

extern "C" int cufe_sife_decrypt(const uint32_t *c, const uint32_t *y, const uint32_t *sk, uint32_t *d, int repeat,
                                 int l)
{
    if (check_l(l)) return -1;
    return do_decrypt(c, y, sk, d, repeat, l, nullptr);
}
