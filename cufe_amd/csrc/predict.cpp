// predict.cpp -- resident-key context and batched IPFE-SVM prediction (C ABI
// section 4 of include/cufe_ipfe.h).
//
// A context keeps in HBM: the master keys (mpk in the NTT domain, msk), the
// functional keys of one SVM model (one per support vector, generated once on
// the device instead of once per row as svm_predict_gpu does, main.c:308),
// and a store of encrypted rows in compact form: c_L (balanced int32, the
// operand of the coefficient-0 contraction) and coefficient 0 of c_0..c_{l-1}.
// That is all decryption of coefficient 0 reads; the other 4095 coefficients of
// each c_i are never materialised.
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <algorithm>
#include <vector>

#include "../../include/cufe_ipfe.h"
#include "kernels_predict.h"
#include "runtime.h"

using namespace ipfe;

namespace {
constexpr size_t kPoly = (size_t)3 * IPFE_N;
}

struct cufe_ctx {
    int l_key = 0;                // slots of the master key
    DevBuf<uint32_t> mpk, msk;    // [l_key+1][3][N], [l_key][3][N]
    // model
    int S = 0, nr_class = 0, sv_maxidx = -1;
    DevBuf<int> y_off, y_idx;
    DevBuf<uint32_t> y_val;
    SparseY y{};
    DevBuf<uint32_t> sk;          // [S][3][N]
    DevBuf<double> kd;            // [3][S][kkey]
    int kkey = 0;                 // key columns: N (+ the y_i slots when fused)
    bool fused = false;           // sum part folded into the contraction
    int kb[3] = {0, 0, 0};
    unsigned int maxabs[3] = {0, 0, 0};
    DevBuf<double> coef, rho;
    SvmModelDev mdev{};
    // rows
    int row_l = 0, cap = 0, nrows = 0;
    int ldr = 0;                  // row-store columns: N + row_l rounded up to 16
    DevBuf<int32_t> A;            // [3][cap][ldr]: c_L, then c_i[0] for i < row_l
    DevBuf<double> A64;           // the same rows blocked for k_contract_sgpr: [3][(cap+64)/16][ldr][16]
                                  // (64 pad rows per limb: a tile reads whole 64-row blocks)
    int mode = 1;                 // contraction kernel: 1 = scalar-operand (sgpr), 0 = LDS tiles
    DevBuf<uint32_t> c0;          // [cap][row_l][3]
    // predict scratch
    int chunk_rows = 0;
    DevBuf<uint32_t> D, kval;
    DevBuf<double> labels;
    DevBuf<int> flag;
    hipEvent_t ev[8] = {};
    float last_ms[4] = {0, 0, 0, 0};  // contract, kvalue, decide, total
};

static int ensure_events(cufe_ctx *c)
{
    for (auto &e : c->ev)
        if (!e) HIP_OK(hipEventCreate(&e));
    return 0;
}

extern "C" int cufe_ctx_create(cufe_ctx **out)
{
    if (!out) {
        ipfe_set_error("cufe_ctx_create: null output pointer");
        return -1;
    }
    if (!device()) return -1;
    cufe_ctx *c = new cufe_ctx();
    if (ensure_events(c)) {
        delete c;
        return -1;
    }
    *out = c;
    return 0;
}

extern "C" void cufe_ctx_destroy(cufe_ctx *c)
{
    if (!c) return;
    for (auto &e : c->ev)
        if (e) (void)hipEventDestroy(e);
    delete c;
}

extern "C" int cufe_ctx_load_keys(cufe_ctx *c, const uint32_t *mpk, const uint32_t *msk, int l)
{
    if (!c || l < 1 || l > IPFE_L) {
        ipfe_set_error("cufe_ctx_load_keys: bad context or length %d", l);
        return -1;
    }
    Device *dev = device();
    if (!dev) return -1;
    if (c->mpk.alloc((size_t)(l + 1) * kPoly) || c->msk.alloc((size_t)l * kPoly)) return -1;
    HIP_OK(hipMemcpy(c->mpk.get(), mpk, (size_t)(l + 1) * kPoly * 4, hipMemcpyHostToDevice));
    HIP_OK(hipMemcpy(c->msk.get(), msk, (size_t)l * kPoly * 4, hipMemcpyHostToDevice));
    c->l_key = l;
    c->S = 0;
    return 0;
}

// Generates resident keys on the device: a = sample_polya(seedA), then setup.
extern "C" int cufe_ctx_setup(cufe_ctx *c, const unsigned char seedA[32], const unsigned char seed2[32],
                              const unsigned char seed3[32], int l)
{
    if (!c || l < 1 || l > IPFE_L) {
        ipfe_set_error("cufe_ctx_setup: bad context or length %d", l);
        return -1;
    }
    Device *dev = device();
    if (!dev) return -1;
    hipStream_t st = dev->stream;
    if (c->mpk.alloc((size_t)(l + 1) * kPoly) || c->msk.alloc((size_t)l * kPoly)) return -1;
    DevBuf<uint32_t> d_rk, d_e, d_ap;
    if (d_rk.alloc(60) || d_e.alloc((size_t)l * kPoly) || d_ap.alloc(kPoly)) return -1;
    uint32_t rk[60];
    aes256_expand_key(seedA, rk);
    HIP_OK(hipMemcpyAsync(d_rk.get(), rk, sizeof rk, hipMemcpyHostToDevice, st));
    uint32_t *ahat = c->mpk.get() + (size_t)l * kPoly;
    if (launch_sample_polya(d_rk.get(), ahat, st)) return -1;
    if (launch_ntt_fwd(ahat, ahat, 3, dev->tw_all, st)) return -1;
    if (launch_shoup_prep(ahat, d_ap.get(), kPoly, st)) return -1;
    std::vector<StreamReq> reqs;
    reqs.push_back(make_stream(seed2, 0, 1, l, c->msk.get(), c->msk.get() + kPoly, (long long)kPoly));
    reqs.push_back(make_stream(seed3, 0, 1, l, d_e.get(), d_e.get() + kPoly, (long long)kPoly));
    if (run_gauss(reqs, st)) return -1;
    if (launch_setup_pk(c->msk.get(), d_e.get(), ahat, d_ap.get(), c->mpk.get(), l, dev->tw_all, st)) return -1;
    if (sync_stream(st)) return -1;
    c->l_key = l;
    c->S = 0;
    return 0;
}

extern "C" int cufe_ctx_get_keys(cufe_ctx *c, uint32_t *mpk, uint32_t *msk)
{
    if (!c || c->l_key < 1) {
        ipfe_set_error("cufe_ctx_get_keys: no keys in context");
        return -1;
    }
    if (mpk) HIP_OK(hipMemcpy(mpk, c->mpk.get(), (size_t)(c->l_key + 1) * kPoly * 4, hipMemcpyDeviceToHost));
    if (msk) HIP_OK(hipMemcpy(msk, c->msk.get(), (size_t)c->l_key * kPoly * 4, hipMemcpyDeviceToHost));
    return 0;
}

// Functional keys of every support vector, generated on the device and kept.
// The fold runs over the key length l_key (785 for the reference layout), so
// it ends on zero slots exactly as the reference's dense loop does.
extern "C" int cufe_ctx_set_model(cufe_ctx *c, const cufe_svm_model *m)
{
    if (!c || !m || c->l_key < 1) {
        ipfe_set_error("cufe_ctx_set_model: context has no keys");
        return -1;
    }
    if (m->nr_class < 2 || m->nr_class > SVM_MAX_CLASS || m->l < 1) {
        ipfe_set_error("cufe_ctx_set_model: nr_class %d (2..%d) or l %d unsupported", m->nr_class, SVM_MAX_CLASS, m->l);
        return -1;
    }
    Device *dev = device();
    if (!dev) return -1;
    hipStream_t st = dev->stream;
    const int S = m->l, L = c->l_key;
    std::vector<int> off(S + 1), idx;
    std::vector<uint32_t> val;
    int maxidx = -1;
    for (int v = 0; v < S; ++v) {
        off[v] = (int)idx.size();
        for (int i = 0; i < IPFE_L; ++i) {
            const uint32_t yv = m->sv[(size_t)v * IPFE_L + i];
            if (!yv) continue;
            if (i >= L) {
                ipfe_set_error("cufe_ctx_set_model: SV %d has feature %d beyond the key length %d", v, i, L);
                return -1;
            }
            idx.push_back(i);
            val.push_back(yv);
            maxidx = std::max(maxidx, i);
        }
    }
    off[S] = (int)idx.size();
    if (idx.empty()) {
        idx.push_back(0);
        val.push_back(0);
    }
    int nsv_sum = 0;
    for (int i = 0; i < m->nr_class; ++i) nsv_sum += m->nSV[i];
    if (nsv_sum != S) {
        ipfe_set_error("cufe_ctx_set_model: sum of nSV (%d) != l (%d)", nsv_sum, S);
        return -1;
    }
    // With a zero slot after the last feature the reference's fold ends
    // canonical, so the sum part joins the contraction (key = [-s', y]).
    const bool fused = maxidx < L - 1;
    const int kkey = fused ? IPFE_N + ((maxidx + 1 + 15) / 16) * 16 : IPFE_N;
    if (c->y_off.alloc(off.size()) || c->y_idx.alloc(idx.size()) || c->y_val.alloc(val.size()) ||
        c->sk.alloc((size_t)S * kPoly) || c->kd.alloc((size_t)3 * S * kkey))
        return -1;
    HIP_OK(hipMemcpyAsync(c->y_off.get(), off.data(), off.size() * 4, hipMemcpyHostToDevice, st));
    HIP_OK(hipMemcpyAsync(c->y_idx.get(), idx.data(), idx.size() * 4, hipMemcpyHostToDevice, st));
    HIP_OK(hipMemcpyAsync(c->y_val.get(), val.data(), val.size() * 4, hipMemcpyHostToDevice, st));
    c->y.off = c->y_off.get();
    c->y.idx = c->y_idx.get();
    c->y.val = c->y_val.get();
    c->y.len = L;
    if (launch_keygen(c->y, S, c->msk.get(), c->sk.get(), st)) return -1;
    DevBuf<unsigned int> d_max;
    if (d_max.alloc(3)) return -1;
    HIP_OK(hipMemsetAsync(d_max.get(), 0, 12, st));
    if (launch_contract_keys(c->sk.get(), S, kkey, fused ? -1 : 1, c->y, c->kd.get(), d_max.get(), st)) return -1;
    HIP_OK(hipMemcpyAsync(c->maxabs, d_max.get(), 12, hipMemcpyDeviceToHost, st));
    // model coefficients
    const int npairs = m->nr_class * (m->nr_class - 1) / 2;
    if (c->coef.alloc((size_t)(m->nr_class - 1) * S) || c->rho.alloc(npairs)) return -1;
    HIP_OK(hipMemcpyAsync(c->coef.get(), m->sv_coef, sizeof(double) * (m->nr_class - 1) * S, hipMemcpyHostToDevice, st));
    HIP_OK(hipMemcpyAsync(c->rho.get(), m->rho, sizeof(double) * npairs, hipMemcpyHostToDevice, st));
    if (sync_stream(st)) return -1;
    SvmModelDev md{};
    md.nr_class = m->nr_class;
    int start = 0;
    for (int i = 0; i < m->nr_class; ++i) {
        md.nSV[i] = m->nSV[i];
        md.start[i] = start;
        md.label[i] = m->label[i];
        start += m->nSV[i];
    }
    md.sv_coef = c->coef.get();
    md.rho = c->rho.get();
    c->mdev = md;
    // reduction interval of the fp64 contraction: every partial sum below 2^53
    for (int s = 0; s < 3; ++s) {
        const double q = (double)qmod(s);
        const double bmax = (double)c->maxabs[s];
        const double room = 9007199254740991.0 - 1.5 * q - 1.0;
        double kb = bmax > 0 ? floor(room / (16.0 * floor(q / 2) * bmax)) : 256.0;
        c->kb[s] = (int)std::min(256.0, kb);
    }
    c->S = S;
    c->nr_class = m->nr_class;
    c->sv_maxidx = maxidx;
    c->kkey = kkey;
    c->fused = fused;
    return 0;
}

// Reserves the row store: up to cap rows of l slots (l must cover every SV feature).
extern "C" int cufe_ctx_rows_reserve(cufe_ctx *c, int cap, int l)
{
    if (!c || cap < 1 || l < 1 || l > c->l_key) {
        ipfe_set_error("cufe_ctx_rows_reserve: cap %d / l %d invalid for key length %d", cap, l, c ? c->l_key : 0);
        return -1;
    }
    const int ldr = IPFE_N + ((l + 15) / 16) * 16;
    cap = ((cap + 63) / 64) * 64;  // whole 64-row tiles of the blocked store
    if (c->A.alloc((size_t)3 * cap * ldr) || c->c0.alloc((size_t)cap * l * 3) ||
        c->A64.alloc((size_t)3 * (cap + 64) * ldr))
        return -1;
    HIP_OK(hipMemset(c->A.get(), 0, (size_t)3 * cap * ldr * 4));
    HIP_OK(hipMemset(c->A64.get(), 0, (size_t)3 * (cap + 64) * ldr * 8));
    const char *mode = getenv("CUFE_CONTRACT");
    c->mode = (mode && strcmp(mode, "lds") == 0) ? 0 : 1;
    c->cap = cap;
    c->row_l = l;
    c->ldr = ldr;
    c->nrows = 0;
    return 0;
}

// Encrypts n rows x [n][row_l] into the row store (compact form).  Row b uses
// nonce nonce0 + b, exactly as rlwe_sife_encrypt_gpu numbers its batch.
extern "C" int cufe_ctx_encrypt_rows(cufe_ctx *c, const uint32_t *x, int n, const unsigned char seed2[32],
                                     const unsigned char seed3[32], int nonce0, float *ms)
{
    if (!c || c->cap < 1 || c->l_key < 1) {
        ipfe_set_error("cufe_ctx_encrypt_rows: reserve rows and load keys first");
        return -1;
    }
    if (n < 0 || c->nrows + n > c->cap) {
        ipfe_set_error("cufe_ctx_encrypt_rows: %d rows exceed the reserved capacity %d", c->nrows + n, c->cap);
        return -1;
    }
    if (nonce0 < 0 || nonce0 + n > 65536) {
        ipfe_set_error("cufe_ctx_encrypt_rows: nonce range [%d, %d) exceeds 16 bits", nonce0, nonce0 + n);
        return -1;
    }
    Device *dev = device();
    if (!dev) return -1;
    hipStream_t st = dev->stream;
    const int l = c->row_l;
    HIP_OK(hipEventRecord(c->ev[6], st));
    // rows per group: the S2 stream (r, f_0) keeps per-chunk values, the compact
    // S3 stream only counts
    const size_t per_row = gauss_bytes(2, 2, 0) + gauss_bytes(3, l, 1) + 4 * kPoly * 4 + (size_t)l * 16;
    const int group = (int)std::max<size_t>(1, std::min<size_t>(16384, ((size_t)12 << 30) / per_row));
    DevBuf<uint32_t> d_x, d_r, d_rp, d_cl, d_f0;
    DevBuf<EncMsg> d_msgs;
    DevBuf<EncItem> d_items;
    const int gmax = std::min(group, std::max(n, 1));
    if (d_x.alloc((size_t)gmax * l) || d_r.alloc((size_t)gmax * kPoly) || d_rp.alloc((size_t)gmax * kPoly) ||
        d_cl.alloc((size_t)gmax * kPoly) || d_f0.alloc((size_t)gmax * l * 3) || d_msgs.alloc(gmax) ||
        d_items.alloc(gmax))
        return -1;
    for (int b0 = 0; b0 < n; b0 += group) {
        const int g = std::min(group, n - b0);
        HIP_OK(hipMemcpyAsync(d_x.get(), x + (size_t)b0 * l, (size_t)g * l * 4, hipMemcpyHostToDevice, st));
        std::vector<StreamReq> reqs;
        std::vector<EncMsg> msgs(g);
        std::vector<EncItem> items(g);
        for (int b = 0; b < g; ++b) {
            const uint16_t nonce = (uint16_t)(nonce0 + b0 + b);
            reqs.push_back(make_stream(seed2, nonce, 2, 2, d_r.get() + (size_t)b * kPoly, d_cl.get() + (size_t)b * kPoly,
                                       0));
            reqs.push_back(make_stream(seed3, nonce, 3, l, d_f0.get() + (size_t)b * l * 3, nullptr, 0, 1));
            msgs[b].x = d_x.get() + (size_t)b * l;
            msgs[b].rhat = d_r.get() + (size_t)b * kPoly;
            msgs[b].rhatp = d_rp.get() + (size_t)b * kPoly;
            msgs[b].c = nullptr;
            msgs[b].cl = d_cl.get() + (size_t)b * kPoly;
            msgs[b].len = l;
            items[b] = EncItem{b, l};
        }
        if (run_gauss(reqs, st)) return -1;
        if (launch_ntt_fwd(d_r.get(), d_r.get(), g * 3, dev->tw_all, st)) return -1;
        if (launch_shoup_prep(d_r.get(), d_rp.get(), (size_t)g * kPoly, st)) return -1;
        HIP_OK(hipMemcpyAsync(d_msgs.get(), msgs.data(), sizeof(EncMsg) * g, hipMemcpyHostToDevice, st));
        HIP_OK(hipMemcpyAsync(d_items.get(), items.data(), sizeof(EncItem) * g, hipMemcpyHostToDevice, st));
        // c_L = INTT(a_hat . r_hat) + f_0 against the key's a_hat at index l_key
        if (launch_enc_poly(d_items.get(), g, d_msgs.get(), c->mpk.get(), c->l_key, dev->tw_all, dev->ninv, dev->scale,
                            st))
            return -1;
        if (launch_enc_coef0(d_r.get(), c->mpk.get(), d_x.get(), g, l, d_f0.get(), dev->ninv, dev->scale, dev->barrett,
                             st))
            return -1;
        if (launch_rows_balanced(d_cl.get(), kPoly, g, c->A.get(), c->cap, c->nrows + b0, c->ldr, st)) return -1;
        if (launch_rows_c0(d_f0.get(), g, l, c->A.get(), c->cap, c->nrows + b0, c->ldr, st)) return -1;
        HIP_OK(hipMemcpyAsync(c->c0.get() + (size_t)(c->nrows + b0) * l * 3, d_f0.get(), (size_t)g * l * 3 * 4,
                              hipMemcpyDeviceToDevice, st));
        if (launch_rows_blocked(c->A.get(), c->cap, c->cap + 64, c->ldr, c->nrows + b0, g, c->A64.get(), st)) return -1;
        if (sync_stream(st)) return -1;
    }
    HIP_OK(hipEventRecord(c->ev[7], st));
    HIP_OK(hipEventSynchronize(c->ev[7]));
    float e = 0;
    HIP_OK(hipEventElapsedTime(&e, c->ev[6], c->ev[7]));
    if (ms) *ms = e;
    c->nrows += n;
    return 0;
}

extern "C" int cufe_ctx_rows_count(const cufe_ctx *c) { return c ? c->nrows : -1; }

// Batched prediction of rows [row0, row0 + nrows): labels [nrows]; optional
// kvalues [nrows][S] (the decrypted inner products, as round_extract returns
// them).  ms (optional, 4 floats): contraction, kvalue, decision and total
// device milliseconds of this call.
extern "C" int cufe_ctx_predict(cufe_ctx *c, int row0, int nrows, double *labels, uint32_t *kvalues, float *ms)
{
    if (!c || c->S < 1) {
        ipfe_set_error("cufe_ctx_predict: no model in context");
        return -1;
    }
    if (row0 < 0 || nrows < 0 || row0 + nrows > c->nrows) {
        ipfe_set_error("cufe_ctx_predict: rows [%d, %d) outside the %d stored", row0, row0 + nrows, c->nrows);
        return -1;
    }
    if (c->sv_maxidx >= c->row_l) {
        ipfe_set_error("cufe_ctx_predict: rows hold %d slots but the model uses feature %d", c->row_l, c->sv_maxidx);
        return -1;
    }
    Device *dev = device();
    if (!dev) return -1;
    hipStream_t st = dev->stream;
    const int S = c->S, l = c->row_l;
    // rows per chunk: D (3 words) + kvalue (1 word) per (row, key) within ~4 GiB
    // (a multiple of 64 rows, so every chunk starts on a blocked-store tile)
    const int chunk = (int)std::max<size_t>(
        64, std::min<size_t>(((size_t)nrows + 63) / 64 * 64, ((size_t)1 << 30) / ((size_t)S) / 64 * 64));
    if (c->D.n < (size_t)3 * chunk * S || c->labels.n < (size_t)chunk) {
        if (c->D.alloc((size_t)3 * chunk * S) || c->kval.alloc((size_t)chunk * S) || c->labels.alloc(chunk) ||
            c->flag.alloc(1))
            return -1;
        c->chunk_rows = chunk;
    }
    float t_contract = 0, t_kv = 0, t_dec = 0;
    HIP_OK(hipEventRecord(c->ev[4], st));
    if (launch_fill_i32(c->flag.get(), 0, 1, st)) return -1;
    for (int r0 = row0; r0 < row0 + nrows; r0 += chunk) {
        const int R = std::min(chunk, row0 + nrows - r0);
        HIP_OK(hipEventRecord(c->ev[0], st));
        for (int s = 0; s < 3; ++s) {
            const int32_t *As = c->A.get() + ((size_t)s * c->cap + r0) * c->ldr;
            const double *Bs = c->kd.get() + (size_t)s * S * c->kkey;
            uint32_t *Ds = c->D.get() + (size_t)s * R * S;
            int rc;
            if (c->kb[s] < 1)
                rc = launch_contract_slow(As, Bs, R, S, c->kkey, c->ldr, c->kkey, s, dev->barrett, Ds, st);
            else if (c->mode == 1 && r0 % 16 == 0)
                // kb counts 16-product steps; this kernel's steps are 8 products
                rc = launch_contract_sgpr(c->A64.get() + ((size_t)s * (c->cap + 64) + r0) * c->ldr, Bs, R, S, c->kkey,
                                          c->ldr, c->kkey, s, 2 * c->kb[s], Ds, st);
            else
                rc = launch_contract(As, Bs, R, S, c->kkey, c->ldr, c->kkey, s, c->kb[s], Ds, st);
            if (rc) return -1;
        }
        HIP_OK(hipEventRecord(c->ev[1], st));
        if (c->fused) {
            if (launch_crt_kvalue(c->D.get(), (size_t)R * S, c->kval.get(), st)) return -1;
        } else if (launch_kvalue(c->D.get(), c->c0.get() + (size_t)r0 * l * 3, l, c->y, R, S, c->kval.get(),
                                 c->flag.get(), st)) {
            return -1;
        }
        HIP_OK(hipEventRecord(c->ev[2], st));
        if (launch_decide(c->kval.get(), R, S, c->mdev, c->labels.get(), st)) return -1;
        HIP_OK(hipEventRecord(c->ev[3], st));
        if (labels) HIP_OK(hipMemcpyAsync(labels + (r0 - row0), c->labels.get(), sizeof(double) * R,
                                          hipMemcpyDeviceToHost, st));
        if (kvalues)
            HIP_OK(hipMemcpyAsync(kvalues + (size_t)(r0 - row0) * S, c->kval.get(), sizeof(uint32_t) * R * S,
                                  hipMemcpyDeviceToHost, st));
        HIP_OK(hipEventSynchronize(c->ev[3]));
        float a = 0, b = 0, d = 0;
        HIP_OK(hipEventElapsedTime(&a, c->ev[0], c->ev[1]));
        HIP_OK(hipEventElapsedTime(&b, c->ev[1], c->ev[2]));
        HIP_OK(hipEventElapsedTime(&d, c->ev[2], c->ev[3]));
        t_contract += a;
        t_kv += b;
        t_dec += d;
    }
    int flag = 0;
    HIP_OK(hipMemcpyAsync(&flag, c->flag.get(), 4, hipMemcpyDeviceToHost, st));
    HIP_OK(hipEventRecord(c->ev[5], st));
    HIP_OK(hipEventSynchronize(c->ev[5]));
    float tot = 0;
    HIP_OK(hipEventElapsedTime(&tot, c->ev[4], c->ev[5]));
    c->last_ms[0] = t_contract;
    c->last_ms[1] = t_kv;
    c->last_ms[2] = t_dec;
    c->last_ms[3] = tot;
    if (ms) memcpy(ms, c->last_ms, sizeof c->last_ms);
    if (flag) {
        ipfe_set_error("cufe_ctx_predict: a decryption sum ended on the representative q with a zero product; "
                       "this needs the full-polynomial path (keys must be shorter than the slot count)");
        return -1;
    }
    return 0;
}

// Diagnostics of the loaded model: fp64 reduction interval and largest
// balanced key coefficient per limb.
extern "C" int cufe_ctx_model_info(const cufe_ctx *c, int *kb, unsigned int *maxabs)
{
    if (!c || c->S < 1) {
        ipfe_set_error("cufe_ctx_model_info: no model");
        return -1;
    }
    for (int s = 0; s < 3; ++s) {
        if (kb) kb[s] = c->kb[s];
        if (maxabs) maxabs[s] = c->maxabs[s];
    }
    return 0;
}

// The resident functional keys [S][3][N] (tests).
extern "C" int cufe_ctx_get_model_keys(const cufe_ctx *c, uint32_t *sk)
{
    if (!c || c->S < 1) {
        ipfe_set_error("cufe_ctx_get_model_keys: no model");
        return -1;
    }
    HIP_OK(hipMemcpy(sk, c->sk.get(), (size_t)c->S * kPoly * 4, hipMemcpyDeviceToHost));
    return 0;
}

// The stored compact ciphertext of one row: c_L [3][N] canonical and
// coefficient 0 of c_0..c_{l-1} as [l][3] (tests).
extern "C" int cufe_ctx_get_row(const cufe_ctx *c, int row, uint32_t *cl, uint32_t *c0)
{
    if (!c || row < 0 || row >= c->nrows) {
        ipfe_set_error("cufe_ctx_get_row: row %d not stored", row);
        return -1;
    }
    std::vector<int32_t> a(kPoly);
    for (int s = 0; s < 3; ++s)
        HIP_OK(hipMemcpy(a.data() + (size_t)s * IPFE_N, c->A.get() + ((size_t)s * c->cap + row) * c->ldr,
                         IPFE_N * 4, hipMemcpyDeviceToHost));
    for (int s = 0; s < 3; ++s)
        for (int j = 0; j < IPFE_N; ++j) {
            int64_t v = a[(size_t)s * IPFE_N + j];
            cl[(size_t)s * IPFE_N + j] = (uint32_t)(v < 0 ? v + qmod(s) : v);
        }
    HIP_OK(hipMemcpy(c0, c->c0.get() + (size_t)row * c->row_l * 3, (size_t)c->row_l * 3 * 4, hipMemcpyDeviceToHost));
    return 0;
}
