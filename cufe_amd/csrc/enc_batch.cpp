// enc_batch.cpp -- batched encryption of messages of uneven lengths, kept in
// HBM, with an optional hipGraph replay of the whole batch.
//
// Message b of length l_b is encrypted exactly as rlwe_sife_encrypt
// (src/rlwe_sife.c:112-181) encrypts it under keys of length l_b: r, f_0 from
// the S2 stream (seed2, nonce0 + b), f_0..f_{l_b-1} from the S3 stream (seed3,
// nonce0 + b).  Both streams are read from their start, so the result equals
// the first l_b ciphertext polynomials and c_L of the same message zero-padded
// to any longer key length -- the reference's L = 785 layout included.
//
// A batch plan fixes the lengths once: device buffers, the sampler stream
// tables (grouped so the chunk scratch fits a budget) and the work items of the
// ciphertext kernel.  A run uploads the messages and the round keys, launches
// every group's kernels (eagerly or as one captured graph) and then checks the
// sampler status; a group whose speculative window was short is redone eagerly
// at a wider window, and the graph is captured again on the next run.
#include <string.h>

#include <algorithm>
#include <memory>
#include <vector>

#include "resident.h"

using namespace ipfe;

namespace {
constexpr size_t kPoly = (size_t)3 * IPFE_N;
constexpr size_t kScratchBudget = (size_t)8 << 30;
constexpr int kMaxItemsPerLaunch = 1 << 20;
}  // namespace

namespace ipfe {
const uint32_t *ctx_mpk(const cufe_ctx *c);
int ctx_key_len(const cufe_ctx *c);
}  // namespace ipfe

struct cufe_encbatch {
    const uint32_t *mpk = nullptr;  // [l_key+1][3][N] (index l_key = a_hat)
    int l_key = 0;
    int n = 0;
    std::vector<int> lens;
    std::vector<size_t> xoff, coff;  // words
    size_t x_words = 0, c_words = 0;
    DevBuf<uint32_t> d_x, d_c, d_r, d_rp;
    DevBuf<EncMsg> d_msgs;
    DevBuf<EncItem> d_items;
    struct Group {
        int b0, b1;
        size_t item0, nitems;
        GaussPlan plan;
    };
    std::vector<std::unique_ptr<Group>> groups;
    DevBuf<uint8_t> s_cnt;
    DevBuf<uint64_t> s_vals;
    DevBuf<int> s_pos;
    hipStream_t st = nullptr;
    hipGraphExec_t exec = nullptr;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    DevBuf<unsigned long long> d_sum;
    int redone = 0;  // groups redone at a wider window (diagnostics)

    ~cufe_encbatch()
    {
        if (exec) (void)hipGraphExecDestroy(exec);
        if (ev0) (void)hipEventDestroy(ev0);
        if (ev1) (void)hipEventDestroy(ev1);
        if (st) (void)hipStreamDestroy(st);
    }

    // stream requests of messages [b0, b1): S2 (r -> d_r, f_0 -> c_L) and S3
    // (f_i -> c_i), with the round keys of (seed2, seed3) or zero keys for planning
    std::vector<StreamReq> requests(int b0, int b1, const unsigned char *seed2, const unsigned char *seed3,
                                    int nonce0) const
    {
        static const unsigned char zero[32] = {0};
        std::vector<StreamReq> reqs;
        for (int b = b0; b < b1; ++b) {
            uint32_t *cb = d_c.get() + coff[b];
            const uint16_t nonce = (uint16_t)(nonce0 + b);
            reqs.push_back(make_stream(seed2 ? seed2 : zero, nonce, 2, 2, d_r.get() + (size_t)b * kPoly,
                                       cb + (size_t)lens[b] * kPoly, 0));
            reqs.push_back(make_stream(seed3 ? seed3 : zero, nonce, 3, lens[b], cb, cb + kPoly, (long long)kPoly));
        }
        return reqs;
    }

    // ciphertext kernels of one group: NTT(r), Shoup companions, c_i and c_L
    int enc_group(const Group &g, Device *dev)
    {
        const int cnt = g.b1 - g.b0;
        uint32_t *r = d_r.get() + (size_t)g.b0 * kPoly;
        if (launch_ntt_fwd(r, r, cnt * 3, dev->tw_all, st)) return -1;
        if (launch_shoup_prep(r, d_rp.get() + (size_t)g.b0 * kPoly, (size_t)cnt * kPoly, st)) return -1;
        for (size_t i = 0; i < g.nitems; i += kMaxItemsPerLaunch) {
            const int k = (int)std::min<size_t>(kMaxItemsPerLaunch, g.nitems - i);
            if (launch_enc_poly(d_items.get() + g.item0 + i, k, d_msgs.get(), mpk, l_key, dev->tw_all, dev->ninv,
                                dev->scale, st))
                return -1;
        }
        return 0;
    }

    int launch_all(Device *dev)
    {
        for (auto &g : groups)
            if (g->plan.launch(st) || enc_group(*g, dev)) return -1;
        return 0;
    }
};

// Plans a batch of n messages with lengths lens[b] (1 <= l_b <= key length)
// against the context's master keys; the ciphertexts [l_b+1][3][N] stay on the
// device, packed in message order.
extern "C" int cufe_encbatch_create(const cufe_ctx *ctx, const int *lens, int n, cufe_encbatch **out)
{
    if (!ctx || !lens || !out || n < 1 || n > 65536 || ctx_key_len(ctx) < 1) {
        ipfe_set_error("cufe_encbatch_create: needs a keyed context and 1..65536 messages");
        return -1;
    }
    Device *dev = device();
    if (!dev) return -1;
    auto eb = std::make_unique<cufe_encbatch>();
    eb->mpk = ctx_mpk(ctx);
    eb->l_key = ctx_key_len(ctx);
    eb->n = n;
    eb->lens.assign(lens, lens + n);
    eb->xoff.resize(n);
    eb->coff.resize(n);
    for (int b = 0; b < n; ++b) {
        if (lens[b] < 1 || lens[b] > eb->l_key) {
            ipfe_set_error("cufe_encbatch_create: message %d has length %d (key length %d)", b, lens[b], eb->l_key);
            return -1;
        }
        eb->xoff[b] = eb->x_words;
        eb->coff[b] = eb->c_words;
        eb->x_words += lens[b];
        eb->c_words += (size_t)(lens[b] + 1) * kPoly;
    }
    size_t nitems = eb->c_words / kPoly;
    if (eb->d_x.alloc(eb->x_words) || eb->d_c.alloc(eb->c_words) || eb->d_r.alloc((size_t)n * kPoly) ||
        eb->d_rp.alloc((size_t)n * kPoly) || eb->d_msgs.alloc(n) || eb->d_items.alloc(nitems) || eb->d_sum.alloc(1))
        return -1;
    std::vector<EncMsg> msgs(n);
    std::vector<EncItem> items;
    items.reserve(nitems);
    for (int b = 0; b < n; ++b) {
        EncMsg &m = msgs[b];
        m.x = eb->d_x.get() + eb->xoff[b];
        m.rhat = eb->d_r.get() + (size_t)b * kPoly;
        m.rhatp = eb->d_rp.get() + (size_t)b * kPoly;
        m.c = eb->d_c.get() + eb->coff[b];
        m.cl = m.c + (size_t)lens[b] * kPoly;
        m.len = lens[b];
        m.pad = 0;
        for (int i = 0; i <= lens[b]; ++i) items.push_back(EncItem{b, i});
    }
    HIP_OK(hipMemcpy(eb->d_msgs.get(), msgs.data(), sizeof(EncMsg) * n, hipMemcpyHostToDevice));
    HIP_OK(hipMemcpy(eb->d_items.get(), items.data(), sizeof(EncItem) * nitems, hipMemcpyHostToDevice));
    // groups of consecutive messages whose sampler scratch fits the budget
    size_t item = 0;
    long long max_chunks = 0, max_full = 0;
    for (int b0 = 0; b0 < n;) {
        int b1 = b0;
        size_t bytes = 0;
        long long chunks = 0;
        while (b1 < n) {
            const size_t by = gauss_bytes(2, 2, 0) + gauss_bytes(3, lens[b1], 0);
            const long long ch = gauss_window(2, 2, 0) + gauss_window(3, lens[b1], 0);
            if (b1 > b0 && (bytes + by > kScratchBudget || chunks + ch > (1ll << 28))) break;
            bytes += by;
            chunks += ch;
            ++b1;
        }
        auto g = std::make_unique<cufe_encbatch::Group>();
        g->b0 = b0;
        g->b1 = b1;
        g->item0 = item;
        g->nitems = 0;
        for (int b = b0; b < b1; ++b) g->nitems += (size_t)lens[b] + 1;
        item += g->nitems;
        std::vector<StreamReq> reqs = eb->requests(b0, b1, nullptr, nullptr, 0);
        if (g->plan.build(reqs.data(), (int)reqs.size(), 0)) return -1;
        max_chunks = std::max(max_chunks, g->plan.chunks);
        max_full = std::max(max_full, g->plan.full_chunks);
        eb->groups.push_back(std::move(g));
        b0 = b1;
    }
    if (eb->s_cnt.alloc((size_t)max_chunks) || eb->s_vals.alloc((size_t)max_full * 8 + 1) ||
        eb->s_pos.alloc((size_t)max_full + 1))
        return -1;
    for (auto &g : eb->groups) g->plan.bind_scratch(eb->s_cnt.get(), eb->s_vals.get(), eb->s_pos.get());
    HIP_OK(hipStreamCreateWithFlags(&eb->st, hipStreamNonBlocking));
    HIP_OK(hipEventCreate(&eb->ev0));
    HIP_OK(hipEventCreate(&eb->ev1));
    *out = eb.release();
    return 0;
}

// Encrypts the packed messages x (message b = lens[b] words at its offset;
// entries below the plaintext bound) with message b on nonce nonce0 + b.
// use_graph: replay the captured graph (captured on first use).  ms: device
// time of the run, upload included.
extern "C" int cufe_encbatch_run(cufe_encbatch *eb, const uint32_t *x, const unsigned char seed2[32],
                                 const unsigned char seed3[32], int nonce0, int use_graph, float *ms)
{
    if (!eb || !x || nonce0 < 0 || nonce0 + eb->n > 65536) {
        ipfe_set_error("cufe_encbatch_run: bad batch or nonce range");
        return -1;
    }
    Device *dev = device();
    if (!dev) return -1;
    hipStream_t st = eb->st;
    HIP_OK(hipEventRecord(eb->ev0, st));
    HIP_OK(hipMemcpyAsync(eb->d_x.get(), x, eb->x_words * 4, hipMemcpyHostToDevice, st));
    for (auto &g : eb->groups) {
        std::vector<StreamReq> reqs = eb->requests(g->b0, g->b1, seed2, seed3, nonce0);
        g->plan.rekey(reqs.data());
        if (g->plan.upload(st)) return -1;
    }
    if (use_graph) {
        if (!eb->exec) {
            hipGraph_t graph = nullptr;
            HIP_OK(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
            const int rc = eb->launch_all(dev);
            const hipError_t ec = hipStreamEndCapture(st, &graph);
            if (rc || ec != hipSuccess) {
                if (graph) (void)hipGraphDestroy(graph);
                if (!rc) ipfe_set_error("hipStreamEndCapture failed: %s", hipGetErrorString(ec));
                return -1;
            }
            const hipError_t ei = hipGraphInstantiate(&eb->exec, graph, nullptr, nullptr, 0);
            (void)hipGraphDestroy(graph);
            if (ei != hipSuccess) {
                eb->exec = nullptr;
                ipfe_set_error("hipGraphInstantiate failed: %s", hipGetErrorString(ei));
                return -1;
            }
        }
        HIP_OK(hipGraphLaunch(eb->exec, st));
    } else if (eb->launch_all(dev)) {
        return -1;
    }
    HIP_OK(hipEventRecord(eb->ev1, st));
    for (auto &g : eb->groups) {
        int overflow = 0, need_full = 0;
        if (g->plan.check(st, &overflow, &need_full)) return -1;
        // a short window or a chunk that needed every AES block: redo this
        // group eagerly (wider windows / all blocks) and its ciphertexts
        int level = g->plan.level;
        while (overflow || need_full) {
            if (overflow && ++level > 6) {
                ipfe_set_error("Gaussian sampler: speculative window exhausted");
                return -1;
            }
            const int full = g->plan.full | need_full;
            std::vector<StreamReq> reqs = eb->requests(g->b0, g->b1, seed2, seed3, nonce0);
            if (g->plan.build(reqs.data(), (int)reqs.size(), level) || g->plan.alloc_scratch()) return -1;
            g->plan.full = full;
            if (g->plan.upload(st) || g->plan.launch(st) || g->plan.check(st, &overflow, &need_full)) return -1;
            if (need_full && full) need_full = 0;  // already ran with every block
            if (!overflow && !need_full && eb->enc_group(*g, dev)) return -1;
            eb->redone++;
            if (eb->exec) {
                (void)hipGraphExecDestroy(eb->exec);  // its kernels hold the old tables
                eb->exec = nullptr;
            }
        }
    }
    HIP_OK(hipStreamSynchronize(st));
    float e = 0;
    HIP_OK(hipEventElapsedTime(&e, eb->ev0, eb->ev1));
    if (ms) *ms = e;
    return 0;
}

// Ciphertexts of messages [b0, b0+nb) to host memory, packed [l_b+1][3][N].
extern "C" int cufe_encbatch_export(const cufe_encbatch *eb, int b0, int nb, uint32_t *c)
{
    if (!eb || !c || b0 < 0 || nb < 0 || b0 + nb > eb->n) {
        ipfe_set_error("cufe_encbatch_export: range [%d, %d) outside the batch", b0, b0 + nb);
        return -1;
    }
    if (nb == 0) return 0;
    const size_t end = (b0 + nb < eb->n) ? eb->coff[b0 + nb] : eb->c_words;
    HIP_OK(hipMemcpy(c, eb->d_c.get() + eb->coff[b0], (end - eb->coff[b0]) * 4, hipMemcpyDeviceToHost));
    return 0;
}

// Position-weighted 64-bit checksum of every ciphertext word of the batch,
// computed on the device; info (optional, 4 values): messages, total words,
// sampler groups, groups redone at a wider window so far.
extern "C" int cufe_encbatch_checksum(cufe_encbatch *eb, uint64_t *sum, long long *info)
{
    if (!eb) {
        ipfe_set_error("cufe_encbatch_checksum: null batch");
        return -1;
    }
    if (sum) {
        if (launch_fill_i32((int *)eb->d_sum.get(), 0, 2, eb->st)) return -1;
        if (launch_checksum(eb->d_c.get(), eb->c_words, eb->d_sum.get(), eb->st)) return -1;
        unsigned long long h = 0;
        HIP_OK(hipMemcpyAsync(&h, eb->d_sum.get(), 8, hipMemcpyDeviceToHost, eb->st));
        HIP_OK(hipStreamSynchronize(eb->st));
        *sum = h;
    }
    if (info) {
        info[0] = eb->n;
        info[1] = (long long)eb->c_words;
        info[2] = (long long)eb->groups.size();
        info[3] = eb->redone;
    }
    return 0;
}

extern "C" void cufe_encbatch_destroy(cufe_encbatch *eb) { delete eb; }
