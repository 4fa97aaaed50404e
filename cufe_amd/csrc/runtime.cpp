// runtime.cpp -- device state, error reporting and the Gaussian-sampler driver.
#include "runtime.h"

#include "pset.h"

#include <stdarg.h>
#include <stdio.h>
#include <string.h>

#include <math.h>

#include <algorithm>
#include <map>
#include <mutex>

namespace ipfe {

const uint32_t *g_aes_te = nullptr;

static thread_local char t_err[1024];

}  // namespace ipfe

void ipfe_set_error(const char *fmt, ...)
{
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(ipfe::t_err, sizeof ipfe::t_err, fmt, ap);
    va_end(ap);
}

extern "C" const char *cufe_last_error(void) { return ipfe::t_err; }

namespace ipfe {

// The constant tables live in static device arrays of this file, so that the
// library's set-up allocates no device memory.  Host code reaches them only
// through the addresses hipGetSymbolAddress returns.
__device__ __align__(256) uint32_t d_tw_all_tab[4 * 3 * IPFE_N];
__device__ __align__(256) uint32_t d_aes_te_tab[1024];

static Device *init_device()
{
    const HostTables &ht = host_tables();
    static Device dev;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1) {
        ipfe_set_error("no HIP device available");
        return nullptr;
    }
    if (hipGetDevice(&dev.id) != hipSuccess ||
        hipGetSymbolAddress((void **)&dev.tw_all, HIP_SYMBOL(d_tw_all_tab)) != hipSuccess ||
        hipGetSymbolAddress((void **)&dev.aes_te, HIP_SYMBOL(d_aes_te_tab)) != hipSuccess) {
        ipfe_set_error("device addresses of the constant tables unavailable");
        return nullptr;
    }
    static_assert(sizeof(d_tw_all_tab) == 4 * 3 * IPFE_N * 4 && sizeof(d_aes_te_tab) == 1024 * 4, "table sizes");
    if (ht.tw_all.size() * 4 != sizeof(d_tw_all_tab) || ht.aes_te.size() * 4 != sizeof(d_aes_te_tab) ||
        hipMemcpy(dev.tw_all, ht.tw_all.data(), sizeof(d_tw_all_tab), hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(dev.aes_te, ht.aes_te.data(), sizeof(d_aes_te_tab), hipMemcpyHostToDevice) != hipSuccess) {
        ipfe_set_error("upload of constant tables failed");
        return nullptr;
    }
    dev.ninv = ht.ninv;
    dev.scale = ht.scale;
    dev.barrett = ht.barrett;
    dev.stream = nullptr;  // default stream
    g_aes_te = dev.aes_te;
    return &dev;
}

Device *device()
{
    static std::once_flag once;
    static Device *d = nullptr;
    static char err[1024];
    std::call_once(once, [] {
        d = init_device();
        if (!d) snprintf(err, sizeof err, "%s", cufe_last_error());
    });
    if (!d) ipfe_set_error("device initialisation failed: %s", err);
    return d;
}

int sync_stream(hipStream_t st)
{
    HIP_OK(hipStreamSynchronize(st));
    return 0;
}

StreamReq make_stream(const uint8_t key[32], uint16_t nonce, int which, int ncalls, uint32_t *out0, uint32_t *out1,
                      long long out_stride, int coef0_only)
{
    StreamReq r;
    aes256_expand_key(key, r.rk);
    r.nonce = nonce;
    r.which = which;
    r.ncalls = ncalls;
    r.out0 = out0;
    r.out1 = out1;
    r.out_stride = out_stride;
    r.coef0_only = coef0_only;
    r.slen = IPFE_N;
    r.nmod = 3;
    for (int s = 0; s < 4; ++s) {
        r.q[s] = s < 3 ? qmod(s) : 0u;
        r.fold[s] = s < 3 ? (fold_a(s) | (fold_b(s) << 8)) : 0u;
    }
    return r;
}

StreamReq make_stream_ps(const PsetSpec &ps, const uint8_t key[32], uint16_t nonce, int which, int ncalls,
                         uint32_t *out0, uint32_t *out1, long long out_stride)
{
    StreamReq r = make_stream(key, nonce, ps.width[which], ncalls, out0, out1, out_stride, 0);
    r.slen = ps.N;
    r.nmod = ps.nmod;
    for (int s = 0; s < 4; ++s) {
        r.q[s] = s < ps.nmod ? ps.q[s] : 0u;
        r.fold[s] = s < ps.nmod ? (ps.fa[s] | (ps.fb[s] << 8)) : 0u;
    }
    return r;
}

// Chunks per 4096 samples of each sampler width, measured on the reference
// algorithm: S1 accepts 68.1% of its candidates (751.6 chunks per 4096
// samples, standard deviation 6.6); the Bernoulli step of S2 / S3 is distorted
// for large k (the int-to-double trick of src/gauss.c:368-369 is wrong above
// 2^52), so S2 accepts 53.8% (951.8 chunks, sd 10.1) and S3 38.7% (1323.4
// chunks, sd 16.2).  The N = 2048 set's widths 262147 and 524287 stay below
// that range and accept 68.0%, as S1 (752.8 chunks per 4096, sd 6.6-6.8).  A
// call of slen samples takes slen / 4096 of that (sd scaled by its square
// root).  The window takes the mean plus 1.5% plus 8 standard deviations;
// level m doubles the margin m times (a short window is detected and retried,
// never wrong).
long long gauss_window(int which, int slen, int ncalls, int m)
{
    static const double mean[6] = {0.0, 751.6, 951.8, 1323.4, 752.8, 753.4};
    static const double sd[6] = {0.0, 6.6, 10.1, 16.2, 6.6, 6.8};
    const int w = (which >= 1 && which <= 5) ? which : 3;
    const double n = (double)ncalls * (double)slen / 4096.0;
    double margin = 0.015 * mean[w] * n + 8.0 * sd[w] * sqrt(n) + 64.0;
    for (int i = 0; i < m; ++i) margin *= 2.0;
    return (long long)ceil(mean[w] * n + margin);
}

long long gauss_window(int which, int ncalls, int m) { return gauss_window(which, IPFE_N, ncalls, m); }

size_t gauss_bytes(int which, int slen, int ncalls, int coef0_only)
{
    const long long ch = gauss_window(which, slen, ncalls, 0);
    return coef0_only ? (size_t)ch * 1 + (size_t)ncalls * 8 : (size_t)ch * (1 + 64 + 4);
}

size_t gauss_bytes(int which, int ncalls, int coef0_only) { return gauss_bytes(which, IPFE_N, ncalls, coef0_only); }

int GaussPlan::build(const StreamReq *reqs, int count, int lvl, bool alloc_meta)
{
    constexpr int kTile = GAUSS_TILE_CHUNKS;
    n = count;
    level = lvl;
    hs.assign(n, GaussStream{});
    tile_prefix.assign(n, 0);
    chunk_prefix.assign(n, 0);
    first_prefix.assign(n, 0);
    tiles = chunks = full_chunks = nfirst = 0;
    // full streams first, so per-chunk values and labels cover only their chunks
    std::vector<int> order;
    for (int i = 0; i < n; ++i)
        if (!reqs[i].coef0_only) order.push_back(i);
    for (int i = 0; i < n; ++i)
        if (reqs[i].coef0_only) order.push_back(i);
    for (int i = 0; i < n; ++i) {
        const StreamReq &rq = reqs[order[i]];
        GaussStream &g = hs[i];
        memcpy(g.rk, rq.rk, sizeof g.rk);
        g.nonce = rq.nonce;
        g.which = rq.which;
        g.ncalls = rq.ncalls;
        g.nchunks = (int)gauss_window(rq.which, rq.slen, rq.ncalls, level);
        g.chunk_base = chunks;
        g.out0 = rq.out0;
        g.out1 = rq.out1;
        g.out_stride = rq.out_stride;
        g.coef0_only = rq.coef0_only;
        g.slen = rq.slen;
        g.first_base = nfirst;
        for (int t = 0; t < 4; ++t) {
            g.q[t] = rq.q[t];
            g.fold[t] = rq.fold[t];
        }
        g.nmod = rq.nmod;
        g.pad = 0;
        tile_prefix[i] = tiles;
        chunk_prefix[i] = chunks;
        first_prefix[i] = nfirst;
        tiles += (g.nchunks + kTile - 1) / kTile;
        chunks += g.nchunks;
        if (g.coef0_only) nfirst += g.ncalls;
        else full_chunks = chunks;
    }
    if (!alloc_meta) return 0;
    if (d_streams.alloc(n) || d_tp.alloc(n) || d_cp.alloc(n) || d_fp.alloc(n) || d_status.alloc(n + 1) ||
        d_first.alloc((size_t)nfirst + 1))
        return -1;
    m_streams = d_streams.get();
    m_tp = d_tp.get();
    m_cp = d_cp.get();
    m_fp = d_fp.get();
    m_first = d_first.get();
    m_status = d_status.get();
    return 0;
}

static size_t align256(size_t b) { return (b + 255) & ~(size_t)255; }

size_t GaussPlan::meta_bytes() const
{
    return align256(sizeof(GaussStream) * n) + 3 * align256(sizeof(long long) * n) +
           align256(sizeof(long long) * ((size_t)nfirst + 1)) + align256(sizeof(int) * ((size_t)n + 1));
}

int GaussPlan::bind_meta(uint8_t *base)
{
    m_streams = (GaussStream *)base;
    base += align256(sizeof(GaussStream) * n);
    m_tp = (long long *)base;
    base += align256(sizeof(long long) * n);
    m_cp = (long long *)base;
    base += align256(sizeof(long long) * n);
    m_fp = (long long *)base;
    base += align256(sizeof(long long) * n);
    m_first = (long long *)base;
    base += align256(sizeof(long long) * ((size_t)nfirst + 1));
    m_status = (int *)base;
    return 0;
}

size_t GaussPlan::bound_scratch_bytes() const
{
    return align256((size_t)chunks) + align256((size_t)full_chunks * 64 + 8) + align256((size_t)full_chunks * 4 + 4);
}

int GaussPlan::bind_scratch_region(uint8_t *base)
{
    uint8_t *cnt = base;
    base += align256((size_t)chunks);
    uint64_t *vals = (uint64_t *)base;
    base += align256((size_t)full_chunks * 64 + 8);
    return bind_scratch(cnt, vals, (int *)base);
}

size_t GaussPlan::scratch_bytes() const { return (size_t)chunks + (size_t)full_chunks * (64 + 4) + 64; }

int GaussPlan::bind_scratch(uint8_t *cnt, uint64_t *vals, int *pos)
{
    s_cnt = cnt;
    s_vals = vals;
    s_pos = pos;
    return 0;
}

int GaussPlan::alloc_scratch()
{
    if (own_cnt.alloc((size_t)chunks) || own_vals.alloc((size_t)full_chunks * 8 + 1) ||
        own_pos.alloc((size_t)full_chunks + 1))
        return -1;
    return bind_scratch(own_cnt.get(), own_vals.get(), own_pos.get());
}

// Re-keys the streams in place (same shapes, new round keys / nonces).
void GaussPlan::rekey(const StreamReq *reqs)
{
    std::vector<int> order;
    for (int i = 0; i < n; ++i)
        if (!reqs[i].coef0_only) order.push_back(i);
    for (int i = 0; i < n; ++i)
        if (reqs[i].coef0_only) order.push_back(i);
    for (int i = 0; i < n; ++i) {
        memcpy(hs[i].rk, reqs[order[i]].rk, sizeof hs[i].rk);
        hs[i].nonce = reqs[order[i]].nonce;
    }
}

int GaussPlan::upload(hipStream_t st)
{
    HIP_OK(hipMemcpyAsync(m_streams, hs.data(), sizeof(GaussStream) * n, hipMemcpyHostToDevice, st));
    HIP_OK(hipMemcpyAsync(m_tp, tile_prefix.data(), sizeof(long long) * n, hipMemcpyHostToDevice, st));
    HIP_OK(hipMemcpyAsync(m_cp, chunk_prefix.data(), sizeof(long long) * n, hipMemcpyHostToDevice, st));
    HIP_OK(hipMemcpyAsync(m_fp, first_prefix.data(), sizeof(long long) * n, hipMemcpyHostToDevice, st));
    return 0;
}

// Kernels only (fills, candidates, walk, scatter, firsts): capturable.
int GaussPlan::launch(hipStream_t st)
{
    if (launch_fill_i32(s_pos, -1, (size_t)full_chunks, st) || launch_fill_i32(m_status, 0, n + 1, st)) return -1;
    if (launch_gauss_candidates(m_streams, n, m_tp, tiles, s_cnt, s_vals, m_status + n, full, st)) return -1;
    if (launch_gauss_walk(m_streams, n, s_cnt, s_pos, m_first, m_status, st)) return -1;
    if (launch_gauss_scatter(m_streams, n, m_cp, full_chunks, s_cnt, s_vals, s_pos, st)) return -1;
    if (launch_gauss_firsts(m_streams, n, m_fp, m_first, nfirst, m_status + n, full, st)) return -1;
    return 0;
}

int GaussPlan::check(hipStream_t st, int *overflow, int *need_full)
{
    std::vector<int> status(n + 1);
    HIP_OK(hipMemcpyAsync(status.data(), m_status, sizeof(int) * (n + 1), hipMemcpyDeviceToHost, st));
    HIP_OK(hipStreamSynchronize(st));
    if (status[n] & 2) {
        ipfe_set_error("Gaussian sampler: a chunk needed more than 46 uniform words (reference reads past its buffer)");
        return -1;
    }
    *need_full = (status[n] & 4) ? 1 : 0;
    *overflow = 0;
    for (int i = 0; i < n; ++i) *overflow |= status[i];
    return 0;
}

static int run_gauss_group(const StreamReq *reqs, int n, int level, hipStream_t st, int *overflow)
{
    GaussPlan p;
    if (p.build(reqs, n, level) || p.alloc_scratch()) return -1;
    for (;;) {
        int need_full = 0;
        if (p.upload(st) || p.launch(st) || p.check(st, overflow, &need_full)) return -1;
        if (!need_full || p.full) return 0;
        p.full = 1;  // a chunk needed the AES blocks the default set omits
    }
}

int run_gauss(const std::vector<StreamReq> &reqs, hipStream_t st, size_t scratch_budget)
{
    size_t i = 0;
    while (i < reqs.size()) {
        // group streams so that the chunk scratch stays within budget
        // and the chunk count stays below 2^28 (grid sizes, 32-bit work-item ids)
        size_t j = i, bytes = 0;
        long long nch = 0;
        while (j < reqs.size()) {
            size_t b = gauss_bytes(reqs[j].which, reqs[j].slen, reqs[j].ncalls, reqs[j].coef0_only);
            long long c = gauss_window(reqs[j].which, reqs[j].slen, reqs[j].ncalls, 0);
            if (j > i && (bytes + b > scratch_budget || nch + c > (1ll << 28))) break;
            bytes += b;
            nch += c;
            ++j;
        }
        int level = 0;
        for (;;) {
            int overflow = 0;
            if (run_gauss_group(reqs.data() + i, (int)(j - i), level, st, &overflow)) return -1;
            if (!overflow) break;
            if (++level > 6) {
                ipfe_set_error("Gaussian sampler: speculative window exhausted");
                return -1;
            }
        }
        i = j;
    }
    return 0;
}

}  // namespace ipfe

// ---------------------------------------------------------------------------
// Pinned host memory handed out by the library (hipHostMalloc), registered so
// that the server can DMA request buffers that lie inside it directly.
// ---------------------------------------------------------------------------
namespace {
std::mutex g_pin_mu;
std::map<uintptr_t, size_t> g_pinned;  // start -> bytes
}  // namespace

namespace ipfe {
bool host_pinned(const void *p, size_t bytes)
{
    const uintptr_t a = (uintptr_t)p;
    std::lock_guard<std::mutex> lk(g_pin_mu);
    auto it = g_pinned.upper_bound(a);
    if (it == g_pinned.begin()) return false;
    --it;
    return a >= it->first && a + bytes <= it->first + it->second;
}
}  // namespace ipfe

extern "C" int cufe_host_alloc(size_t bytes, void **out)
{
    if (!out || bytes == 0) {
        ipfe_set_error("cufe_host_alloc: bad arguments");
        return -1;
    }
    if (!ipfe::device()) return -1;
    void *p = nullptr;
    const hipError_t e = hipHostMalloc(&p, bytes, hipHostMallocDefault);
    if (e != hipSuccess) {
        ipfe_set_error("hipHostMalloc of %zu bytes failed: %s", bytes, hipGetErrorString(e));
        return -1;
    }
    std::lock_guard<std::mutex> lk(g_pin_mu);
    g_pinned[(uintptr_t)p] = bytes;
    *out = p;
    return 0;
}

extern "C" void cufe_host_free(void *p)
{
    if (!p) return;
    {
        std::lock_guard<std::mutex> lk(g_pin_mu);
        g_pinned.erase((uintptr_t)p);
    }
    (void)hipHostFree(p);
}
