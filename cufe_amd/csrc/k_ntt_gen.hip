// k_ntt_gen.hip -- forward / inverse NTT kernels for every parameter set (N =
// 2048, 4096, 8192; any limb count), in several shapes, and the autotuner that
// picks a shape per (set, direction) on the device it runs on.
//
// A shape is the split of the log2(N) radix-2 stages into register phases,
// the threads per transform (so the residues each thread holds), the
// transforms per workgroup, and whether the twiddles are read from HBM (L2)
// or staged in LDS first.  Every shape computes the same result bit for bit
// (CT_forward / GS_reverse, src/ntt.c:38-114); the tuner times each on
// synthetic residues, checks that all agree (a checksum of every output word)
// and keeps the fastest.  For the reference set the hand-scheduled radix-16^3
// kernel of k_ntt.hip is one of the candidates.
#include <stdio.h>
#include <stdlib.h>

#include <mutex>
#include <string>
#include <vector>

#include "kernels.h"
#include "ntt_gen.h"
#include "ntt_variants.h"

namespace ipfe {

namespace {

template <int LOGN, int T, int LPW, typename RAD, bool TWLDS>
struct NttKernelShape {
    using G = NttGen<LOGN, T, RAD>;
    static constexpr int N = G::N, E = G::E;
    static constexpr int LDS_WORDS = LPW * G::LDS_WORDS + (TWLDS ? 2 * N : 0);
    static_assert(!TWLDS || LPW == 1, "LDS twiddles are staged for one limb per workgroup");
    static_assert(T * LPW <= 1024, "workgroup size");
};

// Stages the limb's twiddles (w and Shoup companions) in LDS, or returns the
// HBM tables.
template <int N, bool TWLDS>
__device__ __forceinline__ void twiddles(uint32_t *lds_tw, const uint32_t *tw, const uint32_t *twp,
                                         const uint32_t *&t0, const uint32_t *&t1)
{
    if constexpr (TWLDS) {
        for (int i = threadIdx.x; i < N; i += blockDim.x) {
            lds_tw[i] = tw[i];
            lds_tw[N + i] = twp[i];
        }
        __syncthreads();
        t0 = lds_tw;
        t1 = lds_tw + N;
    } else {
        t0 = tw;
        t1 = twp;
    }
}

template <int LOGN, int T, int LPW, typename RAD, bool TWLDS>
__global__ __launch_bounds__(T *LPW) void k_ntt_gen_fwd(const uint32_t *__restrict__ in, uint32_t *__restrict__ out,
                                                        int ntrans, const uint32_t *__restrict__ tw_all, NttArgs a)
{
    using K = NttKernelShape<LOGN, T, LPW, RAD, TWLDS>;
    using G = typename K::G;
    constexpr int N = K::N, E = K::E;
    __shared__ uint32_t lds[K::LDS_WORDS];
    const int slot = threadIdx.x / T, t = threadIdx.x % T;
    const int z = blockIdx.x * LPW + slot;
    const bool valid = z < ntrans;
    const int s = valid ? z % a.nmod : 0;
    const uint32_t q = a.q[s];
    const uint32_t *tw, *twp;
    twiddles<N, TWLDS>(lds + LPW * G::LDS_WORDS, tw_all + (size_t)s * N, tw_all + (size_t)(a.nmod + s) * N, tw, twp);
    uint32_t x[E];
    bool hq = false;
    const uint32_t *src = in + (size_t)(valid ? z : 0) * N;
#pragma unroll
    for (int e = 0; e < E; ++e) {
        x[e] = valid ? src[G::first_idx(t, e)] : 0u;
        hq |= x[e] == q;
    }
    uint32_t *my = lds + slot * G::LDS_WORDS;
    if (__syncthreads_or(hq))
        G::template fwd<true>(x, my, tw, twp, q, t);
    else
        G::template fwd<false>(x, my, tw, twp, q, t);
    if (!valid) return;
    uint32_t *dst = out + (size_t)z * N;
    constexpr int RL = 1 << PhaseOf<RAD, G::NP - 1>::lr;
    if constexpr (RL % 4 == 0) {
#pragma unroll
        for (int e = 0; e < E; e += 4)
            *(uint4 *)(dst + G::last_idx(t, e)) = make_uint4(x[e], x[e + 1], x[e + 2], x[e + 3]);
    } else {
#pragma unroll
        for (int e = 0; e < E; ++e) dst[G::last_idx(t, e)] = x[e];
    }
}

template <int LOGN, int T, int LPW, typename RAD, bool TWLDS>
__global__ __launch_bounds__(T *LPW) void k_ntt_gen_inv(const uint32_t *__restrict__ in, uint32_t *__restrict__ out,
                                                        int ntrans, const uint32_t *__restrict__ tw_all, NttArgs a)
{
    using K = NttKernelShape<LOGN, T, LPW, RAD, TWLDS>;
    using G = typename K::G;
    constexpr int N = K::N, E = K::E;
    __shared__ uint32_t lds[K::LDS_WORDS];
    const int slot = threadIdx.x / T, t = threadIdx.x % T;
    const int z = blockIdx.x * LPW + slot;
    const bool valid = z < ntrans;
    const int s = valid ? z % a.nmod : 0;
    const uint32_t q = a.q[s];
    const uint32_t *tw, *twp;
    twiddles<N, TWLDS>(lds + LPW * G::LDS_WORDS, tw_all + (size_t)(2 * a.nmod + s) * N,
                       tw_all + (size_t)(3 * a.nmod + s) * N, tw, twp);
    uint32_t x[E];
    bool hq = false;
    const uint32_t *src = in + (size_t)(valid ? z : 0) * N;
    constexpr int RL = 1 << PhaseOf<RAD, G::NP - 1>::lr;
    if constexpr (RL % 4 == 0) {
#pragma unroll
        for (int e = 0; e < E; e += 4) {
            uint4 v = valid ? *(const uint4 *)(src + G::last_idx(t, e)) : make_uint4(0, 0, 0, 0);
            x[e] = v.x;
            x[e + 1] = v.y;
            x[e + 2] = v.z;
            x[e + 3] = v.w;
        }
    } else {
#pragma unroll
        for (int e = 0; e < E; ++e) x[e] = valid ? src[G::last_idx(t, e)] : 0u;
    }
#pragma unroll
    for (int e = 0; e < E; ++e) hq |= x[e] == q;
    uint32_t *my = lds + slot * G::LDS_WORDS;
    uint32_t *dst = out + (size_t)z * N;
    if (__syncthreads_or(hq)) {
        G::template inv<true>(x, my, tw, twp, q, t);
        if (!valid) return;
#pragma unroll
        for (int e = 0; e < E; ++e)
            dst[G::first_idx(t, e)] = scale_ninv<true>(x[e], a.ninv[s], a.ninvp[s], a.ninvf[s], q);
    } else {
        G::template inv<false>(x, my, tw, twp, q, t);
        if (!valid) return;
#pragma unroll
        for (int e = 0; e < E; ++e)
            dst[G::first_idx(t, e)] = scale_ninv<false>(x[e], a.ninv[s], a.ninvp[s], a.ninvf[s], q);
    }
}

// Synthetic residues for tuning: word i of a [ntrans][N] buffer, limb (i / N)
// mod nmod, from a 32-bit mix of i.
__global__ void k_fill_residues(uint32_t *p, size_t n, int logn, NttArgs a)
{
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) {
        uint32_t h = (uint32_t)i * 0x9E3779B1u + 0x7F4A7C15u;
        h ^= h >> 15;
        h *= 0x2C1B3C6Du;
        h ^= h >> 12;
        const int s = (int)((i >> logn) % (size_t)a.nmod);
        p[i] = h % a.q[s];
    }
}

template <int LOGN, int T, int LPW, typename RAD, bool TWLDS>
int run_gen(bool inverse, const uint32_t *in, uint32_t *out, int ntrans, const NttArgs &a, const uint32_t *tw_all,
            hipStream_t st)
{
    if (ntrans <= 0) return 0;
    const dim3 grid((unsigned)((ntrans + LPW - 1) / LPW)), block(T * LPW);
    if (inverse)
        hipLaunchKernelGGL((k_ntt_gen_inv<LOGN, T, LPW, RAD, TWLDS>), grid, block, 0, st, in, out, ntrans, tw_all, a);
    else
        hipLaunchKernelGGL((k_ntt_gen_fwd<LOGN, T, LPW, RAD, TWLDS>), grid, block, 0, st, in, out, ntrans, tw_all, a);
    HIP_OK(hipGetLastError());
    return 0;
}

// The hand-scheduled radix-16^3 kernels of k_ntt.hip (reference set only).
int run_hand(bool inverse, const uint32_t *in, uint32_t *out, int ntrans, const NttArgs &a, const uint32_t *tw_all,
             hipStream_t st)
{
    NinvConsts nc;
    for (int s = 0; s < 3; ++s) {
        nc.ninv[s] = a.ninv[s];
        nc.ninvp[s] = a.ninvp[s];
        nc.flag[s] = a.ninvf[s];
    }
    return inverse ? launch_ntt_inv(in, out, ntrans, tw_all, nc, st) : launch_ntt_fwd(in, out, ntrans, tw_all, st);
}

struct Variant {
    const char *name;
    int logn;
    bool ref_only;  // the hand kernel: reference set (three fixed limbs) only
    NttRunFn run;
};

// Shapes per N: "radices/threads xLPW", "tl" = twiddles staged in LDS.
const Variant kVariants[] = {
    {"16.16.8/128", 11, false, run_gen<11, 128, 1, Rad<4, 4, 3>, false>},
    {"16.16.8/128x2", 11, false, run_gen<11, 128, 2, Rad<4, 4, 3>, false>},
    {"16.16.8/128x4", 11, false, run_gen<11, 128, 4, Rad<4, 4, 3>, false>},
    {"16.16.8/128tl", 11, false, run_gen<11, 128, 1, Rad<4, 4, 3>, true>},
    {"8.8.8.4/256", 11, false, run_gen<11, 256, 1, Rad<3, 3, 3, 2>, false>},
    {"32.8.8/64x4", 11, false, run_gen<11, 64, 4, Rad<5, 3, 3>, false>},
    {"16.16.16/256 hand", 12, true, run_hand},
    {"16.16.16/256", 12, false, run_gen<12, 256, 1, Rad<4, 4, 4>, false>},
    {"16.16.16/256x2", 12, false, run_gen<12, 256, 2, Rad<4, 4, 4>, false>},
    {"16.16.16/256tl", 12, false, run_gen<12, 256, 1, Rad<4, 4, 4>, true>},
    {"16.16.16/128x2", 12, false, run_gen<12, 128, 2, Rad<4, 4, 4>, false>},
    {"32.16.8/128x2", 12, false, run_gen<12, 128, 2, Rad<5, 4, 3>, false>},
    {"8.8.8.8/512", 12, false, run_gen<12, 512, 1, Rad<3, 3, 3, 3>, false>},
    {"16.16.16/64x4", 12, false, run_gen<12, 64, 4, Rad<4, 4, 4>, false>},
    {"32.16.16/256", 13, false, run_gen<13, 256, 1, Rad<5, 4, 4>, false>},
    {"32.16.16/256x2", 13, false, run_gen<13, 256, 2, Rad<5, 4, 4>, false>},
    {"32.16.16/256tl", 13, false, run_gen<13, 256, 1, Rad<5, 4, 4>, true>},
    {"16.8.8.8/512", 13, false, run_gen<13, 512, 1, Rad<4, 3, 3, 3>, false>},
    {"32.16.16/128", 13, false, run_gen<13, 128, 1, Rad<5, 4, 4>, false>},
    {"16.16.16.2/512", 13, false, run_gen<13, 512, 1, Rad<4, 4, 4, 1>, false>},
};
constexpr int kNumVariants = (int)(sizeof(kVariants) / sizeof(kVariants[0]));

// Index in kVariants of variant v of logn, or -1.
int variant_index(int logn, int v)
{
    int k = 0;
    for (int i = 0; i < kNumVariants; ++i)
        if (kVariants[i].logn == logn && k++ == v) return i;
    return -1;
}

struct Tuned {
    int variant[PSET_SLOTS][2];  // per set and direction; -1 = not tuned yet
    std::string report;
    std::mutex mu;
    Tuned()
    {
        for (auto &v : variant) v[0] = v[1] = -1;
    }
};
Tuned &tuned()
{
    static Tuned t;
    return t;
}

struct Timer {
    hipEvent_t a = nullptr, b = nullptr;
    Timer()
    {
        (void)hipEventCreate(&a);
        (void)hipEventCreate(&b);
    }
    ~Timer()
    {
        (void)hipEventDestroy(a);
        (void)hipEventDestroy(b);
    }
};

}  // namespace

int ntt_variant_count(int logn)
{
    int k = 0;
    for (int i = 0; i < kNumVariants; ++i) k += kVariants[i].logn == logn;
    return k;
}

const char *ntt_variant_name(int logn, int v)
{
    const int i = variant_index(logn, v);
    return i < 0 ? nullptr : kVariants[i].name;
}

bool ntt_variant_ok(int logn, int v, bool ref_set)
{
    const int i = variant_index(logn, v);
    return i >= 0 && (ref_set || !kVariants[i].ref_only);
}

int ntt_run_variant(int logn, int v, bool inverse, const uint32_t *in, uint32_t *out, int ntrans, const NttArgs &a,
                    const uint32_t *tw_all, bool ref_set, hipStream_t st)
{
    const int i = variant_index(logn, v);
    if (i < 0 || (kVariants[i].ref_only && !ref_set)) {
        ipfe_set_error("NTT: no variant %d for N = %d on this parameter set", v, 1 << logn);
        return -1;
    }
    return kVariants[i].run(inverse, in, out, ntrans, a, tw_all, st);
}

// The tuned variant of (set, direction): on first use, every valid variant
// runs on ntune synthetic transforms, twice in alternating order, with HIP
// events around 3 launches; outputs must agree word for word (checksum).
int ntt_tuned_variant(int set_slot, int logn, bool inverse, const NttArgs &a, const uint32_t *tw_all, bool ref_set,
                      hipStream_t st)
{
    if (const char *e = getenv("CUFE_NTT_VARIANT")) {
        const int v = atoi(e);
        if (ntt_variant_ok(logn, v, ref_set)) return v;
    }
    Tuned &T = tuned();
    std::lock_guard<std::mutex> lk(T.mu);
    int &slot = T.variant[set_slot][inverse ? 1 : 0];
    if (slot >= 0) return slot;
    const int N = 1 << logn;
    const int ntune = (int)(((size_t)256 << 20) / ((size_t)N * 4) / (size_t)a.nmod * (size_t)a.nmod);  // 256 MiB
    const size_t words = (size_t)ntune * N;
    uint32_t *src = nullptr, *dst = nullptr;
    unsigned long long *sum = nullptr;
    if (hipMalloc((void **)&src, words * 4) != hipSuccess || hipMalloc((void **)&dst, words * 4) != hipSuccess ||
        hipMalloc((void **)&sum, 8) != hipSuccess) {
        (void)hipFree(src);
        (void)hipFree(dst);
        ipfe_set_error("NTT autotune: allocation of %zu bytes failed", 2 * words * 4);
        return -1;
    }
    auto cleanup = [&] {
        (void)hipFree(src);
        (void)hipFree(dst);
        (void)hipFree(sum);
    };
    hipLaunchKernelGGL(k_fill_residues, dim3(4096), dim3(256), 0, st, src, words, logn, a);
    const int nv = ntt_variant_count(logn);
    std::vector<float> best(nv, 1e30f);
    std::vector<unsigned long long> check(nv, 0);
    std::vector<int> valid(nv, 0);
    Timer tm;
    for (int pass = 0; pass < 2; ++pass)
        for (int k = 0; k < nv; ++k) {
            const int v = pass ? nv - 1 - k : k;
            if (!ntt_variant_ok(logn, v, ref_set)) continue;
            if (ntt_run_variant(logn, v, inverse, src, dst, ntune, a, tw_all, ref_set, st)) {
                cleanup();
                return -1;
            }
            if (pass == 0) {
                unsigned long long h = 0;
                (void)hipMemsetAsync(sum, 0, 8, st);
                if (launch_checksum(dst, words, sum, st) ||
                    hipMemcpyAsync(&h, sum, 8, hipMemcpyDeviceToHost, st) != hipSuccess ||
                    hipStreamSynchronize(st) != hipSuccess) {
                    cleanup();
                    ipfe_set_error("NTT autotune: checksum failed");
                    return -1;
                }
                check[v] = h;
            }
            (void)hipEventRecord(tm.a, st);
            for (int it = 0; it < 3; ++it) (void)ntt_run_variant(logn, v, inverse, src, dst, ntune, a, tw_all, ref_set, st);
            (void)hipEventRecord(tm.b, st);
            (void)hipEventSynchronize(tm.b);
            float ms = 0;
            (void)hipEventElapsedTime(&ms, tm.a, tm.b);
            best[v] = std::min(best[v], ms / 3);
            valid[v] = 1;
        }
    cleanup();
    // every variant must agree with the first valid one
    int first = -1, pick = -1;
    char line[256];
    std::string rep;
    snprintf(line, sizeof line, "NTT N=%d limbs=%d %s (%d transforms):", N, a.nmod, inverse ? "inv" : "fwd", ntune);
    rep += line;
    for (int v = 0; v < nv; ++v) {
        if (!valid[v]) continue;
        if (first < 0) first = v;
        const bool same = check[v] == check[first];
        snprintf(line, sizeof line, " %s %.3f%s", ntt_variant_name(logn, v), best[v], same ? "" : " MISMATCH");
        rep += line;
        if (same && (pick < 0 || best[v] < best[pick])) pick = v;
    }
    snprintf(line, sizeof line, " ms -> %s\n", pick >= 0 ? ntt_variant_name(logn, pick) : "none");
    rep += line;
    T.report += rep;
    for (int v = 0; v < nv; ++v)
        if (valid[v] && check[v] != check[first]) {
            ipfe_set_error("NTT autotune: variant %s disagrees with %s", ntt_variant_name(logn, v),
                           ntt_variant_name(logn, first));
            return -1;
        }
    slot = pick;
    return pick;
}

std::string ntt_autotune_report()
{
    Tuned &T = tuned();
    std::lock_guard<std::mutex> lk(T.mu);
    return T.report;
}

}  // namespace ipfe
