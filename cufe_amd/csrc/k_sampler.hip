// k_sampler.hip -- discrete Gaussian sampling on gfx950, bit-exact with
// gaussian_sampler_S1_ori / S2_ori / S3 (src/gauss.c:705-785, :915-993,
// :1086-1163) on the AES-256-CTR stream of src/aes256ctr.c.
//
// The reference draws sample after sample from one sequential stream.  Each
// 384-byte refill ("chunk", 24 AES blocks) yields 8 candidates; a call takes
// accepted candidates in order until it has 4096, drops the rest of its last
// chunk, and the next call on the stream starts on the next chunk.  Here that
// becomes three data-parallel passes:
//   1. candidates: every chunk of a speculative window, in parallel: AES (LDS
//      T-tables), uniform/CDT/Bernoulli per candidate, wave ballot -> per-chunk
//      accept count and the accepted values compacted in order;
//   2. walk: one wave per stream turns counts into call boundaries with
//      1024-chunk windows and a wave prefix scan, and labels every chunk with
//      (call, first output position);
//   3. scatter: every accepted candidate of a used chunk writes its three
//      residues to its call's polynomial.
#include <stdlib.h>

#include <algorithm>
#include <type_traits>

#include "aes_dev.h"
#include "kernels.h"

namespace ipfe {

// Sampler constants (src/gauss.c:11-45): k per width, the Barrett factor
// floor(2^32/k), the uniform acceptance bound k*floor(2^32/k) and -1/k^2.
// Widths 1-3 are the reference's S1/S2/S3; 4 and 5 are the S2/S3 widths of the
// N = 2048 parameter set (the same algorithm with another k).
#define IPFE_SAMPLER_K {1ull, 266ull, 304214978ull, 608429953ull, 262147ull, 524287ull}
constexpr uint64_t kSamplerK[6] = IPFE_SAMPLER_K;
__constant__ uint64_t c_K[6] = IPFE_SAMPLER_K;
__constant__ uint64_t c_BF[6] = {0ull, (1ull << 32) / kSamplerK[1], (1ull << 32) / kSamplerK[2],
                                 (1ull << 32) / kSamplerK[3], (1ull << 32) / kSamplerK[4],
                                 (1ull << 32) / kSamplerK[5]};
__constant__ uint64_t c_UQ[6] = {0ull, kSamplerK[1] * ((1ull << 32) / kSamplerK[1]),
                                 kSamplerK[2] * ((1ull << 32) / kSamplerK[2]),
                                 kSamplerK[3] * ((1ull << 32) / kSamplerK[3]),
                                 kSamplerK[4] * ((1ull << 32) / kSamplerK[4]),
                                 kSamplerK[5] * ((1ull << 32) / kSamplerK[5])};
__constant__ double c_K2INV[6] = {0.0, -1.0 / (double)(long long)(kSamplerK[1] * kSamplerK[1]),
                                  -1.0 / (double)(long long)(kSamplerK[2] * kSamplerK[2]),
                                  -1.0 / (double)(long long)(kSamplerK[3] * kSamplerK[3]),
                                  -1.0 / (double)(long long)(kSamplerK[4] * kSamplerK[4]),
                                  -1.0 / (double)(long long)(kSamplerK[5] * kSamplerK[5])};

// Residue of a candidate on limb s of the stream's parameter set (src/gauss.c:773-779).
__device__ __forceinline__ uint32_t stream_residue(const GaussStream *S, int s, uint64_t z, bool neg)
{
    const uint32_t q = S->q[s], a = S->fold[s] & 255u, b = S->fold[s] >> 8;
    uint32_t m = mod_prime_gen(z, q, a, b);
    if (neg) m = mod_prime_gen((uint32_t)(q - m), q, a, b);
    return m;
}

// CDT rows (low word, high word), src/gauss.c:62-72.
__constant__ uint64_t c_CDT[9][2] = {{2200310400551559144ull, 3327841033070651387ull},
                                     {7912151619254726620ull, 380075531178589176ull},
                                     {5167367257772081627ull, 11604843442081400ull},
                                     {5081592746475748971ull, 90134450315532ull},
                                     {6522074513864805092ull, 175786317361ull},
                                     {2579734681240182346ull, 85801740ull},
                                     {8175784047440310133ull, 10472ull},
                                     {2947787991558061753ull, 0ull},
                                     {22489665999543ull, 0ull}};

// Coefficients of the 2^a polynomial, src/gauss.c:82-103 (IEEE-754 bit patterns).
__constant__ uint64_t c_EXP[10] = {0x3e833b70ffa2c5d4ull, 0x3eb4a480fda7e6e1ull, 0x3ef01b254493363full,
                                   0x3f242e0e0aa273ccull, 0x3f55d8a2334ed31bull, 0x3f83b2aa56db0f1aull,
                                   0x3fac6b08e11fc57eull, 0x3fcebfbdff556072ull, 0x3fe62e42fefa7fe6ull,
                                   0x3ff0000000000000ull};

extern const uint32_t *g_aes_te;  // device T-tables, set by the context

constexpr int CAND_THREADS = 256;
constexpr int TILE_CHUNKS = GAUSS_TILE_CHUNKS;
constexpr int CHUNK_WORDS = 96;

__device__ __forceinline__ int find_stream(const long long *prefix, int n, long long key)
{
    int lo = 0, hi = n - 1;
    while (lo < hi) {
        int mid = (lo + hi + 1) >> 1;
        if (prefix[mid] <= key) lo = mid; else hi = mid - 1;
    }
    return lo;
}

// Bernoulli decision (src/gauss.c:204-264 and the S2/S3 copies): the
// reference's exact sequence of IEEE double operations and integer bit tricks.
// rb points at the 36 Bernoulli bytes of the candidate's group.
__device__ __forceinline__ bool bernoulli_bit(uint64_t b, double k2inv, const uint32_t *rb, int ln)
{
    double vx = __longlong_as_double((long long)(b | 0x4330000000000000ull)) - 4503599627370496.0;
    vx = __dmul_rn(vx, k2inv);
    double fl = floor(vx);
    double sh = __dadd_rn(fl, 6755399441055744.0);
    uint64_t t = ((uint64_t)__double_as_longlong(sh) - 0x4338000000000000ull) << 52;
    double a = __dsub_rn(vx, fl);
    double s = __fma_rn(__longlong_as_double((long long)c_EXP[0]), a, __longlong_as_double((long long)c_EXP[1]));
#pragma unroll
    for (int i = 2; i < 10; ++i) s = __fma_rn(s, a, __longlong_as_double((long long)c_EXP[i]));
    uint64_t res = t + (uint64_t)__double_as_longlong(s);
    uint64_t res_mant = (res & ((1ull << 52) - 1)) | (1ull << 52);
    uint64_t e = (res >> 52) + (uint64_t)(long long)-1003;
    uint64_t res_exp = (e < 64) ? (1ull << e) : 0ull;
    uint64_t rm = ((uint64_t)rb[2 * ln + 1] << 32) | rb[2 * ln];
    uint64_t r_exp = (rm >> 53) | ((uint64_t)((rb[8] >> (8 * ln)) & 255u) << 11);
    uint64_t r_mant = rm & ((1ull << 53) - 1);
    uint64_t bits = (r_mant - res_mant) & (r_exp - res_exp);
    return (bits >> 63) != 0 || res == 0x3FF0000000000000ull;
}

// One candidate of a chunk held in LDS (96 little-endian words).  Lanes 8t..8t+7
// of a wave evaluate candidates 0..7 of one chunk together: the uniform words
// are shared through ballots, so every lane of the wave must call this.
struct Cand {
    uint64_t z;
    uint32_t sgn;
    bool acc;
};

// AES blocks of a chunk computed by default: 0..15 (bytes 0..255) and 17
// (bytes 272..287, holding the sign byte 280).  Candidates read bytes 0..199
// (CDT and Bernoulli words), the uniform words from byte 200 and byte 280; the
// uniform scan stays below byte 256 unless it needs more than 14 words (seven
// or more rejections, about 1e-11 per chunk).  Such a chunk sets status bit 4
// and the caller reruns the group with all 24 blocks.
constexpr int PART_BLOCKS = 17;
__device__ __forceinline__ int chunk_block(int j, int full) { return (full || j < 16) ? j : 17; }

// Sign word (bytes 280..283) read from the chunk in LDS.
struct SignInLds {
    __device__ uint32_t operator()(const uint32_t *ch) const { return ch[70]; }
};

// ulim: uniform words a partial chunk may scan (14 covers its blocks; lower
// values exist to exercise the rerun).  sign_word(ch) returns bytes 280..283
// of the chunk, which hold the candidate's sign bit; with sign_if_zero it is
// asked only for a zero candidate (the only one whose acceptance depends on
// its sign), and the sign of every other candidate is reported as 0.
template <typename SignWord = SignInLds>
__device__ __forceinline__ Cand eval_candidate(const uint32_t *ch, int c, int which, bool valid, int *status,
                                               int full, int ulim, SignWord sign_word = SignWord(),
                                               bool sign_if_zero = false)
{
    const uint64_t K = c_K[which], BF = c_BF[which];
    const uint32_t UQ = (uint32_t)c_UQ[which];
    const double K2INV = c_K2INV[which];
    const int sh = (threadIdx.x & 63) & 56;
    // uniform_sampler (src/gauss.c:609-703): the c-th word below UQ from byte 200 on
    const uint32_t w0 = ch[50 + c], w1 = ch[58 + c], w2 = ch[66 + c];
    const uint64_t m0 = __ballot(w0 < UQ), m1 = __ballot(w1 < UQ), m2 = __ballot(w2 < UQ);
    const uint32_t mask = (uint32_t)((m0 >> sh) & 0xffu) | ((uint32_t)((m1 >> sh) & 0xffu) << 8) |
                          ((uint32_t)((m2 >> sh) & 0xffu) << 16);
    int pos = c;
    if ((mask & 0xffu) != 0xffu) {
        int seen = 0;
        bool partial_miss = false;
        pos = -1;
        for (int u = 0; u < 46; ++u) {
            if (!full && u >= ulim) {
                partial_miss = true;  // word 64 on lies in a block not computed
                break;
            }
            const bool kept = (u < 24) ? ((mask >> u) & 1u) : (ch[50 + u] < UQ);
            if (kept) {
                if (seen == c) { pos = u; break; }
                ++seen;
            }
        }
        if (pos < 0) {
            if (valid) atomicOr(status, partial_miss ? 4 : 2);
            pos = 0;
        }
    }
    const uint64_t w = ch[50 + pos];
    uint64_t y = w - ((((w * BF) >> 32) + 1) * K);
    y = y + (y >> 63) * K;

    // cdt_sampler (src/gauss.c:122-154)
    const int g = c >> 2, ln = c & 3;
    const uint32_t *gb = ch + 25 * g;
    const uint64_t r1 = (((uint64_t)gb[2 * ln + 1] << 32) | gb[2 * ln]) & 0x7fffffffffffffffull;
    const uint64_t r2 = (((uint64_t)gb[8 + 2 * ln + 1] << 32) | gb[8 + 2 * ln]) & 0x7fffffffffffffffull;
    uint32_t xs = 0;
#pragma unroll
    for (int i = 0; i < 9; ++i) xs += (r2 < c_CDT[i][1]) | ((r2 == c_CDT[i][1]) & (r1 < c_CDT[i][0]));

    const uint64_t xk = (uint64_t)xs * (uint64_t)(uint32_t)K;
    Cand out;
    out.z = xk + y;
    const uint64_t bin = (uint64_t)(uint32_t)(out.z + xk) * (uint64_t)(uint32_t)y;
    const bool bit = bernoulli_bit(bin, K2INV, gb + 16, ln);
    out.sgn = (!sign_if_zero || out.z == 0) ? (sign_word(ch) >> c) & 1u : 0u;
    out.acc = bit && (out.z != 0 || out.sgn != 0);
    return out;
}

// AES of nch chunks of one stream into LDS: chunk k at ks + 96 k uses
// counters 24 (chunk_of(k)) .. +23.
// nb: blocks per chunk -- 24 (all), PART_BLOCKS (0..15 and 17) or 16 (0..15,
// for streams whose values are not stored: the sign byte then matters only for
// a zero candidate, and that block is computed where it is needed).
// TM: the T-table layout in te -- 0: four 1 KiB tables; 1: T0 replicated per
// bank (AesTabRep); 2: T0 and T2 replicated per bank (AesTabRep2).
// win (TM 1 and 2; may be null): the counter windows w0, w0 + 1, ... of
// aes256_ctr_window, 8 words each, covering every counter.
template <int TM>
__device__ __forceinline__ void aes_block(const GaussStream *S, const uint32_t *te, const uint32_t *win, uint32_t w0,
                                          uint32_t ctr, uint32_t (&o)[4])
{
    if constexpr (TM == 0) {
        aes256_ctr_block(S->rk, te, S->nonce, ctr, o);
    } else {
        using Tab = std::conditional_t<TM == 2, AesTabRep2, AesTabRep>;
        const Tab tb(te);
        if (win)
            aes256_ctr_block_win(tb, S->rk, win + 8 * ((ctr >> 8) - w0), ctr, o);
        else
            aes256_ctr_block_t(tb, S->rk, S->nonce, ctr, o);
    }
}

// cw: words per chunk in ks (CHUNK_WORDS, or fewer when only the first blocks
// are stored).
template <int TM, int NT, typename ChunkOf>
__device__ __forceinline__ void aes_chunks(const GaussStream *S, const uint32_t *te, int nch, ChunkOf chunk_of,
                                           uint32_t *ks, int cw, int full, int nb, const uint32_t *win, uint32_t w0)
{
    for (int g = threadIdx.x; g < nch * nb; g += NT) {
        const int k = g / nb, blk = chunk_block(g - nb * k, full);
        const uint32_t ctr = (uint32_t)(chunk_of(k) * 24 + blk);
        uint32_t o[4];
        aes_block<TM>(S, te, win, w0, ctr, o);
        uint32_t *d = ks + k * cw + blk * 4;
        d[0] = o[0]; d[1] = o[1]; d[2] = o[2]; d[3] = o[3];
    }
}

// TM (aes_block): the four 1 KiB T-tables, T0 replicated per bank (32 KiB of
// LDS; NT = 512, so that two workgroups per CU hold 16 waves) or T0 and T2
// replicated (64 KiB).  With the 64 KiB tables two workgroups fit a CU's LDS
// when the keystream buffer holds 64 words per chunk: streams that need only
// blocks 0..15 run a tile in one pass; the others run two passes of 32 chunks
// (72 words per chunk for blocks 0..15 and the sign word, 96 for all 24
// blocks).
template <int TM, int NT>
__global__ __launch_bounds__(NT) void k_gauss_candidates(const GaussStream *__restrict__ streams, int nstreams,
                                                                    const long long *__restrict__ tile_prefix,
                                                                    long long ntiles,
                                                                    const uint32_t *__restrict__ te_global,
                                                                    uint8_t *__restrict__ cnt, uint64_t *__restrict__ vals,
                                                                    int *__restrict__ status, int full, int ulim)
{
    // windows of 256 counters one tile spans (its 24 TILE_CHUNKS counters)
    constexpr int WINDOWS = TILE_CHUNKS * 24 / 256 + 2;
    constexpr int KS_WORDS = TM == 2 ? TILE_CHUNKS * 64 : TILE_CHUNKS * CHUNK_WORDS;
    constexpr int HALF = 32;  // chunks per pass of the two-pass tiles
    static_assert(TM != 2 || (HALF * CHUNK_WORDS <= KS_WORDS && 2 * HALF >= TILE_CHUNKS && NT >= HALF * 9 &&
                              TILE_CHUNKS * 8 <= NT),
                  "pass sizes of the two-table kernel");
    __shared__ uint32_t te[TM == 2 ? AES_REP2_WORDS : TM == 1 ? AES_REP_WORDS : AES_TABLE_WORDS];
    __shared__ uint32_t ks[KS_WORDS];
    __shared__ uint32_t win[TM ? WINDOWS * 8 : 8];
    if (TM == 2)
        aes_stage_tables_rep2(te, te_global);
    else if (TM == 1)
        aes_stage_tables_rep(te, te_global);
    else
        aes_stage_tables(te, te_global);
    const int sh = (threadIdx.x & 63) & 56;
    for (long long tile = blockIdx.x; tile < ntiles; tile += gridDim.x) {
        const int si = __builtin_amdgcn_readfirstlane(find_stream(tile_prefix, nstreams, tile));
        const GaussStream *S = streams + si;
        const long long chunk0 = (tile - tile_prefix[si]) * TILE_CHUNKS;
        const int nch = (int)min((long long)TILE_CHUNKS, (long long)S->nchunks - chunk0);
        const int which = S->which;
        const bool store_vals = !S->coef0_only;
        // without stored values only a zero candidate needs its sign (a zero
        // with sign 0 is rejected, src/gauss.c:1086-1163)
        const bool lazy_sign = !store_vals && !full;
        __syncthreads();  // tables staged / previous tile's reads of ks and win done
        // counter windows of the tile (unless its counters wrap at 2^32)
        const uint32_t c_lo = (uint32_t)(chunk0 * 24), w0 = c_lo >> 8;
        const bool use_win = TM && chunk0 * 24 + (long long)nch * 24 <= (1ll << 32);
        const int nw = (int)((c_lo + (uint32_t)(nch * 24 - 1)) >> 8) - (int)w0 + 1;
        if constexpr (TM != 0) {
            using Tab = std::conditional_t<TM == 2, AesTabRep2, AesTabRep>;
            if (use_win && (int)threadIdx.x < nw)
                aes256_ctr_window(Tab(te), S->rk, S->nonce, w0 + threadIdx.x, win + 8 * threadIdx.x);
        }
        // The two-table kernel's partial chunks of stored streams: a pass of
        // HALF chunks computes blocks 0..15 only (HALF 16 = 512 blocks); the
        // HALF 8 evaluating lanes take every sign as 1 while the next HALF
        // threads, otherwise idle, compute block 17 -- the sign word -- of each
        // chunk, and after a barrier the evaluating lanes settle sign and
        // acceptance.
        const bool split_sign = TM == 2 && !full && store_vals;
        // chunks per pass and words per chunk in ks
        const int sub = (TM == 2 && (full || store_vals)) ? HALF : TILE_CHUNKS;
        const int cw = TM != 2 ? CHUNK_WORDS : full ? CHUNK_WORDS : store_vals ? 72 : 64;
        for (int h = 0; h < nch; h += sub) {
            __syncthreads();  // windows written / the previous pass's reads of ks done
            const long long c0 = chunk0 + h;
            const int nsub = min(sub, nch - h);
            aes_chunks<TM, NT>(S, te, nsub, [c0](int k) { return c0 + k; }, ks, cw, full,
                               full ? 24 : (lazy_sign || split_sign) ? 16 : PART_BLOCKS, use_win ? win : nullptr, w0);
            __syncthreads();
            // ballot of the accept bits and the chunk's count and values
            auto commit = [&](const Cand &cd, int k, int c, bool valid) {
                const uint64_t am = __ballot(cd.acc);
                const uint32_t gm = (uint32_t)((am >> sh) & 0xffu);
                if (valid) {
                    const long long gc = S->chunk_base + c0 + k;
                    if (c == 0) cnt[gc] = (uint8_t)__popc(gm);
                    if (store_vals && cd.acc)
                        vals[gc * 8 + __popc(gm & ((1u << c) - 1u))] = cd.z | ((uint64_t)cd.sgn << 63);
                }
            };
            if (split_sign) {
                const int t = threadIdx.x, k = t >> 3, c = t & 7;
                const bool eval = t < HALF * 8;
                Cand cd{};
                if (eval) {
                    cd = eval_candidate(ks + k * cw, c, which, k < nsub, status, full, ulim,
                                        [](const uint32_t *) { return ~0u; });
                } else if (t < HALF * 8 + nsub) {
                    const int kb = t - HALF * 8;
                    uint32_t o[4];
                    aes_block<TM>(S, te, use_win ? win : nullptr, w0, (uint32_t)((c0 + kb) * 24 + 17), o);
                    ks[kb * cw + 70] = o[2];
                }
                __syncthreads();
                if (eval) {
                    // with every sign 1, cd.acc was the Bernoulli bit alone
                    cd.sgn = (ks[k * cw + 70] >> c) & 1u;
                    cd.acc = cd.acc && (cd.z != 0 || cd.sgn != 0);
                    commit(cd, k, c, k < nsub);
                }
                continue;
            }
            for (int slot = threadIdx.x; slot < sub * 8; slot += NT) {
                const int k = slot >> 3, c = slot & 7;
                // with lazy_sign, block 17 of chunk k is computed by the asking
                // lane alone (a zero candidate: about 1e-9 of S2/S3 candidates)
                auto sign_word = [&](const uint32_t *ch) -> uint32_t {
                    if (!lazy_sign) return ch[70];
                    uint32_t o[4];
                    aes_block<TM>(S, te, nullptr, 0, (uint32_t)((c0 + k) * 24 + 17), o);
                    return o[2];
                };
                commit(eval_candidate(ks + k * cw, c, which, k < nsub, status, full, ulim, sign_word, lazy_sign),
                       k, c, k < nsub);
            }
        }
    }
}

// Compact streams: sample 0 of every call.  Item t (a call of a compact
// stream) has its first non-empty chunk in firsts[t]; the chunk is recomputed
// and its first accepted candidate written as 3 residues at out0 + 3 call.
constexpr int FIRST_TILE = 32;

__global__ __launch_bounds__(CAND_THREADS) void k_gauss_firsts(const GaussStream *__restrict__ streams, int nstreams,
                                                                const long long *__restrict__ first_prefix,
                                                                const long long *__restrict__ firsts, long long nitems,
                                                                const uint32_t *__restrict__ te_global,
                                                                int *__restrict__ status, int full, int ulim)
{
    __shared__ uint32_t te[AES_TABLE_WORDS];
    __shared__ uint32_t ks[FIRST_TILE * CHUNK_WORDS];
    __shared__ int s_stream[FIRST_TILE];
    __shared__ long long s_chunk[FIRST_TILE];
    aes_stage_tables(te, te_global);
    const long long t0 = (long long)blockIdx.x * FIRST_TILE;
    if (threadIdx.x < FIRST_TILE) {
        const long long t = t0 + threadIdx.x;
        int si = -1;
        long long ch = 0;
        if (t < nitems) {
            si = find_stream(first_prefix, nstreams, t);
            ch = firsts[t] - streams[si].chunk_base;
        }
        s_stream[threadIdx.x] = si;
        s_chunk[threadIdx.x] = ch;
    }
    __syncthreads();
    // AES: 24 blocks per item, each item possibly from a different stream
    const int nb = full ? 24 : PART_BLOCKS;
    for (int g = threadIdx.x; g < FIRST_TILE * nb; g += CAND_THREADS) {
        const int k = g / nb, blk = chunk_block(g - nb * k, full);
        const int si = s_stream[k];
        if (si < 0) continue;
        const GaussStream *S = streams + si;
        uint32_t o[4];
        aes256_ctr_block(S->rk, te, S->nonce, (uint32_t)(s_chunk[k] * 24 + blk), o);
        uint32_t *d = ks + k * CHUNK_WORDS + blk * 4;
        d[0] = o[0]; d[1] = o[1]; d[2] = o[2]; d[3] = o[3];
    }
    __syncthreads();
    const int sh = (threadIdx.x & 63) & 56;
    const int k = threadIdx.x >> 3, c = threadIdx.x & 7;  // 32 items x 8 candidates = 256 threads
    const int si = s_stream[k];
    const bool valid = si >= 0;
    const int which = valid ? streams[si].which : 1;
    const Cand cd = eval_candidate(ks + k * CHUNK_WORDS, c, which, valid, status, full, ulim);
    const uint64_t am = __ballot(cd.acc && valid);
    const uint32_t gm = (uint32_t)((am >> sh) & 0xffu);
    if (valid && cd.acc && (gm & ((1u << c) - 1u)) == 0) {
        const GaussStream *S = streams + si;
        const long long call = t0 + k - first_prefix[si];
        uint32_t *out = S->out0 + (long long)S->nmod * call;
        for (int s = 0; s < S->nmod; ++s) out[s] = stream_residue(S, s, cd.z, cd.sgn != 0);
    }
}

// One wave per stream.  Window of 64 lanes x 16 chunks.  Full streams label
// every used chunk with (call << 13 | first output position); compact streams
// record only the first non-empty chunk of each call in firsts.
__global__ __launch_bounds__(64) void k_gauss_walk(const GaussStream *__restrict__ streams, const uint8_t *__restrict__ cnt,
                                                    int *__restrict__ pos, long long *__restrict__ firsts,
                                                    int *__restrict__ status)
{
    const GaussStream *S = streams + blockIdx.x;
    const int lane = threadIdx.x;
    const long long base = S->chunk_base;
    const int nchunks = S->nchunks, ncalls = S->ncalls;
    const bool compact = S->coef0_only != 0;
    int c = 0, call = 0, filled = 0;
    bool first_done = false;
    while (call < ncalls) {
        if (c >= nchunks) {
            if (lane == 0) status[blockIdx.x] = 1;
            return;
        }
        uint32_t v[16];
        uint32_t lsum = 0;
#pragma unroll
        for (int r = 0; r < 16; ++r) {
            const int ci = c + 16 * lane + r;
            v[r] = (ci < nchunks) ? cnt[base + ci] : 0u;
            lsum += v[r];
        }
        // inclusive wave scan of lsum
        uint32_t incl = lsum;
#pragma unroll
        for (int d = 1; d < 64; d <<= 1) {
            uint32_t o = __shfl_up(incl, d, 64);
            if (lane >= d) incl += o;
        }
        const uint32_t excl = incl - lsum;
        const uint32_t total = __shfl(incl, 63, 64);
        if (compact && !first_done) {
            const uint64_t nz = __ballot(lsum > 0);
            if (nz) {
                const int fl = __ffsll((unsigned long long)nz) - 1;
                if (lane == fl) {
                    int rr = 0;
                    while (rr < 15 && v[rr] == 0) ++rr;
                    firsts[S->first_base + call] = base + c + 16 * lane + rr;
                }
                first_done = true;
            }
        }
        const uint32_t need = (uint32_t)S->slen - (uint32_t)filled;
        const uint64_t hit = __ballot(incl >= need);
        int e_chunk = -1;  // last chunk of this call, relative to c
        if (hit) {
            const int hl = __ffsll((unsigned long long)hit) - 1;
            if (lane == hl) {
                uint32_t run = excl;
#pragma unroll
                for (int r = 0; r < 16; ++r) {
                    run += v[r];
                    if (run >= need && e_chunk < 0) e_chunk = 16 * lane + r;
                }
            }
            e_chunk = __shfl(e_chunk, hl, 64);
        }
        if (!compact) {
            const int last = (e_chunk >= 0) ? e_chunk : 1023;
            uint32_t run = excl;
#pragma unroll
            for (int r = 0; r < 16; ++r) {
                const int rel = 16 * lane + r;
                const int ci = c + rel;
                if (rel <= last && ci < nchunks) pos[base + ci] = (call << 13) | (int)((uint32_t)filled + run);
                run += v[r];
            }
        }
        if (e_chunk >= 0) {
            ++call;
            filled = 0;
            first_done = false;
            c += e_chunk + 1;
        } else {
            filled += (int)total;
            c += 1024;
        }
    }
}

__global__ void k_gauss_scatter(const GaussStream *__restrict__ streams, int nstreams,
                                const long long *__restrict__ chunk_prefix, long long nchunks_total,
                                const uint8_t *__restrict__ cnt, const uint64_t *__restrict__ vals,
                                const int *__restrict__ pos)
{
    // The 8 chunks of a wave (64 consecutive t) nearly always belong to one
    // stream: it is found once per wave from the wave's first chunk, with
    // wave-uniform loads, and per lane only in a wave that crosses a stream.
    auto one = [&](const GaussStream *S, long long gc, int slot) {
        const int info = pos[gc];
        if (info < 0) return;
        const int n = cnt[gc];
        const int call = info >> 13, off = info & 0x1fff;
        if (slot >= n || off + slot >= S->slen) return;
        const uint64_t v = vals[gc * 8 + slot];
        const uint64_t z = v & 0x7fffffffffffffffull;
        const bool neg = (v >> 63) != 0;
        // output residues (src/gauss.c:773-779)
        uint32_t *out = ((call == 0) ? S->out0 : S->out1 + (long long)(call - 1) * S->out_stride) + off + slot;
        for (int s = 0; s < S->nmod; ++s) out[(long long)s * S->slen] = stream_residue(S, s, z, neg);
    };
    for (long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x; t < nchunks_total * 8;
         t += (long long)gridDim.x * blockDim.x) {
        const long long gc = t >> 3;
        const int slot = (int)(t & 7);
        // the lowest t of the wave is its first active lane's
        const long long gc0 = ((long long)__builtin_amdgcn_readfirstlane((int)(gc >> 32)) << 32) |
                              (uint32_t)__builtin_amdgcn_readfirstlane((int)(uint32_t)gc);
        const int sw = __builtin_amdgcn_readfirstlane(find_stream(chunk_prefix, nstreams, gc0));
        const long long end = sw + 1 < nstreams ? chunk_prefix[sw + 1] : nchunks_total;
        if (__all(gc < end))
            one(streams + sw, gc, slot);
        else
            one(streams + find_stream(chunk_prefix, nstreams, gc), gc, slot);
    }
}

__global__ void k_fill_i32(int *p, int v, size_t n)
{
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) p[i] = v;
}

// sample_polya (src/sample.c:56-128): coefficient i = LE64(keystream[8i..8i+8))
// mod q_j, keystream of AES-256-CTR(seed, nonce 0) from counter 0.  Any
// parameter set: n coefficients on nmod limbs, a [nmod][n].
struct PolyaArgs {
    int n, nmod;
    uint32_t q[4];
};
__global__ void k_sample_polya(const uint32_t *__restrict__ rk, const uint32_t *__restrict__ te_global,
                               uint32_t *__restrict__ a, PolyaArgs pa)
{
    __shared__ uint32_t te[AES_TABLE_WORDS];
    aes_stage_tables(te, te_global);
    __syncthreads();
    const uint32_t blk = blockIdx.x * blockDim.x + threadIdx.x;  // n / 2 blocks
    if (blk >= (uint32_t)pa.n / 2) return;
    uint32_t o[4];
    aes256_ctr_block(rk, te, 0u, blk, o);
#pragma unroll
    for (int h = 0; h < 2; ++h) {
        const uint64_t v = ((uint64_t)o[2 * h + 1] << 32) | o[2 * h];
        const uint32_t i = 2 * blk + h;
        for (int s = 0; s < pa.nmod; ++s) a[(size_t)s * pa.n + i] = (uint32_t)(v % pa.q[s]);
    }
}

__global__ void k_aes_ctr(const uint32_t *__restrict__ rk, const uint32_t *__restrict__ te_global, uint32_t nonce,
                          uint32_t ctr0, uint32_t nblocks, uint32_t *__restrict__ out)
{
    __shared__ uint32_t te[AES_TABLE_WORDS];
    aes_stage_tables(te, te_global);
    __syncthreads();
    const uint32_t b = blockIdx.x * blockDim.x + threadIdx.x;
    if (b >= nblocks) return;
    uint32_t o[4];
    aes256_ctr_block(rk, te, nonce, ctr0 + b, o);
    out[4 * b + 0] = o[0]; out[4 * b + 1] = o[1]; out[4 * b + 2] = o[2]; out[4 * b + 3] = o[3];
}

// ---------------------------------------------------------------------------
// 0: four T-tables, 256 threads; 1: T0 replicated, 512 threads; 2: T0
// replicated, 256 threads; 3 (default): T0 and T2 replicated, 512 threads,
// keystream in half-tile passes (CUFE_AES_REP)
static int aes_rep_mode()
{
    const char *e = getenv("CUFE_AES_REP");
    return e && *e ? atoi(e) : 3;
}

// Uniform words a partial chunk may scan before it asks for all 24 blocks:
// 14 (all its blocks hold); CUFE_GAUSS_SCAN in [8, 14] lowers it so tests can
// drive the rerun path.
static int scan_limit()
{
    const char *e = getenv("CUFE_GAUSS_SCAN");
    const int v = e && *e ? atoi(e) : 14;
    return v < 8 ? 8 : (v > 14 ? 14 : v);
}

int launch_gauss_candidates(const GaussStream *d_streams, int nstreams, const long long *d_tile_prefix,
                            long long ntiles, uint8_t *cnt, uint64_t *vals, int *status, int full, hipStream_t st)
{
    if (ntiles <= 0) return 0;
    const unsigned grid = (unsigned)std::min<long long>(ntiles, 1ll << 20);
    const int rep = aes_rep_mode();
    if (rep >= 3)
        hipLaunchKernelGGL((k_gauss_candidates<2, 512>), dim3(grid), dim3(512), 0, st, d_streams, nstreams,
                           d_tile_prefix, ntiles, g_aes_te, cnt, vals, status, full, scan_limit());
    else if (rep == 2)
        hipLaunchKernelGGL((k_gauss_candidates<1, 256>), dim3(grid), dim3(256), 0, st, d_streams, nstreams,
                           d_tile_prefix, ntiles, g_aes_te, cnt, vals, status, full, scan_limit());
    else if (rep == 1)
        hipLaunchKernelGGL((k_gauss_candidates<1, 512>), dim3(grid), dim3(512), 0, st, d_streams, nstreams,
                           d_tile_prefix, ntiles, g_aes_te, cnt, vals, status, full, scan_limit());
    else
        hipLaunchKernelGGL((k_gauss_candidates<0, CAND_THREADS>), dim3(grid), dim3(CAND_THREADS), 0, st,
                           d_streams, nstreams, d_tile_prefix, ntiles, g_aes_te, cnt, vals, status, full,
                           scan_limit());
    HIP_OK(hipGetLastError());
    return 0;
}

int launch_gauss_walk(const GaussStream *d_streams, int nstreams, const uint8_t *cnt, int *pos, long long *firsts,
                      int *status, hipStream_t st)
{
    if (nstreams <= 0) return 0;
    hipLaunchKernelGGL(k_gauss_walk, dim3(nstreams), dim3(64), 0, st, d_streams, cnt, pos, firsts, status);
    HIP_OK(hipGetLastError());
    return 0;
}

int launch_gauss_scatter(const GaussStream *d_streams, int nstreams, const long long *d_chunk_prefix,
                         long long nchunks_total, const uint8_t *cnt, const uint64_t *vals, const int *pos,
                         hipStream_t st)
{
    if (nchunks_total <= 0) return 0;
    const long long nthreads = nchunks_total * 8;
    const unsigned grid = (unsigned)std::min<long long>((nthreads + 255) / 256, 1ll << 20);
    hipLaunchKernelGGL(k_gauss_scatter, dim3(grid), dim3(256), 0, st, d_streams, nstreams, d_chunk_prefix,
                       nchunks_total, cnt, vals, pos);
    HIP_OK(hipGetLastError());
    return 0;
}

int launch_gauss_firsts(const GaussStream *d_streams, int nstreams, const long long *d_first_prefix,
                        const long long *firsts, long long nitems, int *status, int full, hipStream_t st)
{
    if (nitems <= 0) return 0;
    hipLaunchKernelGGL(k_gauss_firsts, dim3((unsigned)((nitems + FIRST_TILE - 1) / FIRST_TILE)), dim3(CAND_THREADS), 0,
                       st, d_streams, nstreams, d_first_prefix, firsts, nitems, g_aes_te, status, full, scan_limit());
    HIP_OK(hipGetLastError());
    return 0;
}

int launch_fill_i32(int *p, int v, size_t n, hipStream_t st)
{
    if (n == 0) return 0;
    const unsigned grid = (unsigned)std::min<size_t>((n + 255) / 256, (size_t)1 << 20);
    hipLaunchKernelGGL(k_fill_i32, dim3(grid), dim3(256), 0, st, p, v, n);
    HIP_OK(hipGetLastError());
    return 0;
}

int launch_sample_polya(const uint32_t *d_rk, uint32_t *a, hipStream_t st)
{
    const uint32_t q[3] = {qmod(0), qmod(1), qmod(2)};
    return launch_sample_polya_ps(d_rk, a, IPFE_N, 3, q, st);
}

int launch_sample_polya_ps(const uint32_t *d_rk, uint32_t *a, int n, int nmod, const uint32_t *q, hipStream_t st)
{
    if (n < 512 || (n & (n - 1)) || nmod < 1 || nmod > 4) {
        ipfe_set_error("sample_polya: unsupported shape n=%d nmod=%d", n, nmod);
        return -1;
    }
    PolyaArgs pa{};
    pa.n = n;
    pa.nmod = nmod;
    for (int s = 0; s < nmod; ++s) pa.q[s] = q[s];
    hipLaunchKernelGGL(k_sample_polya, dim3((unsigned)(n / 2 / 256)), dim3(256), 0, st, d_rk, g_aes_te, a, pa);
    HIP_OK(hipGetLastError());
    return 0;
}

int launch_aes_ctr(const uint32_t *d_rk, uint32_t nonce, uint32_t ctr0, uint32_t nblocks, uint8_t *out, hipStream_t st)
{
    if (nblocks == 0) return 0;
    hipLaunchKernelGGL(k_aes_ctr, dim3((nblocks + 255) / 256), dim3(256), 0, st, d_rk, g_aes_te, nonce, ctr0, nblocks,
                       (uint32_t *)out);
    HIP_OK(hipGetLastError());
    return 0;
}

}  // namespace ipfe
