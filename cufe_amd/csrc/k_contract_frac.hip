// k_contract_frac.hip -- the coefficient-0 decryption of many rows against many
// keys as ONE int8 matrix-core contraction over fixed-point fractions of the
// CRT value, exact through a per-pair check.
//
// For row r and key v the limbs' residues d_j = sum_k A_j[k] B[k] mod q_j
// (A_j: the row's residues in limb j, B: the key's integer coefficients, the
// same in every limb) give x = CRT(d_0, d_1, d_2) in [0, Q) and the key value
// round(p x / Q) (src/crt.c:46-82, src/rlwe_sife.c:288-332).  With
// u_j = (Q / q_j)^-1 mod q_j, x / Q = sum_j (d_j u_j mod q_j) / q_j mod 1, so
// with alpha[k] = sum_j (A_j[k] u_j mod q_j) / q_j mod 1 -- the CRT value of
// the row's column k over Q --
//     x / Q = sum_k alpha[k] B[k]   (mod 1).
// Each alpha[k] is stored rounded to 72 bits as nine signed base-256 digits,
// alpha~ = sum_{d=1..9} a_d 256^-d with a_2..a_9 in [-128, 127] and a_1 taken
// mod 256 (only alpha mod 1 matters), each key value as two digits,
// B = b_0 + 256 b_1 (|B| <= 32,512).  Then
//     2^72 sum_k alpha~ B = sum_{d, e} P_de 2^(8 (9 - d + e)),
//     P_de = sum_k a_d[k] b_e[k]  (one int8 GEMM each),
// and mod 2^72 the pair (d, e) = (1, 1), of weight 2^72, drops out: 17 digit
// GEMMs (9 with e = 0, 8 with e = 1) instead of the 22 of the three limbs'
// exact contractions (4 x 2 + 4 x 2 + 3 x 2), in one launch that writes key
// values and no residue planes.  Products of equal weight share an accumulator
// C_c, c = d - e - 1 in 0..8: |C_c| <= 2 K 2^14 < 2^31 for K <= 2^15.
//
// The epilogue forms V = sum_c C_c 2^(8 (8 - c)) mod 2^72 exactly (as two
// 64-bit words), so V / 2^72 = sum_k alpha~[k] B[k] mod 1, and the true x / Q lies
// within M_v 2^-72 of it, M_v = ceil(s / 2) + ceil(s / 2^16) + 1 with
// s = sum_k |B_v[k]| (the rounding of alpha~, below).  The key
// value floor((p V + 2^71) / 2^72) is round(p x / Q) unless that interval
// holds a point where the rounding changes:
//   * a half-integer of p x / Q: |(p V mod 2^72) - 2^71| <= p M_v;
//   * the wrap of x / Q from 1 to 0, where the reference's value jumps from p
//     to 0 (V < M_v or V > 2^72 - M_v): a zero inner product whose noise is
//     below about 2^36.
// Such a pair is listed with its V, and k_frac_exact first refines it to 88
// bits: each row also keeps the 16 bits of alpha below the nine digits as one
// int16 per column, D[k] = alpha~88[k] - 2^16 alpha~72[k] in [-2^15, 2^15)
// (alpha~72 is alpha~88 rounded to 72 bits), so
//     V88 = 2^16 V + sum_k D[k] B[k]   (mod 2^88)
// is 2^88 sum_k alpha~88 B mod 1, within M88_v = ceil(s / 2) + ceil(s / 2^7) + 1
// of 2^88 x / Q, and the same two tests at 88 bits settle all but the pairs
// whose x / Q lies within about 2^-66 of a rounding change (at a wrap, noise
// below about 2^20).  Those few are computed from the limbs' residues exactly
// (64-bit sums of the balanced row and key values).  At a wrap only the sign
// of x' (x or x - Q, |x'| <= 2 M88_v Q 2^-88 < 2^60 for M88_v < 2^61) is open,
// and limbs 1 and 2 fix it: x' mod q_1 q_2 (~2^62).
//
// The margins: alpha itself is known to 3 2^-97 (below), alpha~88 rounds it
// to the nearest 2^-88 and alpha~72 rounds alpha~88 to the nearest 2^-72, so
// |alpha~72 - alpha| <= 2^-73 + 2^-89 + 3 2^-97 and M_v = ceil(s / 2) +
// ceil(s / 2^16) + 1 bounds s times that in units of 2^-72.
#include <hip/hip_runtime.h>
#include <stdio.h>

#include <algorithm>

#include "crt_dev.h"
#include "i8_frag.h"
#include "kernels_predict.h"

namespace ipfe {

constexpr int FRAC_PLANES = 9;  // row digits a_1..a_9 (plane d - 1)
constexpr int FRAC_PAIRS = 17;
constexpr int FRAC_ACC = 9;

// Pair i of a step's issue order: plane (d - 1) and key digit e.  Plane 0
// once (e = 0), then each plane p = 1..8 twice, e alternating so that two
// consecutive MFMAs never add into the same accumulator c = plane - e.
__host__ __device__ constexpr int frac_plane(int i) { return i == 0 ? 0 : (i + 1) / 2; }
__host__ __device__ constexpr int frac_kdig(int i)
{
    return i == 0 ? 0 : ((((i & 1) == 1) == ((frac_plane(i) & 1) == 1)) ? 0 : 1);
}
__host__ __device__ constexpr int frac_acc(int i) { return frac_plane(i) - frac_kdig(i); }
// the pair after which plane p is no longer used in a step
__host__ __device__ constexpr int frac_last_use(int p) { return p == 0 ? 0 : 2 * p; }

constexpr bool frac_order_ok()
{
    bool seen[FRAC_PLANES][2] = {};
    for (int i = 0; i < FRAC_PAIRS; ++i) {
        const int p = frac_plane(i), e = frac_kdig(i);
        if (p < 0 || p >= FRAC_PLANES || e < 0 || e > 1 || seen[p][e] || (p == 0 && e == 1)) return false;
        seen[p][e] = true;
        if (i > 0 && frac_acc(i) == frac_acc(i - 1)) return false;
        for (int j = i + 1; j < FRAC_PAIRS; ++j)
            if (frac_plane(j) == p && j > frac_last_use(p)) return false;
    }
    return true;
}
static_assert(frac_order_ok(), "each (plane, key digit) pair once, (0, 1) left out, used planes read after use");

// ---------------------------------------------------------------------------
// Row digits.  One thread per (row, 4 consecutive columns of the ksteps x 32):
// the words of the three limbs (c_L from cl [R][3][N], coefficient 0 of
// c_0..c_{src_l-1} from c0 [R][src_l][3]; words in [0, q], q taken as 0),
// alpha = sum_j r_j / q_j mod 1 with r_j = w_j u_j mod q_j, formed as
// sum_j r_j floor(2^128 / q_j) mod 2^128 (error below 3 2^-97), rounded to
// 88 bits (alpha~88), that rounded to 72 bits (alpha~72) and split into nine
// digits.  Store: [(cap + 128) / 32 row blocks][ksteps][9 planes][64 lanes]
// [16 B], the MFMA fragment layout; the difference D = alpha~88 -
// 2^16 alpha~72 in [-2^15, 2^15) to ref [cap][ksteps x 32] (int16, row-major).
// ---------------------------------------------------------------------------
__global__ void k_rows_frac(const uint32_t *__restrict__ cl, const uint32_t *__restrict__ c0, int src_l, int R, int r0,
                            int ksteps, FracConsts fc, int8_t *__restrict__ dig, int16_t *__restrict__ ref)
{
    const int quads = ksteps * 8;
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= (size_t)R * quads) return;
    const int rr = (int)(i / quads), row = r0 + rr, k = 4 * (int)(i % quads);
    int8_t a[4][FRAC_PLANES];
    int16_t dref[4];
#pragma unroll
    for (int u = 0; u < 4; ++u) {
        const int col = k + u;
        u128 acc = 0;
#pragma unroll
        for (int j = 0; j < 3; ++j) {
            const uint32_t q = qmod(j);
            uint32_t w = 0;
            if (col < IPFE_N)
                w = cl[((size_t)rr * 3 + j) * IPFE_N + col];
            else if (col - IPFE_N < src_l)
                w = c0[((size_t)rr * src_l + (col - IPFE_N)) * 3 + j];
            if (w >= q) w -= q;
            const uint32_t r = (uint32_t)(((uint64_t)w * fc.u[j]) % q);
            acc += (u128)r * (((u128)fc.rhi[j] << 64) | fc.rlo[j]);
        }
        const u128 t88 = (acc + ((u128)1 << 39)) >> 40;  // alpha~88, units of 2^-88
        u128 t = (t88 + ((u128)1 << 15)) >> 16;          // alpha~72 (2^72 taken as 0 below)
        dref[u] = (int16_t)(int)(__int128)(t88 - (t << 16));
#pragma unroll
        for (int d = FRAC_PLANES - 1; d >= 1; --d) {
            const int v = (((int)(t & 255) + 128) & 255) - 128;
            a[u][d] = (int8_t)v;
            t = (t - (u128)(__int128)v) >> 8;
        }
        a[u][0] = (int8_t)(uint8_t)(t & 255);
    }
    const int ks = k >> 5, kk = k & 31;
    const int ln = (kk >> 4) * 32 + (row & 31), j = kk & 15;
    int8_t *base = dig + ((size_t)(row >> 5) * ksteps + ks) * FRAC_PLANES * 1024 + ln * 16 + j;
#pragma unroll
    for (int p = 0; p < FRAC_PLANES; ++p) {
        uint32_t word = 0;
#pragma unroll
        for (int u = 0; u < 4; ++u) word |= (uint32_t)(uint8_t)a[u][p] << (8 * u);
        *reinterpret_cast<uint32_t *>(base + (size_t)p * 1024) = word;
    }
    const uint64_t dw = (uint64_t)(uint16_t)dref[0] | ((uint64_t)(uint16_t)dref[1] << 16) |
                        ((uint64_t)(uint16_t)dref[2] << 32) | ((uint64_t)(uint16_t)dref[3] << 48);
    *reinterpret_cast<uint64_t *>(ref + (size_t)row * (ksteps * 32) + k) = dw;
}

// Keys: whether every key coefficient is the same integer in the three limbs
// (kd [3][S][K] balanced; *neq set otherwise), the margins M_v = ceil(s / 2) +
// ceil(s / 2^16) + 1 (72 bits) and M88_v = ceil(s / 2) + ceil(s / 2^7) + 1
// (88 bits) with s = sum_k |B_v[k]|, each times 2^widen up to 2^62 (widen > 0
// only to test the exact kernel on many pairs), and limb 0's values as int16
// k16 [S][K] for the exact kernel (two key digits: |B| <= 32,512).
__global__ void k_keys_frac(const double *__restrict__ kd, int S, int K, int widen, uint64_t *__restrict__ margin,
                            uint64_t *__restrict__ margin88, int16_t *__restrict__ k16, int *__restrict__ neq)
{
    __shared__ unsigned long long part[256];
    const int v = blockIdx.x;
    unsigned long long sum = 0;
    bool bad = false;
    for (int k = threadIdx.x; k < K; k += blockDim.x) {
        const double b0 = kd[(size_t)v * K + k], b1 = kd[((size_t)S + v) * K + k], b2 = kd[((size_t)2 * S + v) * K + k];
        bad = bad || b0 != b1 || b0 != b2 || fabs(b0) > 32767.0;
        sum += (unsigned long long)fabs(b0);
        k16[(size_t)v * K + k] = (int16_t)b0;
    }
    if (bad) atomicOr(neq, 1);
    part[threadIdx.x] = sum;
    __syncthreads();
    for (int h = blockDim.x / 2; h > 0; h >>= 1) {
        if ((int)threadIdx.x < h) part[threadIdx.x] += part[threadIdx.x + h];
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        const unsigned long long s = part[0], cap = 1ull << 62;
        const unsigned long long m = (s + 1) / 2 + (s + (1ull << 16) - 1) / (1ull << 16) + 1;
        const unsigned long long m88 = (s + 1) / 2 + (s + 127) / 128 + 1;
        margin[v] = m > (cap >> widen) ? cap : m << widen;
        margin88[v] = m88 > (cap >> widen) ? cap : m88 << widen;
    }
}

// V = sum_c C_c 2^(8 (8 - c)) mod 2^72 for output i of a wave's accumulators,
// as lo + 2^64 hi: equal-width pairs first in 64-bit integers (|C_c| < 2^31,
// so every partial sum stays below 2^57), V = q0 2^40 + q1 2^8 + C_8, whose
// 64-bit words are the wrapped low words plus, above them, the arithmetic
// high parts q0 >> 24, q1 >> 56, C_8 >> 63 and the two carries.
static_assert(FRAC_ACC == 9, "the merge tree below has nine accumulators");
__device__ __forceinline__ void frac_value(const v16i (&acc)[FRAC_ACC], int i, uint64_t &lo, uint32_t &hi)
{
    const long long p0 = (long long)acc[0][i] * 256 + acc[1][i];  // weight 2^56
    const long long p1 = (long long)acc[2][i] * 256 + acc[3][i];  // 2^40
    const long long p2 = (long long)acc[4][i] * 256 + acc[5][i];  // 2^24
    const long long p3 = (long long)acc[6][i] * 256 + acc[7][i];  // 2^8
    const long long q0 = p0 * 65536 + p1;                          // 2^40
    const long long q1 = p2 * 65536 + p3;                          // 2^8
    const long long c8 = acc[8][i];
    const uint64_t a = (uint64_t)q0 << 40, b = (uint64_t)q1 << 8, s1 = a + b, s2 = s1 + (uint64_t)c8;
    lo = s2;
    hi = ((uint32_t)(q0 >> 24) + (uint32_t)(q1 >> 56) + (uint32_t)(c8 >> 63) + (uint32_t)(s1 < a) +
          (uint32_t)(s2 < s1)) & 255u;
}

// ---------------------------------------------------------------------------
// The contraction.  A block of WR x WK waves computes a (32 WR) x (32 WK)
// tile, one 32 x 32 block per wave with nine accumulators.  Per k step the
// block streams WR x 9 row and WK x 2 key fragments (the first FR waves one
// load more than the others; each wave's wait counts its own loads) from
// global memory into an LDS ring of STAGES steps
// (global_load_lds_dwordx4), as in k_contract_i8_il: the 17 MFMAs of a step
// are interleaved by hand with the wait and barrier for the next slot (after
// MFMA 1), the next ring loads (after MFMA 2), the next step's key fragments
// (after MFMA 3, double-buffered) and its row fragments, each read into the
// same registers right after its plane's last MFMA (one register set).
// ---------------------------------------------------------------------------
template <int WR, int WK, int STAGES, int OCC>
__global__ __launch_bounds__(64 * WR * WK, OCC) void k_contract_frac(const v4i *__restrict__ A,
                                                                   const v4i *__restrict__ B, int R, int S,
                                                                   int ksteps, long long a_stride, long long b_stride,
                                                                   FracOut out, int ntk, int ntiles, int gr, int kg,
                                                                   int key_fast)
{
    constexpr int W = WR * WK;
    constexpr int FA = WR * FRAC_PLANES, FB = WK * 2;
    // fragments per step: waves w < FR load FW of them, the others FW - 1
    constexpr int FW = (FA + FB + W - 1) / W, FR = FA + FB - (FW - 1) * W;
    constexpr int STEP_V4 = (FA + FB) * FRAG_V4;
    static_assert(STAGES >= 3, "the main loop waits for the step after next");
    __shared__ v4i ring[STAGES * STEP_V4];
    const int lane = threadIdx.x & 63;
    const int w = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int wr = w / WK, wk = w % WK;
    // one tile per block; block b runs on XCD b mod 8, which takes a
    // contiguous share of the tiles
    const int t = (int)(blockIdx.x % 8) * (int)(gridDim.x / 8) + (int)(blockIdx.x / 8);
    if (t >= ntiles) return;
    const bool wfull = w < FR;  // this wave loads FW fragments per step
    int tr, tk;
    if (key_fast)
        tile_kr(t, ntk, ntiles / ntk, gr, kg, tr, tk);
    else
        tile_rc(t, ntk, ntiles / ntk, gr, kg, tr, tk);
    const v4i *base[FW];
    long long step[FW];
#pragma unroll
    for (int u = 0; u < FW; ++u) {
        const int f = w + u * W;
        if (f < FA) {
            base[u] = A + (size_t)(tr * WR + f / FRAC_PLANES) * a_stride + (f % FRAC_PLANES) * FRAG_V4 + lane;
            step[u] = FRAC_PLANES * FRAG_V4;
        } else if (f < FA + FB) {
            const int fb = f - FA;
            base[u] = B + (size_t)(tk * WK + fb / 2) * b_stride + (fb % 2) * FRAG_V4 + lane;
            step[u] = 2 * FRAG_V4;
        } else {
            base[u] = B;  // a wave's last slot past the step's fragments: never loaded
            step[u] = 0;
        }
    }
    const int G = ksteps;
    int it_ks = 0, it_slot = 0;  // next step to load
    auto issue_next = [&]() {
        v4i *slot = ring + it_slot * STEP_V4;
#pragma unroll
        for (int u = 0; u < FW; ++u) {
            const int f = w + u * W;
            if (u < FW - 1 || wfull)
                __builtin_amdgcn_global_load_lds((glb_void_t *)(base[u] + (size_t)it_ks * step[u]),
                                                 (lds_void_t *)(slot + f * FRAG_V4), 16, 0, 0);
        }
        it_slot = it_slot + 1 == STAGES ? 0 : it_slot + 1;
        ++it_ks;
    };
    const int a_off = wr * FRAC_PLANES * FRAG_V4 + lane;
    const int b_off = (FA + wk * 2) * FRAG_V4 + lane;
    v16i acc[FRAC_ACC];
#pragma unroll
    for (int c = 0; c < FRAC_ACC; ++c) acc[c] = v16i{};
#pragma unroll
    for (int p = 0; p < STAGES; ++p)
        if (p < G) issue_next();
    if (G < STAGES)
        wait_vmcnt<0>();
    else if (wfull)
        wait_vmcnt<(STAGES - 1) * FW>();
    else
        wait_vmcnt<(STAGES - 1) * (FW - 1)>();
    asm volatile("s_barrier" ::: "memory");
    v4i fa[FRAC_PLANES], fb0[2], fb1[2];
#pragma unroll
    for (int p = 0; p < FRAC_PLANES; ++p) fa[p] = ring[a_off + p * FRAG_V4];
#pragma unroll
    for (int e = 0; e < 2; ++e) fb0[e] = ring[b_off + e * FRAG_V4];
    int rd_slot = 1 % STAGES;  // slot of step g + 1
    // One step g.  In the last step the reads of "step G" fetch a slot no one
    // loads any more: in bounds, and never used.  Every wave runs the same
    // number of barriers.
    auto body = [&](auto main, int g, v4i (&fb)[2], v4i (&nb)[2]) {
        constexpr bool M = decltype(main)::value;
        const v4i *nx = ring + rd_slot * STEP_V4;
#pragma unroll
        for (int i = 0; i < FRAC_PAIRS; ++i) {
            const int c = frac_acc(i);
            acc[c] = __builtin_amdgcn_mfma_i32_32x32x32_i8(fa[frac_plane(i)], fb[frac_kdig(i)], acc[c], 0, 0, 0);
            __builtin_amdgcn_sched_barrier(0);
            if (i == 1) {
                if constexpr (!M)
                    wait_vmcnt<0>();
                else if (wfull)
                    wait_vmcnt<(STAGES - 2) * FW>();
                else
                    wait_vmcnt<(STAGES - 2) * (FW - 1)>();
                asm volatile("s_barrier" ::: "memory");
            } else if (i == 2) {
                if constexpr (M)
                    issue_next();
                else if (g + STAGES < G)
                    issue_next();
            } else if (i == 3) {
#pragma unroll
                for (int e = 0; e < 2; ++e) nb[e] = nx[b_off + e * FRAG_V4];
            }
            // next step's planes, each after its last use here (and after
            // the barrier that makes the next slot visible)
#pragma unroll
            for (int p = 0; p < FRAC_PLANES; ++p)
                if (i == (frac_last_use(p) < 2 ? 2 : frac_last_use(p))) fa[p] = nx[a_off + p * FRAG_V4];
            __builtin_amdgcn_sched_barrier(0);
        }
        rd_slot = rd_slot + 1 == STAGES ? 0 : rd_slot + 1;
    };
    using Main = std::integral_constant<bool, true>;
    using Tail = std::integral_constant<bool, false>;
    int g = 0;
    // main: steps g and g + 1 both still load step + STAGES
    for (; g + 1 + STAGES < G; g += 2) {
        body(Main{}, g, fb0, fb1);
        body(Main{}, g + 1, fb1, fb0);
    }
    for (; g < G; g += 2) {
        body(Tail{}, g, fb0, fb1);
        if (g + 1 < G) body(Tail{}, g + 1, fb1, fb0);
    }
    // epilogue: lane (row0 + 8 (i / 4) + i mod 4, key) for i < 16.  Every
    // output's key value is stored unless it is near a rounding change; those
    // are marked in amb (bit i; bit 16 + i: near a half-integer) and listed
    // afterwards by the waves that have any (rare), so the common path has no
    // atomics and no branch per output.
    //
    // In 64-bit words: V = hi 2^64 + lo (hi < 2^8), p V = T_hi 2^64 + T_lo
    // (T_hi < 2^24), the key value floor((p V + 2^71) / 2^72) =
    // (T_hi + 128) >> 8.  The wrap test is exact (M < 2^64): V < M is hi = 0
    // and lo < M, V > 2^72 - M is hi = 255 and 2^64 - lo < M.  The half-integer
    // test compares the top 40 bits of p V mod 2^72 (units of 2^32) with the
    // margin rounded outwards: p V mod 2^72 - 2^71 = 2^32 D + r with
    // D = (T_hi mod 2^8 - 128) 2^32 + (T_lo >> 32) and 0 <= r < 2^32, so
    // |p V mod 2^72 - 2^71| <= p M implies |D| <= floor(p M / 2^32) + 2.  The
    // pairs this wider test adds lie within 2^34 of the exact window (valid
    // ciphertexts put p V far from 2^71), and are settled exactly like the
    // others.
    const int key = tk * (32 * WK) + wk * 32 + (lane & 31);
    if (key >= S) return;
    const int row0 = tr * (32 * WR) + wr * 32 + (lane >> 5) * 4;
    const uint64_t mv = out.margin[key];
    const long long pmu = (long long)(((u128)IPFE_P * mv) >> 32);
    uint32_t amb = 0;
#pragma unroll
    for (int i = 0; i < 16; ++i) {
        const int row = row0 + (i >> 2) * 8 + (i & 3);
        uint64_t lo;
        uint32_t hi;
        frac_value(acc, i, lo, hi);
        const uint64_t t = (uint64_t)IPFE_P * (uint32_t)lo;                 // p lo_0
        const uint64_t u = (uint64_t)IPFE_P * (uint32_t)(lo >> 32) + (t >> 32);  // p lo_1 + carry
        const uint32_t thi = IPFE_P * hi + (uint32_t)(u >> 32);
        const uint32_t tlo_hi = (uint32_t)u;  // bits 32..63 of T_lo
        const long long D = ((long long)(thi & 255u) - 128) * 4294967296ll + tlo_hi;
        const bool halfway = D <= pmu + 2 && D >= -pmu - 2;
        const bool wrap = (hi == 0u && lo < mv) || (hi == 255u && lo != 0 && 0 - lo < mv);
        const bool near = wrap || halfway;
        if (row < R && !near) out.kval[(size_t)row * S + key] = (thi + 128u) >> 8;
        amb |= (row < R && near ? 1u << i : 0u) | (row < R && halfway ? 1u << (16 + i) : 0u);
    }
    if (__ballot(amb != 0) == 0) return;
#pragma unroll
    for (int i = 0; i < 16; ++i) {
        if (!((amb >> i) & 1u)) continue;
        // bit 31 of the key word: near a half-integer (else only the wrap);
        // then V's 72 bits for the refinement
        uint64_t lo;
        uint32_t hi;
        frac_value(acc, i, lo, hi);
        const uint32_t n = atomicAdd(out.amb_count, 1u);
        if ((int)n < out.amb_cap) {
            uint32_t *e = out.amb + (size_t)FRAC_AMB_WORDS * n;
            e[0] = (uint32_t)(row0 + (i >> 2) * 8 + (i & 3));
            e[1] = (uint32_t)key | (((amb >> (16 + i)) & 1u) ? 0x80000000u : 0u);
            e[2] = (uint32_t)lo;
            e[3] = (uint32_t)(lo >> 32);
            e[4] = hi;
        }
    }
}

// The listed pairs.  One wave per pair.  With refine set, it first forms
// V88 = 2^16 V + sum_k D[r][k] B_v[k] mod 2^88 (ref: the rows' D values, row
// stride kf; 16-byte loads of eight columns) and writes the key value when
// neither test fires at 88 bits.  Otherwise -- and for every pair without
// refine, a switch for tests -- it sums A_j[r][k] B_v[k] over k in 64-bit
// integers (|A| < 2^30, |B| < 2^15, K < 2^13: below 2^58) and reduces mod
// q_j.  A pair open only at the wrap whose margin allows it (M88_v < 2^61, or
// without refine M_v < 2^44) needs limbs 1 and 2: x' = the centred value of
// x mod q_1 q_2 (Garner), key value 0 for x' >= 0 and p below; the others
// decode all three limbs.  A: the balanced row store [3][cap][ldr] at the
// chunk's first row, ref the D store at the same row; k16: the key values
// [S][K] (the same integers in every limb).  A count above the list's
// capacity sets *overflow (those pairs' key values are then missing).
//
// Each lane takes groups of four (exact sums) or eight (refinement) columns,
// eight groups a lane in flight; K, kf, ldr and the limb stride are multiples
// of 8.  The listed pairs rarely share a row, so the exact sums read a row's
// 34 KB (two limbs) or 51 KB (three) from HBM, the refinement its 8.4 KB of D.
template <int LIMBS>
__device__ __forceinline__ void frac_exact_sums(const int4 *__restrict__ a0, const int4 *__restrict__ a1,
                                                const int4 *__restrict__ a2, const int2 *__restrict__ b, int K4,
                                                int lane, long long &s0, long long &s1, long long &s2)
{
    auto dot4 = [](const int4 v, const int2 w) {
        return (long long)v.x * (int)(short)w.x + (long long)v.y * (w.x >> 16) + (long long)v.z * (int)(short)w.y +
               (long long)v.w * (w.y >> 16);
    };
#pragma unroll 8
    for (int g = lane; g < K4; g += 64) {
        const int2 w = b[g];
        s1 += dot4(a1[g], w);
        s2 += dot4(a2[g], w);
        if (LIMBS == 3) s0 += dot4(a0[g], w);
    }
}

// sum of eight int16 products of two 16-byte groups (|sum| < 2^33)
__device__ __forceinline__ long long dot8_i16(const int4 x, const int4 y)
{
    auto two = [](int u, int v) { return (long long)((int)(short)u * (int)(short)v) + (long long)((u >> 16) * (v >> 16)); };
    return two(x.x, y.x) + two(x.y, y.y) + two(x.z, y.z) + two(x.w, y.w);
}

__global__ __launch_bounds__(256) void k_frac_exact(const int32_t *__restrict__ A, size_t limb_stride, int ldr,
                                                    const int16_t *__restrict__ ref, int kf,
                                                    const int16_t *__restrict__ k16, int S, int K,
                                                    const uint64_t *__restrict__ margin,
                                                    const uint64_t *__restrict__ margin88, int refine, uint32_t inv12,
                                                    const uint32_t *__restrict__ amb,
                                                    const uint32_t *__restrict__ amb_count, int amb_cap,
                                                    uint32_t *__restrict__ kval, int *__restrict__ overflow,
                                                    unsigned long long *__restrict__ total)
{
    const uint32_t cnt = *amb_count;
    const int n = (int)std::min<uint32_t>(cnt, (uint32_t)amb_cap);
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        if (cnt > (uint32_t)amb_cap) *overflow = 1;
        if (total) atomicAdd(total, (unsigned long long)cnt);
    }
    const int lane = threadIdx.x & 63;
    const int nw = (int)(gridDim.x * blockDim.x / 64);
    for (int i = (int)((blockIdx.x * blockDim.x + threadIdx.x) / 64); i < n; i += nw) {
        const uint32_t *e = amb + (size_t)FRAC_AMB_WORDS * i;
        const int row = (int)e[0];
        const uint32_t kw = e[1];
        const int key = (int)(kw & 0x7fffffffu);
        bool two;
        if (refine) {
            const int4 *d8 = reinterpret_cast<const int4 *>(ref + (size_t)row * kf);
            const int4 *b8 = reinterpret_cast<const int4 *>(k16 + (size_t)key * K);
            long long r = 0;
#pragma unroll 8
            for (int g = lane; g < (K >> 3); g += 64) r += dot8_i16(d8[g], b8[g]);
#pragma unroll
            for (int o = 32; o > 0; o >>= 1) r += __shfl_xor(r, o);
            // every lane holds the sum: the tests below are the same on all
            const u128 mask88 = ((u128)1 << 88) - 1, half = (u128)1 << 87, one = (u128)1 << 88;
            const u128 V = (u128)e[2] | ((u128)e[3] << 32) | ((u128)(e[4] & 255u) << 64);
            const u128 V88 = ((V << 16) + (u128)(__int128)r) & mask88;
            const uint64_t mv = margin88[key];
            const u128 pv = (u128)IPFE_P * V88, fr = pv & mask88, pm = (u128)IPFE_P * mv;
            const bool wrap = V88 < mv || V88 > one - mv;
            const bool halfway = fr + pm >= half && fr <= half + pm;
            if (!wrap && !halfway) {
                if (lane == 0) kval[(size_t)row * S + key] = (uint32_t)((pv + half) >> 88);
                continue;
            }
            two = !halfway && mv < (1ull << 61);
        } else {
            two = !(kw >> 31) && margin[key] < (1ull << 44);
        }
        const int2 *b = reinterpret_cast<const int2 *>(k16 + (size_t)key * K);
        const int32_t *r0p = A + (size_t)row * ldr;
        const int4 *a0 = reinterpret_cast<const int4 *>(r0p);
        const int4 *a1 = reinterpret_cast<const int4 *>(r0p + limb_stride);
        const int4 *a2 = reinterpret_cast<const int4 *>(r0p + 2 * limb_stride);
        long long s0 = 0, s1 = 0, s2 = 0;
        if (two)
            frac_exact_sums<2>(a0, a1, a2, b, K >> 2, lane, s0, s1, s2);
        else
            frac_exact_sums<3>(a0, a1, a2, b, K >> 2, lane, s0, s1, s2);
#pragma unroll
        for (int o = 32; o > 0; o >>= 1) {
            s0 += __shfl_xor(s0, o);
            s1 += __shfl_xor(s1, o);
            s2 += __shfl_xor(s2, o);
        }
        if (lane == 0) {
            auto canon = [](long long x, long long q) {
                const long long m = x % q;
                return (uint32_t)(m < 0 ? m + q : m);
            };
            const uint32_t q1 = qmod(1), q2 = qmod(2);
            const uint32_t d1 = canon(s1, q1), d2 = canon(s2, q2);
            uint32_t kv;
            if (two) {
                // x mod q1 q2 = d1 + q1 t, t = (d2 - d1) q1^-1 mod q2
                const uint64_t t = (uint64_t)((d2 + (uint64_t)q2 - d1 % q2) % q2) * inv12 % q2;
                const uint64_t x12 = d1 + (uint64_t)q1 * t, m12 = (uint64_t)q1 * q2;
                kv = x12 > m12 / 2 ? IPFE_P : 0u;
            } else {
                kv = crt_decode_canon(canon(s0, qmod(0)), d1, d2);
            }
            kval[(size_t)row * S + key] = kv;
        }
    }
}

// ---------------------------------------------------------------------------
// Host side
// ---------------------------------------------------------------------------
static uint32_t powmod(uint64_t a, uint64_t e, uint64_t q)
{
    uint64_t r = 1;
    a %= q;
    while (e) {
        if (e & 1) r = r * a % q;
        a = a * a % q;
        e >>= 1;
    }
    return (uint32_t)r;
}

FracConsts frac_consts()
{
    FracConsts fc{};
    for (int j = 0; j < 3; ++j) {
        const uint64_t q = qmod(j);
        uint64_t m = 1;  // Q / q_j mod q_j
        for (int i = 0; i < 3; ++i)
            if (i != j) m = m * (qmod(i) % q) % q;
        fc.u[j] = powmod(m, q - 2, q);
        const u128 r = ~(u128)0 / q;  // floor(2^128 / q): q is odd
        fc.rlo[j] = (uint64_t)r;
        fc.rhi[j] = (uint64_t)(r >> 64);
    }
    return fc;
}

size_t frac_digit_bytes(int cap, int ksteps) { return (size_t)(cap + 128) / 32 * ksteps * FRAC_PLANES * 1024; }

int launch_rows_frac(const uint32_t *cl, const uint32_t *c0, int src_l, int R, int r0, int cap, int ksteps,
                     int8_t *dig, int16_t *ref, hipStream_t st)
{
    const size_t n = (size_t)R * ksteps * 8;
    if (n == 0) return 0;
    if (r0 < 0 || r0 + R > cap || ksteps * 32 < IPFE_N + src_l) {
        ipfe_set_error("fraction digits: rows [%d, %d) of %d, %d k steps for %d slots", r0, r0 + R, cap, ksteps, src_l);
        return -1;
    }
    static const FracConsts fc = frac_consts();
    hipLaunchKernelGGL(k_rows_frac, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, cl, c0, src_l, R, r0, ksteps,
                       fc, dig, ref);
    HIP_OK(hipGetLastError());
    return 0;
}

int launch_keys_frac(const double *kd, int S, int K, int widen, uint64_t *margin, uint64_t *margin88, int16_t *k16,
                     int *neq, hipStream_t st)
{
    if (S <= 0) return 0;
    if (widen < 0 || widen > 62) {
        ipfe_set_error("fraction keys: widening 2^%d out of range", widen);
        return -1;
    }
    hipLaunchKernelGGL(k_keys_frac, dim3((unsigned)S), dim3(256), 0, st, kd, S, K, widen, margin, margin88, k16,
                       neq);
    HIP_OK(hipGetLastError());
    return 0;
}

// Block tiles of the fraction contraction: waves, ring depth, waves per SIMD
// the registers are budgeted for, tile order (gr x kg groups, kg = 0:
// row-major; key_fast: the key tile fastest in a group).  A row tile carries
// 9 digit planes per column, a key tile 2, so the orders that keep a row
// tile's blocks together fetch the least (the row-fastest 8 x 8 groups of
// 64 x 64 tiles fetched twice the bytes of row-major order and ran slowest).
// An XCD runs 32 of the 8-wave blocks at a time on consecutive tiles: for
// 64 x 128 tiles, sub-blocks of 4 row x 8 key tiles balance the row digits
// re-read per sub-block of key tiles against the key digits re-read per
// sub-block of row tiles, and fetch the least of that shape.
struct FracVariant {
    int wr, wk, stages, occ, gr, kg, key_fast;
};
static const FracVariant kFracVariants[] = {
    {2, 2, 3, 2, 1, 0, 0},  {1, 4, 3, 2, 16, 4, 0}, {1, 4, 4, 2, 16, 4, 0}, {2, 4, 3, 2, 8, 4, 0},
    {4, 2, 3, 2, 4, 8, 0},  {4, 2, 3, 2, 1, 0, 0},  {2, 2, 3, 2, 2, 32, 1}, {1, 4, 3, 2, 4, 16, 1},
    {2, 4, 3, 2, 2, 16, 1}, {4, 2, 3, 2, 2, 16, 1}, {2, 4, 3, 2, 4, 8, 1}, {2, 4, 3, 2, 4, 8, 0},
};
constexpr int kNumFracVariants = sizeof(kFracVariants) / sizeof(kFracVariants[0]);

int frac_variant_count() { return kNumFracVariants; }
void frac_shape(int v, int *wr, int *wk)
{
    *wr = kFracVariants[v].wr;
    *wk = kFracVariants[v].wk;
}
// "fWRxWK/STAGES" plus "gGRxKG" for the grouped tile order ("kGRxKG" with
// the key tile fastest)
const char *frac_variant_label(int v)
{
    struct Labels {
        char s[kNumFracVariants][24];
        Labels()
        {
            for (int i = 0; i < kNumFracVariants; ++i) {
                const FracVariant &x = kFracVariants[i];
                if (x.kg)
                    snprintf(s[i], sizeof s[i], "f%dx%d/%d%c%dx%d", x.wr, x.wk, x.stages, x.key_fast ? 'k' : 'g', x.gr,
                             x.kg);
                else
                    snprintf(s[i], sizeof s[i], "f%dx%d/%d", x.wr, x.wk, x.stages);
            }
        }
    };
    static const Labels labels;
    return v >= 0 && v < kNumFracVariants ? labels.s[v] : "?";
}

int launch_contract_frac(const int8_t *A, int ksteps_a, const int8_t *B, int ksteps, int R, int S, const FracOut &out,
                         hipStream_t st, int variant)
{
    if (R <= 0 || S <= 0) return 0;
    if (variant < 0 || variant >= kNumFracVariants || ksteps < 1 || ksteps > ksteps_a || ksteps * 32 > (1 << 15)) {
        ipfe_set_error("fraction contraction: variant %d, %d k steps (row store %d)", variant, ksteps, ksteps_a);
        return -1;
    }
    const FracVariant &x = kFracVariants[variant];
    // row tiles stay within the 128 padding rows, key tiles within the
    // 128-key padding
    const int row_tile = 32 * x.wr, key_tile = 32 * x.wk;
    const int ntk = (S + key_tile - 1) / key_tile, ntr = (R + row_tile - 1) / row_tile, ntiles = ntk * ntr;
    const int gr = x.kg ? x.gr : 1, kg = x.kg ? std::min(x.kg, ntk) : ntk;
    const dim3 grid((unsigned)((ntiles + 7) / 8 * 8)), block(64 * x.wr * x.wk);
    const v4i *a = reinterpret_cast<const v4i *>(A);
    const v4i *b = reinterpret_cast<const v4i *>(B);
    const long long as = (long long)ksteps_a * FRAC_PLANES * FRAG_V4, bs = (long long)ksteps * 2 * FRAG_V4;
#define CFR(wr, wk, stg, occ)                                                                                      \
    hipLaunchKernelGGL((k_contract_frac<wr, wk, stg, occ>), grid, block, 0, st, a, b, R, S, ksteps, as, bs, out, ntk, \
                       ntiles, gr, kg, x.key_fast)
    if (x.wr == 2 && x.wk == 2)
        CFR(2, 2, 3, 2);
    else if (x.wr == 1 && x.stages == 3)
        CFR(1, 4, 3, 2);
    else if (x.wr == 1)
        CFR(1, 4, 4, 2);
    else if (x.wr == 2)
        CFR(2, 4, 3, 2);
    else
        CFR(4, 2, 3, 2);
#undef CFR
    HIP_OK(hipGetLastError());
    return 0;
}

int launch_frac_exact(const int32_t *A, size_t limb_stride, int ldr, const int16_t *ref, int kf, bool refine,
                      const int16_t *k16, int S, int K, const FracOut &out, int *overflow, unsigned long long *total,
                      hipStream_t st)
{
    static const uint32_t inv12 = powmod(qmod(1) % qmod(2), qmod(2) - 2, qmod(2));
    // 16-byte loads: every row, limb and key starts on a multiple of 8
    // elements, in 16-byte aligned stores
    if (K % 8 || ldr % 8 || kf % 8 || limb_stride % 8 || K > ldr || K > kf || ((uintptr_t)A & 15) ||
        ((uintptr_t)k16 & 15) || ((uintptr_t)ref & 15)) {
        ipfe_set_error("fraction exact kernel: %d key columns, row strides %d and %d, limb stride %zu", K, ldr, kf,
                       limb_stride);
        return -1;
    }
    hipLaunchKernelGGL(k_frac_exact, dim3(2048), dim3(256), 0, st, A, limb_stride, ldr, ref, kf, k16, S, K, out.margin,
                       out.margin88, refine ? 1 : 0, inv12, out.amb, out.amb_count, out.amb_cap, out.kval, overflow,
                       total);
    HIP_OK(hipGetLastError());
    return 0;
}

}  // namespace ipfe
