// k_contract_i8.hip -- the coefficient-0 contraction on int8 matrix cores,
// exact.
//
// D[r][v] = sum_k A[r][k] B[v][k] mod q with balanced operands: row values
// |a| <= (q-1)/2 < 2^30 and key values |b| <= 32512 (two key digits) or
// |b| <= 8322943 (three).  Each value is split into signed base-256 digits,
// a = sum_d a_d 256^d (4 digits) and b = sum_e b_e 256^e (NB digits), every
// a_d, b_e in [-128, 127].  Then
//     sum_k a b = sum_e 256^e T_e,   T_e = sum_d 256^d P_de,
//     P_de = sum_k a_d[k] b_e[k],
// and each P_de is one int8 GEMM with exact int32 accumulation:
// |P_de| <= K * 2^14 < 2^31 for K < 2^17.  The epilogue forms each T_e
// exactly in doubles (|T_e| < K 2^38 <= 2^52 for K <= 2^14), reduces it mod q, and
// combines the residues with the weights 256^e < q, so D is the canonical
// residue of the exact integer dot product -- the value the fp64 kernels
// produce.
//
// Digit stores are laid out as MFMA fragments so every load is contiguous:
//   rows, per limb: [row block of 32][k step of 32][digit d][lane][16 B]
//   keys, per limb: [key block of 32][k step of 32][digit e < NB][lane][16 B]
// lane l holds row / key (l mod 32) and k = 16 (l / 32) + j, j < 16 -- the
// operand layout of v_mfma_i32_32x32x32_i8, whose result lane l, register i
// is D[8 (i / 4) + 4 (l / 32) + i mod 4][l mod 32].
#include <stdio.h>

#include <algorithm>
#include <map>
#include <mutex>
#include <type_traits>

#include "crt_dev.h"
#include "i8_frag.h"
#include "kernels_predict.h"

namespace ipfe {

constexpr int DG_A = 4;  // digits per row value (key values: NB = 2 or 3)

// Where a contraction writes: the limb's residues D[R][S], or -- for the last
// limb, with kval set -- the rounded decryption of every pair, combined by CRT
// with the residues d0, d1 of limbs 0 and 1 (what k_crt_kvalue computes from
// three D planes, without writing or re-reading this limb's plane).
struct I8Out {
    uint32_t *D;
    const uint32_t *d0, *d1;
    uint32_t *kval;
};

// Epilogue shared by the kernels: lane (row0, key) holds acc[d][e][i] =
// P_de for row row0 + 8 (i / 4) + i mod 4.  T_e = sum_d 256^d P_de is formed
// in doubles by Horner steps, all exact: |P_de| <= K 2^14, so every partial
// value stays below K 2^38 <= 2^52 (the launcher limits K to 2^14).  Each T_e
// is reduced mod q (the double quotient is off by at most one; fma keeps the
// remainder exact), then the residues are combined with the weights 256^e
// (values below 2^40).
template <int DA, int NB>
__device__ __forceinline__ void i8_store(const v16i (&acc)[DA][NB], int row0, int key, int R, int S, uint32_t q,
                                         const I8Out &out)
{
    const double qd = (double)q, qinv = 1.0 / qd;
    auto reduce = [&](double t) {
        double m = fma(-floor(t * qinv), qd, t);
        m += (m < 0.0) ? qd : 0.0;
        m -= (m >= qd) ? qd : 0.0;
        return m;
    };
#pragma unroll
    for (int i = 0; i < 16; ++i) {
        const int row = row0 + (i >> 2) * 8 + (i & 3);
        if (row >= R) continue;
        double tot = 0.0;
#pragma unroll
        for (int e = NB - 1; e >= 0; --e) {
            double T = (double)acc[DA - 1][e][i];
#pragma unroll
            for (int d = DA - 2; d >= 0; --d) T = fma(T, 256.0, (double)acc[d][e][i]);
            const double m = reduce(T);
            tot = (e == NB - 1) ? m : reduce(fma(tot, 256.0, m));
        }
        const uint32_t res = (uint32_t)tot;
        const size_t idx = (size_t)row * S + key;
        if (out.kval)
            out.kval[idx] = crt_decode_canon(out.d0[idx], out.d1[idx], res);
        else
            out.D[idx] = res;
    }
}

// Epilogue of the merged form: acc[s][i] = C_s = sum over d + e = s of P_de
// (products of equal weight 256^s share one accumulator; |C_s| <= min(DA, NB)
// K 2^14 <= 3 2^28 < 2^31 for K <= 2^14).  sum_s 256^s C_s is formed from the
// top by Horner steps in doubles, reduced mod q after every two steps: a
// reduced value is below 2^31, so two steps stay below 2^48 and the first
// (from |C| < 2^30) below 2^46 -- every value exact.
// LB: outputs whose limb-0 and limb-1 residues are loaded together in the CRT
// epilogue (fewer for kernels budgeted for more waves per SIMD).  With DA = 3
// (limb 0) the epilogue never does the CRT, which runs on limb 2.
template <int DA, int NB, int LB = 8>
__device__ __forceinline__ void i8_store_merged(const v16i (&acc)[DA + NB - 1], int row0, int key, int R, int S,
                                                uint32_t q, const I8Out &out)
{
    constexpr int NS = DA + NB - 1;
    const double qd = (double)q, qinv = 1.0 / qd;
    auto reduce = [&](double t) {
        double m = fma(-floor(t * qinv), qd, t);
        m += (m < 0.0) ? qd : 0.0;
        m -= (m >= qd) ? qd : 0.0;
        return m;
    };
    auto residue = [&](int i) {
        double T = (double)acc[NS - 1][i];
        int steps = 0;
#pragma unroll
        for (int s = NS - 2; s >= 0; --s) {
            T = fma(T, 256.0, (double)acc[s][i]);
            if (++steps == 2 || s == 0) {
                T = reduce(T);
                steps = 0;
            }
        }
        return (uint32_t)T;
    };
    auto index = [&](int i) {  // rows past R read row R - 1 and store nothing
        const int row = min(row0 + (i >> 2) * 8 + (i & 3), R - 1);
        return (size_t)row * S + key;
    };
    auto live = [&](int i) { return row0 + (i >> 2) * 8 + (i & 3) < R; };
    if constexpr (LB == 1) {
        // one output at a time: the fewest live registers
#pragma unroll
        for (int i = 0; i < 16; ++i) {
            if (!live(i)) continue;
            const uint32_t res = residue(i);
            const size_t idx = index(i);
            if (DA == 4 && out.kval)
                out.kval[idx] = crt_decode_canon(out.d0[idx], out.d1[idx], res);
            else
                out.D[idx] = res;
        }
    } else if (DA == 4 && out.kval) {
        // in groups of LB outputs, the 2 LB loads of limbs 0 and 1 are issued
        // before the first is used (one memory wait per group, not per
        // output); the scheduling barrier keeps the groups apart so that the
        // loaded values fit the registers the k loop leaves
        static_assert(16 % LB == 0, "groups of outputs");
#pragma unroll
        for (int h = 0; h < 16; h += LB) {
            uint32_t r0[LB], r1[LB];
#pragma unroll
            for (int i = 0; i < LB; ++i) {
                r0[i] = out.d0[index(h + i)];
                r1[i] = out.d1[index(h + i)];
            }
#pragma unroll
            for (int i = 0; i < LB; ++i) {
                const uint32_t v = crt_decode_canon(r0[i], r1[i], residue(h + i));
                if (live(h + i)) out.kval[index(h + i)] = v;
            }
            __builtin_amdgcn_sched_barrier(0);
        }
    } else {
#pragma unroll
        for (int i = 0; i < 16; ++i) {
            const uint32_t v = residue(i);
            if (live(i)) out.D[index(i)] = v;
        }
    }
}

// ---------------------------------------------------------------------------
// Block tile: WR row blocks x WK key blocks of 32, one wave per 32 x 32 output
// block with DG_A x DG_B int32 accumulators.  Per k step the block stages the
// row digits (WR x 4 KB) and key digits (WK x NB KB) in LDS, double buffered,
// with 16-byte loads spread over the block's threads.  Tiles are ordered so
// that the blocks one XCD receives take consecutive tiles (the key tiles of one
// row tile), sharing the row digits in that XCD's L2.
// ---------------------------------------------------------------------------
template <int WR, int WK, int NB>
constexpr int i8_min_waves()
{
    // three waves per SIMD where the accumulators and prefetch registers fit
    // in 168 VGPRs, else two
    return (NB == 2 && WR * WK <= 4 && (WR * 256 + WK * NB * 64 + 64 * WR * WK - 1) / (64 * WR * WK) <= 3) ? 3 : 2;
}

template <int WR, int WK, int NB>
__global__ __launch_bounds__(64 * WR * WK, (i8_min_waves<WR, WK, NB>())) void k_contract_i8(const v4i *__restrict__ A, const v4i *__restrict__ B,
                                                               int R, int S, int ksteps, long long a_stride,
                                                               long long b_stride, uint32_t q,
                                                               I8Out out, int ntk, int ntiles)
{
    constexpr int NT = 64 * WR * WK;
    constexpr int SA = WR * DG_A * FRAG_V4, SB = WK * NB * FRAG_V4;  // v4i per k step
    constexpr int NSL = (SA + SB + NT - 1) / NT;                       // loads per thread
    __shared__ v4i lds[2][SA + SB];
    const int tid = threadIdx.x, lane = tid & 63;
    const int w = __builtin_amdgcn_readfirstlane(tid >> 6);
    const int per = gridDim.x / 8;
    const int t = (int)(blockIdx.x % 8) * per + (int)(blockIdx.x / 8);
    if (t >= ntiles) return;
    const int tk = t % ntk, tr = t / ntk;
    const int wr = w / WK, wk = w % WK;
    // load slot u of this thread: slot tid + u NT; slots [0, SA) are row digits
    // (row block slot / (DG_A 64)), the rest key digits
    const v4i *src[NSL];
    long long step[NSL];
#pragma unroll
    for (int u = 0; u < NSL; ++u) {
        const int sl = tid + u * NT;
        if (sl < SA) {
            src[u] = A + (size_t)(tr * WR + sl / (DG_A * FRAG_V4)) * a_stride + sl % (DG_A * FRAG_V4);
            step[u] = DG_A * FRAG_V4;
        } else {
            const int sb = sl < SA + SB ? sl - SA : 0;
            src[u] = B + (size_t)(tk * WK + sb / (NB * FRAG_V4)) * b_stride + sb % (NB * FRAG_V4);
            step[u] = NB * FRAG_V4;
        }
    }
    v16i acc[DG_A][NB];
#pragma unroll
    for (int d = 0; d < DG_A; ++d)
#pragma unroll
        for (int e = 0; e < NB; ++e) acc[d][e] = v16i{};
    v4i pre[NSL];
#pragma unroll
    for (int u = 0; u < NSL; ++u) {
        pre[u] = src[u][0];
        if (tid + u * NT < SA + SB) lds[0][tid + u * NT] = pre[u];
    }
    __syncthreads();
    for (int ks = 0; ks < ksteps; ++ks) {
        const int cur = ks & 1;
        const bool more = ks + 1 < ksteps;
        if (more) {
#pragma unroll
            for (int u = 0; u < NSL; ++u) pre[u] = src[u][(size_t)(ks + 1) * step[u]];
        }
        v4i fa[DG_A], fb[NB];
#pragma unroll
        for (int d = 0; d < DG_A; ++d) fa[d] = lds[cur][wr * (DG_A * FRAG_V4) + d * FRAG_V4 + lane];
#pragma unroll
        for (int e = 0; e < NB; ++e) fb[e] = lds[cur][SA + wk * (NB * FRAG_V4) + e * FRAG_V4 + lane];
#pragma unroll
        for (int d = 0; d < DG_A; ++d)
#pragma unroll
            for (int e = 0; e < NB; ++e)
                acc[d][e] = __builtin_amdgcn_mfma_i32_32x32x32_i8(fa[d], fb[e], acc[d][e], 0, 0, 0);
        if (more) {
#pragma unroll
            for (int u = 0; u < NSL; ++u)
                if (tid + u * NT < SA + SB) lds[cur ^ 1][tid + u * NT] = pre[u];
        }
        __syncthreads();
    }
    const int key = tk * (32 * WK) + wk * 32 + (lane & 31);
    if (key >= S) return;
    const int row0 = tr * (32 * WR) + wr * 32 + (lane >> 5) * 4;
    i8_store<DG_A, NB>(acc, row0, key, R, S, q, out);
}

// ---------------------------------------------------------------------------
// Pipelined variant: the digit fragments go from global memory straight into
// an LDS ring of STAGES k steps (global_load_lds_dwordx4: lane l's 16 bytes
// land at the wave's LDS base + 16 l, which is the fragment layout), so loads
// for the next STAGES - 1 k steps are in flight while a step computes, and no
// registers hold them.  Every wave issues the same number of fragment loads
// per step (F = 4 WR + NB WK fragments over WR WK waves), so one
// s_waitcnt vmcnt(loads issued after step ks) and a barrier make step ks
// visible to all waves.
// ---------------------------------------------------------------------------

template <int WR, int WK, int NB, int STAGES, int KPB>
__global__ __launch_bounds__(64 * WR * WK, 2) void k_contract_i8_dma(const v4i *__restrict__ A,
                                                                     const v4i *__restrict__ B, int R, int S,
                                                                     int ksteps, long long a_stride,
                                                                     long long b_stride, uint32_t q,
                                                                     I8Out out, int ntk, int ntiles)
{
    constexpr int W = WR * WK;
    constexpr int F = WR * DG_A + WK * NB;  // fragments per k step
    static_assert(F % W == 0, "fragments must split evenly over the waves");
    constexpr int FW = F / W;  // fragment loads per wave per k step
    constexpr int STEP_V4 = F * FRAG_V4;
    constexpr int SLOT_V4 = KPB * STEP_V4;  // a ring slot holds KPB k steps
    __shared__ v4i ring[STAGES * SLOT_V4];
    const int lane = threadIdx.x & 63;
    const int w = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int per = gridDim.x / 8;
    const int t = (int)(blockIdx.x % 8) * per + (int)(blockIdx.x / 8);
    if (t >= ntiles) return;
    const int tk = t % ntk, tr = t / ntk;
    const int wr = w / WK, wk = w % WK;
    // this wave's fragments: f = w + u W
    const v4i *src[FW];
    long long step[FW];
#pragma unroll
    for (int u = 0; u < FW; ++u) {
        const int f = w + u * W;
        if (f < WR * DG_A) {
            src[u] = A + (size_t)(tr * WR + f / DG_A) * a_stride + (f % DG_A) * FRAG_V4 + lane;
            step[u] = DG_A * FRAG_V4;
        } else {
            const int fb = f - WR * DG_A;
            src[u] = B + (size_t)(tk * WK + fb / NB) * b_stride + (fb % NB) * FRAG_V4 + lane;
            step[u] = NB * FRAG_V4;
        }
    }
    const int ngroups = (ksteps + KPB - 1) / KPB;
    // group g = k steps [g KPB, g KPB + KPB) into ring slot g mod STAGES
    auto issue = [&](int g) {
        v4i *slot = ring + (g % STAGES) * SLOT_V4;
#pragma unroll
        for (int j = 0; j < KPB; ++j) {
            const int ks = g * KPB + j;
            if (ks >= ksteps) break;
#pragma unroll
            for (int u = 0; u < FW; ++u) {
                const int f = w + u * W;
                __builtin_amdgcn_global_load_lds((glb_void_t *)(src[u] + (size_t)ks * step[u]),
                                                 (lds_void_t *)(slot + j * STEP_V4 + f * FRAG_V4), 16, 0, 0);
            }
        }
    };
    v16i acc[DG_A][NB];
#pragma unroll
    for (int d = 0; d < DG_A; ++d)
#pragma unroll
        for (int e = 0; e < NB; ++e) acc[d][e] = v16i{};
#pragma unroll
    for (int p = 0; p < STAGES - 1; ++p)
        if (p < ngroups) issue(p);
    for (int g = 0; g < ngroups; ++g) {
        // loads issued after group g: the next STAGES - 2 groups, all full
        // unless they include the last group (then wait for everything)
        if (g + STAGES - 2 < ngroups - 1)
            wait_vmcnt<(STAGES - 2) * FW * KPB>();
        else
            wait_vmcnt<0>();
        // group g in LDS for every wave; slot (g - 1) read by all.  A bare
        // barrier: a workgroup fence would also wait for the later groups'
        // loads.  The LDS reads of group g - 1 completed before their MFMAs.
        asm volatile("s_barrier" ::: "memory");
        if (g + STAGES - 1 < ngroups) issue(g + STAGES - 1);
        const v4i *slot = ring + (g % STAGES) * SLOT_V4;
#pragma unroll
        for (int j = 0; j < KPB; ++j) {
            if (g * KPB + j >= ksteps) break;
            const v4i *st = slot + j * STEP_V4;
            v4i fa[DG_A], fb[NB];
#pragma unroll
            for (int d = 0; d < DG_A; ++d) fa[d] = st[(wr * DG_A + d) * FRAG_V4 + lane];
#pragma unroll
            for (int e = 0; e < NB; ++e) fb[e] = st[(WR * DG_A + wk * NB + e) * FRAG_V4 + lane];
#pragma unroll
            for (int d = 0; d < DG_A; ++d)
#pragma unroll
                for (int e = 0; e < NB; ++e)
                    acc[d][e] = __builtin_amdgcn_mfma_i32_32x32x32_i8(fa[d], fb[e], acc[d][e], 0, 0, 0);
        }
    }
    const int key = tk * (32 * WK) + wk * 32 + (lane & 31);
    if (key >= S) return;
    const int row0 = tr * (32 * WR) + wr * 32 + (lane >> 5) * 4;
    i8_store<DG_A, NB>(acc, row0, key, R, S, q, out);
}

// ---------------------------------------------------------------------------
// Pipelined variant with the step's work interleaved by hand.  Of a step's
// DA x NB x KW MFMAs, the first two are issued before the wait for the next
// ring slot and the barrier; the third is followed by the ring loads of step
// ks + STAGES, the fourth by the LDS reads of the next step's key fragments,
// and each of the next DA by the read of one next row fragment.  Scheduling
// barriers keep that order, so the barrier and the LDS latency fall under
// queued MFMAs instead of between steps, and a second register set holds the
// next step's fragments.
//
// The products of equal weight 256^(d + e) add into one accumulator
// (DA + NB - 1 of them instead of DA NB), which frees the registers for
// KW = 2: each wave then computes a 32 x 64 output block from 4 key
// fragments and DA row fragments per step -- a third fewer LDS reads per
// MFMA than 32 x 32 -- and, at two 4-wave blocks per CU, one block's
// epilogue runs while the other's MFMAs keep the matrix pipe busy.  OCC is
// the waves per SIMD the registers are budgeted for.
//
// RW = 2 gives each wave RW x KW = 2 x 2 blocks (64 x 64 outputs, 24 MFMAs
// per step from 6 row and 4 key fragments with DA = 3): fewer LDS reads per
// MFMA than 32 x 64, a 128 x 128 block tile (half the operand bytes per MFMA
// from L2), and 16 accumulators of 16 registers -- all 256 accumulation
// registers of a wave at one wave per SIMD.  (With DA = 4 the 20 accumulators
// do not fit them, and the compiler moves accumulators between register files
// on every step; only DA = 3, limb 0, runs this shape.)
//
// DA = 3 uses only row digits 0..2: valid when every row value lies in
// [-8421504, 8355711] (digit 3 is then zero), which the row digit store
// guarantees for limb 0 (q < 2^24).  The zero digit plane is still loaded, so
// every wave issues the same number of ring loads.
//
// The main loop runs while the ring still has steps to load; the last
// STAGES + 1 steps (at most) wait with the conservative vmcnt(0).
// ---------------------------------------------------------------------------
template <int WR, int WK, int NB, int STAGES, int DA, bool PT, int OCC, int KW, int RW = 1>
__global__ __launch_bounds__(64 * WR * WK, OCC) void k_contract_i8_il(const v4i *__restrict__ A,
                                                                    const v4i *__restrict__ B, int R, int S,
                                                                    int ksteps, long long a_stride,
                                                                    long long b_stride, uint32_t q,
                                                                    I8Out out, int ntk, int ntiles, int gr,
                                                                    int kg)
{
    constexpr int W = WR * WK;
    constexpr int FA = WR * RW * DG_A;  // row fragments per step (every row digit plane)
    constexpr int F = FA + WK * KW * NB;
    static_assert(F % W == 0, "fragments must split evenly over the waves");
    static_assert(DA * NB >= 5 && DA <= DG_A, "the interleave needs five MFMAs per step");
    constexpr int FW = F / W;
    constexpr int STEP_V4 = F * FRAG_V4;
    constexpr int NM = RW * DA * NB * KW;  // MFMAs per step
    __shared__ v4i ring[STAGES * STEP_V4];
    const int lane = threadIdx.x & 63;
    const int w = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int wr = w / WK, wk = w % WK;
    // This block's tiles t0, t0 + tstride, ... (nt of them).  The tiles are
    // split into 8 contiguous ranges, one per XCD (block b runs on XCD b mod
    // 8), so the blocks of one XCD work on neighbouring tiles in tile_rc's
    // order and share their digits in that XCD's L2.
    int t0, tstride, nt;
    if (PT) {
        // each XCD's share a whole number of rounds of its bx blocks, so the
        // tiles its blocks run at once are one aligned run of tile_rc's order
        const int bx = gridDim.x / 8, per = ((ntiles + 7) / 8 + bx - 1) / bx * bx;
        const int lo = (int)(blockIdx.x % 8) * per, hi = min(ntiles, lo + per);
        t0 = lo + (int)(blockIdx.x / 8);
        tstride = bx;
        nt = t0 < hi ? (hi - t0 + bx - 1) / bx : 0;
    } else {
        t0 = (int)(blockIdx.x % 8) * (int)(gridDim.x / 8) + (int)(blockIdx.x / 8);
        tstride = 0;
        nt = t0 < ntiles ? 1 : 0;
    }
    if (nt == 0) return;
    // the block's k steps in order: step g = tile g / ksteps, k step g mod ksteps
    const int G = nt * ksteps;
    // fragment bases and strides are wave-uniform (scalar registers); each
    // load adds the lane
    long long step[FW];
    const v4i *base[FW];
    auto set_base = [&](int j) {
        int tr, tk;
        tile_rc(t0 + j * tstride, ntk, ntiles / ntk, gr, kg, tr, tk);
#pragma unroll
        for (int u = 0; u < FW; ++u) {
            const int f = w + u * W;
            if (f < FA) {
                base[u] = A + (size_t)(tr * WR * RW + f / DG_A) * a_stride + (f % DG_A) * FRAG_V4;
                step[u] = DG_A * FRAG_V4;
            } else {
                const int fb = f - FA;
                base[u] = B + (size_t)(tk * WK * KW + fb / NB) * b_stride + (fb % NB) * FRAG_V4;
                step[u] = NB * FRAG_V4;
            }
        }
    };
    set_base(0);
    int it_j = 0, it_ks = 0, it_slot = 0;  // next step to load
    auto issue_next = [&]() {
        v4i *slot = ring + it_slot * STEP_V4;
#pragma unroll
        for (int u = 0; u < FW; ++u) {
            const int f = w + u * W;
            __builtin_amdgcn_global_load_lds((glb_void_t *)(base[u] + (size_t)it_ks * step[u] + lane),
                                             (lds_void_t *)(slot + f * FRAG_V4), 16, 0, 0);
        }
        it_slot = it_slot + 1 == STAGES ? 0 : it_slot + 1;
        if (++it_ks == ksteps) {
            it_ks = 0;
            if (++it_j < nt) set_base(it_j);
        }
    };
    // row fragment x = (row block x / DA, digit x % DA) of the wave
    const int a_off = wr * RW * DG_A * FRAG_V4 + lane;
    auto a_frag = [&](int x) { return a_off + ((x / DA) * DG_A + x % DA) * FRAG_V4; };
    const int b_off = (FA + wk * KW * NB) * FRAG_V4 + lane;
    // per (row block, key block) of the wave, one accumulator per weight
    // 256^(d + e)
    constexpr int NACC = DA + NB - 1;
    constexpr int NBLK = RW * KW;
    v16i acc[NBLK * NACC];
#pragma unroll
    for (int c = 0; c < NBLK * NACC; ++c) acc[c] = v16i{};
#pragma unroll
    for (int p = 0; p < STAGES; ++p)
        if (p < G) issue_next();
    if (G >= STAGES)
        wait_vmcnt<(STAGES - 1) * FW>();
    else
        wait_vmcnt<0>();
    asm volatile("s_barrier" ::: "memory");
    constexpr int NFA = RW * DA;
    v4i fa0[NFA], fb0[KW * NB], fa1[NFA], fb1[KW * NB];
#pragma unroll
    for (int x = 0; x < NFA; ++x) fa0[x] = ring[a_frag(x)];
#pragma unroll
    for (int e = 0; e < KW * NB; ++e) fb0[e] = ring[b_off + e * FRAG_V4];
    int rd_slot = 1 % STAGES;  // slot of step g + 1
    // One step g.  In the last step the reads of "step G" fetch a slot no one
    // loads any more: in bounds, and never used.  Every wave runs the same
    // number of barriers.
    auto body = [&](auto main, int g, v4i (&fa)[NFA], v4i (&fb)[KW * NB], v4i (&na)[NFA], v4i (&nb)[KW * NB]) {
        constexpr bool M = decltype(main)::value;
        const v4i *nx = ring + rd_slot * STEP_V4;
#pragma unroll
        for (int i = 0; i < NM; ++i) {
            // MFMA i: key block kb (alternating), row block rb, row digit d,
            // key digit e (reversed on odd d): two consecutive MFMAs never
            // add into the same accumulator
            const int kb = i % KW, rb = (i / KW) % RW, j = i / (KW * RW);
            const int d = j / NB, e = (d & 1) ? NB - 1 - j % NB : j % NB;
            const int c = (rb * KW + kb) * NACC + d + e;
            acc[c] = __builtin_amdgcn_mfma_i32_32x32x32_i8(fa[rb * DA + d], fb[kb * NB + e], acc[c], 0, 0, 0);
            __builtin_amdgcn_sched_barrier(0);
            if (i == 1) {
                if constexpr (M)
                    wait_vmcnt<(STAGES - 2) * FW>();
                else
                    wait_vmcnt<0>();
                asm volatile("s_barrier" ::: "memory");
            } else if (i == 2) {
                if constexpr (M)
                    issue_next();
                else if (g + STAGES < G)
                    issue_next();
            } else if (i == 3) {
#pragma unroll
                for (int e = 0; e < KW * NB; ++e) nb[e] = nx[b_off + e * FRAG_V4];
            } else if (i >= 4) {
                // row fragment i - 4; the last MFMA also takes any left over
#pragma unroll
                for (int x = 0; x < NFA; ++x)
                    if (x == i - 4 || (i == NM - 1 && x > i - 4)) na[x] = nx[a_frag(x)];
            }
            __builtin_amdgcn_sched_barrier(0);
        }
        rd_slot = rd_slot + 1 == STAGES ? 0 : rd_slot + 1;
    };
    auto store = [&](int j) {
        int tr, tk;
        tile_rc(t0 + j * tstride, ntk, ntiles / ntk, gr, kg, tr, tk);
#pragma unroll
        for (int rb = 0; rb < RW; ++rb) {
            const int row0 = tr * (32 * WR * RW) + (wr * RW + rb) * 32 + (lane >> 5) * 4;
#pragma unroll
            for (int kb = 0; kb < KW; ++kb) {
                const int key = tk * (32 * WK * KW) + (wk * KW + kb) * 32 + (lane & 31);
                if (key >= S) continue;
                // a persistent block stores with the next tile's fragments
                // live, a 64 x 64 wave with the accumulators of 4 blocks
                constexpr int LB = RW > 1 ? 1 : OCC < 3 ? 8 : PT ? 1 : 4;
                i8_store_merged<DA, NB, LB>(reinterpret_cast<const v16i(&)[NACC]>(acc[(rb * KW + kb) * NACC]),
                                            row0, key, R, S, q, out);
            }
        }
    };
    // after step g of a persistent block: at the end of a tile, store it and
    // restart the sums (a one-tile block stores after its loop)
    int c_j = 0, c_ks = 0;
    auto tile_end = [&]() {
        if constexpr (PT) {
            if (++c_ks < ksteps) return;
            store(c_j);
#pragma unroll
            for (int c = 0; c < NBLK * NACC; ++c) acc[c] = v16i{};
            c_ks = 0;
            ++c_j;
        }
    };
    using Main = std::integral_constant<bool, true>;
    using Tail = std::integral_constant<bool, false>;
    int g = 0;
    // main: steps g and g + 1 both still load step + STAGES
    for (; g + 1 + STAGES < G; g += 2) {
        body(Main{}, g, fa0, fb0, fa1, fb1);
        tile_end();
        body(Main{}, g + 1, fa1, fb1, fa0, fb0);
        tile_end();
    }
    for (; g < G; g += 2) {
        body(Tail{}, g, fa0, fb0, fa1, fb1);
        tile_end();
        if (g + 1 < G) {
            body(Tail{}, g + 1, fa1, fb1, fa0, fb0);
            tile_end();
        }
    }
    if constexpr (!PT) store(0);
}

// ---------------------------------------------------------------------------
// Split-K form for small row counts (serving batches): a few row tiles leave
// most CUs idle and each launch waits for its whole k loop, so the three limbs
// and nsplit slices of the k steps run as separate blocks of one launch --
// block (limb, slice, tile) -- with the interleaved pipeline of the kernel
// above over its slice.  Each block writes the residue of its partial sum,
// Dpart[limb][slice][R][S]; k_split_finish adds the slices mod q and does the
// CRT and rounding.  Two key digits, four row digits on every limb.
// ---------------------------------------------------------------------------
template <int WR, int WK, int STAGES>
__global__ __launch_bounds__(64 * WR * WK, 3) void k_contract_i8_split(
    const v4i *__restrict__ A, long long a_limb, const v4i *__restrict__ B, long long b_limb, int R, int S, int ksteps,
    long long a_stride, long long b_stride, int nsplit, int kper, int ntk, int ntiles, uint32_t *__restrict__ Dpart)
{
    constexpr int NB = 2, DA = DG_A;
    constexpr int W = WR * WK;
    constexpr int F = WR * DG_A + WK * NB;
    static_assert(F % W == 0, "fragments must split evenly over the waves");
    constexpr int FW = F / W;
    constexpr int STEP_V4 = F * FRAG_V4;
    constexpr int NM = DA * NB;
    __shared__ v4i ring[STAGES * STEP_V4];
    const int lane = threadIdx.x & 63;
    const int w = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int wr = w / WK, wk = w % WK;
    const int t = (int)(blockIdx.x % ntiles), rest = (int)(blockIdx.x / ntiles);
    const int slice = rest % nsplit, s = rest / nsplit;
    const int k0 = slice * kper, G = min(kper, ksteps - k0);
    if (G <= 0) return;  // whole block
    const int tk = t % ntk, tr = t / ntk;
    const v4i *Al = A + s * a_limb, *Bl = B + s * b_limb;
    const v4i *base[FW];
    long long step[FW];
#pragma unroll
    for (int u = 0; u < FW; ++u) {
        const int f = w + u * W;
        if (f < WR * DG_A) {
            base[u] = Al + (size_t)(tr * WR + f / DG_A) * a_stride + (f % DG_A) * FRAG_V4 + lane;
            step[u] = DG_A * FRAG_V4;
        } else {
            const int fb = f - WR * DG_A;
            base[u] = Bl + (size_t)(tk * WK + fb / NB) * b_stride + (fb % NB) * FRAG_V4 + lane;
            step[u] = NB * FRAG_V4;
        }
    }
    int it_ks = k0, it_slot = 0;
    auto issue_next = [&]() {
        v4i *slot = ring + it_slot * STEP_V4;
#pragma unroll
        for (int u = 0; u < FW; ++u) {
            const int f = w + u * W;
            __builtin_amdgcn_global_load_lds((glb_void_t *)(base[u] + (size_t)it_ks * step[u]),
                                             (lds_void_t *)(slot + f * FRAG_V4), 16, 0, 0);
        }
        it_slot = it_slot + 1 == STAGES ? 0 : it_slot + 1;
        ++it_ks;
    };
    const int a_off = wr * DG_A * FRAG_V4 + lane;
    const int b_off = (WR * DG_A + wk * NB) * FRAG_V4 + lane;
    constexpr int NACC = DA + NB - 1;  // one accumulator per weight 256^(d + e)
    v16i acc[NACC];
#pragma unroll
    for (int c = 0; c < NACC; ++c) acc[c] = v16i{};
#pragma unroll
    for (int p = 0; p < STAGES; ++p)
        if (p < G) issue_next();
    if (G >= STAGES)
        wait_vmcnt<(STAGES - 1) * FW>();
    else
        wait_vmcnt<0>();
    asm volatile("s_barrier" ::: "memory");
    v4i fa0[DA], fb0[NB], fa1[DA], fb1[NB];
#pragma unroll
    for (int d = 0; d < DA; ++d) fa0[d] = ring[a_off + d * FRAG_V4];
#pragma unroll
    for (int e = 0; e < NB; ++e) fb0[e] = ring[b_off + e * FRAG_V4];
    int rd_slot = 1 % STAGES;
    // as in k_contract_i8_il: the reads of "step G" in the last step fetch a
    // slot no one loads any more (in bounds, unused)
    auto body = [&](auto main, int g, v4i (&fa)[DA], v4i (&fb)[NB], v4i (&na)[DA], v4i (&nb)[NB]) {
        constexpr bool M = decltype(main)::value;
        const v4i *nx = ring + rd_slot * STEP_V4;
#pragma unroll
        for (int i = 0; i < NM; ++i) {
            const int d = i / NB, e = (d & 1) ? NB - 1 - i % NB : i % NB;
            acc[d + e] = __builtin_amdgcn_mfma_i32_32x32x32_i8(fa[d], fb[e], acc[d + e], 0, 0, 0);
            __builtin_amdgcn_sched_barrier(0);
            if (i == 1) {
                if constexpr (M)
                    wait_vmcnt<(STAGES - 2) * FW>();
                else
                    wait_vmcnt<0>();
                asm volatile("s_barrier" ::: "memory");
            } else if (i == 2) {
                if constexpr (M)
                    issue_next();
                else if (g + STAGES < G)
                    issue_next();
            } else if (i == 3) {
#pragma unroll
                for (int e = 0; e < NB; ++e) nb[e] = nx[b_off + e * FRAG_V4];
            } else if (i >= 4) {
#pragma unroll
                for (int d = 0; d < DA; ++d)
                    if (d == i - 4 || (i == NM - 1 && d > i - 4)) na[d] = nx[a_off + d * FRAG_V4];
            }
            __builtin_amdgcn_sched_barrier(0);
        }
        rd_slot = rd_slot + 1 == STAGES ? 0 : rd_slot + 1;
    };
    using Main = std::integral_constant<bool, true>;
    using Tail = std::integral_constant<bool, false>;
    int g = 0;
    for (; g + 1 + STAGES < G; g += 2) {
        body(Main{}, g, fa0, fb0, fa1, fb1);
        body(Main{}, g + 1, fa1, fb1, fa0, fb0);
    }
    for (; g < G; g += 2) {
        body(Tail{}, g, fa0, fb0, fa1, fb1);
        if (g + 1 < G) body(Tail{}, g + 1, fa1, fb1, fa0, fb0);
    }
    const int key = tk * (32 * WK) + wk * 32 + (lane & 31);
    if (key >= S) return;
    const I8Out out = {Dpart + (size_t)(s * nsplit + slice) * R * S, nullptr, nullptr, nullptr};
    i8_store_merged<DA, NB>(acc, tr * (32 * WR) + wr * 32 + (lane >> 5) * 4, key, R, S, qmod(s), out);
}

// Sums the slices of k_contract_i8_split mod q per limb, then CRT and rounding.
// All 3 nsplit loads of a thread are issued before the first is used (the
// loop is unrolled to the 16-slice maximum), so the kernel waits for memory
// once, not once per slice.
constexpr int kMaxSlices = 16;
__global__ void k_split_finish(const uint32_t *__restrict__ Dpart, int nsplit, size_t n, uint32_t *__restrict__ kval)
{
    const size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n) return;
    uint32_t v[3][kMaxSlices];
#pragma unroll
    for (int s = 0; s < 3; ++s)
#pragma unroll
        for (int j = 0; j < kMaxSlices; ++j) v[s][j] = j < nsplit ? Dpart[((size_t)s * nsplit + j) * n + t] : 0u;
    uint32_t d[3];
#pragma unroll
    for (int s = 0; s < 3; ++s) {
        uint64_t sum = 0;  // at most 16 residues below 2^31
#pragma unroll
        for (int j = 0; j < kMaxSlices; ++j) sum += v[s][j];
        d[s] = (uint32_t)(sum % qmod(s));
    }
    kval[t] = crt_decode_canon(d[0], d[1], d[2]);
}

// ---------------------------------------------------------------------------
// Hybrid variant: one row block shared by WK waves, one key block per wave.
// The row digits (4 fragments per k step, used by every wave) go through the
// LDS ring by global_load_lds; each wave's key digits (used by that wave only)
// are loaded straight into registers, STAGES - 1 steps ahead.  LDS then
// carries 4 KB per wave and step instead of 4 + NB KB, which keeps it below
// the MFMA time.  Loads of step j are issued together (row fragment DMA first,
// then the key loads), so s_waitcnt vmcnt counts whole steps.
// ---------------------------------------------------------------------------
template <int WK, int NB, int STAGES>
__global__ __launch_bounds__(64 * WK, 2) void k_contract_i8_hyb(const v4i *__restrict__ A,
                                                               const v4i *__restrict__ B, int R, int S,
                                                               int ksteps, long long a_stride, long long b_stride,
                                                               uint32_t q, I8Out out, int ntk,
                                                               int ntiles)
{
    static_assert(DG_A % WK == 0, "row fragments must split evenly over the waves");
    constexpr int FA = DG_A / WK;  // row fragment loads per wave per step
    constexpr int P = STAGES - 1;  // steps in flight
    constexpr int SLOT_V4 = DG_A * FRAG_V4;
    __shared__ v4i ring[STAGES * SLOT_V4];
    const int lane = threadIdx.x & 63;
    const int w = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int per = gridDim.x / 8;
    const int t = (int)(blockIdx.x % 8) * per + (int)(blockIdx.x / 8);
    if (t >= ntiles) return;
    const int tk = t % ntk, tr = t / ntk;
    const v4i *asrc = A + (size_t)tr * a_stride + lane;          // + (f + 4 ks) 64
    const v4i *bsrc = B + (size_t)(tk * WK + w) * b_stride + lane;  // + (e + NB ks) 64
    v4i breg[P][NB];
    auto issue = [&](int ks, v4i (&bdst)[NB]) {
        v4i *slot = ring + (ks % STAGES) * SLOT_V4;
#pragma unroll
        for (int u = 0; u < FA; ++u) {
            const int f = w + u * WK;
            __builtin_amdgcn_global_load_lds((glb_void_t *)(asrc + ((size_t)ks * DG_A + f) * FRAG_V4),
                                             (lds_void_t *)(slot + f * FRAG_V4), 16, 0, 0);
        }
#pragma unroll
        for (int e = 0; e < NB; ++e) bdst[e] = bsrc[((size_t)ks * NB + e) * FRAG_V4];
    };
    v16i acc[DG_A][NB];
#pragma unroll
    for (int d = 0; d < DG_A; ++d)
#pragma unroll
        for (int e = 0; e < NB; ++e) acc[d][e] = v16i{};
#pragma unroll
    for (int p = 0; p < P; ++p)
        if (p < ksteps) issue(p, breg[p]);
    for (int ks0 = 0; ks0 < ksteps; ks0 += P) {
#pragma unroll
        for (int j = 0; j < P; ++j) {
            const int ks = ks0 + j;
            if (ks >= ksteps) break;
            // steps issued after ks: ks+1 .. ks+P-1 (FA + NB loads each)
            if (ks + P - 1 < ksteps)
                wait_vmcnt<(P - 1) * (FA + NB)>();
            else
                wait_vmcnt<0>();
            asm volatile("s_barrier" ::: "memory");  // row fragments of step ks in LDS
            const v4i *st = ring + (ks % STAGES) * SLOT_V4;
            v4i fa[DG_A], fb[NB];
#pragma unroll
            for (int d = 0; d < DG_A; ++d) fa[d] = st[d * FRAG_V4 + lane];
#pragma unroll
            for (int e = 0; e < NB; ++e) fb[e] = breg[j][e];
            if (ks + P < ksteps) issue(ks + P, breg[j]);
#pragma unroll
            for (int d = 0; d < DG_A; ++d)
#pragma unroll
                for (int e = 0; e < NB; ++e)
                    acc[d][e] = __builtin_amdgcn_mfma_i32_32x32x32_i8(fa[d], fb[e], acc[d][e], 0, 0, 0);
        }
    }
    const int key = tk * (32 * WK) + w * 32 + (lane & 31);
    if (key >= S) return;
    const int row0 = tr * 32 + (lane >> 5) * 4;
    i8_store<DG_A, NB>(acc, row0, key, R, S, q, out);
}

constexpr int kThreeDigitMax = 127 * (1 + 256 + 65536);  // 8355711

__device__ __forceinline__ int digit_split(int &x)
{
    const int d = ((x + 128) & 255) - 128;
    x = (x - d) / 256;
    return d;
}

// Compact rows -> every row-store form in one pass (RowStore::put): rows
// [r0, r0 + R) from c_L [R][3][N] and coefficient 0 of c_0..c_{src_l-1},
// c0 [R][src_l][3] (words in [0, q]; q is taken as 0):
//   A [3][cap][ldr]    balanced int32 (c_L, then the c0 columns, zero-padded),
//   c0s [cap][l][3]    the c0 words as given (columns < src_l),
//   dig                the int8 digit store (limb 0 in its three-digit window).
// One thread per (limb, row, 4 consecutive columns of the ksteps x 32).
__global__ void k_rows_put(const uint32_t *__restrict__ cl, const uint32_t *__restrict__ c0, int src_l, int R, int r0,
                           int cap, int ldr, int l_store, int ksteps, size_t limb_bytes, int32_t *__restrict__ A,
                           uint32_t *__restrict__ c0s, int8_t *__restrict__ dig)
{
    const int quads = ksteps * 8;
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= (size_t)3 * R * quads) return;
    const int s = (int)(i / ((size_t)R * quads));
    const size_t rem = i % ((size_t)R * quads);
    const int rr = (int)(rem / quads), row = r0 + rr, k = 4 * (int)(rem % quads);
    const uint32_t q = qmod(s);
    int v[4];
#pragma unroll
    for (int u = 0; u < 4; ++u) {
        const int col = k + u;
        uint32_t t = 0;
        if (col < IPFE_N) {
            t = cl[((size_t)rr * 3 + s) * IPFE_N + col];
        } else if (col - IPFE_N < src_l) {
            t = c0[((size_t)rr * src_l + (col - IPFE_N)) * 3 + s];
            c0s[((size_t)row * l_store + (col - IPFE_N)) * 3 + s] = t;
        }
        if (t >= q) t -= q;
        v[u] = (t > q / 2) ? (int)((int64_t)t - q) : (int)t;
        if (col < ldr) A[((size_t)s * cap + row) * ldr + col] = v[u];
    }
    // limb 0 (q < 2^24): the representative in [-8421504, 8355711], the range
    // of three balanced base-256 digits, so digit 3 is zero
    if (s == 0) {
        const int q0 = (int)qmod(0);
#pragma unroll
        for (int u = 0; u < 4; ++u) v[u] -= (v[u] > kThreeDigitMax) ? q0 : 0;
    }
    const int ks = k >> 5, kk = k & 31;
    const int ln = (kk >> 4) * 32 + (row & 31), j = kk & 15;
    int8_t *base = dig + s * limb_bytes + ((size_t)(row >> 5) * ksteps + ks) * DG_A * 1024 + ln * 16 + j;
#pragma unroll
    for (int d = 0; d < DG_A; ++d) {
        uint32_t word = 0;
#pragma unroll
        for (int u = 0; u < 4; ++u) word |= (uint32_t)(uint8_t)(int8_t)digit_split(v[u]) << (8 * u);
        *reinterpret_cast<uint32_t *>(base + (size_t)d * 1024) = word;
    }
}

// Balanced fp64 keys [3][S][K] -> key digit store [3][spad/32][ksteps][nb][..],
// keys >= S and k >= K zero.  Values must fit nb signed digits.
__global__ void k_keys_digits(const double *__restrict__ kd, int S, int K, int spad, int ksteps, int nb,
                              int8_t *__restrict__ dig)
{
    const int quads = ksteps * 8;
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= (size_t)3 * spad * quads) return;
    const int s = (int)(i / ((size_t)spad * quads));
    const size_t rem = i % ((size_t)spad * quads);
    const int key = (int)(rem / quads), k = 4 * (int)(rem % quads);
    int v[4];
#pragma unroll
    for (int u = 0; u < 4; ++u)
        v[u] = (key < S && k + u < K) ? (int)kd[((size_t)s * S + key) * K + k + u] : 0;
    const int ks = k >> 5, kk = k & 31;
    const int ln = (kk >> 4) * 32 + (key & 31), j = kk & 15;
    const size_t limb_bytes = (size_t)spad / 32 * ksteps * nb * 1024;
    int8_t *base = dig + s * limb_bytes + ((size_t)(key >> 5) * ksteps + ks) * nb * 1024 + ln * 16 + j;
    for (int e = 0; e < nb; ++e) {
        uint32_t word = 0;
#pragma unroll
        for (int u = 0; u < 4; ++u) word |= (uint32_t)(uint8_t)(int8_t)digit_split(v[u]) << (8 * u);
        *reinterpret_cast<uint32_t *>(base + (size_t)e * 1024) = word;
    }
}

// 128 rows past the capacity: a tile of up to 4 row blocks may start at any
// 32-row boundary below the capacity
size_t rows_digit_limb_bytes(int cap, int ksteps) { return (size_t)(cap + 128) / 32 * ksteps * DG_A * 1024; }
size_t keys_digit_limb_bytes(int spad, int ksteps, int nb) { return (size_t)spad / 32 * ksteps * nb * 1024; }

int launch_rows_put(const uint32_t *cl, const uint32_t *c0, int src_l, int R, int r0, int cap, int ldr, int l_store,
                    int ksteps, int32_t *A, uint32_t *c0s, int8_t *dig, hipStream_t st)
{
    const size_t n = (size_t)3 * R * ksteps * 8;
    if (n == 0) return 0;
    if (ksteps * 32 < ldr || ldr < IPFE_N + src_l || src_l > l_store || r0 < 0 || r0 + R > cap) {
        ipfe_set_error("row store: %d k steps for %d columns (%d slots of %d), rows [%d, %d) of %d", ksteps, ldr,
                       src_l, l_store, r0, r0 + R, cap);
        return -1;
    }
    hipLaunchKernelGGL(k_rows_put, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, cl, c0, src_l, R, r0, cap,
                       ldr, l_store, ksteps, rows_digit_limb_bytes(cap, ksteps), A, c0s, dig);
    HIP_OK(hipGetLastError());
    return 0;
}

int launch_keys_digits(const double *kd, int S, int K, int spad, int ksteps, int nb, int8_t *dig, hipStream_t st)
{
    const size_t n = (size_t)3 * spad * ksteps * 8;
    if (n == 0) return 0;
    if (spad % 128 || spad < S || ksteps * 32 < K || nb < 2 || nb > 3) {
        ipfe_set_error("key digits: %d keys padded to %d, %d k steps for K=%d", S, spad, ksteps, K);
        return -1;
    }
    hipLaunchKernelGGL(k_keys_digits, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, kd, S, K, spad, ksteps, nb,
                       dig);
    HIP_OK(hipGetLastError());
    return 0;
}

// Kernel variants the launcher can run: row blocks x key blocks per block
// tile, the LDS ring depth (0: the register-staged double-buffered kernel),
// the kernel, the row digits the MFMAs use (3: limb 0 only), and for the
// interleaved kernels the waves per SIMD the registers are budgeted for and
// the key blocks per wave (block tile: wr x wk kw blocks of 32), and their
// tile order.
enum I8Kind { I8_REG, I8_RING, I8_RING_K2, I8_HYBRID, I8_INTERLEAVED, I8_PERSISTENT };
struct I8Variant {
    int wr, wk, stages, kind, da;
    int occ = 2;
    int kw = 1;
    int gr = 0, kg = 0;  // tile order of the interleaved kernels: gr x kg groups (0: row-major)
    int rw = 1;          // row blocks per wave (interleaved kernels)
};
static const I8Variant kVariants[] = {
    // register-staged, LDS rings and the hybrid (also the three-key-digit set)
    {2, 2, 0, I8_REG, 4}, {1, 4, 0, I8_REG, 4}, {2, 4, 0, I8_REG, 4}, {1, 4, 4, I8_RING, 4}, {1, 4, 3, I8_RING, 4},
    {1, 4, 2, I8_RING_K2, 4}, {1, 4, 3, I8_HYBRID, 4},
    // interleaved, merged accumulators, 32 x 32 per wave
    {2, 2, 4, I8_INTERLEAVED, 4, 3}, {1, 4, 4, I8_INTERLEAVED, 4, 3}, {2, 2, 4, I8_INTERLEAVED, 3, 3},
    {1, 4, 4, I8_INTERLEAVED, 3, 3}, {2, 2, 3, I8_INTERLEAVED, 4, 3}, {2, 2, 3, I8_INTERLEAVED, 3, 4},
    {1, 4, 3, I8_INTERLEAVED, 3, 4},
    // the same, persistent
    {2, 2, 4, I8_PERSISTENT, 4, 3}, {2, 2, 4, I8_PERSISTENT, 3, 3}, {2, 4, 4, I8_PERSISTENT, 4, 2},
    {2, 4, 4, I8_PERSISTENT, 3, 2},
    // interleaved, 32 x 64 per wave
    {4, 1, 3, I8_INTERLEAVED, 4, 2, 2}, {4, 1, 3, I8_INTERLEAVED, 3, 2, 2}, {2, 2, 4, I8_INTERLEAVED, 4, 2, 2},
    {2, 2, 4, I8_INTERLEAVED, 3, 2, 2}, {2, 2, 3, I8_INTERLEAVED, 4, 2, 2}, {2, 2, 5, I8_INTERLEAVED, 4, 2, 2},
    {2, 2, 5, I8_INTERLEAVED, 3, 2, 2},
    // the same in grouped tile order (gr row tiles x kg key tiles)
    {2, 2, 4, I8_INTERLEAVED, 4, 2, 2, 8, 8}, {2, 2, 5, I8_INTERLEAVED, 4, 2, 2, 8, 8},
    {2, 2, 4, I8_INTERLEAVED, 3, 2, 2, 8, 8}, {4, 1, 3, I8_INTERLEAVED, 4, 2, 2, 4, 16},
    {4, 1, 3, I8_INTERLEAVED, 3, 2, 2, 4, 16}, {1, 4, 4, I8_INTERLEAVED, 3, 3, 1, 16, 6},
    {2, 2, 4, I8_INTERLEAVED, 4, 3, 1, 8, 12},
    // interleaved, 64 x 64 per wave, one wave per SIMD (block tile 128 x 128),
    // three row digits
    {2, 2, 4, I8_INTERLEAVED, 3, 1, 2, 4, 8, 2}, {2, 2, 3, I8_INTERLEAVED, 3, 1, 2, 4, 8, 2},
};
constexpr int kNumVariants = sizeof(kVariants) / sizeof(kVariants[0]);

int i8_shape_count() { return kNumVariants; }
void i8_shape(int v, int *wr, int *wk)
{
    *wr = kVariants[v].wr * kVariants[v].rw;
    *wk = kVariants[v].wk * kVariants[v].kw;
}
// "WRxWK" for the register-staged kernel; "WRxWK/STAGES" plus "k2" (two k
// steps per barrier), "h" (hybrid), "i" (interleaved) or "ip" (interleaved,
// persistent, each with "m": merged accumulators) for the others; "w" for
// waves of two key blocks ("q": 2 x 2 blocks), "3" for three row digits and
// "gGRxKG" for the grouped tile order
const char *i8_variant_label(int v)
{
    struct Labels {
        char s[kNumVariants][24];
        Labels()
        {
            for (int i = 0; i < kNumVariants; ++i) {
                const I8Variant &x = kVariants[i];
                const char *suffix = x.kind == I8_RING_K2       ? "k2"
                                     : x.kind == I8_HYBRID      ? "h"
                                     : x.kind == I8_INTERLEAVED ? "im"
                                     : x.kind == I8_PERSISTENT  ? "ipm"
                                                                : "";
                if (x.kind == I8_REG)
                    snprintf(s[i], sizeof s[i], "%dx%d", x.wr, x.wk);
                else if (x.gr)
                    snprintf(s[i], sizeof s[i], "%dx%d/%d%s%s%sg%dx%d", x.wr * x.rw, x.wk * x.kw, x.stages, suffix,
                             x.rw > 1 ? "q" : x.kw > 1 ? "w" : "", x.da == 3 ? "3" : "", x.gr, x.kg);
                else
                    snprintf(s[i], sizeof s[i], "%dx%d/%d%s%s%s", x.wr, x.wk * x.kw, x.stages, suffix,
                             x.kw > 1 ? "w" : "", x.da == 3 ? "3" : "");
            }
        }
    };
    static const Labels labels;  // initialised once, thread-safe
    return v >= 0 && v < kNumVariants ? labels.s[v] : "?";
}
// Variants instantiated for nb key digits on limb s (three row digits only on
// limb 0, whose row digit store holds a zero digit 3).  With three key digits:
// the register-staged kernels and the non-persistent 1x4 tiles of the others
// (the 1x4/3 hybrid only), whose ring slots split evenly over the waves.
bool i8_variant_ok(int v, int nb, int s)
{
    if (v < 0 || v >= kNumVariants || nb < 2 || nb > 3) return false;
    const I8Variant &x = kVariants[v];
    if (x.da == 3 && s != 0) return false;
    if (nb == 2 || x.kind == I8_REG) return true;
    if (x.kind == I8_HYBRID) return x.stages == 3;
    return x.wr == 1 && x.kw == 1 && x.kind != I8_PERSISTENT;
}

// Grid of a persistent kernel: every CU filled once at the kernel's
// occupancy, a multiple of 8 (one share per XCD), no more than the tiles.
static int persistent_grid(const void *fn, int block_threads, int ntiles)
{
    static std::mutex mu;
    static std::map<const void *, int> per_cu;
    static int cus = 0;
    int nb = 0;
    {
        std::lock_guard<std::mutex> lk(mu);
        if (!cus && hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, 0) != hipSuccess) cus = 256;
        auto it = per_cu.find(fn);
        if (it == per_cu.end()) {
            if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, fn, block_threads, 0) != hipSuccess || nb < 1) nb = 1;
            per_cu[fn] = nb;
        } else {
            nb = it->second;
        }
    }
    const int tiles8 = (ntiles + 7) / 8 * 8;
    return std::max(8, std::min(nb * cus, tiles8) / 8 * 8);
}

// All three limbs of a small row range in one split-K launch plus the finish
// kernel.  A: row digit store of limb 0 at row r0 (limb s at + s a_limb_bytes),
// B: key digit store of limb 0 (limb s at + s b_limb_bytes), two key digits.
// Dpart: 3 nsplit R S words of scratch, nsplit from split_k_slices.
int split_k_slices(int R, int S, int ksteps)
{
    const int wr = R <= 32 ? 1 : 2, wk = R <= 32 ? 4 : 2;
    const int tiles = ((R + 32 * wr - 1) / (32 * wr)) * ((S + 32 * wk - 1) / (32 * wk));
    // about 2,048 blocks, at least 8 k steps per slice; every slice non-empty
    const int n = std::max(1, std::min({2048 / std::max(1, 3 * tiles), kMaxSlices, ksteps / 8}));
    const int kper = (ksteps + n - 1) / n;
    return (ksteps + kper - 1) / kper;
}

int launch_contract_i8_split(const int8_t *A, size_t a_limb_bytes, int ksteps_a, const int8_t *B,
                             size_t b_limb_bytes, int ksteps, int R, int S, uint32_t *Dpart, uint32_t *kval,
                             hipStream_t st)
{
    if (R <= 0 || S <= 0) return 0;
    if (ksteps < 1 || ksteps > ksteps_a || ksteps * 32 > (1 << 14) || R > 4096) {
        ipfe_set_error("split int8 contraction: %d rows, %d k steps (row store %d)", R, ksteps, ksteps_a);
        return -1;
    }
    const int nsplit = split_k_slices(R, S, ksteps), kper = (ksteps + nsplit - 1) / nsplit;
    const bool narrow = R <= 32;
    const int wr = narrow ? 1 : 2, wk = narrow ? 4 : 2;
    const int ntk = (S + 32 * wk - 1) / (32 * wk), ntr = (R + 32 * wr - 1) / (32 * wr), ntiles = ntk * ntr;
    const dim3 grid((unsigned)(3 * nsplit * ntiles)), block(64 * wr * wk);
    const long long as = (long long)ksteps_a * DG_A * FRAG_V4, bs = (long long)ksteps * 2 * FRAG_V4;
    const v4i *a = reinterpret_cast<const v4i *>(A);
    const v4i *b = reinterpret_cast<const v4i *>(B);
    const long long al = (long long)(a_limb_bytes / 16), bl = (long long)(b_limb_bytes / 16);
    if (narrow)
        hipLaunchKernelGGL((k_contract_i8_split<1, 4, 3>), grid, block, 0, st, a, al, b, bl, R, S, ksteps, as, bs,
                           nsplit, kper, ntk, ntiles, Dpart);
    else
        hipLaunchKernelGGL((k_contract_i8_split<2, 2, 3>), grid, block, 0, st, a, al, b, bl, R, S, ksteps, as, bs,
                           nsplit, kper, ntk, ntiles, Dpart);
    HIP_OK(hipGetLastError());
    const size_t n = (size_t)R * S;
    hipLaunchKernelGGL(k_split_finish, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, Dpart, nsplit, n, kval);
    HIP_OK(hipGetLastError());
    return 0;
}

// A: row digit store of one limb at row r0 (r0 a multiple of 32; the store
// holds 128 rows past its capacity), ksteps_a its k steps; B: key digit store
// of one limb with nb digits, ksteps k steps (<= ksteps_a), keys padded to a
// multiple of 128.  variant: index into kVariants.
int launch_contract_i8(const int8_t *A, int ksteps_a, const int8_t *B, int ksteps, int nb, int R, int S, int s,
                       uint32_t *D_out, hipStream_t st, int variant, const uint32_t *crt_d0, const uint32_t *crt_d1,
                       uint32_t *kval)
{
    if (R <= 0 || S <= 0) return 0;
    if (ksteps < 1 || ksteps > ksteps_a || ksteps * 32 > (1 << 14) || !i8_variant_ok(variant, nb, s)) {
        ipfe_set_error("int8 contraction: %d k steps (row store %d), variant %d, %d key digits", ksteps, ksteps_a,
                       variant, nb);
        return -1;
    }
    const int WR = kVariants[variant].wr * kVariants[variant].rw, WK = kVariants[variant].wk * kVariants[variant].kw;
    // row tiles never reach past the 128 padding rows; key tiles stay within
    // the 128-key padding
    const int key_tile = 32 * WK, row_tile = 32 * WR;
    if (key_tile > 128 || row_tile > 128) {
        ipfe_set_error("int8 contraction: variant %d tiles past the 128-row / 128-key padding", variant);
        return -1;
    }
    const int ntk = (S + key_tile - 1) / key_tile, ntr = (R + row_tile - 1) / row_tile;
    const int ntiles = ntk * ntr;
    const dim3 grid(((ntiles + 7) / 8) * 8), block(64 * kVariants[variant].wr * kVariants[variant].wk);
    const v4i *a = reinterpret_cast<const v4i *>(A);
    const v4i *b = reinterpret_cast<const v4i *>(B);
    const long long as = (long long)ksteps_a * DG_A * FRAG_V4, bs = (long long)ksteps * nb * FRAG_V4;
    const uint32_t q = qmod(s);
    if (kval && (s != 2 || !crt_d0 || !crt_d1)) {
        ipfe_set_error("int8 contraction: the CRT epilogue runs on limb 2 with limbs 0 and 1 given");
        return -1;
    }
    const I8Out D = {kval ? nullptr : D_out, crt_d0, crt_d1, kval};
    const int gr = kVariants[variant].gr ? kVariants[variant].gr : 1;
    const int kg = kVariants[variant].gr ? std::min(kVariants[variant].kg, ntk) : ntk;
#define CI8(wr, wk, n) \
    hipLaunchKernelGGL((k_contract_i8<wr, wk, n>), grid, block, 0, st, a, b, R, S, ksteps, as, bs, q, D, ntk, ntiles)
#define CI8D(wr, wk, n, stg, kpb)                                                                                  \
    hipLaunchKernelGGL((k_contract_i8_dma<wr, wk, n, stg, kpb>), grid, block, 0, st, a, b, R, S, ksteps, as, bs, q, D, \
                       ntk, ntiles)
#define CI8I(wr, wk, n, stg, da, occ, kw)                                                                            \
    hipLaunchKernelGGL((k_contract_i8_il<wr, wk, n, stg, da, false, occ, kw>), grid, block, 0, st, a, b, R, S, ksteps, \
                       as, bs, q, D, ntk, ntiles, gr, kg)
#define CI8Q(wr, wk, n, stg, da, occ, kw, rw)                                                                         \
    hipLaunchKernelGGL((k_contract_i8_il<wr, wk, n, stg, da, false, occ, kw, rw>), grid, block, 0, st, a, b, R, S,     \
                       ksteps, as, bs, q, D, ntk, ntiles, gr, kg)
#define CI8P(wr, wk, n, stg, da, occ, kw)                                                                            \
    hipLaunchKernelGGL(                                                                                              \
        (k_contract_i8_il<wr, wk, n, stg, da, true, occ, kw>),                                                       \
        dim3(persistent_grid(reinterpret_cast<const void *>(&k_contract_i8_il<wr, wk, n, stg, da, true, occ, kw>),   \
                             block.x, ntiles)),                                                                      \
        block, 0, st, a, b, R, S, ksteps, as, bs, q, D, ntk, ntiles, gr, kg)
#define CI8H(wk, n, stg)                                                                                             \
    hipLaunchKernelGGL((k_contract_i8_hyb<wk, n, stg>), grid, block, 0, st, a, b, R, S, ksteps, as, bs, q, D, ntk,    \
                       ntiles)
    if (nb == 2) {
        switch (variant) {
        case 0: CI8(2, 2, 2); break;
        case 1: CI8(1, 4, 2); break;
        case 2: CI8(2, 4, 2); break;
        case 3: CI8D(1, 4, 2, 4, 1); break;
        case 4: CI8D(1, 4, 2, 3, 1); break;
        case 5: CI8D(1, 4, 2, 2, 2); break;
        case 6: CI8H(4, 2, 3); break;
        case 7: CI8I(2, 2, 2, 4, 4, 3, 1); break;
        case 8: CI8I(1, 4, 2, 4, 4, 3, 1); break;
        case 9: CI8I(2, 2, 2, 4, 3, 3, 1); break;
        case 10: CI8I(1, 4, 2, 4, 3, 3, 1); break;
        case 11: CI8I(2, 2, 2, 3, 4, 3, 1); break;
        case 12: CI8I(2, 2, 2, 3, 3, 4, 1); break;
        case 13: CI8I(1, 4, 2, 3, 3, 4, 1); break;
        case 14: CI8P(2, 2, 2, 4, 4, 3, 1); break;
        case 15: CI8P(2, 2, 2, 4, 3, 3, 1); break;
        case 16: CI8P(2, 4, 2, 4, 4, 2, 1); break;
        case 17: CI8P(2, 4, 2, 4, 3, 2, 1); break;
        case 18: CI8I(4, 1, 2, 3, 4, 2, 2); break;
        case 19: CI8I(4, 1, 2, 3, 3, 2, 2); break;
        case 20: CI8I(2, 2, 2, 4, 4, 2, 2); break;
        case 21: CI8I(2, 2, 2, 4, 3, 2, 2); break;
        case 22: CI8I(2, 2, 2, 3, 4, 2, 2); break;
        case 23: CI8I(2, 2, 2, 5, 4, 2, 2); break;
        case 24: CI8I(2, 2, 2, 5, 3, 2, 2); break;
        case 25: CI8I(2, 2, 2, 4, 4, 2, 2); break;
        case 26: CI8I(2, 2, 2, 5, 4, 2, 2); break;
        case 27: CI8I(2, 2, 2, 4, 3, 2, 2); break;
        case 28: CI8I(4, 1, 2, 3, 4, 2, 2); break;
        case 29: CI8I(4, 1, 2, 3, 3, 2, 2); break;
        case 30: CI8I(1, 4, 2, 4, 3, 3, 1); break;
        case 31: CI8I(2, 2, 2, 4, 4, 3, 1); break;
        case 32: CI8Q(2, 2, 2, 4, 3, 1, 2, 2); break;
        default: CI8Q(2, 2, 2, 3, 3, 1, 2, 2); break;  // variant 33
        }
    } else {
        switch (variant) {
        case 0: CI8(2, 2, 3); break;
        case 1: CI8(1, 4, 3); break;
        case 2: CI8(2, 4, 3); break;
        case 3: CI8D(1, 4, 3, 4, 1); break;
        case 4: CI8D(1, 4, 3, 3, 1); break;
        case 5: CI8D(1, 4, 3, 2, 2); break;
        case 6: CI8H(4, 3, 3); break;
        case 8: CI8I(1, 4, 3, 4, 4, 2, 1); break;
        case 10: CI8I(1, 4, 3, 4, 3, 2, 1); break;
        case 30: CI8I(1, 4, 3, 4, 3, 2, 1); break;
        default: CI8I(1, 4, 3, 3, 3, 2, 1); break;  // variant 13
        }
    }
#undef CI8
#undef CI8D
#undef CI8H
#undef CI8I
#undef CI8P
#undef CI8Q
    HIP_OK(hipGetLastError());
    return 0;
}

}  // namespace ipfe
