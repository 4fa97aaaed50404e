// k_predict.hip -- batched IPFE-SVM prediction: every (row, support vector)
// pair of svm_predict_cpu / svm_predict_gpu (main.c:137-390) at once.
//
// Only coefficient 0 of each decryption is used (main.c:205, src/rlwe_sife.c:380),
// so for row r and key v, limb s:
//   d0 = sub_mod( fold_i y_v[i] c_i[0],  (c_L * sk_v)[0] )
// where the negacyclic coefficient 0 is a plain dot product
//   (c_L * sk)[0] = c_0 sk_0 - sum_{j>=1} c_j sk_{N-j} = sum_j c_j s'_j,
//   s'_0 = sk_0,  s'_j = -sk_{N-j}.
// Over all rows and keys that is a rows x keys x 4096 modular contraction per
// limb.  It runs on the fp64 FMA pipe: operands are mapped to the balanced
// range (-q/2, q/2], every partial sum stays an exact integer below 2^53, and
// the accumulators are reduced exactly (fma(-rint(acc/q), q, acc)) every
// `kb` k-steps, where kb comes from the largest key coefficient.  Functional
// keys of small vectors y built on the sigma_1 master key are small, which makes
// kb large.  The sum part folds only the non-zero y_i with the reference's
// sequential mod_prime rule; with the reference's length L = 785 it always ends
// on zero slots and is canonical, so d0 does not depend on which
// representative the reference's inverse NTT returns for the product.
#include "crt_dev.h"
#include <stdlib.h>
#include <string.h>

#include "kernels_predict.h"

namespace ipfe {

// ---------------------------------------------------------------------------
// Keys: sk [S][3][N] canonical -> B [3][S][K] as balanced doubles (K >= N):
//   columns j < N:   sign * s'_j  with s'_0 = sk_0, s'_j = -sk_{N-j};
//   columns N + i:   y_v[i] mod q_s (the sum part of decryption), zero past the
//                    SV support and when K == N.
// sign = -1 gives D = sum_i y_i c_i[0] - (c_L * sk)[0] = d0 from one contraction
// of the row [c_L, c_0[0], .., c_{l-1}[0]].  maxabs[s] = largest |B| per limb.
// ---------------------------------------------------------------------------
__global__ void k_contract_keys(const uint32_t *__restrict__ sk, int S, int K, int sign, SparseY y,
                                double *__restrict__ kd, unsigned int *__restrict__ maxabs)
{
    const int v = blockIdx.y, s = blockIdx.z;
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    const uint32_t q = qmod(s);
    int64_t val = 0;
    if (j < IPFE_N) {
        const uint32_t *src = sk + ((size_t)v * 3 + s) * IPFE_N;
        uint32_t t = (j == 0) ? src[0] : src[IPFE_N - j];
        if (t >= q) t -= q;
        val = (j == 0) ? (int64_t)t : -(int64_t)t;
        val *= sign;
    } else if (j < K) {
        const int slot = j - IPFE_N;
        for (int e = y.off[v]; e < y.off[v + 1]; ++e)
            if (y.idx[e] == slot) val = mod_prime_exact(y.val[e], s);
        if (val >= (int64_t)q) val -= q;
    }
    if (val > (int64_t)(q / 2)) val -= q;
    if (val < -(int64_t)(q / 2)) val += q;
    if (j < K) kd[((size_t)s * S + v) * K + j] = (double)val;
    unsigned int a = (unsigned int)(val < 0 ? -val : val);
    // wave-level max then one atomic per wave
    for (int d = 32; d >= 1; d >>= 1) a = max(a, (unsigned int)__shfl_xor((int)a, d, 64));
    if ((threadIdx.x & 63) == 0) atomicMax(maxabs + s, a);
}


// ---------------------------------------------------------------------------
// Contraction: D[r][v] = sum_j A[r][j] B[v][j] mod q for one limb.
// Block tile 128 rows x 128 keys, k-step 16, thread tile 8 x 8 (rows ty+16i,
// keys tx+16j), fp64 accumulators reduced every kb k-steps.
// ---------------------------------------------------------------------------
constexpr int CT_BM = 128, CT_BN = 128, CT_BK = 16;

__device__ __forceinline__ double reduce_bal(double acc, double qd, double qinv)
{
    return __fma_rn(-rint(acc * qinv), qd, acc);
}

__global__ __launch_bounds__(256) void k_contract(const int32_t *__restrict__ A, const double *__restrict__ B, int R,
                                                   int S, int K, int lda, int ldb, uint32_t q, int kb,
                                                   uint32_t *__restrict__ D)
{
    __shared__ double As[CT_BK][CT_BM + 1];
    __shared__ double Bs[CT_BK][CT_BN + 1];
    const int tid = threadIdx.x;
    const int tx = tid & 15, ty = tid >> 4;
    const int n0 = blockIdx.x * CT_BN, m0 = blockIdx.y * CT_BM;
    const double qd = (double)q, qinv = 1.0 / (double)q;

    double acc[8][8];
#pragma unroll
    for (int i = 0; i < 8; ++i)
#pragma unroll
        for (int j = 0; j < 8; ++j) acc[i][j] = 0.0;

    // loader mapping: 2 threads per row/key, 8 consecutive k each
    const int lrow = tid >> 1, lk = (tid & 1) * 8;
    const bool arow_ok = (m0 + lrow) < R;
    const bool bkey_ok = (n0 + lrow) < S;
    const int32_t *Ap = A + (size_t)(arow_ok ? m0 + lrow : 0) * lda + lk;
    const double *Bp = B + (size_t)(bkey_ok ? n0 + lrow : 0) * ldb + lk;

    int steps = 0;
    for (int k0 = 0; k0 < K; k0 += CT_BK) {
        {
            int4 a0 = *reinterpret_cast<const int4 *>(Ap + k0);
            int4 a1 = *reinterpret_cast<const int4 *>(Ap + k0 + 4);
            const int av[8] = {a0.x, a0.y, a0.z, a0.w, a1.x, a1.y, a1.z, a1.w};
#pragma unroll
            for (int t = 0; t < 8; ++t) As[lk + t][lrow] = arow_ok ? (double)av[t] : 0.0;
            const double2 *bp2 = reinterpret_cast<const double2 *>(Bp + k0);
#pragma unroll
            for (int t = 0; t < 4; ++t) {
                double2 b = bp2[t];
                Bs[lk + 2 * t][lrow] = bkey_ok ? b.x : 0.0;
                Bs[lk + 2 * t + 1][lrow] = bkey_ok ? b.y : 0.0;
            }
        }
        __syncthreads();
#pragma unroll
        for (int kk = 0; kk < CT_BK; ++kk) {
            double a[8], b[8];
#pragma unroll
            for (int i = 0; i < 8; ++i) a[i] = As[kk][ty + 16 * i];
#pragma unroll
            for (int j = 0; j < 8; ++j) b[j] = Bs[kk][tx + 16 * j];
#pragma unroll
            for (int i = 0; i < 8; ++i)
#pragma unroll
                for (int j = 0; j < 8; ++j) acc[i][j] = __fma_rn(a[i], b[j], acc[i][j]);
        }
        __syncthreads();
        if (++steps == kb) {
            steps = 0;
#pragma unroll
            for (int i = 0; i < 8; ++i)
#pragma unroll
                for (int j = 0; j < 8; ++j) acc[i][j] = reduce_bal(acc[i][j], qd, qinv);
        }
    }
#pragma unroll
    for (int i = 0; i < 8; ++i) {
        const int r = m0 + ty + 16 * i;
        if (r >= R) continue;
#pragma unroll
        for (int j = 0; j < 8; ++j) {
            const int v = n0 + tx + 16 * j;
            if (v >= S) continue;
            double x = reduce_bal(acc[i][j], qd, qinv);
            long long t = (long long)x;
            t %= (long long)q;
            if (t < 0) t += q;
            D[(size_t)r * S + v] = (uint32_t)t;
        }
    }
}


// ---------------------------------------------------------------------------
// Contraction with the row operand in scalar registers.
// Each wave owns 16 rows; their values at step k are wave-uniform, so they are
// read with scalar loads (two s_load_dwordx16 per k-pair from a row store laid
// out per 16-row block as [k][16] doubles) and enter v_fma_f64 as SGPR operands.
// Each lane owns 4 keys read from LDS (two ds_read_b128 per k).  Block tile:
// 64 rows (4 waves) x 256 keys; k-step 8 with the next B tile prefetched into
// registers.  LDS traffic is 0.5 B per FMA, a quarter of the LDS-only tiling.
// Tiles are ordered so that the blocks one XCD receives (block id mod 8) take
// consecutive tiles: the key tiles of one row block share that XCD's L2.
// ---------------------------------------------------------------------------
constexpr int SW_ROWS = 16, SW_KEYS = 4, SBK = 8, SBN = 256, SBM = 64;

__global__ __launch_bounds__(256) void k_contract_sgpr(const double *__restrict__ A, const double *__restrict__ B,
                                                        int R, int S, int K, int lda, int ldb, uint32_t q, int kb,
                                                        uint32_t *__restrict__ D, int ntn, int ntiles)
{
    __shared__ double Bs[2][SBK][SBN];
    const int tid = threadIdx.x, lane = tid & 63;
    const int w = __builtin_amdgcn_readfirstlane(tid >> 6);
    const int per = gridDim.x / 8;
    const int t = (int)(blockIdx.x % 8) * per + (int)(blockIdx.x / 8);
    if (t >= ntiles) return;
    const int tn = t % ntn, tm = t / ntn;
    const int n0 = tn * SBN, m0 = tm * SBM + w * SW_ROWS;
    const double qd = (double)q, qinv = 1.0 / (double)q;
    const double *Ablk = A + (size_t)(m0 / SW_ROWS) * lda * SW_ROWS;
    double acc[SW_ROWS][SW_KEYS];
#pragma unroll
    for (int i = 0; i < SW_ROWS; ++i)
#pragma unroll
        for (int j = 0; j < SW_KEYS; ++j) acc[i][j] = 0.0;
    const int bkey = n0 + tid;
    const bool bok = bkey < S;
    const double *Bp = B + (size_t)(bok ? bkey : 0) * ldb;
    double2 pre[4];
#pragma unroll
    for (int u = 0; u < 4; ++u) pre[u] = reinterpret_cast<const double2 *>(Bp)[u];
    int buf = 0, steps = 0;
    for (int k0 = 0; k0 < K; k0 += SBK) {
#pragma unroll
        for (int u = 0; u < 4; ++u) {
            Bs[buf][2 * u][tid] = bok ? pre[u].x : 0.0;
            Bs[buf][2 * u + 1][tid] = bok ? pre[u].y : 0.0;
        }
        __syncthreads();
        if (k0 + SBK < K) {
#pragma unroll
            for (int u = 0; u < 4; ++u) pre[u] = reinterpret_cast<const double2 *>(Bp + k0 + SBK)[u];
        }
#pragma unroll 1
        for (int kk = 0; kk < SBK; kk += 2) {
            const double *ap = Ablk + (size_t)(k0 + kk) * SW_ROWS;
            double a0[SW_ROWS], a1[SW_ROWS];
#pragma unroll
            for (int i = 0; i < SW_ROWS; ++i) {
                a0[i] = ap[i];
                a1[i] = ap[SW_ROWS + i];
            }
#pragma unroll
            for (int hh = 0; hh < 2; ++hh) {
                const double2 b01 = *reinterpret_cast<const double2 *>(&Bs[buf][kk + hh][lane * SW_KEYS]);
                const double2 b23 = *reinterpret_cast<const double2 *>(&Bs[buf][kk + hh][lane * SW_KEYS + 2]);
                const double bv[4] = {b01.x, b01.y, b23.x, b23.y};
#pragma unroll
                for (int i = 0; i < SW_ROWS; ++i) {
                    const double a = hh ? a1[i] : a0[i];
#pragma unroll
                    for (int j = 0; j < SW_KEYS; ++j) acc[i][j] = __fma_rn(a, bv[j], acc[i][j]);
                }
            }
        }
        buf ^= 1;
        if (++steps == kb) {
            steps = 0;
#pragma unroll
            for (int i = 0; i < SW_ROWS; ++i)
#pragma unroll
                for (int j = 0; j < SW_KEYS; ++j) acc[i][j] = reduce_bal(acc[i][j], qd, qinv);
        }
    }
#pragma unroll
    for (int i = 0; i < SW_ROWS; ++i) {
        const int r = m0 + i;
        if (r >= R) continue;
#pragma unroll
        for (int j = 0; j < SW_KEYS; ++j) {
            const int v = n0 + lane * SW_KEYS + j;
            if (v >= S) continue;
            const double x = reduce_bal(acc[i][j], qd, qinv);
            long long tt = (long long)x;
            tt %= (long long)q;
            if (tt < 0) tt += q;
            D[(size_t)r * S + v] = (uint32_t)tt;
        }
    }
}

// int32 row store [3][cap][ldr] -> blocked fp64 store [3][cap64/16][ldr][16]
// for rows [r0, r0 + R).
__global__ void k_rows_blocked(const int32_t *__restrict__ a32, int cap, int cap64, int ldr, int r0, int R,
                               double *__restrict__ a64)
{
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= (size_t)3 * R * ldr) return;
    const int s = (int)(i / ((size_t)R * ldr));
    const size_t rem = i % ((size_t)R * ldr);
    const int r = r0 + (int)(rem / ldr), k = (int)(rem % ldr);
    const size_t blk = (size_t)s * (cap64 / SW_ROWS) + r / SW_ROWS;
    a64[(blk * ldr + k) * SW_ROWS + (r % SW_ROWS)] = (double)a32[((size_t)s * cap + r) * ldr + k];
}

// Exact integer contraction for key sets whose coefficients are too large for
// exact fp64 partial sums: one thread per output, 64-bit products reduced one
// by one.
__global__ void k_contract_slow(const int32_t *__restrict__ A, const double *__restrict__ B, int R, int S, int K,
                                int lda, int ldb, int s, BarrettConsts bc, uint32_t *__restrict__ D)
{
    const size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (size_t)R * S) return;
    const int r = (int)(t / S), v = (int)(t % S);
    const uint32_t q = qmod(s);
    const int32_t *a = A + (size_t)r * lda;
    const double *b = B + (size_t)v * ldb;
    uint64_t acc = 0;
    for (int j = 0; j < K; ++j) {
        const uint32_t au = a[j] < 0 ? (uint32_t)(a[j] + (int64_t)q) : (uint32_t)a[j];
        const long long bl = (long long)b[j];
        const uint32_t bu = bl < 0 ? (uint32_t)(bl + (long long)q) : (uint32_t)bl;
        acc += reduce64((uint64_t)au * bu, q, bc.mu_hi[s], bc.mu_lo[s]);
    }
    D[t] = (uint32_t)(acc % q);
}

// ---------------------------------------------------------------------------
// kvalue[r][v] = round(p * CRT(d0) / Q) with d0_s = sub_mod(fold_i y_v[i] c0[r][i][s], D_s[r][v]).
// c0 [R][l][3]: coefficient 0 of c_0 .. c_{l-1} per row.  Sets *flag when a
// sum part ends on q and the product residue is 0 (the one case where the
// reference's representative of the product would matter).
// ---------------------------------------------------------------------------
__global__ void k_kvalue(const uint32_t *__restrict__ D, const uint32_t *__restrict__ c0, int l, SparseY y, int R,
                         int S, uint32_t *__restrict__ kval, int *__restrict__ flag)
{
    const size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (size_t)R * S) return;
    const int r = (int)(t / S), v = (int)(t % S);
    const uint32_t *cr = c0 + (size_t)r * l * 3;
    uint32_t d[3];
#pragma unroll
    for (int s = 0; s < 3; ++s) {
        const uint32_t q = qmod(s);
        uint32_t acc = 0;
        int last = -1;
        for (int e = y.off[v]; e < y.off[v + 1]; ++e) {
            const int i = y.idx[e];
            if (i > last + 1 && acc == q) acc = 0;
            const uint64_t ys = mod_prime_exact(y.val[e], s);
            acc = mod_prime_exact(ys * cr[(size_t)i * 3 + s] + acc, s);
            last = i;
        }
        if (last < y.len - 1 && acc == q) acc = 0;
        const uint32_t dd = D[((size_t)s * R + r) * S + v];
        if (acc == q && dd == 0) atomicOr(flag, 1);
        d[s] = sub_q(acc, dd, q);
    }
    kval[t] = crt_decode(d[0], d[1], d[2]);
}

// kvalue[r][v] = round(p * CRT(D_0, D_1, D_2) / Q) when the contraction already
// produced d0.
__global__ void k_crt_kvalue(const uint32_t *__restrict__ D, size_t n, uint32_t *__restrict__ kval)
{
    const size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n) return;
    kval[t] = crt_decode_canon(D[t], D[n + t], D[2 * n + t]);  // canonical residues
}

// ---------------------------------------------------------------------------
// Decision per row (main.c:338-389): products rounded, then summed in the
// reference's order, one thread per row.
// ---------------------------------------------------------------------------
__global__ void k_decide(const uint32_t *__restrict__ kval, int R, int S, SvmModelDev m, double *__restrict__ labels)
{
    const int r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= R) return;
    const uint32_t *kv = kval + (size_t)r * S;
    int vote[SVM_MAX_CLASS];
    for (int i = 0; i < m.nr_class; ++i) vote[i] = 0;
    int p = 0;
    for (int i = 0; i < m.nr_class; ++i)
        for (int j = i + 1; j < m.nr_class; ++j) {
            double sum = 0.0;
            const double *coef1 = m.sv_coef + (size_t)(j - 1) * S;
            const double *coef2 = m.sv_coef + (size_t)i * S;
            for (int k = 0; k < m.nSV[i]; ++k) {
                const int sv = m.start[i] + k;
                sum = __dadd_rn(sum, __dmul_rn(coef1[sv], (double)kv[sv]));
            }
            for (int k = 0; k < m.nSV[j]; ++k) {
                const int sv = m.start[j] + k;
                sum = __dadd_rn(sum, __dmul_rn(coef2[sv], (double)kv[sv]));
            }
            sum = __dsub_rn(sum, m.rho[p]);
            if (sum > 0) ++vote[i]; else ++vote[j];
            ++p;
        }
    int best = 0;
    for (int i = 1; i < m.nr_class; ++i)
        if (vote[i] > vote[best]) best = i;
    labels[r] = (double)m.label[best];
}

// ---------------------------------------------------------------------------
// The same decision with one wave per row.  The products are independent of
// the order, so the 64 lanes form them in parallel; only the sum is ordered.
// EXACT_INT (integer coefficients with sum |coef| * p < 2^53): every product
// and partial sum is an exact integer, so any order gives the same bits and
// the lanes reduce in a tree (with two classes, over 16-byte loads of the
// row).  Otherwise the lanes stage the rounded products
// of 512 terms at a time in LDS and lane 0 adds them in the reference order.
// ---------------------------------------------------------------------------
template <bool EXACT_INT>
__global__ __launch_bounds__(256) void k_decide_wave(const uint32_t *__restrict__ kval, int R, int S, SvmModelDev m,
                                                     double *__restrict__ labels)
{
    constexpr int CH = 512;
    __shared__ double prod[4][EXACT_INT ? 1 : CH];
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    const int r = blockIdx.x * 4 + wv;
    if (r >= R) return;  // whole wave; no block barriers below
    const uint32_t *kv = kval + (size_t)r * S;
    if (EXACT_INT && m.nr_class == 2 && (S & 3) == 0) {
        // Two classes: SV t (of either class) has coefficient sv_coef[t], so
        // the decision value is one dot product over the row; each lane loads
        // 4 consecutive key values (16 B) and their 4 coefficients (16 B as
        // int32, summed in 64-bit integers, else 32 B of doubles).
        const uint4 *kv4 = reinterpret_cast<const uint4 *>(kv);
        double part = 0.0;
        if (m.sv_icoef) {
            const int4 *c4 = reinterpret_cast<const int4 *>(m.sv_icoef);
            long long acc = 0;
#pragma unroll 4
            for (int t = lane; t < (S >> 2); t += 64) {
                const uint4 k = kv4[t];
                const int4 a = c4[t];
                acc += (long long)a.x * k.x + (long long)a.y * k.y + (long long)a.z * k.z + (long long)a.w * k.w;
            }
#pragma unroll
            for (int off = 32; off > 0; off >>= 1) acc += __shfl_xor(acc, off, 64);
            part = (double)acc;  // |acc| < 2^53: exact
        } else {
            const double2 *c2 = reinterpret_cast<const double2 *>(m.sv_coef);
#pragma unroll 4
            for (int t = lane; t < (S >> 2); t += 64) {
                const uint4 k = kv4[t];
                const double2 a = c2[2 * t], b = c2[2 * t + 1];
                part += (__dmul_rn(a.x, (double)k.x) + __dmul_rn(a.y, (double)k.y)) +
                        (__dmul_rn(b.x, (double)k.z) + __dmul_rn(b.y, (double)k.w));
            }
#pragma unroll
            for (int off = 32; off > 0; off >>= 1) part += __shfl_xor(part, off, 64);
        }
        if (lane == 0) labels[r] = (double)m.label[__dsub_rn(part, m.rho[0]) > 0 ? 0 : 1];
        return;
    }
    int vote[SVM_MAX_CLASS];
    for (int i = 0; i < m.nr_class; ++i) vote[i] = 0;
    int p = 0;
    for (int i = 0; i < m.nr_class; ++i)
        for (int j = i + 1; j < m.nr_class; ++j) {
            const double *coef1 = m.sv_coef + (size_t)(j - 1) * S;
            const double *coef2 = m.sv_coef + (size_t)i * S;
            const int n1 = m.nSV[i], total = m.nSV[i] + m.nSV[j];
            // term t: SV start[i] + t (coef1) for t < n1, else start[j] + t - n1 (coef2)
            auto term = [&](int t) {
                const int sv = t < n1 ? m.start[i] + t : m.start[j] + (t - n1);
                const double c = t < n1 ? coef1[sv] : coef2[sv];
                return __dmul_rn(c, (double)kv[sv]);
            };
            double sum = 0.0;
            if (EXACT_INT) {
                double part = 0.0;
                // several loads in flight per lane: small batches wait on
                // memory latency, not bandwidth
#pragma unroll 8
                for (int t = lane; t < total; t += 64) part += term(t);
#pragma unroll
                for (int off = 32; off > 0; off >>= 1) part += __shfl_xor(part, off, 64);
                sum = part;
            } else {
                for (int b0 = 0; b0 < total; b0 += CH) {
                    const int cnt = min(CH, total - b0);
                    for (int t = lane; t < cnt; t += 64) prod[wv][t] = term(b0 + t);
                    __builtin_amdgcn_wave_barrier();
                    if (lane == 0) {
#pragma unroll 16
                        for (int k = 0; k < cnt; ++k) sum = __dadd_rn(sum, prod[wv][k]);
                    }
                    __builtin_amdgcn_wave_barrier();
                }
            }
            sum = __dsub_rn(sum, m.rho[p]);
            if (sum > 0) ++vote[i]; else ++vote[j];
            ++p;
        }
    if (lane == 0) {
        int best = 0;
        for (int i = 1; i < m.nr_class; ++i)
            if (vote[i] > vote[best]) best = i;
        labels[r] = (double)m.label[best];
    }
}

// ---------------------------------------------------------------------------
// Compact encryption: coefficient 0 of c_i for i < l and every row,
//   c_i[0] = add_mod( N^-1 sum_k pk_i[k] r_hat[k], f_i[0] ) + floor(Q/p) x_i  (mod q)
// (coefficient 0 of GS_reverse is N^-1 times the sum of its inputs).  One block
// per (row, limb); r_hat staged in LDS; one wave per i.
// f0buf [R][l][3] holds f_i[0] on entry and c_i[0] on exit.
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_enc_coef0(const uint32_t *__restrict__ rhat, const uint32_t *__restrict__ mpk,
                                                    const uint32_t *__restrict__ x, int l, uint32_t *__restrict__ f0buf,
                                                    NinvConsts nc, ScaleConsts sc, BarrettConsts bc)
{
    __shared__ uint32_t rs[IPFE_N];
    const int r = blockIdx.x, s = blockIdx.y;
    const uint32_t q = qmod(s);
    const uint32_t *rh = rhat + ((size_t)r * 3 + s) * IPFE_N;
    for (int k = threadIdx.x; k < IPFE_N; k += blockDim.x) rs[k] = rh[k];
    __syncthreads();
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    for (int i = wave; i < l; i += blockDim.x / 64) {
        const uint32_t *pk = mpk + ((size_t)i * 3 + s) * IPFE_N;
        // partial sums of products below 2^62, reduced before they can overflow
        uint64_t acc = 0;
        for (int k = lane; k < IPFE_N; k += 64) {
            uint64_t pr = (uint64_t)pk[k] * rs[k];
            acc += reduce64(pr, q, bc.mu_hi[s], bc.mu_lo[s]);
        }
        // acc < 64 q < 2^37
        for (int d = 32; d >= 1; d >>= 1) acc += __shfl_xor(acc, d, 64);
        if (lane == 0) {
            uint32_t sum = (uint32_t)(acc % q);
            uint32_t v = mul_shoup(sum, nc.ninv[s], nc.ninvp[s], q);
            uint32_t *cell = f0buf + ((size_t)r * l + i) * 3 + s;
            uint32_t t = add_q(v, *cell, q);
            const uint32_t mj = mod_prime_exact(x[(size_t)r * l + i], s);
            const uint32_t madd = mod_prime_exact((uint64_t)mj * sc.scale[s], s);
            t += madd;
            *cell = t >= q ? t - q : t;
        }
    }
}

// Gather f_i[0] for every (row, i) from full noise polys is avoided: the compact
// sampler writes coefficient 0 only (see predict.cpp).

// ---------------------------------------------------------------------------
int launch_contract_keys(const uint32_t *sk, int S, int K, int sign, const SparseY &y, double *kd, unsigned int *maxabs,
                         hipStream_t st)
{
    if (S <= 0) return 0;
    hipLaunchKernelGGL(k_contract_keys, dim3((K + 255) / 256, S, 3), dim3(256), 0, st, sk, S, K, sign, y, kd, maxabs);
    HIP_OK(hipGetLastError());
    return 0;
}


int launch_contract(const int32_t *A, const double *B, int R, int S, int K, int lda, int ldb, int s, int kb, uint32_t *D,
                    hipStream_t st)
{
    if (R <= 0 || S <= 0) return 0;
    if (kb < 1 || K % CT_BK || lda % 4 || ldb % 2 || K > lda || K > ldb) {
        ipfe_set_error("contraction: bad shape K=%d lda=%d ldb=%d kb=%d", K, lda, ldb, kb);
        return -1;
    }
    dim3 grid((S + CT_BN - 1) / CT_BN, (R + CT_BM - 1) / CT_BM);
    hipLaunchKernelGGL(k_contract, grid, dim3(256), 0, st, A, B, R, S, K, lda, ldb, qmod(s), kb, D);
    HIP_OK(hipGetLastError());
    return 0;
}

int launch_contract_sgpr(const double *A, const double *B, int R, int S, int K, int lda, int ldb, int s, int kb,
                         uint32_t *D, hipStream_t st)
{
    if (R <= 0 || S <= 0) return 0;
    if (kb < 1 || K % SBK || ldb % 2 || K > lda || K > ldb) {
        ipfe_set_error("contraction: bad shape K=%d lda=%d ldb=%d kb=%d", K, lda, ldb, kb);
        return -1;
    }
    const int ntn = (S + SBN - 1) / SBN, ntm = (R + SBM - 1) / SBM;
    const int ntiles = ntn * ntm;
    const int grid = ((ntiles + 7) / 8) * 8;
    hipLaunchKernelGGL(k_contract_sgpr, dim3(grid), dim3(256), 0, st, A, B, R, S, K, lda, ldb, qmod(s), kb, D, ntn,
                       ntiles);
    HIP_OK(hipGetLastError());
    return 0;
}

int launch_rows_blocked(const int32_t *a32, int cap, int cap64, int ldr, int r0, int R, double *a64, hipStream_t st)
{
    const size_t n = (size_t)3 * R * ldr;
    if (n == 0) return 0;
    hipLaunchKernelGGL(k_rows_blocked, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, a32, cap, cap64, ldr, r0, R,
                       a64);
    HIP_OK(hipGetLastError());
    return 0;
}

int launch_contract_slow(const int32_t *A, const double *B, int R, int S, int K, int lda, int ldb, int s,
                         const BarrettConsts &bc, uint32_t *D, hipStream_t st)
{
    const size_t n = (size_t)R * S;
    if (n == 0) return 0;
    hipLaunchKernelGGL(k_contract_slow, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, A, B, R, S, K, lda, ldb, s,
                       bc, D);
    HIP_OK(hipGetLastError());
    return 0;
}

int launch_kvalue(const uint32_t *D, const uint32_t *c0, int l, const SparseY &y, int R, int S, uint32_t *kval,
                  int *flag, hipStream_t st)
{
    const size_t n = (size_t)R * S;
    if (n == 0) return 0;
    hipLaunchKernelGGL(k_kvalue, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, D, c0, l, y, R, S, kval, flag);
    HIP_OK(hipGetLastError());
    return 0;
}

int launch_crt_kvalue(const uint32_t *D, size_t n, uint32_t *kval, hipStream_t st)
{
    if (n == 0) return 0;
    hipLaunchKernelGGL(k_crt_kvalue, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, D, n, kval);
    HIP_OK(hipGetLastError());
    return 0;
}

int launch_decide(const uint32_t *kval, int R, int S, const SvmModelDev &m, double *labels, hipStream_t st)
{
    if (R <= 0) return 0;
    const char *e = getenv("CUFE_DECIDE");
    if (e && strcmp(e, "thread") == 0)
        hipLaunchKernelGGL(k_decide, dim3((R + 127) / 128), dim3(128), 0, st, kval, R, S, m, labels);
    else if (m.exact_int)
        hipLaunchKernelGGL(k_decide_wave<true>, dim3((R + 3) / 4), dim3(256), 0, st, kval, R, S, m, labels);
    else
        hipLaunchKernelGGL(k_decide_wave<false>, dim3((R + 3) / 4), dim3(256), 0, st, kval, R, S, m, labels);
    HIP_OK(hipGetLastError());
    return 0;
}

int launch_enc_coef0(const uint32_t *rhat, const uint32_t *mpk, const uint32_t *x, int R, int l, uint32_t *f0buf,
                     const NinvConsts &nc, const ScaleConsts &sc, const BarrettConsts &bc, hipStream_t st)
{
    if (R <= 0 || l <= 0) return 0;
    hipLaunchKernelGGL(k_enc_coef0, dim3(R, 3), dim3(256), 0, st, rhat, mpk, x, l, f0buf, nc, sc, bc);
    HIP_OK(hipGetLastError());
    return 0;
}

}  // namespace ipfe
