// k_micro.hip -- microbenchmarks that give the measured ceilings the roofline
// fractions are read against: fp64 FMA rate, 32-bit integer multiply rate,
// HBM copy bandwidth and raw AES-256-CTR keystream rate on this GPU.
#include <stdio.h>

#include <vector>

#include "../../include/cufe_bench.h"
#include "../../include/cufe_ipfe.h"
#include "aes_dev.h"
#include "ntt_core.h"
#include "runtime.h"

namespace ipfe {

// 16 independent FMA chains per thread; the result is stored so nothing is
// optimised away.
__global__ __launch_bounds__(256) void k_mb_fma64(double *out, int iters, double b, double c)
{
    double acc[16];
#pragma unroll
    for (int i = 0; i < 16; ++i) acc[i] = (double)(threadIdx.x + i);
    for (int it = 0; it < iters; ++it) {
#pragma unroll
        for (int i = 0; i < 16; ++i) acc[i] = __fma_rn(acc[i], b, c);
    }
    double s = 0;
#pragma unroll
    for (int i = 0; i < 16; ++i) s += acc[i];
    out[blockIdx.x * blockDim.x + threadIdx.x] = s;
}

// 8 independent chains of (mul_lo, mul_hi, add): 3 int32 ops per step.
__global__ __launch_bounds__(256) void k_mb_mul32(uint32_t *out, int iters, uint32_t b)
{
    uint32_t x[8], y[8];
#pragma unroll
    for (int i = 0; i < 8; ++i) {
        x[i] = threadIdx.x * 2654435761u + i;
        y[i] = i;
    }
    for (int it = 0; it < iters; ++it) {
#pragma unroll
        for (int i = 0; i < 8; ++i) {
            uint32_t lo = x[i] * b;
            uint32_t hi = __umulhi(x[i], b);
            x[i] = lo + y[i];
            y[i] = hi;
        }
    }
    uint32_t s = 0;
#pragma unroll
    for (int i = 0; i < 8; ++i) s ^= x[i] ^ y[i];
    out[blockIdx.x * blockDim.x + threadIdx.x] = s;
}

__global__ void k_mb_copy(const uint4 *__restrict__ in, uint4 *__restrict__ out, size_t n)
{
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
        out[i] = in[i];
}

__global__ __launch_bounds__(256) void k_mb_aes(const uint32_t *__restrict__ rk, const uint32_t *__restrict__ te_global,
                                                uint32_t nblocks_per_thread, uint32_t *__restrict__ out)
{
    __shared__ uint32_t te[AES_TABLE_WORDS];
    aes_stage_tables(te, te_global);
    __syncthreads();
    uint32_t x = 0;
    const uint32_t base = (blockIdx.x * blockDim.x + threadIdx.x) * nblocks_per_thread;
    for (uint32_t i = 0; i < nblocks_per_thread; ++i) {
        uint32_t o[4];
        aes256_ctr_block(rk, te, 7u, base + i, o);
        x ^= o[0] ^ o[1] ^ o[2] ^ o[3];
    }
    out[blockIdx.x * blockDim.x + threadIdx.x] = x;
}

__global__ __launch_bounds__(256) void k_mb_aes_rep(const uint32_t *__restrict__ rk,
                                                    const uint32_t *__restrict__ te_global, uint32_t nblocks_per_thread,
                                                    uint32_t *__restrict__ out)
{
    __shared__ uint32_t te[AES_REP_WORDS];
    aes_stage_tables_rep(te, te_global);
    __syncthreads();
    uint32_t x = 0;
    const uint32_t base = (blockIdx.x * blockDim.x + threadIdx.x) * nblocks_per_thread;
    for (uint32_t i = 0; i < nblocks_per_thread; ++i) {
        uint32_t o[4];
        aes256_ctr_block_rep(rk, te, 7u, base + i, o);
        x ^= o[0] ^ o[1] ^ o[2] ^ o[3];
    }
    out[blockIdx.x * blockDim.x + threadIdx.x] = x;
}

// T0 and T2 bank-replicated (64 KiB): the sampler's default AES, in
// 512-thread workgroups as the sampler runs it (two per CU).
__global__ __launch_bounds__(512) void k_mb_aes_rep2(const uint32_t *__restrict__ rk,
                                                     const uint32_t *__restrict__ te_global,
                                                     uint32_t nblocks_per_thread, uint32_t *__restrict__ out)
{
    __shared__ uint32_t te[AES_REP2_WORDS];
    aes_stage_tables_rep2(te, te_global);
    __syncthreads();
    const AesTabRep2 tb(te);
    uint32_t x = 0;
    const uint32_t base = (blockIdx.x * blockDim.x + threadIdx.x) * nblocks_per_thread;
    for (uint32_t i = 0; i < nblocks_per_thread; ++i) {
        uint32_t o[4];
        aes256_ctr_block_t(tb, rk, 7u, base + i, o);
        x ^= o[0] ^ o[1] ^ o[2] ^ o[3];
    }
    out[blockIdx.x * blockDim.x + threadIdx.x] = x;
}

typedef int mb_v4i __attribute__((ext_vector_type(4)));
typedef int mb_v16i __attribute__((ext_vector_type(16)));
typedef double mb_v4d __attribute__((ext_vector_type(4)));

// Two independent v_mfma_i32_32x32x32_i8 chains per wave: the int8 matrix
// rate with operands in registers.
__global__ __launch_bounds__(256) void k_mb_mfma_i8(int *out, int iters)
{
    const mb_v4i a = {(int)threadIdx.x, 1, 2, 3}, b = {3, 2, 1, (int)threadIdx.x};
    mb_v16i c0 = {}, c1 = {};
    for (int it = 0; it < iters; ++it) {
        c0 = __builtin_amdgcn_mfma_i32_32x32x32_i8(a, b, c0, 0, 0, 0);
        c1 = __builtin_amdgcn_mfma_i32_32x32x32_i8(a, b, c1, 0, 0, 0);
    }
    int x = 0;
#pragma unroll
    for (int r = 0; r < 16; ++r) x ^= c0[r] ^ c1[r];
    out[blockIdx.x * blockDim.x + threadIdx.x] = x;
}

// Four v_mfma_i32_32x32x32_i8 chains per wave on random operands, each MFMA
// taking a different operand pair than the one before it: the int8 rate with
// the data-dependent switching of real digits (constant operands draw less
// power, and the clock stays higher).
__global__ __launch_bounds__(256) void k_mb_mfma_i8_rand(const mb_v4i *__restrict__ src, int *out, int iters)
{
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    const mb_v4i a0 = src[4 * t], b0 = src[4 * t + 1], a1 = src[4 * t + 2], b1 = src[4 * t + 3];
    mb_v16i c0 = {}, c1 = {}, c2 = {}, c3 = {};
    for (int it = 0; it < iters; ++it) {
        c0 = __builtin_amdgcn_mfma_i32_32x32x32_i8(a0, b0, c0, 0, 0, 0);
        c1 = __builtin_amdgcn_mfma_i32_32x32x32_i8(a1, b1, c1, 0, 0, 0);
        c2 = __builtin_amdgcn_mfma_i32_32x32x32_i8(a0, b1, c2, 0, 0, 0);
        c3 = __builtin_amdgcn_mfma_i32_32x32x32_i8(a1, b0, c3, 0, 0, 0);
    }
    int x = 0;
#pragma unroll
    for (int r = 0; r < 16; ++r) x ^= c0[r] ^ c1[r] ^ c2[r] ^ c3[r];
    out[t] = x;
}

// Four independent v_mfma_f64_16x16x4_f64 chains per wave.
__global__ __launch_bounds__(256) void k_mb_mfma_f64(double *out, int iters)
{
    const double a = threadIdx.x * 1e-3, b = 0.999;
    mb_v4d c0 = {}, c1 = {}, c2 = {}, c3 = {};
    for (int it = 0; it < iters; ++it) {
        c0 = __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c0, 0, 0, 0);
        c1 = __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c1, 0, 0, 0);
        c2 = __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c2, 0, 0, 0);
        c3 = __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c3, 0, 0, 0);
    }
    const mb_v4d sum = c0 + c1 + c2 + c3;
    out[blockIdx.x * blockDim.x + threadIdx.x] = sum[0] + sum[1] + sum[2] + sum[3];
}

// Register-resident radix-16 CT butterflies (the NTT's butterfly code, twiddle
// pairs held in registers): the butterfly rate the NTT is measured against.
__global__ __launch_bounds__(256) void k_mb_bfly(uint32_t *out, int iters, uint32_t w, uint32_t wp, uint32_t q)
{
    uint32_t x[16];
#pragma unroll
    for (int r = 0; r < 16; ++r) x[r] = (threadIdx.x * 2654435761u + 977u * r) % q;
    for (int it = 0; it < iters; ++it) {
#pragma unroll
        for (int h = 8; h >= 1; h >>= 1)
#pragma unroll
            for (int r = 0; r < 16; ++r)
                if ((r & h) == 0) ct_bfly<false>(x[r], x[r + h], w, wp, q);
    }
    uint32_t acc = 0;
#pragma unroll
    for (int r = 0; r < 16; ++r) acc ^= x[r];
    out[blockIdx.x * blockDim.x + threadIdx.x] = acc;
}

// Checks that the AES variants produce the keystream words of the four-table
// AES: the replicated T0 table (rep2 = 0) or T0 and T2 (rep2 = 1), whole
// blocks and counter windows (the window of the block's counter is computed by
// the same lane).
__global__ void k_mb_aes_check(const uint32_t *__restrict__ rk, const uint32_t *__restrict__ te_global, int rep2,
                               uint32_t *__restrict__ bad)
{
    __shared__ uint32_t te[AES_TABLE_WORDS];
    __shared__ uint32_t rep[AES_REP2_WORDS];
    aes_stage_tables(te, te_global);
    if (rep2)
        aes_stage_tables_rep2(rep, te_global);
    else
        aes_stage_tables_rep(rep, te_global);
    __syncthreads();
    uint32_t o1[4], o2[4], o3[4], w[8];
    const uint32_t ctr = blockIdx.x * blockDim.x + threadIdx.x;
    aes256_ctr_block(rk, te, 9u, ctr, o1);
    if (rep2) {
        const AesTabRep2 tb(rep);
        aes256_ctr_block_t(tb, rk, 9u, ctr, o2);
        aes256_ctr_window(tb, rk, 9u, ctr >> 8, w);
        aes256_ctr_block_win(tb, rk, w, ctr, o3);
    } else {
        const AesTabRep tb(rep);
        aes256_ctr_block_rep(rk, rep, 9u, ctr, o2);
        aes256_ctr_window(tb, rk, 9u, ctr >> 8, w);
        aes256_ctr_block_win(tb, rk, w, ctr, o3);
    }
    bool ok = true;
#pragma unroll
    for (int j = 0; j < 4; ++j) ok = ok && o1[j] == o2[j] && o1[j] == o3[j];
    if (!ok) atomicAdd(bad, 1u);
}

}  // namespace ipfe

using namespace ipfe;

// kind 0: fp64 FMA (TFLOP/s), 1: int32 mul_lo+mul_hi+add (T int-ops/s),
// 2: device copy (GB/s of read+write), 3: AES-256 CTR blocks (G blocks/s),
// 4: the same with the bank-replicated table (and an equality check),
// 5: int8 MFMA 32x32x32 (T ops/s), 6: fp64 MFMA 16x16x4 (TFLOP/s),
// 7: NTT butterflies in registers (G butterflies/s), 8: int8 MFMA on random
// operands (T ops/s), 9: AES with T0 and T2 replicated (G blocks/s, and an
// equality check).
extern "C" int cufe_microbench(int kind, double *result)
{
    Device *dev = device();
    if (!dev) return -1;
    hipStream_t st = dev->stream;
    hipEvent_t a, b;
    HIP_OK(hipEventCreate(&a));
    HIP_OK(hipEventCreate(&b));
    int cus = 256;
    (void)hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, 0);
    const int blocks = cus * 8;
    float ms = 0;
    double work = 0;
    if (kind == 8) {
        // random operands from a fixed xorshift stream; 8 launches of ~5 ms
        // back to back so the clock settles to its sustained value
        const size_t nthreads = (size_t)blocks * 256;
        std::vector<uint32_t> h(nthreads * 16);
        uint32_t x = 2463534242u;
        for (auto &v : h) {
            x ^= x << 13;
            x ^= x >> 17;
            x ^= x << 5;
            v = x;
        }
        DevBuf<uint32_t> src, out;
        if (src.alloc(h.size()) || out.alloc(nthreads)) return -1;
        HIP_OK(hipMemcpy(src.get(), h.data(), h.size() * 4, hipMemcpyHostToDevice));
        const int iters = 4096, reps = 8;
        for (int rep = 0; rep < 2; ++rep) {
            HIP_OK(hipEventRecord(a, st));
            for (int r = 0; r < reps; ++r)
                hipLaunchKernelGGL(k_mb_mfma_i8_rand, dim3(blocks), dim3(256), 0, st,
                                   reinterpret_cast<const mb_v4i *>(src.get()), (int *)out.get(), iters);
            HIP_OK(hipGetLastError());
            HIP_OK(hipEventRecord(b, st));
            HIP_OK(hipEventSynchronize(b));
        }
        HIP_OK(hipEventElapsedTime(&ms, a, b));
        work = (double)blocks * 4 * iters * 4 * reps * (32.0 * 32 * 32 * 2);
    } else if (kind == 7) {
        DevBuf<uint32_t> out;
        if (out.alloc((size_t)blocks * 256)) return -1;
        const int iters = 2048;
        const uint32_t q = qmod(1), w = 1689795u;  // a limb-1 root and its Shoup companion
        const uint32_t wp = (uint32_t)(((uint64_t)w << 32) / q);
        for (int rep = 0; rep < 2; ++rep) {
            HIP_OK(hipEventRecord(a, st));
            hipLaunchKernelGGL(k_mb_bfly, dim3(blocks), dim3(256), 0, st, out.get(), iters, w, wp, q);
            HIP_OK(hipGetLastError());
            HIP_OK(hipEventRecord(b, st));
            HIP_OK(hipEventSynchronize(b));
        }
        HIP_OK(hipEventElapsedTime(&ms, a, b));
        work = (double)blocks * 256 * iters * 32;  // butterflies
    } else if (kind == 5 || kind == 6) {
        DevBuf<double> out;
        if (out.alloc((size_t)blocks * 256)) return -1;
        const int iters = 4096;
        for (int rep = 0; rep < 2; ++rep) {
            HIP_OK(hipEventRecord(a, st));
            if (kind == 5)
                hipLaunchKernelGGL(k_mb_mfma_i8, dim3(blocks), dim3(256), 0, st, (int *)out.get(), iters);
            else
                hipLaunchKernelGGL(k_mb_mfma_f64, dim3(blocks), dim3(256), 0, st, out.get(), iters);
            HIP_OK(hipGetLastError());
            HIP_OK(hipEventRecord(b, st));
            HIP_OK(hipEventSynchronize(b));
        }
        HIP_OK(hipEventElapsedTime(&ms, a, b));
        const double waves = (double)blocks * 4;
        work = kind == 5 ? waves * iters * 2 * (32.0 * 32 * 32 * 2) : waves * iters * 4 * (16.0 * 16 * 4 * 2);
    } else if (kind == 0 || kind == 1 || kind == 3 || kind == 4 || kind == 9) {
        DevBuf<uint32_t> out;
        if (out.alloc((size_t)blocks * 256 * 2)) return -1;
        const int iters = 4096;
        DevBuf<uint32_t> rk;
        uint32_t hrk[60];
        uint8_t key[32];
        for (int i = 0; i < 32; ++i) key[i] = (uint8_t)i;
        aes256_expand_key(key, hrk);
        if (rk.alloc(60)) return -1;
        HIP_OK(hipMemcpy(rk.get(), hrk, sizeof hrk, hipMemcpyHostToDevice));
        for (int rep = 0; rep < 2; ++rep) {
            HIP_OK(hipEventRecord(a, st));
            if (kind == 0)
                hipLaunchKernelGGL(k_mb_fma64, dim3(blocks), dim3(256), 0, st, (double *)out.get(), iters, 0.999999,
                                   1e-9);
            else if (kind == 1)
                hipLaunchKernelGGL(k_mb_mul32, dim3(blocks), dim3(256), 0, st, out.get(), iters, 2654435761u);
            else if (kind == 3)
                hipLaunchKernelGGL(k_mb_aes, dim3(blocks), dim3(256), 0, st, rk.get(), g_aes_te, 64u, out.get());
            else if (kind == 4)
                hipLaunchKernelGGL(k_mb_aes_rep, dim3(blocks), dim3(256), 0, st, rk.get(), g_aes_te, 64u, out.get());
            else
                hipLaunchKernelGGL(k_mb_aes_rep2, dim3(blocks / 2), dim3(512), 0, st, rk.get(), g_aes_te, 64u,
                                   out.get());
            HIP_OK(hipGetLastError());
            HIP_OK(hipEventRecord(b, st));
            HIP_OK(hipEventSynchronize(b));
        }
        HIP_OK(hipEventElapsedTime(&ms, a, b));
        const double threads = (double)blocks * 256;
        if (kind == 0) work = threads * iters * 16 * 2;       // FLOPs
        else if (kind == 1) work = threads * iters * 8 * 3;   // int ops
        else work = threads * 64;                             // AES blocks
        if (kind == 4 || kind == 9) {
            DevBuf<uint32_t> bad;
            uint32_t hb = 0;
            if (bad.alloc(1)) return -1;
            HIP_OK(hipMemset(bad.get(), 0, 4));
            hipLaunchKernelGGL(k_mb_aes_check, dim3(64), dim3(256), 0, st, rk.get(), g_aes_te, kind == 9 ? 1 : 0,
                               bad.get());
            HIP_OK(hipMemcpy(&hb, bad.get(), 4, hipMemcpyDeviceToHost));
            if (hb) {
                ipfe_set_error("bank-replicated AES disagrees with the T-table AES on %u blocks", hb);
                return -1;
            }
        }
    } else if (kind == 2) {
        const size_t bytes = (size_t)1 << 30;
        DevBuf<uint4> x, y;
        if (x.alloc(bytes / 16) || y.alloc(bytes / 16)) return -1;
        HIP_OK(hipMemset(x.get(), 1, bytes));
        for (int rep = 0; rep < 3; ++rep) {
            HIP_OK(hipEventRecord(a, st));
            hipLaunchKernelGGL(k_mb_copy, dim3(cus * 16), dim3(256), 0, st, x.get(), y.get(), bytes / 16);
            HIP_OK(hipGetLastError());
            HIP_OK(hipEventRecord(b, st));
            HIP_OK(hipEventSynchronize(b));
        }
        HIP_OK(hipEventElapsedTime(&ms, a, b));
        work = 2.0 * bytes;
    } else {
        ipfe_set_error("cufe_microbench: unknown kind %d", kind);
        return -1;
    }
    (void)hipEventDestroy(a);
    (void)hipEventDestroy(b);
    const double s = ms * 1e-3;
    *result = (kind == 0 || kind == 1 || kind == 5 || kind == 6 || kind == 8) ? work / s / 1e12 : work / s / 1e9;
    return 0;
}
