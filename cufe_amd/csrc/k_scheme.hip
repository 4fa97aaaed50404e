// k_scheme.hip -- functional-key generation, the sum part of decryption and
// CRT + rounding, bit-exact with src/rlwe_sife.c:215-393 and src/crt.c:46-82.
//
// Keygen and the first half of decryption fold  acc <- mod_prime(y_i v_i + acc)
// over every slot i < L in order (src/rlwe_sife.c:232-240, :265-273).  A slot
// with y_i = 0 only maps acc = q to 0, so the kernels visit the non-zero y_i of
// a sparse list and canonicalise acc wherever zero slots were skipped; the
// result is identical to the dense loop.
#include <algorithm>

#include "crt_dev.h"
#include "kernels.h"

namespace ipfe {

__device__ __forceinline__ uint32_t sparse_fold(const int *__restrict__ idx, const uint32_t *__restrict__ val, int e0,
                                                int e1, int len, const uint32_t *__restrict__ base, size_t slot_stride,
                                                int s)
{
    const uint32_t q = qmod(s);
    uint32_t acc = 0;
    int last = -1;
    for (int e = e0; e < e1; ++e) {
        const int i = idx[e];
        if (i > last + 1 && acc == q) acc = 0;
        const uint64_t ys = mod_prime_exact(val[e], s);
        acc = mod_prime_exact(ys * base[(size_t)i * slot_stride] + acc, s);
        last = i;
    }
    if (last < len - 1 && acc == q) acc = 0;
    return acc;
}

// sk[v][s][k]; grid (N/256 * 3, nvec)
__global__ void k_keygen(SparseY y, const uint32_t *__restrict__ msk, uint32_t *__restrict__ sk)
{
    const int v = blockIdx.y;
    const int s = blockIdx.x / (IPFE_N / 256);
    const int k = (blockIdx.x % (IPFE_N / 256)) * 256 + threadIdx.x;
    const uint32_t *base = msk + (size_t)s * IPFE_N + k;
    sk[((size_t)v * 3 + s) * IPFE_N + k] = sparse_fold(y.idx, y.val, y.off[v], y.off[v + 1], y.len, base, 3 * IPFE_N, s);
}

// d[v][s][k] = fold_i y_v[i] c_i[s][k] for one ciphertext c [len+1][3][N].
__global__ void k_dec_sum(SparseY y, const uint32_t *__restrict__ c, uint32_t *__restrict__ d)
{
    const int v = blockIdx.y;
    const int s = blockIdx.x / (IPFE_N / 256);
    const int k = (blockIdx.x % (IPFE_N / 256)) * 256 + threadIdx.x;
    const uint32_t *base = c + (size_t)s * IPFE_N + k;
    d[((size_t)v * 3 + s) * IPFE_N + k] = sparse_fold(y.idx, y.val, y.off[v], y.off[v + 1], y.len, base, 3 * IPFE_N, s);
}

__global__ void k_poly_sub(uint32_t *__restrict__ d, const uint32_t *__restrict__ prod, size_t n_total)
{
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n_total) return;
    const uint32_t q = qmod((int)((i / IPFE_N) % 3));
    d[i] = sub_q(d[i], prod[i], q);
}

__global__ void k_crt_round(const uint32_t *__restrict__ d, int nvec, int coef, uint32_t *__restrict__ r)
{
    const int v = blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= nvec) return;
    const uint32_t *dv = d + (size_t)v * 3 * IPFE_N + coef;
    r[v] = crt_decode(dv[0], dv[IPFE_N], dv[2 * IPFE_N]);
}

int launch_keygen(const SparseY &y, int nvec, const uint32_t *msk, uint32_t *sk, hipStream_t st)
{
    if (nvec <= 0) return 0;
    hipLaunchKernelGGL(k_keygen, dim3(3 * IPFE_N / 256, nvec), dim3(256), 0, st, y, msk, sk);
    HIP_OK(hipGetLastError());
    return 0;
}

int launch_dec_sum(const SparseY &y, int nvec, const uint32_t *c, uint32_t *d, hipStream_t st)
{
    if (nvec <= 0) return 0;
    hipLaunchKernelGGL(k_dec_sum, dim3(3 * IPFE_N / 256, nvec), dim3(256), 0, st, y, c, d);
    HIP_OK(hipGetLastError());
    return 0;
}

int launch_poly_sub(uint32_t *d, const uint32_t *prod, size_t n_total, hipStream_t st)
{
    if (n_total == 0) return 0;
    hipLaunchKernelGGL(k_poly_sub, dim3((unsigned)((n_total + 255) / 256)), dim3(256), 0, st, d, prod, n_total);
    HIP_OK(hipGetLastError());
    return 0;
}

int launch_crt_round(const uint32_t *d, int nvec, int coef, uint32_t *r, hipStream_t st)
{
    if (nvec <= 0) return 0;
    hipLaunchKernelGGL(k_crt_round, dim3((nvec + 255) / 256), dim3(256), 0, st, d, nvec, coef, r);
    HIP_OK(hipGetLastError());
    return 0;
}

// Position-weighted checksum of a word array: sum of (2i+1) * w_i (mod 2^64),
// so a moved, dropped or changed word changes it.  Grid-stride; one 64-bit
// atomic add per workgroup.
__global__ __launch_bounds__(256) void k_checksum(const uint32_t *__restrict__ w, size_t n,
                                                  unsigned long long *__restrict__ out)
{
    __shared__ unsigned long long part[256];
    unsigned long long acc = 0;
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
        acc += (2ull * i + 1ull) * (unsigned long long)w[i];
    part[threadIdx.x] = acc;
    __syncthreads();
    for (int off = 128; off > 0; off >>= 1) {
        if ((int)threadIdx.x < off) part[threadIdx.x] += part[threadIdx.x + off];
        __syncthreads();
    }
    if (threadIdx.x == 0) atomicAdd(out, part[0]);
}

int launch_checksum(const uint32_t *w, size_t n, unsigned long long *out, hipStream_t st)
{
    if (n == 0) return 0;
    const size_t blocks = std::min<size_t>((n + 255) / 256, 4096);
    hipLaunchKernelGGL(k_checksum, dim3((unsigned)blocks), dim3(256), 0, st, w, n, out);
    HIP_OK(hipGetLastError());
    return 0;
}

}  // namespace ipfe
