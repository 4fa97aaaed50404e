// k_ntt.hip -- NTT-domain kernels: plain forward/inverse transforms, the fused
// public-key kernel of setup, and the fused ciphertext kernel of encryption.
#include "kernels.h"
#include "ntt_core.h"

namespace ipfe {

__device__ __forceinline__ const uint32_t *tw_fwd(const uint32_t *tw_all, int s) { return tw_all + (0 * 3 + s) * IPFE_N; }
__device__ __forceinline__ const uint32_t *twp_fwd(const uint32_t *tw_all, int s) { return tw_all + (1 * 3 + s) * IPFE_N; }
__device__ __forceinline__ const uint32_t *tw_inv(const uint32_t *tw_all, int s) { return tw_all + (2 * 3 + s) * IPFE_N; }
__device__ __forceinline__ const uint32_t *twp_inv(const uint32_t *tw_all, int s) { return tw_all + (3 * 3 + s) * IPFE_N; }

__device__ __forceinline__ bool any_eq_q(const uint32_t (&x)[16], uint32_t q)
{
    bool f = false;
#pragma unroll
    for (int r = 0; r < 16; ++r) f |= (x[r] == q);
    return f;
}

// ---------------------------------------------------------------------------
// Plain transforms over polys laid out [npoly][3][N]: block z works on limb z % 3.
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(NTT_THREADS) void k_ntt_fwd(const uint32_t *__restrict__ in, uint32_t *__restrict__ out,
                                                          const uint32_t *__restrict__ tw_all)
{
    __shared__ uint32_t lds[NTT_LDS_WORDS];
    const int z = blockIdx.x, s = z % 3, tid = threadIdx.x;
    const uint32_t q = qmod(s);
    const uint32_t *src = in + (size_t)z * IPFE_N;
    uint32_t x[16];
#pragma unroll
    for (int r = 0; r < 16; ++r) x[r] = src[posA(tid, r)];
    if (__syncthreads_or(any_eq_q(x, q)))
        ntt_fwd_block<true>(x, lds, tw_fwd(tw_all, s), twp_fwd(tw_all, s), q);
    else
        ntt_fwd_block<false>(x, lds, tw_fwd(tw_all, s), twp_fwd(tw_all, s), q);
    uint32_t *dst = out + (size_t)z * IPFE_N;
#pragma unroll
    for (int r = 0; r < 16; ++r) dst[posA(tid, r)] = x[r];
}

__global__ __launch_bounds__(NTT_THREADS) void k_ntt_inv(const uint32_t *__restrict__ in, uint32_t *__restrict__ out,
                                                          const uint32_t *__restrict__ tw_all, NinvConsts nc)
{
    __shared__ uint32_t lds[NTT_LDS_WORDS];
    const int z = blockIdx.x, s = z % 3, tid = threadIdx.x;
    const uint32_t q = qmod(s);
    const uint32_t *src = in + (size_t)z * IPFE_N;
    uint32_t x[16];
#pragma unroll
    for (int r = 0; r < 16; ++r) x[r] = src[posA(tid, r)];
    uint32_t *dst = out + (size_t)z * IPFE_N;
    if (__syncthreads_or(any_eq_q(x, q))) {
        ntt_inv_block<true>(x, lds, tw_inv(tw_all, s), twp_inv(tw_all, s), q);
#pragma unroll
        for (int r = 0; r < 16; ++r) dst[posA(tid, r)] = scale_ninv<true>(x[r], nc.ninv[s], nc.ninvp[s], nc.flag[s], q);
    } else {
        ntt_inv_block<false>(x, lds, tw_inv(tw_all, s), twp_inv(tw_all, s), q);
#pragma unroll
        for (int r = 0; r < 16; ++r) dst[posA(tid, r)] = scale_ninv<false>(x[r], nc.ninv[s], nc.ninvp[s], nc.flag[s], q);
    }
}

// ---------------------------------------------------------------------------
// Shoup companions floor(v * 2^32 / q) of a residue array [npoly][3][N], for
// operands that are reused as the fixed factor of many pointwise products.
// ---------------------------------------------------------------------------
__global__ void k_shoup_prep(const uint32_t *__restrict__ v, uint32_t *__restrict__ vp, size_t n_total)
{
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n_total) return;
    const int s = (int)((i / IPFE_N) % 3);
    const uint64_t q = qmod(s);
    uint64_t w = v[i];
    if (w >= q) w -= q;  // the value q stands for 0 in the fast product
    vp[i] = (uint32_t)((w << 32) / q);
}

// pointwise mul_mod(a, b) for b with Shoup companion bp; exact for a, b in [0, q].
template <bool EXACT>
__device__ __forceinline__ uint32_t pmul(uint32_t a, uint32_t b, uint32_t bp, int s)
{
    const uint32_t q = qmod(s);
    if (!EXACT) return mul_shoup(a, b, bp, q);
    if (a == q || b == q) return (a == 0 || b == 0) ? 0u : mod_prime_exact((uint64_t)a * b, s);
    return mul_shoup(a, b, bp, q);
}

// ---------------------------------------------------------------------------
// Setup (src/rlwe_sife.c:61-77): for slot i and limb s
//   pk_i = a_hat (.) NTT(s_i) + NTT(e_i)
// msk [l][3][N] coefficient domain (read), e [l][3][N] (read), ahat/ahatp [3][N].
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(NTT_THREADS) void k_setup_pk(const uint32_t *__restrict__ msk, const uint32_t *__restrict__ e,
                                                           const uint32_t *__restrict__ ahat,
                                                           const uint32_t *__restrict__ ahatp,
                                                           uint32_t *__restrict__ mpk, const uint32_t *__restrict__ tw_all)
{
    __shared__ uint32_t lds[NTT_LDS_WORDS];
    const int z = blockIdx.x, s = z % 3, tid = threadIdx.x;
    const uint32_t q = qmod(s);
    const uint32_t *tw = tw_fwd(tw_all, s), *twp = twp_fwd(tw_all, s);
    uint32_t x[16], y[16];
#pragma unroll
    for (int r = 0; r < 16; ++r) x[r] = msk[(size_t)z * IPFE_N + posA(tid, r)];
#pragma unroll
    for (int r = 0; r < 16; ++r) y[r] = e[(size_t)z * IPFE_N + posA(tid, r)];
    uint32_t av[16], avp[16];
#pragma unroll
    for (int r = 0; r < 16; ++r) {
        av[r] = ahat[s * IPFE_N + posA(tid, r)];
        avp[r] = ahatp[s * IPFE_N + posA(tid, r)];
    }
    const bool exact = __syncthreads_or(any_eq_q(x, q) || any_eq_q(y, q) || any_eq_q(av, q));
    if (exact) {
        ntt_fwd_block<true>(x, lds, tw, twp, q);
        ntt_fwd_block<true>(y, lds, tw, twp, q);
#pragma unroll
        for (int r = 0; r < 16; ++r) x[r] = add_q(pmul<true>(x[r], av[r], avp[r], s), y[r], q);
    } else {
        ntt_fwd_block<false>(x, lds, tw, twp, q);
        ntt_fwd_block<false>(y, lds, tw, twp, q);
#pragma unroll
        for (int r = 0; r < 16; ++r) x[r] = add_q(pmul<false>(x[r], av[r], avp[r], s), y[r], q);
    }
#pragma unroll
    for (int r = 0; r < 16; ++r) mpk[(size_t)z * IPFE_N + posA(tid, r)] = x[r];
}

// ---------------------------------------------------------------------------
// Encryption tail (src/rlwe_sife.c:150-176) for one polynomial of one message:
//   c_i = INTT(pk_i (.) r_hat) + f_i, then + floor(Q/p) x_i mod q (i < l),
//   c_L = INTT(a_hat (.) r_hat) + f_0.
// The noise f was scattered into c beforehand by the sampler.  Work items are
// (message, poly) pairs; poly == len means c_L.
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(NTT_THREADS) void k_enc_poly(const EncItem *__restrict__ items, const EncMsg *__restrict__ msgs,
                                                           const uint32_t *__restrict__ mpk, int mpk_l,
                                                           const uint32_t *__restrict__ tw_all, NinvConsts nc,
                                                           ScaleConsts sc)
{
    __shared__ uint32_t lds[NTT_LDS_WORDS];
    const EncItem it = items[blockIdx.x / 3];
    const int s = blockIdx.x % 3, tid = threadIdx.x;
    const uint32_t q = qmod(s);
    const EncMsg m = msgs[it.msg];
    const int i = it.poly;
    const bool is_cl = (i == m.len);
    const uint32_t *pk = mpk + ((size_t)(is_cl ? mpk_l : i) * 3 + s) * IPFE_N;
    const uint32_t *rh = m.rhat + (size_t)s * IPFE_N;
    const uint32_t *rhp = m.rhatp + (size_t)s * IPFE_N;
    uint32_t *c = is_cl ? m.cl + (size_t)s * IPFE_N : m.c + ((size_t)i * 3 + s) * IPFE_N;
    uint32_t x[16], rv[16], rvp[16];
#pragma unroll
    for (int r = 0; r < 16; ++r) {
        x[r] = pk[posA(tid, r)];
        rv[r] = rh[posA(tid, r)];
        rvp[r] = rhp[posA(tid, r)];
    }
    uint32_t madd = 0;
    if (!is_cl) {
        uint32_t mj = mod_prime_exact(m.x[i], s);
        madd = mod_prime_exact((uint64_t)mj * sc.scale[s], s);
    }
    const bool exact = __syncthreads_or(any_eq_q(x, q) || any_eq_q(rv, q));
    if (exact) {
#pragma unroll
        for (int r = 0; r < 16; ++r) x[r] = pmul<true>(x[r], rv[r], rvp[r], s);
        ntt_inv_block<true>(x, lds, tw_inv(tw_all, s), twp_inv(tw_all, s), q);
#pragma unroll
        for (int r = 0; r < 16; ++r) x[r] = scale_ninv<true>(x[r], nc.ninv[s], nc.ninvp[s], nc.flag[s], q);
    } else {
#pragma unroll
        for (int r = 0; r < 16; ++r) x[r] = pmul<false>(x[r], rv[r], rvp[r], s);
        ntt_inv_block<false>(x, lds, tw_inv(tw_all, s), twp_inv(tw_all, s), q);
#pragma unroll
        for (int r = 0; r < 16; ++r) x[r] = scale_ninv<false>(x[r], nc.ninv[s], nc.ninvp[s], nc.flag[s], q);
    }
#pragma unroll
    for (int r = 0; r < 16; ++r) {
        uint32_t f = c[posA(tid, r)];
        uint32_t t = add_q(x[r], f, q);
        if (!is_cl) {
            t = t + madd;  // < 2q, so mod_prime is one conditional subtract
            t = t >= q ? t - q : t;
        }
        c[posA(tid, r)] = t;
    }
}

// ---------------------------------------------------------------------------
// Pointwise product of two NTT-domain limbs, both read at [z][N] (z = poly*3+s):
// out = mul_mod(a, b), exact for operands in [0, q].
// ---------------------------------------------------------------------------
__global__ void k_point_mul(const uint32_t *__restrict__ a, const uint32_t *__restrict__ b, uint32_t *__restrict__ out,
                            size_t n_total, BarrettConsts bc)
{
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n_total) return;
    const int s = (int)((i / IPFE_N) % 3);
    out[i] = mul_exact(a[i], b[i], s, bc.mu_hi[s], bc.mu_lo[s]);
}

// ---------------------------------------------------------------------------
// Launchers
// ---------------------------------------------------------------------------
int launch_ntt_fwd(const uint32_t *in, uint32_t *out, int npoly_limbs, const uint32_t *tw_all, hipStream_t st)
{
    if (npoly_limbs <= 0) return 0;
    hipLaunchKernelGGL(k_ntt_fwd, dim3(npoly_limbs), dim3(NTT_THREADS), 0, st, in, out, tw_all);
    HIP_OK(hipGetLastError());
    return 0;
}

int launch_ntt_inv(const uint32_t *in, uint32_t *out, int npoly_limbs, const uint32_t *tw_all, const NinvConsts &nc,
                   hipStream_t st)
{
    if (npoly_limbs <= 0) return 0;
    hipLaunchKernelGGL(k_ntt_inv, dim3(npoly_limbs), dim3(NTT_THREADS), 0, st, in, out, tw_all, nc);
    HIP_OK(hipGetLastError());
    return 0;
}

int launch_shoup_prep(const uint32_t *v, uint32_t *vp, size_t n_total, hipStream_t st)
{
    if (n_total == 0) return 0;
    const int bs = 256;
    hipLaunchKernelGGL(k_shoup_prep, dim3((unsigned)((n_total + bs - 1) / bs)), dim3(bs), 0, st, v, vp, n_total);
    HIP_OK(hipGetLastError());
    return 0;
}

int launch_setup_pk(const uint32_t *msk, const uint32_t *e, const uint32_t *ahat, const uint32_t *ahatp, uint32_t *mpk,
                    int l, const uint32_t *tw_all, hipStream_t st)
{
    if (l <= 0) return 0;
    hipLaunchKernelGGL(k_setup_pk, dim3(l * 3), dim3(NTT_THREADS), 0, st, msk, e, ahat, ahatp, mpk, tw_all);
    HIP_OK(hipGetLastError());
    return 0;
}

int launch_enc_poly(const EncItem *items, int nitems, const EncMsg *msgs, const uint32_t *mpk, int mpk_l,
                    const uint32_t *tw_all, const NinvConsts &nc, const ScaleConsts &sc, hipStream_t st)
{
    if (nitems <= 0) return 0;
    hipLaunchKernelGGL(k_enc_poly, dim3(nitems * 3), dim3(NTT_THREADS), 0, st, items, msgs, mpk, mpk_l, tw_all, nc, sc);
    HIP_OK(hipGetLastError());
    return 0;
}

int launch_point_mul(const uint32_t *a, const uint32_t *b, uint32_t *out, size_t n_total, const BarrettConsts &bc,
                     hipStream_t st)
{
    if (n_total == 0) return 0;
    const int bs = 256;
    hipLaunchKernelGGL(k_point_mul, dim3((unsigned)((n_total + bs - 1) / bs)), dim3(bs), 0, st, a, b, out, n_total, bc);
    HIP_OK(hipGetLastError());
    return 0;
}

}  // namespace ipfe
