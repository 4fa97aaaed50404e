// k_pset.hip -- the scheme's kernels for any parameter set (N = 2048, 4096,
// 8192; 2 to 4 limbs): setup's public-key kernel, the encryption tail, the
// keygen / decryption folds, pointwise products and CRT + rounding.  The
// reference set's batched paths use the specialised kernels of k_ntt.hip,
// k_scheme.hip and k_predict.hip; these serve the single-operation API of
// every set.
#include <stdio.h>

#include "kernels_pset.h"
#include "ntt_gen.h"

namespace ipfe {

namespace {

// Shape of the transforms inside the fused kernels, per log2 N.
template <int LOGN>
struct FusedShape;
template <>
struct FusedShape<11> {
    static constexpr int T = 128;
    using RAD = Rad<4, 4, 3>;
};
template <>
struct FusedShape<12> {
    static constexpr int T = 256;
    using RAD = Rad<4, 4, 4>;
};
template <>
struct FusedShape<13> {
    static constexpr int T = 256;
    using RAD = Rad<5, 4, 4>;
};

// mul_mod(a, b) for b with Shoup companion bp; exact for a, b in [0, q].
template <bool EXACT>
__device__ __forceinline__ uint32_t ps_pmul(const PsetDev &ps, uint32_t a, uint32_t b, uint32_t bp, int s)
{
    const uint32_t q = ps.q[s];
    if (!EXACT) return mul_shoup(a, b, bp, q);
    if (a == q || b == q) return (a == 0 || b == 0) ? 0u : ps_mod_prime(ps, (uint64_t)a * b, s);
    return mul_shoup(a, b, bp, q);
}

// rlwe_sife_setup's loop body (src/rlwe_sife.c:61-77) for slot i, limb s:
// pk_i = a_hat (.) NTT(s_i) + NTT(e_i).  Block z = i * nmod + s; msk, e and
// mpk are [l][nmod][N]; ahat / ahatp [nmod][N] (NTT domain, Shoup companions).
template <int LOGN>
__global__ __launch_bounds__(FusedShape<LOGN>::T) void k_ps_setup_pk(const uint32_t *__restrict__ msk,
                                                                   const uint32_t *__restrict__ err,
                                                                   const uint32_t *__restrict__ ahat,
                                                                   const uint32_t *__restrict__ ahatp,
                                                                   uint32_t *__restrict__ mpk,
                                                                   const uint32_t *__restrict__ tw_all, PsetDev ps)
{
    using G = NttGen<LOGN, FusedShape<LOGN>::T, typename FusedShape<LOGN>::RAD>;
    constexpr int N = G::N, E = G::E;
    __shared__ uint32_t lds[G::LDS_WORDS];
    const int z = blockIdx.x, s = z % ps.nmod, t = threadIdx.x;
    const uint32_t q = ps.q[s];
    const uint32_t *tw = tw_all + (size_t)s * N, *twp = tw_all + (size_t)(ps.nmod + s) * N;
    uint32_t x[E], y[E], av[E], avp[E];
    bool hq = false;
#pragma unroll
    for (int e = 0; e < E; ++e) {
        x[e] = msk[(size_t)z * N + G::first_idx(t, e)];
        y[e] = err[(size_t)z * N + G::first_idx(t, e)];
        av[e] = ahat[(size_t)s * N + G::last_idx(t, e)];
        avp[e] = ahatp[(size_t)s * N + G::last_idx(t, e)];
        hq |= x[e] == q || y[e] == q || av[e] == q;
    }
    uint32_t *dst = mpk + (size_t)z * N;
    if (__syncthreads_or(hq)) {
        G::template fwd<true>(x, lds, tw, twp, q, t);
        G::template fwd<true>(y, lds, tw, twp, q, t);
#pragma unroll
        for (int e = 0; e < E; ++e) dst[G::last_idx(t, e)] = add_q(ps_pmul<true>(ps, x[e], av[e], avp[e], s), y[e], q);
    } else {
        G::template fwd<false>(x, lds, tw, twp, q, t);
        G::template fwd<false>(y, lds, tw, twp, q, t);
#pragma unroll
        for (int e = 0; e < E; ++e)
            dst[G::last_idx(t, e)] = add_q(ps_pmul<false>(ps, x[e], av[e], avp[e], s), y[e], q);
    }
}

// rlwe_sife_encrypt's tail (src/rlwe_sife.c:150-176) for message b, poly i,
// limb s (block z = (b (l+1) + i) nmod + s):
//   c_i = INTT(pk_i (.) r_hat_b) + f_i, then + floor(Q/p) x_b[i] mod q (i < l),
//   c_L = INTT(a_hat (.) r_hat_b) + f_0 (i = l, pk = mpk[l] = a_hat).
// c [repeat][l+1][nmod][N] holds the noise on entry; rhat / rhatp
// [repeat][nmod][N]; x [repeat][l].
template <int LOGN>
__global__ __launch_bounds__(FusedShape<LOGN>::T) void k_ps_enc_poly(const uint32_t *__restrict__ mpk,
                                                                   const uint32_t *__restrict__ rhat,
                                                                   const uint32_t *__restrict__ rhatp,
                                                                   const uint32_t *__restrict__ x_msg,
                                                                   uint32_t *__restrict__ c, int l,
                                                                   const uint32_t *__restrict__ tw_all, PsetDev ps)
{
    using G = NttGen<LOGN, FusedShape<LOGN>::T, typename FusedShape<LOGN>::RAD>;
    constexpr int N = G::N, E = G::E;
    __shared__ uint32_t lds[G::LDS_WORDS];
    const int z = blockIdx.x, s = z % ps.nmod, t = threadIdx.x;
    const int poly = z / ps.nmod, b = poly / (l + 1), i = poly % (l + 1);
    const uint32_t q = ps.q[s];
    const uint32_t *tw = tw_all + (size_t)(2 * ps.nmod + s) * N, *twp = tw_all + (size_t)(3 * ps.nmod + s) * N;
    const uint32_t *pk = mpk + ((size_t)i * ps.nmod + s) * N;
    const uint32_t *rh = rhat + ((size_t)b * ps.nmod + s) * N;
    const uint32_t *rhp = rhatp + ((size_t)b * ps.nmod + s) * N;
    uint32_t v[E], rv[E], rvp[E];
    bool hq = false;
#pragma unroll
    for (int e = 0; e < E; ++e) {
        v[e] = pk[G::last_idx(t, e)];
        rv[e] = rh[G::last_idx(t, e)];
        rvp[e] = rhp[G::last_idx(t, e)];
        hq |= v[e] == q || rv[e] == q;
    }
    uint32_t madd = 0;
    if (i < l) {
        const uint32_t mj = ps_mod_prime(ps, x_msg[(size_t)b * l + i], s);
        madd = ps_mod_prime(ps, (uint64_t)mj * ps.scale[s], s);
    }
    uint32_t *cz = c + (size_t)z * N;
    const bool exact = __syncthreads_or(hq);
    if (exact) {
#pragma unroll
        for (int e = 0; e < E; ++e) v[e] = ps_pmul<true>(ps, v[e], rv[e], rvp[e], s);
        G::template inv<true>(v, lds, tw, twp, q, t);
#pragma unroll
        for (int e = 0; e < E; ++e) v[e] = scale_ninv<true>(v[e], ps.ninv[s], ps.ninvp[s], ps.ninvf[s], q);
    } else {
#pragma unroll
        for (int e = 0; e < E; ++e) v[e] = ps_pmul<false>(ps, v[e], rv[e], rvp[e], s);
        G::template inv<false>(v, lds, tw, twp, q, t);
#pragma unroll
        for (int e = 0; e < E; ++e) v[e] = scale_ninv<false>(v[e], ps.ninv[s], ps.ninvp[s], ps.ninvf[s], q);
    }
#pragma unroll
    for (int e = 0; e < E; ++e) {
        const int idx = G::first_idx(t, e);
        uint32_t w = add_q(v[e], cz[idx], q);
        if (i < l) w = ps_mod_prime(ps, (uint64_t)(uint32_t)(w + madd), s);  // mod_prime(c + m), src/rlwe_sife.c:173
        cz[idx] = w;
    }
}

// out[v][s][k] = fold over i < l of mod_prime(y_v[i] P[i][s][k] + acc) -- the
// dense loops of rlwe_sife_keygen (src/rlwe_sife.c:232-240, P = msk) and of
// rlwe_sife_decrypt_gmp (:265-273, P = c), literally, per limb.
__global__ void k_ps_fold(const uint32_t *__restrict__ y, int nvec, int l, const uint32_t *__restrict__ P,
                          uint32_t *__restrict__ out, PsetDev ps)
{
    const size_t per_vec = (size_t)ps.nmod * ps.N;
    const size_t total = (size_t)nvec * per_vec;
    for (size_t w = (size_t)blockIdx.x * blockDim.x + threadIdx.x; w < total; w += (size_t)gridDim.x * blockDim.x) {
        const int v = (int)(w / per_vec);
        const size_t sk = w % per_vec;
        const int s = (int)(sk / ps.N);
        const uint32_t *yv = y + (size_t)v * l;
        uint32_t acc = 0;
        for (int i = 0; i < l; ++i) {
            const uint64_t yi = ps_mod_prime(ps, yv[i], s);
            acc = ps_mod_prime(ps, yi * P[(size_t)i * per_vec + sk] + acc, s);
        }
        out[w] = acc;
    }
}

// Shoup companions floor(v 2^32 / q) of [n][nmod][N] residues (q stands for 0).
__global__ void k_ps_shoup(const uint32_t *__restrict__ v, uint32_t *__restrict__ vp, size_t total, PsetDev ps)
{
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < total; i += (size_t)gridDim.x * blockDim.x) {
        const int s = (int)((i / ps.N) % ps.nmod);
        const uint64_t q = ps.q[s];
        uint64_t w = v[i];
        if (w >= q) w -= q;
        vp[i] = (uint32_t)((w << 32) / q);
    }
}

// out = mul_mod(a, b) (point_mul, src/ntt.c:28-36), exact for operands in [0, q].
__global__ void k_ps_point_mul(const uint32_t *__restrict__ a, const uint32_t *__restrict__ b,
                               uint32_t *__restrict__ out, size_t total, PsetDev ps)
{
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < total; i += (size_t)gridDim.x * blockDim.x) {
        const int s = (int)((i / ps.N) % ps.nmod);
        out[i] = ps_mul_exact(ps, a[i], b[i], s);
    }
}

// d = sub_mod(d, prod) (poly_sub_mod, src/arith_rns.c:95-103).
__global__ void k_ps_sub(uint32_t *__restrict__ d, const uint32_t *__restrict__ prod, size_t total, PsetDev ps)
{
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < total; i += (size_t)gridDim.x * blockDim.x) {
        const int s = (int)((i / ps.N) % ps.nmod);
        d[i] = sub_q(d[i], prod[i], ps.q[s]);
    }
}

// CRT + rounding of coefficient `coef` (all coefficients when coef < 0) of
// d [nvec][nmod][N]: r = round_extract(crt_reverse(d)) -> r [nvec] or [nvec][N].
__global__ void k_ps_crt_round(const uint32_t *__restrict__ d, int nvec, int coef, uint32_t *__restrict__ r,
                               PsetDev ps)
{
    const size_t per = coef < 0 ? (size_t)ps.N : 1;
    const size_t total = (size_t)nvec * per;
    for (size_t w = (size_t)blockIdx.x * blockDim.x + threadIdx.x; w < total; w += (size_t)gridDim.x * blockDim.x) {
        const int v = (int)(w / per);
        const int k = coef < 0 ? (int)(w % per) : coef;
        uint32_t res[PSET_MAX_LIMBS];
        for (int s = 0; s < ps.nmod; ++s) res[s] = d[((size_t)v * ps.nmod + s) * ps.N + k];
        r[w] = ps_crt_round_value(ps, res);
    }
}

template <int LOGN>
int launch_setup_pk_t(const PsetState &st, const uint32_t *msk, const uint32_t *e, const uint32_t *ahat,
                      const uint32_t *ahatp, uint32_t *mpk, int l, hipStream_t s)
{
    hipLaunchKernelGGL(k_ps_setup_pk<LOGN>, dim3((unsigned)(l * st.dev.nmod)), dim3(FusedShape<LOGN>::T), 0, s, msk, e,
                       ahat, ahatp, mpk, st.tw, st.dev);
    HIP_OK(hipGetLastError());
    return 0;
}

template <int LOGN>
int launch_enc_poly_t(const PsetState &st, const uint32_t *mpk, const uint32_t *rhat, const uint32_t *rhatp,
                      const uint32_t *x, uint32_t *c, int repeat, int l, hipStream_t s)
{
    hipLaunchKernelGGL(k_ps_enc_poly<LOGN>, dim3((unsigned)((size_t)repeat * (l + 1) * st.dev.nmod)),
                       dim3(FusedShape<LOGN>::T), 0, s, mpk, rhat, rhatp, x, c, l, st.tw, st.dev);
    HIP_OK(hipGetLastError());
    return 0;
}

unsigned grid_for(size_t n)
{
    const size_t g = (n + 255) / 256;
    return (unsigned)(g < 65536 ? (g ? g : 1) : 65536);
}

}  // namespace

int ps_launch_setup_pk(const PsetState &st, const uint32_t *msk, const uint32_t *e, const uint32_t *ahat,
                       const uint32_t *ahatp, uint32_t *mpk, int l, hipStream_t s)
{
    if (l <= 0) return 0;
    switch (st.dev.logN) {
    case 11: return launch_setup_pk_t<11>(st, msk, e, ahat, ahatp, mpk, l, s);
    case 12: return launch_setup_pk_t<12>(st, msk, e, ahat, ahatp, mpk, l, s);
    case 13: return launch_setup_pk_t<13>(st, msk, e, ahat, ahatp, mpk, l, s);
    }
    ipfe_set_error("setup: unsupported N");
    return -1;
}

int ps_launch_enc_poly(const PsetState &st, const uint32_t *mpk, const uint32_t *rhat, const uint32_t *rhatp,
                       const uint32_t *x, uint32_t *c, int repeat, int l, hipStream_t s)
{
    if (repeat <= 0) return 0;
    switch (st.dev.logN) {
    case 11: return launch_enc_poly_t<11>(st, mpk, rhat, rhatp, x, c, repeat, l, s);
    case 12: return launch_enc_poly_t<12>(st, mpk, rhat, rhatp, x, c, repeat, l, s);
    case 13: return launch_enc_poly_t<13>(st, mpk, rhat, rhatp, x, c, repeat, l, s);
    }
    ipfe_set_error("encrypt: unsupported N");
    return -1;
}

int ps_launch_fold(const PsetState &st, const uint32_t *y, int nvec, int l, const uint32_t *P, uint32_t *out,
                   hipStream_t s)
{
    const size_t total = (size_t)nvec * st.dev.nmod * st.dev.N;
    if (total == 0) return 0;
    hipLaunchKernelGGL(k_ps_fold, dim3(grid_for(total)), dim3(256), 0, s, y, nvec, l, P, out, st.dev);
    HIP_OK(hipGetLastError());
    return 0;
}

int ps_launch_shoup(const PsetState &st, const uint32_t *v, uint32_t *vp, size_t total, hipStream_t s)
{
    if (total == 0) return 0;
    hipLaunchKernelGGL(k_ps_shoup, dim3(grid_for(total)), dim3(256), 0, s, v, vp, total, st.dev);
    HIP_OK(hipGetLastError());
    return 0;
}

int ps_launch_point_mul(const PsetState &st, const uint32_t *a, const uint32_t *b, uint32_t *out, size_t total,
                        hipStream_t s)
{
    if (total == 0) return 0;
    hipLaunchKernelGGL(k_ps_point_mul, dim3(grid_for(total)), dim3(256), 0, s, a, b, out, total, st.dev);
    HIP_OK(hipGetLastError());
    return 0;
}

int ps_launch_sub(const PsetState &st, uint32_t *d, const uint32_t *prod, size_t total, hipStream_t s)
{
    if (total == 0) return 0;
    hipLaunchKernelGGL(k_ps_sub, dim3(grid_for(total)), dim3(256), 0, s, d, prod, total, st.dev);
    HIP_OK(hipGetLastError());
    return 0;
}

int ps_launch_crt_round(const PsetState &st, const uint32_t *d, int nvec, int coef, uint32_t *r, hipStream_t s)
{
    const size_t total = (size_t)nvec * (coef < 0 ? st.dev.N : 1);
    if (total == 0) return 0;
    hipLaunchKernelGGL(k_ps_crt_round, dim3(grid_for(total)), dim3(256), 0, s, d, nvec, coef, r, st.dev);
    HIP_OK(hipGetLastError());
    return 0;
}

}  // namespace ipfe
