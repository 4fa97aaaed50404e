// ntt_core.h -- negacyclic NTT / INTT of one 4096-point limb by one 256-thread
// workgroup (4 waves of 64), bit-exact with CT_forward / GS_reverse
// (src/ntt.c:38-67, :70-114).
//
// Data flow.  Each thread holds 16 residues in registers.  The 12 radix-2
// stages run as three radix-16 register phases; between phases the workgroup
// exchanges data through LDS (4096 words plus one pad word per 16, so that the
// stride-16 and stride-17 access patterns hit distinct banks):
//   phase A  x[r] = a[tid + 256 r]               stages t = 2048 .. 256
//   phase B  x[r] = a[256 b + (tid&15) + 16 r]   stages t = 128 .. 16   (b = tid>>4)
//   phase C  x[r] = a[16 tid + r]                stages t = 8 .. 1
// In phase P with position p (0, b, tid) the twiddle of stage k (k = 0..3) and
// group g is psi[m0 * 2^k + 2^k * p + g] with m0 = 1, 16, 256 -- the psi[m + i]
// of the reference loop.
//
// Twiddle tables (device, per limb, 4096 words each): w | flag << 31 and the
// Shoup companion floor(w 2^32 / q).  flag = 1 when mod_prime(q * w) returns q.
#pragma once
#include "ipfe_common.h"

namespace ipfe {

constexpr int NTT_THREADS = 256;
constexpr int NTT_LDS_WORDS = IPFE_N + IPFE_N / 16;

__device__ __forceinline__ uint32_t lds_idx(uint32_t j) { return j + (j >> 4); }

// One CT butterfly: V = mul_mod(v, w); u' = add_mod(u, V); v' = sub_mod(u, V).
template <bool EXACT>
__device__ __forceinline__ void ct_bfly(uint32_t &u, uint32_t &v, uint32_t w, uint32_t wp, uint32_t q)
{
    uint32_t t = mul_shoup(v, w & 0x7fffffffu, wp, q);
    if (EXACT && v == q && (w >> 31)) t = q;
    uint32_t U = u;
    u = add_q(U, t, q);
    v = sub_q(U, t, q);
}

// One GS butterfly: u' = add_mod(u, v); v' = mul_mod(sub_mod(u, v), w).
template <bool EXACT>
__device__ __forceinline__ void gs_bfly(uint32_t &u, uint32_t &v, uint32_t w, uint32_t wp, uint32_t q)
{
    uint32_t U = u, V = v;
    u = add_q(U, V, q);
    uint32_t d = sub_q(U, V, q);
    uint32_t t = mul_shoup(d, w & 0x7fffffffu, wp, q);
    if (EXACT && d == q && (w >> 31)) t = q;
    v = t;
}

// Radix-16 CT on registers: stage k pairs r and r + (8 >> k); group g uses
// twiddle index base_k + g with base_k = (m0 + p) << k.
template <bool EXACT>
__device__ __forceinline__ void ct16(uint32_t (&x)[16], const uint32_t *__restrict__ tw,
                                     const uint32_t *__restrict__ twp, uint32_t m0, uint32_t p, uint32_t q)
{
    const uint32_t base = m0 + p;
    {
        uint32_t w = tw[base], wp = twp[base];
#pragma unroll
        for (int r = 0; r < 8; ++r) ct_bfly<EXACT>(x[r], x[r + 8], w, wp, q);
    }
#pragma unroll
    for (int g = 0; g < 2; ++g) {
        uint32_t i = 2 * base + g;
        uint32_t w = tw[i], wp = twp[i];
#pragma unroll
        for (int r = 0; r < 4; ++r) ct_bfly<EXACT>(x[8 * g + r], x[8 * g + r + 4], w, wp, q);
    }
#pragma unroll
    for (int g = 0; g < 4; ++g) {
        uint32_t i = 4 * base + g;
        uint32_t w = tw[i], wp = twp[i];
#pragma unroll
        for (int r = 0; r < 2; ++r) ct_bfly<EXACT>(x[4 * g + r], x[4 * g + r + 2], w, wp, q);
    }
#pragma unroll
    for (int g = 0; g < 8; ++g) {
        uint32_t i = 8 * base + g;
        ct_bfly<EXACT>(x[2 * g], x[2 * g + 1], tw[i], twp[i], q);
    }
}

// Radix-16 GS on registers, the mirror image of ct16.
template <bool EXACT>
__device__ __forceinline__ void gs16(uint32_t (&x)[16], const uint32_t *__restrict__ tw,
                                    const uint32_t *__restrict__ twp, uint32_t m0, uint32_t p, uint32_t q)
{
    const uint32_t base = m0 + p;
#pragma unroll
    for (int g = 0; g < 8; ++g) {
        uint32_t i = 8 * base + g;
        gs_bfly<EXACT>(x[2 * g], x[2 * g + 1], tw[i], twp[i], q);
    }
#pragma unroll
    for (int g = 0; g < 4; ++g) {
        uint32_t i = 4 * base + g;
        uint32_t w = tw[i], wp = twp[i];
#pragma unroll
        for (int r = 0; r < 2; ++r) gs_bfly<EXACT>(x[4 * g + r], x[4 * g + r + 2], w, wp, q);
    }
#pragma unroll
    for (int g = 0; g < 2; ++g) {
        uint32_t i = 2 * base + g;
        uint32_t w = tw[i], wp = twp[i];
#pragma unroll
        for (int r = 0; r < 4; ++r) gs_bfly<EXACT>(x[8 * g + r], x[8 * g + r + 4], w, wp, q);
    }
    {
        uint32_t w = tw[base], wp = twp[base];
#pragma unroll
        for (int r = 0; r < 8; ++r) gs_bfly<EXACT>(x[r], x[r + 8], w, wp, q);
    }
}

__device__ __forceinline__ uint32_t posA(int tid, int r) { return (uint32_t)(tid + 256 * r); }
__device__ __forceinline__ uint32_t posB(int tid, int r) { return (uint32_t)(256 * (tid >> 4) + (tid & 15) + 16 * r); }
__device__ __forceinline__ uint32_t posC(int tid, int r) { return (uint32_t)(16 * tid + r); }

// Forward transform.  In: x in phase-A layout.  Out: x in phase-A layout (the
// bit-reversed-order result, element j at thread j & 255, register j >> 8).
template <bool EXACT>
__device__ void ntt_fwd_block(uint32_t (&x)[16], uint32_t *lds, const uint32_t *tw, const uint32_t *twp, uint32_t q)
{
    const int tid = threadIdx.x;
    ct16<EXACT>(x, tw, twp, 1u, 0u, q);
#pragma unroll
    for (int r = 0; r < 16; ++r) lds[lds_idx(posA(tid, r))] = x[r];
    __syncthreads();
#pragma unroll
    for (int r = 0; r < 16; ++r) x[r] = lds[lds_idx(posB(tid, r))];
    ct16<EXACT>(x, tw, twp, 16u, (uint32_t)(tid >> 4), q);
#pragma unroll
    for (int r = 0; r < 16; ++r) lds[lds_idx(posB(tid, r))] = x[r];
    __syncthreads();
#pragma unroll
    for (int r = 0; r < 16; ++r) x[r] = lds[lds_idx(posC(tid, r))];
    ct16<EXACT>(x, tw, twp, 256u, (uint32_t)tid, q);
#pragma unroll
    for (int r = 0; r < 16; ++r) lds[lds_idx(posC(tid, r))] = x[r];
    __syncthreads();
#pragma unroll
    for (int r = 0; r < 16; ++r) x[r] = lds[lds_idx(posA(tid, r))];
    __syncthreads();  // lds may be reused by the caller
}

// Inverse transform without the final N^-1 scaling.  In and out: phase-A layout.
template <bool EXACT>
__device__ void ntt_inv_block(uint32_t (&x)[16], uint32_t *lds, const uint32_t *tw, const uint32_t *twp, uint32_t q)
{
    const int tid = threadIdx.x;
#pragma unroll
    for (int r = 0; r < 16; ++r) lds[lds_idx(posA(tid, r))] = x[r];
    __syncthreads();
#pragma unroll
    for (int r = 0; r < 16; ++r) x[r] = lds[lds_idx(posC(tid, r))];
    gs16<EXACT>(x, tw, twp, 256u, (uint32_t)tid, q);
#pragma unroll
    for (int r = 0; r < 16; ++r) lds[lds_idx(posC(tid, r))] = x[r];
    __syncthreads();
#pragma unroll
    for (int r = 0; r < 16; ++r) x[r] = lds[lds_idx(posB(tid, r))];
    gs16<EXACT>(x, tw, twp, 16u, (uint32_t)(tid >> 4), q);
#pragma unroll
    for (int r = 0; r < 16; ++r) lds[lds_idx(posB(tid, r))] = x[r];
    __syncthreads();
#pragma unroll
    for (int r = 0; r < 16; ++r) x[r] = lds[lds_idx(posA(tid, r))];
    gs16<EXACT>(x, tw, twp, 1u, 0u, q);
    __syncthreads();
}

// Final x N^-1 of GS_reverse (src/ntt.c:110-113): mul_mod with the q case.
template <bool EXACT>
__device__ __forceinline__ uint32_t scale_ninv(uint32_t v, uint32_t ninv, uint32_t ninvp, uint32_t ninv_flag, uint32_t q)
{
    uint32_t t = mul_shoup(v, ninv, ninvp, q);
    if (EXACT && v == q && ninv_flag) t = q;
    return t;
}

}  // namespace ipfe
