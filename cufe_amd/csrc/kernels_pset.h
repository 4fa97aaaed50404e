// kernels_pset.h -- launchers of the parameter-set kernels (k_pset.hip) and
// the CRT + rounding they share with the host.
#pragma once
#include <hip/hip_runtime.h>
#include <math.h>
#include <stddef.h>
#include <stdint.h>

#include "pset.h"

namespace ipfe {

// x mod q for a 128-bit x, by 32-bit words from the top.
__host__ __device__ inline uint32_t mod128(unsigned __int128 x, uint32_t q)
{
    uint64_t r = 0;
    for (int k = 3; k >= 0; --k) r = ((r << 32) | (uint32_t)(x >> (32 * k))) % q;
    return (uint32_t)r;
}

// crt_reverse_gmp (src/crt.c:46-82) then round_extract_gmp
// (src/rlwe_sife.c:288-332) for one coefficient of a set with Q p < 2^128:
// Garner from the limb-0 residue as given, u_i = ((v_i - x) C_i) mod q_i
// (non-negative, as mpz_mod_ui), x += u_i q_0 .. q_{i-1}; then
// r = floor(x p / Q) + [x p mod Q > floor(Q/2)], no reduction mod p.  The
// quotient is estimated in doubles and corrected with exact 128-bit products.
__host__ __device__ inline uint32_t ps_crt_round_value(const PsetDev &ps, const uint32_t *res)
{
    unsigned __int128 x = res[0];
    for (int i = 1; i < ps.nmod; ++i) {
        const uint32_t q = ps.q[i];
        const uint32_t xm = mod128(x, q);
        const uint64_t diff = ((uint64_t)(res[i] % q) + q - xm) % q;
        const uint64_t u = diff * ps.garner_c[i] % q;
        const unsigned __int128 M = ((unsigned __int128)ps.garner_m_hi[i] << 64) | ps.garner_m_lo[i];
        x += (unsigned __int128)u * M;
    }
    const unsigned __int128 Q = ((unsigned __int128)ps.Q_hi << 64) | ps.Q_lo;
    const unsigned __int128 half = ((unsigned __int128)ps.halfQ_hi << 64) | ps.halfQ_lo;
    const unsigned __int128 xp = x * ps.p;
    const double xd = (double)(uint64_t)(xp >> 64) * 18446744073709551616.0 + (double)(uint64_t)xp;
    const double Qd = (double)ps.Q_hi * 18446744073709551616.0 + (double)ps.Q_lo;
    uint64_t quo = (uint64_t)floor(xd / Qd);
    unsigned __int128 prod = (unsigned __int128)quo * Q;
    while (prod > xp) {
        --quo;
        prod -= Q;
    }
    while (xp - prod >= Q) {
        ++quo;
        prod += Q;
    }
    if (xp - prod > half) ++quo;
    return (uint32_t)quo;
}

int ps_launch_setup_pk(const PsetState &st, const uint32_t *msk, const uint32_t *e, const uint32_t *ahat,
                       const uint32_t *ahatp, uint32_t *mpk, int l, hipStream_t s);
int ps_launch_enc_poly(const PsetState &st, const uint32_t *mpk, const uint32_t *rhat, const uint32_t *rhatp,
                       const uint32_t *x, uint32_t *c, int repeat, int l, hipStream_t s);
int ps_launch_fold(const PsetState &st, const uint32_t *y, int nvec, int l, const uint32_t *P, uint32_t *out,
                   hipStream_t s);
int ps_launch_shoup(const PsetState &st, const uint32_t *v, uint32_t *vp, size_t total, hipStream_t s);
int ps_launch_point_mul(const PsetState &st, const uint32_t *a, const uint32_t *b, uint32_t *out, size_t total,
                        hipStream_t s);
int ps_launch_sub(const PsetState &st, uint32_t *d, const uint32_t *prod, size_t total, hipStream_t s);
// ntrans limb transforms ([npoly][nmod][N], transform z on limb z % nmod) of
// set st with its tuned shape (in == out allowed).
int ps_ntt(const PsetState &st, bool inverse, const uint32_t *in, uint32_t *out, int ntrans, hipStream_t s);
// coef < 0: every coefficient (r [nvec][N]); else coefficient coef (r [nvec]).
int ps_launch_crt_round(const PsetState &st, const uint32_t *d, int nvec, int coef, uint32_t *r, hipStream_t s);

}  // namespace ipfe
