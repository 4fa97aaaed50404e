// pset.h -- the library's parameter sets and their device tables.
//
//   set 0: N = 2048, 2 limbs (2^31-2^17+1, 2^31-2^24+1), sampler widths
//          k = 266, 262147, 524287;
//   set 1: the reference's set (src/params.h:14-28): N = 4096, 3 limbs,
//          k = 266, 304214978, 608429953 -- the set every specialised kernel
//          of the library serves;
//   set 2: N = 8192, 4 limbs (the reference's three primes and 2^26-2^16+1),
//          the reference's sampler widths.
// L = 785, B_x = 4, B_y = 16 and p = 50241 in every set.  The reference names
// other levels only in a comment (src/sample.c:16-19: N = 2048 with 64 bits of
// Q, N = 8192 with more); sets 0 and 2 are this library's choice.  Their
// primes are of the form 2^a - 2^b + 1 with q = 1 mod 2N, so that mod_prime's
// fold (src/arith_rns.c:61-83) applies; set 0's widths keep 15 standard
// deviations of the worst-case decryption noise (|y| = B_y sqrt(L)) below
// Q / (2p) and are not powers of two (the uniform bound k floor(2^32/k) must
// fit the 32-bit comparison); their roots are w = g^((q-1)/2N) with g the smallest primitive
// root mod q (set 1 keeps the reference's roots, src/consts.h:7-11); Q p <
// 2^128 in every set, so CRT and rounding fit 128-bit integers.
#pragma once
#include <stdint.h>

#include "ipfe_common.h"

namespace ipfe {

constexpr int PSET_COUNT = 3;
constexpr int PSET_MAX_LIMBS = 4;

struct PsetSpec {
    int id, N, logN, nmod, L;
    uint32_t q[PSET_MAX_LIMBS], fa[PSET_MAX_LIMBS], fb[PSET_MAX_LIMBS], root[PSET_MAX_LIMBS];
    uint64_t k[4];  // sampler widths k[1..3]
    int width[4];   // index of width k[w] in the sampler's device tables (1..5)
    uint32_t p;
};

// Passed to kernels by value.
struct PsetDev {
    int N, logN, nmod, id;
    uint32_t q[PSET_MAX_LIMBS], fa[PSET_MAX_LIMBS], fb[PSET_MAX_LIMBS];
    uint32_t ninv[PSET_MAX_LIMBS], ninvp[PSET_MAX_LIMBS], ninvf[PSET_MAX_LIMBS];
    uint32_t scale[PSET_MAX_LIMBS];           // floor(Q/p) mod q_j
    uint32_t mu_hi[PSET_MAX_LIMBS], mu_lo[PSET_MAX_LIMBS];  // Barrett floor(2^64/q)
    // Garner: C_i = (q_0 .. q_{i-1})^-1 mod q_i and M_i = q_0 .. q_{i-1} (128-bit)
    uint32_t garner_c[PSET_MAX_LIMBS];
    uint64_t garner_m_lo[PSET_MAX_LIMBS], garner_m_hi[PSET_MAX_LIMBS];
    uint64_t Q_lo, Q_hi, halfQ_lo, halfQ_hi;  // Q and floor(Q/2)
    uint32_t p, pad;
};

const PsetSpec *pset_spec(int id);  // nullptr if id is not a set
// Device state of a set (twiddles uploaded on first use); nullptr on error.
struct PsetState {
    PsetDev dev;
    uint32_t *tw = nullptr;  // device [4][nmod][N]: psi|flag, psi', psi_inv|flag, psi_inv'
};
const PsetState *pset_state(int id);

// mod_prime of limb j of a set, on host or device.
__host__ __device__ inline uint32_t ps_mod_prime(const PsetDev &ps, uint64_t m, int j)
{
    return mod_prime_gen(m, ps.q[j], ps.fa[j], ps.fb[j]);
}

// Canonical a*b mod q_j with the reference's representative: the product is 0
// mod q only if a factor is 0 or q, and only then is the fold run literally.
__device__ inline uint32_t ps_mul_exact(const PsetDev &ps, uint32_t a, uint32_t b, int j)
{
    const uint64_t m = (uint64_t)a * b;
    uint32_t r = reduce64(m, ps.q[j], ps.mu_hi[j], ps.mu_lo[j]);
    if (r == 0 && m != 0) r = ps_mod_prime(ps, m, j);
    return r;
}

}  // namespace ipfe
