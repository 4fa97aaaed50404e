// ipfe_common.h -- parameters and exact RNS arithmetic shared by the gfx950 kernels.
//
// The one parameter set of the reference (src/params.h:14-28): N = 4096, three
// NTT-friendly primes q_j = 2^a - 2^b + 1, L = 785, p = 50241.
//
// Representation rule.  The reference's mod_prime (src/arith_rns.c:61-83) maps
// m to its canonical residue, except that it returns q (not 0) when its fold
// stops exactly at 2q; that can only happen when m is a non-zero multiple of q.
// Every kernel here therefore computes canonical residues with fast reductions
// and reproduces the q representative exactly where it can arise:
//   * a product of two values in [0, q) is 0 mod q only if one factor is 0, and
//     then mod_prime(0) = 0, so q never appears from canonical operands;
//   * an operand equal to q is detected (a workgroup-wide vote in the NTT and
//     pointwise kernels) and that workgroup takes an exact path that runs the
//     fold literally.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#define IPFE_N 4096
#define IPFE_LOGN 12
#define IPFE_NMOD 3
#define IPFE_L 785
#define IPFE_P 50241u

namespace ipfe {

// src/params.h:19
__host__ __device__ constexpr uint32_t qmod(int s)
{
    return s == 0 ? 16760833u : (s == 1 ? 2147352577u : 2130706433u);
}
// q_j = 2^a - 2^b + 1 (src/arith_rns.c:7-16)
__host__ __device__ constexpr uint32_t fold_a(int s) { return s == 0 ? 24u : 31u; }
__host__ __device__ constexpr uint32_t fold_b(int s) { return s == 0 ? 14u : (s == 1 ? 17u : 24u); }

// Literal restatement of mod_prime: fold while m > 2q, then one conditional
// subtract.  Used wherever the q representative can arise.
__host__ __device__ inline uint32_t mod_prime_exact(uint64_t m, int s)
{
    const uint64_t q = qmod(s);
    const uint32_t a = fold_a(s), b = fold_b(s);
    const uint64_t lo = ((uint64_t)1 << a) - 1;
    while (m > 2 * q) {
        uint64_t hi = m >> a;
        m = (m & lo) + ((hi << b) - hi);
    }
    if (m >= q) m -= q;
    return (uint32_t)m;
}

// mod_prime for any prime q = 2^a - 2^b + 1 (the other parameter sets).
__host__ __device__ inline uint32_t mod_prime_gen(uint64_t m, uint32_t q32, uint32_t a, uint32_t b)
{
    const uint64_t q = q32;
    const uint64_t lo = ((uint64_t)1 << a) - 1;
    while (m > 2 * q) {
        uint64_t hi = m >> a;
        m = (m & lo) + ((hi << b) - hi);
    }
    if (m >= q) m -= q;
    return (uint32_t)m;
}

// add_mod_ntt / sub_mod_ntt (src/arith_rns.c:22-48).  Operands lie in [0, q]
// with q < 2^31, so the 32-bit sums cannot overflow.
__host__ __device__ inline uint32_t add_q(uint32_t a, uint32_t b, uint32_t q)
{
    uint32_t c = a + b;
    return c >= q ? c - q : c;
}
__host__ __device__ inline uint32_t sub_q(uint32_t a, uint32_t b, uint32_t q)
{
    uint32_t c = a + q - b;
    return c >= q ? c - q : c;
}

// Shoup multiplication by a constant w with w' = floor(w * 2^32 / q): returns
// a*w mod q in [0, q) for any a < 2^32 (needs q < 2^31).
__device__ inline uint32_t mul_shoup(uint32_t a, uint32_t w, uint32_t wp, uint32_t q)
{
    uint32_t quo = __umulhi(a, wp);
    uint32_t r = a * w - quo * q;
    return r >= q ? r - q : r;
}

// Canonical a*b mod q for variable operands (a, b < 2^32) by Barrett on the
// 64-bit product with mu = floor(2^64 / q) split in two 32-bit words.
struct Barrett {
    uint32_t q, mu_hi, mu_lo;
};
__host__ inline Barrett make_barrett(uint32_t q)
{
    unsigned __int128 mu = (((unsigned __int128)1) << 64) / q;
    Barrett b;
    b.q = q;
    b.mu_hi = (uint32_t)(mu >> 32);
    b.mu_lo = (uint32_t)mu;
    return b;
}
__device__ inline uint32_t reduce64(uint64_t x, uint32_t q, uint32_t mu_hi, uint32_t mu_lo)
{
    // quotient estimate floor(x * mu / 2^64), low error <= 2
    uint32_t xl = (uint32_t)x, xh = (uint32_t)(x >> 32);
    uint64_t t = (uint64_t)__umulhi(xl, mu_lo);
    t += (uint64_t)xl * mu_hi;
    uint64_t t2 = (uint64_t)xh * mu_lo;
    uint64_t mid = (t & 0xffffffffull) + (t2 & 0xffffffffull);
    uint64_t est = (uint64_t)xh * mu_hi + (t >> 32) + (t2 >> 32) + (mid >> 32);
    uint64_t r = x - est * q;
    while (r >= q) r -= q;
    return (uint32_t)r;
}

// mul_mod_ntt (src/arith_rns.c:50-59) for operands in [0, q]: canonical fast
// path; the product is 0 mod q only if an operand is 0 or q, and only then is
// the fold run literally.
__device__ inline uint32_t mul_exact(uint32_t a, uint32_t b, int s, uint32_t mu_hi, uint32_t mu_lo)
{
    const uint32_t q = qmod(s);
    uint64_t m = (uint64_t)a * b;
    uint32_t r = reduce64(m, q, mu_hi, mu_lo);
    if (r == 0 && m != 0) r = mod_prime_exact(m, s);
    return r;
}

}  // namespace ipfe

#define HIP_OK(expr)                                                              \
    do {                                                                          \
        hipError_t _e = (expr);                                                   \
        if (_e != hipSuccess) {                                                   \
            ipfe_set_error("%s failed: %s (%s:%d)", #expr, hipGetErrorString(_e), \
                           __FILE__, __LINE__);                                   \
            return -1;                                                            \
        }                                                                         \
    } while (0)

// Thread-local error string for the status-returning API (defined in ipfe_capi.cpp).
void ipfe_set_error(const char *fmt, ...);
