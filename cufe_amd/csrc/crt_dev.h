// crt_dev.h -- exact CRT reconstruction and rounding, shared by the full
// decryption path and the batched prediction path.
//
// crt_reverse_gmp (src/crt.c:46-82) is Garner's method with GMP integers:
//   x = v0;  x += ((v1 - x) C1 mod q1) q0;  x += ((v2 - x) C2 mod q2) q0 q1
// with mpz_mod_ui's non-negative remainder.  v0 enters as given, so a limb-0
// residue equal to q0 yields x in [Q, Q + q0) rather than x mod Q.
// round_extract_gmp (src/rlwe_sife.c:288-332) then returns floor(p x / Q) plus one
// when the remainder exceeds floor(Q/2) -- never reduced mod p.
#pragma once
#include <math.h>

#include "ipfe_common.h"

namespace ipfe {

typedef unsigned __int128 u128;

// (q0 q1 q2), q0 q1, q0^-1 mod q1, (q0 q1)^-1 mod q2 (residues of
// SIFE_CRT_CONSTS, src/params.h:17).
constexpr uint64_t CRT_Q01 = 16760833ull * 2147352577ull;
constexpr uint64_t CRT_C1 = 206923011ull;
constexpr uint64_t CRT_C2 = 2935204199007202395ull % 2130706433ull;

__host__ __device__ inline u128 crt_Q() { return (u128)CRT_Q01 * 2130706433ull; }

// Shoup's constant for multiplying by w < q mod q: floor(w 2^32 / q).
__host__ __device__ constexpr uint32_t crt_shoup(uint64_t w, uint64_t q) { return (uint32_t)((w << 32) / q); }

// a w mod q for a < 2^32, w < q < 2^31 (Shoup: the quotient estimate is low by
// at most one).
__host__ __device__ inline uint32_t crt_mulmod(uint32_t a, uint32_t w, uint32_t ws, uint32_t q)
{
    const uint32_t t = (uint32_t)(((uint64_t)a * ws) >> 32);
    uint32_t r = a * w - t * q;
    return r >= q ? r - q : r;
}

// round(p x / Q) for Garner's x = x01 + u2 Q01 (x01 < Q01 + 2^32, u2 < q2,
// Q = Q01 q2).  With p u2 = A q2 + B (B < q2):
//     p x = A Q + N',   N' = B Q01 + p x01 < Q + p (Q01 + 2^32) < 2 Q,
// so floor(p x / Q) = A + [N' >= Q], and rounding adds one when the
// remainder exceeds floor(Q/2) (Q is odd: when twice the remainder exceeds Q):
//     round(p x / Q) = A + [2 N' >= Q] + [2 N' >= 3 Q].
// N' / Q = B / q2 + p x01 / Q is first evaluated in doubles (absolute error
// below 1e-15); only within 1e-9 of a threshold is it decided in 128 bits.
__host__ __device__ inline uint32_t crt_round_split(uint64_t x01, uint32_t u2)
{
    const uint64_t q2 = qmod(2);
    const uint64_t pu = (uint64_t)u2 * IPFE_P;  // < 2^47, exact in a double
    int64_t A = (int64_t)((double)pu * (1.0 / (double)q2));
    int64_t B = (int64_t)pu - A * (int64_t)q2;
    if (B < 0) { --A; B += (int64_t)q2; }
    if (B >= (int64_t)q2) { ++A; B -= (int64_t)q2; }
    const double Qd = 76687145727357674227351553.0;  // Q to 53 bits
    const double f2 = 2.0 * ((double)B * (1.0 / (double)q2) + (double)x01 * ((double)IPFE_P / Qd));
    if (fabs(f2 - 1.0) > 1e-9 && fabs(f2 - 3.0) > 1e-9)
        return (uint32_t)(A + (f2 > 1.0 ? 1 : 0) + (f2 > 3.0 ? 1 : 0));
    const u128 n2 = 2 * ((u128)(uint64_t)B * CRT_Q01 + (u128)x01 * IPFE_P);  // 2 N' < 2^88
    const u128 Q = crt_Q();
    return (uint32_t)(A + (n2 >= Q ? 1 : 0) + (n2 >= 3 * Q ? 1 : 0));
}

// round_extract(crt_reverse(v0, v1, v2)) for canonical residues v_s < q_s,
// in 32-bit Shoup multiplications: u1 = (v1 - v0) C1 mod q1, x01 = v0 + u1 q0,
// x01 mod q2 = (v0 + u1 q0) mod q2, u2 = (v2 - x01) C2 mod q2.
__host__ __device__ inline uint32_t crt_decode_canon(uint32_t v0, uint32_t v1, uint32_t v2)
{
    constexpr uint32_t q0 = qmod(0), q1 = qmod(1), q2 = qmod(2);
    constexpr uint32_t c1 = (uint32_t)CRT_C1, c2 = (uint32_t)CRT_C2;
    constexpr uint32_t c1s = crt_shoup(c1, q1), c2s = crt_shoup(c2, q2), q0s = crt_shoup(q0, q2);
    const uint32_t d1 = v1 >= v0 ? v1 - v0 : v1 + q1 - v0;  // v0 < q0 < q1
    const uint32_t u1 = crt_mulmod(d1, c1, c1s, q1);
    uint32_t x2 = v0 + crt_mulmod(u1, q0, q0s, q2);  // < 2 q2
    x2 = x2 >= q2 ? x2 - q2 : x2;
    const uint32_t d2 = v2 >= x2 ? v2 - x2 : v2 + q2 - x2;
    const uint32_t u2 = crt_mulmod(d2, c2, c2s, q2);
    return crt_round_split((uint64_t)v0 + (uint64_t)u1 * q0, u2);
}

// round_extract(crt_reverse(v0, v1, v2)) for any words (the reference's
// inputs may hold the representative q, and v0 enters Garner unreduced).
__host__ __device__ inline uint32_t crt_decode(uint32_t v0, uint32_t v1, uint32_t v2)
{
    const uint64_t q0 = qmod(0), q1 = qmod(1), q2 = qmod(2);
    if (v0 < q0) return crt_decode_canon(v0, (uint32_t)(v1 % q1), (uint32_t)(v2 % q2));
    uint64_t x = v0;
    const uint64_t a1 = v1 % q1, x1 = x % q1;
    const uint64_t u1 = ((a1 + q1 - x1) % q1) * CRT_C1 % q1;
    x += u1 * q0;  // x01 < Q01 + 2^32 < 2^56
    const uint64_t a2 = v2 % q2, x2 = x % q2;
    const uint64_t u2 = ((a2 + q2 - x2) % q2) * CRT_C2 % q2;
    return crt_round_split(x, (uint32_t)u2);
}

}  // namespace ipfe
