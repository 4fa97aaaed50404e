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
#include "ipfe_common.h"

namespace ipfe {

typedef unsigned __int128 u128;

// (q0 q1 q2), q0 q1, q0^-1 mod q1, (q0 q1)^-1 mod q2 (residues of
// SIFE_CRT_CONSTS, src/params.h:17).
constexpr uint64_t CRT_Q01 = 16760833ull * 2147352577ull;
constexpr uint64_t CRT_C1 = 206923011ull;
constexpr uint64_t CRT_C2 = 2935204199007202395ull % 2130706433ull;

__host__ __device__ inline u128 crt_Q() { return (u128)CRT_Q01 * 2130706433ull; }

__host__ __device__ inline u128 crt_garner(uint32_t v0, uint32_t v1, uint32_t v2)
{
    const uint64_t q0 = qmod(0), q1 = qmod(1), q2 = qmod(2);
    uint64_t x = v0;                       // < 2^25
    uint64_t a1 = v1 % q1, x1 = x % q1;
    uint64_t u1 = ((a1 + q1 - x1) % q1) * CRT_C1 % q1;
    x += u1 * q0;                          // < 2^56
    uint64_t a2 = v2 % q2, x2 = x % q2;
    uint64_t u2 = ((a2 + q2 - x2) % q2) * CRT_C2 % q2;
    return (u128)x + (u128)u2 * CRT_Q01;
}

// round(p x / Q) as the reference computes it; x < 2^88.
__host__ __device__ inline uint32_t crt_round(u128 x)
{
    const u128 Q = crt_Q();
    const u128 a = x * (u128)IPFE_P;
    // quotient estimate from doubles, then exact correction
    double xd = (double)(uint64_t)(x >> 64) * 18446744073709551616.0 + (double)(uint64_t)x;
    const double Qd = 76687145727357674227351553.0;
    uint64_t quo = (uint64_t)(xd * (double)IPFE_P / Qd);
    u128 prod = (u128)quo * Q;
    while (prod > a) { --quo; prod -= Q; }
    u128 rem = a - prod;
    while (rem >= Q) { ++quo; rem -= Q; }
    if (rem > Q / 2) ++quo;
    return (uint32_t)quo;
}

}  // namespace ipfe
