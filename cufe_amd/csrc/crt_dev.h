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

// round_extract(crt_reverse(v0, v1, v2)) for one coefficient, without
// 128-bit division.  Garner's x = x01 + u2 Q01 (x01 = v0 + u1 q0 < Q01 +
// 2^32 since v0 enters unreduced, u2 < q2, Q = Q01 q2).  With p u2 = A q2 + B
// (B < q2):
//     p x = A Q + N',   N' = B Q01 + p x01 < Q + p (Q01 + 2^32) < 2 Q,
// so floor(p x / Q) = A + [N' >= Q], and rounding adds one when the
// remainder exceeds floor(Q/2) (Q is odd: when twice the remainder exceeds Q):
//     round(p x / Q) = A + [2 N' >= Q] + [2 N' >= 3 Q].
__host__ __device__ inline uint32_t crt_decode(uint32_t v0, uint32_t v1, uint32_t v2)
{
    const uint64_t q0 = qmod(0), q1 = qmod(1), q2 = qmod(2);
    uint64_t x = v0;
    uint64_t a1 = v1 % q1, x1 = x % q1;
    uint64_t u1 = ((a1 + q1 - x1) % q1) * CRT_C1 % q1;
    x += u1 * q0;  // x01 < Q01 + 2^32 < 2^56
    uint64_t a2 = v2 % q2, x2 = x % q2;
    const uint64_t u2 = ((a2 + q2 - x2) % q2) * CRT_C2 % q2;
    const uint64_t pu = u2 * IPFE_P;  // < 2^47
    const uint64_t A = pu / q2, B = pu - A * q2;
    const u128 n2 = 2 * ((u128)B * CRT_Q01 + (u128)x * IPFE_P);  // 2 N' < 2^88
    const u128 Q = crt_Q();
    return (uint32_t)(A + (n2 >= Q ? 1 : 0) + (n2 >= 3 * Q ? 1 : 0));
}

}  // namespace ipfe
