// aes_dev.h -- AES-256 encryption of CTR counter blocks with T-tables in LDS.
//
// Counter block of stream (key, nonce) at counter t is BE16(nonce) || 0^10 ||
// BE32(t) (the block layout src/aes256ctr.c:27-40 produces).  Round keys are
// the FIPS-197 big-endian words; the four T-tables (1 KiB each) are staged in
// LDS by the calling workgroup.  The output block is returned as four
// little-endian words, i.e. word w holds bytes 4w..4w+3 of the keystream, which
// is how every consumer reads it (LE32 / LE64 loads).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace ipfe {

constexpr int AES_TABLE_WORDS = 4 * 256;

// a ^ b ^ c as one v_bitop3_b32 (truth table 0x96); the compiler does not
// form it from two XORs by itself.
__device__ __forceinline__ uint32_t xor3(uint32_t a, uint32_t b, uint32_t c)
{
    return __builtin_amdgcn_bitop3_b32(a, b, c, 0x96);
}

__device__ __forceinline__ void aes_stage_tables(uint32_t *lds_te, const uint32_t *__restrict__ te_global)
{
    for (int i = threadIdx.x; i < AES_TABLE_WORDS; i += blockDim.x) lds_te[i] = te_global[i];
}

// rk: 60 round-key words, uniform across the workgroup (scalar loads).
__device__ __forceinline__ void aes256_ctr_block(const uint32_t *__restrict__ rk, const uint32_t *te, uint32_t nonce,
                                                 uint32_t ctr, uint32_t (&o)[4])
{
    const uint32_t *T0 = te, *T1 = te + 256, *T2 = te + 512, *T3 = te + 768;
    uint32_t s0 = (nonce << 16) ^ rk[0];
    uint32_t s1 = rk[1];
    uint32_t s2 = rk[2];
    uint32_t s3 = ctr ^ rk[3];
#pragma unroll
    for (int r = 1; r < 14; ++r) {
        uint32_t t0 = xor3(xor3(T0[s0 >> 24], T1[(s1 >> 16) & 255], T2[(s2 >> 8) & 255]), T3[s3 & 255], rk[4 * r]);
        uint32_t t1 = xor3(xor3(T0[s1 >> 24], T1[(s2 >> 16) & 255], T2[(s3 >> 8) & 255]), T3[s0 & 255], rk[4 * r + 1]);
        uint32_t t2 = xor3(xor3(T0[s2 >> 24], T1[(s3 >> 16) & 255], T2[(s0 >> 8) & 255]), T3[s1 & 255], rk[4 * r + 2]);
        uint32_t t3 = xor3(xor3(T0[s3 >> 24], T1[(s0 >> 16) & 255], T2[(s1 >> 8) & 255]), T3[s2 & 255], rk[4 * r + 3]);
        s0 = t0; s1 = t1; s2 = t2; s3 = t3;
    }
    // Final round (SubBytes, ShiftRows, AddRoundKey): the S-box byte sits in
    // byte 3 of T2, byte 2 of T3, byte 1 of T0 and byte 0 of T1.
    uint32_t f0 = (T2[s0 >> 24] & 0xff000000u) ^ (T3[(s1 >> 16) & 255] & 0x00ff0000u) ^
                  (T0[(s2 >> 8) & 255] & 0x0000ff00u) ^ (T1[s3 & 255] & 0x000000ffu) ^ rk[56];
    uint32_t f1 = (T2[s1 >> 24] & 0xff000000u) ^ (T3[(s2 >> 16) & 255] & 0x00ff0000u) ^
                  (T0[(s3 >> 8) & 255] & 0x0000ff00u) ^ (T1[s0 & 255] & 0x000000ffu) ^ rk[57];
    uint32_t f2 = (T2[s2 >> 24] & 0xff000000u) ^ (T3[(s3 >> 16) & 255] & 0x00ff0000u) ^
                  (T0[(s0 >> 8) & 255] & 0x0000ff00u) ^ (T1[s1 & 255] & 0x000000ffu) ^ rk[58];
    uint32_t f3 = (T2[s3 >> 24] & 0xff000000u) ^ (T3[(s0 >> 16) & 255] & 0x00ff0000u) ^
                  (T0[(s1 >> 8) & 255] & 0x0000ff00u) ^ (T1[s2 & 255] & 0x000000ffu) ^ rk[59];
    o[0] = __builtin_bswap32(f0);
    o[1] = __builtin_bswap32(f1);
    o[2] = __builtin_bswap32(f2);
    o[3] = __builtin_bswap32(f3);
}

// Bank-replicated variant: T0 stored 32 times so that lane l reads copy
// l mod 32 -- entry x of copy c at word 32 x + c, i.e. in bank c, so the 32
// lanes of a half-wave never collide.  T1..T3 are rotations of T0.
constexpr int AES_REP_WORDS = 256 * 32;

__device__ __forceinline__ void aes_stage_tables_rep(uint32_t *lds_rep, const uint32_t *__restrict__ te_global)
{
    for (int i = threadIdx.x; i < AES_REP_WORDS; i += blockDim.x) lds_rep[i] = te_global[i >> 5];
}

__device__ __forceinline__ uint32_t ror32(uint32_t x, int n) { return __builtin_amdgcn_alignbit(x, x, n); }

__device__ __forceinline__ void aes256_ctr_block_rep(const uint32_t *__restrict__ rk, const uint32_t *te_rep,
                                                     uint32_t nonce, uint32_t ctr, uint32_t (&o)[4])
{
    const uint32_t *T = te_rep + (threadIdx.x & 31);
#define TL(x) T[(x) << 5]
    uint32_t s0 = (nonce << 16) ^ rk[0];
    uint32_t s1 = rk[1];
    uint32_t s2 = rk[2];
    uint32_t s3 = ctr ^ rk[3];
#pragma unroll
    for (int r = 1; r < 14; ++r) {
        uint32_t t0 = xor3(xor3(TL(s0 >> 24), ror32(TL((s1 >> 16) & 255), 8), ror32(TL((s2 >> 8) & 255), 16)),
                           ror32(TL(s3 & 255), 24), rk[4 * r]);
        uint32_t t1 = xor3(xor3(TL(s1 >> 24), ror32(TL((s2 >> 16) & 255), 8), ror32(TL((s3 >> 8) & 255), 16)),
                           ror32(TL(s0 & 255), 24), rk[4 * r + 1]);
        uint32_t t2 = xor3(xor3(TL(s2 >> 24), ror32(TL((s3 >> 16) & 255), 8), ror32(TL((s0 >> 8) & 255), 16)),
                           ror32(TL(s1 & 255), 24), rk[4 * r + 2]);
        uint32_t t3 = xor3(xor3(TL(s3 >> 24), ror32(TL((s0 >> 16) & 255), 8), ror32(TL((s1 >> 8) & 255), 16)),
                           ror32(TL(s2 & 255), 24), rk[4 * r + 3]);
        s0 = t0; s1 = t1; s2 = t2; s3 = t3;
    }
    // S-box byte of T0[x] is byte 1 (and byte 2)
    uint32_t f0 = ((TL(s0 >> 24) << 16) & 0xff000000u) ^ (TL((s1 >> 16) & 255) & 0x00ff0000u) ^
                  (TL((s2 >> 8) & 255) & 0x0000ff00u) ^ ((TL(s3 & 255) >> 8) & 0xffu) ^ rk[56];
    uint32_t f1 = ((TL(s1 >> 24) << 16) & 0xff000000u) ^ (TL((s2 >> 16) & 255) & 0x00ff0000u) ^
                  (TL((s3 >> 8) & 255) & 0x0000ff00u) ^ ((TL(s0 & 255) >> 8) & 0xffu) ^ rk[57];
    uint32_t f2 = ((TL(s2 >> 24) << 16) & 0xff000000u) ^ (TL((s3 >> 16) & 255) & 0x00ff0000u) ^
                  (TL((s0 >> 8) & 255) & 0x0000ff00u) ^ ((TL(s1 & 255) >> 8) & 0xffu) ^ rk[58];
    uint32_t f3 = ((TL(s3 >> 24) << 16) & 0xff000000u) ^ (TL((s0 >> 16) & 255) & 0x00ff0000u) ^
                  (TL((s1 >> 8) & 255) & 0x0000ff00u) ^ ((TL(s2 & 255) >> 8) & 0xffu) ^ rk[59];
#undef TL
    o[0] = __builtin_bswap32(f0);
    o[1] = __builtin_bswap32(f1);
    o[2] = __builtin_bswap32(f2);
    o[3] = __builtin_bswap32(f3);
}

// Table accessors for the generic round code below: t0(w, k) = T0[byte k of
// w] and t2(w, k) = T2[byte k of w] = ror16(T0[byte k of w]).  A round word
//   T0[a] ^ ror8(T0[b]) ^ ror16(T0[c]) ^ ror24(T0[d])
//     = (T0[a] ^ T2[c]) ^ ror8(T0[b] ^ T2[d])
// then costs one rotation when T2 is a table of its own.
struct AesTabRep {  // T0 only, bank-replicated (32 KiB): entry x of copy c at word 32 x + c
    const uint32_t *T;
    __device__ explicit AesTabRep(const uint32_t *te) : T(te + (threadIdx.x & 31)) {}
    __device__ uint32_t t0(uint32_t w, int k) const { return T[((w >> (8 * k)) & 255) << 5]; }
    __device__ uint32_t t2(uint32_t w, int k) const { return ror32(t0(w, k), 16); }
};

// T0 and T2, bank-replicated (64 KiB): entry x of copy c at word 64 x + c
// (T0) and 64 x + 32 + c (T2).  The byte address 256 x + 4 c (+ 128) is one
// v_perm_b32: byte 0 from the lane's offset, byte 1 = byte k of w.
constexpr int AES_REP2_WORDS = 256 * 64;
struct AesTabRep2 {
    const char *base;
    uint32_t l0, l2;
    __device__ explicit AesTabRep2(const uint32_t *te)
        : base(reinterpret_cast<const char *>(te)), l0(4u * (threadIdx.x & 31)), l2(128u + 4u * (threadIdx.x & 31))
    {
    }
    __device__ uint32_t t0(uint32_t w, int k) const
    {
        return *reinterpret_cast<const uint32_t *>(base + __builtin_amdgcn_perm(w, l0, 0x0C0C0400u | ((4u + k) << 8)));
    }
    __device__ uint32_t t2(uint32_t w, int k) const
    {
        return *reinterpret_cast<const uint32_t *>(base + __builtin_amdgcn_perm(w, l2, 0x0C0C0400u | ((4u + k) << 8)));
    }
};

__device__ __forceinline__ void aes_stage_tables_rep2(uint32_t *lds, const uint32_t *__restrict__ te_global)
{
    for (int i = threadIdx.x; i < AES_REP2_WORDS; i += blockDim.x) {
        const uint32_t v = te_global[i >> 6];
        lds[i] = (i & 32) ? ror32(v, 16) : v;
    }
}

// One middle round with round-key words rk[0..3].
template <class Tab>
__device__ __forceinline__ void aes_round(const Tab &tb, uint32_t &s0, uint32_t &s1, uint32_t &s2, uint32_t &s3,
                                          const uint32_t *rk)
{
    const uint32_t t0 = xor3(tb.t0(s0, 3), tb.t2(s2, 1), rk[0]) ^ ror32(tb.t0(s1, 2) ^ tb.t2(s3, 0), 8);
    const uint32_t t1 = xor3(tb.t0(s1, 3), tb.t2(s3, 1), rk[1]) ^ ror32(tb.t0(s2, 2) ^ tb.t2(s0, 0), 8);
    const uint32_t t2 = xor3(tb.t0(s2, 3), tb.t2(s0, 1), rk[2]) ^ ror32(tb.t0(s3, 2) ^ tb.t2(s1, 0), 8);
    const uint32_t t3 = xor3(tb.t0(s3, 3), tb.t2(s1, 1), rk[3]) ^ ror32(tb.t0(s0, 2) ^ tb.t2(s2, 0), 8);
    s0 = t0;
    s1 = t1;
    s2 = t2;
    s3 = t3;
}

// Rounds 3..14 from the state after round 2; the final round takes the S-box
// byte from byte 3 of T2, byte 2 and 1 of T0 and byte 0 of T2.
template <class Tab>
__device__ __forceinline__ void aes_rounds_from3(const Tab &tb, uint32_t s0, uint32_t s1, uint32_t s2, uint32_t s3,
                                                 const uint32_t *__restrict__ rk, uint32_t (&o)[4])
{
#pragma unroll
    for (int r = 3; r < 14; ++r) aes_round(tb, s0, s1, s2, s3, rk + 4 * r);
    const uint32_t f0 = (tb.t2(s0, 3) & 0xff000000u) ^ (tb.t0(s1, 2) & 0x00ff0000u) ^ (tb.t0(s2, 1) & 0x0000ff00u) ^
                        (tb.t2(s3, 0) & 0x000000ffu) ^ rk[56];
    const uint32_t f1 = (tb.t2(s1, 3) & 0xff000000u) ^ (tb.t0(s2, 2) & 0x00ff0000u) ^ (tb.t0(s3, 1) & 0x0000ff00u) ^
                        (tb.t2(s0, 0) & 0x000000ffu) ^ rk[57];
    const uint32_t f2 = (tb.t2(s2, 3) & 0xff000000u) ^ (tb.t0(s3, 2) & 0x00ff0000u) ^ (tb.t0(s0, 1) & 0x0000ff00u) ^
                        (tb.t2(s1, 0) & 0x000000ffu) ^ rk[58];
    const uint32_t f3 = (tb.t2(s3, 3) & 0xff000000u) ^ (tb.t0(s0, 2) & 0x00ff0000u) ^ (tb.t0(s1, 1) & 0x0000ff00u) ^
                        (tb.t2(s2, 0) & 0x000000ffu) ^ rk[59];
    o[0] = __builtin_bswap32(f0);
    o[1] = __builtin_bswap32(f1);
    o[2] = __builtin_bswap32(f2);
    o[3] = __builtin_bswap32(f3);
}

// Block at counter ctr of stream (rk, nonce).
template <class Tab>
__device__ __forceinline__ void aes256_ctr_block_t(const Tab &tb, const uint32_t *__restrict__ rk, uint32_t nonce,
                                                   uint32_t ctr, uint32_t (&o)[4])
{
    uint32_t s0 = (nonce << 16) ^ rk[0], s1 = rk[1], s2 = rk[2], s3 = ctr ^ rk[3];
    aes_round(tb, s0, s1, s2, s3, rk + 4);
    aes_round(tb, s0, s1, s2, s3, rk + 8);
    aes_rounds_from3(tb, s0, s1, s2, s3, rk, o);
}

// Counter-mode shortcut.  The counter sits in state word 3, so in round 1
// only the lookup of its low byte changes between the 256 counters that share
// bits 8..31: t0 = C0 ^ ror8(T2[(ctr ^ rk3) & 255]), and t1..t3 are constant.
// In round 2 each output word then has one lookup that depends on t0 and the
// rest do not.  aes256_ctr_window computes, once per such window of counters,
// w = {C0, t1, t2, t3, D0..D3} with D_j the constant part of round-2 word j;
// aes256_ctr_block_win finishes a block from it with 5 lookups in rounds 1-2
// instead of 32.  Both give the block of aes256_ctr_block_t.
template <class Tab>
__device__ __forceinline__ void aes256_ctr_window(const Tab &tb, const uint32_t *__restrict__ rk, uint32_t nonce,
                                                  uint32_t ctr_hi, uint32_t *w)
{
    const uint32_t s0 = (nonce << 16) ^ rk[0], s1 = rk[1], s2 = rk[2], s3 = (ctr_hi << 8) ^ rk[3];
    const uint32_t t1 = tb.t0(s1, 3) ^ tb.t2(s3, 1) ^ ror32(tb.t0(s2, 2) ^ tb.t2(s0, 0), 8) ^ rk[5];
    const uint32_t t2 = tb.t0(s2, 3) ^ tb.t2(s0, 1) ^ ror32(tb.t0(s3, 2) ^ tb.t2(s1, 0), 8) ^ rk[6];
    const uint32_t t3 = tb.t0(s3, 3) ^ tb.t2(s1, 1) ^ ror32(tb.t0(s0, 2) ^ tb.t2(s2, 0), 8) ^ rk[7];
    w[0] = tb.t0(s0, 3) ^ tb.t2(s2, 1) ^ ror32(tb.t0(s1, 2), 8) ^ rk[4];
    w[1] = t1;
    w[2] = t2;
    w[3] = t3;
    w[4] = tb.t2(t2, 1) ^ ror32(tb.t0(t1, 2) ^ tb.t2(t3, 0), 8) ^ rk[8];
    w[5] = tb.t0(t1, 3) ^ tb.t2(t3, 1) ^ ror32(tb.t0(t2, 2), 8) ^ rk[9];
    w[6] = tb.t0(t2, 3) ^ ror32(tb.t0(t3, 2) ^ tb.t2(t1, 0), 8) ^ rk[10];
    w[7] = tb.t0(t3, 3) ^ tb.t2(t1, 1) ^ ror32(tb.t2(t2, 0), 8) ^ rk[11];
}

template <class Tab>
__device__ __forceinline__ void aes256_ctr_block_win(const Tab &tb, const uint32_t *__restrict__ rk,
                                                     const uint32_t *w, uint32_t ctr, uint32_t (&o)[4])
{
    const uint32_t t0 = w[0] ^ ror32(tb.t2(ctr ^ rk[3], 0), 8);
    const uint32_t s0 = tb.t0(t0, 3) ^ w[4];
    const uint32_t s1 = ror32(tb.t2(t0, 0), 8) ^ w[5];
    const uint32_t s2 = tb.t2(t0, 1) ^ w[6];
    const uint32_t s3 = ror32(tb.t0(t0, 2), 8) ^ w[7];
    aes_rounds_from3(tb, s0, s1, s2, s3, rk, o);
}

}  // namespace ipfe
