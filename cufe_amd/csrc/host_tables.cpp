// host_tables.cpp -- see host_tables.h.
#include "host_tables.h"

#include <mutex>

#include "crt_dev.h"

namespace ipfe {

namespace {

uint8_t gmul(uint8_t a, uint8_t b)
{
    uint8_t r = 0;
    while (b) {
        if (b & 1) r ^= a;
        a = (uint8_t)((a << 1) ^ ((a & 0x80) ? 0x1b : 0));
        b >>= 1;
    }
    return r;
}

uint32_t pw(uint64_t b, uint64_t e, uint64_t q)
{
    uint64_t r = 1;
    b %= q;
    while (e) {
        if (e & 1) r = r * b % q;
        b = b * b % q;
        e >>= 1;
    }
    return (uint32_t)r;
}

uint32_t brev12(uint32_t k)
{
    uint32_t r = 0;
    for (int i = 0; i < 12; ++i) r |= ((k >> i) & 1u) << (11 - i);
    return r;
}

// primitive 2N-th roots of unity whose bit-reversed powers are the reference
// twiddle tables psi / psi_inv (src/consts.h:9-11)
const uint32_t kRoot[3] = {3689u, 1689795u, 456478u};

HostTables build()
{
    HostTables t;
    // S-box from the multiplicative inverse in GF(2^8) and the affine map
    for (int x = 0; x < 256; ++x) {
        uint8_t inv = 0;
        for (int y = 1; x && y < 256; ++y)
            if (gmul((uint8_t)x, (uint8_t)y) == 1) { inv = (uint8_t)y; break; }
        uint8_t s = inv;
        for (int k = 1; k <= 4; ++k) s ^= (uint8_t)((inv << k) | (inv >> (8 - k)));
        t.sbox[x] = (uint8_t)(s ^ 0x63);
    }
    t.aes_te.resize(1024);
    for (int x = 0; x < 256; ++x) {
        uint8_t s = t.sbox[x];
        uint32_t w = ((uint32_t)gmul(s, 2) << 24) | ((uint32_t)s << 16) | ((uint32_t)s << 8) | gmul(s, 3);
        t.aes_te[x] = w;
        t.aes_te[256 + x] = (w >> 8) | (w << 24);
        t.aes_te[512 + x] = (w >> 16) | (w << 16);
        t.aes_te[768 + x] = (w >> 24) | (w << 8);
    }
    t.tw_all.resize(4 * 3 * IPFE_N);
    for (int s = 0; s < 3; ++s) {
        const uint64_t q = qmod(s);
        const uint32_t winv = pw(kRoot[s], q - 2, q);
        for (uint32_t k = 0; k < IPFE_N; ++k) {
            uint32_t w = pw(kRoot[s], brev12(k), q);
            uint32_t wi = pw(winv, brev12(k), q);
            uint32_t fw = mod_prime_exact(q * w, s) == q ? 1u : 0u;
            uint32_t fwi = mod_prime_exact(q * wi, s) == q ? 1u : 0u;
            t.tw_all[(0 * 3 + s) * IPFE_N + k] = w | (fw << 31);
            t.tw_all[(1 * 3 + s) * IPFE_N + k] = (uint32_t)(((uint64_t)w << 32) / q);
            t.tw_all[(2 * 3 + s) * IPFE_N + k] = wi | (fwi << 31);
            t.tw_all[(3 * 3 + s) * IPFE_N + k] = (uint32_t)(((uint64_t)wi << 32) / q);
        }
        const uint32_t ninv = pw(IPFE_N, q - 2, q);
        t.ninv.ninv[s] = ninv;
        t.ninv.ninvp[s] = (uint32_t)(((uint64_t)ninv << 32) / q);
        t.ninv.flag[s] = mod_prime_exact(q * ninv, s) == q ? 1u : 0u;
        t.scale.scale[s] = (uint32_t)((crt_Q() / IPFE_P) % q);
        Barrett b = make_barrett((uint32_t)q);
        t.barrett.mu_hi[s] = b.mu_hi;
        t.barrett.mu_lo[s] = b.mu_lo;
    }
    return t;
}

}  // namespace

const HostTables &host_tables()
{
    static std::once_flag once;
    static HostTables tables;
    std::call_once(once, [] { tables = build(); });
    return tables;
}

// FIPS-197 AES-256 key expansion (the schedule src/aes256ctr.c:83-139 derives
// with aeskeygenassist), big-endian words.
void aes256_expand_key(const uint8_t key[32], uint32_t rk[60])
{
    static const uint8_t rcon[7] = {0x01, 0x02, 0x04, 0x08, 0x10, 0x20, 0x40};
    const uint8_t *sb = host_tables().sbox;
    auto sub = [sb](uint32_t w) {
        return ((uint32_t)sb[w >> 24] << 24) | ((uint32_t)sb[(w >> 16) & 255] << 16) |
               ((uint32_t)sb[(w >> 8) & 255] << 8) | sb[w & 255];
    };
    for (int i = 0; i < 8; ++i)
        rk[i] = ((uint32_t)key[4 * i] << 24) | ((uint32_t)key[4 * i + 1] << 16) | ((uint32_t)key[4 * i + 2] << 8) |
                key[4 * i + 3];
    for (int i = 8; i < 60; ++i) {
        uint32_t t = rk[i - 1];
        if (i % 8 == 0)
            t = sub((t << 8) | (t >> 24)) ^ ((uint32_t)rcon[i / 8 - 1] << 24);
        else if (i % 8 == 4)
            t = sub(t);
        rk[i] = rk[i - 8] ^ t;
    }
}

}  // namespace ipfe
