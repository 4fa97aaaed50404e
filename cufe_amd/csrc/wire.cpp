// wire.cpp -- a byte format for serving requests, ciphertexts and keys.
//
// The reference moves only in-memory [L+1][3][N] arrays between its host code
// and the GPU module.  A request that crosses a process or a network needs a
// self-describing container; this one is a 64-byte little-endian header and
// the residue arrays exactly as the C ABI takes them, so a parsed blob can be
// handed to the API without copying.
//
// Header (offsets in bytes):
//   0  magic "CUFEIPFE"        8  u16 version (1)     10 u16 kind
//   12 u16 limbs (3)           14 u16 reserved (0)    16 u32 N (4096)
//   20 u32 count               24 u32 l (slots)       28 u32 q[3]
//   40 u32 tag (caller's: model id, nonce base ...)   44 u32 CRC-32 of payload
//   48 u64 payload bytes       56 u64 reserved (0)
// Payload by kind:
//   1 rows:        c_L [count][3][N], then c0 [count][l][3]   (compact rows)
//   2 ciphertexts: [count][l+1][3][N]                          (index l = c_L)
//   3 public key:  [l+1][3][N], count = 1                      (index l = a_hat)
//   4 func. keys:  [count][3][N]
// Every residue must lie in [0, q_s] for its limb (q itself is kept: it is a
// value mod_prime returns, and the folds of keygen and decryption treat it
// differently from 0).
#include <stdio.h>
#include <string.h>

#include "../../include/cufe_ipfe.h"
#include "ipfe_common.h"

using ipfe::qmod;

namespace {

constexpr size_t kHeader = 64;
constexpr size_t kPoly = (size_t)3 * IPFE_N;
const char kMagic[8] = {'C', 'U', 'F', 'E', 'I', 'P', 'F', 'E'};

// CRC-32 (IEEE 802.3, reflected polynomial 0xEDB88320), table-driven.
struct Crc32 {
    uint32_t t[256];
    Crc32()
    {
        for (uint32_t i = 0; i < 256; ++i) {
            uint32_t c = i;
            for (int k = 0; k < 8; ++k) c = (c & 1) ? 0xEDB88320u ^ (c >> 1) : c >> 1;
            t[i] = c;
        }
    }
    uint32_t operator()(const uint8_t *p, size_t n) const
    {
        uint32_t c = 0xFFFFFFFFu;
        for (size_t i = 0; i < n; ++i) c = t[(c ^ p[i]) & 0xFF] ^ (c >> 8);
        return c ^ 0xFFFFFFFFu;
    }
};
const Crc32 &crc32() {
    static const Crc32 c;
    return c;
}

void put16(uint8_t *p, uint16_t v) { p[0] = (uint8_t)v; p[1] = (uint8_t)(v >> 8); }
void put32(uint8_t *p, uint32_t v) { for (int i = 0; i < 4; ++i) p[i] = (uint8_t)(v >> (8 * i)); }
void put64(uint8_t *p, uint64_t v) { for (int i = 0; i < 8; ++i) p[i] = (uint8_t)(v >> (8 * i)); }
uint16_t get16(const uint8_t *p) { return (uint16_t)(p[0] | (p[1] << 8)); }
uint32_t get32(const uint8_t *p) { uint32_t v = 0; for (int i = 0; i < 4; ++i) v |= (uint32_t)p[i] << (8 * i); return v; }
uint64_t get64(const uint8_t *p) { uint64_t v = 0; for (int i = 0; i < 8; ++i) v |= (uint64_t)p[i] << (8 * i); return v; }

// payload words of a kind, or 0 when the shape is invalid
size_t payload_words(int kind, int count, int l)
{
    if (count < 1 || l < 0 || l > IPFE_L) return 0;
    switch (kind) {
    case CUFE_WIRE_ROWS: return l >= 1 ? (size_t)count * (kPoly + (size_t)l * 3) : 0;
    case CUFE_WIRE_CIPHERTEXTS: return (size_t)count * (l + 1) * kPoly;
    case CUFE_WIRE_PUBLIC_KEY: return count == 1 ? (size_t)(l + 1) * kPoly : 0;
    case CUFE_WIRE_FUNCTIONAL_KEYS: return (size_t)count * kPoly;
    default: return 0;
    }
}

// limb of payload word w for each layout (to check residues against q)
int limb_of(int kind, size_t w, int count)
{
    switch (kind) {
    case CUFE_WIRE_ROWS: {
        const size_t cl_words = (size_t)count * kPoly;
        return w < cl_words ? (int)((w / IPFE_N) % 3) : (int)((w - cl_words) % 3);
    }
    default: return (int)((w / IPFE_N) % 3);
    }
}

}  // namespace

extern "C" size_t cufe_wire_size(int kind, int count, int l)
{
    const size_t w = payload_words(kind, count, l);
    return w ? kHeader + 4 * w : 0;
}

// Writes header + payload to out (cap bytes).  For rows, data = c_L
// [count][3][N] and data2 = c0 [count][l][3]; other kinds take data only.
// A value above its limb's modulus is an error.
extern "C" int cufe_wire_pack(int kind, const uint32_t *data, const uint32_t *data2, int count, int l, uint32_t tag,
                              void *out, size_t cap)
{
    const size_t words = payload_words(kind, count, l);
    if (!words || !data || !out || (kind == CUFE_WIRE_ROWS && !data2)) {
        ipfe_set_error("cufe_wire_pack: kind %d, count %d, l %d is not a valid shape", kind, count, l);
        return -1;
    }
    if (cap < kHeader + 4 * words) {
        ipfe_set_error("cufe_wire_pack: %zu bytes needed, %zu given", kHeader + 4 * words, cap);
        return -1;
    }
    uint8_t *h = (uint8_t *)out;
    uint8_t *pay = h + kHeader;
    const size_t first = kind == CUFE_WIRE_ROWS ? (size_t)count * kPoly : words;
    for (size_t w = 0; w < words; ++w) {
        const uint32_t v = w < first ? data[w] : data2[w - first];
        const uint32_t q = qmod(limb_of(kind, w, count));
        if (v > q) {
            ipfe_set_error("cufe_wire_pack: word %zu = %u is not a residue mod %u", w, v, q);
            return -1;
        }
        put32(pay + 4 * w, v);
    }
    memset(h, 0, kHeader);
    memcpy(h, kMagic, 8);
    put16(h + 8, 1);
    put16(h + 10, (uint16_t)kind);
    put16(h + 12, 3);
    put32(h + 16, IPFE_N);
    put32(h + 20, (uint32_t)count);
    put32(h + 24, (uint32_t)l);
    for (int s = 0; s < 3; ++s) put32(h + 28 + 4 * s, qmod(s));
    put32(h + 40, tag);
    put32(h + 44, crc32()(pay, 4 * words));
    put64(h + 48, 4 * (uint64_t)words);
    return 0;
}

// Validates a blob (magic, version, parameter set, shape, length, CRC, every
// residue within [0, q]) and returns its description and a pointer to the
// payload inside the blob (for rows: info->c0 points at the c0 block).
extern "C" int cufe_wire_parse(const void *blob, size_t len, cufe_wire_info *info)
{
    if (!blob || !info || len < kHeader) {
        ipfe_set_error("cufe_wire_parse: %zu bytes is shorter than the 64-byte header", len);
        return -1;
    }
    const uint8_t *h = (const uint8_t *)blob;
    if (memcmp(h, kMagic, 8) != 0) {
        ipfe_set_error("cufe_wire_parse: bad magic");
        return -1;
    }
    if (get16(h + 8) != 1) {
        ipfe_set_error("cufe_wire_parse: unsupported version %u", get16(h + 8));
        return -1;
    }
    if (get16(h + 12) != 3 || get32(h + 16) != IPFE_N || get32(h + 28) != qmod(0) || get32(h + 32) != qmod(1) ||
        get32(h + 36) != qmod(2)) {
        ipfe_set_error("cufe_wire_parse: blob is for another parameter set");
        return -1;
    }
    const int kind = get16(h + 10);
    const uint32_t count = get32(h + 20), l = get32(h + 24);
    if (count > (1u << 30) || l > (uint32_t)IPFE_L) {
        ipfe_set_error("cufe_wire_parse: count %u / l %u out of range", count, l);
        return -1;
    }
    const size_t words = payload_words(kind, (int)count, (int)l);
    if (!words || get64(h + 48) != 4 * (uint64_t)words || len < kHeader + 4 * words) {
        ipfe_set_error("cufe_wire_parse: kind %d, count %u, l %u does not match %zu payload bytes", kind, count, l,
                       len - kHeader);
        return -1;
    }
    const uint8_t *pay = h + kHeader;
    if (crc32()(pay, 4 * words) != get32(h + 44)) {
        ipfe_set_error("cufe_wire_parse: payload CRC mismatch");
        return -1;
    }
    for (size_t w = 0; w < words; ++w) {
        const uint32_t v = get32(pay + 4 * w);
        if (v > qmod(limb_of(kind, w, (int)count))) {
            ipfe_set_error("cufe_wire_parse: word %zu = %u is not a residue", w, v);
            return -1;
        }
    }
    info->kind = kind;
    info->count = (int)count;
    info->l = (int)l;
    info->tag = get32(h + 40);
    info->data = (const uint32_t *)pay;
    info->c0 = kind == CUFE_WIRE_ROWS ? (const uint32_t *)pay + (size_t)count * kPoly : nullptr;
    return 0;
}
