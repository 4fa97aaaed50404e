// host_tables.h -- host-side constant tables of the product library: AES
// T-tables and key schedule (FIPS-197), NTT twiddles from their closed form
// w^(+-bitrev12(k)) with Shoup companions and q-representative flags, and the
// N^-1 / floor(Q/p) residues (src/params.h:14-28, src/consts.h:7-11).
#pragma once
#include <stdint.h>
#include <vector>

#include "kernels.h"

namespace ipfe {

struct HostTables {
    std::vector<uint32_t> aes_te;  // 4 x 256
    std::vector<uint32_t> tw_all;  // [4][3][N]: psi|flag, psi', psi_inv|flag, psi_inv'
    NinvConsts ninv;
    ScaleConsts scale;
    BarrettConsts barrett;
    uint8_t sbox[256];
};

const HostTables &host_tables();
void aes256_expand_key(const uint8_t key[32], uint32_t rk[60]);

}  // namespace ipfe
