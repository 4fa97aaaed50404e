// i8_frag.h -- pieces shared by the int8 matrix-core contractions
// (k_contract_i8.hip, k_contract_frac.hip): the MFMA operand types, the
// fragment size, the global-to-LDS load types, the memory-wait helper and the
// tile order of a launch.
#pragma once

namespace ipfe {

typedef int v4i __attribute__((ext_vector_type(4)));
typedef int v16i __attribute__((ext_vector_type(16)));

// One operand fragment of v_mfma_i32_32x32x32_i8: 64 lanes x 16 bytes.  Lane l
// holds row / key (l mod 32) and k = 16 (l / 32) + j, j < 16; the result lane
// l, register i is D[8 (i / 4) + 4 (l / 32) + i mod 4][l mod 32].
constexpr int FRAG_V4 = 64;

typedef __attribute__((address_space(3))) void lds_void_t;
typedef __attribute__((address_space(1))) void glb_void_t;

// Waits until at most N of this wave's vector-memory loads are outstanding.
template <int N>
__device__ __forceinline__ void wait_vmcnt()
{
    static_assert(N >= 0 && N < 64, "vmcnt range");
    asm volatile("s_waitcnt vmcnt(%0)" ::"n"(N) : "memory");
}

// Tile t of a launch -> (row tile tr, key tile tk).  Tiles run in groups of
// gr row tiles; a group runs in sub-blocks of kg key tiles with the row tile
// fastest.  The blocks one XCD runs at a time take neighbouring t, so they
// share key tiles (consecutive t) and row tiles (t apart by gr): with gr x kg
// about the blocks resident on an XCD, the row and key digits of the k steps
// in flight stay in that XCD's L2.  gr = 1, kg = ntk is row-major order.
__device__ __forceinline__ void tile_rc(int t, int ntk, int ntr, int gr, int kg, int &tr, int &tk)
{
    const int g = t / (gr * ntk), i = t - g * gr * ntk;
    const int rg = min(gr, ntr - g * gr);  // row tiles of group g
    const int sb = i / (rg * kg), j = i - sb * rg * kg;
    tr = g * gr + j % rg;
    tk = sb * kg + j / rg;
}

// The same with the key tile fastest: within a group of gr row tiles, sub-
// blocks of kg key tiles run the key tiles of one row tile consecutively, so
// neighbouring blocks share the row tile (the heavier operand when rows carry
// more digit planes than keys) and blocks kg apart share key tiles.
__device__ __forceinline__ void tile_kr(int t, int ntk, int ntr, int gr, int kg, int &tr, int &tk)
{
    const int g = t / (gr * ntk), i = t - g * gr * ntk;
    const int rg = min(gr, ntr - g * gr);  // row tiles of group g
    const int sb = i / (rg * kg), j = i - sb * rg * kg;
    const int kw = min(kg, ntk - sb * kg);  // key tiles of sub-block sb
    tk = sb * kg + j % kw;
    tr = g * gr + j / kw;
}

}  // namespace ipfe
