// kernels.h -- launch wrappers of the gfx950 kernels (host side, C++ linkage).
// Every launcher returns 0 or -1 with the error text set through ipfe_set_error.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>
#include <stddef.h>

#include "ipfe_common.h"

namespace ipfe {

struct NinvConsts {
    uint32_t ninv[3], ninvp[3], flag[3];
};
struct ScaleConsts {
    uint32_t scale[3];  // floor(Q/p) mod q_j
};
struct BarrettConsts {
    uint32_t mu_hi[3], mu_lo[3];
};

// ----- NTT domain (k_ntt.hip) -----
struct EncMsg {
    const uint32_t *x;      // message [len]
    const uint32_t *rhat;   // NTT(r) [3][N]
    const uint32_t *rhatp;  // Shoup companions of rhat
    uint32_t *c;            // c_0 .. c_{len-1} [len][3][N]; hold the noise f on entry
    uint32_t *cl;           // c_L [3][N]; holds f_0 on entry
    int len;
    int pad;
};
struct EncItem {
    int msg, poly;
};

int launch_ntt_fwd(const uint32_t *in, uint32_t *out, int npoly_limbs, const uint32_t *tw_all, hipStream_t st);
int launch_ntt_inv(const uint32_t *in, uint32_t *out, int npoly_limbs, const uint32_t *tw_all, const NinvConsts &nc,
                   hipStream_t st);
int launch_shoup_prep(const uint32_t *v, uint32_t *vp, size_t n_total, hipStream_t st);
int launch_setup_pk(const uint32_t *msk, const uint32_t *e, const uint32_t *ahat, const uint32_t *ahatp, uint32_t *mpk,
                    int l, const uint32_t *tw_all, hipStream_t st);
int launch_enc_poly(const EncItem *items, int nitems, const EncMsg *msgs, const uint32_t *mpk, int mpk_l,
                    const uint32_t *tw_all, const NinvConsts &nc, const ScaleConsts &sc, hipStream_t st);
int launch_point_mul(const uint32_t *a, const uint32_t *b, uint32_t *out, size_t n_total, const BarrettConsts &bc,
                     hipStream_t st);

// ----- Sampler (k_sampler.hip) -----
// One AES-256-CTR Gaussian stream cut into 384-byte chunks (24 AES blocks).
// Chunk c of the stream uses counters 24c .. 24c+23.  Calls take slen accepted
// candidates each; call 0 writes to out0, call k >= 1 to out1 + (k-1)*out_stride.
struct GaussStream {
    uint32_t rk[60];         // expanded AES-256 key
    uint32_t nonce;          // 16-bit nonce
    int which;               // sampler width index into the device tables: 1, 2, 3 (S1, S2, S3 of the
                             // reference set), 4, 5 (the S2 / S3 widths of the N = 2048 set)
    int ncalls;              // number of calls
    int nchunks;             // speculative window length (chunks)
    long long chunk_base;    // index of this stream's chunk 0 in the chunk arrays
    uint32_t *out0;          // [nmod][slen] destination of call 0
    uint32_t *out1;          // destination of call 1
    long long out_stride;    // words between consecutive calls from call 1 on
    int coef0_only;          // 1: write only sample 0 of each call, nmod words at out0 + nmod*call
    int slen;                // samples per call (the parameter set's N)
    long long first_base;    // compact streams: index of call 0 in the per-call first-chunk array
    uint32_t q[4];           // output moduli, limbs 0 .. nmod-1
    uint32_t fold[4];        // mod_prime fold shifts per limb: a | b << 8 (q = 2^a - 2^b + 1)
    int nmod;                // output limbs
    int pad;
};

// Chunks per tile of the candidate pass (the host's tile_prefix counts tiles of
// this size): 63, so that a tile's 16 AES blocks per chunk take two rounds of
// 512 threads and its keystream (64 words per chunk) fits the LDS left beside
// the 64 KiB T-tables of two workgroups.
constexpr int GAUSS_TILE_CHUNKS = 63;

// Candidate pass: for every chunk of every stream, AES + CDT + Bernoulli;
// writes the accepted count and the accepted (z | sign << 63) values in order.
// full: compute all 24 AES blocks of each chunk (else the 17 that candidates
// read unless a chunk's uniform scan runs long, which sets status bit 4)
int launch_gauss_candidates(const GaussStream *d_streams, int nstreams, const long long *d_tile_prefix,
                            long long ntiles, uint8_t *cnt, uint64_t *vals, int *status, int full, hipStream_t st);
// Walk: per stream, the call boundaries from the counts; writes per chunk
// (call << 13 | offset) or -1.  status[stream] = 1 when the window was short.
int launch_gauss_walk(const GaussStream *d_streams, int nstreams, const uint8_t *cnt, int *pos, long long *firsts,
                      int *status, hipStream_t st);
// Compact streams: recompute the first non-empty chunk of every call and write sample 0.
int launch_gauss_firsts(const GaussStream *d_streams, int nstreams, const long long *d_first_prefix,
                        const long long *firsts, long long nitems, int *status, int full, hipStream_t st);
int launch_gauss_scatter(const GaussStream *d_streams, int nstreams, const long long *d_chunk_prefix,
                         long long nchunks_total, const uint8_t *cnt, const uint64_t *vals, const int *pos,
                         hipStream_t st);
int launch_fill_i32(int *p, int v, size_t n, hipStream_t st);

// Uniform polynomial a (src/sample.c:56-128) on device from the expanded key.
int launch_sample_polya(const uint32_t *d_rk, uint32_t *a, hipStream_t st);
// The same for any parameter set: n coefficients on nmod limbs of moduli q.
int launch_sample_polya_ps(const uint32_t *d_rk, uint32_t *a, int n, int nmod, const uint32_t *q, hipStream_t st);
// Raw keystream (tests): nblocks AES blocks from counter ctr0.
int launch_aes_ctr(const uint32_t *d_rk, uint32_t nonce, uint32_t ctr0, uint32_t nblocks, uint8_t *out, hipStream_t st);

// ----- Keys, decryption, CRT (k_scheme.hip) -----
// Sparse description of vectors y: for vector v, entries [off[v], off[v+1]) of
// (idx, val) with idx ascending; len = the slot count the reference loops over.
struct SparseY {
    const int *off;
    const int *idx;
    const uint32_t *val;
    int len;
};
// sk[v][3][N] = sum_i y_v[i] msk[i] with the reference's sequential fold.
int launch_keygen(const SparseY &y, int nvec, const uint32_t *msk, uint32_t *sk, hipStream_t st);
// d[v][3][N] = sum_i y_v[i] c_i (fold)  (first half of rlwe_sife_decrypt_gmp).
int launch_dec_sum(const SparseY &y, int nvec, const uint32_t *c, uint32_t *d, hipStream_t st);
// d[v] = sub_mod(d[v], prod[v]) elementwise.
int launch_poly_sub(uint32_t *d, const uint32_t *prod, size_t n_total, hipStream_t st);
// CRT + rounding of coefficient k of d[v][3][N] -> r[v]  (k = 0 for kvalues).
int launch_crt_round(const uint32_t *d, int nvec, int coef, uint32_t *r, hipStream_t st);
// out += sum over i of (2i+1) * w[i] mod 2^64 (out must be cleared first)
int launch_checksum(const uint32_t *w, size_t n, unsigned long long *out, hipStream_t st);

}  // namespace ipfe
