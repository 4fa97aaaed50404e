// resident.h -- device-resident objects of the batched prediction path:
// functional keys of one SVM model, a store of compact encrypted rows, and the
// scratch of one prediction pass.  Used by the context API (predict.cpp) and
// the serving loop (serve.cpp).
#pragma once
#include <stdint.h>

#include <vector>

#include "../../include/cufe_ipfe.h"
#include "kernels_predict.h"
#include "runtime.h"

namespace ipfe {

// Functional keys of one linear SVM model, generated on the device from the
// master secret key (rlwe_sife_keygen over the key length, src/rlwe_sife.c:215-245),
// plus the contraction operand [3][S][kkey] and the vote parameters.
struct Model {
    int S = 0, nr_class = 0, maxidx = -1, l_key = 0;
    int kkey = 0;       // key columns: N, plus the y slots when fused
    bool fused = false; // sum part of decryption folded into the contraction
    int kb[3] = {0, 0, 0};
    unsigned int maxabs[3] = {0, 0, 0};
    DevBuf<int> y_off, y_idx;
    DevBuf<uint32_t> y_val;
    SparseY y{};
    DevBuf<uint32_t> sk;  // [S][3][N]
    DevBuf<double> kd;    // [3][S][kkey]
    DevBuf<double> coef, rho;
    DevBuf<int32_t> icoef;  // integer coefficients for the exact two-class vote (else empty)
    SvmModelDev mdev{};
    // int8-digit key store [3][spad/32][ksteps][nb][64][16] when every key
    // coefficient fits nb = 2 or 3 signed digits (nb = 0: fp64 kernels only)
    bool i8 = false;
    int nb = 0, spad = 0, ksteps = 0;
    DevBuf<int8_t> kdig;
    // the fraction-form contraction applies (two key digits, fused keys, every
    // key coefficient the same integer in the three limbs); margin [S] holds
    // each key's error bound (about sum_k |B_v[k]| / 2, units of 2^-72)
    bool frac = false;
    DevBuf<uint64_t> margin, margin88;  // the same bound at 88 bits (units of 2^-88) for the refinement
    DevBuf<int16_t> k16;  // [S][kkey] key values for the fraction form's exact kernel

    int build(const uint32_t *msk_dev, int l_key, const cufe_svm_model *m, hipStream_t st);
    size_t bytes() const;
};

// Compact encrypted rows of l slots: c_L and coefficient 0 of c_0..c_{l-1},
// kept as [3][cap][ldr] int32 (balanced; ldr = N + l rounded up to 16), the
// same rows blocked for the scalar-operand contraction ([3][(cap+64)/16][ldr][16]
// fp64), as int8 digit fragments for the matrix-core contraction
// ([3][(cap+128)/32][ksteps][4][64][16], ksteps = (N + l rounded up to 32) / 32),
// optionally as the nine fraction digits of each column's CRT value
// ([(cap+128)/32][ksteps][9][64][16], for the fraction-form contraction) with
// the next 16 bits of each as one int16 ([cap][ksteps x 32], for the 88-bit
// refinement of pairs near a rounding change), and [cap][l][3] canonical
// coefficient-0 residues.
struct RowStore {
    int l = 0, cap = 0, ldr = 0, ksteps = 0;
    DevBuf<int32_t> A;
    DevBuf<int8_t> Adig;
    DevBuf<int8_t> Fdig;   // empty unless reserved with fractions
    DevBuf<int16_t> Fref;  // likewise
    DevBuf<uint32_t> c0;
    // fp64 blocked copy, built on demand (only the fp64 fallback kernel reads
    // it): allocated at first use, converted per 64-row block when stale
    mutable DevBuf<double> A64;
    mutable std::vector<uint8_t> a64_ok;

    int reserve(int cap, int l, bool fractions = false);
    // rows [r0, r0+R) of the fp64 blocked copy current
    int ensure_a64(int r0, int R, hipStream_t st) const;
    // rows [r0, r0+n) from device buffers: cl [n][3][N] canonical, c0
    // [n][src_l][3] with src_l <= l slots (default l); slots past src_l read 0
    int put(int r0, int n, const uint32_t *cl_dev, const uint32_t *c0_dev, hipStream_t st, int src_l = -1);
    static size_t bytes_per_row(int l, bool fractions = false);
};

struct Scratch {
    DevBuf<uint32_t> D, kval;
    DevBuf<uint32_t> Dpart;  // split-K slices of small row ranges
    DevBuf<double> labels;
    DevBuf<int> flag;
    // fraction form: the pairs left to the exact kernel, their count (per
    // chunk), a list-overflow flag and the listed total (per predict_rows)
    DevBuf<uint32_t> amb, amb_count;
    DevBuf<int> amb_overflow;
    DevBuf<unsigned long long> amb_total;
    int frac_chunks = 0;  // chunks of the last predict_rows call that ran the fraction form
    // small row ranges may use the split-K launch (a server turns it off while
    // its other slot runs: there it would slow that slot's contraction)
    bool split_ok = true;
    int reserve(int rows, int S);
    // device bytes of Dpart that reserve(rows, S) allocates
    static size_t split_bytes(int rows, int S);
};

// Predicts rows [r0, r0+R) of the store against the model: labels and
// (optional) kvalues copied to host memory asynchronously on st.  Chunked so
// that the scratch stays within `scratch_rows` rows.  ms (optional, 3 floats,
// only when timing events are passed): contraction / key-value / vote time.
// ev[0..3]: before the contraction, after it, after the key values, after
// the vote; ev[4], ev[5]: after the contraction of limb 0 and of limb 1 (the
// fraction form: after its contraction and after the exact kernel).
// ms: contraction, key values, vote; limb_ms: the contraction per limb (the
// fraction form: its contraction, the exact kernel, 0).
struct PredictTimers {
    hipEvent_t ev[6];
    float ms[3];
    float limb_ms[3];
    bool pending = false;  // the last chunk's events are read after the caller's wait
};
// Adds the event intervals of a finished chunk to ms / limb_ms.
int predict_timers_collect(PredictTimers *tm);
int predict_rows(const Model &m, const RowStore &rs, int r0, int R, Scratch &sc, int scratch_rows, int mode,
                 double *labels_host, uint32_t *kval_host, PredictTimers *tm, hipStream_t st);
// Whether R rows of m are a range too large for the split-K launch (more
// than 256 rows or 2^19 (row, key) pairs): predict_rows runs it as full-GPU
// per-limb or fraction-form launches.  Smaller ranges take the split-K launch
// when the model's keys allow it (fused, two digits) and the scratch does
// (Scratch::split_ok), else small per-limb launches.
bool large_range(const Model &m, int R);

// Encrypts n rows x [n][l] (row b with nonce nonce0 + b) into store rows
// [r0, r0+n), against master keys mpk [l_key+1][3][N] on the device.
int encrypt_rows_compact(const uint32_t *mpk_dev, int l_key, RowStore &rs, int r0, const uint32_t *x, int n,
                         const unsigned char seed2[32], const unsigned char seed3[32], int nonce0, hipStream_t st);

// Contraction kernel choice: 3 = the fraction form where it applies, else the
// int8 matrix cores per limb (default), 2 = int8 per limb
// (CUFE_CONTRACT=limbs), 1 = fp64 scalar-operand (CUFE_CONTRACT=sgpr), 0 = fp64
// LDS tiles (CUFE_CONTRACT=lds).  The int8 kernels need key coefficients
// within two (fraction form) or three digits and rows starting on a 32-row
// boundary; otherwise the fp64 kernels run.  A fraction-form pass whose list
// of pairs for the exact kernel overflowed sets Scratch::amb_overflow; the
// caller then runs the pass again per limb.
int contract_mode();

}  // namespace ipfe
