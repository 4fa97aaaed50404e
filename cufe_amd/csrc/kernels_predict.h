// kernels_predict.h -- launchers of the batched prediction kernels (k_predict.hip).
#pragma once
#include "kernels.h"

namespace ipfe {

constexpr int SVM_MAX_CLASS = 16;

// Device view of the plain SVM model (main.c:338-389 reads these fields).
struct SvmModelDev {
    int nr_class;
    int nSV[SVM_MAX_CLASS];
    int start[SVM_MAX_CLASS];
    int label[SVM_MAX_CLASS];
    const double *sv_coef;  // [nr_class-1][S]
    const double *rho;      // [nr_class*(nr_class-1)/2]
    int exact_int;          // integer coefficients, sum |coef| * p < 2^53: sums exact in any order
    const int *sv_icoef;    // exact_int with every |coef| < 2^31: the coefficients as int32 (else null)
};

int launch_contract_keys(const uint32_t *sk, int S, int K, int sign, const SparseY &y, double *kd, unsigned int *maxabs,
                         hipStream_t st);
int launch_contract_slow(const int32_t *A, const double *B, int R, int S, int K, int lda, int ldb, int s,
                         const BarrettConsts &bc, uint32_t *D, hipStream_t st);
int launch_contract(const int32_t *A, const double *B, int R, int S, int K, int lda, int ldb, int s, int kb, uint32_t *D,
                    hipStream_t st);
// Row operand in scalar registers: A is the blocked fp64 store (rows in groups
// of 16, [k][16] per group, cap a multiple of 64), lda = its column count.
int launch_contract_sgpr(const double *A, const double *B, int R, int S, int K, int lda, int ldb, int s, int kb,
                         uint32_t *D, hipStream_t st);
// int8-digit contraction (k_contract_i8.hip): digit stores and the exact
// MFMA contraction of one limb.
size_t rows_digit_limb_bytes(int cap, int ksteps);
size_t keys_digit_limb_bytes(int spad, int ksteps, int nb);
int launch_rows_put(const uint32_t *cl, const uint32_t *c0, int src_l, int R, int r0, int cap, int ldr, int l_store,
                    int ksteps, int32_t *A, uint32_t *c0s, int8_t *dig, hipStream_t st);
int launch_keys_digits(const double *kd, int S, int K, int spad, int ksteps, int nb, int8_t *dig, hipStream_t st);
int launch_contract_i8(const int8_t *A, int ksteps_a, const int8_t *B, int ksteps, int nb, int R, int S, int s,
                       uint32_t *D, hipStream_t st, int variant = 0, const uint32_t *crt_d0 = nullptr,
                       const uint32_t *crt_d1 = nullptr, uint32_t *kval = nullptr);
// Small row ranges: the three limbs and slices of the k steps as blocks of one
// launch (two key digits), then the slices' sum, CRT and rounding into kval.
int split_k_slices(int R, int S, int ksteps);
int launch_contract_i8_split(const int8_t *A, size_t a_limb_bytes, int ksteps_a, const int8_t *B,
                             size_t b_limb_bytes, int ksteps, int R, int S, uint32_t *Dpart, uint32_t *kval,
                             hipStream_t st);
int i8_shape_count();
void i8_shape(int variant, int *wr, int *wk);
const char *i8_variant_label(int variant);
bool i8_variant_ok(int variant, int nb, int s);
// largest key coefficients the two- and three-digit key splits hold: the top
// digit (|b| + 128 (1 + 256 + ...)) / 256^(nb-1) must stay within [-128, 127]
constexpr unsigned int I8_KEY_MAX2 = 32512;
constexpr unsigned int I8_KEY_MAX3 = 8322943;
// Fraction-form contraction (k_contract_frac.hip): every limb in one int8
// contraction over the CRT value's 72-bit fractions, key values written
// directly; pairs too close to a rounding change are listed for k_frac_exact.
struct FracConsts {
    uint32_t u[3];            // (Q / q_j)^-1 mod q_j
    uint64_t rlo[3], rhi[3];  // floor(2^128 / q_j)
};
// words of a listed pair: row, key (bit 31: near a half-integer), V's 72 bits
constexpr int FRAC_AMB_WORDS = 5;
struct FracOut {
    uint32_t *kval;            // [R][S] key values
    const uint64_t *margin;    // [S]: error bound of each key's sum, units of 2^-72
    const uint64_t *margin88;  // [S]: the same at 88 bits, units of 2^-88
    uint32_t *amb;             // [amb_cap][FRAC_AMB_WORDS]: the pairs left to the exact kernel
    uint32_t *amb_count;       // zeroed before the launch
    int amb_cap;
};
FracConsts frac_consts();
// bytes of the fraction digit store [(cap + 128) / 32][ksteps][9][64][16]
size_t frac_digit_bytes(int cap, int ksteps);
// digits of rows [r0, r0 + R) into dig, and their 16 bits below the digits
// (int16 D) into ref [cap][ksteps x 32]
int launch_rows_frac(const uint32_t *cl, const uint32_t *c0, int src_l, int R, int r0, int cap, int ksteps,
                     int8_t *dig, int16_t *ref, hipStream_t st);
// keys: *neq set unless every coefficient is the same integer (within int16)
// in all limbs; margin[v] = the bound M_v on |x / Q - V 2^-72| in units of
// 2^-72 (M_v ~ sum_k |kd[0][v][k]| / 2), margin88[v] the same at 88 bits,
// both times 2^widen, at most 2^62 (widen <= 62; > 0 only in tests, to send
// many pairs to the exact kernel); k16 [S][K] = limb 0's values
int launch_keys_frac(const double *kd, int S, int K, int widen, uint64_t *margin, uint64_t *margin88, int16_t *k16,
                     int *neq, hipStream_t st);
// A: fraction digits at the first row (a multiple of 32), B: limb 0's
// two-digit key store
int launch_contract_frac(const int8_t *A, int ksteps_a, const int8_t *B, int ksteps, int R, int S, const FracOut &out,
                         hipStream_t st, int variant);
// the listed pairs: refined to 88 bits from ref (the rows' D values at the
// first row, row stride kf) unless refine is false, the rest from the
// balanced row store (A at the first row, limb j at + j limb_stride), with the
// key values k16 [S][K]; *overflow set when the list overflowed; *total (if
// given) += the listed pairs
int launch_frac_exact(const int32_t *A, size_t limb_stride, int ldr, const int16_t *ref, int kf, bool refine,
                      const int16_t *k16, int S, int K, const FracOut &out, int *overflow, unsigned long long *total,
                      hipStream_t st);
int frac_variant_count();
void frac_shape(int variant, int *wr, int *wk);
const char *frac_variant_label(int variant);
int launch_rows_blocked(const int32_t *a32, int cap, int cap64, int ldr, int r0, int R, double *a64, hipStream_t st);
int launch_crt_kvalue(const uint32_t *D, size_t n, uint32_t *kval, hipStream_t st);
int launch_kvalue(const uint32_t *D, const uint32_t *c0, int l, const SparseY &y, int R, int S, uint32_t *kval,
                  int *flag, hipStream_t st);
int launch_decide(const uint32_t *kval, int R, int S, const SvmModelDev &m, double *labels, hipStream_t st);
int launch_enc_coef0(const uint32_t *rhat, const uint32_t *mpk, const uint32_t *x, int R, int l, uint32_t *f0buf,
                     const NinvConsts &nc, const ScaleConsts &sc, const BarrettConsts &bc, hipStream_t st);

}  // namespace ipfe
