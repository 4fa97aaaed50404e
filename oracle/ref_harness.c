/*
 * ref_harness.c -- calls into the reference's own compiled C path (built by
 * oracle/Makefile from the sources where they lie) for the checks that need
 * GMP values or the libsvm model struct, which ctypes cannot build directly.
 *
 * TEST INFRASTRUCTURE ONLY: linked into oracle/_ref/libref.so with the
 * reference sources; loaded by tests/ and bench.py's cpu_baseline leg.  It
 * holds no arithmetic of its own: every value comes from a reference function.
 *
 *   refh_crt_round        crt_reverse_gmp (src/crt.c:46-82), then
 *                         round_extract_gmp (src/rlwe_sife.c:288-332)
 *   refh_decrypt_round    rlwe_sife_decrypt_gmp (src/rlwe_sife.c:248-286),
 *                         then round_extract_gmp
 *   refh_svm_predict_cpu  svm_predict_cpu (main.c:137-267) on a model given
 *                         as dense support vectors
 */
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#include <gmp.h>

#include "params.h"
#include "crt.h"
#include "rlwe_sife.h"
#include "svm/svm.h"

/* main.c:137 (PERF form; main.c has no header) */
double svm_predict_cpu(const struct svm_model *model, uint32_t c[SIFE_L + 1][SIFE_NMODULI][SIFE_N],
                       const uint32_t msk[SIFE_L][SIFE_NMODULI][SIFE_N], double *CLOCK_predict_part1,
                       double *CLOCK_predict_part2, double *CLOCK_kp, double *CLOCK_dec, double *CLOCK_extract);

/* Exports a non-negative mpz below 2^96 as three little-endian 32-bit words. */
static void export96(const mpz_t v, uint32_t w[3])
{
    size_t cnt = 0;
    w[0] = w[1] = w[2] = 0;
    mpz_export(w, &cnt, -1, sizeof(uint32_t), 0, 0, v);
}

/* d [3][N] residues -> x [N][3] (the CRT value, 96-bit little-endian words)
 * and r [N] (round(p x / Q), no reduction mod p). */
void refh_crt_round(const uint32_t *d, uint32_t *x, uint32_t *r)
{
    mpz_t *a = malloc(sizeof(mpz_t) * SIFE_N);
    for (int k = 0; k < SIFE_N; ++k) mpz_init(a[k]);
    crt_reverse_gmp(a, (const uint32_t(*)[SIFE_N])d);
    for (int k = 0; k < SIFE_N; ++k) export96(a[k], x + 3 * k);
    round_extract_gmp(a);
    for (int k = 0; k < SIFE_N; ++k) {
        r[k] = (uint32_t)mpz_get_ui(a[k]);
        mpz_clear(a[k]);
    }
    free(a);
}

/* rlwe_sife_decrypt_gmp of c [786][3][N] with y [785] and sk [3][N], then
 * round_extract_gmp: r [N]. */
void refh_decrypt_round(uint32_t *c, const uint32_t *y, uint32_t *sk, uint32_t *r)
{
    mpz_t *dy = malloc(sizeof(mpz_t) * SIFE_N);
    double t = 0;
    for (int k = 0; k < SIFE_N; ++k) mpz_init(dy[k]);
    rlwe_sife_decrypt_gmp((uint32_t(*)[SIFE_NMODULI][SIFE_N])c, y, (uint32_t(*)[SIFE_N])sk, dy, &t);
    round_extract_gmp(dy);
    for (int k = 0; k < SIFE_N; ++k) {
        r[k] = (uint32_t)mpz_get_ui(dy[k]);
        mpz_clear(dy[k]);
    }
    free(dy);
}

/* svm_predict_cpu for one ciphertext c [786][3][N] against a model whose
 * support vectors are dense rows sv [l][785] (the non-zero entries become the
 * svm_node list convertToarray reads, main.c:30-41).  sv_coef [nr_class-1][l],
 * rho [nr_class (nr_class-1)/2].  Returns the label. */
double refh_svm_predict_cpu(uint32_t *c, const uint32_t *msk, int nr_class, int l, const int *nSV,
                            const int *label, const double *sv_coef, const double *rho, const uint32_t *sv)
{
    struct svm_model m;
    memset(&m, 0, sizeof(m));
    m.nr_class = nr_class;
    m.l = l;
    m.SV = malloc(sizeof(struct svm_node *) * (size_t)l);
    for (int s = 0; s < l; ++s) {
        int nnz = 0;
        for (int i = 0; i < SIFE_L; ++i) nnz += sv[(size_t)s * SIFE_L + i] != 0;
        struct svm_node *nd = malloc(sizeof(struct svm_node) * (size_t)(nnz + 1));
        int k = 0;
        for (int i = 0; i < SIFE_L; ++i)
            if (sv[(size_t)s * SIFE_L + i]) {
                nd[k].index = i;
                nd[k].value = (double)sv[(size_t)s * SIFE_L + i];
                ++k;
            }
        nd[k].index = -1;
        m.SV[s] = nd;
    }
    m.sv_coef = malloc(sizeof(double *) * (size_t)(nr_class - 1));
    for (int k = 0; k < nr_class - 1; ++k) m.sv_coef[k] = (double *)sv_coef + (size_t)k * l;
    m.rho = (double *)rho;
    m.label = (int *)label;
    m.nSV = (int *)nSV;
    double t1 = 0, t2 = 0, t3 = 0, t4 = 0, t5 = 0;
    double lab = svm_predict_cpu(&m, (uint32_t(*)[SIFE_NMODULI][SIFE_N])c, (const uint32_t(*)[SIFE_NMODULI][SIFE_N])msk,
                                 &t1, &t2, &t3, &t4, &t5);
    for (int s = 0; s < l; ++s) free(m.SV[s]);
    free(m.SV);
    free(m.sv_coef);
    return lab;
}
