/* A C caller of the reference GPU entry points, written the way the
 * reference's own code calls them (src/rlwe_sife.c:85-109 and :184-212 for
 * setup / encrypt, main.c:270-390 for keygen / decrypt per row), linked against
 * libcufe_ipfe.so instead of the CUDA objects.  It checks the decryption
 * identity round(p CRT(d) / Q) = <x, y> (mod p) for one message and two keys.
 * Prints "abi ok" and returns 0 on success. */
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "cufe_ipfe.h"

#define L CUFE_SIFE_L
#define NM CUFE_SIFE_NMODULI
#define NN CUFE_SIFE_N

int main(void)
{
    static uint32_t mpk[L + 1][NM][NN], msk[L][NM][NN];
    unsigned char seedA[32], seed2[32], seed3[32];
    for (int i = 0; i < 32; ++i) {
        seedA[i] = (unsigned char)(i + 1);
        seed2[i] = (unsigned char)(3 * i + 7);
        seed3[i] = (unsigned char)(5 * i + 11);
    }
    float t = 0.0f;
    /* rlwe_sife_setup_gui: a = sample_polya(seed) into mpk[L], then the GPU setup */
    if (cufe_sample_polya(seedA, &mpk[L][0][0])) return fprintf(stderr, "%s\n", cufe_last_error()), 1;
    rlwe_sife_setup_gpu(mpk, msk, seed2, seed3, &t);

    uint32_t *m = calloc(L, sizeof *m);
    uint32_t *c = malloc(sizeof(uint32_t) * (size_t)(L + 1) * NM * NN);
    uint32_t *y = calloc(2 * L, sizeof *y);
    uint32_t *sk = malloc(sizeof(uint32_t) * 2 * NM * NN);
    uint32_t *dy = malloc(sizeof(uint32_t) * 2 * NM * NN);
    if (!m || !c || !y || !sk || !dy) return 1;
    for (int i = 1; i < 124; i += 3) m[i] = (uint32_t)(i % 5);
    for (int i = 1; i < 124; i += 7) {
        y[i] = (uint32_t)(1 + i % 16);
        y[L + i] = 1;
    }
    rlwe_sife_encrypt_gpu(m, mpk, c, seed2, seed3, 1, &t);
    rlwe_sife_keygen_gpu(y, msk, sk, 2, &t);
    rlwe_sife_decrypt_gmp_gpu((uint32_t(*)[NM][NN])c, y, sk, dy, 2, &t);
    for (int k = 0; k < 2; ++k) {
        long long ip = 0;
        for (int i = 0; i < L; ++i) ip += (long long)m[i] * y[(size_t)k * L + i];
        const double r = cufe_round_extract2(dy + (size_t)k * NM * NN);
        if (((long long)r) % 50241 != ip % 50241) {
            fprintf(stderr, "key %d: decrypted %.0f, inner product %lld\n", k, r, ip);
            return 1;
        }
    }
    printf("abi ok (device time %.2f ms)\n", t);
    free(m); free(c); free(y); free(sk); free(dy);
    return 0;
}
