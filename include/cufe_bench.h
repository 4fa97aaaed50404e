/*
 * cufe_bench.h -- C ABI of libcufe_bench.so, the measurement library that sits
 * beside libcufe_ipfe.so.  It holds the microbenchmarks that give the measured
 * ceilings the rooflines are stated against; the drop-in library does not
 * contain them.  It links libcufe_ipfe.so (device state, AES tables) and uses
 * the same device.  No counterpart in the reference, which has no
 * measurement code beyond rdtsc timers (main.c:629-680).
 */
#ifndef CUFE_BENCH_H
#define CUFE_BENCH_H

#ifdef __cplusplus
extern "C" {
#endif

/* Measured ceilings of this GPU: kind 0 = fp64 FMA (TFLOP/s), 1 = 32-bit
 * mul_lo + mul_hi + add (T int-ops/s), 2 = device copy (GB/s read + write),
 * 3 = AES-256-CTR keystream (G blocks/s), 4 = the same with the
 * bank-replicated table, 5 = int8 MFMA 32x32x32 (T ops/s), 6 = fp64 MFMA
 * 16x16x4 (TFLOP/s), 7 = NTT butterflies in registers (G/s), 8 = int8 MFMA
 * on random operands, sustained over 8 back-to-back launches (T ops/s), 9 =
 * AES-256-CTR with T0 and T2 replicated per bank (G blocks/s).  Kinds 4 and 9
 * also check every table layout against the four-table AES and fail if they
 * differ.  Returns 0, or -1 with the reason in cufe_last_error(). */
int cufe_microbench(int kind, double *result);

#ifdef __cplusplus
}
#endif
#endif
