// ntt_variants.h -- host interface of the NTT shapes in k_ntt_gen.hip and of
// their autotuner.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <string>

namespace ipfe {

// Per-limb constants of a transform batch: polys are [npoly][nmod][N] and
// transform z works on limb z % nmod; twiddle tables are [4][nmod][N]
// (psi | flag, psi', psi_inv | flag, psi_inv').
struct NttArgs {
    uint32_t q[4], ninv[4], ninvp[4], ninvf[4];
    int nmod;
};

using NttRunFn = int (*)(bool inverse, const uint32_t *in, uint32_t *out, int ntrans, const NttArgs &a,
                         const uint32_t *tw_all, hipStream_t st);

constexpr int PSET_SLOTS = 3;  // tuning slots: one per parameter set

int ntt_variant_count(int logn);
const char *ntt_variant_name(int logn, int v);
// ref_set: the reference's three-limb set, for which the hand kernel exists.
bool ntt_variant_ok(int logn, int v, bool ref_set);
int ntt_run_variant(int logn, int v, bool inverse, const uint32_t *in, uint32_t *out, int ntrans, const NttArgs &a,
                    const uint32_t *tw_all, bool ref_set, hipStream_t st);
// The fastest agreeing variant of set `set_slot` for this direction (tuned on
// first use; CUFE_NTT_VARIANT=v forces v).  Returns the variant or -1.
int ntt_tuned_variant(int set_slot, int logn, bool inverse, const NttArgs &a, const uint32_t *tw_all, bool ref_set,
                      hipStream_t st);
std::string ntt_autotune_report();

}  // namespace ipfe
