// ntt_gen.h -- negacyclic NTT / INTT of one N-point limb (N = 2^LOGN) by T
// threads, for any N and any split of the LOGN radix-2 stages into register
// phases; bit-exact with CT_forward / GS_reverse (src/ntt.c:38-114).
//
// Shape.  Every thread holds E = N / T residues.  Phase p covers LR_p stages
// (radix R_p = 2^LR_p, R_p <= E) after S_p earlier stages; its smallest
// butterfly half-size is 2^LT with LT = LOGN - S_p - LR_p.  The N / R_p
// element groups of the phase are (b, o) with b < 2^S_p, o < 2^LT; group
// (b, o) holds the elements ((b R_p + k) << LT) + o, k < R_p, and uses
// twiddle (2^S_p + b) << st + grp at stage st of the phase -- the psi[m + i]
// of the reference loop.  Thread t owns groups gid = t + T g, g < E / R_p.
// Between phases the workgroup exchanges data through LDS (N words plus one
// pad word per 16).  The first forward phase (S = 0) reads its groups
// straight from HBM (element k of group gid is k 2^LT + gid, coalesced); the
// last one (LT = 0) holds R consecutive outputs per group and writes them as
// one contiguous run.  The inverse runs the phases in reverse order with the
// mirrored (Gentleman-Sande) butterflies, reading contiguous runs first and
// writing coalesced words last.
//
// Exactness: as in ntt_core.h, canonical inputs never produce the value q, and
// a workgroup whose inputs hold q runs the EXACT instantiation (twiddle flag
// bit 31 marks mod_prime(q w) == q).
#pragma once
#include "ntt_core.h"

namespace ipfe {

template <int... LR>
struct Rad {
    static constexpr int count = sizeof...(LR);
    static constexpr int total = (0 + ... + LR);
};

// Log radix and stages-before of phase P of a radix list.
template <typename RAD, int P>
struct PhaseOf;
template <int L0, int... LR, int P>
struct PhaseOf<Rad<L0, LR...>, P> {
    static constexpr int lr = PhaseOf<Rad<LR...>, P - 1>::lr;
    static constexpr int s = L0 + PhaseOf<Rad<LR...>, P - 1>::s;
};
template <int L0, int... LR>
struct PhaseOf<Rad<L0, LR...>, 0> {
    static constexpr int lr = L0;
    static constexpr int s = 0;
};

// Radix-2^LR Cooley-Tukey butterflies on registers (the general form of ct16).
template <int LR, bool EXACT>
__device__ __forceinline__ void ct_radix(uint32_t *x, const uint32_t *__restrict__ tw,
                                         const uint32_t *__restrict__ twp, uint32_t base, uint32_t q)
{
    constexpr int R = 1 << LR;
#pragma unroll
    for (int st = 0; st < LR; ++st) {
        const int half = R >> (st + 1);
#pragma unroll
        for (int grp = 0; grp < (1 << st); ++grp) {
            const uint32_t i = (base << st) + (uint32_t)grp;
            const uint32_t w = tw[i], wp = twp[i];
#pragma unroll
            for (int r = 0; r < half; ++r) ct_bfly<EXACT>(x[grp * 2 * half + r], x[grp * 2 * half + r + half], w, wp, q);
        }
    }
}

// Mirror image: Gentleman-Sande butterflies, stages in reverse order.
template <int LR, bool EXACT>
__device__ __forceinline__ void gs_radix(uint32_t *x, const uint32_t *__restrict__ tw,
                                         const uint32_t *__restrict__ twp, uint32_t base, uint32_t q)
{
    constexpr int R = 1 << LR;
#pragma unroll
    for (int st = LR - 1; st >= 0; --st) {
        const int half = R >> (st + 1);
#pragma unroll
        for (int grp = 0; grp < (1 << st); ++grp) {
            const uint32_t i = (base << st) + (uint32_t)grp;
            const uint32_t w = tw[i], wp = twp[i];
#pragma unroll
            for (int r = 0; r < half; ++r) gs_bfly<EXACT>(x[grp * 2 * half + r], x[grp * 2 * half + r + half], w, wp, q);
        }
    }
}

template <int LOGN, int T, typename RAD>
struct NttGen {
    static constexpr int N = 1 << LOGN;
    static constexpr int E = N / T;
    static constexpr int LDS_WORDS = N + N / 16;
    static constexpr int NP = RAD::count;
    static_assert(RAD::total == LOGN, "phases must cover every stage");
    static_assert(E * T == N && E >= 2, "threads per transform");

    // Index of register e of the first forward phase (and of the last inverse
    // phase): group g = e / R0 holds k = e % R0 at (k << LT0) + t + T g.
    __device__ static __forceinline__ int first_idx(int t, int e)
    {
        constexpr int LR0 = PhaseOf<RAD, 0>::lr, R0 = 1 << LR0, LT0 = LOGN - LR0;
        return ((e % R0) << LT0) + t + T * (e / R0);
    }
    // Index of register e of the last forward phase (and of the first inverse
    // phase): group g = e / RL holds the RL consecutive outputs of gid = t + T g.
    __device__ static __forceinline__ int last_idx(int t, int e)
    {
        constexpr int RL = 1 << PhaseOf<RAD, NP - 1>::lr;
        return (t + T * (e / RL)) * RL + (e % RL);
    }

    template <int P, bool EXACT>
    __device__ static __forceinline__ void fwd_phase(uint32_t (&x)[E], uint32_t *lds, const uint32_t *tw,
                                                     const uint32_t *twp, uint32_t q, int t)
    {
        constexpr int LR = PhaseOf<RAD, P>::lr, S = PhaseOf<RAD, P>::s, R = 1 << LR, G = E / R;
        constexpr int LT = LOGN - S - LR;
        static_assert(E % R == 0, "radix larger than the registers per thread");
        if constexpr (P > 0) {
#pragma unroll
            for (int g = 0; g < G; ++g) {
                const int gid = t + T * g, b = gid >> LT, o = gid & ((1 << LT) - 1);
#pragma unroll
                for (int k = 0; k < R; ++k) x[g * R + k] = lds[lds_idx((uint32_t)((((b * R) + k) << LT) + o))];
            }
        }
#pragma unroll
        for (int g = 0; g < G; ++g) {
            const int gid = t + T * g, b = gid >> LT;
            ct_radix<LR, EXACT>(x + g * R, tw, twp, (1u << S) + (uint32_t)b, q);
        }
        if constexpr (P < NP - 1) {
#pragma unroll
            for (int g = 0; g < G; ++g) {
                const int gid = t + T * g, b = gid >> LT, o = gid & ((1 << LT) - 1);
#pragma unroll
                for (int k = 0; k < R; ++k) lds[lds_idx((uint32_t)((((b * R) + k) << LT) + o))] = x[g * R + k];
            }
        }
    }

    template <int P, bool EXACT>
    __device__ static __forceinline__ void inv_phase(uint32_t (&x)[E], uint32_t *lds, const uint32_t *tw,
                                                     const uint32_t *twp, uint32_t q, int t)
    {
        constexpr int LR = PhaseOf<RAD, P>::lr, S = PhaseOf<RAD, P>::s, R = 1 << LR, G = E / R;
        constexpr int LT = LOGN - S - LR;
        if constexpr (P < NP - 1) {
#pragma unroll
            for (int g = 0; g < G; ++g) {
                const int gid = t + T * g, b = gid >> LT, o = gid & ((1 << LT) - 1);
#pragma unroll
                for (int k = 0; k < R; ++k) x[g * R + k] = lds[lds_idx((uint32_t)((((b * R) + k) << LT) + o))];
            }
        }
#pragma unroll
        for (int g = 0; g < G; ++g) {
            const int gid = t + T * g, b = gid >> LT;
            gs_radix<LR, EXACT>(x + g * R, tw, twp, (1u << S) + (uint32_t)b, q);
        }
        if constexpr (P > 0) {
#pragma unroll
            for (int g = 0; g < G; ++g) {
                const int gid = t + T * g, b = gid >> LT, o = gid & ((1 << LT) - 1);
#pragma unroll
                for (int k = 0; k < R; ++k) lds[lds_idx((uint32_t)((((b * R) + k) << LT) + o))] = x[g * R + k];
            }
        }
    }

    template <int P, bool EXACT>
    __device__ static __forceinline__ void fwd_from(uint32_t (&x)[E], uint32_t *lds, const uint32_t *tw,
                                                    const uint32_t *twp, uint32_t q, int t)
    {
        fwd_phase<P, EXACT>(x, lds, tw, twp, q, t);
        if constexpr (P + 1 < NP) {
            __syncthreads();
            fwd_from<P + 1, EXACT>(x, lds, tw, twp, q, t);
        }
    }

    template <int P, bool EXACT>
    __device__ static __forceinline__ void inv_from(uint32_t (&x)[E], uint32_t *lds, const uint32_t *tw,
                                                    const uint32_t *twp, uint32_t q, int t)
    {
        inv_phase<P, EXACT>(x, lds, tw, twp, q, t);
        if constexpr (P > 0) {
            __syncthreads();
            inv_from<P - 1, EXACT>(x, lds, tw, twp, q, t);
        }
    }

    // Forward transform.  In: x[e] = element first_idx(t, e).  Out: x[e] =
    // element last_idx(t, e) of the bit-reversed result.  Every thread of the
    // workgroup must call it (barriers); lds is free again on return.
    template <bool EXACT>
    __device__ static __forceinline__ void fwd(uint32_t (&x)[E], uint32_t *lds, const uint32_t *tw,
                                               const uint32_t *twp, uint32_t q, int t)
    {
        fwd_from<0, EXACT>(x, lds, tw, twp, q, t);
        __syncthreads();
    }

    // Inverse transform without the final N^-1.  In: x[e] = element
    // last_idx(t, e).  Out: x[e] = element first_idx(t, e).
    template <bool EXACT>
    __device__ static __forceinline__ void inv(uint32_t (&x)[E], uint32_t *lds, const uint32_t *tw,
                                               const uint32_t *twp, uint32_t q, int t)
    {
        inv_from<NP - 1, EXACT>(x, lds, tw, twp, q, t);
        __syncthreads();
    }
};

}  // namespace ipfe
