"""Writes tests/golden/ref_vectors.json from the reference's own compiled code.

Run in the container that has the reference mounted (oracle/_ref/libref.so is
built from /root/reference/src by oracle/Makefile).  The fixtures are data only:
inputs, outputs and SHA-256 digests of outputs.

Sources of each vector:
  * AES keystream, Gaussian samplers, mod_prime, NTT, twiddle tables: the
    reference functions aes256ctr_*, gaussian_sampler_S1_ori / S2_ori / S3,
    mod_prime, CT_forward, GS_reverse and the psi / psi_inv tables.
  * setup / encrypt / keygen / decrypt at a small vector length: the call
    sequence of src/rlwe_sife.c (setup :32-82, encrypt :112-181, keygen
    :215-245, decrypt :248-280) driven from Python over those same reference
    leaves (point_mul, poly_add_mod, poly_sub_mod, poly_mul_ntt, mod_prime).
    sample_polya needs GMP and is not built; its closed form (LE64 of the
    nonce-0 keystream mod q_j, every draw kept because 2^64 < Q) is applied to
    the reference keystream.  Seeds follow the convention of the library
    (secret stream = seedS nonce 0, error = seedE nonce 0; encryption message b
    uses nonce b on seed2 and seed3).
"""
from __future__ import annotations

import ctypes as C
import hashlib
import json
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, os.path.join(ROOT, "oracle"))
import pyoracle as po  # noqa: E402

N = 4096
Q = (16760833, 2147352577, 2130706433)


def sha(a) -> str:
    return hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()


def key_of(tag: str) -> bytes:
    return hashlib.sha256(tag.encode()).digest()


def ref_sample_polya(seed: bytes) -> np.ndarray:
    ks = po.RefStream(seed, 0).squeeze(352)
    v = np.frombuffer(ks[: 8 * N], dtype="<u8")
    return np.stack([(v % np.uint64(q)).astype(np.uint32) for q in Q])


def ref_ct(a: np.ndarray, s: int) -> np.ndarray:
    a = np.ascontiguousarray(a, dtype=np.uint32).copy()
    po.ref().CT_forward(a, s)
    return a


def ref_gs(a: np.ndarray, s: int) -> np.ndarray:
    a = np.ascontiguousarray(a, dtype=np.uint32).copy()
    po.ref().GS_reverse(a, s)
    return a


def ref_binop(fn, a, b, s) -> np.ndarray:
    out = np.empty(N, dtype=np.uint32)
    fn(np.ascontiguousarray(a, dtype=np.uint32), np.ascontiguousarray(b, dtype=np.uint32), out, s)
    return out


def ref_mod_prime_vec(m: np.ndarray, s: int) -> np.ndarray:
    mp = po.ref().mod_prime
    return np.array([mp(int(v), s) for v in m.tolist()], dtype=np.uint32)


def ref_setup(seedA, seedS, seedE, l):
    r = po.ref()
    mpk = np.zeros((l + 1, 3, N), dtype=np.uint32)
    msk = np.zeros((l, 3, N), dtype=np.uint32)
    a = ref_sample_polya(seedA)
    for s in range(3):
        mpk[l, s] = ref_ct(a[s], s)
    S, E = po.RefStream(seedS, 0), po.RefStream(seedE, 0)
    for i in range(l):
        msk[i] = S.gauss(1)
        e = E.gauss(1)
        for s in range(3):
            sh = ref_ct(msk[i, s], s)
            eh = ref_ct(e[s], s)
            t = ref_binop(r.point_mul, mpk[l, s], sh, s)
            mpk[i, s] = ref_binop(r.poly_add_mod, t, eh, s)
    return a, mpk, msk


def ref_encrypt(x, mpk, seed2, seed3, nonce):
    r = po.ref()
    l = len(x)
    scale = [int((Q[0] * Q[1] * Q[2]) // 50241 % q) for q in Q]
    mcrt = np.zeros((3, l), dtype=np.uint64)
    for i in range(l):
        for s in range(3):
            mj = r.mod_prime(int(x[i]), s)
            mcrt[s, i] = r.mod_prime(mj * scale[s], s)
    S2, S3 = po.RefStream(seed2, nonce), po.RefStream(seed3, nonce)
    rr = S2.gauss(2)
    f0 = S2.gauss(2)
    c = np.zeros((l + 1, 3, N), dtype=np.uint32)
    rh = np.stack([ref_ct(rr[s], s) for s in range(3)])
    for s in range(3):
        t = ref_gs(ref_binop(r.point_mul, mpk[l, s], rh[s], s), s)
        c[l, s] = ref_binop(r.poly_add_mod, t, f0[s], s)
    for i in range(l):
        f = S3.gauss(3)
        for s in range(3):
            t = ref_gs(ref_binop(r.point_mul, mpk[i, s], rh[s], s), s)
            t = ref_binop(r.poly_add_mod, t, f[s], s)
            c[i, s] = ref_mod_prime_vec(t.astype(np.uint64) + np.uint64(mcrt[s, i]), s)
    return c


def ref_keygen(y, msk):
    r = po.ref()
    sk = np.zeros((3, N), dtype=np.uint32)
    for i in range(len(y)):
        for s in range(3):
            ys = r.mod_prime(int(y[i]), s)
            sk[s] = ref_mod_prime_vec(np.uint64(ys) * msk[i, s].astype(np.uint64) + sk[s].astype(np.uint64), s)
    return sk


def ref_decrypt(c, y, sk):
    r = po.ref()
    l = len(y)
    d = np.zeros((3, N), dtype=np.uint32)
    for i in range(l):
        for s in range(3):
            ys = r.mod_prime(int(y[i]), s)
            d[s] = ref_mod_prime_vec(np.uint64(ys) * c[i, s].astype(np.uint64) + d[s].astype(np.uint64), s)
    for s in range(3):
        prod = ref_binop(r.poly_mul_ntt, c[l, s], sk[s], s)
        d[s] = ref_binop(r.poly_sub_mod, d[s], prod, s)
    return d


def main() -> None:
    if po.ref() is None:
        raise SystemExit("oracle/_ref/libref.so is not built (make -C oracle ref)")
    out: dict = {"provenance": "reference sources aes256ctr.c arith_rns.c gauss.c ntt.c compiled with -DAVX2"}

    # AES-256-CTR keystream
    key = bytes(range(32))
    st = po.RefStream(key, 0x1234)
    ks = st.squeeze(64)
    out["aes"] = {"key": key.hex(), "nonce": 0x1234, "first384": ks[:384].hex(),
                  "sha256_8192": hashlib.sha256(ks).hexdigest()}

    # Gaussian samplers: 4 consecutive calls on one stream
    out["gauss"] = {}
    for which in (1, 2, 3):
        k = key_of(f"cufe-golden-S{which}")
        rs = po.RefStream(k, 7)
        calls = [rs.gauss(which) for _ in range(4)]
        out["gauss"][str(which)] = {
            "key": k.hex(), "nonce": 7,
            "sha256_calls": [sha(cc) for cc in calls],
            "call0_head": [cc.tolist() for cc in calls[0][:, :16]],
            "call3_tail": [cc.tolist() for cc in calls[3][:, -16:]],
        }

    # mod_prime: random inputs and inputs that land on 2q
    rng = np.random.default_rng(20261202)
    ms = [int(v) for v in rng.integers(0, 2**63, 1500, dtype=np.uint64)]
    ms += [int(v) << 1 for v in rng.integers(0, 2**62, 500, dtype=np.uint64)]
    ms += [1023 * 2**24 + 16761857, 2 * Q[0], 2 * Q[1], 2 * Q[2], Q[0], Q[1], Q[2], 0, 2**64 - 1]
    ms += [k * q for q in Q for k in range(1, 40)]
    out["mod_prime"] = {"m": [str(m) for m in ms],
                        "r": [[po.ref().mod_prime(m, s) for m in ms] for s in range(3)]}

    # NTT and the twiddle tables
    psi = np.ctypeslib.as_array((C.c_uint32 * (3 * N)).in_dll(po.ref(), "psi")).reshape(3, N)
    psi_inv = np.ctypeslib.as_array((C.c_uint32 * (3 * N)).in_dll(po.ref(), "psi_inv")).reshape(3, N)
    out["tables"] = {"psi_sha256": sha(psi), "psi_inv_sha256": sha(psi_inv),
                     "psi_head": psi[:, :8].tolist(), "psi_inv_head": psi_inv[:, :8].tolist()}
    ntt = {}
    for s in range(3):
        x = np.random.default_rng(100 + s).integers(0, Q[s], N, dtype=np.uint32)
        fx = ref_ct(x, s)
        ix = ref_gs(x, s)
        ntt[str(s)] = {"input_seed": 100 + s, "fwd_sha256": sha(fx), "inv_sha256": sha(ix),
                       "fwd_head": fx[:8].tolist(), "inv_head": ix[:8].tolist()}
    out["ntt"] = ntt

    # scheme at l = 3, two messages, two key vectors
    l = 3
    seedA, seedS, seedE = key_of("gA"), key_of("gS"), key_of("gE")
    seed2, seed3 = key_of("g2"), key_of("g3")
    a, mpk, msk = ref_setup(seedA, seedS, seedE, l)
    xs = [np.array([1, 4, 0], dtype=np.uint32), np.array([3, 2, 2], dtype=np.uint32)]
    cs = [ref_encrypt(x, mpk, seed2, seed3, b) for b, x in enumerate(xs)]
    ys = [np.array([5, 0, 16], dtype=np.uint32), np.array([0, 7, 1], dtype=np.uint32)]
    sks = [ref_keygen(y, msk) for y in ys]
    ds = [ref_decrypt(cs[0], y, sk) for y, sk in zip(ys, sks)]
    out["scheme"] = {
        "l": l, "seedA": seedA.hex(), "seedS": seedS.hex(), "seedE": seedE.hex(),
        "seed2": seed2.hex(), "seed3": seed3.hex(),
        "a_sha256": sha(a), "mpk_sha256": sha(mpk), "msk_sha256": sha(msk),
        "x": [x.tolist() for x in xs], "c_sha256": [sha(c) for c in cs],
        "c_head": [c[:, :, :4].tolist() for c in cs],
        "y": [y.tolist() for y in ys], "sk_sha256": [sha(sk) for sk in sks],
        "d_sha256": [sha(d) for d in ds],
        "d_coef0": [d[:, 0].tolist() for d in ds],
    }
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "ref_vectors.json")
    with open(path, "w") as f:
        json.dump(out, f, indent=1)
    print("wrote", path)


if __name__ == "__main__":
    main()
