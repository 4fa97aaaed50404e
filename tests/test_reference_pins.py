"""The oracle against the reference's own GMP code path (oracle/_ref, compiled
from src/sample.c, src/crt.c, src/rlwe_sife.c and main.c where they lie, with
GMP from the image): sample_polya, crt_reverse_gmp + round_extract_gmp,
round_extract_gmp2, rlwe_sife_keygen, rlwe_sife_decrypt_gmp and the whole
svm_predict_cpu (main.c:137-267), including a three-class vote.  Skipped where
the reference build is absent."""
import ctypes
import hashlib
import os
import sys

import numpy as np
import pytest

from test_oracle import VOTE3, vote3_by_hand

HERE = os.path.dirname(os.path.abspath(__file__))
N, L = 4096, 785
Q = (16760833, 2147352577, 2130706433)
QQ = Q[0] * Q[1] * Q[2]
P = 50241


def h(t: bytes) -> bytes:
    return hashlib.sha256(t).digest()


def gaussian_msk(rng, l, width=900):
    z = rng.integers(-width, width + 1, (l, N))
    msk = np.empty((l, 3, N), np.uint32)
    for j in range(3):
        msk[:, j, :] = np.where(z < 0, Q[j] + z, z).astype(np.uint32)
    return msk


def test_sample_polya_vs_reference(po, ref):
    for tag in (b"a", b"polya-2", b"smoke-A"):
        want = np.empty((3, N), np.uint32)
        ref.sample_polya(po.seed_bytes(h(tag)), want)
        assert np.array_equal(po.sample_polya(h(tag)), want), tag


def test_crt_round_vs_reference(po, ref):
    """crt_reverse_gmp + round_extract_gmp on residues at the edges (0, 1, q-1,
    the representative q, values near Q/p multiples) and at random."""
    rng = np.random.default_rng(3)
    d = np.empty((3, N), np.uint32)
    for j in range(3):
        d[j] = rng.integers(0, Q[j], N)
    edges = [0, 1, Q[0] - 1, Q[0]]
    for k, e in enumerate(edges):  # limb 0 at its edges, limbs 1-2 random
        d[0, k] = e
    for j in (1, 2):
        d[j, 4:8] = [0, 1, Q[j] - 1, Q[j]]
    # residues of x = m floor(Q/p) + noise, and of values just below Q
    scale = QQ // P
    for k, x in enumerate([5 * scale, 5 * scale + 1000, 5 * scale - 1000, QQ - 1, QQ - 2000, 0, scale // 2,
                           scale // 2 + 1]):
        for j in range(3):
            d[j, 8 + k] = x % Q[j]
    x_ref = np.empty((N, 3), np.uint32)
    r_ref = np.empty(N, np.uint32)
    ref.refh_crt_round(np.ascontiguousarray(d), x_ref, r_ref)
    for k in range(N):
        xv = int(x_ref[k, 0]) | (int(x_ref[k, 1]) << 32) | (int(x_ref[k, 2]) << 64)
        assert po.crt_value(int(d[0, k]), int(d[1, k]), int(d[2, k])) == xv, k
        assert po.crt_round(int(d[0, k]), int(d[1, k]), int(d[2, k])) == int(r_ref[k]), k
    assert po.orc().orc_round_extract_coef0(np.ascontiguousarray(d)) == ref.round_extract_gmp2(np.ascontiguousarray(d))


def test_keygen_vs_reference(po, ref):
    """rlwe_sife_keygen over the 785 slots, with a key whose last slot is
    non-zero (its fold can end on the representative q)."""
    rng = np.random.default_rng(4)
    msk = gaussian_msk(rng, L)
    for trial in range(3):
        y = np.zeros(L, np.uint32)
        idx = rng.choice(L, 20, replace=False)
        y[idx] = rng.integers(1, 17, 20)
        if trial == 2:
            y[L - 1] = 16
        want = np.empty((3, N), np.uint32)
        t = ctypes.c_double(0.0)
        ref.rlwe_sife_keygen(y, msk, want, ctypes.byref(t))
        assert np.array_equal(po.keygen(y, msk), want), trial


def test_decrypt_round_vs_reference(po, ref):
    """rlwe_sife_decrypt_gmp + round_extract_gmp on every coefficient equals the
    oracle's RNS decryption followed by its CRT and rounding."""
    rng = np.random.default_rng(5)
    msk = gaussian_msk(rng, L)
    c = np.empty((L + 1, 3, N), np.uint32)
    for j in range(3):
        c[:, j, :] = rng.integers(0, Q[j], (L + 1, N))
    c[3, 1, 0] = Q[1]
    y = np.zeros(L, np.uint32)
    y[1:124:7] = rng.integers(0, 17, len(range(1, 124, 7)))
    sk = po.keygen(y, msk)
    r_ref = np.empty(N, np.uint32)
    ref.refh_decrypt_round(c, y, sk, r_ref)
    d = po.decrypt(c, y, sk)
    got = np.array([po.crt_round(int(d[0, k]), int(d[1, k]), int(d[2, k])) for k in range(N)], np.uint32)
    assert np.array_equal(got, r_ref)


def noise_free_ciphertext(x):
    """c_i = floor(Q/p) x_i mod q_j in every coefficient, c_L = 0: decrypts to
    <x, y> exactly, so the key values of a prediction are set by x."""
    c = np.zeros((L + 1, 3, N), np.uint32)
    for j in range(3):
        s = (QQ // P) % Q[j]
        for i, v in enumerate(x):
            c[i, j, :] = (s * int(v)) % Q[j]
    return c


def test_svm_predict_cpu_three_classes_vs_reference(po, ref):
    """The reference's svm_predict_cpu, the oracle's restatement and the
    hand-computed vote agree on the three-class model with differing
    coefficient rows (key values 1..5 and two other patterns)."""
    rng = np.random.default_rng(6)
    msk = gaussian_msk(rng, L)
    sv = np.zeros((5, L), np.uint32)
    for v in range(5):
        sv[v, v + 1] = 1
    nSV = np.array(VOTE3["nSV"], np.int32)
    label = np.array(VOTE3["label"], np.int32)
    coef = np.ascontiguousarray(np.array(VOTE3["coef"], np.float64).ravel())
    rho = np.array(VOTE3["rho"], np.float64)
    for kv, want in (([1, 2, 3, 4, 5], 10), ([1, 1, 4, 1, 1], 20), ([4, 1, 1, 1, 1], 30)):
        x = np.zeros(L, np.uint32)
        x[1:6] = kv
        c = noise_free_ciphertext(x)
        lab_ref = ref.refh_svm_predict_cpu(c, msk, 3, 5, nSV, label, coef, rho, sv)
        kvo = np.empty(5)
        lab_orc = po.orc().orc_svm_predict_cpu(c, L, msk, 3, 5, nSV, label, coef, rho, sv, kvo, 0)
        assert lab_ref == lab_orc == want == vote3_by_hand(kv), kv
        assert kvo.tolist() == kv


def test_svm_predict_cpu_real_coefficients_vs_reference(po, ref):
    """A two-class model with real-valued coefficients taken from the a1a model
    (its first 12 support vectors), on three a1a test rows."""
    from cufe_amd import svm_io

    model = svm_io.read_model(os.path.join(HERE, "golden", "a1a.txt.model"))
    _, xd = svm_io.read_problem(os.path.join(HERE, "golden", "a1a_rows101_400.t"), limit=3)
    S = 12
    nSV = np.array([6, 6], np.int32)
    idx = list(range(6)) + list(range(model.nSV[0], model.nSV[0] + 6))
    sv = np.ascontiguousarray(model.sv[idx])
    coef = np.ascontiguousarray(model.sv_coef[:, idx].ravel())
    label = np.array(model.label, np.int32)
    rho = np.ascontiguousarray(model.rho, np.float64)
    msk = gaussian_msk(np.random.default_rng(7), L)
    for r in range(3):
        c = noise_free_ciphertext(xd[r])
        lab_ref = ref.refh_svm_predict_cpu(c, msk, 2, S, nSV, label, coef, rho, sv)
        kvo = np.empty(S)
        lab_orc = po.orc().orc_svm_predict_cpu(c, L, msk, 2, S, nSV, label, coef, rho, sv, kvo, 0)
        assert lab_ref == lab_orc, r
        assert np.array_equal(kvo, (xd[r].astype(np.int64) @ sv.astype(np.int64).T).astype(np.float64))
