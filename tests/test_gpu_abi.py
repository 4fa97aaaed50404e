"""The four reference drop-in symbols at the reference's full size (L = 785,
layouts [786][3][4096] / [785][3][4096]) and the plain svm_predict_gpu caller,
against the CPU oracle."""
import ctypes as C
import hashlib

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

N, L = 4096, 785


def h(t: bytes) -> bytes:
    return hashlib.sha256(t).digest()


def seed(t):
    return np.frombuffer(h(t), dtype=np.uint8).copy()


def note(msg):
    print(f"[gpu-abi] {msg}", flush=True)


@pytest.fixture(scope="module")
def full_keys(gpu, po):
    note("rlwe_sife_setup_gpu L=785")
    a = po.sample_polya(h(b"abiA"))
    mpk = np.zeros((L + 1, 3, N), np.uint32)
    mpk[L] = a
    msk = np.zeros((L, 3, N), np.uint32)
    t = C.c_float(0.0)
    gpu.lib().rlwe_sife_setup_gpu(mpk, msk, seed(b"abiS"), seed(b"abiE"), C.byref(t))
    assert t.value > 0
    return a, mpk, msk


def test_setup_full_size(po, full_keys):
    a, mpk, msk = full_keys
    mpk_o, msk_o = po.setup_from_a(a, h(b"abiS"), h(b"abiE"), L)
    assert np.array_equal(msk, msk_o)
    assert np.array_equal(mpk, mpk_o)


def test_encrypt_keygen_decrypt_full_size(gpu, po, full_keys):
    a, mpk, msk = full_keys
    rng = np.random.default_rng(61)
    m = np.zeros((2, L), np.uint32)
    m[:, 1:124] = rng.integers(0, 5, (2, 123))
    c = np.empty((2, L + 1, 3, N), np.uint32)
    note("rlwe_sife_encrypt_gpu repeat=2")
    gpu.lib().rlwe_sife_encrypt_gpu(m, mpk, c, seed(b"abi2"), seed(b"abi3"), 2, None)
    for b in range(2):
        assert np.array_equal(c[b], po.encrypt(m[b], mpk, h(b"abi2"), h(b"abi3"), b))
    y = np.zeros((3, L), np.uint32)
    y[:, 1:124] = rng.integers(0, 17, (3, 123)) * (rng.random((3, 123)) < 0.12)
    y[2, 784] = 3  # a non-zero last slot: the fold may end on the representative q
    sk = np.empty((3, 3, N), np.uint32)
    note("rlwe_sife_keygen_gpu repeat=3")
    gpu.lib().rlwe_sife_keygen_gpu(y, msk, sk, 3, None)
    for v in range(3):
        assert np.array_equal(sk[v], po.keygen(y[v], msk))
    d = np.empty((3, 3, N), np.uint32)
    note("rlwe_sife_decrypt_gmp_gpu repeat=3")
    gpu.lib().rlwe_sife_decrypt_gmp_gpu(np.ascontiguousarray(c[1]), y, sk, d, 3, None)
    for v in range(3):
        assert np.array_equal(d[v], po.decrypt(c[1], y[v], sk[v]))
        k = gpu.round_extract2(d[v])
        assert k == po.orc().orc_round_extract_coef0(np.ascontiguousarray(d[v]))
        assert int(k) % 50241 == int(m[1].astype(np.int64) @ y[v].astype(np.int64)) % 50241


def test_svm_predict_plain(gpu, po, full_keys):
    """cufe_svm_predict_gpu (svm_predict_gpu, main.c:270-390) vs the oracle's
    svm_predict_cpu (main.c:137-267) on one ciphertext and a 6-SV model."""
    a, mpk, msk = full_keys
    rng = np.random.default_rng(62)
    x = np.zeros(L, np.uint32)
    x[1:124] = rng.integers(0, 2, 123)
    c = np.empty((1, L + 1, 3, N), np.uint32)
    gpu.lib().rlwe_sife_encrypt_gpu(x[None, :].copy(), mpk, c, seed(b"pl2"), seed(b"pl3"), 1, None)
    S = 6
    sv = np.zeros((S, L), np.uint32)
    for v in range(S):
        sv[v, rng.choice(np.arange(1, 124), 14, replace=False)] = 1
    coef = np.array([[0.75, -1.0, 0.5, -0.25, 1.0, -1.0]])
    model = gpu.SvmModel(2, [3, 3], [1, -1], coef, [0.125], sv)
    note("svm_predict plain")
    lab, kv = gpu.svm_predict_plain(model, c[0], msk)
    kv_o = np.empty(S)
    lab_o = po.orc().orc_svm_predict_cpu(np.ascontiguousarray(c[0]), L, msk, 2, S, model.nSV, model.label,
                                         np.ascontiguousarray(coef.ravel()), model.rho, sv, kv_o, 0)
    assert np.array_equal(kv, kv_o)
    assert lab == lab_o
    assert np.array_equal(kv.astype(np.int64) % 50241, (sv.astype(np.int64) @ x.astype(np.int64)) % 50241)
