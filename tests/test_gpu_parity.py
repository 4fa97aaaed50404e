"""Bit-exact parity of the HIP path (libcufe_ipfe.so on the MI355X) with the CPU
oracle on identical seeds and inputs.  Integer work: every word must match."""
import ctypes
import hashlib
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

N = 4096
Q = (16760833, 2147352577, 2130706433)


def h(t: bytes) -> bytes:
    return hashlib.sha256(t).digest()


def note(msg):
    print(f"[gpu-parity] {msg}", flush=True)
    sys.stdout.flush()


def rand_polys(rng, npoly, with_q=False):
    a = np.stack([rng.integers(0, q + (1 if with_q else 0), (npoly, N), dtype=np.uint32) for q in Q], axis=1)
    return np.ascontiguousarray(a)


def test_device_is_gfx950(gpu):
    info = gpu.device_info()
    note(str(info))
    assert "gfx950" in info["name"]


def test_aes_ctr_keystream(gpu, po):
    note("aes ctr")
    key = h(b"aes")
    for nonce, ctr0, nb in ((0, 0, 64), (0xABCD, 8, 40), (7, 2**32 - 5, 10)):
        got = gpu.aes256ctr(key, nonce, ctr0, nb)
        st = po.stream(key, nonce)
        st.ctr = ctr0
        out = np.empty(128 * ((nb + 7) // 8), dtype=np.uint8)
        po.orc().orc_stream_squeeze(ctypes.byref(st), out, (nb + 7) // 8)
        assert got == out.tobytes()[: 16 * nb]


def test_sample_polya(gpu, po):
    note("sample_polya")
    for t in (b"p1", b"p2"):
        assert np.array_equal(gpu.sample_polya(h(t)), po.sample_polya(h(t)))


@pytest.mark.parametrize("inverse", [False, True])
def test_ntt_random(gpu, po, inverse):
    note(f"ntt inverse={inverse}")
    rng = np.random.default_rng(1 + inverse)
    x = rand_polys(rng, 6)
    got = (gpu.ntt_inverse if inverse else gpu.ntt_forward)(x)
    want = (po.ntt_inverse if inverse else po.ntt_forward)(x)
    assert np.array_equal(got, want)


@pytest.mark.parametrize("inverse", [False, True])
def test_ntt_with_q_representatives(gpu, po, inverse):
    """Inputs holding the value q (which mod_prime can return) take the exact path."""
    note(f"ntt q-values inverse={inverse}")
    rng = np.random.default_rng(3 + inverse)
    x = rand_polys(rng, 4)
    for p in range(4):
        for s in range(3):
            idx = rng.integers(0, N, 40)
            x[p, s, idx] = Q[s]
    x[3, :, :] = np.array(Q, dtype=np.uint32)[:, None]  # all q
    got = (gpu.ntt_inverse if inverse else gpu.ntt_forward)(x)
    want = (po.ntt_inverse if inverse else po.ntt_forward)(x)
    assert np.array_equal(got, want)


def test_ntt_roundtrip_large_batch(gpu):
    note("ntt roundtrip 512 polys")
    rng = np.random.default_rng(5)
    x = rand_polys(rng, 512)
    assert np.array_equal(gpu.ntt_inverse(gpu.ntt_forward(x)), x)


@pytest.mark.parametrize("which", [1, 2, 3])
def test_gauss_single_stream_multi_call(gpu, po, which):
    note(f"gauss S{which} 3 calls")
    key = h(b"g%d" % which)
    got = gpu.gauss_sample([key], [11], which, 3)[0]
    want = po.gauss_calls(key, 11, which, 3)
    assert np.array_equal(got, want)


@pytest.mark.parametrize("which", [1, 2, 3])
def test_gauss_many_streams(gpu, po, which):
    note(f"gauss S{which} 24 streams")
    keys = [h(b"m%d-%d" % (which, t)) for t in range(24)]
    nonces = list(range(100, 124))
    got = gpu.gauss_sample(keys, nonces, which, 1)
    for t in (0, 7, 23):
        assert np.array_equal(got[t, 0], po.gauss_calls(keys[t], nonces[t], which, 1)[0])


def test_gauss_golden(gpu, golden):
    note("gauss golden")
    for which in (1, 2, 3):
        g = golden["gauss"][str(which)]
        out = gpu.gauss_sample([bytes.fromhex(g["key"])], [g["nonce"]], which, 4)[0]
        assert [hashlib.sha256(np.ascontiguousarray(c).tobytes()).hexdigest() for c in out] == g["sha256_calls"]


def test_setup_small(gpu, po):
    note("setup l=7")
    l = 7
    a = po.sample_polya(h(b"A"))
    mpk, msk = gpu.setup(a, h(b"S"), h(b"E"), l)
    mpk_o, msk_o = po.setup_from_a(a, h(b"S"), h(b"E"), l)
    assert np.array_equal(msk, msk_o)
    assert np.array_equal(mpk, mpk_o)


def test_scheme_golden(gpu, golden):
    note("scheme golden l=3")
    g = golden["scheme"]
    l = g["l"]
    b = bytes.fromhex
    sha = lambda a: hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()
    a = gpu.sample_polya(b(g["seedA"]))
    assert sha(a) == g["a_sha256"]
    mpk, msk = gpu.setup(a, b(g["seedS"]), b(g["seedE"]), l)
    assert sha(mpk) == g["mpk_sha256"] and sha(msk) == g["msk_sha256"]
    c = gpu.encrypt(np.array(g["x"], np.uint32), mpk, b(g["seed2"]), b(g["seed3"]))
    assert [sha(ci) for ci in c] == g["c_sha256"]
    y = np.array(g["y"], np.uint32)
    sk = gpu.keygen(y, msk)
    assert [sha(s) for s in sk] == g["sk_sha256"]
    d = gpu.decrypt(c[0], y, sk)
    assert [sha(di) for di in d] == g["d_sha256"]


def test_encrypt_keygen_decrypt_small(gpu, po):
    note("encrypt/keygen/decrypt l=16")
    l = 16
    a = po.sample_polya(h(b"A2"))
    mpk, msk = po.setup_from_a(a, h(b"S2"), h(b"E2"), l)
    rng = np.random.default_rng(21)
    x = rng.integers(0, 5, (3, l)).astype(np.uint32)
    c = gpu.encrypt(x, mpk, h(b"s2"), h(b"s3"), nonce0=40)
    for bb in range(3):
        assert np.array_equal(c[bb], po.encrypt(x[bb], mpk, h(b"s2"), h(b"s3"), 40 + bb))
    y = rng.integers(0, 17, (5, l)).astype(np.uint32)
    y[1] = 0
    y[2, -1] = 0
    y[3, 0] = 2**32 - 1  # reduced per limb as the reference does
    sk = gpu.keygen(y, msk)
    for v in range(5):
        assert np.array_equal(sk[v], po.keygen(y[v], msk))
    d = gpu.decrypt(c[1], y, sk)
    for v in range(5):
        assert np.array_equal(d[v], po.decrypt(c[1], y[v], sk[v]))
    r = gpu.crt_round(d)
    for v in range(5):
        assert int(r[v]) == po.crt_round(int(d[v, 0, 0]), int(d[v, 1, 0]), int(d[v, 2, 0]))
    for v in (0, 4):
        assert int(r[v]) % 50241 == int(np.dot(x[1].astype(np.int64), y[v].astype(np.int64))) % 50241


def test_crt_round_edges(gpu, po):
    note("crt round edges")
    d = np.zeros((6, 3, N), dtype=np.uint32)
    d[1, :, 0] = Q  # value q in every limb
    d[2, 0, 0] = Q[0]
    d[2, 1, 0] = 5
    QQ = Q[0] * Q[1] * Q[2]
    d[3, :, 0] = [(QQ - 1) % q for q in Q]
    d[4, :, 0] = [(QQ // 2) % q for q in Q]
    d[5, :, 0] = [(QQ // 2 + 1) % q for q in Q]
    r = gpu.crt_round(d)
    for v in range(6):
        assert int(r[v]) == po.crt_round(int(d[v, 0, 0]), int(d[v, 1, 0]), int(d[v, 2, 0]))
