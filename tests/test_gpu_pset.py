"""The library's three parameter sets on the GPU against the oracle's
restatement (oracle/pset_oracle.c): every NTT shape of the autotuner and the
tuned one, Gaussian sampling, sample_polya, setup / encrypt / keygen / decrypt
and CRT + rounding -- word for word.  At set 1 (the reference's set) the
generic path also equals the library's specialised reference-set path.  Sets
0 and 2 have no reference: the oracle is pinned by mathematics
(test_pset_oracle.py)."""
import hashlib

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def h(t: bytes) -> bytes:
    return hashlib.sha256(t).digest()


def note(msg):
    print(f"[gpu-pset] {msg}", flush=True)


def random_polys(rng, ps, n, with_q=True):
    a = np.stack([np.stack([rng.integers(0, q, ps.N) for q in ps.q]) for _ in range(n)]).astype(np.uint32)
    if with_q:  # the value q itself on every limb (mod_prime's representative of 0)
        for j, q in enumerate(ps.q):
            a[0, j, [0, 5, ps.N - 1]] = q
    return a


@pytest.mark.parametrize("pid", [0, 1, 2])
def test_every_ntt_shape_matches_oracle(gpu, po, pid):
    ps = gpu.PSet(pid)
    rng = np.random.default_rng(100 + pid)
    a = random_polys(rng, ps, 3)
    fwd = po.ps_ntt_forward(pid, a)
    inv = po.ps_ntt_inverse(pid, a)
    names = ps.ntt_variants()
    note(f"set {pid}: N={ps.N} limbs={ps.nmod} shapes {names}")
    for v, name in enumerate(names):
        if name == "-":
            continue
        assert np.array_equal(ps.ntt(a, False, v), fwd), (pid, name, "fwd")
        assert np.array_equal(ps.ntt(a, True, v), inv), (pid, name, "inv")
    assert np.array_equal(ps.ntt(a), fwd) and np.array_equal(ps.ntt(a, True), inv)
    rep = gpu.ntt_autotune_report()
    assert f"N={ps.N} limbs={ps.nmod} fwd" in rep and f"N={ps.N} limbs={ps.nmod} inv" in rep
    assert "MISMATCH" not in rep


@pytest.mark.parametrize("pid", [0, 2])
@pytest.mark.parametrize("which", [1, 2, 3])
def test_gauss_matches_oracle(gpu, po, pid, which):
    ps = gpu.PSet(pid)
    keys = [h(b"k%d-%d" % (pid, t)) for t in range(3)]
    nonces = [0, 7, 65535]
    got = ps.gauss_sample(keys, nonces, which, 2)
    for t in range(3):
        assert np.array_equal(got[t], po.ps_gauss_calls(pid, keys[t], nonces[t], which, 2)), t


@pytest.mark.parametrize("pid", [0, 2])
def test_sample_polya_matches_oracle(gpu, po, pid):
    ps = gpu.PSet(pid)
    for tag in (b"pa1", b"pa2"):
        assert np.array_equal(ps.sample_polya(h(tag)), po.ps_sample_polya(pid, h(tag)))


def scheme_round_trip(gpu, po, pid, l, nmsg):
    ps = gpu.PSet(pid)
    rng = np.random.default_rng(200 + pid + l)
    a = po.ps_sample_polya(pid, h(b"sA%d" % l))
    note(f"set {pid}: setup l={l}")
    mpk, msk = ps.setup(a, h(b"sS"), h(b"sE"), l)
    mpk_o, msk_o = po.ps_setup_from_a(pid, a, h(b"sS"), h(b"sE"), l)
    assert np.array_equal(mpk, mpk_o) and np.array_equal(msk, msk_o)
    x = rng.integers(0, 5, (nmsg, l)).astype(np.uint32)
    note(f"set {pid}: encrypt {nmsg} x l={l}")
    c = ps.encrypt(x, mpk, h(b"s2"), h(b"s3"), nonce0=3)
    for b in range(nmsg):
        assert np.array_equal(c[b], po.ps_encrypt(pid, x[b], mpk, h(b"s2"), h(b"s3"), 3 + b)), b
    y = rng.integers(0, 17, (2, l)).astype(np.uint32)
    y[1, l - 1] = 16  # a key whose fold ends on a non-zero slot
    sk = ps.keygen(y, msk)
    for v in range(2):
        assert np.array_equal(sk[v], po.ps_keygen(pid, y[v], msk)), v
    d = ps.decrypt(c[0], y, sk)
    r = ps.crt_round(d, coef=-1)
    for v in range(2):
        dv = po.ps_decrypt(pid, c[0], y[v], sk[v])
        assert np.array_equal(d[v], dv), v
        ip = int(x[0].astype(np.int64) @ y[v].astype(np.int64))
        for k in (0, 1, ps.N // 2, ps.N - 1):
            assert int(r[v, k]) == po.ps_crt_round(pid, dv[:, k]), (v, k)
        assert np.all(r[v].astype(np.int64) % ps.p == ip % ps.p)
    assert np.array_equal(ps.crt_round(d, coef=0), r[:, 0])
    return ps, mpk, msk, c


@pytest.mark.parametrize("pid", [0, 1, 2])
def test_scheme_matches_oracle(gpu, po, pid):
    scheme_round_trip(gpu, po, pid, 5, 2)


@pytest.mark.parametrize("pid", [0, 2])
def test_single_operations_at_l123(gpu, po, pid):
    """One setup / encrypt / keygen / decrypt at l = 123, the single-operation
    size of the benchmark, word for word against the oracle."""
    scheme_round_trip(gpu, po, pid, 123, 1)


def test_set1_generic_path_equals_specialised(gpu):
    """At the reference's set the generic kernels give the specialised
    reference-set path's outputs."""
    ps = gpu.PSet(1)
    a = gpu.sample_polya(h(b"g1A"))
    assert np.array_equal(ps.sample_polya(h(b"g1A")), a)
    l = 9
    mpk, msk = gpu.setup(a, h(b"g1S"), h(b"g1E"), l)
    mpk2, msk2 = ps.setup(a, h(b"g1S"), h(b"g1E"), l)
    assert np.array_equal(mpk, mpk2) and np.array_equal(msk, msk2)
    x = np.arange(2 * l, dtype=np.uint32).reshape(2, l) % 5
    c = gpu.encrypt(x, mpk, h(b"g12"), h(b"g13"), nonce0=11)
    assert np.array_equal(c, ps.encrypt(x, mpk, h(b"g12"), h(b"g13"), nonce0=11))
    y = (np.arange(2 * l, dtype=np.uint32).reshape(2, l) * 7) % 17
    sk = gpu.keygen(y, msk)
    assert np.array_equal(sk, ps.keygen(y, msk))
    assert np.array_equal(gpu.decrypt(c[1], y, sk), ps.decrypt(c[1], y, sk))
    keys = [h(b"g1k%d" % t) for t in range(2)]
    for which in (1, 2, 3):
        assert np.array_equal(gpu.gauss_sample(keys, [1, 2], which, 2), ps.gauss_sample(keys, [1, 2], which, 2))
