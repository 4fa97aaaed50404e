"""The library's CRT and rounding (crt_decode, the code the GPU kernels and
the host share), through its host-only entry point cufe_round_extract2,
against the oracle and the closed form round(p x / Q) (src/crt.c:46-82,
src/rlwe_sife.c:334-393): every residue edge (0, 1, q-1, q, q+1 in each limb,
unreduced words), values beside each rounding boundary x = (2m+1) Q / (2p)
and around the distance where the fast double estimate hands over to exact
128-bit arithmetic, and random residues.  No GPU: the call runs on the host."""
import numpy as np
import pytest

Q = (16760833, 2147352577, 2130706433)
QQ = Q[0] * Q[1] * Q[2]
P = 50241
N = 4096


@pytest.fixture(scope="module")
def lib_host():
    import cufe_amd

    return cufe_amd


def decode(cf, v):
    d = np.zeros((3, N), np.uint32)
    d[:, 0] = v
    return int(cf.round_extract2(d))


def closed_form(x):
    return (x * P) // QQ + (1 if (x * P) % QQ > QQ // 2 else 0)


def test_edges_match_oracle(lib_host, po):
    edge = sorted({e for q in Q for e in (q - 1, q, q + 1)} | {0, 1, 2, 2**31 - 1, 2**31, 2**32 - 1})
    n = 0
    for a in edge:
        for b in edge:
            for c in edge:
                assert decode(lib_host, (a, b, c)) == po.crt_round(a, b, c), (a, b, c)
                n += 1
    assert n == len(edge) ** 3


def test_rounding_boundaries(lib_host):
    """Beside each boundary, and at offsets around 7.6e11 where the library
    switches from its double estimate to 128-bit integers (|2 N'/Q - 1| =
    1e-9), on both sides."""
    rng = np.random.default_rng(41)
    for _ in range(3000):
        m = int(rng.integers(0, P))
        x0 = ((2 * m + 1) * QQ) // (2 * P)
        offsets = [-1, 0, 1, 2] + [int(s * 10 ** e) for s in (-1, 1) for e in (11.5, 11.9, 12.0, 13.0, 15.0)]
        for dx in offsets:
            x = min(max(x0 + dx, 0), QQ - 1)
            assert decode(lib_host, tuple(x % q for q in Q)) == closed_form(x), x
    # x just below Q rounds to p (no mod p), and 0 stays 0
    assert decode(lib_host, tuple((QQ - 1) % q for q in Q)) == P
    assert decode(lib_host, (0, 0, 0)) == 0


def test_random_residues(lib_host, po):
    rng = np.random.default_rng(42)
    for _ in range(5000):
        v = tuple(int(rng.integers(0, q + 1)) for q in Q)
        assert decode(lib_host, v) == po.crt_round(*v), v
