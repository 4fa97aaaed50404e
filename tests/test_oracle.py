"""Pins of the CPU oracle (oracle/ipfe_oracle.c).

Each oracle function is checked against something independent of it:
  * the reference's own compiled code (oracle/_ref, built from the reference
    sources where they lie; skipped where the reference is absent);
  * golden vectors that the reference code produced (tests/golden);
  * standards and closed forms (FIPS-197 known answer, w^bitrev twiddles,
    Python big-integer CRT / rounding, the scheme's decryption identity).
"""
import hashlib

import numpy as np
import pytest

N = 4096
Q = (16760833, 2147352577, 2130706433)
QQ = Q[0] * Q[1] * Q[2]
P = 50241


def sha(a):
    return hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()


# ---- AES-256-CTR ---------------------------------------------------------------
def test_aes_fips197_known_answer(po):
    """FIPS-197 Appendix C.3 (AES-256)."""
    rk = np.empty(60, dtype=np.uint32)
    po.orc().orc_aes256_expand(np.arange(32, dtype=np.uint8), rk)
    out = np.empty(16, dtype=np.uint8)
    po.orc().orc_aes256_encrypt_block(rk, np.frombuffer(bytes.fromhex("00112233445566778899aabbccddeeff"),
                                                        dtype=np.uint8).copy(), out)
    assert out.tobytes().hex() == "8ea2b7ca516745bfeafc49904b496089"


def test_aes_ctr_stream_golden(po, golden):
    g = golden["aes"]
    ks = po.keystream(bytes.fromhex(g["key"]), g["nonce"], 64)
    assert ks[:384].hex() == g["first384"]
    assert hashlib.sha256(ks).hexdigest() == g["sha256_8192"]


def test_aes_ctr_counter_block_layout(po):
    """Block t encrypts BE16(nonce) || 0^10 || BE32(t): check block 9 of a stream
    against a direct single-block encryption."""
    key = bytes(range(1, 33))
    ks = po.keystream(key, 0xBEEF, 2)
    rk = np.empty(60, dtype=np.uint32)
    po.orc().orc_aes256_expand(np.frombuffer(key, dtype=np.uint8).copy(), rk)
    blk = np.zeros(16, dtype=np.uint8)
    blk[0], blk[1], blk[15] = 0xBE, 0xEF, 9
    out = np.empty(16, dtype=np.uint8)
    po.orc().orc_aes256_encrypt_block(rk, blk, out)
    assert ks[16 * 9:16 * 10] == out.tobytes()


def test_aes_ctr_vs_reference(po, ref):
    for nonce in (0, 1, 0xFFFF):
        key = hashlib.sha256(b"k%d" % nonce).digest()
        rs = po.RefStream(key, nonce)
        assert po.keystream(key, nonce, 9) == rs.squeeze(9)


# ---- mod_prime ------------------------------------------------------------------
def test_mod_prime_golden(po, golden):
    g = golden["mod_prime"]
    for s in range(3):
        got = [po.orc().orc_mod_prime(int(m), s) for m in g["m"]]
        assert got == g["r"][s]


def test_mod_prime_landing_case_returns_q(po):
    """m = 1023*2^24 + 16761857 folds to exactly 2*q0, so mod_prime returns q0."""
    assert po.orc().orc_mod_prime(1023 * 2**24 + 16761857, 0) == Q[0]
    assert po.orc().orc_mod_prime(2 * Q[1], 1) == Q[1]


def test_mod_prime_congruence(po):
    rng = np.random.default_rng(5)
    for m in rng.integers(0, 2**63, 3000, dtype=np.uint64).tolist():
        for s in range(3):
            r = po.orc().orc_mod_prime(m, s)
            assert r <= Q[s] and r % Q[s] == m % Q[s]


def test_mod_prime_vs_reference(po, ref):
    rng = np.random.default_rng(6)
    for m in rng.integers(0, 2**64 - 1, 20000, dtype=np.uint64).tolist():
        s = m % 3
        assert po.orc().orc_mod_prime(m, s) == ref.mod_prime(m, s)


# ---- twiddles and NTT -------------------------------------------------------------
def _bitrev12(k):
    return int(format(k, "012b")[::-1], 2)


def test_twiddles_closed_form_and_golden(po, golden):
    import ctypes as C

    roots = (3689, 1689795, 456478)
    psi = np.stack([np.ctypeslib.as_array(po.orc().orc_psi(s), (N,)) for s in range(3)]).astype(np.uint32)
    psi_inv = np.stack([np.ctypeslib.as_array(po.orc().orc_psi_inv(s), (N,)) for s in range(3)]).astype(np.uint32)
    for s in range(3):
        w = roots[s]
        assert pow(w, N, Q[s]) == Q[s] - 1  # primitive 2N-th root
        for k in (0, 1, 2, 3, 100, 2047, 4095):
            assert int(psi[s, k]) == pow(w, _bitrev12(k), Q[s])
            assert int(psi_inv[s, k]) * int(psi[s, k]) % Q[s] == 1
    assert sha(psi) == golden["tables"]["psi_sha256"]
    assert sha(psi_inv) == golden["tables"]["psi_inv_sha256"]


def test_ntt_golden(po, golden):
    for s in range(3):
        g = golden["ntt"][str(s)]
        x = np.random.default_rng(g["input_seed"]).integers(0, Q[s], N, dtype=np.uint32)
        a = x.copy()
        po.orc().orc_ct_forward(a, s)
        assert sha(a) == g["fwd_sha256"]
        b = x.copy()
        po.orc().orc_gs_reverse(b, s)
        assert sha(b) == g["inv_sha256"]


def test_ntt_closed_form(po):
    """CT output k = sum_j a_j w^(j (2 bitrev12(k) + 1)); GS inverts CT."""
    roots = (3689, 1689795, 456478)
    rng = np.random.default_rng(7)
    for s in range(3):
        q, w = Q[s], roots[s]
        x = rng.integers(0, q, N, dtype=np.uint32)
        a = x.copy()
        po.orc().orc_ct_forward(a, s)
        xs = [int(v) for v in x]
        for k in (0, 1, 5, 4095):
            e = 2 * _bitrev12(k) + 1
            wk = pow(w, e, q)
            acc, pw = 0, 1
            for j in range(N):
                acc = (acc + xs[j] * pw) % q
                pw = pw * wk % q
            assert int(a[k]) == acc
        b = a.copy()
        po.orc().orc_gs_reverse(b, s)
        assert np.array_equal(b, x)


def test_negacyclic_coefficient0_identity(po):
    """Coefficient 0 of the negacyclic product a*b is a0 b0 - sum_{j>=1} a_j b_{N-j}."""
    rng = np.random.default_rng(8)
    for s in range(3):
        q = Q[s]
        a = rng.integers(0, q, N, dtype=np.uint32)
        b = rng.integers(0, q, N, dtype=np.uint32)
        c = np.empty(N, dtype=np.uint32)
        po.orc().orc_poly_mul_ntt(a, b, c, s)
        ai, bi = [int(v) for v in a], [int(v) for v in b]
        want = (ai[0] * bi[0] - sum(ai[j] * bi[N - j] for j in range(1, N))) % q
        assert int(c[0]) == want


def test_ntt_vs_reference(po, ref):
    rng = np.random.default_rng(9)
    for s in range(3):
        for trial in range(3):
            x = rng.integers(0, Q[s] + 1, N, dtype=np.uint32)  # includes the value q
            a, b = x.copy(), x.copy()
            po.orc().orc_ct_forward(a, s)
            ref.CT_forward(b, s)
            assert np.array_equal(a, b)
            po.orc().orc_gs_reverse(a, s)
            ref.GS_reverse(b, s)
            assert np.array_equal(a, b)


# ---- samplers -----------------------------------------------------------------------
@pytest.mark.parametrize("which", [1, 2, 3])
def test_gauss_golden(po, golden, which):
    g = golden["gauss"][str(which)]
    calls = po.gauss_calls(bytes.fromhex(g["key"]), g["nonce"], which, 4)
    assert [sha(c) for c in calls] == g["sha256_calls"]
    assert calls[0][:, :16].tolist() == g["call0_head"]
    assert calls[3][:, -16:].tolist() == g["call3_tail"]


@pytest.mark.parametrize("which", [1, 2, 3])
def test_gauss_vs_reference(po, ref, which):
    for nonce in (0, 3):
        key = hashlib.sha256(b"gs%d%d" % (which, nonce)).digest()
        rs = po.RefStream(key, nonce)
        mine = po.gauss_calls(key, nonce, which, 3)
        for i in range(3):
            assert np.array_equal(mine[i], rs.gauss(which))


@pytest.mark.parametrize("which,k", [(1, 266), (2, 304214978), (3, 608429953)])
def test_gauss_statistics(po, which, k):
    """Mean 0, standard deviation k * sqrt(1/(2 ln 2)), signed residues consistent across limbs."""
    x = po.gauss_calls(hashlib.sha256(b"stat").digest(), which, which, 2).astype(np.int64)
    v = np.where(x[:, 0] > Q[0] // 2, x[:, 0] - Q[0], x[:, 0]).ravel() if which == 1 else None
    signed = []
    for s in range(3):
        xs = x[:, s].ravel()
        signed.append(np.where(xs > Q[s] // 2, xs - Q[s], xs))
    if which == 1:
        assert np.array_equal(signed[0], signed[1]) and np.array_equal(signed[1], signed[2])
        v = signed[0]
        sigma = k * np.sqrt(1 / (2 * np.log(2)))
        assert abs(v.mean()) < 4 * sigma / np.sqrt(v.size)
        assert abs(v.std() / sigma - 1) < 0.05
    else:
        # residues of one integer: CRT-consistent where |z| < q0/2 in every limb
        small = np.abs(signed[1]) < Q[0] // 2
        assert np.array_equal(signed[0][small], signed[1][small])


# ---- sample_polya (closed form; the reference function needs GMP) -------------------
def test_sample_polya_closed_form(po):
    seed = hashlib.sha256(b"polya").digest()
    a = po.sample_polya(seed)
    ks = po.keystream(seed, 0, 256)
    v = [int.from_bytes(ks[8 * i:8 * i + 8], "little") for i in range(N)]
    for s in range(3):
        assert a[s].tolist() == [x % Q[s] for x in v]


# ---- CRT, rounding, parameters ---------------------------------------------------------
def test_parameter_constants(po):
    assert QQ == 76687145727357674227351553
    assert QQ // P == 1526385735302993058007
    scale = (13798054, 441557681, 1912932552)  # SIFE_SCALE_M_MOD_Q_I, src/params.h:28
    ninv = (16756741, 2146828321, 2130186241)  # SIFE_NTT_NINV, src/consts.h:7
    for s in range(3):
        assert po.orc().orc_scale_mod(s) == (QQ // P) % Q[s] == scale[s]
        assert po.orc().orc_ninv(s) == ninv[s]
    # SIFE_CRT_CONSTS (src/params.h:17): q0^-1 mod q1 and (q0 q1)^-1 mod q2
    assert 206923011 * Q[0] % Q[1] == 1
    assert 2935204199007202395 % Q[2] * (Q[0] * Q[1]) % Q[2] == 1


def test_crt_round_closed_form(po):
    rng = np.random.default_rng(10)
    for _ in range(3000):
        x = int(rng.integers(0, 2**62)) * int(rng.integers(1, 2**24)) % QQ
        d = [x % q for q in Q]
        assert po.crt_value(*d) == x
        want = (x * P) // QQ + (1 if (x * P) % QQ > QQ // 2 else 0)
        assert po.crt_round(*d) == want
    # rounding is not reduced mod p: x just below Q rounds to p
    assert po.crt_round(*[(QQ - 1) % q for q in Q]) == P
    # a limb-0 value of q0 stays q0 in Garner's start value (mpz semantics)
    x = po.crt_value(Q[0], 5, 7)
    assert x % Q[0] == 0 and x % Q[1] == 5 and x % Q[2] == 7 and x >= Q[0]


# ---- scheme -----------------------------------------------------------------------------
def test_scheme_golden(po, golden):
    g = golden["scheme"]
    l = g["l"]
    h = bytes.fromhex
    mpk, msk = po.setup(h(g["seedA"]), h(g["seedS"]), h(g["seedE"]), l)
    assert sha(po.sample_polya(h(g["seedA"]))) == g["a_sha256"]
    assert sha(mpk) == g["mpk_sha256"]
    assert sha(msk) == g["msk_sha256"]
    cs = [po.encrypt(np.array(x, np.uint32), mpk, h(g["seed2"]), h(g["seed3"]), b) for b, x in enumerate(g["x"])]
    assert [sha(c) for c in cs] == g["c_sha256"]
    sks = [po.keygen(np.array(y, np.uint32), msk) for y in g["y"]]
    assert [sha(sk) for sk in sks] == g["sk_sha256"]
    ds = [po.decrypt(cs[0], np.array(y, np.uint32), sk) for y, sk in zip(g["y"], sks)]
    assert [sha(d) for d in ds] == g["d_sha256"]


def test_scheme_decrypts_inner_product(po):
    """The IPFE identity: round(p * CRT(d_y) / Q)[0] = <x, y> (mod p, with the
    reference's unreduced rounding: 0 may come back as p)."""
    l = 6
    h = lambda t: hashlib.sha256(t).digest()
    mpk, msk = po.setup(h(b"A"), h(b"S"), h(b"E"), l)
    rng = np.random.default_rng(11)
    for b in range(2):
        x = rng.integers(0, 5, l).astype(np.uint32)
        c = po.encrypt(x, mpk, h(b"2"), h(b"3"), b)
        for _ in range(2):
            y = rng.integers(0, 17, l).astype(np.uint32)
            d = po.decrypt(c, y, po.keygen(y, msk))
            r = po.crt_round(int(d[0, 0]), int(d[1, 0]), int(d[2, 0]))
            assert r % P == int(np.dot(x.astype(np.int64), y.astype(np.int64))) % P


def test_encrypt_prefix_property(po):
    """A length-l encryption equals the first l polynomials and c_L of a longer one."""
    h = lambda t: hashlib.sha256(t).digest()
    mpk, msk = po.setup(h(b"A"), h(b"S"), h(b"E"), 5)
    x = np.array([1, 2, 3, 0, 4], np.uint32)
    c5 = po.encrypt(x, mpk, h(b"2"), h(b"3"), 9)
    mpk3 = np.concatenate([mpk[:3], mpk[5:6]])
    c3 = po.encrypt(x[:3], mpk3, h(b"2"), h(b"3"), 9)
    assert np.array_equal(c3[:3], c5[:3]) and np.array_equal(c3[3], c5[5])
    mpk2, msk2 = po.setup(h(b"A"), h(b"S"), h(b"E"), 2)
    assert np.array_equal(mpk2[:2], mpk[:2]) and np.array_equal(msk2, msk[:2]) and np.array_equal(mpk2[2], mpk[5])


def test_keygen_closed_form(po):
    h = lambda t: hashlib.sha256(t).digest()
    l = 4
    mpk, msk = po.setup(h(b"A"), h(b"S"), h(b"E"), l)
    y = np.array([3, 0, 16, 1], np.uint32)
    sk = po.keygen(y, msk)
    want = sum(int(y[i]) * msk[i].astype(object) for i in range(l))
    for s in range(3):
        assert np.array_equal(sk[s].astype(object), want[s] % Q[s])


def test_svm_decision_matches_libsvm_vote(po):
    """Two-class vote: label[0] iff sum coef*k - rho > 0 (main.c:338-389)."""
    import ctypes as C  # noqa: F401

    nSV = np.array([2, 3], np.int32)
    label = np.array([1, -1], np.int32)
    coef = np.array([[0.5, 1.0, -0.25, -1.0, -0.5]], np.float64)
    rho = np.array([0.75], np.float64)
    for kv, want in (([1, 1, 1, 0, 0], 1.0), ([0, 0, 4, 1, 1], -1.0), ([1, 0, 1, 0, 0], -1.0)):
        k = np.array(kv, np.float64)
        got = po.orc().orc_svm_decide(2, 5, nSV, label, np.ascontiguousarray(coef.ravel()), rho, k, 0)
        assert got == want


# A hand-computed three-class model whose two sv_coef rows differ, so that the
# pair (class 0, class 2) reads different coefficients depending on which row
# main.c:357-358 takes for which class: coef1 = sv_coef[j-1] on class i's SVs,
# coef2 = sv_coef[i] on class j's SVs.  Classes 10 / 20 / 30 with 2 / 1 / 2 SVs.
VOTE3 = {
    "nSV": [2, 1, 2], "label": [10, 20, 30],
    "coef": [[1.0, 1.0, -1.0, 2.0, 0.0], [-1.0, -1.0, 1.0, -2.0, 0.5]],
    "rho": [-0.5, 1.0, 0.0],
}


def vote3_by_hand(kv):
    """The three pairwise sums written out for VOTE3 (main.c:348-376):
    (0,1): row0 on class 0 + row0 on class 1; (0,2): row1 on class 0 + row0 on
    class 2; (1,2): row1 on class 1 + row1 on class 2."""
    c0, c1 = VOTE3["coef"]
    rho = VOTE3["rho"]
    d01 = c0[0] * kv[0] + c0[1] * kv[1] + c0[2] * kv[2] - rho[0]
    d02 = c1[0] * kv[0] + c1[1] * kv[1] + c0[3] * kv[3] + c0[4] * kv[4] - rho[1]
    d12 = c1[2] * kv[2] + c1[3] * kv[3] + c1[4] * kv[4] - rho[2]
    votes = [0, 0, 0]
    votes[0 if d01 > 0 else 1] += 1
    votes[0 if d02 > 0 else 2] += 1
    votes[1 if d12 > 0 else 2] += 1
    best = max(range(3), key=lambda i: (votes[i], -i))  # first class with most votes
    return VOTE3["label"][best]


def test_svm_vote_three_classes_by_hand(po):
    """Pins the one-vs-one vote with differing coefficient rows: for
    kvalues 1..5 the right reading gives votes (2, 0, 1) -> label 10; reading
    the two rows the other way round gives (1, 0, 2) -> 30."""
    nSV = np.array(VOTE3["nSV"], np.int32)
    label = np.array(VOTE3["label"], np.int32)
    coef = np.ascontiguousarray(np.array(VOTE3["coef"], np.float64).ravel())
    rho = np.array(VOTE3["rho"], np.float64)
    # the last two: a three-way tie goes to the first class; class 30 by two votes
    cases = [([1, 2, 3, 4, 5], 10), ([0, 0, 3, 0, 0], 20), ([0, 0, 0, 1, 0], 10), ([2, 0, 0, 0, 4], 10),
             ([0, 5, 1, 0, 3], 10), ([3, 0, 0, 0, 0], 30)]
    for kv, want in cases:
        assert vote3_by_hand(kv) == want, kv
        got = po.orc().orc_svm_decide(3, 5, nSV, label, coef, rho, np.array(kv, np.float64), 0)
        assert got == want, (kv, got)
    # the swapped reading (coef1 = sv_coef[i], coef2 = sv_coef[j-1]) flips the
    # (0,2) decision of the first case, so the checks above would catch it
    c0, c1 = VOTE3["coef"]
    kv = [1, 2, 3, 4, 5]
    d02_swapped = c0[0] * kv[0] + c0[1] * kv[1] + c1[3] * kv[3] + c1[4] * kv[4] - VOTE3["rho"][1]
    assert d02_swapped < 0


def test_kvalues_coef0_equals_full_decryption(po):
    """orc_kvalues_coef0 (coefficient 0 of the decryption only) equals the full
    RNS decryption followed by round_extract of coefficient 0, for random
    ciphertexts and keys whose fold ends on zero slots (the row is shorter than
    the key length), including ciphertext words equal to q."""
    rng = np.random.default_rng(11)
    l, ly, R, S = 5, 8, 3, 4
    qs = np.array(Q, np.uint64)[None, :, None]
    c = (rng.integers(0, 1 << 62, (R, ly + 1, 3, N), dtype=np.uint64) % qs).astype(np.uint32)
    c[0, 1, 2, 0] = Q[2]  # a coefficient-0 word equal to q
    c[1, ly, 0, 17] = Q[0]  # a c_L word equal to q
    msk = np.empty((ly, 3, N), np.uint32)
    z = rng.integers(-900, 901, (ly, N))
    for j in range(3):
        msk[:, j, :] = np.where(z < 0, Q[j] + z, z).astype(np.uint32)
    y = np.zeros((S, ly), np.uint32)
    y[:, :l] = rng.integers(0, 17, (S, l))
    sk = np.stack([po.keygen(y[v], msk) for v in range(S)])
    kv = po.kvalues_coef0(c[:, ly], c[:, :l, :, 0], y, sk)
    for r in range(R):
        for v in range(S):
            d = po.decrypt(c[r], y[v], sk[v])
            assert kv[r, v] == int(po.orc().orc_round_extract_coef0(np.ascontiguousarray(d))), (r, v)
