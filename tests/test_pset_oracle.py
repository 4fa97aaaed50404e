"""The oracle's restatement for the library's three parameter sets
(oracle/pset_oracle.c), pinned without a reference for sets 0 and 2 (the
reference defines only set 1, src/params.h:14-28; its comment on other levels,
src/sample.c:16-19, gives no constants -- parity for sets 0 and 2 is unpinned
by the reference and pinned here by mathematics):

* at set 1 every function equals the fixed-set oracle, which is pinned against
  the reference build (test_oracle.py, test_reference_pins.py);
* the constants: primes of the form 2^a - 2^b + 1 with q = 1 mod 2N, roots of
  order exactly 2N, Q p < 2^128, widths whose uniform step is exact;
* mod_prime: congruence, and the landing case that returns q for every prime;
* NTT: the closed form A_k = sum_j a_j w^(j (2 bitrev(k) + 1)), GS o CT = id,
  the negacyclic product;
* the sampler: at set 2 (the reference's widths, 8192 samples per call) the
  first 4096 samples of a call are the reference-width sampler's call; at set
  0's widths, the discrete Gaussian's moments;
* CRT and rounding: Python big integers; the library's host CRT + rounding
  equals the oracle's;
* the scheme: every coefficient decrypts to <x, y> (mod p) -- including the
  worst-case noise of set 0 at full length -- the prefix property and the
  keygen closed form."""
import hashlib
import math

import numpy as np
import pytest

N_OF = {0: 2048, 1: 4096, 2: 8192}


def h(t: bytes) -> bytes:
    return hashlib.sha256(t).digest()


def isprime(n):
    if n < 2:
        return False
    for p in (2, 3, 5, 7, 11, 13, 17, 19, 23, 29, 31, 37):
        if n % p == 0:
            return n == p
    d, s = n - 1, 0
    while d % 2 == 0:
        d //= 2
        s += 1
    for a in (2, 3, 5, 7, 11, 13, 17, 19, 23, 29, 31, 37):
        x = pow(a, d, n)
        if x in (1, n - 1):
            continue
        for _ in range(s - 1):
            x = x * x % n
            if x == n - 1:
                break
        else:
            return False
    return True


def bitrev(k, bits):
    return int(format(k, f"0{bits}b")[::-1], 2)


@pytest.mark.parametrize("pid", [0, 1, 2])
def test_constants(po, pid):
    ps = po.pset(pid)
    assert ps.N == N_OF[pid] and ps.logN == N_OF[pid].bit_length() - 1 and ps.L == 785 and ps.p == 50241
    for j in range(ps.nmod):
        q, a, b, w = int(ps.q[j]), int(ps.fa[j]), int(ps.fb[j]), int(ps.root[j])
        assert isprime(q) and q == 2 ** a - 2 ** b + 1 and q < 2 ** 31 and (q - 1) % (2 * ps.N) == 0
        assert pow(w, ps.N, q) == q - 1  # order exactly 2N (N a power of two)
    assert ps.Q * ps.p < 2 ** 128
    if pid == 1:
        assert ps.qs == [16760833, 2147352577, 2130706433] and list(ps.k)[1:] == [266, 304214978, 608429953]


@pytest.mark.parametrize("k", [266, 262147, 524287, 304214978, 608429953])
def test_uniform_step_exact(k):
    """The uniform sampler's reduction (src/gauss.c:609-703) w - (floor(w BF /
    2^32) + 1) k, plus k if negative, equals w mod k for every kept w < UQ."""
    bf = (1 << 32) // k
    uq = k * bf
    assert uq < 2 ** 32  # the bound fits the 32-bit comparison of src/gauss.c:609-703 (no power-of-two k)
    rng = np.random.default_rng(k)
    ws = list(rng.integers(0, uq, 20000)) + [0, 1, k - 1, k, uq - 1, uq - k, uq - k - 1]
    for w in ws:
        w = int(w)
        y = w - ((((w * bf) >> 32) + 1) * k)
        if y < 0:
            y += k
        assert y == w % k


@pytest.mark.parametrize("pid", [0, 2])
def test_mod_prime_congruence_and_landing(po, pid):
    ps = po.pset(pid)
    rng = np.random.default_rng(pid)
    for j in range(ps.nmod):
        q, a, b = int(ps.q[j]), int(ps.fa[j]), int(ps.fb[j])
        for m in list(rng.integers(0, 2 ** 62, 2000)) + [0, q, 2 * q, 2 * q + 1, q * q]:
            r = po.ps_mod_prime(pid, int(m), j)
            assert r % q == int(m) % q and 0 <= r <= q
        # m whose fold stops exactly at 2q: returned as q, not 0
        hi = (2 * q) // (2 ** b - 1)
        lo = 2 * q - hi * (2 ** b - 1)
        m = hi * 2 ** a + lo
        assert lo < 2 ** a and m > 2 * q
        assert po.ps_mod_prime(pid, m, j) == q


@pytest.mark.parametrize("pid", [0, 2])
def test_ntt_closed_form_and_inverse(po, pid):
    ps = po.pset(pid)
    N = ps.N
    rng = np.random.default_rng(10 + pid)
    a = np.stack([rng.integers(0, q, N) for q in ps.qs]).astype(np.uint32)
    A = po.ps_ntt_forward(pid, a)
    for j, q in enumerate(ps.qs):
        w = int(ps.root[j])
        aj = [int(v) for v in a[j]]
        for k in (0, 1, N // 3, N - 1):
            e = 2 * bitrev(k, ps.logN) + 1
            want = sum(aj[i] * pow(w, i * e, q) for i in range(N)) % q
            assert int(A[j, k]) == want, (j, k)
    assert np.array_equal(po.ps_ntt_inverse(pid, A), a)
    # negacyclic product: coefficient k = sum_{i+j=k} a_i b_j - sum_{i+j=k+N} a_i b_j
    b = np.stack([rng.integers(0, q, N) for q in ps.qs]).astype(np.uint32)
    c = po.ps_ntt_inverse(pid, (po.ps_ntt_forward(pid, a).astype(object) * po.ps_ntt_forward(pid, b).astype(object)
                                % np.array(ps.qs, dtype=object)[:, None]).astype(np.uint32))
    for j, q in enumerate(ps.qs):
        aj, bj = [int(v) for v in a[j]], [int(v) for v in b[j]]
        for k in (0, 5, N - 1):
            want = (sum(aj[i] * bj[k - i] for i in range(k + 1)) - sum(aj[i] * bj[k + N - i] for i in range(k + 1, N)))
            assert int(c[j, k]) == want % q, (j, k)


def test_set1_equals_fixed_set_oracle(po):
    """At the reference's parameters the generic restatement gives the fixed-set
    oracle's outputs: NTT both ways, sampler calls, sample_polya, the scheme."""
    rng = np.random.default_rng(1)
    a = np.stack([rng.integers(0, q, 4096) for q in po.pset(1).qs]).astype(np.uint32)
    assert np.array_equal(po.ps_ntt_forward(1, a), po.ntt_forward(a).reshape(a.shape))
    assert np.array_equal(po.ps_ntt_inverse(1, a), po.ntt_inverse(a).reshape(a.shape))
    for which in (1, 2, 3):
        assert np.array_equal(po.ps_gauss_calls(1, h(b"g"), 5, which, 2), po.gauss_calls(h(b"g"), 5, which, 2))
    assert np.array_equal(po.ps_sample_polya(1, h(b"pa")), po.sample_polya(h(b"pa")))
    l = 3
    apoly = po.sample_polya(h(b"A"))
    mpk, msk = po.setup_from_a(apoly, h(b"S"), h(b"E"), l)
    mpk2, msk2 = po.ps_setup_from_a(1, apoly, h(b"S"), h(b"E"), l)
    assert np.array_equal(mpk, mpk2) and np.array_equal(msk, msk2)
    x = np.array([3, 0, 4], np.uint32)
    c = po.encrypt(x, mpk, h(b"2"), h(b"3"), 4)
    assert np.array_equal(c, po.ps_encrypt(1, x, mpk, h(b"2"), h(b"3"), 4))
    y = np.array([16, 1, 0], np.uint32)
    sk = po.keygen(y, msk)
    assert np.array_equal(sk, po.ps_keygen(1, y, msk))
    d = po.decrypt(c, y, sk)
    assert np.array_equal(d, po.ps_decrypt(1, c, y, sk))
    for k in (0, 7, 4095):
        assert po.ps_crt_round(1, d[:, k]) == po.crt_round(int(d[0, k]), int(d[1, k]), int(d[2, k]))
        assert po.ps_crt_value(1, d[:, k]) == po.crt_value(int(d[0, k]), int(d[1, k]), int(d[2, k]))


def test_set2_call_starts_as_reference_call(po):
    """Set 2 keeps the reference's widths with 8192 samples per call: the first
    4096 samples of each stream's first call are the reference sampler's first
    call, on the three limbs the sets share."""
    for which in (1, 2, 3):
        big = po.ps_gauss_calls(2, h(b"s2"), 9, which, 1)[0]
        ref = po.gauss_calls(h(b"s2"), 9, which, 1)[0]
        assert np.array_equal(big[:3, :4096], ref)


@pytest.mark.parametrize("which,k", [(2, 262147), (3, 524287)])
def test_set0_sampler_moments(po, which, k):
    """Set 0's widths: mean 0 and standard deviation k sigma_0, sigma_0 =
    sqrt(1 / (2 ln 2)), within sampling error; both limbs hold the same value."""
    ps = po.pset(0)
    out = po.ps_gauss_calls(0, h(b"m%d" % which), 0, which, 8)
    q0, q1 = ps.qs
    v0 = out[:, 0, :].astype(np.int64).ravel()
    v1 = out[:, 1, :].astype(np.int64).ravel()
    z0 = np.where(v0 > q0 // 2, v0 - q0, v0)
    z1 = np.where(v1 > q1 // 2, v1 - q1, v1)
    assert np.array_equal(z0, z1)
    sigma = k * math.sqrt(1 / (2 * math.log(2)))
    n = z0.size
    assert abs(z0.mean()) < 5 * sigma / math.sqrt(n)
    assert abs(z0.std() / sigma - 1) < 0.03


@pytest.mark.parametrize("pid", [0, 1, 2])
def test_crt_round_closed_form(po, gpu_lib_optional, pid):
    """Garner + rounding against Python big integers, on random residues and
    on residues of values next to multiples of floor(Q/p) and next to Q; the
    library's host CRT + rounding gives the same."""
    ps = po.pset(pid)
    Q, p = ps.Q, ps.p
    rng = np.random.default_rng(30 + pid)
    xs = [int(v) % Q for v in rng.integers(0, 2 ** 62, 300)]
    xs += [int.from_bytes(rng.bytes(16), "little") % Q for _ in range(300)]
    scale = Q // p
    xs += [m * scale + d for m in (0, 1, 7, p - 1) for d in (-5, 0, 5, scale // 2 - 1, scale // 2, scale // 2 + 1)
           if 0 <= m * scale + d < Q]
    xs += [Q - 1, Q - 2, 0, 1]
    for x in xs:
        res = np.array([x % q for q in ps.qs], np.uint32)
        assert po.ps_crt_value(pid, res) == x
        want = (x * p) // Q + (1 if (x * p) % Q > Q // 2 else 0)
        assert po.ps_crt_round(pid, res) == want
        if gpu_lib_optional is not None:
            assert gpu_lib_optional.PSet(pid).crt_round_host(res) == want
    # the representative q on limb 0: Garner starts from q0 as given, as
    # crt_reverse_gmp does, and the value stays = 0 mod Q
    res = np.array([ps.qs[0]] + [0] * (ps.nmod - 1), np.uint32)
    x = po.ps_crt_value(pid, res)
    assert x % Q == 0 and x >= ps.qs[0]
    want = (x * p) // Q + (1 if (x * p) % Q > Q // 2 else 0)
    assert po.ps_crt_round(pid, res) == want
    if gpu_lib_optional is not None:
        assert gpu_lib_optional.PSet(pid).crt_round_host(res) == want


@pytest.mark.parametrize("pid", [0, 2])
def test_scheme_decrypts_inner_product(po, pid):
    """Setup, encryption of two messages, keygen and decryption at l = 6: every
    coefficient rounds to <x, y> mod p; a shorter encryption is a prefix; the
    key is the closed form sum y_i msk_i."""
    ps = po.pset(pid)
    l = 6
    a = po.ps_sample_polya(pid, h(b"A%d" % pid))
    mpk, msk = po.ps_setup_from_a(pid, a, h(b"S"), h(b"E"), l)
    rng = np.random.default_rng(50 + pid)
    y = rng.integers(0, 17, l).astype(np.uint32)
    sk = po.ps_keygen(pid, y, msk)
    want_sk = sum(int(y[i]) * msk[i].astype(object) for i in range(l))
    for j, q in enumerate(ps.qs):
        assert np.array_equal(sk[j].astype(object), want_sk[j] % q)
    for nonce in (0, 3):
        x = rng.integers(0, 5, l).astype(np.uint32)
        c = po.ps_encrypt(pid, x, mpk, h(b"2"), h(b"3"), nonce)
        d = po.ps_decrypt(pid, c, y, sk)
        ip = int(x.astype(np.int64) @ y.astype(np.int64))
        for k in range(0, ps.N, 97):
            assert po.ps_crt_round(pid, d[:, k]) % ps.p == ip % ps.p, k
        mpk3 = np.ascontiguousarray(np.concatenate([mpk[:3], mpk[l:l + 1]]))
        c3 = po.ps_encrypt(pid, x[:3], mpk3, h(b"2"), h(b"3"), nonce)
        assert np.array_equal(c3[:3], c[:3]) and np.array_equal(c3[3], c[l])


def test_set0_worst_case_noise(po):
    """Set 0 at full length with the largest inputs (x_i = B_x = 4, y_i = B_y =
    16 on all 785 slots): <x, y> = p - 1 is recovered on every coefficient, so
    set 0's noise widths leave room at the largest |y|."""
    pid, l = 0, 785
    ps = po.pset(pid)
    a = po.ps_sample_polya(pid, h(b"wcA"))
    mpk, msk = po.ps_setup_from_a(pid, a, h(b"wcS"), h(b"wcE"), l)
    x = np.full(l, 4, np.uint32)
    y = np.full(l, 16, np.uint32)
    c = po.ps_encrypt(pid, x, mpk, h(b"wc2"), h(b"wc3"), 1)
    d = po.ps_decrypt(pid, c, y, po.ps_keygen(pid, y, msk))
    for k in range(0, ps.N, 31):
        assert po.ps_crt_round(pid, d[:, k]) == ps.p - 1, k
