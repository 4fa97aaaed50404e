"""The arithmetic of the fraction-form contraction (k_contract_frac.hip), on
the CPU against the oracle: for real ciphertexts and functional keys, the
72-bit fixed-point sum V = 2^72 sum_k alpha~[k] B[k] mod 2^72 built from nine
signed digits per column lies within the kernel's margin M_v of
2^72 x / Q (x = CRT of the limbs' exact residues), and floor((p V + 2^71) /
2^72) equals the oracle's rounding wherever the margin test passes; the
dropped digit pair (1, 1) contributes only multiples of 2^72.  The exact
kernel's refinement V88 = 2^16 V + sum_k D[k] B[k] mod 2^88 (D: the 16 bits
below the digits) lies within M88_v of 2^88 x / Q and decides every pair the
same way at 88 bits.  The epilogue's 64-bit word arithmetic equals the
128-bit one (its half-integer test a superset) on random and boundary
accumulators."""
import hashlib
from fractions import Fraction

import numpy as np

import conftest  # noqa: F401  (oracle path)
import pyoracle as po

QS = [16760833, 2147352577, 2130706433]
Q = QS[0] * QS[1] * QS[2]
P = 50241
N = 4096


def h(t: bytes) -> bytes:
    return hashlib.sha256(t).digest()


def bal(w, q):
    w = int(w) % q
    return w - q if w > q // 2 else w


def alpha_digits(words):
    """The kernel's k_rows_frac: alpha = sum_j (w_j u_j mod q_j) floor(2^128 / q_j)
    mod 2^128, rounded to 88 bits, that rounded to 72 bits, as nine digits
    (a_2..a_9 balanced, a_1 mod 256), and D = alpha~88 - 2^16 alpha~72."""
    acc = 0
    for j, q in enumerate(QS):
        u = pow(Q // q % q, q - 2, q)
        w = int(words[j])
        w = w - q if w >= q else w
        acc = (acc + (w * u % q) * (((1 << 128) - 1) // q)) % (1 << 128)
    t88 = ((acc + (1 << 39)) % (1 << 128)) >> 40
    t = (t88 + (1 << 15)) >> 16
    dref = t88 - (t << 16)
    assert -(1 << 15) <= dref < (1 << 15)
    a = [0] * 9
    for d in range(8, 0, -1):
        v = ((t & 255) + 128) % 256 - 128
        a[d] = v
        t = (t - v) >> 8
    a[0] = (t & 255) - 256 if (t & 255) >= 128 else (t & 255)
    return a, dref


def decide(V, M, bits):
    """The kernel's tests at `bits` bits: the key value, or None near a
    rounding change (a half-integer of p x / Q, or the wrap of x / Q)."""
    one, half = 1 << bits, 1 << (bits - 1)
    pv = P * V
    if V < M or V > one - M or abs(pv % one - half) <= P * M:
        return None
    return (pv + half) >> bits


def test_fraction_sum_within_margin_and_rounds_like_the_oracle():
    l = 24
    mpk, msk = po.setup(h(b"mA"), h(b"mS"), h(b"mE"), l)
    rng = np.random.default_rng(3)
    rows = rng.integers(0, 3, (3, l)).astype(np.uint32)
    svs = np.zeros((4, l), np.uint32)
    for v in range(4):
        svs[v, rng.choice(l - 1, 5, replace=False)] = rng.integers(1, 3, 5)
    # SV 0 meets every row in zeros: inner product 0, key value 0 or p by the
    # noise's sign (the wrap of x / Q)
    rows[:, :5] = 0
    svs[0] = 0
    svs[0, 1:4] = 1
    checked = wraps = 0
    for r in range(rows.shape[0]):
        c = po.encrypt(rows[r], mpk, h(b"m2"), h(b"m3"), r)
        cols = [c[l, :, k] for k in range(N)] + [c[i, :, 0] for i in range(l)]
        parts = [alpha_digits(w) for w in cols]
        digits = [a for a, _ in parts]
        dref = [dd for _, dd in parts]
        # with the pair (d, e) = (1, 1) left out, as the kernel does
        hi = [a[0] for a in digits]
        lo_all = [sum(a[d] * 256 ** (8 - d) for d in range(9)) for a in digits]
        for v in range(svs.shape[0]):
            sk = po.keygen(svs[v], msk)
            key = []
            for k in range(N):
                vals = [bal(sk[j, 0] if k == 0 else (QS[j] - sk[j, N - k]) % QS[j], QS[j]) for j in range(3)]
                assert vals[0] == vals[1] == vals[2]
                key.append(-vals[0])
            key += [int(y) for y in svs[v]]
            b1 = [(b - (((b + 128) & 255) - 128)) // 256 for b in key]
            V = sum(x * b for x, b in zip(lo_all, key)) % (1 << 72)
            V17 = (V - sum(256 ** 9 * a1 * e1 for a1, e1 in zip(hi, b1))) % (1 << 72)
            assert V17 == V  # the (1, 1) products are multiples of 2^72
            d = [sum(bal(cols[k][j], QS[j]) * key[k] for k in range(len(key))) % QS[j] for j in range(3)]
            x = po.crt_value(*d)
            s = sum(abs(b) for b in key)
            M = (s + 1) // 2 + (s + (1 << 16) - 1) // (1 << 16) + 1
            M88 = (s + 1) // 2 + (s + 127) // 128 + 1
            T = Fraction(x * (1 << 72), Q)
            err = abs(T - V)
            err = min(err, (1 << 72) - err)
            assert err < M
            ref = po.crt_round(*d)
            kv = decide(V, M, 72)
            if kv is not None:
                assert kv == ref
                checked += 1
                wraps += kv in (0, P)
            V88 = (V * (1 << 16) + sum(dd * b for dd, b in zip(dref, key))) % (1 << 88)
            T88 = Fraction(x * (1 << 88), Q)
            err88 = abs(T88 - V88)
            err88 = min(err88, (1 << 88) - err88)
            assert err88 < M88
            # a pair open at 72 bits is settled at 88 unless x / Q lies within
            # about 2^-66 of a rounding change
            assert decide(V88, M88, 88) == ref
    assert checked >= 10 and wraps >= 3


def words_of_v(acc):
    """The kernel's frac_value: V = sum_c C_c 2^(8 (8 - c)) mod 2^72 as two
    words (lo 64 bits, hi 8 bits), from pairwise 64-bit merges and carries."""
    m64 = (1 << 64) - 1
    p0, p1 = acc[0] * 256 + acc[1], acc[2] * 256 + acc[3]
    p2, p3 = acc[4] * 256 + acc[5], acc[6] * 256 + acc[7]
    q0, q1, c8 = p0 * 65536 + p1, p2 * 65536 + p3, acc[8]
    assert abs(q0) < 1 << 57 and abs(q1) < 1 << 57
    a, b = (q0 << 40) & m64, (q1 << 8) & m64
    s1 = (a + b) & m64
    s2 = (s1 + (c8 & m64)) & m64
    hi = ((q0 >> 24) + (q1 >> 56) + (c8 >> 63) + (s1 < a) + (s2 < s1)) & 255
    return s2, hi


def decide_words(lo, hi, m):
    """The kernel's epilogue tests on 64-bit words: (near a half-integer, near
    the wrap, key value)."""
    t = P * (lo & 0xFFFFFFFF)
    u = P * (lo >> 32) + (t >> 32)
    thi = (P * hi + (u >> 32)) & 0xFFFFFFFF
    d = ((thi & 255) - 128) * (1 << 32) + (u & 0xFFFFFFFF)
    pmu = (P * m) >> 32
    halfway = -pmu - 2 <= d <= pmu + 2
    wrap = (hi == 0 and lo < m) or (hi == 255 and lo != 0 and (1 << 64) - lo < m)
    return halfway, wrap, (thi + 128) >> 8


def test_epilogue_words_against_exact_integers():
    """The 64-bit epilogue against 128-bit integer arithmetic: V's words, the
    key value wherever no test fires, the wrap test exactly, and a
    half-integer test that fires whenever the exact one does.  Accumulators
    are random or built so that V sits on and around 0, 2^72 and the
    half-integers of p V / 2^72, for margins from 1 to the 2^62 cap."""
    rng = np.random.default_rng(8)
    one = 1 << 72

    def balanced(v):
        # nine accumulators summing to v mod 2^72, then 2^8 r moved between
        # neighbours so that they are large (as after 4,224 products)
        c = [0] * 9
        for k in range(8, 0, -1):
            c[k] = ((v & 255) + 128) % 256 - 128
            v = (v - c[k]) >> 8
        c[0] = ((v & 255) + 128) % 256 - 128
        for k in range(8, 0, -1):
            r = int(rng.integers(-(1 << 22), 1 << 22))
            c[k] += 256 * r
            c[k - 1] -= r
        return c

    cases = [[int(x) for x in rng.integers(-(1 << 31) + 1, 1 << 31, 9)] for _ in range(3000)]
    for m in (1, 1 << 21, 1 << 33, 1 << 50, 1 << 62):
        for base in (0, one):
            for d in (-m - 1, -m, -m + 1, -2, -1, 0, 1, 2, m - 1, m, m + 1, 1 << 32, -(1 << 32)):
                cases.append((balanced((base + d) % one), m))
        for k in range(0, P, 4999):
            v0 = ((2 * k + 1) << 71) // P  # p V / 2^72 close to k + 1/2
            for d in (-3, -1, 0, 1, 3, m, -m, 1 << 32):
                cases.append((balanced((v0 + d) % one), m))
    margins = (1, 7, 1 << 21, 1 << 40, 1 << 62)
    for n, case in enumerate(cases):
        acc, m = case if isinstance(case, tuple) else (case, margins[n % len(margins)])
        v = sum(c << (8 * (8 - k)) for k, c in enumerate(acc)) % one
        lo, hi = words_of_v(acc)
        assert lo + (hi << 64) == v
        halfway, wrap, kv = decide_words(lo, hi, m)
        fr = P * v % one
        e_wrap = v < m or v > one - m
        e_half = fr + P * m >= 1 << 71 and fr <= (1 << 71) + P * m
        assert wrap == e_wrap, (acc, m)
        assert halfway or not e_half, (acc, m)
        if not (halfway or wrap):
            assert kv == (P * v + (1 << 71)) >> 72
