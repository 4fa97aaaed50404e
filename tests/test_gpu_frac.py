"""The fraction form of the batched decryption: every limb in one int8
contraction over 72-bit fractions of each column's CRT value, pairs near a
rounding change refined to 88 bits and, where that is not enough, recomputed
exactly.  It must give the key values and labels of
the per-limb exact contraction on every pair: with the default margins, with
margins widened so that many pairs (including the zero inner products whose
key value is 0 or p) go to the exact kernel, and with margins so wide that the
exact kernel's list overflows and the call runs again per limb."""
import hashlib
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

P = 50241


def h(t: bytes) -> bytes:
    return hashlib.sha256(t).digest()


def note(msg):
    print(f"[gpu-frac] {msg}", flush=True)


def binary_model(gpu, rng, S, dim, ones=10):
    sv = np.zeros((S, 785), np.uint32)
    for v in range(S):
        sv[v, rng.choice(np.arange(1, dim), size=ones, replace=False)] = 1
    coef = rng.integers(-16, 17, (1, S)).astype(np.float64)
    return gpu.SvmModel(2, [S // 2, S - S // 2], [1, -1], coef, [0.5], sv)


class Env:
    def __init__(self, **kv):
        self.kv = kv

    def __enter__(self):
        self.old = {k: os.environ.get(k) for k in self.kv}
        os.environ.update(self.kv)

    def __exit__(self, *a):
        for k, v in self.old.items():
            if v is None:
                os.environ.pop(k, None)
            else:
                os.environ[k] = v


def rows_ctx(gpu, rng, R, l, tag):
    ctx = gpu.Context()
    ctx.setup(h(tag + b"A"), h(tag + b"S"), h(tag + b"E"), 785)
    # sparse rows: many zero inner products (key value 0 or p)
    x = np.zeros((R, l), np.uint32)
    for r in range(R):
        x[r, rng.choice(np.arange(1, l), size=6, replace=False)] = rng.integers(1, 5, 6)
    ctx.rows_reserve(R, l)
    ctx.encrypt_rows(x, h(tag + b"2"), h(tag + b"3"), nonce0=0)
    return ctx, x


def per_limb(ctx, r0, n):
    with Env(CUFE_CONTRACT="limbs"):
        lab, kv, _ = ctx.predict(r0, n, want_kvalues=True)
    assert ctx.contract_form()[0] == 1
    return lab, kv


def test_fraction_form_equals_per_limb(gpu):
    rng = np.random.default_rng(41)
    l, R, S = 124, 1500, 320
    ctx, x = rows_ctx(gpu, rng, R, l, b"f")
    model = binary_model(gpu, rng, S, l)
    ctx.set_model(model)
    assert ctx.contract_digits() == 2
    lab, kv, _ = ctx.predict(0, R, want_kvalues=True)
    form, exact = ctx.contract_form()
    note(f"form {form}, pairs recomputed exactly {exact} of {R * S}")
    assert form == 2
    lab_l, kv_l = per_limb(ctx, 0, R)
    assert np.array_equal(kv, kv_l) and np.array_equal(lab, lab_l)
    ip = (x.astype(np.int64) @ model.sv[:, :l].astype(np.int64).T) % P
    assert np.array_equal(kv.astype(np.int64) % P, ip)
    zero_as_p = int(np.sum(kv == P))
    note(f"zero inner products decrypted as p: {zero_as_p}")
    assert zero_as_p > 0
    # ragged ranges (rows from a 32-row boundary, then kept), every shape
    for r0, n in ((0, 1), (5, 40), (31, 33), (64, 257), (100, 1399)):
        lab_p, kv_p, _ = ctx.predict(r0, n, want_kvalues=True)
        assert np.array_equal(kv_p, kv[r0:r0 + n]) and np.array_equal(lab_p, lab[r0:r0 + n]), (r0, n)
    for v in range(gpu.frac_variants()):
        with Env(CUFE_FRAC_VARIANT=str(v)):
            lab_v, kv_v, _ = ctx.predict(0, R, want_kvalues=True)
        assert ctx.contract_form()[0] == 2
        assert np.array_equal(kv_v, kv) and np.array_equal(lab_v, lab), v
    ctx.close()


@pytest.mark.parametrize("refine", [True, False])
@pytest.mark.parametrize("widen", [24, 30])
def test_exact_kernel_on_many_pairs(gpu, widen, refine):
    """Margins 2^widen times wider: the zero inner products (wrap of x / Q)
    and a share of the others go to the exact kernel, which must give the
    per-limb key values, 0 and p included -- through its 88-bit refinement
    (whose widened margins still leave many zero inner products to the
    limbs' residues), or, with the refinement switched off, every listed
    pair from the limbs' residues (near a half-integer: all three limbs)."""
    rng = np.random.default_rng(42 + widen)
    l, R, S = 124, 600, 200
    ctx, x = rows_ctx(gpu, rng, R, l, b"w")
    with Env(CUFE_FRAC_WIDEN=str(widen)):
        ctx.set_model(binary_model(gpu, rng, S, l))
    env = {} if refine else {"CUFE_FRAC_NO_REFINE": "1"}
    with Env(**env):
        lab, kv, _ = ctx.predict(0, R, want_kvalues=True)
    form, exact = ctx.contract_form()
    note(f"widen 2^{widen}, refinement {refine}: {exact} of {R * S} pairs listed for the exact kernel")
    assert form == 2 and 0 < exact < R * S
    lab_l, kv_l = per_limb(ctx, 0, R)
    assert np.array_equal(kv, kv_l) and np.array_equal(lab, lab_l)
    ctx.close()


def test_chunked_calls_equal_one_chunk(gpu):
    """A call whose rows run in several chunks (each its own contraction, list
    and exact kernel) gives the one-chunk call's results; with margins wide
    enough that each chunk lists pairs, the totals add up."""
    rng = np.random.default_rng(45)
    l, R, S = 124, 1836, 150  # chunks of 512: 512 x 3 + 300 (every chunk above the split-K size)
    ctx, x = rows_ctx(gpu, rng, R, l, b"c")
    with Env(CUFE_FRAC_WIDEN="26"):
        ctx.set_model(binary_model(gpu, rng, S, l))
    lab, kv, _ = ctx.predict(0, R, want_kvalues=True)
    form, exact = ctx.contract_form()
    assert form == 2 and exact > 0
    with Env(CUFE_PREDICT_CHUNK_ROWS="512"):
        lab_c, kv_c, _ = ctx.predict(0, R, want_kvalues=True)
        form_c, exact_c = ctx.contract_form()
    note(f"one chunk: {exact} listed; chunks of 512 rows: {exact_c}")
    assert form_c == 2 and exact_c == exact
    assert np.array_equal(kv_c, kv) and np.array_equal(lab_c, lab)
    lab_l, kv_l = per_limb(ctx, 0, R)
    assert np.array_equal(kv, kv_l) and np.array_equal(lab, lab_l)
    ctx.close()


def test_list_overflow_runs_per_limb(gpu):
    """Every pair near a rounding change and more of them than the exact
    kernel's list holds (2^20): the call runs again per limb, same results."""
    rng = np.random.default_rng(44)
    l, R, S = 40, 4200, 256
    ctx, x = rows_ctx(gpu, rng, R, l, b"o")
    with Env(CUFE_FRAC_WIDEN="62"):  # margins 2^62: p M > 2^72, every pair listed
        ctx.set_model(binary_model(gpu, rng, S, l, ones=6))
    lab, kv, _ = ctx.predict(0, R, want_kvalues=True)
    form, exact = ctx.contract_form()
    note(f"form {form}, listed {exact} of {R * S}")
    assert form == 1 and exact == R * S
    lab_l, kv_l = per_limb(ctx, 0, R)
    assert np.array_equal(kv, kv_l) and np.array_equal(lab, lab_l)
    ctx.close()
