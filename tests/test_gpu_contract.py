"""The coefficient-0 contraction kernels agree bit for bit: the fraction form
(every limb in one int8 contraction, in every block shape), the per-limb int8
matrix-core kernel in every block shape (and its autotuned choice), with the
CRT and rounding fused into the last limb's epilogue or run as a separate
kernel, the fp64 scalar-operand kernel and the fp64 LDS kernel, on one model
and one row store.  Keys too large for the two-digit split run on fp64 and
still match the oracle."""
import hashlib
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

N = 4096


def h(t: bytes) -> bytes:
    return hashlib.sha256(t).digest()


def note(msg):
    print(f"[gpu-contract] {msg}", flush=True)


def model_with(gpu, rng, S, dim, hi):
    sv = np.zeros((S, 785), np.uint32)
    for v in range(S):
        idx = rng.choice(np.arange(1, dim), size=8, replace=False)
        sv[v, idx] = rng.integers(1, hi, 8)
    coef = rng.integers(-16, 17, (1, S)).astype(np.float64)
    return gpu.SvmModel(2, [S // 2, S - S // 2], [1, -1], coef, [0.5], sv)


def predict_with(ctx, R, env):
    old = {k: os.environ.get(k) for k in env}
    try:
        for k, v in env.items():
            os.environ[k] = v
        lab, kv, _ = ctx.predict(0, R, want_kvalues=True)
    finally:
        for k, v in old.items():
            if v is None:
                os.environ.pop(k, None)
            else:
                os.environ[k] = v
    return lab, kv


@pytest.mark.parametrize("hi,digits", [(2, 2), (17, 3)])
def test_kernels_agree(gpu, hi, digits):
    """Binary features give key coefficients within two digits; features up to
    16 need three."""
    rng = np.random.default_rng(31)
    l, R, S = 124, 700, 300
    ctx = gpu.Context()
    ctx.setup(h(b"cA"), h(b"cS"), h(b"cE"), 785)
    x = rng.integers(0, 5, (R, l)).astype(np.uint32)
    ctx.rows_reserve(R, l)
    ctx.encrypt_rows(x, h(b"c2"), h(b"c3"), nonce0=0)
    model = model_with(gpu, rng, S, l, hi)
    ctx.set_model(model)
    kb, maxabs = ctx.model_info()
    note(f"maxabs {maxabs} digits {ctx.contract_digits()}")
    assert ctx.contract_digits() == digits
    lab0, kv0 = predict_with(ctx, R, {})
    ip = (x.astype(np.int64) @ model.sv[:, :l].astype(np.int64).T) % 50241
    assert np.array_equal(kv0.astype(np.int64) % 50241, ip)
    variants = gpu.contract_variants(digits)
    note(f"variants {variants}")
    assert len(variants) >= 6
    # every per-limb variant (the fused CRT epilogue on limb 2), the separate
    # CRT kernel, the fp64 kernels and every fraction-form variant (two key
    # digits only)
    limbs = [{"CUFE_I8_VARIANT": str(v), "CUFE_CONTRACT": "limbs"} for v in variants]
    fracs = [{"CUFE_FRAC_VARIANT": str(v)} for v in range(gpu.frac_variants())] if digits == 2 else []
    for env in limbs + fracs + [{"CUFE_NO_FUSED_CRT": "1", "CUFE_CONTRACT": "limbs"}, {"CUFE_CONTRACT": "sgpr"},
                                {"CUFE_CONTRACT": "lds"}]:
        note(f"kernel {env}")
        lab, kv = predict_with(ctx, R, env)
        assert np.array_equal(kv, kv0), env
        assert np.array_equal(lab, lab0), env
    note(gpu.autotune_report())
    ctx.close()


def test_longest_rows_every_variant(gpu):
    """Rows of all 785 slots and keys whose support reaches slot 783 give the
    longest contraction (K = 4,880 columns, 153 k steps): every int8 variant,
    the split-K form and the fp64 kernel agree, and every key value decrypts
    <x, y> mod p."""
    rng = np.random.default_rng(34)
    l, R, S = 785, 200, 150
    ctx = gpu.Context()
    ctx.setup(h(b"wA"), h(b"wS"), h(b"wE"), 785)
    x = rng.integers(0, 5, (R, l)).astype(np.uint32)
    ctx.rows_reserve(R, l)
    ctx.encrypt_rows(x, h(b"w2"), h(b"w3"), nonce0=0)
    sv = np.zeros((S, 785), np.uint32)
    for v in range(S):
        sv[v, rng.choice(np.arange(1, 784), 12, replace=False)] = 1
    sv[0, 783] = 1
    coef = rng.integers(-16, 17, (1, S)).astype(np.float64)
    model = gpu.SvmModel(2, [S // 2, S - S // 2], [1, -1], coef, [0.5], sv)
    ctx.set_model(model)
    assert ctx.contract_digits() == 2
    lab0, kv0 = predict_with(ctx, R, {})
    ip = (x.astype(np.int64) @ sv.astype(np.int64).T) % 50241
    assert np.array_equal(kv0.astype(np.int64) % 50241, ip)
    for env in ([{"CUFE_I8_VARIANT": str(v), "CUFE_CONTRACT": "limbs"} for v in gpu.contract_variants(2)] +
                [{"CUFE_FRAC_VARIANT": str(v)} for v in range(gpu.frac_variants())] + [{"CUFE_CONTRACT": "sgpr"}]):
        lab, kv = predict_with(ctx, R, env)
        assert np.array_equal(kv, kv0), env
        assert np.array_equal(lab, lab0), env
    lab_s, kv_s, _ = ctx.predict(10, 100, want_kvalues=True)  # split-K: at most 256 rows
    assert np.array_equal(kv_s, kv0[10:110]) and np.array_equal(lab_s, lab0[10:110])
    note("785-slot rows: every variant equal")
    ctx.close()


def test_split_k_small_batches_agree(gpu):
    """Row ranges of at most 256 rows run all limbs and slices of the k steps
    as one split-K launch plus a finish kernel (sum of slices mod q, CRT);
    every size, including ragged and unaligned ones, gives the bits of the
    per-limb kernels."""
    rng = np.random.default_rng(33)
    l, R, S = 124, 1100, 300
    ctx = gpu.Context()
    ctx.setup(h(b"kA"), h(b"kS"), h(b"kE"), 785)
    x = rng.integers(0, 5, (R, l)).astype(np.uint32)
    ctx.rows_reserve(R, l)
    ctx.encrypt_rows(x, h(b"k2"), h(b"k3"), nonce0=0)
    ctx.set_model(model_with(gpu, rng, S, l, 2))
    assert ctx.contract_digits() == 2
    _, kv_all = predict_with(ctx, R, {})  # 1,100 rows: one chunk above the split size
    for r0, n in ((0, 1), (0, 32), (0, 33), (64, 200), (37, 5), (0, 256), (70, 180), (0, 1024)):
        lab_s, kv_s, _ = ctx.predict(r0, n, want_kvalues=True)
        old = os.environ.get("CUFE_NO_SPLIT_K")
        os.environ["CUFE_NO_SPLIT_K"] = "1"
        try:
            lab_p, kv_p, _ = ctx.predict(r0, n, want_kvalues=True)
        finally:
            if old is None:
                os.environ.pop("CUFE_NO_SPLIT_K", None)
            else:
                os.environ["CUFE_NO_SPLIT_K"] = old
        note(f"split-K rows [{r0}, {r0 + n})")
        assert np.array_equal(kv_s, kv_p) and np.array_equal(lab_s, lab_p), (r0, n)
        assert np.array_equal(kv_s, kv_all[r0:r0 + n]), (r0, n)
    ctx.close()


def test_large_keys_fall_back_to_fp64(gpu, po):
    rng = np.random.default_rng(32)
    l, R, S = 40, 5, 6
    ctx = gpu.Context()
    ctx.setup(h(b"LA"), h(b"LS"), h(b"LE"), 785)
    mpk, msk = ctx.get_keys()
    x = rng.integers(0, 5, (R, l)).astype(np.uint32)
    ctx.rows_reserve(R, l)
    ctx.encrypt_rows(x, h(b"L2"), h(b"L3"), nonce0=3)
    model = model_with(gpu, rng, S, l, 60000)
    ctx.set_model(model)
    kb, maxabs = ctx.model_info()
    note(f"maxabs {maxabs} kb {kb}")
    assert ctx.contract_digits() == 0 and max(maxabs) > 8322943
    _, kv, _ = ctx.predict(0, R, want_kvalues=True)
    _, kv_lds = predict_with(ctx, R, {"CUFE_CONTRACT": "lds"})
    assert np.array_equal(kv, kv_lds)
    mpk_l = np.concatenate([mpk[:l], mpk[785:786]])
    for r in (0, 4):
        c_full = po.encrypt(x[r], mpk_l, h(b"L2"), h(b"L3"), 3 + r)
        for v in (0, 5):
            y_pad = np.zeros(l + 1, np.uint32)
            y_pad[:l] = model.sv[v, :l]
            sk = po.keygen(y_pad, msk[: l + 1])
            c_pad = np.zeros((l + 2, 3, N), np.uint32)
            c_pad[:l] = c_full[:l]
            c_pad[l + 1] = c_full[l]
            d = po.decrypt(c_pad, y_pad, sk)
            assert int(kv[r, v]) == po.crt_round(int(d[0, 0]), int(d[1, 0]), int(d[2, 0]))
    ctx.close()


@pytest.mark.parametrize("nr_class,integer_coef", [(2, True), (2, False), (4, False), (3, True)])
def test_decision_kernels_agree(gpu, nr_class, integer_coef):
    """The wave-per-row vote (exact-integer tree sums, or lane-0 ordered sums of
    rounded products) gives the labels of the one-thread-per-row vote that
    follows main.c:338-389 literally."""
    rng = np.random.default_rng(40 + nr_class)
    l, R, S = 64, 300, 700
    ctx = gpu.Context()
    ctx.setup(h(b"dA"), h(b"dS"), h(b"dE"), 785)
    x = rng.integers(0, 5, (R, l)).astype(np.uint32)
    ctx.rows_reserve(R, l)
    ctx.encrypt_rows(x, h(b"d2"), h(b"d3"), nonce0=0)
    sv = np.zeros((S, 785), np.uint32)
    for v in range(S):
        sv[v, rng.choice(np.arange(1, l), 6, replace=False)] = rng.integers(1, 4, 6)
    base = S // nr_class
    nSV = [base] * (nr_class - 1) + [S - base * (nr_class - 1)]
    npair = nr_class * (nr_class - 1) // 2
    if integer_coef:
        coef = rng.integers(-9, 10, (nr_class - 1, S)).astype(np.float64)
        rho = rng.integers(-50, 50, npair) + 0.5
    else:
        coef = rng.normal(0, 1, (nr_class - 1, S))
        rho = rng.normal(0, 30, npair)
    ctx.set_model(gpu.SvmModel(nr_class, nSV, list(range(1, nr_class + 1)), coef, rho, sv))
    lab, kv = predict_with(ctx, R, {})
    lab_t, kv_t = predict_with(ctx, R, {"CUFE_DECIDE": "thread"})
    assert np.array_equal(kv, kv_t)
    assert np.array_equal(lab, lab_t)
    from cufe_amd import svm_io

    model = gpu.SvmModel(nr_class, nSV, list(range(1, nr_class + 1)), coef, rho, sv)
    for r in range(0, R, 37):
        assert lab[r] == svm_io.decide(model, kv[r].astype(np.float64)), r
    note(f"decision nr_class={nr_class} integer={integer_coef}: labels {np.unique(lab, return_counts=True)}")
    ctx.close()
