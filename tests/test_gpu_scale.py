"""Size-independent properties of batched prediction at the benchmark's key size:
2,000 support vectors of the benchmark's kind (123-dimensional sparse binary
rows, integer coefficients) against 16,416 encrypted rows (513 row blocks of 32,
so the last 64-row tile is ragged) -- 3.3e7 (row, SV) pairs, every one checked:

1. decryption identity: each key value is <x, y> mod p (the reference's
   rounding has no final mod p, so 0 may come out as p; main.c:205 uses it as is);
2. the labels equal the vote of main.c:338-389 recomputed from those key values
   (exact: integer coefficients, |sum| < 2^31, rho = 0.5 never ties);
3. any row range gives the same bits as the whole store (unaligned start, both
   ends inside 64-row tiles);
4. the CRT fused into the last limb's contraction equals the separate CRT
   kernel, and the int8 contraction equals the fp64 kernel on a 2,048-row slice.

The pair-by-pair oracle comparison at small sizes is in test_gpu_predict.py."""
import hashlib
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

P = 50241
R, S, WIDTH = 16416, 2000, 124


def h(t: bytes) -> bytes:
    return hashlib.sha256(t).digest()


def note(msg):
    print(f"[gpu-scale] {msg}", flush=True)


def sparse_binary(rng, n):
    x = np.zeros((n, WIDTH), np.uint32)
    for r in range(n):
        x[r, 1 + rng.choice(123, 14, replace=False)] = 1
    return x


def predict_env(ctx, r0, n, env):
    old = {k: os.environ.get(k) for k in env}
    try:
        os.environ.update(env)
        lab, kv, _ = ctx.predict(r0, n, want_kvalues=True)
    finally:
        for k, v in old.items():
            if v is None:
                os.environ.pop(k, None)
            else:
                os.environ[k] = v
    return lab, kv


@pytest.fixture(scope="module")
def scale(gpu):
    rng = np.random.default_rng(20261202)
    x = sparse_binary(rng, R)
    sv = np.zeros((S, 785), np.uint32)
    sv[:, :WIDTH] = sparse_binary(rng, S)
    coef = rng.choice(np.concatenate([np.arange(-16, 0), np.arange(1, 17)]), S).astype(np.float64)[None, :]
    model = gpu.SvmModel(2, [S // 2, S - S // 2], [1, -1], coef, [0.5], sv)
    ctx = gpu.Context()
    note(f"setup, encrypt {R} rows")
    ctx.setup(h(b"scale-A"), h(b"scale-S"), h(b"scale-E"), 785)
    ctx.rows_reserve(R, WIDTH)
    ctx.encrypt_rows(x, h(b"scale-2"), h(b"scale-3"), nonce0=0)
    ctx.set_model(model)
    assert ctx.contract_digits() == 2
    note("predict all rows")
    lab, kv, _ = ctx.predict(0, R, want_kvalues=True)
    yield ctx, x, sv, coef, lab, kv
    ctx.close()


def test_decryption_identity_every_pair(scale):
    _, x, sv, _, _, kv = scale
    ip = (x.astype(np.int64) @ sv[:, :WIDTH].astype(np.int64).T) % P
    assert np.array_equal(kv.astype(np.int64) % P, ip)
    assert kv.max() <= P
    note(f"{kv.size} pairs; {int(np.sum(kv == P))} zeros decrypted as p")


def test_labels_equal_recomputed_vote(scale):
    _, _, _, coef, lab, kv = scale
    dec = kv.astype(np.float64) @ coef[0] - 0.5
    assert np.array_equal(lab, np.where(dec > 0, 1.0, -1.0))


def test_row_ranges_equal_whole(scale):
    ctx, _, _, _, lab, kv = scale
    for r0, n in ((1000, 3333), (16383, 33), (31, 1), (64 * 200 + 17, 2000)):
        note(f"rows [{r0}, {r0 + n})")
        lab_p, kv_p = predict_env(ctx, r0, n, {})
        assert np.array_equal(kv_p, kv[r0:r0 + n]), (r0, n)
        assert np.array_equal(lab_p, lab[r0:r0 + n]), (r0, n)


def test_separate_crt_and_fp64_agree(scale):
    ctx, _, _, _, lab, kv = scale
    note("separate CRT kernel")
    lab_s, kv_s = predict_env(ctx, 0, R, {"CUFE_NO_FUSED_CRT": "1", "CUFE_CONTRACT": "limbs"})
    assert np.array_equal(kv_s, kv) and np.array_equal(lab_s, lab)
    note("fp64 contraction, rows [0, 2048)")
    lab_f, kv_f = predict_env(ctx, 0, 2048, {"CUFE_CONTRACT": "sgpr"})
    assert np.array_equal(kv_f, kv[:2048]) and np.array_equal(lab_f, lab[:2048])
