"""Batched IPFE-SVM prediction on the GPU against the CPU oracle, pair by pair.

The oracle follows svm_predict_cpu (main.c:137-267) for each (row, support
vector): dense keygen, full decryption, CRT + rounding of coefficient 0, vote.
The reference runs those loops over L = 785 slots with zero padding; the oracle
here uses the row's l slots plus one zero slot, which gives the same folds
(every slot after the last non-zero y_i is a zero step that only canonicalises
the accumulator)."""
import hashlib
import os
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

N = 4096
HERE = os.path.dirname(os.path.abspath(__file__))


def h(t: bytes) -> bytes:
    return hashlib.sha256(t).digest()


def note(msg):
    print(f"[gpu-predict] {msg}", flush=True)


@pytest.fixture(scope="module")
def keyed_ctx(gpu):
    note("ctx setup l=785")
    ctx = gpu.Context()
    ctx.setup(h(b"pA"), h(b"pS"), h(b"pE"), 785)
    mpk, msk = ctx.get_keys()
    return ctx, mpk, msk


def oracle_pair(po, c_full, y_row, msk, l):
    """kvalue of one (row, SV) pair with the L-slot semantics."""
    y_pad = np.zeros(l + 1, np.uint32)
    y_pad[:l] = y_row[:l]
    sk = po.keygen(y_pad, msk[: l + 1])
    c_pad = np.zeros((l + 2, 3, N), np.uint32)
    c_pad[:l] = c_full[:l]
    c_pad[l + 1] = c_full[l]
    d = po.decrypt(c_pad, y_pad, sk)
    return po.crt_round(int(d[0, 0]), int(d[1, 0]), int(d[2, 0])), sk


def make_model(gpu, rng, S, l, nr_class=2, nnz=5, integer_coef=True):
    sv = np.zeros((S, 785), np.uint32)
    for v in range(S):
        idx = rng.choice(np.arange(1, l), size=nnz, replace=False)
        sv[v, idx] = rng.integers(1, 17, nnz)
    base = S // nr_class
    nSV = [base] * (nr_class - 1) + [S - base * (nr_class - 1)]
    if integer_coef:
        coef = rng.integers(-16, 17, (nr_class - 1, S)).astype(np.float64)
        rho = np.full(nr_class * (nr_class - 1) // 2, 0.5)
    else:
        coef = rng.normal(0, 1, (nr_class - 1, S))
        rho = rng.normal(0, 10, nr_class * (nr_class - 1) // 2)
    labels = list(range(1, nr_class + 1))
    return gpu.SvmModel(nr_class, nSV, labels, coef, rho, sv)


@pytest.mark.parametrize("nr_class", [2, 3])
def test_predict_matches_oracle(gpu, po, keyed_ctx, nr_class):
    from cufe_amd import svm_io

    ctx, mpk, msk = keyed_ctx
    rng = np.random.default_rng(40 + nr_class)
    l, R, S = 24, 5, 36
    note(f"predict nr_class={nr_class} rows={R} svs={S} l={l}")
    model = make_model(gpu, rng, S, l, nr_class=nr_class, integer_coef=(nr_class == 2))
    x = rng.integers(0, 5, (R, l)).astype(np.uint32)
    ctx.rows_reserve(R, l)
    ctx.encrypt_rows(x, h(b"e2"), h(b"e3"), nonce0=7)
    ctx.set_model(model)
    labels, kv, ms = ctx.predict(0, R, want_kvalues=True)
    mpk_l = np.concatenate([mpk[:l], mpk[785:786]])
    sk_gpu = ctx.model_keys()
    for r in range(R):
        c_full = po.encrypt(x[r], mpk_l, h(b"e2"), h(b"e3"), 7 + r)
        cl, c0 = ctx.get_row(r)
        assert np.array_equal(cl, c_full[l])
        assert np.array_equal(c0, c_full[:l, :, 0])
        want_kv = np.zeros(S)
        for v in range(S):
            k, sk = oracle_pair(po, c_full, model.sv[v], msk, l)
            if r == 0:
                assert np.array_equal(sk_gpu[v], sk)
            want_kv[v] = k
            assert int(kv[r, v]) == k, (r, v)
            assert k % 50241 == int(x[r].astype(np.int64) @ model.sv[v, :l].astype(np.int64)) % 50241
        assert labels[r] == svm_io.decide(model, want_kv)


def test_predict_row_subsets_and_chunking(gpu, keyed_ctx):
    """Predicting any row range gives the same labels and kvalues."""
    ctx, mpk, msk = keyed_ctx
    rng = np.random.default_rng(50)
    l, R, S = 16, 300, 20
    model = make_model(gpu, rng, S, l)
    x = rng.integers(0, 5, (R, l)).astype(np.uint32)
    ctx.rows_reserve(R, l)
    ctx.encrypt_rows(x[:130], h(b"f2"), h(b"f3"), nonce0=0)
    ctx.encrypt_rows(x[130:], h(b"f2"), h(b"f3"), nonce0=130)
    ctx.set_model(model)
    lab_all, kv_all, _ = ctx.predict(0, R, want_kvalues=True)
    lab_part, kv_part, _ = ctx.predict(77, 101, want_kvalues=True)
    assert np.array_equal(lab_all[77:178], lab_part) and np.array_equal(kv_all[77:178], kv_part)
    # an unaligned range on the fp64 LDS kernel gives the same values
    os.environ["CUFE_CONTRACT"] = "lds"
    try:
        lab_lds, kv_lds, _ = ctx.predict(33, 200, want_kvalues=True)
    finally:
        os.environ.pop("CUFE_CONTRACT", None)
    assert np.array_equal(kv_all[33:233], kv_lds) and np.array_equal(lab_all[33:233], lab_lds)
    # decryption recovers <x, y> (mod p, with p standing for 0)
    ip = (x.astype(np.int64) @ model.sv[:, :l].astype(np.int64).T) % 50241
    assert np.array_equal(kv_all.astype(np.int64) % 50241, ip)


def test_empty_and_ragged_ranges(gpu, keyed_ctx):
    """Zero rows encrypt and predict as no-ops; one unaligned row, the last
    rows of the store and the whole store agree; ranges outside the stored
    rows are errors."""
    from cufe_amd import CufeError

    ctx, _, _ = keyed_ctx
    rng = np.random.default_rng(51)
    l, R, S = 16, 70, 9
    model = make_model(gpu, rng, S, l)
    ctx.rows_reserve(R, l)
    ctx.encrypt_rows(np.zeros((0, l), np.uint32), h(b"g2"), h(b"g3"), nonce0=0)
    assert ctx.rows_count() == 0
    x = rng.integers(0, 5, (R, l)).astype(np.uint32)
    ctx.encrypt_rows(x, h(b"g2"), h(b"g3"), nonce0=0)
    assert ctx.rows_count() == R
    ctx.set_model(model)
    lab0, kv0, _ = ctx.predict(0, 0, want_kvalues=True)
    assert lab0.shape == (0,) and kv0.shape == (0, S)
    lab_all, kv_all, _ = ctx.predict(0, R, want_kvalues=True)
    for r0, n in ((37, 1), (R - 1, 1), (64, R - 64), (31, 2)):
        lab, kv, _ = ctx.predict(r0, n, want_kvalues=True)
        assert np.array_equal(lab_all[r0:r0 + n], lab) and np.array_equal(kv_all[r0:r0 + n], kv), (r0, n)
    for r0, n in ((0, R + 1), (R, 1), (-1, 1)):
        with pytest.raises(CufeError):
            ctx.predict(r0, n)
    ip = (x.astype(np.int64) @ model.sv[:, :l].astype(np.int64).T) % 50241
    assert np.array_equal(kv_all.astype(np.int64) % 50241, ip)


def test_predict_a1a_rows(gpu, po, keyed_ctx):
    """Real data: the reference's a1a model (591 SVs) on 40 of its test rows;
    labels equal the plaintext SVM wherever the decryption is exact."""
    from cufe_amd import svm_io

    ctx, mpk, msk = keyed_ctx
    model = svm_io.read_model(os.path.join(HERE, "golden", "a1a.txt.model"))
    ytrue, xd = svm_io.read_problem(os.path.join(HERE, "golden", "a1a_rows101_400.t"), limit=40)
    l = 124
    x = xd[:, :l]
    note(f"a1a rows={x.shape[0]} svs={model.l}")
    ctx.rows_reserve(x.shape[0], l)
    ctx.encrypt_rows(x, h(b"a2"), h(b"a3"), nonce0=0)
    ctx.set_model(model)
    labels, kv, _ = ctx.predict(0, x.shape[0], want_kvalues=True)
    ip = (x.astype(np.int64) @ model.sv[:, :l].astype(np.int64).T)
    # kvalue = <x, y> exactly, or p when the inner product is 0 and the noise is negative
    ok = (kv.astype(np.int64) == ip) | ((ip == 0) & (kv.astype(np.int64) == 50241))
    assert ok.all()
    want = np.array([svm_io.decide(model, kv[r].astype(np.float64)) for r in range(x.shape[0])])
    assert np.array_equal(labels, want)
    # spot-check two rows pair by pair against the oracle
    mpk_l = np.concatenate([mpk[:l], mpk[785:786]])
    for r in (0, 17):
        c_full = po.encrypt(x[r], mpk_l, h(b"a2"), h(b"a3"), r)
        for v in (0, 1, 290, 590):
            k, _ = oracle_pair(po, c_full, model.sv[v], msk, l)
            assert int(kv[r, v]) == k


def test_predict_key_without_trailing_zero_slot(gpu, po):
    """Key length = last feature + 1: the reference's fold can end on the
    representative q, so the sum part runs as a separate exact fold."""
    rng = np.random.default_rng(70)
    l, R, S = 12, 3, 10
    ctx = gpu.Context()
    ctx.setup(h(b"uA"), h(b"uS"), h(b"uE"), l)
    mpk, msk = ctx.get_keys()
    model = make_model(gpu, rng, S, l)
    model.sv[0, l - 1] = 5  # feature in the last key slot
    model = gpu.SvmModel(model.nr_class, model.nSV, model.label, model.sv_coef, model.rho, model.sv)
    x = rng.integers(0, 5, (R, l)).astype(np.uint32)
    ctx.rows_reserve(R, l)
    ctx.encrypt_rows(x, h(b"u2"), h(b"u3"), nonce0=0)
    ctx.set_model(model)
    labels, kv, _ = ctx.predict(0, R, want_kvalues=True)
    for r in range(R):
        c_full = po.encrypt(x[r], mpk, h(b"u2"), h(b"u3"), r)
        for v in range(S):
            sk = po.keygen(model.sv[v, :l], msk)
            d = po.decrypt(c_full, model.sv[v, :l], sk)
            assert int(kv[r, v]) == po.crt_round(int(d[0, 0]), int(d[1, 0]), int(d[2, 0]))
    ctx.close()
