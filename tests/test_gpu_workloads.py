"""Whole workloads on the HIP path against the CPU oracle:

* the CPU configuration of the benchmark (svm_predict3, main.c:137-267):
  1,000 synthetic sparse binary 123-dim rows x 600 support vectors with integer
  coefficients, rho = 0.5;
* the reference's own a1a model on all 300 committed test rows;
* a hand-computed three-class model (the one-vs-one vote of main.c:338-389 with
  differing coefficient rows);
* uneven-length batched encryption with lengths up to the full key length 785.

Key values are compared with the oracle's coefficient-0 decryption
(orc_kvalues_coef0, pinned against the full decryption in test_oracle.py) on
the GPU's own ciphertexts and functional keys, which are themselves checked
against the oracle's encryption and keygen on sampled rows and keys; labels are
compared with the oracle's vote (orc_svm_decide) on the GPU's key values."""
import hashlib
import os
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

N = 4096
P = 50241
HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, HERE)


def h(t: bytes) -> bytes:
    return hashlib.sha256(t).digest()


def note(msg):
    print(f"[gpu-workloads] {msg}", flush=True)


@pytest.fixture(scope="module")
def keyed_ctx(gpu):
    note("ctx setup l=785")
    ctx = gpu.Context()
    ctx.setup(h(b"wA"), h(b"wS"), h(b"wE"), 785)
    mpk, msk = ctx.get_keys()
    yield ctx, mpk, msk
    ctx.close()


def sparse_binary(rng, n, dim=123, ones=14, width=124):
    x = np.zeros((n, width), np.uint32)
    for r in range(n):
        x[r, 1 + rng.choice(dim, ones, replace=False)] = 1
    return x


def decide_all(po, model, kv):
    nSV = np.array(model.nSV, np.int32)
    label = np.array(model.label, np.int32)
    coef = np.ascontiguousarray(model.sv_coef.ravel())
    rho = np.ascontiguousarray(model.rho, np.float64)
    return np.array([po.orc().orc_svm_decide(model.nr_class, model.l, nSV, label, coef, rho,
                                             np.ascontiguousarray(kv[r], np.float64), 0)
                     for r in range(kv.shape[0])])


def check_rows_against_oracle(po, ctx, mpk, x, rows, seed2, seed3, l):
    """Stored rows equal the oracle's encryption (c_L and coefficient 0)."""
    mpk_l = np.ascontiguousarray(np.concatenate([mpk[:l], mpk[785:786]]))
    for r in rows:
        c_full = po.encrypt(x[r], mpk_l, seed2, seed3, r)
        cl, c0 = ctx.get_row(r)
        assert np.array_equal(cl, c_full[l]), r
        assert np.array_equal(c0, c_full[:l, :, 0]), r


def test_cpu_config_1000_rows_600_svs(gpu, po, keyed_ctx):
    """1,000 rows x 600 SVs (nSV 300/300, coefficients U{-16..16}\\{0},
    rho 0.5): every pair decrypts to <x, y> (p standing for 0), 120,000 pairs
    equal the oracle's key values bit for bit, 2 pairs equal the oracle's whole
    svm_predict_cpu restatement, and every label equals the oracle's vote."""
    ctx, mpk, msk = keyed_ctx
    rng = np.random.default_rng(20261203)
    R, S, l = 1000, 600, 124
    x = sparse_binary(rng, R)
    sv = np.zeros((S, 785), np.uint32)
    sv[:, :l] = sparse_binary(rng, S)
    coef = rng.choice(np.concatenate([np.arange(-16, 0), np.arange(1, 17)]), S).astype(np.float64)[None, :]
    model = gpu.SvmModel(2, [300, 300], [1, -1], coef, [0.5], sv)
    note(f"C1 workload: encrypt {R} rows, keygen {S} SVs, predict")
    ctx.rows_reserve(R, l)
    ctx.encrypt_rows(x, h(b"c1-2"), h(b"c1-3"), nonce0=0)
    ctx.set_model(model)
    labels, kv, _ = ctx.predict(0, R, want_kvalues=True)
    ip = x.astype(np.int64) @ sv[:, :l].astype(np.int64).T
    kv64 = kv.astype(np.int64)
    assert ((kv64 == ip) | ((ip == 0) & (kv64 == P))).all()
    note("C1: rows and keys against the oracle")
    check_rows_against_oracle(po, ctx, mpk, x, (0, 511, 999), h(b"c1-2"), h(b"c1-3"), l)
    sk = ctx.model_keys()
    for v in (0, 299, 300, 599):
        assert np.array_equal(sk[v], po.keygen(sv[v], msk)), v
    note("C1: key values of 200 rows x 600 SVs against the oracle")
    rows = np.arange(0, R, 5)
    cl, c0 = ctx.export_rows(0, R)
    want = po.kvalues_coef0(cl[rows], c0[rows], sv, sk)
    assert np.array_equal(kv[rows], want)
    # the oracle's whole per-row path (dense keygen, full decryption, CRT of
    # every coefficient) on two pairs
    mpk_l = np.ascontiguousarray(np.concatenate([mpk[:l], mpk[785:786]]))
    c_pad = np.zeros((786, 3, N), np.uint32)
    c_full = po.encrypt(x[7], mpk_l, h(b"c1-2"), h(b"c1-3"), 7)
    c_pad[:l] = c_full[:l]
    c_pad[785] = c_full[l]
    kvo = np.empty(2)
    po.orc().orc_svm_predict_cpu(c_pad, 785, msk, 2, 2, np.array([1, 1], np.int32), np.array([1, -1], np.int32),
                                 np.array([1.0, -1.0]), np.array([0.0]), np.ascontiguousarray(sv[[3, 404]]), kvo, 0)
    assert kvo[0] == kv[7, 3] and kvo[1] == kv[7, 404]
    assert np.array_equal(labels, decide_all(po, model, kv))
    assert len(set(labels.tolist())) == 2


def test_a1a_all_rows_against_oracle(gpu, po, keyed_ctx):
    """The reference's a1a model (591 SVs, real coefficients) on all 300
    committed test rows: every key value equals the oracle's, and every label
    equals the oracle's vote on those key values."""
    from cufe_amd import svm_io

    ctx, mpk, msk = keyed_ctx
    model = svm_io.read_model(os.path.join(HERE, "golden", "a1a.txt.model"))
    _, xd = svm_io.read_problem(os.path.join(HERE, "golden", "a1a_rows101_400.t"))
    l = 124
    x = np.ascontiguousarray(xd[:, :l])
    R = x.shape[0]
    assert R == 300
    note(f"a1a: {R} rows x {model.l} SVs")
    ctx.rows_reserve(R, l)
    ctx.encrypt_rows(x, h(b"a1a-2"), h(b"a1a-3"), nonce0=0)
    ctx.set_model(model)
    labels, kv, _ = ctx.predict(0, R, want_kvalues=True)
    check_rows_against_oracle(po, ctx, mpk, x, (0, 299), h(b"a1a-2"), h(b"a1a-3"), l)
    sk = ctx.model_keys()
    for v in (0, 590):
        assert np.array_equal(sk[v], po.keygen(model.sv[v], msk)), v
    cl, c0 = ctx.export_rows(0, R)
    want = po.kvalues_coef0(cl, c0, model.sv, sk)
    assert np.array_equal(kv, want)
    assert np.array_equal(labels, decide_all(po, model, kv))
    ip = x.astype(np.int64) @ model.sv[:, :l].astype(np.int64).T
    assert ((kv.astype(np.int64) == ip) | ((ip == 0) & (kv.astype(np.int64) == P))).all()


def test_three_class_vote_by_hand(gpu, po, keyed_ctx):
    """The hand-computed three-class model of test_oracle.py on rows whose key
    values are set by construction: SV v is the unit vector of slot v + 1, so
    kvalue v = x[v + 1].  The batched path and the plain per-call path
    (svm_predict_gpu) both give the hand-computed labels."""
    from test_oracle import VOTE3, vote3_by_hand

    ctx, mpk, msk = keyed_ctx
    l = 8
    sv = np.zeros((5, 785), np.uint32)
    for v in range(5):
        sv[v, v + 1] = 1
    model = gpu.SvmModel(3, VOTE3["nSV"], VOTE3["label"], np.array(VOTE3["coef"]), VOTE3["rho"], sv)
    cases = [([1, 2, 3, 4, 5], 10), ([1, 1, 4, 1, 1], 20), ([4, 1, 1, 1, 1], 30)]
    x = np.zeros((len(cases), l), np.uint32)
    for r, (kv, want) in enumerate(cases):
        x[r, 1:6] = kv
        assert vote3_by_hand(kv) == want
    ctx.rows_reserve(len(cases), l)
    ctx.encrypt_rows(x, h(b"v3-2"), h(b"v3-3"), nonce0=0)
    ctx.set_model(model)
    labels, kv, _ = ctx.predict(0, len(cases), want_kvalues=True)
    assert np.array_equal(kv, x[:, 1:6])
    assert labels.tolist() == [w for _, w in cases]
    assert np.array_equal(labels, decide_all(po, model, kv))
    # the plain per-call form on full ciphertexts of the same messages
    xf = np.zeros((len(cases), 785), np.uint32)
    xf[:, :l] = x
    c = gpu.encrypt(xf, mpk, h(b"v3p-2"), h(b"v3p-3"))
    for r, (_, want) in enumerate(cases):
        lab, kvp = gpu.svm_predict_plain(model, c[r], msk)
        assert lab == want and np.array_equal(kvp, x[r, 1:6].astype(np.float64))


def test_uneven_batch_up_to_full_length(gpu, po, keyed_ctx):
    """Batched encryption of messages of lengths 16, 400 and 785 (the key
    length) equals the oracle's encryption of each message at its length."""
    ctx, mpk, _ = keyed_ctx
    lens = [16, 400, 785]
    rng = np.random.default_rng(785)
    xs = [rng.integers(0, 5, n).astype(np.uint32) for n in lens]
    note(f"encbatch lens={lens}")
    eb = gpu.EncBatch(ctx, lens)
    eb.run(np.concatenate(xs), h(b"u785-2"), h(b"u785-3"), nonce0=40)
    cts = eb.export(0, len(lens))
    eb.close()
    for b, (xb, c) in enumerate(zip(xs, cts)):
        n = len(xb)
        mpk_b = np.ascontiguousarray(np.concatenate([mpk[:n], mpk[785:786]]))
        want = po.encrypt(xb, mpk_b, h(b"u785-2"), h(b"u785-3"), 40 + b)
        assert np.array_equal(c, want), f"message {b} (l={n})"
