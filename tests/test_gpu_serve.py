"""Continuous-batching server: a request returns the same labels and key values
alone, mixed with other models' requests, one request per batch, and under an
open-loop Poisson replay -- equal, bit for bit, to cufe_ctx_predict on the same
stored rows (which the predict tests pin to the oracle pair by pair).

Requests larger than a batch slot are split into slot-sized parts and return
the same bits.  Also the server's error paths, forced from the host: unknown
model, rows too short for the model, models added after start, and an HBM
budget below what the models need."""
import hashlib

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

N = 4096


def h(t: bytes) -> bytes:
    return hashlib.sha256(t).digest()


def note(msg):
    print(f"[gpu-serve] {msg}", flush=True)


def make_model(gpu, rng, S, dim, nr_class, integer_coef, last_feature=None):
    sv = np.zeros((S, 785), np.uint32)
    for v in range(S):
        idx = rng.choice(np.arange(1, dim), size=6, replace=False)
        sv[v, idx] = rng.integers(1, 17, 6)
    if last_feature is not None:
        sv[0, last_feature] = 3
    base = S // nr_class
    nSV = [base] * (nr_class - 1) + [S - base * (nr_class - 1)]
    if integer_coef:
        coef = rng.integers(-16, 17, (nr_class - 1, S)).astype(np.float64)
        rho = np.full(nr_class * (nr_class - 1) // 2, 0.5)
    else:
        coef = rng.normal(0, 1, (nr_class - 1, S))
        rho = rng.normal(0, 10, nr_class * (nr_class - 1) // 2)
    return gpu.SvmModel(nr_class, nSV, list(range(1, nr_class + 1)), coef, rho, sv)


@pytest.fixture(scope="module")
def env(gpu):
    """Keys at L = 785; three row pools (dims 124, 300, 785) encrypted and
    exported in compact form; four models with context-computed references."""
    note("keys l=785")
    keys = gpu.Context()
    keys.setup(h(b"sA"), h(b"sS"), h(b"sE"), 785)
    mpk, msk = keys.get_keys()
    rng = np.random.default_rng(2024)
    specs = [  # (S, dim of rows, nr_class, integer coef, last feature)
        (40, 124, 2, True, None),
        (25, 300, 3, False, None),
        (60, 124, 2, True, None),
        (12, 785, 2, True, 784),  # uses slot 784: the unfused key path
    ]
    pools = {}
    for dim, rows in ((124, 96), (300, 80), (785, 24)):
        note(f"pool dim={dim} rows={rows}")
        c = gpu.Context()
        c.load_keys(mpk, msk)
        c.rows_reserve(rows, dim)
        x = rng.integers(0, 5, (rows, dim)).astype(np.uint32)
        c.encrypt_rows(x, h(b"r2%d" % dim), h(b"r3%d" % dim), nonce0=11)
        cl, c0 = c.export_rows(0, rows)
        pools[dim] = (c, cl, c0)
    models, refs = [], []
    for S, dim, nc, ic, lf in specs:
        m = make_model(gpu, rng, S, dim, nc, ic, lf)
        c = pools[dim][0]
        c.set_model(m)
        lab, kv, _ = c.predict(0, pools[dim][1].shape[0], want_kvalues=True)
        models.append((m, dim))
        refs.append((lab, kv))
    return keys, pools, models, refs


def start_server(gpu, env, **kw):
    keys, pools, models, refs = env
    srv = gpu.Server(keys, **kw)
    for m, _ in models:
        srv.add_model(m)
    srv.start()
    return srv


def check(env, mid, r0, n, labels, kv):
    _, _, _, refs = env
    lab_ref, kv_ref = refs[mid]
    assert np.array_equal(labels, lab_ref[r0:r0 + n]), f"labels differ, model {mid} rows {r0}+{n}"
    if kv is not None:
        assert np.array_equal(kv, kv_ref[r0:r0 + n]), f"key values differ, model {mid} rows {r0}+{n}"


def request(env, mid, r0, n):
    _, pools, models, _ = env
    dim = models[mid][1]
    _, cl, c0 = pools[dim]
    return cl[r0:r0 + n], c0[r0:r0 + n]


def test_alone_then_mixed(gpu, env):
    srv = start_server(gpu, env, max_batch_rows=512)
    try:
        _, pools, models, _ = env
        note("alone")
        for mid, r0, n in ((0, 3, 7), (1, 0, 64), (2, 5, 1), (3, 0, 24)):
            cl, c0 = request(env, mid, r0, n)
            t = srv.submit(mid, cl, c0, want_kvalues=True)
            lab, kv, lat = srv.wait(t, 60000)
            check(env, mid, r0, n, lab, kv)
        note("mixed")
        rng = np.random.default_rng(7)
        reqs = []
        for _ in range(60):
            mid = int(rng.integers(0, len(models)))
            pool_rows = pools[models[mid][1]][1].shape[0]
            n = int(rng.integers(1, min(64, pool_rows) + 1))
            r0 = int(rng.integers(0, pool_rows - n + 1))
            cl, c0 = request(env, mid, r0, n)
            reqs.append((srv.submit(mid, cl, c0, want_kvalues=bool(rng.integers(0, 2))), mid, r0, n))
        for t, mid, r0, n in reqs:
            lab, kv, _ = srv.wait(t, 60000)
            check(env, mid, r0, n, lab, kv)
        st = srv.stats()
        note(f"stats {st}")
        assert st["requests"] == 64 and st["failed"] == 0
        assert st["batches"] < st["requests"], "no two requests shared a batch"
    finally:
        srv.close()


def test_one_request_per_batch(gpu, env):
    srv = start_server(gpu, env, max_batch_rows=128, batching=False)
    try:
        reqs = []
        for mid, r0, n in ((1, 10, 33), (0, 0, 64), (3, 4, 9), (2, 90, 6)):
            cl, c0 = request(env, mid, r0, n)
            reqs.append((srv.submit(mid, cl, c0, want_kvalues=True), mid, r0, n))
        for t, mid, r0, n in reqs:
            lab, kv, _ = srv.wait(t, 60000)
            check(env, mid, r0, n, lab, kv)
        st = srv.stats()
        assert st["batches"] == st["requests"] == 4
    finally:
        srv.close()


def test_pinned_requests(gpu, env):
    """Rows in pinned host memory are copied by DMA directly: same results."""
    srv = start_server(gpu, env, max_batch_rows=256)
    try:
        reqs = []
        for mid, r0, n in ((0, 5, 40), (1, 0, 17), (3, 2, 20), (2, 30, 64)):
            cl, c0 = request(env, mid, r0, n)
            pcl, pc0 = gpu.host_empty(cl.shape), gpu.host_empty(c0.shape)
            pcl[...] = cl
            pc0[...] = c0
            reqs.append((srv.submit(mid, pcl, pc0, want_kvalues=True), mid, r0, n))
        for t, mid, r0, n in reqs:
            lab, kv, _ = srv.wait(t, 60000)
            check(env, mid, r0, n, lab, kv)
    finally:
        srv.close()


def test_wire_requests(gpu, env):
    """Requests packed in the wire format and parsed back are served from the
    blob's own memory with the same results."""
    srv = start_server(gpu, env, max_batch_rows=256)
    try:
        reqs = []
        for mid, r0, n in ((0, 3, 20), (1, 40, 33), (3, 0, 24)):
            cl, c0 = request(env, mid, r0, n)
            p = gpu.wire_parse(gpu.wire_pack(gpu.WIRE_ROWS, cl, c0, tag=mid))
            assert p["tag"] == mid and np.array_equal(p["cl"], cl)
            reqs.append((srv.submit(mid, p["cl"], p["c0"], want_kvalues=True), mid, r0, n, p))
        for t, mid, r0, n, _ in reqs:
            lab, kv, _ = srv.wait(t, 60000)
            check(env, mid, r0, n, lab, kv)
    finally:
        srv.close()


def test_poisson_replay(gpu, env):
    srv = start_server(gpu, env, max_batch_rows=1024)
    try:
        _, pools, models, _ = env
        rng = np.random.default_rng(99)
        n = 120
        arrival = np.cumsum(rng.exponential(1 / 2000.0, n))
        mids, rows, spans = [], [], []
        for _ in range(n):
            mid = int(rng.integers(0, len(models)))
            pool_rows = pools[models[mid][1]][1].shape[0]
            k = int(rng.integers(1, min(64, pool_rows) + 1))
            r0 = int(rng.integers(0, pool_rows - k + 1))
            mids.append(mid)
            rows.append(request(env, mid, r0, k))
            spans.append((r0, k))
        note(f"replay {n} requests")
        labels, lat, makespan, failed = srv.replay(arrival, mids, rows)
        assert failed == 0
        for i in range(n):
            check(env, mids[i], spans[i][0], spans[i][1], labels[i], None)
        assert np.all(lat > 0) and makespan > 0
        note(f"p50 {np.percentile(lat, 50):.3f} ms p99 {np.percentile(lat, 99):.3f} ms makespan {makespan:.3f} s")
    finally:
        srv.close()


def test_server_errors(gpu, env):
    from cufe_amd import CufeError

    keys, pools, models, _ = env
    srv = gpu.Server(keys, max_batch_rows=64)
    srv.add_model(models[0][0])
    with pytest.raises(CufeError):
        srv.submit(0, *request(env, 0, 0, 1))  # not started
    srv.start()
    try:
        with pytest.raises(CufeError):
            srv.add_model(models[1][0])
        with pytest.raises(CufeError):
            srv.submit(5, *request(env, 0, 0, 1))
        cl, c0 = request(env, 0, 0, 2)
        with pytest.raises(CufeError):
            srv.submit(0, cl, np.ascontiguousarray(c0[:, :3]))  # 3 slots < the model's features
        t = srv.submit(0, cl, c0)
        assert srv.wait(t, 60000) is not None
    finally:
        srv.close()
    small = gpu.Server(keys, hbm_budget=1 << 20)
    small.add_model(models[0][0])
    with pytest.raises(CufeError):
        small.start()
    small.close()


def test_oversize_requests_are_split(gpu, env):
    """With 64-row slots, requests of 96 and 80 rows run as slot-sized parts
    (64 + 32, 64 + 16), interleaved with small requests of other models; every
    request equals cufe_ctx_predict, labels and key values."""
    _, pools, models, _ = env
    srv = start_server(gpu, env, max_batch_rows=64)
    try:
        info = srv.info()
        note(f"server info {info}")
        assert 2 <= info["slots"] <= 4 and srv.stats()["cap_rows"] == 64
        plan = [(0, 0, 96), (2, 3, 5), (1, 0, 80), (3, 1, 20), (0, 10, 64), (2, 0, 96)]
        reqs = []
        for mid, r0, n in plan:
            cl, c0 = request(env, mid, r0, n)
            reqs.append((srv.submit(mid, cl, c0, want_kvalues=True), mid, r0, n))
        for t, mid, r0, n in reqs:
            lab, kv, _ = srv.wait(t, 60000)
            check(env, mid, r0, n, lab, kv)
        st = srv.stats()
        assert st["requests"] == len(plan) and st["failed"] == 0 and st["max_batch_rows"] <= 64
        info = srv.info()
        assert info["split_requests"] == 3
        assert 0 < info["peak_inflight_bytes"] <= info["slots"] * 64 * info["bytes_per_row"]
    finally:
        srv.close()
