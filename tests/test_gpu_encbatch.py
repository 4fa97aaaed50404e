"""Batched encryption of uneven-length messages on the GPU against the oracle
(rlwe_sife_encrypt, src/rlwe_sife.c:112-181, per message with its own length
and nonce), and the captured-graph replay against the eager run: several
replays, with eager runs of other seeds in between, each checked word for word."""
import hashlib

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

N = 4096


def h(t: bytes) -> bytes:
    return hashlib.sha256(t).digest()


def note(msg):
    print(f"[gpu-encbatch] {msg}", flush=True)


@pytest.fixture(scope="module")
def keyed(gpu):
    ctx = gpu.Context()
    ctx.setup(h(b"ebA"), h(b"ebS"), h(b"ebE"), 64)
    mpk, msk = ctx.get_keys()
    return ctx, mpk


def messages(rng, lens):
    return [rng.integers(0, 5, int(l)).astype(np.uint32) for l in lens]


def test_uneven_batch_matches_oracle(gpu, po, keyed):
    ctx, mpk = keyed
    lens = [1, 5, 17, 3, 40, 2, 64]
    rng = np.random.default_rng(5)
    xs = messages(rng, lens)
    note(f"encbatch lens={lens}")
    eb = gpu.EncBatch(ctx, lens)
    eb.run(np.concatenate(xs), h(b"b2"), h(b"b3"), nonce0=100)
    cts = eb.export(0, len(lens))
    for b, (x, c) in enumerate(zip(xs, cts)):
        l = len(x)
        mpk_b = np.ascontiguousarray(np.concatenate([mpk[:l], mpk[64:65]]))
        want = po.encrypt(x, mpk_b, h(b"b2"), h(b"b3"), 100 + b)
        assert np.array_equal(c, want), f"message {b} (l={l}) differs from the oracle"
    eb.close()


def test_graph_replays_match_eager(gpu, keyed):
    ctx, _ = keyed
    rng = np.random.default_rng(9)
    lens = rng.integers(1, 65, 24)
    x1 = np.concatenate(messages(rng, lens))
    x2 = np.concatenate(messages(rng, lens))
    eb = gpu.EncBatch(ctx, lens)
    note("eager reference runs")
    eb.run(x1, h(b"g2"), h(b"g3"), nonce0=0, use_graph=False)
    ref1 = [c.copy() for c in eb.export(0, len(lens))]
    sum1, info = eb.checksum()
    eb.run(x2, h(b"k2"), h(b"k3"), nonce0=500, use_graph=False)
    ref2 = [c.copy() for c in eb.export(0, len(lens))]
    sum2, _ = eb.checksum()
    assert sum1 != sum2
    for rep in range(3):
        note(f"graph replay {rep}")
        eb.run(x1, h(b"g2"), h(b"g3"), nonce0=0, use_graph=True)
        got = eb.export(0, len(lens))
        assert all(np.array_equal(a, b) for a, b in zip(got, ref1)), f"replay {rep} (seeds 1) differs"
        assert eb.checksum()[0] == sum1
        eb.run(x2, h(b"k2"), h(b"k3"), nonce0=500, use_graph=True)
        got = eb.export(0, len(lens))
        assert all(np.array_equal(a, b) for a, b in zip(got, ref2)), f"replay {rep} (seeds 2) differs"
        eb.run(x1, h(b"g2"), h(b"g3"), nonce0=0, use_graph=False)  # eager work between replays
        assert eb.checksum()[0] == sum1
    note(f"info {eb.checksum()[1]}")
    eb.close()


def test_encbatch_errors(gpu, keyed):
    from cufe_amd import CufeError

    ctx, _ = keyed
    with pytest.raises(CufeError):
        gpu.EncBatch(ctx, [3, 65])  # longer than the key
    with pytest.raises(CufeError):
        gpu.EncBatch(ctx, [0])
    eb = gpu.EncBatch(ctx, [2, 2])
    with pytest.raises(CufeError):
        eb.run(np.zeros(4, np.uint32), h(b"a"), h(b"b"), nonce0=65535)  # nonce range overflows 16 bits
    eb.close()
