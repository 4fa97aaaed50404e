"""IPFE-SVM prediction benchmark on one MI355X.

Workload (the batched-SVM configuration): synthetic sparse binary rows of
dimension 123 (about 14 ones per row at uniform positions from a seed), encrypted
with the reference parameter set (N = 4096, three RNS limbs, key length 785)
into the device-resident row store; a linear SVM with 2,000 support vectors
drawn the same way, integer coefficients in [-16, 16] \\ {0}, rho = 0.5.

A step is one batched prediction pass over every resident row: the modular
coefficient-0 contraction against all functional keys, CRT + rounding of each
inner product, and the one-vs-one vote, with labels copied back to the host.
Ciphertexts and keys are in HBM before the timed region starts.  Encryption of
the rows and functional-key generation are timed separately and reported.

Extra fields: the plain per-call GPU form (the four reference C-ABI calls per
row, svm_predict_gpu), the roofline of the dominant kernel (the int8
matrix-core contraction, timed with HIP events on its stream; the fp64
contraction when keys need it), and svm_predict_cpu on a bounded sample: the
reference's own C path compiled from its sources (oracle/_ref) where that build
is present, else the oracle's restatement of it.
"""
from __future__ import annotations

import argparse
import ctypes as C
import hashlib
import json
import os
import re
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

FP64_PEAK_TFLOPS = 78.6  # MI355X published fp64 rate
P_MOD = 50241  # plaintext modulus of the reference parameter set
INT8_PEAK_TOPS = 5000.0  # MI355X published dense int8 matrix rate


def h(t: bytes) -> bytes:
    return hashlib.sha256(t).digest()


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def sparse_binary_rows(rng, n, dim=123, ones=14, width=124):
    x = np.zeros((n, width), np.uint32)
    for r in range(n):
        x[r, 1 + rng.choice(dim, ones, replace=False)] = 1
    return x


def _svm_predict_cpu_fn():
    """svm_predict_cpu (main.c:137-267) for one ciphertext against k SVs: the
    reference's own compiled C path (oracle/_ref/libref.so, built from the
    reference sources with GMP) when it is present, else the oracle's
    restatement.  Returns (kind, fn(c, msk, sv))."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import pyoracle as po

    ref = po.ref()

    def run(c, msk, sv):
        k = sv.shape[0]
        nSV, lab = np.array([k, 0], np.int32), np.array([1, -1], np.int32)
        coef, rho = np.ascontiguousarray(np.ones(k)), np.array([0.0])
        if ref is not None:
            return ref.refh_svm_predict_cpu(c, msk, 2, k, nSV, lab, coef, rho, sv)
        return po.orc().orc_svm_predict_cpu(c, msk.shape[0], msk, 2, k, nSV, lab, coef, rho, sv, np.empty(k), 0)

    return ("reference" if ref is not None else "port"), run


def cpu_baseline(cf, ctx, model, x_row, seconds: float):
    """svm_predict_cpu on one core, on blocks of 4 support vectors of one row
    (keygen, full decryption with GMP CRT, rounding of all 4096 coefficients
    per SV) for about `seconds`; extrapolated to a full prediction (all SVs)."""
    kind, run = _svm_predict_cpu_fn()
    mpk, msk = ctx.get_keys()
    L = msk.shape[0]
    xf = np.zeros((1, L), np.uint32)
    xf[0, : x_row.size] = x_row
    c = cf.encrypt(xf, mpk, h(b"cpu2"), h(b"cpu3"))[0]  # bit-identical to the CPU encryption
    S = model.l
    done = 0
    t0 = time.perf_counter()
    while done < S and time.perf_counter() - t0 < seconds:
        k = min(4, S - done)
        run(c, msk, np.ascontiguousarray(model.sv[done: done + k]))
        done += k
    dt = time.perf_counter() - t0
    per_pair = dt / done
    what = "reference build of svm_predict_cpu (gcc -O3 -DAVX2, GMP)" if kind == "reference" else \
        "oracle restatement of svm_predict_cpu"
    return {"value": 1.0 / (per_pair * S), "unit": "predictions/s", "cores": 1, "kind": kind,
            "sample": f"1 row x {done} of {S} SVs in {dt:.1f} s, {what}; {per_pair * 1e3:.1f} ms per (row, SV) "
                      f"pair, extrapolated to {S} SVs"}


def _cpu_worker(args):
    """One spawned process: times svm_predict_cpu on support vectors
    [v0, v0 + k) repeatedly for `seconds`; returns pairs done.  It loads only
    the CPU libraries (never the HIP runtime)."""
    path, v0, k, seconds = args
    _, run = _svm_predict_cpu_fn()
    d = np.load(path)
    c, msk, sv = d["c"], d["msk"], np.ascontiguousarray(d["sv"][v0:v0 + k])
    done = 0
    t0 = time.perf_counter()
    while time.perf_counter() - t0 < seconds:
        run(c, msk, sv)
        done += k
    return done, time.perf_counter() - t0


def cpu_baseline_all_cores(cf, ctx, model, x_row, seconds: float):
    """The same svm_predict_cpu on every host core the box grants (at most 16),
    one spawned process per core, each on its own slice of the support vectors."""
    import multiprocessing as mp
    import tempfile

    kind, _ = _svm_predict_cpu_fn()
    mpk, msk = ctx.get_keys()
    L = msk.shape[0]
    xf = np.zeros((1, L), np.uint32)
    xf[0, : x_row.size] = x_row
    c = cf.encrypt(xf, mpk, h(b"cpu2"), h(b"cpu3"))[0]
    cores = max(1, min(16, os.cpu_count() or 1))
    with tempfile.TemporaryDirectory() as tmp:
        path = os.path.join(tmp, "cpu_inputs.npz")
        np.savez(path, c=c, msk=msk, sv=model.sv)
        k = 2
        jobs = [(path, (i * k) % (model.l - k), k, seconds) for i in range(cores)]
        # close + join: the workers exit on their own once the map is done.
        # Leaving the `with` block first would call Pool.terminate(), which
        # signals every worker (under a profiler that signal aborts them).
        with mp.get_context("spawn").Pool(cores) as pool:
            res = pool.map(_cpu_worker, jobs)
            pool.close()
            pool.join()
    pairs = sum(r[0] for r in res)
    wall = max(r[1] for r in res)
    rate_pairs = pairs / wall
    return {"value": rate_pairs / model.l, "unit": "predictions/s", "cores": cores, "kind": kind,
            "sample": f"{cores} processes x {seconds:.0f} s, {pairs} (row, SV) pairs of svm_predict_cpu "
                      f"({rate_pairs:.1f} pairs/s), extrapolated to {model.l} SVs per prediction"}


def plain_gpu(cf, ctx, model, x, nrows: int):
    """The plain per-call form: rlwe_sife_encrypt_gpu per row is not timed; per row
    svm_predict_gpu = keygen_gpu + decrypt_gmp_gpu + host round_extract2 + vote."""
    mpk, msk = ctx.get_keys()
    L = msk.shape[0]
    times = []
    for r in range(nrows):
        xf = np.zeros((1, L), np.uint32)
        xf[0, : x.shape[1]] = x[r]
        c = cf.encrypt(xf, mpk, h(b"pl2"), h(b"pl3"), nonce0=r)[0]
        t0 = time.perf_counter()
        cf.svm_predict_plain(model, c, msk)
        times.append(time.perf_counter() - t0)
    return {"value": 1.0 / float(np.mean(times)), "unit": "predictions/s", "rows": nrows,
            "ms_per_prediction": 1e3 * float(np.mean(times))}


class ServingBench:
    """Open-loop Poisson replays through the continuous-batching server.  Four
    models (100, 600 and 2000 SVs over 123-dim rows; 1000 SVs over rows of a
    dimension drawn from U{16..785}); request sizes U{1..64} rows; model
    uniform per request.  Every served label is checked against
    cufe_ctx_predict on the same rows."""

    def __init__(self, cf, ctx, model, x, args, rng):
        self.cf, self.ctx, self.args = cf, ctx, args
        mpk, msk = ctx.get_keys()
        dim2 = int(rng.integers(16, 786))
        pool1 = min(1024, x.shape[0])
        cl1, c01 = ctx.export_rows(0, pool1)
        self.ctx2 = ctx2 = cf.Context()
        ctx2.load_keys(mpk, msk)
        pool2 = 512
        ctx2.rows_reserve(pool2, dim2)
        x2 = np.zeros((pool2, dim2), np.uint32)
        for r in range(pool2):
            x2[r, rng.choice(dim2, min(14, dim2), replace=False)] = 1
        ctx2.encrypt_rows(x2, h(b"serve-2"), h(b"serve-3"), nonce0=0)
        cl2, c02 = ctx2.export_rows(0, pool2)

        def pinned(a):
            # request rows live in pinned host memory, so the server DMAs them
            # straight from the caller's buffers
            p = cf.host_empty(a.shape, a.dtype)
            p[...] = a
            return p

        cl1, c01, cl2, c02 = pinned(cl1), pinned(c01), pinned(cl2), pinned(c02)

        def svm(S, dim, width):
            svx = np.zeros((S, width), np.uint32)
            for v in range(S):
                svx[v, rng.choice(dim, min(14, dim), replace=False)] = 1
            co = rng.choice(np.concatenate([np.arange(-16, 0), np.arange(1, 17)]), S).astype(np.float64)[None, :]
            return cf.SvmModel(2, [S // 2, S - S // 2], [1, -1], co, [0.5], svx)

        self.models = [(svm(100, 123, 124), 1), (svm(600, 123, 124), 1), (model, 1), (svm(1000, dim2, dim2), 2)]
        pools = {1: (ctx, cl1, c01, pool1), 2: (ctx2, cl2, c02, pool2)}
        self.refs = []
        for m, p in self.models:
            c = pools[p][0]
            c.set_model(m)
            self.refs.append(c.predict(0, pools[p][3])[0])
        ctx.set_model(model)

        n = args.serve_requests
        self.arrival = np.cumsum(rng.exponential(1.0 / args.serve_rate, n))
        self.mids, self.rows, self.spans = [], [], []
        for _ in range(n):
            mid = int(rng.integers(0, len(self.models)))
            _, cl, c0, pr = pools[self.models[mid][1]]
            k = int(rng.integers(1, 65))
            r0 = int(rng.integers(0, pr - k + 1))
            self.mids.append(mid)
            self.rows.append((cl[r0:r0 + k], c0[r0:r0 + k]))
            self.spans.append((r0, k))
        self.total_rows = int(sum(k for _, k in self.spans))
        self.info = {"arrival_rate_req_s": args.serve_rate, "requests": n, "rows": self.total_rows,
                     "models": [{"svs": m.l, "row_slots": int(pools[p][2].shape[1])} for m, p in self.models]}

    def run(self, mode: str, batching: bool, speed: float) -> dict:
        cf, n, args = self.cf, len(self.arrival), self.args
        arrival, mids, rows = self.arrival, self.mids, self.rows
        srv = cf.Server(self.ctx, max_batch_rows=4096, batching=batching)
        for m, _ in self.models:
            srv.add_model(m)
        srv.start()
        log(f"serving replay ({mode}): {n} requests, {self.total_rows} rows at {args.serve_rate * speed:.0f} req/s")
        # untimed warm-up: the first 200 requests once, so that one-time costs
        # of the process (first launches, first use of host buffers) stay out
        # of every mode's latencies alike
        w = min(200, n)
        _, _, _, wfailed = srv.replay(arrival[:w] / speed, mids[:w], rows[:w])
        if wfailed:
            raise RuntimeError(f"serving warm-up ({mode}): {wfailed} of {w} requests failed")
        srv.reset_stats()  # every stat below covers the timed replay only
        labels, lat, makespan, failed = srv.replay(arrival / speed, mids, rows)
        st = srv.stats()
        srv.close()
        bad = sum(int(not np.array_equal(labels[i], self.refs[mids[i]][self.spans[i][0]:self.spans[i][0] +
                                                                         self.spans[i][1]]))
                  for i in range(n))
        r = {"arrival_rate_req_s": args.serve_rate * speed, "p50_ms": float(np.percentile(lat, 50)),
             "p99_ms": float(np.percentile(lat, 99)), "p999_ms": float(np.percentile(lat, 99.9)),
             "mean_ms": float(np.mean(lat)), "throughput_req_s": n / makespan,
             "throughput_rows_s": self.total_rows / makespan, "makespan_s": makespan, "batches": st["batches"],
             "max_batch_rows": st["max_batch_rows"], "slot_cap_rows": st["cap_rows"], "failed": failed,
             "requests_differing_from_ctx_predict": bad}
        log(f"  p50 {r['p50_ms']:.2f} ms p99 {r['p99_ms']:.2f} ms {r['throughput_req_s']:.0f} req/s, "
            f"{st['batches']} batches, mismatches {bad}")
        return r

    def close(self):
        self.ctx2.close()


def serving(sb: ServingBench, args) -> dict:
    """Batched replay at the offered rate, the same arrivals with one request
    per batch, and both again at serve_burst times the rate."""
    res = dict(sb.info)
    runs = [("batched", True, 1), ("one_request_per_batch", False, 1)]
    if args.serve_burst > 1:
        runs += [(f"batched_x{args.serve_burst:g}", True, args.serve_burst),
                 (f"one_request_per_batch_x{args.serve_burst:g}", False, args.serve_burst)]
    for mode, batching, speed in runs:
        res[mode] = sb.run(mode, batching, speed)
    return res


def single_ops(cf, ctx, rng):
    """Single-operation kernels at the reference parameter set: forward and
    inverse NTT of 10,000 polynomials (30,000 limb transforms) on synthetic
    residues, 10,000 independent one-call Gaussian streams per sampler, one
    10,000-call S1 stream, and one encrypt / keygen / decrypt at l = 123 through
    the host-buffer entry points.  The NTT is measured against the butterfly
    rate of its own butterfly code held in registers (cufe_microbench(7)) and
    against HBM (32 KiB read + written per limb transform)."""
    bfly_peak = cf.microbench(7) * 1e9
    out = {"butterfly_peak_per_s": bfly_peak}
    for name, inv in (("ntt_forward", False), ("ntt_inverse", True)):
        ms = cf.time_ntt(10000, inv, 20)
        limbs = 30000 / (ms * 1e-3)
        bf = limbs * 24576
        out[name] = {"limb_transforms_per_s": limbs, "ms_per_10000_polys": ms, "butterflies_per_s": bf,
                     "frac_of_butterfly_peak": bf / bfly_peak, "hbm_gb_s": limbs * 32768 / 1e9,
                     "frac_of_hbm": limbs * 32768 / 8e12}
    for which in (1, 2, 3):
        ms = cf.time_gauss(10000, which, 2)
        out[f"gauss_S{which}_10000_streams"] = {"ms": ms, "polys_per_s": 10000 / (ms * 1e-3)}
    ms = cf.time_gauss_calls(1, 10000, 1, 1)
    out["gauss_S1_one_stream_10000_calls"] = {"ms": ms, "polys_per_s": 10000 / (ms * 1e-3)}
    mpk, msk = ctx.get_keys()
    l = 123
    mpk_l = np.ascontiguousarray(np.concatenate([mpk[:l], mpk[785:786]]))
    x = rng.integers(0, 2, (1, l)).astype(np.uint32)
    y = rng.integers(0, 2, (1, l)).astype(np.uint32)
    t0 = time.perf_counter()
    c = cf.encrypt(x, mpk_l, h(b"so2"), h(b"so3"))[0]
    t1 = time.perf_counter()
    sk = cf.keygen(y, msk[:l])
    t2 = time.perf_counter()
    d = cf.decrypt(c, y, sk)
    t3 = time.perf_counter()
    kv = cf.lib().cufe_round_extract2(np.ascontiguousarray(d[0]))
    out["scheme_l123_host_buffers_ms"] = {"encrypt": 1e3 * (t1 - t0), "keygen": 1e3 * (t2 - t1),
                                          "decrypt": 1e3 * (t3 - t2),
                                          "decrypted_equals_inner_product_mod_p":
                                              int(kv) % P_MOD == int(x[0].astype(np.int64) @ y[0].astype(np.int64))}
    out["parameter_sets"] = {str(pid): pset_single_ops(cf, pid, rng) for pid in (0, 1, 2)}
    out["ntt_autotune"] = cf.ntt_autotune_report().strip().split("\n")
    return out


def pset_single_ops(cf, pid, rng):
    """The single-operation kernels at one parameter set (0: N = 2048 / 2 limbs,
    1: the reference set, 2: N = 8192 / 4 limbs): forward and inverse NTT of
    10,000 polynomials with the autotuned shape (HBM: 8 N bytes per limb
    transform), 10,000 one-call Gaussian streams per sampler, and one setup /
    encrypt / keygen / decrypt + CRT at l = 123 on host buffers, the decryption
    checked against <x, y> on every coefficient."""
    ps = cf.PSet(pid)
    out = {"N": ps.N, "limbs": ps.nmod}
    for name, inv in (("ntt_forward", False), ("ntt_inverse", True)):
        ms = ps.time_ntt(10000, inv, 10)
        limbs = 10000 * ps.nmod / (ms * 1e-3)
        out[name] = {"limb_transforms_per_s": limbs, "ms_per_10000_polys": ms,
                     "hbm_gb_s": limbs * 8 * ps.N / 1e9, "frac_of_hbm": limbs * 8 * ps.N / 8e12}
    for which in (1, 2, 3):
        ms = ps.time_gauss(10000, which, 2)
        out[f"gauss_S{which}_10000_streams"] = {"ms": ms, "polys_per_s": 10000 / (ms * 1e-3)}
    l = 123
    a = ps.sample_polya(h(b"pso-a%d" % pid))
    t0 = time.perf_counter()
    mpk, msk = ps.setup(a, h(b"pso-s"), h(b"pso-e"), l)
    t1 = time.perf_counter()
    x = rng.integers(0, 5, (1, l)).astype(np.uint32)
    y = rng.integers(0, 17, (1, l)).astype(np.uint32)
    c = ps.encrypt(x, mpk, h(b"pso-2"), h(b"pso-3"))
    t2 = time.perf_counter()
    sk = ps.keygen(y, msk)
    t3 = time.perf_counter()
    d = ps.decrypt(c[0], y, sk)
    r = ps.crt_round(d, coef=-1)
    t4 = time.perf_counter()
    ip = int(x[0].astype(np.int64) @ y[0].astype(np.int64))
    out["scheme_l123_host_buffers_ms"] = {"setup": 1e3 * (t1 - t0), "encrypt": 1e3 * (t2 - t1),
                                          "keygen": 1e3 * (t3 - t2), "decrypt_and_crt": 1e3 * (t4 - t3),
                                          "every_coefficient_decrypts_to_inner_product":
                                              bool(np.all(r.astype(np.int64) % ps.p == ip % ps.p))}
    return out


def batched_encrypt(cf, ctx, args, rng):
    """Uneven-length batched encryption: n messages with l_b ~ U{16..785},
    entries U{0..4}, message b on nonce b; one batch call (eager), the same
    batch replayed as a captured hipGraph, and the plain reference ABI
    (rlwe_sife_encrypt_gpu, one call per message at L = 785) timed on a few
    messages.  Bit-exactness: message 0 of the batch equals the plain call's
    c_0..c_{l_0-1} and c_L; graph and eager runs have equal device checksums."""
    n = args.enc_messages
    lens = rng.integers(16, 786, n).astype(np.int32)
    x = rng.integers(0, 5, int(lens.sum())).astype(np.uint32)
    s2, s3 = h(b"enc-batch-2"), h(b"enc-batch-3")
    eb = cf.EncBatch(ctx, lens)
    log(f"batched encrypt: {n} messages, {int(lens.sum())} slots")
    ms_eager = eb.run(x, s2, s3, 0, use_graph=False)
    sum_eager, info = eb.checksum()
    t0 = time.perf_counter()
    eb.run(x, s2, s3, 0, use_graph=True)  # capture + first replay
    t_capture = time.perf_counter() - t0
    ms_graph = eb.run(x, s2, s3, 0, use_graph=True)
    sum_graph, info = eb.checksum()
    c0_batch = eb.export(0, 1)[0]
    eb.close()
    mpk, _ = ctx.get_keys()
    L = mpk.shape[0] - 1
    k = max(1, args.enc_plain)
    lib = cf.lib()
    c_full = np.empty((1, L + 1, 3, 4096), np.uint32)
    times = []
    for b in range(k):
        m = np.zeros((1, L), np.uint32)
        m[0, :lens[b]] = x[int(np.sum(lens[:b])):int(np.sum(lens[:b + 1]))]
        t1 = time.perf_counter()
        lib.rlwe_sife_encrypt_gpu(m, mpk, c_full, cf._seed(s2), cf._seed(s3), 1, None)
        times.append(time.perf_counter() - t1)
        if b == 0:
            l0 = int(lens[0])
            same = bool(np.array_equal(c0_batch[:l0], c_full[0, :l0]) and np.array_equal(c0_batch[l0], c_full[0, L]))
    per_call = float(np.mean(times))
    slots = int(lens.sum())
    polys = slots + 2 * n  # Gaussian polynomials sampled (f_i, plus r and f_0 per message)
    return {"messages": n, "lens": "U{16..785}", "slots": slots, "ciphertext_bytes": int(info["words"]) * 4,
            "eager_ms": ms_eager, "graph_ms": ms_graph, "graph_capture_and_first_replay_s": t_capture,
            "encryptions_per_s_graph": n / (ms_graph * 1e-3), "gaussian_polys_per_s": polys / (ms_graph * 1e-3),
            "sampler_groups": info["groups"], "groups_redone": info["redone"],
            "graph_equals_eager": bool(sum_eager == sum_graph),
            "plain_abi": {"calls_timed": k, "ms_per_call": per_call * 1e3,
                          "extrapolated_s_for_batch": per_call * n,
                          "note": "rlwe_sife_encrypt_gpu at L=785 per message, mpk upload and ciphertext "
                                  "download included as in the reference ABI"},
            "batched_vs_plain": per_call * n / (ms_graph * 1e-3),
            "message0_equals_plain_abi_prefix": same}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--rows", type=int, default=100_000)
    ap.add_argument("--svs", type=int, default=2_000)
    ap.add_argument("--plain-rows", type=int, default=2)
    ap.add_argument("--cpu-seconds", type=float, default=12.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--serve-requests", type=int, default=4000)
    ap.add_argument("--enc-messages", type=int, default=4096)
    ap.add_argument("--enc-plain", type=int, default=4)
    ap.add_argument("--no-single-ops", action="store_true")
    ap.add_argument("--serve-rate", type=float, default=2000.0)
    ap.add_argument("--copy-calibration", action="store_true",
                    help="run a 1 GiB device copy first (known bytes, to scale HBM counters)")
    ap.add_argument("--serve-burst", type=float, default=10.0,
                    help="also replay the same requests this many times faster")
    ap.add_argument("--detail", default=os.path.join("gpurun_out", "bench_detail.json"),
                    help="file that receives the full results (the stdout line is a summary)")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py measures one GPU: there is no multi-GPU path (rows would shard across GPUs, each holding "
              "every functional key)", file=sys.stderr)
        sys.exit(2)

    import cufe_amd as cf

    info = cf.device_info()
    log(f"device {info}")
    if args.copy_calibration:
        log(f"copy calibration kernel: 1 GiB read + 1 GiB written per launch, {cf.microbench(2):.0f} GB/s")
    R, S = args.rows, args.svs
    rng = np.random.default_rng(20261202)
    x = sparse_binary_rows(rng, R)
    sv = sparse_binary_rows(rng, S)
    coef = rng.choice(np.concatenate([np.arange(-16, 0), np.arange(1, 17)]), S).astype(np.float64)[None, :]
    model = cf.SvmModel(2, [S // 2, S - S // 2], [1, -1], coef, [0.5], sv)

    ctx = cf.Context()
    t0 = time.perf_counter()
    ctx.setup(h(b"bench-A"), h(b"bench-S"), h(b"bench-E"), 785)
    t_setup = time.perf_counter() - t0
    log(f"setup L=785 {t_setup:.2f}s")
    l = x.shape[1]
    ctx.rows_reserve(R, l)
    t0 = time.perf_counter()
    enc_ms = 0.0
    for b0 in range(0, R, 65536):  # 16-bit nonces: one seed pair per 65536 rows
        b1 = min(R, b0 + 65536)
        enc_ms += ctx.encrypt_rows(x[b0:b1], h(b"bench-2-%d" % b0), h(b"bench-3-%d" % b0), nonce0=0)
        log(f"encrypted {b1}/{R} rows, {time.perf_counter() - t0:.1f}s")
    t_enc = time.perf_counter() - t0
    t0 = time.perf_counter()
    ctx.set_model(model)
    t_keygen = time.perf_counter() - t0
    kb, maxabs = ctx.model_info()
    log(f"keygen {S} SVs {t_keygen * 1e3:.1f} ms; contraction kb={kb} max|s'|={maxabs}")

    for _ in range(args.warmup):
        labels, _, ms = ctx.predict(0, R)
    # predict() synchronises the device before it returns, so the timed region
    # starts and ends with an idle GPU
    contract_ms = 0.0
    limb_ms = [0.0, 0.0, 0.0]
    exact_pairs = 0
    t0 = time.perf_counter()
    for _ in range(args.steps):
        labels, _, ms = ctx.predict(0, R)
        contract_ms += ms[0]
        limb_ms = [a + b for a, b in zip(limb_ms, ctx.contract_limb_ms())]
        exact_pairs += ctx.contract_form()[1]
    elapsed = time.perf_counter() - t0
    form = ctx.contract_form()[0]
    ms_per_step = 1e3 * elapsed / args.steps
    value = R * args.steps / elapsed

    # decryption check: every decrypted inner product equals the plaintext one
    # mod p (the reference's rounding has no final mod p, so 0 may come out as
    # p), and the timed form's key values equal the per-limb exact form's
    chk = min(R, 4096)
    _, kv, _ = ctx.predict(0, chk, want_kvalues=True)
    ip = x[:chk].astype(np.int64) @ sv.astype(np.int64).T
    dec_ok = float(np.mean((kv.astype(np.int64) % P_MOD) == ip))
    dec_p = int(np.sum(kv == P_MOD))
    os.environ["CUFE_CONTRACT"] = "limbs"
    try:
        _, kv_limbs, _ = ctx.predict(0, chk, want_kvalues=True)
    finally:
        os.environ.pop("CUFE_CONTRACT", None)
    dec_same = float(np.mean(kv == kv_limbs))

    chunk_rows = max(64, (1 << 30) // S)
    chunks = (R + chunk_rows - 1) // chunk_rows
    K = 4096 + (max(int(np.nonzero(sv.any(axis=0))[0].max()) + 1, 1) + 15) // 16 * 16  # key columns
    nb = ctx.contract_digits()
    i8 = nb > 0 and os.environ.get("CUFE_CONTRACT", "") not in ("sgpr", "lds")
    # The dominant kernel: the fraction form's contraction (every limb in one
    # launch per chunk), or, per limb, the contraction of limbs 1 and 2 (two of
    # the three launches per chunk; limb 0, whose residues fit three int8
    # digits, runs a kernel with 3/4 of the MFMA work and is reported beside it).
    tuned = {}
    for line in cf.autotune_report().splitlines():
        if line.startswith(f"S={S} R={R} ") and "-> variant " in line:
            which = "fraction" if "fraction form:" in line else "limb 0" if " limb 0:" in line else "limbs 1-2"
            tuned[which] = line.split("-> variant ")[1].split("(")[1].rstrip(")")
    tfile = os.path.join(ROOT, "profiles", "k_contract_i8_traffic.json")
    tdata = {}
    if os.path.exists(tfile):
        with open(tfile) as f:
            t = json.load(f)
        if t.get("rows") == R and t.get("svs") == S:
            tdata = t.get("variants", {})

    def roofline(limbs, digits_a):
        launches = len(limbs) * chunks * args.steps
        avg = sum(limb_ms[s] for s in limbs) / launches
        if i8:
            ops = 2.0 * digits_a * nb * R * S * K / chunks  # digits_a x nb int8 GEMMs, one MAC = 2 ops
            label = tuned.get("limb 0" if limbs == [0] else "limbs 1-2")
            r = {"bound": "mfma", "kernel": tdata.get(label, {}).get("kernel", "k_contract_i8"), "variant": label,
                 "peak": INT8_PEAK_TOPS, "unit": "TOP/s",
                 "work_per_launch": f"{R} rows x {S} keys x {K} columns x {digits_a * nb} digit products "
                                    f"(int8 MFMA) for one limb"}
            traffic = tdata.get(label, {}).get("hbm_bytes_per_launch")
        else:
            ops = 2.0 * R * S * K / chunks  # one fp64 FMA per modular MAC
            r = {"bound": "valu-fp64", "kernel": "k_contract_sgpr", "peak": FP64_PEAK_TFLOPS, "unit": "TFLOP/s",
                 "work_per_launch": f"{R} rows x {S} keys x {K} columns (1 fp64 FMA each) for one limb"}
            traffic = None
        achieved = ops / (avg * 1e-3) / 1e12
        r.update({"achieved": achieved, "frac": achieved / r["peak"], "traffic": traffic, "avg_launch_ms": avg})
        return r

    if form == 2:
        launches = chunks * args.steps
        avg = limb_ms[0] / launches
        label = tuned.get("fraction")
        ops = 2.0 * 17 * R * S * K / chunks  # 17 int8 digit GEMMs, one MAC = 2 ops
        achieved = ops / (avg * 1e-3) / 1e12
        roof = {"bound": "mfma", "kernel": tdata.get(label, {}).get("kernel", "k_contract_frac"), "variant": label,
                "peak": INT8_PEAK_TOPS, "unit": "TOP/s",
                "work_per_launch": f"{R} rows x {S} keys x {K} columns x 17 digit products (int8 MFMA), "
                                   f"all three limbs",
                "achieved": achieved, "frac": achieved / INT8_PEAK_TOPS,
                "traffic": tdata.get(label, {}).get("hbm_bytes_per_launch"), "avg_launch_ms": avg,
                "exact_kernel_ms": limb_ms[1] / launches, "exact_pairs_per_step": exact_pairs / args.steps}
    else:
        roof = roofline([1, 2], 4)
        roof["limb0"] = roofline([0], 3 if i8 and three_row_digits(tuned.get("limb 0", "")) else 4)
    roof["modular_macs_per_s"] = 3.0 * R * S * K * args.steps / (contract_ms * 1e-3)
    out = {
        "metric": "IPFE-SVM predictions/sec on 1 MI355X, batched vs plain (bit-exact); p99 ms",
        "value": value,
        "unit": "predictions/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "u32 residues (exact int8-digit MFMA contraction)" if form != 2 else
                 "u32 residues (int8-digit MFMA contraction over 72-bit CRT fractions, exact by per-pair check)",
        "data": "synthetic",
        "config": {"workload": f"batched IPFE-SVM prediction: {R} sparse binary 123-dim rows x {S} SVs, "
                               f"N=4096, 3 RNS limbs, key length 785",
                   "rows": R, "svs": S, "row_slots": l, "parallelism": "single GPU"},
        "roofline": roof,
        "phases": {"setup_s": t_setup, "encrypt_s": t_enc, "encrypt_rows_per_s": R / t_enc,
                   "keygen_ms": t_keygen * 1e3,
                   "contract_ms_per_limb": [round(t / args.steps, 4) for t in limb_ms],
                   "contract_form": {2: "fraction (all limbs in one launch; per-limb list: contraction, exact "
                                        "kernel, 0)", 1: "int8 per limb", 0: "fp64"}[form],
                   "last_step_device_ms": {"contract": ms[0], "kvalue": ms[1],
                                                                         "vote": ms[2], "total": ms[3]},
                   "predictions_per_s_with_encryption": R / (t_enc + ms_per_step / 1e3)},
        "decryption_check": {"pairs": chk * S, "fraction_equal_mod_p": dec_ok, "zero_decrypted_as_p": dec_p,
                             "fraction_equal_per_limb_form": dec_same},
    }
    # serving right after the prediction leg, and the batched replay once more
    # after the single-operation, batched-encryption and plain legs (about 82 GB
    # of ciphertexts allocated and released in between): both p99s are reported
    sb = None
    if args.serve_requests > 0:
        sb = ServingBench(cf, ctx, model, x, args, rng)
        out["serving"] = serving(sb, args)
        out["p50_ms"] = out["serving"]["batched"]["p50_ms"]
        out["p99_ms"] = out["serving"]["batched"]["p99_ms"]
    if not args.no_single_ops:
        log("single-operation kernels")
        out["single_ops"] = single_ops(cf, ctx, rng)
    if args.enc_messages > 0:
        out["batched_encrypt"] = batched_encrypt(cf, ctx, args, rng)
    if args.plain_rows > 0:
        log("plain per-call GPU form")
        out["plain_gpu"] = plain_gpu(cf, ctx, model, x, args.plain_rows)
        out["batched_vs_plain"] = value / out["plain_gpu"]["value"]
    if sb is not None:
        out["serving"]["batched_after_other_legs"] = sb.run("batched_after_other_legs", True, 1)
        sb.close()
    if not args.no_cpu:
        log("cpu baseline (svm_predict_cpu, 1 core)")
        out["cpu_baseline"] = cpu_baseline(cf, ctx, model, x[0], args.cpu_seconds)
        log("cpu baseline (svm_predict_cpu, all granted cores)")
        out["cpu_baseline_all_cores"] = cpu_baseline_all_cores(cf, ctx, model, x[0], args.cpu_seconds / 2)
    ctx.close()
    out["autotune"] = cf.autotune_report().strip().split("\n")  # every shape tuned, serving batches included
    detail = write_detail(out, args.detail)
    print(compact_line(out, detail), flush=True)


def three_row_digits(label: str) -> bool:
    """Whether a contraction variant label names a kernel that uses three row
    digits: after the ring depth, the kernel letters end in "3", before the
    optional tile-order suffix "gGRxKG"."""
    return re.search(r"/\d+[a-z]+3(g\d+x\d+)?$", label) is not None


def write_detail(out: dict, path: str) -> str:
    """The full result (autotuner report, every leg) as JSON in a file; the
    stdout line carries only the summary."""
    path = path if os.path.isabs(path) else os.path.join(ROOT, path)
    os.makedirs(os.path.dirname(path), exist_ok=True)
    with open(path, "w") as f:
        json.dump(out, f, indent=1)
    log(f"full results in {path}")
    return os.path.relpath(path, ROOT)


LINE_LIMIT = 4000  # bytes of the final stdout line


def _r(v, nd=4):
    return float(f"{v:.{nd}g}") if isinstance(v, float) else v


def compact_line(out: dict, detail: str | None = None) -> str:
    """The one JSON line the bench prints last: the headline metric, config,
    roofline of the dominant kernel, CPU baseline, the plain per-call form,
    serving p50/p99 and a few per-leg figures -- under LINE_LIMIT bytes."""
    keep = ("metric", "value", "unit", "n_gpus", "steps", "warmup", "ms_per_step", "higher_is_better", "scaling",
            "vs_baseline", "dtype", "data", "config", "p50_ms", "p99_ms", "batched_vs_plain")
    line = {k: _r(out[k]) for k in keep if k in out}
    line["value"] = out["value"]
    line["ms_per_step"] = out["ms_per_step"]
    roof = out.get("roofline")
    if roof:
        rk = ("bound", "kernel", "variant", "achieved", "peak", "unit", "frac", "traffic", "avg_launch_ms",
              "work_per_launch", "exact_kernel_ms", "exact_pairs_per_step")
        line["roofline"] = {k: _r(roof[k]) for k in rk if k in roof}
        if "limb0" in roof:
            line["roofline"]["limb0"] = {k: _r(roof["limb0"][k]) for k in ("variant", "achieved", "frac",
                                                                          "avg_launch_ms") if k in roof["limb0"]}
    for key in ("cpu_baseline", "cpu_baseline_all_cores"):
        if key in out:
            line[key] = {k: _r(v) for k, v in out[key].items()}
    if "plain_gpu" in out:
        line["plain_gpu"] = {k: _r(out["plain_gpu"][k]) for k in ("value", "unit", "ms_per_prediction")}
    sv = out.get("serving")
    if sv:
        line["serving"] = {m: {k: _r(sv[m][k]) for k in ("p50_ms", "p99_ms", "throughput_req_s", "batches",
                                                          "failed", "requests_differing_from_ctx_predict")}
                           for m in sv if isinstance(sv[m], dict) and "p99_ms" in sv[m]}
    ph = out.get("phases")
    if ph:
        line["phases"] = {k: _r(ph[k]) for k in ("encrypt_rows_per_s", "keygen_ms", "contract_ms_per_limb",
                                                 "contract_form", "predictions_per_s_with_encryption") if k in ph}
    if "decryption_check" in out:
        line["decryption_check"] = out["decryption_check"]
    be = out.get("batched_encrypt")
    if be:
        line["batched_encrypt"] = {k: _r(be[k]) for k in ("encryptions_per_s_graph", "eager_ms", "graph_ms",
                                                          "batched_vs_plain", "graph_equals_eager",
                                                          "message0_equals_plain_abi_prefix") if k in be}
    so = out.get("single_ops")
    if so:
        line["single_ops"] = {k: _r(v["limb_transforms_per_s"] if isinstance(v, dict) and
                                    "limb_transforms_per_s" in v else v.get("polys_per_s"))
                              for k, v in so.items() if isinstance(v, dict) and
                              ("limb_transforms_per_s" in v or "polys_per_s" in v)}
        for pid, v in so.get("parameter_sets", {}).items():
            line["single_ops"][f"set{pid}_N{v['N']}_ntt_fwd_inv"] = [_r(v["ntt_forward"]["limb_transforms_per_s"]),
                                                                    _r(v["ntt_inverse"]["limb_transforms_per_s"])]
            line["single_ops"][f"set{pid}_decrypts_ok"] = \
                v["scheme_l123_host_buffers_ms"]["every_coefficient_decrypts_to_inner_product"]
    if detail:
        line["detail"] = detail
    text = json.dumps(line, separators=(",", ":"))
    # shed the optional blocks, least important first, until the line fits
    for key in ("single_ops", "batched_encrypt", "decryption_check", "phases", "serving", "cpu_baseline_all_cores"):
        if len(text.encode()) <= LINE_LIMIT:
            break
        line.pop(key, None)
        text = json.dumps(line, separators=(",", ":"))
    return text


if __name__ == "__main__":
    main()
