"""bench.py's final stdout line: a compact JSON object that a reader keeping
only the last few KB of stdout can parse, carrying the headline metric and the
fields the measurement needs (roofline, CPU baseline, plain form, p50/p99)."""
import json
import os
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)

import bench  # noqa: E402


def representative_result():
    serve = {"p50_ms": 0.1834, "p99_ms": 0.2871, "p999_ms": 0.41, "mean_ms": 0.19, "throughput_req_s": 1998.7,
             "throughput_rows_s": 65000.0, "makespan_s": 2.0, "batches": 3900, "max_batch_rows": 128,
             "slot_cap_rows": 4096, "failed": 0, "requests_differing_from_ctx_predict": 0,
             "arrival_rate_req_s": 2000.0}
    return {
        "metric": "IPFE-SVM predictions/sec on 1 MI355X, batched vs plain (bit-exact); p99 ms",
        "value": 6212345.678, "unit": "predictions/s", "n_gpus": 1, "steps": 20, "warmup": 5,
        "ms_per_step": 16.0972, "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "u32 residues (exact int8-digit MFMA contraction)", "data": "synthetic",
        "config": {"workload": "batched IPFE-SVM prediction: 100000 sparse binary 123-dim rows x 2000 SVs, "
                               "N=4096, 3 RNS limbs, key length 785", "rows": 100000, "svs": 2000,
                   "row_slots": 124, "parallelism": "single GPU"},
        "roofline": {"bound": "mfma", "kernel": "k_contract_i8_il", "variant": "2x2/4imw", "peak": 5000.0,
                     "unit": "TOP/s", "work_per_launch": "100000 rows x 2000 keys x 4224 columns x 8 digit "
                                                        "products (int8 MFMA) for one limb",
                     "achieved": 2380.123456, "frac": 0.476024, "traffic": 19234567890, "avg_launch_ms": 5.68,
                     "limb0": {"variant": "1x4/4im3", "achieved": 2200.5, "frac": 0.44, "avg_launch_ms": 4.4,
                               "bound": "mfma", "peak": 5000.0},
                     "modular_macs_per_s": 1.6e14},
        "autotune": ["S=2000 R=100000 ksteps=132 key digits=2 limbs 1-2: " + "2x2 0.054 " * 40] * 120,
        "phases": {"setup_s": 1.2, "encrypt_s": 6.2, "encrypt_rows_per_s": 16100.0, "keygen_ms": 40.0,
                   "contract_ms_per_limb": [4.4, 5.6, 5.8],
                   "last_step_device_ms": {"contract": 15.8, "kvalue": 0.0, "vote": 0.18, "total": 16.0},
                   "predictions_per_s_with_encryption": 16000.0},
        "decryption_check": {"pairs": 512000, "fraction_equal_mod_p": 1.0, "zero_decrypted_as_p": 41959},
        "serving": {"arrival_rate_req_s": 2000.0, "requests": 4000, "rows": 130000,
                    "models": [{"svs": 100, "row_slots": 124}] * 4,
                    "batched": serve, "one_request_per_batch": serve, "batched_x10": serve,
                    "one_request_per_batch_x10": serve, "batched_after_other_legs": serve},
        "p50_ms": 0.1834, "p99_ms": 0.2871,
        "single_ops": {"butterfly_peak_per_s": 3.2e12,
                       "ntt_forward": {"limb_transforms_per_s": 8.9e7, "ms_per_10000_polys": 0.33},
                       "ntt_inverse": {"limb_transforms_per_s": 8.8e7, "ms_per_10000_polys": 0.34},
                       "gauss_S1_10000_streams": {"ms": 6.1, "polys_per_s": 1.6e6},
                       "gauss_S3_10000_streams": {"ms": 11.0, "polys_per_s": 9.1e5},
                       "parameter_sets": {str(pid): {"N": 2048 << pid, "limbs": 2 + pid,
                                                     "ntt_forward": {"limb_transforms_per_s": 9.1e7},
                                                     "ntt_inverse": {"limb_transforms_per_s": 9.0e7},
                                                     "scheme_l123_host_buffers_ms": {
                                                         "every_coefficient_decrypts_to_inner_product": True}}
                                          for pid in range(3)}},
        "batched_encrypt": {"messages": 4096, "eager_ms": 1332.5, "graph_ms": 1331.6,
                            "encryptions_per_s_graph": 3076.0, "batched_vs_plain": 49.0,
                            "graph_equals_eager": True, "message0_equals_plain_abi_prefix": True,
                            "plain_abi": {"note": "x" * 300}},
        "plain_gpu": {"value": 7.0, "unit": "predictions/s", "rows": 2, "ms_per_prediction": 142.8},
        "batched_vs_plain": 887477.9,
        "cpu_baseline": {"value": 0.0198, "unit": "predictions/s", "cores": 1, "kind": "reference",
                         "sample": "1 row x 120 of 2000 SVs of svm_predict_cpu (reference build, -O3 -DAVX2) in "
                                   "12.0 s; 100.0 ms per (row, SV) pair, extrapolated to 2000 SVs"},
        "cpu_baseline_all_cores": {"value": 0.31, "unit": "predictions/s", "cores": 16, "kind": "port",
                                   "sample": "16 processes x 6 s, 900 (row, SV) pairs of svm_predict_cpu"},
    }


def test_line_is_compact_and_complete():
    out = representative_result()
    text = bench.compact_line(out, "gpurun_out/bench_detail.json")
    assert "\n" not in text
    assert len(text.encode()) <= bench.LINE_LIMIT
    line = json.loads(text)
    for k in ("metric", "value", "unit", "n_gpus", "steps", "warmup", "ms_per_step", "higher_is_better",
              "scaling", "vs_baseline", "dtype", "data", "config", "roofline", "cpu_baseline", "plain_gpu",
              "batched_vs_plain", "p50_ms", "p99_ms"):
        assert k in line, k
    assert line["value"] == out["value"] and line["ms_per_step"] == out["ms_per_step"]
    for k in ("bound", "achieved", "peak", "unit", "frac", "traffic", "avg_launch_ms"):
        assert k in line["roofline"], k
    for k in ("value", "unit", "cores", "kind", "sample"):
        assert k in line["cpu_baseline"], k
    assert "autotune" not in line
    assert line["detail"] == "gpurun_out/bench_detail.json"
    assert line["single_ops"]["set2_N8192_ntt_fwd_inv"] == [9.1e7, 9.0e7] and line["single_ops"]["set0_decrypts_ok"]


def test_line_sheds_optional_blocks_when_too_long():
    out = representative_result()
    out["cpu_baseline"]["sample"] = "y" * 2500
    text = bench.compact_line(out)
    assert len(text.encode()) <= bench.LINE_LIMIT
    line = json.loads(text)
    assert "roofline" in line and "cpu_baseline" in line and "p99_ms" in line


def test_three_row_digit_labels():
    for label in ("2x2/4im3", "2x4/4imw3", "4x2/3imw3g4x16", "2x4/4imw3g8x8", "2x4/4ipmw3g8x8"):
        assert bench.three_row_digits(label), label
    for label in ("2x2", "1x4/3", "2x4/5imw", "2x4/5imwg8x8", "1x4/2k2", "2x4/4ipmwg8x8", "2x2/4img8x12"):
        assert not bench.three_row_digits(label), label
