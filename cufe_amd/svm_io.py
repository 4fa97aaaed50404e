"""libsvm text formats for the SVM caller of the IPFE path.

read_model follows the model-file grammar of svm_load_model (svm/svm.cpp:2974):
a header of "key value" lines up to "SV", then one line per support vector with
nr_class-1 coefficients followed by index:value pairs.  read_problem reads the
sparse "label index:value ..." rows the reference driver parses (main.c:461-512).
Features go to the dense slot of their index (convertToarray, main.c:30-41),
with values converted to unsigned integers as the reference's assignment does.
"""
from __future__ import annotations

import numpy as np

from . import L, SvmModel


def _dense_row(pairs, width=L):
    row = np.zeros(width, dtype=np.uint32)
    for idx, val in pairs:
        if not 0 <= idx < width:
            raise ValueError(f"feature index {idx} outside 0..{width - 1}")
        row[idx] = np.uint32(int(val))
    return row


def read_model(path: str) -> SvmModel:
    header = {}
    coefs, rows = [], []
    with open(path) as f:
        for line in f:
            line = line.strip()
            if line == "SV":
                break
            key, _, val = line.partition(" ")
            header[key] = val.split()
        if header.get("kernel_type", ["linear"])[0] != "linear":
            raise ValueError("only linear kernels have an inner-product decision")
        nr_class = int(header["nr_class"][0])
        for line in f:
            parts = line.split()
            if not parts:
                continue
            coefs.append([float(v) for v in parts[: nr_class - 1]])
            pairs = []
            for tok in parts[nr_class - 1:]:
                i, v = tok.split(":")
                pairs.append((int(i), float(v)))
            rows.append(_dense_row(pairs))
    total = int(header["total_sv"][0])
    if len(rows) != total:
        raise ValueError(f"model lists {len(rows)} SVs, header says {total}")
    sv_coef = np.array(coefs, dtype=np.float64).T.copy()  # [nr_class-1][l]
    return SvmModel(nr_class, [int(v) for v in header["nr_sv"]], [int(v) for v in header["label"]], sv_coef,
                    [float(v) for v in header["rho"]], np.stack(rows))


def read_problem(path: str, skip: int = 0, limit: int | None = None):
    """(labels [n], dense rows [n][785]) from a libsvm data file."""
    labels, rows = [], []
    with open(path) as f:
        for k, line in enumerate(f):
            if k < skip:
                continue
            if limit is not None and len(rows) >= limit:
                break
            parts = line.split()
            if not parts:
                continue
            labels.append(float(parts[0]))
            pairs = []
            for tok in parts[1:]:
                i, v = tok.split(":")
                pairs.append((int(i), float(v)))
            rows.append(_dense_row(pairs))
    return np.array(labels), np.stack(rows) if rows else np.zeros((0, L), np.uint32)


def plaintext_predict(model: SvmModel, x_dense) -> np.ndarray:
    """svm_predict_plain (main.c:55-134): kvalue = <x, SV> in double, then the vote."""
    x = np.asarray(x_dense, dtype=np.float64)
    kv = x @ model.sv.astype(np.float64).T
    out = np.empty(x.shape[0])
    for r in range(x.shape[0]):
        out[r] = decide(model, kv[r])
    return out


def decide(model: SvmModel, kvalue) -> float:
    """One-vs-one vote of main.c:338-389, products rounded then summed in order."""
    start = np.concatenate([[0], np.cumsum(model.nSV)[:-1]]).astype(int)
    vote = [0] * model.nr_class
    p = 0
    for i in range(model.nr_class):
        for j in range(i + 1, model.nr_class):
            s = 0.0
            c1, c2 = model.sv_coef[j - 1], model.sv_coef[i]
            for k in range(model.nSV[i]):
                s = s + float(c1[start[i] + k]) * float(kvalue[start[i] + k])
            for k in range(model.nSV[j]):
                s = s + float(c2[start[j] + k]) * float(kvalue[start[j] + k])
            s -= float(model.rho[p])
            if s > 0:
                vote[i] += 1
            else:
                vote[j] += 1
            p += 1
    best = 0
    for i in range(1, model.nr_class):
        if vote[i] > vote[best]:
            best = i
    return float(model.label[best])
