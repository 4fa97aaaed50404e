"""The C ABI library loads and exports every function include/cufe_ipfe.h
declares (no GPU needed: nothing is called)."""
import ctypes as C
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "cufe_ipfe.h")
LIB = os.path.join(ROOT, "cufe_amd", "libcufe_ipfe.so")


def declared_functions(header=HEADER):
    text = open(header).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    names = re.findall(r"^\s*(?:const\s+)?[A-Za-z_][A-Za-z0-9_\s\*]*?\b([A-Za-z_][A-Za-z0-9_]*)\s*\(", text, flags=re.M)
    return sorted({n for n in names if n not in ("defined",)})


def test_header_declares_reference_entry_points():
    names = declared_functions()
    for ref_name in ("rlwe_sife_setup_gpu", "rlwe_sife_encrypt_gpu", "rlwe_sife_keygen_gpu",
                     "rlwe_sife_decrypt_gmp_gpu"):
        assert ref_name in names


def test_library_exports_every_declared_symbol():
    if not os.path.exists(LIB):
        pytest.fail("libcufe_ipfe.so is not built")
    lib = C.CDLL(LIB)
    missing = [n for n in declared_functions() if not hasattr(lib, n)]
    assert not missing, missing


def test_bench_library_exports_its_header():
    """The microbenchmarks live in libcufe_bench.so (include/cufe_bench.h), not
    in the drop-in library."""
    bench_lib = os.path.join(ROOT, "cufe_amd", "libcufe_bench.so")
    if not os.path.exists(bench_lib):
        pytest.fail("libcufe_bench.so is not built")
    names = declared_functions(os.path.join(ROOT, "include", "cufe_bench.h"))
    assert names == ["cufe_microbench"]
    assert hasattr(C.CDLL(bench_lib), "cufe_microbench")
    assert not hasattr(C.CDLL(LIB), "cufe_microbench")


def test_python_binding_rejects_missing_library(monkeypatch):
    import cufe_amd

    monkeypatch.setattr(cufe_amd, "_lib", None)
    monkeypatch.setattr(cufe_amd, "LIB_PATH", "/nonexistent/libcufe_ipfe.so")
    with pytest.raises(cufe_amd.CufeError):
        cufe_amd.lib()


def test_bench_refuses_more_than_one_gpu():
    """bench.py measures one GPU: --gpus above 1 exits with status 2 and says
    so, before it touches a device."""
    import subprocess
    import sys

    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"], capture_output=True,
                       text=True, timeout=120, cwd=ROOT)
    assert r.returncode == 2
    assert "one GPU" in r.stderr and r.stdout == ""
