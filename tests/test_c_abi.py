"""The drop-in boundary from C: tests/c_abi/abi_smoke.c calls the four
reference entry points (rlwe_sife_setup_gpu / encrypt / keygen /
decrypt_gmp_gpu) with the reference's own array types, plus the two host
helpers, linked against libcufe_ipfe.so the way the reference's Makefile would
link its CUDA objects.  Compiling and linking needs no GPU; running does."""
import os
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIBDIR = os.path.join(ROOT, "cufe_amd")


def build(tmp_path):
    exe = str(tmp_path / "abi_smoke")
    cmd = ["gcc", "-O2", "-std=c11", "-Wall", "-Werror", "-I", os.path.join(ROOT, "include"),
           os.path.join(ROOT, "tests", "c_abi", "abi_smoke.c"), "-L", LIBDIR, "-lcufe_ipfe",
           "-Wl,-rpath," + LIBDIR, "-o", exe]
    subprocess.run(cmd, check=True, capture_output=True, text=True)
    return exe


def test_c_caller_compiles_and_links(tmp_path):
    if not os.path.exists(os.path.join(LIBDIR, "libcufe_ipfe.so")):
        pytest.skip("library not built")
    exe = build(tmp_path)
    assert os.path.exists(exe)


@pytest.mark.gpu
def test_c_caller_runs(tmp_path):
    exe = build(tmp_path)
    print("[gpu-c-abi] running abi_smoke", flush=True)
    r = subprocess.run([exe], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    assert r.stdout.startswith("abi ok"), r.stdout
