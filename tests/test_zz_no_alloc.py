"""cufe_gauss_sample_noalloc against cufe_gauss_sample through the C ABI: two
calls in a row (the second smaller, so stale scratch from the first is
present), each equal word for word to the allocating entry point; oversized
streams, host outputs and short outputs are rejected before any launch.  The
device output buffer is the caller's (hipMalloc through the HIP runtime)."""
import ctypes as C
import hashlib
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

N = 4096


def note(msg):
    print(f"[no-alloc] {msg}", flush=True)


@pytest.fixture(scope="module")
def hip():
    rt = C.CDLL(os.path.join("/opt/rocm/lib", "libamdhip64.so"))
    rt.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
    rt.hipFree.argtypes = [C.c_void_p]
    rt.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
    return rt


def keys_nonces(tag, n):
    keys = np.frombuffer(b"".join(hashlib.sha256(tag + bytes([i])).digest() for i in range(n)), np.uint8).copy()
    return keys, np.arange(7, 7 + n, dtype=np.uint16)


def sample_noalloc(gpu, hip, keys, nonces, which, ncalls, d_out):
    lb = gpu.lib()
    lb.cufe_gauss_sample_noalloc.restype = C.c_int
    lb.cufe_gauss_sample_noalloc.argtypes = [C.c_void_p, C.c_void_p, C.c_int, C.c_int, C.c_int, C.c_void_p]
    n = len(nonces)
    return lb.cufe_gauss_sample_noalloc(keys.ctypes.data, nonces.ctypes.data, n, which, ncalls, d_out)


def test_noalloc_matches_allocating_entry_point(gpu, hip):
    big = 8 * 2 * 3 * N * 4
    d_out = C.c_void_p()
    assert hip.hipMalloc(C.byref(d_out), big) == 0
    try:
        for tag, n, which, ncalls in ((b"first", 8, 3, 2), (b"second", 3, 1, 1)):
            note(f"call {tag.decode()}: {n} streams, sampler {which}, {ncalls} calls")
            keys, nonces = keys_nonces(tag, n)
            assert sample_noalloc(gpu, hip, keys, nonces, which, ncalls, d_out) == 0, gpu.lib().cufe_last_error()
            got = np.empty((n, ncalls, 3, N), np.uint32)
            assert hip.hipMemcpy(got.ctypes.data, d_out, got.nbytes, 2) == 0  # device to host
            want = gpu.gauss_sample(keys.reshape(n, 32), nonces, which, ncalls)
            assert np.array_equal(got, want.reshape(got.shape)), tag
    finally:
        hip.hipFree(d_out)


def test_noalloc_rejects_what_does_not_fit(gpu, hip):
    keys, nonces = keys_nonces(b"reject", 1)
    small = C.c_void_p()
    assert hip.hipMalloc(C.byref(small), 3 * N * 4) == 0
    try:
        # a stream whose window needs more than the 64 MiB scratch
        assert sample_noalloc(gpu, hip, keys, nonces, 3, 5000, small) == -1
        assert b"scratch" in gpu.lib().cufe_last_error()
        # an output shorter than [1][2][3][N]
        assert sample_noalloc(gpu, hip, keys, nonces, 1, 2, small) == -1
        assert b"needs" in gpu.lib().cufe_last_error()
        # host memory as the output
        host = np.empty(3 * N, np.uint32)
        assert sample_noalloc(gpu, hip, keys, nonces, 1, 1, C.c_void_p(host.ctypes.data)) == -1
        assert b"device memory" in gpu.lib().cufe_last_error()
    finally:
        hip.hipFree(small)
