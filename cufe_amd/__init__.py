"""cufe_amd -- Python bindings of libcufe_ipfe.so, the MI355X (gfx950) RLWE-IPFE
("rlwe_sife") library.

The product is the HIP library behind the C ABI in include/cufe_ipfe.h; this
module only marshals numpy buffers to it.  There is no CPU fallback: if the
library is missing, importing a function raises.

Names mirror the reference's entry points (src/rlwe_sife.h, src/rlwe_sife_gpu.cu):
setup / encrypt / keygen / decrypt, sample_polya, round_extract, svm_predict.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

N = 4096
NMODULI = 3
L = 785
P = 50241
Q = (16760833, 2147352577, 2130706433)

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libcufe_ipfe.so")
BENCH_LIB_PATH = os.path.join(_HERE, "libcufe_bench.so")  # measurement library (microbenchmarks)
_lib = None
_bench_lib = None

_u32p = np.ctypeslib.ndpointer(dtype=np.uint32, flags="C_CONTIGUOUS")
_u8p = np.ctypeslib.ndpointer(dtype=np.uint8, flags="C_CONTIGUOUS")
_u16p = np.ctypeslib.ndpointer(dtype=np.uint16, flags="C_CONTIGUOUS")
_f64p = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")
_i32p = np.ctypeslib.ndpointer(dtype=np.int32, flags="C_CONTIGUOUS")


class CufeError(RuntimeError):
    pass


def lib() -> C.CDLL:
    """Load the HIP library (raises if it has not been built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise CufeError(f"{LIB_PATH} is missing: build it with `python -c 'import __graft_entry__ as g; g.build()'`")
    lb = C.CDLL(LIB_PATH)
    sig = {
        "cufe_last_error": (C.c_char_p, []),
        "cufe_device_info": (C.c_int, [C.c_char_p, C.c_int, C.POINTER(C.c_int), C.POINTER(C.c_uint64)]),
        "cufe_ntt_forward": (C.c_int, [_u32p, C.c_int]),
        "cufe_ntt_inverse": (C.c_int, [_u32p, C.c_int]),
        "cufe_gauss_sample": (C.c_int, [_u8p, _u16p, C.c_int, C.c_int, C.c_int, _u32p]),
        "cufe_gauss_sample_first": (C.c_int, [_u8p, _u16p, C.c_int, C.c_int, C.c_int, _u32p]),
        "cufe_aes256ctr": (C.c_int, [_u8p, C.c_uint16, C.c_uint32, C.c_uint32, _u8p]),
        "cufe_crt_round": (C.c_int, [_u32p, C.c_int, C.c_int, _u32p]),
        "cufe_sample_polya": (C.c_int, [_u8p, _u32p]),
        "cufe_round_extract2": (C.c_double, [_u32p]),
        "cufe_sife_setup": (C.c_int, [_u32p, _u32p, _u8p, _u8p, C.c_int]),
        "cufe_sife_encrypt": (C.c_int, [_u32p, _u32p, _u32p, _u8p, _u8p, C.c_int, C.c_int, C.c_int]),
        "cufe_sife_keygen": (C.c_int, [_u32p, _u32p, _u32p, C.c_int, C.c_int]),
        "cufe_sife_decrypt": (C.c_int, [_u32p, _u32p, _u32p, _u32p, C.c_int, C.c_int]),
        "cufe_time_ntt": (C.c_int, [C.c_int, C.c_int, C.c_int, C.POINTER(C.c_float)]),
        "cufe_time_gauss": (C.c_int, [C.c_int, C.c_int, C.c_int, C.POINTER(C.c_float)]),
        "cufe_time_gauss_calls": (C.c_int, [C.c_int, C.c_int, C.c_int, C.c_int, C.POINTER(C.c_float)]),
        "cufe_autotune_report": (C.c_int, [C.c_char_p, C.c_int]),
        "cufe_host_alloc": (C.c_int, [C.c_size_t, C.POINTER(C.c_void_p)]),
        "cufe_contract_variants": (C.c_int, [C.c_int, C.POINTER(C.c_ulonglong)]),
        "cufe_contract_frac_variants": (C.c_int, []),
        "cufe_wire_size": (C.c_size_t, [C.c_int, C.c_int, C.c_int]),
        "cufe_wire_pack": (C.c_int, [C.c_int, _u32p, C.c_void_p, C.c_int, C.c_int, C.c_uint32, C.c_void_p,
                                     C.c_size_t]),
        "cufe_wire_parse": (C.c_int, [C.c_void_p, C.c_size_t, C.c_void_p]),
        "cufe_host_free": (None, [C.c_void_p]),
        "rlwe_sife_setup_gpu": (None, [_u32p, _u32p, _u8p, _u8p, C.POINTER(C.c_float)]),
        "rlwe_sife_encrypt_gpu": (None, [_u32p, _u32p, _u32p, _u8p, _u8p, C.c_int, C.POINTER(C.c_float)]),
        "rlwe_sife_keygen_gpu": (None, [_u32p, _u32p, _u32p, C.c_int, C.POINTER(C.c_float)]),
        "rlwe_sife_decrypt_gmp_gpu": (None, [_u32p, _u32p, _u32p, _u32p, C.c_int, C.POINTER(C.c_float)]),
    }
    for name, (res, args) in sig.items():
        f = getattr(lb, name)
        f.restype = res
        f.argtypes = args
    _lib = lb
    return lb


def _check(rc: int, what: str) -> None:
    if rc != 0:
        msg = lib().cufe_last_error()
        raise CufeError(f"{what}: {msg.decode() if msg else 'unknown error'}")


def _seed(s) -> np.ndarray:
    a = np.frombuffer(bytes(s), dtype=np.uint8).copy()
    if a.size != 32:
        raise ValueError("seeds are 32 bytes")
    return a


def _u32(a, shape=None) -> np.ndarray:
    a = np.ascontiguousarray(a, dtype=np.uint32)
    if shape is not None and a.shape != tuple(shape):
        a = a.reshape(shape)
    return a


def device_info() -> dict:
    name = C.create_string_buffer(256)
    cu = C.c_int()
    hbm = C.c_uint64()
    _check(lib().cufe_device_info(name, 256, C.byref(cu), C.byref(hbm)), "device_info")
    return {"name": name.value.decode(), "cu": cu.value, "hbm_bytes": hbm.value}


# ---- single operations -------------------------------------------------------
def ntt_forward(polys) -> np.ndarray:
    """CT_forward (src/ntt.c:38-67) on [npoly][3][N] residues."""
    a = _u32(polys).copy()
    npoly = a.size // (NMODULI * N)
    _check(lib().cufe_ntt_forward(a, npoly), "ntt_forward")
    return a


def ntt_inverse(polys) -> np.ndarray:
    """GS_reverse (src/ntt.c:70-114) including N^-1."""
    a = _u32(polys).copy()
    npoly = a.size // (NMODULI * N)
    _check(lib().cufe_ntt_inverse(a, npoly), "ntt_inverse")
    return a


def gauss_sample(keys, nonces, which: int, ncalls: int) -> np.ndarray:
    """ncalls consecutive gaussian_sampler_S{which} calls on each stream ->
    [nstreams][ncalls][3][N]."""
    k = np.ascontiguousarray(np.frombuffer(b"".join(bytes(x) for x in keys), dtype=np.uint8))
    n = np.ascontiguousarray(nonces, dtype=np.uint16)
    out = np.empty((len(n), ncalls, NMODULI, N), dtype=np.uint32)
    _check(lib().cufe_gauss_sample(k, n, len(n), which, ncalls, out), "gauss_sample")
    return out


def gauss_sample_first(keys, nonces, which: int, ncalls: int) -> np.ndarray:
    """Sample 0 of each of ncalls consecutive calls on each stream, sampled
    the way the resident row encryption samples its f_i -> [nstreams][ncalls][3]."""
    k = np.ascontiguousarray(np.frombuffer(b"".join(bytes(x) for x in keys), dtype=np.uint8))
    n = np.ascontiguousarray(nonces, dtype=np.uint16)
    out = np.empty((len(n), ncalls, NMODULI), dtype=np.uint32)
    _check(lib().cufe_gauss_sample_first(k, n, len(n), which, ncalls, out), "gauss_sample_first")
    return out


def aes256ctr(key, nonce: int, ctr0: int, nblocks: int) -> bytes:
    out = np.empty(nblocks * 16, dtype=np.uint8)
    _check(lib().cufe_aes256ctr(_seed(key), nonce, ctr0, nblocks, out), "aes256ctr")
    return out.tobytes()


def sample_polya(seed) -> np.ndarray:
    a = np.empty((NMODULI, N), dtype=np.uint32)
    _check(lib().cufe_sample_polya(_seed(seed), a), "sample_polya")
    return a


def crt_round(d, coef: int = 0) -> np.ndarray:
    d = _u32(d)
    nvec = d.size // (NMODULI * N)
    r = np.empty(nvec, dtype=np.uint32)
    _check(lib().cufe_crt_round(d, nvec, coef, r), "crt_round")
    return r


def round_extract2(d_y) -> float:
    return lib().cufe_round_extract2(_u32(d_y))


# ---- scheme at vector length l -------------------------------------------------
def setup(a_coeff, seed2, seed3, l: int = L):
    """rlwe_sife_setup_gpu at length l: returns (mpk [l+1][3][N], msk [l][3][N])."""
    mpk = np.zeros((l + 1, NMODULI, N), dtype=np.uint32)
    mpk[l] = _u32(a_coeff, (NMODULI, N))
    msk = np.zeros((max(l, 1), NMODULI, N), dtype=np.uint32)
    _check(lib().cufe_sife_setup(mpk, msk, _seed(seed2), _seed(seed3), l), "setup")
    return mpk, msk[:l]


def encrypt(m, mpk, seed2, seed3, nonce0: int = 0):
    """Encrypt rows m [repeat][l] -> c [repeat][l+1][3][N] (message b uses nonce nonce0+b)."""
    m = _u32(np.atleast_2d(m))
    repeat, l = m.shape
    mpk = _u32(mpk)
    assert mpk.shape == (l + 1, NMODULI, N)
    c = np.empty((repeat, l + 1, NMODULI, N), dtype=np.uint32)
    _check(lib().cufe_sife_encrypt(m if m.size else np.zeros(1, np.uint32), mpk, c, _seed(seed2), _seed(seed3),
                                   repeat, l, nonce0), "encrypt")
    return c


def keygen(y, msk):
    y = _u32(np.atleast_2d(y))
    repeat, l = y.shape
    sk = np.empty((repeat, NMODULI, N), dtype=np.uint32)
    _check(lib().cufe_sife_keygen(y, _u32(msk), sk, repeat, l), "keygen")
    return sk


def decrypt(c, y, sk):
    """One ciphertext c [l+1][3][N] against keys sk [repeat][3][N] -> d_y [repeat][3][N]."""
    y = _u32(np.atleast_2d(y))
    repeat, l = y.shape
    d = np.empty((repeat, NMODULI, N), dtype=np.uint32)
    _check(lib().cufe_sife_decrypt(_u32(c), y, _u32(sk), d, repeat, l), "decrypt")
    return d


def time_ntt(npoly: int, inverse: bool, iters: int) -> float:
    ms = C.c_float()
    _check(lib().cufe_time_ntt(npoly, int(inverse), iters, C.byref(ms)), "time_ntt")
    return ms.value


def microbench(kind: int) -> float:
    """0: fp64 FMA TFLOP/s, 1: int32 mul T ops/s, 2: copy GB/s, 3: AES G blocks/s,
    4: AES (replicated table) G blocks/s, 5: int8 MFMA T ops/s, 6: fp64 MFMA TFLOP/s,
    7: NTT butterflies in registers G/s, 8: int8 MFMA T ops/s on random operands,
    sustained over 8 back-to-back launches, 9: AES (T0 and T2 replicated) G blocks/s."""
    r = C.c_double()
    _check(bench_lib().cufe_microbench(kind, C.byref(r)), "microbench")
    return r.value


def bench_lib() -> C.CDLL:
    """The measurement library beside the product library (raises if missing)."""
    global _bench_lib
    if _bench_lib is None:
        lib()
        if not os.path.exists(BENCH_LIB_PATH):
            raise CufeError(f"{BENCH_LIB_PATH} is missing: build it with `make -C cufe_amd/csrc`")
        lb = C.CDLL(BENCH_LIB_PATH)
        lb.cufe_microbench.restype = C.c_int
        lb.cufe_microbench.argtypes = [C.c_int, C.POINTER(C.c_double)]
        _bench_lib = lb
    return _bench_lib


class _PinnedBlock:
    """Owner of one cufe_host_alloc block; freed when the last array using it goes."""

    def __init__(self, nbytes: int):
        p = C.c_void_p()
        _check(lib().cufe_host_alloc(nbytes, C.byref(p)), "host_alloc")
        self.ptr = p.value

    def __del__(self):
        try:
            lib().cufe_host_free(self.ptr)
        except Exception:
            pass


def host_empty(shape, dtype=np.uint32) -> np.ndarray:
    """Uninitialised array in pinned host memory (server requests inside it are
    copied to the GPU by DMA without staging)."""
    dt = np.dtype(dtype)
    count = int(np.prod(shape))
    owner = _PinnedBlock(max(1, count * dt.itemsize))
    cbuf = (C.c_uint8 * (count * dt.itemsize)).from_address(owner.ptr)
    cbuf._owner = owner
    return np.frombuffer(cbuf, dtype=dt, count=count).reshape(shape)


# ---- wire format (requests, ciphertexts, keys) ------------------------------------
WIRE_ROWS, WIRE_CIPHERTEXTS, WIRE_PUBLIC_KEY, WIRE_FUNCTIONAL_KEYS = 1, 2, 3, 4


class _WireInfo(C.Structure):
    _fields_ = [("kind", C.c_int), ("count", C.c_int), ("l", C.c_int), ("tag", C.c_uint32),
                ("data", C.c_void_p), ("c0", C.c_void_p)]


def wire_pack(kind: int, data, data2=None, tag: int = 0) -> np.ndarray:
    """Blob (uint8 array) of one kind: rows (data = c_L [n][3][N], data2 = c0
    [n][l][3]), ciphertexts [n][l+1][3][N], public key [l+1][3][N] or
    functional keys [n][3][N]."""
    data = _u32(data)
    if kind == WIRE_ROWS:
        data2 = _u32(data2)
        count, l = data2.shape[0], data2.shape[1]
        assert data.shape == (count, NMODULI, N) and data2.shape == (count, l, NMODULI)
    elif kind == WIRE_CIPHERTEXTS:
        count, l = data.shape[0], data.shape[1] - 1
    elif kind == WIRE_PUBLIC_KEY:
        count, l = 1, data.shape[0] - 1
    else:
        count, l = data.shape[0], 0
    size = lib().cufe_wire_size(kind, count, l)
    if size == 0:
        raise CufeError(f"wire_pack: invalid shape for kind {kind}")
    out = np.empty(size, np.uint8)
    _check(lib().cufe_wire_pack(kind, data, data2.ctypes.data if data2 is not None else None, count, l, tag,
                                out.ctypes.data, size), "wire_pack")
    return out


def wire_parse(blob) -> dict:
    """Validate a blob and return its fields; arrays are views into the blob
    (copied first into an aligned uint8 array if it is not one)."""
    b = np.frombuffer(blob, np.uint8) if isinstance(blob, (bytes, bytearray)) else np.asarray(blob, np.uint8)
    if b.ctypes.data % 4:
        b = b.copy()
    info = _WireInfo()
    _check(lib().cufe_wire_parse(b.ctypes.data, b.size, C.byref(info)), "wire_parse")
    base = b.ctypes.data
    words = b[info.data - base:].view(np.uint32)
    k, n, l = info.kind, info.count, info.l
    out = {"kind": k, "count": n, "l": l, "tag": info.tag, "blob": b}
    if k == WIRE_ROWS:
        out["cl"] = words[: n * NMODULI * N].reshape(n, NMODULI, N)
        out["c0"] = b[info.c0 - base:].view(np.uint32)[: n * l * NMODULI].reshape(n, l, NMODULI)
    elif k == WIRE_CIPHERTEXTS:
        out["data"] = words[: n * (l + 1) * NMODULI * N].reshape(n, l + 1, NMODULI, N)
    elif k == WIRE_PUBLIC_KEY:
        out["data"] = words[: (l + 1) * NMODULI * N].reshape(l + 1, NMODULI, N)
    else:
        out["data"] = words[: n * NMODULI * N].reshape(n, NMODULI, N)
    return out


def contract_variants(nb: int) -> list:
    """Indices of the int8 contraction variants that run with nb key digits."""
    m = C.c_ulonglong()
    n = lib().cufe_contract_variants(nb, C.byref(m))
    return [v for v in range(n) if (m.value >> v) & 1]


def frac_variants() -> int:
    """Number of block shapes of the fraction-form contraction
    (CUFE_FRAC_VARIANT=v, 0 <= v < this, forces one)."""
    return int(lib().cufe_contract_frac_variants())


def autotune_report() -> str:
    """Block-shape choices of the int8 contraction so far."""
    n = lib().cufe_autotune_report(None, 0)
    buf = C.create_string_buffer(n + 1)
    lib().cufe_autotune_report(buf, n + 1)
    return buf.value.decode()


def time_gauss_calls(nstreams: int, ncalls: int, which: int, iters: int) -> float:
    """ms per run of nstreams streams x ncalls calls of sampler `which`."""
    ms = C.c_float()
    _check(lib().cufe_time_gauss_calls(nstreams, ncalls, which, iters, C.byref(ms)), "time_gauss_calls")
    return ms.value


def time_gauss(nstreams: int, which: int, iters: int) -> float:
    ms = C.c_float()
    _check(lib().cufe_time_gauss(nstreams, which, iters, C.byref(ms)), "time_gauss")
    return ms.value


# ---- resident context and batched prediction -------------------------------------
# ---- parameter sets -------------------------------------------------------------
def _ps_sigs(lb):
    fp = C.POINTER(C.c_float)
    sig = {
        "cufe_pset_info": (C.c_int, [C.c_int, C.POINTER(C.c_int), C.POINTER(C.c_int), C.POINTER(C.c_int),
                                     C.POINTER(C.c_uint32), C.POINTER(C.c_uint32)]),
        "cufe_ps_ntt": (C.c_int, [C.c_int, C.c_int, _u32p, C.c_int]),
        "cufe_ps_ntt_variant": (C.c_int, [C.c_int, C.c_int, C.c_int, _u32p, C.c_int]),
        "cufe_ntt_variants": (C.c_int, [C.c_int, C.c_char_p, C.c_int]),
        "cufe_ntt_autotune_report": (C.c_int, [C.c_char_p, C.c_int]),
        "cufe_ps_sample_polya": (C.c_int, [C.c_int, _u8p, _u32p]),
        "cufe_ps_gauss_sample": (C.c_int, [C.c_int, _u8p, _u16p, C.c_int, C.c_int, C.c_int, _u32p]),
        "cufe_ps_setup": (C.c_int, [C.c_int, _u32p, _u32p, _u8p, _u8p, C.c_int]),
        "cufe_ps_encrypt": (C.c_int, [C.c_int, _u32p, _u32p, _u32p, _u8p, _u8p, C.c_int, C.c_int, C.c_int]),
        "cufe_ps_keygen": (C.c_int, [C.c_int, _u32p, _u32p, _u32p, C.c_int, C.c_int]),
        "cufe_ps_decrypt": (C.c_int, [C.c_int, _u32p, _u32p, _u32p, _u32p, C.c_int, C.c_int]),
        "cufe_ps_crt_round": (C.c_int, [C.c_int, _u32p, C.c_int, C.c_int, _u32p]),
        "cufe_ps_crt_round_host": (C.c_int, [C.c_int, _u32p, C.POINTER(C.c_uint32)]),
        "cufe_time_ps_ntt": (C.c_int, [C.c_int, C.c_int, C.c_int, C.c_int, fp]),
        "cufe_time_ps_gauss": (C.c_int, [C.c_int, C.c_int, C.c_int, C.c_int, fp]),
    }
    for name, (res, args) in sig.items():
        f = getattr(lb, name)
        f.restype = res
        f.argtypes = args


def _pslib():
    lb = lib()
    if not getattr(lb, "_ps_sigs", False):
        _ps_sigs(lb)
        lb._ps_sigs = True
    return lb


class PSet:
    """One of the library's parameter sets: 0 (N = 2048, 2 limbs), 1 (the
    reference's N = 4096, 3 limbs), 2 (N = 8192, 4 limbs).  Arrays are
    [..][nmod][N] as for the reference set."""

    def __init__(self, pset: int):
        n, nm, l = C.c_int(), C.c_int(), C.c_int()
        q = (C.c_uint32 * 4)()
        p = C.c_uint32()
        _check(_pslib().cufe_pset_info(pset, C.byref(n), C.byref(nm), C.byref(l), q, C.byref(p)), "pset_info")
        self.id, self.N, self.nmod, self.L, self.p = pset, n.value, nm.value, l.value, p.value
        self.q = tuple(int(q[j]) for j in range(self.nmod))

    def ntt(self, polys, inverse: bool = False, variant: int | None = None) -> np.ndarray:
        a = np.array(polys, dtype=np.uint32, copy=True, order="C")
        npoly = a.size // (self.nmod * self.N)
        assert a.size == npoly * self.nmod * self.N
        if variant is None:
            _check(_pslib().cufe_ps_ntt(self.id, int(inverse), a, npoly), "ps_ntt")
        else:
            _check(_pslib().cufe_ps_ntt_variant(self.id, variant, int(inverse), a, npoly), "ps_ntt_variant")
        return a

    def ntt_variants(self) -> list:
        buf = C.create_string_buffer(4096)
        n = _pslib().cufe_ntt_variants(self.id, buf, 4096)
        _check(0 if n >= 0 else -1, "ntt_variants")
        return buf.value.decode().splitlines()[:n]

    def sample_polya(self, seed) -> np.ndarray:
        a = np.empty((self.nmod, self.N), np.uint32)
        _check(_pslib().cufe_ps_sample_polya(self.id, _seed(seed), a), "ps_sample_polya")
        return a

    def gauss_sample(self, keys, nonces, which: int, ncalls: int) -> np.ndarray:
        k = np.ascontiguousarray(np.frombuffer(b"".join(bytes(x) for x in keys), dtype=np.uint8))
        n = np.ascontiguousarray(nonces, dtype=np.uint16)
        out = np.empty((len(n), ncalls, self.nmod, self.N), np.uint32)
        _check(_pslib().cufe_ps_gauss_sample(self.id, k, n, len(n), which, ncalls, out), "ps_gauss_sample")
        return out

    def setup(self, a_coeff, seed2, seed3, l: int):
        mpk = np.zeros((l + 1, self.nmod, self.N), np.uint32)
        mpk[l] = _u32(a_coeff, (self.nmod, self.N))
        msk = np.zeros((l, self.nmod, self.N), np.uint32)
        _check(_pslib().cufe_ps_setup(self.id, mpk, msk, _seed(seed2), _seed(seed3), l), "ps_setup")
        return mpk, msk

    def encrypt(self, m, mpk, seed2, seed3, nonce0: int = 0) -> np.ndarray:
        m = _u32(np.atleast_2d(m))
        repeat, l = m.shape
        mpk = _u32(mpk)
        assert mpk.shape == (l + 1, self.nmod, self.N)
        c = np.empty((repeat, l + 1, self.nmod, self.N), np.uint32)
        _check(_pslib().cufe_ps_encrypt(self.id, m if m.size else np.zeros(1, np.uint32), mpk, c, _seed(seed2),
                                        _seed(seed3), repeat, l, nonce0), "ps_encrypt")
        return c

    def keygen(self, y, msk) -> np.ndarray:
        y = _u32(np.atleast_2d(y))
        repeat, l = y.shape
        sk = np.empty((repeat, self.nmod, self.N), np.uint32)
        _check(_pslib().cufe_ps_keygen(self.id, y, _u32(msk), sk, repeat, l), "ps_keygen")
        return sk

    def decrypt(self, c, y, sk) -> np.ndarray:
        y = _u32(np.atleast_2d(y))
        repeat, l = y.shape
        d = np.empty((repeat, self.nmod, self.N), np.uint32)
        _check(_pslib().cufe_ps_decrypt(self.id, _u32(c), y, _u32(sk), d, repeat, l), "ps_decrypt")
        return d

    def crt_round(self, d, coef: int = 0) -> np.ndarray:
        d = _u32(d)
        nvec = d.size // (self.nmod * self.N)
        r = np.empty((nvec,) if coef >= 0 else (nvec, self.N), np.uint32)
        _check(_pslib().cufe_ps_crt_round(self.id, d, nvec, coef, r), "ps_crt_round")
        return r

    def crt_round_host(self, res) -> int:
        r = C.c_uint32()
        _check(_pslib().cufe_ps_crt_round_host(self.id, _u32(res), C.byref(r)), "ps_crt_round_host")
        return r.value

    def time_ntt(self, npoly: int, inverse: bool, iters: int) -> float:
        ms = C.c_float()
        _check(_pslib().cufe_time_ps_ntt(self.id, npoly, int(inverse), iters, C.byref(ms)), "time_ps_ntt")
        return ms.value

    def time_gauss(self, nstreams: int, which: int, iters: int) -> float:
        ms = C.c_float()
        _check(_pslib().cufe_time_ps_gauss(self.id, nstreams, which, iters, C.byref(ms)), "time_ps_gauss")
        return ms.value


def ntt_autotune_report() -> str:
    buf = C.create_string_buffer(1 << 16)
    _pslib().cufe_ntt_autotune_report(buf, 1 << 16)
    return buf.value.decode()


class SvmModel:
    """Plain linear SVM model: the fields svm_predict_gpu reads (svm/svm.h:52-72),
    support vectors dense [l][785] (feature index = slot)."""

    def __init__(self, nr_class, nSV, label, sv_coef, rho, sv_dense):
        self.nr_class = int(nr_class)
        self.nSV = np.ascontiguousarray(nSV, dtype=np.int32)
        self.label = np.ascontiguousarray(label, dtype=np.int32)
        self.sv_coef = np.ascontiguousarray(sv_coef, dtype=np.float64).reshape(self.nr_class - 1, -1)
        self.rho = np.ascontiguousarray(rho, dtype=np.float64)
        sv = np.ascontiguousarray(sv_dense, dtype=np.uint32)
        if sv.shape[1] < L:
            sv = np.concatenate([sv, np.zeros((sv.shape[0], L - sv.shape[1]), np.uint32)], axis=1)
        self.sv = np.ascontiguousarray(sv)
        self.l = self.sv.shape[0]

    def as_struct(self):
        return _SvmModelC(self.nr_class, self.l, self.nSV.ctypes.data_as(C.POINTER(C.c_int)),
                          self.label.ctypes.data_as(C.POINTER(C.c_int)),
                          self.sv_coef.ctypes.data_as(C.POINTER(C.c_double)),
                          self.rho.ctypes.data_as(C.POINTER(C.c_double)),
                          self.sv.ctypes.data_as(C.POINTER(C.c_uint32)))


class _SvmModelC(C.Structure):
    _fields_ = [("nr_class", C.c_int), ("l", C.c_int), ("nSV", C.POINTER(C.c_int)), ("label", C.POINTER(C.c_int)),
                ("sv_coef", C.POINTER(C.c_double)), ("rho", C.POINTER(C.c_double)), ("sv", C.POINTER(C.c_uint32))]


def _ctx_sigs(lb):
    p = C.c_void_p
    sig = {
        "cufe_ctx_create": (C.c_int, [C.POINTER(C.c_void_p)]),
        "cufe_ctx_destroy": (None, [p]),
        "cufe_ctx_setup": (C.c_int, [p, _u8p, _u8p, _u8p, C.c_int]),
        "cufe_ctx_load_keys": (C.c_int, [p, _u32p, _u32p, C.c_int]),
        "cufe_ctx_get_keys": (C.c_int, [p, _u32p, _u32p]),
        "cufe_ctx_set_model": (C.c_int, [p, C.POINTER(_SvmModelC)]),
        "cufe_ctx_rows_reserve": (C.c_int, [p, C.c_int, C.c_int]),
        "cufe_ctx_encrypt_rows": (C.c_int, [p, _u32p, C.c_int, _u8p, _u8p, C.c_int, C.POINTER(C.c_float)]),
        "cufe_ctx_rows_count": (C.c_int, [p]),
        "cufe_ctx_predict": (C.c_int, [p, C.c_int, C.c_int, _f64p, C.c_void_p, C.POINTER(C.c_float)]),
        "cufe_ctx_model_info": (C.c_int, [p, C.POINTER(C.c_int), C.POINTER(C.c_uint)]),
        "cufe_ctx_get_model_keys": (C.c_int, [p, _u32p]),
        "cufe_ctx_get_row": (C.c_int, [p, C.c_int, _u32p, _u32p]),
        "cufe_ctx_export_rows": (C.c_int, [p, C.c_int, C.c_int, _u32p, _u32p]),
        "cufe_ctx_contract_digits": (C.c_int, [p]),
        "cufe_ctx_contract_limb_ms": (C.c_int, [p, C.POINTER(C.c_float)]),
        "cufe_ctx_contract_form": (C.c_int, [p, C.POINTER(C.c_ulonglong)]),
        "cufe_svm_predict_gpu": (C.c_double, [C.POINTER(_SvmModelC), _u32p, _u32p, C.c_void_p]),
    }
    for name, (res, args) in sig.items():
        f = getattr(lb, name)
        f.restype = res
        f.argtypes = args


class Context:
    """Resident keys, functional keys and encrypted rows on the GPU."""

    def __init__(self):
        lb = lib()
        if not getattr(lb, "_ctx_sigs", False):
            _ctx_sigs(lb)
            lb._ctx_sigs = True
        h = C.c_void_p()
        _check(lb.cufe_ctx_create(C.byref(h)), "ctx_create")
        self._h = h
        self.l_key = 0
        self.row_l = 0
        self.model = None

    def close(self):
        if self._h:
            lib().cufe_ctx_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def setup(self, seedA, seed2, seed3, l: int = L):
        _check(lib().cufe_ctx_setup(self._h, _seed(seedA), _seed(seed2), _seed(seed3), l), "ctx_setup")
        self.l_key = l

    def load_keys(self, mpk, msk):
        mpk, msk = _u32(mpk), _u32(msk)
        l = msk.shape[0]
        _check(lib().cufe_ctx_load_keys(self._h, mpk, msk, l), "ctx_load_keys")
        self.l_key = l

    def get_keys(self):
        mpk = np.empty((self.l_key + 1, NMODULI, N), np.uint32)
        msk = np.empty((self.l_key, NMODULI, N), np.uint32)
        _check(lib().cufe_ctx_get_keys(self._h, mpk, msk), "ctx_get_keys")
        return mpk, msk

    def set_model(self, model: SvmModel):
        self._mstruct = model.as_struct()
        _check(lib().cufe_ctx_set_model(self._h, C.byref(self._mstruct)), "ctx_set_model")
        self.model = model

    def rows_reserve(self, cap: int, l: int):
        _check(lib().cufe_ctx_rows_reserve(self._h, cap, l), "ctx_rows_reserve")
        self.row_l = l

    def encrypt_rows(self, x, seed2, seed3, nonce0: int = 0) -> float:
        x = _u32(np.atleast_2d(x))
        assert x.shape[1] == self.row_l
        ms = C.c_float()
        _check(lib().cufe_ctx_encrypt_rows(self._h, x, x.shape[0], _seed(seed2), _seed(seed3), nonce0, C.byref(ms)),
               "ctx_encrypt_rows")
        return ms.value

    def rows_count(self) -> int:
        return lib().cufe_ctx_rows_count(self._h)

    def predict(self, row0: int, nrows: int, want_kvalues: bool = False):
        labels = np.empty(nrows, np.float64)
        kv = np.empty((nrows, self.model.l), np.uint32) if want_kvalues else None
        ms = (C.c_float * 4)()
        _check(lib().cufe_ctx_predict(self._h, row0, nrows, labels, kv.ctypes.data if kv is not None else None, ms),
               "ctx_predict")
        return labels, kv, list(ms)

    def contract_limb_ms(self):
        """Contraction time per limb (ms) of the last predict call."""
        ms = (C.c_float * 3)()
        _check(lib().cufe_ctx_contract_limb_ms(self._h, ms), "ctx_contract_limb_ms")
        return list(ms)

    def contract_form(self):
        """(form, exact_pairs) of the last predict call: form 2 = the fraction
        form (every limb in one int8 contraction), 1 = int8 per limb, 0 = fp64;
        exact_pairs = the (row, key) pairs its exact kernel recomputed."""
        n = C.c_ulonglong()
        form = lib().cufe_ctx_contract_form(self._h, C.byref(n))
        _check(form if form < 0 else 0, "ctx_contract_form")
        return form, n.value

    def model_info(self):
        kb = (C.c_int * 3)()
        mx = (C.c_uint * 3)()
        _check(lib().cufe_ctx_model_info(self._h, kb, mx), "ctx_model_info")
        return list(kb), list(mx)

    def contract_digits(self) -> int:
        """Key digits of the int8 contraction (2 or 3), 0 for the fp64 kernels."""
        return lib().cufe_ctx_contract_digits(self._h)

    def model_keys(self):
        sk = np.empty((self.model.l, NMODULI, N), np.uint32)
        _check(lib().cufe_ctx_get_model_keys(self._h, sk), "ctx_get_model_keys")
        return sk

    def get_row(self, row: int):
        cl = np.empty((NMODULI, N), np.uint32)
        c0 = np.empty((self.row_l, NMODULI), np.uint32)
        _check(lib().cufe_ctx_get_row(self._h, row, cl, c0), "ctx_get_row")
        return cl, c0

    def export_rows(self, row0: int, n: int):
        """Compact ciphertexts of stored rows: c_L [n][3][N], c0 [n][l][3]."""
        cl = np.empty((n, NMODULI, N), np.uint32)
        c0 = np.empty((n, self.row_l, NMODULI), np.uint32)
        _check(lib().cufe_ctx_export_rows(self._h, row0, n, cl, c0), "ctx_export_rows")
        return cl, c0


def _encbatch_sigs(lb):
    p = C.c_void_p
    sig = {
        "cufe_encbatch_create": (C.c_int, [p, _i32p, C.c_int, C.POINTER(C.c_void_p)]),
        "cufe_encbatch_run": (C.c_int, [p, _u32p, _u8p, _u8p, C.c_int, C.c_int, C.POINTER(C.c_float)]),
        "cufe_encbatch_export": (C.c_int, [p, C.c_int, C.c_int, _u32p]),
        "cufe_encbatch_checksum": (C.c_int, [p, C.POINTER(C.c_uint64), C.POINTER(C.c_longlong)]),
        "cufe_encbatch_destroy": (None, [p]),
    }
    for name, (res, args) in sig.items():
        f = getattr(lb, name)
        f.restype = res
        f.argtypes = args


class EncBatch:
    """Batched encryption of messages of uneven lengths against a Context's keys;
    ciphertexts stay on the GPU (export() copies some back)."""

    def __init__(self, ctx: "Context", lens):
        lb = lib()
        if not getattr(lb, "_encbatch_sigs", False):
            _encbatch_sigs(lb)
            lb._encbatch_sigs = True
        self.lens = np.ascontiguousarray(lens, np.int32)
        self.offsets = np.concatenate([[0], np.cumsum(self.lens)]).astype(np.int64)
        h = C.c_void_p()
        _check(lb.cufe_encbatch_create(ctx._h, self.lens, int(self.lens.size), C.byref(h)), "encbatch_create")
        self._h = h
        self._ctx = ctx

    def run(self, x_packed, seed2, seed3, nonce0: int = 0, use_graph: bool = False) -> float:
        x = _u32(x_packed)
        assert x.size == self.offsets[-1]
        ms = C.c_float()
        _check(lib().cufe_encbatch_run(self._h, x, _seed(seed2), _seed(seed3), nonce0, int(use_graph), C.byref(ms)),
               "encbatch_run")
        return ms.value

    def export(self, b0: int, nb: int):
        """List of ciphertexts [l_b+1][3][N] for messages b0..b0+nb-1."""
        ls = self.lens[b0:b0 + nb]
        words = int(np.sum(ls.astype(np.int64) + 1)) * NMODULI * N
        flat = np.empty(words, np.uint32)
        _check(lib().cufe_encbatch_export(self._h, b0, nb, flat), "encbatch_export")
        out, o = [], 0
        for l in ls:
            k = (int(l) + 1) * NMODULI * N
            out.append(flat[o:o + k].reshape(int(l) + 1, NMODULI, N))
            o += k
        return out

    def checksum(self):
        s = C.c_uint64()
        info = (C.c_longlong * 4)()
        _check(lib().cufe_encbatch_checksum(self._h, C.byref(s), info), "encbatch_checksum")
        return s.value, dict(zip(("messages", "words", "groups", "redone"), list(info)))

    def close(self):
        if self._h:
            lib().cufe_encbatch_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def _server_sigs(lb):
    p = C.c_void_p
    pp = C.POINTER(C.c_void_p)
    sig = {
        "cufe_server_create": (C.c_int, [p, C.c_size_t, C.c_int, C.c_int, pp]),
        "cufe_server_add_model": (C.c_int, [p, C.POINTER(_SvmModelC), C.POINTER(C.c_int)]),
        "cufe_server_start": (C.c_int, [p]),
        "cufe_server_submit": (C.c_int, [p, C.c_int, C.c_int, C.c_int, _u32p, _u32p, _f64p, C.c_void_p,
                                         C.POINTER(C.c_uint64)]),
        "cufe_server_wait": (C.c_int, [p, C.c_uint64, C.c_double, C.POINTER(C.c_double)]),
        "cufe_server_replay": (C.c_int, [p, C.c_int, _f64p, _i32p, _i32p, _i32p, pp, pp, pp, _f64p,
                                         C.POINTER(C.c_double)]),
        "cufe_server_stats": (C.c_int, [p, C.POINTER(C.c_longlong)]),
        "cufe_server_reset_stats": (C.c_int, [p]),
        "cufe_server_info": (C.c_int, [p, C.POINTER(C.c_int), C.POINTER(C.c_longlong), C.POINTER(C.c_longlong),
                                       C.POINTER(C.c_longlong)]),
        "cufe_server_destroy": (None, [p]),
    }
    for name, (res, args) in sig.items():
        f = getattr(lb, name)
        f.restype = res
        f.argtypes = args


class Server:
    """Continuous-batching prediction server over the keys of a Context.

    Requests carry rows in compact ciphertext form (Context.export_rows):
    c_L [n][3][N] and coefficient 0 of c_0..c_{l-1} as [n][l][3]."""

    def __init__(self, ctx: Context, hbm_budget: int = 0, max_batch_rows: int = 0, batching: bool = True):
        lb = lib()
        if not getattr(lb, "_server_sigs", False):
            _server_sigs(lb)
            lb._server_sigs = True
        h = C.c_void_p()
        _check(lb.cufe_server_create(ctx._h, hbm_budget, max_batch_rows, int(bool(batching)), C.byref(h)),
               "server_create")
        self._h = h
        self._ctx = ctx
        self._models = []
        self._structs = []
        self._inflight = {}

    def add_model(self, model: SvmModel) -> int:
        st = model.as_struct()
        mid = C.c_int()
        _check(lib().cufe_server_add_model(self._h, C.byref(st), C.byref(mid)), "server_add_model")
        self._models.append(model)
        self._structs.append(st)
        return mid.value

    def start(self):
        _check(lib().cufe_server_start(self._h), "server_start")

    def submit(self, model: int, cl, c0, want_kvalues: bool = False) -> int:
        cl = _u32(cl)
        c0 = _u32(c0)
        n, l = c0.shape[0], c0.shape[1]
        assert cl.shape == (n, NMODULI, N) and c0.shape == (n, l, NMODULI)
        labels = np.empty(n, np.float64)
        kv = np.empty((n, self._models[model].l), np.uint32) if want_kvalues else None
        t = C.c_uint64()
        _check(lib().cufe_server_submit(self._h, model, n, l, cl, c0, labels,
                                        kv.ctypes.data if kv is not None else None, C.byref(t)), "server_submit")
        self._inflight[t.value] = (cl, c0, labels, kv)
        return t.value

    def wait(self, ticket: int, timeout_ms: float = -1.0):
        """(labels, kvalues, latency_ms), or None while pending; raises on failure."""
        lat = C.c_double()
        rc = lib().cufe_server_wait(self._h, ticket, timeout_ms, C.byref(lat))
        if rc == 0:
            return None
        bufs = self._inflight.pop(ticket)
        if rc < 0:
            raise CufeError(f"server_wait: {lib().cufe_last_error().decode()}")
        return bufs[2], bufs[3], lat.value

    def replay(self, arrival_s, model, rows):
        """Open-loop replay; rows[i] = (cl, c0) of request i.  Returns
        (labels list, latency_ms array, makespan_s, failed count)."""
        n = len(arrival_s)
        arr = np.ascontiguousarray(arrival_s, np.float64)
        mdl = np.ascontiguousarray(model, np.int32)
        keep = [(_u32(cl), _u32(c0)) for cl, c0 in rows]
        nr = np.array([c0.shape[0] for _, c0 in keep], np.int32)
        ls = np.array([c0.shape[1] for _, c0 in keep], np.int32)
        labels = [np.empty(k, np.float64) for k in nr]
        clp = (C.c_void_p * n)(*[a.ctypes.data for a, _ in keep])
        c0p = (C.c_void_p * n)(*[b.ctypes.data for _, b in keep])
        lbp = (C.c_void_p * n)(*[x.ctypes.data for x in labels])
        lat = np.empty(n, np.float64)
        mk = C.c_double()
        failed = lib().cufe_server_replay(self._h, n, arr, mdl, nr, ls, clp, c0p, lbp, lat, C.byref(mk))
        if failed < 0:
            raise CufeError(f"server_replay: {lib().cufe_last_error().decode()}")
        return labels, lat, mk.value, failed

    def stats(self) -> dict:
        out = (C.c_longlong * 8)()
        _check(lib().cufe_server_stats(self._h, out), "server_stats")
        keys = ("requests", "rows", "batches", "failed", "max_batch_rows", "cap_rows", "model_bytes", "slot_bytes")
        return dict(zip(keys, list(out)))

    def info(self) -> dict:
        """Slots, requests split into parts, peak HBM bytes of the batches in
        flight, and the device bytes one batch row occupies."""
        ns = C.c_int()
        sp, pk, br = C.c_longlong(), C.c_longlong(), C.c_longlong()
        _check(lib().cufe_server_info(self._h, C.byref(ns), C.byref(sp), C.byref(pk), C.byref(br)), "server_info")
        return {"slots": ns.value, "split_requests": sp.value, "peak_inflight_bytes": pk.value,
                "bytes_per_row": br.value}

    def reset_stats(self):
        """Zero the request, row, batch, failure and largest-batch counters."""
        _check(lib().cufe_server_reset_stats(self._h), "server_reset_stats")

    def close(self):
        if self._h:
            lib().cufe_server_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def svm_predict_plain(model: SvmModel, c_full, msk) -> tuple:
    """svm_predict_gpu (main.c:270-390) through the reference ABI on one full
    ciphertext c [786][3][N] with msk [785][3][N]: (label, kvalues)."""
    lb = lib()
    if not getattr(lb, "_ctx_sigs", False):
        _ctx_sigs(lb)
        lb._ctx_sigs = True
    ms = model.as_struct()
    kv = np.empty(model.l, np.float64)
    lab = lb.cufe_svm_predict_gpu(C.byref(ms), _u32(c_full), _u32(msk), kv.ctypes.data)
    msg = lb.cufe_last_error()
    return lab, kv
