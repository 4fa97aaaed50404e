"""Python handle on the checker libraries.  TEST INFRASTRUCTURE ONLY: imported by
tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg, never by the
product (cufe_amd).

  liboracle.so   -- ipfe_oracle.c, the CPU restatement of the reference path
  _ref/libref.so -- the reference's own C path (src/*.c, main.c) compiled
                    against GMP, plus ref_harness.c; built where the
                    reference is mounted (it travels with the tree once built)
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ORACLE_SO = os.path.join(HERE, "liboracle.so")
REF_SO = os.path.join(HERE, "_ref", "libref.so")
N = 4096

_u32p = np.ctypeslib.ndpointer(dtype=np.uint32, flags="C_CONTIGUOUS")
_u8p = np.ctypeslib.ndpointer(dtype=np.uint8, flags="C_CONTIGUOUS")
_f64p = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")
_i32p = np.ctypeslib.ndpointer(dtype=np.int32, flags="C_CONTIGUOUS")


class Stream(C.Structure):
    _fields_ = [("rk", C.c_uint32 * 60), ("ctr", C.c_uint32), ("nonce", C.c_uint16)]


def build() -> None:
    """Compile liboracle.so (and _ref when the reference sources are present)."""
    subprocess.run(["make", "-s", "-C", HERE, "oracle"], check=True)
    if os.path.isdir("/root/reference/src"):
        subprocess.run(["make", "-s", "-C", HERE, "ref"], check=True)


_orc = None
_ref = None


def orc() -> C.CDLL:
    global _orc
    if _orc is None:
        if not os.path.exists(ORACLE_SO):
            build()
        o = C.CDLL(ORACLE_SO)
        o.orc_mod_prime.restype = C.c_uint32
        o.orc_mod_prime.argtypes = [C.c_uint64, C.c_uint32]
        for f in ("orc_add_mod", "orc_sub_mod", "orc_mul_mod"):
            getattr(o, f).restype = C.c_uint32
            getattr(o, f).argtypes = [C.c_uint32, C.c_uint32, C.c_uint32]
        for f in ("orc_modulus", "orc_ninv", "orc_scale_mod"):
            getattr(o, f).restype = C.c_uint32
            getattr(o, f).argtypes = [C.c_uint32]
        o.orc_psi.restype = C.POINTER(C.c_uint32)
        o.orc_psi_inv.restype = C.POINTER(C.c_uint32)
        o.orc_stream_init.argtypes = [C.POINTER(Stream), _u8p, C.c_uint16]
        o.orc_stream_squeeze.argtypes = [C.POINTER(Stream), _u8p, C.c_uint64]
        o.orc_aes256_expand.argtypes = [_u8p, _u32p]
        o.orc_aes256_encrypt_block.argtypes = [_u32p, _u8p, _u8p]
        o.orc_sample_polya.argtypes = [_u8p, _u32p]
        o.orc_gauss.argtypes = [C.POINTER(Stream), C.c_int, _u32p, C.c_uint32]
        o.orc_gauss_chunk.argtypes = [_u8p, C.c_int, np.ctypeslib.ndpointer(np.uint64, flags="C_CONTIGUOUS"), _u8p, _u8p]
        o.orc_ct_forward.argtypes = [_u32p, C.c_uint32]
        o.orc_gs_reverse.argtypes = [_u32p, C.c_uint32]
        o.orc_poly_mul_ntt.argtypes = [_u32p, _u32p, _u32p, C.c_uint32]
        o.orc_setup.argtypes = [_u8p, _u8p, _u8p, C.c_int, _u32p, _u32p]
        o.orc_setup_from_a.argtypes = [_u32p, _u8p, _u8p, C.c_int, _u32p, _u32p]
        o.orc_encrypt.argtypes = [_u32p, C.c_int, _u32p, _u8p, _u8p, C.c_uint16, _u32p]
        o.orc_keygen.argtypes = [_u32p, C.c_int, _u32p, _u32p]
        o.orc_decrypt.argtypes = [_u32p, _u32p, C.c_int, _u32p, _u32p]
        o.orc_crt_reverse.argtypes = [C.c_uint32, C.c_uint32, C.c_uint32, C.POINTER(C.c_uint64), C.POINTER(C.c_uint64)]
        o.orc_round_extract.restype = C.c_uint32
        o.orc_round_extract.argtypes = [C.c_uint64, C.c_uint64]
        o.orc_round_extract_coef0.restype = C.c_double
        o.orc_round_extract_coef0.argtypes = [_u32p]
        o.orc_svm_decide.restype = C.c_double
        o.orc_svm_decide.argtypes = [C.c_int, C.c_int, _i32p, _i32p, _f64p, _f64p, _f64p, C.c_int]
        o.orc_kvalues_coef0.restype = C.c_int
        o.orc_kvalues_coef0.argtypes = [_u32p, _u32p, C.c_int, C.c_int, _u32p, C.c_int, _u32p, C.c_int, _u32p]
        psp = C.POINTER(PSet)
        o.orc_pset_get.argtypes = [C.c_int, psp]
        o.orc_ps_mod_prime.restype = C.c_uint32
        o.orc_ps_mod_prime.argtypes = [psp, C.c_uint64, C.c_int]
        o.orc_ps_psi.argtypes = [psp, C.c_int, _u32p, _u32p]
        o.orc_ps_ninv.restype = C.c_uint32
        o.orc_ps_ninv.argtypes = [psp, C.c_int]
        o.orc_ps_scale.restype = C.c_uint32
        o.orc_ps_scale.argtypes = [psp, C.c_int]
        o.orc_ps_ct_forward.argtypes = [psp, _u32p, C.c_int]
        o.orc_ps_gs_reverse.argtypes = [psp, _u32p, C.c_int]
        o.orc_ps_sample_polya.argtypes = [psp, _u8p, _u32p]
        o.orc_ps_gauss.argtypes = [psp, C.POINTER(Stream), C.c_int, _u32p]
        o.orc_ps_setup_from_a.argtypes = [psp, _u32p, _u8p, _u8p, C.c_int, _u32p, _u32p]
        o.orc_ps_encrypt.argtypes = [psp, _u32p, C.c_int, _u32p, _u8p, _u8p, C.c_uint16, _u32p]
        o.orc_ps_keygen.argtypes = [psp, _u32p, C.c_int, _u32p, _u32p]
        o.orc_ps_decrypt.argtypes = [psp, _u32p, _u32p, C.c_int, _u32p, _u32p]
        o.orc_ps_crt.argtypes = [psp, _u32p, C.POINTER(C.c_uint64), C.POINTER(C.c_uint64)]
        o.orc_ps_round.restype = C.c_uint32
        o.orc_ps_round.argtypes = [psp, C.c_uint64, C.c_uint64]
        o.orc_svm_predict_cpu.restype = C.c_double
        o.orc_svm_predict_cpu.argtypes = [_u32p, C.c_int, _u32p, C.c_int, C.c_int, _i32p, _i32p, _f64p, _f64p, _u32p,
                                          _f64p, C.c_int]
        _orc = o
    return _orc


def ref():
    """The reference's own compiled subset, or None when it is not built here."""
    global _ref
    if _ref is None and os.path.exists(REF_SO):
        # lazy binding: the reference functions that call its CUDA module
        # (never called here) keep their symbols unresolved
        r = C.CDLL(REF_SO, mode=os.RTLD_LAZY | os.RTLD_LOCAL)
        r.sample_polya.argtypes = [_u8p, _u32p]
        r.round_extract_gmp2.restype = C.c_double
        r.round_extract_gmp2.argtypes = [_u32p]
        r.rlwe_sife_keygen.argtypes = [_u32p, _u32p, _u32p, C.POINTER(C.c_double)]
        r.refh_crt_round.argtypes = [_u32p, _u32p, _u32p]
        r.refh_decrypt_round.argtypes = [_u32p, _u32p, _u32p, _u32p]
        r.refh_svm_predict_cpu.restype = C.c_double
        r.refh_svm_predict_cpu.argtypes = [_u32p, _u32p, C.c_int, C.c_int, _i32p, _i32p, _f64p, _f64p, _u32p]
        r.mod_prime.restype = C.c_uint32
        r.mod_prime.argtypes = [C.c_uint64, C.c_uint32]
        r.aes256ctr_init.argtypes = [C.c_void_p, _u8p, C.c_uint16]
        r.aes256ctr_squeezeblocks.argtypes = [_u8p, C.c_ulonglong, C.c_void_p]
        for f in ("gaussian_sampler_S1_ori", "gaussian_sampler_S2_ori", "gaussian_sampler_S3"):
            getattr(r, f).argtypes = [C.c_void_p, _u32p, C.c_uint32]
        r.CT_forward.argtypes = [_u32p, C.c_uint32]
        r.GS_reverse.argtypes = [_u32p, C.c_uint32]
        r.point_mul.argtypes = [_u32p, _u32p, _u32p, C.c_uint32]
        r.poly_add_mod.argtypes = [_u32p, _u32p, _u32p, C.c_uint32]
        r.poly_sub_mod.argtypes = [_u32p, _u32p, _u32p, C.c_uint32]
        r.poly_mul_ntt.argtypes = [_u32p, _u32p, _u32p, C.c_uint32]
        _ref = r
    return _ref


def kvalues_coef0(cl, c0, y, sk) -> np.ndarray:
    """kvalue of every (row, key) pair from coefficient 0 of the decryption
    (orc_kvalues_coef0): cl [R][3][N], c0 [R][l][3], y [S][ly], sk [S][3][N]
    -> [R][S] uint32.  Raises if a pair depends on the INTT's representative."""
    cl = np.ascontiguousarray(cl, np.uint32)
    c0 = np.ascontiguousarray(c0, np.uint32)
    y = np.ascontiguousarray(y, np.uint32)
    sk = np.ascontiguousarray(sk, np.uint32)
    R, l, S, ly = c0.shape[0], c0.shape[1], y.shape[0], y.shape[1]
    assert cl.shape == (R, 3, N) and sk.shape == (S, 3, N)
    kv = np.empty((R, S), np.uint32)
    amb = orc().orc_kvalues_coef0(cl, c0, R, l, y, ly, sk, S, kv)
    if amb:
        raise ValueError(f"{amb} pairs end their fold on a non-zero slot")
    return kv


def seed_bytes(s) -> np.ndarray:
    return np.frombuffer(bytes(s), dtype=np.uint8).copy()


# ---- oracle conveniences ------------------------------------------------------
def stream(key, nonce: int) -> Stream:
    st = Stream()
    orc().orc_stream_init(C.byref(st), seed_bytes(key), nonce)
    return st


def keystream(key, nonce: int, nblocks128: int) -> bytes:
    st = stream(key, nonce)
    out = np.empty(128 * nblocks128, dtype=np.uint8)
    orc().orc_stream_squeeze(C.byref(st), out, nblocks128)
    return out.tobytes()


def gauss_calls(key, nonce: int, which: int, ncalls: int) -> np.ndarray:
    st = stream(key, nonce)
    out = np.empty((ncalls, 3, N), dtype=np.uint32)
    for i in range(ncalls):
        buf = np.empty((3, N), dtype=np.uint32)
        orc().orc_gauss(C.byref(st), which, buf, N)
        out[i] = buf
    return out


def ntt_forward(polys) -> np.ndarray:
    a = np.ascontiguousarray(polys, dtype=np.uint32).copy().reshape(-1, 3, N)
    for p in range(a.shape[0]):
        for s in range(3):
            row = np.ascontiguousarray(a[p, s])
            orc().orc_ct_forward(row, s)
            a[p, s] = row
    return a


def ntt_inverse(polys) -> np.ndarray:
    a = np.ascontiguousarray(polys, dtype=np.uint32).copy().reshape(-1, 3, N)
    for p in range(a.shape[0]):
        for s in range(3):
            row = np.ascontiguousarray(a[p, s])
            orc().orc_gs_reverse(row, s)
            a[p, s] = row
    return a


def sample_polya(seed) -> np.ndarray:
    a = np.empty((3, N), dtype=np.uint32)
    orc().orc_sample_polya(seed_bytes(seed), a)
    return a


def setup(seedA, seedS, seedE, l: int):
    mpk = np.zeros((l + 1, 3, N), dtype=np.uint32)
    msk = np.zeros((max(l, 1), 3, N), dtype=np.uint32)
    orc().orc_setup(seed_bytes(seedA), seed_bytes(seedS), seed_bytes(seedE), l, mpk, msk)
    return mpk, msk[:l]


def setup_from_a(a, seedS, seedE, l: int):
    mpk = np.zeros((l + 1, 3, N), dtype=np.uint32)
    msk = np.zeros((max(l, 1), 3, N), dtype=np.uint32)
    orc().orc_setup_from_a(np.ascontiguousarray(a, dtype=np.uint32), seed_bytes(seedS), seed_bytes(seedE), l, mpk, msk)
    return mpk, msk[:l]


def encrypt(x, mpk, seed2, seed3, nonce: int) -> np.ndarray:
    x = np.ascontiguousarray(x, dtype=np.uint32)
    l = x.size
    c = np.empty((l + 1, 3, N), dtype=np.uint32)
    orc().orc_encrypt(x if l else np.zeros(1, np.uint32), l, np.ascontiguousarray(mpk, dtype=np.uint32),
                      seed_bytes(seed2), seed_bytes(seed3), nonce, c)
    return c


def keygen(y, msk) -> np.ndarray:
    y = np.ascontiguousarray(y, dtype=np.uint32)
    sk = np.empty((3, N), dtype=np.uint32)
    orc().orc_keygen(y, y.size, np.ascontiguousarray(msk, dtype=np.uint32), sk)
    return sk


def decrypt(c, y, sk) -> np.ndarray:
    y = np.ascontiguousarray(y, dtype=np.uint32)
    d = np.empty((3, N), dtype=np.uint32)
    orc().orc_decrypt(np.ascontiguousarray(c, dtype=np.uint32), y, y.size, np.ascontiguousarray(sk, dtype=np.uint32), d)
    return d


def crt_round(d0: int, d1: int, d2: int) -> int:
    lo, hi = C.c_uint64(), C.c_uint64()
    orc().orc_crt_reverse(d0, d1, d2, C.byref(lo), C.byref(hi))
    return int(orc().orc_round_extract(lo.value, hi.value))


def crt_value(d0: int, d1: int, d2: int) -> int:
    lo, hi = C.c_uint64(), C.c_uint64()
    orc().orc_crt_reverse(d0, d1, d2, C.byref(lo), C.byref(hi))
    return (hi.value << 64) | lo.value


# ---- parameter sets (pset_oracle.c) ---------------------------------------------
class PSet(C.Structure):
    _fields_ = [("id", C.c_int), ("N", C.c_int), ("logN", C.c_int), ("nmod", C.c_int), ("L", C.c_int),
                ("q", C.c_uint32 * 4), ("fa", C.c_uint32 * 4), ("fb", C.c_uint32 * 4), ("root", C.c_uint32 * 4),
                ("k", C.c_uint64 * 4), ("p", C.c_uint32)]

    @property
    def qs(self):
        return [int(self.q[j]) for j in range(self.nmod)]

    @property
    def Q(self):
        out = 1
        for q in self.qs:
            out *= q
        return out


_ps_cache = {}


def pset(pid: int) -> PSet:
    if pid not in _ps_cache:
        ps = PSet()
        if orc().orc_pset_get(pid, C.byref(ps)) != 0:
            raise ValueError(f"no parameter set {pid}")
        _ps_cache[pid] = ps
    return _ps_cache[pid]


def _psp(pid):
    return C.byref(pset(pid))


def ps_mod_prime(pid: int, m: int, j: int) -> int:
    return int(orc().orc_ps_mod_prime(_psp(pid), m, j))


def ps_psi(pid: int, j: int):
    ps = pset(pid)
    a, b = np.empty(ps.N, np.uint32), np.empty(ps.N, np.uint32)
    orc().orc_ps_psi(_psp(pid), j, a, b)
    return a, b


def ps_ntt_forward(pid: int, polys) -> np.ndarray:
    """[..][nmod][N] -> CT_forward of every limb."""
    ps = pset(pid)
    a = np.array(polys, dtype=np.uint32, copy=True).reshape(-1, ps.nmod, ps.N)
    for p in range(a.shape[0]):
        for j in range(ps.nmod):
            row = np.ascontiguousarray(a[p, j])
            orc().orc_ps_ct_forward(_psp(pid), row, j)
            a[p, j] = row
    return a.reshape(np.shape(polys))


def ps_ntt_inverse(pid: int, polys) -> np.ndarray:
    ps = pset(pid)
    a = np.array(polys, dtype=np.uint32, copy=True).reshape(-1, ps.nmod, ps.N)
    for p in range(a.shape[0]):
        for j in range(ps.nmod):
            row = np.ascontiguousarray(a[p, j])
            orc().orc_ps_gs_reverse(_psp(pid), row, j)
            a[p, j] = row
    return a.reshape(np.shape(polys))


def ps_sample_polya(pid: int, seed) -> np.ndarray:
    ps = pset(pid)
    a = np.empty((ps.nmod, ps.N), np.uint32)
    orc().orc_ps_sample_polya(_psp(pid), seed_bytes(seed), a)
    return a


def ps_gauss_calls(pid: int, key, nonce: int, which: int, ncalls: int) -> np.ndarray:
    ps = pset(pid)
    st = stream(key, nonce)
    out = np.empty((ncalls, ps.nmod, ps.N), np.uint32)
    for c in range(ncalls):
        buf = np.empty((ps.nmod, ps.N), np.uint32)
        orc().orc_ps_gauss(_psp(pid), C.byref(st), which, buf)
        out[c] = buf
    return out


def ps_setup_from_a(pid: int, a, seedS, seedE, l: int):
    ps = pset(pid)
    mpk = np.zeros((l + 1, ps.nmod, ps.N), np.uint32)
    msk = np.zeros((max(l, 1), ps.nmod, ps.N), np.uint32)
    orc().orc_ps_setup_from_a(_psp(pid), np.ascontiguousarray(a, np.uint32), seed_bytes(seedS), seed_bytes(seedE),
                              l, mpk, msk)
    return mpk, msk[:l]


def ps_encrypt(pid: int, x, mpk, seed2, seed3, nonce: int) -> np.ndarray:
    ps = pset(pid)
    x = np.ascontiguousarray(x, np.uint32)
    l = x.size
    c = np.empty((l + 1, ps.nmod, ps.N), np.uint32)
    orc().orc_ps_encrypt(_psp(pid), x if l else np.zeros(1, np.uint32), l, np.ascontiguousarray(mpk, np.uint32),
                         seed_bytes(seed2), seed_bytes(seed3), nonce, c)
    return c


def ps_keygen(pid: int, y, msk) -> np.ndarray:
    ps = pset(pid)
    y = np.ascontiguousarray(y, np.uint32)
    sk = np.empty((ps.nmod, ps.N), np.uint32)
    orc().orc_ps_keygen(_psp(pid), y, y.size, np.ascontiguousarray(msk, np.uint32), sk)
    return sk


def ps_decrypt(pid: int, c, y, sk) -> np.ndarray:
    ps = pset(pid)
    y = np.ascontiguousarray(y, np.uint32)
    d = np.empty((ps.nmod, ps.N), np.uint32)
    orc().orc_ps_decrypt(_psp(pid), np.ascontiguousarray(c, np.uint32), y, y.size, np.ascontiguousarray(sk, np.uint32),
                         d)
    return d


def ps_crt_value(pid: int, res) -> int:
    lo, hi = C.c_uint64(), C.c_uint64()
    orc().orc_ps_crt(_psp(pid), np.ascontiguousarray(res, np.uint32), C.byref(lo), C.byref(hi))
    return (hi.value << 64) | lo.value


def ps_crt_round(pid: int, res) -> int:
    lo, hi = C.c_uint64(), C.c_uint64()
    orc().orc_ps_crt(_psp(pid), np.ascontiguousarray(res, np.uint32), C.byref(lo), C.byref(hi))
    return int(orc().orc_ps_round(_psp(pid), lo.value, hi.value))


# ---- reference-library conveniences ---------------------------------------------
class RefStream:
    """An aes256ctr_ctx of the reference library (272 bytes, 16-byte aligned)."""

    def __init__(self, key, nonce: int):
        self._raw = (C.c_uint8 * (272 + 64))()
        addr = C.addressof(self._raw)
        self.ptr = C.c_void_p((addr + 63) // 64 * 64)
        ref().aes256ctr_init(self.ptr, seed_bytes(key), nonce)

    def squeeze(self, nblocks128: int) -> bytes:
        out = np.empty(128 * nblocks128, dtype=np.uint8)
        ref().aes256ctr_squeezeblocks(out, nblocks128, self.ptr)
        return out.tobytes()

    def gauss(self, which: int) -> np.ndarray:
        fn = {1: ref().gaussian_sampler_S1_ori, 2: ref().gaussian_sampler_S2_ori, 3: ref().gaussian_sampler_S3}[which]
        out = np.empty((3, N), dtype=np.uint32)
        fn(self.ptr, out, N)
        return out
