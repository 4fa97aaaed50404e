"""Wire format for requests, ciphertexts and keys (host code, no GPU needed):
round trips of every kind, the header at its documented offsets, the payload
CRC against zlib's CRC-32, and rejection of corrupt, truncated, foreign and
out-of-range blobs."""
import struct
import zlib

import numpy as np
import pytest

N = 4096
Q = (16760833, 2147352577, 2130706433)


@pytest.fixture(scope="module")
def cf():
    import cufe_amd

    cufe_amd.lib()
    return cufe_amd


def residues(rng, shape_prefix, limb_axis_last=False):
    """uint32 residues below q of each limb; limb axis is the second-to-last
    ([..., 3, N]) or the last ([..., 3])."""
    if limb_axis_last:
        a = np.stack([rng.integers(0, Q[s], shape_prefix, dtype=np.uint64) for s in range(3)], axis=-1)
    else:
        a = np.stack([rng.integers(0, Q[s], shape_prefix + (N,), dtype=np.uint64) for s in range(3)], axis=-2)
    return a.astype(np.uint32)


def test_rows_round_trip_and_header(cf):
    rng = np.random.default_rng(1)
    n, l = 5, 124
    cl = residues(rng, (n,))
    c0 = residues(rng, (n, l), limb_axis_last=True)
    cl[0, 1, 7] = Q[1]  # the representative q travels unchanged
    blob = cf.wire_pack(cf.WIRE_ROWS, cl, c0, tag=77)
    assert blob.size == 64 + 4 * (n * 3 * N + n * l * 3) == cf.lib().cufe_wire_size(1, n, l)
    hdr = blob[:64].tobytes()
    magic, ver, kind, limbs, _, nn, count, ll, q0, q1, q2, tag, crc, nbytes = struct.unpack("<8sHHHHIII3IIIQ", hdr[:56])
    assert (magic, ver, kind, limbs, nn, count, ll, (q0, q1, q2), tag) == (b"CUFEIPFE", 1, 1, 3, N, n, l, Q, 77)
    assert nbytes == blob.size - 64
    assert crc == zlib.crc32(blob[64:].tobytes())
    p = cf.wire_parse(blob)
    assert p["kind"] == 1 and p["count"] == n and p["l"] == l and p["tag"] == 77
    assert np.array_equal(p["cl"], cl) and np.array_equal(p["c0"], c0)
    assert np.array_equal(cf.wire_parse(blob.tobytes())["cl"], cl)


@pytest.mark.parametrize("kind", [2, 3, 4])
def test_other_kinds_round_trip(cf, kind):
    rng = np.random.default_rng(kind)
    if kind == 2:
        data = residues(rng, (2, 8))     # 2 ciphertexts of l = 7
    elif kind == 3:
        data = residues(rng, (6,))       # public key of l = 5
    else:
        data = residues(rng, (3,))       # 3 functional keys
    blob = cf.wire_pack(kind, data)
    assert struct.unpack("<I", blob[44:48].tobytes())[0] == zlib.crc32(blob[64:].tobytes())
    p = cf.wire_parse(blob)
    assert p["kind"] == kind and np.array_equal(p["data"], data)


def test_rejects_bad_blobs(cf):
    from cufe_amd import CufeError

    rng = np.random.default_rng(3)
    cl = residues(rng, (2,))
    c0 = residues(rng, (2, 10), limb_axis_last=True)
    blob = cf.wire_pack(cf.WIRE_ROWS, cl, c0)
    bad = blob.copy()
    bad[64 + 12345] ^= 1  # payload bit flip
    with pytest.raises(CufeError, match="CRC"):
        cf.wire_parse(bad)
    with pytest.raises(CufeError):
        cf.wire_parse(blob[:-4])  # truncated
    bad = blob.copy()
    bad[0] = ord("X")
    with pytest.raises(CufeError, match="magic"):
        cf.wire_parse(bad)
    bad = blob.copy()
    bad[8] = 2  # version
    with pytest.raises(CufeError, match="version"):
        cf.wire_parse(bad)
    bad = blob.copy()
    bad[28:32] = np.frombuffer(struct.pack("<I", 12289), np.uint8)  # another modulus
    with pytest.raises(CufeError, match="parameter set"):
        cf.wire_parse(bad)
    # a word above q with a matching CRC is still refused
    bad = blob.copy()
    bad[64:68] = np.frombuffer(struct.pack("<I", Q[0] + 1), np.uint8)
    bad[44:48] = np.frombuffer(struct.pack("<I", zlib.crc32(bad[64:].tobytes())), np.uint8)
    with pytest.raises(CufeError, match="residue"):
        cf.wire_parse(bad)
    cl2 = cl.copy()
    cl2[1, 2, 3] = Q[2] + 5
    with pytest.raises(CufeError):
        cf.wire_pack(cf.WIRE_ROWS, cl2, c0)
