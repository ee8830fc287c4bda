"""Chunk record format (no GPU): the library's reader against records built here from the format's
description (64-byte little-endian header, maps in bit order padded to 16 bytes, check = sum of the
payload's 32-bit words times (2 i + 1) modulo 2^32), the record sizes in closed form, the check in
closed form (a payload of words equal to 1 has check n^2 mod 2^32), and the reader's refusals."""
import ctypes as C
import struct

import numpy as np
import pytest

from superterrainplus_amd import capi


def pad(n):
    return (n + 15) // 16 * 16


def build(x, y, w, h, maps, biome=None, height=None, fixed=None, completeness=3, check=None):
    parts = []
    if maps & 1:
        parts.append(np.ascontiguousarray(biome, "<u2").tobytes())
    if maps & 2:
        parts.append(np.ascontiguousarray(height, "<f4").tobytes())
    if maps & 4:
        parts.append(np.ascontiguousarray(fixed, "<u2").tobytes())
    payload = b"".join(p + b"\0" * (pad(len(p)) - len(p)) for p in parts)
    words = np.frombuffer(payload, "<u4").astype(np.uint64)
    if check is None:
        check = int((words * (2 * np.arange(words.size, dtype=np.uint64) + 1) % (1 << 32)).sum() % (1 << 32))
    header = b"STPC" + struct.pack("<HHiiiiIIIIQI12x", 1, 64, x, y, x * w, y * h, w, h, completeness, maps, len(payload),
                                   check)
    assert len(header) == 64
    return header + payload


def record_bytes(w, h, maps):
    out = C.c_uint64()
    capi.call("stp_chunk_record_bytes", w, h, maps, C.byref(out))
    return out.value


def test_record_sizes():
    assert record_bytes(4, 4, 7) == 64 + 32 + 64 + 32
    assert record_bytes(3, 3, 1) == 64 + 32
    assert record_bytes(512, 512, 6) == 64 + 512 * 512 * 6
    with pytest.raises(capi.STPError):
        record_bytes(4, 4, 8)


def test_read_records_built_from_the_description():
    rng = np.random.default_rng(3)
    w, h = 13, 7
    recs = []
    for i, maps in enumerate((7, 1, 6, 4)):
        b = rng.integers(0, 9, (h, w)).astype(np.uint16)
        f = rng.normal(size=(h, w)).astype(np.float32)
        u = rng.integers(0, 65536, (h, w)).astype(np.uint16)
        recs.append((maps, b, f, u, build(i, -i, w, h, maps, b, f, u, completeness=i % 4)))
    stream = b"".join(r[-1] for r in recs)
    got = capi.read_chunk_records(stream)
    assert len(got) == 4
    for i, (g, (maps, b, f, u, raw)) in enumerate(zip(got, recs)):
        assert (g["x"], g["y"], g["maps"], g["completeness"]) == (i, -i, maps, i % 4)
        assert len(raw) == record_bytes(w, h, maps)
        for key, want, bit in (("biome", b, 1), ("height", f, 2), ("fixed", u, 4)):
            assert (g[key] is None) == (not maps & bit)
            if maps & bit:
                assert (g[key] == want).all()


def test_check_closed_form():
    # a payload of 32-bit words equal to 1: 16 x 16 UINT16 pixels of 1 and 0 alternately read as
    # little-endian words (1, 0) -> word 1; n = 128 words, check = n^2
    fixed = np.tile(np.array([1, 0], np.uint16), (16, 8))
    raw = build(0, 0, 16, 16, 4, fixed=fixed)
    assert struct.unpack_from("<I", raw, 48)[0] == 128 ** 2
    assert capi.read_chunk_records(raw)[0]["check"] == 128 ** 2


def test_reader_refuses_damaged_records():
    rng = np.random.default_rng(4)
    f = rng.normal(size=(8, 8)).astype(np.float32)
    good = build(1, 2, 8, 8, 2, height=f)
    capi.read_chunk_records(good)
    bad = bytearray(good)
    bad[100] ^= 1  # a payload bit
    with pytest.raises(capi.STPError):
        capi.read_chunk_records(bytes(bad))
    with pytest.raises(capi.STPError):
        capi.read_chunk_records(good[:-16])  # truncated
    with pytest.raises(capi.STPError):
        capi.read_chunk_records(b"STPX" + good[4:])  # magic
    with pytest.raises(capi.STPError):
        capi.read_chunk_records(build(1, 2, 8, 8, 2, height=f, check=12345))


def test_reader_refuses_sizes_beyond_the_side_limit():
    """A header whose width x height x 4 wraps 64 bits (2^31 x 2^31, FP32 and biome maps: the wrapped
    payload size is 0) must not pass as a record of no pixels; sides above 65536 are refused, by
    the reader and by the size query."""
    header = b"STPC" + struct.pack("<HHiiiiIIIIQI12x", 1, 64, 0, 0, 0, 0, 1 << 31, 1 << 31, 3, 3, 0, 0)
    with pytest.raises(capi.STPError):
        capi.read_chunk_records(header)
    header = b"STPC" + struct.pack("<HHiiiiIIIIQI12x", 1, 64, 0, 0, 0, 0, 65537, 1, 3, 4, pad(65537 * 2), 0)
    with pytest.raises(capi.STPError):
        capi.read_chunk_records(header + b"\0" * pad(65537 * 2))
    with pytest.raises(capi.STPError):
        record_bytes(65537, 1, 1)
    assert record_bytes(65536, 1, 4) == 64 + 65536 * 2
