"""The C-ABI library: it loads, exports every symbol include/stp_amd.h declares, and its host-side
entry points (no device work) agree with the oracle and the reference's test data (CPU only)."""
import ctypes as C
import math
import os
import re
import subprocess

import numpy as np
import pytest

from superterrainplus_amd import capi, workloads as W

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    names = set()
    for f in os.listdir(os.path.join(ROOT, "include")):
        if f.endswith(".h"):
            text = open(os.path.join(ROOT, "include", f)).read()
            names |= set(re.findall(r"^\s*(?:int|const char\*)\s+(stp_[a-z0-9_]+)\s*\(", text, re.M))
    return names


def test_library_exports_every_declared_symbol(stp):
    declared = declared_symbols()
    assert len(declared) >= 40
    out = subprocess.run(["nm", "-D", "--defined-only", capi.LIB_PATH], capture_output=True, text=True,
                         check=True).stdout
    exported = set(re.findall(r" T (stp_[a-z0-9_]+)$", out, re.M))
    assert declared <= exported, declared - exported
    for name in declared:
        assert hasattr(stp.lib(), name)
    assert set(capi._SIGNATURES) == declared


def test_permutation_matches_oracle(stp, oracle):
    for eng in (0, 1):
        for seed in (0, 42, 6666, 2 ** 63 + 5):
            p, g = capi.permutation_host(seed, 12, W.GRADIENT_OFFSET, eng)
            po, go, _ = oracle.permutation(seed, 12, W.GRADIENT_OFFSET, eng)
            assert (p == po).all() and (g == go).all()


def test_permutation_reference_kat_through_abi(stp):
    _, g = capi.permutation_host(6666, 10, math.radians(6.5), 0)
    eps = float(np.finfo(np.float32).eps)
    for k, deg in ((0, 6.5), (3, 114.5), (7, 258.5)):
        assert abs(g[2 * k] - math.cos(math.radians(deg))) <= 10 * eps
        assert abs(g[2 * k + 1] - math.sin(math.radians(deg))) <= 10 * eps


@pytest.mark.parametrize("dim,r", [(512, 3), (1536, 3), (97, 4), (2048, 1), (64, 6)])
def test_brush_matches_oracle(stp, oracle, dim, r):
    idx, w = capi.brush(dim, r)
    io, wo = oracle.brush(dim, r)
    assert (idx == io).all() and (w == wo).all()


def test_settings_validation(stp):
    s = W.erosion_setting(100)
    capi.call("stp_raindrop_setting_validate", C.byref(s))
    for field, bad in (("RainDropCount", 0), ("Inertia", 1.5), ("Gravity", 0.0), ("ErosionBrushRadius", 0),
                       ("Friction", -0.1)):
        b = W.erosion_setting(100)
        setattr(b, field, bad)
        with pytest.raises(capi.STPError) as e:
            capi.call("stp_raindrop_setting_validate", C.byref(b))
        assert e.value.code == 1
    # unbounded drop life is rejected (no evaporation, no lifetime cap)
    b = W.erosion_setting(100, lifetime=0)
    b.EvaporateSpeed = 0.0
    with pytest.raises(capi.STPError):
        capi.call("stp_raindrop_setting_validate", C.byref(b))
    n = capi.SimplexNoiseSetting(1, 0, 0.0, 0)
    with pytest.raises(capi.STPError) as e:
        capi.call("stp_simplex_noise_setting_validate", C.byref(n))
    assert e.value.code == 1
    c = capi.ChunkSetting()
    c.ChunkSize[:] = [8, 8]
    c.MapSize[:] = [512, 512]
    c.ChunkScale[:] = [1.0, 1.0]
    for f in ("DiversityNearestNeighbour", "ErosionNearestNeighbour", "SplatNearestNeighbour", "RenderDistance"):
        getattr(c, f)[:] = [3, 3]
    capi.call("stp_chunk_setting_validate", C.byref(c))
    c.ErosionNearestNeighbour[:] = [2, 3]
    with pytest.raises(capi.STPError):
        capi.call("stp_chunk_setting_validate", C.byref(c))


def test_chunk_math_reference_kat(stp):
    # SuperTest+/SuperTerrain+/STPTest2DChunk.cpp:30-89
    out = (C.c_int32 * 2)()
    capi.call("stp_chunk_world_coordinate", (C.c_double * 3)(-573.74, 679.5, 845.982), (C.c_uint32 * 2)(8, 5),
              (C.c_double * 2)(2.5, 4.0), out)
    assert list(out) == [-232, 210]
    d = (C.c_double * 2)()
    capi.call("stp_chunk_map_offset", (C.c_int32 * 2)(-232, 210), (C.c_uint32 * 2)(8, 5), (C.c_uint32 * 2)(512, 256),
              (C.c_double * 2)(25.5, -57.5), d)
    assert list(d) == [-14822.5, 10694.5]
    capi.call("stp_chunk_offset", (C.c_int32 * 2)(30, -30), (C.c_uint32 * 2)(8, 5), (C.c_int32 * 2)(3, -2), out)
    assert list(out) == [54, -40]
    u = (C.c_uint32 * 2)()
    for cid, want in ((7, [2, 1]), (12, [2, 2]), (20, [0, 4])):
        capi.call("stp_chunk_local_coordinate", cid, (C.c_uint32 * 2)(5, 7), u)
        assert list(u) == want
    capi.call("stp_chunk_local_origin", (C.c_int32 * 2)(-24, 50), (C.c_uint32 * 2)(8, 5), (C.c_uint32 * 2)(5, 7), out)
    assert list(out) == [-40, 35]
    arr = (C.c_int32 * 70)()
    capi.call("stp_chunk_neighbour_offset", (C.c_uint32 * 2)(8, 5), (C.c_uint32 * 2)(5, 7), arr, 35)
    pairs = [tuple(arr[2 * i:2 * i + 2]) for i in range(35)]
    assert pairs[0] == (-16, -15) and pairs[9] == (16, -10) and pairs[17] == (0, 0)
    assert pairs[23] == (8, 5) and pairs[34] == (16, 15)


def test_errors_are_codes_not_crashes(stp):
    # no device in the CPU container: device calls report STP_ERROR_HIP with a message
    with pytest.raises(capi.STPError) as e:
        capi.call("stp_format_heightmap", None, None, 4, 4, None)
    assert e.value.code in (3, 6)
    assert stp.lib().stp_last_error()
    with pytest.raises(capi.STPError) as e:
        capi.call("stp_filter_buffer_create", 0, C.byref(C.c_void_p()))
    assert e.value.code == 6
