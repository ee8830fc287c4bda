"""stp_heightfield_erode_device_noalloc (the merged map in the library's fixed device scratch)
against stp_heightfield_erode_device (the merged map from the generator's memory pool), through
the C ABI. Serialized erosion, so both must agree bit for bit: two calls in a row, the second on
a smaller neighbourhood (the scratch keeps the first call's contents beyond it), then a map too
large for the scratch, which must be refused."""
import ctypes as C

import numpy as np
import pytest

from superterrainplus_amd import capi, workloads as W

from helpers import log
from test_gpu_generator import NOISE, Generator

pytestmark = pytest.mark.gpu

SCRATCH_PIXELS = 4 << 20


def chunk_setting(side, nn=3):
    c = capi.ChunkSetting()
    c.ChunkSize[:] = [side, side]
    c.MapSize[:] = [side, side]
    c.ChunkScale[:] = [1.0, 1.0]
    c.DiversityNearestNeighbour[:] = [1, 1]
    c.ErosionNearestNeighbour[:] = [nn, nn]
    c.SplatNearestNeighbour[:] = [1, 1]
    c.RenderDistance[:] = [3, 3]
    return c


def plugin(side):
    d = C.c_void_p()
    biome = capi.BiomeProperty(*W.Fractal().biome())
    capi.call("stp_diversity_simplex_create", C.byref(NOISE), C.byref(biome), (C.c_uint32 * 2)(side, side), C.byref(d))
    return d.value


def neighbourhood(gen, side, nn):
    """nn x nn device chunks of fractal noise at their world offsets, plus UINT16 outputs."""
    heights = [capi.DeviceArray((side, side), np.float32) for _ in range(nn * nn)]
    fixed = [capi.DeviceArray((side, side), np.uint16) for _ in range(nn * nn)]
    none_biome = (C.c_void_p * 1)(None)
    h = nn // 2
    for i, d in enumerate(heights):
        cx, cy = i % nn - h, i // nn - h
        capi.call("stp_heightfield_generate_device", C.c_void_p(gen.h), d.vp, none_biome, float(cx * side),
                  float(cy * side), None)
    return heights, fixed


def erode(name, gen, heights, fixed):
    n = len(heights)
    capi.call(name, C.c_void_p(gen.h), (C.c_void_p * n)(*[d.ptr for d in heights]),
              (C.c_void_p * n)(*[d.ptr for d in fixed]), None)
    capi.synchronize()
    return [d.download() for d in heights], [d.download() for d in fixed]


@pytest.mark.parametrize("sides", [(256, 96)], ids=["768px-then-288px"])
def test_noalloc_erosion_matches_pool_erosion(gpu, sides):
    for k, side in enumerate(sides):
        print(f"[case] noalloc erosion, call {k}, 3x3 chunks of {side}^2", flush=True)
        log(f"noalloc erosion call {k} side {side}")
        rain = W.erosion_setting(3000, seed=11 + k)
        p = plugin(side)
        pool_gen = Generator(p, chunk_setting(side), rain, mode=1)
        free_gen = Generator(p, chunk_setting(side), rain, mode=1)
        before_h, before_f = neighbourhood(pool_gen, side, 3)
        start = [d.download() for d in before_h]
        want_h, want_f = erode("stp_heightfield_erode_device", pool_gen, before_h, before_f)
        heights, fixed = neighbourhood(free_gen, side, 3)
        assert all(d.download().tobytes() == s.tobytes() for d, s in zip(heights, start))
        got_h, got_f = erode("stp_heightfield_erode_device_noalloc", free_gen, heights, fixed)
        # the erosion moved material: the comparison is not of untouched maps
        assert any(not np.array_equal(g, s) for g, s in zip(got_h, start))
        for i in range(9):
            assert got_h[i].tobytes() == want_h[i].tobytes(), (k, i)
            assert np.array_equal(got_f[i], want_f[i]), (k, i)
        del pool_gen, free_gen
        capi.call("stp_diversity_destroy", C.c_void_p(p))


def test_noalloc_erosion_refuses_maps_beyond_the_scratch(gpu):
    side = 1024  # 3 x 3 x 1024^2 = 9.4 M pixels > the scratch's 4 Mi
    assert 9 * side * side > SCRATCH_PIXELS
    log("noalloc erosion beyond scratch")
    p = plugin(side)
    gen = Generator(p, chunk_setting(side), W.erosion_setting(10), mode=1)
    tiny = [capi.DeviceArray((1, 1), np.float32) for _ in range(9)]
    tiny_f = [capi.DeviceArray((1, 1), np.uint16) for _ in range(9)]
    rc = capi.lib().stp_heightfield_erode_device_noalloc(C.c_void_p(gen.h), (C.c_void_p * 9)(*[d.ptr for d in tiny]),
                                                        (C.c_void_p * 9)(*[d.ptr for d in tiny_f]), None)
    assert rc == 4, rc  # STP_ERROR_CAPACITY
    assert "allocation-free erosion holds at most" in capi.lib().stp_last_error().decode()
    del gen
    capi.call("stp_diversity_destroy", C.c_void_p(p))
