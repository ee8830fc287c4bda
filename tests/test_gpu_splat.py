"""Splatmap generator on the GPU against the oracle, bit for bit: rendered areas of 3 x 2 chunks with
smooth random heightfields and patchy biome maps, rule databases with overlapping and tied rules
and biomes without rules, chunks splatted in any order and subset (the others untouched), the
area's clamped borders, and more chunks than one launch holds."""
import ctypes as C

import numpy as np
import pytest

from superterrainplus_amd import capi, workloads as W

from helpers import log

pytestmark = pytest.mark.gpu


def smooth_field(rng, h, w, cells=6):
    coarse = rng.uniform(0, 65535, (cells + 1, cells + 1))
    ys, xs = np.linspace(0, cells, h), np.linspace(0, cells, w)
    y0, x0 = np.minimum(ys.astype(int), cells - 1), np.minimum(xs.astype(int), cells - 1)
    fy, fx = (ys - y0)[:, None], (xs - x0)[None, :]
    c = coarse
    v = (c[y0][:, x0] * (1 - fy) * (1 - fx) + c[y0 + 1][:, x0] * fy * (1 - fx) + c[y0][:, x0 + 1] * (1 - fy) * fx
         + c[y0 + 1][:, x0 + 1] * fy * fx)
    return (v + rng.normal(0, 300, v.shape)).clip(0, 65535).astype(np.uint16)


def random_rules(rng, samples):
    altitude, gradient = [], []
    for s in samples:
        for _ in range(rng.integers(0, 4)):
            altitude.append((s, float(rng.choice([0.25, 0.5, 0.75, rng.uniform(0, 1)])), int(rng.integers(0, 40))))
        for _ in range(rng.integers(0, 3)):
            lo, hi = sorted(rng.uniform(-0.2, 1.0, 2))
            a0, a1 = sorted(rng.uniform(0, 1, 2))
            gradient.append((s, float(lo), float(hi), float(a0), float(a1), int(rng.integers(40, 80))))
    return altitude, gradient


def database(altitude, gradient):
    a = (capi.AltitudeRule * max(len(altitude), 1))(*[capi.AltitudeRule(*r) for r in altitude])
    g = (capi.GradientRule * max(len(gradient), 1))(*[capi.GradientRule(*r) for r in gradient])
    h = C.c_void_p()
    capi.call("stp_splat_database_create", a, len(altitude), g, len(gradient), C.byref(h))
    return h


def splat_gpu(table, db, biome, height, map_size, chunks):
    H, W_ = height.shape
    b, h = capi.DeviceArray.from_host(biome), capi.DeviceArray.from_host(height)
    out = capi.DeviceArray.from_host(np.zeros((H, W_), np.uint8))
    c = (capi.SplatChunk * len(chunks))(*[capi.SplatChunk(*r) for r in chunks])
    capi.call("stp_splatmap_generate", C.byref(table.view), db, b.vp, h.vp, out.vp, W_, H, map_size[0], map_size[1], c,
              len(chunks), None)
    capi.synchronize()
    return out.download()


@pytest.mark.parametrize("seed", [1, 2, 3])
def test_splatmap_bit_exact(gpu, oracle, seed):
    log(f"splatmap seed {seed}")
    rng = np.random.default_rng(seed)
    mx, my = 96, 64
    H, W_ = 2 * my, 3 * mx
    height = smooth_field(rng, H, W_)
    biome = np.repeat(np.repeat(rng.integers(0, 8, (H // 16, W_ // 16)), 16, 0), 16, 1).astype(np.uint16)
    altitude, gradient = random_rules(rng, [1, 2, 4, 5, 7])
    db = database(altitude, gradient)
    table = capi.PermutationTable(42, W.DISTRIBUTION, W.GRADIENT_OFFSET, W.ENGINE_MT19937)
    try:
        chunks = [(2, 1, 192.0, 64.0), (0, 0, -96.0, -64.0), (1, 1, 0.0, 64.0), (2, 0, 96.0, -64.0)]
        gpu = splat_gpu(table, db, biome, height, (mx, my), chunks)
        p, g32, _ = oracle.permutation(42, W.DISTRIBUTION, W.GRADIENT_OFFSET, W.ENGINE_MT19937)
        ref = oracle.splatmap(p, g32, altitude, gradient, biome, height, (mx, my), chunks)
        assert (gpu == ref).all(), np.argwhere(gpu != ref)[:5]
        # the unlisted chunks (1, 0) and (0, 1) keep their zeros; several regions were taken
        assert (gpu[:my, mx:2 * mx] == 0).all() and (gpu[my:, :mx] == 0).all()
        assert len(np.unique(gpu)) > 3
    finally:
        capi.lib().stp_splat_database_destroy(db)


def test_splatmap_many_chunks_and_empty_database(gpu, oracle):
    """70 chunks of 16 x 16 (two launches) with no rules at all: every pixel has no region."""
    log("splatmap many chunks")
    mx = my = 16
    H, W_ = 7 * my, 10 * mx
    rng = np.random.default_rng(5)
    height = smooth_field(rng, H, W_)
    biome = np.zeros((H, W_), np.uint16)
    db = database([], [])
    table = capi.PermutationTable(42, W.DISTRIBUTION, W.GRADIENT_OFFSET, W.ENGINE_MT19937)
    try:
        chunks = [(x, y, 16.0 * x, 16.0 * y) for y in range(7) for x in range(10)]
        gpu = splat_gpu(table, db, biome, height, (mx, my), chunks)
        assert (gpu == 255).all()
    finally:
        capi.lib().stp_splat_database_destroy(db)


def test_splatmap_rejects_bad_input(gpu):
    db = database([(0, 0.5, 1)], [])
    table = capi.PermutationTable(42, W.DISTRIBUTION, W.GRADIENT_OFFSET, W.ENGINE_MT19937)
    try:
        biome = np.zeros((32, 32), np.uint16)
        with pytest.raises(capi.STPError):
            splat_gpu(table, db, biome, biome, (16, 16), [(2, 0, 0.0, 0.0)])  # outside the area
        bad = (capi.GradientRule * 1)(capi.GradientRule(0, 0.5, 0.1, 0.0, 1.0, 1))
        h = C.c_void_p()
        with pytest.raises(capi.STPError):
            capi.call("stp_splat_database_create", None, 0, bad, 1, C.byref(h))
    finally:
        capi.lib().stp_splat_database_destroy(db)


def test_runtime_splatmap_script_matches_builtin(gpu, oracle):
    """The hipRTC-compiled splatmap script (the demo's runtime-compiled generator) against the
    built-in kernel and the oracle, bit for bit; 70 chunks exceed the built-in kernel's 64 per
    launch, the script takes them in one."""
    log("runtime splatmap script")
    rng = np.random.default_rng(11)
    mx = my = 24
    H, W_ = 7 * my, 10 * mx
    height = smooth_field(rng, H, W_)
    biome = np.repeat(np.repeat(rng.integers(0, 8, (H // 8, W_ // 8)), 8, 0), 8, 1).astype(np.uint16)
    altitude, gradient = random_rules(rng, [0, 1, 3, 4, 6])
    db = database(altitude, gradient)
    table = capi.PermutationTable(42, W.DISTRIBUTION, W.GRADIENT_OFFSET, W.ENGINE_MT19937)
    rt = C.c_void_p()
    try:
        capi.call("stp_splatmap_runtime_create", C.byref(table.view), db, mx, my, None, None, 0, C.byref(rt))
        chunks = [(x, y, 24.0 * x - 100.0, 24.0 * y + 7.0) for y in range(7) for x in range(10)]
        builtin = splat_gpu(table, db, biome, height, (mx, my), chunks)
        b, h = capi.DeviceArray.from_host(biome), capi.DeviceArray.from_host(height)
        out = capi.DeviceArray.from_host(np.zeros((H, W_), np.uint8))
        c = (capi.SplatChunk * len(chunks))(*[capi.SplatChunk(*r) for r in chunks])
        capi.call("stp_splatmap_runtime_generate", rt, b.vp, h.vp, out.vp, W_, H, c, len(chunks), None)
        capi.synchronize()
        got = out.download()
        assert (got == builtin).all(), np.argwhere(got != builtin)[:5]
        p, g32, _ = oracle.permutation(42, W.DISTRIBUTION, W.GRADIENT_OFFSET, W.ENGINE_MT19937)
        assert (got == oracle.splatmap(p, g32, altitude, gradient, biome, height, (mx, my), chunks)).all()
        assert len(np.unique(got)) > 3
        with pytest.raises(capi.STPError):  # a chunk outside the area
            capi.call("stp_splatmap_runtime_generate", rt, b.vp, h.vp, out.vp, W_, H,
                      (capi.SplatChunk * 1)(capi.SplatChunk(10, 0, 0.0, 0.0)), 1, None)
        bad = C.c_void_p()
        with pytest.raises(capi.STPError) as e:
            capi.call("stp_splatmap_runtime_create", C.byref(table.view), db, mx, my, b"__global__ void k( {", None, 0,
                      C.byref(bad))
        assert e.value.code == 5
    finally:
        if rt.value:
            capi.lib().stp_splatmap_runtime_destroy(rt)
        capi.lib().stp_splat_database_destroy(db)
