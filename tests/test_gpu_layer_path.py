"""The heightfield path with its biome maps from a layer chain, the reference's arrangement (the world
pipeline generates each chunk's biome map with the biome factory, SuperTerrain+/SuperTerrain+/Private/World/STPWorldPipeline.cpp:411-430,
and the heightmap from the biome maps of its neighbourhood, :437-476).

Without erosion every chunk is bit-exact against the oracle: its halo biome map is the oracle's
layer chain over the chunk's halo window (at the chunk's step between samples), its heights the
oracle's filter and blend over that halo. Ids of a chain at or beyond the batch's biome count are
written as 0 and flagged."""
import ctypes as C

import numpy as np
import pytest

from superterrainplus_amd import capi, workloads as W

from helpers import log

pytestmark = pytest.mark.gpu

NOISE = capi.SimplexNoiseSetting(42, W.DISTRIBUTION, W.GRADIENT_OFFSET, W.ENGINE_MT19937)
RADIUS, COUNT, SEED, GLOBAL = 8, 4, 7, 6666


def props(count=COUNT):
    return [(p[0], 8, p[2], p[3], p[4], p[5]) for p in W.biome_table(count, SEED)]


def factory(nodes=None, registry=None):
    nodes = nodes or capi.demo_layer_program()
    registry = registry or capi.demo_biome_registry()
    arr = (capi.LayerNode * len(nodes))(*nodes)
    h = C.c_void_p()
    capi.call("stp_layer_factory_create", arr, len(nodes), GLOBAL, C.byref(registry), 16, 16, C.byref(h))
    return h, nodes, registry


def oracle_halo(oracle, nodes, registry, origin, w, step=1):
    o = [oracle.layer_node(n.Kind, n.Salt, [p for p in n.Parent if p >= 0], n.Param, n.CacheSize) for n in nodes]
    reg = (np.array(registry.ID[:]), np.array(registry.Temperature[:]), np.array(registry.Precipitation[:]))
    full = oracle.layer_biomemap(o, GLOBAL, reg, [origin], (w - 1) * step + 1, (w - 1) * step + 1)[0]
    return np.ascontiguousarray(full[::step, ::step])


def oracle_heights(oracle, halo, side, origin, step=1):
    item, weight, off = oracle.shf_window(halo, (RADIUS, RADIUS), (side, side), RADIUS)
    perm, g32, _ = oracle.permutation(42, W.DISTRIBUTION, W.GRADIENT_OFFSET, W.ENGINE_MT19937)
    return oracle.multi_biome(perm, g32, props(), item, weight, off, (side, side),
                              (origin[0] + side * step / 2, origin[1] + side * step / 2), step=step)


def make_batch(layers, sides, origins, steps, count=COUNT):
    setting = capi.BatchSetting(SEED, count, 96, RADIUS, 8, 0.0, layers)
    chunks = (capi.BatchChunk * len(sides))(*[capi.BatchChunk(s, (C.c_int32 * 2)(*o), st)
                                               for s, o, st in zip(sides, origins, steps)])
    h = C.c_void_p()
    capi.call("stp_batch_create", C.byref(NOISE), capi.biome_array(props(count)), count,
              C.byref(W.erosion_setting(1)), C.byref(setting), chunks, len(sides), C.byref(h))
    return h


def read(h, i, side):
    w = side + 2 * RADIUS
    height, fixed, biome = np.zeros((side, side), np.float32), np.zeros((side, side), np.uint16), np.zeros((w, w), np.uint16)
    capi.call("stp_batch_read_chunk", h, i, capi.ptr(height), capi.ptr(fixed), capi.ptr(biome), None)
    return height, fixed, biome


def test_batch_with_demo_layer_biomes_bit_exact(gpu, oracle):
    log("batch with layer biomes")
    f, nodes, reg = factory()
    sides, origins, steps = [64, 128, 96, 64, 64], [(32, 420), (0, 360), (-300, 77), (1000, -5000), (16, 400)], [1, 1, 1, 2, 4]
    h = make_batch(f, sides, origins, steps)
    try:
        seen = set()
        for mode in (0, 1, 2):
            capi.call("stp_batch_run", h, mode, None)
            capi.synchronize()
            for i, (s, o, st) in enumerate(zip(sides, origins, steps)):
                height, fixed, biome = read(h, i, s)
                halo = oracle_halo(oracle, nodes, reg, (o[0] - RADIUS * st, o[1] - RADIUS * st), s + 2 * RADIUS, st)
                assert (biome == halo).all(), (mode, i)
                seen |= set(np.unique(halo).tolist())
                assert height.tobytes() == oracle_heights(oracle, halo, s, o, st).tobytes(), (mode, i)
                assert (fixed == oracle.format_heightmap(height)).all()
        assert len(seen) >= 2
        status = C.c_uint32()
        capi.call("stp_layer_factory_status", f, C.byref(status))
        assert status.value == 0
    finally:
        capi.lib().stp_batch_destroy(h)
        capi.lib().stp_layer_factory_destroy(f)


def test_batch_with_layer_biomes_pipelined_graph(gpu):
    """A binding above 8 Mi pixels with layer-chain biomes: its graph run evaluates the chain for four
    chunk groups concurrently (each map's nodes in its own part of the plan's arena). Every chunk's
    halo biome map and heights equal those of the plain run, which takes the chunks one at a time."""
    log("batch with layer biomes, pipelined graph (9 chunks of 1024^2)")
    layers, _, _ = factory()
    try:
        sides = [1024] * 9
        origins = [((k % 3 - 1) * 1024 + 13, (k // 3 - 1) * 1024 - 5) for k in range(9)]
        h = make_batch(layers, sides, origins, [1] * 9)
        try:
            out = {}
            for mode in (0, 2):
                capi.call("stp_batch_run", h, mode, None)
                capi.synchronize()
                got = []
                for i, side in enumerate(sides):
                    hh = np.zeros((side, side), np.float32)
                    bb = np.zeros((side + 2 * RADIUS, side + 2 * RADIUS), np.uint16)
                    capi.call("stp_batch_read_chunk", h, i, capi.ptr(hh), None, capi.ptr(bb), None)
                    got.append((hh, bb))
                out[mode] = got
            for i in range(len(sides)):
                assert (out[0][i][1] == out[2][i][1]).all(), i
                assert out[0][i][0].tobytes() == out[2][i][0].tobytes(), i
        finally:
            capi.lib().stp_batch_destroy(h)
    finally:
        capi.lib().stp_layer_factory_destroy(layers)


def test_layer_ids_beyond_the_biome_table_are_flagged(gpu, oracle):
    """A chain of ocean-temperature noise draws ids 0, 2 and 5 of a registry with distinct ids; with
    4 biomes, id 5 is replaced by 0 and flagged."""
    log("layer ids beyond the biome table")
    reg = capi.BiomeRegistry()
    for i in range(35):
        reg.ID[i], reg.Temperature[i], reg.Precipitation[i] = i, 1.5, 1.5
    nodes = [capi.LayerNode(15, 0, (C.c_int32 * 2)(-1, -1), 99, 0), capi.LayerNode(1, 0, (C.c_int32 * 2)(0, -1), 5, 0)]
    f, nodes, reg = factory(nodes, reg)
    h = make_batch(f, [64], [(100, -40)], [1])
    try:
        capi.call("stp_batch_run", h, 0, None)
        capi.synchronize()
        _, _, biome = read(h, 0, 64)
        ref = oracle_halo(oracle, nodes, reg, (100 - RADIUS, -40 - RADIUS), 64 + 2 * RADIUS)
        assert (ref >= COUNT).any() and (ref < COUNT).any()
        assert (biome == np.where(ref >= COUNT, 0, ref)).all()
        status = C.c_uint32()
        capi.call("stp_layer_factory_status", f, C.byref(status))
        assert status.value == 2
        capi.call("stp_layer_factory_status", f, C.byref(status))
        assert status.value == 0
    finally:
        capi.lib().stp_batch_destroy(h)
        capi.lib().stp_layer_factory_destroy(f)


def test_world_with_layer_biomes(gpu, oracle):
    """World chunks without erosion, and the rendered area's biome map, from the demo chain."""
    log("world with layer biomes")
    f, nodes, reg = factory()
    m = 64
    batch = capi.BatchSetting(SEED, COUNT, 96, RADIUS, 8, 0.0, f)
    setting = capi.WorldSetting(m, 3, 16, 4)
    w = C.c_void_p()
    capi.call("stp_world_create", C.byref(NOISE), capi.biome_array(props()), COUNT, C.byref(W.erosion_setting(1)),
              C.byref(batch), C.byref(setting), C.byref(w))
    try:
        side = 3 * m
        biome, height = capi.DeviceArray((side, side), np.uint16), capi.DeviceArray((side, side), np.uint16)
        capi.call("stp_world_render_area", w, 0, 6, 1, biome.vp, height.vp)
        assert (biome.download() == oracle_halo(oracle, nodes, reg, (-m, 5 * m), side)).all()
        for x, y in [(-1, 5), (0, 6), (1, 7)]:
            hh = np.zeros((m, m), np.float32)
            capi.call("stp_world_read", w, x, y, capi.ptr(hh), None)
            halo = oracle_halo(oracle, nodes, reg, (x * m - RADIUS, y * m - RADIUS), m + 2 * RADIUS)
            assert hh.tobytes() == oracle_heights(oracle, halo, m, (x * m, y * m)).tobytes(), (x, y)
        # exported records carry the chain's biome maps
        size, written = C.c_uint64(), C.c_uint64()
        capi.call("stp_chunk_record_bytes", m, m, 1, C.byref(size))
        buf = (C.c_uint8 * (2 * size.value))()
        capi.call("stp_world_export", w, (C.c_int32 * 4)(-1, 5, 1, 7), 2, 1, buf, len(buf), C.byref(written))
        for r, (x, y) in zip(capi.read_chunk_records(bytes(buf)), [(-1, 5), (1, 7)]):
            assert (r["biome"] == oracle_halo(oracle, nodes, reg, (x * m, y * m), m)).all()
    finally:
        capi.lib().stp_world_destroy(w)
        capi.lib().stp_layer_factory_destroy(f)
