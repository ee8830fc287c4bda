"""World pipeline on resident device state: heightmaps of every neighbourhood chunk first, then
each requested chunk eroded free-slip as the centre of its 3 x 3 neighbourhood (drops in the centre
chunk, the whole merged map eroded and written back), centres with overlapping neighbourhoods in
different batches (SuperTerrain+/SuperTerrain+/Private/World/STPWorldPipeline.cpp:339-663).

Without erosion every chunk is bit-exact against the oracle. With erosion each neighbourhood is
checked, as every parallel erosion here, against the serialized FP32 oracle erosion of the same
merged map with the same drops (seed, stream word of the centre): RMS within 1.5 x the RMS between
two serialized runs with different drops, sediment moved within 4/sqrt(drops) + 0.12 of the
serialized run, UINT16 maps the format of the heights."""
import ctypes as C

import numpy as np
import pytest

from superterrainplus_amd import capi, workloads as W

from helpers import log, oracle_rain

pytestmark = pytest.mark.gpu

NOISE = capi.SimplexNoiseSetting(42, W.DISTRIBUTION, W.GRADIENT_OFFSET, W.ENGINE_MT19937)
RADIUS, COUNT, CELL, SEED, MAP, NN = 8, 4, 96, 7, 128, 3


def props():
    return [(p[0], 8, p[2], p[3], p[4], p[5]) for p in W.biome_table(COUNT, SEED)]


class World:
    def __init__(self, dpp, generate=8, erode=4):
        batch = capi.BatchSetting(SEED, COUNT, CELL, RADIUS, 8, dpp)
        setting = capi.WorldSetting(MAP, NN, generate, erode)
        h = C.c_void_p()
        capi.call("stp_world_create", C.byref(NOISE), capi.biome_array(props()), COUNT,
                  C.byref(W.erosion_setting(1)), C.byref(batch), C.byref(setting), C.byref(h))
        self.h = h.value

    def request(self, chunks):
        flat = (C.c_int32 * (2 * len(chunks)))(*[v for c in chunks for v in c])
        cap = 9 * len(chunks) + 1
        out, n = (C.c_int32 * (2 * cap))(), C.c_uint32()
        capi.call("stp_world_request", C.c_void_p(self.h), flat, len(chunks), out, cap, C.byref(n))
        return [(out[2 * i], out[2 * i + 1]) for i in range(n.value)]

    def read(self, x, y):
        h = np.zeros((MAP, MAP), np.float32)
        f = np.zeros((MAP, MAP), np.uint16)
        capi.call("stp_world_read", C.c_void_p(self.h), x, y, capi.ptr(h), capi.ptr(f))
        return h, f

    def state(self, x, y):
        s = C.c_uint32()
        capi.call("stp_world_state", C.c_void_p(self.h), x, y, C.byref(s))
        return s.value

    def merged(self, cx, cy):
        rows = [np.concatenate([self.read(cx + dx, cy + dy)[0] for dx in (-1, 0, 1)], axis=1) for dy in (-1, 0, 1)]
        return np.ascontiguousarray(np.concatenate(rows, axis=0))

    def __del__(self):
        if getattr(self, "h", None):
            capi.lib().stp_world_destroy(C.c_void_p(self.h))


def oracle_chunk(oracle, x, y):
    """Generated (pre-erosion) heights of chunk (x, y): world pixels from (x, y) * MAP."""
    r = RADIUS
    o = (x * MAP, y * MAP)
    halo = oracle.voronoi_map(SEED, COUNT, CELL, (o[0] - r, o[1] - r), MAP + 2 * r, MAP + 2 * r)
    item, weight, off = oracle.shf_window(halo, (r, r), (MAP, MAP), r)
    perm, g32, _ = oracle.permutation(42, W.DISTRIBUTION, W.GRADIENT_OFFSET, W.ENGINE_MT19937)
    return oracle.multi_biome(perm, g32, props(), item, weight, off, (MAP, MAP), (o[0] + MAP / 2, o[1] + MAP / 2))


def stream_word(x, y):
    w = C.c_uint32()
    capi.call("stp_world_stream_word", x, y, C.byref(w))
    return w.value


def neighbourhood(cx, cy):
    return [(cx + dx, cy + dy) for dy in (-1, 0, 1) for dx in (-1, 0, 1)]


def check_region(oracle, before, after, centre, dpp):
    """The erosion of one merged neighbourhood against the serialized oracle (see the module doc)."""
    setting = W.erosion_setting(1)
    drops = int(MAP * MAP * dpp)
    ser = before.copy()
    sled, _ = oracle.erode(ser, (NN, NN), (MAP, MAP), oracle_rain(oracle, setting), drops, setting.Seed,
                           stream_id=stream_word(*centre))
    other = before.copy()
    oracle.erode(other, (NN, NN), (MAP, MAP), oracle_rain(oracle, setting), drops, setting.Seed,
                 stream_id=stream_word(*centre) ^ 0x5A5A5A5A)
    moved = float(sled.astype(np.float64).sum())
    gpu_moved = -float(after.astype(np.float64).sum() - before.astype(np.float64).sum())
    assert moved > 0 and abs(gpu_moved - moved) <= (4.0 / np.sqrt(drops) + 0.12) * moved, (centre, gpu_moved, moved)
    floor_rms = np.sqrt(np.mean((ser.astype(np.float64) - other) ** 2))
    diff_rms = np.sqrt(np.mean((after.astype(np.float64) - ser) ** 2))
    assert diff_rms <= 1.5 * floor_rms, (centre, diff_rms, floor_rms)
    # drops spawn in the centre chunk but erode across its borders (free slip)
    changed = np.abs(after - before) > 0
    assert changed[:MAP].any() or changed[2 * MAP:].any() or changed[:, :MAP].any() or changed[:, 2 * MAP:].any()


def test_world_generation_bit_exact(gpu, oracle):
    """dpp 0: request two far-apart chunks; their neighbourhoods are generated, the requested
    chunks complete, and every chunk equals the oracle bit for bit."""
    log("world generation")
    w = World(0.0)
    changed = w.request([(0, 0), (7, -3)])
    assert sorted(changed) == sorted(neighbourhood(0, 0) + neighbourhood(7, -3))
    for x, y in changed:
        h, f = w.read(x, y)
        assert h.tobytes() == oracle_chunk(oracle, x, y).tobytes(), (x, y)
        assert (f == oracle.format_heightmap(h)).all()
        assert w.state(x, y) == (2 if (x, y) in [(0, 0), (7, -3)] else 1)
    assert w.state(1, 2) == 0
    assert w.request([(0, 0)]) == []  # complete already


def test_world_overlapping_centres_in_order(gpu, oracle):
    """Centre (0, 0), then (1, 0) whose neighbourhood overlaps it: each erosion starts from the
    maps the previous one left (resident chunks) plus freshly generated ones."""
    log("world overlapping centres")
    dpp = 0.5
    w = World(dpp)
    before = np.ascontiguousarray(np.concatenate(
        [np.concatenate([oracle_chunk(oracle, x, y) for x in (-1, 0, 1)], axis=1) for y in (-1, 0, 1)], axis=0))
    assert sorted(w.request([(0, 0)])) == sorted(neighbourhood(0, 0))
    after = w.merged(0, 0)
    check_region(oracle, before, after, (0, 0), dpp)
    for x, y in neighbourhood(0, 0):
        assert (w.read(x, y)[1] == oracle.format_heightmap(w.read(x, y)[0])).all()
    # (1, 0): its column x = 2 is new, the rest is what the first erosion left
    rows = []
    for y in (-1, 0, 1):
        rows.append(np.concatenate([w.read(x, y)[0] if x < 2 else oracle_chunk(oracle, x, y) for x in (0, 1, 2)],
                                   axis=1))
    before = np.ascontiguousarray(np.concatenate(rows, axis=0))
    assert sorted(w.request([(1, 0)])) == sorted(neighbourhood(1, 0))
    check_region(oracle, before, w.merged(1, 0), (1, 0), dpp)
    assert w.state(0, 0) == 2 and w.state(1, 0) == 2 and w.state(2, 0) == 1


def test_world_disjoint_centres_share_one_erosion(gpu, oracle):
    """Three centres 3 chunks apart have disjoint neighbourhoods: one batch erodes them together;
    each neighbourhood is checked on its own."""
    log("world disjoint centres")
    dpp = 0.5
    w = World(dpp)
    centres = [(20, 5), (23, 5), (20, 8)]
    before = {c: np.ascontiguousarray(np.concatenate(
        [np.concatenate([oracle_chunk(oracle, c[0] + dx, c[1] + dy) for dx in (-1, 0, 1)], axis=1)
         for dy in (-1, 0, 1)], axis=0)) for c in centres}
    changed = w.request(centres)
    assert len(changed) == 27
    for c in centres:
        check_region(oracle, before[c], w.merged(*c), c, dpp)


def test_world_block_of_centres(gpu):
    """A 3 x 2 block of adjacent centres (every pair conflicts: six batches, in request order): all
    complete, the 5 x 4 chunks around them changed and finite."""
    log("world block of centres")
    w = World(0.5, generate=5, erode=2)
    centres = [(40 + i, 40 + j) for j in range(2) for i in range(3)]
    changed = w.request(centres)
    assert sorted(changed) == sorted({(x, y) for x in range(39, 44) for y in range(39, 43)})
    for x, y in changed:
        assert w.state(x, y) == (2 if (x, y) in centres else 1)
        h, f = w.read(x, y)
        assert np.isfinite(h).all()


def test_world_render_area_and_splatmap(gpu, oracle):
    """The rendered area around (3, -2) (3 x 3 chunks, made complete): the chunks' UINT16 maps side
    by side, the area's biome map equal to the oracle's Voronoi map, and the splatmap of its chunks
    bit-exact against the oracle on those buffers."""
    log("world render area")
    w = World(0.5)
    n, side = 3, 3 * MAP
    biome, height = capi.DeviceArray((side, side), np.uint16), capi.DeviceArray((side, side), np.uint16)
    capi.call("stp_world_render_area", C.c_void_p(w.h), 3, -2, 1, biome.vp, height.vp)
    hb, hh = biome.download(), height.download()
    for ly in range(n):
        for lx in range(n):
            x, y = 3 - 1 + lx, -2 - 1 + ly
            assert w.state(x, y) == 2
            assert (hh[ly * MAP:(ly + 1) * MAP, lx * MAP:(lx + 1) * MAP] == w.read(x, y)[1]).all(), (x, y)
    assert (hb == oracle.voronoi_map(SEED, COUNT, CELL, (2 * MAP, -3 * MAP), side, side)).all()
    altitude = [(b, u, 3 * b + k) for b in range(COUNT) for k, u in enumerate((0.4, 0.7, 1.0))]
    gradient = [(b, 0.05, 1.0, 0.0, 1.0, 20 + b) for b in range(COUNT)]
    a = (capi.AltitudeRule * len(altitude))(*[capi.AltitudeRule(*r) for r in altitude])
    g = (capi.GradientRule * len(gradient))(*[capi.GradientRule(*r) for r in gradient])
    db = C.c_void_p()
    capi.call("stp_splat_database_create", a, len(altitude), g, len(gradient), C.byref(db))
    table = capi.PermutationTable(42, W.DISTRIBUTION, W.GRADIENT_OFFSET, W.ENGINE_MT19937)
    try:
        out = capi.DeviceArray.from_host(np.zeros((side, side), np.uint8))
        chunks = [(lx, ly, float((3 - 1 + lx) * MAP), float((-2 - 1 + ly) * MAP)) for ly in range(n) for lx in range(n)]
        c = (capi.SplatChunk * len(chunks))(*[capi.SplatChunk(*r) for r in chunks])
        capi.call("stp_splatmap_generate", C.byref(table.view), db, biome.vp, height.vp, out.vp, side, side, MAP, MAP, c,
                  len(chunks), None)
        capi.synchronize()
        p, g32, _ = oracle.permutation(42, W.DISTRIBUTION, W.GRADIENT_OFFSET, W.ENGINE_MT19937)
        ref = oracle.splatmap(p, g32, altitude, gradient, hb, hh, (MAP, MAP), chunks)
        gpu = out.download()
        assert (gpu == ref).all()
        assert len(np.unique(gpu)) > 2
    finally:
        capi.lib().stp_splat_database_destroy(db)


def export(w, chunks, maps):
    size = C.c_uint64()
    capi.call("stp_chunk_record_bytes", MAP, MAP, maps, C.byref(size))
    buf = (C.c_uint8 * (size.value * len(chunks)))()
    flat = (C.c_int32 * (2 * len(chunks)))(*[v for c in chunks for v in c])
    written = C.c_uint64()
    capi.call("stp_world_export", C.c_void_p(w.h), flat, len(chunks), maps, buf, len(buf), C.byref(written))
    assert written.value == len(buf)
    return bytes(buf)


def test_world_export_records(gpu, oracle):
    """Chunk records of a complete chunk and of neighbourhood chunks with heightmaps only: the maps
    equal the world's reads and the oracle's biome map, the completeness follows the chunk's state,
    every record passes the reader's check; a chunk without a map and a short buffer are refused."""
    log("world export records")
    w = World(0.5)
    w.request([(2, -1)])
    chunks = [(2, -1), (1, -2), (3, 0)]
    recs = capi.read_chunk_records(export(w, chunks, 7))
    assert len(recs) == 3
    for (x, y), r in zip(chunks, recs):
        h, f = w.read(x, y)
        assert (r["x"], r["y"], r["origin"]) == (x, y, (x * MAP, y * MAP))
        assert r["completeness"] == (3 if (x, y) == (2, -1) else 2)
        assert r["height"].tobytes() == h.tobytes() and (r["fixed"] == f).all()
        assert (r["biome"] == oracle.voronoi_map(SEED, COUNT, CELL, (x * MAP, y * MAP), MAP, MAP)).all()
    only = capi.read_chunk_records(export(w, chunks[:1], 4))[0]
    assert only["biome"] is None and only["height"] is None and (only["fixed"] == w.read(2, -1)[1]).all()
    flat = (C.c_int32 * 2)(9, 9)
    buf, written = (C.c_uint8 * (1 << 20))(), C.c_uint64()
    with pytest.raises(capi.STPError):
        capi.call("stp_world_export", C.c_void_p(w.h), flat, 1, 7, buf, len(buf), C.byref(written))
    flat = (C.c_int32 * 2)(2, -1)
    with pytest.raises(capi.STPError):
        capi.call("stp_world_export", C.c_void_p(w.h), flat, 1, 7, buf, 100, C.byref(written))


def test_world_rejects_bad_settings(gpu):
    batch = capi.BatchSetting(SEED, COUNT, CELL, RADIUS, 8, 0.0)
    for bad in [capi.WorldSetting(MAP, 2, 4, 4), capi.WorldSetting(4, 3, 4, 4), capi.WorldSetting(MAP, 3, 0, 4)]:
        h = C.c_void_p()
        with pytest.raises(capi.STPError):
            capi.call("stp_world_create", C.byref(NOISE), capi.biome_array(props()), COUNT,
                      C.byref(W.erosion_setting(1)), C.byref(batch), C.byref(bad), C.byref(h))
    w = World(0.0)
    with pytest.raises(capi.STPError):
        w.read(99, 99)


def test_world_load_swaps_buffers(gpu):
    """STPWorldPipeline::load on a worker thread (STPWorldPipeline.cpp:1504-1633): a new centre
    starts a load (back buffer busy); calls while it runs are ignored; once it finished, a load of
    the same centre swaps the buffers and the front buffer holds the rendered area exactly as the
    synchronous stp_world_render_area makes it (no erosion: bit-exact); a load of the same centre
    then changes nothing; a new centre starts a new load while the front buffer stays usable."""
    import time
    log("world asynchronous load")
    w, ref = World(0.0), World(0.0)
    radius, side = 1, 3 * MAP

    def load(x, y):
        st = C.c_uint32()
        capi.call("stp_world_load", C.c_void_p(w.h), x, y, radius, C.byref(st))
        return st.value

    def front():
        b, h, n = C.c_void_p(), C.c_void_p(), C.c_uint32()
        capi.call("stp_world_front", C.c_void_p(w.h), C.byref(b), C.byref(h), C.byref(n))
        return b.value, h.value, n.value

    def download(ptr):
        out = np.zeros((side, side), np.uint16)
        capi.call("stp_memcpy_async", capi.ptr(out), C.c_void_p(ptr), out.nbytes, None)
        capi.synchronize()
        return out

    assert front() == (None, None, 0)
    for centre in [(2, -1), (9, 4)]:
        assert load(*centre) == 1
        deadline, status = time.time() + 60.0, 1
        while status == 1 and time.time() < deadline:
            time.sleep(0.002)
            status = load(*centre)
        assert status == 2, status
        assert load(*centre) == 0
        b, h, n = front()
        assert n == side and b and h
        rb, rh = capi.DeviceArray((side, side), np.uint16), capi.DeviceArray((side, side), np.uint16)
        capi.call("stp_world_render_area", C.c_void_p(ref.h), centre[0], centre[1], radius, rb.vp, rh.vp)
        capi.synchronize()
        assert (download(b) == rb.download()).all() and (download(h) == rh.download()).all()
        # the chunks of the area are complete in the world itself
        assert w.state(*centre) == 2

