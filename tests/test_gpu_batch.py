"""Batched path (one launch per stage for all chunks, hipGraph replay) against the oracle and
against the plain per-chunk path: bit-exact without erosion, statistically equal with it."""
import ctypes as C

import numpy as np
import pytest

from superterrainplus_amd import capi, workloads as W

from helpers import height_histogram, log, oracle_rain

pytestmark = pytest.mark.gpu

NOISE = capi.SimplexNoiseSetting(42, W.DISTRIBUTION, W.GRADIENT_OFFSET, W.ENGINE_MT19937)
RADIUS, COUNT, CELL, SEED = 8, 4, 96, 7
SIDES = [64, 128, 64, 96, 256, 64]
ORIGINS = [(-1000, 500), (0, 0), (333, -777), (4096, 4096), (-2048, 1024), (64, 64)]


def props(count=COUNT):
    return [(p[0], 8, p[2], p[3], p[4], p[5]) for p in W.biome_table(count, SEED)]


class Batch:
    def __init__(self, dpp, sides=SIDES, origins=ORIGINS, radius=RADIUS, cap=8, count=COUNT, biomes=None):
        self.sides, self.origins, self.radius = sides, origins, radius
        setting = capi.BatchSetting(SEED, count, CELL, radius, cap, dpp)
        chunks = (capi.BatchChunk * len(sides))(*[capi.BatchChunk(s, (C.c_int32 * 2)(*o))
                                                   for s, o in zip(sides, origins)])
        h = C.c_void_p()
        capi.call("stp_batch_create", C.byref(NOISE), capi.biome_array(biomes or props(count)), count,
                  C.byref(W.erosion_setting(1)), C.byref(setting), chunks, len(sides), C.byref(h))
        self.h = h.value

    def run(self, mode):
        capi.call("stp_batch_run", C.c_void_p(self.h), mode, None)
        capi.synchronize()

    def read(self, i):
        s, r = self.sides[i], self.radius
        h = np.zeros((s, s), np.float32)
        f = np.zeros((s, s), np.uint16)
        b = np.zeros((s + 2 * r, s + 2 * r), np.uint16)
        capi.call("stp_batch_read_chunk", C.c_void_p(self.h), i, capi.ptr(h), capi.ptr(f), capi.ptr(b), None)
        return h, f, b

    def ledger(self, i, drops):
        out = np.zeros(max(drops, 1), np.float32)
        capi.call("stp_batch_read_chunk", C.c_void_p(self.h), i, None, None, None, capi.ptr(out))
        return out[:drops]

    def tuning(self, tile=None, threads=256):
        t = capi.BatchTuning()
        if tile is not None:
            capi.call("stp_batch_set_tuning", C.c_void_p(self.h), C.byref(capi.BatchTuning(tile, threads)))
        capi.call("stp_batch_get_tuning", C.c_void_p(self.h), C.byref(t))
        return t.ErosionTile, t.ErosionThreads

    def counters(self):
        out = (C.c_uint64 * 6)()
        capi.call("stp_batch_counters", C.c_void_p(self.h), out, None)
        return list(out)

    def __del__(self):
        if getattr(self, "h", None):
            capi.lib().stp_batch_destroy(C.c_void_p(self.h))


def oracle_chunk(oracle, i, sides=SIDES, origins=ORIGINS, count=COUNT):
    s, o, r = sides[i], origins[i], RADIUS
    halo = oracle.voronoi_map(SEED, count, CELL, (o[0] - r, o[1] - r), s + 2 * r, s + 2 * r)
    item, weight, off = oracle.shf_window(halo, (r, r), (s, s), r)
    perm, g32, _ = oracle.permutation(42, W.DISTRIBUTION, W.GRADIENT_OFFSET, W.ENGINE_MT19937)
    h = oracle.multi_biome(perm, g32, props(count), item, weight, off, (s, s), (o[0] + s / 2, o[1] + s / 2))
    return halo, h


# histogram capacities: 3 and 8 take the biome-indexed item passes (3: an odd capacity, whose
# dense output the horizontal pass stores slot by slot; these chunks hold at most 3 bins per
# window), 16 keeps the vertical accumulator in registers and the horizontal one in LDS, 33 puts
# both in LDS
@pytest.mark.parametrize("cap", [3, 8, 16, 33])
def test_batch_without_erosion_bit_exact(gpu, oracle, cap):
    log(f"batch no erosion cap={cap}")
    b = Batch(0.0, cap=cap)
    results = {}
    for mode in (0, 1, 2):
        b.run(mode)
        results[mode] = [b.read(i) for i in range(len(SIDES))]
    for i in range(len(SIDES)):
        halo, h = oracle_chunk(oracle, i)
        gh, gf, gb = results[0][i]
        assert (gb == halo).all(), i
        assert gh.tobytes() == h.tobytes(), i
        assert (gf == oracle.format_heightmap(h)).all(), i
        for mode in (1, 2):
            assert all((x == y).all() for x, y in zip(results[mode][i], results[0][i])), (mode, i)


# (biomes, bin capacity): the column histograms' slot groups hold 8 slots at capacity 8 and 4 at
# capacity 4, of which the horizontal pass loads 4 with at most 4 biomes (the chunks below hold at
# most 4 bins per window)
@pytest.mark.parametrize("count,cap", [(4, 8), (8, 8), (8, 4)])
def test_batch_split_lines_bit_exact(gpu, oracle, count, cap):
    """A binding of few lines splits every filter line of its large chunks into segments (here
    16 of 64 outputs for the 1024-pixel chunk, 5 uneven ones for the 333-pixel chunk), each
    resumed from the state folded over the segments before it: bit-exact against the oracle, for
    the 4- and the 8-biome kernels, in graph and plain mode."""
    log(f"batch split lines count={count} cap={cap}")
    sides, origins = [1024, 333], [(-5000, 700), (123, -456)]
    b = Batch(0.0, sides=sides, origins=origins, count=count, cap=cap)
    for mode in (0, 2):
        b.run(mode)
        for i in range(len(sides)):
            gh, gf, gb = b.read(i)
            halo, h = oracle_chunk(oracle, i, sides, origins, count)
            assert (gb == halo).all(), (mode, i)
            assert gh.tobytes() == h.tobytes(), (mode, i)
            assert (gf == oracle.format_heightmap(h)).all(), (mode, i)


def test_batch_biome_map_every_level_of_detail(gpu, oracle):
    """The biome map stage draws each 32 x 32 tile's Voronoi sites once and measures 32-bit
    distances while the tile spans less than 2^15 world units, and falls back to one Voronoi
    search per sample beyond that (steps 1024 and 4096 here): the halo maps of every level are
    bit-exact against the oracle."""
    log("batch biome map levels of detail")
    sides, steps = [64, 96, 64, 64], [1, 16, 1024, 4096]
    origins = [(-777, 123), (5000, -9000), (1 << 20, -(1 << 21)), (-(1 << 24), 1 << 23)]
    setting = capi.BatchSetting(SEED, COUNT, CELL, RADIUS, 8, 0.0)
    chunks = (capi.BatchChunk * len(sides))(*[capi.BatchChunk(s, (C.c_int32 * 2)(*o), st)
                                               for s, o, st in zip(sides, origins, steps)])
    h = C.c_void_p()
    capi.call("stp_batch_create", C.byref(NOISE), capi.biome_array(props()), COUNT,
              C.byref(W.erosion_setting(1)), C.byref(setting), chunks, len(sides), C.byref(h))
    try:
        capi.call("stp_batch_run", h, 0, None)
        capi.synchronize()
        for i, (side, o, st) in enumerate(zip(sides, origins, steps)):
            w = side + 2 * RADIUS
            b = np.zeros((w, w), np.uint16)
            capi.call("stp_batch_read_chunk", h, i, None, None, capi.ptr(b), None)
            halo = oracle.voronoi_map(SEED, COUNT, CELL, (o[0] - RADIUS * st, o[1] - RADIUS * st), w, w, step=st)
            assert (b == halo).all(), (side, st)
    finally:
        capi.lib().stp_batch_destroy(h)


def test_batch_graph_replays_are_stable(gpu):
    """Replays of the captured graph with eager work in between give identical results."""
    log("batch graph replays")
    b = Batch(0.0)
    b.run(0)
    first = [b.read(i) for i in range(len(SIDES))]
    for _ in range(3):
        b.run(1)
        b.run(0)
        again = [b.read(i) for i in range(len(SIDES))]
        assert all((x == y).all() for c0, c1 in zip(first, again) for x, y in zip(c0, c1))


# erosion launch shapes: the default LDS-tiled one (128-cell tiles: the 256-cell chunk has four,
# the others one), the smallest tile the colour batches allow at lifetime 30 and r = 3 (72: the
# 96- and 128-cell chunks have four tiles, the last ones cut short, windows clipped at the chunk
# borders), the same tile with 768-thread workgroups (the 25-tap kernel at six waves per SIMD), and
# the global-atomic kernel
TUNINGS = [None, (72, 256), (72, 768), (0, 256)]


@pytest.mark.parametrize("tuning", TUNINGS, ids=["default", "tile72", "tile72x768", "global"])
def test_batch_erosion_against_serialized_oracle(gpu, oracle, tuning):
    """Per chunk: parallel erosion inside the batch vs the serialized FP32 oracle erosion of the
    same pre-erosion heightfield, same seed and stream id (= chunk index).

    Erosion is stochastic and chaotic, so the yardstick is how much two serialized runs with
    DIFFERENT drops differ (measured here per chunk; on these chunks their RMS difference is
    0.4-0.5 of the erosion RMS and the sediment moved varies by ~1/sqrt(drops), 1 sigma):
      * mass: sum(h_after - h_before) == -sum(ledger) to 2e-4 relative (FP32 atomics rounding);
      * sediment moved within 4/sqrt(drops) (4 sigma) + 0.12 of the serialized run; the 0.12 is
        the concurrency deficit of parallel erosion (every drop starts on the unmodified terrain),
        measured with a CPU lockstep emulation at -11% on the 64x64 chunk at (64, 64);
      * RMS(parallel - serialized) <= 1.5 x RMS(serialized - serialized with other drops);
      * height-histogram L1 <= 2 x the same floor + 0.5% of the pixels.
    """
    log(f"batch erosion statistics {tuning}")
    dpp = 0.25
    pre = Batch(0.0)
    pre.run(0)
    b = Batch(dpp)
    if tuning is not None:
        assert b.tuning(*tuning) == tuning
    else:
        assert b.tuning()[0] > 0
    b.run(0)
    cnt = b.counters()
    drops_total = sum(int(s * s * dpp) for s in SIDES)
    assert cnt[0] == sum(s * s for s in SIDES) and cnt[1] == drops_total
    assert cnt[2] <= 30 * drops_total and cnt[5] <= drops_total
    setting = W.erosion_setting(1)
    for i, s in enumerate(SIDES):
        drops = int(s * s * dpp)
        before = pre.read(i)[0]
        after, fixed, _ = b.read(i)
        ledger = b.ledger(i, drops)
        moved = float(ledger.astype(np.float64).sum())
        assert abs((after.astype(np.float64).sum() - before.astype(np.float64).sum()) + moved) <= 2e-4 * moved + 1e-6
        ser = before.copy()
        sled, _ = oracle.erode(ser, (1, 1), (s, s), oracle_rain(oracle, setting), drops, setting.Seed, stream_id=i)
        other = before.copy()
        oracle.erode(other, (1, 1), (s, s), oracle_rain(oracle, setting), drops, setting.Seed, stream_id=1000 + i)
        assert abs(moved - float(sled.astype(np.float64).sum())) <= (4.0 / np.sqrt(drops) + 0.12) * moved, i
        floor_rms = np.sqrt(np.mean((ser.astype(np.float64) - other) ** 2))
        diff_rms = np.sqrt(np.mean((after.astype(np.float64) - ser) ** 2))
        assert diff_rms <= 1.5 * floor_rms, (i, diff_rms, floor_rms)
        floor_hist = np.abs(height_histogram(ser) - height_histogram(other)).sum()
        assert np.abs(height_histogram(after) - height_histogram(ser)).sum() <= 2 * floor_hist + 0.005 * after.size
        assert (fixed == oracle.format_heightmap(after)).all()


def test_batch_erosion_refunds_balance(gpu):
    """Steep biomes (variation 0.9) beside nearly flat ones at depth 0 (variation 1e-4): drops
    running off a slope erode with brush taps on flat cells lower than what the taps want, which
    take the whole cell (the reference's min(height, amount * weight)) through the LDS window's
    take-and-refund path and leave it at exactly 0. Such cells must exist, the heights stay at or
    above 0 (up to one tap's amount, from a refund in flight), and the drops' ledger balances the
    heights to 2e-4 relative."""
    log("batch erosion refunds on flat cells")
    biomes = [(60.0, 6, 0.5, 2.0, 0.0, 0.9), (60.0, 6, 0.5, 2.0, 0.0, 1e-4)] * (COUNT // 2)
    dpp = 1.0
    pre = Batch(0.0, biomes=biomes)
    pre.run(0)
    b = Batch(dpp, biomes=biomes)
    assert b.tuning()[0] > 0
    b.run(0)
    zeroed = 0
    for i, s in enumerate(SIDES):
        before = pre.read(i)[0]
        after = b.read(i)[0]
        assert before.min() >= 0.0
        moved = float(b.ledger(i, int(s * s * dpp)).astype(np.float64).sum())
        assert abs((after.astype(np.float64).sum() - before.astype(np.float64).sum()) + moved) <= 2e-4 * moved + 1e-6, i
        assert after.min() >= -1e-3, i
        zeroed += int(((before > 0) & (after == 0)).sum())
    log(f"cells taken to 0: {zeroed}")
    assert zeroed > 0


def test_batch_negative_heights_use_global_erosion(gpu):
    """A biome reaching below 0 (depth -0.4): the reference's update can drive such terrain
    without bound, which an FP32 map follows and the fixed-point LDS window cannot hold, so the
    batch erodes it with the global-atomic kernel and refuses a tiled tuning."""
    biomes = [(60.0, 6, 0.5, 2.0, -0.4, 0.9)] * COUNT
    b = Batch(0.25, biomes=biomes)
    assert b.tuning()[0] == 0
    with pytest.raises(RuntimeError):
        b.tuning(72, 256)
    b.run(0)


def test_batch_bin_total_matches_oracle(gpu, oracle):
    """stp_batch_bin_total (a measurement aid: the blend's (pixel, biome) noise evaluations) equals
    the oracle histograms' bin count summed over every chunk's pixels; an odd side leaves padding
    pixels after its chunk, which must not count."""
    sides, origins = [61, 96, 64], [(7, -3), (100, 200), (-640, 320)]
    b = Batch(0.0, sides=sides, origins=origins)
    b.run(0)
    total = C.c_uint64()
    capi.call("stp_batch_bin_total", C.c_void_p(b.h), C.byref(total), None)
    expect = 0
    for s, o in zip(sides, origins):
        r = RADIUS
        halo = oracle.voronoi_map(SEED, COUNT, CELL, (o[0] - r, o[1] - r), s + 2 * r, s + 2 * r)
        _, _, off = oracle.shf_window(halo, (r, r), (s, s), r)
        expect += int(np.asarray(off)[-1])
    assert total.value == expect
    status = C.c_uint32(7)
    capi.call("stp_batch_status", C.c_void_p(b.h), C.byref(status), None)
    assert status.value == 0


def test_batch_status_clear_after_tiled_erosion(gpu):
    """The tiled erosion's window status stays clear on the demo's height range (it flags a window
    cell that ends outside [-4, 7.5], where the fixed-point window would wrap)."""
    b = Batch(1.0)
    assert b.tuning()[0] > 0
    b.run(0)
    status = C.c_uint32(7)
    capi.call("stp_batch_status", C.c_void_p(b.h), C.byref(status), None)
    assert status.value == 0


def random_binding(seed):
    """A seeded binding: 1-8 biomes, an even filter radius 2-24, a bin capacity from the biome
    count up to 8 (the item passes; a window never holds more bins than biomes), 3-6 chunks of
    sides 8-300 at levels of detail 1, 2 or 4 and origins anywhere in +-2^20."""
    rng = np.random.default_rng(seed)
    count = int(rng.integers(1, 9))
    radius = 2 * int(rng.integers(1, 13))
    cap = int(rng.integers(count, 9))
    cell = int(rng.integers(16, 200))
    n = int(rng.integers(3, 7))
    sides = [int(v) for v in rng.integers(8, 301, n)]
    steps = [int(v) for v in rng.choice([1, 2, 4], n)]
    origins = [(int(a), int(b)) for a, b in rng.integers(-(1 << 20), 1 << 20, (n, 2))]
    return count, radius, cap, cell, sides, steps, origins


@pytest.mark.parametrize("seed", [101, 202, 303, 404])
def test_batch_random_bindings_bit_exact(gpu, oracle, seed):
    """Seeded random bindings through the item passes (every biome count, slot groups of 4 and 8,
    odd capacities, segmented and unsplit lines, levels of detail): every chunk's biome map,
    heights and UINT16 copy are bit-exact against the oracle in graph and plain mode."""
    count, radius, cap, cell, sides, steps, origins = random_binding(seed)
    log(f"batch random binding seed={seed}: {count} biomes r={radius} cap={cap} cell={cell} sides={sides} steps={steps}")
    setting = capi.BatchSetting(SEED, count, cell, radius, cap, 0.0)
    chunks = (capi.BatchChunk * len(sides))(*[capi.BatchChunk(s, (C.c_int32 * 2)(*o), st)
                                               for s, o, st in zip(sides, origins, steps)])
    h = C.c_void_p()
    capi.call("stp_batch_create", C.byref(NOISE), capi.biome_array(props(count)), count,
              C.byref(W.erosion_setting(1)), C.byref(setting), chunks, len(sides), C.byref(h))
    try:
        perm, g32, _ = oracle.permutation(42, W.DISTRIBUTION, W.GRADIENT_OFFSET, W.ENGINE_MT19937)
        want = []
        for side, o, st in zip(sides, origins, steps):
            w = side + 2 * radius
            halo = oracle.voronoi_map(SEED, count, cell, (o[0] - radius * st, o[1] - radius * st), w, w, step=st)
            item, weight, off = oracle.shf_window(halo, (radius, radius), (side, side), radius)
            half = side * st / 2.0
            hm = oracle.multi_biome(perm, g32, props(count), item, weight, off, (side, side),
                                    (o[0] + half, o[1] + half), step=st)
            want.append((halo, hm))
        for mode in (0, 2):
            capi.call("stp_batch_run", h, mode, None)
            capi.synchronize()
            for i, side in enumerate(sides):
                w = side + 2 * radius
                gh = np.zeros((side, side), np.float32)
                gf = np.zeros((side, side), np.uint16)
                gb = np.zeros((w, w), np.uint16)
                capi.call("stp_batch_read_chunk", h, i, capi.ptr(gh), capi.ptr(gf), capi.ptr(gb), None)
                halo, hm = want[i]
                assert (gb == halo).all(), (mode, i)
                assert gh.tobytes() == hm.tobytes(), (mode, i)
                assert (gf == oracle.format_heightmap(hm)).all(), (mode, i)
    finally:
        capi.lib().stp_batch_destroy(h)


RAGGED_SIDES = [8, 9, 13, 37, 71, 73, 145]
RAGGED_ORIGINS = [(-3, 5), (1000, -1000), (-4097, 33), (250, 250), (-71, -73), (7777, -1), (0, 145)]


@pytest.mark.parametrize("tuning", [(72, 256), (72, 768), (0, 256)], ids=["tile72", "tile72x768", "global"])
def test_batch_ragged_sides(gpu, oracle, tuning):
    """Chunks of the smallest side the erosion brush allows (2r + 2 = 8 at r = 3), odd sides, and
    sides one below, one above and twice the erosion tile plus one (partial tiles on both edges):
    without drops every chunk is bit-exact against the oracle in graph and plain mode; with drops
    (0.25 per pixel, 16 on the smallest chunk) each chunk's ledger balances its heights to 2e-4
    relative, the counters add up, the status words stay clear and the eroded map formats
    bit-exactly."""
    log(f"batch ragged sides {tuning}")
    pre = Batch(0.0, sides=RAGGED_SIDES, origins=RAGGED_ORIGINS)
    before = {}
    for mode in (0, 2):
        pre.run(mode)
        for i in range(len(RAGGED_SIDES)):
            gh, gf, gb = pre.read(i)
            halo, h = oracle_chunk(oracle, i, RAGGED_SIDES, RAGGED_ORIGINS)
            assert (gb == halo).all(), (mode, i)
            assert gh.tobytes() == h.tobytes(), (mode, i)
            assert (gf == oracle.format_heightmap(h)).all(), (mode, i)
            before[i] = gh
    dpp = 0.25
    b = Batch(dpp, sides=RAGGED_SIDES, origins=RAGGED_ORIGINS)
    assert b.tuning(*tuning) == tuning
    b.run(0)
    drops_total = sum(int(s * s * dpp) for s in RAGGED_SIDES)
    cnt = b.counters()
    assert cnt[0] == sum(s * s for s in RAGGED_SIDES) and cnt[1] == drops_total
    assert 0 < cnt[2] <= 30 * drops_total and cnt[5] <= drops_total
    status = C.c_uint32(7)
    capi.call("stp_batch_status", C.c_void_p(b.h), C.byref(status), None)
    assert status.value == 0
    for i, s in enumerate(RAGGED_SIDES):
        after, fixed, _ = b.read(i)
        moved = float(b.ledger(i, int(s * s * dpp)).astype(np.float64).sum())
        assert np.isfinite(after).all(), i
        assert abs((after.astype(np.float64).sum() - before[i].astype(np.float64).sum()) + moved) <= 2e-4 * moved + 1e-6, i
        assert (fixed == oracle.format_heightmap(after)).all(), i


def test_batch_pipelined_graph(gpu, oracle):
    """A binding of at least 8 Mi pixels replays as two chunk groups (split after about half the
    pixels: chunks 0-4 and 5-8 here) whose whole pipelines are concurrent chains of the graph.
    Without drops the chunks on both sides of the split are bit-exact against the oracle, in two
    replays; with drops every chunk's ledger balances its heights to 2e-4 relative and the status
    words stay clear."""
    log("batch pipelined graph, 9 chunks of 1024^2")
    sides = [1024] * 9
    origins = [((k % 3 - 1) * 1024 + 37, (k // 3 - 1) * 1024 - 11) for k in range(9)]
    pre = Batch(0.0, sides=sides, origins=origins)
    for replay in range(2):
        pre.run(0)
        for i in (0, 4, 5, 8):
            h, f, b = pre.read(i)
            halo, want = oracle_chunk(oracle, i, sides, origins)
            assert (b == halo).all(), (replay, i)
            assert h.tobytes() == want.tobytes(), (replay, i)
            assert (f == oracle.format_heightmap(want)).all(), (replay, i)
    dpp = 0.25
    er = Batch(dpp, sides=sides, origins=origins)
    assert er.tuning()[0] > 0
    er.run(0)
    for i, side in enumerate(sides):
        before, after = pre.read(i)[0], er.read(i)[0]
        moved = float(er.ledger(i, int(side * side * dpp)).astype(np.float64).sum())
        assert moved > 0.0, i
        assert abs((after.astype(np.float64).sum() - before.astype(np.float64).sum()) + moved) <= 2e-4 * moved, i
    status = C.c_uint32()
    capi.call("stp_batch_status", C.c_void_p(er.h), C.byref(status), None)
    assert status.value == 0


def test_batch_plain_matches_graph_statistically(gpu, oracle):
    """The plain per-chunk run and the graph run erode with the same drops; only the order of
    concurrent atomic updates differs. Yardstick as above: two serialized runs with different
    drops (RMS) and 4/sqrt(drops) for the sediment moved."""
    log("batch plain vs graph")
    pre = Batch(0.0)
    pre.run(0)
    b = Batch(0.25)
    b.run(2)
    plain = [b.read(i)[0] for i in range(len(SIDES))]
    plain_led = [b.ledger(i, int(s * s * 0.25)).astype(np.float64).sum() for i, s in enumerate(SIDES)]
    lat = (C.c_double * len(SIDES))()
    capi.call("stp_batch_chunk_latencies", C.c_void_p(b.h), lat, len(SIDES))
    assert all(v > 0 for v in lat)
    b.run(0)
    setting = W.erosion_setting(1)
    for i, s in enumerate(SIDES):
        drops = int(s * s * 0.25)
        graph = b.read(i)[0]
        moved = b.ledger(i, drops).astype(np.float64).sum()
        assert abs(moved - plain_led[i]) <= 4.0 / np.sqrt(drops) * moved, i
        before = pre.read(i)[0]
        sa, sb = before.copy(), before.copy()
        oracle.erode(sa, (1, 1), (s, s), oracle_rain(oracle, setting), drops, setting.Seed, stream_id=i)
        oracle.erode(sb, (1, 1), (s, s), oracle_rain(oracle, setting), drops, setting.Seed, stream_id=1000 + i)
        floor_rms = np.sqrt(np.mean((sa.astype(np.float64) - sb) ** 2))
        assert np.sqrt(np.mean((plain[i].astype(np.float64) - graph) ** 2)) <= 1.5 * floor_rms, i


def test_batch_stage_timing(gpu):
    b = Batch(0.25)
    capi.call("stp_batch_set_timing", C.c_void_p(b.h), 1)
    for mode in (0, 1):
        b.run(mode)
        ms = (C.c_float * 7)()
        capi.call("stp_batch_stage_times", C.c_void_p(b.h), ms, 7)
        assert all(v >= 0 for v in ms) and sum(ms) > 0


def test_batch_tuning(gpu):
    """Invalid shapes are refused; the autotuner picks a valid candidate and keeps it; switching
    shapes between graph runs recaptures the graph."""
    b = Batch(0.25)
    default = b.tuning()
    assert default == (128, 1024)
    # 64 < 2 x halo (30 + 3 + 1): windows of one colour batch would overlap
    for bad in [(128, 100), (8, 256), (64, 256), (100000, 256)]:
        with pytest.raises(capi.STPError):
            b.tuning(*bad)
    assert b.tuning() == default
    b.run(0)
    first = b.read(0)[0]
    b.tuning(80, 512)
    b.run(0)
    assert b.tuning() == (80, 512)
    out = capi.BatchTuning()
    capi.call("stp_batch_autotune", C.c_void_p(b.h), 1, None, C.byref(out))
    assert out.ErosionTile > 0 and b.tuning() == (out.ErosionTile, out.ErosionThreads)
    b.run(0)
    capi.synchronize()
    # both runs erode the same terrain with the same drops: comparable mean heights
    again = b.read(0)[0]
    assert np.isfinite(again).all() and abs(again.mean() - first.mean()) < 5e-3


def test_batch_768_threads_any_brush(gpu):
    """768-thread workgroups run a kernel capped at 80 VGPRs: the 25-tap drop fits it, the drop of
    every other brush size spills, so the autotuner offers 768 threads only for the 25-tap brush.
    With a 45-tap brush (r = 4) the autotuner never returns 768 threads, and the shape still runs
    when set by hand: its ledger balances the heights and the status words stay clear."""
    log("batch 768 threads with a 45-tap brush")
    dpp = 0.25
    rain = W.erosion_setting(1)
    rain.ErosionBrushRadius = 4
    setting = capi.BatchSetting(SEED, COUNT, CELL, RADIUS, 8, dpp)
    chunks = (capi.BatchChunk * len(SIDES))(*[capi.BatchChunk(s, (C.c_int32 * 2)(*o)) for s, o in zip(SIDES, ORIGINS)])
    pre_setting = capi.BatchSetting(SEED, COUNT, CELL, RADIUS, 8, 0.0)
    handles = []
    try:
        for st in (pre_setting, setting):
            h = C.c_void_p()
            capi.call("stp_batch_create", C.byref(NOISE), capi.biome_array(props()), COUNT, C.byref(rain),
                      C.byref(st), chunks, len(SIDES), C.byref(h))
            handles.append(h)
        pre, b = handles
        out = capi.BatchTuning()
        capi.call("stp_batch_autotune", b, 1, None, C.byref(out))
        assert out.ErosionThreads != 768 and out.ErosionTile > 0
        capi.call("stp_batch_set_tuning", b, C.byref(capi.BatchTuning(72, 768)))
        for h in handles:
            capi.call("stp_batch_run", h, 0, None)
        capi.synchronize()
        status = C.c_uint32(7)
        capi.call("stp_batch_status", b, C.byref(status), None)
        assert status.value == 0
        for i, side in enumerate(SIDES):
            before = np.zeros((side, side), np.float32)
            after = np.zeros((side, side), np.float32)
            capi.call("stp_batch_read_chunk", pre, i, capi.ptr(before), None, None, None)
            capi.call("stp_batch_read_chunk", b, i, capi.ptr(after), None, None, None)
            drops = int(side * side * dpp)
            ledger = np.zeros(drops, np.float32)
            capi.call("stp_batch_read_chunk", b, i, None, None, None, capi.ptr(ledger))
            moved = float(ledger.astype(np.float64).sum())
            assert np.isfinite(after).all(), i
            assert abs((after.astype(np.float64).sum() - before.astype(np.float64).sum()) + moved) <= 2e-4 * moved + 1e-6, i
    finally:
        for h in handles:
            capi.lib().stp_batch_destroy(h)


def test_batch_tiled_needs_lifetime(gpu):
    """MaxLifetime = 0 (unlimited drops, the reference's behaviour) cannot be tiled: the batch
    starts on the global-atomic kernel and refuses tiled shapes."""
    setting = capi.BatchSetting(SEED, COUNT, CELL, RADIUS, 8, 0.25)
    chunks = (capi.BatchChunk * 1)(capi.BatchChunk(64, (C.c_int32 * 2)(0, 0)))
    rain = W.erosion_setting(1)
    rain.MaxLifetime = 0
    h = C.c_void_p()
    capi.call("stp_batch_create", C.byref(NOISE), capi.biome_array(props()), COUNT, C.byref(rain), C.byref(setting),
              chunks, 1, C.byref(h))
    try:
        t = capi.BatchTuning()
        capi.call("stp_batch_get_tuning", h, C.byref(t))
        assert t.ErosionTile == 0
        with pytest.raises(capi.STPError):
            capi.call("stp_batch_set_tuning", h, C.byref(capi.BatchTuning(64, 256)))
        capi.call("stp_batch_run", h, 0, None)
        capi.synchronize()
    finally:
        capi.lib().stp_batch_destroy(h)


def test_batch_rejects_bad_settings(gpu):
    with pytest.raises(capi.STPError):
        Batch(0.25, radius=3)
    with pytest.raises(capi.STPError):
        Batch(0.25, sides=[4], origins=[(0, 0)])
    # one chunk below the brush's 2r + 2 in a binding of valid ones, a side beyond 16384, no chunks
    with pytest.raises(capi.STPError):
        Batch(0.25, sides=[64, 7], origins=[(0, 0), (64, 0)])
    with pytest.raises(capi.STPError):
        Batch(0.25, sides=[16385], origins=[(0, 0)])
    with pytest.raises(capi.STPError):
        Batch(0.25, sides=[], origins=[])
