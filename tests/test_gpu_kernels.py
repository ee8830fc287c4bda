"""GPU parity of the HIP kernels against the CPU oracle, through the C ABI.

Bars: bit-exact for integer/index work (permutation indices, histograms and their bin order,
free-slip mapping, UINT16 format) and for the FP32 noise and serialized erosion (the oracle's
FP32 restatement uses the same operation order, no FMA contraction); <= 1e-5 max-abs against
the FP64 oracle for noise near the origin; statistical bounds for parallel erosion.
"""
import ctypes as C

import numpy as np
import pytest

from superterrainplus_amd import capi, workloads as W

from helpers import (fractal_info, gpu_erode, gpu_noise, height_histogram, log, oracle_desc,
                     oracle_rain)

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def table(gpu):
    return capi.PermutationTable(42, W.DISTRIBUTION, W.GRADIENT_OFFSET, W.ENGINE_MT19937)


@pytest.fixture(scope="module")
def host_table(oracle):
    perm, g32, g64 = oracle.permutation(42, W.DISTRIBUTION, W.GRADIENT_OFFSET, W.ENGINE_MT19937)
    return perm, g32, g64


# ------------------------------------------------------------------ format
@pytest.mark.parametrize("shape", [(1, 1), (5, 3), (512, 512), (7, 1000), (1536, 1536)])
def test_format_bit_exact(gpu, oracle, shape):
    log(f"format {shape}")
    rng = np.random.default_rng(shape[0] * 7 + shape[1])
    h = rng.uniform(-0.25, 1.25, shape).astype(np.float32)
    flat = h.ravel()
    ties = (np.arange(flat.size, dtype=np.float32) % 65535 + np.float32(0.5)) / np.float32(65535)
    flat[::3] = ties[::3]
    flat[:min(6, flat.size)] = np.array([0, 1, np.nan, np.inf, -np.inf, -0.0], np.float32)[:min(6, flat.size)]
    d_in = capi.DeviceArray.from_host(h)
    d_out = capi.DeviceArray(shape, np.uint16)
    capi.call("stp_format_heightmap", d_in.vp, d_out.vp, shape[1], shape[0], None)
    got = d_out.download()
    assert (got == oracle.format_heightmap(h)).all()


# ------------------------------------------------------------------ noise
@pytest.mark.parametrize("size,offset,octave", [
    ((512, 512), (0.0, 0.0), 8), ((512, 512), (-512.0, 512.0), 8), ((512, 512), (8948.5, -955.5), 8),
    ((1, 1), (3.0, -7.0), 8), ((65, 33), (-1000.25, 77.5), 3), ((300, 2), (0.0, 0.0), 1)])
def test_simplex_heightmap_bit_exact(gpu, oracle, table, host_table, size, offset, octave):
    log(f"simplex {size} {offset} oct={octave}")
    perm, g32, g64 = host_table
    f = W.Fractal(octave=octave)
    d = gpu_noise(table, size[0], size[1], offset, f)
    got = d.download()
    want = oracle.fractal_map(perm, g32, oracle_desc(oracle, size[0], size[1], offset, f), size[0], size[1])
    assert got.tobytes() == want.tobytes()
    if max(abs(offset[0]), abs(offset[1])) + max(size) <= 4096:
        want64 = oracle.fractal_map(perm, g64, oracle_desc(oracle, size[0], size[1], offset, f), size[0], size[1],
                                    f64=True)
        assert np.abs(got - want64).max() <= 1e-5


# ------------------------------------------------------------------ free-slip merge
@pytest.mark.parametrize("dtype", [np.float32, np.uint16])
def test_nn_merge_reference_kat_and_round_trip(gpu, oracle, dtype):
    log(f"nn merge {dtype}")
    chunks = [np.full((2, 2), i, dtype) for i in range(9)]
    d_chunks = [capi.DeviceArray.from_host(c) for c in chunks]
    ptrs = (C.c_void_p * 9)(*[d.ptr for d in d_chunks])
    d_merged = capi.DeviceArray((6, 6), dtype)
    nn = capi.nn_info((2, 2), (3, 3))
    capi.call("stp_nn_merge", ptrs, d_merged.vp, np.dtype(dtype).itemsize, C.byref(nn), None)
    merged = d_merged.download()
    want = np.array([[(x // 2) + (y // 2) * 3 for x in range(6)] for y in range(6)], dtype)
    assert (merged == want).all()
    rng = np.random.default_rng(1)
    chunks = [rng.integers(0, 60000, (33, 17)).astype(dtype) for _ in range(15)]
    d_chunks = [capi.DeviceArray.from_host(c) for c in chunks]
    ptrs = (C.c_void_p * 15)(*[d.ptr for d in d_chunks])
    nn = capi.nn_info((17, 33), (5, 3))
    d_merged = capi.DeviceArray((99, 85), dtype)
    capi.call("stp_nn_merge", ptrs, d_merged.vp, np.dtype(dtype).itemsize, C.byref(nn), None)
    merged = d_merged.download()
    assert (merged == oracle.nn_merge(chunks, (5, 3))).all()
    outs = [capi.DeviceArray((33, 17), dtype) for _ in range(15)]
    optrs = (C.c_void_p * 15)(*[d.ptr for d in outs])
    capi.call("stp_nn_unmerge", d_merged.vp, optrs, np.dtype(dtype).itemsize, C.byref(nn), None)
    assert all((o.download() == c).all() for o, c in zip(outs, chunks))


# ------------------------------------------------------------------ erosion
def test_erosion_serialized_bit_exact_plain_chunk(gpu, oracle, table, host_table):
    """The plain single-chunk workload: 512x512, 8 octaves, 8192 serialized drops."""
    log("erosion serialized 512x512 8192 drops")
    perm, g32, _ = host_table
    cfg = W.plain_chunk()
    d = gpu_noise(table, 512, 512, (0.0, 0.0))
    s = W.erosion_setting(cfg.drops, cfg.erosion_seed)
    led, info = gpu_erode(d, (512, 512), (1, 1), s, mode=1)
    got = d.download()
    want = oracle.fractal_map(perm, g32, oracle_desc(oracle, 512, 512, (0.0, 0.0)), 512, 512)
    wled, stats = oracle.erode(want, (1, 1), (512, 512), oracle_rain(oracle, s), cfg.drops, cfg.erosion_seed)
    assert got.tobytes() == want.tobytes()
    assert led.tobytes() == wled.tobytes()
    steps = info[:, 0] & 0xFFFF
    assert steps.sum() == stats[0] and (info[:, 0] >> 16).sum() == stats[1]
    assert (info[:, 1] & 0xFFFF).sum() == stats[2] and (info[:, 1] >> 16).sum() == stats[3]


def test_erosion_serialized_bit_exact_free_slip(gpu, oracle, table, host_table):
    """3x3 neighbourhood of 96x96 chunks, drops spawned in the centre chunk, stream id 5."""
    log("erosion serialized 3x3 of 96x96")
    perm, g32, _ = host_table
    d = gpu_noise(table, 288, 288, (-96.0, -96.0))
    s = W.erosion_setting(3000, seed=123456789012345)
    led, _ = gpu_erode(d, (96, 96), (3, 3), s, mode=1, stream_id=5)
    want = oracle.fractal_map(perm, g32, oracle_desc(oracle, 288, 288, (-96.0, -96.0)), 288, 288)
    wled, _ = oracle.erode(want, (3, 3), (96, 96), oracle_rain(oracle, s), 3000, s.Seed, stream_id=5)
    assert d.download().tobytes() == want.tobytes()
    assert led.tobytes() == wled.tobytes()


def test_erosion_parallel_free_slip_statistics(gpu, oracle, table, host_table):
    """The free-slip workload: 3x3 chunks of 512x512 at chunk coordinates (-1..1)^2, each sampled
    at its world offset (cx * 512, cy * 512) with half dimension 256, merged to 1536x1536 by the
    device merge kernel (bit-exact against the oracle's per-chunk maps and free-slip index mapping),
    then 65536 drops in parallel (FP32 atomics) against the serialized FP32 oracle on the same input.

    Tolerances (parallel and serialized orders are both valid erosions, but the algorithm is
    chaotic, so cells differ once drop paths diverge):
      * mass: sum(h_after - h_before) == -sum(ledger) to 2e-4 relative (FP32 atomics rounding);
      * total sediment moved within 2% of the serialized run;
      * RMS(parallel - serialized) <= 50% of the serialized run's own erosion RMS;
      * 256-bin height histograms differ in <= 1% of the pixels.
    """
    log("erosion parallel 1536x1536 65536 drops")
    perm, g32, _ = host_table
    cfg = W.free_slip()
    side = cfg.side * 3
    chunks = [gpu_noise(table, cfg.side, cfg.side, off) for off in cfg.chunk_offsets()]
    d = capi.DeviceArray((side, side), np.float32)
    capi.call("stp_nn_merge", (C.c_void_p * 9)(*[c.ptr for c in chunks]), d.vp, 4,
              C.byref(capi.nn_info((cfg.side, cfg.side), (3, 3))), None)
    capi.synchronize()
    want = oracle.nn_merge([oracle.fractal_map(perm, g32, oracle_desc(oracle, cfg.side, cfg.side, off), cfg.side,
                                               cfg.side) for off in cfg.chunk_offsets()], (3, 3))
    before = d.download().astype(np.float64)
    assert before.astype(np.float32).tobytes() == want.tobytes()
    s = W.erosion_setting(cfg.drops, cfg.erosion_seed)
    led, info = gpu_erode(d, (512, 512), (3, 3), s, mode=0)
    after = d.download().astype(np.float64)
    ser = before.astype(np.float32).copy()
    sled, _ = oracle.erode(ser, (3, 3), (512, 512), oracle_rain(oracle, s), cfg.drops, cfg.erosion_seed)
    moved = led.astype(np.float64).sum()
    assert abs((after.sum() - before.sum()) + moved) <= 2e-4 * moved
    assert abs(moved - sled.astype(np.float64).sum()) <= 0.02 * moved
    erosion_rms = np.sqrt(np.mean((ser.astype(np.float64) - before) ** 2))
    diff_rms = np.sqrt(np.mean((after - ser.astype(np.float64)) ** 2))
    print(f"parallel vs serialized: diff RMS {diff_rms:.3e}, erosion RMS {erosion_rms:.3e}, moved {moved:.4f}")
    assert diff_rms <= 0.5 * erosion_rms
    assert np.abs(height_histogram(after) - height_histogram(ser)).sum() <= 0.01 * after.size
    # drops only ever spawn in the centre chunk; no drop outlives the lifetime cap
    assert ((info[:, 0] & 0xFFFF) <= 30).all()


def test_erosion_rejects_mismatched_brush(gpu):
    d = capi.DeviceArray((64, 64), np.float32)
    s = W.erosion_setting(10)
    idx, w = capi.brush(64, 3)
    d_idx, d_w = capi.DeviceArray.from_host(idx[:10]), capi.DeviceArray.from_host(w[:10])
    with pytest.raises(capi.STPError) as e:
        capi.call("stp_hydraulic_erosion", d.vp, C.byref(capi.nn_info((64, 64), (1, 1))), C.byref(s), d_idx.vp,
                  d_w.vp, 10, 0, 0, None, None, None)
    assert e.value.code == 6


# ------------------------------------------------------------------ Voronoi input generator
def test_voronoi_bit_exact(gpu, oracle):
    log("voronoi 2048x2048")
    mb = W.multi_biome()
    d = capi.DeviceArray((mb.side, mb.side), np.uint16)
    capi.call("stp_voronoi_biomemap", mb.biome_seed, mb.biome_count, mb.cell, 0, 0, d.vp, mb.side, mb.side, mb.side,
              None)
    assert (d.download() == oracle.voronoi_map(mb.biome_seed, mb.biome_count, mb.cell, (0, 0), mb.side,
                                              mb.side)).all()
    d2 = capi.DeviceArray((40, 50), np.uint16)
    capi.call("stp_voronoi_biomemap", 3, 5, 7, -1000, 123, d2.vp, 33, 40, 50, None)
    got = d2.download()[:, :33]
    assert (got == oracle.voronoi_map(3, 5, 7, (-1000, 123), 33, 40)).all()
