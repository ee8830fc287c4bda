"""GPU single histogram filter and multi-biome blend against the oracle (bit-exact items, bin
order, weights and offsets; blended FP32 heights bit-exact vs the FP32 oracle, <= 1e-5 vs FP64)."""
import ctypes as C

import numpy as np
import pytest

from superterrainplus_amd import capi, workloads as W

from helpers import log
from test_oracle import EPS32, SHF_KAT_EXPECTED, SHF_KAT_TEXTURE

pytestmark = pytest.mark.gpu

BIN = np.dtype([("item", np.uint16), ("pad", np.uint16), ("weight", np.float32)])


class FilterBuffer:
    def __init__(self, capacity):
        h = C.c_void_p()
        capi.call("stp_filter_buffer_create", capacity, C.byref(h))
        self.h = h.value

    def __del__(self):
        capi.lib().stp_filter_buffer_destroy(C.c_void_p(self.h))


def host_filter(buf, sample, nn, map_size, radius):
    s = np.ascontiguousarray(sample, np.uint16)
    out = capi.SingleHistogram()
    total = C.c_uint64()
    info = capi.nn_info(map_size, nn)
    capi.call("stp_single_histogram_filter_host", C.c_void_p(buf.h), capi.ptr(s), C.byref(info), radius,
              C.byref(out), C.byref(total))
    n = total.value
    pixels = map_size[0] * map_size[1]
    bins = np.ctypeslib.as_array((C.c_uint8 * (8 * n)).from_address(out.Bin)).view(BIN).copy() if n else np.zeros(0, BIN)
    off = np.ctypeslib.as_array((C.c_uint32 * (pixels + 1)).from_address(out.HistogramStartOffset)).copy()
    return bins["item"], bins["weight"], off


def device_filter(buf, d_sample, nn, map_size, radius):
    out = capi.SingleHistogram()
    total = C.c_uint64()
    info = capi.nn_info(map_size, nn)
    capi.call("stp_single_histogram_filter", C.c_void_p(buf.h), d_sample.vp, C.byref(info), radius, C.byref(out),
              C.byref(total), None)
    return out, total.value


def test_shf_reference_kat(gpu):
    log("shf kat")
    buf = FilterBuffer(16)
    item, weight, off = host_filter(buf, SHF_KAT_TEXTURE, (3, 3), (4, 4), 2)
    assert off.size == 17 and off[-1] == item.size
    for p, exp in SHF_KAT_EXPECTED.items():
        assert item[off[p]:off[p + 1]].tolist() == [e[0] for e in exp]
        for w, (_, ew) in zip(weight[off[p]:off[p + 1]], exp):
            assert abs(w - ew) <= 5 * EPS32 * ew
    # the same result a second time from the same buffer (STPTestHistogram.cpp:146-149)
    item2, weight2, off2 = host_filter(buf, SHF_KAT_TEXTURE, (3, 3), (4, 4), 2)
    assert (item2 == item).all() and (weight2 == weight).all() and (off2 == off).all()


def test_shf_reference_domain_errors(gpu):
    buf = FilterBuffer(16)
    for r in (0, 128, 3):
        with pytest.raises(capi.STPError) as e:
            host_filter(buf, SHF_KAT_TEXTURE, (3, 3), (4, 4), r)
        assert e.value.code == 2


def blocky(rng, shape, k, p_noise=0.2):
    coarse = rng.integers(0, k, (shape[0] // 3 + 2, shape[1] // 3 + 2))
    m = np.kron(coarse, np.ones((3, 3), np.int64))[:shape[0], :shape[1]].astype(np.uint16)
    noise = rng.random(shape) < p_noise
    m[noise] = rng.integers(0, k, noise.sum())
    return m


@pytest.mark.parametrize("chunk,nn,r,k,cap", [
    ((6, 6), (3, 3), 2, 4, 8), ((64, 40), (3, 3), 8, 6, 8), ((100, 100), (3, 3), 32, 5, 8),
    ((128, 128), (3, 3), 16, 12, 16), ((33, 257), (3, 5), 30, 3, 4), ((64, 64), (5, 5), 64, 7, 8)])
def test_shf_matches_oracle(gpu, oracle, chunk, nn, r, k, cap):
    log(f"shf chunk={chunk} nn={nn} r={r} k={k}")
    rng = np.random.default_rng(sum(chunk) + r + k)
    sample = blocky(rng, (chunk[1] * nn[1], chunk[0] * nn[0]), k)
    item, weight, off = oracle.shf(sample, nn, chunk, r)
    buf = FilterBuffer(cap)
    gi, gw, go = host_filter(buf, sample, nn, chunk, r)
    assert (go == off).all()
    assert (gi == item).all()
    assert gw.tobytes() == weight.tobytes()
    # the device-in / device-out call gives the same arrays
    d = capi.DeviceArray.from_host(sample)
    out, total = device_filter(buf, d, nn, chunk, r)
    assert total == item.size


def test_shf_voronoi_multi_biome_scale(gpu, oracle):
    """A 2048x2048 chunk of the multi-biome workload's biome map with its radius-32 halo."""
    log("shf voronoi 2048 r=32")
    mb = W.multi_biome()
    r = mb.radius
    side = mb.side + 2 * r
    sample = oracle.voronoi_map(mb.biome_seed, mb.biome_count, mb.cell, (-r, -r), side, side)
    # lay the halo tile out as the centre of a 3x3 neighbourhood of 2048 chunks would be seen:
    # the filter only reads the (map + 2r)^2 window, so embed it in a merged map
    merged = np.zeros((3 * mb.side, 3 * mb.side), np.uint16)
    merged[mb.side - r:2 * mb.side + r, mb.side - r:2 * mb.side + r] = sample
    buf = FilterBuffer(8)
    gi, gw, go = host_filter(buf, merged, (3, 3), (mb.side, mb.side), r)
    item, weight, off = oracle.shf(merged, (3, 3), (mb.side, mb.side), r)
    assert (go == off).all() and (gi == item).all() and gw.tobytes() == weight.tobytes()


def test_shf_capacity_overflow_is_an_error(gpu):
    rng = np.random.default_rng(9)
    sample = rng.integers(0, 50, (48, 48)).astype(np.uint16)
    buf = FilterBuffer(4)
    with pytest.raises(capi.STPError) as e:
        host_filter(buf, sample, (3, 3), (16, 16), 4)
    assert e.value.code == 4


@pytest.mark.parametrize("side,r,count", [(256, 16, 4), (64, 8, 8)])
def test_multi_biome_blend_bit_exact(gpu, oracle, side, r, count):
    log(f"multi-biome blend side={side} r={r}")
    props = W.biome_table(count, 7)
    sample = oracle.voronoi_map(7, count, 64, (-side, -side), 3 * side, 3 * side)
    perm, g32, g64 = oracle.permutation(42, W.DISTRIBUTION, W.GRADIENT_OFFSET, W.ENGINE_MT19937)
    item, weight, off = oracle.shf(sample, (3, 3), (side, side), r)
    offset = (float(side), -2.0 * side)
    want = oracle.multi_biome(perm, g32, props, item, weight, off, (side, side), offset)
    want64 = oracle.multi_biome(perm, g64, props, item, weight.astype(np.float64), off, (side, side), offset,
                                f64=True)
    table = capi.PermutationTable(42, W.DISTRIBUTION, W.GRADIENT_OFFSET, W.ENGINE_MT19937)
    buf = FilterBuffer(8)
    d = capi.DeviceArray.from_host(sample)
    hist, _ = device_filter(buf, d, (3, 3), (side, side), r)
    d_h = capi.DeviceArray((side, side), np.float32)
    tbl = capi.biome_array(props)
    capi.call("stp_multi_biome_heightmap", C.byref(table.view), tbl, count, C.byref(hist), d_h.vp, side, side,
              offset[0], offset[1], None)
    got = d_h.download()
    assert got.tobytes() == want.tobytes()
    assert np.abs(got - want64).max() <= 1e-5
