"""ctypes front-end of the CPU oracle (test infrastructure only).

Only the tests, ``__graft_entry__.smoke()`` and the CPU-baseline leg of ``bench.py`` may
import this module. It loads ``oracle/_build/liboracle.so`` (built from
``oracle/stp_oracle.cpp`` by ``oracle/Makefile``) and exposes numpy-friendly wrappers.
Each wrapper names the reference function it restates; the citations with line numbers
are in ``stp_oracle.cpp``.
"""
import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "_build", "liboracle.so")

_lib = None


class FractalDesc(C.Structure):
    """STPSimplexNoise::STPFractalSimplexInformation (STPSimplexNoise.cuh:29-52)."""
    _fields_ = [("Persistence", C.c_float), ("Lacunarity", C.c_float), ("Octave", C.c_uint32),
                ("Scale", C.c_float), ("OffsetX", C.c_float), ("OffsetY", C.c_float),
                ("HalfX", C.c_float), ("HalfY", C.c_float), ("Amplitude", C.c_float),
                ("Frequency", C.c_float)]


class RainDropSetting(C.Structure):
    """Float fields of STPRainDropSetting (STPRainDropSetting.h:15-53) plus MaxLifetime."""
    _fields_ = [("Inertia", C.c_float), ("SedimentCapacityFactor", C.c_float),
                ("minSedimentCapacity", C.c_float), ("initWaterVolume", C.c_float),
                ("minWaterVolume", C.c_float), ("Friction", C.c_float), ("initSpeed", C.c_float),
                ("ErodeSpeed", C.c_float), ("DepositSpeed", C.c_float),
                ("EvaporateSpeed", C.c_float), ("Gravity", C.c_float),
                ("ErosionBrushRadius", C.c_uint32), ("MaxLifetime", C.c_uint32)]


class BiomeProperty(C.Structure):
    """STPDemo::STPBiomeProperty (SuperDemo+/World/Biomes/STPBiomeProperty.hpp:10-27)."""
    _fields_ = [("Scale", C.c_float), ("Octave", C.c_uint32), ("Persistence", C.c_float),
                ("Lacunarity", C.c_float), ("Depth", C.c_float), ("Variation", C.c_float)]


def _p(a, t):
    return a.ctypes.data_as(C.POINTER(t))


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"oracle not built: {LIB_PATH} (run `make -C oracle`)")
        L = C.CDLL(LIB_PATH)
        L.orc_permutation.restype = C.c_int
        L.orc_permutation.argtypes = [C.c_uint64, C.c_uint32, C.c_double, C.c_int, C.c_void_p,
                                      C.c_void_p, C.c_void_p]
        L.orc_floori_f32.restype = C.c_int
        L.orc_floori_f32.argtypes = [C.c_float]
        L.orc_simplex2d_f32.restype = C.c_float
        L.orc_simplex2d_f32.argtypes = [C.c_void_p, C.c_void_p, C.c_uint32, C.c_float, C.c_float]
        L.orc_simplex2d_f64.restype = C.c_double
        L.orc_simplex2d_f64.argtypes = [C.c_void_p, C.c_void_p, C.c_uint32, C.c_double, C.c_double]
        for n in ("orc_fractal_map_f32", "orc_fractal_map_f64"):
            getattr(L, n).restype = None
            getattr(L, n).argtypes = [C.c_void_p, C.c_void_p, C.c_uint32, C.POINTER(FractalDesc),
                                      C.c_uint32, C.c_uint32, C.c_void_p]
        for n in ("orc_brush_f32", "orc_brush_f64"):
            getattr(L, n).restype = C.c_uint32
            getattr(L, n).argtypes = [C.c_uint32, C.c_uint32, C.c_void_p, C.c_void_p, C.c_uint32]
        L.orc_philox4x32_10.restype = None
        L.orc_philox4x32_10.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p]
        for n in ("orc_erode_f32", "orc_erode_f64"):
            getattr(L, n).restype = None
            getattr(L, n).argtypes = [C.c_void_p, C.c_uint32, C.c_uint32, C.c_uint32, C.c_uint32,
                                      C.c_uint32, C.c_uint32, C.POINTER(RainDropSetting),
                                      C.c_void_p, C.c_void_p, C.c_uint32, C.c_uint32, C.c_uint64,
                                      C.c_uint32, C.c_void_p, C.c_void_p]
        for n in ("orc_format_f32", "orc_format_f64"):
            getattr(L, n).restype = None
            getattr(L, n).argtypes = [C.c_void_p, C.c_void_p, C.c_uint64]
        L.orc_shf_f32.restype = C.c_int64
        L.orc_shf_f32.argtypes = [C.c_void_p, C.c_uint32, C.c_uint32, C.c_uint32, C.c_uint32,
                                  C.c_uint32, C.c_uint32, C.c_void_p, C.c_void_p, C.c_void_p,
                                  C.c_uint64]
        L.orc_shf_window_f32.restype = C.c_int64
        L.orc_shf_window_f32.argtypes = [C.c_void_p, C.c_uint32, C.c_uint32, C.c_uint32, C.c_uint32,
                                         C.c_uint32, C.c_uint32, C.c_void_p, C.c_void_p, C.c_void_p,
                                         C.c_uint64]
        for n in ("orc_nn_merge", "orc_nn_unmerge"):
            getattr(L, n).restype = None
            getattr(L, n).argtypes = [C.c_void_p, C.c_void_p, C.c_uint32, C.c_uint32, C.c_uint32,
                                      C.c_uint32, C.c_uint32]
        for n in ("orc_multi_biome_f32", "orc_multi_biome_f64"):
            getattr(L, n).restype = None
            getattr(L, n).argtypes = [C.c_void_p, C.c_void_p, C.c_uint32, C.POINTER(BiomeProperty),
                                      C.c_void_p, C.c_void_p, C.c_void_p, C.c_uint32, C.c_uint32,
                                      C.c_float, C.c_float, C.c_uint32, C.c_void_p]
        L.orc_voronoi_map.restype = None
        L.orc_voronoi_map.argtypes = [C.c_uint64, C.c_uint32, C.c_uint32, C.c_int64, C.c_int64,
                                      C.c_uint32, C.c_uint32, C.c_uint32, C.c_void_p]
        L.orc_chunk_world_coord.restype = None
        L.orc_chunk_world_coord.argtypes = [C.c_double, C.c_double, C.c_uint32, C.c_uint32,
                                            C.c_double, C.c_double, C.c_void_p]
        L.orc_chunk_map_offset.restype = None
        L.orc_chunk_map_offset.argtypes = [C.c_int32, C.c_int32, C.c_uint32, C.c_uint32, C.c_uint32,
                                           C.c_uint32, C.c_double, C.c_double, C.c_void_p]
        L.orc_pipeline_threads.restype = C.c_double
        L.orc_pipeline_threads.argtypes = [C.c_void_p, C.c_void_p, C.c_uint32,
                                           C.POINTER(FractalDesc), C.c_uint32, C.c_uint32,
                                           C.POINTER(RainDropSetting), C.c_uint32, C.c_uint64,
                                           C.c_uint32]
        L.orc_splatmap.restype = None
        L.orc_splatmap.argtypes = [C.c_void_p, C.c_void_p, C.c_uint32, C.c_void_p, C.c_uint32, C.c_void_p,
                                   C.c_uint32, C.c_void_p, C.c_void_p, C.c_void_p, C.c_uint32, C.c_uint32,
                                   C.c_uint32, C.c_uint32, C.c_void_p, C.c_uint32]
        _lib = L
    return _lib


# ---------------------------------------------------------------- wrappers

def permutation(seed, distribution, offset, engine=0):
    """STPPermutationGenerator::generate -> (perm[512] u8, grad f32[2N], grad f64[2N])."""
    perm = np.zeros(512, np.uint8)
    g32 = np.zeros(2 * distribution, np.float32)
    g64 = np.zeros(2 * distribution, np.float64)
    rc = lib().orc_permutation(seed, distribution, offset, engine, perm.ctypes.data,
                               g32.ctypes.data, g64.ctypes.data)
    if rc != 0:
        raise ValueError(f"invalid simplex noise setting (code {rc})")
    return perm, g32, g64


def floori(v):
    return lib().orc_floori_f32(v)


def simplex2d(perm, grad, x, y, f64=False):
    n = grad.size // 2
    if f64:
        return lib().orc_simplex2d_f64(perm.ctypes.data, grad.ctypes.data, n, x, y)
    return lib().orc_simplex2d_f32(perm.ctypes.data, grad.ctypes.data, n, x, y)


def fractal_map(perm, grad, desc, W, H, f64=False):
    out = np.zeros((H, W), np.float64 if f64 else np.float32)
    fn = lib().orc_fractal_map_f64 if f64 else lib().orc_fractal_map_f32
    fn(perm.ctypes.data, grad.ctypes.data, grad.size // 2, C.byref(desc), W, H, out.ctypes.data)
    return out


def brush(dim_x, radius, f64=False):
    cap = (2 * radius + 1) ** 2
    idx = np.zeros(cap, np.int32)
    w = np.zeros(cap, np.float64 if f64 else np.float32)
    fn = lib().orc_brush_f64 if f64 else lib().orc_brush_f32
    n = fn(dim_x, radius, idx.ctypes.data, w.ctypes.data, cap)
    return idx[:n].copy(), w[:n].copy()


def philox(ctr, key):
    c = np.asarray(ctr, np.uint32)
    k = np.asarray(key, np.uint32)
    out = np.zeros(4, np.uint32)
    lib().orc_philox4x32_10(c.ctypes.data, k.ctypes.data, out.ctypes.data)
    return out


def erode(height, nn, map_size, setting, drops, seed, stream_id=0, f64=False):
    """Serialized erosion of the merged map `height` (modified in place).

    Returns (ledger[drops], stats[4] = steps, erode steps, deposit steps, killed drops)."""
    H, W = height.shape
    dt = np.float64 if f64 else np.float32
    assert height.dtype == dt and height.flags.c_contiguous
    idx, w = brush(W, setting.ErosionBrushRadius, f64)
    ledger = np.zeros(drops, dt)
    stats = np.zeros(4, np.uint64)
    fn = lib().orc_erode_f64 if f64 else lib().orc_erode_f32
    fn(height.ctypes.data, W, H, nn[0], nn[1], map_size[0], map_size[1], C.byref(setting),
       idx.ctypes.data, w.ctypes.data, idx.size, drops, seed, stream_id, ledger.ctypes.data,
       stats.ctypes.data)
    return ledger, stats


def format_heightmap(h):
    out = np.zeros(h.shape, np.uint16)
    hc = np.ascontiguousarray(h)
    fn = lib().orc_format_f64 if hc.dtype == np.float64 else lib().orc_format_f32
    fn(hc.ctypes.data, out.ctypes.data, hc.size)
    return out


def shf(sample, nn, map_size, radius):
    """STPSingleHistogramFilter::operator() -> (items u16, weights f32, offsets u32)."""
    s = np.ascontiguousarray(sample, np.uint16)
    mx, my = map_size
    cap = mx * my * min((2 * radius + 1) ** 2, 65536) if radius < 64 else mx * my * 1024
    cap = min(cap, mx * my * 4096)
    item = np.zeros(cap, np.uint16)
    weight = np.zeros(cap, np.float32)
    off = np.zeros(mx * my + 1, np.uint32)
    n = lib().orc_shf_f32(s.ctypes.data, s.shape[1], nn[0], nn[1], mx, my, radius, item.ctypes.data,
                          weight.ctypes.data, off.ctypes.data, cap)
    if n == -1:
        raise ValueError("single histogram filter domain error")
    if n < 0:
        raise RuntimeError("oracle SHF output capacity exceeded")
    return item[:n].copy(), weight[:n].copy(), off


def shf_window(sample, start, map_size, radius, cap_per_pixel=64):
    """Single histogram filter of the map_size window whose first pixel is `start` in `sample`."""
    s = np.ascontiguousarray(sample, np.uint16)
    mx, my = map_size
    cap = mx * my * cap_per_pixel
    item = np.zeros(cap, np.uint16)
    weight = np.zeros(cap, np.float32)
    off = np.zeros(mx * my + 1, np.uint32)
    n = lib().orc_shf_window_f32(s.ctypes.data, s.shape[1], start[0], start[1], mx, my, radius,
                                 item.ctypes.data, weight.ctypes.data, off.ctypes.data, cap)
    if n == -1:
        raise ValueError("single histogram filter domain error")
    if n < 0:
        raise RuntimeError("oracle SHF output capacity exceeded")
    return item[:n].copy(), weight[:n].copy(), off


def nn_merge(chunks, nn):
    chunks = [np.ascontiguousarray(c) for c in chunks]
    my, mx = chunks[0].shape
    out = np.zeros((my * nn[1], mx * nn[0]), chunks[0].dtype)
    arr = (C.c_void_p * len(chunks))(*[c.ctypes.data for c in chunks])
    lib().orc_nn_merge(arr, out.ctypes.data, chunks[0].itemsize, mx, my, nn[0], nn[1])
    return out


def nn_unmerge(merged, nn, map_size):
    mx, my = map_size
    merged = np.ascontiguousarray(merged)
    chunks = [np.zeros((my, mx), merged.dtype) for _ in range(nn[0] * nn[1])]
    arr = (C.c_void_p * len(chunks))(*[c.ctypes.data for c in chunks])
    lib().orc_nn_unmerge(merged.ctypes.data, arr, merged.itemsize, mx, my, nn[0], nn[1])
    return chunks


def biome_table(props):
    arr = (BiomeProperty * len(props))()
    for i, p in enumerate(props):
        arr[i] = BiomeProperty(*p)
    return arr


def multi_biome(perm, grad, props, item, weight, offset, map_size, offset_xy, f64=False, step=1):
    """`step`: world units per pixel (pixel x samples x * step; 1 = the reference's sampling)."""
    mx, my = map_size
    out = np.zeros((my, mx), np.float64 if f64 else np.float32)
    tbl = biome_table(props)
    fn = lib().orc_multi_biome_f64 if f64 else lib().orc_multi_biome_f32
    w = np.ascontiguousarray(weight, np.float64 if f64 else np.float32)
    fn(perm.ctypes.data, grad.ctypes.data, grad.size // 2, tbl, item.ctypes.data, w.ctypes.data,
       offset.ctypes.data, mx, my, offset_xy[0], offset_xy[1], step, out.ctypes.data)
    return out


def voronoi_map(seed, biome_count, cell, origin, W, H, step=1):
    out = np.zeros((H, W), np.uint16)
    lib().orc_voronoi_map(seed, biome_count, cell, origin[0], origin[1], W, H, step, out.ctypes.data)
    return out


def chunk_world_coord(pos_xz, chunk_size, scale):
    out = np.zeros(2, np.int32)
    lib().orc_chunk_world_coord(pos_xz[0], pos_xz[1], chunk_size[0], chunk_size[1], scale[0],
                                scale[1], out.ctypes.data)
    return tuple(int(v) for v in out)


def chunk_map_offset(coord, chunk_size, map_size, map_offset):
    out = np.zeros(2, np.float64)
    lib().orc_chunk_map_offset(coord[0], coord[1], chunk_size[0], chunk_size[1], map_size[0],
                               map_size[1], map_offset[0], map_offset[1], out.ctypes.data)
    return tuple(float(v) for v in out)


def pipeline_threads(perm, grad, desc, chunk_count, side, setting, drops, seed, threads):
    """FP32 CPU pipeline (noise -> serialized erosion -> format) over `chunk_count` chunks."""
    return lib().orc_pipeline_threads(perm.ctypes.data, grad.ctypes.data, grad.size // 2,
                                      C.byref(desc), chunk_count, side, C.byref(setting), drops,
                                      seed, threads)


class AltitudeRule(C.Structure):
    _fields_ = [("Sample", C.c_uint16), ("UpperBound", C.c_float), ("Region", C.c_uint32)]


class GradientRule(C.Structure):
    _fields_ = [("Sample", C.c_uint16), ("minGradient", C.c_float), ("maxGradient", C.c_float),
                ("LowerBound", C.c_float), ("UpperBound", C.c_float), ("Region", C.c_uint32)]


class SplatChunk(C.Structure):
    _fields_ = [("LocalX", C.c_uint32), ("LocalY", C.c_uint32), ("OffsetX", C.c_float), ("OffsetY", C.c_float)]


def splatmap(perm, grad, altitude, gradient, biome, height, map_size, chunks, fill=0):
    """The splatmap generator (SuperDemo+/Script/STPSplatmapGenerator.cu:65-137) over the rendered area:
    `altitude` = [(sample, upper bound, region)], `gradient` = [(sample, min, max, lower, upper, region)],
    `chunks` = [(local x, local y, offset x, offset y)]; pixels of unlisted chunks keep `fill`."""
    b = np.ascontiguousarray(biome, np.uint16)
    h = np.ascontiguousarray(height, np.uint16)
    H, W = h.shape
    out = np.full((H, W), fill, np.uint8)
    a = (AltitudeRule * max(len(altitude), 1))(*[AltitudeRule(*r) for r in altitude])
    g = (GradientRule * max(len(gradient), 1))(*[GradientRule(*r) for r in gradient])
    c = (SplatChunk * max(len(chunks), 1))(*[SplatChunk(*r) for r in chunks])
    lib().orc_splatmap(perm.ctypes.data, grad.ctypes.data, grad.size // 2, a, len(altitude), g, len(gradient),
                       b.ctypes.data, h.ctypes.data, out.ctypes.data, W, H, map_size[0], map_size[1], c, len(chunks))
    return out



LAYER_KINDS = ["Continent", "Scale", "Land", "Island", "Voronoi", "BaseBiome", "ClimateSingle", "ClimateModerate",
               "ClimateExtreme", "DeepOcean", "EaseEdge", "EdgeBiome", "Hills", "Noise", "NoiseToRiver", "OceanNoise",
               "OceanTemperate", "OceanExtreme", "OceanTransition", "OceanMix", "RiverErode", "RiverMix", "SmoothScale"]
# registration order of SuperDemo+/World/Biomes/STPBiomeRegistry.cpp:66-104
NAMED_BIOMES = ["Ocean", "DeepOcean", "WarmOcean", "LukewarmOcean", "ColdOcean", "FrozenOcean", "DeepWarmOcean",
                "DeepLukewarmOcean", "DeepColdOcean", "DeepFrozenOcean", "River", "FrozenRiver", "Plains", "Desert",
                "Mountain", "Forest", "Taiga", "SnowyTaiga", "SnowyTundra", "Jungle", "Savannah", "Swamp", "Badlands",
                "DesertHills", "TaigaHills", "WoodedMountain", "SnowyMountain", "ForestHills", "JungleHills",
                "SavannahPlateau", "SwampHills", "BadlandsPlateau", "Beach", "SnowyBeach", "StoneShore"]


class LayerNode(C.Structure):
    _fields_ = [("Kind", C.c_uint32), ("Param", C.c_uint32), ("Parent", C.c_int32 * 2), ("Salt", C.c_uint64),
                ("CacheSize", C.c_uint64)]


def layer_node(kind, salt, parents=(), param=0, cache=0):
    k = LAYER_KINDS.index(kind) if isinstance(kind, str) else int(kind)
    p = list(parents) + [-1] * (2 - len(parents))
    return LayerNode(k, param, (C.c_int32 * 2)(*p), salt, cache)


def mix_seed(s, fac):
    """STPLayer::mixSeed (SuperTerrain+/SuperTerrain+/Private/World/Diversity/STPLayer.cpp:174-178)."""
    lib().orc_layer_mix_seed.restype = C.c_uint64
    lib().orc_layer_mix_seed.argtypes = [C.c_uint64, C.c_int64]
    return int(lib().orc_layer_mix_seed(s, fac))


def layer_biomemap(nodes, global_seed, registry, origins, width, height):
    """Biome maps of a layer chain's root (the last node), the reference's way: every pixel through
    the cached recursion of STPLayer::retrieve (STPBiomeFactory.cpp:24-41). `registry` =
    (ids, temperatures, precipitations) over NAMED_BIOMES; returns (maps, height, width) uint16."""
    ids, temp, prec = registry
    n = (LayerNode * len(nodes))(*nodes)
    o = np.ascontiguousarray(np.asarray(origins, np.int32).reshape(-1, 2))
    out = np.zeros((len(o), height, width), np.uint16)
    i = np.ascontiguousarray(ids, np.uint16)
    t = np.ascontiguousarray(temp, np.float32)
    p = np.ascontiguousarray(prec, np.float32)
    f = lib().orc_layer_biomemap
    f.restype = C.c_int
    f.argtypes = [C.c_void_p, C.c_uint32, C.c_uint64, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_uint32,
                  C.c_uint32, C.c_uint32, C.c_void_p]
    rc = f(n, len(nodes), global_seed, i.ctypes.data, t.ctypes.data, p.ctypes.data, o.ctypes.data, len(o), width, height,
           out.ctypes.data)
    if rc:
        raise ValueError("malformed layer chain")
    return out
