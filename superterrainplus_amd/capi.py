"""ctypes binding of the C ABI in ``include/stp_amd.h``.

This is the harness-side binding a Python caller (the tests, ``bench.py``) uses; it is also the
reference shape for a cgo/JNI/N-API binding: one declaration per entry point, plain pointers
and sizes, status codes checked after every call. It loads the in-tree
``libstp_amd.so`` and raises if it is missing: there is no fallback path.
"""
import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libstp_amd.so")

STATUS = {0: "STP_OK", 1: "STP_ERROR_INVALID_ENVIRONMENT", 2: "STP_ERROR_NUMERIC_DOMAIN",
          3: "STP_ERROR_HIP", 4: "STP_ERROR_CAPACITY", 5: "STP_ERROR_COMPILATION",
          6: "STP_ERROR_INVALID_ARGUMENT", 7: "STP_ERROR_INTERNAL"}


class STPError(RuntimeError):
    def __init__(self, code, message):
        super().__init__(f"{STATUS.get(code, code)}: {message}")
        self.code = code


class SimplexNoiseSetting(C.Structure):
    _fields_ = [("Seed", C.c_uint64), ("Distribution", C.c_uint32), ("Offset", C.c_double),
                ("Engine", C.c_uint32)]


class RainDropSetting(C.Structure):
    _fields_ = [("Seed", C.c_uint64), ("RainDropCount", C.c_uint32), ("Inertia", C.c_float),
                ("SedimentCapacityFactor", C.c_float), ("minSedimentCapacity", C.c_float),
                ("initWaterVolume", C.c_float), ("minWaterVolume", C.c_float),
                ("Friction", C.c_float), ("initSpeed", C.c_float), ("ErodeSpeed", C.c_float),
                ("DepositSpeed", C.c_float), ("EvaporateSpeed", C.c_float),
                ("Gravity", C.c_float), ("ErosionBrushRadius", C.c_uint32),
                ("MaxLifetime", C.c_uint32)]


class ChunkSetting(C.Structure):
    _fields_ = [("ChunkSize", C.c_uint32 * 2), ("MapSize", C.c_uint32 * 2),
                ("ChunkOffset", C.c_double * 3), ("ChunkScale", C.c_double * 2),
                ("MapOffset", C.c_double * 2), ("DiversityNearestNeighbour", C.c_uint32 * 2),
                ("ErosionNearestNeighbour", C.c_uint32 * 2), ("SplatNearestNeighbour", C.c_uint32 * 2),
                ("RenderDistance", C.c_uint32 * 2)]


class NNInformation(C.Structure):
    _fields_ = [("MapSize", C.c_uint32 * 2), ("ChunkNearestNeighbour", C.c_uint32 * 2),
                ("TotalMapSize", C.c_uint32 * 2)]


class FractalInformation(C.Structure):
    _fields_ = [("Persistence", C.c_float), ("Lacunarity", C.c_float), ("Octave", C.c_uint32),
                ("Scale", C.c_float), ("Offset", C.c_float * 2), ("HalfDimension", C.c_float * 2),
                ("Amplitude", C.c_float), ("Frequency", C.c_float)]


class BiomeProperty(C.Structure):
    _fields_ = [("Scale", C.c_float), ("Octave", C.c_uint32), ("Persistence", C.c_float),
                ("Lacunarity", C.c_float), ("Depth", C.c_float), ("Variation", C.c_float)]


class BatchSetting(C.Structure):
    _fields_ = [("BiomeSeed", C.c_uint64), ("BiomeCount", C.c_uint32), ("BiomeCell", C.c_uint32),
                ("Radius", C.c_uint32), ("BinCapacity", C.c_uint32), ("DropsPerPixel", C.c_float),
                ("BiomeLayers", C.c_void_p)]


class BatchChunk(C.Structure):
    _fields_ = [("Side", C.c_uint32), ("Origin", C.c_int32 * 2), ("Step", C.c_uint32)]


class AltitudeRule(C.Structure):
    _fields_ = [("Sample", C.c_uint16), ("UpperBound", C.c_float), ("Region", C.c_uint32)]


class GradientRule(C.Structure):
    _fields_ = [("Sample", C.c_uint16), ("minGradient", C.c_float), ("maxGradient", C.c_float),
                ("LowerBound", C.c_float), ("UpperBound", C.c_float), ("Region", C.c_uint32)]


class SplatChunk(C.Structure):
    _fields_ = [("LocalX", C.c_uint32), ("LocalY", C.c_uint32), ("OffsetX", C.c_float), ("OffsetY", C.c_float)]


class WorldSetting(C.Structure):
    _fields_ = [("MapSize", C.c_uint32), ("ErosionNeighbour", C.c_uint32), ("GenerateChunks", C.c_uint32),
                ("ErodeChunks", C.c_uint32)]


class LayerNode(C.Structure):
    """stp_layer_node: one layer of a biome layer chain."""
    _fields_ = [("Kind", C.c_uint32), ("Param", C.c_uint32), ("Parent", C.c_int32 * 2), ("Salt", C.c_uint64),
                ("CacheSize", C.c_uint64)]


class BiomeRegistry(C.Structure):
    """stp_biome_registry: ids, temperatures, precipitations of the 35 named biomes."""
    _fields_ = [("ID", C.c_uint16 * 35), ("Temperature", C.c_float * 35), ("Precipitation", C.c_float * 35)]


class CustomLayer(C.Structure):
    """stp_custom_layer: a runtime-compiled layer's source and parent footprint."""
    _fields_ = [("Source", C.c_char_p), ("Lo", C.c_int32), ("Hi", C.c_int32), ("Shift", C.c_uint32),
                ("Parents", C.c_uint32), ("Options", C.c_void_p), ("OptionCount", C.c_uint32)]


def demo_layer_program():
    """The demo's 14-layer chain (SuperDemo+/World/Layers/STPAllLayers.cpp:55-110) as LayerNode list."""
    nodes, n = (LayerNode * 64)(), C.c_uint32()
    call("stp_layer_program_demo", nodes, 64, C.byref(n))
    return list(nodes[:n.value])


def demo_biome_registry():
    r = BiomeRegistry()
    call("stp_biome_registry_demo", C.byref(r))
    return r


class ChunkRecord(C.Structure):
    """stp_chunk_record: a parsed chunk record (pointers into the parsed bytes)."""
    _fields_ = [("X", C.c_int32), ("Y", C.c_int32), ("OriginX", C.c_int32), ("OriginY", C.c_int32),
                ("Width", C.c_uint32), ("Height", C.c_uint32), ("Completeness", C.c_uint32), ("Maps", C.c_uint32),
                ("Check", C.c_uint32), ("BiomeMap", C.c_void_p), ("HeightMap", C.c_void_p), ("FixedMap", C.c_void_p),
                ("Bytes", C.c_uint64)]


def read_chunk_records(data):
    """Every record of a byte string of chunk records, validated by the library: a list of dicts
    with the header fields and the present maps as numpy arrays (copies)."""
    import numpy as np
    buf = np.frombuffer(data, np.uint8)
    base = buf.ctypes.data
    out, at = [], 0
    while at < buf.size:
        r = ChunkRecord()
        call("stp_chunk_record_read", C.c_void_p(base + at), buf.size - at, C.byref(r))
        shape = (r.Height, r.Width)
        n = r.Width * r.Height

        def arr(ptr, dtype):
            if not ptr:
                return None
            return np.frombuffer(buf, dtype, n, ptr - base).reshape(shape).copy()
        out.append({"x": r.X, "y": r.Y, "origin": (r.OriginX, r.OriginY), "completeness": r.Completeness,
                    "maps": r.Maps, "check": r.Check, "biome": arr(r.BiomeMap, np.uint16),
                    "height": arr(r.HeightMap, np.float32), "fixed": arr(r.FixedMap, np.uint16)})
        at += r.Bytes
    return out


class ServerSetting(C.Structure):
    _fields_ = [("MemoryBudget", C.c_uint64), ("BatchPixels", C.c_uint64), ("BatchChunks", C.c_uint32),
                ("MaxSide", C.c_uint32), ("Engines", C.c_uint32), ("ChunkWorldSize", C.c_uint32)]


class ChunkRequest(C.Structure):
    _fields_ = [("X", C.c_int32), ("Y", C.c_int32), ("Side", C.c_uint32)]


class ServerStats(C.Structure):
    _fields_ = [("P50", C.c_double), ("P99", C.c_double), ("P999", C.c_double), ("Mean", C.c_double),
                ("Max", C.c_double), ("MpixelPerSecond", C.c_double), ("DeadlineMissFraction", C.c_double),
                ("Requests", C.c_uint64), ("CacheHits", C.c_uint64), ("Batches", C.c_uint64),
                ("Evictions", C.c_uint64), ("MeanBatchChunks", C.c_double), ("PeakResidentBytes", C.c_uint64),
                ("Generated", C.c_uint64), ("GeneratedP50", C.c_double), ("GeneratedP99", C.c_double),
                ("GeneratedP999", C.c_double), ("GeneratedMax", C.c_double),
                ("GeneratedMpixelPerSecond", C.c_double)]


class BatchTuning(C.Structure):
    _fields_ = [("ErosionTile", C.c_uint32), ("ErosionThreads", C.c_uint32)]


class Permutation(C.Structure):
    _fields_ = [("Permutation", C.c_void_p), ("Gradient2D", C.c_void_p),
                ("Gradient2DSize", C.c_uint32)]


class SingleHistogram(C.Structure):
    _fields_ = [("Bin", C.c_void_p), ("HistogramStartOffset", C.c_void_p)]


_P = C.c_void_p
_I = C.c_int
_U32 = C.c_uint32
_U64 = C.c_uint64
_F = C.c_float

_SIGNATURES = {
    "stp_last_error": (C.c_char_p, []),
    "stp_engine_initialise": (_I, [_I]),
    "stp_device_malloc": (_I, [C.POINTER(C.c_void_p), C.c_size_t]),
    "stp_device_free": (_I, [_P]),
    "stp_host_malloc_pinned": (_I, [C.POINTER(C.c_void_p), C.c_size_t]),
    "stp_host_free_pinned": (_I, [_P]),
    "stp_memcpy_async": (_I, [_P, _P, C.c_size_t, _P]),
    "stp_stream_create": (_I, [C.POINTER(C.c_void_p)]),
    "stp_stream_destroy": (_I, [_P]),
    "stp_stream_synchronize": (_I, [_P]),
    "stp_simplex_noise_setting_validate": (_I, [C.POINTER(SimplexNoiseSetting)]),
    "stp_raindrop_setting_validate": (_I, [C.POINTER(RainDropSetting)]),
    "stp_chunk_setting_validate": (_I, [C.POINTER(ChunkSetting)]),
    "stp_permutation_generate": (_I, [C.POINTER(SimplexNoiseSetting), _P, _P, _U32]),
    "stp_permutation_table_create": (_I, [C.POINTER(SimplexNoiseSetting), C.POINTER(C.c_void_p),
                                          C.POINTER(Permutation)]),
    "stp_permutation_table_destroy": (_I, [_P]),
    "stp_erosion_brush_generate": (_I, [_U32, _U32, _P, _P, _U32, C.POINTER(_U32)]),
    "stp_chunk_world_coordinate": (_I, [_P, _P, _P, _P]),
    "stp_chunk_local_coordinate": (_I, [_U32, _P, _P]),
    "stp_chunk_local_origin": (_I, [_P, _P, _P, _P]),
    "stp_chunk_map_offset": (_I, [_P, _P, _P, _P, _P]),
    "stp_chunk_offset": (_I, [_P, _P, _P, _P]),
    "stp_chunk_neighbour_offset": (_I, [_P, _P, _P, _U32]),
    "stp_simplex_heightmap": (_I, [C.POINTER(Permutation), C.POINTER(FractalInformation), _F, _F,
                                   _P, _U32, _U32, _P]),
    "stp_multi_biome_heightmap": (_I, [C.POINTER(Permutation), _P, _U32, C.POINTER(SingleHistogram),
                                       _P, _U32, _U32, _F, _F, _P]),
    "stp_format_heightmap": (_I, [_P, _P, _U32, _U32, _P]),
    "stp_hydraulic_erosion": (_I, [_P, C.POINTER(NNInformation), C.POINTER(RainDropSetting), _P,
                                   _P, _U32, _U32, _I, _P, _P, _P]),
    "stp_nn_merge": (_I, [_P, _P, _U32, C.POINTER(NNInformation), _P]),
    "stp_nn_unmerge": (_I, [_P, _P, _U32, C.POINTER(NNInformation), _P]),
    "stp_voronoi_biomemap": (_I, [_U64, _U32, _U32, C.c_int64, C.c_int64, _P, _U32, _U32, _U32,
                                  _P]),
    "stp_filter_buffer_create": (_I, [_U32, C.POINTER(C.c_void_p)]),
    "stp_filter_buffer_destroy": (_I, [_P]),
    "stp_single_histogram_filter": (_I, [_P, _P, C.POINTER(NNInformation), _U32,
                                         C.POINTER(SingleHistogram), C.POINTER(_U64), _P]),
    "stp_single_histogram_filter_host": (_I, [_P, _P, C.POINTER(NNInformation), _U32,
                                              C.POINTER(SingleHistogram), C.POINTER(_U64)]),
    "stp_diversity_simplex_create": (_I, [C.POINTER(SimplexNoiseSetting), C.POINTER(BiomeProperty),
                                          _P, C.POINTER(C.c_void_p)]),
    "stp_diversity_biomefield_create": (_I, [C.POINTER(SimplexNoiseSetting), _P, _U32, _P, _U32,
                                             _U32, C.POINTER(C.c_void_p)]),
    "stp_diversity_destroy": (_I, [_P]),
    "stp_heightfield_generator_create": (_I, [C.POINTER(ChunkSetting), C.POINTER(RainDropSetting),
                                              _P, C.POINTER(C.c_void_p)]),
    "stp_heightfield_generator_destroy": (_I, [_P]),
    "stp_heightfield_generator_set_erosion_mode": (_I, [_P, _I]),
    "stp_heightfield_generate": (_I, [_P, _P, _P, _F, _F]),
    "stp_heightfield_erode": (_I, [_P, _P, _P]),
    "stp_heightfield_generate_device": (_I, [_P, _P, _P, _F, _F, _P]),
    "stp_heightfield_erode_device": (_I, [_P, _P, _P, _P]),
    "stp_heightfield_erode_device_noalloc": (_I, [_P, _P, _P, _P]),
    "stp_batch_create": (_I, [C.POINTER(SimplexNoiseSetting), _P, _U32, C.POINTER(RainDropSetting),
                              C.POINTER(BatchSetting), _P, _U32, C.POINTER(C.c_void_p)]),
    "stp_batch_destroy": (_I, [_P]),
    "stp_batch_run": (_I, [_P, _I, _P]),
    "stp_batch_set_timing": (_I, [_P, _I]),
    "stp_batch_stage_times": (_I, [_P, _P, _U32]),
    "stp_batch_chunk_latencies": (_I, [_P, _P, _U32]),
    "stp_batch_counters": (_I, [_P, _P, _P]),
    "stp_batch_bin_total": (_I, [_P, _P, _P]),
    "stp_batch_status": (_I, [_P, _P, _P]),
    "stp_batch_read_chunk": (_I, [_P, _U32, _P, _P, _P, _P]),
    "stp_diversity_runtime_create": (_I, [_P, _P, _U32, _P, _U32, _U32, C.c_char_p, _P, _U32, _P]),
    "stp_rtc_compile": (_I, [C.c_char_p, C.c_char_p, _P, _U32, _P, _U32, _P, _P, _U32, C.c_int, _P]),
    "stp_rtc_program_destroy": (_I, [_P]),
    "stp_rtc_program_log": (_I, [_P, C.c_char_p, C.c_uint64, _P]),
    "stp_rtc_lowered_name": (_I, [_P, C.c_char_p, C.c_char_p, C.c_uint64, _P]),
    "stp_rtc_link": (_I, [_P, _P, _U32, _P]),
    "stp_rtc_module_destroy": (_I, [_P]),
    "stp_rtc_module_function": (_I, [_P, C.c_char_p, _P]),
    "stp_rtc_module_global": (_I, [_P, C.c_char_p, _P, _P]),
    "stp_rtc_launch": (_I, [_P, _P, _P, _U32, _P, _P]),
    "stp_server_create": (_I, [_P, _P, _U32, _P, _P, _P, _P]),
    "stp_server_destroy": (_I, [_P]),
    "stp_server_submit": (_I, [_P, _P, _P]),
    "stp_server_poll": (_I, [_P, _P, _P, _U32, _P]),
    "stp_server_flush": (_I, [_P]),
    "stp_server_read": (_I, [_P, _P, _P, _P]),
    "stp_server_replay": (_I, [_P, _P, _P, _U32, C.c_double, _P, _P]),
    "stp_server_memory": (_I, [_P, _P, _P]),
    "stp_batch_set_tuning": (_I, [_P, _P]),
    "stp_batch_get_tuning": (_I, [_P, _P]),
    "stp_batch_autotune": (_I, [_P, _U32, _P, _P]),
    "stp_splat_database_create": (_I, [_P, _U32, _P, _U32, _P]),
    "stp_splat_database_destroy": (_I, [_P]),
    "stp_splatmap_generate": (_I, [_P, _P, _P, _P, _P, _U32, _U32, _U32, _U32, _P, _U32, _P]),
    "stp_world_create": (_I, [_P, _P, _U32, _P, _P, _P, _P]),
    "stp_world_destroy": (_I, [_P]),
    "stp_world_request": (_I, [_P, _P, _U32, _P, _U32, _P]),
    "stp_world_read": (_I, [_P, C.c_int32, C.c_int32, _P, _P]),
    "stp_world_state": (_I, [_P, C.c_int32, C.c_int32, _P]),
    "stp_world_schedule": (_I, [_P, _U32, _U32, _P]),
    "stp_world_render_area": (_I, [_P, C.c_int32, C.c_int32, _U32, _P, _P]),
    "stp_world_stream_word": (_I, [C.c_int32, C.c_int32, _P]),
    "stp_world_load": (_I, [_P, C.c_int32, C.c_int32, _U32, _P]),
    "stp_world_front": (_I, [_P, _P, _P, _P]),
    "stp_server_stream_word": (_I, [C.c_int32, C.c_int32, _U32, _P]),
    "stp_layer_program_demo": (_I, [_P, _U32, _P]),
    "stp_biome_registry_demo": (_I, [_P]),
    "stp_layer_factory_create": (_I, [_P, _U32, _U64, _P, _U32, _U32, _P]),
    "stp_layer_factory_destroy": (_I, [_P]),
    "stp_layer_factory_generate": (_I, [_P, _P, C.c_int32, C.c_int32]),
    "stp_layer_factory_generate_device": (_I, [_P, _P, _P, _U32, _P]),
    "stp_layer_factory_status": (_I, [_P, _P]),
    "stp_layer_factory_create_custom": (_I, [_P, _U32, _U64, _P, _U32, _U32, _P, _U32, _P]),
    "stp_chunk_record_bytes": (_I, [_U32, _U32, _U32, _P]),
    "stp_world_export": (_I, [_P, _P, _U32, _U32, _P, _U64, _P]),
    "stp_chunk_record_read": (_I, [_P, _U64, _P]),
    "stp_splatmap_runtime_create": (_I, [_P, _P, _U32, _U32, C.c_char_p, _P, _U32, _P]),
    "stp_splatmap_runtime_destroy": (_I, [_P]),
    "stp_splatmap_runtime_generate": (_I, [_P, _P, _P, _P, _U32, _U32, _P, _U32, _P]),
}

_lib = None


def lib():
    """The loaded C ABI. Raises if the HIP library has not been built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"HIP library missing: {LIB_PATH}; run __graft_entry__.build()")
        L = C.CDLL(LIB_PATH)
        for name, (res, args) in _SIGNATURES.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


def check(rc):
    if rc != 0:
        raise STPError(rc, lib().stp_last_error().decode())
    return rc


def call(name, *args):
    return check(getattr(lib(), name)(*args))


def ptr(a):
    """Host address of a contiguous numpy array."""
    assert a.flags.c_contiguous
    return C.c_void_p(a.ctypes.data)


# ---------------------------------------------------------------- device memory

class DeviceArray:
    """A device allocation owned by the library's allocator (hipMalloc)."""

    def __init__(self, shape, dtype):
        self.shape = tuple(shape) if isinstance(shape, (tuple, list)) else (int(shape),)
        self.dtype = np.dtype(dtype)
        self.nbytes = int(np.prod(self.shape)) * self.dtype.itemsize
        p = C.c_void_p()
        call("stp_device_malloc", C.byref(p), max(self.nbytes, 1))
        self.ptr = p.value

    @classmethod
    def from_host(cls, a, stream=None):
        a = np.ascontiguousarray(a)
        d = cls(a.shape, a.dtype)
        d.upload(a, stream)
        return d

    def upload(self, a, stream=None):
        a = np.ascontiguousarray(a, self.dtype)
        assert a.nbytes == self.nbytes
        call("stp_memcpy_async", C.c_void_p(self.ptr), ptr(a), self.nbytes, stream)
        call("stp_stream_synchronize", stream)

    def download(self, stream=None):
        out = np.empty(self.shape, self.dtype)
        call("stp_memcpy_async", ptr(out), C.c_void_p(self.ptr), self.nbytes, stream)
        call("stp_stream_synchronize", stream)
        return out

    def offset(self, elements):
        return C.c_void_p(self.ptr + elements * self.dtype.itemsize)

    @property
    def vp(self):
        return C.c_void_p(self.ptr)

    def free(self):
        if self.ptr:
            lib().stp_device_free(C.c_void_p(self.ptr))
            self.ptr = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


def synchronize(stream=None):
    call("stp_stream_synchronize", stream)


# ---------------------------------------------------------------- helpers

def nn_info(map_size, nn):
    return NNInformation((C.c_uint32 * 2)(*map_size), (C.c_uint32 * 2)(*nn),
                         (C.c_uint32 * 2)(map_size[0] * nn[0], map_size[1] * nn[1]))


def permutation_host(seed, distribution, offset, engine=0):
    s = SimplexNoiseSetting(seed, distribution, offset, engine)
    perm = np.zeros(512, np.uint8)
    grad = np.zeros(2 * distribution, np.float32)
    call("stp_permutation_generate", C.byref(s), ptr(perm), ptr(grad), grad.size)
    return perm, grad


class PermutationTable:
    """Device permutation + gradient tables (STPPermutationGenerator::generate)."""

    def __init__(self, seed, distribution, offset, engine=0):
        s = SimplexNoiseSetting(seed, distribution, offset, engine)
        h = C.c_void_p()
        self.view = Permutation()
        call("stp_permutation_table_create", C.byref(s), C.byref(h), C.byref(self.view))
        self.handle = h.value

    def __del__(self):
        try:
            if self.handle:
                lib().stp_permutation_table_destroy(C.c_void_p(self.handle))
        except Exception:
            pass


def brush(dim_x, radius):
    n = _U32()
    call("stp_erosion_brush_generate", dim_x, radius, None, None, 0, C.byref(n))
    idx = np.zeros(n.value, np.int32)
    w = np.zeros(n.value, np.float32)
    call("stp_erosion_brush_generate", dim_x, radius, ptr(idx), ptr(w), n.value, C.byref(n))
    return idx, w


def biome_array(props):
    arr = (BiomeProperty * len(props))()
    for i, p in enumerate(props):
        arr[i] = BiomeProperty(*p)
    return arr


# ---------------------------------------------------------------- runtime compilation (hipRTC)

def _strings(values):
    values = [v.encode() if isinstance(v, str) else v for v in values]
    return (C.c_char_p * max(len(values), 1))(*values), len(values)


class RtcProgram:
    """STPDeviceRuntimeBinary::compile through the C ABI. `headers` maps include names to
    sources; `library_headers` offers the library's device headers. Raises STPError (status
    STP_ERROR_COMPILATION, message holding the compiler log) when the source does not compile."""

    def __init__(self, name, source, options=(), names=(), headers=None, library_headers=False):
        headers = headers or {}
        opt, nopt = _strings(list(options))
        nam, nnam = _strings(list(names))
        hn, nh = _strings(list(headers.keys()))
        hs, _ = _strings(list(headers.values()))
        h = C.c_void_p()
        call("stp_rtc_compile", name.encode(), source.encode(), opt, nopt, nam, nnam, hn, hs, nh,
             1 if library_headers else 0, C.byref(h))
        self.h = h.value

    def _text(self, fn, *args):
        n = C.c_uint64()
        call(fn, C.c_void_p(self.h), *args, None, 0, C.byref(n))
        buf = C.create_string_buffer(n.value + 1)
        call(fn, C.c_void_p(self.h), *args, buf, n.value + 1, C.byref(n))
        return buf.value.decode()

    def log(self):
        return self._text("stp_rtc_program_log")

    def lowered(self, expression):
        return self._text("stp_rtc_lowered_name", expression.encode())

    def __del__(self):
        if getattr(self, "h", None):
            lib().stp_rtc_program_destroy(C.c_void_p(self.h))


class RtcModule:
    """STPDeviceRuntimeProgram::link of compiled programs; kernel and global lookup; launch."""

    CODE_OBJECT, LLVM_BITCODE = 0, 1

    def __init__(self, programs, types):
        self.programs = list(programs)  # kept alive with the module
        arr = (C.c_void_p * len(programs))(*[p.h for p in programs])
        ty = (C.c_int * len(types))(*types)
        h = C.c_void_p()
        call("stp_rtc_link", arr, ty, len(programs), C.byref(h))
        self.h = h.value

    def function(self, lowered):
        f = C.c_void_p()
        call("stp_rtc_module_function", C.c_void_p(self.h), lowered.encode(), C.byref(f))
        return f.value

    def global_(self, lowered):
        p, n = C.c_void_p(), C.c_uint64()
        call("stp_rtc_module_global", C.c_void_p(self.h), lowered.encode(), C.byref(p), C.byref(n))
        return p.value, n.value

    @staticmethod
    def launch(function, grid, block, args, shared=0, stream=None):
        """`args`: ctypes values, one per kernel argument."""
        ptrs = (C.c_void_p * max(len(args), 1))(*[C.cast(C.pointer(a), C.c_void_p) for a in args])
        call("stp_rtc_launch", C.c_void_p(function), (C.c_uint32 * 3)(*grid), (C.c_uint32 * 3)(*block), shared,
             stream, ptrs)

    def __del__(self):
        if getattr(self, "h", None):
            lib().stp_rtc_module_destroy(C.c_void_p(self.h))
