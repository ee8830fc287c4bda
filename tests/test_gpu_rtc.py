"""Runtime compilation on the GPU: a compiled program linked into a module (code object, and
LLVM bitcode through the hipRTC linker), globals filled by lowered name, kernels launched and
checked exactly ((A + B) * s and (A - B) * s in FP32, as SuperTest+/SuperTerrain+/
STPTestRTC.cpp:248-292 checks); and the runtime-compiled multi-biome diversity generator, which
must give the same heightfield as the oracle and the built-in generator, bit for bit."""
import ctypes as C

import numpy as np
import pytest

from superterrainplus_amd import capi, workloads as W

import rtc_matrix as M
from helpers import log
from test_gpu_generator import MAP, NOISE, OFFSETS, Generator, chunk_setting

pytestmark = pytest.mark.gpu

DIM = 4


def matrix_run(module, program, op, a, b, factor):
    """transform<op>(buffer, A, B) then scale(out, buffer, factor) on a 4x4 matrix."""
    dim, nbytes = module.global_(program.lowered("MatrixDimension"))
    assert nbytes >= 8
    dims = np.array([DIM, DIM], np.uint32)
    capi.call("stp_memcpy_async", C.c_void_p(dim), capi.ptr(dims), 8, None)
    da, db = capi.DeviceArray.from_host(a), capi.DeviceArray.from_host(b)
    buf, out = capi.DeviceArray((DIM, DIM), np.float32), capi.DeviceArray((DIM, DIM), np.float32)
    grid, block = (1, 1, 1), (8, 4, 1)
    f = module.function(program.lowered(f"transform<MatrixOp::{op}>"))
    capi.RtcModule.launch(f, grid, block, [C.c_void_p(buf.ptr), C.c_void_p(da.ptr), C.c_void_p(db.ptr)])
    capi.RtcModule.launch(module.function(program.lowered("scale")), grid, block,
                          [C.c_void_p(out.ptr), C.c_void_p(buf.ptr), C.c_float(factor)])
    capi.synchronize()
    return out.download()


@pytest.mark.parametrize("relocatable", [False, True], ids=["code_object", "bitcode"])
def test_matrix_program(gpu, relocatable):
    log(f"rtc matrix program relocatable={relocatable}")
    opts = ["-std=c++17", "-DSTP_TEST_ENABLE"] + (["-fgpu-rdc"] if relocatable else [])
    program = capi.RtcProgram("MatrixOperator", M.SOURCE, opts, M.NAMES)
    module = capi.RtcModule([program], [capi.RtcModule.LLVM_BITCODE if relocatable else capi.RtcModule.CODE_OBJECT])
    rng = np.random.default_rng(6666)
    data = (np.round(rng.uniform(-6666.0, 6666.0, 18) * 10.0) / 10.0).astype(np.float32)
    a = data[:16].reshape(DIM, DIM).copy()
    b = (np.eye(DIM, dtype=np.float32) * data[16]).astype(np.float32)
    s = np.float32(data[17])
    assert matrix_run(module, program, "Add", a, b, s).tobytes() == ((a + b) * s).astype(np.float32).tobytes()
    assert matrix_run(module, program, "Sub", a, b, s).tobytes() == ((a - b) * s).astype(np.float32).tobytes()


def test_link_rejects_mixed_inputs(gpu):
    p = capi.RtcProgram("MatrixOperator", M.SOURCE, ["-DSTP_TEST_ENABLE"], M.NAMES)
    with pytest.raises(capi.STPError):
        capi.RtcModule([p, p], [capi.RtcModule.CODE_OBJECT, capi.RtcModule.CODE_OBJECT])


def runtime_plugin(props, script=None, options=()):
    d = C.c_void_p()
    opt = (C.c_char_p * max(len(options), 1))(*[o.encode() for o in options])
    capi.call("stp_diversity_runtime_create", C.byref(NOISE), capi.biome_array(props), len(props),
              (C.c_uint32 * 2)(MAP, MAP), 16, 8, script.encode() if script else None, opt, len(options), C.byref(d))
    return d.value


def test_runtime_generator_bit_exact(gpu, oracle):
    """The runtime-compiled plugin against the oracle pipeline and the built-in plugin."""
    log("rtc diversity generator")
    props = W.biome_table(4, 7)
    outs = []
    for make in ("runtime", "builtin"):
        if make == "runtime":
            plugin = runtime_plugin(props)
        else:
            d = C.c_void_p()
            capi.call("stp_diversity_biomefield_create", C.byref(NOISE), capi.biome_array(props), 4,
                      (C.c_uint32 * 2)(MAP, MAP), 16, 8, C.byref(d))
            plugin = d.value
        gen = Generator(plugin, chunk_setting(diversity_nn=(3, 3)), W.erosion_setting(100), mode=0)
        biome_chunks = [oracle.voronoi_map(7, 4, 64, (int(off[0]), int(off[1])), MAP, MAP) for off in OFFSETS]
        bp = (C.c_void_p * 9)(*[b.ctypes.data for b in biome_chunks])
        out = np.zeros((MAP, MAP), np.float32)
        capi.call("stp_heightfield_generate", C.c_void_p(gen.h), capi.ptr(out), bp, 128.0, -256.0)
        outs.append(out)
        del gen
        capi.call("stp_diversity_destroy", C.c_void_p(plugin))
    perm, g32, _ = oracle.permutation(42, W.DISTRIBUTION, W.GRADIENT_OFFSET, W.ENGINE_MT19937)
    merged = oracle.nn_merge(biome_chunks, (3, 3))
    item, weight, off = oracle.shf(merged, (3, 3), (MAP, MAP), 16)
    want = oracle.multi_biome(perm, g32, props, item, weight, off, (MAP, MAP), (128.0, -256.0))
    assert outs[0].tobytes() == want.tobytes()
    assert outs[1].tobytes() == outs[0].tobytes()


def test_runtime_generator_ids_beyond_the_table(gpu, oracle):
    """Biome maps with an id the table lacks (ids 0..4, 4 biomes): the segmented filter refuses the
    map and the line-by-line filter hands the script a histogram holding id 4, which the built-in
    script's biome-count guard skips. The oracle's table gets a fifth biome of Depth 0 and
    Variation 0 (its products add +0.0): the same heights, bit for bit."""
    log("rtc diversity generator, ids beyond the table")
    props = W.biome_table(4, 7)
    plugin = runtime_plugin(props)
    try:
        gen = Generator(plugin, chunk_setting(diversity_nn=(3, 3)), W.erosion_setting(100), mode=0)
        biome_chunks = [oracle.voronoi_map(7, 5, 64, (int(off[0]), int(off[1])), MAP, MAP) for off in OFFSETS]
        assert max(int(b.max()) for b in biome_chunks) == 4
        out = np.zeros((MAP, MAP), np.float32)
        capi.call("stp_heightfield_generate", C.c_void_p(gen.h), capi.ptr(out),
                  (C.c_void_p * 9)(*[b.ctypes.data for b in biome_chunks]), 128.0, -256.0)
        del gen
    finally:
        capi.call("stp_diversity_destroy", C.c_void_p(plugin))
    perm, g32, _ = oracle.permutation(42, W.DISTRIBUTION, W.GRADIENT_OFFSET, W.ENGINE_MT19937)
    item, weight, off = oracle.shf(oracle.nn_merge(biome_chunks, (3, 3)), (3, 3), (MAP, MAP), 16)
    zero = (props[0][0], props[0][1], props[0][2], props[0][3], 0.0, 0.0)
    want = oracle.multi_biome(perm, g32, list(props) + [zero], item, weight, off, (MAP, MAP), (128.0, -256.0))
    assert out.tobytes() == want.tobytes()


def test_runtime_generator_script_errors(gpu):
    props = W.biome_table(4, 7)
    # a script that does not compile: the log comes back in the error
    with pytest.raises(capi.STPError) as e:
        runtime_plugin(props, script="__global__ void generateMultiBiomeHeightmap( { }\n")
    assert e.value.code == 5
    # a script without the generator kernel
    with pytest.raises(capi.STPError):
        runtime_plugin(props, script="__constant__ uint2 Dimension[1];\n")
    # a biome table larger than the script's table
    with pytest.raises(capi.STPError) as e:
        runtime_plugin(W.biome_table(8, 7), options=["-DSTP_BIOME_CAPACITY=4"])
    assert e.value.code == 6
    # a smaller table capacity still holding the biomes works
    capi.call("stp_diversity_destroy", C.c_void_p(runtime_plugin(props, options=["-DSTP_BIOME_CAPACITY=4"])))
