"""Biome layer chains on the GPU against the oracle, bit for bit (the reference's STPBiomeFactory over
STPLayer trees, SuperTerrain+/SuperTerrain+/Private/World/Diversity/STPBiomeFactory.cpp:24-41, with the demo's
layers SuperDemo+/World/Layers/*.h). The GPU evaluates every layer over rectangles, bottom-up; the
oracle recurses pixel by pixel through the layer caches, as the reference does.

Cases: the demo chain (SuperDemo+/World/Layers/STPAllLayers.cpp:55-110) with the demo's registry on
maps at negative, odd and far offsets, non-square and one-pixel maps; a chain that uses every one
of the 23 layer kinds (merges included) with a registry of distinct ids and snowy and dry biomes,
so every rule has something to act on; the MSVC seed hash of the Voronoi layer; the host entry
point; and the reference's errors (empty dimension, cache size not a power of 2)."""
import ctypes as C

import numpy as np
import pytest

from superterrainplus_amd import capi

from helpers import log

pytestmark = pytest.mark.gpu

KINDS = ["Continent", "Scale", "Land", "Island", "Voronoi", "BaseBiome", "ClimateSingle", "ClimateModerate",
         "ClimateExtreme", "DeepOcean", "EaseEdge", "EdgeBiome", "Hills", "Noise", "NoiseToRiver", "OceanNoise",
         "OceanTemperate", "OceanExtreme", "OceanTransition", "OceanMix", "RiverErode", "RiverMix", "SmoothScale"]


def node(kind, salt, parents=(), param=0, cache=0):
    p = list(parents) + [-1] * (2 - len(parents))
    return capi.LayerNode(KINDS.index(kind), param, (C.c_int32 * 2)(*p), salt, cache)


def registry_tuple(r):
    return np.array(r.ID[:]), np.array(r.Temperature[:]), np.array(r.Precipitation[:])


def full_registry(oracle):
    """Distinct ids 0..34; snowy biomes cold and wet, dry biomes dry."""
    r = capi.BiomeRegistry()
    snowy = {"FrozenOcean", "DeepFrozenOcean", "FrozenRiver", "SnowyTaiga", "SnowyTundra", "SnowyMountain", "SnowyBeach",
             "Taiga", "ColdOcean"}
    dry = {"Desert", "DesertHills", "Savannah", "SavannahPlateau", "Badlands", "BadlandsPlateau"}
    for i, b in enumerate(oracle.NAMED_BIOMES):
        r.ID[i] = i
        r.Temperature[i] = 0.5 if b in snowy else 1.5
        r.Precipitation[i] = 0.2 if b in dry else 1.5
    return r


def every_kind_chain():
    """All 23 kinds (each cached as the demo caches its layers, which keeps the oracle's recursion
    linear): land and climate, base biomes, an ocean temperature branch mixed in, hills from a
    noise branch, edges, a river branch mixed in, smoothing and a final Voronoi zoom."""
    return [node("Continent", 23457829, cache=2048),          # 0
            node("Scale", 875944, [0], 1, 2048),                      # 1 fuzzy
            node("Land", 5748329, [1], cache=2048),                   # 2
            node("Scale", 8947358941, [2], cache=2048),               # 3
            node("Island", 74368, [3], cache=2048),                   # 4
            node("ClimateSingle", 11, [4], cache=2048),               # 5
            node("ClimateModerate", 12, [5], cache=2048),             # 6
            node("ClimateExtreme", 13, [6], cache=2048),              # 7
            node("Scale", 14, [7], cache=2048),                       # 8
            node("BaseBiome", 15, [8], cache=2048),                   # 9
            node("DeepOcean", 16, [9], cache=2048),                   # 10
            node("OceanNoise", 17, cache=2048),                       # 11
            node("OceanExtreme", 18, [11], cache=2048),               # 12
            node("OceanTemperate", 19, [12], cache=2048),             # 13
            node("OceanTransition", 20, [13], cache=2048),            # 14
            node("OceanMix", 21, [10, 14], cache=2048),               # 15
            node("Noise", 22, [15], cache=2048),                      # 16
            node("Hills", 23, [15, 16], cache=2048),                  # 17
            node("EaseEdge", 24, [17], cache=2048),                   # 18
            node("EdgeBiome", 25, [18], cache=2048),                  # 19
            node("NoiseToRiver", 26, [16], cache=2048),               # 20
            node("RiverErode", 27, [20], cache=2048),                 # 21
            node("RiverMix", 28, [19, 21], cache=2048),               # 22
            node("SmoothScale", 29, [22], cache=2048),                # 23
            node("Voronoi", 30, [23])]                         # 24


class Factory:
    def __init__(self, nodes, seed, registry, dim):
        arr = (capi.LayerNode * len(nodes))(*nodes)
        h = C.c_void_p()
        capi.call("stp_layer_factory_create", arr, len(nodes), seed, C.byref(registry), dim[0], dim[1], C.byref(h))
        self.h, self.dim = h, dim

    def device(self, offsets):
        w, h = self.dim
        out = capi.DeviceArray((len(offsets), h, w), np.uint16)
        flat = (C.c_int32 * (2 * len(offsets)))(*[v for o in offsets for v in o])
        capi.call("stp_layer_factory_generate_device", self.h, out.vp, flat, len(offsets), None)
        capi.synchronize()
        return out.download()

    def host(self, offset):
        w, h = self.dim
        out = np.zeros((h, w), np.uint16)
        capi.call("stp_layer_factory_generate", self.h, capi.ptr(out), offset[0], offset[1])
        return out

    def __del__(self):
        if getattr(self, "h", None):
            capi.lib().stp_layer_factory_destroy(self.h)


def oracle_maps(oracle, nodes, seed, registry, offsets, dim):
    o = [oracle.layer_node(n.Kind, n.Salt, [p for p in n.Parent if p >= 0], n.Param, n.CacheSize) for n in nodes]
    return oracle.layer_biomemap(o, seed, registry_tuple(registry), offsets, dim[0], dim[1])


@pytest.mark.parametrize("dim", [(160, 96), (1, 1), (67, 130)])
def test_demo_chain_bit_exact(gpu, oracle, dim):
    log(f"layer demo chain {dim}")
    nodes, reg = capi.demo_layer_program(), capi.demo_biome_registry()
    offsets = [(0, 400), (-333, 77), (1000, -5001), (-1000001, 999999), (7, -3)]
    f = Factory(nodes, 6666, reg, dim)
    gpu_maps = f.device(offsets)
    ref = oracle_maps(oracle, nodes, 6666, reg, offsets, dim)
    assert (gpu_maps == ref).all(), [i for i in range(len(offsets)) if (gpu_maps[i] != ref[i]).any()]
    if dim[0] > 1:
        assert len(np.unique(ref)) >= 2
    assert (f.host(offsets[1]) == ref[1]).all()


@pytest.mark.parametrize("seed", [6666, 2 ** 63 + 12345])
def test_every_layer_kind_bit_exact(gpu, oracle, seed):
    log(f"layer every kind seed {seed}")
    nodes, reg = every_kind_chain(), full_registry(oracle)
    dim = (96, 80)
    offsets = [(0, 0), (-2000, 350), (4100, -770)]
    gpu_maps = Factory(nodes, seed, reg, dim).device(offsets)
    ref = oracle_maps(oracle, nodes, seed, reg, offsets, dim)
    assert (gpu_maps == ref).all(), [i for i in range(len(offsets)) if (gpu_maps[i] != ref[i]).any()]
    assert len(np.unique(ref)) >= 5


def test_voronoi_msvc_hash_and_branches(gpu, oracle):
    """The second Voronoi seed hash; and every intermediate layer of the every-kind chain as a root of
    its own (a prefix chain), so each rule is compared on its own output."""
    log("layer voronoi hash, prefixes")
    reg = full_registry(oracle)
    nodes = [node("Continent", 5), node("Scale", 6, [0]), node("Voronoi", 7, [1], 1)]
    dim, offsets = (64, 64), [(-31, 45), (640, -128)]
    assert (Factory(nodes, 99, reg, dim).device(offsets) == oracle_maps(oracle, nodes, 99, reg, offsets, dim)).all()
    full = every_kind_chain()
    for k in range(1, len(full) - 1):
        prefix = full[:k + 1]
        got = Factory(prefix, 6666, reg, (48, 40)).device([(-700, 90)])
        assert (got == oracle_maps(oracle, prefix, 6666, reg, [(-700, 90)], (48, 40))).all(), KINDS[full[k].Kind]


def test_layer_factory_errors(gpu):
    reg = capi.demo_biome_registry()
    with pytest.raises(capi.STPError) as e:
        Factory([node("Continent", 1, cache=3)], 0, reg, (4, 4))
    assert e.value.code == 2  # numeric domain, as STPLayerCache's power-of-2 check
    with pytest.raises(capi.STPError) as e:
        Factory([node("Continent", 1)], 0, reg, (0, 4))
    assert e.value.code == 2
    with pytest.raises(capi.STPError):
        Factory([node("Continent", 1), node("Scale", 2, [1])], 0, reg, (4, 4))  # parent after the node
    with pytest.raises(capi.STPError):
        Factory([node("Continent", 1), node("Hills", 2, [0])], 0, reg, (4, 4))  # hills needs two parents


ISLAND = b"""
__device__ unsigned short sample(const SuperTerrainPlus::STPLayerDevice::STPLayerContext& c, int x, int z) {
	using SuperTerrainPlus::STPDiversity::STPNamedBiome;
	const auto& r = c.registry();
	const unsigned short v = c.parent(0, x, z);
	auto s = c.sampler(x, z);
	return r.shallowOcean(v) && r.shallowOcean(c.parent(0, x, z - 1)) && r.shallowOcean(c.parent(0, x + 1, z))
		&& r.shallowOcean(c.parent(0, x, z + 1)) && r.shallowOcean(c.parent(0, x - 1, z)) && s.nextValue(2) == 0u
		? r[STPNamedBiome::Plains] : v;
}
"""

MIX = b"""
__device__ unsigned short sample(const SuperTerrainPlus::STPLayerDevice::STPLayerContext& c, int x, int z) {
	return static_cast<unsigned short>((c.parent(0, x, z) * 7u + c.parent(1, x, z)) & 0xFFu);
}
"""

# reads 5 pixels east of its declared footprint
WIDE = b"""
__device__ unsigned short sample(const SuperTerrainPlus::STPLayerDevice::STPLayerContext& c, int x, int z) {
	return static_cast<unsigned short>(c.parent(0, x + 5, z) + 1u);
}
"""


def custom_factory(nodes, reg, dim, custom):
    arr = (capi.LayerNode * len(nodes))(*nodes)
    cl = (capi.CustomLayer * len(custom))(*[capi.CustomLayer(src, lo, hi, shift, parents, None, 0)
                                            for src, lo, hi, shift, parents in custom])
    h = C.c_void_p()
    capi.call("stp_layer_factory_create_custom", arr, len(nodes), 6666, C.byref(reg), dim[0], dim[1], cl, len(custom),
              C.byref(h))
    f = Factory.__new__(Factory)
    f.h, f.dim = h, dim
    return f


def test_custom_layer_replaces_a_builtin_one(gpu, oracle):
    """The demo chain with its island layer written as a custom layer (runtime compiled): the maps
    equal the demo chain's on the oracle, bit for bit."""
    log("layer custom island")
    nodes, reg = capi.demo_layer_program(), capi.demo_biome_registry()
    custom = list(nodes)
    custom[7] = capi.LayerNode(23, 0, (C.c_int32 * 2)(6, -1), nodes[7].Salt, nodes[7].CacheSize)
    dim, offsets = (96, 80), [(0, 400), (-333, 77), (5000, -7000)]
    got = custom_factory(custom, reg, dim, [(ISLAND, -1, 1, 0, 1)]).device(offsets)
    assert (got == oracle_maps(oracle, nodes, 6666, reg, offsets, dim)).all()


def test_custom_two_parent_layer(gpu):
    """A custom layer of two parents against its parents' own maps: (a * 7 + b) & 255."""
    log("layer custom two parents")
    reg = capi.demo_biome_registry()
    for i in range(35):
        reg.ID[i] = i
    a, b = node("Continent", 11), node("OceanNoise", 12)
    mix = capi.LayerNode(23, 0, (C.c_int32 * 2)(0, 1), 13, 0)
    dim, offsets = (64, 48), [(-50, 70), (900, -3)]
    got = custom_factory([a, b, mix], reg, dim, [(MIX, 0, 0, 0, 2)]).device(offsets)
    ma = Factory([a], 6666, reg, dim).device(offsets).astype(np.uint32)
    mb = Factory([a, b], 6666, reg, dim).device(offsets).astype(np.uint32)
    assert (got == ((ma * 7 + mb) & 0xFF)).all()
    assert len(np.unique(got)) >= 3


def test_custom_layer_outside_its_footprint_and_errors(gpu):
    log("layer custom footprint, errors")
    reg = capi.demo_biome_registry()
    nodes = [node("Continent", 11), capi.LayerNode(23, 0, (C.c_int32 * 2)(0, -1), 13, 0)]
    f = custom_factory(nodes, reg, (32, 32), [(WIDE, 0, 0, 0, 1)])
    got = f.device([(0, 0)])
    status = C.c_uint32()
    capi.call("stp_layer_factory_status", f.h, C.byref(status))
    assert status.value & 1
    assert (got[0][:, 27:] == 1).all()  # reads past the parent's rectangle give 0
    with pytest.raises(capi.STPError) as e:
        custom_factory(nodes, reg, (32, 32), [(b"__device__ unsigned short sample(", 0, 0, 0, 1)])
    assert e.value.code == 5
    with pytest.raises(capi.STPError):
        custom_factory(nodes, reg, (32, 32), [])  # the node names custom layer 0, none given
