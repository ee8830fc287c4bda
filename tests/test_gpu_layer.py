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
