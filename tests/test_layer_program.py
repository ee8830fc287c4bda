"""The C ABI's demo layer chain and registry (no GPU needed), against the reference's definitions:
SuperDemo+/World/Layers/STPAllLayers.cpp:53 (cache size 2048) and :82-105 (salts, scale types, 2D
Voronoi layers, an uncached root), and SuperDemo+/Biome.ini with SuperDemo+/Helpers/STPTerrainParaLoader.cpp:271-294
(ids of ocean, plains, desert, forest; the other biomes keep id 0)."""
from superterrainplus_amd import capi

NAMES = ["Continent", "Scale", "Land", "Island", "Voronoi"]


def test_demo_chain_matches_the_demo():
    expect = [("Continent", 23457829, [], 0), ("Scale", 875944, [0], 1), ("Land", 5748329, [1], 0),
              ("Scale", 8947358941, [2], 0), ("Land", 361249673, [3], 0), ("Land", 8769575, [4], 0),
              ("Land", 43562783426564, [5], 0), ("Island", 74368, [6], 0), ("Scale", 1, [7], 0), ("Scale", 2, [8], 0),
              ("Scale", 3, [9], 0), ("Voronoi", 4, [10], 0), ("Voronoi", 5, [11], 0), ("Voronoi", 6, [12], 0)]
    nodes = capi.demo_layer_program()
    assert len(nodes) == len(expect)
    for i, (n, (kind, salt, parents, param)) in enumerate(zip(nodes, expect)):
        assert NAMES[n.Kind] == kind and n.Salt == salt and n.Param == param, i
        assert [p for p in n.Parent if p >= 0] == parents and all(p == -1 for p in n.Parent[len(parents):]), i
        assert n.CacheSize == (0 if i == len(expect) - 1 else 2048), i


def test_demo_registry(oracle):
    r = capi.demo_biome_registry()
    loaded = {"Ocean": (0, 1.5, 1.5), "Plains": (1, 1.5, 1.5), "Desert": (2, 2.8, 0.2), "Forest": (3, 1.5, 1.8)}
    for i, b in enumerate(oracle.NAMED_BIOMES):
        want = loaded.get(b, (0, 0.0, 0.0))
        assert (r.ID[i], round(r.Temperature[i], 3), round(r.Precipitation[i], 3)) == want, b
