"""Writes tests/golden/golden.npz from the CPU oracle.

Run from the repository root after `make -C oracle`:  python tests/golden/make_golden.py
The fixture freezes oracle outputs for the workloads the tests use, so a later change to the
oracle (or to the libstdc++ shuffle it relies on) shows up as a fixture mismatch. The reference
checkout is not read: every value here comes from the oracle or from the reference's own test
data, which is restated as literals in tests/test_oracle.py.
"""
import hashlib
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "oracle"))

import oracle as O  # noqa: E402
from superterrainplus_amd import workloads as W  # noqa: E402


def c0_desc(side):
    d = O.FractalDesc()
    d.Persistence, d.Lacunarity, d.Octave, d.Scale = 0.5, 2.0, 8, 200.0
    d.OffsetX = d.OffsetY = 0.0
    d.HalfX = d.HalfY = side / 2.0
    d.Amplitude = d.Frequency = 1.0
    return d


def rain(drops):
    s = W.erosion_setting(drops)
    r = O.RainDropSetting()
    for f, _ in O.RainDropSetting._fields_:
        setattr(r, f, getattr(s, f))
    return r


def main():
    out = {}
    for eng, name in ((0, "mt19937_64"), (1, "mt19937")):
        p, g32, g64 = O.permutation(42, W.DISTRIBUTION, W.GRADIENT_OFFSET, eng)
        out[f"perm42_{name}"] = p
        out[f"grad12_f32_{name}"] = g32
        out[f"grad12_f64_{name}"] = g64

    side = 512
    perm, g32, g64 = O.permutation(42, W.DISTRIBUTION, W.GRADIENT_OFFSET, W.ENGINE_MT19937)
    n32 = O.fractal_map(perm, g32, c0_desc(side), side, side)
    n64 = O.fractal_map(perm, g64, c0_desc(side), side, side, f64=True)
    rng = np.random.default_rng(1234)
    idx = np.sort(rng.choice(side * side, 64, replace=False)).astype(np.int64)
    out["c0_sample_index"] = idx
    out["c0_noise_f32"] = n32.ravel()[idx]
    out["c0_noise_f64"] = n64.ravel()[idx]

    h = n32.copy()
    ledger, stats = O.erode(h, (1, 1), (side, side), rain(8192), 8192, 42)
    out["c0_eroded_f32"] = h.ravel()[idx]
    out["c0_eroded_sha256"] = np.frombuffer(hashlib.sha256(h.tobytes()).digest(), np.uint8)
    out["c0_erosion_stats"] = stats
    out["c0_ledger_sum"] = np.array([ledger.astype(np.float64).sum()])

    mb = W.multi_biome()
    vm = O.voronoi_map(mb.biome_seed, mb.biome_count, mb.cell, (0, 0), mb.side, mb.side)
    out["voronoi_c2_sha256"] = np.frombuffer(hashlib.sha256(vm.tobytes()).digest(), np.uint8)
    out["voronoi_c2_histogram"] = np.bincount(vm.ravel(), minlength=mb.biome_count)

    np.savez(os.path.join(HERE, "golden.npz"), **out)
    print("wrote", os.path.join(HERE, "golden.npz"), sorted(out))


if __name__ == "__main__":
    main()
