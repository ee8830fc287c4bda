"""Pins of the splatmap oracle (SuperDemo+/Script/STPSplatmapGenerator.cu:65-137 with the rule lookups of
SuperTerrain+/SuperAlgorithm+/Device/Private/STPTextureSplatRuleWrapper.cu:56-126), by closed forms:

* a constant heightfield has slope factor 1 - 5.5 / |(0, 5.5, 0)| = 0 and altitude v / 65535 +- the
  noise contribution (0.15): rules whose bounds keep clear of that band give a known region; a
  biome without rules takes the next larger sample's rules, past the last sample there is none
  (255); gradient rules outweigh altitude rules; the altitude rule taken is the one with the
  smallest upper bound not below the altitude;
* on a ramp T[x] = k x the four gradient samples land on texels (x - 3, x + 2) and on pairs of
  texels (x - 1, x), so the slope factor is 1 - 5.5 / sqrt((5 k 875.5 / 65535)^2 + 5.5^2).
"""
import numpy as np

from superterrainplus_amd import workloads as W


def perm(oracle):
    p, g32, _ = oracle.permutation(42, W.DISTRIBUTION, W.GRADIENT_OFFSET, W.ENGINE_MT19937)
    return p, g32


def test_rules_on_a_constant_field(oracle):
    p, g = perm(oracle)
    side = 32
    height = np.full((side, 2 * side), 32768, np.uint16)  # altitude 0.5 +- 0.15
    biome = np.zeros_like(height)
    biome[:, :side] = np.arange(side)[None, :] % 6  # biomes 0..5 in the first chunk
    biome[:, side:] = 3
    altitude = [(1, 0.9, 2), (1, 0.2, 1), (3, 1.0, 5), (1, 1.0, 3)]
    gradient = [(3, -0.1, 0.1, 0.0, 1.0, 7), (3, -1.0, 1.0, 0.0, 1.0, 8)]
    out = oracle.splatmap(p, g, altitude, gradient, biome, height, (side, side), [(0, 0, 0.0, 0.0), (1, 0, 32.0, 0.0)])
    expect = {0: 2, 1: 2, 2: 7, 3: 7, 4: 255, 5: 255}  # 0 -> sample 1, 2 -> sample 3, past 3 -> none
    for b, region in expect.items():
        assert (out[:, :side][biome[:, :side] == b] == region).all(), b
    assert (out[:, side:] == 7).all()  # the first matching gradient rule, in the order given


def test_altitude_rule_with_smallest_upper_bound(oracle):
    p, g = perm(oracle)
    side = 16
    biome = np.full((side, side), 4, np.uint16)
    for value, region in [(0, 11), (32768, 12), (65535, 13)]:
        height = np.full((side, side), value, np.uint16)
        # altitudes 0 +0.15 (saturated below), 0.5 +- 0.15, 1 -0.15 (saturated above)
        altitude = [(4, 1.0, 13), (4, 0.16, 11), (4, 0.7, 12)]
        out = oracle.splatmap(p, g, altitude, [], biome, height, (side, side), [(0, 0, 5.0, -3.0)])
        assert (out == region).all(), value


def test_slope_on_a_ramp(oracle):
    p, g = perm(oracle)
    side, k = 64, 300
    height = np.tile((np.arange(side, dtype=np.int64) * k).astype(np.uint16), (side, 1))
    biome = np.zeros_like(height)
    slope = 1.0 - 5.5 / np.sqrt((5.0 * k * 875.5 / 65535.0) ** 2 + 5.5 ** 2)
    gradient = [(0, np.float32(slope - 1e-4), np.float32(slope + 1e-4), -1.0, 2.0, 9)]
    out = oracle.splatmap(p, g, [(0, 2.0, 1)], gradient, biome, height, (side, side), [(0, 0, 0.0, 0.0)])
    assert (out[:, 3:side - 2] == 9).all()  # the clamped border columns see a different slope
    assert (out[:, :2] == 1).all()
