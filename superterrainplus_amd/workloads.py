"""Synthetic workload definitions for the heightfield path.

Each function returns plain parameters (no device state) so the tests, the benchmark and the
oracle checks all build the same inputs:

* ``plain_chunk``      -- one 512x512 chunk, 8-octave simplex, serialized erosion (8192 drops).
* ``free_slip``        -- a 3x3 block of 512x512 chunks at (-1..1)^2, merged to 1536x1536,
                          65536 parallel drops spawned in the centre chunk.
* ``multi_biome``      -- a 2048x2048 Voronoi biome map (64 jittered sites, seed 7, 8 biomes),
                          per-biome 4-8 octave parameters, filter radius 32, 1,048,576 drops.
* ``batched``          -- 256 chunks with sides from {128, 256, 512, 1024} (p = .4/.3/.2/.1,
                          seed 11), each covering a 1024-unit world slot (level of detail
                          1024 / side), 4-biome Voronoi blend with radius 16, 0.25 drops per pixel.

Randomness that decides workload shapes uses Philox4x32-10 (``philox``) so it does not depend
on a library's generator implementation.
"""
from dataclasses import dataclass, field

import numpy as np

from .capi import RainDropSetting

# gradient table of the demo configuration (SuperDemo+/Biome.ini:4-6)
DISTRIBUTION = 12
GRADIENT_OFFSET = 0.2181661565
ENGINE_MT19937_64 = 0
ENGINE_MT19937 = 1


def philox(ctr, key):
    """Philox4x32-10 (Salmon et al., SC'11) on python ints; returns 4 uint32 words."""
    c = [int(v) & 0xFFFFFFFF for v in ctr]
    k0, k1 = int(key[0]) & 0xFFFFFFFF, int(key[1]) & 0xFFFFFFFF
    for r in range(10):
        if r:
            k0 = (k0 + 0x9E3779B9) & 0xFFFFFFFF
            k1 = (k1 + 0xBB67AE85) & 0xFFFFFFFF
        p0 = 0xD2511F53 * c[0]
        p1 = 0xCD9E8D57 * c[2]
        c = [((p1 >> 32) ^ c[1] ^ k0) & 0xFFFFFFFF, p1 & 0xFFFFFFFF,
             ((p0 >> 32) ^ c[3] ^ k1) & 0xFFFFFFFF, p0 & 0xFFFFFFFF]
    return c


def unit(word):
    """(0, 1] from a 32-bit word, the same mapping the raindrop spawn uses."""
    return ((word >> 8) + 1) * (1.0 / 16777216.0)


def erosion_setting(drops, seed=42, lifetime=30):
    """The erosion constants every configuration shares (brush radius 3, inertia 0.05,
    capacity 4, min capacity 0.01, deposit 0.3, erode 0.3, evaporation 0.01, gravity 4,
    lifetime 30, water 1, speed 1; friction 0 and min water 0.01 map the reference's
    loop condition, SuperDemo+/Engine.ini:60)."""
    return RainDropSetting(Seed=seed, RainDropCount=drops, Inertia=0.05,
                           SedimentCapacityFactor=4.0, minSedimentCapacity=0.01,
                           initWaterVolume=1.0, minWaterVolume=0.01, Friction=0.0, initSpeed=1.0,
                           ErodeSpeed=0.3, DepositSpeed=0.3, EvaporateSpeed=0.01, Gravity=4.0,
                           ErosionBrushRadius=3, MaxLifetime=lifetime)


@dataclass
class Fractal:
    scale: float = 200.0
    octave: int = 8
    persistence: float = 0.5
    lacunarity: float = 2.0
    depth: float = 0.0
    variation: float = 1.0

    def biome(self):
        return (self.scale, self.octave, self.persistence, self.lacunarity, self.depth,
                self.variation)


@dataclass
class PlainChunk:
    side: int = 512
    noise_seed: int = 42
    engine: int = ENGINE_MT19937
    fractal: Fractal = field(default_factory=Fractal)
    drops: int = 8192
    erosion_seed: int = 42


def plain_chunk():
    return PlainChunk()


@dataclass
class FreeSlip:
    side: int = 512
    nn: tuple = (3, 3)
    noise_seed: int = 42
    engine: int = ENGINE_MT19937
    fractal: Fractal = field(default_factory=Fractal)
    drops: int = 65536
    erosion_seed: int = 42

    def chunk_offsets(self):
        """World offsets of the chunks in row-major neighbour order, chunk (cx, cy) in (-1..1)^2
        at (cx * side, cy * side) (STPChunk::calcChunkMapOffset with MapOffset 0)."""
        return [(float(cx * self.side), float(cy * self.side))
                for cy in range(-(self.nn[1] // 2), self.nn[1] // 2 + 1)
                for cx in range(-(self.nn[0] // 2), self.nn[0] // 2 + 1)]


def free_slip():
    return FreeSlip()


def biome_table(count, seed):
    """Seeded per-biome parameters: octave 4..8, scale 100..400, variation 0.2..1.0,
    depth in [0, 1 - variation], persistence 0.5, lacunarity 2.0."""
    out = []
    for b in range(count):
        r = philox((b, 0x42494F4D, 0, 0), (seed & 0xFFFFFFFF, seed >> 32))
        octave = 4 + r[0] % 5
        scale = float(np.float32(100.0 + 300.0 * unit(r[1])))
        variation = float(np.float32(0.2 + 0.8 * unit(r[2])))
        depth = float(np.float32((1.0 - variation) * unit(r[3]) * 0.999))
        out.append((scale, octave, 0.5, 2.0, depth, variation))
    return out


@dataclass
class MultiBiome:
    side: int = 2048
    biome_seed: int = 7
    biome_count: int = 8
    cell: int = 256          # 2048 / 256 = 8 -> 8 x 8 = 64 jittered sites over the chunk
    radius: int = 32
    noise_seed: int = 42
    engine: int = ENGINE_MT19937
    drops: int = 1048576
    erosion_seed: int = 42

    def table(self):
        return biome_table(self.biome_count, self.biome_seed)


def multi_biome():
    return MultiBiome()


@dataclass
class Batched:
    chunks: int = 256
    size_seed: int = 11
    sides: tuple = (128, 256, 512, 1024)
    probs: tuple = (0.4, 0.3, 0.2, 0.1)
    biome_seed: int = 7
    biome_count: int = 4
    cell: int = 256
    radius: int = 16
    noise_seed: int = 42
    engine: int = ENGINE_MT19937
    fractal: Fractal = field(default_factory=Fractal)
    drops_per_pixel: float = 0.25
    erosion_seed: int = 42
    slot: int = 1024
    lod: bool = True

    def chunk_sides(self):
        cum = np.cumsum(self.probs)
        out = []
        for k in range(self.chunks):
            u = unit(philox((k, 0x53495A45, 0, 0), (self.size_seed, 0))[0])
            out.append(self.sides[int(np.searchsorted(cum, u, side="left"))] if u < 1.0 else self.sides[-1])
        return out

    def chunk_origins(self):
        """World origin of chunk k: a 16 x 16 grid of 1024-unit slots centred on the origin,
        so no chunks overlap and world coordinates stay within +-8192 (FP32 noise precision
        degrades linearly with |coordinate|)."""
        return [((k % 16 - 8) * self.slot, (k // 16 - 8) * self.slot) for k in range(self.chunks)]

    def chunk_steps(self):
        """Level of detail of chunk k: it covers its whole 1024-unit slot, so a chunk of side s
        samples every 1024 / s world units (1, 2, 4 or 8); lod=False gives every chunk step 1."""
        return [self.slot // s if self.lod else 1 for s in self.chunk_sides()]

    def table(self):
        return biome_table(self.biome_count, self.biome_seed)


def batched():
    return Batched()


@dataclass
class Served:
    """Camera-driven request stream: a random walk over the chunk grid, 0.5 chunk per 16 ms
    frame (seed 3); every frame requests the chunks within a view radius of 8 chunks whose level
    of detail changed or that were never requested; the level of detail follows the distance
    from the camera (2048 px per side next to it down to 64 px at the edge of the view). The
    requests keep the walk's order and arrive as a Poisson process at `rate` per second. The
    chunk grid's world size is 2048, so a chunk of side s samples every 2048 / s world units."""
    seed: int = 3
    step: float = 0.5
    frame_ms: float = 16.0
    view_radius: float = 8.0
    rates: tuple = (200, 1000, 4000)
    duration_s: float = 2.0
    deadline_ms: float = 16.0
    world: int = 2048
    memory_budget: int = 200 << 30
    batch: Batched = field(default_factory=Batched)
    fly_step: float = 2.0
    flythrough_rates: tuple = (1000, 4000, 8000, 16000)
    fly_duration_s: float = 1.0

    @staticmethod
    def side_at(distance):
        for limit, side in ((1.5, 2048), (2.5, 1024), (3.5, 512), (5.0, 256), (6.5, 128)):
            if distance < limit:
                return side
        return 64

    def walk(self, count):
        """The first `count` requests of the walk: list of (x, y, side)."""
        rng = np.random.default_rng(self.seed)
        cam = np.zeros(2)
        have = {}
        out = []
        r = int(np.ceil(self.view_radius))
        while len(out) < count:
            cx, cy = int(np.floor(cam[0])), int(np.floor(cam[1]))
            view = []
            for y in range(cy - r - 1, cy + r + 2):
                for x in range(cx - r - 1, cx + r + 2):
                    d = float(np.hypot(x + 0.5 - cam[0], y + 0.5 - cam[1]))
                    if d <= self.view_radius:
                        view.append((d, x, y))
            for d, x, y in sorted(view):
                side = self.side_at(d)
                if have.get((x, y)) != side:
                    have[(x, y)] = side
                    out.append((x, y, side))
            a = rng.uniform(0.0, 2.0 * np.pi)
            cam = cam + self.step * np.array([np.cos(a), np.sin(a)])
        return out[:count]

    def trace(self, rate, count=None):
        """(requests, arrival seconds) at `rate` requests per second."""
        n = int(count if count is not None else rate * self.duration_s)
        req = self.walk(n)
        return req, self.arrivals(rate, n)

    def arrivals(self, rate, n):
        gaps = np.random.default_rng(self.seed + int(rate)).exponential(1.0 / rate, n)
        return np.cumsum(gaps) - gaps[0]

    def flythrough(self, count):
        """The first `count` requests of a straight flight east at `fly_step` chunks per frame
        (y = 0.5): a chunk is requested when it enters the view or when its level of detail gets
        finer than the one it holds (a renderer keeps the finer map it has while the camera moves
        away), so no request repeats a (chunk, side) pair."""
        cam = np.array([0.0, 0.5])
        have = {}
        out = []
        r = int(np.ceil(self.view_radius))
        while len(out) < count:
            cx, cy = int(np.floor(cam[0])), int(np.floor(cam[1]))
            view = []
            for y in range(cy - r - 1, cy + r + 2):
                for x in range(cx - r - 1, cx + r + 2):
                    d = float(np.hypot(x + 0.5 - cam[0], y + 0.5 - cam[1]))
                    if d <= self.view_radius:
                        view.append((d, x, y))
            for d, x, y in sorted(view):
                side = self.side_at(d)
                if have.get((x, y), 0) < side:
                    have[(x, y)] = side
                    out.append((x, y, side))
            cam = cam + np.array([self.fly_step, 0.0])
        return out[:count]

    def flythrough_trace(self, rate, count=None):
        n = int(count if count is not None else rate * self.duration_s)
        return self.flythrough(n), self.arrivals(rate, n)


def served():
    return Served()


@dataclass
class WorldWalk:
    """The world pipeline's request pattern at the free-slip configuration's scale: 512-pixel
    chunks, 3 x 3 erosion neighbourhoods, 65536 drops per centre (0.25 per pixel of the centre
    chunk, eroding the 1536^2 merged map), the batched workload's biome map and noise. A camera
    walks `steps` chunks, east for the first half, then north; every frame requests the chunks of
    the render area (radius 2: 5 x 5 chunks) around it, as the reference's pipeline does for its
    render distance (STPWorldPipeline.cpp:640-663)."""
    map: int = 512
    nn: int = 3
    render: int = 2
    steps: int = 24
    generate: int = 64
    erode: int = 16
    batch: Batched = field(default_factory=Batched)

    def frames(self, origin=(0, 0)):
        cam = list(origin)
        out = []
        for s in range(self.steps + 1):
            out.append([(cam[0] + dx, cam[1] + dy) for dy in range(-self.render, self.render + 1)
                        for dx in range(-self.render, self.render + 1)])
            cam[0 if s < self.steps // 2 else 1] += 1
        return out


def world_walk():
    return WorldWalk()

