"""Benchmark of the heightfield path on one MI355X.

Workload (the batched configuration): 256 square chunks with sides drawn from {128, 256, 512,
1024} (p = 0.4/0.3/0.2/0.1, seed 11), each through the whole path -- 4-biome Voronoi biome map
with a radius-16 halo, single histogram filter (r = 16), 4-biome 8-octave simplex blend,
parallel erosion at 0.25 drops per pixel, UINT16 format -- all chunks resident in HBM.

* value: eroded Mpixel/s of the batched path (one captured hipGraph per step, one launch per
  stage for all chunks).
* plain: the same kernels with one chunk per launch sequence and a host synchronisation after
  every chunk (the reference's generate()/erode() calling pattern); chunk latencies p50/p99.
* roofline: the dominant kernel (erosion), algorithmic bytes per launch from the per-drop step
  counters, divided by its HIP-event duration inside the timed graph replays.
* cpu_baseline: the FP32 oracle running the same per-chunk pipeline on the host cores, on a
  time-bounded sample of the same chunks.

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W] [--plain-steps P] [--cpu-seconds S]
"""
import argparse
import concurrent.futures as cf
import ctypes as C
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "eroded heightfield Mpixel/s + p50/p99 chunk latency, batched vs plain, 1 GPU"
HBM_PEAK_GBS = 8000.0
STAGES = ["biome", "filter_vertical", "filter_horizontal", "blend", "erosion_bin", "erosion", "format"]
# algorithmic bytes of one drop step (r = 3 brush, 25 taps): two 2x2 bilinear height reads;
# an erode step adds 25 x (4 B read + 8 B atomic read-modify-write); a deposit step 4 x 8 B.
BYTES_STEP, BYTES_ERODE, BYTES_DEPOSIT = 32, 25 * 12, 4 * 8
# LDS integer atomics of the tiled erosion: 25 per erode step (one per brush tap), 4 per deposit step
ATOMICS_ERODE, ATOMICS_DEPOSIT = 25, 4
# algorithmic FLOP of the blend (STPSimplexNoise.cu:5-110 as written, FMA = 2, division = 1): 53 per
# simplex2D sample + 13 per octave of the fractal loop, and per histogram bin 3 more (the weight
# product and the sum); one bin of a pixel costs octave * 66 + 3
FLOP_PER_OCTAVE, FLOP_PER_BIN = 66, 3
FP32_PEAK_TFLOPS = 157.3
# HBM bytes per launch of the tiled erosion kernel (tile 128, 1024 threads) on this workload, from
# rocprofv3 --pmc FETCH_SIZE / WRITE_SIZE scaled by a kernel of known size (profiles/erosion_hbm_traffic.md)
TILED_EROSION_TRAFFIC = 493.0e6
# serving loop: two engines of 8 Mpx / 512 chunks each
SERVER_BATCH_PIXELS, SERVER_BATCH_CHUNKS, SERVER_ENGINES = 8 << 20, 512, 2


def workload():
    from superterrainplus_amd import workloads as W
    cfg = W.batched()
    sides = cfg.chunk_sides()
    origins = cfg.chunk_origins()
    props = [(p[0], cfg.fractal.octave, p[2], p[3], p[4], p[5]) for p in cfg.table()]
    return cfg, sides, origins, props


def make_batch(cfg, sides, origins, props, layers=None):
    from superterrainplus_amd import capi, workloads as W
    noise = capi.SimplexNoiseSetting(cfg.noise_seed, W.DISTRIBUTION, W.GRADIENT_OFFSET, cfg.engine)
    # a pixel's histogram can hold at most biome_count distinct biomes
    setting = capi.BatchSetting(cfg.biome_seed, cfg.biome_count, cfg.cell, cfg.radius, cfg.biome_count,
                                cfg.drops_per_pixel, layers)
    chunks = (capi.BatchChunk * len(sides))(*[capi.BatchChunk(s, (C.c_int32 * 2)(*o)) for s, o in zip(sides, origins)])
    h = C.c_void_p()
    capi.call("stp_batch_create", C.byref(noise), capi.biome_array(props), len(props),
              C.byref(W.erosion_setting(1, seed=cfg.erosion_seed)), C.byref(setting), chunks, len(sides), C.byref(h))
    return h


def cpu_baseline(cfg, sides, origins, props, seconds, threads):
    """FP32 oracle pipeline per chunk on `threads` host threads, chunks in workload order until
    `seconds` of wall time have passed."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as O
    from superterrainplus_amd import workloads as W
    perm, g32, _ = O.permutation(cfg.noise_seed, W.DISTRIBUTION, W.GRADIENT_OFFSET, cfg.engine)
    s = W.erosion_setting(1, seed=cfg.erosion_seed)
    rain = O.RainDropSetting()
    for f, _ in O.RainDropSetting._fields_:
        setattr(rain, f, getattr(s, f))
    r = cfg.radius
    deadline = time.perf_counter() + seconds

    def one(i):
        if time.perf_counter() > deadline:
            return 0
        side, o = sides[i], origins[i]
        halo = O.voronoi_map(cfg.biome_seed, cfg.biome_count, cfg.cell, (o[0] - r, o[1] - r), side + 2 * r, side + 2 * r)
        item, weight, off = O.shf_window(halo, (r, r), (side, side), r, cap_per_pixel=8)
        h = O.multi_biome(perm, g32, props, item, weight, off, (side, side), (o[0] + side / 2, o[1] + side / 2))
        drops = int(side * side * cfg.drops_per_pixel)
        O.erode(h, (1, 1), (side, side), rain, drops, s.Seed, stream_id=i)
        O.format_heightmap(h)
        return side * side

    t0 = time.perf_counter()
    with cf.ThreadPoolExecutor(threads) as ex:
        done = list(ex.map(one, range(len(sides))))
    wall = time.perf_counter() - t0
    px = sum(done)
    n = sum(1 for d in done if d)
    return {"value": px / wall / 1e6, "unit": "Mpixel/s", "cores": threads, "kind": "port",
            "sample": f"first {n} of {len(sides)} chunks ({px / 1e6:.2f} Mpixel) of the same workload, FP32 oracle "
                      f"(noise, filter, blend, serialized erosion, format), {wall:.1f} s wall"}


def cpu_baseline_fp64(cfg, sides, origins, props, seconds):
    """The FP64 oracle pipeline on one thread (noise, filter and blend in FP64, serialized FP64
    erosion, format), chunks in workload order until `seconds` of wall time have passed."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as O
    from superterrainplus_amd import workloads as W
    perm, _, g64 = O.permutation(cfg.noise_seed, W.DISTRIBUTION, W.GRADIENT_OFFSET, cfg.engine)
    s = W.erosion_setting(1, seed=cfg.erosion_seed)
    rain = O.RainDropSetting()
    for f, _ in O.RainDropSetting._fields_:
        setattr(rain, f, getattr(s, f))
    r = cfg.radius
    t0, px, n = time.perf_counter(), 0, 0
    for i, (side, o) in enumerate(zip(sides, origins)):
        if time.perf_counter() - t0 > seconds:
            break
        halo = O.voronoi_map(cfg.biome_seed, cfg.biome_count, cfg.cell, (o[0] - r, o[1] - r), side + 2 * r, side + 2 * r)
        item, weight, off = O.shf_window(halo, (r, r), (side, side), r, cap_per_pixel=8)
        h = O.multi_biome(perm, g64, props, item, weight, off, (side, side), (o[0] + side / 2, o[1] + side / 2), f64=True)
        O.erode(h, (1, 1), (side, side), rain, int(side * side * cfg.drops_per_pixel), s.Seed, stream_id=i, f64=True)
        O.format_heightmap(h)
        px += side * side
        n += 1
    wall = time.perf_counter() - t0
    return {"value": round(px / wall / 1e6, 3), "unit": "Mpixel/s", "cores": 1, "kind": "port",
            "sample": f"first {n} of {len(sides)} chunks ({px / 1e6:.2f} Mpixel) of the same workload, FP64 oracle on one "
                      f"thread, {wall:.1f} s wall"}


def served(seconds_scale=1.0):
    """Camera-walk request stream replayed in real time at each rate on a fresh server; a short
    warm-up trace far from the walk comes first so that the timed replay finds warm engines."""
    from superterrainplus_amd import capi, workloads as W
    sv = W.served()
    cfg = sv.batch
    noise = capi.SimplexNoiseSetting(cfg.noise_seed, W.DISTRIBUTION, W.GRADIENT_OFFSET, cfg.engine)
    props = [(p[0], cfg.fractal.octave, p[2], p[3], p[4], p[5]) for p in cfg.table()]
    batch = capi.BatchSetting(cfg.biome_seed, cfg.biome_count, cfg.cell, cfg.radius, cfg.biome_count,
                              cfg.drops_per_pixel)
    setting = capi.ServerSetting(sv.memory_budget, SERVER_BATCH_PIXELS, SERVER_BATCH_CHUNKS, 2048, SERVER_ENGINES,
                                 sv.world)
    out = {}
    for rate in sv.rates:
        h = C.c_void_p()
        capi.call("stp_server_create", C.byref(noise), capi.biome_array(props), len(props),
                  C.byref(W.erosion_setting(1, seed=cfg.erosion_seed)), C.byref(batch), C.byref(setting), C.byref(h))
        try:
            def replay(req, arr):
                r = (capi.ChunkRequest * len(req))(*[capi.ChunkRequest(x, y, side) for x, y, side in req])
                a = (C.c_double * len(req))(*arr)
                lat = (C.c_double * len(req))()
                st = capi.ServerStats()
                capi.call("stp_server_replay", h, r, a, len(req), sv.deadline_ms, lat, C.byref(st))
                return st
            wreq, warr = sv.trace(rate, count=min(200, int(rate * sv.duration_s)))
            replay([(x + 4096, y, side) for x, y, side in wreq], warr)
            req, arr = sv.trace(rate, count=int(rate * sv.duration_s * seconds_scale))
            st = replay(req, arr)
        finally:
            capi.lib().stp_server_destroy(h)
        out[str(rate)] = {"requests": int(st.Requests), "p50_ms": round(st.P50, 3), "p99_ms": round(st.P99, 3),
                          "p99_9_ms": round(st.P999, 3), "max_ms": round(st.Max, 3),
                          "mpixel_s": round(st.MpixelPerSecond, 1),
                          "deadline_miss": round(st.DeadlineMissFraction, 4), "batches": int(st.Batches),
                          "chunks_per_batch": round(st.MeanBatchChunks, 1), "cache_hits": int(st.CacheHits),
                          "evictions": int(st.Evictions)}
        print(f"[bench] served {rate}/s: p50 {st.P50:.2f} p99 {st.P99:.2f} p99.9 {st.P999:.2f} ms, "
              f"{st.MpixelPerSecond:.0f} Mpixel/s, {st.Batches} batches", file=sys.stderr, flush=True)
    return out


def world(plain_frames=6):
    """The world pipeline over a camera walk (workloads.WorldWalk): each frame's request makes the
    render area complete; a warm-up walk elsewhere comes first. The same walk with one chunk per
    batch (its first `plain_frames` frames) is the plain pattern."""
    from superterrainplus_amd import capi, workloads as W
    ww = W.world_walk()
    cfg = ww.batch
    noise = capi.SimplexNoiseSetting(cfg.noise_seed, W.DISTRIBUTION, W.GRADIENT_OFFSET, cfg.engine)
    props = [(p[0], cfg.fractal.octave, p[2], p[3], p[4], p[5]) for p in cfg.table()]
    batch = capi.BatchSetting(cfg.biome_seed, cfg.biome_count, cfg.cell, cfg.radius, cfg.biome_count,
                              cfg.drops_per_pixel)

    export = {}

    def walk(generate, erode, frames, measure_export=False):
        h = C.c_void_p()
        capi.call("stp_world_create", C.byref(noise), capi.biome_array(props), len(props),
                  C.byref(W.erosion_setting(1, seed=cfg.erosion_seed)), C.byref(batch),
                  C.byref(capi.WorldSetting(ww.map, ww.nn, generate, erode)), C.byref(h))
        try:
            def request(chunks):
                flat = (C.c_int32 * (2 * len(chunks)))(*[v for c in chunks for v in c])
                n = C.c_uint32()
                capi.call("stp_world_request", h, flat, len(chunks), None, 0, C.byref(n))
                return n.value
            for f in ww.frames(origin=(5000, 5000))[:3]:
                request(f)
            done, lat, t_all = set(), [], time.perf_counter()
            for f in frames:
                new = [c for c in f if c not in done]
                t0 = time.perf_counter()
                request(f)
                lat.append((time.perf_counter() - t0) * 1e3)
                done.update(new)
            wall = time.perf_counter() - t_all
            if measure_export:
                # chunk records of the last frame's render area, every map, into pinned host memory
                last = frames[-1]
                size = C.c_uint64()
                capi.call("stp_chunk_record_bytes", ww.map, ww.map, 7, C.byref(size))
                total = size.value * len(last)
                host = C.c_void_p()
                capi.call("stp_host_malloc_pinned", C.byref(host), total)
                flat = (C.c_int32 * (2 * len(last)))(*[v for c in last for v in c])
                written = C.c_uint64()
                try:
                    capi.call("stp_world_export", h, flat, len(last), 7, host, total, C.byref(written))
                    t0 = time.perf_counter()
                    reps = 5
                    for _ in range(reps):
                        capi.call("stp_world_export", h, flat, len(last), 7, host, total, C.byref(written))
                    ms = (time.perf_counter() - t0) / reps * 1e3
                finally:
                    capi.lib().stp_host_free_pinned(host)
                export.update({"workload": f"chunk records of the {len(last)} render-area chunks (biome, FP32 and UINT16 "
                                           f"heights, {total / 1e6:.1f} MB) into pinned host memory",
                               "ms": round(ms, 3), "gb_s": round(total / ms / 1e6, 2)})
        finally:
            capi.lib().stp_world_destroy(h)
        return len(done), np.array(lat), wall

    frames = ww.frames()
    n, lat, wall = walk(ww.generate, ww.erode, frames, measure_export=True)
    pn, plat, pwall = walk(1, 1, frames[:plain_frames])
    px = ww.map * ww.map
    out = {"workload": f"camera walk {ww.steps} chunks (east, then north), render area {2 * ww.render + 1}^2 of "
                       f"{ww.map}^2 chunks, {ww.nn}x{ww.nn} free-slip erosion neighbourhoods, "
                       f"{int(px * cfg.drops_per_pixel)} drops per centre, batches of {ww.generate} heightmaps / "
                       f"{ww.erode} neighbourhoods",
           "frames": len(frames), "chunks_completed": n, "first_frame_ms": round(float(lat[0]), 2),
           "frame_p50_ms": round(float(np.percentile(lat[1:], 50)), 2),
           "frame_p99_ms": round(float(np.percentile(lat[1:], 99)), 2),
           "mpixel_s": round(n * px / wall / 1e6, 1),
           "plain_mpixel_s": round(pn * px / pwall / 1e6, 1),
           "plain_frames": plain_frames}
    out["batched_vs_plain"] = round(out["mpixel_s"] / out["plain_mpixel_s"], 2)
    out["export"] = export
    print(f"[bench] world: first frame {lat[0]:.1f} ms, frames p50 {out['frame_p50_ms']} p99 {out['frame_p99_ms']} ms, "
          f"{out['mpixel_s']} Mpixel/s completed (plain {out['plain_mpixel_s']})", file=sys.stderr, flush=True)
    print(f"[bench] world export: {export.get('ms')} ms, {export.get('gb_s')} GB/s", file=sys.stderr, flush=True)
    return out


def layers(reps=10, maps=64, tuning=None, steps=5):
    """Biome maps from the demo's layer chain (14 layers ending in 3 Voronoi zooms, seed 6666, the demo
    registry): `maps` maps of 512^2 at the chunk offsets of an 8 x 8 block, in one call. The oracle's
    recursive evaluation (the reference's algorithm, one production line, one thread) on one map of
    the same block is timed beside it."""
    from superterrainplus_amd import capi
    side = 512
    nodes, reg = capi.demo_layer_program(), capi.demo_biome_registry()
    arr = (capi.LayerNode * len(nodes))(*nodes)
    h = C.c_void_p()
    capi.call("stp_layer_factory_create", arr, len(nodes), 6666, C.byref(reg), side, side, C.byref(h))
    try:
        offsets = [((i % 8) * side, (i // 8) * side) for i in range(maps)]
        flat = (C.c_int32 * (2 * maps))(*[v for o in offsets for v in o])
        out = capi.DeviceArray((maps, side, side), np.uint16)

        def call():
            capi.call("stp_layer_factory_generate_device", h, out.vp, flat, maps, None)
        call()
        capi.synchronize()
        t0 = time.perf_counter()
        for _ in range(reps):
            call()
        capi.synchronize()
        ms = (time.perf_counter() - t0) / reps * 1e3
        first = out.download()[0]
    finally:
        capi.lib().stp_layer_factory_destroy(h)
    mpx = maps * side * side / ms / 1e3
    cpu = None
    try:
        sys.path.insert(0, os.path.join(ROOT, "oracle"))
        import oracle as O
        o = [O.layer_node(n.Kind, n.Salt, [p for p in n.Parent if p >= 0], n.Param, n.CacheSize) for n in nodes]
        t0 = time.perf_counter()
        ref = O.layer_biomemap(o, 6666, (np.array(reg.ID[:]), np.array(reg.Temperature[:]), np.array(reg.Precipitation[:])),
                               [offsets[0]], side, side)[0]
        cpu_s = time.perf_counter() - t0
        cpu = {"mpixel_s": round(side * side / cpu_s / 1e6, 2), "cores": 1, "kind": "port",
               "sample": "one 512^2 map of the same block, recursive cached evaluation", "matches_gpu": bool((ref == first).all())}
    except OSError:
        pass
    print(f"[bench] biome layers: {ms:.3f} ms per {maps} x 512^2 maps, {mpx:.0f} Mpixel/s"
          + (f" (CPU oracle {cpu['mpixel_s']} Mpixel/s)" if cpu else ""), file=sys.stderr, flush=True)
    # the batched workload with its biome maps (chunks and filter halos) from the demo chain
    cfg, sides, origins, props = workload()
    capi.call("stp_layer_factory_create", arr, len(nodes), 6666, C.byref(reg), side, side, C.byref(h))
    batch = make_batch(cfg, sides, origins, props, layers=h)
    try:
        if tuning is not None:
            capi.call("stp_batch_set_tuning", batch, C.byref(tuning))
        capi.call("stp_batch_set_timing", batch, 1)
        capi.call("stp_batch_run", batch, 0, None)
        capi.synchronize()
        stage = np.zeros(len(STAGES))
        t0 = time.perf_counter()
        for _ in range(steps):
            capi.call("stp_batch_run", batch, 0, None)
            capi.synchronize()
            t = (C.c_float * len(STAGES))()
            capi.call("stp_batch_stage_times", batch, t, len(STAGES))
            stage += np.array(t, np.float64)
        wall = (time.perf_counter() - t0) / steps
        status = C.c_uint32()
        capi.call("stp_layer_factory_status", h, C.byref(status))
    finally:
        capi.lib().stp_batch_destroy(batch)
        capi.lib().stp_layer_factory_destroy(h)
    path_mpx = sum(s * s for s in sides) / wall / 1e6
    stage /= steps
    print(f"[bench] batched path with layer biomes: {path_mpx:.0f} Mpixel/s, biome stage {stage[0]:.3f} ms",
          file=sys.stderr, flush=True)
    return {"workload": f"demo biome layer chain (14 layers, 3 Voronoi zooms), {maps} maps of 512^2 per call",
            "ms_per_call": round(ms, 3), "mpixel_s": round(mpx, 1), "cpu_baseline": cpu,
            "batched_path": {"workload": "the batched workload with chunk and halo biome maps from the demo chain "
                             "(4 biome ids) instead of the Voronoi map", "mpixel_s": round(path_mpx, 1),
                             "stage_ms": {k: round(float(v), 3) for k, v in zip(STAGES, stage)},
                             "flagged_ids": int(status.value)}}


def splat(reps=10):
    """The splatmap generator over a rendered area of 5 x 5 chunks of 512^2 (6.6 Mpx): a smooth
    synthetic UINT16 heightfield, biome patches of the 4 biomes, two altitude and one gradient rule
    per biome; all 25 chunks per call."""
    from superterrainplus_amd import capi, workloads as W
    rng = np.random.default_rng(9)
    m, n = 512, 5
    side = m * n
    coarse = rng.uniform(0, 65535, (n * 4 + 1, n * 4 + 1))
    t = np.linspace(0, n * 4, side)
    i0 = np.minimum(t.astype(int), n * 4 - 1)
    f = t - i0
    rows = coarse[i0] * (1 - f)[:, None] + coarse[i0 + 1] * f[:, None]
    height = (rows[:, i0] * (1 - f)[None, :] + rows[:, i0 + 1] * f[None, :]).astype(np.uint16)
    biome = np.repeat(np.repeat(rng.integers(0, 4, (side // 64, side // 64)), 64, 0), 64, 1).astype(np.uint16)
    alt = [(b, u, 4 * b + k) for b in range(4) for k, u in enumerate((0.45, 1.0))]
    gra = [(b, 0.2, 1.0, 0.0, 1.0, 4 * b + 3) for b in range(4)]
    a = (capi.AltitudeRule * len(alt))(*[capi.AltitudeRule(*r) for r in alt])
    g = (capi.GradientRule * len(gra))(*[capi.GradientRule(*r) for r in gra])
    db = C.c_void_p()
    capi.call("stp_splat_database_create", a, len(alt), g, len(gra), C.byref(db))
    table = capi.PermutationTable(42, W.DISTRIBUTION, W.GRADIENT_OFFSET, W.ENGINE_MT19937)
    try:
        b, h = capi.DeviceArray.from_host(biome), capi.DeviceArray.from_host(height)
        out = capi.DeviceArray((side, side), np.uint8)
        chunks = (capi.SplatChunk * (n * n))(*[capi.SplatChunk(x, y, float(x * m), float(y * m))
                                              for y in range(n) for x in range(n)])

        def call():
            capi.call("stp_splatmap_generate", C.byref(table.view), db, b.vp, h.vp, out.vp, side, side, m, m, chunks,
                      n * n, None)
        call()
        capi.synchronize()
        t0 = time.perf_counter()
        for _ in range(reps):
            call()
        capi.synchronize()
        ms = (time.perf_counter() - t0) / reps * 1e3
    finally:
        capi.lib().stp_splat_database_destroy(db)
    mpx = side * side / ms / 1e3
    print(f"[bench] splatmap: {ms:.3f} ms per 5x5 x 512^2 area, {mpx:.0f} Mpixel/s", file=sys.stderr, flush=True)
    return {"workload": "splatmap of a 5x5 rendered area of 512^2 chunks (6.6 Mpx), 3-octave noise, 4 biomes x "
                        "(2 altitude + 1 gradient) rules", "ms_per_area": round(ms, 3), "mpixel_s": round(mpx, 1)}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--plain-steps", type=int, default=2)
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-autotune", action="store_true", help="keep the default erosion launch shape")
    ap.add_argument("--no-served", action="store_true", help="leave out the camera-walk serving replays")
    ap.add_argument("--no-world", action="store_true", help="leave out the world-pipeline camera walk")
    args = ap.parse_args()
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU; multi-GPU runs are not part of it",
                          "n_gpus": args.gpus}))
        sys.exit(2)

    from superterrainplus_amd import capi
    capi.call("stp_engine_initialise", 0)
    cfg, sides, origins, props = workload()
    print(f"[bench] {len(sides)} chunks, {sum(s * s for s in sides) / 1e6:.1f} Mpixel", file=sys.stderr, flush=True)
    batch = make_batch(cfg, sides, origins, props)
    stream = C.c_void_p()
    capi.call("stp_stream_create", C.byref(stream))
    tuning = capi.BatchTuning()
    if args.no_autotune:
        capi.call("stp_batch_get_tuning", batch, C.byref(tuning))
    else:
        capi.call("stp_batch_autotune", batch, 2, stream, C.byref(tuning))
    print(f"[bench] erosion tile {tuning.ErosionTile}, {tuning.ErosionThreads} threads"
          f"{'' if args.no_autotune else ' (autotuned)'}", file=sys.stderr, flush=True)

    # timed region: one graph per stage, HIP events recorded on the stream between the graph
    # launches (so the erosion kernel's duration is measured in the same run as the throughput)
    capi.call("stp_batch_set_timing", batch, 1)
    for _ in range(args.warmup):
        capi.call("stp_batch_run", batch, 0, stream)
    capi.call("stp_stream_synchronize", stream)
    print("[bench] warm-up done", file=sys.stderr, flush=True)
    stage_ms = np.zeros(len(STAGES))
    t0 = time.perf_counter()
    for _ in range(args.steps):
        capi.call("stp_batch_run", batch, 0, stream)
        capi.call("stp_stream_synchronize", stream)
        ms = (C.c_float * len(STAGES))()
        capi.call("stp_batch_stage_times", batch, ms, len(STAGES))
        stage_ms += np.array(ms, np.float64)
    wall = time.perf_counter() - t0
    stage_ms /= args.steps

    # the same pipeline as one captured graph (no events), for reference
    capi.call("stp_batch_set_timing", batch, 0)
    capi.call("stp_batch_run", batch, 0, stream)
    capi.call("stp_stream_synchronize", stream)
    t1 = time.perf_counter()
    for _ in range(args.steps):
        capi.call("stp_batch_run", batch, 0, stream)
    capi.call("stp_stream_synchronize", stream)
    single_wall = (time.perf_counter() - t1) / args.steps
    counters = (C.c_uint64 * 6)()
    capi.call("stp_batch_counters", batch, counters, stream)
    pixels, drops, steps, erode, deposit, killed = [int(v) for v in counters]
    bins = C.c_uint64()
    capi.call("stp_batch_bin_total", batch, C.byref(bins), stream)
    bins = int(bins.value)
    ms_per_step = wall / args.steps * 1e3
    batched_mpx = pixels / (wall / args.steps) / 1e6
    print(f"[bench] batched {batched_mpx:.1f} Mpixel/s ({ms_per_step:.2f} ms/step)", file=sys.stderr, flush=True)

    # plain: one chunk per launch sequence + host sync per chunk (reported beside the batched
    # value; --plain-steps 0 leaves it out, e.g. for a kernel profile of the batched launches only)
    plain_mpx, lat = None, np.zeros(0)
    if args.plain_steps > 0:
        lat = []
        t0 = time.perf_counter()
        for _ in range(args.plain_steps):
            capi.call("stp_batch_run", batch, 2, stream)
            arr = (C.c_double * len(sides))()
            capi.call("stp_batch_chunk_latencies", batch, arr, len(sides))
            lat.extend(arr)
        plain_wall = (time.perf_counter() - t0) / args.plain_steps
        plain_mpx = pixels / plain_wall / 1e6
        lat = np.array(lat)
        print(f"[bench] plain {plain_mpx:.1f} Mpixel/s", file=sys.stderr, flush=True)

    erosion_bytes = steps * BYTES_STEP + erode * BYTES_ERODE + deposit * BYTES_DEPOSIT
    # the tiled erosion launches once per colour batch; the global-atomic kernel once
    kernel = f"erodeTileKERNEL<{tuning.ErosionThreads}>" if tuning.ErosionTile else "erosionKERNEL"
    launches = 4 if tuning.ErosionTile else 1
    erosion_s = stage_ms[STAGES.index("erosion")] / 1e3
    achieved = erosion_bytes / erosion_s / 1e9 if erosion_s > 0 else 0.0
    # every biome of the workload has cfg.fractal.octave octaves
    blend_flop = bins * (cfg.fractal.octave * FLOP_PER_OCTAVE + FLOP_PER_BIN)
    blend_s = stage_ms[STAGES.index("blend")] / 1e3
    blend_tflops = blend_flop / blend_s / 1e12 if blend_s > 0 else 0.0
    atomics = erode * ATOMICS_ERODE + deposit * ATOMICS_DEPOSIT
    result = {
        "metric": METRIC, "value": round(batched_mpx, 2), "unit": "Mpixel/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": round(ms_per_step, 3), "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f32", "data": "synthetic",
        "config": {
            "workload": "batched: 256 chunks, sides {128,256,512,1024} p={.4,.3,.2,.1} seed 11, 4-biome Voronoi "
                        "(seed 7, cell 256) blend r=16, 8-octave simplex, 0.25 drops/px parallel erosion, UINT16",
            "chunks": len(sides), "mpixel": round(pixels / 1e6, 3), "drops": drops,
            "plain_mpixel_s": round(plain_mpx, 2) if plain_mpx else None,
            "batched_vs_plain": round(batched_mpx / plain_mpx, 2) if plain_mpx else None,
            "plain_chunk_latency_ms": {"p50": round(float(np.percentile(lat, 50)), 3),
                                       "p99": round(float(np.percentile(lat, 99)), 3)} if len(lat) else None,
            "batched_step_latency_ms": round(ms_per_step, 3),
            "single_graph_mpixel_s": round(pixels / single_wall / 1e6, 2),
            "stage_ms": {k: round(float(v), 3) for k, v in zip(STAGES, stage_ms)},
            "erosion_tuning": {"tile": tuning.ErosionTile, "threads": tuning.ErosionThreads,
                               "autotuned": not args.no_autotune},
            "erosion_counters": {"steps": steps, "erode_steps": erode, "deposit_steps": deposit, "killed": killed},
            "erosion_atomics": {"per_step": atomics, "per_s": round(atomics / erosion_s, 1) if erosion_s > 0 else None,
                                "note": "LDS integer atomics (25 per erode step, 4 per deposit step) over the "
                                        "erosion stage's event time"},
            "blend_roofline": {"bound": "fp32 valu", "achieved": round(blend_tflops, 2), "peak": FP32_PEAK_TFLOPS,
                               "unit": "TFLOP/s", "frac": round(blend_tflops / FP32_PEAK_TFLOPS, 4),
                               "bins": bins, "flop_per_step": blend_flop,
                               "note": "algorithmic FLOP = bins x (octaves x 66 + 3); the peak counts packed "
                                       "FP32, which this build does not emit"},
        },
        "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": TILED_EROSION_TRAFFIC if (tuning.ErosionTile, tuning.ErosionThreads) == (128, 1024) else None,
                     "kernel": kernel, "launches_per_step": launches,
                     "launch_ms_avg": round(stage_ms[STAGES.index("erosion")] / launches, 4),
                     "algorithmic_bytes_per_step": erosion_bytes,
                     "note": "algorithmic bytes = the drop steps' height reads and updates (32 B per step, "
                             "+300 B per erode step, +32 B per deposit step); the LDS-tiled kernel serves "
                             "them from LDS, so they can exceed the HBM peak"},
    }
    if not args.no_served:
        result["config"]["served"] = {
            "workload": "camera random walk 0.5 chunk / 16 ms (seed 3), view radius 8, LOD 64..2048 px by distance, "
                        "Poisson arrivals, 16 ms deadline, whole path per chunk as the batched workload, "
                        "2 engines x 8 Mpx, 200 GB budget with LRU",
            "rates": served()}
    if not args.no_world:
        result["config"]["world"] = world()
        result["config"]["splatmap"] = splat()
        result["config"]["biome_layers"] = layers(tuning=tuning)
    if not args.no_cpu:
        threads = min(16, os.cpu_count() or 1)
        result["cpu_baseline"] = cpu_baseline(cfg, sides, origins, props, args.cpu_seconds, threads)
        result["config"]["cpu_baseline_fp64_1core"] = cpu_baseline_fp64(cfg, sides, origins, props,
                                                                         min(8.0, args.cpu_seconds / 2))
        print(f"[bench] cpu: FP32 {result['cpu_baseline']['value']:.2f} Mpixel/s on {threads} threads, FP64 "
              f"{result['config']['cpu_baseline_fp64_1core']['value']} Mpixel/s on one", file=sys.stderr, flush=True)
    capi.call("stp_batch_destroy", batch)
    print(json.dumps(result))


if __name__ == "__main__":
    main()
