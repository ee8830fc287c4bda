"""Shared helpers of the GPU parity tests: build inputs, run the HIP path through the C ABI
and the same steps through the oracle."""
import ctypes as C
import sys

import numpy as np

from superterrainplus_amd import capi, workloads as W


def log(msg):
    """Names the case before a GPU step (so a failing run says which case it was in)."""
    print(f"[case] {msg}", flush=True)
    sys.stdout.flush()


def fractal_info(side_x, side_y, offset, f=None):
    f = f or W.Fractal()
    info = capi.FractalInformation()
    info.Persistence, info.Lacunarity, info.Octave, info.Scale = f.persistence, f.lacunarity, f.octave, f.scale
    info.Offset[:] = list(offset)
    info.HalfDimension[:] = [side_x / 2.0, side_y / 2.0]
    info.Amplitude = info.Frequency = 1.0
    return info


def oracle_desc(oracle, side_x, side_y, offset, f=None):
    f = f or W.Fractal()
    d = oracle.FractalDesc()
    d.Persistence, d.Lacunarity, d.Octave, d.Scale = f.persistence, f.lacunarity, f.octave, f.scale
    d.OffsetX, d.OffsetY = offset
    d.HalfX, d.HalfY = side_x / 2.0, side_y / 2.0
    d.Amplitude = d.Frequency = 1.0
    return d


def oracle_rain(oracle, s):
    r = oracle.RainDropSetting()
    for f, _ in oracle.RainDropSetting._fields_:
        setattr(r, f, getattr(s, f))
    return r


def gpu_noise(table, side_x, side_y, offset, f=None):
    d = capi.DeviceArray((side_y, side_x), np.float32)
    capi.call("stp_simplex_heightmap", C.byref(table.view), C.byref(fractal_info(side_x, side_y, offset, f)),
              1.0, 0.0, d.vp, side_x, side_y, None)
    return d


def gpu_erode(d_height, map_size, nn, setting, mode, stream_id=0, ledger=True, info=True):
    H, W_ = d_height.shape
    assert (W_, H) == (map_size[0] * nn[0], map_size[1] * nn[1])
    idx, w = capi.brush(W_, setting.ErosionBrushRadius)
    d_idx, d_w = capi.DeviceArray.from_host(idx), capi.DeviceArray.from_host(w)
    d_led = capi.DeviceArray(setting.RainDropCount, np.float32) if ledger else None
    d_inf = capi.DeviceArray((setting.RainDropCount, 2), np.uint32) if info else None
    nninfo = capi.nn_info(map_size, nn)
    capi.call("stp_hydraulic_erosion", d_height.vp, C.byref(nninfo), C.byref(setting), d_idx.vp, d_w.vp, idx.size,
              stream_id, mode, d_led.vp if d_led else None, d_inf.vp if d_inf else None, None)
    capi.synchronize()
    return (d_led.download() if d_led else None), (d_inf.download() if d_inf else None)


def height_histogram(h, bins=256):
    return np.histogram(np.clip(h, 0, 1), bins=bins, range=(0.0, 1.0))[0]
