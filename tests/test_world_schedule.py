"""The world pipeline's erosion schedule (host logic, no GPU): centres in one batch have disjoint
neighbourhoods, as the reference's scheduleChunk guarantees
(SuperTerrain+/SuperTerrain+/Private/World/STPWorldPipeline.cpp:339-404), and every centre takes
the first batch free of conflicts (first fit), checked by brute force."""
import ctypes as C

import numpy as np
import pytest

from superterrainplus_amd import capi


def schedule(centres, nn):
    flat = (C.c_int32 * (2 * len(centres)))(*[v for c in centres for v in c])
    out = (C.c_uint32 * max(len(centres), 1))()
    capi.call("stp_world_schedule", flat, len(centres), nn, out)
    return list(out)[:len(centres)]


def conflict(a, b, nn):
    return abs(a[0] - b[0]) < nn and abs(a[1] - b[1]) < nn


@pytest.mark.parametrize("nn", [1, 3, 5])
@pytest.mark.parametrize("seed", [0, 1, 2])
def test_schedule_is_conflict_free_first_fit(nn, seed):
    rng = np.random.default_rng(seed)
    centres = [tuple(int(v) for v in rng.integers(-6, 7, 2)) for _ in range(40)]
    centres = list(dict.fromkeys(centres))  # the world dedupes its centres
    batch = schedule(centres, nn)
    for i, a in enumerate(centres):
        for j in range(i):
            if batch[i] == batch[j]:
                assert not conflict(a, centres[j], nn), (a, centres[j])
        # first fit: every smaller batch holds an earlier centre that conflicts with this one
        for b in range(batch[i]):
            assert any(batch[j] == b and conflict(a, centres[j], nn) for j in range(i)), (i, b)


def test_schedule_examples():
    # a 3 x 3 neighbourhood: adjacent centres conflict, centres 3 apart share a batch
    assert schedule([(0, 0), (1, 0), (3, 0), (0, 3), (6, 6)], 3) == [0, 1, 0, 0, 0]
    assert schedule([(0, 0), (2, 2), (-2, 2), (4, 4)], 3) == [0, 1, 1, 0]
    # nn = 1: no two distinct chunks conflict
    assert schedule([(0, 0), (1, 0), (0, 1)], 1) == [0, 0, 0]
    assert schedule([], 3) == []


def test_schedule_rejects_even_neighbourhood():
    with pytest.raises(capi.STPError):
        schedule([(0, 0)], 2)


def test_stream_word_is_a_pure_function_of_the_chunk():
    def word(x, y):
        w = C.c_uint32()
        capi.call("stp_world_stream_word", x, y, C.byref(w))
        return w.value
    words = {(x, y): word(x, y) for x in range(-4, 5) for y in range(-4, 5)}
    assert len(set(words.values())) == len(words)
    assert word(3, -2) == words[(3, -2)]
