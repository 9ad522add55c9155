"""Runs longer than 255 slots.

The fast rebuild window keeps each listed quotient's old and new run
length in one byte; a region whose batch makes or touches a longer run is
redone in the wider windows, which keep 16 bits.  One quotient carrying all
256 remainders (a 256-slot run), then more counts on it and on its
neighbours, then deletes: every step must leave the canonical image that
the spec model builds from the multiset (runs at max(q, previous end + 1))
and brute-force counts.  Every test prints its case first."""
import numpy as np
import pytest
import torch

from tests import spec_model as S

pytestmark = pytest.mark.gpu

KEY_IS_HASH = 0x08


def _say(*a):
    print(*a, flush=True)


def _check(f, oracle, q, ms):
    from cqf_amd.gqf import to_device_u64, to_numpy_u64
    torch.cuda.synchronize()
    assert f.export_image().tobytes() == S.build_image(q, ms)
    hs = np.array(sorted(ms), dtype=np.uint64)
    mask = (1 << (q + 8)) - 1
    keys = np.array([oracle.hash_64i(int(h), mask) for h in hs], dtype=np.uint64)  # hash_64(key) = h
    got = to_numpy_u64(f.bulk_get(to_device_u64(keys)))
    assert np.array_equal(got, np.array([ms[int(h)] for h in hs], dtype=np.uint64))
    assert f.status() == 0 and f.layout_errors() == 0


@pytest.mark.parametrize("home", [5, 700, 1100])
def test_runs_past_255_slots(gpu, oracle, home):
    from cqf_amd.gqf import GQF, to_device_u64
    q = 11
    _say(f"case long runs q={q}: quotient {home} with 256 remainders, counts on it and its neighbours, deletes")
    f = GQF(q)
    ms = {}

    def apply(hs, d):
        for h in hs:
            h = int(h)
            ms[h] = ms.get(h, 0) + d
            if ms[h] == 0:
                del ms[h]

    try:
        # a 256-slot run made from empty
        first = np.array([(home << 8) | r for r in range(256)], dtype=np.uint64)
        f.bulk_insert(to_device_u64(first), flags=KEY_IS_HASH)
        apply(first, 1)
        _check(f, oracle, q, ms)
        # a second count on every fourth remainder (the run grows to 320
        # slots) and short runs on the next quotients, which the long run pushes
        more = np.array([(home << 8) | r for r in range(0, 256, 4)] +
                        [((home + k) << 8) | r for k in range(1, 6) for r in (2, 77, 200)], dtype=np.uint64)
        f.bulk_insert(to_device_u64(more), flags=KEY_IS_HASH)
        apply(more, 1)
        _check(f, oracle, q, ms)
        # deletes on the long run and after it
        gone = np.array([(home << 8) | r for r in range(1, 256, 3)] + [((home + 2) << 8) | 77], dtype=np.uint64)
        f.bulk_delete(to_device_u64(gone), flags=KEY_IS_HASH)
        apply(gone, -1)
        _check(f, oracle, q, ms)
        # everything out: the zero image
        rest = np.array([h for h, c in sorted(ms.items()) for _ in range(c)], dtype=np.uint64)
        f.bulk_delete(to_device_u64(rest), flags=KEY_IS_HASH)
        torch.cuda.synchronize()
        assert not f.export_image().any()
        assert f.status() == 0 and f.layout_errors() == 0
    finally:
        f.close()
