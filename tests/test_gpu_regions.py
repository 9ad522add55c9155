"""Region sizes below 32 blocks.  Regions of one parity are 2 R blocks
apart and rebuild concurrently, so a region's window must stay inside its
own blocks and those of the next region (2 R blocks): a 64-block window at
R = 16 or 24 would push into the region two ahead while that region's wave
rewrites it.  These cases fill q = 20 to 95% load (long clusters, pushes
across whole regions) in several batches, with plain and group inserts,
and compare every image with the oracle's.  Every test prints its case
before it touches the GPU."""
import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu


def _say(*a):
    print(*a, flush=True)


@pytest.mark.parametrize("region", [16, 24])
def test_small_regions_plain_insert_high_load(gpu, oracle, region):
    from cqf_amd.gqf import GQF, to_device_u64
    q = 20
    n = int(0.95 * (1 << q))
    keys = oracle.splitmix64(300 + region, n)
    _say(f"case plain insert q={q} load 0.95 region={region} keys={n} in 4 batches")
    f = GQF(q)
    o = oracle.OracleQF(q)
    try:
        f.set_tuning(region_blocks=region)
        for part in np.array_split(keys, 4):
            f.bulk_insert(to_device_u64(part))
        o.bulk_insert(keys.copy())
        torch.cuda.synchronize()
        assert np.array_equal(f.export_image(), o.image())
        assert f.status() == 0 and f.layout_errors() == 0
        # and a delete of a third, with the same tuning
        dele = keys[::3].copy()
        f.bulk_delete(to_device_u64(dele))
        o.bulk_delete(dele.copy())
        torch.cuda.synchronize()
        assert np.array_equal(f.export_image(), o.image())
        assert f.status() == 0 and f.layout_errors() == 0
    finally:
        f.close()
        o.close()


@pytest.mark.parametrize("region", [16, 24])
def test_small_regions_group_insert_high_load(gpu, oracle, region):
    from cqf_amd.gqf import GQF, GQFGroup, to_device_u64
    qs = [20, 18, 19]
    loads = [0.95, 0.93, 0.9]
    segs = [oracle.splitmix64(400 + i + region, int(ld * (1 << q))) for i, (q, ld) in enumerate(zip(qs, loads))]
    offsets = np.cumsum([0] + [s.size for s in segs]).tolist()
    _say(f"case group insert q={qs} loads={loads} region={region} keys={offsets[-1]}")
    filters = [GQF(q) for q in qs]
    group = GQFGroup(filters)
    try:
        for f in filters:
            f.set_tuning(region_blocks=region)
        group.bulk_insert(to_device_u64(np.concatenate(segs)), offsets)
        torch.cuda.synchronize()
        for i, q in enumerate(qs):
            o = oracle.OracleQF(q)
            o.bulk_insert(segs[i].copy())
            assert np.array_equal(filters[i].export_image(), o.image()), i
            assert filters[i].status() == 0 and filters[i].layout_errors() == 0
            o.close()
    finally:
        group.close()
        for f in filters:
            f.close()


def test_group_wait_bound_zero_sends_odd_regions_to_wide_pass(gpu, oracle):
    """With a wait bound of 0 every odd region's wait for its even
    neighbours gives up at once; such a region is not rebuilt in the
    persistent launch but by the wide pass after it, and every kernel exits
    normally.  The images still equal the oracle's, for graphs on and off,
    and a second call with the default bound agrees too."""
    from cqf_amd.gqf import GQF, GQFGroup, to_device_u64
    qs = [14, 16, 12, 15]
    segs = [oracle.splitmix64(500 + i, int(0.9 * (1 << q))) for i, q in enumerate(qs)]
    offsets = np.cumsum([0] + [s.size for s in segs]).tolist()
    allk = np.concatenate(segs)
    want = []
    for i, q in enumerate(qs):
        o = oracle.OracleQF(q)
        o.bulk_insert(segs[i].copy())
        want.append(o.image())
        o.close()
    for graphs in (True, False):
        _say(f"case group wait bound 0 graphs={graphs} q={qs} keys={allk.size}")
        filters = [GQF(q) for q in qs]
        group = GQFGroup(filters, graphs=graphs)
        try:
            group.set_wait_bound(0)
            group.bulk_insert(to_device_u64(allk), offsets)
            torch.cuda.synchronize()
            for i in range(len(qs)):
                assert np.array_equal(filters[i].export_image(), want[i]), i
                assert filters[i].status() == 0 and filters[i].layout_errors() == 0
            for f in filters:
                f.reset()
            group.set_wait_bound(1 << 22)
            group.bulk_insert(to_device_u64(allk), offsets)
            torch.cuda.synchronize()
            for i in range(len(qs)):
                assert np.array_equal(filters[i].export_image(), want[i]), i
        finally:
            group.close()
            for f in filters:
                f.close()
