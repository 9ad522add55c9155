"""Filter groups: many filters filled by one tagged sort and ONE persistent
rebuild launch (captured into a hipGraph), queried by one launch.  Each
filter must end exactly as if its segment had gone through bulk_insert
alone: the oracle's image, counts and hashed keys.  Every test prints its
case before it touches the GPU."""
import time

import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu

KEY_IS_HASH = 0x08


def _say(*a):
    print(*a, flush=True)


def _setup(nfilters, q_lo, q_hi, seed=5):
    from cqf_amd import workloads as W
    from cqf_amd.gqf import GQF
    params, keys, offsets = W.kmer_workload(nfilters, seed=seed, q_lo=q_lo, q_hi=q_hi)
    filters = [GQF(q) for q, *_ in params]
    return params, keys, offsets, filters


def _close(filters, group=None):
    if group is not None:
        group.close()
    for f in filters:
        f.close()


def test_group_insert_matches_oracle(gpu, oracle):
    from cqf_amd.gqf import GQFGroup, to_numpy_u64
    params, keys, offsets, filters = _setup(12, 10, 14)
    _say(f"case group insert 12 filters q10-14 keys={keys.numel()}")
    group = GQFGroup(filters)
    try:
        dk = keys.clone().cuda()
        group.bulk_insert(dk, offsets)
        got_keys = to_numpy_u64(dk)
        rets = to_numpy_u64(group.bulk_get(keys.cuda(), offsets))
        allk = keys.numpy().view(np.uint64)
        for f, (q, load, n, _) in enumerate(params):
            seg = allk[offsets[f]:offsets[f + 1]].copy()
            o = oracle.OracleQF(q)
            o.bulk_insert(seg)  # leaves the ascending hashes in seg
            assert np.array_equal(filters[f].export_image(), o.image()), f
            assert np.array_equal(got_keys[offsets[f]:offsets[f + 1]], seg), f
            assert np.array_equal(rets[offsets[f]:offsets[f + 1]], o.bulk_get(allk[offsets[f]:offsets[f + 1]])), f
            assert filters[f].status() == 0 and filters[f].layout_errors() == 0
            o.close()
    finally:
        _close(filters, group)


def test_group_graph_replays_and_eager_agree(gpu, oracle):
    """The captured graph replayed several times (eager work in between)
    and the eager path leave the same images."""
    from cqf_amd.gqf import GQF, GQFGroup
    params, keys, offsets, filters = _setup(6, 11, 13, seed=7)
    _say(f"case group graph replay 6 filters keys={keys.numel()}")
    group = GQFGroup(filters, graphs=True)
    other = GQF(12)
    try:
        dk = torch.empty_like(keys, device="cuda")
        ref = None
        for rep in range(3):
            for f in filters:
                f.reset()
            dk.copy_(keys.cuda())
            group.bulk_insert(dk, offsets)
            other.bulk_insert(keys[:1000].clone().cuda())  # eager work between replays
            torch.cuda.synchronize()
            imgs = [f.export_image() for f in filters]
            if ref is None:
                ref = imgs
            for a, b in zip(imgs, ref):
                assert np.array_equal(a, b), rep
        eager = GQFGroup(filters, graphs=False)
        for f in filters:
            f.reset()
        dk.copy_(keys.cuda())
        eager.bulk_insert(dk, offsets)
        torch.cuda.synchronize()
        for f, r in zip(filters, ref):
            assert np.array_equal(f.export_image(), r)
        eager.close()
        # and the oracle agrees with the first filter
        q = params[0][0]
        o = oracle.OracleQF(q)
        o.bulk_insert(keys[offsets[0]:offsets[1]].numpy().view(np.uint64).copy())
        assert np.array_equal(ref[0], o.image())
        o.close()
    finally:
        other.close()
        _close(filters, group)


def test_group_prehashed_out_of_range_flagged(gpu, oracle):
    from cqf_amd.gqf import GQF, GQFGroup
    _say("case group pre-hashed with out-of-range hashes")
    qs = [10, 12, 11]
    filters = [GQF(q) for q in qs]
    group = GQFGroup(filters)
    try:
        segs = []
        for i, q in enumerate(qs):
            h = oracle.hash_many(oracle.splitmix64(40 + i, 800), (1 << (q + 8)) - 1)
            if i == 1:
                h[::50] = (1 << (q + 8)) + np.arange(h[::50].size, dtype=np.uint64)
            segs.append(h)
        offsets = [0]
        for s_ in segs:
            offsets.append(offsets[-1] + s_.size)
        dk = torch.from_numpy(np.concatenate(segs).view(np.int64)).cuda()
        group.bulk_insert(dk, offsets, flags=KEY_IS_HASH)
        torch.cuda.synchronize()
        for i, q in enumerate(qs):
            o = oracle.OracleQF(q)
            ok = segs[i][segs[i] < (1 << (q + 8))]
            o.bulk_insert(ok.copy(), KEY_IS_HASH)
            assert np.array_equal(filters[i].export_image(), o.image()), i
            assert filters[i].status() == (0x4 if i == 1 else 0)
            o.close()
    finally:
        _close(filters, group)


def test_group_exact_counts_larger(gpu):
    """32 filters of 2^16..2^20 slots: every inserted k-mer reads back the
    exact number of instances sharing its masked key (size-independent)."""
    from cqf_amd import workloads as W
    from cqf_amd.gqf import GQF, GQFGroup
    params, keys, offsets = W.kmer_workload(32, seed=9, q_lo=16, q_hi=20, device="cuda")
    _say(f"case group exact counts 32 filters keys={keys.numel()}")
    filters = [GQF(q) for q, *_ in params]
    group = GQFGroup(filters)
    try:
        group.bulk_insert(keys.clone(), offsets)
        got = group.bulk_get(keys, offsets)
        for f, (q, *_r) in enumerate(params):
            seg = keys[offsets[f]:offsets[f + 1]]
            masked = seg & ((1 << (q + 8)) - 1)
            _, inv, cnt = torch.unique(masked, return_inverse=True, return_counts=True)
            assert torch.equal(got[offsets[f]:offsets[f + 1]], cnt[inv]), f
            assert filters[f].status() == 0 and filters[f].layout_errors() == 0
    finally:
        _close(filters, group)


def test_group_get_empty_and_short_segments(gpu):
    """Group queries whose segments are empty (first, middle and last
    filter), one key long, or end inside a workgroup's chunk of keys: every
    key's count equals its own filter's bulk_get."""
    from cqf_amd.gqf import GQF, GQFGroup
    qs = [10, 12, 11, 13, 10, 12]
    sizes = [0, 700, 0, 1, 5000, 0]
    _say(f"case group get segments {sizes} over q {qs}")
    gen = torch.Generator().manual_seed(11)
    filters = [GQF(q) for q in qs]
    group = GQFGroup(filters)
    try:
        pools = []
        for f, q in zip(filters, qs):
            ins = torch.randint(0, 2**62, (1 << (q - 1),), generator=gen, dtype=torch.int64)
            f.bulk_insert(ins.clone().cuda())
            pools.append(ins)
        segs = []
        for f, n in enumerate(sizes):
            # half inserted keys (some twice), half fresh ones
            pick = pools[f][torch.randint(0, pools[f].numel(), (n // 2,), generator=gen)]
            fresh = torch.randint(0, 2**62, (n - n // 2,), generator=gen, dtype=torch.int64)
            segs.append(torch.cat([pick, fresh]))
        offsets = [0]
        for sg in segs:
            offsets.append(offsets[-1] + sg.numel())
        keys = torch.cat(segs).cuda()
        got = group.bulk_get(keys, offsets)
        for f, sg in enumerate(segs):
            if sg.numel():
                want = filters[f].bulk_get(sg.cuda())
                assert torch.equal(got[offsets[f]:offsets[f + 1]], want), f
    finally:
        _close(filters, group)


def test_autotune_choices_keep_results(gpu, oracle):
    from cqf_amd import gqf
    _say("case autotune q=14 load 0.9")
    r, d = gqf.autotune(14, 0.9)
    assert 16 <= r <= 56 and d in (0, 6, 7, 8, 9, 10)
    # a choice recorded with gqf_autotune_set (as the committed table is
    # loaded) is returned without timing anything
    assert gqf.load_library().gqf_autotune_set(15, 0.55, 40, 8) == 0
    t0 = time.perf_counter()
    assert gqf.autotune(15, 0.55) == (40, 8)
    assert time.perf_counter() - t0 < 0.05
    params, keys, offsets, filters = _setup(4, 12, 14, seed=12)
    group = gqf.GQFGroup(filters)
    try:
        group.autotune(offsets)
        assert 1 <= group.workgroups_per_cu <= 16
        group.bulk_insert(keys.clone().cuda(), offsets)
        torch.cuda.synchronize()
        allk = keys.numpy().view(np.uint64)
        for f, (q, *_r) in enumerate(params):
            o = oracle.OracleQF(q)
            o.bulk_insert(allk[offsets[f]:offsets[f + 1]].copy())
            assert np.array_equal(filters[f].export_image(), o.image()), f
            o.close()
    finally:
        _close(filters, group)


def test_group_workgroups_per_cu_keep_results(gpu, oracle):
    """The persistent rebuild's grid (workgroups per CU) changes which
    workgroup takes which region, never the result: 1 per CU (each
    workgroup loops over many regions), 3, and the most the GPU keeps
    resident (found by asking for more), replayed from graphs, against the
    oracle."""
    from cqf_amd.gqf import GQFGroup
    params, keys, offsets, filters = _setup(10, 12, 15, seed=21)
    _say(f"case group workgroups per CU 10 filters keys={keys.numel()}")
    group = GQFGroup(filters)
    try:
        assert group.workgroups_per_cu == 8  # the default, when the GPU keeps 8 resident
        with pytest.raises(ValueError):
            group.workgroups_per_cu = 1000
        allk = keys.numpy().view(np.uint64)
        want = []
        for f, (q, *_r) in enumerate(params):
            o = oracle.OracleQF(q)
            o.bulk_insert(allk[offsets[f]:offsets[f + 1]].copy())
            want.append(o.image())
            o.close()
        dk = torch.empty_like(keys, device="cuda")
        most = 0  # the most workgroups per CU the GPU keeps resident (the setter refuses more)
        for pc in range(16, 0, -1):
            try:
                group.workgroups_per_cu = pc
                most = pc
                break
            except ValueError:
                pass
        assert most >= 8
        for pc in (1, 3, most):
            group.workgroups_per_cu = pc
            assert group.workgroups_per_cu <= pc
            for f in filters:
                f.reset()
            dk.copy_(keys.cuda())
            group.bulk_insert(dk, offsets)
            torch.cuda.synchronize()
            for f in range(len(filters)):
                assert np.array_equal(filters[f].export_image(), want[f]), (pc, f)
                assert filters[f].status() == 0 and filters[f].layout_errors() == 0
    finally:
        _close(filters, group)


@pytest.mark.parametrize("big", [1 << 24, (1 << 24) + 4099])
def test_group_segment_sort_limits(gpu, oracle, big):
    """Segments of at most 2^24 keys are sorted one by one with the
    two-level sort, which ORs each filter's tag into its keys; a larger
    segment sends the whole
    group to the tagged LSD sort.  Both give every filter the oracle's
    image and leave each segment holding its ascending hashes."""
    from cqf_amd.gqf import GQF, GQFGroup, to_device_u64, to_numpy_u64
    sizes = [(25, big), (12, 3000), (16, 40000)]
    _say(f"case group segment sizes {[n for _, n in sizes]}")
    segs = [oracle.splitmix64(90 + i, n) for i, (_, n) in enumerate(sizes)]
    offsets = np.cumsum([0] + [n for _, n in sizes]).tolist()
    filters = [GQF(q) for q, _ in sizes]
    group = GQFGroup(filters)
    try:
        dk = to_device_u64(np.concatenate(segs))
        group.bulk_insert(dk, offsets)
        got = to_numpy_u64(dk)
        for f, (q, n) in enumerate(sizes):
            seg = segs[f].copy()
            o = oracle.OracleQF(q)
            o.bulk_insert(seg)
            assert np.array_equal(filters[f].export_image(), o.image()), f
            assert np.array_equal(got[offsets[f]:offsets[f + 1]], seg), f
            assert filters[f].status() == 0 and filters[f].layout_errors() == 0
            o.close()
    finally:
        _close(filters, group)
