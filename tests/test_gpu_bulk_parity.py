"""GPU parity of the bulk API against the CPU oracle (bit-exact).

For the same keys, the GPU library (libgqf.so through its C ABI) and the
oracle's sequential restatement of the reference host path must leave the
same 81-byte block image and return the same counts.  Every test prints the
case before it touches the GPU.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _say(*a):
    print(*a, flush=True)


def _gpu_and_oracle(oracle, nbits):
    from cqf_amd.gqf import GQF
    return GQF(nbits), oracle.OracleQF(nbits)


def _assert_images_equal(img_gpu, img_orc, what):
    if not np.array_equal(img_gpu, img_orc):
        d = np.nonzero(img_gpu != img_orc)[0]
        blk = d[0] // 81
        raise AssertionError(f"{what}: {d.size} bytes differ, first at byte {d[0]} "
                             f"(block {blk}, field byte {d[0] % 81}): gpu {img_gpu[d[0]]} "
                             f"oracle {img_orc[d[0]]}")


@pytest.mark.parametrize("nbits,load,seed", [(10, 0.5, 1), (12, 0.9, 2), (14, 0.95, 3), (16, 0.9, 4),
                                             (17, 0.9, 5), (20, 0.9, 1)])
def test_bulk_insert_image_and_counts(gpu, oracle, nbits, load, seed):
    from cqf_amd.gqf import to_device_u64, to_numpy_u64
    n = int(load * (1 << nbits))
    _say(f"case bulk_insert q={nbits} n={n} seed={seed}")
    keys = oracle.splitmix64(seed, n)
    g, o = _gpu_and_oracle(oracle, nbits)
    try:
        dk = to_device_u64(keys)
        g.bulk_insert(dk)
        gpu.cuda.synchronize()
        hk = keys.copy()
        o.bulk_insert(hk)
        # keys are left as the ascending hashes (src/gqf.cu:3206-3209)
        assert np.array_equal(to_numpy_u64(dk), hk)
        _assert_images_equal(g.export_image(), o.image(), "image after bulk_insert")
        assert g.status() == 0 and g.layout_errors() == 0
        # counts of the inserted keys and of disjoint keys
        q = np.concatenate([keys, oracle.splitmix64(seed + 100, n)])
        got = to_numpy_u64(g.bulk_get(to_device_u64(q)))
        want = o.bulk_get(q)
        assert np.array_equal(got, want)
        assert (got[:n] >= 1).all()
    finally:
        g.close()
        o.close()


@pytest.mark.parametrize("nbits,batches", [(12, 4), (16, 8), (18, 5)])
def test_multi_batch_insert(gpu, oracle, nbits, batches):
    from cqf_amd.gqf import to_device_u64
    n = int(0.9 * (1 << nbits))
    _say(f"case multi-batch q={nbits} n={n} batches={batches}")
    keys = oracle.splitmix64(9, n)
    g, o = _gpu_and_oracle(oracle, nbits)
    try:
        for part in np.array_split(keys, batches):
            g.bulk_insert(to_device_u64(part))
            o.bulk_insert(part.copy())
        gpu.cuda.synchronize()
        _assert_images_equal(g.export_image(), o.image(), "image after batches")
    finally:
        g.close()
        o.close()


@pytest.mark.parametrize("nbits", [10, 13, 16])
def test_skewed_counts_reduce_and_plain(gpu, oracle, nbits):
    """Extended counters: a small key universe drawn with heavy repetition."""
    from cqf_amd.gqf import to_device_u64, to_numpy_u64
    rng = np.random.default_rng(nbits)
    univ = oracle.splitmix64(77, (1 << nbits) // 8)
    draws = univ[rng.zipf(1.3, size=(1 << nbits) * 2) % univ.size]
    _say(f"case skew q={nbits} draws={draws.size}")
    g1, o = _gpu_and_oracle(oracle, nbits)
    from cqf_amd.gqf import GQF
    g2 = GQF(nbits)
    try:
        for part in np.array_split(draws, 3):
            g1.bulk_insert(to_device_u64(part))
            g2.bulk_insert_reduce(to_device_u64(part))
            o.bulk_insert(part.copy())
        gpu.cuda.synchronize()
        img = o.image()
        _assert_images_equal(g1.export_image(), img, "skew bulk_insert")
        _assert_images_equal(g2.export_image(), img, "skew bulk_insert_reduce")
        got = to_numpy_u64(g2.bulk_get(to_device_u64(univ)))
        assert np.array_equal(got, o.bulk_get(univ))
    finally:
        g1.close()
        g2.close()
        o.close()


@pytest.mark.parametrize("nbits", [12, 16, 20])
def test_delete_half_then_all(gpu, oracle, nbits):
    from cqf_amd.gqf import to_device_u64
    n = int(0.95 * (1 << nbits))
    _say(f"case delete q={nbits} n={n}")
    keys = oracle.splitmix64(3, n)
    sel = oracle.splitmix64(7, n) & np.uint64(1)
    gone, keep = keys[sel == 0], keys[sel == 1]
    g, o = _gpu_and_oracle(oracle, nbits)
    try:
        g.bulk_insert(to_device_u64(keys))
        o.bulk_insert(keys.copy())
        pre = None
        g.bulk_delete(to_device_u64(gone))
        o.bulk_delete(gone.copy())
        gpu.cuda.synchronize()
        _assert_images_equal(g.export_image(), o.image(), "image after deleting half")
        g.bulk_delete(to_device_u64(keep))
        gpu.cuda.synchronize()
        img = g.export_image()
        assert not img.any(), "full delete must restore the all-zero image"
        del pre
    finally:
        g.close()
        o.close()


def test_delete_absent_and_repeated(gpu, oracle):
    from cqf_amd.gqf import to_device_u64
    nbits = 12
    keys = oracle.splitmix64(21, 1500)
    absent = oracle.splitmix64(22, 500)
    dels = np.concatenate([keys[:300], keys[:300], keys[:100], absent])
    _say("case delete absent/repeated")
    g, o = _gpu_and_oracle(oracle, nbits)
    try:
        ins = np.concatenate([keys, keys[:200]])
        g.bulk_insert(to_device_u64(ins))
        o.bulk_insert(ins.copy())
        g.bulk_delete(to_device_u64(dels))
        o.bulk_delete(dels.copy())
        gpu.cuda.synchronize()
        _assert_images_equal(g.export_image(), o.image(), "image after mixed deletes")
    finally:
        g.close()
        o.close()


def test_empty_and_single_batches(gpu, oracle):
    from cqf_amd.gqf import to_device_u64
    _say("case empty/single")
    g, o = _gpu_and_oracle(oracle, 12)
    try:
        empty = to_device_u64(np.zeros(0, dtype=np.uint64))
        g.bulk_insert(empty)
        g.bulk_delete(empty)
        g.bulk_insert_reduce(empty)
        one = np.array([12345], dtype=np.uint64)
        g.bulk_insert(to_device_u64(one))
        o.bulk_insert(one.copy())
        gpu.cuda.synchronize()
        _assert_images_equal(g.export_image(), o.image(), "single key")
        assert g.bulk_get_misses(to_device_u64(one)) == 0
    finally:
        g.close()
        o.close()


def test_misses_wrappers(gpu, oracle):
    from cqf_amd.gqf import to_device_u64
    nbits = 16
    n = int(0.9 * (1 << nbits))
    _say("case misses wrappers")
    keys = oracle.splitmix64(5, n)
    other = oracle.splitmix64(6, n)
    g, o = _gpu_and_oracle(oracle, nbits)
    try:
        g.bulk_insert(to_device_u64(keys))
        o.bulk_insert(keys.copy())
        assert g.bulk_get_misses(to_device_u64(keys)) == 0
        assert g.bulk_get_misses(to_device_u64(other)) == o.bulk_get_misses(other)
        assert g.point_get_misses(to_device_u64(other)) == o.bulk_get_misses(other)
    finally:
        g.close()
        o.close()


@pytest.mark.parametrize("region_blocks,digit", [(16, 6), (32, 8), (40, 10), (56, 0)])
def test_tuning_does_not_change_results(gpu, oracle, region_blocks, digit):
    """Region size and radix digit are speed knobs only: images and counts
    stay the oracle's."""
    from cqf_amd.gqf import to_device_u64, to_numpy_u64
    nbits = 16
    _say(f"case tuning R={region_blocks} digit={digit}")
    keys = oracle.splitmix64(31, int(0.93 * (1 << nbits)))
    g, o = _gpu_and_oracle(oracle, nbits)
    try:
        g.set_tuning(region_blocks, digit)
        assert g.info["region_blocks"] == region_blocks
        for part in np.array_split(keys, 3):
            g.bulk_insert(to_device_u64(part))
            o.bulk_insert(part.copy())
        _assert_images_equal(g.export_image(), o.image(), "tuned insert")
        half = keys[::2]
        g.bulk_delete(to_device_u64(half))
        o.bulk_delete(half.copy())
        _assert_images_equal(g.export_image(), o.image(), "tuned delete")
        got = to_numpy_u64(g.bulk_get(to_device_u64(keys)))
        assert np.array_equal(got, o.bulk_get(keys))
        assert g.status() == 0 and g.layout_errors() == 0
    finally:
        g.close()
        o.close()


@pytest.mark.parametrize("nbits", [16, 22])
def test_wide_window_fallback_at_high_load(gpu, oracle, nbits):
    """56-block regions leave an 8-block push window, so at 95% load many
    regions outrun it and are redone by the wide pass: from its list of
    flagged regions at 2^16 slots (19 regions), and at 2^22 (1,171
    regions) by the class launch, whose classes redo the flagged regions
    with disjoint widest windows at once, one class after another.  Inserts and deletes must still
    give the oracle's image."""
    from cqf_amd.gqf import to_device_u64, to_numpy_u64
    _say(f"case wide fallback q={nbits} R=56 load 0.95")
    keys = oracle.splitmix64(77, int(0.95 * (1 << nbits)))
    g, o = _gpu_and_oracle(oracle, nbits)
    try:
        g.set_tuning(56, 0)
        for part in np.array_split(keys, 4):
            g.bulk_insert(to_device_u64(part))
            o.bulk_insert(part.copy())
            _assert_images_equal(g.export_image(), o.image(), "insert batch")
        half = keys[1::2]
        g.bulk_delete(to_device_u64(half))
        o.bulk_delete(half.copy())
        _assert_images_equal(g.export_image(), o.image(), "delete half")
        got = to_numpy_u64(g.bulk_get(to_device_u64(keys)))
        assert np.array_equal(got, o.bulk_get(keys))
        assert g.status() == 0 and g.layout_errors() == 0
    finally:
        g.close()
        o.close()


def test_wide_class_wait_bound_zero_reports_overflow(gpu, oracle):
    """The wide pass's classes wait for each other with a bounded poll.  At
    bound 0 every class after the first gives up at once: its flagged
    regions stay as they were, the filter reports CLUSTER_OVERFLOW, every
    kernel exits, and the table's layout stays valid (the first class and
    the phases still rebuilt theirs).  The default bound is restored
    afterwards, and the same batches then give the oracle's image."""
    from cqf_amd import gqf as G
    from cqf_amd.gqf import to_device_u64
    nbits = 22
    _say(f"case wide class wait bound 0 q={nbits} R=56 load 0.95")
    keys = oracle.splitmix64(77, int(0.95 * (1 << nbits)))
    g = G.GQF(nbits)
    prev = G.set_wide_wait_bound(0)
    try:
        g.set_tuning(56, 0)
        for part in np.array_split(keys, 4):
            g.bulk_insert(to_device_u64(part))
        assert g.status() & G.QF_STATUS_CLUSTER_OVERFLOW
        assert g.layout_errors() == 0
    finally:
        G.set_wide_wait_bound(prev)
        g.close()
    assert G.set_wide_wait_bound(prev) == prev  # restored
    g, o = _gpu_and_oracle(oracle, nbits)
    try:
        g.set_tuning(56, 0)
        for part in np.array_split(keys, 4):
            g.bulk_insert(to_device_u64(part))
            o.bulk_insert(part.copy())
        _assert_images_equal(g.export_image(), o.image(), "insert after the bound is restored")
        assert g.status() == 0 and g.layout_errors() == 0
    finally:
        g.close()
        o.close()


@pytest.mark.parametrize("nbits", [20, 22])
def test_sparse_batches_after_dense_fill(gpu, oracle, nbits):
    """Batches of at most one key per point-lock region take the point path
    (one lane per key under the region locks); the image must equal the
    oracle's after each, including keys repeated within a batch and keys
    already present, through bulk_insert and bulk_insert_reduce."""
    from cqf_amd.gqf import to_device_u64
    from cqf_amd.gqf import GQF
    _say(f"case sparse batches q={nbits}")
    keys = oracle.splitmix64(31, int(0.85 * (1 << nbits)))
    g, o = _gpu_and_oracle(oracle, nbits)
    try:
        g.bulk_insert(to_device_u64(keys))
        o.bulk_insert(keys.copy())
        limit = min(g.info["nregions"] // 2, g.info["num_locks"])
        extra = oracle.splitmix64(32, 40 * limit)
        for b in range(20):
            part = extra[b * limit:(b + 1) * limit - 3]
            part = np.concatenate([part, part[:2], keys[b:b + 1]])[:limit]  # repeats and a present key
            if b % 2:
                g.bulk_insert(to_device_u64(part))
            else:
                g.bulk_insert_reduce(to_device_u64(part))
            o.bulk_insert(part.copy())
        gpu.cuda.synchronize()
        _assert_images_equal(g.export_image(), o.image(), "image after sparse batches")
        assert g.status() == 0 and g.layout_errors() == 0
    finally:
        g.close()
        o.close()


@pytest.mark.parametrize("nbits,sizes", [(22, [1 << 21]), (23, [int(0.9 * (1 << 23))]),
                                         (22, [1 << 20, 1 << 21])])
def test_large_batches_two_level_sort(gpu, oracle, nbits, sizes):
    """Batches of 2^20..2^24 keys take the two-level sort (top-bit buckets,
    then one LDS sort per bucket): keys must come back as the ascending
    hashes and the image must equal the oracle's."""
    from cqf_amd.gqf import GQF, to_device_u64, to_numpy_u64
    _say(f"case two-level sort q={nbits} batches={sizes}")
    g, o = GQF(nbits), oracle.OracleQF(nbits)
    try:
        off = 0
        for n in sizes:
            keys = oracle.splitmix64(41, n, skip=off)
            off += n
            dk = to_device_u64(keys)
            g.bulk_insert(dk)
            hk = keys.copy()
            o.bulk_insert(hk)
            assert np.array_equal(to_numpy_u64(dk), hk)
        _assert_images_equal(g.export_image(), o.image(), "image after two-level sorted batches")
        assert g.status() == 0 and g.layout_errors() == 0
        q = oracle.splitmix64(41, 1 << 16)
        assert np.array_equal(to_numpy_u64(g.bulk_get(to_device_u64(q))), o.bulk_get(q))
    finally:
        g.close()
        o.close()


def test_two_level_sort_repeated_keys(gpu, oracle):
    """2^20 copies of 16 keys: each key's bucket holds far more keys than the
    LDS sort takes, so those buckets are sorted in global memory."""
    from cqf_amd.gqf import GQF, to_device_u64, to_numpy_u64
    nbits = 20
    _say("case two-level sort, 16 keys x 2^16 copies")
    univ = oracle.splitmix64(43, 16)
    keys = np.repeat(univ, 1 << 16)
    keys = keys[oracle.splitmix64(44, keys.size) % np.uint64(keys.size)]  # shuffled with repeats
    g, o = GQF(nbits), oracle.OracleQF(nbits)
    try:
        dk = to_device_u64(keys)
        g.bulk_insert(dk)
        hk = keys.copy()
        o.bulk_insert(hk)
        assert np.array_equal(to_numpy_u64(dk), hk)
        _assert_images_equal(g.export_image(), o.image(), "image after repeated keys")
        assert np.array_equal(to_numpy_u64(g.bulk_get(to_device_u64(univ))), o.bulk_get(univ))
        g.bulk_delete(to_device_u64(keys))
        gpu.cuda.synchronize()
        assert not g.export_image().any()
    finally:
        g.close()
        o.close()


def test_two_level_sort_moderate_repeats(gpu, oracle):
    """2^20 draws over 2^16 keys: buckets fit in LDS, but some of their
    sub-buckets hold more than 64 keys, so those buckets take the stable
    8-bit passes instead of the per-wave bitonic sorts."""
    from cqf_amd.gqf import GQF, to_device_u64, to_numpy_u64
    nbits = 21
    _say("case two-level sort, 2^20 draws over 2^16 keys")
    univ = oracle.splitmix64(45, 1 << 16)
    keys = univ[oracle.splitmix64(46, (1 << 20) + 777) % np.uint64(univ.size)]
    g, o = GQF(nbits), oracle.OracleQF(nbits)
    try:
        dk = to_device_u64(keys)
        g.bulk_insert(dk)
        hk = keys.copy()
        o.bulk_insert(hk)
        assert np.array_equal(to_numpy_u64(dk), hk)
        _assert_images_equal(g.export_image(), o.image(), "image after moderate repeats")
        assert np.array_equal(to_numpy_u64(g.bulk_get(to_device_u64(univ))), o.bulk_get(univ))
    finally:
        g.close()
        o.close()
