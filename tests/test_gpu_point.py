"""GPU parity of the device point API (include/gqf_device.h) against the CPU
oracle: point_insert, point_insert_not_exists, point_query and
point_query_concurrent run one lane per key under the region locks, and the
images they leave are compared byte for byte.  Every test prints its case
before it touches the GPU."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

QF_ITEM_INSERTED, QF_ITEM_FOUND, QF_FULL = 0, 1, 2
KEY_IS_HASH = 0x08


def _say(*a):
    print(*a, flush=True)


def _ms(hashes):
    ms = {}
    for h in hashes:
        ms[int(h)] = ms.get(int(h), 0) + 1
    return ms


@pytest.mark.parametrize("nbits,load,seed", [(12, 0.9, 1), (16, 0.9, 2), (16, 0.95, 3)])
def test_point_bulk_insert_matches_bulk_image(gpu, oracle, nbits, load, seed):
    from cqf_amd.gqf import GQF, to_device_u64
    n = int(load * (1 << nbits))
    _say(f"case point_bulk_insert q={nbits} n={n}")
    keys = oracle.splitmix64(seed, n)
    g = GQF(nbits, bulk_config=False)
    o = oracle.OracleQF(nbits)
    try:
        dk = to_device_u64(keys)
        g.point_bulk_insert(dk)
        o.bulk_insert(keys.copy())
        assert np.array_equal(g.export_image(), o.image())
        assert g.status() == 0 and g.layout_errors() == 0
        other = oracle.splitmix64(seed + 50, n)
        assert g.point_get_misses(to_device_u64(keys)) == 0
        assert g.point_get_misses(to_device_u64(other)) == o.bulk_get_misses(other)
    finally:
        g.close()
        o.close()


def test_point_insert_counts_concurrent(gpu, oracle):
    """Many lanes increment the same counters at once: the result is the
    canonical image of the multiset and every return is QF_ITEM_INSERTED."""
    from cqf_amd.gqf import GQF, to_device_u64, to_numpy_u64
    nbits = 14
    _say(f"case point_insert duplicates q={nbits}")
    rng = np.random.default_rng(3)
    univ = oracle.splitmix64(4, 3000)
    draws = univ[rng.zipf(1.3, size=12000) % univ.size]
    g = GQF(nbits)
    o = oracle.OracleQF(nbits)
    try:
        rets = g.point_insert(to_device_u64(draws)).cpu().numpy()
        assert (rets == QF_ITEM_INSERTED).all()
        o.bulk_insert(draws.copy())
        assert np.array_equal(g.export_image(), o.image())
        probe = np.concatenate([univ, oracle.splitmix64(5, 3000)])
        want = o.bulk_get(probe)
        for conc in (False, True):
            got = to_numpy_u64(g.point_query(to_device_u64(probe), concurrent=conc))
            assert np.array_equal(got, want), conc
    finally:
        g.close()
        o.close()


def test_point_insert_not_exists_matches_oracle(gpu, oracle):
    """After counted bulk inserts, one point_insert_not_exists per quotient
    (distinct quotients, so the order of lanes does not matter) leaves the
    oracle's image and return codes, including the slot-wise search over
    multi-slot counters."""
    from cqf_amd.gqf import GQF, to_device_u64
    nbits = 12
    mask = (1 << (nbits + 8)) - 1
    _say(f"case point_insert_not_exists q={nbits}")
    rng = np.random.default_rng(5)
    base = oracle.splitmix64(6, 1500)
    reps = rng.choice([1, 1, 2, 3, 5, 40, 300], size=base.size)
    counted = np.repeat(base, reps)
    hb = oracle.hash_many(base, mask)
    # one probe per distinct quotient: either an inserted hash or the same
    # quotient with another remainder
    _, first = np.unique(hb >> np.uint64(8), return_index=True)
    probes_h = hb[first].copy()
    flip = rng.random(probes_h.size) < 0.6
    probes_h[flip] = (probes_h[flip] & ~np.uint64(0xFF)) | rng.integers(0, 256, flip.sum()).astype(np.uint64)
    g = GQF(nbits)
    o = oracle.OracleQF(nbits)
    try:
        g.bulk_insert(to_device_u64(counted))
        o.bulk_insert(counted.copy())
        assert np.array_equal(g.export_image(), o.image())
        rets = g.point_insert_not_exists(to_device_u64(probes_h), flags=KEY_IS_HASH).cpu().numpy()
        want = np.array([o.insert_not_exists(int(h), KEY_IS_HASH)[0] for h in probes_h])
        assert np.array_equal(rets, want)
        assert np.array_equal(g.export_image(), o.image())
    finally:
        g.close()
        o.close()


def test_point_insert_full_table(gpu, oracle):
    """Inserting four times the capacity: lanes past the last empty slot get
    QF_FULL, the status records it, and the image is exactly the canonical
    image of the keys that were accepted."""
    from cqf_amd.gqf import GQF, to_device_u64
    nbits = 8
    _say(f"case point_insert full q={nbits}")
    keys = oracle.splitmix64(8, 4 << nbits)
    g = GQF(nbits)
    o = oracle.OracleQF(nbits)
    try:
        rets = g.point_insert(to_device_u64(keys)).cpu().numpy()
        assert (rets == QF_FULL).any() and (rets == QF_ITEM_INSERTED).any()
        assert set(np.unique(rets)) <= {QF_ITEM_INSERTED, QF_FULL}
        assert g.status() & 0x8
        o.bulk_insert(keys[rets == QF_ITEM_INSERTED].copy())
        assert np.array_equal(g.export_image(), o.image())
    finally:
        g.close()
        o.close()
