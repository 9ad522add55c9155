"""Image export / import on the GPU: the device keeps 128-byte blocks with
exact offsets, and converts to / from the reference's 81-byte image
(include/gqf_int.cuh:38-57) whose offset byte saturates at 255.

The images come from the oracle (its canonical builder and its sequential
inserts), including clusters long enough to saturate offsets, so import has
to rebuild exact offsets that the image does not hold."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _say(*a):
    print(*a, flush=True)


def _long_cluster(oracle, nbits, q0, nq, per_q, count, seed):
    """Hashes of nq consecutive quotients from q0, per_q remainders each,
    all with `count`: one cluster of about nq * per_q * len(counter) slots."""
    rng = np.random.default_rng(seed)
    hs, cs = [], []
    for q in range(q0, q0 + nq):
        for r in sorted(rng.choice(256, size=per_q, replace=False)):
            hs.append((q << 8) | int(r))
            cs.append(count)
    return np.array(hs, dtype=np.uint64), np.array(cs, dtype=np.uint64)


def test_import_saturated_offsets_then_query_and_insert(gpu, oracle):
    from cqf_amd.gqf import GQF, to_device_u64, to_numpy_u64
    nbits = 14
    _say("case import q=14 canonical image with a saturated cluster")
    hs, cs = _long_cluster(oracle, nbits, 320, 64, 8, 10 ** 6, 1)
    img = oracle.canonical_image(nbits, hs, cs)
    assert (img[::81] == 255).sum() > 10  # offsets saturate across the cluster
    g = GQF(nbits)
    o = oracle.OracleQF(nbits)
    try:
        g.import_image(img)
        o.set_image(img)
        assert np.array_equal(g.export_image(), img)
        # bulk_get hashes its keys: query with the hashes' pre-images
        pre = np.array([oracle.hash_64i(int(h), (1 << (nbits + 8)) - 1) for h in hs], dtype=np.uint64)
        got = to_numpy_u64(g.bulk_get(to_device_u64(pre)))
        assert np.array_equal(got, cs)
        # inserts landing in and around the saturated cluster
        keys = np.concatenate([pre[::3], oracle.splitmix64(2, 3000)])
        g.bulk_insert(to_device_u64(keys))
        o.bulk_insert(keys.copy())
        gpu.cuda.synchronize()
        assert np.array_equal(g.export_image(), o.image())
        assert g.status() == 0 and g.layout_errors() == 0
    finally:
        g.close()
        o.close()


@pytest.mark.parametrize("nbits", [12, 16])
def test_export_import_round_trip_device(gpu, oracle, nbits):
    from cqf_amd.gqf import GQF, to_device_u64, to_numpy_u64
    _say(f"case export/import device q={nbits}")
    keys = oracle.splitmix64(7, int(0.9 * (1 << nbits)))
    g = GQF(nbits)
    h = GQF(nbits)
    try:
        g.bulk_insert(to_device_u64(keys))
        dimg = g.export_image_device()
        gpu.cuda.synchronize()
        host = g.export_image()
        assert np.array_equal(dimg.cpu().numpy(), host)
        h.import_image_device(dimg)
        assert np.array_equal(h.export_image(), host)
        probe = np.concatenate([keys[:2000], oracle.splitmix64(8, 2000)])
        a = to_numpy_u64(g.bulk_get(to_device_u64(probe)))
        b = to_numpy_u64(h.bulk_get(to_device_u64(probe)))
        assert np.array_equal(a, b)
        assert (a[:2000] >= 1).all()
    finally:
        g.close()
        h.close()


def test_import_rejects_inconsistent_offsets(gpu, oracle):
    from cqf_amd.gqf import GQF
    nbits = 12
    _say("case import rejects a wrong offset byte")
    keys = oracle.splitmix64(9, int(0.9 * (1 << nbits)))
    o = oracle.OracleQF(nbits)
    o.bulk_insert(keys.copy())
    img = o.image().copy()
    o.close()
    bad = img.copy()
    blk = int(np.nonzero(bad[::81][1:])[0][0]) + 1  # a block with a nonzero offset
    bad[blk * 81] = (int(bad[blk * 81]) + 1) % 255
    g = GQF(nbits)
    try:
        with pytest.raises(RuntimeError, match="rc=-4"):
            g.import_image(bad)
        g.import_image(img)  # the table is still usable
        assert np.array_equal(g.export_image(), img)
    finally:
        g.close()
