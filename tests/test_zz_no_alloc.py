"""bulk_insert_noalloc -- the entry point that allocates nothing -- against
bulk_insert, through the C ABI.  Two calls in a row on each filter, the
second smaller, so the shared scratch array keeps the first call's
contents when the second starts; the table images and the keys left behind
(the ascending hashes) must equal bulk_insert's byte for byte, and the
images the oracle's.  Also: a two-level-sort batch (2^20 keys), a sparse
batch, pre-hashed keys, and the size limits (rejected, nothing launched).
Every test prints its case before it touches the GPU."""
import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu

KEY_IS_HASH = 0x08


def _say(*a):
    print(*a, flush=True)


@pytest.mark.parametrize("q,n1,n2", [(16, 50000, 9000), (22, 1 << 20, 300000), (20, 700000, 3)])
def test_noalloc_equals_bulk_insert(gpu, oracle, q, n1, n2):
    from cqf_amd.gqf import GQF, to_device_u64, to_numpy_u64
    _say(f"case bulk_insert_noalloc q={q} calls of {n1} then {n2} keys")
    k1 = oracle.splitmix64(700 + q, n1)
    k2 = np.concatenate([oracle.splitmix64(800 + q, n2 - n2 // 3), k1[: n2 // 3]])  # repeats
    a, b = GQF(q), GQF(q)
    o = oracle.OracleQF(q)
    try:
        for keys in (k1, k2):
            da, db = to_device_u64(keys), to_device_u64(keys)
            a.bulk_insert(da)
            assert b.bulk_insert_noalloc(db) == 0
            torch.cuda.synchronize()
            assert np.array_equal(to_numpy_u64(db), to_numpy_u64(da))
            o.bulk_insert(keys.copy())
            img = b.export_image()
            assert np.array_equal(img, a.export_image())
            assert np.array_equal(img, o.image())
        assert a.status() == 0 and a.layout_errors() == 0 and b.status() == 0 and b.layout_errors() == 0
    finally:
        a.close()
        b.close()
        o.close()


def test_noalloc_prehashed_and_limits(gpu, oracle):
    from cqf_amd.gqf import GQF, to_device_u64, to_numpy_u64
    q = 14
    _say(f"case bulk_insert_noalloc pre-hashed q={q}, then the size limits")
    h1 = oracle.hash_many(oracle.splitmix64(901, 12000), (1 << (q + 8)) - 1)
    h2 = oracle.hash_many(oracle.splitmix64(902, 500), (1 << (q + 8)) - 1)
    a, b = GQF(q), GQF(q)
    try:
        for h in (h1, h2):
            da, db = to_device_u64(h), to_device_u64(h)
            a.bulk_insert(da, KEY_IS_HASH)
            assert b.bulk_insert_noalloc(db, KEY_IS_HASH) == 0
            torch.cuda.synchronize()
            assert np.array_equal(to_numpy_u64(db), to_numpy_u64(da))
            assert np.array_equal(b.export_image(), a.export_image())
        before = b.export_image()
        big = torch.zeros((1 << 20) + 1, dtype=torch.int64, device="cuda")
        assert b.bulk_insert_noalloc(big) == -2  # over the per-call limit: nothing launched
        torch.cuda.synchronize()
        assert np.array_equal(b.export_image(), before)
        assert b.bulk_insert_noalloc(big[:0]) == 0  # empty: a no-op
    finally:
        a.close()
        b.close()
