"""Hazard H5 on the device: deletes in a table whose offsets exceed 255.

The reference's remove path stores re-derived block offsets in a u8 field
without saturating (src/gqf.cu:1172-1190), so once a cluster pushes an
offset past 255, a delete leaves offset bytes that no longer describe the
runs (tests/test_oracle.py::test_remove_path_truncates_offsets_h5 pins the
oracle's restatement of it byte by byte).  The device table keeps exact
offsets, so parity here is to the canonical image -- runs at max(q, previous
end + 1), offsets min(exact, 255) -- that the spec model builds from the
multiset, and to brute-force counts.  Every test prints its case first."""
import numpy as np
import pytest
import torch

from tests import spec_model as S

pytestmark = pytest.mark.gpu

KEY_IS_HASH = 0x08


def _say(*a):
    print(*a, flush=True)


def test_delete_with_offsets_above_255_keeps_canonical_image(gpu, oracle):
    from cqf_amd.gqf import GQF, to_device_u64, to_numpy_u64
    q = 10
    mask = (1 << (q + 8)) - 1
    hashes = np.array([(qq << 8) | r for qq in range(4) for r in range(200)], dtype=np.uint64)
    _say(f"case H5 q={q}: 800 hashes on quotients 0..3, one delete, then inserts on blocks 4..12")
    ms = {int(h): 1 for h in hashes}
    keys = np.array([oracle.hash_64i(int(h), mask) for h in hashes], dtype=np.uint64)  # hash_64(key) = h
    f = GQF(q)
    try:
        f.bulk_insert(to_device_u64(hashes), flags=KEY_IS_HASH)
        torch.cuda.synchronize()
        img = f.export_image()
        assert img.tobytes() == S.build_image(q, ms)
        assert [int(img[81 * b]) for b in range(10)] == [0] + [255] * 8 + [224]
        # one delete: the reference would now write 223, 159, 95, 31, 0, ...
        gone = (1 << 8) | 7
        f.bulk_delete(to_device_u64(np.array([gone], dtype=np.uint64)), flags=KEY_IS_HASH)
        del ms[gone]
        torch.cuda.synchronize()
        img = f.export_image()
        canon = S.build_image(q, ms)
        assert img.tobytes() == canon
        o = oracle.OracleQF(q)
        for h in hashes:
            o.insert(int(h), 1, KEY_IS_HASH)
        o.remove(gone, 1, KEY_IS_HASH)
        ref_img = o.image()
        o.close()
        diff = np.nonzero(ref_img != np.frombuffer(canon, dtype=np.uint8))[0]
        assert diff.size and all(p % 81 == 0 for p in diff)  # the reference differs in offset bytes only
        # brute-force counts through the hashing path (keys hash to the hashes)
        got = to_numpy_u64(f.bulk_get(to_device_u64(keys)))
        want = np.array([ms.get(int(h), 0) for h in hashes], dtype=np.uint64)
        assert np.array_equal(got, want)
        # inserts on quotients whose blocks the reference's bytes misdescribe
        more = np.array([(qq << 8) | r for qq in (300, 301, 450, 700) for r in (3, 90, 91)], dtype=np.uint64)
        f.bulk_insert(to_device_u64(more), flags=KEY_IS_HASH)
        for h in more:
            ms[int(h)] = ms.get(int(h), 0) + 1
        torch.cuda.synchronize()
        assert f.export_image().tobytes() == S.build_image(q, ms)
        # an image whose offset bytes disagree with its runs (what the
        # reference wrote) is rejected by import
        with pytest.raises(RuntimeError):
            f.import_image(ref_img)
        assert f.export_image().tobytes() == S.build_image(q, ms)  # unchanged by the rejected import
        # and everything out again: the zero image
        rest = np.array(sorted(ms), dtype=np.uint64)
        f.bulk_delete(to_device_u64(rest), flags=KEY_IS_HASH)
        torch.cuda.synchronize()
        assert not f.export_image().any()
        assert f.status() == 0 and f.layout_errors() == 0
    finally:
        f.close()
