"""Whole-table operations on the GPU against the oracle: enumeration (the
reference's iterator, src/gqf.cu:3655-3840), table statistics, the cluster
census (src/cluster_length_test.cu:37-120), the qfmetadata file format
(include/gqf_int.cuh:96-113), growth (qf_resize_malloc, src/gqf.cu:2301-2330)
and merges (qf_merge / qf_multi_merge, src/gqf.cu:3896-4008).

Every GPU result is compared with the oracle on the exported 81-byte image
(orc_enumerate, orc_cluster_census) or with an image the oracle builds
from the multiset (orc_build_canonical), and with brute-force counts."""
import struct

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

MAGIC = 1018874902021329732


def _say(*a):
    print(*a, flush=True)


def _filled(oracle, nbits, load, seed, dup_every=0):
    """GQF with splitmix64(seed) keys (every dup_every-th key inserted 3 more
    times when set, giving multi-slot counters); returns (gqf, keys)."""
    from cqf_amd.gqf import GQF, to_device_u64
    keys = oracle.splitmix64(seed, int(load * (1 << nbits)))
    if dup_every:
        extra = np.repeat(keys[::dup_every], 3)
        keys = np.concatenate([keys, extra])
    g = GQF(nbits)
    g.bulk_insert(to_device_u64(keys))
    assert g.status() == 0 and g.layout_errors() == 0, "the test's keys overflow the table"
    return g, keys


# (nbits, load, seed, dup_every): slots used stay below xnslots (a duplicated
# key's counter of 4 takes 3 slots)
CASES = [(10, 0.5, 1, 0), (12, 0.9, 2, 0), (12, 0.7, 3, 7), (16, 0.9, 4, 0), (14, 0.45, 5, 2)]


@pytest.mark.parametrize("nbits,load,seed,dup", CASES)
def test_enumerate_matches_oracle(gpu, oracle, nbits, load, seed, dup):
    from cqf_amd.gqf import QF_ENUM_KEY, to_numpy_u64
    _say(f"case enumerate q={nbits} load={load} dup={dup}")
    g, keys = _filled(oracle, nbits, load, seed, dup)
    o = oracle.OracleQF(nbits)
    try:
        img = g.export_image()
        o.set_image(img)
        oh, oc = o.enumerate()
        h, c = g.enumerate()
        assert np.array_equal(to_numpy_u64(h), oh)
        assert np.array_equal(to_numpy_u64(c), oc)
        # the multiset: each masked key's hash with its multiplicity
        mask = (1 << (nbits + 8)) - 1
        ms = {}
        for x in oracle.hash_many(keys, mask).tolist():
            ms[x] = ms.get(x, 0) + 1
        assert oh.tolist() == sorted(ms) and oc.tolist() == [ms[k] for k in sorted(ms)]
        # keys: hash_64i of the hashes (qfi_get_key)
        k, c2 = g.enumerate(form=QF_ENUM_KEY)
        want = np.array([oracle.hash_64i(int(x), mask) for x in oh[:3000]], dtype=np.uint64)
        assert np.array_equal(to_numpy_u64(k)[:3000], want)
        assert np.array_equal(to_numpy_u64(c2), oc)
    finally:
        g.close()
        o.close()


def test_enumerate_keys_recover_small_keys(gpu, oracle):
    """Keys below 2^(q+8) are stored exactly: enumeration gives them back."""
    from cqf_amd.gqf import GQF, QF_ENUM_KEY, to_device_u64, to_numpy_u64
    nbits = 14
    _say("case enumerate keys < 2^(q+8)")
    rng = np.random.default_rng(3)
    keys = np.unique(rng.integers(0, 1 << (nbits + 8), size=5000).astype(np.uint64))
    counts = rng.integers(1, 5, size=keys.size)
    g = GQF(nbits)
    try:
        g.bulk_insert(to_device_u64(np.repeat(keys, counts)))
        assert g.status() == 0 and g.layout_errors() == 0
        k, c = g.enumerate(form=QF_ENUM_KEY)
        got = dict(zip(to_numpy_u64(k).tolist(), to_numpy_u64(c).tolist()))
        assert got == dict(zip(keys.tolist(), counts.tolist()))
    finally:
        g.close()


def test_enumerate_ranges_concatenate(gpu, oracle):
    from cqf_amd.gqf import to_numpy_u64
    nbits = 16
    _say("case enumerate quotient ranges q=16")
    g, _ = _filled(oracle, nbits, 0.65, 9, 5)
    try:
        h, c = g.enumerate()
        h, c = to_numpy_u64(h), to_numpy_u64(c)
        cuts = [0, 1000, 1001, 5000, 4096 * 3, 40000, 1 << nbits]
        parts_h, parts_c = [], []
        for a, b in zip(cuts[:-1], cuts[1:]):
            ph, pc = g.enumerate(a, b)
            ph = to_numpy_u64(ph)
            assert ((ph >> np.uint64(8)) >= a).all() and ((ph >> np.uint64(8)) < b).all()
            parts_h.append(ph)
            parts_c.append(to_numpy_u64(pc))
        assert np.array_equal(np.concatenate(parts_h), h)
        assert np.array_equal(np.concatenate(parts_c), c)
        empty_h, _ = g.enumerate(500, 500)
        assert empty_h.numel() == 0
    finally:
        g.close()


@pytest.mark.parametrize("nbits,load,seed,dup", CASES)
def test_stats_and_census_match_oracle(gpu, oracle, nbits, load, seed, dup):
    from cqf_amd.gqf import to_numpy_u64
    _say(f"case stats/census q={nbits} load={load} dup={dup}")
    g, _ = _filled(oracle, nbits, load, seed, dup)
    o = oracle.OracleQF(nbits)
    try:
        o.set_image(g.export_image())
        oh, oc = o.enumerate()
        lens = o.cluster_census()
        st = g.stats()
        assert st == {"nelts": int(oc.sum()), "ndistinct_elts": int(oh.size), "noccupied_slots": int(lens.sum())}
        glens, hist, mx = g.cluster_census(nbins=64)
        assert np.array_equal(to_numpy_u64(glens), lens)
        want = np.bincount(np.minimum(lens, 63).astype(np.int64), minlength=64)
        assert np.array_equal(hist.cpu().numpy(), want)
        assert mx == int(lens.max())
    finally:
        g.close()
        o.close()


def test_census_empty_and_saturated(gpu, oracle):
    from cqf_amd.gqf import GQF, to_numpy_u64
    nbits = 14
    _say("case census empty / saturated cluster")
    g = GQF(nbits)
    o = oracle.OracleQF(nbits)
    try:
        lens, hist, mx = g.cluster_census(nbins=8)
        assert lens.numel() == 0 and int(hist.sum()) == 0 and mx == 0
        assert g.stats() == {"nelts": 0, "ndistinct_elts": 0, "noccupied_slots": 0}
        # one long cluster of multi-slot counters (offsets past 255)
        rng = np.random.default_rng(1)
        hs = []
        for q in range(320, 384):
            for r in sorted(rng.choice(256, size=8, replace=False)):
                hs.append((q << 8) | int(r))
        hs = np.array(hs, dtype=np.uint64)
        cs = np.full(hs.size, 10 ** 6, dtype=np.uint64)
        img = oracle.canonical_image(nbits, hs, cs)
        g.import_image(img)
        o.set_image(img)
        lens, _, mx = g.cluster_census()
        assert np.array_equal(to_numpy_u64(lens), o.cluster_census())
        h, c = g.enumerate()
        assert np.array_equal(to_numpy_u64(h), hs) and np.array_equal(to_numpy_u64(c), cs)
    finally:
        g.close()
        o.close()


def _meta(blob):
    f = struct.unpack("<QiIQIIQQQQQQQQQQQ", blob[:120])
    names = ["magic", "hash_mode", "reserved", "total_size", "seed", "pad", "nslots", "xnslots", "key_bits",
             "value_bits", "key_remainder_bits", "bits_per_slot", "range", "nblocks", "nelts", "ndistinct",
             "noccupied"]
    return dict(zip(names, f))


def test_save_load_round_trip(gpu, oracle, tmp_path):
    from cqf_amd.gqf import GQF, to_device_u64, to_numpy_u64
    nbits = 14
    _say("case save/load q=14")
    g, keys = _filled(oracle, nbits, 0.6, 11, 4)
    path = tmp_path / "f.cqf"
    try:
        n = g.save(str(path))
        blob = path.read_bytes()
        assert n == len(blob)
        m = _meta(blob)
        sz = oracle.sizes(nbits)
        st = g.stats()
        assert m["magic"] == MAGIC and m["hash_mode"] == 1 and m["seed"] == 0
        assert (m["nslots"], m["xnslots"], m["nblocks"]) == (1 << nbits, sz["xnslots"], sz["nblocks"])
        assert (m["key_bits"], m["value_bits"], m["key_remainder_bits"], m["bits_per_slot"]) == (nbits + 8, 0, 8, 8)
        assert m["range"] == 1 << (nbits + 8) and m["total_size"] == sz["nblocks"] * 81
        assert (m["nelts"], m["ndistinct"], m["noccupied"]) == (st["nelts"], st["ndistinct_elts"],
                                                               st["noccupied_slots"])
        img = g.export_image()
        assert blob[120:] == img.tobytes()
        h = GQF.load(str(path))
        try:
            assert np.array_equal(h.export_image(), img)
            probe = to_device_u64(np.concatenate([keys[:3000], oracle.splitmix64(12, 3000)]))
            assert np.array_equal(to_numpy_u64(h.bulk_get(probe)), to_numpy_u64(g.bulk_get(probe)))
        finally:
            h.close()
    finally:
        g.close()


def test_load_file_built_from_the_format(gpu, oracle, tmp_path):
    """A file assembled here from the struct layout and an oracle image."""
    from cqf_amd.gqf import GQF
    nbits = 12
    _say("case load hand-built file q=12")
    sz = oracle.sizes(nbits)
    keys = oracle.splitmix64(13, 3000)
    o = oracle.OracleQF(nbits)
    o.bulk_insert(keys.copy())
    img = o.image()
    o.close()
    hdr = struct.pack("<QiIQIIQQQQQQQQQQQ", MAGIC, 1, 0, sz["nblocks"] * 81, 0, 0, 1 << nbits, sz["xnslots"],
                      nbits + 8, 0, 8, 8, 1 << (nbits + 8), sz["nblocks"], 0, 0, 0)
    good = tmp_path / "good.cqf"
    good.write_bytes(hdr + img.tobytes())
    g = GQF.load(str(good))
    try:
        assert np.array_equal(g.export_image(), img)
    finally:
        g.close()
    bad = tmp_path / "bad.cqf"
    bad.write_bytes(struct.pack("<Q", MAGIC + 1) + hdr[8:] + img.tobytes())
    with pytest.raises(RuntimeError):
        GQF.load(str(bad))
    short = tmp_path / "short.cqf"
    short.write_bytes(hdr + img.tobytes()[:-1])
    with pytest.raises(RuntimeError):
        GQF.load(str(short))


@pytest.mark.parametrize("nbits,grow,dup", [(12, 1, 0), (12, 2, 5), (14, 1, 3)])
def test_resize_keeps_counts_and_layout(gpu, oracle, nbits, grow, dup):
    from cqf_amd.gqf import GQF, to_device_u64, to_numpy_u64
    _say(f"case resize q={nbits} +{grow} dup={dup}")
    g, keys = _filled(oracle, nbits, 0.5 if dup else 0.95, 20 + grow, dup)
    probe = to_device_u64(np.concatenate([keys[:4000], oracle.splitmix64(30, 4000)]))
    k = None
    try:
        before = to_numpy_u64(g.bulk_get(probe))
        oh, oc = (to_numpy_u64(t) for t in g.enumerate())
        moved = g.resize(nbits + grow)
        assert moved == oh.size
        assert g.info["nslots"] == 1 << (nbits + grow) and g.info["value_bits"] == grow
        assert g.info["key_bits"] == nbits + 8 and g.info["key_remainder_bits"] == 8 - grow
        assert np.array_equal(to_numpy_u64(g.bulk_get(probe)), before)
        # the layout is the canonical one of the slot hashes hash << grow
        want = oracle.canonical_image(nbits + grow, oh << np.uint64(grow), oc)
        assert np.array_equal(g.export_image(), want)
        h2, c2 = g.enumerate()
        assert np.array_equal(to_numpy_u64(h2), oh) and np.array_equal(to_numpy_u64(c2), oc)
        # the same as a filter made with those value bits and the same keys
        k = GQF(nbits + grow, value_bits=grow)
        k.bulk_insert(to_device_u64(keys))
        assert np.array_equal(k.export_image(), want)
        # further inserts and deletes keep brute-force counts; a full delete empties it
        more = oracle.splitmix64(31, 1 << nbits)
        g.bulk_insert(to_device_u64(more))
        allk = np.concatenate([keys, more])
        mask = np.uint64((1 << (nbits + 8)) - 1)
        u, cnt = np.unique(allk & mask, return_counts=True)
        got = to_numpy_u64(g.bulk_get(to_device_u64(u)))
        assert np.array_equal(got, cnt.astype(np.uint64))
        g.bulk_delete(to_device_u64(allk))
        gpu.cuda.synchronize()
        assert not g.export_image().any()
        assert g.status() == 0 and g.layout_errors() == 0
    finally:
        g.close()
        if k is not None:
            k.close()


def test_merge_equals_inserting_everything(gpu, oracle):
    from cqf_amd.gqf import GQF, to_device_u64, to_numpy_u64
    nbits = 14
    _say("case merge / multi-merge q=14")
    parts = [oracle.splitmix64(40 + i, 4000 + 1000 * i) for i in range(3)]
    parts[1] = np.concatenate([parts[1], parts[0][:500], parts[0][:500]])  # shared keys: counts add
    fs = []
    try:
        for p in parts:
            f = GQF(nbits)
            f.bulk_insert(to_device_u64(p))
            fs.append(f)
        ref = GQF(nbits)
        fs.append(ref)
        ref.bulk_insert(to_device_u64(np.concatenate(parts)))
        c = GQF(nbits)
        fs.append(c)
        c.merge_from(fs[0], fs[1])
        ref2 = GQF(nbits)
        fs.append(ref2)
        ref2.bulk_insert(to_device_u64(np.concatenate(parts[:2])))
        assert np.array_equal(c.export_image(), ref2.export_image())
        d = GQF(nbits)
        fs.append(d)
        d.bulk_insert(to_device_u64(parts[2][:100]))  # a non-empty target keeps its items
        d.merge_from(fs[0], fs[1])
        d.bulk_insert(to_device_u64(parts[2][100:]))
        assert np.array_equal(d.export_image(), ref.export_image())
        # into a larger table (value bits): same counts
        e = GQF(nbits + 1, value_bits=1)
        fs.append(e)
        e.merge_from(fs[0], fs[1], fs[2])
        probe = to_device_u64(np.concatenate([np.concatenate(parts)[:5000], oracle.splitmix64(49, 5000)]))
        assert np.array_equal(to_numpy_u64(e.bulk_get(probe)), to_numpy_u64(ref.bulk_get(probe)))
        # mismatched hashes are refused
        bad = GQF(nbits + 1)
        fs.append(bad)
        with pytest.raises(RuntimeError):
            bad.merge_from(fs[0])
        with pytest.raises(RuntimeError):
            fs[0].resize(nbits - 1)
        assert all(f.status() == 0 for f in fs)
    finally:
        for f in fs:
            f.close()
