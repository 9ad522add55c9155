"""CPU tests that pin the oracle (oracle/cqf_oracle.c) before it is trusted.

The reference publishes no golden vectors or known-answer tests for the
GQF (its own tests only assert misses == 0 after inserts and misses == nvals
after a full delete, src/gqf_verify.cu:404,439), and its host path cannot be
built here without stand-ins for cuda.h and Thrust.  So every oracle routine
is checked against something independent of its C code:
  - hash_64: a multiplication-form restatement, bijectivity, the reference's
    own inverse hash_64i, and dependence on key & mask only (H1);
  - counter codec: an encoder written from the format comment, decode of
    every encoding, and the sequential increment path (insert1) reaching the
    same bytes;
  - table image: a pure-Python layout builder from the format's invariants
    (tests/spec_model.py) compared byte for byte with the oracle's sequential
    bulk insert / reduce / delete;
  - counts: brute-force multisets of masked keys;
  - enumeration and cluster census: the multiset and the used-slot stretches
    of layouts built by the spec model;
  - sizing: the closed-form table of xnslots / nblocks / bytes / locks.
"""
import numpy as np
import pytest

from tests import spec_model as S


# ---------------------------------------------------------------------------
# hashing

@pytest.mark.parametrize("bits", [16, 20, 28, 36, 38, 64])
def test_hash_matches_multiplicative_form(oracle, bits):
    mask = (1 << bits) - 1 if bits < 64 else (1 << 64) - 1
    keys = oracle.splitmix64(11, 2000)
    for k in keys[:2000]:
        assert oracle.hash_64(int(k), mask) == S.hash_64(int(k), mask)


def test_hash_bijective_on_16_bits(oracle):
    mask = (1 << 16) - 1
    out = oracle.hash_many(np.arange(1 << 16, dtype=np.uint64), mask)
    assert np.unique(out).size == 1 << 16
    assert int(out.max()) <= mask


@pytest.mark.parametrize("bits", [16, 24, 28, 36, 64])
def test_hash_inverse_and_masking(oracle, bits):
    mask = (1 << bits) - 1 if bits < 64 else (1 << 64) - 1
    for k in oracle.splitmix64(12, 500):
        k = int(k)
        h = oracle.hash_64(k, mask)
        assert oracle.hash_64i(h, mask) == (k & mask)
        assert oracle.hash_64(k & mask, mask) == h  # depends on key & mask only


# ---------------------------------------------------------------------------
# counter codec

@pytest.mark.parametrize("rem", [0, 1, 2, 3, 7, 100, 127, 128, 200, 253, 254, 255])
def test_encode_matches_format(oracle, rem):
    counts = list(range(0, 3000)) + [254 ** 2 - 1, 254 ** 2, 254 ** 2 + 3, 10 ** 6, 1031 ** 2,
                                     10 ** 8 + 7, 2 ** 32 + 5, 2 ** 40 + 123, 2 ** 63]
    for c in counts:
        assert list(oracle.encode_counter(rem, c)) == S.encode(rem, c), (rem, c)


@pytest.mark.parametrize("rem", [0, 1, 5, 128, 254, 255])
def test_decode_inverts_encode(oracle, rem):
    for c in list(range(1, 20000, 7)) + [10 ** 7 + 1, 2 ** 33]:
        run = oracle.encode_counter(rem, c)
        end, r, got = oracle.decode_counter_run(run, 0)
        assert (end, r, got) == (len(run) - 1, rem, c), (rem, c)


def test_decode_delimits_counters_in_a_run(oracle):
    # a run of several counters: every counter decodes to its own bytes
    rng = np.random.default_rng(3)
    for _ in range(300):
        rems = sorted(rng.choice(256, size=rng.integers(1, 6), replace=False).tolist())
        cnts = [int(rng.choice([1, 2, 3, 4, 5, 300, 70000])) for _ in rems]
        run = b"".join(oracle.encode_counter(r, c) for r, c in zip(rems, cnts))
        idx, got = 0, []
        while idx < len(run):
            end, r, c = oracle.decode_counter_run(run, idx)
            got.append((r, c))
            idx = end + 1
        assert got == list(zip(rems, cnts))


@pytest.mark.parametrize("rem", [0, 1, 2, 77, 254, 255])
def test_insert1_increments_reach_encode(oracle, rem):
    """qf_insert with count 1 repeated (the in-place digit increment of
    src/gqf.cu:1650-1706) leaves the bytes encode_counter gives."""
    o = oracle.OracleQF(8)
    q = 40
    h = (q << 8) | rem
    for c in range(1, 1500):
        o.insert(h, 1, oracle.KEY_IS_HASH)
        img = o.image()
        enc = S.encode(rem, c)
        assert list(img[17 + q:17 + q + len(enc)]) == enc, (rem, c)
        assert o.count(h, oracle.KEY_IS_HASH) == c
    o.close()


# ---------------------------------------------------------------------------
# sizing

@pytest.mark.parametrize("nbits,xn,nb,size,locks", [
    (16, 68096, 1064, 86184, 10), (17, 134692, 2105, 170505, 18),
    (20, 1058816, 16544, 1340064, 131), (22, 4214784, 65856, 5334336, 516),
    (24, 16818176, 262784, 21285504, 2055), (28, 268599296, 4196864, 339945984, 32790),
    (30, 1074069504, 16782336, 1359369216, 131114)])
def test_sizes_closed_form(oracle, nbits, xn, nb, size, locks):
    s = oracle.sizes(nbits)
    assert (s["xnslots"], s["nblocks"], s["total_size_in_bytes"], s["num_locks"]) == (xn, nb, size, locks)
    assert s["range"] == 1 << (nbits + 8) and s["key_remainder_bits"] == 8


def test_sizes_odd_q(oracle):
    for nbits in (9, 11, 13, 15, 17, 19, 21, 23):
        n = 1 << nbits
        assert oracle.sizes(nbits)["xnslots"] == n + int(10 * np.sqrt(float(n)))


# ---------------------------------------------------------------------------
# images: the oracle's sequential paths against the spec model

def _multiset(hashes):
    ms = {}
    for h in hashes:
        ms[int(h)] = ms.get(int(h), 0) + 1
    return ms


@pytest.mark.parametrize("nbits,load,seed", [(8, 0.9, 1), (10, 0.5, 2), (10, 0.95, 3), (12, 0.9, 4)])
def test_sequential_insert_matches_spec(oracle, nbits, load, seed):
    n = int(load * (1 << nbits))
    keys = oracle.splitmix64(seed, n)
    o = oracle.OracleQF(nbits)
    k = keys.copy()
    o.bulk_insert(k)
    spec = S.build_image(nbits, _multiset(k))
    assert bytes(o.image()) == spec
    o.close()


@pytest.mark.parametrize("nbits", [9, 11])
def test_skewed_insert_reduce_delete_match_spec(oracle, nbits):
    rng = np.random.default_rng(nbits)
    univ = oracle.splitmix64(5, 96)
    draws = univ[rng.zipf(1.4, size=3000) % univ.size]
    mask = (1 << (nbits + 8)) - 1
    hs = [S.hash_64(int(x), mask) for x in draws]
    ms = _multiset(hs)
    o1 = oracle.OracleQF(nbits)
    o2 = oracle.OracleQF(nbits)
    for part in np.array_split(draws, 4):
        o1.bulk_insert(part.copy())
        o2.bulk_insert_reduce(part.copy())
    spec = S.build_image(nbits, ms)
    assert bytes(o1.image()) == spec
    assert bytes(o2.image()) == spec
    dels = draws[rng.integers(0, draws.size, size=2500)]
    for x in dels:
        h = S.hash_64(int(x), mask)
        if ms.get(h, 0):
            ms[h] -= 1
    o1.bulk_delete(dels.copy())
    assert bytes(o1.image()) == S.build_image(nbits, ms)
    o1.close()
    o2.close()


def test_full_delete_restores_zero_image(oracle):
    nbits = 12
    keys = oracle.splitmix64(8, int(0.95 * (1 << nbits)))
    o = oracle.OracleQF(nbits)
    o.bulk_insert(keys.copy())
    assert o.image().any()
    o.bulk_delete(keys.copy())
    assert not o.image().any()
    o.close()


def test_canonical_builder_matches_spec(oracle):
    rng = np.random.default_rng(9)
    nbits = 10
    hs = np.unique(rng.integers(0, 1 << (nbits + 8), size=250).astype(np.uint64))
    cs = rng.choice([1, 2, 3, 4, 9, 300, 65000], size=hs.size).astype(np.uint64)
    img = oracle.canonical_image(nbits, hs, cs)
    assert bytes(img) == S.build_image(nbits, {int(h): int(c) for h, c in zip(hs, cs)})


# ---------------------------------------------------------------------------
# enumeration and cluster census against layouts built from the invariants

def _spec_multiset(rng, nbits, n, counts=(1,)):
    hs = np.unique(rng.integers(0, 1 << (nbits + 8), size=n).astype(np.uint64))
    cs = rng.choice(counts, size=hs.size)
    return {int(h): int(c) for h, c in zip(hs, cs)}


@pytest.mark.parametrize("nbits,frac,counts", [(8, 0.5, (1,)), (10, 0.9, (1,)), (10, 0.25, (1, 2, 3, 7, 1000)),
                                               (12, 0.95, (1,)), (11, 0.3, (1, 70000))])
def test_enumerate_returns_the_multiset(oracle, nbits, frac, counts):
    rng = np.random.default_rng(nbits * 100 + len(counts))
    ms = _spec_multiset(rng, nbits, int(frac * (1 << nbits)), counts)
    try:
        img = S.build_image(nbits, ms)
    except ValueError:
        pytest.skip("layout does not fit")
    o = oracle.OracleQF(nbits)
    o.set_image(np.frombuffer(img, dtype=np.uint8))
    h, c = o.enumerate()
    assert h.tolist() == sorted(ms)
    assert c.tolist() == [ms[k] for k in sorted(ms)]
    o.close()


@pytest.mark.parametrize("nbits,frac,counts", [(8, 0.9, (1,)), (10, 0.5, (1,)), (10, 0.95, (1,)),
                                               (12, 0.3, (1, 2, 3, 400)), (9, 0.2, (1, 70000))])
def test_cluster_census_matches_layout(oracle, nbits, frac, counts):
    rng = np.random.default_rng(nbits * 7 + len(counts))
    ms = _spec_multiset(rng, nbits, int(frac * (1 << nbits)), counts)
    try:
        img = S.build_image(nbits, ms)
    except ValueError:
        pytest.skip("layout does not fit")
    o = oracle.OracleQF(nbits)
    o.set_image(np.frombuffer(img, dtype=np.uint8))
    lens = o.cluster_census().tolist()
    assert lens == S.clusters(S.used_slots(nbits, ms)[:S.sizes(nbits)[1]])
    assert sum(lens) == sum(S.used_slots(nbits, ms))
    o.close()


def test_cluster_census_special_cases(oracle):
    o = oracle.OracleQF(10)
    assert o.cluster_census().tolist() == []
    # two runs that touch form one cluster; a gap splits them
    img = S.build_image(10, {(5 << 8) | 1: 1, (5 << 8) | 2: 1, (6 << 8) | 9: 3, (20 << 8): 1})
    o.set_image(np.frombuffer(img, dtype=np.uint8))
    # quotient 5: slots 5,6; quotient 6: slots 7..9 (0 9 9 -> r 0 r); quotient 20: slot 20
    assert o.cluster_census().tolist() == [5, 1]
    o.close()


# ---------------------------------------------------------------------------
# counts: brute force over the masked multiset

@pytest.mark.parametrize("nbits", [10, 14])
def test_counts_bruteforce(oracle, nbits):
    rng = np.random.default_rng(nbits)
    base = oracle.splitmix64(21, 400)
    keys = np.concatenate([base, base[:50], base[:10], oracle.splitmix64(22, 300)])
    rng.shuffle(keys)
    mask = (1 << (nbits + 8)) - 1
    o = oracle.OracleQF(nbits)
    o.bulk_insert(keys.copy())
    masked = {}
    for k in keys:
        masked[int(k) & mask] = masked.get(int(k) & mask, 0) + 1
    probe = np.concatenate([keys, oracle.splitmix64(23, 3000)])
    got = o.bulk_get(probe)
    want = np.array([masked.get(int(k) & mask, 0) for k in probe], dtype=np.uint64)
    assert np.array_equal(got, want)
    o.close()


def test_false_positives_are_exactly_mask_collisions(oracle):
    nbits = 12
    mask = (1 << (nbits + 8)) - 1
    keys = oracle.splitmix64(31, 3000)
    other = oracle.splitmix64(32, 200000)
    o = oracle.OracleQF(nbits)
    o.bulk_insert(keys.copy())
    got = o.bulk_get(other) > 0
    inserted = set(int(k) & mask for k in keys)
    want = np.array([int(k) & mask in inserted for k in other])
    assert np.array_equal(got, want)
    o.close()


def test_table_full_returns_no_space(oracle):
    nbits = 8
    o = oracle.OracleQF(nbits)
    n_ok = 0
    rets = []
    for k in oracle.splitmix64(41, 2 * o.xnslots):
        r = o.insert(int(k))
        rets.append(r)
        if r >= 0:
            n_ok += 1
    assert -1 in rets  # QF_NO_SPACE once full
    o.close()


def test_insert_not_exists_semantics(oracle):
    o = oracle.OracleQF(10)
    k = 123456789
    assert o.insert_not_exists(k) == (0, 77)   # inserted; returnedVal untouched
    ret, val = o.insert_not_exists(k)
    assert (ret, val) == (1, 0)                # found; returnedVal is always 0 (H13)
    assert o.count(k) == 1
    o.close()


# ---------------------------------------------------------------------------
# insert1_if_not_exists (src/gqf.cu:1364-1518), pinned two ways:
#   1. on count-1 tables the slot-wise search is the counter-wise one, so
#      not-exists inserts of any hash sequence leave exactly the canonical
#      image of the SET of hashes (tests/spec_model.build_image);
#   2. on runs with multi-slot counters the search is slot-wise: hand-derived
#      run bytes from the reference lines, and a model of the search on run
#      byte lists (tests/spec_model.insert_not_exists) whose layout is the
#      spec model's, over random mixed tables.

KEY_IS_HASH = 0x08
INSERTED, FOUND = 0, 1


def test_not_exists_count1_tables_are_the_canonical_set(oracle):
    import random
    rnd = random.Random(11)
    for nbits, n in ((8, 230), (10, 900), (12, 3800)):
        o = oracle.OracleQF(nbits)
        rng = 1 << (nbits + 8)
        hs = [rnd.randrange(rng) for _ in range(n)]
        hs += [hs[rnd.randrange(len(hs))] for _ in range(n // 4)]  # repeats: FOUND
        seen = set()
        for h in hs:
            ret, val = o.insert_not_exists(h, KEY_IS_HASH)
            if h in seen:
                assert (ret, val) == (FOUND, 0), h  # returnedVal is always 0 (H13)
            else:
                assert (ret, val) == (INSERTED, 77), h  # returnedVal untouched
            seen.add(h)
        want = S.build_image(nbits, {h: 1 for h in seen})
        assert o.image().tobytes() == want, nbits
        # a found key changes nothing
        before = o.image()
        for h in list(seen)[:50]:
            assert o.insert_not_exists(h, KEY_IS_HASH)[0] == FOUND
        assert np.array_equal(o.image(), before)
        o.close()


def _mixed_table(oracle, nbits, runs_spec):
    """A table holding {quotient: [(remainder, count), ...]} (inserted with
    counts through the oracle, whose images the spec model pins), and its
    run bytes."""
    o = oracle.OracleQF(nbits)
    ms = {}
    for q, rcs in runs_spec.items():
        for r, c in rcs:
            h = (q << 8) | r
            assert o.insert(h, c, KEY_IS_HASH) >= 0
            ms[h] = ms.get(h, 0) + c
    runs = S.runs_of(ms)
    assert o.image().tobytes() == S.image_from_runs(nbits, runs)
    return o, runs


def test_not_exists_multislot_known_answers(oracle):
    """Hand-derived cases on a count-300 counter of remainder 5 at quotient
    126 (the run [5, 2, 45, 5]: 297 = 1 x 254 + 43, digits 1 + 1 = 2 and
    43 + 1 = 44, which is >= 5 and becomes 45) and a count-3 counter of
    remainder 0 at quotient 127 ([0, 0, 0]), inside a cluster that crosses
    into block 2, so each insert shifts the runs after it and moves block
    2's offset."""
    assert S.encode(5, 300) == [5, 2, 45, 5]
    assert S.encode(0, 3) == [0, 0, 0]
    spec = {q: [(17, 1), (90, 1)] for q in range(118, 126)}
    spec[126] = [(5, 300)]
    spec[127] = [(0, 3)]
    spec.update({q: [(33, 1), (60, 1), (61, 1)] for q in range(128, 134)})
    cases = [
        # :1413-1419 skips 5 and 2 (both < 7) and stops at 45 >= 7; 45 != 7,
        # so :1431-1436 (operation 2) puts 7 before 45, inside the counter
        ((126 << 8) | 7, INSERTED, {126: [5, 2, 7, 45, 5]}),
        # the skip stops at 45 == 45: :1437-1444 returns QF_ITEM_FOUND (a
        # digit, not a remainder), nothing moves
        ((126 << 8) | 45, FOUND, {}),
        # every slot is below 200: the index passes the run end, :1425-1429
        # (operation 1) appends after it and moves the run end bit
        ((126 << 8) | 200, INSERTED, {126: [5, 2, 45, 5, 200]}),
        # the first slot, 5, is >= 3 and != 3: operation 2 at the run start
        ((126 << 8) | 3, INSERTED, {126: [3, 5, 2, 45, 5]}),
        # the zero counter: slot 0 == 0 at the run start, found
        ((127 << 8) | 0, FOUND, {}),
        # 1 is above every slot of [0, 0, 0]: appended
        ((127 << 8) | 1, INSERTED, {127: [0, 0, 0, 1]}),
        # an empty quotient past the cluster (slot 140 is used by it): a new
        # run at the cluster's end (operation 0, :1447-1466)
        ((134 << 8) | 9, INSERTED, {134: [9]}),
    ]
    for h, ret, change in cases:
        o, runs = _mixed_table(oracle, 10, spec)
        before = o.image()
        got, val = o.insert_not_exists(h, KEY_IS_HASH)
        assert got == ret, hex(h)
        if ret == FOUND:
            assert val == 0
            assert np.array_equal(o.image(), before), hex(h)
        else:
            runs.update(change)
            assert o.image().tobytes() == S.image_from_runs(10, runs), hex(h)
        o.close()


@pytest.mark.parametrize("seed", [1, 2, 3])
def test_not_exists_slotwise_model_on_random_mixed_tables(oracle, seed):
    """Random tables with counters of counts 1..70000 (multi-slot), then 300
    not-exists inserts, most aimed at occupied quotients: return codes and
    the images equal the slot-wise model on run bytes."""
    import random
    rnd = random.Random(seed)
    nbits = 10
    qs = rnd.sample(range(1 << nbits), 120)
    spec = {}
    for q in qs:
        rems = sorted(rnd.sample(range(256), rnd.choice((1, 1, 2, 3))))
        spec[q] = [(r, rnd.choice((1, 1, 2, 3, 4, 17, 254, 300, 70000))) for r in rems]
    o, runs = _mixed_table(oracle, nbits, spec)
    for i in range(300):
        q = rnd.choice(qs) if i % 4 else rnd.randrange(1 << nbits)
        h = (q << 8) | rnd.randrange(256)
        want = S.insert_not_exists(runs, h)
        got, val = o.insert_not_exists(h, KEY_IS_HASH)
        assert got == (FOUND if want == "found" else INSERTED), (i, hex(h))
        if i % 20 == 0 or want == "inserted":
            assert o.image().tobytes() == S.image_from_runs(nbits, runs), (i, hex(h))
    o.close()


def test_remove_path_truncates_offsets_h5(oracle):
    """Hazard H5.  The reference's remove path stores each re-derived block
    offset in the u8 field without saturating (src/gqf.cu:1172-1190), while
    its inserts saturate at 255 (:1748-1753).  Quotients 0..3 with 200
    remainders each fill slots 0..799: blocks 1..8 have exact offsets
    736 - 64 (b - 1) > 255 (stored 255), block 9 has 224.  Removing one key
    ends the runs of quotients < 64 at 798, and the loop writes, derived by
    hand from those lines:
      block 1: 798 - 127 + 64 = 735, stored as 735 mod 256 = 223;
      block 2: run_end(127) reads block 1's 223 > 63 with no occupied
               quotient, so it is 64 + 223 - 1 = 286: offset 286 - 127 = 159;
      block 3: 286 - 191 = 95;  block 4: 286 - 255 = 31;
      block 5: run_end(319) reads 31 <= 63: 319, which is in block 4, so 0;
      blocks 6..12: 0 likewise (their stored 255 / 224 / 160 / 96 / 32 differ
               from 0); block 13 already holds 0 and the loop stops.
    Slots, occupieds and run ends equal the canonical image; only those
    offset bytes differ from it (255 x 8, 223, 159, 95, 31).  The device
    table keeps exact offsets and exports the canonical image instead."""
    o = oracle.OracleQF(10)
    ms = {}
    for q in range(4):
        for r in range(200):
            h = (q << 8) | r
            assert o.insert(h, 1, KEY_IS_HASH) >= 0
            ms[h] = 1
    img = o.image()
    assert img.tobytes() == S.build_image(10, ms)
    assert [int(img[81 * b]) for b in range(14)] == [0] + [255] * 8 + [224, 160, 96, 32, 0]
    h = (1 << 8) | 7
    assert o.remove(h, 1, KEY_IS_HASH) >= 0
    del ms[h]
    img = o.image()
    canon = np.frombuffer(S.build_image(10, ms), dtype=np.uint8)
    assert [int(img[81 * b]) for b in range(14)] == [0, 223, 159, 95, 31] + [0] * 9
    assert [int(canon[81 * b]) for b in range(14)] == [0] + [255] * 8 + [223, 159, 95, 31, 0]
    diff = np.nonzero(img != canon)[0]
    assert diff.size and all(p % 81 == 0 for p in diff)  # offset bytes only
    o.close()
