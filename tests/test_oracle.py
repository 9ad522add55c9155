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
