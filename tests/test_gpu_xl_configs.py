"""The benchmark configurations at their stated sizes (one MI355X).

  config 2: 2^28 slots; 241,591,910 splitmix64(3) keys in 14 batches of
            2^24 and one of 6,710,886; the 81-byte image and bulk_get counts
            of every inserted key and of 2^28 other keys equal the oracle's;
            the seeded half delete leaves the oracle's image; deleting the
            rest leaves the all-zero image (the checks of
            src/gqf_verify.cu:368-441 and src/test.cu:478-484, made exact);
  config 3: 2^28 Zipf(1.5) draws over 2^24 ids into 2^28 slots with one
            bulk_insert_reduce: every id reads its exact multiplicity and
            the image is the canonical image of those counts;
  config 4: the 256 seed-5 k-mer filters (q 16..24, loads 0.5..0.95) in one
            autotuned group insert and one group query: per-filter counts
            equal a brute-force count of the masked keys, and every image is
            the canonical image of its filter's multiset;
  config 5: the 64-filter (q 22..30) served stream under an HBM budget that
            forces evictions: every get returns what the requests applied
            one at a time in arrival order give.

The oracle runs on the host's cores (its queries on 16 threads: they only
read the table).  Each test prints its case and its phases."""
import concurrent.futures as cf
import time

import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu

Q2 = 28
N2 = 241_591_910            # floor(0.9 * 2^28)
B2 = 1 << 24


def _say(*a):
    print(*a, flush=True)


def _oracle_counts(o, vals, threads=16):
    """o.bulk_get(vals) computed in `threads` slices at once."""
    vals = np.ascontiguousarray(vals, dtype=np.uint64)
    out = np.empty_like(vals)
    cuts = np.linspace(0, vals.size, threads + 1).astype(np.int64)

    def part(k):
        out[cuts[k]:cuts[k + 1]] = o.bulk_get(vals[cuts[k]:cuts[k + 1]])

    with cf.ThreadPoolExecutor(threads) as ex:
        list(ex.map(part, range(threads)))
    return out


@pytest.fixture(scope="module")
def cfg2(gpu, oracle):
    from cqf_amd import gqf
    t0 = time.time()
    _say(f"case config 2: q={Q2}, {N2} keys of splitmix64(3) in batches of 2^24")
    f = gqf.GQF(Q2)
    f.reserve(B2)
    master = torch.empty(N2, dtype=torch.int64, device="cuda")
    gqf.fill_splitmix64(master, 3, 0)
    work = master.clone()
    bounds = list(range(0, N2, B2)) + [N2]
    assert len(bounds) == 16 and bounds[-1] - bounds[-2] == 6_710_886
    for i in range(len(bounds) - 1):
        f.bulk_insert(work[bounds[i]:bounds[i + 1]])
    torch.cuda.synchronize()
    keys = master.cpu().numpy().view(np.uint64)
    _say(f"  gpu inserts done ({time.time() - t0:.1f} s); oracle inserting")
    o = oracle.OracleQF(Q2)
    assert o.bulk_insert(keys.copy()) == 0
    _say(f"  oracle inserts done ({time.time() - t0:.1f} s)")
    st = {"f": f, "o": o, "master": master, "work": work, "keys": keys, "bounds": bounds}
    yield st
    f.close()
    o.close()


def test_config2_insert_image_and_sorted_keys(cfg2, oracle):
    f, o = cfg2["f"], cfg2["o"]
    _say("  config 2: image after the inserts")
    assert np.array_equal(f.export_image(), o.image())
    assert f.status() == 0 and f.layout_errors() == 0
    # each batch's keys are left as its ascending hashes (first and last batch)
    mask = (1 << (Q2 + 8)) - 1
    b = cfg2["bounds"]
    for i in (0, len(b) - 2):
        want = np.sort(oracle.hash_many(cfg2["keys"][b[i]:b[i + 1]], mask))
        got = cfg2["work"][b[i]:b[i + 1]].cpu().numpy().view(np.uint64)
        assert np.array_equal(got, want), i


def test_config2_counts_of_inserted_and_other_keys(cfg2, oracle):
    from cqf_amd import gqf
    f, o = cfg2["f"], cfg2["o"]
    t0 = time.time()
    _say(f"  config 2: bulk_get of the {N2} inserted keys and of 2^28 others")
    got_ins = f.bulk_get(cfg2["master"]).cpu().numpy().view(np.uint64)
    other = torch.empty(1 << Q2, dtype=torch.int64, device="cuda")
    gqf.fill_splitmix64(other, 3, N2)  # outputs N2 .. N2 + 2^28 - 1 of the same stream
    got_oth = f.bulk_get(other).cpu().numpy().view(np.uint64)
    assert f.bulk_get_misses(cfg2["master"]) == 0  # src/test.cu:478-484
    _say(f"  gpu queries done ({time.time() - t0:.1f} s); oracle queries on 16 threads")
    want_ins = _oracle_counts(o, cfg2["keys"])
    assert np.array_equal(got_ins, want_ins)
    del want_ins
    want_oth = _oracle_counts(o, other.cpu().numpy().view(np.uint64))
    assert np.array_equal(got_oth, want_oth)
    # false positives at the closed form 0.9 / 256 of 2^28 (binomial, 6 sigma)
    fp = int((got_oth > 0).sum())
    assert abs(fp - 943_718) < 6 * 972, fp
    _say(f"  config 2 counts equal ({time.time() - t0:.1f} s); false positives {fp}")


def test_config2_half_then_full_delete(cfg2):
    from cqf_amd import gqf
    f, o = cfg2["f"], cfg2["o"]
    t0 = time.time()
    master = cfg2["master"]
    sel = torch.empty_like(master)
    gqf.fill_splitmix64(sel, 7, 0)
    half = master[(sel & 1) == 1].contiguous()
    kept = master[(sel & 1) == 0].contiguous()
    del sel
    _say(f"  config 2: delete the seeded half ({half.numel()} keys) in batches of 2^24")
    work = half.clone()
    for i in range(0, half.numel(), B2):
        f.bulk_delete(work[i:i + B2])
    torch.cuda.synchronize()
    assert o.bulk_delete(half.cpu().numpy().view(np.uint64).copy()) == 0
    _say(f"  deletes done on both ({time.time() - t0:.1f} s)")
    assert np.array_equal(f.export_image(), o.image())
    assert bool((f.bulk_get(kept) > 0).all())
    _say(f"  config 2: delete the other {kept.numel()} keys")
    work = kept.clone()
    for i in range(0, kept.numel(), B2):
        f.bulk_delete(work[i:i + B2])
    torch.cuda.synchronize()
    assert not f.export_image().any()
    assert f.bulk_get_misses(master) == N2  # src/gqf_verify.cu:439
    assert f.status() == 0 and f.layout_errors() == 0


def test_config3_zipf_exact_multiplicities(gpu, oracle):
    from cqf_amd import gqf
    from cqf_amd import workloads as W
    q, n_ids, n_draws = 28, 1 << 24, 1 << 28
    t0 = time.time()
    _say(f"case config 3: {n_draws} Zipf(1.5) draws over {n_ids} ids (seed 4), q={q}, bulk_insert_reduce")
    keys, draws, mult = W.zipf_workload(n_ids, n_draws, s=1.5, seed=4, distinct_bits=36, device="cuda")
    assert int(mult.max()) > 10 ** 7  # multi-digit counters
    f = gqf.GQF(q)
    try:
        f.bulk_insert_reduce(draws)
        got = f.bulk_get(keys)
        assert torch.equal(got, mult)
        assert f.status() == 0 and f.layout_errors() == 0
        _say(f"  exact multiplicities ({time.time() - t0:.1f} s); canonical image")
        mask = (1 << (q + 8)) - 1
        hk = oracle.hash_many(keys.cpu().numpy().view(np.uint64), mask)
        m = mult.cpu().numpy().view(np.uint64)
        order = np.argsort(hk)
        hk, m = hk[order], m[order]
        keep = m > 0
        want = oracle.canonical_image(q, hk[keep], m[keep])
        assert np.array_equal(f.export_image(), want)
        # the draws are left as the ascending hashes, each repeated
        d = draws.cpu().numpy().view(np.uint64)
        assert bool(np.all(d[1:] >= d[:-1]))
        assert np.array_equal(np.unique(d), hk[keep])
    finally:
        f.close()


def test_config4_256_kmer_filters(gpu, oracle):
    from cqf_amd import gqf
    from cqf_amd import workloads as W
    t0 = time.time()
    params, keys, offsets = W.kmer_workload(256, seed=5, device="cuda")
    _say(f"case config 4: 256 k-mer filters, {keys.numel()} keys ({time.time() - t0:.1f} s to generate)")
    filters = [gqf.GQF(q) for q, *_ in params]
    group = gqf.GQFGroup(filters)
    try:
        group.autotune(offsets)
        dk = keys.clone()
        group.bulk_insert(dk, offsets)
        got = group.bulk_get(keys, offsets)
        torch.cuda.synchronize()
        _say(f"  group insert + query done ({time.time() - t0:.1f} s); brute-force counts and images")
        for fi, (q, *_r) in enumerate(params):
            a, b = offsets[fi], offsets[fi + 1]
            masked = keys[a:b] & ((1 << (q + 8)) - 1)
            _, inv, cnt = torch.unique(masked, return_inverse=True, return_counts=True)
            assert torch.equal(got[a:b], cnt[inv]), fi
            assert filters[fi].status() == 0 and filters[fi].layout_errors() == 0, fi
            # the segment holds its ascending hashes: their run lengths are
            # the multiset, whose canonical image the filter must hold
            h, c = torch.unique_consecutive(dk[a:b], return_counts=True)
            want = oracle.canonical_image(q, h.cpu().numpy().view(np.uint64),
                                          c.cpu().numpy().astype(np.uint64))
            assert np.array_equal(filters[fi].export_image(), want), fi
        _say(f"  config 4 checked ({time.time() - t0:.1f} s)")
    finally:
        group.close()
        for f in filters:
            f.close()


def test_config5_served_stream_equals_alone(gpu):
    from cqf_amd import serving as S
    from cqf_amd import workloads as W
    t0 = time.time()
    nbits, spec = W.serve_workload(2000, nfilters=64, seed=6, q_lo=22, q_hi=30, median_log2=14, sigma=1.0,
                                   max_log2=22, rate_per_s=2.0e5)
    total = sum(S.filter_bytes(nb, 1 << 22) for nb in nbits.values())
    budget = total // 3  # a third of the filters fit: admission must evict
    _say(f"case config 5: 64 filters q22-30, 2000 requests, budget {budget / 2**30:.1f} of {total / 2**30:.1f} GiB")
    # every third get re-queries its filter's previous insert, so gets hit
    last_ins = {}
    for i, (t, fid, op, n, skip) in enumerate(spec):
        if op == "insert":
            last_ins[fid] = (n, skip)
        elif op == "get" and fid in last_ins and i % 3 == 0:
            spec[i] = (t, fid, op) + last_ins[fid]

    def make():
        return [S.Request(i, fid, op, W.splitmix64(6, n, skip=skip, device="cuda"), t)
                for i, (t, fid, op, n, skip) in enumerate(spec)]

    reqs = make()
    torch.cuda.synchronize()
    srv = S.Server(budget_bytes=budget, n_streams=4, max_batch_keys=1 << 22)
    for fid, nb in nbits.items():
        srv.add_filter(fid, nb)
    srv.run(reqs)
    torch.cuda.synchronize()
    assert srv.evictions > 0
    assert srv.batches < len(reqs)
    assert not any(S.statuses(srv).values())
    srv.close()
    _say(f"  served ({time.time() - t0:.1f} s, {srv.evictions} evictions); replaying one at a time")
    alone = S.replay_alone(nbits, make())
    hits = 0
    for r in reqs:
        if r.op == "get":
            got = r.result.cpu().numpy()
            assert np.array_equal(got, alone[r.rid]), r.rid
            hits += int((got > 0).sum())
    assert hits > 0
    _say(f"  config 5 results equal ({time.time() - t0:.1f} s)")
