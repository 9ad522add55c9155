"""Sequenced batches (qf_bulk_mutate_seq + qf_bulk_get_seq_adjust) against
the oracle applying the same requests one at a time in order: the 81-byte
image after the mutation and every query's count must be bit-exact.  Keys
come from a small universe so hashes repeat across requests, deletes hit
absent and already-deleted keys (the count floors at 0) and inserts follow
deletes of the same key.  Every test prints its case before it touches the
GPU."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _say(*a):
    print(*a, flush=True)


def _requests(oracle, rng, univ, nreq, max_size, p_get, p_del):
    """[(op, keys)] with ops drawn per request; sizes 0..max_size."""
    out = []
    for _ in range(nreq):
        u = rng.random()
        op = "get" if u < p_get else ("delete" if u < p_get + p_del else "insert")
        k = int(rng.integers(0, max_size + 1))
        out.append((op, univ[rng.integers(0, univ.size, size=k)].astype(np.uint64)))
    return out


def _run_seq(g, reqs, dev):
    """One sequenced batch on the GPU; returns {request index: counts}."""
    import torch
    from cqf_amd.gqf import to_device_u64
    gets = [(i, k) for i, (op, k) in enumerate(reqs) if op == "get"]
    muts = [(i, op, k) for i, (op, k) in enumerate(reqs) if op != "get"]
    out = None
    if gets:
        gk = to_device_u64(np.concatenate([k for _, k in gets]))
        g_ends = np.cumsum([k.size for _, k in gets]).tolist()
        out = g.bulk_get(gk)
    if muts:
        mk = to_device_u64(np.concatenate([k for _, _, k in muts]))
        m_ends = np.cumsum([k.size for _, _, k in muts]).tolist()
        g.mutate_seq(mk, m_ends, [(i << 1) | (op == "delete") for i, op, _ in muts])
        if gets:
            g.get_seq_adjust(gk, g_ends, [i for i, _ in gets], out, mk)
        torch.cuda.synchronize()
        # the keys are left sorted by hash, requests in order within a hash
        s = mk.cpu().numpy().view(np.uint64)
        h = s & np.uint64((1 << 48) - 1)
        assert (np.diff(h.astype(np.int64)) >= 0).all()
        same = h[1:] == h[:-1]
        seq = (s >> np.uint64(48)) & np.uint64(0x7FFF)
        assert (seq[1:][same] >= seq[:-1][same]).all()
    res = {}
    if gets:
        c = out.cpu().numpy().view(np.uint64)
        o = 0
        for i, k in gets:
            res[i] = c[o:o + k.size]
            o += k.size
    return res


def _run_oracle(o, reqs):
    res = {}
    for i, (op, k) in enumerate(reqs):
        if op == "get":
            res[i] = o.bulk_get(k)
        elif op == "insert":
            o.bulk_insert(k.copy())
        else:
            o.bulk_delete(k.copy())
    return res


@pytest.mark.parametrize("nbits,prefill,nreq,max_size,univ_n,seed", [
    (12, 0.3, 40, 60, 300, 1),        # dense repeats, small table
    (16, 0.85, 120, 400, 5000, 2),    # loaded table, mixed sizes
    (13, 0.0, 8, 3000, 2500, 3),      # regions beyond the LDS item cache
    (20, 0.9, 256, 200, 40000, 4),    # the request limit, sparse batch on a full table
    (14, 0.5, 1, 500, 800, 5),        # a single request
])
def test_seq_batch_equals_one_by_one(gpu, oracle, nbits, prefill, nreq, max_size, univ_n, seed):
    from cqf_amd.gqf import GQF, to_device_u64
    rng = np.random.default_rng(seed)
    univ = oracle.splitmix64(100 + seed, univ_n)
    reqs = _requests(oracle, rng, univ, nreq, max_size, p_get=0.4, p_del=0.3)
    if nreq == 1:
        reqs = [("insert", reqs[0][1])]
    if len(reqs) > 2:
        reqs[len(reqs) // 2] = ("delete", np.zeros(0, dtype=np.uint64))  # an empty request
    _say(f"case seq q={nbits} prefill={prefill} requests={len(reqs)} "
         f"keys={sum(k.size for _, k in reqs)} seed={seed}")
    g, o = GQF(nbits), oracle.OracleQF(nbits)
    try:
        pre = oracle.splitmix64(200 + seed, int(prefill * (1 << nbits)))
        if pre.size:
            pre = np.concatenate([pre, univ[: univ.size // 3]])
            g.bulk_insert(to_device_u64(pre))
            o.bulk_insert(pre.copy())
        got = _run_seq(g, reqs, gpu)
        want = _run_oracle(o, reqs)
        assert g.status() == 0 and g.layout_errors() == 0
        img_g, img_o = g.export_image(), o.image()
        assert np.array_equal(img_g, img_o), f"{np.count_nonzero(img_g != img_o)} image bytes differ"
        for i in want:
            assert np.array_equal(got[i], want[i]), f"request {i}"
        # a second batch on the result, deletes only
        reqs2 = _requests(oracle, rng, univ, 30, max_size, p_get=0.3, p_del=0.7)
        got2 = _run_seq(g, reqs2, gpu)
        want2 = _run_oracle(o, reqs2)
        assert np.array_equal(g.export_image(), o.image())
        for i in want2:
            assert np.array_equal(got2[i], want2[i]), f"second batch request {i}"
    finally:
        g.close()
        o.close()


def test_seq_rejects_bad_requests(gpu):
    from cqf_amd.gqf import GQF, to_device_u64
    _say("case seq argument checks")
    g = GQF(12)
    try:
        k = to_device_u64(np.arange(10, dtype=np.uint64))
        with pytest.raises(ValueError):
            g.mutate_seq(k, [5, 9], [0, 2])          # ends short of n
        with pytest.raises(ValueError):
            g.mutate_seq(k, [5, 10], [2, 2])         # sequence numbers not ascending
        with pytest.raises(ValueError):
            g.mutate_seq(k, [6, 5, 10], [0, 2, 4])   # ends descending
        with pytest.raises(ValueError):
            g.mutate_seq(k, list(range(1, 258)), [2 * i for i in range(257)])  # too many requests
        assert not g.export_image().any()  # nothing was applied
    finally:
        g.close()


def test_seq_batch_with_value_bits(gpu, oracle):
    """A filter with 2 value bits (the layout a resize leaves): slot hash =
    hash_64(key under 2^key_bits) << 2.  The expected counts apply the
    requests one at a time to a {slot hash: count} map (deletes floor at 0);
    the expected image is the canonical layout of the final map."""
    from cqf_amd.gqf import GQF, to_device_u64
    nbits, vb = 14, 2
    key_bits = nbits + 8 - vb
    _say(f"case seq with value bits q={nbits} vb={vb}")
    rng = np.random.default_rng(9)
    univ = oracle.splitmix64(109, 3000)
    reqs = _requests(oracle, rng, univ, 60, 300, p_get=0.4, p_del=0.3)
    slot = lambda k: oracle.hash_many(k, (1 << key_bits) - 1) << np.uint64(vb)
    counts, want = {}, {}
    for i, (op, k) in enumerate(reqs):
        hs = slot(k)
        if op == "get":
            want[i] = np.array([counts.get(int(h), 0) for h in hs], dtype=np.uint64)
        else:
            for h in hs.tolist():
                c = counts.get(h, 0)
                counts[h] = c + 1 if op == "insert" else max(0, c - 1)
    g = GQF(nbits, value_bits=vb)
    try:
        assert g.info["value_bits"] == vb
        got = _run_seq(g, reqs, gpu)
        assert g.status() == 0 and g.layout_errors() == 0
        for i in want:
            assert np.array_equal(got[i], want[i]), f"request {i}"
        live = {h: c for h, c in counts.items() if c}
        hs = np.array(sorted(live), dtype=np.uint64)
        cs = np.array([live[int(h)] for h in hs], dtype=np.uint64)
        assert np.array_equal(g.export_image(), oracle.canonical_image(nbits, hs, cs))
    finally:
        g.close()
