"""The benchmark key streams (cqf_amd/workloads.py) on the CPU: splitmix64
against the oracle's C implementation, distinctness of the Zipf ids in the
bits the filter hashes, and the sampled distribution against its closed
form."""
import numpy as np
import torch

from cqf_amd import workloads as W


def test_splitmix64_matches_oracle(oracle):
    for seed, skip in [(0, 0), (3, 0), (3, 241591910), (7, 12345), (2 ** 63 + 5, 99)]:
        a = W.splitmix64(seed, 5000, skip=skip).numpy().view(np.uint64)
        b = oracle.splitmix64(seed % (1 << 64), 5000, skip=skip)
        assert np.array_equal(a, b), (seed, skip)


def test_distinct_keys_low_bits():
    k = W.distinct_keys(4, 50000, 20)  # 20 bits: collisions are common here
    low = (k & ((1 << 20) - 1)).numpy()
    assert np.unique(low).size == k.numel() == 50000
    # stream order kept: the first output always survives
    assert k[0] == W.splitmix64(4, 1)[0]


def test_zipf_workload_distribution():
    n_ids, n_draws, s = 1 << 12, 1 << 20, 1.5
    keys, draws, mult = W.zipf_workload(n_ids, n_draws, s=s, seed=4, distinct_bits=36)
    assert int(mult.sum()) == n_draws
    assert torch.equal(torch.sort(draws).values, torch.sort(torch.repeat_interleave(keys, mult)).values)
    h = sum(k ** -s for k in range(1, n_ids + 1))
    p1 = 1.0 / h
    sd = (n_draws * p1 * (1 - p1)) ** 0.5
    assert abs(int(mult[0]) - n_draws * p1) < 5 * sd
    # reproducible
    _, d2, m2 = W.zipf_workload(n_ids, n_draws, s=s, seed=4, distinct_bits=36)
    assert torch.equal(draws, d2) and torch.equal(mult, m2)


def test_kmers_are_canonical_31mers():
    km = W.kmers(20000, 5000, 11)
    assert km.numel() == 20000 and int(km.min()) >= 0 and int(km.max()) < (1 << 62)

    def dec(v):
        return [(v >> (2 * (30 - i))) & 3 for i in range(31)]

    def enc(bb):
        return sum(x << (2 * (30 - i)) for i, x in enumerate(bb))

    for v in km[:300].tolist():
        b = dec(v)
        assert v == min(enc(b), enc([3 - x for x in b[::-1]]))
    # consecutive k-mers of one read overlap by 30 bases (one strand or the other)
    a, b = dec(int(km[0])), dec(int(km[1]))
    rc = lambda x: [3 - y for y in x[::-1]]
    assert any(p[1:] == q[:-1] for p in (a, rc(a)) for q in (b, rc(b)))
    # 4x coverage of a 5000-base genome: far fewer distinct k-mers than instances
    assert torch.unique(km).numel() < 0.7 * km.numel()


def test_kmer_params_ranges():
    p = W.kmer_filter_params()
    assert len(p) == 256
    assert all(16 <= q <= 24 and 0.5 <= ld <= 0.95 and n == int(ld * (1 << q)) for q, ld, n, _ in p)
    assert len({q for q, *_ in p}) == 9
