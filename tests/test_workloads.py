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
