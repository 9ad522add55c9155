"""Skewed counting through bulk_insert_reduce (the Zipf configuration at
reduced scale): bulk_get must return the exact multiplicity of every id
(ids are distinct in the q + 8 bits the hash sees), and the image must be
the oracle's.  Every test prints its case before it touches the GPU."""
import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu


def _say(*a):
    print(*a, flush=True)


@pytest.mark.parametrize("batches", [1, 4])
def test_zipf_reduce_exact_and_image(gpu, oracle, batches):
    from cqf_amd import workloads as W
    from cqf_amd.gqf import GQF, to_numpy_u64
    nbits, n_ids, n_draws = 20, 1 << 14, 1 << 20
    _say(f"case zipf reduce q={nbits} ids={n_ids} draws={n_draws} batches={batches}")
    keys, draws, mult = W.zipf_workload(n_ids, n_draws, s=1.5, seed=4, distinct_bits=nbits + 8)
    g = GQF(nbits)
    o = oracle.OracleQF(nbits)
    try:
        for part in torch.tensor_split(draws, batches):
            g.bulk_insert_reduce(part.clone().cuda())
            o.bulk_insert_reduce(part.numpy().view(np.uint64).copy())
        got = to_numpy_u64(g.bulk_get(keys.cuda())).astype(np.int64)
        assert np.array_equal(got, mult.numpy())
        assert int(mult[0]) > 100000  # multi-digit counters in play
        assert np.array_equal(g.export_image(), o.image())
        assert g.status() == 0
    finally:
        g.close()
        o.close()


def test_zipf_reduce_exact_q24(gpu):
    from cqf_amd import workloads as W
    from cqf_amd.gqf import GQF
    nbits, n_ids, n_draws = 24, 1 << 20, 1 << 24
    _say(f"case zipf reduce q={nbits} ids={n_ids} draws={n_draws}")
    keys, draws, mult = W.zipf_workload(n_ids, n_draws, s=1.5, seed=4, distinct_bits=nbits + 8, device="cuda")
    g = GQF(nbits)
    try:
        g.bulk_insert_reduce(draws.clone())
        got = g.bulk_get(keys)
        assert torch.equal(got, mult)
        assert g.status() == 0
    finally:
        g.close()
