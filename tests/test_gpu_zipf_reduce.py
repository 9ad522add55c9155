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
        assert g.status() == 0 and g.layout_errors() == 0
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
        assert g.status() == 0 and g.layout_errors() == 0
    finally:
        g.close()


def _path_runs(gqf):
    return gqf.profile_read("reduce_count")[1], gqf.profile_read("reduce_finish")[1]


def test_reduce_counting_path_matches_oracle(gpu, oracle):
    """A duplicate-heavy batch of 2^22 draws takes the counting reduce
    (hash table, sort of the distinct hashes, expansion): the image, the
    keys left sorted and every multiplicity must be the oracle's, and a
    second batch into the same filter must take it again (the table is
    left empty)."""
    from cqf_amd import gqf
    from cqf_amd import workloads as W
    from cqf_amd.gqf import GQF, to_numpy_u64
    nbits, n_ids, n_draws = 24, 1 << 18, 1 << 22
    _say(f"case counting reduce q={nbits} ids={n_ids} draws={n_draws} x2")
    keys, draws, mult = W.zipf_workload(n_ids, n_draws, s=1.5, seed=11, distinct_bits=nbits + 8)
    g = GQF(nbits)
    o = oracle.OracleQF(nbits)
    try:
        for rep in range(2):
            gqf.profile_enable(True)
            dk = draws.clone().cuda()
            g.bulk_insert_reduce(dk)
            torch.cuda.synchronize()
            assert _path_runs(gqf) == (1, 1), rep
            gqf.profile_enable(False)
            ok = draws.numpy().view(np.uint64).copy()
            o.bulk_insert_reduce(ok)
            assert np.array_equal(to_numpy_u64(dk), ok), f"keys left by batch {rep}"
            assert np.array_equal(g.export_image(), o.image()), f"image after batch {rep}"
        got = to_numpy_u64(g.bulk_get(keys.cuda())).astype(np.int64)
        assert np.array_equal(got, 2 * mult.numpy())
        assert g.status() == 0 and g.layout_errors() == 0
    finally:
        gqf.profile_enable(False)
        g.close()
        o.close()


def test_reduce_counting_falls_back_when_table_overflows(gpu, oracle):
    """60% of 2^23 keys are one key and 40% distinct: the sample admits the
    counting reduce, but 3.4M distinct hashes overflow its table, so the
    call falls back to the sort (same image and keys as the oracle); a
    Zipf batch into the same filter then takes the counting path again."""
    from cqf_amd import gqf
    from cqf_amd import workloads as W
    from cqf_amd.gqf import GQF, to_device_u64, to_numpy_u64
    nbits, n = 24, 1 << 23
    _say(f"case counting reduce fallback q={nbits} keys={n}")
    rng = np.random.default_rng(13)
    ks = oracle.splitmix64(31, n)
    ks[rng.random(n) < 0.6] = np.uint64(12345)
    g = GQF(nbits)
    o = oracle.OracleQF(nbits)
    try:
        gqf.profile_enable(True)
        dk = to_device_u64(ks)
        g.bulk_insert_reduce(dk)
        torch.cuda.synchronize()
        assert _path_runs(gqf) == (1, 0)
        gqf.profile_enable(False)
        ok = ks.copy()
        o.bulk_insert_reduce(ok)
        assert np.array_equal(to_numpy_u64(dk), ok)
        assert np.array_equal(g.export_image(), o.image())
        _, draws, _ = W.zipf_workload(1 << 16, 1 << 22, s=1.5, seed=12, distinct_bits=nbits + 8)
        gqf.profile_enable(True)
        dk = draws.clone().cuda()
        g.bulk_insert_reduce(dk)
        torch.cuda.synchronize()
        assert _path_runs(gqf) == (1, 1)
        gqf.profile_enable(False)
        ok = draws.numpy().view(np.uint64).copy()
        o.bulk_insert_reduce(ok)
        assert np.array_equal(to_numpy_u64(dk), ok)
        assert np.array_equal(g.export_image(), o.image())
        assert g.status() == 0 and g.layout_errors() == 0
    finally:
        gqf.profile_enable(False)
        g.close()
        o.close()


def test_reduce_of_distinct_keys_keeps_the_sort(gpu, oracle):
    """A batch of 2^22 distinct keys fails the sample check: no counting
    kernel runs and the sort path gives the oracle's image."""
    from cqf_amd import gqf
    from cqf_amd.gqf import GQF, to_device_u64, to_numpy_u64
    nbits, n = 23, 1 << 22
    _say(f"case reduce of distinct keys q={nbits} keys={n}")
    ks = oracle.splitmix64(41, n)
    g = GQF(nbits)
    o = oracle.OracleQF(nbits)
    try:
        gqf.profile_enable(True)
        dk = to_device_u64(ks)
        g.bulk_insert_reduce(dk)
        torch.cuda.synchronize()
        assert _path_runs(gqf) == (0, 0)
        gqf.profile_enable(False)
        ok = ks.copy()
        o.bulk_insert_reduce(ok)
        assert np.array_equal(to_numpy_u64(dk), ok)
        assert np.array_equal(g.export_image(), o.image())
    finally:
        gqf.profile_enable(False)
        g.close()
        o.close()
