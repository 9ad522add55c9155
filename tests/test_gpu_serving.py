"""The serving loop: many filters, continuous batching over several
streams, admission with eviction to host images.  Every query's result
must equal the one from applying the same requests one at a time in
arrival order ("alone").  Every test prints its case before it touches the
GPU."""
import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu


def _say(*a):
    print(*a, flush=True)


def _requests(spec, device="cuda"):
    from cqf_amd import serving as S
    from cqf_amd import workloads as W
    out = []
    for i, (t, fid, op, n, skip) in enumerate(spec):
        keys = W.splitmix64(6, n, skip=skip, device=device)
        out.append(S.Request(i, fid, op, keys, t))
    return out


@pytest.mark.parametrize("budget_filters", [100, 3])
def test_served_results_equal_alone(gpu, budget_filters):
    from cqf_amd import serving as S
    from cqf_amd import workloads as W
    # filters large enough never to fill: a full table drops keys by region,
    # so which ones depends on how the inserts were batched
    nbits, spec = W.serve_workload(700, nfilters=8, q_lo=15, q_hi=17, median_log2=8, sigma=1.5,
                                   max_log2=14, rate_per_s=2.0e4)
    _say(f"case serving 8 filters 700 requests budget={budget_filters} filters")
    # make the gets hit: every third get re-queries the previous insert's keys
    last_ins = {}
    for i, (t, fid, op, n, skip) in enumerate(spec):
        if op == "insert":
            last_ins[fid] = (n, skip)
        elif op == "get" and fid in last_ins and i % 3 == 0:
            spec[i] = (t, fid, op) + last_ins[fid]
    reqs = _requests(spec)
    torch.cuda.synchronize()
    per = max(S.filter_bytes(nb, 1 << 14) for nb in nbits.values())
    srv = S.Server(budget_bytes=budget_filters * per, n_streams=3, max_batch_keys=1 << 14)
    for fid, nb in nbits.items():
        srv.add_filter(fid, nb)
    srv.run(reqs)
    torch.cuda.synchronize()
    if budget_filters == 3:
        assert srv.evictions > 0
    assert srv.batches < len(reqs)  # requests were coalesced
    alone = S.replay_alone(nbits, _requests(spec))
    hits = 0
    for r in reqs:
        if r.op == "get":
            got = r.result.cpu().numpy()
            assert np.array_equal(got, alone[r.rid]), r.rid
            hits += int((got > 0).sum())
    assert hits > 0
    assert not any(S.statuses(srv).values())  # status words and used-slot masks of the resident filters
    srv.close()
    # and the plain per-request path serves the same results
    reqs2 = _requests(spec)
    torch.cuda.synchronize()
    S.run_plain(nbits, reqs2, 1 << 14)
    for r in reqs2:
        if r.op == "get":
            assert np.array_equal(r.result.cpu().numpy(), alone[r.rid]), r.rid
