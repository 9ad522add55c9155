"""Facts of the counter format that the rebuild's straight-line merges rely
on (lean_read for inserts, lean_delete for deletes, gqf_rebuild.hip): they
take a run of at most 8 strictly ascending bytes as count-1 counters, one
remainder per byte.  The format is the reference's encode_counter
(src/gqf.cu:1216-1272), restated in tests/spec_model.py."""
import random

from tests import spec_model as S


def test_counts_above_one_repeat_their_remainder():
    """encode(r, c) for c >= 2 repeats r (a zero remainder repeats 0), so it
    is never strictly ascending.  A run is its counters' encodings in
    ascending remainder order, so a strictly ascending run holds count-1
    counters only."""
    counts = list(range(2, 600)) + [254 ** k + d for k in (1, 2, 3) for d in (-1, 0, 1, 2, 3, 4)]
    for r in range(256):
        for c in counts:
            b = S.encode(r, c)
            assert any(b[i + 1] <= b[i] for i in range(len(b) - 1)), (r, c, b)


def test_delete_from_plain_run_drops_matching_bytes():
    """Removing one instance per item (an absent key's removal is a no-op)
    from a run of count-1 counters leaves the run without the items'
    remainders: the bytes lean_delete keeps."""
    rng = random.Random(7)
    for _ in range(20000):
        rems = sorted(rng.sample(range(256), rng.randint(1, 8)))
        run = [x for r in rems for x in S.encode(r, 1)]
        assert all(run[i + 1] > run[i] for i in range(len(run) - 1))
        items = sorted(rng.choice(rems + [rng.randrange(256)]) for _ in range(rng.randint(1, 2)))
        counts = dict.fromkeys(rems, 1)
        for r in items:
            if counts.get(r, 0):
                counts[r] -= 1
        want = [x for r in sorted(counts) if counts[r] for x in S.encode(r, counts[r])]
        assert [b for b in run if b not in items] == want, (run, items)
