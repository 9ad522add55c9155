"""Pure-Python model of the CQF image, written from the format's definition.

Independent of oracle/cqf_oracle.c: it knows nothing of insert1 / shifting;
it builds the image a multiset of hashes must have from the invariants the
reference's readers rely on (PASSIONLab/cqf):
  - block layout: offset u8, occupieds u64 LE, runends u64 LE, 64 slot bytes
    (include/gqf_int.cuh:38-57);
  - counter format (the comment at src/gqf.cu:1207-1215):
        1: x   2: x x   3: x 0 x (x > 0) / 0 0 0   >3: x [0] c..c x / 0 c..c 0 0
    with digits c in base 2^8-2 (x > 0; values 1..255 skipping x) or 2^8-1
    (x = 0; values 1..255), most significant first, and the leading 0 when
    the top digit is >= x;
  - the run of quotient q starts at max(q, end of previous run + 1) and holds
    its remainders' counters in ascending remainder order;
  - block i's offset is run_end(64 i - 1) - 64 i + 1 floored at 0, stored
    saturated at 255.
Used by the CPU tests to pin the oracle; slow, so only for small tables.
"""


def encode(rem, count):
    """Slots of the counter (rem, count), in slot order, from the format."""
    if count == 0:
        return []
    if count == 1:
        return [rem]
    if count == 2:
        return [rem, rem]
    if count == 3:
        return [rem, 0, rem] if rem else [0, 0, 0]
    if rem:
        base, c = 254, count - 3
    else:
        base, c = 255, count - 4
    digits = []
    while True:
        d = c % base + 1
        if rem and d >= rem:
            d += 1
        digits.append(d)
        c //= base
        if c == 0:
            break
    digits.reverse()  # most significant first
    if rem:
        lead = [0] if digits[0] >= rem else []
        return [rem] + lead + digits + [rem]
    return [0] + digits + [0, 0]


def sizes(nbits):
    nslots = 1 << nbits
    # the reference computes nslots + 10*sqrt(nslots) in double precision and
    # truncates (src/gqf.cu:1950); for nslots = 2^k the product is exact
    # enough that isqrt of 100*nslots reproduces it
    import math
    xnslots = nslots + math.isqrt(100 * nslots)
    nblocks = (xnslots + 63) // 64
    return nslots, xnslots, nblocks


def build_image(nbits, multiset):
    """multiset: dict hash -> count (hash < 2^(nbits+8)).  Returns bytes."""
    return image_from_runs(nbits, runs_of(multiset))


def runs_of(multiset):
    """{quotient: slot bytes of its run}: the counters of its remainders in
    ascending remainder order."""
    by_q = {}
    for h, c in multiset.items():
        if c > 0:
            by_q.setdefault(h >> 8, []).append((h & 0xFF, c))
    out = {}
    for q, rcs in by_q.items():
        run = []
        for rem, c in sorted(rcs):
            run += encode(rem, c)
        out[q] = run
    return out


def image_from_runs(nbits, runs):
    """The image whose quotient q holds the slot bytes runs[q] (any bytes:
    the layout does not read them), each run at max(q, end of the previous
    run + 1).  Returns bytes."""
    nslots, xnslots, nblocks = sizes(nbits)
    slots = [0] * (nblocks * 64)
    runend = [0] * (nblocks * 64)
    occ = [0] * (nblocks * 64)
    prev_end = -1
    ends = []  # (quotient, end) in quotient order
    for q in sorted(q for q, r in runs.items() if r):
        run = list(runs[q])
        start = max(q, prev_end + 1)
        end = start + len(run) - 1
        if end >= xnslots:
            raise ValueError("does not fit")
        slots[start:end + 1] = run
        runend[end] = 1
        occ[q] = 1
        ends.append((q, end))
        prev_end = end
    img = bytearray(nblocks * 81)
    k = 0
    last_end = -1
    for b in range(nblocks):
        # end of the runs with quotient <= 64 b - 1
        while k < len(ends) and ends[k][0] <= 64 * b - 1:
            last_end = ends[k][1]
            k += 1
        off = max(0, last_end - (64 * b - 1)) if b > 0 else 0
        base = b * 81
        img[base] = min(off, 255)
        o = sum(occ[64 * b + i] << i for i in range(64))
        r = sum(runend[64 * b + i] << i for i in range(64))
        img[base + 1:base + 9] = o.to_bytes(8, "little")
        img[base + 9:base + 17] = r.to_bytes(8, "little")
        img[base + 17:base + 81] = bytes(slots[64 * b:64 * b + 64])
    return bytes(img)


def insert_not_exists(runs, h):
    """A not-exists insert of hash h on {quotient: run bytes}, searched slot
    by slot as the format's readers would NOT (they walk counters): skip the
    run's slots while the slot is below the remainder; past the run end,
    append; at a larger slot, insert the remainder before it; at an equal
    slot, report it found and change nothing.  A quotient without a run
    gets the run [remainder].  Returns "inserted" or "found"."""
    q, r = h >> 8, h & 0xFF
    run = runs.get(q)
    if not run:
        runs[q] = [r]
        return "inserted"
    i = 0
    while i < len(run) and run[i] < r:
        i += 1
    if i < len(run) and run[i] == r:
        return "found"
    run.insert(i, r)
    return "inserted"


def used_slots(nbits, multiset):
    """0/1 per slot of the table (nblocks * 64): 1 where a run of the layout
    the multiset fixes lies (runs placed as in build_image)."""
    nslots, xnslots, nblocks = sizes(nbits)
    used = [0] * (nblocks * 64)
    by_q = {}
    for h, c in multiset.items():
        if c > 0:
            by_q.setdefault(h >> 8, []).append((h & 0xFF, c))
    prev_end = -1
    for q in sorted(by_q):
        n = sum(len(encode(rem, c)) for rem, c in by_q[q])
        start = max(q, prev_end + 1)
        for s in range(start, start + n):
            used[s] = 1
        prev_end = start + n - 1
    return used


def clusters(used):
    """Lengths of the maximal stretches of 1s, left to right."""
    out, run = [], 0
    for u in used:
        if u:
            run += 1
        elif run:
            out.append(run)
            run = 0
    if run:
        out.append(run)
    return out


def hash_64(key, mask):
    """Thomas Wang's 64-bit mix written with multiplications
    ((key << 21) - key - 1, *265, *21, *(2^31 + 1)), not the reference's
    shift-add form (src/hashutil.cu:132-142)."""
    M = (1 << 64) - 1
    key &= M
    key = (((key << 21) - key - 1) & M) & mask
    key ^= key >> 24
    key = (key * 265) & M & mask
    key ^= key >> 14
    key = (key * 21) & M & mask
    key ^= key >> 28
    key = (key * ((1 << 31) + 1)) & M & mask
    return key
