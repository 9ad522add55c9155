"""Synthetic key streams of the benchmark configurations, in torch (CPU or
GPU), reproducible from their seeds.

splitmix64(seed): state = seed; each output adds the golden gamma
0x9E3779B97F4A7C15 to the state and mixes it with two xor-shift-multiply
rounds (Steele, Lea & Flood 2014).  Output i (0-based) is the mix of
seed + (i + 1) * gamma, so any range of the stream is computed directly.

zipf_workload: n_draws samples of Zipf(s) over n_ids ids by inverse CDF
(ids ranked 1..n_ids with weight k^-s), each id mapped to a distinct uint64
key -- distinct in its low `distinct_bits` bits, so that a filter whose
hash sees only key & (2^(q+8) - 1) counts every id separately.
"""
import torch

GAMMA = 0x9E3779B97F4A7C15
M1 = 0xBF58476D1CE4E5B9
M2 = 0x94D049BB133111EB


def _i64(c):
    """uint64 constant -> the int64 with the same bits."""
    return c - (1 << 64) if c >= (1 << 63) else c


def _srl(x, k):
    """logical right shift of int64 bits"""
    return (x >> k) & ((1 << (64 - k)) - 1)


def splitmix64(seed, n, skip=0, device="cpu"):
    """Outputs skip .. skip+n-1 of splitmix64(seed), as int64 tensors holding
    the uint64 bits (torch int64 arithmetic wraps modulo 2^64)."""
    i = torch.arange(skip + 1, skip + n + 1, dtype=torch.int64, device=device)
    z = _i64(seed % (1 << 64)) + i * _i64(GAMMA)
    z = (z ^ _srl(z, 30)) * _i64(M1)
    z = (z ^ _srl(z, 27)) * _i64(M2)
    return z ^ _srl(z, 31)


def distinct_keys(seed, n, distinct_bits, device="cpu"):
    """The first n outputs of splitmix64(seed) whose low `distinct_bits`
    bits differ from those of every earlier output (stream order kept)."""
    mask = (1 << distinct_bits) - 1
    extra = max(1024, int(4 * n * n / float(1 << distinct_bits)) + 1024)
    while True:
        x = splitmix64(seed, n + extra, device=device)
        low = x & mask
        _, perm = torch.sort(low, stable=True)
        ls = low[perm]
        dup = torch.zeros_like(x, dtype=torch.bool)
        dup[perm[1:][ls[1:] == ls[:-1]]] = True
        keep = x[~dup]
        if keep.numel() >= n:
            return keep[:n].contiguous()
        extra *= 2


def zipf_workload(n_ids, n_draws, s=1.5, seed=4, distinct_bits=36, device="cpu", chunk=1 << 24):
    """(keys[n_ids], draws[n_draws], multiplicity[n_ids]).

    Uniforms are u = (splitmix64(seed) output (2^40 + i) >> 11 + 0.5) / 2^53;
    the draw is the smallest id whose normalised CDF is >= u."""
    keys = distinct_keys(seed, n_ids, distinct_bits, device=device)
    w = torch.arange(1, n_ids + 1, dtype=torch.float64, device=device).pow(-s)
    cdf = torch.cumsum(w, 0)
    cdf /= cdf[-1].clone()
    draws = torch.empty(n_draws, dtype=torch.int64, device=device)
    mult = torch.zeros(n_ids, dtype=torch.int64, device=device)
    for b in range(0, n_draws, chunk):
        m = min(chunk, n_draws - b)
        r = splitmix64(seed, m, skip=(1 << 40) + b, device=device)
        u = (_srl(r, 11).to(torch.float64) + 0.5) * (2.0 ** -53)
        ids = torch.searchsorted(cdf, u).clamp_(max=n_ids - 1)
        draws[b:b + m] = keys[ids]
        mult += torch.bincount(ids, minlength=n_ids)
    return keys, draws, mult


def _bits(seed, n, skip, device):
    return splitmix64(seed, n, skip=skip, device=device)


def kmer_filter_params(nfilters=256, seed=5, q_lo=16, q_hi=24, load_lo=0.5, load_hi=0.95):
    """Per filter: (log2 slots, load factor, k-mer instances, genome length).
    q is uniform in [q_lo, q_hi], the load uniform in [load_lo, load_hi]
    (splitmix64(seed) outputs 2f and 2f + 1); the genome is a quarter of
    the k-mer instances long (about 4x coverage), at least 1000 bases."""
    u = splitmix64(seed, 2 * nfilters).tolist()
    out = []
    for f in range(nfilters):
        a = u[2 * f] % (1 << 64)
        b = u[2 * f + 1] % (1 << 64)
        q = q_lo + a % (q_hi - q_lo + 1)
        load = load_lo + (load_hi - load_lo) * ((b >> 11) / float(1 << 53))
        n = int(load * (1 << q))
        out.append((q, load, n, max(1000, n // 4)))
    return out


def kmers(n, genome_len, seed, read_len=150, k=31, err_per_100=1, device="cpu", chunk_reads=1 << 16):
    """n canonical k-mers (2-bit packed, min of forward and reverse
    complement, 62 bits) from reads of read_len bases sampled uniformly from
    a random genome of genome_len bases, each base substituted with
    probability err_per_100 / 100.  All randomness is splitmix64(seed):
    genome bases from outputs [0, ...), read starts from [2^40, ...), error
    draws from [2^41, ...)."""
    per_read = read_len - k + 1
    n_reads = (n + per_read - 1) // per_read
    words = (genome_len + 31) // 32
    g = _bits(seed, words, 0, device)
    shifts = torch.arange(0, 64, 2, device=device)
    genome = (g[:, None] >> shifts) & 3  # the sign fill of >> never reaches the 2 kept bits
    genome = genome.reshape(-1)[:genome_len].to(torch.int64)
    w_fwd = torch.tensor([4 ** (k - 1 - i) for i in range(k)], dtype=torch.int64, device=device)
    w_rev = torch.tensor([4 ** i for i in range(k)], dtype=torch.int64, device=device)
    out = torch.empty(n_reads * per_read, dtype=torch.int64, device=device)
    pos = torch.arange(read_len, device=device)
    for r0 in range(0, n_reads, chunk_reads):
        m = min(chunk_reads, n_reads - r0)
        st = _srl(_bits(seed, m, (1 << 40) + r0, device), 1) % (genome_len - read_len + 1)
        reads = genome[st[:, None] + pos]
        e = _srl(_bits(seed, m * read_len, (1 << 41) + r0 * read_len, device), 1).reshape(m, read_len)
        hit = (e % 100) < err_per_100
        sub = (reads + 1 + (e >> 8) % 3) % 4
        reads = torch.where(hit, sub, reads)
        win = reads.unfold(1, k, 1)                      # m x per_read x k
        fwd = (win * w_fwd).sum(-1)
        rev = ((3 - win) * w_rev).sum(-1)
        out[r0 * per_read:(r0 + m) * per_read] = torch.minimum(fwd, rev).reshape(-1)
    return out[:n].contiguous()


def kmer_workload(nfilters=256, seed=5, device="cpu", **kw):
    """(params, keys, offsets): every filter's k-mer instances, concatenated
    in filter order; filter f draws from splitmix64(seed * 1000003 + f)."""
    params = kmer_filter_params(nfilters, seed, **kw)
    parts, offsets = [], [0]
    for f, (q, load, n, glen) in enumerate(params):
        parts.append(kmers(n, glen, seed * 1000003 + f, device=device))
        offsets.append(offsets[-1] + n)
    return params, torch.cat(parts), offsets


def serve_workload(n_requests, nfilters=64, seed=6, q_lo=22, q_hi=30, median_log2=14, sigma=1.0,
                   max_log2=22, rate_per_s=1.0e5):
    """A served request stream.  Filter f has 2^(q_lo + a_f % (q_hi - q_lo + 1))
    slots; request i picks a filter uniformly, an operation 60% get / 30%
    insert / 10% delete, a size lognormal(ln 2^median_log2, sigma) clipped
    to [1, 2^max_log2], and arrives after an exponential gap of mean
    1/rate_per_s.  All draws come from splitmix64(seed) (filter sizes from
    outputs [0, nfilters), request draws from [2^32, ...)); request i's keys
    are outputs [2^40 + offset_i, ...) of the same stream, offsets running.
    Returns ({fid: log2 slots}, [(t_arrival, fid, op, n_keys, key_skip)])."""
    import math
    a = splitmix64(seed, nfilters).tolist()
    nbits = {f: q_lo + (a[f] % (1 << 64)) % (q_hi - q_lo + 1) for f in range(nfilters)}
    u = splitmix64(seed, 5 * n_requests, skip=1 << 32).tolist()
    unit = lambda x: ((x % (1 << 64)) >> 11) * (2.0 ** -53) + 2.0 ** -54
    reqs, t, off = [], 0.0, 0
    for i in range(n_requests):
        u0, u1, u2, u3, u4 = (unit(x) for x in u[5 * i:5 * i + 5])
        t += -math.log(u0) / rate_per_s
        fid = min(nfilters - 1, int(u1 * nfilters))
        op = "get" if u2 < 0.6 else ("insert" if u2 < 0.9 else "delete")
        z = math.sqrt(-2.0 * math.log(u3)) * math.cos(2.0 * math.pi * u4)
        n = int(round(math.exp(median_log2 * math.log(2.0) + sigma * z)))
        n = max(1, min(1 << max_log2, n))
        reqs.append((t, fid, op, n, (1 << 40) + off))
        off += n
    return nbits, reqs
