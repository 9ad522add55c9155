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
