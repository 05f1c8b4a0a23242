"""Seeded synthetic sparse tensors for the benchmark and test workloads.

Every generator takes an explicit seed (numpy PCG64) and returns
(ndims, inds, vals) with distinct coordinates (duplicates removed before the
values are drawn), index arrays uint32 and values float32.  Values are drawn
from the open interval (0, 1) where a uniform distribution is asked for,
because the .tns reader drops exact zeros.
"""
from __future__ import annotations

import os

import numpy as np

_HEAD = 1 << 20


def _workers():
    try:
        n = len(os.sched_getaffinity(0))
    except AttributeError:
        n = os.cpu_count() or 1
    return max(1, min(16, n))


class ZipfSampler:
    """Bounded Zipf(s) over ranks 1..n, mapped through a seeded permutation of
    ids [0, n).  Ranks up to 2^20 are sampled exactly from the normalised
    k^-s table; beyond it the rank is drawn by inverting the integral of
    x^-s, which matches the tail mass to within the error of the integral
    approximation (well under 1% per rank there)."""

    def __init__(self, n: int, s: float, rng: np.random.Generator):
        self.n = int(n)
        self.s = float(s)
        head = min(self.n, _HEAD)
        k = np.arange(1, head + 1, dtype=np.float64)
        w = k ** (-self.s)
        head_mass = w.sum()
        if self.n > head:
            a = head + 0.5
            b = self.n + 0.5
            tail_mass = (b ** (1 - self.s) - a ** (1 - self.s)) / (1 - self.s)
        else:
            tail_mass = 0.0
        total = head_mass + tail_mass
        self.p_head = head_mass / total
        self.cdf = np.cumsum(w) / head_mass
        self.head = head
        self.perm = rng.permutation(self.n).astype(np.uint32)

    def ranks(self, size: int, rng: np.random.Generator) -> np.ndarray:
        u = rng.random(size)
        out = np.empty(size, dtype=np.int64)
        in_head = u < self.p_head
        uh = u[in_head] / self.p_head
        out[in_head] = np.minimum(np.searchsorted(self.cdf, uh, side="right"), self.head - 1)
        nt = size - int(in_head.sum())
        if nt:
            ut = rng.random(nt)
            a = self.head + 0.5
            b = self.n + 0.5
            e = 1 - self.s
            x = (a ** e + ut * (b ** e - a ** e)) ** (1 / e)
            out[~in_head] = np.clip(np.floor(x - 0.5).astype(np.int64), self.head, self.n - 1)
        return out

    def sample(self, size: int, rng: np.random.Generator) -> np.ndarray:
        """Draw `size` ids.  Large draws are split into fixed 2^20 chunks, each
        with its own generator seeded from `rng`, and run on a thread pool;
        the result depends only on `rng`'s state, not on the thread count."""
        chunk = 1 << 20
        if size <= chunk:
            return self.perm[self.ranks(size, rng)]
        nch = (size + chunk - 1) // chunk
        seeds = rng.integers(0, 2**63 - 1, size=nch)
        out = np.empty(size, dtype=np.uint32)

        def work(c):
            a = c * chunk
            b = min(size, a + chunk)
            out[a:b] = self.perm[self.ranks(b - a, np.random.default_rng(int(seeds[c])))]

        from concurrent.futures import ThreadPoolExecutor
        with ThreadPoolExecutor(max_workers=_workers()) as ex:
            list(ex.map(work, range(nch)))
        return out


def dedup(inds: list[np.ndarray], sort_fn=None) -> list[np.ndarray]:
    """Sort coordinates lexicographically and drop repeats.  sort_fn(inds)
    may supply a faster lexicographic sort returning the sorted arrays."""
    n = len(inds[0])
    if n == 0:
        return inds
    if sort_fn is not None:
        s = sort_fn(inds)
    else:
        order = np.lexsort(inds[::-1])
        s = [a[order] for a in inds]
    keep = np.ones(n, dtype=bool)
    diff = np.zeros(n - 1, dtype=bool)
    for a in s:
        diff |= a[1:] != a[:-1]
    keep[1:] = diff
    return [a[keep] for a in s]


def unique_coords(ndims, nnz, draw, rng, sort_fn=None, max_rounds=12):
    """Draw coordinates with draw(m, size) until nnz distinct ones exist, then
    keep a seeded random subset of exactly nnz (in lexicographic order)."""
    nm = len(ndims)
    have = [np.zeros(0, dtype=np.uint32) for _ in range(nm)]
    want = int(nnz)
    batch = want
    for _ in range(max_rounds):
        new = [draw(m, batch).astype(np.uint32) for m in range(nm)]
        have = dedup([np.concatenate([h, x]) for h, x in zip(have, new)], sort_fn)
        got = len(have[0])
        if got >= want:
            break
        frac = max(got / max(batch, 1), 0.05)
        batch = int((want - got) / frac * 1.1) + 1024
    got = len(have[0])
    if got > want:
        pick = np.sort(rng.choice(got, size=want, replace=False))
        have = [h[pick] for h in have]
    return have


def open_unit(rng, n):
    v = rng.random(n, dtype=np.float32)
    v[v == 0] = np.float32(0.5)
    return v


def uniform_tensor(ndims, nnz, seed, sort_fn=None):
    """Uniform random coordinates, values U(0,1) float32."""
    rng = np.random.default_rng(seed)
    inds = unique_coords(ndims, nnz, lambda m, k: rng.integers(0, ndims[m], size=k), rng, sort_fn)
    return np.asarray(ndims, dtype=np.uint32), inds, open_unit(rng, len(inds[0]))


def zipf_tensor(ndims, nnz, s, seed, values="lognormal", sort_fn=None):
    """Per-mode bounded Zipf(s) indices over permuted ids, distinct
    coordinates, values lognormal(0,1) (or U(0,1))."""
    rng = np.random.default_rng(seed)
    samplers = [ZipfSampler(d, s, rng) for d in ndims]
    inds = unique_coords(ndims, nnz, lambda m, k: samplers[m].sample(k, rng), rng, sort_fn)
    n = len(inds[0])
    if values == "lognormal":
        vals = rng.lognormal(0.0, 1.0, size=n).astype(np.float32)
    else:
        vals = open_unit(rng, n)
    return np.asarray(ndims, dtype=np.uint32), inds, vals


def planted_values(inds, factors, noise, rng, chunk=1 << 20):
    """v = sum_r prod_m U_m[i_m][r] plus Gaussian noise of std noise*rms."""
    n = len(inds[0])
    out = np.empty(n, dtype=np.float64)
    for a in range(0, n, chunk):
        b = min(n, a + chunk)
        p = np.ones((b - a, factors[0].shape[1]), dtype=np.float64)
        for m, U in enumerate(factors):
            p *= U[inds[m][a:b]]
        out[a:b] = p.sum(axis=1)
    rms = float(np.sqrt(np.mean(out * out))) if n else 0.0
    out += rng.normal(0.0, noise * rms, size=n)
    return out.astype(np.float32), rms


def planted_tensor(ndims, nnz, rank, s, seed, noise=0.01, sort_fn=None):
    """Zipf(s) coordinates with values from a planted U[0,1) rank-`rank`
    Kruskal model plus noise*rms Gaussian noise.  Returns
    (ndims, inds, vals, planted_factors)."""
    rng = np.random.default_rng(seed)
    samplers = [ZipfSampler(d, s, rng) for d in ndims]
    inds = unique_coords(ndims, nnz, lambda m, k: samplers[m].sample(k, rng), rng, sort_fn)
    factors = [rng.random((d, rank)) for d in ndims]
    vals, _ = planted_values(inds, factors, noise, rng)
    return np.asarray(ndims, dtype=np.uint32), inds, vals, factors


def uniform_factors(ndims, rank, seed):
    """Factor matrices U[0,1) float32, one per mode."""
    rng = np.random.default_rng(seed)
    return [rng.random((int(d), rank), dtype=np.float32) for d in ndims]
