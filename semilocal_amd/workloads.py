"""Seeded synthetic workloads (numpy PCG64 streams; the input SHA-256 is
recorded next to every golden vector so that any drift is caught).

Shapes:
  * plain pair: a, b i.i.d. uniform over an alphabet of size sigma, uint8;
  * binary pair: uniform bits;
  * similar pair: b = mutate(a): a fraction of positions substituted by a
    different symbol, single-symbol insertions/deletions at uniform positions
    (half insertions, half deletions);
  * ragged batch: lengths round(LogNormal(ln median, sigma_ln)) clipped, a and
    b independent, half of the pairs mutated copies;
  * serving stream: Poisson arrivals, log-normal lengths, a share of follow-up
    window queries on earlier comparisons.
"""
from __future__ import annotations

import hashlib
from dataclasses import dataclass, field

import numpy as np


def rng(seed: int) -> np.random.Generator:
    return np.random.Generator(np.random.PCG64(seed))


def sha256(*arrays) -> str:
    h = hashlib.sha256()
    for a in arrays:
        a = np.ascontiguousarray(a)
        h.update(str(a.dtype).encode())
        h.update(np.int64(a.size).tobytes())
        h.update(a.tobytes())
    return h.hexdigest()


def uniform_string(g: np.random.Generator, length: int, sigma: int) -> np.ndarray:
    return g.integers(0, sigma, size=length, dtype=np.uint8)


def mutate(g: np.random.Generator, a: np.ndarray, sigma: int, sub: float, indel: float) -> np.ndarray:
    """Copy of a with round(sub*len) substitutions (always to a different
    symbol) and round(indel*len) single-symbol indels, half insertions."""
    b = a.copy()
    L = len(b)
    ns = int(round(sub * L))
    if ns and sigma > 1:
        pos = g.choice(L, size=ns, replace=False)
        delta = g.integers(1, sigma, size=ns, dtype=np.int64)
        b[pos] = ((b[pos].astype(np.int64) + delta) % sigma).astype(np.uint8)
    nid = int(round(indel * L))
    n_ins, n_del = nid - nid // 2, nid // 2
    if n_del:
        keep = np.ones(len(b), bool)
        keep[g.choice(len(b), size=min(n_del, len(b) - 1), replace=False)] = False
        b = b[keep]
    if n_ins:
        pos = np.sort(g.integers(0, len(b) + 1, size=n_ins))
        syms = g.integers(0, sigma, size=n_ins, dtype=np.uint8)
        b = np.insert(b, pos, syms)
    return b.astype(np.uint8)


def lognormal_lengths(g: np.random.Generator, count: int, median: float, sigma_ln: float, lo: int, hi: int):
    x = np.round(g.lognormal(mean=np.log(median), sigma=sigma_ln, size=count))
    return np.clip(x, lo, hi).astype(np.int64)


# --- the five workloads ------------------------------------------------------

def plain_pair(seed: int = 1, length: int = 10_000, sigma: int = 4):
    """One pair, a and b each `length` uint8 symbols uniform over `sigma`."""
    g = rng(seed)
    return uniform_string(g, length, sigma), uniform_string(g, length, sigma)


def binary_pair(seed: int = 2, length: int = 1 << 20):
    """One pair of uniform bit strings (uint8 0/1)."""
    g = rng(seed)
    return uniform_string(g, length, 2), uniform_string(g, length, 2)


def similar_pair(seed: int = 3, length: int = 1_000_000, sigma: int = 4, sub: float = 0.05, indel: float = 0.01):
    """a uniform; b = a with `sub` substitutions and `indel` single-symbol indels."""
    g = rng(seed)
    a = uniform_string(g, length, sigma)
    return a, mutate(g, a, sigma, sub, indel)


def ragged_batch(seed: int = 4, count: int = 4096, median: float = 2000, sigma_ln: float = 1.0,
                 lo: int = 100, hi: int = 50_000, sigma: int = 20, mutated_share: float = 0.5,
                 sub: float = 0.10, indel: float = 0.02):
    """`count` pairs: lengths log-normal(median, sigma_ln) clipped to [lo, hi],
    independent for a and b; a `mutated_share` of the pairs has b = mutate(a)."""
    g = rng(seed)
    la = lognormal_lengths(g, count, median, sigma_ln, lo, hi)
    lb = lognormal_lengths(g, count, median, sigma_ln, lo, hi)
    mut = g.random(count) < mutated_share
    pairs = []
    for i in range(count):
        a = uniform_string(g, int(la[i]), sigma)
        if mut[i]:
            b = mutate(g, a, sigma, sub, indel)
        else:
            b = uniform_string(g, int(lb[i]), sigma)
        pairs.append((a, b))
    return pairs


@dataclass
class Request:
    rid: int
    arrival_s: float
    kind: str                      # "compare" or "query"
    a: np.ndarray | None = None
    b: np.ndarray | None = None
    target: int = -1               # rid of the comparison a query reads
    window: tuple = field(default_factory=tuple)   # (start, stop) of the substring
    qkind: str = ""                # "substring_string": LCS(a[start:stop), b); "string_substring": LCS(a, b[start:stop))


def serving_stream(seed: int = 5, count: int = 2000, rate_hz: float = 500.0, median: float = 5000,
                   sigma_ln: float = 1.0, lo: int = 200, hi: int = 200_000, sigma: int = 4,
                   query_share: float = 0.3, substring_string_share: float = 0.5, long_pairs=()):
    """Poisson arrivals at `rate_hz`; a `query_share` of requests are follow-up
    LCS queries on an earlier comparison (uniform window start, then uniform
    stop): a `substring_string_share` of them substring-vs-string,
    LCS(a[start:stop), b), the rest string-vs-substring, LCS(a, b[start:stop)).
    `long_pairs`: (request index, m, n) triples forcing those requests to be
    comparisons of these lengths (e.g. the clip maximum)."""
    g = rng(seed)
    gaps = g.exponential(1.0 / rate_hz, size=count)
    t = np.cumsum(gaps)
    forced = {int(i): (int(m), int(n)) for i, m, n in long_pairs}
    reqs: list[Request] = []
    compares: list[int] = []
    for i in range(count):
        if i not in forced and compares and g.random() < query_share:
            tgt = compares[int(g.integers(0, len(compares)))]
            if g.random() < substring_string_share:
                qkind, length = "substring_string", len(reqs[tgt].a)
            else:
                qkind, length = "string_substring", len(reqs[tgt].b)
            start = int(g.integers(0, length + 1))
            stop = int(g.integers(start, length + 1))
            reqs.append(Request(i, float(t[i]), "query", target=tgt, window=(start, stop), qkind=qkind))
        else:
            la, lb = lognormal_lengths(g, 2, median, sigma_ln, lo, hi)
            if i in forced:
                la, lb = forced[i]
            a = uniform_string(g, int(la), sigma)
            b = uniform_string(g, int(lb), sigma)
            reqs.append(Request(i, float(t[i]), "compare", a=a, b=b))
            compares.append(i)
    return reqs


def serving_parity_stream(count: int = 320, rate_hz: float = 2000.0):
    """The request stream of the serving parity fixture: config 5's law
    (Poisson arrivals, lengths log-normal median 5,000 clipped to [200,
    200,000], alphabet 4, 30% follow-up queries of both kinds), seed 55, with
    two comparisons at the clip maximum early in the stream: 200,000 x 200,000
    (request 3) and 200,000 x 700 (request 9)."""
    return serving_stream(seed=55, count=count, rate_hz=rate_hz,
                          long_pairs=((3, 200_000, 200_000), (9, 200_000, 700)))
