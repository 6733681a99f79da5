"""Integer helpers of the FFT planner.

``prime_factors`` / ``group_primes`` / ``pad_dim`` keep the reference's
behaviour (vkdispatch/fft/prime_utils.py:12-64): factors come out 2s first,
then odd primes ascending; ``pad_dim`` rounds a size up until every prime
factor is at most ``max_register_count`` (default 16, i.e. primes <= 13).
"""
from __future__ import annotations

from typing import List


def default_register_limit() -> int:
    return 16


def default_max_prime() -> int:
    return 13


def prime_factors(n: int) -> List[int]:
    if n < 1:
        raise ValueError("n must be positive")
    out: List[int] = []
    while n % 2 == 0:
        out.append(2)
        n //= 2
    f = 3
    while f * f <= n:
        while n % f == 0:
            out.append(f)
            n //= f
        f += 2
    if n > 1:
        out.append(n)
    return out


def group_primes(primes, register_count: int) -> List[List[int]]:
    groups: List[List[int]] = []
    for p in primes:
        if groups:
            prod = 1
            for q in groups[-1]:
                prod *= q
            if prod * p <= register_count:
                groups[-1].append(p)
                continue
        groups.append([p])
    return groups


def pad_dim(dim: int, max_register_count: int = None) -> int:
    if max_register_count is None:
        max_register_count = default_register_limit()
    if dim <= 0:
        raise ValueError("Dimension must be greater than 0")
    d = dim
    while any(p > max_register_count for p in prime_factors(d)):
        d += 1
    return d


def is_prime(n: int) -> bool:
    return n >= 2 and prime_factors(n) == [n]
