"""Paired Shell sorts of the reference (SRC/dsort2.c), test oracle only.

Each sorts `a1` ascending by a key and applies the same swaps to `a2`, in
place, with the gap sequence n/2, n/4, ..., 1.  The swap order matters for
ties, so the sequence and comparisons follow the reference exactly.
"""
import math


def _shell(n, a1, a2, key):
    gap = n // 2
    while gap > 0:
        for i in range(gap, n):
            j = i - gap
            while j >= 0:
                if key(a1[j]) > key(a1[j + gap]):
                    a1[j], a1[j + gap] = a1[j + gap], a1[j]
                    a2[j], a2[j + gap] = a2[j + gap], a2[j]
                    j -= gap
                else:
                    break
        gap //= 2


def dsort2(n, a1, a2):
    """dsort2 (dsort2.c:11): ascending by value."""
    _shell(n, a1, a2, lambda x: x)


def dsort2a(n, a1, a2):
    """dsort2a (dsort2.c:66): ascending by |value|."""
    _shell(n, a1, a2, lambda x: abs(x))


def dsort2s(n, s, a1, a2):
    """dsort2s (dsort2.c:121): ascending by |value - s|."""
    _shell(n, a1, a2, lambda x: abs(x - s))


def dsort2su(n, s, a1, a2):
    """dsort2su_ (dsort2.c:179): values above s first, by distance."""
    if n <= 0:
        return
    maxd = max(abs(a1[i]) for i in range(n))

    def key(x):
        v = abs(x)
        d = v - s
        return maxd + v if d < 0.0 else d

    _shell(n, a1, a2, key)


def dsort2sd(n, s, a1, a2):
    """dsort2sd (dsort2.c:252): values below s first, by distance."""
    if n <= 0:
        return
    maxd = max(abs(a1[i]) for i in range(n)) + 1.0

    def key(x):
        v = abs(x)
        d = s - v
        return s + (maxd - v) if d < 0.0 else d

    _shell(n, a1, a2, key)
