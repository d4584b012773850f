"""Seeded shapes of the batched and served (T) workloads.

Shapes are drawn with SplitMix64 (one draw per (seed, counter)), so a
workload is reproducible from its seed alone:
  * batched: n systems, No ~ U{4..16}, Nv ~ U{20..120}, system s uses the
    tensor seed 1000 + s;
  * served: request r has No and Nv log-uniform in [8, 40] and [40, 320]
    (rounded to integers), arrival times from a Poisson process of the given
    rate, and tensor seed 5000 + r.
"""
from __future__ import annotations

import math

MASK = (1 << 64) - 1


def mix64(x: int) -> int:
    """SplitMix64 output for state x."""
    z = (x + 0x9E3779B97F4A7C15) & MASK
    z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & MASK
    z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & MASK
    return z ^ (z >> 31)


def draw(seed: int, counter: int) -> int:
    return mix64((seed * 0x632BE59BD9B4E019 + counter) & MASK)


def u01(seed: int, counter: int) -> float:
    """Uniform in (0, 1) from the top 53 bits."""
    return ((draw(seed, counter) >> 11) + 0.5) / float(1 << 53)


def batch_shapes(n: int = 256, seed: int = 0, no=(4, 16), nv=(20, 120)):
    """[(No, Nv, tensor_seed)] of the batched workload."""
    out = []
    for s in range(n):
        a = no[0] + draw(seed, 2 * s) % (no[1] - no[0] + 1)
        b = nv[0] + draw(seed, 2 * s + 1) % (nv[1] - nv[0] + 1)
        out.append((int(a), int(b), 1000 + s))
    return out


def serve_requests(n: int, rate: float = 20.0, seed: int = 0, no=(8, 40),
                   nv=(40, 320)):
    """[(arrival_s, No, Nv, tensor_seed)] of the served workload."""
    out = []
    t = 0.0
    for r in range(n):
        t += -math.log(u01(seed, 3 * r)) / rate
        lno = math.log(no[0]) + u01(seed, 3 * r + 1) * (math.log(no[1]) - math.log(no[0]))
        lnv = math.log(nv[0]) + u01(seed, 3 * r + 2) * (math.log(nv[1]) - math.log(nv[0]))
        out.append((t, int(round(math.exp(lno))), int(round(math.exp(lnv))), 5000 + r))
    return out


def doubles_flop(no: int, nv: int) -> float:
    """Algorithmic doubles FLOP of a full system: 12 No^3 (No+Nv) per tuple."""
    n = nv * (nv + 1) * (nv + 2) // 6 - nv
    return 12.0 * no ** 3 * (no + nv) * n


def n_tuples(nv: int) -> int:
    return nv * (nv + 1) * (nv + 2) // 6 - nv
