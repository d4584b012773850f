"""The counter-based synthetic generator (include/atrip_synth.h), CPU only.

Pins: SplitMix64 outputs against an independent Python implementation of the
published recurrence; moments of each distribution; the Tabij = Tbaji and
Vabij = Vbaji symmetries; determinism across thread counts.
"""
import numpy as np

MASK = (1 << 64) - 1


def _mix(z):
    z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & MASK
    z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & MASK
    return z ^ (z >> 31)


def test_uniform_stream_matches_splitmix64(oracle):
    # ea = 0.1 + 2.9 u, u = ((raw >> 11) + 0.5) / 2^53, raw = mix(s + (n+1) g)
    seed, nv = 5, 64
    s = _mix(_mix((seed + 0x9E3779B97F4A7C15) & MASK) ^ ((1 + 1) * 0xD1B54A32D192ED03 & MASK))
    t = oracle.synthetic(3, nv, seed)
    for n in range(nv):
        raw = _mix((s + (n + 1) * 0x9E3779B97F4A7C15) & MASK)
        u = ((raw >> 11) + 0.5) / 2.0 ** 53
        assert abs(t["ea"][n] - (0.1 + 2.9 * u)) <= 1e-15


def test_moments_and_ranges(oracle):
    no, nv = 12, 30
    t = oracle.synthetic(no, nv, seed=1, nthreads=4)
    assert t["ei"].min() >= -1.5 and t["ei"].max() <= -0.3
    assert t["ea"].min() >= 0.1 and t["ea"].max() <= 3.0
    for name, sd in [("Tai", 1e-3), ("Tabij", 1e-2), ("Vabij", 5e-2), ("Vijka", 5e-2),
                     ("Vabci", 5e-2)]:
        x = t[name]
        n = x.size
        assert abs(x.mean()) < 5 * sd / np.sqrt(n) + 1e-18, name
        assert abs(x.std() / sd - 1) < 5 / np.sqrt(n) + 0.02, name
        # normal tails: fraction beyond 2 sigma ~ 4.55 %
        frac = np.mean(np.abs(x) > 2 * sd)
        assert abs(frac - 0.0455) < 0.02, name


def test_symmetries(oracle):
    no, nv = 5, 7
    t = oracle.synthetic(no, nv, seed=2)
    for name in ("Tabij", "Vabij"):
        X = t[name].reshape((nv, nv, no, no), order="F")
        assert np.array_equal(X, X.transpose(1, 0, 3, 2))
    Vp = t["Vabci"].reshape((nv, nv, nv, no), order="F")
    assert not np.allclose(Vp, Vp.transpose(1, 0, 2, 3))


def test_deterministic_and_seeded(oracle):
    a = oracle.synthetic(4, 9, seed=3)
    b = oracle.synthetic(4, 9, seed=3, nthreads=3)
    c = oracle.synthetic(4, 9, seed=4)
    for k in a:
        assert np.array_equal(a[k], b[k])
        assert not np.array_equal(a[k], c[k])
