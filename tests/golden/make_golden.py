"""Writes tests/golden/c1_seed1.json: E(T) and sampled per-tuple energies of the
CPU-runnable configuration (No=10, Nv=50, seed 1) from the oracle.

The values come from the oracle's direct-loop form and are cross-checked here
against its DGEMM form; the oracle itself is pinned by the full-space identity
and planted-mistake tests in tests/test_oracle.py.  Run from the repo root:
    python tests/golden/make_golden.py
"""
import json
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)
from oracle import oracle as O  # noqa: E402

no, nv, seed = 10, 50, 1
t = O.synthetic(no, nv, seed)
e, ect, pt, _ = O.run(no, nv, t, per_tuple=True)
O.blas_init()
e2, _, pt2, _ = O.run(no, nv, t, use_gemm=True, per_tuple=True)
assert abs(e - e2) <= 1e-12 * abs(e)
assert np.max(np.abs(pt - pt2)) <= 1e-12 * np.max(np.abs(pt))
idx = [0, 1, 2, 48, 49, 1225, 5000, 12345, 19599, len(pt) - 1]
out = {
    "config": {"No": no, "Nv": nv, "seed": seed, "generator": "include/atrip_synth.h"},
    "n_tuples": int(len(pt)),
    "energy": e,
    "ct_energy": ect,
    "max_abs_tuple_energy": float(np.max(np.abs(pt))),
    "tuple_energies": {str(i): float(pt[i]) for i in idx},
}
path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "c1_seed1.json")
with open(path, "w") as f:
    json.dump(out, f, indent=1)
print(path, e)
