"""MI355X-native (T) perturbative-triples hot path of atrip.

`atrip_amd.hip` binds the C-ABI library (libatrip_hip.so); `atrip_amd.api`
mirrors the reference's atrip::Atrip interface (Input builder, run, Output)
on top of it.
"""
from . import hip  # noqa: F401
