"""pas-cg_amd: MI355X-native per-step loop of the PAS coarse-grained protein
model (drop-in for the reference's per-force-term loop).

The package directory name is not a Python identifier; load it with
`pcg_import.load()` (repo root) or through tests/conftest.py, which register
it as `pas_cg_amd`.
"""
from . import params, model  # noqa: F401
