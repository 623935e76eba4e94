"""Registers the `pas-cg_amd/` package directory as the module `pas_cg_amd`."""
import importlib.util
import os
import sys

ROOT = os.path.dirname(os.path.abspath(__file__))


def load():
    if "pas_cg_amd" in sys.modules:
        return sys.modules["pas_cg_amd"]
    pkg_dir = os.path.join(ROOT, "pas-cg_amd")
    spec = importlib.util.spec_from_file_location(
        "pas_cg_amd", os.path.join(pkg_dir, "__init__.py"),
        submodule_search_locations=[pkg_dir])
    mod = importlib.util.module_from_spec(spec)
    sys.modules["pas_cg_amd"] = mod
    spec.loader.exec_module(mod)
    return mod
