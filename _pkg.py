"""Import helper: the package directory ``topopt.jl_amd`` has a dot in its
name, so it is registered under the module name ``topopt_jl_amd``."""
from __future__ import annotations

import importlib.util
import os
import sys

ROOT = os.path.dirname(os.path.abspath(__file__))
PKG_DIR = os.path.join(ROOT, "topopt.jl_amd")


def load_package():
    if "topopt_jl_amd" in sys.modules:
        return sys.modules["topopt_jl_amd"]
    spec = importlib.util.spec_from_file_location(
        "topopt_jl_amd", os.path.join(PKG_DIR, "__init__.py"),
        submodule_search_locations=[PKG_DIR])
    mod = importlib.util.module_from_spec(spec)
    sys.modules["topopt_jl_amd"] = mod
    spec.loader.exec_module(mod)
    return mod


def load_oracle():
    """The oracle (test infrastructure only: tests, smoke check, CPU baseline)."""
    if ROOT not in sys.path:
        sys.path.insert(0, ROOT)
    from oracle import oracle  # noqa: E402
    return oracle
