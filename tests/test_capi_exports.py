"""The C-ABI library loads and exports every symbol include/aprl_gpu.h declares
(no compute calls; CPU-only)."""
import ctypes
import os
import re

import harness as H


def declared_symbols():
    text = open(os.path.join(H.ROOT, "include", "aprl_gpu.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(aprl_[a-z0-9_]+)\s*\(", text)))


def test_library_exports_every_declared_symbol():
    import aprl_amd
    lib = ctypes.CDLL(aprl_amd._capi.LIB_PATH)
    names = declared_symbols()
    assert len(names) >= 20
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing


def test_python_binding_lists_every_declared_symbol():
    import aprl_amd
    assert sorted(set(aprl_amd._capi.EXPORTED)) == declared_symbols()
