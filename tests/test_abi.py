"""The C-ABI library loads and exports every symbol include/astg.h declares."""
import ctypes
import os
import re

import astg

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    src = open(os.path.join(ROOT, "include", "astg.h")).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    return sorted(set(re.findall(r"\b(astg_[a-z0-9_]+)\s*\(", src)))


def test_library_exports_header_symbols():
    lib = ctypes.CDLL(astg.LIB_PATH)
    names = declared_symbols()
    assert len(names) >= 19
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing
    assert sorted(astg.EXPORTS) == names


def test_host_only_context_rejects_device_calls():
    c = astg.Context(-1)
    import numpy as np
    with __import__("pytest").raises(astg.AstgError) as e:
        c.pcg_solve_csr(np.array([0, 1]), np.array([0]), np.array([1.0]), np.array([1.0]))
    assert e.value.code == astg.ASTG_ERR_STATE


def test_bad_arguments_are_reported():
    c = astg.Context(-1)
    import numpy as np
    with __import__("pytest").raises(astg.AstgError) as e:
        c.set_topology(np.full(8, 5, np.int32), 2, 2, 0)
    assert e.value.code == astg.ASTG_ERR_ARG
