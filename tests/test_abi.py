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


SHIM = os.path.join(ROOT, "ast-grid_amd", "astg", "cuda_solver_shim")


def write_csr(path, rs, col, val, b):
    import numpy as np
    with open(path, "wb") as f:
        np.array([len(rs) - 1, len(col)], np.int32).tofile(f)
        np.asarray(rs, np.int32).tofile(f)
        np.asarray(col, np.int32).tofile(f)
        np.asarray(val, np.float64).tofile(f)
        np.asarray(b, np.float64).tofile(f)


def test_cuda_solver_shim_is_built_and_rejects_truncated_input(tmp_path):
    """The C++ CudaSolver-shaped shim (include/astg_cuda_solver.hpp) compiles,
    links against libastg.so and validates its input before any device call."""
    import subprocess
    assert os.access(SHIM, os.X_OK)
    bad = tmp_path / "bad.bin"
    bad.write_bytes(b"\x03\x00\x00\x00")
    p = subprocess.run([SHIM, str(bad), str(tmp_path / "out.bin")], capture_output=True, text=True)
    assert p.returncode == 2 and "truncated" in p.stderr


def test_batch_and_tuning_entry_points_validate_arguments():
    """Batched solve, tuner and tuning setter: null / empty batches and bad
    launch shapes are refused before any device work; a host-only context
    has no device."""
    import numpy as np
    import pytest
    L = astg.lib()
    c = astg.Context(-1)
    prm = astg.default_params()
    res = (astg.PcgResult * 1)()
    assert L.astg_pcg_solve_batch(None, 1, None, None, None, None, None, None, ctypes.byref(prm), res) == \
        astg.ASTG_ERR_ARG
    assert L.astg_pcg_solve_batch(c.h, 0, None, None, None, None, None, None, ctypes.byref(prm), res) == \
        astg.ASTG_ERR_ARG
    with pytest.raises(astg.AstgError) as e:
        c.pcg_solve_batch([(np.array([0, 1]), np.array([0]), np.array([1.0]), np.array([1.0]))])
    assert e.value.code == astg.ASTG_ERR_STATE
    t = astg.BatchTuning()
    L.astg_batch_tuning_default(ctypes.byref(t))
    assert (t.merge_levels, t.stream_blocks_per_cu, t.lanes_per_row) == (2, 8, 8)
    for bad in ((0, 8, 8), (9, 8, 8), (2, -1, 8), (2, 65, 8), (2, 8, 6)):
        with pytest.raises(astg.AstgError) as e:
            c.set_batch_tuning(*bad)
        assert e.value.code == astg.ASTG_ERR_ARG
    c.set_batch_tuning(1, 4, 16)
    best = astg.BatchTuning()
    assert L.astg_pcg_tune_batch(c.h, 1, None, None, None, None, None, ctypes.byref(prm), 20,
                                 ctypes.byref(best)) == astg.ASTG_ERR_ARG


def test_frame_entry_points_need_a_topology():
    import pytest
    c = astg.Context(-1)
    with pytest.raises(astg.AstgError) as e:
        c.write_frame(layout="/nonexistent/x.bin")
    assert e.value.code == astg.ASTG_ERR_STATE
    with pytest.raises(astg.AstgError) as e:
        c.read_frame(layout="/nonexistent/x.bin")
    assert e.value.code == astg.ASTG_ERR_STATE
    assert astg.lib().astg_write_frame(None, None, None, None, None, None) == astg.ASTG_ERR_ARG
