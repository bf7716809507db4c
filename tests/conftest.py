"""Shared test set-up: import paths, markers and golden-fixture loading."""
import glob
import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "ast-grid_amd"))

# projection fixtures (the adv_*.npz advection and adj_*.npz tilt-adaptivity
# fixtures build on them; smoke64.npz is the multi-step smoke configuration,
# frame32.npz the on-disk frame formats, col_*.npz the collider scenes,
# ast*.npz the uniform-layout AST projection)
GOLDEN = sorted(p for p in glob.glob(os.path.join(ROOT, "tests", "golden", "*.npz"))
                if not os.path.basename(p).startswith(("adv_", "adj_", "smoke", "frame", "col_", "ast")))
AST_GOLDEN = sorted(glob.glob(os.path.join(ROOT, "tests", "golden", "ast*.npz")))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs through the C ABI on device 0)")


def load_golden(path):
    z = np.load(path)
    return {k: z[k] for k in z.files}


@pytest.fixture(params=GOLDEN, ids=[os.path.basename(p)[:-4] for p in GOLDEN])
def golden(request):
    return load_golden(request.param)


def make_context(g, device):
    """Context with the fixture's topology, tilt openings and input velocities."""
    import astg
    c = astg.Context(device)
    R, L = int(g["R"]), int(g["levels"])
    c.set_topology(g["split"].astype(np.int32), R, L, int(g["closed"]))
    if int(g["open_tilts"]):
        c.set_tilt_open(0, g["open_mask_l0"])
    return c
