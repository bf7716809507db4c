"""Caller-owned device fields (PyTorch-ROCm tensors) through the C ABI.

The projection of a golden scene run on torch tensors (astg_project_fields:
device pointers, the caller's current HIP stream) gives bit-identical
velocities to the same projection through the host-array path
(astg_set_velocity / astg_project / astg_get_velocity), also when the
caller's work runs on a non-default torch stream; scalar fields round-trip
unchanged through astg_fields_upload / astg_fields_download.
"""
import numpy as np
import pytest

from conftest import load_golden, make_context, ROOT

pytestmark = pytest.mark.gpu

SCENE = "s32_l4_open_tilts"


def flat(ctx, per_level, width):
    """Per-level (nnode, width) arrays -> the flat component-major layout."""
    G, off = ctx.field_layout()
    out = np.zeros(width * G)
    for l, a in enumerate(per_level):
        for k in range(width):
            out[k * G + off[l]:k * G + off[l + 1]] = a[:, k]
    return out


def unflat(ctx, x, width):
    G, off = ctx.field_layout()
    return [np.stack([x[k * G + off[l]:k * G + off[l + 1]] for k in range(width)], axis=1)
            for l in range(len(off) - 1)]


def test_torch_projection_equals_host_path():
    import torch
    import astg
    g = load_golden("%s/tests/golden/%s.npz" % (ROOT, SCENE))
    L = int(g["levels"])
    prm = astg.default_params(tol=1e-8, max_iterations=3000)
    dt, rho = float(g["dt"]), float(g["rho"])
    # host-array path
    a = make_context(g, 0)
    for l in range(L):
        a.set_velocity(l, g["u_mac_in_%d" % l], g["u_tilt_in_%d" % l])
    ra = a.project(dt, rho, prm)
    host = [a.velocity(l) for l in range(L)]
    # torch path, on a side stream of the caller
    b = make_context(g, 0)
    mac = torch.tensor(flat(b, [g["u_mac_in_%d" % l] for l in range(L)], 3), device="cuda")
    tilt = torch.tensor(flat(b, [g["u_tilt_in_%d" % l] for l in range(L)], 8), device="cuda")
    side = torch.cuda.Stream()
    side.wait_stream(torch.cuda.current_stream())
    with torch.cuda.stream(side):
        print("torch path: project_fields on a side stream", flush=True)
        rb = b.project_fields(dt, rho, prm, mac, tilt)
        mac2 = mac * 1.0  # queued after the projection on the same stream
    torch.cuda.synchronize()
    assert rb.iterations == ra.iterations and rb.converged == 1
    m = unflat(b, mac2.cpu().numpy(), 3)
    t = unflat(b, tilt.cpu().numpy(), 8)
    for l in range(L):
        assert np.array_equal(m[l], host[l][0]), l
        assert np.array_equal(t[l], host[l][1]), l
    a.close()
    b.close()


def test_scalar_fields_round_trip():
    import torch
    g = load_golden("%s/tests/golden/%s.npz" % (ROOT, SCENE))
    c = make_context(g, 0)
    L = int(g["levels"])
    G, off = c.field_layout()
    rng = np.random.default_rng(11)
    d = [torch.tensor(rng.random(G), device="cuda") for _ in range(4)]
    c.upload_fields(density_cell=d[0], density_vert=d[1], heat_cell=d[2], heat_vert=d[3])
    # the host path sees the same values per level
    for l in range(L):
        o, v = c.scalar(0, l)
        assert np.array_equal(o, d[0].cpu().numpy()[off[l]:off[l + 1]])
        assert np.array_equal(v, d[1].cpu().numpy()[off[l]:off[l + 1]])
    out = [torch.zeros(G, dtype=torch.float64, device="cuda") for _ in range(4)]
    c.download_fields(density_cell=out[0], density_vert=out[1], heat_cell=out[2], heat_vert=out[3])
    torch.cuda.synchronize()
    for x, y in zip(out, d):
        assert torch.equal(x, y)
    # a host pointer mixed with a device tensor is refused
    with pytest.raises(ValueError):
        c.upload_fields(density_cell=d[0], density_vert=np.zeros(G))
    c.close()
