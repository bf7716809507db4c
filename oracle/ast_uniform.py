"""CPU restatement of the reference's uniform-layout AST projection.

TEST INFRASTRUCTURE ONLY: the tests use it as the checker; nothing in the
product imports it.

The reference is ast::ASTSolver<d>::project (include/ast_solver.h:322-498) with
the build's own configuration (include/ast_general.h:11-15: CORRECTION_MAC,
EDGE_NODE_ADAPTIVITY false), under which edge nodes are never unknowns
(isSkipClosed, ast_iterator.h:108-113) and every edge-node status is false:
  * numbering (updateIndex, ast_solver.h:599-621): every cell in box order
    (x fastest), then every open tilt node in box order;
  * the boundary tilt nodes are fixed closed (initData 79-87,
    ASTGridLayout::constrainFixedTiltNode, ast_grid_layout.h:179-186);
  * cell rows (327-402): open corner tilts first (each cuts the d MAC faces it
    touches), then the 2d faces; a face whose remaining area is not above the
    machine epsilon is no coupling; a face on the y walls (axis 1) is a
    Dirichlet face; zero diagonals become 1;
  * tilt rows (403-428): the 2^d neighbouring cells;
  * b scaled by rho/dt (471), MIC(0)-PCG with tau 0.97, min-diagonal ratio
    0.95, 5000 iterations, stopping at an absolute infinity-norm residual of
    1e-5 (initialize 40; PCGSolver::solve, pcg_solver.h:302-358, whose
    std::vector form does not scale the tolerance by the initial residual);
  * correctVelocityByPressure (283-320) and constrainVelocity (623-643).
Arrays follow the reference's grids, x fastest: cells prod(res), tilt nodes
prod(res + 1), MAC faces of axis a prod(res + e_a), tilt velocities node-major
(prod(res + 1), 4 or 8).
"""
import numpy as np

from . import oracle as orc

SQRT2 = 1.4142135623730951  # macro.h:16
SQRT3 = 1.7320508075688772  # macro.h:17
EPS = np.finfo(np.float64).eps  # general.h:48

# corner tilts of a cell and the cells of a tilt (constants.h:71-77;
# ast_constant.h:62-70), the MAC faces each corner tilt cuts (ast_constant.h:38-56)
OCT_NB_TILT = {2: [(1, 1), (1, 0), (0, 0), (0, 1)],
               3: [(1, 1, 1), (1, 0, 1), (0, 0, 1), (0, 1, 1), (1, 1, 0), (1, 0, 0), (0, 0, 0), (0, 1, 0)]}
TILT_NB_OCT = {2: [(0, 0), (0, -1), (-1, -1), (-1, 0)],
               3: [(0, 0, 0), (0, -1, 0), (-1, -1, 0), (-1, 0, 0),
                   (0, 0, -1), (0, -1, -1), (-1, -1, -1), (-1, 0, -1)]}
TILT_NB_MAC = {2: [(0, 2), (0, 3), (1, 3), (1, 2)],
               3: [(0, 2, 4), (0, 3, 4), (1, 3, 4), (1, 2, 4), (0, 2, 5), (0, 3, 5), (1, 3, 5), (1, 2, 5)]}


def opposite(d, k):
    """oppositeTiltIdx (ast_util.h:27-30)."""
    return (k + 2) % 4 if d == 2 else ((8 if k & 1 else 6) - k)


def lin(coords, shape):
    """Box-order linear index, x fastest."""
    i = np.zeros_like(coords[0])
    m = 1
    for a in range(len(shape)):
        i = i + coords[a] * m
        m *= shape[a]
    return i


def grid(shape):
    """Coordinates of every point of a box, x fastest, as a list of arrays."""
    g = np.meshgrid(*[np.arange(s) for s in shape], indexing="ij")
    return [c.reshape(-1, order="F") for c in g]


def constrain_layout(res, open_):
    """Boundary tilt nodes closed (constrainFixedTiltNode)."""
    d = len(res)
    tshape = [r + 1 for r in res]
    o = np.asarray(open_, dtype=np.uint8).reshape(-1).copy()
    tc = grid(tshape)
    edge = np.zeros(o.size, bool)
    for a in range(d):
        edge |= (tc[a] == 0) | (tc[a] == res[a])
    o[edge] = 0
    return o


def numbering(res, open_):
    """(cell index per cell, unknown index per tilt node or -1, n)."""
    nc = int(np.prod(res))
    o = constrain_layout(res, open_)
    tilt = np.full(o.size, -1, np.int64)
    idx = np.nonzero(o)[0]
    tilt[idx] = nc + np.arange(idx.size)
    return np.arange(nc), tilt, nc + idx.size


def assemble(res, dx, open_, mac, tilt_u, dt, rho):
    """A (CSR, rows sorted by column), b of ASTSolver::project."""
    d = len(res)
    res = [int(r) for r in res]
    half = dx / 2.0
    tod = half * (SQRT2 if d == 2 else SQRT3)
    tarea = half * SQRT2 if d == 2 else SQRT3 * 0.5 * half * half
    cut = half if d == 2 else 0.5 * half * half
    full = dx if d == 2 else dx * dx
    kT = 4 if d == 2 else 8
    tshape = [r + 1 for r in res]
    _, tid, n = numbering(res, open_)
    nc = int(np.prod(res))
    tu = np.asarray(tilt_u, dtype=np.float64).reshape(-1, kT)
    rows, cols, vals = [], [], []
    b = np.zeros(n)
    # cell rows
    cc = grid(res)
    area = np.zeros((2 * d, nc))
    center = np.zeros(nc)
    tb = np.zeros(nc)
    for k in range(kT):
        tcoor = [cc[a] + OCT_NB_TILT[d][k][a] for a in range(d)]
        t = lin(tcoor, tshape)
        op = tid[t] >= 0
        for i in range(d):
            area[TILT_NB_MAC[d][k][i], op] -= cut
        coeff = tarea / tod
        rows.append(np.nonzero(op)[0])
        cols.append(tid[t][op])
        vals.append(np.full(op.sum(), -coeff))
        center[op] += coeff
        sign = -1.0 if k <= (1 if d == 2 else 3) else 1.0
        tb[op] += sign * tarea * tu[t[op], opposite(d, k)]
    for k in range(2 * d):
        area[k] += full
        ax = k // 2
        up = (k & 1) == 0
        nb = [c.copy() for c in cc]
        nb[ax] = nb[ax] + (1 if up else -1)
        valid = (nb[ax] >= 0) & (nb[ax] < res[ax])
        open_face = area[k] > EPS
        fshape = [r + (1 if a == ax else 0) for a, r in enumerate(res)]
        fc = [c.copy() for c in cc]
        fc[ax] = fc[ax] + (1 if up else 0)
        u = np.asarray(mac[ax], dtype=np.float64).reshape(-1)[lin(fc, fshape)]
        s = -1.0 if up else 1.0
        m = open_face & valid
        coeff = area[k] / dx
        rows.append(np.nonzero(m)[0])
        cols.append(lin(nb, res)[m])
        vals.append(-coeff[m])
        center[m] += coeff[m]
        tb[m] += s * area[k][m] * u[m]
        if ax == 1:
            m2 = open_face & ~valid
            center[m2] += coeff[m2]
            tb[m2] += s * area[k][m2] * u[m2]
    center[center < EPS] = 1.0
    rows.append(np.arange(nc))
    cols.append(np.arange(nc))
    vals.append(center)
    b[:nc] = tb
    # tilt rows
    ot = np.nonzero(tid >= 0)[0]
    if ot.size:
        tcs = [c[ot] for c in grid(tshape)]
        r = tid[ot]
        center = np.zeros(ot.size)
        tb = np.zeros(ot.size)
        for k in range(kT):
            ncoor = [tcs[a] + TILT_NB_OCT[d][k][a] for a in range(d)]
            coeff = tarea / tod
            rows.append(r)
            cols.append(lin(ncoor, res))
            vals.append(np.full(ot.size, -coeff))
            center += coeff
            sign = -1.0 if k <= (1 if d == 2 else 3) else 1.0
            tb += sign * tarea * tu[ot, k]
        rows.append(r)
        cols.append(r)
        vals.append(center)
        b[r] = tb
    b = b * (rho / dt)
    R = np.concatenate(rows)
    C = np.concatenate(cols)
    V = np.concatenate(vals)
    order = np.lexsort((C, R))
    R, C, V = R[order], C[order], V[order]
    rs = np.zeros(n + 1, np.int64)
    np.add.at(rs, R + 1, 1)
    return np.cumsum(rs).astype(np.int32), C.astype(np.int32), V, b


def correct(res, dx, open_, mac, tilt_u, p, dt, rho):
    """correctVelocityByPressure then constrainVelocity; returns new arrays."""
    d = len(res)
    res = [int(r) for r in res]
    half = dx / 2.0
    tod = half * (SQRT2 if d == 2 else SQRT3)
    kT = 4 if d == 2 else 8
    tshape = [r + 1 for r in res]
    _, tid, _ = numbering(res, open_)
    tu = np.asarray(tilt_u, dtype=np.float64).reshape(-1, kT).copy()
    ot = np.nonzero(tid >= 0)[0]
    if ot.size:
        tcs = [c[ot] for c in grid(tshape)]
        pc = p[tid[ot]]
        for k in range(kT):
            nb = lin([tcs[a] + TILT_NB_OCT[d][k][a] for a in range(d)], res)
            s = -1.0 if k < kT // 2 else 1.0
            tu[ot, k] += s * dt / rho * (p[nb] - pc) / tod
    out = []
    for ax in range(d):
        fshape = [r + (1 if a == ax else 0) for a, r in enumerate(res)]
        fc = grid(fshape)
        u = np.asarray(mac[ax], dtype=np.float64).reshape(-1).copy()
        v1 = np.all([(fc[a] >= 0) & (fc[a] < res[a]) for a in range(d)], axis=0)
        lo = [c.copy() for c in fc]
        lo[ax] = lo[ax] - 1
        v0 = np.all([(lo[a] >= 0) & (lo[a] < res[a]) for a in range(d)], axis=0)
        p1 = np.zeros(u.size)
        p0 = np.zeros(u.size)
        p1[v1] = p[lin([c[v1] for c in fc], res)]
        p0[v0] = p[lin([c[v0] for c in lo], res)]
        u -= dt / rho * (p1 - p0) / dx
        out.append(u)
    nx, ny = res[0], res[1]
    if d == 2:
        for x in (0, nx):
            out[0][x + np.arange(ny) * (nx + 1)] = 0.0
    else:
        nz = res[2]
        y, z = np.meshgrid(np.arange(ny), np.arange(nz), indexing="ij")
        for x in (0, nx):
            out[0][(x + y * (nx + 1) + z * (nx + 1) * ny).reshape(-1)] = 0.0
        # the reference runs x over [0, ny) here (ast_solver.h:634), with the
        # y-face grid's own linear index
        x, z = np.meshgrid(np.arange(ny), np.arange(nz), indexing="ij")
        for yy in (0, ny):
            out[1][(x + yy * nx + z * nx * (ny + 1)).reshape(-1)] = 0.0
    return out, tu


def project(res, dx, open_, mac, tilt_u, dt, rho=2.0, tol=1e-5, maxit=5000, tau=0.97, ratio=0.95):
    """The whole projection: system, pressure, corrected velocity and the
    solve's iterations (absolute tolerance, as the reference)."""
    rs, col, val, b = assemble(res, dx, open_, mac, tilt_u, dt, rho)
    bmax = np.max(np.abs(b)) if b.size else 0.0
    if bmax == 0:
        p, iters, resid = np.zeros(b.size), 0, 0.0
    else:
        p, resid, iters, ok = orc.pcg(rs, col, val, b, tol=tol / bmax, maxit=maxit, tau=tau, ratio=ratio)
    mac2, tu2 = correct(res, dx, open_, mac, tilt_u, p, dt, rho)
    return {"rowstart": rs, "col": col, "val": val, "b": b, "p": p, "iters": iters, "residual": resid,
            "mac": mac2, "tilt": tu2}
