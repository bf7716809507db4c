"""The uniform-layout AST solver's projection on the device.

Mirrors the reference's ast::ASTSolver<d> (include/ast_solver.h:20-673), which
its pybind module exposes as ASTSolver2 / ASTSolver3(resolution, dx,
adaptive_dofs, name) (python_bindings/ast_solver.cc:43-54), for the projection
path: the tilt layout and the velocity field go in, project(dt) runs
ASTSolver::project (ast_solver.h:322-498) on the GPU through the C ABI
(astg_ast_*), and the corrected velocity, the system and the pressure come
back. Arrays follow the reference's grids, x fastest: tilt nodes prod(res + 1),
MAC faces of axis a prod(res + e_a), tilt velocities (prod(res + 1), 4 or 8).
"""
import ctypes

import numpy as np

from . import AstgError, PcgParams, PcgResult, _ptr, lib


def default_params(**kw):
    """ASTSolver::initialize's solver settings (ast_solver.h:40): tol 1e-5 as
    an absolute bound on the infinity norm of the residual, 5000 iterations,
    MIC(0) 0.97, min-diagonal ratio 0.95."""
    p = PcgParams()
    lib().astg_ast_params_default(ctypes.byref(p))
    for k, v in kw.items():
        setattr(p, k, v)
    return p


class ASTSolver:
    """ASTSolver<dim>(res, dx) after initialize(): fields zero, tilts closed."""

    def __init__(self, dim, res, dx, device=0):
        self._L = lib()
        self.dim = int(dim)
        self.res = tuple(int(r) for r in res)
        assert len(self.res) == self.dim
        self.dx = float(dx)
        self.kt = 4 if self.dim == 2 else 8
        h = ctypes.c_void_p()
        r = np.array(self.res, np.int32)
        rc = self._L.astg_ast_create(ctypes.byref(h), device, self.dim, _ptr(r), self.dx)
        self.h = h if h.value else None
        if rc:
            msg = self._L.astg_ast_last_error(self.h).decode() if self.h else ""
            self.close()
            raise AstgError(rc, "astg_ast_create: %s" % msg)
        self.n = self.nnz = 0

    def close(self):
        if self.h:
            self._L.astg_ast_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _chk(self, rc, what):
        if rc:
            raise AstgError(rc, "%s: %s" % (what, self._L.astg_ast_last_error(self.h).decode()))

    def tilt_count(self):
        return int(np.prod([r + 1 for r in self.res]))

    def mac_count(self, a):
        return int(np.prod([r + (1 if i == a else 0) for i, r in enumerate(self.res)]))

    def set_layout(self, tilt_open):
        o = np.ascontiguousarray(tilt_open, np.uint8).reshape(-1)
        if o.size != self.tilt_count():
            raise ValueError("tilt_open needs %d flags" % self.tilt_count())
        self._chk(self._L.astg_ast_set_layout(self.h, _ptr(o)), "set_layout")

    def set_velocity(self, mac, tilt):
        m = [np.ascontiguousarray(mac[a], np.float64).reshape(-1) for a in range(self.dim)]
        t = np.ascontiguousarray(tilt, np.float64).reshape(-1)
        for a in range(self.dim):
            if m[a].size != self.mac_count(a):
                raise ValueError("MAC axis %d needs %d faces" % (a, self.mac_count(a)))
        if t.size != self.tilt_count() * self.kt:
            raise ValueError("tilt needs %d values" % (self.tilt_count() * self.kt))
        self._chk(self._L.astg_ast_set_velocity(self.h, _ptr(m[0]), _ptr(m[1]), _ptr(m[2]) if self.dim == 3 else None,
                                                _ptr(t)), "set_velocity")

    def velocity(self):
        m = [np.empty(self.mac_count(a)) for a in range(self.dim)]
        t = np.empty(self.tilt_count() * self.kt)
        self._chk(self._L.astg_ast_get_velocity(self.h, _ptr(m[0]), _ptr(m[1]), _ptr(m[2]) if self.dim == 3 else None,
                                                _ptr(t)), "velocity")
        return m, t.reshape(-1, self.kt)

    def project(self, dt, rho=2.0, params=None):
        prm = params if params is not None else default_params()
        res = PcgResult()
        self._chk(self._L.astg_ast_project(self.h, float(dt), float(rho), ctypes.byref(prm), ctypes.byref(res)),
                  "project")
        self.n, self.nnz = res.n, res.nnz
        return res

    def system(self):
        """(rowstart, col, val, b, p) of the last projection."""
        rs = np.empty(self.n + 1, np.int32)
        col = np.empty(self.nnz, np.int32)
        val = np.empty(self.nnz)
        b = np.empty(self.n)
        p = np.empty(self.n)
        self._chk(self._L.astg_ast_system(self.h, self.n, self.nnz, _ptr(rs), _ptr(col), _ptr(val), _ptr(b), _ptr(p)),
                  "system")
        return rs, col, val, b, p

    def tilt_index(self):
        t = np.empty(self.tilt_count(), np.int32)
        self._chk(self._L.astg_ast_tilt_index(self.h, _ptr(t)), "tilt_index")
        return t


def ASTSolver2(resolution, dx, device=0):
    """The reference's ASTSolver2(resolution, dx, ...) shape."""
    return ASTSolver(2, resolution, dx, device)


def ASTSolver3(resolution, dx, device=0):
    """The reference's ASTSolver3(resolution, dx, ...) shape."""
    return ASTSolver(3, resolution, dx, device)
