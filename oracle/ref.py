"""ctypes wrapper of oracle/_ref/libastgref.so (TEST INFRASTRUCTURE ONLY).

The library is the reference solver compiled from its own sources plus
oracle/ref_harness.cpp (see oracle/ref.mk). It exists only in a container that
has the reference tree; everything here raises ``RefUnavailable`` otherwise.
"""
import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "_ref", "libastgref.so")


class RefUnavailable(RuntimeError):
    pass


_lib = None


def available():
    return os.path.exists(LIB_PATH)


def lib():
    global _lib
    if _lib is not None:
        return _lib
    if not available():
        raise RefUnavailable("oracle/_ref/libastgref.so not built (needs the reference sources)")
    L = ctypes.CDLL(LIB_PATH)
    P, I, D = ctypes.c_void_p, ctypes.c_int, ctypes.c_double
    sig = {
        "ref_last_error": (ctypes.c_char_p, []),
        "ref_quiet": (None, []),
        "ref_num_workers": (I, []),
        "ref_create": (P, [I, D, I, P, I]),
        "ref_destroy": (None, [P]),
        "ref_level_size": (I, [P, I, I]),
        "ref_nodes": (None, [P, I, P, P]),
        "ref_tilt_nodes": (None, [P, I, P, P, P]),
        "ref_half_tilts": (I, [P, P]),
        "ref_set_tilt_open": (I, [P, I, P]),
        "ref_set_velocity": (None, [P, I, P, P]),
        "ref_get_velocity": (None, [P, I, P, P]),
        "ref_get_weights": (None, [P, I, P, P]),
        "ref_number": (I, [P]),
        "ref_get_pindex": (None, [P, I, P, P]),
        "ref_project": (I, [P, D, D, I]),
        "ref_resolve": (I, [P, P, P, P]),
        "ref_matrix_n": (I, [P]),
        "ref_matrix_nnz": (I, [P]),
        "ref_get_system": (None, [P, P, P, P, P, P]),
        "ref_update_velocity": (I, [P, P, I, D]),
        "ref_mic_factor": (I, [I, P, P, P, D, D, P, P, P, P]),
        "ref_mic_apply": (None, [I, P, P, P, D, D, P, P]),
        "ref_pcg_csr": (I, [I, P, P, P, P, P, D, I, D, D, P, P]),
        "ref_set_scalar": (None, [P, I, I, P, P]),
        "ref_get_scalar": (None, [P, I, I, P, P]),
        "ref_post_project": (I, [P]),
        "ref_scalar_update_ghost": (I, [P, I]),
        "ref_get_dual": (None, [P, I, P, P]),
        "ref_sample_points": (I, [P, I, I, P, D, P]),
        "ref_sample_scalar": (I, [P, I, I, P, P]),
        "ref_advect": (I, [P, D]),
        "ref_adjust_e": (I, [P, D, D, D, D]),
        "ref_set_tilt_adaptivity": (None, [P, I]),
        "ref_tilt_nodes_of": (None, [P, I, I, P, P, P]),
        "ref_set_source": (None, [P, D, D, D, D, D, D, D]),
        "ref_set_buoyancy": (None, [P, D, D]),
        "ref_step": (I, [P, D, D, I]),
        "ref_write_frame": (I, [P] + [ctypes.c_char_p] * 5),
        "ref_restart": (I, [P] + [ctypes.c_char_p] * 5),
        "ref_create_collider": (P, [I, D, I, P, I, D, D, D, D]),
        "ref_get_sdf": (None, [P, P]),
        "ref_sample_sdf": (I, [P, I, P, P]),
        "ref_constrain_and_ghost_serial": (I, [P]),
        "ref_post_project_serial": (I, [P]),
        "ref_wipe_boundary": (I, [P]),
        "ref_benchmark_reset": (None, []),
        "ref_benchmark_report": (I, [ctypes.c_char_p, I]),
        "ast_ref_last_error": (ctypes.c_char_p, []),
        "ast_ref_create": (P, [I, P, D]),
        "ast_ref_destroy": (None, [P]),
        "ast_ref_set_layout": (I, [P, P]),
        "ast_ref_set_velocity": (I, [P, P, P, P, P]),
        "ast_ref_velocity": (I, [P, P, P, P, P]),
        "ast_ref_project": (I, [P, D, D, P, P, I]),
        "ast_ref_system": (I, [P, P, P, P, P, P]),
        "ast_ref_numbering": (I, [P, P, P]),
        "ast_ref_solve_info": (I, [P, P, P]),
    }
    for name, (res, args) in sig.items():
        f = getattr(L, name)
        f.restype = res
        f.argtypes = args
    L.ref_quiet()
    _lib = L
    return L


def benchmark_reset():
    """Clears the reference's section timers (global_benchmark.cpp)."""
    lib().ref_benchmark_reset()


def benchmark_sections():
    """The reference's section timers since the last reset, as
    {"project": (seconds, calls), "project:A setup": (...), ...}: the report of
    BenchmarkTimer::report (timer.h:214-225), whose nesting is shown by 4
    spaces of indent per level, turned back into ':'-joined paths."""
    L = lib()
    n = L.ref_benchmark_report(None, 0)
    buf = ctypes.create_string_buffer(n + 1)
    L.ref_benchmark_report(buf, n + 1)
    out, stack = {}, []
    for line in buf.value.decode().splitlines():
        parts = line.split("\t")
        if len(parts) < 2 or parts[0] == "Full time":
            continue
        name = parts[0]
        depth = (len(name) - len(name.lstrip(" "))) // 4
        stack = stack[:depth] + [name.strip()]
        calls = int(parts[2]) if len(parts) > 2 else 1
        out[":".join(stack)] = (float(parts[1]), calls)
    return out


def _p(a):
    return None if a is None else a.ctypes.data_as(ctypes.c_void_p)


class RefSolver:
    """The reference OctreeOmniGridSolver3 driven through the harness."""

    def __init__(self, split, R, levels, closed_flags, collider=None):
        """collider: (cx, cy, cz, r) of the reference's sphere collider, or None
        for a constant positive boundary SDF (no collider)."""
        self.L = lib()
        split = np.ascontiguousarray(split, np.int32)
        if collider is None:
            self.h = self.L.ref_create(R, 1.0 / R, levels, _p(split), closed_flags)
        else:
            cx, cy, cz, r = collider
            self.h = self.L.ref_create_collider(R, 1.0 / R, levels, _p(split), closed_flags, cx, cy, cz, r)
        if not self.h:
            raise RuntimeError("ref_create failed: %s" % self.L.ref_last_error())
        self.levels = levels
        self.R = R

    def sdf(self):
        out = np.empty(self.R ** 3, np.float64)
        self.L.ref_get_sdf(self.h, _p(out))
        return out

    def sample_sdf(self, xyz):
        xyz = np.ascontiguousarray(xyz, np.float64).reshape(-1, 3)
        out = np.empty(len(xyz))
        self._chk(self.L.ref_sample_sdf(self.h, len(xyz), _p(xyz), _p(out)))
        return out

    def constrain_and_ghost_serial(self):
        self._chk(self.L.ref_constrain_and_ghost_serial(self.h))

    def post_project_serial(self):
        self._chk(self.L.ref_post_project_serial(self.h))

    def wipe_boundary(self):
        self._chk(self.L.ref_wipe_boundary(self.h))

    def close(self):
        if self.h:
            self.L.ref_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def level_size(self, l, ghost=True):
        return self.L.ref_level_size(self.h, l, 1 if ghost else 0)

    def nodes(self, l):
        n = self.level_size(l)
        xyz = np.empty((n, 3), np.int32)
        fl = np.empty(n, np.int32)
        self.L.ref_nodes(self.h, l, _p(xyz), _p(fl))
        return xyz, fl

    def tilt(self, l, which=0):
        """Tilt state per layer node; which=1 reads the state adjustE writes."""
        n = self.level_size(l)
        e = np.empty(n, np.float64)
        bits = np.empty(n, np.int32)
        hd = np.empty(n, np.int32)
        self.L.ref_tilt_nodes_of(self.h, which, l, _p(e), _p(bits), _p(hd))
        return e, bits, hd

    def half_tilts(self):
        n = self.L.ref_half_tilts(self.h, None)
        out = np.empty((n, 4), np.int32)
        self.L.ref_half_tilts(self.h, _p(out))
        return out

    def set_tilt_open(self, l, mask):
        m = np.ascontiguousarray(mask, np.uint8)
        if self.L.ref_set_tilt_open(self.h, l, _p(m)):
            raise RuntimeError(self.L.ref_last_error())

    def number(self):
        n = self.L.ref_number(self.h)
        if n < 0:
            raise RuntimeError(self.L.ref_last_error())
        return n

    def pindex(self, l):
        n = self.level_size(l)
        o = np.empty(n, np.int32)
        t = np.empty(n, np.int32)
        self.L.ref_get_pindex(self.h, l, _p(o), _p(t))
        return o, t

    def weights(self, l):
        n = self.level_size(l)
        m = np.empty((n, 3), np.float64)
        t = np.empty((n, 8), np.float64)
        self.L.ref_get_weights(self.h, l, _p(m), _p(t))
        return m, t

    def set_velocity(self, l, mac, tilt8):
        self.L.ref_set_velocity(self.h, l, _p(np.ascontiguousarray(mac, np.float64)),
                                _p(np.ascontiguousarray(tilt8, np.float64)))

    def velocity(self, l):
        n = self.level_size(l)
        m = np.empty((n, 3), np.float64)
        t = np.empty((n, 8), np.float64)
        self.L.ref_get_velocity(self.h, l, _p(m), _p(t))
        return m, t

    # ---- post-projection chain, sampling, advection ----
    def _chk(self, rc):
        if rc != 0:
            raise RuntimeError(self.L.ref_last_error().decode())

    def set_scalar(self, field, l, oct_, tilt):
        self.L.ref_set_scalar(self.h, field, l, _p(np.ascontiguousarray(oct_, np.float64)),
                              _p(np.ascontiguousarray(tilt, np.float64)))

    def scalar(self, field, l):
        n = self.level_size(l)
        o, t = np.empty(n), np.empty(n)
        self.L.ref_get_scalar(self.h, field, l, _p(o), _p(t))
        return o, t

    def post_project(self):
        self._chk(self.L.ref_post_project(self.h))

    def scalar_update_ghost(self, field):
        self._chk(self.L.ref_scalar_update_ghost(self.h, field))

    def dual(self, l):
        n = self.level_size(l)
        o, t = np.empty((n, 3)), np.empty((n, 3))
        self.L.ref_get_dual(self.h, l, _p(o), _p(t))
        return o, t

    def sample_points(self, what, xyz, dt=0.0):
        xyz = np.ascontiguousarray(xyz, np.float64)
        out = np.empty_like(xyz)
        self._chk(self.L.ref_sample_points(self.h, what, len(xyz), _p(xyz), dt, _p(out)))
        return out

    def sample_scalar(self, field, xyz):
        xyz = np.ascontiguousarray(xyz, np.float64)
        out = np.empty(len(xyz))
        self._chk(self.L.ref_sample_scalar(self.h, field, len(xyz), _p(xyz), _p(out)))
        return out

    def advect(self, dt):
        self._chk(self.L.ref_advect(self.h, dt))

    def adjust_e(self, d_threshold, dt, v_threshold, max_threshold):
        self._chk(self.L.ref_adjust_e(self.h, d_threshold, dt, v_threshold, max_threshold))

    def set_tilt_adaptivity(self, on):
        self.L.ref_set_tilt_adaptivity(self.h, 1 if on else 0)

    def set_source(self, center, radius, density=1.0, heat=1.0, speed=1.0):
        self.L.ref_set_source(self.h, center[0], center[1], center[2], radius, density, heat, speed)

    def set_buoyancy(self, smoke_density_factor, temperature_factor):
        self.L.ref_set_buoyancy(self.h, smoke_density_factor, temperature_factor)

    def write_frame(self, layout=None, tilt=None, density=None, heat=None, velocity=None):
        """The reference's frame dumps (grid_dump.cpp); None skips a file."""
        enc = lambda q: None if q is None else os.fsencode(q)
        self._chk(self.L.ref_write_frame(self.h, enc(layout), enc(tilt), enc(density), enc(heat), enc(velocity)))

    def restart(self, layout, tilt, density, heat, velocity):
        """The reference's restart from frame files (writes a preview image
        under ./astg_ref_harness/ in the working directory)."""
        enc = os.fsencode
        self._chk(self.L.ref_restart(self.h, enc(layout), enc(tilt), enc(density), enc(heat), enc(velocity)))

    def step(self, dt, tol=1e-6, maxit=3000):
        """One onUpdate(dt): sceneSetup, addForce, constrain, project, advect."""
        if self.L.ref_step(self.h, dt, tol, maxit):
            raise RuntimeError(self.L.ref_last_error())

    def project(self, dt, tol=1e-6, maxit=3000):
        if self.L.ref_project(self.h, dt, tol, maxit):
            raise RuntimeError(self.L.ref_last_error())

    def system(self):
        n = self.L.ref_matrix_n(self.h)
        nnz = self.L.ref_matrix_nnz(self.h)
        rs = np.empty(n + 1, np.int32)
        col = np.empty(nnz, np.int32)
        val = np.empty(nnz, np.float64)
        b = np.empty(n, np.float64)
        p = np.empty(n, np.float64)
        self.L.ref_get_system(self.h, _p(rs), _p(col), _p(val), _p(b), _p(p))
        return rs, col, val, b, p

    def resolve(self, n):
        x = np.empty(n, np.float64)
        r = ctypes.c_double()
        it = ctypes.c_int()
        if self.L.ref_resolve(self.h, _p(x), ctypes.byref(r), ctypes.byref(it)):
            raise RuntimeError(self.L.ref_last_error())
        return x, r.value, it.value

    def update_velocity(self, p, dt):
        p = np.ascontiguousarray(p, np.float64)
        if self.L.ref_update_velocity(self.h, _p(p), len(p), dt):
            raise RuntimeError(self.L.ref_last_error())


def mic_factor(rs, col, val, tau=0.97, ratio=0.25):
    L = lib()
    rs = np.ascontiguousarray(rs, np.int32)
    col = np.ascontiguousarray(col, np.int32)
    val = np.ascontiguousarray(val, np.float64)
    n = len(rs) - 1
    nnz = L.ref_mic_factor(n, _p(rs), _p(col), _p(val), tau, ratio, None, None, None, None)
    inv = np.empty(n, np.float64)
    cs = np.empty(n + 1, np.uint32)
    ri = np.empty(nnz, np.uint32)
    v = np.empty(nnz, np.float64)
    L.ref_mic_factor(n, _p(rs), _p(col), _p(val), tau, ratio, _p(inv), _p(cs), _p(ri), _p(v))
    return inv, cs, ri, v


def mic_apply(rs, col, val, r, tau=0.97, ratio=0.25):
    L = lib()
    rs = np.ascontiguousarray(rs, np.int32)
    col = np.ascontiguousarray(col, np.int32)
    val = np.ascontiguousarray(val, np.float64)
    r = np.ascontiguousarray(r, np.float64)
    z = np.empty_like(r)
    L.ref_mic_apply(len(rs) - 1, _p(rs), _p(col), _p(val), tau, ratio, _p(r), _p(z))
    return z


def pcg_csr(rs, col, val, b, tol=1e-6, maxit=3000, tau=0.97, ratio=0.25):
    L = lib()
    rs = np.ascontiguousarray(rs, np.int32)
    col = np.ascontiguousarray(col, np.int32)
    val = np.ascontiguousarray(val, np.float64)
    b = np.ascontiguousarray(b, np.float64)
    x = np.empty_like(b)
    r = ctypes.c_double()
    it = ctypes.c_int()
    ok = L.ref_pcg_csr(len(rs) - 1, _p(rs), _p(col), _p(val), _p(b), _p(x), tol, maxit, tau, ratio,
                       ctypes.byref(r), ctypes.byref(it))
    if ok < 0:
        raise RuntimeError(L.ref_last_error())
    return x, r.value, it.value, bool(ok)


class RefAST:
    """The reference's uniform-layout AST solver (ast::ASTSolver<d>,
    include/ast_solver.h) through oracle/ast_harness.cpp. Shapes follow the
    reference's grids, x fastest: tilt nodes prod(res + 1), MAC faces of axis a
    prod(res + e_a), tilt velocities (prod(res + 1), 4 or 8) node-major."""

    def __init__(self, dim, res, dx):
        self.dim, self.res, self.dx = dim, tuple(int(r) for r in res), float(dx)
        r = np.array(self.res, dtype=np.int32)
        self.h = lib().ast_ref_create(dim, _p(r), self.dx)
        if not self.h:
            raise RuntimeError(lib().ast_ref_last_error().decode())
        self.kt = 4 if dim == 2 else 8

    def close(self):
        if self.h:
            lib().ast_ref_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _chk(self, rc):
        if rc:
            raise RuntimeError(lib().ast_ref_last_error().decode())

    def tilt_count(self):
        return int(np.prod([r + 1 for r in self.res]))

    def mac_count(self, a):
        return int(np.prod([r + (1 if i == a else 0) for i, r in enumerate(self.res)]))

    def set_layout(self, open_):
        o = np.ascontiguousarray(open_, dtype=np.uint8).reshape(-1)
        assert o.size == self.tilt_count()
        self._chk(lib().ast_ref_set_layout(self.h, _p(o)))

    def set_velocity(self, mac, tilt):
        m = [np.ascontiguousarray(mac[a], dtype=np.float64).reshape(-1) for a in range(self.dim)]
        t = np.ascontiguousarray(tilt, dtype=np.float64).reshape(-1)
        assert t.size == self.tilt_count() * self.kt
        for a in range(self.dim):
            assert m[a].size == self.mac_count(a)
        self._chk(lib().ast_ref_set_velocity(self.h, _p(m[0]), _p(m[1]), _p(m[2]) if self.dim == 3 else None, _p(t)))

    def velocity(self):
        m = [np.zeros(self.mac_count(a)) for a in range(self.dim)]
        t = np.zeros(self.tilt_count() * self.kt)
        self._chk(lib().ast_ref_velocity(self.h, _p(m[0]), _p(m[1]), _p(m[2]) if self.dim == 3 else None, _p(t)))
        return m, t.reshape(-1, self.kt)

    def project(self, dt, rho=2.0, report=True):
        """project(dt); report also re-solves to record the iteration count
        (solve_info), which doubles the solve time: pass False to time it."""
        n = np.zeros(1, np.int32)
        nnz = np.zeros(1, np.int64)
        self._chk(lib().ast_ref_project(self.h, float(dt), float(rho), _p(n), _p(nnz), 1 if report else 0))
        self.n, self.nnz = int(n[0]), int(nnz[0])
        return self.n, self.nnz

    def system(self):
        rs = np.zeros(self.n + 1, np.int32)
        col = np.zeros(self.nnz, np.int32)
        val = np.zeros(self.nnz)
        b = np.zeros(self.n)
        x = np.zeros(self.n)
        self._chk(lib().ast_ref_system(self.h, _p(rs), _p(col), _p(val), _p(b), _p(x)))
        return rs, col, val, b, x

    def numbering(self):
        cell = np.zeros(int(np.prod(self.res)), np.int32)
        tilt = np.zeros(self.tilt_count(), np.int32)
        self._chk(lib().ast_ref_numbering(self.h, _p(cell), _p(tilt)))
        return cell, tilt

    def solve_info(self):
        """(iterations, final infinity-norm residual) of the last projection's
        solve; iterations is -1 if the re-solve did not reproduce it."""
        it = np.zeros(1, np.int32)
        err = np.zeros(1)
        self._chk(lib().ast_ref_solve_info(self.h, _p(it), _p(err)))
        return int(it[0]), float(err[0])
