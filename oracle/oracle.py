"""ctypes wrapper of oracle/liboracle.so, the CPU restatement of the
reference's projection (TEST INFRASTRUCTURE ONLY: used by tests/, smoke() and
bench.py's cpu_baseline leg as the checker, never by the product path).
"""
import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")
_lib = None


class OrcLevel(ctypes.Structure):
    _fields_ = [("nnode", ctypes.c_int), ("nleaf", ctypes.c_int), ("res", ctypes.c_int), ("h", ctypes.c_double),
                ("xyz", ctypes.c_void_p), ("flags", ctypes.c_void_p), ("tilt_bits", ctypes.c_void_p),
                ("tilt_hd", ctypes.c_void_p), ("pid_oct", ctypes.c_void_p), ("pid_tilt", ctypes.c_void_p),
                ("tilt_e", ctypes.c_void_p), ("w_mac", ctypes.c_void_p), ("w_tilt", ctypes.c_void_p),
                ("u_mac", ctypes.c_void_p), ("u_tilt", ctypes.c_void_p)]


def lib():
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError("oracle/liboracle.so not built: run `make -C oracle`")
    L = ctypes.CDLL(LIB_PATH)
    P, I, D = ctypes.c_void_p, ctypes.c_int, ctypes.c_double
    L.orc_last_error.restype = ctypes.c_char_p
    L.orc_project.restype = P
    L.orc_project.argtypes = [I, I, P, I, P, I, D, D, D, I, D, D]
    L.orc_result_nnz.restype = ctypes.c_long
    L.orc_result_nnz.argtypes = [P]
    L.orc_result_get.argtypes = [P, P, P, P, P, P, P, P, P]
    L.orc_result_free.argtypes = [P]
    L.orc_mic_factor.restype = ctypes.c_long
    L.orc_mic_factor.argtypes = [I, P, P, P, D, D, P, P, P, P]
    L.orc_mic_apply.argtypes = [I, P, P, P, D, D, P, P]
    L.orc_pcg.restype = I
    L.orc_pcg.argtypes = [I, P, P, P, P, P, D, I, D, D, P, P]
    _lib = L
    return L


def _p(a):
    return a.ctypes.data_as(ctypes.c_void_p)


def project(R, L, closed, levels, half, n, dt, rho, tol=1e-6, maxit=3000, tau=0.97, ratio=0.25):
    """levels: list of dicts with xyz, flags, nleaf, tilt_e, tilt_bits, tilt_hd,
    pid_oct, pid_tilt, w_mac (nnode,3), w_tilt (nnode,8), u_mac, u_tilt (updated
    in place on copies returned in the result)."""
    lb = lib()
    keep = []
    arr = (OrcLevel * L)()
    u_out = []
    for l, d in enumerate(levels):
        def c(a, t):
            a = np.ascontiguousarray(a, t)
            keep.append(a)
            return a
        um = c(np.array(d["u_mac"], np.float64, copy=True), np.float64)
        ut = c(np.array(d["u_tilt"], np.float64, copy=True), np.float64)
        u_out.append((um, ut))
        arr[l] = OrcLevel(len(d["xyz"]), int(d["nleaf"]), R >> l, (1.0 / R) * (1 << l),
                          _p(c(d["xyz"], np.int32)), _p(c(d["flags"], np.int32)), _p(c(d["tilt_bits"], np.int32)),
                          _p(c(d["tilt_hd"], np.int32)), _p(c(d["pid_oct"], np.int32)),
                          _p(c(d["pid_tilt"], np.int32)), _p(c(d["tilt_e"], np.float64)),
                          _p(c(d["w_mac"], np.float64)), _p(c(d["w_tilt"], np.float64)), _p(um), _p(ut))
    half = np.ascontiguousarray(half, np.int32).reshape(-1, 4)
    h = lb.orc_project(L, closed, ctypes.cast(arr, ctypes.c_void_p), len(half), _p(half), n, dt, rho, tol, maxit,
                       tau, ratio)
    if not h:
        raise RuntimeError("oracle: " + lb.orc_last_error().decode())
    nnz = lb.orc_result_nnz(h)
    rs = np.empty(n + 1, np.int32)
    col = np.empty(nnz, np.int32)
    val = np.empty(nnz, np.float64)
    b = np.empty(n, np.float64)
    p = np.empty(n, np.float64)
    it = ctypes.c_int()
    res = ctypes.c_double()
    ok = ctypes.c_int()
    lb.orc_result_get(h, _p(rs), _p(col), _p(val), _p(b), _p(p), ctypes.byref(it), ctypes.byref(res), ctypes.byref(ok))
    lb.orc_result_free(h)
    return {"rowstart": rs, "col": col, "val": val, "b": b, "p": p, "iters": it.value, "residual": res.value,
            "converged": bool(ok.value), "u": u_out}


def levels_from_context(ctx, u_levels=None):
    """Reference-form topology of a product context (host data, via the C ABI)."""
    L = ctx.levels
    out = []
    for l in range(L):
        xyz, fl = ctx.nodes(l)
        e, bits, hd = ctx.tilt(l)
        o, t, n = ctx.numbering(l)
        wm, wt = ctx.weights(l)
        um, ut = u_levels[l] if u_levels is not None else ctx.velocity(l)
        out.append({"xyz": xyz, "flags": fl, "nleaf": ctx.level_info(l)[1], "tilt_e": e, "tilt_bits": bits,
                    "tilt_hd": hd, "pid_oct": o, "pid_tilt": t, "w_mac": wm, "w_tilt": wt, "u_mac": um, "u_tilt": ut})
    return out, n


def levels_from_golden(g, inputs=True):
    L = int(g["levels"])
    out = []
    for l in range(L):
        out.append({"xyz": g["nodes_%d" % l].astype(np.int32), "flags": g["flags_%d" % l].astype(np.int32),
                    "nleaf": int(g["nleaf_%d" % l]), "tilt_e": g["tilt_e_%d" % l],
                    "tilt_bits": g["tilt_bits_%d" % l].astype(np.int32),
                    "tilt_hd": g["tilt_hd_%d" % l].astype(np.int32), "pid_oct": g["pid_oct_%d" % l],
                    "pid_tilt": g["pid_tilt_%d" % l], "w_mac": g["w_mac_%d" % l], "w_tilt": g["w_tilt_%d" % l],
                    "u_mac": g["u_mac_in_%d" % l], "u_tilt": g["u_tilt_in_%d" % l]})
    return out


def project_from_context(ctx, u_levels, dt, rho, tol=1e-6, maxit=3000):
    """Oracle projection of a product context's topology with the given inputs."""
    lv, n = levels_from_context(ctx, u_levels)
    return project(ctx.R, ctx.levels, ctx.closed, lv, ctx.half_tilts(), n, dt, rho, tol, maxit)


def mic_factor(rs, col, val, tau=0.97, ratio=0.25):
    lb = lib()
    rs = np.ascontiguousarray(rs, np.int32)
    col = np.ascontiguousarray(col, np.int32)
    val = np.ascontiguousarray(val, np.float64)
    n = len(rs) - 1
    nnz = lb.orc_mic_factor(n, _p(rs), _p(col), _p(val), tau, ratio, None, None, None, None)
    inv = np.empty(n)
    cs = np.empty(n + 1, np.uint32)
    ri = np.empty(nnz, np.uint32)
    v = np.empty(nnz)
    lb.orc_mic_factor(n, _p(rs), _p(col), _p(val), tau, ratio, _p(inv), _p(cs), _p(ri), _p(v))
    return inv, cs, ri, v


def mic_apply(rs, col, val, r, tau=0.97, ratio=0.25):
    lb = lib()
    rs = np.ascontiguousarray(rs, np.int32)
    col = np.ascontiguousarray(col, np.int32)
    val = np.ascontiguousarray(val, np.float64)
    r = np.ascontiguousarray(r, np.float64)
    z = np.empty_like(r)
    lb.orc_mic_apply(len(rs) - 1, _p(rs), _p(col), _p(val), tau, ratio, _p(r), _p(z))
    return z


def pcg(rs, col, val, b, tol=1e-6, maxit=3000, tau=0.97, ratio=0.25):
    lb = lib()
    rs = np.ascontiguousarray(rs, np.int32)
    col = np.ascontiguousarray(col, np.int32)
    val = np.ascontiguousarray(val, np.float64)
    b = np.ascontiguousarray(b, np.float64)
    x = np.empty_like(b)
    res = ctypes.c_double()
    it = ctypes.c_int()
    ok = lb.orc_pcg(len(rs) - 1, _p(rs), _p(col), _p(val), _p(b), _p(x), tol, maxit, tau, ratio, ctypes.byref(res),
                    ctypes.byref(it))
    return x, res.value, it.value, bool(ok)
