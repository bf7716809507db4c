"""Python bindings of the AST-octree projection library (include/astg.h).

The library is ``libastg.so`` next to this file, built by
``ast-grid_amd/build.py``. Bindings are plain ctypes over the C ABI; numpy
arrays are passed as host pointers. Loading fails loudly when the library is
missing: there is no fallback implementation.
"""
import atexit
import ctypes
import os
import weakref

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libastg.so")

ASTG_OK = 0
ASTG_ERR_ARG = -1
ASTG_ERR_HIP = -2
ASTG_ERR_NOT_CONVERGED = -3
ASTG_ERR_TOPOLOGY = -4
ASTG_ERR_STATE = -5
ASTG_ERR_IO = -6

WALL_XNEG, WALL_XPOS, WALL_ZNEG, WALL_ZPOS, WALL_YNEG, WALL_YPOS = 1, 2, 4, 8, 16, 32
PREC_FP64, PREC_MIXED = 0, 1
ORDER_AUTO, ORDER_REFERENCE, ORDER_BLOCK_COLOUR = 0, 1, 2
TAU_AUTO = -1.0  # mic_tau: 0.97 (the reference's MIC(0)) in FP64, 0 (IC(0)) in the mixed form
NORM_INF, NORM_L2 = 0, 1

# Header symbols, in the order include/astg.h declares them.
EXPORTS = [
    "astg_pcg_params_default", "astg_create", "astg_destroy", "astg_last_error",
    "astg_pcg_solve_csr", "astg_pcg_solve_batch", "astg_batch_tuning_default", "astg_pcg_set_batch_tuning",
    "astg_pcg_tune_batch", "astg_write_frame", "astg_read_frame", "astg_split_spheres", "astg_set_topology_split", "astg_set_tilt_open",
    "astg_level_info", "astg_get_nodes", "astg_get_tilt", "astg_get_half_tilts", "astg_get_numbering",
    "astg_get_weights", "astg_set_velocity", "astg_get_velocity", "astg_project", "astg_get_system",
    "astg_get_mic_factor", "astg_fields_snapshot", "astg_kernel_time", "astg_sweep_timeline",
    "astg_set_scalar", "astg_get_scalar", "astg_post_project", "astg_scalar_update_ghost", "astg_get_dual",
    "astg_sample_points", "astg_sample_points_noalloc", "astg_apply_sphere_source", "astg_add_force", "astg_adjust_e",
    "astg_set_tilt_adaptivity", "astg_advect", "astg_step",
    "astg_set_boundary_sdf", "astg_sphere_sdf", "astg_wipe_boundary", "astg_split_distance",
    "astg_apply_sphere_source_velocity",
    "astg_field_layout", "astg_fields_upload", "astg_fields_download", "astg_project_fields",
    "astg_diff_state_size", "astg_diff_get_state", "astg_diff_set_state", "astg_diff_step",
    "astg_diff_step_adjoint", "astg_diff_apply",
    "astg_ast_params_default", "astg_ast_create", "astg_ast_destroy", "astg_ast_last_error", "astg_ast_set_layout",
    "astg_ast_set_velocity", "astg_ast_get_velocity", "astg_ast_project", "astg_ast_system", "astg_ast_tilt_index",
]
FIELD_DENSITY, FIELD_HEAT = 0, 1
SAMPLE_VELOCITY, SAMPLE_MAC, SAMPLE_BACKTRACE, SAMPLE_DENSITY, SAMPLE_HEAT, SAMPLE_SDF = range(6)
K_SPMV64, K_FWD64, K_BWD64, K_MIC, K_ASSEMBLE, K_SPMV32, K_FWD32, K_BWD32, K_COPY = range(9)


class PcgParams(ctypes.Structure):
    _fields_ = [("tol", ctypes.c_double), ("max_iterations", ctypes.c_int), ("norm", ctypes.c_int),
                ("mic_tau", ctypes.c_double), ("min_diag_ratio", ctypes.c_double),
                ("precision", ctypes.c_int), ("max_outer", ctypes.c_int), ("ordering", ctypes.c_int)]


class BatchTuning(ctypes.Structure):
    _fields_ = [("merge_levels", ctypes.c_int), ("stream_blocks_per_cu", ctypes.c_int),
                ("lanes_per_row", ctypes.c_int), ("ms_per_iteration", ctypes.c_double)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


class Fields(ctypes.Structure):
    """astg_fields: caller-owned field arrays in the context's flat layout."""
    _fields_ = [("mac", ctypes.c_void_p), ("tilt", ctypes.c_void_p), ("density_cell", ctypes.c_void_p),
                ("density_vert", ctypes.c_void_p), ("heat_cell", ctypes.c_void_p), ("heat_vert", ctypes.c_void_p),
                ("device_pointers", ctypes.c_int), ("stream", ctypes.c_void_p)]


class PcgResult(ctypes.Structure):
    _fields_ = [("residual", ctypes.c_double), ("rel_residual", ctypes.c_double),
                ("iterations", ctypes.c_int), ("outer", ctypes.c_int), ("converged", ctypes.c_int),
                ("n", ctypes.c_int), ("nnz", ctypes.c_longlong), ("depth", ctypes.c_int),
                ("ms_assemble", ctypes.c_double), ("ms_factor", ctypes.c_double),
                ("ms_solve", ctypes.c_double), ("ms_velocity", ctypes.c_double),
                ("ms_total", ctypes.c_double)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


_lib = None


def lib():
    """Loads libastg.so (raises if it has not been built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError("libastg.so not built: run `python ast-grid_amd/build.py` (or __graft_entry__.build())")
    # One HIP runtime per process: PyTorch-ROCm ships its own libamdhip64
    # under the same soname, and a second runtime loaded beside it makes the
    # one loaded later see no GPU. Loading torch first lets the library bind
    # to torch's runtime, so torch tensors and streams are valid here.
    try:
        import torch  # noqa: F401
    except Exception:
        pass
    L = ctypes.CDLL(LIB_PATH)
    P = ctypes.c_void_p
    I = ctypes.c_int
    D = ctypes.c_double
    sig = {
        "astg_pcg_params_default": (None, [ctypes.POINTER(PcgParams)]),
        "astg_create": (I, [ctypes.POINTER(P), I, ctypes.c_uint32]),
        "astg_destroy": (I, [P]),
        "astg_last_error": (ctypes.c_char_p, [P]),
        "astg_pcg_solve_csr": (I, [P, I, P, P, P, P, P, ctypes.POINTER(PcgParams), ctypes.POINTER(PcgResult)]),
        "astg_pcg_solve_batch": (I, [P, I, P, P, P, P, P, P, ctypes.POINTER(PcgParams), ctypes.POINTER(PcgResult)]),
        "astg_batch_tuning_default": (None, [ctypes.POINTER(BatchTuning)]),
        "astg_pcg_set_batch_tuning": (I, [P, ctypes.POINTER(BatchTuning)]),
        "astg_pcg_tune_batch": (I, [P, I, P, P, P, P, P, ctypes.POINTER(PcgParams), I,
                                    ctypes.POINTER(BatchTuning)]),
        "astg_split_spheres": (I, [I, I, I, P, P, P]),
        "astg_write_frame": (I, [P, ctypes.c_char_p, ctypes.c_char_p, ctypes.c_char_p, ctypes.c_char_p,
                                 ctypes.c_char_p]),
        "astg_read_frame": (I, [P, ctypes.c_char_p, ctypes.c_char_p, ctypes.c_char_p, ctypes.c_char_p,
                                ctypes.c_char_p]),
        "astg_set_topology_split": (I, [P, I, D, I, P, I]),
        "astg_set_tilt_open": (I, [P, I, P]),
        "astg_level_info": (I, [P, I, P, P]),
        "astg_get_nodes": (I, [P, I, P, P]),
        "astg_get_tilt": (I, [P, I, P, P, P]),
        "astg_get_half_tilts": (I, [P, P, P]),
        "astg_get_numbering": (I, [P, I, P, P, P]),
        "astg_get_weights": (I, [P, I, P, P]),
        "astg_set_velocity": (I, [P, I, P, P]),
        "astg_get_velocity": (I, [P, I, P, P]),
        "astg_project": (I, [P, D, D, ctypes.POINTER(PcgParams), ctypes.POINTER(PcgResult)]),
        "astg_get_system": (I, [P, P, P, P, P, P]),
        "astg_get_mic_factor": (I, [P, P, P, P, P, P]),
        "astg_fields_snapshot": (I, [P, I]),
        "astg_kernel_time": (I, [P, I, I, P, P]),
        "astg_sweep_timeline": (I, [P, I, P, P]),
        "astg_set_scalar": (I, [P, I, I, P, P]),
        "astg_get_scalar": (I, [P, I, I, P, P]),
        "astg_post_project": (I, [P]),
        "astg_scalar_update_ghost": (I, [P, I]),
        "astg_get_dual": (I, [P, I, P, P]),
        "astg_sample_points": (I, [P, I, I, P, D, P]),
        "astg_sample_points_noalloc": (I, [P, I, I, P, D, P]),
        "astg_apply_sphere_source": (I, [P, D, D, D, D, D, D, D, I]),
        "astg_add_force": (I, [P, D, D, D]),
        "astg_adjust_e": (I, [P, D, D, D, D]),
        "astg_set_tilt_adaptivity": (I, [P, I]),
        "astg_advect": (I, [P, D]),
        "astg_step": (I, [P, D, D, ctypes.POINTER(PcgParams), ctypes.POINTER(PcgResult)]),
        "astg_set_boundary_sdf": (I, [P, P, P]),
        "astg_sphere_sdf": (I, [I, D, D, D, D, D, P]),
        "astg_wipe_boundary": (I, [P]),
        "astg_split_distance": (I, [I, I, I, P, P, P]),
        "astg_apply_sphere_source_velocity": (I, [P, D, D, D, D, D, D, P, I]),
        "astg_field_layout": (I, [P, P, P]),
        "astg_fields_upload": (I, [P, ctypes.POINTER(Fields)]),
        "astg_fields_download": (I, [P, ctypes.POINTER(Fields)]),
        "astg_project_fields": (I, [P, D, D, ctypes.POINTER(PcgParams), ctypes.POINTER(Fields),
                                    ctypes.POINTER(PcgResult)]),
        "astg_ast_params_default": (None, [ctypes.POINTER(PcgParams)]),
        "astg_ast_create": (I, [ctypes.POINTER(P), I, I, P, D]),
        "astg_ast_destroy": (I, [P]),
        "astg_ast_last_error": (ctypes.c_char_p, [P]),
        "astg_ast_set_layout": (I, [P, P]),
        "astg_ast_set_velocity": (I, [P, P, P, P, P]),
        "astg_ast_get_velocity": (I, [P, P, P, P, P]),
        "astg_ast_project": (I, [P, D, D, ctypes.POINTER(PcgParams), ctypes.POINTER(PcgResult)]),
        "astg_ast_system": (I, [P, I, ctypes.c_longlong, P, P, P, P, P]),
        "astg_ast_tilt_index": (I, [P, P]),
    }
    for name, (res, args) in sig.items():
        f = getattr(L, name)
        f.restype = res
        f.argtypes = args
    _lib = L
    return L


def _ptr(a):
    return None if a is None else a.ctypes.data_as(ctypes.c_void_p)


class AstgError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__("astg error %d: %s" % (code, msg))
        self.code = code


def default_params(**kw):
    p = PcgParams()
    lib().astg_pcg_params_default(ctypes.byref(p))
    for k, v in kw.items():
        setattr(p, k, v)
    return p


def split_spheres(R, levels, spheres, band):
    """Split map (R^3 int32, x fastest) refined toward a union of spheres."""
    sph = np.ascontiguousarray(np.asarray(spheres, np.float64).reshape(-1, 4))
    band = np.ascontiguousarray(np.asarray(band, np.float64))
    out = np.empty(R * R * R, np.int32)
    rc = lib().astg_split_spheres(R, levels, len(sph), _ptr(sph), _ptr(band), _ptr(out))
    if rc:
        raise AstgError(rc, "split_spheres")
    return out


def split_distance(R, levels, dist, band, shift=0):
    """Split map (R^3 int32, x fastest) refined where the distance field `dist`
    lies within band[l-1] of a level-l leaf, then balanced. dist holds
    (R >> shift)^3 values at the centres of cells 2^shift finest cells wide."""
    d = np.ascontiguousarray(np.asarray(dist, np.float32).reshape(-1))
    dr = R >> shift
    if d.size != dr ** 3:
        raise ValueError("dist must hold (R >> shift)^3 values")
    band = np.ascontiguousarray(np.asarray(band, np.float64))
    out = np.empty(R * R * R, np.int32)
    rc = lib().astg_split_distance(R, levels, shift, _ptr(d), _ptr(band), _ptr(out))
    if rc:
        raise AstgError(rc, "split_distance")
    return out


# Contexts still open when the interpreter exits are closed by an exit hook,
# in creation order, while the HIP runtime and every module are still loaded
# (not by __del__ during interpreter teardown, whose order is undefined).
_live = weakref.WeakValueDictionary()
_next_id = [0]


@atexit.register
def _close_live_contexts():
    for key in sorted(_live.keys()):
        c = _live.get(key)
        if c is not None:
            try:
                c.close()
            except Exception:
                pass


def sphere_sdf(R, center, radius, h0=None):
    """R^3 signed distances of a sphere at the finest cell centres (x fastest)."""
    out = np.empty(R ** 3, np.float64)
    cx, cy, cz = center
    rc = lib().astg_sphere_sdf(R, 1.0 / R if h0 is None else h0, cx, cy, cz, radius, _ptr(out))
    if rc:
        raise AstgError(rc, "sphere_sdf")
    return out


class Context:
    """One library context (device >= 0: HIP stream + device buffers; -1: host only)."""

    def __init__(self, device=0):
        self._L = lib()
        h = ctypes.c_void_p()
        rc = self._L.astg_create(ctypes.byref(h), device, 0)
        if rc:
            raise AstgError(rc, "astg_create(device=%d)" % device)
        self.h = h
        self.levels = 0
        self.device = device
        _next_id[0] += 1
        _live[_next_id[0]] = self

    def close(self):
        if self.h:
            self._L.astg_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _chk(self, rc, what, allow=()):
        if rc and rc not in allow:
            msg = self._L.astg_last_error(self.h)
            raise AstgError(rc, "%s: %s" % (what, msg.decode() if msg else ""))
        return rc

    # ---- topology ----
    def set_topology(self, split, R, levels, closed_flags, h0=None):
        split = np.ascontiguousarray(split, np.int32)
        assert split.size == R ** 3
        self._chk(self._L.astg_set_topology_split(self.h, R, 1.0 / R if h0 is None else h0, levels,
                                                  _ptr(split), closed_flags), "set_topology")
        self.levels = levels
        self.R = R
        self.closed = closed_flags

    def set_boundary_sdf(self, sdf, slip_center=None):
        """Collider SDF on the finest grid (R^3, x fastest, cell centres), or None
        to remove it; slip_center defaults to the reference's (0.5, 0.5, 0.45)."""
        if sdf is None:
            self._chk(self._L.astg_set_boundary_sdf(self.h, None, None), "set_boundary_sdf")
            return
        a = np.ascontiguousarray(sdf, np.float64).ravel()
        assert a.size == self.R ** 3
        c = None if slip_center is None else np.ascontiguousarray(slip_center, np.float64)
        self._chk(self._L.astg_set_boundary_sdf(self.h, _ptr(a), _ptr(c)), "set_boundary_sdf")

    # ---- fields in the caller's memory (flat layout, see field_layout) ----
    def field_layout(self):
        """(G, offsets): every level's nodes concatenated; level l at [offsets[l], offsets[l+1])."""
        G = ctypes.c_longlong()
        off = np.empty(self.levels + 1, np.int64)
        self._chk(self._L.astg_field_layout(self.h, ctypes.byref(G), _ptr(off)), "field_layout")
        return G.value, off

    def _fields(self, arrays, stream):
        f = Fields()
        dev = None
        for k in ("mac", "tilt", "density_cell", "density_vert", "heat_cell", "heat_vert"):
            a = arrays.get(k)
            if a is None:
                continue
            if hasattr(a, "data_ptr"):  # a torch tensor: device memory, float64, contiguous
                if a.dtype.__str__() != "torch.float64" or not a.is_contiguous():
                    raise ValueError("field %s must be a contiguous float64 tensor" % k)
                setattr(f, k, a.data_ptr())
                d = bool(a.is_cuda)
            else:
                if a.dtype != np.float64 or not a.flags.c_contiguous:
                    raise ValueError("field %s must be a contiguous float64 array" % k)
                setattr(f, k, a.ctypes.data)
                d = False
            if dev is not None and dev != d:
                raise ValueError("fields mix host and device memory")
            dev = d
        f.device_pointers = 1 if dev else 0
        f.stream = stream if dev else None
        return f

    @staticmethod
    def _torch_stream(stream):
        if stream is not None:
            return stream
        try:
            import torch
            return torch.cuda.current_stream().cuda_stream
        except Exception:
            return None

    def upload_fields(self, stream=None, **arrays):
        """Copies caller fields (numpy arrays or cuda torch tensors, flat layout) in;
        device copies run in the order of `stream` (default: torch's current stream)."""
        f = self._fields(arrays, self._torch_stream(stream))
        self._chk(self._L.astg_fields_upload(self.h, ctypes.byref(f)), "fields_upload")

    def download_fields(self, stream=None, **arrays):
        f = self._fields(arrays, self._torch_stream(stream))
        self._chk(self._L.astg_fields_download(self.h, ctypes.byref(f)), "fields_download")

    def project_fields(self, dt, rho, params, mac, tilt, stream=None, allow_unconverged=False):
        """project(dt) on caller-owned velocity arrays (updated in place)."""
        f = self._fields({"mac": mac, "tilt": tilt}, self._torch_stream(stream))
        res = PcgResult()
        rc = self._L.astg_project_fields(self.h, dt, rho, ctypes.byref(params), ctypes.byref(f), ctypes.byref(res))
        self._chk(rc, "project_fields", allow=(ASTG_ERR_NOT_CONVERGED,) if allow_unconverged else ())
        return res

    def wipe_boundary(self):
        """wipeBoundary: density and heat inside the collider set to 0."""
        self._chk(self._L.astg_wipe_boundary(self.h), "wipe_boundary")

    def set_tilt_open(self, level, mask):
        m = np.ascontiguousarray(mask, np.uint8)
        self._chk(self._L.astg_set_tilt_open(self.h, level, _ptr(m)), "set_tilt_open")

    def level_info(self, level):
        nn, nl = ctypes.c_int(), ctypes.c_int()
        self._chk(self._L.astg_level_info(self.h, level, ctypes.byref(nn), ctypes.byref(nl)), "level_info")
        return nn.value, nl.value

    def nodes(self, level):
        nn, _ = self.level_info(level)
        xyz = np.empty((nn, 3), np.int32)
        fl = np.empty(nn, np.int32)
        self._chk(self._L.astg_get_nodes(self.h, level, _ptr(xyz), _ptr(fl)), "get_nodes")
        return xyz, fl

    def tilt(self, level):
        nn, _ = self.level_info(level)
        e = np.empty(nn, np.float64)
        bits = np.empty(nn, np.int32)
        hd = np.empty(nn, np.int32)
        self._chk(self._L.astg_get_tilt(self.h, level, _ptr(e), _ptr(bits), _ptr(hd)), "get_tilt")
        return e, bits, hd

    def half_tilts(self):
        cnt = ctypes.c_int()
        self._chk(self._L.astg_get_half_tilts(self.h, None, ctypes.byref(cnt)), "get_half_tilts")
        out = np.empty((cnt.value, 4), np.int32)
        self._chk(self._L.astg_get_half_tilts(self.h, _ptr(out), ctypes.byref(cnt)), "get_half_tilts")
        return out

    def numbering(self, level):
        nn, _ = self.level_info(level)
        o = np.empty(nn, np.int32)
        t = np.empty(nn, np.int32)
        n = ctypes.c_int()
        self._chk(self._L.astg_get_numbering(self.h, level, _ptr(o), _ptr(t), ctypes.byref(n)), "get_numbering")
        return o, t, n.value

    def weights(self, level):
        nn, _ = self.level_info(level)
        m = np.empty((nn, 3), np.float64)
        t = np.empty((nn, 8), np.float64)
        self._chk(self._L.astg_get_weights(self.h, level, _ptr(m), _ptr(t)), "get_weights")
        return m, t

    # ---- fields ----
    def set_velocity(self, level, mac, tilt8):
        mac = np.ascontiguousarray(mac, np.float64)
        tilt8 = np.ascontiguousarray(tilt8, np.float64)
        self._chk(self._L.astg_set_velocity(self.h, level, _ptr(mac), _ptr(tilt8)), "set_velocity")

    def velocity(self, level):
        nn, _ = self.level_info(level)
        m = np.empty((nn, 3), np.float64)
        t = np.empty((nn, 8), np.float64)
        self._chk(self._L.astg_get_velocity(self.h, level, _ptr(m), _ptr(t)), "get_velocity")
        return m, t

    # ---- projection ----
    def project(self, dt, rho, params=None, allow_unconverged=False):
        prm = params if params is not None else default_params()
        res = PcgResult()
        self._chk(self._L.astg_project(self.h, dt, rho, ctypes.byref(prm), ctypes.byref(res)), "project",
                  allow=(ASTG_ERR_NOT_CONVERGED,) if allow_unconverged else ())
        return res

    def system(self, n, nnz):
        rs = np.empty(n + 1, np.int32)
        col = np.empty(nnz, np.int32)
        val = np.empty(nnz, np.float64)
        b = np.empty(n, np.float64)
        p = np.empty(n, np.float64)
        self._chk(self._L.astg_get_system(self.h, _ptr(rs), _ptr(col), _ptr(val), _ptr(b), _ptr(p)), "get_system")
        return rs, col, val, b, p

    def pattern(self):
        n = self.numbering(0)[2]
        rs = np.empty(n + 1, np.int32)
        self._chk(self._L.astg_get_system(self.h, _ptr(rs), None, None, None, None), "get_system")
        col = np.empty(rs[-1], np.int32)
        self._chk(self._L.astg_get_system(self.h, _ptr(rs), _ptr(col), None, None, None), "get_system")
        return rs, col

    def mic_factor(self, n):
        nnz = ctypes.c_longlong()
        self._chk(self._L.astg_get_mic_factor(self.h, None, None, None, None, ctypes.byref(nnz)), "get_mic_factor")
        inv = np.empty(n, np.float64)
        cs = np.empty(n + 1, np.uint32)
        ri = np.empty(nnz.value, np.uint32)
        v = np.empty(nnz.value, np.float64)
        self._chk(self._L.astg_get_mic_factor(self.h, _ptr(inv), _ptr(cs), _ptr(ri), _ptr(v), ctypes.byref(nnz)),
                  "get_mic_factor")
        return inv, cs, ri, v

    def snapshot(self, restore=False):
        self._chk(self._L.astg_fields_snapshot(self.h, 1 if restore else 0), "fields_snapshot")

    def kernel_time(self, which, reps=10):
        ms, by = ctypes.c_double(), ctypes.c_double()
        self._chk(self._L.astg_kernel_time(self.h, which, reps, ctypes.byref(ms), ctypes.byref(by)), "kernel_time")
        return ms.value, by.value

    # ---- the step around the projection and advection ----
    def set_scalar(self, field, level, oct_, tilt):
        o = np.ascontiguousarray(oct_, np.float64)
        t = np.ascontiguousarray(tilt, np.float64)
        self._chk(self._L.astg_set_scalar(self.h, field, level, _ptr(o), _ptr(t)), "set_scalar")

    def scalar(self, field, level):
        nn, _ = self.level_info(level)
        o, t = np.empty(nn), np.empty(nn)
        self._chk(self._L.astg_get_scalar(self.h, field, level, _ptr(o), _ptr(t)), "get_scalar")
        return o, t

    def post_project(self):
        self._chk(self._L.astg_post_project(self.h), "post_project")

    def scalar_update_ghost(self, field):
        self._chk(self._L.astg_scalar_update_ghost(self.h, field), "scalar_update_ghost")

    def dual(self, level):
        nn, _ = self.level_info(level)
        o, t = np.empty((nn, 3)), np.empty((nn, 3))
        self._chk(self._L.astg_get_dual(self.h, level, _ptr(o), _ptr(t)), "get_dual")
        return o, t

    def sample_points(self, what, xyz, dt=0.0, noalloc=False):
        """Point queries; noalloc=True stages them through the library's fixed
        device array (astg_sample_points_noalloc, at most 262144 points)."""
        xyz = np.ascontiguousarray(xyz, np.float64).reshape(-1, 3)
        out = np.empty((len(xyz), 3)) if what <= SAMPLE_BACKTRACE else np.empty(len(xyz))
        f = self._L.astg_sample_points_noalloc if noalloc else self._L.astg_sample_points
        self._chk(f(self.h, what, len(xyz), _ptr(xyz), dt, _ptr(out)), "sample_points")
        return out

    def apply_sphere_source(self, center, radius, density=1.0, heat=1.0, speed=1.0, update_ghost=False,
                            velocity=None):
        """The octree smoke example's sphere source (density, heat, +z inflow at
        `speed`); with `velocity` (3 values) that inflow vector instead."""
        cx, cy, cz = center
        if velocity is None:
            rc = self._L.astg_apply_sphere_source(self.h, cx, cy, cz, radius, density, heat, speed,
                                                  1 if update_ghost else 0)
        else:
            v = np.ascontiguousarray(np.asarray(velocity, np.float64).reshape(3))
            rc = self._L.astg_apply_sphere_source_velocity(self.h, cx, cy, cz, radius, density, heat, _ptr(v),
                                                           1 if update_ghost else 0)
        self._chk(rc, "apply_sphere_source")

    def add_force(self, dt, smoke_density_factor=-0.000625, temperature_factor=5.0):
        """addForce: buoyancy along +z (the reference's default factors)."""
        self._chk(self._L.astg_add_force(self.h, dt, smoke_density_factor, temperature_factor), "add_force")

    def adjust_e(self, d_threshold, dt, v_threshold, max_threshold):
        """adjustE: tilt open/close decision and commit; fields are kept."""
        self._chk(self._L.astg_adjust_e(self.h, d_threshold, dt, v_threshold, max_threshold), "adjust_e")

    def set_tilt_adaptivity(self, on):
        self._chk(self._L.astg_set_tilt_adaptivity(self.h, 1 if on else 0), "set_tilt_adaptivity")

    def advect(self, dt):
        self._chk(self._L.astg_advect(self.h, dt), "advect")

    def step(self, dt, rho, params=None, allow_unconverged=False):
        prm = params if params is not None else default_params()
        res = PcgResult()
        allow = (ASTG_ERR_NOT_CONVERGED,) if allow_unconverged else ()
        self._chk(self._L.astg_step(self.h, dt, rho, ctypes.byref(prm), ctypes.byref(res)), "step", allow)
        return res

    def sweep_timeline(self, which, n):
        """Per-row (gate, deps seen, published) device wall-clock ticks
        (100 MHz) of one FP64 sweep launch, and the rows' dependency levels."""
        t = np.empty((n, 3), np.int64)
        lev = np.empty(n, np.int32)
        self._chk(self._L.astg_sweep_timeline(self.h, which, t.ctypes.data_as(ctypes.c_void_p),
                                              lev.ctypes.data_as(ctypes.c_void_p)), "sweep_timeline")
        return t, lev

    def write_frame(self, layout=None, tilt=None, density=None, heat=None, velocity=None):
        """Writes the reference's binary frame files (None skips a file)."""
        enc = lambda p: None if p is None else os.fsencode(p)
        self._chk(self._L.astg_write_frame(self.h, enc(layout), enc(tilt), enc(density), enc(heat), enc(velocity)),
                  "write_frame")

    def read_frame(self, layout=None, tilt=None, density=None, heat=None, velocity=None):
        """restart from the reference's binary frame files (None skips a file)."""
        enc = lambda p: None if p is None else os.fsencode(p)
        self._chk(self._L.astg_read_frame(self.h, enc(layout), enc(tilt), enc(density), enc(heat), enc(velocity)),
                  "read_frame")

    def pcg_solve_csr(self, rowstart, col, val, b, params=None, allow_unconverged=False):
        rowstart = np.ascontiguousarray(rowstart, np.int32)
        col = np.ascontiguousarray(col, np.int32)
        val = np.ascontiguousarray(val, np.float64)
        b = np.ascontiguousarray(b, np.float64)
        n = len(rowstart) - 1
        x = np.empty(n, np.float64)
        prm = params if params is not None else default_params()
        res = PcgResult()
        self._chk(self._L.astg_pcg_solve_csr(self.h, n, _ptr(rowstart), _ptr(col), _ptr(val), _ptr(b), _ptr(x),
                                             ctypes.byref(prm), ctypes.byref(res)), "pcg_solve_csr",
                  allow=(ASTG_ERR_NOT_CONVERGED,) if allow_unconverged else ())
        return x, res

    @staticmethod
    def _batch_args(systems):
        K = len(systems)
        keep = []
        for rs, col, val, b in systems:
            keep.append((np.ascontiguousarray(rs, np.int32), np.ascontiguousarray(col, np.int32),
                         np.ascontiguousarray(val, np.float64), np.ascontiguousarray(b, np.float64)))
        n = np.array([len(s[0]) - 1 for s in keep], np.int32)
        arr = lambda items: (ctypes.c_void_p * K)(*[a.ctypes.data for a in items])
        cols = [arr([s[j] for s in keep]) for j in range(4)]
        return K, keep, n, arr, cols

    def pcg_solve_batch(self, systems, params=None, allow_unconverged=False):
        """Batched form of pcg_solve_csr: systems = [(rowstart, col, val, b), ...].
        Returns ([x_k], [PcgResult_k])."""
        K, keep, n, arr, cols = self._batch_args(systems)
        xs = [np.empty(int(nk), np.float64) for nk in n]
        res = (PcgResult * K)()
        prm = params if params is not None else default_params()
        self._chk(self._L.astg_pcg_solve_batch(self.h, K, _ptr(n), *cols, arr(xs), ctypes.byref(prm), res),
                  "pcg_solve_batch", allow=(ASTG_ERR_NOT_CONVERGED,) if allow_unconverged else ())
        return xs, list(res)

    def tune_batch(self, systems, params=None, iterations=20):
        """Autotuner over the batched launch shape; sets and returns the best."""
        K, keep, n, arr, cols = self._batch_args(systems)
        best = BatchTuning()
        prm = params if params is not None else default_params()
        self._chk(self._L.astg_pcg_tune_batch(self.h, K, _ptr(n), *cols, ctypes.byref(prm), iterations,
                                              ctypes.byref(best)), "tune_batch")
        return best

    def set_batch_tuning(self, merge_levels=2, stream_blocks_per_cu=8, lanes_per_row=8):
        t = BatchTuning(merge_levels, stream_blocks_per_cu, lanes_per_row, 0.0)
        self._chk(self._L.astg_pcg_set_batch_tuning(self.h, ctypes.byref(t)), "set_batch_tuning")
