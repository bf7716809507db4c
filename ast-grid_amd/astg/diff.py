"""Differentiable smoke steps on the AST octree as a torch.autograd.Function.

The forward runs astg_diff_step on the device (constrain + ghosts, MIC(0)-PCG
projection, post-projection chain, advection of density and velocity) with the
state in a torch tensor; the backward runs astg_diff_step_adjoint, whose
hand-written kernels apply the transposes of the step's linear passes, reuse
the symmetric PCG for the transposed solve and back-propagate through the
advection's interpolation weights and back-traced positions. States are kept
only every `every` steps; a segment is recomputed forward before its adjoint
steps run (checkpointed recomputation).

The state vector (astg_diff_state_size doubles, the context's flat layout):
MAC 3G | tilt 8G | dual cell 3G | dual vertex 3G | density cell G | density
vertex G.
"""
import ctypes

import numpy as np

from . import AstgError, PcgParams, PcgResult, lib


def _bind():
    L = lib()
    P, I, D = ctypes.c_void_p, ctypes.c_int, ctypes.c_double
    sig = {
        "astg_diff_state_size": (I, [P, ctypes.POINTER(ctypes.c_longlong)]),
        "astg_diff_get_state": (I, [P, P, I, P]),
        "astg_diff_set_state": (I, [P, P, I, P]),
        "astg_diff_step": (I, [P, D, D, ctypes.POINTER(PcgParams), ctypes.POINTER(PcgResult)]),
        "astg_diff_step_adjoint": (I, [P, D, D, ctypes.POINTER(PcgParams), P, P, P, I, P]),
        "astg_diff_apply": (I, [P, I, I, D, D, ctypes.POINTER(PcgParams), P, P]),
    }
    for name, (res, args) in sig.items():
        f = getattr(L, name)
        f.restype = res
        f.argtypes = args
    return L


OP_CONSTRAIN, OP_POST, OP_SCALAR_GHOST, OP_PROJECT, OP_ADVECT = range(5)


class DiffSim:
    """One context stepped as a differentiable map (dt, rho and PCG parameters fixed)."""

    def __init__(self, ctx, dt, rho, params):
        self.ctx = ctx
        self.L = _bind()
        self.dt, self.rho, self.params = dt, rho, params
        n = ctypes.c_longlong()
        self._chk(self.L.astg_diff_state_size(ctx.h, ctypes.byref(n)), "diff_state_size")
        self.n = n.value
        self.G = self.n // 19
        G = self.G
        self.regions = {"mac": (0, 3 * G), "tilt": (3 * G, 11 * G), "dual_cell": (11 * G, 14 * G),
                        "dual_vert": (14 * G, 17 * G), "density_cell": (17 * G, 18 * G),
                        "density_vert": (18 * G, 19 * G)}
        self.last = PcgResult()

    def _chk(self, rc, what):
        if rc:
            msg = self.L.astg_last_error(self.ctx.h)
            raise AstgError(rc, "%s: %s" % (what, msg.decode() if msg else ""))

    @staticmethod
    def _stream():
        import torch
        return torch.cuda.current_stream().cuda_stream

    def region(self, state, name):
        a, b = self.regions[name]
        return state[a:b]

    def get_state(self):
        import torch
        s = torch.empty(self.n, dtype=torch.float64, device="cuda")
        self._chk(self.L.astg_diff_get_state(self.ctx.h, s.data_ptr(), 1, self._stream()), "diff_get_state")
        return s

    def set_state(self, s):
        s = s.contiguous()
        self._chk(self.L.astg_diff_set_state(self.ctx.h, s.data_ptr(), 1, self._stream()), "diff_set_state")

    def set_state_host(self, s):
        s = np.ascontiguousarray(s, np.float64)
        self._chk(self.L.astg_diff_set_state(self.ctx.h, s.ctypes.data, 0, None), "diff_set_state")

    def get_state_host(self):
        s = np.empty(self.n)
        self._chk(self.L.astg_diff_get_state(self.ctx.h, s.ctypes.data, 0, None), "diff_get_state")
        return s

    def step(self, s):
        """The state after one step from s (a new tensor)."""
        self.set_state(s)
        self._chk(self.L.astg_diff_step(self.ctx.h, self.dt, self.rho, ctypes.byref(self.params),
                                        ctypes.byref(self.last)), "diff_step")
        return self.get_state()

    def adjoint(self, s_in, adj_out):
        """J^T adj_out, J the step's Jacobian at s_in."""
        import torch
        s_in = s_in.contiguous()
        adj_out = adj_out.contiguous()
        adj_in = torch.empty_like(adj_out)
        self._chk(self.L.astg_diff_step_adjoint(self.ctx.h, self.dt, self.rho, ctypes.byref(self.params),
                                                s_in.data_ptr(), adj_out.data_ptr(), adj_in.data_ptr(), 1,
                                                self._stream()), "diff_step_adjoint")
        return adj_in

    def apply(self, op, x, transpose=False):
        """One operator of the step on a host state vector (checks)."""
        x = np.ascontiguousarray(x, np.float64)
        y = np.empty_like(x)
        self._chk(self.L.astg_diff_apply(self.ctx.h, op, 1 if transpose else 0, self.dt, self.rho,
                                         ctypes.byref(self.params), x.ctypes.data, y.ctypes.data), "diff_apply")
        return y


def _function():
    import torch

    class Rollout(torch.autograd.Function):
        """density_cell(T) = F_T(...F_1(state0 with MAC = u0)); gradient w.r.t. u0."""

        @staticmethod
        def forward(fctx, u0, sim, state0, steps, every):
            a, b = sim.regions["mac"]
            s = state0.clone()
            s[a:b] = u0
            ckpt = {}
            for k in range(steps):
                if k % every == 0:
                    ckpt[k] = s
                s = sim.step(s)
            fctx.sim, fctx.ckpt, fctx.steps, fctx.every = sim, ckpt, steps, every
            c0, c1 = sim.regions["density_cell"]
            return s[c0:c1].clone()

        @staticmethod
        def backward(fctx, grad):
            sim, steps, every = fctx.sim, fctx.steps, fctx.every
            adj = torch.zeros(sim.n, dtype=torch.float64, device=grad.device)
            c0, c1 = sim.regions["density_cell"]
            adj[c0:c1] = grad
            for start in sorted(fctx.ckpt, reverse=True):
                end = min(start + every, steps)
                states = [fctx.ckpt[start]]
                for _ in range(start, end - 1):  # recompute the segment's inputs
                    states.append(sim.step(states[-1]))
                for k in reversed(range(start, end)):
                    adj = sim.adjoint(states[k - start], adj)
            a, b = sim.regions["mac"]
            return adj[a:b].clone(), None, None, None, None

    return Rollout


_ROLLOUT = None


def rollout(u0, sim, state0, steps, every=5):
    """Final cell density after `steps` differentiable steps from state0 with its
    MAC velocity replaced by u0 (a float64 cuda tensor of 3G entries); autograd
    gives d/du0 through the adjoint kernels, with checkpoints every `every` steps."""
    global _ROLLOUT
    if _ROLLOUT is None:
        _ROLLOUT = _function()
    return _ROLLOUT.apply(u0, sim, state0, steps, every)
