"""Writes tests/golden/ast{2,3}_*.npz: the reference's uniform-layout AST
projection (ast::ASTSolver<d>::project, include/ast_solver.h:322-498) on
seeded inputs, through oracle/ast_harness.cpp.

Runs only where oracle/_ref/libastgref.so exists (the reference built from its
own sources by oracle/ref.mk). Each fixture is data only:
  inputs  dim, res, dx, dt, rho, open (tilt-node flags, before the boundary
          tilts are closed), mac_in_<a>, tilt_in (node-major, 4 or 8 per node);
  outputs cell_index, tilt_index (numbering), rowstart, col, val, b (the
          assembled system), p (the solve), iters (its iteration count),
          mac_out_<a>, tilt_out (the corrected and constrained velocity).

    python tests/golden/make_golden_ast.py
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

from oracle import ref  # noqa: E402

# name, dimension, resolution, share of tilt nodes opened, seed
CASES = [
    ("ast2_r64x48", 2, (64, 48), 0.3, 0),
    ("ast2_r40_dense", 2, (40, 40), 0.85, 1),
    ("ast2_r32_closed", 2, (32, 32), 0.0, 2),
    ("ast3_r16x16x12", 3, (16, 16, 12), 0.3, 3),
    ("ast3_r12x10x8_dense", 3, (12, 10, 8), 0.7, 4),
]
DT, RHO = 0.02, 2.0


def main():
    for name, dim, res, frac, seed in CASES:
        rng = np.random.default_rng(seed)
        dx = 1.0 / max(res)
        r = ref.RefAST(dim, res, dx)
        open_ = (rng.random(r.tilt_count()) < frac).astype(np.uint8)
        mac = [rng.standard_normal(r.mac_count(a)) for a in range(dim)]
        tilt = rng.standard_normal((r.tilt_count(), r.kt))
        r.set_layout(open_)
        r.set_velocity(mac, tilt)
        n, nnz = r.project(DT, RHO)
        rs, col, val, b, p = r.system()
        iters, resid = r.solve_info()
        assert iters >= 0, "the reference's solve did not reproduce"
        mac_o, tilt_o = r.velocity()
        cell, tind = r.numbering()
        out = dict(dim=dim, res=np.array(res), dx=dx, dt=DT, rho=RHO, open=open_, tilt_in=tilt,
                   cell_index=cell, tilt_index=tind, rowstart=rs, col=col, val=val, b=b, p=p, iters=iters,
                   residual=resid, tilt_out=tilt_o)
        for a in range(dim):
            out["mac_in_%d" % a] = mac[a]
            out["mac_out_%d" % a] = mac_o[a]
        np.savez_compressed(os.path.join(HERE, name + ".npz"), **out)
        r.close()
        print("%s: n=%d nnz=%d iterations %d residual %.2e" % (name, n, nnz, iters, resid))


if __name__ == "__main__":
    main()
