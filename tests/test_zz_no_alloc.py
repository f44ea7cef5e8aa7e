"""mpf_solve_noalloc (CG with its work vectors in a library-owned device array, no allocation)
against mpf_solve on the same assembled systems, through the C ABI: two calls in a row, the second
on a smaller system, each within 1e-9 of mpf_solve's solution at rtol 1e-13; sizes above the
maximum and other Krylov methods are refused before any device work."""
from __future__ import annotations

import numpy as np
import pytest

from tests.cases import normwise_rel

pytestmark = pytest.mark.gpu


def _assembled(mpf, n, seed):
    X, conn = mpf.mesh_hex(n, n, n, jitter=0.05, seed=seed)
    P = mpf.Problem(mpf.HEX8, X, conn, mpf.matl_from_E_nu(1e3, 0.3))
    P.add_dirichlet(mpf.face_nodes(mpf.HEX8, n, n, n, 0), -1)
    P.add_dirichlet(mpf.face_nodes(mpf.HEX8, n, n, n, 1), 0, 0.1)
    P.prepare_matrix_pattern()
    P.set_boundary_conditions(0.5)
    P.calc_stiffness_and_residual(True, 0.5)
    return P


def test_solve_noalloc_matches_solve(mpf, gpu):
    for n, seed, pc in ((8, 3, mpf.PC_JACOBI), (5, 4, mpf.PC_NONE)):  # the second call is smaller
        print(f"[noalloc] {n}^3 Hex8 pc={pc}", flush=True)
        P = _assembled(mpf, n, seed)
        rep = P.factorise_and_solve(ksp=mpf.KSP_CG, pc=pc, rtol=1e-13, maxit=20000)
        assert rep.converged
        x_ref = P.solution()
        rep2 = P.solve_noalloc(pc=pc, rtol=1e-13, maxit=20000)
        assert rep2.converged, (rep2.iterations, rep2.relres)
        x = P.solution()
        assert normwise_rel(x, x_ref) < 1e-9
        # the residual of the returned solution, from the assembled system on the host
        vals, rhs = P.system()
        pat = P.pattern()
        Ax = np.zeros_like(rhs)
        for r in range(len(rhs)):
            a, b = pat.rowptr[r], pat.rowptr[r + 1]
            Ax[r] = vals[a:b] @ x[pat.colidx[a:b]]
        assert np.linalg.norm(rhs - Ax) / np.linalg.norm(rhs) < 1e-11


def test_solve_noalloc_refusals(mpf, gpu):
    P = _assembled(mpf, 3, 5)
    for ksp, pc in ((mpf.KSP_GMRES, mpf.PC_JACOBI), (mpf.KSP_CG, mpf.PC_BJACOBI3), (mpf.KSP_BAND, mpf.PC_NONE)):
        o = mpf.KrylovOpts(ksp, pc, 30, 100, 1e-8)
        r = mpf.KrylovReport()
        assert mpf.lib().mpf_solve_noalloc(P._h, mpf.C.byref(o), mpf.C.byref(r)) == mpf.ERR_ARG
    # a diagonal system one row above the maximum
    m = mpf.NOALLOC_MAX_DOFS + 1
    S = mpf.System(m, 0, np.arange(m + 1, dtype=np.int32), np.arange(m, dtype=np.int32))
    with pytest.raises(mpf.MpfError):
        S.solve_noalloc()
    S.close()
