"""GPU: flexible GMRES with the two-level p-multigrid preconditioner (PC_PMG) on Tet10 and Tet10/P1
systems, against a sparse direct solve of the same assembled system (scipy) and, inside Newton,
against the oracle's Newton with its direct solve.

Bars: the solve stops at ||b - A x|| / ||b|| < rtol, checked on the returned x; x within 1e-6 of the
direct solution (the mixed system's condition number at n = 8 is about 1e8, so rtol 1e-12 leaves
that margin); converged Newton states within 1e-9 of the oracle's."""
from __future__ import annotations

import numpy as np
import pytest

from tests.cases import apply_bcs, normwise_rel, tet_compress_bcs

pytestmark = pytest.mark.gpu


def _system(mpf, etype, n, Kinv=1e-3, seed=2):
    X, conn = mpf.mesh_tet10(n, jitter=0.1, seed=seed)
    P = mpf.Problem(etype, X, conn, mpf.matl(1.0, Kinv, 3, mixed=etype == mpf.TET10P1))
    apply_bcs(P, tet_compress_bcs(n))
    P.prepare_matrix_pattern()
    P.set_boundary_conditions(1.0)
    P.calc_stiffness_and_residual(True, 1.0)
    return P, X, conn


def _csr(P):
    import scipy.sparse as sp
    vals, rhs = P.system()
    pat = P.pattern()
    return sp.csr_matrix((vals, pat.colidx, pat.rowptr), shape=(len(rhs),) * 2), rhs, pat


def _coarse_dofs(conn, nda):
    verts = np.unique(np.asarray(conn)[:, :4])
    return int((np.asarray(nda).reshape(-1, 3)[verts] >= 0).sum())


@pytest.mark.parametrize("n", [6, 8])
def test_pmg_mixed_vs_direct(mpf, gpu, n):
    import scipy.sparse.linalg as spl
    P, X, conn = _system(mpf, mpf.TET10P1, n)
    A, rhs, pat = _csr(P)
    x_ref = spl.spsolve(A.tocsc(), rhs)
    print(f"Tet10/P1 n={n}: {pat.dispDOF} + {pat.presDOF} dofs", flush=True)
    r = P.factorise_and_solve(ksp=mpf.KSP_GMRES, pc=mpf.PC_PMG, rtol=1e-12, maxit=2000, restart=30)
    st = P.pmg_stats()
    print("PMG:", r.iterations, "its relres", r.relres, "coarse", st.coarse_dofs, "coarse its",
          st.coarse_iterations, "lambda", st.lambda_max, flush=True)
    assert 1.5 < st.lambda_max < 3.0  # B^-1 Kuu of this system: largest eigenvalue 2.23 (scipy eigs)
    assert r.converged, (r.iterations, r.relres)
    du = P.solution()
    assert np.linalg.norm(A @ du - rhs) / np.linalg.norm(rhs) < 1e-11
    assert normwise_rel(du, x_ref) < 1e-6
    assert st.coarse_dofs == _coarse_dofs(conn, pat.NodeDofArray)
    assert st.vcycles == r.iterations
    assert st.coarse_iterations >= st.vcycles
    # far fewer iterations than the node-block Jacobi GMRES needs on the same system
    rb = P.factorise_and_solve(ksp=mpf.KSP_GMRES, pc=mpf.PC_BJACOBI3, rtol=1e-8, maxit=3000, restart=30)
    rp = P.factorise_and_solve(ksp=mpf.KSP_GMRES, pc=mpf.PC_PMG, rtol=1e-8, maxit=3000, restart=30)
    print("to 1e-8: block-Jacobi", rb.iterations, rb.converged, "PMG", rp.iterations, flush=True)
    assert rp.converged and (not rb.converged or rp.iterations * 3 < rb.iterations)


def test_pmg_displacement_tet10_vs_direct(mpf, gpu):
    import scipy.sparse.linalg as spl
    P, X, conn = _system(mpf, mpf.TET10, 6, Kinv=0.05)
    A, rhs, pat = _csr(P)
    assert pat.presDOF == 0
    x_ref = spl.spsolve(A.tocsc(), rhs)
    r = P.factorise_and_solve(ksp=mpf.KSP_GMRES, pc=mpf.PC_PMG, rtol=1e-12, maxit=2000, restart=30)
    print("Tet10 PMG:", r.iterations, r.relres, "lambda", P.pmg_stats().lambda_max, flush=True)
    assert 3.0 < P.pmg_stats().lambda_max < 3.7  # scipy eigs: 3.64
    assert r.converged
    du = P.solution()
    assert np.linalg.norm(A @ du - rhs) / np.linalg.norm(rhs) < 1e-11
    assert normwise_rel(du, x_ref) < 1e-8


def test_pmg_reproducible_and_parameters(mpf, gpu):
    P, _, _ = _system(mpf, mpf.TET10P1, 5)
    sols = []
    for _ in range(2):
        r = P.factorise_and_solve(ksp=mpf.KSP_GMRES, pc=mpf.PC_PMG, rtol=1e-12, maxit=2000, restart=30)
        assert r.converged
        sols.append(P.solution())
    assert np.array_equal(sols[0], sols[1])  # fixed-order reductions and Galerkin sums
    base = P.pmg_stats().vcycles
    P.set_pmg(degree=3, ratio=20.0, coarse_rtol=0.01, coarse_maxit=500)
    r = P.factorise_and_solve(ksp=mpf.KSP_GMRES, pc=mpf.PC_PMG, rtol=1e-12, maxit=2000, restart=30)
    assert r.converged
    print("other parameters vs default:", normwise_rel(P.solution(), sols[0]), flush=True)
    assert normwise_rel(P.solution(), sols[0]) < 1e-6  # two solves to 1e-12 of a system with cond ~1e8
    print("vcycles default", base, "degree 3 / coarse 1e-2", P.pmg_stats().vcycles, flush=True)
    for bad in (dict(degree=0), dict(degree=9), dict(ratio=1.0), dict(coarse_rtol=1.0), dict(coarse_maxit=0)):
        with pytest.raises(mpf.MpfError):
            P.set_pmg(**bad)


def test_pmg_rejected_outside_tet10_gmres(mpf, gpu):
    P, _, _ = _system(mpf, mpf.TET10P1, 3)
    with pytest.raises(mpf.MpfError):
        P.factorise_and_solve(ksp=mpf.KSP_CG, pc=mpf.PC_PMG)
    with pytest.raises(mpf.MpfError):
        P.pmg_stats()  # nothing built yet
    n = 3
    X, conn = mpf.mesh_hex(n, n, n)
    H = mpf.Problem(mpf.HEX8, X, conn, mpf.matl_from_E_nu(1e3, 0.3))
    H.add_dirichlet(mpf.face_nodes(mpf.HEX8, n, n, n, 0), -1, None)
    H.add_dirichlet(mpf.face_nodes(mpf.HEX8, n, n, n, 1), 0, 0.01)
    H.prepare_matrix_pattern()
    H.set_boundary_conditions(1.0)
    H.calc_stiffness_and_residual(True, 1.0)
    with pytest.raises(mpf.MpfError):
        H.factorise_and_solve(ksp=mpf.KSP_GMRES, pc=mpf.PC_PMG)


def test_newton_tet10p1_pmg_vs_oracle(mpf, orc, gpu):
    n = 3
    X, conn = mpf.mesh_tet10(n, jitter=0.1, seed=2)
    P = mpf.Problem(mpf.TET10P1, X, conn, mpf.matl(1.0, 1e-3, 3, mixed=True))
    O = orc.Problem(10, X, conn, 1.0, 1e-3, 3)
    for Q in (P, O):
        apply_bcs(Q, tet_compress_bcs(n))
        Q.prepare_matrix_pattern()
    opts = P.newton_opts(nsteps=2, dt=0.5, itmax=20, rtol=1e-11, ksp=mpf.KSP_GMRES, pc=mpf.PC_PMG,
                         lin_rtol=1e-13, maxit=5000, restart=30)
    iters, rep = P.solve_with_newton_raphson(opts)
    assert rep.steps_done == 2
    st, oit, _ = O.newton(2, 0.5, solver=2, precond=2, lin_rtol=1e-13, rtol=1e-11, itmax=20)
    assert st == 0
    ud, pd = P.state()
    uo, po = O.state()
    print("Tet10/P1 Newton with PMG: u", normwise_rel(ud, uo), "p", normwise_rel(pd, po), flush=True)
    assert normwise_rel(ud, uo) < 1e-9
    assert normwise_rel(pd, po) < 1e-9


def test_pmg_three_levels_vs_two(mpf, gpu):
    """The vertex level's own V-cycle (aggregates of 8 vertices below) against its CG solve: both
    FGMRES solves reach the tolerance on the assembled system; the aggregate level has between one
    and three dofs per aggregate of 8 vertex blocks."""
    import math
    P, X, conn = _system(mpf, mpf.TET10P1, 8)
    A, rhs, pat = _csr(P)
    its = {}
    for lv in (2, 3):
        P.set_pmg(levels=lv)
        r = P.factorise_and_solve(ksp=mpf.KSP_GMRES, pc=mpf.PC_PMG, rtol=1e-10, maxit=2000, restart=30)
        st = P.pmg_stats()
        print(f"levels {lv}: {r.iterations} its, relres {r.relres:.2e}, lowest-level CG its {st.coarse_iterations},"
              f" agg dofs {st.agg_dofs}, vertex lambda {st.lambda_vertex:.3f}", flush=True)
        assert r.converged and st.levels == lv
        du = P.solution()
        assert np.linalg.norm(A @ du - rhs) / np.linalg.norm(rhs) < 1e-9
        its[lv] = (r.iterations, st.coarse_iterations, st.agg_dofs, st.lambda_vertex)
    nda = np.asarray(pat.NodeDofArray).reshape(-1, 3)
    verts = np.unique(np.asarray(conn)[:, :4])
    nvb = int(((nda[verts] >= 0).any(axis=1)).sum())
    nagg = math.ceil(nvb / 8)
    assert nagg <= its[3][2] <= 3 * nagg
    assert 1.0 < its[3][3] < 3.0  # B^-1 Kc of the vertex level: about 1.9 (scipy power iteration)
    assert its[2][3] == 0.0
    with pytest.raises(mpf.MpfError):
        P.set_pmg(levels=4)
