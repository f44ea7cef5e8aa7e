"""Load paths of the device Newton driver (mpf_newton_solve: solveWithNewtonRaphson with MyTime's cut
and stack, femSolidmechanics2.cpp:730-851, MyTime.cpp:26-90) that the plain runs never take, forced
from the host so that every kernel exits normally:
  - the full-scale block's load case (bottom fixed, top UZ = -0.3 lam, nu = 0.49, 15 steps,
    block-Jacobi PCG) on an 8^3 block: every step and its Newton iteration count as the oracle's
    direct-solve Newton, the final state within 1e-9;
  - a load step too large to converge (an inverted element or an indefinite tangent in its first
    iterations): the step is cut and the sub-steps reach the oracle's equilibrium within 1e-9;
  - too few iterations allowed (itmax): cut and reset, same end state;
  - a NaN in the state: the divergence test ends the step, the cuts run down to 1% of the step and
    the driver reports MPF_ERR_DIVERGED without touching the last converged state's size."""
from __future__ import annotations

import numpy as np
import pytest

from tests.cases import normwise_rel

pytestmark = pytest.mark.gpu


def _block(mpf, orc, n, nu, val, jitter=0.0, seed=5, dof=2):
    X, conn = mpf.mesh_hex(n, n, n, jitter=jitter, seed=seed)
    m = mpf.matl_from_E_nu(1e3, nu, tangent=mpf.TANGENT_CONSISTENT)
    bot, top = mpf.face_nodes(mpf.HEX8, n, n, n, 4), mpf.face_nodes(mpf.HEX8, n, n, n, 5)
    P = mpf.Problem(mpf.HEX8, X, conn, m)
    O = orc.Problem(8, X, conn, m.mu, m.Kinv, 3)
    O.set_tangent(True)
    for R in (P, O):
        R.add_dirichlet(bot, -1)
        R.add_dirichlet(top, dof, val)
        R.prepare_matrix_pattern()
    return P, O


def test_full_scale_load_case_reduced(mpf, orc, gpu):
    P, O = _block(mpf, orc, 8, 0.49, -0.3)
    opts = P.newton_opts(nsteps=15, dt=1.0 / 15, itmax=30, rtol=1e-9, ksp=mpf.KSP_CG, pc=mpf.PC_BJACOBI3,
                         lin_rtol=1e-12, maxit=20000)
    iters, rep = P.solve_with_newton_raphson(opts)
    st, oit, _ = O.newton(15, 1.0 / 15, solver=0, rtol=1e-9, itmax=30)
    assert st == 0 and rep.steps_done == 15
    assert np.array_equal(iters, oit), (iters, oit)
    u, _ = P.state()
    uo, _ = O.state()
    assert normwise_rel(u, uo) < 1e-9


@pytest.mark.parametrize("ksp", ["band", "cg"])
def test_oversized_step_is_cut(mpf, orc, gpu, ksp):
    """The top face sheared by 0.8 in one step: the first Newton iterate inverts an element (the
    oracle's Newton reports the negative Jacobian at that step size); the driver cuts the step and
    the two halves converge to the oracle's two-step solution."""
    P, O = _block(mpf, orc, 3, 0.3, 0.8, jitter=0.05, dof=0)
    lin = dict(ksp=mpf.KSP_BAND) if ksp == "band" else dict(ksp=mpf.KSP_CG, pc=mpf.PC_BJACOBI3, lin_rtol=1e-13)
    opts = P.newton_opts(nsteps=1, dt=1.0, itmax=30, rtol=1e-11, **lin)
    iters, rep = P.solve_with_newton_raphson(opts)
    assert rep.steps_done == 2  # the single step of the load case was cut in two
    st, _, _ = O.newton(2, 0.5, solver=0, rtol=1e-12, itmax=30)
    assert st == 0
    assert normwise_rel(P.state()[0], O.state()[0]) < 1e-9


def test_itmax_forces_cut_and_reset(mpf, orc, gpu):
    n = 3
    X, conn = mpf.mesh_hex(n, n, n, jitter=0.05, seed=5)
    m = mpf.matl_from_E_nu(1e3, 0.3, tangent=mpf.TANGENT_CONSISTENT)
    P = mpf.Problem(mpf.HEX8, X, conn, m)
    O = orc.Problem(8, X, conn, m.mu, m.Kinv, 3)
    O.set_tangent(True)
    for R in (P, O):
        R.add_dirichlet(mpf.face_nodes(mpf.HEX8, n, n, n, 0), -1)
        R.add_dirichlet(mpf.face_nodes(mpf.HEX8, n, n, n, 1), 0, 0.1)
        R.prepare_matrix_pattern()
    # a 0.25 step takes 4 assemblies: with 3 allowed every full step fails, is reset and cut
    opts = P.newton_opts(nsteps=4, dt=0.25, itmax=3, rtol=1e-10, ksp=mpf.KSP_BAND)
    iters, rep = P.solve_with_newton_raphson(opts)
    assert rep.steps_done > 4
    st, _, _ = O.newton(8, 0.125, solver=0, rtol=1e-12, itmax=30)
    assert st == 0
    assert normwise_rel(P.state()[0], O.state()[0]) < 1e-9


def test_nan_state_diverges(mpf, orc, gpu):
    P, _ = _block(mpf, orc, 3, 0.3, -0.1)
    u = np.zeros(3 * P.nNode)
    u[5] = np.nan
    P.set_state(disp=u, disp_prev=u)
    opts = P.newton_opts(nsteps=2, dt=0.5, itmax=10, rtol=1e-10, ksp=mpf.KSP_CG, pc=mpf.PC_BJACOBI3)
    with pytest.raises(mpf.MpfError) as e:
        P.solve_with_newton_raphson(opts)
    assert e.value.status == mpf.ERR_DIVERGED
