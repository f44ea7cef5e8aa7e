"""Arc-length continuation on the GPU (csrc/arclength.hip) against the oracle's restatement of
femSolidmechanics::solveWithArclength (femSolidmechanics2.cpp:855-1039; SolverPetsc.cpp:671-757) on
the same cantilever under a dead tip shear: the same converged steps, load factors and arc lengths
within 1e-9, the same final displacement within 1e-9, with the direct solve (one factorisation and a
re-solve per iteration) and with CG."""
from __future__ import annotations

import numpy as np
import pytest

from tests.cases import normwise_rel
from tests.test_oracle import arclength_beam

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("ksp", ["band", "cg"])
def test_arclength_vs_oracle(mpf, orc, gpu, ksp):
    X, conn, m, fixed, quads, O = arclength_beam(orc, mpf)
    lam_o, ds_o, us_o, att_o = O.arclength(0.1, 1.0, 8, 30, 1e-10)
    P = mpf.Problem(mpf.HEX8, X, conn, mpf.matl_from_E_nu(1e4, 0.3, tangent=mpf.TANGENT_CONSISTENT))
    P.add_dirichlet(fixed, -1)
    P.add_traction(quads, 0.0, 400.0)
    P.prepare_matrix_pattern()
    kw = dict(ksp=mpf.KSP_BAND) if ksp == "band" else dict(ksp=mpf.KSP_CG, pc=mpf.PC_BJACOBI3, lin_rtol=1e-14)
    lam, ds, rep = P.solve_with_arclength(0.1, 1.0, 8, max_iters=30, tol=1e-10, **kw)
    assert rep.steps_done == len(lam_o) == 8 and rep.attempts == att_o
    assert normwise_rel(lam, lam_o) < 1e-9 and normwise_rel(ds, ds_o) < 1e-9
    u, _ = P.state()
    assert normwise_rel(u, us_o[-1]) < 1e-9


def test_arclength_refusals(mpf, gpu):
    n = 2
    X, conn = mpf.mesh_hex(n, n, n, jitter=0.0, seed=1)
    P = mpf.Problem(mpf.HEX8, X, conn, mpf.matl_from_E_nu(1e3, 0.3))
    P.add_dirichlet(mpf.face_nodes(mpf.HEX8, n, n, n, 0), -1)
    P.add_dirichlet(mpf.face_nodes(mpf.HEX8, n, n, n, 1), 0, 0.1)  # a nonzero prescribed value
    P.add_traction(mpf.hex_face_quads(n, n, n, 1), 1.0, 0.0)
    P.prepare_matrix_pattern()
    with pytest.raises(mpf.MpfError):
        P.solve_with_arclength(0.1, 1.0, 4)
    Q = mpf.Problem(mpf.HEX8, X, conn, mpf.matl_from_E_nu(1e3, 0.3))
    Q.add_dirichlet(mpf.face_nodes(mpf.HEX8, n, n, n, 0), -1)
    Q.prepare_matrix_pattern()  # no traction: nothing to continue along
    with pytest.raises(mpf.MpfError):
        Q.solve_with_arclength(0.1, 1.0, 4)
