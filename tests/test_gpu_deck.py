"""Input decks run end to end on the GPU (csrc/deck.cpp): a generated 3D deck in the reference's format
(config + Gmsh 2.2 mesh + petsc_options.dat) is read, its problem built and its load steps solved
with the reference's load stepping and absolute tolerance; the converged displacements (and pressures)
agree with the oracle's Newton on the same mesh and loads within 1e-9.  The Tet10/P1 deck follows the
reference's Example 4 (doc/examples/SolidMechanics/Example4/example4.rst): symmetry planes, a
compressive traction on the top face, near-incompressible Neo-Hookean, ELEM_SOLID_3D_MIXED1."""
from __future__ import annotations

import numpy as np
import pytest

from tests.cases import normwise_rel, tet10_boundary_tris, write_config, write_gmsh22

pytestmark = pytest.mark.gpu


def test_deck_hex8_stretch_vs_oracle(mpf, orc, gpu, tmp_path):
    """4% stretch in 4 steps: the deck runs the reference's element tangent (its Newton converges
    linearly here, and not at all at 10% in 4 steps, as the oracle with that tangent); the converged
    state does not depend on the tangent, so the oracle runs the consistent one."""
    n = 4
    X, conn = mpf.mesh_hex(n, n, n, jitter=0.05, seed=2)
    patches = {"left": (2, 3, mpf.hex_face_quads(n, n, n, 0)), "right": (2, 3, mpf.hex_face_quads(n, n, n, 1))}
    write_gmsh22(tmp_path / "blk.msh", X, conn, 5, "block", patches)
    mu, Kinv = 384.6153846153846, 1.2e-3
    write_config(tmp_path / "config", "blk", "ELEM_SOLID_3D_DISP", mu, Kinv,
                 [("left", "fixed"), ("right", "specified", "UX", 0.04, 1)], [], 0.25, 1.0, 1e-9, max_iters=40,
                 petsc=["-ksp_type preonly", "-pc_type lu", "-pc_factor_mat_solver_type mumps"])
    D = mpf.Deck(tmp_path / "config")
    P = D.setup()
    rep = D.run(vtu_prefix=tmp_path / "out")
    assert rep.converged and rep.steps_done == 4 and rep.final_time == pytest.approx(1.0)
    assert rep.final_norm < 1e-9
    u, _ = P.state()
    O = orc.Problem(8, X, conn, mu, Kinv, 3)
    O.set_tangent(True)
    O.add_dirichlet(mpf.face_nodes(mpf.HEX8, n, n, n, 0), -1)
    O.add_dirichlet(mpf.face_nodes(mpf.HEX8, n, n, n, 1), 0, 0.04)
    O.prepare_matrix_pattern()
    st, _, _ = O.newton(4, 0.25, solver=0, rtol=1e-13, itmax=40)
    assert st == 0
    uo, _ = O.state()
    assert normwise_rel(u, uo) < 1e-9
    assert sorted(p.name for p in tmp_path.glob("out-*.vtu")) == [f"out-{k:06d}.vtu" for k in range(5)]
    D.close()


@pytest.mark.parametrize("petsc", [None, ["-ksp_type gmres", "-pc_type gamg", "-ksp_rtol 1e-12"]],
                         ids=["default", "gamg"])
def test_deck_tet10p1_example4_vs_oracle(mpf, orc, gpu, tmp_path, petsc):
    X, conn = mpf.mesh_tet10(2, jitter=0.1, seed=4)
    sym = {f"sym{'XYZ'[a]}": (2, 9, tet10_boundary_tris(X, conn, a, 0.0, -1.0)) for a in range(3)}
    top = tet10_boundary_tris(X, conn, 2, 1.0, 1.0)
    patches = dict(sym, pressure=(2, 9, top))
    write_gmsh22(tmp_path / "block3d.msh", X, conn, 11, "block", patches)
    mu, Kinv = 80.19367957863858, 2.494128406043830e-06
    p0 = -32.0
    bcs = [("symX", "specified", "UX", 0.0, None), ("symY", "specified", "UY", 0.0, None),
           ("symZ", "specified", "UZ", 0.0, None), ("pressure", "specified", "UX", 0.0, None),
           ("pressure", "specified", "UY", 0.0, None)]
    write_config(tmp_path / "config", "block3d", "ELEM_SOLID_3D_MIXED1", mu, Kinv, bcs, [("pressure", p0, 0.0, 1)],
                 0.2, 1.0, 1e-10, max_iters=20, petsc=petsc)
    D = mpf.Deck(tmp_path / "config")
    if petsc:  # a multigrid preconditioner maps to the two-level p-multigrid inside flexible GMRES
        assert D.info().lin.pc == mpf.PC_PMG
    P = D.setup()
    if petsc:
        assert (D.info().lin.ksp, D.info().lin.pc) == (mpf.KSP_GMRES, mpf.PC_PMG)
    rep = D.run()
    assert rep.converged and rep.steps_done == 5 and rep.final_norm < 1e-10
    u, p = P.state()
    O = orc.Problem(10, X, conn, mu, Kinv, 3)
    for name, dof in (("symX", 0), ("symY", 1), ("symZ", 2), ("pressure", 0), ("pressure", 1)):
        O.add_dirichlet(np.unique(patches[name][2]), dof, 0.0)
    O.add_traction(top, p0, 0.0)
    O.prepare_matrix_pattern()
    st, _, _ = O.newton(5, 0.2, solver=2, precond=2, lin_rtol=1e-14, rtol=1e-13, itmax=30)
    assert st == 0
    uo, po = O.state()
    assert normwise_rel(u, uo) < 1e-8 and normwise_rel(p, po) < 1e-8
    # the top face moved down under the compressive traction, the symmetry planes stayed
    assert u.reshape(-1, 3)[np.unique(top), 2].max() < 0
    assert np.abs(u.reshape(-1, 3)[np.unique(sym["symZ"][2]), 2]).max() == 0.0
    D.close()


def test_deck_arclength_vs_oracle(mpf, orc, gpu, tmp_path):
    """solvertype arclength: the deck's cantilever follows the oracle's arc-length path (the loads'
    time function is 1, so Fext formed at time 0 is the full traction).  Decks run the reference's
    element tangent, so the oracle does too: the same steps, cuts and load factors."""
    from tests.test_oracle import arclength_beam
    X, conn, m, fixed, quads, O = arclength_beam(orc, mpf, consistent=False)
    # 6 converged steps (the reference tangent's Newton misses some attempts): the run ends on a
    # converged step, so the final states compare
    lam_o, _, us_o, _ = O.arclength(0.1, 1.0, 6, 30, 1e-10)
    assert len(lam_o) == 6
    nx, ny, nz = 6, 2, 2
    patches = {"clamp": (2, 3, mpf.hex_face_quads(nx, ny, nz, 0)), "tip": (2, 3, quads)}
    write_gmsh22(tmp_path / "beam.msh", X, conn, 5, "block", patches)
    write_config(tmp_path / "config", "beam", "ELEM_SOLID_3D_DISP", m.mu, m.Kinv, [("clamp", "fixed")],
                 [("tip", 0.0, 400.0, 1)], 0.1, 1.0, 1e-10, max_iters=30, max_steps=6, solvertype="arclength",
                 timefunction="   1   0.0   1000.0  1.0  0.0  0.0  0.0  0.0  0.0  0.0  0.0",
                 petsc=["-ksp_type preonly", "-pc_type lu"])
    D = mpf.Deck(tmp_path / "config")
    P = D.setup()
    rep = D.run()
    assert rep.steps_done == len(lam_o) and abs(rep.final_lam - lam_o[-1]) < 1e-9 * lam_o[-1]
    u, _ = P.state()
    assert normwise_rel(u, us_o[-1]) < 1e-9
    D.close()
