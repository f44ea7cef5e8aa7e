"""Batched independent meshes: every mesh of a batched Newton solve against the oracle's Newton on
that mesh alone, graph replay against direct launches, and the autotuner's choice being valid."""
from __future__ import annotations

import numpy as np
import pytest

from tests.cases import normwise_rel

pytestmark = pytest.mark.gpu

EDGES = [2, 3, 4, 3]
E = np.array([1.0e3, 3.0e2, 5.0e3, 1.2e3])
NU = np.array([0.30, 0.42, 0.25, 0.45])
V0 = -0.03 * E
V1 = 0.01 * E


def _batch(mpf):
    B, P = mpf.batch_problem(EDGES, E, NU, V0, V1)
    P.prepare_matrix_pattern()
    return B, P


def _oracle_mesh(orc, mpf, B, m):
    lo, hi = B.node_off[m], B.node_off[m + 1]
    e0, e1 = B.elem_off[m], B.elem_off[m + 1]
    mu = E[m] / (2.0 * (1.0 + NU[m]))
    Kinv = 3.0 * (1.0 - 2.0 * NU[m]) / E[m]
    O = orc.Problem(8, B.X[lo:hi], B.conn[e0:e1] - lo, mu, Kinv, 3)
    O.set_tangent(True)
    O.add_dirichlet(B.bottom[m] - lo, -1)
    O.add_traction(B.top[m] - lo, V0[m], V1[m])
    O.prepare_matrix_pattern()
    return O


def test_batched_newton_matches_each_mesh_alone(mpf, orc, gpu):
    B, P = _batch(mpf)
    iters, rel, rep = P.batch_newton_solve(nsteps=2, dt=0.5, itmax=25, rtol=1e-11, cg_maxit=20000, cg_rtol=1e-13)
    assert rep.steps_done == 2 and np.all(rel < 1e-11), (rel, iters)
    u, _ = P.state()
    for m in range(len(EDGES)):
        O = _oracle_mesh(orc, mpf, B, m)
        st, oit, orel = O.newton(2, 0.5, solver=0, rtol=1e-11, itmax=25)
        assert st == 0
        uo, _ = O.state()
        lo, hi = B.node_off[m], B.node_off[m + 1]
        assert normwise_rel(u[3 * lo:3 * hi], uo) < 1e-9, m


def test_batched_band_newton_matches_each_mesh_alone(mpf, orc, gpu):
    """Batched Newton with exact corrections (banded LDL^T per mesh, one workgroup each): every
    mesh against the oracle's Newton with its direct solve; graph replay equal to direct launches."""
    B, P = _batch(mpf)
    iters, rel, rep = P.batch_newton_solve(nsteps=2, dt=0.5, itmax=25, rtol=1e-11, solver=mpf.BATCH_SOLVER_BAND)
    assert rep.steps_done == 2 and np.all(rel < 1e-11), (rel, iters)
    assert rep.cg_iters_total == 0
    u, _ = P.state()
    for m in range(len(EDGES)):
        O = _oracle_mesh(orc, mpf, B, m)
        st, oit, orel = O.newton(2, 0.5, solver=0, rtol=1e-11, itmax=25)
        assert st == 0
        uo, _ = O.state()
        lo, hi = B.node_off[m], B.node_off[m + 1]
        assert normwise_rel(u[3 * lo:3 * hi], uo) < 1e-9, m
    _, Pe = _batch(mpf)
    ie, re, _ = Pe.batch_newton_solve(nsteps=2, dt=0.5, itmax=25, rtol=1e-11, solver=mpf.BATCH_SOLVER_BAND,
                                      use_graph=False)
    ue, _ = Pe.state()
    assert np.array_equal(iters, ie) and np.array_equal(u, ue)


def test_tangent_increment_solves_the_tangent_system(mpf, gpu):
    """d = K^-1 R(u_1; lam_2) with the factors each mesh's last direct solve left, per mesh against
    the tangent assembled at the converged u_1 (the factored tangent is at the iterate before, one
    Newton correction of relative size ~1e-6 away, hence the 1e-4 bound; factors that were lost
    give O(1)); zero on prescribed dofs; bitwise repeatable; an error before any direct solve."""
    import torch

    load = np.array([1.0, 0.02, 1.0, 2.0])  # meshes that converge after different iteration counts
    B, P = mpf.batch_problem(EDGES, E, NU, V0 * load, V1 * load)
    P.prepare_matrix_pattern()
    d = torch.zeros(3 * P.nNode, dtype=torch.float64, device="cuda")
    with pytest.raises(mpf.MpfError):
        P.batch_tangent_increment(1.0, d)
    iters, rel, rep = P.batch_newton_step(0.5, itmax=25, rtol=1e-11, solver=mpf.BATCH_SOLVER_BAND)
    assert np.all(rel < 1e-11)
    print("assemblies per mesh", iters)
    assert iters.min() < iters.max()  # some mesh's factors were left from an earlier iteration
    P.batch_tangent_increment(1.0, d)
    vals, rhs = P.system()  # K at the converged u_1 (the step's last assembly), R(u_1; 1.0)
    pat = P.pattern()
    nda = pat.NodeDofArray.ravel()
    dh = d.cpu().numpy()
    free = nda >= 0
    assert np.all(dh[~free] == 0.0)
    x = np.zeros(len(rhs))
    x[nda[free]] = dh[free]
    rows = np.repeat(np.arange(len(rhs)), np.diff(pat.rowptr))
    Kx = np.bincount(rows, weights=vals * x[pat.colidx], minlength=len(rhs))
    for m in range(len(EDGES)):
        eq = nda[3 * B.node_off[m]:3 * B.node_off[m + 1]]
        eq = eq[eq >= 0]
        r = rhs[eq]
        assert np.linalg.norm(Kx[eq] - r) / np.linalg.norm(r) < 1e-4, m
    d2 = torch.zeros_like(d)
    P.batch_tangent_increment(1.0, d2)
    assert torch.equal(d, d2)


def test_batched_band_cut_blocks_match_whole_blocks(mpf, gpu):
    """Cantilevers of 32 .. 96 elements (half-bandwidth 167): every beam cut in two at a separator
    (the default) against every beam kept whole -- a batched load step with direct Newton
    corrections and the tangent increment from the factors left in the band."""
    import torch
    from mphyfem_amd import learned

    beams = learned.cantilever_workload(40)
    idx = [int(i) for i in np.argsort(beams.lengths)[[0, 20, 39]]]  # shortest, middle, longest
    out = {}
    for split in (False, True):
        _, P = learned.cantilever_problem(beams, idx)
        P.set_band_split(split)
        iters, rel, rep = P.batch_newton_step(1.0 / learned.NSTEPS, itmax=25, rtol=1e-9,
                                              solver=mpf.BATCH_SOLVER_BAND)  # the first step of the ramp
        print("split", split, "iterations", iters, "relres", rel, flush=True)
        assert np.all(rel < 1e-9), rel
        u, _ = P.state()
        d = torch.zeros(3 * P.nNode, dtype=torch.float64, device="cuda")
        P.batch_tangent_increment(1.0, d)
        out[split] = (iters, u, d.cpu().numpy(), P.band_segments())
    assert out[True][3] == (6, 3) and out[False][3] == (3, 0)
    assert normwise_rel(out[True][1], out[False][1]) < 1e-9
    assert normwise_rel(out[True][2], out[False][2]) < 1e-8


def test_tangent_increment_of_a_mesh_never_factored_is_zero(mpf, gpu):
    """A mesh with no load converges at its start and is never factored: its tangent increment is
    zero (not a solve with an empty band), the others' are finite."""
    import torch

    load = np.array([1.0, 0.0, 1.0, 1.0])
    B, P = mpf.batch_problem(EDGES, E, NU, V0 * load, V1 * load)
    P.prepare_matrix_pattern()
    iters, rel, rep = P.batch_newton_step(0.5, itmax=25, rtol=1e-11, solver=mpf.BATCH_SOLVER_BAND)
    assert iters[1] == 1 and np.all(iters[[0, 2, 3]] > 1), iters
    d = torch.zeros(3 * P.nNode, dtype=torch.float64, device="cuda")
    P.batch_tangent_increment(1.0, d)
    dh = d.cpu().numpy()
    assert np.all(np.isfinite(dh))
    assert np.all(dh[3 * B.node_off[1]:3 * B.node_off[2]] == 0.0)
    assert np.abs(dh[3 * B.node_off[0]:3 * B.node_off[1]]).max() > 0.0


def test_batched_graph_replay_equals_direct_launches(mpf, gpu):
    """Two load steps: the graphs, captured once and replayed at both load factors (read from device
    memory), against the same kernels launched directly; one graph launch per Newton iteration."""
    _, Pg = _batch(mpf)
    _, Pe = _batch(mpf)
    ig, rg, rep = Pg.batch_newton_solve(nsteps=2, dt=0.5, rtol=1e-10, use_graph=True, chunk=8)
    ie, re, repe = Pe.batch_newton_solve(nsteps=2, dt=0.5, rtol=1e-10, use_graph=False, chunk=8)
    ug, _ = Pg.state()
    ue, _ = Pe.state()
    assert np.array_equal(ig, ie)
    assert np.array_equal(ug, ue)  # same kernels, same order: bitwise
    print("newton iterations", rep.newton_iters, "iteration graphs", rep.iteration_graphs, "chunk graphs",
          rep.chunk_graphs, "capture s", rep.seconds_capture, flush=True)
    # one per Newton iteration plus each step's closing test
    assert rep.iteration_graphs == rep.newton_iters + rep.steps_done
    assert rep.graph_launches == rep.iteration_graphs + rep.chunk_graphs
    assert (repe.iteration_graphs, repe.chunk_graphs) == (rep.iteration_graphs, rep.chunk_graphs)


def test_batch_autotune_choice_is_used(mpf, gpu):
    _, P = _batch(mpf)
    T, L, ms = P.batch_autotune()
    assert T in (512, 1024, 2048, 4096) and L in (4, 8, 16) and ms > 0.0
    mode, cap, ams = P.batch_assembly_choice()
    print("autotune: tile", T, "lanes", L, "assembly mode", mode, "cap", cap, "ms", ams, flush=True)
    assert mode in (mpf.ASSEMBLY_GATHER, mpf.ASSEMBLY_GATHER_DIRECT, mpf.ASSEMBLY_COLORED_SCATTER) and ams > 0.0
    assert P.assembly_mode() == mode
    _, rel, rep = P.batch_newton_solve(nsteps=1, dt=1.0, rtol=1e-9)
    assert (rep.tile, rep.lanes_per_row) == (T, L)
    assert np.all(rel < 1e-9)


def test_gather_stage_caps_give_the_same_system(mpf, gpu):
    """The element-batch candidates of the autotuner that share the row-owner gather (chunk kernel
    staging 8..32 distinct elements, or direct record reads) assemble bitwise the same K and R."""
    _, P = _batch(mpf)
    P.set_boundary_conditions(1.0)
    ref = None
    for mode, cap in ((mpf.ASSEMBLY_GATHER, 32), (mpf.ASSEMBLY_GATHER, 8), (mpf.ASSEMBLY_GATHER, 16),
                      (mpf.ASSEMBLY_GATHER_DIRECT, 32)):
        P.set_assembly_mode(mode)
        P.set_chunk_stage_cap(cap)
        P.calc_stiffness_and_residual(True, 1.0)
        vals, rhs = P.system()
        if ref is None:
            ref = (vals.copy(), rhs.copy())
        assert np.array_equal(vals, ref[0]) and np.array_equal(rhs, ref[1]), (mode, cap)


def test_batched_q1p0_matches_each_mesh_alone(mpf, gpu):
    """Hex8 Q1/P0 meshes in one batch (Jacobi-PCG and banded direct solves) against each mesh solved
    alone by the device Newton with the same element, loads and tolerances: displacements and element
    pressures within 1e-9; pressures of meshes that converged early are left alone."""
    E_ = np.array([1.0e3, 3.0e2, 5.0e3])
    NU_ = np.array([0.45, 0.49, 0.3])
    edges = [2, 3, 2]
    V0_, V1_ = -0.02 * E_, 0.005 * E_
    for solver in (mpf.BATCH_SOLVER_PCG, mpf.BATCH_SOLVER_BAND):
        B, P = mpf.batch_problem(edges, E_, NU_, V0_, V1_, elem_type=mpf.HEX8P0)
        P.prepare_matrix_pattern()
        iters, rel, rep = P.batch_newton_solve(nsteps=2, dt=0.5, itmax=25, rtol=1e-11, cg_maxit=20000,
                                               cg_rtol=1e-13, solver=solver)
        assert rep.steps_done == 2 and np.all(rel < 1e-11), (rel, iters)
        u, pr = P.state()
        for m in range(len(edges)):
            lo, hi = B.node_off[m], B.node_off[m + 1]
            e0, e1 = B.elem_off[m], B.elem_off[m + 1]
            S = mpf.Problem(mpf.HEX8P0, B.X[lo:hi], B.conn[e0:e1] - lo,
                            mpf.matl_from_E_nu(float(E_[m]), float(NU_[m]), 3, True, mpf.TANGENT_CONSISTENT))
            S.add_dirichlet(B.bottom[m] - lo, -1)
            S.add_traction(B.top[m] - lo, float(V0_[m]), float(V1_[m]))
            S.prepare_matrix_pattern()
            _, srep = S.solve_with_newton_raphson(S.newton_opts(nsteps=2, dt=0.5, itmax=25, rtol=1e-11,
                                                                ksp=mpf.KSP_BAND))
            assert srep.steps_done == 2
            us, ps = S.state()
            print("Q1/P0 batch", solver, "mesh", m, normwise_rel(u[3 * lo:3 * hi], us), normwise_rel(pr[e0:e1], ps),
                  flush=True)
            assert normwise_rel(u[3 * lo:3 * hi], us) < 1e-9, m
            assert normwise_rel(pr[e0:e1], ps) < 1e-9, m
            S.close()
        P.close()


def test_batch_rejects_shared_nodes(mpf, gpu):
    B = mpf.mesh_batch_hex([2, 2])
    P = mpf.Problem(mpf.HEX8, B.X, B.conn, mpf.matl_from_E_nu(1e3, 0.3))
    bad_off = np.array([0, 20, B.node_off[-1]], dtype=np.int32)  # splits the first mesh's nodes
    with pytest.raises(mpf.MpfError):
        P.set_batch(bad_off, B.elem_off, [1.0, 1.0], [1e-3, 1e-3])


def test_set_batch_twice_is_refused(mpf, gpu):
    """The batch layout, tiles, per-mesh flags and captured graphs are built for one batch: a second
    set_batch on the same problem is refused before anything is reallocated, and the problem still
    solves the first batch."""
    B, P = _batch(mpf)
    mu = E / (2.0 * (1.0 + NU))
    Kinv = 3.0 * (1.0 - 2.0 * NU) / E
    with pytest.raises(mpf.MpfError):
        P.set_batch(B.node_off, B.elem_off, mu, Kinv)
    _, rel, rep = P.batch_newton_solve(nsteps=1, dt=1.0, rtol=1e-9)
    assert rep.steps_done == 1 and np.all(rel < 1e-9)


def test_graph_recaptured_when_tolerances_change(mpf, gpu):
    """Two load steps at the same load factor with different Newton and CG tolerances, each from the
    undeformed state (so both start from the same residual): the graph replay (tolerances are
    kernel arguments frozen at capture) matches direct launches bitwise."""
    runs = []
    for use_graph in (True, False):
        _, P = _batch(mpf)
        zero = np.zeros(3 * P.nNode)
        out = []
        for rtol, cg_rtol, cg_maxit in ((1e-6, 1e-8, 3000), (1e-11, 1e-13, 20000)):
            P.set_state(disp=zero, disp_prev=zero)
            it, rel, _ = P.batch_newton_step(0.5, itmax=25, rtol=rtol, cg_rtol=cg_rtol, cg_maxit=cg_maxit,
                                             use_graph=use_graph, chunk=8)
            out.append((it.copy(), rel.copy(), P.state()[0]))
        runs.append(out)
    for (ig, rg, ug), (ie, re_, ue) in zip(*runs):
        assert np.array_equal(ig, ie) and np.array_equal(rg, re_) and np.array_equal(ug, ue)
    assert np.all(runs[0][1][1] < 1e-11)
