"""GPU parity: the HIP path, called through the C ABI, against the oracle on the same seeded inputs.

Bars (set by the workload definition): CSR row pointers, column indices, DOF numbering and the
colouring bit-exact; K_e, R_e and the global K / R within 1e-12 normwise relative
(max |a - b| / max |b|) in FP64; converged displacements within 1e-9 relative.
"""
from __future__ import annotations

from pathlib import Path

import numpy as np
import pytest

from tests.cases import (apply_bcs, cube_hex, hex_stretch_bcs, normwise_rel, q1p0_newton, random_state, tet10_fan,
                         tet_compress_bcs)

pytestmark = pytest.mark.gpu
GOLD = Path(__file__).resolve().parent / "golden"
TOL_K = 1e-12


def _t(gpu, a, dtype=None):
    import torch
    return torch.as_tensor(np.ascontiguousarray(a), device=gpu, dtype=dtype)


def test_hex8_element_golden(mpf, orc, gpu):
    import torch
    g = np.load(GOLD / "hex8_element.npz")
    X, u = g["X"], g["u"]
    conn = np.arange(8, dtype=np.int32)[None]
    for tangent, Kexp in [(mpf.TANGENT_REFERENCE, g["Ke_consistent"] - g["Ke_volgeo"]),
                          (mpf.TANGENT_CONSISTENT, g["Ke_consistent"])]:
        m = mpf.matl(float(g["mu"]), float(g["Kinv"]), 3, tangent=tangent)
        Ke, Re, bad = mpf.elem_hex8_nh_batch(_t(gpu, X), _t(gpu, u.ravel()), _t(gpu, conn, torch.int32), m)
        assert bad == 0
        assert normwise_rel(Re.cpu().numpy()[0], g["Re"]) < 1e-12
        assert normwise_rel(Ke.cpu().numpy()[0], Kexp) < 1e-7  # FD fixture accuracy
        Ko, Ro, _ = (orc.elem_disp if tangent == 0 else orc.elem_disp_consistent)(X, u, m.mu, m.Kinv, 3)
        assert normwise_rel(Ke.cpu().numpy()[0], Ko) < TOL_K
        assert normwise_rel(Re.cpu().numpy()[0], Ro) < TOL_K


@pytest.mark.parametrize("tangent", [0, 1])
def test_hex8_batch_vs_oracle(mpf, orc, gpu, tangent):
    import torch
    n = 10
    X, conn = cube_hex(n)
    u = random_state(len(X), 0.01, seed=11)
    m = mpf.matl_from_E_nu(1e3, 0.3, tangent=tangent)
    Ke, Re, bad = mpf.elem_hex8_nh_batch(_t(gpu, X), _t(gpu, u), _t(gpu, conn, torch.int32), m)
    assert bad == 0
    Ke, Re = Ke.cpu().numpy(), Re.cpu().numpy()
    f = orc.elem_disp if tangent == 0 else orc.elem_disp_consistent
    worst_k = worst_r = 0.0
    for e in range(0, len(conn), 7):
        Ko, Ro, st = f(X[conn[e]], u.reshape(-1, 3)[conn[e]], m.mu, m.Kinv, 3)
        worst_k = max(worst_k, normwise_rel(Ke[e], Ko))
        worst_r = max(worst_r, normwise_rel(Re[e], Ro))
    assert worst_k < TOL_K and worst_r < TOL_K, (worst_k, worst_r)
    # first / count sub-range
    Ke2, Re2, _ = mpf.elem_hex8_nh_batch(_t(gpu, X), _t(gpu, u), _t(gpu, conn, torch.int32), m, first=13, count=5)
    assert np.array_equal(Ke2.cpu().numpy(), Ke[13:18])


def test_hex8p0_batch_vs_oracle(mpf, orc, gpu):
    """Hex8 Q1/P0, condensed, every element of a jittered block against the oracle (which is pinned
    by finite differences, closed forms and the condensation identity): Kc, Fc, Kup, Kpp, Fp within
    1e-12."""
    import torch
    n = 3
    X, conn = mpf.mesh_hex(n, n, n, jitter=0.1, seed=5)
    u = random_state(len(X), 0.02, seed=7)
    up = 0.5 * u
    pe = np.random.default_rng(8).uniform(-0.5, 0.5, len(conn))
    mu, Kinv = 50.0, 1e-2
    out = mpf.elem_hex8p0_nh_batch(_t(gpu, X), _t(gpu, u), _t(gpu, up), _t(gpu, pe), _t(gpu, conn, torch.int32),
                                   mpf.matl(mu, Kinv, 3))
    Kc, Fc, Kup, Kpp, Fp, bad = [o.cpu().numpy() if hasattr(o, "cpu") else o for o in out]
    assert bad == 0
    ref = [orc.elem_hex8p0(X[c], u.reshape(-1, 3)[c], up.reshape(-1, 3)[c], pe[e], mu, Kinv, 3)
           for e, c in enumerate(conn)]
    assert all(o[8] == 0 for o in ref)
    for got, k in ((Kc, 0), (Fc, 1), (Kup, 4), (Kpp, 6), (Fp, 7)):
        want = np.array([o[k] for o in ref])
        assert normwise_rel(got, want) < 1e-12, (k, normwise_rel(got, want))


def test_hex8p0_newton_traction_vs_oracle(mpf, orc, gpu):
    """Hex8 Q1/P0 under a normal dead-load traction on the top face as well as the end stretch: the
    external force enters the condensed system like the displacement element's (lambda * Fext added
    to the right-hand side); converged displacements and element pressures within 1e-9."""
    from tests.cases import oracle_fext
    n, mu, Kinv = 3, 1.0, 1e-3
    X, conn = cube_hex(n, jitter=0.1, seed=4)
    bcs = hex_stretch_bcs(n)
    faces, v0 = mpf.hex_face_quads(n, n, n, 5), -0.05
    P = mpf.Problem(mpf.HEX8P0, X, conn, mpf.matl(mu, Kinv, 3))
    apply_bcs(P, bcs)
    P.add_traction(faces, v0)
    P.prepare_matrix_pattern()
    opts = P.newton_opts(nsteps=2, dt=0.5, itmax=20, rtol=1e-10, ksp=mpf.KSP_BAND)
    iters, rep = P.solve_with_newton_raphson(opts)
    assert rep.steps_done == 2, iters
    ud, pd = P.state()
    fext = oracle_fext(orc, X, conn, bcs, faces, v0)
    assert np.abs(fext).max() > 0
    uo, po, _ = q1p0_newton(orc, X, conn, bcs, mu, Kinv, nsteps=2, dt=0.5, rtol=1e-10, fext=fext)
    assert normwise_rel(ud, uo) < 1e-9, normwise_rel(ud, uo)
    assert normwise_rel(pd, po) < 1e-9, normwise_rel(pd, po)


def test_hex8p0_relieves_locking_on_gpu(mpf, orc, gpu):
    """Clamped 4^3 block under a small top load at nu = 0.4999: the GPU displacement Hex8 locks (over
    10x less deflection) where the GPU Q1/P0 does not, and the Q1/P0 displacements match the CPU
    Q1/P0 driver within 1e-9."""
    from tests.cases import oracle_fext
    n, nu = 4, 0.4999
    X, conn = mpf.mesh_hex(n, n, n, jitter=0.0, seed=5)
    bot, top = mpf.face_nodes(mpf.HEX8, n, n, n, 4), mpf.face_nodes(mpf.HEX8, n, n, n, 5)
    faces = mpf.hex_face_quads(n, n, n, 5)
    uz = {}
    for name, et in (("hex8", mpf.HEX8), ("q1p0", mpf.HEX8P0)):
        P = mpf.Problem(et, X, conn, mpf.matl_from_E_nu(1.0, nu, tangent=mpf.TANGENT_CONSISTENT))
        P.add_dirichlet(bot, -1)
        P.add_traction(faces, -1e-3)
        P.prepare_matrix_pattern()
        _, rep = P.solve_with_newton_raphson(P.newton_opts(nsteps=1, dt=1.0, itmax=30, rtol=1e-8,
                                                           ksp=mpf.KSP_BAND))
        assert rep.steps_done == 1
        uz[name] = P.state()[0]
        P.close()
    assert uz["q1p0"][3 * top + 2].mean() / uz["hex8"][3 * top + 2].mean() > 10.0
    m = mpf.matl_from_E_nu(1.0, nu)
    bcs = [(bot, -1, None)]
    uo, _, _ = q1p0_newton(orc, X, conn, bcs, m.mu, m.Kinv, 1, 1.0, rtol=1e-8,
                           fext=oracle_fext(orc, X, conn, bcs, faces, -1e-3))
    assert normwise_rel(uz["q1p0"], uo) < 1e-9, normwise_rel(uz["q1p0"], uo)


@pytest.mark.parametrize("mode", [0, 1], ids=["gather", "colored_scatter"])
@pytest.mark.parametrize("ksp", ["band", "cg"])
def test_hex8p0_newton_vs_oracle(mpf, orc, gpu, ksp, mode):
    """Hex8 Q1/P0 problem on the GPU (condensed element matrices through the row-owner gather or
    scattered colour by colour, element pressures recovered after each correction) against the CPU
    Q1/P0 driver on the oracle's element:
    a near-incompressible jittered block stretched 10% in two load steps, converged displacements
    and element pressures within 1e-9."""
    n, mu, Kinv = 3, 1.0, 1e-3
    X, conn = cube_hex(n, jitter=0.1, seed=3)
    bcs = hex_stretch_bcs(n)
    m = mpf.matl(mu, Kinv, 3)
    P = mpf.Problem(mpf.HEX8P0, X, conn, m)
    apply_bcs(P, bcs)
    P.prepare_matrix_pattern()
    assert P.assembly_mode() == mpf.ASSEMBLY_GATHER  # condensed records through the row-owner gather
    P.set_assembly_mode(mode)
    if ksp == "band":
        opts = P.newton_opts(nsteps=2, dt=0.5, itmax=20, rtol=1e-10, ksp=mpf.KSP_BAND)
    else:
        opts = P.newton_opts(nsteps=2, dt=0.5, itmax=20, rtol=1e-10, ksp=mpf.KSP_CG, pc=mpf.PC_BJACOBI3,
                             lin_rtol=1e-13, maxit=50000)
    iters, rep = P.solve_with_newton_raphson(opts)
    assert rep.steps_done == 2, iters
    ud, pd = P.state()
    assert pd.shape == (len(conn),)
    uo, po, oit = q1p0_newton(orc, X, conn, bcs, mu, Kinv, nsteps=2, dt=0.5, rtol=1e-10)
    assert normwise_rel(ud, uo) < 1e-9, normwise_rel(ud, uo)
    assert normwise_rel(pd, po) < 1e-9, normwise_rel(pd, po)
    assert np.abs(po).max() > 1e-3


def test_tet10p1_golden_and_batch(mpf, orc, gpu):
    import torch
    g = np.load(GOLD / "tet10p1_element.npz")
    conn1 = np.arange(10, dtype=np.int32)[None]
    m = mpf.matl(float(g["mu"]), float(g["Kinv"]), 3, mixed=True)
    pres = np.concatenate([g["pres"], np.zeros(6)])
    out = mpf.elem_tet10p1_nh_batch(_t(gpu, g["X"]), _t(gpu, g["u"].ravel()), _t(gpu, g["uprev"].ravel()),
                                    _t(gpu, pres), _t(gpu, conn1, torch.int32), m)
    Kuu, Kup, Kpu, Kpp, Fu, Fp, bad = [o.cpu().numpy() if hasattr(o, "cpu") else o for o in out]
    assert bad == 0
    R = np.concatenate([Fu[0], Fp[0]])
    K = np.block([[Kuu[0], Kup[0]], [Kpu[0], Kpp[0]]])
    assert normwise_rel(R, g["R"]) < 1e-12
    assert normwise_rel(K, g["K"]) < 1e-6
    # batch vs oracle
    n = 3
    X, conn = mpf.mesh_tet10(n, jitter=0.1, seed=2)
    u = random_state(len(X), 0.01, seed=5)
    up = 0.6 * u
    p = np.random.default_rng(6).uniform(-1, 1, len(X))
    out = mpf.elem_tet10p1_nh_batch(_t(gpu, X), _t(gpu, u), _t(gpu, up), _t(gpu, p), _t(gpu, conn, torch.int32),
                                    mpf.matl(1.0, 1e-3, 3, mixed=True))
    Kuu, Kup, Kpu, Kpp, Fu, Fp, bad = [o.cpu().numpy() if hasattr(o, "cpu") else o for o in out]
    assert bad == 0
    worst = 0.0
    for e in range(0, len(conn), 5):
        c = conn[e]
        o = orc.elem_tet10p1(X[c], u.reshape(-1, 3)[c], up.reshape(-1, 3)[c], p[c[:4]], 1.0, 1e-3, 3)
        Ko = np.block([[o[0], o[1]], [o[2], o[3]]])
        Kg = np.block([[Kuu[e], Kup[e]], [Kpu[e], Kpp[e]]])
        worst = max(worst, normwise_rel(Kg, Ko), normwise_rel(np.concatenate([Fu[e], Fp[e]]),
                                                              np.concatenate([o[4], o[5]])))
    assert worst < TOL_K, worst


def test_negative_jacobian_flag(mpf, gpu):
    import torch
    X, conn = cube_hex(3, jitter=0.0)
    u = np.zeros((len(X), 3))
    e_bad = 13
    # collapse element 13 through itself: move its top nodes below its bottom nodes
    for a in range(4, 8):
        u[conn[e_bad][a], 2] = -1.5 / 3
    Ke, Re, bad = mpf.elem_hex8_nh_batch(_t(gpu, X), _t(gpu, u.ravel()), _t(gpu, conn, torch.int32),
                                         mpf.matl(10.0, 1e-2))
    assert bad >= 1 and bad - 1 <= e_bad
    P = mpf.Problem(mpf.HEX8, X, conn, mpf.matl(10.0, 1e-2))
    P.prepare_matrix_pattern()
    P.set_state(disp=u.ravel())
    with pytest.raises(mpf.NegativeJacobian):
        P.calc_stiffness_and_residual(False, 1.0)


def test_material_batch_vs_oracle(mpf, orc, gpu):
    rng = np.random.default_rng(4)
    Fs = np.eye(3)[None] + rng.uniform(-0.1, 0.1, (32, 3, 3))
    ps = rng.uniform(-3, 3, 32)
    for mixed, tangent, Ut in [(False, 0, 3), (False, 1, 3), (True, 0, 3), (False, 0, 1), (False, 0, 4), (False, 0, 7)]:
        m = mpf.matl(12.0, 1e-2, Ut, mixed=mixed, tangent=tangent)
        sig, Cm = mpf.neohooke_stress_tangent(_t(gpu, Fs), _t(gpu, ps), m)
        sig, Cm = sig.cpu().numpy(), Cm.cpu().numpy()
        for i in range(0, 32, 3):
            so, Co = orc.neohooke(Fs[i], 12.0, 1e-2, Ut, mixed=mixed, pres=ps[i])
            if tangent == 1:
                J = np.linalg.det(Fs[i])
                U1 = orc.lib().orc_dUdJ(Ut, 100.0, J)
                mp = [(0, 0), (1, 0), (2, 0), (0, 1), (1, 1), (2, 1), (0, 2), (1, 2), (2, 2)]
                Co = Co + np.array([[U1 * (j == l) * (i_ == k) for (k, l) in mp] for (i_, j) in mp])
            assert normwise_rel(sig[i], so) < TOL_K
            assert normwise_rel(Cm[i], Co) < TOL_K


def _hex_problem(mpf, n, m, traction=None):
    X, conn = cube_hex(n)
    P = mpf.Problem(mpf.HEX8, X, conn, m)
    apply_bcs(P, hex_stretch_bcs(n))
    if traction is not None:
        P.add_traction(mpf.hex_face_quads(n, n, n, 5), *traction)
    P.prepare_matrix_pattern()
    return X, conn, P


def _orc_hex_problem(orc, mpf, X, conn, n, m, traction=None):
    O = orc.Problem(8, X, conn, m.mu, m.Kinv, m.Utype)
    O.set_tangent(m.tangent == 1)
    apply_bcs(O, hex_stretch_bcs(n))
    if traction is not None:
        O.add_traction(mpf.hex_face_quads(n, n, n, 5), *traction)
    O.prepare_matrix_pattern()
    return O


def test_pattern_and_colouring_bitexact_hex(mpf, orc, gpu):
    n = 10
    m = mpf.matl_from_E_nu(1e3, 0.3)
    X, conn, P = _hex_problem(mpf, n, m)
    O = _orc_hex_problem(orc, mpf, X, conn, n, m)
    pat = P.pattern()
    rowptr, colidx, nda, idp = O.pattern()
    assert (pat.dispDOF, pat.presDOF, len(pat.colidx)) == (3509, 0, 231601)
    assert np.array_equal(pat.rowptr, rowptr)
    assert np.array_equal(pat.colidx, colidx)
    assert np.array_equal(pat.NodeDofArray, nda)
    col, ncol = O.color()
    assert pat.ncolors == ncol
    assert np.array_equal(pat.color, col)


def test_pattern_and_colouring_bitexact_tet(mpf, orc, gpu):
    n = 4
    X, conn = mpf.mesh_tet10(n, jitter=0.1, seed=2)
    m = mpf.matl(1.0, 1e-3, 3, mixed=True)
    P = mpf.Problem(mpf.TET10P1, X, conn, m)
    O = orc.Problem(10, X, conn, 1.0, 1e-3, 3)
    for Q in (P, O):
        apply_bcs(Q, tet_compress_bcs(n))
        Q.prepare_matrix_pattern()
    pat = P.pattern()
    rowptr, colidx, nda, idp = O.pattern()
    assert np.array_equal(pat.rowptr, rowptr) and np.array_equal(pat.colidx, colidx)
    assert np.array_equal(pat.NodeDofArray, nda) and np.array_equal(pat.ID_pres, idp)
    col, ncol = O.color()
    assert pat.ncolors == ncol and np.array_equal(pat.color, col)


@pytest.mark.parametrize("mode", [0, 1, 2], ids=["gather", "colored_scatter", "gather_direct"])
@pytest.mark.parametrize("tangent,traction", [(0, None), (1, (-20.0, 5.0))])
def test_global_assembly_hex_vs_oracle(mpf, orc, gpu, tangent, traction, mode):
    n = 10
    m = mpf.matl_from_E_nu(1e3, 0.3, tangent=tangent)
    X, conn, P = _hex_problem(mpf, n, m, traction)
    P.set_assembly_mode(mode)
    O = _orc_hex_problem(orc, mpf, X, conn, n, m, traction)
    u = random_state(len(X), 0.005, seed=21)
    for Q in (P, O):
        Q.set_state(disp=u, disp_prev=u)
        Q.set_boundary_conditions(0.7)
    for first in (True, False):
        P.calc_stiffness_and_residual(first, 0.7)
        vals, rhs = P.system()
        ov, orhs, st = O.assemble(first, 0.7)
        assert st == 0
        assert normwise_rel(vals, ov) < TOL_K
        assert normwise_rel(rhs, orhs) < TOL_K
        assert abs(P.rhs_norm() - np.linalg.norm(orhs)) < 1e-12 * np.linalg.norm(orhs)


@pytest.mark.parametrize("mode", [0, 1, 2], ids=["gather", "colored_scatter", "gather_direct"])
def test_global_assembly_tet_vs_oracle(mpf, orc, gpu, mode):
    n = 4
    X, conn = mpf.mesh_tet10(n, jitter=0.1, seed=2)
    m = mpf.matl(1.0, 1e-3, 3, mixed=True)
    P = mpf.Problem(mpf.TET10P1, X, conn, m)
    P.set_assembly_mode(mode)
    O = orc.Problem(10, X, conn, 1.0, 1e-3, 3)
    u = random_state(len(X), 0.004, seed=31)
    up = 0.5 * u
    p = np.random.default_rng(32).uniform(-0.5, 0.5, len(X))
    for Q in (P, O):
        apply_bcs(Q, tet_compress_bcs(n))
        Q.prepare_matrix_pattern()
        Q.set_state(disp=u, disp_prev=up, pres=p, pres_prev=p)
        Q.set_boundary_conditions(1.0)
    for first in (True, False):
        P.calc_stiffness_and_residual(first, 1.0)
        vals, rhs = P.system()
        ov, orhs, st = O.assemble(first, 1.0)
        assert normwise_rel(vals, ov) < TOL_K
        assert normwise_rel(rhs, orhs) < TOL_K
    assert P.assembly_mode() == mode


@pytest.mark.parametrize("mode,cap", [(0, 32), (0, 12), (2, 32), (1, 32)],
                         ids=["gather", "gather_partly_staged", "gather_direct", "colored_scatter"])
def test_global_assembly_hex8p0_vs_oracle(mpf, orc, gpu, mode, cap):
    """Hex8 Q1/P0 condensed system (Kc = Kuu - Kup Kpu / Kpp, Fc = Fu - Kup Fp / Kpp) through the
    row-owner gather (chunk kernel staging the condensed records in LDS, some chunks reading them
    directly, or every chunk reading directly) and through the coloured scatter: global K and R
    against the oracle's condensed assembly, with the first-iteration lift and without; the gather
    bitwise repeatable."""
    nx, ny, nz = 7, 5, 6
    X, conn = mpf.mesh_hex(nx, ny, nz, L=(1.0, 0.8, 0.9), jitter=0.1, seed=6)
    mu, Kinv = 1.0, 1e-3
    bcs = [(mpf.face_nodes(mpf.HEX8, nx, ny, nz, 0), -1, None), (mpf.face_nodes(mpf.HEX8, nx, ny, nz, 1), 0, 0.05)]
    P = mpf.Problem(mpf.HEX8P0, X, conn, mpf.matl(mu, Kinv, 3))
    O = orc.Problem(9, X, conn, mu, Kinv, 3)
    u = random_state(len(X), 0.004, seed=41)
    pe = np.random.default_rng(42).uniform(-0.5, 0.5, len(conn))
    for Q in (P, O):
        apply_bcs(Q, bcs)
        Q.prepare_matrix_pattern()
        Q.set_state(disp=u, disp_prev=0.5 * u, pres=pe, pres_prev=pe)
        Q.set_boundary_conditions(0.8)
    P.set_assembly_mode(mode)
    P.set_chunk_stage_cap(cap)
    for first in (True, False):
        P.calc_stiffness_and_residual(first, 0.8)
        vals, rhs = P.system()
        P.calc_stiffness_and_residual(first, 0.8)
        v2, r2 = P.system()
        if mode != 1:
            assert np.array_equal(vals, v2) and np.array_equal(rhs, r2)
        ov, orhs, st = O.assemble(first, 0.8)
        assert st == 0
        print("Q1/P0 mode", mode, "cap", cap, "first", first, normwise_rel(vals, ov), normwise_rel(rhs, orhs),
              flush=True)
        assert normwise_rel(vals, ov) < TOL_K
        assert normwise_rel(rhs, orhs) < TOL_K


@pytest.mark.parametrize("cap", [32, 20, 0], ids=["staged", "mixed", "direct_reads"])
def test_hex_chunk_staging_vs_oracle(mpf, orc, gpu, cap):
    """Hex8 chunked matrix kernel on a jittered box that is not a cube: chunks staged in LDS (cap 32),
    some staged and some reading the records directly (cap 20), every chunk reading directly (0)."""
    nx, ny, nz = 9, 6, 11
    X, conn = mpf.mesh_hex(nx, ny, nz, L=(1.0, 0.7, 1.6), jitter=0.1, seed=4)
    m = mpf.matl_from_E_nu(1e3, 0.3, tangent=1)
    bcs = [(mpf.face_nodes(mpf.HEX8, nx, ny, nz, 0), -1, None), (mpf.face_nodes(mpf.HEX8, nx, ny, nz, 1), 0, 0.1)]
    P = mpf.Problem(mpf.HEX8, X, conn, m)
    O = orc.Problem(8, X, conn, m.mu, m.Kinv, m.Utype)
    O.set_tangent(True)
    u = random_state(len(X), 0.005, seed=23)
    for Q in (P, O):
        apply_bcs(Q, bcs)
        Q.prepare_matrix_pattern()
        Q.set_state(disp=u, disp_prev=u)
        Q.set_boundary_conditions(0.6)
    P.set_chunk_stage_cap(cap)
    for first in (True, False):
        P.calc_stiffness_and_residual(first, 0.6)
        vals, rhs = P.system()
        P.calc_stiffness_and_residual(first, 0.6)
        v2, r2 = P.system()
        assert np.array_equal(vals, v2) and np.array_equal(rhs, r2)
        ov, orhs, st = O.assemble(first, 0.6)
        assert st == 0
        assert normwise_rel(vals, ov) < TOL_K
        assert normwise_rel(rhs, orhs) < TOL_K
    with pytest.raises(mpf.MpfError):
        P.set_chunk_stage_cap(33)


def test_high_valence_node_falls_back_to_scatter(mpf, orc, gpu):
    """A node in 80 elements exceeds the gather tables (64 elements, 128 coupled nodes per node):
    the problem switches itself to the coloured scatter, and K and R still match the oracle."""
    X, conn, outer = tet10_fan(1)
    bcs = [(outer, 0, None), (outer, 1, None), (outer, 2, -0.02)]
    P = mpf.Problem(mpf.TET10P1, X, conn, mpf.matl(1.0, 1e-3, 3, mixed=True))
    O = orc.Problem(10, X, conn, 1.0, 1e-3, 3)
    u = random_state(len(X), 0.003, seed=61)
    p = np.random.default_rng(62).uniform(-0.2, 0.2, len(X))
    for Q in (P, O):
        apply_bcs(Q, bcs)
        Q.prepare_matrix_pattern()
        Q.set_state(disp=u, disp_prev=0.5 * u, pres=p, pres_prev=p)
        Q.set_boundary_conditions(1.0)
    assert P.assembly_mode() == mpf.ASSEMBLY_COLORED_SCATTER
    pat = P.pattern()
    rowptr, colidx, _, _ = O.pattern()
    assert np.array_equal(pat.rowptr, rowptr) and np.array_equal(pat.colidx, colidx)
    for first in (True, False):
        P.calc_stiffness_and_residual(first, 1.0)
        vals, rhs = P.system()
        ov, orhs, st = O.assemble(first, 1.0)
        assert st == 0
        assert normwise_rel(vals, ov) < TOL_K
        assert normwise_rel(rhs, orhs) < TOL_K


@pytest.mark.parametrize("etype", ["hex8", "tet10p1"])
def test_gather_matrix_exactly_symmetric(mpf, gpu, etype):
    """The gather sums each pair of column blocks once and writes it to both places, so the
    assembled matrix (u-u, u-p, p-u and p-p blocks alike) equals its transpose bit for bit."""
    import scipy.sparse as sp

    if etype == "hex8":
        n = 6
        X, conn = cube_hex(n)
        P = mpf.Problem(mpf.HEX8, X, conn, mpf.matl_from_E_nu(1e3, 0.3, tangent=1))
        apply_bcs(P, hex_stretch_bcs(n))
    else:
        n = 4
        X, conn = mpf.mesh_tet10(n, jitter=0.1, seed=2)
        P = mpf.Problem(mpf.TET10P1, X, conn, mpf.matl(1.0, 1e-3, 3, mixed=True))
        apply_bcs(P, tet_compress_bcs(n))
    P.prepare_matrix_pattern()
    u = random_state(len(X), 0.004, seed=51)
    P.set_state(disp=u, disp_prev=0.5 * u)
    P.set_boundary_conditions(1.0)
    P.calc_stiffness_and_residual(True, 1.0)
    vals, _ = P.system()
    pat = P.pattern()
    nd = len(pat.rowptr) - 1
    A = sp.csr_matrix((vals, pat.colidx, pat.rowptr), shape=(nd, nd))
    D = (A - A.T).tocsr()
    D.eliminate_zeros()
    assert D.nnz == 0
    assert abs(A).max() > 0


def test_gather_matches_scatter_bitwise_reproducible(mpf, gpu):
    """The gather path writes each CSR value once in a fixed order: repeated assemblies are
    bit-identical, and both algorithms agree to rounding."""
    n = 6
    X, conn = mpf.mesh_tet10(n, jitter=0.1, seed=2)
    P = mpf.Problem(mpf.TET10P1, X, conn, mpf.matl(1.0, 1e-3, 3, mixed=True))
    apply_bcs(P, tet_compress_bcs(n))
    P.prepare_matrix_pattern()
    u = random_state(len(X), 0.003, seed=41)
    P.set_state(disp=u, disp_prev=0.5 * u)
    P.set_boundary_conditions(1.0)
    P.calc_stiffness_and_residual(True, 1.0)
    v1, r1 = P.system()
    P.calc_stiffness_and_residual(True, 1.0)
    v2, r2 = P.system()
    assert np.array_equal(v1, v2) and np.array_equal(r1, r2)
    P.set_assembly_mode(mpf.ASSEMBLY_COLORED_SCATTER)
    P.calc_stiffness_and_residual(True, 1.0)
    v3, r3 = P.system()
    assert normwise_rel(v3, v1) < TOL_K and normwise_rel(r3, r1) < TOL_K


def test_krylov_vs_oracle(mpf, orc, gpu):
    n = 8
    m = mpf.matl_from_E_nu(1e3, 0.3, tangent=1)
    X, conn, P = _hex_problem(mpf, n, m)
    O = _orc_hex_problem(orc, mpf, X, conn, n, m)
    for Q in (P, O):
        Q.set_boundary_conditions(1.0)
    P.calc_stiffness_and_residual(True, 1.0)
    vals, rhs = P.system()
    x_ref, st = O.direct_solve(vals, rhs)
    assert st == 0
    for ksp, pc in [(mpf.KSP_CG, mpf.PC_BJACOBI3), (mpf.KSP_CG, mpf.PC_JACOBI), (mpf.KSP_GMRES, mpf.PC_BJACOBI3)]:
        r = P.factorise_and_solve(ksp=ksp, pc=pc, rtol=1e-11, maxit=5000, restart=40)
        assert r.converged, (ksp, pc, r.iterations, r.relres)
        du = P.solution()
        assert normwise_rel(du, x_ref) < 1e-8
        assert np.linalg.norm(O.spmv(vals, du) - rhs) / np.linalg.norm(rhs) < 2e-11


def test_band_direct_solve_vs_oracle(mpf, orc, gpu):
    """Banded LDL^T on the GPU (bandwidth-reducing renumbering) against the oracle's banded
    LDL^T on the same assembled system, and bitwise repeatable."""
    n = 8
    m = mpf.matl_from_E_nu(1e3, 0.3, tangent=1)
    X, conn, P = _hex_problem(mpf, n, m)
    O = _orc_hex_problem(orc, mpf, X, conn, n, m)
    for Q in (P, O):
        Q.set_boundary_conditions(1.0)
    P.calc_stiffness_and_residual(True, 1.0)
    vals, rhs = P.system()
    x_ref, st = O.direct_solve(vals, rhs)
    assert st == 0
    r = P.factorise_and_solve(ksp=mpf.KSP_BAND)
    assert r.converged
    du = P.solution()
    assert normwise_rel(du, x_ref) < 1e-10
    assert np.linalg.norm(O.spmv(vals, du) - rhs) / np.linalg.norm(rhs) < 1e-12
    P.factorise_and_solve(ksp=mpf.KSP_BAND)
    assert np.array_equal(P.solution(), du)
    nblk, bmax, ndof, flops = P.band_stats()  # one block: the whole displacement system
    assert nblk == 1 and ndof == len(rhs) and 0 < bmax < ndof
    assert flops == float(ndof) * bmax * (bmax + 1)


@pytest.mark.parametrize("shape", ["cube", "beam", "wide"])
def test_band_cut_block_vs_oracle(mpf, orc, gpu, shape):
    """A block cut in two at a separator (elimination order [first half, second half reversed,
    separator]) against the oracle's banded LDL^T and against the same block kept whole: the
    8^3 cube takes the wide-band panel path (b > 176), a 40 x 3 x 3 beam the one-wavefront panel
    path with the look-ahead, a 12 x 11 x 11 block (b > 448) also the solve's second pass over
    coupling rows.  Different rounding from the whole block, bitwise repeatable."""
    m = mpf.matl_from_E_nu(1e3, 0.3, tangent=1)
    if shape == "cube":
        n = 8
        X, conn, P = _hex_problem(mpf, n, m)
        O = _orc_hex_problem(orc, mpf, X, conn, n, m)
    else:
        dims, size = ((40, 3, 3), (4.0, 0.3, 0.3)) if shape == "beam" else ((12, 11, 11), (1.2, 1.1, 1.1))
        X, conn = mpf.mesh_hex(*dims, L=size, jitter=0.05, seed=3)
        bcs = [(mpf.face_nodes(mpf.HEX8, *dims, 0), -1, None), (mpf.face_nodes(mpf.HEX8, *dims, 1), 0, 0.05)]
        P = mpf.Problem(mpf.HEX8, X, conn, m)
        O = orc.Problem(8, X, conn, m.mu, m.Kinv, m.Utype)
        O.set_tangent(True)
        for Q in (P, O):
            apply_bcs(Q, bcs)
            Q.prepare_matrix_pattern()
    for Q in (P, O):
        Q.set_boundary_conditions(1.0)
    P.calc_stiffness_and_residual(True, 1.0)
    vals, rhs = P.system()
    x_ref, st = O.direct_solve(vals, rhs)
    assert st == 0
    assert P.factorise_and_solve(ksp=mpf.KSP_BAND).converged
    cut = P.solution()
    nblk, bmax, ndof, flops = P.band_stats()
    assert (nblk, ndof) == (1, len(rhs)) and (bmax > 176) == (shape != "beam")
    assert (bmax > 448) == (shape == "wide") and bmax <= 493
    assert P.band_segments() == (2, 1)
    P.factorise_and_solve(ksp=mpf.KSP_BAND)
    assert np.array_equal(P.solution(), cut)
    P.set_band_split(False)
    assert P.factorise_and_solve(ksp=mpf.KSP_BAND).converged
    whole = P.solution()
    assert P.band_segments() == (1, 0)
    for du in (cut, whole):
        assert normwise_rel(du, x_ref) < 1e-10
        assert np.linalg.norm(O.spmv(vals, du) - rhs) / np.linalg.norm(rhs) < 1e-12
    assert normwise_rel(cut, whole) < 1e-12 and not np.array_equal(cut, whole)


def test_band_rejects_mixed_system(mpf, gpu):
    X, conn = mpf.mesh_tet10(2, jitter=0.1, seed=2)
    P = mpf.Problem(mpf.TET10P1, X, conn, mpf.matl(1.0, 1e-3, 3, mixed=True))
    apply_bcs(P, tet_compress_bcs(2))
    P.prepare_matrix_pattern()
    P.set_boundary_conditions(1.0)
    P.calc_stiffness_and_residual(True, 1.0)
    with pytest.raises(mpf.MpfError):
        P.factorise_and_solve(ksp=mpf.KSP_BAND)


def test_gmres_mixed_vs_oracle(mpf, orc, gpu):
    import scipy.sparse as sp
    import scipy.sparse.linalg as spl
    n = 3
    X, conn = mpf.mesh_tet10(n, jitter=0.1, seed=2)
    P = mpf.Problem(mpf.TET10P1, X, conn, mpf.matl(1.0, 1e-3, 3, mixed=True))
    apply_bcs(P, tet_compress_bcs(n))
    P.prepare_matrix_pattern()
    P.set_boundary_conditions(1.0)
    P.calc_stiffness_and_residual(True, 1.0)
    vals, rhs = P.system()
    pat = P.pattern()
    A = sp.csr_matrix((vals, pat.colidx, pat.rowptr), shape=(len(rhs),) * 2)
    x_ref = spl.spsolve(A.tocsc(), rhs)
    r = P.factorise_and_solve(ksp=mpf.KSP_GMRES, pc=mpf.PC_BJACOBI3, rtol=1e-10, maxit=20000, restart=100)
    assert r.converged, (r.iterations, r.relres)
    du = P.solution()
    assert np.linalg.norm(A @ du - rhs) / np.linalg.norm(rhs) < 1e-9
    assert normwise_rel(du, x_ref) < 1e-6


def test_newton_cfg1_vs_oracle(mpf, orc, gpu):
    """10^3 Hex8 cube, 10% stretch over 10 steps, converged to ||R||/||R0|| < 1e-9 each step."""
    n = 10
    m = mpf.matl_from_E_nu(1e3, 0.3, tangent=mpf.TANGENT_CONSISTENT)
    X, conn, P = _hex_problem(mpf, n, m)
    opts = P.newton_opts(nsteps=10, dt=0.1, itmax=25, rtol=1e-9, ksp=mpf.KSP_CG, pc=mpf.PC_BJACOBI3,
                         lin_rtol=1e-12)
    iters, rep = P.solve_with_newton_raphson(opts)
    assert rep.steps_done == 10 and np.all(iters <= 6), iters
    O = _orc_hex_problem(orc, mpf, X, conn, n, m)
    st, oit, rel = O.newton(10, 0.1, solver=0, rtol=1e-9, itmax=25)
    assert st == 0
    ud, _ = P.state()
    uo, _ = O.state()
    assert normwise_rel(ud, uo) < 1e-9
    assert abs(ud[3 * M_face1(mpf, n)].mean() - 0.1) < 1e-12


def test_newton_cfg1_direct_vs_oracle(mpf, orc, gpu):
    """The plain method as specified: full Newton with a direct solve (banded LDL^T on the GPU)."""
    n = 10
    m = mpf.matl_from_E_nu(1e3, 0.3, tangent=mpf.TANGENT_CONSISTENT)
    X, conn, P = _hex_problem(mpf, n, m)
    opts = P.newton_opts(nsteps=10, dt=0.1, itmax=25, rtol=1e-9, ksp=mpf.KSP_BAND)
    iters, rep = P.solve_with_newton_raphson(opts)
    assert rep.steps_done == 10 and np.all(iters <= 6), iters
    O = _orc_hex_problem(orc, mpf, X, conn, n, m)
    st, oit, rel = O.newton(10, 0.1, solver=0, rtol=1e-9, itmax=25)
    assert st == 0
    ud, _ = P.state()
    uo, _ = O.state()
    assert normwise_rel(ud, uo) < 1e-9


def M_face1(mpf, n):
    return mpf.face_nodes(mpf.HEX8, n, n, n, 1)


def test_newton_tet_one_step_vs_oracle(mpf, orc, gpu):
    n = 3
    X, conn = mpf.mesh_tet10(n, jitter=0.1, seed=2)
    m = mpf.matl(1.0, 1e-3, 3, mixed=True)
    P = mpf.Problem(mpf.TET10P1, X, conn, m)
    O = orc.Problem(10, X, conn, 1.0, 1e-3, 3)
    for Q in (P, O):
        apply_bcs(Q, tet_compress_bcs(n))
        Q.prepare_matrix_pattern()
    opts = P.newton_opts(nsteps=2, dt=0.5, itmax=20, rtol=1e-11, ksp=mpf.KSP_GMRES, pc=mpf.PC_BJACOBI3,
                         lin_rtol=1e-13, maxit=50000, restart=100)
    iters, rep = P.solve_with_newton_raphson(opts)
    assert rep.steps_done == 2
    st, oit, _ = O.newton(2, 0.5, solver=2, precond=2, lin_rtol=1e-13, rtol=1e-11, itmax=20)
    assert st == 0
    ud, pd = P.state()
    uo, po = O.state()
    print("Tet10/P1 Newton: u", normwise_rel(ud, uo), "p", normwise_rel(pd, po), flush=True)
    assert normwise_rel(ud, uo) < 1e-9
    assert normwise_rel(pd, po) < 1e-9


# ---------------------------------------------------------------------------------------------
# Tet10 displacement element (Elem_Solid_3D_Disp with npElem = 10, Elem_Solid_3D_Disp.cpp:47-51)
# ---------------------------------------------------------------------------------------------
@pytest.mark.parametrize("tangent", [0, 1])
def test_tet10_golden_and_batch(mpf, orc, gpu, tangent):
    import torch
    g = np.load(GOLD / "tet10_element.npz")
    X1, u1 = g["X"], g["u"]
    Kexp = g["Ke_consistent"] - (g["Ke_volgeo"] if tangent == 0 else 0.0)
    m = mpf.matl(float(g["mu"]), float(g["Kinv"]), 3, tangent=tangent)
    Ke, Re, bad = mpf.elem_tet10_nh_batch(_t(gpu, X1), _t(gpu, u1.ravel()),
                                          _t(gpu, np.arange(10, dtype=np.int32)[None], torch.int32), m)
    assert bad == 0
    assert normwise_rel(Re.cpu().numpy()[0], g["Re"]) < 1e-12
    assert normwise_rel(Ke.cpu().numpy()[0], Kexp) < 1e-7  # FD fixture accuracy
    n = 3
    X, conn = mpf.mesh_tet10(n, jitter=0.1, seed=2)
    u = random_state(len(X), 0.01, seed=5)
    m = mpf.matl_from_E_nu(1e3, 0.3, tangent=tangent)
    Ke, Re, bad = mpf.elem_tet10_nh_batch(_t(gpu, X), _t(gpu, u), _t(gpu, conn, torch.int32), m)
    assert bad == 0
    Ke, Re = Ke.cpu().numpy(), Re.cpu().numpy()
    f = orc.elem_disp if tangent == 0 else orc.elem_disp_consistent
    worst = 0.0
    for e in range(0, len(conn), 5):
        Ko, Ro, st = f(X[conn[e]], u.reshape(-1, 3)[conn[e]], m.mu, m.Kinv, 3)
        assert st == 0
        worst = max(worst, normwise_rel(Ke[e], Ko), normwise_rel(Re[e], Ro))
    assert worst < TOL_K, worst
    Ke2, Re2, _ = mpf.elem_tet10_nh_batch(_t(gpu, X), _t(gpu, u), _t(gpu, conn, torch.int32), m, first=7, count=9)
    assert np.array_equal(Ke2.cpu().numpy(), Ke[7:16]) and np.array_equal(Re2.cpu().numpy(), Re[7:16])


def _tet10_problems(mpf, orc, n, tangent, mode=0):
    X, conn = mpf.mesh_tet10(n, jitter=0.1, seed=2)
    m = mpf.matl_from_E_nu(1e3, 0.3, tangent=tangent)
    P = mpf.Problem(mpf.TET10, X, conn, m)
    P.set_assembly_mode(mode)
    O = orc.Problem(11, X, conn, m.mu, m.Kinv, 3)
    O.set_tangent(tangent == 1)
    for Q in (P, O):
        apply_bcs(Q, tet_compress_bcs(n))
        Q.prepare_matrix_pattern()
    return X, conn, P, O


def test_pattern_and_colouring_bitexact_tet10(mpf, orc, gpu):
    _, _, P, O = _tet10_problems(mpf, orc, 4, 1)
    pat = P.pattern()
    rowptr, colidx, nda, _ = O.pattern()
    assert pat.presDOF == 0
    assert np.array_equal(pat.rowptr, rowptr) and np.array_equal(pat.colidx, colidx)
    assert np.array_equal(pat.NodeDofArray, nda)
    col, ncol = O.color()
    assert pat.ncolors == ncol and np.array_equal(pat.color, col)


@pytest.mark.parametrize("mode", [0, 1, 2], ids=["gather", "colored_scatter", "gather_direct"])
@pytest.mark.parametrize("tangent", [0, 1])
def test_global_assembly_tet10_vs_oracle(mpf, orc, gpu, tangent, mode):
    X, conn, P, O = _tet10_problems(mpf, orc, 4, tangent, mode)
    u = random_state(len(X), 0.004, seed=33)
    for Q in (P, O):
        Q.set_state(disp=u, disp_prev=u)
        Q.set_boundary_conditions(0.8)
    for first in (True, False):
        P.calc_stiffness_and_residual(first, 0.8)
        vals, rhs = P.system()
        P.calc_stiffness_and_residual(first, 0.8)
        v2, r2 = P.system()
        assert np.array_equal(vals, v2) and np.array_equal(rhs, r2)
        ov, orhs, st = O.assemble(first, 0.8)
        assert st == 0
        assert normwise_rel(vals, ov) < TOL_K
        assert normwise_rel(rhs, orhs) < TOL_K
    assert P.assembly_mode() == mode


@pytest.mark.parametrize("ksp", ["band", "pcg"])
def test_newton_tet10_vs_oracle(mpf, orc, gpu, ksp):
    """Two load steps of 5% compression with the Tet10 displacement element: full Newton with the
    banded direct solve, and with block-Jacobi PCG, against the oracle's Newton (direct solve)."""
    X, conn, P, O = _tet10_problems(mpf, orc, 3, 1)
    k = mpf.KSP_BAND if ksp == "band" else mpf.KSP_CG
    opts = P.newton_opts(nsteps=2, dt=0.5, itmax=20, rtol=1e-11, ksp=k, pc=mpf.PC_BJACOBI3, lin_rtol=1e-13,
                         maxit=50000)
    iters, rep = P.solve_with_newton_raphson(opts)
    assert rep.steps_done == 2
    st, oit, _ = O.newton(2, 0.5, solver=0, rtol=1e-11, itmax=20)
    assert st == 0
    assert np.array_equal(iters, oit), (iters, oit)
    assert normwise_rel(P.state()[0], O.state()[0]) < 1e-9
