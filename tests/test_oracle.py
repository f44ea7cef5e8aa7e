"""CPU pins of the oracle (oracle/oracle.cpp), the restatement of the reference path.

The reference ships no tests and no golden vectors for this path, so each oracle function is
pinned by mathematics that a plausible mistake (dropped term, wrong sign or index, transposed
operand) would fail:

  quadrature          exactness on polynomials of the design degree; volumes 8 and 1/6
  basis functions     partition of unity, zero-sum derivatives, Kronecker property at the nodes
  det_inv3            numpy.linalg on random matrices
  neohooke            F = I gives sigma = 0 and the isotropic tangent mu(d_il d_jk + d_ik d_jl)
                      + (K - 2mu/3) d_ij d_kl; frame indifference; sigma against the energy FD
                      fixture; the reference tangent minus the consistent one is U' d_jl d_ik
  Hex8 element        golden fixture (independent numpy R_e; K = -dR/du by FD; K_ref = K_FD - K_volgeo),
                      rigid modes at F = I, symmetry, homogeneous-deformation patch test
  Tet10/P1 element    golden fixture (independent numpy residual; full 34x34 tangent by FD)
  Tet10 element       golden fixture (independent numpy residual; K = -dR/du by FD, both tangent
                      modes), rigid modes, symmetry
  numbering/pattern   brute-force dense coupling on tiny meshes (bit-exact), nnz of the
                      CPU-runnable 10^3 workload = 231 601, symmetric and sorted rows
  colouring           validity (no shared node within a colour), determinism, greedy order
  solvers             scipy sparse direct solve
  Newton              quadratic convergence with the consistent tangent, and the same converged
                      displacement from both tangents where the reference tangent converges
  Hex8 Q1/P0 element  FD of the uncondensed 25x25 mixed tangent, reference-state closed forms,
                      the condensation identity
  Q1/P0 driver        ends at the uncondensed mixed equilibrium (Fu = 0, Fp = 0 per element),
                      relieves volumetric locking as nu -> 1/2; the oracle's Q1/P0 problem
                      reproduces it
  Quad4 traction      closed forms on a flat face: v0 A / 4 per node along n = dx/deta x dx/dxi (the
                      order of the face's nodes fixes its sign), the unnormalised tangent term
                      v1 (h/2) A / 4 per node, zero net force and moment on a closed surface
  Dirichlet lift      the first Newton iterate after the lift is the linearised solution: its error
                      to the converged displacement is O(delta^2) for a prescribed stretch delta,
                      while without the lift it stays O(delta)
"""
from __future__ import annotations

from pathlib import Path

import numpy as np
import pytest

from tests.cases import apply_bcs, cube_hex, hex_stretch_bcs, normwise_rel, q1p0_newton

GOLD = Path(__file__).resolve().parent / "golden"


def test_quadrature_exactness(orc):
    xi, w = orc.gauss3d(8)
    assert len(w) == 8 and abs(w.sum() - 8.0) < 1e-15
    # 2-point Gauss is exact up to degree 3 per axis; the reference literal 0.577350269189626 is 15 digits
    for a, b, c in [(2, 0, 0), (0, 2, 2), (2, 2, 2), (3, 1, 0)]:
        exact = np.prod([0.0 if k % 2 else 2.0 / (k + 1) for k in (a, b, c)])
        val = np.sum(w * xi[:, 0] ** a * xi[:, 1] ** b * xi[:, 2] ** c)
        assert abs(val - exact) < 1e-13
    xi, w = orc.gauss3d(10)
    assert len(w) == 4 and abs(w.sum() - 1.0 / 6.0) < 1e-16
    # degree-2 exactness on the unit tet: int x^2 = 1/60, int xy = 1/120
    assert abs(np.sum(w * xi[:, 0] ** 2) - 1.0 / 60.0) < 1e-15
    assert abs(np.sum(w * xi[:, 0] * xi[:, 1]) - 1.0 / 120.0) < 1e-15


def test_basis_functions(orc):
    rng = np.random.default_rng(0)
    for _ in range(5):
        xi = rng.uniform(-1, 1, 3)
        N, dN = orc.basis_hex8(xi)
        assert abs(N.sum() - 1) < 1e-15 and np.abs(dN.sum(1)).max() < 1e-15
        t = rng.dirichlet(np.ones(4))[1:]
        N, dN = orc.basis_tet(10, t)
        assert abs(N.sum() - 1) < 1e-14 and np.abs(dN.sum(1)).max() < 1e-14
    nodes = [(-1, -1, -1), (1, -1, -1), (1, 1, -1), (-1, 1, -1), (-1, -1, 1), (1, -1, 1), (1, 1, 1), (-1, 1, 1)]
    for a, x in enumerate(nodes):
        N, _ = orc.basis_hex8(np.array(x, float))
        assert np.allclose(N, np.eye(8)[a], atol=0)
    V = np.array([[0, 0, 0], [1, 0, 0], [0, 1, 0], [0, 0, 1]], float)
    tet = list(V) + [(V[a] + V[b]) / 2 for a, b in [(0, 1), (1, 2), (0, 2), (0, 3), (2, 3), (1, 3)]]
    for a, x in enumerate(tet):
        N, _ = orc.basis_tet(10, x)
        assert np.allclose(N, np.eye(10)[a], atol=1e-15)


def test_det_inv3(orc):
    rng = np.random.default_rng(1)
    for _ in range(20):
        A = rng.normal(size=(3, 3))
        d, Ai = orc.det_inv3(A)
        assert abs(d - np.linalg.det(A)) < 1e-13 * max(1, abs(d))
        assert np.abs(Ai - np.linalg.inv(A)).max() < 1e-12 * np.abs(np.linalg.inv(A)).max()


def test_neohooke_reference_state(orc):
    mu, Kinv = 384.6153846153846, 1.2e-3
    K = 1.0 / Kinv
    s, Cm = orc.neohooke(np.eye(3), mu, Kinv, 3)
    assert np.abs(s).max() < 1e-12
    mp = [(0, 0), (1, 0), (2, 0), (0, 1), (1, 1), (2, 1), (0, 2), (1, 2), (2, 2)]
    d = np.eye(3)
    Ciso = np.array([[mu * (d[i, l] * d[j, k] + d[i, k] * d[j, l]) + (K - 2 * mu / 3) * d[i, j] * d[k, l]
                      for (k, l) in mp] for (i, j) in mp])
    assert normwise_rel(Cm, Ciso) < 1e-14


def test_neohooke_frame_indifference_and_quirk(orc):
    rng = np.random.default_rng(2)
    mu, Kinv = 10.0, 1e-2
    F = np.eye(3) + rng.uniform(-0.1, 0.1, (3, 3))
    Q, _ = np.linalg.qr(rng.normal(size=(3, 3)))
    if np.linalg.det(Q) < 0:
        Q[:, 0] *= -1
    s1, _ = orc.neohooke(F, mu, Kinv, 3)
    s2, _ = orc.neohooke(Q @ F, mu, Kinv, 3)
    S1 = s1.reshape(3, 3).T  # column-major 9-map -> matrix
    S2 = s2.reshape(3, 3).T
    assert normwise_rel(S2, Q @ S1 @ Q.T) < 1e-13
    # the reference tangent carries -U' d_jl d_ik in its volumetric part (utilitiesmaterialHyperelastic.cpp:280-281)
    _, Cref = orc.neohooke(F, mu, Kinv, 3)
    J = np.linalg.det(F)
    U1 = (J - 1.0) / Kinv
    mp = [(0, 0), (1, 0), (2, 0), (0, 1), (1, 1), (2, 1), (0, 2), (1, 2), (2, 2)]
    d = np.eye(3)
    term = np.array([[-U1 * d[j, l] * d[i, k] for (k, l) in mp] for (i, j) in mp])
    # Cref - term must be symmetric in the way the consistent tangent is (major symmetry) and
    # equal the element-level consistent K (checked in the element test); here check the term
    # is the only difference between mixed-with-p=U' and displacement forms' volumetric parts
    _, Cmix = orc.neohooke(F, mu, Kinv, 3, mixed=True, pres=U1)
    # mixed form: p (d_ij d_kl - d_jk d_il); displacement: (J U''+U') d_ij d_kl - U' d_il d_jk - U' d_jl d_ik
    U2 = 1.0 / Kinv
    extra = np.array([[J * U2 * d[i, j] * d[k, l] for (k, l) in mp] for (i, j) in mp])
    assert normwise_rel(Cref - term, Cmix + extra) < 1e-13


def test_neohooke_stress_golden(orc):
    g = np.load(GOLD / "hex8_element.npz")
    mu, Kinv = float(g["mu"]), float(g["Kinv"])
    rng = np.random.default_rng(3)
    F = np.eye(3) + rng.uniform(-0.05, 0.05, (3, 3))
    s, _ = orc.neohooke(F, mu, Kinv, 3)
    J = np.linalg.det(F)
    b = F @ F.T
    ref = mu * J ** (-5 / 3) * (b - np.trace(b) / 3 * np.eye(3)) + (J - 1) / Kinv * np.eye(3)
    assert normwise_rel(s.reshape(3, 3).T, ref) < 1e-13


def test_hex8_element_golden(orc):
    g = np.load(GOLD / "hex8_element.npz")
    X, u, mu, Kinv = g["X"], g["u"], float(g["mu"]), float(g["Kinv"])
    Ke, Re, st = orc.elem_disp(X, u, mu, Kinv, 3)
    assert st == 0
    assert normwise_rel(Re, g["Re"]) < 1e-12
    # reference tangent = consistent (FD) - volumetric geometric term; FD accuracy ~1e-8
    assert normwise_rel(Ke, g["Ke_consistent"] - g["Ke_volgeo"]) < 1e-7
    Kc, Rc, _ = orc.elem_disp_consistent(X, u, mu, Kinv, 3)
    assert normwise_rel(Kc, g["Ke_consistent"]) < 1e-7
    assert np.array_equal(Rc, Re)
    assert normwise_rel(Ke, Ke.T) < 1e-14
    # the volumetric term is not negligible: a dropped term would not pass
    assert normwise_rel(Kc, Ke) > 1e-4


def test_hex8_rigid_modes(orc):
    g = np.load(GOLD / "hex8_element.npz")
    X = g["X"]
    Ke, Re, _ = orc.elem_disp(X, np.zeros((8, 3)), 100.0, 1e-2, 3)
    assert np.abs(Re).max() < 1e-12
    modes = []
    for d in range(3):
        t = np.zeros((8, 3)); t[:, d] = 1.0
        modes.append(t.ravel())
    for w in np.eye(3):
        modes.append(np.cross(w, X).ravel())
    for r in modes:
        assert np.abs(Ke @ r).max() < 1e-10 * np.abs(Ke).max()


def test_tet10p1_element_golden(orc):
    g = np.load(GOLD / "tet10p1_element.npz")
    Kuu, Kup, Kpu, Kpp, Fu, Fp, st = orc.elem_tet10p1(g["X"], g["u"], g["uprev"], g["pres"], float(g["mu"]),
                                                       float(g["Kinv"]), 3)
    assert st == 0
    R = np.concatenate([Fu, Fp])
    assert normwise_rel(R, g["R"]) < 1e-12
    K = np.block([[Kuu, Kup], [Kpu, Kpp]])
    assert normwise_rel(K, g["K"]) < 1e-6
    assert normwise_rel(Kup, Kpu.T) < 1e-14


def test_patch_test_homogeneous(orc, mpf):
    # u = (F0 - I) X on a jittered mesh: interior nodes carry zero net internal force
    n = 3
    X, conn = cube_hex(n, jitter=0.1, seed=7)
    F0 = np.array([[1.05, 0.02, 0.0], [0.01, 0.97, 0.03], [0.0, -0.02, 1.02]])
    u = (X @ (F0 - np.eye(3)).T).ravel()
    P = orc.Problem(8, X, conn, 50.0, 1e-2, 3)
    P.prepare_matrix_pattern()
    P.set_state(disp=u)
    vals, rhs, st = P.assemble(False, 1.0)
    assert st == 0
    _, _, nda, _ = P.pattern()
    interior = [i for i in range(len(X)) if np.all((X[i] > 1e-12) & (X[i] < 1 - 1e-12))]
    rows = nda[interior].ravel()
    assert np.abs(rhs[rows]).max() < 1e-11 * np.abs(rhs).max()


@pytest.mark.parametrize("key", ["hex_111", "hex_222", "hex_321"])
def test_pattern_bruteforce(orc, key):
    g = np.load(GOLD / "patterns.npz")
    X, conn = g[key + "_X"], g[key + "_conn"]
    P = orc.Problem(8, X, conn, 1.0, 1e-2, 3)
    x0 = np.flatnonzero(X[:, 0] == 0.0)
    x1 = np.flatnonzero(X[:, 0] == 1.0)
    P.add_dirichlet(x0, -1)
    P.add_dirichlet(x1, 0, 0.1)
    P.prepare_matrix_pattern()
    rowptr, colidx, nda, _ = P.pattern()
    assert np.array_equal(rowptr, g[key + "_rowptr"])
    assert np.array_equal(colidx, g[key + "_colidx"])
    assert np.array_equal(nda, g[key + "_nda"])


@pytest.mark.parametrize("etype", [10, 11], ids=["tet10p1", "tet10"])
def test_pattern_tet10_bruteforce(orc, mpf, etype):
    from tests.golden.make_golden import brute_pattern
    X, conn = mpf.mesh_tet10(2, jitter=0.1, seed=2)
    bot = mpf.face_nodes(mpf.TET10P1, 2, 2, 2, 4)
    top = mpf.face_nodes(mpf.TET10P1, 2, 2, 2, 5)
    P = orc.Problem(etype, X, conn, 1.0, 1e-3, 3)
    P.add_dirichlet(bot, -1)
    P.add_dirichlet(top, 2, -0.05)
    P.prepare_matrix_pattern()
    rowptr, colidx, nda, idp = P.pattern()
    fixed = {3 * int(v) + d for v in bot for d in range(3)} | {3 * int(v) + 2 for v in top}
    rp, ci, nd, ip = brute_pattern(len(X), conn, fixed, etype == 10)
    assert np.array_equal(rowptr, rp) and np.array_equal(colidx, ci)
    assert np.array_equal(nda, nd)
    if etype == 10:
        assert np.array_equal(idp, ip)
    else:
        assert P.sizes()[1] == 0  # no pressure dofs


def test_tet10_element_golden(orc):
    """Tet10 displacement element (npElem 10, 4-point rule): R_e against the independent numpy
    residual, K_e against its FD Jacobian in both tangent modes, symmetry, rigid-body modes."""
    g = np.load(GOLD / "tet10_element.npz")
    X, u, mu, Kinv = g["X"], g["u"], float(g["mu"]), float(g["Kinv"])
    Ke, Re, st = orc.elem_disp(X, u, mu, Kinv, 3)
    assert st == 0 and Ke.shape == (30, 30)
    assert normwise_rel(Re, g["Re"]) < 1e-12
    assert normwise_rel(Ke, g["Ke_consistent"] - g["Ke_volgeo"]) < 1e-7
    Kc, Rc, _ = orc.elem_disp_consistent(X, u, mu, Kinv, 3)
    assert normwise_rel(Kc, g["Ke_consistent"]) < 1e-7
    assert np.array_equal(Rc, Re) and normwise_rel(Ke, Ke.T) < 1e-14
    assert normwise_rel(Kc, Ke) > 1e-4
    K0, R0, _ = orc.elem_disp(X, np.zeros((10, 3)), 100.0, 1e-2, 3)
    assert np.abs(R0).max() < 1e-12
    modes = [np.tile(np.eye(3)[d], 10) for d in range(3)] + [np.cross(w, X).ravel() for w in np.eye(3)]
    for r in modes:
        assert np.abs(K0 @ r).max() < 1e-10 * np.abs(K0).max()


def test_pattern_cfg1_counts(orc):
    n = 10
    X, conn = cube_hex(n)
    P = orc.Problem(8, X, conn, 1.0, 1e-2, 3)
    apply_bcs(P, hex_stretch_bcs(n))
    P.prepare_matrix_pattern()
    d, p, nnz = P.sizes()
    assert (d, p, nnz) == (3509, 0, 231601)
    rowptr, colidx, _, _ = P.pattern()
    for r in range(0, d, 97):
        row = colidx[rowptr[r]:rowptr[r + 1]]
        assert np.all(np.diff(row) > 0)
    # structural symmetry
    import scipy.sparse as sp
    A = sp.csr_matrix((np.ones(nnz), colidx, rowptr), shape=(d, d))
    assert (A - A.T).nnz == 0


def _greedy_color(conn, prio):
    nE = conn.shape[0]
    node_el = {}
    for e, c in enumerate(conn):
        for v in c:
            node_el.setdefault(int(v), []).append(e)
    color = -np.ones(nE, int)
    for e in sorted(range(nE), key=lambda e: -prio[e]):
        used = {color[f] for v in conn[e] for f in node_el[int(v)] if color[f] >= 0}
        c = 0
        while c in used:
            c += 1
        color[e] = c
    return color


def _splitmix_prio(e):
    z = (e + 0x9E3779B97F4A7C15) & 0xFFFFFFFFFFFFFFFF
    z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & 0xFFFFFFFFFFFFFFFF
    z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & 0xFFFFFFFFFFFFFFFF
    return z ^ (z >> 31)


def test_color_priority(orc):
    for e in [0, 1, 2, 12345, 2 ** 31 - 1]:
        assert orc.color_priority(e) == _splitmix_prio(e)


def test_coloring(orc, mpf):
    X, conn = cube_hex(4)
    P = orc.Problem(8, X, conn, 1.0, 1e-2, 3)
    P.prepare_matrix_pattern()
    c1, n1 = P.color()
    c2, _ = P.color()
    assert np.array_equal(c1, c2)
    assert np.array_equal(c1, _greedy_color(conn, [_splitmix_prio(e) for e in range(len(conn))]))
    for c in range(n1):
        nodes = conn[c1 == c].ravel()
        assert len(nodes) == len(np.unique(nodes))


@pytest.mark.parametrize("etype", [8, 10])
def test_colored_parallel_assembly_matches_serial(orc, mpf, etype):
    """The oracle's colour-parallel loop (the all-cores CPU baseline) assembles the serial loop's
    K and R up to summation order."""
    if etype == 8:
        X, conn = cube_hex(4, jitter=0.1, seed=3)
        P = orc.Problem(8, X, conn, 50.0, 1e-2, 3)
    else:
        X, conn = mpf.mesh_tet10(2, jitter=0.1, seed=2)
        P = orc.Problem(10, X, conn, 1.0, 1e-3, 3)
    P.prepare_matrix_pattern()
    rng = np.random.default_rng(5)
    P.set_state(disp=0.01 * rng.standard_normal(3 * len(X)))
    d, pp, nnz = P.sizes()
    v1, r1 = np.zeros(nnz), np.zeros(d + pp)
    v2, r2 = np.zeros(nnz), np.zeros(d + pp)
    assert P.assemble_range(0, len(conn), v1, r1) == 0
    color, _ = P.color()
    assert P.assemble_colored(color, 4, v2, r2) == 0
    assert normwise_rel(v2, v1) < 1e-13 and normwise_rel(r2, r1) < 1e-13


def test_solvers_vs_scipy(orc):
    import scipy.sparse as sp
    import scipy.sparse.linalg as spl
    n = 5
    X, conn = cube_hex(n)
    P = orc.Problem(8, X, conn, 384.6, 1.2e-3, 3)
    apply_bcs(P, hex_stretch_bcs(n))
    P.prepare_matrix_pattern()
    P.set_boundary_conditions(0.5)
    vals, rhs, st = P.assemble(True, 0.5)
    rowptr, colidx, _, _ = P.pattern()
    A = sp.csr_matrix((vals, colidx, rowptr), shape=(len(rhs), len(rhs)))
    x_ref = spl.spsolve(A.tocsc(), rhs)
    x_d, st = P.direct_solve(vals, rhs)
    assert st == 0 and normwise_rel(x_d, x_ref) < 1e-10
    x_c, it, rr = P.pcg(vals, rhs, 2, 1e-12)
    assert rr < 1e-12 and normwise_rel(x_c, x_ref) < 1e-9
    x_g, it, rr = P.gmres(vals, rhs, 2, 30, 1e-12)
    assert rr < 1e-11 and normwise_rel(x_g, x_ref) < 1e-9
    assert normwise_rel(P.spmv(vals, x_ref), A @ x_ref) < 1e-15


def test_newton_consistent_quadratic_and_tangent_independent_solution(orc):
    n = 4
    X, conn = cube_hex(n)
    sols = []
    for consistent in (True, False):
        P = orc.Problem(8, X, conn, 384.6153846153846, 1.2e-3, 3)
        P.set_tangent(consistent)
        apply_bcs(P, hex_stretch_bcs(n))
        P.prepare_matrix_pattern()
        st, iters, rel = P.newton(3, 0.1, solver=0, rtol=1e-11, itmax=40)
        assert st == 0, (consistent, iters)
        if consistent:
            assert np.all(iters <= 5)
        sols.append(P.state()[0])
    assert normwise_rel(sols[0], sols[1]) < 1e-9


def _hex8_element(jitter, seed):
    X, conn = cube_hex(1, jitter=jitter, seed=seed)
    return X[conn[0]]


def test_hex8p0_element_closed_forms(orc):
    """Hex8 Q1/P0 in its reference state (unit cube, V0 = 1, F = I, so the stress is p I):
    Kpp = -V0/K, Fp = p V0/K, Kpu = Kup^T = int grad N dV with zero sum per direction, Fu = -p Kup."""
    Xe = _hex8_element(0.0, 0)
    u = np.zeros(24)
    pe, mu, Kinv = 0.7, 50.0, 1e-2
    Kc, Fc, Kuu, Fu, Kup, Kpu, Kpp, Fp, st = orc.elem_hex8p0(Xe, u, u, pe, mu, Kinv)
    assert st == 0
    assert abs(Kpp - (-Kinv)) < 1e-15 and abs(Fp - Kinv * pe) < 1e-15
    assert np.array_equal(Kpu, Kup)
    assert np.abs(Kup.reshape(8, 3).sum(axis=0)).max() < 1e-14
    assert normwise_rel(Fu, -pe * Kup) < 1e-13


def _hex8p0_state(orc):
    rng = np.random.default_rng(11)
    Xe = _hex8_element(0.0, 3) + 0.06 * rng.standard_normal((8, 3))  # a general hexahedron
    u = 0.03 * rng.standard_normal(24)
    up = 0.01 * rng.standard_normal(24)
    return Xe, u, up, 0.4, 50.0, 1e-2


def test_hex8p0_element_fd_tangent(orc):
    """[[Kuu, Kup], [Kpu, Kpp]] against central differences of -[Fu; Fp] in (u, p)."""
    Xe, u, up, pe, mu, Kinv = _hex8p0_state(orc)
    _, _, Kuu, Fu, Kup, Kpu, Kpp, Fp, st = orc.elem_hex8p0(Xe, u, up, pe, mu, Kinv)
    assert st == 0
    K = np.zeros((25, 25))
    K[:24, :24], K[:24, 24], K[24, :24], K[24, 24] = Kuu, Kup, Kpu, Kpp

    def resid(z):
        r = orc.elem_hex8p0(Xe, z[:24], up, z[24], mu, Kinv)
        return np.concatenate([r[3], [r[7]]])

    z0 = np.concatenate([u, [pe]])
    h = 1e-6
    Kfd = np.zeros((25, 25))
    for j in range(25):
        e = np.zeros(25)
        e[j] = h
        Kfd[:, j] = -(resid(z0 + e) - resid(z0 - e)) / (2 * h)
    assert normwise_rel(Kfd, K) < 1e-6


def test_hex8p0_condensation_identity(orc):
    """Solving the condensed system gives the displacement of the full mixed system, and the
    pressure recovery (Fp - Kpu du) / Kpp its pressure increment."""
    Xe, u, up, pe, mu, Kinv = _hex8p0_state(orc)
    Kc, Fc, Kuu, Fu, Kup, Kpu, Kpp, Fp, st = orc.elem_hex8p0(Xe, u, up, pe, mu, Kinv)
    assert st == 0
    K = np.zeros((25, 25))
    K[:24, :24], K[:24, 24], K[24, :24], K[24, 24] = Kuu, Kup, Kpu, Kpp
    keep = np.arange(3, 24)  # node 0 held: the element alone is singular in its rigid modes
    full = np.concatenate([keep, [24]])
    sol = np.linalg.solve(K[np.ix_(full, full)], np.concatenate([Fu, [Fp]])[full])
    du = np.linalg.solve(Kc[np.ix_(keep, keep)], Fc[keep])
    assert normwise_rel(du, sol[:-1]) < 1e-10
    dfull = np.zeros(24)
    dfull[keep] = du
    assert abs((Fp - Kpu @ dfull) / Kpp - sol[-1]) <= 1e-10 * max(1.0, abs(sol[-1]))


def test_q1p0_newton_driver_reaches_mixed_equilibrium(orc):
    """The CPU Q1/P0 driver (condensed solve, element pressure recovery after each correction) ends
    where the uncondensed mixed equations hold: Fu = 0 on the free dofs and Fp = 0 in every element
    (a wrong recovery leaves Fp nonzero while the condensed residual still vanishes), the prescribed
    face sits at its value, and Newton converges quadratically."""
    n, mu, Kinv = 2, 1.0, 1e-3
    X, conn = cube_hex(n, jitter=0.1, seed=3)
    bcs = hex_stretch_bcs(n)
    u, p, iters = q1p0_newton(orc, X, conn, bcs, mu, Kinv, nsteps=2, dt=0.5, rtol=1e-11)
    assert max(iters) <= 6, iters
    assert np.allclose(u[3 * bcs[1][0]], 0.1, atol=1e-13)
    fixed = set((3 * np.asarray(bcs[0][0])[:, None] + np.arange(3)).ravel().tolist()) | set((3 * bcs[1][0]).tolist())
    Fu_all = np.zeros_like(u)
    fp_max = 0.0
    for e, c in enumerate(conn):
        _, _, _, Fu, _, _, Kpp, Fp, st = orc.elem_hex8p0(X[c], u.reshape(-1, 3)[c], u.reshape(-1, 3)[c], p[e], mu, Kinv)
        assert st == 0
        Fu_all[(3 * np.asarray(c)[:, None] + np.arange(3)).ravel()] += Fu
        fp_max = max(fp_max, abs(Fp / Kpp))
    free = np.array([d for d in range(len(u)) if d not in fixed])
    assert np.linalg.norm(Fu_all[free]) < 1e-9 * mu
    assert fp_max < 1e-9 * max(1.0, np.abs(p).max())
    assert np.abs(p).max() > 1e-3  # a stretched near-incompressible block carries pressure


def _clamped_block_deflection(orc, mpf, nu, elem):
    """Mean top-face deflection of a 4^3 unit block, bottom clamped, top under a small normal dead
    load (E = 1, linear regime), by the oracle's Hex8 problem or the CPU Q1/P0 driver."""
    from tests.cases import oracle_fext
    n = 4
    X, conn = mpf.mesh_hex(n, n, n, jitter=0.0, seed=5)
    bot, top = mpf.face_nodes(mpf.HEX8, n, n, n, 4), mpf.face_nodes(mpf.HEX8, n, n, n, 5)
    faces = mpf.hex_face_quads(n, n, n, 5)
    m = mpf.matl_from_E_nu(1.0, nu)
    if elem == "hex8":
        Q = orc.Problem(8, X, conn, m.mu, m.Kinv, 3)
        Q.set_tangent(True)
        Q.add_dirichlet(bot, -1)
        Q.add_traction(faces, -1e-3)
        Q.prepare_matrix_pattern()
        st, _, _ = Q.newton(1, 1.0, solver=0, precond=2, lin_rtol=1e-12, rtol=1e-8, itmax=30)
        assert st == 0
        u, _ = Q.state()
    else:
        bcs = [(bot, -1, None)]
        fext = oracle_fext(orc, X, conn, bcs, faces, -1e-3)
        u, _, _ = q1p0_newton(orc, X, conn, bcs, m.mu, m.Kinv, 1, 1.0, rtol=1e-8, fext=fext)
    return u[3 * top + 2].mean()


def test_q1p0_relieves_volumetric_locking(orc, mpf):
    """What the Q1/P0 element is for: in a clamped block under a top load the displacement Hex8
    stiffens without bound as nu -> 1/2 (volumetric locking: 26x less deflection at nu = 0.4999 than
    at 0.3), while the Q1/P0 deflection tends to a finite incompressible limit; both agree in the
    compressible range."""
    h3, p3 = (_clamped_block_deflection(orc, mpf, 0.3, e) for e in ("hex8", "q1p0"))
    assert abs(p3 / h3 - 1.0) < 0.02
    h5, p5 = (_clamped_block_deflection(orc, mpf, 0.4999, e) for e in ("hex8", "q1p0"))
    assert p5 / h5 > 10.0
    p4 = _clamped_block_deflection(orc, mpf, 0.499, "q1p0")
    assert abs(p5 / p4 - 1.0) < 0.01
    assert h5 / h3 < 0.1 and p5 / p3 > 0.8


def test_q1p0_oracle_problem_matches_driver(orc):
    """The C oracle's Q1/P0 problem (elem_type 9: condensed scatter, per-element pressure recovery in
    its Newton update) reproduces the pinned CPU Q1/P0 driver."""
    n = 3
    X, conn = cube_hex(n, jitter=0.1, seed=3)
    bcs = hex_stretch_bcs(n)
    Q = orc.Problem(9, X, conn, 1.0, 1e-3, 3)
    apply_bcs(Q, bcs)
    Q.prepare_matrix_pattern()
    st, it, _ = Q.newton(2, 0.5, solver=0, rtol=1e-10, itmax=20)
    assert st == 0
    uq, pq = Q.state()
    u, p, its = q1p0_newton(orc, X, conn, bcs, 1.0, 1e-3, 2, 0.5, rtol=1e-10)
    assert list(it) == its
    assert normwise_rel(uq, u) < 1e-12 and normwise_rel(pq, p) < 1e-12


# ---------------------------------------------------------------------------------------------
# Quad4 dead-load traction (femSolidmechanics1.cpp:1026-1070; BasisFunctionsLagrange.cpp:1688-1732)
# and the first-iteration Dirichlet lift (ElementBase.cpp:567-594), pinned by closed forms.
# ---------------------------------------------------------------------------------------------
def _cube_element(h):
    """One Hex8 element [0,h]^3 in the reference's node order."""
    X = h * np.array([[0, 0, 0], [1, 0, 0], [1, 1, 0], [0, 1, 0], [0, 0, 1], [1, 0, 1], [1, 1, 1], [0, 1, 1]], float)
    return X, np.arange(8, dtype=np.int32)[None, :]


def test_traction_flat_face_normal_load_closed_form(orc):
    """A flat h x h face under v0 alone: every node receives v0 A / 4 along n = dx/deta x dx/dxi.
    For the top face listed 4,5,6,7 (counter-clockwise seen from +z) dx/dxi = +x h/2 and
    dx/deta = +y h/2, so n = y x x = -z; listed 4,7,6,5 it is +z.  A normal taken as du x dv
    instead, or left unnormalised, fails one of the two."""
    from tests.cases import oracle_fext
    h, v0 = 0.7, 3.5
    X, conn = _cube_element(h)
    A = h * h
    for quad, nz in (([4, 5, 6, 7], -1.0), ([4, 7, 6, 5], 1.0)):
        f = oracle_fext(orc, X, conn, [], np.array([quad], dtype=np.int32), v0, 0.0).reshape(-1, 3)
        # the bottom nodes carry only the rounding-level internal force of the stress-free state
        assert np.abs(f[:4]).max() < 1e-13
        np.testing.assert_allclose(f[4:], np.tile([0.0, 0.0, v0 * A / 4 * nz], (4, 1)), rtol=0, atol=1e-13)


def test_traction_tangent_term_is_unnormalised(orc):
    """The v1 term uses tangent1 = dx/dxi without normalising it: on an h x h face it adds
    v1 (h/2) A e_xi in total, v1 (h/2) A / 4 per node; a unit tangent would give v1 A."""
    from tests.cases import oracle_fext
    h, v1 = 0.6, -2.0
    X, conn = _cube_element(h)
    A = h * h
    f = oracle_fext(orc, X, conn, [], np.array([[4, 5, 6, 7]], dtype=np.int32), 0.0, v1).reshape(-1, 3)
    np.testing.assert_allclose(f[4:], np.tile([v1 * (h / 2) * A / 4, 0.0, 0.0], (4, 1)), rtol=0, atol=1e-13)
    # dx/dxi runs along +y when the same face is listed 4,7,6,5
    f = oracle_fext(orc, X, conn, [], np.array([[4, 7, 6, 5]], dtype=np.int32), 0.0, v1).reshape(-1, 3)
    np.testing.assert_allclose(f[4:], np.tile([0.0, v1 * (h / 2) * A / 4, 0.0], (4, 1)), rtol=0, atol=1e-13)
    # the two terms superpose
    f0 = oracle_fext(orc, X, conn, [], np.array([[4, 5, 6, 7]], dtype=np.int32), 1.25, 0.0)
    f1 = oracle_fext(orc, X, conn, [], np.array([[4, 5, 6, 7]], dtype=np.int32), 0.0, v1)
    f01 = oracle_fext(orc, X, conn, [], np.array([[4, 5, 6, 7]], dtype=np.int32), 1.25, v1)
    np.testing.assert_allclose(f01, f0 + f1, rtol=0, atol=1e-13)


def test_traction_closed_surface_zero_force_and_moment(mpf, orc):
    """Uniform normal traction on all six faces of a jittered block (boundary faces planar, their
    quads distorted): with outward normals every face's force is v0 times its area vector, and over
    the closed surface the net force and the net moment about the origin vanish (2x2 Gauss points
    integrate x times the bilinear surface Jacobian exactly on planar quads)."""
    from tests.cases import oracle_fext
    nx, ny, nz = 3, 2, 4
    L = (1.2, 0.8, 1.6)
    X, conn = mpf.mesh_hex(nx, ny, nz, L=L, jitter=0.15, seed=7)
    v0 = 2.5
    total = np.zeros((len(X), 3))
    area = {0: L[1] * L[2], 1: L[1] * L[2], 2: L[0] * L[2], 3: L[0] * L[2], 4: L[0] * L[1], 5: L[0] * L[1]}
    for face in range(6):
        q = mpf.hex_face_quads(nx, ny, nz, face)
        f = oracle_fext(orc, X, conn, [], q, v0, 0.0).reshape(-1, 3)
        n = np.zeros(3)
        n[face // 2] = 1.0 if face % 2 else -1.0
        np.testing.assert_allclose(f.sum(0), v0 * area[face] * n, rtol=0, atol=1e-13)
        total += f
    assert np.abs(total.sum(0)).max() < 1e-13
    moment = np.cross(X, total).sum(0)
    assert np.abs(moment).max() < 1e-13


def _first_iterate_error(orc, mpf, delta, lift=True):
    """Relative distance between the first Newton iterate of one load step (from u = 0, with the
    step's prescribed increment lifted into the right-hand side when lift) and the converged u."""
    from tests.cases import apply_bcs
    n = 3
    X, conn = cube_hex(n, jitter=0.05, seed=11)
    m = mpf.matl_from_E_nu(1e3, 0.3)
    bcs = [(mpf.face_nodes(mpf.HEX8, n, n, n, 0), -1, None), (mpf.face_nodes(mpf.HEX8, n, n, n, 1), 0, delta)]
    O = orc.Problem(8, X, conn, m.mu, m.Kinv, 3)
    O.set_tangent(True)
    apply_bcs(O, bcs)
    O.prepare_matrix_pattern()
    O.set_boundary_conditions(1.0)
    vals, rhs, st = O.assemble(lift, 1.0)
    assert st == 0
    du, st = O.direct_solve(vals, rhs)
    assert st == 0
    _, _, nda, _ = O.pattern()
    d = nda.ravel()
    u1 = np.zeros(3 * len(X))
    u1[d >= 0] = du[d[d >= 0]]
    right = mpf.face_nodes(mpf.HEX8, n, n, n, 1)
    u1[3 * right] = delta  # prescribed dofs take their increment in the first update
    C = orc.Problem(8, X, conn, m.mu, m.Kinv, 3)
    C.set_tangent(True)
    apply_bcs(C, bcs)
    C.prepare_matrix_pattern()
    st, _, _ = C.newton(1, 1.0, solver=0, rtol=1e-13, itmax=30)
    assert st == 0
    uc = C.state()[0]
    return np.linalg.norm(u1 - uc) / np.linalg.norm(uc)


def test_dirichlet_lift_first_iterate_is_second_order(orc, mpf):
    """The lift makes the first Newton iterate the linearised solution: its distance to the converged
    displacement is O(delta^2) for a prescribed stretch delta, so halving delta quarters it.  Without
    the lift (or with its sign flipped) the free dofs do not move in the first iteration and the
    error stays O(1) relative, i.e. O(delta) absolute."""
    e = [_first_iterate_error(orc, mpf, d) for d in (0.02, 0.01, 0.005)]
    assert e[0] < 0.05
    r1, r2 = e[0] / e[1], e[1] / e[2]
    assert 1.8 < r1 < 2.2 and 1.8 < r2 < 2.2, e  # relative error O(delta): absolute O(delta^2)
    e_nolift = [_first_iterate_error(orc, mpf, d, lift=False) for d in (0.02, 0.01)]
    assert min(e_nolift) > 0.2 and abs(e_nolift[0] / e_nolift[1] - 1.0) < 0.1, e_nolift


# ---------------------------------------------------------------------------------------------
# Post-processing: Gauss-point Cauchy stress, von Mises, element and nodal means
# (Elem_Solid_3D_Disp.cpp:480-557; femSolidmechanics3.cpp:918-960)
# ---------------------------------------------------------------------------------------------
def _vm_from_j2(sig):
    """von Mises as sqrt(3 J2) from the deviator (another form of the reference's component formula)."""
    s = sig - np.trace(sig) / 3.0 * np.eye(3)
    return np.sqrt(1.5 * np.sum(s * s))


def _sig3(stre):
    """3x3 from the column-major 9-index map."""
    return np.asarray(stre).reshape(3, 3).T


@pytest.mark.parametrize("etype", [8, 11], ids=["hex8", "tet10"])
def test_stress_homogeneous_deformation(orc, mpf, etype):
    """u = (F0 - I) X on a jittered mesh: every Gauss point, element and node carries sigma(F0) of the
    material routine, von Mises equal to sqrt(3 J2) and tr(sigma)/3 the volumetric stress U'(J)."""
    F0 = np.array([[1.08, 0.05, -0.02], [0.03, 0.96, 0.04], [-0.01, 0.02, 1.05]])
    mu, Kinv = 50.0, 2e-3
    if etype == 8:
        X, conn = mpf.mesh_hex(3, 3, 3, jitter=0.1, seed=3)
    else:
        X, conn = mpf.mesh_tet10(2, jitter=0.1, seed=3)
    u = (X @ (F0 - np.eye(3)).T).ravel()
    P = orc.Problem(etype, X, conn, mu, Kinv, 3)
    P.prepare_matrix_pattern()
    P.set_state(disp=u)
    g = P.stress_gauss()
    s, _ = orc.neohooke(F0, mu, Kinv, 3)
    sig = _sig3(s)
    assert np.abs(g[..., :9] - s).max() < 1e-12 * np.abs(s).max()
    assert np.abs(g[..., 9] - _vm_from_j2(sig)).max() < 1e-12 * np.abs(s).max()
    J = np.linalg.det(F0)
    assert np.abs(g[..., 10] - orc.lib().orc_dUdJ(3, 1.0 / Kinv, J)).max() < 1e-10
    ev, nv = P.stress_project()
    assert np.abs(ev - g[:, 0, :]).max() < 1e-12 * np.abs(s).max()
    assert np.abs(nv - g[0, 0, :]).max() < 1e-12 * np.abs(s).max()


def test_stress_tet10_quadratic_field_closed_form(orc, mpf):
    """A quadratic displacement on affine (unjittered) Tet10 elements is reproduced exactly, so the
    Gauss-point F is I + grad u evaluated at the point's physical position; the stress there is the
    material's sigma(F).  A wrong Gauss-point order, a transposed gradient or current-configuration
    gradients fail it."""
    X, conn = mpf.mesh_tet10(2, jitter=0.0, seed=1)
    a = np.array([0.02, -0.03, 0.015])

    def ufield(x):
        return np.stack([a[0] * x[..., 0] * x[..., 1], a[1] * x[..., 2] ** 2 + 0.01 * x[..., 0],
                         a[2] * x[..., 0] * x[..., 2] - 0.02 * x[..., 1]], axis=-1)

    def grad(x):
        G = np.zeros((3, 3))
        G[0] = [a[0] * x[1], a[0] * x[0], 0.0]
        G[1] = [0.01, 0.0, 2 * a[1] * x[2]]
        G[2] = [a[2] * x[2], -0.02, a[2] * x[0]]
        return G

    mu, Kinv = 10.0, 1e-2
    P = orc.Problem(11, X, conn, mu, Kinv, 3)
    P.prepare_matrix_pattern()
    P.set_state(disp=ufield(X).ravel())
    g = P.stress_gauss()
    xi, _ = orc.gauss3d(10)
    worst = 0.0
    for e in range(len(conn)):
        V = X[conn[e, :4]]
        for gp in range(4):
            x1, x2, x3 = xi[gp]
            xg = (1 - x1 - x2 - x3) * V[0] + x1 * V[1] + x2 * V[2] + x3 * V[3]
            s, _ = orc.neohooke(np.eye(3) + grad(xg), mu, Kinv, 3)
            worst = max(worst, np.abs(g[e, gp, :9] - s).max() / np.abs(s).max())
    assert worst < 1e-12


def test_stress_hex8_trilinear_field_closed_form(orc, mpf):
    """A trilinear displacement on box-shaped (unjittered) Hex8 elements is reproduced exactly: the
    Gauss-point F is I + grad u at the point's position."""
    nx, ny, nz = 2, 2, 3
    L = (1.0, 0.8, 1.2)
    X, conn = mpf.mesh_hex(nx, ny, nz, L=L, jitter=0.0, seed=1)
    c = 0.02

    def ufield(x):
        return np.stack([c * x[..., 0] * x[..., 1] * x[..., 2], -c * x[..., 1] * x[..., 2] + 0.01 * x[..., 0],
                         0.5 * c * x[..., 0] * x[..., 1]], axis=-1)

    def grad(x):
        return np.array([[c * x[1] * x[2], c * x[0] * x[2], c * x[0] * x[1]],
                         [0.01, -c * x[2], -c * x[1]],
                         [0.5 * c * x[1], 0.5 * c * x[0], 0.0]])

    mu, Kinv = 10.0, 1e-2
    P = orc.Problem(8, X, conn, mu, Kinv, 3)
    P.prepare_matrix_pattern()
    P.set_state(disp=ufield(X).ravel())
    g = P.stress_gauss()
    xi, _ = orc.gauss3d(8)
    worst = 0.0
    for e in range(len(conn)):
        lo, hi = X[conn[e]].min(0), X[conn[e]].max(0)
        for gp in range(8):
            xg = lo + 0.5 * (np.asarray(xi[gp]) + 1.0) * (hi - lo)
            s, _ = orc.neohooke(np.eye(3) + grad(xg), mu, Kinv, 3)
            worst = max(worst, np.abs(g[e, gp, :9] - s).max() / np.abs(s).max())
    assert worst < 1e-12


def test_stress_means_over_points_and_elements(orc, mpf):
    """Element value = mean over its Gauss points; nodal value = mean of the element values of the
    elements containing the node (a node of one element takes that element's value)."""
    X, conn = mpf.mesh_hex(2, 2, 2, jitter=0.1, seed=4)
    u = np.random.default_rng(2).uniform(-0.02, 0.02, X.size)
    P = orc.Problem(8, X, conn, 10.0, 1e-2, 3)
    P.prepare_matrix_pattern()
    P.set_state(disp=u)
    g = P.stress_gauss()
    ev, nv = P.stress_project()
    assert np.allclose(ev, g.mean(axis=1), rtol=0, atol=1e-13 * np.abs(g).max())
    for n in range(len(X)):
        els = np.nonzero((conn == n).any(axis=1))[0]
        assert np.allclose(nv[n], ev[els].mean(axis=0), rtol=0, atol=1e-13 * np.abs(g).max())
    corner = int(conn[0, 0])  # node 0 belongs to element 0 only
    assert np.array_equal(nv[corner], ev[0])


# ---------------------------------------------------------------------------------------------
# Tria6 dead-load traction (the Tet10 boundary faces; BasisFunctionsLagrange.cpp:364-418,
# 1688-1732; QuadratureUtil.cpp:562-570), pinned by closed forms
# ---------------------------------------------------------------------------------------------
def _tri6_patch(P0, P1, P2):
    """One flat Tria6 face (straight edges, mid-edge nodes at the midpoints) as a 6-node mesh."""
    P = np.array([P0, P1, P2, 0.5 * (P0 + P1), 0.5 * (P1 + P2), 0.5 * (P2 + P0)], float)
    return P, np.arange(6, dtype=np.int32)[None, :]


def _tri6_fext(orc, X, faces, v0, v1):
    """The external force of one Tria6 traction patch at load factor 1 (addExternalForces alone)."""
    O = orc.Problem(11, X, np.zeros((0, 10), dtype=np.int32), 1.0, 1e-3, 3)
    O.add_traction(faces, v0, v1)
    return O.external_force(1.0)


def test_traction_tri6_flat_face_closed_form(orc):
    """Uniform v0 on a flat quadratic triangle of area A: the consistent nodal loads are 0 at the
    corners and v0 A / 3 n at the mid-edge nodes, n = (P2 - P0) x (P1 - P0) / |.|; the tangent term
    adds v1 (P1 - P0) A / 3 at the mid-edge nodes (dx/dxi1 = P1 - P0, not normalised)."""
    P0, P1, P2 = np.array([0.2, 0.1, 0.5]), np.array([1.3, 0.4, 0.2]), np.array([0.5, 1.1, 0.9])
    X, f = _tri6_patch(P0, P1, P2)
    A = 0.5 * np.linalg.norm(np.cross(P1 - P0, P2 - P0))
    n = np.cross(P2 - P0, P1 - P0)
    n /= np.linalg.norm(n)
    v0, v1 = 2.5, -0.75
    F = _tri6_fext(orc, X, f, v0, 0.0).reshape(-1, 3)
    np.testing.assert_allclose(F[:3], 0.0, atol=1e-13)
    np.testing.assert_allclose(F[3:], np.tile(v0 * A / 3 * n, (3, 1)), rtol=0, atol=1e-13)
    F = _tri6_fext(orc, X, f, 0.0, v1).reshape(-1, 3)
    np.testing.assert_allclose(F[:3], 0.0, atol=1e-13)
    np.testing.assert_allclose(F[3:], np.tile(v1 * (P1 - P0) * A / 3, (3, 1)), rtol=0, atol=1e-13)
    # reversing the corner order flips the normal
    fr = f[:, [0, 2, 1, 5, 4, 3]]
    Fr = _tri6_fext(orc, X, fr, v0, 0.0).reshape(-1, 3)
    np.testing.assert_allclose(Fr.sum(0), -v0 * A * n, rtol=0, atol=1e-13)


def test_traction_tri6_closed_surface(orc, mpf):
    """Uniform normal traction over all six boundary planes of a Tet10 block, faces oriented outward:
    zero net force and zero net moment; each plane's force is v0 times its area vector."""
    from tests.cases import tet10_boundary_tris
    X, conn = mpf.mesh_tet10(2, jitter=0.1, seed=5)
    v0 = 1.7
    total = np.zeros_like(X)
    for axis in range(3):
        for value, outward in ((0.0, -1.0), (1.0, 1.0)):
            f = tet10_boundary_tris(X, conn, axis, value, outward)
            assert len(f) == 8  # 2 x 2 squares, 2 triangles each
            F = _tri6_fext(orc, X, f, v0, 0.0).reshape(-1, 3)
            e = np.zeros(3)
            e[axis] = outward
            np.testing.assert_allclose(F.sum(0), v0 * 1.0 * e, rtol=0, atol=1e-13)
            total += F
    assert np.abs(total.sum(0)).max() < 1e-13
    assert np.abs(np.cross(X, total).sum(0)).max() < 1e-13


# ---------------------------------------------------------------------------------------------
# Arc-length continuation (femSolidmechanics2.cpp:855-1039; SolverPetsc.cpp:671-757)
# ---------------------------------------------------------------------------------------------
def arclength_beam(orc, mpf, consistent=True):
    """Cantilever 3 x 0.5 x 0.5 (6 x 2 x 2 Hex8, x = 0 clamped) under a dead shear traction on the tip
    face; E = 1e4, nu = 0.3; the load factor's scale makes the path bend visibly within 8 steps."""
    nx, ny, nz = 6, 2, 2
    X, conn = mpf.mesh_hex(nx, ny, nz, L=(3.0, 0.5, 0.5), jitter=0.05, seed=3)
    m = mpf.matl_from_E_nu(1e4, 0.3)
    P = orc.Problem(8, X, conn, m.mu, m.Kinv, 3)
    P.set_tangent(consistent)
    fixed = mpf.face_nodes(mpf.HEX8, nx, ny, nz, 0)
    quads = mpf.hex_face_quads(nx, ny, nz, 1)
    P.add_dirichlet(fixed, -1)
    P.add_traction(quads, 0.0, 400.0)
    P.prepare_matrix_pattern()
    return X, conn, m, fixed, quads, P


def test_arclength_constraint_and_equilibrium(orc, mpf):
    """Every converged step lies on the equilibrium path and on its arc-length sphere:
    ||u_k - u_(k-1)||^2 + (lam_k - lam_(k-1))^2 F.F = ds_k^2 (psi = 1); the first step is load controlled
    at lam = dt; the last state equals a plain Newton solve at the same load factor.  A dropped
    F.F factor or a wrong sign in the load update fails the sphere; a wrong residual fails Newton."""
    X, conn, m, fixed, quads, P = arclength_beam(orc, mpf)
    F = P.external_force(1.0)
    FtF = F @ F
    lam, ds, us, att = P.arclength(0.1, 1.0, 8, 30, 1e-10)
    assert len(lam) == 8 and att == 8 and lam[0] == 0.1
    assert np.all(np.diff(lam) > 0.09) and lam[-1] > 0.8
    for k in range(1, len(lam)):
        du = us[k] - us[k - 1]
        lhs = du @ du + (lam[k] - lam[k - 1]) ** 2 * FtF
        assert abs(lhs - ds[k] ** 2) < 1e-8 * ds[k] ** 2, (k, lhs, ds[k] ** 2)
    Q = orc.Problem(8, X, conn, m.mu, m.Kinv, 3)
    Q.set_tangent(True)
    Q.add_dirichlet(fixed, -1)
    Q.add_traction(quads, 0.0, 400.0 * lam[-1])
    Q.prepare_matrix_pattern()
    st, _, _ = Q.newton(8, 0.125, solver=0, rtol=1e-10, itmax=30)
    assert st == 0
    uq, _ = Q.state()
    assert normwise_rel(us[-1], uq) < 1e-9
