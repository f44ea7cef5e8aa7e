"""Generate the golden fixtures under tests/golden/ from independent numpy restatements.

The reference (chennachaos/mphyfem) ships no tests and no golden vectors, so the fixtures here
come from the mathematics, computed without the oracle:
  * hex8_element.npz  -- a jittered, deformed Hex8 element: R_e from the Neo-Hookean energy
    (Cauchy stress obtained by finite differences of Psi(F) = mu/2 (J^-2/3 I1 - 3) + K/2 (J-1)^2,
    sigma = (1/J) dPsi/dF F^T), K_consistent = -dR/du by central differences of that residual,
    and K_volgeo = U'(J) int grad N_a . grad N_b delta_ik dv, the term the reference's tangent
    leaves out (utilitiesmaterialHyperelastic.cpp:280-285), so K_reference = K_consistent - K_volgeo.
  * tet10p1_element.npz -- the same for a Tet10/P1 mixed element: Fu, Fp from sigma(p) and
    Fp = -int Np (J - 1 - p/K) dV0, and the full 34x34 tangent by central differences.
  * tet10_element.npz -- a jittered, deformed Tet10 displacement element (4-point rule): R_e from
    the closed-form Cauchy stress (checked against the energy FD), K_consistent = -dR/du by central
    differences and K_volgeo, as for the Hex8 element.
  * patterns.npz -- CSR rows of tiny meshes by brute-force dense coupling (every (i, j) pair of
    free dofs sharing an element, plus the u-p couplings), for the reference's numbering.
Run:  python tests/golden/make_golden.py
"""
from __future__ import annotations

import sys
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
sys.path.insert(0, str(HERE.parent.parent))

G2 = 0.577350269189626
HEX_GP = np.array([[-1, -1, -1], [1, -1, -1], [1, 1, -1], [-1, 1, -1], [-1, -1, 1], [1, -1, 1], [1, 1, 1], [-1, 1, 1]]) * G2
HEX_NODES = np.array([[-1, -1, -1], [1, -1, -1], [1, 1, -1], [-1, 1, -1], [-1, -1, 1], [1, -1, 1], [1, 1, 1], [-1, 1, 1]], float)


def hex_dN(xi):
    """d N_a / d xi (3 x 8) for trilinear shape functions N_a = prod (1 + xi_a xi)/2."""
    s = HEX_NODES
    f = 0.5 * (1 + s * xi)  # 8 x 3
    d = np.empty((3, 8))
    d[0] = 0.5 * s[:, 0] * f[:, 1] * f[:, 2]
    d[1] = 0.5 * s[:, 1] * f[:, 0] * f[:, 2]
    d[2] = 0.5 * s[:, 2] * f[:, 0] * f[:, 1]
    return d


def psi(F, mu, K):
    J = np.linalg.det(F)
    return 0.5 * mu * (J ** (-2.0 / 3.0) * np.trace(F @ F.T) - 3.0) + 0.5 * K * (J - 1.0) ** 2


def cauchy_fd(F, mu, K, h=1e-7):
    P = np.zeros((3, 3))
    for i in range(3):
        for j in range(3):
            Fp = F.copy(); Fp[i, j] += h
            Fm = F.copy(); Fm[i, j] -= h
            P[i, j] = (psi(Fp, mu, K) - psi(Fm, mu, K)) / (2 * h)
    return P @ F.T / np.linalg.det(F)


def cauchy_closed(F, mu, K, p=None):
    J = np.linalg.det(F)
    b = F @ F.T
    dev = b - np.trace(b) / 3.0 * np.eye(3)
    vol = K * (J - 1.0) if p is None else p
    return mu * J ** (-5.0 / 3.0) * dev + vol * np.eye(3)


def hex_residual(X, u, mu, K, sigma=cauchy_closed):
    R = np.zeros((8, 3))
    x = X + u
    for xi in HEX_GP:
        d = hex_dN(xi)
        gX = np.linalg.solve(d @ X, d)       # grad_X N (3 x 8)
        F = np.eye(3) + u.T @ gX.T
        Bc = d @ x
        g = np.linalg.solve(Bc, d)           # grad_x N
        dv = np.linalg.det(Bc)
        s = sigma(F, mu, K)
        R -= dv * (g.T @ s)                  # R_a,i = -sum_m g_a,m sigma_i,m dv
    return R.ravel()


def hex_volgeo(X, u, mu, K):
    Kg = np.zeros((24, 24))
    x = X + u
    for xi in HEX_GP:
        d = hex_dN(xi)
        gX = np.linalg.solve(d @ X, d)
        F = np.eye(3) + u.T @ gX.T
        J = np.linalg.det(F)
        Bc = d @ x
        g = np.linalg.solve(Bc, d)
        dv = np.linalg.det(Bc)
        G = g.T @ g
        Kg += K * (J - 1.0) * dv * np.kron(G, np.eye(3))
    return Kg


def fd_jac(fun, z, h):
    n = z.size
    r0 = fun(z)
    Jm = np.zeros((r0.size, n))
    for j in range(n):
        zp = z.copy(); zp[j] += h
        zm = z.copy(); zm[j] -= h
        Jm[:, j] = -(fun(zp) - fun(zm)) / (2 * h)
    return Jm


TET_GP = np.array([[0.1381966011250105] * 3] * 4)
TET_GP[1, 0] = TET_GP[2, 1] = TET_GP[3, 2] = 0.5854101966249685
TET_W = 0.25 / 6.0
TET_EDGES = [(0, 1), (1, 2), (0, 2), (0, 3), (2, 3), (1, 3)]


def tet10_N_dN(xi):
    l = np.array([1 - xi.sum(), xi[0], xi[1], xi[2]])
    dl = np.array([[-1, -1, -1], [1, 0, 0], [0, 1, 0], [0, 0, 1]], float)  # d l_v / d xi
    N = np.empty(10)
    dN = np.empty((3, 10))
    for v in range(4):
        N[v] = l[v] * (2 * l[v] - 1)
        dN[:, v] = (4 * l[v] - 1) * dl[v]
    for q, (a, b) in enumerate(TET_EDGES):
        N[4 + q] = 4 * l[a] * l[b]
        dN[:, 4 + q] = 4 * (dl[a] * l[b] + l[a] * dl[b])
    return N, dN, l


def tet10_residual(X, u, mu, K, sigma=cauchy_closed):
    """Tet10 displacement element: R_a,i = -sum_gp w det(dx/dxi) sum_m g_a,m sigma_i,m."""
    R = np.zeros((10, 3))
    x = X + u
    for xi in TET_GP:
        _, d, _ = tet10_N_dN(xi)
        gX = np.linalg.solve(d @ X, d)
        F = np.eye(3) + u.T @ gX.T
        Bc = d @ x
        g = np.linalg.solve(Bc, d)
        dv = TET_W * np.linalg.det(Bc)
        R -= dv * (g.T @ sigma(F, mu, K))
    return R.ravel()


def tet10_volgeo(X, u, mu, K):
    Kg = np.zeros((30, 30))
    x = X + u
    for xi in TET_GP:
        _, d, _ = tet10_N_dN(xi)
        gX = np.linalg.solve(d @ X, d)
        J = np.linalg.det(np.eye(3) + u.T @ gX.T)
        Bc = d @ x
        g = np.linalg.solve(Bc, d)
        dv = TET_W * np.linalg.det(Bc)
        Kg += K * (J - 1.0) * dv * np.kron(g.T @ g, np.eye(3))
    return Kg


def tet10p1_residual(X, u, up, p, mu, K):
    Fu = np.zeros((10, 3))
    Fp = np.zeros(4)
    x = X + u
    for xi in TET_GP:
        _, d, Np = tet10_N_dN(xi)
        B0 = d @ X
        gX = np.linalg.solve(B0, d)
        dv0 = TET_W * np.linalg.det(B0)
        F = np.eye(3) + u.T @ gX.T
        Fn = np.eye(3) + up.T @ gX.T
        Jn = np.linalg.det(Fn)
        Jhat = Jn - K * (Jn - 1.0) / K   # Utype 3
        theta = 1.0 / K
        J = np.linalg.det(F)
        Bc = d @ x
        g = np.linalg.solve(Bc, d)
        dv = TET_W * np.linalg.det(Bc)
        pres = Np @ p
        s = cauchy_closed(F, mu, K, p=pres)
        Fu -= dv * (g.T @ s)
        Fp -= dv0 * Np * (J - Jhat - theta * pres)
    return np.concatenate([Fu.ravel(), Fp])


def brute_pattern(nNode, conn, fixed_dofs, mixed):
    """Reference numbering + dense coupling -> (rowptr, colidx, nda, idp)."""
    fixed = np.zeros(nNode * 3, bool)
    fixed[list(fixed_dofs)] = True
    nda = -np.ones(nNode * 3, int)
    nda[~fixed] = np.arange((~fixed).sum())
    nd = int((~fixed).sum())
    idp = -np.ones(nNode, int)
    if mixed:
        pn = np.unique(conn[:, :4])
        idp[pn] = np.arange(len(pn))
    ntot = nd + int((idp >= 0).sum())
    A = np.zeros((ntot, ntot), bool)
    for e in conn:
        fav = nda[(3 * e[:, None] + np.arange(3)).ravel()]
        fav = fav[fav >= 0]
        A[np.ix_(fav, fav)] = True
        if mixed:
            pv = idp[e[:4]] + nd
            A[np.ix_(pv, fav)] = True
            A[np.ix_(fav, pv)] = True
            A[np.ix_(pv, pv)] = True
    rowptr = np.concatenate([[0], np.cumsum(A.sum(1))]).astype(np.int32)
    colidx = np.concatenate([np.flatnonzero(r) for r in A]).astype(np.int32)
    return rowptr, colidx, nda.reshape(nNode, 3).astype(np.int32), idp.astype(np.int32)


def hex_grid(nx, ny, nz):
    X = np.array([[i, j, k] for k in range(nz + 1) for j in range(ny + 1) for i in range(nx + 1)], float)
    X /= np.array([nx, ny, nz], float)
    nid = lambda i, j, k: i + (nx + 1) * (j + (ny + 1) * k)
    conn = np.array([[nid(i, j, k), nid(i + 1, j, k), nid(i + 1, j + 1, k), nid(i, j + 1, k),
                      nid(i, j, k + 1), nid(i + 1, j, k + 1), nid(i + 1, j + 1, k + 1), nid(i, j + 1, k + 1)]
                     for k in range(nz) for j in range(ny) for i in range(nx)], np.int32)
    return X, conn


def main():
    rng = np.random.default_rng(20261202)
    mu, K = 384.6153846153846, 1.0 / 1.2e-3
    # ---- Hex8 element ----
    X = (HEX_NODES + 1) * 0.5 * np.array([0.1, 0.12, 0.09]) + rng.uniform(-0.004, 0.004, (8, 3))
    u = rng.uniform(-0.006, 0.006, (8, 3))
    R = hex_residual(X, u, mu, K)
    R_energy = hex_residual(X, u, mu, K, sigma=cauchy_fd)
    assert np.abs(R - R_energy).max() < 1e-6 * np.abs(R).max(), "closed-form stress disagrees with energy FD"
    Kc = fd_jac(lambda z: hex_residual(X, z.reshape(8, 3), mu, K), u.ravel(), 1e-7)
    Kg = hex_volgeo(X, u, mu, K)
    np.savez(HERE / "hex8_element.npz", X=X, u=u, mu=mu, Kinv=1.2e-3, Re=R, Ke_consistent=Kc, Ke_volgeo=Kg)
    # ---- Tet10/P1 element ----
    V = np.array([[0, 0, 0], [1, 0, 0], [0, 1, 0], [0, 0, 1]], float) * 0.2
    Xt = np.vstack([V, [(V[a] + V[b]) / 2 for a, b in TET_EDGES]]) + rng.uniform(-0.004, 0.004, (10, 3))
    ut = rng.uniform(-0.005, 0.005, (10, 3))
    upt = 0.5 * ut
    pt = rng.uniform(-2.0, 2.0, 4)
    muT, KT = 1.0, 1.0e3
    R34 = tet10p1_residual(Xt, ut, upt, pt, muT, KT)
    z = np.concatenate([ut.ravel(), pt])
    K34 = fd_jac(lambda zz: tet10p1_residual(Xt, zz[:30].reshape(10, 3), upt, zz[30:], muT, KT), z, 1e-7)
    np.savez(HERE / "tet10p1_element.npz", X=Xt, u=ut, uprev=upt, pres=pt, mu=muT, Kinv=1.0 / KT, R=R34, K=K34)
    # ---- Tet10 displacement element (own generator: the fixtures above keep their draws) ----
    rng2 = np.random.default_rng(20261203)
    V = np.array([[0, 0, 0], [1, 0, 0], [0, 1, 0], [0, 0, 1]], float) * 0.25
    Xd = np.vstack([V, [(V[a] + V[b]) / 2 for a, b in TET_EDGES]]) + rng2.uniform(-0.005, 0.005, (10, 3))
    ud = rng2.uniform(-0.007, 0.007, (10, 3))
    muD, KD = 384.6153846153846, 1.0 / 1.2e-3
    Rd = tet10_residual(Xd, ud, muD, KD)
    Rd_energy = tet10_residual(Xd, ud, muD, KD, sigma=cauchy_fd)
    assert np.abs(Rd - Rd_energy).max() < 1e-6 * np.abs(Rd).max(), "closed-form stress disagrees with energy FD"
    Kd = fd_jac(lambda z: tet10_residual(Xd, z.reshape(10, 3), muD, KD), ud.ravel(), 1e-7)
    Kgd = tet10_volgeo(Xd, ud, muD, KD)
    np.savez(HERE / "tet10_element.npz", X=Xd, u=ud, mu=muD, Kinv=1.2e-3, Re=Rd, Ke_consistent=Kd, Ke_volgeo=Kgd)
    # ---- patterns ----
    out = {}
    for (nx, ny, nz) in [(1, 1, 1), (2, 2, 2), (3, 2, 1)]:
        Xh, ch = hex_grid(nx, ny, nz)
        x0 = [n for n in range(len(Xh)) if Xh[n, 0] == 0.0]
        x1 = [n for n in range(len(Xh)) if Xh[n, 0] == 1.0]
        fixed = {3 * n + d for n in x0 for d in range(3)} | {3 * n for n in x1}
        rp, ci, nda, idp = brute_pattern(len(Xh), ch, fixed, False)
        key = f"hex_{nx}{ny}{nz}"
        out[key + "_X"], out[key + "_conn"] = Xh, ch
        out[key + "_rowptr"], out[key + "_colidx"], out[key + "_nda"] = rp, ci, nda
    np.savez(HERE / "patterns.npz", **out)
    print("wrote", sorted(p.name for p in HERE.glob("*.npz")))


if __name__ == "__main__":
    main()
