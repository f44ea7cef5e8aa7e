"""The SolverPetsc-shaped system over a caller-supplied pattern (csrc/system.hip) through the C ABI:
element matrices summed into CSR in element order, as PETSc's ADD_VALUES loop does one element after
another (SolverPetsc.cpp:357-378, 434-502), so the device result equals a sequential host
accumulation (numpy's unbuffered add.at, which applies the contributions in index order) bit for
bit; -1 rows and columns skipped; an entry outside the pattern refused; the solve on the result."""
from __future__ import annotations

import numpy as np
import pytest

from tests.cases import normwise_rel

pytestmark = pytest.mark.gpu


def _pattern(mpf, n=3):
    X, conn = mpf.mesh_hex(n, n, n, jitter=0.05, seed=5)
    P = mpf.Problem(mpf.HEX8, X, conn, mpf.matl_from_E_nu(1e3, 0.3))
    P.add_dirichlet(mpf.face_nodes(mpf.HEX8, n, n, n, 0), -1)
    P.add_dirichlet(mpf.face_nodes(mpf.HEX8, n, n, n, 1), 0, 0.1)
    P.prepare_matrix_pattern()
    pat = P.pattern()
    d, p, nnz, _ = P.sizes()
    return X, conn, pat, d, p, nnz


def _host_scatter(rowptr, colidx, rows, cols, K, F, n):
    """Sequential ADD_VALUES in element order (the reference's loop), skipping -1."""
    vals = np.zeros(len(colidx))
    rhs = np.zeros(n)
    slot_idx, slot_val, r_idx, r_val = [], [], [], []
    for e in range(rows.shape[0]):
        for i, r in enumerate(rows[e]):
            if r < 0:
                continue
            r_idx.append(r)
            r_val.append(F[e, i])
            for j, c in enumerate(cols[e]):
                if c < 0:
                    continue
                k = rowptr[r] + np.searchsorted(colidx[rowptr[r]:rowptr[r + 1]], c)
                assert colidx[k] == c
                slot_idx.append(k)
                slot_val.append(K[e, i, j])
    np.add.at(vals, np.array(slot_idx, dtype=np.int64), np.array(slot_val))
    np.add.at(rhs, np.array(r_idx, dtype=np.int64), np.array(r_val))
    return vals, rhs


def test_system_scatter_equals_sequential_add_values(mpf, gpu):
    import torch
    X, conn, pat, d, p, nnz = _pattern(mpf)
    rows = pat.NodeDofArray[conn].reshape(len(conn), 24).astype(np.int32)
    rng = np.random.default_rng(3)
    K = rng.standard_normal((len(conn), 24, 24))
    F = rng.standard_normal((len(conn), 24))
    S = mpf.System(d, p, pat.rowptr, pat.colidx, pat.NodeDofArray)
    dev = lambda a: torch.from_numpy(np.ascontiguousarray(a)).to(gpu)
    rt = dev(rows)
    mid = S.map_elements(rt, rt)
    S.add_elements(mid, dev(K), dev(F))
    vals, rhs = S.system()
    hv, hr = _host_scatter(pat.rowptr, pat.colidx, rows, rows, K, F, d + p)
    assert np.array_equal(vals, hv) and np.array_equal(rhs, hr)
    # zeroMtx, then the same map again: the same bits
    S.zero_mtx()
    S.add_elements(mid, dev(K), dev(F))
    v2, r2 = S.system()
    assert np.array_equal(v2, hv) and np.array_equal(r2, hr)
    # a second add without zeroing accumulates onto the first (ADD_VALUES)
    S.add_elements(mid, dev(K), None)
    v3, r3 = S.system()
    v_seq, _ = _host_scatter(pat.rowptr, pat.colidx, np.concatenate([rows, rows]), np.concatenate([rows, rows]),
                             np.concatenate([K, K]), np.concatenate([F, F]), d + p)
    assert np.array_equal(v3, v_seq) and np.array_equal(r3, hr)
    S.close()


def test_system_offsets_skips_and_pattern_check(mpf, gpu):
    import torch
    X, conn, pat, d, p, nnz = _pattern(mpf)
    S = mpf.System(d, p, pat.rowptr, pat.colidx, pat.NodeDofArray)
    dev = lambda a: torch.from_numpy(np.ascontiguousarray(a)).to(gpu)
    # rows given relative to an offset: -1 stays skipped, others move by r_off
    rows = pat.NodeDofArray[conn].reshape(len(conn), 24).astype(np.int32)
    shifted = np.where(rows >= 0, rows - 5, -1).astype(np.int32)
    shifted[shifted < -1] = -1  # dofs 0..4 dropped from this call
    rng = np.random.default_rng(4)
    K = rng.standard_normal((len(conn), 24, 24))
    F = rng.standard_normal((len(conn), 24))
    st = dev(shifted)
    mid = S.map_elements(st, st, 5, 5)
    S.add_elements(mid, dev(K), dev(F))
    vals, rhs = S.system()
    back = np.where(shifted >= 0, shifted + 5, -1)
    hv, hr = _host_scatter(pat.rowptr, pat.colidx, back, back, K, F, d + p)
    assert np.array_equal(vals, hv) and np.array_equal(rhs, hr)
    # a column outside the pattern (first and last free dofs of far-apart nodes) is refused
    bad = np.array([[0, d - 1]], dtype=np.int32)
    with pytest.raises(mpf.MpfError):
        S.map_elements(dev(bad), dev(bad))
    S.close()


def test_system_solve_matches_problem_solve(mpf, gpu):
    """The caller-pattern system filled from the problem's own assembled CSR values (one 'element' per
    row) solves to the same answer as the mesh problem's solve, and BJACOBI3 without a NodeDofArray
    is refused."""
    import torch
    n = 4
    X, conn = mpf.mesh_hex(n, n, n, jitter=0.05, seed=6)
    P = mpf.Problem(mpf.HEX8, X, conn, mpf.matl_from_E_nu(1e3, 0.3))
    P.add_dirichlet(mpf.face_nodes(mpf.HEX8, n, n, n, 0), -1)
    P.add_dirichlet(mpf.face_nodes(mpf.HEX8, n, n, n, 1), 0, 0.1)
    P.prepare_matrix_pattern()
    P.set_boundary_conditions(0.5)
    P.calc_stiffness_and_residual(True, 0.5)
    vals, rhs = P.system()
    pat = P.pattern()
    d, p, nnz, _ = P.sizes()
    P.factorise_and_solve(pc=mpf.PC_BJACOBI3, rtol=1e-13)
    x_ref = P.solution()
    # each CSR row as a 1 x rowlen "element"
    rp = pat.rowptr
    L = int(np.diff(rp).max())
    nr = d + p
    rows = np.arange(nr, dtype=np.int32)[:, None]
    cols = -np.ones((nr, L), dtype=np.int32)
    Kr = np.zeros((nr, 1, L))
    for r in range(nr):
        m = rp[r + 1] - rp[r]
        cols[r, :m] = pat.colidx[rp[r]:rp[r + 1]]
        Kr[r, 0, :m] = vals[rp[r]:rp[r + 1]]
    dev = lambda a: torch.from_numpy(np.ascontiguousarray(a)).to(gpu)
    S = mpf.System(d, p, pat.rowptr, pat.colidx, pat.NodeDofArray)
    mid = S.map_elements(dev(rows), dev(cols))
    S.add_elements(mid, dev(Kr), dev(rhs[:, None]))
    v2, r2 = S.system()
    assert np.array_equal(v2, vals) and np.array_equal(r2, rhs)
    rep = S.factorise_and_solve(pc=mpf.PC_BJACOBI3, rtol=1e-13)
    assert rep.converged
    assert normwise_rel(S.solution(), x_ref) < 1e-9
    S.close()
    S2 = mpf.System(d, p, pat.rowptr, pat.colidx)
    mid = S2.map_elements(dev(rows), dev(cols))
    S2.add_elements(mid, dev(Kr), dev(rhs[:, None]))
    with pytest.raises(mpf.MpfError):
        S2.factorise_and_solve(pc=mpf.PC_BJACOBI3)
    rep = S2.factorise_and_solve(pc=mpf.PC_JACOBI, rtol=1e-13)
    assert rep.converged and normwise_rel(S2.solution(), x_ref) < 1e-9
    S2.close()
