"""Shared problem builders for the parity tests (same seeded inputs for oracle and device)."""
from __future__ import annotations

import numpy as np

import mphyfem_amd as M


def cube_hex(n: int, jitter: float = 0.05, seed: int = 1):
    """Unit cube of n^3 Hex8 elements (the structured block of the CPU-runnable workload)."""
    return M.mesh_hex(n, n, n, jitter=jitter, seed=seed)


def hex_stretch_bcs(n: int):
    """x=0 face fixed, x=1 face given UX = 0.1 * lambda."""
    return [(M.face_nodes(M.HEX8, n, n, n, 0), -1, None), (M.face_nodes(M.HEX8, n, n, n, 1), 0, 0.1)]


def tet_compress_bcs(n: int, load: float = -0.05):
    """z=0 face fixed, z=1 face given UZ = load * lambda."""
    return [(M.face_nodes(M.TET10P1, n, n, n, 4), -1, None), (M.face_nodes(M.TET10P1, n, n, n, 5), 2, load)]


def apply_bcs(prob, bcs):
    for nodes, dof, val in bcs:
        prob.add_dirichlet(nodes, dof, val)


def random_state(nNode: int, scale: float, seed: int):
    rng = np.random.default_rng(seed)
    return rng.uniform(-scale, scale, nNode * 3)


def normwise_rel(a, b) -> float:
    """max |a - b| / max |b| (normwise relative error; entrywise is meaningless near zero entries)."""
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    den = np.abs(b).max()
    return float(np.abs(a - b).max() / (den if den > 0 else 1.0))
