"""CPU check of the algebra behind the banded direct solver's cut blocks (band.hip): a block of n
rows and half-bandwidth b eliminated in the order [A1, A2 reversed, S], with the separator S of
b <= s < b + 16 rows in the middle chosen by the same rule as the solver's layout (A1 and A2 whole
16-column panels), gives the solution of the whole system.  The two segments are factored as plain
bands up to S, segment 1's partial Schur complement is added into segment 0's separator triangle
(rows reversed, same band offsets), S is factored at the end of segment 0, and the solve runs
forward to S in both, S forward and back, back from S in both -- the steps of k_band_factor,
k_band_schur, k_band_sep and k_band_solve, in band storage entry (r, c) at c (b + 1) + (r - c)."""
from __future__ import annotations

import numpy as np
import pytest

KW = 16


def _spd_band(rng, n, b):
    A = np.zeros((n, n))
    for d in range(1, b + 1):
        v = rng.standard_normal(n - d) * 0.3 * (rng.random(n - d) < 0.7)
        A += np.diag(v, -d) + np.diag(v, d)
    return A + np.eye(n) * (2 * b + 5)


def _segments(n, b):
    a1 = KW * ((n - b) // (2 * KW))
    a2 = KW * ((n - a1 - b) // KW)
    s = n - a1 - a2
    seg0 = dict(base=0, dir=1, n=a1 + s, nf=a1, nend=a1 + s, zc=a1 + s)
    seg1 = dict(base=n - 1, dir=-1, n=a2 + s, nf=a2, nend=a2, zc=a2)
    return seg0, seg1, s


def _fill(K, g, b):
    B1 = b + 1
    L = np.zeros(g["n"] * B1)
    for k in range(min(g["zc"], g["n"])):
        p = g["base"] + g["dir"] * k
        for pp in np.nonzero(K[p])[0]:
            i = g["dir"] * (pp - g["base"])
            if k <= i < g["n"] and i - k <= b:
                L[k * B1 + i - k] = K[p, pp]
    return L


def _factor(L, n, b, c0, c1):
    B1 = b + 1
    for c in range(c0, c1):
        d = L[c * B1]
        rows = range(c + 1, min(n, c + b + 1))
        for r in rows:
            L[c * B1 + r - c] /= d
        for cc in rows:
            lcc = L[c * B1 + cc - c]
            for r in range(cc, min(n, c + b + 1)):
                L[cc * B1 + r - cc] -= L[c * B1 + r - c] * d * lcc


def _forward(L, y, n, b, c0, c1):
    B1 = b + 1
    for c in range(c0, c1):
        for r in range(c + 1, min(n, c + b + 1)):
            y[r] -= L[c * B1 + r - c] * y[c]
    for c in range(c0, c1):
        y[c] /= L[c * B1]


def _backward(L, y, n, b, c0, c1):
    B1 = b + 1
    for c in range(c1 - 1, c0 - 1, -1):
        for r in range(c + 1, min(n, c + b + 1)):
            y[c] -= L[c * B1 + r - c] * y[r]


@pytest.mark.parametrize("n,b", [(3 * 20 + 64, 20), (517, 37), (700, 47)])
def test_cut_block_solves_the_whole_system(n, b):
    rng = np.random.default_rng(n + b)
    K = _spd_band(rng, n, b)
    rhs = rng.standard_normal(n)
    g0, g1, s = _segments(n, b)
    assert b <= s < b + KW and g0["nf"] % KW == 0 and g1["nf"] % KW == 0 and g0["nf"] > 0 and g1["nf"] > 0
    B1 = b + 1
    L0, L1 = _fill(K, g0, b), _fill(K, g1, b)
    _factor(L0, g0["n"], b, 0, g0["nf"])
    _factor(L1, g1["n"], b, 0, g1["nf"])
    for i in range(s):  # k_band_schur
        for d in range(min(b, i) + 1):
            L0[(g0["nf"] + i - d) * B1 + d] += L1[(g1["n"] - 1 - i) * B1 + d]
    _factor(L0, g0["n"], b, g0["nf"], g0["nend"])
    y0 = np.array([rhs[k] for k in range(g0["n"])])
    y1 = np.array([rhs[n - 1 - k] if k < g1["zc"] else 0.0 for k in range(g1["n"])])
    _forward(L0, y0, g0["n"], b, 0, g0["nf"])
    _forward(L1, y1, g1["n"], b, 0, g1["nf"])
    y0[g0["nf"]:] += y1[::-1][:s]  # k_band_sep, copy 0
    _forward(L0, y0, g0["n"], b, g0["nf"], g0["nend"])
    _backward(L0, y0, g0["n"], b, g0["nf"], g0["nend"])
    y1[g1["nf"]:] = y0[g0["nf"]:][::-1]  # k_band_sep, copy 1
    _backward(L0, y0, g0["n"], b, 0, g0["nf"])
    _backward(L1, y1, g1["n"], b, 0, g1["nf"])
    x = np.empty(n)
    x[: g0["zc"]] = y0[: g0["zc"]]
    x[n - g1["zc"]:] = y1[: g1["zc"]][::-1]
    x_ref = np.linalg.solve(K, rhs)
    assert np.abs(x - x_ref).max() <= 1e-12 * np.abs(x_ref).max()
    # a dropped separator term (no Schur sum) must not pass
    L0b = _fill(K, g0, b)
    _factor(L0b, g0["n"], b, 0, g0["nend"])
    yb = rhs[: g0["n"]].copy()
    _forward(L0b, yb, g0["n"], b, 0, g0["nend"])
    _backward(L0b, yb, g0["n"], b, 0, g0["nend"])
    assert np.abs(yb[: g0["nf"]] - x_ref[: g0["nf"]]).max() > 1e-6 * np.abs(x_ref).max()
