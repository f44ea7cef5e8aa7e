"""The CPU restatement of the learned-guess network, pinned by finite differences (float64) and by
the adjoint identity of the graph mean; the GPU kernels are checked against it in test_gpu_mpnn.py."""
from __future__ import annotations

import numpy as np
import pytest

from oracle import mpnn as O


def node_graph(conn, nnode):
    """NB(i): nodes sharing an element with i (i included), sorted -- the problem's coupled-node lists."""
    nb = [set() for _ in range(nnode)]
    for e in conn:
        for a in e:
            nb[a].update(int(b) for b in e)
    ptr = np.zeros(nnode + 1, dtype=np.int64)
    ptr[1:] = np.cumsum([len(s) for s in nb])
    return ptr, np.concatenate([sorted(s) for s in nb]).astype(np.int64)


def small_case(seed=0):
    rng = np.random.default_rng(seed)
    conn = np.array([[0, 1, 2, 3], [2, 3, 4, 5], [4, 5, 6, 7], [1, 3, 5, 7]])
    ptr, nb = node_graph(conn, 8)
    w = rng.standard_normal(O.offsets()["N"]) * 0.2
    x = rng.standard_normal((8, O.FIN))
    dy = rng.standard_normal((8, O.FOUT))
    return w, x, dy, ptr, nb


def test_mean_aggregation_adjoint():
    _, _, _, ptr, nb = small_case()
    rng = np.random.default_rng(1)
    h = rng.standard_normal((8, 5))
    d = rng.standard_normal((8, 5))
    assert np.isclose((O.mean_agg(h, ptr, nb) * d).sum(), (h * O.mean_agg_T(d, ptr, nb)).sum(), rtol=1e-13)


def test_pack_unpack_roundtrip():
    w = np.arange(O.offsets()["N"], dtype=np.float64)
    assert np.array_equal(O.pack(O.unpack(w)), w)


@pytest.mark.parametrize("seed", [0, 1])
def test_backward_matches_finite_differences(seed):
    w, x, dy, ptr, nb = small_case(seed)
    y, cache = O.forward(w, x, ptr, nb, bf16_ops=False)
    g = O.backward(w, cache, dy, ptr, nb, bf16_ops=False)
    rng = np.random.default_rng(seed + 10)
    idx = rng.choice(len(w), 60, replace=False)
    h = 1e-6
    for i in idx:
        wp, wm = w.copy(), w.copy()
        wp[i] += h
        wm[i] -= h
        fp = (O.forward(wp, x, ptr, nb, False)[0] * dy).sum()
        fm = (O.forward(wm, x, ptr, nb, False)[0] * dy).sum()
        fd = (fp - fm) / (2 * h)
        assert abs(fd - g[i]) <= 1e-6 * max(1.0, abs(fd)), (i, fd, g[i])


def test_bf16_rounding_is_nearest_even():
    vals = np.array([1.0, 1.0 + 2 ** -8, 1.0 + 2 ** -7 + 2 ** -8, -3.1415926, 0.0])
    r = O.bf16(vals)
    assert r[0] == 1.0 and r[1] == 1.0 and r[2] == 1.0 + 2 ** -6 and r[4] == 0.0
    assert abs(r[3] + 3.140625) < 1e-12
