"""The learned-guess network's BF16 MFMA kernels (forward and backward) against the CPU oracle
network with the same BF16 rounding points, and the guess path against its closed form."""
from __future__ import annotations

import numpy as np
import pytest

from oracle import mpnn as O
from tests.test_oracle_mpnn import node_graph

pytestmark = pytest.mark.gpu
# BF16 operands rounded at the same points on both sides and FP32 accumulation on the GPU (FP64 in
# the oracle): the outputs agree to ~1e-7 of their largest entry (every entry within TOL, 99% within
# BULK); the weight gradients, sums over every row in FP32, within TOL_GRAD.  A dropped term moves
# the outputs by ~1e-2 (a planted mistake in the oracle below), far outside these bars.
TOL = 1e-5
BULK = 1e-6
TOL_GRAD = 1e-5  # weight gradients: sums over all rows in FP32 (measured ~2e-7)


def _problem(mpf):
    B = mpf.mesh_batch_hex([2, 3])
    P = mpf.Problem(mpf.HEX8, B.X, B.conn, mpf.matl_from_E_nu(1e3, 0.3, tangent=mpf.TANGENT_CONSISTENT))
    P.set_batch(B.node_off, B.elem_off, [300.0, 400.0], [1e-3, 2e-3])
    for m in range(2):
        P.add_dirichlet(B.bottom[m], -1)
    P.prepare_matrix_pattern()
    return B, P


def _rel(a, b):
    return np.abs(a - b).max() / max(np.abs(b).max(), 1e-30)


def _bulk(a, b, q=99.0):
    """The q-th percentile of |a - b| / max |b|."""
    return float(np.percentile(np.abs(a - b), q) / max(np.abs(b).max(), 1e-30))


def test_mpnn_forward_backward_vs_oracle(mpf, gpu):
    import torch
    B, P = _problem(mpf)
    n = len(B.X)
    ptr, nb = node_graph(B.conn, n)
    net = mpf.Mpnn(seed=7)
    w, _ = net.params()
    rng = np.random.default_rng(3)
    # nonzero biases (the initialisation sets them to 0), so that every term of every layer is live
    o = O.offsets()
    bias = [slice(o["B0"], o["WL"]), slice(o["BO"], o["N"])]
    bias += [slice(o["WL"] + l * o["LSZ"] + 2 * O.W * O.W, o["WL"] + (l + 1) * o["LSZ"]) for l in range(O.LAYERS)]
    for sl in bias:
        w[sl] = (0.1 * rng.standard_normal(sl.stop - sl.start)).astype(np.float32)
    net.set_params(w)
    x = O.bf16(rng.standard_normal((n, O.FIN)))
    xb = (x.astype(np.float32).view(np.uint32) >> 16).astype(np.uint16)
    xt = torch.from_numpy(xb.view(np.int16)).to(gpu)
    yt = torch.zeros((n, O.FOUT), dtype=torch.float32, device=gpu)
    net.forward(P, xt, yt)
    y_ref, cache = O.forward(w.astype(np.float64), x, ptr, nb, bf16_ops=True)
    y = yt.cpu().numpy()
    print(f"[mpnn] forward: max {_rel(y, y_ref):.2e}, 99% {_bulk(y, y_ref):.2e}", flush=True)
    assert _rel(y, y_ref) < TOL and _bulk(y, y_ref) < BULK
    # the same metric on the oracle with a planted mistake (the first layer's bias dropped): far
    # outside the bars, so a missing term cannot hide under them
    w_bad = w.astype(np.float64).copy()
    w_bad[o["B0"]:o["WL"]] = 0.0
    y_bad, _ = O.forward(w_bad, x, ptr, nb, bf16_ops=True)
    assert _bulk(y, y_bad) > 10 * BULK and _rel(y, y_bad) > TOL
    dy = rng.standard_normal((n, O.FOUT)).astype(np.float32)
    net.zero_grad()
    net.backward(P, torch.from_numpy(dy).to(gpu))
    _, g = net.params()
    g_ref = O.backward(w.astype(np.float64), cache, dy.astype(np.float64), ptr, nb, bf16_ops=True)
    Gg, Gr = O.unpack(g), O.unpack(g_ref)
    for k in Gr:
        print(f"[mpnn] grad {k}: max {_rel(Gg[k], Gr[k]):.2e}, 99% {_bulk(Gg[k], Gr[k]):.2e}", flush=True)
        assert _rel(Gg[k], Gr[k]) < TOL_GRAD, k


def test_mpnn_adam_moves_loss_down(mpf, gpu):
    import torch
    B, P = _problem(mpf)
    n = len(B.X)
    rng = np.random.default_rng(5)
    uk = torch.from_numpy(rng.standard_normal(3 * n) * 1e-3).to(gpu)
    ukm1 = uk * 0.5
    # next state = linear extrapolation + a smooth field of the position (learnable from the features)
    unext = uk * 1.5 + torch.from_numpy(1e-3 * B.X.ravel() ** 2).to(gpu)
    net = mpf.Mpnn(seed=9)
    net.features(P, uk, ukm1, 1.0, 1.0, 1.0)
    losses = []
    for _ in range(30):
        net.forward(P)
        losses.append(net.loss(P, unext, uk, ukm1, 1.0))
        net.backward(P)
        net.adam(lr=3e-3)
    assert losses[-1] < 0.5 * losses[0], losses[::5]


def test_guess_without_network_is_linear_extrapolation(mpf, gpu):
    import torch
    B, P = _problem(mpf)
    n = len(B.X)
    rng = np.random.default_rng(6)
    uk = torch.from_numpy(rng.standard_normal(3 * n)).to(gpu)
    ukm1 = torch.from_numpy(rng.standard_normal(3 * n)).to(gpu)
    net = mpf.Mpnn(seed=1)
    net.features(P, uk, ukm1, 0.5, 1.0, 1.0)
    out = torch.zeros(3 * n, dtype=torch.float64, device=gpu)
    net.guess(P, uk, ukm1, 0.5, False, out)
    ref = uk + 0.5 * (uk - ukm1)
    assert torch.allclose(out, ref, rtol=1e-15, atol=1e-15)
