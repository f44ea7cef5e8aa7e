"""CPU restatement of the learned-guess network (mphyfem_amd/csrc/mpnn.hip), numpy.

TEST INFRASTRUCTURE ONLY: tests/ compare the GPU forward and backward kernels against it;
the product never imports it.  No counterpart exists in the reference (the learned initial guess
has none).

`bf16=True` reproduces the GPU's arithmetic contract: every GEMM operand rounded to bfloat16
(round to nearest even), products accumulated in float64 here (float32 on the GPU), the residual
stream kept unrounded.  `bf16=False` is the plain float64 network, pinned by finite differences.
"""
from __future__ import annotations

import numpy as np

FIN, W, LAYERS, FOUT = 32, 64, 3, 32


def offsets():
    """Offsets of the flat parameter vector (the GPU's MpnnParams layout)."""
    o = {"W0": 0}
    o["B0"] = o["W0"] + FIN * W
    o["WL"] = o["B0"] + W
    o["LSZ"] = 2 * W * W + W
    o["WO"] = o["WL"] + LAYERS * o["LSZ"]
    o["BO"] = o["WO"] + W * FOUT
    o["N"] = o["BO"] + FOUT
    return o


def unpack(w):
    o = offsets()
    w = np.asarray(w, dtype=np.float64)
    P = {"W0": w[o["W0"]:o["B0"]].reshape(FIN, W), "b0": w[o["B0"]:o["WL"]]}
    for l in range(LAYERS):
        base = o["WL"] + l * o["LSZ"]
        P[f"W{l + 1}"] = w[base:base + 2 * W * W].reshape(2 * W, W)
        P[f"b{l + 1}"] = w[base + 2 * W * W:base + o["LSZ"]]
    P["Wo"] = w[o["WO"]:o["BO"]].reshape(W, FOUT)
    P["bo"] = w[o["BO"]:o["N"]]
    return P


def pack(G):
    o = offsets()
    g = np.zeros(o["N"])
    g[o["W0"]:o["B0"]] = G["W0"].ravel()
    g[o["B0"]:o["WL"]] = G["b0"]
    for l in range(LAYERS):
        base = o["WL"] + l * o["LSZ"]
        g[base:base + 2 * W * W] = G[f"W{l + 1}"].ravel()
        g[base + 2 * W * W:base + o["LSZ"]] = G[f"b{l + 1}"]
    g[o["WO"]:o["BO"]] = G["Wo"].ravel()
    g[o["BO"]:o["N"]] = G["bo"]
    return g


def bf16(a):
    """Round to bfloat16 (nearest even), returned as float64."""
    f = np.asarray(a, dtype=np.float32)
    u = f.view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) >> 16
    return (u.astype(np.uint32) << 16).view(np.float32).astype(np.float64)


def mean_agg(h, nbptr, nb):
    """a_i = mean over NB(i) of h_j."""
    deg = np.diff(nbptr)
    rows = np.repeat(np.arange(len(deg)), deg)
    out = np.zeros_like(h)
    np.add.at(out, rows, h[nb])
    return out / np.maximum(deg, 1)[:, None]


def mean_agg_T(d, nbptr, nb):
    """Adjoint of mean_agg: g_j = sum over i with j in NB(i) of d_i / |NB(i)|."""
    deg = np.diff(nbptr)
    rows = np.repeat(np.arange(len(deg)), deg)
    out = np.zeros_like(d)
    np.add.at(out, nb, (d / np.maximum(deg, 1)[:, None])[rows])
    return out


def forward(w, x, nbptr, nb, bf16_ops=True):
    """Returns (y, cache)."""
    R = bf16 if bf16_ops else (lambda a: np.asarray(a, dtype=np.float64))
    P = unpack(w)
    xb = R(x)
    z0 = xb @ R(P["W0"]) + P["b0"]
    h = np.maximum(z0, 0.0)
    cache = {"x": xb, "z": [z0], "cat": []}
    for l in range(1, LAYERS + 1):
        hb = R(h)
        a = R(mean_agg(hb, nbptr, nb)) if bf16_ops else mean_agg(hb, nbptr, nb)
        cat = np.concatenate([hb, a], axis=1)
        z = cat @ R(P[f"W{l}"]) + P[f"b{l}"]
        cache["cat"].append(cat)
        cache["z"].append(z)
        h = h + np.maximum(z, 0.0)
    hb = R(h)
    cache["hL"] = hb
    y = hb @ R(P["Wo"]) + P["bo"]
    return y, cache


def backward(w, cache, dy, nbptr, nb, bf16_ops=True):
    """Gradient of sum(dy * y) w.r.t. the flat parameters (the GPU's rounding points when bf16_ops)."""
    R = bf16 if bf16_ops else (lambda a: np.asarray(a, dtype=np.float64))
    P = unpack(w)
    G = {}
    dyb = R(dy)
    G["Wo"] = cache["hL"].T @ dyb
    G["bo"] = dy.sum(axis=0)
    dh = dyb @ R(P["Wo"]).T
    for l in range(LAYERS, 0, -1):
        dr = dh * (cache["z"][l] > 0)
        drb = R(dr)
        G[f"W{l}"] = cache["cat"][l - 1].T @ drb
        G[f"b{l}"] = dr.sum(axis=0)
        dcat = drb @ R(P[f"W{l}"]).T
        dh = dh + dcat[:, :W] + mean_agg_T(dcat[:, W:], nbptr, nb)
    dz0 = dh * (cache["z"][0] > 0)
    dz0b = R(dz0)
    G["W0"] = cache["x"].T @ dz0b
    G["b0"] = dz0.sum(axis=0)
    return pack(G)
