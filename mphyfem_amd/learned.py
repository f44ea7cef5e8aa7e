"""Learned Newton initial guess: cantilever workload, trajectories from the plain batched solver,
training of the message-passing network (BF16 MFMA kernels, csrc/mpnn.hip) and the comparison of
plain Newton, linear extrapolation, the tangent-secant predictor and the learned guess on held-out
beams.

Start points of load step k+1 (equal load increments h):
  plain     u_k (the reference's Newton start);
  extrap    u_k + (u_k - u_{k-1});
  tangent   u_k + 2 d_k - (u_k - u_{k-1}), d_k = K^-1 R(u_k; lam_{k+1}) with the factors of step k's
            last direct solve (= h u'(lam_k) to first order): the second-order Taylor predictor
            with u'' from the secant (batched direct solver only);
  learned   the network's correction on top of the tangent predictor (of the extrapolation with the
            iterative solver), kept per mesh only where it lowers the start residual.
The network sees the base predictor through a virtual previous state v_k with u_k + (u_k - v_k)
equal to the base (v_k = u_{k-1} for the extrapolation, 2 u_k - u_{k-1} - 2 d_k for the tangent
predictor), so its features, loss and guess kernels are the same for either base.

Everything numerical runs in libmpf (the batched Newton/PCG solver and the network's kernels);
torch only holds device memory for the trajectories.
"""
from __future__ import annotations

import sys
import time
from dataclasses import dataclass, field

import numpy as np

from . import BATCH_SOLVER_BAND, BATCH_SOLVER_PCG, ERR_DIVERGED, ERR_NEG_JACOBIAN, ERR_NOT_CONVERGED, HEX8, \
    TANGENT_CONSISTENT, HexBatch, MpfError, Mpnn, Problem, SplitMix64, face_nodes, hex_face_quads, matl, mesh_hex

H_ELEM, W_ELEM = 8, 4      # beam height (z) and width (y) in elements, element size 1/8
NSTEPS = 20                # tip load ramped over 20 load steps, lambda = k / 20
NU = 0.3
TIP_KAPPA = 0.075          # tip shear tau = TIP_KAPPA * E / length^2 (large-deflection regime)


@dataclass
class Beams:
    lengths: np.ndarray    # elements along x
    mu: np.ndarray


def cantilever_workload(nbeams: int = 2000, seed: int = 4) -> Beams:
    """Beams of length L = round(U[4, 12] * 8) elements, mu log-uniform in [1, 100] (SplitMix64)."""
    g = SplitMix64(seed)
    L, mu = [], []
    for _ in range(nbeams):
        L.append(int(round((4.0 + 8.0 * g.uniform()) * 8.0)))
        mu.append(10.0 ** (2.0 * g.uniform()))
    return Beams(np.array(L, dtype=np.int32), np.array(mu))


def cantilever_problem(beams: Beams, idx) -> tuple[HexBatch, Problem]:
    """Batched problem of the selected beams: x = 0 face fixed, tip face (x = length) under a dead
    shear traction along -z, in the reference's (normal, tangent1) form: on the tip face tangent1 =
    dx/dxi = +z h/2, so v1 = -2 tau / h gives the shear tau."""
    h = 1.0 / H_ELEM
    Xs, Cs, bottom, top, node_off, elem_off = [], [], [], [], [0], [0]
    mus, kinv, v1 = [], [], []
    for i in idx:
        n = int(beams.lengths[i])
        X, conn = mesh_hex(n, W_ELEM, H_ELEM, L=(n * h, W_ELEM * h, H_ELEM * h))
        o = node_off[-1]
        Xs.append(X)
        Cs.append(conn + o)
        bottom.append(face_nodes(HEX8, n, W_ELEM, H_ELEM, 0) + o)
        top.append(hex_face_quads(n, W_ELEM, H_ELEM, 1) + o)
        node_off.append(o + len(X))
        elem_off.append(elem_off[-1] + len(conn))
        mu = float(beams.mu[i])
        E = 2.0 * mu * (1.0 + NU)
        mus.append(mu)
        kinv.append(3.0 * (1.0 - 2.0 * NU) / E)
        tau = TIP_KAPPA * E / (n * h) ** 2
        v1.append(-2.0 * tau / h)
    B = HexBatch(np.concatenate(Xs), np.concatenate(Cs).astype(np.int32), np.array(node_off, dtype=np.int32),
                 np.array(elem_off, dtype=np.int32), beams.lengths[list(idx)], bottom, top)
    P = Problem(HEX8, B.X, B.conn, matl(mus[0], kinv[0], 3, False, TANGENT_CONSISTENT))
    P.set_batch(B.node_off, B.elem_off, mus, kinv)
    for m in range(len(idx)):
        P.add_dirichlet(bottom[m], -1)
        P.add_traction(top[m], 0.0, v1[m])
    P.prepare_matrix_pattern()
    return B, P


@dataclass
class StepLog:
    iters: list = field(default_factory=list)      # per step: assemblies per mesh (np.int32 array)
    start_rel: list = field(default_factory=list)  # per step: mean ||R|| / ||R0|| at the start point
    kept: list = field(default_factory=list)       # per step: meshes keeping the network's correction
    seconds: list = field(default_factory=list)    # per step: wall time of the batched step
    newton: list = field(default_factory=list)     # per step: batch Newton iterations (slowest mesh)


def _progress(msg: str):
    print(f"[learned] {msg}", file=sys.stderr, flush=True)


def uses_tangent(solver: dict) -> bool:
    """The tangent predictor needs the factors of the batched direct solver."""
    return solver.get("solver", BATCH_SOLVER_PCG) == BATCH_SOLVER_BAND


def run_steps(P: Problem, mode: str, net: Mpnn | None = None, nsteps: int = NSTEPS, store=None, vstore=None,
              **solver):
    """Ramp the load over nsteps with Newton started from: 'plain' the previous converged solution,
    'extrap' the linear extrapolation, 'tangent' the tangent-secant predictor, 'learned' the network's
    guess on the base predictor (tangent with the direct solver, else extrapolation).  store:
    optional device tensor [nsteps + 1, nNode * 3] receiving u_0 .. u_nsteps; vstore: likewise the
    virtual previous states v_k of the tangent predictor (k >= 1; v_0 = 0).  One progress line per
    load step on stderr."""
    import torch

    n3 = 3 * P.nNode
    P.set_state(np.zeros(n3))
    dev = torch.device("cuda")
    hist = [torch.zeros(n3, dtype=torch.float64, device=dev) for _ in range(2)]  # u_{k-1}, u_k
    guess = torch.zeros(n3, dtype=torch.float64, device=dev)
    tang = uses_tangent(solver)
    if mode == "tangent" and not tang:
        raise ValueError("the tangent predictor needs solver=BATCH_SOLVER_BAND")
    dtan = torch.zeros(n3, dtype=torch.float64, device=dev) if tang else None
    vprev = torch.zeros(n3, dtype=torch.float64, device=dev)
    if store is not None:
        store[0].zero_()
    if vstore is not None:
        vstore[0].zero_()
        vstore[1].zero_()
    log = StepLog()
    for k in range(nsteps):
        lam = (k + 1) / nsteps
        g = None
        torch.cuda.synchronize()
        t0 = time.perf_counter()  # the guess is part of the step's wall time
        want_v = k >= 1 and tang and (mode in ("tangent", "learned") or vstore is not None)
        if want_v:  # v_k = 2 u_k - u_{k-1} - 2 d_k
            P.batch_tangent_increment(lam, dtan)
            torch.add(hist[1], hist[1], out=vprev).sub_(hist[0]).sub_(dtan, alpha=2.0)
            if vstore is not None:
                vstore[k].copy_(vprev)
        if mode != "plain" and k >= 1:
            rho = 1.0  # equal load increments
            prev = vprev if (tang and mode in ("tangent", "learned")) else hist[0]
            net.features(P, hist[1], prev, rho, lam, k / (k + 1.0))
            if mode == "learned":
                net.forward(P)
                log.kept.append(net.guarded_guess(P, lam, hist[1], prev, rho, guess))
            else:
                net.guess(P, hist[1], prev, rho, False, guess)
            g = guess
        iters, rel, rep = P.batch_newton_step(lam, guess=g, **solver)
        log.seconds.append(time.perf_counter() - t0)
        log.iters.append(iters.copy())
        log.newton.append(rep.newton_iters)
        log.start_rel.append(rep.start_relres_mean)
        _progress(f"{mode} step {k + 1}/{nsteps}: {rep.newton_iters} batch Newton iterations, "
                  f"{iters.mean():.2f} assemblies per mesh, {log.seconds[-1]:.3f} s")
        hist[0].copy_(hist[1])
        P.copy_state_device(hist[1])
        if store is not None:
            store[k + 1].copy_(hist[1])
    return log


def train(net: Mpnn, P: Problem, traj, epochs: int = 60, lr: float = 2e-3, residual_epochs: int = 0,
          log_every: int = 10, vtraj=None):
    """Full-batch steps over the stored trajectories u_0..u_N: sample k predicts u_{k+1} from u_k and
    the previous state u_{k-1} (k = 1 .. N-1), or the virtual previous state vtraj[k] of the tangent
    predictor when given.  The first `epochs` minimise the displacement error; the following
    `residual_epochs` the linearised start residual ||K e|| / ||dlam f|| (what Newton's convergence
    depends on; one tangent assembly per sample).  Returns the mean loss per epoch."""
    nsteps = traj.shape[0] - 1
    losses = []
    total = epochs + residual_epochs
    for ep in range(total):
        tot = 0.0
        res = ep >= epochs
        for k in range(1, nsteps):
            lam = (k + 1) / nsteps
            prev = traj[k - 1] if vtraj is None else vtraj[k]
            net.features(P, traj[k], prev, 1.0, lam, k / (k + 1.0))
            net.forward(P)
            if res:
                tot += net.residual_loss(P, lam, 1.0 / nsteps, traj[k + 1], traj[k], prev, 1.0)
            else:
                tot += net.loss(P, traj[k + 1], traj[k], prev, 1.0)
            net.backward(P)
            net.adam(lr=lr * (0.5 ** ((ep - (epochs if res else 0)) / max((residual_epochs if res else epochs) / 3, 1))))
        losses.append(tot / (nsteps - 1))
        if log_every and (ep % log_every == 0 or ep == total - 1 or ep == epochs):
            kind = "start residual^2" if res else "displacement"
            _progress(f"epoch {ep}: {kind} loss {losses[-1]:.4e}")
    return losses


def summarize(log: StepLog, first: int = 1) -> dict:
    """Mean assemblies per mesh and step, batch Newton iterations and wall time over steps >= first
    (step 0 has no history, every method starts it from the previous solution)."""
    it = np.array(log.iters[first:], dtype=np.float64)
    return {"assemblies_per_mesh_step": round(float(it.mean()), 3),
            "batch_newton_iters": int(np.sum(log.newton[first:])),
            "seconds": round(float(np.sum(log.seconds[first:])), 4),
            "seconds_all_steps": round(float(np.sum(log.seconds)), 4),
            "start_relres_geomean": float(np.exp(np.mean(np.log(np.maximum(log.start_rel[first:], 1e-300))))),
            "meshes_keeping_correction_per_step": [int(x) for x in log.kept]}


def experiment(nbeams: int = 2000, ntrain: int = 1800, ntest: int = 200, epochs: int = 60, lr: float = 2e-3,
               nsteps: int = NSTEPS, seed: int = 4, autotune: bool = True, residual_epochs: int = 0,
               **solver) -> dict:
    """Trajectories of the plain solver on the training beams, training, then plain / extrapolated /
    learned Newton on the held-out beams.  Returns the measurements."""
    import torch

    beams = cantilever_workload(nbeams, seed)
    train_idx = list(range(ntrain))
    test_idx = list(range(nbeams - ntest, nbeams))
    out = {"beams": nbeams, "train": ntrain, "test": ntest, "steps": nsteps, "epochs": epochs}
    t0 = time.perf_counter()
    _, Ptr = cantilever_problem(beams, train_idx)
    if autotune:
        out["tile"], out["lanes_per_row"], _ = Ptr.batch_autotune()
    traj = torch.zeros((nsteps + 1, 3 * Ptr.nNode), dtype=torch.float64, device="cuda")
    tang = uses_tangent(solver)
    vtraj = torch.zeros_like(traj) if tang else None
    log_tr = run_steps(Ptr, "plain", None, nsteps, store=traj, vstore=vtraj, **solver)
    out["train_trajectories_seconds"] = round(time.perf_counter() - t0, 2)
    out["train_plain"] = summarize(log_tr)
    net = Mpnn(seed=seed)
    t1 = time.perf_counter()
    losses = train(net, Ptr, traj, epochs=epochs, lr=lr, residual_epochs=residual_epochs, vtraj=vtraj)
    out["train_seconds"] = round(time.perf_counter() - t1, 2)
    out["loss_first"], out["loss_last"] = losses[0], losses[epochs - 1]
    out["residual_epochs"] = residual_epochs
    if residual_epochs:
        out["start_residual_loss_first"], out["start_residual_loss_last"] = losses[epochs], losses[-1]
    out["base_predictor"] = "tangent" if tang else "extrap"
    del traj, vtraj, Ptr
    _, Pte = cantilever_problem(beams, test_idx)
    if autotune:
        Pte.batch_autotune()
    modes = ("plain", "extrap", "tangent", "learned") if tang else ("plain", "extrap", "learned")
    for mode in modes:
        out[mode] = summarize(run_steps(Pte, mode, net, nsteps, **solver))
    for mode in modes[1:]:
        out[f"speedup_{mode}_vs_plain"] = round(out["plain"]["seconds"] / max(out[mode]["seconds"], 1e-12), 3)
    return out


# ---------------------------------------------------------------------------------------------
# single large problem: the 128^3 compression block
# ---------------------------------------------------------------------------------------------
BLOCK_STEPS = 15


def compression_block(n: int, elem_type: int = HEX8) -> Problem:
    """n^3 Hex8 unit block, E = 1e3, nu = 0.49, bottom fixed, top UZ = -0.3 lambda (15 load steps);
    elem_type HEX8 (displacement) or HEX8P0 (Q1/P0, one condensed pressure per element)."""
    from . import HEX8P0, TANGENT_CONSISTENT as TC, matl_from_E_nu

    if elem_type not in (HEX8, HEX8P0):
        raise ValueError("compression_block: elem_type HEX8 or HEX8P0")
    X, conn = mesh_hex(n, n, n, jitter=0.0, seed=5)
    P = Problem(elem_type, X, conn, matl_from_E_nu(1e3, 0.49, tangent=TC))
    P.add_dirichlet(face_nodes(HEX8, n, n, n, 4), -1)
    P.add_dirichlet(face_nodes(HEX8, n, n, n, 5), 2, -0.3)
    P.prepare_matrix_pattern()
    return P


def block_newton_opts(P: Problem):
    from . import KSP_CG, PC_BJACOBI3
    # every linear solve to 1e-6: the first (lifted) iteration must carry the top-face increment into
    # the block accurately, or the element layer under the face inverts (a loose first solve does)
    return P.newton_opts(nsteps=1, dt=1.0 / BLOCK_STEPS, itmax=30, rtol=1e-9, ksp=KSP_CG, pc=PC_BJACOBI3,
                         lin_rtol=1e-6, maxit=100000, forcing=False)


_CUT_STATUS = (ERR_NEG_JACOBIAN, ERR_NOT_CONVERGED, ERR_DIVERGED)


def newton_with_cut(P: Problem, lam0: float, lam1: float, opts, guess=None, depth: int = 0, max_depth: int = 6,
                    log=None):
    """Newton from the converged state at lam0 to lam1, started from guess (or plainly).  On a negative
    Jacobian, divergence or no convergence the state is restored and the increment halved, both
    halves started plainly (MyTime::cut, MyTime.cpp:26-106; reset, SolutionDataSolid.cpp:157-167),
    down to 1/64 of the increment.  Returns (assemblies and Krylov iterations of the converged
    solves, number of cuts).  log: called after every failed attempt and converged sub-step.
    The restored state holds the pressures too (nodal for Tet10/P1, per element for Q1/P0)."""
    ustart, pstart = P.state()
    try:
        it, rep = P.newton_step(lam1, opts, guess)
        if log and depth > 0:
            log(f"  sub-step to lambda {lam1:.6f} converged: {it} assemblies, {rep.krylov_iters} PCG iterations")
        return it, rep.krylov_iters, 0
    except MpfError as e:
        if e.status not in _CUT_STATUS or depth >= max_depth:
            raise
        if log:
            log(f"  Newton to lambda {lam1:.6f} failed ({e}); cutting the increment in half")
    P.set_state(ustart, pres=pstart)
    mid = 0.5 * (lam0 + lam1)
    a1, k1, c1 = newton_with_cut(P, lam0, mid, opts, None, depth + 1, max_depth, log)
    a2, k2, c2 = newton_with_cut(P, mid, lam1, opts, None, depth + 1, max_depth, log)
    return a1 + a2, k1 + k2, 1 + c1 + c2


def block_steps(P: Problem, mode: str, nsteps: int, net: Mpnn | None = None, store=None, log=print):
    """Load steps 1..nsteps of the block from 'plain', 'extrap' or 'learned' start points."""
    import torch

    n3 = 3 * P.nNode
    P.set_state(np.zeros(n3))
    opts = block_newton_opts(P)
    dev = torch.device("cuda")
    hist = [torch.zeros(n3, dtype=torch.float64, device=dev) for _ in range(2)]
    guess = torch.zeros(n3, dtype=torch.float64, device=dev)
    if store is not None:
        store[0].zero_()
    steps = []
    for k in range(nsteps):
        lam = (k + 1) / BLOCK_STEPS
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        g = None
        kept = None
        if mode != "plain" and k >= 1:
            net.features(P, hist[1], hist[0], 1.0, lam, k / (k + 1.0))
            if mode == "learned":
                net.forward(P)
                kept = net.guarded_guess(P, lam, hist[1], hist[0], 1.0, guess)
            else:
                net.guess(P, hist[1], hist[0], 1.0, False, guess)
            g = guess.cpu().numpy()
        try:
            it, kry, cuts = newton_with_cut(P, k / BLOCK_STEPS, lam, opts, g, log=log)
        except MpfError as e:
            if e.status not in _CUT_STATUS:
                raise
            # not reached at the smallest increment: recorded, and the load path ends here
            steps.append({"step": k + 1, "failed": str(e), "seconds": round(time.perf_counter() - t0, 3)})
            if log:
                log(f"block {mode} step {k + 1}: not reached ({e})")
            break
        dt = time.perf_counter() - t0
        hist[0].copy_(hist[1])
        P.copy_state_device(hist[1])
        if store is not None:
            store[k + 1].copy_(hist[1])
        steps.append({"step": k + 1, "assemblies": it, "krylov_iterations": kry, "seconds": round(dt, 3),
                      "cuts": cuts, "network_correction_kept": kept})
        if log:
            log(f"block {mode} step {k + 1}: {it} assemblies, {kry} PCG iterations, {cuts} cuts, {dt:.2f} s")
    return steps


def block_experiment(n_train: int = 48, n_eval: int = 128, eval_steps: int = 3, epochs: int = 40,
                     lr: float = 2e-3, log=print, modes=("plain", "extrap", "learned"), elem_type: int = HEX8) -> dict:
    """Train the network on the trajectory of an n_train^3 block (all 15 steps, plain Newton), then run
    the first eval_steps steps of the n_eval^3 block from the start points in modes (plain,
    extrapolated, learned; the network is trained only when 'learned' is among them)."""
    import torch

    out = {"train_block": n_train, "eval_block": n_eval, "eval_steps": eval_steps, "epochs": epochs,
           "element": "Hex8 Q1/P0" if elem_type != HEX8 else "Hex8"}
    net = Mpnn(seed=5)
    if elem_type != HEX8 and modes != ("plain",):
        raise ValueError("block_experiment: the Q1/P0 block runs from the plain start only")
    if "learned" not in modes:
        Pe = compression_block(n_eval, elem_type)
        for mode in modes:
            out[mode] = block_steps(Pe, mode, eval_steps, net, log=log)
        Pe.close()
        return out
    Pt = compression_block(n_train)
    traj = torch.zeros((BLOCK_STEPS + 1, 3 * Pt.nNode), dtype=torch.float64, device="cuda")
    t0 = time.perf_counter()
    block_steps(Pt, "plain", BLOCK_STEPS, store=traj, log=None)
    out["train_trajectory_seconds"] = round(time.perf_counter() - t0, 2)
    t1 = time.perf_counter()
    losses = train(net, Pt, traj, epochs=epochs, lr=lr, log_every=0)
    out["train_seconds"] = round(time.perf_counter() - t1, 2)
    out["loss_first"], out["loss_last"] = losses[0], losses[-1]
    del traj, Pt
    Pe = compression_block(n_eval)
    for mode in modes:
        out[mode] = block_steps(Pe, mode, eval_steps, net, log=log)
    Pe.close()
    return out
