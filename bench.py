"""Benchmark: Neo-Hookean tangent + residual assembly into FP64 CSR on one MI355X.

Default workload (the headline configuration, Tet10/P1): 60^3 cubes split into 6 Kuhn tets each =
1 296 000 Tet10 elements (1 771 561 nodes), P1 pressure on the 226 981 vertices, interior
vertices jittered +-0.1h (seed 2), mu = 1, K = 1e3 (Kinv = 1e-3, Utype 3), 4-point rule; z=0 fixed,
z=1 given UZ = -0.05 (5% compression) applied in one load step.  One "step" = one full Newton
tangent + residual assembly (zeroMtx, every colour of the atomic-free element scatter, Dirichlet
lift) into the CSR matrix, inputs resident in HBM.  After the timed steps the linear solve of that
system (flexible GMRES with the two-level p-multigrid preconditioner, to ||b - A x|| / ||b|| < 1e-8) is
run once and reported beside it.

Beside the headline line the default run also reports "hex8" (the 128^3 Hex8 block's assembly, with
its roofline) and "load_step" (seconds per converged load step on the 200 held-out cantilevers:
plain, extrapolated, tangent-secant and learned starts, with training in the run); --no-secondary
skips both.

--workload hex8: the 128^3 Hex8 block (2 097 152 elements, nu = 0.49, bottom fixed, top UZ = -0.3),
block-Jacobi PCG beside it.

    python bench.py [--gpus 1] [--steps 10] [--warmup 2] [--workload tet10|hex8]

Prints one JSON line.  Only N = 1 is supported (the path is single-GPU).
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

# Algorithmic bytes per element (element-scatter model): npe*(4 conn + 24 X + 24 u)
# + nK*(8 read + 8 write CSR value + 4 CSR-slot index) + nR*16 (+ nP*12 pressure gather/slot).
# Hex8 Q1/P0 adds the element pressure and u_prev of its nodes (8 + 192) and writes Kup, Kpp, Fp (208).
BYTES_PER_ELEM = {"tet10": 24232, "hex8": 12320, "hex8p0": 12728}
HBM_PEAK_GBS = 8000.0
# algorithmic FP64 work of one element's tangent + residual (per Gauss point: K_e 64*72 + Gc 8*162 +
# stress 8*18 + material ~900 + geometry ~700 for Hex8, x 8 points; Tet10/P1 with its u-p blocks, x 4)
# Q1/P0: the Hex8 work plus Kup/Kpp/Fp (8 points x 30) and the condensation (576 x 3)
FLOPS_PER_ELEM = {"tet10": 4.6e4, "hex8": 6.1e4, "hex8p0": 6.3e4}
FP64_PEAK_TFLOPS = 78.6


def log(msg: str) -> None:
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def build_problem(M, n: int, workload: str):
    X, conn, _, bcs = build_problem_host(M, n, workload)
    if workload == "hex8":
        P = M.Problem(M.HEX8, X, conn, M.matl_from_E_nu(1e3, 0.49, tangent=M.TANGENT_CONSISTENT))
    elif workload == "hex8p0":
        P = M.Problem(M.HEX8P0, X, conn, M.matl_from_E_nu(1e3, 0.49))
    else:
        P = M.Problem(M.TET10P1, X, conn, M.matl(1.0, 1e-3, 3, mixed=True))
    for nodes, dof, val in bcs:
        P.add_dirichlet(nodes, dof, val)
    return X, conn, P, bcs


def cpu_baseline(M, workload: str, budget_s: float):
    """Oracle (serial C++ restatement, -O3, one core) assembling a bounded sample of the same workload."""
    from oracle import oracle as O
    n = 16 if workload == "tet10" else 32
    X, conn, _, bcs = build_problem_host(M, n, workload)
    Q = O.Problem({"tet10": 10, "hex8": 8, "hex8p0": 9}[workload], X, conn, *material_host(M, workload))
    for nodes, dof, val in bcs:
        Q.add_dirichlet(nodes, dof, val)
    Q.prepare_matrix_pattern()
    Q.set_boundary_conditions(1.0)
    d, p, nnz = Q.sizes()
    vals = np.zeros(nnz)
    rhs = np.zeros(d + p)
    done = 0
    t0 = time.perf_counter()
    while time.perf_counter() - t0 < budget_s:
        Q.assemble_range(0, len(conn), vals, rhs)
        done += len(conn)
    dt = time.perf_counter() - t0
    shape = {"tet10": f"{n}^3x6 = {len(conn)}-element Tet10/P1", "hex8": f"{n}^3 = {len(conn)}-element Hex8",
             "hex8p0": f"{n}^3 = {len(conn)}-element Hex8 Q1/P0"}[workload]
    out = {"value": round(done / dt / 1e6, 5), "unit": "Melem/s", "cores": 1, "kind": "port",
           "sample": f"oracle (serial C++, g++ -O3) assembling {done} elements: repeated full sweeps of a "
                     f"{shape} mesh of the same generator, material and boundary conditions, {dt:.1f} s"}
    # all cores: the same element loop colour by colour, OpenMP threads inside a colour
    nthr = max(1, min(os.cpu_count() or 1, int(os.environ.get("OMP_NUM_THREADS", "16") or 16)))
    color, ncol = Q.color()
    done = 0
    t0 = time.perf_counter()
    while time.perf_counter() - t0 < budget_s / 2:
        Q.assemble_colored(color, nthr, vals, rhs)
        done += len(conn)
    dt = time.perf_counter() - t0
    out["all_cores"] = {"value": round(done / dt / 1e6, 5), "unit": "Melem/s", "cores": nthr,
                        "sample": f"the same oracle loop over the {ncol} element colours with {nthr} OpenMP threads "
                                  f"inside a colour, {done} elements in {dt:.1f} s"}
    out["host"] = host_cpu()
    return out


def host_cpu() -> dict:
    """nproc, the cores this process may use, and the CPU model (as lscpu's "Model name")."""
    model = ""
    try:
        for line in Path("/proc/cpuinfo").read_text().splitlines():
            if line.startswith("model name"):
                model = line.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    try:
        usable = len(os.sched_getaffinity(0))
    except (AttributeError, OSError):
        usable = os.cpu_count() or 1
    return {"nproc": os.cpu_count(), "usable_cores": usable, "model": model}


def build_problem_host(M, n, workload):
    if workload in ("hex8", "hex8p0"):
        X, conn = M.mesh_hex(n, n, n, jitter=0.0, seed=5)
        bcs = [(M.face_nodes(M.HEX8, n, n, n, 4), -1, None), (M.face_nodes(M.HEX8, n, n, n, 5), 2, -0.3)]
    else:
        X, conn = M.mesh_tet10(n, L=1.0, jitter=0.1, seed=2)
        bcs = [(M.face_nodes(M.TET10P1, n, n, n, 4), -1, None), (M.face_nodes(M.TET10P1, n, n, n, 5), 2, -0.05)]
    return X, conn, None, bcs


def material_host(M, workload):
    if workload in ("hex8", "hex8p0"):
        m = M.matl_from_E_nu(1e3, 0.49)
        return m.mu, m.Kinv, 3
    return 1.0, 1e-3, 3


def read_traffic(workload: str):
    """HBM bytes per assembly of the gather kernels, measured with rocprofv3 FETCH_SIZE / WRITE_SIZE and
    scaled by a copy calibration (profiles/traffic.json, written from the runs its .md files list)."""
    f = ROOT / "profiles" / "traffic.json"
    if not f.exists():
        return None
    try:
        return json.loads(f.read_text()).get(workload, {}).get("bytes_per_launch")
    except Exception:
        return None


SOLVER = dict(itmax=25, rtol=1e-9, cg_maxit=20000, cg_rtol=1e-10, chunk=32)


def solver_opts(M, a) -> dict:
    """Batched Newton options: Jacobi-PCG (inexact Newton) or the banded LDL^T direct solve per mesh."""
    return dict(SOLVER, solver=M.BATCH_SOLVER_BAND if a.solver == "band" else M.BATCH_SOLVER_PCG)


def bench_batch(M, a) -> int:
    """One load step (lambda = 1) of every block of the batched workload, converged per block to
    ||R||/||R0|| < 1e-9, after the tile autotuner."""
    import torch

    edges, E, nu, v0, v1 = M.batch_workload(a.nmesh)
    t0 = time.perf_counter()
    B, P = M.batch_problem(edges, E, nu, v0, v1)
    P.prepare_matrix_pattern()
    t_setup = time.perf_counter() - t0
    nElem, nNode = len(B.conn), len(B.X)
    log(f"batch: {a.nmesh} blocks, {nElem} elements, {nNode} nodes, setup {t_setup:.2f} s")
    T, L, ms_it = P.batch_autotune()
    amode, acap, ams = P.batch_assembly_choice()
    aname = {M.ASSEMBLY_GATHER: f"gather, chunk kernel staging <= {acap} elements",
             M.ASSEMBLY_GATHER_DIRECT: "gather, direct record reads",
             M.ASSEMBLY_COLORED_SCATTER: "coloured scatter"}[amode]
    log(f"autotune: {aname} ({ams:.2f} ms per assembly); tile {T}, {L} lanes per row, {ms_it:.3f} ms per CG iteration")
    torch.cuda.synchronize()
    # warm-up steps from the undeformed state (graph capture, the first solve's CG count), then the
    # timed step from the undeformed state again
    zero = np.zeros(3 * nNode)
    capture = 0.0
    for _ in range(max(a.warmup, 1)):
        _, _, wrep = P.batch_newton_step(1.0, **solver_opts(M, a))
        capture += wrep.seconds_capture
        P.set_state(disp=zero, disp_prev=zero)
    iters, rel, rep = P.batch_newton_step(1.0, **solver_opts(M, a))
    log(f"newton: {rep.newton_iters} batch iterations, max CG {rep.cg_iters_max}, {rep.seconds:.3f} s "
        f"({rep.iteration_graphs} iteration graphs, {rep.chunk_graphs} chunk graphs; capture {capture:.3f} s "
        f"in the warm-up)")
    lin = "banded LDL^T" if a.solver == "band" else "Jacobi-PCG"
    out = {"metric": f"s per converged load step ({a.nmesh} Hex8 blocks, batched Newton + {lin})",
           "value": round(rep.seconds, 4), "unit": "s", "n_gpus": 1, "steps": 1, "warmup": max(a.warmup, 1),
           "ms_per_step": round(rep.seconds * 1e3, 2), "higher_is_better": False, "scaling": "weak",
           "vs_baseline": None, "dtype": "f64", "data": "synthetic",
           "config": {"workload": f"{a.nmesh} unit-cube Hex8 blocks, edge U{{8..32}}, E log-U[1e2,1e4], "
                                  f"nu U[0.2,0.45], top dead traction U[0.01,0.1] E, bottom fixed, seed 3",
                      "elements": nElem, "nodes": nNode, "tile": T, "lanes_per_row": L,
                      "assembly": aname, "ms_per_assembly": round(ams, 3)},
           "newton": {"batch_iterations": rep.newton_iters, "assemblies_per_block_mean": float(iters.mean()),
                      "assemblies_per_block_max": int(iters.max()), "cg_iterations_slowest": rep.cg_iters_max,
                      "cg_iterations_all_blocks": rep.cg_iters_total, "graph_launches": rep.graph_launches,
                      "iteration_graphs": rep.iteration_graphs, "chunk_graphs": rep.chunk_graphs,
                      "seconds_iteration_graphs": round(rep.seconds_assembly, 4),
                      "seconds_chunk_graphs": round(rep.seconds_solve, 4),
                      "seconds_capture_warmup": round(capture, 4), "max_relres": float(rel.max())}}
    P.close()
    print(json.dumps(out), flush=True)
    return 0


def bench_learned(M, a) -> int:
    from mphyfem_amd import learned

    r = learned.experiment(nbeams=a.beams, ntrain=a.train, ntest=a.test, epochs=a.epochs,
                           residual_epochs=a.residual_epochs, **solver_opts(M, a))
    out = {"metric": "s per converged load step, learned vs plain Newton start (held-out cantilevers; learned = "
                     "network correction on the tangent-secant predictor, guarded)",
           "value": round(r["learned"]["seconds"] / max(r["steps"] - 1, 1), 4), "unit": "s",
           "n_gpus": 1, "steps": r["steps"], "warmup": 0, "higher_is_better": False, "scaling": "weak",
           "vs_baseline": None, "dtype": "f64 solve, bf16/fp32 network", "data": "synthetic",
           "config": {"workload": "Hex8 cantilevers H=8 W=4 L=round(U[4,12]*8), mu log-U[1,100], nu=0.3, "
                                  "tip shear ramped over 20 steps, seed 4",
                      "linear_solver": "banded LDL^T per beam" if a.solver == "band" else "Jacobi-PCG per beam"},
           "learned": r}
    print(json.dumps(out), flush=True)
    return 0


def bench_direct(M, a) -> int:
    """Plain load steps of the learned workload's 200 held-out cantilevers, every Newton correction
    by the banded LDL^T direct solve of each beam (the reference's full Newton with a direct solve)."""
    import torch

    from mphyfem_amd import learned

    beams = learned.cantilever_workload(2000)
    _, P = learned.cantilever_problem(beams, list(range(1800, 2000)))
    P.set_state(np.zeros(3 * P.nNode))
    opts = dict(SOLVER, solver=M.BATCH_SOLVER_BAND)
    nsteps = learned.NSTEPS
    warm = min(max(a.warmup, 1), nsteps - 1)
    steps = min(max(a.steps, 1), nsteps - warm)
    for k in range(warm):  # band layout, graph capture
        P.batch_newton_step((k + 1) / nsteps, **opts)
    nblk, bmax, ndof, flops = P.band_stats()
    nseg, ncut = P.band_segments()
    log(f"direct: {nblk} beams ({ncut} cut in two, {nseg} segments), {ndof} free dofs, half-bandwidth <= {bmax}, "
        f"{flops / 1e9:.1f} GFLOP per factorisation")
    rep = M.BatchReport()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for k in range(warm, warm + steps):
        P.batch_newton_step((k + 1) / nsteps, rep=rep, **opts)
    torch.cuda.synchronize()
    dt = time.perf_counter() - t0
    s_solve = P.batch_time_direct(5) * 1e-3  # the Newton iterations' graphs hold assembly and solve together
    achieved = flops / s_solve / 1e12
    out = {"metric": "s per converged load step (200 Hex8 cantilevers, batched full Newton + banded LDL^T)",
           "value": round(dt / steps, 4), "unit": "s", "n_gpus": 1, "steps": steps, "warmup": warm,
           "ms_per_step": round(dt / steps * 1e3, 2), "higher_is_better": False, "scaling": "weak",
           "vs_baseline": None, "dtype": "f64", "data": "synthetic",
           "config": {"workload": "the learned workload's 200 held-out Hex8 cantilevers (H=8, W=4, L=round(U[4,12]*8)), "
                                  "plain start, load steps of the 20-step tip-shear ramp",
                      "beams": nblk, "free_dofs": ndof, "max_half_bandwidth": bmax, "band_segments": nseg,
                      "beams_cut_at_a_separator": ncut},
           "roofline": {"bound": "mfma", "achieved": round(achieved, 3), "peak": 78.6, "unit": "TFLOP/s",
                        "frac": round(achieved / 78.6, 4), "traffic": None,
                        "kernel": "k_band_fill + k_band_factor + k_band_schur + k_band_solve + k_band_sep + "
                                  "k_band_scatter (one direct solve of every beam; FP64 MFMA peak)",
                        "flops_per_solve": flops, "ms_per_solve": round(s_solve * 1e3, 3)},
           "newton": {"batch_iterations": rep.newton_iters, "graph_launches": rep.graph_launches,
                      "seconds_iteration_graphs": round(rep.seconds_assembly, 4)}}
    P.close()
    print(json.dumps(out), flush=True)
    return 0


def bench_full(M, a) -> int:
    """Full-scale Newton: 128^3 Hex8, E = 1e3, nu = 0.49, bottom fixed, top UZ = -0.3 lambda over 15 load
    steps; block-Jacobi PCG to 1e-6 per Newton iteration, ||R||/||R0|| < 1e-9 per step.  The first
    --full-steps steps from the previous converged state (plain), from the extrapolation of the last
    two converged steps, and from the learned guess (network trained on a 48^3 block's trajectory)."""
    from mphyfem_amd import learned

    n = a.n or 128
    modes = tuple(m for m in a.full_modes.split(",") if m)
    p0 = a.full_elem == "hex8p0"
    if p0:
        modes = ("plain",)
    r = learned.block_experiment(n_train=48, n_eval=n, eval_steps=a.full_steps, epochs=a.epochs, log=log,
                                 modes=modes, elem_type=M.HEX8P0 if p0 else M.HEX8)

    def per_step(mode):
        later = [s for s in r[mode][1:] if "failed" not in s]  # step 1: every mode starts plainly
        return sum(s["seconds"] for s in later) / max(len(later), 1)

    t = {m: per_step(m) for m in modes}
    first = modes[0]
    elem = "Hex8 Q1/P0" if p0 else "Hex8"
    out = {"metric": f"s per converged load step ({n}^3 {elem}, nu=0.49, block-Jacobi PCG Newton)",
           "value": round(t[first], 3), "unit": "s", "n_gpus": 1, "steps": a.full_steps, "warmup": 0,
           "higher_is_better": False, "scaling": "weak", "vs_baseline": None, "dtype": "f64", "data": "synthetic",
           "config": {"workload": f"{n}^3 {elem} block, E=1e3, nu=0.49, bottom fixed, top UZ=-0.3 lambda, 15 steps; "
                                  f"first {a.full_steps} measured, start points {','.join(modes)} (value: {first})"},
           "seconds_per_step": {m: round(v, 3) for m, v in t.items()},
           "steps_converged": {m: sum(1 for s in r[m] if "failed" not in s) for m in modes},
           "speedup_vs_plain": {m: round(t["plain"] / v, 3) for m, v in t.items() if "plain" in t and m != "plain" and v},
           "full": r}
    print(json.dumps(out), flush=True)
    return 0


def measured_copy_gbs(reps: int = 5) -> float:
    """Device-to-device copy bandwidth (read + write bytes per second) of 1 GiB, HIP events."""
    import torch

    n = 1 << 27
    x = torch.ones(n, dtype=torch.float64, device="cuda")
    y = torch.empty_like(x)
    y.copy_(x)
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record()
    for _ in range(reps):
        y.copy_(x)
    e1.record()
    torch.cuda.synchronize()
    ms = e0.elapsed_time(e1) / reps
    del x, y
    return 2.0 * n * 8 / (ms * 1e-3) / 1e9


def copy_calibration() -> int:
    """Three device-to-device copies of 1 GiB: what FETCH_SIZE / WRITE_SIZE read for known traffic."""
    import torch

    n = 1 << 27
    x = torch.ones(n, dtype=torch.float64, device="cuda")
    y = torch.empty_like(x)
    for _ in range(3):
        y.copy_(x)
    torch.cuda.synchronize()
    log("copied 3 x 1 GiB")
    return 0


def assembly_bench(M, a, wl: str, n: int):
    """The timed assembly steps of one workload: (bench line dict, problem, host data)."""
    import torch

    log(f"workload {wl}, n = {n}")
    X, conn, P, _ = build_problem(M, n, wl)
    nElem = len(conn)
    t0 = time.perf_counter()
    P.set_assembly_mode({"gather": M.ASSEMBLY_GATHER, "direct": M.ASSEMBLY_GATHER_DIRECT,
                         "scatter": M.ASSEMBLY_COLORED_SCATTER}[a.assembly])
    P.prepare_matrix_pattern()
    t_pattern = time.perf_counter() - t0
    dispDOF, presDOF, nnz, ncolors = P.sizes()
    log(f"pattern: {dispDOF}+{presDOF} dofs, nnz {nnz}, {ncolors} colours, {t_pattern:.2f} s")
    P.set_boundary_conditions(1.0)
    for _ in range(a.warmup):
        P.calc_stiffness_and_residual(True, 1.0)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(a.steps):
        P.calc_stiffness_and_residual(True, 1.0)
    torch.cuda.synchronize()
    dt = time.perf_counter() - t0
    ms_step = dt / a.steps * 1e3
    value = nElem / (ms_step * 1e-3) / 1e6
    log(f"assembly {ms_step:.3f} ms/step -> {value:.1f} Melem/s")
    # live HIP-event timing of the element-kernel launches on the problem's stream
    ms_all, ms_kern = P.time_assembly(max(5, a.steps), first_iter=True, lam=1.0)
    bpe = BYTES_PER_ELEM[wl]
    achieved = bpe * nElem / (ms_kern * 1e-3) / 1e9
    if a.assembly != "scatter":
        kname = "k_stage + k_flat + k_rhs + k_lift (row-owner gather, 4 launches"
    else:
        kname = {"tet10": "tet10_kernel<4,false>", "hex8": "hex8_kernel<8,false>",
                 "hex8p0": "hex8p0_kernel<true>"}[wl] + f" ({ncolors} colour launches"
    # secondaries: the unique bytes an assembly must move (CSR values written once, mesh and state
    # read once, residual written once) and the FP64 rate of the element work
    npe = conn.shape[1]
    unique = nnz * 8 + nElem * npe * 4 + len(X) * 48 + (dispDOF + presDOF) * 8 + presDOF * 8
    flops = FLOPS_PER_ELEM[wl] * nElem
    roofline = {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": round(achieved / HBM_PEAK_GBS, 4),
                "traffic": read_traffic(wl if a.assembly == "gather" else wl + "_" + a.assembly),
                "kernel": f"{kname} per assembly)", "assembly": a.assembly,
                "kernel_ms_per_assembly": round(ms_kern, 4), "bytes_per_elem": bpe,
                "unique_bytes": unique,
                "unique_frac": round(unique / (ms_kern * 1e-3) / 1e9 / HBM_PEAK_GBS, 4),
                "fp64_tflops": round(flops / (ms_kern * 1e-3) / 1e12, 3),
                "fp64_frac": round(flops / (ms_kern * 1e-3) / 1e12 / FP64_PEAK_TFLOPS, 4)}
    copy_gbs = measured_copy_gbs()
    roofline["measured_copy_gbs"] = round(copy_gbs, 1)
    roofline["frac_of_measured_copy"] = round(achieved / copy_gbs, 4)
    if wl == "tet10":
        metric = "Melem/s assembly (Tet10/P1 tangent+residual into FP64 CSR)"
        desc = f"Tet10/P1 mixed Neo-Hookean, {n}^3x6 Kuhn tets ({nElem} elements), one tangent+residual " \
               f"assembly at 5% compression"
    elif wl == "hex8p0":
        metric = "Melem/s assembly (Hex8 Q1/P0 condensed tangent+residual into FP64 CSR)"
        desc = f"Hex8 Q1/P0 mixed Neo-Hookean nu=0.49, {n}^3 block ({nElem} elements), one condensed " \
               f"tangent+residual assembly"
    else:
        metric = "Melem/s assembly (Hex8 tangent+residual into FP64 CSR)"
        desc = f"Hex8 Neo-Hookean nu=0.49, {n}^3 block ({nElem} elements), one tangent+residual assembly"
    out = {"metric": metric, "value": round(value, 2), "unit": "Melem/s", "n_gpus": 1, "steps": a.steps,
           "warmup": a.warmup, "ms_per_step": round(ms_step, 4), "higher_is_better": True, "scaling": "weak",
           "vs_baseline": None, "dtype": "f64", "data": "synthetic",
           "config": {"workload": desc, "elements": nElem, "dofs": dispDOF + presDOF, "nnz": nnz,
                      "colours": ncolors, "pattern_seconds": round(t_pattern, 3)},
           "roofline": roofline}
    return out, P, {"X": X, "conn": conn, "dispDOF": dispDOF, "presDOF": presDOF, "nnz": nnz}


def main(argv=None) -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--workload", choices=["tet10", "hex8", "hex8p0", "batch", "learned", "direct", "full"],
                    default="tet10",
                    help="tet10/hex8: assembly; batch: 256 independent Hex8 blocks, batched Newton + Jacobi-PCG; "
                         "learned: cantilever beams, plain vs learned Newton initial guess; "
                         "direct: the 200 held-out cantilevers, plain full Newton with the banded direct solve; "
                         "full: 128^3 Hex8 nu=0.49 Newton load steps, plain vs extrapolated start")
    ap.add_argument("--full-steps", type=int, default=2, help="full: load steps measured (of 15)")
    ap.add_argument("--full-elem", choices=["hex8", "hex8p0"], default="hex8",
                    help="full: displacement Hex8, or Hex8 Q1/P0 (plain start only)")
    ap.add_argument("--full-modes", default="plain,extrap,learned",
                    help="full: comma-separated start points to run (plain, extrap, learned)")
    ap.add_argument("--nmesh", type=int, default=256, help="batch: number of blocks")
    ap.add_argument("--beams", type=int, default=2000, help="learned: beams (the last --test are held out)")
    ap.add_argument("--train", type=int, default=1800)
    ap.add_argument("--test", type=int, default=200)
    ap.add_argument("--epochs", type=int, default=60)
    ap.add_argument("--residual-epochs", type=int, default=0,
                    help="learned: extra epochs on the linearised start residual")
    ap.add_argument("--solver", choices=["pcg", "band"], default=None,
                    help="batch/learned: per-mesh linear solve, Jacobi-PCG (inexact Newton; batch default) or "
                         "banded LDL^T (exact Newton corrections, the reference's plain method; learned default)")
    ap.add_argument("--n", type=int, default=0, help="cubes per edge (default 60 for tet10, 128 for hex8)")
    ap.add_argument("--krylov-maxit", type=int, default=300)
    ap.add_argument("--no-krylov", action="store_true")
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-secondary", action="store_true",
                    help="tet10: skip the Hex8 assembly and load-step objects beside the headline line")
    ap.add_argument("--assembly", choices=["gather", "direct", "scatter"], default="gather",
                    help="row-owner gather assembly (default; Hex8 stages chunks of element records in LDS), "
                         "the gather with every item reading the records directly, or coloured element scatter")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="only copy 1 GiB device-to-device 3 times (reference run for the HBM byte counters)")
    a = ap.parse_args(argv)
    if a.copy_calibration:
        return copy_calibration()
    if a.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py: this path runs on one GPU only (no multi-GPU code); run with --gpus 1", file=sys.stderr)
        return 2

    import torch

    import mphyfem_amd as M

    if not torch.cuda.is_available() or M.device_count() == 0:
        print("bench.py: no HIP device", file=sys.stderr)
        return 1
    torch.cuda.init()
    wl = a.workload
    if a.solver is None:
        a.solver = "band" if wl == "learned" else "pcg"
    if wl == "batch":
        return bench_batch(M, a)
    if wl == "learned":
        return bench_learned(M, a)
    if wl == "direct":
        return bench_direct(M, a)
    if wl == "full":
        return bench_full(M, a)
    n = a.n or (60 if wl == "tet10" else 128)
    out, P, info = assembly_bench(M, a, wl, n)
    X, conn = info["X"], info["conn"]
    dispDOF, presDOF, nnz = info["dispDOF"], info["presDOF"], info["nnz"]
    if not a.no_krylov:
        if wl == "tet10":
            restart = 100
            r = P.factorise_and_solve(ksp=M.KSP_GMRES, pc=M.PC_PMG, rtol=1e-8, maxit=a.krylov_maxit, restart=restart)
            name = "gmres"
            pc = ("flexible GMRES(100), three-level p-multigrid (degree-1 Chebyshev node-block smoothing; Galerkin "
                  "vertex level with degree-1 Chebyshev vertex-block smoothing; aggregates of 8 vertices by "
                  "block-Jacobi CG to 0.1), pressure block diag(Kpp)(1 + K/mu)")
        else:
            r = P.factorise_and_solve(ksp=M.KSP_CG, pc=M.PC_BJACOBI3, rtol=1e-8, maxit=a.krylov_maxit)
            name = "pcg"
            pc = "block-Jacobi (3x3 node blocks)"
        ms_it = r.seconds / max(r.iterations, 1) * 1e3
        out[name] = {"pc": pc, "iterations": r.iterations, "relres": r.relres, "converged": bool(r.converged),
                     "seconds": round(r.seconds, 4), "ms_per_iteration": round(ms_it, 4)}
        if name == "gmres":
            st = P.pmg_stats()
            # the same solve again: the transfer tables and aggregates (built once per pattern, on the
            # host) are then in place, as for every Newton iteration after the first
            r2 = P.factorise_and_solve(ksp=M.KSP_GMRES, pc=M.PC_PMG, rtol=1e-8, maxit=a.krylov_maxit, restart=restart)
            st2 = P.pmg_stats()
            out[name].update({"seconds_repeat": round(r2.seconds, 4), "setup_seconds_repeat": round(st2.setup_seconds, 4),
                              "iterations_repeat": r2.iterations})
            its = max(r.iterations, 1)
            # per outer iteration: A z (12 bytes per nonzero), the V-cycle's 2 x degree displacement-
            # row sweeps over the FP32 copy of the values (counted as full matrix reads, 8 bytes per
            # nonzero), the lowest CG's matrix (FP32 values and int32 columns, 8 bytes per nonzero)
            # and ~18 vector streams per iteration, and the Arnoldi step's CGS2 over on average
            # m/2 + 1 basis vectors (4 reads each) plus ~10 vector streams; with three levels the
            # vertex level takes 2 x degree + 1 products of its FP32 matrix per V-cycle
            deg = st.degree
            fine = 12 * nnz + 2 * deg * 8 * nnz
            if st.levels == 3:
                coarse = ((2 * deg + 1) * (8 * st.coarse_nnz + 64 * st.coarse_dofs)
                          + st.coarse_iterations / its * (8 * st.agg_nnz + 144 * st.agg_dofs))
            else:
                coarse = st.coarse_iterations / its * (8 * st.coarse_nnz + 144 * st.coarse_dofs)
            ortho = (4 * (restart / 2 + 1) + 10) * 8 * (dispDOF + presDOF)
            bpi = fine + coarse + ortho
            out[name].update({"levels": st.levels, "vcycles": st.vcycles, "coarse_dofs": st.coarse_dofs,
                              "coarse_nnz": st.coarse_nnz, "agg_dofs": st.agg_dofs, "agg_nnz": st.agg_nnz,
                              "lowest_cg_iterations": st.coarse_iterations,
                              "lowest_cg_per_vcycle": round(st.coarse_iterations / max(st.vcycles, 1), 1),
                              "setup_seconds": round(st.setup_seconds, 4),
                              "smoother_lambda_max": round(st.lambda_max, 4),
                              "vertex_lambda_max": round(st.lambda_vertex, 4)})
            out[name]["roofline"] = {"bound": "hbm", "bytes_per_iteration": int(bpi),
                                     "split": {"fine": int(fine), "coarse": int(coarse), "orthogonalisation": int(ortho)},
                                     "achieved": round(bpi / (ms_it * 1e-3) / 1e9, 1), "peak": HBM_PEAK_GBS,
                                     "unit": "GB/s", "frac": round(bpi / (ms_it * 1e-3) / 1e9 / HBM_PEAK_GBS, 4)}
        if name == "pcg":  # CSR values + int32 columns once, ~6 vector streams of n doubles per iteration
            bpi = 12 * nnz + 144 * (dispDOF + presDOF)
            out[name]["roofline"] = {"bound": "hbm", "bytes_per_iteration": bpi,
                                     "achieved": round(bpi / (ms_it * 1e-3) / 1e9, 1), "peak": HBM_PEAK_GBS,
                                     "unit": "GB/s", "frac": round(bpi / (ms_it * 1e-3) / 1e9 / HBM_PEAK_GBS, 4)}
        log(f"{name} {r.iterations} its relres {r.relres:.2e} {r.seconds:.2f} s")
    if not a.no_cpu:
        log("cpu baseline")
        out["cpu_baseline"] = cpu_baseline(M, wl, a.cpu_seconds)
    P.close()
    if wl == "tet10" and not a.no_secondary:
        # the full-scale Hex8 block's assembly beside the headline one
        log("hex8 assembly, 128^3 block")
        h, HP, _ = assembly_bench(M, a, "hex8", 128)
        HP.close()
        out["hex8"] = {k: h[k] for k in ("metric", "value", "unit", "ms_per_step", "config", "roofline")}
        # the second headline metric: seconds per converged load step, learned vs plain
        log("load step: learned vs plain on the held-out cantilevers")
        out["load_step"] = load_step_summary(M, a)
    print(json.dumps(out), flush=True)
    return 0


def load_step_summary(M, a) -> dict:
    """Seconds per converged load step (||R|| / ||R0|| < 1e-9, R0 at the plain start) on the held-out
    cantilevers: plain (previous converged state), linear extrapolation, the tangent-secant
    predictor, and the learned guess (the network's correction on that predictor, guarded); full
    Newton with the banded direct solve per beam; the network trained on the plain solver's own
    trajectories of the training beams in the same run."""
    from mphyfem_amd import learned

    r = learned.experiment(nbeams=a.beams, ntrain=a.train, ntest=a.test, epochs=a.epochs,
                           residual_epochs=a.residual_epochs, **dict(SOLVER, solver=M.BATCH_SOLVER_BAND))
    nst = max(r["steps"] - 1, 1)
    modes = [m for m in ("plain", "extrap", "tangent", "learned") if m in r]
    out = {"metric": "s per converged load step, learned vs plain Newton start (||R||/||R0|| < 1e-9)",
           "unit": "s", "higher_is_better": False,
           "config": {"workload": f"{a.beams} Hex8 cantilevers H=8 W=4 L=round(U[4,12]*8), mu log-U[1,100], "
                                  f"nu=0.3, tip shear ramped over {r['steps']} steps, seed 4; {a.train} trained on, "
                                  f"{a.test} held out; steps 2-{r['steps']} timed",
                      "linear_solver": "banded LDL^T per beam (full Newton)", "epochs": a.epochs}}
    for m in modes:
        out[m] = {"s_per_step": round(r[m]["seconds"] / nst, 4),
                  "assemblies_per_mesh_step": r[m]["assemblies_per_mesh_step"],
                  "batch_newton_iters": r[m]["batch_newton_iters"]}
    out["value"] = out["learned"]["s_per_step"]
    out["speedup_vs_plain"] = {m: round(r["plain"]["seconds"] / max(r[m]["seconds"], 1e-12), 3) for m in modes}
    if "tangent" in r:
        out["learned_vs_tangent"] = round(r["tangent"]["seconds"] / max(r["learned"]["seconds"], 1e-12), 3)
    out["meshes_keeping_correction_per_step"] = r["learned"]["meshes_keeping_correction_per_step"]
    out["train_seconds"] = r.get("train_seconds")
    out["loss_first_last"] = [r.get("loss_first"), r.get("loss_last")]
    return out


if __name__ == "__main__":
    raise SystemExit(main())
