"""Benchmark: Tet10/P1 mixed Neo-Hookean tangent + residual assembly into FP64 CSR on one MI355X.

Workload (the headline configuration): 60^3 cubes split into 6 Kuhn tets each = 1 296 000 Tet10
elements (1 771 561 nodes), P1 pressure on the 226 981 vertices, interior vertices jittered
+-0.1h (seed 2), mu = 1, K = 1e3 (Kinv = 1e-3, Utype 3), 4-point rule; z=0 fixed, z=1 given
UZ = -0.05 (5% compression) applied in one load step.  One "step" = one full Newton tangent +
residual assembly (zeroMtx, every colour of the atomic-free element scatter, Dirichlet lift) into
the CSR matrix, inputs resident in HBM.  After the timed steps the block-Jacobi GMRES solve of
that system is run once and reported beside it.

    python bench.py [--gpus 1] [--steps 10] [--warmup 2]

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

BYTES_PER_ELEM_TET10P1 = 24232   # npe*(4+24+24) + 1156*(8+8+4) + 34*16 + 4*12
HBM_PEAK_GBS = 8000.0


def log(msg: str) -> None:
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def build_problem(M, n: int):
    X, conn = M.mesh_tet10(n, L=1.0, jitter=0.1, seed=2)
    P = M.Problem(M.TET10P1, X, conn, M.matl(1.0, 1e-3, 3, mixed=True))
    P.add_dirichlet(M.face_nodes(M.TET10P1, n, n, n, 4), -1)
    P.add_dirichlet(M.face_nodes(M.TET10P1, n, n, n, 5), 2, -0.05)
    return X, conn, P


def cpu_baseline(M, n_sample: int, budget_s: float):
    """Oracle (serial C++ restatement, -O3) assembling a bounded sample of the same element family."""
    from oracle import oracle as O
    X, conn = M.mesh_tet10(n_sample, L=1.0, jitter=0.1, seed=2)
    Q = O.Problem(10, X, conn, 1.0, 1e-3, 3)
    Q.add_dirichlet(M.face_nodes(M.TET10P1, n_sample, n_sample, n_sample, 4), -1)
    Q.add_dirichlet(M.face_nodes(M.TET10P1, n_sample, n_sample, n_sample, 5), 2, -0.05)
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
    return {"value": done / dt / 1e6, "unit": "Melem/s", "cores": 1, "kind": "port",
            "sample": f"oracle (serial C++, g++ -O3) assembling {done} Tet10/P1 elements: repeated sweeps of a "
                      f"{n_sample}^3x6 = {len(conn)}-element mesh of the same generator, material and load, "
                      f"{dt:.1f} s"}


def read_traffic(name: str):
    f = ROOT / "profiles" / name
    if not f.exists():
        return None
    try:
        return json.loads(f.read_text()).get("bytes_per_launch")
    except Exception:
        return None


def main(argv=None) -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--n", type=int, default=60, help="cubes per edge (60 -> 1 296 000 Tet10)")
    ap.add_argument("--gmres-maxit", type=int, default=300)
    ap.add_argument("--no-gmres", action="store_true")
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    ap.add_argument("--no-cpu", action="store_true")
    a = ap.parse_args(argv)
    if a.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py: this path runs on one GPU only (no multi-GPU code); run with --gpus 1", file=sys.stderr)
        return 2

    import torch

    import mphyfem_amd as M

    if not torch.cuda.is_available() or M.device_count() == 0:
        print("bench.py: no HIP device", file=sys.stderr)
        return 1
    torch.cuda.init()
    n = a.n
    log(f"mesh {n}^3 x 6 Tet10")
    X, conn, P = build_problem(M, n)
    nElem = len(conn)
    t0 = time.perf_counter()
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
    achieved = BYTES_PER_ELEM_TET10P1 * nElem / (ms_kern * 1e-3) / 1e9
    traffic = read_traffic("tet10_scatter_traffic.json")
    roofline = {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic,
                "kernel": "tet10_kernel<EPW,false> (all colour launches of one assembly)",
                "kernel_ms_per_assembly": round(ms_kern, 4), "launches_per_assembly": ncolors,
                "bytes_per_elem": BYTES_PER_ELEM_TET10P1}
    out = {"metric": "Melem/s assembly (Tet10/P1 tangent+residual into FP64 CSR)", "value": round(value, 2),
           "unit": "Melem/s", "n_gpus": 1, "steps": a.steps, "warmup": a.warmup, "ms_per_step": round(ms_step, 4),
           "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f64", "data": "synthetic",
           "config": {"workload": f"Tet10/P1 mixed Neo-Hookean, {n}^3x6 Kuhn tets ({nElem} elements), "
                                  f"one tangent+residual assembly at 5% compression",
                      "elements": nElem, "dofs": dispDOF + presDOF, "nnz": nnz, "colours": ncolors,
                      "pattern_seconds": round(t_pattern, 3)},
           "roofline": roofline}
    if not a.no_gmres:
        r = P.factorise_and_solve(ksp=M.KSP_GMRES, pc=M.PC_BJACOBI3, rtol=1e-8, maxit=a.gmres_maxit, restart=50)
        out["gmres"] = {"pc": "block-Jacobi (3x3 node blocks, 1x1 pressure)", "restart": 50,
                        "iterations": r.iterations, "relres": r.relres, "converged": bool(r.converged),
                        "seconds": round(r.seconds, 3),
                        "ms_per_iteration": round(r.seconds / max(r.iterations, 1) * 1e3, 3)}
        log(f"gmres {r.iterations} its relres {r.relres:.2e} {r.seconds:.2f} s")
    if not a.no_cpu:
        log("cpu baseline")
        out["cpu_baseline"] = cpu_baseline(M, 16, a.cpu_seconds)
    P.close()
    print(json.dumps(out), flush=True)
    return 0


if __name__ == "__main__":
    raise SystemExit(main())
