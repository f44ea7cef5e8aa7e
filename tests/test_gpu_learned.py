"""End to end on a few cantilevers: trajectories from the plain batched solver, training of the
BF16 network, and the held-out comparison (the learned start must save Newton iterations)."""
from __future__ import annotations

import pytest

pytestmark = pytest.mark.gpu


def test_learned_guess_saves_newton_iterations(mpf, gpu):
    from mphyfem_amd import learned
    r = learned.experiment(nbeams=12, ntrain=8, ntest=4, epochs=12, nsteps=8, autotune=False, itmax=25, rtol=1e-9,
                           cg_maxit=20000, cg_rtol=1e-10, chunk=16)
    assert r["loss_last"] < r["loss_first"]
    assert r["learned"]["assemblies_per_mesh_step"] < r["plain"]["assemblies_per_mesh_step"], r


def test_tangent_predictor_and_learned_guess_with_direct_solves(mpf, gpu):
    """With exact Newton corrections (banded direct solve per beam) the tangent-secant predictor
    saves Newton iterations over the extrapolation, which saves them over the plain start; the
    guarded learned guess is built on the tangent predictor, is never worse than the extrapolation,
    and starts no further from the solution than the predictor alone."""
    from mphyfem_amd import learned
    r = learned.experiment(nbeams=12, ntrain=8, ntest=4, epochs=12, nsteps=8, autotune=False, itmax=25, rtol=1e-9,
                           solver=mpf.BATCH_SOLVER_BAND)
    assert r["base_predictor"] == "tangent"
    a = {m: r[m]["assemblies_per_mesh_step"] for m in ("plain", "extrap", "tangent", "learned")}
    assert a["tangent"] < a["extrap"] < a["plain"], a
    assert a["learned"] <= a["extrap"], a
    # the guard weights the network's correction per beam by the minimiser of the start residual
    # linearised between the predictor and the whole correction: the start point is never further
    # (to that model) than the predictor's
    rel = {m: r[m]["start_relres_geomean"] for m in ("tangent", "learned")}
    print("start residual geomean", rel, "weights > 0 per step", r["learned"]["meshes_keeping_correction_per_step"],
          flush=True)
    assert rel["learned"] <= rel["tangent"] * (1.0 + 1e-3), rel
