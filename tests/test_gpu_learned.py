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
