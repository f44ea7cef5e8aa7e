"""CPU checks of the product library: it loads, exports every symbol declared in include/*.h,
the Python binding covers them, and the host-side mesh generators are valid.  No compute call
needs a device here."""
from __future__ import annotations

import ctypes as C
import re
import subprocess
from pathlib import Path

import numpy as np
import pytest

ROOT = Path(__file__).resolve().parent.parent


def declared_symbols():
    names = set()
    for h in (ROOT / "include").glob("*.h"):
        txt = h.read_text()
        txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
        for m in re.finditer(r"\b(mpf_[a-z0-9_]+)\s*\(", txt):
            names.add(m.group(1))
    return names


def test_library_exports_every_declared_symbol(mpf):
    names = declared_symbols()
    assert len(names) > 25
    out = subprocess.run(["nm", "-D", "--defined-only", str(mpf.LIB_PATH)], capture_output=True, text=True).stdout
    exported = {l.split()[-1] for l in out.splitlines() if l.strip()}
    missing = sorted(names - exported)
    assert not missing, missing
    bound = {s[0] for s in mpf.SIGNATURES}
    assert not sorted(names - bound), sorted(names - bound)


def test_library_has_gfx950_code(mpf):
    out = subprocess.run(["strings", str(mpf.LIB_PATH)], capture_output=True, text=True).stdout
    assert "gfx950" in out


def test_oracle_not_linked_into_product(mpf):
    out = subprocess.run(["nm", "-D", str(mpf.LIB_PATH)], capture_output=True, text=True).stdout
    assert "orc_" not in out
    ldd = subprocess.run(["ldd", str(mpf.LIB_PATH)], capture_output=True, text=True).stdout
    assert "oracle" not in ldd


def test_status_strings(mpf):
    assert mpf.status_string(3) == "Negative Jacobian in the element"
    assert mpf.status_string(0) == "ok"


def _tet_volume(X, c):
    a, b, cc, d = X[c[0]], X[c[1]], X[c[2]], X[c[3]]
    return np.linalg.det(np.stack([b - a, cc - a, d - a])) / 6.0


def test_mesh_hex(mpf):
    X, conn = mpf.mesh_hex(4, 3, 2, L=(1.0, 0.6, 0.5), jitter=0.05, seed=9)
    assert X.shape == (5 * 4 * 3, 3) and conn.shape == (24, 8)
    # faces stay planar under the per-axis interior jitter
    assert np.all(X[mpf.face_nodes(mpf.HEX8, 4, 3, 2, 1)][:, 0] == 1.0)
    assert np.all(X[mpf.face_nodes(mpf.HEX8, 4, 3, 2, 4)][:, 2] == 0.0)
    # total volume by the divergence-free corner decomposition: split hexes into 6 tets each
    tets = [(0, 1, 2, 6), (0, 2, 3, 6), (0, 3, 7, 6), (0, 7, 4, 6), (0, 4, 5, 6), (0, 5, 1, 6)]
    vol = sum(_tet_volume(X, e[list(t)]) for e in conn for t in tets)
    assert abs(vol - 0.3) < 1e-12
    X2, _ = mpf.mesh_hex(4, 3, 2, L=(1.0, 0.6, 0.5), jitter=0.05, seed=9)
    assert np.array_equal(X, X2)


def test_mesh_tet10(mpf):
    n = 3
    X, conn = mpf.mesh_tet10(n, jitter=0.1, seed=2)
    assert conn.shape == (6 * n ** 3, 10) and X.shape == ((2 * n + 1) ** 3, 3)
    vols = np.array([_tet_volume(X, c) for c in conn])
    assert np.all(vols > 0)
    assert abs(vols.sum() - 1.0) < 1e-12
    edges = [(0, 1), (1, 2), (0, 2), (0, 3), (2, 3), (1, 3)]
    for c in conn[:50]:
        for q, (a, b) in enumerate(edges):
            assert np.allclose(X[c[4 + q]], 0.5 * (X[c[a]] + X[c[b]]))
    # conforming: every face is shared by at most two tets
    faces = {}
    for e, c in enumerate(conn):
        for f in [(0, 1, 2), (0, 1, 3), (0, 2, 3), (1, 2, 3)]:
            key = tuple(sorted(c[list(f)]))
            faces[key] = faces.get(key, 0) + 1
    assert max(faces.values()) == 2
    nb = sum(1 for v in faces.values() if v == 1)
    assert nb == 6 * 2 * n * n


def test_hex_face_quads_outward(mpf):
    n = 3
    X, _ = mpf.mesh_hex(n, n, n)
    for face in range(6):
        q = mpf.hex_face_quads(n, n, n, face)
        assert q.shape == (n * n, 4)
        axis, hi = face // 2, face % 2
        for quad in q:
            P = X[quad]
            du = 0.5 * (P[1] - P[0] + P[2] - P[3])
            dv = 0.5 * (P[3] - P[0] + P[2] - P[1])
            nrm = np.cross(dv, du)
            assert np.sign(nrm[axis]) == (1 if hi else -1)


def test_no_device_is_an_error_not_a_fallback(mpf):
    if mpf.device_count() > 0:
        pytest.skip("device present")
    X, conn = mpf.mesh_hex(1, 1, 1)
    with pytest.raises(mpf.MpfError) as ei:
        mpf.Problem(mpf.HEX8, X, conn, mpf.matl(1.0, 1e-2))
    assert ei.value.status == mpf.ERR_NO_DEVICE


class _CutProblem:
    """Stand-in for Problem.newton_step: a negative Jacobian whenever the load increment exceeds
    max_inc, and a record of the load factors reached."""

    def __init__(self, max_inc):
        self.max_inc, self.lam, self.u, self.p, self.calls = max_inc, 0.0, np.zeros(3), np.zeros(2), []

    def state(self):
        return self.u.copy(), self.p.copy()

    def set_state(self, u, pres=None):
        self.u = np.array(u)
        if pres is not None:
            self.p = np.array(pres)

    def newton_step(self, lam, opts, guess):
        import types

        import mphyfem_amd as M
        self.calls.append(lam)
        if lam - self.lam > self.max_inc + 1e-12:
            self.u = self.u + 1.0  # a failed attempt leaves a spoiled state behind
            self.p = self.p - 5.0
            raise M.NegativeJacobian(M.ERR_NEG_JACOBIAN, "newton_step")
        self.p = self.p + (lam - self.lam)  # pressures accumulate over converged increments
        self.lam = lam
        self.u = np.full(3, lam)
        return 3, types.SimpleNamespace(krylov_iters=10)


def test_newton_step_cut_reaches_the_target_by_halving():
    from mphyfem_amd.learned import newton_with_cut
    P = _CutProblem(max_inc=0.3)
    it, kry, cuts = newton_with_cut(P, 0.0, 1.0, None)
    assert P.lam == 1.0 and np.all(P.u == 1.0)  # restored after each failure, then sub-stepped
    assert np.allclose(P.p, 1.0)  # the pressure state is restored with the displacements
    assert cuts == 3 and it == 4 * 3 and kry == 4 * 10  # 1 -> 0.5 -> 0.25 x2, then 0.25 x2
    assert P.calls == [1.0, 0.5, 0.25, 0.5, 1.0, 0.75, 1.0]


def test_newton_step_cut_gives_up_below_the_smallest_increment():
    import mphyfem_amd as M
    from mphyfem_amd.learned import newton_with_cut
    P = _CutProblem(max_inc=0.001)
    with pytest.raises(M.NegativeJacobian):
        newton_with_cut(P, 0.0, 1.0, None)


def test_host_layer_header_compiles_and_caller_links(mpf):
    """include/mpf_host.hpp is self-contained C++17 over mpf.h (no HIP headers), and the compiled
    caller links against libmpf.so and loads without touching a device."""
    r = subprocess.run(["g++", "-std=c++17", "-Wall", "-Werror", "-fsyntax-only", "-x", "c++", "-I",
                        str(ROOT / "include"), "-"], input='#include "mpf_host.hpp"\nint main() { return 0; }\n',
                       capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    caller = ROOT / "tests" / "host_caller"
    assert caller.exists(), "run `python -m mphyfem_amd.build`"
    r = subprocess.run([str(caller), "--link-check"], capture_output=True, text=True, timeout=60)
    assert r.returncode == 0 and "Negative Jacobian in the element" in r.stdout, (r.stdout, r.stderr)
