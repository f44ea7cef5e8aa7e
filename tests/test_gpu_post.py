"""Post-processing on the GPU (csrc/post.hip) against the oracle: Gauss-point Cauchy stress, von
Mises and tr(sigma)/3 for every element type within 1e-12 (Elem_Solid_3D_Disp::projectStress,
Elem_Solid_3D_Disp.cpp:480-557), element and nodal means within 1e-12 (femSolidmechanics3.cpp:
918-960), and the .vtu file (femSolidmechanics3.cpp:267-776): ASCII and raw-appended binary
carry the same arrays, points at X + u, the reference's Tet10 node map."""
from __future__ import annotations

import xml.etree.ElementTree as ET

import numpy as np
import pytest

from tests.cases import normwise_rel

pytestmark = pytest.mark.gpu


def _problem(mpf, orc, etype, seed=7):
    rng = np.random.default_rng(seed)
    if etype in (mpf.HEX8, mpf.HEX8P0):
        X, conn = mpf.mesh_hex(3, 3, 2, jitter=0.1, seed=seed)
    else:
        X, conn = mpf.mesh_tet10(2, jitter=0.1, seed=seed)
    mixed = etype in (mpf.TET10P1, mpf.HEX8P0)
    m = mpf.matl(20.0, 2e-3, 3, mixed=mixed)
    P = mpf.Problem(etype, X, conn, m)
    P.prepare_matrix_pattern()
    O = orc.Problem(etype, X, conn, 20.0, 2e-3, 3)
    O.prepare_matrix_pattern()
    u = rng.uniform(-0.03, 0.03, X.size)
    npres = len(conn) if etype == mpf.HEX8P0 else len(X)
    pres = rng.uniform(-0.5, 0.5, npres) if mixed else None
    P.set_state(disp=u, disp_prev=u, pres=pres)
    O.set_state(disp=u, disp_prev=u, pres=pres)
    return X, conn, u, pres, P, O


@pytest.mark.parametrize("etype", [8, 9, 10, 11], ids=["hex8", "hex8p0", "tet10p1", "tet10"])
def test_stress_gauss_and_means_vs_oracle(mpf, orc, gpu, etype):
    X, conn, u, pres, P, O = _problem(mpf, orc, etype)
    g, og = P.stress_gauss(), O.stress_gauss()
    assert g.shape == og.shape
    assert normwise_rel(g, og) < 1e-12
    ev, nv = P.stress_project()
    oev, onv = O.stress_project()
    assert normwise_rel(ev, oev) < 1e-12 and normwise_rel(nv, onv) < 1e-12


def _arrays(path):
    """{(section, name): array} of a .vtu file, ASCII or raw appended."""
    raw = path.read_bytes()
    head, sep, tail = raw.partition(b"<AppendedData encoding=\"raw\">")
    app = None
    if sep:
        app = tail[tail.index(b"_") + 1:]
    xml = head + (b"</VTKFile>\n" if sep else b"")
    root = ET.fromstring(xml)
    out = {}
    dt = {"Float64": np.float64, "Int64": np.int64, "UInt8": np.uint8}
    for sec in ("PointData", "CellData", "Points", "Cells"):
        for node in root.iter(sec):
            for da in node.iter("DataArray"):
                t = dt[da.get("type")]
                if da.get("format") == "appended":
                    off = int(da.get("offset"))
                    nb = int(np.frombuffer(app[off:off + 8], dtype=np.uint64)[0])
                    a = np.frombuffer(app[off + 8:off + 8 + nb], dtype=t)
                else:
                    a = np.array(da.text.split(), dtype=t)
                out[(sec, da.get("Name"))] = a
    time = float(next(root.iter("FieldData")).find("DataArray").text)
    return out, time


@pytest.mark.parametrize("etype", [8, 10], ids=["hex8", "tet10p1"])
def test_write_vtu(mpf, orc, gpu, etype, tmp_path):
    X, conn, u, pres, P, O = _problem(mpf, orc, etype, seed=9)
    P.write_vtu(tmp_path / "a.vtu", time=0.25, binary=False)
    P.write_vtu(tmp_path / "b.vtu", time=0.25, binary=True)
    A, ta = _arrays(tmp_path / "a.vtu")
    B, tb = _arrays(tmp_path / "b.vtu")
    assert ta == tb == 0.25 and A.keys() == B.keys()
    for k in A:
        assert np.array_equal(A[k], B[k]), k
    npe = conn.shape[1]
    assert np.array_equal(A[("Points", "Points")].reshape(-1, 3), X + u.reshape(-1, 3))
    assert np.array_equal(A[("PointData", "disp")], u)
    cells = A[("Cells", "connectivity")].reshape(-1, npe)
    ref = conn.copy()
    if npe == 10:
        ref[:, [8, 9]] = conn[:, [9, 8]]  # tetra10nodemap (femSolidmechanics3.cpp:324-325)
    assert np.array_equal(cells, ref)
    assert np.array_equal(A[("Cells", "offsets")], npe * np.arange(1, len(conn) + 1))
    assert np.all(A[("Cells", "types")] == (12 if npe == 8 else 24))
    oev, onv = O.stress_project()
    assert normwise_rel(A[("PointData", "vonMises")], onv[:, 9]) < 1e-12
    assert normwise_rel(A[("CellData", "stress")].reshape(-1, 9), oev[:, :9]) < 1e-12
    if etype == mpf.TET10P1:
        pv = A[("PointData", "pressure")]
        corners = np.unique(conn[:, :4])
        assert np.array_equal(pv[corners], pres[corners])
        # mid-edge node 4 of element 0 lies on edge (0, 1)
        assert pv[conn[0, 4]] == 0.5 * pres[conn[0, 0]] + 0.5 * pres[conn[0, 1]]
