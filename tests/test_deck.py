"""Input decks on the CPU (csrc/deck.cpp; no device call): the reference's own sample deck and mesh
(cook's membrane, 2D: sample-inputs/solidmechanics/cooksmembrane2d/inputs, kept as data under
tests/golden/decks/cooks2d) parses block by block and is then refused for the 3D path; generated 3D
Hex8 and Tet10 decks give the mesh, the patches' node sets and the solver settings the writer put in."""
from __future__ import annotations

import shutil
from pathlib import Path

import numpy as np
import pytest

from tests.cases import tet10_boundary_tris, write_config, write_gmsh22

GOLD = Path(__file__).resolve().parent / "golden" / "decks"


def test_reference_sample_deck_parses(mpf, tmp_path):
    # the config names a mesh "cookmembrane2d-Q1-nelem2"; the reference ships the nelem8 file
    for f in (GOLD / "cooks2d").iterdir():
        shutil.copy(f, tmp_path / f.name)
    (tmp_path / "cookmembrane2d-Q1-nelem8.msh").rename(tmp_path / "cookmembrane2d-Q1-nelem2.msh")
    D = mpf.Deck(tmp_path / "config")
    i = D.info()
    assert (i.ndim, i.nNode, i.npe, i.gmsh_type) == (2, 81, 4, 3)
    assert i.nElem == 64  # quads of the "solid" domain: 80 records minus 16 boundary lines
    assert (i.nmat, i.ntf, i.nbc, i.ntrac) == (4, 1, 1, 1)
    assert (i.max_iters, i.max_steps) == (2, 1) and i.tol == 1e-7 and i.dt == 1.0 and i.final_time == 1.0
    # petsc_options.dat: preonly + lu (MUMPS) -> the direct solve, rtol 1e-6, 1000 iterations
    assert i.lin_from_file == 1 and i.lin.ksp == mpf.KSP_BAND and i.lin.rtol == 1e-6 and i.lin.maxit == 1000
    P = D.patches()
    assert set(P) == {"leftedge", "rightedge"}
    assert all(dim == 1 for dim, _ in P.values())
    assert len(P["leftedge"][1]) == 9 and len(P["rightedge"][1]) == 9
    with pytest.raises(mpf.MpfError):  # 2D: refused before any device work
        D.setup()
    D.close()


def test_unknown_block_is_refused(mpf, tmp_path):
    X, conn = mpf.mesh_hex(1, 1, 1, jitter=0.0, seed=1)
    write_gmsh22(tmp_path / "m.msh", X, conn, 5, "block", {})
    (tmp_path / "config").write_text("Files\n{\n  mesh : m\n}\n\nNot A Block\n{\n}\n")
    with pytest.raises(mpf.MpfError):
        mpf.Deck(tmp_path / "config")


@pytest.mark.parametrize("kind", ["hex8", "tet10p1"])
def test_generated_3d_deck_reads_back(mpf, tmp_path, kind):
    if kind == "hex8":
        n = 3
        X, conn = mpf.mesh_hex(n, n, n, jitter=0.05, seed=2)
        q0 = mpf.hex_face_quads(n, n, n, 0)
        q1 = mpf.hex_face_quads(n, n, n, 1)
        patches = {"left": (2, 3, q0), "right": (2, 3, q1)}
        write_gmsh22(tmp_path / "blk.msh", X, conn, 5, "block", patches)
        write_config(tmp_path / "config", "blk", "ELEM_SOLID_3D_DISP", 384.6, 1.2e-3,
                     [("left", "fixed"), ("right", "specified", "UX", 0.1, 1)], [], 0.25, 1.0, 1e-9,
                     petsc=["-ksp_type cg", "-pc_type jacobi", "-ksp_rtol 1e-12"])
        faces = {"left": q0, "right": q1}
    else:
        X, conn = mpf.mesh_tet10(2, jitter=0.1, seed=3)
        t0 = tet10_boundary_tris(X, conn, 2, 0.0, -1.0)
        t1 = tet10_boundary_tris(X, conn, 2, 1.0, 1.0)
        patches = {"bottom": (2, 9, t0), "top": (2, 9, t1)}
        write_gmsh22(tmp_path / "blk.msh", X, conn, 11, "block", patches)
        write_config(tmp_path / "config", "blk", "ELEM_SOLID_3D_MIXED1", 80.0, 1e-4, [("bottom", "fixed")],
                     [("top", -2.0, 0.0, 1)], 0.5, 1.0, 1e-8)
        faces = {"bottom": t0, "top": t1}
    D = mpf.Deck(tmp_path / "config")
    i = D.info()
    assert (i.ndim, i.nNode, i.nElem, i.npe) == (3, len(X), len(conn), conn.shape[1])
    assert i.nbc == (2 if kind == "hex8" else 1) and i.ntrac == (0 if kind == "hex8" else 1)
    if kind == "hex8":
        assert i.lin_from_file and i.lin.ksp == mpf.KSP_CG and i.lin.pc == mpf.PC_JACOBI and i.lin.rtol == 1e-12
    else:
        assert not i.lin_from_file and i.lin.ksp == mpf.KSP_CG and i.lin.pc == mpf.PC_BJACOBI3
    P = D.patches()
    for name, f in faces.items():
        assert np.array_equal(P[name][1], np.unique(f))
    D.close()


@pytest.mark.parametrize("pc, want", [("gamg", "PC_PMG"), ("mg", "PC_PMG"), ("fieldsplit", "PC_PMG"),
                                      ("ilu", "PC_BJACOBI3"), ("jacobi", "PC_JACOBI"), ("none", "PC_NONE")])
def test_petsc_pc_type_mapping(mpf, tmp_path, pc, want):
    """-pc_type in petsc_options.dat: multigrid and field-split preconditioners map to the two-level
    p-multigrid (which mpf_deck_setup keeps for Tet10 meshes only, inside flexible GMRES), the
    incomplete factorisations and block methods to the 3x3 node-block Jacobi."""
    X, conn = mpf.mesh_hex(2, 2, 2, jitter=0.0, seed=1)
    write_gmsh22(tmp_path / "blk.msh", X, conn, 5, "block", {"left": (2, 3, mpf.hex_face_quads(2, 2, 2, 0))})
    write_config(tmp_path / "config", "blk", "ELEM_SOLID_3D_DISP", 1.0, 1e-3, [("left", "fixed")], [], 1.0, 1.0,
                 1e-9, petsc=["-ksp_type gmres", f"-pc_type {pc}"])
    D = mpf.Deck(tmp_path / "config")
    assert D.info().lin.pc == getattr(mpf, want) and D.info().lin.ksp == mpf.KSP_GMRES
    D.close()
