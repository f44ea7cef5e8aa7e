"""Shared problem builders for the parity tests (same seeded inputs for oracle and device)."""
from __future__ import annotations

from pathlib import Path

import numpy as np

import mphyfem_amd as M


def cube_hex(n: int, jitter: float = 0.05, seed: int = 1):
    """Unit cube of n^3 Hex8 elements (the structured block of the CPU-runnable workload)."""
    return M.mesh_hex(n, n, n, jitter=jitter, seed=seed)


def hex_stretch_bcs(n: int):
    """x=0 face fixed, x=1 face given UX = 0.1 * lambda."""
    return [(M.face_nodes(M.HEX8, n, n, n, 0), -1, None), (M.face_nodes(M.HEX8, n, n, n, 1), 0, 0.1)]


def tet_compress_bcs(n: int, load: float = -0.05):
    """z=0 face fixed, z=1 face given UZ = load * lambda."""
    return [(M.face_nodes(M.TET10P1, n, n, n, 4), -1, None), (M.face_nodes(M.TET10P1, n, n, n, 5), 2, load)]


def apply_bcs(prob, bcs):
    for nodes, dof, val in bcs:
        prob.add_dirichlet(nodes, dof, val)


def random_state(nNode: int, scale: float, seed: int):
    rng = np.random.default_rng(seed)
    return rng.uniform(-scale, scale, nNode * 3)


def normwise_rel(a, b) -> float:
    """max |a - b| / max |b| (normwise relative error; entrywise is meaningless near zero entries)."""
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    den = np.abs(b).max()
    return float(np.abs(a - b).max() / (den if den > 0 else 1.0))


def tet10_fan(subdiv: int = 1, radius: float = 1.0):
    """Tet10 mesh of 20 * 4^subdiv tetrahedra sharing one central vertex: the faces of a subdivided
    icosahedron joined to its centre (80 tets around one node for subdiv = 1).  Mid-edge nodes at
    the edge midpoints, Tet10 local order 4:(0,1) 5:(1,2) 6:(0,2) 7:(0,3) 8:(2,3) 9:(1,3).
    Returns X, conn (int32) and the indices of the nodes on the outer sphere."""
    t = (1.0 + 5.0 ** 0.5) / 2.0
    V = [(-1, t, 0), (1, t, 0), (-1, -t, 0), (1, -t, 0), (0, -1, t), (0, 1, t), (0, -1, -t), (0, 1, -t),
         (t, 0, -1), (t, 0, 1), (-t, 0, -1), (-t, 0, 1)]
    V = [np.array(v, dtype=float) / np.linalg.norm(v) for v in V]
    F = [(0, 11, 5), (0, 5, 1), (0, 1, 7), (0, 7, 10), (0, 10, 11), (1, 5, 9), (5, 11, 4), (11, 10, 2), (10, 7, 6),
         (7, 1, 8), (3, 9, 4), (3, 4, 2), (3, 2, 6), (3, 6, 8), (3, 8, 9), (4, 9, 5), (2, 4, 11), (6, 2, 10),
         (8, 6, 7), (9, 8, 1)]
    for _ in range(subdiv):
        mids = {}

        def mid(i, j):
            k = (min(i, j), max(i, j))
            if k not in mids:
                m = V[i] + V[j]
                V.append(m / np.linalg.norm(m))
                mids[k] = len(V) - 1
            return mids[k]
        F2 = []
        for a, b, c in F:
            ab, bc, ca = mid(a, b), mid(b, c), mid(c, a)
            F2 += [(a, ab, ca), (b, bc, ab), (c, ca, bc), (ab, bc, ca)]
        F = F2
    X = [np.zeros(3)] + [radius * v for v in V]  # node 0 is the centre
    outer = list(range(1, len(X)))
    edge_node = {}

    def enode(i, j):
        k = (min(i, j), max(i, j))
        if k not in edge_node:
            X.append(0.5 * (X[i] + X[j]))
            edge_node[k] = len(X) - 1
            if i != 0 and j != 0:
                outer.append(len(X) - 1)
        return edge_node[k]
    conn = []
    for a, b, c in F:
        v = [0, a + 1, b + 1, c + 1]
        if np.linalg.det(np.array([X[v[1]] - X[v[0]], X[v[2]] - X[v[0]], X[v[3]] - X[v[0]]])) < 0:
            v[2], v[3] = v[3], v[2]
        e = [(0, 1), (1, 2), (0, 2), (0, 3), (2, 3), (1, 3)]
        conn.append(v + [enode(v[p], v[q]) for p, q in e])
    return np.array(X), np.array(conn, dtype=np.int32), np.array(outer, dtype=np.int32)


def q1p0_newton(orc, X, conn, bcs, mu, Kinv, nsteps, dt, rtol=1e-10, itmax=20, Utype=3, fext=None):
    """CPU Newton driver for Hex8 Q1/P0 on a small mesh (dense global system), built on the oracle's
    condensed element (orc.elem_hex8p0) and following the reference's order of operations:
    saveSolution, setBoundaryConditions (dispApplied = value * lambda - disp on prescribed dofs,
    femSolidmechanics1.cpp:864-890), calcStiffnessAndResidual with the first-iteration lift
    (ElementBase.cpp:567-594), ||R|| / ||R0|| < rtol checked after each assembly, then the update
    (femSolidmechanics2.cpp:1137-1186) and solveForPressure per element
    (Elem_Solid_3D_Mixed0.cpp:112-199).  fext: external force on every dof for lambda = 1, added as
    lambda * fext (femSolidmechanics2.cpp:785-789).  Returns (disp [nNode*3], element pressures
    [nElem], Newton iterations per step)."""
    X = np.asarray(X, dtype=np.float64)
    conn = np.asarray(conn)
    nn, ne = len(X), len(conn)
    spec = {}
    for nodes, dof, val in bcs:
        vals = np.broadcast_to(0.0 if val is None else val, np.shape(nodes))
        for nd, v in zip(np.asarray(nodes).tolist(), vals.tolist()):
            for c in (range(3) if dof < 0 else (dof,)):
                spec[3 * nd + c] = 0.0 if dof < 0 else v
    presc = np.array(sorted(spec), dtype=np.int64)
    free = np.setdiff1d(np.arange(3 * nn), presc)
    pval = np.array([spec[d] for d in presc.tolist()])
    u, p = np.zeros(3 * nn), np.zeros(ne)
    iters = []
    for step in range(1, nsteps + 1):
        lam = step * dt
        uprev = u.copy()
        applied = pval * lam - u[presc]
        R0 = -1.0
        for it in range(1, itmax + 1):
            first = it == 1
            K, F = np.zeros((3 * nn, 3 * nn)), np.zeros(3 * nn)
            Kup, Kpp, Fp = np.zeros((ne, 24)), np.zeros(ne), np.zeros(ne)
            for e, c in enumerate(conn):
                Kc, Fc, _, _, kup, kpu, kpp, fp, st = orc.elem_hex8p0(X[c], u.reshape(-1, 3)[c],
                                                                     uprev.reshape(-1, 3)[c], p[e], mu, Kinv, Utype)
                assert st == 0
                d = (3 * np.asarray(c)[:, None] + np.arange(3)).ravel()
                K[np.ix_(d, d)] += Kc
                F[d] += Fc
                Kup[e], Kpp[e], Fp[e] = kpu, kpp, fp
            if fext is not None:
                F += lam * fext
            R = F[free]
            if first:
                R = R - K[np.ix_(free, presc)] @ applied
            rn = np.linalg.norm(R)
            if R0 < 0:
                R0 = rn
            if R0 == 0.0 or rn / R0 < rtol:
                break
            inc = np.zeros(3 * nn)
            inc[free] = np.linalg.solve(K[np.ix_(free, free)], R)
            if first:
                inc[presc] = applied
            u += inc
            for e, c in enumerate(conn):
                d = (3 * np.asarray(c)[:, None] + np.arange(3)).ravel()
                p[e] += (Fp[e] - Kup[e] @ inc[d]) / Kpp[e]
        else:
            raise RuntimeError(f"Q1/P0 Newton did not converge at step {step}")
        iters.append(it)
    return u, p, iters


def oracle_fext(orc, X, conn, bcs, faces, v0, v1=0.0):
    """External force of a Quad4 dead-load traction patch for lambda = 1 on every dof (0 on prescribed
    ones), from the oracle's Hex8 problem assembled at the stress-free state, where the internal force
    vanishes and the right-hand side is the external force alone."""
    m = M.matl_from_E_nu(1e3, 0.3)
    O = orc.Problem(8, X, conn, m.mu, m.Kinv, 3)
    for nodes, dof, val in bcs:
        O.add_dirichlet(nodes, dof, val)
    O.add_traction(faces, v0, v1)
    O.prepare_matrix_pattern()
    _, rhs, st = O.assemble(False, 1.0)
    assert st == 0
    _, _, nda, _ = O.pattern()
    f = np.zeros(3 * len(X))
    d = nda.ravel()
    f[d >= 0] = rhs[d[d >= 0]]
    return f


def tet10_boundary_tris(X, conn, axis: int, value: float, outward: float, tol: float = 1e-9):
    """Tria6 boundary faces of a Tet10 mesh on the plane X[axis] == value, ordered so that the
    reference's normal dx/dxi2 x dx/dxi1 points along outward * e_axis: corners (c0, c1, c2), then
    the mid-edge nodes of (c0,c1), (c1,c2), (c2,c0)."""
    X = np.asarray(X)
    edges = {(0, 1): 4, (1, 2): 5, (0, 2): 6, (0, 3): 7, (2, 3): 8, (1, 3): 9}
    tris = []
    for e in conn:
        on = [a for a in range(4) if abs(X[e[a], axis] - value) < tol]
        if len(on) != 3:
            continue
        c = list(on)
        P = X[e[c]]
        n = np.cross(P[2] - P[0], P[1] - P[0])  # dv x du for du = P1 - P0, dv = P2 - P0
        if n[axis] * outward < 0:
            c = [c[0], c[2], c[1]]
        mid = [e[edges[tuple(sorted((c[i], c[(i + 1) % 3])))]] for i in range(3)]
        tris.append([e[c[0]], e[c[1]], e[c[2]]] + mid)
    return np.asarray(tris, dtype=np.int32)


def write_gmsh22(path, X, conn, gmsh_type, domain, patches):
    """A Gmsh 2.2 ASCII mesh as femBase::readInputGMSH reads it (femBase.cpp:121-364): physical names
    (the domain with dim 3, each patch with its dim), nodes (1-based), boundary elements with the
    patch's physical tag, then the domain elements.  patches: {name: (dim, gmsh_type, faces [nf, k])}."""
    lines = ["$MeshFormat", "2.2 0 8", "$EndMeshFormat", "$PhysicalNames", str(len(patches) + 1)]
    tags = {}
    for k, (name, (dim, _, _)) in enumerate(patches.items()):
        tags[name] = k + 1
        lines.append(f'{dim} {k + 1} "{name}"')
    dom_tag = len(patches) + 1
    lines.append(f'3 {dom_tag} "{domain}"')
    lines += ["$EndPhysicalNames", "$Nodes", str(len(X))]
    lines += [f"{i + 1} {x[0]:.17g} {x[1]:.17g} {x[2]:.17g}" for i, x in enumerate(np.asarray(X))]
    lines += ["$EndNodes", "$Elements"]
    body = []
    for name, (dim, ft, faces) in patches.items():
        for f in np.asarray(faces):
            body.append((ft, tags[name], f))
    for c in np.asarray(conn):
        body.append((gmsh_type, dom_tag, c))
    lines.append(str(len(body)))
    for i, (t, tag, nodes) in enumerate(body):
        lines.append(f"{i + 1} {t} 2 {tag} {tag} " + " ".join(str(int(v) + 1) for v in nodes))
    lines.append("$EndElements")
    Path(path).write_text("\n".join(lines) + "\n")


def write_config(path, mesh, elem_type, mu, Kinv, bcs, tractions, dt, final_time, tol, max_iters=20,
                 max_steps=1000, Utype=3, petsc=None, solvertype="newton",
                 timefunction="   1   0.0   1000.0  0.0  1.0  0.0  0.0  0.0  0.0  0.0  0.0"):
    """A config deck in the reference's block format (doc/examples/SolidMechanics/Example4):
    bcs: [(patch, 'fixed') | (patch, 'specified', 'UX', value, timefunction or None)],
    tractions: [(patch, v0, v1, timefunction)].  Time function 1 is lam(t) = t unless given."""
    out = ["Files", "{", f"  mesh : {mesh}", "}", "", "Model Type", "{", "  dimension : 3", "}", "",
           "Domains", "{", "  block", "  {", "    Material : 1", "    Element  : 1", "  }", "}", "",
           "Material", "{", "  id : 1", "  name : Matl_NeoHookean", "  density : 1.0",
           f"  data_deviatoric : {mu:.17g}", f"  data_volumetric : {Utype} {Kinv:.17g}", "}", "",
           "Element", "{", "  id : 1", f"  type : {elem_type}", "}", "", "Boundary Conditions", "{"]
    for bc in bcs:
        out += [f"    {bc[0]}", "    {", f"        type : {bc[1]}"]
        if bc[1] == "specified":
            out += [f"        dof : {bc[2]}", f"        value : {bc[3]:.17g}"]
            if bc[4] is not None:
                out.append(f"        timefunction : {bc[4]}")
        out.append("    }")
    out += ["}", "", "Tractions", "{"]
    for name, v0, v1, tf in tractions:
        out += [f"    {name}", "    {", "        type : traction", f"        value : {v0:.17g} {v1:.17g} 0.0",
                f"        timefunction : {tf}", "    }"]
    out += ["}", "", "Time Functions", "{", "! id t0 t1 p1 p2 p3 p4 p5 p6 p7 p8",
            timefunction, "}", "",
            "Solver", "{", f"solvertype : {solvertype}", "timescheme : STEADY", f"finalTime : {final_time:.17g}",
            f"timeStep : {dt:.17g}", f"maximumSteps : {max_steps}", f"maximumIterations : {max_iters}",
            f"tolerance : {tol:.17g}", "debug : 0", "}"]
    Path(path).write_text("\n".join(out) + "\n")
    if petsc is not None:
        (Path(path).parent / "petsc_options.dat").write_text("\n".join(petsc) + "\n")
