"""ctypes binding of the CPU oracle (oracle/liboracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, the smoke entry point and the CPU-baseline
leg of bench.py.  The product package never imports this module.
"""
from __future__ import annotations

import ctypes as C
from pathlib import Path

import numpy as np

LIB_PATH = Path(__file__).resolve().parent / "liboracle.so"
_lib = None

_D = C.POINTER(C.c_double)
_I = C.POINTER(C.c_int32)


def lib() -> C.CDLL:
    global _lib
    if _lib is None:
        if not LIB_PATH.exists():
            from mphyfem_amd.build import build_oracle
            build_oracle()
        L = C.CDLL(str(LIB_PATH))
        sig = {
            "orc_gauss3d": (C.c_int, [C.c_int, _D, _D]),
            "orc_gauss_quad4": (C.c_int, [_D, _D]),
            "orc_basis_hex8": (None, [_D, _D, _D]),
            "orc_basis_tet": (None, [C.c_int, _D, _D, _D]),
            "orc_basis_quad4": (None, [_D, _D, _D]),
            "orc_det_inv3": (None, [_D, _D, _D]),
            "orc_dUdJ": (C.c_double, [C.c_int, C.c_double, C.c_double]),
            "orc_d2UdJ2": (C.c_double, [C.c_int, C.c_double, C.c_double]),
            "orc_vol_constants": (None, [C.c_int, C.c_int, C.c_double, C.c_double, _D, _D]),
            "orc_neohooke": (C.c_int, [C.c_int, C.c_int, C.c_double, C.c_double, _D, C.c_double, _D, _D]),
            "orc_elem_disp": (C.c_int, [C.c_int, _D, _D, C.c_double, C.c_double, C.c_int, _D, _D]),
            "orc_elem_disp_consistent": (C.c_int, [C.c_int, _D, _D, C.c_double, C.c_double, C.c_int, _D, _D]),
            "orc_set_tangent": (None, [C.c_void_p, C.c_int]),
            "orc_elem_hex8p0": (C.c_int, [_D, _D, _D, C.c_double, C.c_double, C.c_double, C.c_int,
                                          _D, _D, _D, _D, _D, _D, _D, _D]),
            "orc_elem_tet10p1": (C.c_int, [_D, _D, _D, _D, C.c_double, C.c_double, C.c_int,
                                           _D, _D, _D, _D, _D, _D]),
            "orc_problem_create": (C.c_void_p, [C.c_int, C.c_int, C.c_int, _D, _I, C.c_double, C.c_double, C.c_int]),
            "orc_problem_destroy": (None, [C.c_void_p]),
            "orc_add_dirichlet": (C.c_int, [C.c_void_p, C.c_int, _I, C.c_int, _D]),
            "orc_add_traction": (C.c_int, [C.c_void_p, C.c_int, _I, C.c_double, C.c_double]),
            "orc_add_traction_tri6": (C.c_int, [C.c_void_p, C.c_int, _I, C.c_double, C.c_double]),
            "orc_external_force": (None, [C.c_void_p, C.c_double, _D]),
            "orc_prepare_pattern": (C.c_int, [C.c_void_p]),
            "orc_sizes": (None, [C.c_void_p, _I, _I, C.POINTER(C.c_int64)]),
            "orc_get_pattern": (None, [C.c_void_p, _I, _I, _I, _I]),
            "orc_color": (C.c_int, [C.c_void_p, _I]),
            "orc_set_state": (None, [C.c_void_p, _D, _D, _D, _D]),
            "orc_get_state": (None, [C.c_void_p, _D, _D]),
            "orc_set_bc": (None, [C.c_void_p, C.c_double]),
            "orc_assemble": (C.c_int, [C.c_void_p, C.c_int, C.c_double, _D, _D]),
            "orc_assemble_range": (C.c_int, [C.c_void_p, C.c_int, C.c_int, _D, _D]),
            "orc_assemble_colored": (C.c_int, [C.c_void_p, _I, C.c_int, _D, _D]),
            "orc_pcg": (C.c_int, [C.c_void_p, _D, _D, _D, C.c_int, C.c_double, C.c_int, _D]),
            "orc_gmres": (C.c_int, [C.c_void_p, _D, _D, _D, C.c_int, C.c_int, C.c_double, C.c_int, _D]),
            "orc_direct_solve": (C.c_int, [C.c_void_p, _D, _D, _D]),
            "orc_spmv": (None, [C.c_void_p, _D, _D, _D]),
            "orc_newton": (C.c_int, [C.c_void_p, C.c_int, C.c_double, C.c_int, C.c_int, C.c_double, C.c_double,
                                     C.c_int, _I, _D]),
            "orc_color_priority": (C.c_uint64, [C.c_int64]),
            "orc_stress_gauss": (C.c_int, [C.c_void_p, _D]),
            "orc_arclength": (C.c_int, [C.c_void_p, C.c_double, C.c_double, C.c_int, C.c_int, C.c_double, _D, _D, _D,
                                        _I]),
            "orc_stress_project": (None, [C.c_void_p, _D, _D]),
        }
        for name, (res, args) in sig.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def _d(a):
    return None if a is None else a.ctypes.data_as(_D)


def _i(a):
    return None if a is None else a.ctypes.data_as(_I)


def f64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


def gauss3d(npe: int):
    xi = np.zeros((27, 3))
    w = np.zeros(27)
    n = lib().orc_gauss3d(npe, _d(xi), _d(w))
    return xi[:n], w[:n]


def basis_hex8(xi):
    N = np.zeros(8)
    dN = np.zeros((3, 8))
    lib().orc_basis_hex8(_d(f64(xi)), _d(N), _d(dN))
    return N, dN


def basis_tet(npe: int, xi):
    N = np.zeros(npe)
    dN = np.zeros((3, npe))
    lib().orc_basis_tet(npe, _d(f64(xi)), _d(N), _d(dN))
    return N, dN


def det_inv3(A):
    det = C.c_double()
    Ai = np.zeros((3, 3))
    lib().orc_det_inv3(_d(f64(A)), C.byref(det), _d(Ai))
    return det.value, Ai


def neohooke(F, mu, Kinv, Utype=3, mixed=False, pres=0.0):
    s = np.zeros(9)
    Cm = np.zeros((9, 9))
    lib().orc_neohooke(int(mixed), Utype, Kinv, mu, _d(f64(F)), pres, _d(s), _d(Cm))
    return s, Cm


def elem_disp(X, u, mu, Kinv, Utype=3):
    X, u = f64(X), f64(u)
    npe = X.shape[0]
    Ke = np.zeros((3 * npe, 3 * npe))
    Re = np.zeros(3 * npe)
    st = lib().orc_elem_disp(npe, _d(X), _d(u), mu, Kinv, Utype, _d(Ke), _d(Re))
    return Ke, Re, st


def elem_disp_consistent(X, u, mu, Kinv, Utype=3):
    X, u = f64(X), f64(u)
    npe = X.shape[0]
    Ke = np.zeros((3 * npe, 3 * npe))
    Re = np.zeros(3 * npe)
    st = lib().orc_elem_disp_consistent(npe, _d(X), _d(u), mu, Kinv, Utype, _d(Ke), _d(Re))
    return Ke, Re, st


def elem_hex8p0(X, u, uprev, pe, mu, Kinv, Utype=3):
    """Hex8 Q1/P0: (Kc, Fc, Kuu, Fu, Kup, Kpu, Kpp, Fp, status)."""
    Kc, Kuu = np.zeros((24, 24)), np.zeros((24, 24))
    Fc, Fu, Kup, Kpu = np.zeros(24), np.zeros(24), np.zeros(24), np.zeros(24)
    Kpp, Fp = np.zeros(1), np.zeros(1)
    st = lib().orc_elem_hex8p0(_d(f64(X)), _d(f64(u)), _d(f64(uprev)), float(pe), mu, Kinv, Utype, _d(Kc), _d(Fc),
                               _d(Kuu), _d(Fu), _d(Kup), _d(Kpu), _d(Kpp), _d(Fp))
    return Kc, Fc, Kuu, Fu, Kup, Kpu, float(Kpp[0]), float(Fp[0]), st


def elem_tet10p1(X, u, uprev, pres, mu, Kinv, Utype=3):
    out = [np.zeros((30, 30)), np.zeros((30, 4)), np.zeros((4, 30)), np.zeros((4, 4)), np.zeros(30), np.zeros(4)]
    st = lib().orc_elem_tet10p1(_d(f64(X)), _d(f64(u)), _d(f64(uprev)), _d(f64(pres)), mu, Kinv, Utype,
                                *[_d(o) for o in out])
    return (*out, st)


class Problem:
    """CPU restatement of femSolidmechanics' numbering, pattern, assembly and Newton driver."""

    def __init__(self, elem_type, X, conn, mu, Kinv, Utype=3):
        self.elem_type = elem_type
        self.X = f64(X)
        self.conn = np.ascontiguousarray(conn, dtype=np.int32)
        self.nNode, self.nElem = self.X.shape[0], self.conn.shape[0]
        self.h = lib().orc_problem_create(elem_type, self.nNode, self.nElem, _d(self.X), _i(self.conn), mu, Kinv,
                                          Utype)

    def __del__(self):
        if getattr(self, "h", None):
            lib().orc_problem_destroy(self.h)
            self.h = None

    def set_tangent(self, consistent: bool):
        lib().orc_set_tangent(self.h, int(consistent))

    def add_dirichlet(self, nodes, dof=-1, value=None):
        nodes = np.ascontiguousarray(nodes, dtype=np.int32)
        val = None if value is None else f64(np.broadcast_to(value, nodes.shape))
        lib().orc_add_dirichlet(self.h, len(nodes), _i(nodes), dof, _d(val))

    def add_traction(self, faces, v0, v1=0.0):
        """faces [nf, 4] (Quad4) or [nf, 6] (Tria6)."""
        faces = np.ascontiguousarray(faces, dtype=np.int32)
        f = lib().orc_add_traction_tri6 if faces.shape[1] == 6 else lib().orc_add_traction
        f(self.h, faces.shape[0], _i(faces), v0, v1)

    def prepare_matrix_pattern(self):
        lib().orc_prepare_pattern(self.h)

    def external_force(self, lam=1.0):
        """The traction patches' external force [nNode * 3] at load factor lam."""
        F = np.zeros(3 * self.nNode)
        lib().orc_external_force(self.h, lam, _d(F))
        return F

    def sizes(self):
        d, p = C.c_int32(), C.c_int32()
        nnz = C.c_int64()
        lib().orc_sizes(self.h, C.byref(d), C.byref(p), C.byref(nnz))
        return d.value, p.value, nnz.value

    def pattern(self):
        d, p, nnz = self.sizes()
        rowptr = np.zeros(d + p + 1, dtype=np.int32)
        colidx = np.zeros(nnz, dtype=np.int32)
        nda = np.zeros((self.nNode, 3), dtype=np.int32)
        idp = np.zeros(self.nNode, dtype=np.int32)
        lib().orc_get_pattern(self.h, _i(rowptr), _i(colidx), _i(nda), _i(idp))
        return rowptr, colidx, nda, idp

    def color(self):
        c = np.zeros(self.nElem, dtype=np.int32)
        n = lib().orc_color(self.h, _i(c))
        return c, n

    def set_state(self, disp=None, disp_prev=None, pres=None, pres_prev=None):
        a = [None if x is None else f64(x) for x in (disp, disp_prev, pres, pres_prev)]
        lib().orc_set_state(self.h, *[_d(x) for x in a])

    def state(self):
        """(disp [nNode*3], pres [nNode], or [nElem] for elem_type 9, Hex8 Q1/P0)."""
        disp = np.zeros(self.nNode * 3)
        pres = np.zeros(self.nElem if self.elem_type == 9 else self.nNode)
        lib().orc_get_state(self.h, _d(disp), _d(pres))
        return disp, pres

    def set_boundary_conditions(self, lam):
        lib().orc_set_bc(self.h, lam)

    def assemble(self, first_iter, lam):
        d, p, nnz = self.sizes()
        vals = np.zeros(nnz)
        rhs = np.zeros(d + p)
        st = lib().orc_assemble(self.h, int(first_iter), lam, _d(vals), _d(rhs))
        return vals, rhs, st

    def assemble_range(self, first, count, vals, rhs):
        return lib().orc_assemble_range(self.h, first, count, _d(vals), _d(rhs))

    def assemble_colored(self, color, nthreads, vals, rhs):
        """All elements colour by colour on nthreads OpenMP threads (timing; sums in another order)."""
        return lib().orc_assemble_colored(self.h, _i(np.ascontiguousarray(color, dtype=np.int32)), nthreads,
                                          _d(vals), _d(rhs))

    def pcg(self, vals, b, precond=2, rtol=1e-10, maxit=100000):
        x = np.zeros_like(b)
        rr = C.c_double()
        it = lib().orc_pcg(self.h, _d(vals), _d(f64(b)), _d(x), precond, rtol, maxit, C.byref(rr))
        return x, it, rr.value

    def gmres(self, vals, b, precond=2, restart=50, rtol=1e-10, maxit=100000):
        x = np.zeros_like(b)
        rr = C.c_double()
        it = lib().orc_gmres(self.h, _d(vals), _d(f64(b)), _d(x), precond, restart, rtol, maxit, C.byref(rr))
        return x, it, rr.value

    def direct_solve(self, vals, b):
        x = np.zeros_like(b)
        st = lib().orc_direct_solve(self.h, _d(vals), _d(f64(b)), _d(x))
        return x, st

    def spmv(self, vals, x):
        y = np.zeros_like(x)
        lib().orc_spmv(self.h, _d(vals), _d(f64(x)), _d(y))
        return y

    def arclength(self, dt, final_time, max_steps, max_iters, tol):
        """(load factors, arc lengths, displacements [k, nNode*3], attempts) of the k converged steps;
        the state is the last step's."""
        lam = np.zeros(max_steps)
        ds = np.zeros(max_steps)
        us = np.zeros((max_steps, 3 * self.nNode))
        att = np.zeros(1, dtype=np.int32)
        k = lib().orc_arclength(self.h, dt, final_time, max_steps, max_iters, tol, _d(lam), _d(ds), _d(us), _i(att))
        if k < 0:
            raise RuntimeError("orc_arclength: a solve failed")
        return lam[:k], ds[:k], us[:k], int(att[0])

    def stress_gauss(self):
        """[nElem, nGP, 11]: sigma (9, column-major map), von Mises, tr(sigma)/3 at the Gauss points."""
        ngp = 8 if self.elem_type in (8, 9) else 4
        out = np.zeros((self.nElem, ngp, 11))
        lib().orc_stress_gauss(self.h, _d(out))
        return out

    def stress_project(self):
        """(element means [nElem, 11], nodal means over adjacent elements [nNode, 11])."""
        ev = np.zeros((self.nElem, 11))
        nv = np.zeros((self.nNode, 11))
        lib().orc_stress_project(self.h, _d(ev), _d(nv))
        return ev, nv

    def newton(self, nsteps, dt, solver=0, precond=2, lin_rtol=1e-12, rtol=1e-9, itmax=30):
        iters = np.zeros(nsteps, dtype=np.int32)
        rel = C.c_double()
        st = lib().orc_newton(self.h, nsteps, dt, solver, precond, lin_rtol, rtol, itmax, _i(iters), C.byref(rel))
        return st, iters, rel.value


def color_priority(e: int) -> int:
    return int(lib().orc_color_priority(e))
