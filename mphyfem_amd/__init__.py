"""mphyfem_amd -- MI355X (gfx950) femSolidmechanics assembly and Newton-Krylov path.

Python view of the C ABI in include/mpf.h (libmpf.so, built in-tree by
``python -m mphyfem_amd.build``).  The names mirror the reference's API
(chennachaos/mphyfem src/femSolidmechanics/femSolidmechanics.h,
src/utilities/ElementBase.h, src/mySolvers/SolverPetsc.h).

There is no CPU fallback: importing works without a GPU (the library loads and
exports its symbols), but every compute call raises ``MpfError`` when no HIP
device is present, and ``lib()`` raises when libmpf.so has not been built.
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass
from pathlib import Path

import numpy as np

_PKG = Path(__file__).resolve().parent
LIB_PATH = _PKG / "libmpf.so"

HEX8 = 8
HEX8P0 = 9  # Hex8 Q1/P0: one element pressure, condensed
TET10P1 = 10
TET10 = 11  # Tet10 displacement element (Elem_Solid_3D_Disp with npElem = 10)
KSP_CG, KSP_GMRES, KSP_BAND = 0, 1, 2
PC_NONE, PC_JACOBI, PC_BJACOBI3, PC_PMG = 0, 1, 2, 3
ASSEMBLY_GATHER, ASSEMBLY_COLORED_SCATTER, ASSEMBLY_GATHER_DIRECT = 0, 1, 2
NOALLOC_MAX_DOFS = 524288  # MPF_NOALLOC_MAX_DOFS

OK = 0
ERR_ARG, ERR_HIP, ERR_NEG_JACOBIAN, ERR_STATE, ERR_NOT_CONVERGED, ERR_DIVERGED, ERR_NO_DEVICE = range(1, 8)


class MpfError(RuntimeError):
    def __init__(self, status: int, what: str):
        self.status = status
        super().__init__(f"{what}: {status_string(status)} (status {status})")


class NegativeJacobian(MpfError):
    """runtime_error("Negative Jacobian in the element") of the reference (Elem_Solid_3D_Disp.cpp:277)."""


TANGENT_REFERENCE, TANGENT_CONSISTENT = 0, 1


class Matl(C.Structure):
    _fields_ = [("mu", C.c_double), ("Kinv", C.c_double), ("Utype", C.c_int32), ("mixed", C.c_int32),
                ("tangent", C.c_int32), ("pad_", C.c_int32)]


class KrylovOpts(C.Structure):
    _fields_ = [("ksp", C.c_int32), ("pc", C.c_int32), ("restart", C.c_int32), ("maxit", C.c_int32),
                ("rtol", C.c_double)]


class PmgReport(C.Structure):
    _fields_ = [("coarse_dofs", C.c_int64), ("coarse_nnz", C.c_int64), ("vcycles", C.c_int64),
                ("coarse_iterations", C.c_int64), ("setup_seconds", C.c_double), ("lambda_max", C.c_double),
                ("lambda_vertex", C.c_double), ("agg_dofs", C.c_int64), ("agg_nnz", C.c_int64),
                ("levels", C.c_int32), ("degree", C.c_int32)]


class KrylovReport(C.Structure):
    _fields_ = [("iterations", C.c_int32), ("converged", C.c_int32), ("relres", C.c_double),
                ("seconds", C.c_double)]


class NewtonOpts(C.Structure):
    _fields_ = [("nsteps", C.c_int32), ("dt", C.c_double), ("itmax", C.c_int32), ("rtol", C.c_double),
                ("lin", KrylovOpts), ("forcing", C.c_int32)]


class NewtonReport(C.Structure):
    _fields_ = [("steps_done", C.c_int32), ("newton_iters", C.c_int32), ("krylov_iters", C.c_int32),
                ("final_relres", C.c_double), ("seconds", C.c_double), ("seconds_assembly", C.c_double),
                ("seconds_solve", C.c_double)]


class BatchOpts(C.Structure):
    _fields_ = [("nsteps", C.c_int32), ("dt", C.c_double), ("itmax", C.c_int32), ("rtol", C.c_double),
                ("cg_maxit", C.c_int32), ("cg_rtol", C.c_double), ("use_graph", C.c_int32), ("chunk", C.c_int32),
                ("tile", C.c_int32), ("lanes_per_row", C.c_int32), ("solver", C.c_int32)]


BATCH_SOLVER_PCG, BATCH_SOLVER_BAND = 0, 1


class BatchReport(C.Structure):
    _fields_ = [("steps_done", C.c_int32), ("newton_iters", C.c_int32), ("cg_iters_total", C.c_int64),
                ("cg_iters_max", C.c_int64), ("graph_launches", C.c_int32), ("tile", C.c_int32),
                ("lanes_per_row", C.c_int32), ("chunk", C.c_int32), ("seconds", C.c_double),
                ("seconds_assembly", C.c_double), ("seconds_solve", C.c_double),
                ("start_relres_mean", C.c_double), ("start_relres_max", C.c_double),
                ("iteration_graphs", C.c_int32), ("chunk_graphs", C.c_int32), ("seconds_capture", C.c_double)]


_P = C.c_void_p
_I32P = C.POINTER(C.c_int32)
_I64P = C.POINTER(C.c_int64)
_DP = C.POINTER(C.c_double)

class ArclengthOpts(C.Structure):
    _fields_ = [("dt", C.c_double), ("final_time", C.c_double), ("max_steps", C.c_int32), ("max_iters", C.c_int32),
                ("tol", C.c_double), ("lin", KrylovOpts)]


class ArclengthReport(C.Structure):
    _fields_ = [("steps_done", C.c_int32), ("attempts", C.c_int32), ("newton_iters", C.c_int32),
                ("krylov_iters", C.c_int32), ("final_lam", C.c_double), ("final_ds", C.c_double),
                ("seconds", C.c_double)]


class DeckInfo(C.Structure):
    _fields_ = [("ndim", C.c_int32), ("nNode", C.c_int32), ("nElem", C.c_int32), ("npe", C.c_int32),
                ("gmsh_type", C.c_int32), ("elem_type", C.c_int32), ("npatch", C.c_int32), ("nbc", C.c_int32),
                ("ntrac", C.c_int32), ("nmat", C.c_int32), ("ntf", C.c_int32), ("arclength", C.c_int32),
                ("dt", C.c_double), ("dt_min", C.c_double), ("dt_max", C.c_double), ("final_time", C.c_double),
                ("max_steps", C.c_int32), ("max_iters", C.c_int32), ("tol", C.c_double), ("lin", KrylovOpts),
                ("lin_from_file", C.c_int32), ("pad_", C.c_int32)]


class DeckReport(C.Structure):
    _fields_ = [("steps_done", C.c_int32), ("newton_iters", C.c_int32), ("krylov_iters", C.c_int32),
                ("cuts", C.c_int32), ("converged", C.c_int32), ("pad_", C.c_int32), ("final_time", C.c_double),
                ("final_lam", C.c_double), ("final_norm", C.c_double), ("seconds", C.c_double)]


# (name, restype, argtypes) for every symbol of include/mpf.h
SIGNATURES = [
    ("mpf_status_string", C.c_char_p, [C.c_int32]),
    ("mpf_device_count", C.c_int32, []),
    ("mpf_neohooke_stress_tangent_batch", C.c_int, [_P, _P, C.POINTER(Matl), C.c_int32, _P, _P, _P]),
    ("mpf_elem_hex8_nh_batch", C.c_int, [_P, _P, _P, C.POINTER(Matl), C.c_int32, C.c_int32, _P, _P, _P, _P]),
    ("mpf_elem_tet10_nh_batch", C.c_int, [_P, _P, _P, C.POINTER(Matl), C.c_int32, C.c_int32, _P, _P, _P, _P]),
    ("mpf_elem_hex8p0_nh_batch", C.c_int, [_P, _P, _P, _P, _P, C.POINTER(Matl), C.c_int32, C.c_int32,
                                            _P, _P, _P, _P, _P, _P, _P]),
    ("mpf_elem_tet10p1_nh_batch", C.c_int, [_P, _P, _P, _P, _P, C.POINTER(Matl), C.c_int32, C.c_int32,
                                             _P, _P, _P, _P, _P, _P, _P, _P]),
    ("mpf_problem_create", C.c_int, [C.POINTER(_P), C.c_int32, C.c_int32, C.c_int32, _DP, _I32P,
                                     C.POINTER(Matl)]),
    ("mpf_problem_destroy", C.c_int, [_P]),
    ("mpf_add_dirichlet", C.c_int, [_P, C.c_int32, _I32P, C.c_int32, _DP]),
    ("mpf_add_traction", C.c_int, [_P, C.c_int32, _I32P, C.c_double, C.c_double]),
    ("mpf_prepare_pattern", C.c_int, [_P]),
    ("mpf_get_sizes", C.c_int, [_P, _I32P, _I32P, _I64P, _I32P]),
    ("mpf_get_pattern", C.c_int, [_P, _I32P, _I32P, _I32P, _I32P, _I32P]),
    ("mpf_set_state", C.c_int, [_P, _DP, _DP, _DP, _DP]),
    ("mpf_get_state", C.c_int, [_P, _DP, _DP]),
    ("mpf_set_bc", C.c_int, [_P, C.c_double]),
    ("mpf_assemble", C.c_int, [_P, C.c_int32, C.c_double]),
    ("mpf_system_device_ptrs", C.c_int, [_P, C.POINTER(_P), C.POINTER(_P), C.POINTER(_P), C.POINTER(_P)]),
    ("mpf_get_system", C.c_int, [_P, _DP, _DP]),
    ("mpf_rhs_norm", C.c_int, [_P, _DP]),
    ("mpf_solve", C.c_int, [_P, C.POINTER(KrylovOpts), C.POINTER(KrylovReport)]),
    ("mpf_get_solution", C.c_int, [_P, _DP]),
    ("mpf_solve_noalloc", C.c_int, [_P, C.POINTER(KrylovOpts), C.POINTER(KrylovReport)]),
    ("mpf_add_traction_tri6", C.c_int, [_P, C.c_int32, _I32P, C.c_double, C.c_double]),
    ("mpf_arclength_solve", C.c_int, [_P, C.POINTER(ArclengthOpts), _DP, _DP, C.POINTER(ArclengthReport)]),
    ("mpf_deck_load", C.c_int, [C.c_char_p, C.POINTER(_P)]),
    ("mpf_deck_destroy", C.c_int, [_P]),
    ("mpf_deck_get_info", C.c_int, [_P, C.POINTER(DeckInfo)]),
    ("mpf_deck_patch", C.c_int, [_P, C.c_int32, C.c_char_p, C.c_int32, _I32P, _I32P, _I32P]),
    ("mpf_deck_setup", C.c_int, [_P, C.POINTER(_P)]),
    ("mpf_deck_run", C.c_int, [_P, C.c_char_p, C.POINTER(DeckReport)]),
    ("mpf_stress_gauss", C.c_int, [_P, _DP]),
    ("mpf_stress_project", C.c_int, [_P, _DP, _DP]),
    ("mpf_write_vtu", C.c_int, [_P, C.c_char_p, C.c_double, C.c_int32]),
    ("mpf_update", C.c_int, [_P, C.c_int32]),
    ("mpf_newton_solve", C.c_int, [_P, C.POINTER(NewtonOpts), _I32P, C.POINTER(NewtonReport)]),
    ("mpf_newton_step", C.c_int, [_P, C.c_double, C.POINTER(NewtonOpts), _DP, _I32P, C.POINTER(NewtonReport)]),
    ("mpf_time_assembly", C.c_int, [_P, C.c_int32, C.c_int32, C.c_double, _DP, _DP]),
    ("mpf_set_batch", C.c_int, [_P, C.c_int32, _I32P, _I32P, _DP, _DP]),
    ("mpf_batch_autotune", C.c_int, [_P, _I32P, _I32P, _DP]),
    ("mpf_batch_newton_solve", C.c_int, [_P, C.POINTER(BatchOpts), _I32P, _DP, C.POINTER(BatchReport)]),
    ("mpf_batch_tangent_increment", C.c_int, [_P, C.c_double, _P]),
    ("mpf_batch_newton_step", C.c_int, [_P, C.c_double, C.POINTER(BatchOpts), _P, _I32P, _DP,
                                        C.POINTER(BatchReport)]),
    ("mpf_state_device_ptrs", C.c_int, [_P, C.POINTER(C.c_void_p), C.POINTER(C.c_void_p)]),
    ("mpf_copy_state_device", C.c_int, [_P, _P]),
    ("mpf_mpnn_num_params", C.c_int32, []),
    ("mpf_mpnn_create", C.c_int, [C.POINTER(C.c_void_p), C.c_uint64]),
    ("mpf_mpnn_destroy", C.c_int, [_P]),
    ("mpf_mpnn_get_params", C.c_int, [_P, _P, _P]),
    ("mpf_mpnn_set_params", C.c_int, [_P, _P]),
    ("mpf_mpnn_zero_grad", C.c_int, [_P]),
    ("mpf_mpnn_features", C.c_int, [_P, _P, _P, _P, C.c_double, C.c_double, C.c_double]),
    ("mpf_mpnn_forward", C.c_int, [_P, _P, _P, _P]),
    ("mpf_mpnn_loss", C.c_int, [_P, _P, _P, _P, _P, C.c_double, _DP]),
    ("mpf_mpnn_residual_loss", C.c_int, [_P, _P, C.c_double, C.c_double, _P, _P, _P, C.c_double, _DP]),
    ("mpf_mpnn_backward", C.c_int, [_P, _P, _P]),
    ("mpf_mpnn_adam", C.c_int, [_P, C.c_float, C.c_float, C.c_float, C.c_float]),
    ("mpf_mpnn_guess", C.c_int, [_P, _P, _P, _P, C.c_double, C.c_int32, _P]),
    ("mpf_mpnn_guarded_guess", C.c_int, [_P, _P, C.c_double, _P, _P, C.c_double, _P, _I32P]),
    ("mpf_set_assembly_mode", C.c_int, [_P, C.c_int32]),
    ("mpf_get_assembly_mode", C.c_int, [_P, _I32P]),
    ("mpf_set_chunk_stage_cap", C.c_int, [_P, C.c_int32]),
    ("mpf_band_stats", C.c_int, [_P, _I32P, _I32P, C.POINTER(C.c_int64), _DP]),
    ("mpf_set_band_split", C.c_int, [_P, C.c_int32]),
    ("mpf_band_segments", C.c_int, [_P, _I32P, _I32P]),
    ("mpf_set_pmg", C.c_int, [_P, C.c_int32, C.c_double, C.c_double, C.c_int32]),
    ("mpf_set_pmg_levels", C.c_int, [_P, C.c_int32]),
    ("mpf_batch_assembly_choice", C.c_int, [_P, _I32P, _I32P, _DP]),
    ("mpf_batch_time_direct", C.c_int, [_P, C.c_int32, _DP]),
    ("mpf_pmg_stats", C.c_int, [_P, C.POINTER(PmgReport)]),
    ("mpf_mesh_hex_sizes", None, [C.c_int32, C.c_int32, C.c_int32, _I32P, _I32P]),
    ("mpf_mesh_hex", None, [C.c_int32, C.c_int32, C.c_int32, C.c_double, C.c_double, C.c_double, C.c_double,
                            C.c_uint64, _DP, _I32P]),
    ("mpf_mesh_tet10_sizes", None, [C.c_int32, _I32P, _I32P]),
    ("mpf_mesh_tet10", None, [C.c_int32, C.c_double, C.c_double, C.c_uint64, _DP, _I32P]),
    ("mpf_mesh_face_nodes", C.c_int32, [C.c_int32, C.c_int32, C.c_int32, C.c_int32, C.c_int32, _I32P]),
    ("mpf_mesh_hex_face_quads", C.c_int32, [C.c_int32, C.c_int32, C.c_int32, C.c_int32, _I32P]),
    ("mpf_system_create", C.c_int, [C.POINTER(_P), C.c_int32, C.c_int32, _I32P, _I32P, C.c_int32, _I32P]),
    ("mpf_system_zero", C.c_int, [_P]),
    ("mpf_system_map_elements", C.c_int, [_P, C.c_int32, C.c_int32, C.c_int32, _P, _P, C.c_int32, C.c_int32,
                                          _I32P]),
    ("mpf_system_add_elements", C.c_int, [_P, C.c_int32, _P, _P]),
    ("mpf_system_add_rhs", C.c_int, [_P, _P]),
    ("mpf_elem_apply_dirichlet_batch", C.c_int, [C.c_int32, C.c_int32, _P, _P, _P, _P, _P, _P]),
    ("mpf_malloc", C.c_int, [C.POINTER(_P), C.c_size_t]),
    ("mpf_free", C.c_int, [_P]),
    ("mpf_memcpy_htod", C.c_int, [_P, _P, C.c_size_t]),
    ("mpf_memcpy_dtoh", C.c_int, [_P, _P, C.c_size_t]),
]

_lib = None


def lib() -> C.CDLL:
    """Load libmpf.so (raises if it has not been built)."""
    global _lib
    if _lib is None:
        if not LIB_PATH.exists():
            raise RuntimeError(f"{LIB_PATH} is missing: run `python -m mphyfem_amd.build`")
        # torch ships its own libamdhip64.so.7 / libhsa-runtime64.so.1 with the same sonames as
        # /opt/rocm's.  Whichever loads first serves the whole process; load torch's first so
        # one HIP runtime owns the device and torch tensors can be passed to libmpf.
        try:
            import torch  # noqa: F401
        except ImportError:
            pass
        L = C.CDLL(str(LIB_PATH))
        for name, res, args in SIGNATURES:
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def status_string(st: int) -> str:
    return lib().mpf_status_string(st).decode()


def _check(st: int, what: str) -> None:
    if st == OK:
        return
    if st == ERR_NEG_JACOBIAN:
        raise NegativeJacobian(st, what)
    raise MpfError(st, what)


def device_count() -> int:
    return int(lib().mpf_device_count())


def _ptr(a: np.ndarray, t):
    return a.ctypes.data_as(C.POINTER(t))


def _dp(a):
    return None if a is None else _ptr(a, C.c_double)


def _ip(a):
    return None if a is None else _ptr(a, C.c_int32)


def matl(mu: float, Kinv: float, Utype: int = 3, mixed: bool = False, tangent: int = TANGENT_REFERENCE) -> Matl:
    return Matl(float(mu), float(Kinv), int(Utype), int(mixed), int(tangent), 0)


def matl_from_E_nu(E: float, nu: float, Utype: int = 3, mixed: bool = False,
                   tangent: int = TANGENT_REFERENCE) -> Matl:
    """mu = E / (2 (1 + nu)), Kinv = 3 (1 - 2 nu) / E."""
    return matl(E / (2.0 * (1.0 + nu)), 3.0 * (1.0 - 2.0 * nu) / E, Utype, mixed, tangent)


# ---------------------------------------------------------------------------------------------
# meshes
# ---------------------------------------------------------------------------------------------
def mesh_hex(nx: int, ny: int, nz: int, L=(1.0, 1.0, 1.0), jitter: float = 0.0, seed: int = 1):
    nn, ne = C.c_int32(), C.c_int32()
    lib().mpf_mesh_hex_sizes(nx, ny, nz, C.byref(nn), C.byref(ne))
    X = np.zeros((nn.value, 3))
    conn = np.zeros((ne.value, 8), dtype=np.int32)
    lib().mpf_mesh_hex(nx, ny, nz, float(L[0]), float(L[1]), float(L[2]), float(jitter), seed, _dp(X), _ip(conn))
    return X, conn


def mesh_tet10(n: int, L: float = 1.0, jitter: float = 0.0, seed: int = 2):
    nn, ne = C.c_int32(), C.c_int32()
    lib().mpf_mesh_tet10_sizes(n, C.byref(nn), C.byref(ne))
    X = np.zeros((nn.value, 3))
    conn = np.zeros((ne.value, 10), dtype=np.int32)
    lib().mpf_mesh_tet10(n, float(L), float(jitter), seed, _dp(X), _ip(conn))
    return X, conn


def face_nodes(elem_type: int, nx: int, ny: int, nz: int, face: int) -> np.ndarray:
    cnt = lib().mpf_mesh_face_nodes(elem_type, nx, ny, nz, face, None)
    out = np.zeros(cnt, dtype=np.int32)
    lib().mpf_mesh_face_nodes(elem_type, nx, ny, nz, face, _ip(out))
    return out


def hex_face_quads(nx: int, ny: int, nz: int, face: int) -> np.ndarray:
    cnt = lib().mpf_mesh_hex_face_quads(nx, ny, nz, face, None)
    out = np.zeros((cnt, 4), dtype=np.int32)
    lib().mpf_mesh_hex_face_quads(nx, ny, nz, face, _ip(out))
    return out


class SplitMix64:
    """The project's seeded generator: SplitMix64, u = (x >> 11) * 2^-53 (same as csrc/mesh.cpp)."""

    def __init__(self, seed: int):
        self.x = seed & 0xFFFFFFFFFFFFFFFF

    def next(self) -> int:
        self.x = (self.x + 0x9E3779B97F4A7C15) & 0xFFFFFFFFFFFFFFFF
        z = self.x
        z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & 0xFFFFFFFFFFFFFFFF
        z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & 0xFFFFFFFFFFFFFFFF
        return z ^ (z >> 31)

    def uniform(self) -> float:
        return (self.next() >> 11) * (1.0 / 9007199254740992.0)


@dataclass
class HexBatch:
    """Union of independent unit-cube Hex8 blocks (mesh m: nodes node_off[m]:node_off[m+1], ...)."""
    X: np.ndarray
    conn: np.ndarray
    node_off: np.ndarray
    elem_off: np.ndarray
    edges: np.ndarray
    bottom: list          # per mesh: node ids of the z = 0 face (global numbering)
    top: list             # per mesh: Quad4 faces of the z = 1 face, outward normal (global numbering)


def mesh_batch_hex(edges, jitter: float = 0.0, seed: int = 3) -> HexBatch:
    Xs, Cs, bottom, top = [], [], [], []
    node_off, elem_off = [0], [0]
    for m, n in enumerate(edges):
        n = int(n)
        X, conn = mesh_hex(n, n, n, jitter=jitter, seed=seed * 1000003 + m)
        o = node_off[-1]
        Xs.append(X)
        Cs.append(conn + o)
        bottom.append(face_nodes(HEX8, n, n, n, 4) + o)
        top.append(hex_face_quads(n, n, n, 5) + o)
        node_off.append(o + len(X))
        elem_off.append(elem_off[-1] + len(conn))
    return HexBatch(np.concatenate(Xs), np.concatenate(Cs).astype(np.int32), np.array(node_off, dtype=np.int32),
                    np.array(elem_off, dtype=np.int32), np.array(edges, dtype=np.int32), bottom, top)


def batch_workload(nmesh: int = 256, nmin: int = 8, nmax: int = 32, seed: int = 3):
    """The batched workload: edge n ~ U{nmin..nmax}, E log-uniform [1e2, 1e4], nu ~ U[0.2, 0.45],
    top-face dead traction of magnitude tau ~ U[0.01, 0.1] E at angle theta ~ U[0, 30 deg] from the
    inward normal, in the reference's (normal, tangent1) form: v0 = -tau cos(theta), v1 = tau sin(theta).
    Returns (edges, E, nu, v0, v1)."""
    g = SplitMix64(seed)
    edges, E, nu, v0, v1 = [], [], [], [], []
    for _ in range(nmesh):
        edges.append(nmin + min(int(g.uniform() * (nmax - nmin + 1)), nmax - nmin))
        e = 10.0 ** (2.0 + 2.0 * g.uniform())
        E.append(e)
        nu.append(0.2 + 0.25 * g.uniform())
        tau = (0.01 + 0.09 * g.uniform()) * e
        th = g.uniform() * np.pi / 6.0
        v0.append(-tau * np.cos(th))
        v1.append(tau * np.sin(th))
    return np.array(edges), np.array(E), np.array(nu), np.array(v0), np.array(v1)


def batch_problem(edges, E, nu, v0, v1, tangent: int = TANGENT_CONSISTENT, Utype: int = 3, elem_type: int = HEX8):
    """Problem holding the batch: bottom faces fixed, top faces under their dead traction (HEX8, or
    HEX8P0 for Q1/P0 meshes)."""
    B = mesh_batch_hex(edges)
    P = Problem(elem_type, B.X, B.conn, matl_from_E_nu(float(E[0]), float(nu[0]), Utype, elem_type == HEX8P0, tangent))
    E, nu = np.asarray(E, dtype=np.float64), np.asarray(nu, dtype=np.float64)
    P.set_batch(B.node_off, B.elem_off, E / (2.0 * (1.0 + nu)), 3.0 * (1.0 - 2.0 * nu) / E)
    for m in range(len(edges)):
        P.add_dirichlet(B.bottom[m], -1)
        P.add_traction(B.top[m], float(v0[m]), float(v1[m]))
    return B, P


# ---------------------------------------------------------------------------------------------
# element and material operators on device tensors (torch provides the device memory)
# ---------------------------------------------------------------------------------------------
def _stream():
    import torch
    return C.c_void_p(torch.cuda.current_stream().cuda_stream)


def elem_hex8_nh_batch(X, u, conn, m: Matl, first: int = 0, count: int | None = None):
    """Elem_Solid_3D_Disp::calcStiffnessAndResidual for a batch of Hex8 elements.

    X [nNode,3], u [nNode*3] float64 and conn [nElem,8] int32 are CUDA tensors.  Returns
    (Ke [count,24,24], Re [count,24], bad) where bad = 1 + first element id with det F < 0 or 0.
    """
    import torch
    count = conn.shape[0] - first if count is None else count
    dev = X.device
    Ke = torch.empty((count, 24, 24), dtype=torch.float64, device=dev)
    Re = torch.empty((count, 24), dtype=torch.float64, device=dev)
    bad = torch.zeros(1, dtype=torch.int32, device=dev)
    st = lib().mpf_elem_hex8_nh_batch(X.data_ptr(), u.data_ptr(), conn.data_ptr(), C.byref(m), first, count,
                                      Ke.data_ptr(), Re.data_ptr(), bad.data_ptr(), _stream())
    _check(st, "mpf_elem_hex8_nh_batch")
    return Ke, Re, int(bad.item())


def elem_hex8p0_nh_batch(X, u, u_prev, pres_elem, conn, m: Matl, first: int = 0, count: int | None = None):
    """Elem_Solid_3D_Mixed0 (Hex8 Q1/P0), condensed, for a batch of elements:
    (Kc [n,24,24], Fc [n,24], Kup [n,24], Kpp [n], Fp [n], bad)."""
    import torch
    count = conn.shape[0] - first if count is None else count
    dev = X.device
    f64 = dict(dtype=torch.float64, device=dev)
    Kc = torch.empty((count, 24, 24), **f64)
    Fc = torch.empty((count, 24), **f64)
    Kup = torch.empty((count, 24), **f64)
    Kpp = torch.empty((count,), **f64)
    Fp = torch.empty((count,), **f64)
    bad = torch.zeros(1, dtype=torch.int32, device=dev)
    st = lib().mpf_elem_hex8p0_nh_batch(X.data_ptr(), u.data_ptr(), u_prev.data_ptr(), pres_elem.data_ptr(),
                                        conn.data_ptr(), C.byref(m), first, count, Kc.data_ptr(), Fc.data_ptr(),
                                        Kup.data_ptr(), Kpp.data_ptr(), Fp.data_ptr(), bad.data_ptr(), _stream())
    _check(st, "mpf_elem_hex8p0_nh_batch")
    return Kc, Fc, Kup, Kpp, Fp, int(bad.item())


def elem_tet10_nh_batch(X, u, conn, m: Matl, first: int = 0, count: int | None = None):
    """Elem_Solid_3D_Disp::calcStiffnessAndResidual (Tet10, 4 Gauss points) for a batch of elements on the
    device: (Ke [count, 30, 30], Re [count, 30], bad)."""
    import torch
    count = conn.shape[0] - first if count is None else count
    Ke = torch.empty((count, 30, 30), dtype=torch.float64, device=X.device)
    Re = torch.empty((count, 30), dtype=torch.float64, device=X.device)
    bad = torch.zeros(1, dtype=torch.int32, device=X.device)
    st = lib().mpf_elem_tet10_nh_batch(X.data_ptr(), u.data_ptr(), conn.data_ptr(), C.byref(m), first, count,
                                       Ke.data_ptr(), Re.data_ptr(), bad.data_ptr(), _stream())
    _check(st, "mpf_elem_tet10_nh_batch")
    return Ke, Re, int(bad.item())


def elem_tet10p1_nh_batch(X, u, u_prev, pres, conn, m: Matl, first: int = 0, count: int | None = None):
    """ElementBase::calcStiffnessAndResidualMixed3D (Tet10/P1) for a batch of elements."""
    import torch
    count = conn.shape[0] - first if count is None else count
    dev = X.device
    f64 = dict(dtype=torch.float64, device=dev)
    Kuu = torch.empty((count, 30, 30), **f64)
    Kup = torch.empty((count, 30, 4), **f64)
    Kpu = torch.empty((count, 4, 30), **f64)
    Kpp = torch.empty((count, 4, 4), **f64)
    Fu = torch.empty((count, 30), **f64)
    Fp = torch.empty((count, 4), **f64)
    bad = torch.zeros(1, dtype=torch.int32, device=dev)
    st = lib().mpf_elem_tet10p1_nh_batch(X.data_ptr(), u.data_ptr(), u_prev.data_ptr(), pres.data_ptr(),
                                         conn.data_ptr(), C.byref(m), first, count, Kuu.data_ptr(), Kup.data_ptr(),
                                         Kpu.data_ptr(), Kpp.data_ptr(), Fu.data_ptr(), Fp.data_ptr(),
                                         bad.data_ptr(), _stream())
    _check(st, "mpf_elem_tet10p1_nh_batch")
    return Kuu, Kup, Kpu, Kpp, Fu, Fp, int(bad.item())


def neohooke_stress_tangent(F, p, m: Matl):
    """computeStressAndTangent_NeoHooke for a batch: F [n,3,3] (row-major) -> (sig [n,9], C [n,9,9])."""
    import torch
    n = F.shape[0]
    sig = torch.empty((n, 9), dtype=torch.float64, device=F.device)
    Cm = torch.empty((n, 9, 9), dtype=torch.float64, device=F.device)
    st = lib().mpf_neohooke_stress_tangent_batch(F.data_ptr(), p.data_ptr() if p is not None else None,
                                                 C.byref(m), n, sig.data_ptr(), Cm.data_ptr(), _stream())
    _check(st, "mpf_neohooke_stress_tangent_batch")
    return sig, Cm


# ---------------------------------------------------------------------------------------------
# problem (femSolidmechanics-shaped)
# ---------------------------------------------------------------------------------------------
@dataclass
class Pattern:
    rowptr: np.ndarray
    colidx: np.ndarray
    NodeDofArray: np.ndarray
    ID_pres: np.ndarray
    color: np.ndarray
    dispDOF: int
    presDOF: int
    ncolors: int


class Problem:
    """Device-resident problem mirroring femSolidmechanics' driver calls."""

    def __init__(self, elem_type: int, X: np.ndarray, conn: np.ndarray, m: Matl):
        self.elem_type = elem_type
        self.X = np.ascontiguousarray(X, dtype=np.float64)
        self.conn = np.ascontiguousarray(conn, dtype=np.int32)
        self.nNode, self.nElem = self.X.shape[0], self.conn.shape[0]
        self._h = C.c_void_p()
        st = lib().mpf_problem_create(C.byref(self._h), elem_type, self.nNode, self.nElem, _dp(self.X),
                                      _ip(self.conn), C.byref(m))
        _check(st, "mpf_problem_create")

    def close(self):
        if self._h:
            lib().mpf_problem_destroy(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    # setSpecifiedDOFs_Displacement / setBoundaryConditions
    def add_dirichlet(self, nodes, dof: int = -1, value=None):
        nodes = np.ascontiguousarray(nodes, dtype=np.int32)
        val = None if value is None else np.ascontiguousarray(np.broadcast_to(value, nodes.shape), dtype=np.float64)
        _check(lib().mpf_add_dirichlet(self._h, len(nodes), _ip(nodes), dof, _dp(val)), "mpf_add_dirichlet")

    def add_traction(self, faces, v0: float, v1: float = 0.0):
        """faces [nf, 4] (Quad4) or [nf, 6] (Tria6: corners, then the mid-edge nodes of (0,1), (1,2), (2,0))."""
        faces = np.ascontiguousarray(faces, dtype=np.int32)
        if faces.ndim != 2 or faces.shape[1] not in (4, 6):
            raise ValueError("faces must be [nf, 4] or [nf, 6]")
        f = lib().mpf_add_traction_tri6 if faces.shape[1] == 6 else lib().mpf_add_traction
        _check(f(self._h, faces.shape[0], _ip(faces), v0, v1), "mpf_add_traction")

    def prepare_matrix_pattern(self):
        _check(lib().mpf_prepare_pattern(self._h), "mpf_prepare_pattern")

    def sizes(self):
        d, p, c = C.c_int32(), C.c_int32(), C.c_int32()
        nnz = C.c_int64()
        _check(lib().mpf_get_sizes(self._h, C.byref(d), C.byref(p), C.byref(nnz), C.byref(c)), "mpf_get_sizes")
        return d.value, p.value, nnz.value, c.value

    def pattern(self) -> Pattern:
        d, p, nnz, nc = self.sizes()
        rowptr = np.zeros(d + p + 1, dtype=np.int32)
        colidx = np.zeros(nnz, dtype=np.int32)
        nda = np.zeros((self.nNode, 3), dtype=np.int32)
        idp = np.zeros(self.nNode, dtype=np.int32)
        color = np.zeros(self.nElem, dtype=np.int32)
        _check(lib().mpf_get_pattern(self._h, _ip(rowptr), _ip(colidx), _ip(nda), _ip(idp), _ip(color)),
               "mpf_get_pattern")
        return Pattern(rowptr, colidx, nda, idp, color, d, p, nc)

    def _npres(self) -> int:
        return self.nElem if self.elem_type == HEX8P0 else self.nNode

    def set_state(self, disp=None, disp_prev=None, pres=None, pres_prev=None):
        """disp/disp_prev: [nNode*3]; pres/pres_prev: [nNode], or [nElem] for HEX8P0 (the C call copies
        exactly that many doubles, so shorter arrays are refused here)."""
        def c(a, n, name):
            if a is None:
                return None
            a = np.ascontiguousarray(a, dtype=np.float64).reshape(-1)
            if a.size != n:
                raise ValueError(f"{name} has {a.size} entries, expected {n}")
            return a
        n3, npres = 3 * self.nNode, self._npres()
        a = [c(disp, n3, "disp"), c(disp_prev, n3, "disp_prev"), c(pres, npres, "pres"),
             c(pres_prev, npres, "pres_prev")]
        _check(lib().mpf_set_state(self._h, *[_dp(x) for x in a]), "mpf_set_state")

    def state(self):
        """(disp [nNode*3], pres [nNode], or [nElem] for HEX8P0)."""
        disp = np.zeros(self.nNode * 3)
        pres = np.zeros(self._npres())
        _check(lib().mpf_get_state(self._h, _dp(disp), _dp(pres)), "mpf_get_state")
        return disp, pres

    def set_boundary_conditions(self, lam: float):
        _check(lib().mpf_set_bc(self._h, lam), "mpf_set_bc")

    def calc_stiffness_and_residual(self, first_iter: bool, lam: float):
        _check(lib().mpf_assemble(self._h, int(first_iter), lam), "mpf_assemble")

    def system(self):
        d, p, nnz, _ = self.sizes()
        vals = np.zeros(nnz)
        rhs = np.zeros(d + p)
        _check(lib().mpf_get_system(self._h, _dp(vals), _dp(rhs)), "mpf_get_system")
        return vals, rhs

    def rhs_norm(self) -> float:
        v = C.c_double()
        _check(lib().mpf_rhs_norm(self._h, C.byref(v)), "mpf_rhs_norm")
        return v.value

    def factorise_and_solve(self, ksp=KSP_CG, pc=PC_BJACOBI3, rtol=1e-10, maxit=10000, restart=50):
        o = KrylovOpts(ksp, pc, restart, maxit, rtol)
        r = KrylovReport()
        st = lib().mpf_solve(self._h, C.byref(o), C.byref(r))
        if st not in (OK, ERR_NOT_CONVERGED):
            _check(st, "mpf_solve")
        return r

    def solve_with_arclength(self, dt, final_time, max_steps, max_iters=20, tol=1e-9, ksp=KSP_BAND, pc=PC_BJACOBI3,
                             lin_rtol=1e-12, maxit=20000):
        """Arc-length continuation (femSolidmechanics::solveWithArclength); returns (load factors, arc
        lengths, report) of the converged steps."""
        o = ArclengthOpts(dt, final_time, max_steps, max_iters, tol, KrylovOpts(ksp, pc, 50, maxit, lin_rtol))
        lam = np.zeros(max(max_steps, 1))
        ds = np.zeros(max(max_steps, 1))
        r = ArclengthReport()
        _check(lib().mpf_arclength_solve(self._h, C.byref(o), _dp(lam), _dp(ds), C.byref(r)), "mpf_arclength_solve")
        return lam[:r.steps_done], ds[:r.steps_done], r

    def stress_gauss(self):
        """[nElem, nGP, 11]: Cauchy stress (9, column-major map), von Mises, tr(sigma)/3."""
        ngp = 8 if self.elem_type in (HEX8, HEX8P0) else 4
        out = np.zeros((self.nElem, ngp, 11))
        _check(lib().mpf_stress_gauss(self._h, _dp(out)), "mpf_stress_gauss")
        return out

    def stress_project(self):
        """(element means [nElem, 11], nodal means over the adjacent elements [nNode, 11])."""
        ev, nv = np.zeros((self.nElem, 11)), np.zeros((self.nNode, 11))
        _check(lib().mpf_stress_project(self._h, _dp(ev), _dp(nv)), "mpf_stress_project")
        return ev, nv

    def write_vtu(self, path, time: float = 0.0, binary: bool = True):
        _check(lib().mpf_write_vtu(self._h, str(path).encode(), time, int(binary)), "mpf_write_vtu")

    def solve_noalloc(self, pc=PC_JACOBI, rtol=1e-10, maxit=10000):
        """mpf_solve_noalloc: Jacobi or unpreconditioned CG with no device allocation (library-owned
        work array; at most NOALLOC_MAX_DOFS rows)."""
        o = KrylovOpts(KSP_CG, pc, 0, maxit, rtol)
        r = KrylovReport()
        st = lib().mpf_solve_noalloc(self._h, C.byref(o), C.byref(r))
        if st not in (OK, ERR_NOT_CONVERGED):
            _check(st, "mpf_solve_noalloc")
        return r

    def solution(self):
        d, p, _, _ = self.sizes()
        du = np.zeros(d + p)
        _check(lib().mpf_get_solution(self._h, _dp(du)), "mpf_get_solution")
        return du

    def update(self, first_iter: bool):
        _check(lib().mpf_update(self._h, int(first_iter)), "mpf_update")

    @staticmethod
    def newton_opts(nsteps=10, dt=0.1, itmax=20, rtol=1e-9, ksp=KSP_CG, pc=PC_BJACOBI3, lin_rtol=1e-10,
                    maxit=20000, restart=50, forcing=False) -> NewtonOpts:
        return NewtonOpts(nsteps, dt, itmax, rtol, KrylovOpts(ksp, pc, restart, maxit, lin_rtol), int(forcing))

    def solve_with_newton_raphson(self, opts: NewtonOpts):
        iters = np.zeros(max(opts.nsteps, 1), dtype=np.int32)
        rep = NewtonReport()
        st = lib().mpf_newton_solve(self._h, C.byref(opts), _ip(iters), C.byref(rep))
        _check(st, "mpf_newton_solve")
        return iters, rep

    def newton_step(self, lam: float, opts: NewtonOpts, guess=None):
        it = C.c_int32()
        rep = NewtonReport()
        g = None if guess is None else np.ascontiguousarray(guess, dtype=np.float64).reshape(-1)
        if g is not None and g.size != 3 * self.nNode:
            raise ValueError(f"guess has {g.size} entries, expected {3 * self.nNode}")
        st = lib().mpf_newton_step(self._h, lam, C.byref(opts), _dp(g), C.byref(it), C.byref(rep))
        _check(st, "mpf_newton_step")
        return it.value, rep

    def set_assembly_mode(self, mode: int):
        """ASSEMBLY_GATHER (row-owner, default), ASSEMBLY_COLORED_SCATTER or ASSEMBLY_GATHER_DIRECT."""
        _check(lib().mpf_set_assembly_mode(self._h, mode), "mpf_set_assembly_mode")

    def set_chunk_stage_cap(self, cap: int):
        """Hex8 gather: most distinct elements a chunk of 8 wavefronts stages in LDS (0..32)."""
        _check(lib().mpf_set_chunk_stage_cap(self._h, cap), "mpf_set_chunk_stage_cap")

    def band_stats(self):
        """(blocks, largest half-bandwidth, free dofs, factorisation flops) of the banded direct solver."""
        nb, bm, nd, fl = C.c_int32(), C.c_int32(), C.c_int64(), C.c_double()
        _check(lib().mpf_band_stats(self._h, C.byref(nb), C.byref(bm), C.byref(nd), C.byref(fl)), "mpf_band_stats")
        return nb.value, bm.value, nd.value, fl.value

    def set_band_split(self, enable=True):
        """Banded direct solver: cut long blocks in two at a separator (default) or keep them whole."""
        _check(lib().mpf_set_band_split(self._h, 1 if enable else 0), "mpf_set_band_split")

    def band_segments(self):
        """(segments, blocks cut in two) of the banded direct solver's layout."""
        ns, nc = C.c_int32(), C.c_int32()
        _check(lib().mpf_band_segments(self._h, C.byref(ns), C.byref(nc)), "mpf_band_segments")
        return ns.value, nc.value

    def set_pmg(self, degree=1, ratio=10.0, coarse_rtol=0.1, coarse_maxit=1000, levels=3):
        """Parameters of the p-multigrid preconditioner PC_PMG (Chebyshev smoothing degree and interval
        ratio, the lowest level's CG relative tolerance and iteration limit, and the number of levels:
        2, the vertex level solved by CG; 3, a V-cycle on the vertex level with vertex aggregates below)."""
        _check(lib().mpf_set_pmg(self._h, degree, ratio, coarse_rtol, coarse_maxit), "mpf_set_pmg")
        _check(lib().mpf_set_pmg_levels(self._h, levels), "mpf_set_pmg_levels")

    def pmg_stats(self) -> PmgReport:
        """Coarse level size and the V-cycles / coarse CG iterations of the last PC_PMG solve."""
        r = PmgReport()
        _check(lib().mpf_pmg_stats(self._h, C.byref(r)), "mpf_pmg_stats")
        return r

    def assembly_mode(self) -> int:
        v = C.c_int32()
        _check(lib().mpf_get_assembly_mode(self._h, C.byref(v)), "mpf_get_assembly_mode")
        return v.value

    # ---- batched independent meshes ----
    def set_batch(self, node_off, elem_off, mu, Kinv):
        """The problem's mesh is the union of len(mu) independent meshes (Hex8), one material each."""
        node_off = np.ascontiguousarray(node_off, dtype=np.int32)
        elem_off = np.ascontiguousarray(elem_off, dtype=np.int32)
        mu = np.ascontiguousarray(mu, dtype=np.float64)
        Kinv = np.ascontiguousarray(Kinv, dtype=np.float64)
        self.nmesh = len(mu)
        _check(lib().mpf_set_batch(self._h, self.nmesh, _ip(node_off), _ip(elem_off), _dp(mu), _dp(Kinv)),
               "mpf_set_batch")

    def batch_autotune(self):
        """(tile, lanes per row, ms per CG iteration) chosen by timing the candidates on this batch."""
        t, l, ms = C.c_int32(), C.c_int32(), C.c_double()
        _check(lib().mpf_batch_autotune(self._h, C.byref(t), C.byref(l), C.byref(ms)), "mpf_batch_autotune")
        return t.value, l.value, ms.value

    def batch_time_direct(self, reps: int = 5) -> float:
        """ms per direct solve of every mesh of the last assembled system (batched solver BAND)."""
        ms = C.c_double()
        _check(lib().mpf_batch_time_direct(self._h, reps, C.byref(ms)), "mpf_batch_time_direct")
        return ms.value

    def batch_assembly_choice(self):
        """(ASSEMBLY_* mode, staged-element cap, ms per assembly) picked by batch_autotune."""
        m, c, ms = C.c_int32(), C.c_int32(), C.c_double()
        _check(lib().mpf_batch_assembly_choice(self._h, C.byref(m), C.byref(c), C.byref(ms)),
               "mpf_batch_assembly_choice")
        return m.value, c.value, ms.value

    def batch_newton_solve(self, nsteps=1, dt=1.0, itmax=20, rtol=1e-9, cg_maxit=5000, cg_rtol=1e-10,
                           use_graph=True, chunk=0, tile=0, lanes_per_row=0, solver=BATCH_SOLVER_PCG):
        """Newton per mesh over nsteps load steps; returns (iters[nsteps][nmesh], relres[nmesh], BatchReport).
        solver: BATCH_SOLVER_PCG (Jacobi-PCG, inexact Newton) or BATCH_SOLVER_BAND (banded LDL^T per mesh)."""
        o = BatchOpts(nsteps, dt, itmax, rtol, cg_maxit, cg_rtol, int(use_graph), chunk, tile, lanes_per_row, solver)
        iters = np.zeros((nsteps, self.nmesh), dtype=np.int32)
        rel = np.zeros(self.nmesh)
        rep = BatchReport()
        st = lib().mpf_batch_newton_solve(self._h, C.byref(o), _ip(iters), _dp(rel), C.byref(rep))
        _check(st, "mpf_batch_newton_solve")
        return iters, rel, rep

    def batch_newton_step(self, lam, guess=None, rep=None, itmax=20, rtol=1e-9, cg_maxit=5000, cg_rtol=1e-10,
                          use_graph=True, chunk=0, tile=0, lanes_per_row=0, solver=BATCH_SOLVER_PCG):
        """One load step of every mesh; guess: device tensor [nNode*3] (float64) or None.
        Returns (iters[nmesh], relres[nmesh], rep) with rep accumulated across calls."""
        o = BatchOpts(1, 1.0, itmax, rtol, cg_maxit, cg_rtol, int(use_graph), chunk, tile, lanes_per_row, solver)
        iters = np.zeros(self.nmesh, dtype=np.int32)
        rel = np.zeros(self.nmesh)
        rep = BatchReport() if rep is None else rep
        gp = None
        if guess is not None:
            import torch
            assert guess.dtype == torch.float64 and guess.is_contiguous() and guess.numel() == 3 * self.nNode
            torch.cuda.current_stream().synchronize()  # the problem runs on its own stream
            gp = C.c_void_p(guess.data_ptr())
        st = lib().mpf_batch_newton_step(self._h, lam, C.byref(o), gp, _ip(iters), _dp(rel), C.byref(rep))
        _check(st, "mpf_batch_newton_step")
        return iters, rel, rep

    def batch_tangent_increment(self, lam, out):
        """out (device float64 tensor [nNode*3]) = K^-1 R(u; lam) with the factors of each mesh's last
        direct solve (batched solver BATCH_SOLVER_BAND): the tangent increment to load factor lam."""
        import torch
        assert out.dtype == torch.float64 and out.is_contiguous() and out.numel() == 3 * self.nNode
        torch.cuda.current_stream().synchronize()
        _check(lib().mpf_batch_tangent_increment(self._h, lam, C.c_void_p(out.data_ptr())),
               "mpf_batch_tangent_increment")

    def state_device_ptr(self) -> int:
        """Device address of the displacement [nNode*3] (float64), ordered on the problem's stream."""
        d, pr = C.c_void_p(), C.c_void_p()
        _check(lib().mpf_state_device_ptrs(self._h, C.byref(d), C.byref(pr)), "mpf_state_device_ptrs")
        return d.value

    def copy_state_device(self, dst):
        """Copy the displacement into a device tensor dst (float64, nNode * 3 entries)."""
        import torch
        assert dst.dtype == torch.float64 and dst.is_contiguous() and dst.numel() == 3 * self.nNode
        torch.cuda.current_stream().synchronize()
        _check(lib().mpf_copy_state_device(self._h, C.c_void_p(dst.data_ptr())), "mpf_copy_state_device")

    def time_assembly(self, reps: int = 20, first_iter: bool = True, lam: float = 1.0):
        """(ms per full assembly, ms of the element-kernel launches per assembly), HIP events."""
        a, s = C.c_double(), C.c_double()
        _check(lib().mpf_time_assembly(self._h, reps, int(first_iter), lam, C.byref(a), C.byref(s)),
               "mpf_time_assembly")
        return a.value, s.value


class _ProblemView(Problem):
    """A Problem whose handle another object owns (a Deck's problem): the calls work, close() leaves
    the handle alone."""

    def __init__(self, h, elem_type: int, nNode: int, nElem: int):
        self._h = h
        self.elem_type, self.nNode, self.nElem = elem_type, nNode, nElem
        self.X = self.conn = None

    def close(self):
        self._h = C.c_void_p()


# ---------------------------------------------------------------------------------------------
# input decks (csrc/deck.cpp): config + Gmsh 2.2 mesh + petsc_options.dat
# ---------------------------------------------------------------------------------------------
class Deck:
    """A reference input deck (config file, the Gmsh mesh it names, petsc_options.dat beside it)."""

    def __init__(self, config_path):
        self._h = C.c_void_p()
        _check(lib().mpf_deck_load(str(config_path).encode(), C.byref(self._h)), "mpf_deck_load")

    def close(self):
        if self._h:
            lib().mpf_deck_destroy(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def info(self) -> DeckInfo:
        i = DeckInfo()
        _check(lib().mpf_deck_get_info(self._h, C.byref(i)), "mpf_deck_get_info")
        return i

    def patches(self):
        """{name: (dim, sorted unique node ids)}."""
        out = {}
        for k in range(self.info().npatch):
            name = C.create_string_buffer(256)
            dim, nn = C.c_int32(), C.c_int32()
            _check(lib().mpf_deck_patch(self._h, k, name, 256, C.byref(dim), C.byref(nn), None), "mpf_deck_patch")
            nodes = np.zeros(nn.value, dtype=np.int32)
            _check(lib().mpf_deck_patch(self._h, k, None, 0, None, None, _ip(nodes)), "mpf_deck_patch")
            out[name.value.decode()] = (dim.value, nodes)
        return out

    def setup(self):
        """Builds the problem; returns a Problem view of it (owned by the deck)."""
        h = C.c_void_p()
        _check(lib().mpf_deck_setup(self._h, C.byref(h)), "mpf_deck_setup")
        i = self.info()
        return _ProblemView(h, i.elem_type, i.nNode, i.nElem)

    def run(self, vtu_prefix=None) -> DeckReport:
        r = DeckReport()
        st = lib().mpf_deck_run(self._h, None if vtu_prefix is None else str(vtu_prefix).encode(), C.byref(r))
        _check(st, "mpf_deck_run")
        return r


# ---------------------------------------------------------------------------------------------
# SolverPetsc-shaped linear system over a caller-supplied pattern (csrc/system.hip)
# ---------------------------------------------------------------------------------------------
class System:
    """SolverPetsc's assemble / solve calls (SolverPetsc.h:16-96) on the device, for a caller that
    keeps its own pattern and element loop.  Element matrices arrive as device tensors
    K [count, nrow, ncol] and F [count, nrow] with their dof lists rows [count, nrow], cols
    [count, ncol] (-1 = skip); map_elements builds the CSR-slot map once, add_elements sums in element
    order (bitwise the sequential ADD_VALUES loop)."""

    def __init__(self, dispDOF: int, presDOF: int, rowptr, colidx, NodeDofArray=None):
        rowptr = np.ascontiguousarray(rowptr, dtype=np.int32)
        colidx = np.ascontiguousarray(colidx, dtype=np.int32)
        if rowptr.size != dispDOF + presDOF + 1 or colidx.size != int(rowptr[-1]):
            raise ValueError("rowptr / colidx sizes do not match the dof counts")
        self.n, self.nnz = dispDOF + presDOF, int(rowptr[-1])
        nda = None if NodeDofArray is None else np.ascontiguousarray(NodeDofArray, dtype=np.int32).reshape(-1, 3)
        self._h = C.c_void_p()
        st = lib().mpf_system_create(C.byref(self._h), dispDOF, presDOF, _ip(rowptr), _ip(colidx),
                                     0 if nda is None else len(nda), None if nda is None else _ip(nda))
        _check(st, "mpf_system_create")

    def close(self):
        if self._h:
            lib().mpf_problem_destroy(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def zero_mtx(self):
        _check(lib().mpf_system_zero(self._h), "mpf_system_zero")

    def map_elements(self, rows, cols, r_off: int = 0, c_off: int = 0) -> int:
        """rows / cols: contiguous device int32 tensors [count, nrow] / [count, ncol]."""
        import torch
        count, nrow = rows.shape
        if cols.shape[0] != count or rows.dtype != torch.int32 or cols.dtype != torch.int32:
            raise ValueError("rows / cols: int32 [count, n] with the same count")
        torch.cuda.current_stream().synchronize()
        mid = C.c_int32()
        _check(lib().mpf_system_map_elements(self._h, count, nrow, cols.shape[1], _dptr(rows.contiguous()),
                                             _dptr(cols.contiguous()), r_off, c_off, C.byref(mid)),
               "mpf_system_map_elements")
        return mid.value

    def add_elements(self, map_id: int, K=None, F=None):
        import torch
        torch.cuda.current_stream().synchronize()
        _check(lib().mpf_system_add_elements(self._h, map_id, None if K is None else _dptr(K),
                                             None if F is None else _dptr(F)), "mpf_system_add_elements")

    def add_rhs(self, f):
        import torch
        torch.cuda.current_stream().synchronize()
        _check(lib().mpf_system_add_rhs(self._h, _dptr(f)), "mpf_system_add_rhs")

    def system(self):
        vals, rhs = np.zeros(self.nnz), np.zeros(self.n)
        _check(lib().mpf_get_system(self._h, _dp(vals), _dp(rhs)), "mpf_get_system")
        return vals, rhs

    def factorise_and_solve(self, ksp=KSP_CG, pc=PC_JACOBI, rtol=1e-10, maxit=10000, restart=50):
        o = KrylovOpts(ksp, pc, restart, maxit, rtol)
        r = KrylovReport()
        st = lib().mpf_solve(self._h, C.byref(o), C.byref(r))
        if st not in (OK, ERR_NOT_CONVERGED):
            _check(st, "mpf_solve")
        return r

    def solve_noalloc(self, pc=PC_JACOBI, rtol=1e-10, maxit=10000):
        o = KrylovOpts(KSP_CG, pc, 0, maxit, rtol)
        r = KrylovReport()
        st = lib().mpf_solve_noalloc(self._h, C.byref(o), C.byref(r))
        if st not in (OK, ERR_NOT_CONVERGED):
            _check(st, "mpf_solve_noalloc")
        return r

    def solution(self):
        x = np.zeros(self.n)
        _check(lib().mpf_get_solution(self._h, _dp(x)), "mpf_get_solution")
        return x


def apply_dirichlet_batch(forAssy, gdof, dispApplied, K, F):
    """ElementBase::applyDirichletBCs for a batch of element matrices (device tensors, F in place)."""
    import torch
    count, nsize = forAssy.shape
    torch.cuda.current_stream().synchronize()
    _check(lib().mpf_elem_apply_dirichlet_batch(count, nsize, _dptr(forAssy), _dptr(gdof), _dptr(dispApplied),
                                                _dptr(K), _dptr(F), None), "mpf_elem_apply_dirichlet_batch")
    torch.cuda.synchronize()


# ---------------------------------------------------------------------------------------------
# learned initial guess (message-passing network with BF16 MFMA kernels)
# ---------------------------------------------------------------------------------------------
def _dptr(t) -> C.c_void_p:
    return C.c_void_p(t.data_ptr() if hasattr(t, "data_ptr") else int(t))


class Mpnn:
    """The learned-guess network on a problem's node graph (csrc/mpnn.hip)."""

    FIN, WIDTH, LAYERS, FOUT = 32, 64, 3, 32

    def __init__(self, seed: int = 4):
        self._h = C.c_void_p()
        _check(lib().mpf_mpnn_create(C.byref(self._h), seed), "mpf_mpnn_create")
        self.nparams = lib().mpf_mpnn_num_params()

    def close(self):
        if self._h:
            lib().mpf_mpnn_destroy(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def params(self):
        w = np.zeros(self.nparams, dtype=np.float32)
        g = np.zeros(self.nparams, dtype=np.float32)
        _check(lib().mpf_mpnn_get_params(self._h, w.ctypes.data_as(C.c_void_p), g.ctypes.data_as(C.c_void_p)),
               "mpf_mpnn_get_params")
        return w, g

    def set_params(self, w):
        w = np.ascontiguousarray(w, dtype=np.float32)
        _check(lib().mpf_mpnn_set_params(self._h, w.ctypes.data_as(C.c_void_p)), "mpf_mpnn_set_params")

    def zero_grad(self):
        _check(lib().mpf_mpnn_zero_grad(self._h), "mpf_mpnn_zero_grad")

    def features(self, P: Problem, uk, ukm1, rho: float, lam_next: float, lam_ratio: float):
        _check(lib().mpf_mpnn_features(self._h, P._h, _dptr(uk), _dptr(ukm1), rho, lam_next, lam_ratio),
               "mpf_mpnn_features")

    def forward(self, P: Problem, x=None, y=None):
        """x: device int16/bfloat16 tensor [nNode, 32] or None (the features); y: device float32 [nNode, 32]."""
        import torch
        torch.cuda.current_stream().synchronize()
        _check(lib().mpf_mpnn_forward(self._h, P._h, None if x is None else _dptr(x), None if y is None else _dptr(y)),
               "mpf_mpnn_forward")

    def loss(self, P: Problem, unext, uk, ukm1, rho: float) -> float:
        v = C.c_double()
        _check(lib().mpf_mpnn_loss(self._h, P._h, _dptr(unext), _dptr(uk), _dptr(ukm1), rho, C.byref(v)),
               "mpf_mpnn_loss")
        return v.value

    def residual_loss(self, P: Problem, lam_next: float, dlam: float, unext, uk, ukm1, rho: float) -> float:
        """Linearised relative start residual (squared, mean over meshes) of the last prediction."""
        v = C.c_double()
        _check(lib().mpf_mpnn_residual_loss(self._h, P._h, lam_next, dlam, _dptr(unext), _dptr(uk), _dptr(ukm1),
                                            rho, C.byref(v)), "mpf_mpnn_residual_loss")
        return v.value

    def backward(self, P: Problem, dy=None):
        import torch
        torch.cuda.current_stream().synchronize()
        _check(lib().mpf_mpnn_backward(self._h, P._h, None if dy is None else _dptr(dy)), "mpf_mpnn_backward")

    def adam(self, lr: float = 1e-3, beta1: float = 0.9, beta2: float = 0.999, eps: float = 1e-8):
        _check(lib().mpf_mpnn_adam(self._h, lr, beta1, beta2, eps), "mpf_mpnn_adam")

    def guess(self, P: Problem, uk, ukm1, rho: float, use_net: bool, out):
        import torch
        torch.cuda.current_stream().synchronize()
        _check(lib().mpf_mpnn_guess(self._h, P._h, _dptr(uk), _dptr(ukm1), rho, int(use_net), _dptr(out)),
               "mpf_mpnn_guess")

    def guarded_guess(self, P: Problem, lam: float, uk, ukm1, rho: float, out) -> int:
        """Per mesh: the base predictor plus the network's correction scaled by the weight in [0, 1]
        that minimises the linearised start residual.  Returns the number of meshes with a weight > 0."""
        import torch
        torch.cuda.current_stream().synchronize()
        k = C.c_int32()
        _check(lib().mpf_mpnn_guarded_guess(self._h, P._h, lam, _dptr(uk), _dptr(ukm1), rho, _dptr(out), C.byref(k)),
               "mpf_mpnn_guarded_guess")
        return k.value
