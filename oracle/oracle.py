"""ctypes wrapper of the CPU oracle (liborc_pd.so).

TEST INFRASTRUCTURE ONLY: the parity checker and the timed CPU baseline.
Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg load
this module; the product path never does.

Arrays follow the reference's host layout: positions, velocities and
forces are real4 (4 doubles, w lane zero) as in core/definitions.h:37-39,
and 3x3 blocks are row-major Mat3x3 (core/definitions.h:144-148).
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "liborc_pd.so")
MAX_BODIES = 8

_dp = C.POINTER(C.c_double)
_ip = C.POINTER(C.c_int32)
_lp = C.POINTER(C.c_int64)
_fp = C.POINTER(C.c_float)


class Material(C.Structure):
    _fields_ = [("horizon", C.c_double), ("radius", C.c_double),
                ("c_bond", C.c_double), ("mass", C.c_double)]


class System(C.Structure):
    _fields_ = [
        ("n", C.c_int32), ("nbody", C.c_int32),
        ("mat", Material * MAX_BODIES),
        ("x", _dp), ("v", _dp), ("x0", _dp), ("sol", _dp), ("f", _dp),
        ("b", _dp), ("fc", _dp), ("vbc", _dp),
        ("s", _dp), ("s_new", _dp), ("s00", _dp),
        ("act", _ip), ("body", _ip), ("bc", _ip),
        ("row_start", _lp), ("col", _ip), ("top", _ip),
        ("val", _dp), ("diag", _dp), ("minv", _dp),
        ("val32", _fp), ("diag32", _fp), ("minv32", _fp),
    ]


class StepParams(C.Structure):
    _fields_ = [
        ("dt", C.c_double), ("clock_scale", C.c_double), ("gravity", C.c_double),
        ("integrator", C.c_int32), ("precond", C.c_int32), ("rtol", C.c_double),
        ("max_iter", C.c_int32), ("walls", C.c_int32),
        ("bmin", C.c_double * 3), ("bmax", C.c_double * 3),
        ("wall_radius", C.c_double), ("restitution", C.c_double),
        ("contact_radius", C.c_double), ("contact_stiffness", C.c_double),
        ("contact_volume", C.c_double), ("contact_cell", C.c_double),
        ("precision", C.c_int32), ("inner_rtol", C.c_double),
        ("inner_max_iter", C.c_int32), ("max_refine", C.c_int32),
        ("contact_mode", C.c_int32),
    ]


class StepStats(C.Structure):
    _fields_ = [
        ("cg_iterations", C.c_int32), ("converged", C.c_int32),
        ("residual_norm", C.c_double), ("b_norm", C.c_double),
        ("broken_bonds", C.c_int64), ("live_bonds", C.c_int64),
        ("min_gap", C.c_double), ("gap_p", C.c_int32), ("gap_q", C.c_int32),
        ("refinements", C.c_int32),
    ]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


class RandState(C.Structure):
    _fields_ = [("ring", C.c_uint32 * 34), ("i", C.c_int64)]


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            raise RuntimeError(f"oracle library missing: {_LIB_PATH} (run make -C oracle)")
        L = C.CDLL(_LIB_PATH)
        L.orc_rand_seed.argtypes = [C.POINTER(RandState), C.c_uint32]
        L.orc_rand_next.argtypes = [C.POINTER(RandState)]
        L.orc_rand_next.restype = C.c_int32
        L.orc_getstretch.argtypes = [C.POINTER(RandState), C.c_float]
        L.orc_getstretch.restype = C.c_double
        L.orc_stretch_init.argtypes = [C.c_int32, C.c_float, C.c_uint32, _dp]
        L.orc_bonds_kdtree.argtypes = [C.c_int32, _dp, C.c_double, C.c_int32,
                                       C.c_int32, _ip, _ip]
        L.orc_bonds_kdtree.restype = C.c_int64
        L.orc_bonds_brute.argtypes = [C.c_int32, _dp, _ip, _dp, C.c_int32, _ip, _ip]
        L.orc_bonds_brute.restype = C.c_int64
        L.orc_dot.argtypes = [C.c_int32, _dp, _dp]
        L.orc_dot.restype = C.c_double
        L.orc_dot_f32.argtypes = [C.c_int32, _fp, _fp]
        L.orc_dot_f32.restype = C.c_double
        for name in ("orc_assemble", "orc_cg", "orc_cg_mixed"):
            getattr(L, name).argtypes = [C.POINTER(System), C.POINTER(StepParams),
                                         C.POINTER(StepStats)]
        L.orc_step.argtypes = [C.POINTER(System), C.POINTER(StepParams),
                               C.POINTER(StepStats)]
        L.orc_step.restype = C.c_int32
        L.orc_spmv.argtypes = [C.POINTER(System), _dp, _dp]
        L.orc_block_jacobi.argtypes = [C.POINTER(System)]
        L.orc_mixed_matrix.argtypes = [C.POINTER(System)]
        L.orc_round_blocks.argtypes = [C.POINTER(System)]
        L.orc_spmv_wide.argtypes = [C.POINTER(System), _dp, _dp]
        for name in ("orc_update", "orc_contact"):
            getattr(L, name).argtypes = [C.POINTER(System), C.POINTER(StepParams)]
        _lib = L
    return _lib


def _ptr(a, t):
    return a.ctypes.data_as(t)


# ---------------------------------------------------------------- rand
class GlibcRand:
    """glibc rand() restatement (TYPE_3 additive feedback)."""

    def __init__(self, seed: int):
        self.st = RandState()
        lib().orc_rand_seed(C.byref(self.st), seed)

    def next(self) -> int:
        return lib().orc_rand_next(C.byref(self.st))

    def getstretch(self, s0: float) -> float:
        return lib().orc_getstretch(C.byref(self.st), s0)


def stretch_init(n: int, s0: float, seed: int = 1973) -> np.ndarray:
    out = np.zeros(n, dtype=np.float64)
    lib().orc_stretch_init(n, s0, seed, _ptr(out, _dp))
    return out


# ---------------------------------------------------------------- bonds
def bonds_kdtree(x4, horizon, max_bonds=528, leaf_size=100):
    x4 = np.ascontiguousarray(x4, dtype=np.float64)
    n = x4.shape[0]
    counts = np.zeros(n, dtype=np.int32)
    lists = np.zeros((n, max_bonds), dtype=np.int32)
    dropped = lib().orc_bonds_kdtree(n, _ptr(x4, _dp), horizon, max_bonds,
                                     leaf_size, _ptr(counts, _ip), _ptr(lists, _ip))
    return counts, lists, dropped


def bonds_brute(x4, horizons, body=None, max_bonds=528):
    x4 = np.ascontiguousarray(x4, dtype=np.float64)
    n = x4.shape[0]
    counts = np.zeros(n, dtype=np.int32)
    lists = np.zeros((n, max_bonds), dtype=np.int32)
    h = np.ascontiguousarray(np.atleast_1d(horizons), dtype=np.float64)
    bp = None if body is None else np.ascontiguousarray(body, dtype=np.int32)
    total = lib().orc_bonds_brute(n, _ptr(x4, _dp),
                                  None if bp is None else _ptr(bp, _ip),
                                  _ptr(h, _dp), max_bonds,
                                  _ptr(counts, _ip), _ptr(lists, _ip))
    return counts, lists, total


def lists_to_csr(counts, lists):
    """(counts, padded lists) -> (row_start int64[n+1], col int32[nnz])."""
    counts = np.asarray(counts, dtype=np.int64)
    row_start = np.zeros(len(counts) + 1, dtype=np.int64)
    np.cumsum(counts, out=row_start[1:])
    mask = np.arange(lists.shape[1])[None, :] < counts[:, None]
    col = np.ascontiguousarray(lists[mask], dtype=np.int32)
    return row_start, col


def dot_f32(a4, b4) -> float:
    a4 = np.ascontiguousarray(a4, dtype=np.float32)
    b4 = np.ascontiguousarray(b4, dtype=np.float32)
    return lib().orc_dot_f32(a4.shape[0], _ptr(a4, _fp), _ptr(b4, _fp))


def dot(a4, b4) -> float:
    a4 = np.ascontiguousarray(a4, dtype=np.float64)
    b4 = np.ascontiguousarray(b4, dtype=np.float64)
    return lib().orc_dot(a4.shape[0], _ptr(a4, _dp), _ptr(b4, _dp))


# ---------------------------------------------------------------- system
def make_params(**kw) -> StepParams:
    p = StepParams()
    defaults = dict(dt=1e-4, clock_scale=1.0, gravity=9.8, integrator=0, precond=1,
                    rtol=1e-6, max_iter=100000, walls=0, bmin=(-1e30,) * 3,
                    bmax=(1e30,) * 3, wall_radius=0.0, restitution=0.001,
                    contact_radius=0.0, contact_stiffness=0.0, contact_volume=1.0,
                    contact_cell=0.0, precision=0, inner_rtol=1e-5, max_refine=20,
                    contact_mode=0)
    defaults.update(kw)
    # the engine's operator choice (stored or recomputed blocks) gives the
    # same results: nothing of it to restate
    defaults.pop("matrix_free", None)
    defaults.setdefault("inner_max_iter", defaults["max_iter"])
    for k, v in defaults.items():
        if k in ("bmin", "bmax"):
            getattr(p, k)[:] = [float(t) for t in v]
        else:
            setattr(p, k, v)
    return p


class OracleSystem:
    """Owns numpy arrays for every field of orc_system."""

    def __init__(self, x4, v4, s00, row_start, col, materials, body=None,
                 act=None, bc=None, vbc=None, s=None):
        n = x4.shape[0]
        self.n = n
        f64 = lambda a: np.ascontiguousarray(a, dtype=np.float64).copy()
        i32 = lambda a: np.ascontiguousarray(a, dtype=np.int32).copy()
        self.x = f64(x4)
        self.v = f64(v4)
        self.x0 = f64(x4)
        self.sol = np.zeros((n, 4))
        self.f = np.zeros((n, 4))
        self.b = np.zeros((n, 4))
        self.fc = np.zeros((n, 4))
        self.vbc = np.zeros((n, 4)) if vbc is None else f64(vbc)
        self.s00 = f64(s00)
        self.s = f64(s00 if s is None else s)
        self.s_new = np.zeros(n)
        self.act = np.zeros(n, np.int32) if act is None else i32(act)
        self.body = np.zeros(n, np.int32) if body is None else i32(body)
        self.bc = np.zeros(n, np.int32) if bc is None else i32(bc)
        self.row_start = np.ascontiguousarray(row_start, dtype=np.int64).copy()
        self.col = i32(col)
        self.top = (np.diff(self.row_start) - 1).astype(np.int32)
        nnz = int(self.row_start[-1])
        self.val = np.zeros((max(nnz, 1), 9))
        self.diag = np.zeros((n, 9))
        self.minv = np.zeros((n, 9))
        self.val32 = np.zeros((max(nnz, 1), 9), np.float32)
        self.diag32 = np.zeros((n, 9), np.float32)
        self.minv32 = np.zeros((n, 9), np.float32)
        self.materials = list(materials)
        self._struct = System()
        self._refresh()

    def _refresh(self):
        s = self._struct
        s.n = self.n
        s.nbody = len(self.materials)
        for i, m in enumerate(self.materials):
            s.mat[i] = Material(*[float(t) for t in m])
        for name in ("x", "v", "x0", "sol", "f", "b", "fc", "vbc", "s", "s_new",
                     "s00", "val", "diag", "minv"):
            setattr(s, name, _ptr(getattr(self, name), _dp))
        for name in ("val32", "diag32", "minv32"):
            setattr(s, name, _ptr(getattr(self, name), _fp))
        for name in ("act", "body", "bc", "col", "top"):
            setattr(s, name, _ptr(getattr(self, name), _ip))
        s.row_start = _ptr(self.row_start, _lp)

    def assemble(self, prm):
        st = StepStats()
        lib().orc_assemble(C.byref(self._struct), C.byref(prm), C.byref(st))
        return st

    def spmv(self, x4):
        x4 = np.ascontiguousarray(x4, dtype=np.float64)
        y = np.zeros((self.n, 4))
        lib().orc_spmv(C.byref(self._struct), _ptr(x4, _dp), _ptr(y, _dp))
        return y

    def block_jacobi(self):
        lib().orc_block_jacobi(C.byref(self._struct))

    def cg(self, prm):
        st = StepStats()
        lib().orc_cg(C.byref(self._struct), C.byref(prm), C.byref(st))
        return st

    def cg_mixed(self, prm):
        st = StepStats()
        lib().orc_cg_mixed(C.byref(self._struct), C.byref(prm), C.byref(st))
        return st

    def mixed_matrix(self):
        lib().orc_mixed_matrix(C.byref(self._struct))

    def spmv_wide(self, x4):
        x4 = np.ascontiguousarray(x4, dtype=np.float64)
        y = np.zeros((self.n, 4))
        lib().orc_spmv_wide(C.byref(self._struct), _ptr(x4, _dp), _ptr(y, _dp))
        return y

    def solve(self, prm):
        """orc_cg or orc_cg_mixed by prm.precision."""
        return self.cg_mixed(prm) if prm.precision == 1 else self.cg(prm)

    def update(self, prm):
        lib().orc_update(C.byref(self._struct), C.byref(prm))

    def contact(self, prm):
        lib().orc_contact(C.byref(self._struct), C.byref(prm))

    def swap_stretch(self):
        self.s[:] = self.s_new

    def step(self, prm):
        st = StepStats()
        lib().orc_step(C.byref(self._struct), C.byref(prm), C.byref(st))
        return st

    def live_bond_sets(self):
        """Per row: the live bond columns in list order."""
        return [self.col[self.row_start[p]:self.row_start[p] + self.top[p] + 1]
                for p in range(self.n)]

    def dense_matrix(self):
        """Dense 3n x 3n system matrix of the live bonds (small n only)."""
        n = self.n
        A = np.zeros((3 * n, 3 * n))
        for p in range(n):
            A[3 * p:3 * p + 3, 3 * p:3 * p + 3] += self.diag[p].reshape(3, 3)
            rs = self.row_start[p]
            for k in range(self.top[p] + 1):
                q = self.col[rs + k]
                A[3 * p:3 * p + 3, 3 * q:3 * q + 3] += self.val[rs + k].reshape(3, 3)
        return A
