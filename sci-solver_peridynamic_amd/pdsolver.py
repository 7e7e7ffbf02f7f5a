"""ctypes binding of libpdsolver.so (the C ABI of include/pd_solver.h).

This is plumbing for Python callers (tests, bench): every computation runs
in the HIP library.  Loading fails loudly when the library is missing; there
is no CPU fallback.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libpdsolver.so")
MAX_BODIES = 8

PD_OK = 0
PD_ERR_HIP = 1
PD_ERR_ARG = 2
PD_ERR_NOT_CONVERGED = 3
PD_ERR_BOND_OVERFLOW = 4
PD_ERR_STATE = 5
PD_ERR_NO_DEVICE = 6

FIELD = {
    "position": 0, "velocity": 1, "original_position": 2, "stretch": 3,
    "original_stretch": 4, "activity": 5, "body": 6, "bc_flag": 7, "bc_velocity": 8,
    "solution": 9, "force": 10, "rhs": 11, "new_stretch": 12, "bond_top": 13,
    "contact_force": 14,
}
_FIELD_KIND = {  # (dtype, components)
    0: (np.float64, 4), 1: (np.float64, 4), 2: (np.float64, 4), 3: (np.float64, 1),
    4: (np.float64, 1), 5: (np.int32, 1), 6: (np.int32, 1), 7: (np.int32, 1),
    8: (np.float64, 4), 9: (np.float64, 4), 10: (np.float64, 4), 11: (np.float64, 4),
    12: (np.float64, 1), 13: (np.int32, 1), 14: (np.float64, 4),
}


class Material(C.Structure):
    _fields_ = [("horizon", C.c_double), ("radius", C.c_double),
                ("c_bond", C.c_double), ("mass", C.c_double)]


class Config(C.Structure):
    _fields_ = [("num_particles", C.c_int32), ("device", C.c_int32),
                ("max_bonds", C.c_int32), ("num_bodies", C.c_int32),
                ("material", Material * MAX_BODIES)]


class StepParams(C.Structure):
    _fields_ = [
        ("dt", C.c_double), ("clock_scale", C.c_double), ("gravity", C.c_double),
        ("integrator", C.c_int32), ("precond", C.c_int32), ("rtol", C.c_double),
        ("max_iter", C.c_int32), ("walls", C.c_int32),
        ("bmin", C.c_double * 3), ("bmax", C.c_double * 3),
        ("wall_radius", C.c_double), ("restitution", C.c_double),
        ("contact_radius", C.c_double), ("contact_stiffness", C.c_double),
        ("contact_volume", C.c_double), ("contact_cell", C.c_double),
        ("track_gap", C.c_int32), ("timing", C.c_int32),
        ("precision", C.c_int32), ("max_refine", C.c_int32),
        ("inner_rtol", C.c_double), ("inner_max_iter", C.c_int32),
        ("contact_mode", C.c_int32), ("spmv_unroll", C.c_int32), ("graph", C.c_int32),
        ("learned_fallback", C.c_double), ("matrix_free", C.c_int32),
        ("learned_probe", C.c_int32),
    ]


class StepStats(C.Structure):
    _fields_ = [
        ("cg_iterations", C.c_int32), ("converged", C.c_int32),
        ("residual_norm", C.c_double), ("b_norm", C.c_double),
        ("broken_bonds", C.c_int64), ("live_bonds", C.c_int64),
        ("min_gap", C.c_double), ("gap_p", C.c_int32), ("gap_q", C.c_int32),
        ("ms_contact", C.c_double), ("ms_assemble", C.c_double), ("ms_solve", C.c_double),
        ("ms_update", C.c_double), ("ms_spmv_total", C.c_double),
        ("spmv_launches", C.c_int32), ("host_syncs", C.c_int32),
        ("refinements", C.c_int32), ("ms_resid_total", C.c_double),
        ("resid_launches", C.c_int32), ("precond_fallbacks", C.c_int32),
        ("ms_spmv_store", C.c_double), ("precond_probe", C.c_int32),
        ("operator_mf", C.c_int32), ("contact_grid", C.c_int32),
    ]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


class SpecimenStats(C.Structure):
    _fields_ = [("cg_iterations", C.c_int32), ("converged", C.c_int32),
                ("residual_norm", C.c_double), ("b_norm", C.c_double),
                ("refinements", C.c_int32)]


class BondStats(C.Structure):
    _fields_ = [("total_bonds", C.c_int64), ("min_bonds", C.c_int32),
                ("max_bonds", C.c_int32), ("slots", C.c_int64), ("ms_build", C.c_double)]


DEFAULT_PARAMS = dict(dt=1e-4, clock_scale=1.0, gravity=9.8, integrator=0, precond=1,
                      rtol=1e-6, max_iter=100000, walls=0, bmin=(-1e30,) * 3,
                      bmax=(1e30,) * 3, wall_radius=0.0, restitution=0.001,
                      contact_radius=0.0, contact_stiffness=0.0, contact_volume=1.0,
                      contact_cell=0.0, track_gap=0, timing=0, precision=0, max_refine=20,
                      inner_rtol=1e-5, contact_mode=0, spmv_unroll=0, graph=0,
                      learned_fallback=0.0, matrix_free=0, learned_probe=0)


def make_params(**kw) -> StepParams:
    p = StepParams()
    d = dict(DEFAULT_PARAMS)
    d.update(kw)
    d.setdefault("inner_max_iter", d["max_iter"])
    for k, v in d.items():
        if k in ("bmin", "bmax"):
            getattr(p, k)[:] = [float(t) for t in v]
        else:
            setattr(p, k, v)
    return p


class PDError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"pd error {code}: {msg}")
        self.code = code


_lib = None


def lib():
    """Load libpdsolver.so; raises if it is not built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"HIP library not built: {LIB_PATH} (run build.py)")
        L = C.CDLL(LIB_PATH)
        vp = C.c_void_p
        L.pd_ctx_create.argtypes = [C.POINTER(Config), C.POINTER(vp)]
        L.pd_ctx_destroy.argtypes = [vp]
        L.pd_last_error.argtypes = [vp]
        L.pd_last_error.restype = C.c_char_p
        L.pd_upload.argtypes = [vp, C.c_int32, vp, C.c_size_t]
        L.pd_download.argtypes = [vp, C.c_int32, vp, C.c_size_t]
        L.pd_build_bonds.argtypes = [vp, C.POINTER(BondStats)]
        L.pd_set_bonds.argtypes = [vp, vp, vp, C.c_int32]
        L.pd_get_bonds.argtypes = [vp, vp, vp, C.c_int32]
        L.pd_get_matrix.argtypes = [vp, vp, vp, C.c_int32]
        for name in ("pd_step", "pd_assemble", "pd_solve"):
            getattr(L, name).argtypes = [vp, C.POINTER(StepParams), C.POINTER(StepStats)]
        L.pd_update.argtypes = [vp, C.POINTER(StepParams)]
        L.pd_commit_stretch.argtypes = [vp]
        L.pd_spmv.argtypes = [vp, vp, vp]
        L.pd_backup.argtypes = [vp]
        L.pd_restore.argtypes = [vp]
        L.pd_synchronize.argtypes = [vp]
        L.pd_bench_spmv.argtypes = [vp, C.c_int32, C.POINTER(C.c_double), C.POINTER(C.c_double)]
        L.pd_spmv_bytes.argtypes = [vp, C.POINTER(C.c_double)]
        L.pd_check_division.argtypes = [vp, C.c_int64, C.c_uint64, C.POINTER(C.c_int64)]
        L.pd_set_specimens.argtypes = [vp, C.c_int32, vp, vp]
        L.pd_tune_spmv.argtypes = [vp, C.c_int32, vp, C.POINTER(C.c_int32)]
        L.pd_learned_set.argtypes = [vp, vp, C.c_int32]
        L.pd_get_precond.argtypes = [vp, vp]
        L.pd_trainer_create.argtypes = [C.c_int32, C.c_int64, C.c_uint64, C.POINTER(vp)]
        L.pd_trainer_destroy.argtypes = [vp]
        L.pd_trainer_last_error.argtypes = [vp]
        L.pd_trainer_last_error.restype = C.c_char_p
        L.pd_trainer_add.argtypes = [vp, vp, C.c_uint64, C.c_int32, C.POINTER(C.c_int64)]
        L.pd_trainer_fit.argtypes = [vp, C.c_int32, C.c_int32, C.c_float, C.c_uint64, vp]
        L.pd_trainer_fit_noalloc.argtypes = [vp, C.c_int32, C.c_int32, C.c_float, C.c_uint64, vp]
        L.pd_trainer_params.argtypes = [vp, vp, C.c_int32]
        L.pd_trainer_samples.argtypes = [vp, C.c_int64, C.c_int64, vp]
        L.pd_get_specimen_stats.argtypes = [vp, C.c_int32, C.POINTER(SpecimenStats)]
        for name in ("pd_ctx_create", "pd_ctx_destroy", "pd_upload", "pd_download",
                     "pd_build_bonds", "pd_set_bonds", "pd_get_bonds", "pd_get_matrix",
                     "pd_step", "pd_assemble", "pd_solve", "pd_update", "pd_commit_stretch",
                     "pd_spmv", "pd_backup", "pd_restore", "pd_synchronize", "pd_bench_spmv",
                     "pd_spmv_bytes", "pd_check_division", "pd_set_specimens", "pd_get_specimen_stats",
                     "pd_tune_spmv", "pd_learned_set", "pd_get_precond", "pd_trainer_create",
                     "pd_trainer_destroy", "pd_trainer_add", "pd_trainer_fit",
                     "pd_trainer_fit_noalloc", "pd_trainer_params", "pd_trainer_samples"):
            getattr(L, name).restype = C.c_int
        _lib = L
    return _lib


def _ptr(a):
    return a.ctypes.data_as(C.c_void_p)


class Context:
    """One particle system resident on a GPU (pd_ctx)."""

    def __init__(self, n, materials, device=0, max_bonds=528):
        cfg = Config()
        cfg.num_particles = int(n)
        cfg.device = int(device)
        cfg.max_bonds = int(max_bonds)
        cfg.num_bodies = len(materials)
        for i, m in enumerate(materials):
            cfg.material[i] = Material(*[float(t) for t in m])
        self.n = int(n)
        self.max_bonds = int(max_bonds)
        self._h = C.c_void_p()
        L = lib()
        rc = L.pd_ctx_create(C.byref(cfg), C.byref(self._h))
        if rc != PD_OK:
            raise PDError(rc, L.pd_last_error(None).decode())

    def _check(self, rc):
        if rc != PD_OK:
            raise PDError(rc, lib().pd_last_error(self._h).decode())

    def close(self):
        if self._h:
            lib().pd_ctx_destroy(self._h)
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

    # -- fields
    def upload(self, field, arr):
        f = FIELD[field]
        dt, comp = _FIELD_KIND[f]
        a = np.ascontiguousarray(arr, dtype=dt)
        if a.size != self.n * comp:
            raise ValueError(f"{field}: expected {self.n * comp} values, got {a.size}")
        self._check(lib().pd_upload(self._h, f, _ptr(a), a.nbytes))

    def download(self, field):
        f = FIELD[field]
        dt, comp = _FIELD_KIND[f]
        shape = (self.n, comp) if comp > 1 else (self.n,)
        a = np.zeros(shape, dtype=dt)
        self._check(lib().pd_download(self._h, f, _ptr(a), a.nbytes))
        return a

    # -- bonds
    def build_bonds(self):
        st = BondStats()
        self._check(lib().pd_build_bonds(self._h, C.byref(st)))
        return st

    def set_bonds(self, counts, lists):
        counts = np.ascontiguousarray(counts, dtype=np.int32)
        lists = np.ascontiguousarray(lists, dtype=np.int32)
        self._check(lib().pd_set_bonds(self._h, _ptr(counts), _ptr(lists), lists.shape[1]))

    def get_bonds(self, stride=None):
        stride = stride or self.max_bonds
        cap = np.zeros(self.n, np.int32)
        lists = np.zeros((self.n, stride), np.int32)
        self._check(lib().pd_get_bonds(self._h, _ptr(cap), _ptr(lists), stride))
        return cap, lists

    def get_matrix(self, stride=None):
        stride = stride or self.max_bonds
        diag = np.zeros((self.n, 9))
        vals = np.zeros((self.n, stride, 9))
        self._check(lib().pd_get_matrix(self._h, _ptr(diag), _ptr(vals), stride))
        return diag, vals

    # -- step pieces
    def assemble(self, prm):
        st = StepStats()
        self._check(lib().pd_assemble(self._h, C.byref(prm), C.byref(st)))
        return st

    def commit_stretch(self):
        self._check(lib().pd_commit_stretch(self._h))

    def solve(self, prm, allow_fail=False):
        st = StepStats()
        rc = lib().pd_solve(self._h, C.byref(prm), C.byref(st))
        if not (allow_fail and rc == PD_ERR_NOT_CONVERGED):
            self._check(rc)
        return st

    def update(self, prm):
        self._check(lib().pd_update(self._h, C.byref(prm)))

    def step(self, prm, allow_fail=False):
        st = StepStats()
        rc = lib().pd_step(self._h, C.byref(prm), C.byref(st))
        if not (allow_fail and rc == PD_ERR_NOT_CONVERGED):
            self._check(rc)
        return st

    def tune_spmv(self, iters=5):
        """Times the SpMV variants on the resident matrix; returns (ms per
        variant for 2/4/8 slots per trip, chosen slots per trip)."""
        ms = np.zeros(3)
        best = C.c_int32()
        self._check(lib().pd_tune_spmv(self._h, iters, _ptr(ms), C.byref(best)))
        return ms, best.value

    def learned_set(self, params):
        params = np.ascontiguousarray(params, np.float32)
        self._check(lib().pd_learned_set(self._h, _ptr(params), len(params)))

    def get_precond(self):
        out = np.zeros((self.n, 9))
        self._check(lib().pd_get_precond(self._h, _ptr(out)))
        return out

    def set_specimens(self, first, count):
        first = np.ascontiguousarray(first, np.int32)
        count = np.ascontiguousarray(count, np.int32)
        self._check(lib().pd_set_specimens(self._h, len(first), _ptr(first), _ptr(count)))

    def specimen_stats(self, nspec):
        out = (SpecimenStats * nspec)()
        self._check(lib().pd_get_specimen_stats(self._h, nspec, out))
        return [dict(cg_iterations=o.cg_iterations, converged=o.converged,
                     residual_norm=o.residual_norm, b_norm=o.b_norm, refinements=o.refinements)
                for o in out]

    def spmv(self, x4):
        x4 = np.ascontiguousarray(x4, dtype=np.float64)
        y = np.zeros((self.n, 4))
        self._check(lib().pd_spmv(self._h, _ptr(x4), _ptr(y)))
        return y

    def backup(self):
        self._check(lib().pd_backup(self._h))

    def restore(self):
        self._check(lib().pd_restore(self._h))

    def synchronize(self):
        self._check(lib().pd_synchronize(self._h))

    def bench_spmv(self, iters):
        ms = C.c_double()
        by = C.c_double()
        self._check(lib().pd_bench_spmv(self._h, iters, C.byref(ms), C.byref(by)))
        return ms.value, by.value

    def check_division(self, samples, seed=1):
        """Bitwise mismatches of the matrix-free operator's shared-reciprocal
        quotients against IEEE division on `samples` random operand pairs."""
        bad = C.c_int64()
        self._check(lib().pd_check_division(self._h, samples, seed, C.byref(bad)))
        return bad.value

    def spmv_bytes(self):
        by = C.c_double()
        self._check(lib().pd_spmv_bytes(self._h, C.byref(by)))
        return by.value


HOST_LIB_PATH = os.path.join(HERE, "libpdhost.so")
CONFIG_DIR = os.path.join(HERE, "configs")
_hlib = None


LEARNED_PARAMS = 5801
FIT_NOALLOC_MAX_BATCH = 4096  # PD_FIT_NOALLOC_MAX_BATCH
FIT_NOALLOC_MAX_STEPS = 16384  # PD_FIT_NOALLOC_MAX_STEPS
LEARNED_LAYOUT = dict(W1=(0, (64, 16)), b1=(1024, (64,)), W2=(1088, (64, 64)), b2=(5184, (64,)),
                      W3=(5248, (8, 64)), b3=(5760, (8,)), mean=(5768, (16,)),
                      istd=(5784, (16,)), eps=(5800, ()))


class Trainer:
    """pd_trainer: samples of assembled systems and the network fit."""

    def __init__(self, capacity, device=0, seed=1):
        self._h = C.c_void_p()
        rc = lib().pd_trainer_create(device, int(capacity), seed, C.byref(self._h))
        if rc != PD_OK:
            raise PDError(rc, lib().pd_trainer_last_error(None).decode())

    def _check(self, rc):
        if rc != PD_OK:
            raise PDError(rc, lib().pd_trainer_last_error(self._h).decode())

    def add(self, ctx, seed, stride):
        n = C.c_int64()
        self._check(lib().pd_trainer_add(self._h, ctx._h, seed, stride, C.byref(n)))
        return n.value

    def fit(self, steps, batch=8192, lr=1e-3, seed=7, noalloc=False):
        """Adam steps on the samples; noalloc: pd_trainer_fit_noalloc (no
        device allocation, batch and steps bounded by FIT_NOALLOC_MAX_*)."""
        losses = np.zeros(steps, np.float32)
        fn = lib().pd_trainer_fit_noalloc if noalloc else lib().pd_trainer_fit
        self._check(fn(self._h, steps, batch, lr, seed, _ptr(losses)))
        return losses

    def params(self):
        out = np.zeros(LEARNED_PARAMS, np.float32)
        self._check(lib().pd_trainer_params(self._h, _ptr(out), LEARNED_PARAMS))
        return out

    def samples(self, first, count):
        out = np.zeros((count, 28), np.float32)
        self._check(lib().pd_trainer_samples(self._h, first, count, _ptr(out)))
        return out

    def close(self):
        if self._h:
            lib().pd_trainer_destroy(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def host_lib():
    """libpdhost.so: the C++ SystemParameter / Scene classes (no device code)."""
    global _hlib
    if _hlib is None:
        if not os.path.exists(HOST_LIB_PATH):
            raise RuntimeError(f"host library not built: {HOST_LIB_PATH} (run build.py)")
        L = C.CDLL(HOST_LIB_PATH)
        vp = C.c_void_p
        L.pdh_scene.argtypes = [C.c_char_p, C.POINTER(C.c_int32), vp, vp, vp, vp, vp]
        L.pdh_scene.restype = C.c_int
        L.pdh_describe.argtypes = [C.c_char_p, C.c_char_p, C.c_int32]
        L.pdh_describe.restype = C.c_int
        L.pdh_stretch_init.argtypes = [C.c_int32, C.c_float, C.c_uint32, vp]
        L.pdh_load_obj.argtypes = [C.c_char_p, C.POINTER(C.c_int32), C.POINTER(C.c_int32), vp, vp]
        L.pdh_scene_extra.argtypes = [C.c_char_p, C.POINTER(C.c_int32), C.POINTER(C.c_int32), vp,
                                      vp, vp, vp]
        L.pdh_scene_extra.restype = C.c_int
        L.pdh_stretch_init_each.argtypes = [C.c_int32, vp, C.c_uint32, vp]
        L.pdh_load_obj.restype = C.c_int
        L.pdh_mesh_inside.argtypes = [C.c_int32, vp, C.c_int32, vp, C.c_int32, vp, vp]
        L.pdh_orient.argtypes = [C.c_int32, vp, vp, vp, vp, vp, vp]
        _hlib = L
    return _hlib


def load_obj(path):
    """OBJ mesh as read by mesh-filled scenes: (float32 vertices [nv,3],
    int32 triangles [nt,3])."""
    L = host_lib()
    nv, nf = C.c_int32(), C.c_int32()
    if L.pdh_load_obj(path.encode(), C.byref(nv), C.byref(nf), None, None) != 0:
        raise RuntimeError(f"cannot read {path}")
    v = np.zeros((nv.value, 3), np.float32)
    f = np.zeros((nf.value, 3), np.int32)
    L.pdh_load_obj(path.encode(), C.byref(nv), C.byref(nf), _ptr(v), _ptr(f))
    return v, f


def mesh_inside(xyz, tri, pts):
    """Exact point-in-mesh flags of pts [n,3] for the mesh (xyz [nv,3], tri [nt,3])."""
    xyz = np.ascontiguousarray(xyz, np.float64)
    tri = np.ascontiguousarray(tri, np.int32)
    pts = np.ascontiguousarray(pts, np.float64)
    out = np.zeros(len(pts), np.int32)
    host_lib().pdh_mesh_inside(len(xyz), _ptr(xyz), len(tri), _ptr(tri), len(pts), _ptr(pts),
                               _ptr(out))
    return out


def orient(a, b, c, d):
    """Exact signs of orient2d(a, b, c) and orient3d(a, b, c, d) per row."""
    a, b, c, d = (np.ascontiguousarray(t, np.float64) for t in (a, b, c, d))
    o2 = np.zeros(len(a), np.int32)
    o3 = np.zeros(len(a), np.int32)
    host_lib().pdh_orient(len(a), _ptr(a), _ptr(b), _ptr(c), _ptr(d), _ptr(o2), _ptr(o3))
    return o2, o3


def with_obj(params_path, obj_file):
    """A copy of the params file naming obj_file (the reference's driver
    takes the OBJ path on its command line, examples/main.cpp:37)."""
    import tempfile
    fd, path = tempfile.mkstemp(suffix=".txt", prefix="params_")
    with os.fdopen(fd, "w") as f, open(params_path) as g:
        f.write(g.read().rstrip("\n") + f"\nobj_file {obj_file}\n")
    return path


def describe(params_path, obj_file=None):
    """Derived run settings of a params file (materials, dt, solver keys)."""
    import json
    if obj_file is not None:
        tmp = with_obj(params_path, obj_file)
        try:
            return describe(tmp)
        finally:
            os.unlink(tmp)
    buf = C.create_string_buffer(16384)
    n = host_lib().pdh_describe(params_path.encode(), buf, len(buf))
    if n < 0:
        raise RuntimeError(f"cannot parse {params_path}")
    return json.loads(buf.value.decode())


def scene(params_path, obj_file=None):
    """Particles of the params file's scene: dict of x4, v4, body, bc, vbc."""
    if obj_file is not None:
        tmp = with_obj(params_path, obj_file)
        try:
            return scene(tmp)
        finally:
            os.unlink(tmp)
    L = host_lib()
    n = C.c_int32()
    if L.pdh_scene(params_path.encode(), C.byref(n), None, None, None, None, None) != 0:
        raise RuntimeError(f"cannot build the scene of {params_path}")
    n = n.value
    out = dict(x4=np.zeros((n, 4)), v4=np.zeros((n, 4)), body=np.zeros(n, np.int32),
               bc=np.zeros(n, np.int32), vbc=np.zeros((n, 4)))
    rc = L.pdh_scene(params_path.encode(), C.byref(C.c_int32()), _ptr(out["x4"]),
                     _ptr(out["v4"]), _ptr(out["body"]), _ptr(out["bc"]), _ptr(out["vbc"]))
    if rc != 0:
        raise RuntimeError(f"cannot build the scene of {params_path}")
    return out


def scene_extra(params_path):
    """Activity, per-particle s0 and the specimen table (first, count) of the
    params file's scene."""
    L = host_lib()
    n, ns = C.c_int32(), C.c_int32()
    if L.pdh_scene_extra(params_path.encode(), C.byref(n), C.byref(ns), None, None, None,
                         None) != 0:
        raise RuntimeError(f"cannot build the scene of {params_path}")
    act = np.zeros(n.value, np.int32)
    s0 = np.zeros(n.value, np.float32)
    first = np.zeros(ns.value, np.int32)
    count = np.zeros(ns.value, np.int32)
    L.pdh_scene_extra(params_path.encode(), C.byref(n), C.byref(ns), _ptr(act), _ptr(s0),
                      _ptr(first), _ptr(count))
    return dict(act=act, s0=s0, first=first, count=count)


def stretch_init_each(s0, seed=1973):
    """getstretch() limits with one s0 per particle (float32 array)."""
    s0 = np.ascontiguousarray(s0, np.float32)
    out = np.zeros(len(s0))
    host_lib().pdh_stretch_init_each(len(s0), _ptr(s0), seed, _ptr(out))
    return out


def stretch_init(n, s0, seed=1973):
    """getstretch() limits of the host path (C library rand())."""
    out = np.zeros(n)
    host_lib().pdh_stretch_init(n, s0, seed, _ptr(out))
    return out


def step_kwargs(desc, step):
    """pd_step_params fields of step `step` (1-based) of a described run."""
    ph = desc["phase_in"]
    clock = 1.0 if ph <= 0 else min(step, ph) / float(ph)
    return dict(dt=desc["dt"], clock_scale=clock, gravity=desc["gravity"],
                integrator=desc["integrator"], precond=desc["precond"], rtol=desc["rtol"],
                max_iter=desc["max_iter"], walls=desc["walls"], bmin=desc["bmin"],
                bmax=desc["bmax"], wall_radius=desc["wall_radius"],
                restitution=desc["restitution"], contact_radius=desc["contact_radius"],
                contact_stiffness=desc["contact_stiffness"],
                contact_volume=desc["contact_volume"], contact_cell=desc["contact_radius"],
                precision=desc.get("precision", 0), inner_rtol=desc.get("inner_rtol", 1e-5),
                max_refine=desc.get("max_refine", 20), inner_max_iter=desc["max_iter"],
                contact_mode=desc.get("contact_mode", 0),
                matrix_free=desc.get("matrix_free", 0))


def materials_of(desc):
    return [tuple(m) for m in desc["materials"][:desc["num_bodies"]]]


def load_state(ctx: Context, x4, v4, s00, s=None, body=None, act=None, bc=None, vbc=None,
               x0=None):
    """Upload the particle state the reference's makeReady() uploads."""
    n = ctx.n
    ctx.upload("position", x4)
    ctx.upload("original_position", x4 if x0 is None else x0)
    ctx.upload("velocity", v4)
    ctx.upload("original_stretch", s00)
    ctx.upload("stretch", s00 if s is None else s)
    ctx.upload("body", np.zeros(n, np.int32) if body is None else body)
    ctx.upload("activity", np.zeros(n, np.int32) if act is None else act)
    ctx.upload("bc_flag", np.zeros(n, np.int32) if bc is None else bc)
    ctx.upload("bc_velocity", np.zeros((n, 4)) if vbc is None else vbc)
    ctx.upload("solution", np.zeros((n, 4)))
