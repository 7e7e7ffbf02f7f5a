"""Pins the oracle's mixed-precision solve (orc_cg_mixed, orc_spmv_wide,
orc_dot_f32): iterative refinement in FP64 around a float PCG.  The
reference solves in FP64 only (core/simulator.cu:735-799); the checks here
come from the mathematics of refinement, not from the oracle's own code:
dense numpy solves of the FP64 system (refinement's residual is that of the
FP64 matrix, so it converges to the FP64 solution, not to the solution of the
float-rounded matrix its inner solves use), the FP64 stopping rule on the
true residual, exactness of FP64 products of floats, and the accuracy floor
that float arithmetic puts on a single inner solve."""
import numpy as np
import pytest

import oracle as orc
from pdhelpers import lattice, make_system, material


def deformed_system(nx=5, ny=4, nz=4, seed=0, bc=None, vbc=None):
    h = 0.01
    x4 = lattice(nx, ny, nz, h, 0.05, seed)
    sys = make_system(x4, [material(h)], s0=10.0, bc=bc, vbc=vbc)
    rng = np.random.default_rng(seed + 100)
    sys.x[:, :3] += rng.uniform(-1, 1, (sys.n, 3)) * 0.02 * h
    sys.v[:, :3] = rng.uniform(-1, 1, (sys.n, 3))
    return sys


def dense_a32(sys):
    """A32 as the refinement sees it: FP64 diagonal, float bonds widened."""
    n = sys.n
    A = np.zeros((3 * n, 3 * n))
    for p in range(n):
        A[3 * p:3 * p + 3, 3 * p:3 * p + 3] += sys.diag[p].reshape(3, 3)
        rs = sys.row_start[p]
        for k in range(sys.top[p] + 1):
            q = sys.col[rs + k]
            A[3 * p:3 * p + 3, 3 * q:3 * q + 3] += sys.val32[rs + k].astype(np.float64).reshape(3, 3)
    return A


def mixed(**kw):
    kw.setdefault("precision", 1)
    kw.setdefault("dt", 1e-3)
    return orc.make_params(**kw)


def test_dot_f32_exact_on_integers():
    rng = np.random.default_rng(1)
    a = np.zeros((50000, 4), np.float32)
    b = np.zeros((50000, 4), np.float32)
    a[:, :3] = rng.integers(-4096, 4096, (50000, 3))
    b[:, :3] = rng.integers(-4096, 4096, (50000, 3))
    exact = int(np.sum(a[:, :3].astype(np.int64) * b[:, :3].astype(np.int64)))
    assert orc.dot_f32(a, b) == float(exact)


def test_dot_f32_products_are_exact_in_fp64():
    # a float times a float has at most 48 significant bits, so the FP64
    # product is exact and the float dot equals the FP64 dot of the widened
    # vectors bit for bit
    rng = np.random.default_rng(2)
    a = np.zeros((3000, 4), np.float32)
    b = np.zeros((3000, 4), np.float32)
    a[:, :3] = rng.normal(size=(3000, 3)) * np.exp(rng.normal(size=(3000, 3)) * 4)
    b[:, :3] = rng.normal(size=(3000, 3))
    assert orc.dot_f32(a, b) == orc.dot(a.astype(np.float64), b.astype(np.float64))


def test_spmv_wide_matches_dense():
    sys = deformed_system()
    sys.assemble(mixed())
    sys.mixed_matrix()
    assert np.array_equal(sys.val32[:len(sys.col)], sys.val[:len(sys.col)].astype(np.float32))
    A = dense_a32(sys)
    x = np.zeros((sys.n, 4))
    x[:, :3] = np.random.default_rng(3).normal(size=(sys.n, 3))
    y = sys.spmv_wide(x)
    np.testing.assert_allclose(y[:, :3].reshape(-1), A @ x[:, :3].reshape(-1), rtol=1e-12,
                               atol=1e-15 * np.abs(A).max())


@pytest.mark.parametrize("precond", [0, 1])
def test_refinement_reaches_fp64_tolerance(precond):
    # rtol 1e-12 is out of reach of float arithmetic: only FP64 residuals and
    # FP64 corrections get there, in more than one refinement
    sys = deformed_system()
    prm = mixed(precond=precond, rtol=1e-12)
    sys.assemble(prm)
    st = sys.cg_mixed(prm)
    assert st.converged == 1
    assert st.refinements >= 2
    A = sys.dense_matrix()
    b = sys.b[:, :3].reshape(-1)
    x = sys.sol[:, :3].reshape(-1)
    r = b - A @ x
    assert np.linalg.norm(r) <= 1e-12 * np.linalg.norm(b) * (1 + 1e-3)
    assert abs(st.residual_norm - np.linalg.norm(r)) <= 1e-14 * np.linalg.norm(b)
    assert st.b_norm == pytest.approx(np.linalg.norm(b), rel=1e-14)
    xs = np.linalg.solve(A, b)
    np.testing.assert_allclose(x, xs, rtol=0, atol=1e-10 * np.abs(xs).max())


def test_mixed_solution_is_the_fp64_systems():
    """Refined to rtol 1e-13 the mixed solution agrees with the FP64 solve
    to the tolerance, far closer than the solution of the float-rounded
    matrix (x_a32), which a residual formed with that matrix would reach."""
    sys = deformed_system(seed=4)
    p64 = orc.make_params(dt=1e-3, rtol=1e-13)
    sys.assemble(p64)
    sys.cg(p64)
    x64 = sys.sol[:, :3].reshape(-1).copy()
    A64 = sys.dense_matrix()
    sys.sol[:] = 0
    sys.cg_mixed(mixed(rtol=1e-13))
    x32 = sys.sol[:, :3].reshape(-1).copy()
    A32 = dense_a32(sys)
    b = sys.b[:, :3].reshape(-1)
    x_a32 = np.linalg.solve(A32, b)
    scale = np.abs(x64).max()
    assert np.abs(x32 - x64).max() <= 1e-10 * scale
    assert np.abs(x_a32 - x64).max() > 1e3 * np.abs(x32 - x64).max()
    assert np.linalg.norm(b - A64 @ x32) <= 1e-13 * np.linalg.norm(b) * 1.001


def test_single_refinement_stops_at_the_float_floor():
    # one inner solve asked for 1e-10: its recursive float residual gets
    # there, but the true residual of x + d stays near float accuracy
    sys = deformed_system(seed=5)
    prm = mixed(rtol=1e-14, inner_rtol=1e-10, max_refine=1, inner_max_iter=300)
    sys.assemble(prm)
    st = sys.cg_mixed(prm)
    assert st.refinements == 1 and st.converged == 0
    r0 = np.linalg.norm(sys.b[:, :3])  # cold start: r0 = S b
    assert 1e-10 * r0 < st.residual_norm < 1e-5 * r0


def test_inner_reduction_per_refinement():
    sys = deformed_system(seed=6)
    prm = mixed(rtol=1e-14, inner_rtol=1e-3, max_refine=1)
    sys.assemble(prm)
    st = sys.cg_mixed(prm)
    r0 = np.linalg.norm(sys.b[:, :3])
    # float recursion and true residual agree far above the float floor
    assert st.residual_norm <= 1.01e-3 * r0
    assert st.residual_norm > 1e-7 * r0


def test_block_diagonal_system_one_inner_iteration_per_refinement():
    # no bonds: block-Jacobi PCG solves each refinement in one iteration up
    # to the rounding of the float inverse, so every refinement gains about
    # seven digits
    x4 = lattice(4, 4, 4, 1.0, 0.0, 0)
    n = x4.shape[0]
    sys = orc.OracleSystem(x4, np.ones((n, 4)), np.ones(n), np.zeros(n + 1, np.int64),
                           np.zeros(0, np.int32), [(0.5, 0.25, 1.0, 2.0)])
    rng = np.random.default_rng(7)
    prm = mixed(rtol=1e-13, inner_rtol=0.5, dt=1.0, gravity=0.0)
    sys.assemble(prm)
    for p in range(n):  # SPD blocks
        B = rng.normal(size=(3, 3))
        B = B @ B.T + 3 * np.eye(3)
        B = 0.5 * (B + B.T)
        sys.diag[p] = B.reshape(-1)
    st = sys.cg_mixed(prm)
    assert st.converged == 1
    assert st.cg_iterations == st.refinements
    assert 2 <= st.refinements <= 3


def test_limits():
    sys = deformed_system(seed=8)
    prm = mixed(rtol=1e-30, max_iter=3)
    sys.assemble(prm)
    st = sys.cg_mixed(prm)
    assert st.converged == 0 and st.cg_iterations == 3
    sys.sol[:] = 0
    st = sys.cg_mixed(mixed(rtol=1e-30, max_refine=2, inner_max_iter=2))
    assert st.converged == 0 and st.refinements == 2 and st.cg_iterations == 4


def test_dirichlet_rows_hold_their_values():
    nx, ny, nz = 4, 4, 5
    n = nx * ny * nz
    k = np.arange(n) // (nx * ny)
    bc = ((k == 0) | (k == nz - 1)).astype(np.int32)
    vbc = np.zeros((n, 4))
    vbc[k == nz - 1, 2] = 0.25
    sys = deformed_system(nx, ny, nz, seed=9, bc=bc, vbc=vbc)
    prm = mixed(rtol=1e-11)
    sys.assemble(prm)
    st = sys.cg_mixed(prm)
    assert st.converged == 1
    assert np.array_equal(sys.sol[bc == 1, :3], vbc[bc == 1, :3])
    # the free rows solve A_ff x_f = b_f - A_fd x_d (the FP64 system)
    A = sys.dense_matrix()
    free = np.repeat(bc == 0, 3)
    x = sys.sol[:, :3].reshape(-1)
    r = sys.b[:, :3].reshape(-1) - A @ x
    assert np.linalg.norm(r[free]) <= 1e-11 * np.linalg.norm(sys.b[:, :3].reshape(-1)[free]) * 1.001
