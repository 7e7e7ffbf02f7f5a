"""Pins orc_assemble (core/cg_solver.cuh:200-339 + implicit_euler.cuh:19-86)
with closed forms, finite differences and invariants."""
import numpy as np
import pytest

import oracle as orc
from pdhelpers import lattice, make_system, material


def two_particles(d0, d, direction, c_bond=100.0, horizon=0.03, radius=0.005,
                  mass=1e-3, s=10.0):
    e = np.asarray(direction, float)
    e = e / np.linalg.norm(e)
    x0 = np.zeros((2, 4))
    x0[1, :3] = d0 * e
    sys = orc.OracleSystem(x0, np.zeros((2, 4)), np.full(2, s), np.array([0, 1, 2]),
                           np.array([1, 0]), [(horizon, radius, c_bond, mass)])
    sys.x[1, :3] = d * e
    return sys, e


@pytest.mark.parametrize("d", [0.0101, 0.0099, 0.0123])
def test_two_particle_closed_form(d):
    d0, c, dt, m, g, clock = 0.01, 100.0, 1e-4, 1e-3, 9.8, 0.5
    sys, e = two_particles(d0, d, [1.0, 2.0, -0.5], c_bond=c, mass=m)
    prm = orc.make_params(dt=dt, clock_scale=clock, gravity=g)
    st = sys.assemble(prm)
    u = sys.x[1, :3] - sys.x[0, :3]
    dd = np.linalg.norm(u)
    e = u / float(np.float32(dd))  # unit vector over the float-rounded length
    sigma = dd / d0 - 1.0
    f_expect = c * sigma * e
    f_expect[1] -= clock * g * m
    np.testing.assert_allclose(sys.f[0, :3], f_expect, rtol=1e-13, atol=1e-16)
    f_q = -c * sigma * e
    f_q[1] -= clock * g * m
    np.testing.assert_allclose(sys.f[1, :3], f_q, rtol=1e-13, atol=1e-16)
    J = (c / d0) * ((1 - d0 / dd) * np.eye(3) + (d0 / dd) * np.outer(e, e))
    np.testing.assert_allclose(sys.val[0].reshape(3, 3), -J * dt * dt, rtol=1e-13, atol=1e-22)
    np.testing.assert_allclose(sys.diag[0].reshape(3, 3), J * dt * dt + m * np.eye(3),
                               rtol=1e-13)
    # b = m v + dt f, v = 0
    np.testing.assert_allclose(sys.b[0, :3], dt * sys.f[0, :3], rtol=1e-15)
    # stretch schedule: s_new = (10 - 9c) s00 - 0.25 min(0, sigma)
    assert sys.s_new[0] == (10.0 - 9.0 * clock) * sys.s00[0] - 0.25 * min(0.0, sigma)
    assert st.broken_bonds == 0 and st.live_bonds == 2


def test_volume_correction_on_force_not_jacobian():
    # d0 beyond horizon - r: the force carries vscale, the Jacobian does not
    d0, d, c = 0.027, 0.0272, 100.0
    sys, e = two_particles(d0, d, [0, 0, 1.0], c_bond=c, horizon=0.03, radius=0.005)
    sys.assemble(orc.make_params(gravity=0.0))
    u = sys.x[1, :3] - sys.x[0, :3]
    dd = np.linalg.norm(u)
    e = u / float(np.float32(dd))
    vscale = 0.5 + (0.03 - d0) / 2.0 / 0.005
    np.testing.assert_allclose(sys.f[0, :3], c * (dd / d0 - 1) * vscale * e, rtol=1e-13)
    J = (c / d0) * ((1 - d0 / dd) * np.eye(3) + (d0 / dd) * np.outer(e, e))
    np.testing.assert_allclose(sys.val[0].reshape(3, 3), -J * 1e-8, rtol=1e-12, atol=1e-24)


def test_float_rounded_unit_vector():
    # cg_solver.cuh:256 divides by (float)max(d, 1e-10), not by d
    d0, d = 0.01, 0.0100000123456789
    sys, e = two_particles(d0, d, [1.0, 0.0, 0.0], c_bond=1.0)
    sys.assemble(orc.make_params(gravity=0.0))
    fd = float(np.float32(d))
    ex = d / fd
    assert sys.f[0, 0] == ((d / d0 - 1.0) * 1.0 * ex) * 1.0
    assert ex != 1.0


def _forces(sys, prm):
    sys.top[:] = np.diff(sys.row_start) - 1
    sys.assemble(prm)
    return sys.f[:, :3].copy()


def test_jacobian_matches_finite_differences():
    h = 0.01
    x4 = lattice(4, 4, 4, h, 0.05, 11)
    # choose r so that every bond has d0 <= horizon - r (vscale = 1): the
    # reference tangent leaves vscale out, so FD only matches there
    horizon = 2.2 * h
    c, l, _ = orc.bonds_brute(x4, [horizon])
    d0max = max(np.linalg.norm(x4[l[p, :c[p]], :3] - x4[p, :3], axis=1).max() for p in range(len(x4)))
    mat = (horizon, 0.5 * (horizon - d0max), 100.0, 1e-3)
    sys = make_system(x4, [mat], s0=10.0)
    rng = np.random.default_rng(3)
    sys.x[:, :3] += rng.uniform(-1, 1, (sys.n, 3)) * 0.01 * h
    prm = orc.make_params(dt=1.0, gravity=0.0, clock_scale=1.0)
    sys.assemble(prm)
    A = sys.dense_matrix()  # = m I + K with dt = 1
    K = A - 1e-3 * np.eye(3 * sys.n)
    base = sys.x.copy()
    # the unit vector divides by a float-rounded length (relative step
    # ~6e-8), so the stencil must move bond lengths by far more than that
    eps = 1e-3 * h
    for (q, k) in [(5, 0), (21, 1), (42, 2)]:
        sys.x[:] = base
        sys.x[q, k] += eps
        fp = _forces(sys, prm)
        sys.x[:] = base
        sys.x[q, k] -= eps
        fm = _forces(sys, prm)
        dfdx = ((fp - fm) / (2 * eps)).reshape(-1)
        # K = -df/dx
        np.testing.assert_allclose(-dfdx, K[:, 3 * q + k], rtol=1e-3, atol=1e-4 * np.abs(K).max())


def test_symmetry_bitwise_and_translation():
    h = 0.01
    x4 = lattice(6, 5, 4, h, 0.05, 4)
    sys = make_system(x4, [material(h)], s0=10.0)
    rng = np.random.default_rng(1)
    sys.x[:, :3] += rng.uniform(-1, 1, (sys.n, 3)) * 0.02 * h
    sys.assemble(orc.make_params(dt=1e-4, gravity=0.0))
    blocks = {}
    for p in range(sys.n):
        rs = sys.row_start[p]
        for k in range(sys.top[p] + 1):
            blocks[(p, int(sys.col[rs + k]))] = sys.val[rs + k].reshape(3, 3)
        d = sys.diag[p].reshape(3, 3)
        assert np.array_equal(d, d.T)
    for (p, q), B in blocks.items():
        assert np.array_equal(B, blocks[(q, p)])
        assert np.array_equal(B, B.T)
    A = sys.dense_matrix()
    t = np.tile([0.3, -0.2, 0.7], sys.n)
    np.testing.assert_allclose(A @ t, 1e-3 * t, rtol=1e-9, atol=1e-12)


def test_rotation_equivariance():
    h = 0.01
    x4 = lattice(5, 5, 5, h, 0.05, 8)
    sys = make_system(x4, [material(h)], s0=10.0)
    rng = np.random.default_rng(2)
    disp = rng.uniform(-1, 1, (sys.n, 3)) * 0.02 * h
    sys.x[:, :3] += disp
    prm = orc.make_params(dt=1e-4, gravity=0.0)
    sys.assemble(prm)
    f1 = sys.f[:, :3].copy()
    A1 = sys.dense_matrix()
    th = 0.7
    R = np.array([[np.cos(th), -np.sin(th), 0], [np.sin(th), np.cos(th), 0], [0, 0, 1.0]])
    R = R @ np.array([[1, 0, 0], [0, np.cos(0.3), -np.sin(0.3)], [0, np.sin(0.3), np.cos(0.3)]])
    x4r = x4.copy()
    x4r[:, :3] = x4[:, :3] @ R.T
    sys2 = make_system(x4r, [material(h)], s0=10.0)
    sys2.x[:, :3] = (x4[:, :3] + disp) @ R.T
    sys2.assemble(prm)
    np.testing.assert_allclose(sys2.f[:, :3], f1 @ R.T, rtol=1e-8, atol=1e-9 * np.abs(f1).max())
    Rb = np.kron(np.eye(sys.n), R)
    np.testing.assert_allclose(sys2.dense_matrix(), Rb @ A1 @ Rb.T, rtol=1e-8, atol=1e-14)


def test_breaking_threshold_both_rows():
    d0 = 0.01
    s = np.float32(0.02)
    thr = float(s)  # (double)(float)max(s_p, s_q)
    for sigma, breaks in [(thr * (1 - 1e-9), False), (thr * (1 + 1e-9), True)]:
        sys, e = two_particles(d0, d0 * (1 + sigma), [0.3, 0.1, 1.0], s=float(s))
        st = sys.assemble(orc.make_params(gravity=0.0))
        sig_eff = np.linalg.norm(sys.x[1, :3] - sys.x[0, :3]) / d0 - 1
        assert st.broken_bonds == (2 if breaks else 0)
        assert list(sys.top) == ([-1, -1] if breaks else [0, 0])
        assert st.min_gap < 1e-9 * thr * 10


def test_swap_remove_order():
    # cg_solver.cuh:266-275: a broken bond swaps with the last live bond and
    # the walker stays put, so the swapped-in bond is tested next
    h = 0.01
    x4 = lattice(4, 1, 1, h, 0.0, 0)
    sys = make_system(x4, [material(h, horizon_factor=3.5)], s0=0.02)
    sys.s[:] = 0.02
    assert list(sys.col[sys.row_start[0]:sys.row_start[1]]) == [1, 2, 3]
    sys.x[2, 0] += 0.5 * h  # bond 0-2 stretches 25 %, 0-1 and 0-3 unchanged
    st = sys.assemble(orc.make_params(gravity=0.0))
    assert sys.top[0] == 1
    assert list(sys.col[sys.row_start[0]:sys.row_start[1]]) == [1, 3, 2]
    # row 2 loses bonds to 0 (sigma .25), 1 (sigma .5) and 3 (sigma -.5 kept)
    assert sys.top[2] == 0
    assert int(sys.col[sys.row_start[2]]) == 3
    assert st.broken_bonds == 2 + 2 + 0
