"""Pins of the oracle's Newmark-beta branch (integrator 1) by what the
mathematics fixes, independent of the engine:

* a particle with no bonds has A = m I, so one step is closed form:
  dx = h v + h^2/(2m) f and v' = 2 dx / h - v = v + h f / m
  (core/newmark_beta.cuh:62-86 fill, :89-110 velocity, :113-199 position);
* with bonds, the Newmark system is the implicit-Euler one with h^2/4 in
  place of h^2 (core/newmark_beta.cuh:19-59 against
  core/implicit_euler.cuh:19-59): the bond blocks differ by exactly a factor
  4 (a power of two), the diagonal is m I + K h^2/4, the right-hand side is
  h (m v + h/2 f), and the solve and update agree with a dense solve.

A dropped factor in h^2/4, in b or in v = 2 dx/h - v fails one of these.
"""
import numpy as np

import oracle as orc
from pdhelpers import lattice, make_system, material


def free_particles(m):
    x4 = np.array([[0, 0, 0, 0], [1, 0, 0, 0], [0, 1, 0, 0], [0, 0, 1, 0.]])
    v4 = np.array([[1, 2, 3, 0], [0, 0, 0, 0], [-1, 0, 5, 0], [0.5, -4, 0, 0.]])
    n = x4.shape[0]
    return orc.OracleSystem(x4, v4, np.ones(n), np.zeros(n + 1, np.int64),
                            np.zeros(0, np.int32), [(0.1, 0.05, 1.0, m)])


def test_newmark_free_particle_closed_form():
    m, h, g = 2e-3, 1e-4, 9.8
    sys_ = free_particles(m)
    x = sys_.x[:, :3].copy()
    v = sys_.v[:, :3].copy()
    for step in range(1, 7):
        c = step / 30.0  # phased gravity (core/simulator.cu:1163-1167)
        st = sys_.step(orc.make_params(dt=h, clock_scale=c, gravity=g, integrator=1, precond=1,
                                       rtol=1e-14))
        assert st.converged
        f = np.zeros(3)
        f[1] = -c * g * m
        dx = h * v + h * h / (2.0 * m) * f
        v = v + h * f / m
        x = x + dx
        np.testing.assert_allclose(sys_.v[:, :3], v, rtol=1e-12, atol=1e-15)
        np.testing.assert_allclose(sys_.x[:, :3], x, rtol=1e-13, atol=1e-16)
        # the solution of the Newmark system is the displacement dx
        np.testing.assert_allclose(sys_.sol[:, :3], dx, rtol=1e-12, atol=1e-18)


def test_newmark_free_particle_differs_from_implicit_euler():
    # same state and one step: implicit Euler moves by h v' (v' = v + h f/m),
    # Newmark-beta by h v + h^2/(2m) f, so the y displacements differ by
    # h^2 g / 2 (the closed forms differ exactly by half the gravity term)
    m, h, g = 1e-3, 2e-4, 9.8
    a, b = free_particles(m), free_particles(m)
    a.step(orc.make_params(dt=h, clock_scale=1.0, gravity=g, integrator=0, rtol=1e-14))
    b.step(orc.make_params(dt=h, clock_scale=1.0, gravity=g, integrator=1, rtol=1e-14))
    dy = b.x[:, 1] - a.x[:, 1]
    np.testing.assert_allclose(dy, 0.5 * h * h * g, rtol=1e-6)
    np.testing.assert_allclose(a.v[:, :3], b.v[:, :3], rtol=1e-12, atol=1e-15)


def bonded_pair():
    """A jittered lattice with random velocities and a random displacement
    (so the force, and every bond block, is nonzero), stretch limits high
    enough that nothing breaks."""
    hsp = 0.01
    x4 = lattice(6, 5, 4, hsp, 0.05, 7)
    rng = np.random.default_rng(8)
    v4 = np.zeros_like(x4)
    v4[:, :3] = rng.uniform(-1, 1, (x4.shape[0], 3))
    disp = rng.uniform(-1, 1, (x4.shape[0], 3)) * 0.03 * hsp
    out = []
    for _ in range(2):
        s = make_system(x4, [material(hsp)], v4=v4)
        s.x[:, :3] += disp
        s.s[:] = 10.0
        out.append(s)
    return out


def test_newmark_system_is_the_quarter_scaled_euler_system():
    h, m = 1e-4, 1e-3
    ie, nm = bonded_pair()
    kw = dict(dt=h, clock_scale=0.5, gravity=9.8, precond=1, rtol=1e-13)
    ie.assemble(orc.make_params(integrator=0, **kw))
    nm.assemble(orc.make_params(integrator=1, **kw))
    nnz = int(ie.row_start[-1])
    assert nnz > 0 and np.abs(ie.val[:nnz]).max() > 0
    # bond blocks: -J h^2 against -J (h^2 / 4), a power-of-two factor: exact
    assert np.array_equal(ie.val[:nnz], 4.0 * nm.val[:nnz])
    # diagonal: K h^2 + m I against K h^2/4 + m I
    eye = np.eye(3).ravel()
    K_ie = (ie.diag - m * eye) / (h * h)
    K_nm = (nm.diag - m * eye) / ((h * h) / 4.0)
    np.testing.assert_allclose(K_nm, K_ie, rtol=1e-9, atol=1e-9 * np.abs(K_ie).max())
    # the force is the integrator's input, the same for both
    assert np.array_equal(ie.f, nm.f)
    # right-hand sides: m v + h f (Euler), h (m v + h/2 f) (Newmark-beta)
    f, v = nm.f[:, :3], nm.v[:, :3]
    np.testing.assert_allclose(ie.b[:, :3], m * v + h * f, rtol=1e-14, atol=1e-20)
    np.testing.assert_allclose(nm.b[:, :3], h * (m * v + 0.5 * h * f), rtol=1e-14, atol=1e-24)
    assert np.abs(nm.b[:, :3] - h * m * v).max() > 0  # the force term matters


def test_newmark_step_against_dense_solve():
    h = 1e-4
    _, nm = bonded_pair()
    x_before = nm.x[:, :3].copy()
    v_before = nm.v[:, :3].copy()
    kw = dict(dt=h, clock_scale=0.5, gravity=9.8, integrator=1, precond=1, rtol=1e-13)
    nm.assemble(orc.make_params(**kw))
    A = nm.dense_matrix()
    b = nm.b[:, :3].ravel().copy()
    dx = np.linalg.solve(A, b).reshape(-1, 3)
    st = nm.cg(orc.make_params(**kw))
    assert st.converged
    np.testing.assert_allclose(nm.sol[:, :3], dx, rtol=1e-9, atol=1e-12 * np.abs(dx).max())
    nm.update(orc.make_params(**kw))
    # core/newmark_beta.cuh:106 and :130-131
    np.testing.assert_allclose(nm.v[:, :3], 2.0 / h * dx - v_before, rtol=1e-8,
                               atol=1e-8 * np.abs(v_before).max())
    np.testing.assert_allclose(nm.x[:, :3], x_before + dx, rtol=1e-14, atol=1e-16)
    # and the full step equals assemble + solve + update
    _, nm2 = bonded_pair()
    nm2.step(orc.make_params(**kw))
    assert np.array_equal(nm2.x, nm.x) and np.array_equal(nm2.v, nm.v)
