"""Pins orc_spmv (core/cg_solver.cuh:343-383), orc_cg (core/simulator.cu:
735-799, CUSP PCG structure), orc_dot and orc_update / orc_contact."""
import numpy as np
import pytest

import oracle as orc
from pdhelpers import lattice, make_system, material


def deformed_system(nx=5, ny=4, nz=4, seed=0, bc=None):
    h = 0.01
    x4 = lattice(nx, ny, nz, h, 0.05, seed)
    sys = make_system(x4, [material(h)], s0=10.0, bc=bc)
    rng = np.random.default_rng(seed + 100)
    sys.x[:, :3] += rng.uniform(-1, 1, (sys.n, 3)) * 0.02 * h
    sys.v[:, :3] = rng.uniform(-1, 1, (sys.n, 3))
    return sys


def test_spmv_matches_dense():
    sys = deformed_system()
    sys.assemble(orc.make_params(dt=1e-4))
    A = sys.dense_matrix()
    rng = np.random.default_rng(9)
    x = np.zeros((sys.n, 4))
    x[:, :3] = rng.normal(size=(sys.n, 3))
    y = sys.spmv(x)
    np.testing.assert_allclose(y[:, :3].reshape(-1), A @ x[:, :3].reshape(-1), rtol=1e-12,
                               atol=1e-15)
    assert np.all(y[:, 3] == 0.0)


def _tree_dot_py(a, b):
    n = a.shape[0]
    terms = (a[:, 0] * b[:, 0] + a[:, 1] * b[:, 1]) + a[:, 2] * b[:, 2]

    def bfly(v):
        v = list(v)
        m = 32
        while m >= 1:
            v = [v[l] + v[l ^ m] for l in range(64)]
            m //= 2
        return v[0]

    def tile(v):
        w = [bfly(v[64 * k:64 * k + 64]) for k in range(4)]
        return (w[0] + w[1]) + (w[2] + w[3])

    g = (n + 255) // 256
    parts = []
    for t in range(g):
        v = [float(terms[i]) if i < n else 0.0 for i in range(256 * t, 256 * t + 256)]
        parts.append(tile(v))
    fin = []
    for t in range(256):
        acc = 0.0
        for k in range(t, g, 256):
            acc = acc + parts[k]
        fin.append(acc)
    return tile(fin)


@pytest.mark.parametrize("n", [1, 255, 256, 257, 1000, 70000])
def test_dot_tree_order(n):
    rng = np.random.default_rng(n)
    a = np.zeros((n, 4))
    b = np.zeros((n, 4))
    a[:, :3] = rng.normal(size=(n, 3)) * np.exp(rng.normal(size=(n, 3)) * 5)
    b[:, :3] = rng.normal(size=(n, 3))
    got = orc.dot(a, b)
    if n <= 2000:
        assert got == _tree_dot_py(a, b)
    ref = float(np.sum(a[:, :3] * b[:, :3], dtype=np.longdouble))
    assert abs(got - ref) <= 1e-12 * np.sum(np.abs(a[:, :3] * b[:, :3]))


def test_dot_exact_on_integers():
    rng = np.random.default_rng(0)
    a = np.zeros((100000, 4))
    b = np.zeros((100000, 4))
    a[:, :3] = rng.integers(-1000, 1000, (100000, 3))
    b[:, :3] = rng.integers(-1000, 1000, (100000, 3))
    assert orc.dot(a, b) == float(np.sum(a[:, :3].astype(np.int64) * b[:, :3].astype(np.int64)))


@pytest.mark.parametrize("precond", [0, 1])
def test_cg_matches_dense_solve(precond):
    sys = deformed_system()
    prm = orc.make_params(dt=1e-4, precond=precond, rtol=1e-13)
    sys.assemble(prm)
    st = sys.cg(prm)
    assert st.converged == 1 and st.cg_iterations > 0
    A = sys.dense_matrix()
    x = np.linalg.solve(A, sys.b[:, :3].reshape(-1))
    np.testing.assert_allclose(sys.sol[:, :3].reshape(-1), x, rtol=1e-10, atol=1e-12 * np.abs(x).max())
    r = sys.b[:, :3].reshape(-1) - A @ sys.sol[:, :3].reshape(-1)
    assert np.linalg.norm(r) <= 1e-12 * np.linalg.norm(sys.b[:, :3])


def test_block_jacobi_cuts_iterations():
    sys = deformed_system(6, 6, 6)
    prm0 = orc.make_params(dt=1e-3, precond=0, rtol=1e-8)
    sys.assemble(prm0)
    it0 = sys.cg(prm0).cg_iterations
    sys.sol[:] = 0
    prm1 = orc.make_params(dt=1e-3, precond=1, rtol=1e-8)
    it1 = sys.cg(prm1).cg_iterations
    assert it1 <= it0


def _block_system(blocks):
    n = len(blocks)
    sys = orc.OracleSystem(np.zeros((n, 4)), np.zeros((n, 4)), np.ones(n),
                           np.zeros(n + 1, np.int64), np.zeros(0, np.int32),
                           [(1.0, 0.5, 1.0, 1.0)])
    for i, B in enumerate(blocks):
        sys.diag[i] = np.asarray(B, float).reshape(-1)
    return sys


def test_cg_terminates_within_distinct_eigenvalues():
    # identity-preconditioned CG on an operator with 3 distinct eigenvalues
    lams = [1.0, 2.5, 7.0]
    n = 300
    sys = _block_system([np.eye(3) * lams[i % 3] for i in range(n)])
    rng = np.random.default_rng(1)
    sys.b[:, :3] = rng.normal(size=(n, 3))
    st = sys.cg(orc.make_params(precond=0, rtol=1e-10))
    assert st.converged and st.cg_iterations <= 3


def test_block_jacobi_on_block_diagonal_is_one_iteration():
    rng = np.random.default_rng(2)
    blocks = []
    for _ in range(500):
        M = rng.normal(size=(3, 3))
        blocks.append(M @ M.T + 3 * np.eye(3))
    sys = _block_system(blocks)
    sys.b[:, :3] = rng.normal(size=(500, 3))
    st = sys.cg(orc.make_params(precond=1, rtol=1e-12))
    assert st.converged and st.cg_iterations == 1
    for i in range(500):
        np.testing.assert_allclose(blocks[i] @ sys.sol[i, :3], sys.b[i, :3], rtol=1e-12, atol=1e-12)


def test_cg_iteration_limit_reports_failure():
    sys = deformed_system()
    prm = orc.make_params(dt=1e-3, precond=0, rtol=1e-14, max_iter=2)
    sys.assemble(prm)
    st = sys.cg(prm)
    assert st.converged == 0 and st.cg_iterations == 2


def test_cg_dirichlet_rows_fixed():
    n = 5 * 4 * 4
    bc = np.zeros(n, np.int32)
    bc[:20] = 1
    sys = deformed_system(bc=bc)
    sys.vbc[:20, 2] = 1.5
    prm = orc.make_params(dt=1e-4, precond=1, rtol=1e-12)
    sys.assemble(prm)
    st = sys.cg(prm)
    assert st.converged
    assert np.all(sys.sol[:20, :3] == [0.0, 0.0, 1.5])
    A = sys.dense_matrix()
    free = np.repeat(bc == 0, 3)
    r = sys.b[:, :3].reshape(-1) - A @ sys.sol[:, :3].reshape(-1)
    assert np.linalg.norm(r[free]) <= 1e-12 * np.linalg.norm(sys.b[:, :3].reshape(-1)[free]) * 1.0001


def test_free_fall_closed_form():
    # no bonds: A = m I, b = m v + dt f with f = -c g m y^
    n, m, dt, g = 4, 2e-3, 1e-4, 9.8
    sys = orc.OracleSystem(np.array([[0, 0, 0, 0], [1, 0, 0, 0], [0, 1, 0, 0], [0, 0, 1, 0.]]),
                           np.array([[1, 2, 3, 0], [0, 0, 0, 0], [-1, 0, 5, 0], [0, 0, 0, 0.]]),
                           np.ones(n), np.zeros(n + 1, np.int64), np.zeros(0, np.int32),
                           [(0.1, 0.05, 1.0, m)])
    x_start = sys.x.copy()
    v_start = sys.v.copy()
    for step in range(1, 6):
        c = step / 30.0
        prm = orc.make_params(dt=dt, clock_scale=c, gravity=g, precond=1, rtol=1e-14)
        st = sys.step(prm)
        assert st.converged
    vexp = v_start.copy()
    vexp[:, 1] -= dt * g * sum(k / 30.0 for k in range(1, 6))
    np.testing.assert_allclose(sys.v[:, :3], vexp[:, :3], rtol=1e-13, atol=1e-15)
    xexp = x_start.copy()
    vv = v_start.copy()
    for k in range(1, 6):
        vv[:, 1] -= dt * g * k / 30.0
        xexp += vv * dt
    np.testing.assert_allclose(sys.x[:, :3], xexp[:, :3], rtol=1e-13, atol=1e-15)


def test_wall_clamp_closed_form():
    # implicit_euler.cuh:126-168: overshoot beyond max-r reflects by R
    sys = orc.OracleSystem(np.array([[0.95, 0.5, 0.5, 0]]), np.array([[10.0, -20.0, 0, 0]]),
                           np.ones(1), np.zeros(2, np.int64), np.zeros(0, np.int32),
                           [(0.1, 0.05, 1.0, 1.0)])
    sys.sol[0] = [10.0, -20.0, 0.0, 0.0]
    prm = orc.make_params(dt=0.01, walls=1, bmin=(0, 0, 0), bmax=(1, 1, 1), wall_radius=0.02,
                          restitution=0.001)
    sys.update(prm)
    newx = 0.95 + 10.0 * 0.01
    hi = 1 - 0.02
    assert sys.x[0, 0] == hi - 0.001 * (newx - hi)
    assert sys.v[0, 0] == 10.0 * -0.001
    newy = 0.5 + (-20.0) * 0.01
    assert sys.x[0, 1] == newy and sys.v[0, 1] == -20.0


def test_contact_two_particles():
    Rc, kc, vol = 0.0075, 2.5e8, 1.0
    x4 = np.array([[0, 0, 0, 0], [0.004, 0.003, 0.001, 0.0], [0.5, 0.5, 0.5, 0.0]])
    sys = orc.OracleSystem(x4, np.zeros((3, 4)), np.ones(3), np.zeros(4, np.int64),
                           np.zeros(0, np.int32), [(0.1, 0.05, 1.0, 1.0)] * 2,
                           body=np.array([0, 1, 1]))
    prm = orc.make_params(contact_radius=Rc, contact_stiffness=kc, contact_volume=vol,
                          contact_cell=0.01)
    sys.contact(prm)
    pq = x4[0, :3] - x4[1, :3]
    r = np.linalg.norm(pq)
    f = kc * (r - Rc) ** 2 / vol * pq / r
    np.testing.assert_allclose(sys.fc[0, :3], f, rtol=1e-14)
    np.testing.assert_allclose(sys.fc[1, :3], -f, rtol=1e-14)
    assert np.all(sys.fc[2] == 0)
    sys.body[:] = 0  # same body: no contact between bodies ...
    sys.contact(prm)
    assert np.all(sys.fc == 0)
    # ... but the reference's pair set (all pairs) has it
    sys.contact(orc.make_params(contact_radius=Rc, contact_stiffness=kc, contact_volume=vol,
                                contact_cell=0.01, contact_mode=1))
    np.testing.assert_allclose(sys.fc[0, :3], f, rtol=1e-14)
    np.testing.assert_allclose(sys.fc[1, :3], -f, rtol=1e-14)
    assert np.all(sys.fc[2] == 0)


def test_contact_all_pairs_brute_force():
    # every pair closer than Rc, by brute force in numpy, on a compressed cloud
    rng = np.random.default_rng(12)
    n = 400
    x4 = np.zeros((n, 4))
    x4[:, :3] = rng.uniform(0, 0.05, (n, 3))
    body = rng.integers(0, 2, n).astype(np.int32)
    sys = orc.OracleSystem(x4, np.zeros((n, 4)), np.ones(n), np.zeros(n + 1, np.int64),
                           np.zeros(0, np.int32), [(0.1, 0.05, 1.0, 1.0)] * 2, body=body)
    Rc, kc = 0.006, 3.0e6
    for mode in (0, 1):
        sys.contact(orc.make_params(contact_radius=Rc, contact_stiffness=kc, contact_volume=2.0,
                                    contact_cell=0.007, contact_mode=mode))
        d = x4[:, None, :3] - x4[None, :, :3]
        r = np.linalg.norm(d, axis=2)
        mask = (r < Rc) & (r > 0)
        if mode == 0:
            mask &= body[:, None] != body[None, :]
        with np.errstate(invalid="ignore", divide="ignore"):
            mag = np.where(mask, kc * (r - Rc) ** 2 / 2.0 / np.where(r > 0, r, 1), 0.0)
        f = (mag[:, :, None] * d).sum(1)
        assert mask.sum() > 20
        np.testing.assert_allclose(sys.fc[:, :3], f, rtol=1e-12, atol=1e-12 * np.abs(f).max())


def test_reference_timestep_units():
    # dt = pd_time_step * 2^-23: x + v*dt equals x + (v*2^-23)*ts bitwise
    ts = 128
    dt = ts * (1.0 / 8388608.0)
    rng = np.random.default_rng(0)
    v = rng.normal(size=10000)
    assert np.array_equal(v * dt, (v * (1.0 / 8388608.0)) * ts)
