"""GPU parity at the edge of the path: particles with no bonds at all (every
list empty, no bond slots allocated), including a single particle.  Then
A = m I and one step has a closed form (core/implicit_euler.cuh:19-176,
core/newmark_beta.cuh:19-199):

* implicit Euler: v' = v + h f / m, x' = x + h v';
* Newmark-beta: dx = h v + h^2/(2m) f, v' = 2 dx / h - v = v + h f / m;

with f = -c g m y^ (phased gravity, core/simulator.cu:1163-1167).  Each
operator (stored blocks, recomputed blocks, mixed precision) must equal the
oracle bit for bit and the closed form to rounding."""
import numpy as np
import pytest

import oracle as orc
from pdhelpers import make_pair

pytestmark = pytest.mark.gpu
pds = pytest.importorskip("pdsolver")

X4 = np.array([[0, 0, 0, 0], [1, 0, 0, 0], [0, 1, 0, 0], [0, 0, 1, 0.]])
V4 = np.array([[1, 2, 3, 0], [0, 0, 0, 0], [-1, 0, 5, 0], [0.5, -4, 0, 0.]])
OPERATORS = {"stored": dict(), "recomputed": dict(matrix_free=1), "mixed": dict(precision=1)}


def say(*a):
    print(*a, flush=True)


@pytest.mark.parametrize("integrator", [0, 1])
@pytest.mark.parametrize("op", sorted(OPERATORS))
@pytest.mark.parametrize("n", [1, 4])
def test_free_particles_closed_form(n, op, integrator):
    say("case free particles", n, op, integrator)
    m, h, g = 2e-3, 1e-4, 9.8
    sys_, ctx = make_pair(X4[:n].copy(), [(0.1, 0.05, 1.0, m)], v4=V4[:n].copy())
    x = sys_.x[:, :3].copy()
    v = sys_.v[:, :3].copy()
    for step in range(1, 7):
        c = step / 30.0
        kw = dict(dt=h, clock_scale=c, gravity=g, integrator=integrator, precond=1,
                  rtol=1e-14, **OPERATORS[op])
        st_o = sys_.step(orc.make_params(**kw))
        st_g = ctx.step(pds.make_params(**kw))
        assert st_o.converged and st_g.converged
        assert (st_g.cg_iterations, st_g.live_bonds, st_g.broken_bonds) == \
            (st_o.cg_iterations, 0, 0), step
        xg = ctx.download("position")[:, :3]
        vg = ctx.download("velocity")[:, :3]
        assert np.array_equal(xg, sys_.x[:, :3]), step
        assert np.array_equal(vg, sys_.v[:, :3]), step
        f = np.zeros(3)
        f[1] = -c * g * m
        if integrator == 0:
            v = v + h * f / m
            x = x + h * v
        else:
            x = x + h * v + h * h / (2.0 * m) * f
            v = v + h * f / m
        np.testing.assert_allclose(vg, v, rtol=1e-12, atol=1e-15)
        np.testing.assert_allclose(xg, x, rtol=1e-13, atol=1e-16)
    ctx.close()
