"""Stored-block FP64 assemblies leave the bond blocks to the solve's first
product, which recomputes and writes them: the first CG product when the
solve's start is fused into the assembly (pd_step), the start's y = A x0
otherwise (pd_assemble then pd_solve).  The two contexts must hold the same
matrix, bit for bit, after every step, and the same trajectory; also when
the solve converges at its start and the CG product only writes blocks.
(The blocks that pd_get_matrix writes itself after a bare pd_assemble are
compared with the oracle's by the assembly parity test.)"""
import sys

import numpy as np
import pytest

from pdhelpers import config0_state, make_pair, material

pytestmark = pytest.mark.gpu
pds = pytest.importorskip("pdsolver")


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


@pytest.mark.parametrize("rtol", [1e-8, 1e10])
def test_step_matrix_equals_assembled_matrix(rtol):
    say("case first product writes blocks, rtol", rtol)
    x4, bc, vbc, v4 = config0_state(8, 8, 10)
    mats = [material(0.01)]
    _, stepped = make_pair(x4, mats, v4=v4, bc=bc, vbc=vbc)
    _, split = make_pair(x4, mats, v4=v4, bc=bc, vbc=vbc)
    broken = 0
    for k in range(1, 31):
        kw = dict(dt=1e-4, clock_scale=k / 30.0, precond=1, rtol=rtol)
        prm = pds.make_params(**kw)
        st = stepped.step(prm)
        sa = split.assemble(prm)
        split.commit_stretch()
        ss = split.solve(prm)
        split.update(prm)
        assert st.cg_iterations == ss.cg_iterations and st.broken_bonds == sa.broken_bonds, k
        if rtol > 1.0:
            assert st.cg_iterations == 0 and st.converged
        ds, vs = stepped.get_matrix()
        da, va = split.get_matrix()
        assert np.array_equal(ds, da), k
        assert np.array_equal(vs, va), k
        assert np.array_equal(stepped.download("position"), split.download("position")), k
        broken += st.broken_bonds
    say(dict(broken=broken))
    if rtol < 1.0:
        assert broken > 0, "the tension case must break bonds"
    stepped.close()
    split.close()


def test_blocks_follow_the_assembly_not_later_uploads():
    """After a bare pd_assemble the blocks are still unwritten; replacing the
    positions must not change the matrix the assembly defined."""
    say("case deferred blocks and uploads")
    x4, bc, vbc, v4 = config0_state(6, 6, 8)
    mats = [material(0.01)]
    _, a = make_pair(x4, mats, v4=v4, bc=bc, vbc=vbc)
    _, b = make_pair(x4, mats, v4=v4, bc=bc, vbc=vbc)
    prm = pds.make_params(dt=1e-4, clock_scale=0.5, precond=1)
    a.assemble(prm)
    b.assemble(prm)
    moved = a.download("position")
    moved[:, :3] += 1e-4
    a.upload("position", moved)
    da, va = a.get_matrix()
    db, vb = b.get_matrix()
    assert np.array_equal(da, db) and np.array_equal(va, vb)
    a.close()
    b.close()
