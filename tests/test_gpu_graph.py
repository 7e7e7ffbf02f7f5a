"""Captured steps (hipGraph replay, graph = 1) and the SpMV autotuner's
variants must give results identical to the eager path: every replay is
checked against an eager context, with downloads and host waits between
replays, and switching between the two paths mid-run changes nothing."""
import os
import sys

import numpy as np
import pytest

import oracle as orc
from pdhelpers import deformed_pair

pytestmark = pytest.mark.gpu
pds = pytest.importorskip("pdsolver")

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BATCH = os.path.join(ROOT, "sci-solver_peridynamic_amd", "configs", "batch_small.txt")


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def batch_context():
    desc = pds.describe(BATCH)
    sc = pds.scene(BATCH)
    ex = pds.scene_extra(BATCH)
    n = desc["num_pd_particle"]
    ctx = pds.Context(n, pds.materials_of(desc))
    pds.load_state(ctx, sc["x4"], sc["v4"], pds.stretch_init_each(ex["s0"], desc["stretch_seed"]),
                   body=sc["body"], act=ex["act"], bc=sc["bc"], vbc=sc["vbc"])
    ctx.build_bonds()
    ctx.set_specimens(ex["first"], ex["count"])
    return desc, ex, ctx


@pytest.mark.parametrize("precision", [0, 1])
def test_graph_replays_match_eager_steps(precision):
    say("case graph precision", precision)
    desc, ex, eager = batch_context()
    _, _, graph = batch_context()
    nspec = len(ex["first"])
    for k in range(1, 41):
        kw = dict(pds.step_kwargs(desc, k), precision=precision)
        # switch the graph context to eager for a few steps mid-run
        use_graph = 0 if 18 <= k <= 20 else 1
        se = eager.step(pds.make_params(**kw))
        sg = graph.step(pds.make_params(graph=use_graph, **kw))
        assert (sg.cg_iterations, sg.broken_bonds, sg.live_bonds) == \
            (se.cg_iterations, se.broken_bonds, se.live_bonds), k
        assert eager.specimen_stats(nspec) == graph.specimen_stats(nspec), k
        assert np.array_equal(eager.download("position"), graph.download("position")), k
        assert np.array_equal(eager.download("velocity"), graph.download("velocity")), k
        if k % 10 == 0:
            say(dict(step=k, it=sg.cg_iterations, syncs=sg.host_syncs, broken=sg.broken_bonds))
    eager.close()
    graph.close()


def test_graph_rejects_contact():
    say("case graph contact")
    desc, ex, ctx = batch_context()
    kw = dict(pds.step_kwargs(desc, 1), contact_radius=0.004, contact_stiffness=1.0,
              contact_cell=0.005, graph=1)
    with pytest.raises(pds.PDError) as e:
        ctx.step(pds.make_params(**kw))
    assert e.value.code == pds.PD_ERR_ARG
    ctx.close()


@pytest.mark.parametrize("precision", [0, 1])
def test_spmv_variants_identical(precision):
    say("case spmv variants precision", precision)
    sys_, ctx = deformed_pair(s_fixed=10.0)
    kw = dict(dt=2e-4, precond=1, rtol=1e-10, precision=precision)
    sys_.assemble(orc.make_params(**kw))
    (sys_.cg_mixed if precision else sys_.cg)(orc.make_params(**kw))
    x0 = ctx.download("solution")
    for u in (2, 4, 8):
        ctx.upload("solution", x0)
        ctx.assemble(pds.make_params(spmv_unroll=u, **kw))
        st = ctx.solve(pds.make_params(spmv_unroll=u, **kw))
        assert st.converged
        assert np.array_equal(ctx.download("solution")[:, :3], sys_.sol[:, :3]), u
    ms, best = ctx.tune_spmv(3)
    say(dict(ms=ms.tolist(), best=best))
    assert best in (2, 4, 8) and np.all(ms > 0)
    ctx.close()
