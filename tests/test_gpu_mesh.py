"""The reference's own example on the engine: the bunny mesh filled with
particles (configs/bunny_break.txt, examples/params.txt of the reference:
scaled units, identity CG to rtol 1e-20, walls, gravity phase-in), stepped
on the GPU against the oracle, and driven end to end by peri_run.  Also the
all-pairs contact mode (the pair set of collidePDParticles) on the GPU."""
import json
import os
import subprocess
import sys
import tempfile

import numpy as np
import pytest

import oracle as orc
from pdhelpers import make_pair

pytestmark = pytest.mark.gpu
pds = pytest.importorskip("pdsolver")

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BUNNY = os.path.join(ROOT, "tests", "golden", "bunny.obj")
CFG = os.path.join(ROOT, "sci-solver_peridynamic_amd", "configs", "bunny_break.txt")
EXE = os.path.join(ROOT, "sci-solver_peridynamic_amd", "peri_run")


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


@pytest.mark.parametrize("mf", [0, 1])
def test_bunny_steps_match_oracle(mf):
    """The bunny scene (irregular, boundary-heavy neighbour lists) stepped
    with stored (mf 0) and recomputed (mf 1) bond blocks."""
    say("case bunny 30 steps, matrix_free", mf)
    desc = pds.describe(CFG, BUNNY)
    sc = pds.scene(CFG, BUNNY)
    n = desc["num_pd_particle"]
    mats = pds.materials_of(desc)
    assert n == 4782
    c, l, total = orc.bonds_brute(sc["x4"], [m[0] for m in mats])
    rs, col = orc.lists_to_csr(c, l)
    s00 = orc.stretch_init(n, desc["s0"], desc["stretch_seed"])
    S = orc.OracleSystem(sc["x4"], sc["v4"], s00, rs, col, mats)
    ctx = pds.Context(n, mats)
    pds.load_state(ctx, sc["x4"], sc["v4"], pds.stretch_init(n, desc["s0"], desc["stretch_seed"]))
    bst = ctx.build_bonds()
    assert bst.total_bonds == total
    for k in range(1, 31):
        kw = pds.step_kwargs(desc, k)
        so = S.step(orc.make_params(**kw))
        sg = ctx.step(pds.make_params(**dict(kw, matrix_free=mf)))
        assert (sg.cg_iterations, sg.broken_bonds) == (so.cg_iterations, so.broken_bonds), k
        assert np.array_equal(ctx.download("position")[:, :3], S.x[:, :3]), k
        if k % 10 == 0:
            say(dict(step=k, it=sg.cg_iterations, residual=sg.residual_norm, b=sg.b_norm))
    ctx.close()


def test_bunny_driver_runs():
    say("case bunny peri_run 200 steps")
    text = open(CFG).read() + "\nfinal_step 199\nsteps_per_frame 50\nprint_steps 1\n"
    with tempfile.TemporaryDirectory() as out:
        p = os.path.join(out, "params.txt")
        with open(p, "w") as f:
            f.write(text)
        r = subprocess.run([EXE, p, BUNNY, out], capture_output=True, text=True, timeout=300)
        if r.returncode != 0:
            say(r.stdout[-3000:], r.stderr[-3000:])
        assert r.returncode == 0
        recs = [json.loads(l) for l in r.stdout.splitlines() if l.startswith("{")]
        assert len(recs) == 200 and all(x["converged"] for x in recs)
        frames = sorted(os.listdir(os.path.join(out, "DATA", "PD_POSITION")),
                        key=lambda s: int(s.split(".")[1]))
        x = np.fromfile(os.path.join(out, "DATA", "PD_POSITION", frames[-1])).reshape(-1, 4)
        assert x.shape[0] == 4782 and np.isfinite(x).all()
        say(dict(frames=len(frames), it_first=recs[0]["cg_iterations"],
                 it_last=recs[-1]["cg_iterations"]))


def test_all_pairs_contact_bit_exact():
    say("case contact all pairs")
    rng = np.random.default_rng(21)
    n = 3000
    x4 = np.zeros((n, 4))
    x4[:, :3] = rng.uniform(0, 0.08, (n, 3))
    body = rng.integers(0, 2, n).astype(np.int32)
    mats = [(0.004, 0.002, 1.0, 1.0)] * 2
    sys_, ctx = make_pair(x4, mats, body=body)
    for mode in (0, 1):
        kw = dict(contact_radius=0.004, contact_stiffness=2.5e6, contact_volume=1.0,
                  contact_cell=0.005, contact_mode=mode)
        sys_.contact(orc.make_params(**kw))
        ctx.assemble(pds.make_params(**kw))
        fc = ctx.download("contact_force")
        assert np.abs(sys_.fc).max() > 0
        assert np.array_equal(fc[:, :3], sys_.fc[:, :3]), mode
    ctx.close()
