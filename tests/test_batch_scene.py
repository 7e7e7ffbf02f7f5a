"""The batched-specimen scene (scene batch): specimen tables, padding rows,
draws within their ranges, and no bond across specimens (brute force)."""
import os

import numpy as np

import oracle as orc
import pdsolver as pds

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CFG = os.path.join(ROOT, "sci-solver_peridynamic_amd", "configs")


def test_batch_scene_layout():
    p = os.path.join(CFG, "batch_64.txt")
    desc = pds.describe(p)
    ex = pds.scene_extra(p)
    first, count, act = ex["first"], ex["count"], ex["act"]
    n = desc["num_pd_particle"]
    assert len(first) == 64 and first[0] == 0
    assert np.all(first % 256 == 0)
    assert np.all(first[1:] >= first[:-1] + count[:-1])
    assert np.all(first[1:] - (first[:-1] + count[:-1]) < 256)
    assert first[-1] + count[-1] == n
    inside = np.zeros(n, bool)
    for f, c in zip(first, count):
        inside[f:f + c] = True
    assert np.all(act[inside] == 0) and np.all(act[~inside] == 3)
    assert count.min() >= 0.9 * 5000 and count.max() <= 1.1 * 60000
    s0 = np.array([ex["s0"][f] for f in first])
    assert s0.min() >= 0.01 and s0.max() <= 0.03
    assert len(np.unique(s0)) == 64
    # log-uniform counts: the median sits near the geometric mean
    assert 0.5 * np.sqrt(5000 * 60000) < np.median(count) < 2.0 * np.sqrt(5000 * 60000)


def test_batch_small_has_no_cross_specimen_bonds():
    p = os.path.join(CFG, "batch_small.txt")
    desc = pds.describe(p)
    sc = pds.scene(p)
    ex = pds.scene_extra(p)
    keep = ex["act"] == 0
    spec = np.full(len(keep), -1)
    for k, (f, c) in enumerate(zip(ex["first"], ex["count"])):
        spec[f:f + c] = k
    idx = np.nonzero(keep)[0]
    counts, lists, total = orc.bonds_brute(sc["x4"][idx], [desc["materials"][0][0]])
    assert total > 0
    for r in range(len(idx)):
        nb = idx[lists[r, :counts[r]]]
        assert np.all(spec[nb] == spec[idx[r]])


def test_cg_operator_key(tmp_path):
    """cg_operator matrix_free reaches the step parameters (FP64 only)."""
    base = open(os.path.join(CFG, "tension_20.txt")).read()
    for text, want in ((base, 0), (base + "\ncg_operator matrix_free\n", 1),
                       (base + "\ncg_operator matrix_free\ncg_precision mixed\n", 0)):
        p = tmp_path / "params.txt"
        p.write_text(text)
        desc = pds.describe(str(p))
        assert desc["matrix_free"] == want
        assert pds.make_params(**pds.step_kwargs(desc, 1)).matrix_free == want
