"""Train the learned block preconditioner on systems the simulator assembles.

Runs the 1M-particle impact workload (configs/impact_100.txt) with `--runs`
different scene seeds (lattice, projectile and stretch-limit seeds
101, 102, ...; the shipped config uses 1973 and 2024, so evaluation runs
are unseen), collects (features, D_p, (A e)_p, e_p) samples on every
`--stride`-th row at the snapshot steps, fits the network with the engine's
trainer and writes the parameters (one float per line, the layout of
PD_LEARNED_PARAMS) to --out.

    python train_precond.py [--runs 20] [--steps 120] [--snapshots 40,80,120]
                            [--stride 64] [--fit-steps 3000] [--batch 8192]
                            [--lr 1e-3] [--out networks/impact_precond.txt]
"""
from __future__ import annotations

import argparse
import os
import sys
import tempfile
import time

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, HERE)

import numpy as np  # noqa: E402

import pdsolver as pds  # noqa: E402

CONFIG = os.path.join(HERE, "configs", "impact_100.txt")


def seeded_config(seed):
    text = open(CONFIG).read().rstrip("\n")
    text += f"\nlattice_seed {seed}\nprojectile_seed {seed}\nstretch_seed {seed}\n"
    fd, path = tempfile.mkstemp(suffix=".txt", prefix="impact_seed_")
    with os.fdopen(fd, "w") as f:
        f.write(text)
    return path


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--runs", type=int, default=20)
    ap.add_argument("--steps", type=int, default=120)
    ap.add_argument("--snapshots", default="40,80,120")
    ap.add_argument("--stride", type=int, default=64)
    ap.add_argument("--fit-steps", type=int, default=3000)
    ap.add_argument("--batch", type=int, default=8192)
    ap.add_argument("--lr", type=float, default=1e-3)
    ap.add_argument("--out", default=os.path.join(HERE, "networks", "impact_precond.txt"))
    a = ap.parse_args()
    snaps = {int(s) for s in a.snapshots.split(",")}
    trainer = None
    t0 = time.perf_counter()
    for run in range(a.runs):
        seed = 101 + run
        path = seeded_config(seed)
        try:
            desc = pds.describe(path)
            sc = pds.scene(path)
        finally:
            os.unlink(path)
        n = desc["num_pd_particle"]
        if trainer is None:
            cap = a.runs * len(snaps) * (n // a.stride + 1)
            trainer = pds.Trainer(cap, seed=1)
        ctx = pds.Context(n, pds.materials_of(desc))
        pds.load_state(ctx, sc["x4"], sc["v4"], pds.stretch_init(n, desc["s0"], desc["stretch_seed"]),
                       body=sc["body"], bc=sc["bc"], vbc=sc["vbc"])
        ctx.build_bonds()
        its = []
        for k in range(1, a.steps + 1):
            st = ctx.step(pds.make_params(**pds.step_kwargs(desc, k)))
            its.append(st.cg_iterations)
            if k in snaps:
                trainer.add(ctx, seed * 1000 + k, a.stride)
        ctx.close()
        log(f"run {run + 1}/{a.runs} (seed {seed}): {n} particles, mean CG iterations "
            f"{np.mean(its):.2f}, {time.perf_counter() - t0:.0f} s")
    losses = trainer.fit(a.fit_steps, a.batch, a.lr, seed=7)
    log(f"fit: loss {losses[:50].mean():.4e} -> {losses[-50:].mean():.4e} over {a.fit_steps} steps")
    p = trainer.params()
    trainer.close()
    os.makedirs(os.path.dirname(a.out), exist_ok=True)
    with open(a.out, "w") as f:
        for v in p:
            f.write(f"{float(v):.9g}\n")
    log(f"wrote {a.out}")
    print(f'{{"runs": {a.runs}, "loss_first": {float(losses[:50].mean())}, '
          f'"loss_last": {float(losses[-50:].mean())}, "seconds": {time.perf_counter() - t0:.1f}}}')


if __name__ == "__main__":
    main()
