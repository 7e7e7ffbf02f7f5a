"""Benchmark: implicit steps/s of the 1M-particle impact workload on one GPU.

Workload (sci-solver_peridynamic_amd/configs/impact_100.txt): 100^3 jittered
lattice + 5,000-particle spherical projectile at 20 m/s, implicit Euler
dt = 5e-5 s, block-Jacobi CG to rtol 1e-6, FP64, 200 steps in all
(warmup + timed steps; the defaults cover the 200).

Prints one JSON line with the metric, the roofline of the dominant kernel
(the CG SpMV, timed with HIP events on the engine's stream) and a CPU
baseline (the CPU oracle on a bounded sample of the same workload).

    python bench.py [--gpus 1] [--steps 195] [--warmup 5] [--no-cpu-baseline]
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
PKG = os.path.join(ROOT, "sci-solver_peridynamic_amd")
sys.path.insert(0, PKG)

import numpy as np  # noqa: E402

import pdsolver as pds  # noqa: E402

CONFIG = os.path.join(PKG, "configs", "impact_100.txt")
HBM_PEAK_GBS = 8000.0
TRAFFIC_FILE = os.path.join(ROOT, "profiles", "cg_spmv_impact_100_traffic.txt")


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def read_traffic():
    """HBM bytes per CG SpMV launch from the committed rocprof counter runs."""
    try:
        with open(TRAFFIC_FILE) as f:
            for line in f:
                if line.startswith("hbm_bytes_per_launch"):
                    return float(line.split()[1])
    except OSError:
        pass
    return None


def gpu_run(steps, warmup, copy_calibration=0):
    desc = pds.describe(CONFIG)
    sc = pds.scene(CONFIG)
    n = desc["num_pd_particle"]
    mats = pds.materials_of(desc)
    s00 = pds.stretch_init(n, desc["s0"], desc["stretch_seed"])
    ctx = pds.Context(n, mats)
    pds.load_state(ctx, sc["x4"], sc["v4"], s00, body=sc["body"], bc=sc["bc"], vbc=sc["vbc"])
    bst = ctx.build_bonds()
    log(f"particles {n}, directed bonds {bst.total_bonds} ({bst.min_bonds}..{bst.max_bonds} per "
        f"particle), bond build {bst.ms_build:.1f} ms")
    step = 0
    for _ in range(warmup):
        step += 1
        ctx.step(pds.make_params(**pds.step_kwargs(desc, step)))
    ctx.synchronize()
    spmv_ms = 0.0
    spmv_n = 0
    iters = []
    broken = 0
    phase = dict(contact=0.0, assemble=0.0, solve=0.0, update=0.0)
    t0 = time.perf_counter()
    for _ in range(steps):
        step += 1
        st = ctx.step(pds.make_params(timing=1, **pds.step_kwargs(desc, step)))
        spmv_ms += st.ms_spmv_total
        spmv_n += st.spmv_launches
        iters.append(st.cg_iterations)
        broken += st.broken_bonds
        phase["contact"] += st.ms_contact
        phase["assemble"] += st.ms_assemble
        phase["solve"] += st.ms_solve
        phase["update"] += st.ms_update
    ctx.synchronize()
    elapsed = time.perf_counter() - t0
    spmv_bytes = ctx.spmv_bytes()
    live = st.live_bonds
    # after the timed region: the plain SpMV kernel y = A x alone
    plain_ms, plain_bytes = ctx.bench_spmv(20)
    for _ in range(copy_calibration):
        # device-to-device copies of known size (positions and velocities
        # 32 B per particle, bond tops 4 B): the scale of the HBM counters
        ctx.backup()
    if copy_calibration:
        log(f"copy calibration: {copy_calibration} x (2 copies of {32 * n} B + 1 of {4 * n} B)")
    ctx.close()
    return dict(n=n, bonds=int(bst.total_bonds), live=int(live), elapsed=elapsed, steps=steps,
                spmv_ms=spmv_ms, spmv_n=spmv_n, iters=iters, broken=int(broken),
                spmv_bytes=spmv_bytes, phase=phase, build_ms=bst.ms_build,
                plain_ms=plain_ms, plain_bytes=plain_bytes)


def cpu_baseline(sample_steps=2):
    """The CPU oracle on the same workload: kd-tree bonds per body (not
    timed), then `sample_steps` steps timed with a wall clock."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as orc
    desc = pds.describe(CONFIG)
    sc = pds.scene(CONFIG)
    n = desc["num_pd_particle"]
    mats = pds.materials_of(desc)
    body = sc["body"]
    counts = np.zeros(n, np.int32)
    lists = np.zeros((n, 528), np.int32)
    for b in range(len(mats)):
        idx = np.nonzero(body == b)[0]
        c, l, dropped = orc.bonds_kdtree(sc["x4"][idx], mats[b][0])
        counts[idx] = c
        lists[idx] = np.where(np.arange(528)[None, :] < c[:, None], idx[l], 0)
        del l
    rs, col = orc.lists_to_csr(counts, lists)
    del lists
    s00 = orc.stretch_init(n, desc["s0"], desc["stretch_seed"])
    S = orc.OracleSystem(sc["x4"], sc["v4"], s00, rs, col, mats, body=body, bc=sc["bc"],
                         vbc=sc["vbc"])
    # at least `sample_steps` steps, continuing up to ~10 s of CPU work
    t0 = time.perf_counter()
    k = 0
    while k < sample_steps or (time.perf_counter() - t0 < 10.0 and k < 20):
        k += 1
        S.step(orc.make_params(**pds.step_kwargs(desc, k)))
    dt = time.perf_counter() - t0
    sample_steps = k
    threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    return dict(value=sample_steps / dt, unit="steps/s", cores=threads, kind="port",
                sample=f"steps 1-{sample_steps} of the 200-step impact_100 run "
                       f"({n} particles), CPU oracle (C/OpenMP restatement) after its "
                       f"own kd-tree bond build, wall clock {dt:.2f} s")


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=195)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-steps", type=int, default=2)
    ap.add_argument("--copy-calibration", type=int, default=0,
                    help="after the run, N rounds of device copies of known size")
    a = ap.parse_args()
    if a.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py runs on one GPU: the implicit peridynamics path is built and "
              "measured on a single MI355X; multi-GPU sharding is design-only.",
              file=sys.stderr)
        sys.exit(2)
    r = gpu_run(a.steps, a.warmup, a.copy_calibration)
    ms_step = 1000.0 * r["elapsed"] / r["steps"]
    spmv_avg_ms = r["spmv_ms"] / max(r["spmv_n"], 1)
    achieved = r["spmv_bytes"] / (spmv_avg_ms * 1e-3) / 1e9
    traffic = read_traffic()
    out = {
        "metric": "implicit steps/s @1M particles, CG rtol 1e-6; SpMV HBM GB/s vs ~8 TB/s peak",
        "value": r["steps"] / r["elapsed"],
        "unit": "steps/s",
        "n_gpus": 1,
        "steps": r["steps"],
        "warmup": a.warmup,
        "ms_per_step": ms_step,
        "higher_is_better": True,
        "scaling": "strong",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (seeded jittered lattice + projectile scene)",
        "config": {
            "workload": "impact_100: 100^3 jittered lattice + 5000-particle projectile at "
                        "20 m/s, implicit Euler dt 5e-5 s, block-Jacobi CG rtol 1e-6, FP64",
            "particles": r["n"],
            "directed_bonds": r["bonds"],
            "live_bonds_end": r["live"],
            "cg_iterations_mean": float(np.mean(r["iters"])),
            "cg_iterations_max": int(np.max(r["iters"])),
            "broken_bonds": r["broken"],
            "phase_ms_per_step": {k: v / r["steps"] for k, v in r["phase"].items()},
            "bond_build_ms": r["build_ms"],
            "plain_spmv_gbs": r["plain_bytes"] / (r["plain_ms"] * 1e-3) / 1e9,
        },
        "roofline": {
            "kernel": "k_cg_spmv (CG SpMV y = A p + <y,p> partial)",
            "bound": "hbm",
            "achieved": achieved,
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBS,
            "traffic": traffic,
            "bytes_per_launch": r["spmv_bytes"],
            "avg_launch_ms": spmv_avg_ms,
            "launches": r["spmv_n"],
        },
    }
    if not a.no_cpu_baseline:
        log("cpu baseline: CPU oracle on the same workload ...")
        out["cpu_baseline"] = cpu_baseline(a.cpu_steps)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
