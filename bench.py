"""Benchmark: implicit steps/s of the 1M-particle impact workload on one GPU.

Workload (sci-solver_peridynamic_amd/configs/impact_100.txt): 100^3 jittered
lattice + 5,000-particle spherical projectile at 20 m/s, implicit Euler
dt = 5e-5 s, block-Jacobi CG to rtol 1e-6, FP64, 200 steps in all.

The projectile reaches the target at step 37 (first contact force) and the
first bonds break at step 41.  The steps before the impact run untimed (at
least --warmup of them), so the timed window of --steps steps starts at the
impact: contact and bond breaking are always inside it.  A separate
"full_run" object times all 200 steps from the first.

Prints one JSON line with the metric, the roofline of the dominant kernel
(the CG SpMV, timed with HIP events on the engine's stream) and a CPU
baseline (the CPU oracle on the same timed window of the same workload).

    python bench.py [--gpus 1] [--steps 164] [--warmup 5] [--no-cpu-baseline]
                    [--precision fp64|mixed] [--no-mixed]

The FP64 line also carries a "mixed_precision" object (the same run with the
float inner CG and FP64 refinement, its SpMV roofline, and its per-step
position error against the FP64 path over all 200 steps), a "matrix_free"
object (the same FP64 run with the blocks recomputed in every SpMV: identical
results), the learned preconditioner rows and a "batched_specimens" object
(64 independent tension specimens in one context, aggregate particle-steps/s:
FP64 with the matrix-free operator, and the mixed solver captured and eager).
"""
from __future__ import annotations

import argparse
import hashlib
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
# impact_100: the first step with a contact force (the CPU oracle's run:
# contact from step 37, first broken bonds at step 41); the timed window
# starts here unless --warmup asks for more untimed steps
IMPACT_STEP = 37
RUN_STEPS = 200
HBM_PEAK_GBS = 8000.0
TRAFFIC_FILES = {0: os.path.join(ROOT, "profiles", "cg_spmv_impact_100_traffic.txt"),
                 1: os.path.join(ROOT, "profiles", "mx_spmv_impact_100_traffic.txt"),
                 "mf": os.path.join(ROOT, "profiles", "mf_spmv_impact_100_traffic.txt")}


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def read_traffic(precision=0):
    """HBM bytes per CG SpMV launch (FP64 k_cg_spmv, the float inner
    k_mx_spmv for precision 1, the recomputing k_cg_spmv_mf for "mf") from
    the committed rocprof counter runs."""
    try:
        with open(TRAFFIC_FILES[precision]) as f:
            for line in f:
                if line.startswith("hbm_bytes_per_launch"):
                    return float(line.split()[1])
    except OSError:
        pass
    return None


def read_counters(key, kernel):
    """The counter row of `kernel` in a committed rocprof summary (the
    columns of its utilisation table), as fractions and per-launch counts;
    None when the summary holds no such row."""
    try:
        with open(TRAFFIC_FILES[key]) as f:
            for line in f:
                cells = [c.strip() for c in line.strip().strip("|").split("|")]
                if len(cells) == 12 and cells[0] == f"`pd::{kernel}`":
                    v = [float(c) for c in cells[1:]]
                    return {"valu_busy": v[2] / 100.0, "waves_waiting": v[3] / 100.0,
                            "valu_instr_per_launch": v[4] * 1e6,
                            "fp64_transcendental_per_launch": v[5] * 1e6,
                            "ta_busy": v[7] / 100.0, "td_busy": v[8] / 100.0,
                            "source": os.path.relpath(TRAFFIC_FILES[key], ROOT)}
    except (OSError, ValueError):
        pass
    return None


def new_context(desc, sc):
    n = desc["num_pd_particle"]
    s00 = pds.stretch_init(n, desc["s0"], desc["stretch_seed"])
    ctx = pds.Context(n, pds.materials_of(desc))
    pds.load_state(ctx, sc["x4"], sc["v4"], s00, body=sc["body"], bc=sc["bc"], vbc=sc["vbc"])
    return ctx, ctx.build_bonds()


def step_params(desc, step, precision, **kw):
    return pds.make_params(**dict(pds.step_kwargs(desc, step), precision=precision, **kw))


NETWORK = os.path.join(PKG, "networks", "impact_precond.txt")
# why a block-diagonal correction cannot cut iterations here (the derivation
# is in the package's design notes): for block-diagonal M the probe loss is
# minimised by M*_p = D_p (D_p^2 + sum_q A_pq A_qp)^-1; at impact_100,
# dt^2 sum_q K_pq ~ 0.8 m and each off-diagonal block ~ dt^2 k ~ 2.5e-5, so
# sum_q A_pq^2 / D_p^2 ~ 0.02 and M* lies within ~2 % of D_p^-1 (block Jacobi)
LEARNED_BOUND = ("block-diagonal M: the probe-loss optimum M*_p = D_p (D_p^2 + sum_q A_pq A_qp)^-1 "
                 "is within ~2% of block Jacobi D_p^-1 at impact_100 (sum_q A_pq^2 / D_p^2 ~ 0.02), "
                 "so the network cannot cut the 4-6 block-Jacobi iterations per step")


def gpu_run(steps, untimed, precision=0, copy_calibration=0, precond=None, network=None,
            matrix_free=0, tune=True):
    """`untimed` steps, then `steps` timed steps (steps untimed + 1 ..
    untimed + steps of the run)."""
    desc = pds.describe(CONFIG)
    sc = pds.scene(CONFIG)
    n = desc["num_pd_particle"]
    ctx, bst = new_context(desc, sc)
    extra = {}
    if precond is not None:
        extra["precond"] = precond
    if matrix_free:
        extra["matrix_free"] = matrix_free
    if network is not None:
        ctx.learned_set(network)
    opname = {0: "", 1: " matrix-free", 2: " operator per step"}[matrix_free]
    log(f"[{'mixed' if precision else 'fp64'}{opname} "
        f"precond {extra.get('precond', desc['precond'])}] "
        f"particles {n}, directed bonds {bst.total_bonds} "
        f"({bst.min_bonds}..{bst.max_bonds} per particle), bond build {bst.ms_build:.1f} ms, "
        f"timed steps {untimed + 1}..{untimed + steps}")
    step = 0
    for _ in range(untimed):
        step += 1
        ctx.step(step_params(desc, step, precision, **extra))
    ctx.synchronize()
    # the autotuner picks the SpMV variant on the resident matrix (outside
    # the timed region); spmv_unroll = 0 then uses it
    # (the matrix-free operator has no stored-block variants to tune)
    tune_ms, unroll = (ctx.tune_spmv(10) if tune and untimed > 0 and matrix_free == 0
                       else (np.zeros(3), 0))
    spmv_ms = 0.0
    spmv_n = 0
    store_ms = 0.0
    resid_ms = 0.0
    resid_n = 0
    iters = []
    refinements = []
    fallbacks = 0
    probes = 0
    mf_steps = 0
    broken = 0
    phase = dict(contact=0.0, assemble=0.0, solve=0.0, update=0.0)
    first_timed = step + 1
    # the step parameters are built before the timed region, as a compiled
    # caller holds them; the statistics are summed after it
    prms = [step_params(desc, first_timed + i, precision, timing=1, **extra)
            for i in range(steps)]
    t0 = time.perf_counter()
    stats = [ctx.step(prm) for prm in prms]
    ctx.synchronize()
    elapsed = time.perf_counter() - t0
    step += steps
    for st in stats:
        fallbacks += st.precond_fallbacks
        probes += st.precond_probe
        mf_steps += st.operator_mf
        spmv_ms += st.ms_spmv_total
        spmv_n += st.spmv_launches
        store_ms += st.ms_spmv_store
        resid_ms += st.ms_resid_total
        resid_n += st.resid_launches
        iters.append(st.cg_iterations)
        refinements.append(st.refinements)
        broken += st.broken_bonds
        phase["contact"] += st.ms_contact
        phase["assemble"] += st.ms_assemble
        phase["solve"] += st.ms_solve
        phase["update"] += st.ms_update
    spmv_bytes = ctx.spmv_bytes()
    live = st.live_bonds
    # final state digest: equal digests = bit-identical trajectories
    digest = hashlib.sha256(ctx.download("position").tobytes() +
                            ctx.download("velocity").tobytes()).hexdigest()[:16]
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
                plain_ms=plain_ms, plain_bytes=plain_bytes, resid_ms=resid_ms, resid_n=resid_n,
                refinements=refinements, unroll=unroll, tune_ms=[float(t) for t in tune_ms],
                fallbacks=fallbacks, probes=probes, digest=digest, store_ms=store_ms,
                mf_steps=mf_steps,
                timed_steps=[first_timed, step])


BATCH_CONFIG = os.path.join(PKG, "configs", "batch_64.txt")


def batched_run(steps, warmup, graph, **over):
    """batch_64: 64 independent tension specimens in one context (the
    config's float inner CG with FP64 refinement unless `over` sets
    precision / matrix_free); aggregate particle-steps/s over the
    specimens' particles (padding rows excluded)."""
    desc = pds.describe(BATCH_CONFIG)
    sc = pds.scene(BATCH_CONFIG)
    ex = pds.scene_extra(BATCH_CONFIG)
    n = desc["num_pd_particle"]
    ctx = pds.Context(n, pds.materials_of(desc))
    pds.load_state(ctx, sc["x4"], sc["v4"], pds.stretch_init_each(ex["s0"], desc["stretch_seed"]),
                   body=sc["body"], act=ex["act"], bc=sc["bc"], vbc=sc["vbc"])
    bst = ctx.build_bonds()
    ctx.set_specimens(ex["first"], ex["count"])
    nspec = len(ex["first"])
    real = int(ex["count"].sum())
    step = 0
    for _ in range(warmup):
        step += 1
        ctx.step(pds.make_params(**dict(pds.step_kwargs(desc, step), graph=graph, **over)))
    ctx.synchronize()
    # the matrix-free operator has no stored-block variants to tune
    tune_ms, unroll = (ctx.tune_spmv(10) if not over.get("matrix_free")
                       else (np.zeros(3), 0))
    its, spec_its, broken, syncs = [], [], 0, 0
    t0 = time.perf_counter()
    for _ in range(steps):
        step += 1
        st = ctx.step(pds.make_params(**dict(pds.step_kwargs(desc, step), graph=graph, **over)))
        its.append(st.cg_iterations)
        broken += st.broken_bonds
        syncs += st.host_syncs
    ctx.synchronize()
    elapsed = time.perf_counter() - t0
    per = ctx.specimen_stats(nspec)
    spec_its = [p["cg_iterations"] for p in per]
    ctx.close()
    log(f"[batch graph={graph} {over}] {nspec} specimens, {real} particles ({n} rows), "
        f"{steps / elapsed:.1f} steps/s")
    return dict(specimens=nspec, particles=real, rows=n, bonds=int(bst.total_bonds),
                steps=steps, elapsed=elapsed, value=real * steps / elapsed,
                ms_per_step=1000.0 * elapsed / steps, cg_iterations_max_mean=float(np.mean(its)),
                last_step_iterations_min_max=[min(spec_its), max(spec_its)], broken_bonds=broken,
                host_syncs_per_step=syncs / steps, spmv_unroll=unroll,
                tune_ms=[float(t) for t in tune_ms],
                count_min_max=[int(ex["count"].min()), int(ex["count"].max())])


def mixed_error(steps):
    """Steps 1..steps of the FP64 and the mixed-precision path side by side:
    per-step max |x_mixed - x_fp64| / max |x_fp64| over all coordinates."""
    desc = pds.describe(CONFIG)
    sc = pds.scene(CONFIG)
    c64, _ = new_context(desc, sc)
    c32, _ = new_context(desc, sc)
    errs = []
    broken = [0, 0]
    split = None  # first step whose broken-bond count differs
    for k in range(1, steps + 1):
        s64 = c64.step(step_params(desc, k, 0))
        s32 = c32.step(step_params(desc, k, 1))
        broken[0] += s64.broken_bonds
        broken[1] += s32.broken_bonds
        if split is None and s64.broken_bonds != s32.broken_bonds:
            split = k
        x64 = c64.download("position")[:, :3]
        x32 = c32.download("position")[:, :3]
        errs.append(float(np.abs(x32 - x64).max() / np.abs(x64).max()))
        if k % 50 == 0:
            log(f"mixed vs fp64 step {k}: rel position error {errs[-1]:.3e}")
    c64.close()
    c32.close()
    before = errs[:split - 1] if split else errs
    return dict(max_rel=max(errs), final_rel=errs[-1], steps=steps,
                first_step_broken_counts_differ=split,
                max_rel_before_that_step=max(before) if before else None,
                per_step_rel_every_10=errs[9::10], broken_bonds_fp64=broken[0],
                broken_bonds_mixed=broken[1])


def spmv_roofline(r, kernel, traffic):
    avg_ms = r["spmv_ms"] / max(r["spmv_n"], 1)
    achieved = r["spmv_bytes"] / (avg_ms * 1e-3) / 1e9
    return {
        "kernel": kernel,
        "bound": "hbm",
        "achieved": achieved,
        "peak": HBM_PEAK_GBS,
        "unit": "GB/s",
        "frac": achieved / HBM_PEAK_GBS,
        "traffic": traffic,
        "bytes_per_launch": r["spmv_bytes"],
        "avg_launch_ms": avg_ms,
        "launches": r["spmv_n"],
    }


def step_roofline(r):
    """Whole-step algorithmic HBM bytes over the step time, stored FP64
    blocks: B_step = B_asm + iterations x (B_spmv + B_vec) + N x 96 with
    B_asm = live blocks x (6 x 8 + 4) + N x (24 + 24 + 8 + 8 + 24 + 8) (blocks
    and columns written / read; x, x0, s, s00 read, f, s_new written) and
    B_vec = N x (10 x 24 + 72) per iteration (the fused CG vector work)."""
    n, live = r["n"], r["live"]
    its = float(np.mean(r["iters"]))
    b_asm = live * (6 * 8 + 4) + n * (24 + 24 + 8 + 8 + 24 + 8)
    b_vec = n * (10 * 24 + 72)
    b_step = b_asm + its * (r["spmv_bytes"] + b_vec) + n * 96
    t_step = r["elapsed"] / r["steps"]
    achieved = b_step / t_step / 1e9
    return {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBS, "bytes_per_step": b_step,
            "ms_per_step": 1e3 * t_step, "cg_iterations_mean": its}


def mf_step_roofline(r):
    """The same whole-step view with the matrix-free operator: B_asm = live
    bonds x (4 + 8) + N x 250 (columns and reference lengths read; per row
    the record, x, v, s, s00, b, diagonal, inverse, f, s_new and the fused
    start), each product's algorithmic bytes (no stored blocks), B_vec as
    above."""
    n, live = r["n"], r["live"]
    its = float(np.mean(r["iters"]))
    b_asm = live * (4 + 8) + n * 250
    b_vec = n * (10 * 24 + 72)
    b_step = b_asm + its * (r["spmv_bytes"] + b_vec) + n * 96
    t_step = r["elapsed"] / r["steps"]
    achieved = b_step / t_step / 1e9
    return {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBS, "bytes_per_step": b_step,
            "ms_per_step": 1e3 * t_step, "cg_iterations_mean": its}


# FP64 operations per live bond of the matrix-free SpMV (bond_dir 12,
# bond_jac 27, the dt^2 scaling 12, the 3x3 product 15, the row sum 3;
# a division or square root counts as one) and per row (diagonal product
# 15, the <y, p> term 5)
MF_FLOPS_BOND = 69.0
MF_FLOPS_ROW = 20.0
FP64_PEAK_TFLOPS = 78.6
# MI355X: 256 CUs of four SIMDs, 2.4 GHz peak engine clock
SIMDS = 1024
CLOCK_HZ = 2.4e9


def mf_roofline(r):
    avg_ms = r["spmv_ms"] / max(r["spmv_n"], 1)
    flops = MF_FLOPS_BOND * r["live"] + MF_FLOPS_ROW * r["n"]
    tflops = flops / (avg_ms * 1e-3) / 1e12
    gbs = r["spmv_bytes"] / (avg_ms * 1e-3) / 1e9
    intensity = flops / r["spmv_bytes"]
    ridge = FP64_PEAK_TFLOPS * 1e3 / HBM_PEAK_GBS
    cnt = read_counters("mf", "k_cg_spmv_mf")
    floor_ms = None
    if cnt is not None:
        cycles = 4.0 * cnt["valu_instr_per_launch"] + 12.0 * cnt["fp64_transcendental_per_launch"]
        floor_ms = cycles / (SIMDS * CLOCK_HZ) * 1e3
    return {
        "kernel": "k_cg_spmv_mf (matrix-free CG SpMV: blocks recomputed from positions)",
        # its FP64 operations per algorithmic byte lie below the ridge of
        # the FP64 and HBM peaks, so HBM is its roofline ceiling; measured,
        # the latency of its dependent column -> record gathers holds it
        # below that (texture address / data units 84 / 91 % busy, VALU 68 %)
        "bound": "hbm",
        "achieved": gbs,
        "peak": HBM_PEAK_GBS,
        "unit": "GB/s",
        "frac": gbs / HBM_PEAK_GBS,
        # HBM bytes per launch from the calibrated counter runs
        "traffic": read_traffic("mf"),
        "hbm_bytes_per_launch": r["spmv_bytes"],
        "flops_per_launch": flops,
        "flops_per_byte": intensity,
        "ridge_flops_per_byte": ridge,
        "fp64_tflops": tflops,
        "fp64_frac": tflops / FP64_PEAK_TFLOPS,
        "avg_launch_ms": avg_ms,
        "launches": r["spmv_n"],
        # what holds it below that ceiling, from the committed counter runs
        "counters": cnt,
        # the time its VALU instructions alone need: a wave64 instruction
        # issues over four cycles of a SIMD, an FP64 reciprocal or square
        # root over sixteen, on 1,024 SIMDs at 2.4 GHz
        "valu_issue_floor_ms": floor_ms,
        "valu_issue_frac": floor_ms / avg_ms if floor_ms is not None else None,
    }


def cpu_model():
    """CPU model name and logical CPU count of this host."""
    model = "unknown"
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    return model, os.cpu_count()


def cpu_baseline(first, count):
    """The CPU oracle on the same workload and the GPU line's timed window:
    kd-tree bonds per body and steps 1 .. first - 1 untimed, then steps
    first .. first + count - 1 timed with a wall clock."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as orc
    desc = pds.describe(CONFIG)
    sc = pds.scene(CONFIG)
    n = desc["num_pd_particle"]
    mats = pds.materials_of(desc)
    body = sc["body"]
    counts = np.zeros(n, np.int32)
    lists = np.zeros((n, 128), np.int32)
    for b in range(len(mats)):
        idx = np.nonzero(body == b)[0]
        c, l, dropped = orc.bonds_kdtree(sc["x4"][idx], mats[b][0], max_bonds=128)
        assert dropped == 0
        counts[idx] = c
        lists[idx] = np.where(np.arange(128)[None, :] < c[:, None], idx[l], 0)
        del l
    rs, col = orc.lists_to_csr(counts, lists)
    del lists
    s00 = orc.stretch_init(n, desc["s0"], desc["stretch_seed"])
    S = orc.OracleSystem(sc["x4"], sc["v4"], s00, rs, col, mats, body=body, bc=sc["bc"],
                         vbc=sc["vbc"])
    t_pre = time.perf_counter()
    for k in range(1, first):
        S.step(orc.make_params(**pds.step_kwargs(desc, k)))
    t_pre = time.perf_counter() - t_pre
    broken = 0
    t0 = time.perf_counter()
    for k in range(first, first + count):
        broken += S.step(orc.make_params(**pds.step_kwargs(desc, k))).broken_bonds
    dt = time.perf_counter() - t0
    threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    model, logical = cpu_model()
    return dict(value=count / dt, unit="steps/s", cores=threads, kind="port",
                cpu_model=model, logical_cpus=logical,
                sample=f"steps {first}-{first + count - 1} of the 200-step impact_100 run "
                       f"({n} particles; contact from step {IMPACT_STEP}, {broken} bonds broken "
                       f"in the sample), CPU oracle (C/OpenMP restatement, {threads} threads) "
                       f"after its own kd-tree bond build and steps 1-{first - 1} untimed "
                       f"({t_pre:.1f} s); wall clock {dt:.2f} s")


OPERATORS = {"stored": 0, "mf": 1, "auto": 2}


def operator_row(x, ref_value, ref_digest):
    """A comparison row: the same FP64 run with another operator."""
    rate = x["steps"] / x["elapsed"]
    row = {
        "value": rate,
        "unit": "steps/s",
        "speedup_vs_timed_line": rate / ref_value,
        "ms_per_step": 1000.0 * x["elapsed"] / x["steps"],
        "identical_to_timed_line": x["digest"] == ref_digest,
        "final_state_digest": x["digest"],
        "cg_iterations_mean": float(np.mean(x["iters"])),
        "phase_ms_per_step": {k: v / x["steps"] for k, v in x["phase"].items()},
    }
    return row


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=RUN_STEPS - IMPACT_STEP + 1)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--precision", choices=["fp64", "mixed"], default="fp64",
                    help="solver of the timed line (mixed: float inner CG, FP64 refinement)")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-mixed", action="store_true",
                    help="skip the mixed-precision report of the fp64 line")
    ap.add_argument("--no-batched", action="store_true",
                    help="skip the batched-specimen report")
    ap.add_argument("--no-learned", action="store_true",
                    help="skip the learned-preconditioner report")
    ap.add_argument("--no-stored", action="store_true",
                    help="skip the stored-block FP64 run (the HBM SpMV half of the metric)")
    ap.add_argument("--operator", choices=list(OPERATORS), default="auto",
                    help="FP64 operator of the timed line: auto (per step the cheaper of the "
                         "two, measured), stored (3x3 BSR blocks streamed by every SpMV) or mf "
                         "(blocks recomputed in every SpMV); the results are identical")
    ap.add_argument("--cpu-steps", type=int, default=20,
                    help="steps of the timed window the CPU oracle times (at most --steps)")
    ap.add_argument("--no-full-run", action="store_true",
                    help="skip the separately timed run of all 200 steps")
    ap.add_argument("--copy-calibration", type=int, default=0,
                    help="after the run, N rounds of device copies of known size")
    a = ap.parse_args()
    if a.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py runs on one GPU: the implicit peridynamics path is built and "
              "measured on a single MI355X; multi-GPU sharding is design-only.",
              file=sys.stderr)
        sys.exit(2)
    prec = 1 if a.precision == "mixed" else 0
    op = OPERATORS[a.operator] if not prec else 0
    # the pre-impact steps run untimed: the timed window starts at the impact
    untimed = max(a.warmup, IMPACT_STEP - 1)
    r = gpu_run(a.steps, untimed, prec, a.copy_calibration, matrix_free=op)
    # the dominant kernel of the timed steps: the recomputing CG SpMV when
    # most steps used it, else the streaming one
    mf = r["mf_steps"] * 2 > r["steps"]
    ms_step = 1000.0 * r["elapsed"] / r["steps"]
    kernel = ("k_mx_spmv (float inner CG SpMV y = A32 p + <y,p> partial)" if prec else
              "k_cg_spmv (CG SpMV y = A p + <y,p> partial)")
    opdesc = {0: "stored 3x3 BSR blocks", 1: "blocks recomputed in every SpMV",
              2: "per step the cheaper of stored and recomputed blocks (measured)"}[op]
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
        "dtype": "f32+f64" if prec else "f64",
        "data": "synthetic (seeded jittered lattice + projectile scene)",
        "config": {
            "workload": "impact_100: 100^3 jittered lattice + 5000-particle projectile at "
                        "20 m/s, implicit Euler dt 5e-5 s, block-Jacobi CG rtol 1e-6, " +
                        ("float inner CG with FP64 residual / refinement" if prec else
                         "FP64, operator: " + opdesc),
            "untimed_steps": untimed,
            "timed_steps": r["timed_steps"],
            "particles": r["n"],
            "directed_bonds": r["bonds"],
            "live_bonds_end": r["live"],
            "cg_iterations_mean": float(np.mean(r["iters"])),
            "cg_iterations_max": int(np.max(r["iters"])),
            "broken_bonds": r["broken"],
            "steps_recomputing_blocks": r["mf_steps"],
            "phase_ms_per_step": {k: v / r["steps"] for k, v in r["phase"].items()},
            "bond_build_ms": r["build_ms"],
        },
        "roofline": (mf_roofline(r) if mf else
                     spmv_roofline(r, kernel, read_traffic(prec))),
    }
    if not prec and mf:
        out["step_roofline"] = mf_step_roofline(r)
    if not prec and not mf:
        out["step_roofline"] = step_roofline(r)
        out["config"]["first_product_with_block_stores_ms"] = r["store_ms"] / r["steps"]
        out["config"]["spmv_unroll"] = r["unroll"]
        out["config"]["spmv_tune_ms_2_4_8"] = r["tune_ms"]
    if prec:
        out["config"]["refinements_mean"] = float(np.mean(r["refinements"]))
        out["config"]["resid_spmv_ms_mean"] = r["resid_ms"] / max(r["resid_n"], 1)
        out["config"]["plain_spmv_gbs"] = r["plain_bytes"] / (r["plain_ms"] * 1e-3) / 1e9
    if not prec and op != 0 and not a.no_stored:
        # the same run on stored blocks: every product streams the 3x3 BSR
        # matrix from HBM -- the SpMV GB/s half of the metric
        sb = gpu_run(a.steps, untimed, 0, matrix_free=0)
        row = operator_row(sb, out["value"], r["digest"])
        row["roofline"] = spmv_roofline(sb, "k_cg_spmv (CG SpMV y = A p + <y,p> partial, "
                                            "streamed FP64 blocks)", read_traffic(0))
        row["step_roofline"] = step_roofline(sb)
        row["first_product_with_block_stores_ms"] = sb["store_ms"] / sb["steps"]
        row["plain_spmv_gbs"] = sb["plain_bytes"] / (sb["plain_ms"] * 1e-3) / 1e9
        row["spmv_unroll"] = sb["unroll"]
        row["spmv_tune_ms_2_4_8"] = sb["tune_ms"]
        out["stored_blocks"] = row
        out["spmv_hbm_gbs"] = row["roofline"]["achieved"]
    elif not prec and not mf:
        out["spmv_hbm_gbs"] = out["roofline"]["achieved"]
    if not prec and op == 0 and not a.no_stored:
        x = gpu_run(a.steps, untimed, 0, matrix_free=2)
        out["operator_per_step"] = dict(operator_row(x, out["value"], r["digest"]),
                                        steps_recomputing_blocks=x["mf_steps"],
                                        roofline=mf_roofline(x))
    if not a.no_full_run:
        # all 200 steps of the workload timed from the first (no untimed
        # steps, no tuning: the SpMV runs its default 4 slots per trip)
        fr = gpu_run(RUN_STEPS, 0, prec, matrix_free=op, tune=False)
        out["full_run"] = {
            "steps": fr["steps"],
            "value": fr["steps"] / fr["elapsed"],
            "unit": "steps/s",
            "ms_per_step": 1000.0 * fr["elapsed"] / fr["steps"],
            "cg_iterations_mean": float(np.mean(fr["iters"])),
            "cg_iterations_max": int(np.max(fr["iters"])),
            "broken_bonds": fr["broken"],
            "live_bonds_end": fr["live"],
            "steps_recomputing_blocks": fr["mf_steps"],
            "phase_ms_per_step": {k: v / fr["steps"] for k, v in fr["phase"].items()},
        }
    m_run = None
    if not prec and not a.no_mixed:
        m = gpu_run(a.steps, untimed, 1)
        m_run = m
        m_rate = m["steps"] / m["elapsed"]
        out["mixed_precision"] = {
            "value": m_rate,
            "unit": "steps/s",
            "speedup_vs_fp64": m_rate / out["value"],
            "ms_per_step": 1000.0 * m["elapsed"] / m["steps"],
            "cg_iterations_mean": float(np.mean(m["iters"])),
            "refinements_mean": float(np.mean(m["refinements"])),
            "phase_ms_per_step": {k: v / m["steps"] for k, v in m["phase"].items()},
            "roofline": spmv_roofline(m, "k_mx_spmv (float inner CG SpMV)", read_traffic(1)),
            "resid_spmv_ms_mean": m["resid_ms"] / max(m["resid_n"], 1),
            # the first inner product, which writes the float blocks the
            # assembly leaves to it
            "first_inner_product_with_block_stores_ms": m["store_ms"] / m["steps"],
            "plain_spmv_gbs": m["plain_bytes"] / (m["plain_ms"] * 1e-3) / 1e9,
            "spmv_unroll": m["unroll"],
            "spmv_tune_ms_2_4_8": m["tune_ms"],
            "position_error_vs_fp64": mixed_error(RUN_STEPS),
        }
    if not a.no_learned and os.path.exists(NETWORK):
        net = np.loadtxt(NETWORK, dtype=np.float32)
        rows = {}
        for lprec, name in ((0, "fp64"), (1, "mixed")):
            if lprec == 1 and m_run is None and not prec:
                continue
            L = gpu_run(a.steps, untimed, lprec, precond=2, network=net,
                        matrix_free=op if lprec == 0 else 0)
            base = r if lprec == prec else m_run
            rows[name] = {
                "value": L["steps"] / L["elapsed"],
                "unit": "steps/s",
                "speedup_vs_block_jacobi_same_precision": (L["steps"] / L["elapsed"]) /
                                                          (base["steps"] / base["elapsed"]),
                "speedup_vs_fp64_block_jacobi": (L["steps"] / L["elapsed"]) /
                                                (r["steps"] / r["elapsed"]) if not prec else None,
                "cg_iterations_mean": float(np.mean(L["iters"])),
                "block_jacobi_iterations_mean": float(np.mean(base["iters"])),
                "fallbacks": L["fallbacks"],
                "block_jacobi_probe_steps": L["probes"],
                "phase_ms_per_step": {k: v / L["steps"] for k, v in L["phase"].items()},
            }
        out["learned_preconditioner"] = {
            "network": "16-64-64-8 MLP, BF16 weights on MFMA, FP32 accumulation; trained by "
                       "train_precond.py on samples of 20 impact_100 runs with scene seeds "
                       "101-120; evaluated here on the shipped seeds (1973 / 2024)",
            "fallback": "block Jacobi after 1.5x the iteration count of the context's latest "
                        "block-Jacobi solve; every 16th precond-2 step (and the first, and the "
                        "one after a fallback) is a block-Jacobi probe that refreshes that count",
            "bound_on_gain": LEARNED_BOUND,
            **rows,
        }
    if not a.no_batched:
        bw = max(1, min(a.warmup, 5))
        bmf = batched_run(95, bw, 1, precision=0, matrix_free=1)
        graph = batched_run(95, bw, 1)
        eager = batched_run(95, bw, 0)
        out["batched_specimens"] = {
            "metric": "aggregate particle-steps/s",
            "value": bmf["value"],
            "unit": "particle-steps/s",
            "workload": "batch_64: 64 independent tension specimens in one context, counts "
                        "log-uniform in [5000, 60000], aspect 1..4, s0 U[0.01, 0.03], dt 1e-4 s, "
                        "block-Jacobi CG rtol 1e-6, one launch per phase for all specimens, "
                        "hipGraph-captured steps; value: FP64 with the matrix-free operator",
            "fp64_matrix_free_graph": bmf,
            "mixed_graph": dict(graph, solver="float inner CG with FP64 refinement, "
                                              "autotuned SpMV slots per trip"),
            "mixed_eager": eager,
        }
    if not a.no_cpu_baseline:
        log("cpu baseline: CPU oracle on the same workload ...")
        out["cpu_baseline"] = cpu_baseline(r["timed_steps"][0], max(1, min(a.cpu_steps, a.steps)))
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
