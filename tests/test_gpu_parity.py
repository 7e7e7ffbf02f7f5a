"""GPU parity: the HIP path through the C ABI against the CPU oracle on the
same seeded inputs.  Integer / index results must match exactly; floating
point is compared bit for bit where the kernels follow the oracle's order of
operations, with the parity contract's tolerance (1e-9 relative) as the bar."""
import sys

import numpy as np
import pytest

import oracle as orc
from pdhelpers import config0_state, deformed_pair, lattice, make_pair, make_system, material

pytestmark = pytest.mark.gpu

pds = pytest.importorskip("pdsolver")


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def gpu_lists(ctx):
    cap, lists = ctx.get_bonds()
    top = ctx.download("bond_top")
    return cap, lists, top


def compare_bonds(sys_, ctx):
    cap, lists, top = gpu_lists(ctx)
    rs = sys_.row_start
    assert np.array_equal(cap, np.diff(rs))
    assert np.array_equal(top, sys_.top)
    for p in range(sys_.n):
        assert np.array_equal(lists[p, :cap[p]], sys_.col[rs[p]:rs[p + 1]]), p


@pytest.mark.parametrize("case", ["lattice", "cloud", "bodies", "exact_lattice"])
def test_bond_lists_bit_exact(case):
    say("case", case)
    rng = np.random.default_rng(17)
    body = None
    if case == "lattice":
        x4 = lattice(13, 11, 9, 0.01, 0.05, 3)
        mats = [material(0.01)]
    elif case == "cloud":
        x4 = np.zeros((3000, 4))
        x4[:, :3] = rng.uniform(-0.2, 0.3, (3000, 3))
        mats = [(0.05, 0.01, 1.0, 1.0)]
    elif case == "bodies":
        a = lattice(10, 10, 6, 0.01, 0.05, 4)
        b = lattice(7, 7, 7, 0.0047, 0.05, 5, origin=(0.02, 0.02, 0.03))
        x4 = np.concatenate([a, b])
        body = np.r_[np.zeros(len(a)), np.ones(len(b))].astype(np.int32)
        mats = [material(0.01), material(0.0047)]
    else:  # distances exactly at the horizon are excluded (strict <)
        x4 = lattice(9, 8, 7, 0.5, 0.0, 0)
        mats = [(1.5, 0.25, 1.0, 1.0)]
    hs = [m[0] for m in mats]
    counts, lists, total = orc.bonds_brute(x4, hs, body=body)
    ctx = pds.Context(x4.shape[0], mats)
    pds.load_state(ctx, x4, np.zeros_like(x4), np.full(x4.shape[0], 0.02), body=body)
    st = ctx.build_bonds()
    assert st.total_bonds == total
    cap, glists, top = gpu_lists(ctx)
    assert np.array_equal(cap, counts)
    assert np.array_equal(top, counts - 1)
    for p in range(x4.shape[0]):
        assert np.array_equal(glists[p, :cap[p]], lists[p, :counts[p]]), p
    ctx.close()


def test_bond_overflow_is_an_error():
    say("case overflow")
    x4 = lattice(6, 6, 6, 0.01, 0.0, 0)
    ctx = pds.Context(x4.shape[0], [(1.0, 0.1, 1.0, 1.0)], max_bonds=16)
    pds.load_state(ctx, x4, np.zeros_like(x4), np.ones(x4.shape[0]))
    with pytest.raises(pds.PDError) as e:
        ctx.build_bonds()
    assert e.value.code == pds.PD_ERR_BOND_OVERFLOW
    ctx.close()


@pytest.mark.parametrize("integrator,amp", [(0, 0.03), (1, 0.03), (0, 0.3)])
def test_assemble_bit_exact(integrator, amp):
    """amp 0.3: displacements of 0.3 spacings break a large share of the
    bonds, so rows hold several breaking bonds at every position of the
    assembly's three-slot trips and of its tail."""
    say("case assemble integrator", integrator, "amp", amp)
    sys_, ctx = deformed_pair(amp=amp)
    kw = dict(dt=1e-4, clock_scale=0.4, integrator=integrator)
    live0 = int((sys_.top + 1).sum())
    st_o = sys_.assemble(orc.make_params(**kw))
    st_g = ctx.assemble(pds.make_params(track_gap=1, **kw))
    assert st_o.broken_bonds > 0, "the case must exercise breaking"
    if amp > 0.1:
        assert st_o.broken_bonds > live0 // 5, (st_o.broken_bonds, live0)
    assert st_g.broken_bonds == st_o.broken_bonds
    assert st_g.live_bonds == st_o.live_bonds
    assert st_g.min_gap == st_o.min_gap and (st_g.gap_p, st_g.gap_q) == (st_o.gap_p, st_o.gap_q)
    compare_bonds(sys_, ctx)
    assert np.array_equal(ctx.download("force")[:, :3], sys_.f[:, :3])
    assert np.array_equal(ctx.download("rhs")[:, :3], sys_.b[:, :3])
    assert np.array_equal(ctx.download("new_stretch"), sys_.s_new)
    diag, vals = ctx.get_matrix()
    assert np.array_equal(diag, sys_.diag)
    for p in range(sys_.n):
        rs = sys_.row_start[p]
        k = sys_.top[p] + 1
        assert np.array_equal(vals[p, :k], sys_.val[rs:rs + k]), p
    ctx.close()


def test_assemble_ties_exact_lattice():
    """An unjittered, undisplaced lattice: every bond of the interior has the
    same stretch, so the nearest-to-threshold bond is decided by the order
    in which bonds are processed (the first one wins)."""
    say("case assemble ties")
    h = 0.01
    x4 = lattice(9, 8, 7, h, 0.0, 0)
    sys_, ctx = make_pair(x4, [material(h)])
    kw = dict(dt=1e-4, clock_scale=0.4)
    st_o = sys_.assemble(orc.make_params(**kw))
    st_g = ctx.assemble(pds.make_params(track_gap=1, **kw))
    assert st_g.min_gap == st_o.min_gap and (st_g.gap_p, st_g.gap_q) == (st_o.gap_p, st_o.gap_q)
    compare_bonds(sys_, ctx)
    assert np.array_equal(ctx.download("force")[:, :3], sys_.f[:, :3])
    assert np.array_equal(ctx.download("rhs")[:, :3], sys_.b[:, :3])
    ctx.close()


def test_spmv_bit_exact():
    say("case spmv")
    sys_, ctx = deformed_pair()
    sys_.assemble(orc.make_params())
    ctx.assemble(pds.make_params())
    rng = np.random.default_rng(3)
    x = np.zeros((sys_.n, 4))
    x[:, :3] = rng.normal(size=(sys_.n, 3))
    assert np.array_equal(ctx.spmv(x)[:, :3], sys_.spmv(x)[:, :3])
    ctx.close()


@pytest.mark.parametrize("precond", [0, 1])
def test_cg_bit_exact(precond):
    say("case cg precond", precond)
    sys_, ctx = deformed_pair(s_fixed=10.0)
    kw = dict(dt=2e-4, precond=precond, rtol=1e-10)
    sys_.assemble(orc.make_params(**kw))
    ctx.assemble(pds.make_params(**kw))
    st_o = sys_.cg(orc.make_params(**kw))
    st_g = ctx.solve(pds.make_params(**kw))
    assert st_o.converged and st_g.converged
    assert st_g.cg_iterations == st_o.cg_iterations
    assert st_g.b_norm == st_o.b_norm
    assert st_g.residual_norm == st_o.residual_norm
    assert np.array_equal(ctx.download("solution")[:, :3], sys_.sol[:, :3])
    ctx.close()


def test_cg_iteration_limit():
    say("case cg limit")
    sys_, ctx = deformed_pair(s_fixed=10.0)
    kw = dict(dt=2e-4, precond=0, rtol=1e-14, max_iter=3)
    ctx.assemble(pds.make_params(**kw))
    st = ctx.solve(pds.make_params(**kw), allow_fail=True)
    assert st.converged == 0 and st.cg_iterations == 3
    with pytest.raises(pds.PDError) as e:
        ctx.solve(pds.make_params(**kw))
    assert e.value.code == pds.PD_ERR_NOT_CONVERGED
    ctx.close()


def run_both(sys_, ctx, steps, params_for_step, check_every=1):
    """Advance both; returns per-step records."""
    rec = []
    for k in range(1, steps + 1):
        kw = params_for_step(k)
        st_o = sys_.step(orc.make_params(**kw))
        st_g = ctx.step(pds.make_params(track_gap=1, **kw))
        x_g = ctx.download("position")
        scale = np.abs(sys_.x[:, :3]).max()
        disp = np.abs(sys_.x[:, :3] - sys_.x0[:, :3]).max()
        dx = np.abs(x_g[:, :3] - sys_.x[:, :3]).max()
        rec.append(dict(step=k, it_o=st_o.cg_iterations, it_g=st_g.cg_iterations,
                        br_o=st_o.broken_bonds, br_g=st_g.broken_bonds, gap=st_o.min_gap,
                        rel=dx / scale, rel_disp=dx / max(disp, 1e-300), exact=dx == 0.0))
    return rec


def test_tension_steps_small():
    say("case tension small")
    x4, bc, vbc, v4 = config0_state(8, 8, 10)
    mats = [material(0.01)]
    sys_, ctx = make_pair(x4, mats, v4=v4, bc=bc, vbc=vbc)
    rec = run_both(sys_, ctx, 40, lambda k: dict(dt=1e-4, clock_scale=min(k, 30) / 30.0,
                                                 precond=1, rtol=1e-6))
    for r in rec:
        say(r)
        assert r["it_g"] == r["it_o"]
        assert r["br_g"] == r["br_o"]
        assert r["rel"] <= 1e-9
    assert sum(r["br_o"] for r in rec) > 0, "the tension case must break bonds"
    ctx.close()


def contact_scene(case):
    """(x4, body, materials, contact kwargs) of a contact case.  two_bodies and
    three_bodies build their per-step grid in one workgroup (few candidates,
    few cells; two_bodies' grid has at most 256 cells, one radix pass,
    three_bodies' more, two passes); split_far's grid spans more than 65536
    cells and interleaved selects more than 16384 candidates, so both take
    the multi-launch build."""
    if case == "two_bodies":
        a = lattice(8, 8, 4, 0.01, 0.05, 1)
        b = lattice(4, 4, 4, 0.0047, 0.05, 2, origin=(0.02, 0.02, 0.0335))
        parts, mats, rc = [a, b], [material(0.01), material(0.0047)], 0.0075
    elif case in ("three_bodies", "split_far"):
        # a slab with a small block just above one corner and another just
        # below the opposite corner (no bonds: horizon under the spacing)
        a = lattice(30, 30, 10, 0.01, 0.05, 4)
        b = lattice(3, 3, 3, 0.01, 0.05, 5, origin=(0.0, 0.0, 0.093))
        c = lattice(3, 3, 3, 0.01, 0.05, 6, origin=(0.27, 0.27, -0.023))
        parts = [a, b, c]
        mats = [material(0.01, horizon_factor=0.5)] * 3
        rc = 0.0075 if case == "three_bodies" else 0.004
    else:  # interleaved: two bodies on the two sublattices of one lattice
        x4 = lattice(28, 28, 28, 0.01, 0.05, 3)
        i = np.arange(x4.shape[0])
        body = ((i % 28 + (i // 28) % 28 + i // 784) % 2).astype(np.int32)
        kw = dict(contact_radius=0.012, contact_stiffness=2.5e8, contact_volume=1.0,
                  contact_cell=0.012)
        return x4, body, [material(0.01, horizon_factor=0.5)] * 2, kw
    x4 = np.concatenate(parts)
    body = np.concatenate([np.full(len(p), k) for k, p in enumerate(parts)]).astype(np.int32)
    kw = dict(contact_radius=rc, contact_stiffness=2.5e8, contact_volume=1.0, contact_cell=rc)
    return x4, body, mats, kw


@pytest.mark.parametrize("case", ["two_bodies", "three_bodies", "split_far", "interleaved"])
def test_contact_bit_exact(case):
    say("case contact", case)
    x4, body, mats, kw = contact_scene(case)
    sys_, ctx = make_pair(x4, mats, body=body)
    sys_.contact(orc.make_params(**kw))
    st = ctx.assemble(pds.make_params(**kw))
    fc = ctx.download("contact_force")
    assert np.abs(sys_.fc).max() > 0, "the case must have contacts"
    assert np.array_equal(fc[:, :3], sys_.fc[:, :3])
    # the grid build each case is meant to cover: one workgroup with one
    # radix pass (<= 256 cells), with two, and the multi-launch build
    say("contact grid build", st.contact_grid)
    want = {"two_bodies": 1, "three_bodies": 2, "split_far": 3, "interleaved": 3}[case]
    assert st.contact_grid == want, (case, st.contact_grid)
    ctx.close()


def test_newmark_steps():
    say("case newmark")
    x4, bc, vbc, v4 = config0_state(6, 6, 8)
    sys_, ctx = make_pair(x4, [material(0.01)], v4=v4, bc=bc, vbc=vbc)
    rec = run_both(sys_, ctx, 10, lambda k: dict(dt=1e-4, clock_scale=min(k, 30) / 30.0,
                                                 integrator=1, precond=1, rtol=1e-8))
    for r in rec:
        say(r)
        assert r["it_g"] == r["it_o"] and r["rel"] <= 1e-9
    ctx.close()


def test_reference_units_identity_cg():
    """The reference's own solver settings: identity preconditioner, rtol
    1e-20, dt = 128 * 2^-23 in scaled units (lattice spacing 0.0005), walls,
    gravity phase-in; bunny-like scaled material constants."""
    say("case reference units")
    h = 0.0005
    x4 = lattice(8, 8, 8, h, 0.01, 9, origin=(0.01, 0.01, 0.01))
    r = 0.00025
    horizon = r * 5.001
    c_times_v = 18.0 / np.pi / 2.0 / 1296.0 / 0.05  # pd_C_times_V with an unscaled radius
    mats = [(horizon, r, c_times_v * (5.0 * 2.94315e9), 1000.0)]
    sys_, ctx = make_pair(x4, mats, s0=1e-3)
    dt = 128 * (1.0 / 8388608.0)
    rec = run_both(sys_, ctx, 5, lambda k: dict(dt=dt, clock_scale=k / 30.0, precond=0,
                                                rtol=1e-20, walls=1, bmin=(0, 0, 0),
                                                bmax=(0.02, 0.02, 0.04), wall_radius=r))
    for rr in rec:
        say(rr)
        assert rr["it_g"] == rr["it_o"] and rr["rel"] <= 1e-9
    ctx.close()


@pytest.mark.parametrize("precision", [0, 1])
def test_solve_above_alpha_fold_threshold(precision):
    """More than 8192 tiles of 256 rows in one specimen: the CG updates take
    alpha from the one-workgroup fold ahead of them (the per-workgroup fold
    reads grow with the square of the tile count above that size) -- FP64
    and the float inner solve; iterations, residual and solution equal the
    oracle's bit for bit."""
    say("case alpha fold above threshold, precision", precision)
    h = 0.01
    x4 = lattice(130, 130, 125, h, 0.05, 21)  # 2,112,500 rows: 8,252 tiles
    n = x4.shape[0]
    assert (n + 255) // 256 > 8192
    mats = [material(h, horizon_factor=1.5)]
    counts, lists, dropped = orc.bonds_kdtree(x4, mats[0][0], max_bonds=32)
    assert dropped == 0
    rs, col = orc.lists_to_csr(counts, lists)
    del lists
    rng = np.random.default_rng(22)
    v4 = np.zeros_like(x4)
    v4[:, :3] = rng.uniform(-1, 1, (n, 3))
    s00 = orc.stretch_init(n, 0.02, 1973)
    S = orc.OracleSystem(x4, v4, s00, rs, col, mats)
    S.x[:, :3] += rng.uniform(-1, 1, (n, 3)) * 0.03 * h
    S.s[:] = 10.0
    ctx = pds.Context(n, mats)
    pds.load_state(ctx, x4, v4, s00)
    bst = ctx.build_bonds()
    assert bst.total_bonds == int(counts.sum())
    ctx.upload("position", S.x)
    ctx.upload("stretch", S.s)
    kw = dict(dt=2e-4, precond=1, rtol=1e-10, precision=precision)
    S.assemble(orc.make_params(**kw))
    ctx.assemble(pds.make_params(**kw))
    so = S.solve(orc.make_params(**kw))
    sg = ctx.solve(pds.make_params(**kw))
    say(dict(n=n, tiles=(n + 255) // 256, it=(so.cg_iterations, sg.cg_iterations),
             rel=sg.residual_norm / sg.b_norm))
    assert so.converged and sg.converged and sg.cg_iterations >= 3
    assert (sg.cg_iterations, sg.residual_norm) == (so.cg_iterations, so.residual_norm)
    assert np.array_equal(ctx.download("solution")[:, :3], S.sol[:, :3])
    ctx.close()
