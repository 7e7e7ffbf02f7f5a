"""The learned block preconditioner (precond 2): features and the BF16 MFMA
network against a numpy float32 restatement (BF16 rounding at the same
points), the preconditioned solve against the oracle given the same blocks
(bit for bit), the block-Jacobi fallback, and the trainer's samples and
loss."""
import os
import sys

import numpy as np
import pytest

import oracle as orc
from pdhelpers import config0_state, deformed_pair, make_pair, material

pytestmark = pytest.mark.gpu
pds = pytest.importorskip("pdsolver")


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def bf16(x):
    """float32 -> nearest bfloat16 (ties to even), returned as float32"""
    u = np.ascontiguousarray(x, np.float32).view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000
    return u.astype(np.uint32).view(np.float32)


def random_net(seed, w3_scale=0.3, eps=1e-3):
    rng = np.random.default_rng(seed)
    p = np.zeros(pds.LEARNED_PARAMS, np.float32)

    def put(name, val):
        off, shape = pds.LEARNED_LAYOUT[name]
        p[off:off + int(np.prod(shape))] = np.asarray(val, np.float32).ravel()

    put("W1", rng.normal(size=(64, 16)) * 0.5)
    put("b1", rng.normal(size=64) * 0.1)
    put("W2", rng.normal(size=(64, 64)) * 0.2)
    put("b2", rng.normal(size=64) * 0.1)
    put("W3", rng.normal(size=(8, 64)) * w3_scale * 0.1)
    put("b3", rng.normal(size=8) * 0.1)
    put("mean", rng.normal(size=16) * 0.1)
    put("istd", np.abs(rng.normal(size=16)) + 0.5)
    put("eps", eps)
    return p


def get(p, name):
    off, shape = pds.LEARNED_LAYOUT[name]
    return p[off:off + int(np.prod(shape))].reshape(shape)


def features(S, mats, rows=None, body=None):
    """restated features of the oracle system after assembly (rows: the
    rows to fill, default all; body: per-row body index into mats)"""
    n = S.n
    f = np.zeros((n, 16), np.float32)
    for p in (range(n) if rows is None else rows):
        h = mats[0 if body is None else body[p]][0]
        m = mats[0 if body is None else body[p]][3]
        rs = S.row_start[p]
        top = S.top[p]
        cap = S.row_start[p + 1] - rs
        live = S.col[rs:rs + top + 1]
        d = np.linalg.norm(S.x[live, :3] - S.x[p, :3], axis=1)
        d0 = np.linalg.norm(S.x0[live, :3] - S.x0[p, :3], axis=1)
        st = d / d0 - 1.0
        f[p, 0] = (top + 1) * 0.01
        f[p, 1] = np.float32(st.mean()) * 100 if len(st) else 0
        f[p, 2] = np.float32(st.max()) * 100 if len(st) else 0
        D = S.diag[p]
        f[p, 3:9] = [(D[0] - m) / m, D[1] / m, D[2] / m, (D[4] - m) / m, D[5] / m, (D[8] - m) / m]
        broken = S.col[rs + top + 1:rs + cap]
        db = np.linalg.norm(S.x[broken, :3] - S.x[p, :3], axis=1) if len(broken) else []
        f[p, 9] = min([2 * h] + list(db)) / h
    return f


def mlp(p, f):
    x = bf16((f - get(p, "mean")) * get(p, "istd"))
    h1 = bf16(np.maximum(x @ bf16(get(p, "W1")).T + get(p, "b1"), 0))
    h2 = bf16(np.maximum(h1 @ bf16(get(p, "W2")).T + get(p, "b2"), 0))
    return h2 @ bf16(get(p, "W3")).T + get(p, "b3")


def blocks(p, o, diag):
    eps = float(get(p, "eps"))
    out = np.zeros((len(o), 9))
    for i in range(len(o)):
        d = np.clip(o[i, [0, 2, 5]], -4, 4)
        L = np.array([[np.exp(d[0]), 0, 0], [o[i, 1], np.exp(d[1]), 0], [o[i, 3], o[i, 4], np.exp(d[2])]])
        S = L @ L.T + eps * np.eye(3)
        C = np.linalg.cholesky(diag[i].reshape(3, 3))
        Ci = np.linalg.inv(C)
        out[i] = (Ci.T @ S @ Ci).ravel()
    return out


@pytest.mark.parametrize("precision", [0, 1])
def test_network_blocks_match_restatement(precision):
    say("case learned blocks precision", precision)
    sys_, ctx = deformed_pair(amp=0.05, s_fixed=0.01)
    mats = [material(0.01)]
    net = random_net(3)
    ctx.learned_set(net)
    kw = dict(dt=1e-4, precond=2, rtol=1e-8, precision=precision)
    so = sys_.assemble(orc.make_params(**kw))
    assert so.broken_bonds > 0, "the case must have broken bonds (feature 9)"
    ctx.assemble(pds.make_params(**kw))
    ctx.solve(pds.make_params(**kw))
    M = ctx.get_precond()
    f = features(sys_, mats)
    ref = blocks(net, mlp(net, f), sys_.diag)
    bj = np.array([np.linalg.inv(d.reshape(3, 3)).ravel() for d in sys_.diag])
    scale = np.abs(ref).max(axis=1, keepdims=True)
    err = np.abs(M - ref) / scale
    say(dict(max_rel=float(err.max()), median_rel=float(np.median(err)),
             departure_from_block_jacobi=float((np.abs(ref - bj) / scale).max())))
    assert (np.abs(ref - bj) / scale).max() > 0.1, "the network must change the blocks"
    assert np.median(err) < 2e-3 and err.max() < 5e-2
    ctx.close()


@pytest.mark.parametrize("precision", [0, 1])
def test_learned_solve_matches_oracle_given_blocks(precision):
    say("case learned solve precision", precision)
    sys_, ctx = deformed_pair(s_fixed=10.0)
    ctx.learned_set(random_net(5))
    kw = dict(dt=2e-4, precond=2, rtol=1e-10, precision=precision)
    sys_.assemble(orc.make_params(**kw))
    ctx.assemble(pds.make_params(**kw))
    st = ctx.solve(pds.make_params(**kw))
    sys_.minv[:] = ctx.get_precond()
    so = (sys_.cg_mixed if precision else sys_.cg)(orc.make_params(**kw))
    assert st.converged and so.converged
    assert st.cg_iterations == so.cg_iterations and st.residual_norm == so.residual_norm
    assert np.array_equal(ctx.download("solution")[:, :3], sys_.sol[:, :3])
    ctx.close()


def test_learned_solve_matrix_free_equals_stored():
    """The learned preconditioner over the matrix-free operator: the same
    blocks, iterations and solution as over the stored blocks."""
    say("case learned solve matrix-free")
    _, a = deformed_pair(s_fixed=10.0)
    _, b = deformed_pair(s_fixed=10.0)
    net = random_net(5)
    a.learned_set(net)
    b.learned_set(net)
    kw = dict(dt=2e-4, precond=2, rtol=1e-10)
    a.assemble(pds.make_params(**kw))
    b.assemble(pds.make_params(matrix_free=1, **kw))
    sa = a.solve(pds.make_params(**kw))
    sb = b.solve(pds.make_params(matrix_free=1, **kw))
    assert np.array_equal(a.get_precond(), b.get_precond())
    assert (sa.cg_iterations, sa.residual_norm) == (sb.cg_iterations, sb.residual_norm)
    assert np.array_equal(a.download("solution"), b.download("solution"))
    a.close()
    b.close()


def test_fallback_to_block_jacobi():
    say("case learned fallback")
    x4, bc, vbc, v4 = config0_state(8, 8, 10)
    sys_, ctx = make_pair(x4, [material(0.01)], v4=v4, bc=bc, vbc=vbc)
    kw = dict(dt=1e-4, rtol=1e-8)
    s1 = ctx.step(pds.make_params(precond=1, **kw))  # records the block-Jacobi count
    bad = random_net(9, w3_scale=40.0)  # far from block Jacobi
    ctx.learned_set(bad)
    s2 = ctx.step(pds.make_params(precond=2, learned_fallback=1.0, **kw))
    say(dict(bj=s1.cg_iterations, learned=s2.cg_iterations, fallbacks=s2.precond_fallbacks))
    assert s2.converged and s2.residual_norm <= 1e-8 * s2.b_norm
    assert s2.precond_fallbacks == 1
    ctx.close()


def mix64(z):
    z = np.asarray(z, np.uint64)
    z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
    z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
    return z ^ (z >> np.uint64(31))


def test_trainer_samples_and_fit():
    say("case trainer")
    sys_, ctx = deformed_pair(amp=0.05, s_fixed=0.01)
    kw = dict(dt=1e-4, precond=1, rtol=1e-8)
    sys_.assemble(orc.make_params(**kw))
    ctx.assemble(pds.make_params(**kw))
    tr = pds.Trainer(100000, seed=3)
    seed, stride = 11, 3
    added = tr.add(ctx, seed=seed, stride=stride)
    smp = tr.samples(0, added)
    off = int(mix64(np.uint64(seed)) % np.uint64(stride))
    rows = np.arange(off, sys_.n, stride)
    assert added == len(rows)
    # probes: +-1 from the hash of (seed, p); v = (A e)_p with the FP64 matrix
    with np.errstate(over="ignore"):
        h = mix64(np.uint64(seed) * np.uint64(0x9E3779B97F4A7C15) + np.arange(sys_.n, dtype=np.uint64))
    bits = np.stack([(h >> np.uint64(k)) & np.uint64(1) for k in range(3)], 1)
    E = np.where(bits == 1, 1.0, -1.0)
    V = (sys_.dense_matrix() @ E.ravel()).reshape(-1, 3)
    assert np.array_equal(smp[:, 25:28], E[rows].astype(np.float32))
    np.testing.assert_allclose(smp[:, 22:25], V[rows], rtol=1e-5, atol=1e-6 * np.abs(V).max())
    np.testing.assert_allclose(smp[:, 16:22], sys_.diag[rows][:, [0, 1, 2, 4, 5, 8]], rtol=1e-6)
    f = features(sys_, [material(0.01)])
    np.testing.assert_allclose(smp[:, :16], f[rows], rtol=1e-4, atol=1e-5)
    losses = tr.fit(300, batch=1024, lr=3e-3)
    say(dict(first=float(losses[:10].mean()), last=float(losses[-10:].mean())))
    assert np.all(np.isfinite(losses))
    assert losses[-10:].mean() < losses[:10].mean()
    p = tr.params()
    assert np.all(np.isfinite(p)) and p.shape == (pds.LEARNED_PARAMS,)
    ctx.learned_set(p)
    st = ctx.solve(pds.make_params(dt=1e-4, precond=2, rtol=1e-8))
    assert st.converged
    tr.close()
    ctx.close()


ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
NETWORK = os.path.join(ROOT, "sci-solver_peridynamic_amd", "networks", "impact_precond.txt")
IMPACT_30 = os.path.join(ROOT, "sci-solver_peridynamic_amd", "configs", "impact_30.txt")


@pytest.mark.parametrize("precision", [0, 1])
def test_shipped_network_on_impact_matches_oracle(precision):
    """The learned preconditioner on its own workload: the shipped network
    (networks/impact_precond.txt) steps the impact_30 scene with precond 2
    for 60 steps, through the projectile's first contact and the first
    broken bonds.  Each step the oracle takes the GPU's preconditioner
    blocks (block Jacobi on the probe steps, which refresh the fallback's
    baseline) and must give the same broken bonds, CG iterations, residual
    and positions bit for bit.  On two steps, one before the contact and one
    after bonds broke, the network's blocks are checked against the numpy
    restatement of the features (broken-bond distance included) and the
    MLP.  The reference solve the blocks precondition is the identity-
    preconditioned CG at core/simulator.cu:763-770."""
    from test_gpu_configs import setup_pair
    say("case shipped network impact_30 precision", precision)
    net = np.loadtxt(NETWORK, dtype=np.float32)
    assert net.shape == (pds.LEARNED_PARAMS,)
    desc, sc, S, ctx = setup_pair(IMPACT_30)
    mats = pds.materials_of(desc)
    ctx.learned_set(net)
    steps = 60
    probes = learned_steps = fallbacks = 0
    bj = 0  # iterations of the latest block-Jacobi probe: the fallback's baseline
    its = {"learned": [], "probe": []}
    checked = []
    first_contact = first_break = None
    broken_total = 0
    for k in range(1, steps + 1):
        kw = dict(pds.step_kwargs(desc, k), precision=precision, precond=2)
        sg = ctx.step(pds.make_params(**kw))
        prm = orc.make_params(**kw)
        S.contact(prm)
        sa = S.assemble(prm)
        S.swap_stretch()
        if sg.precond_probe:
            so = S.solve(orc.make_params(**dict(kw, precond=1)))
            probes += 1
            bj = sg.cg_iterations
            its["probe"].append(sg.cg_iterations)
        else:
            assert bj > 0
            M = ctx.get_precond()  # the learned blocks, also after a fallback
            budget = int(np.ceil(1.5 * bj))
            S.minv[:] = M
            so = S.solve(orc.make_params(**dict(kw, max_iter=budget)))
            if sg.precond_fallbacks:
                # the learned solve used its budget; block Jacobi goes on from
                # its iterate with the rest of the iteration limit
                assert not so.converged and so.cg_iterations == budget, k
                s3 = S.solve(orc.make_params(**dict(kw, precond=1,
                                                    max_iter=kw["max_iter"] - budget)))
                so.cg_iterations = budget + s3.cg_iterations
                so.converged, so.residual_norm = s3.converged, s3.residual_norm
                fallbacks += 1
            learned_steps += 1
            its["learned"].append(sg.cg_iterations)
            want = (k == 25) or (first_break is not None and len(checked) < 2 and k >= 50)
            if want:
                rows = np.arange(0, S.n, 7)
                brk = np.nonzero(S.top < np.diff(S.row_start) - 1)[0]
                rows = np.union1d(rows, brk)
                f = features(S, mats, rows=rows, body=sc["body"])
                ref = blocks(net, mlp(net, f[rows]), S.diag[rows])
                scale = np.abs(ref).max(axis=1, keepdims=True)
                err = np.abs(M[rows] - ref) / scale
                say(dict(step=k, rows=len(rows), rows_with_broken=len(brk),
                         max_rel=float(err.max()), median_rel=float(np.median(err)),
                         feature9_min=float(f[rows, 9].min())))
                assert np.median(err) < 2e-3 and err.max() < 5e-2
                if k >= 50:
                    assert len(brk) > 0 and f[brk, 9].min() < 2.0
                checked.append(k)
        S.update(prm)
        assert sa.broken_bonds == sg.broken_bonds, (k, sa.broken_bonds, sg.broken_bonds)
        assert so.converged and sg.converged
        assert (so.cg_iterations, so.residual_norm) == (sg.cg_iterations, sg.residual_norm), k
        assert np.array_equal(ctx.download("solution")[:, :3], S.sol[:, :3]), k
        assert np.array_equal(ctx.download("position")[:, :3], S.x[:, :3]), k
        if first_contact is None and np.abs(S.fc[:, :3]).max() > 0:
            first_contact = k
        if first_break is None and sa.broken_bonds > 0:
            first_break = k
        broken_total += sa.broken_bonds
    assert probes >= 2 and learned_steps >= 50
    assert first_contact is not None and first_break is not None and len(checked) == 2
    say(dict(precision=precision, probes=probes, learned_steps=learned_steps, fallbacks=fallbacks,
             first_contact=first_contact, first_break=first_break, broken=broken_total,
             learned_iterations_mean=float(np.mean(its["learned"])),
             probe_iterations=its["probe"], checked_steps=checked))
    ctx.close()
