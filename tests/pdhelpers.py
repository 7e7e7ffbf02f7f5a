"""Small input builders shared by the tests (no reference code involved)."""
import numpy as np

import oracle as orc


def lattice(nx, ny, nz, h, jitter, seed, origin=(0.0, 0.0, 0.0)):
    """x-fastest simple-cubic lattice with uniform jitter in +-jitter*h."""
    rng = np.random.default_rng(seed)
    k, j, i = np.meshgrid(np.arange(nz), np.arange(ny), np.arange(nx), indexing="ij")
    g = np.stack([i, j, k], -1).reshape(-1, 3).astype(np.float64) * h
    g += rng.uniform(-1, 1, g.shape) * jitter * h
    g += np.asarray(origin)
    x4 = np.zeros((g.shape[0], 4))
    x4[:, :3] = g
    return x4


def material(h, horizon_factor=3.0, stiffness=1e4, mass=1e-3):
    """(horizon, radius, c_bond, mass) for spacing h: r = h/2, c = k h."""
    return (horizon_factor * h, 0.5 * h, stiffness * h, mass)


def make_system(x4, materials, body=None, s0=0.02, seed=1973, v4=None, **kw):
    n = x4.shape[0]
    hs = [m[0] for m in materials]
    counts, lists, total = orc.bonds_brute(x4, hs, body=body)
    assert total >= 0
    rs, col = orc.lists_to_csr(counts, lists)
    s00 = orc.stretch_init(n, s0, seed)
    v4 = np.zeros((n, 4)) if v4 is None else v4
    return orc.OracleSystem(x4, v4, s00, rs, col, materials, body=body, **kw)


def make_pair(x4, materials, body=None, s0=0.02, seed=1973, v4=None, bc=None, vbc=None,
              act=None):
    """The same initial state in the oracle and in a GPU context (bonds built
    on the GPU); returns (oracle_system, gpu_context)."""
    import pdsolver as pds
    n = x4.shape[0]
    sys_ = make_system(x4, materials, body=body, s0=s0, seed=seed, v4=v4, bc=bc, vbc=vbc,
                       act=act)
    ctx = pds.Context(n, materials)
    pds.load_state(ctx, x4, sys_.v, sys_.s00, body=body, act=act, bc=bc, vbc=vbc)
    ctx.build_bonds()
    return sys_, ctx


def config0_state(nx=20, ny=20, nz=20, h=0.01, jitter=0.05, seed=1973, top_speed=1.0):
    """Tension specimen: x-fastest jittered lattice, bottom z-layer fixed,
    top z-layer moved at +top_speed along z."""
    x4 = lattice(nx, ny, nz, h, jitter, seed)
    n = x4.shape[0]
    k = np.arange(n) // (nx * ny)
    bc = ((k == 0) | (k == nz - 1)).astype(np.int32)
    vbc = np.zeros((n, 4))
    vbc[k == nz - 1, 2] = top_speed
    v4 = vbc.copy()
    return x4, bc, vbc, v4


def deformed_pair(nx=10, ny=9, nz=8, amp=0.03, seed=0, s_fixed=0.02):
    """Oracle / GPU pair on a jittered lattice with random velocities and a
    random displacement of amp * spacing; stretch limits set to s_fixed."""
    h = 0.01
    x4 = lattice(nx, ny, nz, h, 0.05, seed)
    rng = np.random.default_rng(seed + 1)
    v4 = np.zeros_like(x4)
    v4[:, :3] = rng.uniform(-1, 1, (x4.shape[0], 3))
    sys_, ctx = make_pair(x4, [material(h)], v4=v4)
    disp = rng.uniform(-1, 1, (x4.shape[0], 3)) * amp * h
    sys_.x[:, :3] += disp
    ctx.upload("position", sys_.x)
    if s_fixed is not None:
        sys_.s[:] = s_fixed
        ctx.upload("stretch", sys_.s)
    return sys_, ctx
