"""Scene seeding: the reference's particle seeding routines and the benchmark workloads.

* ``reference_scene(variant)`` restates SPHSystem::add_box_particle (SPH-3D-2/sph_system.cpp:
  247-264; SPH-3D-4/sph_system.cpp:248-289) and add_init_particle (SPH-2D/sph_system.cpp:
  148-162): float32-accumulated lattice loops, x outermost, exactly as the C++ loops run
  them (7220, 41850 and 1953 particles).
* ``dam_break_2d`` / ``adaptive_3d2`` / ``adaptive_3d4_collide`` build the seeded
  workloads: a 64x64 2D lattice in physical units, a jittered 64^3 3D-2 block at the floor
  corner of a 64^3-cell box, and two 128^3 3D-4 blocks moving toward each other in a
  512^3-cell box.

Random numbers come from a counter-based SplitMix64 (Steele, Lea, Flood 2014): value k of
stream s under seed S is splitmix64(S * 0x100000001B3 + (s << 40) + k), and a uniform in
[0, 1) is its top 24 bits / 2^24. Every value is a pure function of (S, s, k), so scenes are
identical on any machine and independent of generation order.
"""
import math

import numpy as np

from . import _abi
from ._abi import V2D, V3D2, V3D4, PARTICLE2D_DTYPE, PARTICLE3D_DTYPE

SEED_BASE = 0x5CA2012

_M64 = (1 << 64) - 1


def _splitmix64(x):
    x = np.asarray(x, dtype=np.uint64)
    with np.errstate(over="ignore"):
        z = x + np.uint64(0x9E3779B97F4A7C15)
        z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
        z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
    return z ^ (z >> np.uint64(31))


def uniform(seed, stream, n):
    """n uniforms in [0, 1) (float64 holding 24-bit values) of (seed, stream)."""
    base = (int(seed) * 0x100000001B3 + (int(stream) << 40)) & _M64
    with np.errstate(over="ignore"):
        ctr = np.uint64(base) + np.arange(n, dtype=np.uint64)
    z = _splitmix64(ctr)
    return (z >> np.uint64(40)).astype(np.float64) / float(1 << 24)


def _fseq(start, step, cond):
    """float32-accumulated loop values: v = start; while cond(v): yield v; v += step."""
    out = []
    v = np.float32(start)
    step = np.float32(step)
    while cond(v):
        out.append(v)
        v = np.float32(v + step)
    return np.array(out, dtype=np.float32)


def _f(x):
    return np.float32(x)


def _new_particles(variant, n):
    return np.zeros(n, dtype=PARTICLE2D_DTYPE if variant == V2D else PARTICLE3D_DTYPE)


def reference_scene(variant):
    """(params, particles) of the reference application's start-up scene."""
    p = _abi.param_defaults(variant)
    if variant == V2D:
        W, D = _f(p["world_width"][0]), _f(p["world_depth"][0])
        r, h = _f(p["large_radius"][0]), _f(p["large_kernel"][0])
        xs = _fseq(r / _f(2), h * _f(0.5), lambda x: x < W / _f(2))
        ys = _fseq(r / _f(2), h * _f(0.5), lambda y: D - y >= r / _f(2))
        X, Y = np.meshgrid(xs, ys, indexing="ij")
        a = _new_particles(V2D, X.size)
        a["pos"][:, 0] = X.ravel()
        a["pos"][:, 1] = Y.ravel()
        a["level"] = 1
        return p, a
    W, H, L = (_f(p[k][0]) for k in ("world_width", "world_height", "world_length"))
    if variant == V3D2:
        h = _f(p["large_kernel"][0])
        half = h / _f(2)
        st = h * _f(0.8)
        xs = _fseq(half, st, lambda x: x <= W / _f(2))
        ys = _fseq(half, st, lambda y: H - y >= half)
        zs = _fseq(half, st, lambda z: L - z >= half)
        level = 1
    else:
        sp = _f(p["small_kernel"][0])
        st = sp * _f(0.5)
        xs = _fseq(sp, st, lambda x: W - x >= sp)
        ys = _fseq(sp, st, lambda y: y < H / _f(4))
        zs = _fseq(sp, st, lambda z: L - z >= sp)
        level = 2
    X, Y, Z = np.meshgrid(xs, ys, zs, indexing="ij")
    a = _new_particles(variant, X.size)
    a["pos"][:, 0] = X.ravel()
    a["pos"][:, 1] = Y.ravel()
    a["pos"][:, 2] = Z.ravel()
    a["level"] = level
    return p, a


def new_particles(variant, p):
    """The particles SPHSystem::add_new_particle appends (the reference's '1' key):
    3D-2 (SPH-3D-2/sph_system.cpp:266-283) a lattice of small particles at spacing
    0.4 h_S filling the box 4 units in from the walls, x outermost; 3D-4
    (SPH-3D-4/sph_system.cpp:291-331) a ball of radius 5 around (24, 18, 12) falling at 10,
    lattice spacing 0.5 h_S over [-10, 10]^3, z outermost, one particle per site (the
    reference's second count increment is not reproduced). float32 loops as the C++ runs them."""
    W, H, L = (_f(p[k][0]) for k in ("world_width", "world_height", "world_length"))
    hl, hs = _f(p["large_kernel"][0]), _f(p["small_kernel"][0])
    if variant == V3D2:
        half = hl / _f(2)
        st = hs * _f(0.4)
        four = _f(4)
        xs = _fseq(half + four, st, lambda x: W - four - x >= half)
        ys = _fseq(half + four, st, lambda y: H - four - y >= half)
        zs = _fseq(half + four, st, lambda z: L - four - z >= half)
        X, Y, Z = np.meshgrid(xs, ys, zs, indexing="ij")
        a = _new_particles(V3D2, X.size)
        a["pos"][:, 0], a["pos"][:, 1], a["pos"][:, 2] = X.ravel(), Y.ravel(), Z.ravel()
        a["level"] = 2
        return a
    r = _f(10.0)
    sp = _f(float(hs) * 0.5)
    vs = _fseq(-r, sp, lambda v: v <= r)
    Z, Y, X = np.meshgrid(vs, vs, vs, indexing="ij")  # z outermost, x innermost
    X, Y, Z = X.ravel(), Y.ravel(), Z.ravel()
    ln = np.sqrt(((X * X + Y * Y).astype(np.float32) + Z * Z).astype(np.float32)).astype(np.float32)
    keep = ln <= _f(5.0)
    a = _new_particles(V3D4, int(keep.sum()))
    a["pos"][:, 0] = (_f(24.0) + X[keep]).astype(np.float32)
    a["pos"][:, 1] = (_f(18.0) + Y[keep]).astype(np.float32)
    a["pos"][:, 2] = (_f(12.0) + Z[keep]).astype(np.float32)
    a["vel"][:, 1] = -10.0
    a["level"] = 2
    return a


def lattice_density_sum(h, spacing, eps):
    """sum of (h^2 - r^2)^3 over an infinite square lattice around one particle (2D),
    self term h^6 included, pairs with r^2 < eps skipped (float64)."""
    k = int(math.ceil(h / spacing)) + 1
    s = h ** 6
    for i in range(-k, k + 1):
        for j in range(-k, k + 1):
            if i == 0 and j == 0:
                continue
            r2 = (i * spacing) ** 2 + (j * spacing) ** 2
            if r2 < eps or r2 >= h * h:
                continue
            s += (h * h - r2) ** 3
    return s


def dam_break_2d(n_side=64, spacing=0.01, h=0.02, rest_density=1000.0, stiffness=3.0,
                 viscosity=0.1, gravity=-9.8, dt=5e-4):
    """2D dam break in physical units: an n_side^2 lattice (spacing 0.01) from (0.005, 0.005)
    in a 1x1 box, h = 0.02 (so 50x50 cells), radius 0.01; mass calibrated so that the
    interior lattice density equals the rest density under the reference's density sum."""
    p = _abi.param_defaults(V2D)
    p["world_width"] = 1.0
    p["world_depth"] = 1.0
    p["large_kernel"] = h
    p["small_kernel"] = 0.9 * h
    p["large_radius"] = spacing
    p["small_radius"] = 0.8 * spacing
    p["rest_density"] = rest_density
    p["gas_constant"] = stiffness
    p["viscosity"] = viscosity
    p["gravity"] = (0.0, gravity)
    p["time_step"] = dt
    _abi.param_derive(V2D, p)
    poly6 = float(p["large_poly6"][0])
    s = lattice_density_sum(float(np.float32(h)), spacing, 1e-12)
    m = rest_density / (poly6 * s)
    p["large_mass"] = m
    p["small_mass"] = 0.5 * m
    i = np.arange(n_side, dtype=np.float64)
    xs = (0.5 * spacing + i * spacing).astype(np.float32)
    X, Y = np.meshgrid(xs, xs, indexing="ij")
    a = _new_particles(V2D, X.size)
    a["pos"][:, 0] = X.ravel()
    a["pos"][:, 1] = Y.ravel()
    a["level"] = 1
    return p, a


def adaptive_3d2(n_side=64, seed=SEED_BASE + 2, jitter=0.1, cells=64, dens_mul=4):
    """3D-2 dam break: an n_side^3 lattice of large particles at spacing 0.8 h with
    +-jitter*spacing uniform jitter per axis, starting at 1.0 against the floor corner of a
    cells^3 box (reference-native units, h_L = 1, so cells^3 grid cells); density grid
    cells*dens_mul + 1 voxels per axis."""
    p = _abi.param_defaults(V3D2)
    p["world_width"] = cells
    p["world_height"] = cells
    p["world_length"] = cells
    p["dens_mul"] = dens_mul
    _abi.param_derive(V3D2, p)
    h = float(p["large_kernel"][0])
    sp = 0.8 * h
    n = n_side ** 3
    idx = np.arange(n_side, dtype=np.float64)
    X, Y, Z = np.meshgrid(idx, idx, idx, indexing="ij")
    a = _new_particles(V3D2, n)
    for ax, G in enumerate((X, Y, Z)):
        u = uniform(seed, ax, n)
        a["pos"][:, ax] = (1.0 + G.ravel() * sp + (2.0 * u - 1.0) * jitter * sp).astype(np.float32)
    a["level"] = 1
    return p, a


def adaptive_3d4_collide(n_side=128, seed=SEED_BASE + 3, cells=512, dens_mul=1, gap=16.0,
                         drift=0.5):
    """3D-4 collision: two n_side^3 lattices of small particles (spacing 0.5 h_S) resting on
    the floor, gap apart along x around the box centre, each particle with a uniform
    U(-1, 1) velocity per component plus a drift of +-drift along x toward the other block;
    cells^3 grid cells of size h_L = 1.2."""
    p = _abi.param_defaults(V3D4)
    cell = float(np.float32(p["large_kernel"][0]))
    for k in ("world_width", "world_height", "world_length"):
        p[k] = (cells - 0.5) * cell
    p["dens_mul"] = dens_mul
    _abi.param_derive(V3D4, p)
    W = float(p["world_width"][0])
    sp = 0.5 * float(p["small_kernel"][0])
    ext = (n_side - 1) * sp
    n1 = n_side ** 3
    idx = np.arange(n_side, dtype=np.float64)
    X, Y, Z = np.meshgrid(idx, idx, idx, indexing="ij")
    a = _new_particles(V3D4, 2 * n1)
    x0 = [0.5 * W - 0.5 * gap - ext, 0.5 * W + 0.5 * gap]
    z0 = 0.5 * W - 0.5 * ext
    for b in range(2):
        sl = slice(b * n1, (b + 1) * n1)
        a["pos"][sl, 0] = (x0[b] + X.ravel() * sp).astype(np.float32)
        a["pos"][sl, 1] = (1.5 + Y.ravel() * sp).astype(np.float32)
        a["pos"][sl, 2] = (z0 + Z.ravel() * sp).astype(np.float32)
        for ax in range(3):
            u = uniform(seed, 10 * b + ax, n1)
            v = 2.0 * u - 1.0
            if ax == 0:
                v = v + (drift if b == 0 else -drift)
            a["vel"][sl, ax] = v.astype(np.float32)
    a["ev"] = a["vel"]
    a["level"] = 2
    return p, a


def tank_3d2(n, seed, scene, jitter=0.1, dens_mul=1):
    """One batched-workload tank (3D-2 rules, reference-native units): n large particles on
    the first n sites of a bx x by x bz lattice (bx = bz = ceil(cbrt n), x fastest, then z,
    then y) at spacing 0.8 h with +-jitter*spacing jitter, from 1 + jitter*spacing (clear of
    the walls at the particle radius 1) in the floor corner of a box
    2x as long as the block along x, 1.5x as high and one cell wider along z; each particle
    moves with a seeded velocity U(-a, a)^3, a ~ U(0, 1) per scene."""
    p = _abi.param_defaults(V3D2)
    h = float(p["large_kernel"][0])
    sp = 0.8 * h
    bx = int(math.ceil(round(n ** (1.0 / 3.0), 9)))
    bz = bx
    by = -(-n // (bx * bz))
    p["world_width"] = math.ceil(2.0 * bx * sp) + 2.0
    p["world_height"] = math.ceil(1.5 * by * sp) + 2.0
    p["world_length"] = math.ceil(bz * sp) + 2.0
    p["dens_mul"] = dens_mul
    _abi.param_derive(V3D2, p)
    k = np.arange(n)
    ix, iz, iy = k % bx, (k // bx) % bz, k // (bx * bz)
    a = _new_particles(V3D2, n)
    for ax, G in enumerate((ix, iy, iz)):
        u = uniform(seed, 16 * scene + ax, n)
        a["pos"][:, ax] = (1.0 + jitter * sp + G * sp + (2.0 * u - 1.0) * jitter * sp).astype(np.float32)
    amp = uniform(seed, 16 * scene + 3, 1)[0]
    for ax in range(3):
        u = uniform(seed, 16 * scene + 4 + ax, n)
        a["vel"][:, ax] = ((2.0 * u - 1.0) * amp).astype(np.float32)
    a["ev"] = a["vel"]
    a["level"] = 1
    return p, a


def batched_3d2(n_scenes=256, seed=SEED_BASE + 4, n_min=2048, n_max=65536):
    """The batched workload: n_scenes tanks (tank_3d2) with N_s = round(exp(U(ln n_min,
    ln n_max))) particles each. Returns (sizes, [(params, particles)] per scene)."""
    u = uniform(seed, 4095, n_scenes)
    sizes = np.rint(np.exp(math.log(n_min) + u * (math.log(n_max) - math.log(n_min)))).astype(int)
    return sizes, [tank_3d2(int(sizes[s]), seed, s) for s in range(n_scenes)]


def rollout_block(n_side=32, seed=SEED_BASE + 5, box=(48, 32, 32), gravity=None, jitter=0.1,
                  time_step=None):
    """Differentiable-rollout scene (3D-2 rules, reference-native units): an n_side^3
    lattice of large particles at spacing 0.8 h with +-jitter*spacing jitter, its low
    corner at 1 + jitter*spacing in a box of `box` cells; `gravity` (the y component) and
    `time_step` override the reference's values when given."""
    p = _abi.param_defaults(V3D2)
    p["world_width"], p["world_height"], p["world_length"] = box
    p["dens_mul"] = 1
    if gravity is not None:
        p["gravity"][0][1] = gravity
    if time_step is not None:
        p["time_step"] = time_step
    _abi.param_derive(V3D2, p)
    sp = 0.8 * float(p["large_kernel"][0])
    n = n_side ** 3
    idx = np.arange(n_side, dtype=np.float64)
    X, Y, Z = np.meshgrid(idx, idx, idx, indexing="ij")
    a = _new_particles(V3D2, n)
    for ax, G in enumerate((X, Y, Z)):
        u = uniform(seed, ax, n)
        a["pos"][:, ax] = (1.0 + jitter * sp + G.ravel() * sp + (2.0 * u - 1.0) * jitter * sp).astype(np.float32)
    a["level"] = 1
    return p, a


def field_theta(nf, seed, amp=1.0):
    """A seeded learnable field: nf^3 nodes x 3 components, U(-amp, amp), x fastest."""
    return ((2.0 * uniform(seed, 77, nf ** 3 * 3) - 1.0) * amp).astype(np.float32)


# workload table used by the benchmark and the tests
WORKLOADS = {
    "dam_break_2d_4096": dict(variant=V2D, build=dam_break_2d, capacity=8192, flags=0,
                              mc_every=0, steps=1000),
    "adaptive_3d2_262k": dict(variant=V3D2, build=adaptive_3d2, capacity=1000000,
                              flags=_abi.STEP_SPLIT | _abi.STEP_MERGE, mc_every=10, steps=2000),
    "adaptive_3d4_4m": dict(variant=V3D4, build=adaptive_3d4_collide, capacity=16000000,
                            flags=_abi.STEP_SPLIT | _abi.STEP_MERGE, mc_every=10, steps=500),
    # differentiable rollout: 32^3 particles falling into a 48 x 32 x 32 tank, split and
    # merge on (about 32k splits and 1.2k merge groups over the run), 8^3 x 3 field, 64^3
    # density sample, 200 steps of dt = 0.01 (at the reference's dt = 0.1 a 200-step
    # rollout is chaotic: its finite differences do not converge)
    "rollout_3d2_32k": dict(variant=V3D2, build=lambda: rollout_block(time_step=0.01),
                            capacity=None, flags=_abi.STEP_SPLIT | _abi.STEP_MERGE,
                            mc_every=0, steps=200, rollout=dict(field_nodes=8, sample_n=64,
                                                                checkpoint_every=20)),
    # 256 independent tanks stepped as one batch (capacity: 4x each tank's start count)
    "batched_3d2_256": dict(variant=V3D2, build=batched_3d2, capacity=None, capacity_factor=4,
                            flags=_abi.STEP_SPLIT | _abi.STEP_MERGE, mc_every=0, steps=1000,
                            batched=True),
}


def step_flags(workload, step_index):
    """Flags for step `step_index` (0-based): marching cubes on every mc_every-th step."""
    w = WORKLOADS[workload]
    f = w["flags"]
    if w["mc_every"] and (step_index + 1) % w["mc_every"] == 0:
        f |= _abi.STEP_MC
    return f
