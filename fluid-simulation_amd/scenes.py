"""Synthetic tank set-ups of the benchmark workloads (host-side seeding).

Every workload is a unit-cube tank with origin 0 and h = 1/n (at 32^3 this is
the reference's H = 2/64 tank scaled, and for power-of-two n the division and
multiply forms of the cell index agree exactly).  Particles are seeded in every
cell whose centre lies inside the fluid region: with 8 per cell, one uniform
sample in each of the 2x2x2 sub-cells (stratified jitter); with fewer, uniform
samples over the cell.  Positions are clamped into [1.05h, 1-1.05h] as
generate_particles does (src/gpu/fluid.rs:1852-1861).  The random source is
numpy's PCG64 (`numpy.random.default_rng(seed)`), so a seed fixes the layout.

Physics shared by all workloads: rho 1000, g = (0,-9.81,0), wall friction 0,
cfl 2, volume stiffness 0.45, drift clamp 0.5, rest density = particles per
cell, surface dilation 0 at 8 per cell (1 below: src/scene/mod.rs:321-324),
liquid threshold 1, warm start on, CG cap 200 with relative tolerance 1e-5,
sort every substep.
"""
from __future__ import annotations

import numpy as np


def _cells_in(n, pred):
    c = (np.arange(n, dtype=np.float64) + 0.5) / n
    k, j, i = np.meshgrid(c, c, c, indexing="ij")  # i fastest in linear order
    m = pred(i, j, k)
    idx = np.nonzero(m.reshape(-1))[0]
    ii = idx % n
    jj = (idx // n) % n
    kk = idx // (n * n)
    return np.stack([ii, jj, kk], 1)


def _seed(cells, n, ppc, rng):
    h = 1.0 / n
    nc = len(cells)
    if ppc == 8:
        sub = np.array([[sx, sy, sz] for sz in (0, 1) for sy in (0, 1) for sx in (0, 1)], np.float64)
        u = rng.random((nc, 8, 3))
        frac = (sub[None, :, :] + u) * 0.5
    else:
        frac = rng.random((nc, ppc, 3))
    pos = (cells[:, None, :].astype(np.float64) + frac) * h
    pos = pos.reshape(-1, 3)
    lo, hi = 1.05 * h, 1.0 - 1.05 * h
    return np.clip(pos, lo, hi).astype(np.float32)


def base_physics(n, ppc, dt, flip=0.95):
    return dict(
        h=np.float32(1.0) / np.float32(n), dt=dt, rho=1000.0, flip_blend=flip, wall_friction=0.0,
        gravity=(0.0, -9.81, 0.0), rest_density=float(ppc), volume_stiffness=0.45, drift_clamp=0.5,
        liquid_threshold=1, surface_dilation=0 if ppc >= 8 else 1, cfl=2.0,
        pressure_iterations=200, pressure_warm_start=1, pressure_residual_tolerance=1e-5,
        particle_sort=1, particle_sort_period=1, scene_density=float(ppc),
    )


def dam_break_32(seed=0):
    """32^3 dam break: block x<0.5, y<0.6, 8 per cell, v = 0, dt 1/120, 100 steps."""
    n = 32
    cells = _cells_in(n, lambda x, y, z: (x < 0.5) & (y < 0.6))
    rng = np.random.default_rng(seed)
    pos = _seed(cells, n, 8, rng)
    vel = np.zeros_like(pos)
    return dict(name="dam_break_32", n=n, pos=pos, vel=vel, steps=100,
                physics=base_physics(n, 8, 1.0 / 120.0))


def high_density_128(seed=1):
    """128^3 tank, lower 40% filled at 8 per cell (~6.7M particles), velocity
    U[-0.5,0.5]^3, dt 1/240, 50 steps."""
    n = 128
    cells = _cells_in(n, lambda x, y, z: y < 0.4)
    rng = np.random.default_rng(seed)
    pos = _seed(cells, n, 8, rng)
    vel = (rng.random(pos.shape) - 0.5).astype(np.float32)
    return dict(name="high_density_128", n=n, pos=pos, vel=vel, steps=50,
                physics=base_physics(n, 8, 1.0 / 240.0))


def large_splash_256(seed=2):
    """256^3: pool y<0.25 plus a sphere r=0.15 at (0.5,0.75,0.5) moving at
    (0,-2,0); 8 per cell (~35.4M particles), dt 1/240, 30 steps."""
    n = 256
    pool = lambda x, y, z: y < 0.25
    sph = lambda x, y, z: (x - 0.5) ** 2 + (y - 0.75) ** 2 + (z - 0.5) ** 2 < 0.15 ** 2
    rng = np.random.default_rng(seed)
    c_pool = _cells_in(n, pool)
    c_sph = _cells_in(n, lambda x, y, z: sph(x, y, z) & ~pool(x, y, z))
    p1 = _seed(c_pool, n, 8, rng)
    p2 = _seed(c_sph, n, 8, rng)
    pos = np.concatenate([p1, p2])
    vel = np.zeros_like(pos)
    vel[len(p1):, 1] = -2.0
    return dict(name="large_splash_256", n=n, pos=pos, vel=vel, steps=30,
                physics=base_physics(n, 8, 1.0 / 240.0))


def control_64(seed=4):
    """64^3 block x<0.5, y<0.5, 8 per cell (524,288 particles), dt 1/120."""
    n = 64
    cells = _cells_in(n, lambda x, y, z: (x < 0.5) & (y < 0.5))
    rng = np.random.default_rng(seed)
    pos = _seed(cells, n, 8, rng)
    vel = np.zeros_like(pos)
    return dict(name="control_64", n=n, pos=pos, vel=vel, steps=60,
                physics=base_physics(n, 8, 1.0 / 120.0))


def batched_tanks(count=64, seed=3):
    """64 tanks: edge ~ U{24,32,48,64,96}, fill height ~ U[0.2,0.6],
    particles per cell ~ U{4..8}."""
    rng = np.random.default_rng(seed)
    tanks = []
    for t in range(count):
        n = int(rng.choice([24, 32, 48, 64, 96]))
        fill = float(rng.uniform(0.2, 0.6))
        ppc = int(rng.integers(4, 9))
        cells = _cells_in(n, lambda x, y, z, f=fill: y < f)
        pos = _seed(cells, n, ppc, rng)
        tanks.append(dict(name=f"tank{t}", n=n, pos=pos, vel=np.zeros_like(pos), fill=fill, ppc=ppc,
                          physics=base_physics(n, ppc, 1.0 / 120.0)))
    return tanks


def small_tank(n=16, seed=5, fill=0.5, ppc=8, vel_scale=0.5, dt=1.0 / 120.0):
    """Small random tank for parity tests."""
    cells = _cells_in(n, lambda x, y, z: y < fill)
    rng = np.random.default_rng(seed)
    pos = _seed(cells, n, ppc, rng)
    vel = ((rng.random(pos.shape) - 0.5) * 2 * vel_scale).astype(np.float32)
    return dict(name=f"small_{n}", n=n, pos=pos, vel=vel, steps=5, physics=base_physics(n, ppc, dt))
