"""Seeded scene generators for the snow workloads.

Counter-based RNG (numpy Philox) with recorded seeds, never rand(): the
reference's initializer draws from the platform's rand()
(realtime-deformations/material_point_method.cpp:18-63,
realtime-deformations/utils.h:110-127), so its particle placement cannot be
reproduced across platforms; these generators replace it.

Each generator returns a dict: x (n,3), v (n,3), m (n,), E0 (n,) float32,
params (mpm_amd.Params keyword arguments), dims, h, dt, steps.
"""
from __future__ import annotations

import math

import numpy as np

RHO = 400.0  # kg/m^3


def _rng(seed):
    return np.random.Generator(np.random.Philox(int(seed)))


def _ball(rng, n, centre, radius):
    """n points uniform in a ball (rejection sampling)."""
    out = np.empty((0, 3))
    while out.shape[0] < n:
        k = int((n - out.shape[0]) * 2.2) + 64
        p = rng.uniform(-1.0, 1.0, size=(k, 3))
        p = p[(p * p).sum(1) <= 1.0]
        out = np.concatenate([out, p])
    return (np.asarray(centre) + radius * out[:n]).astype(np.float32)


def material(theta_s=7.5e-3, E=1.4e5, bc_mode=2):
    return dict(E=E, nu=0.2, xi=10.0, theta_c=2.5e-2, theta_s=theta_s, bc_mode=bc_mode)


def snowball(seed=0, n=8192, grid=64, radius=0.08, centre=(0.5, 0.6, 0.5), v0=(0.0, -1.0, 0.0), E=1.4e5,
             dt=1e-4, steps=100):
    """Plain snowball: n particles uniform in a sphere, sticky floor and
    separating walls.  Defaults: the 8,192-particle 64^3 free-fall case."""
    rng = _rng(seed)
    x = _ball(rng, n, centre, radius)
    mp = RHO * (4.0 / 3.0) * math.pi * radius ** 3 / n
    return dict(x=x, v=np.tile(np.float32(v0), (n, 1)), m=np.full(n, mp, np.float32),
                E0=np.full(n, E, np.float32), dims=(grid, grid, grid), h=1.0 / grid, dt=dt, steps=steps,
                params=material(E=E))


def two_snowballs(seed=1, n_total=262144, grid=128, radius=0.1, dt=5e-5, steps=1000):
    """Two colliding snowballs: centres (0.3,0.5,0.5) and (0.7,0.5,0.5),
    velocities (+3,0,0) and (-3,0,0); N/2 particles uniform in each sphere."""
    rng = _rng(seed)
    half = n_total // 2
    xa = _ball(rng, half, (0.3, 0.5, 0.5), radius)
    xb = _ball(rng, n_total - half, (0.7, 0.5, 0.5), radius)
    x = np.concatenate([xa, xb])
    v = np.concatenate([np.tile(np.float32([3, 0, 0]), (half, 1)), np.tile(np.float32([-3, 0, 0]), (n_total - half, 1))])
    mp = RHO * (4.0 / 3.0) * math.pi * radius ** 3 / half
    return dict(x=x, v=v, m=np.full(n_total, mp, np.float32), E0=np.full(n_total, 1.4e5, np.float32),
                dims=(grid, grid, grid), h=1.0 / grid, dt=dt, steps=steps, params=material())


def slab(seed=2, n_total=4194304, grid=256, cells=(128, 32, 128), bottom=0.5, ppc=(4, 12), dt=2.5e-5, steps=200):
    """Snow slab at scale: a cells[0] x cells[1] x cells[2] block of cells
    (0.5 x 0.125 x 0.5 at dx = 1/256) with its bottom at y = `bottom`; each
    cell's particle count drawn uniformly from ppc, then adjusted by seeded
    +-1 steps (within ppc) so the total is exactly n_total; positions uniform
    in their cell; E0 uniform in [1e5, 2e5]; uniform mass rho h^3 / 8."""
    rng = _rng(seed)
    h = 1.0 / grid
    nc = int(np.prod(cells))
    counts = rng.integers(ppc[0], ppc[1] + 1, size=nc)
    diff = int(counts.sum()) - n_total
    while diff != 0:
        idx = rng.integers(0, nc, size=abs(diff) * 2 + 16)
        for i in idx:
            if diff > 0 and counts[i] > ppc[0]:
                counts[i] -= 1
                diff -= 1
            elif diff < 0 and counts[i] < ppc[1]:
                counts[i] += 1
                diff += 1
            if diff == 0:
                break
    origin_cell = np.array([(grid - cells[0]) // 2, int(round(bottom / h)), (grid - cells[2]) // 2])
    cell_ids = np.repeat(np.arange(nc), counts)
    ci = np.stack(np.unravel_index(cell_ids, cells), axis=1) + origin_cell
    x = ((ci + rng.uniform(0.0, 1.0, size=(n_total, 3))) * h).astype(np.float32)
    E0 = rng.uniform(1e5, 2e5, size=n_total).astype(np.float32)
    return dict(x=x, v=np.zeros((n_total, 3), np.float32), m=np.full(n_total, RHO * h ** 3 / 8, np.float32), E0=E0,
                dims=(grid, grid, grid), h=h, dt=dt, steps=steps, params=material())


def landing_slab(seed=2, n_total=4194304, grid=256, speed=2.0, dt=2.5e-5, steps=200):
    """The 4,194,304-particle slab of slab() moved down onto the sticky floor:
    its bottom cell row starts at node row 3, so the floor band (node rows
    j < 3) is inside the bottom particles' stencils from the first substep,
    and it falls at `speed` m/s.  The impact stops the bottom layers and sends
    a compression wave up the slab (about 24 m/s, 100 substeps for 15 cells):
    F_E leaves the identity, the SVD rotates and the clamp turns strain into
    plastic flow, the deforming counterpart of the free-falling slab."""
    out = slab(seed=seed, n_total=n_total, grid=grid, bottom=3.0 / grid, dt=dt, steps=steps)
    out["v"][:, 1] = -float(speed)
    return out


def compressing(sc, rate, centre=None):
    """Adds a uniform compression rate * (centre - x) (1/s) to a scene's
    velocities (centre: the particles' mean by default), so the body
    deforms from the first substep: stress, hardening and the clamp act."""
    x = sc["x"].astype(np.float64)
    c = x.mean(0) if centre is None else np.asarray(centre, np.float64)
    sc = dict(sc)
    sc["v"] = (sc["v"].astype(np.float64) + rate * (c - x)).astype(np.float32)
    return sc


def reference_scene(seed=0, n=2147, grid=20, h=0.05, origin=(0.5, 0.6, 0.5), velocity=(0.0, -200.0, 0.0),
                    width=0.2, mass=6e-5):
    """The reference demo's placement (material_point_method.cpp:18-63):
    eight quarter-offset candidates per cell of the bounding cube, each
    jittered uniformly in a ball of radius h/4, kept if within `width` of the
    origin; filled from the back of the particle array.  Reference constants
    (theta_s = 5e-3, no domain walls), grid 20^3 at h = 0.05."""
    rng = _rng(seed)
    oc = (np.asarray(origin, np.float32) / np.float32(h)).astype(np.int32)
    wc = int(np.float32(width) / np.float32(h))
    ds = np.array([[1, 1, 1], [1, 1, 3], [1, 3, 1], [1, 3, 3], [3, 1, 1], [3, 1, 3], [3, 3, 1], [3, 3, 3]], np.float32)
    pts = []
    for i in range(oc[0] - wc, oc[0] + wc):
        for j in range(oc[1] - wc, oc[1] + wc):
            for k in range(oc[2] - wc, oc[2] + wc):
                for d in ds:
                    u = rng.uniform(-1, 1, 3)
                    while (u * u).sum() > 1:
                        u = rng.uniform(-1, 1, 3)
                    p = (np.float32([i, j, k]) + d * np.float32(0.25) + np.float32(0.25) * u.astype(np.float32)) * np.float32(h)
                    if np.linalg.norm(p - np.asarray(origin, np.float32)) > width:
                        continue
                    pts.append(p.astype(np.float32))
    pts = np.array(pts[:n], np.float32)
    n = pts.shape[0]
    x = pts[::-1].copy()
    return dict(x=x, v=np.tile(np.float32(velocity), (n, 1)), m=np.full(n, mass, np.float32),
                E0=np.full(n, 1.4e5, np.float32), dims=(grid, grid, grid), h=h, dt=1e-5, steps=100,
                params=dict(E=1.4e5, nu=0.2, xi=10.0, theta_c=2.5e-2, theta_s=5e-3, bc_mode=0))


def reference_colliders():
    """main.cpp's three static box colliders (realtime-deformations/main.cpp:119-156):
    floor box T(0.5,-0.2,0.5) S(0.4); two 45-degree boxes at x = 0 and x = 1."""
    s45, c45 = math.sin(math.radians(45.0) / 2), math.cos(math.radians(45.0) / 2)
    q45 = (c45, 0.0, 0.0, s45)
    return [
        dict(translation=(0.5, -0.2, 0.5), rotation=(1.0, 0.0, 0.0, 0.0), scale=(0.4, 0.4, 0.4)),
        dict(translation=(0.0, 0.3, 0.5), rotation=q45, scale=(0.2, 0.2, 0.3)),
        dict(translation=(1.0, 0.3, 0.5), rotation=q45, scale=(0.2, 0.2, 0.3)),
    ]


def scene_batch(seed=3, nscenes=64, n_range=(4096, 262144), e_range=(5e4, 5e5), v_range=3.0, ppc=8.0,
                small_grid=64, large_grid=128, split=32768, dt=1e-4, steps=500):
    """Many independent snowballs stepped together: per scene N_s
    log-uniform in n_range, grid small_grid^3 for N_s <= split else
    large_grid^3, a sphere whose radius gives `ppc` particles per cell, centre
    uniform in [0.3, 0.7]^3, E0 log-uniform in e_range, initial velocity
    uniform in [-v_range, v_range]^3, uniform density RHO.  Returns a list of
    single-scene dicts (same keys as snowball())."""
    rng = _rng(seed)
    out = []
    for s in range(nscenes):
        n = int(round(math.exp(rng.uniform(math.log(n_range[0]), math.log(n_range[1])))))
        grid = small_grid if n <= split else large_grid
        h = 1.0 / grid
        radius = (3.0 * n * h ** 3 / (ppc * 4.0 * math.pi)) ** (1.0 / 3.0)
        centre = rng.uniform(0.3, 0.7, 3)
        E = float(math.exp(rng.uniform(math.log(e_range[0]), math.log(e_range[1]))))
        v0 = rng.uniform(-v_range, v_range, 3).astype(np.float32)
        x = _ball(rng, n, centre, radius)
        mp = RHO * (4.0 / 3.0) * math.pi * radius ** 3 / n
        out.append(dict(x=x, v=np.tile(v0, (n, 1)), m=np.full(n, mp, np.float32), E0=np.full(n, E, np.float32),
                        dims=(grid, grid, grid), h=h, dt=dt, steps=steps, params=material(E=E)))
    return out


def adjoint_batch(seed=4, nscenes=16, n=32768, rate=1.2):
    """The differentiable workload: `nscenes` snowballs of n particles on 64^3
    (dt 1e-4, 200 substeps, uniform initial velocity in [-3, 3]^3 per scene),
    each compressed at `rate` 1/s on top of it so the stress acts from the
    start (dL/dE0 != 0), and a seeded target point per scene 0.02 from its
    centre; -> (scenes, targets).  At 1.2/s the singular values of F_E stay
    within [0.9757, 0.9985] over the 200 substeps (scenes 0 and 1, FP64
    oracle), inside the clamp band [1 - 2.5e-2, 1 + 7.5e-3]: the loss is
    smooth (no plastic-clamp kinks) where finite differences check the
    gradients.  1.4/s reaches the band's lower edge."""
    scs = [compressing(sc, rate) for sc in scene_batch(seed=seed, nscenes=nscenes, n_range=(n, n), split=n,
                                                        steps=200)]
    rng = np.random.default_rng(seed + 100)
    targets = [sc["x"].astype(np.float64).mean(0) + rng.normal(0.0, 0.02, 3) for sc in scs]
    return scs, targets


def com_loss(xs, vs, targets):
    """Per scene |centre of mass - target|^2 + 1e-3 mean |v|^2 (uniform
    masses: the centre of mass is the mean position), summed over scenes;
    -> (L, [dL/dx per scene], [dL/dv per scene]) for the final particles."""
    L, gx, gv = 0.0, [], []
    for x, v, t in zip(xs, vs, targets):
        x64, v64 = np.asarray(x, np.float64), np.asarray(v, np.float64)
        d = x64.mean(0) - t
        L += float(d @ d) + 1e-3 * float((v64 * v64).sum(1).mean())
        gx.append(np.tile(2.0 * d / x64.shape[0], (x64.shape[0], 1)).astype(np.float32))
        gv.append((2e-3 * v64 / x64.shape[0]).astype(np.float32))
    return L, gx, gv
