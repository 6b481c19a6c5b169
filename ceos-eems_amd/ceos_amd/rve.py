"""Synthetic voxel RVEs for the FE^2 workloads (periodic unit cells with spherical inclusions).

Each RVE is the unit cube [0,1]^3 cut into n^3 voxel hex8 elements (n in {8, 12, 16}); a voxel is
inclusion (phase 1) when its centre lies inside one of the RVE's spheres (periodic distance),
matrix (phase 0) otherwise.  Spheres are placed by random sequential addition without overlap
(periodic), radii uniform in [R_MIN, R_MAX], until their volume reaches the target fraction; the
last radius is shrunk to hit it.  Everything is drawn from numpy PCG64 seeded by (seed, rve id),
so an RVE is reproducible from its id alone.

Phase arrays are uint8 of length n^3, voxel (i, j, k) at i + n (j + n k) (x fastest), the element
numbering of meshgen.cube(n).

Material of the workload (Neo-Hookean, model 3 of ModConstitutiveModelLibrary.f90:39-56, props
C10 = mu/2, BulkModulus = K): matrix mu = 1, K = 50; inclusion 10x stiffer, mu = 10, K = 500.
"""
import numpy as np

SIDES = (8, 12, 16)
VOLUME_FRACTION = 0.2
R_MIN, R_MAX = 0.12, 0.2
SEED = 20261203
MATRIX_PROPS = (0.5, 50.0)     # C10, K
INCLUSION_PROPS = (5.0, 500.0)


def sides(n_rve, seed=SEED, choices=SIDES):
    rng = np.random.Generator(np.random.PCG64([seed, 0x5eed]))
    return rng.choice(np.asarray(choices, dtype=np.int32), size=n_rve)


def spheres(rve_id, seed=SEED, vf=VOLUME_FRACTION, max_tries=20000):
    """Non-overlapping periodic spheres (centres in [0,1)^3, radii) of total volume vf."""
    rng = np.random.Generator(np.random.PCG64([seed, int(rve_id)]))
    cen, rad = [], []
    vol = 0.0
    tries = 0
    while vol < vf - 1e-12 and tries < max_tries:
        tries += 1
        r = rng.uniform(R_MIN, R_MAX)
        c = rng.uniform(0.0, 1.0, size=3)
        need = vf - vol
        if 4.0 / 3.0 * np.pi * r ** 3 > need:
            r = (need * 3.0 / (4.0 * np.pi)) ** (1.0 / 3.0)
        ok = True
        for c2, r2 in zip(cen, rad):
            d = c - c2
            d -= np.round(d)
            if np.dot(d, d) < (r + r2) ** 2:
                ok = False
                break
        if ok:
            cen.append(c)
            rad.append(r)
            vol += 4.0 / 3.0 * np.pi * r ** 3
    return np.array(cen).reshape(-1, 3), np.array(rad)


def phases(n, rve_id, seed=SEED, vf=VOLUME_FRACTION):
    """uint8 phase of every voxel (1 = inclusion) of an n^3 RVE."""
    cen, rad = spheres(rve_id, seed, vf)
    g = (np.arange(n) + 0.5) / n
    z, y, x = np.meshgrid(g, g, g, indexing="ij")
    P = np.stack([x.ravel(), y.ravel(), z.ravel()], axis=1)
    ph = np.zeros(n ** 3, dtype=np.uint8)
    for c, r in zip(cen, rad):
        d = P - c
        d -= np.round(d)
        ph[np.einsum("ij,ij->i", d, d) < r * r] = 1
    return ph


def batch(n_rve, seed=SEED, choices=SIDES):
    """(sides int32 [n_rve], phases uint8 concatenated) of a seeded batch."""
    s = sides(n_rve, seed, choices)
    return s, np.concatenate([phases(int(n), i, seed) for i, n in enumerate(s)])


def torus_elements(n):
    """Element connectivity of the periodic n^3 voxel mesh: (ne, 8) 0-based torus node ids,
    hex8 node order, and the unwrapped local corner offsets (8, 3) in voxel units."""
    d = np.array([[0, 0, 0], [1, 0, 0], [1, 1, 0], [0, 1, 0],
                  [0, 0, 1], [1, 0, 1], [1, 1, 1], [0, 1, 1]])
    k, j, i = np.meshgrid(np.arange(n), np.arange(n), np.arange(n), indexing="ij")
    ijk = np.stack([i.ravel(), j.ravel(), k.ravel()], axis=1)
    c = (ijk[:, None, :] + d[None, :, :]) % n
    return (c[..., 0] + n * (c[..., 1] + n * c[..., 2])).astype(np.int32), d
