"""Seeded synthetic inputs for the sheet workloads (no datasets exist).

Terrain is fBm built from the reference's own value noise ``relief_value``
(src/fluid_core.rs:4639-4668: hash01 lattice, smoothstep interpolation, range
-1..1): five octaves, base wavelength 8 m, lacunarity 2, gain 0.5, scaled to a
peak of +-0.3 m.  Puddles are Gaussian depth bumps.
"""
from __future__ import annotations

import numpy as np

U32 = np.uint32


def hash01(ix, iz, seed):
    """hash01 (src/fluid_core.rs:4639-4650) on integer arrays."""
    with np.errstate(over="ignore"):
        n = (ix.astype(np.int64).astype(U32) * U32(0x27BB2EE6)
             + iz.astype(np.int64).astype(U32) * U32(0x165667B1)) ^ U32(seed & 0xFFFFFFFF)
        n = n * U32(0x9E3779B1)
        n ^= n >> U32(16)
        n = n * U32(0x85EBCA6B)
        n ^= n >> U32(13)
    return (n >> U32(8)).astype(np.float32) * np.float32(1.0 / 16777216.0)


def relief_value(x, z, seed):
    """relief_value (src/fluid_core.rs:4652-4668): value noise in -1..1."""
    x = np.asarray(x, np.float32)
    z = np.asarray(z, np.float32)
    x0 = np.floor(x)
    z0 = np.floor(z)
    fx = x - x0
    fz = z - z0
    sx = fx * fx * (np.float32(3.0) - np.float32(2.0) * fx)
    sz = fz * fz * (np.float32(3.0) - np.float32(2.0) * fz)
    ix = x0.astype(np.int64)
    iz = z0.astype(np.int64)
    v00 = hash01(ix, iz, seed)
    v10 = hash01(ix + 1, iz, seed)
    v01 = hash01(ix, iz + 1, seed)
    v11 = hash01(ix + 1, iz + 1, seed)
    a = v00 + (v10 - v00) * sx
    b = v01 + (v11 - v01) * sx
    return (a + (b - a) * sz) * np.float32(2.0) - np.float32(1.0)


def cell_centres(res, cell, origin):
    """World x, z of cell centres, origin + (i + 0.5) cell (src/fluid_core.rs:9483-9488)."""
    c = (np.arange(res, dtype=np.float64) + 0.5) * cell
    x = (origin[0] + c).astype(np.float32)
    z = (origin[1] + c).astype(np.float32)
    return np.meshgrid(x, z)  # X[j, i], Z[j, i]


def fbm_terrain(res, cell, origin, seed, amplitude=0.3, octaves=5, wavelength=8.0, rows=None):
    """Seeded fBm heightfield in metres, peak |value| scaled to ``amplitude``
    over the full grid.  ``rows`` (j0, j1) builds only those rows with the same
    scale (for grids generated in stripes)."""
    X, Z = cell_centres(res, cell, origin)
    if rows is not None:
        X, Z = X[rows[0]:rows[1]], Z[rows[0]:rows[1]]
    f = np.zeros(X.shape, np.float32)
    amp, lam = 1.0, wavelength
    norm = 0.0
    for k in range(octaves):
        f += np.float32(amp) * relief_value(X / np.float32(lam), Z / np.float32(lam), seed + k)
        norm += amp
        amp *= 0.5
        lam *= 0.5
    # Scale by the octave weight sum, so the peak can reach +-amplitude.
    return (f * np.float32(amplitude / norm)).astype(np.float32)


def gaussian(res, cell, origin, cx, cz, depth, sigma):
    X, Z = cell_centres(res, cell, origin)
    r2 = (X - np.float32(cx)) ** 2 + (Z - np.float32(cz)) ** 2
    return (np.float32(depth) * np.exp(-r2 / np.float32(2.0 * sigma * sigma))).astype(np.float32)


def window_scene(seed=1, res=200, cell=0.2):
    """The 40 m window workload: fBm bed (5 octaves, 0.3 m), a Gaussian water
    puddle (0.10 m, sigma 4 m) at the centre, 0.02 m of slurry on a 10 x 10 m
    band at 0.5 Pa s.  Depths below 1e-6 m are cut to zero (a dry cell)."""
    origin = (-res * cell / 2.0, -res * cell / 2.0)
    bed = fbm_terrain(res, cell, origin, seed)
    water = gaussian(res, cell, origin, 0.0, 0.0, 0.10, 4.0)
    water[water < 1.0e-6] = 0.0
    X, Z = cell_centres(res, cell, origin)
    slurry = np.where((np.abs(X) <= 5.0) & (np.abs(Z) <= 5.0), np.float32(0.02), np.float32(0.0)).astype(np.float32)
    return {"res": res, "cell": cell, "origin": origin, "bed": bed.ravel(), "water": water.ravel(),
            "slurry": slurry.ravel(), "nu_water": 0.9, "nu_slurry": 0.5}


def large_scene(res=4096, cell=0.2, seed=7, puddles=64):
    """The 4096^2 (819 m) workload: fBm bed, a 0.05 m water film with
    ``puddles`` seeded Gaussian puddles (0.10 m, sigma 4 m) on top, and a
    0.02 m slurry sheet at 0.5 Pa s.  Every cell starts wet."""
    origin = (-res * cell / 2.0, -res * cell / 2.0)
    rng = np.random.default_rng(seed)
    bed = np.empty((res, res), np.float32)
    water = np.full((res, res), 0.05, np.float32)
    step = 256
    ext = res * cell
    centres = rng.uniform(origin[0], origin[0] + ext, size=(puddles, 2)).astype(np.float32)
    c = (np.arange(res, dtype=np.float64) + 0.5) * cell
    xs = (origin[0] + c).astype(np.float32)
    for j0 in range(0, res, step):
        j1 = min(res, j0 + step)
        bed[j0:j1] = fbm_terrain(res, cell, origin, seed, rows=(j0, j1))
        zs = (origin[1] + c[j0:j1]).astype(np.float32)
        for cx, cz in centres:
            dz2 = (zs - cz) ** 2
            near = dz2 < (6 * 4.0) ** 2
            if not near.any():
                continue
            dx2 = (xs - cx) ** 2
            water[j0:j1][near] += 0.10 * np.exp(-(dz2[near][:, None] + dx2[None, :]) / (2 * 16.0))
    slurry = np.full((res, res), 0.02, np.float32)
    return {"res": res, "cell": cell, "origin": origin, "bed": bed.ravel(), "water": water.ravel(),
            "slurry": slurry.ravel(), "nu_water": 0.9, "nu_slurry": 0.5}


def batch_scene(n=64, seed=11, lo=128, hi=1024, cell=0.2, film=0.0):
    """Independent sheet windows: side drawn uniformly from lo..hi cells
    (seeded), each with its own fBm bed (seed + k) and 1 to 3 Gaussian
    puddles (0.10 m, sigma 4 m) at seeded positions, on a film of ``film``
    metres (0: puddles on a dry bed, whose wet set may be several bodies)."""
    rng = np.random.default_rng(seed)
    sides = rng.integers(lo, hi + 1, size=n)
    out = []
    for k, res in enumerate(sides):
        res = int(res)
        ext = res * cell
        origin = (-ext / 2.0, -ext / 2.0)
        bed = fbm_terrain(res, cell, origin, seed + 1 + k).ravel()
        water = np.full(res * res, film, np.float32)
        for _ in range(int(rng.integers(1, 4))):
            cx, cz = rng.uniform(origin[0], origin[0] + ext, size=2)
            water += gaussian(res, cell, origin, np.float32(cx), np.float32(cz), 0.10, 4.0).ravel()
        water = water.astype(np.float32)
        water[water < 1.0e-6] = 0.0
        out.append({"res": res, "cell": cell, "origin": origin, "bed": bed, "water": water})
    return out
