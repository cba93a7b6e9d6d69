"""Point generators (src/discretization/point_generators.jl) and the seeded lattice jitter.

Positions are returned as Julia's ``3 x n`` layout (row d = coordinate d)."""
import math

import numpy as np


def _grid(l, dx):
    # range((-l + dx)/2, (l - dx)/2; step=dx) .- mean  (point_generators.jl:46-51)
    start = (-l + dx) / 2
    stop = (l - dx) / 2
    n = int(math.floor((stop - start) / dx + 1e-10)) + 1
    g = start + np.arange(n, dtype=np.float64) * dx
    return g - np.sum(g) / n


def uniform_box(lx, ly, lz, dx, center=(0.0, 0.0, 0.0)):
    """uniform_box(lx, ly, lz, ΔX0; center): point_generators.jl:45-60.

    Points ordered x fastest, then y, then z (Iterators.product).  Returns
    (position 3 x n, volume n)."""
    gx, gy, gz = _grid(lx, dx), _grid(ly, dx), _grid(lz, dx)
    Z, Y, X = np.meshgrid(gz, gy, gx, indexing="ij")
    pos = np.stack([X.reshape(-1), Y.reshape(-1), Z.reshape(-1)])
    vol = np.full(pos.shape[1], dx ** 3)
    for d in range(3):
        if abs(center[d]) > np.finfo(np.float64).eps:
            pos[d] += center[d]
    return pos, vol


_M64 = np.uint64(0xFFFFFFFFFFFFFFFF)


def splitmix64(counter, seed):
    """Counter-based SplitMix64: output k of the stream started at `seed` (uint64 arrays)."""
    with np.errstate(over="ignore"):
        z = np.uint64(seed) + (np.asarray(counter, np.uint64) + np.uint64(1)) * np.uint64(
            0x9E3779B97F4A7C15)
        z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
        z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
        z = z ^ (z >> np.uint64(31))
    return z


def jitter(pos, dx, amplitude=0.05, seed=0x5EED):
    """Add a uniform perturbation in [-amplitude*dx, amplitude*dx) to every coordinate.

    Coordinate d of point i uses SplitMix64 output 3*i + d of `seed`, so the same body is
    generated in any language."""
    pos = np.array(pos, dtype=np.float64, copy=True)
    n = pos.shape[1]
    k = np.arange(3 * n, dtype=np.uint64)
    u = (splitmix64(k, seed) >> np.uint64(11)).astype(np.float64) * (1.0 / 9007199254740992.0)
    off = (2.0 * u - 1.0) * (amplitude * dx)
    pos += off.reshape(n, 3).T
    return pos


def lattice_block(nx, ny, nz, dx, jitter_amplitude=0.0, seed=0x5EED):
    """nx x ny x nz lattice of spacing dx (uniform_box ordering), optionally jittered."""
    pos, vol = uniform_box(nx * dx, ny * dx, nz * dx, dx)
    assert pos.shape[1] == nx * ny * nz, (pos.shape, nx, ny, nz)
    if jitter_amplitude:
        pos = jitter(pos, dx, jitter_amplitude, seed)
    return pos, vol
