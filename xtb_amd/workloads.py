"""Seeded synthetic molecules of the benchmark workloads.

* water clusters: centres on a jittered cubic grid (3.0 A spacing, uniform
  +-0.3 A jitter), O-H 0.96 A, H-O-H 104.5 deg, uniformly random orientation;
  32 waters fill a 4x4x2 grid.
* all-trans linear alkanes CnH(2n+2): C-C 1.54 A, C-H 1.09 A, tetrahedral
  angles, uniform +-0.02 A jitter on every coordinate.
* the 512-molecule batch: water cluster or alkane with probability 0.5, size
  log-uniform in [10, 200] atoms.

Positions are returned in bohr (1 bohr = 0.52917726 A, src/mctc/convert.f90).
"""
import math

import numpy as np

AATOAU = 1.0 / 0.52917726


def _rotation(rng):
    # uniformly random rotation from a unit quaternion (Shoemake's method)
    u1, u2, u3 = rng.random(3)
    a, b = math.sqrt(1 - u1), math.sqrt(u1)
    x, y, z, w = a * math.sin(2 * math.pi * u2), a * math.cos(2 * math.pi * u2), \
        b * math.sin(2 * math.pi * u3), b * math.cos(2 * math.pi * u3)
    return np.array([
        [1 - 2 * (y * y + z * z), 2 * (x * y - z * w), 2 * (x * z + y * w)],
        [2 * (x * y + z * w), 1 - 2 * (x * x + z * z), 2 * (y * z - x * w)],
        [2 * (x * z - y * w), 2 * (y * z + x * w), 1 - 2 * (x * x + y * y)]])


def water_cluster(nwater, seed, spacing=3.0, jitter=0.3, grid=None):
    """(numbers, positions[bohr]) of a seeded water cluster."""
    rng = np.random.default_rng(seed)
    if grid is None:
        grid = (4, 4, 2) if nwater == 32 else (math.ceil(nwater ** (1 / 3)),) * 3
    sites = [(i, j, k) for k in range(grid[2]) for j in range(grid[1]) for i in range(grid[0])][:nwater]
    if len(sites) < nwater:
        raise ValueError("grid too small")
    half = math.radians(104.5) / 2
    mono = np.array([[0.0, 0.0, 0.0],
                     [0.96 * math.sin(half), 0.0, 0.96 * math.cos(half)],
                     [-0.96 * math.sin(half), 0.0, 0.96 * math.cos(half)]])
    numbers, pos = [], []
    for s in sites:
        centre = np.array(s, dtype=float) * spacing + rng.uniform(-jitter, jitter, 3)
        rot = _rotation(rng)
        for z, r in zip((8, 1, 1), mono):
            numbers.append(z)
            pos.append(centre + rot @ r)
    return np.array(numbers, dtype=np.int32), np.array(pos) * AATOAU


def large_water_cluster(seed=2024):
    """700 waters on a 10 x 10 x 7 grid (2100 atoms, 4200 AOs)."""
    return water_cluster(700, seed, grid=(10, 10, 7))


def alkane(ncarbon, seed, jitter=0.02):
    """(numbers, positions[bohr]) of a seeded all-trans CnH(2n+2)."""
    rng = np.random.default_rng(seed)
    tet = math.acos(-1.0 / 3.0)
    dx, dy = 1.54 * math.sin(tet / 2), 1.54 * math.cos(tet / 2)
    c = np.array([[i * dx, (i % 2) * dy, 0.0] for i in range(-1, ncarbon + 1)])  # with virtual ends
    numbers = [6] * ncarbon
    pos = [c[i + 1] for i in range(ncarbon)]
    zhat = np.array([0.0, 0.0, 1.0])
    for i in range(1, ncarbon + 1):
        ci = c[i]
        nb = [c[i - 1], c[i + 1]]
        if ncarbon == 1:
            dirs = _tetra_methane()
        else:
            virt = []
            if i == 1:
                virt.append(c[0])
            if i == ncarbon:
                virt.append(c[ncarbon + 1])
            b = -((nb[0] - ci) / 1.54 + (nb[1] - ci) / 1.54)
            b /= np.linalg.norm(b)
            dirs = [(v - ci) / 1.54 for v in virt]
            for sgn in (1.0, -1.0):
                dirs.append(b * math.cos(tet / 2) + sgn * zhat * math.sin(tet / 2))
        for d in dirs:
            numbers.append(1)
            pos.append(ci + 1.09 * np.asarray(d))
    pos = np.array(pos) + rng.uniform(-jitter, jitter, (len(pos), 3))
    return np.array(numbers, dtype=np.int32), pos * AATOAU


def _tetra_methane():
    s = 1.0 / math.sqrt(3.0)
    return [np.array(v) * s for v in ((1, 1, 1), (1, -1, -1), (-1, 1, -1), (-1, -1, 1))]


def mixed_batch(nmol=512, seed=2024, lo=10, hi=200):
    """The batched workload: list of (numbers, positions) tuples."""
    rng = np.random.default_rng(seed)
    mols = []
    for m in range(nmol):
        size = math.exp(rng.uniform(math.log(lo), math.log(hi)))
        sub = int(rng.integers(1 << 62))
        if rng.random() < 0.5:
            mols.append(water_cluster(max(1, round(size / 3)), sub))
        else:
            mols.append(alkane(max(1, round((size - 2) / 3)), sub))
    return mols


def ao_count(numbers):
    """AOs of a H/C/O molecule (H: 1s; C, O: 2s2p) -- sizes only."""
    numbers = np.asarray(numbers)
    return int(np.sum(np.where(numbers == 1, 1, 4)))
