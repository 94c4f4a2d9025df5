"""Seeded synthetic molecules for the benchmark configurations.

All generators return atomic numbers (int64) and positions in bohr
(float64), built with ``numpy.random.default_rng(seed)``.

* ``alkane``       -- all-anti zig-zag C_nH_{2n+2} (C-C 1.53 A, C-H 1.09 A,
                      tetrahedral angles), Gaussian noise on every coordinate.
* ``chain``        -- an alkane whose carbons are replaced by N (one H fewer)
                      or O (two H fewer) with probability 0.05 each.
* ``water_cluster``-- k waters on a simple-cubic lattice (3.1 A) with noise,
                      O-H 0.96 A, H-O-H 104.5 deg, random orientations.
* ``config0`` .. ``config4`` -- the benchmark workloads.
"""

from __future__ import annotations

import math

import numpy as np

from .constants import AA2AU

TET = math.acos(-1.0 / 3.0)  # 109.47 deg


def _unit(v):
    return v / np.linalg.norm(v)


def alkane_geometry(nc: int, rcc=1.53, rch=1.09):
    """Ideal all-anti alkane in Angstrom: (numbers, positions, H owners)."""
    if nc == 1:  # methane
        d = rch / math.sqrt(3.0)
        hs = np.array([[d, d, d], [-d, -d, d], [-d, d, -d], [d, -d, -d]])
        return (np.array([6, 1, 1, 1, 1], dtype=np.int64),
                np.vstack([np.zeros(3), hs]), np.zeros(4, dtype=np.int64))
    dx = rcc * math.sin(TET / 2)
    dy = rcc * math.cos(TET / 2)
    cpos = np.array([[i * dx, (i % 2) * dy, 0.0] for i in range(nc)])
    nums = [6] * nc
    pos = [c for c in cpos]
    owner = []
    half = TET / 2
    for i in range(nc):
        c = cpos[i]
        # virtual neighbours continue the zig-zag at the chain ends
        prev = cpos[i - 1] if i > 0 else np.array([-dx, dy, 0.0])
        nxt = cpos[i + 1] if i < nc - 1 else np.array([nc * dx, (nc % 2) * dy, 0.0])
        u1, u2 = _unit(prev - c), _unit(nxt - c)
        bis = -_unit(u1 + u2)
        nrm = _unit(np.cross(u1, u2))
        hs = [c + rch * (math.cos(half) * bis + math.sin(half) * nrm),
              c + rch * (math.cos(half) * bis - math.sin(half) * nrm)]
        if i == 0:
            hs.append(c + rch * u1)
        if i == nc - 1:
            hs.append(c + rch * u2)
        for h in hs:
            nums.append(1)
            pos.append(h)
            owner.append(i)
    return np.array(nums, dtype=np.int64), np.array(pos), np.array(owner)


def chain(rng: np.random.Generator, nc: int, p_n=0.05, p_o=0.05, sigma=0.03):
    """Alkane with random N/O substitution, noise sigma (A); returns bohr."""
    nums, pos, owner = alkane_geometry(nc)
    keep = np.ones(len(nums), dtype=bool)
    for i in range(nc):
        u = rng.random()
        if u < p_n:
            nums[i], drop = 7, 1
        elif u < p_n + p_o:
            nums[i], drop = 8, 2
        else:
            continue
        hidx = np.nonzero(owner == i)[0] + nc
        keep[hidx[len(hidx) - drop:]] = False
    nums, pos = nums[keep], pos[keep]
    pos = pos + rng.normal(0.0, sigma, size=pos.shape)
    return nums, pos * AA2AU


def _lattice_sites(shape=(9, 9, 8), order="lex"):
    idx = np.array([(i, j, k) for i in range(shape[0]) for j in range(shape[1])
                    for k in range(shape[2])], dtype=np.float64)
    if order == "compact":
        d = np.linalg.norm(idx, axis=1)
        idx = idx[np.lexsort((idx[:, 2], idx[:, 1], idx[:, 0], d))]
    return idx


def _random_rotation(rng):
    q = rng.normal(size=4)
    q /= np.linalg.norm(q)
    a, b, c, d = q
    return np.array([
        [a * a + b * b - c * c - d * d, 2 * (b * c - a * d), 2 * (b * d + a * c)],
        [2 * (b * c + a * d), a * a - b * b + c * c - d * d, 2 * (c * d - a * b)],
        [2 * (b * d - a * c), 2 * (c * d + a * b), a * a - b * b - c * c + d * d],
    ])


def water_cluster(rng: np.random.Generator, k: int, spacing=3.1, sigma=0.1, roh=0.96,
                  angle=104.5, order="lex"):
    """k waters on a simple-cubic lattice; returns bohr."""
    sites = _lattice_sites(order=order)[:k] * spacing
    sites = sites + rng.normal(0.0, sigma, size=sites.shape)
    t = math.radians(angle) / 2
    h1 = roh * np.array([math.sin(t), math.cos(t), 0.0])
    h2 = roh * np.array([-math.sin(t), math.cos(t), 0.0])
    nums, pos = [], []
    for o in sites:
        R = _random_rotation(rng)
        nums += [8, 1, 1]
        pos += [o, o + R @ h1, o + R @ h2]
    return np.array(nums, dtype=np.int64), np.array(pos) * AA2AU


def config0(seed=0):
    rng = np.random.default_rng(seed)
    return chain(rng, 8, p_n=0.0, p_o=0.0)


def config1(seed=1, count=256, nmin=2, nmax=20):
    rng = np.random.default_rng(seed)
    return [chain(rng, int(rng.integers(nmin, nmax + 1))) for _ in range(count)]


def config2(seed=2, count=512):
    return config1(seed=seed, count=count, nmin=5, nmax=40)


def config3(seed=3, nwater=600):
    return water_cluster(np.random.default_rng(seed), nwater)


def config4(seed=4, count=20000, rate=None, smin=5, smax=200):
    """Serving stream: list of (arrival_time_s, numbers, positions)."""
    rng = np.random.default_rng(seed)
    t = 0.0
    out = []
    for _ in range(count):
        if rate is not None:
            t += rng.exponential(1.0 / rate)
        s = math.exp(rng.uniform(math.log(smin), math.log(smax)))
        if rng.random() < 0.7:
            nums, pos = chain(rng, max(1, int(round((s - 2) / 3))))
        else:
            nums, pos = water_cluster(rng, max(1, int(round(s / 3))), order="compact")
        out.append((t, nums, pos))
    return out


def pad(mols):
    """Stack molecules into the padded-batch convention (0 = padding)."""
    nmax = max(len(z) for z, _ in mols)
    B = len(mols)
    nums = np.zeros((B, nmax), dtype=np.int64)
    pos = np.zeros((B, nmax, 3))
    for b, (z, p) in enumerate(mols):
        nums[b, : len(z)] = z
        pos[b, : len(z)] = p
    return nums, pos
