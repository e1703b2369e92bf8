"""Seeded molecular geometries for the benchmark workloads (angstrom in, bohr out).

Each generator takes a ``numpy.random.Generator`` so a workload is fully reproducible
from its seed.  Geometries:

* water: O at the origin, O-H = 0.958 A + U(-0.02, 0.02), H-O-H = 104.5 deg + U(-2, 2);
* benzene: regular hexagon C-C 1.39 A, radial C-H 1.09 A, N(0, 0.02 A) on every coordinate;
* n-alkane CnH2n+2: all-trans zig-zag, C-C 1.54 A, C-H 1.09 A, tetrahedral angles,
  N(0, 0.02 A) jitter;
* naphthalene: two fused regular hexagons (C-C 1.40 A, C-H 1.09 A) + N(0, 0.02 A);
* carbon cloud: n C atoms uniform in a cube, rejection-sampled to a minimum separation.
"""
from __future__ import annotations

import math
from typing import List, Tuple

import numpy as np

from .basis import ANGSTROM_TO_BOHR

Geometry = Tuple[List[str], np.ndarray]  # symbols, coords in bohr


def water(rng: np.random.Generator) -> Geometry:
    r1 = 0.958 + rng.uniform(-0.02, 0.02)
    r2 = 0.958 + rng.uniform(-0.02, 0.02)
    th = math.radians(104.5 + rng.uniform(-2.0, 2.0))
    xyz = np.array([[0.0, 0.0, 0.0], [r1, 0.0, 0.0], [r2 * math.cos(th), r2 * math.sin(th), 0.0]])
    return ["O", "H", "H"], xyz * ANGSTROM_TO_BOHR


def benzene(rng: np.random.Generator, jitter: float = 0.02) -> Geometry:
    sym, xyz = [], []
    for k in range(6):
        a = k * math.pi / 3
        sym.append("C")
        xyz.append([1.39 * math.cos(a), 1.39 * math.sin(a), 0.0])
    for k in range(6):
        a = k * math.pi / 3
        sym.append("H")
        xyz.append([(1.39 + 1.09) * math.cos(a), (1.39 + 1.09) * math.sin(a), 0.0])
    xyz = np.array(xyz) + rng.normal(0.0, jitter, size=(12, 3))
    return sym, xyz * ANGSTROM_TO_BOHR


def alkane(rng: np.random.Generator, n_c: int, jitter: float = 0.02) -> Geometry:
    """All-trans CnH2n+2 in the xy plane zig-zag, hydrogens tetrahedral."""
    cc, ch = 1.54, 1.09
    half = math.radians(109.47) / 2
    dx = cc * math.sin(half)
    dy = cc * math.cos(half)
    carbons = np.array([[i * dx, (i % 2) * dy, 0.0] for i in range(n_c)])
    sym = ["C"] * n_c
    hs = []
    # two out-of-plane H per carbon, pointing away from the chain's local bisector
    for i in range(n_c):
        side = -1.0 if i % 2 == 0 else 1.0
        base = np.array([0.0, side * math.cos(half), 0.0])
        for sz in (-1.0, 1.0):
            v = base + np.array([0.0, 0.0, sz * math.sin(half)])
            v = v / np.linalg.norm(v)
            hs.append(carbons[i] + ch * v)
    # terminal H along the chain direction
    if n_c == 1:
        hs = [carbons[0] + ch * np.array(v) / np.linalg.norm(v)
              for v in ([1, 1, 1], [1, -1, -1], [-1, 1, -1], [-1, -1, 1])]
    else:
        for end, nb in ((0, 1), (n_c - 1, n_c - 2)):
            v = carbons[end] - carbons[nb]
            v[1] = -v[1]
            v = v / np.linalg.norm(v)
            hs.append(carbons[end] + ch * v)
    xyz = np.vstack([carbons, np.array(hs)])
    sym = sym + ["H"] * len(hs)
    xyz = xyz + rng.normal(0.0, jitter, size=xyz.shape)
    return sym, xyz * ANGSTROM_TO_BOHR


def naphthalene(rng: np.random.Generator, jitter: float = 0.02) -> Geometry:
    cc, ch = 1.40, 1.09
    s3 = math.sqrt(3.0)
    c = [[0.0, cc / 2, 0.0], [0.0, -cc / 2, 0.0]]
    for sx in (1.0, -1.0):
        c += [[sx * cc * s3 / 2, cc, 0.0], [sx * cc * s3, cc / 2, 0.0],
              [sx * cc * s3, -cc / 2, 0.0], [sx * cc * s3 / 2, -cc, 0.0]]
    c = np.array(c)
    hs = []
    for sx in (1.0, -1.0):
        hs.append([sx * cc * s3 / 2, cc + ch, 0.0])
        hs.append([sx * (cc * s3 + ch * s3 / 2), cc / 2 + ch / 2, 0.0])
        hs.append([sx * (cc * s3 + ch * s3 / 2), -cc / 2 - ch / 2, 0.0])
        hs.append([sx * cc * s3 / 2, -cc - ch, 0.0])
    xyz = np.vstack([c, np.array(hs)]) + rng.normal(0.0, jitter, size=(18, 3))
    return ["C"] * 10 + ["H"] * 8, xyz * ANGSTROM_TO_BOHR


def carbon_cloud(rng: np.random.Generator, n: int = 40, box: float = 12.0,
                 min_sep: float = 1.3) -> Geometry:
    pts: List[np.ndarray] = []
    while len(pts) < n:
        p = rng.uniform(0.0, box, size=3)
        if all(np.linalg.norm(p - q) >= min_sep for q in pts):
            pts.append(p)
    return ["C"] * n, np.array(pts) * ANGSTROM_TO_BOHR
