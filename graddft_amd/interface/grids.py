"""Atom-centred Becke integration grids (radial x Lebedev, pruned, fuzzy-cell weights).

The reference takes ``mf.grids`` (coords, weights) from PySCF's ``dft.Grids`` at a given
``level`` (``grad_dft/interface/pyscf.py:54-60``, ``:1048``).  PySCF is absent, so this
module builds grids with the same published ingredients PySCF's defaults use:

* Treutler-Ahlrichs M4 radial quadrature with element scaling xi
  (Treutler & Ahlrichs, J. Chem. Phys. 102, 346 (1995));
* Lebedev-Laikov angular rules (``scipy.integrate.lebedev_rule``);
* NWChem-style radial pruning of the angular order;
* Becke fuzzy-cell partition (3 iterations of p(mu) = (3 mu - mu^3)/2) with Treutler's
  square-root Bragg-radius size adjustment;
* a level table giving (radial points, Lebedev order) per element row, and removal of
  points whose weight is below 1e-15.

Point sets are *not* bit-identical to PySCF's; the workloads only need grids of the
stated level and size class.  ``molecular_grid(..., weights_on="device")`` evaluates the Becke
partition (the O(points x atoms^2) part) with the ``gdx_becke_weights`` kernel; the atomic rules
stay on the host.
"""
from __future__ import annotations

import math
from functools import lru_cache
from typing import Sequence, Tuple

import numpy as np
from scipy.integrate import lebedev_rule

# Bragg radii in angstrom, indexed by atomic number (0 unused).
_BRAGG_ANGSTROM = [0.0, 0.35, 1.40, 1.45, 1.05, 0.85, 0.70, 0.65, 0.60, 0.50, 1.50]
BRAGG_BOHR = np.array(_BRAGG_ANGSTROM) / 0.52917721092

# Treutler-Ahlrichs xi per element (H..Ne).
_TA_XI = [1.0, 0.8, 0.9, 1.8, 1.4, 1.3, 1.1, 0.9, 0.9, 0.9, 0.9]

# level -> (radial points, Lebedev order) for rows 1 (H-He) and 2 (Li-Ne)
_RAD = {0: (10, 15), 1: (30, 40), 2: (40, 60), 3: (50, 75), 4: (60, 90), 5: (70, 105),
        6: (80, 120), 7: (90, 135), 8: (100, 150), 9: (200, 200)}
_ANG_ORDER = {0: (11, 15), 1: (17, 23), 2: (23, 29), 3: (29, 29), 4: (35, 41), 5: (41, 47),
              6: (47, 53), 7: (59, 59), 8: (71, 71), 9: (131, 131)}

# Lebedev point counts and the polynomial order of each rule.
_LEB_NPTS = [1, 6, 14, 26, 38, 50, 74, 86, 110, 146, 170, 194, 230, 266, 302, 350, 434,
             590, 770, 974, 1202, 1454, 1730, 2030, 2354, 2702, 3074, 3470, 3890, 4334,
             4802, 5294, 5810]
_LEB_ORDER = [0, 3, 5, 7, 9, 11, 13, 15, 17, 19, 21, 23, 25, 27, 29, 31, 35, 41, 47, 53,
              59, 65, 71, 77, 83, 89, 95, 101, 107, 113, 119, 125, 131]


def _row(z: int) -> int:
    return 0 if z <= 2 else 1


@lru_cache(maxsize=None)
def lebedev(npts: int) -> Tuple[np.ndarray, np.ndarray]:
    """Unit-sphere points (npts, 3) and weights summing to 4 pi."""
    order = _LEB_ORDER[_LEB_NPTS.index(npts)]
    x, w = lebedev_rule(order)
    return np.ascontiguousarray(x.T), np.ascontiguousarray(w)


def treutler_ahlrichs(n: int, z: int) -> Tuple[np.ndarray, np.ndarray]:
    """M4 radial nodes r and Jacobian dr (ascending r)."""
    xi = _TA_XI[z]
    i = np.arange(1, n + 1)
    step = math.pi / (n + 1)
    x = np.cos(i * step)
    ln2 = xi / math.log(2.0)
    r = -ln2 * (1 + x) ** 0.6 * np.log((1 - x) / 2)
    dr = step * np.sin(i * step) * ln2 * (1 + x) ** 0.6 * (
        -0.6 / (1 + x) * np.log((1 - x) / 2) + 1 / (1 - x))
    return r[::-1].copy(), dr[::-1].copy()


def nwchem_prune(z: int, rads: np.ndarray, n_ang: int) -> np.ndarray:
    """Angular point count per radial shell (NWChem partitioning of r / R_Bragg)."""
    alphas = np.array(((0.25, 0.5, 1.0, 4.5), (0.1667, 0.5, 0.9, 3.5), (0.1, 0.4, 0.8, 2.5)))
    leb = np.array(_LEB_NPTS[4:])
    if n_ang < 50:
        return np.full(len(rads), n_ang)
    if n_ang == 50:
        leb_l = np.array([1, 2, 2, 2, 1])
    else:
        idx = int(np.where(leb == n_ang)[0][0])
        leb_l = np.array([1, 3, idx - 1, idx, idx - 1])
    r_atom = BRAGG_BOHR[z] + 1e-200
    a = alphas[0] if z <= 2 else (alphas[1] if z <= 10 else alphas[2])
    place = ((rads / r_atom)[:, None] > a[None, :]).sum(axis=1)
    return leb[leb_l[place]]


def atomic_grid(z: int, level: int) -> Tuple[np.ndarray, np.ndarray]:
    """Atom-centred points (relative to the nucleus) and their quadrature volume."""
    nrad = _RAD[level][_row(z)]
    n_ang = _LEB_NPTS[_LEB_ORDER.index(_ANG_ORDER[level][_row(z)])]
    r, dr = treutler_ahlrichs(nrad, z)
    rw = r * r * dr
    angs = nwchem_prune(z, r, n_ang)
    coords, vol = [], []
    for n in sorted(set(angs.tolist())):
        pts, w = lebedev(int(n))
        idx = np.where(angs == n)[0]
        coords.append((r[idx][None, :, None] * pts[:, None, :]).reshape(-1, 3))
        vol.append((rw[idx][None, :] * w[:, None]).ravel())
    return np.concatenate(coords), np.concatenate(vol)


def becke_weights(coords: np.ndarray, owner: int, atom_xyz: np.ndarray,
                  charges: Sequence[int]) -> np.ndarray:
    """Becke cell function of atom ``owner`` at ``coords``, normalised over atoms."""
    natm = len(atom_xyz)
    if natm == 1:
        return np.ones(len(coords))
    adist, a = becke_adjustment(charges, atom_xyz)
    dist = np.linalg.norm(coords[None, :, :] - atom_xyz[:, None, :], axis=2)
    p = np.ones((natm, len(coords)))
    for i in range(natm):
        for j in range(i):
            g = (dist[i] - dist[j]) / adist[i, j]
            g = g + a[i, j] * (1.0 - g * g)
            for _ in range(3):
                g = (3.0 - g * g) * g * 0.5
            p[i] *= 0.5 * (1.0 - g)
            p[j] *= 0.5 * (1.0 + g)
    return p[owner] / p.sum(axis=0)


def becke_adjustment(charges: Sequence[int], atom_xyz: np.ndarray) -> Tuple[np.ndarray, np.ndarray]:
    """Inter-atomic distances |R_i - R_j| and the size adjustment a_ij of the Becke partition."""
    rad = np.sqrt(BRAGG_BOHR[np.asarray(charges)]) + 1e-200
    rr = rad[:, None] / rad[None, :]
    a = np.clip(0.25 * (rr.T - rr), -0.5, 0.5)
    adist = np.linalg.norm(atom_xyz[:, None, :] - atom_xyz[None, :, :], axis=2)
    return adist, a


def becke_weights_device(coords: np.ndarray, owner: np.ndarray, vol: np.ndarray, atom_xyz: np.ndarray,
                         charges: Sequence[int], device: str = "cuda") -> np.ndarray:
    """vol * Becke cell function of each point's owner atom, evaluated by ``gdx_becke_weights``."""
    import torch
    from .. import _lib
    lib = _lib.load()
    natm = len(atom_xyz)
    adist, a = becke_adjustment(charges, atom_xyz)
    dev = lambda x, dt=torch.float64: torch.as_tensor(np.ascontiguousarray(x), dtype=dt, device=device)  # noqa
    c, o, v = dev(coords), dev(owner, torch.int32), dev(vol)
    at, ad, aa = dev(atom_xyz), dev(adist), dev(a)
    w = torch.empty(len(coords), dtype=torch.float64, device=device)
    st = torch.cuda.current_stream(w.device).cuda_stream
    _lib.check(lib.gdx_becke_weights(st, len(coords), c.data_ptr(), o.data_ptr(), v.data_ptr(), natm, at.data_ptr(),
                                     ad.data_ptr(), aa.data_ptr(), w.data_ptr()), "gdx_becke_weights")
    return w.cpu().numpy()


def molecular_grid(charges: Sequence[int], atom_xyz_bohr: np.ndarray, level: int,
                   cutoff: float = 1e-15, weights_on: str = "host") -> Tuple[np.ndarray, np.ndarray]:
    """All atoms' points with Becke-partitioned weights; tiny weights removed.  weights_on:
    "host" (numpy) or "device" (the gdx_becke_weights kernel; same formulas, same order)."""
    if weights_on not in ("host", "device"):
        raise ValueError(f"weights_on must be 'host' or 'device', not {weights_on!r}")
    atom_xyz = np.asarray(atom_xyz_bohr, dtype=np.float64)
    cs, vs, owners = [], [], []
    for ia, z in enumerate(charges):
        c, v = atomic_grid(int(z), level)
        cs.append(c + atom_xyz[ia])
        vs.append(v)
        owners.append(np.full(len(v), ia, dtype=np.int32))
    if weights_on == "device":
        coords = np.concatenate(cs)
        weights = becke_weights_device(coords, np.concatenate(owners), np.concatenate(vs), atom_xyz, charges)
    else:
        ws = [v * becke_weights(c, ia, atom_xyz, charges) for ia, (c, v) in enumerate(zip(cs, vs))]
        coords = np.concatenate(cs)
        weights = np.concatenate(ws)
    keep = np.abs(weights) > cutoff
    return np.ascontiguousarray(coords[keep]), np.ascontiguousarray(weights[keep])
