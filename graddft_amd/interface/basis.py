"""Contracted Gaussian basis functions: normalisation, shell packing, component order.

Stands in for the PySCF ``Mole`` basis machinery the reference relies on
(``grad_dft/interface/pyscf.py:824-827``).  A :class:`BasisSet` is a flat list of shells;
each shell carries its centre (bohr), angular momentum ``l``, primitive exponents and
fully normalised contraction coefficients, and whether it is expanded in real spherical
harmonics (``2l+1`` functions) or Cartesian monomials (``(l+1)(l+2)/2`` functions).

Conventions (fixed here, shared by every consumer):

* Cartesian monomials of a shell are ordered ``x^a y^b z^c`` with ``a`` descending, then
  ``b`` descending (``xx, xy, xz, yy, yz, zz`` for d), and each is individually
  normalised to 1.
* Spherical shells use real solid harmonics ordered ``m = -l..l``, except p shells which
  are ordered ``x, y, z``; each is normalised to 1.
* Coefficients multiply ``exp(-alpha r^2)``; the radial normalisation of the pure ``x^l``
  monomial is folded into them, so an AO is ``sum_k T[k] * P_k(r - R) * sum_i c_i
  exp(-alpha_i |r - R|^2)`` with ``T`` the shell's component-to-monomial matrix.
"""
from __future__ import annotations

import math
from dataclasses import dataclass, field
from typing import Dict, List, Sequence, Tuple

import numpy as np

from .basis_data import ATOMIC_NUMBER, BASIS_SETS

ANGSTROM_TO_BOHR = 1.0 / 0.52917721092  # grad_dft/utils/types.py:37 (Bohr2Angstroms)


def _dfact(n: int) -> float:
    """Double factorial with (-1)!! = 0!! = 1."""
    r = 1.0
    while n > 1:
        r *= n
        n -= 2
    return r


def cart_exponents(l: int) -> List[Tuple[int, int, int]]:
    """Monomial exponents (a, b, c) of a degree-l shell in the fixed order."""
    out = []
    for a in range(l, -1, -1):
        for b in range(l - a, -1, -1):
            out.append((a, b, l - a - b))
    return out


def _sphere_moment(a: int, b: int, c: int) -> float:
    """Integral of x^a y^b z^c over the unit sphere."""
    if a % 2 or b % 2 or c % 2:
        return 0.0
    return 4.0 * math.pi * _dfact(a - 1) * _dfact(b - 1) * _dfact(c - 1) / _dfact(a + b + c + 1)


def _poly(terms: Dict[Tuple[int, int, int], float], l: int) -> np.ndarray:
    cart = cart_exponents(l)
    v = np.zeros(len(cart))
    for k, coef in terms.items():
        v[cart.index(k)] = coef
    return v


def _solid_harmonic_rows(l: int) -> np.ndarray:
    """Unnormalised real solid harmonics r^l Y_lm as monomial coefficients."""
    if l == 0:
        rows = [{(0, 0, 0): 1.0}]
    elif l == 1:
        rows = [{(1, 0, 0): 1.0}, {(0, 1, 0): 1.0}, {(0, 0, 1): 1.0}]
    elif l == 2:
        rows = [
            {(1, 1, 0): 1.0},                                   # m=-2  xy
            {(0, 1, 1): 1.0},                                   # m=-1  yz
            {(0, 0, 2): 2.0, (2, 0, 0): -1.0, (0, 2, 0): -1.0},  # m=0   2zz-xx-yy
            {(1, 0, 1): 1.0},                                   # m=1   xz
            {(2, 0, 0): 1.0, (0, 2, 0): -1.0},                  # m=2   xx-yy
        ]
    elif l == 3:
        rows = [
            {(2, 1, 0): 3.0, (0, 3, 0): -1.0},                  # m=-3  3xxy-yyy
            {(1, 1, 1): 1.0},                                   # m=-2  xyz
            {(0, 1, 2): 4.0, (2, 1, 0): -1.0, (0, 3, 0): -1.0},  # m=-1  y(4zz-xx-yy)
            {(0, 0, 3): 2.0, (2, 0, 1): -3.0, (0, 2, 1): -3.0},  # m=0   z(2zz-3xx-3yy)
            {(1, 0, 2): 4.0, (3, 0, 0): -1.0, (1, 2, 0): -1.0},  # m=1   x(4zz-xx-yy)
            {(2, 0, 1): 1.0, (0, 2, 1): -1.0},                  # m=2   z(xx-yy)
            {(3, 0, 0): 1.0, (1, 2, 0): -3.0},                  # m=3   xxx-3xyy
        ]
    else:
        raise NotImplementedError(f"angular momentum {l} > 3 is not tabulated")
    return np.array([_poly(r, l) for r in rows])


def monomial_gram(l: int) -> np.ndarray:
    cart = cart_exponents(l)
    g = np.empty((len(cart), len(cart)))
    for i, (a, b, c) in enumerate(cart):
        for j, (d, e, f) in enumerate(cart):
            g[i, j] = _sphere_moment(a + d, b + e, c + f)
    return g


def component_matrix(l: int, cart: bool) -> np.ndarray:
    """T (ncomp, ncart): AO component k = sum_j T[k, j] * monomial_j * radial.

    Rows are scaled so that every component has unit norm given a radial part that
    normalises the pure x^l monomial (whose angular moment is 4 pi / (2l + 1))."""
    ref = 4.0 * math.pi / (2 * l + 1)
    if cart:
        rows = np.eye(len(cart_exponents(l)))
    else:
        rows = _solid_harmonic_rows(l)
    g = monomial_gram(l)
    norms = np.einsum("ki,ij,kj->k", rows, g, rows)
    return rows * np.sqrt(ref / norms)[:, None]


def primitive_norm(l: int, alpha: np.ndarray) -> np.ndarray:
    """Normalisation of x^l exp(-alpha r^2)."""
    return np.sqrt((2.0 * alpha / math.pi) ** 1.5 * (4.0 * alpha) ** l / _dfact(2 * l - 1))


@dataclass
class Shell:
    atom: int
    center: np.ndarray
    l: int
    exps: np.ndarray
    coefs: np.ndarray  # normalised, radial norm folded in
    cart: bool
    ao_offset: int = 0

    @property
    def nfunc(self) -> int:
        return (self.l + 1) * (self.l + 2) // 2 if self.cart else 2 * self.l + 1

    @property
    def ncart(self) -> int:
        return (self.l + 1) * (self.l + 2) // 2


def contract(l: int, exps: Sequence[float], coefs: Sequence[float]) -> Tuple[np.ndarray, np.ndarray]:
    """Fold primitive normalisation into ``coefs`` and normalise the contraction."""
    a = np.asarray(exps, dtype=np.float64)
    c = np.asarray(coefs, dtype=np.float64) * primitive_norm(l, a)
    ap = a[:, None] + a[None, :]
    s = (math.pi / ap) ** 1.5 * _dfact(2 * l - 1) / (2.0 * ap) ** l
    norm = float(c @ s @ c)
    return a, c / math.sqrt(norm)


@dataclass
class BasisSet:
    name: str
    shells: List[Shell] = field(default_factory=list)
    nao: int = 0

    def component_matrices(self) -> Dict[Tuple[int, bool], np.ndarray]:
        return {(s.l, s.cart): component_matrix(s.l, s.cart) for s in self.shells}

    def ao_labels(self) -> List[Tuple[int, int, int]]:
        out = []
        for si, s in enumerate(self.shells):
            for k in range(s.nfunc):
                out.append((s.atom, si, k))
        return out


def build_basis(symbols: Sequence[str], coords_bohr: np.ndarray, basis: str) -> BasisSet:
    """Shells for the atoms in order, each atom's shells in table order."""
    key = basis.lower()
    if key not in BASIS_SETS:
        raise KeyError(f"unknown basis {basis!r}; known: {sorted(BASIS_SETS)}")
    table, cart = BASIS_SETS[key]
    bs = BasisSet(name=key)
    off = 0
    for ia, (sym, xyz) in enumerate(zip(symbols, np.asarray(coords_bohr, dtype=np.float64))):
        if sym not in table:
            raise KeyError(f"basis {basis} has no entry for element {sym}")
        for l, prims in table[sym]:
            e, c = contract(l, [p[0] for p in prims], [p[1] for p in prims])
            sh = Shell(atom=ia, center=xyz.copy(), l=l, exps=e, coefs=c,
                       cart=cart and l >= 2, ao_offset=off)
            off += sh.nfunc
            bs.shells.append(sh)
    bs.nao = off
    return bs


def nuclear_repulsion(charges: Sequence[float], coords_bohr: np.ndarray) -> float:
    e = 0.0
    r = np.asarray(coords_bohr)
    for i in range(len(charges)):
        for j in range(i):
            e += charges[i] * charges[j] / np.linalg.norm(r[i] - r[j])
    return float(e)


def atomic_numbers(symbols: Sequence[str]) -> List[int]:
    return [ATOMIC_NUMBER[s] for s in symbols]


# ---- flat descriptors consumed by the device AO kernel ---------------------------------

MAX_L = 3
MAX_NCART = 10


def pack_shells(bs: BasisSet) -> Dict[str, np.ndarray]:
    """Flat arrays describing every shell for the device AO-evaluation kernel.

    ``shell_i32``  (nshell, 6): l, nprim, prim_offset, ao_offset, nfunc, tmat_offset
    ``shell_f64``  (nshell, 3): centre
    ``prim_f64``   (nprim, 2):  exponent, coefficient
    ``tmat_f64``   (sum nfunc*ncart,): component matrices, row-major per shell
    """
    si, sf, pf, tm = [], [], [], []
    poff = 0
    toff = 0
    for s in bs.shells:
        T = component_matrix(s.l, s.cart)
        si.append([s.l, len(s.exps), poff, s.ao_offset, s.nfunc, toff])
        sf.append(s.center)
        for a, c in zip(s.exps, s.coefs):
            pf.append([a, c])
        tm.append(T.ravel())
        poff += len(s.exps)
        toff += T.size
    return {
        "shell_i32": np.ascontiguousarray(np.array(si, dtype=np.int32)),
        "shell_f64": np.ascontiguousarray(np.array(sf, dtype=np.float64)),
        "prim_f64": np.ascontiguousarray(np.array(pf, dtype=np.float64)),
        "tmat_f64": np.ascontiguousarray(np.concatenate(tm).astype(np.float64)),
    }
