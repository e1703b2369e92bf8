"""Oracle for the semi-numerical exact exchange -- TEST INFRASTRUCTURE ONLY.

Restates, in numpy float64:

* ``nu_on_grid``: nu_ab(r) = int dr' phi_a(r') f(|r - r'|) phi_b(r'), f = 1/|r-r'| (omega = 0) or
  erf(omega |r-r'|)/|r-r'|.  The reference takes these from PySCF (``int1e_grids_sph`` under
  ``mol.with_range_coulomb(omega)``, ``grad_dft/external/_hf_density.py:18-67``); here they come
  from the McMurchie-Davidson formula, vectorised over grid points, with the Boys function from
  scipy's regularised incomplete gamma function (F_n(T) = gamma(n + 1/2, T) / (2 T^(n + 1/2))).
* ``chi_tensor``: ``generate_chi_tensor`` (``grad_dft/interface/pyscf.py:1067-1126``),
  chi[r, s, a] = sum_bd D_s[b][d] phi_b(r) nu_da(r) (``chi_make``, one omega).
* ``hf_energy_density``: ``HF_energy_density`` (``grad_dft/molecule.py:507-541``),
  e[s, r] = -1/2 sum_ac chi[r, s, c] D_s[a][c] phi_a(r).
* ``hf_fock``: ``HF_density_grad_2_Fock`` (``grad_dft/molecule.py:545-613``) for a given
  gr = dE_xc/de_HF: V[s, a, b] = -1/2 sum_r gr[s, r] phi_a(r) chi[r, s, b].

Pinned by (``tests/test_oracle_hf.py``): the closed form of s-type nu integrals (Boys F_0 from
math.erf, plain and attenuated), the host integral library's nuclear-attraction integrals of
a unit charge (an independent McMurchie-Davidson code with a series Boys function) for shells
up to f, the far-field limit nu_ab(C) -> S_ab / |C - P|, the small-omega limit
nu -> 2 omega S / sqrt(pi), and the exchange energy: the grid quadrature of e_HF at D = D0
reproduces -1/2 sum D0 D0 (ab|cd) from the packed ERI to the quadrature error.
"""
from __future__ import annotations

import math

import numpy as np
import scipy.special


def boys(nmax: int, T: np.ndarray) -> np.ndarray:
    """F_n(T) for n = 0..nmax, shape (nmax + 1, len(T))."""
    T = np.asarray(T, dtype=np.float64)
    out = np.empty((nmax + 1,) + T.shape)
    small = T < 1e-3
    Ts = np.where(small, 1.0, T)
    for n in range(nmax + 1):
        a = n + 0.5
        big = scipy.special.gammainc(a, Ts) * scipy.special.gamma(a) / (2.0 * Ts ** a)
        # series for small T: sum_k (-T)^k / (k! (2n + 2k + 1))
        ser = np.zeros_like(T)
        term = np.ones_like(T)
        for k in range(12):
            ser = ser + term / (2 * n + 2 * k + 1)
            term = term * (-T) / (k + 1)
        out[n] = np.where(small, ser, big)
    return out


def _hermite_e(la: int, lb: int, a: float, b: float, xab: float) -> np.ndarray:
    """Hermite expansion coefficients E[i, j, t] of one Cartesian factor x_A^i x_B^j of the
    Gaussian product (Helgaker, Jorgensen, Olsen eq. 9.5.6-9.5.7)."""
    p = a + b
    mu = a * b / p
    xpa, xpb = -b * xab / p, a * xab / p
    E = np.zeros((la + 1, lb + 1, la + lb + 2))
    E[0, 0, 0] = math.exp(-mu * xab * xab)
    for i in range(la + 1):
        for j in range(lb + 1):
            if i == 0 and j == 0:
                continue
            for t in range(i + j + 1):
                if i > 0:
                    v = xpa * E[i - 1, j, t] + (t + 1) * E[i - 1, j, t + 1]
                    if t > 0:
                        v += E[i - 1, j, t - 1] / (2 * p)
                else:
                    v = xpb * E[i, j - 1, t] + (t + 1) * E[i, j - 1, t + 1]
                    if t > 0:
                        v += E[i, j - 1, t - 1] / (2 * p)
                E[i, j, t] = v
    return E


def _hermite_r(L: int, p: float, PC: np.ndarray, F: np.ndarray) -> dict:
    """R^0_tuv(p, PC) for t + u + v <= L at every point (PC (N, 3), F = boys(L, p |PC|^2))."""
    X, Y, Z = PC[:, 0], PC[:, 1], PC[:, 2]
    R = {}
    for n in range(L + 1):
        R[(n, 0, 0, 0)] = (-2.0 * p) ** n * F[n]

    def get(n, t, u, v):
        if t < 0 or u < 0 or v < 0:
            return 0.0
        key = (n, t, u, v)
        if key not in R:
            if t > 0:
                R[key] = (t - 1) * get(n + 1, t - 2, u, v) + X * get(n + 1, t - 1, u, v)
            elif u > 0:
                R[key] = (u - 1) * get(n + 1, t, u - 2, v) + Y * get(n + 1, t, u - 1, v)
            else:
                R[key] = (v - 1) * get(n + 1, t, u, v - 2) + Z * get(n + 1, t, u, v - 1)
        return R[key]

    return {(t, u, v): get(0, t, u, v) for t in range(L + 1) for u in range(L + 1 - t) for v in range(L + 1 - t - u)}


def nu_on_grid(basis, coords: np.ndarray, omega: float = 0.0) -> np.ndarray:
    """nu (N, n, n) at the points `coords` (N, 3), bohr."""
    from graddft_amd.interface.basis import cart_exponents, component_matrix

    coords = np.asarray(coords, dtype=np.float64)
    N, n = len(coords), basis.nao
    nu = np.zeros((N, n, n))
    shells = basis.shells
    for ia, A in enumerate(shells):
        for ib in range(ia + 1):
            B = shells[ib]
            ea, eb = cart_exponents(A.l), cart_exponents(B.l)
            L = A.l + B.l
            cart = np.zeros((N, len(ea), len(eb)))
            AB = np.asarray(A.center) - np.asarray(B.center)
            for a, ca in zip(A.exps, A.coefs):
                for b, cb in zip(B.exps, B.coefs):
                    p = a + b
                    P = (a * np.asarray(A.center) + b * np.asarray(B.center)) / p
                    Ex = [_hermite_e(A.l, B.l, a, b, AB[d]) for d in range(3)]
                    PC = P[None, :] - coords
                    pe, scale = p, 1.0
                    if omega > 0.0:
                        pe = p * omega ** 2 / (p + omega ** 2)
                        scale = math.sqrt(omega ** 2 / (p + omega ** 2))
                    F = boys(L, pe * np.einsum("ij,ij->i", PC, PC))
                    R = _hermite_r(L, pe, PC, F)
                    pref = 2.0 * math.pi / p * ca * cb * scale
                    for i, (ix, iy, iz) in enumerate(ea):
                        for j, (jx, jy, jz) in enumerate(eb):
                            acc = np.zeros(N)
                            for t in range(ix + jx + 1):
                                for u in range(iy + jy + 1):
                                    for v in range(iz + jz + 1):
                                        acc += Ex[0][ix, jx, t] * Ex[1][iy, jy, u] * Ex[2][iz, jz, v] * R[(t, u, v)]
                            cart[:, i, j] += pref * acc
            TA, TB = component_matrix(A.l, A.cart), component_matrix(B.l, B.cart)
            blk = np.einsum("ai,nij,bj->nab", TA, cart, TB)
            sa = slice(A.ao_offset, A.ao_offset + A.nfunc)
            sb = slice(B.ao_offset, B.ao_offset + B.nfunc)
            nu[:, sa, sb] = blk
            nu[:, sb, sa] = blk.transpose(0, 2, 1)
    return nu


def chi_tensor(D0: np.ndarray, ao: np.ndarray, nu: np.ndarray) -> np.ndarray:
    """chi (N, nspin, n): chi_make of generate_chi_tensor, einsum("...bd,b,da->...a")."""
    return np.einsum("sbd,rb,rda->rsa", D0, ao, nu)


def hf_energy_density(D: np.ndarray, ao: np.ndarray, chi: np.ndarray) -> np.ndarray:
    """e_HF (nspin, N): -1/2 einsum("wsc,sac,a->ws") per point, one omega."""
    return -0.5 * np.einsum("rsc,sac,ra->sr", chi, D, ao)


def hf_fock(gr: np.ndarray, ao: np.ndarray, chi: np.ndarray) -> np.ndarray:
    """V (nspin, n, n) = -1/2 einsum("rws,wsr,ra->wsa") over points for one omega."""
    return -0.5 * np.einsum("sr,ra,rsb->sab", gr, ao, chi)
