"""Oracle SCF path (numpy float64) -- TEST INFRASTRUCTURE ONLY.

Restates, term by term:

* non-XC energy and Coulomb potential -- ``grad_dft/molecule.py:694-811``
  (nonXC, one_body_energy, coulomb_energy, coulomb_potential) on D = D_up + D_dn;
* the energy/Fock predictor -- ``grad_dft/train.py:123-216`` (predict): Fock = h1e + J +
  dE_xc/dD, abs_clip, symmetrise, abs_clip, plus the explicit exact-exchange term
  V + V^T (``:200-206``), clips;
* B3LYP exact exchange: the reference's semi-numerical term built from the *initial*
  density (``molecule.py:507-613``, ``interface/pyscf.py:152-160``) is taken here from the
  ERI as K[D0] ("stale" exchange): E_x^HF = -(a0/2) sum_s tr(D_s K[D0_s]), Fock term
  -a0 K[D0_s];
* ``JittableDiis`` -- ``grad_dft/evaluate.py:1041-1205`` with its exact buffer semantics
  (the write at cycle == max_diis is dropped; shifting starts at cycle > max_diis; border
  +1 for active slots, diagonal 1 for inactive ones; x = B^-1 e_0);
* ``safe_fock_solver`` / ``safe_general_eigh`` -- ``grad_dft/utils/eigenproblem.py:110-149``;
* ``get_occ``, ``make_rdm1`` -- ``grad_dft/molecule.py:813-889``;
* ``diff_scf_loop`` -- ``grad_dft/evaluate.py:917-1038`` (the discarded final iteration
  is not computed: it does not change the returned molecule).

Pinning: H2/STO-3G textbook energies (Szabo & Ostlund) for the integrals and RHF limit;
brute-force loops for J and K; idempotency D S D = D and tr(D S) = N_e/2 per spin;
DIIS quirks checked against hand-built buffers in ``tests/test_oracle_scf.py``.
"""
from __future__ import annotations

from dataclasses import dataclass, field
from typing import Callable, List, Optional, Tuple

import numpy as np

CLIP = 1e-30


def abs_clip(x, thr=CLIP):
    return np.where(np.abs(x) > thr, x, 0.0)


def coulomb_potential(D, eri):
    """molecule.py:788-811: J_pq = sum_rt (pq|rt) D_rt."""
    return np.einsum("pqrt,rt->pq", eri, D)


def exchange_potential(D, eri):
    """K_ac = sum_bd (ab|cd) D_bd."""
    return np.einsum("abcd,bd->ac", eri, D)


def non_xc(D, h1e, eri, e_nuc):
    """molecule.py:694-733 on D = D_up + D_dn."""
    J = coulomb_potential(D, eri)
    return e_nuc + np.einsum("ij,ij->", D, h1e) + np.einsum("pq,pq->", D, J) / 2.0


def hf_exchange_energy(rdm1, K0, a0):
    return -0.5 * a0 * sum(np.einsum("ij,ij->", rdm1[s], K0[s]) for s in range(2))


@dataclass
class SCFInputs:
    h1e: np.ndarray
    s1e: np.ndarray
    eri: np.ndarray            # dense (n,n,n,n)
    e_nuc: float
    K0: Optional[np.ndarray]   # (2,n,n) exchange of the initial density, or None
    a0: float = 0.2


def predict(xc_value_and_vxc: Callable, inp: SCFInputs, rdm1: np.ndarray):
    """train.py:123-216 with the stale-exchange B3LYP term: returns (E, fock (2,n,n))."""
    exc, vxc = xc_value_and_vxc(rdm1)
    Dt = rdm1.sum(axis=0)
    J = coulomb_potential(Dt, inp.eri)
    energy = exc + non_xc(Dt, inp.h1e, inp.eri, inp.e_nuc)
    fock = inp.h1e[None] + J[None] + vxc
    fock = abs_clip(fock)
    fock = 0.5 * (fock + fock.transpose(0, 2, 1))
    fock = abs_clip(fock)
    if inp.K0 is not None:
        energy += hf_exchange_energy(rdm1, inp.K0, inp.a0)
        v = -0.5 * inp.a0 * inp.K0
        fock = fock + v + v.transpose(0, 2, 1)
        fock = abs_clip(fock)
    fock = abs_clip(fock)
    return energy, fock


@dataclass
class JittableDiis:
    """evaluate.py:1041-1205 (A = I)."""
    s1e: np.ndarray
    max_diis: int = 10

    def zeros(self, n):
        m = self.max_diis
        return [np.zeros((m, 2, n, n)), np.zeros((m, 2, n, n)), np.zeros(m), np.zeros((m, 2, n, n))]

    def update(self, rdm1, fock, energy, data, cycle):
        dens, focks, energies, errs = [d.copy() for d in data]
        fds = np.einsum("sjk,skl,lm->sjm", fock, rdm1, self.s1e)
        err = fds - fds.transpose(0, 2, 1)
        if cycle > self.max_diis:
            errs = np.concatenate([errs, err[None]])[1:]
            dens = np.concatenate([dens, rdm1[None]])[1:]
            focks = np.concatenate([focks, fock[None]])[1:]
            energies = np.concatenate([energies, [energy]])[1:]
        elif cycle < self.max_diis:
            errs[cycle] = err
            dens[cycle] = rdm1
            focks[cycle] = fock
            energies[cycle] = energy
        # cycle == max_diis: the out-of-bounds .at[].set is dropped by JAX
        return [dens, focks, energies, errs]

    def coefficients(self, errs, cycle):
        m = errs.shape[0]
        gram = np.einsum("iskl,jskl->sij", errs, errs)
        B = np.zeros((2, m + 1, m + 1))
        B[:, 1:, 1:] = gram
        for i in range(m + 2):
            v = 1.0 if i <= cycle else 0.0
            if i + 1 <= m:
                B[:, 0, i + 1] = v
                B[:, i + 1, 0] = v
        for i in range(m + 2):
            if i + 1 <= m:
                B[:, i + 1, i + 1] = gram[:, i, i] if i <= cycle else 1.0
        C = np.zeros(m + 1)
        C[0] = 1.0
        x = np.stack([np.linalg.inv(B[s]) @ C for s in range(2)])
        return x[:, 1:]

    def run(self, rdm1, fock, energy, data, cycle):
        data = self.update(rdm1, fock, energy, data, cycle)
        x = self.coefficients(data[3], cycle)
        F = np.einsum("si,isjk->sjk", x, data[1])
        return F, data


def safe_general_eigh(A, B):
    """eigenproblem.py:110-129."""
    L = np.linalg.cholesky(B)
    Linv = np.linalg.inv(L)
    C = Linv @ A @ Linv.T
    w, V = np.linalg.eigh(C)
    return w, Linv.T @ V


def safe_fock_solver(fock, s1e):
    """eigenproblem.py:132-149."""
    out = [safe_general_eigh(fock[s], s1e) for s in range(2)]
    return np.stack([o[0] for o in out]), np.stack([o[1] for o in out])


def get_occ(mo_energy, nelecs):
    """molecule.py:851-889: the lowest nelec_s orbitals (stable argsort) get 1."""
    occ = np.zeros_like(mo_energy)
    for s in range(2):
        idx = np.argsort(mo_energy[s], kind="stable")
        occ[s, idx[: int(nelecs[s])]] = 1.0
    return occ


def make_rdm1(mo_coeff, mo_occ):
    """molecule.py:816-846."""
    return np.einsum("sij,sj,skj->sik", mo_coeff, mo_occ, mo_coeff)


@dataclass
class SCFState:
    rdm1: np.ndarray
    fock: np.ndarray
    energy: float
    mo_energy: np.ndarray
    mo_coeff: np.ndarray
    mo_occ: np.ndarray
    energies: List[float] = field(default_factory=list)


def diff_scf_loop(xc_value_and_vxc: Callable, inp: SCFInputs, rdm1: np.ndarray, mo_occ: np.ndarray,
                  cycles: int) -> SCFState:
    """evaluate.py:938-1038."""
    e, fock = predict(xc_value_and_vxc, inp, rdm1)
    diis = JittableDiis(inp.s1e, 10)
    data = diis.zeros(rdm1.shape[-1])
    nelecs = mo_occ.sum(axis=1).astype(np.int64)
    st = SCFState(rdm1, fock, e, None, None, mo_occ, [e])
    for cycle in range(cycles):
        F, data = diis.run(st.rdm1, st.fock, st.energy, data, cycle)
        w, C = safe_fock_solver(F, inp.s1e)
        occ = get_occ(w, nelecs)
        D = make_rdm1(C, occ)
        e, fock = predict(xc_value_and_vxc, inp, D)
        st = SCFState(D, fock, e, w, C, occ, st.energies + [e])
    return st
