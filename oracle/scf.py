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

import dataclasses
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


def _pair_index(n):
    """pair(i, j) = i (i + 1) / 2 + j for i >= j, symmetric (n, n)."""
    i, j = np.tril_indices(n)
    pair = np.zeros((n, n), dtype=np.int64)
    pair[i, j] = np.arange(len(i))
    pair[j, i] = pair[i, j]
    return pair


def eri_pair_matrix(eri_packed, n):
    """The 8-fold packed ERI as the symmetric (P, P) matrix E[pair(p,q), pair(r,s)] = (pq|rs),
    P = n (n + 1) / 2 (memory 8 P^2 bytes instead of the dense tensor's 8 n^4)."""
    P = n * (n + 1) // 2
    E = np.zeros((P, P))
    for I in range(P):
        E[I, : I + 1] = eri_packed[I * (I + 1) // 2: I * (I + 1) // 2 + I + 1]
    E = E + E.T
    E[np.diag_indices(P)] *= 0.5
    return E


def coulomb_potential_pairs(D, E):
    """J_pq = sum_rt (pq|rt) D_rt from the pair matrix (same as coulomb_potential)."""
    n = D.shape[-1]
    pair = _pair_index(n)
    i, j = np.tril_indices(n)
    d = np.where(i == j, 1.0, 2.0) * 0.5 * (D[i, j] + D[j, i])
    return (E @ d)[pair]


def exchange_potential_pairs(D, E):
    """K_ac = sum_bd (ab|cd) D_bd from the pair matrix (same as exchange_potential)."""
    n = D.shape[-1]
    pair = _pair_index(n)
    K = np.empty((n, n))
    for a in range(n):
        Ma = E[pair[a]][:, pair]            # (b, c, d) = (ab|cd)
        K[a] = np.einsum("bcd,bd->c", Ma, D)
    return K


def exchange_ordered_matrix(eri_packed, n):
    """M[a n + c, b n + d] = (ab|cd) over ordered pairs from the 8-fold packed ERI: the
    exchange operator as an (n^2, n^2) matrix, K = M vec(D) (the layout of gdx_exchange_xo)."""
    pair = _pair_index(n)
    P1 = pair[:, None, :, None]                  # pair(a, b) at [a, c, b, d]
    P2 = pair[None, :, None, :]                  # pair(c, d)
    hi, lo = np.maximum(P1, P2), np.minimum(P1, P2)
    return eri_packed[hi * (hi + 1) // 2 + lo].reshape(n * n, n * n)


def exchange_potential_ordered(D, M):
    """K_ac = sum_bd (ab|cd) D_bd as the product M vec(D) (same as exchange_potential)."""
    n = D.shape[-1]
    return (M @ D.reshape(n * n)).reshape(n, n)


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
    eri_pairs: Optional[np.ndarray] = None   # (P, P) pair matrix in place of the dense eri

    def coulomb(self, D):
        if self.eri_pairs is not None:
            return coulomb_potential_pairs(D, self.eri_pairs)
        return coulomb_potential(D, self.eri)


def predict(xc_value_and_vxc: Callable, inp: SCFInputs, rdm1: np.ndarray):
    """train.py:123-216 with the stale-exchange B3LYP term: returns (E, fock (2,n,n))."""
    exc, vxc = xc_value_and_vxc(rdm1)
    Dt = rdm1.sum(axis=0)
    J = inp.coulomb(Dt)
    energy = exc + inp.e_nuc + np.einsum("ij,ij->", Dt, inp.h1e) + np.einsum("pq,pq->", Dt, J) / 2.0
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
                  cycles: int, exchange_update: Optional[Callable] = None) -> SCFState:
    """evaluate.py:938-1038.  exchange_update(D) -> K, when given, rebuilds the exact-exchange
    matrices from every density before it is predicted at (the "current" exchange of the
    device loop; the reference's chi stays at the initial density)."""
    if exchange_update is not None:
        inp = dataclasses.replace(inp, K0=exchange_update(rdm1))
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
        if exchange_update is not None:
            inp = dataclasses.replace(inp, K0=exchange_update(D))
        e, fock = predict(xc_value_and_vxc, inp, D)
        st = SCFState(D, fock, e, w, C, occ, st.energies + [e])
    return st


# ---------------------------------------------------------------------------------------
# scf_loop / simple_scf_loop and their pieces (grad_dft/evaluate.py:131-592, 1211-1545;
# grad_dft/molecule.py:344-373).
# ---------------------------------------------------------------------------------------

HARTREE2KCALMOL = 627.50947   # utils/types.py:35


def damping(s, d, f, factor):
    """evaluate.py:1477-1503."""
    dm_vir = np.eye(s.shape[0]) - s @ d
    f0 = dm_vir @ f @ d @ s
    f0 = (f0 + f0.T) * (factor / (factor + 1.0))
    return f - f0


def level_shift(s, d, f, factor):
    """evaluate.py:1506-1545."""
    return f + (s - s @ d @ s) * factor


def orbital_grad(mo_coeff, mo_occ, F):
    """molecule.py:344-373: (n, n), summed over spins."""
    out = np.zeros(mo_coeff.shape[1:])
    for s in range(2):
        occ = mo_occ[s] > 0
        vir = mo_occ[s] == 0
        Cocc = mo_coeff[s] * occ[None, :]
        Cvir = mo_coeff[s] * vir[None, :]
        out += Cvir.T @ F[s] @ Cocc
    return out


def _bfgs_weights(costf_and_grad, nx):
    """jax.scipy.optimize.minimize(costf, ones(nx), method="BFGS", tol=1e-9) -> (c, fun)."""
    from scipy.optimize import minimize
    res = minimize(costf_and_grad, np.ones(nx), jac=True, method="BFGS", tol=1e-9)
    return res.x ** 2 / np.sum(res.x ** 2), float(res.fun)


def _weights_cost(lin, quad):
    """cost(x) = c.lin + c^T quad c with c = x^2 / |x|^2, and its gradient in x."""
    def f(x):
        n2 = float(np.sum(x * x))
        c = x * x / n2
        val = float(lin @ c + c @ quad @ c)
        dc = lin + quad @ c + quad.T @ c                     # d cost / d c
        jac = (2.0 * x / n2) * dc - (2.0 * x / n2) * float(dc @ c)
        return val, jac
    return f


@dataclass
class Diis:
    """evaluate.py:1211-1474 with Python lists for the history."""
    s1e: np.ndarray
    A: np.ndarray
    max_diis: int = 8
    diis_method: str = "EDIIS2"
    ediis2_threshold: float = 1e-2
    adiis2_threshold: float = 1e-2

    def update(self, rdm1, fock, energy, data):
        dens, focks, energies, errs = [list(d) for d in data]
        fds = np.stack([self.A @ fock[s] @ rdm1[s] @ self.s1e @ self.A.T for s in range(2)])
        errs.append(fds - fds.transpose(0, 2, 1))
        dens.append(rdm1)
        focks.append(fock)
        energies.append(float(energy))
        if len(errs) > self.max_diis:
            dens, focks, energies, errs = dens[1:], focks[1:], energies[1:], errs[1:]
        return dens, focks, energies, errs

    def run(self, rdm1, fock, energy, data, cycle=0):
        data = self.update(rdm1, fock, energy, data)
        dens, focks, energies, errs = data
        if len(errs) == 1:
            return focks[0], data
        rel = (energies[-1] - energies[-2]) / energies[-2]
        m = self.diis_method
        if m == "CDIIS" or (m == "EDIIS2" and rel < self.ediis2_threshold) or (m == "ADIIS2" and rel < self.adiis2_threshold):
            x = self.cdiis_minimize(np.stack(errs))
            F = np.einsum("si,isjk->sjk", x, np.stack(focks))
            return np.einsum("ji,sjk,kl->sil", self.A, F, self.A), data
        if m in ("EDIIS", "EDIIS2"):
            x, _ = self.ediis_minimize(np.stack(dens), np.stack(focks), np.array(energies))
        elif m in ("ADIIS", "ADIIS2"):
            x, _ = self.adiis_minimize(np.stack(dens), np.stack(focks), cycle % self.max_diis)
        else:
            raise ValueError(m)
        return np.einsum("i,isjk->sjk", x, np.stack(focks)), data

    @staticmethod
    def cdiis_minimize(errs):
        k = errs.shape[0]
        xs = []
        for s in range(2):
            B = np.zeros((k + 1, k + 1))
            B[1:, 1:] = np.einsum("ikl,jkl->ij", errs[:, s], errs[:, s])
            B[0, 1:] = 1.0
            B[1:, 0] = 1.0
            C = np.zeros(k + 1)
            C[0] = 1.0
            w, v = np.linalg.eigh(B)
            xs.append(v @ np.diag(1.0 / w) @ v.T @ C)
        x = np.stack(xs)
        assert not np.any(np.isnan(x))
        return x[:, 1:]

    @staticmethod
    def _df(dens, focks):
        k = dens.shape[0]
        return np.array([[sum(np.trace(dens[i, s] @ focks[j, s]) for s in range(dens.shape[1]))
                          for j in range(k)] for i in range(k)])

    def ediis_minimize(self, dens, focks, energies):
        df = self._df(dens, focks)
        d = np.diag(df)
        df = d[:, None] + d[None, :] - df - df.T
        return _bfgs_weights(_weights_cost(energies, -df), len(energies))

    def adiis_minimize(self, dens, focks, idnewest):
        df = self._df(dens, focks)
        d_fn = df[:, idnewest]
        dn_f = df[idnewest]
        dn_fn = df[idnewest, idnewest]
        dd_fn = d_fn - dn_fn
        df = df - d_fn[:, None] - dn_f[None, :] + dn_fn
        return _bfgs_weights(_weights_cost(2.0 * dd_fn, df), dens.shape[0])


def smearing_occ(mo_energy, nelectron, smearing, sigma):
    """evaluate.py:467-497 with one chemical potential over both spins' orbitals, (2, n)."""
    from scipy.optimize import bisect
    from scipy.special import erfc
    es = np.hstack(mo_energy)
    if smearing == "gaussian":
        occ = lambda m, e: 0.5 * erfc((e - m) / sigma)  # noqa: E731
    else:
        def occ(m, e):
            with np.errstate(over="ignore"):
                return 1.0 / (np.exp((e - m) / sigma) + 1.0)
    mu = bisect(lambda m: occ(m, es).sum() - nelectron, es.min(), es.max(), xtol=1e-10, rtol=1e-10, maxiter=10000)
    return occ(mu, es).reshape(mo_energy.shape)


def scf_loop(xc_value_and_vxc: Callable, inp: SCFInputs, rdm1: np.ndarray, mo_occ: np.ndarray, *,
             level_shift_factor=(0.0, 0.0), damp_factor=(0.0, 0.0), cycles=50, diis_start_cycle=0,
             e_conv=1e-5, g_conv=1e-5, diis_method="CDIIS", smearing=None, smearing_sigma=0.0,
             carried_fock=None, nelectron=None, exchange_update=None):
    """evaluate.py:387-589.  Returns (SCFState, cycles run).  The orbital gradient uses the Fock
    matrix the molecule carried in (carried_fock), or the current one when there is none.
    exchange_update(D) -> K0, when given, is the chi update of evaluate.py:510-522 and 565-576:
    the exact-exchange matrix is rebuilt from every new density before it is predicted at."""
    S = inp.s1e
    n = S.shape[0]
    if nelectron is None:
        nelectron = float(mo_occ.sum())
    e, fock = predict(xc_value_and_vxc, inp, rdm1)
    old_e, norm_gorb, cycle = np.inf, np.inf, 0
    diis = Diis(S, np.eye(n), 10, diis_method)
    data = ([], [], [], [])
    D, occ, w, C = rdm1, mo_occ, None, None

    def solve(F, occ):
        w, C = safe_fock_solver(F, S)
        occ = get_occ(w, np.array([int(v) for v in occ.sum(axis=1)]))
        if smearing:
            occ = smearing_occ(w, nelectron, smearing, smearing_sigma)
        return w, C, occ, make_rdm1(C, occ)

    def gnorm(C, occ, F):
        return np.linalg.norm(orbital_grad(C, occ, carried_fock if carried_fock is not None else F))

    while (abs(e - old_e) * HARTREE2KCALMOL > e_conv or norm_gorb > g_conv) and cycle < cycles:
        old_e = e
        if 0 <= cycle < diis_start_cycle - 1 and abs(damp_factor[0]) + abs(damp_factor[1]) > 1e-4:
            fock = np.stack([damping(S, D[s], fock[s], damp_factor[s]) for s in range(2)])
        if cycle >= diis_start_cycle:
            fock, data = diis.run(D, fock, e, data, cycle)
        if abs(level_shift_factor[0]) + abs(level_shift_factor[1]) > 1e-4:
            fock = np.stack([level_shift(S, D[s], fock[s], level_shift_factor[s]) for s in range(2)])
        w, C, occ, D = solve(fock, occ)
        if exchange_update is not None:
            inp = dataclasses.replace(inp, K0=exchange_update(D))
        e, fock = predict(xc_value_and_vxc, inp, D)
        norm_gorb = gnorm(C, occ, fock)
        cycle += 1
    if abs(e - old_e) * HARTREE2KCALMOL < e_conv and norm_gorb < g_conv:
        w, C, occ, D = solve(fock, occ)
        if exchange_update is not None:
            inp = dataclasses.replace(inp, K0=exchange_update(D))
        e, fock = predict(xc_value_and_vxc, inp, D)
    return SCFState(D, fock, e, w, C, occ), cycle


def simple_scf_loop(xc_value_and_vxc: Callable, inp: SCFInputs, rdm1, mo_coeff, mo_energy, mo_occ,
                    mixing_factor=0.4, cycles=50) -> SCFState:
    """evaluate.py:176-252: cycle 0 keeps the given orbitals (the density used is the given
    rdm1), later cycles diagonalise the last Fock matrix and mix the densities."""
    S = inp.s1e
    D, C, w, occ = rdm1, mo_coeff, mo_energy, mo_occ
    energies = []
    fock = None
    for cycle in range(cycles):
        if cycle > 0:
            w, C = safe_fock_solver(fock, S)
        occ = get_occ(w, np.array([int(v) for v in occ.sum(axis=1)]))
        if cycle == 0:
            old = make_rdm1(C, occ)
        else:
            D = (1 - mixing_factor) * old + mixing_factor * make_rdm1(C, occ)
            old = D
        e, fock = predict(xc_value_and_vxc, inp, D)
        energies.append(e)
    return SCFState(D, fock, e, w, C, occ, energies)
