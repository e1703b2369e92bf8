"""Oracle for reverse mode through E_xc and through the SCF loop (torch-CPU float64) --
TEST INFRASTRUCTURE ONLY.

Restates:

* ``xc_hvp`` -- the second-order quantity jax.grad forms when it differentiates the Fock
  matrix's XC term (``grad_dft/train.py:86-121``) inside ``diff_scf_loop``: with
  E(theta, D) = E_xc, Phi = g E + <dD, dE/dD>, and dPhi/dD, dPhi/dtheta, by double backward
  through ``oracle.xc``;
* ``diff_simple_scf_energy`` -- the linear-mixing loop of ``evaluate.py:257-352``, same pieces
  without DIIS;
* ``diff_scf_energy`` -- the differentiable loop of ``grad_dft/evaluate.py:917-1038`` in
  torch ops: ``predict`` (``train.py:123-216``; V_xc by autograd with create_graph, the
  stale-exchange B3LYP term as in ``oracle.scf.predict``), ``JittableDiis``
  (``evaluate.py:1111-1205``), ``safe_general_eigh`` (``utils/eigenproblem.py:110-129``) with
  the Lorentzian-broadened reverse mode of ``safe_eigh`` (``:26-103``), ``get_occ`` and
  ``make_rdm1`` (``molecule.py:816-889``).

Pinning (``tests/test_oracle_scf_grad.py``): central finite differences of the oracle's own
values -- E_xc along dD for edot; dE/dD and dE/dtheta along dD for the HVP; the final SCF
energy along a parameter direction for the loop gradient -- and the torch loop's energies
equal ``oracle.scf.diff_scf_loop``'s (the numpy restatement pinned in
``tests/test_oracle_scf.py``); the mixing loop reaches the DIIS loop's self-consistent energy
(the common fixed point) and its parameter gradient matches central differences.
"""
from __future__ import annotations

from dataclasses import dataclass
from typing import Callable, Optional

import numpy as np
import torch

from . import xc as X

DT = torch.float64
CLIP = 1e-30
DEGEN_TOL = 1e-6    # utils/eigenproblem.py:23
BROADENING = 1e-10  # utils/eigenproblem.py:24


def abs_clip(x):
    return torch.where(x.abs() > CLIP, x, torch.zeros_like(x))


def xc_hvp(kind: str, inp: X.XCInputs, rdm1: np.ndarray, params: Optional[np.ndarray], spec,
           direction: np.ndarray, g: float):
    """(E, edot = <dD, dE/dD>, dPhi/dD, dPhi/dtheta) for Phi = g E + edot."""
    D = torch.tensor(rdm1, dtype=DT, requires_grad=True)
    th = None if params is None else torch.tensor(params, dtype=DT, requires_grad=True)
    E = X.xc_energy_from_rdm1(kind, inp, D, th, spec)
    gD, = torch.autograd.grad(E, [D], create_graph=True)
    edot = (gD * torch.tensor(direction, dtype=DT)).sum()
    phi = g * E + edot
    wrt = [D] + ([th] if th is not None else [])
    grads = torch.autograd.grad(phi, wrt, allow_unused=True)
    dD = grads[0].numpy()
    dth = None
    if th is not None:
        dth = grads[1].numpy() if grads[1] is not None else np.zeros_like(params)
    return float(E.detach()), float(edot.detach()), dD, dth


class SafeEigh(torch.autograd.Function):
    """eigh with the reverse mode of grad_dft/utils/eigenproblem.py:54-103."""

    @staticmethod
    def forward(ctx, A):
        w, V = torch.linalg.eigh(A)
        ctx.save_for_backward(w, V)
        return w, V

    @staticmethod
    def backward(ctx, gw, gV):
        w, V = ctx.saved_tensors
        if gw is None:
            gw = torch.zeros_like(w)
        if gV is None:
            gV = torch.zeros_like(V)
        diff = w[None, :] - w[:, None]
        degen = diff.abs() < DEGEN_TOL
        regular = torch.nan_to_num(1.0 / diff)
        broadened = diff / (diff * diff + BROADENING)
        F = 0.5 * torch.where(degen, broadened, regular)
        F = F - torch.diag(torch.diag(F))
        Vt = V.T
        grad = torch.linalg.inv(Vt) @ (0.5 * torch.diag(gw) + F * (Vt @ gV)) @ Vt
        return grad + grad.T


def safe_general_eigh(A, L_inv):
    C = L_inv @ A @ L_inv.T
    w, V = SafeEigh.apply(C)
    return w, L_inv.T @ V


@dataclass
class TorchSCFInputs:
    h1e: torch.Tensor
    s1e: torch.Tensor
    eri: torch.Tensor          # dense (n, n, n, n)
    e_nuc: float
    K0: Optional[torch.Tensor]
    a0: float = 0.2
    # exact exchange rebuilt from the density being predicted at: K_s = sum_bd (ab|cd) D_s[b][d]
    # on the tape (the device's "current" exchange); False keeps K0
    current_exchange: bool = False

    @staticmethod
    def from_numpy(inp, current_exchange: bool = False) -> "TorchSCFInputs":
        t = lambda a: None if a is None else torch.tensor(a, dtype=DT)  # noqa: E731
        return TorchSCFInputs(t(inp.h1e), t(inp.s1e), t(inp.eri), float(inp.e_nuc), t(inp.K0), inp.a0,
                              current_exchange)


def predict(xc_energy: Callable, inp: TorchSCFInputs, rdm1: torch.Tensor):
    """train.py:123-216 with V_xc = dE_xc/dD kept on the autograd tape."""
    if not rdm1.requires_grad:
        rdm1 = rdm1.detach().requires_grad_(True)
        exc = xc_energy(rdm1)
        vxc, = torch.autograd.grad(exc, [rdm1], create_graph=True)
    else:
        exc = xc_energy(rdm1)
        vxc, = torch.autograd.grad(exc, [rdm1], create_graph=True)
    Dt = rdm1.sum(0)
    J = torch.einsum("pqrt,rt->pq", inp.eri, Dt)
    energy = exc + inp.e_nuc + (Dt * inp.h1e).sum() + (Dt * J).sum() / 2.0
    fock = inp.h1e[None] + J[None] + vxc
    fock = abs_clip(fock)
    fock = 0.5 * (fock + fock.transpose(1, 2))
    fock = abs_clip(fock)
    if inp.K0 is not None:
        K = torch.einsum("abcd,sbd->sac", inp.eri, rdm1) if inp.current_exchange else inp.K0
        energy = energy - 0.5 * inp.a0 * (rdm1 * K).sum()
        v = -0.5 * inp.a0 * K
        fock = abs_clip(fock + v + v.transpose(1, 2))
    fock = abs_clip(fock)
    return energy, fock


class JittableDiis:
    """evaluate.py:1111-1205 (A = I), torch ops."""

    def __init__(self, s1e, max_diis=10):
        self.s1e, self.max_diis = s1e, max_diis

    def run(self, rdm1, fock, data, cycle):
        focks, errs = data
        fds = fock @ rdm1 @ self.s1e
        err = fds - fds.transpose(1, 2)
        m = self.max_diis
        if cycle > m:
            errs = torch.cat([errs, err[None]])[1:]
            focks = torch.cat([focks, fock[None]])[1:]
        elif cycle < m:
            errs = torch.cat([errs[:cycle], err[None], errs[cycle + 1:]])
            focks = torch.cat([focks[:cycle], fock[None], focks[cycle + 1:]])
        gram = torch.einsum("iskl,jskl->sij", errs, errs)
        rows = []
        for s in range(2):
            B = torch.zeros((m + 1, m + 1), dtype=DT)
            act = torch.tensor([1.0 if i <= cycle else 0.0 for i in range(m)], dtype=DT)
            B = B.index_put((torch.zeros(m, dtype=torch.long), torch.arange(1, m + 1)), act)
            B = B.index_put((torch.arange(1, m + 1), torch.zeros(m, dtype=torch.long)), act)
            inner = torch.where(torch.eye(m, dtype=torch.bool),
                                torch.where(act[:, None] > 0, gram[s], torch.ones_like(gram[s])), gram[s])
            B = torch.cat([B[:1], torch.cat([B[1:, :1], inner], dim=1)], dim=0)
            e0 = torch.zeros(m + 1, dtype=DT)
            e0[0] = 1.0
            rows.append(torch.linalg.inv(B) @ e0)
        x = torch.stack(rows)[:, 1:]
        F = torch.einsum("si,isjk->sjk", x, focks)
        return F, (focks, errs)


def diff_scf_energy(xc_energy: Callable, inp: TorchSCFInputs, rdm1: torch.Tensor, mo_occ: torch.Tensor,
                    cycles: int) -> torch.Tensor:
    """evaluate.py:938-1038: the final predicted energy, on the autograd tape."""
    n = rdm1.shape[-1]
    L = torch.linalg.cholesky(inp.s1e)
    L_inv = torch.linalg.inv(L)
    e, fock = predict(xc_energy, inp, rdm1)
    diis = JittableDiis(inp.s1e, 10)
    data = (torch.zeros((10, 2, n, n), dtype=DT), torch.zeros((10, 2, n, n), dtype=DT))
    nelecs = [int(round(float(v))) for v in mo_occ.sum(dim=1)]
    D = rdm1
    for cycle in range(cycles):
        F, data = diis.run(D, fock, data, cycle)
        outs = [safe_general_eigh(F[s], L_inv) for s in range(2)]
        occ = torch.zeros((2, n), dtype=DT)
        for s in range(2):
            idx = torch.argsort(outs[s][0].detach(), stable=True)
            occ[s, idx[: nelecs[s]]] = 1.0
        C = torch.stack([o[1] for o in outs])
        D = torch.einsum("sij,sj,skj->sik", C, occ, C)
        e, fock = predict(xc_energy, inp, D)
    return e


def diff_simple_scf_energy(xc_energy: Callable, inp: TorchSCFInputs, rdm1: torch.Tensor, mo_occ: torch.Tensor,
                           cycles: int, mixing_factor: float = 0.4) -> torch.Tensor:
    """evaluate.py:257-352 (diff_simple_scf_loop): predict, then per cycle diagonalise the
    current Fock matrix (safe_fock_solver), aufbau occupations, linear density mixing
    D = (1 - m) D_old + m D_new, predict; the final predicted energy, on the autograd tape."""
    n = rdm1.shape[-1]
    L = torch.linalg.cholesky(inp.s1e)
    L_inv = torch.linalg.inv(L)
    e, fock = predict(xc_energy, inp, rdm1)
    nelecs = [int(round(float(v))) for v in mo_occ.sum(dim=1)]
    D = rdm1
    for _ in range(cycles):
        outs = [safe_general_eigh(fock[s], L_inv) for s in range(2)]
        occ = torch.zeros((2, n), dtype=DT)
        for s in range(2):
            idx = torch.argsort(outs[s][0].detach(), stable=True)
            occ[s, idx[: nelecs[s]]] = 1.0
        C = torch.stack([o[1] for o in outs])
        Dnew = torch.einsum("sij,sj,skj->sik", C, occ, C)
        D = (1.0 - mixing_factor) * D + mixing_factor * Dnew
        e, fock = predict(xc_energy, inp, D)
    return e


# ---------------------------------------------------------------------------------------
# Orbital optimizers (grad_dft/evaluate.py:595-914): the energy as a function of a matrix W
# whose QR factor Q gives orthonormal orbitals C = (Q D)^T, minimised by Adam.
# ---------------------------------------------------------------------------------------

def whitening_matrix(S: np.ndarray, whitening: str = "PCA") -> np.ndarray:
    """evaluate.py:729-740: D with D^T D = S^-1.  PCA diag(w^-1/2) V^T, Cholesky chol(S^-1)^T,
    ZCA V diag(w^-1/2) V^T; eigenvector signs fixed so each eigenvector's largest-magnitude
    component is positive."""
    if whitening == "Cholesky":
        return np.linalg.cholesky(np.linalg.inv(S)).T
    w, V = np.linalg.eigh(S)
    idx = np.argmax(np.abs(V), axis=0)
    V = V * np.sign(V[idx, np.arange(V.shape[1])])[None, :]
    if whitening == "PCA":
        return np.diag(w ** -0.5) @ V.T
    if whitening == "ZCA":
        return V @ np.diag(w ** -0.5) @ V.T
    raise ValueError(whitening)


def energy_only(xc_energy: Callable, inp: TorchSCFInputs, rdm1: torch.Tensor) -> torch.Tensor:
    """train.py:123-216's energy (no Fock matrix)."""
    Dt = rdm1.sum(0)
    J = torch.einsum("pqrt,rt->pq", inp.eri, Dt)
    e = xc_energy(rdm1) + inp.e_nuc + (Dt * inp.h1e).sum() + (Dt * J).sum() / 2.0
    if inp.K0 is not None:
        e = e - 0.5 * inp.a0 * (rdm1 * inp.K0).sum()
    return e


def orbital_energy(xc_energy: Callable, inp: TorchSCFInputs, W: torch.Tensor, D: torch.Tensor,
                   mo_occ: torch.Tensor) -> torch.Tensor:
    """molecule_orbitals_iterator (evaluate.py:642-699): Q = qr(W_s), C_s = (Q_s D)^T,
    rdm1 = C diag(occ) C^T, predicted energy."""
    Q = torch.stack([torch.linalg.qr(W[s])[0] for s in range(2)])
    C = torch.einsum("sij,jk->ski", Q, D)
    rdm1 = torch.einsum("sij,sj,skj->sik", C, mo_occ, C)
    return energy_only(xc_energy, inp, rdm1)


def mol_orb_optimizer(xc_energy: Callable, inp: TorchSCFInputs, mo_coeff: torch.Tensor, mo_occ: torch.Tensor,
                      lr: float, cycles: int, e_conv: Optional[float] = 1e-7, whitening: str = "PCA",
                      b1: float = 0.9, b2: float = 0.999, eps: float = 1e-8):
    """neural_iterator of evaluate.py:701-775 (e_conv given) or of jitted_mol_orb_optimizer
    (evaluate.py:871-912, e_conv None: a fixed number of cycles), with optax.adam(lr).
    Returns the list of predicted energies, one per cycle (the last is the returned one)."""
    D = torch.tensor(whitening_matrix(inp.s1e.numpy(), whitening), dtype=DT)
    W = torch.einsum("sji,jk->sik", mo_coeff, torch.linalg.inv(D))
    mu = torch.zeros_like(W)
    nu = torch.zeros_like(W)
    out = []
    old_e = float("inf")
    e = float(energy_only(xc_energy, inp, torch.einsum("sij,sj,skj->sik", mo_coeff, mo_occ, mo_coeff)))
    for t in range(1, cycles + 1):
        if e_conv is not None and not abs(e - old_e) * 627.50947 > e_conv:
            break
        old_e = e
        Wg = W.detach().requires_grad_(True)
        E = orbital_energy(xc_energy, inp, Wg, D, mo_occ)
        g, = torch.autograd.grad(E, [Wg])
        e = float(E.detach())
        out.append(e)
        mu = b1 * mu + (1 - b1) * g
        nu = b2 * nu + (1 - b2) * g * g
        W = W - lr * (mu / (1 - b1 ** t)) / (torch.sqrt(nu / (1 - b2 ** t)) + eps)
    return out
