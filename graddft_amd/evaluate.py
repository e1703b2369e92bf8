"""Self-consistent field driver (``grad_dft/evaluate.py:917-1205``).

``diff_scf_loop(functional, cycles)(params, molecule) -> molecule`` runs a fixed number of
CDIIS iterations exactly as the reference's jitted loop: predict on the input density, then
per cycle DIIS extrapolation of the Fock matrix, generalised diagonalisation, aufbau
occupations, new density, predict.  ``JittableDiis`` keeps the reference's buffer semantics
(``evaluate.py:1111-1205``): slot ``cycle`` is written while ``cycle < max_diis``, the write at
``cycle == max_diis`` is dropped (JAX's out-of-bounds scatter), later cycles shift the
buffers; the bordered matrix has +1 borders for active slots and 1 on the diagonal of
inactive ones, and the coefficients are B^-1 e_0 (hand-written solve, ``gdx_diis_solve``).
The reference's extra diagonalisation after the loop is discarded by it and not computed.

The loop is differentiable w.r.t. params (reverse mode through every cycle, as jax.grad
through the reference's fori_loop): the predictor, the DIIS solve and the eigensolver are
autograd Functions over the device kernels when autograd is recording.
"""
from __future__ import annotations

from dataclasses import dataclass
from typing import Callable, List, Optional

import torch

from . import _lib
from .eigenproblem import cholesky_inv, orthogonal_eigh, scf_step
from .functional import Functional
from .molecule import Molecule
from .train import energy_predictor


def _diis_solve(B: torch.Tensor) -> torch.Tensor:
    B = B.detach().contiguous()
    x = torch.empty(B.shape[:2], dtype=torch.float64, device=B.device)
    _lib.check(_lib.load().gdx_diis_solve(torch.cuda.current_stream(B.device).cuda_stream, B.shape[0],
                                          B.shape[1], B.data_ptr(), x.data_ptr()), "gdx_diis_solve")
    return x


def _diis_coefficients(errs: torch.Tensor, cycle: int):
    errs = errs.detach().contiguous()
    m, ns = errs.shape[0], errs.shape[1]
    B = torch.empty((ns, m + 1, m + 1), dtype=torch.float64, device=errs.device)
    x = torch.empty((ns, m + 1), dtype=torch.float64, device=errs.device)
    _lib.check(_lib.load().gdx_diis_coefficients(torch.cuda.current_stream(errs.device).cuda_stream, ns, m,
                                                 errs[0, 0].numel(), errs.data_ptr(), int(cycle), B.data_ptr(),
                                                 x.data_ptr()), "gdx_diis_coefficients")
    return B, x


class DiisCoefficientsFn(torch.autograd.Function):
    """JittableDiis coefficients from the stored errors (m, nspin, n, n) in one device call
    (Gram matrix, bordered matrix, solve).  Reverse: Bbar = -B^-T xbar x^T as for the solve;
    the Gram entries get Bbar's inner block (diagonal only for active slots, the inactive ones
    are the constant 1), and errs get (G + G^T) E per spin."""

    @staticmethod
    def forward(ctx, errs, cycle):
        B, x = _diis_coefficients(errs, cycle)
        ctx.save_for_backward(errs, B, x)
        ctx.cycle = cycle
        return x[:, 1:]

    @staticmethod
    @torch.autograd.function.once_differentiable
    def backward(ctx, gc):
        errs, B, x = ctx.saved_tensors
        m, ns = errs.shape[0], errs.shape[1]
        gx = torch.zeros_like(x)
        gx[:, 1:] = gc
        y = torch.linalg.solve(B.transpose(1, 2), gx.unsqueeze(-1)).squeeze(-1)
        G = -(y.unsqueeze(-1) * x.unsqueeze(-2))[:, 1:, 1:]
        inactive = torch.arange(m, device=errs.device) > ctx.cycle
        G = G.clone()
        G[:, inactive, inactive] = 0.0
        E = errs.reshape(m, ns, -1).transpose(0, 1)                 # (ns, m, n2)
        Ebar = torch.matmul(G + G.transpose(1, 2), E)               # (ns, m, n2)
        return Ebar.transpose(0, 1).reshape(errs.shape), None


class DiisSolveFn(torch.autograd.Function):
    """x_s = B_s^-1 e_0; reverse: Bbar_s = -B_s^-T xbar_s x_s^T (B_s is symmetric)."""

    @staticmethod
    def forward(ctx, B):
        x = _diis_solve(B)
        ctx.save_for_backward(B, x)
        return x

    @staticmethod
    @torch.autograd.function.once_differentiable
    def backward(ctx, gx):
        B, x = ctx.saved_tensors
        y = torch.linalg.solve(B.transpose(1, 2), gx.unsqueeze(-1)).squeeze(-1)
        return -y.unsqueeze(-1) * x.unsqueeze(-2)


@dataclass
class JittableDiis:
    overlap_matrix: torch.Tensor
    max_diis: int = 10

    def zeros(self, n: int, device):
        m = self.max_diis
        z = lambda *s: torch.zeros(s, dtype=torch.float64, device=device)  # noqa: E731
        return [z(m, 2, n, n), z(m, 2, n, n), z(m), z(m, 2, n, n)]

    def update(self, rdm1, fock, energy, data, cycle: int):
        dens, focks, energies, errs = data
        fds = fock @ rdm1 @ self.overlap_matrix
        err = fds - fds.transpose(1, 2)
        if cycle > self.max_diis:
            errs = torch.cat([errs[1:], err[None]])
            dens = torch.cat([dens[1:], rdm1[None]])
            focks = torch.cat([focks[1:], fock[None]])
            energies = torch.cat([energies[1:], energy.reshape(1)])
        elif cycle < self.max_diis:
            if torch.is_grad_enabled() and any(t.requires_grad for t in (err, rdm1, fock, energy)):
                # functional update: tensors saved for the reverse pass must not change
                errs, dens, focks, energies = errs.clone(), dens.clone(), focks.clone(), energies.clone()
            errs[cycle] = err
            dens[cycle] = rdm1
            focks[cycle] = fock
            energies[cycle] = energy
        return [dens, focks, energies, errs]

    def coefficients(self, errs: torch.Tensor, cycle: int) -> torch.Tensor:
        # Gram matrix, bordered matrix (+1 borders for slots <= cycle, 1 on the diagonal of the
        # others) and x = B^-1 e_0 in one device call (gdx_diis_coefficients)
        return DiisCoefficientsFn.apply(errs, cycle)

    def run(self, rdm1, fock, energy, data, cycle: int):
        data = self.update(rdm1, fock, energy, data, cycle)
        x = self.coefficients(data[3], cycle)
        F = torch.einsum("si,isjk->sjk", x, data[1])
        return F, data


def non_scf_predictor(functional: Functional, precision: str = "fp64", **kwargs) -> Callable:
    """``non_scf_predictor`` (evaluate.py:88-128): energy and Fock at the molecule's density."""
    compute_energy = energy_predictor(functional, precision=precision, **kwargs)

    def predictor(params: Optional[torch.Tensor], molecule: Molecule) -> Molecule:
        e, fock = compute_energy(params, molecule)
        return molecule.replace(fock=fock, energy=e)

    return predictor


def diff_simple_scf_loop(functional: Functional, cycles: int = 25, mixing_factor: float = 0.4,
                         precision: str = "fp64", **kwargs) -> Callable:
    """``diff_simple_scf_loop`` (evaluate.py:257-352): predict, then per cycle diagonalise the
    current Fock matrix (safe_fock_solver), aufbau occupations, linear density mixing
    D = (1 - mixing_factor) D_old + mixing_factor D_new, predict.  Differentiable w.r.t. params
    like diff_scf_loop (the same predictor and eigensolver Functions); the eigensolver is
    warm-started from the previous cycle's eigenvectors."""
    compute_energy = energy_predictor(functional, precision=precision, **kwargs)

    def iterator(params: Optional[torch.Tensor], molecule: Molecule) -> Molecule:
        Linv = molecule.cache.get("Linv")
        if Linv is None:
            _, Linv = cholesky_inv(molecule.s1e)
            molecule.cache["Linv"] = Linv
        predicted_e, fock = compute_energy(params, molecule)
        molecule = molecule.replace(fock=fock, energy=predicted_e)
        occ = _aufbau_occupations(molecule)
        V = None
        for _ in range(cycles):
            mo_energy, mo_coeff, V = orthogonal_eigh(molecule.fock, Linv, V)
            new_rdm1 = torch.einsum("sij,sj,skj->sik", mo_coeff, occ, mo_coeff)
            rdm1 = (1.0 - mixing_factor) * molecule.rdm1 + mixing_factor * new_rdm1
            molecule = molecule.replace(mo_coeff=mo_coeff, mo_energy=mo_energy, mo_occ=occ, rdm1=rdm1,
                                        restricted=molecule.restricted and functional.spin_symmetric)
            predicted_e, fock = compute_energy(params, molecule)
            molecule = molecule.replace(fock=fock)
        return molecule.replace(energy=predicted_e)

    return iterator


def _aufbau_occupations(molecule: Molecule) -> torch.Tensor:
    """get_occ for ascending eigenvalues: the first N_e,s orbitals of each spin (cached)."""
    occ = molecule.cache.get("aufbau_occ")
    if occ is None:
        nelecs = [int(round(float(v))) for v in molecule.mo_occ.sum(dim=1).tolist()]
        occ = torch.zeros((2, molecule.nao), dtype=torch.float64, device=molecule.s1e.device)
        for s in range(2):
            occ[s, : nelecs[s]] = 1.0
        molecule.cache["aufbau_occ"] = occ
    return occ


def diff_scf_loop(functional: Functional, cycles: int = 25, precision: str = "fp64", graph: bool = False,
                  refine_cycles: Optional[int] = None, **kwargs) -> Callable:
    """graph=True (forward only, autograd not recording): the whole loop is captured once per
    molecule as one HIP graph (after an eager run that sets up plans and workspaces) and
    replayed on later calls -- no host work between the kernels, as the reference's jitted
    fori_loop.  The rdm1 and params of each call are copied into the captured inputs.

    refine_cycles (mixed precisions; default 1, 0 for fp64): the last refine_cycles cycles,
    and the energy they return, use the FP64 grid path.  The returned energy is then an FP64
    evaluation, and a loop stopped before self-consistency -- first-order sensitive to the
    Fock matrices of its last cycles -- has the FP32 perturbations of the earlier cycles damped
    by the iteration (naphthalene def2-SVP, 20 cycles: |dE| vs fp64 1e-6 Ha -> 2e-8 Ha)."""
    if refine_cycles is None:
        refine_cycles = 0 if precision == "fp64" else 1
    compute_energy = energy_predictor(functional, precision=precision, **kwargs)
    compute_energy64 = energy_predictor(functional, precision="fp64", **kwargs)

    def run_loop(params: Optional[torch.Tensor], molecule: Molecule) -> Molecule:
        S = molecule.s1e
        Linv = molecule.cache.get("Linv")
        if Linv is None:
            _, Linv = cholesky_inv(S)
            molecule.cache["Linv"] = Linv
        predicted_e, fock = compute_energy(params, molecule)
        molecule = molecule.replace(fock=fock)
        diis = JittableDiis(S, max_diis=10)
        data = diis.zeros(molecule.nao, S.device)
        occ = _aufbau_occupations(molecule)
        V = None
        # nothing recorded for reverse mode: extrapolation, eigenproblem and density in one call
        fused = not (torch.is_grad_enabled() and (
            (params is not None and params.requires_grad) or molecule.rdm1.requires_grad))
        for cycle in range(cycles):
            predictor = compute_energy64 if cycle >= cycles - refine_cycles else compute_energy
            if fused:
                data = diis.update(molecule.rdm1, molecule.fock, predicted_e, data, cycle)
                _, x = _diis_coefficients(data[3], cycle)
                F, mo_energy, mo_coeff, V, rdm1 = scf_step(data[1], x, Linv, V, occ)
            else:
                F, data = diis.run(molecule.rdm1, molecule.fock, predicted_e, data, cycle)
                # safe_fock_solver, warm-started from the previous cycle's orthonormal eigenvectors
                mo_energy, mo_coeff, V = orthogonal_eigh(F, Linv, V)
                # eigenvalues ascend, so the aufbau occupation of get_occ is the first nelec_s
                rdm1 = torch.einsum("sij,sj,skj->sik", mo_coeff, occ, mo_coeff)
            molecule = molecule.replace(fock=F, mo_coeff=mo_coeff, mo_energy=mo_energy, mo_occ=occ,
                                        rdm1=rdm1, restricted=molecule.restricted and functional.spin_symmetric)
            predicted_e, fock = predictor(params, molecule)
            molecule = molecule.replace(fock=fock)
        return molecule.replace(energy=predicted_e)

    def scf_iterator(params: Optional[torch.Tensor], molecule: Molecule) -> Molecule:
        recording = torch.is_grad_enabled() and (
            (params is not None and params.requires_grad) or molecule.rdm1.requires_grad)
        if not graph or recording:
            return run_loop(params, molecule)
        key = ("scf_graph", id(functional), cycles, precision, refine_cycles, params is None)
        entry = molecule.cache.get(key)
        if entry is None:
            out = run_loop(params, molecule)          # plans, workspaces, Cholesky, K0
            static_p = None if params is None else params.detach().clone()
            static_D = molecule.rdm1.detach().clone()
            m_in = molecule.replace(rdm1=static_D)
            side = torch.cuda.Stream()
            side.wait_stream(torch.cuda.current_stream())
            with torch.cuda.stream(side):
                run_loop(static_p, m_in)
            torch.cuda.current_stream().wait_stream(side)
            g = torch.cuda.CUDAGraph()
            with torch.cuda.graph(g):
                m_out = run_loop(static_p, m_in)
            molecule.cache[key] = (g, static_p, static_D, m_out)
            return out
        g, static_p, static_D, m_out = entry
        if static_p is not None:
            static_p.copy_(params.detach())
        static_D.copy_(molecule.rdm1.detach())
        g.replay()
        return molecule.replace(energy=m_out.energy.clone(), rdm1=m_out.rdm1.clone(), fock=m_out.fock.clone(),
                                mo_coeff=m_out.mo_coeff.clone(), mo_energy=m_out.mo_energy.clone(),
                                mo_occ=m_out.mo_occ.clone(), restricted=m_out.restricted)

    return scf_iterator
