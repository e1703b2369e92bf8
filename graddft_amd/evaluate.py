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

import ctypes
import gc
from dataclasses import dataclass
from typing import Callable, List, Optional

import torch

from . import _lib
from .eigenproblem import cholesky_inv, eigh, householder_q, orthogonal_eigh, raise_on_status, scf_status, scf_step
from .functional import Functional
from .molecule import Molecule, make_rdm1, orbital_grad
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
        # solve_ex: no error check, so no host synchronisation (the reverse pass is capturable)
        y = torch.linalg.solve_ex(B.transpose(1, 2), gx.unsqueeze(-1))[0].squeeze(-1)
        G = -(y.unsqueeze(-1) * x.unsqueeze(-2))[:, 1:, 1:]
        G = G.clone()
        G.diagonal(dim1=1, dim2=2)[:, ctx.cycle + 1:] = 0.0     # inactive slots: constant 1
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
        y = torch.linalg.solve_ex(B.transpose(1, 2), gx.unsqueeze(-1))[0].squeeze(-1)
        return -y.unsqueeze(-1) * x.unsqueeze(-2)


class _GraphedLoopFn(torch.autograd.Function):
    """A captured diff_scf_loop: forward replays the loop's graph (its inputs already copied in)
    and returns its energy; backward writes the energy's cotangent into the reverse graph's
    input and replays it."""

    @staticmethod
    def forward(ctx, params, g_fwd, g_bwd, static_g, energy_out, d_params_out):
        g_fwd.replay()
        ctx.g_bwd, ctx.static_g, ctx.d_params_out = g_bwd, static_g, d_params_out
        return energy_out.detach().clone()

    @staticmethod
    @torch.autograd.function.once_differentiable
    def backward(ctx, g):
        ctx.static_g.copy_(g.detach().reshape(()))
        ctx.g_bwd.replay()
        return ctx.d_params_out.clone(), None, None, None, None, None


def capture_graph(g: "torch.cuda.CUDAGraph", fn: Callable, pool=None):
    """Capture fn() into g (in the memory pool of another graph when pool is given) with
    Python's cyclic garbage collector run first and paused during the capture.  Molecules and their XC plans reference each other, so a dead molecule --
    with the graphs of its captured loops in its cache -- is freed only by the cyclic
    collector; a CUDAGraph destroyed while another capture is in progress makes calls a
    capturing stream does not permit, and its destructor then aborts the process."""
    gc.collect()
    enabled = gc.isenabled()
    gc.disable()
    try:
        with torch.cuda.graph(g, pool=pool):
            return fn()
    finally:
        if enabled:
            gc.enable()


def _diis_error_device(F, D, S):
    ns, n = F.shape[0], F.shape[-1]
    nbytes = ctypes.c_size_t(0)
    _lib.check(_lib.load().gdx_diis_error_workspace_size(ns, n, ctypes.byref(nbytes)),
               "gdx_diis_error_workspace_size")
    # per call (the caching allocator; a captured graph gets its own): loops of different
    # molecules may run concurrently on several streams
    ws = torch.empty(int(nbytes.value) // 8, dtype=torch.float64, device=F.device)
    err = torch.empty_like(F)
    _lib.check(_lib.load().gdx_diis_error(torch.cuda.current_stream(F.device).cuda_stream, ns, n, F.data_ptr(),
                                          D.data_ptr(), S.data_ptr(), err.data_ptr(), ws.data_ptr()),
               "gdx_diis_error")
    return err


class DiisErrorFn(torch.autograd.Function):
    """err = F D S - (F D S)^T per spin (JittableDiis.update, evaluate.py:1111-1123) in one device
    call (gdx_diis_error).  Reverse, with H = errbar - errbar^T: Fbar = H (D S)^T,
    Dbar = F^T H S^T."""

    @staticmethod
    def forward(ctx, F, D, S):
        F, D, S = F.contiguous(), D.contiguous(), S.contiguous()
        ctx.save_for_backward(F, D, S)
        return _diis_error_device(F, D, S)

    @staticmethod
    @torch.autograd.function.once_differentiable
    def backward(ctx, gE):
        F, D, S = ctx.saved_tensors
        H = gE - gE.transpose(-1, -2)
        gF = gD = None
        if ctx.needs_input_grad[0]:
            gF = H @ (D @ S).transpose(-1, -2)
        if ctx.needs_input_grad[1]:
            gD = F.transpose(-1, -2) @ H @ S.transpose(-1, -2)
        return gF, gD, None


class DiisExtrapolateFn(torch.autograd.Function):
    """F_s = sum_i x_si F_is (JittableDiis.run's extrapolation, evaluate.py:1198-1205).  Reverse:
    Fbar_is = x_si Fbar_s, and xbar_si = <Fbar_s, F_is> as an elementwise product reduced over
    the n^2 entries (written as a product with an n^2-long inner dimension and m x nspin
    outputs, the library GEMM runs it in a handful of workgroups)."""

    @staticmethod
    def forward(ctx, x, focks):
        ctx.save_for_backward(x, focks)
        return torch.einsum("si,isjk->sjk", x, focks)

    @staticmethod
    @torch.autograd.function.once_differentiable
    def backward(ctx, gF):
        x, focks = ctx.saved_tensors
        gx = gfocks = None
        if ctx.needs_input_grad[0]:
            gx = (focks * gF.unsqueeze(0)).sum(dim=(2, 3)).transpose(0, 1)
        if ctx.needs_input_grad[1]:
            gfocks = x.transpose(0, 1)[:, :, None, None] * gF.unsqueeze(0)
        return gx, gfocks


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
        err = DiisErrorFn.apply(fock, rdm1, self.overlap_matrix)
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
        F = DiisExtrapolateFn.apply(x, data[1])
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
        # get_occ (molecule.py:851-889) casts the occupation sums to integers (truncation)
        nelecs = [int(float(v)) for v in molecule.mo_occ.sum(dim=1).tolist()]
        occ = torch.zeros((2, molecule.nao), dtype=torch.float64, device=molecule.s1e.device)
        for s in range(2):
            occ[s, : nelecs[s]] = 1.0
        molecule.cache["aufbau_occ"] = occ
    return occ


def _status_buffers(molecule: Molecule):
    """(status word, Jacobi sweeps per spin) device int32 buffers of a molecule's loop; kept in
    the molecule cache so a captured loop writes the same buffers on every replay."""
    buf = molecule.cache.get("scf_status")
    if buf is None:
        dev = molecule.s1e.device
        buf = (torch.zeros(1, dtype=torch.int32, device=dev), torch.zeros(2, dtype=torch.int32, device=dev))
        molecule.cache["scf_status"] = buf
    return buf


def diff_scf_loop(functional: Functional, cycles: int = 25, precision: str = "fp64", graph: bool = False,
                  refine_cycles: Optional[int] = None, check_status: bool = True, **kwargs) -> Callable:
    """graph=True: the whole loop is captured once per molecule as one HIP graph (after an eager
    run that sets up plans and workspaces) and replayed on later calls -- no host work between
    the kernels, as the reference's jitted fori_loop.  The rdm1 and params of each call are
    copied into the captured inputs.  When autograd records (params requiring grad), the reverse
    pass through the loop is captured as a second graph in the same memory pool: the call
    returns an energy whose backward replays it (d energy / d params; one backward per call,
    the forward replay overwrites the saved intermediates).  rdm1 requiring grad takes the eager
    path.

    refine_cycles (mixed precisions; default 1, 0 for fp64): the last refine_cycles cycles,
    and the energy they return, use the FP64 grid path.  The returned energy is then an FP64
    evaluation, and a loop stopped before self-consistency -- first-order sensitive to the
    Fock matrices of its last cycles -- has the FP32 perturbations of the earlier cycles damped
    by the iteration (naphthalene def2-SVP, 20 cycles: |dE| vs fp64 1e-6 Ha -> 2e-8 Ha).

    Every cycle ORs its checks into a device status word (gdx_scf_status: eigensolver
    convergence, a finite energy, tr(D S) equal to the occupation count); the word is read once
    after the loop (or graph replay) and, with check_status, a set flag raises GdxError."""
    if graph and getattr(functional, "custom", False):
        raise NotImplementedError("graph=True: a custom coefficient function runs the eager loop")
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
        status, sweeps = _status_buffers(molecule)
        status.zero_()
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
                F, mo_energy, mo_coeff, V, rdm1 = scf_step(data[1], x, Linv, V, occ, sweeps)
            else:
                F, data = diis.run(molecule.rdm1, molecule.fock, predicted_e, data, cycle)
                # safe_fock_solver, warm-started from the previous cycle's orthonormal eigenvectors
                mo_energy, mo_coeff, V = orthogonal_eigh(F, Linv, V, sweeps)
                # eigenvalues ascend, so the aufbau occupation of get_occ is the first nelec_s
                rdm1 = torch.einsum("sij,sj,skj->sik", mo_coeff, occ, mo_coeff)
            molecule = molecule.replace(fock=F, mo_coeff=mo_coeff, mo_energy=mo_energy, mo_occ=occ,
                                        rdm1=rdm1, restricted=molecule.restricted and functional.spin_symmetric)
            predicted_e, fock = predictor(params, molecule)
            molecule = molecule.replace(fock=fock)
            scf_status(molecule.rdm1, S, occ, sweeps, predicted_e, status)
        return molecule.replace(energy=predicted_e)

    def checked(out: Molecule) -> Molecule:
        if check_status:
            raise_on_status(_status_buffers(out)[0], "diff_scf_loop")
        return out

    def graphed_reverse(params: torch.Tensor, molecule: Molecule) -> Molecule:
        key = ("scf_graph_ad", id(functional), cycles, precision, refine_cycles)
        entry = molecule.cache.get(key)
        if entry is None:
            out = run_loop(params, molecule)          # plans, workspaces, Cholesky, K0
            static_p = params.detach().clone().requires_grad_(True)
            static_D = molecule.rdm1.detach().clone()
            static_g = torch.ones((), dtype=torch.float64, device=static_D.device)
            m_in = molecule.replace(rdm1=static_D)
            side = torch.cuda.Stream()
            side.wait_stream(torch.cuda.current_stream())
            with torch.cuda.stream(side):            # warm-up of both passes off the capture
                w = run_loop(static_p, m_in)
                torch.autograd.grad(w.energy, [static_p], grad_outputs=static_g)
            torch.cuda.current_stream().wait_stream(side)
            g_fwd, g_bwd = torch.cuda.CUDAGraph(), torch.cuda.CUDAGraph()
            m_out = capture_graph(g_fwd, lambda: run_loop(static_p, m_in))
            d_p, = capture_graph(g_bwd, lambda: torch.autograd.grad(m_out.energy, [static_p], grad_outputs=static_g,
                                                                    retain_graph=True), pool=g_fwd.pool())
            molecule.cache[key] = (g_fwd, g_bwd, static_p, static_D, static_g, m_out, d_p)
            return checked(out)
        g_fwd, g_bwd, static_p, static_D, static_g, m_out, d_p = entry
        with torch.no_grad():
            static_p.copy_(params.detach())
            static_D.copy_(molecule.rdm1.detach())
        energy = _GraphedLoopFn.apply(params, g_fwd, g_bwd, static_g, m_out.energy, d_p)
        return checked(molecule.replace(energy=energy, rdm1=m_out.rdm1.detach().clone(),
                                        fock=m_out.fock.detach().clone(), mo_coeff=m_out.mo_coeff.detach().clone(),
                                        mo_energy=m_out.mo_energy.detach().clone(), mo_occ=m_out.mo_occ.clone(),
                                        restricted=m_out.restricted))

    def scf_iterator(params: Optional[torch.Tensor], molecule: Molecule) -> Molecule:
        recording = torch.is_grad_enabled() and (
            (params is not None and params.requires_grad) or molecule.rdm1.requires_grad)
        if graph and recording and params is not None and params.requires_grad and not molecule.rdm1.requires_grad:
            return graphed_reverse(params, molecule)
        if not graph or recording:
            return checked(run_loop(params, molecule))
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
            m_out = capture_graph(g, lambda: run_loop(static_p, m_in))
            molecule.cache[key] = (g, static_p, static_D, m_out)
            return checked(out)
        g, static_p, static_D, m_out = entry
        if static_p is not None:
            static_p.copy_(params.detach())
        static_D.copy_(molecule.rdm1.detach())
        g.replay()
        return checked(molecule.replace(energy=m_out.energy.clone(), rdm1=m_out.rdm1.clone(), fock=m_out.fock.clone(),
                                        mo_coeff=m_out.mo_coeff.clone(), mo_energy=m_out.mo_energy.clone(),
                                        mo_occ=m_out.mo_occ.clone(), restricted=m_out.restricted))

    return scf_iterator


# ---------------------------------------------------------------------------------------
# Non-differentiable drivers: scf_loop (CDIIS / EDIIS / ADIIS, damping, level shift,
# smearing) and simple_scf_loop (linear mixing), on the same device predictor and
# eigensolver.  Their control flow is host-side, as in the reference (Python while loop,
# scipy's bisect, the BFGS minimisation of EDIIS / ADIIS over at most max_diis weights).
# ---------------------------------------------------------------------------------------

HARTREE2KCALMOL = 627.50947   # utils/types.py:35


def damping(s: torch.Tensor, d: torch.Tensor, f: torch.Tensor, factor: float) -> torch.Tensor:
    """Fock damping (``evaluate.py:1477-1503``, pyscf's ``hf.damping``):
    f - (f0 + f0^T) factor / (factor + 1) with f0 = (I - S D) F D S.  In an S-orthonormal
    orbital basis that diagonalises D it scales the occupied-virtual blocks by 1/(1+factor)."""
    eye = torch.eye(s.shape[-1], dtype=s.dtype, device=s.device)
    f0 = (eye - s @ d) @ f @ d @ s
    return f - (f0 + f0.transpose(-1, -2)) * (factor / (factor + 1.0))


def level_shift(s: torch.Tensor, d: torch.Tensor, f: torch.Tensor, factor: float) -> torch.Tensor:
    """Level shift of the virtual space (``evaluate.py:1506-1545``): F + factor (S - S D S)."""
    return f + (s - s @ d @ s) * factor


def _simplex_minimize(lin, quad):
    """Minimise c.lin + c^T quad c over c = x^2 / sum(x^2) by BFGS from x = 1 with tol 1e-9
    (the reference's ``jax.scipy.optimize.minimize(costf, ones, method="BFGS", tol=1e-9)``,
    here scipy's BFGS with the analytic gradient).  Returns (c, cost)."""
    import numpy as np
    from scipy.optimize import minimize

    sym = quad + quad.T

    def cost(x):
        s = x @ x
        c = x * x / s
        gc = lin + sym @ c
        return c @ lin + c @ quad @ c, 2.0 * x / s * (gc - gc @ c)

    res = minimize(cost, np.ones(len(lin)), jac=True, method="BFGS", tol=1e-9)
    return res.x ** 2 / (res.x ** 2).sum(), float(res.fun)


@dataclass
class Diis:
    """DIIS extrapolation with the reference's variants (``evaluate.py:1211-1474``).

    History tensors grow by one entry per ``run`` to ``max_diis`` and then drop the oldest.
    CDIIS: per spin, errors e_i = A (F_i D_i S - S D_i F_i) A^T, the bordered Gram matrix
    (+1 borders, 0 corner) solved for x on the device (``gdx_diis_solve``), and
    F = A^T (sum_i x_i F_i) A.  EDIIS (JCP 116, 8255) and ADIIS (JCP 132, 054109): one weight
    vector for both spins from the (k, k) matrix tr(D_i F_j) (reduced on the device) and the
    reference's cost functions, F = sum_i x_i F_i.  EDIIS2 / ADIIS2 switch to CDIIS when the
    relative energy change (E_k - E_{k-1}) / E_{k-1} is below the threshold.  ADIIS takes the
    newest entry at index cycle % max_diis, as the reference does."""
    overlap_matrix: torch.Tensor
    A: Optional[torch.Tensor] = None
    max_diis: int = 8
    diis_method: str = "EDIIS2"
    ediis2_threshold: float = 1e-2
    adiis2_threshold: float = 1e-2

    METHODS = ("CDIIS", "EDIIS", "ADIIS", "EDIIS2", "ADIIS2")

    def _a(self) -> torch.Tensor:
        S = self.overlap_matrix
        if self.A is None:
            return torch.eye(S.shape[0], dtype=S.dtype, device=S.device)
        return self.A

    def empty(self):
        n = self.overlap_matrix.shape[0]
        z = lambda *s: torch.zeros(s, dtype=torch.float64, device=self.overlap_matrix.device)  # noqa: E731
        return (z(0, 2, n, n), z(0, 2, n, n), z(0), z(0, 2, n, n))

    def update(self, new_data, diis_data):
        rdm1, fock, energy = new_data
        dens, focks, energies, errs = diis_data
        A = self._a()
        fds = A @ fock @ rdm1 @ self.overlap_matrix @ A.T
        err = fds - fds.transpose(1, 2)
        errs = torch.cat([errs, err[None]])
        dens = torch.cat([dens, rdm1[None]])
        focks = torch.cat([focks, fock[None]])
        energy = torch.as_tensor(energy, dtype=torch.float64, device=energies.device).reshape(1)
        energies = torch.cat([energies, energy])
        if errs.shape[0] > self.max_diis:
            errs, dens, focks, energies = errs[1:], dens[1:], focks[1:], energies[1:]
        return dens, focks, energies, errs

    def run(self, new_data, diis_data, cycle: int = 0):
        if self.diis_method not in self.METHODS:
            raise ValueError(f"unknown DIIS method {self.diis_method!r}; one of {self.METHODS}")
        diis_data = self.update(new_data, diis_data)
        dens, focks, energies, errs = diis_data
        if errs.shape[0] == 1:
            return focks[0], diis_data
        method = self.diis_method
        rel = None
        if method in ("EDIIS2", "ADIIS2"):
            e_prev, e_last = energies[-2:].tolist()
            rel = (e_last - e_prev) / e_prev
        if (method == "CDIIS" or (method == "EDIIS2" and rel < self.ediis2_threshold)
                or (method == "ADIIS2" and rel < self.adiis2_threshold)):
            x = self.cdiis_minimize(errs)
            F = torch.einsum("si,isjk->sjk", x, focks)
            A = self._a()
            return A.T @ F @ A, diis_data
        if method in ("EDIIS", "EDIIS2"):
            x, _ = self.ediis_minimize(dens, focks, energies)
        else:
            x, _ = self.adiis_minimize(dens, focks, cycle % self.max_diis)
        F = torch.einsum("i,isjk->sjk", x, focks)
        return F, diis_data

    def cdiis_minimize(self, errs: torch.Tensor) -> torch.Tensor:
        """(2, k) weights: x = B^-1 e_0 with B = [[0, 1^T], [1, G_s]], G_s the Gram matrix."""
        k = errs.shape[0]
        E = errs.transpose(0, 1).reshape(2, k, -1)
        B = torch.zeros((2, k + 1, k + 1), dtype=torch.float64, device=errs.device)
        B[:, 1:, 1:] = E @ E.transpose(1, 2)
        B[:, 0, 1:] = 1.0
        B[:, 1:, 0] = 1.0
        if k + 1 <= 16:
            x = _diis_solve(B)
        else:
            e0 = torch.zeros((2, k + 1, 1), dtype=torch.float64, device=errs.device)
            e0[:, 0] = 1.0
            x = torch.linalg.solve(B, e0)[..., 0]
        if not bool(torch.isfinite(x).all()):
            raise AssertionError("CDIIS coefficients are not finite (singular error Gram matrix)")
        return x[:, 1:]

    @staticmethod
    def _trace_matrix(dens: torch.Tensor, focks: torch.Tensor) -> torch.Tensor:
        # df_ij = sum_s tr(D_i,s F_j,s), one (k, 2n^2) x (2n^2, k) product on the device
        k = dens.shape[0]
        return dens.reshape(k, -1) @ focks.transpose(2, 3).reshape(k, -1).T

    def ediis_minimize(self, dens, focks, energies):
        df = self._trace_matrix(dens, focks)
        diag = df.diagonal()
        df = diag[:, None] + diag - df - df.T
        c, fun = _simplex_minimize(energies.cpu().numpy(), -df.cpu().numpy())
        return torch.as_tensor(c, dtype=torch.float64, device=dens.device), fun

    def adiis_minimize(self, dens, focks, idnewest: int):
        df = self._trace_matrix(dens, focks)
        d_fn = df[:, idnewest]
        dn_f = df[idnewest]
        dn_fn = df[idnewest, idnewest]
        dd_fn = d_fn - dn_fn
        df = df - d_fn[:, None] - dn_f + dn_fn
        c, fun = _simplex_minimize(2.0 * dd_fn.cpu().numpy(), df.cpu().numpy())
        return torch.as_tensor(c, dtype=torch.float64, device=dens.device), fun


def smearing_occupations(mo_energy: torch.Tensor, nelectron: float, smearing: str, sigma: float) -> torch.Tensor:
    """Fractional occupations of ``scf_loop``'s smearing (``evaluate.py:467-497``): one chemical
    potential mu for all spin orbitals, found by bisection (xtol = rtol = 1e-10) between the
    lowest and highest orbital energy so that the occupations sum to nelectron; "gaussian":
    erfc((e - mu) / sigma) / 2, "fermi-dirac": 1 / (exp((e - mu) / sigma) + 1)."""
    import numpy as np
    from scipy.optimize import bisect
    from scipy.special import erfc

    e = mo_energy.detach().double().cpu().numpy().reshape(-1)
    if smearing == "gaussian":
        occf = lambda m: 0.5 * erfc((e - m) / sigma)  # noqa: E731
    elif smearing == "fermi-dirac":
        def occf(m):
            with np.errstate(over="ignore"):
                return 1.0 / (np.exp((e - m) / sigma) + 1.0)
    else:
        raise ValueError(f"unknown smearing {smearing!r}; 'gaussian' or 'fermi-dirac'")
    mu = bisect(lambda m: occf(m).sum() - nelectron, e.min(), e.max(), xtol=1e-10, rtol=1e-10, maxiter=10000)
    return torch.as_tensor(occf(mu).reshape(tuple(mo_energy.shape)), dtype=torch.float64, device=mo_energy.device)


def _check_charge(functional: Functional, molecule: Molecule, precision: str, nelectron: float) -> None:
    """The loops' charge-conservation assert: sum_r w_r (rho_up + rho_dn)(r) = N_e within
    atol 1e-3 (and jnp.isclose's rtol 1e-5); rho from the device density kernel."""
    rho = functional.plan(molecule, precision).densities(molecule.rdm1.detach())[:, :, 0].sum(dim=1)
    charge = float((molecule.grid.weights * rho).sum())
    if not abs(nelectron - charge) <= 1e-3 + 1e-5 * abs(charge):
        raise AssertionError("Total charge is not conserved. given electrons: %.3f, computed electrons: %.3f"
                             % (nelectron, charge))


def _linv(molecule: Molecule) -> torch.Tensor:
    Linv = molecule.cache.get("Linv")
    if Linv is None:
        _, Linv = cholesky_inv(molecule.s1e)
        molecule.cache["Linv"] = Linv
    return Linv


def scf_loop(functional: Functional, level_shift_factor=(0.0, 0.0), damp_factor=(0.0, 0.0),
             chunk_size: int = 1024, cycles: int = 50, diis_start_cycle: int = 0, e_conv: float = 1e-5,
             g_conv: float = 1e-5, diis_method: str = "CDIIS", smearing: Optional[str] = None,
             smearing_sigma: float = 0.0, verbose: int = 0, precision: str = "fp64", **kwargs) -> Callable:
    """``scf_loop`` (``evaluate.py:355-592``): predict at the molecule's density, then while
    (|dE| in kcal/mol > e_conv or |g_orb| > g_conv) and cycle < cycles: damping of the Fock
    matrix for cycle < diis_start_cycle - 1 (when the factors are nonzero), DIIS from
    diis_start_cycle on (``Diis``, max_diis 10, A = I), level shift, diagonalisation, aufbau
    (or smeared) occupations, new density, charge check, predict.  A converged loop ends with
    one more unshifted diagonalisation and prediction.

    The orbital-gradient norm is taken, as in the reference, with the Fock matrix the molecule
    carried in (``Molecule.get_mo_grads`` on the loop's molecule, whose ``fock`` field the loop
    does not update); a molecule without one uses the loop's current Fock matrix.  Smearing
    uses one chemical potential for the spin orbitals of both spins and keeps the (2, n)
    occupation layout.  A molecule built with ``omegas`` has its ``chi`` tensor rebuilt on the
    device from every new density (``evaluate.py:510-522, 565-576``), so its semi-numerical
    exact exchange follows the current density; without chi the exchange is the packed-ERI
    term of the initial density.  Not differentiable (runs without autograd)."""
    compute_energy = energy_predictor(functional, precision=precision, **kwargs)
    spin_equal = damp_factor[0] == damp_factor[1] and level_shift_factor[0] == level_shift_factor[1]

    def solve(fock, Linv, V, molecule, nelectron):
        mo_energy, mo_coeff, V = orthogonal_eigh(fock, Linv, V)
        molecule = molecule.replace(mo_coeff=mo_coeff, mo_energy=mo_energy)
        mo_occ = molecule.get_occ()
        if smearing:
            mo_occ = smearing_occupations(mo_energy, nelectron, smearing, smearing_sigma)
        molecule = molecule.replace(mo_occ=mo_occ)
        molecule = molecule.replace(rdm1=molecule.make_rdm1(),
                                    restricted=molecule.restricted and functional.spin_symmetric and spin_equal)
        if molecule.omegas is not None and molecule.chi_grid is not None:
            from .hf import generate_chi_tensor
            molecule = molecule.replace(chi_grid=generate_chi_tensor(molecule, molecule.rdm1, molecule.omegas))
        return molecule, V

    def scf_iterator(params: Optional[torch.Tensor], molecule: Molecule) -> Molecule:
        with torch.no_grad():
            S = molecule.s1e
            Linv = _linv(molecule)
            carried_fock = molecule.fock
            nelectron = float(molecule.atom_index.sum()) - molecule.charge
            predicted_e, fock = compute_energy(params, molecule)
            e = float(predicted_e)
            old_e = float("inf")
            norm_gorb = float("inf")
            cycle = 0
            diis = Diis(S, None, max_diis=10, diis_method=diis_method)
            data = diis.empty()
            V = None
            damp = abs(damp_factor[0]) + abs(damp_factor[1]) > 1e-4
            shift = abs(level_shift_factor[0]) + abs(level_shift_factor[1]) > 1e-4

            def gnorm(mol, f):
                g = orbital_grad(mol.mo_coeff, mol.mo_occ, carried_fock if carried_fock is not None else f)
                return float(torch.linalg.norm(g))

            while (abs(e - old_e) * HARTREE2KCALMOL > e_conv or norm_gorb > g_conv) and cycle < cycles:
                old_e = e
                if 0 <= cycle < diis_start_cycle - 1 and damp:
                    fock = torch.stack([damping(S, molecule.rdm1[s], fock[s], damp_factor[s]) for s in range(2)])
                if cycle >= diis_start_cycle:
                    fock, data = diis.run((molecule.rdm1, fock, predicted_e), data, cycle)
                if shift:
                    fock = torch.stack([level_shift(S, molecule.rdm1[s], fock[s], level_shift_factor[s])
                                        for s in range(2)])
                molecule, V = solve(fock, Linv, V, molecule, nelectron)
                _check_charge(functional, molecule, precision, nelectron)
                predicted_e, fock = compute_energy(params, molecule)
                e = float(predicted_e)
                norm_gorb = gnorm(molecule, fock)
                if verbose > 1:
                    print(f"cycle: {cycle}, energy: {e:.7e}, energy difference: {abs(e - old_e):.4e}, "
                          f"norm_gradient_orbitals: {norm_gorb:.2e}")
                cycle += 1

            if abs(e - old_e) * HARTREE2KCALMOL < e_conv and norm_gorb < g_conv:
                # one more diagonalisation of the unshifted Fock matrix
                molecule, V = solve(fock, Linv, V, molecule, nelectron)
                predicted_e, fock = compute_energy(params, molecule)
                e = float(predicted_e)
                norm_gorb = gnorm(molecule, fock)
            if verbose > 1:
                print(f"cycle: {cycle}, predicted energy: {e:.7e}, energy difference: {abs(e - old_e):.4e}, "
                      f"norm_gradient_orbitals: {norm_gorb:.2e}")
            out = molecule.replace(fock=fock, energy=predicted_e)
            out.cache["scf_cycles"] = cycle
            return out

    return scf_iterator


def simple_scf_loop(functional: Functional, mixing_factor: float = 0.4, chunk_size: int = 1024,
                    cycles: int = 50, start_cycle: int = 0, e_conv: float = 1e-5, g_conv: float = 1e-5,
                    verbose: int = 0, precision: str = "fp64", **kwargs) -> Callable:
    """``simple_scf_loop`` (``evaluate.py:131-254``): a fixed number of cycles; cycle 0 keeps
    the molecule's orbitals (aufbau occupations from its mo_energy) and predicts at its
    density, which becomes the mixing reference make_rdm1(); every later cycle diagonalises the
    previous Fock matrix and mixes D = (1 - mixing_factor) D_old + mixing_factor D_new.  Charge
    check each cycle.  Autograd records through it when the caller enables it (the device
    predictor and eigensolver are differentiable), unlike the reference's asserting loop."""
    compute_energy = energy_predictor(functional, precision=precision, **kwargs)

    def simple_scf_iterator(params: Optional[torch.Tensor], atoms: Molecule, clip_cte: float = 1e-30) -> Molecule:
        Linv = _linv(atoms)
        nelectron = float(atoms.atom_index.sum()) - atoms.charge
        fock = atoms.fock
        old_rdm1 = None
        V = None
        predicted_e = None
        for cycle in range(cycles):
            if cycle > 0:
                mo_energy, mo_coeff, V = orthogonal_eigh(fock, Linv, V)
                atoms = atoms.replace(mo_coeff=mo_coeff, mo_energy=mo_energy)
            atoms = atoms.replace(mo_occ=atoms.get_occ())
            if cycle == 0:
                old_rdm1 = atoms.make_rdm1()
            else:
                rdm1 = (1 - mixing_factor) * old_rdm1 + mixing_factor * atoms.make_rdm1()
                atoms = atoms.replace(rdm1=rdm1, restricted=atoms.restricted and functional.spin_symmetric)
                old_rdm1 = rdm1
            _check_charge(functional, atoms, precision, nelectron)
            predicted_e, fock = compute_energy(params, atoms)
            if verbose > 0:
                print(f"cycle: {cycle}, energy: {float(predicted_e):.7e}")
        return atoms.replace(fock=fock, energy=predicted_e)

    return simple_scf_iterator


# ---------------------------------------------------------------------------------------
# Orbital optimizers (evaluate.py:595-914): minimise the predicted energy over orthonormal
# orbitals C = (qr(W) D)^T with an optax-style optimizer on W; dE/dW by reverse mode through
# the device predictor (dE/dD = V_xc + J + h) and torch's QR.
# ---------------------------------------------------------------------------------------

def whitening_matrix(S: torch.Tensor, whitening: str = "PCA") -> torch.Tensor:
    """D with D^T D = S^-1 (``evaluate.py:729-740``): "PCA" diag(w^-1/2) V^T, "Cholesky"
    chol(S^-1)^T, "ZCA" V diag(w^-1/2) V^T.  Eigenvector signs are fixed so that each
    eigenvector's largest-magnitude component is positive (LAPACK leaves them arbitrary; the
    starting orbitals after the QR step depend on them)."""
    if whitening == "Cholesky":
        return torch.linalg.cholesky(torch.linalg.inv(S)).T.contiguous()
    if whitening not in ("PCA", "ZCA"):
        raise ValueError(f"unknown whitening {whitening!r}; 'PCA', 'Cholesky' or 'ZCA'")
    w, V = eigh(S[None].contiguous())
    w, V = w[0], V[0]
    idx = V.abs().argmax(dim=0)
    V = V * torch.sign(V[idx, torch.arange(V.shape[1], device=V.device)])[None, :]
    if whitening == "PCA":
        return (w.rsqrt()[:, None] * V.T).contiguous()
    return (V * w.rsqrt()[None, :]) @ V.T


def _orbital_energy(compute_energy, params, molecule: Molecule, W: torch.Tensor, D: torch.Tensor):
    """molecule_orbitals_iterator (``evaluate.py:642-699``): Q_s = qr(W_s) (LAPACK's
    Householder signs), C_s = (Q_s D)^T, rdm1 = make_rdm1(C, mo_occ), predicted energy."""
    Q = householder_q(W)
    C = torch.einsum("sij,jk->ski", Q, D)
    rdm1 = make_rdm1(C, molecule.mo_occ)
    e, _ = compute_energy(params, molecule.replace(rdm1=rdm1, mo_coeff=C, restricted=False))
    return e


def _orbital_step(compute_energy, params, molecule, W, D, tx, opt_state):
    Wg = W.detach().requires_grad_(True)
    with torch.enable_grad():
        E = _orbital_energy(compute_energy, params, molecule, Wg, D)
        g, = torch.autograd.grad(E, [Wg])
    updates, opt_state = tx.update(g, opt_state, W)
    return E.detach(), W + updates, opt_state


def mol_orb_optimizer(fxc: Functional, tx, chunk_size: int = 1024, cycles: int = 500, e_conv: float = 1e-7,
                      whitening: str = "PCA", verbose: int = 0, precision: str = "fp64", **kwargs) -> Callable:
    """``mol_orb_optimizer`` (``evaluate.py:595-777``, D4FT, ICLR 2023): W = C^T D^-1 from the
    molecule's orbitals, then while |dE| in kcal/mol > e_conv and cycle < cycles: energy and
    dE/dW at W, one optimizer step.  Returns the molecule with the last predicted energy (its
    density and orbitals are left as they came in, as in the reference)."""
    compute_energy = energy_predictor(fxc, precision=precision, **kwargs)

    def neural_iterator(params: Optional[torch.Tensor], molecule: Molecule) -> Molecule:
        with torch.no_grad():
            predicted_e, _ = compute_energy(params, molecule)
            D = whitening_matrix(molecule.s1e, whitening)
            W = torch.einsum("sji,jk->sik", molecule.mo_coeff, torch.linalg.inv(D))
        opt_state = tx.init(W)
        e, old_e, cycle = float(predicted_e), float("inf"), 0
        while abs(e - old_e) * HARTREE2KCALMOL > e_conv and cycle < cycles:
            old_e = e
            predicted_e, W, opt_state = _orbital_step(compute_energy, params, molecule, W, D, tx, opt_state)
            e = float(predicted_e)
            cycle += 1
            if verbose > 1:
                print(f"cycle: {cycle}, predicted energy: {e:.7e}, energy difference: {abs(e - old_e):.4e}")
            molecule = molecule.replace(energy=predicted_e)
        molecule.cache["orb_cycles"] = cycle
        return molecule

    return neural_iterator


def jitted_mol_orb_optimizer(functional: Functional, tx, cycles: int = 500, precision: str = "fp64",
                             **kwargs) -> Callable:
    """``jitted_mol_orb_optimizer`` (``evaluate.py:783-914``): PCA whitening and exactly
    ``cycles`` optimizer steps; returns the molecule with the energy of the last step."""
    compute_energy = energy_predictor(functional, precision=precision, **kwargs)

    def neural_iterator(params: Optional[torch.Tensor], molecule: Molecule) -> Molecule:
        with torch.no_grad():
            predicted_e, _ = compute_energy(params, molecule)
            D = whitening_matrix(molecule.s1e, "PCA")
            W = torch.einsum("sji,jk->sik", molecule.mo_coeff, torch.linalg.inv(D))
        opt_state = tx.init(W)
        for _ in range(cycles):
            predicted_e, W, opt_state = _orbital_step(compute_energy, params, molecule, W, D, tx, opt_state)
        return molecule.replace(energy=predicted_e)

    return neural_iterator


def batched_scf_loop(functional: Functional, cycles: int = 25, precision: str = "fp64", n_streams: int = 4,
                     **kwargs) -> Callable:
    """Forward self-consistent loops of a batch of molecules run concurrently: each molecule's
    ``diff_scf_loop`` is one captured HIP graph (captured on its first call, after one eager
    run), and the replays of different molecules are issued round-robin on ``n_streams``
    streams (default priority), so small molecules -- each filling a fraction of the GPU --
    overlap.  Results equal the molecule-by-molecule loop (``train.py:519-527`` evaluates a
    batch that way).  Forward only; for reverse mode call ``diff_scf_loop`` per molecule.

    ``iterator(params, molecules) -> [molecule]`` with the loops' outputs.  The molecules'
    status words are read once, after every replay has finished (check_status)."""
    check = kwargs.pop("check_status", True)
    loop = diff_scf_loop(functional, cycles=cycles, precision=precision, graph=True, check_status=False, **kwargs)
    streams: List[torch.cuda.Stream] = []
    captured = set()   # (id of the molecule's cache, params given): the cache holds the graph

    def iterator(params: Optional[torch.Tensor], molecules: List[Molecule]) -> List[Molecule]:
        with torch.no_grad():
            for m in molecules:           # first call per molecule: eager run + capture
                tag = (id(m.cache), params is None)
                if tag not in captured:
                    loop(params, m)
                    captured.add(tag)
            if not streams:
                streams.extend(torch.cuda.Stream() for _ in range(max(1, n_streams)))
            main = torch.cuda.current_stream()
            for s in streams:
                s.wait_stream(main)
            # molecules sharing a cache (Molecule.replace keeps it; a molecule listed twice)
            # replay the same captured graph into the same buffers: they share one stream, so
            # their replays run one after another
            lane = {}
            outs = []
            for m in molecules:
                k = lane.setdefault(id(m.cache), len(lane) % len(streams))
                with torch.cuda.stream(streams[k]):
                    outs.append(loop(params, m))
            for s in streams:
                main.wait_stream(s)
            for o in outs:                # outputs allocated on the side streams, used on main
                for t in (o.energy, o.rdm1, o.fock, o.mo_coeff, o.mo_energy, o.mo_occ):
                    t.record_stream(main)
            if check:
                for i, o in enumerate(outs):
                    raise_on_status(_status_buffers(o)[0], f"batched_scf_loop (molecule {i})")
            return outs

    return iterator
