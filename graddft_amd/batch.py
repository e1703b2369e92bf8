"""A batch of molecules through the XC grid path as one evaluation, and the differentiable SCF
loop and losses on top of it.

The reference evaluates a training batch molecule by molecule (``grad_dft/train.py:519-527``,
one ``diff_scf_loop`` each).  Here the molecules of a batch share one XC evaluation per SCF
cycle (``gdx_xc_batch_*``): their grid points are concatenated, the coefficient MLP runs once
over all of them (one GEMM chain instead of one per molecule), and the AO-side kernels (density,
V_xc and their padding and reduction) are single launches over a ragged queue of
(molecule, tile) work items.  The SCF bookkeeping of each molecule (Coulomb matrix, Fock
assembly, DIIS, eigensolver) stays per molecule, on the same device kernels as
``diff_scf_loop``.

Densities, their cotangents and V_xc travel as one concatenated vector: molecule m's (2, n, n)
block starts at ``plan.offsets[m]``.
"""
from __future__ import annotations

import ctypes
from typing import Callable, List, Optional, Sequence

import torch

from . import _lib
from .functional import PRECISIONS, Functional, ao_planes_for
from .molecule import Molecule


class XCBatchPlan:
    """Device plan of one (functional, precision) over a fixed list of molecules."""

    def __init__(self, functional: Functional, molecules: Sequence[Molecule], precision: str = "fp64",
                 restricted: Optional[bool] = None, hvp: bool = True):
        if getattr(functional, "custom", False):
            raise NotImplementedError("custom coefficient functions are evaluated per molecule (energy_predictor)")
        lib = _lib.load()
        if restricted is None:
            restricted = functional.spin_symmetric and all(m.restricted for m in molecules)
        d = _lib.XCDesc()
        d.spin_restricted = 1 if restricted else 0
        d.kind = functional.kind
        d.precision = PRECISIONS[precision]
        spec = functional.spec
        if spec is not None:
            d.mlp_nc, d.mlp_h, d.mlp_nhidden, d.mlp_nd = spec.nc, spec.h, spec.n_hidden, spec.nd
        d.correlation = int(functional.correlation)
        if functional.coef is not None:
            for f, c in enumerate(functional.coef):
                d.coef[f] = float(c)
        self.desc = d
        self.molecules = list(molecules)
        self._ao = [ao_planes_for(m, precision) for m in molecules]   # kept alive with the plan
        arr = (_lib.XCBatchMol * len(molecules))()
        for i, (m, ao) in enumerate(zip(molecules, self._ao)):
            if m.npts_pad % 256:
                raise ValueError(f"molecule {i}: a batch needs npts_pad % 256 == 0 (build it with npts_align=256)")
            arr[i].ao = ao.data_ptr()
            arr[i].weights = m.grid.weights_pad.data_ptr()
            arr[i].npts, arr[i].npts_pad, arr[i].nao, arr[i].nao_pad = m.grid_size, m.npts_pad, m.nao, m.nao_pad
        handle = ctypes.c_void_p(0)
        _lib.check(lib.gdx_xc_batch_create(ctypes.byref(d), len(molecules), arr, 1 if hvp else 0,
                                           ctypes.byref(handle)), "gdx_xc_batch_create")
        self.handle = handle
        off = (ctypes.c_int64 * (len(molecules) + 1))()
        _lib.check(lib.gdx_xc_batch_offsets(handle, off), "gdx_xc_batch_offsets")
        self.offsets = [int(v) for v in off]
        self.device = molecules[0].ao_planes.device
        self.restricted = restricted

    def close(self) -> None:
        """Free the plan's device workspaces now (they are allocated by the library, outside
        torch's caching allocator); the plan is unusable afterwards."""
        h = getattr(self, "handle", None)
        if h is not None and h.value and _lib._lib is not None:
            _lib._lib.gdx_xc_batch_destroy(h)
        self.handle = None

    def __del__(self):
        self.close()

    @property
    def nmol(self) -> int:
        return len(self.molecules)

    def concat(self, mats: Sequence[torch.Tensor]) -> torch.Tensor:
        return torch.cat([m.reshape(-1) for m in mats]).contiguous()

    def split(self, cat: torch.Tensor) -> List[torch.Tensor]:
        return [cat[self.offsets[i]:self.offsets[i + 1]].view(2, m.nao, m.nao)
                for i, m in enumerate(self.molecules)]

    def _stream(self) -> int:
        return torch.cuda.current_stream(self.device).cuda_stream

    @staticmethod
    def _ptr(t: Optional[torch.Tensor]) -> int:
        return 0 if t is None else t.data_ptr()

    def forward(self, D: torch.Tensor, params: Optional[torch.Tensor], exc: torch.Tensor) -> None:
        _lib.check(_lib.load().gdx_xc_batch_forward(self._stream(), self.handle, D.data_ptr(), self._ptr(params),
                                                    exc.data_ptr()), "gdx_xc_batch_forward")

    def backward(self, params, g: torch.Tensor, d_rdm1: Optional[torch.Tensor], d_params: Optional[torch.Tensor]):
        _lib.check(_lib.load().gdx_xc_batch_backward(self._stream(), self.handle, self._ptr(params), g.data_ptr(),
                                                     self._ptr(d_rdm1), self._ptr(d_params)), "gdx_xc_batch_backward")

    def hvp(self, D, params, direction, g, exc=None, edot=None, d_rdm1=None, d_params=None) -> None:
        _lib.check(_lib.load().gdx_xc_batch_hvp(self._stream(), self.handle, D.data_ptr(), self._ptr(params),
                                                direction.data_ptr(), g.data_ptr(), self._ptr(exc), self._ptr(edot),
                                                self._ptr(d_rdm1), self._ptr(d_params)), "gdx_xc_batch_hvp")


class XCBatchEnergyFn(torch.autograd.Function):
    """E_xc of every molecule (nmol,) from the concatenated densities; backward = the batch
    adjoint with one cotangent per molecule."""

    @staticmethod
    def forward(ctx, D, params, plan: XCBatchPlan):
        D = D.contiguous()
        p = None if params is None else params.contiguous()
        exc = torch.empty(plan.nmol, dtype=torch.float64, device=D.device)
        plan.forward(D, p, exc)
        ctx.plan = plan
        ctx.has_params = p is not None
        ctx.save_for_backward(D, p if p is not None else D)
        return exc

    @staticmethod
    @torch.autograd.function.once_differentiable
    def backward(ctx, g):
        D, p = ctx.saved_tensors
        p = p if ctx.has_params else None
        plan = ctx.plan
        # the forward state is the plan's: re-run it (the workspace may have been reused since)
        exc = torch.empty(plan.nmol, dtype=torch.float64, device=D.device)
        plan.forward(D, p, exc)
        dD = torch.empty_like(D) if ctx.needs_input_grad[0] else None
        dp = torch.empty_like(p) if (p is not None and ctx.needs_input_grad[1]) else None
        plan.backward(p, g.detach().to(torch.float64).contiguous(), dD, dp)
        return dD, dp, None


class XCBatchEnergyVxcFn(torch.autograd.Function):
    """(E_xc (nmol,), dE_m/dD_m concatenated) of the batch; backward = the batch second-order
    pass with the cotangent of V_xc as the direction and that of E_xc as the per-molecule g."""

    @staticmethod
    def forward(ctx, D, params, plan: XCBatchPlan):
        D = D.contiguous()
        p = None if params is None else params.contiguous()
        exc = torch.empty(plan.nmol, dtype=torch.float64, device=D.device)
        vxc = torch.empty_like(D)
        plan.forward(D, p, exc)
        plan.backward(p, torch.ones(plan.nmol, dtype=torch.float64, device=D.device), vxc, None)
        ctx.plan = plan
        ctx.has_params = p is not None
        ctx.save_for_backward(D, p if p is not None else D)
        return exc, vxc

    @staticmethod
    @torch.autograd.function.once_differentiable
    def backward(ctx, g_exc, g_vxc):
        D, p = ctx.saved_tensors
        p = p if ctx.has_params else None
        plan = ctx.plan
        g = torch.zeros(plan.nmol, dtype=torch.float64, device=D.device) if g_exc is None else \
            g_exc.detach().to(torch.float64).contiguous()
        direction = torch.zeros_like(D) if g_vxc is None else g_vxc.detach().contiguous()
        dD = torch.empty_like(D) if ctx.needs_input_grad[0] else None
        dp = torch.empty_like(p) if (p is not None and ctx.needs_input_grad[1]) else None
        plan.hvp(D, p, direction, g, d_rdm1=dD, d_params=dp)
        return dD, dp, None


def batch_plan(functional: Functional, molecules: Sequence[Molecule], precision: str = "fp64") -> XCBatchPlan:
    """The batch plan of these molecules (cached in the first molecule's cache; molecules made
    by ``Molecule.replace`` share their cache, so an SCF loop reuses one plan)."""
    restricted = functional.spin_symmetric and all(m.restricted for m in molecules)
    key = ("xcbatch", tuple(id(m.cache) for m in molecules), functional.kind, functional.spec, functional.coef,
           functional.correlation, precision, restricted)
    plan = molecules[0].cache.get(key)
    if plan is None:
        plan = XCBatchPlan(functional, molecules, precision, restricted)
        molecules[0].cache[key] = plan
    plan.molecules = list(molecules)
    return plan


def batch_energy_predictor(functional: Functional, precision: str = "fp64") -> Callable:
    """``energy_predictor`` (train.py:34-218) of a list of molecules with one batched XC
    evaluation: returns (energies, focks), one per molecule."""
    from .scf import CoulombFn, ExchangeFn, FockAssembleFn, NonXCFn, coulomb, exchange, fock_assemble, non_xc_energy

    def predict(params: Optional[torch.Tensor], molecules: Sequence[Molecule]):
        plan = batch_plan(functional, molecules, precision)
        a0 = functional.exact_exchange
        current = bool(a0) and functional.exchange == "current"
        recording = torch.is_grad_enabled() and (
            (params is not None and params.requires_grad) or any(m.rdm1.requires_grad for m in molecules))
        D = plan.concat([m.rdm1 for m in molecules])
        energies, focks = [], []
        if recording:
            exc, vxc = XCBatchEnergyVxcFn.apply(D, params, plan)
            for i, (m, v) in enumerate(zip(molecules, plan.split(vxc))):
                J = CoulombFn.apply(m.rdm1, m)
                K = ExchangeFn.apply(m.rdm1, m) if current else None
                energies.append(exc[i] + NonXCFn.apply(m.rdm1, J, K, m, a0))
                focks.append(FockAssembleFn.apply(J, v, K, m, a0))
            return energies, focks
        p = None if params is None else params.detach().contiguous()
        D = D.detach()
        exc = torch.empty(plan.nmol, dtype=torch.float64, device=D.device)
        vxc = torch.empty_like(D)
        plan.forward(D, p, exc)
        plan.backward(p, torch.ones(plan.nmol, dtype=torch.float64, device=D.device), vxc, None)
        for i, (m, v) in enumerate(zip(molecules, plan.split(vxc))):
            Dm = m.rdm1.detach().contiguous()
            J = coulomb(m, Dm)
            K = exchange(m, Dm) if current else None
            energies.append(exc[i] + non_xc_energy(m, Dm, J, exact_exchange=a0, K=K))
            focks.append(fock_assemble(m, J, v, exact_exchange=a0, K=K))
        return energies, focks

    return predict


def batched_diff_scf_loop(functional: Functional, cycles: int = 25, precision: str = "fp64",
                          refine_cycles: Optional[int] = None, check_status: bool = True) -> Callable:
    """``diff_scf_loop`` (evaluate.py:917-1038) of a list of molecules, differentiable, with the
    XC evaluation of every cycle batched over the molecules (one ``gdx_xc_batch_*`` evaluation,
    whose reverse mode is the batch second-order pass).  Each molecule keeps its own DIIS history,
    eigensolver warm start and status word; the results equal ``diff_scf_loop`` per molecule.
    ``iterator(params, molecules) -> [molecule]``; ``iterator.batched`` is True, so
    ``mse_energy_loss`` evaluates the whole list in one call."""
    from .eigenproblem import cholesky_inv, orthogonal_eigh, raise_on_status, scf_status, scf_step
    from .evaluate import JittableDiis, _aufbau_occupations, _diis_coefficients, _status_buffers
    if refine_cycles is None:
        refine_cycles = 0 if precision == "fp64" else 1
    predict = batch_energy_predictor(functional, precision)
    predict64 = batch_energy_predictor(functional, "fp64")

    def iterator(params: Optional[torch.Tensor], molecules: Sequence[Molecule]) -> List[Molecule]:
        mols = list(molecules)
        recording = torch.is_grad_enabled() and (
            (params is not None and params.requires_grad) or any(m.rdm1.requires_grad for m in mols))
        st = []
        for m in mols:
            Linv = m.cache.get("Linv")
            if Linv is None:
                _, Linv = cholesky_inv(m.s1e)
                m.cache["Linv"] = Linv
            status, sweeps = _status_buffers(m)
            status.zero_()
            diis = JittableDiis(m.s1e, max_diis=10)
            st.append({"Linv": Linv, "diis": diis, "data": diis.zeros(m.nao, m.s1e.device),
                       "occ": _aufbau_occupations(m), "V": None, "status": status, "sweeps": sweeps})
        energies, focks = predict(params, mols)
        mols = [m.replace(fock=f) for m, f in zip(mols, focks)]
        for cycle in range(cycles):
            for i, m in enumerate(mols):
                s = st[i]
                if not recording:
                    s["data"] = s["diis"].update(m.rdm1, m.fock, energies[i], s["data"], cycle)
                    _, x = _diis_coefficients(s["data"][3], cycle)
                    F, mo_energy, mo_coeff, s["V"], rdm1 = scf_step(s["data"][1], x, s["Linv"], s["V"], s["occ"],
                                                                    s["sweeps"])
                else:
                    F, s["data"] = s["diis"].run(m.rdm1, m.fock, energies[i], s["data"], cycle)
                    mo_energy, mo_coeff, s["V"] = orthogonal_eigh(F, s["Linv"], s["V"], s["sweeps"])
                    rdm1 = torch.einsum("sij,sj,skj->sik", mo_coeff, s["occ"], mo_coeff)
                mols[i] = m.replace(fock=F, mo_coeff=mo_coeff, mo_energy=mo_energy, mo_occ=s["occ"], rdm1=rdm1,
                                    restricted=m.restricted and functional.spin_symmetric)
            pred = predict64 if cycle >= cycles - refine_cycles else predict
            energies, focks = pred(params, mols)
            for i, m in enumerate(mols):
                mols[i] = m.replace(fock=focks[i])
                scf_status(mols[i].rdm1, m.s1e, st[i]["occ"], st[i]["sweeps"], energies[i], st[i]["status"])
        out = [m.replace(energy=e) for m, e in zip(mols, energies)]
        if check_status:
            for i, s in enumerate(st):
                raise_on_status(s["status"], f"batched_diff_scf_loop (molecule {i})")
        return out

    iterator.batched = True
    return iterator
