"""Energy and Fock predictor, losses and the training kernel (``grad_dft/train.py``).

``energy_predictor(functional)(params, molecule) -> (E, fock)`` (``train.py:34-218``):
E_xc and its density-matrix gradient come from one forward + adjoint pass of the XC grid
kernels (the ``value_and_grad`` of ``xc_energy_and_grads``, ``train.py:86-121``); the
Coulomb matrix from the packed ERI; the non-XC energy (``molecule.py:694-733``); and the
Fock matrix is assembled with the reference's clips and symmetrisation (``:148-215``),
including the explicit exact-exchange term of B3LYP (``:200-206``).  When autograd is
recording (params or rdm1 require grad) the same device computations run as autograd
Functions, so reverse mode reaches params through dE/dD (second-order pass ``gdx_xc_hvp``).

Training (``train.py:312-359, 478-555``): ``train_kernel(tx, loss)``,
``simple_energy_loss`` (value and gradient of (E - E_ref)^2), ``mse_energy_loss`` (mean
of ((E_m - E_m^ref) / N_e,m)^2 over molecules, differentiable), and ``adam`` /
``apply_updates``, an optax-style gradient transformation over the flat parameter tensor.
"""
from __future__ import annotations

import math
from dataclasses import dataclass
from typing import Callable, List, Optional, Sequence, Tuple, Union

import torch

from .functional import Functional, XCEnergyVxcFn
from .molecule import Molecule
from .scf import CoulombFn, ExchangeFn, FockAssembleFn, NonXCFn, coulomb, exchange, fock_assemble, non_xc_energy


def _recording(*ts) -> bool:
    return torch.is_grad_enabled() and any(t is not None and t.requires_grad for t in ts)


def energy_predictor(functional: Functional, precision: str = "fp64", **kwargs) -> Callable:
    if getattr(functional, "custom", False):
        return _custom_predictor(functional, precision)

    def predict(params: Optional[torch.Tensor], molecule: Molecule) -> Tuple[torch.Tensor, torch.Tensor]:
        plan = functional.plan(molecule, precision)
        D = molecule.rdm1
        a0 = functional.exact_exchange
        current = bool(a0) and functional.exchange == "current"
        if _recording(params, D):
            exc, vxc = XCEnergyVxcFn.apply(D, params, plan)
            J = CoulombFn.apply(D, molecule)
            K = ExchangeFn.apply(D, molecule) if current else None
            energy = exc + NonXCFn.apply(D, J, K, molecule, a0)
            fock = FockAssembleFn.apply(J, vxc, K, molecule, a0)
            return energy, fock
        D = D.detach().contiguous()
        exc = torch.empty((), dtype=torch.float64, device=D.device)
        vxc = torch.empty_like(D)
        one = torch.ones((), dtype=torch.float64, device=D.device)
        p = None if params is None else params.detach().contiguous()
        plan.forward(D, p, exc)
        plan.backward(D, p, one, vxc, None, plan.generation)
        J = coulomb(molecule, D)
        K = exchange(molecule, D) if current else None
        energy = exc + non_xc_energy(molecule, D, J, exact_exchange=a0, K=K)
        fock = fock_assemble(molecule, J, vxc, exact_exchange=a0, K=K)
        return energy, fock

    return predict


class _NoSecondOrderFn(torch.autograd.Function):
    """Passes V_xc of a custom coefficient function through; reverse mode through it would need
    the second derivatives of the user's function, which the device path does not form."""

    @staticmethod
    def forward(ctx, vxc, *inputs):
        return vxc.clone()

    @staticmethod
    def backward(ctx, g):
        raise NotImplementedError("reverse mode through V_xc of a custom coefficient function (an SCF loop "
                                  "or a loss of the Fock matrix) needs its second derivatives; differentiate "
                                  "the energy at a fixed density instead")


def _custom_predictor(functional, precision: str) -> Callable:
    """energy_predictor of a CustomFunctional: E_xc through autograd over the device-formed
    inputs and the user's coefficient function, V_xc = dE_xc/dD by one reverse pass."""
    if precision != "fp64":
        raise NotImplementedError("custom coefficient functions run in FP64")

    def predict(params: Optional[torch.Tensor], molecule: Molecule) -> Tuple[torch.Tensor, torch.Tensor]:
        D = molecule.rdm1
        a0 = functional.exact_exchange
        current = bool(a0) and functional.exchange == "current"
        if _recording(params, D):
            with torch.enable_grad():
                Dg = D if D.requires_grad else D.detach().requires_grad_(True)
                exc = functional.xc_energy(params, molecule, rdm1=Dg)
                vxc, = torch.autograd.grad(exc, [Dg], retain_graph=True)
            ins = [t for t in (params, D) if t is not None and t.requires_grad]
            vxc = _NoSecondOrderFn.apply(vxc, *ins)
            J = CoulombFn.apply(D, molecule)
            K = ExchangeFn.apply(D, molecule) if current else None
            energy = exc + NonXCFn.apply(D, J, K, molecule, a0)
            fock = FockAssembleFn.apply(J, vxc, K, molecule, a0)
            return energy, fock
        D = D.detach().contiguous()
        exc, vxc = functional.exc_vxc(params, molecule, D)
        J = coulomb(molecule, D)
        K = exchange(molecule, D) if current else None
        energy = exc + non_xc_energy(molecule, D, J, exact_exchange=a0, K=K)
        fock = fock_assemble(molecule, J, vxc, exact_exchange=a0, K=K)
        return energy, fock

    return predict


def num_electrons(molecule: Molecule) -> float:
    """sum(atom_index) - charge (train.py:521)."""
    return float(molecule.atom_index.sum().item()) - molecule.charge


def simple_energy_loss(params: torch.Tensor, compute_energy: Callable, atoms: Molecule, truth_energy: float):
    """``simple_energy_loss`` (train.py:532-555), a value_and_grad with aux:
    returns ((cost, E_predict), d cost / d params) with cost = (E - E_ref)^2."""
    p = params.detach().requires_grad_(True)
    with torch.enable_grad():
        out = compute_energy(p, atoms)
        diff = out.energy - truth_energy
        cost = diff * diff
    grads, = torch.autograd.grad(cost, [p])
    return (cost.detach(), out.energy.detach()), grads


def mse_energy_loss(params: torch.Tensor, compute_energy: Callable,
                    atoms_list: Union[Molecule, Sequence[Molecule]], truth_energies: Sequence[float],
                    elec_num_norm: bool = True) -> torch.Tensor:
    """``mse_energy_loss`` (train.py:478-530): sum_m ((E_m - E_m^ref) / N_e,m)^2 / M, with the
    molecules evaluated one after another as in the reference, or all at once when compute_energy
    is batched (``batched_diff_scf_loop``).  Differentiable w.r.t. params."""
    if isinstance(atoms_list, Molecule):
        atoms_list = [atoms_list]
    total = 0.0
    # a batched compute_energy (batched_diff_scf_loop) evaluates the whole list at once
    outs = compute_energy(params, list(atoms_list)) if getattr(compute_energy, "batched", False) else None
    for i, atoms in enumerate(atoms_list):
        out = outs[i] if outs is not None else compute_energy(params, atoms)
        diff = out.energy - truth_energies[i]
        if elec_num_norm:
            diff = diff / num_electrons(atoms)
        total = total + diff * diff
    return total / len(atoms_list)


@dataclass
class AdamState:
    count: int
    mu: torch.Tensor
    nu: torch.Tensor


class adam:  # noqa: N801  (optax naming)
    """optax.adam over one flat parameter tensor: mu = b1 mu + (1-b1) g, nu = b2 nu + (1-b2) g^2,
    update = -lr mu_hat / (sqrt(nu_hat) + eps) with bias-corrected mu_hat, nu_hat."""

    def __init__(self, learning_rate: float, b1: float = 0.9, b2: float = 0.999, eps: float = 1e-8):
        self.lr, self.b1, self.b2, self.eps = learning_rate, b1, b2, eps

    def init(self, params: torch.Tensor) -> AdamState:
        z = torch.zeros_like(params)
        return AdamState(0, z, z.clone())

    def update(self, grads: torch.Tensor, state: AdamState, params: Optional[torch.Tensor] = None):
        count = state.count + 1
        mu = self.b1 * state.mu + (1.0 - self.b1) * grads
        nu = self.b2 * state.nu + (1.0 - self.b2) * grads * grads
        mu_hat = mu / (1.0 - self.b1 ** count)
        nu_hat = nu / (1.0 - self.b2 ** count)
        updates = -self.lr * mu_hat / (torch.sqrt(nu_hat) + self.eps)
        return updates, AdamState(count, mu, nu)


def apply_updates(params: torch.Tensor, updates: torch.Tensor) -> torch.Tensor:
    return params + updates


def train_kernel(tx, loss: Callable) -> Callable:
    """``train_kernel`` (train.py:312-359): loss(params, atoms, truth) returns
    ((cost, predicted_energy), grads); the optimizer step is applied to params."""

    def kernel(params: torch.Tensor, opt_state, atoms, ground_truth_energy, *args):
        (cost_value, predicted_energy), grads = loss(params, atoms, ground_truth_energy, *args)
        updates, opt_state = tx.update(grads, opt_state, params)
        params = apply_updates(params, updates)
        return params, opt_state, cost_value, predicted_energy

    return kernel
