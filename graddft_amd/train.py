"""Energy and Fock predictor (``grad_dft/train.py:34-218``).

``energy_predictor(functional)(params, molecule) -> (E, fock)``:
E_xc and its density-matrix gradient come from one forward + adjoint pass of the XC grid
kernels (the ``value_and_grad`` of ``xc_energy_and_grads``, ``train.py:86-121``); the
Coulomb matrix from the packed ERI; the non-XC energy (``molecule.py:694-733``); and the
Fock matrix is assembled with the reference's clips and symmetrisation (``:148-215``),
including the explicit exact-exchange term of B3LYP (``:200-206``).
"""
from __future__ import annotations

from typing import Callable, Optional, Tuple

import torch

from .functional import Functional
from .molecule import Molecule
from .scf import coulomb, fock_assemble, non_xc_energy


def energy_predictor(functional: Functional, precision: str = "fp64", **kwargs) -> Callable:
    def predict(params: Optional[torch.Tensor], molecule: Molecule) -> Tuple[torch.Tensor, torch.Tensor]:
        plan = functional.plan(molecule, precision)
        D = molecule.rdm1.contiguous()
        exc = torch.empty((), dtype=torch.float64, device=D.device)
        vxc = torch.empty_like(D)
        one = torch.ones((), dtype=torch.float64, device=D.device)
        p = None if params is None else params.contiguous()
        plan.forward(D, p, exc)
        plan.backward(D, p, one, vxc, None, plan.generation)
        J = coulomb(molecule, D)
        a0 = functional.exact_exchange
        energy = exc + non_xc_energy(molecule, D, J, exact_exchange=a0)
        fock = fock_assemble(molecule, J, vxc, exact_exchange=a0)
        return energy, fock

    return predict
