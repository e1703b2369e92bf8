"""graddft_amd: GradDFT's self-consistent exchange-correlation grid path, native on MI355X.

Public API (names follow XanaduAI/GradDFT's ``grad_dft`` package):
``Grid``, ``Molecule``, ``Functional``, ``NeuralFunctional``, ``LinearFunctional``, ``MLPSpec``,
``LSDA``, ``B88``, ``VWN``, ``PW92``, ``LYP``, ``B3LYP``, ``generate_chi_tensor``,
``energy_predictor``, ``non_scf_predictor``, ``diff_scf_loop``, ``diff_simple_scf_loop``,
``scf_loop``, ``simple_scf_loop``, ``Diis``, ``damping``, ``level_shift``, ``mol_orb_optimizer``,
``jitted_mol_orb_optimizer``, ``simple_energy_loss``,
``mse_energy_loss``, ``train_kernel``, ``adam``, ``molecule_from_geometry``.
"""
from .molecule import Grid, Molecule
from .functional import (Functional, LinearFunctional, NeuralFunctional, B3LYPNeuralCorrection, CustomFunctional,
                         MLPSpec, XCPlan, XCEnergyFn, XCEnergyVxcFn)
from .popular_functionals import B3LYP, B3LYP_GRID, B88, LSDA, LYP, PW92, VWN
from .hf import generate_chi_tensor
from .interface.builder import molecule_from_geometry
from .interface.store import load_molecules, save_molecules
from .train import (energy_predictor, simple_energy_loss, mse_energy_loss, train_kernel, adam,
                    apply_updates)
from .evaluate import (Diis, JittableDiis, batched_scf_loop, damping, diff_scf_loop, diff_simple_scf_loop,
                       jitted_mol_orb_optimizer, level_shift, mol_orb_optimizer, non_scf_predictor, scf_loop,
                       simple_scf_loop)
from .eigenproblem import safe_eigh_vjp, safe_fock_solver, safe_general_eigh
from .batch import XCBatchPlan, batch_energy_predictor, batched_diff_scf_loop

__all__ = ["Grid", "Molecule", "Functional", "LinearFunctional", "NeuralFunctional", "B3LYPNeuralCorrection",
           "CustomFunctional",
           "MLPSpec", "XCPlan", "XCEnergyFn", "XCEnergyVxcFn", "B3LYP", "B3LYP_GRID", "LSDA", "B88", "VWN", "PW92",
           "LYP", "generate_chi_tensor", "molecule_from_geometry",
           "energy_predictor", "non_scf_predictor", "simple_energy_loss", "mse_energy_loss",
           "train_kernel", "adam", "apply_updates", "JittableDiis", "Diis", "diff_scf_loop", "batched_scf_loop", "diff_simple_scf_loop",
           "scf_loop", "simple_scf_loop", "damping", "level_shift", "mol_orb_optimizer",
           "jitted_mol_orb_optimizer", "safe_eigh_vjp",
           "safe_fock_solver", "safe_general_eigh", "save_molecules", "load_molecules", "XCBatchPlan",
           "batch_energy_predictor", "batched_diff_scf_loop"]
