"""graddft_amd: GradDFT's self-consistent exchange-correlation grid path, native on MI355X.

Public API (names follow XanaduAI/GradDFT's ``grad_dft`` package):
``Grid``, ``Molecule``, ``Functional``, ``NeuralFunctional``, ``MLPSpec``, ``B3LYP``,
``energy_predictor``, ``diff_scf_loop``, ``molecule_from_geometry``.
"""
from .molecule import Grid, Molecule
from .functional import Functional, NeuralFunctional, MLPSpec, XCPlan, XCEnergyFn
from .popular_functionals import B3LYP, B3LYP_GRID
from .interface.builder import molecule_from_geometry
from .train import energy_predictor
from .evaluate import JittableDiis, diff_scf_loop
from .eigenproblem import safe_eigh_vjp, safe_fock_solver, safe_general_eigh

__all__ = ["Grid", "Molecule", "Functional", "NeuralFunctional", "MLPSpec", "XCPlan",
           "XCEnergyFn", "B3LYP", "B3LYP_GRID", "molecule_from_geometry", "energy_predictor",
           "JittableDiis", "diff_scf_loop", "safe_eigh_vjp", "safe_fock_solver", "safe_general_eigh"]
