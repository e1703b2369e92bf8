"""SCF-side terms on the device: Coulomb matrix, non-XC energy, exchange of the initial
density (the stale exact-exchange term of B3LYP), Fock assembly.

Reference: ``grad_dft/molecule.py:694-811`` (nonXC, coulomb_potential),
``grad_dft/train.py:148-215`` (Fock assembly), ``molecule.py:507-613`` (exact exchange from
the initial density; here from the packed ERI instead of the semi-numerical chi tensor).
"""
from __future__ import annotations

import ctypes
from typing import Optional

import torch

from . import _lib
from .molecule import Molecule


def _stream(t: torch.Tensor) -> int:
    return torch.cuda.current_stream(t.device).cuda_stream


def _workspace(molecule: Molecule, key: str, nbytes: int) -> torch.Tensor:
    ws = molecule.cache.get(key)
    if ws is None or ws.numel() < nbytes:
        ws = torch.empty(max(nbytes, 256), dtype=torch.uint8, device=molecule.ao_planes.device)
        molecule.cache[key] = ws
    return ws


def _require_eri(molecule: Molecule):
    if molecule.rep_tensor is None:
        raise ValueError("this molecule has no ERI (build it with with_eri=True)")


def coulomb(molecule: Molecule, rdm1: Optional[torch.Tensor] = None) -> torch.Tensor:
    """J (n, n) of rdm1[0] + rdm1[1] (Molecule.get_coulomb_potential, molecule.py:108-115)."""
    _require_eri(molecule)
    lib = _lib.load()
    D = (molecule.rdm1 if rdm1 is None else rdm1).contiguous()
    n = molecule.nao
    nbytes = ctypes.c_size_t(0)
    _lib.check(lib.gdx_coulomb_workspace_size(n, ctypes.byref(nbytes)), "gdx_coulomb_workspace_size")
    ws = _workspace(molecule, "coulomb_ws", int(nbytes.value))
    J = torch.empty((n, n), dtype=torch.float64, device=D.device)
    _lib.check(lib.gdx_coulomb_packed(_stream(D), n, molecule.rep_tensor.data_ptr(), D.data_ptr(), J.data_ptr(),
                                      ws.data_ptr()), "gdx_coulomb_packed")
    return J


def exchange(molecule: Molecule, D: torch.Tensor) -> torch.Tensor:
    """K_s[a][c] = sum_bd (ab|cd) D_s[b][d] for D (ns, n, n)."""
    _require_eri(molecule)
    D = D.contiguous()
    ns, n = D.shape[0], D.shape[-1]
    K = torch.empty_like(D)
    _lib.check(_lib.load().gdx_exchange_packed(_stream(D), n, ns, molecule.rep_tensor.data_ptr(), D.data_ptr(),
                                               K.data_ptr()), "gdx_exchange_packed")
    return K


def initial_exchange(molecule: Molecule) -> torch.Tensor:
    """K[D0_s] of the density the molecule was built with (cached): the reference freezes
    its exact-exchange tensor chi at that density (interface/pyscf.py:152-160)."""
    K0 = molecule.cache.get("K0")
    if K0 is None:
        D0 = molecule.cache.get("rdm1_init", molecule.rdm1)
        K0 = exchange(molecule, D0)
        molecule.cache["K0"] = K0
    return K0


def non_xc_energy(molecule: Molecule, rdm1: Optional[torch.Tensor] = None, J: Optional[torch.Tensor] = None,
                  exact_exchange: float = 0.0) -> torch.Tensor:
    """e_nuc + sum D h1e + 1/2 sum D J (+ stale exact exchange when exact_exchange != 0)."""
    D = (molecule.rdm1 if rdm1 is None else rdm1).contiguous()
    if J is None:
        J = coulomb(molecule, D)
    K0 = initial_exchange(molecule) if exact_exchange else None
    out = torch.empty((), dtype=torch.float64, device=D.device)
    _lib.check(_lib.load().gdx_nonxc_energy(_stream(D), molecule.nao, D.data_ptr(), molecule.h1e.data_ptr(),
                                            J.data_ptr(), 0 if K0 is None else K0.data_ptr(), float(exact_exchange),
                                            float(molecule.nuclear_repulsion), out.data_ptr()), "gdx_nonxc_energy")
    return out


def exact_exchange_energy(molecule: Molecule, rdm1: torch.Tensor, a0: float) -> torch.Tensor:
    """-(a0/2) sum_s tr(D_s K0_s)."""
    K0 = initial_exchange(molecule)
    return -0.5 * a0 * (rdm1 * K0).sum()


def fock_assemble(molecule: Molecule, J: torch.Tensor, vxc: torch.Tensor, exact_exchange: float = 0.0) -> torch.Tensor:
    K0 = initial_exchange(molecule) if exact_exchange else None
    n = molecule.nao
    F = torch.empty((2, n, n), dtype=torch.float64, device=J.device)
    _lib.check(_lib.load().gdx_fock_assemble(_stream(J), n, molecule.h1e.data_ptr(), J.data_ptr(),
                                             vxc.contiguous().data_ptr(), 0 if K0 is None else K0.data_ptr(),
                                             float(exact_exchange), F.data_ptr()), "gdx_fock_assemble")
    return F
