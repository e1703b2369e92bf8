"""Semi-numerical exact exchange on the device: the ``chi`` tensor, the Hartree-Fock energy
density and its Fock term.

Reference (``grad_dft``):

* ``generate_chi_tensor`` (``interface/pyscf.py:1067-1126``) with the nu integrals of
  ``external/_hf_density.py:18-103``: chi[r, omega, s, a] = sum_bd D0_s[b][d] phi_b(r) nu_da(r);
* ``Molecule.select_HF_omegas`` / ``HF_energy_density`` (``molecule.py:153-204, 507-541``);
* ``HF_density_grad_2_Fock`` (``molecule.py:206-250, 545-613``).

Device layout: ``chi_grid (nomega, 2, nao, npts_pad)`` float64, grid points contiguous (the
reference's ``(N, nomega, 2, nao)`` is ``Molecule.chi``, a permuted view).  The nu integrals
never exist in memory: ``gdx_hf_chi`` evaluates them per grid point and primitive pair and
contracts them with u = D0^T phi on the fly.

For a functional whose coefficient of e_HF is a constant a0 (B3LYP), the energy and Fock terms
reduce to one matrix per spin, K~_s[a][b] = sum_r w_r phi_a(r) chi[r, s, b] (the quadrature of
the exchange matrix of the initial density): E_HF = -(a0/2) sum_s <D_s, K~_s> and the Fock term
V + V^T with V_s = -(a0/2) K~_s.  That is exactly what the reference computes through
``HF_energy_density`` and ``HF_density_grad_2_Fock`` with gr = a0 w (``popular_functionals.py:
306-372``).
"""
from __future__ import annotations

import ctypes
from typing import Optional, Sequence

import numpy as np
import torch

from . import _lib


class HFPlan:
    """Device plan of the nu integrals for one basis (``gdx_hf_plan_create``)."""

    def __init__(self, basis, screen: float = 1e-16):
        from .interface.basis import pack_shells
        pk = pack_shells(basis)
        self._keep = pk
        self.nao = basis.nao
        handle = ctypes.c_void_p(0)
        ptr = lambda a: a.ctypes.data_as(ctypes.c_void_p)  # noqa: E731
        _lib.check(_lib.load().gdx_hf_plan_create(len(basis.shells), ptr(pk["shell_i32"]), ptr(pk["shell_f64"]),
                                                  ptr(pk["prim_f64"]), ptr(pk["tmat_f64"]), basis.nao,
                                                  float(screen), ctypes.byref(handle)), "gdx_hf_plan_create")
        self.handle = handle

    def info(self):
        ne, nm = ctypes.c_int64(0), ctypes.c_int64(0)
        _lib.check(_lib.load().gdx_hf_plan_info(self.handle, ctypes.byref(ne), ctypes.byref(nm)), "gdx_hf_plan_info")
        return int(ne.value), int(nm.value)

    def __del__(self):
        h = getattr(self, "handle", None)
        if h is not None and h.value and _lib._lib is not None:
            _lib._lib.gdx_hf_plan_destroy(h)
            self.handle = None


def hf_plan(molecule) -> HFPlan:
    p = molecule.cache.get("hf_plan")
    if p is None:
        p = HFPlan(molecule.basis)
        molecule.cache["hf_plan"] = p
    return p


def _stream(t: torch.Tensor) -> int:
    return torch.cuda.current_stream(t.device).cuda_stream


def _phi(molecule) -> torch.Tensor:
    return molecule.ao_planes[0]


def ao_matmul(molecule, M: torch.Tensor, npts: Optional[int] = None) -> torch.Tensor:
    """out (nspin, nao, npts_pad) = (phi M_s)^T over the first npts grid points."""
    M = M.detach().to(torch.float64).contiguous()
    ns, n = int(M.shape[0]), molecule.nao
    N = molecule.grid_size if npts is None else int(npts)
    npad = molecule.npts_pad if npts is None else (N + 63) // 64 * 64
    out = torch.empty((ns, n, npad), dtype=torch.float64, device=M.device)
    _lib.check(_lib.load().gdx_ao_matmul(_stream(M), N, npad, n, molecule.nao_pad, _phi(molecule).data_ptr(), ns,
                                         M.data_ptr(), out.data_ptr()), "gdx_ao_matmul")
    return out


def chi_from_u(molecule, u: torch.Tensor, omega: float, coords: Optional[torch.Tensor] = None) -> torch.Tensor:
    """chi (nspin, nao, npts_pad) = sum_a u_a nu_ab(r; omega) at the grid points (or `coords`)."""
    c = molecule.grid.coords if coords is None else coords
    c = c.detach().to(torch.float64).contiguous()
    ns, n, npad = (int(x) for x in u.shape)
    chi = torch.empty_like(u)
    _lib.check(_lib.load().gdx_hf_chi(_stream(u), hf_plan(molecule).handle, int(c.shape[0]), npad, c.data_ptr(), ns,
                                      u.contiguous().data_ptr(), float(omega), chi.data_ptr()), "gdx_hf_chi")
    return chi


def generate_chi_tensor(molecule, rdm1: torch.Tensor, omegas: Sequence[float]) -> torch.Tensor:
    """chi_grid (nomega, 2, nao, npts_pad) of ``generate_chi_tensor`` (interface/pyscf.py:1067-1126)
    at the density rdm1 (2, n, n).  Equal spin blocks are evaluated once."""
    D = rdm1.detach().to(torch.float64).contiguous()
    same = bool(torch.equal(D[0], D[1]))
    u = ao_matmul(molecule, D[:1] if same else D)
    out = torch.empty((len(omegas), 2, molecule.nao, molecule.npts_pad), dtype=torch.float64, device=D.device)
    for i, w in enumerate(omegas):
        c = chi_from_u(molecule, u, float(w))
        out[i] = c.expand(2, -1, -1) if same else c
    return out


def select_HF_omegas(molecule, omegas: Sequence[float]) -> torch.Tensor:  # noqa: N802 (reference name)
    """Rows of chi_grid for the requested omegas (molecule.py:153-183), same errors."""
    if molecule.chi_grid is None:
        raise ValueError("Precomputed chi tensor has not been loaded.")
    have = list(molecule.omegas)
    for o in omegas:
        if o not in have:
            raise ValueError(f"The molecule.chi tensor does not contain omega value {o}, only {have}")
    return torch.stack([molecule.chi_grid[have.index(o)] for o in omegas])


def hf_energy_density(molecule, omegas: Sequence[float], rdm1: Optional[torch.Tensor] = None) -> torch.Tensor:
    """e_HF (nomega, 2, N) at the molecule's current density (molecule.py:507-541)."""
    chi = select_HF_omegas(molecule, omegas)
    D = (molecule.rdm1 if rdm1 is None else rdm1).detach().to(torch.float64).contiguous()
    v = ao_matmul(molecule, D)
    n, N, npad = molecule.nao, molecule.grid_size, molecule.npts_pad
    out = torch.empty((len(omegas), 2, npad), dtype=torch.float64, device=D.device)
    for i in range(len(omegas)):
        _lib.check(_lib.load().gdx_hf_energy_density(_stream(D), N, npad, n, 2, chi[i].contiguous().data_ptr(),
                                                     v.data_ptr(), out[i].data_ptr()), "gdx_hf_energy_density")
    return out[:, :, :N]


def _fock_ws(molecule, nspin: int) -> torch.Tensor:
    nb = ctypes.c_size_t(0)
    _lib.check(_lib.load().gdx_hf_fock_workspace_size(molecule.grid_size, molecule.nao, nspin, ctypes.byref(nb)),
               "gdx_hf_fock_workspace_size")
    key = ("hf_fock_ws", nspin)
    ws = molecule.cache.get(key)
    if ws is None or ws.numel() < nb.value:
        ws = torch.empty(int(nb.value), dtype=torch.uint8, device=molecule.ao_planes.device)
        molecule.cache[key] = ws
    return ws


def hf_fock_term(molecule, chi_s: torch.Tensor, g: torch.Tensor, scale: float) -> torch.Tensor:
    """out (nspin, n, n) = scale sum_r g_s(r) phi_a(r) chi_s[b](r); g (npts_pad,) shared by the
    spins or (nspin, npts_pad)."""
    chi_s = chi_s.contiguous()
    ns, n, npad = (int(x) for x in chi_s.shape)
    g = g.detach().to(torch.float64).contiguous()
    stride = 0 if g.dim() == 1 else int(g.shape[-1])
    out = torch.empty((ns, n, n), dtype=torch.float64, device=chi_s.device)
    _lib.check(_lib.load().gdx_hf_fock(_stream(chi_s), molecule.grid_size, npad, n, molecule.nao_pad,
                                       _phi(molecule).data_ptr(), ns, chi_s.data_ptr(), g.data_ptr(), stride,
                                       float(scale), out.data_ptr(), _fock_ws(molecule, ns).data_ptr()),
               "gdx_hf_fock")
    return out


def hf_density_grad_2_fock(molecule, gr: torch.Tensor, omegas: Sequence[float]) -> torch.Tensor:
    """V (nomega, 2, n, n) = -1/2 sum_r gr[omega, s, r] phi_a(r) chi[r, omega, s, b]
    (molecule.py:545-613) for gr (nomega, 2, N) = dE_xc / d e_HF."""
    chi = select_HF_omegas(molecule, omegas)
    N, npad = molecule.grid_size, molecule.npts_pad
    g = torch.zeros((len(omegas), 2, npad), dtype=torch.float64, device=chi.device)
    g[:, :, :N] = gr.detach().to(torch.float64)
    return torch.stack([hf_fock_term(molecule, chi[i], g[i], -0.5) for i in range(len(omegas))])


def exchange_matrix(molecule) -> torch.Tensor:
    """K~ (2, n, n) = sum_r w_r phi_a(r) chi[r, omega = 0, s, b]: the semi-numerical exchange
    matrix of the density chi was built from (B3LYP's constant-coefficient HF term)."""
    chi = select_HF_omegas(molecule, [0.0])[0]
    if torch.equal(chi[0], chi[1]):
        k = hf_fock_term(molecule, chi[:1], molecule.grid.weights_pad, 1.0)
        return k.expand(2, -1, -1).contiguous()
    return hf_fock_term(molecule, chi, molecule.grid.weights_pad, 1.0)
