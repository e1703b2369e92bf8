"""SCF-side terms on the device: Coulomb matrix, non-XC energy, exchange of the initial
density (the stale exact-exchange term of B3LYP), Fock assembly.

Reference: ``grad_dft/molecule.py:694-811`` (nonXC, coulomb_potential),
``grad_dft/train.py:148-215`` (Fock assembly), ``molecule.py:507-613`` (exact exchange from
the initial density; here from the packed ERI instead of the semi-numerical chi tensor).

The ``*Fn`` classes are the same device computations as ``torch.autograd.Function``s, for
reverse mode through an SCF loop: J is linear and self-adjoint (the ERI is symmetric under
pair exchange), the non-XC energy's adjoint is h + J/2 per spin plus the J cotangent, and
the Fock assembly's adjoint applies the clip masks and the symmetrisation in reverse.
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


# Largest exchange-ordered ERI copy built per molecule (bytes); it is also kept under a
# quarter of the device memory free when it is built.  Larger molecules use the gather kernel.
XO_MAX_BYTES = 16 << 30


def _exchange_xo(molecule: Molecule, n: int, device):
    """(xo, workspace) of the molecule's exchange-ordered ERI copy, built on first use and
    kept in its cache; None when it would exceed the budget or cannot be built now (under
    graph capture, which must not allocate)."""
    hit = molecule.cache.get("eri_xo")
    if hit is not None and hit[0] is molecule.rep_tensor:
        return hit[1], hit[2]
    if torch.cuda.is_current_stream_capturing():
        return None
    lib = _lib.load()
    nbytes, wbytes = ctypes.c_size_t(0), ctypes.c_size_t(0)
    _lib.check(lib.gdx_exchange_xo_size(n, ctypes.byref(nbytes)), "gdx_exchange_xo_size")
    _lib.check(lib.gdx_exchange_xo_workspace_size(n, ctypes.byref(wbytes)), "gdx_exchange_xo_workspace_size")
    free, _ = torch.cuda.mem_get_info(device)
    if nbytes.value > XO_MAX_BYTES or nbytes.value + wbytes.value > free // 4:
        return None
    xo = torch.empty(nbytes.value // 8, dtype=torch.float64, device=device)
    ws = torch.empty(wbytes.value // 8, dtype=torch.float64, device=device)
    _lib.check(lib.gdx_exchange_xo_build(_stream(xo), n, molecule.rep_tensor.data_ptr(), xo.data_ptr()),
               "gdx_exchange_xo_build")
    molecule.cache["eri_xo"] = (molecule.rep_tensor, xo, ws)
    return xo, ws


def release_exchange_xo(molecule: Molecule) -> None:
    """Drop the molecule's exchange-ordered ERI copy (4x the packed ERI)."""
    molecule.cache.pop("eri_xo", None)


def exchange(molecule: Molecule, D: torch.Tensor) -> torch.Tensor:
    """K_s[a][c] = sum_bd (ab|cd) D_s[b][d] for D (ns, n, n): from the exchange-ordered ERI
    copy (a symmetric product read at streaming rate; built once per molecule) when it fits the
    budget, else gathered from the packed ERI."""
    _require_eri(molecule)
    D = D.contiguous()
    ns, n = D.shape[0], D.shape[-1]
    K = torch.empty_like(D)
    xo = _exchange_xo(molecule, n, D.device)
    if xo is not None:
        _lib.check(_lib.load().gdx_exchange_xo(_stream(D), n, ns, xo[0].data_ptr(), D.data_ptr(), K.data_ptr(),
                                               xo[1].data_ptr()), "gdx_exchange_xo")
        return K
    _lib.check(_lib.load().gdx_exchange_packed(_stream(D), n, ns, molecule.rep_tensor.data_ptr(), D.data_ptr(),
                                               K.data_ptr()), "gdx_exchange_packed")
    return K


def initial_exchange(molecule: Molecule) -> torch.Tensor:
    """Exchange matrix of the density the molecule was built with (cached), K0 (2, n, n).

    With a chi tensor at omega = 0 on the molecule this is the reference's semi-numerical term:
    K~_s[a][b] = sum_r w_r phi_a(r) chi[r, 0, s, b] (``HF_energy_density`` /
    ``HF_density_grad_2_Fock`` with B3LYP's constant coefficient, molecule.py:507-613); otherwise
    the exchange of the initial density from the packed ERI, K[D0_s], which the grid form
    approximates to the quadrature error.  Both are frozen at the initial density, as the
    reference's chi (interface/pyscf.py:152-160)."""
    if molecule.chi_grid is not None and molecule.omegas is not None and 0.0 in molecule.omegas:
        hit = molecule.cache.get("K0_chi")
        if hit is not None and hit[0] is molecule.chi_grid:
            return hit[1]
        from .hf import exchange_matrix
        K0 = exchange_matrix(molecule)
        molecule.cache["K0_chi"] = (molecule.chi_grid, K0)
        return K0
    K0 = molecule.cache.get("K0")
    if K0 is None:
        D0 = molecule.cache.get("rdm1_init", molecule.rdm1)
        had_xo = "eri_xo" in molecule.cache
        K0 = exchange(molecule, D0)
        molecule.cache["K0"] = K0
        if not had_xo:
            # one K: do not keep an exchange-ordered copy (4x the packed ERI) built for it
            release_exchange_xo(molecule)
    return K0


def non_xc_energy(molecule: Molecule, rdm1: Optional[torch.Tensor] = None, J: Optional[torch.Tensor] = None,
                  exact_exchange: float = 0.0, K: Optional[torch.Tensor] = None) -> torch.Tensor:
    """e_nuc + sum D h1e + 1/2 sum D J - (a0/2) sum_s tr(D_s K_s) when exact_exchange = a0 != 0,
    K the exchange matrices given, else those of the initial density (initial_exchange)."""
    D = (molecule.rdm1 if rdm1 is None else rdm1).contiguous()
    if J is None:
        J = coulomb(molecule, D)
    K0 = (K.contiguous() if K is not None else initial_exchange(molecule)) if exact_exchange else None
    out = torch.empty((), dtype=torch.float64, device=D.device)
    _lib.check(_lib.load().gdx_nonxc_energy(_stream(D), molecule.nao, D.data_ptr(), molecule.h1e.data_ptr(),
                                            J.data_ptr(), 0 if K0 is None else K0.data_ptr(), float(exact_exchange),
                                            float(molecule.nuclear_repulsion), out.data_ptr()), "gdx_nonxc_energy")
    return out


def exact_exchange_energy(molecule: Molecule, rdm1: torch.Tensor, a0: float,
                          K: Optional[torch.Tensor] = None) -> torch.Tensor:
    """-(a0/2) sum_s tr(D_s K_s), K the given exchange matrices or K0 of the initial density."""
    K0 = initial_exchange(molecule) if K is None else K
    return -0.5 * a0 * (rdm1 * K0).sum()


def fock_assemble(molecule: Molecule, J: torch.Tensor, vxc: torch.Tensor, exact_exchange: float = 0.0,
                  K: Optional[torch.Tensor] = None) -> torch.Tensor:
    """h + J + vxc - a0 K (symmetrised, clipped), K the given exchange matrices or K0."""
    K0 = (K.contiguous() if K is not None else initial_exchange(molecule)) if exact_exchange else None
    n = molecule.nao
    F = torch.empty((2, n, n), dtype=torch.float64, device=J.device)
    _lib.check(_lib.load().gdx_fock_assemble(_stream(J), n, molecule.h1e.data_ptr(), J.data_ptr(),
                                             vxc.contiguous().data_ptr(), 0 if K0 is None else K0.data_ptr(),
                                             float(exact_exchange), F.data_ptr()), "gdx_fock_assemble")
    return F


CLIP = 1e-30


class CoulombFn(torch.autograd.Function):
    """J(D_up + D_dn); cotangent Jbar -> dD_s = J[Jbar] for both spins."""

    @staticmethod
    def forward(ctx, rdm1, molecule: Molecule):
        ctx.molecule = molecule
        return coulomb(molecule, rdm1)

    @staticmethod
    @torch.autograd.function.once_differentiable
    def backward(ctx, g):
        g = g.detach().contiguous()
        src = torch.stack([g, torch.zeros_like(g)])
        dD = coulomb(ctx.molecule, src)
        return torch.stack([dD, dD]), None


class ExchangeFn(torch.autograd.Function):
    """K[D] from the packed ERI; reverse Dbar = K[Kbar] (the exchange operator is self-adjoint
    under the Frobenius product, by the (ab|cd) = (ba|dc) symmetry)."""

    @staticmethod
    def forward(ctx, rdm1, molecule: Molecule):
        ctx.molecule = molecule
        return exchange(molecule, rdm1.detach())

    @staticmethod
    @torch.autograd.function.once_differentiable
    def backward(ctx, g):
        return exchange(ctx.molecule, g.contiguous()), None


class NonXCFn(torch.autograd.Function):
    """non_xc_energy(D, J, K) with J and K inputs: dE/dD_s = h + J/2 - (a0/2) K_s,
    dE/dJ = D_t/2, dE/dK_s = -(a0/2) D_s.  K = None: the initial density's K0, a constant."""

    @staticmethod
    def forward(ctx, rdm1, J, K, molecule: Molecule, exact_exchange: float):
        ctx.molecule = molecule
        ctx.a0 = exact_exchange
        ctx.has_k = K is not None
        ctx.save_for_backward(rdm1, J, K)
        return non_xc_energy(molecule, rdm1, J, exact_exchange=exact_exchange, K=K)

    @staticmethod
    @torch.autograd.function.once_differentiable
    def backward(ctx, g):
        rdm1, J, K = ctx.saved_tensors
        m = ctx.molecule
        dD = (m.h1e + 0.5 * J).expand(2, -1, -1).clone()
        dK = None
        if ctx.a0:
            dD = dD - 0.5 * ctx.a0 * (K if ctx.has_k else initial_exchange(m))
            if ctx.has_k:
                dK = g * (-0.5 * ctx.a0) * rdm1
        dJ = 0.5 * rdm1.sum(0)
        return g * dD, g * dJ, dK, None, None


class FockAssembleFn(torch.autograd.Function):
    """fock_assemble(J, vxc) with the reverse of its clips and symmetrisation."""

    @staticmethod
    def forward(ctx, J, vxc, K, molecule: Molecule, exact_exchange: float):
        ctx.molecule = molecule
        ctx.a0 = exact_exchange
        ctx.has_k = K is not None
        ctx.save_for_backward(J, vxc, K)
        return fock_assemble(molecule, J, vxc, exact_exchange=exact_exchange, K=K)

    @staticmethod
    @torch.autograd.function.once_differentiable
    def backward(ctx, g):
        J, vxc, K = ctx.saved_tensors
        m = ctx.molecule
        gK = None
        x1 = m.h1e + J + vxc
        m1 = x1.abs() > CLIP
        y1 = torch.where(m1, x1, torch.zeros_like(x1))
        x2 = 0.5 * (y1 + y1.transpose(1, 2))
        m2 = x2.abs() > CLIP
        y = torch.where(m2, x2, torch.zeros_like(x2))
        if ctx.a0:
            X = -0.5 * ctx.a0 * (K if ctx.has_k else initial_exchange(m))
            x3 = y + X + X.transpose(1, 2)
            m3 = x3.abs() > CLIP
            g = torch.where(m3, g, torch.zeros_like(g))
            y = torch.where(m3, x3, torch.zeros_like(x3))
        g = torch.where(y.abs() > CLIP, g, torch.zeros_like(g))
        if ctx.a0 and ctx.has_k:   # x3 holds X + X^T, X = -(a0/2) K
            gK = -0.5 * ctx.a0 * (g + g.transpose(1, 2))
        g = torch.where(m2, g, torch.zeros_like(g))
        g = 0.5 * (g + g.transpose(1, 2))
        g = torch.where(m1, g, torch.zeros_like(g))
        return g.sum(0), g, gK, None, None
