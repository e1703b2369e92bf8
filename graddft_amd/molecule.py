"""``Grid`` and ``Molecule``: the data the XC grid path runs on, resident in device memory.

Mirrors the pytrees of ``grad_dft/molecule.py:34-336`` (field names and meanings kept).
The differences are in layout, chosen for the device kernels:

* AO values are stored as five zero-padded planes ``ao_planes (5, npts_pad, nao_pad)``
  (value, d/dx, d/dy, d/dz, laplacian) instead of ``ao (N,n)``, ``grad_ao (N,n,3)``
  and ``grad_n_ao[2] (N,n,3)``; the reference's arrays are exposed as views.
* ``rep_tensor`` is the 8-fold packed ERI (length P(P+1)/2, P = n(n+1)/2) rather than the
  dense (n,n,n,n) tensor.
* ``restricted`` records that both spin blocks of ``rdm1`` are equal, which lets the
  density kernel form them once.
"""
from __future__ import annotations

import dataclasses
from dataclasses import dataclass, field
from typing import Any, Optional

import torch


@dataclass
class Grid:
    """Integration grid (``grad_dft/molecule.py:34-69``); tensors on the device."""
    coords: torch.Tensor           # (N, 3) bohr
    weights: torch.Tensor          # (N,)
    weights_pad: torch.Tensor      # (npts_pad,), zero past N

    def __len__(self) -> int:
        return int(self.weights.shape[0])

    def integrate(self, vals: torch.Tensor, axis: int = 0) -> torch.Tensor:
        """Grid quadrature, as ``Grid.integrate`` (``molecule.py:45-69``)."""
        return torch.tensordot(self.weights, vals, dims=([0], [axis]))


@dataclass
class Molecule:
    """A molecule on a grid (``grad_dft/molecule.py:72-102``)."""
    grid: Grid
    atom_index: torch.Tensor       # (natm,) atomic numbers
    nuclear_pos: torch.Tensor      # (natm, 3) bohr
    ao_planes: torch.Tensor        # (5, npts_pad, nao_pad)
    rdm1: torch.Tensor             # (2, n, n)
    nuclear_repulsion: float
    h1e: torch.Tensor              # (n, n)
    s1e: torch.Tensor              # (n, n)
    mo_coeff: torch.Tensor         # (2, n, n)
    mo_occ: torch.Tensor           # (2, n)
    mo_energy: torch.Tensor        # (2, n)
    basis: Any = None              # graddft_amd.interface.basis.BasisSet (host)
    rep_tensor: Optional[torch.Tensor] = None  # packed ERI
    fock: Optional[torch.Tensor] = None
    energy: Optional[torch.Tensor] = None
    spin: int = 0
    charge: int = 0
    restricted: bool = True
    grid_level: int = 2
    name: str = ""
    omegas: Optional[tuple] = None          # range-separation parameters chi was built for
    chi_grid: Optional[torch.Tensor] = None  # (nomega, 2, n, npts_pad): chi, grid points contiguous
    cache: dict = field(default_factory=dict, repr=False, compare=False)

    # -- shapes ---------------------------------------------------------------------
    @property
    def nao(self) -> int:
        return int(self.rdm1.shape[-1])

    @property
    def grid_size(self) -> int:
        return len(self.grid)

    @property
    def npts_pad(self) -> int:
        return int(self.ao_planes.shape[1])

    @property
    def nao_pad(self) -> int:
        return int(self.ao_planes.shape[2])

    # -- reference-layout views of the AO planes -----------------------------------
    @property
    def ao(self) -> torch.Tensor:
        return self.ao_planes[0, : self.grid_size, : self.nao]

    @property
    def grad_ao(self) -> torch.Tensor:
        return self.ao_planes[1:4, : self.grid_size, : self.nao].permute(1, 2, 0)

    @property
    def lapl_ao(self) -> torch.Tensor:
        return self.ao_planes[4, : self.grid_size, : self.nao]

    @property
    def chi(self) -> Optional[torch.Tensor]:
        """chi in the reference's layout (N, nomega, 2, n) (molecule.py:82-102), a view."""
        if self.chi_grid is None:
            return None
        return self.chi_grid[..., : self.grid_size].permute(3, 0, 1, 2)

    def replace(self, **kw) -> "Molecule":
        """Functional update, as flax ``struct.dataclass.replace``.  The per-molecule
        cache (device workspaces) is shared with the new instance."""
        new = dataclasses.replace(self, **kw)
        new.cache = self.cache
        return new

    def nelec(self) -> torch.Tensor:
        return self.mo_occ.sum(dim=1)

    # -- grid densities and energy terms (molecule.py:108-151, 298-300) ------------------
    # Values from the device kernels (the density kernel of the XC path, the packed-ERI
    # Coulomb kernel); not on the autograd tape -- differentiable energies go through
    # Functional.energy / energy_predictor.
    def _grid_densities(self) -> torch.Tensor:
        from .popular_functionals import B3LYP_GRID
        with torch.no_grad():
            return B3LYP_GRID.plan(self).densities(self.rdm1.detach())

    def density(self) -> torch.Tensor:
        """rho per spin at each grid point, (N, 2)."""
        return self._grid_densities()[:, :, 0]

    def grad_density(self) -> torch.Tensor:
        """grad rho per spin, (N, 2, 3)."""
        return self._grid_densities()[:, :, 1:4]

    def kinetic_density(self) -> torch.Tensor:
        """tau = 1/2 sum_ab D_ab grad phi_a . grad phi_b per spin, (N, 2)."""
        return self._grid_densities()[:, :, 4]

    def lapl_density(self) -> torch.Tensor:
        """Laplacian of rho per spin, (N, 2)."""
        return self._grid_densities()[:, :, 5]

    def get_coulomb_potential(self) -> torch.Tensor:
        """J of the total density from the packed ERI, (n, n)."""
        from .scf import coulomb
        with torch.no_grad():
            return coulomb(self, self.rdm1.detach())

    def nonXC(self) -> torch.Tensor:  # noqa: N802  (reference name)
        """E_nuc + tr(D h) + 1/2 tr(D J) of the total density."""
        from .scf import non_xc_energy
        with torch.no_grad():
            return non_xc_energy(self, self.rdm1.detach())

    # -- semi-numerical exact exchange (molecule.py:153-250) --------------------------------
    def select_HF_omegas(self, omegas) -> torch.Tensor:  # noqa: N802  (reference name)
        """chi rows for `omegas`, device layout (nomega, 2, n, npts_pad); ValueError when chi is
        missing or lacks an omega, as the reference."""
        from .hf import select_HF_omegas
        return select_HF_omegas(self, omegas)

    def HF_energy_density(self, omegas) -> torch.Tensor:  # noqa: N802
        """e_HF (nomega, 2, N) at the current rdm1 with the stored chi."""
        from .hf import hf_energy_density
        with torch.no_grad():
            return hf_energy_density(self, omegas)

    def HF_density_grad_2_Fock(self, functional=None, params=None, omegas=(0.0,), ehf=None,  # noqa: N802
                               coefficient_inputs=None, densities_wout_hf=None, gr=None) -> torch.Tensor:
        """Fock contribution of the e_HF dependence, (nomega, 2, n, n):
        V = -1/2 sum_r gr phi_a chi_b with gr = dE_xc / d e_HF.  The reference obtains gr by
        differentiating xc_energy; here gr is passed, or, for a functional whose e_HF
        coefficient is the constant a0 = functional.exact_exchange (B3LYP), it is a0 w_r at
        omega = 0 (and 0 for other omegas)."""
        from .hf import hf_density_grad_2_fock
        if gr is None:
            a0 = float(getattr(functional, "exact_exchange", 0.0))
            w = self.grid.weights.to(torch.float64)
            gr = torch.stack([torch.stack([a0 * w if o == 0.0 else 0.0 * w] * 2) for o in omegas])
        with torch.no_grad():
            return hf_density_grad_2_fock(self, gr, omegas)

    # -- orbital helpers (molecule.py:302-330) ------------------------------------------
    def make_rdm1(self) -> torch.Tensor:
        return make_rdm1(self.mo_coeff, self.mo_occ)

    def get_occ(self) -> torch.Tensor:
        nelecs = [int(v) for v in self.mo_occ.sum(dim=1).tolist()]
        return get_occ(self.mo_energy, nelecs)

    def get_mo_grads(self) -> torch.Tensor:
        return orbital_grad(self.mo_coeff, self.mo_occ, self.fock)


def make_rdm1(mo_coeff: torch.Tensor, mo_occ: torch.Tensor) -> torch.Tensor:
    """D_s = C_s diag(occ_s) C_s^T (``molecule.py:816-846``)."""
    return torch.einsum("sij,sj,skj->sik", mo_coeff, mo_occ, mo_coeff)


def get_occ(mo_energy: torch.Tensor, nelecs) -> torch.Tensor:
    """Aufbau occupations (``molecule.py:851-889``): per spin, the nelec_s orbitals of lowest
    energy (ties in index order) get 1, the others 0.  nelecs is truncated to integers, as the
    reference's int64 cast of the occupation sums."""
    occ = torch.zeros_like(mo_energy)
    order = torch.argsort(mo_energy, dim=1, stable=True)
    for s in range(2):
        occ[s, order[s, : int(nelecs[s])]] = 1.0
    return occ


def orbital_grad(mo_coeff: torch.Tensor, mo_occ: torch.Tensor, fock: torch.Tensor) -> torch.Tensor:
    """Orbital gradient (``molecule.py:344-373``): sum over spins of C_vir^T F C_occ, where
    C_occ keeps the columns with occ > 0 and C_vir those with occ == 0 (others zeroed);
    an (n, n) matrix."""
    c_occ = torch.where(mo_occ[:, None, :] > 0, mo_coeff, torch.zeros_like(mo_coeff))
    c_vir = torch.where(mo_occ[:, None, :] == 0, mo_coeff, torch.zeros_like(mo_coeff))
    return torch.einsum("sab,sac,scd->bd", c_vir, fock, c_occ)
