"""Build a device-resident :class:`~graddft_amd.molecule.Molecule` from a geometry.

Plays the role of ``molecule_from_pyscf`` (``grad_dft/interface/pyscf.py:105-201``) and
its packaging step ``_package_outputs`` (``:808-864``) without PySCF:

* basis shells from :mod:`.basis`, integration grid from :mod:`.grids` (atomic rules on the
  host, Becke partition weights on the device by default);
* S, T, V and, optionally, the packed ERI on the device by default (``gdx_int1e_eval``,
  ``gdx_eri_packed``), or from the host integral library;
* the initial density from the core-Hamiltonian guess (lowest ``nelec/2`` orbitals of
  ``h1e C = S C e``), split into two equal spin halves exactly as the reference splits an
  RKS density (``pyscf.py:830-838``: ``rdm1 / 2`` per spin, ``mo_occ / 2``);
* AO planes evaluated on the device by ``gdx_ao_eval``.
"""
from __future__ import annotations

import math
from typing import Optional, Sequence

import numpy as np
import scipy.linalg
import torch

from .. import _lib
from ..molecule import Grid, Molecule
from .basis import atomic_numbers, build_basis, nuclear_repulsion, pack_shells
from .grids import molecular_grid
from .integrals import int1e, int1e_device, int2e_packed, int2e_packed_device


def pad_to(x: int, m: int) -> int:
    return (x + m - 1) // m * m


def eval_ao_device(basis, coords: np.ndarray, npts_pad: int, nao_pad: int,
                   device: str = "cuda") -> torch.Tensor:
    """AO planes (5, npts_pad, nao_pad) on the device via the HIP kernel."""
    lib = _lib.load()
    pk = pack_shells(basis)
    t = {k: torch.as_tensor(v, device=device) for k, v in pk.items()}
    c = torch.as_tensor(np.ascontiguousarray(coords, dtype=np.float64), device=device)
    ao = torch.empty((5, npts_pad, nao_pad), dtype=torch.float64, device=device)
    stream = torch.cuda.current_stream(device).cuda_stream
    rc = lib.gdx_ao_eval(stream, len(basis.shells), t["shell_i32"].data_ptr(), t["shell_f64"].data_ptr(),
                         t["prim_f64"].data_ptr(), t["tmat_f64"].data_ptr(), len(coords), npts_pad,
                         c.data_ptr(), basis.nao, nao_pad, ao.data_ptr())
    _lib.check(rc, "gdx_ao_eval")
    return ao


def core_guess(h1e: np.ndarray, s1e: np.ndarray, nocc: int):
    e, C = scipy.linalg.eigh(h1e, s1e)
    occ = np.zeros(len(e))
    occ[:nocc] = 1.0
    D = (C[:, :nocc] @ C[:, :nocc].T)
    return D, C, e, occ


def molecule_from_geometry(symbols: Sequence[str], coords_bohr: np.ndarray, basis: str,
                           grid_level: int = 2, charge: int = 0, with_eri: bool = False,
                           device: str = "cuda", rdm1: Optional[np.ndarray] = None,
                           name: str = "", npts_align: int = 128,
                           omegas: Optional[Sequence[float]] = None, eri_on: str = "device",
                           int1e_on: str = "device", grid_on: str = "device") -> Molecule:
    """Closed-shell molecule with a core-Hamiltonian (or given) density, AOs on device.  With
    `omegas` the chi tensor of that density is built on the device, as ``molecule_from_pyscf``
    does when given omegas (pyscf.py:152-160); B3LYP's exact exchange then uses it.
    with_eri: the packed ERI, computed on the device (eri_on="device", shells up to d) or by the
    host integral library (eri_on="host", and for bases with f shells).  int1e_on / grid_on:
    "device" (S, T, V and the Becke weights by the HIP kernels) or "host"."""
    for option, v in (("eri_on", eri_on), ("int1e_on", int1e_on), ("grid_on", grid_on)):
        if v not in ("host", "device"):
            raise ValueError(f"{option} must be 'host' or 'device', not {v!r}")
    coords_bohr = np.asarray(coords_bohr, dtype=np.float64)
    Z = atomic_numbers(symbols)
    bs = build_basis(symbols, coords_bohr, basis)
    gc, gw = molecular_grid(Z, coords_bohr, grid_level, weights_on=grid_on)
    if int1e_on == "device":
        S_d, T_d, V_d = int1e_device(bs, Z, coords_bohr, device)
        S, h1e = S_d.cpu().numpy(), (T_d + V_d).cpu().numpy()
    else:
        S, T, V = int1e(bs, Z, coords_bohr)
        h1e = T + V
    nelec = sum(Z) - charge
    if nelec % 2:
        raise ValueError("only closed-shell molecules are supported")
    nocc = nelec // 2
    D, C, e, occ = core_guess(h1e, S, nocc)
    if rdm1 is not None:
        D = np.asarray(rdm1, dtype=np.float64)
    n = bs.nao
    N = len(gw)
    npts_pad = pad_to(N, npts_align)
    nao_pad = pad_to(n, 16)
    dev = lambda a: torch.as_tensor(np.ascontiguousarray(a), dtype=torch.float64, device=device)  # noqa
    wpad = np.zeros(npts_pad)
    wpad[:N] = gw
    grid = Grid(coords=dev(gc), weights=dev(gw), weights_pad=dev(wpad))
    ao = eval_ao_device(bs, gc, npts_pad, nao_pad, device)
    eri = None
    if with_eri:
        if eri_on == "device":
            eri = int2e_packed_device(bs, device)
        if eri is None:
            eri = dev(int2e_packed(bs))
    mol = Molecule(
        grid=grid,
        atom_index=torch.as_tensor(Z, device=device),
        nuclear_pos=dev(coords_bohr),
        ao_planes=ao,
        rdm1=dev(np.stack([D, D])),
        nuclear_repulsion=nuclear_repulsion(Z, coords_bohr),
        h1e=dev(h1e),
        s1e=dev(S),
        mo_coeff=dev(np.stack([C, C])),
        mo_occ=dev(np.stack([occ, occ])),
        mo_energy=dev(np.stack([e, e])),
        basis=bs,
        rep_tensor=eri,
        spin=0,
        charge=charge,
        restricted=True,
        grid_level=grid_level,
        name=name,
    )
    mol.cache["rdm1_init"] = mol.rdm1
    if omegas is not None:
        from ..hf import generate_chi_tensor
        om = tuple(float(o) for o in omegas)
        mol = mol.replace(omegas=om, chi_grid=generate_chi_tensor(mol, mol.rdm1, om))
    return mol


def converged_molecule(mol: Molecule, functional=None, round_cycles: int = 10, max_rounds: int = 6,
                       e_tol: float = 1e-10) -> Molecule:
    """The molecule carrying the density, orbitals and orbital energies of a self-consistent
    calculation with `functional` (B3LYP by default, its exact exchange rebuilt from every
    density as a Kohn-Sham code does): the state ``molecule_from_pyscf(mf)`` packages from a
    mean-field object after ``mf.kernel()`` (interface/pyscf.py:105-201, 823-845), where this
    builder's default is the core-Hamiltonian guess.

    Rounds of `round_cycles` CDIIS cycles of the device loop (``diff_scf_loop``, no reverse
    pass), each restarted from the previous round's density with a fresh DIIS history, until
    two rounds' energies agree to `e_tol` Ha (a history of nearly identical error vectors is a
    badly conditioned DIIS system); a round whose result is not finite is discarded.  The
    result has a fresh cache whose initial density (for K0 and chi) is the converged one."""
    import torch
    from ..evaluate import diff_scf_loop
    from ..popular_functionals import B3LYP
    f = (functional or B3LYP)
    if getattr(f, "exact_exchange", 0.0):
        f = f.with_exchange("current")
    loop = diff_scf_loop(f, cycles=round_cycles, check_status=False)
    cur, e_prev = mol, None
    with torch.no_grad():
        for _ in range(max_rounds):
            out = loop(None, cur)
            e = float(out.energy)
            if not (math.isfinite(e) and bool(torch.isfinite(out.rdm1).all())):
                break
            cur = cur.replace(rdm1=out.rdm1.detach().contiguous(), mo_coeff=out.mo_coeff.detach(),
                              mo_energy=out.mo_energy.detach(), mo_occ=out.mo_occ.detach())
            if e_prev is not None and abs(e - e_prev) <= e_tol:
                break
            e_prev = e
    new = mol.replace(rdm1=cur.rdm1, mo_coeff=cur.mo_coeff, mo_energy=cur.mo_energy, mo_occ=cur.mo_occ)
    new.cache = {"rdm1_init": new.rdm1}
    # the rounds' exchange-ordered ERI copy (4x the packed ERI) lives in the input's cache
    from ..scf import release_exchange_xo
    release_exchange_xo(mol)
    return new
