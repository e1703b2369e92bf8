"""On-disk molecule store: the role of ``saver`` / ``loader`` (``grad_dft/interface/pyscf.py:331-602``).

The reference writes HDF5 groups per molecule (``save_molecule_data``, ``:584-606``) and
reads them back as ``(name, Molecule)`` pairs, optionally in random order (``loader``,
``:430-520``).  h5py is not part of this stack, so the store is one uncompressed NumPy
``.npz`` archive (read without pickle: every entry is a numeric or fixed-width string
array), one key prefix ``m<i>/`` per molecule plus an optional ``energy`` target per
molecule for training batches.

What is stored: geometry, basis name, integration grid, one-electron matrices, packed ERI
(when present), density and orbitals, the initial density of the SCF, scalars.  AO planes
are not stored: ``load_molecules`` rebuilds the basis from its name and the geometry and
evaluates the planes with the device kernel (``gdx_ao_eval``), which is cheaper than
reading them (benzene's planes are 0.56 GB, its ERI-free archive a few MB).
"""
from __future__ import annotations

import os
from typing import Iterator, Optional, Sequence, Tuple

import numpy as np
import torch

from ..molecule import Grid, Molecule
from .basis import build_basis
from .basis_data import ATOMIC_NUMBER
from .builder import eval_ao_device, pad_to

_SYMBOL = {z: s for s, z in ATOMIC_NUMBER.items()}
FORMAT_VERSION = 1


def _np(t) -> np.ndarray:
    return t.detach().cpu().numpy() if isinstance(t, torch.Tensor) else np.asarray(t)


def molecule_arrays(mol: Molecule, energy: Optional[float] = None) -> dict:
    """The stored fields of one molecule as NumPy arrays (no device work)."""
    if mol.basis is None:
        raise ValueError("molecule has no basis description; it cannot be stored")
    out = {
        "atom_index": _np(mol.atom_index).astype(np.int64),
        "nuclear_pos": _np(mol.nuclear_pos).astype(np.float64),
        "basis": np.array(mol.basis.name),
        "grid_coords": _np(mol.grid.coords).astype(np.float64),
        "grid_weights": _np(mol.grid.weights).astype(np.float64),
        "npts_pad": np.array(mol.npts_pad, dtype=np.int64),
        "nao_pad": np.array(mol.nao_pad, dtype=np.int64),
        "h1e": _np(mol.h1e),
        "s1e": _np(mol.s1e),
        "rdm1": _np(mol.rdm1),
        "mo_coeff": _np(mol.mo_coeff),
        "mo_occ": _np(mol.mo_occ),
        "mo_energy": _np(mol.mo_energy),
        "nuclear_repulsion": np.array(float(mol.nuclear_repulsion)),
        "spin": np.array(int(mol.spin)),
        "charge": np.array(int(mol.charge)),
        "restricted": np.array(bool(mol.restricted)),
        "grid_level": np.array(int(mol.grid_level)),
        "name": np.array(mol.name),
    }
    if mol.rep_tensor is not None:
        out["rep_tensor"] = _np(mol.rep_tensor)
    if "rdm1_init" in mol.cache:
        out["rdm1_init"] = _np(mol.cache["rdm1_init"])
    if energy is not None:
        out["energy"] = np.array(float(energy))
    return out


def save_molecules(path: str, molecules: Sequence[Molecule], energies: Optional[Sequence[float]] = None) -> str:
    """Write molecules (and optional target energies) to one ``.npz`` archive; returns the path."""
    if energies is not None and len(energies) != len(molecules):
        raise ValueError("one energy per molecule")
    payload = {"format_version": np.array(FORMAT_VERSION), "count": np.array(len(molecules))}
    for i, mol in enumerate(molecules):
        arrays = molecule_arrays(mol, None if energies is None else energies[i])
        for k, v in arrays.items():
            payload[f"m{i}/{k}"] = v
    if not path.endswith(".npz"):
        path = path + ".npz"
    np.savez(path, **payload)
    return path


def load_arrays(path: str) -> Tuple[int, list]:
    """Read an archive written by ``save_molecules``: (count, [dict of arrays per molecule])."""
    with np.load(path, allow_pickle=False) as z:
        version = int(z["format_version"])
        if version != FORMAT_VERSION:
            raise ValueError(f"unsupported molecule store version {version}")
        count = int(z["count"])
        mols = [dict() for _ in range(count)]
        for key in z.files:
            if "/" not in key:
                continue
            head, field = key.split("/", 1)
            mols[int(head[1:])][field] = z[key]
    return count, mols


def molecule_from_arrays(a: dict, device: str = "cuda") -> Tuple[Molecule, Optional[float]]:
    """Rebuild a device-resident molecule (AO planes evaluated on the device)."""
    Z = [int(z) for z in a["atom_index"]]
    symbols = [_SYMBOL[z] for z in Z]
    coords = np.asarray(a["nuclear_pos"], dtype=np.float64)
    bs = build_basis(symbols, coords, str(a["basis"]))
    gc = np.asarray(a["grid_coords"], dtype=np.float64)
    gw = np.asarray(a["grid_weights"], dtype=np.float64)
    npts_pad = int(a["npts_pad"])
    nao_pad = int(a["nao_pad"])
    if npts_pad < len(gw) or nao_pad != pad_to(bs.nao, 16):
        raise ValueError("stored padding does not match the grid / basis")
    dev = lambda x: torch.as_tensor(np.ascontiguousarray(x), dtype=torch.float64, device=device)  # noqa: E731
    wpad = np.zeros(npts_pad)
    wpad[: len(gw)] = gw
    grid = Grid(coords=dev(gc), weights=dev(gw), weights_pad=dev(wpad))
    ao = eval_ao_device(bs, gc, npts_pad, nao_pad, device)
    mol = Molecule(
        grid=grid,
        atom_index=torch.as_tensor(np.asarray(Z), device=device),
        nuclear_pos=dev(coords),
        ao_planes=ao,
        rdm1=dev(a["rdm1"]),
        nuclear_repulsion=float(a["nuclear_repulsion"]),
        h1e=dev(a["h1e"]),
        s1e=dev(a["s1e"]),
        mo_coeff=dev(a["mo_coeff"]),
        mo_occ=dev(a["mo_occ"]),
        mo_energy=dev(a["mo_energy"]),
        basis=bs,
        rep_tensor=dev(a["rep_tensor"]) if "rep_tensor" in a else None,
        spin=int(a["spin"]),
        charge=int(a["charge"]),
        restricted=bool(a["restricted"]),
        grid_level=int(a["grid_level"]),
        name=str(a["name"]),
    )
    mol.cache["rdm1_init"] = dev(a["rdm1_init"]) if "rdm1_init" in a else mol.rdm1
    energy = float(a["energy"]) if "energy" in a else None
    return mol, energy


def load_molecules(path: str, device: str = "cuda", randomize: bool = False,
                   seed: int = 0) -> Iterator[Tuple[str, Molecule, Optional[float]]]:
    """Yield (name, molecule, target energy or None) in stored order, or shuffled with
    ``randomize`` (the reference loader's ``randomize``, ``pyscf.py:430-520``)."""
    if not os.path.exists(path) and os.path.exists(path + ".npz"):
        path = path + ".npz"
    count, arrays = load_arrays(path)
    order = np.arange(count)
    if randomize:
        np.random.default_rng(seed).shuffle(order)
    for i in order:
        mol, energy = molecule_from_arrays(arrays[i], device)
        yield mol.name, mol, energy
