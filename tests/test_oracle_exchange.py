"""The oracle's exact exchange rebuilt from every density (diff_scf_loop exchange_update, the
"current" exchange of the device loop), pinned to Hartree-Fock: with no grid XC and a0 = 1 the
loop's energy -(1/2) sum_s tr(D_s K[D_s]) + h + J terms is the RHF energy, so its fixed point
is the textbook water STO-3G RHF energy (-74.963 Ha; the independent RHF iteration of
test_oracle_inputs reaches it too).  The frozen exchange of the initial density misses it."""
import math

import numpy as np
import pytest


def _water():
    from graddft_amd.interface.basis import build_basis
    R, th = 1.809, math.radians(104.52)
    xyz = np.array([[0, 0, 0], [R, 0, 0], [R * math.cos(th), R * math.sin(th), 0]])
    return xyz, build_basis(["O", "H", "H"], xyz, "sto-3g")


def test_current_exchange_loop_reaches_the_rhf_energy():
    from graddft_amd.interface.basis import nuclear_repulsion
    from graddft_amd.interface.builder import core_guess
    from graddft_amd.interface.integrals import int1e, int2e_packed, unpack_eri
    from oracle import scf as O
    xyz, bs = _water()
    Z = [8, 1, 1]
    S, T, V = int1e(bs, Z, xyz)
    eri = unpack_eri(int2e_packed(bs), bs.nao)
    D, _, _, occ = core_guess(T + V, S, 5)
    D0, occ0 = np.stack([D, D]), np.stack([occ, occ])
    no_xc = lambda d: (0.0, np.zeros_like(d))  # noqa: E731
    kx = lambda d: np.stack([O.exchange_potential(d[s], eri) for s in range(2)])  # noqa: E731
    inp = O.SCFInputs(T + V, S, eri, nuclear_repulsion(Z, xyz), kx(D0), 1.0)
    cur = O.diff_scf_loop(no_xc, inp, D0, occ0, 30, exchange_update=kx)
    assert cur.energy == pytest.approx(-74.963, abs=1e-3)
    # converged: the last cycles agree, and the total is the RHF expression at that density
    assert abs(cur.energies[-1] - cur.energies[-2]) <= 1e-9
    Dt = cur.rdm1.sum(0)
    J = O.coulomb_potential(Dt, eri)
    e_rhf = (np.sum(Dt * (T + V)) + 0.5 * np.sum(Dt * J)
             - 0.5 * sum(np.sum(cur.rdm1[s] * O.exchange_potential(cur.rdm1[s], eri)) for s in range(2))
             + nuclear_repulsion(Z, xyz))
    assert abs(cur.energy - e_rhf) <= 1e-10
    stale = O.diff_scf_loop(no_xc, inp, D0, occ0, 30)
    assert abs(stale.energy - cur.energy) > 1e-2
