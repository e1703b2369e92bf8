"""Writes tests/golden/*.npz: FP64 oracle outputs on small seeded inputs.

The reference (GradDFT) cannot run here (jax/flax/pyscf are not installed), so these are the
oracle's own outputs, pinned independently by the checks in tests/test_oracle_*.py; they
guard the oracle against regressions and give the device path fixed targets.
Run from the repository root:  python tests/golden/make_golden.py
"""
import math
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)

from graddft_amd.interface import geometry  # noqa: E402
from graddft_amd.interface.basis import atomic_numbers, build_basis, nuclear_repulsion  # noqa: E402
from graddft_amd.interface.builder import core_guess  # noqa: E402
from graddft_amd.interface.grids import molecular_grid  # noqa: E402
from graddft_amd.interface.integrals import int1e, int2e_packed, unpack_eri  # noqa: E402
from oracle import scf as O  # noqa: E402
from oracle.ao import eval_ao  # noqa: E402
from oracle.xc import MLPSpec, init_mlp_params, to_inputs, xc_value_and_grads  # noqa: E402

HERE = os.path.dirname(os.path.abspath(__file__))


def water_case(seed=7, basis="def2-svp", level=1):
    sym, xyz = geometry.water(np.random.default_rng(seed))
    Z = atomic_numbers(sym)
    bs = build_basis(sym, xyz, basis)
    S, T, V = int1e(bs, Z, xyz)
    D, C, e, occ = core_guess(T + V, S, sum(Z) // 2)
    c, w = molecular_grid(Z, xyz, level)
    return sym, xyz, Z, bs, S, T, V, D, C, occ, c, w


def main():
    sym, xyz, Z, bs, S, T, V, D, C, occ, c, w = water_case()
    xin = to_inputs(*eval_ao(bs, c), w)
    rdm1 = np.stack([D, D])
    spec = MLPSpec(6, 32, 3, 4)
    th = init_mlp_params(spec, 5)
    E_b, gD_b, _ = xc_value_and_grads("b3lyp", xin, rdm1)
    E_n, gD_n, gth_n = xc_value_and_grads("neural_b3lyp", xin, rdm1, th, spec)
    eri = unpack_eri(int2e_packed(bs), bs.nao)
    K0 = np.stack([O.exchange_potential(rdm1[s], eri) for s in range(2)])
    inp = O.SCFInputs(T + V, S, eri, nuclear_repulsion(Z, xyz), K0, 0.2)
    xc = lambda d: xc_value_and_grads("b3lyp", xin, d)[:2]  # noqa: E731
    E_p, F_p = O.predict(xc, inp, rdm1)
    st = O.diff_scf_loop(xc, inp, rdm1, np.stack([occ, occ]), 12)
    np.savez(os.path.join(HERE, "water_def2svp_l1.npz"),
             seed=7, basis="def2-svp", level=1, npts=len(w), nao=bs.nao,
             exc_b3lyp=E_b, vxc_b3lyp=gD_b,
             mlp_h=32, mlp_seed=5, exc_neural=E_n, vxc_neural=gD_n, dtheta_neural=gth_n,
             predict_energy=E_p, predict_fock=F_p,
             scf12_energy=st.energy, scf12_energies=np.array(st.energies), scf12_rdm1=st.rdm1)
    print("wrote water_def2svp_l1.npz", len(w), bs.nao, E_b, E_n, E_p, st.energy)


if __name__ == "__main__":
    main()
