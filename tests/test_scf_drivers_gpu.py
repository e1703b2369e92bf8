"""GPU parity of the non-differentiable drivers (scf_loop with CDIIS / EDIIS / ADIIS,
damping, level shift and smearing; simple_scf_loop) and their pieces against the FP64
oracle restatements (oracle/scf.py) on identical inputs."""
import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu


def _say(msg):
    print(msg, flush=True)


def _water(basis="def2-svp", level=1, seed=11):
    from graddft_amd.interface import geometry
    from graddft_amd.interface.builder import molecule_from_geometry
    sym, xyz = geometry.water(np.random.default_rng(seed))
    return molecule_from_geometry(sym, xyz, basis, level, with_eri=True)


def _oracle_inputs(mol):
    from graddft_amd.interface.integrals import unpack_eri
    from oracle.ao import eval_ao
    from oracle.scf import SCFInputs, exchange_potential
    from oracle.xc import to_inputs, xc_value_and_grads
    eri = unpack_eri(mol.rep_tensor.cpu().numpy(), mol.nao)
    D0 = mol.cache["rdm1_init"].cpu().numpy()
    K0 = np.stack([exchange_potential(D0[s], eri) for s in range(2)])
    inp = SCFInputs(mol.h1e.cpu().numpy(), mol.s1e.cpu().numpy(), eri, mol.nuclear_repulsion, K0, 0.2)
    xin = to_inputs(*eval_ao(mol.basis, mol.grid.coords.cpu().numpy()), mol.grid.weights.cpu().numpy())
    return inp, (lambda d: xc_value_and_grads("b3lyp", xin, d)[:2]), xin


def _sym(rng, *shape):
    a = rng.normal(size=shape)
    return a + np.swapaxes(a, -1, -2)


def test_pieces_match_oracle():
    _say("case: damping / level_shift / orbital_grad / get_occ / CDIIS / EDIIS / ADIIS weights, n=9")
    from graddft_amd.evaluate import Diis, damping, level_shift
    from graddft_amd.molecule import get_occ, orbital_grad
    from oracle import scf as O
    rng = np.random.default_rng(5)
    n = 9
    A = rng.normal(size=(n, n))
    S = A @ A.T / n + np.eye(n)
    D, F = _sym(rng, n, n), _sym(rng, n, n)
    t = lambda x: torch.as_tensor(x, device="cuda")  # noqa: E731
    assert np.abs(damping(t(S), t(D), t(F), 0.3).cpu().numpy() - O.damping(S, D, F, 0.3)).max() <= 1e-12
    assert np.abs(level_shift(t(S), t(D), t(F), 0.2).cpu().numpy() - O.level_shift(S, D, F, 0.2)).max() <= 1e-12
    C, F2 = rng.normal(size=(2, n, n)), _sym(rng, 2, n, n)
    e = rng.normal(size=(2, n))
    occ = O.get_occ(e, np.array([4, 3]))
    assert np.array_equal(get_occ(t(e), [4, 3]).cpu().numpy(), occ)
    g = orbital_grad(t(C), t(occ), t(F2)).cpu().numpy()
    assert np.abs(g - O.orbital_grad(C, occ, F2)).max() <= 1e-12
    k = 5
    E = rng.normal(size=(k, 2, n, n))
    errs = E - np.swapaxes(E, -1, -2)
    dev = Diis(t(S), None, 8, "CDIIS")
    ora = O.Diis(S, np.eye(n), 8, "CDIIS")
    x = dev.cdiis_minimize(t(errs)).cpu().numpy()
    assert np.abs(x - ora.cdiis_minimize(errs)).max() <= 1e-10
    # F_i = H + D_i (a linear Fock map): the EDIIS and ADIIS costs are then convex on the
    # simplex, so the minimiser is unique and both BFGS runs must find it
    Ds, Es = _sym(rng, k, 2, n, n), rng.normal(size=k)
    Fs = _sym(rng, 2, n, n)[None] + Ds
    c, f = dev.ediis_minimize(t(Ds), t(Fs), t(Es))
    c0, f0 = ora.ediis_minimize(Ds, Fs, Es)
    assert abs(f - f0) <= 1e-9 * max(1.0, abs(f0)) and np.abs(c.cpu().numpy() - c0).max() <= 1e-5
    c, f = dev.adiis_minimize(t(Ds), t(Fs), 3)
    c0, f0 = ora.adiis_minimize(Ds, Fs, 3)
    assert abs(f - f0) <= 1e-9 * max(1.0, abs(f0)) and np.abs(c.cpu().numpy() - c0).max() <= 1e-5


@pytest.mark.parametrize("opts,tol", [
    (dict(), 1e-8),
    (dict(diis_start_cycle=3, damp_factor=(0.3, 0.3), level_shift_factor=(0.2, 0.2)), 1e-8),
    (dict(diis_method="EDIIS2"), 1e-8),
    (dict(diis_method="ADIIS", cycles=25), 5e-7),
    (dict(smearing="fermi-dirac", smearing_sigma=0.01), 1e-8),
], ids=["cdiis", "damp_shift", "ediis2", "adiis", "fermi"])
def test_scf_loop_matches_oracle(opts, tol):
    _say(f"case: scf_loop {opts} B3LYP water/def2-svp level 1")
    from graddft_amd import B3LYP
    from graddft_amd.evaluate import scf_loop
    from oracle import scf as O
    mol = _water()
    out = scf_loop(B3LYP, **opts)(None, mol)
    inp, xc, _ = _oracle_inputs(mol)
    st, cycles = O.scf_loop(xc, inp, mol.rdm1.cpu().numpy(), mol.mo_occ.cpu().numpy(), **opts)
    assert abs(float(out.energy) - st.energy) <= tol, (float(out.energy), st.energy)
    assert abs(out.cache["scf_cycles"] - cycles) <= (0 if tol <= 1e-8 else 1)
    assert np.abs(out.rdm1.cpu().numpy() - st.rdm1).max() <= 1e-5


def test_simple_scf_loop_matches_oracle():
    _say("case: simple_scf_loop B3LYP water/def2-svp level 1, 10 cycles, mixing 0.4")
    from graddft_amd import B3LYP
    from graddft_amd.evaluate import simple_scf_loop
    from oracle import scf as O
    mol = _water()
    out = simple_scf_loop(B3LYP, cycles=10, mixing_factor=0.4)(None, mol)
    inp, xc, _ = _oracle_inputs(mol)
    st = O.simple_scf_loop(xc, inp, mol.rdm1.cpu().numpy(), mol.mo_coeff.cpu().numpy(),
                           mol.mo_energy.cpu().numpy(), mol.mo_occ.cpu().numpy(), mixing_factor=0.4, cycles=10)
    assert abs(float(out.energy) - st.energy) <= 1e-9, (float(out.energy), st.energy)
    assert np.abs(out.rdm1.cpu().numpy() - st.rdm1).max() <= 1e-7


def test_charge_check_rejects_a_wrong_density():
    _say("case: simple_scf_loop charge check, rdm1 scaled by 1.1")
    from graddft_amd import B3LYP
    from graddft_amd.evaluate import simple_scf_loop
    mol = _water()
    bad = mol.replace(rdm1=mol.rdm1 * 1.1)
    with pytest.raises(AssertionError, match="Total charge is not conserved"):
        simple_scf_loop(B3LYP, cycles=2)(None, bad)


def _generic_orbitals(mol):
    """The molecule's orbitals mixed by a fixed random orthogonal matrix: still orthonormal,
    but with no symmetry zeros, so no Householder pivot of W = C^T D^-1 sits at rounding level
    (where LAPACK's column sign, and with it the starting orbitals, is decided by rounding)."""
    n = mol.nao
    R, _ = np.linalg.qr(np.random.default_rng(3).normal(size=(n, n)))
    C = mol.mo_coeff.cpu().numpy() @ R
    return mol.replace(mo_coeff=torch.as_tensor(C, device="cuda"))


def _orbital_oracle(mol, lr, cycles, e_conv, whitening="PCA"):
    from oracle import scf_grad as G
    from oracle import xc as X
    inp, _, xin = _oracle_inputs(mol)
    ti = G.TorchSCFInputs.from_numpy(inp)
    C = torch.tensor(mol.mo_coeff.cpu().numpy())
    occ = torch.tensor(mol.mo_occ.cpu().numpy())
    return G.mol_orb_optimizer(lambda d: X.xc_energy_from_rdm1("b3lyp", xin, d), ti, C, occ, lr, cycles,
                               e_conv=e_conv, whitening=whitening)


@pytest.mark.parametrize("whitening", ["PCA", "ZCA", "Cholesky"])
def test_mol_orb_optimizer_matches_oracle(whitening):
    _say(f"case: mol_orb_optimizer {whitening}, adam 0.01, 12 cycles, B3LYP water/def2-svp level 1")
    from graddft_amd import B3LYP, adam
    from graddft_amd.evaluate import mol_orb_optimizer
    mol = _generic_orbitals(_water())
    out = mol_orb_optimizer(B3LYP, adam(0.01), cycles=12, e_conv=1e-7, whitening=whitening)(None, mol)
    es = _orbital_oracle(mol, 0.01, 12, 1e-7, whitening)
    assert out.cache["orb_cycles"] == len(es)
    assert abs(float(out.energy) - es[-1]) <= 1e-8, (float(out.energy), es[-1])
    assert es[-1] < es[0]


def test_jitted_mol_orb_optimizer_matches_oracle():
    _say("case: jitted_mol_orb_optimizer, adam 0.01, 8 cycles, B3LYP water/def2-svp level 1")
    from graddft_amd import B3LYP, adam
    from graddft_amd.evaluate import jitted_mol_orb_optimizer
    mol = _generic_orbitals(_water())
    out = jitted_mol_orb_optimizer(B3LYP, adam(0.01), cycles=8)(None, mol)
    es = _orbital_oracle(mol, 0.01, 8, None)
    assert len(es) == 8
    assert abs(float(out.energy) - es[-1]) <= 1e-8, (float(out.energy), es[-1])


def test_scf_loop_neural_functional_matches_oracle():
    _say("case: scf_loop CDIIS + level shift 0.1, neural B3LYP MLP h=32, water/def2-svp level 1")
    from graddft_amd import MLPSpec, NeuralFunctional
    from graddft_amd.evaluate import scf_loop
    from graddft_amd.interface.integrals import unpack_eri
    from oracle import scf as O
    from oracle.ao import eval_ao
    from oracle.xc import MLPSpec as OSpec
    from oracle.xc import to_inputs, xc_value_and_grads
    mol = _water()
    f = NeuralFunctional(MLPSpec(nc=6, h=32, n_hidden=2, nd=4))
    theta = f.init(seed=2)
    out = scf_loop(f, level_shift_factor=(0.1, 0.1))(theta, mol)
    eri = unpack_eri(mol.rep_tensor.cpu().numpy(), mol.nao)
    inp = O.SCFInputs(mol.h1e.cpu().numpy(), mol.s1e.cpu().numpy(), eri, mol.nuclear_repulsion, None, 0.0)
    xin = to_inputs(*eval_ao(mol.basis, mol.grid.coords.cpu().numpy()), mol.grid.weights.cpu().numpy())
    th, spec = theta.cpu().numpy(), OSpec(6, 32, 2, 4)
    st, cycles = O.scf_loop(lambda d: xc_value_and_grads("neural_b3lyp", xin, d, th, spec)[:2], inp,
                            mol.rdm1.cpu().numpy(), mol.mo_occ.cpu().numpy(), level_shift_factor=(0.1, 0.1))
    assert abs(float(out.energy) - st.energy) <= 1e-8, (float(out.energy), st.energy)
    assert out.cache["scf_cycles"] == cycles


def test_batched_scf_loop_matches_molecule_by_molecule():
    """Concurrent graph replays on 3 streams equal the eager per-molecule loops, on two
    batches in a row with eager work and synchronisations between them."""
    _say("case: batched_scf_loop, H2O x2 + CH4 + C2H6 def2-svp level 1, B3LYP 8 cycles, 3 streams")
    from graddft_amd import B3LYP, batched_scf_loop, diff_scf_loop
    from graddft_amd.interface import geometry
    from graddft_amd.interface.builder import molecule_from_geometry
    rng = np.random.default_rng(4)
    mols = [_water(seed=11), _water(seed=12)]
    for nc in (1, 2):
        sym, xyz = geometry.alkane(rng, nc)
        mols.append(molecule_from_geometry(sym, xyz, "def2-svp", 1, with_eri=True))
    ref = [float(diff_scf_loop(B3LYP, cycles=8)(None, m).energy) for m in mols]
    bat = batched_scf_loop(B3LYP, cycles=8, n_streams=3)
    for _ in range(2):
        outs = bat(None, mols)
        torch.cuda.synchronize()
        for o, e in zip(outs, ref):
            assert abs(float(o.energy) - e) <= 1e-10, (float(o.energy), e)
        _ = sum(float(m.rdm1.sum()) for m in mols)   # eager work between the batches
