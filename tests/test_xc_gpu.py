"""GPU parity of the FP64 XC grid path against the FP64 oracle.

Inputs are identical on both sides except the AO values: the device path evaluates them
with its own kernel, the oracle with ``oracle/ao.py``, so these tests cover the AO kernel
too.  Bar (FP64 path): |dE| <= 1e-9 Ha, relative gradient error <= 1e-8 (the tier's
gate is 1e-6 Ha / 1e-5; the FP64 path is held far tighter).
"""
import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu

E_TOL = 1e-9
G_REL = 1e-8


def _say(msg):
    print(msg, flush=True)


def _geom(name, seed):
    from graddft_amd.interface import geometry as G
    rng = np.random.default_rng(seed)
    return {"water": G.water, "benzene": G.benzene}[name](rng)


def _build(name, basis, level, seed=0, rdm1=None, restricted=True):
    from graddft_amd.interface.builder import molecule_from_geometry
    sym, xyz = _geom(name, seed)
    mol = molecule_from_geometry(sym, xyz, basis, level, rdm1=rdm1)
    if rdm1 is not None and not restricted:
        mol = mol.replace(restricted=False)
    return mol


def _oracle_inputs(mol):
    from oracle.ao import eval_ao
    from oracle.xc import to_inputs
    c = mol.grid.coords.cpu().numpy()
    w = mol.grid.weights.cpu().numpy()
    ao, g, h = eval_ao(mol.basis, c)
    return to_inputs(ao, g, h, w)


def _rel(a, b):
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))


@pytest.mark.parametrize("basis", ["def2-tzvp", "6-31g*"])
def test_ao_planes_match_oracle(basis):
    _say(f"case: AO planes water/{basis} level 1")
    from oracle.ao import eval_ao
    mol = _build("water", basis, 1, seed=1)
    c = mol.grid.coords.cpu().numpy()
    ao, g, h = eval_ao(mol.basis, c)
    N, n = ao.shape
    planes = mol.ao_planes.cpu().numpy()
    ref = [ao, g[..., 0], g[..., 1], g[..., 2], h.sum(-1)]
    for q in range(5):
        err = np.abs(planes[q, :N, :n] - ref[q]).max()
        scale = np.abs(ref[q]).max()
        assert err <= 1e-12 * scale, (q, err, scale)
    assert np.all(planes[:, N:, :] == 0) and np.all(planes[:, :, n:] == 0)


def test_densities_match_oracle():
    _say("case: densities water/def2-svp level 1")
    from graddft_amd import B3LYP_GRID
    from oracle import xc as X
    mol = _build("water", "def2-svp", 1, seed=2)
    dens = B3LYP_GRID.plan(mol).densities(mol.rdm1).cpu().numpy()
    inp = _oracle_inputs(mol)
    D = torch.as_tensor(mol.rdm1.cpu().numpy())
    rho = X.density(D, inp.ao).numpy()
    grho = X.grad_density(D, inp.ao, inp.grad_ao).numpy()
    tau = X.kinetic_density(D, inp.grad_ao).numpy()
    lap = X.lapl_density(D, inp.ao, inp.grad_ao, inp.d2_ao).numpy()
    for name, got, ref in [("rho", dens[:, :, 0], rho), ("grad", dens[:, :, 1:4], grho),
                           ("tau", dens[:, :, 4], tau), ("lapl", dens[:, :, 5], lap)]:
        err = np.abs(got - ref).max()
        assert err <= 1e-11 * np.abs(ref).max(), (name, err, np.abs(ref).max())


@pytest.mark.parametrize("basis,level", [("def2-svp", 2), ("6-31g*", 1)])
def test_b3lyp_grid_energy_and_vxc(basis, level):
    _say(f"case: B3LYP grid water/{basis} level {level}")
    from graddft_amd import B3LYP_GRID
    from oracle.xc import xc_value_and_grads
    mol = _build("water", basis, level, seed=2)
    D = mol.rdm1.detach().clone().requires_grad_(True)
    E = B3LYP_GRID.xc_energy(None, mol, rdm1=D)
    (gD,) = torch.autograd.grad(E, [D])
    inp = _oracle_inputs(mol)
    E0, gD0, _ = xc_value_and_grads("b3lyp", inp, mol.rdm1.cpu().numpy())
    assert abs(float(E) - E0) <= E_TOL, (float(E), E0)
    assert _rel(gD.cpu().numpy(), gD0) <= G_REL


def _neural_case(name, basis, level, h, seed, restricted=True, kind="neural_b3lyp"):
    from graddft_amd import B3LYPNeuralCorrection, MLPSpec, NeuralFunctional
    from oracle.xc import MLPSpec as OSpec, xc_value_and_grads
    rdm1 = None
    mol0 = _build(name, basis, level, seed=seed)
    if not restricted:
        # spin-down density with one orbital fewer occupied: a valid (non-negative,
        # idempotent) density different from the spin-up one
        C = mol0.mo_coeff[0].cpu().numpy()
        nocc = int(mol0.mo_occ[0].sum().item())
        D = mol0.rdm1.cpu().numpy().copy()
        D[1] = C[:, : nocc - 1] @ C[:, : nocc - 1].T
        mol = mol0.replace(rdm1=torch.as_tensor(D, device="cuda"), restricted=False)
    else:
        mol = mol0
    spec = MLPSpec(nc=6, h=h, n_hidden=3, nd=4)
    f = NeuralFunctional(spec) if kind == "neural_b3lyp" else B3LYPNeuralCorrection(spec)
    theta = f.init(seed=seed + 7)
    D = mol.rdm1.detach().clone().requires_grad_(True)
    th = theta.detach().clone().requires_grad_(True)
    E = f.xc_energy(th, mol, rdm1=D)
    gD, gth = torch.autograd.grad(E, [D, th])
    inp = _oracle_inputs(mol)
    E0, gD0, gth0 = xc_value_and_grads(kind, inp, mol.rdm1.cpu().numpy(),
                                       theta.cpu().numpy(), OSpec(6, h, 3, 4))
    assert abs(float(E) - E0) <= E_TOL, (float(E), E0)
    assert _rel(gD.cpu().numpy(), gD0) <= G_REL, _rel(gD.cpu().numpy(), gD0)
    assert _rel(gth.cpu().numpy(), gth0) <= G_REL, _rel(gth.cpu().numpy(), gth0)


def test_neural_water_restricted():
    _say("case: neural water/6-31g* level 1 h=32")
    _neural_case("water", "6-31g*", 1, 32, seed=3)


def test_neural_water_unrestricted():
    _say("case: neural water/def2-svp level 1 h=32, D_up != D_dn")
    _neural_case("water", "def2-svp", 1, 32, seed=4, restricted=False)


def test_b3lyp_neural_correction_water():
    _say("case: B3LYP + neural correction water/def2-svp level 1 h=64, D_up != D_dn")
    _neural_case("water", "def2-svp", 1, 64, seed=6, restricted=False, kind="b3lyp_neural_corr")


def test_neural_benzene():
    _say("case: neural benzene/6-31g* level 1 h=64")
    _neural_case("benzene", "6-31g*", 1, 64, seed=5)


@pytest.mark.parametrize("precision", ["fp32", "bf16x6"])
def test_mixed_precision_within_gates(precision):
    """Mixed-precision paths against the device FP64 path on benzene (level 1, MLP 3 x 256,
    whole 128-tiles so the tile GEMMs run): |dE| <= 1e-6 Ha, gradient rel. error <= 1e-5."""
    _say(f"case: {precision} vs fp64 benzene/6-31g* level 1 h=256")
    from graddft_amd import MLPSpec, NeuralFunctional
    mol = _build("benzene", "6-31g*", 1, seed=5)
    f = NeuralFunctional(MLPSpec(nc=6, h=256, n_hidden=3, nd=4))
    theta = f.init(seed=12)
    out = {}
    for prec in ("fp64", precision):
        D = mol.rdm1.detach().clone().requires_grad_(True)
        th = theta.detach().clone().requires_grad_(True)
        E = f.xc_energy(th, mol, precision=prec, rdm1=D)
        gD, gth = torch.autograd.grad(E, [D, th])
        out[prec] = (float(E.detach()), gD.cpu().numpy(), gth.cpu().numpy())
    E0, gD0, gt0 = out["fp64"]
    E1, gD1, gt1 = out[precision]
    assert abs(E1 - E0) <= 1e-6, (E1, E0)
    assert _rel(gD1, gD0) <= 1e-5
    assert _rel(gt1, gt0) <= 1e-5


def test_molecule_density_and_energy_methods():
    """Molecule.density / grad_density / kinetic_density / lapl_density /
    get_coulomb_potential / nonXC (molecule.py:108-151, 298-300) against the oracle."""
    _say("case: Molecule grid-density and energy-term methods, water/def2-svp level 1")
    from graddft_amd.interface.integrals import unpack_eri
    from oracle import scf as O
    from oracle import xc as X
    from graddft_amd.interface import geometry
    from graddft_amd.interface.builder import molecule_from_geometry
    sym, xyz = geometry.water(np.random.default_rng(2))
    mol = molecule_from_geometry(sym, xyz, "def2-svp", 1, with_eri=True)
    inp = _oracle_inputs(mol)
    D = torch.as_tensor(mol.rdm1.cpu().numpy())
    for got, ref in [(mol.density(), X.density(D, inp.ao)),
                     (mol.grad_density(), X.grad_density(D, inp.ao, inp.grad_ao)),
                     (mol.kinetic_density(), X.kinetic_density(D, inp.grad_ao)),
                     (mol.lapl_density(), X.lapl_density(D, inp.ao, inp.grad_ao, inp.d2_ao))]:
        got, ref = got.cpu().numpy(), ref.numpy()
        assert got.shape == ref.shape
        assert np.abs(got - ref).max() <= 1e-11 * np.abs(ref).max()
    eri = unpack_eri(mol.rep_tensor.cpu().numpy(), mol.nao)
    Dt = mol.rdm1.cpu().numpy().sum(0)
    J = mol.get_coulomb_potential().cpu().numpy()
    assert np.abs(J - O.coulomb_potential(Dt, eri)).max() <= 1e-11 * np.abs(J).max()
    e0 = O.non_xc(Dt, mol.h1e.cpu().numpy(), eri, mol.nuclear_repulsion)
    assert abs(float(mol.nonXC()) - e0) <= 1e-10
