"""Workload-level parity cases on the GPU (the bench measures the benzene and water cases).

* training batch of alkanes (def2-SVP): mse_energy_loss over diff_scf_loop for a batch of
  molecules of different sizes, parameter gradient against the oracle (FP64) and the
  mixed-precision gradients against FP64;
* naphthalene def2-SVP (nao 180), packed ERI, 20-cycle diff_scf_loop (mixed paths with their
  default final FP64 cycle): B3LYP energies in FP32 and split-BF16 against FP64; B3LYP plus
  the width-256 neural correction (exact exchange from the ERI): energy and d E / d theta;
* 40 carbon atoms def2-SVP (nao 560), no ERI, D from 120 seeded orthonormal orbitals:
  E_xc, V_xc and dE/dtheta with the 3 x 256 MLP, FP32 and split-BF16 against FP64, plus the
  electron count of the grid densities.

Gates: |dE| <= 1e-6 Ha and relative gradient error <= 1e-5 between precisions; FP64 against
the oracle at 1e-9 Ha / 1e-6 relative.
"""
import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu


def _say(msg):
    print(msg, flush=True)


def _rel(a, b):
    a, b = np.asarray(a), np.asarray(b)
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))


def test_alkane_training_batch_gradient():
    _say("case: alkane batch CH4, C2H6, C3H8 def2-svp level 1, MLP 3x128, 3-cycle diff_scf_loop, mse loss")
    from graddft_amd import MLPSpec, NeuralFunctional, diff_scf_loop, mse_energy_loss
    from graddft_amd.interface import geometry
    from graddft_amd.interface.builder import molecule_from_geometry
    from graddft_amd.interface.integrals import unpack_eri
    from oracle import scf as O
    from oracle import scf_grad as G
    from oracle import xc as X
    from oracle.ao import eval_ao
    rng = np.random.default_rng(21)
    mols = []
    for nc in (1, 2, 3):
        sym, xyz = geometry.alkane(rng, nc)
        mols.append(molecule_from_geometry(sym, xyz, "def2-svp", 1, with_eri=True))
    truth = [-40.4, -79.7, -119.0]
    spec = MLPSpec(nc=6, h=128, n_hidden=3, nd=4)
    f = NeuralFunctional(spec)
    theta = f.init(seed=9)
    grads = {}
    for prec in ("fp64", "fp32"):
        th = theta.detach().clone().requires_grad_(True)
        loss = mse_energy_loss(th, diff_scf_loop(f, cycles=3, precision=prec), mols, truth)
        g, = torch.autograd.grad(loss, [th])
        grads[prec] = (loss.item(), g.cpu().numpy())
    # oracle: the same loss through the torch restatement of the loop
    th0 = torch.tensor(theta.cpu().numpy(), dtype=torch.float64, requires_grad=True)
    total = 0.0
    for mol, t in zip(mols, truth):
        xin = X.to_inputs(*eval_ao(mol.basis, mol.grid.coords.cpu().numpy()), mol.grid.weights.cpu().numpy())
        eri = unpack_eri(mol.rep_tensor.cpu().numpy(), mol.nao)
        inp = O.SCFInputs(mol.h1e.cpu().numpy(), mol.s1e.cpu().numpy(), eri, mol.nuclear_repulsion, None, 0.0)
        E = G.diff_scf_energy(lambda d, xin=xin: X.xc_energy_from_rdm1("neural_b3lyp", xin, d, th0,
                                                                         X.MLPSpec(6, 128, 3, 4)),
                              G.TorchSCFInputs.from_numpy(inp), torch.tensor(mol.rdm1.cpu().numpy()),
                              torch.tensor(mol.mo_occ.cpu().numpy()), 3)
        ne = float(mol.atom_index.sum().item()) - mol.charge
        total = total + ((E - t) / ne) ** 2
    total = total / len(mols)
    g0, = torch.autograd.grad(total, [th0])
    assert abs(grads["fp64"][0] - total.item()) <= 1e-10 * max(1.0, total.item())
    assert _rel(grads["fp64"][1], g0.numpy()) <= 1e-6, _rel(grads["fp64"][1], g0.numpy())
    assert _rel(grads["fp32"][1], grads["fp64"][1]) <= 1e-5, _rel(grads["fp32"][1], grads["fp64"][1])


def _naphthalene():
    from graddft_amd.interface import geometry
    from graddft_amd.interface.builder import molecule_from_geometry
    sym, xyz = geometry.naphthalene(np.random.default_rng(4))
    mol = molecule_from_geometry(sym, xyz, "def2-svp", 1, with_eri=True)
    assert mol.nao == 180
    return mol


def test_naphthalene_b3lyp_scf_energy():
    _say("case: naphthalene def2-svp level 1 (nao 180), B3LYP, packed ERI, 20-cycle diff_scf_loop, "
         "fp64/fp32/bf16x6 energies")
    from graddft_amd import B3LYP, diff_scf_loop
    mol = _naphthalene()
    out = {}
    for prec in ("fp64", "fp32", "bf16x6"):
        res = diff_scf_loop(B3LYP, cycles=20, precision=prec)(None, mol)
        out[prec] = res.energy.item()
        _say(f"  {prec}: E = {out[prec]:.10f}")
    assert -386.0 < out["fp64"] < -383.0, out["fp64"]
    for prec in ("fp32", "bf16x6"):
        assert abs(out[prec] - out["fp64"]) <= 1e-6, (prec, out[prec], out["fp64"])


def test_naphthalene_neural_correction_scf_gradient():
    """B3LYP + the width-256 neural correction (random theta), 20 cycles: the loop is still far
    from self-consistency (||FDS - SDF|| ~ 2e-3), so its energy is first-order sensitive to the
    Fock matrices of the last cycles; the mixed paths run their last cycle in FP64
    (diff_scf_loop's default refine_cycles=1)."""
    _say("case: naphthalene def2-svp level 1 (nao 180), B3LYP + neural correction 3x256, packed ERI, "
         "20-cycle diff_scf_loop, d E / d theta, fp64/fp32/bf16x6")
    from graddft_amd import B3LYPNeuralCorrection, MLPSpec, diff_scf_loop
    mol = _naphthalene()
    f = B3LYPNeuralCorrection(MLPSpec(nc=6, h=256, n_hidden=3, nd=4))
    theta = f.init(seed=5)
    out = {}
    for prec in ("fp64", "fp32", "bf16x6"):
        th = theta.detach().clone().requires_grad_(True)
        res = diff_scf_loop(f, cycles=20, precision=prec)(th, mol)
        g, = torch.autograd.grad(res.energy, [th])
        out[prec] = (res.energy.item(), g.cpu().numpy())
        _say(f"  {prec}: E = {out[prec][0]:.10f}")
    E0, g0 = out["fp64"]
    for prec in ("fp32", "bf16x6"):
        _say(f"  {prec} vs fp64: dE = {out[prec][0] - E0:.3e}, grad rel = {_rel(out[prec][1], g0):.3e}")
    assert -420.0 < E0 < -370.0, E0
    for prec in ("fp32", "bf16x6"):
        E1, g1 = out[prec]
        assert abs(E1 - E0) <= 1e-6, (prec, E1, E0)
        assert _rel(g1, g0) <= 1e-5, (prec, _rel(g1, g0))


def test_carbon_cloud_grid_xc_at_scale():
    _say("case: 40 C def2-svp level 2 (nao 560), D from 120 seeded orbitals, MLP 3x256: E_xc, V_xc, dE/dtheta")
    from graddft_amd import MLPSpec, NeuralFunctional
    from graddft_amd.interface import geometry
    from graddft_amd.interface.basis import build_basis
    from graddft_amd.interface.builder import molecule_from_geometry
    from graddft_amd.interface.integrals import int1e
    rng = np.random.default_rng(5)
    sym, xyz = geometry.carbon_cloud(rng)
    bs = build_basis(sym, xyz, "def2-svp")
    S, _, _ = int1e(bs, [6] * len(sym), xyz)
    C = rng.normal(size=(bs.nao, 120))
    w, V = np.linalg.eigh(C.T @ S @ C)
    C = C @ V @ np.diag(w ** -0.5) @ V.T          # C^T S C = 1
    D = C @ C.T
    mol = molecule_from_geometry(sym, xyz, "def2-svp", 2, rdm1=D)
    assert mol.nao == 560
    f = NeuralFunctional(MLPSpec(nc=6, h=256, n_hidden=3, nd=4))
    theta = f.init(seed=6)
    dens = f.plan(mol, "fp64").densities(mol.rdm1)
    ne = float((mol.grid.weights * dens[:, 0, 0]).sum())
    assert abs(ne - 120.0) <= 1e-3 * 120.0, ne
    out = {}
    for prec in ("fp64", "fp32", "bf16x6"):
        Dg = mol.rdm1.detach().clone().requires_grad_(True)
        th = theta.detach().clone().requires_grad_(True)
        E = f.xc_energy(th, mol, precision=prec, rdm1=Dg)
        gD, gt = torch.autograd.grad(E, [Dg, th])
        out[prec] = (E.item(), gD.cpu().numpy(), gt.cpu().numpy())
        _say(f"  {prec}: E_xc = {out[prec][0]:.10f}")
    E0, gD0, gt0 = out["fp64"]
    for prec in ("fp32", "bf16x6"):
        _say(f"  {prec} vs fp64: dE = {out[prec][0] - E0:.3e}, dE/dD rel = {_rel(out[prec][1], gD0):.3e}, "
             f"dE/dtheta rel = {_rel(out[prec][2], gt0):.3e}")
    for prec in ("fp32", "bf16x6"):
        E1, gD1, gt1 = out[prec]
        assert abs(E1 - E0) <= 1e-6, (prec, E1, E0)
        assert _rel(gD1, gD0) <= 1e-5, (prec, _rel(gD1, gD0))
        assert _rel(gt1, gt0) <= 1e-5, (prec, _rel(gt1, gt0))
