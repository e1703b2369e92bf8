"""The molecule batch (gdx_xc_batch_*: concatenated grid rows, one MLP pass, ragged AO-side
launches) against the same molecules evaluated one at a time, on the GPU.

Molecules of different bases and grids (CH4 / C2H6 / C6H14 def2-SVP, H2O def2-TZVP, level-1
grids; padded bases of 48, 48, 64 and 160, which take the three V_xc tile groups):
E_xc per molecule, dE/dD per molecule and sum_m g_m dE_m/dtheta for per-molecule cotangents;
the second-order pass with per-molecule directions; the differentiable SCF loop with the
batched XC (energies and the parameter gradient of their sum); the batched mse_energy_loss.
Bars: FP64 1e-12 relative for per-point results, 1e-10 for reductions over points, 1e-9 Ha /
1e-8 relative for the loops; FP32 against its own per-molecule path at 1e-9 Ha / 1e-6.
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


def _mols(with_eri=False):
    from graddft_amd.interface import geometry
    from graddft_amd.interface.builder import molecule_from_geometry
    rng = np.random.default_rng(9)
    out = []
    for sym_xyz, basis in ((geometry.alkane(rng, 1), "def2-svp"), (geometry.water(rng), "def2-tzvp"),
                           (geometry.alkane(rng, 2), "def2-svp"), (geometry.alkane(rng, 6), "def2-svp")):
        sym, xyz = sym_xyz
        out.append(molecule_from_geometry(sym, xyz, basis, 1, with_eri=with_eri, npts_align=256))
    return out


def _functional(kind):
    from graddft_amd import B3LYP_GRID, MLPSpec, NeuralFunctional
    if kind == "b3lyp":
        return B3LYP_GRID, None
    f = NeuralFunctional(MLPSpec(nc=6, h=32, n_hidden=3, nd=4))
    return f, f.init(seed=3)


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
@pytest.mark.parametrize("kind", ["b3lyp", "neural"])
def test_batch_xc_matches_per_molecule(kind, precision):
    _say(f"case: batch XC vs per molecule, {kind}, {precision}, CH4/H2O/C2H6/C6H14")
    from graddft_amd.batch import XCBatchPlan
    f, theta = _functional(kind)
    mols = _mols()
    if kind == "neural":
        mols = [m.replace(restricted=False) for m in mols]
    plan = XCBatchPlan(f, mols, precision)
    D = plan.concat([m.rdm1 for m in mols])
    exc = torch.empty(len(mols), dtype=torch.float64, device="cuda")
    vxc = torch.empty_like(D)
    g = torch.tensor([0.3, -1.2, 0.7, 0.45], dtype=torch.float64, device="cuda")
    dth = None if theta is None else torch.empty_like(theta)
    plan.forward(D, theta, exc)
    plan.backward(theta, g, vxc, dth)
    torch.cuda.synchronize()
    e_ref, v_ref, t_ref = [], [], 0.0
    for i, m in enumerate(mols):
        p = f.plan(m, precision)
        e = torch.empty((), dtype=torch.float64, device="cuda")
        dD = torch.empty_like(m.rdm1)
        dt = None if theta is None else torch.empty_like(theta)
        p.forward(m.rdm1.contiguous(), theta, e)
        p.backward(m.rdm1.contiguous(), theta, g[i].reshape(()).contiguous(), dD, dt, p.generation)
        torch.cuda.synchronize()
        e_ref.append(e.item())
        v_ref.append(dD.cpu().numpy())
        if dt is not None:
            t_ref = t_ref + dt.cpu().numpy()
    tol = 1e-12 if precision == "fp64" else 1e-9
    e_b = exc.cpu().numpy()
    _say(f"  E batch {e_b} per molecule {np.array(e_ref)}")
    assert np.abs(e_b - np.array(e_ref)).max() <= tol * np.abs(e_ref).max()
    for vb, vr in zip(plan.split(vxc), v_ref):
        assert _rel(vb.cpu().numpy(), vr) <= (1e-11 if precision == "fp64" else 1e-6)
    if theta is not None:
        assert _rel(dth.cpu().numpy(), t_ref) <= (1e-10 if precision == "fp64" else 1e-6)


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_batch_hvp_matches_per_molecule(precision):
    _say(f"case: batch second-order pass vs per molecule, neural h=32, {precision}")
    from graddft_amd.batch import XCBatchPlan
    f, theta = _functional("neural")
    mols = [m.replace(restricted=False) for m in _mols()]
    plan = XCBatchPlan(f, mols, precision)
    rng = np.random.default_rng(5)
    dirs = [torch.as_tensor(rng.normal(size=(2, m.nao, m.nao)) * 1e-2, device="cuda") for m in mols]
    g = torch.tensor([0.5, 1.5, -0.25, 0.8], dtype=torch.float64, device="cuda")
    D = plan.concat([m.rdm1 for m in mols])
    exc = torch.empty(len(mols), dtype=torch.float64, device="cuda")
    edot = torch.empty_like(exc)
    dD = torch.empty_like(D)
    dth = torch.empty_like(theta)
    plan.hvp(D, theta, plan.concat(dirs), g, exc, edot, dD, dth)
    torch.cuda.synchronize()
    t_ref = 0.0
    for i, m in enumerate(mols):
        p = f.plan(m, precision)
        e = torch.empty((), dtype=torch.float64, device="cuda")
        ed = torch.empty_like(e)
        dR = torch.empty_like(m.rdm1)
        dt = torch.empty_like(theta)
        p.hvp(m.rdm1.contiguous(), theta, dirs[i].contiguous(), g[i].reshape(()).contiguous(), e, ed, dR, dt)
        torch.cuda.synchronize()
        tol = 1e-12 if precision == "fp64" else 1e-9
        assert abs(exc[i].item() - e.item()) <= tol * abs(e.item())
        assert abs(edot[i].item() - ed.item()) <= 1e-9 * max(1.0, abs(ed.item()))
        assert _rel(plan.split(dD)[i].cpu().numpy(), dR.cpu().numpy()) <= (1e-10 if precision == "fp64" else 1e-6)
        t_ref = t_ref + dt.cpu().numpy()
    assert _rel(dth.cpu().numpy(), t_ref) <= (1e-10 if precision == "fp64" else 1e-6)


def test_batched_scf_loop_and_loss_match_per_molecule():
    _say("case: batched_diff_scf_loop + mse_energy_loss vs per molecule, neural h=32, 4 cycles, fp64")
    from graddft_amd import batched_diff_scf_loop, diff_scf_loop, mse_energy_loss
    f, theta = _functional("neural")
    mols = _mols(with_eri=True)
    th = theta.detach().clone().requires_grad_(True)
    outs = batched_diff_scf_loop(f, cycles=4)(th, mols)
    Eb = torch.stack([o.energy for o in outs])
    gb, = torch.autograd.grad(Eb.sum(), [th])
    Er, gr = [], 0.0
    for m in mols:
        t2 = theta.detach().clone().requires_grad_(True)
        o = diff_scf_loop(f, cycles=4)(t2, m)
        g, = torch.autograd.grad(o.energy, [t2])
        Er.append(o.energy.item())
        gr = gr + g.cpu().numpy()
    _say(f"  batched {Eb.detach().cpu().numpy()} per molecule {np.array(Er)}")
    assert np.abs(Eb.detach().cpu().numpy() - np.array(Er)).max() <= 1e-9
    assert _rel(gb.cpu().numpy(), gr) <= 1e-8
    truth = [e - 0.01 for e in Er]
    lb = mse_energy_loss(theta, batched_diff_scf_loop(f, cycles=4), mols, truth)
    lr = mse_energy_loss(theta, diff_scf_loop(f, cycles=4), mols, truth)
    assert abs(lb.item() - lr.item()) <= 1e-10 * abs(lr.item())
