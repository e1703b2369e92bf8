"""Exact exchange from the packed ERI on the device (gdx_exchange_xo, a symmetric product over an
exchange-ordered copy, and gdx_exchange_packed, a gather; both with a fixed summation order)
and the "current" exchange mode (Functional.with_exchange("current"):
E_x = -(a0/2) sum_s tr(D_s K[D_s]), Fock -a0 K[D], rebuilt from every density).

* K against the dense-ERI contraction for a non-symmetric D (the reverse pass applies K to
  cotangents), 1e-12 relative, on both paths; two calls equal bit for bit; the two paths agree
  to 1e-13 on butane and pentane def2-SVP (n 106, 130);
* diff_scf_loop with current exchange against the oracle loop with exchange_update, water
  def2-SVP level 1, 12 cycles: energy 1e-8 Ha, density 1e-6; the graph-captured loop equals the
  eager one; the current and frozen loops differ;
* the parameter gradient through the loop (B3LYP + neural correction, 5 cycles, FP64) against
  the oracle's torch loop with K rebuilt on the tape, 1e-6 relative;
* batched_diff_scf_loop with current exchange equals the per-molecule loops (1e-9 Ha, gradient
  1e-8 relative).
"""
import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu


def _say(msg):
    print(msg, flush=True)


def _water(basis="def2-svp", seed=11):
    from graddft_amd.interface import geometry
    from graddft_amd.interface.builder import molecule_from_geometry
    sym, xyz = geometry.water(np.random.default_rng(seed))
    return molecule_from_geometry(sym, xyz, basis, 1, with_eri=True)


@pytest.mark.parametrize("path", ["xo", "gather"])
def test_exchange_gather_general_density_and_reproducible(path, monkeypatch):
    _say(f"case: K[D] for non-symmetric D, water def2-svp, {path} path")
    from graddft_amd import scf
    from graddft_amd.interface.integrals import unpack_eri
    from graddft_amd.scf import exchange
    if path == "gather":
        monkeypatch.setattr(scf, "XO_MAX_BYTES", 0)
    mol = _water()
    n = mol.nao
    rng = np.random.default_rng(1)
    D = rng.normal(size=(2, n, n))
    eri = unpack_eri(mol.rep_tensor.cpu().numpy(), n)
    K0 = np.einsum("abcd,sbd->sac", eri, D)
    Dt = torch.as_tensor(D, device="cuda")
    K1 = exchange(mol, Dt)
    K2 = exchange(mol, Dt)
    assert torch.equal(K1, K2)
    K1 = K1.cpu().numpy()
    assert np.abs(K1 - K0).max() <= 1e-12 * np.abs(K0).max()
    assert ("eri_xo" in mol.cache) == (path == "xo")


@pytest.mark.parametrize("carbons", [4, 5])
def test_exchange_xo_matches_gather(carbons, monkeypatch):
    """The exchange-ordered product against the gather from the packed ERI (independent index
    paths over the same integrals): butane def2-SVP (n 106, P' = n^2 = 11236: 256-tiles) and
    pentane (n 130, P' = 16900: the wave-per-row reader, both spins in one pass), ragged edge
    tiles; 1e-13 relative; the spins one at a time give the same bits."""
    from graddft_amd import scf
    from graddft_amd.interface import geometry
    from graddft_amd.interface.builder import molecule_from_geometry
    sym, xyz = geometry.alkane(np.random.default_rng(3), carbons)
    mol = molecule_from_geometry(sym, xyz, "def2-svp", 1, with_eri=True)
    n = mol.nao
    _say(f"case: XO vs gather exchange, C{carbons} def2-svp n={n}")
    D = torch.as_tensor(np.random.default_rng(2).normal(size=(2, n, n)), device="cuda")
    K_xo = scf.exchange(mol, D)
    assert "eri_xo" in mol.cache
    assert torch.equal(scf.exchange(mol, D), K_xo)
    # one spin at a time (pentane: the two-spin pass of the wave-per-row reader) -- same bits
    K_sep = torch.cat([scf.exchange(mol, D[0:1]), scf.exchange(mol, D[1:2])])
    assert torch.equal(K_sep, K_xo)
    scf.release_exchange_xo(mol)
    monkeypatch.setattr(scf, "XO_MAX_BYTES", 0)
    K_g = scf.exchange(mol, D)
    err = float((K_xo - K_g).abs().max() / K_g.abs().max())
    _say(f"  max rel diff {err:.2e}")
    assert err <= 1e-13


def _oracle_inputs(mol):
    from oracle.ao import eval_ao
    from oracle.scf import SCFInputs, exchange_potential
    from oracle.xc import to_inputs, xc_value_and_grads
    from graddft_amd.interface.integrals import unpack_eri
    eri = unpack_eri(mol.rep_tensor.cpu().numpy(), mol.nao)
    kx = lambda d: np.stack([exchange_potential(d[s], eri) for s in range(2)])  # noqa: E731
    D0 = mol.cache["rdm1_init"].cpu().numpy()
    inp = SCFInputs(mol.h1e.cpu().numpy(), mol.s1e.cpu().numpy(), eri, mol.nuclear_repulsion, kx(D0), 0.2)
    xin = to_inputs(*eval_ao(mol.basis, mol.grid.coords.cpu().numpy()), mol.grid.weights.cpu().numpy())

    def xc(rdm1):
        e, g, _ = xc_value_and_grads("b3lyp", xin, rdm1)
        return e, g
    return inp, xc, kx


def test_current_exchange_loop_matches_oracle():
    _say("case: diff_scf_loop B3LYP current exchange, water def2-svp level 1, 12 cycles vs oracle")
    from graddft_amd import B3LYP, diff_scf_loop
    from oracle.scf import diff_scf_loop as odiff
    mol = _water()
    f = B3LYP.with_exchange("current")
    out = diff_scf_loop(f, cycles=12)(None, mol)
    inp, xc, kx = _oracle_inputs(mol)
    st = odiff(xc, inp, mol.rdm1.cpu().numpy(), mol.mo_occ.cpu().numpy(), 12, exchange_update=kx)
    _say(f"  device {float(out.energy):.10f} oracle {st.energy:.10f}")
    assert abs(float(out.energy) - st.energy) <= 1e-8
    assert np.abs(out.rdm1.cpu().numpy() - st.rdm1).max() <= 1e-6
    g = diff_scf_loop(f, cycles=12, graph=True)
    for _ in range(3):
        r = g(None, mol)
        torch.cuda.synchronize()
        assert abs(float(r.energy) - float(out.energy)) <= 1e-12
    frozen = diff_scf_loop(B3LYP, cycles=12)(None, mol)
    assert abs(float(frozen.energy) - float(out.energy)) > 1e-6


def test_current_exchange_parameter_gradient_matches_oracle():
    _say("case: d E / d theta through 5 cycles, B3LYP + neural correction, current exchange, fp64, vs oracle")
    from graddft_amd import B3LYPNeuralCorrection, MLPSpec, diff_scf_loop
    from graddft_amd.interface.integrals import unpack_eri
    from oracle import scf as O
    from oracle import scf_grad as G
    from oracle import xc as X
    from oracle.ao import eval_ao
    from oracle.xc import MLPSpec as OSpec
    mol = _water(seed=8)
    f = B3LYPNeuralCorrection(MLPSpec(nc=6, h=32, n_hidden=2, nd=4)).with_exchange("current")
    theta = f.init(seed=4)
    th = theta.detach().clone().requires_grad_(True)
    out = diff_scf_loop(f, cycles=5)(th, mol)
    g, = torch.autograd.grad(out.energy, [th])
    xin = X.to_inputs(*eval_ao(mol.basis, mol.grid.coords.cpu().numpy()), mol.grid.weights.cpu().numpy())
    eri = unpack_eri(mol.rep_tensor.cpu().numpy(), mol.nao)
    D0 = mol.cache["rdm1_init"].cpu().numpy()
    K0 = np.stack([O.exchange_potential(D0[s], eri) for s in range(2)])
    inp = O.SCFInputs(mol.h1e.cpu().numpy(), mol.s1e.cpu().numpy(), eri, mol.nuclear_repulsion, K0, 0.2)
    ti = G.TorchSCFInputs.from_numpy(inp, current_exchange=True)
    th0 = torch.tensor(theta.cpu().numpy(), dtype=torch.float64, requires_grad=True)
    E0 = G.diff_scf_energy(lambda d: X.xc_energy_from_rdm1("b3lyp_neural_corr", xin, d, th0, OSpec(6, 32, 2, 4)),
                           ti, torch.tensor(mol.rdm1.cpu().numpy()), torch.tensor(mol.mo_occ.cpu().numpy()), 5)
    g0, = torch.autograd.grad(E0, [th0])
    rel = float(np.linalg.norm(g.cpu().numpy() - g0.numpy()) / np.linalg.norm(g0.numpy()))
    _say(f"  device E {out.energy.item():.10f} oracle {E0.item():.10f}, gradient rel {rel:.2e}")
    assert abs(out.energy.item() - E0.item()) <= 1e-9
    assert rel <= 1e-6


def test_batched_loop_with_current_exchange_matches_per_molecule():
    _say("case: batched_diff_scf_loop vs per molecule, current exchange, 3 molecules, 4 cycles, fp64")
    from graddft_amd import B3LYPNeuralCorrection, MLPSpec, batched_diff_scf_loop, diff_scf_loop
    from graddft_amd.interface import geometry
    from graddft_amd.interface.builder import molecule_from_geometry
    rng = np.random.default_rng(9)
    mols = []
    for (sym, xyz), basis in ((geometry.alkane(rng, 1), "def2-svp"), (geometry.water(rng), "6-31g*"),
                              (geometry.alkane(rng, 2), "def2-svp")):
        mols.append(molecule_from_geometry(sym, xyz, basis, 1, with_eri=True, npts_align=256))
    f = B3LYPNeuralCorrection(MLPSpec(nc=6, h=16, n_hidden=2, nd=4)).with_exchange("current")
    theta = f.init(seed=2)
    th = theta.detach().clone().requires_grad_(True)
    outs = batched_diff_scf_loop(f, cycles=4)(th, mols)
    Eb = torch.stack([o.energy for o in outs])
    gb, = torch.autograd.grad(Eb.sum(), [th])
    Er, gr = [], 0.0
    for m in mols:
        t2 = theta.detach().clone().requires_grad_(True)
        o = diff_scf_loop(f, cycles=4)(t2, m)
        g, = torch.autograd.grad(o.energy, [t2])
        Er.append(float(o.energy))
        gr = gr + g.cpu().numpy()
    assert np.abs(Eb.detach().cpu().numpy() - np.array(Er)).max() <= 1e-9
    assert np.linalg.norm(gb.cpu().numpy() - gr) <= 1e-8 * np.linalg.norm(gr)


def test_converged_molecule_is_self_consistent():
    """interface.builder.converged_molecule: the returned density is a fixed point of the B3LYP
    loop with the exchange rebuilt from each density (ten more cycles move the energy by less
    than 1e-9 Ha and the density by less than 1e-7), idempotent in the S metric with the
    electron count per spin, and its energy matches the oracle's loop from the core guess (two
    rounds of 15 cycles with exchange_update) to 1e-8 Ha.  Its cache starts afresh with the
    converged density as the initial one."""
    _say("case: converged_molecule water def2-svp level 1 (B3LYP, current exchange)")
    from graddft_amd import B3LYP, diff_scf_loop
    from graddft_amd.interface.builder import converged_molecule
    mol = _water()
    mc = converged_molecule(mol)
    assert mc.cache["rdm1_init"] is mc.rdm1 and mc.cache is not mol.cache
    f = B3LYP.with_exchange("current")
    out = diff_scf_loop(f, cycles=10)(None, mc)
    e0 = diff_scf_loop(f, cycles=1)(None, mc).energy.item()
    _say(f"  E after 1 cycle {e0:.12f}, after 10 {out.energy.item():.12f}")
    assert abs(out.energy.item() - e0) <= 1e-9
    assert (out.rdm1 - mc.rdm1).abs().max().item() <= 1e-7
    S = mc.s1e
    for s in range(2):
        D = mc.rdm1[s]
        assert abs(float(torch.trace(D @ S)) - 5.0) <= 1e-9
        assert float((D @ S @ D - D).abs().max()) <= 1e-9
    from oracle.scf import diff_scf_loop as odiff
    inp, xc, kx = _oracle_inputs(mol)
    occ = mol.mo_occ.cpu().numpy()
    st = odiff(xc, inp, mol.rdm1.cpu().numpy(), occ, 15, exchange_update=kx)
    st = odiff(xc, inp, st.rdm1, occ, 15, exchange_update=kx)      # restarted DIIS history
    _say(f"  oracle, 2 x 15 cycles from the core guess {st.energy:.12f}")
    assert abs(out.energy.item() - st.energy) <= 1e-8
