"""The reference's single-part built-ins (LSDA, B88, VWN, PW92, LYP; popular_functionals.py:375-393)
on the GPU against the oracle: E_xc and dE/dD for equal and for different spin densities, the
second-order pass, the total energy, and a diff_scf_loop with PW92 correlation.

Bars (FP64 path): |dE| <= 1e-9 Ha, relative gradient error <= 1e-8, loop energy 1e-8 Ha.
"""
import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu

KINDS = ["LSDA", "B88", "VWN", "PW92", "LYP"]


def _say(msg):
    print(msg, flush=True)


def _rel(a, b):
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))


def _water(level=1, with_eri=False):
    from graddft_amd.interface import geometry
    from graddft_amd.interface.builder import molecule_from_geometry
    sym, xyz = geometry.water(np.random.default_rng(6))
    return molecule_from_geometry(sym, xyz, "def2-svp", level, with_eri=with_eri)


def _inputs(mol):
    from oracle.ao import eval_ao
    from oracle.xc import to_inputs
    ao, g, h = eval_ao(mol.basis, mol.grid.coords.cpu().numpy())
    return to_inputs(ao, g, h, mol.grid.weights.cpu().numpy())


@pytest.mark.parametrize("name", KINDS)
@pytest.mark.parametrize("polarised", [False, True], ids=["equal_spins", "spin_polarised"])
def test_energy_and_vxc_match_oracle(name, polarised):
    _say(f"case: {name} E_xc + dE/dD water/def2-svp level 1, polarised={polarised}")
    import graddft_amd as G
    from oracle.xc import xc_value_and_grads
    f = getattr(G, name)
    mol = _water()
    D = mol.rdm1.cpu().numpy().copy()
    if polarised:
        D[1] *= 0.8
        mol = mol.replace(rdm1=torch.as_tensor(D, device="cuda"), restricted=False)
    Dt = mol.rdm1.detach().clone().requires_grad_(True)
    E = f.xc_energy(None, mol, rdm1=Dt)
    gD, = torch.autograd.grad(E, [Dt])
    E0, gD0, _ = xc_value_and_grads(name.lower(), _inputs(mol), D)
    _say(f"  E {float(E):.12f} oracle {E0:.12f}")
    assert abs(float(E) - E0) <= 1e-9
    assert _rel(gD.cpu().numpy(), gD0) <= 1e-8


@pytest.mark.parametrize("name", ["PW92", "B88", "LYP"])
def test_second_order_pass_matches_oracle(name):
    _say(f"case: {name} hvp water/def2-svp level 1, D_up != D_dn")
    import graddft_amd as G
    from oracle import scf_grad as SG
    f = getattr(G, name)
    mol = _water()
    D = mol.rdm1.cpu().numpy().copy()
    D[1] *= 0.9
    mol = mol.replace(rdm1=torch.as_tensor(D, device="cuda"), restricted=False)
    n = mol.nao
    rng = np.random.default_rng(3)
    dD = rng.normal(size=(2, n, n)) * 1e-2
    dev = lambda a: torch.as_tensor(a, dtype=torch.float64, device="cuda").contiguous()  # noqa: E731
    exc = torch.empty((), dtype=torch.float64, device="cuda")
    edot = torch.empty_like(exc)
    dR = torch.empty((2, n, n), dtype=torch.float64, device="cuda")
    f.plan(mol, "fp64").hvp(dev(D), None, dev(dD), dev(0.6), exc, edot, dR, None)
    torch.cuda.synchronize()
    E0, ed0, dR0, _ = SG.xc_hvp(name.lower(), _inputs(mol), D, None, None, dD, 0.6)
    assert abs(exc.item() - E0) <= 1e-9
    assert abs(edot.item() - ed0) <= 1e-9 * max(1.0, abs(ed0))
    assert _rel(dR.cpu().numpy(), dR0) <= 1e-8


def test_pw92_total_energy_and_scf_loop_match_oracle():
    _say("case: PW92 Functional.energy + 6-cycle diff_scf_loop, water/def2-svp level 1")
    from graddft_amd import PW92, diff_scf_loop
    from graddft_amd.interface.integrals import unpack_eri
    from oracle import scf as O
    from oracle.xc import xc_value_and_grads
    mol = _water(with_eri=True)
    inp = O.SCFInputs(mol.h1e.cpu().numpy(), mol.s1e.cpu().numpy(), unpack_eri(mol.rep_tensor.cpu().numpy(), mol.nao),
                      mol.nuclear_repulsion, None, 0.0)
    xin = _inputs(mol)
    xc = lambda d: xc_value_and_grads("pw92", xin, d)[:2]  # noqa: E731
    D0 = mol.rdm1.cpu().numpy()
    e_dev = PW92.energy(None, mol).item()
    e_o, _ = O.predict(xc, inp, D0)
    assert abs(e_dev - e_o) <= 1e-9, (e_dev, e_o)
    out = diff_scf_loop(PW92, cycles=6)(None, mol)
    st = O.diff_scf_loop(xc, inp, D0, mol.mo_occ.cpu().numpy(), 6)
    _say(f"  6 cycles {out.energy.item():.12f} oracle {st.energy:.12f}")
    assert abs(out.energy.item() - st.energy) <= 1e-8
