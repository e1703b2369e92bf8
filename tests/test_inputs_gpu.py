"""One-electron integrals (gdx_int1e_eval) and Becke grid weights (gdx_becke_weights) on the
device.

* S, T, V of H2 / STO-3G at R = 1.4 bohr against Szabo & Ostlund's values (section 3.5): an
  independent pin of the device kernels.
* S, T, V against the host integral library (pinned by the textbook and PySCF-fixture tests of
  the host library) for spherical d and f shells (water def2-TZVP), Cartesian d (6-31G*), and
  larger molecules (propane def2-SVP, benzene 6-31G*, 40 carbons def2-SVP with V over 40
  nuclei).  Bar: max |device - host| <= 1e-12 max |host| (same integrals, Obara-Saika overlaps on
  the device against McMurchie-Davidson ones on the host); outputs exactly symmetric; a second
  call reproduces the first bit for bit.
* Becke weights against the numpy partition of interface/grids.py on the same points (water,
  octane, 40 carbons): the kernel evaluates the same formulas in the same order without
  contraction (IEEE-rounded sqrt and division on both sides), bar bitwise equality; the kept
  grids are identical.
* A molecule built with device inputs has the host-built molecule's S, h1e and grid.
"""
import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu


def _say(msg):
    print(msg, flush=True)


def _geometry(name):
    from graddft_amd.interface import geometry
    rng = np.random.default_rng(5)
    if name == "water":
        return geometry.water(rng)
    if name == "propane":
        return geometry.alkane(rng, 3)
    if name == "octane":
        return geometry.alkane(rng, 8)
    if name == "benzene":
        return geometry.benzene(rng)
    return geometry.carbon_cloud(rng, 40)


def test_int1e_textbook_h2():
    from graddft_amd.interface.basis import build_basis
    from graddft_amd.interface.integrals import int1e_device
    _say("int1e H2 STO-3G")
    xyz = np.array([[0, 0, 0], [1.4, 0, 0.0]])
    bs = build_basis(["H", "H"], xyz, "sto-3g")
    S, T, V = (x.cpu().numpy() for x in int1e_device(bs, [1, 1], xyz))
    assert S[0, 0] == pytest.approx(1.0, abs=1e-12)
    assert S[0, 1] == pytest.approx(0.6593, abs=1e-4)
    assert T[0, 0] == pytest.approx(0.7600, abs=1e-4)
    assert T[0, 1] == pytest.approx(0.2365, abs=1e-4)
    assert V[0, 0] == pytest.approx(-1.8804, abs=1e-4)
    assert V[0, 1] == pytest.approx(-1.1948, abs=1e-4)


INT_CASES = [("water", "def2-tzvp"), ("water", "6-31g*"), ("propane", "def2-svp"), ("benzene", "6-31g*"),
             ("cloud", "def2-svp")]


@pytest.mark.parametrize("case", INT_CASES, ids=lambda c: f"{c[0]}-{c[1]}")
def test_int1e_device_matches_host(case):
    from graddft_amd.interface.basis import atomic_numbers, build_basis
    from graddft_amd.interface.integrals import int1e, int1e_device
    name, basis = case
    sym, xyz = _geometry(name)
    Z = atomic_numbers(sym)
    bs = build_basis(sym, xyz, basis)
    _say(f"int1e {name} {basis} nao={bs.nao}")
    host = int1e(bs, Z, xyz)
    dev = [x.cpu().numpy() for x in int1e_device(bs, Z, xyz)]
    again = [x.cpu().numpy() for x in int1e_device(bs, Z, xyz)]
    for label, h, d, d2 in zip("STV", host, dev, again):
        err = np.abs(d - h).max() / np.abs(h).max()
        _say(f"  {label}: max rel diff {err:.2e}")
        assert err <= 1e-12, label
        assert np.array_equal(d, d.T), label
        assert np.array_equal(d, d2), label


GRID_CASES = [("water", 3), ("octane", 2), ("cloud", 1)]


@pytest.mark.parametrize("case", GRID_CASES, ids=lambda c: f"{c[0]}-l{c[1]}")
def test_becke_weights_match_host(case):
    from graddft_amd.interface.basis import atomic_numbers
    from graddft_amd.interface.grids import atomic_grid, becke_weights, becke_weights_device, molecular_grid
    name, level = case
    sym, xyz = _geometry(name)
    Z = atomic_numbers(sym)
    cs, vs, owners, host = [], [], [], []
    for ia, z in enumerate(Z):
        c, v = atomic_grid(int(z), level)
        c = c + xyz[ia]
        cs.append(c)
        vs.append(v)
        owners.append(np.full(len(v), ia, dtype=np.int32))
        host.append(v * becke_weights(c, ia, xyz, Z))
    host = np.concatenate(host)
    _say(f"becke {name} level {level}: {len(host)} points, {len(Z)} atoms")
    dev = becke_weights_device(np.concatenate(cs), np.concatenate(owners), np.concatenate(vs), xyz, Z)
    scale = np.maximum(np.abs(host), 1e-300)
    rel = (np.abs(dev - host) / scale)[np.abs(host) > 1e-15]
    _say(f"  max rel diff {rel.max():.2e}, bitwise equal {np.mean(dev == host):.4f}")
    assert np.array_equal(dev, host)
    ch, wh = molecular_grid(Z, xyz, level)
    cd, wd = molecular_grid(Z, xyz, level, weights_on="device")
    assert np.array_equal(wd, wh) and np.array_equal(cd, ch)


def test_builder_device_inputs_match_host():
    from graddft_amd.interface.builder import molecule_from_geometry
    sym, xyz = _geometry("water")
    _say("builder water def2-svp level 1, device vs host inputs")
    md = molecule_from_geometry(sym, xyz, "def2-svp", grid_level=1)
    mh = molecule_from_geometry(sym, xyz, "def2-svp", grid_level=1, int1e_on="host", grid_on="host")
    for a, b in [(md.s1e, mh.s1e), (md.h1e, mh.h1e)]:
        assert torch.abs(a - b).max().item() <= 1e-12 * torch.abs(b).max().item()
    assert md.grid.weights.shape == mh.grid.weights.shape
    assert torch.equal(md.grid.weights, mh.grid.weights)
    assert torch.abs(md.rdm1 - mh.rdm1).max().item() <= 1e-9
