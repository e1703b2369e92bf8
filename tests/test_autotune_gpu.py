"""The V_xc tile autotuner (gdx_xc_autotune): a measured choice is one of the candidates and is
kept for the shape; the build gives the same E_xc, dE/dD and dE/dtheta with either tile pinned
(64 or 128: different tilings and point chunkings of the same sums).

Cases: propane def2-SVP (n_pad 96: one 128 tile or 2 x 2 tiles of 64) and naphthalene
def2-SVP (n_pad 192: 2 x 2 tiles of 128 or 3 x 3 of 64), level-1 grids; B3LYP (restricted,
one spin) and the neural functional (two spins); FP64 and FP32.  Bars between the two tiles:
FP64 1e-12 relative, FP32 1e-6 relative.
"""
import ctypes

import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu


def _say(msg):
    print(msg, flush=True)


def _mol(name):
    from graddft_amd.interface import geometry
    from graddft_amd.interface.builder import molecule_from_geometry
    rng = np.random.default_rng(17)
    sym, xyz = geometry.alkane(rng, 3) if name == "propane" else geometry.naphthalene(rng)
    return molecule_from_geometry(sym, xyz, "def2-svp", 1)


def _pin(plan, vt):
    from graddft_amd import _lib
    v = ctypes.c_int32(vt)
    st = torch.cuda.current_stream().cuda_stream
    _lib.check(_lib.load().gdx_xc_autotune(st, ctypes.byref(plan.desc), plan._ao().data_ptr(),
                                           plan.workspace.data_ptr(), ctypes.byref(v)), "gdx_xc_autotune")


def _run(plan, mol, theta):
    e = torch.empty((), dtype=torch.float64, device="cuda")
    one = torch.ones((), dtype=torch.float64, device="cuda")
    dD = torch.empty_like(mol.rdm1)
    dth = None if theta is None else torch.empty_like(theta)
    plan.forward(mol.rdm1.contiguous(), theta, e)
    plan.backward(mol.rdm1.contiguous(), theta, one, dD, dth, plan.generation)
    torch.cuda.synchronize()
    return e.item(), dD.cpu().numpy(), None if dth is None else dth.cpu().numpy()


def _rel(a, b):
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
@pytest.mark.parametrize("name", ["propane", "naphthalene"])
def test_vxc_tiles_agree_and_tuner_picks_a_candidate(name, precision, monkeypatch):
    from graddft_amd import B3LYP, MLPSpec, NeuralFunctional
    from graddft_amd import functional as fmod
    mol = _mol(name)
    _say(f"case: V_xc tile 64 vs 128, {name} (n_pad {mol.nao_pad}), {precision}")
    nf = NeuralFunctional(MLPSpec(nc=6, h=32, n_hidden=3, nd=4))
    for f, theta, m in ((B3LYP, None, mol), (nf, nf.init(seed=2), mol.replace(restricted=False))):
        monkeypatch.setattr(fmod, "AUTOTUNE", False)
        plan = fmod.XCPlan(m, f.kind, f.spec, precision, f.coef, f.correlation)
        out = {}
        for vt in (64, 128):
            _pin(plan, vt)
            out[vt] = _run(plan, m, theta)
        tol = 1e-12 if precision == "fp64" else 1e-6
        e64, d64, t64 = out[64]
        e128, d128, t128 = out[128]
        _say(f"  kind {f.kind}: dE {abs(e64 - e128):.2e} dD rel {_rel(d64, d128):.2e}")
        assert abs(e64 - e128) <= tol * abs(e128)
        assert _rel(d64, d128) <= tol
        if theta is not None:
            assert _rel(t64, t128) <= tol
        # a plan created with measuring on keeps the choice already recorded for its shape
        monkeypatch.setattr(fmod, "AUTOTUNE", True)
        plan2 = fmod.XCPlan(m, f.kind, f.spec, precision, f.coef, f.correlation)
        assert plan2.vxc_tile == 128
