"""CPU tests of the host-side logic of the non-differentiable drivers in graddft_amd.evaluate
and graddft_amd.molecule (pure tensor code, no device library): damping, level shift,
orbital gradient, aufbau occupations, smearing, the Diis history and the EDIIS / ADIIS
weights, each against the oracle restatement (oracle/scf.py) on the same inputs."""
import numpy as np
import pytest
import torch

from oracle import scf as O


def _sym(rng, *shape):
    a = rng.normal(size=shape)
    return a + np.swapaxes(a, -1, -2)


def test_damping_level_shift_orbital_grad_get_occ():
    from graddft_amd.evaluate import damping, level_shift
    from graddft_amd.molecule import get_occ, orbital_grad
    rng = np.random.default_rng(0)
    n = 8
    A = rng.normal(size=(n, n))
    S = A @ A.T / n + np.eye(n)
    D, F = _sym(rng, n, n), _sym(rng, n, n)
    t = torch.as_tensor
    assert np.abs(damping(t(S), t(D), t(F), 0.4).numpy() - O.damping(S, D, F, 0.4)).max() <= 1e-12
    assert np.abs(level_shift(t(S), t(D), t(F), 0.3).numpy() - O.level_shift(S, D, F, 0.3)).max() <= 1e-12
    e = rng.normal(size=(2, n))
    e[0, 3] = e[0, 5]                                       # a tie: index order decides
    occ = get_occ(t(e), [4, 2]).numpy()
    assert np.array_equal(occ, O.get_occ(e, np.array([4, 2])))
    C, F2 = rng.normal(size=(2, n, n)), _sym(rng, 2, n, n)
    g = orbital_grad(t(C), t(occ), t(F2)).numpy()
    assert np.abs(g - O.orbital_grad(C, occ, F2)).max() <= 1e-12


@pytest.mark.parametrize("kind", ["gaussian", "fermi-dirac"])
def test_smearing_occupations_match_oracle(kind):
    from graddft_amd.evaluate import smearing_occupations
    rng = np.random.default_rng(1)
    e = np.sort(rng.normal(size=(2, 9)), axis=1)
    occ = smearing_occupations(torch.as_tensor(e), 7.0, kind, 0.1).numpy()
    assert np.abs(occ - O.smearing_occ(e, 7.0, kind, 0.1)).max() <= 1e-9
    with pytest.raises(ValueError):
        smearing_occupations(torch.as_tensor(e), 7.0, "cold", 0.1)


def test_diis_history_and_ediis_adiis_weights():
    from graddft_amd.evaluate import Diis
    rng = np.random.default_rng(2)
    n, k = 5, 4
    S = np.eye(n) + 0.05 * _sym(rng, n, n)
    dev = Diis(torch.as_tensor(S), None, 3, "EDIIS")
    data = dev.empty()
    focks = []
    for c in range(k):
        Dk, Fk = _sym(rng, 2, n, n), _sym(rng, 2, n, n)
        focks.append(Fk)
        data = dev.update((torch.as_tensor(Dk), torch.as_tensor(Fk), torch.tensor(-1.0 - c)), data)
    assert data[1].shape[0] == 3 and np.array_equal(data[1][0].numpy(), focks[1])
    assert data[2].tolist() == [-2.0, -3.0, -4.0]
    # F_i = H + D_i makes both costs convex on the simplex: one minimiser
    Ds = _sym(rng, k, 2, n, n)
    Fs = _sym(rng, 2, n, n)[None] + Ds
    Es = rng.normal(size=k)
    ora = O.Diis(S, np.eye(n), 8, "EDIIS")
    c, f = dev.ediis_minimize(torch.as_tensor(Ds), torch.as_tensor(Fs), torch.as_tensor(Es))
    c0, f0 = ora.ediis_minimize(Ds, Fs, Es)
    assert abs(f - f0) <= 1e-9 * max(1.0, abs(f0)) and np.abs(c.numpy() - c0).max() <= 1e-5
    c, f = dev.adiis_minimize(torch.as_tensor(Ds), torch.as_tensor(Fs), 2)
    c0, f0 = ora.adiis_minimize(Ds, Fs, 2)
    assert abs(f - f0) <= 1e-9 * max(1.0, abs(f0)) and np.abs(c.numpy() - c0).max() <= 1e-5
    with pytest.raises(ValueError):
        Diis(torch.as_tensor(S), None, 3, "DIIS").run((torch.as_tensor(Ds[0]), torch.as_tensor(Fs[0]), -1.0),
                                                      dev.empty(), 0)


class _StubLib:
    """Stands in for libgdx.so in host-logic tests: every entry point succeeds, computes nothing."""

    def __getattr__(self, name):
        if name == "gdx_xc_workspace_size":
            def ws(desc, nbytes):
                nbytes._obj.value = 64
                return 0
            return ws
        return lambda *a: 0


def test_densities_pass_invalidates_pending_backward(monkeypatch):
    """XCPlan.densities reuses the forward workspace: a backward of an earlier forward must
    then refuse to run instead of reading overwritten intermediates."""
    from graddft_amd import _lib
    from graddft_amd.functional import XCPlan
    from graddft_amd.molecule import Grid, Molecule
    monkeypatch.setattr(_lib, "load", lambda *a: _StubLib())
    monkeypatch.setattr(torch.cuda, "current_stream", lambda *a: type("S", (), {"cuda_stream": 0})())
    n, N = 4, 8
    z = lambda *s: torch.zeros(s, dtype=torch.float64)  # noqa: E731
    mol = Molecule(grid=Grid(z(N, 3), z(N), z(N)), atom_index=torch.ones(1), nuclear_pos=z(1, 3),
                   ao_planes=z(5, N, 16), rdm1=z(2, n, n), nuclear_repulsion=0.0, h1e=z(n, n),
                   s1e=z(n, n), mo_coeff=z(2, n, n), mo_occ=z(2, n), mo_energy=z(2, n))
    plan = XCPlan(mol, 0, None, "fp64")
    exc = z()
    plan.forward(mol.rdm1, None, exc)
    gen = plan.generation
    plan.backward(mol.rdm1, None, z(), z(2, n, n), None, gen)      # fine right after forward
    plan.densities(mol.rdm1)
    with pytest.raises(RuntimeError, match="overwritten"):
        plan.backward(mol.rdm1, None, z(), z(2, n, n), None, gen)


def test_status_word_decoding():
    """raise_on_status names every flag of the device status word and is silent for 0."""
    import torch
    from graddft_amd import _lib
    from graddft_amd.eigenproblem import raise_on_status
    raise_on_status(torch.zeros(1, dtype=torch.int32), "loop")
    with pytest.raises(_lib.GdxError) as e:
        raise_on_status(torch.tensor([5], dtype=torch.int32), "loop")
    assert "did not converge" in str(e.value) and "tr(D S)" in str(e.value) and "finite" not in str(e.value)


def test_capture_graph_pauses_the_cyclic_collector(monkeypatch):
    """capture_graph collects garbage first and keeps the cyclic collector paused while the
    capture runs (a CUDAGraph freed by it mid-capture would abort the process), restoring the
    collector's state afterwards, also when the captured function raises."""
    import contextlib
    import gc

    import torch

    from graddft_amd import evaluate

    calls = []
    monkeypatch.setattr(torch.cuda, "graph", lambda g, pool=None: contextlib.nullcontext())
    real_collect = gc.collect
    monkeypatch.setattr(gc, "collect", lambda *a: calls.append("collect") or real_collect(*a))
    seen = {}

    def body():
        seen["enabled"] = gc.isenabled()
        return 7

    assert gc.isenabled()
    assert evaluate.capture_graph(object(), body) == 7
    assert calls == ["collect"] and seen["enabled"] is False and gc.isenabled()

    def boom():
        raise RuntimeError("x")

    with pytest.raises(RuntimeError):
        evaluate.capture_graph(object(), boom)
    assert gc.isenabled()


def test_custom_functional_host_logic(monkeypatch):
    """CustomFunctional: argument checks, a plan of its own that refuses the MLP entry points,
    E = sum_r clip(w) clip(sum_f c_f clip(e_f)) over the device inputs, and the cotangents it
    hands back (dE/de_f = clip(w) c_f where the clipped sum is kept, dE/dx through the user's
    function) -- with the device calls replaced by fixed inputs on the CPU."""
    from graddft_amd import _lib
    from graddft_amd.functional import CustomFunctional, XCPlan
    from graddft_amd.molecule import Grid, Molecule
    monkeypatch.setattr(_lib, "load", lambda *a: _StubLib())
    monkeypatch.setattr(torch.cuda, "current_stream", lambda *a: type("S", (), {"cuda_stream": 0})())
    with pytest.raises(TypeError):
        CustomFunctional(3.0)
    with pytest.raises(ValueError):
        CustomFunctional(lambda p, x: x, correlation="lyp")
    n, N = 4, 6
    z = lambda *s: torch.zeros(s, dtype=torch.float64)  # noqa: E731
    rng = np.random.default_rng(3)
    w = torch.tensor(rng.uniform(0.1, 1.0, N))
    w[2] = 1e-31                                           # clipped weight
    mol = Molecule(grid=Grid(z(N, 3), w, z(N)), atom_index=torch.ones(1), nuclear_pos=z(1, 3),
                   ao_planes=z(5, N, 16), rdm1=z(2, n, n), nuclear_repulsion=0.0, h1e=z(n, n),
                   s1e=z(n, n), mo_coeff=z(2, n, n), mo_occ=z(2, n), mo_energy=z(2, n))
    e = torch.tensor(rng.normal(size=(N, 4)))
    e[4, :] = 0.0                                          # clipped energy density sum
    x = torch.tensor(rng.normal(size=(N, 6)))
    seen = {}

    def fake_inputs(self, rdm1):
        self.generation += 1
        self.fwd_generation = self.generation
        return e.clone(), x.clone()

    def fake_vxc(self, g_e, g_x, gen):
        seen["g_e"], seen["g_x"] = g_e.clone(), g_x.clone()
        return torch.full((2, n, n), float(g_e.sum() + g_x.sum()), dtype=torch.float64)

    monkeypatch.setattr(XCPlan, "custom_inputs", fake_inputs)
    monkeypatch.setattr(XCPlan, "custom_vxc", fake_vxc)
    A = torch.tensor(rng.normal(scale=0.3, size=(6, 4)), requires_grad=True)
    f = CustomFunctional(lambda p, xx: 1.0 + torch.tanh(xx @ p))
    with pytest.raises(NotImplementedError):
        f.plan(mol, "fp32")
    plan = f.plan(mol)
    assert plan.custom and f.plan(mol) is plan
    with pytest.raises(RuntimeError, match="custom"):
        plan.forward(mol.rdm1, None, z())
    D = mol.rdm1.clone().requires_grad_(True)
    E = f.xc_energy(A, mol, rdm1=D)
    c = 1.0 + np.tanh(x.numpy() @ A.detach().numpy())
    eps = (c * e.numpy()).sum(axis=1)
    keep = (np.abs(eps) > 1e-30) & (np.abs(w.numpy()) > 1e-30)
    gD, gA = torch.autograd.grad(E, [D, A])
    assert float(E.detach()) == pytest.approx(float((w.numpy() * eps)[keep].sum()), rel=1e-14)
    ge_ref = np.where(keep[:, None], w.numpy()[:, None] * c, 0.0)
    assert np.allclose(seen["g_e"].numpy(), ge_ref, rtol=1e-14, atol=0)
    # dE/dx_j = sum_f clip(w) e_f dc_f/dx_j, dc/dz = 1 - tanh^2
    dz = np.where(keep[:, None], w.numpy()[:, None] * e.numpy() * (1 - np.tanh(x.numpy() @ A.detach().numpy()) ** 2), 0)
    assert np.allclose(seen["g_x"].numpy(), dz @ A.detach().numpy().T, rtol=1e-12, atol=1e-15)
    assert np.allclose(gA.numpy(), x.numpy().T @ dz, rtol=1e-12, atol=1e-15)
    assert float(gD[0, 0, 0]) == pytest.approx(float(seen["g_e"].sum() + seen["g_x"].sum()), rel=1e-14)
