"""CPU pins of the scf_loop / simple_scf_loop restatements in oracle/scf.py
(grad_dft/evaluate.py:131-592 and 1211-1545, grad_dft/molecule.py:344-373).

Each piece is checked against what the mathematics fixes rather than a retyped formula:
damping scales the occupied-virtual blocks of F by 1/(1+f) in an S-orthonormal orbital
basis; a level shift raises the virtual eigenvalues by f and leaves the occupied ones;
the orbital gradient equals pyscf's compact C_vir^T F C_occ placed at (virtual, occupied)
and vanishes at self-consistency; CDIIS has the two-vector closed form and equals the
(independently pinned) JittableDiis coefficients when every slot is active; the EDIIS /
ADIIS costs equal their published forms (JCP 116, 8255; JCP 132, 054109) written with
explicit density / Fock differences, and the BFGS weights are no worse than a brute-force
search of the simplex; and every loop variant reaches the self-consistent energy of the
oracle diff_scf_loop (the common fixed point)."""
import itertools
import math

import numpy as np
import pytest

from oracle import scf as O


def _orbital_problem(n, nocc, seed):
    rng = np.random.default_rng(seed)
    A = rng.normal(size=(n, n))
    S = A @ A.T / n + np.eye(n)
    F = rng.normal(size=(n, n))
    F = F + F.T
    w, C = O.safe_general_eigh(F, S)
    P = np.zeros(n)
    P[:nocc] = 1.0
    D = C @ np.diag(P) @ C.T
    return rng, S, F, C, w, P, D


def test_damping_scales_occupied_virtual_blocks():
    rng, S, F, C, w, P, D = _orbital_problem(7, 3, 0)
    G = rng.normal(size=(7, 7))
    G = G + G.T
    f = 0.35
    Gm = C.T @ G @ C
    Dm = C.T @ O.damping(S, D, G, f) @ C
    o, v = slice(0, 3), slice(3, 7)
    assert np.allclose(Dm[o, o], Gm[o, o], atol=1e-12)
    assert np.allclose(Dm[v, v], Gm[v, v], atol=1e-12)
    assert np.allclose(Dm[o, v], Gm[o, v] / (1 + f), atol=1e-12)
    assert np.allclose(Dm[v, o], Gm[v, o] / (1 + f), atol=1e-12)
    # at self-consistency (F commutes with D in the S metric) damping changes nothing
    assert np.allclose(O.damping(S, D, F, f), F, atol=1e-11)


def test_level_shift_raises_virtual_eigenvalues():
    rng, S, F, C, w, P, D = _orbital_problem(8, 3, 1)
    f = 0.5
    w2, C2 = O.safe_general_eigh(O.level_shift(S, D, F, f), S)
    assert np.allclose(np.sort(w2), np.sort(w + f * (1 - P)), atol=1e-11)


def test_orbital_grad_is_compact_virtual_occupied_block():
    rng = np.random.default_rng(2)
    n = 6
    C = rng.normal(size=(2, n, n))
    F = rng.normal(size=(2, n, n))
    F = F + F.transpose(0, 2, 1)
    occ = np.array([[1, 0, 1, 0, 0, 1], [1, 1, 0, 0, 0, 0]], dtype=float)
    g = O.orbital_grad(C, occ, F)
    ref = np.zeros((n, n))
    for s in range(2):
        vi = np.flatnonzero(occ[s] == 0)
        oi = np.flatnonzero(occ[s] > 0)
        ref[np.ix_(vi, oi)] += C[s][:, vi].T @ F[s] @ C[s][:, oi]      # pyscf hf.get_grad block
    assert np.allclose(g, ref, atol=1e-12)
    # vanishes at self-consistency
    _, S, F1, C1, w1, P1, D1 = _orbital_problem(6, 2, 3)
    Cs = np.stack([C1, C1])
    g0 = O.orbital_grad(Cs, np.stack([P1, P1]), np.stack([F1, F1]))
    assert np.abs(g0).max() < 1e-11


def _errors(k, n, seed):
    rng = np.random.default_rng(seed)
    E = rng.normal(size=(k, 2, n, n))
    return E - E.transpose(0, 1, 3, 2)


def test_cdiis_two_vector_closed_form_and_full_buffer_equivalence():
    errs = _errors(2, 4, 4)
    x = O.Diis.cdiis_minimize(errs)
    for s in range(2):
        g = np.einsum("ikl,jkl->ij", errs[:, s], errs[:, s])
        x1 = (g[1, 1] - g[0, 1]) / (g[0, 0] - 2 * g[0, 1] + g[1, 1])
        assert x[s, 0] == pytest.approx(x1, rel=1e-10)
        assert x[s, 1] == pytest.approx(1 - x1, rel=1e-10)
    errs = _errors(10, 8, 5)       # 10 independent antisymmetric errors (28 dimensions)
    x = O.Diis.cdiis_minimize(errs)
    xj = O.JittableDiis(np.eye(8), 10).coefficients(errs, 12)
    assert np.allclose(x, xj, atol=1e-10)
    assert np.allclose(x.sum(axis=1), 1.0)


def _density_fock_history(k, n, seed):
    rng = np.random.default_rng(seed)
    D = rng.normal(size=(k, 2, n, n))
    D = D + D.transpose(0, 1, 3, 2)
    F = rng.normal(size=(k, 2, n, n))
    F = F + F.transpose(0, 1, 3, 2)
    E = rng.normal(size=k)
    return D, F, E


def _tr(A, B):
    return sum(np.trace(A[s] @ B[s]) for s in range(A.shape[0]))


def _simplex_grid(k, steps):
    for idx in itertools.product(range(steps + 1), repeat=k - 1):
        if sum(idx) <= steps:
            c = np.array(list(idx) + [steps - sum(idx)], dtype=float) / steps
            yield c


def test_ediis_cost_and_minimum():
    D, F, E = _density_fock_history(3, 3, 6)
    k = 3

    def cost(c):   # E(c) = sum c_i E_i - sum_ij c_i c_j tr((D_i - D_j)(F_i - F_j))
        return c @ E - sum(c[i] * c[j] * _tr(D[i] - D[j], F[i] - F[j]) for i in range(k) for j in range(k))

    diis = O.Diis(np.eye(3), np.eye(3), 8, "EDIIS")
    c, fun = diis.ediis_minimize(D, F, E)
    assert c.min() >= 0 and c.sum() == pytest.approx(1.0)
    assert fun == pytest.approx(cost(c), rel=1e-10, abs=1e-12)
    best = min(cost(g) for g in _simplex_grid(k, 60))
    assert fun <= best + 1e-10


def test_adiis_cost_and_minimum():
    D, F, E = _density_fock_history(3, 3, 7)
    k, nw = 3, 2

    def cost(c):   # 2 sum c_i <D_i - D_n | F_n> + sum c_i c_j <D_i - D_n | F_j - F_n>
        lin = sum(c[i] * _tr(D[i] - D[nw], F[nw]) for i in range(k))
        quad = sum(c[i] * c[j] * _tr(D[i] - D[nw], F[j] - F[nw]) for i in range(k) for j in range(k))
        return 2 * lin + quad

    diis = O.Diis(np.eye(3), np.eye(3), 8, "ADIIS")
    c, fun = diis.adiis_minimize(D, F, nw)
    assert c.min() >= 0 and c.sum() == pytest.approx(1.0)
    assert fun == pytest.approx(cost(c), rel=1e-10, abs=1e-12)
    best = min(cost(g) for g in _simplex_grid(k, 60))
    assert fun <= best + 1e-10


def test_diis_history_trims_to_max_and_first_run_returns_fock():
    n = 3
    diis = O.Diis(np.eye(n), np.eye(n), 3, "CDIIS")
    data = ([], [], [], [])
    rng = np.random.default_rng(8)
    focks = []
    for c in range(5):
        Fk = rng.normal(size=(2, n, n))
        Fk = Fk + Fk.transpose(0, 2, 1)
        Dk = rng.normal(size=(2, n, n))
        Dk = Dk + Dk.transpose(0, 2, 1)
        focks.append(Fk)
        out, data = diis.run(Dk, Fk, -1.0 - c, data, c)
        if c == 0:
            assert np.array_equal(out, Fk)
    assert len(data[1]) == 3 and np.array_equal(data[1][0], focks[2])


def test_smearing_occupations():
    rng = np.random.default_rng(9)
    e = np.sort(rng.normal(size=(2, 8)), axis=1)
    for kind in ("gaussian", "fermi-dirac"):
        occ = O.smearing_occ(e, 6.0, kind, 0.05)
        assert occ.sum() == pytest.approx(6.0, abs=1e-8)
        assert np.all((occ >= 0) & (occ <= 1))
    # a large gap and a small width give the aufbau occupations
    e = np.array([[-2.0, -1.5, 1.0, 2.0], [-2.0, -1.5, 1.0, 2.0]])
    occ = O.smearing_occ(e, 4.0, "fermi-dirac", 0.01)
    assert np.allclose(occ, [[1, 1, 0, 0], [1, 1, 0, 0]], atol=1e-12)


@pytest.fixture(scope="module")
def water():
    from graddft_amd.interface.basis import build_basis, nuclear_repulsion
    from graddft_amd.interface.builder import core_guess
    from graddft_amd.interface.grids import molecular_grid
    from graddft_amd.interface.integrals import int1e, int2e_packed, unpack_eri
    from oracle.ao import eval_ao
    from oracle.xc import to_inputs, xc_value_and_grads
    xyz = np.array([[0, 0, 0], [1.81, 0, 0], [1.81 * math.cos(1.824), 1.81 * math.sin(1.824), 0]])
    bs = build_basis(["O", "H", "H"], xyz, "def2-svp")
    S, T, V = int1e(bs, [8, 1, 1], xyz)
    eri = unpack_eri(int2e_packed(bs), bs.nao)
    D, C, e, occ = core_guess(T + V, S, 5)
    rdm1 = np.stack([D, D])
    K0 = np.stack([O.exchange_potential(rdm1[s], eri) for s in range(2)])
    inp = O.SCFInputs(T + V, S, eri, nuclear_repulsion([8, 1, 1], xyz), K0, 0.2)
    c, w = molecular_grid([8, 1, 1], xyz, 1)
    xin = to_inputs(*eval_ao(bs, c), w)
    xc = lambda d: xc_value_and_grads("b3lyp", xin, d)[:2]  # noqa: E731
    occ2 = np.stack([occ, occ])
    ref = O.diff_scf_loop(xc, inp, rdm1, occ2, 25)
    return dict(inp=inp, xc=xc, rdm1=rdm1, occ=occ2, C=np.stack([C, C]), e=np.stack([e, e]), E=ref.energy,
                dE=abs(ref.energies[-1] - ref.energies[-2]))


@pytest.mark.parametrize("opts", [
    dict(),
    dict(diis_start_cycle=3, damp_factor=(0.3, 0.3), level_shift_factor=(0.2, 0.2)),
    dict(diis_method="EDIIS2"),
    dict(diis_method="ADIIS", e_conv=1e-9, g_conv=1e-9, cycles=60),   # ADIIS alone converges slowly
], ids=["cdiis", "damp_shift", "ediis2", "adiis"])
def test_scf_loop_reaches_the_self_consistent_energy(water, opts):
    assert water["dE"] < 1e-9
    st, cycles = O.scf_loop(water["xc"], water["inp"], water["rdm1"], water["occ"], **opts)
    assert cycles <= opts.get("cycles", 49)
    assert abs(st.energy - water["E"]) < 1e-7, (st.energy, water["E"])
    S = water["inp"].s1e
    for s in range(2):
        assert np.abs(st.rdm1[s] @ S @ st.rdm1[s] - st.rdm1[s]).max() < 1e-9


def test_simple_scf_loop_first_cycle_and_fixed_point(water):
    st = O.simple_scf_loop(water["xc"], water["inp"], water["rdm1"], water["C"], water["e"], water["occ"], cycles=1)
    e0, _ = O.predict(water["xc"], water["inp"], water["rdm1"])
    assert st.energy == e0                       # cycle 0 predicts at the given density
    st = O.simple_scf_loop(water["xc"], water["inp"], water["rdm1"], water["C"], water["e"], water["occ"],
                           mixing_factor=0.5, cycles=40)
    assert abs(st.energy - water["E"]) < 1e-6, (st.energy, water["E"])


@pytest.fixture(scope="module")
def water_sto3g():
    import torch
    from graddft_amd.interface.basis import build_basis, nuclear_repulsion
    from graddft_amd.interface.builder import core_guess
    from graddft_amd.interface.grids import molecular_grid
    from graddft_amd.interface.integrals import int1e, int2e_packed, unpack_eri
    from oracle import scf_grad as G
    from oracle import xc as X
    from oracle.ao import eval_ao
    xyz = np.array([[0, 0, 0], [1.81, 0, 0], [1.81 * math.cos(1.824), 1.81 * math.sin(1.824), 0]])
    bs = build_basis(["O", "H", "H"], xyz, "sto-3g")
    S, T, V = int1e(bs, [8, 1, 1], xyz)
    eri = unpack_eri(int2e_packed(bs), bs.nao)
    D, C, e, occ = core_guess(T + V, S, 5)
    rdm1 = np.stack([D, D])
    K0 = np.stack([O.exchange_potential(rdm1[s], eri) for s in range(2)])
    inp = O.SCFInputs(T + V, S, eri, nuclear_repulsion([8, 1, 1], xyz), K0, 0.2)
    c, w = molecular_grid([8, 1, 1], xyz, 1)
    xin = X.to_inputs(*eval_ao(bs, c), w)
    xcv = lambda d: X.xc_value_and_grads("b3lyp", xin, d)[:2]  # noqa: E731
    E_scf = O.diff_scf_loop(xcv, inp, rdm1, np.stack([occ, occ]), 25).energy
    return dict(inp=G.TorchSCFInputs.from_numpy(inp), S=S, E_scf=E_scf,
                xce=lambda d: X.xc_energy_from_rdm1("b3lyp", xin, d),
                C=torch.tensor(np.stack([C, C])), occ=torch.tensor(np.stack([occ, occ])))


@pytest.mark.parametrize("kind", ["PCA", "Cholesky", "ZCA"])
def test_whitening_matrices(water_sto3g, kind):
    from oracle import scf_grad as G
    S = water_sto3g["S"]
    D = G.whitening_matrix(S, kind)
    assert np.allclose(D.T @ D, np.linalg.inv(S), atol=1e-10)
    assert np.allclose(D @ S @ D.T, np.eye(len(S)), atol=1e-10)       # orbitals (Q D)^T are S-orthonormal


def test_orbital_energy_gradient_and_variational_bound(water_sto3g):
    """dE/dW through the QR factor against central differences; every optimizer energy is the
    energy of an idempotent density, so none lies below the self-consistent minimum, and the
    optimizer descends."""
    import torch
    from oracle import scf_grad as G
    w = water_sto3g
    D = torch.tensor(G.whitening_matrix(w["S"], "PCA"))
    W = torch.einsum("sji,jk->sik", w["C"], torch.linalg.inv(D))
    # a generic point (no Householder pivot near zero, which flips a QR column sign), the same
    # for both spins (zeta = 0: the spin-polarisation term's log2(1 -+ zeta) stays smooth)
    P0 = torch.tensor(np.random.default_rng(1).normal(size=W.shape[1:]))
    W = W + 0.01 * torch.stack([P0, P0])
    Wg = W.clone().requires_grad_(True)
    E = G.orbital_energy(w["xce"], w["inp"], Wg, D, w["occ"])
    g, = torch.autograd.grad(E, [Wg])
    P = torch.tensor(np.random.default_rng(1).normal(size=W.shape))
    h = 1e-5
    fd = (G.orbital_energy(w["xce"], w["inp"], W + h * P, D, w["occ"])
          - G.orbital_energy(w["xce"], w["inp"], W - h * P, D, w["occ"])) / (2 * h)
    assert float((g * P).sum()) == pytest.approx(float(fd), rel=1e-6)
    es = G.mol_orb_optimizer(w["xce"], w["inp"], w["C"], w["occ"], 0.01, 40, e_conv=None)
    assert min(es) >= w["E_scf"] - 1e-8
    assert es[-1] < es[0] - 0.1


def test_householder_q_follows_lapack_signs():
    """The product's QR (used by the orbital optimizers) reproduces LAPACK's Q, column signs
    included, and its gradient: random matrices, and orthogonal ones with exact zero pivots
    (no reflection, the sign is kept)."""
    import torch
    from graddft_amd.eigenproblem import householder_q
    rng = np.random.default_rng(0)
    for n in (1, 2, 7, 30):
        W = torch.tensor(rng.normal(size=(2, n, n)))
        assert (householder_q(W) - torch.linalg.qr(W)[0]).abs().max().item() < 1e-13
    P = torch.tensor([[0.0, 1, 0, 0], [1, 0, 0, 0], [0, 0, -1, 0], [0, 0, 0, 1]])
    W = torch.stack([torch.eye(4, dtype=torch.float64), P.double()])
    assert torch.equal(householder_q(W), torch.linalg.qr(W)[0])
    W = torch.tensor(rng.normal(size=(2, 6, 6)), requires_grad=True)
    Pm = torch.tensor(rng.normal(size=(2, 6, 6)))
    g, = torch.autograd.grad((householder_q(W) * Pm).sum(), [W])
    g0, = torch.autograd.grad((torch.linalg.qr(W)[0] * Pm).sum(), [W])
    assert (g - g0).abs().max().item() < 1e-12
