"""GPU parity of the SCF path (J/K from the packed ERI, Cholesky, eigensolver, DIIS,
energy_predictor, diff_scf_loop) against the FP64 oracle on identical inputs."""
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


def _dense_eri(mol):
    from graddft_amd.interface.integrals import unpack_eri
    return unpack_eri(mol.rep_tensor.cpu().numpy(), mol.nao)


def test_coulomb_and_exchange_match_oracle():
    _say("case: J/K water/def2-svp")
    from graddft_amd.scf import coulomb, exchange
    from oracle.scf import coulomb_potential, exchange_potential
    mol = _water()
    rng = np.random.default_rng(0)
    n = mol.nao
    A = rng.normal(size=(2, n, n))
    D = A + A.transpose(0, 2, 1)
    Dt = torch.as_tensor(D, device="cuda")
    J = coulomb(mol, Dt).cpu().numpy()
    K = exchange(mol, Dt).cpu().numpy()
    eri = _dense_eri(mol)
    J0 = coulomb_potential(D.sum(0), eri)
    K0 = np.stack([exchange_potential(D[s], eri) for s in range(2)])
    assert np.abs(J - J0).max() <= 1e-11 * np.abs(J0).max()
    assert np.abs(K - K0).max() <= 1e-11 * np.abs(K0).max()


@pytest.mark.parametrize("n", [30, 90, 120, 155, 180, 202])
def test_coulomb_packed_matches_dense_product(n):
    """gdx_coulomb_packed on random packed matrices against a row-by-row host product: the small
    tiles (n = 30), the 256-tiles with the interior fast path (n = 90, 120: P < 12000) and the
    wave-per-row reader (n = 155, 180, 202), each with ragged edge tiles."""
    _say(f"case: coulomb packed vs dense product, n={n}")
    import ctypes
    from graddft_amd import _lib
    lib = _lib.load()
    rng = np.random.default_rng(100 + n)
    P = n * (n + 1) // 2
    eri = rng.normal(size=P * (P + 1) // 2)
    D = rng.normal(size=(2, n, n))
    Ds = D.sum(0)
    i, j = np.tril_indices(n)
    Dt = np.where(i == j, Ds[i, j], Ds[i, j] + Ds[j, i])
    Jt = np.zeros(P)
    for I in range(P):
        seg = eri[I * (I + 1) // 2: I * (I + 1) // 2 + I + 1]
        Jt[I] += seg @ Dt[: I + 1]
        Jt[:I] += seg[:I] * Dt[I]
    J0 = np.zeros((n, n))
    J0[i, j] = Jt
    J0[j, i] = Jt
    e = torch.as_tensor(eri, device="cuda")
    d = torch.as_tensor(D, device="cuda")
    J = torch.full((n, n), float("nan"), dtype=torch.float64, device="cuda")
    nb = ctypes.c_size_t(0)
    _lib.check(lib.gdx_coulomb_workspace_size(n, ctypes.byref(nb)), "gdx_coulomb_workspace_size")
    ws = torch.empty(nb.value // 8, dtype=torch.float64, device="cuda")
    st = torch.cuda.current_stream().cuda_stream
    _lib.check(lib.gdx_coulomb_packed(st, n, e.data_ptr(), d.data_ptr(), J.data_ptr(), ws.data_ptr()),
               "gdx_coulomb_packed")
    J = J.cpu().numpy()
    err = np.abs(J - J0).max() / np.abs(J0).max()
    _say(f"  max rel diff {err:.2e}")
    assert err <= 1e-13


@pytest.mark.parametrize("n", [7, 43, 64, 65, 102, 180, 202, 224, 225])
@pytest.mark.parametrize("spectrum", ["random", "degenerate"])
def test_eigh_matches_lapack(n, spectrum):
    """Jacobi (n <= ~80 and n > 224) and the tridiagonal path (Householder, bisection, inverse
    iteration, back-transformation; n in (~80, 224]).  'degenerate': every eigenvalue twice
    (Q diag(l) Q^T with l in pairs), the case that needs the cluster re-orthogonalisation."""
    _say(f"case: eigh n={n} {spectrum}")
    from graddft_amd.eigenproblem import eigh
    rng = np.random.default_rng(n)
    if spectrum == "random":
        A = rng.normal(size=(2, n, n))
        A = A + A.transpose(0, 2, 1)
    else:
        A = np.empty((2, n, n))
        for s in range(2):
            Q = np.linalg.qr(rng.normal(size=(n, n)))[0]
            lam = np.repeat(rng.normal(size=(n + 1) // 2), 2)[:n]
            A[s] = (Q * lam) @ Q.T
            A[s] = 0.5 * (A[s] + A[s].T)
    w, V = eigh(torch.as_tensor(A, device="cuda"))
    w, V = w.cpu().numpy(), V.cpu().numpy()
    for s in range(2):
        w0 = np.linalg.eigvalsh(A[s])
        assert np.abs(w[s] - w0).max() <= 1e-12 * np.abs(w0).max()
        R = A[s] @ V[s] - V[s] * w[s]
        assert np.abs(R).max() <= 1e-11 * np.abs(w0).max()
        assert np.abs(V[s].T @ V[s] - np.eye(n)).max() <= 1e-12


@pytest.mark.parametrize("n", [43, 102, 180])
def test_eigh_warm_start(n):
    """A warm start from the eigenvectors of a nearby matrix gives the same decomposition
    (the SCF loop passes the previous cycle's eigenvectors).  n = 43 runs the Jacobi sweeps with
    every buffer in LDS, and the warm start cuts the sweeps; 102 and 180 run the tridiagonal
    path, which takes no start (sweeps reported 0) and must give the same result with one."""
    _say(f"case: eigh warm start n={n}")
    import ctypes
    from graddft_amd import _lib
    rng = np.random.default_rng(7)
    A = rng.normal(size=(2, n, n))
    A = A + A.transpose(0, 2, 1)
    dA = rng.normal(size=(2, n, n)) * 1e-3
    A1 = A + dA + dA.transpose(0, 2, 1)
    V0 = np.stack([np.linalg.eigh(A[s])[1] for s in range(2)])
    lib = _lib.load()
    nbytes = ctypes.c_size_t(0)
    _lib.check(lib.gdx_eigh_workspace_size(2, n, ctypes.byref(nbytes)), "ws")
    ws = torch.empty(max(int(nbytes.value), 8), dtype=torch.uint8, device="cuda")
    At = torch.as_tensor(A1, device="cuda")
    V0t = torch.as_tensor(V0, device="cuda")
    out = {}
    for name, v0 in (("cold", None), ("warm", V0t)):
        w = torch.empty((2, n), dtype=torch.float64, device="cuda")
        V = torch.empty_like(At)
        sw = torch.zeros(2, dtype=torch.int32, device="cuda")
        _lib.check(lib.gdx_eigh_batched(torch.cuda.current_stream().cuda_stream, 2, n, At.data_ptr(),
                                        0 if v0 is None else v0.data_ptr(), w.data_ptr(), V.data_ptr(),
                                        ws.data_ptr(), 40, 1e-14, sw.data_ptr()), "eigh")
        out[name] = (w.cpu().numpy(), V.cpu().numpy(), sw.cpu().numpy())
    for s in range(2):
        w0 = np.linalg.eigvalsh(A1[s])
        w, V, _ = out["warm"]
        assert np.abs(w[s] - w0).max() <= 1e-12 * np.abs(w0).max()
        assert np.abs(A1[s] @ V[s] - V[s] * w[s]).max() <= 1e-11 * np.abs(w0).max()
        assert np.abs(V[s].T @ V[s] - np.eye(n)).max() <= 1e-12
    if n <= 64:
        assert out["warm"][2].max() < out["cold"][2].min(), (out["warm"][2], out["cold"][2])
    else:
        assert out["warm"][2].max() == 0 and out["cold"][2].max() == 0


def test_cholesky_inverse():
    _say("case: cholesky water/def2-tzvp overlap")
    from graddft_amd.eigenproblem import cholesky_inv
    mol = _water("def2-tzvp")
    S = mol.s1e.cpu().numpy()
    L, Li = cholesky_inv(mol.s1e)
    L, Li = L.cpu().numpy(), Li.cpu().numpy()
    assert np.abs(L - np.linalg.cholesky(S)).max() <= 1e-12
    assert np.abs(Li @ L - np.eye(len(S))).max() <= 1e-10


def test_diis_solve_matches_inverse():
    _say("case: diis solve")
    from graddft_amd import _lib
    rng = np.random.default_rng(3)
    m1 = 11
    B = rng.normal(size=(2, m1, m1))
    B = B + B.transpose(0, 2, 1)
    Bt = torch.as_tensor(B, device="cuda")
    x = torch.empty((2, m1), dtype=torch.float64, device="cuda")
    _lib.check(_lib.load().gdx_diis_solve(torch.cuda.current_stream().cuda_stream, 2, m1, Bt.data_ptr(),
                                          x.data_ptr()), "diis")
    x0 = np.stack([np.linalg.inv(B[s])[:, 0] for s in range(2)])
    assert np.abs(x.cpu().numpy() - x0).max() <= 1e-10 * np.abs(x0).max()


def test_diis_run_matches_oracle_with_gradients():
    """JittableDiis.run (update + gdx_diis_coefficients) over 14 cycles -- slots filling, the
    dropped write at cycle == max_diis, then shifting -- against the oracle's torch restatement:
    extrapolated Fock matrices, and the gradient of a weighted sum of them w.r.t. every input
    Fock and density matrix (reverse pass through the coefficient Function)."""
    _say("case: diis run vs oracle, 14 cycles, n=6, forward and reverse")
    from graddft_amd.evaluate import JittableDiis
    from oracle import scf_grad as G
    rng = np.random.default_rng(12)
    n, cycles = 6, 14
    X = rng.normal(size=(n, n))
    S = X @ X.T + n * np.eye(n)
    focks = [rng.normal(size=(2, n, n)) for _ in range(cycles)]
    focks = [f + f.transpose(0, 2, 1) for f in focks]
    dens = [rng.normal(size=(2, n, n)) * 0.1 for _ in range(cycles)]
    dens = [d + d.transpose(0, 2, 1) for d in dens]
    W = [rng.normal(size=(2, n, n)) for _ in range(cycles)]

    def run(dev, ours):
        Ft = [torch.tensor(f, device=dev, requires_grad=True) for f in focks]
        Dt = [torch.tensor(d, device=dev, requires_grad=True) for d in dens]
        St = torch.tensor(S, device=dev)
        if ours:
            diis = JittableDiis(St, max_diis=10)
            data = diis.zeros(n, dev)
        else:
            diis = G.JittableDiis(St, 10)
            data = (torch.zeros((10, 2, n, n), dtype=torch.float64), torch.zeros((10, 2, n, n), dtype=torch.float64))
        outs, loss = [], 0.0
        for c in range(cycles):
            if ours:
                F, data = diis.run(Dt[c], Ft[c], torch.tensor(0.1 * c, device=dev, dtype=torch.float64), data, c)
            else:
                F, data = diis.run(Dt[c], Ft[c], data, c)
            outs.append(F.detach().cpu().numpy())
            loss = loss + (F * torch.tensor(W[c], device=dev)).sum()
        # the inputs of cycle == max_diis are dropped by the reference's buffer update: no gradient
        gF = torch.autograd.grad(loss, Ft + Dt, allow_unused=True)
        return outs, [np.zeros((2, n, n)) if g is None else g.cpu().numpy() for g in gF]

    o1, g1 = run("cuda", True)
    o0, g0 = run("cpu", False)
    for c in range(cycles):
        assert np.abs(o1[c] - o0[c]).max() <= 1e-10 * max(1.0, np.abs(o0[c]).max()), c
    for a, b in zip(g1, g0):
        assert np.abs(a - b).max() <= 1e-8 * max(1.0, np.abs(b).max())


@pytest.mark.parametrize("n", [43, 50, 102, 180, 202])
def test_scf_step_matches_unfused_path(n):
    """gdx_scf_step (extrapolation, eigenproblem, back-transform, density; n = 43 and 50 in
    one Jacobi workgroup with every buffer in LDS, 102, 180 and 202 through the tiled GEMMs
    and the tridiagonal eigensolver) against the separate path: einsum extrapolation,
    orthogonal_eigh, einsum density -- cold and warm-started."""
    _say(f"case: scf_step fused vs unfused, n={n}")
    from graddft_amd.eigenproblem import orthogonal_eigh, scf_step
    rng = np.random.default_rng(n)
    m = 10
    dev = "cuda"
    X = rng.normal(size=(n, n))
    S = X @ X.T / n + np.eye(n)
    L = np.linalg.cholesky(S)
    Linv = torch.tensor(np.linalg.inv(L), device=dev)
    F = rng.normal(size=(m, 2, n, n))
    focks = torch.tensor(F + F.transpose(0, 1, 3, 2), device=dev)
    x = torch.tensor(rng.normal(size=(2, m + 1)), device=dev)
    occ = torch.zeros((2, n), dtype=torch.float64, device=dev)
    occ[0, :7] = 1.0
    occ[1, :6] = 1.0
    Fr = torch.einsum("si,isjk->sjk", x[:, 1:], focks)
    w0, C0, V0 = orthogonal_eigh(Fr, Linv, None)
    D0 = torch.einsum("sij,sj,skj->sik", C0, occ, C0)
    F1, w1, C1, V1, D1 = scf_step(focks, x, Linv, None, occ)
    tol = 1e-10 * max(1.0, float(w0.abs().max()))
    assert (F1 - Fr).abs().max().item() <= 1e-12 * max(1.0, float(Fr.abs().max()))
    assert (w1 - w0).abs().max().item() <= tol
    assert (D1 - D0).abs().max().item() <= 1e-9
    # eigenvectors agree up to sign; V orthonormal
    sgn = torch.sign((C1 * C0).sum(dim=1, keepdim=True))
    assert (C1 - sgn * C0).abs().max().item() <= 1e-8
    eye = torch.eye(n, dtype=torch.float64, device=dev)
    assert (V1.transpose(1, 2) @ V1 - eye).abs().max().item() <= 1e-12
    # warm start from a perturbed problem's eigenvectors
    x2 = x + 1e-3 * torch.tensor(rng.normal(size=(2, m + 1)), device=dev)
    Fr2 = torch.einsum("si,isjk->sjk", x2[:, 1:], focks)
    w2, C2, _ = orthogonal_eigh(Fr2, Linv, V1)
    _, w3, C3, _, D3 = scf_step(focks, x2, Linv, V1, occ)
    assert (w3 - w2).abs().max().item() <= tol
    D2 = torch.einsum("sij,sj,skj->sik", C2, occ, C2)
    assert (D3 - D2).abs().max().item() <= 1e-9


def _oracle_scf_inputs(mol):
    from oracle.ao import eval_ao
    from oracle.scf import SCFInputs, exchange_potential
    from oracle.xc import to_inputs, xc_value_and_grads
    eri = _dense_eri(mol)
    D0 = mol.cache["rdm1_init"].cpu().numpy()
    K0 = np.stack([exchange_potential(D0[s], eri) for s in range(2)])
    inp = SCFInputs(mol.h1e.cpu().numpy(), mol.s1e.cpu().numpy(), eri, mol.nuclear_repulsion, K0, 0.2)
    xin = to_inputs(*eval_ao(mol.basis, mol.grid.coords.cpu().numpy()), mol.grid.weights.cpu().numpy())

    def xc(rdm1):
        e, g, _ = xc_value_and_grads("b3lyp", xin, rdm1)
        return e, g

    return inp, xc


def test_predict_b3lyp_matches_oracle():
    _say("case: predict B3LYP water/def2-svp level 1")
    from graddft_amd import B3LYP, energy_predictor
    from oracle.scf import predict
    mol = _water()
    E, F = energy_predictor(B3LYP)(None, mol)
    inp, xc = _oracle_scf_inputs(mol)
    E0, F0 = predict(xc, inp, mol.rdm1.cpu().numpy())
    assert abs(float(E) - E0) <= 1e-9, (float(E), E0)
    assert np.abs(F.cpu().numpy() - F0).max() <= 1e-9


def test_diff_scf_loop_matches_oracle():
    _say("case: diff_scf_loop B3LYP water/def2-svp level 1, 12 cycles")
    from graddft_amd import B3LYP, diff_scf_loop
    from oracle.scf import diff_scf_loop as odiff
    mol = _water()
    out = diff_scf_loop(B3LYP, cycles=12)(None, mol)
    inp, xc = _oracle_scf_inputs(mol)
    st = odiff(xc, inp, mol.rdm1.cpu().numpy(), mol.mo_occ.cpu().numpy(), 12)
    assert abs(float(out.energy) - st.energy) <= 1e-8, (float(out.energy), st.energy)
    assert np.abs(out.rdm1.cpu().numpy() - st.rdm1).max() <= 1e-6
    S = mol.s1e.cpu().numpy()
    D = out.rdm1.cpu().numpy()
    for s in range(2):
        assert np.abs(D[s] @ S @ D[s] - D[s]).max() <= 1e-9          # idempotent
        assert abs(np.trace(D[s] @ S) - 5.0) <= 1e-9                 # N_e / 2 per spin


@pytest.mark.parametrize("kind", ["b3lyp", "neural"])
def test_graph_captured_scf_loop_replays_match_eager(kind):
    """diff_scf_loop(graph=True): every replay, with eager device work in between and with
    new inputs copied in, equals the eager loop on the same inputs."""
    _say(f"case: graph-captured diff_scf_loop {kind} water/def2-svp level 1, 6 cycles")
    from graddft_amd import B3LYP, MLPSpec, NeuralFunctional, diff_scf_loop, energy_predictor
    mol = _water()
    if kind == "b3lyp":
        f, thetas = B3LYP, [None, None, None]
    else:
        f = NeuralFunctional(MLPSpec(nc=6, h=32, n_hidden=2, nd=4))
        thetas = [f.init(seed=1), f.init(seed=2), f.init(seed=1)]
    eager = diff_scf_loop(f, cycles=6)
    graphed = diff_scf_loop(f, cycles=6, graph=True)
    graphed(thetas[0], mol)                      # eager run + capture
    for i, th in enumerate(thetas):
        ref = eager(th, mol)
        energy_predictor(f)(th, mol)             # unrelated eager work between replays
        out = graphed(th, mol)
        torch.cuda.synchronize()
        assert abs(out.energy.item() - ref.energy.item()) <= 1e-12, (i, out.energy.item(), ref.energy.item())
        assert torch.allclose(out.rdm1, ref.rdm1, rtol=0, atol=1e-11), i


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_graph_captured_reverse_pass_matches_eager(precision):
    """diff_scf_loop(graph=True) under autograd: the loop and its reverse pass replayed as two
    captured graphs give the eager loop's energy and d loss / d theta for several parameter
    vectors in turn (each replay with new parameters, eager device work in between), and a
    loss whose cotangent is not 1 (the reverse graph's input is the energy's cotangent)."""
    _say(f"case: graph-captured forward + reverse diff_scf_loop, neural, water/def2-svp level 1, 8 cycles, "
         f"{precision}")
    from graddft_amd import MLPSpec, NeuralFunctional, diff_scf_loop, energy_predictor, simple_energy_loss
    mol = _water()
    f = NeuralFunctional(MLPSpec(nc=6, h=32, n_hidden=2, nd=4))
    thetas = [f.init(seed=1), f.init(seed=2), f.init(seed=3), f.init(seed=1)]
    eager = diff_scf_loop(f, cycles=8, precision=precision)
    graphed = diff_scf_loop(f, cycles=8, precision=precision, graph=True)
    simple_energy_loss(thetas[0], graphed, mol, -76.0)       # eager run + both captures
    for i, th in enumerate(thetas):
        (c0, e0), g0 = simple_energy_loss(th, eager, mol, -76.0)
        energy_predictor(f)(th, mol)                           # unrelated eager work between replays
        (c1, e1), g1 = simple_energy_loss(th, graphed, mol, -76.0)
        torch.cuda.synchronize()
        rel = float(torch.linalg.norm(g1 - g0) / torch.linalg.norm(g0))
        _say(f"  theta {i}: E {e1.item():.12f} (eager {e0.item():.12f}), grad rel diff {rel:.2e}")
        assert abs(e1.item() - e0.item()) <= 1e-12, i
        assert rel <= 1e-12, i
    # a cotangent other than the loss's: d (3 E) / d theta, replayed and eager
    th = thetas[1].detach().clone().requires_grad_(True)
    gg, = torch.autograd.grad(3.0 * graphed(th, mol).energy, [th])
    th2 = thetas[1].detach().clone().requires_grad_(True)
    ge, = torch.autograd.grad(3.0 * eager(th2, mol).energy, [th2])
    assert float(torch.linalg.norm(gg - ge) / torch.linalg.norm(ge)) <= 1e-12


def test_batched_loop_with_shared_cache_matches_eager():
    """batched_scf_loop with a molecule listed twice and a replace()d copy (all three share one
    cache, hence one captured graph and its buffers) equals the molecule-by-molecule loop."""
    _say("case: batched_scf_loop, shared cache [mol, mol.replace(rdm1=D2), mol]")
    from graddft_amd import B3LYP, diff_scf_loop
    from graddft_amd.evaluate import batched_scf_loop
    mol = _water()
    eager = diff_scf_loop(B3LYP, cycles=5)
    ref1 = eager(None, mol)
    D2 = (0.9 * mol.rdm1 + 0.1 * ref1.rdm1).contiguous()
    mol2 = mol.replace(rdm1=D2)
    ref2 = eager(None, mol2)
    batched = batched_scf_loop(B3LYP, cycles=5, n_streams=3)
    for _ in range(2):
        outs = batched(None, [mol, mol2, mol])
        torch.cuda.synchronize()
        for o, r in zip(outs, [ref1, ref2, ref1]):
            assert abs(o.energy.item() - r.energy.item()) <= 1e-12, (o.energy.item(), r.energy.item())
            assert torch.allclose(o.rdm1, r.rdm1, rtol=0, atol=1e-11)


def test_scf_status_word_flags_and_loop_raises(monkeypatch):
    """gdx_scf_status sets one bit per failed check (eigensolver at max_sweeps, non-finite energy,
    tr(D S) != occupation count); diff_scf_loop raises when an eigensolve is cut short by a tiny
    sweep limit set from the host (the kernels still finish normally)."""
    _say("case: scf status word (direct calls) and a loop with max_sweeps = 1")
    from graddft_amd import B3LYP, _lib, diff_scf_loop, eigenproblem
    from graddft_amd.eigenproblem import scf_status
    dev = "cuda"
    n = 5
    rng = np.random.default_rng(3)
    C = np.linalg.qr(rng.normal(size=(n, n)))[0]
    occ = np.zeros((2, n))
    occ[:, :2] = 1.0
    D = np.einsum("ij,sj,kj->sik", C, occ, C)
    t = lambda a, dt=torch.float64: torch.as_tensor(a, dtype=dt, device=dev)  # noqa: E731
    S = t(np.eye(n))
    cases = [(D, occ, [1, 2], 0.5, 0),
             (D, occ, [1, eigenproblem.JACOBI_MAX_SWEEPS], 0.5, _lib.GDX_STATUS_EIGH_NOT_CONVERGED),
             (D, occ, [1, 2], float("nan"), _lib.GDX_STATUS_NONFINITE_ENERGY),
             (1.1 * D, occ, [1, 2], 0.5, _lib.GDX_STATUS_CHARGE)]
    for Dm, oc, sw, e, want in cases:
        status = torch.zeros(1, dtype=torch.int32, device=dev)
        scf_status(t(Dm), S, t(oc), t(sw, torch.int32), t(e), status)
        assert int(status.item()) == want, (sw, e, int(status.item()), want)
    mol = _water()
    diff_scf_loop(B3LYP, cycles=3)(None, mol)                      # converged: no flag
    monkeypatch.setattr(eigenproblem, "JACOBI_MAX_SWEEPS", 1)
    with pytest.raises(_lib.GdxError, match="did not converge"):
        diff_scf_loop(B3LYP, cycles=3)(None, mol)


@pytest.mark.parametrize("n", [7, 43, 180])
def test_diis_error_matches_torch(n):
    """gdx_diis_error (F D S - (F D S)^T through the tiled GEMMs) against torch, and the
    DiisErrorFn reverse mode against finite differences (gradcheck, n = 7)."""
    _say(f"case: diis error n={n}")
    from graddft_amd.evaluate import DiisErrorFn
    rng = np.random.default_rng(n)
    t = lambda a: torch.as_tensor(a, device="cuda")  # noqa: E731
    F = rng.normal(size=(2, n, n))
    D = rng.normal(size=(2, n, n))
    X = rng.normal(size=(n, n))
    S = X @ X.T / n + np.eye(n)
    err = DiisErrorFn.apply(t(F), t(D), t(S)).cpu().numpy()
    A = F @ D @ S
    ref = A - A.transpose(0, 2, 1)
    assert np.abs(err - ref).max() <= 1e-12 * max(1.0, np.abs(ref).max())
    if n == 7:
        Ft = t(F).requires_grad_(True)
        Dt = t(D).requires_grad_(True)
        assert torch.autograd.gradcheck(lambda a, b: DiisErrorFn.apply(a, b, t(S)), (Ft, Dt))
