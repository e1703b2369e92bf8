"""Pins of the oracle's safe_eigh reverse mode (``grad_dft/utils/eigenproblem.py:54-103``),
CPU, in the way of the reference's own tests (``tests/unit/test_eigenproblem.py:84-122``):

* non-degenerate random symmetric matrices, n = 2..9, 8 seeds: the reverse-mode Jacobian of
  (eigenvalues, eigenvectors) equals that of the plain eigensolver applied to the symmetrised
  input -- what ``jnp.linalg.eigh`` differentiates (``symmetrize_input=True``) -- to 1e-10;
* a matrix with a duplicated eigenvalue: the Jacobian is NaN-free;
* the Lorentzian branch in closed form: for a 2x2 matrix with gap delta, the eigenvector
  cotangent's gradient is the exact one scaled by 1 (delta >= 1e-6) or by
  delta^2 / (delta^2 + 1e-10) (delta < 1e-6);
* a function of the occupied subspace only (tr(M P), P the projector on a degenerate pair and
  the orbitals below it) gets its exact gradient (central differences) through the broadened
  branch, because the antisymmetric gap matrix cancels inside the subspace.

The seeds are numpy's (the reference's JAX PRNG streams cannot be reproduced here).
"""
import numpy as np
import pytest
import torch

from oracle.scf_grad import SafeEigh

DT = torch.float64
SEEDS = [1984, 1993, 1945, 2001, 10, 29, 101, 1992]


def rand_sym(n, seed):
    a = np.random.default_rng(seed).normal(size=(n, n))
    return 0.5 * (a + a.T)


def degenerate_sym(n, seed):
    rng = np.random.default_rng(seed)
    w, V = np.linalg.eigh(rand_sym(n, seed))
    i = int(rng.integers(0, n))
    w[i] = w[i - 1]
    return V @ np.diag(w) @ V.T


def jac(fn, A):
    return torch.autograd.functional.jacobian(fn, torch.tensor(A, dtype=DT))


def safe(A):
    return SafeEigh.apply(A)


def plain_symmetrised(A):
    return torch.linalg.eigh(0.5 * (A + A.T))


@pytest.mark.parametrize("n", list(range(2, 10)))
def test_non_degenerate_jacobians_match_plain_eigh(n):
    for seed in SEEDS:
        A = rand_sym(n, seed)
        js, jp = jac(safe, A), jac(plain_symmetrised, A)
        assert torch.allclose(js[0], jp[0], rtol=0, atol=1e-10), (n, seed)
        assert torch.allclose(js[1], jp[1], rtol=0, atol=1e-10), (n, seed)


@pytest.mark.parametrize("n", list(range(2, 10)))
def test_degenerate_jacobians_are_nan_free(n):
    for seed in SEEDS:
        js = jac(safe, degenerate_sym(n, seed))
        assert not torch.isnan(js[0]).any() and not torch.isnan(js[1]).any(), (n, seed)


@pytest.mark.parametrize("delta", [1e-3, 2e-6, 9e-7, 3e-7, 1e-8, 0.0])
def test_lorentzian_branch_closed_form(delta):
    """2x2, eigenvalues 1 and 1 + delta, eigenvectors rotated by 0.3 rad; cotangent on one
    eigenvector component.  Exact gradient: V (F o (V^T gV)) V^T symmetrised, F = 1/gap;
    the reverse mode of safe_eigh scales it by delta^2 / (delta^2 + 1e-10) below 1e-6."""
    c, s = np.cos(0.3), np.sin(0.3)
    V = np.array([[c, -s], [s, c]])
    A = V @ np.diag([1.0, 1.0 + delta]) @ V.T
    At = torch.tensor(A, dtype=DT, requires_grad=True)
    w, Vt = safe(At)
    gV = torch.zeros(2, 2, dtype=DT)
    gV[0, 1] = 1.0
    g, = torch.autograd.grad((Vt * gV).sum(), [At])
    # closed form with the eigenvectors the solver returned
    Vn = Vt.detach().numpy()
    X = Vn.T @ gV.numpy()
    gap = w[1].item() - w[0].item()
    if abs(gap) >= 1e-6:
        f = 1.0 / gap
    else:
        f = gap / (gap * gap + 1e-10)
    F = np.array([[0.0, 0.5 * f], [-0.5 * f, 0.0]])
    inner = Vn @ (F * X) @ Vn.T
    expect = inner + inner.T
    assert np.abs(g.numpy() - expect).max() <= 1e-12 * max(1.0, np.abs(expect).max())
    if delta >= 1e-6:
        # the same as the exact derivative of the plain eigensolver
        Ap = torch.tensor(A, dtype=DT, requires_grad=True)
        _, Vp = plain_symmetrised(Ap)
        sign = float(np.sign((Vp.detach().numpy() * Vn).sum(0)[1]))
        gp, = torch.autograd.grad((Vp * gV).sum() * sign, [Ap])
        assert np.abs(g.numpy() - gp.numpy()).max() <= 1e-9 * np.abs(gp.numpy()).max()


def test_subspace_function_exact_through_degenerate_pair():
    """tr(M P) with P the projector on the 3 lowest eigenvectors, where eigenvalues 2 and 3
    coincide (index 1, 2): its gradient through safe_eigh equals central differences."""
    n = 6
    rng = np.random.default_rng(5)
    Q, _ = np.linalg.qr(rng.normal(size=(n, n)))
    w = np.array([-2.0, 0.5, 0.5, 1.7, 2.4, 3.0])
    A = Q @ np.diag(w) @ Q.T
    M = rng.normal(size=(n, n))
    M = M + M.T
    Mt = torch.tensor(M, dtype=DT)

    def f(At):
        _, V = safe(At)
        P = V[:, :3] @ V[:, :3].T
        return (Mt * P).sum()

    At = torch.tensor(A, dtype=DT, requires_grad=True)
    g, = torch.autograd.grad(f(At), [At])
    E = rng.normal(size=(n, n))
    E = E + E.T
    h = 1e-5
    fd = (f(torch.tensor(A + h * E, dtype=DT)) - f(torch.tensor(A - h * E, dtype=DT))).item() / (2 * h)
    assert (g.numpy() * E).sum() == pytest.approx(fd, rel=1e-7, abs=1e-9)
