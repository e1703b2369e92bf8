"""Oracle AO evaluation: phi, grad phi and the diagonal second derivatives on a grid.

Restates what the reference receives from PySCF ``numint.eval_ao(mol, coords, deriv=1)``
and ``ao_grads(mol, coords, order=2)`` (``grad_dft/interface/pyscf.py:824-827, 677-705``):
``ao (N, n)``, ``grad_ao (N, n, 3)`` and ``grad_n_ao[2] (N, n, 3)`` holding
d2/dx2, d2/dy2, d2/dz2 only.

Pinned by: unit norm of every AO by quadrature, analytic overlap S vs grid sum
(``tests/test_oracle_inputs.py``), and central finite differences of phi and grad phi.
"""
from __future__ import annotations

import numpy as np


def _mono_and_derivs(dx, dy, dz, a, b, c):
    """x^a y^b z^c and its first and pure second derivatives (vectorised)."""
    def p(v, k):
        return v ** k if k > 0 else np.ones_like(v)

    def d1(v, k):
        return k * v ** (k - 1) if k > 0 else np.zeros_like(v)

    def d2(v, k):
        return k * (k - 1) * v ** (k - 2) if k > 1 else np.zeros_like(v)

    X, Y, Z = p(dx, a), p(dy, b), p(dz, c)
    X1, Y1, Z1 = d1(dx, a), d1(dy, b), d1(dz, c)
    X2, Y2, Z2 = d2(dx, a), d2(dy, b), d2(dz, c)
    P = X * Y * Z
    return P, (X1 * Y * Z, X * Y1 * Z, X * Y * Z1), (X2 * Y * Z, X * Y2 * Z, X * Y * Z2)


def eval_ao(basis, coords: np.ndarray):
    """Return ``ao (N,n)``, ``grad_ao (N,n,3)``, ``d2_ao (N,n,3)`` in float64."""
    from graddft_amd.interface.basis import cart_exponents, component_matrix

    coords = np.asarray(coords, dtype=np.float64)
    N = len(coords)
    n = basis.nao
    ao = np.zeros((N, n))
    g = np.zeros((N, n, 3))
    h = np.zeros((N, n, 3))
    for sh in basis.shells:
        d = coords - sh.center[None, :]
        r2 = np.einsum("ij,ij->i", d, d)
        e = np.exp(-np.outer(r2, sh.exps))          # (N, nprim)
        R0 = e @ sh.coefs                            # sum c exp
        R1 = e @ (sh.coefs * sh.exps)                # sum c a exp
        R2 = e @ (sh.coefs * sh.exps ** 2)           # sum c a^2 exp
        T = component_matrix(sh.l, sh.cart)
        cart = cart_exponents(sh.l)
        vals = np.zeros((N, len(cart)))
        grads = np.zeros((N, len(cart), 3))
        secs = np.zeros((N, len(cart), 3))
        for k, (a, b, c) in enumerate(cart):
            P, dP, d2P = _mono_and_derivs(d[:, 0], d[:, 1], d[:, 2], a, b, c)
            vals[:, k] = P * R0
            for j in range(3):
                xj = d[:, j]
                # d/dx [P R] with dR0/dx = -2 x R1, dR1/dx = -2 x R2
                grads[:, k, j] = dP[j] * R0 - 2.0 * xj * P * R1
                secs[:, k, j] = d2P[j] * R0 - 2.0 * (P + 2.0 * xj * dP[j]) * R1 \
                    + 4.0 * xj * xj * P * R2
        sl = slice(sh.ao_offset, sh.ao_offset + sh.nfunc)
        ao[:, sl] = vals @ T.T
        g[:, sl, :] = np.einsum("nkj,mk->nmj", grads, T)
        h[:, sl, :] = np.einsum("nkj,mk->nmj", secs, T)
    return ao, g, h
