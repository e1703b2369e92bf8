"""Generalised symmetric eigenproblem on the device (``grad_dft/utils/eigenproblem.py``).

``safe_general_eigh(A, B)``: L = chol(B), Linv = L^-1, C = Linv A Linv^T, eigh(C), Linv^T V
(``eigenproblem.py:110-129``); ``safe_fock_solver`` applies it per spin (``:132-149``).
Cholesky, the triangular inverse and the eigensolver are hand-written kernels
(``gdx_cholesky_inv``, ``gdx_eigh_batched``); the two congruence products are plain GEMMs.
The Lorentzian-broadened reverse mode of ``safe_eigh`` (``:54-103``) is :func:`safe_eigh_vjp`,
attached to the device eigensolver by :class:`EighFn` when autograd is recording.
"""
from __future__ import annotations

import ctypes
from typing import Optional, Tuple

import torch

from . import _lib

DEGEN_TOL = 1e-6   # eigenproblem.py:23
BROADENING = 1e-10  # eigenproblem.py:24
JACOBI_TOL = 1e-14
JACOBI_MAX_SWEEPS = 40


def _stream(t):
    return torch.cuda.current_stream(t.device).cuda_stream


def cholesky_inv(S: torch.Tensor) -> Tuple[torch.Tensor, torch.Tensor]:
    S = S.contiguous()
    n = S.shape[-1]
    L = torch.empty_like(S)
    Linv = torch.empty_like(S)
    status = torch.zeros(1, dtype=torch.int32, device=S.device)
    _lib.check(_lib.load().gdx_cholesky_inv(_stream(S), n, S.data_ptr(), L.data_ptr(), Linv.data_ptr(),
                                            status.data_ptr()), "gdx_cholesky_inv")
    # the reference's cholesky returns NaN for a matrix that is not positive definite; the
    # kernel flags it instead (read here, outside graph capture: one sync per molecule)
    if not torch.cuda.is_current_stream_capturing() and int(status.item()) != 0:
        raise _lib.GdxError("gdx_cholesky_inv: the overlap matrix is not positive definite")
    return L, Linv


def eigh(A: torch.Tensor, V0: Optional[torch.Tensor] = None,
         sweeps: Optional[torch.Tensor] = None) -> Tuple[torch.Tensor, torch.Tensor]:
    """Batched eigh of symmetric A (..., n, n): ascending eigenvalues, eigenvectors as columns.
    V0 (same shape, orthonormal columns): starting eigenvector estimates (warm start).
    sweeps (int32 device tensor, one per matrix): Jacobi sweeps used (JACOBI_MAX_SWEEPS = not
    converged)."""
    shape = A.shape
    n = shape[-1]
    Ab = A.reshape(-1, n, n).contiguous()
    nb = Ab.shape[0]
    w = torch.empty((nb, n), dtype=torch.float64, device=A.device)
    V = torch.empty_like(Ab)
    V0b = None if V0 is None else V0.detach().reshape(-1, n, n).contiguous()
    nbytes = ctypes.c_size_t(0)
    lib = _lib.load()
    _lib.check(lib.gdx_eigh_workspace_size(nb, n, ctypes.byref(nbytes)), "gdx_eigh_workspace_size")
    ws = torch.empty(int(nbytes.value), dtype=torch.uint8, device=A.device)
    _lib.check(lib.gdx_eigh_batched(_stream(A), nb, n, Ab.data_ptr(), 0 if V0b is None else V0b.data_ptr(),
                                    w.data_ptr(), V.data_ptr(), ws.data_ptr(), JACOBI_MAX_SWEEPS, JACOBI_TOL,
                                    0 if sweeps is None else sweeps.data_ptr()),
               "gdx_eigh_batched")
    return w.reshape(shape[:-1]), V.reshape(shape)


class EighFn(torch.autograd.Function):
    """Batched eigh on the device with the reverse mode of safe_eigh (safe_eigh_vjp)."""

    @staticmethod
    def forward(ctx, A, V0=None, sweeps=None):
        w, V = eigh(A, V0, sweeps)
        ctx.save_for_backward(w, V)
        return w, V

    @staticmethod
    @torch.autograd.function.once_differentiable
    def backward(ctx, gw, gV):
        w, V = ctx.saved_tensors
        n = w.shape[-1]
        wb, Vb = w.reshape(-1, n), V.reshape(-1, n, n)
        gw = torch.zeros_like(wb) if gw is None else gw.reshape(-1, n)
        gV = torch.zeros_like(Vb) if gV is None else gV.reshape(-1, n, n)
        return safe_eigh_vjp(wb, Vb, gw, gV).reshape(V.shape), None, None


def scf_step(focks: torch.Tensor, x: torch.Tensor, Linv: torch.Tensor, V0: Optional[torch.Tensor],
             occ: torch.Tensor, sweeps: Optional[torch.Tensor] = None):
    """Forward-only diagonalisation step of the SCF loop in one device call (gdx_scf_step):
    F = sum_i x[:, 1+i] focks[i], the eigenproblem of sym(Linv F Linv^T) warm-started from V0,
    C = Linv^T V and the density C diag(occ) C^T.  Returns (F, w, C, V, rdm1)."""
    focks = focks.detach().contiguous()
    m, ns, n = focks.shape[0], focks.shape[1], focks.shape[-1]
    dev = focks.device
    out = lambda *shape: torch.empty(shape, dtype=torch.float64, device=dev)  # noqa: E731
    F, w, V, C, D = out(ns, n, n), out(ns, n), out(ns, n, n), out(ns, n, n), out(ns, n, n)
    V0c = None if V0 is None else V0.detach().reshape(ns, n, n).contiguous()
    nbytes = ctypes.c_size_t(0)
    lib = _lib.load()
    _lib.check(lib.gdx_scf_step_workspace_size(ns, n, ctypes.byref(nbytes)), "gdx_scf_step_workspace_size")
    ws = torch.empty(int(nbytes.value), dtype=torch.uint8, device=dev)
    xc = x.detach().contiguous()
    Lc = Linv.detach().contiguous()
    oc = occ.detach().contiguous()
    _lib.check(lib.gdx_scf_step(_stream(focks), ns, n, m, focks.data_ptr(), xc.data_ptr(), Lc.data_ptr(),
                                0 if V0c is None else V0c.data_ptr(), oc.data_ptr(), F.data_ptr(), w.data_ptr(),
                                V.data_ptr(), C.data_ptr(), D.data_ptr(), ws.data_ptr(), JACOBI_MAX_SWEEPS,
                                JACOBI_TOL, 0 if sweeps is None else sweeps.data_ptr()), "gdx_scf_step")
    return F, w, C, V, D


def orthogonal_eigh(A: torch.Tensor, Linv: torch.Tensor, V0: Optional[torch.Tensor] = None,
                    sweeps: Optional[torch.Tensor] = None):
    """(w, C, V): eigh of Linv A Linv^T = V diag(w) V^T and C = Linv^T V; V0 warm-starts."""
    C = Linv @ A @ Linv.transpose(-1, -2)
    C = 0.5 * (C + C.transpose(-1, -2))
    if torch.is_grad_enabled() and C.requires_grad:
        w, V = EighFn.apply(C, V0, sweeps)
    else:
        w, V = eigh(C, V0, sweeps)
    return w, Linv.transpose(-1, -2) @ V, V


def safe_general_eigh(A: torch.Tensor, B: torch.Tensor, Linv: torch.Tensor = None):
    if Linv is None:
        _, Linv = cholesky_inv(B)
    w, C, _ = orthogonal_eigh(A, Linv)
    return w, C


def safe_fock_solver(fock: torch.Tensor, overlap: torch.Tensor, Linv: torch.Tensor = None):
    """(mo_energy (2, n), mo_coeff (2, n, n)) of both spin Fock matrices."""
    return safe_general_eigh(fock, overlap, Linv)


def safe_eigh_vjp(evals: torch.Tensor, evecs: torch.Tensor, g_evals: torch.Tensor, g_evecs: torch.Tensor):
    """Reverse mode of safe_eigh (eigenproblem.py:54-103): Lorentzian-broadened gaps for
    near-degenerate pairs, symmetrised result.  Batched over leading dimensions: evals (..., n),
    evecs (..., n, n) with eigenvectors as columns."""
    diff = evals[..., None, :] - evals[..., :, None]
    degen = diff.abs() < DEGEN_TOL
    regular = torch.nan_to_num(1.0 / diff)
    broadened = diff / (diff * diff + BROADENING)
    n = evals.shape[-1]
    off = ~torch.eye(n, dtype=torch.bool, device=evals.device)
    F = torch.where(off, 0.5 * torch.where(degen, broadened, regular), torch.zeros_like(diff))
    Vt = evecs.transpose(-1, -2)
    # inv(V^T) of the reference is V: the eigenvectors are orthonormal to rounding
    grad = evecs @ (0.5 * torch.diag_embed(g_evals) + F * (Vt @ g_evecs)) @ Vt
    return grad + grad.transpose(-1, -2)


def householder_q(W: torch.Tensor) -> torch.Tensor:
    """Q of W = QR for a batch of square matrices, by Householder reflections with LAPACK's
    conventions (dgeqrf / dlarfg / dorgqr), which the reference's ``jnp.linalg.qr`` follows:
    for column j, beta = -sign(alpha) * ||x|| (sign(0) = +1) with alpha the leading entry of
    the sub-column x, no reflection when the part of x below alpha is zero (in particular for
    the last column), Q = H_1 H_2 ... H_n.  The column signs of Q fix which orbitals the
    orbital optimizer starts from, so they follow LAPACK rather than whatever the device QR
    library picks.  Plain differentiable tensor ops (autograd gives dQ/dW)."""
    b, m, n = W.shape
    if m != n:
        raise ValueError("householder_q expects square matrices")
    A = W
    vs, taus = [], []
    for j in range(n):
        x = A[:, j:, j]
        alpha = x[:, 0]
        if j == n - 1:
            vs.append(torch.ones_like(x))
            taus.append(torch.zeros_like(alpha))
            break
        tail = x[:, 1:]
        xnorm2 = (tail * tail).sum(dim=1)
        nz = xnorm2 > 0
        safe2 = torch.where(nz, xnorm2, torch.ones_like(xnorm2))
        sgn = torch.where(alpha >= 0, torch.ones_like(alpha), -torch.ones_like(alpha))
        beta = -sgn * torch.sqrt(alpha * alpha + safe2)
        tau = torch.where(nz, (beta - alpha) / beta, torch.zeros_like(alpha))
        scale = torch.where(nz, 1.0 / (alpha - beta), torch.zeros_like(alpha))
        v = torch.cat([torch.ones_like(alpha)[:, None], tail * scale[:, None]], dim=1)
        sub = A[:, j:, j:]
        w = tau[:, None] * torch.einsum("bi,bik->bk", v, sub)
        sub = sub - v[:, :, None] * w[:, None, :]
        A = torch.cat([A[:, :j, :], torch.cat([A[:, j:, :j], sub], dim=2)], dim=1)
        vs.append(v)
        taus.append(tau)
    Q = torch.eye(m, dtype=W.dtype, device=W.device).expand(b, m, m)
    for j in reversed(range(n)):
        v, tau = vs[j], taus[j]
        sub = Q[:, j:, :]
        w = tau[:, None] * torch.einsum("bi,bik->bk", v, sub)
        Q = torch.cat([Q[:, :j, :], sub - v[:, :, None] * w[:, None, :]], dim=1)
    return Q


STATUS_FLAGS = {_lib.GDX_STATUS_EIGH_NOT_CONVERGED: f"the Jacobi eigensolver did not converge in {JACOBI_MAX_SWEEPS} sweeps",
                _lib.GDX_STATUS_NONFINITE_ENERGY: "the energy is not finite",
                _lib.GDX_STATUS_CHARGE: "tr(D S) differs from the occupation count (orbitals not S-orthonormal)"}


def scf_status(rdm1: torch.Tensor, S: torch.Tensor, occ: torch.Tensor, sweeps: Optional[torch.Tensor],
               energy: Optional[torch.Tensor], status: torch.Tensor, tol: float = 1e-6) -> None:
    """OR this cycle's checks into the device status word (gdx_scf_status); no host sync."""
    ns, n = int(rdm1.shape[0]), int(rdm1.shape[-1])
    D = rdm1.detach().contiguous()
    oc = occ.detach().contiguous()
    e = None if energy is None else energy.detach().reshape(()).contiguous()
    _lib.check(_lib.load().gdx_scf_status(_stream(D), ns, n, D.data_ptr(), S.contiguous().data_ptr(), oc.data_ptr(),
                                          0 if sweeps is None else sweeps.data_ptr(), JACOBI_MAX_SWEEPS,
                                          0 if e is None else e.data_ptr(), float(tol), status.data_ptr()),
               "gdx_scf_status")


def raise_on_status(status: torch.Tensor, what: str) -> None:
    """Read the status word (one host sync) and raise GdxError naming every flag set."""
    bits = int(status.item())
    if bits:
        msgs = [m for b, m in STATUS_FLAGS.items() if bits & b]
        raise _lib.GdxError(f"{what}: " + "; ".join(msgs))
