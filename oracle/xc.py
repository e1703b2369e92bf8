"""Oracle XC path in torch-CPU float64 (autograd stands in for ``jax.grad``).

Restates, expression by expression, the reference:

* grid densities -- ``grad_dft/molecule.py:388-409`` (density), ``:414-440``
  (grad_density), ``:445-474`` (lapl_density), ``:479-502`` (kinetic_density),
  ``:507-541`` (HF_energy_density), ``:687-689`` (abs_clip);
* energy densities -- ``grad_dft/popular_functionals.py:29-50`` (lsda_x_e), ``:52-103``
  (b88_x_e), ``:105-139`` (pw92_c_e), ``:141-195`` (vwn_c_e), ``:197-269`` (lyp_c_e),
  ``:306-338`` (b3lyp_exhf_densities, b3lyp_combine), ``:340-347`` (b3lyp_coefficients);
  ``grad_dft/functional.py:950-979`` / ``:982-1045`` (spin-polarisation corrections);
* integration -- ``grad_dft/functional.py:160-185`` (compute_densities clip),
  ``:219-253`` (xc_energy), ``:316-342`` (_integrate);
* the neural head -- ``grad_dft/functional.py:407-419`` (dense -> sigmoid(x/2) * 2) with
  the log(|x| + 1e-4) input squash of the reference's neural examples
  (``examples/intermediate_notebooks/neural_metaGGA_01.ipynb``, nn_coefficients).

* custom coefficient functions -- ``grad_dft/functional.py:160-253`` with a user
  ``coefficients`` callable over the squashed inputs (custom_energy_from_rdm1).

Pinning: brute-force loops on tiny inputs and the identities
``sum_r w rho_s = tr(D_s S)``, ``lapl = 4 tau + 2 sum D phi lapl(phi)``;
closed forms for LDA exchange and the zeta = 0 / +-1 limits of the polarisation
corrections; central finite differences for every derivative; the custom path against
the neural, B3LYP and PW92 paths it reduces to (``tests/test_oracle_xc.py``).  flax defaults beyond ELU/Glorot are not used.
"""
from __future__ import annotations

import math
from dataclasses import dataclass
from typing import List, Optional, Sequence, Tuple

import numpy as np
import torch

CLIP = 1e-30
DT = torch.float64


def abs_clip(x, thr=CLIP):
    """molecule.py:687-689."""
    return torch.where(torch.abs(x) > thr, x, torch.zeros_like(x))


# ---------------------------------------------------------------- grid densities ----

def density(rdm1, ao):
    """molecule.py:409  einsum('...ab,ra,rb->r...')  -> (N, 2)."""
    return torch.einsum("sab,ra,rb->rs", rdm1, ao, ao)


def grad_density(rdm1, ao, grad_ao):
    """molecule.py:440  2 einsum('...ab,ra,rbj->r...j')  -> (N, 2, 3)."""
    return 2.0 * torch.einsum("sab,ra,rbj->rsj", rdm1, ao, grad_ao)


def lapl_density(rdm1, ao, grad_ao, grad_2_ao):
    """molecule.py:472-474  -> (N, 2)."""
    return 2.0 * torch.einsum("sab,raj,rbj->rs", rdm1, grad_ao, grad_ao) + \
        2.0 * torch.einsum("sab,ra,rbi->rs", rdm1, ao, grad_2_ao)


def kinetic_density(rdm1, grad_ao):
    """molecule.py:502  0.5 einsum('...ab,raj,rbj->r...')  -> (N, 2)."""
    return 0.5 * torch.einsum("sab,raj,rbj->rs", rdm1, grad_ao, grad_ao)


def hf_energy_density(rdm1, ao, chi):
    """molecule.py:537-541  -> (omega, spin, N); chi (N, omega, spin, n)."""
    return -torch.einsum("rwsc,sac,ra->wsr", chi, rdm1, ao) / 2.0


# ------------------------------------------------------------ energy densities ----

def _fzeta_x(z):
    return ((1 - z) ** (4.0 / 3.0) + (1 + z) ** (4.0 / 3.0) - 2) / (2 * (2 ** (1.0 / 3.0) - 1))


def exchange_polarization_correction(e_PF, rho):
    """functional.py:950-979; e_PF (N,2), rho (N,2) -> (N,)."""
    zeta = (rho[:, 0] - rho[:, 1]) / rho.sum(dim=1)
    return e_PF[:, 0] + (e_PF[:, 1] - e_PF[:, 0]) * _fzeta_x(zeta)


# PW92 fits G(rs) = -2A (1 + alpha1 rs) ln(1 + 1 / (2A (beta1 rs^1/2 + beta2 rs + beta3 rs^3/2
# + beta4 rs^2))) as (A, alpha1, beta1, beta2, beta3, beta4): the spin stiffness
# (functional.py:1019-1024, used as alpha_c = -G), the paramagnetic and ferromagnetic gas
# (popular_functionals.py:120-125).  Module-level so that tests can plant a mistake in one.
PW92_ALPHA = (0.016887, 0.11125, 10.357, 3.6231, 0.88026, 0.49671)
PW92_PARA = (0.031091, 0.21370, 7.5957, 3.5876, 1.6382, 0.49294)
PW92_FERRO = (0.015545, 0.20548, 14.1189, 6.1977, 3.3662, 0.62517)

# f''(0) of the correlation zeta function: (4/3)(1/3)*2 / (2(2^(1/3)-1))
FZETA_PP0 = 4.0 / (9.0 * (2 ** (1.0 / 3.0) - 1))


def correlation_polarization_correction(e_tilde_PF, rho, clip_cte=CLIP):
    """functional.py:982-1045."""
    rs_sum = rho.sum(dim=1)
    log_rho = torch.log2(torch.clamp(rs_sum, min=clip_cte))
    log_rs = math.log2((3 / (4 * math.pi)) ** (1 / 3)) - log_rho / 3.0
    zeta = torch.where(rs_sum > clip_cte, (rho[:, 0] - rho[:, 1]) / rs_sum,
                       torch.zeros_like(rs_sum))

    def fzeta(z):
        zm = torch.exp2(4 * torch.log2(1 - z) / 3)
        zp = torch.exp2(4 * torch.log2(1 + z) / 3)
        return (zm + zp - 2) / (2 * (2 ** (1 / 3) - 1))

    A_, alpha1, beta1, beta2, beta3, beta4 = PW92_ALPHA
    ars = torch.exp2(math.log2(alpha1) + log_rs)
    brs_1_2 = torch.exp2(math.log2(beta1) + log_rs / 2)
    brs = torch.exp2(math.log2(beta2) + log_rs)
    brs_3_2 = torch.exp2(math.log2(beta3) + 3 * log_rs / 2)
    brs2 = torch.exp2(math.log2(beta4) + 2 * log_rs)
    alphac = 2 * A_ * (1 + ars) * torch.log(1 + (1 / (2 * A_)) / (brs_1_2 + brs + brs_3_2 + brs2))
    fz = fzeta(zeta)
    z4 = zeta ** 4
    return (e_tilde_PF[:, 0] + alphac * (fz / FZETA_PP0) * (1 - z4)
            + (e_tilde_PF[:, 1] - e_tilde_PF[:, 0]) * fz * z4)


def lsda_x_e(rho, clip_cte=CLIP):
    """popular_functionals.py:29-50."""
    rho = torch.clamp(rho, min=clip_cte)
    c = torch.tensor([[3.0, 6.0]], dtype=rho.dtype)
    lda_es = -3.0 / 4.0 * (c / math.pi) ** (1 / 3) * rho.sum(dim=1, keepdim=True) ** (4 / 3)
    return exchange_polarization_correction(lda_es, rho)


def b88_x_e(rho, grad_rho, clip_cte=CLIP):
    """popular_functionals.py:52-103 (gradient correction only)."""
    beta = 0.0042
    rho = torch.clamp(rho, min=clip_cte)
    log_rho = torch.log2(torch.clamp(rho, min=clip_cte))
    grad_rho_norm_sq = torch.sum(grad_rho ** 2, dim=-1)
    log_grad_rho_norm = torch.log2(torch.clamp(grad_rho_norm_sq, min=clip_cte)) / 2
    log_x_sigma = log_grad_rho_norm - 4 / 3.0 * log_rho
    x_sigma = torch.exp2(log_x_sigma)
    return -(beta * torch.exp2(4 * log_rho / 3 + 2 * log_x_sigma
                               - torch.log2(1 + 6 * beta * x_sigma * torch.asinh(x_sigma)))).sum(dim=1)


def pw92_c_e(rho, clip_cte=CLIP):
    """popular_functionals.py:105-139."""
    t = lambda *v: torch.tensor([list(v)], dtype=rho.dtype)  # noqa: E731
    A_, alpha1, beta1, beta2, beta3, beta4 = (t(p, f) for p, f in zip(PW92_PARA, PW92_FERRO))
    log_rho = torch.log2(torch.clamp(rho.sum(dim=1, keepdim=True), min=clip_cte))
    log_rs = math.log2((3 / (4 * math.pi)) ** (1 / 3)) - log_rho / 3.0
    brs_1_2 = torch.exp2(log_rs / 2 + torch.log2(beta1))
    ars = torch.exp2(log_rs + torch.log2(alpha1))
    brs = torch.exp2(log_rs + torch.log2(beta2))
    brs_3_2 = torch.exp2(3 * log_rs / 2 + torch.log2(beta3))
    brs2 = torch.exp2(2 * log_rs + torch.log2(beta4))
    e_PF = -2 * A_ * (1 + ars) * torch.log(1 + (1 / (2 * A_)) / (brs_1_2 + brs + brs_3_2 + brs2))
    return correlation_polarization_correction(e_PF, rho, clip_cte) * rho.sum(dim=1)


# VWN5 (paramagnetic, ferromagnetic), popular_functionals.py:158-161; A in the reference's
# convention (the formula uses A / 2).
VWN_A = (0.0621814, 0.0621814 / 2)
VWN_B = (3.72744, 7.06042)
VWN_C = (12.9352, 18.0578)
VWN_X0 = (-0.10498, -0.325)


def vwn_c_e(rho, clip_cte=CLIP):
    """popular_functionals.py:141-195 (VWN5 parameters)."""
    t = lambda v: torch.tensor([list(v)], dtype=rho.dtype)  # noqa: E731
    A, b, c, x0 = t(VWN_A), t(VWN_B), t(VWN_C), t(VWN_X0)
    rho = torch.where(rho > clip_cte, rho, torch.zeros_like(rho))
    log_rho = torch.log2(torch.clamp(rho.sum(dim=1, keepdim=True), min=clip_cte))
    log_rs = math.log2((3 / (4 * math.pi)) ** (1 / 3)) - log_rho / 3.0
    log_x = log_rs / 2
    x = torch.exp2(log_x)
    X = torch.exp2(2 * log_x) + torch.exp2(log_x + torch.log2(b)) + c
    X0 = x0 ** 2 + b * x0 + c
    Q = torch.sqrt(4 * c - b ** 2)
    e_PF = A / 2 * (2 * torch.log(x) - torch.log(X) + 2 * b / Q * torch.atan(Q / (2 * x + b))
                    - b * x0 / X0 * (torch.log((x - x0) ** 2 / X)
                                     + 2 * (2 * x0 + b) / Q * torch.atan(Q / (2 * x + b))))
    return correlation_polarization_correction(e_PF, rho, clip_cte) * rho.sum(dim=1)


def lyp_c_e(rho, grad_rho, grad2rho, clip_cte=CLIP):
    """popular_functionals.py:197-269 (original LYP form)."""
    a, b, c, d = 0.04918, 0.132, 0.2533, 0.349
    CF = (3 / 10) * (3 * math.pi ** 2) ** (2 / 3)
    rho = torch.clamp(rho, min=clip_cte)
    gsq = torch.sum(grad_rho ** 2, dim=-1)
    t = (torch.where(rho > clip_cte, gsq / rho, torch.zeros_like(rho)) - grad2rho) / 8.0
    rs = rho.sum(dim=1)
    frac = torch.where(rs > clip_cte, (rho ** 2).sum(dim=1) / rs ** 2, torch.ones_like(rs))
    gamma = 2 * (1 - frac)
    rhos_ts = rs * t.sum(dim=1)
    rho_t = (rho * t).sum(dim=1)
    rho_grad2rho = (rho * grad2rho).sum(dim=1)
    rhom1_3 = rs ** (-1 / 3)
    rho8_3 = (rho ** (8 / 3)).sum(dim=1)
    rhom5_3 = rs ** (-5 / 3)
    exp_factor = torch.where(rs > 0, torch.exp(-c * rhom1_3), torch.zeros_like(rs))
    parenthesis = 2 ** (2 / 3) * CF * rho8_3 - rhos_ts + rho_t / 9 + rho_grad2rho / 18
    braket_m_rho = torch.where(rs > clip_cte, 2 * b * rhom5_3 * parenthesis * exp_factor,
                               torch.zeros_like(rs))
    return -a * torch.where(rs > clip_cte, gamma / (1 + d * rhom1_3) * (rs + braket_m_rho),
                            torch.zeros_like(rs))


def b3lyp_exhf_densities(rho, grad_rho, lapl, clip_cte=CLIP):
    """popular_functionals.py:306-326 -> (N, 4): lda, b88, vwn, lyp."""
    return torch.stack((lsda_x_e(rho, clip_cte), b88_x_e(rho, grad_rho, clip_cte),
                        vwn_c_e(rho, clip_cte), lyp_c_e(rho, grad_rho, lapl, clip_cte)), dim=1)


B3LYP_COEFFS = (1 - 0.2, 0.72, 1 - 0.81, 0.81, 0.2)  # popular_functionals.py:340-347


# ------------------------------------------------------------------ neural head ----

SQUASH_OFFSET = 1e-4


@dataclass
class MLPSpec:
    """Coefficient MLP: nc -> h (ELU) -> ... -> h (ELU) -> nd, head 2*sigmoid(z/2)."""
    nc: int = 6
    h: int = 256
    n_hidden: int = 3
    nd: int = 4

    def shapes(self) -> List[Tuple[int, int]]:
        dims = [self.nc] + [self.h] * self.n_hidden + [self.nd]
        return [(dims[i], dims[i + 1]) for i in range(len(dims) - 1)]

    def n_params(self) -> int:
        return sum(i * o + o for i, o in self.shapes())


def init_mlp_params(spec: MLPSpec, seed: int) -> np.ndarray:
    """Glorot-normal kernels, zero biases, flat [W0, b0, W1, b1, ...], W row-major (in,out)."""
    rng = np.random.default_rng(seed)
    parts = []
    for i, o in spec.shapes():
        parts.append(rng.normal(0.0, math.sqrt(2.0 / (i + o)), size=(i, o)).ravel())
        parts.append(np.zeros(o))
    return np.concatenate(parts)


def unflatten(spec: MLPSpec, flat):
    out = []
    off = 0
    for i, o in spec.shapes():
        W = flat[off:off + i * o].reshape(i, o)
        off += i * o
        b = flat[off:off + o]
        off += o
        out.append((W, b))
    return out


def mlp_coefficients(spec: MLPSpec, flat_params, x, head: str = "sigmoid"):
    """log squash -> dense/ELU stack -> head: 'sigmoid' is 2 sigmoid(z/2) (functional.py:407-419,
    scale factor 2); 'b3lyp_correction' is the linear head c = c_B3LYP + z (B3LYP plus a neural
    correction)."""
    y = torch.log(torch.abs(x) + SQUASH_OFFSET)
    layers = unflatten(spec, flat_params)
    for W, b in layers[:-1]:
        y = torch.nn.functional.elu(y @ W + b)
    W, b = layers[-1]
    z = y @ W + b
    if head == "b3lyp_correction":
        return torch.tensor([B3LYP_COEFFS[:4]], dtype=DT) + z
    return 2.0 * torch.sigmoid(z / 2.0)


def neural_features(rho, grad_rho, tau):
    """Coefficient inputs: rho_up, rho_dn, |grad rho_up|^2, |grad rho_dn|^2, tau_up, tau_dn."""
    g = torch.sum(grad_rho ** 2, dim=-1)
    return torch.cat([rho, g, tau], dim=1)


# ------------------------------------------------------------------- integration ----

def xc_energy(weights, coefficients, densities, clip_cte=CLIP):
    """functional.py:219-253 + _integrate :316-342."""
    e = torch.einsum("rf,rf->r", coefficients.expand_as(densities), densities)
    e = abs_clip(e, clip_cte)
    return torch.einsum("r,r->", abs_clip(weights, clip_cte), abs_clip(e, clip_cte))


@dataclass
class XCInputs:
    ao: torch.Tensor        # (N, n)
    grad_ao: torch.Tensor   # (N, n, 3)
    d2_ao: torch.Tensor     # (N, n, 3)
    weights: torch.Tensor   # (N,)
    chi: Optional[torch.Tensor] = None  # (N, 1, 2, n) for exact exchange


LINEAR_KINDS = {"lsda": ("lda",), "b88": ("lda", "b88"), "vwn": ("vwn",), "pw92": ("pw92",), "lyp": ("lyp",)}


def xc_energy_from_rdm1(kind: str, inp: XCInputs, rdm1, params=None,
                        spec: Optional[MLPSpec] = None):
    """E_xc(D) for ``kind`` in {'b3lyp', 'neural_b3lyp', 'b3lyp_neural_corr'} following compute_densities ->
    compute_coefficient_inputs -> xc_energy (functional.py:255-288 without nonXC)."""
    rho = density(rdm1, inp.ao)
    grho = grad_density(rdm1, inp.ao, inp.grad_ao)
    lap = lapl_density(rdm1, inp.ao, inp.grad_ao, inp.d2_ao)
    dens = b3lyp_exhf_densities(rho, grho, lap)
    if kind == "b3lyp":
        if inp.chi is not None:
            ehf = hf_energy_density(rdm1, inp.ao, inp.chi).detach()  # stop_gradient
            dens = torch.cat([dens, ehf.sum(dim=(0, 1))[:, None]], dim=1)
            coefs = torch.tensor([B3LYP_COEFFS], dtype=DT)
        else:
            coefs = torch.tensor([B3LYP_COEFFS[:4]], dtype=DT)
        dens = abs_clip(dens)
        return xc_energy(inp.weights, coefs, dens)
    if kind in LINEAR_KINDS:
        # LSDA / B88 / VWN / PW92 / LYP (popular_functionals.py:375-393): a unit coefficient
        # broadcast over the functional's energy densities (lsda_density ... lyp_density, :271-304)
        parts = {"lda": lambda: lsda_x_e(rho), "b88": lambda: b88_x_e(rho, grho), "vwn": lambda: vwn_c_e(rho),
                 "pw92": lambda: pw92_c_e(rho), "lyp": lambda: lyp_c_e(rho, grho, lap)}
        dens = abs_clip(torch.stack([parts[k]() for k in LINEAR_KINDS[kind]], dim=1))
        return xc_energy(inp.weights, torch.ones((1, dens.shape[1]), dtype=DT), dens)
    if kind in ("neural_b3lyp", "b3lyp_neural_corr"):
        tau = kinetic_density(rdm1, inp.grad_ao)
        x = neural_features(rho, grho, tau)
        dens = abs_clip(dens)
        coefs = mlp_coefficients(spec, params, x,
                                 head="b3lyp_correction" if kind == "b3lyp_neural_corr" else "sigmoid")
        return xc_energy(inp.weights, coefs, dens)
    raise ValueError(kind)


def xc_value_and_grads(kind: str, inp: XCInputs, rdm1: np.ndarray, params=None,
                       spec: Optional[MLPSpec] = None):
    """(E_xc, dE/dD (2,n,n), dE/dtheta) -- the value_and_grad of train.py:86-121,
    extended with the parameter gradient used for training."""
    D = torch.tensor(rdm1, dtype=DT, requires_grad=True)
    th = None
    if params is not None:
        th = torch.tensor(params, dtype=DT, requires_grad=True)
    E = xc_energy_from_rdm1(kind, inp, D, th, spec)
    wrt = [D] + ([th] if th is not None else [])
    grads = torch.autograd.grad(E, wrt)
    gD = grads[0].numpy()
    gth = grads[1].numpy() if th is not None else None
    return float(E.detach()), gD, gth


def custom_energy_from_rdm1(inp: XCInputs, rdm1, coefficients, params=None, correlation: str = "vwn"):
    """E_xc(D) of Functional(coefficients=user function) (functional.py:160-253): energy
    densities (lda, b88, correlation, lyp) clipped, coefficient inputs x = log(|f| + 1e-4) of
    the six features, c = coefficients(params, x) (N, 4), then xc_energy."""
    rho = density(rdm1, inp.ao)
    grho = grad_density(rdm1, inp.ao, inp.grad_ao)
    lap = lapl_density(rdm1, inp.ao, inp.grad_ao, inp.d2_ao)
    tau = kinetic_density(rdm1, inp.grad_ao)
    corr = vwn_c_e(rho) if correlation == "vwn" else pw92_c_e(rho)
    dens = abs_clip(torch.stack((lsda_x_e(rho), b88_x_e(rho, grho), corr, lyp_c_e(rho, grho, lap)), dim=1))
    x = torch.log(torch.abs(neural_features(rho, grho, tau)) + SQUASH_OFFSET)
    return xc_energy(inp.weights, coefficients(params, x), dens)


def custom_value_and_grads(inp: XCInputs, rdm1: np.ndarray, coefficients, params=None, correlation: str = "vwn"):
    """(E_xc, dE/dD, dE/dtheta) of custom_energy_from_rdm1."""
    D = torch.tensor(rdm1, dtype=DT, requires_grad=True)
    th = None if params is None else torch.tensor(params, dtype=DT, requires_grad=True)
    E = custom_energy_from_rdm1(inp, D, coefficients, th, correlation)
    grads = torch.autograd.grad(E, [D] + ([th] if th is not None else []))
    return float(E.detach()), grads[0].numpy(), (grads[1].numpy() if th is not None else None)


def to_inputs(ao, grad_ao, d2_ao, weights, chi=None) -> XCInputs:
    t = lambda a: None if a is None else torch.as_tensor(np.ascontiguousarray(a), dtype=DT)  # noqa
    return XCInputs(t(ao), t(grad_ao), t(d2_ao), t(weights), t(chi))
