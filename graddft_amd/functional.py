"""Functional plugin API over the gfx950 XC kernels.

Mirrors ``grad_dft/functional.py``: a :class:`Functional` maps a molecule's density
matrix to E_xc; :class:`NeuralFunctional` adds a coefficient MLP.  The reference lets a
functional be any composition of Python callables (``functional.py:128-141``); here a
functional is *declarative* -- a kind (``b3lyp`` energy densities, optionally weighted by
an MLP of fixed architecture) -- because every piece of it runs in hand-written kernels.

Entry points kept from the reference:

* ``Functional.xc_energy(params, molecule)``   -- ``functional.py:219-253``
* ``Functional.energy(params, molecule)``      -- ``functional.py:255-288``
  (E_xc + ``molecule.nonXC()`` when ``is_xc``)
* ``NeuralFunctional.init(seed)``             -- parameter initialisation (Glorot kernels,
  zero biases, as ``flax.linen.Dense`` with a Glorot initialiser)

Both are differentiable through :class:`XCEnergyFn` (a ``torch.autograd.Function``) whose
backward is the hand-written adjoint ``gdx_xc_backward`` (dE/dD and dE/dtheta).
:class:`XCEnergyVxcFn` returns (E_xc, dE/dD) -- the ``value_and_grad`` of
``train.py:86-121`` -- and is differentiable in turn: its backward is the hand-written
second-order pass ``gdx_xc_hvp``, which is what reverse mode through an SCF loop needs.
"""
from __future__ import annotations

import copy
import ctypes
import math
from dataclasses import dataclass
from typing import Callable, Optional

import numpy as np
import torch

from . import _lib
from .molecule import Molecule

PRECISIONS = {"fp64": _lib.GDX_FP64, "fp32": _lib.GDX_FP32, "bf16x6": _lib.GDX_BF16X6}


@dataclass(frozen=True)
class MLPSpec:
    """Coefficient network: log(|x| + 1e-4) -> [dense + ELU] x n_hidden -> dense ->
    2 sigmoid(z / 2) (the NeuralFunctional head, ``functional.py:407-419``)."""
    nc: int = 6
    h: int = 256
    n_hidden: int = 3
    nd: int = 4

    def shapes(self):
        dims = [self.nc] + [self.h] * self.n_hidden + [self.nd]
        return [(dims[i], dims[i + 1]) for i in range(len(dims) - 1)]

    def n_params(self) -> int:
        return sum(i * o + o for i, o in self.shapes())

    def flops_per_point(self) -> int:
        return 2 * sum(i * o for i, o in self.shapes())


# Tile shapes of the V_xc build measured on the device per shape (gdx_xc_autotune) when a plan
# is created; False keeps the fixed rule (64 for n_pad <= 64, else 128).
AUTOTUNE = True


def ao_planes_for(m: Molecule, precision: str) -> torch.Tensor:
    """AO planes in the element type of a precision: the FP64 planes, or their FP32 copy
    (made once and cached on the molecule)."""
    if precision == "fp64":
        return m.ao_planes
    key = ("ao_planes_f32", m.ao_planes.data_ptr())
    a = m.cache.get(key)
    if a is None:
        a = m.ao_planes.to(torch.float32)
        m.cache[key] = a
    return a


class XCPlan:
    """Descriptor + device workspace for one (molecule, functional, precision).

    The workspace carries forward intermediates to the backward call, so a plan serves one
    forward/backward pair at a time; ``generation`` guards against interleaving."""

    custom = False   # plans of custom coefficient functions serve custom_inputs/custom_vxc only

    def _no_custom(self):
        if self.custom:
            raise RuntimeError("this plan belongs to a custom coefficient function: use custom_inputs/custom_vxc")

    def __init__(self, molecule: Molecule, kind: int, spec: Optional[MLPSpec], precision: str,
                 coef=None, correlation: int = 0):
        lib = _lib.load()
        d = _lib.XCDesc()
        d.npts = molecule.grid_size
        d.npts_pad = molecule.npts_pad
        d.nao = molecule.nao
        d.nao_pad = molecule.nao_pad
        d.spin_restricted = 1 if molecule.restricted else 0
        d.kind = kind
        d.precision = PRECISIONS[precision]
        if spec is not None:
            d.mlp_nc, d.mlp_h, d.mlp_nhidden, d.mlp_nd = spec.nc, spec.h, spec.n_hidden, spec.nd
        d.correlation = int(correlation)
        if coef is not None:
            for f, c in enumerate(coef):
                d.coef[f] = float(c)
        self.desc = d
        nbytes = ctypes.c_size_t(0)
        _lib.check(lib.gdx_xc_workspace_size(ctypes.byref(d), ctypes.byref(nbytes)), "gdx_xc_workspace_size")
        self.workspace = torch.empty(int(nbytes.value), dtype=torch.uint8, device=molecule.ao_planes.device)
        self.precision = precision
        self.molecule = molecule
        self.generation = 0
        self.fwd_generation = -1
        self.hvp_workspace = None
        # measured choice of the V_xc output tile for this shape (first plan of a shape in the
        # process measures; later ones reuse it)
        self.vxc_tile = None
        if AUTOTUNE and molecule.ao_planes.is_cuda and not torch.cuda.is_current_stream_capturing():
            vt = ctypes.c_int32(0)
            st = torch.cuda.current_stream(molecule.ao_planes.device).cuda_stream
            _lib.check(lib.gdx_xc_autotune(st, ctypes.byref(d), self._ao().data_ptr(), self.workspace.data_ptr(),
                                           ctypes.byref(vt)), "gdx_xc_autotune")
            self.vxc_tile = int(vt.value)

    @staticmethod
    def _ptr(t: Optional[torch.Tensor]) -> int:
        return 0 if t is None else t.data_ptr()

    def _ao(self) -> torch.Tensor:
        """AO planes in the element type of the path (FP32 copy cached on the molecule)."""
        return ao_planes_for(self.molecule, self.precision)

    def forward(self, rdm1: torch.Tensor, params: Optional[torch.Tensor], exc: torch.Tensor) -> None:
        self._no_custom()
        m = self.molecule
        st = torch.cuda.current_stream(m.ao_planes.device).cuda_stream
        rc = _lib.load().gdx_xc_forward(st, ctypes.byref(self.desc), self._ao().data_ptr(),
                                        m.grid.weights_pad.data_ptr(), rdm1.data_ptr(),
                                        self._ptr(params), exc.data_ptr(), self.workspace.data_ptr())
        _lib.check(rc, "gdx_xc_forward")
        self.generation += 1
        self.fwd_generation = self.generation

    def densities(self, rdm1: torch.Tensor) -> torch.Tensor:
        """(N, 2, 6): rho, grad rho (x, y, z), tau, laplacian per spin."""
        m = self.molecule
        out = torch.empty((m.npts_pad, 2, 6), dtype=torch.float64, device=rdm1.device)
        st = torch.cuda.current_stream(m.ao_planes.device).cuda_stream
        rc = _lib.load().gdx_xc_densities(st, ctypes.byref(self.desc), self._ao().data_ptr(),
                                          rdm1.contiguous().data_ptr(), out.data_ptr(),
                                          self.workspace.data_ptr())
        _lib.check(rc, "gdx_xc_densities")
        self.generation += 1
        # the densities pass overwrites the forward intermediates a backward would read
        self.fwd_generation = -1
        return out[: m.grid_size]

    def backward(self, rdm1, params, g_exc, d_rdm1, d_params, generation: int) -> None:
        self._no_custom()
        if generation != self.fwd_generation:
            raise RuntimeError("XC workspace was overwritten by another forward before this backward")
        m = self.molecule
        st = torch.cuda.current_stream(m.ao_planes.device).cuda_stream
        rc = _lib.load().gdx_xc_backward(st, ctypes.byref(self.desc), self._ao().data_ptr(),
                                         m.grid.weights_pad.data_ptr(), rdm1.data_ptr(),
                                         self._ptr(params), g_exc.data_ptr(), self._ptr(d_rdm1),
                                         self._ptr(d_params), self.workspace.data_ptr())
        _lib.check(rc, "gdx_xc_backward")


    def custom_inputs(self, rdm1: torch.Tensor):
        """(e (N, 4), x (N, 6)): the B3LYP energy densities and the six coefficient inputs at
        rdm1 (gdx_xc_custom_inputs); the plan's workspace keeps the densities for custom_vxc."""
        m = self.molecule
        N = m.grid_size
        e = torch.empty((N, 4), dtype=torch.float64, device=rdm1.device)
        x = torch.empty((N, 6), dtype=torch.float64, device=rdm1.device)
        st = torch.cuda.current_stream(m.ao_planes.device).cuda_stream
        _lib.check(_lib.load().gdx_xc_custom_inputs(st, ctypes.byref(self.desc), self._ao().data_ptr(),
                                                    m.grid.weights_pad.data_ptr(), rdm1.contiguous().data_ptr(),
                                                    self.workspace.data_ptr(), e.data_ptr(), x.data_ptr()),
                   "gdx_xc_custom_inputs")
        self.generation += 1
        self.fwd_generation = self.generation
        return e, x

    def custom_vxc(self, g_e: torch.Tensor, g_x: torch.Tensor, generation: int) -> torch.Tensor:
        """dE/dD (2, n, n) from the cotangents of the custom inputs (gdx_xc_custom_vxc)."""
        if generation != self.fwd_generation:
            raise RuntimeError("XC workspace was overwritten by another forward before this backward")
        m = self.molecule
        out = torch.empty((2, m.nao, m.nao), dtype=torch.float64, device=g_e.device)
        st = torch.cuda.current_stream(m.ao_planes.device).cuda_stream
        _lib.check(_lib.load().gdx_xc_custom_vxc(st, ctypes.byref(self.desc), self._ao().data_ptr(),
                                                 g_e.contiguous().data_ptr(), g_x.contiguous().data_ptr(),
                                                 self.workspace.data_ptr(), out.data_ptr()), "gdx_xc_custom_vxc")
        return out

    def hvp(self, rdm1, params, direction, g_exc, exc=None, edot=None, d_rdm1=None, d_params=None) -> None:
        self._no_custom()
        """gdx_xc_hvp: gradients of g E + <direction, dE/dD> w.r.t. rdm1 and params."""
        m = self.molecule
        lib = _lib.load()
        if self.hvp_workspace is None:
            nbytes = ctypes.c_size_t(0)
            _lib.check(lib.gdx_xc_hvp_workspace_size(ctypes.byref(self.desc), ctypes.byref(nbytes)),
                       "gdx_xc_hvp_workspace_size")
            self.hvp_workspace = torch.empty(int(nbytes.value), dtype=torch.uint8, device=m.ao_planes.device)
        st = torch.cuda.current_stream(m.ao_planes.device).cuda_stream
        rc = lib.gdx_xc_hvp(st, ctypes.byref(self.desc), self._ao().data_ptr(), m.grid.weights_pad.data_ptr(),
                            rdm1.data_ptr(), self._ptr(params), direction.data_ptr(), g_exc.data_ptr(),
                            self._ptr(exc), self._ptr(edot), self._ptr(d_rdm1), self._ptr(d_params),
                            self.hvp_workspace.data_ptr())
        _lib.check(rc, "gdx_xc_hvp")


class XCEnergyFn(torch.autograd.Function):
    """E_xc(rdm1, params) with the hand-written adjoint as its backward."""

    @staticmethod
    def forward(ctx, rdm1, params, plan: XCPlan):
        rdm1 = rdm1.contiguous()
        p = None if params is None else params.contiguous()
        exc = torch.empty((), dtype=torch.float64, device=rdm1.device)
        plan.forward(rdm1, p, exc)
        ctx.plan = plan
        ctx.gen = plan.generation
        ctx.has_params = p is not None
        ctx.save_for_backward(rdm1, p if p is not None else rdm1)
        return exc

    @staticmethod
    def backward(ctx, g):
        rdm1, p = ctx.saved_tensors
        p = p if ctx.has_params else None
        d_rdm1 = torch.empty_like(rdm1) if ctx.needs_input_grad[0] else None
        d_params = torch.empty_like(p) if (p is not None and ctx.needs_input_grad[1]) else None
        g = g.detach().to(torch.float64).contiguous().reshape(())
        ctx.plan.backward(rdm1, p, g, d_rdm1, d_params, ctx.gen)
        return d_rdm1, d_params, None


class XCEnergyVxcFn(torch.autograd.Function):
    """(E_xc, dE_xc/dD) of (rdm1, params); backward = the second-order pass gdx_xc_hvp with
    the cotangent of dE/dD as the direction and the cotangent of E_xc as g."""

    @staticmethod
    def forward(ctx, rdm1, params, plan: XCPlan):
        rdm1 = rdm1.contiguous()
        p = None if params is None else params.contiguous()
        exc = torch.empty((), dtype=torch.float64, device=rdm1.device)
        vxc = torch.empty_like(rdm1)
        one = torch.ones((), dtype=torch.float64, device=rdm1.device)
        plan.forward(rdm1, p, exc)
        plan.backward(rdm1, p, one, vxc, None, plan.generation)
        ctx.plan = plan
        ctx.has_params = p is not None
        ctx.save_for_backward(rdm1, p if p is not None else rdm1)
        return exc, vxc

    @staticmethod
    @torch.autograd.function.once_differentiable
    def backward(ctx, g_exc, g_vxc):
        rdm1, p = ctx.saved_tensors
        p = p if ctx.has_params else None
        g = torch.zeros((), dtype=torch.float64, device=rdm1.device) if g_exc is None else \
            g_exc.detach().to(torch.float64).contiguous().reshape(())
        direction = torch.zeros_like(rdm1) if g_vxc is None else g_vxc.detach().contiguous()
        d_rdm1 = torch.empty_like(rdm1) if ctx.needs_input_grad[0] else None
        d_params = torch.empty_like(p) if (p is not None and ctx.needs_input_grad[1]) else None
        ctx.plan.hvp(rdm1, p, direction, g, d_rdm1=d_rdm1, d_params=d_params)
        return d_rdm1, d_params, None


class CustomInputsFn(torch.autograd.Function):
    """(e, x) = per-point energy densities and coefficient inputs of rdm1 (device kernels);
    backward: dE/dD from the cotangents (g_e, g_x) (point adjoint + V_xc build on the device)."""

    @staticmethod
    def forward(ctx, rdm1, plan: XCPlan):
        e, x = plan.custom_inputs(rdm1.detach())
        ctx.plan = plan
        ctx.gen = plan.generation
        return e, x

    @staticmethod
    @torch.autograd.function.once_differentiable
    def backward(ctx, g_e, g_x):
        plan = ctx.plan
        N = plan.molecule.grid_size
        dev = plan.molecule.ao_planes.device
        g_e = torch.zeros((N, 4), dtype=torch.float64, device=dev) if g_e is None else g_e
        g_x = torch.zeros((N, 6), dtype=torch.float64, device=dev) if g_x is None else g_x
        return plan.custom_vxc(g_e.to(torch.float64), g_x.to(torch.float64), ctx.gen), None


class Functional:
    """Declarative mirror of ``grad_dft.Functional`` for the kinds the kernels implement."""

    kind = _lib.GDX_FUNC_B3LYP
    spec: Optional[MLPSpec] = None
    spin_symmetric = True
    coef = None                       # constant coefficients of GDX_FUNC_LINEAR
    correlation = _lib.GDX_CORR_VWN5  # correlation in the third energy-density slot

    # exact exchange from the initial density ("initial": the reference's frozen chi) or from
    # the density being evaluated ("current", K[D] from the packed ERI every evaluation)
    exchange = "initial"

    def __init__(self, name: str = "B3LYP", is_xc: bool = True, exact_exchange: float = 0.0):
        self.name = name
        self.is_xc = is_xc
        # fraction a0 of exact exchange added outside the grid (B3LYP: 0.2)
        self.exact_exchange = exact_exchange

    def with_exchange(self, mode: str) -> "Functional":
        """A copy of this functional whose exact-exchange term uses K0 of the initial density
        (mode "initial", the reference's semantics: chi is built once from the initial density,
        interface/pyscf.py:152-160) or K[D] of the density being evaluated (mode "current":
        E_x = -(a0/2) sum_s tr(D_s K[D_s]), Fock term -a0 K[D], both differentiated through K)."""
        if mode not in ("initial", "current"):
            raise ValueError(f"exchange mode must be 'initial' or 'current', not {mode!r}")
        f = copy.copy(self)
        f.exchange = mode
        return f

    def plan(self, molecule: Molecule, precision: str = "fp64") -> XCPlan:
        key = ("xcplan", self.kind, self.spec, precision, molecule.restricted, self.coef, self.correlation)
        p = molecule.cache.get(key)
        if p is None or p.molecule.ao_planes is not molecule.ao_planes:
            p = XCPlan(molecule, self.kind, self.spec, precision, self.coef, self.correlation)
            molecule.cache[key] = p
        p.molecule = molecule
        return p

    def xc_energy(self, params: Optional[torch.Tensor], molecule: Molecule,
                  precision: str = "fp64", rdm1: Optional[torch.Tensor] = None) -> torch.Tensor:
        """Grid part of E_xc (functional.py:219-253), differentiable w.r.t. rdm1/params."""
        D = molecule.rdm1 if rdm1 is None else rdm1
        return XCEnergyFn.apply(D, params, self.plan(molecule, precision))

    def energy(self, params: Optional[torch.Tensor], molecule: Molecule, precision: str = "fp64"):
        """Total energy (functional.py:255-288): E_xc (+ nonXC when is_xc).  The exact-exchange
        part of B3LYP (from the molecule's initial density) is added to E_xc."""
        from .scf import ExchangeFn, exact_exchange_energy, non_xc_energy
        e = self.xc_energy(params, molecule, precision)
        if self.exact_exchange:
            K = ExchangeFn.apply(molecule.rdm1, molecule) if self.exchange == "current" else None
            e = e + exact_exchange_energy(molecule, molecule.rdm1, self.exact_exchange, K)
        if self.is_xc:
            e = e + non_xc_energy(molecule, molecule.rdm1)
        return e


class NeuralFunctional(Functional):
    """B3LYP energy densities weighted by a coefficient MLP (``functional.py:345-419``)."""

    kind = _lib.GDX_FUNC_NEURAL_B3LYP
    spin_symmetric = False

    def __init__(self, spec: MLPSpec = MLPSpec(), name: str = "neural_b3lyp", is_xc: bool = True,
                 exact_exchange: float = 0.0):
        super().__init__(name=name, is_xc=is_xc, exact_exchange=exact_exchange)
        self.spec = spec

    def init(self, seed: int = 0, device: str = "cuda") -> torch.Tensor:
        """Glorot-normal kernels and zero biases, flat [W0, b0, W1, b1, ...], W as (in, out)."""
        rng = np.random.default_rng(seed)
        parts = []
        for i, o in self.spec.shapes():
            parts.append(rng.normal(0.0, math.sqrt(2.0 / (i + o)), size=(i, o)).ravel())
            parts.append(np.zeros(o))
        return torch.as_tensor(np.concatenate(parts), dtype=torch.float64, device=device)


class B3LYPNeuralCorrection(NeuralFunctional):
    """B3LYP plus a neural correction: per-point coefficients c_f = c_B3LYP,f + MLP_theta(x)_f
    (linear head) on the four grid energy densities, with B3LYP's exact exchange (a0 = 0.2)
    from the ERI as for ``B3LYP``.  Zero last-layer weights reproduce B3LYP exactly."""

    kind = _lib.GDX_FUNC_B3LYP_NEURAL_CORR

    def __init__(self, spec: MLPSpec = MLPSpec(), name: str = "b3lyp_neural_corr", is_xc: bool = True):
        super().__init__(spec=spec, name=name, is_xc=is_xc, exact_exchange=0.2)


class CustomFunctional(Functional):
    """``Functional(coefficients=..., energy_densities=..., coefficient_inputs=...)`` with a user
    coefficient function (``functional.py:40-253``): E_xc = sum_r clip(w_r) clip(sum_f c_f(x_r)
    clip(e_f(r))) with the four B3LYP energy densities e_f (lda, b88, correlation, lyp) and the
    six coefficient inputs x = log(|.| + 1e-4) of (rho, |grad rho|^2, tau) per spin, both from
    the device kernels.  ``coefficients(params, x)`` is any torch function returning (N, 4)
    (an ``nn.Module`` call, a closure over its own parameters, ...); autograd runs through it,
    and dE/dD comes back through the device's point adjoint and V_xc build.  FP64.  Energies,
    V_xc (``energy_predictor``, the SCF loops' forward, ``scf_loop``) and dE/dtheta at a fixed
    density are supported; reverse mode through an SCF loop needs the second-order pass, which
    a custom function does not have, and raises NotImplementedError."""

    kind = _lib.GDX_FUNC_NEURAL_B3LYP
    spin_symmetric = False
    custom = True

    def __init__(self, coefficients: Callable, name: str = "custom", is_xc: bool = True,
                 exact_exchange: float = 0.0, correlation: str = "vwn"):
        super().__init__(name=name, is_xc=is_xc, exact_exchange=exact_exchange)
        if not callable(coefficients):
            raise TypeError("coefficients must be callable as coefficients(params, x)")
        if correlation not in ("vwn", "pw92"):
            raise ValueError("correlation must be 'vwn' or 'pw92'")
        self.coefficients = coefficients
        self.correlation = _lib.GDX_CORR_PW92 if correlation == "pw92" else _lib.GDX_CORR_VWN5
        # the kind's MLP fields only size the workspace: the smallest valid network
        self.spec = MLPSpec(nc=6, h=16, n_hidden=1, nd=4)

    def plan(self, molecule: Molecule, precision: str = "fp64") -> XCPlan:
        if precision != "fp64":
            raise NotImplementedError("custom coefficient functions run in FP64")
        key = ("xcplan-custom", molecule.restricted, self.correlation)
        p = molecule.cache.get(key)
        if p is None or p.molecule.ao_planes is not molecule.ao_planes:
            p = XCPlan(molecule, self.kind, self.spec, "fp64", None, self.correlation)
            p.custom = True
            molecule.cache[key] = p
        p.molecule = molecule
        return p

    def xc_energy(self, params, molecule: Molecule, precision: str = "fp64",
                  rdm1: Optional[torch.Tensor] = None) -> torch.Tensor:
        if precision != "fp64":
            raise NotImplementedError("custom coefficient functions run in FP64")
        D = molecule.rdm1 if rdm1 is None else rdm1
        e, x = CustomInputsFn.apply(D, self.plan(molecule, "fp64"))
        c = self.coefficients(params, x)
        if tuple(c.shape) != (molecule.grid_size, 4):
            raise ValueError(f"coefficients must return ({molecule.grid_size}, 4), got {tuple(c.shape)}")
        clip = lambda t: torch.where(t.abs() > 1e-30, t, torch.zeros_like(t))  # noqa: E731
        ed = clip((c.to(torch.float64) * clip(e)).sum(dim=1))
        return (clip(molecule.grid.weights) * ed).sum()

    def exc_vxc(self, params, molecule: Molecule, rdm1: torch.Tensor):
        """(E_xc, dE_xc/dD) at rdm1, not recorded for reverse mode through the caller."""
        with torch.enable_grad():
            D = rdm1.detach().requires_grad_(True)
            p = None if params is None else params.detach()
            E = self.xc_energy(p, molecule, rdm1=D)
            vxc, = torch.autograd.grad(E, [D])
        return E.detach(), vxc


class LinearFunctional(Functional):
    """A functional with constant coefficients over the grid energy densities (the reference's
    ``Functional(coefficients=lambda self, *_: jnp.array([[1.0]]), energy_densities=...)``
    built-ins, ``popular_functionals.py:375-393``): E = sum_r w clip(sum_f c_f e_f) with
    c over the slots (lda, b88, correlation, lyp); ``correlation`` is "vwn" (VWN5) or "pw92"."""

    kind = _lib.GDX_FUNC_LINEAR

    def __init__(self, coef, correlation: str = "vwn", name: str = "linear", is_xc: bool = True,
                 exact_exchange: float = 0.0):
        super().__init__(name=name, is_xc=is_xc, exact_exchange=exact_exchange)
        if len(coef) != 4:
            raise ValueError("coef: four coefficients (lda, b88, correlation, lyp)")
        self.coef = tuple(float(c) for c in coef)
        if correlation not in ("vwn", "pw92"):
            raise ValueError("correlation must be 'vwn' or 'pw92'")
        self.correlation = _lib.GDX_CORR_PW92 if correlation == "pw92" else _lib.GDX_CORR_VWN5
