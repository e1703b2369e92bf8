#!/usr/bin/env python3
"""Headline benchmark: XC grid points per second, forward + adjoint, on one MI355X.

Workload: benzene C6H6 at a seeded geometry (C-C 1.39 A, C-H 1.09 A, N(0, 0.02 A)
jitter), 6-31G* with Cartesian d (nao = 102), level-3 Becke grid, per-point features
rho, |grad rho|^2, tau per spin fed to a coefficient MLP (3 hidden layers of width 256,
ELU, Glorot init from the seed) weighting the four B3LYP energy densities.  One step is
E_xc plus its adjoints dE/dtheta and dE/dD over every grid point, with AO values, D and
theta already resident in HBM.

Prints ONE JSON line.  ``roofline`` is for the kernel with the largest share of the step,
timed with HIP events on its own stream inside the timed region; ``cpu_baseline`` is the
FP64 oracle (torch-CPU restatement of the reference) on a bounded sample of the same
grid.
"""
from __future__ import annotations

import argparse
import ctypes
import gc
import json
import math
import os
import sys
import time

import numpy as np

METRIC = "XC grid pts/s (fwd+adjoint) & ms/SCF iter at ≤1e-6 Ha vs FP64 GPU path"
PEAK = {"fp64": ("TFLOP/s", 78.6), "fp32": ("TFLOP/s", 157.3), "bf16x6": ("TFLOP/s", 2500.0)}
DTYPE = {"fp64": "f64", "fp32": "f32", "bf16x6": "bf16x6"}


def scf_roofline(npts, n, nsd, f_mlp, ms_per_iter, precision):
    """ms per SCF iteration against the roofline: algorithmic work per iteration W = N (n_s 16 n^2
    + 2 F_mlp) flops (density phi D and grad-phi D, V_xc, the MLP forward and its adjoint) at
    the precision's peak (FP64 78.6 TF/s; FP32 MFMA 157.3 TF/s for fp32 and bf16x6, whose grid
    GEMMs run in FP32), plus the packed ERI read once (8 n_pack bytes) at 8 TB/s."""
    P = n * (n + 1) // 2
    eri_bytes = 8 * (P * (P + 1) // 2)
    flops = npts * (nsd * 16 * n * n + 2 * f_mlp)
    peak = 78.6e12 if precision == "fp64" else 157.3e12
    t_c, t_m = flops / peak, eri_bytes / 8e12
    return {"flops_per_iter": int(flops), "eri_bytes_per_iter": int(eri_bytes), "peak_tflops": peak / 1e12,
            "roofline_us": round((t_c + t_m) * 1e6, 2), "compute_us": round(t_c * 1e6, 2),
            "hbm_us": round(t_m * 1e6, 2), "measured_us": round(ms_per_iter * 1e3, 2),
            "frac": round((t_c + t_m) / (ms_per_iter / 1e3), 4)}


def parse():
    ap = argparse.ArgumentParser(description=__doc__.splitlines()[0])
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--precision", default="fp32", choices=["fp64", "fp32", "bf16x6"])
    ap.add_argument("--no-scf", action="store_true", help="skip the ms/SCF-iteration measurement")
    ap.add_argument("--scf-cycles", type=int, default=15)
    ap.add_argument("--no-train", action="store_true", help="skip the training-step measurement")
    ap.add_argument("--no-batch", action="store_true", help="skip the concurrent small-molecule batch")
    ap.add_argument("--no-train-batch", action="store_true", help="skip the batched training step")
    ap.add_argument("--no-eri", action="store_true", help="skip the device ERI generation measurement")
    ap.add_argument("--no-config4", action="store_true", help="skip the naphthalene dense-ERI SCF workload")
    ap.add_argument("--no-config5", action="store_true", help="skip the 40-carbon grid-XC workload")
    ap.add_argument("--no-chi", action="store_true", help="skip the semi-numerical exchange (chi) measurement")
    ap.add_argument("--batch-size", type=int, default=64, help="alkanes in the batched training step")
    ap.add_argument("--batch-max-nc", type=int, default=8, help="largest n_C of the batched training step")
    ap.add_argument("--batch-init", default="b3lyp", choices=["b3lyp", "core"],
                    help="initial densities of the training batch: self-consistent B3LYP or the core guess")
    ap.add_argument("--no-graph", action="store_true", help="time the eager step instead of graph replays")
    ap.add_argument("--copy-only", type=int, default=0, metavar="BYTES",
                    help="only run a plain device copy of BYTES bytes (calibrates HBM byte counters)")
    ap.add_argument("--train-cycles", type=int, default=10)
    ap.add_argument("--seed", type=int, default=2)
    ap.add_argument("--grid-level", type=int, default=3)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-sample", type=int, default=16384, help="grid points in the CPU sample")
    return ap.parse_args()


def kernel_flops(name, npts, nao, spec, nsd):
    """Algorithmic flops of ONE launch of the named kernel (N real points)."""
    h = spec.h
    table = {
        "density": nsd * 8 * nao * nao * npts,       # 4 products (phi, dphi) x D
        "vxc": 2 * 8 * nao * nao * npts,             # phi^T Z + 3 dphi^T u dphi, per spin
        "gemm_fwd": 2 * npts * h * h,
        "gemm_bwd_data": 2 * npts * h * h,
        # the last one also forms the first-layer weight gradient (7 columns incl. the bias)
        "gemm_bwd_data_w0": 2 * npts * h * h + 2 * npts * h * (spec.nc + 1),
        "gemm_wgrad": 2 * npts * h * h,
    }
    for suffix in ("_f64", "_f32", "_bf16x6"):
        name = name.replace(suffix, "")
    return table.get(name)


def cpu_baseline(mol, theta, spec, sample):
    """FP64 oracle on the first `sample` grid points: returns (pts/s, threads, description)."""
    import torch
    from oracle.ao import eval_ao
    from oracle.xc import MLPSpec as OSpec, to_inputs, xc_value_and_grads
    nthreads = os.cpu_count() or 1
    torch.set_num_threads(nthreads)
    c = mol.grid.coords[:sample].cpu().numpy()
    w = mol.grid.weights[:sample].cpu().numpy()
    inp = to_inputs(*eval_ao(mol.basis, c), w)
    D = mol.rdm1.cpu().numpy()
    th = theta.cpu().numpy()
    ospec = OSpec(spec.nc, spec.h, spec.n_hidden, spec.nd)
    xc_value_and_grads("neural_b3lyp", inp, D, th, ospec)  # warm-up
    times = []
    t_start = time.perf_counter()
    while len(times) < 5 and (time.perf_counter() - t_start) < 20.0:
        t0 = time.perf_counter()
        xc_value_and_grads("neural_b3lyp", inp, D, th, ospec)
        times.append(time.perf_counter() - t0)
    med = float(np.median(times))
    return len(w) / med, torch.get_num_threads(), f"first {len(w)} points of the same grid, median of {len(times)} runs"


def scf_bench(args):
    """ms per SCF iteration: H2O (seeded geometry) def2-TZVP (nao 43), level-3 grid, B3LYP
    with exact exchange from the packed ERI, diff_scf_loop for args.scf_cycles cycles;
    headline = the graph-captured loop (mean of 5 replays), with the eager loop's time and
    per-kernel breakdown beside it."""
    import torch
    from graddft_amd import B3LYP, _lib, diff_scf_loop
    from graddft_amd.interface import geometry
    from graddft_amd.interface.builder import molecule_from_geometry
    sym, xyz = geometry.water(np.random.default_rng(args.seed))
    mol = molecule_from_geometry(sym, xyz, "def2-tzvp", 3, with_eri=True, name="water")
    out = {"workload": f"H2O def2-TZVP nao={mol.nao}, grid level 3 ({mol.grid_size} pts), B3LYP "
                       f"(exact exchange from packed ERI), diff_scf_loop {args.scf_cycles} cycles",
           "npts": mol.grid_size, "nao": mol.nao}
    reps = 5
    for prec in (["fp64", args.precision] if args.precision != "fp64" else ["fp64"]):
        # eager loop: per-kernel breakdown (HIP events around every launch)
        loop = diff_scf_loop(B3LYP, cycles=args.scf_cycles, precision=prec)
        res = loop(None, mol)  # warm-up (plans, workspaces, Cholesky of S, exchange of D0)
        torch.cuda.synchronize()
        with _lib.KernelTimer() as kt:
            t0 = time.perf_counter()
            res = loop(None, mol)
            torch.cuda.synchronize()
            t1 = time.perf_counter()
        # graph-captured loop: the whole diff_scf_loop replayed as one HIP graph
        gloop = diff_scf_loop(B3LYP, cycles=args.scf_cycles, precision=prec, graph=True)
        gloop(None, mol)       # eager run + capture
        gloop(None, mol)
        torch.cuda.synchronize()
        t2 = time.perf_counter()
        for _ in range(reps):
            gres = gloop(None, mol)
        torch.cuda.synchronize()
        t3 = time.perf_counter()
        # cycles + 1 energy/Fock evaluations (the initial one plus one per cycle)
        P = mol.nao * (mol.nao + 1) // 2
        eri_bytes = 8 * P * (P + 1) // 2
        ct = kt.result.get("coulomb_tile")
        coul = None
        if ct is not None:
            gbs = eri_bytes * ct[1] / (ct[0] / 1e3) / 1e9
            coul = {"eri_bytes": eri_bytes, "avg_launch_us": round(ct[0] / ct[1] * 1e3, 2), "GB/s": round(gbs, 1),
                    "frac_of_8TB/s": round(gbs / 8000.0, 4)}
        ms_it = (t3 - t2) / reps / args.scf_cycles * 1e3
        out[prec] = {"ms_per_iter": round(ms_it, 4),
                     "scf_roofline": scf_roofline(mol.grid_size, mol.nao, 1 if mol.restricted else 2, 0, ms_it, prec),
                     "coulomb_packed_eri": coul,
                     "ms_per_iter_eager": round((t1 - t0) / args.scf_cycles * 1e3, 4),
                     "energy": float(gres.energy),
                     "graph_vs_eager_abs_dE": abs(float(gres.energy) - float(res.energy)),
                     "kernels_ms_per_iter": {k: round(v[0] / args.scf_cycles, 4) for k, v in
                                             sorted(kt.result.items(), key=lambda kv: -kv[1][0])[:8]}}
    if args.precision != "fp64":
        out["abs_dE_vs_fp64_Ha"] = abs(out[args.precision]["energy"] - out["fp64"]["energy"])
    return out


def eri_bench(args):
    """Packed ERI generation for naphthalene (seeded geometry) def2-SVP, nao 180, ~1.3e8
    elements: the device kernels (gdx_eri_packed, HIP events, mean of 3 calls after a warm-up)
    and the plan build on the host, beside the host integral library (OpenMP, one call) on the
    same basis; max |device - host| over all elements."""
    import torch
    from graddft_amd import _lib
    from graddft_amd.interface import geometry
    from graddft_amd.interface.basis import build_basis, pack_shells
    from graddft_amd.interface.integrals import int2e_packed, npair
    sym, xyz = geometry.naphthalene(np.random.default_rng(args.seed))
    bs = build_basis(sym, xyz, "def2-svp")
    lib = _lib.load()
    pk = pack_shells(bs)
    ptr = lambda a: a.ctypes.data  # noqa: E731
    handle = ctypes.c_void_p(0)
    t0 = time.perf_counter()
    _lib.check(lib.gdx_eri_plan_create(len(bs.shells), ptr(pk["shell_i32"]), ptr(pk["shell_f64"]),
                                       ptr(pk["prim_f64"]), ptr(pk["tmat_f64"]), bs.nao, 0.0, ctypes.byref(handle)),
               "gdx_eri_plan_create")
    t_plan = time.perf_counter() - t0
    P = npair(bs.nao)
    nel = P * (P + 1) // 2
    eri = torch.empty(nel, dtype=torch.float64, device="cuda")
    st = torch.cuda.current_stream().cuda_stream
    try:
        _lib.check(lib.gdx_eri_packed(st, handle, eri.data_ptr()), "gdx_eri_packed")
        torch.cuda.synchronize()
        reps = 3
        with _lib.KernelTimer() as kt:
            t0 = time.perf_counter()
            for _ in range(reps):
                _lib.check(lib.gdx_eri_packed(st, handle, eri.data_ptr()), "gdx_eri_packed")
            torch.cuda.synchronize()
            t1 = time.perf_counter()
        nq = ctypes.c_int64(0)
        ne = ctypes.c_int64(0)
        lib.gdx_eri_plan_info(handle, ctypes.byref(nq), ctypes.byref(ne))
    finally:
        lib.gdx_eri_plan_destroy(handle)
    k_ms = kt.result["eri"][0] / reps
    h0 = time.perf_counter()
    host = int2e_packed(bs)
    t_host = time.perf_counter() - h0
    err = float(np.abs(eri.cpu().numpy() - host).max())
    return {"workload": f"naphthalene def2-SVP nao={bs.nao}, {len(bs.shells)} shells, packed ERI {nel} elements",
            "elements": nel, "shell_quartets": int(nq.value), "primitive_pairs": int(ne.value),
            "device_ms": round((t1 - t0) / reps * 1e3, 2), "device_kernel_ms": round(k_ms, 2),
            "device_elements_per_s": round(nel / (k_ms / 1e3), 1), "plan_ms": round(t_plan * 1e3, 1),
            "host_library_ms": round(t_host * 1e3, 1), "host_threads": int(os.environ.get("OMP_NUM_THREADS", os.cpu_count())),
            "speedup_vs_host": round(t_host * 1e3 / ((t1 - t0) / reps * 1e3 + t_plan * 1e3), 2),
            "max_abs_diff_vs_host": err}


def exchange_bench(mol):
    """K[D] (both spins) of the molecule's density from the exchange-ordered ERI copy (its
    one-time build timed separately) and from the gather over the packed ERI, per call, with
    the copy's streaming rate."""
    import torch
    from graddft_amd import scf
    n = mol.nao
    D = mol.rdm1.detach().contiguous()
    scf.release_exchange_xo(mol)

    def timed(fn, reps):
        fn()
        torch.cuda.synchronize()
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record()
        for _ in range(reps):
            r = fn()
        e1.record()
        torch.cuda.synchronize()
        return e0.elapsed_time(e1) / reps, r

    t0 = time.perf_counter()
    built = scf._exchange_xo(mol, n, D.device)
    torch.cuda.synchronize()
    t_build = time.perf_counter() - t0
    res = {"n": n, "xo_bytes": 8 * (n * n * (n * n + 1) // 2)}
    if built is None:
        res["xo"] = "over budget"
    else:
        ms_xo, K_xo = timed(lambda: scf.exchange(mol, D), 5)
        # both spins share one pass over the copy where the wave-per-row reader runs (n^2 >= 12000)
        passes = 1 if (D.shape[0] == 2 and n * n >= 12000) else D.shape[0]
        res.update({"xo_build_ms": round(1e3 * t_build, 1), "xo_ms_per_call": round(ms_xo, 3), "xo_passes": passes,
                    "xo_GB_per_s": round(passes * res["xo_bytes"] / (ms_xo * 1e6), 1)})
    scf.release_exchange_xo(mol)
    keep = scf.XO_MAX_BYTES
    scf.XO_MAX_BYTES = 0
    try:
        ms_g, K_g = timed(lambda: scf.exchange(mol, D), 2)
    finally:
        scf.XO_MAX_BYTES = keep
    res["gather_ms_per_call"] = round(ms_g, 3)
    if built is not None:
        res["speedup"] = round(ms_g / ms_xo, 2)
        res["max_rel_diff"] = float((K_xo - K_g).abs().max() / K_g.abs().max())
    return res


def config4_bench(args):
    """Naphthalene (seeded geometry) def2-SVP, nao 180, level-3 grid, packed ERI generated on
    the device (1.06 GB), starting from its self-consistent B3LYP density, B3LYP plus the
    width-256 neural correction (MLP 3x256): the 20-cycle
    diff_scf_loop forward replayed as one HIP graph (ms per iteration, mean of 3 replays) in
    FP64 and args.precision, with the eager loop's kernel breakdown, J's bandwidth on the packed
    ERI and the iteration's roofline; then forward + reverse (energy and dE/dtheta, eager) in
    both precisions with the energy and gradient differences."""
    import torch
    from graddft_amd import B3LYPNeuralCorrection, MLPSpec, _lib, diff_scf_loop
    from graddft_amd.interface import geometry
    from graddft_amd.interface.builder import molecule_from_geometry
    cycles = 20
    sym, xyz = geometry.naphthalene(np.random.default_rng(args.seed))
    from graddft_amd.interface.builder import converged_molecule
    t0 = time.perf_counter()
    mol = molecule_from_geometry(sym, xyz, "def2-svp", 3, with_eri=True, name="naphthalene")
    torch.cuda.synchronize()
    t_build = time.perf_counter() - t0
    # initial density: self-consistent B3LYP, as the reference's molecules carry a converged
    # mean-field density (molecule_from_pyscf after mf.kernel())
    t0 = time.perf_counter()
    mol = converged_molecule(mol)
    torch.cuda.synchronize()
    t_init = time.perf_counter() - t0
    spec = MLPSpec(nc=6, h=256, n_hidden=3, nd=4)
    f = B3LYPNeuralCorrection(spec)
    theta = f.init(seed=args.seed)
    nsd = 1 if mol.restricted else 2
    out = {"workload": f"naphthalene def2-SVP nao={mol.nao}, grid level 3 ({mol.grid_size} pts), packed ERI "
                       f"{mol.rep_tensor.numel()} elements (device-generated), B3LYP + neural correction MLP 3x256, "
                       f"diff_scf_loop {cycles} cycles",
           "npts": mol.grid_size, "nao": mol.nao, "molecule_build_s": round(t_build, 2),
           "initial_density": "self-consistent B3LYP (device CDIIS, converged_molecule)",
           "initial_density_s": round(t_init, 2)}
    out["exchange"] = exchange_bench(mol)
    P = mol.nao * (mol.nao + 1) // 2
    eri_bytes = 8 * (P * (P + 1) // 2)
    res = {}
    for prec in (["fp64", args.precision] if args.precision != "fp64" else ["fp64"]):
        loop = diff_scf_loop(f, cycles=cycles, precision=prec)
        with torch.no_grad():
            loop(theta, mol)
            torch.cuda.synchronize()
            with _lib.KernelTimer() as kt:
                t0 = time.perf_counter()
                r_e = loop(theta, mol)
                torch.cuda.synchronize()
                t1 = time.perf_counter()
            gloop = diff_scf_loop(f, cycles=cycles, precision=prec, graph=True)
            gloop(theta, mol)
            gloop(theta, mol)
            torch.cuda.synchronize()
            reps = 3
            t2 = time.perf_counter()
            for _ in range(reps):
                r_g = gloop(theta, mol)
            torch.cuda.synchronize()
            t3 = time.perf_counter()
        ms_it = (t3 - t2) / reps / cycles * 1e3
        ct = kt.result.get("coulomb_tile")
        coul = None
        if ct is not None:
            gbs = eri_bytes * ct[1] / (ct[0] / 1e3) / 1e9
            coul = {"eri_bytes": eri_bytes, "avg_launch_us": round(ct[0] / ct[1] * 1e3, 2), "GB/s": round(gbs, 1),
                    "frac_of_8TB/s": round(gbs / 8000.0, 4)}
        # forward + reverse through the loop
        th = theta.detach().clone().requires_grad_(True)
        r = loop(th, mol)
        torch.autograd.grad(r.energy, [th])
        torch.cuda.synchronize()
        th = theta.detach().clone().requires_grad_(True)
        t4 = time.perf_counter()
        r = loop(th, mol)
        g, = torch.autograd.grad(r.energy, [th])
        torch.cuda.synchronize()
        t5 = time.perf_counter()
        res[prec] = (float(r_g.energy), g.detach().cpu().numpy())
        out[prec] = {"ms_per_iter": round(ms_it, 4), "ms_per_iter_eager": round((t1 - t0) / cycles * 1e3, 4),
                     "scf_roofline": scf_roofline(mol.grid_size, mol.nao, nsd, spec.flops_per_point(), ms_it, prec),
                     "coulomb_packed_eri": coul, "energy": float(r_g.energy),
                     "graph_vs_eager_abs_dE": abs(float(r_g.energy) - float(r_e.energy)),
                     "fwd_rev_ms_per_step": round((t5 - t4) * 1e3, 2),
                     "kernels_ms_per_iter": {k: round(v[0] / cycles, 4) for k, v in
                                             sorted(kt.result.items(), key=lambda kv: -kv[1][0])[:10]}}
    if args.precision != "fp64":
        e64, g64 = res["fp64"]
        e1, g1 = res[args.precision]
        out["abs_dE_vs_fp64_Ha"] = abs(e1 - e64)
        out["grad_rel_err_vs_fp64"] = float(np.linalg.norm(g1 - g64) / np.linalg.norm(g64))
        out["speedup_ms_per_iter"] = round(out["fp64"]["ms_per_iter"] / out[args.precision]["ms_per_iter"], 3)
    return out


def config5_bench(args):
    """40 carbon atoms at seeded positions (12 A cube), def2-SVP (nao 560), level-2 grid, D from
    120 seeded S-orthonormal orbitals, the 3x256 neural functional: one E_xc + dE/dD + dE/dtheta
    evaluation (forward + adjoint, eager, mean of 3 after one warm-up) in FP64, FP32 and split-BF16,
    with the differences against FP64 and the step's MFMA work.  Also the input build of this
    molecule: Becke weights and S, T, V on the device against the host paths."""
    import torch
    from graddft_amd import MLPSpec, NeuralFunctional
    from graddft_amd.interface import geometry
    from graddft_amd.interface.basis import atomic_numbers, build_basis
    from graddft_amd.interface.builder import molecule_from_geometry
    from graddft_amd.interface.grids import molecular_grid
    from graddft_amd.interface.integrals import int1e, int1e_device
    rng = np.random.default_rng(5)
    sym, xyz = geometry.carbon_cloud(rng)
    Z = atomic_numbers(sym)
    bs = build_basis(sym, xyz, "def2-svp")
    level = 2
    inputs = {}
    t0 = time.perf_counter()
    molecular_grid(Z, xyz, level, weights_on="device")
    inputs["grid_device_s"] = round(time.perf_counter() - t0, 3)
    t0 = time.perf_counter()
    molecular_grid(Z, xyz, level)
    inputs["grid_host_s"] = round(time.perf_counter() - t0, 3)
    int1e_device(bs, Z, xyz)
    t0 = time.perf_counter()
    S_d, _, _ = int1e_device(bs, Z, xyz)
    inputs["int1e_device_ms"] = round((time.perf_counter() - t0) * 1e3, 2)
    t0 = time.perf_counter()
    S, _, _ = int1e(bs, Z, xyz)
    inputs["int1e_host_ms"] = round((time.perf_counter() - t0) * 1e3, 2)
    inputs["host_threads"] = torch.get_num_threads()
    C = rng.normal(size=(bs.nao, 120))
    w, V = np.linalg.eigh(C.T @ S @ C)
    C = C @ V @ np.diag(w ** -0.5) @ V.T
    t0 = time.perf_counter()
    mol = molecule_from_geometry(sym, xyz, "def2-svp", level, rdm1=C @ C.T, name="carbon40")
    torch.cuda.synchronize()
    inputs["molecule_build_s"] = round(time.perf_counter() - t0, 2)
    # AO planes of this grid (5 x npts_pad x nao_pad doubles, written once)
    from graddft_amd.interface.builder import eval_ao_device
    gc_np = mol.grid.coords.cpu().numpy()
    planes = eval_ao_device(bs, gc_np, mol.npts_pad, mol.nao_pad)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    planes = eval_ao_device(bs, gc_np, mol.npts_pad, mol.nao_pad)
    torch.cuda.synchronize()
    ao_s = time.perf_counter() - t0
    inputs["ao_eval_ms"] = round(ao_s * 1e3, 2)
    inputs["ao_planes_bytes"] = planes.numel() * 8
    inputs["ao_eval_GB_per_s"] = round(planes.numel() * 8 / ao_s / 1e9, 1)
    del planes
    spec = MLPSpec(nc=6, h=256, n_hidden=3, nd=4)
    f = NeuralFunctional(spec)
    theta = f.init(seed=6)
    N, n = mol.grid_size, mol.nao
    nsd = 1 if mol.restricted else 2
    w_pt = 8 * n * n * nsd + 8 * n * n * 2 + 3 * spec.flops_per_point()
    out = {"workload": f"40 C (seeded, 12 A cube) def2-SVP nao={n}, grid level {level} ({N} pts), D from 120 "
                       "orbitals, MLP 3x256: E_xc + dE/dD + dE/dtheta", "npts": N, "nao": n, "inputs": inputs,
           "flop_per_point": w_pt}
    D = mol.rdm1.contiguous()
    one = torch.ones((), dtype=torch.float64, device="cuda")
    res = {}
    reps = 3
    for prec in ("fp64", "fp32", "bf16x6"):
        plan = f.plan(mol, prec)
        exc = torch.empty((), dtype=torch.float64, device="cuda")
        dD = torch.empty_like(D)
        dth = torch.empty_like(theta)
        plan.forward(D, theta, exc)
        plan.backward(D, theta, one, dD, dth, plan.generation)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(reps):
            plan.forward(D, theta, exc)
            plan.backward(D, theta, one, dD, dth, plan.generation)
        torch.cuda.synchronize()
        ms = (time.perf_counter() - t0) / reps * 1e3
        # split-BF16 mixes bf16 products (hidden layers) with FP32 ones: no single peak
        peak = PEAK[prec][1] if prec != "bf16x6" else None
        tf = w_pt * N / (ms / 1e3) / 1e12
        res[prec] = (float(exc), dD.clone(), dth.clone())
        out[prec] = {"ms_per_step": round(ms, 3), "grid_pts_per_s": round(N / (ms / 1e3), 1),
                     "algorithmic_tflops": round(tf, 2), "peak": peak,
                     "frac": round(tf / peak, 4) if peak else None}
        del plan
    e64, d64, t64 = res["fp64"]
    rel = lambda a, b: float(torch.linalg.norm(a - b) / torch.linalg.norm(b))  # noqa: E731
    for prec in ("fp32", "bf16x6"):
        e1, d1, t1 = res[prec]
        out[prec].update({"abs_dE_vs_fp64_Ha": abs(e1 - e64), "rel_err_dE_dD": rel(d1, d64),
                          "rel_err_dE_dtheta": rel(t1, t64),
                          "speedup_vs_fp64": round(out["fp64"]["ms_per_step"] / out[prec]["ms_per_step"], 3)})
    return out


def chi_bench(args):
    """Semi-numerical exchange: chi (generate_chi_tensor at omega = 0) on the device for water
    def2-TZVP (f shells) and benzene 6-31G* on level-3 grids, after one warm-up build (mean of 3;
    kernel split from per-launch events).  Work unit: one (grid point, primitive-pair entry)
    evaluation of the nu integrals (Boys function, Hermite R, contraction).  CPU: the FP64 numpy
    oracle (nu on the grid, then the chi contraction) on the first 128 points, one process."""
    import torch
    from graddft_amd import _lib, hf
    from graddft_amd.interface import geometry
    from graddft_amd.interface.builder import molecule_from_geometry
    from oracle.hf import chi_tensor, nu_on_grid
    out = {}
    for name, sym_xyz, basis in (("water", geometry.water(np.random.default_rng(args.seed)), "def2-tzvp"),
                                 ("benzene", geometry.benzene(np.random.default_rng(args.seed)), "6-31g*")):
        sym, xyz = sym_xyz
        mol = molecule_from_geometry(sym, xyz, basis, 3, name=name)
        N, n = mol.grid_size, mol.nao
        t0 = time.perf_counter()
        plan = hf.hf_plan(mol)
        t_plan = time.perf_counter() - t0
        entries, _ = plan.info()
        hf.generate_chi_tensor(mol, mol.rdm1, [0.0])
        torch.cuda.synchronize()
        reps = 3
        t0 = time.perf_counter()
        for _ in range(reps):
            chi = hf.generate_chi_tensor(mol, mol.rdm1, [0.0])
        torch.cuda.synchronize()
        ms = (time.perf_counter() - t0) / reps * 1e3
        with _lib.KernelTimer() as kt:
            hf.generate_chi_tensor(mol, mol.rdm1, [0.0])
            torch.cuda.synchronize()
        S = 128
        c = mol.grid.coords[:S].cpu().numpy()
        ao = mol.ao_planes[0][:S, :n].cpu().numpy()
        D = mol.rdm1.cpu().numpy()
        t0 = time.perf_counter()
        chi_o = chi_tensor(D[:1], ao, nu_on_grid(mol.basis, c, 0.0))
        t_cpu = time.perf_counter() - t0
        got = chi[0, :1, :, :S].permute(2, 0, 1).cpu().numpy()
        out[name] = {"basis": basis, "nao": n, "npts": N, "primitive_pair_entries": entries,
                     "plan_ms": round(t_plan * 1e3, 1), "ms_per_chi": round(ms, 3),
                     "grid_pts_per_s": round(N / (ms / 1e3), 1),
                     "entry_point_evals_per_s": round(entries * N / (ms / 1e3), 1),
                     "kernels_ms": {k: round(v[0], 3) for k, v in sorted(kt.result.items(), key=lambda kv: -kv[1][0])},
                     "cpu_baseline": {"value": round(S / t_cpu, 1), "unit": "grid_pts/s", "kind": "port",
                                      "cores": 1,
                                      "sample": f"FP64 numpy oracle, chi at the first {S} grid points"},
                     "max_rel_diff_vs_oracle_sample": float(np.abs(got - chi_o).max() / np.abs(chi_o).max())}
        del mol, plan, chi
    return out


def batch_bench(args):
    """A batch of small molecules evaluated concurrently: CH4, C2H6, C3H8, C4H10 (seeded
    geometries) def2-SVP, level-1 grids, B3LYP with exact exchange, diff_scf_loop of
    args.scf_cycles cycles; the molecules' captured loops replayed one after another on one
    stream vs batched_scf_loop (replays round-robin on 4 streams).  Mean of 5 batches each."""
    import torch
    from graddft_amd import B3LYP, batched_scf_loop, diff_scf_loop
    from graddft_amd.interface import geometry
    from graddft_amd.interface.builder import molecule_from_geometry
    rng = np.random.default_rng(args.seed)
    mols = []
    for nc in (1, 2, 3, 4):
        sym, xyz = geometry.alkane(rng, nc)
        mols.append(molecule_from_geometry(sym, xyz, "def2-svp", 1, with_eri=True, name=f"C{nc}"))
    prec = args.precision
    seq = diff_scf_loop(B3LYP, cycles=args.scf_cycles, precision=prec, graph=True)
    bat = batched_scf_loop(B3LYP, cycles=args.scf_cycles, precision=prec, n_streams=4)
    bat(None, mols)                       # eager runs + captures
    bat(None, mols)
    torch.cuda.synchronize()
    reps = 5
    t0 = time.perf_counter()
    for _ in range(reps):
        seq_out = [seq(None, m) for m in mols]
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    for _ in range(reps):
        bat_out = bat(None, mols)
    torch.cuda.synchronize()
    t2 = time.perf_counter()
    dE = max(abs(float(a.energy) - float(b.energy)) for a, b in zip(seq_out, bat_out))
    ms_seq, ms_bat = (t1 - t0) / reps * 1e3, (t2 - t1) / reps * 1e3
    return {"workload": "CH4, C2H6, C3H8, C4H10 def2-SVP (nao " + ", ".join(str(m.nao) for m in mols)
                        + f"), level-1 grids, B3LYP, diff_scf_loop {args.scf_cycles} cycles, {prec}",
            "ms_per_batch_sequential_graphs": round(ms_seq, 3), "ms_per_batch_4_streams": round(ms_bat, 3),
            "speedup": round(ms_seq / ms_bat, 3), "max_abs_dE_Ha": dE}


def train_bench(args):
    """One training step through the differentiable SCF loop: H2O (seeded geometry)
    def2-TZVP, level-3 grid, neural B3LYP with an MLP of 3 x 128, args.train_cycles-cycle
    diff_scf_loop forward + reverse mode to d loss / d theta (simple_energy_loss).  ms_per_step:
    the Python-driven step (mean of 5); ms_per_step_graph: the loop and its reverse pass replayed
    as captured graphs (mean of 5); kernel breakdown and total from one more eager step with
    per-launch events (which slow that step down; it is not timed)."""
    import torch
    from graddft_amd import MLPSpec, NeuralFunctional, _lib, diff_scf_loop, simple_energy_loss
    from graddft_amd.interface import geometry
    from graddft_amd.interface.builder import molecule_from_geometry
    sym, xyz = geometry.water(np.random.default_rng(args.seed))
    mol = molecule_from_geometry(sym, xyz, "def2-tzvp", 3, with_eri=True, name="water")
    f = NeuralFunctional(MLPSpec(nc=6, h=128, n_hidden=3, nd=4))
    theta = f.init(seed=args.seed)
    out = {"workload": f"H2O def2-TZVP nao={mol.nao}, grid level 3 ({mol.grid_size} pts), neural B3LYP "
                       f"MLP 3x128, diff_scf_loop {args.train_cycles} cycles forward + reverse "
                       "(simple_energy_loss)"}
    grads = {}
    reps = 5
    for prec in (["fp64", args.precision] if args.precision != "fp64" else ["fp64"]):
        loop = diff_scf_loop(f, cycles=args.train_cycles, precision=prec)
        simple_energy_loss(theta, loop, mol, -76.4)   # warm-up
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(reps):
            (cost, e), g = simple_energy_loss(theta, loop, mol, -76.4)
        torch.cuda.synchronize()
        ms_eager = (time.perf_counter() - t0) / reps * 1e3
        with _lib.KernelTimer() as kt:
            simple_energy_loss(theta, loop, mol, -76.4)
            torch.cuda.synchronize()
        # the loop and its reverse pass as two captured graphs (graph=True under autograd)
        gloop = diff_scf_loop(f, cycles=args.train_cycles, precision=prec, graph=True)
        simple_energy_loss(theta, gloop, mol, -76.4)  # eager run + captures
        simple_energy_loss(theta, gloop, mol, -76.4)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(reps):
            (cost_g, e_g), g_g = simple_energy_loss(theta, gloop, mol, -76.4)
        torch.cuda.synchronize()
        ms_graph = (time.perf_counter() - t0) / reps * 1e3
        grads[prec] = g_g.cpu().numpy()
        out[prec] = {"ms_per_step": round(ms_eager, 3), "ms_per_step_graph": round(ms_graph, 3),
                     "kernels_total_ms_per_step": round(sum(v[0] for v in kt.result.values()), 3),
                     "kernel_launches_per_step": sum(v[1] for v in kt.result.values()),
                     "energy": float(e_g), "graph_vs_eager_abs_dE": abs(float(e_g) - float(e)),
                     "graph_vs_eager_grad_rel": float(torch.linalg.norm(g_g - g) / torch.linalg.norm(g)),
                     "kernels_ms_per_step": {k: round(v[0], 4) for k, v in
                                             sorted(kt.result.items(), key=lambda kv: -kv[1][0])[:8]}}
    if args.precision != "fp64":
        out["abs_dE_vs_fp64_Ha"] = abs(out[args.precision]["energy"] - out["fp64"]["energy"])
        g0, g1 = grads["fp64"], grads[args.precision]
        out["grad_rel_err_vs_fp64"] = float(np.linalg.norm(g1 - g0) / np.linalg.norm(g0))
    return out


def _release_xc_plans(mols):
    """Drop the XC plans (per molecule and batched) cached on the molecules and return their
    device memory: the batched plan's workspaces are allocated by the library, outside torch's
    caching allocator, and molecules referenced by their plans are freed only by the cyclic
    garbage collector."""
    import gc
    import torch
    for m in mols:
        for k in [k for k in m.cache if isinstance(k, tuple) and k and k[0] in ("xcplan", "xcbatch")]:
            plan = m.cache.pop(k)
            if hasattr(plan, "close"):
                plan.close()
    gc.collect()
    torch.cuda.empty_cache()


def train_batch_bench(args):
    """The ragged training batch: args.batch_size linear alkanes C_nH_2n+2 with n_C ~
    U{1..args.batch_max_nc} from the seed (all-trans geometries, N(0, 0.02 A) jitter),
    def2-SVP, level-2 grids, device-generated packed ERIs, neural B3LYP with an MLP of 3 x 128;
    one training step = mse_energy_loss over args.train_cycles-cycle differentiable SCF loops,
    forward + reverse to d loss / d theta, with the XC evaluation of every cycle batched over
    the molecules (batched_diff_scf_loop: one MLP pass over all grid points, ragged AO-side
    launches).  One warm-up and one timed step per precision.  Roofline: W_pt = n_s 32 n^2 +
    6 F_mlp flops per grid point and cycle (forward, SCF adjoint with the second-order pass;
    n_s = 2) at the precision's peak.  Runs: FP64; the mixed precision in every cycle
    (refine_cycles=0); and the mixed precision with the loops' default last FP64 cycle.  The
    molecules start from self-consistent B3LYP densities (--batch-init b3lyp, as the
    reference's molecules carry a converged PySCF density) or the core guess.  Per molecule,
    |E - E_fp64| is reported: loops far from self-consistency amplify any perturbation."""
    import torch
    from graddft_amd import MLPSpec, NeuralFunctional, batched_diff_scf_loop
    from graddft_amd.interface import geometry
    from graddft_amd.interface.builder import molecule_from_geometry
    rng = np.random.default_rng(args.seed)
    ncs = [int(x) for x in rng.integers(1, args.batch_max_nc + 1, args.batch_size)]
    from graddft_amd.interface.builder import converged_molecule
    t0 = time.perf_counter()
    mols = []
    for nc in ncs:
        sym, xyz = geometry.alkane(rng, nc)
        mols.append(molecule_from_geometry(sym, xyz, "def2-svp", 2, with_eri=True, name=f"C{nc}", npts_align=256))
    torch.cuda.synchronize()
    t_build = time.perf_counter() - t0
    # the reference's molecules carry a converged mean-field density (molecule_from_pyscf after
    # mf.kernel()): self-consistent B3LYP here, unless the core-Hamiltonian guess is asked for
    t0 = time.perf_counter()
    if args.batch_init == "b3lyp":
        mols = [converged_molecule(m) for m in mols]
    torch.cuda.synchronize()
    t_init = time.perf_counter() - t0
    spec = MLPSpec(nc=6, h=128, n_hidden=3, nd=4)
    f = NeuralFunctional(spec)
    theta = f.init(seed=args.seed)
    truth = torch.tensor([-40.0 * nc for nc in ncs], dtype=torch.float64, device="cuda")
    cyc = args.train_cycles
    npts = [m.grid_size for m in mols]
    ns = [m.nao for m in mols]
    flops = sum(N * cyc * (2 * 32 * n * n + 6 * spec.flops_per_point()) for N, n in zip(npts, ns))
    out = {"workload": f"{len(mols)} alkanes C_nH_2n+2, n_C ~ U{{1..{args.batch_max_nc}}} (seeded), def2-SVP (nao "
                       f"{min(ns)}-{max(ns)}), level-2 grids ({sum(npts)} pts in total), neural B3LYP MLP 3x128, "
                       f"mse_energy_loss over {cyc}-cycle diff_scf_loops, forward + reverse, XC batched per cycle",
           "n_molecules": len(mols), "npts_total": sum(npts), "molecule_build_s": round(t_build, 1),
           "initial_density": "self-consistent B3LYP (device CDIIS, converged_molecule)" if args.batch_init == "b3lyp" else "core Hamiltonian",
           "initial_density_s": round(t_init, 1),
           "flops_per_step": int(flops)}
    grads, energies = {}, {}
    runs = [("fp64", "fp64", 0)]
    if args.precision != "fp64":
        # every cycle in the mixed precision, and the loops' default for mixed precisions: the
        # last cycle (and the returned energy) on the FP64 path, which also holds the FP64 batch
        # plan (its second-order workspace is 81 GB at this size)
        runs += [(args.precision, args.precision, 0), (args.precision + "_fp64_last_cycle", args.precision, 1)]
    for key, prec, refine in runs:
        loop = batched_diff_scf_loop(f, cycles=cyc, precision=prec, refine_cycles=refine)

        def step():
            th = theta.detach().clone().requires_grad_(True)
            E = torch.stack([o.energy for o in loop(th, mols)])
            loss = ((E - truth) ** 2).mean()
            g, = torch.autograd.grad(loss, [th])
            return float(loss.detach()), g, E.detach()

        _release_xc_plans(mols)                  # the previous run's plans and workspaces
        free, _ = torch.cuda.mem_get_info()
        print(f"[bench] train_batch {key}: device memory free {free / 2**30:.1f} GiB", file=sys.stderr, flush=True)
        step()                                   # warm-up: plans, workspaces
        torch.cuda.synchronize()
        t1 = time.perf_counter()
        loss, g, E = step()
        torch.cuda.synchronize()
        dt = time.perf_counter() - t1
        peak = 78.6e12 if prec == "fp64" else 157.3e12
        grads[key] = g.cpu().numpy()
        energies[key] = E.cpu().numpy()
        out[key] = {"s_per_step": round(dt, 3), "loss": loss, "peak_tflops": peak / 1e12,
                    "achieved_tflops": round(flops / dt / 1e12, 2), "roofline_frac": round(flops / peak / dt, 4),
                    "grid_pts_cycles_per_s": round(sum(npts) * cyc / dt, 1), "refine_cycles": refine}
        if key != "fp64":
            g0 = grads["fp64"]
            dE = np.abs(energies[key] - energies["fp64"])
            out[key].update({"grad_rel_err_vs_fp64": float(np.linalg.norm(grads[key] - g0) / np.linalg.norm(g0)),
                             "per_molecule_abs_dE_Ha": {"max": float(dE.max()), "median": float(np.median(dE)),
                                                        "within_1e-6": int((dE <= 1e-6).sum()), "of": len(mols)},
                             "step_time_over_fp64": round(dt / out["fp64"]["s_per_step"], 3)})
    _release_xc_plans(mols)
    for m in mols:
        m.cache.clear()
    return out


def measured_traffic():
    """{timer name: HBM bytes per launch} from the committed PMC summaries (profiles/traffic_*.txt,
    lines 'traffic <name> <bytes>'), measured on this workload with rocprofv3."""
    import glob
    out = {}
    for path in sorted(glob.glob(os.path.join(os.path.dirname(os.path.abspath(__file__)), "profiles",
                                              "traffic_*.txt"))):
        for line in open(path):
            parts = line.split()
            if len(parts) == 3 and parts[0] == "traffic":
                out[parts[1]] = int(parts[2])
    return out


def copy_only(nbytes: int):
    """A plain device-to-device copy of a known size, five times: the reference launch for
    scaling the FETCH_SIZE / WRITE_SIZE counters of a profiled kernel to HBM bytes."""
    import torch
    n = nbytes // 4
    x = torch.ones(n, dtype=torch.float32, device="cuda")
    y = torch.empty_like(x)
    for _ in range(5):
        y.copy_(x)
    torch.cuda.synchronize()
    print(json.dumps({"copy_bytes": n * 4, "launches": 5}), flush=True)


def main():
    args = parse()
    if args.gpus != 1:
        print(f"bench.py measures one GPU; --gpus {args.gpus} is not supported in this build "
              "(the path shards grid points across GPUs by design, but only one GPU is used here)",
              file=sys.stderr)
        sys.exit(2)
    if args.copy_only:
        copy_only(args.copy_only)
        return
    import torch
    from graddft_amd import MLPSpec, NeuralFunctional, _lib
    from graddft_amd.evaluate import capture_graph
    from graddft_amd.interface import geometry
    from graddft_amd.interface.builder import molecule_from_geometry

    rng = np.random.default_rng(args.seed)
    sym, xyz = geometry.benzene(rng)
    mol = molecule_from_geometry(sym, xyz, "6-31g*", args.grid_level, name="benzene")
    spec = MLPSpec(nc=6, h=256, n_hidden=3, nd=4)
    f = NeuralFunctional(spec)
    theta = f.init(seed=args.seed)
    N, n = mol.grid_size, mol.nao
    plan = f.plan(mol, args.precision)
    D = mol.rdm1.contiguous()
    exc = torch.empty((), dtype=torch.float64, device="cuda")
    one = torch.ones((), dtype=torch.float64, device="cuda")
    dD = torch.empty_like(D)
    dth = torch.empty_like(theta)

    def step():
        plan.forward(D, theta, exc)
        plan.backward(D, theta, one, dD, dth, plan.generation)

    for _ in range(args.warmup):
        step()
    torch.cuda.synchronize()
    # eager timed region with per-launch HIP events: per-kernel durations and the roofline
    with _lib.KernelTimer() as kt:
        t0 = time.perf_counter()
        for _ in range(args.steps):
            step()
        torch.cuda.synchronize()
        t1 = time.perf_counter()
    ms_eager = (t1 - t0) / args.steps * 1e3
    E = float(exc)
    # graph timed region: the same step captured once as a HIP graph and replayed
    ms = ms_eager
    if not args.no_graph:
        graph = torch.cuda.CUDAGraph()
        side = torch.cuda.Stream()
        side.wait_stream(torch.cuda.current_stream())
        with torch.cuda.stream(side):
            step()
        torch.cuda.current_stream().wait_stream(side)
        capture_graph(graph, step)
        for _ in range(args.warmup):
            graph.replay()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(args.steps):
            graph.replay()
        torch.cuda.synchronize()
        t1 = time.perf_counter()
        ms = (t1 - t0) / args.steps * 1e3
        if abs(float(exc) - E) > 1e-12 * abs(E):
            raise RuntimeError(f"graph replay changed E_xc: {float(exc)} vs {E}")
    value = N / (ms / 1e3)

    # dominant kernel and its roofline
    name, (tot_ms, cnt) = max(kt.result.items(), key=lambda kv: kv[1][0])
    nsd = 1 if mol.restricted else 2
    flops = kernel_flops(name, N, n, spec, nsd)
    unit, peak = PEAK[args.precision]
    roof = None
    if flops is not None:
        avg_s = tot_ms / cnt / 1e3
        achieved = flops / avg_s / 1e12
        tr = measured_traffic().get(name)
        roof = {"bound": "mfma", "kernel": name, "achieved": round(achieved, 3), "peak": peak,
                "unit": unit, "frac": round(achieved / peak, 4), "traffic": tr,
                "traffic_unit": "bytes per launch (HBM, rocprofv3 FETCH_SIZE/WRITE_SIZE scaled by a "
                                "copy-kernel calibration; profiles/traffic_*.txt)" if tr else None,
                "avg_launch_us": round(avg_s * 1e6, 2), "launches_per_step": cnt // args.steps,
                "flops_per_launch": flops}
    # accuracy and speed against the plain FP64 GPU path on the same inputs
    vs64 = None
    if args.precision != "fp64":
        p64 = f.plan(mol, "fp64")
        e64 = torch.empty_like(exc)
        dD64 = torch.empty_like(dD)
        dth64 = torch.empty_like(dth)

        def step64():
            p64.forward(D, theta, e64)
            p64.backward(D, theta, one, dD64, dth64, p64.generation)

        step64()
        torch.cuda.synchronize()
        k64 = max(3, args.steps // 2)
        run64 = step64
        if not args.no_graph:      # same timing method as the measured path
            g64 = torch.cuda.CUDAGraph()
            side = torch.cuda.Stream()
            side.wait_stream(torch.cuda.current_stream())
            with torch.cuda.stream(side):
                step64()
            torch.cuda.current_stream().wait_stream(side)
            capture_graph(g64, step64)
            g64.replay()
            torch.cuda.synchronize()
            run64 = g64.replay
        t0 = time.perf_counter()
        for _ in range(k64):
            run64()
        torch.cuda.synchronize()
        ms64 = (time.perf_counter() - t0) / k64 * 1e3
        rel = lambda a, b: float(torch.linalg.norm(a - b) / torch.linalg.norm(b))  # noqa: E731
        vs64 = {"fp64_ms_per_step": round(ms64, 4), "fp64_value": round(N / (ms64 / 1e3), 1),
                "speedup": round(ms64 / ms, 3), "abs_dE_Ha": abs(E - float(e64)),
                "rel_err_dE_dD": rel(dD, dD64), "rel_err_dE_dtheta": rel(dth, dth64)}
    # streaming (HBM-bound) kernels of the step: algorithmic bytes per launch over the
    # measured launch time, against the 8 TB/s HBM peak
    Np = int(mol.ao_planes.shape[1])
    es = 8 if args.precision == "fp64" else 4
    h, nd, nc = spec.h, spec.nd, spec.nc
    sbytes = {"layer0": Np * (16 + h) * es,
              "head_fwd": Np * (h + nd) * es + Np * (4 + 1) * 8,
              "head_delta": Np * (2 * h + 16) * es,
              "layer0_bwd": Np * (h + nc) * es}
    streaming = {}
    for k, b in sbytes.items():
        if k in kt.result:
            tot, cnt = kt.result[k]
            gbs = b * cnt / (tot / 1e3) / 1e9
            streaming[k] = {"bytes_per_launch": b, "avg_launch_us": round(tot / cnt * 1e3, 2),
                            "GB/s": round(gbs, 1), "frac_of_8TB/s": round(gbs / 8000.0, 3)}
    nsv = 2
    w_pt = 8 * n * n * nsd + 8 * n * n * nsv + 3 * spec.flops_per_point()
    step_tflops = w_pt * N / (ms / 1e3) / 1e12

    def section(run, skip):
        # each section's molecules hold device memory until the cyclic collector frees them
        # (molecules and their XC plans reference each other): collect between sections
        if skip:
            return None
        gc.collect()
        torch.cuda.empty_cache()
        free, total = torch.cuda.mem_get_info()
        print(f"[bench] {run.__name__}: device memory free {free / 2**30:.1f} of {total / 2**30:.1f} GiB, "
              f"torch allocated {torch.cuda.memory_allocated() / 2**30:.1f} GiB", file=sys.stderr, flush=True)
        return run(args)

    scf = section(scf_bench, args.no_scf)
    train = section(train_bench, args.no_train)
    batch = section(batch_bench, args.no_batch or args.no_scf)
    train_batch = section(train_batch_bench, args.no_train_batch or args.no_train)
    eri = section(eri_bench, args.no_eri)
    config4 = section(config4_bench, args.no_config4 or args.no_scf)
    config5 = section(config5_bench, args.no_config5)
    chi = section(chi_bench, args.no_chi)

    cpu = None
    if not args.no_cpu_baseline:
        v, cores, sample = cpu_baseline(mol, theta, spec, min(args.cpu_sample, N))
        cpu = {"value": round(v, 1), "unit": "grid_pts/s", "cores": cores, "kind": "port",
               "sample": f"FP64 torch-CPU oracle, {sample}"}

    out = {
        "metric": METRIC,
        "value": round(value, 1),
        "unit": "grid_pts/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms, 4),
        "ms_per_step_eager": round(ms_eager, 4),
        "timing": ("value/ms_per_step: HIP-graph replays of the step; roofline and kernels_ms_per_step: "
                   "the eager timed region with per-launch HIP events") if not args.no_graph else "eager",
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": DTYPE[args.precision],
        "data": "synthetic",
        "config": {
            "workload": f"benzene 6-31G* (cartesian d) nao={n}, grid level {args.grid_level} "
                        f"({N} pts), neural B3LYP MLP 3x256 ELU: E_xc + dE/dtheta + dE/dD",
            "npts": N, "nao": n, "precision": args.precision, "seed": args.seed,
            "exc": E,
        },
        "roofline": roof,
        "streaming": streaming,
        "step_mfma": {"algorithmic_tflops": round(step_tflops, 3), "peak": peak,
                      "frac": round(step_tflops / peak, 4), "flop_per_point": w_pt},
        "kernels_ms_per_step": {k: round(v[0] / args.steps, 4) for k, v in
                                sorted(kt.result.items(), key=lambda kv: -kv[1][0])},
        "vs_fp64_gpu": vs64,
        "gate_1e-6_Ha": None if vs64 is None else vs64["abs_dE_Ha"] <= 1e-6,
        "scf": scf,
        "train_step": train,
        "batch": batch,
        "train_batch": train_batch,
        "eri": eri,
        "config4": config4,
        "config5": config5,
        "hf_chi": chi,
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
