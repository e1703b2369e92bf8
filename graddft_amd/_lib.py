"""ctypes binding of ``libgdx.so`` (the C-ABI declared in ``include/gdx.h``).

Loading fails loudly: there is no CPU fallback for any entry point.  Device pointers are
passed as integers (``tensor.data_ptr()``) and the HIP stream as the raw handle of
torch's current stream, so torch only provides memory and streams.
"""
from __future__ import annotations

import ctypes
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libgdx.so")

GDX_FP64, GDX_FP32, GDX_BF16X6 = 0, 1, 2
GDX_FUNC_B3LYP, GDX_FUNC_NEURAL_B3LYP, GDX_FUNC_B3LYP_NEURAL_CORR, GDX_FUNC_LINEAR = 0, 1, 2, 3
GDX_CORR_VWN5, GDX_CORR_PW92 = 0, 1
GDX_STATUS_EIGH_NOT_CONVERGED, GDX_STATUS_NONFINITE_ENERGY, GDX_STATUS_CHARGE = 1, 2, 4


class XCDesc(ctypes.Structure):
    _fields_ = [
        ("npts", ctypes.c_int32),
        ("npts_pad", ctypes.c_int32),
        ("nao", ctypes.c_int32),
        ("nao_pad", ctypes.c_int32),
        ("spin_restricted", ctypes.c_int32),
        ("kind", ctypes.c_int32),
        ("precision", ctypes.c_int32),
        ("mlp_nc", ctypes.c_int32),
        ("mlp_h", ctypes.c_int32),
        ("mlp_nhidden", ctypes.c_int32),
        ("mlp_nd", ctypes.c_int32),
        ("correlation", ctypes.c_int32),
        ("coef", ctypes.c_double * 4),
    ]


# name -> (restype, argtypes); every pointer is c_void_p
_P, _I, _Z = ctypes.c_void_p, ctypes.c_int32, ctypes.c_size_t
SIGNATURES = {
    "gdx_last_error": (ctypes.c_char_p, []),
    "gdx_version": (ctypes.c_int, []),
    "gdx_ao_eval": (ctypes.c_int, [_P, _I, _P, _P, _P, _P, _I, _I, _P, _I, _I, _P]),
    "gdx_xc_autotune": (ctypes.c_int, [_P, ctypes.POINTER(XCDesc), _P, _P, _P]),
    "gdx_xc_workspace_size": (ctypes.c_int, [ctypes.POINTER(XCDesc), ctypes.POINTER(_Z)]),
    "gdx_xc_forward": (ctypes.c_int, [_P, ctypes.POINTER(XCDesc), _P, _P, _P, _P, _P, _P]),
    "gdx_coulomb_workspace_size": (ctypes.c_int, [_I, ctypes.POINTER(_Z)]),
    "gdx_coulomb_packed": (ctypes.c_int, [_P, _I, _P, _P, _P, _P]),
    "gdx_coulomb_noalloc_max_n": (ctypes.c_int, [_P]),
    "gdx_coulomb_packed_noalloc": (ctypes.c_int, [_P, _I, _P, _P, _P]),
    "gdx_nonxc_energy": (ctypes.c_int, [_P, _I, _P, _P, _P, _P, ctypes.c_double, ctypes.c_double, _P]),
    "gdx_fock_assemble": (ctypes.c_int, [_P, _I, _P, _P, _P, _P, ctypes.c_double, _P]),
    "gdx_exchange_packed": (ctypes.c_int, [_P, _I, _I, _P, _P, _P]),
    "gdx_exchange_xo_size": (ctypes.c_int, [_I, ctypes.POINTER(ctypes.c_size_t)]),
    "gdx_exchange_xo_build": (ctypes.c_int, [_P, _I, _P, _P]),
    "gdx_exchange_xo_workspace_size": (ctypes.c_int, [_I, ctypes.POINTER(ctypes.c_size_t)]),
    "gdx_exchange_xo": (ctypes.c_int, [_P, _I, _I, _P, _P, _P, _P]),
    "gdx_cholesky_inv": (ctypes.c_int, [_P, _I, _P, _P, _P, _P]),
    "gdx_eigh_workspace_size": (ctypes.c_int, [_I, _I, ctypes.POINTER(_Z)]),
    "gdx_eigh_batched": (ctypes.c_int, [_P, _I, _I, _P, _P, _P, _P, _P, _I, ctypes.c_double, _P]),
    "gdx_diis_solve": (ctypes.c_int, [_P, _I, _I, _P, _P]),
    "gdx_scf_status": (ctypes.c_int, [_P, _I, _I, _P, _P, _P, _P, _I, _P, ctypes.c_double, _P]),
    "gdx_diis_coefficients": (ctypes.c_int, [_P, _I, _I, ctypes.c_int64, _P, _I, _P, _P]),
    "gdx_diis_error_workspace_size": (ctypes.c_int, [_I, _I, ctypes.POINTER(_Z)]),
    "gdx_diis_error": (ctypes.c_int, [_P, _I, _I, _P, _P, _P, _P, _P]),
    "gdx_scf_step_workspace_size": (ctypes.c_int, [_I, _I, ctypes.POINTER(_Z)]),
    "gdx_scf_step": (ctypes.c_int, [_P, _I, _I, _I, _P, _P, _P, _P, _P, _P, _P, _P, _P, _P, _P, _I, ctypes.c_double,
                                    _P]),
    "gdx_eri_plan_create": (ctypes.c_int, [_I, _P, _P, _P, _P, _I, ctypes.c_double, ctypes.POINTER(_P)]),
    "gdx_eri_plan_destroy": (ctypes.c_int, [_P]),
    "gdx_eri_plan_info": (ctypes.c_int, [_P, ctypes.POINTER(ctypes.c_int64), ctypes.POINTER(ctypes.c_int64)]),
    "gdx_eri_packed": (ctypes.c_int, [_P, _P, _P]),
    "gdx_int1e_plan_create": (ctypes.c_int, [_I, _P, _P, _P, _P, _I, _I, _I, _P, _P, ctypes.POINTER(_P)]),
    "gdx_int1e_plan_destroy": (ctypes.c_int, [_P]),
    "gdx_int1e_eval": (ctypes.c_int, [_P, _P, _P, _P, _P]),
    "gdx_becke_weights": (ctypes.c_int, [_P, _I, _P, _P, _P, _I, _P, _P, _P, _P]),
    "gdx_hf_plan_create": (ctypes.c_int, [_I, _P, _P, _P, _P, _I, ctypes.c_double, ctypes.POINTER(_P)]),
    "gdx_hf_plan_destroy": (ctypes.c_int, [_P]),
    "gdx_hf_plan_info": (ctypes.c_int, [_P, ctypes.POINTER(ctypes.c_int64), ctypes.POINTER(ctypes.c_int64)]),
    "gdx_hf_chi": (ctypes.c_int, [_P, _P, _I, _I, _P, _I, _P, ctypes.c_double, _P]),
    "gdx_ao_matmul": (ctypes.c_int, [_P, _I, _I, _I, _I, _P, _I, _P, _P]),
    "gdx_hf_energy_density": (ctypes.c_int, [_P, _I, _I, _I, _I, _P, _P, _P]),
    "gdx_hf_fock_workspace_size": (ctypes.c_int, [_I, _I, _I, ctypes.POINTER(_Z)]),
    "gdx_hf_fock": (ctypes.c_int, [_P, _I, _I, _I, _I, _P, _I, _P, _P, ctypes.c_int64, ctypes.c_double, _P, _P]),
    "gdx_profile_begin": (ctypes.c_int, []),
    "gdx_profile_end": (ctypes.c_int, [_I, ctypes.c_char_p, _P, _P, _P]),
    "gdx_xc_custom_inputs": (ctypes.c_int, [_P, ctypes.POINTER(XCDesc), _P, _P, _P, _P, _P, _P]),
    "gdx_xc_custom_vxc": (ctypes.c_int, [_P, ctypes.POINTER(XCDesc), _P, _P, _P, _P, _P]),
    "gdx_xc_densities": (ctypes.c_int, [_P, ctypes.POINTER(XCDesc), _P, _P, _P, _P]),
    "gdx_xc_backward": (ctypes.c_int, [_P, ctypes.POINTER(XCDesc), _P, _P, _P, _P, _P, _P, _P, _P]),
    "gdx_xc_hvp_workspace_size": (ctypes.c_int, [ctypes.POINTER(XCDesc), ctypes.POINTER(_Z)]),
    "gdx_xc_hvp": (ctypes.c_int, [_P, ctypes.POINTER(XCDesc), _P, _P, _P, _P, _P, _P, _P, _P, _P, _P, _P]),
}

class XCBatchMol(ctypes.Structure):
    _fields_ = [
        ("ao", ctypes.c_void_p),
        ("weights", ctypes.c_void_p),
        ("npts", ctypes.c_int32),
        ("npts_pad", ctypes.c_int32),
        ("nao", ctypes.c_int32),
        ("nao_pad", ctypes.c_int32),
    ]


SIGNATURES.update({
    "gdx_xc_batch_create": (ctypes.c_int, [ctypes.POINTER(XCDesc), _I, ctypes.POINTER(XCBatchMol), _I,
                                           ctypes.POINTER(_P)]),
    "gdx_xc_batch_destroy": (ctypes.c_int, [_P]),
    "gdx_xc_batch_offsets": (ctypes.c_int, [_P, ctypes.POINTER(ctypes.c_int64)]),
    "gdx_xc_batch_forward": (ctypes.c_int, [_P, _P, _P, _P, _P]),
    "gdx_xc_batch_backward": (ctypes.c_int, [_P, _P, _P, _P, _P, _P]),
    "gdx_xc_batch_hvp": (ctypes.c_int, [_P, _P, _P, _P, _P, _P, _P, _P, _P, _P]),
})

_lib = None


def load(path: str = LIB_PATH) -> ctypes.CDLL:
    """Load the HIP library (once).  Raises if it was not built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(path):
        raise RuntimeError(
            f"gdx HIP library not found at {path}; build it with "
            "`python -m graddft_amd.build_ext` (there is no CPU fallback)")
    lib = ctypes.CDLL(path)
    for name, (res, args) in SIGNATURES.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib


class GdxError(RuntimeError):
    pass


def check(rc: int, what: str) -> None:
    if rc != 0:
        msg = load().gdx_last_error()
        raise GdxError(f"{what} failed with status {rc}: {msg.decode() if msg else ''}")


class KernelTimer:
    """Context manager: per-kernel HIP-event timing of everything gdx enqueues inside it.
    ``result`` maps kernel name -> (total_ms, launches) after exit."""

    def __init__(self, max_entries: int = 64):
        self.max_entries = max_entries
        self.result = {}

    def __enter__(self):
        check(load().gdx_profile_begin(), "gdx_profile_begin")
        return self

    def __exit__(self, *exc):
        import numpy as np
        n = self.max_entries
        names = ctypes.create_string_buffer(64 * n)
        tot = np.zeros(n, dtype=np.float64)
        cnt = np.zeros(n, dtype=np.int32)
        nout = np.zeros(1, dtype=np.int32)
        check(load().gdx_profile_end(n, names, tot.ctypes.data, cnt.ctypes.data, nout.ctypes.data),
              "gdx_profile_end")
        raw = names.raw
        for i in range(int(nout[0])):
            nm = raw[64 * i: 64 * (i + 1)].split(b"\0", 1)[0].decode()
            self.result[nm] = (float(tot[i]), int(cnt[i]))
        return False
