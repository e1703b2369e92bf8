"""gdx_coulomb_packed_noalloc (Coulomb matrix from the packed ERI with the library's fixed
device scratch in place of a workspace) against gdx_coulomb_packed, through the C-ABI: two
calls in a row, the second with a smaller basis, each bitwise equal to the workspace entry
point; sizes past the scratch are rejected before any launch."""
import ctypes

import numpy as np
import pytest
import torch


def _inputs(n, seed):
    rng = np.random.default_rng(seed)
    P = n * (n + 1) // 2
    eri = rng.normal(size=P * (P + 1) // 2)
    D = rng.normal(size=(2, n, n))
    D = D + D.transpose(0, 2, 1)
    return eri, D


def _coulomb(lib, n, eri, D, noalloc):
    from graddft_amd import _lib
    dev = "cuda:0"
    e = torch.as_tensor(eri, device=dev)
    d = torch.as_tensor(D, device=dev)
    J = torch.full((n, n), float("nan"), dtype=torch.float64, device=dev)
    st = torch.cuda.current_stream().cuda_stream
    if noalloc:
        _lib.check(lib.gdx_coulomb_packed_noalloc(st, n, e.data_ptr(), d.data_ptr(), J.data_ptr()),
                   "gdx_coulomb_packed_noalloc")
    else:
        nb = ctypes.c_size_t(0)
        _lib.check(lib.gdx_coulomb_workspace_size(n, ctypes.byref(nb)), "gdx_coulomb_workspace_size")
        ws = torch.empty(nb.value // 8, dtype=torch.float64, device=dev)
        _lib.check(lib.gdx_coulomb_packed(st, n, e.data_ptr(), d.data_ptr(), J.data_ptr(), ws.data_ptr()),
                   "gdx_coulomb_packed")
    torch.cuda.synchronize()
    return J.cpu().numpy()


def test_noalloc_rejects_sizes_past_the_scratch():
    from graddft_amd import _lib
    lib = _lib.load()
    nmax = ctypes.c_int32(0)
    assert lib.gdx_coulomb_noalloc_max_n(ctypes.byref(nmax)) == 0
    assert nmax.value == 214
    # rejected before any device call: runs without a GPU
    assert lib.gdx_coulomb_packed_noalloc(None, nmax.value + 1, 8, 8, 8) == -1
    assert b"scratch" in lib.gdx_last_error()


@pytest.mark.gpu
def test_noalloc_coulomb_matches_workspace_entry_point():
    from graddft_amd import _lib
    lib = _lib.load()
    for n, seed in ((180, 1), (43, 2)):       # second call smaller: stale scratch must not matter
        print(f"case: coulomb noalloc vs workspace, n={n}", flush=True)
        eri, D = _inputs(n, seed)
        ref = _coulomb(lib, n, eri, D, noalloc=False)
        got = _coulomb(lib, n, eri, D, noalloc=True)
        assert np.isfinite(ref).all()
        assert np.array_equal(got, ref), np.abs(got - ref).max()
        if n <= 43:                            # and the plain definition J_ij = sum_kl (ij|kl) D_kl
            from graddft_amd.interface.integrals import unpack_eri
            dense = unpack_eri(eri, n)
            Jd = np.einsum("ijkl,kl->ij", dense, D.sum(0))
            assert np.abs(got - Jd).max() <= 1e-10 * np.abs(Jd).max()
