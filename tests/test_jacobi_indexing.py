"""CPU check of the Jacobi rotation pass's work distribution (graddft_amd/csrc/gdx_scf.hip,
jacobi_sweeps): 1024 threads step through the 2x2 blocks (k, l) of a round and then the
rows (l, i) of the eigenvector update by adding the thread count to carried indices instead
of dividing per item.  For every matrix size the kernels accept, each item must be visited
exactly once and every index must stay in bounds."""
import pytest

NTH = 1024


def _visits(n):
    m = n + (n & 1)
    npair = m // 2
    nA, nV = npair * npair, n * npair
    blocks, rows = [], []
    for tid in range(NTH):
        e = tid
        k, l = divmod(e, npair)
        dk, dl = divmod(NTH, npair)
        while e < nA:
            assert 0 <= k < npair and 0 <= l < npair
            blocks.append((k, l))
            l += dl
            k += dk
            if l >= npair:
                l -= npair
                k += 1
            e += NTH
        ev = e - nA
        l, i = divmod(ev, n)
        dlv, di = divmod(NTH, n)
        while ev < nV:
            assert 0 <= l < npair and 0 <= i < n
            rows.append((l, i))
            i += di
            l += dlv
            if i >= n:
                i -= n
                l += 1
            ev += NTH
    return npair, blocks, rows


@pytest.mark.parametrize("n", [1, 2, 3, 43, 64, 65, 106, 154, 511, 512, 1021, 1022])
def test_every_item_once_and_in_bounds(n):
    npair, blocks, rows = _visits(n)
    assert sorted(blocks) == [(k, l) for k in range(npair) for l in range(npair)]
    assert sorted(rows) == [(l, i) for l in range(npair) for i in range(n)]
