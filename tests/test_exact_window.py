"""The assembly stores only the lags whose block is not zero in exact
arithmetic; every block the reference's cull admits but the library drops
must be exactly zero in the oracle (after the 1e-14 cleanup of
calcMatrixBlock.m:77), and the dropped share is substantial."""
import numpy as np

from hostasm import exact_windows, windows


def _check(mesh, Tfin, nT, rows):
    from oracle import Oracle
    o = Oracle(mesh.coordinates, mesh.triangles, Tfin=Tfin, nT=nT)
    L = o.num_blocks()
    lo, hi = windows(o, L)
    lo2, hi2 = exact_windows(o, L)
    kept = np.where(lo2 <= hi2, hi2 - lo2 + 1, 0).sum()
    total = np.where(lo <= hi, hi - lo + 1, 0).sum()
    dropped = 0
    for i in rows:
        for j in range(o.N):
            if lo[i, j] > hi[i, j]:
                continue
            for l in range(lo[i, j], hi[i, j] + 1):
                if lo2[i, j] <= l <= hi2[i, j]:
                    continue
                B, ok = o.pair_block(int(l), int(i), int(j))
                assert ok and not np.any(B), (i, j, l, B)
                dropped += 1
    return kept / total, dropped


def test_dropped_lags_are_zero_screen(screen2):
    frac, dropped = _check(screen2, 1.0, 16, range(0, 128, 3))
    assert dropped > 1000 and frac < 0.9, (frac, dropped)


def test_dropped_lags_are_zero_sphere(sphere2):
    frac, dropped = _check(sphere2, 4.0, 32, range(0, 320, 16))
    assert dropped > 1000 and frac < 0.9, (frac, dropped)
