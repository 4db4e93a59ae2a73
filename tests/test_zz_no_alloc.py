"""ebem_analitic_intg_noalloc (the batch of region routines staged in the
library's fixed per-GPU array, no device allocation) against
ebem_analitic_intg (same kernel, library-allocated staging), through the C
ABI: two calls in a row, the second smaller, bit for bit; over-capacity
batches and bad kinds rejected before any device work."""
import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _calls():
    from oracle import Oracle
    from elastodynamicebem_amd.mesh import screen_uniform
    from test_oracle import _random_cases
    m = screen_uniform(1)
    o = Oracle(m.coordinates, m.triangles, Tfin=1.0, nT=16)
    parts = [[] for _ in range(4)]
    for c, T, t in _random_cases(7, 120):
        n = np.cross(T[1] - T[0], T[2] - T[0])
        n /= np.linalg.norm(n)
        for tt in (t, 0.5 * t, 1.7 * t):
            kind, args, info, B, _ = o.region_trace(c, T, n, tt)
            for p, v in zip(parts, (kind, args, info, B)):
                p.append(v)
    kind, args, info, B = (np.concatenate(p) for p in parts)
    Bc = np.ascontiguousarray(B.transpose(0, 1, 3, 2))  # column-major 3x3 per matrix
    return (np.ascontiguousarray(kind, np.int32), np.ascontiguousarray(args), np.ascontiguousarray(info), Bc)


def _call(lib, h, fn, kind, args, info, B):
    from elastodynamicebem_amd import _lib
    n = kind.size
    out = np.full((n, 36), np.nan)
    rc = getattr(lib, fn)(h, n, kind.ctypes.data_as(_lib._ip), _lib.dptr(args), _lib.dptr(info), _lib.dptr(B),
                          _lib.dptr(out))
    return rc, out


def test_noalloc_matches_allocating_entry_point():
    from elastodynamicebem_amd import _lib
    lib = _lib.load()
    h = C.c_void_p()
    _lib.check(lib.ebem_create(C.byref(h), 0))
    try:
        kind, args, info, B = _calls()
        n1 = kind.size
        assert 1000 < n1 <= 8192
        n2 = n1 // 3
        for n in (n1, n2):  # second call smaller: stale staging from the first must not leak in
            sl = slice(0, n)
            rc, ref = _call(lib, h, "ebem_analitic_intg", kind[sl], args[sl], info[sl], B[sl])
            assert rc == 0
            rc, got = _call(lib, h, "ebem_analitic_intg_noalloc", kind[sl], args[sl], info[sl], B[sl])
            assert rc == 0, lib.ebem_last_error()
            assert np.array_equal(got, ref, equal_nan=False), n
        # over capacity and bad kinds: rejected, output untouched
        big = 8193
        kb = np.zeros(big, np.int32)
        z8, z12, z54 = np.zeros((big, 8)), np.zeros((big, 12)), np.zeros((big, 54))
        rc, out = _call(lib, h, "ebem_analitic_intg_noalloc", kb, z8, z12, z54)
        assert rc == -1 and np.isnan(out).all()
        assert b"at most 8192" in lib.ebem_last_error()
        bad = kind[:4].copy()
        bad[2] = 9
        rc, out = _call(lib, h, "ebem_analitic_intg_noalloc", bad, args[:4], info[:4], B[:4])
        assert rc == -1 and np.isnan(out).all()
    finally:
        lib.ebem_destroy(h)
