"""ctypes wrapper over oracle/libebem_oracle.so (test infrastructure only)."""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = os.path.join(_HERE, "libebem_oracle.so")
_dp = C.POINTER(C.c_double)
_ip = C.POINTER(C.c_int)


class _Param(C.Structure):
    _fields_ = [("rho", C.c_double), ("mu", C.c_double), ("lam", C.c_double),
                ("velP", C.c_double), ("velS", C.c_double), ("Tfin", C.c_double),
                ("nT", C.c_int)]


class _Mesh(C.Structure):
    _fields_ = [("N", C.c_int), ("Nv", C.c_int), ("coords", _dp), ("tri", _ip),
                ("normal", _dp), ("area", _dp), ("center", _dp), ("maxL", _dp)]


def build() -> None:
    """Compile the oracle with its Makefile (gcc, OpenMP)."""
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


def _p(a, t=_dp):
    return a.ctypes.data_as(t)


class Oracle:
    """Problem-bound oracle: material, time grid, mesh and outer GH rule."""

    def __init__(self, coords, tri, rho=1.0, mu=1.0, lam=1.0, Tfin=1.0, nT=16, nGH=7):
        if not os.path.exists(_LIB):
            build()
        self.lib = C.CDLL(_LIB)
        self._sig()
        self.coords = np.ascontiguousarray(coords, np.float64)
        self.tri = np.ascontiguousarray(tri, np.int32)
        N = self.tri.shape[0]
        self.N = N
        self.normal = np.zeros((N, 3))
        self.area = np.zeros(N)
        self.center = np.zeros((N, 3))
        self.maxL = np.zeros(N)
        self.lib.orc_mesh_derive(N, _p(self.coords), _p(self.tri, _ip), _p(self.normal), _p(self.area),
                                 _p(self.center), _p(self.maxL))
        velP = np.sqrt((lam + 2 * mu) / rho)
        velS = np.sqrt(mu / rho)
        self.param = _Param(rho, mu, lam, velP, velS, Tfin, nT)
        self.mesh = _Mesh(N, self.coords.shape[0], _p(self.coords), _p(self.tri, _ip), _p(self.normal),
                          _p(self.area), _p(self.center), _p(self.maxL))
        self.nGH = nGH
        self.ghn = np.zeros((nGH, 3))
        self.ghw = np.zeros(nGH)
        if self.lib.orc_gauss_hammer(nGH, _p(self.ghn), _p(self.ghw)) != 0:
            raise ValueError("unsupported Gauss-Hammer rule")
        self.velP, self.velS, self.rho, self.nT, self.Tfin = velP, velS, rho, nT, Tfin
        self.dt = Tfin / nT

    def _sig(self):
        L = self.lib
        L.orc_mesh_derive.argtypes = [C.c_int, _dp, _ip, _dp, _dp, _dp, _dp]
        L.orc_gauss_hammer.argtypes = [C.c_int, _dp, _dp]
        L.orc_gauss_hammer.restype = C.c_int
        L.orc_num_blocks.argtypes = [C.POINTER(_Param), C.c_int, _dp]
        L.orc_num_blocks.restype = C.c_int
        L.orc_node_ris.argtypes = [C.POINTER(_Param), _dp, _dp, _dp, C.c_double, _dp]
        L.orc_node_ris.restype = C.c_int
        L.orc_intg.argtypes = [C.c_int, _dp, _dp, _dp, _dp]
        L.orc_intg.restype = C.c_int
        L.orc_region_trace.argtypes = [C.POINTER(_Param), _dp, _dp, _dp, C.c_double, C.c_int, _ip, _dp, _dp, _dp,
                                       _dp]
        L.orc_region_trace.restype = C.c_int
        L.orc_pair_block.argtypes = [C.POINTER(_Param), C.POINTER(_Mesh), C.c_int, _dp, _dp, C.c_int,
                                     C.c_int, C.c_int, _dp]
        L.orc_pair_block.restype = C.c_int
        L.orc_matrix_block_rows.argtypes = [C.POINTER(_Param), C.POINTER(_Mesh), C.c_int, _dp, _dp,
                                            C.c_int, C.c_int, C.c_int, _dp, C.c_int]
        L.orc_matrix_block_rows.restype = C.c_int
        L.orc_rhs.argtypes = [C.POINTER(_Param), C.POINTER(_Mesh), C.c_int, _dp, _dp, C.c_int, _dp]
        L.orc_datum.argtypes = [C.c_int, _dp, _dp, C.c_double, _dp]
        L.orc_lu.argtypes = [C.c_int, _dp, _ip]
        L.orc_lu.restype = C.c_int
        L.orc_lu_solve.argtypes = [C.c_int, _dp, _ip, _dp]
        L.orc_mot.argtypes = [C.c_int, C.c_int, C.c_int, _dp, _dp, _dp]
        L.orc_mot.restype = C.c_int
        L.orc_pp_block.argtypes = [C.POINTER(_Param), _dp, _dp, _dp, C.c_double, _dp]
        L.orc_pp_point.argtypes = [C.POINTER(_Param), C.POINTER(_Mesh), _dp, _dp, C.c_double, _dp]
        L.orc_counters.argtypes = [C.POINTER(C.c_longlong)]
        L.orc_field_record.argtypes = [_dp, _dp, _dp]
        L.orc_pp_block_num.argtypes = [C.POINTER(_Param), C.POINTER(_Mesh), C.c_int, _dp, C.c_double, C.c_int,
                                       _dp, _dp, _dp]
        L.orc_pp_point_num.argtypes = [C.POINTER(_Param), C.POINTER(_Mesh), _dp, _dp, C.c_double, C.c_int, _dp,
                                       _dp, _dp]
        L.orc_gh_composite.argtypes = [C.c_int, C.c_int, _dp, _dp]
        L.orc_gh_composite.restype = C.c_int
        L.orc_doppio_gauss.argtypes = [C.c_int, _dp, _dp]
        L.orc_doppio_gauss.restype = C.c_int
        L.orc_nucleo.argtypes = [_dp, C.c_double, C.c_double, C.c_double, C.c_double, _dp]
        L.orc_ghc_pair_block.argtypes = [C.POINTER(_Param), C.POINTER(_Mesh), C.c_int, _dp, _dp, C.c_int, _dp, _dp,
                                         C.c_int, _dp, _dp, C.c_int, C.c_int, C.c_int, _dp]
        L.orc_ghc_pair_block.restype = C.c_int
        L.orc_ghc_matrix_block_rows.argtypes = [C.POINTER(_Param), C.POINTER(_Mesh), C.c_int, _dp, _dp, C.c_int,
                                                _dp, _dp, C.c_int, _dp, _dp, C.c_int, C.c_int, C.c_int, _dp,
                                                C.c_int]
        L.orc_k_basis.argtypes = [_dp, _dp, _dp]
        L.orc_ppP_block.argtypes = [C.POINTER(_Param), C.POINTER(_Mesh), C.c_int, _dp, _dp, _dp, C.c_double,
                                    C.c_int, C.c_int, _dp]
        L.orc_ppG_block.argtypes = [C.POINTER(_Param), C.POINTER(_Mesh), C.c_int, _dp, _dp, _dp, C.c_double,
                                    C.c_int, C.c_int, _dp]
        L.orc_ppD_point.argtypes = [C.POINTER(_Param), C.POINTER(_Mesh), _dp, _dp, _dp, C.c_double, C.c_int, _dp,
                                    _dp, _dp, C.c_int]
        L.orc_k_block.argtypes = [C.POINTER(_Param), C.POINTER(_Mesh), C.c_int, _dp, _dp, C.c_int, _dp, _dp,
                                  C.c_int, _dp, _dp, C.c_int, C.c_int, C.c_int, C.c_int, _dp]
        L.orc_k_block.restype = C.c_int
        L.orc_k_matrix_rows.argtypes = [C.POINTER(_Param), C.POINTER(_Mesh), C.c_int, _dp, _dp, C.c_int, _dp, _dp,
                                        C.c_int, _dp, _dp, C.c_int, C.c_int, C.c_int, _dp, C.c_int]

    # -- pieces -------------------------------------------------------------
    def num_blocks(self) -> int:
        return int(self.lib.orc_num_blocks(C.byref(self.param), self.coords.shape[0], _p(self.coords)))

    def node_ris(self, x, T, n, t):
        """ris1..ris6 (6 x 3 x 3) for outer point x, field triangle T (3x3 rows=vertices), time t."""
        out = np.zeros(54)
        x = np.ascontiguousarray(x, np.float64)
        T = np.ascontiguousarray(T, np.float64)
        n = np.ascontiguousarray(n, np.float64)
        cls = self.lib.orc_node_ris(C.byref(self.param), _p(x), _p(T), _p(n), float(t), _p(out))
        return out.reshape(6, 3, 3).transpose(0, 2, 1).copy(), cls

    def intg(self, kind, args, info, B):
        """One region routine of functions/BEMenerg_analiticIntg/ (kind 0..5 = CC,
        SC_P, SC_SP, TC, TR_P, TR_SP): args (8,), info (12,), B (6, 3, 3) ->
        ris (4, 3, 3) (ris1..ris4, or ris5, ris6 and two zero matrices)."""
        a = np.ascontiguousarray(args, np.float64)
        f = np.ascontiguousarray(info, np.float64)
        b = np.ascontiguousarray(np.asarray(B, np.float64).transpose(0, 2, 1))
        out = np.zeros(36)
        if self.lib.orc_intg(int(kind), _p(a), _p(f), _p(b), _p(out)) != 0:
            raise ValueError("unknown region kind")
        return out.reshape(4, 3, 3).transpose(0, 2, 1).copy()

    def region_trace(self, x, T, n, t, cap=256):
        """Every region call of the node integral of point x over triangle T at
        time t: kinds (k,), args (k, 8), info (k, 12), B (k, 6, 3, 3), ris (k, 4, 3, 3)."""
        x = np.ascontiguousarray(x, np.float64)
        T = np.ascontiguousarray(T, np.float64)
        n = np.ascontiguousarray(n, np.float64)
        kind = np.zeros(cap, np.int32)
        args, info = np.zeros((cap, 8)), np.zeros((cap, 12))
        B, ris = np.zeros((cap, 54)), np.zeros((cap, 36))
        k = self.lib.orc_region_trace(C.byref(self.param), _p(x), _p(T), _p(n), float(t), cap,
                                      kind.ctypes.data_as(_ip), _p(args), _p(info), _p(B), _p(ris))
        k = min(k, cap)
        return (kind[:k].copy(), args[:k].copy(), info[:k].copy(),
                B[:k].reshape(k, 6, 3, 3).transpose(0, 1, 3, 2).copy(),
                ris[:k].reshape(k, 4, 3, 3).transpose(0, 1, 3, 2).copy())

    def pair_block(self, lag, S, F):
        out = np.zeros(9)
        ok = self.lib.orc_pair_block(C.byref(self.param), C.byref(self.mesh), self.nGH, _p(self.ghn),
                                     _p(self.ghw), lag, S, F, _p(out))
        return out.reshape(3, 3).T.copy(), ok

    def matrix_block(self, lag, rows=None, nthreads=None):
        """Dense E_lag (3N x 3N) as a numpy array (row index = test DOF)."""
        N = self.N
        out = np.zeros((3 * N, 3 * N), order="F")
        r0, r1 = (0, N) if rows is None else rows
        nt = nthreads or os.cpu_count() or 1
        any_ = self.lib.orc_matrix_block_rows(C.byref(self.param), C.byref(self.mesh), self.nGH, _p(self.ghn),
                                              _p(self.ghw), lag, r0, r1, out.ctypes.data_as(_dp), nt)
        return np.ascontiguousarray(out), bool(any_)

    def rhs(self, datum):
        out = np.zeros((3 * self.N, self.nT), order="F")
        self.lib.orc_rhs(C.byref(self.param), C.byref(self.mesh), self.nGH, _p(self.ghn), _p(self.ghw),
                         int(datum), out.ctypes.data_as(_dp))
        return np.ascontiguousarray(out)

    def datum(self, kind, x, nrm, t):
        g = np.zeros(3)
        self.lib.orc_datum(int(kind), _p(np.ascontiguousarray(x, np.float64)),
                           _p(np.ascontiguousarray(nrm, np.float64)), float(t), _p(g))
        return g

    def mot(self, E, beta):
        """E: (L, 3N, 3N) numpy (row = test DOF), beta (3N, nT).  Returns alpha (3N, nT)."""
        L, n3, _ = E.shape
        Ef = np.ascontiguousarray(E.transpose(0, 2, 1))  # column-major per lag
        b = np.asfortranarray(beta)
        a = np.zeros_like(b, order="F")
        rc = self.lib.orc_mot(n3, L, b.shape[1], _p(Ef), b.ctypes.data_as(_dp), a.ctypes.data_as(_dp))
        if rc != 0:
            raise np.linalg.LinAlgError("singular E0")
        return np.ascontiguousarray(a)

    def pp_block(self, x, j, tau):
        T = np.ascontiguousarray(self.coords[self.tri[j]])
        out = np.zeros(9)
        self.lib.orc_pp_block(C.byref(self.param), _p(np.ascontiguousarray(x, np.float64)), _p(T),
                              _p(np.ascontiguousarray(self.normal[j])), float(tau), _p(out))
        return out.reshape(3, 3).T.copy()

    def pp_point(self, density, x, t):
        d = np.asfortranarray(density)
        u = np.zeros(3)
        self.lib.orc_pp_point(C.byref(self.param), C.byref(self.mesh), d.ctypes.data_as(_dp),
                              _p(np.ascontiguousarray(x, np.float64)), float(t), _p(u))
        return u

    def field_record(self, T, n):
        rec = np.zeros(32)
        self.lib.orc_field_record(_p(np.ascontiguousarray(T, np.float64)), _p(np.ascontiguousarray(n, np.float64)),
                                  _p(rec))
        return rec

    def counters(self):
        out = (C.c_longlong * 12)()
        self.lib.orc_counters(out)
        return list(out)

    def counters_reset(self):
        self.lib.orc_counters_reset()

    # -- numeric inner quadrature mode (kernelV.cu / calcSubBlockDiagV.m) ----
    def gh_composite(self, nSub, nNod):
        n = nSub * nNod
        nodes, w = np.zeros((n, 3)), np.zeros(n)
        if self.lib.orc_gh_composite(nSub, nNod, _p(nodes), _p(w)) != 0:
            raise ValueError("invalid composite rule")
        return nodes, w

    def doppio_gauss(self, n):
        nodes, w = np.zeros((n, 3)), np.zeros(n)
        if self.lib.orc_doppio_gauss(n, _p(nodes), _p(w)) != 0:
            raise ValueError("invalid collapsed rule")
        return nodes, w

    def nucleo(self, x, t):
        x = np.ascontiguousarray(x, np.float64)
        out = np.zeros(9)
        self.lib.orc_nucleo(_p(x), float(np.linalg.norm(x)), float(t), self.velP, self.velS, _p(out))
        return out.reshape(3, 3).T.copy()

    def set_ghc(self, nSub=64, nSing=3, nDiag=256):
        """Inner rules of the GHC mode ("FN nExt nSub nSing nDiag")."""
        self.ghcn, self.ghcw = self.gh_composite(nSub, nSing)
        self.dgn, self.dgw = self.doppio_gauss(nDiag)

    def ghc_pair_block(self, lag, S, F):
        out = np.zeros(9)
        self.lib.orc_ghc_pair_block(C.byref(self.param), C.byref(self.mesh), self.nGH, _p(self.ghn), _p(self.ghw),
                                    self.ghcw.size, _p(self.ghcn), _p(self.ghcw), self.dgw.size, _p(self.dgn),
                                    _p(self.dgw), int(lag), int(S), int(F), _p(out))
        return out.reshape(3, 3).T.copy()

    def ghc_matrix_block(self, lag, rows=None, nthreads=None):
        N = self.N
        out = np.zeros((3 * N, 3 * N), order="F")
        r0, r1 = (0, N) if rows is None else rows
        nt = nthreads or os.cpu_count() or 1
        self.lib.orc_ghc_matrix_block_rows(C.byref(self.param), C.byref(self.mesh), self.nGH, _p(self.ghn),
                                           _p(self.ghw), self.ghcw.size, _p(self.ghcn), _p(self.ghcw),
                                           self.dgw.size, _p(self.dgn), _p(self.dgw), int(lag), r0, r1,
                                           out.ctypes.data_as(_dp), nt)
        return np.ascontiguousarray(out)

    # -- double-layer K of the direct Dirichlet driver ------------------------
    def k_basis(self, T):
        """P1 basis of triangle T (3x3, vertices row-wise): (mat 3x3, vec 3), phi_k(x) = vec[k] + mat[k] . x."""
        mat, vec = np.zeros(9), np.zeros(3)
        self.lib.orc_k_basis(_p(np.ascontiguousarray(T, np.float64)), _p(mat), _p(vec))
        return mat.reshape(3, 3), vec

    def k_num_blocks(self) -> int:
        """numBlocksK = min(numBlocksV + 1, nT) (timeMarchingD.m:176)."""
        return min(self.num_blocks() + 1, self.nT)

    def k_block(self, lag, S, s, prune=True):
        """Block (test triangle S, vertex s) of K_lag (needs set_ghc); prune=False
        integrates it even where the reference's skip tests drop it."""
        out = np.zeros(9)
        self.lib.orc_k_block(C.byref(self.param), C.byref(self.mesh), self.nGH, _p(self.ghn), _p(self.ghw),
                             self.ghcw.size, _p(self.ghcn), _p(self.ghcw), self.dgw.size, _p(self.dgn), _p(self.dgw),
                             int(lag), int(S), int(s), int(bool(prune)), _p(out))
        return out.reshape(3, 3).T.copy()

    def k_matrix(self, lag, rows=None, nthreads=None):
        """Dense K_lag (3N x 3Nv), test rows restricted to `rows`."""
        N, Nv = self.N, self.coords.shape[0]
        out = np.zeros((3 * N, 3 * Nv), order="F")
        r0, r1 = (0, N) if rows is None else rows
        nt = nthreads or os.cpu_count() or 1
        self.lib.orc_k_matrix_rows(C.byref(self.param), C.byref(self.mesh), self.nGH, _p(self.ghn), _p(self.ghw),
                                   self.ghcw.size, _p(self.ghcn), _p(self.ghcw), self.dgw.size, _p(self.dgn),
                                   _p(self.dgw), int(lag), r0, r1, out.ctypes.data_as(_dp), nt)
        return np.ascontiguousarray(out)

    def beta_k(self, g, rows=None):
        """beta_K (3N x nT) of dir/BEMenerg_dir_calcBetaKgpu.m:7-13 for the vertex datum
        g (3Nv x nT, column j at time (j+1) dt): beta_n = sum_{j<=n} K_{n-j} g_j."""
        LK = self.k_num_blocks()
        K = [self.k_matrix(l, rows) for l in range(LK)]
        out = np.zeros((3 * self.N, self.nT))
        for n in range(self.nT):
            for j in range(n + 1):
                if n - j < LK:
                    out[:, n] += K[n - j] @ g[:, j]
        return out

    # -- direct post-processing u = G alpha - P g (dirPP/BEMenerg_ppD_calc.m) --
    def pp_p_block(self, x, t, y, s):
        """P_y block of vertex s (kernelP.cu), 3x3 (needs set_pp_rule)."""
        out = np.zeros(9)
        x = np.ascontiguousarray(x, np.float64)
        self.lib.orc_ppP_block(C.byref(self.param), C.byref(self.mesh), self.ppw.size, _p(self.ppn), _p(self.ppw),
                               _p(x), float(t), int(y), int(s), _p(out))
        return out.reshape(3, 3).T.copy()

    def pp_g_block(self, x, t, y, F):
        """G_y block of triangle F (kernelG.cu), 3x3 (needs set_pp_rule)."""
        out = np.zeros(9)
        x = np.ascontiguousarray(x, np.float64)
        self.lib.orc_ppG_block(C.byref(self.param), C.byref(self.mesh), self.ppw.size, _p(self.ppn), _p(self.ppw),
                               _p(x), float(t), int(y), int(F), _p(out))
        return out.reshape(3, 3).T.copy()

    def ppD_point(self, density, g, x, t, nthreads=None):
        """u(x, t) = sum_y G_y alpha_y - P_y g_y (BEMenerg_ppD_calc.m)."""
        d = np.asfortranarray(density, np.float64)
        gg = np.asfortranarray(g, np.float64)
        x = np.ascontiguousarray(x, np.float64)
        u = np.zeros(3)
        self.lib.orc_ppD_point(C.byref(self.param), C.byref(self.mesh), d.ctypes.data_as(_dp), gg.ctypes.data_as(_dp),
                               _p(x), float(t), self.ppw.size, _p(self.ppn), _p(self.ppw), _p(u),
                               nthreads or os.cpu_count() or 1)
        return u

    # -- Neumann driver (dir/BEMenerg_dir_timeMarchingN.m), NumPy over the blocks --
    def neumann_datum(self, h):
        """gV (3N x nT) of dir/BEMenerg_dir_calcBoundDataNeumann.m:5-25 for the bar
        of height h: (0, 0, 2 mu tildeP(c_F, t) n_z) at the triangle centres,
        column n at time (n+1) dt, tildeP = sum_k (-1)^k [H(a_k) + H(b_k)]."""
        cP, T = self.velP, self.Tfin
        khat = int(np.ceil((cP * T) / (2 * h))) - 1
        out = np.zeros((3 * self.N, self.nT))
        for n in range(self.nT):
            t = (n + 1) * self.dt
            for F in range(self.N):
                z = self.center[F, 2]
                tp = 0.0
                for k in range(khat + 1):
                    a = cP * t - 2 * h * k - (h - z)
                    b = cP * t - 2 * h * (k + 1) + (h - z)
                    tp += (-1.0) ** k * (float(a > 0) + float(b > 0))
                out[3 * F + 2, n] = 2 * self.param.mu * tp * self.normal[F, 2]
        return out

    def mass_gamma(self):
        """I_Gamma = kron((indSM > 0) .* area / 6, eye(3)) (timeMarchingN.m:275): 3N x 3Nv."""
        Nv = self.coords.shape[0]
        M = np.zeros((3 * self.N, 3 * Nv))
        for S in range(self.N):
            for v in self.tri[S]:
                for c in range(3):
                    M[3 * S + c, 3 * v + c] = self.area[S] / 6
        return M

    def dn_mot(self, V, K, gV):
        """density (3Nv x nT) of timeMarchingN.m:170-296 from dense V_l (3N x 3N,
        l < len(V)), K_l (3N x 3Nv, l < len(K)) and the Neumann datum gV:
        beta_V as calcBetaVgpu.m, then (K_0 + I_Gamma) d_n = beta_V,n + I_Gamma d_{n-1}
        - sum_{l=1}^{min(n, LK-1)} K_l d_{n-l} by least squares."""
        nT = self.nT
        Nv = self.coords.shape[0]
        betaV = np.zeros((3 * self.N, nT))
        for n in range(nT):
            for j in range(n + 1):
                if n - j < len(V):
                    betaV[:, n] += V[n - j] @ gV[:, j]
        I = self.mass_gamma()
        A = K[0] + I
        d = np.zeros((3 * Nv, nT))
        for n in range(nT):
            r = betaV[:, n].copy()
            if n > 0:
                r += I @ d[:, n - 1]
            for l in range(1, min(n + 1, len(K))):
                r -= K[l] @ d[:, n - l]
            d[:, n] = np.linalg.lstsq(A, r, rcond=None)[0]
        return d, betaV

    def set_pp_rule(self, nSub=64, nSing=3):
        """Rule of the numeric post-processing (PPn, PPw)."""
        self.ppn, self.ppw = self.gh_composite(nSub, nSing)

    def pp_block_num(self, x, j, tau):
        out = np.zeros(9)
        self.lib.orc_pp_block_num(C.byref(self.param), C.byref(self.mesh), int(j),
                                  _p(np.ascontiguousarray(x, np.float64)), float(tau), self.ppw.size, _p(self.ppn),
                                  _p(self.ppw), _p(out))
        return out.reshape(3, 3).T.copy()

    def pp_point_num(self, density, x, t):
        d = np.asfortranarray(density)
        u = np.zeros(3)
        self.lib.orc_pp_point_num(C.byref(self.param), C.byref(self.mesh), d.ctypes.data_as(_dp),
                                  _p(np.ascontiguousarray(x, np.float64)), float(t), self.ppw.size, _p(self.ppn),
                                  _p(self.ppw), _p(u))
        return u

