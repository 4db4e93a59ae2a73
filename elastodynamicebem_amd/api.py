"""Python mirror of the reference's MATLAB entry points on the single-layer path.

Same names, argument meaning and error behaviour as the reference functions
cited below; the computation runs in the native HIP library (libebem_hip.so).

  readInputFile                 util/fileRead/readInputFile.m
  readSpaceMesh                 util/fileRead/readSpaceMesh.m
  BEMenerg_setupCore            util/setup/BEMenerg_setupCore.m   ("SA nn" only)
  BEMenerg_dir_calcNumBlocks    functions/BEMenerg_dir/BEMenerg_dir_calcNumBlocks.m
  BEMenerg_coreSA_timeMarching  functions/BEMenerg_indir/BEMenerg_core/BEMenerg_coreSA/BEMenerg_coreSA_timeMarching.m
  BEMenerg_dir_timeMarchingD    functions/BEMenerg_dir/BEMenerg_dir_timeMarchingD.m (single-layer system)
  BEMenerg_postProcA_setup      functions/BEMenerg_indir/BEMenerg_postProc/BEMenerg_postProcA/BEMenerg_postProcA_setup.m
  BEMenerg_ppD_setup            functions/BEMenerg_dirPP/BEMenerg_ppD_setup.m
  BEMenerg_ppD_calc             functions/BEMenerg_dirPP/BEMenerg_ppD_calc.m
  BEMenerg_postProc_loadPoints   functions/BEMenerg_indir/BEMenerg_postProc/BEMenerg_postProc_loadPoints.m
  BEMenerg_postProcGPU_setupLayer / _setupSP
                                functions/BEMenerg_indir/BEMenerg_postProc/BEMenerg_postProcGPU/*.m
  BEMenerg_analiticIntg_{CC,SC_P,SC_SP,TC,TR_P,TR_SP}
                                functions/BEMenerg_analiticIntg/*.m
"""
from __future__ import annotations

import ctypes as C
import math
import os
from dataclasses import dataclass, field

import numpy as np

from . import _lib
from .mesh import SpaceMesh, read_mesh

PREC_FP64 = 0
PREC_MIXED = 1
PREC_SURROGATE = 2  # flag: near-field pairs from the attached surrogate (surrogate.py)
PREC_GHC = 4        # numeric inner quadrature of the reference's GPU path (kernelV.cu), FP64

OPT_FAR_FACTOR = 1        # mixed form: FP32 beyond this many longest edges (default 3)
OPT_SURROGATE_SOLVE = 2   # allow mot_solve on a surrogate-assembled store
OPT_SURROGATE_PROJECT = 3  # moment projection of the surrogate pairs (default on)

AI_CC, AI_SC_P, AI_SC_SP, AI_TC, AI_TR_P, AI_TR_SP = range(6)

DATUM_SCREEN_SIN2 = 0
DATUM_SPHERE_T2EXP = 1
DATUM_REF_SPHERE = 2
DATUM_REF_SCREEN = 3
DATUM_REF_BAR_H1 = 4
DATUM_REF_BAR_H3 = 5
DATUM_REF_SPHERE_WAVE = 6
DATUM_REF_ELEMENTO = 7

# datum hard-coded per domainType in Tn/BEMenerg_core_calcNucleoTn.m:133-150
_DOMAIN_DATUM = {"sphereUniform": DATUM_REF_SPHERE, "sphereNotUniform": DATUM_REF_SPHERE,
                 "screenUniform": DATUM_REF_SCREEN, "screenGraded": DATUM_REF_SCREEN,
                 "barH1": DATUM_REF_BAR_H1, "barH3": DATUM_REF_BAR_H3, "sphereWave": DATUM_REF_SPHERE_WAVE,
                 "elementoIndustriale": DATUM_REF_ELEMENTO}


@dataclass
class PbParam:
    """pbParam of readInputFile.m:10-76."""
    rho: float = 1.0
    mu: float = 1.0
    nu: float = 0.25
    lambda_: float = 1.0
    Tfin: float = 1.0
    nT: int = 16
    domainType: str = "screenUniform"
    lev: int = 2
    BIE: str = "EFIE"
    BOU: str = "DIR"
    MTDTN: str | None = None

    @property
    def velP(self) -> float:
        return math.sqrt((self.lambda_ + 2 * self.mu) / self.rho)

    @property
    def velS(self) -> float:
        return math.sqrt(self.mu / self.rho)


@dataclass
class MethodInfo:
    typeIntg: str = "SA"
    numNodiExt: int = 7
    numSubRegion: int = 0
    numNodiSing: int = 0
    numNodiInt: int = 0
    numNodiDiag: int = 0
    numSubRegionPP: int = 0
    numNodiSingPP: int = 0
    extra: dict = field(default_factory=dict)


def inner_rule_tables(nSub: int, nSing: int, nDiag: int):
    """(GHCn, GHCw, DIAGn, DIAGw) of the numeric form: GaussHammerComposite(nSub, nSing)
    and doppioGauss1D(nDiag), built by the native library."""
    lib = _lib.load()
    inner = np.zeros((nSub * nSing, 4))
    diag = np.zeros((nDiag, 4))
    _lib.check(lib.ebem_inner_rule_tables(nSub, nSing, nDiag, _lib.dptr(inner), _lib.dptr(diag)))
    return inner[:, :3].copy(), inner[:, 3].copy(), diag[:, :3].copy(), diag[:, 3].copy()


def readInputFile(path: str) -> PbParam:
    """Ten '%header'/value line pairs (readInputFile.m:10-76)."""
    if not os.path.exists(path):
        raise FileNotFoundError("Impossibile aprire file coi dati del problema")
    with open(path, "r", errors="replace") as fh:
        lines = [ln.rstrip("\n") for ln in fh]
    vals = [lines[k] for k in range(1, len(lines), 2)]
    if len(vals) < 10:
        raise ValueError("problem file has fewer than 10 header/value pairs")
    p = PbParam(rho=float(vals[0]), mu=float(vals[1]), nu=float(vals[2]), lambda_=float(vals[3]),
                Tfin=float(vals[4]), nT=int(float(vals[5])), domainType=vals[6].strip(), lev=int(float(vals[7])),
                BIE=vals[8].strip(), BOU=vals[9].strip())
    if p.BOU == "TEST" and len(vals) > 10:
        p.MTDTN = vals[10].strip()
    return p


def readSpaceMesh(path: str) -> SpaceMesh:
    """MEDIT mesh with derived data (readSpaceMesh.m:13-151)."""
    return read_mesh(path)


def GaussHammerComposite_1(n: int):
    """Outer rule of GaussHammerComposite(1, n) (same table as the native library)."""
    from . import quad
    return quad.gauss_hammer(n)


def BEMenerg_setupCore(methodSelected: str):
    """Parse "SA nn" and "FN nn nSub nSing nDiag" (BEMenerg_setupCore.m:3-93).
    "FN" is the numeric inner quadrature of the direct driver (kernelV.cu off the
    diagonal, doppioGauss1D on the diagonal); the CPU families MX.* are not
    provided and raise."""
    toks = methodSelected.split()
    if not toks:
        raise ValueError("Metodo non disponibile")
    mi = MethodInfo(typeIntg=toks[0])
    if mi.typeIntg not in ("SA", "FN"):
        raise NotImplementedError(f"method family {mi.typeIntg!r} is not provided (SA and FN are)")
    if len(toks) < 2:
        raise ValueError("Valori nodi esterni non validi")
    mi.numNodiExt = int(toks[1])
    if mi.numNodiExt not in (3, 7, 12, 19):
        raise ValueError("Valori nodi esterni non validi")
    EXTn, EXTw = GaussHammerComposite_1(mi.numNodiExt)
    if mi.typeIntg == "SA":
        return mi, EXTn, EXTw, None, None, None, None
    if len(toks) < 5:
        raise ValueError("Valori nodi interni non validi")
    mi.numSubRegion, mi.numNodiSing, mi.numNodiDiag = int(toks[2]), int(toks[3]), int(toks[4])
    if mi.numSubRegion not in (1, 4, 16, 64) or mi.numNodiSing not in (1, 3, 7, 12, 19):
        raise ValueError("Valori nodi interni non validi")
    if mi.numNodiDiag not in (16, 64, 256, 1024):
        raise ValueError("Valori nodi interni non validi")
    mi.numNodiInt = mi.numSubRegion * mi.numNodiSing
    INTn, INTw, DIAGn, DIAGw = inner_rule_tables(mi.numSubRegion, mi.numNodiSing, mi.numNodiDiag)
    return mi, EXTn, EXTw, INTn, INTw, DIAGn, DIAGw


class Solver:
    """One native handle: mesh + problem + outer rule, device-resident store."""

    def __init__(self, mesh: SpaceMesh, pb: PbParam, nGH: int = 7, device: int = 0):
        self.lib = _lib.load()
        h = C.c_void_p()
        _lib.check(self.lib.ebem_create(C.byref(h), device))
        self.h = h
        self.mesh = mesh
        self.pb = pb
        self.N = mesh.numberTriangles
        self.nT = pb.nT
        coords = np.ascontiguousarray(mesh.coordinates, np.float64)
        tri = np.ascontiguousarray(mesh.triangles, np.int32)
        _lib.check(self.lib.ebem_set_mesh(h, self.N, coords.shape[0], _lib.dptr(coords),
                                          tri.ctypes.data_as(_lib._ip), 0))
        _lib.check(self.lib.ebem_set_problem(h, pb.rho, pb.mu, pb.lambda_, pb.Tfin, pb.nT))
        _lib.check(self.lib.ebem_set_rule(h, nGH))

    def close(self):
        if getattr(self, "h", None):
            self.lib.ebem_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    # -- pieces --------------------------------------------------------------
    def set_inner_rule(self, nSub: int, nSing: int, nDiag: int) -> None:
        """Inner rules of the numeric form (PREC_GHC)."""
        _lib.check(self.lib.ebem_set_inner_rule(self.h, nSub, nSing, nDiag))

    def set_option(self, option: int, value: float) -> None:
        """OPT_FAR_FACTOR / OPT_SURROGATE_SOLVE / OPT_SURROGATE_PROJECT (include/ebem.h)."""
        _lib.check(self.lib.ebem_set_option(self.h, int(option), float(value)))

    def analitic_intg(self, kind, args, info, B) -> np.ndarray:
        """Batch of region routines of functions/BEMenerg_analiticIntg/: kind (n,),
        args (n, 8) = RS, rhoS, RP, rhoP, zeta, THETA1, THETA2, -; info (n, 12)
        infoTri2D fields; B (n, 6, 3, 3).  Returns (n, 4, 3, 3): ris1..ris4, or
        ris5, ris6 and two zero matrices."""
        kind = np.ascontiguousarray(kind, np.int32).ravel()
        n = kind.size
        args = np.ascontiguousarray(args, np.float64).reshape(n, 8)
        info = np.ascontiguousarray(info, np.float64).reshape(n, 12)
        Bc = np.ascontiguousarray(np.asarray(B, np.float64).reshape(n, 6, 3, 3).transpose(0, 1, 3, 2))
        out = np.zeros((n, 4, 3, 3))
        _lib.check(self.lib.ebem_analitic_intg(self.h, n, kind.ctypes.data_as(_lib._ip), _lib.dptr(args),
                                               _lib.dptr(info), _lib.dptr(Bc), _lib.dptr(out)))
        return out.transpose(0, 1, 3, 2).copy()

    def num_blocks(self) -> int:
        L = C.c_int()
        _lib.check(self.lib.ebem_num_blocks(self.h, C.byref(L)))
        return L.value

    def assemble(self, precision: int = PREC_FP64) -> None:
        _lib.check(self.lib.ebem_assemble(self.h, precision))

    def store_info(self) -> dict:
        v = [C.c_int64() for _ in range(4)]
        _lib.check(self.lib.ebem_store_info(self.h, *[C.byref(x) for x in v]))
        return dict(nnzb=v[0].value, pair_lags=v[1].value, items=v[2].value, bytes=v[3].value)

    def lag_blocks(self) -> np.ndarray:
        """Stored 3x3 blocks per lag (length L)."""
        out = np.zeros(self.num_blocks(), np.int64)
        _lib.check(self.lib.ebem_store_lag_blocks(self.h, out.ctypes.data_as(_lib._i64p)))
        return out

    def assembly_stats(self) -> dict:
        a, b = C.c_int64(), C.c_int64()
        _lib.check(self.lib.ebem_assembly_stats(self.h, C.byref(a), C.byref(b)))
        return dict(far_items=a.value, fp64_items=b.value)

    _KP_KEYS = ("ranges", "levels", "bf_ranges", "bf_levels", "fallbacks")

    def assembly_kp_stats(self) -> dict:
        """Level ranges of the last assembly evaluated per pair: kP ranges and their
        levels, between-front ranges and their levels, level-by-level fallbacks."""
        out = np.zeros(5, np.int64)
        _lib.check(self.lib.ebem_assembly_kp_stats(self.h, out.ctypes.data_as(_lib._i64p)))
        return dict(zip(self._KP_KEYS, out.tolist()))

    def pp_kp_stats(self) -> dict:
        """The same for the last post-processing (ranges of m per point-triangle pair)."""
        out = np.zeros(5, np.int64)
        _lib.check(self.lib.ebem_pp_kp_stats(self.h, out.ctypes.data_as(_lib._i64p)))
        return dict(zip(self._KP_KEYS, out.tolist()))

    def surrogate_stats(self) -> dict:
        a, b, ms = C.c_int64(), C.c_int64(), C.c_double()
        _lib.check(self.lib.ebem_surrogate_stats(self.h, C.byref(a), C.byref(b), C.byref(ms)))
        return dict(items=a.value, blocks=b.value, ms=ms.value)

    def get_block(self, lag: int) -> np.ndarray:
        """Dense E_lag (3N x 3N), row index = test DOF."""
        n3 = 3 * self.N
        out = np.zeros((n3, n3), order="F")
        _lib.check(self.lib.ebem_get_block(self.h, lag, out.ctypes.data_as(_lib._dp)))
        return out

    def get_block_rows(self, lag: int, r0: int, r1: int) -> np.ndarray:
        """Rows of test triangles [r0, r1) of E_lag: (3 (r1 - r0)) x 3N."""
        out = np.zeros((3 * (r1 - r0), 3 * self.N), order="F")
        _lib.check(self.lib.ebem_get_block_rows(self.h, lag, r0, r1, out.ctypes.data_as(_lib._dp)))
        return out

    def rhs(self, datum: int) -> np.ndarray:
        out = np.zeros((3 * self.N, self.nT), order="F")
        _lib.check(self.lib.ebem_rhs(self.h, datum, out.ctypes.data_as(_lib._dp)))
        return out

    # -- double-layer K of the direct Dirichlet driver -------------------------
    def k_num_blocks(self) -> int:
        """numBlocksK = min(L + 1, nT)."""
        L = C.c_int()
        _lib.check(self.lib.ebem_k_num_blocks(self.h, C.byref(L)))
        return L.value

    def assemble_k(self) -> None:
        """K_0 .. K_{LK-1} on the device (needs set_inner_rule)."""
        _lib.check(self.lib.ebem_assemble_k(self.h))

    def k_store_info(self) -> dict:
        v = [C.c_int64() for _ in range(3)]
        _lib.check(self.lib.ebem_k_store_info(self.h, *[C.byref(x) for x in v]))
        return dict(pairs=v[0].value, slots=v[1].value, bytes=v[2].value)

    def get_k_block(self, lag: int, rows=None) -> np.ndarray:
        """K_lag (3N x 3Nv), row index = test DOF; with rows=(r0, r1) only those
        test triangles' rows (3 (r1 - r0) x 3Nv)."""
        r0, r1 = (0, self.N) if rows is None else rows
        out = np.zeros((3 * (r1 - r0), 3 * self.mesh.number_nodes), order="F")
        _lib.check(self.lib.ebem_get_k_block_rows(self.h, lag, r0, r1, out.ctypes.data_as(_lib._dp)))
        return out

    def dirichlet_datum(self, datum: int) -> np.ndarray:
        """Vertex datum g (3Nv x nT), column n at time (n+1) dt."""
        out = np.zeros((3 * self.mesh.number_nodes, self.nT), order="F")
        _lib.check(self.lib.ebem_dirichlet_datum(self.h, datum, out.ctypes.data_as(_lib._dp)))
        return out

    def beta_k(self, g: np.ndarray) -> np.ndarray:
        """beta_K (3N x nT) = sum_j K_{n-j} g_j for a vertex datum g (3Nv x nT)."""
        g = np.asfortranarray(g, dtype=np.float64)
        if g.shape != (3 * self.mesh.number_nodes, self.nT):
            raise ValueError("g must be 3Nv x nT")
        out = np.zeros((3 * self.N, self.nT), order="F")
        _lib.check(self.lib.ebem_beta_k(self.h, g.ctypes.data_as(_lib._dp), out.ctypes.data_as(_lib._dp)))
        return out

    def rhs_dirichlet(self, datum: int) -> np.ndarray:
        """Right-hand side rhs / 2 + beta_K of the direct Dirichlet driver, kept on
        the device for mot_solve(None)."""
        out = np.zeros((3 * self.N, self.nT), order="F")
        _lib.check(self.lib.ebem_rhs_dirichlet(self.h, datum, out.ctypes.data_as(_lib._dp)))
        return out

    # -- Neumann driver -----------------------------------------------------------
    def neumann_datum(self, datum: int) -> np.ndarray:
        """Traction datum gV (3N x nT) at the triangle centres, column n at (n+1) dt."""
        out = np.zeros((3 * self.N, self.nT), order="F")
        _lib.check(self.lib.ebem_neumann_datum(self.h, datum, out.ctypes.data_as(_lib._dp)))
        return out

    def beta_v(self, gV: np.ndarray) -> np.ndarray:
        """beta_V (3N x nT) = sum_j V_{n-j} gV_j over the assembled single-layer blocks."""
        g = np.asfortranarray(gV, dtype=np.float64)
        if g.shape != (3 * self.N, self.nT):
            raise ValueError("gV must be 3N x nT")
        out = np.zeros((3 * self.N, self.nT), order="F")
        _lib.check(self.lib.ebem_beta_v(self.h, g.ctypes.data_as(_lib._dp), out.ctypes.data_as(_lib._dp)))
        return out

    def mot_solve_neumann(self, datum: int) -> np.ndarray:
        """Density (3Nv x nT) of the Neumann driver (needs assemble and assemble_k)."""
        out = np.zeros((3 * self.mesh.number_nodes, self.nT), order="F")
        _lib.check(self.lib.ebem_mot_solve_neumann(self.h, datum, out.ctypes.data_as(_lib._dp)))
        return out

    def _density_arg(self, a, what):
        if a is None:
            return None
        a = np.asfortranarray(a, dtype=np.float64)
        if a.shape != (3 * self.N, self.nT):
            raise ValueError(f"{what} must be 3N x nT = {(3 * self.N, self.nT)}, got {a.shape}")
        return a

    def mot_solve(self, beta: np.ndarray | None = None) -> np.ndarray:
        out = np.zeros((3 * self.N, self.nT), order="F")
        b = self._density_arg(beta, "beta")
        _lib.check(self.lib.ebem_mot_solve(self.h, None if b is None else b.ctypes.data_as(_lib._dp),
                                           out.ctypes.data_as(_lib._dp)))
        return out

    def pp_eval(self, x: np.ndarray, t: np.ndarray, density: np.ndarray | None = None) -> np.ndarray:
        """u (3, npts, nt)."""
        x = np.ascontiguousarray(x, np.float64).reshape(-1, 3)
        t = np.ascontiguousarray(t, np.float64).ravel()
        out = np.zeros((t.size, x.shape[0], 3))
        d = self._density_arg(density, "density")
        _lib.check(self.lib.ebem_pp_eval(self.h, None if d is None else d.ctypes.data_as(_lib._dp), x.shape[0],
                                         _lib.dptr(x), t.size, _lib.dptr(t), _lib.dptr(out)))
        return out.transpose(2, 1, 0).copy()

    def pp_prepare(self, x: np.ndarray, t: np.ndarray) -> None:
        """Start the density-independent part of pp_eval(x, t) on the library's
        side stream (ebem_pp_prepare), to overlap the next mot_solve."""
        x = np.ascontiguousarray(x, np.float64).reshape(-1, 3)
        t = np.ascontiguousarray(t, np.float64).ravel()
        _lib.check(self.lib.ebem_pp_prepare(self.h, x.shape[0], _lib.dptr(x), t.size, _lib.dptr(t)))

    def set_pp_rule(self, nSub: int, nSing: int) -> None:
        """Rule of the numeric post-processing (pp_eval_numeric)."""
        _lib.check(self.lib.ebem_set_pp_rule(self.h, nSub, nSing))

    def pp_eval_numeric(self, x: np.ndarray, t: np.ndarray, density: np.ndarray | None = None) -> np.ndarray:
        """u (3, npts, nt) by the reference's numeric GPU post-processing (kernelLayer.cu)."""
        x = np.ascontiguousarray(x, np.float64).reshape(-1, 3)
        t = np.ascontiguousarray(t, np.float64).ravel()
        out = np.zeros((t.size, x.shape[0], 3))
        d = self._density_arg(density, "density")
        _lib.check(self.lib.ebem_pp_eval_numeric(self.h, None if d is None else d.ctypes.data_as(_lib._dp),
                                                 x.shape[0], _lib.dptr(x), t.size, _lib.dptr(t), _lib.dptr(out)))
        return out.transpose(2, 1, 0).copy()

    def pp_eval_direct(self, x: np.ndarray, t: np.ndarray, density: np.ndarray | None = None,
                       g: np.ndarray | None = None) -> np.ndarray:
        """u (3, npts, nt) = sum_y G_y alpha_y - P_y g_y of the direct
        post-processing (BEMenerg_ppD_calc.m; needs set_pp_rule).  density None:
        the device density of mot_solve; g None: the vertex datum of the last
        rhs_dirichlet."""
        x = np.ascontiguousarray(x, np.float64).reshape(-1, 3)
        t = np.ascontiguousarray(t, np.float64).ravel()
        out = np.zeros((t.size, x.shape[0], 3))
        d = self._density_arg(density, "density")
        gg = None if g is None else np.asfortranarray(g, dtype=np.float64)
        if gg is not None and gg.shape != (3 * self.mesh.number_nodes, self.nT):
            raise ValueError("g must be 3Nv x nT")
        _lib.check(self.lib.ebem_pp_eval_direct(self.h, None if d is None else d.ctypes.data_as(_lib._dp),
                                                None if gg is None else gg.ctypes.data_as(_lib._dp), x.shape[0],
                                                _lib.dptr(x), t.size, _lib.dptr(t), _lib.dptr(out)))
        return out.transpose(2, 1, 0).copy()

    def node_integral(self, x, tri: int, t: float) -> np.ndarray:
        x = np.ascontiguousarray(x, np.float64)
        out = np.zeros(9)
        _lib.check(self.lib.ebem_node_integral(self.h, _lib.dptr(x), int(tri), float(t), _lib.dptr(out)))
        return out.reshape(3, 3).T.copy()

    def timings(self) -> dict:
        keys = ["levels_ms", "assembly_ms", "factor_ms", "history_ms", "apply_ms", "mot_ms", "pp_ms",
                "history_launches", "levels_kp_ms", "pp_kp_ms", "levels_bf_ms", "pp_bf_ms", "k_assembly_ms",
                "beta_k_ms", "pp_direct_ms", "mot_neumann_ms", "node_tasks", "node_tasks_reused",
                "node_reuse_fallbacks", "pp_prepare_ms"]
        out = np.zeros(len(keys))
        _lib.check(self.lib.ebem_timings(self.h, _lib.dptr(out), len(keys)))
        return dict(zip(keys, out.tolist()))


def _datum_for(pb: PbParam, datum):
    if datum is not None:
        return datum
    if pb.domainType in _DOMAIN_DATUM:
        return _DOMAIN_DATUM[pb.domainType]
    raise ValueError("Caso ancora non riportato/Errore nei dati")


def BEMenerg_dir_calcNumBlocks(pbParam: PbParam, domainMesh: SpaceMesh) -> int:
    with Solver(domainMesh, pbParam) as s:
        return s.num_blocks()


def BEMenerg_coreSA_timeMarching(pbParam: PbParam, domainMesh: SpaceMesh, methodInfo: MethodInfo, GHn=None,
                                 GHw=None, *, datum=None, precision: int = PREC_FP64, device: int = 0,
                                 save_dir: str | None = None):
    """density (3N x nT): E_0..E_{L-1} assembly, RHS and MOT on the GPU.  With
    save_dir, also written as the reference does (SA_timeMarching.m:116-117):
    <save_dir>/<domainType><lev>_SA_<numNodiExt>_density.mat."""
    with Solver(domainMesh, pbParam, methodInfo.numNodiExt, device) as s:
        s.assemble(precision)
        s.rhs(_datum_for(pbParam, datum))
        density = s.mot_solve(None)
    if save_dir is not None:
        from . import matio
        matio.save_density(matio.output_base(pbParam, "SA", methodInfo, save_dir), density)
    return density


def BEMenerg_dir_timeMarchingD(pbParam, domainMesh, methodInfo, GHn=None, GHw=None, GHCn=None, GHCw=None,
                               DIAGn=None, DIAGw=None, *, datum=None, precision: int = PREC_FP64, device: int = 0,
                               save_dir: str | None = None):
    """Direct Dirichlet driver (dir/BEMenerg_dir_timeMarchingD.m): with methodInfo
    "FN nExt nSub nSing nDiag" the single-layer blocks by the numeric inner
    quadrature of the reference's GPU kernel (kernelV.cu + calcSubBlockDiagV.m,
    PREC_GHC), the double-layer blocks K_0 .. K_{LK-1} (kernelK.cu +
    calcSubBlockSingK.m), the right-hand side rhs / 2 + beta_K (:163-166, 272,
    287) and the MOT with E_0 inverted (:277-296), all on the device.  Raises
    if lambda + mu == 0, like main.m:61-63.  With save_dir, the density is also
    written as the reference does (:300-301):
    <save_dir>/<domainType><lev>_NEW_<ext>_<sub>_<sing>_<diag>_density.mat."""
    if pbParam.lambda_ + pbParam.mu == 0:
        raise ValueError("Problema non adatto a questa formulazione")
    if methodInfo.typeIntg != "FN":
        raise ValueError("the direct Dirichlet driver needs an FN method (numeric inner rules of K)")
    kind = _datum_for(pbParam, datum)
    with Solver(domainMesh, pbParam, methodInfo.numNodiExt, device) as s:
        s.set_inner_rule(methodInfo.numSubRegion, methodInfo.numNodiSing, methodInfo.numNodiDiag)
        s.assemble(PREC_GHC)
        s.assemble_k()
        s.rhs_dirichlet(kind)
        density = s.mot_solve(None)
    if save_dir is not None:
        from . import matio
        matio.save_density(matio.output_base(pbParam, "NEW", methodInfo, save_dir), density)
    return density


def BEMenerg_dir_timeMarchingN(pbParam, domainMesh, methodInfo, GHn=None, GHw=None, GHCn=None, GHCw=None,
                               DIAGn=None, DIAGw=None, *, datum=None, device: int = 0,
                               save_dir: str | None = None):
    """Neumann driver (dir/BEMenerg_dir_timeMarchingN.m): V in the numeric form,
    K, beta_V of the traction datum, and the marching with the least-squares
    solve of (K_0 + I_Gamma); density 3Nv x nT at the vertices.  With save_dir,
    written as the reference does (:299-301)."""
    if pbParam.lambda_ + pbParam.mu == 0:
        raise ValueError("Problema non adatto a questa formulazione")
    if methodInfo.typeIntg != "FN":
        raise ValueError("the Neumann driver needs an FN method (numeric inner rules of V and K)")
    kind = _datum_for(pbParam, datum)
    with Solver(domainMesh, pbParam, methodInfo.numNodiExt, device) as s:
        s.set_inner_rule(methodInfo.numSubRegion, methodInfo.numNodiSing, methodInfo.numNodiDiag)
        s.assemble(PREC_GHC)
        s.assemble_k()
        density = s.mot_solve_neumann(kind)
    if save_dir is not None:
        from . import matio
        matio.save_density(matio.output_base(pbParam, "NEW", methodInfo, save_dir), density)
    return density


def BEMenerg_dir_calcBoundDataNeumann(deltaT, pbParam, domainMesh, *, datum=None, device: int = 0):
    """neuBoundData: list of nT traction data (3N each, at the triangle centres),
    entry n at time (n+1) dt (dir/BEMenerg_dir_calcBoundDataNeumann.m)."""
    if abs(deltaT - pbParam.Tfin / pbParam.nT) > 1e-12 * max(1.0, deltaT):
        raise ValueError("deltaT must be Tfin / nT")
    with Solver(domainMesh, pbParam, 7, device) as s:
        g = s.neumann_datum(_datum_for(pbParam, datum))
    return [g[:, n].copy() for n in range(pbParam.nT)]


def BEMenerg_ppN_setup(pbParam, domainMesh, density, numPoints, xVal, tVal, methodInfo=None, PPn=None, PPw=None,
                       *, datum=None, device: int = 0, save_dir: str | None = None):
    """campoVett of the Neumann post-processing (dirPP/BEMenerg_ppN_setup.m):
    u = sum_y G_y gV_y - P_y d_y, gV the traction datum, d the vertex density
    (BEMenerg_ppN_calc.m:121)."""
    xVal = np.asarray(xVal, float).reshape(-1, 3)
    tVal = np.asarray(tVal, float).ravel()
    if numPoints != xVal.shape[0] * tVal.size:
        raise ValueError("numPoints does not match xVal x tVal")
    nsub, nsing = _pp_rule_of(methodInfo, PPn)
    with Solver(domainMesh, pbParam, 7, device) as s:
        s.set_pp_rule(nsub, nsing)
        gV = s.neumann_datum(_datum_for(pbParam, datum))
        u = s.pp_eval_direct(xVal, tVal, gV, density)
    campo = u.transpose(1, 2, 0)
    if save_dir is not None:
        from . import matio
        matio.save_field(matio.output_base(pbParam, "NEW_PP", None, save_dir), campo)
    return campo


def BEMenerg_dir_calcBoundDataDirichlet(deltaT, pbParam, domainMesh, *, datum=None, device: int = 0):
    """dirBoundData: list of nT vertex data (3Nv each), entry n at time (n+1) dt
    (dir/BEMenerg_dir_calcBoundDataDirichlet.m)."""
    if abs(deltaT - pbParam.Tfin / pbParam.nT) > 1e-12 * max(1.0, deltaT):
        raise ValueError("deltaT must be Tfin / nT")
    with Solver(domainMesh, pbParam, 7, device) as s:
        g = s.dirichlet_datum(_datum_for(pbParam, datum))
    return [g[:, n].copy() for n in range(pbParam.nT)]


def BEMenerg_dir_calcBetaKgpu(deltaT, pbParam, domainMesh, methodInfo, *, datum=None, device: int = 0):
    """betaK: list of nT vectors (3N each), beta_n = sum_{j<=n} K_{n-j} g_j with the
    Dirichlet datum of the domain (dir/BEMenerg_dir_calcBetaKgpu.m); the K blocks
    are assembled here (the reference passes matrixSavedK)."""
    if abs(deltaT - pbParam.Tfin / pbParam.nT) > 1e-12 * max(1.0, deltaT):
        raise ValueError("deltaT must be Tfin / nT")
    with Solver(domainMesh, pbParam, methodInfo.numNodiExt, device) as s:
        s.set_inner_rule(methodInfo.numSubRegion, methodInfo.numNodiSing, methodInfo.numNodiDiag)
        s.assemble_k()
        b = s.beta_k(s.dirichlet_datum(_datum_for(pbParam, datum)))
    return [b[:, n].copy() for n in range(pbParam.nT)]


def BEMenerg_postProcA_setup(pbParam, domainMesh, density, numPoints, xVal, tVal, *, device: int = 0,
                             save_dir: str | None = None):
    """campoVett: xSize x tSize array of 3-vectors (postProcA_setup.m:1-22).  With
    save_dir, also written as the reference does (:21-22):
    <save_dir>/<domainType><lev>_A_campoVett.mat."""
    xVal = np.asarray(xVal, float).reshape(-1, 3)
    tVal = np.asarray(tVal, float).ravel()
    if numPoints != xVal.shape[0] * tVal.size:
        raise ValueError("numPoints does not match xVal x tVal")
    with Solver(domainMesh, pbParam, 7, device) as s:
        u = s.pp_eval(xVal, tVal, density)
    campo = u.transpose(1, 2, 0)  # (xSize, tSize, 3)
    if save_dir is not None:
        from . import matio
        matio.save_field(matio.output_base(pbParam, "A", None, save_dir), campo)
    return campo


def _pp_rule_of(methodInfo, PPn):
    nsub = getattr(methodInfo, "numSubRegionPP", 0) or methodInfo.extra.get("numSubRegionPP", 0)
    nsing = getattr(methodInfo, "numNodiSingPP", 0) or methodInfo.extra.get("numNodiSingPP", 0)
    if not nsub or not nsing:
        if PPn is None:
            raise ValueError("numeric post-processing needs numSubRegionPP / numNodiSingPP")
        nsing = 3
        nsub = int(np.asarray(PPn).shape[0]) // nsing
    return int(nsub), int(nsing)


def BEMenerg_postProc_loadPoints(pbParam: PbParam):
    """(xVal, tVal, iVal, numPoints, typePlot): the post-processing point and time
    sets of each domain type (BEMenerg_postProc_loadPoints.m:1-71).  Grids are
    flattened in MATLAB's column-major order."""
    d, T = pbParam.domainType, pbParam.Tfin
    empty = np.zeros((0, 3))
    typePlot = ""
    if d in ("screenTest", "screenGraded", "sphereUniform", "sphereNotUniform"):
        xVal, tVal, iVal = empty, np.zeros(0), []
    elif d == "screenUniform":
        xVal, tVal, iVal = empty, np.zeros(0), [3]
    elif d == "barH1":
        xVal = np.array([[0, 0, 0.25], [0, 0, 0.50], [0, 0, 0.75]])
        tVal = np.linspace(0, T, int(np.floor(100 * T)))
        iVal, typePlot = [3], "u(x, :)"
    elif d == "barH3":
        xVal = np.array([[0.25, -0.25, 2]])
        tVal = np.linspace(0, T, int(np.floor(100 * T)))
        iVal, typePlot = [3], "u(x, :)"
    elif d == "sphereWave":
        c = np.linspace(-5, 5, 250)
        X2D, Y2D = np.meshgrid(c, c)
        X, Y = X2D.ravel(order="F"), Y2D.ravel(order="F")
        keep = X ** 2 + Y ** 2 > 1
        xVal = np.stack([X[keep], Y[keep], np.zeros(int(keep.sum()))], axis=1)
        tVal = np.array([1.2676, 1.5211, 1.7746, 2.0282, 2.2817, 2.5352, 2.7887, 3.0423])
        iVal, typePlot = [1, 2], "u(:, t)"
    elif d == "elementoIndustriale":
        Y2D, Z2D = np.meshgrid(np.linspace(-0.2, 0.2, 250), np.linspace(-0.4, 0.0, 250))
        Y, Z = Y2D.ravel(order="F"), Z2D.ravel(order="F")
        keep = ~((np.abs(Y) > 0.17) & (Z < -0.01))
        xVal = np.stack([np.zeros(int(keep.sum())), Y[keep], Z[keep]], axis=1)
        tVal = np.array([0.0005, 0.0010, 0.0015, 0.0020, 0.0025, 0.0030, 0.0035, 0.0040, 0.0045, 0.0050, 0.0055,
                         0.0065, 0.0070, 0.0075])
        iVal, typePlot = [1, 2, 3], "u(:, t)"
    else:
        raise ValueError("Caso ancora non riportato/Errore nei dati")
    return xVal, tVal, iVal, xVal.shape[0] * tVal.size, typePlot


def BEMenerg_postProcGPU_setupLayer(pbParam, domainMesh, density, numPoints, xVal, tVal, methodInfo, PPn=None,
                                    PPw=None, *, device: int = 0):
    """campoVett (xSize x tSize of 3-vectors) by the numeric GPU post-processing
    (BEMenerg_postProcGPU_setupLayer.m / kernelLayer.cu; setupSP is the same
    sum point by point): composite rule GaussHammerComposite(numSubRegionPP, numNodiSingPP)."""
    xVal = np.asarray(xVal, float).reshape(-1, 3)
    tVal = np.asarray(tVal, float).ravel()
    if numPoints != xVal.shape[0] * tVal.size:
        raise ValueError("numPoints does not match xVal x tVal")
    nsub, nsing = _pp_rule_of(methodInfo, PPn)
    with Solver(domainMesh, pbParam, 7, device) as s:
        s.set_pp_rule(nsub, nsing)
        u = s.pp_eval_numeric(xVal, tVal, density)
    out = np.empty((xVal.shape[0], tVal.size), dtype=object)
    for a in range(xVal.shape[0]):
        for k in range(tVal.size):
            out[a, k] = u[:, a, k].reshape(3, 1)
    return out


BEMenerg_postProcGPU_setupSP = BEMenerg_postProcGPU_setupLayer


def BEMenerg_ppD_setup(pbParam, domainMesh, density, numPoints, xVal, tVal, methodInfo=None, PPn=None, PPw=None,
                       *, datum=None, device: int = 0, save_dir: str | None = None):
    """campoVett (xSize x tSize x 3) of the direct post-processing
    (dirPP/BEMenerg_ppD_setup.m:1-24): u = sum_y G_y alpha_y - P_y g_y at every
    (x, t) of xVal x tVal, with the Dirichlet datum g of the domain
    (calcBoundDataDirichlet.m; domains without one raise, as there) and the
    rule numSubRegionPP x numNodiSingPP (or PPn / PPw) of methodInfo.  With
    save_dir, also written as the reference does (:23-24):
    <save_dir>/<domainType><lev>_NEW_campoVett.mat."""
    xVal = np.asarray(xVal, float).reshape(-1, 3)
    tVal = np.asarray(tVal, float).ravel()
    if numPoints != xVal.shape[0] * tVal.size:
        raise ValueError("numPoints does not match xVal x tVal")
    kind = _datum_for(pbParam, datum)
    if kind < DATUM_REF_BAR_H1:
        raise ValueError("Problema non codificato")
    nsub, nsing = _pp_rule_of(methodInfo, PPn)
    with Solver(domainMesh, pbParam, 7, device) as s:
        s.set_pp_rule(nsub, nsing)
        g = s.dirichlet_datum(kind)
        u = s.pp_eval_direct(xVal, tVal, density, g)
    campo = u.transpose(1, 2, 0)
    if save_dir is not None:
        from . import matio
        matio.save_field(matio.output_base(pbParam, "NEW_PP", None, save_dir), campo)
    return campo


def BEMenerg_ppD_calc(pbParam, domainMesh, density, gK, methodInfo, x, t, PPn=None, PPw=None, *, device: int = 0):
    """uXT (3,) at one point and time (dirPP/BEMenerg_ppD_calc.m): gK the list
    of nT vertex data of calcBoundDataDirichlet (or a 3Nv x nT array)."""
    g = np.column_stack([np.asarray(v, float).ravel() for v in gK]) if isinstance(gK, (list, tuple)) else gK
    nsub, nsing = _pp_rule_of(methodInfo, PPn)
    with Solver(domainMesh, pbParam, 7, device) as s:
        s.set_pp_rule(nsub, nsing)
        u = s.pp_eval_direct(np.asarray(x, float).reshape(1, 3), np.array([float(t)]), density, g)
    return u[:, 0, 0].copy()


# ---- functions/BEMenerg_analiticIntg/ -----------------------------------------

class _RegionDevice:
    """A mesh-less native handle on one device for the region routines."""

    def __init__(self, device: int = 0):
        self.lib = _lib.load()
        h = C.c_void_p()
        _lib.check(self.lib.ebem_create(C.byref(h), device))
        self.h = h

    def __del__(self):
        try:
            if getattr(self, "h", None):
                self.lib.ebem_destroy(self.h)
                self.h = None
        except Exception:
            pass


_REGION_DEV: dict = {}


def _region_call(fn: str, scalars, infoTri2D, B, nris: int, device: int):
    d = _REGION_DEV.get(device)
    if d is None:
        d = _REGION_DEV[device] = _RegionDevice(device)
    info = _info_array(infoTri2D)
    Bc = np.ascontiguousarray(np.asarray(B, np.float64).reshape(3, 3, 6).transpose(2, 1, 0))  # B(:,:,k) col-major
    out = np.zeros(9 * nris)
    _lib.check(getattr(d.lib, fn)(d.h, *[float(v) for v in scalars], _lib.dptr(info), _lib.dptr(Bc),
                                  _lib.dptr(out)))
    return tuple(out[9 * k:9 * k + 9].reshape(3, 3).T.copy() for k in range(nris))


_INFO_FIELDS = ("a", "b", "c", "cos_alpha", "cos_beta", "cos_gamma", "sin_alpha", "sin_beta", "sin_gamma",
                "alpha", "beta", "gamma")


def _info_array(infoTri2D) -> np.ndarray:
    """infoTri2D as the 12 doubles of the C ABI: a dict / object with the fields
    of calcoloDatiTriangolo.m:10-32, or already a 12-vector."""
    if isinstance(infoTri2D, dict):
        return np.array([float(infoTri2D[k]) for k in _INFO_FIELDS])
    if hasattr(infoTri2D, "cos_gamma"):
        return np.array([float(getattr(infoTri2D, k)) for k in _INFO_FIELDS])
    a = np.ascontiguousarray(infoTri2D, np.float64).ravel()
    if a.size != 12:
        raise ValueError("infoTri2D must carry the 12 fields of calcoloDatiTriangolo")
    return a


def calcoloDatiTriangolo(verticiTriangolo):
    """(tri2D, infoTri2D) of a 2-D triangle (geo/calcoloDatiTriangolo.m:1-39):
    side lengths, cosines by the law of cosines, sines, angles, and the
    triangle rotated so its first edge lies on +x."""
    v = np.asarray(verticiTriangolo, float)
    a = math.hypot(v[1, 0] - v[0, 0], v[1, 1] - v[0, 1])
    b = math.hypot(v[2, 0] - v[0, 0], v[2, 1] - v[0, 1])
    c = math.hypot(v[2, 0] - v[1, 0], v[2, 1] - v[1, 1])
    ca = (b * b + c * c - a * a) / (2 * b * c)
    cb = (a * a + c * c - b * b) / (2 * a * c)
    cg = (a * a + b * b - c * c) / (2 * a * b)
    info = dict(a=a, b=b, c=c, cos_alpha=ca, cos_beta=cb, cos_gamma=cg,
                sin_alpha=math.sqrt(max(0.0, 1 - ca * ca)), sin_beta=math.sqrt(max(0.0, 1 - cb * cb)),
                sin_gamma=math.sqrt(max(0.0, 1 - cg * cg)),
                alpha=math.acos(max(-1.0, min(1.0, ca))), beta=math.acos(max(-1.0, min(1.0, cb))),
                gamma=math.acos(max(-1.0, min(1.0, cg))))
    tri2D = np.array([[0.0, 0.0], [a, 0.0], [b * cg, b * info["sin_gamma"]]])
    return tri2D, info


def BEMenerg_analiticIntg_CC(RS, rhoS, RP, rhoP, zeta, THETA1, THETA2, infoTri2D, B, *, device: int = 0):
    """[ris1, ris2, ris3, ris4] of AI/BEMenerg_analiticIntg_CC.m (B: 3x3x6)."""
    return _region_call("ebem_analiticIntg_CC", (RS, rhoS, RP, rhoP, zeta, THETA1, THETA2), infoTri2D, B, 4, device)


def BEMenerg_analiticIntg_SC_P(RP, rhoP, zeta, THETA1, THETA2, infoTri2D, B, *, device: int = 0):
    """[ris1, ris2, ris3, ris4] of AI/BEMenerg_analiticIntg_SC_P.m."""
    return _region_call("ebem_analiticIntg_SC_P", (RP, rhoP, zeta, THETA1, THETA2), infoTri2D, B, 4, device)


def BEMenerg_analiticIntg_SC_SP(RS, rhoS, zeta, THETA1, THETA2, infoTri2D, B, *, device: int = 0):
    """[ris5, ris6] of AI/BEMenerg_analiticIntg_SC_SP.m."""
    return _region_call("ebem_analiticIntg_SC_SP", (RS, rhoS, zeta, THETA1, THETA2), infoTri2D, B, 2, device)


def BEMenerg_analiticIntg_TC(zeta, RS, rhoS, THETA1, THETA2, infoTri2D, B, *, device: int = 0):
    """[ris1, ris2, ris3, ris4] of AI/BEMenerg_analiticIntg_TC.m."""
    return _region_call("ebem_analiticIntg_TC", (zeta, RS, rhoS, THETA1, THETA2), infoTri2D, B, 4, device)


def BEMenerg_analiticIntg_TR_P(zeta, THETA1, THETA2, infoTri2D, B, *, device: int = 0):
    """[ris1, ris2, ris3, ris4] of AI/BEMenerg_analiticIntg_TR_P.m."""
    return _region_call("ebem_analiticIntg_TR_P", (zeta, THETA1, THETA2), infoTri2D, B, 4, device)


def BEMenerg_analiticIntg_TR_SP(zeta, THETA1, THETA2, infoTri2D, B, *, device: int = 0):
    """[ris5, ris6] of AI/BEMenerg_analiticIntg_TR_SP.m."""
    return _region_call("ebem_analiticIntg_TR_SP", (zeta, THETA1, THETA2), infoTri2D, B, 2, device)
