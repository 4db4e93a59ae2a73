/*
 * ebem_oracle.h -- CPU restatement (FP64) of the reference's semi-analytic (SA)
 * single-layer path, used ONLY as test infrastructure: parity checker for the
 * HIP library and the timed CPU baseline.  Nothing in the shipped library links
 * or calls this code.
 *
 * Conventions
 *   - coordinates are row-major: vertex v at coords[3v .. 3v+2];
 *   - triangles are 0-based vertex triples: tri[3t .. 3t+2];
 *   - dense matrices are column-major (MATLAB layout), e.g. a 3N x 3N E_lag
 *     block matrix has entry (r, c) at out[r + 3N*c];
 *   - a "node integral" is the value the reference accumulates per outer node
 *     and time level in SAc/BEMenerg_coreSA_calcMatrixSubBlock.m:49-93.
 */
#ifndef EBEM_ORACLE_H
#define EBEM_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

/* Material + time discretisation, util/fileRead/readInputFile.m:36-70. */
typedef struct {
    double rho, mu, lambda;
    double velP, velS;   /* sqrt((lambda+2mu)/rho), sqrt(mu/rho) */
    double Tfin;
    int nT;
} orc_param;

/* Mesh with the derived per-triangle data of util/fileRead/readSpaceMesh.m:93-139. */
typedef struct {
    int N, Nv;
    const double *coords;  /* Nv*3 */
    const int *tri;        /* N*3, 0-based */
    double *normal;        /* N*3 */
    double *area;          /* N   */
    double *center;        /* N*3 */
    double *maxL;          /* N   */
} orc_mesh;

/* Fill normal/area/center/maxL from coords/tri (readSpaceMesh.m:93-139). */
void orc_mesh_derive(int N, const double *coords, const int *tri,
                     double *normal, double *area, double *center, double *maxL);

/* Gauss-Hammer rule of GaussHammerComposite(1, n): nodes n x 3 (row-major,
 * barycentric columns in the (b1,b3,b2) order of the composite map) and
 * weights n.  n in {1,3,7,12,19,28}.  Returns 0 or -1. */
int orc_gauss_hammer(int n, double *nodes, double *weights);

/* L = min(nT, ceil(maxVertexDist/(velS*dt)) + 1)   (dir/BEMenerg_dir_calcNumBlocks.m:1-17) */
int orc_num_blocks(const orc_param *p, int Nv, const double *coords);

/* One analytic region routine of functions/BEMenerg_analiticIntg/ (kind 0..5 =
 * CC, SC_P, SC_SP, TC, TR_P, TR_SP): args {RS, rhoS, RP, rhoP, zeta, TH1, TH2, 0},
 * info the 12 fields of infoTri2D (calcoloDatiTriangolo.m:10-32), B 6 x (3x3)
 * column-major; ris 4 x (3x3) column-major (ris1..ris4, or ris5, ris6). */
int orc_intg(int kind, const double *args, const double *info, const double *B, double *ris);
/* Every region call node_integrals makes for point x, triangle T at time t, in
 * call order (arguments, infoTri2D, B, results of orc_intg); returns the number
 * of calls (only the first cap are stored).  Not thread-safe. */
int orc_region_trace(const orc_param *p, const double *x, const double *T, const double *n, double t, int cap,
                     int *kind, double *args, double *info, double *B, double *ris);
/* One outer point x, field triangle (vertices T row-major 3x3, unit normal n),
 * time t (> 0): the six kernel integrals ris1..ris6 of
 * SAc/BEMenerg_coreSA_calcMatrixSubBlock.m:84-97 *after* the inverse Euler
 * rotation of ris2/4/6.  out is 6 x 9 (each 3x3 column-major).  Returns the
 * integration class (1 SP_SP, 2 SP_P, 3 P, 0 skipped). */
int orc_node_ris(const orc_param *p, const double *x, const double *T, const double *n,
                 double t, double *out);

/* 3x3 block E_lag[S, F] for test triangle S and field triangle F, with the
 * cull and 1e-14 cleanup of SAc/BEMenerg_coreSA_calcMatrixBlock.m:41-77.
 * gh_nodes/gh_w: rule from orc_gauss_hammer.  out: 3x3 column-major.
 * Returns 1 if the pair passed the light-cone test, 0 otherwise (out = 0). */
int orc_pair_block(const orc_param *p, const orc_mesh *m, int nGH, const double *gh_nodes,
                   const double *gh_w, int lag, int S, int F, double *out);

/* Full E_lag (3N x 3N column-major), OpenMP over pairs when nthreads > 1.
 * Returns isnotzero (any pair passed the cull). */
int orc_matrix_block(const orc_param *p, const orc_mesh *m, int nGH, const double *gh_nodes,
                     const double *gh_w, int lag, double *out, int nthreads);

/* Same, but only rows [row0, row1) of test triangles; out is 3N x 3N (other
 * rows untouched).  Used for bounded CPU timing samples. */
int orc_matrix_block_rows(const orc_param *p, const orc_mesh *m, int nGH, const double *gh_nodes,
                          const double *gh_w, int lag, int row0, int row1, double *out, int nthreads);

/* Dirichlet data g(x,t).  kind:
 *  0 screen flat:  (0,0,sin^2(pi t/2) H(t))
 *  1 sphere:       n(x) t^2 exp(-4t) H(t)   (n = triangle normal)
 *  2 reference 'sphereUniform': (0,0,0.5 t)      (Tn/BEMenerg_core_calcNucleoTn.m:135-136)
 *  3 reference 'screenUniform': (0,0,[sin^5 t (t<pi/2) | 1] * x1^2) (:141-142) */
void orc_datum(int kind, const double *x, const double *nrm, double t, double *g);

/* beta: 3N x nT column-major, Tn/BEMenerg_core_calcTnBlock.m:27-58. */
void orc_rhs(const orc_param *p, const orc_mesh *m, int nGH, const double *gh_nodes,
             const double *gh_w, int datum, double *beta);

/* Dense LU with partial pivoting, in place, column-major n x n. Returns 0 or
 * -(k+1) on an exact zero pivot at column k. */
int orc_lu(int n, double *A, int *piv);
void orc_lu_solve(int n, const double *LU, const int *piv, double *b);

/* Marching on in time (SA/BEMenerg_coreSA_timeMarching.m:54-113):
 * alpha_n = E0^{-1} (beta_n - sum_{l=1}^{min(n,L)-1} E_l alpha_{n-l}).
 * E: L dense 3N x 3N matrices back to back; beta/alpha 3N x nT column-major. */
int orc_mot(int n3, int L, int nT, const double *E, const double *beta, double *alpha);

/* 3x3 matrix M(x, tau) for one triangle, ppA/BEMenerg_postProcA_calcMatrixBlock.m. */
void orc_pp_block(const orc_param *p, const double *x, const double *T, const double *n,
                  double tau, double *out);

/* u(x,t) for one point (ppA/BEMenerg_postProcA_singlePoint.m), density 3N x nT. */
void orc_pp_point(const orc_param *p, const orc_mesh *m, const double *density,
                  const double *x, double t, double *u);

/* Counters of analytic region routines reached since the last reset
 * (index: 0 CC, 1 SC_P, 2 SC_SP, 3 TC(z>0), 4 TC(z=0), 5 TR_P, 6 TR_SP(z>0),
 * 7 TR_SP(z=0), 8 CC(z=0), 9 SC_SP(z=0), 10 DeltaP branch 2, 11 branch 3). */
/* 32-double triangle record in the HIP library layout (test hook). */
void orc_field_record(const double *T, const double *n, double *rec);

void orc_counters(long long *out12);
void orc_counters_reset(void);

/* ---- numeric inner quadrature mode of the reference's GPU path ----------
 * (functions/BEMenerg_dir/CUDAkernels/kernelV.cu, BEMenerg_dir_calcSubBlockDiagV.m) */

/* GaussHammerComposite(nSub, nNod): nSub*nNod nodes (barycentric, 3 each) and weights. */
int orc_gh_composite(int nSub, int nNod, double *nodes, double *w);
/* doppioGauss1D(nNodes): collapsed Gauss rule on the reference triangle. */
int orc_doppio_gauss(int nNodes, double *nodes, double *w);
/* kernelV.cu nucleo: 3x3 column-major kernel at offset x (r = |x|), time t. */
void orc_nucleo(const double *x, double r, double t, double cP, double cS, double *nu);
/* One block (column-major) / rows [r0, r1) of E_lag, GHC mode. */
int orc_ghc_pair_block(const orc_param *p, const orc_mesh *m, int nGH, const double *ghn, const double *ghw,
                       int nInt, const double *ghcn, const double *ghcw, int nDiag, const double *dgn,
                       const double *dgw, int lag, int S, int F, double *out);
int orc_ghc_matrix_block_rows(const orc_param *p, const orc_mesh *m, int nGH, const double *ghn, const double *ghw,
                              int nInt, const double *ghcn, const double *ghcw, int nDiag, const double *dgn,
                              const double *dgw, int lag, int r0, int r1, double *out, int nthreads);

/* Numeric post-processing (BEMenerg_postProcGPU/CUDAkernels/kernelLayer.cu):
 * one triangle's M(x, tau) by the rule rn (nInt x 3) / rw, and u(x, t). */
void orc_pp_block_num(const orc_param *p, const orc_mesh *m, int F, const double *x, double tau, int nInt,
                      const double *rn, const double *rw, double *out);
void orc_pp_point_num(const orc_param *p, const orc_mesh *m, const double *density, const double *x, double t,
                      int nInt, const double *rn, const double *rw, double *u);

/* ---- double-layer K of the direct Dirichlet driver -------------------- */
/* P1 basis of a triangle (dir/BEMenerg_dir_calcCostantData.m:16-20):
 * phi_k(x) = vec[k] + mat[3k..3k+2] . x, phi_k(V_l) = delta_kl. */
void orc_k_basis(const double *T, double *mat, double *vec);
/* Block (test triangle S, vertex s) of K_lag, 3x3 column-major: kernelK.cu
 * over the field triangles F != S that have vertex s, plus
 * calcSubBlockSingK.m when s is a vertex of S.  `prune` = 1 applies the
 * reference's skip tests, 0 integrates every block.  Returns 1 if computed. */
int orc_k_block(const orc_param *p, const orc_mesh *m, int nGH, const double *ghn, const double *ghw, int nInt,
                const double *ghcn, const double *ghcw, int nDiag, const double *dgn, const double *dgw, int lag,
                int S, int s, int prune, double *out);
/* K_lag restricted to test rows [r0, r1): dense 3N x 3Nv column-major. */
int orc_k_matrix_rows(const orc_param *p, const orc_mesh *m, int nGH, const double *ghn, const double *ghw, int nInt,
                      const double *ghcn, const double *ghcw, int nDiag, const double *dgn, const double *dgw,
                      int lag, int r0, int r1, double *out, int nthreads);

/* ---- direct post-processing u = G alpha - P g (dirPP/BEMenerg_ppD_calc.m) ---- */
void orc_ppP_block(const orc_param *p, const orc_mesh *m, int nInt, const double *rn, const double *rw,
                   const double *x, double t, int y, int s, double *out);
void orc_ppG_block(const orc_param *p, const orc_mesh *m, int nInt, const double *rn, const double *rw,
                   const double *x, double t, int y, int F, double *out);
void orc_ppD_point(const orc_param *p, const orc_mesh *m, const double *density, const double *g, const double *x,
                   double t, int nInt, const double *rn, const double *rw, double *u, int nthreads);

#ifdef __cplusplus
}
#endif
#endif
