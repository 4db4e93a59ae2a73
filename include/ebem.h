/*
 * ebem.h -- C ABI of the MI355X (gfx950) energetic space-time Galerkin BEM
 * single-layer hot path: E_lag assembly, marching-on-in-time (MOT) solve and
 * field post-processing.
 *
 * The reference reaches its device code only through MATLAB
 * (parallel.gpu.CUDAKernel + feval, functions/BEMenerg_dir/BEMenerg_dir_launchKernelV.m:7-14)
 * and through the MATLAB functions named below; each entry point cites the
 * reference interface it replaces.  A MEX gateway (elastodynamicebem_amd/mex/)
 * and a Python ctypes layer (elastodynamicebem_amd/api.py) bind these symbols.
 *
 * Conventions (all match the MATLAB layouts of the reference):
 *   - coordinates: Nv x 3, row-major (vertex v at coords[3v..3v+2]);
 *   - triangles:   N x 3 vertex indices, 0- or 1-based (flag);
 *   - E_lag, beta, density: column-major, DOF index 3*t + component, i.e. the
 *     3N x 3N blocks of BEMenerg_coreSA_calcMatrixBlock.m and the 3N x nT
 *     arrays of BEMenerg_coreSA_timeMarching.m;
 *   - every function returns 0 on success or a negative EBEM_E* code; the
 *     message is available from ebem_last_error() on the calling thread.
 * A handle serialises its own calls with a mutex (the reference calls the GPU
 * from parfor workers, dirPP/BEMenerg_ppD_calc.m:8); distinct handles may be
 * used from distinct threads.
 */
#ifndef EBEM_H
#define EBEM_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define EBEM_OK 0
#define EBEM_EINVAL -1   /* bad argument / state */
#define EBEM_EHIP -2     /* HIP runtime failure */
#define EBEM_ENOMEM -3   /* device allocation failed (dir/BEMenerg_dir_timeMarchingD.m:20-22) */
#define EBEM_ESINGULAR -4 /* E_0 factorisation hit a zero pivot */
#define EBEM_ESOLVER -5  /* rocSOLVER/rocBLAS failure */

/* Integration precision of ebem_assemble. */
#define EBEM_PREC_FP64 0   /* plain form: everything FP64 */
#define EBEM_PREC_MIXED 1  /* far-field pairs (centroid distance > 3h) integrated in FP32 */
#define EBEM_PREC_SURROGATE 2 /* flag: near-field pairs predicted by the attached surrogate MLP */
#define EBEM_PREC_GHC 4    /* numeric inner quadrature of the reference's GPU path (kernelV.cu), FP64 */

/* Dirichlet data of ebem_rhs (Tn/BEMenerg_core_calcNucleoTn.m:133-183). */
#define EBEM_DATUM_SCREEN_SIN2 0    /* (0,0,sin^2(pi t/2) H(t)) */
#define EBEM_DATUM_SPHERE_T2EXP 1   /* n(x) t^2 exp(-4t) H(t), n = triangle normal */
#define EBEM_DATUM_REF_SPHERE 2     /* reference 'sphereUniform': (0,0,0.5 t) */
#define EBEM_DATUM_REF_SCREEN 3     /* reference 'screenUniform': (0,0,[sin^5 t | 1] x1^2) */
/* Data of the direct Dirichlet problems; also have a vertex trace for
 * ebem_dirichlet_datum / ebem_rhs_dirichlet (dir/BEMenerg_dir_calcBoundDataDirichlet.m). */
#define EBEM_DATUM_REF_BAR_H1 4     /* 'barH1': (0,0,u(z,t)), the 1-D bar solution of height 1 */
#define EBEM_DATUM_REF_BAR_H3 5     /* 'barH3': the same for height 3 */
#define EBEM_DATUM_REF_SPHERE_WAVE 6 /* 'sphereWave': (-+exp(-20 (x1 - 2 + cP t - 0.475)^2), 0, 0) */
#define EBEM_DATUM_REF_ELEMENTO 7   /* 'elementoIndustriale': (0,0,-exp(-20 (x3 + 1 - cP t)^2)); vertex trace 0 */
#define EBEM_DATUM_MAX 7

typedef struct ebem_handle ebem_handle;
typedef struct ebem_sur ebem_sur;

/* Thread-local message of the last failing call on this thread. */
const char *ebem_last_error(void);

/* Library/device set-up.  Replaces gpuDevice/reset (dir/BEMenerg_dir_timeMarchingD.m:5-8). */
int ebem_create(ebem_handle **h, int device);
int ebem_destroy(ebem_handle *h);

/* Mesh: replaces readSpaceMesh's derived data (util/fileRead/readSpaceMesh.m:93-139);
 * normals, areas, centres and longest edges are derived here with the same
 * expressions.  coords Nv x 3 row-major; tri N x 3. */
int ebem_set_mesh(ebem_handle *h, int N, int Nv, const double *coords, const int32_t *tri, int one_based);

/* Problem: pbParam of util/fileRead/readInputFile.m:36-70
 * (c_P = sqrt((lambda+2mu)/rho), c_S = sqrt(mu/rho), dt = Tfin/nT). */
int ebem_set_problem(ebem_handle *h, double rho, double mu, double lambda, double Tfin, int nT);

/* Outer Gauss-Hammer rule of methodInfo "SA nn" (util/setup/BEMenerg_setupCore.m:9-21,
 * GaussHammerComposite(1, nn)); nn in {3, 7, 12, 19}. */
int ebem_set_rule(ebem_handle *h, int nGH);

/* Inner rules of the numeric form EBEM_PREC_GHC, method "FN nn nSub nSing nDiag"
 * of util/setup/BEMenerg_setupCore.m:58-93: composite Gauss-Hammer
 * GaussHammerComposite(nSub, nSing) for pairs i != j (kernelV.cu, launched by
 * dir/BEMenerg_dir_timeMarchingD.m:97-161) and the collapsed rule
 * doppioGauss1D(nDiag) on the three children of the diagonal blocks
 * (dir/BEMenerg_dir_calcSubBlockDiagV.m).  nSub in {1,4,16,64}, nSing in
 * {1,3,7,12,19}, nDiag in {16,64,256,1024}. */
int ebem_set_inner_rule(ebem_handle *h, int nSub, int nSing, int nDiag);

/* The two inner rules as tables (GHCn/GHCw, DIAGn/DIAGw outputs of
 * BEMenerg_setupCore.m): inner (nSub*nSing) x 4 and diag nDiag x 4 row-major,
 * columns = three barycentric coordinates and the weight.  Either may be NULL. */
int ebem_inner_rule_tables(int nSub, int nSing, int nDiag, double *inner, double *diag);

/* Number of lag blocks L = min(nT, ceil(maxVertexDist/(c_S dt)) + 1)
 * (dir/BEMenerg_dir_calcNumBlocks.m:1-17). */
int ebem_num_blocks(ebem_handle *h, int *L);

/* Assemble E_0..E_{L-1} into the device-resident block store.  Replaces the
 * SA block loop (SA/BEMenerg_coreSA_timeMarching.m:39,89 ->
 * SAc/BEMenerg_coreSA_calcMatrixBlock.m) and the kernelV launches of
 * dir/BEMenerg_dir_timeMarchingD.m:97-161.  precision: EBEM_PREC_*. */
int ebem_assemble(ebem_handle *h, int precision);

/* Block-store statistics: stored 3x3 blocks (all lags), the pair.lags that
 * pass the reference's light-cone test (calcMatrixBlock.m:57-61), (pair,
 * level) integral items evaluated, bytes of the store.  The store leaves out
 * the lags of a pair whose block is zero in exact arithmetic (the field
 * triangle inside every outer node's S disk at the lag's earliest level, or
 * out of reach of the P front at its latest), so nnzb < pair_lags. */
int ebem_store_info(ebem_handle *h, int64_t *nnzb, int64_t *pair_lags, int64_t *items, int64_t *bytes);

/* Stored blocks per lag: counts[0..L-1] (L of ebem_num_blocks). */
int ebem_store_lag_blocks(ebem_handle *h, int64_t *counts);

/* Items of the last ebem_assemble: (pair, level) integral items evaluated in
 * FP32 (far field, EBEM_PREC_MIXED) and in FP64.  No reference counterpart
 * (the reference has no mixed precision); reports the split of the work. */
int ebem_assembly_stats(ebem_handle *h, int64_t *far_items, int64_t *fp64_items);

/* Level ranges of the last ebem_assemble evaluated per pair rather than per
 * level.  kP ranges: every outer node in class kP of
 * calcMatrixSubBlock.m:58-81 (S sphere short of the field plane, P sphere
 * past the whole field triangle), where W_m = WP + t_m^2 WQ from one
 * evaluation of the t-independent region coefficients.  Between-front ranges
 * (after them, S sphere still short of the field triangle): the region terms
 * that do not depend on the level are evaluated once per few levels and
 * finished per level with the reference's expressions (the same values as
 * the level-by-level evaluation).  out5 = {kP ranges, levels in kP ranges,
 * between-front ranges, levels in them, ranges that fell back to the
 * level-by-level evaluation (0 expected)}.  No reference counterpart. */
int ebem_assembly_kp_stats(ebem_handle *h, int64_t *out5);

/* The same for the last ebem_pp_eval (analytic form), over the ranges of m
 * of each (point, field triangle) pair.  No reference counterpart. */
int ebem_pp_kp_stats(ebem_handle *h, int64_t *out5);

/* Export E_lag as a dense 3N x 3N column-major host matrix (the `matrix` of
 * calcMatrixBlock.m / matrixSavedV{lag+1} of timeMarchingD.m:141-158). */
int ebem_get_block(ebem_handle *h, int lag, double *out);

/* Test rows [r0, r1) of E_lag: dense 3(r1-r0) x 3N column-major (the rows of
 * the `matrix` of calcMatrixBlock.m that belong to those test triangles). */
int ebem_get_block_rows(ebem_handle *h, int lag, int r0, int r1, double *out);

/* Right-hand side beta (3N x nT col-major) on the device and, if beta_out is
 * non-NULL, copied to the host (Tn/BEMenerg_core_calcTnBlock.m:27-58). */
int ebem_rhs(ebem_handle *h, int datum, double *beta_out);

/* ---- double-layer operator K of the direct Dirichlet driver ----------------
 * K_lag (3N x 3Nv) couples test triangle S with the P1 hat function of
 * vertex s.  Needs ebem_set_inner_rule (the FN method's nSub, nSing, nDiag),
 * as the reference's kernelK.cu does. */

/* numBlocksK = min(L + 1, nT) (dir/BEMenerg_dir_timeMarchingD.m:176). */
int ebem_k_num_blocks(ebem_handle *h, int *LK);

/* K_0 .. K_{LK-1} on the device: regular blocks as dir/CUDAkernels/kernelK.cu
 * (isBlockNull skip, 1e-14 clean-up, /(4 pi dt)), blocks of a triangle
 * against its own vertices plus dir/BEMenerg_dir_calcSubBlockSingK.m +
 * calcKIntgSing.m (timeMarchingD.m:173-274, BEMenerg_dir_launchKernelK.m). */
int ebem_assemble_k(ebem_handle *h);

/* Store of the last ebem_assemble_k: (test triangle, vertex) pairs with a
 * nonempty lag window, (pair, lag) slots, device bytes. */
int ebem_k_store_info(ebem_handle *h, int64_t *pairs, int64_t *slots, int64_t *bytes);

/* Test rows [r0, r1) of K_lag: dense 3(r1-r0) x 3Nv column-major (the rows of
 * matrixSavedK{lag+1}, timeMarchingD.m:244-258; lags >= LK are zero). */
int ebem_get_k_block_rows(ebem_handle *h, int lag, int r0, int r1, double *out);

/* Dirichlet datum at the vertices, 3Nv x nT column-major, column n at time
 * (n+1) dt (dir/BEMenerg_dir_calcBoundDataDirichlet.m); datum 4..7. */
int ebem_dirichlet_datum(ebem_handle *h, int datum, double *g_out);

/* beta_K (3N x nT column-major): beta_n = sum_{j<=n} K_{n-j} g_j for a vertex
 * datum g (3Nv x nT column-major) (dir/BEMenerg_dir_calcBetaKgpu.m:7-13). */
int ebem_beta_k(ebem_handle *h, const double *g, double *beta_k_out);

/* Right-hand side of the direct Dirichlet driver on the device,
 * beta = rhs(datum) / 2 + beta_K(g(datum)) (timeMarchingD.m:163-166, 272, 287),
 * for ebem_mot_solve(h, NULL, ...); copied to beta_out if non-NULL. */
int ebem_rhs_dirichlet(ebem_handle *h, int datum, double *beta_out);

/* Direct post-processing (dirPP/BEMenerg_ppD_calc.m:9-123, BEMenerg_ppD_setup.m):
 * u(x_p, t_k) = sum_{y < nHat} G_y alpha_y - P_y g_y, nHat = ceil(t_k / dt), with
 * G_y as dirPP/CUDAkernels/kernelG.cu and P_y as kernelP.cu over the rule of
 * ebem_set_pp_rule (numSubRegionPP x numNodiSingPP); u = 0 for t_k <= 0, an
 * error for t_k past Tfin.  density: host 3N x nT or NULL for the device
 * density of ebem_mot_solve; g: host vertex datum 3Nv x nT (column y at
 * (y+1) dt) or NULL for the one of the last ebem_rhs_dirichlet.  x: npts x 3
 * row-major; u_out: 3 x npts x nt (u_out[3*(p + npts*k) + c]). */
int ebem_pp_eval_direct(ebem_handle *h, const double *density, const double *g, int npts, const double *x, int nt,
                        const double *t, double *u_out);

/* ---- Neumann driver (dir/BEMenerg_dir_timeMarchingN.m) -------------------- */

/* Traction datum at the triangle centres, 3N x nT column-major, column n at
 * time (n+1) dt: (0, 0, 2 mu tildeP n_z) of dir/BEMenerg_dir_calcBoundDataNeumann.m
 * (datum 4 and 5, the bars). */
int ebem_neumann_datum(ebem_handle *h, int datum, double *gV_out);

/* beta_V (3N x nT column-major): beta_n = sum_{j<=n} V_{n-j} gV_j over the
 * assembled single-layer blocks (dir/BEMenerg_dir_calcBetaVgpu.m:7-13). */
int ebem_beta_v(ebem_handle *h, const double *gV, double *beta_v_out);

/* Density (3Nv x nT column-major, at the vertices) of the Neumann driver:
 * (K_0 + I_Gamma) d_n = beta_V,n + I_Gamma d_{n-1} - sum_{l=1}^{min(n,LK-1)} K_l d_{n-l},
 * I_Gamma = kron((indSM > 0) .* area / 6, eye(3)), solved in the least-squares
 * sense as MATLAB's backslash does for the rectangular system
 * (timeMarchingN.m:170-296).  Needs ebem_assemble (V; the reference uses the
 * numeric form EBEM_PREC_GHC) and ebem_assemble_k; kept on the device for
 * ebem_pp_eval_direct (the Neumann post-processing passes the datum as
 * `density` and this density as `g`, BEMenerg_ppN_calc.m:121). */
int ebem_mot_solve_neumann(ebem_handle *h, int datum, double *density_out);

/* MOT: factorise E_0 in FP64 on the GPU and march
 * alpha_n = E_0^{-1}(beta_n - sum_{l=1}^{min(n,L)-1} E_l alpha_{n-l})
 * (SA/BEMenerg_coreSA_timeMarching.m:45-113, FN/BEMenerg_coreFN_timeMarching.m:152-171).
 * beta: host 3N x nT col-major, or NULL to use the device RHS of ebem_rhs.
 * alpha_out: host 3N x nT col-major, or NULL to keep the density on the device. */
int ebem_mot_solve(ebem_handle *h, const double *beta, double *alpha_out);

/* Post-processing u(x_p, t_k) for npts points and nt times (analytic, ppA/
 * BEMenerg_postProcA_singlePoint.m; signature role of BEMenerg_ppD_setup,
 * dirPP/BEMenerg_ppD_setup.m:1).  density: host 3N x nT col-major or NULL for
 * the device density of ebem_mot_solve.  x: npts x 3 row-major; t: nt times.
 * u_out: 3 x npts x nt (u_out[3*(p + npts*k) + c]). */
int ebem_pp_eval(ebem_handle *h, const double *density, int npts, const double *x, int nt, const double *t,
                 double *u_out);

/* No counterpart in the reference (its post-processing runs after the
 * marching).  For callers that post-process at points and times known before
 * the density: starts the part of ebem_pp_eval(x, t) that does not depend on
 * the density (point windows, items, the M_m integrals) on a stream and host
 * thread of its own and returns at once, so that it runs while the caller's
 * next calls -- ebem_rhs, ebem_mot_solve -- run on the library's stream.
 * The next ebem_pp_eval with the same points and times (bit for bit) does
 * only the rest; any other call on the handle waits for the prepared part,
 * and every call but ebem_rhs, ebem_mot_solve, ebem_pp_eval and the
 * statistics and timing queries discards it.  Requests of more than one
 * group of times (t - ceil(t/dt - 1) dt differing by more than 1e-12 dt) or
 * more than one chunk of points are not prepared (ebem_pp_eval then does all
 * of it).  The results are the same bit for bit either way. */
int ebem_pp_prepare(ebem_handle *h, int npts, const double *x, int nt, const double *t);

/* Numeric post-processing of the reference's GPU path
 * (BEMenerg_postProcGPU_setupLayer.m / _calcLayer.m, CUDAkernels/kernelLayer.cu;
 * kernelSP.cu is the same sum for one point): the M(x, tau) of each triangle
 * by the composite rule GaussHammerComposite(nSub, nSing) of methodInfo
 * numSubRegionPP / numNodiSingPP, then the same time sum and layout as
 * ebem_pp_eval. */
int ebem_set_pp_rule(ebem_handle *h, int nSub, int nSing);
int ebem_pp_eval_numeric(ebem_handle *h, const double *density, int npts, const double *x, int nt, const double *t,
                         double *u_out);

/* Unit hook for the analytic inner integration (functions/BEMenerg_analiticIntg via calcMatrixSubBlock.m):
 * the kernel-combined, /(4 pi rho)-scaled 3x3 integral (column-major) over
 * triangle `tri` of the mesh for outer point x at time t. */
int ebem_node_integral(ebem_handle *h, const double *x, int tri, double t, double *out9);

/* ---- analytic region integrals (functions/BEMenerg_analiticIntg/) ----------
 * The six closed-form region routines that the reference's child-triangle
 * integration calls (SAc/BEMenerg_coreSA_calcIntgInt_childTriangle.m:126-234),
 * evaluated on the handle's device with the same code the assembly kernels
 * inline.  infoTri2D: the twelve fields of calcoloDatiTriangolo.m:10-32 in the
 * order a, b, c, cos_alpha, cos_beta, cos_gamma, sin_alpha, sin_beta,
 * sin_gamma, alpha, beta, gamma.  B: the six 3x3 matrices B(:,:,1..6) of
 * AI/coefficientiB/BEMenerg_coreSA_coeffB.m, column-major (54 doubles).
 * Outputs column-major 3x3: ris1..ris4 (36 doubles) for the P-front routines,
 * ris5, ris6 (18 doubles) for the S-front ones.  The handle needs no mesh. */
#define EBEM_AI_CC 0
#define EBEM_AI_SC_P 1
#define EBEM_AI_SC_SP 2
#define EBEM_AI_TC 3
#define EBEM_AI_TR_P 4
#define EBEM_AI_TR_SP 5

/* [ris1, ris2, ris3, ris4] = BEMenerg_analiticIntg_CC(RS, rhoS, RP, rhoP, zeta, THETA1, THETA2, infoTri2D, B)
 * (AI/BEMenerg_analiticIntg_CC.m:1): P annulus between the S and P circles. */
int ebem_analiticIntg_CC(ebem_handle *h, double RS, double rhoS, double RP, double rhoP, double zeta, double THETA1,
                         double THETA2, const double *infoTri2D, const double *B, double *ris1to4);
/* [ris1, ris2, ris3, ris4] = BEMenerg_analiticIntg_SC_P(RP, rhoP, zeta, THETA1, THETA2, infoTri2D, B)
 * (AI/BEMenerg_analiticIntg_SC_P.m:1): P sector from the plane foot. */
int ebem_analiticIntg_SC_P(ebem_handle *h, double RP, double rhoP, double zeta, double THETA1, double THETA2,
                           const double *infoTri2D, const double *B, double *ris1to4);
/* [ris5, ris6] = BEMenerg_analiticIntg_SC_SP(RS, rhoS, zeta, THETA1, THETA2, infoTri2D, B)
 * (AI/BEMenerg_analiticIntg_SC_SP.m:1): S disk sector. */
int ebem_analiticIntg_SC_SP(ebem_handle *h, double RS, double rhoS, double zeta, double THETA1, double THETA2,
                            const double *infoTri2D, const double *B, double *ris56);
/* [ris1, ris2, ris3, ris4] = BEMenerg_analiticIntg_TC(zeta, RS, rhoS, THETA1, THETA2, infoTri2D, B)
 * (AI/BEMenerg_analiticIntg_TC.m:1): P region between the S circle and the edge. */
int ebem_analiticIntg_TC(ebem_handle *h, double zeta, double RS, double rhoS, double THETA1, double THETA2,
                         const double *infoTri2D, const double *B, double *ris1to4);
/* [ris1, ris2, ris3, ris4] = BEMenerg_analiticIntg_TR_P(zeta, THETA1, THETA2, infoTri2D, B)
 * (AI/BEMenerg_analiticIntg_TR_P.m:1): P region between the plane foot and the edge. */
int ebem_analiticIntg_TR_P(ebem_handle *h, double zeta, double THETA1, double THETA2, const double *infoTri2D,
                           const double *B, double *ris1to4);
/* [ris5, ris6] = BEMenerg_analiticIntg_TR_SP(zeta, THETA1, THETA2, infoTri2D, B)
 * (AI/BEMenerg_analiticIntg_TR_SP.m:1): S region bounded by the edge. */
int ebem_analiticIntg_TR_SP(ebem_handle *h, double zeta, double THETA1, double THETA2, const double *infoTri2D,
                            const double *B, double *ris56);
/* Batch form of the six: n calls, kind[k] = EBEM_AI_*, args[8k..8k+7] =
 * {RS, rhoS, RP, rhoP, zeta, THETA1, THETA2, unused} (each routine reads the
 * ones in its argument list), info[12k..], B[54k..], ris[36k..] (ris1..ris4,
 * or ris5, ris6 followed by 18 zeros).  Replaces the 13 call sites of
 * calcIntgInt_childTriangle.m:129-222 for callers that batch them. */
int ebem_analitic_intg(ebem_handle *h, int n, const int32_t *kind, const double *args, const double *info,
                       const double *B, double *ris);

/* The same batch without any device allocation: no counterpart in the
 * reference.  For callers that cannot allocate device memory (code under HIP
 * graph capture, tight serving loops, harnesses that forbid hipMalloc).  The
 * inputs and results are staged in a fixed array that each GPU holds for
 * the life of the process (EBEM_AI_NOALLOC_MAX_CALLS calls, 7.2 MB); larger
 * batches are rejected with EBEM_EINVAL.  Calls on one GPU share that array:
 * the library serialises them (a per-device lock), so calls that overlap on
 * one GPU wait for each other instead of racing; the array keeps the last
 * call's contents.  Same results as ebem_analitic_intg, bit for bit. */
#define EBEM_AI_NOALLOC_MAX_CALLS 8192
int ebem_analitic_intg_noalloc(ebem_handle *h, int n, const int32_t *kind, const double *args, const double *info,
                               const double *B, double *ris);

/* Options of a handle.  No reference counterpart (the reference has neither
 * form these select).
 * EBEM_OPT_FAR_FACTOR: the mixed-precision form (EBEM_PREC_MIXED) evaluates in
 *   FP32 the pairs whose centroid distance exceeds value x (longest edge of the
 *   mesh); default 3; value <= 0 evaluates every pair in FP64.
 * EBEM_OPT_SURROGATE_SOLVE: value != 0 lets ebem_mot_solve run on a store
 *   assembled with EBEM_PREC_SURROGATE; without it that call fails, because
 *   the network's block errors are amplified by the time marching.
 * EBEM_OPT_SURROGATE_PROJECT: value != 0 (default) keeps, for every pair the
 *   surrogate predicts, E_0 and the pair's static level on the analytic
 *   routine and adjusts the predicted lags so that the pair's blocks sum to
 *   zero over its lags and their first moment equals minus its static
 *   integral, as the exact blocks do; 0 stores the network's blocks as they
 *   come (E_0 still analytic).  Without the projection the marching diverges. */
#define EBEM_OPT_FAR_FACTOR 1
#define EBEM_OPT_SURROGATE_SOLVE 2
#define EBEM_OPT_SURROGATE_PROJECT 3
int ebem_set_option(ebem_handle *h, int option, double value);

/* Timing of the last ebem_assemble / ebem_mot_solve / ebem_pp_eval call, in
 * milliseconds, from HIP events recorded on the library's stream; writes the
 * first min(n, EBEM_NUM_TIMINGS) of:
 * out[0] assembly integral kernel (per-level items), out[1] assembly total,
 * out[2] E_0 factorisation, out[3] history-sum kernel total, out[4] E_0 apply
 * total, out[5] MOT total, out[6] post-processing total, out[7] history-sum
 * launches, out[8] assembly kP-range kernel, out[9] post-processing kP-range
 * kernel, out[10] assembly between-front kernel, out[11] post-processing
 * between-front kernel, out[12] ebem_assemble_k total, out[13] beta_K kernel
 * of the last ebem_beta_k / ebem_rhs_dirichlet, out[14] ebem_pp_eval_direct
 * total, out[15] ebem_mot_solve_neumann total; counts of the last
 * ebem_assemble: out[16] outer-node tasks evaluated, out[17] tasks whose
 * static value was reused from an earlier level, out[18] row chunks
 * evaluated again without reuse after a failed check of the reused values
 * (0 expected); out[19] the part of the last ebem_pp_eval prepared by
 * ebem_pp_prepare, on its own stream (0 if none; out[6] is then the rest). */
#define EBEM_NUM_TIMINGS 20
int ebem_timings(ebem_handle *h, double *out, int n);

/* ------------------------------------------------------------------------
 * Learned near-field surrogate.  No reference counterpart: the reference
 * evaluates every pair analytically (functions/BEMenerg_analiticIntg/ via
 * SAc/BEMenerg_coreSA_calcMatrixSubBlock.m).  A surrogate is an MLP
 * 24 -> 256 -> 256 -> 256 -> 256 -> 18 (GELU, BF16 weights and activations,
 * FP32 accumulation, FP32 master weights) that predicts, for a test/field
 * pair and a lag, the P-front and S-front parts of the 3x3 block E_lag in the
 * field triangle's frame, divided by A_S A_F / (4 pi rho h).
 * Features (all lengths / h, h = the mesh's longest edge, frame of the field
 * triangle: origin at its centroid, e1 along its first edge, e3 its normal):
 * the test triangle's vertices (9), the field triangle's vertices (9),
 * dt c_P / h, dt c_S / h, lag, (lag dt c_P - d) / h and (lag dt c_S - d) / h
 * (d the centroid distance), cosine between the normals.
 * Labels come from the library's FP64 analytic routine; training, inference
 * and gradients run on the device with hand-written BF16 MFMA kernels.
 * ------------------------------------------------------------------------ */

/* Number of FP32 parameters (padded layout: per layer W (K x N row-major)
 * then b (N), K = 64, 256, 256, 256, 256 and N = 256, 256, 256, 256, 64;
 * rows/columns beyond 24 inputs and 18 outputs stay zero) and of the
 * normalisation block (x mean 24, x 1/std 24, y mean 18, y std 18,
 * training box lo 24, hi 24). */
#define EBEM_SUR_PARAMS 230464
#define EBEM_SUR_NORM 132

int ebem_sur_create(ebem_sur **s, int device);
/* Fails with EBEM_EINVAL while the surrogate is attached to a handle. */
int ebem_sur_destroy(ebem_sur *s);

/* Append n samples (test triangle, field triangle, lag) of h's mesh and
 * problem: features and FP64 analytic labels computed on the device. */
int ebem_sur_add_samples(ebem_sur *s, ebem_handle *h, int64_t n, const int32_t *test, const int32_t *field,
                         const int32_t *lag);
int ebem_sur_num_samples(ebem_sur *s, int64_t *n);
/* Copy samples [first, first+n): x n x 24, y n x 18 (either may be NULL). */
int ebem_sur_get_samples(ebem_sur *s, int64_t first, int64_t n, float *x, float *y);

/* Normalisation and training box from samples [0, ntrain), He-normal weights. */
int ebem_sur_init(ebem_sur *s, int64_t ntrain, uint64_t seed);
/* Adam (0.9, 0.999), cosine learning-rate decay lr -> lr/100, MSE on
 * standardised labels over samples [0, ntrain), shuffled per epoch (seed);
 * batch a multiple of 64.  loss_out[epochs]: mean training loss per epoch. */
int ebem_sur_train(ebem_sur *s, int64_t ntrain, int epochs, int batch, double lr, uint64_t seed, double *loss_out);
/* Mean standardised MSE over samples [first, first+n). */
int ebem_sur_loss(ebem_sur *s, int64_t first, int64_t n, double *loss);
/* One forward + backward pass on n <= 65536 raw samples (x n x 24, y n x 18):
 * gradient of the MSE w.r.t. every parameter (EBEM_SUR_PARAMS floats). */
int ebem_sur_grad(ebem_sur *s, int n, const float *x, const float *y, float *grads, double *loss);
/* Predicted labels (n x 18, unstandardised) of raw features x (n x 24). */
int ebem_sur_predict(ebem_sur *s, int64_t n, const float *x, float *y);
/* set = 0: copy parameters / normalisation out (either may be NULL);
 * set = 1: load both (resets Adam). */
int ebem_sur_params(ebem_sur *s, int set, float *params, float *norm);

/* Attach (s) or detach (NULL) the surrogate used by ebem_assemble with
 * EBEM_PREC_SURROGATE: pairs i != j with centroid distance <= 3h whose features
 * over their whole lag window lie inside the training box are predicted; all
 * other pairs fall back to the analytic routine. */
int ebem_attach_surrogate(ebem_handle *h, ebem_sur *s);
/* Last assembly: surrogate items, predicted blocks, milliseconds. */
int ebem_surrogate_stats(ebem_handle *h, int64_t *items, int64_t *blocks, double *ms);

#ifdef __cplusplus
}
#endif
#endif
