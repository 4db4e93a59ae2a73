/*
 * ebem_mex.c -- MATLAB MEX gateway over the C ABI of include/ebem.h.
 *
 * Build inside MATLAB (not possible in this repository's CI, which has no MATLAB):
 *   mex -R2018a -I../../include ebem_mex.c -L.. -lebem_hip
 *
 * One persistent handle per MATLAB session (the reference keeps one gpuDevice,
 * functions/BEMenerg_dir/BEMenerg_dir_timeMarchingD.m:5-8).  Usage:
 *   ebem_mex('init', device)
 *   ebem_mex('mesh', coordinates(Nv x 3), triangles(N x 3 or N x 4, 1-based))
 *   ebem_mex('problem', rho, mu, lambda, Tfin, nT)
 *   ebem_mex('rule', numNodiExt)
 *   ebem_mex('inner_rule', numSubRegion, numNodiSing, numNodiDiag)   % method "FN ..."
 *   ebem_mex('pp_rule', numSubRegionPP, numNodiSingPP)
 *   L       = ebem_mex('numblocks')
 *   ebem_mex('assemble', form)                   % 0 FP64, 1 mixed, +2 surrogate, 4 numeric (FN)
 *   E       = ebem_mex('block', lag)              % dense 3N x 3N
 *   beta    = ebem_mex('rhs', datum)              % 3N x nT
 *   density = ebem_mex('mot', beta_or_empty)      % 3N x nT
 *   u       = ebem_mex('pp', density_or_empty, xVal(k x 3), tVal)          % 3 x k x numel(tVal)
 *   u       = ebem_mex('pp_numeric', density_or_empty, xVal(k x 3), tVal)  % numeric GPU post-processing
 *   ebem_mex('free')
 * Every non-zero return code of the library becomes mexErrMsgIdAndTxt.
 */
#include <string.h>

#include "ebem.h"
#include "mex.h"

static ebem_handle *g_h = NULL;
static mwSize g_n3 = 0, g_nT = 0, g_nv = 0; /* 3N and Nv of the mesh, nT of the problem */

static void chk(int rc) {
    if (rc != 0) mexErrMsgIdAndTxt("ebem:lib", "%s (code %d)", ebem_last_error(), rc);
}

static void cleanup(void) {
    if (g_h) ebem_destroy(g_h);
    g_h = NULL;
}

static double scal(const mxArray *a) { return mxGetScalar(a); }

void mexFunction(int nlhs, mxArray *plhs[], int nrhs, const mxArray *prhs[]) {
    char cmd[32];
    if (nrhs < 1 || mxGetString(prhs[0], cmd, sizeof cmd)) mexErrMsgIdAndTxt("ebem:args", "first argument: command");
    if (!strcmp(cmd, "init")) {
        cleanup();
        chk(ebem_create(&g_h, nrhs > 1 ? (int)scal(prhs[1]) : 0));
        mexAtExit(cleanup);
        return;
    }
    if (!strcmp(cmd, "free")) {
        cleanup();
        return;
    }
    if (!g_h) mexErrMsgIdAndTxt("ebem:state", "call ebem_mex('init') first");
    if (!strcmp(cmd, "mesh")) {
        const mxArray *X = prhs[1], *T = prhs[2];
        mwSize Nv = mxGetM(X), N = mxGetM(T);
        double *x = mxGetDoubles(X), *t = mxGetDoubles(T);
        double *xr = (double *)mxMalloc(sizeof(double) * 3 * Nv);
        int32_t *tr = (int32_t *)mxMalloc(sizeof(int32_t) * 3 * N);
        for (mwSize v = 0; v < Nv; v++)
            for (int k = 0; k < 3; k++) xr[3 * v + k] = x[v + Nv * k]; /* column-major -> row-major */
        for (mwSize e = 0; e < N; e++)
            for (int k = 0; k < 3; k++) tr[3 * e + k] = (int32_t)t[e + N * k];
        int rc = ebem_set_mesh(g_h, (int)N, (int)Nv, xr, tr, 1);
        mxFree(xr);
        mxFree(tr);
        chk(rc);
        g_n3 = 3 * N;
        g_nv = Nv;
    } else if (!strcmp(cmd, "problem")) {
        chk(ebem_set_problem(g_h, scal(prhs[1]), scal(prhs[2]), scal(prhs[3]), scal(prhs[4]), (int)scal(prhs[5])));
        g_nT = (mwSize)scal(prhs[5]);
    } else if (!strcmp(cmd, "rule")) {
        chk(ebem_set_rule(g_h, (int)scal(prhs[1])));
    } else if (!strcmp(cmd, "inner_rule")) {
        chk(ebem_set_inner_rule(g_h, (int)scal(prhs[1]), (int)scal(prhs[2]), (int)scal(prhs[3])));
    } else if (!strcmp(cmd, "pp_rule")) {
        chk(ebem_set_pp_rule(g_h, (int)scal(prhs[1]), (int)scal(prhs[2])));
    } else if (!strcmp(cmd, "numblocks")) {
        int L = 0;
        chk(ebem_num_blocks(g_h, &L));
        plhs[0] = mxCreateDoubleScalar(L);
    } else if (!strcmp(cmd, "assemble")) {
        chk(ebem_assemble(g_h, nrhs > 1 ? (int)scal(prhs[1]) : EBEM_PREC_FP64));
    } else if (!strcmp(cmd, "block")) {
        plhs[0] = mxCreateDoubleMatrix(g_n3, g_n3, mxREAL);
        chk(ebem_get_block(g_h, (int)scal(prhs[1]), mxGetDoubles(plhs[0])));
    } else if (!strcmp(cmd, "rhs")) {
        plhs[0] = mxCreateDoubleMatrix(g_n3, g_nT, mxREAL);
        chk(ebem_rhs(g_h, (int)scal(prhs[1]), mxGetDoubles(plhs[0])));
    } else if (!strcmp(cmd, "assemble_k")) {
        chk(ebem_assemble_k(g_h));
    } else if (!strcmp(cmd, "kblock")) {
        plhs[0] = mxCreateDoubleMatrix(g_n3, 3 * g_nv, mxREAL);
        chk(ebem_get_k_block_rows(g_h, (int)scal(prhs[1]), 0, (int)(g_n3 / 3), mxGetDoubles(plhs[0])));
    } else if (!strcmp(cmd, "rhs_dirichlet")) {
        plhs[0] = mxCreateDoubleMatrix(g_n3, g_nT, mxREAL);
        chk(ebem_rhs_dirichlet(g_h, (int)scal(prhs[1]), mxGetDoubles(plhs[0])));
    } else if (!strcmp(cmd, "pp_direct")) {
        /* ebem_mex('pp_direct', density_or_[], g_or_[], X (k x 3), t) */
        const double *dens = mxIsEmpty(prhs[1]) ? NULL : mxGetDoubles(prhs[1]);
        const double *gk = mxIsEmpty(prhs[2]) ? NULL : mxGetDoubles(prhs[2]);
        mwSize k = mxGetM(prhs[3]), nt = mxGetNumberOfElements(prhs[4]);
        double *xv = mxGetDoubles(prhs[3]);
        double *xr = (double *)mxMalloc(sizeof(double) * 3 * k);
        for (mwSize p = 0; p < k; p++)
            for (int c = 0; c < 3; c++) xr[3 * p + c] = xv[p + k * c];
        mwSize dims[3] = {3, k, nt};
        plhs[0] = mxCreateNumericArray(3, dims, mxDOUBLE_CLASS, mxREAL);
        int rc = ebem_pp_eval_direct(g_h, dens, gk, (int)k, xr, (int)nt, mxGetDoubles(prhs[4]), mxGetDoubles(plhs[0]));
        mxFree(xr);
        chk(rc);
    } else if (!strcmp(cmd, "mot_neumann")) {
        plhs[0] = mxCreateDoubleMatrix(3 * g_nv, g_nT, mxREAL);
        chk(ebem_mot_solve_neumann(g_h, (int)scal(prhs[1]), mxGetDoubles(plhs[0])));
    } else if (!strcmp(cmd, "mot")) {
        plhs[0] = mxCreateDoubleMatrix(g_n3, g_nT, mxREAL);
        const double *beta = mxIsEmpty(prhs[1]) ? NULL : mxGetDoubles(prhs[1]);
        chk(ebem_mot_solve(g_h, beta, mxGetDoubles(plhs[0])));
    } else if (!strcmp(cmd, "pp") || !strcmp(cmd, "pp_numeric")) {
        const double *dens = mxIsEmpty(prhs[1]) ? NULL : mxGetDoubles(prhs[1]);
        mwSize k = mxGetM(prhs[2]), nt = mxGetNumberOfElements(prhs[3]);
        double *xv = mxGetDoubles(prhs[2]);
        double *xr = (double *)mxMalloc(sizeof(double) * 3 * k);
        for (mwSize p = 0; p < k; p++)
            for (int c = 0; c < 3; c++) xr[3 * p + c] = xv[p + k * c];
        mwSize dims[3] = {3, k, nt};
        plhs[0] = mxCreateNumericArray(3, dims, mxDOUBLE_CLASS, mxREAL);
        int rc = !strcmp(cmd, "pp")
                     ? ebem_pp_eval(g_h, dens, (int)k, xr, (int)nt, mxGetDoubles(prhs[3]), mxGetDoubles(plhs[0]))
                     : ebem_pp_eval_numeric(g_h, dens, (int)k, xr, (int)nt, mxGetDoubles(prhs[3]),
                                            mxGetDoubles(plhs[0]));
        mxFree(xr);
        chk(rc);
    } else if (strcmp(cmd, "mesh") && strcmp(cmd, "problem") && strcmp(cmd, "rule") && strcmp(cmd, "inner_rule") &&
               strcmp(cmd, "pp_rule") && strcmp(cmd, "numblocks") && strcmp(cmd, "assemble") && strcmp(cmd, "block") &&
               strcmp(cmd, "rhs") && strcmp(cmd, "mot") && strcmp(cmd, "assemble_k") && strcmp(cmd, "kblock") &&
               strcmp(cmd, "rhs_dirichlet") && strcmp(cmd, "pp_direct") && strcmp(cmd, "mot_neumann")) {
        mexErrMsgIdAndTxt("ebem:args", "unknown command '%s'", cmd);
    }
    (void)nlhs;
}
