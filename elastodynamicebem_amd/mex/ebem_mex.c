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
 *   L       = ebem_mex('numblocks')
 *   ebem_mex('assemble', precision)              % 0 FP64, 1 mixed
 *   E       = ebem_mex('block', lag)              % dense 3N x 3N
 *   beta    = ebem_mex('rhs', datum)              % 3N x nT
 *   density = ebem_mex('mot', beta_or_empty)      % 3N x nT
 *   u       = ebem_mex('pp', density_or_empty, xVal(k x 3), tVal)   % 3 x k x numel(tVal)
 *   ebem_mex('free')
 * Every non-zero return code of the library becomes mexErrMsgIdAndTxt.
 */
#include <string.h>

#include "ebem.h"
#include "mex.h"

static ebem_handle *g_h = NULL;

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
    } else if (!strcmp(cmd, "problem")) {
        chk(ebem_set_problem(g_h, scal(prhs[1]), scal(prhs[2]), scal(prhs[3]), scal(prhs[4]), (int)scal(prhs[5])));
    } else if (!strcmp(cmd, "rule")) {
        chk(ebem_set_rule(g_h, (int)scal(prhs[1])));
    } else if (!strcmp(cmd, "numblocks")) {
        int L = 0;
        chk(ebem_num_blocks(g_h, &L));
        plhs[0] = mxCreateDoubleScalar(L);
    } else if (!strcmp(cmd, "assemble")) {
        chk(ebem_assemble(g_h, nrhs > 1 ? (int)scal(prhs[1]) : EBEM_PREC_FP64));
    } else if (!strcmp(cmd, "block") || !strcmp(cmd, "rhs") || !strcmp(cmd, "mot")) {
        int64_t nnzb;
        chk(ebem_store_info(g_h, &nnzb, NULL, NULL, NULL));
    }
    (void)nlhs;
    if (!strcmp(cmd, "block")) {
        /* size from the mesh: ask the library for lag 0 dims via the MOT arrays */
        mwSize n3 = (mwSize)scal(prhs[2]);
        plhs[0] = mxCreateDoubleMatrix(n3, n3, mxREAL);
        chk(ebem_get_block(g_h, (int)scal(prhs[1]), mxGetDoubles(plhs[0])));
    } else if (!strcmp(cmd, "rhs")) {
        mwSize n3 = (mwSize)scal(prhs[2]), nT = (mwSize)scal(prhs[3]);
        plhs[0] = mxCreateDoubleMatrix(n3, nT, mxREAL);
        chk(ebem_rhs(g_h, (int)scal(prhs[1]), mxGetDoubles(plhs[0])));
    } else if (!strcmp(cmd, "mot")) {
        mwSize n3 = (mwSize)scal(prhs[2]), nT = (mwSize)scal(prhs[3]);
        plhs[0] = mxCreateDoubleMatrix(n3, nT, mxREAL);
        const double *beta = mxIsEmpty(prhs[1]) ? NULL : mxGetDoubles(prhs[1]);
        chk(ebem_mot_solve(g_h, beta, mxGetDoubles(plhs[0])));
    } else if (!strcmp(cmd, "pp")) {
        const double *dens = mxIsEmpty(prhs[1]) ? NULL : mxGetDoubles(prhs[1]);
        mwSize k = mxGetM(prhs[2]), nt = mxGetNumberOfElements(prhs[3]);
        double *xv = mxGetDoubles(prhs[2]);
        double *xr = (double *)mxMalloc(sizeof(double) * 3 * k);
        for (mwSize p = 0; p < k; p++)
            for (int c = 0; c < 3; c++) xr[3 * p + c] = xv[p + k * c];
        mwSize dims[3] = {3, k, nt};
        plhs[0] = mxCreateNumericArray(3, dims, mxDOUBLE_CLASS, mxREAL);
        int rc = ebem_pp_eval(g_h, dens, (int)k, xr, (int)nt, mxGetDoubles(prhs[3]), mxGetDoubles(plhs[0]));
        mxFree(xr);
        chk(rc);
    }
}
