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
 *   ebem_mex('option', id, value)                % EBEM_OPT_*: 1 far factor, 2 surrogate solve
 *   L       = ebem_mex('numblocks')
 *   ebem_mex('assemble', form)                   % 0 FP64, 1 mixed, +2 surrogate, 4 numeric (FN)
 *   E       = ebem_mex('block', lag)              % dense 3N x 3N
 *   beta    = ebem_mex('rhs', datum)              % 3N x nT
 *   density = ebem_mex('mot', beta_or_empty)      % 3N x nT
 *   u       = ebem_mex('pp', density_or_empty, xVal(k x 3), tVal)          % 3 x k x numel(tVal)
 *   u       = ebem_mex('pp_numeric', density_or_empty, xVal(k x 3), tVal)  % numeric GPU post-processing
 *   u       = ebem_mex('pp_direct', density_or_empty, g_or_empty, xVal(k x 3), tVal)
 *   g       = ebem_mex('dirichlet_datum', datum)  % vertex datum 3Nv x nT (datum 4..7)
 *   [r1, r2, r3, r4] = ebem_mex('intg', kind, args(1 x 7), infoVec(1 x 12), B(3 x 3 x 6))
 *                                                 % kind 0..5: CC SC_P SC_SP TC TR_P TR_SP
 *   ebem_mex('free')
 * Every argument is checked (count, class, size against the mesh and problem
 * set before) before the library sees a pointer; every non-zero return code of
 * the library becomes mexErrMsgIdAndTxt.
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
    g_n3 = g_nT = g_nv = 0;
}

static void need_args(int nrhs, int n, const char *cmd) {
    if (nrhs < n) mexErrMsgIdAndTxt("ebem:args", "'%s' needs %d argument(s) after the command", cmd, n - 1);
}

/* A real, full (not sparse) double array. */
static void need_double(const mxArray *a, const char *what) {
    if (!mxIsDouble(a) || mxIsComplex(a) || mxIsSparse(a))
        mexErrMsgIdAndTxt("ebem:args", "%s must be a real full double array", what);
}

static double scal(const mxArray *a, const char *what) {
    need_double(a, what);
    if (mxGetNumberOfElements(a) != 1) mexErrMsgIdAndTxt("ebem:args", "%s must be a scalar", what);
    return mxGetScalar(a);
}

static void need_size(const mxArray *a, mwSize m, mwSize n, const char *what) {
    need_double(a, what);
    if (mxGetM(a) != m || mxGetN(a) != n)
        mexErrMsgIdAndTxt("ebem:args", "%s must be %d x %d (got %d x %d)", what, (int)m, (int)n, (int)mxGetM(a),
                          (int)mxGetN(a));
}

static void need_mesh(void) {
    if (!g_n3 || !g_nT) mexErrMsgIdAndTxt("ebem:state", "set 'mesh' and 'problem' first");
}

/* Density-shaped argument (3N x nT) or [] (NULL: the device copy). */
static const double *density_arg(const mxArray *a, const char *what) {
    if (mxIsEmpty(a)) return NULL;
    need_size(a, g_n3, g_nT, what);
    return mxGetDoubles(a);
}

/* Points k x 3 (column-major) -> row-major copy; *k set. */
static double *points_arg(const mxArray *a, mwSize *k) {
    need_double(a, "xVal");
    if (mxGetN(a) != 3 || mxGetM(a) < 1) mexErrMsgIdAndTxt("ebem:args", "xVal must be k x 3");
    *k = mxGetM(a);
    const double *xv = mxGetDoubles(a);
    double *xr = (double *)mxMalloc(sizeof(double) * 3 * (*k));
    for (mwSize p = 0; p < *k; p++)
        for (int c = 0; c < 3; c++) xr[3 * p + c] = xv[p + (*k) * c];
    return xr;
}

static const double *times_arg(const mxArray *a, mwSize *nt) {
    need_double(a, "tVal");
    *nt = mxGetNumberOfElements(a);
    if (*nt < 1) mexErrMsgIdAndTxt("ebem:args", "tVal must not be empty");
    return mxGetDoubles(a);
}

void mexFunction(int nlhs, mxArray *plhs[], int nrhs, const mxArray *prhs[]) {
    char cmd[32];
    if (nrhs < 1 || !mxIsChar(prhs[0]) || mxGetString(prhs[0], cmd, sizeof cmd))
        mexErrMsgIdAndTxt("ebem:args", "first argument: command");
    if (!strcmp(cmd, "init")) {
        cleanup();
        chk(ebem_create(&g_h, nrhs > 1 ? (int)scal(prhs[1], "device") : 0));
        mexAtExit(cleanup);
        return;
    }
    if (!strcmp(cmd, "free")) {
        cleanup();
        return;
    }
    if (!g_h) mexErrMsgIdAndTxt("ebem:state", "call ebem_mex('init') first");
    if (!strcmp(cmd, "mesh")) {
        need_args(nrhs, 3, cmd);
        const mxArray *X = prhs[1], *T = prhs[2];
        need_double(X, "coordinates");
        need_double(T, "triangles");
        mwSize Nv = mxGetM(X), N = mxGetM(T);
        if (mxGetN(X) != 3 || Nv < 3) mexErrMsgIdAndTxt("ebem:args", "coordinates must be Nv x 3");
        if ((mxGetN(T) != 3 && mxGetN(T) != 4) || N < 1) mexErrMsgIdAndTxt("ebem:args", "triangles must be N x 3 or N x 4");
        const double *x = mxGetDoubles(X), *t = mxGetDoubles(T);
        double *xr = (double *)mxMalloc(sizeof(double) * 3 * Nv);
        int32_t *tr = (int32_t *)mxMalloc(sizeof(int32_t) * 3 * N);
        for (mwSize v = 0; v < Nv; v++)
            for (int k = 0; k < 3; k++) xr[3 * v + k] = x[v + Nv * k]; /* column-major -> row-major */
        int bad = 0;
        for (mwSize e = 0; e < N; e++)
            for (int k = 0; k < 3; k++) {
                double w = t[e + N * k];
                if (!(w >= 1 && w <= (double)Nv) || w != (double)(int32_t)w) bad = 1;
                tr[3 * e + k] = bad ? 1 : (int32_t)w;
            }
        if (bad) {
            mxFree(xr);
            mxFree(tr);
            mexErrMsgIdAndTxt("ebem:args", "triangles must hold integer vertex indices in 1..Nv");
        }
        g_n3 = 0; /* a rejected mesh leaves the handle without one */
        int rc = ebem_set_mesh(g_h, (int)N, (int)Nv, xr, tr, 1);
        mxFree(xr);
        mxFree(tr);
        chk(rc);
        g_n3 = 3 * N;
        g_nv = Nv;
    } else if (!strcmp(cmd, "problem")) {
        need_args(nrhs, 6, cmd);
        double nT = scal(prhs[5], "nT");
        if (!(nT >= 1) || nT != (double)(int)nT) mexErrMsgIdAndTxt("ebem:args", "nT must be a positive integer");
        chk(ebem_set_problem(g_h, scal(prhs[1], "rho"), scal(prhs[2], "mu"), scal(prhs[3], "lambda"),
                             scal(prhs[4], "Tfin"), (int)nT));
        g_nT = (mwSize)nT;
    } else if (!strcmp(cmd, "rule")) {
        need_args(nrhs, 2, cmd);
        chk(ebem_set_rule(g_h, (int)scal(prhs[1], "numNodiExt")));
    } else if (!strcmp(cmd, "inner_rule")) {
        need_args(nrhs, 4, cmd);
        chk(ebem_set_inner_rule(g_h, (int)scal(prhs[1], "numSubRegion"), (int)scal(prhs[2], "numNodiSing"),
                                (int)scal(prhs[3], "numNodiDiag")));
    } else if (!strcmp(cmd, "pp_rule")) {
        need_args(nrhs, 3, cmd);
        chk(ebem_set_pp_rule(g_h, (int)scal(prhs[1], "numSubRegionPP"), (int)scal(prhs[2], "numNodiSingPP")));
    } else if (!strcmp(cmd, "option")) {
        need_args(nrhs, 3, cmd);
        chk(ebem_set_option(g_h, (int)scal(prhs[1], "option"), scal(prhs[2], "value")));
    } else if (!strcmp(cmd, "numblocks")) {
        int L = 0;
        chk(ebem_num_blocks(g_h, &L));
        plhs[0] = mxCreateDoubleScalar(L);
    } else if (!strcmp(cmd, "assemble")) {
        chk(ebem_assemble(g_h, nrhs > 1 ? (int)scal(prhs[1], "form") : EBEM_PREC_FP64));
    } else if (!strcmp(cmd, "block")) {
        need_args(nrhs, 2, cmd);
        need_mesh();
        double lag = scal(prhs[1], "lag");
        if (!(lag >= 0) || lag != (double)(int)lag) mexErrMsgIdAndTxt("ebem:args", "lag must be a non-negative integer");
        plhs[0] = mxCreateDoubleMatrix(g_n3, g_n3, mxREAL);
        chk(ebem_get_block(g_h, (int)lag, mxGetDoubles(plhs[0])));
    } else if (!strcmp(cmd, "rhs")) {
        need_args(nrhs, 2, cmd);
        need_mesh();
        plhs[0] = mxCreateDoubleMatrix(g_n3, g_nT, mxREAL);
        chk(ebem_rhs(g_h, (int)scal(prhs[1], "datum"), mxGetDoubles(plhs[0])));
    } else if (!strcmp(cmd, "assemble_k")) {
        chk(ebem_assemble_k(g_h));
    } else if (!strcmp(cmd, "kblock")) {
        need_args(nrhs, 2, cmd);
        need_mesh();
        plhs[0] = mxCreateDoubleMatrix(g_n3, 3 * g_nv, mxREAL);
        chk(ebem_get_k_block_rows(g_h, (int)scal(prhs[1], "lag"), 0, (int)(g_n3 / 3), mxGetDoubles(plhs[0])));
    } else if (!strcmp(cmd, "rhs_dirichlet")) {
        need_args(nrhs, 2, cmd);
        need_mesh();
        plhs[0] = mxCreateDoubleMatrix(g_n3, g_nT, mxREAL);
        chk(ebem_rhs_dirichlet(g_h, (int)scal(prhs[1], "datum"), mxGetDoubles(plhs[0])));
    } else if (!strcmp(cmd, "pp_direct")) {
        /* ebem_mex('pp_direct', density_or_[], g_or_[], X (k x 3), t) */
        need_args(nrhs, 5, cmd);
        need_mesh();
        const double *dens = density_arg(prhs[1], "density");
        const double *gk = NULL;
        if (!mxIsEmpty(prhs[2])) {
            need_size(prhs[2], 3 * g_nv, g_nT, "g");
            gk = mxGetDoubles(prhs[2]);
        }
        mwSize k = 0, nt = 0;
        const double *tv = times_arg(prhs[4], &nt);
        double *xr = points_arg(prhs[3], &k);
        mwSize dims[3] = {3, k, nt};
        plhs[0] = mxCreateNumericArray(3, dims, mxDOUBLE_CLASS, mxREAL);
        int rc = ebem_pp_eval_direct(g_h, dens, gk, (int)k, xr, (int)nt, tv, mxGetDoubles(plhs[0]));
        mxFree(xr);
        chk(rc);
    } else if (!strcmp(cmd, "dirichlet_datum")) {
        need_args(nrhs, 2, cmd);
        need_mesh();
        plhs[0] = mxCreateDoubleMatrix(3 * g_nv, g_nT, mxREAL);
        chk(ebem_dirichlet_datum(g_h, (int)scal(prhs[1], "datum"), mxGetDoubles(plhs[0])));
    } else if (!strcmp(cmd, "mot_neumann")) {
        need_args(nrhs, 2, cmd);
        need_mesh();
        plhs[0] = mxCreateDoubleMatrix(3 * g_nv, g_nT, mxREAL);
        chk(ebem_mot_solve_neumann(g_h, (int)scal(prhs[1], "datum"), mxGetDoubles(plhs[0])));
    } else if (!strcmp(cmd, "mot")) {
        need_args(nrhs, 2, cmd);
        need_mesh();
        const double *beta = density_arg(prhs[1], "beta");
        plhs[0] = mxCreateDoubleMatrix(g_n3, g_nT, mxREAL);
        chk(ebem_mot_solve(g_h, beta, mxGetDoubles(plhs[0])));
    } else if (!strcmp(cmd, "pp") || !strcmp(cmd, "pp_numeric")) {
        need_args(nrhs, 4, cmd);
        need_mesh();
        const double *dens = density_arg(prhs[1], "density");
        mwSize k = 0, nt = 0;
        const double *tv = times_arg(prhs[3], &nt);
        double *xr = points_arg(prhs[2], &k);
        mwSize dims[3] = {3, k, nt};
        plhs[0] = mxCreateNumericArray(3, dims, mxDOUBLE_CLASS, mxREAL);
        int rc = !strcmp(cmd, "pp") ? ebem_pp_eval(g_h, dens, (int)k, xr, (int)nt, tv, mxGetDoubles(plhs[0]))
                                    : ebem_pp_eval_numeric(g_h, dens, (int)k, xr, (int)nt, tv, mxGetDoubles(plhs[0]));
        mxFree(xr);
        chk(rc);
    } else if (!strcmp(cmd, "intg")) {
        /* [r1, r2, (r3, r4)] = ebem_mex('intg', kind, args(7), infoVec(12), B(3x3x6)) */
        need_args(nrhs, 5, cmd);
        double kd = scal(prhs[1], "kind");
        if (!(kd >= EBEM_AI_CC && kd <= EBEM_AI_TR_SP) || kd != (double)(int)kd)
            mexErrMsgIdAndTxt("ebem:args", "kind must be 0..5 (CC SC_P SC_SP TC TR_P TR_SP)");
        need_double(prhs[2], "args");
        need_double(prhs[3], "infoTri2D");
        need_double(prhs[4], "B");
        if (mxGetNumberOfElements(prhs[2]) != 7) mexErrMsgIdAndTxt("ebem:args", "args must hold 7 values");
        if (mxGetNumberOfElements(prhs[3]) != 12) mexErrMsgIdAndTxt("ebem:args", "infoTri2D must hold 12 values");
        if (mxGetNumberOfElements(prhs[4]) != 54) mexErrMsgIdAndTxt("ebem:args", "B must be 3 x 3 x 6");
        const int32_t kind = (int32_t)kd;
        double a[8], ris[36];
        memcpy(a, mxGetDoubles(prhs[2]), 7 * sizeof(double));
        a[7] = 0.0;
        chk(ebem_analitic_intg(g_h, 1, &kind, a, mxGetDoubles(prhs[3]), mxGetDoubles(prhs[4]), ris));
        const int nris = (kind == EBEM_AI_SC_SP || kind == EBEM_AI_TR_SP) ? 2 : 4;
        for (int r = 0; r < nris && r < (nlhs > 0 ? nlhs : 1); r++) {
            plhs[r] = mxCreateDoubleMatrix(3, 3, mxREAL);
            memcpy(mxGetDoubles(plhs[r]), ris + 9 * r, 9 * sizeof(double));
        }
    } else {
        mexErrMsgIdAndTxt("ebem:args", "unknown command '%s'", cmd);
    }
}
