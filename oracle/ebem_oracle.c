/*
 * ebem_oracle.c -- TEST INFRASTRUCTURE ONLY.
 *
 * Plain-C FP64 restatement of the reference's semi-analytic (SA) single-layer
 * path of IvanDB/ElastodynamicEBEM, written to follow the MATLAB statement by
 * statement so that it can serve as the parity oracle of the HIP library and
 * as the timed CPU baseline.  Each function cites the reference file it
 * restates (paths relative to the reference root; "SAc/" below stands for
 * functions/BEMenerg_indir/BEMenerg_core/BEMenerg_coreSA/calcoloMatrice/,
 * "geo/" for SAc/operazioniGeometriche/, "AI/" for
 * functions/BEMenerg_analiticIntg/).
 *
 * MATLAB semantics kept on purpose: sign(0) = 0, min() ignores NaN, all
 * thresholds (1e-6, 1e-10, 1e-14) with the same strict/non-strict comparison.
 * Where MATLAB would step into complex arithmetic (acos/asin/sqrt of an
 * argument pushed just past its domain by rounding) the argument is clamped,
 * which equals the real part MATLAB would carry to within rounding.
 *
 * Also restated, beyond the SA path: the numeric inner quadrature of the
 * reference's GPU path (kernelV.cu, calcSubBlockDiagV.m, kernelLayer.cu), the
 * double-layer operator K of the direct drivers (kernelK.cu,
 * calcSubBlockSingK.m, calcKIntgSing.m) and the direct post-processing
 * u = G alpha - P g (kernelG.cu, kernelP.cu).
 *
 * What pins each part independently of its own formulas (tests/test_oracle*.py):
 * SA -- Gauss-Hammer exactness, an independent polar quadrature of the node
 * integrals, closed-form disk/annulus integrals, light-cone zeros, block
 * symmetry, the RHS closed form; numeric form -- rule exactness, the kernel's
 * trace identity, agreement with SA where the integrand is smooth; K -- the P1
 * basis values, tensor covariance under rotations and reflections, late-time
 * cancellation of every term; direct post-processing -- the rigid-translation
 * (Somigliana) identity inside and outside a closed surface, tensor
 * covariance, causality; the drivers (DD, DN; NumPy in pyoracle.py) -- the 1-D
 * bar's closed-form solution.  Parity unpinned: the MATLAB runtime's sparse
 * LU / QR (UMFPACK / SuiteSparseQR), replaced by LAPACK-style solves checked
 * by residuals.
 *
 * Build with -ffp-contract=off: MATLAB evaluates scalar expressions without
 * fused multiply-add.
 */
#include "ebem_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define PI 3.141592653589793238462643383279502884

static long long g_cnt[12];
#ifdef _OPENMP
#define CNT(i) do { _Pragma("omp atomic") g_cnt[i]++; } while (0)
#else
#define CNT(i) (g_cnt[i]++)
#endif

void orc_counters(long long *out) { for (int i = 0; i < 12; i++) out[i] = g_cnt[i]; }
void orc_counters_reset(void) { memset(g_cnt, 0, sizeof g_cnt); }

static double msign(double x) { return x > 0 ? 1.0 : (x < 0 ? -1.0 : 0.0); }
static double clamp1(double x) { return x > 1.0 ? 1.0 : (x < -1.0 ? -1.0 : x); }
static double sqrt0(double x) { return x < 0 ? 0.0 : sqrt(x); }
static double dot3(const double *a, const double *b) { return a[0] * b[0] + a[1] * b[1] + a[2] * b[2]; }

typedef double mat3[3][3]; /* [row][col] */

static void m3_zero(mat3 A) { memset(A, 0, sizeof(mat3)); }
static void m3_eye(mat3 A, double s) { m3_zero(A); A[0][0] = A[1][1] = A[2][2] = s; }
static void m3_axpy(mat3 Y, double a, mat3 X) {
    for (int i = 0; i < 3; i++) for (int j = 0; j < 3; j++) Y[i][j] += a * X[i][j];
}
static void m3_mul(mat3 C, mat3 A, mat3 B) { /* C = A*B (C may not alias) */
    for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++) C[i][j] = A[i][0] * B[0][j] + A[i][1] * B[1][j] + A[i][2] * B[2][j];
}
static void m3_mul_bt(mat3 C, mat3 A, mat3 B) { /* C = A*B' */
    for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++) C[i][j] = A[i][0] * B[j][0] + A[i][1] * B[j][1] + A[i][2] * B[j][2];
}
static void m3_mul_at(mat3 C, mat3 A, mat3 B) { /* C = A'*B */
    for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++) C[i][j] = A[0][i] * B[0][j] + A[1][i] * B[1][j] + A[2][i] * B[2][j];
}

/* ------------------------------------------------------------------ mesh -- */

/* util/fileRead/readSpaceMesh.m:93-139 */
void orc_mesh_derive(int N, const double *X, const int *tri, double *normal, double *area,
                     double *center, double *maxL) {
    for (int t = 0; t < N; t++) {
        const double *P1 = X + 3 * tri[3 * t], *P2 = X + 3 * tri[3 * t + 1], *P3 = X + 3 * tri[3 * t + 2];
        double vg1[3], vg2[3], vg3[3], n[3];
        for (int k = 0; k < 3; k++) {
            vg1[k] = P1[k] - P3[k];
            vg2[k] = P2[k] - P3[k];
            vg3[k] = P1[k] - P2[k];
        }
        n[0] = vg1[1] * vg2[2] - vg2[1] * vg1[2];
        n[1] = -vg1[0] * vg2[2] + vg2[0] * vg1[2];
        n[2] = vg1[0] * vg2[1] - vg2[0] * vg1[1];
        double a = sqrt(n[0] * n[0] + n[1] * n[1] + n[2] * n[2]) / 2;
        area[t] = a;
        for (int k = 0; k < 3; k++) normal[3 * t + k] = n[k] / (2 * a);
        for (int k = 0; k < 3; k++) center[3 * t + k] = (P1[k] + P2[k] + P3[k]) / 3;
        double l1 = sqrt(vg3[0] * vg3[0] + vg3[1] * vg3[1] + vg3[2] * vg3[2]);
        double l2 = sqrt(vg2[0] * vg2[0] + vg2[1] * vg2[1] + vg2[2] * vg2[2]);
        double l3 = sqrt(vg1[0] * vg1[0] + vg1[1] * vg1[1] + vg1[2] * vg1[2]);
        double mx = l1;
        if (l2 > mx) mx = l2;
        if (l3 > mx) mx = l3;
        maxL[t] = mx;
    }
}

/* ----------------------------------------------------------- quadrature -- */

/* util/quadRules/GaussHammer_base.m:19-255 (rules 1,3,7,12,19) and
 * util/quadRules/GaussHammerComposite.m:5-52 with nSubPart = 1, whose map sends
 * barycentric (b1,b2,b3) to (b1,b3,b2) and doubles the weights. */
static int gh_base(int n, double g[28][3], double gw[28]) {
    memset(g, 0, 28 * sizeof g[0]);
    memset(gw, 0, 28 * sizeof gw[0]);
    if (n == 1) {
        g[0][0] = g[0][1] = g[0][2] = 1.0 / 3;
        gw[0] = 1.0 / 2;
    } else if (n == 3) {
        double v1 = 2.0 / 3, v2 = 1.0 / 6, v3 = 1.0 / 6;
        g[0][0] = v1; g[0][1] = v2; g[0][2] = v3;
        g[1][0] = v3; g[1][1] = v1; g[1][2] = v2;
        g[2][0] = v2; g[2][1] = v3; g[2][2] = v1;
        gw[0] = gw[1] = gw[2] = 1.0 / 6;
    } else if (n == 7) {
        const double a1[7] = {0.1012865073235, 0.7974269853531, 0.1012865073235, 0.4701420641051,
                              0.4701420641051, 0.0597158717898, 0.3333333333333};
        const double a2[7] = {0.1012865073235, 0.1012865073235, 0.7974269853531, 0.0597158717898,
                              0.4701420641051, 0.4701420641051, 0.3333333333333};
        for (int i = 0; i < 7; i++) {
            g[i][0] = a1[i];
            g[i][1] = a2[i];
            g[i][2] = 1 - g[i][0] - g[i][1];
        }
        gw[0] = gw[1] = gw[2] = 0.1259391805448 / 2;
        gw[3] = gw[4] = gw[5] = 0.1323941527885 / 2;
        gw[6] = 0.225 / 2;
    } else if (n == 12) {
        double v[3];
        v[0] = 0.873821971016996; v[1] = (1 - v[0]) / 2; v[2] = v[1];
        g[0][0] = v[0]; g[0][1] = v[1]; g[0][2] = v[2];
        g[1][0] = v[2]; g[1][1] = v[0]; g[1][2] = v[1];
        g[2][0] = v[1]; g[2][1] = v[2]; g[2][2] = v[0];
        v[0] = 0.501426509658179; v[1] = (1 - v[0]) / 2; v[2] = v[1];
        g[3][0] = v[0]; g[3][1] = v[1]; g[3][2] = v[2];
        g[4][0] = v[2]; g[4][1] = v[0]; g[4][2] = v[1];
        g[5][0] = v[1]; g[5][1] = v[2]; g[5][2] = v[0];
        v[0] = 0.636502499121399; v[1] = 0.310352451033784; v[2] = 1 - v[0] - v[1];
        const int perm[6][3] = {{0, 1, 2}, {0, 2, 1}, {1, 0, 2}, {1, 2, 0}, {2, 0, 1}, {2, 1, 0}};
        for (int k = 0; k < 6; k++)
            for (int c = 0; c < 3; c++) g[6 + k][c] = v[perm[k][c]];
        double w1 = 0.050844906370207 / 2, w2 = 0.116786275726379 / 2;
        double w3 = (0.5 - 3 * w1 - 3 * w2) / 6;
        for (int k = 0; k < 3; k++) gw[k] = w1;
        for (int k = 3; k < 6; k++) gw[k] = w2;
        for (int k = 6; k < 12; k++) gw[k] = w3;
    } else if (n == 19) {
        g[0][0] = g[0][1] = g[0][2] = 1.0 / 3;
        gw[0] = 0.09713579628279610 / 2;
        const double pa[4][2] = {{0.02063496160252593, 0.48968251919873700},
                                 {0.1258208170141290, 0.4370895914929355},
                                 {0.6235929287619356, 0.1882035356190322},
                                 {0.91054097321109410, 0.04472951339445297}};
        const double pw[4] = {0.09400410068141950 / 6, 0.2334826230143263 / 6,
                              0.2389432167816273 / 6, 0.07673302697609430 / 6};
        for (int s = 0; s < 4; s++) {
            int b = 1 + 3 * s;
            double v1 = pa[s][0], v2 = pa[s][1];
            g[b][0] = v1; g[b][1] = v2; g[b][2] = v2;
            g[b + 1][0] = v2; g[b + 1][1] = v1; g[b + 1][2] = v2;
            g[b + 2][0] = v2; g[b + 2][1] = v2; g[b + 2][2] = v1;
            gw[b] = gw[b + 1] = gw[b + 2] = pw[s];
        }
        double v[3] = {0.03683841205473626, 0.74119859878449800, 0};
        v[2] = 1 - v[0] - v[1];
        const int perm[6][3] = {{0, 1, 2}, {0, 2, 1}, {1, 0, 2}, {1, 2, 0}, {2, 0, 1}, {2, 1, 0}};
        for (int k = 0; k < 6; k++) {
            for (int c = 0; c < 3; c++) g[13 + k][c] = v[perm[k][c]];
            gw[13 + k] = 0.2597012362637364 / 12;
        }
    } else {
        return -1;
    }
    return 0;
}

int orc_gauss_hammer(int n, double *nodes, double *w) {
    double g[28][3], gw[28];
    if (gh_base(n, g, gw)) return -1;
    /* GaussHammerComposite.m:31-52, nDiv = 1: nodes = nGHstd * [[1 0 0];[0 0 1];[0 1 0]],
     * weights = 2 * wGHstd. */
    for (int i = 0; i < n; i++) {
        nodes[3 * i + 0] = g[i][0];
        nodes[3 * i + 1] = g[i][2];
        nodes[3 * i + 2] = g[i][1];
        w[i] = 2.0 / 1 * gw[i];
    }
    return 0;
}

/* dir/BEMenerg_dir_calcNumBlocks.m:1-17 */
int orc_num_blocks(const orc_param *p, int Nv, const double *X) {
    double dt = p->Tfin / p->nT;
    double mx = 0;
    for (int a = 0; a < Nv; a++)
        for (int b = 0; b < Nv; b++) {
            double d0 = X[3 * a] - X[3 * b], d1 = X[3 * a + 1] - X[3 * b + 1], d2 = X[3 * a + 2] - X[3 * b + 2];
            double d = sqrt(d0 * d0 + d1 * d1 + d2 * d2);
            if (d > mx) mx = d;
        }
    int m = (int)ceil(mx / (p->velS * dt)) + 1;
    return p->nT < m ? p->nT : m;
}

/* ------------------------------------------------------------- geometry -- */

typedef struct {
    double L1, L3, cos_d, sin_d;
    double e1[3], e2[3], e3[3];
} sdr_t;

/* geo/calcoloSistemaRiferimento.m:1-45 */
static void sdr_compute(const double *T, const double *n, sdr_t *s) {
    double vl1[3] = {T[3] - T[0], T[4] - T[1], T[5] - T[2]};
    s->L1 = sqrt(vl1[0] * vl1[0] + vl1[1] * vl1[1] + vl1[2] * vl1[2]);
    for (int k = 0; k < 3; k++) s->e1[k] = vl1[k] / s->L1;
    double vl3[3] = {T[6] - T[0], T[7] - T[1], T[8] - T[2]};
    s->L3 = sqrt(vl3[0] * vl3[0] + vl3[1] * vl3[1] + vl3[2] * vl3[2]);
    for (int k = 0; k < 3; k++) s->e3[k] = vl3[k] / s->L3;
    s->cos_d = dot3(s->e3, s->e1);
    double delta = acos(clamp1(s->cos_d));
    s->sin_d = sin(delta);
    double e2[3] = {n[1] * s->e1[2] - n[2] * s->e1[1], n[2] * s->e1[0] - n[0] * s->e1[2],
                    n[0] * s->e1[1] - n[1] * s->e1[0]};
    double ne = sqrt(e2[0] * e2[0] + e2[1] * e2[1] + e2[2] * e2[2]);
    for (int k = 0; k < 3; k++) s->e2[k] = e2[k] / ne;
}

static double nanmin3(double a, double b, double c) {
    double m = NAN;
    double v[3] = {a, b, c};
    for (int i = 0; i < 3; i++)
        if (!isnan(v[i]) && (isnan(m) || v[i] < m)) m = v[i];
    return m;
}

/* geo/calcoloDistanzaTriangolo.m:1-52 */
static double dist_tri(double t2[3][2]) {
    double D[3];
    for (int i = 0; i < 3; i++) D[i] = sqrt(t2[i][0] * t2[i][0] + t2[i][1] * t2[i][1]);
    const int e[3][2] = {{0, 1}, {1, 2}, {0, 2}};
    double dm[3];
    for (int i = 0; i < 3; i++) {
        const double *P1 = t2[e[i][0]], *P2 = t2[e[i][1]];
        double d1 = D[e[i][0]], d2 = D[e[i][1]];
        double vL[2] = {P2[0] - P1[0], P2[1] - P1[1]};
        double L = sqrt(vL[0] * vL[0] + vL[1] * vL[1]);
        if (-(P1[0] * vL[0] + P1[1] * vL[1]) < -1.0e-6) {
            dm[i] = d1;
        } else if (-(P2[0] * (-vL[0]) + P2[1] * (-vL[1])) < -1.0e-6) {
            dm[i] = d2;
        } else {
            double ca = (-P1[0] / d1) * (vL[0] / L) + (-P1[1] / d1) * (vL[1] / L);
            double sa = sqrt0(1 - ca * ca);
            dm[i] = d1 * sa;
        }
    }
    return nanmin3(dm[0], dm[1], dm[2]);
}

/* geo/preparazioneTriangolo.m:1-115.  ch[k][v][xy]: child k, vertex v (v=0 is
 * the origin). */
static void prep_triangle(const double *sp, const double *T, const double *vn, const sdr_t *s,
                          double *zeta, double ch[3][3][2], double *dMIN, double c[3],
                          double *sign_prod) {
    double d[3] = {sp[0] - T[0], sp[1] - T[1], sp[2] - T[2]};
    double param_s = dot3(vn, d);
    double spp[3];
    for (int k = 0; k < 3; k++) spp[k] = sp[k] - param_s * vn[k];
    double dd[3] = {sp[0] - spp[0], sp[1] - spp[1], sp[2] - spp[2]};
    double sgp = dot3(dd, vn);
    if (fabs(sgp) <= 1.0e-6) sgp = 0;
    *sign_prod = msign(sgp);
    double z = fabs(param_s);
    if (z < 1.0e-6) z = 0;
    *zeta = z;

    double diff[3] = {T[0] - spp[0], T[1] - spp[1], T[2] - spp[2]};
    double t2[3][2];
    t2[0][0] = dot3(diff, s->e1);
    t2[0][1] = dot3(diff, s->e2);
    t2[1][0] = t2[0][0] + s->L1;
    t2[1][1] = t2[0][1] + 0;
    t2[2][0] = t2[0][0] + s->L3 * s->cos_d;
    t2[2][1] = t2[0][1] + s->L3 * s->sin_d;

    memset(ch, 0, sizeof(double) * 18);
    for (int xy = 0; xy < 2; xy++) {
        ch[0][1][xy] = t2[0][xy]; ch[0][2][xy] = t2[1][xy];
        ch[1][1][xy] = t2[1][xy]; ch[1][2][xy] = t2[2][xy];
        ch[2][1][xy] = t2[2][xy]; ch[2][2][xy] = t2[0][xy];
    }
    c[0] = (t2[0][0] * t2[1][1]) - (t2[0][1] * t2[1][0]);
    c[1] = (t2[1][0] * t2[2][1]) - (t2[1][1] * t2[2][0]);
    c[2] = (t2[2][0] * t2[0][1]) - (t2[2][1] * t2[0][0]);

    double A2 = (t2[1][0] - t2[0][0]) * (t2[2][1] - t2[0][1]);
    double eta[3];
    eta[1] = c[2] / A2;
    eta[2] = c[0] / A2;
    eta[0] = 1 - eta[1] - eta[2];
    int in = 1;
    for (int k = 0; k < 3; k++)
        if (!(-1.0e-6 <= eta[k] && eta[k] <= 1)) in = 0;
    *dMIN = in ? 0 : dist_tri(t2);
}

typedef struct {
    double a, b, c;
    double cos_alpha, cos_beta, cos_gamma;
    double sin_alpha, sin_beta, sin_gamma;
    double alpha, beta, gamma;
} tinfo_t;

/* geo/calcoloDatiTriangolo.m:1-39 */
static void tri_data(double v[3][2], double t2[3][2], tinfo_t *f) {
    f->a = sqrt((v[1][0] - v[0][0]) * (v[1][0] - v[0][0]) + (v[1][1] - v[0][1]) * (v[1][1] - v[0][1]));
    f->b = sqrt((v[2][0] - v[0][0]) * (v[2][0] - v[0][0]) + (v[2][1] - v[0][1]) * (v[2][1] - v[0][1]));
    f->c = sqrt((v[2][0] - v[1][0]) * (v[2][0] - v[1][0]) + (v[2][1] - v[1][1]) * (v[2][1] - v[1][1]));
    double a = f->a, b = f->b, c = f->c;
    f->cos_alpha = (b * b + c * c - a * a) / (2 * b * c);
    f->cos_beta = (a * a + c * c - b * b) / (2 * a * c);
    f->cos_gamma = (a * a + b * b - c * c) / (2 * a * b);
    f->sin_alpha = sqrt0(1 - f->cos_alpha * f->cos_alpha);
    f->sin_beta = sqrt0(1 - f->cos_beta * f->cos_beta);
    f->sin_gamma = sqrt0(1 - f->cos_gamma * f->cos_gamma);
    f->alpha = acos(clamp1(f->cos_alpha));
    f->beta = acos(clamp1(f->cos_beta));
    f->gamma = acos(clamp1(f->cos_gamma));
    t2[0][0] = 0; t2[0][1] = 0;
    t2[1][0] = a; t2[1][1] = 0;
    t2[2][0] = b * f->cos_gamma;
    t2[2][1] = b * f->sin_gamma;
}

/* AI/coefficientiB/BEMenerg_coreSA_coeffB.m:1-45 */
static void coeffB(double zeta, const tinfo_t *f, double sign_prod, mat3 B[6]) {
    double A[3][3] = {{1, 0, 0}, {f->cos_gamma, f->sin_gamma, 0}, {0, 0, -sign_prod}};
    double sg = f->sin_gamma, sg2 = f->sin_gamma * f->sin_gamma;
    for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++) {
            B[0][i][j] = zeta * zeta * (A[2][i] * A[2][j]);
            B[1][i][j] = zeta * ((A[0][i] * A[2][j]) + (A[2][i] * A[0][j])) / sg;
            B[2][i][j] = zeta * ((A[1][i] * A[2][j]) + (A[2][i] * A[1][j])) / sg;
            B[3][i][j] = (A[0][i] * A[0][j]) / sg2;
            B[4][i][j] = ((A[0][i] * A[1][j]) + (A[1][i] * A[0][j])) / sg2;
            B[5][i][j] = (A[1][i] * A[1][j]) / sg2;
        }
}

/* ---------------------------------------------------- analytic integrals -- */

typedef struct {
    double p1, p2, p3, p4, p5;
} phi_t;

static phi_t phis(double T1t, double T2t, double gamma, double cos_gamma) {
    phi_t P;
    P.p1 = 2 * sin(T2t / 2) * sin(T1t / 2);
    P.p2 = 2 * sin(gamma - (T2t / 2)) * sin(T1t / 2);
    P.p3 = 1.0 / 2 * (T1t - (cos(T2t) * sin(T1t)));
    P.p4 = -1.0 / 2 * (-(sin(T1t) * cos(gamma - T2t)) + (T1t * cos_gamma));
    P.p5 = 1.0 / 2 * (T1t - (cos(2 * gamma - T2t) * sin(T1t)));
    return P;
}

/* AI/BEMenerg_analiticIntg_CC.m:1-93 */
static void intg_CC(double RS, double rhoS, double RP, double rhoP, double zeta, double TH1,
                    double TH2, const tinfo_t *f, mat3 B[6], mat3 r1, mat3 r2, mat3 r3, mat3 r4) {
    double T1t = TH2 - TH1, T2t = TH2 + TH1;
    double cos_gamma = f->cos_gamma, gamma = f->gamma;
    double I[8];
    mat3 B1, B2;
    m3_zero(B1);
    m3_zero(B2);
    if (zeta > 1.0e-6) {
        CNT(0);
        I[1] = RP - RS;
        I[2] = RP - RS + (zeta * zeta * (1 / RP - 1 / RS));
        I[3] = log((rhoP + RP) / (rhoS + RS)) - (rhoP / RP) + (rhoS / RS);
        I[4] = -(1 / RP) + (1 / RS);
        I[5] = -(1 / RP) + (1 / RS) + (zeta * zeta / 3) * ((1 / (RP * RP * RP)) - (1 / (RS * RS * RS)));
        I[6] = (1 / (3 * zeta * zeta)) * (pow(rhoP / RP, 3) - pow(rhoS / RS, 3));
        I[7] = -1.0 / 3 * ((1 / (RP * RP * RP)) - (1 / (RS * RS * RS)));
        phi_t P = phis(T1t, T2t, gamma, cos_gamma);
        m3_axpy(B1, P.p5, B[3]); m3_axpy(B1, P.p4, B[4]); m3_axpy(B1, P.p3, B[5]);
        m3_axpy(B2, P.p2, B[1]); m3_axpy(B2, P.p1, B[2]);
        m3_eye(r1, T1t * I[1]);
        m3_zero(r2);
        m3_axpy(r2, I[2], B1); m3_axpy(r2, I[3], B2); m3_axpy(r2, T1t * I[4], B[0]);
        m3_eye(r3, T1t * I[4]);
        m3_zero(r4);
        m3_axpy(r4, I[5], B1); m3_axpy(r4, I[6], B2); m3_axpy(r4, T1t * I[7], B[0]);
    } else {
        CNT(8);
        I[1] = RP - RS;
        I[2] = I[1];
        I[4] = -(1 / RP) + (1 / RS);
        I[5] = I[4];
        phi_t P = phis(T1t, T2t, gamma, cos_gamma);
        m3_axpy(B1, P.p5, B[3]); m3_axpy(B1, P.p4, B[4]); m3_axpy(B1, P.p3, B[5]);
        m3_eye(r1, T1t * I[1]);
        m3_zero(r2); m3_axpy(r2, I[2], B1);
        m3_eye(r3, T1t * I[4]);
        m3_zero(r4); m3_axpy(r4, I[5], B1);
    }
}

/* AI/BEMenerg_analiticIntg_SC_P.m:1-54 */
static void intg_SC_P(double RP, double rhoP, double zeta, double TH1, double TH2, const tinfo_t *f,
                      mat3 B[6], mat3 r1, mat3 r2, mat3 r3, mat3 r4) {
    CNT(1);
    double T1t = TH2 - TH1, T2t = TH2 + TH1;
    double I[8];
    I[1] = RP - zeta;
    I[2] = RP + (zeta * zeta / RP) - (2 * zeta);
    I[3] = log((rhoP + RP) / zeta) - (rhoP / RP);
    I[4] = -(1 / RP) + (1 / zeta);
    I[5] = -(1 / RP) + (zeta * zeta / (3 * RP * RP * RP)) + (2 / (3 * zeta));
    I[6] = (1 / (3 * zeta * zeta)) * pow(rhoP / RP, 3);
    I[7] = -(1 / (3 * RP * RP * RP)) + (1 / (3 * zeta * zeta * zeta));
    phi_t P = phis(T1t, T2t, f->gamma, f->cos_gamma);
    mat3 B1, B2;
    m3_zero(B1); m3_zero(B2);
    m3_axpy(B1, P.p5, B[3]); m3_axpy(B1, P.p4, B[4]); m3_axpy(B1, P.p3, B[5]);
    m3_axpy(B2, P.p2, B[1]); m3_axpy(B2, P.p1, B[2]);
    m3_eye(r1, T1t * I[1]);
    m3_zero(r2); m3_axpy(r2, I[2], B1); m3_axpy(r2, I[3], B2); m3_axpy(r2, T1t * I[4], B[0]);
    m3_eye(r3, T1t * I[4]);
    m3_zero(r4); m3_axpy(r4, I[5], B1); m3_axpy(r4, I[6], B2); m3_axpy(r4, T1t * I[7], B[0]);
}

/* AI/BEMenerg_analiticIntg_SC_SP.m:1-75 */
static void intg_SC_SP(double RS, double rhoS, double zeta, double TH1, double TH2, const tinfo_t *f,
                       mat3 B[6], mat3 r5, mat3 r6) {
    double T1t = TH2 - TH1, T2t = TH2 + TH1;
    double I[5];
    mat3 B1, B2;
    m3_zero(B1); m3_zero(B2);
    if (zeta > 1.0e-6) {
        CNT(2);
        I[1] = RS - zeta;
        I[2] = RS + (zeta * zeta / RS) - (2 * zeta);
        I[3] = log((rhoS + RS) / zeta) - (rhoS / RS);
        I[4] = -(1 / RS) + (1 / zeta);
        phi_t P = phis(T1t, T2t, f->gamma, f->cos_gamma);
        m3_axpy(B1, P.p5, B[3]); m3_axpy(B1, P.p4, B[4]); m3_axpy(B1, P.p3, B[5]);
        m3_axpy(B2, P.p2, B[1]); m3_axpy(B2, P.p1, B[2]);
        m3_eye(r5, T1t * I[1]);
        m3_zero(r6); m3_axpy(r6, I[2], B1); m3_axpy(r6, I[3], B2); m3_axpy(r6, T1t * I[4], B[0]);
    } else {
        CNT(9);
        I[1] = RS;
        I[2] = RS;
        phi_t P = phis(T1t, T2t, f->gamma, f->cos_gamma);
        /* Phi(4) is written with its two terms swapped at SC_SP.m:58; same value */
        m3_axpy(B1, P.p5, B[3]); m3_axpy(B1, P.p4, B[4]); m3_axpy(B1, P.p3, B[5]);
        m3_eye(r5, T1t * I[1]);
        m3_zero(r6); m3_axpy(r6, I[2], B1);
    }
}

/* AI/coefficientiG/BEMenerg_analiticIntg_coeffG_TC.m:1-182.  G is 1-based (G[1..19]). */
static void coeffG_TC(double zeta, double RS, double rhoS, double TH1, double TH2, const tinfo_t *f,
                      double G[20]) {
    double a = f->a;
    double sin_beta = f->sin_beta, sin_alpha = f->sin_alpha;
    double cos_alpha = f->cos_alpha, cos_beta = f->cos_beta, cos_gamma = f->cos_gamma;
    double alpha = f->alpha, beta = f->beta, gamma = f->gamma;
    double T1t = TH2 - TH1, T2t = TH2 + TH1;
    double F = a * sin_beta;
    phi_t P = phis(T1t, T2t, gamma, cos_gamma);
    memset(G, 0, sizeof(double) * 20);
    if (zeta > 1.0e-6) {
        double R1 = F / sin(TH1 + beta);
        double R2 = F / sin(TH2 + beta);
        double fz1 = zeta / R1, fz2 = zeta / R2;
        double ADD = zeta * zeta + F * F;
        double g1 = sqrt(1 + fz1 * fz1);
        double g2 = sqrt(1 + fz2 * fz2);
        double c1b = cos(TH1 + beta), c2b = cos(TH2 + beta);
        double g3 = 1.0 / 2 * (log((g1 + c1b) * (g2 - c2b)) - log((g1 - c1b) * (g2 + c2b)));
        double g4 = log((1 / fz1) + sqrt(1 + (1 / fz1) * (1 / fz1)));
        double g5 = log((1 / fz2) + sqrt(1 + (1 / fz2) * (1 / fz2)));
        double g7 = PI - acos(clamp1((-zeta * c2b) / sqrt(ADD))) - acos(clamp1((zeta * c1b) / sqrt(ADD)));
        double g8 = F * F / ADD;
        double lnS = (rhoS / RS) - log((rhoS + RS) / zeta);
        double s1b = sin(TH1 + beta), s2b = sin(TH2 + beta);
        G[1] = -(RS * T1t) + (zeta * g7) + (F * g3);
        G[4] = (1 / RS * T1t) - (1 / zeta * g7);
        G[5] = cos(gamma - TH2) * g5 - cos(gamma - TH1) * g4 - cos_alpha * g3 +
               lnS * (cos(gamma - TH2) - cos(gamma - TH1));
        G[6] = -cos(TH2) * g5 + cos(TH1) * g4 - cos_beta * g3 - lnS * (cos(TH2) - cos(TH1));
        double q = ((RS * RS + zeta * zeta) / RS);
        G[7] = zeta * g7 - q * P.p5 +
               F * (-g1 * cos(alpha - gamma + TH1) + g2 * cos(alpha - gamma + TH2) + g3 * sin_alpha * sin_alpha);
        G[8] = -zeta * cos_gamma * g7 - q * P.p4 +
               F * (-sin_alpha * sin_beta * g3 + g1 * cos(TH1 + alpha) - g2 * cos(TH2 + alpha));
        G[9] = zeta * g7 - q * P.p3 + F * (sin_beta * sin_beta * g3 + g1 * cos(TH1 - beta) - g2 * cos(TH2 - beta));
        G[14] = 1 / (3 * RS * RS * RS) * T1t +
                1 / (3 * zeta * zeta) * (F / ADD * (c1b / g1 - c2b / g2) - 1 / zeta * g7);
        double rS3 = rhoS * rhoS * rhoS / (3 * zeta * zeta * RS * RS * RS);
        G[15] = -rS3 * P.p2 +
                1 / (3 * zeta * zeta) *
                    ((sin_alpha * s2b - g8 * cos_alpha * c2b) / g2 + (-sin_alpha * s1b + g8 * cos_alpha * c1b) / g1);
        G[16] = -rS3 * P.p1 +
                1 / (3 * zeta * zeta) *
                    ((sin_beta * s1b + g8 * cos_beta * c1b) / g1 + (-sin_beta * s2b - g8 * cos_beta * c2b) / g2);
        double tail = ((1 / RS) - (zeta * zeta / (3 * RS * RS * RS)));
        G[17] = 1.0 / 3 * (-(s2b * s2b) * cos(alpha - gamma + TH2) + g8 * (cos_alpha * cos_alpha) * c2b) / (F * g2) -
                1.0 / 3 * (-(s1b * s1b) * cos(alpha - gamma + TH1) + g8 * (cos_alpha * cos_alpha) * c1b) / (F * g1) -
                g7 / (3 * zeta) + tail * P.p5;
        G[18] = 1.0 / 3 * ((s2b * s2b) * cos(alpha + TH2) + g8 * (cos_alpha * cos_beta * c2b)) / (F * g2) -
                1.0 / 3 * ((s1b * s1b) * cos(alpha + TH1) + g8 * (cos_alpha * cos_beta * c1b)) / (F * g1) +
                g7 / (3 * zeta) * cos_gamma + tail * P.p4;
        G[19] = 1.0 / 3 * ((s2b * s2b) * cos(beta - TH2) + g8 * (cos_beta * cos_beta * c2b)) / (F * g2) -
                1.0 / 3 * ((s1b * s1b) * cos(beta - TH1) + g8 * (cos_beta * cos_beta * c1b)) / (F * g1) -
                g7 / (3 * zeta) + tail * P.p3;
    } else {
        double c1b = cos(TH1 + beta), c2b = cos(TH2 + beta);
        double s1b = sin(TH1 + beta), s2b = sin(TH2 + beta);
        double g_3 = 1.0 / 2 * (log((1 + c1b) * (1 - c2b)) - log((1 - c1b) * (1 + c2b)));
        G[1] = (F * g_3) - (RS * T1t);
        G[4] = 1 / F * (c2b - c1b) + 1 / RS * T1t;
        G[7] = -RS * P.p5 + F * (cos(alpha - gamma + TH2) - cos(alpha - gamma + TH1) + sin_alpha * sin_alpha * g_3);
        G[8] = -RS * P.p4 + F * (cos(TH1 + alpha) - cos(TH2 + alpha) - sin_alpha * sin_beta * g_3);
        G[9] = -RS * P.p3 + F * (cos(beta - TH1) - cos(beta - TH2) + sin_beta * sin_beta * g_3);
        G[17] = 1 / (3 * F) *
                    (((1 + cos_alpha * cos_alpha) * c2b - (s2b * s2b) * cos(alpha - gamma + TH2)) -
                     ((1 + cos_alpha * cos_alpha) * c1b - (s1b * s1b) * cos(alpha - gamma + TH1))) +
                1 / RS * P.p5;
        G[18] = 1 / (3 * F) *
                    ((c2b * ((cos_alpha * cos_beta) + cos(alpha + beta)) + (s2b * s2b) * cos(alpha + TH2)) -
                     (c1b * ((cos_alpha * cos_beta) + cos(alpha + beta)) + (s1b * s1b) * cos(alpha + TH1))) +
                1 / RS * P.p4;
        G[19] = 1 / (3 * F) *
                    (((1 + cos_beta * cos_beta) * c2b + (s2b * s2b) * cos(TH2 - beta)) -
                     ((1 + cos_beta * cos_beta) * c1b + (s1b * s1b) * cos(beta - TH1))) +
                1 / RS * P.p3;
    }
}

/* AI/coefficientiG/BEMenerg_analiticIntg_coeffG_TR_P.m:1-120 */
static void coeffG_TR_P(double zeta, double TH1, double TH2, const tinfo_t *f, double G[20]) {
    double a = f->a;
    double sin_beta = f->sin_beta, sin_alpha = f->sin_alpha;
    double cos_alpha = f->cos_alpha, cos_beta = f->cos_beta, cos_gamma = f->cos_gamma;
    double alpha = f->alpha, beta = f->beta, gamma = f->gamma;
    double T1t = TH2 - TH1, T2t = TH2 + TH1;
    double F = a * sin_beta;
    double R1 = F / sin(TH1 + beta), R2 = F / sin(TH2 + beta);
    double fz1 = zeta / R1, fz2 = zeta / R2;
    double ADD = zeta * zeta + F * F;
    phi_t P = phis(T1t, T2t, gamma, cos_gamma);
    double g1 = sqrt(1 + fz1 * fz1);
    double g2 = sqrt(1 + fz2 * fz2);
    double c1b = cos(TH1 + beta), c2b = cos(TH2 + beta);
    double s1b = sin(TH1 + beta), s2b = sin(TH2 + beta);
    double g3 = 1.0 / 2 * (log((g1 + c1b) * (g2 - c2b)) - log((g1 - c1b) * (g2 + c2b)));
    double g4 = log((1 / fz1) + sqrt(1 + (1 / fz1) * (1 / fz1)));
    double g5 = log((1 / fz2) + sqrt(1 + (1 / fz2) * (1 / fz2)));
    double g7 = PI - acos(clamp1((-zeta * c2b) / sqrt(ADD))) - acos(clamp1((zeta * c1b) / sqrt(ADD)));
    double g8 = F * F / ADD;
    memset(G, 0, sizeof(double) * 20);
    G[1] = (zeta * (-T1t + g7)) + (F * g3);
    G[4] = 1 / zeta * (T1t - g7);
    G[5] = cos(gamma - TH2) * g5 - cos(gamma - TH1) * g4 - cos_alpha * g3;
    G[6] = -cos(TH2) * g5 + cos(TH1) * g4 - cos_beta * g3;
    G[7] = zeta * (g7 - 2 * P.p5) +
           F * (-g1 * cos(alpha - gamma + TH1) + g2 * cos(alpha - gamma + TH2) + g3 * sin_alpha * sin_alpha);
    G[8] = -zeta * ((cos_gamma * g7) + 2 * P.p4) +
           F * (-sin_alpha * sin_beta * g3 + g1 * cos(TH1 + alpha) - g2 * cos(TH2 + alpha));
    G[9] = zeta * (g7 - 2 * P.p3) + F * (sin_beta * sin_beta * g3 + g1 * cos(TH1 - beta) - g2 * cos(TH2 - beta));
    G[14] = 1 / (3 * zeta * zeta) * (1 / zeta * (T1t - g7) + (F / ADD * (c1b / g1 - c2b / g2)));
    G[15] = 1 / (3 * zeta * zeta) *
            ((sin_alpha * s2b - g8 * cos_alpha * c2b) / g2 + (-sin_alpha * s1b + g8 * cos_alpha * c1b) / g1);
    G[16] = 1 / (3 * zeta * zeta) *
            ((sin_beta * s1b + g8 * cos_beta * c1b) / g1 + (-sin_beta * s2b - g8 * cos_beta * c2b) / g2);
    G[17] = (1 / (3 * F) *
             ((-(s2b * s2b) * cos(alpha - gamma + TH2) + g8 * (cos_alpha * cos_alpha) * c2b) / g2 -
              (-(s1b * s1b) * cos(alpha - gamma + TH1) + g8 * (cos_alpha * cos_alpha) * c1b) / g1)) -
            (1 / (3 * zeta) * (g7 - 2 * P.p5));
    G[18] = (1 / (3 * F) *
             (((s2b * s2b) * cos(alpha + TH2) + g8 * (cos_alpha * cos_beta * c2b)) / g2 -
              ((s1b * s1b) * cos(alpha + TH1) + g8 * (cos_alpha * cos_beta * c1b)) / g1)) +
            (1 / (3 * zeta) * ((g7 * cos_gamma) + 2 * P.p4));
    G[19] = (1 / (3 * F) *
             (((s2b * s2b) * cos(beta - TH2) + g8 * (cos_beta * cos_beta * c2b)) / g2 -
              ((s1b * s1b) * cos(beta - TH1) + g8 * (cos_beta * cos_beta * c1b)) / g1)) -
            (1 / (3 * zeta) * (g7 - 2 * P.p3));
}

/* AI/coefficientiG/BEMenerg_analiticIntg_coeffG_TR_SP.m:1-111 */
static void coeffG_TR_SP(double zeta, double TH1, double TH2, const tinfo_t *f, double G[20]) {
    double a = f->a;
    double sin_beta = f->sin_beta, sin_alpha = f->sin_alpha;
    double cos_alpha = f->cos_alpha, cos_beta = f->cos_beta, cos_gamma = f->cos_gamma;
    double alpha = f->alpha, beta = f->beta, gamma = f->gamma;
    double T1t = TH2 - TH1, T2t = TH2 + TH1;
    double F = a * sin_beta;
    memset(G, 0, sizeof(double) * 20);
    if (zeta > 1.0e-6) {
        double R1 = F / sin(TH1 + beta), R2 = F / sin(TH2 + beta);
        double fz1 = zeta / R1, fz2 = zeta / R2;
        double ADD = zeta * zeta + F * F;
        phi_t P = phis(T1t, T2t, gamma, cos_gamma);
        double g1 = sqrt(1 + fz1 * fz1);
        double g2 = sqrt(1 + fz2 * fz2);
        double c1b = cos(TH1 + beta), c2b = cos(TH2 + beta);
        double g3 = 1.0 / 2 * (log((g1 + c1b) * (g2 - c2b)) - log((g1 - c1b) * (g2 + c2b)));
        double g4 = log((1 / fz1) + sqrt(1 + (1 / fz1) * (1 / fz1)));
        double g5 = log((1 / fz2) + sqrt(1 + (1 / fz2) * (1 / fz2)));
        double g7 = PI - acos(clamp1((-zeta * c2b) / sqrt(ADD))) - acos(clamp1((zeta * c1b) / sqrt(ADD)));
        G[1] = (zeta * (-T1t + g7)) + (F * g3);
        G[4] = 1 / zeta * (T1t - g7);
        G[5] = cos(gamma - TH2) * g5 - cos(gamma - TH1) * g4 - cos_alpha * g3;
        G[6] = -cos(TH2) * g5 + cos(TH1) * g4 - cos_beta * g3;
        G[7] = zeta * (g7 - 2 * P.p5) +
               F * (-g1 * cos(alpha - gamma + TH1) + g2 * cos(alpha - gamma + TH2) + g3 * sin_alpha * sin_alpha);
        G[8] = -zeta * ((cos_gamma * g7) + 2 * P.p4) +
               F * (-sin_alpha * sin_beta * g3 + g1 * cos(TH1 + alpha) - g2 * cos(TH2 + alpha));
        G[9] = zeta * (g7 - 2 * P.p3) + F * (sin_beta * sin_beta * g3 + g1 * cos(TH1 - beta) - g2 * cos(TH2 - beta));
    } else {
        double c1b = cos(TH1 + beta), c2b = cos(TH2 + beta);
        double g_3 = 1.0 / 2 * (log((1 + c1b) * (1 - c2b)) - log((1 - c1b) * (1 + c2b)));
        G[1] = F * g_3;
        G[7] = F * (-cos(alpha - gamma + TH1) + cos(alpha - gamma + TH2) + sin_alpha * sin_alpha * g_3);
        G[8] = F * (-(sin_alpha * sin_beta * g_3) + cos(TH1 + alpha) - cos(TH2 + alpha));
        G[9] = F * (sin_beta * sin_beta * g_3 + cos(TH1 - beta) - cos(TH2 - beta));
    }
}

/* AI/BEMenerg_analiticIntg_TC.m:1-50 */
static void intg_TC(double zeta, double RS, double rhoS, double TH1, double TH2, const tinfo_t *f,
                    mat3 B[6], mat3 r1, mat3 r2, mat3 r3, mat3 r4) {
    double G[20];
    if (zeta > 1.0e-6) {
        CNT(3);
        coeffG_TC(zeta, RS, rhoS, TH1, TH2, f, G);
        m3_eye(r1, G[1]);
        m3_zero(r2);
        for (int k = 0; k < 6; k++) m3_axpy(r2, G[4 + k], B[k]);
        m3_eye(r3, G[4]);
        m3_zero(r4);
        for (int k = 0; k < 6; k++) m3_axpy(r4, G[14 + k], B[k]);
    } else {
        CNT(4);
        coeffG_TC(0, RS, rhoS, TH1, TH2, f, G);
        m3_eye(r1, G[1]);
        m3_zero(r2);
        for (int k = 3; k < 6; k++) m3_axpy(r2, G[4 + k], B[k]);
        m3_eye(r3, G[4]);
        m3_zero(r4);
        for (int k = 3; k < 6; k++) m3_axpy(r4, G[14 + k], B[k]);
    }
}

/* AI/BEMenerg_analiticIntg_TR_P.m:1-28 */
static void intg_TR_P(double zeta, double TH1, double TH2, const tinfo_t *f, mat3 B[6], mat3 r1, mat3 r2,
                      mat3 r3, mat3 r4) {
    CNT(5);
    double G[20];
    coeffG_TR_P(zeta, TH1, TH2, f, G);
    m3_eye(r1, G[1]);
    m3_zero(r2);
    for (int k = 0; k < 6; k++) m3_axpy(r2, G[4 + k], B[k]);
    m3_eye(r3, G[4]);
    m3_zero(r4);
    for (int k = 0; k < 6; k++) m3_axpy(r4, G[14 + k], B[k]);
}

/* AI/BEMenerg_analiticIntg_TR_SP.m:1-40 */
static void intg_TR_SP(double zeta, double TH1, double TH2, const tinfo_t *f, mat3 B[6], mat3 r5, mat3 r6) {
    double G[20];
    if (zeta > 1.0e-6) {
        CNT(6);
        coeffG_TR_SP(zeta, TH1, TH2, f, G);
        m3_eye(r5, G[1]);
        m3_zero(r6);
        for (int k = 0; k < 6; k++) m3_axpy(r6, G[4 + k], B[k]);
    } else {
        CNT(7);
        coeffG_TR_SP(0, TH1, TH2, f, G);
        m3_eye(r5, G[1]);
        m3_zero(r6);
        for (int k = 3; k < 6; k++) m3_axpy(r6, G[4 + k], B[k]);
    }
}

/* ------------------------------------------- per-region entry and trace -- */

/* Region kinds of orc_intg and the trace: the six routines of
 * functions/BEMenerg_analiticIntg/.  args = {RS, rhoS, RP, rhoP, zeta, TH1, TH2, 0}
 * (each routine reads the ones in its argument list), info = the twelve
 * fields of infoTri2D (calcoloDatiTriangolo.m:10-32: a, b, c, cos_alpha,
 * cos_beta, cos_gamma, sin_alpha, sin_beta, sin_gamma, alpha, beta, gamma),
 * B = B(:,:,1..6) column-major, ris = ris1..ris4 (P routines) or ris5, ris6
 * (S routines) column-major. */
enum { AI_CC = 0, AI_SC_P = 1, AI_SC_SP = 2, AI_TC = 3, AI_TR_P = 4, AI_TR_SP = 5 };

static void info_unpack(const double *info, tinfo_t *f) {
    f->a = info[0]; f->b = info[1]; f->c = info[2];
    f->cos_alpha = info[3]; f->cos_beta = info[4]; f->cos_gamma = info[5];
    f->sin_alpha = info[6]; f->sin_beta = info[7]; f->sin_gamma = info[8];
    f->alpha = info[9]; f->beta = info[10]; f->gamma = info[11];
}

static void info_pack(const tinfo_t *f, double *info) {
    const double v[12] = {f->a, f->b, f->c, f->cos_alpha, f->cos_beta, f->cos_gamma,
                          f->sin_alpha, f->sin_beta, f->sin_gamma, f->alpha, f->beta, f->gamma};
    memcpy(info, v, sizeof v);
}

static void intg_dispatch(int kind, const double *a, const tinfo_t *f, mat3 B[6], mat3 r[4]) {
    for (int k = 0; k < 4; k++) m3_zero(r[k]);
    switch (kind) {
        case AI_CC: intg_CC(a[0], a[1], a[2], a[3], a[4], a[5], a[6], f, B, r[0], r[1], r[2], r[3]); break;
        case AI_SC_P: intg_SC_P(a[2], a[3], a[4], a[5], a[6], f, B, r[0], r[1], r[2], r[3]); break;
        case AI_SC_SP: intg_SC_SP(a[0], a[1], a[4], a[5], a[6], f, B, r[0], r[1]); break;
        case AI_TC: intg_TC(a[4], a[0], a[1], a[5], a[6], f, B, r[0], r[1], r[2], r[3]); break;
        case AI_TR_P: intg_TR_P(a[4], a[5], a[6], f, B, r[0], r[1], r[2], r[3]); break;
        default: intg_TR_SP(a[4], a[5], a[6], f, B, r[0], r[1]); break;
    }
}

int orc_intg(int kind, const double *args, const double *info, const double *Bin, double *ris) {
    if (kind < AI_CC || kind > AI_TR_SP) return -1;
    tinfo_t f;
    info_unpack(info, &f);
    mat3 B[6], r[4];
    for (int k = 0; k < 6; k++)
        for (int i = 0; i < 3; i++)
            for (int j = 0; j < 3; j++) B[k][i][j] = Bin[9 * k + i + 3 * j];
    intg_dispatch(kind, args, &f, B, r);
    for (int k = 0; k < 4; k++)
        for (int i = 0; i < 3; i++)
            for (int j = 0; j < 3; j++) ris[9 * k + i + 3 * j] = r[k][i][j];
    return 0;
}

typedef struct {
    int cap, n;
    int *kind;
    double *args, *info, *B, *ris;
} trace_t;
static trace_t *g_trace; /* set only inside orc_region_trace (single-threaded use) */

static void trace_call(int kind, double RS, double rhoS, double RP, double rhoP, double zeta, double TH1,
                       double TH2, const tinfo_t *f, mat3 B[6]) {
    trace_t *T = g_trace;
    if (!T) return;
    int k = T->n++;
    if (k >= T->cap) return;
    const double a[8] = {RS, rhoS, RP, rhoP, zeta, TH1, TH2, 0};
    T->kind[k] = kind;
    memcpy(T->args + 8 * k, a, sizeof a);
    info_pack(f, T->info + 12 * k);
    for (int q = 0; q < 6; q++)
        for (int i = 0; i < 3; i++)
            for (int j = 0; j < 3; j++) T->B[54 * k + 9 * q + i + 3 * j] = B[q][i][j];
    mat3 r[4];
    intg_dispatch(kind, a, f, B, r);
    for (int q = 0; q < 4; q++)
        for (int i = 0; i < 3; i++)
            for (int j = 0; j < 3; j++) T->ris[36 * k + 9 * q + i + 3 * j] = r[q][i][j];
}

/* ------------------------------------------------- child / field triangle -- */

enum { T_NONE = 0, T_SPSP = 1, T_SPP = 2, T_P = 3 };

/* SAc/BEMenerg_coreSA_calcIntgInt_childTriangle.m:1-234.  R[0..5] = ris1..ris6. */
static void child_triangle(double t2[3][2], const tinfo_t *f, double zeta, double RP, double rhoP, double RS,
                           double rhoS, double sign_prod, int type, mat3 R[6]) {
    double a = f->a, sinBeta = f->sin_beta, beta = f->beta, gamma = f->gamma;
    mat3 B[6];
    coeffB(zeta, f, sign_prod, B);
    double vt0 = (t2[2][0] - t2[1][0]) / f->c;
    double Bm = a * vt0;
    double C = a * a - rhoP * rhoP;
    double DeltaP = Bm * Bm - C;
    C = a * a - rhoS * rhoS;
    double DeltaS = Bm * Bm - C;
    double rP = a * sinBeta / rhoP;
    double rS = a * sinBeta / rhoS;
    if (type == T_P) {
        rS = 0;
        DeltaS = 0;
    }
    for (int k = 0; k < 6; k++) m3_zero(R[k]);
    double TH1, TH2, TH3, TH4, TH5, TH6;
    if (DeltaP <= 1.0e-6) {
        TH1 = 0; TH2 = gamma; TH3 = gamma; TH4 = gamma; TH5 = gamma; TH6 = gamma;
    } else if ((DeltaP > 1.0e-6) && (DeltaS <= 1.0e-6)) {
        CNT(10);
        double as = asin(clamp1(rP));
        double O1 = as - beta, O2 = PI - as - beta;
        double th1 = fmin(O1, O2), th2 = fmax(O1, O2);
        TH1 = 0;
        TH2 = fmin(fmax(0, th1), gamma);
        TH3 = fmax(0, fmin(th2, gamma));
        TH4 = TH3;
        TH5 = TH3;
        TH6 = gamma;
    } else {
        CNT(11);
        double as = asin(clamp1(rP)), bs = asin(clamp1(rS));
        double O1 = as - beta, O2 = PI - as - beta, O3 = bs - beta, O4 = PI - bs - beta;
        double th1 = fmin(O1, O2), th2 = fmax(O1, O2), th3 = fmin(O3, O4), th4 = fmax(O3, O4);
        TH1 = 0;
        TH2 = fmin(fmax(0, th1), gamma);
        TH3 = fmin(fmax(0, th3), gamma);
        TH4 = fmax(0, fmin(th4, gamma));
        TH5 = fmax(0, fmin(th2, gamma));
        TH6 = gamma;
    }
    mat3 q1, q2, q3, q4, q5, q6;
    int sp = (type == T_SPSP);
    if (TH1 < TH2) { /* E1 + E6 */
        if (sp) intg_SC_SP(RS, rhoS, zeta, TH1, TH2, f, B, q5, q6);
        if (type == T_P) intg_SC_P(RP, rhoP, zeta, TH1, TH2, f, B, q1, q2, q3, q4);
        else intg_CC(RS, rhoS, RP, rhoP, zeta, TH1, TH2, f, B, q1, q2, q3, q4);
        if (sp) trace_call(AI_SC_SP, RS, rhoS, RP, rhoP, zeta, TH1, TH2, f, B);
        trace_call(type == T_P ? AI_SC_P : AI_CC, RS, rhoS, RP, rhoP, zeta, TH1, TH2, f, B);
        m3_axpy(R[0], 1, q1); m3_axpy(R[1], 1, q2); m3_axpy(R[2], 1, q3); m3_axpy(R[3], 1, q4);
        if (sp) { m3_axpy(R[4], 1, q5); m3_axpy(R[5], 1, q6); }
    }
    if (TH2 < TH3) { /* E2 + E7 */
        if (sp) intg_SC_SP(RS, rhoS, zeta, TH2, TH3, f, B, q5, q6);
        if (type == T_P) intg_TR_P(zeta, TH2, TH5, f, B, q1, q2, q3, q4);
        else intg_TC(zeta, RS, rhoS, TH2, TH3, f, B, q1, q2, q3, q4);
        if (sp) trace_call(AI_SC_SP, RS, rhoS, RP, rhoP, zeta, TH2, TH3, f, B);
        if (type == T_P) trace_call(AI_TR_P, RS, rhoS, RP, rhoP, zeta, TH2, TH5, f, B);
        else trace_call(AI_TC, RS, rhoS, RP, rhoP, zeta, TH2, TH3, f, B);
        m3_axpy(R[0], 1, q1); m3_axpy(R[1], 1, q2); m3_axpy(R[2], 1, q3); m3_axpy(R[3], 1, q4);
        if (sp) { m3_axpy(R[4], 1, q5); m3_axpy(R[5], 1, q6); }
    }
    if (TH3 < TH4) { /* E3 */
        if (sp) {
            intg_TR_SP(zeta, TH3, TH4, f, B, q5, q6);
            trace_call(AI_TR_SP, RS, rhoS, RP, rhoP, zeta, TH3, TH4, f, B);
            m3_axpy(R[4], 1, q5); m3_axpy(R[5], 1, q6);
        }
    }
    if (TH4 < TH5) { /* E4 + E8 */
        if (sp) intg_SC_SP(RS, rhoS, zeta, TH4, TH5, f, B, q5, q6);
        intg_TC(zeta, RS, rhoS, TH4, TH5, f, B, q1, q2, q3, q4);
        if (sp) trace_call(AI_SC_SP, RS, rhoS, RP, rhoP, zeta, TH4, TH5, f, B);
        trace_call(AI_TC, RS, rhoS, RP, rhoP, zeta, TH4, TH5, f, B);
        m3_axpy(R[0], 1, q1); m3_axpy(R[1], 1, q2); m3_axpy(R[2], 1, q3); m3_axpy(R[3], 1, q4);
        if (sp) { m3_axpy(R[4], 1, q5); m3_axpy(R[5], 1, q6); }
    }
    if (TH5 < TH6) { /* E5 + E9 */
        if (sp) intg_SC_SP(RS, rhoS, zeta, TH5, TH6, f, B, q5, q6);
        if (type == T_P) intg_SC_P(RP, rhoP, zeta, TH5, TH6, f, B, q1, q2, q3, q4);
        else intg_CC(RS, rhoS, RP, rhoP, zeta, TH5, TH6, f, B, q1, q2, q3, q4);
        if (sp) trace_call(AI_SC_SP, RS, rhoS, RP, rhoP, zeta, TH5, TH6, f, B);
        trace_call(type == T_P ? AI_SC_P : AI_CC, RS, rhoS, RP, rhoP, zeta, TH5, TH6, f, B);
        m3_axpy(R[0], 1, q1); m3_axpy(R[1], 1, q2); m3_axpy(R[2], 1, q3); m3_axpy(R[3], 1, q4);
        if (sp) { m3_axpy(R[4], 1, q5); m3_axpy(R[5], 1, q6); }
    }
}

/* SAc/BEMenerg_coreSA_calcIntgInt_fieldTriangle.m:1-85 */
static void field_triangle(double zeta, double RP, double rhoP, double RS, double rhoS, double ch[3][3][2],
                           const double c[3], double sign_prod, int type, mat3 R[6]) {
    for (int k = 0; k < 6; k++) m3_zero(R[k]);
    for (int k = 0; k < 3; k++) {
        double v[3][2];
        for (int i = 0; i < 3; i++) { v[i][0] = ch[k][i][0]; v[i][1] = ch[k][i][1]; }
        double C = c[k];
        if (fabs(C) < 1.0e-06) continue;
        if (C < 0) {
            double tx = v[1][0], ty = v[1][1];
            v[1][0] = v[2][0]; v[1][1] = v[2][1];
            v[2][0] = tx; v[2][1] = ty;
        }
        double th0 = atan2(v[1][1], v[1][0]);
        if (th0 < 0) th0 = fabs(th0);
        else th0 = 2 * PI - th0;
        double t2[3][2];
        tinfo_t f;
        tri_data(v, t2, &f);
        mat3 Rt[6];
        child_triangle(t2, &f, zeta, RP, rhoP, RS, rhoS, sign_prod, type, Rt);
        double ct = cos(th0), st = sin(th0);
        mat3 A = {{ct, st, 0}, {-st, ct, 0}, {0, 0, 1}};
        for (int r = 1; r < 6; r += 2) { /* ris2, ris4, ris6 */
            mat3 tmp, out;
            m3_mul_bt(tmp, Rt[r], A);
            m3_mul(out, A, tmp);
            memcpy(Rt[r], out, sizeof(mat3));
        }
        double coeff = msign(C);
        for (int r = 0; r < 6; r++) m3_axpy(R[r], coeff, Rt[r]);
    }
}

/* geo/rotazioneInversaEulero.m:1-95: ris <- C' * (ris * C) */
static void rot_inv_matrix(const sdr_t *s, const double *Z, mat3 C) {
    const double *X = s->e1;
    double cz[3] = {-Z[1], Z[0], 0}; /* cross([0 0 1], Z) */
    double ncz = sqrt(cz[0] * cz[0] + cz[1] * cz[1] + cz[2] * cz[2]);
    double phi, omega, psi;
    if (ncz < 1.0e-10) {
        double l[3] = {X[0], X[1], X[2]};
        phi = acos(clamp1(l[0]));
        if (l[1] < -1.0e-10) phi = 2 * PI - phi; /* cross([1 0 0], l) * [0 0 1]' = l(2) */
        omega = PI * (Z[2] < 0);
        psi = 0;
    } else {
        double l[3] = {cz[0] / ncz, cz[1] / ncz, cz[2] / ncz};
        phi = acos(clamp1(l[0]));
        if (l[1] < -1.0e-10) phi = 2 * PI - phi;
        omega = acos(clamp1(Z[2]));
        psi = acos(clamp1(dot3(l, X)));
        double cl[3] = {l[1] * X[2] - l[2] * X[1], l[2] * X[0] - l[0] * X[2], l[0] * X[1] - l[1] * X[0]};
        if (dot3(cl, Z) < -1.0e-10) psi = 2 * PI - psi;
    }
    mat3 Ap = {{cos(psi), sin(psi), 0}, {-sin(psi), cos(psi), 0}, {0, 0, 1}};
    mat3 Ao = {{1, 0, 0}, {0, cos(omega), sin(omega)}, {0, -sin(omega), cos(omega)}};
    mat3 Af = {{cos(phi), sin(phi), 0}, {-sin(phi), cos(phi), 0}, {0, 0, 1}};
    mat3 T;
    m3_mul(T, Ap, Ao);
    m3_mul(C, T, Af);
    for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++)
            if (fabs(C[i][j]) < 1.0e-14) C[i][j] = 0;
}

static void rot_inv_apply(mat3 C, mat3 ris) {
    mat3 tmp, out;
    m3_mul(tmp, ris, C);
    m3_mul_at(out, C, tmp);
    memcpy(ris, out, sizeof(mat3));
}

/* SAc/BEMenerg_coreSA_calcMatrixSubBlock.m:58-81 (identical in
 * BEMenerg_postProcA_calcMatrixBlock.m:28-52). */
static int classify(double zeta, double rP, double rS, double dMIN, double *rhoP, double *rhoS) {
    if (zeta == 0) {
        *rhoP = rP;
        *rhoS = rS;
    } else if (zeta < rS) {
        *rhoP = sqrt(rP * rP - zeta * zeta);
        *rhoS = sqrt(rS * rS - zeta * zeta);
    } else if (rS < zeta && zeta < rP) {
        *rhoP = sqrt(rP * rP - zeta * zeta);
        *rhoS = -1;
    } else {
        *rhoP = -1;
        *rhoS = -1;
    }
    if (dMIN < *rhoS) return T_SPSP;
    if (0 < *rhoS && *rhoS <= dMIN && dMIN < *rhoP) return T_SPP;
    if (*rhoS <= 0 && dMIN < *rhoP) return T_P;
    return T_NONE;
}

/* Per (outer point, time) kernel integrals after the inverse rotation. */
static int node_integrals(const orc_param *p, const double *x, const double *T, const double *vn,
                          const sdr_t *s, double t, mat3 R[6]) {
    double rP = p->velP * t, rS = p->velS * t;
    double zeta, ch[3][3][2], dMIN, c[3], sgp, rhoP, rhoS;
    prep_triangle(x, T, vn, s, &zeta, ch, &dMIN, c, &sgp);
    int type = classify(zeta, rP, rS, dMIN, &rhoP, &rhoS);
    for (int k = 0; k < 6; k++) m3_zero(R[k]);
    if (type == T_NONE) return 0;
    field_triangle(zeta, rP, rhoP, rS, rhoS, ch, c, sgp, type, R);
    mat3 C;
    rot_inv_matrix(s, vn, C);
    rot_inv_apply(C, R[1]);
    rot_inv_apply(C, R[3]);
    rot_inv_apply(C, R[5]);
    return type;
}

int orc_region_trace(const orc_param *p, const double *x, const double *T, const double *n, double t, int cap,
                     int *kind, double *args, double *info, double *B, double *ris) {
    trace_t tr = {cap, 0, kind, args, info, B, ris};
    sdr_t s;
    sdr_compute(T, n, &s);
    mat3 R[6];
    g_trace = &tr;
    node_integrals(p, x, T, n, &s, t, R);
    g_trace = NULL;
    return tr.n;
}

int orc_node_ris(const orc_param *p, const double *x, const double *T, const double *n, double t, double *out) {
    sdr_t s;
    sdr_compute(T, n, &s);
    mat3 R[6];
    int type = node_integrals(p, x, T, n, &s, t, R);
    for (int k = 0; k < 6; k++)
        for (int i = 0; i < 3; i++)
            for (int j = 0; j < 3; j++) out[9 * k + i + 3 * j] = R[k][i][j];
    return type;
}

/* Kernel combination of calcMatrixSubBlock.m:101-104 for one time t. */
static void combine(const orc_param *p, double t, mat3 R[6], mat3 out) {
    double vP2 = p->velP * p->velP, vS2 = p->velS * p->velS;
    double den = (p->velP * p->velS) * (p->velP * p->velS);
    for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++)
            out[i][j] = ((R[0][i][j] - R[1][i][j]) / vP2 + t * t * (-R[2][i][j] + 3 * R[3][i][j]) +
                         ((vP2 + vS2) * R[4][i][j] + (vP2 - vS2) * R[5][i][j]) / den) /
                        2;
}

/* ------------------------------------------------------------ assembly -- */

/* SAc/BEMenerg_coreSA_calcMatrixSubBlock.m:1-112 */
static void sub_block(const orc_param *p, const double *TF, const double *vnF, int indTemp, double dt,
                      const sdr_t *sF, int nGH, const double *nodes, const double *wts, mat3 ris) {
    const double coef[3] = {1, -2, 1};
    m3_zero(ris);
    for (int var = 0; var < 3; var++) {
        int ist = indTemp + 1 - var;
        if (ist <= 0) continue;
        mat3 acc[6];
        for (int k = 0; k < 6; k++) m3_zero(acc[k]);
        double currT = ist * dt;
        for (int q = 0; q < nGH; q++) {
            mat3 R[6];
            if (node_integrals(p, nodes + 3 * q, TF, vnF, sF, currT, R) == T_NONE) continue;
            for (int k = 0; k < 6; k++) m3_axpy(acc[k], wts[q], R[k]);
        }
        mat3 rist;
        combine(p, currT, acc, rist);
        m3_axpy(ris, coef[var], rist);
    }
    for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++) ris[i][j] = ris[i][j] / (4 * PI * p->rho);
}

/* BEMenerg_coreSA_calcCostantData.m:1-24: GH nodes/weights of triangle S */
static void tri_gh(const orc_mesh *m, int S, int nGH, const double *ghn, const double *ghw, double *nodes,
                   double *wts) {
    const double *V[3] = {m->coords + 3 * m->tri[3 * S], m->coords + 3 * m->tri[3 * S + 1],
                          m->coords + 3 * m->tri[3 * S + 2]};
    for (int q = 0; q < nGH; q++) {
        for (int k = 0; k < 3; k++)
            nodes[3 * q + k] = ghn[3 * q] * V[0][k] + ghn[3 * q + 1] * V[1][k] + ghn[3 * q + 2] * V[2][k];
        wts[q] = m->area[S] * ghw[q];
    }
}

static void tri_verts(const orc_mesh *m, int F, double *T) {
    for (int v = 0; v < 3; v++)
        for (int k = 0; k < 3; k++) T[3 * v + k] = m->coords[3 * m->tri[3 * F + v] + k];
}

/* Light-cone test of calcMatrixBlock.m:57-61 */
static int pair_passes(const orc_param *p, const orc_mesh *m, int S, int F, int lag, double dt) {
    const double *cS = m->center + 3 * S, *cF = m->center + 3 * F;
    double v0 = cF[0] - cS[0], v1 = cF[1] - cS[1], v2 = cF[2] - cS[2];
    double d = sqrt(v0 * v0 + v1 * v1 + v2 * v2);
    double distMin = d - m->maxL[F] - m->maxL[S];
    double distMax = d + m->maxL[F] + m->maxL[S];
    return ((lag - 1) * p->velS * dt < distMax) && ((lag + 1) * p->velP * dt >= distMin);
}

int orc_pair_block(const orc_param *p, const orc_mesh *m, int nGH, const double *ghn, const double *ghw, int lag,
                   int S, int F, double *out) {
    double dt = p->Tfin / p->nT;
    memset(out, 0, 9 * sizeof(double));
    if (!pair_passes(p, m, S, F, lag, dt)) return 0;
    double nodes[28 * 3], wts[28], T[9];
    tri_gh(m, S, nGH, ghn, ghw, nodes, wts);
    tri_verts(m, F, T);
    sdr_t sF;
    sdr_compute(T, m->normal + 3 * F, &sF);
    mat3 ris;
    sub_block(p, T, m->normal + 3 * F, lag, dt, &sF, nGH, nodes, wts, ris);
    for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++) out[i + 3 * j] = fabs(ris[i][j]) < 1.0e-14 ? 0 : ris[i][j];
    return 1;
}

/* SAc/BEMenerg_coreSA_calcMatrixBlock.m:1-90, restricted to test rows [r0,r1). */
int orc_matrix_block_rows(const orc_param *p, const orc_mesh *m, int nGH, const double *ghn, const double *ghw,
                          int lag, int r0, int r1, double *out, int nthreads) {
    int N = m->N;
    long n3 = 3L * N;
    int any = 0;
    long npairs = (long)(r1 - r0) * N;
#ifdef _OPENMP
    if (nthreads < 1) nthreads = 1;
#pragma omp parallel for schedule(dynamic, 64) num_threads(nthreads) reduction(| : any)
#endif
    for (long ind = 0; ind < npairs; ind++) {
        int S = r0 + (int)(ind % (r1 - r0)); /* ind2sub: column-major */
        int F = (int)(ind / (r1 - r0));
        double blk[9];
        int ok = orc_pair_block(p, m, nGH, ghn, ghw, lag, S, F, blk);
        any |= ok;
        for (int i = 0; i < 3; i++)
            for (int j = 0; j < 3; j++) out[(3L * S + i) + n3 * (3L * F + j)] = blk[i + 3 * j];
    }
    (void)nthreads;
    return any;
}

int orc_matrix_block(const orc_param *p, const orc_mesh *m, int nGH, const double *ghn, const double *ghw, int lag,
                     double *out, int nthreads) {
    return orc_matrix_block_rows(p, m, nGH, ghn, ghw, lag, 0, m->N, out, nthreads);
}

/* ------------------------------------------------------------------ RHS -- */

void orc_datum(int kind, const double *x, const double *nrm, double t, double *g) {
    g[0] = g[1] = g[2] = 0;
    switch (kind) {
    case 0: {
        double s = sin(PI * t / 2);
        g[2] = t > 0 ? s * s : 0;
        break;
    }
    case 1: {
        double f = t > 0 ? t * t * exp(-4 * t) : 0;
        g[0] = nrm[0] * f; g[1] = nrm[1] * f; g[2] = nrm[2] * f;
        break;
    }
    case 2:
        g[2] = 0.5 * t;
        break;
    case 3: {
        double s = sin(t);
        double a = (t < PI / 2) ? s * s * s * s * s : 1;
        g[2] = a * (x[0] * x[0]);
        break;
    }
    default:
        break;
    }
}

/* Tn/BEMenerg_core_calcTnBlock.m:27-58 with BEMenerg_core_calcNucleoTn.m:183 (g(t1) - g(t0)) */
void orc_rhs(const orc_param *p, const orc_mesh *m, int nGH, const double *ghn, const double *ghw, int datum,
             double *beta) {
    double dt = p->Tfin / p->nT;
    long n3 = 3L * m->N;
    for (int n = 0; n < p->nT; n++) {
        double t0 = dt * n, t1 = dt * (n + 1);
        for (int T = 0; T < m->N; T++) {
            double nodes[28 * 3], wts[28], ris[3] = {0, 0, 0};
            tri_gh(m, T, nGH, ghn, ghw, nodes, wts);
            for (int q = 0; q < nGH; q++) {
                double g1[3], g0[3];
                orc_datum(datum, nodes + 3 * q, m->normal + 3 * T, t1, g1);
                orc_datum(datum, nodes + 3 * q, m->normal + 3 * T, t0, g0);
                for (int k = 0; k < 3; k++) ris[k] = ris[k] + (g1[k] - g0[k]) * wts[q];
            }
            for (int k = 0; k < 3; k++) beta[3L * T + k + n3 * n] = fabs(ris[k]) < 1.0e-14 ? 0 : ris[k];
        }
    }
}

/* ---------------------------------------------------------------- solve -- */

int orc_lu(int n, double *A, int *piv) {
    for (int k = 0; k < n; k++) {
        int pr = k;
        double mx = fabs(A[k + (long)n * k]);
        for (int i = k + 1; i < n; i++) {
            double v = fabs(A[i + (long)n * k]);
            if (v > mx) { mx = v; pr = i; }
        }
        piv[k] = pr;
        if (mx == 0) return -(k + 1);
        if (pr != k)
            for (int j = 0; j < n; j++) {
                double t = A[k + (long)n * j];
                A[k + (long)n * j] = A[pr + (long)n * j];
                A[pr + (long)n * j] = t;
            }
        double d = A[k + (long)n * k];
        for (int i = k + 1; i < n; i++) A[i + (long)n * k] /= d;
        for (int j = k + 1; j < n; j++) {
            double a = A[k + (long)n * j];
            if (a == 0) continue;
            double *cj = A + (long)n * j, *ck = A + (long)n * k;
            for (int i = k + 1; i < n; i++) cj[i] -= ck[i] * a;
        }
    }
    return 0;
}

void orc_lu_solve(int n, const double *LU, const int *piv, double *b) {
    for (int k = 0; k < n; k++)
        if (piv[k] != k) { double t = b[k]; b[k] = b[piv[k]]; b[piv[k]] = t; }
    for (int j = 0; j < n; j++) {
        double v = b[j];
        if (v != 0)
            for (int i = j + 1; i < n; i++) b[i] -= LU[i + (long)n * j] * v;
    }
    for (int j = n - 1; j >= 0; j--) {
        b[j] /= LU[j + (long)n * j];
        double v = b[j];
        if (v != 0)
            for (int i = 0; i < j; i++) b[i] -= LU[i + (long)n * j] * v;
    }
}

/* SA/BEMenerg_coreSA_timeMarching.m:45-113 (history sum over l = 1..n-1 in increasing l) */
int orc_mot(int n3, int L, int nT, const double *E, const double *beta, double *alpha) {
    double *LU = (double *)malloc(sizeof(double) * (size_t)n3 * n3);
    int *piv = (int *)malloc(sizeof(int) * n3);
    double *r = (double *)malloc(sizeof(double) * n3);
    memcpy(LU, E, sizeof(double) * (size_t)n3 * n3);
    int rc = orc_lu(n3, LU, piv);
    if (rc != 0) { free(LU); free(piv); free(r); return rc; }
    for (int n = 0; n < nT; n++) {
        memcpy(r, beta + (size_t)n3 * n, sizeof(double) * n3);
        for (int l = 1; l <= n && l < L; l++) {
            const double *El = E + (size_t)l * n3 * n3;
            const double *a = alpha + (size_t)n3 * (n - l);
            for (int j = 0; j < n3; j++) {
                double v = a[j];
                if (v == 0) continue;
                const double *col = El + (size_t)n3 * j;
                for (int i = 0; i < n3; i++) r[i] -= col[i] * v;
            }
        }
        orc_lu_solve(n3, LU, piv, r);
        memcpy(alpha + (size_t)n3 * n, r, sizeof(double) * n3);
    }
    free(LU); free(piv); free(r);
    return 0;
}

/* ------------------------------------------------------ post-processing -- */

/* ppA/BEMenerg_postProcA_calcMatrixBlock.m:1-68 (deltaT argument = tau) */
void orc_pp_block(const orc_param *p, const double *x, const double *T, const double *n, double tau, double *out) {
    sdr_t s;
    sdr_compute(T, n, &s);
    mat3 R[6], ris;
    m3_zero(ris);
    if (node_integrals(p, x, T, n, &s, tau, R) != T_NONE) combine(p, tau, R, ris);
    for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++) {
            double v = ris[i][j] / (4 * PI * p->rho);
            out[i + 3 * j] = fabs(v) < 1.0e-14 ? 0 : v; /* calcMatrix.m:20 cleanup */
        }
}

/* ppA/BEMenerg_postProcA_singlePoint.m:1-62 */
void orc_pp_point(const orc_param *p, const orc_mesh *m, const double *density, const double *x, double t,
                  double *u) {
    u[0] = u[1] = u[2] = 0;
    if (t <= 0) return;
    double dt = p->Tfin / p->nT;
    int nHat = (int)ceil(t / dt) - 1;
    int N = m->N;
    long n3 = 3L * N;
    double *Mc = (double *)malloc(sizeof(double) * 9 * N);
    double *Mn = (double *)malloc(sizeof(double) * 9 * N);
    double T[9];
    for (int j = 0; j < N; j++) {
        tri_verts(m, j, T);
        orc_pp_block(p, x, T, m->normal + 3 * j, t, Mc + 9 * j);
    }
    for (int k = 1; k <= nHat; k++) {
        double diffT = t - (k * dt);
        const double *al = density + n3 * (k - 1);
        for (int j = 0; j < N; j++) {
            tri_verts(m, j, T);
            orc_pp_block(p, x, T, m->normal + 3 * j, diffT, Mn + 9 * j);
        }
        for (int j = 0; j < N; j++)
            for (int r = 0; r < 3; r++)
                for (int c = 0; c < 3; c++)
                    u[r] += (Mc[9 * j + r + 3 * c] - Mn[9 * j + r + 3 * c]) * al[3 * j + c];
        double *tmp = Mc; Mc = Mn; Mn = tmp;
    }
    const double *al = density + n3 * nHat;
    for (int j = 0; j < N; j++)
        for (int r = 0; r < 3; r++)
            for (int c = 0; c < 3; c++) u[r] += Mc[9 * j + r + 3 * c] * al[3 * j + c];
    free(Mc);
    free(Mn);
}

/* Test hook: the per-triangle record laid out as the HIP library's triangle
 * record (v0[3], L1, L3, cos, sin, e1[3], e2[3], n[3], C[9] row-major,
 * center[3], maxL, area, pad[2]) computed with the oracle's own routines. */
void orc_field_record(const double *T, const double *n, double *rec) {
    sdr_t s;
    sdr_compute(T, n, &s);
    mat3 C;
    rot_inv_matrix(&s, n, C);
    memset(rec, 0, 32 * sizeof(double));
    for (int k = 0; k < 3; k++) rec[k] = T[k];
    rec[3] = s.L1; rec[4] = s.L3; rec[5] = s.cos_d; rec[6] = s.sin_d;
    for (int k = 0; k < 3; k++) { rec[7 + k] = s.e1[k]; rec[10 + k] = s.e2[k]; rec[13 + k] = n[k]; }
    for (int i = 0; i < 3; i++) for (int j = 0; j < 3; j++) rec[16 + 3 * i + j] = C[i][j];
    for (int k = 0; k < 3; k++) rec[25 + k] = (T[k] + T[3 + k] + T[6 + k]) / 3;
}

/* ------------------------------------------- numeric inner quadrature (GHC) --
 * The reference's GPU single-layer path: functions/BEMenerg_dir/CUDAkernels/
 * kernelV.cu (off-diagonal pairs, Gauss-Hammer outer x composite
 * Gauss-Hammer inner rule on the kernel "nucleo") and
 * BEMenerg_dir_calcSubBlockDiagV.m (diagonal blocks: the field triangle split
 * at each outer node into three children, each with the doppioGauss1D rule). */

/* util/quadRules/GaussHammerComposite.m:5-52 for any nSubPart (a square). */
int orc_gh_composite(int nSub, int nNod, double *nodes, double *w) {
    double g[28][3], gw[28];
    if (gh_base(nNod, g, gw)) return -1;
    int nDiv = (int)lround(sqrt((double)nSub));
    if (nDiv < 1 || nDiv * nDiv != nSub) return -1;
    double f = 1.0 / nDiv;
    int n = 0;
    for (int i = 1; i <= nDiv; i++) {
        double zMin = (i - 1) * f, zMax = i * f;
        int nTr = 2 * (nDiv - i) + 1;
        double x[3] = {1 - zMin, 1 - zMin, (1 - zMin) - f};
        double y[3] = {0 - f, 0, 0};
        double z[3] = {0, 0, 0};
        for (int j = 1; j <= nTr; j++) {
            double x2 = x[1], x3 = x[2], y2 = y[1], y3 = y[2];
            if (j % 2 == 1) {
                x[0] = x2; x[1] = x3; x[2] = x3;
                y[0] = y2; y[1] = y3; y[2] = y3 + f;
                z[0] = zMin; z[1] = zMax; z[2] = zMin;
            } else {
                x[0] = x2; x[1] = x3; x[2] = x3 - f;
                y[0] = y2; y[1] = y3; y[2] = y3;
                z[0] = zMax; z[1] = zMin; z[2] = zMax;
            }
            for (int q = 0; q < nNod; q++) {
                double *o = nodes + 3 * (n * nNod + q);
                o[0] = g[q][0] * x[0] + g[q][1] * x[1] + g[q][2] * x[2];
                o[1] = g[q][0] * y[0] + g[q][1] * y[1] + g[q][2] * y[2];
                o[2] = g[q][0] * z[0] + g[q][1] * z[1] + g[q][2] * z[2];
                w[n * nNod + q] = (2.0 / nSub) * gw[q];
            }
            n++;
        }
    }
    return 0;
}

/* Gauss-Legendre n-point rule on (-1, 1) (util/quadRules/Gauss1D.m type 1:
 * eigenvalues of the Jacobi matrix; here Newton on P_n, same nodes and weights
 * to rounding), ascending. */
static void gauss_legendre(int n, double *x, double *w) {
    for (int i = 0; i < n; i++) {
        double r = cos(PI * (i + 0.75) / (n + 0.5));
        for (int it = 0; it < 100; it++) {
            double p0 = 1.0, p1 = r;
            for (int k = 2; k <= n; k++) {
                double p2 = ((2.0 * k - 1.0) * r * p1 - (k - 1.0) * p0) / k;
                p0 = p1;
                p1 = p2;
            }
            double dp = n * (r * p1 - p0) / (r * r - 1.0);
            double dr = p1 / dp;
            r -= dr;
            if (fabs(dr) < 1e-16) break;
        }
        double p0 = 1.0, p1 = r;
        for (int k = 2; k <= n; k++) {
            double p2 = ((2.0 * k - 1.0) * r * p1 - (k - 1.0) * p0) / k;
            p0 = p1;
            p1 = p2;
        }
        double dp = n * (r * p1 - p0) / (r * r - 1.0);
        x[n - 1 - i] = r;
        w[n - 1 - i] = 2.0 / ((1.0 - r * r) * dp * dp);
    }
}

/* doppioGauss1D (util/quadRules/doppioGauss1D.m): collapsed
 * tensor rule on the reference triangle, nodes ((1-x_j)(1-x_i), x_j(1-x_i), x_i),
 * weights 2 w_i w_j (1-x_i), x, w the Gauss-Legendre rule on (0, 1);
 * node k = i + n1 j (MATLAB column-major reshape).  nNodes a square. */
int orc_doppio_gauss(int nNodes, double *nodes, double *w) {
    int n1 = (int)lround(sqrt((double)nNodes));
    if (n1 < 1 || n1 * n1 != nNodes || n1 > 64) return -1;
    double x[64], wx[64];
    gauss_legendre(n1, x, wx);
    for (int k = 0; k < n1; k++) {
        x[k] = 0.5 + (x[k] / 2);
        wx[k] = wx[k] / 2;
    }
    for (int j = 0; j < n1; j++)
        for (int i = 0; i < n1; i++) {
            int k = i + n1 * j;
            nodes[3 * k + 0] = (1 - x[j]) * (1 - x[i]);
            nodes[3 * k + 1] = x[j] * (1 - x[i]);
            nodes[3 * k + 2] = x[i];
            w[k] = 2 * wx[j] * (wx[i] * (1 - x[i]));
        }
    return 0;
}

/* kernelV.cu:188-258 nucleo (and the anonymous function nu of
 * calcSubBlockDiagV.m:9-53): the time-integrated kernel at offset x, r = |x|. */
void orc_nucleo(const double *x, double r, double t, double cP, double cS, double *nu) {
    double HP = (t - (r / cP) > 0), HS = (t - (r / cS) > 0);
    double r3 = pow(r, 3), r5 = pow(r, 5);
    double br = (cP * cP * t * t - r * r) / (cP * cP) * HP - (cS * cS * t * t - r * r) / (cS * cS) * HS;
    for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++) {
            double xx = x[i] * x[j];
            double v;
            if (i == j)
                v = (xx / r3) * (1 / (cP * cP)) * HP + ((1 / r) - (xx / r3)) * (1 / (cS * cS)) * HS -
                    0.5 * ((1 / r3) - 3 * (xx / r5)) * br;
            else
                v = (xx / r3) * (1 / (cP * cP)) * HP + (-(xx / r3)) * (1 / (cS * cS)) * HS -
                    0.5 * (-3 * (xx / r5)) * br;
            nu[i + 3 * j] = v;
        }
}

/* kernelV.cu:158-186 isBlockNull: every vertex pair farther than c_S t_{lag-1}. */
static int ghc_block_null(const double *TS, const double *TF, double cS, double dt, int lag) {
    double dmax = 0;
    for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++) {
            double a = TS[3 * i] - TF[3 * j], b = TS[3 * i + 1] - TF[3 * j + 1], c = TS[3 * i + 2] - TF[3 * j + 2];
            double d = sqrt(a * a + b * b + c * c);
            if (d > dmax) dmax = d;
        }
    return (lag - 1) * cS * dt > dmax;
}

/* One 3x3 block of E_lag in the GHC mode (column-major out), with the
 * off-diagonal cleanup |sum| > 1e-14 before the division by 4 pi rho
 * (kernelV.cu:150-157); diagonal blocks uncleaned (calcSubBlockDiagV.m:104). */
int orc_ghc_pair_block(const orc_param *p, const orc_mesh *m, int nGH, const double *ghn, const double *ghw,
                       int nInt, const double *ghcn, const double *ghcw, int nDiag, const double *dgn,
                       const double *dgw, int lag, int S, int F, double *out) {
    double dt = p->Tfin / p->nT, cP = p->velP, cS = p->velS;
    memset(out, 0, 9 * sizeof(double));
    double TS[9], TF[9], xq[28 * 3], wq[28];
    tri_verts(m, S, TS);
    tri_verts(m, F, TF);
    tri_gh(m, S, nGH, ghn, ghw, xq, wq);
    double acc[9];
    memset(acc, 0, sizeof acc);
    if (S != F) {
        if (ghc_block_null(TS, TF, cS, dt, lag)) return 0;
        const double coef[3] = {1, -2, 1}, ct[3] = {-1, 0, 1};
        for (int l = 0; l < nGH; l++)
            for (int pi = 0; pi < nInt; pi++) {
                const double *b = ghcn + 3 * pi;
                double y[3], d[3];
                for (int k = 0; k < 3; k++) y[k] = b[0] * TF[k] + b[1] * TF[3 + k] + b[2] * TF[6 + k];
                for (int k = 0; k < 3; k++) d[k] = y[k] - xq[3 * l + k];
                double r = sqrt(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]);
                double pw = wq[l] * (ghcw[pi] * m->area[F]);
                for (int k = 0; k < 3; k++) {
                    double t = dt * (lag + ct[k]);
                    if (t < 0) continue;
                    double nu[9];
                    orc_nucleo(d, r, t, cP, cS, nu);
                    for (int e = 0; e < 9; e++) acc[e] += pw * coef[k] * nu[e];
                }
            }
        for (int e = 0; e < 9; e++)
            if (fabs(acc[e]) > 1.0e-14) out[e] = acc[e] / (4 * PI * p->rho);
        return 1;
    }
    const double coef[3] = {1, -2, 1};
    for (int v = 0; v < 3; v++) {
        int it = lag + 1 - v;
        if (it <= 0) continue;
        double t = it * dt;
        double ghsum[9];
        memset(ghsum, 0, sizeof ghsum);
        for (int l = 0; l < nGH; l++) {
            const double *P = xq + 3 * l;
            double g2[9];
            memset(g2, 0, sizeof g2);
            /* children (V1,V3,P), (V2,V1,P), (V3,V2,P): calcCostantData.m:49-55 */
            const int ca[3] = {0, 1, 2}, cb[3] = {2, 0, 1};
            for (int c = 0; c < 3; c++) {
                const double *A = TS + 3 * ca[c], *B = TS + 3 * cb[c];
                double u[3] = {A[0] - P[0], A[1] - P[1], A[2] - P[2]}, w2[3] = {B[0] - P[0], B[1] - P[1], B[2] - P[2]};
                double cr[3] = {u[1] * w2[2] - u[2] * w2[1], u[2] * w2[0] - u[0] * w2[2], u[0] * w2[1] - u[1] * w2[0]};
                double ac = sqrt(cr[0] * cr[0] + cr[1] * cr[1] + cr[2] * cr[2]) / 2;
                for (int k = 0; k < nDiag; k++) {
                    const double *b = dgn + 3 * k;
                    double y[3], d[3];
                    for (int e = 0; e < 3; e++) y[e] = b[0] * A[e] + b[1] * B[e] + b[2] * P[e];
                    for (int e = 0; e < 3; e++) d[e] = y[e] - P[e];
                    double r = sqrt(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]);
                    double nu[9];
                    orc_nucleo(d, r, t, cP, cS, nu);
                    for (int e = 0; e < 9; e++) g2[e] += (dgw[k] * ac) * nu[e];
                }
            }
            for (int e = 0; e < 9; e++) ghsum[e] += wq[l] * g2[e];
        }
        for (int e = 0; e < 9; e++) acc[e] += coef[v] * ghsum[e];
    }
    for (int e = 0; e < 9; e++) out[e] = acc[e] / (4 * PI * p->rho);
    return 1;
}

/* Rows [r0, r1) of E_lag in the GHC mode (timeMarchingD.m:97-161), 3N x 3N column-major. */
int orc_ghc_matrix_block_rows(const orc_param *p, const orc_mesh *m, int nGH, const double *ghn, const double *ghw,
                              int nInt, const double *ghcn, const double *ghcw, int nDiag, const double *dgn,
                              const double *dgw, int lag, int r0, int r1, double *out, int nthreads) {
    int N = m->N;
    long n3 = 3L * N;
    long npairs = (long)(r1 - r0) * N;
#ifdef _OPENMP
    if (nthreads < 1) nthreads = 1;
#pragma omp parallel for schedule(dynamic, 16) num_threads(nthreads)
#endif
    for (long ind = 0; ind < npairs; ind++) {
        int S = r0 + (int)(ind % (r1 - r0));
        int F = (int)(ind / (r1 - r0));
        double blk[9];
        orc_ghc_pair_block(p, m, nGH, ghn, ghw, nInt, ghcn, ghcw, nDiag, dgn, dgw, lag, S, F, blk);
        for (int i = 0; i < 3; i++)
            for (int j = 0; j < 3; j++) out[(3L * S + i) + n3 * (3L * F + j)] = blk[i + 3 * j];
    }
    (void)nthreads;
    return 0;
}

/* Numeric post-processing of the reference's GPU path
 * (BEMenerg_postProcGPU/CUDAkernels/kernelLayer.cu:23-103): M(x, tau) of one
 * triangle by the rule (nInt nodes, barycentric b0..b2 and weight), divided by
 * 4 pi rho, entries below 1e-14 zeroed; zero when c_P tau < |x - c_F| - maxL_F
 * (isBlockNull).  out column-major. */
void orc_pp_block_num(const orc_param *p, const orc_mesh *m, int F, const double *x, double tau, int nInt,
                      const double *rn, const double *rw, double *out) {
    memset(out, 0, 9 * sizeof(double));
    const double *c = m->center + 3 * F;
    double v[3] = {x[0] - c[0], x[1] - c[1], x[2] - c[2]};
    double dmin = sqrt(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]) - m->maxL[F];
    if (tau * p->velP < dmin) return;
    double T[9];
    tri_verts(m, F, T);
    double acc[9];
    memset(acc, 0, sizeof acc);
    for (int k = 0; k < nInt; k++) {
        const double *b = rn + 3 * k;
        double y[3], d[3];
        for (int e = 0; e < 3; e++) y[e] = b[0] * T[e] + b[1] * T[3 + e] + b[2] * T[6 + e];
        for (int e = 0; e < 3; e++) d[e] = y[e] - x[e];
        double r = sqrt(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]);
        double nu[9];
        orc_nucleo(d, r, tau, p->velP, p->velS, nu);
        double w = rw[k] * m->area[F];
        for (int e = 0; e < 9; e++) acc[e] += w * nu[e];
    }
    for (int e = 0; e < 9; e++) {
        double val = acc[e] / (4 * PI * p->rho);
        out[e] = fabs(val) < 1.0e-14 ? 0 : val;
    }
}

/* u(x, t) of the numeric post-processing: sum_{c=0}^{nHat} M(x, t - c dt)
 * (alpha_c - alpha_{c-1}) (kernelLayer.cu:85-103, postProcGPU_calcLayer.m). */
void orc_pp_point_num(const orc_param *p, const orc_mesh *m, const double *density, const double *x, double t,
                      int nInt, const double *rn, const double *rw, double *u) {
    u[0] = u[1] = u[2] = 0;
    if (t <= 0) return;
    double dt = p->Tfin / p->nT;
    int nHat = (int)ceil(t / dt) - 1;
    long n3 = 3L * m->N;
    for (int j = 0; j < m->N; j++)
        for (int c = 0; c <= nHat; c++) {
            double M[9];
            orc_pp_block_num(p, m, j, x, t - c * dt, nInt, rn, rw, M);
            const double *ap = density + n3 * c + 3 * j;
            double d[3];
            for (int e = 0; e < 3; e++) d[e] = ap[e] - (c > 0 ? ap[e - n3] : 0.0);
            for (int r = 0; r < 3; r++) u[r] += M[r] * d[0] + M[r + 3] * d[1] + M[r + 6] * d[2];
        }
}


/* ======================================================================
 * Double-layer operator K of the direct Dirichlet driver
 * (dir/BEMenerg_dir_timeMarchingD.m:173-274): regular blocks as
 * dirK/kernelK.cu computes them, the blocks of a triangle with its own
 * vertices as dir/BEMenerg_dir_calcSubBlockSingK.m + calcKIntgSing.m, and
 * beta_K of dir/BEMenerg_dir_calcBetaKgpu.m.
 * ====================================================================== */

/* P1 basis of triangle T (vertices row-wise, 9 doubles):
 * matCoeff = [-1 -1 0; 1 0 0; 0 1 0] / [l21, l31, n] (MATLAB mrdivide: X B = A),
 * vetCoeff = [1;0;0] - matCoeff V1, phi_k(x) = vetCoeff[k] + matCoeff[k][.] x
 * (dir/BEMenerg_dir_calcCostantData.m:16-20).  mat row-major [k][c]. */
void orc_k_basis(const double *T, double *mat, double *vec) {
    double l21[3], l31[3], n[3];
    for (int c = 0; c < 3; c++) {
        l21[c] = T[3 + c] - T[c];
        l31[c] = T[6 + c] - T[c];
    }
    n[0] = l21[1] * l31[2] - l21[2] * l31[1];
    n[1] = l21[2] * l31[0] - l21[0] * l31[2];
    n[2] = l21[0] * l31[1] - l21[1] * l31[0];
    /* B = [l21 l31 n] (columns); row k of X solves B^T x = A(k,:)^T */
    const double A[3][3] = {{-1, -1, 0}, {1, 0, 0}, {0, 1, 0}};
    for (int k = 0; k < 3; k++) {
        double M[3][4];
        for (int r = 0; r < 3; r++) {
            const double *col = r == 0 ? l21 : (r == 1 ? l31 : n);
            for (int c = 0; c < 3; c++) M[r][c] = col[c];
            M[r][3] = A[k][r];
        }
        for (int c = 0; c < 3; c++) {
            int piv = c;
            for (int r = c + 1; r < 3; r++)
                if (fabs(M[r][c]) > fabs(M[piv][c])) piv = r;
            if (piv != c)
                for (int e = 0; e < 4; e++) {
                    double tmp = M[c][e];
                    M[c][e] = M[piv][e];
                    M[piv][e] = tmp;
                }
            for (int r = c + 1; r < 3; r++) {
                double f = M[r][c] / M[c][c];
                for (int e = c; e < 4; e++) M[r][e] -= f * M[c][e];
            }
        }
        double x[3];
        for (int c = 2; c >= 0; c--) {
            double s = M[c][3];
            for (int e = c + 1; e < 3; e++) s -= M[c][e] * x[e];
            x[c] = s / M[c][c];
        }
        for (int c = 0; c < 3; c++) mat[3 * k + c] = x[c];
    }
    for (int k = 0; k < 3; k++)
        vec[k] = (k == 0 ? 1.0 : 0.0) - (mat[3 * k] * T[0] + mat[3 * k + 1] * T[1] + mat[3 * k + 2] * T[2]);
}

static const int K_EPS[3][3][3] = {{{0, 0, 0}, {0, 0, 1}, {0, -1, 0}},
                                   {{0, 0, -1}, {0, 0, 0}, {1, 0, 0}},
                                   {{0, 1, 0}, {-1, 0, 0}, {0, 0, 0}}};

/* R_j part of the kernel, nu[i][k] (kernelK.cu nucleoKRj; `typo` = 1 keeps its
 * Heaviside (r - r/cP) > 0 in the P part of the lambda-mu term, as the GPU
 * kernel has it; calcKIntgSing.m nuKsmRj has (t - r/cP) > 0). */
static void k_nuRj(const double *x, double r, double t, const orc_param *p, int j, int typo, double nu[3][3]) {
    const double cP = p->velP, cS = p->velS, lam = p->lambda, mu = p->mu, rho = p->rho;
    const double HP = (t - r / cP) > 0, HS = (t - r / cS) > 0;
    const double Hcub = (t - r / cS) * (t - r / cS) * (t + 2 * r / cS) / 6 * HS -
                        (t - r / cP) * (t - r / cP) * (t + 2 * r / cP) / 6 * HP;
    const double Dlin = ((t - r / cS) * HS / (cS * cS)) - ((t - r / cP) * HP / (cP * cP));
    const double HPt = typo ? (double)((r - r / cP) > 0) : HP;
    const double Dlam = ((t - r / cS) * HS / (cS * cS)) - ((t - r / cP) * HPt / (cP * cP));
    const double r3 = pow(r, 3), r5 = pow(r, 5);
    const double a = 2 * mu / rho, b = lam * mu / (rho * (lam + mu));
    for (int i = 0; i < 3; i++) {
        /* the two tensor components of row i: (i, i+1) and (i, i+2) in the order kernelK.cu uses */
        for (int k = 0; k < 3; k++) {
            double h = 0, d = 0;
            for (int q = 0; q < 3; q++) {
                if (q == i) continue;
                const double e = K_EPS[j][i][q];
                h += e * (3 * x[q] * x[k] / r5 - (q == k ? 1.0 / r3 : 0.0));
                d += e * (x[q] * x[k] / r3);
            }
            double v = a * h * Hcub + a * d * Dlin;
            if (i != k) v = b * K_EPS[j][k][i] / r * Dlam + v;
            nu[i][k] = v;
        }
    }
}

/* KL + KT part (kernelK.cu nucleoKL + nucleoKT with `dotterm` = 1; calcKIntgSing.m
 * nuKL + nuKT without the x.n diagonal terms with `dotterm` = 0). */
static void k_nuLT(const double *x, const double *n, double r, double t, const orc_param *p, int dotterm,
                   double nu[3][3]) {
    const double cP = p->velP, cS = p->velS, lam = p->lambda, mu = p->mu;
    const double HP = (t - r / cP) > 0, HS = (t - r / cS) > 0;
    const double fL = (t - r / cP) * HP - (t - r / cS) * HS, fT = HP / cP - HS / cS;
    const double cl = lam / (lam + mu), cm = mu / (lam + mu);
    const double r2 = pow(r, 2), r3 = pow(r, 3);
    const double xn = x[0] * n[0] + x[1] * n[1] + x[2] * n[2];
    for (int i = 0; i < 3; i++)
        for (int k = 0; k < 3; k++) {
            double L = cl * x[i] * n[k] / r3 * fL + cm * x[k] * n[i] / r3 * fL;
            double T = cl * x[i] * n[k] / r2 * fT + cm * x[k] * n[i] / r2 * fT;
            if (dotterm && i == k) {
                L += xn / r3 * (t - r / cS) * HS;
                T += xn / r2 * (HS / cS);
            }
            nu[i][k] = L + T;
        }
}

/* kernel value for outer point xo, inner point y on field triangle F with
 * normal nF and basis (mat, vec), local vertex v (0..2):
 * -sum_j (nF x grad phi_v)_j R_j + phi_v(y) (KL + KT) */
static void k_kernel(const double *xo, const double *y, const double *nF, const double *mat, const double *vec, int v,
                     double t, const orc_param *p, int regular, double out[3][3]) {
    double x[3] = {xo[0] - y[0], xo[1] - y[1], xo[2] - y[2]};
    double r = sqrt(x[0] * x[0] + x[1] * x[1] + x[2] * x[2]);
    const double *g = mat + 3 * v;
    double w[3] = {nF[1] * g[2] - nF[2] * g[1], -nF[0] * g[2] + nF[2] * g[0], nF[0] * g[1] - nF[1] * g[0]};
    double acc[3][3] = {{0}};
    for (int j = 0; j < 3; j++) {
        double R[3][3];
        k_nuRj(x, r, t, p, j, regular, R);
        for (int i = 0; i < 3; i++)
            for (int k = 0; k < 3; k++) acc[i][k] -= w[j] * R[i][k];
    }
    double LT[3][3];
    k_nuLT(x, nF, r, t, p, regular, LT);
    double phi = vec[v] + g[0] * y[0] + g[1] * y[1] + g[2] * y[2];
    for (int i = 0; i < 3; i++)
        for (int k = 0; k < 3; k++) out[i][k] = acc[i][k] + LT[i][k] * phi;
}

/* isBlockNull of kernelK.cu: test triangle S against vertex s */
static int k_block_null(const orc_mesh *m, int S, int s, int lag, double dt, double cP, double cS, double maxLen) {
    double dmax = 0, dmin = 100;
    const double *P = m->coords + 3 * s;
    for (int v = 0; v < 3; v++) {
        const double *V = m->coords + 3 * m->tri[3 * S + v];
        double d = sqrt((V[0] - P[0]) * (V[0] - P[0]) + (V[1] - P[1]) * (V[1] - P[1]) + (V[2] - P[2]) * (V[2] - P[2]));
        if (d > dmax) dmax = d;
        if (d < dmin) dmin = d;
    }
    return ((lag - 2) * cS * dt > dmax + maxLen) || ((lag + 1) * cP * dt < dmin - 2 * maxLen);
}

/* Block (test triangle S, vertex s) of K_lag, 3x3 column-major.  `prune` = 1
 * applies the reference's skip tests (isBlockNull, calcSubBlockSingK.m:12);
 * 0 integrates every block (used to show those tests drop zeros only). */
int orc_k_block(const orc_param *p, const orc_mesh *m, int nGH, const double *ghn, const double *ghw, int nInt,
                const double *ghcn, const double *ghcw, int nDiag, const double *dgn, const double *dgw, int lag,
                int S, int s, int prune, double *out) {
    const double dt = p->Tfin / p->nT, cP = p->velP, cS = p->velS;
    memset(out, 0, 9 * sizeof(double));
    double maxLen = 0;
    for (int t = 0; t < m->N; t++)
        if (m->maxL[t] > maxLen) maxLen = m->maxL[t];
    double xq[28 * 3], wq[28];
    tri_gh(m, S, nGH, ghn, ghw, xq, wq);
    const int ct[4] = {-2, -1, 0, 1};
    const double cn[4] = {-1, 3, -3, 1};
    int any = 0;
    /* regular part: field triangles F != S with vertex s (kernelK.cu:63-160) */
    if (!prune || !k_block_null(m, S, s, lag, dt, cP, cS, maxLen)) {
        double acc[3][3] = {{0}};
        for (int F = 0; F < m->N; F++) {
            if (F == S) continue;
            int v = -1;
            for (int k = 0; k < 3; k++)
                if (m->tri[3 * F + k] == s) v = k;
            if (v < 0) continue;
            double TF[9], mat[9], vec[3];
            tri_verts(m, F, TF);
            orc_k_basis(TF, mat, vec);
            const double *nF = m->normal + 3 * F;
            for (int pi = 0; pi < nInt; pi++) {
                const double *b = ghcn + 3 * pi;
                double y[3];
                for (int e = 0; e < 3; e++) y[e] = b[0] * TF[e] + b[1] * TF[3 + e] + b[2] * TF[6 + e];
                const double wi = ghcw[pi] * m->area[F];
                for (int l = 0; l < nGH; l++)
                    for (int k = 0; k < 4; k++) {
                        const double t = dt * (lag + ct[k]);
                        if (t <= 0) continue;
                        double K[3][3];
                        k_kernel(xq + 3 * l, y, nF, mat, vec, v, t, p, 1, K);
                        for (int i = 0; i < 3; i++)
                            for (int j = 0; j < 3; j++) acc[i][j] += wq[l] * wi * cn[k] * K[i][j];
                    }
            }
        }
        for (int i = 0; i < 3; i++)
            for (int j = 0; j < 3; j++)
                if (fabs(acc[i][j]) > 1.0e-14) out[i + 3 * j] = acc[i][j] / (4 * PI * dt);
        any = 1;
    }
    /* the test triangle's own vertex: calcSubBlockSingK.m + calcKIntgSing.m */
    int vS = -1;
    for (int k = 0; k < 3; k++)
        if (m->tri[3 * S + k] == s) vS = k;
    if (vS >= 0 && (!prune || !(lag >= (int)ceil(m->maxL[S] / (cS * dt)) + 2))) {
        double TS[9], mat[9], vec[3];
        tri_verts(m, S, TS);
        orc_k_basis(TS, mat, vec);
        const double *nS = m->normal + 3 * S;
        double blk[3][3] = {{0}};
        for (int k = 0; k < 4; k++) {
            const int it = lag + ct[k];
            if (it <= 0) continue;
            const double t = it * dt;
            double tot[3][3] = {{0}};
            for (int l = 0; l < nGH; l++) {
                const double *P = xq + 3 * l;
                double in[3][3] = {{0}};
                /* children [V_c, V_{mod(c+1,3)+1}, P] (calcCostantData.m:49-55) */
                const int ca[3] = {0, 1, 2}, cb[3] = {2, 0, 1};
                for (int c = 0; c < 3; c++) {
                    const double *A = TS + 3 * ca[c], *B = TS + 3 * cb[c];
                    double u[3] = {A[0] - P[0], A[1] - P[1], A[2] - P[2]}, w2[3] = {B[0] - P[0], B[1] - P[1], B[2] - P[2]};
                    double cr[3] = {u[1] * w2[2] - u[2] * w2[1], u[2] * w2[0] - u[0] * w2[2], u[0] * w2[1] - u[1] * w2[0]};
                    double ac = sqrt(cr[0] * cr[0] + cr[1] * cr[1] + cr[2] * cr[2]) / 2;
                    for (int d = 0; d < nDiag; d++) {
                        const double *b = dgn + 3 * d;
                        double y[3];
                        for (int e = 0; e < 3; e++) y[e] = b[0] * A[e] + b[1] * B[e] + b[2] * P[e];
                        double K[3][3];
                        k_kernel(P, y, nS, mat, vec, vS, t, p, 0, K);
                        for (int i = 0; i < 3; i++)
                            for (int j = 0; j < 3; j++) in[i][j] += (dgw[d] * ac) * K[i][j];
                    }
                }
                for (int i = 0; i < 3; i++)
                    for (int j = 0; j < 3; j++) tot[i][j] += wq[l] * in[i][j];
            }
            for (int i = 0; i < 3; i++)
                for (int j = 0; j < 3; j++) blk[i][j] += cn[k] * tot[i][j];
        }
        for (int i = 0; i < 3; i++)
            for (int j = 0; j < 3; j++) out[i + 3 * j] += blk[i][j] / (4 * PI * dt);
        any = 1;
    }
    return any;
}

/* K_lag for test rows [r0, r1) as a dense 3N x 3Nv column-major matrix. */
int orc_k_matrix_rows(const orc_param *p, const orc_mesh *m, int nGH, const double *ghn, const double *ghw, int nInt,
                      const double *ghcn, const double *ghcw, int nDiag, const double *dgn, const double *dgw,
                      int lag, int r0, int r1, double *out, int nthreads) {
    const long n3 = 3L * m->N;
    memset(out, 0, sizeof(double) * n3 * 3L * m->Nv);
#pragma omp parallel for schedule(dynamic, 1) num_threads(nthreads > 0 ? nthreads : 1)
    for (long q = 0; q < (long)(r1 - r0) * m->Nv; q++) {
        int S = r0 + (int)(q / m->Nv), s = (int)(q % m->Nv);
        double blk[9];
        orc_k_block(p, m, nGH, ghn, ghw, nInt, ghcn, ghcw, nDiag, dgn, dgw, lag, S, s, 1, blk);
        for (int i = 0; i < 3; i++)
            for (int j = 0; j < 3; j++) out[(3L * S + i) + n3 * (3L * s + j)] = blk[i + 3 * j];
    }
    return 0;
}

/* ======================================================================
 * Direct post-processing u(x, t) = sum_y G_y alpha_y - P_y g_y
 * (dirPP/BEMenerg_ppD_calc.m:1-124 with CUDAkernels/kernelG.cu and kernelP.cu).
 * ====================================================================== */

/* kernelP.cu nucleo (:169-246): the time-integrated traction kernel at
 * offset x = source - node, r = |x|, normal n, time t (before the basis
 * function), 3x3 in nu[i][j]. */
static void ppP_nucleo(const double *x, const double *n, double r, double t, const orc_param *p, double nu[3][3]) {
    const double lam = p->lambda, mu = p->mu, cP = p->velP, cS = p->velS;
    const double HP = (t - (r / cP)) > 0, HS = (t - (r / cS)) > 0;
    const double dot = x[0] * n[0] + x[1] * n[1] + x[2] * n[2];
    const double r2 = pow(r, 2), r3 = pow(r, 3), r4 = pow(r, 4), r5 = pow(r, 5), r7 = pow(r, 7);
    const double Hc = (t - (r / cP)) * (t - (r / cP)) * (t + (2 * r / cP)) / 6 * HP -
                      (t - (r / cS)) * (t - (r / cS)) * (t + (2 * r / cS)) / 6 * HS;
    for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++) {
            const int d = i == j;
            const double A = ((lam * x[i] * n[j] / r2) + (2 * mu * x[j] * x[i] / r4 * dot)) * (HP / pow(cP, 3));
            const double b2 = d ? ((1 / r2) - (2 * x[j] * x[i] / r4)) : (-(2 * x[j] * x[i] / r4));
            const double B = mu * ((x[j] * n[i] / r2) + (b2 * dot)) * (HS / pow(cS, 3));
            const double c2 = d ? ((1 / r3) - (6 * x[j] * x[i] / r5)) : (-(6 * x[j] * x[i] / r5));
            const double C = (((lam - 2 * mu) * x[i] * n[j] / r3) - (2 * mu * x[j] * n[i] / r3) - (2 * mu * c2 * dot)) *
                             ((t - (r / cP)) * HP / pow(cP, 2));
            const double d2 = d ? ((1 / r3) - (4 * x[j] * x[i] / r5)) : (-(4 * x[j] * x[i] / r5));
            const double D = mu * ((2 * x[i] * n[j] / r3) + (3 * x[j] * n[i] / r3) + (3 * d2 * dot)) *
                             ((t - (r / cS)) * HS / pow(cS, 2));
            const double e2 = d ? ((1 / r5) - (5 * x[j] * x[i] / r7)) : (-(5 * x[j] * x[i] / r7));
            const double E = 6 * mu * ((x[i] * n[j] / r5) + (x[j] * n[i] / r5) + (e2 * dot)) * Hc;
            nu[i][j] = A + B + C + D - E;
        }
}

/* P_y block (3x3 column-major) of vertex s for the point x at time t:
 * levels t - (y + k) dt, k = 0, 1, 2, weights 1, -2, 1 (non-positive skipped),
 * every triangle with vertex s, the composite rule (rn barycentric, rw
 * weights) scaled by the triangle's area, kernel times the hat function of s;
 * entries cleaned below 1e-14 and divided by 4 pi rho dt (kernelP.cu:7-147). */
void orc_ppP_block(const orc_param *p, const orc_mesh *m, int nInt, const double *rn, const double *rw,
                   const double *x, double t, int y, int s, double *out) {
    const double dt = p->Tfin / p->nT;
    double acc[3][3] = {{0}};
    const double cn[3] = {1, -2, 1};
    for (int k = 0; k < 3; k++) {
        const double tau = t - (double)(y + k) * dt;
        if (tau <= 0) continue;
        for (int F = 0; F < m->N; F++) {
            int v = -1;
            for (int q = 0; q < 3; q++)
                if (m->tri[3 * F + q] == s) v = q;
            if (v < 0) continue;
            double T[9], mat[9], vec[3];
            tri_verts(m, F, T);
            orc_k_basis(T, mat, vec);
            const double *n = m->normal + 3 * F;
            for (int q = 0; q < nInt; q++) {
                const double *b = rn + 3 * q;
                double yv[3], xv[3];
                for (int e = 0; e < 3; e++) yv[e] = b[0] * T[e] + b[1] * T[3 + e] + b[2] * T[6 + e];
                const double w = rw[q] * m->area[F];
                for (int e = 0; e < 3; e++) xv[e] = x[e] - yv[e];
                const double r = sqrt(xv[0] * xv[0] + xv[1] * xv[1] + xv[2] * xv[2]);
                double nu[3][3];
                ppP_nucleo(xv, n, r, tau, p, nu);
                const double phi = vec[v] + mat[3 * v] * yv[0] + mat[3 * v + 1] * yv[1] + mat[3 * v + 2] * yv[2];
                for (int i = 0; i < 3; i++)
                    for (int j = 0; j < 3; j++) acc[i][j] += cn[k] * w * (nu[i][j] * phi);
            }
        }
    }
    for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++)
            out[i + 3 * j] = fabs(acc[i][j]) > 1.0e-14 ? acc[i][j] / (4 * PI * p->rho * dt) : 0.0;
}

/* G_y block (3x3 column-major) of triangle F: levels t - (y + k) dt, k = 0, 1,
 * weights 1, -1, the single-layer kernel, cleaned below 1e-14 and divided by
 * 4 pi rho (kernelG.cu:4-116). */
void orc_ppG_block(const orc_param *p, const orc_mesh *m, int nInt, const double *rn, const double *rw,
                   const double *x, double t, int y, int F, double *out) {
    const double dt = p->Tfin / p->nT;
    double acc[9] = {0};
    const double cn[2] = {1, -1};
    double T[9];
    tri_verts(m, F, T);
    for (int k = 0; k < 2; k++) {
        const double tau = t - (double)(y + k) * dt;
        if (tau <= 0) continue;
        for (int q = 0; q < nInt; q++) {
            const double *b = rn + 3 * q;
            double d[3];
            for (int e = 0; e < 3; e++) d[e] = (b[0] * T[e] + b[1] * T[3 + e] + b[2] * T[6 + e]) - x[e];
            const double r = sqrt(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]);
            double nu[9];
            orc_nucleo(d, r, tau, p->velP, p->velS, nu);
            const double w = rw[q] * m->area[F];
            for (int e = 0; e < 9; e++) acc[e] += cn[k] * w * nu[e];
        }
    }
    for (int e = 0; e < 9; e++) out[e] = fabs(acc[e]) > 1.0e-14 ? acc[e] / (4 * PI * p->rho) : 0.0;
}

/* u(x, t) = sum_{y < nHat} G_y alpha_y - P_y g_y, nHat = ceil(t / dt); density
 * 3N x nT and g 3Nv x nT column-major (column y at time (y + 1) dt); u = 0 for
 * t <= 0 (BEMenerg_ppD_calc.m:9-13, 113-123).  OpenMP over the blocks. */
void orc_ppD_point(const orc_param *p, const orc_mesh *m, const double *density, const double *g, const double *x,
                   double t, int nInt, const double *rn, const double *rw, double *u, int nthreads) {
    u[0] = u[1] = u[2] = 0;
    if (t <= 0) return;
    const double dt = p->Tfin / p->nT;
    const int nHat = (int)ceil(t / dt);
    const long n3 = 3L * m->N, nv3 = 3L * m->Nv;
    const long nG = (long)m->N * nHat, nP = (long)m->Nv * nHat;
    double *part = (double *)calloc(3 * (size_t)(nG + nP), sizeof(double));
#pragma omp parallel for schedule(dynamic, 4) num_threads(nthreads > 0 ? nthreads : 1)
    for (long q = 0; q < nG + nP; q++) {
        double B[9];
        const double *v;
        if (q < nG) {
            const int y = (int)(q / m->N), F = (int)(q % m->N);
            orc_ppG_block(p, m, nInt, rn, rw, x, t, y, F, B);
            v = density + n3 * y + 3 * F;
        } else {
            const int y = (int)((q - nG) / m->Nv), s = (int)((q - nG) % m->Nv);
            orc_ppP_block(p, m, nInt, rn, rw, x, t, y, s, B);
            v = g + nv3 * y + 3 * s;
        }
        for (int r = 0; r < 3; r++) part[3 * q + r] = B[r] * v[0] + B[r + 3] * v[1] + B[r + 6] * v[2];
    }
    for (long q = 0; q < nG + nP; q++)
        for (int r = 0; r < 3; r++) u[r] += (q < nG ? 1.0 : -1.0) * part[3 * q + r];
    free(part);
}
