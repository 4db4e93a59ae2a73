/* Test-only driver: runs the CPU oracle end to end on a small flat screen
 * (4 x 4 squares, 32 triangles, 8 steps) and on the region hook, built with
 * AddressSanitizer and UndefinedBehaviorSanitizer by tests/test_sanitizers.py.
 * Exits 0 when every call returns normally; the sanitizers abort otherwise. */
#include <math.h>
#include <stdio.h>
#include <stdlib.h>

#include "../../oracle/ebem_oracle.h"

int main(void) {
    const int n = 4, Nv = (n + 1) * (n + 1), N = 2 * n * n;
    double *X = malloc(sizeof(double) * 3 * Nv);
    int *T = malloc(sizeof(int) * 3 * N);
    for (int j = 0; j <= n; j++)
        for (int i = 0; i <= n; i++) {
            double *v = X + 3 * (j * (n + 1) + i);
            v[0] = -0.5 + (double)i / n;
            v[1] = -0.5 + (double)j / n;
            v[2] = 0.0;
        }
    int t = 0;
    for (int j = 0; j < n; j++)
        for (int i = 0; i < n; i++) {
            int a = j * (n + 1) + i, b = a + 1, c = a + n + 1, d = c + 1;
            T[3 * t] = a; T[3 * t + 1] = b; T[3 * t + 2] = d; t++;
            T[3 * t] = a; T[3 * t + 1] = d; T[3 * t + 2] = c; t++;
        }
    double *nrm = malloc(sizeof(double) * 3 * N), *area = malloc(sizeof(double) * N);
    double *ctr = malloc(sizeof(double) * 3 * N), *maxL = malloc(sizeof(double) * N);
    orc_mesh_derive(N, X, T, nrm, area, ctr, maxL);
    orc_mesh m = {N, Nv, X, T, nrm, area, ctr, maxL};
    orc_param p = {1.0, 1.0, 1.0, sqrt(3.0), 1.0, 0.5, 8};
    double ghn[7 * 3], ghw[7];
    if (orc_gauss_hammer(7, ghn, ghw)) return 1;
    int L = orc_num_blocks(&p, Nv, X);
    const int n3 = 3 * N;
    double *E = calloc((size_t)L * n3 * n3, sizeof(double));
    for (int l = 0; l < L; l++) orc_matrix_block(&p, &m, 7, ghn, ghw, l, E + (size_t)l * n3 * n3, 1);
    double *beta = malloc(sizeof(double) * n3 * p.nT), *alpha = malloc(sizeof(double) * n3 * p.nT);
    orc_rhs(&p, &m, 7, ghn, ghw, 0, beta);
    if (orc_mot(n3, L, p.nT, E, beta, alpha)) return 2;
    double x[3] = {0.1, -0.05, 0.2}, u[3];
    orc_pp_point(&p, &m, alpha, x, 0.43, u);
    /* region hook over every region routine */
    int kind[256];
    double *args = malloc(sizeof(double) * 256 * 8), *info = malloc(sizeof(double) * 256 * 12);
    double *B = malloc(sizeof(double) * 256 * 54), *ris = malloc(sizeof(double) * 256 * 36);
    double Tt[9];
    for (int k = 0; k < 3; k++)
        for (int c = 0; c < 3; c++) Tt[3 * k + c] = X[3 * T[3 * 5 + k] + c];
    for (int s = 0; s < 20; s++) {
        double y[3] = {0.3 * sin(s), 0.2 * cos(3.0 * s), 0.05 * (s % 3)};
        int k = orc_region_trace(&p, y, Tt, nrm + 15, 0.05 + 0.02 * s, 256, kind, args, info, B, ris);
        for (int c = 0; c < k && c < 256; c++) orc_intg(kind[c], args + 8 * c, info + 12 * c, B + 54 * c, ris + 36 * c);
    }
    printf("oracle ok: L=%d u=(%g, %g, %g)\n", L, u[0], u[1], u[2]);
    free(X); free(T); free(nrm); free(area); free(ctr); free(maxL); free(E); free(beta); free(alpha);
    free(args); free(info); free(B); free(ris);
    return 0;
}
