// ebem_k_math.hpp -- double-layer operator K of the direct Dirichlet driver
// (functions/BEMenerg_dir/BEMenerg_dir_timeMarchingD.m:173-274) and its
// right-hand side beta_K = sum_j K_{n-j} g_j (BEMenerg_dir_calcBetaKgpu.m).
//
// K_lag is 3N x 3Nv: block (S, s) couples test triangle S with the P1 hat
// function of vertex s.  Its regular part (kernelK.cu:15-224) integrates the
// kernel over every field triangle F != S that has s as a vertex, with the
// outer Gauss-Hammer rule on S and the composite inner rule on F; the
// reference's skip test isBlockNull (kernelK.cu:226-260) drops it, and its
// sum is cleaned below 1e-14 before the division by 4 pi dt.  When s is a
// vertex of S, the singular part (calcSubBlockSingK.m, calcKIntgSing.m) adds
// the integral over S itself with the collapsed rule on the three children
// (V_c, V_c', P) around each outer node P, for lags below
// ceil(maxL_S / (cS dt)) + 2.  Both parts combine the four time levels
// lag-2 .. lag+1 with weights -1, 3, -3, 1 per evaluation, as the reference.
//
// Store: per (S, s) pair a window [lo, hi) of lags (the union of the regular
// and the singular lags, which is an interval), values 9 per (pair, lag)
// slot, column-major 3x3, slots row-major over pairs (S * Nv + s).
//
// This header holds the math shared by the device kernels (ebem_k.hpp) and
// host builds; it has no kernels.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "ebem_sa.hpp"

namespace ebem {

// P1 basis of a triangle: phi_v(y) = vec[v] + mat[3v..3v+2] . y
// (calcCostantData.m:16-20: matCoeff rows are the tangential gradients).
struct KTri {
    double mat[9];
    double vec[3];
    double n[3];  // unit normal (the mesh's)
    double area;
};
static_assert(sizeof(KTri) == 128, "KTri must stay 128 B");

// Basis of triangle (V0, V1, V2): grad phi_1 = (e2 x n)/|n|^2, grad phi_2 =
// (n x e1)/|n|^2, grad phi_0 = -(grad phi_1 + grad phi_2) with e1 = V1 - V0,
// e2 = V2 - V0, n = e1 x e2 (the solution of calcCostantData.m:18's
// mrdivide); vec[v] = delta_v0 - grad phi_v . V0.  `nrm`, `area`: the mesh's.
__host__ __device__ inline void k_tri_basis(const double* V0, const double* V1, const double* V2, const double* nrm,
                                            double area, KTri& B) {
#pragma clang fp contract(off)
    double e1[3], e2[3], n[3], g[3][3];
    for (int c = 0; c < 3; c++) {
        e1[c] = V1[c] - V0[c];
        e2[c] = V2[c] - V0[c];
    }
    n[0] = e1[1] * e2[2] - e1[2] * e2[1];
    n[1] = e1[2] * e2[0] - e1[0] * e2[2];
    n[2] = e1[0] * e2[1] - e1[1] * e2[0];
    const double nn = n[0] * n[0] + n[1] * n[1] + n[2] * n[2];
    g[1][0] = (e2[1] * n[2] - e2[2] * n[1]) / nn;
    g[1][1] = (e2[2] * n[0] - e2[0] * n[2]) / nn;
    g[1][2] = (e2[0] * n[1] - e2[1] * n[0]) / nn;
    g[2][0] = (n[1] * e1[2] - n[2] * e1[1]) / nn;
    g[2][1] = (n[2] * e1[0] - n[0] * e1[2]) / nn;
    g[2][2] = (n[0] * e1[1] - n[1] * e1[0]) / nn;
    for (int c = 0; c < 3; c++) g[0][c] = -(g[1][c] + g[2][c]);
    for (int v = 0; v < 3; v++) {
        for (int c = 0; c < 3; c++) B.mat[3 * v + c] = g[v][c];
        B.vec[v] = (v == 0 ? 1.0 : 0.0) - (g[v][0] * V0[0] + g[v][1] * V0[1] + g[v][2] * V0[2]);
        B.n[v] = nrm[v];
    }
    B.area = area;
}

struct KParams {
    int N, Nv, LK, nInt, nDiag;
    double dt, cP, cS, lam, mu, rho, maxLen, fourpidt;
};

// isBlockNull (kernelK.cu:226-260): test triangle S (vS: 9 doubles, vertex
// rows) against the vertex P at lag `lag`.
__host__ __device__ __attribute__((always_inline)) inline bool k_block_null(const double* vS, const double* P, int lag,
                                                                    const KParams& K) {
#pragma clang fp contract(off)
    double dmax = 0, dmin = 100;
#pragma unroll
    for (int i = 0; i < 3; i++) {
        const double a = vS[3 * i] - P[0], b = vS[3 * i + 1] - P[1], c = vS[3 * i + 2] - P[2];
        const double d = sqrt(a * a + b * b + c * c);
        if (d > dmax) dmax = d;
        if (d < dmin) dmin = d;
    }
    return ((double)(lag - 2) * K.cS * K.dt > dmax + K.maxLen) ||
           ((double)(lag + 1) * K.cP * K.dt < dmin - 2 * K.maxLen);
}

// Singular part's lag cut-off (calcSubBlockSingK.m:12).
__host__ __device__ __attribute__((always_inline)) inline int k_sing_cut(double maxL, const KParams& K) {
#pragma clang fp contract(off)
    return (int)ceil(maxL / (K.cS * K.dt)) + 2;
}

// Time-independent factors of the kernel at offset x = outer - inner, r = |x|:
// g5[q][k] = 3 x_q x_k / r^5 (- 1/r^3 if q == k), g3[q][k] = x_q x_k / r^3,
// and the KL / KT factors a3, b3 (1/r^3) and a2, b2 (1/r^2) of nucleoKL/KT.
struct KGeom {
    double g5[3][3], g3[3][3];
    double a3[3][3], b3[3][3], a2[3][3], b2[3][3];
    double d3, d2;  // (x . n) / r^3, (x . n) / r^2
    double r;
};

__host__ __device__ __attribute__((always_inline)) inline void k_geom(const double x[3], double r, const double n[3],
                                                              const KParams& K, KGeom& G) {
#pragma clang fp contract(off)
    const double r2 = r * r, r3 = r2 * r, r5 = r3 * r2;
    const double cl = K.lam / (K.lam + K.mu), cm = K.mu / (K.lam + K.mu);
    G.r = r;
#pragma unroll
    for (int q = 0; q < 3; q++)
#pragma unroll
        for (int k = 0; k < 3; k++) {
            const double t5 = 3 * x[q] * x[k] / r5;
            G.g5[q][k] = (q == k) ? t5 - 1 / r3 : t5;
            G.g3[q][k] = x[q] * x[k] / r3;
            G.a3[q][k] = cl * x[q] * n[k] / r3;
            G.b3[q][k] = cm * x[k] * n[q] / r3;
            G.a2[q][k] = cl * x[q] * n[k] / r2;
            G.b2[q][k] = cm * x[k] * n[q] / r2;
        }
    const double xn = x[0] * n[0] + x[1] * n[1] + x[2] * n[2];
    G.d3 = xn / r3;
    G.d2 = xn / r2;
}

// k_geom with the powers of 1/r taken once and multiplied (values within
// an ulp or two of k_geom's quotients; used by the level-sum path, whose
// sums already differ from the reference's order).  G.r keeps r; ir = 1/r is
// returned for the lambda-mu term.
__host__ __device__ __attribute__((always_inline)) inline void k_geom_fast(const double x[3], double r,
                                                                           const double n[3], const KParams& K,
                                                                           KGeom& G, double& ir) {
#pragma clang fp contract(off)
    ir = 1 / r;
    const double ir2 = ir * ir, ir3 = ir2 * ir, ir5 = ir3 * ir2;
    const double cl = K.lam / (K.lam + K.mu), cm = K.mu / (K.lam + K.mu);
    G.r = r;
#pragma unroll
    for (int q = 0; q < 3; q++)
#pragma unroll
        for (int k = 0; k < 3; k++) {
            const double t5 = 3 * x[q] * x[k] * ir5;
            G.g5[q][k] = (q == k) ? t5 - ir3 : t5;
            G.g3[q][k] = x[q] * x[k] * ir3;
            G.a3[q][k] = cl * x[q] * n[k] * ir3;
            G.b3[q][k] = cm * x[k] * n[q] * ir3;
            G.a2[q][k] = cl * x[q] * n[k] * ir2;
            G.b2[q][k] = cm * x[k] * n[q] * ir2;
        }
    const double xn = x[0] * n[0] + x[1] * n[1] + x[2] * n[2];
    G.d3 = xn * ir3;
    G.d2 = xn * ir2;
}

// Kernel of vertex basis function v at time t, 3x3 row-major in out[i][k]:
// -sum_j w_j R_j + phi (KL + KT), w = n x grad phi_v (kernelK.cu:296-459).
// REG: the regular form of kernelK.cu (the x.n diagonal terms of KL / KT and
// the Heaviside (r - r/cP) > 0 of its lambda-mu term); otherwise the singular
// form of calcKIntgSing.m (no x.n terms, (t - r/cP) > 0).  R_j has a zero row
// j, and each other row i one nonzero permutation entry e = eps(j, i, q),
// q = 3 - i - j; the reference's sums over q reduce to that term exactly.
template <bool REG>
__host__ __device__ __attribute__((always_inline)) inline void k_kernel(const KGeom& G, double t, const double w[3],
                                                                double phi, const KParams& K, double out[3][3]) {
#pragma clang fp contract(off)
    const double r = G.r, cP = K.cP, cS = K.cS;
    const double HP = (t - (r / cP)) > 0 ? 1.0 : 0.0;
    const double HS = (t - (r / cS)) > 0 ? 1.0 : 0.0;
    const double HPl = REG ? ((r - (r / cP)) > 0 ? 1.0 : 0.0) : HP;
    const double Hcub = (t - (r / cS)) * (t - (r / cS)) * (t + (2 * r / cS)) / 6 * HS -
                        (t - (r / cP)) * (t - (r / cP)) * (t + (2 * r / cP)) / 6 * HP;
    const double Dlin = ((t - (r / cS)) * HS / (cS * cS)) - ((t - (r / cP)) * HP / (cP * cP));
    const double Dlam = ((t - (r / cS)) * HS / (cS * cS)) - ((t - (r / cP)) * HPl / (cP * cP));
    const double a = 2 * K.mu / K.rho;
    const double b = K.lam * K.mu / (K.rho * (K.lam + K.mu));
    double tmp[3][3];
#pragma unroll
    for (int i = 0; i < 3; i++)
#pragma unroll
        for (int k = 0; k < 3; k++) tmp[i][k] = 0.0;
    // eps(j, i, q) for q = 3 - i - j, and eps(j, k, i)
#pragma unroll
    for (int j = 0; j < 3; j++) {
#pragma unroll
        for (int i = 0; i < 3; i++) {
            if (i == j) continue;
            const int q = 3 - i - j;
            // eps(j, i, q) = +1 for cyclic (j, i, q)
            const double e = ((i - j + 3) % 3 == 1) ? 1.0 : -1.0;
#pragma unroll
            for (int k = 0; k < 3; k++) {
                double R;
                const double hp = a * (e * G.g5[q][k]) * Hcub;
                const double dp = a * (e * G.g3[q][k]) * Dlin;
                if (k != i && k != j) {
                    // eps(j, k, i) with k = q: -e
                    R = b * (-e) / r * Dlam + hp + dp;
                } else {
                    R = hp + dp;
                }
                tmp[i][k] = tmp[i][k] - w[j] * R;
            }
        }
    }
    const double fL = (t - (r / cP)) * HP - (t - (r / cS)) * HS;
    const double fT = (HP / cP) - (HS / cS);
#pragma unroll
    for (int i = 0; i < 3; i++)
#pragma unroll
        for (int k = 0; k < 3; k++) {
            double L = G.a3[i][k] * fL + G.b3[i][k] * fL;
            double T = G.a2[i][k] * fT + G.b2[i][k] * fT;
            if (REG && i == k) {
                L = L + G.d3 * (t - (r / cS)) * HS;
                T = T + G.d2 * (HS / cS);
            }
            out[i][k] = tmp[i][k] + (L + T) * phi;
        }
}

// Inputs of one block (S, s, lag) shared by the device kernel and the host
// harness of the tests: rules (rin: nInt x 4, rdg: nDiag x 4, barycentric
// b0, b1, b2 and weight), outer nodes of S (g: x[32], y[32], z[32], w[32]).
struct KBlockArgs {
    const KTri* kt;
    const double* verts;  // N x 9, vertex rows
    const int32_t* vadj_off;
    const int32_t* vadj;  // F * 4 + local vertex, F ascending per vertex
    const double* rin;
    const double* rdg;
    const double* g;
    int nGH;
};

// Lane `lane` of `nl` lanes: its part of the regular sum over the field
// triangles F != S with vertex s (inner nodes k = lane, lane + nl, ...), the
// outer nodes and the four levels, 3x3 row-major (kernelK.cu:63-160).
__host__ __device__ inline void k_reg_lane(const KBlockArgs& A, const KParams& K, int S, int s, int lag, int lane,
                                           int nl, double acc[3][3]) {
#pragma clang fp contract(off)
    const int ct[4] = {-2, -1, 0, 1};
    const double cn[4] = {-1, 3, -3, 1};
    for (int i = 0; i < 3; i++)
        for (int c = 0; c < 3; c++) acc[i][c] = 0.0;
    for (int e = A.vadj_off[s]; e < A.vadj_off[s + 1]; e++) {
        const int F = A.vadj[e] >> 2, v = A.vadj[e] & 3;
        if (F == S) continue;
        const KTri& B = A.kt[F];
        const double* gv = B.mat + 3 * v;
        const double w[3] = {+B.n[1] * gv[2] - B.n[2] * gv[1], -B.n[0] * gv[2] + B.n[2] * gv[0],
                             +B.n[0] * gv[1] - B.n[1] * gv[0]};
        const double* vf = A.verts + (int64_t)F * 9;
        const double aF = B.area;
        for (int q = 0; q < A.nGH; q++) {
            const double xq[3] = {A.g[q], A.g[32 + q], A.g[64 + q]};
            const double wq = A.g[96 + q];
            for (int k = lane; k < K.nInt; k += nl) {
                const double* b = A.rin + 4 * k;
                double y[3];
                for (int c = 0; c < 3; c++) y[c] = b[0] * vf[c] + b[1] * vf[3 + c] + b[2] * vf[6 + c];
                const double pw = wq * (b[3] * aF);
                const double x[3] = {xq[0] - y[0], xq[1] - y[1], xq[2] - y[2]};
                const double r = sqrt(x[0] * x[0] + x[1] * x[1] + x[2] * x[2]);
                const double phi = B.vec[v] + gv[0] * y[0] + gv[1] * y[1] + gv[2] * y[2];
                KGeom G;
                k_geom(x, r, B.n, K, G);
#pragma unroll
                for (int l = 0; l < 4; l++) {
                    const double t = K.dt * (double)(lag + ct[l]);
                    if (t <= 0) continue;
                    double kv[3][3];
                    k_kernel<true>(G, t, w, phi, K, kv);
#pragma unroll
                    for (int i = 0; i < 3; i++)
#pragma unroll
                        for (int c = 0; c < 3; c++) acc[i][c] += pw * cn[l] * kv[i][c];
                }
            }
        }
    }
}

// Lags per wave of the regular K blocks: consecutive lags of a pair share
// three of their four time levels and every point's geometry.
#ifndef EBEM_K_CHUNK
#define EBEM_K_CHUNK 4
#endif
constexpr int kKChunk = EBEM_K_CHUNK;

// k_reg_lane for the nA <= kKChunk consecutive lags lagA .. lagA+nA-1 at
// once: each point's geometry once, each level lagA-2 .. lagA+nA once, its
// value added to every lag whose four levels contain it with that lag's
// weight.  Per lag the additions come in the order of k_reg_lane (field
// triangle, outer node, inner node, level ascending), so every acc[j] equals
// k_reg_lane's for lag lagA + j bit for bit.
__host__ __device__ inline void k_reg_lane_chunk(const KBlockArgs& A, const KParams& K, int S, int s, int lagA,
                                                 int nA, int lane, int nl, double acc[kKChunk][3][3]) {
#pragma clang fp contract(off)
    const double cn[4] = {-1, 3, -3, 1};
    for (int j = 0; j < kKChunk; j++)
        for (int i = 0; i < 3; i++)
            for (int c = 0; c < 3; c++) acc[j][i][c] = 0.0;
    for (int e = A.vadj_off[s]; e < A.vadj_off[s + 1]; e++) {
        const int F = A.vadj[e] >> 2, v = A.vadj[e] & 3;
        if (F == S) continue;
        const KTri& B = A.kt[F];
        const double* gv = B.mat + 3 * v;
        const double w[3] = {+B.n[1] * gv[2] - B.n[2] * gv[1], -B.n[0] * gv[2] + B.n[2] * gv[0],
                             +B.n[0] * gv[1] - B.n[1] * gv[0]};
        const double* vf = A.verts + (int64_t)F * 9;
        const double aF = B.area;
        for (int q = 0; q < A.nGH; q++) {
            const double xq[3] = {A.g[q], A.g[32 + q], A.g[64 + q]};
            const double wq = A.g[96 + q];
            for (int k = lane; k < K.nInt; k += nl) {
                const double* b = A.rin + 4 * k;
                double y[3];
                for (int c = 0; c < 3; c++) y[c] = b[0] * vf[c] + b[1] * vf[3 + c] + b[2] * vf[6 + c];
                const double pw = wq * (b[3] * aF);
                const double x[3] = {xq[0] - y[0], xq[1] - y[1], xq[2] - y[2]};
                const double r = sqrt(x[0] * x[0] + x[1] * x[1] + x[2] * x[2]);
                const double phi = B.vec[v] + gv[0] * y[0] + gv[1] * y[1] + gv[2] * y[2];
                KGeom G;
                k_geom(x, r, B.n, K, G);
#pragma unroll
                for (int li = 0; li < kKChunk + 3; li++) {
                    if (li >= nA + 3) break;
                    const int il = lagA - 2 + li;
                    const double t = K.dt * (double)il;
                    if (t <= 0) continue;
                    double kv[3][3];
                    k_kernel<true>(G, t, w, phi, K, kv);
#pragma unroll
                    for (int j = 0; j < kKChunk; j++) {
                        const int d = li - j;  // level index of il within lag lagA + j
                        if (j >= nA || d < 0 || d > 3) continue;
#pragma unroll
                        for (int i = 0; i < 3; i++)
#pragma unroll
                            for (int c = 0; c < 3; c++) acc[j][i][c] += pw * cn[d] * kv[i][c];
                    }
                }
            }
        }
    }
}

// Lane `lane` of `nl` lanes: its part of one level's singular sum
// sum_q w_q sum_children sum_d (w_d area_child) K (calcKIntgSing.m:142-171)
// for local vertex vl of S, time t.
__host__ __device__ inline void k_sing_lane(const KBlockArgs& A, const KParams& K, int S, int vl, double t, int lane,
                                            int nl, double lev[3][3]) {
#pragma clang fp contract(off)
    const KTri& B = A.kt[S];
    const double* vS = A.verts + (int64_t)S * 9;
    const double* gv = B.mat + 3 * vl;
    const double w[3] = {B.n[1] * gv[2] - B.n[2] * gv[1], B.n[2] * gv[0] - B.n[0] * gv[2],
                         B.n[0] * gv[1] - B.n[1] * gv[0]};
    const int ca[3] = {0, 1, 2}, cb[3] = {2, 0, 1};
    for (int i = 0; i < 3; i++)
        for (int c = 0; c < 3; c++) lev[i][c] = 0.0;
    for (int q = 0; q < A.nGH; q++) {
        const double Pq[3] = {A.g[q], A.g[32 + q], A.g[64 + q]};
        const double wq = A.g[96 + q];
        double in[3][3];
        for (int i = 0; i < 3; i++)
            for (int c = 0; c < 3; c++) in[i][c] = 0.0;
        for (int ch = 0; ch < 3; ch++) {
            const double* Av = vS + 3 * ca[ch];
            const double* Bv = vS + 3 * cb[ch];
            const double u0 = Av[0] - Pq[0], u1 = Av[1] - Pq[1], u2 = Av[2] - Pq[2];
            const double v0 = Bv[0] - Pq[0], v1 = Bv[1] - Pq[1], v2 = Bv[2] - Pq[2];
            const double c0 = u1 * v2 - u2 * v1, c1 = u2 * v0 - u0 * v2, c2 = u0 * v1 - u1 * v0;
            const double ac = sqrt(c0 * c0 + c1 * c1 + c2 * c2) / 2;
            for (int d = lane; d < K.nDiag; d += nl) {
                const double* b = A.rdg + 4 * d;
                double y[3];
                for (int e = 0; e < 3; e++) y[e] = b[0] * Av[e] + b[1] * Bv[e] + b[2] * Pq[e];
                const double x[3] = {Pq[0] - y[0], Pq[1] - y[1], Pq[2] - y[2]};
                const double r = sqrt(x[0] * x[0] + x[1] * x[1] + x[2] * x[2]);
                const double phi = B.vec[vl] + gv[0] * y[0] + gv[1] * y[1] + gv[2] * y[2];
                KGeom G;
                k_geom(x, r, B.n, K, G);
                double kv[3][3];
                k_kernel<false>(G, t, w, phi, K, kv);
                const double pw = b[3] * ac;
#pragma unroll
                for (int i = 0; i < 3; i++)
#pragma unroll
                    for (int c = 0; c < 3; c++) in[i][c] += pw * kv[i][c];
            }
        }
        for (int i = 0; i < 3; i++)
            for (int c = 0; c < 3; c++) lev[i][c] += wq * in[i][c];
    }
}

// The vertex-independent parts of the regular kernel (k_kernel<true>) at time
// t: the R_j entries R[j][i][k] (rows i == j unused) and KL + KT.  For a hat
// function with w = n x grad phi and value phi,
//   ((0 - w_0 R_0) - w_1 R_1) - w_2 R_2 + (KL + KT) phi   (terms j == i left out)
// equals k_kernel<true>'s entry bit for bit.
__host__ __device__ __attribute__((always_inline)) inline void k_kernel_parts(const KGeom& G, double t,
                                                                              const KParams& K, double R[3][3][3],
                                                                              double LT[3][3]) {
#pragma clang fp contract(off)
    const double r = G.r, cP = K.cP, cS = K.cS;
    const double HP = (t - (r / cP)) > 0 ? 1.0 : 0.0;
    const double HS = (t - (r / cS)) > 0 ? 1.0 : 0.0;
    const double HPl = (r - (r / cP)) > 0 ? 1.0 : 0.0;
    const double Hcub = (t - (r / cS)) * (t - (r / cS)) * (t + (2 * r / cS)) / 6 * HS -
                        (t - (r / cP)) * (t - (r / cP)) * (t + (2 * r / cP)) / 6 * HP;
    const double Dlin = ((t - (r / cS)) * HS / (cS * cS)) - ((t - (r / cP)) * HP / (cP * cP));
    const double Dlam = ((t - (r / cS)) * HS / (cS * cS)) - ((t - (r / cP)) * HPl / (cP * cP));
    const double a = 2 * K.mu / K.rho;
    const double b = K.lam * K.mu / (K.rho * (K.lam + K.mu));
#pragma unroll
    for (int j = 0; j < 3; j++)
#pragma unroll
        for (int i = 0; i < 3; i++) {
            if (i == j) continue;
            const int q = 3 - i - j;
            const double e = ((i - j + 3) % 3 == 1) ? 1.0 : -1.0;
#pragma unroll
            for (int k = 0; k < 3; k++) {
                const double hp = a * (e * G.g5[q][k]) * Hcub;
                const double dp = a * (e * G.g3[q][k]) * Dlin;
                R[j][i][k] = (k != i && k != j) ? b * (-e) / r * Dlam + hp + dp : hp + dp;
            }
        }
    const double fL = (t - (r / cP)) * HP - (t - (r / cS)) * HS;
    const double fT = (HP / cP) - (HS / cS);
#pragma unroll
    for (int i = 0; i < 3; i++)
#pragma unroll
        for (int k = 0; k < 3; k++) {
            double L = G.a3[i][k] * fL + G.b3[i][k] * fL;
            double T = G.a2[i][k] * fT + G.b2[i][k] * fT;
            if (i == k) {
                L = L + G.d3 * (t - (r / cS)) * HS;
                T = T + G.d2 * (HS / cS);
            }
            LT[i][k] = L + T;
        }
}

// k_kernel_parts with the divisions by the wave speeds, by 6 and by r taken
// as products with reciprocals (ir = 1/r from k_geom_fast): same terms,
// values within a few ulps; r / cP and r / cS stay quotients, so every
// Heaviside front test decides as k_kernel_parts'.
__host__ __device__ __attribute__((always_inline)) inline void k_kernel_parts_fast(const KGeom& G, double ir,
                                                                                   double t, const KParams& K,
                                                                                   double R[3][3][3],
                                                                                   double LT[3][3]) {
#pragma clang fp contract(off)
    const double r = G.r;
    const double icP = 1 / K.cP, icS = 1 / K.cS, icP2 = 1 / (K.cP * K.cP), icS2 = 1 / (K.cS * K.cS);
    // the quotients themselves: the front tests t - r/c > 0 decide as the reference's
    const double rcP = r / K.cP, rcS = r / K.cS;
    const double tP = t - rcP, tS = t - rcS;
    const double HP = tP > 0 ? 1.0 : 0.0;
    const double HS = tS > 0 ? 1.0 : 0.0;
    const double HPl = (r - rcP) > 0 ? 1.0 : 0.0;
    const double Hcub = (tS * tS * (t + 2 * rcS) * HS - tP * tP * (t + 2 * rcP) * HP) * (1.0 / 6);
    const double Dlin = tS * HS * icS2 - tP * HP * icP2;
    const double Dlam = tS * HS * icS2 - tP * HPl * icP2;
    const double a = 2 * K.mu / K.rho;
    const double bir = K.lam * K.mu / (K.rho * (K.lam + K.mu)) * ir;
#pragma unroll
    for (int j = 0; j < 3; j++)
#pragma unroll
        for (int i = 0; i < 3; i++) {
            if (i == j) continue;
            const int q = 3 - i - j;
            const double e = ((i - j + 3) % 3 == 1) ? 1.0 : -1.0;
#pragma unroll
            for (int k = 0; k < 3; k++) {
                const double hp = a * (e * G.g5[q][k]) * Hcub;
                const double dp = a * (e * G.g3[q][k]) * Dlin;
                R[j][i][k] = (k != i && k != j) ? (-e) * bir * Dlam + hp + dp : hp + dp;
            }
        }
    const double fL = tP * HP - tS * HS;
    const double fT = HP * icP - HS * icS;
#pragma unroll
    for (int i = 0; i < 3; i++)
#pragma unroll
        for (int k = 0; k < 3; k++) {
            double L = G.a3[i][k] * fL + G.b3[i][k] * fL;
            double T = G.a2[i][k] * fT + G.b2[i][k] * fT;
            if (i == k) {
                L = L + G.d3 * tS * HS;
                T = T + G.d2 * (HS * icS);
            }
            LT[i][k] = L + T;
        }
}

// Regular kernel for the three hat functions of one field triangle at once:
// out[v] equals k_kernel<true>(G, t, w[v], phi[v], K, .) bit for bit (FAST:
// to a few ulps).
// FAST: the parts from k_kernel_parts_fast (ir = 1/r), equal to a few ulps.
template <bool FAST = false>
__host__ __device__ __attribute__((always_inline)) inline void k_kernel3(const KGeom& G, double t,
                                                                         const double w[3][3], const double phi[3],
                                                                         const KParams& K, double out[3][3][3],
                                                                         double ir = 0.0) {
#pragma clang fp contract(off)
    double R[3][3][3], LT[3][3];
    if (FAST)
        k_kernel_parts_fast(G, ir, t, K, R, LT);
    else
        k_kernel_parts(G, t, K, R, LT);
#pragma unroll
    for (int v = 0; v < 3; v++)
#pragma unroll
        for (int i = 0; i < 3; i++)
#pragma unroll
            for (int k = 0; k < 3; k++) {
                double tmp = 0.0;
#pragma unroll
                for (int j = 0; j < 3; j++)
                    if (j != i) tmp = tmp - w[v][j] * R[j][i][k];
                out[v][i][k] = tmp + LT[i][k] * phi[v];
            }
}

// Time levels per wave of the field-triangle items of the regular K part
// (one: k_ksf then fits two waves per SIMD, which beats sharing each point's
// geometry between two levels at one wave).
#ifndef EBEM_KSF_CHUNK
#define EBEM_KSF_CHUNK 1
#endif
constexpr int kKsfChunk = EBEM_KSF_CHUNK;

// Positive levels of a lag window [a, b): the lags a .. b-1 use the levels
// lag-2 .. lag+1, of which those m >= 1 (t = m dt > 0) are nonzero.
__host__ __device__ __attribute__((always_inline)) inline int k_sf_mlo(int a) { return a - 2 > 1 ? a - 2 : 1; }
__host__ __device__ __attribute__((always_inline)) inline int k_sf_nlev(int a, int b) {
    return b > a ? b - k_sf_mlo(a) + 1 : 0;
}

// Lane `lane` of `nl`: the level sums of test triangle S against field
// triangle F != S for the hat functions of F's three vertices at the nL <=
// kKsfChunk levels m = mA .. mA+nL-1 (all >= 1): acc[v][l][i + 3c] = sum over
// the outer nodes of S and the inner nodes k = lane, lane + nl, ... of F of
// w_q w_k |F| times the regular kernel of vertex v at t = m dt
// (kernelK.cu:63-160 restricted to F and one level, in the reciprocal forms
// k_geom_fast / k_kernel_parts_fast); each point's geometry once.  The block
// of lag l is sum_d c_d W_{l-2+d}, c = (-1, 3, -3, 1).
__host__ __device__ inline void k_sf_lane(const KBlockArgs& A, const KParams& K, int F, int mA, int nL, int lane,
                                          int nl, double acc[3][kKsfChunk][9]) {
#pragma clang fp contract(off)
    for (int v = 0; v < 3; v++)
        for (int j = 0; j < kKsfChunk; j++)
            for (int e = 0; e < 9; e++) acc[v][j][e] = 0.0;
    const KTri& B = A.kt[F];
    double w[3][3];
    for (int v = 0; v < 3; v++) {
        const double* gv = B.mat + 3 * v;
        w[v][0] = +B.n[1] * gv[2] - B.n[2] * gv[1];
        w[v][1] = -B.n[0] * gv[2] + B.n[2] * gv[0];
        w[v][2] = +B.n[0] * gv[1] - B.n[1] * gv[0];
    }
    const double* vf = A.verts + (int64_t)F * 9;
    const double aF = B.area;
    for (int q = 0; q < A.nGH; q++) {
        const double xq[3] = {A.g[q], A.g[32 + q], A.g[64 + q]};
        const double wq = A.g[96 + q];
        for (int k = lane; k < K.nInt; k += nl) {
            const double* bb = A.rin + 4 * k;
            double y[3];
            for (int c = 0; c < 3; c++) y[c] = bb[0] * vf[c] + bb[1] * vf[3 + c] + bb[2] * vf[6 + c];
            const double pw = wq * (bb[3] * aF);
            const double x[3] = {xq[0] - y[0], xq[1] - y[1], xq[2] - y[2]};
            const double r = sqrt(x[0] * x[0] + x[1] * x[1] + x[2] * x[2]);
            double phi[3];
            for (int v = 0; v < 3; v++)
                phi[v] = B.vec[v] + B.mat[3 * v] * y[0] + B.mat[3 * v + 1] * y[1] + B.mat[3 * v + 2] * y[2];
            KGeom G;
            double ir;
            k_geom_fast(x, r, B.n, K, G, ir);
#pragma unroll
            for (int l = 0; l < kKsfChunk; l++) {
                if (l >= nL) break;
                const double t = K.dt * (double)(mA + l);
                double kv[3][3][3];
                k_kernel3<true>(G, t, w, phi, K, kv, ir);
#pragma unroll
                for (int v = 0; v < 3; v++)
#pragma unroll
                    for (int i = 0; i < 3; i++)
#pragma unroll
                        for (int c = 0; c < 3; c++) acc[v][l][i + 3 * c] += pw * kv[v][i][c];
            }
        }
    }
}

// Local vertex of s in S (or -1).
__host__ __device__ inline int k_local_vertex(const int32_t* T, int S, int s) {
    return (T[3 * S] == s) ? 0 : (T[3 * S + 1] == s) ? 1 : (T[3 * S + 2] == s) ? 2 : -1;
}

}  // namespace ebem
