// ebem_ppd_math.hpp -- math of the direct post-processing u = G alpha - P g
// (functions/BEMenerg_dirPP/BEMenerg_ppD_calc.m:1-124): the double-layer part
// P of CUDAkernels/kernelP.cu, shared by the device kernels (ebem_ppd.hpp)
// and host builds; no kernels here.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "ebem_k_math.hpp"

namespace ebem {

// Time-independent factors of kernelP.cu's nucleo (:169-246) at offset
// x = source - node, r = |x|, normal n: entry (i, j) is
//   A_ij HP/cP^3 + B_ij HS/cS^3 + C_ij (t - r/cP) HP/cP^2 + D_ij (t - r/cS) HS/cS^2 - E_ij Hc,
// Hc = (t - r/cP)^2 (t + 2r/cP)/6 HP - (t - r/cS)^2 (t + 2r/cS)/6 HS,
// each factor with the reference's expression.
struct PGeom {
    double A[3][3], B[3][3], C[3][3], D[3][3], E[3][3];
    double r;
    double rcP, rcS;  // r / cP, r / cS (p_geom_fast only)
};

// The wave speeds and the reciprocals of their powers for the fast forms.
struct PInv {
    double cP, cS, icP2, icS2, icP3, icS3;
};

__host__ __device__ __attribute__((always_inline)) inline PInv p_inv(double cP, double cS) {
    return PInv{cP, cS, 1 / (cP * cP), 1 / (cS * cS), 1 / (cP * cP * cP), 1 / (cS * cS * cS)};
}

__host__ __device__ __attribute__((always_inline)) inline void p_geom(const double x[3], double r, const double n[3],
                                                                      double lam, double mu, PGeom& G) {
#pragma clang fp contract(off)
    const double dot = x[0] * n[0] + x[1] * n[1] + x[2] * n[2];
    const double r2 = r * r, r3 = r2 * r, r4 = r2 * r2, r5 = r4 * r, r7 = r5 * r2;
    G.r = r;
#pragma unroll
    for (int i = 0; i < 3; i++)
#pragma unroll
        for (int j = 0; j < 3; j++) {
            const bool d = i == j;
            G.A[i][j] = (lam * x[i] * n[j] / r2) + (2 * mu * x[j] * x[i] / r4 * dot);
            const double b2 = d ? ((1 / r2) - (2 * x[j] * x[i] / r4)) : (-(2 * x[j] * x[i] / r4));
            G.B[i][j] = mu * ((x[j] * n[i] / r2) + (b2 * dot));
            const double c2 = d ? ((1 / r3) - (6 * x[j] * x[i] / r5)) : (-(6 * x[j] * x[i] / r5));
            G.C[i][j] = ((lam - 2 * mu) * x[i] * n[j] / r3) - (2 * mu * x[j] * n[i] / r3) - (2 * mu * c2 * dot);
            const double d2 = d ? ((1 / r3) - (4 * x[j] * x[i] / r5)) : (-(4 * x[j] * x[i] / r5));
            G.D[i][j] = mu * ((2 * x[i] * n[j] / r3) + (3 * x[j] * n[i] / r3) + (3 * d2 * dot));
            const double e2 = d ? ((1 / r5) - (5 * x[j] * x[i] / r7)) : (-(5 * x[j] * x[i] / r7));
            G.E[i][j] = 6 * mu * ((x[i] * n[j] / r5) + (x[j] * n[i] / r5) + (e2 * dot));
        }
}

// p_geom with the powers of 1/r taken once and multiplied (values within a
// few ulps of p_geom's quotients; used by the level-sum path).  r / cP and
// r / cS stay quotients, so every Heaviside front test decides as p_kernel's.
__host__ __device__ __attribute__((always_inline)) inline void p_geom_fast(const double x[3], double r,
                                                                           const double n[3], double lam, double mu,
                                                                           const PInv& c, PGeom& G) {
#pragma clang fp contract(off)
    const double dot = x[0] * n[0] + x[1] * n[1] + x[2] * n[2];
    const double ir = 1 / r, ir2 = ir * ir, ir3 = ir2 * ir, ir4 = ir2 * ir2, ir5 = ir4 * ir, ir7 = ir5 * ir2;
    G.r = r;
    // the quotients themselves: the front tests t - r/c > 0 decide as the reference's
    G.rcP = r / c.cP;
    G.rcS = r / c.cS;
#pragma unroll
    for (int i = 0; i < 3; i++)
#pragma unroll
        for (int j = 0; j < 3; j++) {
            const bool d = i == j;
            const double xx = x[j] * x[i];
            G.A[i][j] = (lam * x[i] * n[j] * ir2) + (2 * mu * xx * ir4 * dot);
            const double b2 = d ? (ir2 - (2 * xx * ir4)) : (-(2 * xx * ir4));
            G.B[i][j] = mu * ((x[j] * n[i] * ir2) + (b2 * dot));
            const double c2 = d ? (ir3 - (6 * xx * ir5)) : (-(6 * xx * ir5));
            G.C[i][j] = ((lam - 2 * mu) * x[i] * n[j] * ir3) - (2 * mu * x[j] * n[i] * ir3) - (2 * mu * c2 * dot);
            const double d2 = d ? (ir3 - (4 * xx * ir5)) : (-(4 * xx * ir5));
            G.D[i][j] = mu * ((2 * x[i] * n[j] * ir3) + (3 * x[j] * n[i] * ir3) + (3 * d2 * dot));
            const double e2 = d ? (ir5 - (5 * xx * ir7)) : (-(5 * xx * ir7));
            G.E[i][j] = 6 * mu * ((x[i] * n[j] * ir5) + (x[j] * n[i] * ir5) + (e2 * dot));
        }
}

// p_kernel from p_geom_fast's factors with the reciprocal powers of the wave
// speeds (no phi factor).
__host__ __device__ __attribute__((always_inline)) inline void p_kernel_fast(const PGeom& G, double t, const PInv& c,
                                                                             double out[3][3]) {
#pragma clang fp contract(off)
    const double tP = t - G.rcP, tS = t - G.rcS;
    const double HP = tP > 0 ? 1.0 : 0.0;
    const double HS = tS > 0 ? 1.0 : 0.0;
    const double f1 = HP * c.icP3, f2 = HS * c.icS3;
    const double f3 = tP * HP * c.icP2, f4 = tS * HS * c.icS2;
    const double Hc = (tP * tP * (t + 2 * G.rcP) * HP - tS * tS * (t + 2 * G.rcS) * HS) * (1.0 / 6);
#pragma unroll
    for (int i = 0; i < 3; i++)
#pragma unroll
        for (int j = 0; j < 3; j++)
            out[i][j] = G.A[i][j] * f1 + G.B[i][j] * f2 + G.C[i][j] * f3 + G.D[i][j] * f4 - G.E[i][j] * Hc;
}

// nucleo of kernelP.cu at time t times the hat function value phi.
__host__ __device__ __attribute__((always_inline)) inline void p_kernel(const PGeom& G, double t, double cP, double cS,
                                                                        double phi, double out[3][3]) {
#pragma clang fp contract(off)
    const double r = G.r;
    const double HP = (t - (r / cP)) > 0 ? 1.0 : 0.0;
    const double HS = (t - (r / cS)) > 0 ? 1.0 : 0.0;
    const double f1 = HP / (cP * cP * cP), f2 = HS / (cS * cS * cS);
    const double f3 = (t - (r / cP)) * HP / (cP * cP), f4 = (t - (r / cS)) * HS / (cS * cS);
    const double Hc = (t - (r / cP)) * (t - (r / cP)) * (t + (2 * r / cP)) / 6 * HP -
                      (t - (r / cS)) * (t - (r / cS)) * (t + (2 * r / cS)) / 6 * HS;
#pragma unroll
    for (int i = 0; i < 3; i++)
#pragma unroll
        for (int j = 0; j < 3; j++)
            out[i][j] = (G.A[i][j] * f1 + G.B[i][j] * f2 + G.C[i][j] * f3 + G.D[i][j] * f4 - G.E[i][j] * Hc) * phi;
}

struct PPDArgs {
    const KTri* kt;
    const double* verts;  // N x 9
    const int32_t* vadj_off;
    const int32_t* vadj;  // F * 4 + local vertex, F ascending per vertex
    const double* rule;   // nInt x 4: barycentric b0, b1, b2 and weight
    int nInt;
    double lam, mu, cP, cS, dt;
};

// Lane `lane` of `nl`: its part of the P_y block of vertex s for the point x at
// time t (kernelP.cu:27-107): the three levels t - (y + k) dt with weights
// 1, -2, 1 (non-positive skipped), every triangle with vertex s, the rule's
// nodes k = lane, lane + nl, ...; 3x3 row-major.
__host__ __device__ inline void ppd_p_lane(const PPDArgs& A, const double* x, double t, int y, int s, int lane, int nl,
                                           double acc[3][3]) {
#pragma clang fp contract(off)
    const double cn[3] = {1, -2, 1};
    for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++) acc[i][j] = 0.0;
    for (int e = A.vadj_off[s]; e < A.vadj_off[s + 1]; e++) {
        const int F = A.vadj[e] >> 2, v = A.vadj[e] & 3;
        const KTri& B = A.kt[F];
        const double* gv = B.mat + 3 * v;
        const double* vf = A.verts + (int64_t)F * 9;
        for (int q = lane; q < A.nInt; q += nl) {
            const double* b = A.rule + 4 * q;
            double yv[3];
            for (int c = 0; c < 3; c++) yv[c] = b[0] * vf[c] + b[1] * vf[3 + c] + b[2] * vf[6 + c];
            const double w = b[3] * B.area;
            const double xv[3] = {x[0] - yv[0], x[1] - yv[1], x[2] - yv[2]};
            const double r = sqrt(xv[0] * xv[0] + xv[1] * xv[1] + xv[2] * xv[2]);
            const double phi = B.vec[v] + gv[0] * yv[0] + gv[1] * yv[1] + gv[2] * yv[2];
            PGeom G;
            p_geom(xv, r, B.n, A.lam, A.mu, G);
#pragma unroll
            for (int k = 0; k < 3; k++) {
                const double tau = t - (double)(y + k) * A.dt;
                if (tau <= 0) continue;
                double nu[3][3];
                p_kernel(G, tau, A.cP, A.cS, phi, nu);
#pragma unroll
                for (int i = 0; i < 3; i++)
#pragma unroll
                    for (int j = 0; j < 3; j++) acc[i][j] += cn[k] * w * nu[i][j];
            }
        }
    }
}

// Time blocks per chunk of the double-layer post-processing: consecutive
// blocks share two of their three levels and every point's geometry.
#ifndef EBEM_PPD_CHUNK
#define EBEM_PPD_CHUNK 8
#endif
constexpr int kPpdChunk = EBEM_PPD_CHUNK;

// ppd_p_lane for the nA <= kPpdChunk blocks yA .. yA+nA-1 at once: each point's
// geometry once, each level t - l dt (l = yA .. yA+nA+1) once, added to every
// block whose three levels contain it with that block's weight; per block the
// additions come in ppd_p_lane's order, so acc[j] equals ppd_p_lane's for
// y = yA + j bit for bit.
__host__ __device__ inline void ppd_p_lane_chunk(const PPDArgs& A, const double* x, double t, int yA, int nA, int s,
                                                 int lane, int nl, double acc[kPpdChunk][3][3]) {
#pragma clang fp contract(off)
    const double cn[3] = {1, -2, 1};
    for (int j = 0; j < kPpdChunk; j++)
        for (int i = 0; i < 3; i++)
            for (int c = 0; c < 3; c++) acc[j][i][c] = 0.0;
    for (int e = A.vadj_off[s]; e < A.vadj_off[s + 1]; e++) {
        const int F = A.vadj[e] >> 2, v = A.vadj[e] & 3;
        const KTri& B = A.kt[F];
        const double* gv = B.mat + 3 * v;
        const double* vf = A.verts + (int64_t)F * 9;
        for (int q = lane; q < A.nInt; q += nl) {
            const double* b = A.rule + 4 * q;
            double yv[3];
            for (int c = 0; c < 3; c++) yv[c] = b[0] * vf[c] + b[1] * vf[3 + c] + b[2] * vf[6 + c];
            const double w = b[3] * B.area;
            const double xv[3] = {x[0] - yv[0], x[1] - yv[1], x[2] - yv[2]};
            const double r = sqrt(xv[0] * xv[0] + xv[1] * xv[1] + xv[2] * xv[2]);
            const double phi = B.vec[v] + gv[0] * yv[0] + gv[1] * yv[1] + gv[2] * yv[2];
            PGeom G;
            p_geom(xv, r, B.n, A.lam, A.mu, G);
#pragma unroll
            for (int li = 0; li < kPpdChunk + 2; li++) {
                if (li >= nA + 2) break;
                const double tau = t - (double)(yA + li) * A.dt;
                if (tau <= 0) continue;
                double nu[3][3];
                p_kernel(G, tau, A.cP, A.cS, phi, nu);
#pragma unroll
                for (int j = 0; j < kPpdChunk; j++) {
                    const int d = li - j;
                    if (j >= nA || d < 0 || d > 2) continue;
#pragma unroll
                    for (int i = 0; i < 3; i++)
#pragma unroll
                        for (int c = 0; c < 3; c++) acc[j][i][c] += cn[d] * w * nu[i][c];
                }
            }
        }
    }
}

// Levels per wave of the field-triangle form of the double-layer part.
#ifndef EBEM_PPF_CHUNK
#define EBEM_PPF_CHUNK 4
#endif
constexpr int kPpfChunk = EBEM_PPF_CHUNK;
// Lanes per item (a wave takes 64 / kPpfLanes items): the rules have a few
// hundred nodes, so narrower groups cut the butterflies per item.
constexpr int kPpfLanes = 16;

// Lane `lane` of `nl`: the level sums of field triangle F for the point x at
// time t, the nL <= kPpfChunk levels l = lA .. lA+nL-1 (tau = t - l dt) and
// F's three hat functions: acc[v][j][i + 3c] = sum over the rule's nodes k =
// lane, lane + nl, ... of w_k |F| nucleo(tau) phi_v (kernelP.cu:27-107 for one
// level and triangle, in the reciprocal forms p_geom_fast / p_kernel_fast);
// each node's geometry once, nucleo without phi once per level, then times
// each phi_v.  The block y of vertex s is sum_{F with vertex s} W_y -
// 2 W_{y+1} + W_{y+2}.
__host__ __device__ inline void ppf_lane(const PPDArgs& A, const double* x, double t, int F, int lA, int nL, int lane,
                                         int nl, double acc[3][kPpfChunk][9]) {
#pragma clang fp contract(off)
    for (int v = 0; v < 3; v++)
        for (int j = 0; j < kPpfChunk; j++)
            for (int e = 0; e < 9; e++) acc[v][j][e] = 0.0;
    const KTri& B = A.kt[F];
    const double* vf = A.verts + (int64_t)F * 9;
    const PInv ci = p_inv(A.cP, A.cS);
    for (int q = lane; q < A.nInt; q += nl) {
        const double* b = A.rule + 4 * q;
        double yv[3];
        for (int e = 0; e < 3; e++) yv[e] = b[0] * vf[e] + b[1] * vf[3 + e] + b[2] * vf[6 + e];
        const double w = b[3] * B.area;
        const double xv[3] = {x[0] - yv[0], x[1] - yv[1], x[2] - yv[2]};
        const double r = sqrt(xv[0] * xv[0] + xv[1] * xv[1] + xv[2] * xv[2]);
        double phi[3];
        for (int v = 0; v < 3; v++)
            phi[v] = B.vec[v] + B.mat[3 * v] * yv[0] + B.mat[3 * v + 1] * yv[1] + B.mat[3 * v + 2] * yv[2];
        PGeom G;
        p_geom_fast(xv, r, B.n, A.lam, A.mu, ci, G);
#pragma unroll
        for (int j = 0; j < kPpfChunk; j++) {
            if (j >= nL) break;
            const double tau = t - (double)(lA + j) * A.dt;
            if (tau <= 0) continue;
            double nu[3][3];
            p_kernel_fast(G, tau, ci, nu);
#pragma unroll
            for (int v = 0; v < 3; v++)
#pragma unroll
                for (int i = 0; i < 3; i++)
#pragma unroll
                    for (int c = 0; c < 3; c++) acc[v][j][i + 3 * c] += w * (nu[i][c] * phi[v]);
        }
    }
}

}  // namespace ebem
