// ebem_ghc.hpp -- numeric inner quadrature form of the single-layer blocks:
// the reference's GPU path (functions/BEMenerg_dir/CUDAkernels/kernelV.cu for
// pairs i != j, BEMenerg_dir_calcSubBlockDiagV.m for the diagonal blocks).
// Unlike the analytic form, a block is summed per evaluation with its three
// time levels combined (kernelV.cu:98-110), as the reference does: level sums
// shared between lags would carry the rounding of the large static part of
// each level into the small differences of late lags.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "ebem_kernels.hpp"

namespace ebem {

// kernelV.cu:188-258 (nucleo): the time-integrated elastodynamic kernel at
// offset x, r = |x|, as the 6 entries (00, 01, 02, 11, 12, 22) of the
// symmetric 3x3, same expressions and Heaviside products as the reference
// (r^3 and r^5 as products instead of pow calls).
__device__ __attribute__((always_inline)) inline void nucleo6(const double x[3], double r, double t, double cP,
                                                              double cS, double o[6]) {
#pragma clang fp contract(off)
    const double HP = (t - (r / cP) > 0) ? 1.0 : 0.0;
    const double HS = (t - (r / cS) > 0) ? 1.0 : 0.0;
    // pow(r, 3) and pow(r, 5) of the reference as products (within an ulp of pow)
    const double r2 = r * r;
    const double r3 = r2 * r, r5 = r3 * r2;
    const double br = (cP * cP * t * t - r * r) / (cP * cP) * HP - (cS * cS * t * t - r * r) / (cS * cS) * HS;
    const int ii[6] = {0, 0, 0, 1, 1, 2}, jj[6] = {0, 1, 2, 1, 2, 2};
#pragma unroll
    for (int q = 0; q < 6; q++) {
        const double xx = x[ii[q]] * x[jj[q]];
        if (ii[q] == jj[q])
            o[q] = (xx / r3) * (1 / (cP * cP)) * HP + ((1 / r) - (xx / r3)) * (1 / (cS * cS)) * HS -
                   0.5 * ((1 / r3) - 3 * (xx / r5)) * br;
        else
            o[q] = (xx / r3) * (1 / (cP * cP)) * HP + (-(xx / r3)) * (1 / (cS * cS)) * HS -
                   0.5 * (-3 * (xx / r5)) * br;
    }
}

// nucleo6 with 1/r (ir) taken once and the reciprocal squares of the wave
// speeds (icP2, icS2) as products; r / cP and r / cS (rcP, rcS) stay
// quotients, so the front tests decide as nucleo6's.  Values within a few
// ulps of nucleo6's (used by the direct post-processing and the off-diagonal
// numeric blocks).
__device__ __attribute__((always_inline)) inline void nucleo6_fast(const double x[3], double r, double ir, double t,
                                                                   double rcP, double rcS, double icP2, double icS2,
                                                                   double o[6]) {
#pragma clang fp contract(off)
    const double HP = (t - rcP > 0) ? 1.0 : 0.0;
    const double HS = (t - rcS > 0) ? 1.0 : 0.0;
    const double ir2 = ir * ir, ir3 = ir2 * ir, ir5 = ir3 * ir2;
    const double br = (t * t - r * r * icP2) * HP - (t * t - r * r * icS2) * HS;
    const int ii[6] = {0, 0, 0, 1, 1, 2}, jj[6] = {0, 1, 2, 1, 2, 2};
#pragma unroll
    for (int q = 0; q < 6; q++) {
        const double xx = x[ii[q]] * x[jj[q]];
        const double a3 = xx * ir3, a5 = xx * ir5;
        if (ii[q] == jj[q])
            o[q] = a3 * icP2 * HP + (ir - a3) * icS2 * HS - 0.5 * (ir3 - 3 * a5) * br;
        else
            o[q] = a3 * icP2 * HP + (-a3) * icS2 * HS - 0.5 * (-3 * a5) * br;
    }
}

// Lags per wave of the numeric off-diagonal blocks: consecutive lags share
// two of their three time levels.
#ifndef EBEM_GHC_CHUNK
#define EBEM_GHC_CHUNK 4
#endif
constexpr int kGhcChunk = EBEM_GHC_CHUNK;

struct GhcRules {
    const double* inner;  // nInt x 4: barycentric b0, b1, b2 and weight (composite Gauss-Hammer)
    const double* diag;   // nDiag x 4: collapsed rule (doppioGauss1D)
    int nInt, nDiag;
};

// One wave per (pair, level m) item of a row chunk (grid-stride over waves);
// the item stands for the block of lag l = m - 1 (items with l below the
// pair's window are skipped), written raw (not divided by 4 pi rho) into the
// item's slot W[item*6 ..] for the emit pass.  Lanes split the inner nodes and
// the wave reduces the six sums.  Pairs i != j: kernelV.cu:34-110 (inner node p
// of the field triangle with weight ghc_w[p] * area_F; per evaluation the
// levels t = (l-1, l, l+1) dt with weights 1, -2, 1, negative times skipped).
// Pairs i == j: calcSubBlockDiagV.m:55-100 (per level, for each outer node P
// the children (V1, V3, P), (V2, V1, P), (V3, V2, P) of calcCostantData.m:49-55
// with the collapsed rule scaled by the child's area; the level sums are then
// combined with 1, -2, 1, non-positive levels skipped).  Rules in LDS.
template <int NGH>
__global__ void __launch_bounds__(256) k_ghc_blocks(const FieldTri* __restrict__ tri, const double* __restrict__ gh,
                                                    const double* __restrict__ verts, AsmParams P, GhcRules R, int r0,
                                                    int64_t nitems, const int32_t* __restrict__ item_pair,
                                                    const int64_t* __restrict__ off, const int16_t* __restrict__ lo,
                                                    double* __restrict__ W) {
    extern __shared__ double rule[];  // (nInt + nDiag) x 4
    for (int k = threadIdx.x; k < 4 * (R.nInt + R.nDiag); k += blockDim.x)
        rule[k] = k < 4 * R.nInt ? R.inner[k] : R.diag[k - 4 * R.nInt];
    __syncthreads();
    const double* rin = rule;
    const double* rdg = rule + 4 * R.nInt;
    const int lane = threadIdx.x & 63;
    const int64_t wave = (blockIdx.x * (int64_t)blockDim.x + threadIdx.x) >> 6;
    const int64_t nwaves = ((int64_t)gridDim.x * blockDim.x) >> 6;
    for (int64_t it = wave; it < nitems; it += nwaves) {
        const int32_t p = item_pair[it];
        const int i = r0 + p / P.N, j = p % P.N;
        const int l0 = lo[p];
        const int lag0 = (l0 - 1 < 1 ? 1 : l0 - 1) + (int)(it - off[p]) - 1;
        if (lag0 < l0) continue;                      // wave-uniform
        if ((lag0 - l0) % kGhcChunk != 0) continue;   // chunk leaders only (wave-uniform)
        int nc = 1;  // the pair's items are consecutive, one lag each
        while (nc < kGhcChunk && it + nc < nitems && item_pair[it + nc] == p) nc++;
        const double* g = gh + (int64_t)i * 128;
        const double* vf = verts + (int64_t)j * 9;
        if (i != j) {
            // each level lag0-1 .. lag0+nc once, added to every lag of the chunk
            // whose three levels contain it (per lag in kernelV.cu's order; the
            // kernel in nucleo6_fast's reciprocal form)
            double acc[kGhcChunk][6];
#pragma unroll
            for (int u = 0; u < kGhcChunk; u++)
#pragma unroll
                for (int c = 0; c < 6; c++) acc[u][c] = 0.0;
            const double aF = tri[j].area;
            const double coef[3] = {1, -2, 1};
            const double icP2 = 1 / (P.cP * P.cP), icS2 = 1 / (P.cS * P.cS);
            for (int q = 0; q < NGH; q++) {
                const double x0 = g[q], x1 = g[32 + q], x2 = g[64 + q], wq = g[96 + q];
                for (int k = lane; k < R.nInt; k += 64) {
                    const double* b = rin + 4 * k;
                    double d[3];
                    d[0] = (b[0] * vf[0] + b[1] * vf[3] + b[2] * vf[6]) - x0;
                    d[1] = (b[0] * vf[1] + b[1] * vf[4] + b[2] * vf[7]) - x1;
                    d[2] = (b[0] * vf[2] + b[1] * vf[5] + b[2] * vf[8]) - x2;
                    const double r = sqrt(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]);
                    const double w = wq * (b[3] * aF);
                    const double ir = 1 / r, rcP = r / P.cP, rcS = r / P.cS;
#pragma unroll
                    for (int li = 0; li < kGhcChunk + 2; li++) {
                        if (li >= nc + 2) break;
                        const double t = P.dt * (double)(lag0 - 1 + li);
                        if (t < 0) continue;
                        double nu[6];
                        nucleo6_fast(d, r, ir, t, rcP, rcS, icP2, icS2, nu);
#pragma unroll
                        for (int u = 0; u < kGhcChunk; u++) {
                            const int v = li - u;
                            if (u >= nc || v < 0 || v > 2) continue;
#pragma unroll
                            for (int c = 0; c < 6; c++) acc[u][c] += w * coef[v] * nu[c];
                        }
                    }
                }
            }
#pragma unroll
            for (int u = 0; u < kGhcChunk; u++) {
                if (u >= nc) break;
#pragma unroll
                for (int c = 0; c < 6; c++) {
                    double a = acc[u][c];
                    for (int s = 32; s >= 1; s >>= 1) a += __shfl_xor(a, s);
                    if (lane == 0) W[(it + u) * 6 + c] = a;
                }
            }
            continue;
        }
        for (int u = 0; u < nc; u++) {
            const int lag = lag0 + u;
            double acc[6] = {0, 0, 0, 0, 0, 0};
            const int ca[3] = {0, 1, 2}, cb[3] = {2, 0, 1};
            const double coef[3] = {1, -2, 1};
            for (int v = 0; v < 3; v++) {
                const int lv = lag + 1 - v;
                if (lv <= 0) continue;
                const double t = (double)lv * P.dt;
                double lev[6] = {0, 0, 0, 0, 0, 0};
                for (int q = 0; q < NGH; q++) {
                    const double Pq[3] = {g[q], g[32 + q], g[64 + q]};
                    const double wq = g[96 + q];
                    for (int c = 0; c < 3; c++) {
                        const double* A = vf + 3 * ca[c];
                        const double* B = vf + 3 * cb[c];
                        const double u0 = A[0] - Pq[0], u1 = A[1] - Pq[1], u2 = A[2] - Pq[2];
                        const double v0 = B[0] - Pq[0], v1 = B[1] - Pq[1], v2 = B[2] - Pq[2];
                        const double c0 = u1 * v2 - u2 * v1, c1 = u2 * v0 - u0 * v2, c2 = u0 * v1 - u1 * v0;
                        const double ac = sqrt(c0 * c0 + c1 * c1 + c2 * c2) / 2;
                        for (int k = lane; k < R.nDiag; k += 64) {
                            const double* b = rdg + 4 * k;
                            double d[3];
#pragma unroll
                            for (int e = 0; e < 3; e++) d[e] = (b[0] * A[e] + b[1] * B[e] + b[2] * Pq[e]) - Pq[e];
                            const double r = sqrt(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]);
                            double nu[6];
                            nucleo6(d, r, t, P.cP, P.cS, nu);
                            const double w = wq * (b[3] * ac);
#pragma unroll
                            for (int e = 0; e < 6; e++) lev[e] += w * nu[e];
                        }
                    }
                }
#pragma unroll
                for (int e = 0; e < 6; e++) {
                    for (int s = 32; s >= 1; s >>= 1) lev[e] += __shfl_xor(lev[e], s);
                    acc[e] += coef[v] * lev[e];
                }
            }
            if (lane == 0) {
                double* o = W + (it + u) * 6;
#pragma unroll
                for (int c = 0; c < 6; c++) o[c] = acc[c];
            }
        }
    }
}

// Numeric post-processing (BEMenerg_postProcGPU/CUDAkernels/kernelLayer.cu:
// 23-83): one thread per (point, triangle, m) item of the post-processing
// windows, M_m = sum_p (w_p area_F) G(y_p - x, tau_m) over the rule in LDS,
// divided by 4 pi rho, entries below 1e-14 zeroed (kernelLayer.cu:84-93);
// tau_m as the analytic M kernel.  Writes M[item*6 ..].
__global__ void __launch_bounds__(256) k_pp_M_num(const FieldTri* __restrict__ tri, const double* __restrict__ verts,
                                                  const double* __restrict__ xs, int p0, PPParams P,
                                                  const double* __restrict__ rule_g, int nInt, int64_t nitems,
                                                  const int32_t* __restrict__ item_q, const int64_t* __restrict__ off,
                                                  const int32_t* __restrict__ ms, double fourpirho,
                                                  double* __restrict__ M) {
    extern __shared__ double rule[];  // nInt x 4
    for (int k = threadIdx.x; k < 4 * nInt; k += blockDim.x) rule[k] = rule_g[k];
    __syncthreads();
    for (int64_t it = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; it < nitems;
         it += (int64_t)gridDim.x * blockDim.x) {
        const int32_t q = item_q[it];
        const int p = p0 + q / P.N, j = q % P.N;
        const int m = ms[q] + (int)(it - off[q]);
        double tau;
        {
#pragma clang fp contract(off)
            tau = (m == P.nrep) ? P.trep : P.trep - (double)(P.nrep - m) * P.dt;
        }
        double acc[6] = {0, 0, 0, 0, 0, 0};
        if (tau > 0.0) {
            const double* vf = verts + (int64_t)j * 9;
            const double aF = tri[j].area;
            const double x0 = xs[3 * p], x1 = xs[3 * p + 1], x2 = xs[3 * p + 2];
            for (int k = 0; k < nInt; k++) {
                const double* b = rule + 4 * k;
                double d[3];
                d[0] = (b[0] * vf[0] + b[1] * vf[3] + b[2] * vf[6]) - x0;
                d[1] = (b[0] * vf[1] + b[1] * vf[4] + b[2] * vf[7]) - x1;
                d[2] = (b[0] * vf[2] + b[1] * vf[5] + b[2] * vf[8]) - x2;
                const double r = sqrt(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]);
                double nu[6];
                nucleo6(d, r, tau, P.cP, P.cS, nu);
                const double w = b[3] * aF;
#pragma unroll
                for (int c = 0; c < 6; c++) acc[c] += w * nu[c];
            }
        }
        double* o = M + it * 6;
#pragma unroll
        for (int c = 0; c < 6; c++) {
            const double v = acc[c] / fourpirho;
            o[c] = fabs(v) < 1.0e-14 ? 0.0 : v;
        }
    }
}

}  // namespace ebem
