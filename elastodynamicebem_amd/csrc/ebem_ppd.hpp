// ebem_ppd.hpp -- device kernels of the direct post-processing
// u(x, t) = sum_{y < nHat} G_y alpha_y - P_y g_y, nHat = ceil(t / dt)
// (functions/BEMenerg_dirPP/BEMenerg_ppD_calc.m:9-123; G as
// CUDAkernels/kernelG.cu, P as CUDAkernels/kernelP.cu, math of P in
// ebem_ppd_math.hpp).  One wave per (point-time pair q, triangle F) for G and
// per (q, vertex s) for P: the wave walks y = 0 .. nHat-1, the lanes split the
// rule's nodes, each block is butterflied, cleaned below 1e-14 and divided as
// the reference does, and the wave keeps sum_y block * vector in order; a
// last kernel sums each pair's waves in a fixed order.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "ebem_ghc.hpp"
#include "ebem_ppd_math.hpp"

namespace ebem {

struct PPDParams {
    int N, Nv, nT, nInt;
    double dt, cP, cS, lam, mu, rho, maxLen;
};

__device__ __attribute__((always_inline)) inline double wave_sum(double a) {
    for (int m = 32; m >= 1; m >>= 1) a += __shfl_xor(a, m);
    return a;
}

// nHat = ceil(t / dt) of BEMenerg_ppD_calc.m:18 (0 for t <= 0).
__device__ __attribute__((always_inline)) inline int ppd_nhat(double t, double dt) {
#pragma clang fp contract(off)
    return t <= 0 ? 0 : (int)ceil(t / dt);
}

// Single-layer part: one group of kPpfLanes lanes per (q, F), q = pair index
// of this launch (pairs q0 + q of pq: point, time index).  out[(q * N + F) * 3 + c] = sum_y (G_y alpha_y)_c.
// G_y (kernelG.cu:26-116): levels t - y dt and t - (y + 1) dt with weights 1, -1
// (non-positive skipped), clean-up below 1e-14, / (4 pi rho); each level is
// evaluated once and its lane partial kept for the next block (the reference
// combines the two levels node by node: the rounding differs, not its size),
// the kernel in nucleo6_fast's reciprocal form.  A block whose
// triangle lies beyond the P front of its later level, |x - c_F| - maxL_F >
// cP (t - y dt), is zero in every entry, and so are those of larger y.
__global__ void __launch_bounds__(256) k_ppd_G(const FieldTri* __restrict__ tri, const double* __restrict__ verts,
                                               const double* __restrict__ xs, const double* __restrict__ ts,
                                               const int32_t* __restrict__ pq, int q0, int nq, PPDParams P,
                                               const double* __restrict__ rule_g, const double* __restrict__ alpha,
                                               double* __restrict__ out) {
#pragma clang fp contract(off)
    extern __shared__ double rule[];
    for (int k = threadIdx.x; k < 4 * P.nInt; k += blockDim.x) rule[k] = rule_g[k];
    __syncthreads();
    const int lane = threadIdx.x & (kPpfLanes - 1);
    const int64_t grp = (blockIdx.x * (int64_t)blockDim.x + threadIdx.x) / kPpfLanes;
    const int64_t ngrp = ((int64_t)gridDim.x * blockDim.x) / kPpfLanes;
    const double fourpirho = 4 * kPi * P.rho;
    const double icP2 = 1 / (P.cP * P.cP), icS2 = 1 / (P.cS * P.cS);
    for (int64_t it = grp; it < (int64_t)nq * P.N; it += ngrp) {
        const int q = (int)(it / P.N), F = (int)(it % P.N);
        const int p = pq[2 * (q0 + q)], kt = pq[2 * (q0 + q) + 1];
        const double x[3] = {xs[3 * p], xs[3 * p + 1], xs[3 * p + 2]};
        const double t = ts[kt];
        const int nHat = ppd_nhat(t, P.dt);
        const double* vf = verts + (int64_t)F * 9;
        const double aF = tri[F].area;
        const double c0 = x[0] - tri[F].center[0], c1 = x[1] - tri[F].center[1], c2 = x[2] - tri[F].center[2];
        const double dmin = sqrt(c0 * c0 + c1 * c1 + c2 * c2) - tri[F].maxL;
        double u[3] = {0, 0, 0};
        // lane partial of level l: sum over its nodes of w nucleo(t - l dt)
        auto level = [&](int l, double L[6]) {
            for (int c = 0; c < 6; c++) L[c] = 0.0;
            const double tau = t - (double)l * P.dt;
            if (tau <= 0) return;
            for (int k = lane; k < P.nInt; k += kPpfLanes) {
                const double* b = rule + 4 * k;
                double d[3];
                for (int e = 0; e < 3; e++) d[e] = (b[0] * vf[e] + b[1] * vf[3 + e] + b[2] * vf[6 + e]) - x[e];
                const double r = sqrt(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]);
                const double w = b[3] * aF;
                double nu[6];
                nucleo6_fast(d, r, 1 / r, tau, r / P.cP, r / P.cS, icP2, icS2, nu);
#pragma unroll
                for (int c = 0; c < 6; c++) L[c] += w * nu[c];
            }
        };
        // levels l and l + 1 in one pass over the nodes (each node's geometry
        // once; per level the same sums, in the same order, as level())
        auto level2 = [&](int l, double L1[6], double L2[6]) {
            for (int c = 0; c < 6; c++) L1[c] = L2[c] = 0.0;
            const double tau1 = t - (double)l * P.dt, tau2 = t - (double)(l + 1) * P.dt;
            if (tau1 <= 0) return;
            for (int k = lane; k < P.nInt; k += kPpfLanes) {
                const double* b = rule + 4 * k;
                double d[3];
                for (int e = 0; e < 3; e++) d[e] = (b[0] * vf[e] + b[1] * vf[3 + e] + b[2] * vf[6 + e]) - x[e];
                const double r = sqrt(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]);
                const double w = b[3] * aF;
                const double ir = 1 / r, rcP = r / P.cP, rcS = r / P.cS;
                double nu[6];
                nucleo6_fast(d, r, ir, tau1, rcP, rcS, icP2, icS2, nu);
#pragma unroll
                for (int c = 0; c < 6; c++) L1[c] += w * nu[c];
                if (tau2 > 0) {
                    nucleo6_fast(d, r, ir, tau2, rcP, rcS, icP2, icS2, nu);
#pragma unroll
                    for (int c = 0; c < 6; c++) L2[c] += w * nu[c];
                }
            }
        };
        double Lc[6] = {0, 0, 0, 0, 0, 0}, Lnext[6] = {0, 0, 0, 0, 0, 0};
        if (nHat > 0 && !(dmin > P.cP * t * (1 + 1e-12))) level(0, Lc);
        for (int y = 0; y < nHat; y++) {
            const double t0 = t - (double)y * P.dt;
            if (dmin > P.cP * t0 * (1 + 1e-12)) break;  // group-uniform; exact zeros from here on
            double Ln[6];
            if ((y & 1) == 0)
                level2(y + 1, Ln, Lnext);  // levels y + 1 and y + 2
            else
                for (int c = 0; c < 6; c++) Ln[c] = Lnext[c];
            double G[6];
#pragma unroll
            for (int c = 0; c < 6; c++) {
                double a = Lc[c] - Ln[c];
                for (int m = kPpfLanes / 2; m >= 1; m >>= 1) a += __shfl_xor(a, m);
                G[c] = fabs(a) > 1.0e-14 ? a / fourpirho : 0.0;
                Lc[c] = Ln[c];
            }
            const double* al = alpha + (int64_t)y * 3 * P.N + 3 * F;
            const double a0 = al[0], a1 = al[1], a2 = al[2];
            u[0] += G[0] * a0 + G[1] * a1 + G[2] * a2;
            u[1] += G[1] * a0 + G[3] * a1 + G[4] * a2;
            u[2] += G[2] * a0 + G[4] * a1 + G[5] * a2;
        }
        if (lane == 0) {
            double* o = out + it * 3;
            o[0] = u[0];
            o[1] = u[1];
            o[2] = u[2];
        }
    }
}

// Double-layer part: wave per (q, s).  out[(q * Nv + s) * 3 + c] =
// sum_y (P_y g_y)_c with g_y = gT[s][y] (the datum at (y + 1) dt).  A vertex
// whose triangles all lie beyond the P front of the latest level,
// |x - X_s| - maxLen > cP (t - y dt), gives zero blocks from that y on.
__global__ void __launch_bounds__(256) k_ppd_P(const KTri* __restrict__ kt, const double* __restrict__ verts,
                                               const double* __restrict__ X, const int32_t* __restrict__ vadj_off,
                                               const int32_t* __restrict__ vadj, const double* __restrict__ xs,
                                               const double* __restrict__ ts, const int32_t* __restrict__ pq, int q0,
                                               int nq, PPDParams P, const double* __restrict__ rule_g,
                                               const double* __restrict__ gT, double* __restrict__ out) {
#pragma clang fp contract(off)
    extern __shared__ double rule[];
    for (int k = threadIdx.x; k < 4 * P.nInt; k += blockDim.x) rule[k] = rule_g[k];
    __syncthreads();
    const int lane = threadIdx.x & 63;
    const int64_t wave = (blockIdx.x * (int64_t)blockDim.x + threadIdx.x) >> 6;
    const int64_t nwaves = ((int64_t)gridDim.x * blockDim.x) >> 6;
    const double scale = 4 * kPi * P.rho * P.dt;
    PPDArgs A{kt, verts, vadj_off, vadj, rule, P.nInt, P.lam, P.mu, P.cP, P.cS, P.dt};
    for (int64_t it = wave; it < (int64_t)nq * P.Nv; it += nwaves) {
        const int q = (int)(it / P.Nv), s = (int)(it % P.Nv);
        const int p = pq[2 * (q0 + q)], kk = pq[2 * (q0 + q) + 1];
        const double x[3] = {xs[3 * p], xs[3 * p + 1], xs[3 * p + 2]};
        const double t = ts[kk];
        const int nHat = ppd_nhat(t, P.dt);
        const double d0 = x[0] - X[3 * s], d1 = x[1] - X[3 * s + 1], d2 = x[2] - X[3 * s + 2];
        const double dmin = sqrt(d0 * d0 + d1 * d1 + d2 * d2) - P.maxLen;
        double u[3] = {0, 0, 0};
        for (int y0 = 0; y0 < nHat; y0 += kPpdChunk) {
            // blocks of this chunk before the first one beyond the P front
            int nA = 0;
            while (nA < kPpdChunk && y0 + nA < nHat && !(dmin > P.cP * (t - (double)(y0 + nA) * P.dt) * (1 + 1e-12)))
                nA++;
            if (nA == 0) break;  // wave-uniform
            double acc[kPpdChunk][3][3];
            ppd_p_lane_chunk(A, x, t, y0, nA, s, lane, 64, acc);
#pragma unroll
            for (int j = 0; j < kPpdChunk; j++) {
                if (j >= nA) break;
                double B[3][3];
#pragma unroll
                for (int i = 0; i < 3; i++)
#pragma unroll
                    for (int c = 0; c < 3; c++) {
                        const double a = wave_sum(acc[j][i][c]);
                        B[i][c] = fabs(a) > 1.0e-14 ? a / scale : 0.0;
                    }
                const double* g = gT + ((int64_t)s * P.nT + y0 + j) * 3;
                const double g0 = g[0], g1 = g[1], g2 = g[2];
#pragma unroll
                for (int i = 0; i < 3; i++) u[i] += B[i][0] * g0 + B[i][1] * g1 + B[i][2] * g2;
            }
            if (nA < kPpdChunk) break;
        }
        if (lane == 0) {
            double* o = out + it * 3;
            o[0] = u[0];
            o[1] = u[1];
            o[2] = u[2];
        }
    }
}

// u of pair q0 + q = sum_F G-part - sum_s P-part, one workgroup per pair,
// fixed strided order and tree: u_out[3 * (p + npts * k) + c].
__global__ void __launch_bounds__(256) k_ppd_sum(const double* __restrict__ outG, const double* __restrict__ outP,
                                                 const int32_t* __restrict__ pq, int q0, int N, int Nv, int npts,
                                                 double* __restrict__ u_out) {
    __shared__ double red[3][256];
    const int q = blockIdx.x;
    double s[3] = {0, 0, 0};
    for (int j = threadIdx.x; j < N; j += blockDim.x)
        for (int c = 0; c < 3; c++) s[c] += outG[((int64_t)q * N + j) * 3 + c];
    for (int j = threadIdx.x; j < Nv; j += blockDim.x)
        for (int c = 0; c < 3; c++) s[c] -= outP[((int64_t)q * Nv + j) * 3 + c];
    for (int c = 0; c < 3; c++) red[c][threadIdx.x] = s[c];
    __syncthreads();
    for (int d = 128; d > 0; d >>= 1) {
        if ((int)threadIdx.x < d)
            for (int c = 0; c < 3; c++) red[c][threadIdx.x] += red[c][threadIdx.x + d];
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        const int p = pq[2 * (q0 + q)], k = pq[2 * (q0 + q) + 1];
        for (int c = 0; c < 3; c++) u_out[3 * ((int64_t)p + (int64_t)npts * k) + c] = red[c][0];
    }
}

// ---- Double-layer part by field triangle and time level --------------------
// For pair q (point x, time t, nHat = ceil(t / dt)) and triangle F the levels
// l = 0 .. nHat-1 (tau = t - l dt > 0) up to the first one whose P front has
// not reached F, |x - c_F| - maxL_F > cP tau (all its kernels vanish there and
// at every later level): lF(q, F) levels.

// Number of levels lF of pair q0 + qq against F (see above).
__device__ __attribute__((always_inline)) inline int ppf_levels(const FieldTri& T, const double* x, double t,
                                                                const PPDParams& P) {
#pragma clang fp contract(off)
    const int nHat = ppd_nhat(t, P.dt);
    const double c0 = x[0] - T.center[0], c1 = x[1] - T.center[1], c2 = x[2] - T.center[2];
    const double dmin = sqrt(c0 * c0 + c1 * c1 + c2 * c2) - T.maxL;
    int l = 0;
    while (l < nHat && !(dmin > P.cP * (t - (double)l * P.dt) * (1 + 1e-12))) l++;
    return l;
}

// One workgroup per pair q of [0, nq): lF[q * N + F] and the pair's totals
// cnt[q] = sum_F lF, cnt[nq + q] = sum_F ceil(lF / kPpfChunk).
__global__ void __launch_bounds__(256) k_ppf_win(const FieldTri* __restrict__ tri, const double* __restrict__ xs,
                                                 const double* __restrict__ ts, const int32_t* __restrict__ pq, int nq,
                                                 PPDParams P, int16_t* __restrict__ lF, int64_t* __restrict__ cnt) {
    __shared__ long long part[256], partc[256];
    const int q = blockIdx.x;
    const int p = pq[2 * q], kk = pq[2 * q + 1];
    const double x[3] = {xs[3 * p], xs[3 * p + 1], xs[3 * p + 2]};
    const double t = ts[kk];
    long long a = 0, c = 0;
    for (int F = threadIdx.x; F < P.N; F += blockDim.x) {
        const int l = ppf_levels(tri[F], x, t, P);
        lF[(int64_t)q * P.N + F] = (int16_t)l;
        a += l;
        c += (l + kPpfChunk - 1) / kPpfChunk;
    }
    part[threadIdx.x] = a;
    partc[threadIdx.x] = c;
    __syncthreads();
    for (int d = 128; d > 0; d >>= 1) {
        if ((int)threadIdx.x < d) {
            part[threadIdx.x] += part[threadIdx.x + d];
            partc[threadIdx.x] += partc[threadIdx.x + d];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        cnt[q] = part[0];
        cnt[nq + q] = partc[0];
    }
}

// One workgroup per pair qq of the batch (pairs q0 + qq): level offsets
// woff[qq * N + F] from base[qq] and the item list (itemP = qq * N + F,
// itemC = chunk) from base[n + qq], triangles in ascending order.
__global__ void __launch_bounds__(256) k_ppf_items(const int16_t* __restrict__ lF, int N, int n,
                                                   const int64_t* __restrict__ base, int64_t* __restrict__ woff,
                                                   int32_t* __restrict__ itemP, int16_t* __restrict__ itemC) {
    __shared__ long long part[256], partc[256];
    const int qq = blockIdx.x;
    const int nper = (N + blockDim.x - 1) / blockDim.x;
    const int j0 = threadIdx.x * nper;
    const int64_t pb = (int64_t)qq * N;
    long long s = 0, c = 0;
    for (int k = 0; k < nper; k++) {
        const int j = j0 + k;
        if (j < N) {
            const int l = lF[pb + j];
            s += l;
            c += (l + kPpfChunk - 1) / kPpfChunk;
        }
    }
    part[threadIdx.x] = s;
    partc[threadIdx.x] = c;
    __syncthreads();
    for (int d = 1; d < 256; d <<= 1) {
        long long v = ((int)threadIdx.x >= d) ? part[threadIdx.x - d] : 0;
        long long vc = ((int)threadIdx.x >= d) ? partc[threadIdx.x - d] : 0;
        __syncthreads();
        part[threadIdx.x] += v;
        partc[threadIdx.x] += vc;
        __syncthreads();
    }
    long long run = base[qq] + (threadIdx.x ? part[threadIdx.x - 1] : 0);
    long long runc = base[n + qq] + (threadIdx.x ? partc[threadIdx.x - 1] : 0);
    for (int k = 0; k < nper; k++) {
        const int j = j0 + k;
        if (j >= N) break;
        const int l = lF[pb + j];
        woff[pb + j] = run;
        const int nc = (l + kPpfChunk - 1) / kPpfChunk;
        for (int u = 0; u < nc; u++) {
            itemP[runc + u] = (int32_t)(pb + j);
            itemC[runc + u] = (int16_t)u;
        }
        run += l;
        runc += nc;
    }
}

// One group of kPpfLanes lanes per item: the level sums of (pair q0 + qq, F)
// for F's three vertices at the item's levels, butterflied in the group:
// W[(woff + l) * 27 + 9 v + (i + 3 c)].
// waves per SIMD that k_ppf is compiled for (registers per wave <= 512 / this)
#ifndef EBEM_PPF_WAVES
#define EBEM_PPF_WAVES 1
#endif
__global__ void __launch_bounds__(256, EBEM_PPF_WAVES) k_ppf(const KTri* __restrict__ kt, const double* __restrict__ verts,
                                             const double* __restrict__ xs, const double* __restrict__ ts,
                                             const int32_t* __restrict__ pq, int q0, PPDParams P,
                                             const double* __restrict__ rule_g, int64_t nitems,
                                             const int32_t* __restrict__ itemP, const int16_t* __restrict__ itemC,
                                             const int16_t* __restrict__ lF, const int64_t* __restrict__ woff,
                                             double* __restrict__ W) {
#pragma clang fp contract(off)
    extern __shared__ double rule[];
    for (int k = threadIdx.x; k < 4 * P.nInt; k += blockDim.x) rule[k] = rule_g[k];
    __syncthreads();
    const int lane = threadIdx.x & (kPpfLanes - 1);
    const int64_t grp = (blockIdx.x * (int64_t)blockDim.x + threadIdx.x) / kPpfLanes;
    const int64_t ngrp = ((int64_t)gridDim.x * blockDim.x) / kPpfLanes;
    PPDArgs A{kt, verts, nullptr, nullptr, rule, P.nInt, P.lam, P.mu, P.cP, P.cS, P.dt};
    for (int64_t it = grp; it < nitems; it += ngrp) {
        const int32_t qf = itemP[it];
        const int qq = qf / P.N, F = qf % P.N;
        const int p = pq[2 * (q0 + qq)], kk = pq[2 * (q0 + qq) + 1];
        const double x[3] = {xs[3 * p], xs[3 * p + 1], xs[3 * p + 2]};
        const double t = ts[kk];
        const int lA = kPpfChunk * itemC[it];
        const int nL = min(kPpfChunk, (int)lF[qf] - lA);
        double acc[3][kPpfChunk][9];
        ppf_lane(A, x, t, F, lA, nL, lane, kPpfLanes, acc);
        double* o = W + (woff[qf] + lA) * 27;
#pragma unroll
        for (int j = 0; j < kPpfChunk; j++) {
            if (j >= nL) break;
#pragma unroll
            for (int v = 0; v < 3; v++)
#pragma unroll
                for (int e = 0; e < 9; e++) {
                    double a = acc[v][j][e];
                    for (int m = kPpfLanes / 2; m >= 1; m >>= 1) a += __shfl_xor(a, m);
                    if (lane == 0) o[j * 27 + v * 9 + e] = a;
                }
        }
    }
}

// One wave per (pair q0 + qq, vertex s), lanes over the time blocks y <
// nHat: block P_y of s = sum over the triangles F with vertex s (ascending)
// of W_y - 2 W_{y+1} + W_{y+2} (levels at or past lF(q, F) vanish), cleaned
// below 1e-14 and divided by 4 pi rho dt (kernelP.cu:97-106), times g_y; the
// lanes' sums are butterflied: out[(qq * Nv + s) * 3 + c].
__global__ void __launch_bounds__(256) k_ppf_gather(const int32_t* __restrict__ vadj_off,
                                                    const int32_t* __restrict__ vadj, const double* __restrict__ ts,
                                                    const int32_t* __restrict__ pq, int q0, int n, PPDParams P,
                                                    const int16_t* __restrict__ lF, const int64_t* __restrict__ woff,
                                                    const double* __restrict__ W, const double* __restrict__ gT,
                                                    double* __restrict__ out) {
#pragma clang fp contract(off)
    const int lane = threadIdx.x & 63;
    const int64_t wave = (blockIdx.x * (int64_t)blockDim.x + threadIdx.x) >> 6;
    const int64_t nwaves = ((int64_t)gridDim.x * blockDim.x) >> 6;
    const double scale = 4 * kPi * P.rho * P.dt;
    const double cn[3] = {1, -2, 1};
    for (int64_t it = wave; it < (int64_t)n * P.Nv; it += nwaves) {
        const int qq = (int)(it / P.Nv), s = (int)(it % P.Nv);
        const double t = ts[pq[2 * (q0 + qq) + 1]];
        const int nHat = ppd_nhat(t, P.dt);
        int ymax = 0;
        for (int e = vadj_off[s]; e < vadj_off[s + 1]; e++) ymax = max(ymax, (int)lF[(int64_t)qq * P.N + (vadj[e] >> 2)]);
        ymax = min(ymax, nHat);
        double u[3] = {0, 0, 0};
        for (int y = lane; y < ymax; y += 64) {
            double B[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
            for (int e = vadj_off[s]; e < vadj_off[s + 1]; e++) {
                const int F = vadj[e] >> 2, v = vadj[e] & 3;
                const int64_t qf = (int64_t)qq * P.N + F;
                const int L = lF[qf];
                const double* Wq = W + woff[qf] * 27 + v * 9;
                double b[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
                for (int k = 0; k < 3; k++) {
                    const int l = y + k;
                    if (l >= L) break;
                    for (int c = 0; c < 9; c++) b[c] += cn[k] * Wq[(int64_t)l * 27 + c];
                }
                for (int c = 0; c < 9; c++) B[c] += b[c];
            }
            const double* g = gT + ((int64_t)s * P.nT + y) * 3;
            const double g0 = g[0], g1 = g[1], g2 = g[2];
            double Bc[9];
            for (int c = 0; c < 9; c++) Bc[c] = fabs(B[c]) > 1.0e-14 ? B[c] / scale : 0.0;
            // B[i + 3 c] is entry (i, c)
            for (int i = 0; i < 3; i++) u[i] += Bc[i] * g0 + Bc[i + 3] * g1 + Bc[i + 6] * g2;
        }
        for (int i = 0; i < 3; i++) u[i] = wave_sum(u[i]);
        if (lane == 0) {
            double* o = out + it * 3;
            o[0] = u[0];
            o[1] = u[1];
            o[2] = u[2];
        }
    }
}

}  // namespace ebem
