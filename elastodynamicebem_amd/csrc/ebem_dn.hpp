// ebem_dn.hpp -- device kernels of the Neumann driver
// (functions/BEMenerg_dir/BEMenerg_dir_timeMarchingN.m): the traction datum
// (BEMenerg_dir_calcBoundDataNeumann.m), beta_V = sum_j V_{n-j} gV_j over the
// single-layer store (BEMenerg_dir_calcBetaVgpu.m), the system matrix
// K_0 + I_Gamma and the right-hand side of each marching step
// beta_V,n + I_Gamma d_{n-1} - sum_{l>=1} K_l d_{n-l} (:275-296).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "ebem_k.hpp"

namespace ebem {

// Traction datum at the triangle centres, gVT[(F * nT + n) * 3 + c] at time
// (n + 1) dt: (0, 0, 2 mu tildeP(c_F, t) n_z) for the bars of height h,
// tildeP = sum_{k=0}^{khat} (-1)^k [H(cP t - 2hk - (h - z)) + H(cP t - 2h(k+1) + (h - z))]
// (calcBoundDataNeumann.m:6-13).
__global__ void k_neu_datum(const FieldTri* __restrict__ tri, int N, int nT, double dt, double cP, double Tfin,
                            double mu, double h, double* __restrict__ gVT) {
#pragma clang fp contract(off)
    const int64_t id = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (id >= (int64_t)N * nT) return;
    const int F = (int)(id / nT), n = (int)(id % nT);
    const double t = (double)(n + 1) * dt, z = tri[F].center[2];
    const int khat = (int)ceil((cP * Tfin) / (2 * h)) - 1;
    double tp = 0.0;
    for (int k = 0; k <= khat; k++) {
        const double a = cP * t - 2 * h * k - (h - z);
        const double b = cP * t - 2 * h * (k + 1) + (h - z);
        const double sg = (k & 1) ? -1.0 : 1.0;
        tp = tp + sg * ((a > 0 ? 1.0 : 0.0) + (b > 0 ? 1.0 : 0.0));
    }
    gVT[id * 3] = 0.0;
    gVT[id * 3 + 1] = 0.0;
    gVT[id * 3 + 2] = 2 * mu * tp * tri[F].n[2];
}

// beta_V rows of test triangle i = blockIdx.x: out[n*3N + 3i + c] = sum over the
// stored blocks (i, j) of E_l (l <= n) of E_l gV_j((n - l + 1) dt); threads own
// time steps n = tid + 256 k.
constexpr int kBetaVPer = 16;
__global__ void __launch_bounds__(256) k_betav(Store S, int L, int N, int nT, const double* __restrict__ gVT,
                                               double* __restrict__ out) {
    const int i = blockIdx.x;
    double acc[kBetaVPer][3];
#pragma unroll
    for (int k = 0; k < kBetaVPer; k++) acc[k][0] = acc[k][1] = acc[k][2] = 0.0;
    for (int l = 0; l < L; l++) {
        for (int64_t b = S.seg[(int64_t)i * L + l]; b < S.seg[(int64_t)i * L + l + 1]; b++) {
            const double e0 = S.v[0][b], e1 = S.v[1][b], e2 = S.v[2][b], e3 = S.v[3][b], e4 = S.v[4][b],
                         e5 = S.v[5][b];
            const double* gj = gVT + (int64_t)S.col[b] * nT * 3;
#pragma unroll
            for (int k = 0; k < kBetaVPer; k++) {
                const int n = threadIdx.x + 256 * k;
                if (n >= nT || n < l) continue;
                const double* g = gj + (int64_t)(n - l) * 3;
                const double g0 = g[0], g1 = g[1], g2 = g[2];
                acc[k][0] += e0 * g0 + e1 * g1 + e2 * g2;
                acc[k][1] += e1 * g0 + e3 * g1 + e4 * g2;
                acc[k][2] += e2 * g0 + e4 * g1 + e5 * g2;
            }
        }
    }
#pragma unroll
    for (int k = 0; k < kBetaVPer; k++) {
        const int n = threadIdx.x + 256 * k;
        if (n >= nT) continue;
        double* o = out + (int64_t)n * 3 * N + 3 * i;
        o[0] = acc[k][0];
        o[1] = acc[k][1];
        o[2] = acc[k][2];
    }
}

// System matrix A = K_0 + I_Gamma (3N x 3Nv column-major, zeroed before):
// K_0 blocks of the pairs whose window starts at lag 0, then area_S / 6 on the
// diagonal of the blocks (S, s) for the vertices s of S (timeMarchingN.m:275-279).
__global__ void k_dn_sysmat(const int16_t* __restrict__ lo, const int16_t* __restrict__ hi,
                            const int64_t* __restrict__ off, const double* __restrict__ val,
                            const int32_t* __restrict__ T, const FieldTri* __restrict__ tri, int N, int Nv,
                            double* __restrict__ A) {
    const int64_t p = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (p >= (int64_t)N * Nv) return;
    const int S = (int)(p / Nv), s = (int)(p % Nv);
    const int64_t ld = 3 * (int64_t)N;
    double blk[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
    if (lo[p] == 0 && hi[p] > 0)
        for (int e = 0; e < 9; e++) blk[e] = val[off[p] * 9 + e];
    if (T[3 * S] == s || T[3 * S + 1] == s || T[3 * S + 2] == s) {
        const double m = tri[S].area / 6;
        blk[0] = blk[0] + m;
        blk[4] = blk[4] + m;
        blk[8] = blk[8] + m;
    }
    for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++) A[(3 * S + i) + ld * (3 * s + j)] = blk[i + 3 * j];
}

// Right-hand side of marching step n for test row S = blockIdx.x:
// tnf = beta_V,n + I_Gamma d_{n-1} - sum_{l=1}^{min(n, LK-1)} K_l d_{n-l}
// (timeMarchingN.m:286-296); d 3Nv x nT column-major.  Threads stride over
// the row's pairs; LDS tree in a fixed order.
__global__ void __launch_bounds__(256) k_dn_rhs(const int16_t* __restrict__ lo, const int16_t* __restrict__ hi,
                                                const int64_t* __restrict__ off, const double* __restrict__ val,
                                                const int32_t* __restrict__ T, const FieldTri* __restrict__ tri,
                                                int N, int Nv, int LK, int n, const double* __restrict__ betaV,
                                                const double* __restrict__ d, double* __restrict__ tnf) {
    __shared__ double red[3][256];
    const int S = blockIdx.x;
    const int64_t n3v = 3 * (int64_t)Nv;
    double a[3] = {0, 0, 0};
    for (int s = threadIdx.x; s < Nv; s += blockDim.x) {
        const int64_t p = (int64_t)S * Nv + s;
        const int l0 = max((int)lo[p], 1), l1 = min((int)hi[p], min(n + 1, LK));
        for (int l = l0; l < l1; l++) {
            const double* K = val + (off[p] + (l - lo[p])) * 9;
            const double* x = d + (int64_t)(n - l) * n3v + 3 * s;
            for (int i = 0; i < 3; i++) a[i] -= K[i] * x[0] + K[i + 3] * x[1] + K[i + 6] * x[2];
        }
    }
    for (int c = 0; c < 3; c++) red[c][threadIdx.x] = a[c];
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w)
            for (int c = 0; c < 3; c++) red[c][threadIdx.x] += red[c][threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        double r[3];
        for (int c = 0; c < 3; c++) r[c] = betaV[(int64_t)n * 3 * N + 3 * S + c];
        if (n > 0) {
            const double m = tri[S].area / 6;
            for (int v = 0; v < 3; v++) {
                const double* x = d + (int64_t)(n - 1) * n3v + 3 * T[3 * S + v];
                for (int c = 0; c < 3; c++) r[c] = r[c] + m * x[c];
            }
        }
        for (int c = 0; c < 3; c++) tnf[3 * S + c] = r[c] + red[c][0];
    }
}

}  // namespace ebem
