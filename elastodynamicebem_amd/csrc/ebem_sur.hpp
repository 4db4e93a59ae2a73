// ebem_sur.hpp -- learned near-field surrogate.
//
//   * pair features (24) and analytic labels (P-front and S-front parts of the
//     E_lag block, 9 + 9 values) computed by the library's own FP64 routine;
//   * a BF16 MFMA GEMM (FP32 accumulate) with the epilogues the MLP needs
//     (bias + GELU forward, GELU backward, split-K weight gradients);
//   * loss, column sums, Adam, and the scatter of predictions into the
//     assembly's level slots.
// MLP: 24 -> 256 -> 256 -> 256 -> 256 -> 18, GELU, BF16 weights and
// activations, FP32 accumulation and FP32 master weights.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "ebem_sa.hpp"

namespace ebem {

constexpr int kSurIn = 24;
constexpr int kSurInPad = 64;  // padded so every GEMM dimension is a multiple of 64
constexpr int kSurHid = 256;
constexpr int kSurOut = 18;
constexpr int kSurOutPad = 64;
constexpr int kSurLayers = 5;

// Parameter layout (FP32 masters, and a BF16 copy of the same layout):
// layer l: W_l (K_l x N_l row-major) then b_l (N_l).
__host__ __device__ constexpr int sur_K(int l) { return l == 0 ? kSurInPad : kSurHid; }
__host__ __device__ constexpr int sur_N(int l) { return l == kSurLayers - 1 ? kSurOutPad : kSurHid; }
__host__ __device__ constexpr int64_t sur_woff(int l) {
    return l == 0 ? 0 : sur_woff(l - 1) + (int64_t)sur_K(l - 1) * sur_N(l - 1) + sur_N(l - 1);
}
__host__ __device__ constexpr int64_t sur_boff(int l) { return sur_woff(l) + (int64_t)sur_K(l) * sur_N(l); }
constexpr int64_t kSurParams = sur_woff(kSurLayers);

// ------------------------------------------------------------------ bf16 --

__host__ __device__ inline uint32_t f_bits(float f) {
    uint32_t u;
    __builtin_memcpy(&u, &f, 4);
    return u;
}
__host__ __device__ inline float bits_f(uint32_t u) {
    float f;
    __builtin_memcpy(&f, &u, 4);
    return f;
}
// Round to nearest even (finite inputs; NaN stays NaN).
__host__ __device__ inline uint16_t f2bf(float f) {
    uint32_t u = f_bits(f);
    if ((u & 0x7f800000u) == 0x7f800000u) return (uint16_t)((u >> 16) | ((u & 0xffffu) ? 0x40u : 0u));
    u += 0x7fffu + ((u >> 16) & 1u);
    return (uint16_t)(u >> 16);
}
__host__ __device__ inline float bf2f(uint16_t b) { return bits_f((uint32_t)b << 16); }

__device__ inline float gelu(float z) { return 0.5f * z * (1.0f + erff(z * 0.70710678118654752f)); }
__device__ inline float gelu_d(float z) {
    return 0.5f * (1.0f + erff(z * 0.70710678118654752f)) + z * 0.39894228040143268f * expf(-0.5f * z * z);
}

// ---------------------------------------------------------------- features --

// Pair-local frame of the field triangle F: rows e1 (its first edge), n x e1, n.
__host__ __device__ inline void sur_frame(const FieldTri& F, double R[3][3]) {
    for (int k = 0; k < 3; k++) {
        R[0][k] = F.e1[k];
        R[2][k] = F.n[k];
    }
    R[1][0] = F.n[1] * F.e1[2] - F.n[2] * F.e1[1];
    R[1][1] = F.n[2] * F.e1[0] - F.n[0] * F.e1[2];
    R[1][2] = F.n[0] * F.e1[1] - F.n[1] * F.e1[0];
}

// Features of (test S, field F, lag), lengths divided by the mesh size h and
// expressed in F's frame with origin at F's centroid:
//   0..8   the test triangle's vertices, 9..17 the field triangle's vertices,
//   18 dt cP / h, 19 dt cS / h, 20 lag,
//   21 (lag dt cP - d) / h and 22 (lag dt cS - d) / h: where the P and the S
//   front stand relative to the centroid distance d (products a network
//   would otherwise have to learn), 23 cosine between the two normals.
// vs / vf: the two triangles' vertices (3 x 3 row-major, mesh order).
__host__ __device__ inline void sur_features(const FieldTri& S, const double* vs, const FieldTri& F, const double* vf,
                                             int lag, double dt, double cP, double cS, double h, float* x) {
    double R[3][3];
    sur_frame(F, R);
    const double* o = F.center;
    const double ih = 1.0 / h;
    for (int v = 0; v < 3; v++)
        for (int r = 0; r < 3; r++) {
            double a = 0, b = 0;
            for (int k = 0; k < 3; k++) {
                a += R[r][k] * (vs[3 * v + k] - o[k]);
                b += R[r][k] * (vf[3 * v + k] - o[k]);
            }
            x[3 * v + r] = (float)(a * ih);
            x[9 + 3 * v + r] = (float)(b * ih);
        }
    x[18] = (float)(dt * cP * ih);
    x[19] = (float)(dt * cS * ih);
    x[20] = (float)lag;
    double dc[3] = {S.center[0] - o[0], S.center[1] - o[1], S.center[2] - o[2]};
    const double d = sqrt(dc[0] * dc[0] + dc[1] * dc[1] + dc[2] * dc[2]);
    x[21] = (float)(((double)lag * dt * cP - d) * ih);
    x[22] = (float)(((double)lag * dt * cS - d) * ih);
    x[23] = (float)(S.n[0] * F.n[0] + S.n[1] * F.n[1] + S.n[2] * F.n[2]);
}

// Output scale of a pair: A_S A_F / (4 pi rho h).
__host__ __device__ inline double sur_scale(const FieldTri& S, const FieldTri& F, double rho, double h) {
    return S.area * F.area / (4.0 * kPi * rho * h);
}

// --------------------------------------------------------------- labels --

// E_lag[S, F] split into its P-front part (the ris1..ris4 terms, integrals over
// the annulus c_S t < r < c_P t) and its S-front part (the ris5, ris6 terms,
// over the disk r < c_S t) of calcMatrixSubBlock.m:101-104, each summed over the
// levels lag+1, lag, lag-1 with weights 1, -2, 1 and divided by 4 pi rho
// (calcMatrixSubBlock.m:101-111).  out: 9 P values then 9 S values, 3 x 3
// row-major in the global frame.
template <int NGH>
__device__ inline void pair_parts(const FieldTri& F, const double* gh, int lag, double dt, double cP, double cS,
                                  double inv4pirho, double* out) {
#pragma clang fp contract(off)
    double EP[6] = {0, 0, 0, 0, 0, 0}, ES[6] = {0, 0, 0, 0, 0, 0};
    const double coef[3] = {1.0, -2.0, 1.0};
    for (int v = 0; v < 3; v++) {
        const int m = lag + 1 - v;
        if (m <= 0) continue;
        const Level K = make_level((double)m * dt, cP, cS);
        Level KP = K, KS = K;
        KP.k5 = KP.k6 = 0.0;
        KS.k1 = KS.k2 = KS.k3 = KS.k4 = 0.0;
        NodeAcc aP, aS;
        aP.s = aS.s = 0.0;
        for (int q = 0; q < 6; q++) aP.m[q] = aS.m[q] = 0.0;
#pragma unroll 1
        for (int q = 0; q < NGH; q++) {
            double x[3] = {gh[q], gh[32 + q], gh[64 + q]};
            NodeGeo G = node_geo(x, F);
            node_eval<double>(G, KP, gh[96 + q], aP);
            node_eval<double>(G, KS, gh[96 + q], aS);
        }
        double WP[6], WS[6];
        finish_rotation(aP, F.C, WP);
        finish_rotation(aS, F.C, WS);
        for (int q = 0; q < 6; q++) {
            EP[q] = EP[q] + coef[v] * WP[q];
            ES[q] = ES[q] + coef[v] * WS[q];
        }
    }
    const int qq[9] = {0, 1, 2, 1, 3, 4, 2, 4, 5};
    for (int k = 0; k < 9; k++) {
        out[k] = EP[qq[k]] * inv4pirho;
        out[9 + k] = ES[qq[k]] * inv4pirho;
    }
}

// One thread per sample (test si, field sj, lag sl): raw features x (n x 24)
// and labels y (n x 18): the two parts rotated into F's frame (R M R^T) and
// divided by the pair scale.
template <int NGH>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(2)))
k_sur_samples(const FieldTri* __restrict__ tri, const double* __restrict__ gh, const double* __restrict__ verts,
              int64_t n, const int32_t* __restrict__ si, const int32_t* __restrict__ sj,
              const int32_t* __restrict__ sl, double dt, double cP, double cS, double rho, double h,
              float* __restrict__ x, float* __restrict__ y) {
    const int64_t k = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (k >= n) return;
    const int i = si[k], j = sj[k], l = sl[k];
    const FieldTri& S = tri[i];
    const FieldTri& F = tri[j];
    sur_features(S, verts + 9 * (int64_t)i, F, verts + 9 * (int64_t)j, l, dt, cP, cS, h, x + (int64_t)kSurIn * k);
    double E[18];
    pair_parts<NGH>(F, gh + (int64_t)i * 128, l, dt, cP, cS, 1.0 / (4.0 * kPi * rho), E);
    double R[3][3];
    sur_frame(F, R);
    const double is = 1.0 / sur_scale(S, F, rho, h);
    for (int part = 0; part < 2; part++) {
        const double* M = E + 9 * part;
        for (int a = 0; a < 3; a++)
            for (int b = 0; b < 3; b++) {
                double v = 0;
                for (int c = 0; c < 3; c++)
                    for (int d = 0; d < 3; d++) v += R[a][c] * M[3 * c + d] * R[b][d];
                y[(int64_t)kSurOut * k + 9 * part + 3 * a + b] = (float)(v * is);
            }
    }
}

// -------------------------------------------------------------- BF16 GEMM --

typedef __bf16 bf16x8_t __attribute__((ext_vector_type(8)));
typedef float f32x4_t __attribute__((ext_vector_type(4)));

enum SurEpi {
    kEpiStore = 0,     // C[z] = acc                          (split-K partials)
    kEpiBiasGelu = 1,  // Z = acc + b (if C), A = bf16(gelu(Z))
    kEpiBias = 2,      // C = acc + b
    kEpiGeluBwd = 3,   // G = acc * gelu'(Zin); C = G (if C), Cb = bf16(G)
};

struct EpiArgs {
    float* C;
    int ldc;
    int64_t cz;  // stride between split-K partials
    uint16_t* Cb;
    int ldcb;
    const float* bias;
    const float* Zin;
    int ldz;
};

constexpr int kGemmLds = 40;  // bf16 row stride of the LDS tiles (32 + 8 pad)

// C (M x N) = op(A) (M x K) * op(B) (K x N) with BF16 operands and FP32
// accumulation on the 16x16x32 BF16 MFMA.  op(A)(i,k) = TA ? A[k lda + i] :
// A[i lda + k]; op(B)(k,j) = TB ? B[j ldb + k] : B[k ldb + j].
// Grid (N/64, M/64, splits), 256 threads: wave w owns the 32x32 quarter
// (w/2, w%2) of the 64x64 tile as 2x2 MFMA tiles; blockIdx.z covers K range
// [z ksplit, min(K, (z+1) ksplit)).  The host guarantees M, N % 64 == 0,
// K, ksplit % 32 == 0, lda, ldb % 8 == 0 and 16-byte aligned operands.
// MFMA lane layout: lane l supplies row (l % 16) of the A tile and column
// (l % 16) of the B tile, 8 consecutive k each starting at 8 (l / 16); it
// receives rows 4 (l / 16) + r, r < 4, of column l % 16 of the result.
template <bool TA, bool TB, int EPI>
__global__ void __launch_bounds__(256) k_gemm_bf16(const uint16_t* __restrict__ A, int lda,
                                                   const uint16_t* __restrict__ B, int ldb, int K, int ksplit,
                                                   EpiArgs e) {
    __shared__ __attribute__((aligned(16))) uint16_t As[64 * kGemmLds];
    __shared__ __attribute__((aligned(16))) uint16_t Bs[64 * kGemmLds];
    const int m0 = blockIdx.y * 64, n0 = blockIdx.x * 64;
    const int kb = blockIdx.z * ksplit;
    const int ke = kb + ksplit < K ? kb + ksplit : K;
    const int t = threadIdx.x, lane = t & 63, w = t >> 6;
    const int wr = (w >> 1) * 32, wc = (w & 1) * 32;
    f32x4_t acc[2][2];
    for (int a = 0; a < 2; a++)
        for (int b = 0; b < 2; b++) acc[a][b] = f32x4_t{0.f, 0.f, 0.f, 0.f};
    for (int k0 = kb; k0 < ke; k0 += 32) {
        if (!TA) {
            const int r = t >> 2, kq = (t & 3) * 8;
            uint4 v = *reinterpret_cast<const uint4*>(A + (int64_t)(m0 + r) * lda + k0 + kq);
            *reinterpret_cast<uint4*>(As + r * kGemmLds + kq) = v;
        } else {
            const int kk = t >> 3, iq = (t & 7) * 8;
            uint4 v = *reinterpret_cast<const uint4*>(A + (int64_t)(k0 + kk) * lda + m0 + iq);
            const uint32_t u[4] = {v.x, v.y, v.z, v.w};
#pragma unroll
            for (int s = 0; s < 8; s++) As[(iq + s) * kGemmLds + kk] = (uint16_t)(u[s >> 1] >> (16 * (s & 1)));
        }
        if (TB) {
            const int c = t >> 2, kq = (t & 3) * 8;
            uint4 v = *reinterpret_cast<const uint4*>(B + (int64_t)(n0 + c) * ldb + k0 + kq);
            *reinterpret_cast<uint4*>(Bs + c * kGemmLds + kq) = v;
        } else {
            const int kk = t >> 3, jq = (t & 7) * 8;
            uint4 v = *reinterpret_cast<const uint4*>(B + (int64_t)(k0 + kk) * ldb + n0 + jq);
            const uint32_t u[4] = {v.x, v.y, v.z, v.w};
#pragma unroll
            for (int s = 0; s < 8; s++) Bs[(jq + s) * kGemmLds + kk] = (uint16_t)(u[s >> 1] >> (16 * (s & 1)));
        }
        __syncthreads();
        bf16x8_t fa[2], fb[2];
#pragma unroll
        for (int q = 0; q < 2; q++) {
            uint4 va = *reinterpret_cast<const uint4*>(As + (wr + q * 16 + (lane & 15)) * kGemmLds + 8 * (lane >> 4));
            uint4 vb = *reinterpret_cast<const uint4*>(Bs + (wc + q * 16 + (lane & 15)) * kGemmLds + 8 * (lane >> 4));
            fa[q] = __builtin_bit_cast(bf16x8_t, va);
            fb[q] = __builtin_bit_cast(bf16x8_t, vb);
        }
#pragma unroll
        for (int a = 0; a < 2; a++)
#pragma unroll
            for (int b = 0; b < 2; b++) acc[a][b] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(fa[a], fb[b], acc[a][b], 0, 0, 0);
        __syncthreads();
    }
#pragma unroll
    for (int a = 0; a < 2; a++)
#pragma unroll
        for (int b = 0; b < 2; b++)
#pragma unroll
            for (int r = 0; r < 4; r++) {
                const int row = m0 + wr + a * 16 + 4 * (lane >> 4) + r;
                const int col = n0 + wc + b * 16 + (lane & 15);
                const float v = acc[a][b][r];
                if (EPI == kEpiStore) {
                    e.C[blockIdx.z * e.cz + (int64_t)row * e.ldc + col] = v;
                } else if (EPI == kEpiBiasGelu) {
                    const float z = v + e.bias[col];
                    if (e.C) e.C[(int64_t)row * e.ldc + col] = z;
                    e.Cb[(int64_t)row * e.ldcb + col] = f2bf(gelu(z));
                } else if (EPI == kEpiBias) {
                    e.C[(int64_t)row * e.ldc + col] = v + e.bias[col];
                } else {
                    const float g = v * gelu_d(e.Zin[(int64_t)row * e.ldz + col]);
                    if (e.C) e.C[(int64_t)row * e.ldc + col] = g;
                    e.Cb[(int64_t)row * e.ldcb + col] = f2bf(g);
                }
            }
}

// ----------------------------------------------------- training helpers --

// Rows b < nb of the batch: standardised features of sample idx[b] (bf16,
// 64 columns, zero padded) and standardised targets (fp32, 18 columns); rows
// nb..B-1 are zero.
// Block (64, 4): threadIdx.x is the column.
__global__ void k_sur_gather(const float* __restrict__ X, const float* __restrict__ Y, const int32_t* __restrict__ idx,
                             int nb, int B, const float* __restrict__ norm, uint16_t* __restrict__ A0,
                             float* __restrict__ T) {
    const int b = blockIdx.x * blockDim.y + threadIdx.y;
    const int c = threadIdx.x;  // 0..63
    if (b >= B) return;
    float xv = 0.f, tv = 0.f;
    if (b < nb) {
        const int64_t s = idx[b];
        if (c < kSurIn) xv = (X[s * kSurIn + c] - norm[c]) * norm[kSurIn + c];
        if (c < kSurOut) tv = (Y[s * kSurOut + c] - norm[2 * kSurIn + c]) / norm[2 * kSurIn + kSurOut + c];
    }
    A0[(int64_t)b * kSurInPad + c] = f2bf(xv);
    if (c < kSurOut) T[(int64_t)b * kSurOut + c] = tv;
}

// dY = 2 (Y - T) / (nb * 18) on the first nb rows and 18 columns, 0 elsewhere
// (fp32 for the bias gradient and bf16 for the GEMMs); loss += sum (Y - T)^2 / (nb * 18).
__global__ void k_sur_loss(const float* __restrict__ Y, const float* __restrict__ T, int nb, int B,
                           float* __restrict__ dY, uint16_t* __restrict__ dYb, double* __restrict__ loss) {
    __shared__ double part[256];
    const int64_t id = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    double l = 0.0;
    if (id < (int64_t)B * kSurOutPad) {
        const int b = (int)(id / kSurOutPad), c = (int)(id % kSurOutPad);
        float g = 0.f;
        if (b < nb && c < kSurOut) {
            const float d = Y[id] - T[(int64_t)b * kSurOut + c];
            const float inv = 1.0f / ((float)nb * (float)kSurOut);
            g = 2.0f * d * inv;
            l = (double)d * (double)d * (double)inv;
        }
        if (dY) dY[id] = g;
        if (dYb) dYb[id] = f2bf(g);
    }
    part[threadIdx.x] = l;
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        if ((int)threadIdx.x < s) part[threadIdx.x] += part[threadIdx.x + s];
        __syncthreads();
    }
    if (threadIdx.x == 0 && loss) atomicAdd(loss, part[0]);
}

// out[c] = sum_r G[r ld + c], grid N/64 blocks of 256 threads (4 row groups).
__global__ void __launch_bounds__(256) k_colsum(const float* __restrict__ G, int rows, int ld, float* __restrict__ out) {
    __shared__ float part[4][64];
    const int c = blockIdx.x * 64 + (threadIdx.x & 63), g = threadIdx.x >> 6;
    float s = 0.f;
    for (int r = g; r < rows; r += 4) s += G[(int64_t)r * ld + c];
    part[g][threadIdx.x & 63] = s;
    __syncthreads();
    if (g == 0) out[c] = (part[0][threadIdx.x] + part[1][threadIdx.x]) + (part[2][threadIdx.x] + part[3][threadIdx.x]);
}

// out[i] = sum_s part[s * n + i]
__global__ void k_sum_splits(const float* __restrict__ part, int splits, int64_t n, float* __restrict__ out) {
    const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (i >= n) return;
    float s = 0.f;
    for (int z = 0; z < splits; z++) s += part[z * n + i];
    out[i] = s;
}

// Adam on the FP32 masters and refresh of their BF16 copy.
__global__ void k_adam(float* __restrict__ p, const float* __restrict__ g, float* __restrict__ m, float* __restrict__ v,
                       uint16_t* __restrict__ pb, int64_t n, float lr, float b1, float b2, float eps, float bc1,
                       float bc2) {
    const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float gi = g[i];
    const float mi = b1 * m[i] + (1.f - b1) * gi;
    const float vi = b2 * v[i] + (1.f - b2) * gi * gi;
    m[i] = mi;
    v[i] = vi;
    const float pi = p[i] - lr * (mi / bc1) / (sqrtf(vi / bc2) + eps);
    p[i] = pi;
    pb[i] = f2bf(pi);
}

__global__ void k_to_bf16(const float* __restrict__ p, uint16_t* __restrict__ pb, int64_t n) {
    const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (i < n) pb[i] = f2bf(p[i]);
}

// ------------------------------------------------------ assembly inference --

// Surrogate pair flag of every pair of a row chunk: not coincident, centroid
// distance <= sqrt(near2), a lag window that ends before the last lag L - 1
// (so that its level hi + 1 is static, see k_emit), and the features of the
// window inside the training box [box[c], box[24 + c]].
__global__ void __launch_bounds__(256) k_sur_flag(const FieldTri* __restrict__ tri, const double* __restrict__ verts,
                                                  int N, int L, int project, int r0, int rows,
                                                  const int16_t* __restrict__ lo,
                                                  const int16_t* __restrict__ hi, double dt, double cP, double cS,
                                                  double h, double near2, const float* __restrict__ box,
                                                  uint8_t* __restrict__ flag, unsigned long long* __restrict__ nblk) {
    __shared__ unsigned int cnt;
    if (threadIdx.x == 0) cnt = 0;
    __syncthreads();
    const int64_t p = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    uint8_t f = 0;
    int l0 = 0, l1 = -1, i = 0, j = 0;
    if (p < (int64_t)rows * N) {
        i = r0 + (int)(p / N);
        j = (int)(p % N);
        l0 = lo[p];
        l1 = hi[p];
    }
    if (p < (int64_t)rows * N && i != j && l0 <= l1 && l1 < L - 1) {
        const FieldTri& S = tri[i];
        const FieldTri& F = tri[j];
        double v0 = F.center[0] - S.center[0], v1 = F.center[1] - S.center[1], v2 = F.center[2] - S.center[2];
        if (v0 * v0 + v1 * v1 + v2 * v2 <= near2) {
            // the lag enters features 20..22 monotonically: test both window ends
            float x[kSurIn], x1[kSurIn];
            sur_features(S, verts + 9 * (int64_t)i, F, verts + 9 * (int64_t)j, l0, dt, cP, cS, h, x);
            sur_features(S, verts + 9 * (int64_t)i, F, verts + 9 * (int64_t)j, l1, dt, cP, cS, h, x1);
            bool in = true;
            for (int c = 0; c < kSurIn; c++)
                in = in && x[c] >= box[c] && x[c] <= box[kSurIn + c] && x1[c] >= box[c] && x1[c] <= box[kSurIn + c];
            f = in ? 1 : 0;
        }
    }
    if (p < (int64_t)rows * N) flag[p] = f;
    // predicted blocks: lags max(lo,1)..hi-1 with the projection (E_0 and E_hi
    // then come from the analytic routine), max(lo,1)..hi without it
    const int np = f ? l1 - (l0 > 1 ? l0 : 1) + (project ? 0 : 1) : 0;
    if (np > 0) atomicAdd(&cnt, (unsigned int)np);
    __syncthreads();
    if (threadIdx.x == 0 && cnt) atomicAdd(nblk, (unsigned long long)cnt);
}

// Standardised bf16 features of the surrogate items perm[0..n) (pair p,
// level m, lag m - 1); slot[k] = item index when lag m - 1 lies in the pair's
// window, else -1 (the item only pads the batch).  Rows n..B-1 are zero.
__global__ void __launch_bounds__(256) k_sur_feats(const FieldTri* __restrict__ tri, const double* __restrict__ verts,
                                                   int N, int r0, int64_t n, int B, const int32_t* __restrict__ perm,
                                                   const int32_t* __restrict__ item_pair,
                                                   const int64_t* __restrict__ off, const int16_t* __restrict__ lo,
                                                   const int16_t* __restrict__ hi, double dt, double cP, double cS,
                                                   double h, const float* __restrict__ norm,
                                                   uint16_t* __restrict__ A0, int64_t* __restrict__ slot) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= B) return;
    float x[kSurIn];
    int64_t s = -1;
    if (k < n) {
        const int64_t it = perm[k];
        const int32_t p = item_pair[it];
        const int i = r0 + p / N, j = p % N;
        const int l0 = lo[p], l1 = hi[p];
        const int mlo = l0 - 1 < 1 ? 1 : l0 - 1;
        const int lag = mlo + (int)(it - off[p]) - 1;
        if (lag >= l0 && lag <= l1) {
            s = it;
            sur_features(tri[i], verts + 9 * (int64_t)i, tri[j], verts + 9 * (int64_t)j, lag, dt, cP, cS, h, x);
        }
    }
    slot[k] = s;
    uint16_t* a = A0 + (int64_t)k * kSurInPad;
    for (int c = 0; c < kSurInPad; c++)
        a[c] = (s >= 0 && c < kSurIn) ? f2bf((x[c] - norm[c]) * norm[kSurIn + c]) : (uint16_t)0;
}

// Predicted block of every valid row into its level slot: y -> unstandardised
// local P + S parts, symmetrised, rotated back to the global frame (R^T M R),
// times the pair scale and 4 pi rho (the emit pass divides by 4 pi rho).
__global__ void __launch_bounds__(256) k_sur_scatter(const FieldTri* __restrict__ tri, int N, int r0, int B,
                                                     const int64_t* __restrict__ slot,
                                                     const int32_t* __restrict__ item_pair, const float* __restrict__ Y,
                                                     const float* __restrict__ norm, double rho, double h,
                                                     double* __restrict__ W) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= B) return;
    const int64_t it = slot[k];
    if (it < 0) return;
    const int32_t p = item_pair[it];
    const int i = r0 + p / N, j = p % N;
    const FieldTri& S = tri[i];
    const FieldTri& F = tri[j];
    const float* ym = norm + 2 * kSurIn;
    const float* ys = ym + kSurOut;
    double M[3][3];
    for (int a = 0; a < 3; a++)
        for (int b = 0; b < 3; b++) {
            const int c = 3 * a + b;
            M[a][b] = (double)(Y[(int64_t)k * kSurOutPad + c] * ys[c] + ym[c]) +
                      (double)(Y[(int64_t)k * kSurOutPad + 9 + c] * ys[9 + c] + ym[9 + c]);
        }
    double R[3][3];
    sur_frame(F, R);
    const double sc = sur_scale(S, F, rho, h) * 4.0 * kPi * rho;
    const int ii[6] = {0, 0, 0, 1, 1, 2}, jj[6] = {0, 1, 2, 1, 2, 2};
    double* o = W + it * 6;
    for (int q = 0; q < 6; q++) {
        double v = 0;
        for (int a = 0; a < 3; a++)
            for (int b = 0; b < 3; b++) v += R[a][ii[q]] * 0.5 * (M[a][b] + M[b][a]) * R[b][jj[q]];
        o[q] = v * sc;
    }
}

}  // namespace ebem
