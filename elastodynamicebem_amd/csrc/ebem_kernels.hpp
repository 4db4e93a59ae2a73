// ebem_kernels.hpp -- gfx950 kernels of the single-layer hot path.
//
// Assembly (replaces SAc/BEMenerg_coreSA_calcMatrixBlock.m and the kernelV
// launches of dir/BEMenerg_dir_timeMarchingD.m):
//   k_windows      per (test i, field j) pair: the contiguous lag window that
//                  passes the light-cone test of calcMatrixBlock.m:57-61;
//   k_rowcount     per test row: number of blocks per (row, lag) and of
//                  (pair, level) items;
//   k_pair_items   per row: item offsets of every pair and the item -> pair map;
//   k_levels       per (pair, level m) item: W_m = sum_q w_q * integral at
//                  t_m = m dt over the field triangle (the time-level term of
//                  calcMatrixSubBlock.m:20-111), rotated to the global frame.
//                  Each W_m serves the three lags m-1, m, m+1, so the
//                  analytic work is a third of the reference's;
//   k_emit         per row: E_l = (W_{l+1} - 2 W_l + W_{l-1}) / (4 pi rho) with
//                  the 1e-14 cleanup (calcMatrixBlock.m:77), written into the
//                  row-lag-segmented block store (symmetric 3x3 -> 6 values).
// MOT (SA/BEMenerg_coreSA_timeMarching.m:70-113):
//   k_history_far  the lags of the history sum whose densities precede a block of
//                  kHistBlock steps, for all steps of the block in one pass;
//   k_history      r_n = beta_n - (far part) - the remaining lags, one workgroup
//                  per test row, streaming the row's contiguous lag segments;
//   k_gemv_rows    alpha_n = E_0^{-1} r_n with a row-major inverse, one wave
//                  per row.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "ebem_sa.hpp"

namespace ebem {

// Block store: six value arrays (w00,w01,w02,w11,w12,w22) of nnzb entries,
// column (field triangle) index per block, and segment offsets
// seg[i*L + l] .. seg[i*L + l + 1] of the blocks of test row i at lag l.
struct Store {
    double* v[6];
    int32_t* col;
    const int64_t* seg;
};

struct AsmParams {
    int N, L;
    double dt, cP, cS, rho;
    int nGH;
    double far2;  // squared centroid distance above which a pair is far field (mixed precision); <0: none
    int kp;       // 1: levels of a pair in the kP state go through k_levels_kp (kp_range)
    int sur_project;  // 1: surrogate pairs keep E_0 and the static level analytic and k_emit projects
};

// (lag-1)*cS*dt < dmax && (lag+1)*cP*dt >= dmin  (calcMatrixBlock.m:61), same
// operation order, no contraction.
__device__ inline bool cull_pass(int lag, double cS, double cP, double dt, double dmin, double dmax) {
#pragma clang fp contract(off)
    return ((double)(lag - 1) * cS * dt < dmax) && ((double)(lag + 1) * cP * dt >= dmin);
}

__device__ inline void pair_dist(const FieldTri& S, const FieldTri& F, double& dmin, double& dmax) {
#pragma clang fp contract(off)
    double v0 = F.center[0] - S.center[0], v1 = F.center[1] - S.center[1], v2 = F.center[2] - S.center[2];
    double d = sqrt(v0 * v0 + v1 * v1 + v2 * v2);
    dmin = d - F.maxL - S.maxL;
    dmax = d + F.maxL + S.maxL;
}

// Lag window [lo, hi] (clipped to [0, L-1]); lo > hi when empty.
__device__ inline void lag_window(const AsmParams& P, double dmin, double dmax, int& lo, int& hi) {
    // upper end: largest lag with (lag-1) cS dt < dmax
    double uh = dmax / (P.cS * P.dt) + 1.0;
    int h = (int)fmin(uh, (double)(P.L + 2));
    if (h > P.L - 1) h = P.L - 1;
    while (h >= 0 && !((double)(h - 1) * P.cS * P.dt < dmax)) h--;
    while (h + 1 <= P.L - 1 && ((double)(h) * P.cS * P.dt < dmax)) h++;
    // lower end: smallest lag with (lag+1) cP dt >= dmin
    double ul = dmin / (P.cP * P.dt) - 1.0;
    int l = (int)fmax(0.0, fmin(ul, (double)(P.L + 2)));
    while (l > 0 && ((double)(l) * P.cP * P.dt >= dmin)) l--;
    while (l <= P.L - 1 && !((double)(l + 1) * P.cP * P.dt >= dmin)) l++;
    lo = l;
    hi = h;
}

__device__ inline int64_t pair_items(int lo, int hi) {
    if (lo > hi) return 0;
    int mlo = lo - 1 < 1 ? 1 : lo - 1;
    return (int64_t)(hi + 1 - mlo + 1);
}

// grid: ceil(rows*N/256); pairs p = (i - r0)*N + j for rows [r0, r1).
// Lags whose block is zero in exact arithmetic although the centroid test of
// calcMatrixBlock.m:57-61 lets the pair through:
//  * after the S front: W_m = sum_q w_q I(x_q, t_m) is the same for every
//    level m with c_S t_m > Dmax = max over outer nodes x_q and field vertices
//    v of |x_q - v| (every node sees the whole field triangle inside its S disk,
//    the P annulus misses it, and the disk integrals no longer depend on t), so
//    E_l = W_{l+1} - 2 W_l + W_{l-1} = 0 once level l-1 is such a level;
//  * before the P front: W_m = 0 while c_P t_m < Dlo = min_q |x_q - c_F| - R_F
//    (R_F the field triangle's circumradius about its centroid), so E_l = 0
//    while W_{l+1} = 0.
// The analytic form reproduces these zeros to rounding, far below the 1e-14
// cleanup of calcMatrixBlock.m:77, so dropping the lags leaves the matrices
// unchanged; the relative margins 1e-12 keep ties on the computed side.
__device__ inline void exact_window(const AsmParams& P, const double* __restrict__ gS, const double* __restrict__ vF,
                                    const double* cF, int& l, int& h) {
    double rf2 = 0.0;
    for (int v = 0; v < 3; v++) {
        double a = vF[3 * v] - cF[0], b = vF[3 * v + 1] - cF[1], c = vF[3 * v + 2] - cF[2];
        rf2 = fmax(rf2, a * a + b * b + c * c);
    }
    double dmax2 = 0.0, dlo = 1e300;
    for (int q = 0; q < P.nGH; q++) {
        const double x = gS[q], y = gS[32 + q], z = gS[64 + q];
        for (int v = 0; v < 3; v++) {
            double a = x - vF[3 * v], b = y - vF[3 * v + 1], c = z - vF[3 * v + 2];
            dmax2 = fmax(dmax2, a * a + b * b + c * c);
        }
        double a = x - cF[0], b = y - cF[1], c = z - cF[2];
        dlo = fmin(dlo, sqrt(a * a + b * b + c * c));
    }
    dlo = fmax(0.0, dlo - sqrt(rf2));
    // last lag with a non-static level l-1: m* = floor(Dmax/(c_S dt)) + 1 is the
    // first static level, so E_l = 0 for l >= m* + 1
    const double ms = floor(sqrt(dmax2) * (1.0 + 1e-12) / (P.cS * P.dt)) + 1.0;
    if (ms < (double)h) h = (int)ms;
    // first lag whose level l+1 can reach the field triangle
    const double ml = ceil(dlo * (1.0 - 1e-12) / (P.cP * P.dt) - 1.0);
    if (ml > (double)l) l = (int)fmin(ml, 32000.0);
}

__global__ void __launch_bounds__(256) k_windows(const FieldTri* __restrict__ tri, const double* __restrict__ gh,
                                                 const double* __restrict__ verts, AsmParams P, int r0, int r1,
                                                 int16_t* __restrict__ lo, int16_t* __restrict__ hi,
                                                 int16_t* __restrict__ ka, int16_t* __restrict__ kb,
                                                 int16_t* __restrict__ sa, int16_t* __restrict__ sb,
                                                 unsigned long long* __restrict__ culled) {
    __shared__ unsigned long long part;
    if (threadIdx.x == 0) part = 0;
    __syncthreads();
    unsigned long long my = 0;
    int64_t np = (int64_t)(r1 - r0) * P.N;
    for (int64_t p = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; p < np; p += (int64_t)gridDim.x * blockDim.x) {
        int i = r0 + (int)(p / P.N), j = (int)(p % P.N);
        double dmin, dmax;
        pair_dist(tri[i], tri[j], dmin, dmax);
        int l, h;
        lag_window(P, dmin, dmax, l, h);
        if (l <= h) {
            my += (unsigned long long)(h - l + 1);  // pair.lags of the reference's test
            exact_window(P, gh + (int64_t)i * 128, verts + (int64_t)j * 9, tri[j].center, l, h);
        }
        if (P.kp) {
            int a = 1, b = 0;
            int c = 1, d = 0;
            if (l <= h)
                kp_range(P.nGH, P.cP, P.cS, P.dt, gh + (int64_t)i * 128, tri[j], verts + (int64_t)j * 9, l, h, a, b, c,
                         d);
            sa[p] = (int16_t)c;
            sb[p] = (int16_t)d;
            ka[p] = (int16_t)a;
            kb[p] = (int16_t)b;
        }
        lo[p] = (int16_t)l;
        hi[p] = (int16_t)h;
    }
    atomicAdd(&part, my);
    __syncthreads();
    if (threadIdx.x == 0) atomicAdd(culled, part);
}

// One workgroup per row (rows [r0, r1)): cnt[(i-r0)*L + l] = blocks of row i at lag l,
// items[i - r0] = (pair, level) items of the row.
__global__ void __launch_bounds__(256) k_rowcount(const int16_t* __restrict__ lo, const int16_t* __restrict__ hi,
                                                  int N, int L, int r0, int r1, int64_t* __restrict__ cnt,
                                                  int64_t* __restrict__ items) {
    extern __shared__ int sdiff[];  // L + 1 ints
    __shared__ long long sitems;
    int row = r0 + blockIdx.x;
    if (row >= r1) return;
    for (int l = threadIdx.x; l <= L; l += blockDim.x) sdiff[l] = 0;
    if (threadIdx.x == 0) sitems = 0;
    __syncthreads();
    long long my = 0;
    const int16_t* plo = lo + (int64_t)blockIdx.x * N;
    const int16_t* phi = hi + (int64_t)blockIdx.x * N;
    for (int j = threadIdx.x; j < N; j += blockDim.x) {
        int a = plo[j], b = phi[j];
        if (a <= b) {
            atomicAdd(&sdiff[a], 1);
            atomicAdd(&sdiff[b + 1], -1);
            my += pair_items(a, b);
        }
    }
    atomicAdd((unsigned long long*)&sitems, (unsigned long long)my);
    __syncthreads();
    if (threadIdx.x == 0) {
        int run = 0;
        for (int l = 0; l < L; l++) {
            run += sdiff[l];
            cnt[(int64_t)blockIdx.x * L + l] = run;
        }
        items[blockIdx.x] = sitems;
    }
}

// One workgroup (256 threads) per row of the chunk: item offsets per pair
// (base[row] + exclusive prefix within the row) and the item -> pair map.
__global__ void __launch_bounds__(256) k_pair_items(const int16_t* __restrict__ lo, const int16_t* __restrict__ hi,
                                                    int N, const int64_t* __restrict__ rowbase,
                                                    int64_t* __restrict__ off, int32_t* __restrict__ item_pair) {
    __shared__ long long part[256];
    const int r = blockIdx.x;
    const int nper = (N + blockDim.x - 1) / blockDim.x;
    const int j0 = threadIdx.x * nper;
    const int64_t pbase = (int64_t)r * N;
    long long s = 0;
    for (int k = 0; k < nper; k++) {
        int j = j0 + k;
        if (j < N) s += pair_items(lo[pbase + j], hi[pbase + j]);
    }
    part[threadIdx.x] = s;
    __syncthreads();
    // inclusive scan over 256 partials (Hillis-Steele)
    for (int d = 1; d < 256; d <<= 1) {
        long long v = (threadIdx.x >= d) ? part[threadIdx.x - d] : 0;
        __syncthreads();
        part[threadIdx.x] += v;
        __syncthreads();
    }
    long long run = rowbase[r] + (threadIdx.x ? part[threadIdx.x - 1] : 0);
    for (int k = 0; k < nper; k++) {
        int j = j0 + k;
        if (j >= N) break;
        int64_t p = pbase + j;
        int64_t n = pair_items(lo[p], hi[p]);
        off[p] = run;
        for (int64_t q = 0; q < n; q++) item_pair[run + q] = (int32_t)p;
        run += n;
    }
}

// Front-state key of an item: where the P and the S front of level m sit
// relative to the span [d - e, d + e] of the pair (d centroid distance, e sum of
// longest edges), each in kFrontBins + 2 classes (not arrived, kFrontBins equal
// parts of the span, passed).  Items with equal keys take the same region
// branches in most outer-node evaluations, so sorting by key keeps wavefronts
// convergent.
#ifndef EBEM_FRONT_BINS
#define EBEM_FRONT_BINS 3
#endif
constexpr int kFrontBins = EBEM_FRONT_BINS;
constexpr int kFrontClasses = kFrontBins + 2;
constexpr int kFrontKeys = kFrontClasses * kFrontClasses;
// Key order = order of the sorted permutation: far-field items (mixed
// precision) 0..kFrontKeys-1, the per-level items of coplanar pairs (every
// outer node in the field plane: cheap coplanar branches, evaluated item per
// lane), the other per-level items (evaluated by outer node), the first level
// of each kP range (one k_levels_kp item per pair) and its other levels (no
// work of their own), the same for the between-front ranges (k_levels_bf;
// their first levels keyed by the number of chunks, so that a wavefront's
// ranges take similar time), surrogate pairs last.
constexpr int kCopKeys = kFrontKeys;     // kFrontKeys .. 2 kFrontKeys - 1
constexpr int kNearKeys = 2 * kFrontKeys;  // .. 3 kFrontKeys - 1
// Range heads of far-field pairs (mixed precision, region integrals in FP32)
// precede those of the other pairs, and so do their other levels.
constexpr int kKpFarKey = 3 * kFrontKeys;
constexpr int kKpKey = kKpFarKey + 1;
constexpr int kKpFarRestKey = kKpKey + 1;
constexpr int kKpRestKey = kKpFarRestKey + 1;
constexpr int kBfHeadKeys = 4;
constexpr int kBfFarKey = kKpRestKey + 1;  // .. kBfFarKey + kBfHeadKeys - 1
constexpr int kBfKey = kBfFarKey + kBfHeadKeys;  // .. kBfKey + kBfHeadKeys - 1
constexpr int kBfFarRestKey = kBfKey + kBfHeadKeys;
constexpr int kBfRestKey = kBfFarRestKey + 1;
constexpr int kSurKey = kBfRestKey + 1;
constexpr int kNumKeys = kSurKey + 1;

// The post-processing's between-front ranges (k_pp_M_bf) in chunks of 8
// levels (8.2 ms against 9.0 with 6 on 4096 points x 1280 triangles).
#ifndef EBEM_PP_BF_CHUNK
#define EBEM_PP_BF_CHUNK 8
#endif
constexpr int kPpBfChunk = EBEM_PP_BF_CHUNK;

// Range key by the number of chunks of C levels (waves of ranges with equal
// chunk counts run convergent).
template <int C>
__device__ inline int bf_head_key(int a, int b, int base = kBfKey) {
    const int chunks = (b - a + C) / C;
    return base + (chunks < kBfHeadKeys ? chunks : kBfHeadKeys) - 1;
}
static_assert(kNumKeys <= 256, "keys are stored as uint8");

// Test triangle S within 1e-6 of the plane of F at all three vertices: every
// outer node of the pair then has zeta = 0 in node_geo (the nodes are convex
// combinations of the vertices).  Only routes the pair's items to a kernel;
// both evaluate any node.
__device__ inline bool coplanar_pair(const FieldTri& S, const FieldTri& F) {
#pragma unroll
    for (int v = 0; v < 3; v++) {
        const double a = (v == 1) ? S.L1 : (v == 2 ? S.L3 * S.cosd : 0.0);
        const double b = (v == 2) ? S.L3 * S.sind : 0.0;
        double dn = 0.0;
#pragma unroll
        for (int c = 0; c < 3; c++) dn += F.n[c] * (S.v0[c] + a * S.e1[c] + b * S.e2[c] - F.v0[c]);
        if (fabs(dn) >= 1.0e-6) return false;
    }
    return true;
}

__device__ inline int front_class(double r, double lo, double span) {
    double u = (r - lo) / span;
    int q = (int)floor(u * (double)kFrontBins);
    return q < -1 ? -1 : (q > kFrontBins ? kFrontBins : q);
}

// Pass 1 of the counting sort: block b keys its contiguous item range
// [b*per, (b+1)*per) and writes its histogram to bh[key*G + b].
__global__ void __launch_bounds__(256) k_item_keys(const FieldTri* __restrict__ tri, AsmParams P, int r0,
                                                   int64_t nitems, int64_t per, const int32_t* __restrict__ item_pair,
                                                   const int64_t* __restrict__ off, const int16_t* __restrict__ lo,
                                                   const int16_t* __restrict__ hi,
                                                   const int16_t* __restrict__ ka, const int16_t* __restrict__ kb,
                                                   const int16_t* __restrict__ sa, const int16_t* __restrict__ sb,
                                                   const uint8_t* __restrict__ sflag, uint8_t* __restrict__ key,
                                                   unsigned int* __restrict__ bh) {
    __shared__ unsigned int sh[kNumKeys];
    for (int k = threadIdx.x; k < kNumKeys; k += blockDim.x) sh[k] = 0;
    __syncthreads();
    const int64_t a = blockIdx.x * per;
    const int64_t e = a + per < nitems ? a + per : nitems;
    for (int64_t it = a + threadIdx.x; it < e; it += blockDim.x) {
        const int32_t p = item_pair[it];
        const int i = r0 + p / P.N, j = p % P.N;
        const int l0 = lo[p];
        const int m = (l0 - 1 < 1 ? 1 : l0 - 1) + (int)(it - off[p]);
        const FieldTri& S = tri[i];
        const FieldTri& F = tri[j];
        double v0 = F.center[0] - S.center[0], v1 = F.center[1] - S.center[1], v2 = F.center[2] - S.center[2];
        double d2 = v0 * v0 + v1 * v1 + v2 * v2;
        double d = sqrt(d2);
        double ex = F.maxL + S.maxL;
        double t = (double)m * P.dt;
        int qp = front_class(P.cP * t, d - ex, 2.0 * ex) + 1;
        int qs = front_class(P.cS * t, d - ex, 2.0 * ex) + 1;
        const int near_key = qp * kFrontClasses + qs + (coplanar_pair(S, F) ? kCopKeys : kNearKeys);
        const bool far = P.far2 >= 0.0 && d2 > P.far2;
        int k = far ? qp * kFrontClasses + qs : near_key;
        if (ka && m >= ka[p] && m <= kb[p])
            k = (m == ka[p]) ? (far ? kKpFarKey : kKpKey) : (far ? kKpFarRestKey : kKpRestKey);
        if (sa && m >= sa[p] && m <= sb[p])
            k = (m == sa[p]) ? (far ? bf_head_key<kBfChunkF32>(sa[p], sb[p], kBfFarKey)
                                    : bf_head_key<kBfChunk>(sa[p], sb[p], kBfKey))
                             : (far ? kBfFarRestKey : kBfRestKey);
        // Surrogate pairs: the network predicts the lags max(lo,1)..hi-1; the
        // analytic routine (FP64, level by level) keeps E_0 = W_1 / (4 pi rho),
        // which the marching inverts, and the static level hi+1, which the
        // moment conditions of k_emit need.
        if (sflag && sflag[p])
            k = ((m == 1 && l0 == 0) || (P.sur_project && m == hi[p] + 1)) ? near_key : kSurKey;
        key[it] = (uint8_t)k;
        atomicAdd(&sh[k], 1u);
    }
    __syncthreads();
    for (int k = threadIdx.x; k < kNumKeys; k += blockDim.x) bh[(int64_t)k * gridDim.x + blockIdx.x] = sh[k];
}

// Pass 2: block b scatters its range starting from the exclusive offsets
// bo[key*G + b] (ranks inside a block come from LDS atomics).
__global__ void __launch_bounds__(256) k_item_scatter(int64_t nitems, int64_t per, const uint8_t* __restrict__ key,
                                                      const unsigned long long* __restrict__ bo,
                                                      int32_t* __restrict__ perm) {
    __shared__ unsigned long long cur[kNumKeys];
    for (int k = threadIdx.x; k < kNumKeys; k += blockDim.x) cur[k] = bo[(int64_t)k * gridDim.x + blockIdx.x];
    __syncthreads();
    const int64_t a = blockIdx.x * per;
    const int64_t e = a + per < nitems ? a + per : nitems;
    for (int64_t it = a + threadIdx.x; it < e; it += blockDim.x) {
        unsigned long long pos = atomicAdd(&cur[key[it]], 1ull);
        perm[pos] = (int32_t)it;
    }
}

// Waves per SIMD the per-level, node-task, kP-range and post-processing
// integral kernels are built for.  The node evaluation needs about 380
// registers per lane: at one wave per SIMD it runs without spills (256 VGPRs
// + AGPRs), at two it spills ~450 B per lane to scratch (~980 B for the kP
// ranges).  One wave is measured faster on gfx950 (1280-triangle sphere, 128
// steps: per-level stage 46.6 vs 47.7 ms, post-processing 52.0 vs 54.0 ms,
// kP ranges 3.2 vs 4.6 ms and 1.8 vs 3.1 ms).
#ifndef EBEM_LEVELS_WAVES
#define EBEM_LEVELS_WAVES 1
#endif

// Persistent grid-stride kernel over (pair, level) items of a row chunk.
// The whole node evaluation is inlined (an out-of-line call spills ~1 KB of
// stack per lane; out-of-line region routines are 16% slower).
// W[item*6 + q]: symmetric 3x3 in the global frame, NOT yet divided by 4 pi rho.
template <int NGH, typename R>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(EBEM_LEVELS_WAVES))) k_levels(const FieldTri* __restrict__ tri, const double* __restrict__ gh,
                                                AsmParams P, int r0, int64_t nitems,
                                                const int32_t* __restrict__ item_pair,
                                                const int64_t* __restrict__ off, const int16_t* __restrict__ lo,
                                                const int32_t* __restrict__ perm, double* __restrict__ W) {
    for (int64_t k = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; k < nitems;
         k += (int64_t)gridDim.x * blockDim.x) {
        const int64_t it = perm ? (int64_t)perm[k] : k;
        const int32_t p = item_pair[it];
        const int i = r0 + p / P.N, j = p % P.N;
        const int l0 = lo[p];
        const int mlo = l0 - 1 < 1 ? 1 : l0 - 1;
        const int m = mlo + (int)(it - off[p]);
        const FieldTri& F = tri[j];
        const double* g = gh + (int64_t)i * 4 * 32;  // 28 nodes max: x,y,z rows then weights
        Level K = make_level((double)m * P.dt, P.cP, P.cS);
        NodeAcc acc;
        acc.s = 0.0;
        for (int q = 0; q < 6; q++) acc.m[q] = 0.0;
#pragma unroll 1
        for (int q = 0; q < NGH; q++) {
            double x[3] = {g[q], g[32 + q], g[64 + q]};
            NodeGeo G = node_geo(x, F);
            node_eval<R>(G, K, g[96 + q], acc);
        }
        double Wv[6];
        finish_rotation(acc, F.C, Wv);
        double* o = W + it * 6;
#pragma unroll
        for (int q = 0; q < 6; q++) o[q] = Wv[q];
    }
}

// ---- per-level items by outer node --------------------------------------
//
// k_levels gives a lane a whole item, so at each step of its node loop a
// wavefront runs the region code of every front state its lanes' nodes are
// in.  Here the outer nodes of the per-level items are the units of work:
// k_node_keys keys each (item, node) task by the node's front state
// (classify: S and P circles crossing, only P, or P alone, and where each
// circle's radius falls between the node's nearest and farthest distance to
// the field triangle), tasks whose node no front reaches are dropped, a
// counting sort by key orders the rest, k_node_eval integrates one node per
// lane into its own slot, and k_node_reduce adds an item's node results in
// node order and applies the rotation.  Task it * NGH + q is node q of item
// it (the chunk's item index); its slot is part[(it * NGH + q) * 8 ..]
// (7 values and one pad, a 64-byte line per task).
//
// Static nodes: once both circles of a node lie beyond every child's edge
// (node_static_pattern), node_eval returns the same bits at every later level
// of the pair.  A task whose node is static with a margin (node_static_margin)
// at its level and at the level before is not evaluated: its key says how
// many levels back the node's first static level is, whose slot the
// reduction takes.  k_node_keys checks the exact pattern at each (pair,
// node)'s first static level; a failed check raises *bad, and the caller
// then keys the chunk again without reuse.
constexpr int kNodeKeys = 48;
// post-processing items keyed by point state: keys 0..kNodeKeys, below the range keys
static_assert(kNodeKeys < kKpFarKey, "point-state keys must precede the range keys");
constexpr uint8_t kNodeSkip = 255;  // no front reaches the node: nothing to add
// Reused task: key kNodeReuse + d, its value is the slot of the item d levels
// before it (1 <= d <= kNodeReuseMax); a task further than kNodeReuseMax
// levels from the node's first static level is evaluated again.
constexpr int kNodeReuse = 64;
constexpr int kNodeReuseMax = 254 - kNodeReuse;
static_assert(kNodeKeys <= kNodeReuse, "reuse keys follow the state keys");

__device__ inline int node_task_key(const NodeGeo& G, const Level& K) {
    double rhoP, rhoS;
    const int c = classify(G.zeta, K.rP, K.rS, G.dMIN, rhoP, rhoS);
    if (c == kNone) return -1;
    double dmax2 = 0.0;
#pragma unroll
    for (int k = 0; k < 3; k++) dmax2 = fmax(dmax2, G.t[k][0] * G.t[k][0] + G.t[k][1] * G.t[k][1]);
    const double dmin = G.dMIN, dmax = sqrt(dmax2);
    auto bin = [&](double rho) {
        if (rho < dmin) return 0;
        if (rho >= dmax) return 3;
        const int b = 1 + (int)(2.0 * (rho - dmin) / (dmax - dmin));
        return b > 2 ? 2 : b;
    };
    return (c - 1) * 16 + bin(rhoP) * 4 + bin(rhoS);
}

// Level m of item it of pair p, and the pair's first level.
__device__ inline int item_level(const int16_t* __restrict__ lo, const int64_t* __restrict__ off, int32_t p,
                                 int64_t it, int& mfirst) {
    const int l0 = lo[p];
    mfirst = l0 - 1 < 1 ? 1 : l0 - 1;
    return mfirst + (int)(it - off[p]);
}

// First level in [mfirst, mlast] at which node_static_margin holds (it holds
// at every later level: it only grows with t), or mlast + 1 if none: the
// level where c_S t clears the threshold, then corrected by evaluating the
// predicate itself on either side.
__device__ inline int node_first_static(const NodeGeo& G, int mfirst, int mlast, const AsmParams& P) {
    double dmax2 = 0.0;
    for (int k = 0; k < 3; k++) dmax2 = fmax(dmax2, G.t[k][0] * G.t[k][0] + G.t[k][1] * G.t[k][1]);
    const double rs = sqrt(dmax2 * 1.002 + 1.0e-5 + G.zeta * G.zeta);
    double est = ceil(rs / (P.cS * P.dt));
    int m = est < (double)mfirst ? mfirst : (est > (double)(mlast + 1) ? mlast + 1 : (int)est);
    auto stat = [&](int l) { return node_static_margin(G, make_level((double)l * P.dt, P.cP, P.cS)); };
    while (m > mfirst && stat(m - 1)) m--;
    while (m <= mlast && !stat(m)) m++;
    return m;
}

// An item k_item_keys put among the per-level items evaluated by node: item
// keys [klo, khi) (the FP64 near field, or the far field of the mixed form).
__device__ inline bool node_path_item(uint8_t ikey, int klo, int khi) { return ikey >= klo && ikey < khi; }

// (pair, node) tasks of the chunk's npairs pairs: block b takes pairs
// [b per, (b + 1) per) with all their nodes; a thread computes its node's
// projection and first static level once and keys the node at each of the
// pair's levels that are evaluated by node (ikey), counting the tasks per key
// in bh[key * gridDim + b] (key kNodeKeys: the reused tasks).  reuse: 0 none,
// 1 static reuse with the exact check at the first static level, 2 the same
// with every check reported failed (the caller's fallback, forced).
template <int NGH>
__global__ void __launch_bounds__(256) k_node_keys(const FieldTri* __restrict__ tri, const double* __restrict__ gh,
                                                   AsmParams P, int r0, int64_t npairs, int64_t per,
                                                   const int64_t* __restrict__ off, const int16_t* __restrict__ lo,
                                                   const int16_t* __restrict__ hi, const uint8_t* __restrict__ ikey,
                                                   int klo, int khi, const uint8_t* __restrict__ sflag, int reuse,
                                                   uint8_t* __restrict__ tkey, unsigned int* __restrict__ bh,
                                                   unsigned int* __restrict__ bad) {
    __shared__ unsigned int sh[kNodeKeys + 1];  // + reused tasks
    for (int k = threadIdx.x; k <= kNodeKeys; k += blockDim.x) sh[k] = 0;
    __syncthreads();
    const int64_t a = blockIdx.x * per;
    const int64_t e = a + per < npairs ? a + per : npairs;
    for (int64_t k = a * NGH + threadIdx.x; k < e * NGH; k += blockDim.x) {
        const int64_t p = k / NGH;
        const int q = (int)(k - p * NGH);
        const int l0 = lo[p], l1 = hi[p];
        if (l0 > l1) continue;
        const int mfirst = l0 - 1 < 1 ? 1 : l0 - 1, mlast = l1 + 1;
        const int64_t it0 = off[p];
        const int i = r0 + (int)(p / P.N), j = (int)(p % P.N);
        const double* g = gh + (int64_t)i * 4 * 32;
        const double x[3] = {g[q], g[32 + q], g[64 + q]};
        const NodeGeo G = node_geo(x, tri[j]);
        const bool can_reuse = reuse && !(sflag && sflag[p]);
        const int ms = can_reuse ? node_first_static(G, mfirst, mlast, P) : mlast + 1;
        // the levels after ms take the value evaluated at ms: the exact
        // pattern must hold there
        if (ms < mlast && node_path_item(ikey[it0 + (ms - mfirst)], klo, khi) &&
            (reuse == 2 || !node_static_pattern(G, make_level((double)ms * P.dt, P.cP, P.cS))))
            atomicOr(bad, 1u);
        for (int m = mfirst; m <= mlast; m++) {
            const int64_t it = it0 + (m - mfirst);
            if (!node_path_item(ikey[it], klo, khi)) continue;
            const int key = node_task_key(G, make_level((double)m * P.dt, P.cP, P.cS));
            int tk = key < 0 ? kNodeSkip : key;
            const int d = m - ms;
            if (key >= 0 && m > mfirst && d >= 1 && d <= kNodeReuseMax) tk = kNodeReuse + d;
            tkey[it * NGH + q] = (uint8_t)tk;
            if (tk < kNodeKeys) atomicAdd(&sh[tk], 1u);
            else if (tk != kNodeSkip) atomicAdd(&sh[kNodeKeys], 1u);
        }
    }
    __syncthreads();
    for (int k = threadIdx.x; k <= kNodeKeys; k += blockDim.x) bh[(int64_t)k * gridDim.x + blockIdx.x] = sh[k];
}

// Block b places the tasks k_node_keys' block b keyed (same partition).
template <int NGH>
__global__ void __launch_bounds__(256) k_node_scatter(int64_t npairs, int64_t per, const int64_t* __restrict__ off,
                                                      const int16_t* __restrict__ lo,
                                                      const int16_t* __restrict__ hi,
                                                      const uint8_t* __restrict__ ikey, int klo, int khi,
                                                      const uint8_t* __restrict__ tkey,
                                                      const unsigned long long* __restrict__ bo,
                                                      int32_t* __restrict__ tperm) {
    __shared__ unsigned long long cur[kNodeKeys];
    for (int k = threadIdx.x; k < kNodeKeys; k += blockDim.x) cur[k] = bo[(int64_t)k * gridDim.x + blockIdx.x];
    __syncthreads();
    const int64_t a = blockIdx.x * per;
    const int64_t e = a + per < npairs ? a + per : npairs;
    for (int64_t k = a * NGH + threadIdx.x; k < e * NGH; k += blockDim.x) {
        const int64_t p = k / NGH;
        const int q = (int)(k - p * NGH);
        const int l0 = lo[p], l1 = hi[p];
        if (l0 > l1) continue;
        const int mfirst = l0 - 1 < 1 ? 1 : l0 - 1, mlast = l1 + 1;
        const int64_t it0 = off[p];
        for (int m = mfirst; m <= mlast; m++) {
            const int64_t it = it0 + (m - mfirst);
            if (!node_path_item(ikey[it], klo, khi)) continue;
            const uint8_t key = tkey[it * NGH + q];
            if (key >= kNodeKeys) continue;
            const unsigned long long pos = atomicAdd(&cur[key], 1ull);
            tperm[pos] = (int32_t)(it * NGH + q);
        }
    }
}

// R: the arithmetic of the closed forms (float for the far field of the
// mixed form; the front cutting stays FP64).
template <int NGH, typename R>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(EBEM_LEVELS_WAVES)))
k_node_eval(const FieldTri* __restrict__ tri, const double* __restrict__ gh, AsmParams P, int r0, int64_t ntask,
            const int32_t* __restrict__ item_pair, const int64_t* __restrict__ off, const int16_t* __restrict__ lo,
            const int32_t* __restrict__ tperm, double* __restrict__ part) {
    for (int64_t s = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; s < ntask;
         s += (int64_t)gridDim.x * blockDim.x) {
        const int32_t task = tperm[s];
        const int64_t it = task / NGH;
        const int q = task - (int)it * NGH;
        const int32_t p = item_pair[it];
        const int i = r0 + p / P.N, j = p % P.N;
        int mfirst;
        const int m = item_level(lo, off, p, it, mfirst);
        const FieldTri& F = tri[j];
        const double* g = gh + (int64_t)i * 4 * 32;
        const Level K = make_level((double)m * P.dt, P.cP, P.cS);
        const double x[3] = {g[q], g[32 + q], g[64 + q]};
        const NodeGeo G = node_geo(x, F);
        NodeAcc acc;
        acc.s = 0.0;
        for (int e = 0; e < 6; e++) acc.m[e] = 0.0;
        node_eval<R>(G, K, g[96 + q], acc);
        double* o = part + (int64_t)task * 8;
        o[0] = acc.s;
#pragma unroll
        for (int e = 0; e < 6; e++) o[1 + e] = acc.m[e];
        o[7] = 0.0;
    }
}

template <int NGH>
__global__ void __launch_bounds__(256) k_node_reduce(const FieldTri* __restrict__ tri, AsmParams P, int64_t nitems,
                                                     const int32_t* __restrict__ item_pair,
                                                     const int32_t* __restrict__ perm,
                                                     const uint8_t* __restrict__ tkey,
                                                     const double* __restrict__ part, double* __restrict__ W) {
    const int64_t k = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (k >= nitems) return;
    const int64_t it = perm[k];
    const int32_t p = item_pair[it];
    const int j = p % P.N;
    NodeAcc acc;
    acc.s = 0.0;
    for (int e = 0; e < 6; e++) acc.m[e] = 0.0;
    for (int q = 0; q < NGH; q++) {
        const int key = tkey[it * NGH + q];
        if (key == kNodeSkip) continue;
        // a reused task takes the slot of its node's first static level
        const int64_t src = key >= kNodeReuse ? it - (key - kNodeReuse) : it;
        const double* o = part + (src * NGH + q) * 8;
        acc.s += o[0];
#pragma unroll
        for (int e = 0; e < 6; e++) acc.m[e] += o[1 + e];
    }
    double Wv[6];
    finish_rotation(acc, tri[j].C, Wv);
    double* w = W + it * 6;
#pragma unroll
    for (int e = 0; e < 6; e++) w[e] = Wv[e];
}

// One thread per kP range (the items of key kKpFarKey or kKpKey, the first level ka of the
// range): WP and WQ of node_eval_kp at the ends ka, kb, then
// W_m = WP + t_m^2 WQ for m = ka..kb into the range's W slots (consecutive
// items of the pair).  If some node is not in the assumed state, the range is
// evaluated level by level as in k_levels.  Region integrals in R (FP32 for
// the far-field pairs of the mixed-precision form).
template <int NGH, typename R>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(EBEM_LEVELS_WAVES))) k_levels_kp(
    const FieldTri* __restrict__ tri, const double* __restrict__ gh, AsmParams P, int r0, int64_t nitems,
    const int32_t* __restrict__ item_pair, const int16_t* __restrict__ ka, const int16_t* __restrict__ kb,
    const int32_t* __restrict__ perm, double* __restrict__ W, unsigned int* __restrict__ nfallback) {
    for (int64_t k = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; k < nitems;
         k += (int64_t)gridDim.x * blockDim.x) {
        const int64_t it = perm[k];
        const int32_t p = item_pair[it];
        const int i = r0 + p / P.N, j = p % P.N;
        const int ma = ka[p], mb = kb[p];
        const FieldTri& F = tri[j];
        const double* g = gh + (int64_t)i * 4 * 32;
        const Level KA = make_level((double)ma * P.dt, P.cP, P.cS);
        const Level KB = make_level((double)mb * P.dt, P.cP, P.cS);
        NodeAcc aP, aQ;
        aP.s = aQ.s = 0.0;
        for (int q = 0; q < 6; q++) aP.m[q] = aQ.m[q] = 0.0;
        bool ok = true;
#pragma unroll 1
        for (int q = 0; q < NGH && ok; q++) {
            double x[3] = {g[q], g[32 + q], g[64 + q]};
            NodeGeo G = node_geo(x, F);
            ok = node_eval_kp<R>(G, KA, KB, g[96 + q], aP, aQ);
        }
        if (ok) {
            double WP[6], WQ[6];
            finish_rotation(aP, F.C, WP);
            finish_rotation(aQ, F.C, WQ);
            for (int m = ma; m <= mb; m++) {
                const double t = (double)m * P.dt;
                const double t2 = t * t;
                double* o = W + (it + (m - ma)) * 6;
#pragma unroll
                for (int q = 0; q < 6; q++) o[q] = WP[q] + t2 * WQ[q];
            }
        } else {
            atomicAdd(nfallback, 1u);
            for (int m = ma; m <= mb; m++) {
                Level K = make_level((double)m * P.dt, P.cP, P.cS);
                NodeAcc acc;
                acc.s = 0.0;
                for (int q = 0; q < 6; q++) acc.m[q] = 0.0;
#pragma unroll 1
                for (int q = 0; q < NGH; q++) {
                    double x[3] = {g[q], g[32 + q], g[64 + q]};
                    NodeGeo G = node_geo(x, F);
                    node_eval<R>(G, K, g[96 + q], acc);
                }
                double Wv[6];
                finish_rotation(acc, F.C, Wv);
                double* o = W + (it + (m - ma)) * 6;
#pragma unroll
                for (int q = 0; q < 6; q++) o[q] = Wv[q];
            }
        }
    }
}

// Waves per SIMD the between-front kernels are compiled for.  One wave gives
// them 256 VGPRs + 256 AGPRs (the AGPRs hold what would spill to scratch at
// two waves: 1056 B/lane there): 10.3 vs 16.4 ms (assembly) and 8.8 vs 19.4 ms
// (post-processing) on the 1280-triangle sphere.
#ifndef EBEM_BF_WAVES
#define EBEM_BF_WAVES 1
#endif

// One thread per between-front range (the head items of kBfFarKey or kBfKey, the first
// level sa of the range): levels sa..sb in chunks of kBfChunk through
// node_eval_bf, which yields k_levels' values with the level-independent
// region terms evaluated once per chunk.  From a chunk whose state check
// fails, the rest of the range is evaluated level by level as in k_levels.
// Region integrals in R, as in k_levels_kp.
template <int NGH, typename R>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(EBEM_BF_WAVES))) k_levels_bf(
    const FieldTri* __restrict__ tri, const double* __restrict__ gh, AsmParams P, int r0, int64_t nitems,
    const int32_t* __restrict__ item_pair, const int16_t* __restrict__ sa, const int16_t* __restrict__ sb,
    const int32_t* __restrict__ perm, double* __restrict__ W, unsigned int* __restrict__ nfallback) {
    for (int64_t k = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; k < nitems;
         k += (int64_t)gridDim.x * blockDim.x) {
        const int64_t it = perm[k];
        const int32_t p = item_pair[it];
        const int i = r0 + p / P.N, j = p % P.N;
        const int ma = sa[p], mb = sb[p];
        const FieldTri& F = tri[j];
        const double* g = gh + (int64_t)i * 4 * 32;
        constexpr int CH = sizeof(R) == 4 ? kBfChunkF32 : kBfChunk;
        int m = ma;
        for (; m <= mb; m += CH) {
            const int n = mb - m + 1 < CH ? mb - m + 1 : CH;
            double t[CH];
#pragma unroll
            for (int l = 0; l < CH; l++) t[l] = (double)(m + (l < n ? l : 0)) * P.dt;
            NodeAcc acc[CH];
#pragma unroll
            for (int l = 0; l < CH; l++) {
                acc[l].s = 0.0;
                for (int q = 0; q < 6; q++) acc[l].m[q] = 0.0;
            }
            bool ok = true;
#pragma unroll 1
            for (int q = 0; q < NGH && ok; q++) {
                double x[3] = {g[q], g[32 + q], g[64 + q]};
                NodeGeo G = node_geo(x, F);
                ok = node_eval_bf<CH, R>(G, t, n, P.cP, P.cS, g[96 + q], acc);
            }
            if (!ok) break;
#pragma unroll
            for (int l = 0; l < CH; l++) {
                if (l < n) {
                    double Wv[6];
                    finish_rotation(acc[l], F.C, Wv);
                    double* o = W + (it + (m + l - ma)) * 6;
                    for (int q = 0; q < 6; q++) o[q] = Wv[q];
                }
            }
        }
        if (m <= mb) {
            atomicAdd(nfallback, 1u);
            for (; m <= mb; m++) {
                Level K = make_level((double)m * P.dt, P.cP, P.cS);
                NodeAcc acc;
                acc.s = 0.0;
                for (int q = 0; q < 6; q++) acc.m[q] = 0.0;
#pragma unroll 1
                for (int q = 0; q < NGH; q++) {
                    double x[3] = {g[q], g[32 + q], g[64 + q]};
                    NodeGeo G = node_geo(x, F);
                    node_eval<R>(G, K, g[96 + q], acc);
                }
                double Wv[6];
                finish_rotation(acc, F.C, Wv);
                double* o = W + (it + (m - ma)) * 6;
                for (int q = 0; q < 6; q++) o[q] = Wv[q];
            }
        }
    }
}

// One wave (64 threads) per row of the chunk.  Writes the blocks of lags
// lo..hi of every pair of the row into the segmented store.
//
// Surrogate pairs (sflag set): slot l+1 holds the network's 4 pi rho E_l for
// the lags max(lo,1)..hi-1, while slot 1 (if lo = 0: W_1 = 4 pi rho E_0) and
// slot hi+1 (W_{hi+1}, the static level: every outer node already sees the
// whole field triangle inside its S disk) come from the analytic routine.
// The exact blocks of every pair satisfy sum_l E_l = W_{hi+1} - W_hi = 0 and
// sum_l l E_l = -W_hi = -W_static (the second difference telescopes); the
// first makes z = 1 an exact root of the marching's symbol, and a predicted
// sequence that breaks it moves that root off the unit circle (the solve
// then grows geometrically).  So the predicted lags get one constant shift
// and E_hi is set so that both sums hold exactly.
__global__ void __launch_bounds__(64) k_emit(const int16_t* __restrict__ lo, const int16_t* __restrict__ hi,
                                             const int64_t* __restrict__ off, const double* __restrict__ W,
                                             const uint8_t* __restrict__ sflag, int N, int L, int r0,
                                             double inv4pirho, double fourpirho, int ghc, int sur_project,
                                             Store S) {
    extern __shared__ int64_t run[];  // L entries
    const int r = blockIdx.x;
    const int i = r0 + r;
    const int lane = threadIdx.x;
    for (int l = lane; l < L; l += 64) run[l] = 0;
    __syncthreads();
    const int64_t pbase = (int64_t)r * N;
    for (int j0 = 0; j0 < N; j0 += 64) {
        int j = j0 + lane;
        int a = 1 << 30, b = -1;
        int64_t o = 0;
        bool direct = false, sur = false;
        if (j < N) {
            a = lo[pbase + j];
            b = hi[pbase + j];
            o = off[pbase + j];
            direct = ghc != 0;
            sur = !direct && sflag && sflag[pbase + j];
        }
        int amin = a, bmax = b;
        for (int d = 32; d >= 1; d >>= 1) {
            amin = min(amin, __shfl_xor(amin, d));
            bmax = max(bmax, __shfl_xor(bmax, d));
        }
        const int mlo = a - 1 < 1 ? 1 : a - 1;
        double shift[6] = {0, 0, 0, 0, 0, 0}, ehi[6] = {0, 0, 0, 0, 0, 0};
        if (sur && sur_project && b >= 1) {
            double S0[6], S1[6];
            const double* e0 = W + o * 6;  // slot of level 1 (mlo = 1 when a = 0)
#pragma unroll
            for (int q = 0; q < 6; q++) {
                S0[q] = (a == 0) ? e0[q] : 0.0;
                S1[q] = 0.0;
            }
            int n = 0;
            double m1 = 0.0;
            for (int l = (a > 1 ? a : 1); l < b; l++) {
                const double* e = W + (o + (l + 1 - mlo)) * 6;
#pragma unroll
                for (int q = 0; q < 6; q++) {
                    S0[q] += e[q];
                    S1[q] += (double)l * e[q];
                }
                n++;
                m1 += (double)l;
            }
            const double* ws = W + (o + (b + 1 - mlo)) * 6;
            const double den = m1 - (double)b * (double)n;  // < 0 whenever n > 0
#pragma unroll
            for (int q = 0; q < 6; q++) {
                shift[q] = n > 0 ? ((double)b * S0[q] - S1[q] - ws[q]) / den : 0.0;
                ehi[q] = -S0[q] - (double)n * shift[q];
            }
        }
        for (int l = amin; l <= bmax; l++) {
            bool in = (a <= l && l <= b);
            unsigned long long mask = __ballot(in);
            int64_t base = run[l];
            if (in) {
                int rank = __popcll(mask & ((1ull << lane) - 1ull));
                int64_t pos = S.seg[(int64_t)i * L + l] + base + rank;
                const double* w1 = W + (o + (l + 1 - mlo)) * 6;
                double e[6];
#pragma unroll
                for (int q = 0; q < 6; q++) e[q] = 1.0 * w1[q];
                if (direct) {
                    // numeric form: the slot already holds 4 pi rho E_l
                } else if (sur && !sur_project) {
                    // raw network output (E_0 still analytic)
                } else if (sur) {
                    if (l == b && l >= 1) {
#pragma unroll
                        for (int q = 0; q < 6; q++) e[q] = ehi[q];
                    } else if (l >= 1) {
#pragma unroll
                        for (int q = 0; q < 6; q++) e[q] = e[q] + shift[q];
                    }  // l = 0: W_1 = 4 pi rho E_0 from the analytic routine
                } else if (l >= 1) {
                    const double* w0 = W + (o + (l - mlo)) * 6;
#pragma unroll
                    for (int q = 0; q < 6; q++) e[q] = e[q] + (-2.0) * w0[q];
                }
                if (!direct && !sur && l >= 2) {
                    const double* wm = W + (o + (l - 1 - mlo)) * 6;
#pragma unroll
                    for (int q = 0; q < 6; q++) e[q] = e[q] + wm[q];
                }
                if (!ghc) {  // calcMatrixBlock.m:77
#pragma unroll
                    for (int q = 0; q < 6; q++) {
                        double v = e[q] * inv4pirho;
                        S.v[q][pos] = (fabs(v) < 1.0e-14) ? 0.0 : v;
                    }
                } else {  // kernelV.cu:150-157 off the diagonal, none on it (calcSubBlockDiagV.m)
#pragma unroll
                    for (int q = 0; q < 6; q++)
                        S.v[q][pos] = (j == i || fabs(e[q]) > 1.0e-14) ? e[q] / fourpirho : 0.0;
                }
                S.col[pos] = j;
            }
            __syncthreads();
            if (lane == 0) run[l] = base + __popcll(mask);
            __syncthreads();
        }
    }
}

// One thread per analytic region call (ai_region): kind[k], args 8, info 12,
// B 54 and ris 36 doubles per call.
__global__ void __launch_bounds__(64) k_ai_regions(int n, const int32_t* __restrict__ kind,
                                                   const double* __restrict__ args, const double* __restrict__ info,
                                                   const double* __restrict__ B, double* __restrict__ ris) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n) return;
    ai_region(kind[k], args + 8 * (int64_t)k, info + 12 * (int64_t)k, B + 54 * (int64_t)k, ris + 36 * (int64_t)k);
}

// Dense export of lag l: out (3N x 3N col-major), one thread per block of the segment.
// Rows [r0, r0 + gridDim.y) of E_lag into a dense column-major array with
// leading dimension ld (3 x rows by 3N).
__global__ void k_dense_lag(Store S, int N, int L, int lag, int r0, int64_t ld, double* __restrict__ out) {
    int i = r0 + blockIdx.y;
    int64_t b0 = S.seg[(int64_t)i * L + lag], b1 = S.seg[(int64_t)i * L + lag + 1];
    for (int64_t b = b0 + blockIdx.x * (int64_t)blockDim.x + threadIdx.x; b < b1; b += (int64_t)gridDim.x * blockDim.x) {
        int j = S.col[b];
        const int ii[9] = {0, 0, 0, 1, 1, 1, 2, 2, 2}, jj[9] = {0, 1, 2, 0, 1, 2, 0, 1, 2};
        const int qq[9] = {0, 1, 2, 1, 3, 4, 2, 4, 5};
        for (int k = 0; k < 9; k++)
            out[(3 * (int64_t)blockIdx.y + ii[k]) + ld * (3 * (int64_t)j + jj[k])] = S.v[qq[k]][b];
    }
}

// ---------------------------------------------------------------- MOT --

// r = beta_n - sum_{l=1}^{lmax} E_l alpha_{n-l}; one workgroup of 256 per test row.
// alpha: 3N x nT column-major (column n contiguous).
// Workgroup sum of v (all 256 threads) returned on every thread of wave 0.
__device__ inline double wg_sum(double v, double* red4) {
    for (int d = 32; d >= 1; d >>= 1) v += __shfl_xor(v, d);
    const int w = threadIdx.x >> 6;
    __syncthreads();
    if ((threadIdx.x & 63) == 0) red4[w] = v;
    __syncthreads();
    return (red4[0] + red4[1]) + (red4[2] + red4[3]);
}

// The history sum is taken kHistBlock steps at a time.  For steps
// n = n0 + b (b < nb <= kHistBlock) the terms with lag l > b only involve
// densities alpha_{n-l} with n - l < n0, known before the block starts:
// k_history_far streams the store once for the whole block and leaves
// part_b = sum_{l=b+1}^{min(n0+b, L-1)} E_l alpha_{n0+b-l};
// k_history then adds the remaining lags 1..min(b, L-1) step by step.
// Store traffic per block: one pass plus lags 1..b per step, instead of one
// full pass per step.
#ifndef EBEM_HIST_BLOCK
#define EBEM_HIST_BLOCK 8
#endif
constexpr int kHistBlock = EBEM_HIST_BLOCK;

// Far part: one workgroup per test row.  The store keeps a row's lag
// segments contiguous, so the row's lags 1..lmax are one flat range
// [seg(i, 1), seg(i, lmax + 1)) that the workgroup walks with stride
// blockDim; each thread tracks the lag of its block from the row's segment
// starts, staged in LDS (kFlatSeg lags at a time).  Against a loop over the
// lags with a strided loop over each segment (segments of a few hundred
// blocks, each ending in a partly filled pass): 10.6 ms instead of 11.7 ms
// for the 1280-triangle sphere at 128 steps.  Taking two blocks per thread
// and iteration (all loads first), splitting a row between two workgroups,
// and blocks of 12, 16 or 24 steps were measured no faster.  alpha stays
// step-major: the blocks of a row segment are in column order, so a wave's
// reads of one step share cache lines (a triangle-major copy was 2.5x
// slower).
constexpr int kFlatSeg = 1024;

__global__ void __launch_bounds__(256) k_history_far(Store S, int N, int L, int n0, int nb,
                                                     const double* __restrict__ alpha, double* __restrict__ part) {
    __shared__ double red4[4];
    __shared__ int64_t sseg[kFlatSeg + 1];
    const int i = blockIdx.x;
    const int64_t n3 = 3 * (int64_t)N;
    double acc[kHistBlock][3];
#pragma unroll
    for (int q = 0; q < kHistBlock; q++) acc[q][0] = acc[q][1] = acc[q][2] = 0.0;
    const int lmax = min(L - 1, n0 + nb - 1);
    for (int lc = 1; lc <= lmax; lc += kFlatSeg) {
        const int nl = min(lmax - lc + 1, kFlatSeg);  // lags lc .. lc + nl - 1
        __syncthreads();
        for (int k = threadIdx.x; k <= nl; k += blockDim.x) sseg[k] = S.seg[(int64_t)i * L + lc + k];
        __syncthreads();
        const int64_t r1 = sseg[nl];
        int li = 0;  // the thread's current block is in lag lc + li
        for (int64_t b = sseg[0] + threadIdx.x; b < r1; b += blockDim.x) {
            while (b >= sseg[li + 1]) li++;
            const int l = lc + li;
            const int bmin = max(0, l - n0), bmax = min(nb - 1, l - 1);
            const int j = S.col[b];
            const double w00 = S.v[0][b], w01 = S.v[1][b], w02 = S.v[2][b];
            const double w11 = S.v[3][b], w12 = S.v[4][b], w22 = S.v[5][b];
#pragma unroll
            for (int q = 0; q < kHistBlock; q++) {
                if (q < bmin || q > bmax) continue;
                const double* al = alpha + (int64_t)(n0 + q - l) * n3 + 3 * (int64_t)j;
                const double x0 = al[0], x1 = al[1], x2 = al[2];
                acc[q][0] += w00 * x0 + w01 * x1 + w02 * x2;
                acc[q][1] += w01 * x0 + w11 * x1 + w12 * x2;
                acc[q][2] += w02 * x0 + w12 * x1 + w22 * x2;
            }
        }
    }
#pragma unroll
    for (int q = 0; q < kHistBlock; q++)
#pragma unroll
        for (int c = 0; c < 3; c++) {
            const double v = wg_sum(acc[q][c], red4);
            if (threadIdx.x == 0 && q < nb) part[(int64_t)q * n3 + 3 * i + c] = v;
        }
}

// r = beta_n - part - sum_{l=1}^{lmax} E_l alpha_{n-l} (part may be null);
// one workgroup of 256 per test row.
__global__ void __launch_bounds__(256) k_history(Store S, int N, int L, int n, int lmax,
                                                 const double* __restrict__ beta, const double* __restrict__ alpha,
                                                 const double* __restrict__ part, double* __restrict__ r) {
    const int i = blockIdx.x;
    const int64_t n3 = 3 * (int64_t)N;
    double a0 = 0.0, a1 = 0.0, a2 = 0.0;
    for (int l = 1; l <= lmax; l++) {
        const int64_t b0 = S.seg[(int64_t)i * L + l], b1 = S.seg[(int64_t)i * L + l + 1];
        const double* al = alpha + (int64_t)(n - l) * n3;
        for (int64_t b = b0 + threadIdx.x; b < b1; b += blockDim.x) {
            const int j = S.col[b];
            const double x0 = al[3 * j], x1 = al[3 * j + 1], x2 = al[3 * j + 2];
            const double w00 = S.v[0][b], w01 = S.v[1][b], w02 = S.v[2][b];
            const double w11 = S.v[3][b], w12 = S.v[4][b], w22 = S.v[5][b];
            a0 += w00 * x0 + w01 * x1 + w02 * x2;
            a1 += w01 * x0 + w11 * x1 + w12 * x2;
            a2 += w02 * x0 + w12 * x1 + w22 * x2;
        }
    }
    // workgroup reduction (4 waves)
    for (int d = 32; d >= 1; d >>= 1) {
        a0 += __shfl_xor(a0, d);
        a1 += __shfl_xor(a1, d);
        a2 += __shfl_xor(a2, d);
    }
    __shared__ double red[4][3];
    const int w = threadIdx.x >> 6;
    if ((threadIdx.x & 63) == 0) {
        red[w][0] = a0;
        red[w][1] = a1;
        red[w][2] = a2;
    }
    __syncthreads();
    if (threadIdx.x < 3) {
        int c = threadIdx.x;
        double s = red[0][c] + red[1][c] + red[2][c] + red[3][c];
        if (part) s += part[3 * i + c];
        r[3 * i + c] = beta[(int64_t)n * n3 + 3 * i + c] - s;
    }
}

// y[row] = sum_k A[row][k] x[k]; A row-major n x n; one wave per row, 4 rows per block.
// If yT is set, the value also goes to yT[((row / 3) (nT + 1) + 1 + step) 3 +
// row % 3] (the density's triangle-major copy, each triangle's row led by one
// zero step) and the same place of yD receives the step difference
// y[row] - (y - n)[row] (the previous step's value; 0 at step 0), both read by
// k_pp_conv.
__global__ void __launch_bounds__(256) k_gemv_rows(const double* __restrict__ A, const double* __restrict__ x,
                                                   double* __restrict__ y, int n, double* __restrict__ yT,
                                                   double* __restrict__ yD, int nT, int step) {
    const int row = blockIdx.x * 4 + (threadIdx.x >> 6);
    const int lane = threadIdx.x & 63;
    if (row >= n) return;
    const double* a = A + (int64_t)row * n;
    double s0 = 0.0, s1 = 0.0;
    int k = lane;
    for (; k + 64 < n; k += 128) {
        s0 += a[k] * x[k];
        s1 += a[k + 64] * x[k + 64];
    }
    if (k < n) s0 += a[k] * x[k];
    double s = s0 + s1;
    for (int d = 32; d >= 1; d >>= 1) s += __shfl_xor(s, d);
    if (lane == 0) {
        y[row] = s;
        if (yT) {
            const int64_t o = ((int64_t)(row / 3) * (nT + 1) + 1 + step) * 3 + row % 3;
            yT[o] = s;
            yD[o] = s - (step > 0 ? y[(int64_t)row - n] : 0.0);
        }
    }
}

// A[i + n i] = v for a zeroed n x n array (identity for v = 1).
__global__ void k_set_diag(double* __restrict__ A, int n, double v) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) A[(int64_t)i * n + i] = v;
}

// E_0 dense (col-major) from the lag-0 segments; one thread per block.
__global__ void k_e0_dense(Store S, int N, int L, double* __restrict__ A) {
    int i = blockIdx.y;
    int64_t b0 = S.seg[(int64_t)i * L + 0], b1 = S.seg[(int64_t)i * L + 1];
    int64_t n3 = 3 * (int64_t)N;
    for (int64_t b = b0 + blockIdx.x * (int64_t)blockDim.x + threadIdx.x; b < b1; b += (int64_t)gridDim.x * blockDim.x) {
        int j = S.col[b];
        const int ii[9] = {0, 0, 0, 1, 1, 1, 2, 2, 2}, jj[9] = {0, 1, 2, 0, 1, 2, 0, 1, 2};
        const int qq[9] = {0, 1, 2, 1, 3, 4, 2, 4, 5};
        for (int k = 0; k < 9; k++) A[(3 * (int64_t)i + ii[k]) + n3 * (3 * (int64_t)j + jj[k])] = S.v[qq[k]][b];
    }
}

// out (row-major) = in (col-major) ; n x n, 32x32 tiles through LDS.
__global__ void k_transpose(const double* __restrict__ in, double* __restrict__ out, int n) {
    __shared__ double tile[32][33];
    int bx = blockIdx.x * 32, by = blockIdx.y * 32;
    int tx = threadIdx.x, ty = threadIdx.y;  // 32 x 8
    for (int k = ty; k < 32; k += 8) {
        int r = by + k, c = bx + tx;  // in: element (row=c? ) -- read in[c_row + n * col]
        if (r < n && c < n) tile[k][tx] = in[(int64_t)r * n + c];  // column r, row c of the col-major matrix
    }
    __syncthreads();
    for (int k = ty; k < 32; k += 8) {
        int r = bx + k, c = by + tx;
        if (r < n && c < n) out[(int64_t)r * n + c] = tile[tx][k];
    }
}

// Single point / single triangle: M(x, t) of ppA/BEMenerg_postProcA_calcMatrixBlock.m
// (one outer point, weight 1, no cleanup), symmetric 6 values scaled by 1/(4 pi rho).
__global__ void k_point_block(const FieldTri* __restrict__ tri, int j, const double* __restrict__ x, double t,
                              double cP, double cS, double inv4pirho, double* __restrict__ out6) {
    if (threadIdx.x != 0) return;
    const FieldTri& F = tri[j];
    Level K = make_level(t, cP, cS);
    NodeAcc acc;
    acc.s = 0.0;
    for (int q = 0; q < 6; q++) acc.m[q] = 0.0;
    double xx[3] = {x[0], x[1], x[2]};
    NodeGeo G = node_geo(xx, F);
    node_eval(G, K, 1.0, acc);
    double W[6];
    finish_rotation(acc, F.C, W);
    for (int q = 0; q < 6; q++) out6[q] = W[q] * inv4pirho;
}

// ---------------------------------------------------------- post-process --
//
// ppA/BEMenerg_postProcA_singlePoint.m:
//   u(x,t) = sum_{c=0}^{nhat} M(x, t - c dt) (alpha_c - alpha_{c-1}),  nhat = ceil(t/dt) - 1,
// rewritten with tau_m = delta + m dt (delta = t - nhat dt, m = nhat - c) as
//   u = sum_m D_m alpha_{nhat - m},  D_m = M_m - M_{m-1},  M_m = M(x, tau_m),
// where D_m vanishes before the P front reaches the triangle and once the S
// front has passed it.  Times sharing delta share every M_m.

struct PPParams {
    int N, nT;
    double dt, cP, cS, delta, inv4pirho;
    int mmax;
    double trep;  // representative time of the group (largest nhat)
    int nrep;     // its nhat: tau_m = trep - (nrep - m) dt, as singlePoint.m:37
    int kp;       // 1: kP ranges of m through k_pp_M_kp (pp_kp_range)
};

__device__ inline double pp_tau(const PPParams& P, int m) {
#pragma clang fp contract(off)
    return (m == P.nrep) ? P.trep : P.trep - (double)(P.nrep - m) * P.dt;
}

// kP range [ka, kb] of m for point x and field triangle F within the window
// [lo, hi], and the between-front range [sa, sb] after it (kp_range with the
// single outer point x and tau_m ~ delta + m dt; the 1e-7 margins cover the
// difference to pp_tau).  first > last: none.
__device__ inline void pp_kp_range(const PPParams& P, double x0, double x1, double x2, const FieldTri& F,
                                   const double* __restrict__ vF, int lo, int hi, int& ka, int& kb, int& sa,
                                   int& sb) {
#pragma clang fp contract(off)
    ka = 1;
    kb = 0;
    sa = 1;
    sb = 0;
    double cmin2 = 1e300, dmax2 = 0.0;
    for (int v = 0; v < 3; v++) {
        const int w = (v + 1) % 3;
        double a = vF[3 * w] - vF[3 * v], b = vF[3 * w + 1] - vF[3 * v + 1], c = vF[3 * w + 2] - vF[3 * v + 2];
        cmin2 = fmin(cmin2, a * a + b * b + c * c);
        double d0 = x0 - vF[3 * v], d1 = x1 - vF[3 * v + 1], d2 = x2 - vF[3 * v + 2];
        dmax2 = fmax(dmax2, d0 * d0 + d1 * d1 + d2 * d2);
    }
    if (!(0.25 * cmin2 > 1.0e-5)) return;
    double d0 = x0 - F.v0[0], d1 = x1 - F.v0[1], d2 = x2 - F.v0[2];
    double z = fabs(F.n[0] * d0 + F.n[1] * d1 + F.n[2] * d2);
    if (z < 1.0e-6) z = 0.0;
    const double d3 = tri_dist3(x0, x1, x2, vF);
    const double fa = ceil((sqrt(dmax2) * (1.0 + 1e-7) / P.cP - P.delta) / P.dt);
    const double fb = ceil((z * (1.0 - 1e-7) / P.cS - P.delta) / P.dt) - 1.0;
    const double fs = ceil(((d3 * (1.0 - 1e-7) - 1e-9) / P.cS - P.delta) / P.dt) - 1.0;
    const double a = fmax(fa, (double)lo), b = fmin(fb, (double)hi);
    if (b - a + 1.0 >= 3.0) {
        ka = (int)a;
        kb = (int)b;
    }
    const double s0 = ka <= kb ? (double)(kb + 1) : a, s1 = fmin(fs, (double)hi);
    if (s1 - s0 + 1.0 >= 2.0) {
        sa = (int)s0;
        sb = (int)s1;
    }
}

// Window [ms, me] of m per (point, triangle); count = me - ms + 1 (0 if empty).
//  * M_m = 0 while c_P tau_m < Dlo = |x - c_F| - R_F (R_F the triangle's
//    circumradius about its centroid, Dlo a lower bound of the distance from
//    x to the triangle): ms is the first m that may reach it, so M_{ms-1} = 0;
//  * M_m takes its static value once c_S tau_m > Dmax = max_v |x - v| (the
//    whole triangle inside the S sphere: the P shell misses it and the S-ball
//    integrals no longer depend on tau), so D_m = 0 for m > me with me the
//    first such m.
// Relative margins 1e-12 keep ties on the computed side.
__global__ void __launch_bounds__(256) k_pp_windows(const FieldTri* __restrict__ tri, const double* __restrict__ verts,
                                                    const double* __restrict__ xs, int p0, int np, PPParams P,
                                                    int32_t* __restrict__ ms, int32_t* __restrict__ me,
                                                    int32_t* __restrict__ ka, int32_t* __restrict__ kb,
                                                    int32_t* __restrict__ sa, int32_t* __restrict__ sb) {
    int64_t tot = (int64_t)np * P.N;
    for (int64_t q = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; q < tot; q += (int64_t)gridDim.x * blockDim.x) {
        int p = p0 + (int)(q / P.N), j = (int)(q % P.N);
        const FieldTri& F = tri[j];
        const double* v = verts + (int64_t)j * 9;
        const double x0 = xs[3 * p], x1 = xs[3 * p + 1], x2 = xs[3 * p + 2];
        double rf2 = 0.0, dmax2 = 0.0;
        for (int k = 0; k < 3; k++) {
            double a0 = v[3 * k] - F.center[0], a1 = v[3 * k + 1] - F.center[1], a2 = v[3 * k + 2] - F.center[2];
            rf2 = fmax(rf2, a0 * a0 + a1 * a1 + a2 * a2);
            double b0 = x0 - v[3 * k], b1 = x1 - v[3 * k + 1], b2 = x2 - v[3 * k + 2];
            dmax2 = fmax(dmax2, b0 * b0 + b1 * b1 + b2 * b2);
        }
        double d0 = x0 - F.center[0], d1 = x1 - F.center[1], d2 = x2 - F.center[2];
        const double dlo = fmax(0.0, sqrt(d0 * d0 + d1 * d1 + d2 * d2) - sqrt(rf2));
        const double a = ceil((dlo * (1.0 - 1e-12) / P.cP - P.delta) / P.dt);
        int lo = (int)fmax(0.0, fmin(a, (double)P.mmax + 1.0));
        const double b = floor((sqrt(dmax2) * (1.0 + 1e-12) / P.cS - P.delta) / P.dt) + 1.0;
        int hi = (int)fmax(-1.0, fmin(b, (double)P.mmax));
        ms[q] = lo;
        me[q] = hi;
        if (P.kp) {
            int a0 = 1, b0 = 0, c0 = 1, e0 = 0;
            if (lo <= hi) pp_kp_range(P, x0, x1, x2, F, v, lo, hi, a0, b0, c0, e0);
            ka[q] = a0;
            kb[q] = b0;
            sa[q] = c0;
            sb[q] = e0;
        }
    }
}

// One workgroup per point of the chunk: exclusive prefix of the window sizes
// (item offsets) and the item -> (point, triangle) map.
__global__ void __launch_bounds__(256) k_pp_items(const int32_t* __restrict__ ms, const int32_t* __restrict__ me, int N,
                                                  const int64_t* __restrict__ rowbase, int64_t* __restrict__ off,
                                                  int32_t* __restrict__ item_q) {
    __shared__ long long part[256];
    const int r = blockIdx.x;
    const int nper = (N + blockDim.x - 1) / blockDim.x;
    const int j0 = threadIdx.x * nper;
    const int64_t base = (int64_t)r * N;
    long long s = 0;
    for (int k = 0; k < nper; k++) {
        int j = j0 + k;
        if (j < N) s += max(0, me[base + j] - ms[base + j] + 1);
    }
    part[threadIdx.x] = s;
    __syncthreads();
    for (int d = 1; d < 256; d <<= 1) {
        long long v = (threadIdx.x >= d) ? part[threadIdx.x - d] : 0;
        __syncthreads();
        part[threadIdx.x] += v;
        __syncthreads();
    }
    long long run = rowbase[r] + (threadIdx.x ? part[threadIdx.x - 1] : 0);
    for (int k = 0; k < nper; k++) {
        int j = j0 + k;
        if (j >= N) break;
        int64_t q = base + j;
        int n = max(0, me[q] - ms[q] + 1);
        off[q] = run;
        for (int c = 0; c < n; c++) item_q[run + c] = (int32_t)q;
        run += n;
    }
}

// M_m(x_p, tri j) for every item; 6 symmetric values, /(4 pi rho), entries
// below 1e-14 zeroed (ppA/BEMenerg_postProcA_calcMatrix.m:20).
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(EBEM_LEVELS_WAVES))) k_pp_M(const FieldTri* __restrict__ tri, const double* __restrict__ xs, int p0,
                                              PPParams P, int64_t nitems, const int32_t* __restrict__ item_q,
                                              const int64_t* __restrict__ off, const int32_t* __restrict__ ms,
                                              const int32_t* __restrict__ perm, double* __restrict__ M) {
    for (int64_t kx = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; kx < nitems;
         kx += (int64_t)gridDim.x * blockDim.x) {
        const int64_t it = perm ? (int64_t)perm[kx] : kx;
        const int32_t q = item_q[it];
        const int p = p0 + q / P.N, j = q % P.N;
        const int m = ms[q] + (int)(it - off[q]);
        const double tau = pp_tau(P, m);
        double W[6] = {0, 0, 0, 0, 0, 0};
        if (tau > 0.0) {
            const FieldTri& F = tri[j];
            Level K = make_level(tau, P.cP, P.cS);
            NodeAcc acc;
            acc.s = 0.0;
            for (int k = 0; k < 6; k++) acc.m[k] = 0.0;
            double x[3] = {xs[3 * p], xs[3 * p + 1], xs[3 * p + 2]};
            NodeGeo G = node_geo(x, F);
            if (node_eval(G, K, 1.0, acc) != kNone) finish_rotation(acc, F.C, W);
        }
        double* o = M + it * 6;
#pragma unroll
        for (int k = 0; k < 6; k++) {
            double v = W[k] * P.inv4pirho;
            o[k] = fabs(v) < 1.0e-14 ? 0.0 : v;
        }
    }
}

// One thread per kP range of m (the items of key kKpKey, m = ka): WP, WQ of
// node_eval_kp for the single point at tau_ka, tau_kb, then
// M_m = (WP + tau_m^2 WQ) / (4 pi rho) with the 1e-14 cleanup for m = ka..kb
// (consecutive items of the point-triangle pair).  A range whose state check
// fails is evaluated per m as in k_pp_M.
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(EBEM_LEVELS_WAVES))) k_pp_M_kp(
    const FieldTri* __restrict__ tri, const double* __restrict__ xs, int p0, PPParams P, int64_t nheads,
    const int32_t* __restrict__ item_q, const int32_t* __restrict__ ka, const int32_t* __restrict__ kb,
    const int32_t* __restrict__ perm, double* __restrict__ M, unsigned int* __restrict__ nfallback) {
    for (int64_t kx = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; kx < nheads;
         kx += (int64_t)gridDim.x * blockDim.x) {
        const int64_t it = perm[kx];
        const int32_t q = item_q[it];
        const int p = p0 + q / P.N, j = q % P.N;
        const int ma = ka[q], mb = kb[q];
        const FieldTri& F = tri[j];
        double x[3] = {xs[3 * p], xs[3 * p + 1], xs[3 * p + 2]};
        const NodeGeo G = node_geo(x, F);
        const Level KA = make_level(pp_tau(P, ma), P.cP, P.cS);
        const Level KB = make_level(pp_tau(P, mb), P.cP, P.cS);
        NodeAcc aP, aQ;
        aP.s = aQ.s = 0.0;
        for (int k = 0; k < 6; k++) aP.m[k] = aQ.m[k] = 0.0;
        const bool ok = node_eval_kp(G, KA, KB, 1.0, aP, aQ);
        double WP[6], WQ[6];
        if (ok) {
            finish_rotation(aP, F.C, WP);
            finish_rotation(aQ, F.C, WQ);
        } else {
            atomicAdd(nfallback, 1u);
        }
        for (int m = ma; m <= mb; m++) {
            const double tau = pp_tau(P, m);
            double W[6];
            if (ok) {
                const double t2 = tau * tau;
#pragma unroll
                for (int k = 0; k < 6; k++) W[k] = WP[k] + t2 * WQ[k];
            } else {
                for (int k = 0; k < 6; k++) W[k] = 0.0;
                Level K = make_level(tau, P.cP, P.cS);
                NodeAcc acc;
                acc.s = 0.0;
                for (int k = 0; k < 6; k++) acc.m[k] = 0.0;
                if (node_eval(G, K, 1.0, acc) != kNone) finish_rotation(acc, F.C, W);
            }
            double* o = M + (it + (m - ma)) * 6;
#pragma unroll
            for (int k = 0; k < 6; k++) {
                double v = W[k] * P.inv4pirho;
                o[k] = fabs(v) < 1.0e-14 ? 0.0 : v;
            }
        }
    }
}

// One thread per between-front range of m (the items of key kBfKey, m = sa):
// M_m for m = sa..sb through node_eval_bf (one outer point, weight 1) in
// chunks of kPpBfChunk, the same values as k_pp_M; from a chunk whose state
// check fails, per m as in k_pp_M.
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(EBEM_BF_WAVES))) k_pp_M_bf(
    const FieldTri* __restrict__ tri, const double* __restrict__ xs, int p0, PPParams P, int64_t nheads,
    const int32_t* __restrict__ item_q, const int32_t* __restrict__ sa, const int32_t* __restrict__ sb,
    const int32_t* __restrict__ perm, double* __restrict__ M, unsigned int* __restrict__ nfallback) {
    for (int64_t kx = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; kx < nheads;
         kx += (int64_t)gridDim.x * blockDim.x) {
        const int64_t it = perm[kx];
        const int32_t q = item_q[it];
        const int p = p0 + q / P.N, j = q % P.N;
        const int ma = sa[q], mb = sb[q];
        const FieldTri& F = tri[j];
        double x[3] = {xs[3 * p], xs[3 * p + 1], xs[3 * p + 2]};
        const NodeGeo G = node_geo(x, F);
        int m = ma;
        for (; m <= mb; m += kPpBfChunk) {
            const int n = mb - m + 1 < kPpBfChunk ? mb - m + 1 : kPpBfChunk;
            double t[kPpBfChunk];
#pragma unroll
            for (int l = 0; l < kPpBfChunk; l++) t[l] = pp_tau(P, m + (l < n ? l : 0));
            NodeAcc acc[kPpBfChunk];
#pragma unroll
            for (int l = 0; l < kPpBfChunk; l++) {
                acc[l].s = 0.0;
                for (int k = 0; k < 6; k++) acc[l].m[k] = 0.0;
            }
            if (!node_eval_bf<kPpBfChunk>(G, t, n, P.cP, P.cS, 1.0, acc)) break;
#pragma unroll
            for (int l = 0; l < kPpBfChunk; l++) {
                if (l < n) {
                    double W[6];
                    finish_rotation(acc[l], F.C, W);
                    double* o = M + (it + (m + l - ma)) * 6;
                    for (int k = 0; k < 6; k++) {
                        double v = W[k] * P.inv4pirho;
                        o[k] = fabs(v) < 1.0e-14 ? 0.0 : v;
                    }
                }
            }
        }
        if (m <= mb) {
            atomicAdd(nfallback, 1u);
            for (; m <= mb; m++) {
                const double tau = pp_tau(P, m);
                double W[6] = {0, 0, 0, 0, 0, 0};
                if (tau > 0.0) {
                    Level K = make_level(tau, P.cP, P.cS);
                    NodeAcc acc;
                    acc.s = 0.0;
                    for (int k = 0; k < 6; k++) acc.m[k] = 0.0;
                    if (node_eval(G, K, 1.0, acc) != kNone) finish_rotation(acc, F.C, W);
                }
                double* o = M + (it + (m - ma)) * 6;
                for (int k = 0; k < 6; k++) {
                    double v = W[k] * P.inv4pirho;
                    o[k] = fabs(v) < 1.0e-14 ? 0.0 : v;
                }
            }
        }
    }
}

// One workgroup per group of kConvGroup points of the chunk, one wave per
// slice of triangles (j = s, s + S, s + 2S, ...); lane l of a wave takes the
// requested times kb + l + 64 i, i < kConvPerLane, so one pass covers
// 64 kConvPerLane times.  The time sum of postProcA_singlePoint.m,
// sum_m (M_m - M_{m-1}) alpha[n - m] over a (point, triangle) window [a, b]
// (M_{a-1} = 0; beyond b the differences vanish), is taken by parts:
//     sum_{m=a..b-1} M_m dalpha[n - m] + M_b alpha[n - b],
// dalpha[k] = alpha[k] - alpha[k-1], alpha at negative steps 0.  The density
// differences a lane loads at step n - m serve every point of the group whose
// window holds m (the points are ordered along a space-filling curve, so the
// windows of a group overlap), and each (point, m) costs the 9 FMAs of the
// symmetric 3x3 product with M in scalar registers.  Window bounds and M
// values are wave-uniform (scalar loads); only the density reads and the
// accumulation are per lane.  The slice partials are reduced through LDS,
// point by point.  aT, dT: alpha and dalpha as [triangle][1 + nT][3], each
// triangle's row led by one zero step, which a lane reads whenever its step
// n - m (or n - b) is negative: the reads need no branch, and the idle lanes
// of a wave all read that one zero step.  The point of chunk row r is
// written to column pout[p0 + r] of u.
#ifndef EBEM_CONV_SLICES
#define EBEM_CONV_SLICES 8
#endif
constexpr int kConvSlices = EBEM_CONV_SLICES;
constexpr int kConvPerLane = 2;
#ifndef EBEM_CONV_GROUP
#define EBEM_CONV_GROUP 2
#endif
constexpr int kConvGroup = EBEM_CONV_GROUP;

__global__ void __launch_bounds__(64 * kConvSlices) k_pp_conv(const double* __restrict__ M,
                                                            const int64_t* __restrict__ off,
                                                            const int32_t* __restrict__ ms,
                                                            const int32_t* __restrict__ me,
                                                            const double* __restrict__ aT,
                                                            const double* __restrict__ dT, int N, int nT, int p0,
                                                            int np, const int32_t* __restrict__ pout,
                                                            const int32_t* __restrict__ nhat,
                                                            const int32_t* __restrict__ kidx, int nk, int npts,
                                                            double* __restrict__ u) {
    constexpr int KP = 64 * kConvPerLane;  // times per pass
    constexpr int GP = kConvGroup;
    __shared__ double red[kConvSlices][KP][3];
    const int r0 = blockIdx.x * GP;  // chunk row of the group's first point
    const int ng = np - r0 < GP ? np - r0 : GP;
    const int lane = threadIdx.x % 64;
    const int slice = __builtin_amdgcn_readfirstlane((int)(threadIdx.x / 64));  // blockDim = 64 * kConvSlices
    for (int kb = 0; kb < nk; kb += KP) {
        int nh[kConvPerLane];
        double acc[GP][kConvPerLane][3];
#pragma unroll
        for (int i = 0; i < kConvPerLane; i++) {
            const int kk = kb + i * 64 + lane;
            nh[i] = kk < nk ? nhat[kk] : -1;
        }
#pragma unroll
        for (int g = 0; g < GP; g++)
#pragma unroll
            for (int i = 0; i < kConvPerLane; i++) acc[g][i][0] = acc[g][i][1] = acc[g][i][2] = 0.0;
        for (int j = slice; j < N; j += kConvSlices) {
            int a[GP], b[GP];
            int64_t o[GP];
            int A = nT, B = -1;  // union of the group's windows
#pragma unroll
            for (int g = 0; g < GP; g++) {
                a[g] = 1;
                b[g] = 0;
                o[g] = 0;
                if (g < ng) {
                    const int64_t q = (int64_t)(r0 + g) * N + j;
                    a[g] = ms[q];
                    b[g] = me[q];
                    o[g] = off[q];
                    if (a[g] <= b[g]) {
                        A = a[g] < A ? a[g] : A;
                        B = b[g] > B ? b[g] : B;
                    }
                }
            }
            if (A > B) continue;
            const double* dj = dT + ((int64_t)j * (nT + 1) + 1) * 3;  // step 0 of row j
            for (int m = A; m < B; m++) {
                double x[kConvPerLane][3];
#pragma unroll
                for (int i = 0; i < kConvPerLane; i++) {
                    const int k = nh[i] - m < -1 ? -1 : nh[i] - m;
                    const double* al = dj + k * 3;
                    x[i][0] = al[0];
                    x[i][1] = al[1];
                    x[i][2] = al[2];
                }
#pragma unroll
                for (int g = 0; g < GP; g++) {
                    if (m < a[g] || m >= b[g]) continue;
                    const double* Mm = M + (o[g] + (m - a[g])) * 6;
                    const double d0 = Mm[0], d1 = Mm[1], d2 = Mm[2], d3 = Mm[3], d4 = Mm[4], d5 = Mm[5];
#pragma unroll
                    for (int i = 0; i < kConvPerLane; i++) {
                        acc[g][i][0] = fma(d2, x[i][2], fma(d1, x[i][1], fma(d0, x[i][0], acc[g][i][0])));
                        acc[g][i][1] = fma(d4, x[i][2], fma(d3, x[i][1], fma(d1, x[i][0], acc[g][i][1])));
                        acc[g][i][2] = fma(d5, x[i][2], fma(d4, x[i][1], fma(d2, x[i][0], acc[g][i][2])));
                    }
                }
            }
            // M_b alpha[n - b]
            const double* aj = aT + ((int64_t)j * (nT + 1) + 1) * 3;
#pragma unroll
            for (int g = 0; g < GP; g++) {
                if (a[g] > b[g]) continue;
                const double* Mb = M + (o[g] + (b[g] - a[g])) * 6;
                const double d0 = Mb[0], d1 = Mb[1], d2 = Mb[2], d3 = Mb[3], d4 = Mb[4], d5 = Mb[5];
#pragma unroll
                for (int i = 0; i < kConvPerLane; i++) {
                    const int k = nh[i] - b[g] < -1 ? -1 : nh[i] - b[g];
                    const double* al = aj + k * 3;
                    const double x0 = al[0], x1 = al[1], x2 = al[2];
                    acc[g][i][0] = fma(d2, x2, fma(d1, x1, fma(d0, x0, acc[g][i][0])));
                    acc[g][i][1] = fma(d4, x2, fma(d3, x1, fma(d1, x0, acc[g][i][1])));
                    acc[g][i][2] = fma(d5, x2, fma(d4, x1, fma(d2, x0, acc[g][i][2])));
                }
            }
        }
#pragma unroll
        for (int g = 0; g < GP; g++) {
            if (g >= ng) break;
#pragma unroll
            for (int i = 0; i < kConvPerLane; i++) {
                red[slice][i * 64 + lane][0] = acc[g][i][0];
                red[slice][i * 64 + lane][1] = acc[g][i][1];
                red[slice][i * 64 + lane][2] = acc[g][i][2];
            }
            __syncthreads();
            const int64_t col = pout[p0 + r0 + g];
            for (int t = threadIdx.x; t < KP; t += blockDim.x) {
                const int kk = kb + t;
                if (kk >= nk) continue;
                double s0 = 0.0, s1 = 0.0, s2 = 0.0;
                for (int sl = 0; sl < kConvSlices; sl++) {
                    s0 += red[sl][t][0];
                    s1 += red[sl][t][1];
                    s2 += red[sl][t][2];
                }
                const int64_t k = kidx[kk];
                u[3 * (col + (int64_t)npts * k) + 0] = s0;
                u[3 * (col + (int64_t)npts * k) + 1] = s1;
                u[3 * (col + (int64_t)npts * k) + 2] = s2;
            }
            __syncthreads();
        }
    }
}

// Per point of the chunk: total window size (number of M items).
__global__ void __launch_bounds__(256) k_pp_counts(const int32_t* __restrict__ ms, const int32_t* __restrict__ me,
                                                   int N, int64_t* __restrict__ cnt) {
    __shared__ long long part[256];
    const int r = blockIdx.x;
    long long sum = 0;
    for (int j = threadIdx.x; j < N; j += blockDim.x) sum += max(0, me[(int64_t)r * N + j] - ms[(int64_t)r * N + j] + 1);
    part[threadIdx.x] = sum;
    __syncthreads();
    for (int d = 128; d >= 1; d >>= 1) {
        if (threadIdx.x < d) part[threadIdx.x] += part[threadIdx.x + d];
        __syncthreads();
    }
    if (threadIdx.x == 0) cnt[r] = part[0];
}

// Keys of the post-processing items, pass 1 of the counting sort: block b
// takes the (point, triangle) pairs [b per, (b + 1) per) -- items
// [off[b per], off[(b + 1) per]) -- and a thread walks one pair's window of m.
// Items of a kP range get kKpKey (its head) or kKpRestKey, items of a
// between-front range their head key or kBfRestKey; an item of k_pp_M is one
// point at one level, keyed by that point's front state (node_task_key, the
// key of the assembly's node tasks, with the point's projection computed once
// per pair), so that a wavefront's lanes take the same region branches;
// points no front reaches (nothing to integrate) get key kNodeKeys.
__global__ void __launch_bounds__(256) k_pp_keys(const FieldTri* __restrict__ tri, const double* __restrict__ xs,
                                                 int p0, PPParams P, int64_t npairs, int64_t per,
                                                 const int64_t* __restrict__ off, const int32_t* __restrict__ ms,
                                                 const int32_t* __restrict__ me, const int32_t* __restrict__ ka,
                                                 const int32_t* __restrict__ kb, const int32_t* __restrict__ sa,
                                                 const int32_t* __restrict__ sb, uint8_t* __restrict__ key,
                                                 unsigned int* __restrict__ bh) {
    __shared__ unsigned int sh[kNumKeys];
    for (int k = threadIdx.x; k < kNumKeys; k += blockDim.x) sh[k] = 0;
    __syncthreads();
    const int64_t qa = blockIdx.x * per;
    const int64_t qe = qa + per < npairs ? qa + per : npairs;
    for (int64_t q = qa + threadIdx.x; q < qe; q += blockDim.x) {
        const int a = ms[q], b = me[q];
        if (a > b) continue;
        const int p = p0 + (int)(q / P.N), j = (int)(q % P.N);
        const int64_t it0 = off[q];
        const int kpa = ka ? ka[q] : 1, kpb = ka ? kb[q] : 0;
        const int bfa = sa ? sa[q] : 1, bfb = sa ? sb[q] : 0;
        bool have = false;
        NodeGeo G;
        for (int m = a; m <= b; m++) {
            int k;
            if (m >= kpa && m <= kpb) {
                k = (m == kpa) ? kKpKey : kKpRestKey;
            } else if (m >= bfa && m <= bfb) {
                k = (m == bfa) ? bf_head_key<kPpBfChunk>(bfa, bfb) : kBfRestKey;
            } else {
                const double tau = pp_tau(P, m);
                k = kNodeKeys;
                if (tau > 0.0) {
                    if (!have) {
                        const double x[3] = {xs[3 * p], xs[3 * p + 1], xs[3 * p + 2]};
                        G = node_geo(x, tri[j]);
                        have = true;
                    }
                    const int nk = node_task_key(G, make_level(tau, P.cP, P.cS));
                    if (nk >= 0) k = nk;
                }
            }
            key[it0 + (m - a)] = (uint8_t)k;
            atomicAdd(&sh[k], 1u);
        }
    }
    __syncthreads();
    for (int k = threadIdx.x; k < kNumKeys; k += blockDim.x) bh[(int64_t)k * gridDim.x + blockIdx.x] = sh[k];
}

// Pass 2 for k_pp_keys' partition: block b places the items of its pairs
// [b per, (b + 1) per), starting from the exclusive offsets bo[key*G + b].
__global__ void __launch_bounds__(256) k_pp_scatter(int64_t npairs, int64_t per, const int64_t* __restrict__ off,
                                                    int64_t nitems, const uint8_t* __restrict__ key,
                                                    const unsigned long long* __restrict__ bo,
                                                    int32_t* __restrict__ perm) {
    __shared__ unsigned long long cur[kNumKeys];
    for (int k = threadIdx.x; k < kNumKeys; k += blockDim.x) cur[k] = bo[(int64_t)k * gridDim.x + blockIdx.x];
    __syncthreads();
    const int64_t qa = blockIdx.x * per;
    if (qa >= npairs) return;
    const int64_t qe = qa + per < npairs ? qa + per : npairs;
    const int64_t a = off[qa];
    const int64_t e = qe < npairs ? off[qe] : nitems;
    for (int64_t it = a + threadIdx.x; it < e; it += blockDim.x) {
        unsigned long long pos = atomicAdd(&cur[key[it]], 1ull);
        perm[pos] = (int32_t)it;
    }
}

// aT[(j*(nT+1) + 1 + n)*3 + c] = alpha[n*3N + 3j + c] and dT at the same
// place alpha[n*3N + 3j + c] - alpha[(n-1)*3N + 3j + c] (0 for n = 0); the
// leading zero step of each row is cleared by the caller.
__global__ void k_alpha_T(const double* __restrict__ alpha, int N, int nT, double* __restrict__ aT,
                          double* __restrict__ dT) {
    int64_t id = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    int64_t tot = (int64_t)N * nT * 3;
    if (id >= tot) return;
    int c = (int)(id % 3);
    int64_t r = id / 3;
    int n = (int)(r % nT), j = (int)(r / nT);
    const double v = alpha[(int64_t)n * 3 * N + 3 * j + c];
    const int64_t o = ((int64_t)j * (nT + 1) + 1 + n) * 3 + c;
    aT[o] = v;
    dT[o] = v - (n > 0 ? alpha[(int64_t)(n - 1) * 3 * N + 3 * j + c] : 0.0);
}

// ------------------------------------------------------------------ RHS --

// 1-D bar solution of height h (calcNucleoTn.m:148-164, calcBoundDataDirichlet.m:6-13):
// sum_{k=0}^{khat} (-1)^k [(cP t - 2hk - (h - z))_+ - (cP t - 2h(k+1) + (h - z))_+],
// khat = ceil(cP Tfin / (2h)) - 1.
__host__ __device__ inline double bar_solution(double h, double cP, double Tfin, double z, double t) {
#pragma clang fp contract(off)
    const int khat = (int)ceil((cP * Tfin) / (2 * h)) - 1;
    double s = 0.0;
    for (int k = 0; k <= khat; k++) {
        const double a = cP * t - 2 * h * k - (h - z);
        const double b = cP * t - 2 * h * (k + 1) + (h - z);
        const double sg = (k & 1) ? -1.0 : 1.0;
        s = s + sg * (a * (a > 0 ? 1.0 : 0.0) - b * (b > 0 ? 1.0 : 0.0));
    }
    return s;
}

// Boundary datum of the right-hand side, g(x, t) of calcNucleoTn.m ('DIR'
// branch for kinds 2..7).
__device__ inline void datum(int kind, const double* x, const double* nrm, double t, double cP, double Tfin,
                             double* g) {
    g[0] = g[1] = g[2] = 0.0;
    if (kind == 0) {
        double s = sin(kPi * t / 2.0);
        g[2] = t > 0.0 ? s * s : 0.0;
    } else if (kind == 1) {
        double f = t > 0.0 ? t * t * exp(-4.0 * t) : 0.0;
        g[0] = nrm[0] * f;
        g[1] = nrm[1] * f;
        g[2] = nrm[2] * f;
    } else if (kind == 2) {
        g[2] = 0.5 * t;
    } else if (kind == 3) {
        double s = sin(t);
        double a = (t < kPi / 2.0) ? s * s * s * s * s : 1.0;
        g[2] = a * (x[0] * x[0]);
    } else if (kind == 4 || kind == 5) {  // barH1, barH3
        g[2] = bar_solution(kind == 4 ? 1.0 : 3.0, cP, Tfin, x[2], t);
    } else if (kind == 6) {  // sphereWave: calcNucleoTn.m:166-169 (minus sign)
        const double d = x[0] - 2 + cP * t - 0.475;
        g[0] = -exp(-20 * (d * d));
    } else if (kind == 7) {  // elementoIndustriale: calcNucleoTn.m:171-174
        const double d = x[2] + 1 - cP * t;
        g[2] = -exp(-20 * (d * d));
    }
}

// Dirichlet datum at a vertex (BEMenerg_dir_calcBoundDataDirichlet.m:5-25):
// the bars as above, sphereWave with a plus sign, elementoIndustriale zero.
__host__ __device__ inline void dir_datum(int kind, const double* x, double t, double cP, double Tfin, double* g) {
#pragma clang fp contract(off)
    g[0] = g[1] = g[2] = 0.0;
    if (kind == 4 || kind == 5) {
        g[2] = bar_solution(kind == 4 ? 1.0 : 3.0, cP, Tfin, x[2], t);
    } else if (kind == 6) {
        const double d = x[0] - 2 + cP * t - 0.475;
        g[0] = exp(-20 * (d * d));
    }
}

// beta (3N x nT col-major): Tn/BEMenerg_core_calcTnBlock.m:27-58.
__global__ void k_rhs(const FieldTri* __restrict__ tri, const double* __restrict__ gh, int N, int nT, int nGH,
                      double dt, double cP, double Tfin, int kind, double* __restrict__ beta) {
    int64_t id = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (id >= (int64_t)N * nT) return;
    int T = (int)(id % N), n = (int)(id / N);
    double t0 = dt * n, t1 = dt * (n + 1);
    const double* g = gh + (int64_t)T * 4 * 32;
    double ris[3] = {0.0, 0.0, 0.0};
    for (int q = 0; q < nGH; q++) {
        double x[3] = {g[q], g[32 + q], g[64 + q]};
        double a[3], b[3];
        datum(kind, x, tri[T].n, t1, cP, Tfin, a);
        datum(kind, x, tri[T].n, t0, cP, Tfin, b);
        for (int k = 0; k < 3; k++) ris[k] = ris[k] + (a[k] - b[k]) * g[96 + q];
    }
    for (int k = 0; k < 3; k++) beta[(int64_t)n * 3 * N + 3 * T + k] = fabs(ris[k]) < 1.0e-14 ? 0.0 : ris[k];
}

}  // namespace ebem
