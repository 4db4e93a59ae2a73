// ebem_k.hpp -- device kernels of the double-layer operator K (math in
// ebem_k_math.hpp): pair windows, slot map, blocks, Dirichlet datum, beta_K.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "ebem_k_math.hpp"
#include "ebem_kernels.hpp"

namespace ebem {

// Windows of the (S, s) pairs of row S = blockIdx.x; rowcnt[S] = slots of the
// row, rowcnt[N + S] = its pairs with a nonempty window.
__global__ void __launch_bounds__(256) k_kwin(const double* __restrict__ verts, const double* __restrict__ X,
                                              const int32_t* __restrict__ T, const FieldTri* __restrict__ tri,
                                              KParams K, int16_t* __restrict__ lo, int16_t* __restrict__ hi,
                                              int64_t* __restrict__ rowcnt) {
    __shared__ long long part[256], partp[256];
    const int S = blockIdx.x;
    const double* vS = verts + (int64_t)S * 9;
    const int cut = k_sing_cut(tri[S].maxL, K);
    const int t0 = T[3 * S], t1 = T[3 * S + 1], t2 = T[3 * S + 2];
    long long cnt = 0, npr = 0;
    for (int s = threadIdx.x; s < K.Nv; s += blockDim.x) {
        const double* P = X + 3 * (int64_t)s;
        const bool own = (s == t0) || (s == t1) || (s == t2);
        int l0 = -1, l1 = -1;
        for (int lag = 0; lag < K.LK; lag++) {
            const bool nz = !k_block_null(vS, P, lag, K) || (own && lag < cut);
            if (nz) {
                if (l0 < 0) l0 = lag;
                l1 = lag + 1;
            }
        }
        const int64_t p = (int64_t)S * K.Nv + s;
        lo[p] = (int16_t)(l0 < 0 ? 0 : l0);
        hi[p] = (int16_t)(l0 < 0 ? 0 : l1);
        cnt += l0 < 0 ? 0 : l1 - l0;
        npr += l0 < 0 ? 0 : 1;
    }
    part[threadIdx.x] = cnt;
    partp[threadIdx.x] = npr;
    __syncthreads();
    for (int d = 128; d > 0; d >>= 1) {
        if ((int)threadIdx.x < d) {
            part[threadIdx.x] += part[threadIdx.x + d];
            partp[threadIdx.x] += partp[threadIdx.x + d];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        rowcnt[S] = part[0];
        rowcnt[K.N + S] = partp[0];
    }
}

// Slot offsets of the pairs of row S = blockIdx.x and the slot -> pair map.
__global__ void __launch_bounds__(256) k_kslots(const int16_t* __restrict__ lo, const int16_t* __restrict__ hi,
                                                int Nv, const int64_t* __restrict__ rowbase,
                                                int64_t* __restrict__ off, int32_t* __restrict__ slot_pair) {
    __shared__ long long part[256];
    const int S = blockIdx.x;
    const int nper = (Nv + blockDim.x - 1) / blockDim.x;
    const int j0 = threadIdx.x * nper;
    const int64_t pbase = (int64_t)S * Nv;
    long long s = 0;
    for (int k = 0; k < nper; k++) {
        const int j = j0 + k;
        if (j < Nv) s += hi[pbase + j] - lo[pbase + j];
    }
    part[threadIdx.x] = s;
    __syncthreads();
    for (int d = 1; d < 256; d <<= 1) {
        long long v = ((int)threadIdx.x >= d) ? part[threadIdx.x - d] : 0;
        __syncthreads();
        part[threadIdx.x] += v;
        __syncthreads();
    }
    long long run = rowbase[S] + (threadIdx.x ? part[threadIdx.x - 1] : 0);
    for (int k = 0; k < nper; k++) {
        const int j = j0 + k;
        if (j >= Nv) break;
        const int64_t p = pbase + j;
        const int n = hi[p] - lo[p];
        off[p] = run;
        for (int q = 0; q < n; q++) slot_pair[run + q] = (int32_t)p;
        run += n;
    }
}

// One wave per chunk of up to kKChunk consecutive slots of a pair (the wave
// of a slot whose offset in its pair's window is a multiple of kKChunk; the
// others pass): the blocks (S, s) of K_lag for those lags.  Regular part
// (k_reg_lane_chunk over the chunk's lags that isBlockNull keeps, an
// interval), each lag's lanes summed by an xor butterfly, cleaned below
// 1e-14 and divided by 4 pi dt (kernelK.cu:212-221); singular part per lag,
// its levels combined with -1, 3, -3, 1 and divided by 4 pi dt
// (calcSubBlockSingK.m:17-39), added (timeMarchingD.m:244-255).
__global__ void __launch_bounds__(256) k_kblocks(const KTri* __restrict__ kt, const FieldTri* __restrict__ tri,
                                                 const double* __restrict__ verts, const double* __restrict__ gh,
                                                 const double* __restrict__ X, const int32_t* __restrict__ T,
                                                 const int32_t* __restrict__ vadj_off,
                                                 const int32_t* __restrict__ vadj, const double* __restrict__ rin_g,
                                                 const double* __restrict__ rdg_g, int nGH, KParams K, int64_t nslots,
                                                 const int32_t* __restrict__ slot_pair,
                                                 const int64_t* __restrict__ off, const int16_t* __restrict__ lo,
                                                 const int16_t* __restrict__ hi, double* __restrict__ val) {
#pragma clang fp contract(off)
    extern __shared__ double rule[];  // (nInt + nDiag) x 4
    for (int k = threadIdx.x; k < 4 * (K.nInt + K.nDiag); k += blockDim.x)
        rule[k] = k < 4 * K.nInt ? rin_g[k] : rdg_g[k - 4 * K.nInt];
    __syncthreads();
    KBlockArgs A{kt, verts, vadj_off, vadj, rule, rule + 4 * K.nInt, nullptr, nGH};
    const int lane = threadIdx.x & 63;
    const int64_t wave = (blockIdx.x * (int64_t)blockDim.x + threadIdx.x) >> 6;
    const int64_t nwaves = ((int64_t)gridDim.x * blockDim.x) >> 6;
    const int ct[4] = {-2, -1, 0, 1};
    const double cn[4] = {-1, 3, -3, 1};
    for (int64_t it = wave; it < nslots; it += nwaves) {
        const int32_t p = slot_pair[it];
        const int rel = (int)(it - off[p]);
        if (rel % kKChunk != 0) continue;  // wave-uniform
        const int S = p / K.Nv, s = p % K.Nv;
        const int lag0 = lo[p] + rel;
        const int nc = min(kKChunk, (int)hi[p] - lag0);
        A.g = gh + (int64_t)S * 128;
        const double* vS = verts + (int64_t)S * 9;
        const double* Ps = X + 3 * (int64_t)s;
        // regular parts first, written by lane 0 (zero for the lags the skip test drops)
        int ra = -1, rb = -1;  // the chunk's lags that isBlockNull keeps (an interval)
        for (int j = 0; j < nc; j++)
            if (!k_block_null(vS, Ps, lag0 + j, K)) {
                if (ra < 0) ra = j;
                rb = j + 1;
            }
        if (lane == 0)
            for (int j = 0; j < nc; j++)
                if (j < ra || j >= rb)
                    for (int e = 0; e < 9; e++) val[(it + j) * 9 + e] = 0.0;
        if (ra >= 0) {  // wave-uniform
            double acc[kKChunk][3][3];
            k_reg_lane_chunk(A, K, S, s, lag0 + ra, rb - ra, lane, 64, acc);
#pragma unroll
            for (int j = 0; j < kKChunk; j++) {
                if (j >= rb - ra) break;
#pragma unroll
                for (int i = 0; i < 3; i++)
#pragma unroll
                    for (int c = 0; c < 3; c++) {
                        double a = acc[j][i][c];
                        for (int m = 32; m >= 1; m >>= 1) a += __shfl_xor(a, m);
                        if (lane == 0) val[(it + ra + j) * 9 + i + 3 * c] = fabs(a) > 1.0e-14 ? a / K.fourpidt : 0.0;
                    }
            }
        }
        const int vl = k_local_vertex(T, S, s);
        if (vl >= 0) {
            const int cut = k_sing_cut(tri[S].maxL, K);
            for (int j = 0; j < nc; j++) {
                const int lag = lag0 + j;
                if (lag >= cut) break;  // wave-uniform
                double blk[3][3];
                for (int i = 0; i < 3; i++)
                    for (int c = 0; c < 3; c++) blk[i][c] = 0.0;
                for (int l = 0; l < 4; l++) {
                    const int il = lag + ct[l];
                    if (il <= 0) continue;
                    double lev[3][3];
                    k_sing_lane(A, K, S, vl, (double)il * K.dt, lane, 64, lev);
#pragma unroll
                    for (int i = 0; i < 3; i++)
#pragma unroll
                        for (int c = 0; c < 3; c++) {
                            double a = lev[i][c];
                            for (int m = 32; m >= 1; m >>= 1) a += __shfl_xor(a, m);
                            blk[i][c] += cn[l] * a;
                        }
                }
                if (lane == 0)
                    for (int i = 0; i < 3; i++)
                        for (int c = 0; c < 3; c++) val[(it + j) * 9 + i + 3 * c] += blk[i][c] / K.fourpidt;
            }
        }
    }
}

// ---- regular part by field triangle: one kernel evaluation for the three
// vertices of F, then a gather per (S, s, lag) over the triangles of s ----

// Rows r0 + blockIdx.x: the lag window of every (S, F != S) pair, the union of
// the lags isBlockNull keeps for F's three vertices; rowcnt[r] = partial slots
// of the row, rowcnt[rows + r] = its chunks of kKsfChunk lags.
__global__ void __launch_bounds__(256) k_ksf_win(const double* __restrict__ verts, const double* __restrict__ X,
                                                 const int32_t* __restrict__ T, KParams K, int r0, int rows,
                                                 int16_t* __restrict__ sflo, int16_t* __restrict__ sfhi,
                                                 int64_t* __restrict__ rowcnt) {
    __shared__ long long part[256], partc[256];
    const int r = blockIdx.x, S = r0 + r;
    const double* vS = verts + (int64_t)S * 9;
    long long cnt = 0, chk = 0;
    for (int F = threadIdx.x; F < K.N; F += blockDim.x) {
        int a = K.LK, b = 0;
        if (F != S)
            for (int v = 0; v < 3; v++) {
                const double* P = X + 3 * (int64_t)T[3 * F + v];
                for (int lag = 0; lag < K.LK; lag++)
                    if (!k_block_null(vS, P, lag, K)) {
                        a = min(a, lag);
                        b = max(b, lag + 1);
                    }
            }
        if (b <= a) a = b = 0;
        const int64_t q = (int64_t)r * K.N + F;
        sflo[q] = (int16_t)a;
        sfhi[q] = (int16_t)b;
        const int nlev = k_sf_nlev(a, b);
        cnt += nlev;
        chk += (nlev + kKsfChunk - 1) / kKsfChunk;
    }
    part[threadIdx.x] = cnt;
    partc[threadIdx.x] = chk;
    __syncthreads();
    for (int d = 128; d > 0; d >>= 1) {
        if ((int)threadIdx.x < d) {
            part[threadIdx.x] += part[threadIdx.x + d];
            partc[threadIdx.x] += partc[threadIdx.x + d];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        rowcnt[r] = part[0];
        rowcnt[rows + r] = partc[0];
    }
}

// Per row: partial-slot offsets pOff of the (S, F) pairs and the chunk items
// (pair index, chunk number); rowbase[r], rowbase[rows + r] the row's bases.
__global__ void __launch_bounds__(256) k_ksf_items(const int16_t* __restrict__ sflo, const int16_t* __restrict__ sfhi,
                                                   int N, int rows, const int64_t* __restrict__ rowbase,
                                                   int64_t* __restrict__ pOff, int32_t* __restrict__ itemP,
                                                   int16_t* __restrict__ itemC) {
    __shared__ long long part[256], partc[256];
    const int r = blockIdx.x;
    const int nper = (N + blockDim.x - 1) / blockDim.x;
    const int j0 = threadIdx.x * nper;
    const int64_t pbase = (int64_t)r * N;
    long long s = 0, c = 0;
    for (int k = 0; k < nper; k++) {
        const int j = j0 + k;
        if (j < N) {
            const int w = k_sf_nlev(sflo[pbase + j], sfhi[pbase + j]);
            s += w;
            c += (w + kKsfChunk - 1) / kKsfChunk;
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
    long long run = rowbase[r] + (threadIdx.x ? part[threadIdx.x - 1] : 0);
    long long runc = rowbase[rows + r] + (threadIdx.x ? partc[threadIdx.x - 1] : 0);
    for (int k = 0; k < nper; k++) {
        const int j = j0 + k;
        if (j >= N) break;
        const int64_t q = pbase + j;
        const int w = k_sf_nlev(sflo[q], sfhi[q]);
        pOff[q] = run;
        const int nc = (w + kKsfChunk - 1) / kKsfChunk;
        for (int u = 0; u < nc; u++) {
            itemP[runc + u] = (int32_t)q;
            itemC[runc + u] = (int16_t)u;
        }
        run += w;
        runc += nc;
    }
}

// One wave per chunk item: the level sums of pair (S, F) for F's three
// vertices and the item's levels, butterflied:
// part[(pOff + m - mlo) * 27 + 9 v + (i + 3 c)], mlo = k_sf_mlo(sflo).
// waves per SIMD that k_ksf is compiled for (registers per wave <= 512 / this)
#ifndef EBEM_KSF_WAVES
#define EBEM_KSF_WAVES 2
#endif
__global__ void __launch_bounds__(256, EBEM_KSF_WAVES) k_ksf(const KTri* __restrict__ kt, const double* __restrict__ verts,
                                             const double* __restrict__ gh, const double* __restrict__ rin_g, int nGH,
                                             KParams K, int r0, int64_t nitems, const int32_t* __restrict__ itemP,
                                             const int16_t* __restrict__ itemC, const int16_t* __restrict__ sflo,
                                             const int16_t* __restrict__ sfhi, const int64_t* __restrict__ pOff,
                                             double* __restrict__ part) {
#pragma clang fp contract(off)
    extern __shared__ double rule[];  // nInt x 4
    for (int k = threadIdx.x; k < 4 * K.nInt; k += blockDim.x) rule[k] = rin_g[k];
    __syncthreads();
    KBlockArgs A{kt, verts, nullptr, nullptr, rule, rule, nullptr, nGH};
    const int lane = threadIdx.x & 63;
    const int64_t wave = (blockIdx.x * (int64_t)blockDim.x + threadIdx.x) >> 6;
    const int64_t nwaves = ((int64_t)gridDim.x * blockDim.x) >> 6;
    for (int64_t it = wave; it < nitems; it += nwaves) {
        const int32_t q = itemP[it];
        const int S = r0 + q / K.N, F = q % K.N;
        const int mlo = k_sf_mlo(sflo[q]);
        const int mA = mlo + kKsfChunk * itemC[it];
        const int nL = min(kKsfChunk, (int)sfhi[q] + 1 - mA);
        A.g = gh + (int64_t)S * 128;
        double acc[3][kKsfChunk][9];
        k_sf_lane(A, K, F, mA, nL, lane, 64, acc);
        double* o = part + (pOff[q] + (mA - mlo)) * 27;
#pragma unroll
        for (int j = 0; j < kKsfChunk; j++) {
            if (j >= nL) break;
#pragma unroll
            for (int v = 0; v < 3; v++)
#pragma unroll
                for (int e = 0; e < 9; e++) {
                    double a = acc[v][j][e];
                    for (int m = 32; m >= 1; m >>= 1) a += __shfl_xor(a, m);
                    if (lane == 0) o[j * 27 + v * 9 + e] = a;
                }
        }
    }
}

// One thread per store slot of rows [r0, r0 + rows): the regular block
// (S, s, lag) = sum over the triangles F != S of s (ascending) of
// -W_{lag-2} + 3 W_{lag-1} - 3 W_lag + W_{lag+1} (levels m <= 0 left out) of
// their level sums, zero where isBlockNull drops it, cleaned below 1e-14 and
// divided by 4 pi dt (kernelK.cu:212-221).
__global__ void k_kgather(const double* __restrict__ verts, const double* __restrict__ X, KParams K, int r0,
                          int64_t s0, int64_t ns, const int32_t* __restrict__ slot_pair,
                          const int64_t* __restrict__ off, const int16_t* __restrict__ lo,
                          const int32_t* __restrict__ vadj_off, const int32_t* __restrict__ vadj,
                          const int16_t* __restrict__ sflo, const int16_t* __restrict__ sfhi,
                          const int64_t* __restrict__ pOff, const double* __restrict__ part, double* __restrict__ val) {
#pragma clang fp contract(off)
    const int64_t id = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (id >= ns) return;
    const int64_t slot = s0 + id;
    const int32_t p = slot_pair[slot];
    const int S = p / K.Nv, s = p % K.Nv;
    const int lag = lo[p] + (int)(slot - off[p]);
    double o[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
    if (!k_block_null(verts + (int64_t)S * 9, X + 3 * (int64_t)s, lag, K)) {
        double a[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
        for (int e = vadj_off[s]; e < vadj_off[s + 1]; e++) {
            const int F = vadj[e] >> 2, v = vadj[e] & 3;
            if (F == S) continue;
            const int64_t q = (int64_t)(S - r0) * K.N + F;
            if (lag < sflo[q] || lag >= sfhi[q]) continue;
            const int mlo = k_sf_mlo(sflo[q]);
            const double cn[4] = {-1, 3, -3, 1};
            double b[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
            for (int d = 0; d < 4; d++) {
                const int m = lag - 2 + d;
                if (m < 1) continue;
                const double* pp = part + (pOff[q] + (m - mlo)) * 27 + v * 9;
                for (int c = 0; c < 9; c++) b[c] += cn[d] * pp[c];
            }
            for (int c = 0; c < 9; c++) a[c] += b[c];
        }
        for (int c = 0; c < 9; c++) o[c] = fabs(a[c]) > 1.0e-14 ? a[c] / K.fourpidt : 0.0;
    }
    double* out = val + slot * 9;
    for (int c = 0; c < 9; c++) out[c] = o[c];
}

// One wave per (test triangle S, local vertex v): the singular part of the
// blocks (S, T[S][v], lag) for the lags of the pair's window below the
// cut-off, per lag and level (calcSubBlockSingK.m:17-39), added to the store.
__global__ void __launch_bounds__(256) k_ksing(const KTri* __restrict__ kt, const FieldTri* __restrict__ tri,
                                               const double* __restrict__ verts, const double* __restrict__ gh,
                                               const int32_t* __restrict__ T, const double* __restrict__ rdg_g, int nGH,
                                               KParams K, const int64_t* __restrict__ off,
                                               const int16_t* __restrict__ lo, const int16_t* __restrict__ hi,
                                               double* __restrict__ val) {
#pragma clang fp contract(off)
    extern __shared__ double rule[];  // nDiag x 4
    for (int k = threadIdx.x; k < 4 * K.nDiag; k += blockDim.x) rule[k] = rdg_g[k];
    __syncthreads();
    KBlockArgs A{kt, verts, nullptr, nullptr, rule, rule, nullptr, nGH};
    const int lane = threadIdx.x & 63;
    const int64_t wave = (blockIdx.x * (int64_t)blockDim.x + threadIdx.x) >> 6;
    const int64_t nwaves = ((int64_t)gridDim.x * blockDim.x) >> 6;
    const int ct[4] = {-2, -1, 0, 1};
    const double cn[4] = {-1, 3, -3, 1};
    for (int64_t it = wave; it < 3 * (int64_t)K.N; it += nwaves) {
        const int S = (int)(it / 3), vl = (int)(it % 3);
        const int s = T[3 * S + vl];
        const int64_t p = (int64_t)S * K.Nv + s;
        const int cut = min(k_sing_cut(tri[S].maxL, K), (int)hi[p]);
        A.g = gh + (int64_t)S * 128;
        for (int lag = lo[p]; lag < cut; lag++) {
            double blk[3][3];
            for (int i = 0; i < 3; i++)
                for (int c = 0; c < 3; c++) blk[i][c] = 0.0;
            for (int l = 0; l < 4; l++) {
                const int il = lag + ct[l];
                if (il <= 0) continue;
                double lev[3][3];
                k_sing_lane(A, K, S, vl, (double)il * K.dt, lane, 64, lev);
#pragma unroll
                for (int i = 0; i < 3; i++)
#pragma unroll
                    for (int c = 0; c < 3; c++) {
                        double a = lev[i][c];
                        for (int m = 32; m >= 1; m >>= 1) a += __shfl_xor(a, m);
                        blk[i][c] += cn[l] * a;
                    }
            }
            if (lane == 0) {
                double* o = val + (off[p] + (lag - lo[p])) * 9;
                for (int i = 0; i < 3; i++)
                    for (int c = 0; c < 3; c++) o[i + 3 * c] += blk[i][c] / K.fourpidt;
            }
        }
    }
}

// Dirichlet datum at the vertices (BEMenerg_dir_calcBoundDataDirichlet.m):
// gT[(s * nT + n) * 3 + c] = g_c(X_s, (n + 1) dt).
__global__ void k_dir_datum(const double* __restrict__ X, int Nv, int nT, double dt, double cP, double Tfin,
                            int kind, double* __restrict__ gT) {
    const int64_t id = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (id >= (int64_t)Nv * nT) return;
    const int s = (int)(id / nT), n = (int)(id % nT);
    double g[3];
    dir_datum(kind, X + 3 * (int64_t)s, (double)(n + 1) * dt, cP, Tfin, g);
    for (int c = 0; c < 3; c++) gT[id * 3 + c] = g[c];
}

// Vertex datum (3Nv x nT column-major) -> gT layout [s][n][c].
__global__ void k_dir_transpose(const double* __restrict__ g, int Nv, int nT, double* __restrict__ gT) {
    const int64_t id = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (id >= (int64_t)Nv * nT * 3) return;
    const int c = (int)(id % 3);
    const int64_t sn = id / 3;
    const int s = (int)(sn / nT), n = (int)(sn % nT);
    gT[id] = g[(int64_t)n * 3 * Nv + 3 * s + c];
}

// beta_K rows of test triangle S = blockIdx.x: out[n*3N + 3S + i] =
// sum over the row's slots (pair s, lag l <= n) of K[S,s,l] g_s((n - l + 1) dt);
// threads own time steps n = tid + 256 k (kKBetaPer per thread).
constexpr int kKBetaPer = 16;
__global__ void __launch_bounds__(256) k_kbeta(const int16_t* __restrict__ lo, const int16_t* __restrict__ hi,
                                               const int64_t* __restrict__ off, const double* __restrict__ val,
                                               const double* __restrict__ gT, int N, int Nv, int nT,
                                               double* __restrict__ out) {
    const int S = blockIdx.x;
    double acc[kKBetaPer][3];
#pragma unroll
    for (int k = 0; k < kKBetaPer; k++) acc[k][0] = acc[k][1] = acc[k][2] = 0.0;
    const int64_t pbase = (int64_t)S * Nv;
    for (int s = 0; s < Nv; s++) {
        const int l0 = lo[pbase + s], l1 = hi[pbase + s];
        if (l1 <= l0) continue;
        const double* vb = val + off[pbase + s] * 9;
        const double* gs = gT + (int64_t)s * nT * 3;
        for (int l = l0; l < l1; l++) {
            const double* Kb = vb + (int64_t)(l - l0) * 9;  // column-major 3x3
            const double k00 = Kb[0], k10 = Kb[1], k20 = Kb[2], k01 = Kb[3], k11 = Kb[4], k21 = Kb[5], k02 = Kb[6],
                         k12 = Kb[7], k22 = Kb[8];
#pragma unroll
            for (int k = 0; k < kKBetaPer; k++) {
                const int n = threadIdx.x + 256 * k;
                if (n >= nT || n < l) continue;
                const double* gg = gs + (int64_t)(n - l) * 3;
                const double g0 = gg[0], g1 = gg[1], g2 = gg[2];
                acc[k][0] += k00 * g0 + k01 * g1 + k02 * g2;
                acc[k][1] += k10 * g0 + k11 * g1 + k12 * g2;
                acc[k][2] += k20 * g0 + k21 * g1 + k22 * g2;
            }
        }
    }
#pragma unroll
    for (int k = 0; k < kKBetaPer; k++) {
        const int n = threadIdx.x + 256 * k;
        if (n >= nT) continue;
        double* o = out + (int64_t)n * 3 * N + 3 * S;
        o[0] = acc[k][0];
        o[1] = acc[k][1];
        o[2] = acc[k][2];
    }
}

// beta = beta / 2 + betaK (timeMarchingD.m:287).
__global__ void k_dd_combine(double* __restrict__ beta, const double* __restrict__ betak, int64_t n) {
    const int64_t id = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (id < n) beta[id] = beta[id] / 2 + betak[id];
}

}  // namespace ebem
