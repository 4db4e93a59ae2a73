// Test-only host build of the device K math (ebem_k.hpp: k_reg_lane and
// k_sing_lane are __host__ __device__): one block (S, s) of K_lag computed by
// the same code as the device kernel, the 64 lanes' parts summed by the same
// xor butterfly, so the CPU test-suite can check it against the oracle.
#include <cmath>
#include <cstring>
#include <vector>

#include "../../elastodynamicebem_amd/csrc/ebem_k_math.hpp"
#include "../../elastodynamicebem_amd/csrc/ebem_ppd_math.hpp"

using namespace ebem;

namespace {
void butterfly(double a[64][9]) {
    for (int m = 32; m >= 1; m >>= 1) {
        double b[64][9];
        for (int l = 0; l < 64; l++)
            for (int e = 0; e < 9; e++) b[l][e] = a[l][e] + a[l ^ m][e];
        std::memcpy(a, b, sizeof b);
    }
}
// 0: the regular part per (S, s) pair as k_kblocks (k_reg_lane); 1: by field
// triangle as k_ksf + k_kgather (k_sf_lane level sums of every F != S with
// vertex s, ascending F, combined with -1, 3, -3, 1).
int g_sf_mode = 0;
}  // namespace

extern "C" {

void k_host_set_sf_mode(int mode) { g_sf_mode = mode; }

// Block (S, s) of K_lag, column-major, for a mesh (X: Nv x 3, T: N x 3,
// normals, areas, longest edges), material, outer rule (barycentric ghn, ghw)
// and inner rules (rin: nInt x 4, rdg: nDiag x 4).
int k_block_host(int N, int Nv, const double* X, const int32_t* T, const double* nrm, const double* area,
                 const double* maxL, double rho, double mu, double lam, double Tfin, int nT, int nGH,
                 const double* ghn, const double* ghw, int nInt, const double* rin, int nDiag, const double* rdg,
                 int lag, int S, int s, double* out9) {
#pragma clang fp contract(off)
    std::vector<KTri> kt(N);
    std::vector<double> verts((size_t)N * 9);
    std::vector<int32_t> voff(Nv + 1, 0), vadj(3 * (size_t)N);
    double maxLen = 0;
    for (int F = 0; F < N; F++) {
        const double* V[3] = {X + 3 * T[3 * F], X + 3 * T[3 * F + 1], X + 3 * T[3 * F + 2]};
        k_tri_basis(V[0], V[1], V[2], nrm + 3 * F, area[F], kt[F]);
        for (int v = 0; v < 3; v++) {
            for (int c = 0; c < 3; c++) verts[9 * (size_t)F + 3 * v + c] = V[v][c];
            voff[T[3 * F + v] + 1]++;
        }
        if (maxL[F] > maxLen) maxLen = maxL[F];
    }
    for (int v = 0; v < Nv; v++) voff[v + 1] += voff[v];
    std::vector<int32_t> pos(voff.begin(), voff.end() - 1);
    for (int F = 0; F < N; F++)
        for (int v = 0; v < 3; v++) vadj[pos[T[3 * F + v]]++] = F * 4 + v;
    double g[128] = {0};
    const double* VS[3] = {X + 3 * T[3 * S], X + 3 * T[3 * S + 1], X + 3 * T[3 * S + 2]};
    for (int q = 0; q < nGH; q++) {
        for (int k = 0; k < 3; k++)
            g[32 * k + q] = ghn[3 * q] * VS[0][k] + ghn[3 * q + 1] * VS[1][k] + ghn[3 * q + 2] * VS[2][k];
        g[96 + q] = area[S] * ghw[q];
    }
    KParams K{};
    K.N = N;
    K.Nv = Nv;
    K.LK = nT;
    K.nInt = nInt;
    K.nDiag = nDiag;
    K.dt = Tfin / nT;
    K.cP = std::sqrt((lam + 2 * mu) / rho);
    K.cS = std::sqrt(mu / rho);
    K.lam = lam;
    K.mu = mu;
    K.rho = rho;
    K.maxLen = maxLen;
    K.fourpidt = 4 * kPi * K.dt;
    KBlockArgs A{kt.data(), verts.data(), voff.data(), vadj.data(), rin, rdg, g, nGH};
    double out[9] = {0};
    static double lanes[64][9];
    if (g_sf_mode == 1 && !k_block_null(verts.data() + 9 * (size_t)S, X + 3 * s, lag, K)) {
        const double cn[4] = {-1, 3, -3, 1};
        double a[9] = {0};
        for (int e = voff[s]; e < voff[s + 1]; e++) {
            const int F = vadj[e] >> 2, v = vadj[e] & 3;
            if (F == S) continue;
            double b[9] = {0};
            for (int d = 0; d < 4; d++) {
                const int m = lag - 2 + d;
                if (m < 1) continue;
                for (int l = 0; l < 64; l++) {
                    double acc[3][kKsfChunk][9];
                    k_sf_lane(A, K, F, m, 1, l, 64, acc);
                    for (int c = 0; c < 9; c++) lanes[l][c] = acc[v][0][c];
                }
                butterfly(lanes);
                for (int c = 0; c < 9; c++) b[c] += cn[d] * lanes[0][c];
            }
            for (int c = 0; c < 9; c++) a[c] += b[c];
        }
        for (int e = 0; e < 9; e++) out[e] = std::fabs(a[e]) > 1.0e-14 ? a[e] / K.fourpidt : 0.0;
    } else if (!k_block_null(verts.data() + 9 * (size_t)S, X + 3 * s, lag, K)) {
        for (int l = 0; l < 64; l++) {
            double acc[3][3];
            k_reg_lane(A, K, S, s, lag, l, 64, acc);
            for (int i = 0; i < 3; i++)
                for (int c = 0; c < 3; c++) lanes[l][i + 3 * c] = acc[i][c];
        }
        butterfly(lanes);
        for (int e = 0; e < 9; e++) out[e] = std::fabs(lanes[0][e]) > 1.0e-14 ? lanes[0][e] / K.fourpidt : 0.0;
    }
    const int vl = k_local_vertex(T, S, s);
    if (vl >= 0 && lag < k_sing_cut(maxL[S], K)) {
        const int ct[4] = {-2, -1, 0, 1};
        const double cn[4] = {-1, 3, -3, 1};
        double blk[9] = {0};
        for (int l = 0; l < 4; l++) {
            const int il = lag + ct[l];
            if (il <= 0) continue;
            for (int ln = 0; ln < 64; ln++) {
                double lev[3][3];
                k_sing_lane(A, K, S, vl, (double)il * K.dt, ln, 64, lev);
                for (int i = 0; i < 3; i++)
                    for (int c = 0; c < 3; c++) lanes[ln][i + 3 * c] = lev[i][c];
            }
            butterfly(lanes);
            for (int e = 0; e < 9; e++) blk[e] += cn[l] * lanes[0][e];
        }
        for (int e = 0; e < 9; e++) out[e] += blk[e] / K.fourpidt;
    }
    std::memcpy(out9, out, sizeof out);
    return 0;
}

// Regular parts (before clean-up) of the blocks (S, s) of lags lagA ..
// lagA+nA-1 by k_reg_lane_chunk (the kernel's chunked form) and, lag by lag,
// by k_reg_lane: out[0 .. 9 nA) and out[9 nA .. 18 nA), column-major, each
// butterflied over 64 lanes.  The two must agree bit for bit.
int k_chunk_vs_single_host(int N, int Nv, const double* X, const int32_t* T, const double* nrm, const double* area,
                           const double* maxL, double rho, double mu, double lam, double Tfin, int nT, int nGH,
                           const double* ghn, const double* ghw, int nInt, const double* rin, int lagA, int nA, int S,
                           int s, double* out) {
#pragma clang fp contract(off)
    if (nA < 1 || nA > kKChunk) return -1;
    std::vector<KTri> kt(N);
    std::vector<double> verts((size_t)N * 9);
    std::vector<int32_t> voff(Nv + 1, 0), vadj(3 * (size_t)N);
    double maxLen = 0;
    for (int F = 0; F < N; F++) {
        const double* V[3] = {X + 3 * T[3 * F], X + 3 * T[3 * F + 1], X + 3 * T[3 * F + 2]};
        k_tri_basis(V[0], V[1], V[2], nrm + 3 * F, area[F], kt[F]);
        for (int v = 0; v < 3; v++) {
            for (int c = 0; c < 3; c++) verts[9 * (size_t)F + 3 * v + c] = V[v][c];
            voff[T[3 * F + v] + 1]++;
        }
        if (maxL[F] > maxLen) maxLen = maxL[F];
    }
    for (int v = 0; v < Nv; v++) voff[v + 1] += voff[v];
    std::vector<int32_t> pos(voff.begin(), voff.end() - 1);
    for (int F = 0; F < N; F++)
        for (int v = 0; v < 3; v++) vadj[pos[T[3 * F + v]]++] = F * 4 + v;
    double g[128] = {0};
    const double* VS[3] = {X + 3 * T[3 * S], X + 3 * T[3 * S + 1], X + 3 * T[3 * S + 2]};
    for (int q = 0; q < nGH; q++) {
        for (int k = 0; k < 3; k++)
            g[32 * k + q] = ghn[3 * q] * VS[0][k] + ghn[3 * q + 1] * VS[1][k] + ghn[3 * q + 2] * VS[2][k];
        g[96 + q] = area[S] * ghw[q];
    }
    KParams K{};
    K.N = N;
    K.Nv = Nv;
    K.LK = nT;
    K.nInt = nInt;
    K.nDiag = 1;
    K.dt = Tfin / nT;
    K.cP = std::sqrt((lam + 2 * mu) / rho);
    K.cS = std::sqrt(mu / rho);
    K.lam = lam;
    K.mu = mu;
    K.rho = rho;
    K.maxLen = maxLen;
    K.fourpidt = 4 * kPi * K.dt;
    KBlockArgs A{kt.data(), verts.data(), voff.data(), vadj.data(), rin, rin, g, nGH};
    static double lanes[kKChunk][64][9];
    for (int l = 0; l < 64; l++) {
        double acc[kKChunk][3][3];
        k_reg_lane_chunk(A, K, S, s, lagA, nA, l, 64, acc);
        for (int j = 0; j < nA; j++)
            for (int i = 0; i < 3; i++)
                for (int c = 0; c < 3; c++) lanes[j][l][i + 3 * c] = acc[j][i][c];
    }
    for (int j = 0; j < nA; j++) {
        butterfly(lanes[j]);
        for (int e = 0; e < 9; e++) out[9 * j + e] = lanes[j][0][e];
    }
    for (int j = 0; j < nA; j++) {
        for (int l = 0; l < 64; l++) {
            double acc[3][3];
            k_reg_lane(A, K, S, s, lagA + j, l, 64, acc);
            for (int i = 0; i < 3; i++)
                for (int c = 0; c < 3; c++) lanes[0][l][i + 3 * c] = acc[i][c];
        }
        butterfly(lanes[0]);
        for (int e = 0; e < 9; e++) out[9 * nA + 9 * j + e] = lanes[0][0][e];
    }
    return 0;
}

// P_y block of vertex s for the point x at time t (the direct
// post-processing's kernelP part), column-major, by the device math with the
// 64-lane butterfly; rule rin: nInt x 4.
int ppp_block_host(int N, int Nv, const double* X, const int32_t* T, const double* nrm, const double* area,
                   double rho, double mu, double lam, double Tfin, int nT, int nInt, const double* rin,
                   const double* x, double t, int y, int s, double* out9) {
    // out9[0..8]: per vertex as k_ppd_P; out9[9..17]: by field triangle and level
#pragma clang fp contract(off)
    std::vector<KTri> kt(N);
    std::vector<double> verts((size_t)N * 9);
    std::vector<int32_t> voff(Nv + 1, 0), vadj(3 * (size_t)N);
    for (int F = 0; F < N; F++) {
        const double* V[3] = {X + 3 * T[3 * F], X + 3 * T[3 * F + 1], X + 3 * T[3 * F + 2]};
        k_tri_basis(V[0], V[1], V[2], nrm + 3 * F, area[F], kt[F]);
        for (int v = 0; v < 3; v++) {
            for (int c = 0; c < 3; c++) verts[9 * (size_t)F + 3 * v + c] = V[v][c];
            voff[T[3 * F + v] + 1]++;
        }
    }
    for (int v = 0; v < Nv; v++) voff[v + 1] += voff[v];
    std::vector<int32_t> pos(voff.begin(), voff.end() - 1);
    for (int F = 0; F < N; F++)
        for (int v = 0; v < 3; v++) vadj[pos[T[3 * F + v]]++] = F * 4 + v;
    const double dt = Tfin / nT;
    PPDArgs A{kt.data(), verts.data(), voff.data(), vadj.data(), rin, nInt, lam, mu,
              std::sqrt((lam + 2 * mu) / rho), std::sqrt(mu / rho), dt};
    static double lanes[64][9];
    for (int l = 0; l < 64; l++) {
        double acc[3][3];
        ppd_p_lane(A, x, t, y, s, l, 64, acc);
        // the kernel's chunked form must give the same sum bit for bit
        double accc[kPpdChunk][3][3];
        const int j = y % kPpdChunk;
        ppd_p_lane_chunk(A, x, t, y - j, j + 1, s, l, 64, accc);
        for (int i = 0; i < 3; i++)
            for (int c = 0; c < 3; c++)
                if (accc[j][i][c] != acc[i][c]) return 1;
        for (int i = 0; i < 3; i++)
            for (int jj = 0; jj < 3; jj++) lanes[l][i + 3 * jj] = acc[i][jj];
    }
    butterfly(lanes);
    const double scale = 4 * kPi * rho * dt;
    for (int e = 0; e < 9; e++) out9[e] = std::fabs(lanes[0][e]) > 1.0e-14 ? lanes[0][e] / scale : 0.0;
    // the field-triangle form (k_ppf + k_ppf_gather): level sums W_l of every
    // triangle F with vertex s, combined W_y - 2 W_{y+1} + W_{y+2}
    const double cn[3] = {1, -2, 1};
    double B[9] = {0};
    for (int e = voff[s]; e < voff[s + 1]; e++) {
        const int F = vadj[e] >> 2, v = vadj[e] & 3;
        double b[9] = {0};
        for (int k = 0; k < 3; k++) {
            // kPpfLanes lanes per item; the zero lanes above them add exact zeros
            // in the first butterfly steps, then the steps match the device's
            for (int l = 0; l < 64; l++) {
                for (int c = 0; c < 9; c++) lanes[l][c] = 0.0;
                if (l >= kPpfLanes) continue;
                double acc[3][kPpfChunk][9];
                ppf_lane(A, x, t, F, y + k, 1, l, kPpfLanes, acc);
                for (int c = 0; c < 9; c++) lanes[l][c] = acc[v][0][c];
            }
            butterfly(lanes);
            for (int c = 0; c < 9; c++) b[c] += cn[k] * lanes[0][c];
        }
        for (int c = 0; c < 9; c++) B[c] += b[c];
    }
    for (int e = 0; e < 9; e++) out9[9 + e] = std::fabs(B[e]) > 1.0e-14 ? B[e] / scale : 0.0;
    return 0;
}

}  // extern "C"
