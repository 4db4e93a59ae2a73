// ebem_lib.hip -- host side of the C ABI declared in include/ebem.h.
//
// One handle owns one device, one HIP stream (default priority) and the
// device-resident problem: triangle records, Gauss-Hammer nodes, the block
// store of E_0..E_{L-1}, the right-hand side and the density.
#include <hip/hip_runtime.h>
#include <rocblas/rocblas.h>
#include <rocsolver/rocsolver.h>

#include <algorithm>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <climits>
#include <cstdint>
#include <cstdlib>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/ebem.h"
#include "ebem_kernels.hpp"

using namespace ebem;

namespace {

thread_local std::string g_err;

int fail(int code, const char* fmt, ...) {
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    g_err = buf;
    return code;
}

#define HIPC(call)                                                                              \
    do {                                                                                        \
        hipError_t e_ = (call);                                                                 \
        if (e_ != hipSuccess)                                                                   \
            return fail(e_ == hipErrorOutOfMemory ? EBEM_ENOMEM : EBEM_EHIP, "%s failed: %s (%s:%d)", \
                        #call, hipGetErrorString(e_), __FILE__, __LINE__);                     \
    } while (0)

template <class T>
struct DevBuf {
    T* p = nullptr;
    size_t n = 0;
    ~DevBuf() { reset(); }
    void reset() {
        if (p) (void)hipFree(p);
        p = nullptr;
        n = 0;
    }
    hipError_t alloc(size_t count) {
        if (count == n && p) return hipSuccess;
        reset();
        if (count == 0) return hipSuccess;
        hipError_t e = hipMalloc((void**)&p, count * sizeof(T));
        if (e == hipSuccess) n = count;
        else p = nullptr;
        return e;
    }
};

// Gauss-Hammer tables of util/quadRules/GaussHammer_base.m:19-255 mapped by
// GaussHammerComposite(1, n) (util/quadRules/GaussHammerComposite.m:31-52):
// barycentric columns (b1,b3,b2), weights doubled.
int gauss_hammer(int n, double* nodes, double* w) {
    double g[28][3] = {};
    double gw[28] = {};
    if (n == 3) {
        const double v1 = 2.0 / 3, v2 = 1.0 / 6, v3 = 1.0 / 6;
        const double t[3][3] = {{v1, v2, v3}, {v3, v1, v2}, {v2, v3, v1}};
        for (int i = 0; i < 3; i++) {
            for (int c = 0; c < 3; c++) g[i][c] = t[i][c];
            gw[i] = 1.0 / 6;
        }
    } else if (n == 7) {
        const double a1[7] = {0.1012865073235, 0.7974269853531, 0.1012865073235, 0.4701420641051,
                              0.4701420641051, 0.0597158717898, 0.3333333333333};
        const double a2[7] = {0.1012865073235, 0.1012865073235, 0.7974269853531, 0.0597158717898,
                              0.4701420641051, 0.4701420641051, 0.3333333333333};
        for (int i = 0; i < 7; i++) {
            g[i][0] = a1[i];
            g[i][1] = a2[i];
            g[i][2] = 1 - g[i][0] - g[i][1];
        }
        gw[0] = gw[1] = gw[2] = 0.1259391805448 / 2;
        gw[3] = gw[4] = gw[5] = 0.1323941527885 / 2;
        gw[6] = 0.225 / 2;
    } else if (n == 12) {
        const int perm3[3][3] = {{0, 1, 2}, {2, 0, 1}, {1, 2, 0}};
        const double c1[2] = {0.873821971016996, 0.501426509658179};
        for (int s = 0; s < 2; s++) {
            double v[3] = {c1[s], (1 - c1[s]) / 2, 0};
            v[2] = v[1];
            for (int k = 0; k < 3; k++)
                for (int c = 0; c < 3; c++) g[3 * s + k][c] = v[perm3[k][c]];
        }
        double v[3] = {0.636502499121399, 0.310352451033784, 0};
        v[2] = 1 - v[0] - v[1];
        const int perm6[6][3] = {{0, 1, 2}, {0, 2, 1}, {1, 0, 2}, {1, 2, 0}, {2, 0, 1}, {2, 1, 0}};
        for (int k = 0; k < 6; k++)
            for (int c = 0; c < 3; c++) g[6 + k][c] = v[perm6[k][c]];
        double w1 = 0.050844906370207 / 2, w2 = 0.116786275726379 / 2;
        double w3 = (0.5 - 3 * w1 - 3 * w2) / 6;
        for (int k = 0; k < 12; k++) gw[k] = k < 3 ? w1 : (k < 6 ? w2 : w3);
    } else if (n == 19) {
        g[0][0] = g[0][1] = g[0][2] = 1.0 / 3;
        gw[0] = 0.09713579628279610 / 2;
        const double pa[4][2] = {{0.02063496160252593, 0.48968251919873700},
                                 {0.1258208170141290, 0.4370895914929355},
                                 {0.6235929287619356, 0.1882035356190322},
                                 {0.91054097321109410, 0.04472951339445297}};
        const double pw[4] = {0.09400410068141950 / 6, 0.2334826230143263 / 6, 0.2389432167816273 / 6,
                              0.07673302697609430 / 6};
        for (int s = 0; s < 4; s++) {
            int b = 1 + 3 * s;
            double v1 = pa[s][0], v2 = pa[s][1];
            const double t[3][3] = {{v1, v2, v2}, {v2, v1, v2}, {v2, v2, v1}};
            for (int k = 0; k < 3; k++) {
                for (int c = 0; c < 3; c++) g[b + k][c] = t[k][c];
                gw[b + k] = pw[s];
            }
        }
        double v[3] = {0.03683841205473626, 0.74119859878449800, 0};
        v[2] = 1 - v[0] - v[1];
        const int perm6[6][3] = {{0, 1, 2}, {0, 2, 1}, {1, 0, 2}, {1, 2, 0}, {2, 0, 1}, {2, 1, 0}};
        for (int k = 0; k < 6; k++) {
            for (int c = 0; c < 3; c++) g[13 + k][c] = v[perm6[k][c]];
            gw[13 + k] = 0.2597012362637364 / 12;
        }
    } else {
        return -1;
    }
    for (int i = 0; i < n; i++) {
        nodes[3 * i + 0] = g[i][0];
        nodes[3 * i + 1] = g[i][2];
        nodes[3 * i + 2] = g[i][1];
        w[i] = 2.0 * gw[i];
    }
    return 0;
}

double clamp1h(double x) { return x > 1.0 ? 1.0 : (x < -1.0 ? -1.0 : x); }

// Triangle records: readSpaceMesh.m:93-139, calcoloSistemaRiferimento.m:1-45
// and the rotation matrix of rotazioneInversaEulero.m:47-130.
void build_tri(const double* X, const int* tri, int N, std::vector<FieldTri>& out) {
#pragma clang fp contract(off)
    out.resize(N);
    for (int t = 0; t < N; t++) {
        const double *P1 = X + 3 * tri[3 * t], *P2 = X + 3 * tri[3 * t + 1], *P3 = X + 3 * tri[3 * t + 2];
        FieldTri& F = out[t];
        std::memset(&F, 0, sizeof F);
        double vg1[3], vg2[3], vg3[3], n[3];
        for (int k = 0; k < 3; k++) {
            vg1[k] = P1[k] - P3[k];
            vg2[k] = P2[k] - P3[k];
            vg3[k] = P1[k] - P2[k];
        }
        n[0] = vg1[1] * vg2[2] - vg2[1] * vg1[2];
        n[1] = -vg1[0] * vg2[2] + vg2[0] * vg1[2];
        n[2] = vg1[0] * vg2[1] - vg2[0] * vg1[1];
        double a = std::sqrt(n[0] * n[0] + n[1] * n[1] + n[2] * n[2]) / 2;
        F.area = a;
        for (int k = 0; k < 3; k++) F.n[k] = n[k] / (2 * a);
        for (int k = 0; k < 3; k++) F.center[k] = (P1[k] + P2[k] + P3[k]) / 3;
        double l1 = std::sqrt(vg3[0] * vg3[0] + vg3[1] * vg3[1] + vg3[2] * vg3[2]);
        double l2 = std::sqrt(vg2[0] * vg2[0] + vg2[1] * vg2[1] + vg2[2] * vg2[2]);
        double l3 = std::sqrt(vg1[0] * vg1[0] + vg1[1] * vg1[1] + vg1[2] * vg1[2]);
        F.maxL = std::max(std::max(l1, l2), l3);
        for (int k = 0; k < 3; k++) F.v0[k] = P1[k];
        // frame
        double v1[3] = {P2[0] - P1[0], P2[1] - P1[1], P2[2] - P1[2]};
        F.L1 = std::sqrt(v1[0] * v1[0] + v1[1] * v1[1] + v1[2] * v1[2]);
        for (int k = 0; k < 3; k++) F.e1[k] = v1[k] / F.L1;
        double v3[3] = {P3[0] - P1[0], P3[1] - P1[1], P3[2] - P1[2]};
        F.L3 = std::sqrt(v3[0] * v3[0] + v3[1] * v3[1] + v3[2] * v3[2]);
        double e3[3];
        for (int k = 0; k < 3; k++) e3[k] = v3[k] / F.L3;
        F.cosd = e3[0] * F.e1[0] + e3[1] * F.e1[1] + e3[2] * F.e1[2];
        F.sind = std::sin(std::acos(clamp1h(F.cosd)));
        const double* vn = F.n;
        double e2[3] = {vn[1] * F.e1[2] - vn[2] * F.e1[1], vn[2] * F.e1[0] - vn[0] * F.e1[2],
                        vn[0] * F.e1[1] - vn[1] * F.e1[0]};
        double ne = std::sqrt(e2[0] * e2[0] + e2[1] * e2[1] + e2[2] * e2[2]);
        for (int k = 0; k < 3; k++) F.e2[k] = e2[k] / ne;
        // inverse Euler rotation
        const double* Xa = F.e1;
        const double* Z = F.n;
        double cz[3] = {-Z[1], Z[0], 0};
        double ncz = std::sqrt(cz[0] * cz[0] + cz[1] * cz[1] + cz[2] * cz[2]);
        double phi, omega, psi;
        if (ncz < 1.0e-10) {
            phi = std::acos(clamp1h(Xa[0]));
            if (Xa[1] < -1.0e-10) phi = 2 * kPi - phi;
            omega = kPi * (Z[2] < 0);
            psi = 0;
        } else {
            double l[3] = {cz[0] / ncz, cz[1] / ncz, cz[2] / ncz};
            phi = std::acos(clamp1h(l[0]));
            if (l[1] < -1.0e-10) phi = 2 * kPi - phi;
            omega = std::acos(clamp1h(Z[2]));
            psi = std::acos(clamp1h(l[0] * Xa[0] + l[1] * Xa[1] + l[2] * Xa[2]));
            double cl[3] = {l[1] * Xa[2] - l[2] * Xa[1], l[2] * Xa[0] - l[0] * Xa[2], l[0] * Xa[1] - l[1] * Xa[0]};
            if (cl[0] * Z[0] + cl[1] * Z[1] + cl[2] * Z[2] < -1.0e-10) psi = 2 * kPi - psi;
        }
        double Ap[3][3] = {{std::cos(psi), std::sin(psi), 0}, {-std::sin(psi), std::cos(psi), 0}, {0, 0, 1}};
        double Ao[3][3] = {{1, 0, 0}, {0, std::cos(omega), std::sin(omega)}, {0, -std::sin(omega), std::cos(omega)}};
        double Af[3][3] = {{std::cos(phi), std::sin(phi), 0}, {-std::sin(phi), std::cos(phi), 0}, {0, 0, 1}};
        double T[3][3], C[3][3];
        for (int i = 0; i < 3; i++)
            for (int j = 0; j < 3; j++) T[i][j] = Ap[i][0] * Ao[0][j] + Ap[i][1] * Ao[1][j] + Ap[i][2] * Ao[2][j];
        for (int i = 0; i < 3; i++)
            for (int j = 0; j < 3; j++) {
                C[i][j] = T[i][0] * Af[0][j] + T[i][1] * Af[1][j] + T[i][2] * Af[2][j];
                if (std::fabs(C[i][j]) < 1.0e-14) C[i][j] = 0;
                F.C[3 * i + j] = C[i][j];
            }
    }
}

// Gauss-Hammer nodes/weights per triangle (BEMenerg_coreSA_calcCostantData.m:5-24),
// layout gh[t*128 + {0,32,64,96} + q] = x, y, z, weight.
void build_gh(const double* X, const int* tri, const std::vector<FieldTri>& F, int nGH, std::vector<double>& gh) {
#pragma clang fp contract(off)
    double ghn[28 * 3], ghw[28];
    gauss_hammer(nGH, ghn, ghw);
    int N = (int)F.size();
    gh.assign((size_t)N * 128, 0.0);
    for (int t = 0; t < N; t++) {
        const double* V[3] = {X + 3 * tri[3 * t], X + 3 * tri[3 * t + 1], X + 3 * tri[3 * t + 2]};
        for (int q = 0; q < nGH; q++) {
            for (int k = 0; k < 3; k++)
                gh[(size_t)t * 128 + 32 * k + q] = ghn[3 * q] * V[0][k] + ghn[3 * q + 1] * V[1][k] + ghn[3 * q + 2] * V[2][k];
            gh[(size_t)t * 128 + 96 + q] = F[t].area * ghw[q];
        }
    }
}

struct Timer {
    hipEvent_t a = nullptr, b = nullptr;
    Timer() {
        (void)hipEventCreate(&a);
        (void)hipEventCreate(&b);
    }
    ~Timer() {
        if (a) (void)hipEventDestroy(a);
        if (b) (void)hipEventDestroy(b);
    }
    void start(hipStream_t s) { (void)hipEventRecord(a, s); }
    void stop(hipStream_t s) { (void)hipEventRecord(b, s); }
    double ms() {
        float f = 0.f;
        (void)hipEventSynchronize(b);
        (void)hipEventElapsedTime(&f, a, b);
        return (double)f;
    }
};

}  // namespace

struct ebem_handle {
    std::mutex lock;
    int device = 0;
    hipStream_t stream = nullptr;
    rocblas_handle blas = nullptr;
    int CUs = 256;
    // problem
    int N = 0, Nv = 0;
    std::vector<double> X;
    std::vector<int> T;
    std::vector<FieldTri> tri;
    bool have_mesh = false, have_problem = false;
    double rho = 1, mu = 1, lambda = 1, Tfin = 1, cP = 0, cS = 0, dt = 0;
    int nT = 0;
    int nGH = 7;
    int L = 0;
    DevBuf<FieldTri> d_tri;
    DevBuf<double> d_gh;
    // store
    bool assembled = false;
    int store_L = 0;
    int64_t nnzb = 0, pair_lags = 0, items = 0;
    std::vector<int64_t> seg;
    DevBuf<int64_t> d_seg;
    DevBuf<int32_t> d_col;
    DevBuf<double> d_val;  // 6 * nnzb
    // MOT
    DevBuf<double> d_beta, d_alpha, d_r, d_A, d_Ainv;
    DevBuf<rocblas_int> d_ipiv, d_info;
    bool have_beta = false, have_alpha = false;
    double tim[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    bool sort_items = true;
    int64_t far_items = 0;
    int precision = 0;

    Store store() {
        Store S;
        for (int q = 0; q < 6; q++) S.v[q] = d_val.p + (size_t)q * nnzb;
        S.col = d_col.p;
        S.seg = d_seg.p;
        return S;
    }
};

namespace {

int check_handle(ebem_handle* h) {
    if (!h) return fail(EBEM_EINVAL, "null handle");
    return 0;
}

int upload_geometry(ebem_handle* h) {
    std::vector<double> gh;
    build_gh(h->X.data(), h->T.data(), h->tri, h->nGH, gh);
    HIPC(h->d_tri.alloc(h->N));
    HIPC(hipMemcpyAsync(h->d_tri.p, h->tri.data(), sizeof(FieldTri) * h->N, hipMemcpyHostToDevice, h->stream));
    HIPC(h->d_gh.alloc(gh.size()));
    HIPC(hipMemcpyAsync(h->d_gh.p, gh.data(), sizeof(double) * gh.size(), hipMemcpyHostToDevice, h->stream));
    HIPC(hipStreamSynchronize(h->stream));
    return 0;
}

int compute_L(ebem_handle* h) {
#pragma clang fp contract(off)
    double mx = 0;
    const double* X = h->X.data();
    for (int a = 0; a < h->Nv; a++)
        for (int b = 0; b < h->Nv; b++) {
            double d0 = X[3 * a] - X[3 * b], d1 = X[3 * a + 1] - X[3 * b + 1], d2 = X[3 * a + 2] - X[3 * b + 2];
            double d = std::sqrt(d0 * d0 + d1 * d1 + d2 * d2);
            if (d > mx) mx = d;
        }
    int m = (int)std::ceil(mx / (h->cS * h->dt)) + 1;
    h->L = std::min(h->nT, m);
    return 0;
}

template <int NGH, typename R>
void launch_levels(ebem_handle* h, AsmParams P, int r0, int64_t nitems, const int32_t* item_pair, const int64_t* off,
                   const int16_t* lo, const int32_t* perm, double* W) {
    if (nitems <= 0) return;
    int64_t blocks = (nitems + 255) / 256;
    int64_t cap = (int64_t)h->CUs * 16;
    if (blocks > cap) blocks = cap;
    if (blocks < 1) blocks = 1;
    hipLaunchKernelGGL((k_levels<NGH, R>), dim3((unsigned)blocks), dim3(256), 0, h->stream, h->d_tri.p, h->d_gh.p, P,
                       r0, nitems, item_pair, off, lo, perm, W);
}

// Far-field items (keys below kFrontKeys, a prefix of the sorted permutation)
// in R_far, the rest in FP64.
template <int NGH>
void launch_levels_split(ebem_handle* h, AsmParams P, int r0, int64_t nitems, int64_t nfar, const int32_t* item_pair,
                         const int64_t* off, const int16_t* lo, const int32_t* perm, double* W) {
    if (perm && nfar > 0) {
        launch_levels<NGH, float>(h, P, r0, nfar, item_pair, off, lo, perm, W);
        launch_levels<NGH, double>(h, P, r0, nitems - nfar, item_pair, off, lo, perm + nfar, W);
    } else {
        launch_levels<NGH, double>(h, P, r0, nitems, item_pair, off, lo, perm, W);
    }
}

int assemble(ebem_handle* h, int precision) {
    if (!h->have_mesh || !h->have_problem) return fail(EBEM_EINVAL, "mesh and problem must be set before ebem_assemble");
    if (precision != EBEM_PREC_FP64 && precision != EBEM_PREC_MIXED)
        return fail(EBEM_EINVAL, "unknown precision %d", precision);
    compute_L(h);
    const int N = h->N, L = h->L;
    if (L > 4096) return fail(EBEM_EINVAL, "too many lag blocks (%d > 4096)", L);
    AsmParams P;
    P.N = N;
    P.L = L;
    P.dt = h->dt;
    P.cP = h->cP;
    P.cS = h->cS;
    P.rho = h->rho;
    P.nGH = h->nGH;
    {
        // far field: centroid distance > 3h, h = longest edge of the mesh
        double hmax = 0;
        for (int t = 0; t < N; t++) hmax = std::max(hmax, h->tri[t].maxL);
        P.far2 = (precision == EBEM_PREC_MIXED) ? (3.0 * hmax) * (3.0 * hmax) : -1.0;
    }
    Timer tall, tlev;
    double lev_ms = 0;
    tall.start(h->stream);
    const int64_t np = (int64_t)N * N;
    h->far_items = 0;
    h->precision = precision;
    DevBuf<int16_t> d_lo, d_hi;
    HIPC(d_lo.alloc(np));
    HIPC(d_hi.alloc(np));
    {
        int64_t blocks = std::min<int64_t>((np + 255) / 256, (int64_t)h->CUs * 64);
        hipLaunchKernelGGL(k_windows, dim3((unsigned)blocks), dim3(256), 0, h->stream, h->d_tri.p, P, 0, N, d_lo.p,
                           d_hi.p);
        HIPC(hipGetLastError());
    }
    DevBuf<int64_t> d_cnt, d_items;
    HIPC(d_cnt.alloc((size_t)N * L));
    HIPC(d_items.alloc(N));
    hipLaunchKernelGGL(k_rowcount, dim3(N), dim3(256), (L + 1) * sizeof(int), h->stream, d_lo.p, d_hi.p, N, L, 0, N,
                       d_cnt.p, d_items.p);
    HIPC(hipGetLastError());
    std::vector<int64_t> cnt((size_t)N * L), rowitems(N);
    HIPC(hipMemcpyAsync(cnt.data(), d_cnt.p, sizeof(int64_t) * cnt.size(), hipMemcpyDeviceToHost, h->stream));
    HIPC(hipMemcpyAsync(rowitems.data(), d_items.p, sizeof(int64_t) * N, hipMemcpyDeviceToHost, h->stream));
    HIPC(hipStreamSynchronize(h->stream));
    h->seg.assign((size_t)N * L + 1, 0);
    for (size_t k = 0; k < cnt.size(); k++) h->seg[k + 1] = h->seg[k] + cnt[k];
    h->nnzb = h->seg.back();
    h->pair_lags = h->nnzb;
    h->items = 0;
    for (int i = 0; i < N; i++) h->items += rowitems[i];
    HIPC(h->d_seg.alloc(h->seg.size()));
    HIPC(hipMemcpyAsync(h->d_seg.p, h->seg.data(), sizeof(int64_t) * h->seg.size(), hipMemcpyHostToDevice, h->stream));
    h->d_val.reset();
    h->d_col.reset();
    HIPC(h->d_val.alloc((size_t)6 * std::max<int64_t>(h->nnzb, 1)));
    HIPC(h->d_col.alloc((size_t)std::max<int64_t>(h->nnzb, 1)));
    Store S = h->store();

    // row chunks bounded by the W scratch budget
    size_t freeb = 0, totb = 0;
    HIPC(hipMemGetInfo(&freeb, &totb));
    int64_t budget_items = (int64_t)std::min<double>(256.0e6, 0.35 * (double)freeb / (6 * 8 + 4 + 4 + 1));
    if (budget_items < 1) return fail(EBEM_ENOMEM, "not enough device memory for assembly scratch");
    int64_t maxrow_items = *std::max_element(rowitems.begin(), rowitems.end());
    if (maxrow_items > budget_items) budget_items = maxrow_items;
    DevBuf<double> d_W;
    DevBuf<int32_t> d_ip;
    DevBuf<int64_t> d_off, d_rowbase;
    int64_t chunk_cap = std::min<int64_t>(budget_items, std::max<int64_t>(h->items, 1));
    HIPC(d_W.alloc((size_t)chunk_cap * 6));
    HIPC(d_ip.alloc((size_t)chunk_cap));
    DevBuf<int32_t> d_perm;
    DevBuf<uint8_t> d_key;
    DevBuf<unsigned int> d_hist;
    DevBuf<unsigned long long> d_cur;
    const int G = h->CUs * 8;  // sort blocks
    HIPC(d_perm.alloc((size_t)chunk_cap));
    HIPC(d_key.alloc((size_t)chunk_cap));
    HIPC(d_hist.alloc((size_t)kNumKeys * G));
    HIPC(d_cur.alloc((size_t)kNumKeys * G));
    std::vector<unsigned int> hist((size_t)kNumKeys * G);
    std::vector<unsigned long long> cur((size_t)kNumKeys * G);
    const double inv4pirho = 1.0 / (4 * kPi * h->rho);
    int r0 = 0;
    while (r0 < N) {
        int r1 = r0;
        int64_t acc = 0;
        while (r1 < N && acc + rowitems[r1] <= chunk_cap) acc += rowitems[r1++];
        if (r1 == r0) r1 = r0 + 1, acc = rowitems[r0];
        const int rows = r1 - r0;
        std::vector<int64_t> rb(rows);
        int64_t run = 0;
        for (int r = 0; r < rows; r++) {
            rb[r] = run;
            run += rowitems[r0 + r];
        }
        HIPC(d_rowbase.alloc(rows));
        HIPC(hipMemcpyAsync(d_rowbase.p, rb.data(), sizeof(int64_t) * rows, hipMemcpyHostToDevice, h->stream));
        HIPC(d_off.alloc((size_t)rows * N));
        const int16_t* lo = d_lo.p + (int64_t)r0 * N;
        const int16_t* hi = d_hi.p + (int64_t)r0 * N;
        hipLaunchKernelGGL(k_pair_items, dim3(rows), dim3(256), 0, h->stream, lo, hi, N, d_rowbase.p, d_off.p, d_ip.p);
        HIPC(hipGetLastError());
        if (acc > 0) {
            // bucket items by front-state key (counting sort) for convergent wavefronts
            const int64_t per = (acc + G - 1) / G;
            hipLaunchKernelGGL(k_item_keys, dim3(G), dim3(256), 0, h->stream, h->d_tri.p, P, r0, acc, per, d_ip.p,
                               d_off.p, lo, d_key.p, d_hist.p);
            HIPC(hipGetLastError());
            HIPC(hipMemcpyAsync(hist.data(), d_hist.p, sizeof(unsigned int) * hist.size(), hipMemcpyDeviceToHost,
                                h->stream));
            HIPC(hipStreamSynchronize(h->stream));
            unsigned long long runk = 0;
            for (size_t k = 0; k < hist.size(); k++) {  // key-major, block-minor
                cur[k] = runk;
                runk += hist[k];
            }
            HIPC(hipMemcpyAsync(d_cur.p, cur.data(), sizeof(unsigned long long) * cur.size(), hipMemcpyHostToDevice,
                                h->stream));
            hipLaunchKernelGGL(k_item_scatter, dim3(G), dim3(256), 0, h->stream, acc, per, d_key.p, d_cur.p,
                               d_perm.p);
            HIPC(hipGetLastError());
            const int32_t* perm = (h->sort_items || precision == EBEM_PREC_MIXED) ? d_perm.p : nullptr;
            int64_t nfar = 0;
            for (size_t k = 0; k < (size_t)kFrontKeys * G; k++) nfar += hist[k];
            h->far_items += nfar;
            tlev.start(h->stream);
            switch (h->nGH) {
                case 3: launch_levels_split<3>(h, P, r0, acc, nfar, d_ip.p, d_off.p, lo, perm, d_W.p); break;
                case 7: launch_levels_split<7>(h, P, r0, acc, nfar, d_ip.p, d_off.p, lo, perm, d_W.p); break;
                case 12: launch_levels_split<12>(h, P, r0, acc, nfar, d_ip.p, d_off.p, lo, perm, d_W.p); break;
                default: launch_levels_split<19>(h, P, r0, acc, nfar, d_ip.p, d_off.p, lo, perm, d_W.p); break;
            }
            HIPC(hipGetLastError());
            tlev.stop(h->stream);
            lev_ms += tlev.ms();
        }
        hipLaunchKernelGGL(k_emit, dim3(rows), dim3(64), L * sizeof(int64_t), h->stream, lo, hi, d_off.p, d_W.p, N, L,
                           r0, inv4pirho, S);
        HIPC(hipGetLastError());
        r0 = r1;
    }
    tall.stop(h->stream);
    h->tim[1] = tall.ms();
    h->tim[0] = lev_ms;
    HIPC(hipStreamSynchronize(h->stream));
    h->assembled = true;
    h->store_L = L;
    h->have_alpha = false;
    return 0;
}

int rocb(rocblas_status s, const char* what) {
    if (s != rocblas_status_success) return fail(EBEM_ESOLVER, "%s failed: rocblas status %d", what, (int)s);
    return 0;
}

int factor_e0(ebem_handle* h) {
    const int n3 = 3 * h->N;
    const int L = h->store_L;
    HIPC(h->d_A.alloc((size_t)n3 * n3));
    HIPC(h->d_Ainv.alloc((size_t)n3 * n3));
    HIPC(h->d_ipiv.alloc(n3));
    HIPC(h->d_info.alloc(1));
    HIPC(hipMemsetAsync(h->d_A.p, 0, sizeof(double) * (size_t)n3 * n3, h->stream));
    hipLaunchKernelGGL(k_e0_dense, dim3(4, h->N), dim3(64), 0, h->stream, h->store(), h->N, L, h->d_A.p);
    HIPC(hipGetLastError());
    if (!h->blas) {
        if (int e = rocb(rocblas_create_handle(&h->blas), "rocblas_create_handle")) return e;
    }
    if (int e = rocb(rocblas_set_stream(h->blas, h->stream), "rocblas_set_stream")) return e;
    if (int e = rocb(rocsolver_dgetrf(h->blas, n3, n3, h->d_A.p, n3, h->d_ipiv.p, h->d_info.p), "rocsolver_dgetrf"))
        return e;
    rocblas_int info = 0;
    HIPC(hipMemcpyAsync(&info, h->d_info.p, sizeof(info), hipMemcpyDeviceToHost, h->stream));
    HIPC(hipStreamSynchronize(h->stream));
    if (info != 0) return fail(EBEM_ESINGULAR, "E_0 is singular (zero pivot at %d)", (int)info);
    if (int e = rocb(rocsolver_dgetri(h->blas, n3, h->d_A.p, n3, h->d_ipiv.p, h->d_info.p), "rocsolver_dgetri"))
        return e;
    HIPC(hipMemcpyAsync(&info, h->d_info.p, sizeof(info), hipMemcpyDeviceToHost, h->stream));
    HIPC(hipStreamSynchronize(h->stream));
    if (info != 0) return fail(EBEM_ESINGULAR, "E_0 inversion failed (info %d)", (int)info);
    dim3 grid((n3 + 31) / 32, (n3 + 31) / 32);
    hipLaunchKernelGGL(k_transpose, grid, dim3(32, 8), 0, h->stream, h->d_A.p, h->d_Ainv.p, n3);
    HIPC(hipGetLastError());
    return 0;
}

int mot(ebem_handle* h, const double* beta, double* alpha_out) {
    if (!h->assembled) return fail(EBEM_EINVAL, "ebem_assemble must run before ebem_mot_solve");
    const int n3 = 3 * h->N, nT = h->nT;
    if (beta) {
        HIPC(h->d_beta.alloc((size_t)n3 * nT));
        HIPC(hipMemcpyAsync(h->d_beta.p, beta, sizeof(double) * (size_t)n3 * nT, hipMemcpyHostToDevice, h->stream));
        h->have_beta = true;
    } else if (!h->have_beta) {
        return fail(EBEM_EINVAL, "no right-hand side: pass beta or call ebem_rhs first");
    }
    HIPC(h->d_alpha.alloc((size_t)n3 * nT));
    HIPC(h->d_r.alloc(n3));
    Timer tall, tf;
    tall.start(h->stream);
    tf.start(h->stream);
    if (int e = factor_e0(h)) return e;
    tf.stop(h->stream);
    h->tim[2] = tf.ms();
    Store S = h->store();
    hipEvent_t ev[4];
    for (auto& e : ev) HIPC(hipEventCreate(&e));
    double th = 0, ts = 0;
    for (int n = 0; n < nT; n++) {
        HIPC(hipEventRecord(ev[0], h->stream));
        hipLaunchKernelGGL(k_history, dim3(h->N), dim3(256), 0, h->stream, S, h->N, h->store_L, n, h->d_beta.p,
                           h->d_alpha.p, h->d_r.p);
        HIPC(hipEventRecord(ev[1], h->stream));
        hipLaunchKernelGGL(k_gemv_rows, dim3((n3 + 3) / 4), dim3(256), 0, h->stream, h->d_Ainv.p, h->d_r.p,
                           h->d_alpha.p + (size_t)n * n3, n3);
        HIPC(hipEventRecord(ev[2], h->stream));
        HIPC(hipEventSynchronize(ev[2]));
        float a = 0, b = 0;
        HIPC(hipEventElapsedTime(&a, ev[0], ev[1]));
        HIPC(hipEventElapsedTime(&b, ev[1], ev[2]));
        th += a;
        ts += b;
    }
    HIPC(hipGetLastError());
    for (auto& e : ev) (void)hipEventDestroy(e);
    tall.stop(h->stream);
    h->tim[5] = tall.ms();
    h->tim[3] = th;
    h->tim[4] = ts;
    h->tim[7] = nT;
    if (alpha_out)
        HIPC(hipMemcpyAsync(alpha_out, h->d_alpha.p, sizeof(double) * (size_t)n3 * nT, hipMemcpyDeviceToHost, h->stream));
    HIPC(hipStreamSynchronize(h->stream));
    h->have_alpha = true;
    return 0;
}

// Post-processing: times grouped by delta = t - nhat dt (shared M_m), points
// chunked so that the M scratch stays within budget.
int pp_eval(ebem_handle* h, const double* density, int npts, const double* x, int nt, const double* t,
            double* u_out) {
#pragma clang fp contract(off)
    if (!h->have_mesh || !h->have_problem) return fail(EBEM_EINVAL, "mesh and problem must be set");
    if (npts <= 0 || nt <= 0 || !x || !t || !u_out) return fail(EBEM_EINVAL, "bad post-processing arguments");
    if (!density && !h->have_alpha) return fail(EBEM_EINVAL, "no density: pass one or run ebem_mot_solve first");
    const int N = h->N, nT = h->nT;
    const double dt = h->dt;
    std::vector<int> nhat(nt);
    std::vector<double> delta(nt);
    for (int k = 0; k < nt; k++) {
        if (!(t[k] > 0)) {
            nhat[k] = -1;
            delta[k] = 0;
            continue;
        }
        nhat[k] = (int)std::ceil(t[k] / dt) - 1;  // singlePoint.m:21
        if (nhat[k] > nT - 1) return fail(EBEM_EINVAL, "time %g beyond the density's last step", t[k]);
        delta[k] = t[k] - nhat[k] * dt;
    }
    Timer tall;
    tall.start(h->stream);
    DevBuf<double> d_x, d_aT, d_u, d_dens, d_M;
    HIPC(d_x.alloc(3 * (size_t)npts));
    HIPC(hipMemcpyAsync(d_x.p, x, sizeof(double) * 3 * npts, hipMemcpyHostToDevice, h->stream));
    const double* dens = h->d_alpha.p;
    if (density) {
        HIPC(d_dens.alloc((size_t)3 * N * nT));
        HIPC(hipMemcpyAsync(d_dens.p, density, sizeof(double) * 3 * N * nT, hipMemcpyHostToDevice, h->stream));
        dens = d_dens.p;
    }
    HIPC(d_aT.alloc((size_t)3 * N * nT));
    {
        int64_t tot = (int64_t)3 * N * nT;
        hipLaunchKernelGGL(k_alpha_T, dim3((unsigned)((tot + 255) / 256)), dim3(256), 0, h->stream, dens, N, nT,
                           d_aT.p);
        HIPC(hipGetLastError());
    }
    HIPC(d_u.alloc((size_t)3 * npts * nt));
    HIPC(hipMemsetAsync(d_u.p, 0, sizeof(double) * 3 * npts * nt, h->stream));
    // group times by delta
    std::vector<int> order;
    for (int k = 0; k < nt; k++)
        if (nhat[k] >= 0) order.push_back(k);
    std::sort(order.begin(), order.end(), [&](int a, int b) { return delta[a] < delta[b]; });
    const int P0 = std::min(npts, 1024);  // points per chunk
    const int G = h->CUs * 8;             // counting-sort blocks
    DevBuf<int32_t> d_ms, d_me, d_iq, d_nh, d_kidx, d_perm;
    DevBuf<int64_t> d_off, d_rb, d_cnt;
    DevBuf<uint8_t> d_key;
    DevBuf<unsigned int> d_bh;
    DevBuf<unsigned long long> d_bo;
    HIPC(d_ms.alloc((size_t)P0 * N));
    HIPC(d_me.alloc((size_t)P0 * N));
    HIPC(d_off.alloc((size_t)P0 * N));
    HIPC(d_rb.alloc(P0));
    HIPC(d_cnt.alloc(P0));
    HIPC(d_bh.alloc((size_t)kNumKeys * G));
    HIPC(d_bo.alloc((size_t)kNumKeys * G));
    std::vector<unsigned int> bh((size_t)kNumKeys * G);
    std::vector<unsigned long long> bo((size_t)kNumKeys * G);
    std::vector<int64_t> cnt(P0), rb(P0);
    size_t g0 = 0;
    while (g0 < order.size()) {
        size_t g1 = g0 + 1;
        while (g1 < order.size() && delta[order[g1]] - delta[order[g0]] <= 1e-12 * dt) g1++;
        std::vector<int32_t> nh, kid;
        int mmax = -1, krep = order[g0];
        for (size_t q = g0; q < g1; q++) {
            nh.push_back(nhat[order[q]]);
            kid.push_back(order[q]);
            if (nhat[order[q]] > mmax) {
                mmax = nhat[order[q]];
                krep = order[q];
            }
        }
        PPParams P;
        P.N = N;
        P.nT = nT;
        P.dt = dt;
        P.cP = h->cP;
        P.cS = h->cS;
        P.delta = delta[krep];
        P.inv4pirho = 1.0 / (4 * kPi * h->rho);
        P.mmax = mmax;
        P.trep = t[krep];
        P.nrep = mmax;
        HIPC(d_nh.alloc(nh.size()));
        HIPC(d_kidx.alloc(kid.size()));
        HIPC(hipMemcpyAsync(d_nh.p, nh.data(), sizeof(int32_t) * nh.size(), hipMemcpyHostToDevice, h->stream));
        HIPC(hipMemcpyAsync(d_kidx.p, kid.data(), sizeof(int32_t) * kid.size(), hipMemcpyHostToDevice, h->stream));
        for (int p0 = 0; p0 < npts; p0 += P0) {
            const int np = std::min(P0, npts - p0);
            int64_t tot = (int64_t)np * N;
            int64_t blocks = std::min<int64_t>((tot + 255) / 256, (int64_t)h->CUs * 64);
            hipLaunchKernelGGL(k_pp_windows, dim3((unsigned)blocks), dim3(256), 0, h->stream, h->d_tri.p, d_x.p, p0,
                               np, P, d_ms.p, d_me.p);
            HIPC(hipGetLastError());
            hipLaunchKernelGGL(k_pp_counts, dim3(np), dim3(256), 0, h->stream, d_ms.p, d_me.p, N, d_cnt.p);
            HIPC(hipGetLastError());
            HIPC(hipMemcpyAsync(cnt.data(), d_cnt.p, sizeof(int64_t) * np, hipMemcpyDeviceToHost, h->stream));
            HIPC(hipStreamSynchronize(h->stream));
            int64_t run = 0;
            for (int r = 0; r < np; r++) {
                rb[r] = run;
                run += cnt[r];
            }
            const int64_t nitems = run;
            if (nitems == 0) continue;
            if (nitems > INT32_MAX) return fail(EBEM_EINVAL, "too many post-processing items in one chunk");
            HIPC(hipMemcpyAsync(d_rb.p, rb.data(), sizeof(int64_t) * np, hipMemcpyHostToDevice, h->stream));
            HIPC(d_iq.alloc((size_t)nitems));
            HIPC(d_M.alloc((size_t)nitems * 6));
            HIPC(d_key.alloc((size_t)nitems));
            HIPC(d_perm.alloc((size_t)nitems));
            hipLaunchKernelGGL(k_pp_items, dim3(np), dim3(256), 0, h->stream, d_ms.p, d_me.p, N, d_rb.p, d_off.p,
                               d_iq.p);
            HIPC(hipGetLastError());
            // bucket items by front-state key (two-pass counting sort)
            const int64_t per = (nitems + G - 1) / G;
            hipLaunchKernelGGL(k_pp_keys, dim3(G), dim3(256), 0, h->stream, h->d_tri.p, d_x.p, p0, P, nitems, per,
                               d_iq.p, d_off.p, d_ms.p, d_key.p, d_bh.p);
            HIPC(hipGetLastError());
            HIPC(hipMemcpyAsync(bh.data(), d_bh.p, sizeof(unsigned int) * bh.size(), hipMemcpyDeviceToHost,
                                h->stream));
            HIPC(hipStreamSynchronize(h->stream));
            unsigned long long runk = 0;
            for (size_t k = 0; k < bh.size(); k++) {
                bo[k] = runk;
                runk += bh[k];
            }
            HIPC(hipMemcpyAsync(d_bo.p, bo.data(), sizeof(unsigned long long) * bo.size(), hipMemcpyHostToDevice,
                                h->stream));
            hipLaunchKernelGGL(k_item_scatter, dim3(G), dim3(256), 0, h->stream, nitems, per, d_key.p, d_bo.p,
                               d_perm.p);
            HIPC(hipGetLastError());
            int64_t mb = std::min<int64_t>((nitems + 255) / 256, (int64_t)h->CUs * 16);
            hipLaunchKernelGGL(k_pp_M, dim3((unsigned)mb), dim3(256), 0, h->stream, h->d_tri.p, d_x.p, p0, P, nitems,
                               d_iq.p, d_off.p, d_ms.p, h->sort_items ? d_perm.p : nullptr, d_M.p);
            HIPC(hipGetLastError());
            hipLaunchKernelGGL(k_pp_conv, dim3(np), dim3(128 * kConvSlices), 0, h->stream, d_M.p, d_off.p, d_ms.p,
                               d_me.p, d_aT.p, N, nT, p0, d_nh.p, d_kidx.p, (int)nh.size(), npts, d_u.p);
            HIPC(hipGetLastError());
        }
        g0 = g1;
    }
    tall.stop(h->stream);
    HIPC(hipMemcpyAsync(u_out, d_u.p, sizeof(double) * 3 * npts * nt, hipMemcpyDeviceToHost, h->stream));
    HIPC(hipStreamSynchronize(h->stream));
    h->tim[6] = tall.ms();
    return 0;
}

}  // namespace

// ================================================================= C ABI ==

extern "C" {

const char* ebem_last_error(void) { return g_err.c_str(); }

int ebem_create(ebem_handle** out, int device) {
    if (!out) return fail(EBEM_EINVAL, "null output pointer");
    *out = nullptr;
    int ndev = 0;
    HIPC(hipGetDeviceCount(&ndev));
    if (device < 0 || device >= ndev) return fail(EBEM_EINVAL, "device %d not available (%d devices)", device, ndev);
    HIPC(hipSetDevice(device));
    ebem_handle* h = new ebem_handle();
    h->device = device;
    hipError_t e = hipStreamCreate(&h->stream);
    if (e != hipSuccess) {
        delete h;
        return fail(EBEM_EHIP, "hipStreamCreate failed: %s", hipGetErrorString(e));
    }
    int cus = 0;
    if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, device) == hipSuccess && cus > 0) h->CUs = cus;
    // EBEM_ITEM_ORDER=pair keeps the assembly items in pair-major order instead
    // of bucketing them by front-state key (same results, for A/B timing).
    const char* order = std::getenv("EBEM_ITEM_ORDER");
    h->sort_items = !(order && std::strcmp(order, "pair") == 0);
    *out = h;
    return 0;
}

int ebem_destroy(ebem_handle* h) {
    if (!h) return 0;
    {
        std::lock_guard<std::mutex> lk(h->lock);
        (void)hipSetDevice(h->device);
        (void)hipStreamSynchronize(h->stream);
        if (h->blas) rocblas_destroy_handle(h->blas);
        h->d_tri.reset();
        h->d_gh.reset();
        h->d_seg.reset();
        h->d_col.reset();
        h->d_val.reset();
        h->d_beta.reset();
        h->d_alpha.reset();
        h->d_r.reset();
        h->d_A.reset();
        h->d_Ainv.reset();
        h->d_ipiv.reset();
        h->d_info.reset();
        (void)hipStreamDestroy(h->stream);
    }
    delete h;
    return 0;
}

int ebem_set_mesh(ebem_handle* h, int N, int Nv, const double* coords, const int32_t* tri, int one_based) {
    if (int e = check_handle(h)) return e;
    std::lock_guard<std::mutex> lk(h->lock);
    if (N <= 0 || Nv <= 0 || !coords || !tri) return fail(EBEM_EINVAL, "invalid mesh (N=%d, Nv=%d)", N, Nv);
    if ((int64_t)N * N > (int64_t)1 << 31) return fail(EBEM_EINVAL, "mesh too large (N=%d)", N);
    HIPC(hipSetDevice(h->device));
    h->N = N;
    h->Nv = Nv;
    h->X.assign(coords, coords + 3 * (size_t)Nv);
    h->T.resize(3 * (size_t)N);
    for (size_t k = 0; k < h->T.size(); k++) {
        int v = tri[k] - (one_based ? 1 : 0);
        if (v < 0 || v >= Nv) return fail(EBEM_EINVAL, "triangle %zu has vertex index %d out of range", k / 3, tri[k]);
        h->T[k] = v;
    }
    build_tri(h->X.data(), h->T.data(), N, h->tri);
    for (int t = 0; t < N; t++)
        if (!(h->tri[t].area > 0)) return fail(EBEM_EINVAL, "triangle %d is degenerate", t);
    h->have_mesh = true;
    h->assembled = false;
    h->have_beta = h->have_alpha = false;
    return upload_geometry(h);
}

int ebem_set_problem(ebem_handle* h, double rho, double mu, double lambda, double Tfin, int nT) {
    if (int e = check_handle(h)) return e;
    std::lock_guard<std::mutex> lk(h->lock);
    if (!(rho > 0) || !(mu > 0) || !(lambda + 2 * mu > 0) || !(Tfin > 0) || nT <= 0)
        return fail(EBEM_EINVAL, "invalid problem parameters");
    h->rho = rho;
    h->mu = mu;
    h->lambda = lambda;
    h->Tfin = Tfin;
    h->nT = nT;
    h->cP = std::sqrt((lambda + 2 * mu) / rho);
    h->cS = std::sqrt(mu / rho);
    h->dt = Tfin / nT;
    h->have_problem = true;
    h->assembled = false;
    h->have_beta = h->have_alpha = false;
    return 0;
}

int ebem_set_rule(ebem_handle* h, int nGH) {
    if (int e = check_handle(h)) return e;
    std::lock_guard<std::mutex> lk(h->lock);
    if (nGH != 3 && nGH != 7 && nGH != 12 && nGH != 19)
        return fail(EBEM_EINVAL, "Valori nodi esterni non validi: %d (allowed 3, 7, 12, 19)", nGH);
    HIPC(hipSetDevice(h->device));
    h->nGH = nGH;
    h->assembled = false;
    if (h->have_mesh) return upload_geometry(h);
    return 0;
}

int ebem_num_blocks(ebem_handle* h, int* L) {
    if (int e = check_handle(h)) return e;
    std::lock_guard<std::mutex> lk(h->lock);
    if (!h->have_mesh || !h->have_problem || !L) return fail(EBEM_EINVAL, "mesh and problem must be set");
    compute_L(h);
    *L = h->L;
    return 0;
}

int ebem_assemble(ebem_handle* h, int precision) {
    if (int e = check_handle(h)) return e;
    std::lock_guard<std::mutex> lk(h->lock);
    HIPC(hipSetDevice(h->device));
    return assemble(h, precision);
}

int ebem_store_info(ebem_handle* h, int64_t* nnzb, int64_t* pair_lags, int64_t* items, int64_t* bytes) {
    if (int e = check_handle(h)) return e;
    std::lock_guard<std::mutex> lk(h->lock);
    if (!h->assembled) return fail(EBEM_EINVAL, "not assembled");
    if (nnzb) *nnzb = h->nnzb;
    if (pair_lags) *pair_lags = h->pair_lags;
    if (items) *items = h->items;
    if (bytes) *bytes = h->nnzb * (6 * 8 + 4) + (int64_t)h->seg.size() * 8;
    return 0;
}

int ebem_assembly_stats(ebem_handle* h, int64_t* far_items, int64_t* fp64_items) {
    if (int e = check_handle(h)) return e;
    std::lock_guard<std::mutex> lk(h->lock);
    if (!h->assembled) return fail(EBEM_EINVAL, "not assembled");
    if (far_items) *far_items = h->far_items;
    if (fp64_items) *fp64_items = h->items - h->far_items;
    return 0;
}

int ebem_get_block(ebem_handle* h, int lag, double* out) {
    if (int e = check_handle(h)) return e;
    std::lock_guard<std::mutex> lk(h->lock);
    if (!h->assembled) return fail(EBEM_EINVAL, "not assembled");
    if (!out) return fail(EBEM_EINVAL, "null output");
    HIPC(hipSetDevice(h->device));
    const size_t n3 = 3 * (size_t)h->N;
    if (lag < 0 || lag >= h->store_L) {
        if (lag >= 0) {
            std::memset(out, 0, sizeof(double) * n3 * n3);
            return 0;  // beyond L: exactly zero
        }
        return fail(EBEM_EINVAL, "negative lag");
    }
    DevBuf<double> d;
    HIPC(d.alloc(n3 * n3));
    HIPC(hipMemsetAsync(d.p, 0, sizeof(double) * n3 * n3, h->stream));
    hipLaunchKernelGGL(k_dense_lag, dim3(4, h->N), dim3(64), 0, h->stream, h->store(), h->N, h->store_L, lag, d.p);
    HIPC(hipGetLastError());
    HIPC(hipMemcpyAsync(out, d.p, sizeof(double) * n3 * n3, hipMemcpyDeviceToHost, h->stream));
    HIPC(hipStreamSynchronize(h->stream));
    return 0;
}

int ebem_rhs(ebem_handle* h, int datum_kind, double* beta_out) {
    if (int e = check_handle(h)) return e;
    std::lock_guard<std::mutex> lk(h->lock);
    if (!h->have_mesh || !h->have_problem) return fail(EBEM_EINVAL, "mesh and problem must be set");
    if (datum_kind < 0 || datum_kind > 3) return fail(EBEM_EINVAL, "unknown datum %d", datum_kind);
    HIPC(hipSetDevice(h->device));
    const size_t n3 = 3 * (size_t)h->N;
    HIPC(h->d_beta.alloc(n3 * h->nT));
    int64_t tot = (int64_t)h->N * h->nT;
    hipLaunchKernelGGL(k_rhs, dim3((unsigned)((tot + 255) / 256)), dim3(256), 0, h->stream, h->d_tri.p, h->d_gh.p, h->N,
                       h->nT, h->nGH, h->dt, datum_kind, h->d_beta.p);
    HIPC(hipGetLastError());
    if (beta_out)
        HIPC(hipMemcpyAsync(beta_out, h->d_beta.p, sizeof(double) * n3 * h->nT, hipMemcpyDeviceToHost, h->stream));
    HIPC(hipStreamSynchronize(h->stream));
    h->have_beta = true;
    return 0;
}

int ebem_mot_solve(ebem_handle* h, const double* beta, double* alpha_out) {
    if (int e = check_handle(h)) return e;
    std::lock_guard<std::mutex> lk(h->lock);
    HIPC(hipSetDevice(h->device));
    return mot(h, beta, alpha_out);
}

int ebem_node_integral(ebem_handle* h, const double* x, int tri, double t, double* out9) {
    if (int e = check_handle(h)) return e;
    std::lock_guard<std::mutex> lk(h->lock);
    if (!h->have_mesh || !h->have_problem) return fail(EBEM_EINVAL, "mesh and problem must be set");
    if (tri < 0 || tri >= h->N || !x || !out9) return fail(EBEM_EINVAL, "bad argument");
    HIPC(hipSetDevice(h->device));
    DevBuf<double> d;
    HIPC(d.alloc(12));
    HIPC(hipMemcpyAsync(d.p, x, 3 * sizeof(double), hipMemcpyHostToDevice, h->stream));
    hipLaunchKernelGGL(k_point_block, dim3(1), dim3(64), 0, h->stream, h->d_tri.p, tri, d.p, t, h->cP, h->cS,
                       1.0 / (4 * kPi * h->rho), d.p + 3);
    HIPC(hipGetLastError());
    double w[6];
    HIPC(hipMemcpyAsync(w, d.p + 3, 6 * sizeof(double), hipMemcpyDeviceToHost, h->stream));
    HIPC(hipStreamSynchronize(h->stream));
    const int qq[9] = {0, 1, 2, 1, 3, 4, 2, 4, 5};
    for (int k = 0; k < 9; k++) out9[k] = w[qq[k]];
    return 0;
}

int ebem_pp_eval(ebem_handle* h, const double* density, int npts, const double* x, int nt, const double* t,
                 double* u_out) {
    if (int e = check_handle(h)) return e;
    std::lock_guard<std::mutex> lk(h->lock);
    HIPC(hipSetDevice(h->device));
    return pp_eval(h, density, npts, x, nt, t, u_out);
}

int ebem_timings(ebem_handle* h, double* out8) {
    if (int e = check_handle(h)) return e;
    std::lock_guard<std::mutex> lk(h->lock);
    if (!out8) return fail(EBEM_EINVAL, "null output");
    for (int k = 0; k < 8; k++) out8[k] = h->tim[k];
    return 0;
}

}  // extern "C"
