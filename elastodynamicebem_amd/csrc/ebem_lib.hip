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
#include <condition_variable>
#include <mutex>
#include <thread>
#include <string>
#include <vector>

#include "../../include/ebem.h"
#include "ebem_kernels.hpp"
#include "ebem_sur.hpp"
#include "ebem_ghc.hpp"
#include "ebem_k.hpp"
#include "ebem_ppd.hpp"
#include "ebem_dn.hpp"

using namespace ebem;

// Staging of ebem_analitic_intg_noalloc: a fixed device array per GPU (each
// GPU the library runs on holds its own copy until the process ends), so that
// entry point allocates nothing.  Layout for kAiScratchCalls calls: args (8
// doubles per call), infoTri2D (12), B (54), ris (36), then the kinds (int32).
// Host code never uses the array's name as an address: ai_scratch_ptr[device]
// holds the address hipGetSymbolAddress returned on that device.
constexpr int kAiScratchCalls = 8192;
constexpr size_t kAiScratchBytes = (size_t)kAiScratchCalls * (sizeof(double) * (8 + 12 + 54 + 36) + sizeof(int32_t));
static __device__ __align__(256) unsigned char scratch[kAiScratchBytes];
static_assert(kAiScratchCalls == EBEM_AI_NOALLOC_MAX_CALLS, "staging capacity of ebem_analitic_intg_noalloc");

namespace {

constexpr int kMaxDevices = 64;
std::mutex ai_scratch_lock[kMaxDevices];     // serialises the calls that share one device's array
void* ai_scratch_ptr[kMaxDevices] = {};      // device address of `scratch` per device (set once)
std::mutex ai_scratch_init_lock;

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
    // At least count elements, reallocating only to grow (per-call scratch kept
    // in the handle, so repeated calls do not pay hipMalloc/hipFree).
    hipError_t grow(size_t count) {
        if (count <= n && p) return hipSuccess;
        return alloc(count);
    }
};

// Gauss-Hammer tables of util/quadRules/GaussHammer_base.m:19-255 mapped by
// GaussHammerComposite(1, n) (util/quadRules/GaussHammerComposite.m:31-52):
// barycentric columns (b1,b3,b2), weights doubled.
// util/quadRules/GaussHammer_base.m rows for n in {1, 3, 7, 12, 19}: raw
// barycentric nodes and weights (summing to 1/2).
int gh_base(int n, double g[28][3], double gw[28]) {
    for (int i = 0; i < 28; i++) g[i][0] = g[i][1] = g[i][2] = gw[i] = 0.0;
    if (n == 1) {
        g[0][0] = g[0][1] = g[0][2] = 1.0 / 3;
        gw[0] = 1.0 / 2;
    } else if (n == 3) {
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
    return 0;
}

int gauss_hammer(int n, double* nodes, double* w) {
    double g[28][3], gw[28];
    if (gh_base(n, g, gw)) return -1;
    for (int i = 0; i < n; i++) {
        nodes[3 * i + 0] = g[i][0];
        nodes[3 * i + 1] = g[i][2];
        nodes[3 * i + 2] = g[i][1];
        w[i] = 2.0 * gw[i];
    }
    return 0;
}

// GaussHammerComposite.m:5-52: the reference triangle cut into nSub = nDiv^2
// congruent sub-triangles (rows of 2(nDiv - i) + 1 alternating triangles), the
// nNod-point rule on each; weights (2/nSub) w.  out: nSub*nNod x 4 (b0, b1, b2, w).
int gh_composite(int nSub, int nNod, std::vector<double>& out) {
#pragma clang fp contract(off)
    double g[28][3], gw[28];
    if (gh_base(nNod, g, gw)) return -1;
    int nDiv = (int)std::lround(std::sqrt((double)nSub));
    if (nDiv < 1 || nDiv * nDiv != nSub) return -1;
    out.assign((size_t)4 * nSub * nNod, 0.0);
    const double f = 1.0 / nDiv;
    int n = 0;
    for (int i = 1; i <= nDiv; i++) {
        const double zMin = (i - 1) * f, zMax = i * f;
        const int nTr = 2 * (nDiv - i) + 1;
        double x[3] = {1 - zMin, 1 - zMin, (1 - zMin) - f}, y[3] = {0 - f, 0, 0}, z[3] = {0, 0, 0};
        for (int j = 1; j <= nTr; j++) {
            const double x2 = x[1], x3 = x[2], y2 = y[1], y3 = y[2];
            if (j % 2 == 1) {
                x[0] = x2; x[1] = x3; x[2] = x3;
                y[0] = y2; y[1] = y3; y[2] = y3 + f;
                z[0] = zMin; z[1] = zMax; z[2] = zMin;
            } else {
                x[0] = x2; x[1] = x3; x[2] = x3 - f;
                y[0] = y2; y[1] = y3; y[2] = y3;
                z[0] = zMax; z[1] = zMin; z[2] = zMax;
            }
            for (int q = 0; q < nNod; q++) {
                double* o = out.data() + 4 * ((size_t)n * nNod + q);
                o[0] = g[q][0] * x[0] + g[q][1] * x[1] + g[q][2] * x[2];
                o[1] = g[q][0] * y[0] + g[q][1] * y[1] + g[q][2] * y[2];
                o[2] = g[q][0] * z[0] + g[q][1] * z[1] + g[q][2] * z[2];
                o[3] = (2.0 / nSub) * gw[q];
            }
            n++;
        }
    }
    return 0;
}

// Gauss-Legendre rule on (-1, 1), ascending (Newton on P_n).
void gauss_legendre(int n, double* x, double* w) {
    for (int i = 0; i < n; i++) {
        double r = std::cos(kPi * (i + 0.75) / (n + 0.5));
        double dp = 1.0;
        for (int it = 0; it < 100; it++) {
            double p0 = 1.0, p1 = r;
            for (int k = 2; k <= n; k++) {
                double p2 = ((2.0 * k - 1.0) * r * p1 - (k - 1.0) * p0) / k;
                p0 = p1;
                p1 = p2;
            }
            dp = n * (r * p1 - p0) / (r * r - 1.0);
            double dr = p1 / dp;
            r -= dr;
            if (std::fabs(dr) < 1e-16) break;
        }
        double p0 = 1.0, p1 = r;
        for (int k = 2; k <= n; k++) {
            double p2 = ((2.0 * k - 1.0) * r * p1 - (k - 1.0) * p0) / k;
            p0 = p1;
            p1 = p2;
        }
        dp = n * (r * p1 - p0) / (r * r - 1.0);
        x[n - 1 - i] = r;
        w[n - 1 - i] = 2.0 / ((1.0 - r * r) * dp * dp);
    }
}

// doppioGauss1D (util/quadRules/doppioGauss1D.m): collapsed tensor
// Gauss rule, nodes ((1-x_j)(1-x_i), x_j(1-x_i), x_i), weights 2 w_i w_j (1-x_i)
// from the Gauss-Legendre rule on (0, 1).  out: nNodes x 4.
int doppio_gauss(int nNodes, std::vector<double>& out) {
#pragma clang fp contract(off)
    const int n1 = (int)std::lround(std::sqrt((double)nNodes));
    if (n1 < 1 || n1 * n1 != nNodes) return -1;
    std::vector<double> x(n1), w(n1);
    gauss_legendre(n1, x.data(), w.data());
    for (int k = 0; k < n1; k++) {
        x[k] = 0.5 + (x[k] / 2);
        w[k] = w[k] / 2;
    }
    out.assign((size_t)4 * nNodes, 0.0);
    for (int j = 0; j < n1; j++)
        for (int i = 0; i < n1; i++) {
            double* o = out.data() + 4 * ((size_t)i + (size_t)n1 * j);
            o[0] = (1 - x[j]) * (1 - x[i]);
            o[1] = x[j] * (1 - x[i]);
            o[2] = x[i];
            o[3] = 2 * w[j] * (w[i] * (1 - x[i]));
        }
    return 0;
}

double clamp1h(double x) { return x > 1.0 ? 1.0 : (x < -1.0 ? -1.0 : x); }

// Triangle records: readSpaceMesh.m:93-139, calcoloSistemaRiferimento.m:1-45
// and the rotation matrix of rotazioneInversaEulero.m:19-91.
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
    DevBuf<double> d_verts;  // N x 9: the three vertices of every triangle, mesh order
    double hmax = 0;         // longest edge of the mesh
    // inner rules of the numeric form (EBEM_PREC_GHC), ebem_set_inner_rule
    int ghc_nSub = 0, ghc_nSing = 0, ghc_nDiag = 0;
    DevBuf<double> d_ghc_in, d_ghc_dg;
    // rule of the numeric post-processing, ebem_set_pp_rule
    int pp_nInt = 0;
    DevBuf<double> d_pp_rule;
    // surrogate attached by ebem_attach_surrogate, and what the last assembly gave it
    ebem_sur* sur = nullptr;
    int64_t sur_items = 0, sur_pairs_blocks = 0;
    double sur_ms = 0;
    // store
    bool assembled = false;
    int store_L = 0;
    int64_t nnzb = 0, pair_lags = 0, items = 0;
    std::vector<int64_t> seg;
    DevBuf<int64_t> d_seg;
    DevBuf<int32_t> d_col;
    DevBuf<double> d_val;  // 6 * nnzb
    // MOT
    DevBuf<double> d_beta, d_alpha, d_alphaT, d_dalphaT, d_r, d_A, d_Ainv, d_part;  // alphaT, dalphaT: k_pp_conv
    DevBuf<rocblas_int> d_ipiv, d_info;
    bool have_beta = false, have_alpha = false;
    bool e0_pivoted = false;  // the last factorisation needed partial pivoting
    double tim[EBEM_NUM_TIMINGS] = {};
    // ebem_pp_prepare: the density-independent part of one post-processing
    // request, run by a worker thread on a stream of its own so that it
    // overlaps whatever the caller runs next on the library's stream
    struct PPPrep {
        std::thread worker;
        bool active = false;  // worker started and not yet joined
        bool ready = false;   // the pw buffers hold the prepared part of (x, t)
        int err = 0;
        std::string msg;
        std::vector<double> x, t;
        hipStream_t stream = nullptr;
        double ms = 0;  // span of the prepared part on its stream
        // the worker starts its kernels once E_0 is factorised (the
        // factorisation's many small rocSOLVER kernels slow down 4x beside
        // them) or when a call joins it
        std::mutex go_m;
        std::condition_variable go_cv;
        bool go = false;
    } prep;
    // scratch of ebem_assemble and ebem_pp_eval, kept across calls (grow-only)
    struct {
        DevBuf<uint8_t> flag, key;
        DevBuf<unsigned long long> nsb, culled, cur;
        DevBuf<int16_t> lo, hi, ka, kb, sa, sb;
        DevBuf<unsigned int> nfb, hist;
        DevBuf<int64_t> cnt, items, off, rowbase;
        DevBuf<double> W;
        DevBuf<int32_t> ip, perm;
    } aw;
    struct {
        DevBuf<double> x, aT, u, dens, M;
        DevBuf<int32_t> ms, me, iq, nh, kidx, perm, ka, kb, sa, sb, pout;
        DevBuf<unsigned int> nfb, bh;
        DevBuf<int64_t> off, rb, cnt;
        DevBuf<uint8_t> key;
        DevBuf<unsigned long long> bo;
        int64_t nitems = 0;  // items of the last chunk
    } pw;
    std::vector<hipEvent_t> mot_ev;  // per-step events of ebem_mot_solve
    // double-layer K of the direct Dirichlet driver (ebem_assemble_k):
    // pair windows [klo, khi) over (S, s) = S * Nv + s, slot offsets koff,
    // values 9 per slot; rows S occupy slots [k_rowbase[S], k_rowbase[S+1])
    bool k_ready = false;
    bool kgeo_ready = false;  // d_ktri, d_T, d_X, vertex lists of the current mesh
    bool have_gT = false;     // d_gT holds the vertex datum of the last ebem_rhs_dirichlet
    int k_LK = 0;
    int64_t k_pairs = 0, k_slots = 0;
    std::vector<int64_t> k_rowbase;
    DevBuf<KTri> d_ktri;
    DevBuf<int32_t> d_T, d_vadj_off, d_vadj, d_kslot;
    DevBuf<double> d_X, d_kval, d_gT, d_betak;
    DevBuf<int16_t> d_klo, d_khi;
    DevBuf<int64_t> d_koff, d_krow;
    // Neumann driver: traction datum [triangle][step][3], beta_V, density 3Nv x nT
    DevBuf<double> d_gVT, d_betaV, d_dn;
    bool have_dn = false;
    bool sort_items = true;
    bool kp_split = true;  // kP ranges through k_levels_kp (EBEM_KP_SPLIT=0 turns it off)
    bool node_tasks = true;  // FP64 per-level items by outer node (EBEM_NODE_TASKS=0: k_levels, item per lane)
    struct {
        DevBuf<uint8_t> key;
        DevBuf<int32_t> perm;
        DevBuf<double> part;
        DevBuf<unsigned int> hist;
        DevBuf<unsigned long long> cur;
        DevBuf<unsigned int> bad;
    } nw;  // node-task scratch of ebem_assemble (grow-only)
    // static node values reused across levels: 1 (default), 0 evaluated per
    // level (EBEM_NODE_REUSE=0), 2 every chunk's check reported failed so that
    // it falls back to 0 (EBEM_NODE_REUSE=fallback, for tests of that path)
    int node_reuse = 1;
    // node tasks of the last assembly: evaluated, reused (static), chunks
    // evaluated again without reuse after a failed check of the reused values
    int64_t node_stats[3] = {0, 0, 0};
    bool k_by_field = true;  // regular K part by (S, F) pairs + gather (EBEM_K_BY_FIELD=0: by (S, s) slots)
    bool ppd_by_field = true;  // direct PP double layer by (pair, F) + gather (EBEM_PPD_BY_FIELD=0: by vertex)
    // kP and between-front ranges of the last assembly / post-processing:
    // kP ranges, kP levels, between-front ranges, their levels, fallbacks
    int64_t kp_stats[5] = {0, 0, 0, 0, 0};
    int64_t pp_kp_stats[5] = {0, 0, 0, 0, 0};
    int64_t far_items = 0;
    int precision = 0;
    double far_factor = 3.0;        // mixed form: FP32 beyond far_factor x longest edge (EBEM_OPT_FAR_FACTOR)
    bool allow_sur_solve = false;   // EBEM_OPT_SURROGATE_SOLVE: MOT / PP on a surrogate-assembled store
    bool sur_project = true;        // EBEM_OPT_SURROGATE_PROJECT: moment projection of surrogate pairs

    Store store() {
        Store S;
        for (int q = 0; q < 6; q++) S.v[q] = d_val.p + (size_t)q * nnzb;
        S.col = d_col.p;
        S.seg = d_seg.p;
        return S;
    }
};

struct ebem_sur {
    std::mutex lock;
    int device = 0;
    hipStream_t stream = nullptr;
    int users = 0;  // handles it is attached to
    // samples: raw features (n x 24) and labels (n x 18)
    int64_t n = 0, cap = 0;
    DevBuf<float> d_X, d_Y;
    // parameters (FP32 masters, BF16 copy), gradients, Adam moments, and the
    // normalisation block: x mean, x 1/std, y mean, y std, box lo, box hi
    DevBuf<float> d_p, d_g, d_m, d_v, d_norm, d_part;
    DevBuf<uint16_t> d_pb;
    std::vector<float> norm;
    bool ready = false;
    int64_t adam_t = 0;
    // batch workspace
    int B = 0;
    bool wsZ = false;
    DevBuf<uint16_t> d_A0, d_A[4], d_dYb, d_dZb[2];
    DevBuf<float> d_Z[4], d_Yo, d_T, d_dY, d_dZ[2];
    DevBuf<int32_t> d_idx;
    DevBuf<double> d_loss;
    DevBuf<int64_t> d_slot;
};

namespace {

constexpr int kSurNorm = 2 * kSurIn + 2 * kSurOut + 2 * kSurIn;  // 132 floats
static_assert(kSurNorm == EBEM_SUR_NORM, "normalisation block size");
static_assert(kSurParams == EBEM_SUR_PARAMS, "parameter count");

int sur_workspace(ebem_sur* s, int B, bool train) {
    if (B == s->B && (!train || s->wsZ)) return 0;
    s->B = B;
    s->wsZ = train;
    HIPC(s->d_A0.alloc((size_t)B * kSurInPad));
    for (int l = 0; l < 4; l++) {
        HIPC(s->d_A[l].alloc((size_t)B * kSurHid));
        if (train) HIPC(s->d_Z[l].alloc((size_t)B * kSurHid));
        else s->d_Z[l].reset();
    }
    HIPC(s->d_Yo.alloc((size_t)B * kSurOutPad));
    HIPC(s->d_T.alloc((size_t)B * kSurOut));
    HIPC(s->d_loss.alloc(1));
    if (train) {
        HIPC(s->d_dY.alloc((size_t)B * kSurOutPad));
        HIPC(s->d_dYb.alloc((size_t)B * kSurOutPad));
        for (int q = 0; q < 2; q++) {
            HIPC(s->d_dZ[q].alloc((size_t)B * kSurHid));
            HIPC(s->d_dZb[q].alloc((size_t)B * kSurHid));
        }
    }
    return 0;
}

template <bool TA, bool TB, int EPI>
int gemm(hipStream_t st, const uint16_t* A, int lda, const uint16_t* B, int ldb, int M, int N, int K, int splits,
         const EpiArgs& e) {
    if (M % 64 || N % 64 || K % 32 || lda % 8 || ldb % 8 || splits < 1)
        return fail(EBEM_EINVAL, "gemm shape M=%d N=%d K=%d not tileable", M, N, K);
    int ksplit = ((K / splits + 31) / 32) * 32;
    if (ksplit < 32) ksplit = 32;
    const int z = (K + ksplit - 1) / ksplit;
    hipLaunchKernelGGL((k_gemm_bf16<TA, TB, EPI>), dim3(N / 64, M / 64, z), dim3(256), 0, st, A, lda, B, ldb, K, ksplit,
                       e);
    HIPC(hipGetLastError());
    return 0;
}

// Forward pass of B rows from s->d_A0; keepZ stores the pre-activations.
int sur_forward(ebem_sur* s, hipStream_t st, int B, bool keepZ) {
    for (int l = 0; l < kSurLayers; l++) {
        const uint16_t* in = l == 0 ? s->d_A0.p : s->d_A[l - 1].p;
        const int K = sur_K(l), N = sur_N(l);
        EpiArgs e{};
        e.bias = s->d_p.p + sur_boff(l);
        if (l < kSurLayers - 1) {
            e.C = keepZ ? s->d_Z[l].p : nullptr;
            e.ldc = N;
            e.Cb = s->d_A[l].p;
            e.ldcb = N;
            if (int r = gemm<false, false, kEpiBiasGelu>(st, in, K, s->d_pb.p + sur_woff(l), N, B, N, K, 1, e)) return r;
        } else {
            e.C = s->d_Yo.p;
            e.ldc = N;
            if (int r = gemm<false, false, kEpiBias>(st, in, K, s->d_pb.p + sur_woff(l), N, B, N, K, 1, e)) return r;
        }
    }
    return 0;
}

// Backward pass from dY (s->d_dY, s->d_dYb) into the gradient s->d_g.
int sur_backward(ebem_sur* s, hipStream_t st, int B) {
    const float* cur_f = s->d_dY.p;
    const uint16_t* cur_b = s->d_dYb.p;
    int q = 0;
    for (int l = kSurLayers - 1; l >= 0; l--) {
        const uint16_t* in = l == 0 ? s->d_A0.p : s->d_A[l - 1].p;
        const int K = sur_K(l), N = sur_N(l);
        // dW_l = in^T cur (split over the batch), db_l = column sums of cur
        const int tiles = (K / 64) * (N / 64);
        int splits = std::max(1, std::min(B / 32, 256 / tiles));
        EpiArgs e{};
        e.C = s->d_part.p;
        e.ldc = N;
        e.cz = (int64_t)K * N;
        int ksplit = ((B / splits + 31) / 32) * 32;
        const int z = (B + ksplit - 1) / ksplit;
        if (int r = gemm<true, false, kEpiStore>(st, in, K, cur_b, N, K, N, B, splits, e)) return r;
        const int64_t kn = (int64_t)K * N;
        hipLaunchKernelGGL(k_sum_splits, dim3((unsigned)((kn + 255) / 256)), dim3(256), 0, st, s->d_part.p, z, kn,
                           s->d_g.p + sur_woff(l));
        hipLaunchKernelGGL(k_colsum, dim3(N / 64), dim3(256), 0, st, cur_f, B, N, s->d_g.p + sur_boff(l));
        HIPC(hipGetLastError());
        if (l > 0) {
            EpiArgs g{};
            g.C = s->d_dZ[q].p;
            g.ldc = K;
            g.Cb = s->d_dZb[q].p;
            g.ldcb = K;
            g.Zin = s->d_Z[l - 1].p;
            g.ldz = K;
            if (int r = gemm<false, true, kEpiGeluBwd>(st, cur_b, N, s->d_pb.p + sur_woff(l), N, B, K, N, 1, g)) return r;
            cur_f = s->d_dZ[q].p;
            cur_b = s->d_dZb[q].p;
            q ^= 1;
        }
    }
    return 0;
}

// Forward (and loss) of nb rows of samples idx (device), padded to B rows.
int sur_batch_loss(ebem_sur* s, hipStream_t st, const float* X, const float* Y, const int32_t* idx, int nb, int B,
                   bool backward) {
    hipLaunchKernelGGL(k_sur_gather, dim3((B + 3) / 4), dim3(64, 4), 0, st, X, Y, idx, nb, B, s->d_norm.p, s->d_A0.p,
                       s->d_T.p);
    HIPC(hipGetLastError());
    if (int r = sur_forward(s, st, B, backward)) return r;
    hipLaunchKernelGGL(k_sur_loss, dim3((unsigned)(((int64_t)B * kSurOutPad + 255) / 256)), dim3(256), 0, st, s->d_Yo.p,
                       s->d_T.p, nb, B, backward ? s->d_dY.p : nullptr, backward ? s->d_dYb.p : nullptr,
                       s->d_loss.p);
    HIPC(hipGetLastError());
    if (backward) return sur_backward(s, st, B);
    return 0;
}

int sur_alloc_params(ebem_sur* s) {
    HIPC(s->d_p.alloc(kSurParams));
    HIPC(s->d_pb.alloc(kSurParams));
    HIPC(s->d_g.alloc(kSurParams));
    HIPC(s->d_m.alloc(kSurParams));
    HIPC(s->d_v.alloc(kSurParams));
    HIPC(s->d_norm.alloc(kSurNorm));
    HIPC(s->d_part.alloc((size_t)256 * 64 * 64));  // largest splits * K * N of the weight gradients
    return 0;
}

// Surrogate prediction of the items perm[0..n) of a chunk into their level
// slots of W (k_sur_feats -> MLP -> k_sur_scatter), in batches.
int sur_assemble_items(ebem_handle* h, ebem_sur* s, int r0, int64_t n, const int32_t* perm, const int32_t* item_pair,
                       const int64_t* off, const int16_t* lo, const int16_t* hi, double* W) {
    const int Bi = 65536;
    if (int r = sur_workspace(s, Bi, false)) return r;
    HIPC(s->d_slot.alloc(Bi));
    hipStream_t st = h->stream;
    for (int64_t k0 = 0; k0 < n; k0 += Bi) {
        const int nb = (int)std::min<int64_t>(Bi, n - k0);
        const int B = ((nb + 63) / 64) * 64;
        hipLaunchKernelGGL(k_sur_feats, dim3((B + 255) / 256), dim3(256), 0, st, h->d_tri.p, h->d_verts.p, h->N, r0,
                           (int64_t)nb, B, perm + k0, item_pair, off, lo, hi, h->dt, h->cP, h->cS, h->hmax,
                           s->d_norm.p, s->d_A0.p, s->d_slot.p);
        HIPC(hipGetLastError());
        if (int r = sur_forward(s, st, B, false)) return r;
        hipLaunchKernelGGL(k_sur_scatter, dim3((B + 255) / 256), dim3(256), 0, st, h->d_tri.p, h->N, r0, B,
                           s->d_slot.p, item_pair, s->d_Yo.p, s->d_norm.p, h->rho, h->hmax, W);
        HIPC(hipGetLastError());
    }
    return 0;
}

}  // namespace

namespace {

int check_handle(ebem_handle* h) {
    if (!h) return fail(EBEM_EINVAL, "null handle");
    return 0;
}

int upload_geometry(ebem_handle* h) {
    std::vector<double> gh;
    build_gh(h->X.data(), h->T.data(), h->tri, h->nGH, gh);
    std::vector<double> verts((size_t)h->N * 9);
    h->hmax = 0;
    for (int t = 0; t < h->N; t++) {
        for (int v = 0; v < 3; v++)
            for (int k = 0; k < 3; k++) verts[(size_t)t * 9 + 3 * v + k] = h->X[3 * (size_t)h->T[3 * t + v] + k];
        h->hmax = std::max(h->hmax, h->tri[t].maxL);
    }
    HIPC(h->d_verts.alloc(verts.size()));
    HIPC(hipMemcpyAsync(h->d_verts.p, verts.data(), sizeof(double) * verts.size(), hipMemcpyHostToDevice, h->stream));
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

// FP64 per-level items perm[0..n) of a chunk of nchunk items by outer node:
// keys, counting sort of the node tasks, one node per lane, reduction per
// item (k_node_* in ebem_kernels.hpp).  Task ids it * NGH + q (it < nchunk)
// must stay below 2^31 (the chunk budget keeps it so).  With static-node
// reuse on, a failed check of the reused values (k_node_keys' *bad) makes
// the chunk be keyed again without reuse.
template <int NGH, typename R>
int launch_nodes(ebem_handle* h, AsmParams P, int r0, int64_t npairs, int64_t nchunk, int64_t n,
                 const int32_t* item_pair, const int64_t* off, const int16_t* lo, const int16_t* hi,
                 const uint8_t* ikey, int klo, int khi, const uint8_t* sflag, const int32_t* perm, double* W) {
    if (n <= 0) return 0;
    const int64_t ntask = nchunk * NGH;
    if (ntask > INT32_MAX) return fail(EBEM_EINVAL, "too many node tasks in one chunk");
    auto& nw = h->nw;
    const int G = h->CUs * 8;
    HIPC(nw.key.grow((size_t)ntask));
    HIPC(nw.perm.grow((size_t)n * NGH));
    HIPC(nw.part.grow((size_t)ntask * 8));
    HIPC(nw.hist.grow((size_t)(kNodeKeys + 1) * G));
    HIPC(nw.cur.grow((size_t)kNodeKeys * G));
    HIPC(nw.bad.grow(1));
    const int64_t per = (npairs + G - 1) / G;  // pairs per sort block
    for (int reuse = h->node_reuse;; reuse = 0) {
        HIPC(hipMemsetAsync(nw.bad.p, 0, sizeof(unsigned int), h->stream));
        hipLaunchKernelGGL(k_node_keys<NGH>, dim3(G), dim3(256), 0, h->stream, h->d_tri.p, h->d_gh.p, P, r0, npairs,
                           per, off, lo, hi, ikey, klo, khi, sflag, reuse, nw.key.p, nw.hist.p, nw.bad.p);
        HIPC(hipGetLastError());
        std::vector<unsigned int> hist((size_t)(kNodeKeys + 1) * G);
        std::vector<unsigned long long> cur((size_t)kNodeKeys * G);
        unsigned int bad = 0;
        HIPC(hipMemcpyAsync(hist.data(), nw.hist.p, sizeof(unsigned int) * hist.size(), hipMemcpyDeviceToHost,
                            h->stream));
        HIPC(hipMemcpyAsync(&bad, nw.bad.p, sizeof bad, hipMemcpyDeviceToHost, h->stream));
        HIPC(hipStreamSynchronize(h->stream));
        if (reuse && bad) {
            h->node_stats[2]++;
            continue;
        }
        unsigned long long run = 0;
        for (size_t k = 0; k < cur.size(); k++) {  // key-major, block-minor
            cur[k] = run;
            run += hist[k];
        }
        int64_t reused = 0;
        for (int b = 0; b < G; b++) reused += hist[(size_t)kNodeKeys * G + b];
        HIPC(hipMemcpyAsync(nw.cur.p, cur.data(), sizeof(unsigned long long) * cur.size(), hipMemcpyHostToDevice,
                            h->stream));
        // block b of the scatter takes the tasks of the items block b keyed
        hipLaunchKernelGGL(k_node_scatter<NGH>, dim3(G), dim3(256), 0, h->stream, npairs, per, off, lo, hi, ikey,
                           klo, khi, nw.key.p, nw.cur.p, nw.perm.p);
        HIPC(hipGetLastError());
        if (run > 0) {
            int64_t blocks = std::min<int64_t>(((int64_t)run + 255) / 256, (int64_t)h->CUs * 16);
            hipLaunchKernelGGL((k_node_eval<NGH, R>), dim3((unsigned)std::max<int64_t>(blocks, 1)), dim3(256), 0,
                               h->stream, h->d_tri.p, h->d_gh.p, P, r0, (int64_t)run, item_pair, off, lo,
                               nw.perm.p, nw.part.p);
            HIPC(hipGetLastError());
        }
        hipLaunchKernelGGL(k_node_reduce<NGH>, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, h->stream,
                           h->d_tri.p, P, n, item_pair, perm, nw.key.p, nw.part.p, W);
        HIPC(hipGetLastError());
        h->node_stats[0] += (int64_t)run;
        h->node_stats[1] += reused;
        return 0;
    }
}

// The per-level items perm[0..nitems) of a chunk of nchunk items, in key
// order: far-field items (mixed precision, nfar) by outer node in FP32,
// coplanar pairs (ncop) by k_levels in FP64, the rest by outer node in FP64.
// Unsorted items (perm null) or EBEM_NODE_TASKS=0: all by k_levels.
template <int NGH>
int launch_per_level(ebem_handle* h, AsmParams P, int r0, int64_t npairs, int64_t nchunk, int64_t nitems,
                     int64_t nfar, int64_t ncop, const int32_t* item_pair, const int64_t* off, const int16_t* lo,
                     const int16_t* hi, const uint8_t* ikey, const uint8_t* sflag, const int32_t* perm, double* W) {
    if (!(perm && h->node_tasks)) {
        launch_levels_split<NGH>(h, P, r0, nitems, nfar, item_pair, off, lo, perm, W);
        return 0;
    }
    if (nfar > 0) {
        if (int e = launch_nodes<NGH, float>(h, P, r0, npairs, nchunk, nfar, item_pair, off, lo, hi, ikey, 0,
                                             kFrontKeys, sflag, perm, W))
            return e;
    }
    launch_levels<NGH, double>(h, P, r0, ncop, item_pair, off, lo, perm + nfar, W);
    return launch_nodes<NGH, double>(h, P, r0, npairs, nchunk, nitems - nfar - ncop, item_pair, off, lo, hi, ikey,
                                     kNearKeys, kNearKeys + kFrontKeys, sflag, perm + nfar + ncop, W);
}

int assemble(ebem_handle* h, int precision) {
    if (!h->have_mesh || !h->have_problem) return fail(EBEM_EINVAL, "mesh and problem must be set before ebem_assemble");
    if (precision < 0 || precision > (EBEM_PREC_MIXED | EBEM_PREC_SURROGATE | EBEM_PREC_GHC))
        return fail(EBEM_EINVAL, "unknown precision %d", precision);
    const bool ghc = (precision & EBEM_PREC_GHC) != 0;
    if (ghc && precision != EBEM_PREC_GHC)
        return fail(EBEM_EINVAL, "EBEM_PREC_GHC is an FP64 form of its own (no mixed precision or surrogate)");
    if (ghc && !h->ghc_nSub) return fail(EBEM_EINVAL, "EBEM_PREC_GHC needs ebem_set_inner_rule first");
    const bool mixed = (precision & EBEM_PREC_MIXED) != 0;
    ebem_sur* sur = (precision & EBEM_PREC_SURROGATE) ? h->sur : nullptr;
    if ((precision & EBEM_PREC_SURROGATE) && !sur)
        return fail(EBEM_EINVAL, "EBEM_PREC_SURROGATE needs a surrogate attached with ebem_attach_surrogate");
    std::unique_lock<std::mutex> slk;
    if (sur) {
        slk = std::unique_lock<std::mutex>(sur->lock);
        if (!sur->ready) return fail(EBEM_EINVAL, "the attached surrogate has no parameters (init, train or set them)");
    }
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
        P.far2 = (mixed && h->far_factor > 0.0) ? (h->far_factor * hmax) * (h->far_factor * hmax) : -1.0;
    }
    // kP ranges need the key-sorted item order
    const bool kp = !ghc && h->sort_items && h->kp_split;
    P.kp = kp ? 1 : 0;
    P.sur_project = h->sur_project ? 1 : 0;
    for (int k = 0; k < 5; k++) h->kp_stats[k] = 0;
    for (int k = 0; k < 3; k++) h->node_stats[k] = 0;
    const double near2 = (3.0 * h->hmax) * (3.0 * h->hmax);  // surrogate pairs: centroid distance <= 3h
    Timer tall, tlev, tkp, tbf;
    double lev_ms = 0, kp_ms = 0, bf_ms = 0;
    tall.start(h->stream);
    const int64_t np = (int64_t)N * N;
    h->far_items = 0;
    h->sur_items = 0;
    h->sur_pairs_blocks = 0;
    h->sur_ms = 0;
    h->precision = precision;
    auto& d_flag = h->aw.flag;
    auto& d_nsb = h->aw.nsb;
    if (sur) {
        HIPC(d_nsb.grow(1));
        HIPC(hipMemsetAsync(d_nsb.p, 0, sizeof(unsigned long long), h->stream));
    }
    Timer tsur;
    auto &d_lo = h->aw.lo, &d_hi = h->aw.hi, &d_ka = h->aw.ka, &d_kb = h->aw.kb, &d_sa = h->aw.sa,
         &d_sb = h->aw.sb;
    HIPC(d_lo.grow(np));
    HIPC(d_hi.grow(np));
    if (kp) {
        HIPC(d_ka.grow(np));
        HIPC(d_kb.grow(np));
        HIPC(d_sa.grow(np));
        HIPC(d_sb.grow(np));
    }
    auto& d_culled = h->aw.culled;
    HIPC(d_culled.grow(1));
    HIPC(hipMemsetAsync(d_culled.p, 0, sizeof(unsigned long long), h->stream));
    auto& d_nfb = h->aw.nfb;
    HIPC(d_nfb.grow(1));
    HIPC(hipMemsetAsync(d_nfb.p, 0, sizeof(unsigned int), h->stream));
    {
        int64_t blocks = std::min<int64_t>((np + 255) / 256, (int64_t)h->CUs * 64);
        hipLaunchKernelGGL(k_windows, dim3((unsigned)blocks), dim3(256), 0, h->stream, h->d_tri.p, h->d_gh.p,
                           h->d_verts.p, P, 0, N, d_lo.p, d_hi.p, d_ka.p, d_kb.p, d_sa.p, d_sb.p, d_culled.p);
        HIPC(hipGetLastError());
    }
    unsigned long long culled = 0;
    HIPC(hipMemcpyAsync(&culled, d_culled.p, sizeof culled, hipMemcpyDeviceToHost, h->stream));
    auto &d_cnt = h->aw.cnt, &d_items = h->aw.items;
    HIPC(d_cnt.grow((size_t)N * L));
    HIPC(d_items.grow(N));
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
    h->pair_lags = (int64_t)culled;  // the reference's culled pair.lags; nnzb leaves out exact zeros
    h->items = 0;
    for (int i = 0; i < N; i++) h->items += rowitems[i];
    HIPC(h->d_seg.grow(h->seg.size()));
    HIPC(hipMemcpyAsync(h->d_seg.p, h->seg.data(), sizeof(int64_t) * h->seg.size(), hipMemcpyHostToDevice, h->stream));
    HIPC(h->d_val.grow((size_t)6 * std::max<int64_t>(h->nnzb, 1)));
    HIPC(h->d_col.grow((size_t)std::max<int64_t>(h->nnzb, 1)));
    Store S = h->store();

    // row chunks bounded by the W scratch budget
    size_t freeb = 0, totb = 0;
    HIPC(hipMemGetInfo(&freeb, &totb));
    // per item: W, pair index, permutation, key; with node tasks also a
    // 64-byte slot, a key and a task index per outer node
    const double item_bytes = 6 * 8 + 4 + 4 + 1 + (h->node_tasks ? h->nGH * (64.0 + 1 + 4) : 0.0);
    int64_t budget_items = (int64_t)std::min<double>(256.0e6, 0.35 * (double)freeb / item_bytes);
    if (h->node_tasks) budget_items = std::min<int64_t>(budget_items, (int64_t)INT32_MAX / h->nGH);
    if (budget_items < 1) return fail(EBEM_ENOMEM, "not enough device memory for assembly scratch");
    int64_t maxrow_items = *std::max_element(rowitems.begin(), rowitems.end());
    if (maxrow_items > budget_items) budget_items = maxrow_items;
    auto& d_W = h->aw.W;
    auto& d_ip = h->aw.ip;
    auto &d_off = h->aw.off, &d_rowbase = h->aw.rowbase;
    int64_t chunk_cap = std::min<int64_t>(budget_items, std::max<int64_t>(h->items, 1));
    HIPC(d_W.grow((size_t)chunk_cap * 6));
    HIPC(d_ip.grow((size_t)chunk_cap));
    auto& d_perm = h->aw.perm;
    auto& d_key = h->aw.key;
    auto& d_hist = h->aw.hist;
    auto& d_cur = h->aw.cur;
    const int G = h->CUs * 8;  // sort blocks
    HIPC(d_perm.grow((size_t)chunk_cap));
    HIPC(d_key.grow((size_t)chunk_cap));
    HIPC(d_hist.grow((size_t)kNumKeys * G));
    HIPC(d_cur.grow((size_t)kNumKeys * G));
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
        HIPC(d_rowbase.grow(rows));
        HIPC(hipMemcpyAsync(d_rowbase.p, rb.data(), sizeof(int64_t) * rows, hipMemcpyHostToDevice, h->stream));
        HIPC(d_off.grow((size_t)rows * N));
        const int16_t* lo = d_lo.p + (int64_t)r0 * N;
        const int16_t* hi = d_hi.p + (int64_t)r0 * N;
        hipLaunchKernelGGL(k_pair_items, dim3(rows), dim3(256), 0, h->stream, lo, hi, N, d_rowbase.p, d_off.p, d_ip.p);
        HIPC(hipGetLastError());
        const uint8_t* sflag = nullptr;
        if (sur) {
            HIPC(d_flag.grow((size_t)rows * N));
            const int64_t npr = (int64_t)rows * N;
            hipLaunchKernelGGL(k_sur_flag, dim3((unsigned)((npr + 255) / 256)), dim3(256), 0, h->stream, h->d_tri.p,
                               h->d_verts.p, N, L, P.sur_project, r0, rows, lo, hi, h->dt, h->cP, h->cS, h->hmax,
                               near2,
                               sur->d_norm.p + 2 * kSurIn + 2 * kSurOut, d_flag.p, d_nsb.p);
            HIPC(hipGetLastError());
            sflag = d_flag.p;
        }
        if (acc > 0) {
            // bucket items by front-state key (counting sort) for convergent wavefronts
            const int64_t per = (acc + G - 1) / G;
            const int16_t* ka = kp ? d_ka.p + (int64_t)r0 * N : nullptr;
            const int16_t* kb = kp ? d_kb.p + (int64_t)r0 * N : nullptr;
            const int16_t* sa = kp ? d_sa.p + (int64_t)r0 * N : nullptr;
            const int16_t* sb = kp ? d_sb.p + (int64_t)r0 * N : nullptr;
            hipLaunchKernelGGL(k_item_keys, dim3(G), dim3(256), 0, h->stream, h->d_tri.p, P, r0, acc, per, d_ip.p,
                               d_off.p, lo, hi, ka, kb, sa, sb, sflag, d_key.p, d_hist.p);
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
            const int32_t* perm = (h->sort_items || mixed || sur) ? d_perm.p : nullptr;
            int64_t nfar = 0, ncop = 0, nsur = 0, nkp = 0, nkrest = 0, nbf = 0, nbrest = 0;
            int64_t nkpfar = 0, nkfrest = 0, nbffar = 0, nbfrest = 0;  // far-field ranges (FP32)
            for (size_t k = 0; k < (size_t)kFrontKeys * G; k++) nfar += hist[k];
            for (size_t k = (size_t)kCopKeys * G; k < (size_t)kNearKeys * G; k++) ncop += hist[k];
            for (int b = 0; b < G; b++) {
                nsur += hist[(size_t)kSurKey * G + b];
                nkpfar += hist[(size_t)kKpFarKey * G + b];
                nkfrest += hist[(size_t)kKpFarRestKey * G + b];
                nkp += hist[(size_t)kKpFarKey * G + b] + hist[(size_t)kKpKey * G + b];
                nkrest += hist[(size_t)kKpFarRestKey * G + b] + hist[(size_t)kKpRestKey * G + b];
                for (int c = 0; c < kBfHeadKeys; c++) {
                    nbffar += hist[(size_t)(kBfFarKey + c) * G + b];
                    nbf += hist[(size_t)(kBfFarKey + c) * G + b] + hist[(size_t)(kBfKey + c) * G + b];
                }
                nbfrest += hist[(size_t)kBfFarRestKey * G + b];
                nbrest += hist[(size_t)kBfFarRestKey * G + b] + hist[(size_t)kBfRestKey * G + b];
            }
            h->far_items += nfar + nkpfar + nkfrest + nbffar + nbfrest;
            h->sur_items += nsur;
            h->kp_stats[0] += nkp;
            h->kp_stats[1] += nkp + nkrest;
            h->kp_stats[2] += nbf;
            h->kp_stats[3] += nbf + nbrest;
            // perm: [per-level items | kP range heads (far ones first) | other kP levels |
            //        between-front range heads (far ones first) | their other levels | surrogate items]
            const int64_t nlev = acc - nsur - nkp - nkrest - nbf - nbrest;
            tlev.start(h->stream);
            if (ghc) {
                GhcRules R{h->d_ghc_in.p, h->d_ghc_dg.p, h->ghc_nSub * h->ghc_nSing, h->ghc_nDiag};
                const size_t lds = sizeof(double) * 4 * (size_t)(R.nInt + R.nDiag);  // <= 71.7 KB
                if (lds > 65536) {
                    HIPC(hipFuncSetAttribute((const void*)k_ghc_blocks<3>, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
                    HIPC(hipFuncSetAttribute((const void*)k_ghc_blocks<7>, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
                    HIPC(hipFuncSetAttribute((const void*)k_ghc_blocks<12>, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
                    HIPC(hipFuncSetAttribute((const void*)k_ghc_blocks<19>, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
                }
                const int64_t blocks = std::max<int64_t>(1, std::min<int64_t>((acc + 3) / 4, (int64_t)h->CUs * 32));
#define GHC_LAUNCH(NG)                                                                                              \
    hipLaunchKernelGGL(k_ghc_blocks<NG>, dim3((unsigned)blocks), dim3(256), lds, h->stream, h->d_tri.p, h->d_gh.p, \
                       h->d_verts.p, P, R, r0, acc, d_ip.p, d_off.p, lo, d_W.p)
                switch (h->nGH) {
                    case 3: GHC_LAUNCH(3); break;
                    case 7: GHC_LAUNCH(7); break;
                    case 12: GHC_LAUNCH(12); break;
                    default: GHC_LAUNCH(19); break;
                }
#undef GHC_LAUNCH
            } else {
                int e = 0;
#define PL_LAUNCH(NG)                                                                                     \
    launch_per_level<NG>(h, P, r0, (int64_t)rows * N, acc, nlev, nfar, ncop, d_ip.p, d_off.p, lo, hi, d_key.p, \
                         sflag, perm, d_W.p)
                switch (h->nGH) {
                    case 3: e = PL_LAUNCH(3); break;
                    case 7: e = PL_LAUNCH(7); break;
                    case 12: e = PL_LAUNCH(12); break;
                    default: e = PL_LAUNCH(19); break;
                }
#undef PL_LAUNCH
                if (e) return e;
            }
            HIPC(hipGetLastError());
            tlev.stop(h->stream);
            lev_ms += tlev.ms();
            if (nkp > 0) {
                tkp.start(h->stream);
                // far-field ranges (mixed precision) in FP32, then the rest in FP64
                for (int part = 0; part < 2; part++) {
                    const int64_t n = part ? nkp - nkpfar : nkpfar;
                    if (n == 0) continue;
                    const int32_t* pk = d_perm.p + nlev + (part ? nkpfar : 0);
                    const int64_t blocks = std::max<int64_t>(1, std::min<int64_t>((n + 255) / 256, (int64_t)h->CUs * 16));
#define KP_LAUNCH(NG, R)                                                                                              \
    hipLaunchKernelGGL((k_levels_kp<NG, R>), dim3((unsigned)blocks), dim3(256), 0, h->stream, h->d_tri.p, h->d_gh.p, \
                       P, r0, n, d_ip.p, ka, kb, pk, d_W.p, d_nfb.p)
#define KP_LAUNCH_R(R)                    \
    switch (h->nGH) {                     \
        case 3: KP_LAUNCH(3, R); break;   \
        case 7: KP_LAUNCH(7, R); break;   \
        case 12: KP_LAUNCH(12, R); break; \
        default: KP_LAUNCH(19, R); break; \
    }
                    if (part) {
                        KP_LAUNCH_R(double);
                    } else {
                        KP_LAUNCH_R(float);
                    }
#undef KP_LAUNCH_R
#undef KP_LAUNCH
                    HIPC(hipGetLastError());
                }
                tkp.stop(h->stream);
                kp_ms += tkp.ms();
            }
            if (nbf > 0) {
                tbf.start(h->stream);
                for (int part = 0; part < 2; part++) {
                    const int64_t n = part ? nbf - nbffar : nbffar;
                    if (n == 0) continue;
                    const int32_t* pb = d_perm.p + nlev + nkp + nkrest + (part ? nbffar : 0);
                    const int64_t blocks = std::max<int64_t>(1, std::min<int64_t>((n + 255) / 256, (int64_t)h->CUs * 16));
#define BF_LAUNCH(NG, R)                                                                                              \
    hipLaunchKernelGGL((k_levels_bf<NG, R>), dim3((unsigned)blocks), dim3(256), 0, h->stream, h->d_tri.p, h->d_gh.p, \
                       P, r0, n, d_ip.p, sa, sb, pb, d_W.p, d_nfb.p)
#define BF_LAUNCH_R(R)                    \
    switch (h->nGH) {                     \
        case 3: BF_LAUNCH(3, R); break;   \
        case 7: BF_LAUNCH(7, R); break;   \
        case 12: BF_LAUNCH(12, R); break; \
        default: BF_LAUNCH(19, R); break; \
    }
                    if (part) {
                        BF_LAUNCH_R(double);
                    } else {
                        BF_LAUNCH_R(float);
                    }
#undef BF_LAUNCH_R
#undef BF_LAUNCH
                    HIPC(hipGetLastError());
                }
                tbf.stop(h->stream);
                bf_ms += tbf.ms();
            }
            if (nsur > 0) {
                tsur.start(h->stream);
                if (int r = sur_assemble_items(h, sur, r0, nsur, d_perm.p + (acc - nsur), d_ip.p, d_off.p, lo, hi, d_W.p))
                    return r;
                tsur.stop(h->stream);
                h->sur_ms += tsur.ms();
            }
        }
        hipLaunchKernelGGL(k_emit, dim3(rows), dim3(64), L * sizeof(int64_t), h->stream, lo, hi, d_off.p, d_W.p, sflag,
                           N, L, r0, inv4pirho, 4 * kPi * h->rho, ghc ? 1 : 0, P.sur_project, S);
        HIPC(hipGetLastError());
        r0 = r1;
    }
    tall.stop(h->stream);
    h->tim[1] = tall.ms();
    h->tim[0] = lev_ms;
    for (int k = 0; k < 3; k++) h->tim[16 + k] = (double)h->node_stats[k];
    h->tim[8] = kp_ms;
    h->tim[10] = bf_ms;
    {
        unsigned int nfb = 0;
        HIPC(hipMemcpyAsync(&nfb, d_nfb.p, sizeof nfb, hipMemcpyDeviceToHost, h->stream));
        HIPC(hipStreamSynchronize(h->stream));
        h->kp_stats[4] = nfb;
    }
    if (sur) {
        unsigned long long nb = 0;
        HIPC(hipMemcpyAsync(&nb, d_nsb.p, sizeof nb, hipMemcpyDeviceToHost, h->stream));
        HIPC(hipStreamSynchronize(h->stream));
        h->sur_pairs_blocks = (int64_t)nb;
    }
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

// Inverse of E_0 (rocSOLVER getrf + getrs on the identity) into the row-major d_Ainv.
// E_0 is dominated by its diagonal (self-interaction) blocks, so LU without
// pivoting is tried first (its panels need a fraction of the kernel launches
// of the pivoted factorisation: 21 vs 48 ms at 3840 unknowns); the inverse is
// accepted when E_0 (A_inv v) reproduces a fixed random v to 1e-10 relative,
// otherwise the factorisation is redone with partial pivoting.  Either way the
// result agrees with a pivoted LU solve to rounding.
int factor_e0(ebem_handle* h) {
    const int n3 = 3 * h->N;
    const int L = h->store_L;
    HIPC(h->d_A.alloc((size_t)n3 * n3));
    HIPC(h->d_Ainv.alloc((size_t)n3 * n3));
    HIPC(h->d_ipiv.alloc(n3));
    HIPC(h->d_info.alloc(1));
    if (!h->blas) {
        if (int e = rocb(rocblas_create_handle(&h->blas), "rocblas_create_handle")) return e;
    }
    if (int e = rocb(rocblas_set_stream(h->blas, h->stream), "rocblas_set_stream")) return e;
    dim3 grid((n3 + 31) / 32, (n3 + 31) / 32);
    h->e0_pivoted = false;
    // EBEM_E0_PIVOT=1 skips the unpivoted attempt (for comparisons)
    const char* force = std::getenv("EBEM_E0_PIVOT");
    for (int attempt = (force && std::strcmp(force, "1") == 0) ? 1 : 0; attempt < 2; attempt++) {
        const bool pivot = attempt == 1;
        HIPC(hipMemsetAsync(h->d_A.p, 0, sizeof(double) * (size_t)n3 * n3, h->stream));
        hipLaunchKernelGGL(k_e0_dense, dim3(4, h->N), dim3(64), 0, h->stream, h->store(), h->N, L, h->d_A.p);
        HIPC(hipGetLastError());
        rocblas_int info = 0;
        if (pivot) {
            if (int e = rocb(rocsolver_dgetrf(h->blas, n3, n3, h->d_A.p, n3, h->d_ipiv.p, h->d_info.p),
                             "rocsolver_dgetrf"))
                return e;
        } else {
            if (int e = rocb(rocsolver_dgetrf_npvt(h->blas, n3, n3, h->d_A.p, n3, h->d_info.p), "rocsolver_dgetrf_npvt"))
                return e;
            std::vector<rocblas_int> iota(n3);
            for (int k = 0; k < n3; k++) iota[k] = k + 1;
            HIPC(hipMemcpyAsync(h->d_ipiv.p, iota.data(), sizeof(rocblas_int) * n3, hipMemcpyHostToDevice, h->stream));
        }
        HIPC(hipMemcpyAsync(&info, h->d_info.p, sizeof(info), hipMemcpyDeviceToHost, h->stream));
        HIPC(hipStreamSynchronize(h->stream));
        if (info != 0) {
            if (!pivot) continue;
            return fail(EBEM_ESINGULAR, "E_0 is singular (zero pivot at %d)", (int)info);
        }
        // inverse = LU solve of the identity (two triangular solves with n
        // right-hand sides, rocBLAS trsm) into d_Ainv (column-major), then
        // transposed to row-major through d_A
        HIPC(hipMemsetAsync(h->d_Ainv.p, 0, sizeof(double) * (size_t)n3 * n3, h->stream));
        hipLaunchKernelGGL(k_set_diag, dim3((n3 + 255) / 256), dim3(256), 0, h->stream, h->d_Ainv.p, n3, 1.0);
        HIPC(hipGetLastError());
        if (int e = rocb(rocsolver_dgetrs(h->blas, rocblas_operation_none, n3, n3, h->d_A.p, n3, h->d_ipiv.p,
                                          h->d_Ainv.p, n3),
                         "rocsolver_dgetrs"))
            return e;
        hipLaunchKernelGGL(k_transpose, grid, dim3(32, 8), 0, h->stream, h->d_Ainv.p, h->d_A.p, n3);
        HIPC(hipGetLastError());
        std::swap(h->d_A.p, h->d_Ainv.p);
        if (pivot) {
            h->e0_pivoted = true;
            return 0;
        }
        // check: E_0 (A_inv v) = v, E_0 rebuilt row-major in d_A (its transpose)
        std::vector<double> v(n3), z(n3);
        uint64_t st = 0x243f6a8885a308d3ull;
        for (int k = 0; k < n3; k++) {
            st = st * 6364136223846793005ull + 1442695040888963407ull;
            v[k] = (double)(st >> 11) / 9007199254740992.0 - 0.5;
        }
        DevBuf<double> dv;
        HIPC(dv.alloc(2 * (size_t)n3));
        HIPC(hipMemcpyAsync(dv.p, v.data(), sizeof(double) * n3, hipMemcpyHostToDevice, h->stream));
        hipLaunchKernelGGL(k_gemv_rows, dim3((n3 + 3) / 4), dim3(256), 0, h->stream, h->d_Ainv.p, dv.p, dv.p + n3, n3,
                           nullptr, nullptr, 0, 0);
        HIPC(hipMemsetAsync(h->d_A.p, 0, sizeof(double) * (size_t)n3 * n3, h->stream));
        hipLaunchKernelGGL(k_e0_dense, dim3(4, h->N), dim3(64), 0, h->stream, h->store(), h->N, L, h->d_A.p);
        // d_A holds E_0 column-major: E_0 y by rocBLAS dgemv
        HIPC(hipGetLastError());
        {
            const double one = 1.0, zero = 0.0;
            if (int e = rocb(rocblas_dgemv(h->blas, rocblas_operation_none, n3, n3, &one, h->d_A.p, n3, dv.p + n3, 1,
                                           &zero, dv.p, 1),
                             "rocblas_dgemv"))
                return e;
        }
        HIPC(hipMemcpyAsync(z.data(), dv.p, sizeof(double) * n3, hipMemcpyDeviceToHost, h->stream));
        HIPC(hipStreamSynchronize(h->stream));
        double num = 0, den = 0;
        for (int k = 0; k < n3; k++) {
            num += (z[k] - v[k]) * (z[k] - v[k]);
            den += v[k] * v[k];
        }
        if (std::isfinite(num) && std::sqrt(num / den) < 1e-10) return 0;
    }
    return fail(EBEM_ESINGULAR, "E_0 factorisation failed");
}

void pp_release(ebem_handle* h);

int mot(ebem_handle* h, const double* beta, double* alpha_out) {
    if (!h->assembled) return fail(EBEM_EINVAL, "ebem_assemble must run before ebem_mot_solve");
    if ((h->precision & EBEM_PREC_SURROGATE) && !h->allow_sur_solve)
        return fail(EBEM_EINVAL,
                    "the store holds surrogate-predicted blocks, whose errors the time marching amplifies; "
                    "set EBEM_OPT_SURROGATE_SOLVE to solve on it anyway");
    const int n3 = 3 * h->N, nT = h->nT;
    if (beta) {
        HIPC(h->d_beta.alloc((size_t)n3 * nT));
        HIPC(hipMemcpyAsync(h->d_beta.p, beta, sizeof(double) * (size_t)n3 * nT, hipMemcpyHostToDevice, h->stream));
        h->have_beta = true;
    } else if (!h->have_beta) {
        return fail(EBEM_EINVAL, "no right-hand side: pass beta or call ebem_rhs first");
    }
    HIPC(h->d_alpha.alloc((size_t)n3 * nT));
    // triangle-major copies of the density and of its step differences for
    // the post-processing, each triangle's row led by one zero step
    HIPC(h->d_alphaT.alloc((size_t)n3 * (nT + 1)));
    HIPC(h->d_dalphaT.alloc((size_t)n3 * (nT + 1)));
    HIPC(hipMemsetAsync(h->d_alphaT.p, 0, sizeof(double) * h->d_alphaT.n, h->stream));
    HIPC(hipMemsetAsync(h->d_dalphaT.p, 0, sizeof(double) * h->d_dalphaT.n, h->stream));
    HIPC(h->d_r.alloc(n3));
    Timer tall, tf;
    tall.start(h->stream);
    tf.start(h->stream);
    if (int e = factor_e0(h)) return e;
    tf.stop(h->stream);
    h->tim[2] = tf.ms();  // waits for the factorisation
    if (h->prep.active) pp_release(h);  // a prepared post-processing may start beside the marching
    Store S = h->store();
    const int L = h->store_L;
    HIPC(h->d_part.alloc((size_t)kHistBlock * n3));
    // events: [start of history work, end of history work = start of apply] per step
    // (the far pass of a block is timed with its block's first step)
    std::vector<hipEvent_t>& ev = h->mot_ev;
    while (ev.size() < 3 * (size_t)nT) {
        hipEvent_t e;
        HIPC(hipEventCreate(&e));
        ev.push_back(e);
    }
    int launches = 0;
    for (int n0 = 0; n0 < nT; n0 += kHistBlock) {
        const int nb = std::min(kHistBlock, nT - n0);
        HIPC(hipEventRecord(ev[3 * n0], h->stream));
        if (n0 > 0) {
            hipLaunchKernelGGL(k_history_far, dim3(h->N), dim3(256), 0, h->stream, S, h->N, L, n0, nb,
                               h->d_alpha.p, h->d_part.p);
            launches++;
        }
        for (int b = 0; b < nb; b++) {
            const int n = n0 + b;
            if (b > 0) HIPC(hipEventRecord(ev[3 * n], h->stream));
            hipLaunchKernelGGL(k_history, dim3(h->N), dim3(256), 0, h->stream, S, h->N, L, n, std::min(b, L - 1),
                               h->d_beta.p, h->d_alpha.p, n0 > 0 ? h->d_part.p + (size_t)b * n3 : nullptr, h->d_r.p);
            launches++;
            HIPC(hipEventRecord(ev[3 * n + 1], h->stream));
            hipLaunchKernelGGL(k_gemv_rows, dim3((n3 + 3) / 4), dim3(256), 0, h->stream, h->d_Ainv.p, h->d_r.p,
                               h->d_alpha.p + (size_t)n * n3, n3, h->d_alphaT.p, h->d_dalphaT.p, nT, n);
            HIPC(hipEventRecord(ev[3 * n + 2], h->stream));
        }
    }
    HIPC(hipGetLastError());
    HIPC(hipStreamSynchronize(h->stream));
    double th = 0, ts = 0;
    for (int n = 0; n < nT; n++) {
        float a = 0, b = 0;
        HIPC(hipEventElapsedTime(&a, ev[3 * n], ev[3 * n + 1]));
        HIPC(hipEventElapsedTime(&b, ev[3 * n + 1], ev[3 * n + 2]));
        th += a;
        ts += b;
    }
    tall.stop(h->stream);
    h->tim[5] = tall.ms();
    h->tim[3] = th;
    h->tim[4] = ts;
    h->tim[7] = launches;
    if (alpha_out)
        HIPC(hipMemcpyAsync(alpha_out, h->d_alpha.p, sizeof(double) * (size_t)n3 * nT, hipMemcpyDeviceToHost, h->stream));
    HIPC(hipStreamSynchronize(h->stream));
    h->have_alpha = true;
    return 0;
}

// Points per chunk of the post-processing: as many as keep every chunk's
// (point, triangle, m) items below 2^31 even if every window held all nT + 1
// levels (at least 1024); fewer chunks, fewer launches and host round trips.
int pp_chunk_points(const ebem_handle* h, int npts) {
    const int64_t pmax = (int64_t)INT32_MAX / ((int64_t)h->N * (h->nT + 1));
    return (int)std::min<int64_t>(npts, std::max<int64_t>(1024, pmax));
}

// Post-processing: times grouped by delta = t - nhat dt (shared M_m), points
// chunked so that the M scratch stays within budget.  stage 0: the whole
// evaluation on the library's stream.  stage 1 (ebem_pp_prepare, a request of
// one time group and one chunk of points, analytic form): the part that does
// not depend on the density -- points, windows, items, keys, the M integrals
// -- on stream st, complete when it returns; u_out and density are not used.
// stage 2: the rest of such a request after stage 1 (density copies and the
// convolution) on the library's stream.
int pp_eval(ebem_handle* h, const double* density, int npts, const double* x, int nt, const double* t,
            double* u_out, bool numeric = false, int stage = 0, hipStream_t st = nullptr) {
#pragma clang fp contract(off)
    if (!h->have_mesh || !h->have_problem) return fail(EBEM_EINVAL, "mesh and problem must be set");
    if (npts <= 0 || nt <= 0 || !x || !t || (!u_out && stage != 1))
        return fail(EBEM_EINVAL, "bad post-processing arguments");
    if (stage != 1 && !density && !h->have_alpha)
        return fail(EBEM_EINVAL, "no density: pass one or run ebem_mot_solve first");
    if (numeric && !h->pp_nInt) return fail(EBEM_EINVAL, "numeric post-processing needs ebem_set_pp_rule first");
    hipStream_t s0 = stage == 1 ? st : h->stream;
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
    Timer tall, tppk, tppb;
    double ppk_ms = 0, ppb_ms = 0;
    tall.start(s0);
    auto &d_x = h->pw.x, &d_aT = h->pw.aT, &d_u = h->pw.u, &d_dens = h->pw.dens, &d_M = h->pw.M;
    auto& d_pout = h->pw.pout;
    // points in the order of a Morton curve over their bounding box, so that
    // the point groups of k_pp_conv are neighbours; pout maps back
    std::vector<int32_t> pout(npts);
    std::vector<double> xs((size_t)3 * npts);
    if (stage != 2) {
        double lo[3], hi[3];
        for (int c = 0; c < 3; c++) lo[c] = hi[c] = x[c];
        for (int p = 0; p < npts; p++)
            for (int c = 0; c < 3; c++) {
                lo[c] = std::min(lo[c], x[3 * p + c]);
                hi[c] = std::max(hi[c], x[3 * p + c]);
            }
        std::vector<uint32_t> code(npts);
        for (int p = 0; p < npts; p++) {
            uint32_t z = 0;
            for (int c = 0; c < 3; c++) {
                const double e = hi[c] - lo[c];
                const uint32_t qv = e > 0 ? (uint32_t)std::min(1023.0, std::floor((x[3 * p + c] - lo[c]) / e * 1024.0)) : 0u;
                for (int bit = 0; bit < 10; bit++) z |= ((qv >> bit) & 1u) << (3 * bit + c);
            }
            code[p] = z;
            pout[p] = p;
        }
        std::stable_sort(pout.begin(), pout.end(), [&](int32_t a, int32_t b) { return code[a] < code[b]; });
        for (int r = 0; r < npts; r++)
            for (int c = 0; c < 3; c++) xs[3 * (size_t)r + c] = x[3 * (size_t)pout[r] + c];
        HIPC(d_x.grow(3 * (size_t)npts));
        HIPC(hipMemcpyAsync(d_x.p, xs.data(), sizeof(double) * 3 * npts, hipMemcpyHostToDevice, s0));
        HIPC(d_pout.grow(npts));
        HIPC(hipMemcpyAsync(d_pout.p, pout.data(), sizeof(int32_t) * npts, hipMemcpyHostToDevice, s0));
    }
    const double* dens = h->d_alpha.p;
    if (density && stage != 1) {
        HIPC(d_dens.grow((size_t)3 * N * nT));
        HIPC(hipMemcpyAsync(d_dens.p, density, sizeof(double) * 3 * N * nT, hipMemcpyHostToDevice, h->stream));
        dens = d_dens.p;
    }
    // the density triangle-major: the MOT's copy, or a transpose of the one passed in
    const double* aT = h->d_alphaT.p;
    const double* dT = h->d_dalphaT.p;
    if (density && stage != 1) {
        const int64_t tot = (int64_t)3 * N * nT, row = (int64_t)3 * N * (nT + 1);
        HIPC(d_aT.grow((size_t)2 * row));
        HIPC(hipMemsetAsync(d_aT.p, 0, sizeof(double) * 2 * row, h->stream));
        hipLaunchKernelGGL(k_alpha_T, dim3((unsigned)((tot + 255) / 256)), dim3(256), 0, h->stream, dens, N, nT,
                           d_aT.p, d_aT.p + row);
        HIPC(hipGetLastError());
        aT = d_aT.p;
        dT = d_aT.p + row;
    }
    if (stage != 1) {
        HIPC(d_u.grow((size_t)3 * npts * nt));
        HIPC(hipMemsetAsync(d_u.p, 0, sizeof(double) * 3 * npts * nt, h->stream));
    }
    // group times by delta
    std::vector<int> order;
    for (int k = 0; k < nt; k++)
        if (nhat[k] >= 0) order.push_back(k);
    std::sort(order.begin(), order.end(), [&](int a, int b) { return delta[a] < delta[b]; });
    const int P0 = pp_chunk_points(h, npts);
    const int G = h->CUs * 8;  // counting-sort blocks
    // kP ranges of m through k_pp_M_kp (analytic form with key-sorted items)
    const bool kp = !numeric && h->sort_items && h->kp_split;
    auto &d_ms = h->pw.ms, &d_me = h->pw.me, &d_iq = h->pw.iq, &d_nh = h->pw.nh, &d_kidx = h->pw.kidx;
    auto &d_perm = h->pw.perm, &d_ka = h->pw.ka, &d_kb = h->pw.kb, &d_sa = h->pw.sa, &d_sb = h->pw.sb;
    auto& d_nfb = h->pw.nfb;
    HIPC(d_nfb.grow(1));
    if (stage != 2) HIPC(hipMemsetAsync(d_nfb.p, 0, sizeof(unsigned int), s0));
    auto &d_off = h->pw.off, &d_rb = h->pw.rb, &d_cnt = h->pw.cnt;
    auto& d_key = h->pw.key;
    auto& d_bh = h->pw.bh;
    auto& d_bo = h->pw.bo;
    HIPC(d_ms.grow((size_t)P0 * N));
    HIPC(d_me.grow((size_t)P0 * N));
    if (kp) {
        HIPC(d_ka.grow((size_t)P0 * N));
        HIPC(d_kb.grow((size_t)P0 * N));
        HIPC(d_sa.grow((size_t)P0 * N));
        HIPC(d_sb.grow((size_t)P0 * N));
    }
    if (stage != 2)
        for (int k = 0; k < 5; k++) h->pp_kp_stats[k] = 0;
    HIPC(d_off.grow((size_t)P0 * N));
    HIPC(d_rb.grow(P0));
    HIPC(d_cnt.grow(P0));
    HIPC(d_bh.grow((size_t)kNumKeys * G));
    HIPC(d_bo.grow((size_t)kNumKeys * G));
    std::vector<unsigned int> bh((size_t)kNumKeys * G);
    std::vector<unsigned long long> bo((size_t)kNumKeys * G);
    std::vector<int64_t> cnt(P0), rb(P0);
    size_t g0 = 0;
    while (g0 < order.size()) {
        size_t g1 = g0 + 1;
        while (g1 < order.size() && delta[order[g1]] - delta[order[g0]] <= 1e-12 * dt) g1++;
        // a group's times by step, so that k_pp_conv's lanes read consecutive
        // density steps
        std::sort(order.begin() + g0, order.begin() + g1, [&](int a, int b) { return nhat[a] < nhat[b]; });
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
        P.kp = kp ? 1 : 0;
        if (stage == 1 && (g1 != order.size() || P0 < npts || numeric))
            return fail(EBEM_EINVAL, "a prepared post-processing request must be one time group and one chunk");
        if (stage != 2) {
            HIPC(d_nh.grow(nh.size()));
            HIPC(d_kidx.grow(kid.size()));
            HIPC(hipMemcpyAsync(d_nh.p, nh.data(), sizeof(int32_t) * nh.size(), hipMemcpyHostToDevice, s0));
            HIPC(hipMemcpyAsync(d_kidx.p, kid.data(), sizeof(int32_t) * kid.size(), hipMemcpyHostToDevice, s0));
        }
        for (int p0 = 0; p0 < npts; p0 += P0) {
            const int np = std::min(P0, npts - p0);
            if (stage == 2) {  // the prepared items and M of this (only) chunk
                if (h->pw.nitems > 0)
                    hipLaunchKernelGGL(k_pp_conv, dim3((np + kConvGroup - 1) / kConvGroup), dim3(64 * kConvSlices), 0,
                                       h->stream, d_M.p, h->pw.off.p, h->pw.ms.p, h->pw.me.p, aT, dT, N, nT, p0, np,
                                       d_pout.p, d_nh.p, d_kidx.p, (int)nh.size(), npts, d_u.p);
                HIPC(hipGetLastError());
                continue;
            }
            int64_t tot = (int64_t)np * N;
            int64_t blocks = std::min<int64_t>((tot + 255) / 256, (int64_t)h->CUs * 64);
            hipLaunchKernelGGL(k_pp_windows, dim3((unsigned)blocks), dim3(256), 0, s0, h->d_tri.p,
                               h->d_verts.p, d_x.p, p0, np, P, d_ms.p, d_me.p, d_ka.p, d_kb.p, d_sa.p, d_sb.p);
            HIPC(hipGetLastError());
            hipLaunchKernelGGL(k_pp_counts, dim3(np), dim3(256), 0, s0, d_ms.p, d_me.p, N, d_cnt.p);
            HIPC(hipGetLastError());
            HIPC(hipMemcpyAsync(cnt.data(), d_cnt.p, sizeof(int64_t) * np, hipMemcpyDeviceToHost, s0));
            HIPC(hipStreamSynchronize(s0));
            int64_t run = 0;
            for (int r = 0; r < np; r++) {
                rb[r] = run;
                run += cnt[r];
            }
            const int64_t nitems = run;
            h->pw.nitems = nitems;
            if (nitems == 0) continue;
            if (nitems > INT32_MAX) return fail(EBEM_EINVAL, "too many post-processing items in one chunk");
            HIPC(hipMemcpyAsync(d_rb.p, rb.data(), sizeof(int64_t) * np, hipMemcpyHostToDevice, s0));
            HIPC(d_iq.grow((size_t)nitems));
            HIPC(d_M.grow((size_t)nitems * 6));
            HIPC(d_key.grow((size_t)nitems));
            HIPC(d_perm.grow((size_t)nitems));
            hipLaunchKernelGGL(k_pp_items, dim3(np), dim3(256), 0, s0, d_ms.p, d_me.p, N, d_rb.p, d_off.p,
                               d_iq.p);
            HIPC(hipGetLastError());
            // bucket items by front-state key (two-pass counting sort; blocks
            // take ranges of (point, triangle) pairs)
            const int64_t per = (tot + G - 1) / G;
            hipLaunchKernelGGL(k_pp_keys, dim3(G), dim3(256), 0, s0, h->d_tri.p, d_x.p, p0, P, tot, per, d_off.p,
                               d_ms.p, d_me.p, kp ? d_ka.p : nullptr, kp ? d_kb.p : nullptr,
                               kp ? d_sa.p : nullptr, kp ? d_sb.p : nullptr, d_key.p, d_bh.p);
            HIPC(hipGetLastError());
            HIPC(hipMemcpyAsync(bh.data(), d_bh.p, sizeof(unsigned int) * bh.size(), hipMemcpyDeviceToHost,
                                s0));
            HIPC(hipStreamSynchronize(s0));
            unsigned long long runk = 0;
            for (size_t k = 0; k < bh.size(); k++) {
                bo[k] = runk;
                runk += bh[k];
            }
            HIPC(hipMemcpyAsync(d_bo.p, bo.data(), sizeof(unsigned long long) * bo.size(), hipMemcpyHostToDevice,
                                s0));
            hipLaunchKernelGGL(k_pp_scatter, dim3(G), dim3(256), 0, s0, tot, per, d_off.p, nitems, d_key.p, d_bo.p,
                               d_perm.p);
            HIPC(hipGetLastError());
            // perm: [per-m items | kP range heads | other kP items |
            //        between-front range heads | their other items]
            int64_t nkp = 0, nkrest = 0, nbf = 0, nbrest = 0;
            if (kp)
                for (int b = 0; b < G; b++) {
                    nkp += bh[(size_t)kKpKey * G + b];
                    nkrest += bh[(size_t)kKpRestKey * G + b];
                    for (int c = 0; c < kBfHeadKeys; c++) nbf += bh[(size_t)(kBfKey + c) * G + b];
                    nbrest += bh[(size_t)kBfRestKey * G + b];
                }
            h->pp_kp_stats[0] += nkp;
            h->pp_kp_stats[1] += nkp + nkrest;
            h->pp_kp_stats[2] += nbf;
            h->pp_kp_stats[3] += nbf + nbrest;
            const int64_t nper = nitems - nkp - nkrest - nbf - nbrest;
            int64_t mb = std::min<int64_t>(std::max<int64_t>(1, (nper + 255) / 256), (int64_t)h->CUs * 16);
            if (numeric)
                hipLaunchKernelGGL(k_pp_M_num, dim3((unsigned)mb), dim3(256), sizeof(double) * 4 * h->pp_nInt,
                                   s0, h->d_tri.p, h->d_verts.p, d_x.p, p0, P, h->d_pp_rule.p, h->pp_nInt,
                                   nitems, d_iq.p, d_off.p, d_ms.p, 4 * kPi * h->rho, d_M.p);
            else
                hipLaunchKernelGGL(k_pp_M, dim3((unsigned)mb), dim3(256), 0, s0, h->d_tri.p, d_x.p, p0, P,
                                   nper, d_iq.p, d_off.p, d_ms.p, h->sort_items ? d_perm.p : nullptr, d_M.p);
            HIPC(hipGetLastError());
            if (nkp > 0) {
                tppk.start(s0);
                const int64_t kblocks = std::min<int64_t>((nkp + 255) / 256, (int64_t)h->CUs * 16);
                hipLaunchKernelGGL(k_pp_M_kp, dim3((unsigned)kblocks), dim3(256), 0, s0, h->d_tri.p, d_x.p, p0, P,
                                   nkp, d_iq.p, d_ka.p, d_kb.p, d_perm.p + nper, d_M.p, d_nfb.p);
                HIPC(hipGetLastError());
                tppk.stop(s0);
                ppk_ms += tppk.ms();
            }
            if (nbf > 0) {
                tppb.start(s0);
                const int64_t bblocks = std::min<int64_t>((nbf + 255) / 256, (int64_t)h->CUs * 16);
                hipLaunchKernelGGL(k_pp_M_bf, dim3((unsigned)bblocks), dim3(256), 0, s0, h->d_tri.p, d_x.p,
                                   p0, P, nbf, d_iq.p, d_sa.p, d_sb.p, d_perm.p + nper + nkp + nkrest, d_M.p,
                                   d_nfb.p);
                HIPC(hipGetLastError());
                tppb.stop(s0);
                ppb_ms += tppb.ms();
            }
            if (stage == 1) continue;
            hipLaunchKernelGGL(k_pp_conv, dim3((np + kConvGroup - 1) / kConvGroup), dim3(64 * kConvSlices), 0,
                               s0, d_M.p, d_off.p, d_ms.p, d_me.p, aT, dT, N, nT, p0, np, d_pout.p, d_nh.p,
                               d_kidx.p, (int)nh.size(), npts, d_u.p);
            HIPC(hipGetLastError());
        }
        g0 = g1;
    }
    tall.stop(s0);
    if (stage == 1) {  // complete on return: the caller joins and stage 2 runs on another stream
        HIPC(hipStreamSynchronize(s0));
        h->prep.ms = tall.ms();
        h->tim[9] = ppk_ms;
        h->tim[11] = ppb_ms;
        return 0;
    }
    HIPC(hipMemcpyAsync(u_out, d_u.p, sizeof(double) * 3 * npts * nt, hipMemcpyDeviceToHost, h->stream));
    unsigned int nfb = 0;
    HIPC(hipMemcpyAsync(&nfb, d_nfb.p, sizeof nfb, hipMemcpyDeviceToHost, h->stream));
    HIPC(hipStreamSynchronize(h->stream));
    h->pp_kp_stats[4] = nfb;
    h->tim[6] = tall.ms();
    if (stage == 0) {
        h->tim[9] = ppk_ms;
        h->tim[11] = ppb_ms;
        h->tim[19] = 0;
    } else {
        h->tim[19] = h->prep.ms;
    }
    return 0;
}

// Whether a post-processing request can be prepared by ebem_pp_prepare: the
// analytic form, one group of times (one delta = t - nhat dt) and one chunk
// of points.
bool pp_preparable(const ebem_handle* h, int npts, int nt, const double* t) {
    if (npts <= 0 || nt <= 0 || pp_chunk_points(h, npts) < npts) return false;
    double dmin = 0, dmax = 0;
    int nv = 0;
    for (int k = 0; k < nt; k++) {
        if (!(t[k] > 0)) continue;
        const int nh = (int)std::ceil(t[k] / h->dt) - 1;
        if (nh > h->nT - 1) return false;
        const double d = t[k] - nh * h->dt;
        dmin = nv ? std::min(dmin, d) : d;
        dmax = nv ? std::max(dmax, d) : d;
        nv++;
    }
    return nv > 0 && dmax - dmin <= 1e-12 * h->dt;
}

// Waits for the worker of ebem_pp_prepare, if one runs, and reports its
// error; keep = false also drops the prepared request (state it depends on
// may change, or its buffers may be reused).
void pp_release(ebem_handle* h) {
    std::lock_guard<std::mutex> g(h->prep.go_m);
    h->prep.go = true;
    h->prep.go_cv.notify_all();
}

int pp_join(ebem_handle* h, bool keep = false) {
    if (h->prep.active) {
        pp_release(h);
        h->prep.worker.join();
        h->prep.active = false;
    }
    if (!keep) h->prep.ready = false;
    if (h->prep.err) {
        const int e = h->prep.err;
        h->prep.err = 0;
        h->prep.ready = false;
        return fail(e, "%s", h->prep.msg.c_str());
    }
    return 0;
}

}  // namespace

// ================================================================= C ABI ==


// ---------------------------------------------------------------------------
// Double-layer K of the direct Dirichlet driver (timeMarchingD.m:173-274)
// ---------------------------------------------------------------------------

// numBlocksK = min(numBlocksV + 1, nT) (timeMarchingD.m:176)
int k_num_blocks(ebem_handle* h) {
    compute_L(h);
    return std::min(h->L + 1, h->nT);
}

// P1 bases and the vertex -> (triangle, local vertex) lists, triangles
// ascending, on the device (kept until the mesh changes).
int k_geometry(ebem_handle* h) {
    if (h->kgeo_ready) return 0;
    const int N = h->N, Nv = h->Nv;
    std::vector<KTri> kt(N);
    std::vector<int32_t> voff(Nv + 1, 0), vadj(3 * (size_t)N);
    for (int F = 0; F < N; F++) {
        const double* V[3] = {&h->X[3 * (size_t)h->T[3 * F]], &h->X[3 * (size_t)h->T[3 * F + 1]],
                              &h->X[3 * (size_t)h->T[3 * F + 2]]};
        k_tri_basis(V[0], V[1], V[2], h->tri[F].n, h->tri[F].area, kt[F]);
        for (int v = 0; v < 3; v++) voff[h->T[3 * F + v] + 1]++;
    }
    for (int s = 0; s < Nv; s++) voff[s + 1] += voff[s];
    {
        std::vector<int32_t> pos(voff.begin(), voff.end() - 1);
        for (int F = 0; F < N; F++)
            for (int v = 0; v < 3; v++) vadj[pos[h->T[3 * F + v]]++] = F * 4 + v;
    }
    HIPC(h->d_ktri.grow(N));
    HIPC(h->d_T.grow(3 * (size_t)N));
    HIPC(h->d_X.grow(3 * (size_t)Nv));
    HIPC(h->d_vadj_off.grow(Nv + 1));
    HIPC(h->d_vadj.grow(3 * (size_t)N));
    HIPC(hipMemcpyAsync(h->d_ktri.p, kt.data(), sizeof(KTri) * N, hipMemcpyHostToDevice, h->stream));
    HIPC(hipMemcpyAsync(h->d_T.p, h->T.data(), sizeof(int32_t) * 3 * N, hipMemcpyHostToDevice, h->stream));
    HIPC(hipMemcpyAsync(h->d_X.p, h->X.data(), sizeof(double) * 3 * Nv, hipMemcpyHostToDevice, h->stream));
    HIPC(hipMemcpyAsync(h->d_vadj_off.p, voff.data(), sizeof(int32_t) * (Nv + 1), hipMemcpyHostToDevice, h->stream));
    HIPC(hipMemcpyAsync(h->d_vadj.p, vadj.data(), sizeof(int32_t) * 3 * N, hipMemcpyHostToDevice, h->stream));
    HIPC(hipStreamSynchronize(h->stream));
    h->kgeo_ready = true;
    return 0;
}

int assemble_k(ebem_handle* h) {
#pragma clang fp contract(off)
    if (!h->have_mesh || !h->have_problem) return fail(EBEM_EINVAL, "mesh and problem must be set");
    if (!h->ghc_nSub) return fail(EBEM_EINVAL, "ebem_assemble_k needs ebem_set_inner_rule first");
    if (h->lambda + h->mu == 0) return fail(EBEM_EINVAL, "Problema non adatto a questa formulazione");
    const int N = h->N, Nv = h->Nv;
    if ((int64_t)N * Nv >= (int64_t)INT32_MAX) return fail(EBEM_EINVAL, "mesh too large for the K store");
    const int LK = k_num_blocks(h);
    if (LK > 32767) return fail(EBEM_EINVAL, "too many lags for the K store");
    Timer tall;
    tall.start(h->stream);
    if (int e = k_geometry(h)) return e;
    KParams K{};
    K.N = N;
    K.Nv = Nv;
    K.LK = LK;
    K.nInt = h->ghc_nSub * h->ghc_nSing;
    K.nDiag = h->ghc_nDiag;
    K.dt = h->dt;
    K.cP = h->cP;
    K.cS = h->cS;
    K.lam = h->lambda;
    K.mu = h->mu;
    K.rho = h->rho;
    K.maxLen = h->hmax;
    K.fourpidt = 4 * kPi * h->dt;
    // pair windows and per-row slot counts
    const int64_t np = (int64_t)N * Nv;
    HIPC(h->d_klo.grow(np));
    HIPC(h->d_khi.grow(np));
    HIPC(h->d_krow.grow(2 * (size_t)N));
    hipLaunchKernelGGL(k_kwin, dim3(N), dim3(256), 0, h->stream, h->d_verts.p, h->d_X.p, h->d_T.p, h->d_tri.p, K,
                       h->d_klo.p, h->d_khi.p, h->d_krow.p);
    HIPC(hipGetLastError());
    std::vector<int64_t> rc(2 * (size_t)N);  // slots per row, then pairs per row
    HIPC(hipMemcpyAsync(rc.data(), h->d_krow.p, sizeof(int64_t) * 2 * N, hipMemcpyDeviceToHost, h->stream));
    HIPC(hipStreamSynchronize(h->stream));
    h->k_rowbase.assign(N + 1, 0);
    for (int S = 0; S < N; S++) h->k_rowbase[S + 1] = h->k_rowbase[S] + rc[S];
    const int64_t slots = h->k_rowbase[N];
    {
        size_t freeb = 0, totb = 0;
        HIPC(hipMemGetInfo(&freeb, &totb));
        const double need = (double)slots * (9 * 8 + 4) + (double)np * 8;
        if (need > 0.8 * (double)freeb)
            return fail(EBEM_ENOMEM, "K store needs %.1f GB of device memory", need / 1e9);
    }
    HIPC(h->d_koff.grow(np));
    HIPC(h->d_kslot.grow(std::max<int64_t>(slots, 1)));
    HIPC(h->d_kval.grow((size_t)9 * std::max<int64_t>(slots, 1)));
    DevBuf<int64_t> d_rb;
    HIPC(d_rb.alloc(N));
    HIPC(hipMemcpyAsync(d_rb.p, h->k_rowbase.data(), sizeof(int64_t) * N, hipMemcpyHostToDevice, h->stream));
    hipLaunchKernelGGL(k_kslots, dim3(N), dim3(256), 0, h->stream, h->d_klo.p, h->d_khi.p, Nv, d_rb.p, h->d_koff.p,
                       h->d_kslot.p);
    HIPC(hipGetLastError());
    if (slots > 0 && !h->k_by_field) {
        const size_t lds = sizeof(double) * 4 * (size_t)(K.nInt + K.nDiag);
        if (lds > 160 * 1024) return fail(EBEM_EINVAL, "inner rules too large for LDS");
        int64_t blocks = std::min<int64_t>((slots + 3) / 4, (int64_t)h->CUs * 8);
        HIPC(hipFuncSetAttribute((const void*)k_kblocks, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
        hipLaunchKernelGGL(k_kblocks, dim3((unsigned)blocks), dim3(256), lds, h->stream, h->d_ktri.p, h->d_tri.p,
                           h->d_verts.p, h->d_gh.p, h->d_X.p, h->d_T.p, h->d_vadj_off.p, h->d_vadj.p, h->d_ghc_in.p,
                           h->d_ghc_dg.p, h->nGH, K, slots, h->d_kslot.p, h->d_koff.p, h->d_klo.p, h->d_khi.p,
                           h->d_kval.p);
        HIPC(hipGetLastError());
    } else if (slots > 0) {
        // regular part by (S, F) pairs: windows of all rows, then row chunks
        // bounded by the partial-sum buffer, each gathered into the store
        DevBuf<int16_t> sflo, sfhi, itemC;
        DevBuf<int64_t> sfrow, pOff, rb;
        DevBuf<int32_t> itemP;
        DevBuf<double> part;
        const int64_t nsf = (int64_t)N * N;
        HIPC(sflo.alloc(nsf));
        HIPC(sfhi.alloc(nsf));
        HIPC(sfrow.alloc(2 * (size_t)N));
        hipLaunchKernelGGL(k_ksf_win, dim3(N), dim3(256), 0, h->stream, h->d_verts.p, h->d_X.p, h->d_T.p, K, 0, N,
                           sflo.p, sfhi.p, sfrow.p);
        HIPC(hipGetLastError());
        std::vector<int64_t> src(2 * (size_t)N);
        HIPC(hipMemcpyAsync(src.data(), sfrow.p, sizeof(int64_t) * 2 * N, hipMemcpyDeviceToHost, h->stream));
        HIPC(hipStreamSynchronize(h->stream));
        size_t freeb = 0, totb = 0;
        HIPC(hipMemGetInfo(&freeb, &totb));
        int64_t cap = (int64_t)(0.3 * (double)freeb / (27 * 8));
        int64_t maxrow = 0;
        for (int S = 0; S < N; S++) maxrow = std::max(maxrow, src[S]);
        if (cap < maxrow) cap = maxrow;
        const size_t lds_in = sizeof(double) * 4 * (size_t)K.nInt;
        if (lds_in > 160 * 1024) return fail(EBEM_EINVAL, "inner rule too large for LDS");
        HIPC(hipFuncSetAttribute((const void*)k_ksf, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds_in));
        int r0 = 0;
        while (r0 < N) {
            int r1 = r0;
            int64_t acc = 0, accI = 0;
            // (rows of a chunk) x N stays below 2^31: the item map holds int32 pair indices
            while (r1 < N && (r1 == r0 || (acc + src[r1] <= cap && (int64_t)(r1 - r0 + 1) * N <= INT32_MAX))) {
                acc += src[r1];
                accI += src[N + r1];
                r1++;
            }
            const int rows = r1 - r0;
            std::vector<int64_t> base(2 * (size_t)rows);
            int64_t run = 0, runI = 0;
            for (int r = 0; r < rows; r++) {
                base[r] = run;
                base[rows + r] = runI;
                run += src[r0 + r];
                runI += src[N + r0 + r];
            }
            HIPC(rb.grow(2 * (size_t)rows));
            HIPC(hipMemcpyAsync(rb.p, base.data(), sizeof(int64_t) * 2 * rows, hipMemcpyHostToDevice, h->stream));
            HIPC(pOff.grow((size_t)rows * N));
            HIPC(itemP.grow(std::max<int64_t>(accI, 1)));
            HIPC(itemC.grow(std::max<int64_t>(accI, 1)));
            HIPC(part.grow(27 * (size_t)std::max<int64_t>(acc, 1)));
            const int16_t* lo_c = sflo.p + (int64_t)r0 * N;
            const int16_t* hi_c = sfhi.p + (int64_t)r0 * N;
            hipLaunchKernelGGL(k_ksf_items, dim3(rows), dim3(256), 0, h->stream, lo_c, hi_c, N, rows, rb.p, pOff.p,
                               itemP.p, itemC.p);
            HIPC(hipGetLastError());
            if (accI > 0) {
                const int64_t blocks = std::min<int64_t>((accI + 3) / 4, (int64_t)h->CUs * 8);
                hipLaunchKernelGGL(k_ksf, dim3((unsigned)blocks), dim3(256), lds_in, h->stream, h->d_ktri.p,
                                   h->d_verts.p, h->d_gh.p, h->d_ghc_in.p, h->nGH, K, r0, accI, itemP.p, itemC.p,
                                   lo_c, hi_c, pOff.p, part.p);
                HIPC(hipGetLastError());
            }
            const int64_t s0 = h->k_rowbase[r0], ns = h->k_rowbase[r1] - s0;
            if (ns > 0) {
                hipLaunchKernelGGL(k_kgather, dim3((unsigned)((ns + 255) / 256)), dim3(256), 0, h->stream,
                                   h->d_verts.p, h->d_X.p, K, r0, s0, ns, h->d_kslot.p, h->d_koff.p, h->d_klo.p,
                                   h->d_vadj_off.p, h->d_vadj.p, lo_c, hi_c, pOff.p, part.p, h->d_kval.p);
                HIPC(hipGetLastError());
            }
            HIPC(hipStreamSynchronize(h->stream));  // part, pOff and the item map are reused by the next chunk
            r0 = r1;
        }
        const size_t lds_dg = sizeof(double) * 4 * (size_t)K.nDiag;
        HIPC(hipFuncSetAttribute((const void*)k_ksing, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds_dg));
        const int64_t blocks = std::min<int64_t>((3 * (int64_t)N + 3) / 4, (int64_t)h->CUs * 8);
        hipLaunchKernelGGL(k_ksing, dim3((unsigned)blocks), dim3(256), lds_dg, h->stream, h->d_ktri.p, h->d_tri.p,
                           h->d_verts.p, h->d_gh.p, h->d_T.p, h->d_ghc_dg.p, h->nGH, K, h->d_koff.p, h->d_klo.p,
                           h->d_khi.p, h->d_kval.p);
        HIPC(hipGetLastError());
    }
    tall.stop(h->stream);
    h->tim[12] = tall.ms();
    HIPC(hipStreamSynchronize(h->stream));
    h->k_pairs = 0;
    for (int S = 0; S < N; S++) h->k_pairs += rc[N + S];
    h->k_slots = slots;
    h->k_LK = LK;
    h->k_ready = true;
    return 0;
}

// Vertex datum (3Nv x nT) of kind `kind` into h->d_gT ([s][n][c]).
int k_datum_device(ebem_handle* h, int kind) {
    const int64_t n = (int64_t)h->Nv * h->nT;
    HIPC(h->d_X.grow(3 * (size_t)h->Nv));
    HIPC(hipMemcpyAsync(h->d_X.p, h->X.data(), sizeof(double) * 3 * h->Nv, hipMemcpyHostToDevice, h->stream));
    HIPC(h->d_gT.grow(3 * (size_t)n));
    hipLaunchKernelGGL(k_dir_datum, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, h->stream, h->d_X.p, h->Nv,
                       h->nT, h->dt, h->cP, h->Tfin, kind, h->d_gT.p);
    HIPC(hipGetLastError());
    return 0;
}

// beta_K (3N x nT) into h->d_betak from the datum in h->d_gT.
int k_beta_device(ebem_handle* h) {
    if (h->nT > 256 * kKBetaPer) return fail(EBEM_EINVAL, "beta_K supports at most %d time steps", 256 * kKBetaPer);
    HIPC(h->d_betak.grow(3 * (size_t)h->N * h->nT));
    Timer tb;
    tb.start(h->stream);
    hipLaunchKernelGGL(k_kbeta, dim3(h->N), dim3(256), 0, h->stream, h->d_klo.p, h->d_khi.p, h->d_koff.p,
                       h->d_kval.p, h->d_gT.p, h->N, h->Nv, h->nT, h->d_betak.p);
    HIPC(hipGetLastError());
    tb.stop(h->stream);
    h->tim[13] = tb.ms();
    return 0;
}


// ---------------------------------------------------------------------------
// Neumann driver (timeMarchingN.m:170-296)
// ---------------------------------------------------------------------------

int neu_datum_device(ebem_handle* h, int kind) {
    const int64_t n = (int64_t)h->N * h->nT;
    HIPC(h->d_gVT.grow(3 * (size_t)n));
    hipLaunchKernelGGL(k_neu_datum, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, h->stream, h->d_tri.p, h->N,
                       h->nT, h->dt, h->cP, h->Tfin, h->mu, kind == EBEM_DATUM_REF_BAR_H1 ? 1.0 : 3.0, h->d_gVT.p);
    HIPC(hipGetLastError());
    return 0;
}

int betav_device(ebem_handle* h) {
    if (h->nT > 256 * kBetaVPer) return fail(EBEM_EINVAL, "beta_V supports at most %d time steps", 256 * kBetaVPer);
    HIPC(h->d_betaV.grow(3 * (size_t)h->N * h->nT));
    hipLaunchKernelGGL(k_betav, dim3(h->N), dim3(256), 0, h->stream, h->store(), h->store_L, h->N, h->nT, h->d_gVT.p,
                       h->d_betaV.p);
    HIPC(hipGetLastError());
    return 0;
}

// (K_0 + I_Gamma) d_n = beta_V,n + I_Gamma d_{n-1} - sum_{l>=1} K_l d_{n-l} in the
// least-squares sense (MATLAB's backslash on the 3N x 3Nv system): A = QR once
// (rocSOLVER geqrf), P = R^{-1} Q^T (ormqr on the identity, trsm), then one
// gemv per step.
int mot_neumann(ebem_handle* h, int kind) {
    const int N = h->N, Nv = h->Nv, nT = h->nT;
    const int m = 3 * N, n = 3 * Nv;
    if (m < n) return fail(EBEM_EINVAL, "the Neumann system needs 3N >= 3Nv (N = %d, Nv = %d)", N, Nv);
    if (!h->blas) {
        if (int e = rocb(rocblas_create_handle(&h->blas), "rocblas_create_handle")) return e;
    }
    if (int e = rocb(rocblas_set_stream(h->blas, h->stream), "rocblas_set_stream")) return e;
    Timer tall;
    tall.start(h->stream);
    if (int e = neu_datum_device(h, kind)) return e;
    if (int e = betav_device(h)) return e;
    DevBuf<double> A, tau, Pm, tnf;
    HIPC(A.alloc((size_t)m * n));
    HIPC(tau.alloc(n));
    HIPC(Pm.alloc((size_t)m * m));
    HIPC(tnf.alloc(m));
    HIPC(hipMemsetAsync(A.p, 0, sizeof(double) * (size_t)m * n, h->stream));
    const int64_t np = (int64_t)N * Nv;
    hipLaunchKernelGGL(k_dn_sysmat, dim3((unsigned)((np + 255) / 256)), dim3(256), 0, h->stream, h->d_klo.p,
                       h->d_khi.p, h->d_koff.p, h->d_kval.p, h->d_T.p, h->d_tri.p, N, Nv, A.p);
    HIPC(hipGetLastError());
    if (int e = rocb(rocsolver_dgeqrf(h->blas, m, n, A.p, m, tau.p), "rocsolver_dgeqrf")) return e;
    HIPC(hipMemsetAsync(Pm.p, 0, sizeof(double) * (size_t)m * m, h->stream));
    hipLaunchKernelGGL(k_set_diag, dim3((m + 255) / 256), dim3(256), 0, h->stream, Pm.p, m, 1.0);
    HIPC(hipGetLastError());
    if (int e = rocb(rocsolver_dormqr(h->blas, rocblas_side_left, rocblas_operation_transpose, m, m, n, A.p, m, tau.p,
                                      Pm.p, m),
                     "rocsolver_dormqr"))
        return e;
    const double one = 1.0, zero = 0.0;
    if (int e = rocb(rocblas_dtrsm(h->blas, rocblas_side_left, rocblas_fill_upper, rocblas_operation_none,
                                   rocblas_diagonal_non_unit, n, m, &one, A.p, m, Pm.p, m),
                     "rocblas_dtrsm"))
        return e;
    HIPC(h->d_dn.grow((size_t)n * nT));
    for (int s = 0; s < nT; s++) {
        hipLaunchKernelGGL(k_dn_rhs, dim3(N), dim3(256), 0, h->stream, h->d_klo.p, h->d_khi.p, h->d_koff.p,
                           h->d_kval.p, h->d_T.p, h->d_tri.p, N, Nv, h->k_LK, s, h->d_betaV.p, h->d_dn.p, tnf.p);
        HIPC(hipGetLastError());
        if (int e = rocb(rocblas_dgemv(h->blas, rocblas_operation_none, n, m, &one, Pm.p, m, tnf.p, 1, &zero,
                                       h->d_dn.p + (size_t)s * n, 1),
                         "rocblas_dgemv"))
            return e;
    }
    tall.stop(h->stream);
    HIPC(hipStreamSynchronize(h->stream));
    h->tim[15] = tall.ms();
    h->have_dn = true;
    return 0;
}

extern "C" {

const char* ebem_last_error(void) { return g_err.c_str(); }

int ebem_create(ebem_handle** out, int device) {
    if (!out) return fail(EBEM_EINVAL, "null output pointer");
    *out = nullptr;
    int ndev = 0;
    HIPC(hipGetDeviceCount(&ndev));
    if (device < 0 || device >= ndev) return fail(EBEM_EINVAL, "device %d not available (%d devices)", device, ndev);
    HIPC(hipSetDevice(device));
    if (device >= kMaxDevices) return fail(EBEM_EINVAL, "device %d beyond the library's %d", device, kMaxDevices);
    {
        // address of this device's staging array, once per device, with the device current
        std::lock_guard<std::mutex> lk(ai_scratch_init_lock);
        if (!ai_scratch_ptr[device]) {
            void* ptr = nullptr;
            hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
            if (e != hipSuccess || !ptr)
                return fail(EBEM_EHIP, "hipGetSymbolAddress(scratch) failed on device %d: %s", device,
                            hipGetErrorString(e));
            ai_scratch_ptr[device] = ptr;
        }
    }
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
    const char* kps = std::getenv("EBEM_KP_SPLIT");
    h->kp_split = !(kps && std::strcmp(kps, "0") == 0);
    const char* nts = std::getenv("EBEM_NODE_TASKS");
    h->node_tasks = !(nts && std::strcmp(nts, "0") == 0);
    const char* nru = std::getenv("EBEM_NODE_REUSE");
    h->node_reuse = (nru && std::strcmp(nru, "0") == 0) ? 0 : (nru && std::strcmp(nru, "fallback") == 0) ? 2 : 1;
    const char* kbf = std::getenv("EBEM_K_BY_FIELD");
    h->k_by_field = !(kbf && std::strcmp(kbf, "0") == 0);
    const char* pbf = std::getenv("EBEM_PPD_BY_FIELD");
    h->ppd_by_field = !(pbf && std::strcmp(pbf, "0") == 0);
    *out = h;
    return 0;
}

int ebem_destroy(ebem_handle* h) {
    if (!h) return 0;
    {
        std::lock_guard<std::mutex> lk(h->lock);
        (void)pp_join(h);
        (void)hipSetDevice(h->device);
        (void)hipStreamSynchronize(h->stream);
        if (h->prep.stream) (void)hipStreamDestroy(h->prep.stream);
        if (h->blas) rocblas_destroy_handle(h->blas);
        if (h->sur) {
            std::lock_guard<std::mutex> lk2(h->sur->lock);
            h->sur->users--;
            h->sur = nullptr;
        }
        h->d_tri.reset();
        h->d_gh.reset();
        h->d_seg.reset();
        h->d_col.reset();
        h->d_val.reset();
        h->d_beta.reset();
        h->d_alpha.reset();
        h->d_alphaT.reset();
        h->d_dalphaT.reset();
        h->d_r.reset();
        h->d_A.reset();
        h->d_Ainv.reset();
        h->d_part.reset();
        h->d_ipiv.reset();
        h->d_info.reset();
        for (hipEvent_t e : h->mot_ev) (void)hipEventDestroy(e);
        h->mot_ev.clear();
        (void)hipStreamDestroy(h->stream);
    }
    delete h;  // the scratch buffers (aw, pw) are freed by their destructors
    return 0;
}

int ebem_set_mesh(ebem_handle* h, int N, int Nv, const double* coords, const int32_t* tri, int one_based) {
    if (int e = check_handle(h)) return e;
    std::lock_guard<std::mutex> lk(h->lock);
    if (int e = pp_join(h)) return e;
    if (N <= 0 || Nv <= 0 || !coords || !tri) return fail(EBEM_EINVAL, "invalid mesh (N=%d, Nv=%d)", N, Nv);
    if ((int64_t)N * N > (int64_t)1 << 31) return fail(EBEM_EINVAL, "mesh too large (N=%d)", N);
    HIPC(hipSetDevice(h->device));
    // Everything derived from the previous mesh (store, factors, densities,
    // device geometry) is invalid from here on, whether or not this mesh is
    // accepted: a rejected mesh leaves the handle without a mesh.
    h->have_mesh = false;
    h->assembled = false;
    h->k_ready = false;
    h->kgeo_ready = false;
    h->have_gT = false;
    h->have_dn = false;
    h->have_beta = h->have_alpha = false;
    // validate into locals; the handle takes the mesh only once it is valid
    std::vector<double> X(coords, coords + 3 * (size_t)Nv);
    std::vector<int> T(3 * (size_t)N);
    for (size_t k = 0; k < T.size(); k++) {
        int v = tri[k] - (one_based ? 1 : 0);
        if (v < 0 || v >= Nv) return fail(EBEM_EINVAL, "triangle %zu has vertex index %d out of range", k / 3, tri[k]);
        T[k] = v;
    }
    for (size_t k = 0; k < X.size(); k++)
        if (!std::isfinite(X[k])) return fail(EBEM_EINVAL, "vertex %zu has a non-finite coordinate", k / 3);
    std::vector<FieldTri> F;
    build_tri(X.data(), T.data(), N, F);
    for (int t = 0; t < N; t++)
        if (!(F[t].area > 0)) return fail(EBEM_EINVAL, "triangle %d is degenerate", t);
    h->N = N;
    h->Nv = Nv;
    h->X.swap(X);
    h->T.swap(T);
    h->tri.swap(F);
    if (int e = upload_geometry(h)) return e;
    h->have_mesh = true;
    return 0;
}

int ebem_set_problem(ebem_handle* h, double rho, double mu, double lambda, double Tfin, int nT) {
    if (int e = check_handle(h)) return e;
    std::lock_guard<std::mutex> lk(h->lock);
    if (int e = pp_join(h)) return e;
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
    h->k_ready = false;
    h->have_gT = false;
    h->have_beta = h->have_alpha = false;
    return 0;
}

int ebem_set_rule(ebem_handle* h, int nGH) {
    if (int e = check_handle(h)) return e;
    std::lock_guard<std::mutex> lk(h->lock);
    if (int e = pp_join(h)) return e;
    if (nGH != 3 && nGH != 7 && nGH != 12 && nGH != 19)
        return fail(EBEM_EINVAL, "Valori nodi esterni non validi: %d (allowed 3, 7, 12, 19)", nGH);
    HIPC(hipSetDevice(h->device));
    h->nGH = nGH;
    h->assembled = false;
    h->k_ready = false;
    if (h->have_mesh) return upload_geometry(h);
    return 0;
}

int ebem_set_inner_rule(ebem_handle* h, int nSub, int nSing, int nDiag) {
    if (int e = check_handle(h)) return e;
    std::lock_guard<std::mutex> lk(h->lock);
    if (int e = pp_join(h)) return e;
    // BEMenerg_setupCore.m:58-93 (method families GPU / FN)
    if (!(nSub == 1 || nSub == 4 || nSub == 16 || nSub == 64) ||
        !(nSing == 1 || nSing == 3 || nSing == 7 || nSing == 12 || nSing == 19))
        return fail(EBEM_EINVAL, "Valori nodi interni non validi (%d sub-regions x %d nodes)", nSub, nSing);
    if (!(nDiag == 16 || nDiag == 64 || nDiag == 256 || nDiag == 1024))
        return fail(EBEM_EINVAL, "Valori nodi interni non validi (%d diagonal nodes)", nDiag);
    HIPC(hipSetDevice(h->device));
    std::vector<double> in, dg;
    if (gh_composite(nSub, nSing, in) || doppio_gauss(nDiag, dg)) return fail(EBEM_EINVAL, "inner rule construction failed");
    HIPC(h->d_ghc_in.alloc(in.size()));
    HIPC(h->d_ghc_dg.alloc(dg.size()));
    HIPC(hipMemcpyAsync(h->d_ghc_in.p, in.data(), sizeof(double) * in.size(), hipMemcpyHostToDevice, h->stream));
    HIPC(hipMemcpyAsync(h->d_ghc_dg.p, dg.data(), sizeof(double) * dg.size(), hipMemcpyHostToDevice, h->stream));
    HIPC(hipStreamSynchronize(h->stream));
    h->ghc_nSub = nSub;
    h->ghc_nSing = nSing;
    h->ghc_nDiag = nDiag;
    h->assembled = false;
    h->k_ready = false;
    return 0;
}

int ebem_inner_rule_tables(int nSub, int nSing, int nDiag, double* inner, double* diag) {
    std::vector<double> in, dg;
    if (gh_composite(nSub, nSing, in) || doppio_gauss(nDiag, dg)) return fail(EBEM_EINVAL, "invalid inner rule");
    if (inner) std::memcpy(inner, in.data(), sizeof(double) * in.size());
    if (diag) std::memcpy(diag, dg.data(), sizeof(double) * dg.size());
    return 0;
}

int ebem_num_blocks(ebem_handle* h, int* L) {
    if (int e = check_handle(h)) return e;
    std::lock_guard<std::mutex> lk(h->lock);
    if (int e = pp_join(h, true)) return e;
    if (!h->have_mesh || !h->have_problem || !L) return fail(EBEM_EINVAL, "mesh and problem must be set");
    compute_L(h);
    *L = h->L;
    return 0;
}

int ebem_assemble(ebem_handle* h, int precision) {
    if (int e = check_handle(h)) return e;
    std::lock_guard<std::mutex> lk(h->lock);
    if (int e = pp_join(h)) return e;
    HIPC(hipSetDevice(h->device));
    return assemble(h, precision);
}

int ebem_store_info(ebem_handle* h, int64_t* nnzb, int64_t* pair_lags, int64_t* items, int64_t* bytes) {
    if (int e = check_handle(h)) return e;
    std::lock_guard<std::mutex> lk(h->lock);
    if (int e = pp_join(h, true)) return e;
    if (!h->assembled) return fail(EBEM_EINVAL, "not assembled");
    if (nnzb) *nnzb = h->nnzb;
    if (pair_lags) *pair_lags = h->pair_lags;
    if (items) *items = h->items;
    if (bytes) *bytes = h->nnzb * (6 * 8 + 4) + (int64_t)h->seg.size() * 8;
    return 0;
}

int ebem_store_lag_blocks(ebem_handle* h, int64_t* counts) {
    if (int e = check_handle(h)) return e;
    std::lock_guard<std::mutex> lk(h->lock);
    if (int e = pp_join(h, true)) return e;
    if (!h->assembled) return fail(EBEM_EINVAL, "not assembled");
    if (!counts) return fail(EBEM_EINVAL, "null output");
    const int L = h->store_L;
    for (int l = 0; l < L; l++) counts[l] = 0;
    for (int i = 0; i < h->N; i++)
        for (int l = 0; l < L; l++) counts[l] += h->seg[(size_t)i * L + l + 1] - h->seg[(size_t)i * L + l];
    return 0;
}

int ebem_assembly_kp_stats(ebem_handle* h, int64_t* out5) {
    if (int e = check_handle(h)) return e;
    std::lock_guard<std::mutex> lk(h->lock);
    if (int e = pp_join(h, true)) return e;
    if (!h->assembled) return fail(EBEM_EINVAL, "not assembled");
    if (!out5) return fail(EBEM_EINVAL, "null output");
    for (int k = 0; k < 5; k++) out5[k] = h->kp_stats[k];
    return 0;
}

int ebem_pp_kp_stats(ebem_handle* h, int64_t* out5) {
    if (int e = check_handle(h)) return e;
    std::lock_guard<std::mutex> lk(h->lock);
    if (int e = pp_join(h, true)) return e;
    if (!out5) return fail(EBEM_EINVAL, "null output");
    for (int k = 0; k < 5; k++) out5[k] = h->pp_kp_stats[k];
    return 0;
}

int ebem_assembly_stats(ebem_handle* h, int64_t* far_items, int64_t* fp64_items) {
    if (int e = check_handle(h)) return e;
    std::lock_guard<std::mutex> lk(h->lock);
    if (int e = pp_join(h, true)) return e;
    if (!h->assembled) return fail(EBEM_EINVAL, "not assembled");
    if (far_items) *far_items = h->far_items;
    if (fp64_items) *fp64_items = h->items - h->far_items - h->sur_items;
    return 0;
}

int ebem_get_block_rows(ebem_handle* h, int lag, int r0, int r1, double* out) {
    if (int e = check_handle(h)) return e;
    std::lock_guard<std::mutex> lk(h->lock);
    if (int e = pp_join(h, true)) return e;
    if (!h->assembled) return fail(EBEM_EINVAL, "not assembled");
    if (!out) return fail(EBEM_EINVAL, "null output");
    if (lag < 0) return fail(EBEM_EINVAL, "negative lag");
    if (r0 < 0 || r1 > h->N || r0 >= r1) return fail(EBEM_EINVAL, "row range [%d, %d) outside [0, %d)", r0, r1, h->N);
    HIPC(hipSetDevice(h->device));
    const size_t ld = 3 * (size_t)(r1 - r0), n3 = 3 * (size_t)h->N;
    if (lag >= h->store_L) {
        std::memset(out, 0, sizeof(double) * ld * n3);
        return 0;  // beyond L: exactly zero
    }
    DevBuf<double> d;
    HIPC(d.alloc(ld * n3));
    HIPC(hipMemsetAsync(d.p, 0, sizeof(double) * ld * n3, h->stream));
    hipLaunchKernelGGL(k_dense_lag, dim3(4, r1 - r0), dim3(64), 0, h->stream, h->store(), h->N, h->store_L, lag, r0,
                       (int64_t)ld, d.p);
    HIPC(hipGetLastError());
    HIPC(hipMemcpyAsync(out, d.p, sizeof(double) * ld * n3, hipMemcpyDeviceToHost, h->stream));
    HIPC(hipStreamSynchronize(h->stream));
    return 0;
}

int ebem_get_block(ebem_handle* h, int lag, double* out) {
    int N = 0;
    {
        if (int e = check_handle(h)) return e;
        std::lock_guard<std::mutex> lk(h->lock);
        if (int e = pp_join(h, true)) return e;
        N = h->N;
    }
    return ebem_get_block_rows(h, lag, 0, N, out);
}

int ebem_rhs(ebem_handle* h, int datum_kind, double* beta_out) {
    if (int e = check_handle(h)) return e;
    std::lock_guard<std::mutex> lk(h->lock);
    if (!h->have_mesh || !h->have_problem) return fail(EBEM_EINVAL, "mesh and problem must be set");
    if (datum_kind < 0 || datum_kind > EBEM_DATUM_MAX) return fail(EBEM_EINVAL, "unknown datum %d", datum_kind);
    HIPC(hipSetDevice(h->device));
    const size_t n3 = 3 * (size_t)h->N;
    HIPC(h->d_beta.alloc(n3 * h->nT));
    int64_t tot = (int64_t)h->N * h->nT;
    hipLaunchKernelGGL(k_rhs, dim3((unsigned)((tot + 255) / 256)), dim3(256), 0, h->stream, h->d_tri.p, h->d_gh.p, h->N,
                       h->nT, h->nGH, h->dt, h->cP, h->Tfin, datum_kind, h->d_beta.p);
    HIPC(hipGetLastError());
    if (beta_out)
        HIPC(hipMemcpyAsync(beta_out, h->d_beta.p, sizeof(double) * n3 * h->nT, hipMemcpyDeviceToHost, h->stream));
    HIPC(hipStreamSynchronize(h->stream));
    h->have_beta = true;
    return 0;
}

int ebem_k_num_blocks(ebem_handle* h, int* LK) {
    if (int e = check_handle(h)) return e;
    std::lock_guard<std::mutex> lk(h->lock);
    if (int e = pp_join(h, true)) return e;
    if (!h->have_mesh || !h->have_problem) return fail(EBEM_EINVAL, "mesh and problem must be set");
    if (!LK) return fail(EBEM_EINVAL, "bad argument");
    *LK = k_num_blocks(h);
    return 0;
}

int ebem_assemble_k(ebem_handle* h) {
    if (int e = check_handle(h)) return e;
    std::lock_guard<std::mutex> lk(h->lock);
    if (int e = pp_join(h)) return e;
    HIPC(hipSetDevice(h->device));
    return assemble_k(h);
}

int ebem_k_store_info(ebem_handle* h, int64_t* pairs, int64_t* slots, int64_t* bytes) {
    if (int e = check_handle(h)) return e;
    std::lock_guard<std::mutex> lk(h->lock);
    if (int e = pp_join(h, true)) return e;
    if (!h->k_ready) return fail(EBEM_EINVAL, "ebem_assemble_k has not run");
    if (pairs) *pairs = h->k_pairs;
    if (slots) *slots = h->k_slots;
    if (bytes) *bytes = h->k_slots * (int64_t)(9 * sizeof(double)) + (int64_t)h->N * h->Nv * (2 * 2 + 8);
    return 0;
}

int ebem_get_k_block_rows(ebem_handle* h, int lag, int r0, int r1, double* out) {
    if (int e = check_handle(h)) return e;
    std::lock_guard<std::mutex> lk(h->lock);
    if (int e = pp_join(h, true)) return e;
    if (!h->k_ready) return fail(EBEM_EINVAL, "ebem_assemble_k has not run");
    if (!out || r0 < 0 || r1 > h->N || r0 >= r1) return fail(EBEM_EINVAL, "bad argument");
    if (lag < 0 || lag >= h->nT) return fail(EBEM_EINVAL, "lag %d out of range [0, %d)", lag, h->nT);
    HIPC(hipSetDevice(h->device));
    const int Nv = h->Nv;
    const int64_t rows = r1 - r0, ld = 3 * rows;
    std::memset(out, 0, sizeof(double) * ld * 3 * Nv);
    if (lag >= h->k_LK) return 0;  // the reference's null blocks (timeMarchingD.m:267-269)
    const int64_t p0 = (int64_t)r0 * Nv, np = rows * Nv;
    const int64_t s0 = h->k_rowbase[r0], ns = h->k_rowbase[r1] - s0;
    std::vector<int16_t> lo(np), hi(np);
    std::vector<int64_t> off(np);
    std::vector<double> v((size_t)9 * ns);
    HIPC(hipMemcpyAsync(lo.data(), h->d_klo.p + p0, sizeof(int16_t) * np, hipMemcpyDeviceToHost, h->stream));
    HIPC(hipMemcpyAsync(hi.data(), h->d_khi.p + p0, sizeof(int16_t) * np, hipMemcpyDeviceToHost, h->stream));
    HIPC(hipMemcpyAsync(off.data(), h->d_koff.p + p0, sizeof(int64_t) * np, hipMemcpyDeviceToHost, h->stream));
    if (ns > 0)
        HIPC(hipMemcpyAsync(v.data(), h->d_kval.p + 9 * s0, sizeof(double) * 9 * ns, hipMemcpyDeviceToHost,
                            h->stream));
    HIPC(hipStreamSynchronize(h->stream));
    for (int64_t q = 0; q < np; q++) {
        if (lag < lo[q] || lag >= hi[q]) continue;
        const int64_t r = q / Nv, s = q % Nv;
        const double* b = v.data() + 9 * (off[q] - s0 + (lag - lo[q]));
        for (int i = 0; i < 3; i++)
            for (int j = 0; j < 3; j++) out[(3 * r + i) + ld * (3 * s + j)] = b[i + 3 * j];
    }
    return 0;
}

int ebem_dirichlet_datum(ebem_handle* h, int kind, double* g_out) {
    if (int e = check_handle(h)) return e;
    std::lock_guard<std::mutex> lk(h->lock);
    if (int e = pp_join(h)) return e;
    if (!h->have_mesh || !h->have_problem) return fail(EBEM_EINVAL, "mesh and problem must be set");
    if (kind < EBEM_DATUM_REF_BAR_H1 || kind > EBEM_DATUM_MAX)
        return fail(EBEM_EINVAL, "Problema non codificato (datum %d has no Dirichlet trace)", kind);
    if (!g_out) return fail(EBEM_EINVAL, "bad argument");
    HIPC(hipSetDevice(h->device));
    if (int e = k_datum_device(h, kind)) return e;
    const int64_t n = (int64_t)h->Nv * h->nT;
    std::vector<double> gT(3 * (size_t)n);
    HIPC(hipMemcpyAsync(gT.data(), h->d_gT.p, sizeof(double) * 3 * n, hipMemcpyDeviceToHost, h->stream));
    HIPC(hipStreamSynchronize(h->stream));
    for (int s = 0; s < h->Nv; s++)
        for (int t = 0; t < h->nT; t++)
            for (int c = 0; c < 3; c++)
                g_out[(int64_t)t * 3 * h->Nv + 3 * s + c] = gT[((int64_t)s * h->nT + t) * 3 + c];
    return 0;
}

int ebem_beta_k(ebem_handle* h, const double* g, double* beta_k_out) {
    if (int e = check_handle(h)) return e;
    std::lock_guard<std::mutex> lk(h->lock);
    if (int e = pp_join(h)) return e;
    if (!h->k_ready) return fail(EBEM_EINVAL, "ebem_assemble_k has not run");
    if (!g || !beta_k_out) return fail(EBEM_EINVAL, "bad argument");
    HIPC(hipSetDevice(h->device));
    const int64_t n = (int64_t)h->Nv * h->nT;
    DevBuf<double> d_g;
    HIPC(d_g.alloc(3 * (size_t)n));
    HIPC(hipMemcpyAsync(d_g.p, g, sizeof(double) * 3 * n, hipMemcpyHostToDevice, h->stream));
    HIPC(h->d_gT.grow(3 * (size_t)n));
    hipLaunchKernelGGL(k_dir_transpose, dim3((unsigned)((3 * n + 255) / 256)), dim3(256), 0, h->stream, d_g.p, h->Nv,
                       h->nT, h->d_gT.p);
    HIPC(hipGetLastError());
    h->have_gT = false;
    if (int e = k_beta_device(h)) return e;
    HIPC(hipMemcpyAsync(beta_k_out, h->d_betak.p, sizeof(double) * 3 * (size_t)h->N * h->nT, hipMemcpyDeviceToHost,
                        h->stream));
    HIPC(hipStreamSynchronize(h->stream));
    return 0;
}

int ebem_rhs_dirichlet(ebem_handle* h, int kind, double* beta_out) {
    if (int e = check_handle(h)) return e;
    std::lock_guard<std::mutex> lk(h->lock);
    if (int e = pp_join(h)) return e;
    if (!h->k_ready) return fail(EBEM_EINVAL, "ebem_assemble_k has not run");
    if (kind < EBEM_DATUM_REF_BAR_H1 || kind > EBEM_DATUM_MAX)
        return fail(EBEM_EINVAL, "Problema non codificato (datum %d has no Dirichlet trace)", kind);
    HIPC(hipSetDevice(h->device));
    const size_t n3 = 3 * (size_t)h->N;
    HIPC(h->d_beta.alloc(n3 * h->nT));
    const int64_t tot = (int64_t)h->N * h->nT;
    hipLaunchKernelGGL(k_rhs, dim3((unsigned)((tot + 255) / 256)), dim3(256), 0, h->stream, h->d_tri.p, h->d_gh.p, h->N,
                       h->nT, h->nGH, h->dt, h->cP, h->Tfin, kind, h->d_beta.p);
    HIPC(hipGetLastError());
    if (int e = k_datum_device(h, kind)) return e;
    h->have_gT = true;
    if (int e = k_beta_device(h)) return e;
    const int64_t nb = (int64_t)n3 * h->nT;
    hipLaunchKernelGGL(k_dd_combine, dim3((unsigned)((nb + 255) / 256)), dim3(256), 0, h->stream, h->d_beta.p,
                       h->d_betak.p, nb);
    HIPC(hipGetLastError());
    if (beta_out)
        HIPC(hipMemcpyAsync(beta_out, h->d_beta.p, sizeof(double) * nb, hipMemcpyDeviceToHost, h->stream));
    HIPC(hipStreamSynchronize(h->stream));
    h->have_beta = true;
    return 0;
}

int ebem_pp_eval_direct(ebem_handle* h, const double* density, const double* g, int npts, const double* x, int nt,
                        const double* t, double* u_out) {
    if (int e = check_handle(h)) return e;
    std::lock_guard<std::mutex> lk(h->lock);
    if (int e = pp_join(h)) return e;
    if (!h->have_mesh || !h->have_problem) return fail(EBEM_EINVAL, "mesh and problem must be set");
    if (!h->pp_nInt) return fail(EBEM_EINVAL, "the direct post-processing needs ebem_set_pp_rule first");
    if (npts < 0 || nt < 0 || (npts > 0 && !x) || (nt > 0 && !t) || !u_out) return fail(EBEM_EINVAL, "bad argument");
    if (!density && !h->have_alpha) return fail(EBEM_EINVAL, "no density: pass one or run ebem_mot_solve first");
    if (!g && !h->have_gT) return fail(EBEM_EINVAL, "no vertex datum: pass g or run ebem_rhs_dirichlet first");
    const int N = h->N, Nv = h->Nv, nT = h->nT;
    std::memset(u_out, 0, sizeof(double) * 3 * (size_t)npts * nt);
    // (point, time) pairs with t > 0; nHat = ceil(t / dt) must not pass nT (density(:, nHat))
    std::vector<int32_t> pq;
    for (int k = 0; k < nt; k++) {
        if (!(t[k] > 0)) continue;  // BEMenerg_ppD_calc.m:9-13
        const int nHat = (int)std::ceil(t[k] / h->dt);
        if (nHat > nT) return fail(EBEM_EINVAL, "time %g is past Tfin: the density has no step %d", t[k], nHat);
        for (int p = 0; p < npts; p++) pq.push_back(p), pq.push_back(k);
    }
    const int nq = (int)(pq.size() / 2);
    if (nq == 0) return 0;
    HIPC(hipSetDevice(h->device));
    if (int e = k_geometry(h)) return e;
    Timer tall;
    tall.start(h->stream);
    DevBuf<double> d_x, d_t, d_a, d_g, d_gT, d_oG, d_oP, d_u;
    DevBuf<int32_t> d_pq;
    HIPC(d_x.alloc(3 * (size_t)std::max(npts, 1)));
    HIPC(d_t.alloc(std::max(nt, 1)));
    HIPC(d_pq.alloc(pq.size()));
    HIPC(d_u.alloc(3 * (size_t)npts * nt));
    HIPC(hipMemcpyAsync(d_x.p, x, sizeof(double) * 3 * npts, hipMemcpyHostToDevice, h->stream));
    HIPC(hipMemcpyAsync(d_t.p, t, sizeof(double) * nt, hipMemcpyHostToDevice, h->stream));
    HIPC(hipMemcpyAsync(d_pq.p, pq.data(), sizeof(int32_t) * pq.size(), hipMemcpyHostToDevice, h->stream));
    HIPC(hipMemsetAsync(d_u.p, 0, sizeof(double) * 3 * (size_t)npts * nt, h->stream));
    const double* alpha = h->d_alpha.p;
    if (density) {
        HIPC(d_a.alloc(3 * (size_t)N * nT));
        HIPC(hipMemcpyAsync(d_a.p, density, sizeof(double) * 3 * (size_t)N * nT, hipMemcpyHostToDevice, h->stream));
        alpha = d_a.p;
    }
    const double* gT = h->d_gT.p;
    if (g) {
        const int64_t n = (int64_t)Nv * nT;
        HIPC(d_g.alloc(3 * (size_t)n));
        HIPC(d_gT.alloc(3 * (size_t)n));
        HIPC(hipMemcpyAsync(d_g.p, g, sizeof(double) * 3 * n, hipMemcpyHostToDevice, h->stream));
        hipLaunchKernelGGL(k_dir_transpose, dim3((unsigned)((3 * n + 255) / 256)), dim3(256), 0, h->stream, d_g.p, Nv,
                           nT, d_gT.p);
        HIPC(hipGetLastError());
        gT = d_gT.p;
    }
    PPDParams P{};
    P.N = N;
    P.Nv = Nv;
    P.nT = nT;
    P.nInt = h->pp_nInt;
    P.dt = h->dt;
    P.cP = h->cP;
    P.cS = h->cS;
    P.lam = h->lambda;
    P.mu = h->mu;
    P.rho = h->rho;
    P.maxLen = h->hmax;
    const size_t lds = sizeof(double) * 4 * (size_t)h->pp_nInt;
    // pairs per launch bounded by the per-wave output buffers (64 MB)
    const int qc = (int)std::max<int64_t>(1, std::min<int64_t>(nq, (int64_t)(64e6 / (24.0 * (N + Nv)))));
    HIPC(d_oG.alloc(3 * (size_t)qc * N));
    HIPC(d_oP.alloc(3 * (size_t)qc * Nv));
    if (h->ppd_by_field) {
        // double-layer part by (pair, triangle) and level: level windows of
        // every pair, then batches of pairs whose level sums fit the budget
        if (nT > 32767) return fail(EBEM_EINVAL, "too many time steps for the level windows (%d)", nT);
        DevBuf<int16_t> lF, itemC;
        DevBuf<int64_t> cnt, woff, base;
        DevBuf<int32_t> itemP;
        DevBuf<double> W;
        HIPC(lF.alloc((size_t)nq * N));
        HIPC(cnt.alloc(2 * (size_t)nq));
        hipLaunchKernelGGL(k_ppf_win, dim3(nq), dim3(256), 0, h->stream, h->d_tri.p, d_x.p, d_t.p, d_pq.p, nq, P, lF.p,
                           cnt.p);
        HIPC(hipGetLastError());
        std::vector<int64_t> hc(2 * (size_t)nq);
        HIPC(hipMemcpyAsync(hc.data(), cnt.p, sizeof(int64_t) * 2 * nq, hipMemcpyDeviceToHost, h->stream));
        HIPC(hipStreamSynchronize(h->stream));
        size_t freeb = 0, totb = 0;
        HIPC(hipMemGetInfo(&freeb, &totb));
        int64_t cap = (int64_t)(0.3 * (double)freeb / (27 * 8));
        for (int q = 0; q < nq; q++) cap = std::max(cap, hc[q]);
        HIPC(hipFuncSetAttribute((const void*)k_ppf, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
        int q0 = 0;
        while (q0 < nq) {
            int q1 = q0;
            int64_t lev = 0, items = 0;
            while (q1 < nq && q1 - q0 < qc && (q1 == q0 || lev + hc[q1] <= cap)) {
                lev += hc[q1];
                items += hc[nq + q1];
                q1++;
            }
            const int n = q1 - q0;
            std::vector<int64_t> hb(2 * (size_t)n);
            int64_t run = 0, runI = 0;
            for (int r = 0; r < n; r++) {
                hb[r] = run;
                hb[n + r] = runI;
                run += hc[q0 + r];
                runI += hc[nq + q0 + r];
            }
            HIPC(base.grow(2 * (size_t)n));
            HIPC(hipMemcpyAsync(base.p, hb.data(), sizeof(int64_t) * 2 * n, hipMemcpyHostToDevice, h->stream));
            HIPC(woff.grow((size_t)n * N));
            HIPC(itemP.grow(std::max<int64_t>(items, 1)));
            HIPC(itemC.grow(std::max<int64_t>(items, 1)));
            HIPC(W.grow(27 * (size_t)std::max<int64_t>(lev, 1)));
            const int16_t* lFb = lF.p + (int64_t)q0 * N;
            hipLaunchKernelGGL(k_ppf_items, dim3(n), dim3(256), 0, h->stream, lFb, N, n, base.p, woff.p, itemP.p,
                               itemC.p);
            HIPC(hipGetLastError());
            if (items > 0) {
                const int64_t bI = std::min<int64_t>((items + 256 / kPpfLanes - 1) / (256 / kPpfLanes),
                                                     (int64_t)h->CUs * 8);
                hipLaunchKernelGGL(k_ppf, dim3((unsigned)bI), dim3(256), lds, h->stream, h->d_ktri.p, h->d_verts.p,
                                   d_x.p, d_t.p, d_pq.p, q0, P, h->d_pp_rule.p, items, itemP.p, itemC.p, lFb, woff.p,
                                   W.p);
                HIPC(hipGetLastError());
            }
            const int64_t bP = std::min<int64_t>(((int64_t)n * Nv + 3) / 4, (int64_t)h->CUs * 8);
            hipLaunchKernelGGL(k_ppf_gather, dim3((unsigned)bP), dim3(256), 0, h->stream, h->d_vadj_off.p,
                               h->d_vadj.p, d_t.p, d_pq.p, q0, n, P, lFb, woff.p, W.p, gT, d_oP.p);
            HIPC(hipGetLastError());
            const int64_t bG = std::min<int64_t>(((int64_t)n * N + 15) / 16, (int64_t)h->CUs * 8);
            hipLaunchKernelGGL(k_ppd_G, dim3((unsigned)bG), dim3(256), lds, h->stream, h->d_tri.p, h->d_verts.p, d_x.p,
                               d_t.p, d_pq.p, q0, n, P, h->d_pp_rule.p, alpha, d_oG.p);
            HIPC(hipGetLastError());
            hipLaunchKernelGGL(k_ppd_sum, dim3(n), dim3(256), 0, h->stream, d_oG.p, d_oP.p, d_pq.p, q0, N, Nv, npts,
                               d_u.p);
            HIPC(hipGetLastError());
            HIPC(hipStreamSynchronize(h->stream));  // hb, the level sums and the item map are reused
            q0 = q1;
        }
    }
    for (int q0 = 0; q0 < nq && !h->ppd_by_field; q0 += qc) {
        const int n = std::min(qc, nq - q0);
        const int64_t bG = std::min<int64_t>(((int64_t)n * N + 15) / 16, (int64_t)h->CUs * 8);
        hipLaunchKernelGGL(k_ppd_G, dim3((unsigned)bG), dim3(256), lds, h->stream, h->d_tri.p, h->d_verts.p, d_x.p,
                           d_t.p, d_pq.p, q0, n, P, h->d_pp_rule.p, alpha, d_oG.p);
        HIPC(hipGetLastError());
        const int64_t bP = std::min<int64_t>(((int64_t)n * Nv + 3) / 4, (int64_t)h->CUs * 8);
        hipLaunchKernelGGL(k_ppd_P, dim3((unsigned)bP), dim3(256), lds, h->stream, h->d_ktri.p, h->d_verts.p, h->d_X.p,
                           h->d_vadj_off.p, h->d_vadj.p, d_x.p, d_t.p, d_pq.p, q0, n, P, h->d_pp_rule.p, gT, d_oP.p);
        HIPC(hipGetLastError());
        hipLaunchKernelGGL(k_ppd_sum, dim3(n), dim3(256), 0, h->stream, d_oG.p, d_oP.p, d_pq.p, q0, N, Nv, npts,
                           d_u.p);
        HIPC(hipGetLastError());
    }
    tall.stop(h->stream);
    HIPC(hipMemcpyAsync(u_out, d_u.p, sizeof(double) * 3 * (size_t)npts * nt, hipMemcpyDeviceToHost, h->stream));
    HIPC(hipStreamSynchronize(h->stream));
    h->tim[14] = tall.ms();
    return 0;
}

int ebem_neumann_datum(ebem_handle* h, int datum, double* gV_out) {
    if (int e = check_handle(h)) return e;
    std::lock_guard<std::mutex> lk(h->lock);
    if (int e = pp_join(h)) return e;
    if (!h->have_mesh || !h->have_problem) return fail(EBEM_EINVAL, "mesh and problem must be set");
    if (datum != EBEM_DATUM_REF_BAR_H1 && datum != EBEM_DATUM_REF_BAR_H3)
        return fail(EBEM_EINVAL, "Problema non codificato (datum %d has no Neumann trace)", datum);
    if (!gV_out) return fail(EBEM_EINVAL, "bad argument");
    HIPC(hipSetDevice(h->device));
    if (int e = neu_datum_device(h, datum)) return e;
    const int64_t n = (int64_t)h->N * h->nT;
    std::vector<double> gT(3 * (size_t)n);
    HIPC(hipMemcpyAsync(gT.data(), h->d_gVT.p, sizeof(double) * 3 * n, hipMemcpyDeviceToHost, h->stream));
    HIPC(hipStreamSynchronize(h->stream));
    for (int F = 0; F < h->N; F++)
        for (int t = 0; t < h->nT; t++)
            for (int c = 0; c < 3; c++)
                gV_out[(int64_t)t * 3 * h->N + 3 * F + c] = gT[((int64_t)F * h->nT + t) * 3 + c];
    return 0;
}

int ebem_beta_v(ebem_handle* h, const double* gV, double* beta_v_out) {
    if (int e = check_handle(h)) return e;
    std::lock_guard<std::mutex> lk(h->lock);
    if (int e = pp_join(h)) return e;
    if (!h->assembled) return fail(EBEM_EINVAL, "ebem_assemble must run before ebem_beta_v");
    if (!gV || !beta_v_out) return fail(EBEM_EINVAL, "bad argument");
    HIPC(hipSetDevice(h->device));
    const int N = h->N, nT = h->nT;
    const int64_t n = (int64_t)N * nT;
    std::vector<double> gT(3 * (size_t)n);
    for (int F = 0; F < N; F++)
        for (int t = 0; t < nT; t++)
            for (int c = 0; c < 3; c++) gT[((int64_t)F * nT + t) * 3 + c] = gV[(int64_t)t * 3 * N + 3 * F + c];
    HIPC(h->d_gVT.grow(3 * (size_t)n));
    HIPC(hipMemcpyAsync(h->d_gVT.p, gT.data(), sizeof(double) * 3 * n, hipMemcpyHostToDevice, h->stream));
    if (int e = betav_device(h)) return e;
    HIPC(hipMemcpyAsync(beta_v_out, h->d_betaV.p, sizeof(double) * 3 * (size_t)n, hipMemcpyDeviceToHost, h->stream));
    HIPC(hipStreamSynchronize(h->stream));
    return 0;
}

int ebem_mot_solve_neumann(ebem_handle* h, int datum, double* density_out) {
    if (int e = check_handle(h)) return e;
    std::lock_guard<std::mutex> lk(h->lock);
    if (int e = pp_join(h)) return e;
    if (!h->assembled) return fail(EBEM_EINVAL, "ebem_assemble must run before ebem_mot_solve_neumann");
    if (!h->k_ready) return fail(EBEM_EINVAL, "ebem_assemble_k must run before ebem_mot_solve_neumann");
    if (datum != EBEM_DATUM_REF_BAR_H1 && datum != EBEM_DATUM_REF_BAR_H3)
        return fail(EBEM_EINVAL, "Problema non codificato (datum %d has no Neumann trace)", datum);
    HIPC(hipSetDevice(h->device));
    if (int e = mot_neumann(h, datum)) return e;
    if (density_out) {
        HIPC(hipMemcpyAsync(density_out, h->d_dn.p, sizeof(double) * 3 * (size_t)h->Nv * h->nT,
                            hipMemcpyDeviceToHost, h->stream));
        HIPC(hipStreamSynchronize(h->stream));
    }
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
    if (int e = pp_join(h)) return e;
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

int ebem_analitic_intg(ebem_handle* h, int n, const int32_t* kind, const double* args, const double* info,
                       const double* B, double* ris) {
    if (int e = check_handle(h)) return e;
    std::lock_guard<std::mutex> lk(h->lock);
    if (int e = pp_join(h)) return e;
    if (n < 0 || (n > 0 && (!kind || !args || !info || !B || !ris))) return fail(EBEM_EINVAL, "bad argument");
    if (n > (1 << 24)) return fail(EBEM_EINVAL, "too many region calls in one batch (%d > %d)", n, 1 << 24);
    for (int k = 0; k < n; k++)
        if (kind[k] < EBEM_AI_CC || kind[k] > EBEM_AI_TR_SP) return fail(EBEM_EINVAL, "unknown region kind %d", kind[k]);
    if (n == 0) return 0;
    HIPC(hipSetDevice(h->device));
    DevBuf<int32_t> dk;
    DevBuf<double> d;
    const size_t nn = (size_t)n;
    HIPC(dk.alloc(nn));
    HIPC(d.alloc(nn * (8 + 12 + 54 + 36)));
    double *da = d.p, *di = da + 8 * nn, *dB = di + 12 * nn, *dr = dB + 54 * nn;
    HIPC(hipMemcpyAsync(dk.p, kind, sizeof(int32_t) * nn, hipMemcpyHostToDevice, h->stream));
    HIPC(hipMemcpyAsync(da, args, sizeof(double) * 8 * nn, hipMemcpyHostToDevice, h->stream));
    HIPC(hipMemcpyAsync(di, info, sizeof(double) * 12 * nn, hipMemcpyHostToDevice, h->stream));
    HIPC(hipMemcpyAsync(dB, B, sizeof(double) * 54 * nn, hipMemcpyHostToDevice, h->stream));
    hipLaunchKernelGGL(k_ai_regions, dim3((unsigned)((n + 63) / 64)), dim3(64), 0, h->stream, n, dk.p, da, di, dB, dr);
    HIPC(hipGetLastError());
    HIPC(hipMemcpyAsync(ris, dr, sizeof(double) * 36 * nn, hipMemcpyDeviceToHost, h->stream));
    HIPC(hipStreamSynchronize(h->stream));
    return 0;
}

int ebem_analitic_intg_noalloc(ebem_handle* h, int n, const int32_t* kind, const double* args, const double* info,
                               const double* B, double* ris) {
    if (int e = check_handle(h)) return e;
    std::lock_guard<std::mutex> lk(h->lock);
    if (int e = pp_join(h)) return e;
    if (n < 0 || (n > 0 && (!kind || !args || !info || !B || !ris))) return fail(EBEM_EINVAL, "bad argument");
    if (n > kAiScratchCalls)
        return fail(EBEM_EINVAL, "ebem_analitic_intg_noalloc takes at most %d calls per batch (got %d)",
                    kAiScratchCalls, n);
    for (int k = 0; k < n; k++)
        if (kind[k] < EBEM_AI_CC || kind[k] > EBEM_AI_TR_SP) return fail(EBEM_EINVAL, "unknown region kind %d", kind[k]);
    if (n == 0) return 0;
    void* base = ai_scratch_ptr[h->device];
    if (!base) return fail(EBEM_EHIP, "no staging array on device %d", h->device);
    HIPC(hipSetDevice(h->device));
    std::lock_guard<std::mutex> dlk(ai_scratch_lock[h->device]);
    const size_t nn = (size_t)n, cap = (size_t)kAiScratchCalls;
    double* da = static_cast<double*>(base);
    double *di = da + 8 * cap, *dB = di + 12 * cap, *dr = dB + 54 * cap;
    int32_t* dk = reinterpret_cast<int32_t*>(dr + 36 * cap);
    HIPC(hipMemcpyAsync(dk, kind, sizeof(int32_t) * nn, hipMemcpyHostToDevice, h->stream));
    HIPC(hipMemcpyAsync(da, args, sizeof(double) * 8 * nn, hipMemcpyHostToDevice, h->stream));
    HIPC(hipMemcpyAsync(di, info, sizeof(double) * 12 * nn, hipMemcpyHostToDevice, h->stream));
    HIPC(hipMemcpyAsync(dB, B, sizeof(double) * 54 * nn, hipMemcpyHostToDevice, h->stream));
    hipLaunchKernelGGL(k_ai_regions, dim3((unsigned)((n + 63) / 64)), dim3(64), 0, h->stream, n, dk, da, di, dB, dr);
    HIPC(hipGetLastError());
    HIPC(hipMemcpyAsync(ris, dr, sizeof(double) * 36 * nn, hipMemcpyDeviceToHost, h->stream));
    HIPC(hipStreamSynchronize(h->stream));
    return 0;
}

static int ai_one(ebem_handle* h, int kind, double RS, double rhoS, double RP, double rhoP, double zeta, double TH1,
                  double TH2, const double* info, const double* B, double* ris, int nris) {
    if (!info || !B || !ris) return fail(EBEM_EINVAL, "bad argument");
    const double a[8] = {RS, rhoS, RP, rhoP, zeta, TH1, TH2, 0.0};
    const int32_t k = kind;
    double r[36];
    if (int e = ebem_analitic_intg(h, 1, &k, a, info, B, r)) return e;
    std::memcpy(ris, r, sizeof(double) * 9 * nris);
    return 0;
}

int ebem_analiticIntg_CC(ebem_handle* h, double RS, double rhoS, double RP, double rhoP, double zeta, double THETA1,
                         double THETA2, const double* infoTri2D, const double* B, double* ris1to4) {
    return ai_one(h, EBEM_AI_CC, RS, rhoS, RP, rhoP, zeta, THETA1, THETA2, infoTri2D, B, ris1to4, 4);
}

int ebem_analiticIntg_SC_P(ebem_handle* h, double RP, double rhoP, double zeta, double THETA1, double THETA2,
                           const double* infoTri2D, const double* B, double* ris1to4) {
    return ai_one(h, EBEM_AI_SC_P, 0.0, 0.0, RP, rhoP, zeta, THETA1, THETA2, infoTri2D, B, ris1to4, 4);
}

int ebem_analiticIntg_SC_SP(ebem_handle* h, double RS, double rhoS, double zeta, double THETA1, double THETA2,
                            const double* infoTri2D, const double* B, double* ris56) {
    return ai_one(h, EBEM_AI_SC_SP, RS, rhoS, 0.0, 0.0, zeta, THETA1, THETA2, infoTri2D, B, ris56, 2);
}

int ebem_analiticIntg_TC(ebem_handle* h, double zeta, double RS, double rhoS, double THETA1, double THETA2,
                         const double* infoTri2D, const double* B, double* ris1to4) {
    return ai_one(h, EBEM_AI_TC, RS, rhoS, 0.0, 0.0, zeta, THETA1, THETA2, infoTri2D, B, ris1to4, 4);
}

int ebem_analiticIntg_TR_P(ebem_handle* h, double zeta, double THETA1, double THETA2, const double* infoTri2D,
                           const double* B, double* ris1to4) {
    return ai_one(h, EBEM_AI_TR_P, 0.0, 0.0, 0.0, 0.0, zeta, THETA1, THETA2, infoTri2D, B, ris1to4, 4);
}

int ebem_analiticIntg_TR_SP(ebem_handle* h, double zeta, double THETA1, double THETA2, const double* infoTri2D,
                            const double* B, double* ris56) {
    return ai_one(h, EBEM_AI_TR_SP, 0.0, 0.0, 0.0, 0.0, zeta, THETA1, THETA2, infoTri2D, B, ris56, 2);
}

int ebem_set_option(ebem_handle* h, int option, double value) {
    if (int e = check_handle(h)) return e;
    std::lock_guard<std::mutex> lk(h->lock);
    if (int e = pp_join(h)) return e;
    switch (option) {
        case EBEM_OPT_FAR_FACTOR:
            if (!std::isfinite(value)) return fail(EBEM_EINVAL, "far-field factor must be finite");
            h->far_factor = value;
            h->assembled = false;
            h->have_alpha = false;
            return 0;
        case EBEM_OPT_SURROGATE_SOLVE:
            h->allow_sur_solve = value != 0.0;
            return 0;
        case EBEM_OPT_SURROGATE_PROJECT:
            h->sur_project = value != 0.0;
            h->assembled = false;
            h->have_alpha = false;
            return 0;
        default: return fail(EBEM_EINVAL, "unknown option %d", option);
    }
}

int ebem_pp_eval(ebem_handle* h, const double* density, int npts, const double* x, int nt, const double* t,
                 double* u_out) {
    if (int e = check_handle(h)) return e;
    std::lock_guard<std::mutex> lk(h->lock);
    if (int e = pp_join(h, true)) return e;
    HIPC(hipSetDevice(h->device));
    // the request ebem_pp_prepare prepared (same points and times, bit for
    // bit): only its density-dependent rest
    auto& P = h->prep;
    const bool prepared = P.ready && x && t && npts > 0 && nt > 0 && P.x.size() == 3 * (size_t)npts &&
                          P.t.size() == (size_t)nt &&
                          std::memcmp(P.x.data(), x, sizeof(double) * 3 * (size_t)npts) == 0 &&
                          std::memcmp(P.t.data(), t, sizeof(double) * (size_t)nt) == 0;
    P.ready = false;
    return pp_eval(h, density, npts, x, nt, t, u_out, false, prepared ? 2 : 0);
}

int ebem_pp_prepare(ebem_handle* h, int npts, const double* x, int nt, const double* t) {
    if (int e = check_handle(h)) return e;
    std::lock_guard<std::mutex> lk(h->lock);
    if (int e = pp_join(h)) return e;
    if (!h->have_mesh || !h->have_problem) return fail(EBEM_EINVAL, "mesh and problem must be set");
    if (npts <= 0 || nt <= 0 || !x || !t) return fail(EBEM_EINVAL, "bad post-processing arguments");
    HIPC(hipSetDevice(h->device));
    if (!pp_preparable(h, npts, nt, t)) return 0;  // ebem_pp_eval then does all of it
    auto& P = h->prep;
    if (!P.stream) HIPC(hipStreamCreateWithFlags(&P.stream, hipStreamNonBlocking));
    P.x.assign(x, x + 3 * (size_t)npts);
    P.t.assign(t, t + (size_t)nt);
    P.err = 0;
    P.go = false;
    try {
        P.worker = std::thread([h, npts, nt] {
            auto& Q = h->prep;
            {
                std::unique_lock<std::mutex> g(Q.go_m);
                Q.go_cv.wait(g, [&Q] { return Q.go; });
            }
            hipError_t e0 = hipSetDevice(h->device);
            const int e = e0 != hipSuccess ? fail(EBEM_EHIP, "hipSetDevice failed: %s", hipGetErrorString(e0))
                                           : pp_eval(h, nullptr, npts, Q.x.data(), nt, Q.t.data(), nullptr, false,
                                                     1, Q.stream);
            if (e) {
                Q.err = e;
                Q.msg = g_err;
            } else {
                Q.ready = true;
            }
        });
    } catch (const std::exception& ex) {  // no worker: ebem_pp_eval does all of it
        return fail(EBEM_EINVAL, "ebem_pp_prepare: cannot start its worker thread (%s)", ex.what());
    }
    P.active = true;
    return 0;
}

int ebem_set_pp_rule(ebem_handle* h, int nSub, int nSing) {
    if (int e = check_handle(h)) return e;
    std::lock_guard<std::mutex> lk(h->lock);
    if (int e = pp_join(h)) return e;
    if (!(nSub == 1 || nSub == 4 || nSub == 16 || nSub == 64) ||
        !(nSing == 1 || nSing == 3 || nSing == 7 || nSing == 12 || nSing == 19))
        return fail(EBEM_EINVAL, "Valori nodi interni non validi (%d sub-regions x %d nodes)", nSub, nSing);
    HIPC(hipSetDevice(h->device));
    std::vector<double> r;
    if (gh_composite(nSub, nSing, r)) return fail(EBEM_EINVAL, "rule construction failed");
    HIPC(h->d_pp_rule.alloc(r.size()));
    HIPC(hipMemcpyAsync(h->d_pp_rule.p, r.data(), sizeof(double) * r.size(), hipMemcpyHostToDevice, h->stream));
    HIPC(hipStreamSynchronize(h->stream));
    h->pp_nInt = nSub * nSing;
    return 0;
}

int ebem_pp_eval_numeric(ebem_handle* h, const double* density, int npts, const double* x, int nt, const double* t,
                         double* u_out) {
    if (int e = check_handle(h)) return e;
    std::lock_guard<std::mutex> lk(h->lock);
    if (int e = pp_join(h)) return e;
    HIPC(hipSetDevice(h->device));
    return pp_eval(h, density, npts, x, nt, t, u_out, true);
}

int ebem_timings(ebem_handle* h, double* out, int n) {
    if (int e = check_handle(h)) return e;
    std::lock_guard<std::mutex> lk(h->lock);
    if (int e = pp_join(h, true)) return e;
    if (!out || n < 0) return fail(EBEM_EINVAL, "null output or negative count");
    for (int k = 0; k < n && k < EBEM_NUM_TIMINGS; k++) out[k] = h->tim[k];
    return 0;
}


// ------------------------------------------------------------ surrogate --

int ebem_sur_create(ebem_sur** out, int device) {
    if (!out) return fail(EBEM_EINVAL, "null output pointer");
    *out = nullptr;
    int ndev = 0;
    HIPC(hipGetDeviceCount(&ndev));
    if (device < 0 || device >= ndev) return fail(EBEM_EINVAL, "device %d not available (%d devices)", device, ndev);
    HIPC(hipSetDevice(device));
    ebem_sur* s = new ebem_sur();
    s->device = device;
    hipError_t e = hipStreamCreate(&s->stream);
    if (e != hipSuccess) {
        delete s;
        return fail(EBEM_EHIP, "hipStreamCreate failed: %s", hipGetErrorString(e));
    }
    if (int r = sur_alloc_params(s)) {
        (void)hipStreamDestroy(s->stream);
        delete s;
        return r;
    }
    s->norm.assign(kSurNorm, 0.f);
    *out = s;
    return 0;
}

int ebem_sur_destroy(ebem_sur* s) {
    if (!s) return 0;
    {
        std::lock_guard<std::mutex> lk(s->lock);
        if (s->users > 0) return fail(EBEM_EINVAL, "surrogate still attached to %d handle(s)", s->users);
        (void)hipSetDevice(s->device);
        (void)hipStreamSynchronize(s->stream);
        (void)hipStreamDestroy(s->stream);
    }
    delete s;
    return 0;
}

int ebem_sur_add_samples(ebem_sur* s, ebem_handle* h, int64_t n, const int32_t* test, const int32_t* field,
                         const int32_t* lag) {
    if (!s || !h) return fail(EBEM_EINVAL, "null surrogate or handle");
    std::lock_guard<std::mutex> lk(h->lock);
    std::lock_guard<std::mutex> lk2(s->lock);
    if (!h->have_mesh || !h->have_problem) return fail(EBEM_EINVAL, "mesh and problem must be set");
    if (h->device != s->device) return fail(EBEM_EINVAL, "surrogate and handle live on different devices");
    if (n < 0 || (n > 0 && (!test || !field || !lag))) return fail(EBEM_EINVAL, "bad sample arrays");
    for (int64_t k = 0; k < n; k++)
        if (test[k] < 0 || test[k] >= h->N || field[k] < 0 || field[k] >= h->N || lag[k] < 0 || lag[k] > 4096)
            return fail(EBEM_EINVAL, "sample %lld out of range (test %d, field %d, lag %d)", (long long)k, test[k],
                        field[k], lag[k]);
    if (n == 0) return 0;
    HIPC(hipSetDevice(h->device));
    if (s->n + n > s->cap) {
        int64_t cap = std::max<int64_t>(s->n + n, 2 * s->cap);
        DevBuf<float> X, Y;
        HIPC(X.alloc((size_t)cap * kSurIn));
        HIPC(Y.alloc((size_t)cap * kSurOut));
        if (s->n) {
            HIPC(hipMemcpyAsync(X.p, s->d_X.p, sizeof(float) * s->n * kSurIn, hipMemcpyDeviceToDevice, h->stream));
            HIPC(hipMemcpyAsync(Y.p, s->d_Y.p, sizeof(float) * s->n * kSurOut, hipMemcpyDeviceToDevice, h->stream));
        }
        HIPC(hipStreamSynchronize(h->stream));
        std::swap(s->d_X.p, X.p);
        std::swap(s->d_X.n, X.n);
        std::swap(s->d_Y.p, Y.p);
        std::swap(s->d_Y.n, Y.n);
        s->cap = cap;
    }
    DevBuf<int32_t> idx;
    HIPC(idx.alloc((size_t)3 * n));
    HIPC(hipMemcpyAsync(idx.p, test, sizeof(int32_t) * n, hipMemcpyHostToDevice, h->stream));
    HIPC(hipMemcpyAsync(idx.p + n, field, sizeof(int32_t) * n, hipMemcpyHostToDevice, h->stream));
    HIPC(hipMemcpyAsync(idx.p + 2 * n, lag, sizeof(int32_t) * n, hipMemcpyHostToDevice, h->stream));
    const unsigned blocks = (unsigned)((n + 255) / 256);
    float* X = s->d_X.p + s->n * kSurIn;
    float* Y = s->d_Y.p + s->n * kSurOut;
#define SUR_SAMPLES(NG)                                                                                               \
    hipLaunchKernelGGL(k_sur_samples<NG>, dim3(blocks), dim3(256), 0, h->stream, h->d_tri.p, h->d_gh.p, h->d_verts.p, \
                       n, idx.p, idx.p + n, idx.p + 2 * n, h->dt, h->cP, h->cS, h->rho, h->hmax, X, Y)
    switch (h->nGH) {
        case 3: SUR_SAMPLES(3); break;
        case 7: SUR_SAMPLES(7); break;
        case 12: SUR_SAMPLES(12); break;
        default: SUR_SAMPLES(19); break;
    }
#undef SUR_SAMPLES
    HIPC(hipGetLastError());
    HIPC(hipStreamSynchronize(h->stream));
    s->n += n;
    return 0;
}

int ebem_sur_num_samples(ebem_sur* s, int64_t* n) {
    if (!s || !n) return fail(EBEM_EINVAL, "null argument");
    std::lock_guard<std::mutex> lk(s->lock);
    *n = s->n;
    return 0;
}

int ebem_sur_get_samples(ebem_sur* s, int64_t first, int64_t n, float* x, float* y) {
    if (!s) return fail(EBEM_EINVAL, "null surrogate");
    std::lock_guard<std::mutex> lk(s->lock);
    if (first < 0 || n < 0 || first + n > s->n) return fail(EBEM_EINVAL, "sample range out of bounds");
    HIPC(hipSetDevice(s->device));
    if (x) HIPC(hipMemcpyAsync(x, s->d_X.p + first * kSurIn, sizeof(float) * n * kSurIn, hipMemcpyDeviceToHost, s->stream));
    if (y) HIPC(hipMemcpyAsync(y, s->d_Y.p + first * kSurOut, sizeof(float) * n * kSurOut, hipMemcpyDeviceToHost, s->stream));
    HIPC(hipStreamSynchronize(s->stream));
    return 0;
}

int ebem_sur_init(ebem_sur* s, int64_t ntrain, uint64_t seed) {
    if (!s) return fail(EBEM_EINVAL, "null surrogate");
    std::lock_guard<std::mutex> lk(s->lock);
    if (ntrain <= 1 || ntrain > s->n) return fail(EBEM_EINVAL, "ntrain %lld outside [2, %lld]", (long long)ntrain, (long long)s->n);
    HIPC(hipSetDevice(s->device));
    std::vector<float> X((size_t)ntrain * kSurIn), Y((size_t)ntrain * kSurOut);
    HIPC(hipMemcpyAsync(X.data(), s->d_X.p, sizeof(float) * X.size(), hipMemcpyDeviceToHost, s->stream));
    HIPC(hipMemcpyAsync(Y.data(), s->d_Y.p, sizeof(float) * Y.size(), hipMemcpyDeviceToHost, s->stream));
    HIPC(hipStreamSynchronize(s->stream));
    std::vector<float>& nm = s->norm;
    nm.assign(kSurNorm, 0.f);
    float* box = nm.data() + 2 * kSurIn + 2 * kSurOut;
    for (int c = 0; c < kSurIn; c++) {
        double m = 0, q = 0;
        float lo = X[c], hi = X[c];
        for (int64_t k = 0; k < ntrain; k++) {
            const double v = X[(size_t)k * kSurIn + c];
            m += v;
            q += v * v;
            lo = std::min(lo, X[(size_t)k * kSurIn + c]);
            hi = std::max(hi, X[(size_t)k * kSurIn + c]);
        }
        m /= (double)ntrain;
        const double sd = std::sqrt(std::max(0.0, q / (double)ntrain - m * m));
        nm[c] = (float)m;
        nm[kSurIn + c] = sd > 1e-6 ? (float)(1.0 / sd) : 1.0f;
        // a small margin: features of congruent meshes agree to rounding only
        const float pad = 1e-4f * (hi - lo) + 1e-6f * std::max(std::fabs(lo), std::fabs(hi));
        box[c] = lo - pad;
        box[kSurIn + c] = hi + pad;
    }
    for (int c = 0; c < kSurOut; c++) {
        double m = 0, q = 0;
        for (int64_t k = 0; k < ntrain; k++) {
            const double v = Y[(size_t)k * kSurOut + c];
            m += v;
            q += v * v;
        }
        m /= (double)ntrain;
        const double sd = std::sqrt(std::max(0.0, q / (double)ntrain - m * m));
        nm[2 * kSurIn + c] = (float)m;
        nm[2 * kSurIn + kSurOut + c] = sd > 1e-12 ? (float)sd : 1.0f;
    }
    // He-normal weights over the real fan-in, zero biases and zero padding
    std::vector<float> p(kSurParams, 0.f);
    uint64_t st = seed ^ 0x9e3779b97f4a7c15ull;
    auto next = [&]() {  // splitmix64 -> uniform (0, 1)
        uint64_t z = (st += 0x9e3779b97f4a7c15ull);
        z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
        z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
        z ^= z >> 31;
        return ((double)(z >> 11) + 0.5) / 9007199254740992.0;
    };
    for (int l = 0; l < kSurLayers; l++) {
        const int K = sur_K(l), N = sur_N(l);
        const int kin = l == 0 ? kSurIn : K, nout = l == kSurLayers - 1 ? kSurOut : N;
        const double sd = (l == kSurLayers - 1) ? std::sqrt(1.0 / kin) : std::sqrt(2.0 / kin);
        for (int a = 0; a < kin; a++)
            for (int b = 0; b < nout; b++) {
                const double u1 = next(), u2 = next();
                p[sur_woff(l) + (int64_t)a * N + b] =
                    (float)(sd * std::sqrt(-2.0 * std::log(u1)) * std::cos(2.0 * kPi * u2));
            }
    }
    HIPC(hipMemcpyAsync(s->d_p.p, p.data(), sizeof(float) * kSurParams, hipMemcpyHostToDevice, s->stream));
    HIPC(hipMemcpyAsync(s->d_norm.p, nm.data(), sizeof(float) * kSurNorm, hipMemcpyHostToDevice, s->stream));
    HIPC(hipMemsetAsync(s->d_m.p, 0, sizeof(float) * kSurParams, s->stream));
    HIPC(hipMemsetAsync(s->d_v.p, 0, sizeof(float) * kSurParams, s->stream));
    hipLaunchKernelGGL(k_to_bf16, dim3((unsigned)((kSurParams + 255) / 256)), dim3(256), 0, s->stream, s->d_p.p,
                       s->d_pb.p, kSurParams);
    HIPC(hipGetLastError());
    HIPC(hipStreamSynchronize(s->stream));
    s->adam_t = 0;
    s->ready = true;
    return 0;
}

int ebem_sur_train(ebem_sur* s, int64_t ntrain, int epochs, int batch, double lr, uint64_t seed, double* loss_out) {
    if (!s) return fail(EBEM_EINVAL, "null surrogate");
    std::lock_guard<std::mutex> lk(s->lock);
    if (!s->ready) return fail(EBEM_EINVAL, "surrogate parameters not initialised (ebem_sur_init)");
    if (ntrain <= 0 || ntrain > s->n || epochs <= 0 || batch < 64 || batch % 64 || batch > 65536 || !(lr > 0))
        return fail(EBEM_EINVAL, "bad training arguments");
    HIPC(hipSetDevice(s->device));
    if (int r = sur_workspace(s, batch, true)) return r;
    const int64_t steps_per = (ntrain + batch - 1) / batch;
    const int64_t total = steps_per * epochs;
    std::vector<int32_t> perm(ntrain);
    DevBuf<int32_t> d_perm;
    HIPC(d_perm.alloc(ntrain));
    uint64_t st = seed;
    auto next = [&]() {
        uint64_t z = (st += 0x9e3779b97f4a7c15ull);
        z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
        z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
        return z ^ (z >> 31);
    };
    const float b1 = 0.9f, b2 = 0.999f, eps = 1e-8f;
    int64_t step = 0;
    for (int ep = 0; ep < epochs; ep++) {
        for (int64_t k = 0; k < ntrain; k++) perm[k] = (int32_t)k;
        for (int64_t k = ntrain - 1; k > 0; k--) std::swap(perm[k], perm[next() % (uint64_t)(k + 1)]);
        HIPC(hipMemcpyAsync(d_perm.p, perm.data(), sizeof(int32_t) * ntrain, hipMemcpyHostToDevice, s->stream));
        HIPC(hipMemsetAsync(s->d_loss.p, 0, sizeof(double), s->stream));
        for (int64_t b0 = 0; b0 < ntrain; b0 += batch, step++) {
            const int nb = (int)std::min<int64_t>(batch, ntrain - b0);
            if (int r = sur_batch_loss(s, s->stream, s->d_X.p, s->d_Y.p, d_perm.p + b0, nb, batch, true)) return r;
            // cosine decay from lr to lr / 100
            const double frac = (double)step / (double)std::max<int64_t>(1, total - 1);
            const double lrt = lr * (0.01 + 0.99 * 0.5 * (1.0 + std::cos(kPi * frac)));
            s->adam_t++;
            const float bc1 = (float)(1.0 - std::pow((double)b1, (double)s->adam_t));
            const float bc2 = (float)(1.0 - std::pow((double)b2, (double)s->adam_t));
            hipLaunchKernelGGL(k_adam, dim3((unsigned)((kSurParams + 255) / 256)), dim3(256), 0, s->stream, s->d_p.p,
                               s->d_g.p, s->d_m.p, s->d_v.p, s->d_pb.p, kSurParams, (float)lrt, b1, b2, eps, bc1, bc2);
            HIPC(hipGetLastError());
        }
        double l = 0;
        HIPC(hipMemcpyAsync(&l, s->d_loss.p, sizeof(double), hipMemcpyDeviceToHost, s->stream));
        HIPC(hipStreamSynchronize(s->stream));
        // the loss kernel averages per batch; report the mean over the epoch's batches
        if (loss_out) loss_out[ep] = l / (double)steps_per;
    }
    return 0;
}

int ebem_sur_loss(ebem_sur* s, int64_t first, int64_t n, double* loss) {
    if (!s || !loss) return fail(EBEM_EINVAL, "null argument");
    std::lock_guard<std::mutex> lk(s->lock);
    if (!s->ready) return fail(EBEM_EINVAL, "surrogate parameters not set");
    if (first < 0 || n <= 0 || first + n > s->n) return fail(EBEM_EINVAL, "sample range out of bounds");
    HIPC(hipSetDevice(s->device));
    const int B = 16384;
    if (int r = sur_workspace(s, B, false)) return r;
    std::vector<int32_t> idx(B);
    DevBuf<int32_t> d_idx;
    HIPC(d_idx.alloc(B));
    double tot = 0;
    for (int64_t k0 = 0; k0 < n; k0 += B) {
        const int nb = (int)std::min<int64_t>(B, n - k0);
        for (int k = 0; k < nb; k++) idx[k] = (int32_t)(first + k0 + k);
        HIPC(hipMemcpyAsync(d_idx.p, idx.data(), sizeof(int32_t) * nb, hipMemcpyHostToDevice, s->stream));
        HIPC(hipMemsetAsync(s->d_loss.p, 0, sizeof(double), s->stream));
        if (int r = sur_batch_loss(s, s->stream, s->d_X.p, s->d_Y.p, d_idx.p, nb, ((nb + 63) / 64) * 64, false)) return r;
        double l = 0;
        HIPC(hipMemcpyAsync(&l, s->d_loss.p, sizeof(double), hipMemcpyDeviceToHost, s->stream));
        HIPC(hipStreamSynchronize(s->stream));
        tot += l * (double)nb;
    }
    *loss = tot / (double)n;
    return 0;
}

int ebem_sur_grad(ebem_sur* s, int n, const float* x, const float* y, float* grads, double* loss) {
    if (!s || !x || !y) return fail(EBEM_EINVAL, "null argument");
    std::lock_guard<std::mutex> lk(s->lock);
    if (!s->ready) return fail(EBEM_EINVAL, "surrogate parameters not set");
    if (n <= 0 || n > 65536) return fail(EBEM_EINVAL, "n must be in [1, 65536]");
    HIPC(hipSetDevice(s->device));
    const int B = ((n + 63) / 64) * 64;
    if (int r = sur_workspace(s, B, true)) return r;
    DevBuf<float> dx, dy;
    DevBuf<int32_t> di;
    HIPC(dx.alloc((size_t)n * kSurIn));
    HIPC(dy.alloc((size_t)n * kSurOut));
    HIPC(di.alloc(n));
    std::vector<int32_t> idx(n);
    for (int k = 0; k < n; k++) idx[k] = k;
    HIPC(hipMemcpyAsync(dx.p, x, sizeof(float) * n * kSurIn, hipMemcpyHostToDevice, s->stream));
    HIPC(hipMemcpyAsync(dy.p, y, sizeof(float) * n * kSurOut, hipMemcpyHostToDevice, s->stream));
    HIPC(hipMemcpyAsync(di.p, idx.data(), sizeof(int32_t) * n, hipMemcpyHostToDevice, s->stream));
    HIPC(hipMemsetAsync(s->d_loss.p, 0, sizeof(double), s->stream));
    if (int r = sur_batch_loss(s, s->stream, dx.p, dy.p, di.p, n, B, true)) return r;
    if (grads) HIPC(hipMemcpyAsync(grads, s->d_g.p, sizeof(float) * kSurParams, hipMemcpyDeviceToHost, s->stream));
    double l = 0;
    HIPC(hipMemcpyAsync(&l, s->d_loss.p, sizeof(double), hipMemcpyDeviceToHost, s->stream));
    HIPC(hipStreamSynchronize(s->stream));
    if (loss) *loss = l;
    return 0;
}

int ebem_sur_predict(ebem_sur* s, int64_t n, const float* x, float* y) {
    if (!s || (n > 0 && (!x || !y))) return fail(EBEM_EINVAL, "null argument");
    std::lock_guard<std::mutex> lk(s->lock);
    if (!s->ready) return fail(EBEM_EINVAL, "surrogate parameters not set");
    if (n < 0) return fail(EBEM_EINVAL, "negative count");
    HIPC(hipSetDevice(s->device));
    const int B = 16384;
    if (int r = sur_workspace(s, B, false)) return r;
    DevBuf<float> dx, dyz;
    DevBuf<int32_t> di;
    HIPC(dx.alloc((size_t)B * kSurIn));
    HIPC(dyz.alloc((size_t)B * kSurOut));
    HIPC(di.alloc(B));
    std::vector<int32_t> idx(B);
    for (int k = 0; k < B; k++) idx[k] = k;
    HIPC(hipMemcpyAsync(di.p, idx.data(), sizeof(int32_t) * B, hipMemcpyHostToDevice, s->stream));
    HIPC(hipMemsetAsync(dyz.p, 0, sizeof(float) * B * kSurOut, s->stream));
    std::vector<float> Yo((size_t)B * kSurOutPad);
    const float* ym = s->norm.data() + 2 * kSurIn;
    const float* ys = ym + kSurOut;
    for (int64_t k0 = 0; k0 < n; k0 += B) {
        const int nb = (int)std::min<int64_t>(B, n - k0);
        const int Bp = ((nb + 63) / 64) * 64;
        HIPC(hipMemcpyAsync(dx.p, x + k0 * kSurIn, sizeof(float) * nb * kSurIn, hipMemcpyHostToDevice, s->stream));
        if (int r = sur_batch_loss(s, s->stream, dx.p, dyz.p, di.p, nb, Bp, false)) return r;
        HIPC(hipMemcpyAsync(Yo.data(), s->d_Yo.p, sizeof(float) * (size_t)Bp * kSurOutPad, hipMemcpyDeviceToHost,
                            s->stream));
        HIPC(hipStreamSynchronize(s->stream));
        for (int k = 0; k < nb; k++)
            for (int c = 0; c < kSurOut; c++) y[(k0 + k) * kSurOut + c] = Yo[(size_t)k * kSurOutPad + c] * ys[c] + ym[c];
    }
    return 0;
}

int ebem_sur_params(ebem_sur* s, int set, float* params, float* norm) {
    if (!s) return fail(EBEM_EINVAL, "null surrogate");
    std::lock_guard<std::mutex> lk(s->lock);
    HIPC(hipSetDevice(s->device));
    if (set) {
        if (!params || !norm) return fail(EBEM_EINVAL, "set needs params and norm");
        s->norm.assign(norm, norm + kSurNorm);
        HIPC(hipMemcpyAsync(s->d_p.p, params, sizeof(float) * kSurParams, hipMemcpyHostToDevice, s->stream));
        HIPC(hipMemcpyAsync(s->d_norm.p, norm, sizeof(float) * kSurNorm, hipMemcpyHostToDevice, s->stream));
        HIPC(hipMemsetAsync(s->d_m.p, 0, sizeof(float) * kSurParams, s->stream));
        HIPC(hipMemsetAsync(s->d_v.p, 0, sizeof(float) * kSurParams, s->stream));
        hipLaunchKernelGGL(k_to_bf16, dim3((unsigned)((kSurParams + 255) / 256)), dim3(256), 0, s->stream, s->d_p.p,
                           s->d_pb.p, kSurParams);
        HIPC(hipGetLastError());
        HIPC(hipStreamSynchronize(s->stream));
        s->adam_t = 0;
        s->ready = true;
        return 0;
    }
    if (!s->ready) return fail(EBEM_EINVAL, "surrogate parameters not set");
    if (params)
        HIPC(hipMemcpyAsync(params, s->d_p.p, sizeof(float) * kSurParams, hipMemcpyDeviceToHost, s->stream));
    if (norm) std::memcpy(norm, s->norm.data(), sizeof(float) * kSurNorm);
    HIPC(hipStreamSynchronize(s->stream));
    return 0;
}

int ebem_attach_surrogate(ebem_handle* h, ebem_sur* s) {
    if (int e = check_handle(h)) return e;
    std::lock_guard<std::mutex> lk(h->lock);
    if (int e = pp_join(h)) return e;
    if (s && s->device != h->device) return fail(EBEM_EINVAL, "surrogate and handle live on different devices");
    if (h->sur) {
        std::lock_guard<std::mutex> lk2(h->sur->lock);
        h->sur->users--;
    }
    h->sur = s;
    if (s) {
        std::lock_guard<std::mutex> lk2(s->lock);
        s->users++;
    }
    return 0;
}

int ebem_surrogate_stats(ebem_handle* h, int64_t* items, int64_t* blocks, double* ms) {
    if (int e = check_handle(h)) return e;
    std::lock_guard<std::mutex> lk(h->lock);
    if (int e = pp_join(h, true)) return e;
    if (!h->assembled) return fail(EBEM_EINVAL, "not assembled");
    if (items) *items = h->sur_items;
    if (blocks) *blocks = h->sur_pairs_blocks;
    if (ms) *ms = h->sur_ms;
    return 0;
}

}  // extern "C"
