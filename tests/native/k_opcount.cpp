// Op-counting host build of the device K math (elastodynamicebem_amd/csrc/ebem_k_math.hpp),
// as tests/native/opcount.cpp does for the SA math: every double becomes
// `cdouble`, whose arithmetic operators and sqrt each add 1 to a counter.
// Used to fix the algorithmic flops per evaluation point of k_kblocks.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdint>
#include <cstring>

static thread_local long long g_ops = 0;

struct cdouble {
    double v;
    cdouble() = default;
    constexpr cdouble(double x) : v(x) {}
    explicit operator int() const { return (int)v; }
    explicit operator double() const { return v; }
    cdouble& operator+=(cdouble o) { g_ops++; v += o.v; return *this; }
    cdouble& operator-=(cdouble o) { g_ops++; v -= o.v; return *this; }
    cdouble& operator*=(cdouble o) { g_ops++; v *= o.v; return *this; }
    cdouble operator-() const { return cdouble(-v); }
    cdouble operator+() const { return *this; }
};
static_assert(sizeof(cdouble) == 8, "cdouble must stay 8 bytes");
inline cdouble operator+(cdouble a, cdouble b) { g_ops++; return cdouble(a.v + b.v); }
inline cdouble operator-(cdouble a, cdouble b) { g_ops++; return cdouble(a.v - b.v); }
inline cdouble operator*(cdouble a, cdouble b) { g_ops++; return cdouble(a.v * b.v); }
inline cdouble operator/(cdouble a, cdouble b) { g_ops++; return cdouble(a.v / b.v); }
inline bool operator<(cdouble a, cdouble b) { return a.v < b.v; }
inline bool operator>(cdouble a, cdouble b) { return a.v > b.v; }
inline bool operator<=(cdouble a, cdouble b) { return a.v <= b.v; }
inline bool operator>=(cdouble a, cdouble b) { return a.v >= b.v; }
inline bool operator==(cdouble a, cdouble b) { return a.v == b.v; }
inline bool operator!=(cdouble a, cdouble b) { return a.v != b.v; }
#define CD1(f) inline cdouble f(cdouble a) { g_ops++; return cdouble(std::f(a.v)); }
CD1(sqrt) CD1(sin) CD1(cos) CD1(log) CD1(acos) CD1(asin) CD1(exp)
// sin and cos of one argument: one libm call
inline void sincos(cdouble a, cdouble* s, cdouble* c) {
    g_ops++;
    *s = cdouble(std::sin(a.v));
    *c = cdouble(std::cos(a.v));
}
inline cdouble fabs(cdouble a) { return cdouble(std::fabs(a.v)); }
inline cdouble floor(cdouble a) { return cdouble(std::floor(a.v)); }
inline cdouble ceil(cdouble a) { return cdouble(std::ceil(a.v)); }
inline cdouble atan2(cdouble a, cdouble b) { g_ops++; return cdouble(std::atan2(a.v, b.v)); }
inline cdouble fmin(cdouble a, cdouble b) { return cdouble(std::fmin(a.v, b.v)); }
inline cdouble fmax(cdouble a, cdouble b) { return cdouble(std::fmax(a.v, b.v)); }
inline bool cd_isnan(cdouble a) { return std::isnan(a.v); }

#define double cdouble
#define __builtin_isnan cd_isnan
#include "../../elastodynamicebem_amd/csrc/ebem_k_math.hpp"
#undef __builtin_isnan
#undef double

using namespace ebem;

namespace {
// Two triangles sharing vertex 0: S = (0, 1, 2) and F = (0, 3, 4), apart.
const cdouble kX[15] = {0, 0, 0, 1, 0, 0, 0, 1, 0, -1, 0.2, 0.3, -0.3, -1, 0.1};
const int32_t kT[6] = {0, 1, 2, 0, 3, 4};
}  // namespace

extern "C" {

// Flops of one regular evaluation point (one outer node, one inner node of
// one field triangle) at lag `lag` (4 levels from lag 2 on).
long long k_opcount_reg_point(int lag) {
    KTri kt[2];
    const cdouble n[3] = {0, 0, 1};
    k_tri_basis(kX + 0, kX + 3, kX + 6, n, 0.5, kt[0]);
    k_tri_basis(kX + 0, kX + 9, kX + 12, n, 0.5, kt[1]);
    cdouble verts[18];
    for (int t = 0; t < 2; t++)
        for (int v = 0; v < 3; v++)
            for (int c = 0; c < 3; c++) verts[9 * t + 3 * v + c] = kX[3 * kT[3 * t + v] + c];
    const int32_t voff[2] = {0, 1};
    const int32_t vadj[1] = {1 * 4 + 0};  // vertex 0 in F = 1 (local 0)
    const cdouble rin[4] = {1.0 / 3, 1.0 / 3, 1.0 / 3, 1.0};
    cdouble g[128];
    for (int k = 0; k < 128; k++) g[k] = 0.0;
    g[0] = 0.3, g[32] = 0.3, g[64] = 0.0, g[96] = 0.5;
    KParams K{};
    K.N = 2, K.Nv = 5, K.LK = 64, K.nInt = 1, K.nDiag = 1;
    K.dt = 0.1, K.cP = 2, K.cS = 1, K.lam = 2, K.mu = 1, K.rho = 1, K.maxLen = 1.5, K.fourpidt = 1.2566;
    KBlockArgs A{kt, verts, voff, vadj, rin, rin, g, 1};
    cdouble acc[3][3];
    g_ops = 0;
    k_reg_lane(A, K, 0, 0, lag, 0, 1, acc);
    return g_ops;
}

// Flops of one evaluation point of the field-triangle form (one outer node,
// one inner node of F, F's three hat functions) over the nL levels from mA
// (this file is built with EBEM_KSF_CHUNK=2: the counted form shares a
// point's geometry between two levels).
long long k_opcount_sf_point(int mA, int nL) {
    KTri kt[2];
    const cdouble n[3] = {0, 0, 1};
    k_tri_basis(kX + 0, kX + 3, kX + 6, n, 0.5, kt[0]);
    k_tri_basis(kX + 0, kX + 9, kX + 12, n, 0.5, kt[1]);
    cdouble verts[18];
    for (int t = 0; t < 2; t++)
        for (int v = 0; v < 3; v++)
            for (int c = 0; c < 3; c++) verts[9 * t + 3 * v + c] = kX[3 * kT[3 * t + v] + c];
    const cdouble rin[4] = {1.0 / 3, 1.0 / 3, 1.0 / 3, 1.0};
    cdouble g[128];
    for (int k = 0; k < 128; k++) g[k] = 0.0;
    g[0] = 0.3, g[32] = 0.3, g[64] = 0.0, g[96] = 0.5;
    KParams K{};
    K.N = 2, K.Nv = 5, K.LK = 64, K.nInt = 1, K.nDiag = 1;
    K.dt = 0.1, K.cP = 2, K.cS = 1, K.lam = 2, K.mu = 1, K.rho = 1, K.maxLen = 1.5, K.fourpidt = 1.2566;
    KBlockArgs A{kt, verts, nullptr, nullptr, rin, rin, g, 1};
    cdouble acc[3][kKsfChunk][9];
    g_ops = 0;
    k_sf_lane(A, K, 1, mA, nL, 0, 1, acc);
    return g_ops;
}

// Flops of one level of the singular sum for one outer node and one node of
// the collapsed rule in each of the three children (3 evaluation points).
long long k_opcount_sing_level() {
    KTri kt[1];
    const cdouble n[3] = {0, 0, 1};
    k_tri_basis(kX + 0, kX + 3, kX + 6, n, 0.5, kt[0]);
    cdouble verts[9];
    for (int v = 0; v < 3; v++)
        for (int c = 0; c < 3; c++) verts[3 * v + c] = kX[3 * kT[v] + c];
    const cdouble rdg[4] = {0.2, 0.3, 0.5, 1.0};
    cdouble g[128];
    for (int k = 0; k < 128; k++) g[k] = 0.0;
    g[0] = 0.3, g[32] = 0.3, g[64] = 0.0, g[96] = 0.5;
    KParams K{};
    K.N = 1, K.Nv = 3, K.LK = 64, K.nInt = 1, K.nDiag = 1;
    K.dt = 0.1, K.cP = 2, K.cS = 1, K.lam = 2, K.mu = 1, K.rho = 1, K.maxLen = 1.5, K.fourpidt = 1.2566;
    KBlockArgs A{kt, verts, nullptr, nullptr, rdg, rdg, g, 1};
    cdouble lev[3][3];
    g_ops = 0;
    k_sing_lane(A, K, 0, 0, 0.7, 0, 1, lev);
    return g_ops;
}
}
