// Op-counting host build of the device SA math (elastodynamicebem_amd/csrc/ebem_sa.hpp).
// Every double in that header becomes `cdouble`, whose arithmetic operators,
// sqrt and libm calls each add 1 to a global counter (a*b+c counts 2).
// Comparisons, conversions and copies are not counted.  Used to fix the
// algorithmic flops per (pair, level) item of the assembly kernel.
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
#include "../../elastodynamicebem_amd/csrc/ebem_sa.hpp"
#undef __builtin_isnan
#undef double

using namespace ebem;

extern "C" {

// Flops of one (pair, level) item: field triangle record rec (32 doubles),
// nGH outer nodes (x,y,z,w rows of 32), level m.
long long opcount_item(const double* rec, const double* gh, int nGH, double t, double cP, double cS) {
    FieldTri F;
    static_assert(sizeof(FieldTri) == 256, "layout");
    std::memcpy((void*)&F, rec, sizeof(FieldTri));
    g_ops = 0;
    Level K = make_level(cdouble(t), cdouble(cP), cdouble(cS));
    long long setup = g_ops;  // per-level constants: shared by the whole item
    NodeAcc acc;
    acc.s = 0.0;
    for (int q = 0; q < 6; q++) acc.m[q] = 0.0;
    for (int q = 0; q < nGH; q++) {
        cdouble x[3] = {gh[q], gh[32 + q], gh[64 + q]};
        NodeGeo G = node_geo(x, F);
        node_eval(G, K, cdouble(gh[96 + q]), acc);
    }
    cdouble W[6];
    finish_rotation(acc, F.C, W);
    return g_ops - setup + setup;
}

// The same item as the node-task path evaluates it: an outer node whose
// static value the previous level of the pair already holds (tprev >= 0 and
// node_static_margin at t and tprev) is not evaluated and costs nothing
// (the predicate is bookkeeping, not counted).
long long opcount_item_reuse(const double* rec, const double* gh, int nGH, double t, double tprev, double cP,
                             double cS) {
    FieldTri F;
    std::memcpy((void*)&F, rec, sizeof(FieldTri));
    g_ops = 0;
    Level K = make_level(cdouble(t), cdouble(cP), cdouble(cS));
    NodeAcc acc;
    acc.s = 0.0;
    for (int q = 0; q < 6; q++) acc.m[q] = 0.0;
    for (int q = 0; q < nGH; q++) {
        const long long before = g_ops;
        cdouble x[3] = {gh[q], gh[32 + q], gh[64 + q]};
        NodeGeo G = node_geo(x, F);
        if (tprev >= 0.0) {
            const long long keep = g_ops;
            const bool reused = node_static_margin(G, K) &&
                                node_static_margin(G, make_level(cdouble(tprev), cdouble(cP), cdouble(cS)));
            g_ops = keep;
            if (reused) {
                g_ops = before;
                continue;
            }
        }
        node_eval(G, K, cdouble(gh[96 + q]), acc);
    }
    cdouble W[6];
    finish_rotation(acc, F.C, W);
    return g_ops;
}

// Flops of one kP range [ta, tb] of k_levels_kp: node_eval_kp over the nGH
// nodes and the two rotations (the writes W_m = WP + t_m^2 WQ add 13 flops per
// level, counted by the caller).  -1 if a node fails the state check.
long long opcount_kp(const double* rec, const double* gh, int nGH, double ta, double tb, double cP, double cS) {
    FieldTri F;
    std::memcpy((void*)&F, rec, sizeof(FieldTri));
    Level KA = make_level(cdouble(ta), cdouble(cP), cdouble(cS));
    Level KB = make_level(cdouble(tb), cdouble(cP), cdouble(cS));
    g_ops = 0;
    NodeAcc aP, aQ;
    aP.s = aQ.s = 0.0;
    for (int q = 0; q < 6; q++) aP.m[q] = aQ.m[q] = 0.0;
    for (int q = 0; q < nGH; q++) {
        cdouble x[3] = {gh[q], gh[32 + q], gh[64 + q]};
        NodeGeo G = node_geo(x, F);
        if (!node_eval_kp(G, KA, KB, cdouble(gh[96 + q]), aP, aQ)) return -1;
    }
    cdouble WP[6], WQ[6];
    finish_rotation(aP, F.C, WP);
    finish_rotation(aQ, F.C, WQ);
    return g_ops;
}

// Flops of one between-front range of n levels from m0 as k_levels_bf
// evaluates it (node_eval_bf in chunks of kBfChunk, the rotations per level).
// -1 if a chunk fails the state check.
long long opcount_bf(const double* rec, const double* gh, int nGH, int m0, int n, double dt, double cP, double cS) {
    FieldTri F;
    std::memcpy((void*)&F, rec, sizeof(FieldTri));
    g_ops = 0;
    for (int c0 = 0; c0 < n; c0 += kBfChunk) {
        const int nc = n - c0 < kBfChunk ? n - c0 : kBfChunk;
        cdouble t[kBfChunk];
        for (int l = 0; l < kBfChunk; l++) t[l] = cdouble((double)(m0 + c0 + (l < nc ? l : 0)) * dt);
        NodeAcc acc[kBfChunk];
        for (int l = 0; l < kBfChunk; l++) {
            acc[l].s = 0.0;
            for (int q = 0; q < 6; q++) acc[l].m[q] = 0.0;
        }
        for (int q = 0; q < nGH; q++) {
            cdouble x[3] = {gh[q], gh[32 + q], gh[64 + q]};
            NodeGeo G = node_geo(x, F);
            if (!node_eval_bf<kBfChunk>(G, t, nc, cdouble(cP), cdouble(cS), cdouble(gh[96 + q]), acc)) return -1;
        }
        for (int l = 0; l < nc; l++) {
            cdouble W[6];
            finish_rotation(acc[l], F.C, W);
        }
    }
    return g_ops;
}
}
