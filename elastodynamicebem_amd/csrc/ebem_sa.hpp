// ebem_sa.hpp -- device-side semi-analytic (SA) inner integration of the
// single-layer retarded potential over one flat field triangle, for one outer
// point and one time level.
//
// Mathematical content follows the reference's SA path
// (functions/BEMenerg_indir/BEMenerg_core/BEMenerg_coreSA/calcoloMatrice/ and
// functions/BEMenerg_analiticIntg/); the evaluation is reorganised for the GPU:
//
//  * the six kernel integrals ris1..ris6 of calcMatrixSubBlock.m:34-39 are
//    never formed.  ris1/ris3/ris5 are multiples of the identity and ris2/4/6
//    are linear in the six basis matrices B_1..B_6 of coeffB.m:32-42, so each
//    region routine folds its contribution straight into one scalar (identity
//    part) and six basis coefficients, already weighted by the level
//    coefficients of calcMatrixSubBlock.m:101-104;
//  * B_1..B_6 are sparse; the child matrix is built once per child from the six
//    coefficients (6 symmetric entries, not 6 dense 3x3 products) and rotated
//    once by the child angle theta0 (fieldTriangle.m:57-70);
//  * the inverse Euler rotation (rotazioneInversaEulero.m) depends only on the
//    field triangle: it is applied once per (pair, level) by the caller, after
//    the Gauss-Hammer sum over outer nodes, instead of 3x per node.
//
// Branch structure, thresholds and the region decomposition E1..E9 are those of
// calcIntgInt_childTriangle.m:62-234 (kept exactly: they define the result).
#pragma once
#include <hip/hip_runtime.h>

// The reference evaluates every expression unfused; several closed forms here
// are differences of nearly equal terms (thin child fans, 1/(3F) prefactors of
// coeffG_TC.m:164-180), where a fused multiply-add changes the result by up to
// 1e-8 relative.  Contraction is therefore off for the whole header (restored
// at its end); libm-level calls keep their own internal FMAs.
#pragma clang fp contract(off)

namespace ebem {

constexpr double kPi = 3.141592653589793238462643383279502884;

// Per-triangle record (256 B, one cache-line pair), built on the host.
struct FieldTri {
    double v0[3];               // first vertex
    double L1, L3, cosd, sind;  // calcoloSistemaRiferimento.m: |P2-P1|, |P3-P1|, cos/sin(angle e1,e3)
    double e1[3], e2[3], n[3];  // in-plane frame and unit normal
    double C[9];                // inverse Euler rotation, row-major, entries < 1e-14 zeroed
    double center[3];
    double maxL;                // longest edge
    double area;
    double pad[2];
};
static_assert(sizeof(FieldTri) == 256, "FieldTri must stay 256 B");

// Level (time) constants.  s = k1*r1 + k3*r3 + k5*r5 and
// M = k2*ris2 + k4*ris4 + k6*ris6 reproduce calcMatrixSubBlock.m:101-104.
template <typename R>
struct LevelT {
    R t, rP, rS;
    R k1, k3, k5, k2, k4, k6;
};
using Level = LevelT<double>;

__host__ __device__ inline Level make_level(double t, double cP, double cS) {
    Level K;
    double cP2 = cP * cP, cS2 = cS * cS, den = (cP * cS) * (cP * cS);
    K.t = t;
    K.rP = cP * t;
    K.rS = cS * t;
    K.k1 = 0.5 / cP2;
    K.k2 = -0.5 / cP2;
    K.k3 = -0.5 * t * t;
    K.k4 = 1.5 * t * t;
    K.k5 = 0.5 * (cP2 + cS2) / den;
    K.k6 = 0.5 * (cP2 - cS2) / den;
    return K;
}

// Geometry of one child triangle (calcoloDatiTriangolo.m), angles at the
// origin vertex (gamma), second vertex (beta), third vertex (alpha).
template <typename R>
struct ChildGeoT {
    R a, b, c;
    R ca, cb, cg, sa, sb, sg;
    R al, be, ga;
};
using ChildGeo = ChildGeoT<double>;

template <typename R>
__host__ __device__ inline ChildGeoT<R> cast_geo(const ChildGeo& g) {
    return ChildGeoT<R>{R(g.a), R(g.b), R(g.c), R(g.ca), R(g.cb), R(g.cg), R(g.sa), R(g.sb), R(g.sg),
                        R(g.al), R(g.be), R(g.ga)};
}

template <typename R>
__host__ __device__ inline LevelT<R> cast_level(const Level& K) {
    return LevelT<R>{R(K.t), R(K.rP), R(K.rS), R(K.k1), R(K.k3), R(K.k5), R(K.k2), R(K.k4), R(K.k6)};
}

__host__ __device__ inline double clamp1(double x) { return fmin(1.0, fmax(-1.0, x)); }
__host__ __device__ inline double sqrt0(double x) { return x < 0.0 ? 0.0 : sqrt(x); }
__host__ __device__ inline double msign(double x) { return x > 0.0 ? 1.0 : (x < 0.0 ? -1.0 : 0.0); }
template <typename R>
__host__ __device__ inline R clampR(R x) { return fmin(R(1), fmax(R(-1), x)); }

// Angular integrals Phi(1..5) of the region routines (e.g. CC.m:34-38).
template <typename R>
struct Phi {
    R p1, p2, p3, p4, p5;
};


template <typename R>
__host__ __device__ __attribute__((always_inline)) inline Phi<R> make_phi(R th1, R th2, const ChildGeoT<R>& g) {
    R T1t = th2 - th1, T2t = th2 + th1;
    R s1t = sin(T1t);
    R sh1 = sin(T1t * R(0.5));
    Phi<R> P;
    P.p1 = R(2.0) * sin(T2t * R(0.5)) * sh1;
    P.p2 = R(2.0) * sin(g.ga - T2t * R(0.5)) * sh1;
    P.p3 = R(0.5) * (T1t - cos(T2t) * s1t);
    P.p4 = -R(0.5) * (-(s1t * cos(g.ga - T2t)) + T1t * g.cg);
    P.p5 = R(0.5) * (T1t - cos(R(2.0) * g.ga - T2t) * s1t);
    return P;
}

// sin and cos of one argument from one argument reduction: the same bits as
// separate sin and cos calls (checked on the device), at about half the cost.
__host__ __device__ __attribute__((always_inline)) inline void sincos_r(double x, double& s, double& c) {
    sincos(x, &s, &c);
}
__host__ __device__ __attribute__((always_inline)) inline void sincos_r(float x, float& s, float& c) {
    sincosf(x, &s, &c);
}

// Accumulator of one child: identity part and coefficients of B_1..B_6.
struct ChildAcc {
    double s;
    double b[6];
};

// ---- sector / annulus routines -------------------------------------------

// BEMenerg_analiticIntg_CC.m (P annulus between the S and P circles, inner
// radius RS/rhoS) and BEMenerg_analiticIntg_SC_P.m (inner radius collapsed to
// the plane distance: RS = zeta, rhoS = 0), folded with the level weights.
template <typename R>
__host__ __device__ inline void region_annulus(bool cc, bool zpos, R RS, R rhoS, R RP, R rhoP,
                                      R zeta, R th1, R th2, const Phi<R>& P,
                                      const LevelT<R>& K, ChildAcc& A) {
    R T1t = th2 - th1;
    if (cc && !zpos) {
        // CC.m:59-91, coplanar
        R I1 = RP - RS;
        R I4 = -(R(1.0) / RP) + (R(1.0) / RS);
        A.s += K.k1 * (T1t * I1) + K.k3 * (T1t * I4);
        R w = K.k2 * I1 + K.k4 * I4;
        A.b[3] += w * P.p5;
        A.b[4] += w * P.p4;
        A.b[5] += w * P.p3;
        return;
    }
    R I1, I2, I3, I4, I5, I6, I7;
    R iRP = R(1.0) / RP, iRP3 = iRP * iRP * iRP, z2 = zeta * zeta;
    R qP = rhoP * iRP;
    if (cc) {  // CC.m:25-31
        R iRS = R(1.0) / RS, iRS3 = iRS * iRS * iRS;
        R qS = rhoS * iRS;
        I1 = RP - RS;
        I2 = RP - RS + z2 * (iRP - iRS);
        I3 = log((rhoP + RP) / (rhoS + RS)) - qP + qS;
        I4 = -iRP + iRS;
        I5 = -iRP + iRS + (z2 / R(3.0)) * (iRP3 - iRS3);
        I6 = (R(1.0) / (R(3.0) * z2)) * (qP * qP * qP - qS * qS * qS);
        I7 = -R(1.0) / R(3.0) * (iRP3 - iRS3);
    } else {  // SC_P.m:20-26
        I1 = RP - zeta;
        I2 = RP + z2 * iRP - R(2.0) * zeta;
        I3 = log((rhoP + RP) / zeta) - qP;
        I4 = -iRP + R(1.0) / zeta;
        I5 = -iRP + z2 / (R(3.0) * RP * RP * RP) + R(2.0) / (R(3.0) * zeta);
        I6 = (R(1.0) / (R(3.0) * z2)) * (qP * qP * qP);
        I7 = -(R(1.0) / (R(3.0) * RP * RP * RP)) + R(1.0) / (R(3.0) * zeta * z2);
    }
    A.s += K.k1 * (T1t * I1) + K.k3 * (T1t * I4);
    A.b[0] += K.k2 * (T1t * I4) + K.k4 * (T1t * I7);
    R w3 = K.k2 * I3 + K.k4 * I6;
    A.b[1] += w3 * P.p2;
    A.b[2] += w3 * P.p1;
    R w2 = K.k2 * I2 + K.k4 * I5;
    A.b[3] += w2 * P.p5;
    A.b[4] += w2 * P.p4;
    A.b[5] += w2 * P.p3;
}

// BEMenerg_analiticIntg_SC_SP.m: S disk sector, kernels 1/r and r_i r_j/r^3.
template <typename R>
__host__ __device__ inline void region_sdisk(bool zpos, R RS, R rhoS, R zeta, R th1, R th2,
                                    const Phi<R>& P, const LevelT<R>& K, ChildAcc& A) {
    R T1t = th2 - th1;
    if (zpos) {
        R I1 = RS - zeta;
        R I2 = RS + zeta * zeta / RS - R(2.0) * zeta;
        R I3 = log((rhoS + RS) / zeta) - rhoS / RS;
        R I4 = -(R(1.0) / RS) + R(1.0) / zeta;
        A.s += K.k5 * (T1t * I1);
        A.b[0] += K.k6 * (T1t * I4);
        A.b[1] += K.k6 * I3 * P.p2;
        A.b[2] += K.k6 * I3 * P.p1;
        A.b[3] += K.k6 * I2 * P.p5;
        A.b[4] += K.k6 * I2 * P.p4;
        A.b[5] += K.k6 * I2 * P.p3;
    } else {
        A.s += K.k5 * (T1t * RS);
        A.b[3] += K.k6 * RS * P.p5;
        A.b[4] += K.k6 * RS * P.p4;
        A.b[5] += K.k6 * RS * P.p3;
    }
}

// ---- regions bounded by the child's far edge (Milroy-type G_i) ------------

// Edge-bounded terms shared by coeffG_TC / coeffG_TR_P / coeffG_TR_SP for
// zeta > 1e-6: g1..g8 and the angle functions of TH1, TH2.
template <typename R>
struct EdgeG {
    R F, g1, g2, g3, g4, g5, g7, g8, ADD;
    R c1b, c2b, s1b, s2b;
};

template <typename R>
__host__ __device__ inline EdgeG<R> make_edge(R zeta, R th1, R th2, const ChildGeoT<R>& g) {
    EdgeG<R> E;
    E.F = g.a * g.sb;
    sincos_r(th1 + g.be, E.s1b, E.c1b);
    sincos_r(th2 + g.be, E.s2b, E.c2b);
    R R1 = E.F / E.s1b, R2 = E.F / E.s2b;
    R fz1 = zeta / R1, fz2 = zeta / R2;
    E.ADD = zeta * zeta + E.F * E.F;
    E.g1 = sqrt(R(1.0) + fz1 * fz1);
    E.g2 = sqrt(R(1.0) + fz2 * fz2);
    E.g3 = R(0.5) * (log((E.g1 + E.c1b) * (E.g2 - E.c2b)) - log((E.g1 - E.c1b) * (E.g2 + E.c2b)));
    R i1 = R(1.0) / fz1, i2 = R(1.0) / fz2;
    E.g4 = log(i1 + sqrt(R(1.0) + i1 * i1));
    E.g5 = log(i2 + sqrt(R(1.0) + i2 * i2));
    R sq = sqrt(E.ADD);
    E.g7 = R(kPi) - acos(clampR((-zeta * E.c2b) / sq)) - acos(clampR((zeta * E.c1b) / sq));
    E.g8 = E.F * E.F / E.ADD;
    return E;
}

// The parts of region_edgeP that do not depend on the level (the angular
// interval, the child and the plane distance fix them): evaluated once, they
// serve every level whose region structure is the same.
template <typename R>
struct EdgePre {
    R T1t, F, z2, iz, g3, g7, h5, h6, dcgt, dct, f7, f8, f9, u15, u16, v17, v18, v19, w14;
};

template <typename R>
__host__ __device__ inline EdgePre<R> edge_pre(R zeta, R th1, R th2, const ChildGeoT<R>& g) {
    EdgePre<R> X;
    X.T1t = th2 - th1;
    EdgeG<R> E = make_edge<R>(zeta, th1, th2, g);
    const R F = E.F;
    X.F = F;
    X.g3 = E.g3;
    X.g7 = E.g7;
    R ca_g_1 = cos(g.al - g.ga + th1), ca_g_2 = cos(g.al - g.ga + th2);
    R cta1 = cos(th1 + g.al), cta2 = cos(th2 + g.al);
    R ctb1 = cos(th1 - g.be), ctb2 = cos(th2 - g.be);
    R cgt1 = cos(g.ga - th1), cgt2 = cos(g.ga - th2);
    R ct1 = cos(th1), ct2 = cos(th2);
    X.z2 = zeta * zeta;
    X.iz = R(1.0) / zeta;
    X.h5 = cgt2 * E.g5 - cgt1 * E.g4 - g.ca * E.g3;
    X.h6 = -ct2 * E.g5 + ct1 * E.g4 - g.cb * E.g3;
    X.dcgt = cgt2 - cgt1;
    X.dct = ct2 - ct1;
    X.f7 = F * (-E.g1 * ca_g_1 + E.g2 * ca_g_2 + E.g3 * g.sa * g.sa);
    X.f8 = F * (-g.sa * g.sb * E.g3 + E.g1 * cta1 - E.g2 * cta2);
    X.f9 = F * (g.sb * g.sb * E.g3 + E.g1 * ctb1 - E.g2 * ctb2);
    X.u15 = (g.sa * E.s2b - E.g8 * g.ca * E.c2b) / E.g2 + (-g.sa * E.s1b + E.g8 * g.ca * E.c1b) / E.g1;
    X.u16 = (g.sb * E.s1b + E.g8 * g.cb * E.c1b) / E.g1 + (-g.sb * E.s2b - E.g8 * g.cb * E.c2b) / E.g2;
    X.v17 = (-(E.s2b * E.s2b) * ca_g_2 + E.g8 * (g.ca * g.ca) * E.c2b) / E.g2 -
            (-(E.s1b * E.s1b) * ca_g_1 + E.g8 * (g.ca * g.ca) * E.c1b) / E.g1;
    X.v18 = ((E.s2b * E.s2b) * cta2 + E.g8 * (g.ca * g.cb * E.c2b)) / E.g2 -
            ((E.s1b * E.s1b) * cta1 + E.g8 * (g.ca * g.cb * E.c1b)) / E.g1;
    X.v19 = ((E.s2b * E.s2b) * ctb2 + E.g8 * (g.cb * g.cb * E.c2b)) / E.g2 -
            ((E.s1b * E.s1b) * ctb1 + E.g8 * (g.cb * g.cb * E.c1b)) / E.g1;
    X.w14 = F / E.ADD * (E.c1b / E.g1 - E.c2b / E.g2);
    return X;
}

// The level-dependent rest of region_edgeP: TC (P region between the S circle
// and the edge, coeffG_TC.m) when tc is true, TR_P (P region between the plane
// foot and the edge, coeffG_TR_P.m) otherwise; zeta > 1e-6 branch of both.
template <typename R>
__host__ __device__ inline void edge_fin(bool tc, R RS, R rhoS, R zeta, const EdgePre<R>& X, const ChildGeoT<R>& g,
                                         const Phi<R>& P, const LevelT<R>& K, ChildAcc& A) {
    const R T1t = X.T1t, F = X.F, z2 = X.z2, iz = X.iz;
    R G1, G4, G5, G6, G7, G8, G9, G14, G15, G16, G17, G18, G19;
    if (tc) {  // coeffG_TC.m:44-136
        R lnS = (rhoS / RS) - log((rhoS + RS) * iz);
        R q = (RS * RS + z2) / RS;
        R iRS3 = R(1.0) / (R(3.0) * RS * RS * RS);
        R rS3 = rhoS * rhoS * rhoS / (R(3.0) * z2 * RS * RS * RS);
        R tail = (R(1.0) / RS) - z2 * iRS3;
        G1 = -(RS * T1t) + zeta * X.g7 + F * X.g3;
        G4 = (R(1.0) / RS) * T1t - iz * X.g7;
        G5 = X.h5 + lnS * X.dcgt;
        G6 = X.h6 - lnS * X.dct;
        G7 = zeta * X.g7 - q * P.p5 + X.f7;
        G8 = -zeta * g.cg * X.g7 - q * P.p4 + X.f8;
        G9 = zeta * X.g7 - q * P.p3 + X.f9;
        G14 = iRS3 * T1t + R(1.0) / (R(3.0) * z2) * (X.w14 - iz * X.g7);
        G15 = -rS3 * P.p2 + R(1.0) / (R(3.0) * z2) * X.u15;
        G16 = -rS3 * P.p1 + R(1.0) / (R(3.0) * z2) * X.u16;
        G17 = R(1.0) / R(3.0) * X.v17 / F - X.g7 / (R(3.0) * zeta) + tail * P.p5;
        G18 = R(1.0) / R(3.0) * X.v18 / F + X.g7 / (R(3.0) * zeta) * g.cg + tail * P.p4;
        G19 = R(1.0) / R(3.0) * X.v19 / F - X.g7 / (R(3.0) * zeta) + tail * P.p3;
    } else {  // coeffG_TR_P.m:63-119
        G1 = zeta * (-T1t + X.g7) + F * X.g3;
        G4 = iz * (T1t - X.g7);
        G5 = X.h5;
        G6 = X.h6;
        G7 = zeta * (X.g7 - R(2.0) * P.p5) + X.f7;
        G8 = -zeta * (g.cg * X.g7 + R(2.0) * P.p4) + X.f8;
        G9 = zeta * (X.g7 - R(2.0) * P.p3) + X.f9;
        G14 = R(1.0) / (R(3.0) * z2) * (iz * (T1t - X.g7) + X.w14);
        G15 = R(1.0) / (R(3.0) * z2) * X.u15;
        G16 = R(1.0) / (R(3.0) * z2) * X.u16;
        G17 = R(1.0) / (R(3.0) * F) * X.v17 - R(1.0) / (R(3.0) * zeta) * (X.g7 - R(2.0) * P.p5);
        G18 = R(1.0) / (R(3.0) * F) * X.v18 + R(1.0) / (R(3.0) * zeta) * (X.g7 * g.cg + R(2.0) * P.p4);
        G19 = R(1.0) / (R(3.0) * F) * X.v19 - R(1.0) / (R(3.0) * zeta) * (X.g7 - R(2.0) * P.p3);
    }
    // TC.m:18-29 / TR_P.m:65-76
    A.s += K.k1 * G1 + K.k3 * G4;
    A.b[0] += K.k2 * G4 + K.k4 * G14;
    A.b[1] += K.k2 * G5 + K.k4 * G15;
    A.b[2] += K.k2 * G6 + K.k4 * G16;
    A.b[3] += K.k2 * G7 + K.k4 * G17;
    A.b[4] += K.k2 * G8 + K.k4 * G18;
    A.b[5] += K.k2 * G9 + K.k4 * G19;
}

template <typename R>
__host__ __device__ inline void region_edgeP(bool tc, R RS, R rhoS, R zeta, R th1, R th2,
                                   const ChildGeoT<R>& g, const Phi<R>& P, const LevelT<R>& K, ChildAcc& A) {
    const EdgePre<R> X = edge_pre<R>(zeta, th1, th2, g);
    edge_fin<R>(tc, RS, rhoS, zeta, X, g, P, K, A);
}

// Coplanar (zeta <= 1e-6) branch of coeffG_TC.m:138-181 (TC.m:31-49), split
// like region_edgeP into the level-independent terms and the rest: each G is
// (level-independent term) +/- (term in RS), the same expression tree as the
// one-piece form.
template <typename R>
struct CopPre {
    R T1t, Fg3, X4, Y7, Y8, Y9, Z17, Z18, Z19;
};

template <typename R>
__host__ __device__ inline CopPre<R> cop_pre(R th1, R th2, const ChildGeoT<R>& g) {
    CopPre<R> X;
    X.T1t = th2 - th1;
    const R F = g.a * g.sb;
    R c1b, c2b, s1b, s2b;
    sincos_r(th1 + g.be, s1b, c1b);
    sincos_r(th2 + g.be, s2b, c2b);
    R g_3 = R(0.5) * (log((R(1.0) + c1b) * (R(1.0) - c2b)) - log((R(1.0) - c1b) * (R(1.0) + c2b)));
    R ca_g_1 = cos(g.al - g.ga + th1), ca_g_2 = cos(g.al - g.ga + th2);
    R cta1 = cos(th1 + g.al), cta2 = cos(th2 + g.al);
    R ctb1 = cos(g.be - th1), ctb2 = cos(g.be - th2);
    X.Fg3 = F * g_3;
    X.X4 = R(1.0) / F * (c2b - c1b);
    X.Y7 = F * (ca_g_2 - ca_g_1 + g.sa * g.sa * g_3);
    X.Y8 = F * (cta1 - cta2 - g.sa * g.sb * g_3);
    X.Y9 = F * (ctb1 - ctb2 + g.sb * g.sb * g_3);
    R i3F = R(1.0) / (R(3.0) * F);
    X.Z17 = i3F * (((R(1.0) + g.ca * g.ca) * c2b - (s2b * s2b) * ca_g_2) -
                   ((R(1.0) + g.ca * g.ca) * c1b - (s1b * s1b) * ca_g_1));
    R cab = g.ca * g.cb + cos(g.al + g.be);
    X.Z18 = i3F * ((c2b * cab + (s2b * s2b) * cta2) - (c1b * cab + (s1b * s1b) * cta1));
    X.Z19 = i3F * (((R(1.0) + g.cb * g.cb) * c2b + (s2b * s2b) * ctb2) -
                   ((R(1.0) + g.cb * g.cb) * c1b + (s1b * s1b) * ctb1));
    return X;
}

template <typename R>
__host__ __device__ inline void cop_fin(R RS, const CopPre<R>& X, const Phi<R>& P, const LevelT<R>& K,
                                        ChildAcc& A) {
    R iRS = R(1.0) / RS;
    R G1 = X.Fg3 - RS * X.T1t;
    R G4 = X.X4 + iRS * X.T1t;
    R G7 = -RS * P.p5 + X.Y7;
    R G8 = -RS * P.p4 + X.Y8;
    R G9 = -RS * P.p3 + X.Y9;
    R G17 = X.Z17 + iRS * P.p5;
    R G18 = X.Z18 + iRS * P.p4;
    R G19 = X.Z19 + iRS * P.p3;
    A.s += K.k1 * G1 + K.k3 * G4;
    A.b[3] += K.k2 * G7 + K.k4 * G17;
    A.b[4] += K.k2 * G8 + K.k4 * G18;
    A.b[5] += K.k2 * G9 + K.k4 * G19;
}

template <typename R>
__host__ __device__ inline void region_tc_coplanar(R RS, R th1, R th2, const ChildGeoT<R>& g,
                                          const Phi<R>& P, const LevelT<R>& K, ChildAcc& A) {
    const CopPre<R> X = cop_pre<R>(th1, th2, g);
    cop_fin<R>(RS, X, P, K, A);
}

// TR_SP (S disk part bounded by the edge, coeffG_TR_SP.m, TR_SP.m).
template <typename R>
__host__ __device__ inline void region_edgeS(bool zpos, R zeta, R th1, R th2, const ChildGeoT<R>& g, const Phi<R>& P,
                                    const LevelT<R>& K, ChildAcc& A) {
    R T1t = th2 - th1;
    if (zpos) {
        EdgeG<R> E = make_edge<R>(zeta, th1, th2, g);
        const R F = E.F;
        R G1 = zeta * (-T1t + E.g7) + F * E.g3;
        R G4 = R(1.0) / zeta * (T1t - E.g7);
        R G5 = cos(g.ga - th2) * E.g5 - cos(g.ga - th1) * E.g4 - g.ca * E.g3;
        R G6 = -cos(th2) * E.g5 + cos(th1) * E.g4 - g.cb * E.g3;
        R G7 = zeta * (E.g7 - R(2.0) * P.p5) +
                    F * (-E.g1 * cos(g.al - g.ga + th1) + E.g2 * cos(g.al - g.ga + th2) + E.g3 * g.sa * g.sa);
        R G8 = -zeta * (g.cg * E.g7 + R(2.0) * P.p4) +
                    F * (-g.sa * g.sb * E.g3 + E.g1 * cos(th1 + g.al) - E.g2 * cos(th2 + g.al));
        R G9 = zeta * (E.g7 - R(2.0) * P.p3) +
                    F * (g.sb * g.sb * E.g3 + E.g1 * cos(th1 - g.be) - E.g2 * cos(th2 - g.be));
        A.s += K.k5 * G1;
        A.b[0] += K.k6 * G4;
        A.b[1] += K.k6 * G5;
        A.b[2] += K.k6 * G6;
        A.b[3] += K.k6 * G7;
        A.b[4] += K.k6 * G8;
        A.b[5] += K.k6 * G9;
    } else {
        const R F = g.a * g.sb;
        R c1b = cos(th1 + g.be), c2b = cos(th2 + g.be);
        R g_3 = R(0.5) * (log((R(1.0) + c1b) * (R(1.0) - c2b)) - log((R(1.0) - c1b) * (R(1.0) + c2b)));
        R G1 = F * g_3;
        R G7 = F * (-cos(g.al - g.ga + th1) + cos(g.al - g.ga + th2) + g.sa * g.sa * g_3);
        R G8 = F * (-(g.sa * g.sb * g_3) + cos(th1 + g.al) - cos(th2 + g.al));
        R G9 = F * (g.sb * g.sb * g_3 + cos(th1 - g.be) - cos(th2 - g.be));
        A.s += K.k5 * G1;
        A.b[3] += K.k6 * G7;
        A.b[4] += K.k6 * G8;
        A.b[5] += K.k6 * G9;
    }
}

// One region routine of functions/BEMenerg_analiticIntg/ on its own, with the
// argument lists of BEMenerg_analiticIntg_{CC,SC_P,SC_SP,TC,TR_P,TR_SP}.m:
// kind 0..5 in that order; a = {RS, rhoS, RP, rhoP, zeta, TH1, TH2, -} (each
// routine reads the ones its .m file takes); info = the twelve fields of
// infoTri2D (calcoloDatiTriangolo.m:10-32) in ChildGeo order; B = B(:,:,1..6)
// of coeffB.m, column-major.  ris = ris1..ris4 (P routines) or ris5, ris6 (S
// routines, the last two matrices zero), column-major.  Each ris is the region
// routine run with one level weight set to 1 and the others 0, so its
// identity part and B-basis coefficients are exactly the G_i / I_i * Phi_i of
// the .m file; ris = s I + sum_q b_q B_{q+1}.
__host__ __device__ inline void ai_region(int kind, const double* a, const double* info, const double* B,
                                          double* ris) {
    const ChildGeo g{info[0], info[1], info[2], info[3], info[4], info[5],
                     info[6], info[7], info[8], info[9], info[10], info[11]};
    const double RS = a[0], rhoS = a[1], RP = a[2], rhoP = a[3], zeta = a[4], th1 = a[5], th2 = a[6];
    const bool zpos = zeta > 1.0e-6;
    const bool sreg = (kind == 2 || kind == 5);
    const Phi<double> P = make_phi<double>(th1, th2, g);
    for (int r = 0; r < 4; r++) {
        double* o = ris + 9 * r;
        for (int q = 0; q < 9; q++) o[q] = 0.0;
        if (r >= (sreg ? 2 : 4)) continue;
        Level K{};
        if (sreg) (r == 0 ? K.k5 : K.k6) = 1.0;
        else (r == 0 ? K.k1 : r == 1 ? K.k2 : r == 2 ? K.k3 : K.k4) = 1.0;
        ChildAcc A{0.0, {0.0, 0.0, 0.0, 0.0, 0.0, 0.0}};
        switch (kind) {
            case 0: region_annulus<double>(true, zpos, RS, rhoS, RP, rhoP, zeta, th1, th2, P, K, A); break;
            case 1: region_annulus<double>(false, zpos, RS, rhoS, RP, rhoP, zeta, th1, th2, P, K, A); break;
            case 2: region_sdisk<double>(zpos, RS, rhoS, zeta, th1, th2, P, K, A); break;
            case 3:
                if (zpos) region_edgeP<double>(true, RS, rhoS, zeta, th1, th2, g, P, K, A);
                else region_tc_coplanar<double>(RS, th1, th2, g, P, K, A);
                break;
            case 4: region_edgeP<double>(false, RS, rhoS, zeta, th1, th2, g, P, K, A); break;
            default: region_edgeS<double>(zpos, zeta, th1, th2, g, P, K, A); break;
        }
        if (r == 0 || r == 2) {  // ris1, ris3, ris5: multiples of the identity
            o[0] = o[4] = o[8] = A.s;
        } else {
            for (int q = 0; q < 6; q++)
                for (int e = 0; e < 9; e++) o[e] = o[e] + A.b[q] * B[9 * q + e];
        }
    }
}

enum : int { kNone = 0, kSPSP = 1, kSPP = 2, kP = 3 };

// calcoloDatiTriangolo.m:1-39 for the child (O, V2, V3), O at the origin.
__host__ __device__ inline ChildGeo child_geo(double v2x, double v2y, double v3x, double v3y) {
    // No contraction: cos(gamma) of a thin child (distant outer point) comes from
    // the law of cosines with heavy cancellation; an FMA here changes the
    // result by ~1e-8 relative where the unfused form matches the reference.
#pragma clang fp contract(off)
    ChildGeo g;
    g.a = sqrt(v2x * v2x + v2y * v2y);
    g.b = sqrt(v3x * v3x + v3y * v3y);
    double dx = v3x - v2x, dy = v3y - v2y;
    g.c = sqrt(dx * dx + dy * dy);
    const double a = g.a, b = g.b, c = g.c;
    g.ca = (b * b + c * c - a * a) / (2.0 * b * c);
    g.cb = (a * a + c * c - b * b) / (2.0 * a * c);
    g.cg = (a * a + b * b - c * c) / (2.0 * a * b);
    g.sa = sqrt0(1.0 - g.ca * g.ca);
    g.sb = sqrt0(1.0 - g.cb * g.cb);
    g.sg = sqrt0(1.0 - g.cg * g.cg);
    g.al = acos(clamp1(g.ca));
    g.be = acos(clamp1(g.cb));
    g.ga = acos(clamp1(g.cg));
    return g;
}

// calcIntgInt_childTriangle.m:27-50: the angles TH2..TH5 that bound the
// regions of a child (TH1 = 0, TH6 = gamma), from the in-plane front radii;
// front cutting is always FP64.  Rotated child: V2 = (a,0), V3 = (b cg, b sg).
struct FrontAngles {
    double TH2, TH3, TH4, TH5;
};

__host__ __device__ __attribute__((always_inline)) inline FrontAngles front_angles(const ChildGeo& g, double rhoP,
                                                                                   double rhoS, int type) {
    const double a = g.a;
    double vt0 = (g.b * g.cg - a) / g.c;
    double Bm = a * vt0;
    double DeltaP = Bm * Bm - (a * a - rhoP * rhoP);
    double DeltaS = Bm * Bm - (a * a - rhoS * rhoS);
    double rP = a * g.sb / rhoP;
    double rS = a * g.sb / rhoS;
    if (type == kP) {
        rS = 0.0;
        DeltaS = 0.0;
    }
    const double gamma = g.ga, beta = g.be;
    FrontAngles f;
    if (DeltaP <= 1.0e-6) {
        f.TH2 = f.TH3 = f.TH4 = f.TH5 = gamma;
    } else if (DeltaS <= 1.0e-6) {
        double as = asin(clamp1(rP));
        double O1 = as - beta, O2 = kPi - as - beta;
        double th1 = fmin(O1, O2), th2 = fmax(O1, O2);
        f.TH2 = fmin(fmax(0.0, th1), gamma);
        f.TH3 = fmax(0.0, fmin(th2, gamma));
        f.TH4 = f.TH3;
        f.TH5 = f.TH3;
    } else {
        double as = asin(clamp1(rP)), bs = asin(clamp1(rS));
        double O1 = as - beta, O2 = kPi - as - beta, O3 = bs - beta, O4 = kPi - bs - beta;
        f.TH2 = fmin(fmax(0.0, fmin(O1, O2)), gamma);
        f.TH3 = fmin(fmax(0.0, fmin(O3, O4)), gamma);
        f.TH4 = fmax(0.0, fmin(fmax(O3, O4), gamma));
        f.TH5 = fmax(0.0, fmin(fmax(O1, O2), gamma));
    }
    return f;
}

// calcIntgInt_childTriangle.m:9-234 for one child, result folded into A.
template <typename R>
__host__ __device__ inline void child_regions(const ChildGeo& g, double zeta, double RP, double rhoP, double RS,
                                     double rhoS, int type, const Level& Kd, ChildAcc& A) {
    const FrontAngles fa = front_angles(g, rhoP, rhoS, type);
    const double TH1 = 0.0, TH2 = fa.TH2, TH3 = fa.TH3, TH4 = fa.TH4, TH5 = fa.TH5, TH6 = g.ga;
    const bool sp = (type == kSPSP);
    const bool zpos = zeta > 1.0e-6;
    // Closed-form region integrals in R (FP64 for the plain form, FP32 for the
    // far field of the mixed-precision form).
    const ChildGeoT<R> gr = cast_geo<R>(g);
    const LevelT<R> K{R(Kd.t), R(Kd.rP), R(Kd.rS), R(Kd.k1), R(Kd.k3), R(Kd.k5), R(Kd.k2), R(Kd.k4), R(Kd.k6)};
    const R rs = R(RS), rhs = R(rhoS), rp = R(RP), rhp = R(rhoP), z = R(zeta);
    // One Phi per angular interval, shared by the S-sector and P routines on it
    // (identical arguments, identical values).
    // The two sector + annulus intervals (E1/E6 on [TH1, TH2], E5/E9 on
    // [TH5, TH6]) run the same code, and so do the two intervals bounded by
    // the edge (E2/E7 on [TH2, TH3], E4/E8 on [TH4, TH5]; TH5 == TH3 in class
    // P, where E4 is empty): each pair is one loop whose body a wavefront
    // executes as often as its busiest lane needs it, instead of once per
    // interval that any lane needs (region branches are the main divergence
    // of the assembly and post-processing kernels).  The regions are added in
    // the order E1, E5, E2, E4, E3.
    const int nA = (TH1 < TH2 ? 1 : 0) + (TH5 < TH6 ? 1 : 0);
#pragma unroll 1
    for (int it = 0; it < nA; it++) {
        const bool first = (it == 0) && (TH1 < TH2);
        const R lo = R(first ? TH1 : TH5), hi = R(first ? TH2 : TH6);
        Phi<R> P = make_phi<R>(lo, hi, gr);
        if (sp) region_sdisk<R>(zpos, rs, rhs, z, lo, hi, P, K, A);
        region_annulus<R>(type != kP, zpos, rs, rhs, rp, rhp, z, lo, hi, P, K, A);
    }
    const int nB = (TH2 < TH3 ? 1 : 0) + (TH4 < TH5 ? 1 : 0);
#pragma unroll 1
    for (int it = 0; it < nB; it++) {
        const bool first = (it == 0) && (TH2 < TH3);
        const R lo = R(first ? TH2 : TH4), hi = R(first ? TH3 : TH5);
        Phi<R> P = make_phi<R>(lo, hi, gr);
        if (sp) region_sdisk<R>(zpos, rs, rhs, z, lo, hi, P, K, A);
        if (type == kP) region_edgeP<R>(false, rs, rhs, z, lo, hi, gr, P, K, A);
        else if (!zpos) region_tc_coplanar<R>(rs, lo, hi, gr, P, K, A);
        else region_edgeP<R>(true, rs, rhs, z, lo, hi, gr, P, K, A);
    }
    if (TH3 < TH4 && sp) {  // E3 (S region bounded by the edge)
        Phi<R> P = make_phi<R>(R(TH3), R(TH4), gr);
        region_edgeS<R>(zpos, z, R(TH3), R(TH4), gr, P, K, A);
    }
}

// Result of one outer point: identity coefficient and symmetric local-frame
// matrix (m00, m01, m02, m11, m12, m22), before the inverse Euler rotation.
struct NodeAcc {
    double s;
    double m[6];
};

// geo/calcoloDistanzaTriangolo.m:1-52 (MATLAB min ignores NaN).
__host__ __device__ inline double dist_tri(const double t[3][2]) {
#pragma clang fp contract(off)
    double D[3];
    for (int i = 0; i < 3; i++) D[i] = sqrt(t[i][0] * t[i][0] + t[i][1] * t[i][1]);
    double best = __builtin_nan("");
    const int ea[3] = {0, 1, 0}, eb[3] = {1, 2, 2};
    for (int i = 0; i < 3; i++) {
        const double* P1 = t[ea[i]];
        const double* P2 = t[eb[i]];
        double d1 = D[ea[i]], d2 = D[eb[i]];
        double vx = P2[0] - P1[0], vy = P2[1] - P1[1];
        double L = sqrt(vx * vx + vy * vy);
        double dm;
        if (-(P1[0] * vx + P1[1] * vy) < -1.0e-6) {
            dm = d1;
        } else if (-(P2[0] * (-vx) + P2[1] * (-vy)) < -1.0e-6) {
            dm = d2;
        } else {
            double ca = (-P1[0] / d1) * (vx / L) + (-P1[1] / d1) * (vy / L);
            dm = d1 * sqrt0(1.0 - ca * ca);
        }
        if (!__builtin_isnan(dm) && (__builtin_isnan(best) || dm < best)) best = dm;
    }
    return best;
}

// Projection of the outer point on the field plane and the fan of three child
// triangles around it (geo/preparazioneTriangolo.m:1-115).
struct NodeGeo {
    double zeta, sgn, dMIN;
    double t[3][2];  // field triangle vertices in the plane frame, origin at the projection
    double c[3];     // signed (doubled) child areas
};

__host__ __device__ inline NodeGeo node_geo(const double* x, const FieldTri& F) {
#pragma clang fp contract(off)
    NodeGeo G;
    double d0 = x[0] - F.v0[0], d1 = x[1] - F.v0[1], d2 = x[2] - F.v0[2];
    double ps = F.n[0] * d0 + F.n[1] * d1 + F.n[2] * d2;
    double sp0 = x[0] - ps * F.n[0], sp1 = x[1] - ps * F.n[1], sp2 = x[2] - ps * F.n[2];
    double sgp = (x[0] - sp0) * F.n[0] + (x[1] - sp1) * F.n[1] + (x[2] - sp2) * F.n[2];
    if (fabs(sgp) <= 1.0e-6) sgp = 0.0;
    G.sgn = msign(sgp);
    double z = fabs(ps);
    G.zeta = (z < 1.0e-6) ? 0.0 : z;
    double f0 = F.v0[0] - sp0, f1 = F.v0[1] - sp1, f2 = F.v0[2] - sp2;
    G.t[0][0] = f0 * F.e1[0] + f1 * F.e1[1] + f2 * F.e1[2];
    G.t[0][1] = f0 * F.e2[0] + f1 * F.e2[1] + f2 * F.e2[2];
    G.t[1][0] = G.t[0][0] + F.L1;
    G.t[1][1] = G.t[0][1] + 0.0;
    G.t[2][0] = G.t[0][0] + F.L3 * F.cosd;
    G.t[2][1] = G.t[0][1] + F.L3 * F.sind;
    G.c[0] = (G.t[0][0] * G.t[1][1]) - (G.t[0][1] * G.t[1][0]);
    G.c[1] = (G.t[1][0] * G.t[2][1]) - (G.t[1][1] * G.t[2][0]);
    G.c[2] = (G.t[2][0] * G.t[0][1]) - (G.t[2][1] * G.t[0][0]);
    double A2 = (G.t[1][0] - G.t[0][0]) * (G.t[2][1] - G.t[0][1]);
    double e2 = G.c[2] / A2, e3 = G.c[0] / A2;
    double e1 = 1.0 - e2 - e3;
    bool in = (-1.0e-6 <= e1 && e1 <= 1.0) && (-1.0e-6 <= e2 && e2 <= 1.0) && (-1.0e-6 <= e3 && e3 <= 1.0);
    G.dMIN = in ? 0.0 : dist_tri(G.t);
    return G;
}

// calcMatrixSubBlock.m:58-81: in-plane front radii and integration class.
__host__ __device__ inline int classify(double zeta, double rP, double rS, double dMIN, double& rhoP, double& rhoS) {
#pragma clang fp contract(off)
    if (zeta == 0.0) {
        rhoP = rP;
        rhoS = rS;
    } else if (zeta < rS) {
        rhoP = sqrt(rP * rP - zeta * zeta);
        rhoS = sqrt(rS * rS - zeta * zeta);
    } else if (rS < zeta && zeta < rP) {
        rhoP = sqrt(rP * rP - zeta * zeta);
        rhoS = -1.0;
    } else {
        rhoP = -1.0;
        rhoS = -1.0;
    }
    if (dMIN < rhoS) return kSPSP;
    if (0.0 < rhoS && rhoS <= dMIN && dMIN < rhoP) return kSPP;
    if (rhoS <= 0.0 && dMIN < rhoP) return kP;
    return kNone;
}

// The fan of three children around the projection, child k = (O, V_k,
// V_{k+1 mod 3}) in counter-clockwise order (fieldTriangle.m:22-33), walked
// child by child with the vertices and doubled areas rotated in registers, so
// that a loop over k indexes no array (an indexed fan is kept in scratch
// memory).  next() gives the child false if it is degenerate (|doubled area|
// < 1e-6), which the reference skips, and always moves on to the next.
struct Fan {
    double ax, ay, bx, by, cx, cy;  // V_k, V_{k+1}, V_{k+2}
    double ca, cb, cc;              // their doubled child areas c[k], c[k+1], c[k+2]
    __host__ __device__ __attribute__((always_inline)) explicit Fan(const NodeGeo& G)
        : ax(G.t[0][0]), ay(G.t[0][1]), bx(G.t[1][0]), by(G.t[1][1]), cx(G.t[2][0]), cy(G.t[2][1]),
          ca(G.c[0]), cb(G.c[1]), cc(G.c[2]) {}
    __host__ __device__ __attribute__((always_inline)) bool next(double& C, double& v2x, double& v2y, double& v3x,
                                                                 double& v3y) {
        C = ca;
        v2x = ax;
        v2y = ay;
        v3x = bx;
        v3y = by;
        const double tx = ax, ty = ay, tc = ca;
        ax = bx;
        ay = by;
        ca = cb;
        bx = cx;
        by = cy;
        cb = cc;
        cx = tx;
        cy = ty;
        cc = tc;
        if (fabs(C) < 1.0e-06) return false;
        if (C < 0.0) {
            const double sx = v2x, sy = v2y;
            v2x = v3x;
            v2y = v3y;
            v3x = sx;
            v3y = sy;
        }
        return true;
    }
};

// Folds one child's coefficients into the node accumulator: child matrix
// sum_i b_i B_i (coeffB.m:20-42, symmetric, child frame) rotated by
// A(theta0) * M * A(theta0)' (fieldTriangle.m:37-70), theta0 the clockwise
// angle of V2 (cos = V2x/|V2|, sin = -V2y/|V2|), weighted by w sign(C).
__host__ __device__ __attribute__((always_inline)) inline void add_child(const ChildAcc& A, const ChildGeo& g,
                                                                         double v2x, double v2y, double zs,
                                                                         double ws, NodeAcc& acc) {
    double isg = 1.0 / g.sg;
    double isg2 = isg * isg;
    double m00 = (A.b[3] + 2.0 * g.cg * A.b[4] + g.cg * g.cg * A.b[5]) * isg2;
    double m01 = (A.b[4] + g.cg * A.b[5]) * isg;
    double m11 = A.b[5];
    double m02 = -zs * (A.b[1] + g.cg * A.b[2]) * isg;
    double m12 = -zs * A.b[2];
    double m22 = zs * zs * A.b[0];
    double ct = v2x / g.a, st = -v2y / g.a;
    double r00 = ct * ct * m00 + 2.0 * ct * st * m01 + st * st * m11;
    double r01 = -ct * st * m00 + (ct * ct - st * st) * m01 + ct * st * m11;
    double r11 = st * st * m00 - 2.0 * ct * st * m01 + ct * ct * m11;
    double r02 = ct * m02 + st * m12;
    double r12 = -st * m02 + ct * m12;
    acc.s += ws * A.s;
    acc.m[0] += ws * r00;
    acc.m[1] += ws * r01;
    acc.m[2] += ws * r02;
    acc.m[3] += ws * r11;
    acc.m[4] += ws * r12;
    acc.m[5] += ws * m22;
}

// One outer point at one level: adds w * (s, M_local) into acc.
// Returns the integration class (0 when the point is skipped).
template <typename R = double>
__host__ __device__ __attribute__((always_inline)) inline int node_eval(const NodeGeo& G, const Level& K, double w, NodeAcc& acc) {
    double rhoP, rhoS;
    int type = classify(G.zeta, K.rP, K.rS, G.dMIN, rhoP, rhoS);
    if (type == kNone) return kNone;
    const double zeta = G.zeta;
    const double sgn = G.sgn;
    Fan fan(G);
#pragma unroll 1
    for (int k = 0; k < 3; k++) {
        double C, v2x, v2y, v3x, v3y;
        if (!fan.next(C, v2x, v2y, v3x, v3y)) continue;
        ChildGeo g = child_geo(v2x, v2y, v3x, v3y);
        ChildAcc A;
        A.s = 0.0;
        for (int i = 0; i < 6; i++) A.b[i] = 0.0;
        child_regions<R>(g, zeta, K.rP, rhoP, K.rS, rhoS, type, K, A);
        add_child(A, g, v2x, v2y, zeta * sgn, w * msign(C), acc);
    }
    return type;
}

// Static outer point: class kSPSP with both circles beyond the edge of every
// child of the fan (TH2 = TH3 = 0, TH4 = TH5 = gamma), so that child_regions
// evaluates the single region E3 over [0, gamma] (coeffG_TR_SP.m) with the
// level weights k5 and k6, which do not depend on t: node_eval then returns
// the same bits at every level where this holds.  Children node_eval skips
// (degenerate) are skipped here too.
__host__ __device__ inline bool node_static_pattern(const NodeGeo& G, const Level& K) {
    double rhoP, rhoS;
    if (classify(G.zeta, K.rP, K.rS, G.dMIN, rhoP, rhoS) != kSPSP) return false;
    Fan fan(G);
#pragma unroll 1
    for (int k = 0; k < 3; k++) {
        double C, v2x, v2y, v3x, v3y;
        if (!fan.next(C, v2x, v2y, v3x, v3y)) continue;
        const ChildGeo g = child_geo(v2x, v2y, v3x, v3y);
        const FrontAngles f = front_angles(g, rhoP, rhoS, kSPSP);
        if (!(f.TH2 == 0.0 && f.TH3 == 0.0 && f.TH4 == g.ga && f.TH5 == g.ga && g.ga > 0.0)) return false;
    }
    return true;
}

// Sufficient condition for node_static_pattern with a margin, from the
// projection's distances alone: class kSPSP and rho_S^2 above the largest
// squared vertex distance by 0.2% plus 1e-5.  Every child's vertices then lie
// strictly inside the S circle (and the P circle, which is larger), so the
// circle meets the child's edge line outside the edge on both sides and both
// discriminants of the front cutting exceed 1e-6.  It only grows with t.
__host__ __device__ inline bool node_static_margin(const NodeGeo& G, const Level& K) {
    double rhoP, rhoS;
    if (classify(G.zeta, K.rP, K.rS, G.dMIN, rhoP, rhoS) != kSPSP) return false;
    double dmax2 = 0.0;
    for (int k = 0; k < 3; k++) dmax2 = fmax(dmax2, G.t[k][0] * G.t[k][0] + G.t[k][1] * G.t[k][1]);
    return rhoS * rhoS > dmax2 * 1.002 + 1.0e-5;
}

// The same sum over the fan for an outer point that stays, over the levels
// [KA.t, KB.t], in the class kP of classify() (S sphere short of the field
// plane) with the P circle covering each child's far edge (TH2 = 0 and
// TH3 = TH5 = gamma in child_regions).  Each child then has the one region
// E7 over [0, gamma] (coeffG_TR_P.m), whose G coefficients do not depend on t,
// and every accumulation is linear in the level weights, so the node adds
// w (accP + t^2 accQ) at each of these levels: accP with the weights (k1, k2)
// and k3 = k4 = 0, accQ with k1 = k2 = 0 and (k3, k4) / t^2 = (-1/2, 3/2).
// The class is checked at both ends and the covering at KA (it only grows
// with t); returns false, leaving accP and accQ untouched, if either fails.
// The region integrals run in R (the front cutting and the sums stay FP64).
template <typename R = double>
__host__ __device__ __attribute__((always_inline)) inline bool node_eval_kp(const NodeGeo& G, const Level& KA,
                                                                            const Level& KB, double w,
                                                                            NodeAcc& accP, NodeAcc& accQ) {
    double rhoP, rhoS, rhoP2, rhoS2;
    if (classify(G.zeta, KA.rP, KA.rS, G.dMIN, rhoP, rhoS) != kP) return false;
    if (classify(G.zeta, KB.rP, KB.rS, G.dMIN, rhoP2, rhoS2) != kP) return false;
    const double zeta = G.zeta;
    const double zs = zeta * G.sgn;
    Level KPw = KA, KQw = KA;
    KPw.k3 = 0.0;
    KPw.k4 = 0.0;
    KQw.k1 = 0.0;
    KQw.k2 = 0.0;
    KQw.k3 = -0.5;
    KQw.k4 = 1.5;
    NodeAcc nP, nQ;
    nP.s = nQ.s = 0.0;
    for (int q = 0; q < 6; q++) nP.m[q] = nQ.m[q] = 0.0;
    Fan fan(G);
#pragma unroll 1
    for (int k = 0; k < 3; k++) {
        double C, v2x, v2y, v3x, v3y;
        if (!fan.next(C, v2x, v2y, v3x, v3y)) continue;
        ChildGeo g = child_geo(v2x, v2y, v3x, v3y);
        // front cutting of child_regions for class kP at KA
        const double a = g.a;
        const double vt0 = (g.b * g.cg - a) / g.c;
        const double Bm = a * vt0;
        const double DeltaP = Bm * Bm - (a * a - rhoP * rhoP);
        if (!(DeltaP > 1.0e-6)) return false;
        const double as = asin(clamp1(a * g.sb / rhoP));
        const double O1 = as - g.be, O2 = kPi - as - g.be;
        if (fmin(fmax(0.0, fmin(O1, O2)), g.ga) != 0.0 || fmax(0.0, fmin(fmax(O1, O2), g.ga)) != g.ga) return false;
        const ChildGeoT<R> gr = cast_geo<R>(g);
        const Phi<R> P = make_phi<R>(R(0.0), gr.ga, gr);
        ChildAcc AP, AQ;
        AP.s = AQ.s = 0.0;
        for (int i = 0; i < 6; i++) AP.b[i] = AQ.b[i] = 0.0;
        region_edgeP<R>(false, R(KA.rS), R(-1.0), R(zeta), R(0.0), gr.ga, gr, P, cast_level<R>(KPw), AP);
        region_edgeP<R>(false, R(KA.rS), R(-1.0), R(zeta), R(0.0), gr.ga, gr, P, cast_level<R>(KQw), AQ);
        const double ws = w * msign(C);
        add_child(AP, g, v2x, v2y, zs, ws, nP);
        add_child(AQ, g, v2x, v2y, zs, ws, nQ);
    }
    accP.s += nP.s;
    accQ.s += nQ.s;
    for (int q = 0; q < 6; q++) {
        accP.m[q] += nP.m[q];
        accQ.m[q] += nQ.m[q];
    }
    return true;
}

// Levels sharing one evaluation of the level-independent region terms in
// node_eval_bf for the assembly's between-front ranges (k_levels_bf at one wave
// per SIMD, 1280-triangle sphere: 6 levels 9.0 ms with 292 B/lane of spills,
// 8 levels 10.3 ms with 704 B/lane, 5 levels 9.6 ms, 4 levels without spills
// 10.7 ms: the shared level-independent terms against the spills).
#ifndef EBEM_BF_CHUNK
#define EBEM_BF_CHUNK 6
#endif
constexpr int kBfChunk = EBEM_BF_CHUNK;
// The same for the FP32 ranges of the mixed-precision form's far field
// (k_levels_bf<7, float>, 1280-triangle sphere: 10 levels 5.5 ms without
// spills, 6 levels 6.2 ms, 8 levels 6.2 ms with 100 B/lane of spills, 12
// levels 6.7 ms with 128 B/lane).
#ifndef EBEM_BF_CHUNK_F32
#define EBEM_BF_CHUNK_F32 10
#endif
constexpr int kBfChunkF32 = EBEM_BF_CHUNK_F32;

// Outer point between the fronts at each of the n <= C levels t[0..n) (in
// increasing order): the P circle covers each child's far edge (TH2 = 0 and
// TH5 = gamma in child_regions; checked at t[0], where the circle is smallest)
// and the S sphere is short of the field triangle (class kP or kSPP of
// classify() at every level), so each child has exactly one region, [0, gamma]:
// E7 by coeffG_TR_P.m in class kP, by coeffG_TC.m in class kSPP, by its
// coplanar branch when zeta = 0.  The level-independent terms (child
// geometry, angular integrals, edge functions) are evaluated once per child
// and finished per level with node_eval's own expressions, so acc[l] receives
// exactly what node_eval adds at t[l].  Returns false (acc then partly
// updated) if a level's class or the covering check does not hold.  The
// region integrals run in R (the front cutting and the sums stay FP64).
template <int C, typename R = double>
__host__ __device__ __attribute__((always_inline)) inline bool node_eval_bf(const NodeGeo& G, const double* t, int n,
                                                                            double cP, double cS, double w,
                                                                            NodeAcc* acc) {
    int type[C];
    double rhoS[C], rhoP0 = 0.0;
#pragma unroll
    for (int l = 0; l < C; l++) {
        type[l] = kP;
        rhoS[l] = 0.0;
        if (l < n) {
            const Level K = make_level(t[l], cP, cS);
            double rp, rs;
            type[l] = classify(G.zeta, K.rP, K.rS, G.dMIN, rp, rs);
            if (type[l] != kP && type[l] != kSPP) return false;
            if (l == 0) rhoP0 = rp;
            rhoS[l] = rs;
        }
    }
    const double zeta = G.zeta;
    const double zs = zeta * G.sgn;
    const bool zpos = zeta > 1.0e-6;
    Fan fan(G);
#pragma unroll 1
    for (int k = 0; k < 3; k++) {
        double Cs, v2x, v2y, v3x, v3y;
        if (!fan.next(Cs, v2x, v2y, v3x, v3y)) continue;
        const ChildGeo g = child_geo(v2x, v2y, v3x, v3y);
        // front cutting of child_regions by the P circle at t[0]
        const double a = g.a;
        const double vt0 = (g.b * g.cg - a) / g.c;
        const double Bm = a * vt0;
        const double DeltaP = Bm * Bm - (a * a - rhoP0 * rhoP0);
        if (!(DeltaP > 1.0e-6)) return false;
        const double as = asin(clamp1(a * g.sb / rhoP0));
        const double O1 = as - g.be, O2 = kPi - as - g.be;
        if (fmin(fmax(0.0, fmin(O1, O2)), g.ga) != 0.0 || fmax(0.0, fmin(fmax(O1, O2), g.ga)) != g.ga) return false;
        const ChildGeoT<R> gr = cast_geo<R>(g);
        const Phi<R> P = make_phi<R>(R(0.0), gr.ga, gr);
        const double ws = w * msign(Cs);
        if (zpos) {
            const EdgePre<R> X = edge_pre<R>(R(zeta), R(0.0), gr.ga, gr);
#pragma unroll
            for (int l = 0; l < C; l++) {
                if (l < n) {
                    const Level K = make_level(t[l], cP, cS);
                    ChildAcc A;
                    A.s = 0.0;
                    for (int i = 0; i < 6; i++) A.b[i] = 0.0;
                    edge_fin<R>(type[l] != kP, R(K.rS), R(rhoS[l]), R(zeta), X, gr, P, cast_level<R>(K), A);
                    add_child(A, g, v2x, v2y, zs, ws, acc[l]);
                }
            }
        } else {
            const CopPre<R> X = cop_pre<R>(R(0.0), gr.ga, gr);
#pragma unroll
            for (int l = 0; l < C; l++) {
                if (l < n) {
                    const Level K = make_level(t[l], cP, cS);
                    ChildAcc A;
                    A.s = 0.0;
                    for (int i = 0; i < 6; i++) A.b[i] = 0.0;
                    cop_fin<R>(R(K.rS), X, P, cast_level<R>(K), A);
                    add_child(A, g, v2x, v2y, zs, ws, acc[l]);
                }
            }
        }
    }
    return true;
}

// Distance from x to the triangle (v0, v1, v2) in 3D (vF: 9 doubles): the
// plane distance if the projection falls inside, else the nearest edge.
__host__ __device__ inline double tri_dist3(double x0, double x1, double x2, const double* vF) {
    double e[3][3], d[3];
    for (int c = 0; c < 3; c++) {
        e[0][c] = vF[3 + c] - vF[c];
        e[1][c] = vF[6 + c] - vF[3 + c];
        e[2][c] = vF[c] - vF[6 + c];
    }
    const double n0 = e[0][1] * e[1][2] - e[0][2] * e[1][1];
    const double n1 = e[0][2] * e[1][0] - e[0][0] * e[1][2];
    const double n2 = e[0][0] * e[1][1] - e[0][1] * e[1][0];
    bool inside = true;
    double best = 1e300;
    for (int k = 0; k < 3; k++) {
        const double* A = vF + 3 * k;
        const double p0 = x0 - A[0], p1 = x1 - A[1], p2 = x2 - A[2];
        // side of the edge within the plane: (e_k x p) . n
        const double c0 = e[k][1] * p2 - e[k][2] * p1, c1 = e[k][2] * p0 - e[k][0] * p2,
                     c2 = e[k][0] * p1 - e[k][1] * p0;
        if (c0 * n0 + c1 * n1 + c2 * n2 < 0.0) inside = false;
        const double ee = e[k][0] * e[k][0] + e[k][1] * e[k][1] + e[k][2] * e[k][2];
        double u = (p0 * e[k][0] + p1 * e[k][1] + p2 * e[k][2]) / ee;
        u = fmin(1.0, fmax(0.0, u));
        const double q0 = p0 - u * e[k][0], q1 = p1 - u * e[k][1], q2 = p2 - u * e[k][2];
        d[k] = sqrt(q0 * q0 + q1 * q1 + q2 * q2);
        best = fmin(best, d[k]);
    }
    if (inside) {
        const double nn = sqrt(n0 * n0 + n1 * n1 + n2 * n2);
        return fabs(((x0 - vF[0]) * n0 + (x1 - vF[1]) * n1 + (x2 - vF[2]) * n2) / nn);
    }
    return best;
}

// Levels [ka, kb] of a (test, field) pair whose W_m = WP + t_m^2 WQ
// (node_eval_kp), and the levels [sa, sb] after them that node_eval_bf
// evaluates (S sphere of every node short of the field triangle,
// c_S t_m < dist(x_q, T)): every outer node q in class kP, c_S t_m < zeta_q (its
// distance to the field plane, zeroed below 1e-6 as in node_geo), and the P
// sphere past every field vertex, c_P t_m > Dmax.  Relative margins 1e-7 keep
// the levels clear of both ends (node_eval_kp checks the state again).  Each
// field edge must have (c/2)^2 > 1e-5: P covering an edge of length c gives
// DeltaP >= (c/2)^2 in child_regions, clear of its 1e-6 threshold.  The range
// is cut to the pair's item levels [max(1, l-1), h+1] and dropped below 3
// levels.  gS: the test triangle's outer nodes as x, y, z rows of 32.
__host__ __device__ inline void kp_range(int nGH, double cP, double cS, double dt, const double* gS,
                                         const FieldTri& F, const double* vF, int l, int h, int& ka, int& kb,
                                         int& sa, int& sb) {
    ka = 1;
    kb = 0;
    sa = 1;
    sb = 0;
    double cmin2 = 1e300, dmax2 = 0.0, zmin = 1e300, d3min = 1e300;
    for (int v = 0; v < 3; v++) {
        const int w = (v + 1) % 3;
        double a = vF[3 * w] - vF[3 * v], b = vF[3 * w + 1] - vF[3 * v + 1], c = vF[3 * w + 2] - vF[3 * v + 2];
        cmin2 = fmin(cmin2, a * a + b * b + c * c);
    }
    if (!(0.25 * cmin2 > 1.0e-5)) return;
    for (int q = 0; q < nGH; q++) {
        const double x = gS[q], y = gS[32 + q], z = gS[64 + q];
        for (int v = 0; v < 3; v++) {
            double a = x - vF[3 * v], b = y - vF[3 * v + 1], c = z - vF[3 * v + 2];
            dmax2 = fmax(dmax2, a * a + b * b + c * c);
        }
        double d0 = x - F.v0[0], d1 = y - F.v0[1], d2 = z - F.v0[2];
        double zq = fabs(F.n[0] * d0 + F.n[1] * d1 + F.n[2] * d2);
        zmin = fmin(zmin, zq < 1.0e-6 ? 0.0 : zq);
        d3min = fmin(d3min, tri_dist3(x, y, z, vF));
    }
    const double fa = ceil(sqrt(dmax2) * (1.0 + 1e-7) / (cP * dt));
    const double fb = ceil(zmin * (1.0 - 1e-7) / (cS * dt)) - 1.0;
    const double fs = ceil((d3min * (1.0 - 1e-7) - 1e-9) / (cS * dt)) - 1.0;
    const int mlo = l - 1 < 1 ? 1 : l - 1;
    const double a = fmax(fa, (double)mlo), b = fmin(fb, (double)(h + 1));
    if (b - a + 1.0 >= 3.0) {
        ka = (int)a;
        kb = (int)b;
    }
    // levels after the kP range (or from the P covering on) with the S sphere
    // still short of the field triangle: node_eval_bf, two levels at least
    const double s0 = ka <= kb ? (double)(kb + 1) : a, s1 = fmin(fs, (double)(h + 1));
    if (s1 - s0 + 1.0 >= 2.0) {
        sa = (int)s0;
        sb = (int)s1;
    }
}

// W = s I + C' M C  (rotazioneInversaEulero.m:94), symmetric output
// (w00, w01, w02, w11, w12, w22).
__host__ __device__ inline void finish_rotation(const NodeAcc& acc, const double* Cm, double* W) {
    double M[3][3] = {{acc.m[0], acc.m[1], acc.m[2]}, {acc.m[1], acc.m[3], acc.m[4]}, {acc.m[2], acc.m[4], acc.m[5]}};
    double T[3][3];  // T = M C
    for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++) T[i][j] = M[i][0] * Cm[0 * 3 + j] + M[i][1] * Cm[1 * 3 + j] + M[i][2] * Cm[2 * 3 + j];
    const int ii[6] = {0, 0, 0, 1, 1, 2}, jj[6] = {0, 1, 2, 1, 2, 2};
    for (int q = 0; q < 6; q++) {
        int i = ii[q], j = jj[q];
        double v = Cm[0 * 3 + i] * T[0][j] + Cm[1 * 3 + i] * T[1][j] + Cm[2 * 3 + i] * T[2][j];
        W[q] = v + (i == j ? acc.s : 0.0);
    }
}

}  // namespace ebem

#pragma clang fp contract(fast)
