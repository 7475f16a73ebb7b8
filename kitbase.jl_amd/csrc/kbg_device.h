// kbg_device.h -- device-side physics and reduction helpers shared by the kbgpu kernels.
//
// Scalar macroscopic work (conserve_prim, Gaussian moments, pdf slopes, UGKS/KCU time
// coefficients) is done in FP64 on every lane of a wavefront; per-velocity work is done in the
// storage type T (float or double).  Formulas follow KitBase.jl v0.9.31; the citation is given on
// each helper.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <utility>

#define KBG_HD __host__ __device__ __forceinline__
#define KBG_D __device__ __forceinline__

namespace kbg {

constexpr double kPi = 3.14159265358979323846;

enum Flux { KFVS = 0, KCU = 1, UGKS = 2 };
enum Coll { BGK = 0, SHAKHOV = 1, FSM = 2 };
enum Limiter { VANLEER = 0, MINMOD = 1, SUPERBEE = 2, VANALBABA = 3, LINEAR = 4 };

// global-address-space view of a pointer read from a descriptor (so loads and stores are
// global_* instructions rather than flat_*)
template <class T> using gptr = __attribute__((address_space(1))) T*;
template <class T> KBG_D gptr<T> to_global(T* p) { return (gptr<T>)p; }
template <class T> KBG_D gptr<const T> to_global(const T* p) { return (gptr<const T>)p; }

// compile-time loop: f(std::integral_constant<int, i>) for i = 0 .. N-1, so register arrays are
// indexed by constant expressions from the front end on
template <class F, int... I> KBG_D void static_for_impl(F& f, std::integer_sequence<int, I...>) {
    (f(std::integral_constant<int, I>{}), ...);
}
template <int N, class F> KBG_D void static_for(F&& f) { static_for_impl(f, std::make_integer_sequence<int, N>{}); }

// Julia sign(): -1, +1, or y itself for +-0 and NaN (src/Math/math_general.jl:20 fortsign)
template <class T> KBG_HD T jsign(T y) { return y > T(0) ? T(1) : (y < T(0) ? T(-1) : y); }

// limiters   src/Reconstruction/recon_limiter.jl:10-105
template <class T> KBG_HD T limiter(int kind, T sL, T sR) {
    const T aL = sL < T(0) ? -sL : sL, aR = sR < T(0) ? -sR : sR;
    switch (kind) {
    case MINMOD: return T(0.5) * (jsign(sL) + jsign(sR)) * (aR < aL ? aR : aL);
    case SUPERBEE:
        if (sR >= T(0.5) * sL && sR <= T(2) * sL) return T(0.5) * (jsign(sL) + jsign(sR)) * (aL > aR ? aL : aR);
        if (sR < T(0.5) * sL && sR > T(2) * sL) return (jsign(sL) + jsign(sR)) * (aL < aR ? aL : aR);
        return T(0);
    case VANALBABA: return (sL * sL * sR + sL * sR * sR) / (sL * sL + sR * sR + T(1e-7));
    case LINEAR: return T(0.5) * (sL + sR);
    default: return (jsign(sL) + jsign(sR)) * aL * aR / (aL + aR + T(1e-7));
    }
}

// division in the storage precision: the hardware reciprocal for FP32, IEEE division for FP64
template <class T> KBG_D T tdiv(T a, T b);
template <> KBG_D float tdiv<float>(float a, float b) { return __fdividef(a, b); }
template <> KBG_D double tdiv<double>(double a, double b) { return a / b; }

// limiter evaluated per velocity; van Leer (the KitBase default) takes the fast division
template <class T> KBG_D T limf(int kind, T sL, T sR) {
    if (kind == VANLEER) {
        // (sign sL + sign sR) |sL| |sR| written without sign branches: sL |sR| + |sL| sR
        const T aL = fabs(sL), aR = fabs(sR);
        return tdiv<T>(sL * aR + aL * sR, aL + aR + T(1e-7));
    }
    return limiter<T>(kind, sL, sR);
}

// Failure detection: the health of a cell's new primitive variables -- 1 not finite, 2 rho <= 0,
// 4 lambda <= 0 (a NaN fails all three).  A bad cell raises these bits in its solve's flag word; the
// step itself completes.
template <int N> KBG_HD int prim_health(const double* p) {
    bool fin = true;
    for (int m = 0; m < N; m++) fin = fin && isfinite(p[m]);
    return (fin ? 0 : 1) | (p[0] > 0.0 ? 0 : 2) | (p[N - 1] > 0.0 ? 0 : 4);
}
KBG_D void flag_health(int* h, int bits) {
    if (bits && h) __hip_atomic_fetch_or(h, bits, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// conserve_prim (theory_macro.jl:185-206) for n = 3 (1D), 4 (2D), 5 (3V); 1D and 2D with one
// division by rho and one for lambda: U = m / rho, lambda = rho / (2 (gamma - 1) (E - m U / 2))
template <int N> KBG_HD void conserve_prim(const double* W, double gam, double* p) {
    if constexpr (N == 3) {
        const double r = 1.0 / W[0];
        p[0] = W[0];
        p[1] = W[1] * r;
        p[2] = 0.5 * W[0] / ((gam - 1.0) * (W[2] - 0.5 * (W[1] * p[1])));
    } else if constexpr (N == 4) {
        const double r = 1.0 / W[0];
        p[0] = W[0];
        p[1] = W[1] * r;
        p[2] = W[2] * r;
        p[3] = 0.5 * W[0] / ((gam - 1.0) * (W[3] - 0.5 * (W[1] * p[1] + W[2] * p[2])));
    } else {
        p[0] = W[0];
        p[1] = W[1] / W[0];
        p[2] = W[2] / W[0];
        p[3] = W[3] / W[0];
        p[4] = 0.5 * W[0] / (gam - 1.0) / (W[4] - 0.5 * (W[1] * W[1] + W[2] * W[2] + W[3] * W[3]) / W[0]);
    }
}

// vhs_collision_time   theory_relaxation.jl:28
KBG_HD double vhs_tau(double rho, double lam, double mu, double om) { return mu * 2.0 * pow(lam, 1.0 - om) / rho; }

// dt / tau of the BGK/Shakhov relaxation in the storage precision: FP64 keeps the library pow,
// FP32 evaluates lambda^(1-omega) with the hardware log2/exp2
template <class T> KBG_D T relax_ratio(double rho, double lam, double mu, double om, double dt);
template <> KBG_D double relax_ratio<double>(double rho, double lam, double mu, double om, double dt) {
    return dt / vhs_tau(rho, lam, mu, om);
}
template <> KBG_D float relax_ratio<float>(double rho, double lam, double mu, double om, double dt) {
    const float p = __expf(float(1.0 - om) * __logf(float(lam)));
    return float(dt) * float(rho) / (float(mu * 2.0) * p);
}

// Half-range Gaussian moments up to order 6 (theory_moments.jl:10-57) plus the internal-energy
// moments Mxi (:64-86).  NP is the primitive-vector length (3, 4 or 5).
template <int NP> struct GMom {
    double Mu[7], MuL[7], MuR[7], Mv[7], Mw[7], Mxi[3];
};

template <int NP> KBG_HD void gauss_moments(const double* prim, double K, GMom<NP>& g) {
    const double lam = prim[NP - 1], U = prim[1];
    const double ex = 0.5 * exp(-lam * (U * U)) / sqrt(kPi * lam);
    g.MuL[0] = 0.5 * erfc(-sqrt(lam) * U);
    g.MuL[1] = U * g.MuL[0] + ex;
    g.MuR[0] = 0.5 * erfc(sqrt(lam) * U);
    g.MuR[1] = U * g.MuR[0] - ex;
    for (int i = 2; i <= 6; i++) {
        g.MuL[i] = U * g.MuL[i - 1] + 0.5 * (i - 1) * g.MuL[i - 2] / lam;
        g.MuR[i] = U * g.MuR[i - 1] + 0.5 * (i - 1) * g.MuR[i - 2] / lam;
    }
    for (int i = 0; i <= 6; i++) g.Mu[i] = g.MuL[i] + g.MuR[i];
    if constexpr (NP >= 4) {
        const double V = prim[2];
        g.Mv[0] = 1.0;
        g.Mv[1] = V;
        for (int i = 2; i <= 6; i++) g.Mv[i] = V * g.Mv[i - 1] + 0.5 * (i - 1) * g.Mv[i - 2] / lam;
    }
    if constexpr (NP == 5) {
        const double W = prim[3];
        g.Mw[0] = 1.0;
        g.Mw[1] = W;
        for (int i = 2; i <= 6; i++) g.Mw[i] = W * g.Mw[i - 1] + 0.5 * (i - 1) * g.Mw[i - 2] / lam;
    }
    g.Mxi[0] = 1.0;
    g.Mxi[1] = 0.5 * K / lam;
    g.Mxi[2] = (K * K + 2.0 * K) / (4.0 * (lam * lam));
}

// moments_conserve 1D (theory_moments.jl:98-110)
KBG_HD void mc1(const double* Mu, const double* Mxi, int a, int d, double* o) {
    o[0] = Mu[a] * Mxi[d / 2];
    o[1] = Mu[a + 1] * Mxi[d / 2];
    o[2] = 0.5 * (Mu[a + 2] * Mxi[d / 2] + Mu[a] * Mxi[(d + 2) / 2]);
}
// moments_conserve with internal motion (theory_moments.jl:123-131)
KBG_HD void mc2(const double* Mu, const double* Mv, const double* Mxi, int a, int b, int d, double* o) {
    o[0] = Mu[a] * Mv[b] * Mxi[d / 2];
    o[1] = Mu[a + 1] * Mv[b] * Mxi[d / 2];
    o[2] = Mu[a] * Mv[b + 1] * Mxi[d / 2];
    o[3] = 0.5 * (Mu[a + 2] * Mv[b] * Mxi[d / 2] + Mu[a] * Mv[b + 2] * Mxi[d / 2] + Mu[a] * Mv[b] * Mxi[(d + 2) / 2]);
}
// moments_conserve 3V (theory_moments.jl:132-142)
KBG_HD void mc3(const double* Mu, const double* Mv, const double* Mw, int a, int b, int d, double* o) {
    o[0] = Mu[a] * Mv[b] * Mw[d];
    o[1] = Mu[a + 1] * Mv[b] * Mw[d];
    o[2] = Mu[a] * Mv[b + 1] * Mw[d];
    o[3] = Mu[a] * Mv[b] * Mw[d + 1];
    o[4] = 0.5 * (Mu[a + 2] * Mv[b] * Mw[d] + Mu[a] * Mv[b + 2] * Mw[d] + Mu[a] * Mv[b] * Mw[d + 2]);
}
// moments_conserve_slope 1D (theory_moments.jl:335-339)
KBG_HD void mcs1(const double* a, const double* Mu, const double* Mxi, int al, double* o) {
    double t0[3], t1[3], t2[3], t3[3];
    mc1(Mu, Mxi, al, 0, t0);
    mc1(Mu, Mxi, al + 1, 0, t1);
    mc1(Mu, Mxi, al + 2, 0, t2);
    mc1(Mu, Mxi, al, 2, t3);
    for (int m = 0; m < 3; m++) o[m] = a[0] * t0[m] + a[1] * t1[m] + 0.5 * a[2] * t2[m] + 0.5 * a[2] * t3[m];
}
// moments_conserve_slope 2D with internal motion (theory_moments.jl:344-359)
KBG_HD void mcs2(const double* a, const double* Mu, const double* Mv, const double* Mxi, int al, int be, double* o) {
    double t0[4], t1[4], t2[4], t3[4], t4[4], t5[4];
    mc2(Mu, Mv, Mxi, al, be, 0, t0);
    mc2(Mu, Mv, Mxi, al + 1, be, 0, t1);
    mc2(Mu, Mv, Mxi, al, be + 1, 0, t2);
    mc2(Mu, Mv, Mxi, al + 2, be, 0, t3);
    mc2(Mu, Mv, Mxi, al, be + 2, 0, t4);
    mc2(Mu, Mv, Mxi, al, be, 2, t5);
    for (int m = 0; m < 4; m++)
        o[m] = a[0] * t0[m] + a[1] * t1[m] + a[2] * t2[m] + 0.5 * a[3] * t3[m] + 0.5 * a[3] * t4[m] + 0.5 * a[3] * t5[m];
}

// pdf_slope   theory_pdf.jl:11-49
template <int N> KBG_HD void pdf_slope(const double* p, const double* sw, double K, double* sl) {
    if constexpr (N == 3) {
        sl[2] = 4.0 * (p[2] * p[2]) / (K + 1.0) / p[0] *
                (2.0 * sw[2] - 2.0 * p[1] * sw[1] + sw[0] * ((p[1] * p[1]) - 0.5 * (K + 1.0) / p[2]));
        sl[1] = 2.0 * p[2] / p[0] * (sw[1] - p[1] * sw[0]) - p[1] * sl[2];
        sl[0] = sw[0] / p[0] - p[1] * sl[1] - 0.5 * ((p[1] * p[1]) + 0.5 * (K + 1.0) / p[2]) * sl[2];
    } else {
        sl[3] = 4.0 * (p[3] * p[3]) / (K + 2.0) / p[0] *
                (2.0 * sw[3] - 2.0 * p[1] * sw[1] - 2.0 * p[2] * sw[2] +
                 sw[0] * ((p[1] * p[1]) + (p[2] * p[2]) - 0.5 * (K + 2.0) / p[3]));
        sl[2] = 2.0 * p[3] / p[0] * (sw[2] - p[2] * sw[0]) - p[2] * sl[3];
        sl[1] = 2.0 * p[3] / p[0] * (sw[1] - p[1] * sw[0]) - p[1] * sl[3];
        sl[0] = sw[0] / p[0] - p[1] * sl[1] - p[2] * sl[2] -
                0.5 * ((p[1] * p[1]) + (p[2] * p[2]) + 0.5 * (K + 2.0) / p[3]) * sl[3];
    }
}

// Interface state of the KCU / UGKS fluxes from the reconstructed left/right conserved states:
// w = rho_L <u>_{u>0} + rho_R <u>_{u<0} (flux_kcu.jl:262-274, flux_ugks.jl:46-54), prim = prim(w).
template <int NW> KBG_HD void interface_state(const double* wL, const double* wR, double K, double gam,
                                              double* pL, double* pR, double* w, double* prim) {
    conserve_prim<NW>(wL, gam, pL);
    conserve_prim<NW>(wR, gam, pR);
    GMom<NW> g1, g2;
    gauss_moments<NW>(pL, K, g1);
    gauss_moments<NW>(pR, K, g2);
    double m1[5], m2[5];
    if constexpr (NW == 3) {
        mc1(g1.MuL, g1.Mxi, 0, 0, m1);
        mc1(g2.MuR, g2.Mxi, 0, 0, m2);
    } else if constexpr (NW == 4) {
        mc2(g1.MuL, g1.Mv, g1.Mxi, 0, 0, 0, m1);
        mc2(g2.MuR, g2.Mv, g2.Mxi, 0, 0, 0, m2);
    } else {
        mc3(g1.MuL, g1.Mv, g1.Mw, 0, 0, 0, m1);
        mc3(g2.MuR, g2.Mv, g2.Mw, 0, 0, 0, m2);
    }
    for (int m = 0; m < NW; m++) w[m] = pL[0] * m1[m] + pR[0] * m2[m];
    conserve_prim<NW>(w, gam, prim);
}

// KCU relaxation-time correction   flux_kcu.jl:276-280
KBG_HD double kcu_tau_extra(double rL, double lL, double rR, double lR, double dt) {
    const double a = rL / lL, b = rR / lR;
    return fabs(a - b) / (a + b) * dt * 2.0;
}

// UGKS integral time coefficients Mt[0..4] (Julia Mt[1..5])   flux_ugks.jl:66-71
KBG_HD void ugks_mt(double tau, double dt, double* Mt) {
    const double e = exp(-dt / tau);
    Mt[3] = tau * (1.0 - e);
    Mt[4] = -tau * dt * e + tau * Mt[3];
    Mt[0] = dt - Mt[3];
    Mt[1] = -tau * Mt[0] + Mt[4];
    Mt[2] = (dt * dt) / 2.0 - tau * Mt[0];
}

// Face coefficients of the 1D UGKS flux (flux_ugks.jl:46-79): interface prim, slopes aL/aR/aT,
// Mt, and the Gaussian part of fw.  Stored FP64.
struct UgksFace1 {
    double prim[3], aL[3], aR[3], aT[3], Mt[5], fw0[3], Mxi2;
};
KBG_HD void ugks_face_1d(const double* wL, const double* wR, double K, double gam, double mu, double om,
                         double dt, double dxL, double dxR, UgksFace1& F) {
    double pL[3], pR[3], w[3], t[3], MauL[3], MauR[3], MauT[3], Muv[3];
    interface_state<3>(wL, wR, K, gam, pL, pR, w, F.prim);
    const double* prim = F.prim;
    for (int m = 0; m < 3; m++) t[m] = (w[m] - wL[m]) / dxL;
    pdf_slope<3>(prim, t, K, F.aL);
    for (int m = 0; m < 3; m++) t[m] = (wR[m] - w[m]) / dxR;
    pdf_slope<3>(prim, t, K, F.aR);
    GMom<3> g;
    gauss_moments<3>(prim, K, g);
    mcs1(F.aL, g.MuL, g.Mxi, 1, MauL);
    mcs1(F.aR, g.MuR, g.Mxi, 1, MauR);
    for (int m = 0; m < 3; m++) t[m] = -prim[0] * (MauL[m] + MauR[m]);
    pdf_slope<3>(prim, t, K, F.aT);
    ugks_mt(vhs_tau(prim[0], prim[2], mu, om), dt, F.Mt);
    mc1(g.Mu, g.Mxi, 1, 0, Muv);
    mcs1(F.aL, g.MuL, g.Mxi, 2, MauL);
    mcs1(F.aR, g.MuR, g.Mxi, 2, MauR);
    mcs1(F.aT, g.Mu, g.Mxi, 1, MauT);
    for (int m = 0; m < 3; m++)
        F.fw0[m] = F.Mt[0] * prim[0] * Muv[m] + F.Mt[1] * prim[0] * (MauL[m] + MauR[m]) + F.Mt[2] * prim[0] * MauT[m];
    F.Mxi2 = g.Mxi[2];
}

// 2D version (flux_ugks.jl:161-194), prim/slopes of length 4, fw0 not yet multiplied by len
struct UgksFace2 {
    double prim[4], aL[4], aR[4], aT[4], Mt[5], fw0[4], Mxi2;
};
KBG_HD void ugks_face_2d(const double* wL, const double* wR, double K, double gam, double mu, double om,
                         double dt, double dxL, double dxR, UgksFace2& F) {
    double pL[4], pR[4], w[4], t[4], MauL[4], MauR[4], MauT[4], Muv[4];
    interface_state<4>(wL, wR, K, gam, pL, pR, w, F.prim);
    const double* prim = F.prim;
    for (int m = 0; m < 4; m++) t[m] = (w[m] - wL[m]) / dxL;
    pdf_slope<4>(prim, t, K, F.aL);
    for (int m = 0; m < 4; m++) t[m] = (wR[m] - w[m]) / dxR;
    pdf_slope<4>(prim, t, K, F.aR);
    GMom<4> g;
    gauss_moments<4>(prim, K, g);
    mcs2(F.aL, g.MuL, g.Mv, g.Mxi, 1, 0, MauL);
    mcs2(F.aR, g.MuR, g.Mv, g.Mxi, 1, 0, MauR);
    for (int m = 0; m < 4; m++) t[m] = -prim[0] * (MauL[m] + MauR[m]);
    pdf_slope<4>(prim, t, K, F.aT);
    ugks_mt(vhs_tau(prim[0], prim[3], mu, om), dt, F.Mt);
    mc2(g.Mu, g.Mv, g.Mxi, 1, 0, 0, Muv);
    mcs2(F.aL, g.MuL, g.Mv, g.Mxi, 2, 0, MauL);
    mcs2(F.aR, g.MuR, g.Mv, g.Mxi, 2, 0, MauR);
    mcs2(F.aT, g.Mu, g.Mv, g.Mxi, 1, 0, MauT);
    for (int m = 0; m < 4; m++)
        F.fw0[m] = F.Mt[0] * prim[0] * Muv[m] + F.Mt[1] * prim[0] * (MauL[m] + MauR[m]) + F.Mt[2] * prim[0] * MauT[m];
    F.Mxi2 = g.Mxi[2];
}

// Face coefficients of the KCU flux: interface prim, Mt1/Mt2 and the Gaussian part Mt1 rho <u psi>
template <int NW> struct KcuFace {
    double prim[NW], Mt1, Mt2, fw0[NW];
};
template <int NW> KBG_HD void kcu_face(const double* wL, const double* wR, double K, double gam, double mu, double om,
                                       double dt, KcuFace<NW>& F) {
    double pL[NW], pR[NW], w[NW], Muv[NW];
    interface_state<NW>(wL, wR, K, gam, pL, pR, w, F.prim);
    const double tau = vhs_tau(F.prim[0], F.prim[NW - 1], mu, om) + kcu_tau_extra(pL[0], pL[NW - 1], pR[0], pR[NW - 1], dt);
    F.Mt2 = tau * (1.0 - exp(-dt / tau));
    F.Mt1 = dt - F.Mt2;
    GMom<NW> g;
    gauss_moments<NW>(F.prim, K, g);
    if constexpr (NW == 3) mc1(g.Mu, g.Mxi, 1, 0, Muv);
    else if constexpr (NW == 4) mc2(g.Mu, g.Mv, g.Mxi, 1, 0, 0, Muv);
    else mc3(g.Mu, g.Mv, g.Mw, 1, 0, 0, Muv);
    for (int m = 0; m < NW; m++) F.fw0[m] = F.Mt1 * F.prim[0] * Muv[m];
}

// local/global frame rotations   geo_general.jl:6-26, :130-150 (4-vectors)
KBG_HD void local_frame4(const double* w, double c, double s, double* o) {
    o[0] = w[0];
    o[1] = w[1] * c + w[2] * s;
    o[2] = w[2] * c - w[1] * s;
    o[3] = w[3];
}
KBG_HD void global_frame4(const double* w, double c, double s, double* o) {
    o[0] = w[0];
    o[1] = w[1] * c - w[2] * s;
    o[2] = w[1] * s + w[2] * c;
    o[3] = w[3];
}

// exp in the storage precision
KBG_D float texp(float x) { return __expf(x); }
KBG_D double texp(double x) { return exp(x); }
KBG_D float tsqrt(float x) { return sqrtf(x); }
KBG_D double tsqrt(double x) { return sqrt(x); }

// Raw buffer access: lane byte offset + uniform (scalar) offset.  Offsets past
// the buffer read 0 and drop stores.  Word 3 of the resource is the GFX9 raw 32-bit data format.
using rsrc_t = __amdgpu_buffer_rsrc_t;
KBG_D rsrc_t make_rsrc(const void* p, unsigned bytes) { return __builtin_amdgcn_make_buffer_rsrc((void*)p, (short)0, (int)bytes, 0x00020000); }
KBG_D float bload(float*, rsrc_t r, int voff, int soff) { return __builtin_bit_cast(float, __builtin_amdgcn_raw_buffer_load_b32(r, voff, soff, 0)); }
KBG_D double bload(double*, rsrc_t r, int voff, int soff) { return __builtin_bit_cast(double, __builtin_amdgcn_raw_buffer_load_b64(r, voff, soff, 0)); }
// AUX: the instruction's cache-policy bits (0: default)
template <int AUX = 0> KBG_D void bstore(float x, rsrc_t r, int voff, int soff) {
    __builtin_amdgcn_raw_buffer_store_b32(__builtin_bit_cast(unsigned, x), r, voff, soff, AUX);
}
template <int AUX = 0> KBG_D void bstore(double x, rsrc_t r, int voff, int soff) {
    __builtin_amdgcn_raw_buffer_store_b64(__builtin_bit_cast(__attribute__((ext_vector_type(2))) unsigned, x), r, voff, soff, AUX);
}

// ---- reductions -------------------------------------------------------------------------------

// all-reduce of N doubles across the 64 lanes of a wavefront (butterfly; every lane gets the sums)
template <int N> KBG_D void wave_allreduce(double* v) {
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) {
#pragma unroll
        for (int m = 0; m < N; m++) v[m] += __shfl_xor(v[m], off, 64);
    }
}

// Reduce-scatter of V values (V a power of two, 2..64) across the 64 lanes of a wavefront (a
// pairwise tree in the value type).
// Afterwards v[0] of lane l holds the wavefront total of value index l >> (6 - log2 V), i.e. value i
// lives in lanes [i << (6 - log2 V), (i + 1) << (6 - log2 V)).  Cost: V - 1 + 6 - log2 V shuffles,
// against 6 V for V independent all-reductions.
template <int V> constexpr int ilog2c() { return V <= 1 ? 0 : 1 + ilog2c<V / 2>(); }
template <int V, class T> KBG_D T wave_reduce_scatter(T* v) {
    constexpr int LV = ilog2c<V>();
    static_assert((1 << LV) == V && V <= 64, "V must be a power of two <= 64");
    const int lane = threadIdx.x & 63;
#pragma unroll
    for (int s = 0; s < LV; s++) {
        const int m = 32 >> s;
        const int h = (V >> s) / 2;
        const bool up = (lane & m) != 0;
#pragma unroll
        for (int i = 0; i < h; i++) {
            // the empty register constraints keep the two values in registers: without them the
            // compiler turns the lane-dependent select of two array elements into a load through a
            // selected address, which moves the whole array to scratch memory
            T a = v[i], b = v[i + h];
            asm volatile("" : "+v"(a), "+v"(b));
            const T send = up ? a : b;
            const T keep = up ? b : a;
            v[i] = keep + __shfl_xor(send, m, 64);
        }
    }
#pragma unroll
    for (int m = (32 >> LV); m >= 1; m >>= 1) v[0] += __shfl_xor(v[0], m, 64);
    return v[0];
}
template <int V> KBG_D int scatter_lane(int i) { return i << (6 - ilog2c<V>()); }

// broadcast of a wave-uniform lane's value (lane index must be uniform)
KBG_D float bcast(float x, int l) { return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(x), l)); }
KBG_D double bcast(double x, int l) {
    const long long b = __double_as_longlong(x);
    const int lo = __builtin_amdgcn_readlane((int)(b & 0xffffffffLL), l);
    const int hi = __builtin_amdgcn_readlane((int)(b >> 32), l);
    return __longlong_as_double(((long long)hi << 32) | (long long)(unsigned)lo);
}

// all-reduce of N doubles across a workgroup of up to 16 wavefronts, wavefront totals added in
// wavefront order; scratch must hold 16*N doubles and is reused by the caller only after the
// trailing barrier.
template <int N> KBG_D void block_allreduce(double* v, double* scratch) {
    wave_allreduce<N>(v);
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6, nwv = blockDim.x >> 6;
    if (lane == 0) {
#pragma unroll
        for (int m = 0; m < N; m++) scratch[wv * N + m] = v[m];
    }
    __syncthreads();
#pragma unroll
    for (int m = 0; m < N; m++) {
        double t = scratch[m];
        for (int w = 1; w < nwv; w++) t += scratch[w * N + m];
        v[m] = t;
    }
    __syncthreads();
}

// all-reduce of N doubles across a 256-thread block; scratch must hold 4*N doubles and is reused
// by the caller only after the trailing barrier.
template <int N> KBG_D void block_allreduce256(double* v, double* scratch) {
    wave_allreduce<N>(v);
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    if (lane == 0) {
#pragma unroll
        for (int m = 0; m < N; m++) scratch[wv * N + m] = v[m];
    }
    __syncthreads();
#pragma unroll
    for (int m = 0; m < N; m++) v[m] = ((scratch[m] + scratch[N + m]) + scratch[2 * N + m]) + scratch[3 * N + m];
    __syncthreads();
}

// Kahan-compensated accumulator in the storage precision (used where one thread sums many terms)
template <class T> struct KSum {
    T s = T(0), c = T(0);
    KBG_D void add(T x) {
        T y = x - c;
        T t = s + y;
        c = (t - s) - y;
        s = t;
    }
    KBG_D double val() const { return double(s) - double(c); }
};

}  // namespace kbg
