/*
 * kitbase_oracle.c -- FP64 CPU restatement of KitBase.jl (v0.9.31) for the discrete-velocity time
 * step: reconstruct! -> evolve! (KFVS / KCU / UGKS fluxes, Maxwell walls) -> update! (BGK /
 * Shakhov step!).  It also restates the setup helpers the step depends on (velocity quadratures,
 * the structured 2D mesh geometry).
 *
 * TEST INFRASTRUCTURE ONLY: the HIP path is tested against it.  Only tests/,
 * __graft_entry__.smoke() and bench.py's cpu_baseline leg may load it; the product library never
 * does.  OpenMP threads the cell and face loops the way KitBase threads them with @threads, so the
 * same code doubles as the CPU baseline.
 *
 * Parity unpinned (no golden vector, closed form or brute-force check fixes them, only their
 * limits and consistency): the UGKS time coefficients Mt at intermediate tau (pinned at tau -> 0 and
 * tau -> infinity and against the discrete moments), and multi-step trajectories with all of
 * KitBase's quirks combined.
 *
 * Citations are relative to the KitBase.jl source tree.  Julia evaluates broadcast expressions
 * left to right; the C below keeps the same operation order so the FP64 results agree with KitBase
 * to rounding (summation order of `sum` is Julia's pairwise/SIMD order in KitBase and sequential
 * here, a difference of a few ulp).
 */
#define _DEFAULT_SOURCE /* j1 (Bessel J1) from libm, for the fast-spectral kernel modes */
#include "kitbase_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define PI_ 3.14159265358979323846

/* ------------------------------------------------------------------------------------------ */
/* math helpers                                                                                 */
/* ------------------------------------------------------------------------------------------ */

/* heaviside(x) = ifelse(x >= 0, 1, 0)   src/Math/math_general.jl:13 */
double orc_heaviside(double x) { return x >= 0.0 ? 1.0 : 0.0; }

/* Julia's sign: -1, +1, or x itself for +-0 and NaN */
static double jsign(double y) {
    if (y > 0.0) return 1.0;
    if (y < 0.0) return -1.0;
    return y;
}

/* fortsign(x, y) = abs(x) * sign(y)   src/Math/math_general.jl:20 (Julia sign(0) == 0) */
double orc_fortsign(double x, double y) { return fabs(x) * jsign(y); }

/* limiters   src/Reconstruction/recon_limiter.jl */
double orc_limiter(int kind, double sL, double sR) {
    switch (kind) {
    case ORC_VANLEER: /* :17-19 */
        return (orc_fortsign(1.0, sL) + orc_fortsign(1.0, sR)) * fabs(sL) * fabs(sR) /
               (fabs(sL) + fabs(sR) + 1.e-7);
    case ORC_MINMOD: /* :57 */
        return 0.5 * (orc_fortsign(1.0, sL) + orc_fortsign(1.0, sR)) * fmin(fabs(sR), fabs(sL));
    case ORC_SUPERBEE: /* :90-98 */
        if (sR >= 0.5 * sL && sR <= 2.0 * sL)
            return 0.5 * (orc_fortsign(1.0, sL) + orc_fortsign(1.0, sR)) * fmax(fabs(sL), fabs(sR));
        else if (sR < 0.5 * sL && sR > 2.0 * sL)
            return (orc_fortsign(1.0, sL) + orc_fortsign(1.0, sR)) * fmin(fabs(sL), fabs(sR));
        return 0.0;
    case ORC_VANALBABA: /* :105 */
        return (sL * sL * sR + sL * sR * sR) / (sL * sL + sR * sR + 1.e-7);
    case ORC_LINEAR: /* :10 */
        return 0.5 * (sL + sR);
    }
    return 0.0;
}

/* ------------------------------------------------------------------------------------------ */
/* setup: quadratures and mesh                                                                  */
/* ------------------------------------------------------------------------------------------ */

/* newton_cotes   src/Phase/phase_quadrature.jl:276-288 (Julia `%` is a remainder with the sign of
 * the dividend; idx >= 1 here so C `%` agrees) */
double orc_newton_cotes(int idx, int num) {
    if (idx == 1 || idx == num) return 14.0 / 45.0;
    if ((idx - 5) % 4 == 0) return 28.0 / 45.0;
    if ((idx - 3) % 4 == 0) return 24.0 / 45.0;
    return 64.0 / 45.0;
}

/* VSpace1D   src/Phase/phase_velocity.jl:23-65.  Arrays hold nu + 2 ng nodes, the node with Julia
 * index i (1-ng .. nu+ng) at position i - 1 + ng. */
void orc_vspace1d(int type, double U0, double U1, int nu, int ng, double *u, double *du,
                  double *weights) {
    double d = (U1 - U0) / nu;
    if (type == ORC_ALGEBRA) { /* :54-59 (ng must be 0) */
        int n1 = nu + 1;
        double *uu = (double *)malloc(sizeof(double) * n1);
        for (int i = 1; i <= n1; i++) {
            double a = (double)(-n1 + 1 + 2 * (i - 1));
            double p = (double)(n1 - 1);
            uu[i - 1] = U1 / (p * p * p) * (a * a * a);
        }
        for (int i = 0; i < nu; i++) {
            u[i] = (uu[i] + uu[i + 1]) / 2.0;
            du[i] = uu[i + 1] - uu[i];
            weights[i] = du[i];
        }
        free(uu);
        return;
    }
    for (int i = 1 - ng; i <= nu + ng; i++) {
        int k = i - 1 + ng;
        u[k] = U0 + (i - 0.5) * d;
        du[k] = d;
        weights[k] = (type == ORC_NEWTON) ? orc_newton_cotes(i + ng, nu + ng * 2) * d : d;
    }
}

/* Cyclic Jacobi eigen-decomposition of a small symmetric matrix (n <= 32), column eigenvectors.
 * Used for the Golub-Welsch step of maxwell_quadrature in place of LAPACK's symmetric eigen. */
static void jacobi_eigen(int n, double *A, double *V) {
    for (int i = 0; i < n; i++)
        for (int j = 0; j < n; j++) V[i * n + j] = (i == j) ? 1.0 : 0.0;
    for (int sweep = 0; sweep < 100; sweep++) {
        double off = 0.0;
        for (int p = 0; p < n; p++)
            for (int q = p + 1; q < n; q++) off += A[p * n + q] * A[p * n + q];
        if (off < 1e-60) break;
        for (int p = 0; p < n; p++) {
            for (int q = p + 1; q < n; q++) {
                double apq = A[p * n + q];
                if (fabs(apq) < 1e-300) continue;
                double theta = (A[q * n + q] - A[p * n + p]) / (2.0 * apq);
                double t = (theta >= 0 ? 1.0 : -1.0) / (fabs(theta) + sqrt(theta * theta + 1.0));
                double c = 1.0 / sqrt(t * t + 1.0), s = t * c;
                /* A <- J^T A J with J = I + (c-1)(e_p e_p^T + e_q e_q^T) + s(e_p e_q^T - e_q e_p^T) */
                for (int k = 0; k < n; k++) {
                    double akp = A[k * n + p], akq = A[k * n + q];
                    A[k * n + p] = c * akp - s * akq;
                    A[k * n + q] = s * akp + c * akq;
                }
                for (int k = 0; k < n; k++) {
                    double apk = A[p * n + k], aqk = A[q * n + k];
                    A[p * n + k] = c * apk - s * aqk;
                    A[q * n + k] = s * apk + c * aqk;
                }
                for (int k = 0; k < n; k++) {
                    double vkp = V[k * n + p], vkq = V[k * n + q];
                    V[k * n + p] = c * vkp - s * vkq;
                    V[k * n + q] = s * vkp + c * vkq;
                }
            }
        }
    }
}

/* maxwell_quadrature(N2, C)   src/Phase/phase_quadrature.jl:13-44: half-range Gauss rule for the
 * weight exp(-x^2) on [0, inf) by Golub-Welsch, mirrored to the full line, weights multiplied by
 * exp(x^2) * C and nodes by C.  Returns 0, or -1 if n2 is not an even number in [4, 64]. */
int orc_maxwell_quadrature(int n2, double C, double *x, double *w) {
    int N = n2 / 2;
    if (n2 % 2 != 0 || N < 2 || N > 32) return -1;
    double a[32], b[32], A[32 * 32], V[32 * 32], ev[32], ew[32];
    memset(a, 0, sizeof a);
    memset(b, 0, sizeof b);
    a[0] = 1.0 / sqrt(PI_);
    a[1] = 2.0 / sqrt(PI_) / (PI_ - 2.0);
    b[1] = a[0] / (a[0] + a[1]) / 2.0;
    for (int i = 3; i <= N; i++) { /* Julia indices i, arrays 1-based */
        b[i - 1] = (i - 2) + 1.0 / 2.0 - b[i - 2] - a[i - 2] * a[i - 2];
        a[i - 1] = ((double)(i - 1) * (i - 1) / (4.0 * b[i - 1]) - b[i - 2] - 1.0 / 2) / a[i - 2] - a[i - 2];
    }
    memset(A, 0, sizeof(double) * N * N);
    for (int i = 0; i < N; i++) A[i * N + i] = a[i];
    for (int i = 1; i < N; i++) {
        A[i * N + i - 1] = sqrt(b[i]);
        A[(i - 1) * N + i] = sqrt(b[i]);
    }
    jacobi_eigen(N, A, V);
    for (int i = 0; i < N; i++) {
        ev[i] = A[i * N + i];
        ew[i] = V[0 * N + i] * V[0 * N + i] * sqrt(PI_) / 2.0;
    }
    /* sort ascending by node (sortslices by=x->x[1]) */
    for (int i = 1; i < N; i++) {
        double kv = ev[i], kw = ew[i];
        int j = i - 1;
        while (j >= 0 && ev[j] > kv) {
            ev[j + 1] = ev[j];
            ew[j + 1] = ew[j];
            j--;
        }
        ev[j + 1] = kv;
        ew[j + 1] = kw;
    }
    for (int i = 0; i < N; i++) {
        x[i] = -ev[N - 1 - i];
        w[i] = ew[N - 1 - i];
        x[N + i] = ev[i];
        w[N + i] = ew[i];
    }
    for (int i = 0; i < 2 * N; i++) {
        w[i] *= exp(x[i] * x[i]) * C;
        x[i] *= C;
    }
    return 0;
}

static double norm2v(double a, double b) { return sqrt(a * a + b * b); }

/* unit_normal(p1, p2)   src/Geometry/geo_general.jl:230-235 (note the +1e-6 in the length) */
static void unit_normal(const double *p1, const double *p2, double *n) {
    double d0 = p2[0] - p1[0], d1 = p2[1] - p1[1];
    double l = norm2v(d0, d1) + 1e-6;
    n[0] = -d1 / l;
    n[1] = d0 / l;
}

/* PSpace2D   src/Geometry/geo_struct.jl:64-175 (no ghost layers).  Cell c = (i-1) + nx (j-1);
 * areas[4 c + k] and normals[8 c + 2 k + {0,1}] for the edges k = 0..3 (KitBase edges 1..4:
 * y-, x+, y+, x-). */
void orc_pspace2d(double X0, double X1, int nx, double Y0, double Y1, int ny, double *x, double *y,
                  double *dx, double *dy, double *areas, double *normals) {
    double ddx = (X1 - X0) / nx, ddy = (Y1 - Y0) / ny;
    for (int j = 1; j <= ny; j++)
        for (int i = 1; i <= nx; i++) {
            int c = (i - 1) + nx * (j - 1);
            x[c] = X0 + (i - 0.5) * ddx;
            y[c] = Y0 + (j - 0.5) * ddy;
            dx[c] = ddx;
            dy[c] = ddy;
            double vt[4][2];
            vt[0][0] = x[c] - 0.5 * dx[c];
            vt[1][0] = x[c] + 0.5 * dx[c];
            vt[2][0] = x[c] + 0.5 * dx[c];
            vt[3][0] = x[c] - 0.5 * dx[c];
            vt[0][1] = y[c] - 0.5 * dy[c];
            vt[1][1] = y[c] - 0.5 * dy[c];
            vt[2][1] = y[c] + 0.5 * dy[c];
            vt[3][1] = y[c] + 0.5 * dy[c];
            for (int k = 0; k < 4; k++) {
                const double *p1 = vt[k], *p2 = vt[(k + 1) % 4];
                areas[4 * c + k] = norm2v(p1[0] - p2[0], p1[1] - p2[1]);
                double n[2];
                unit_normal(p1, p2, n);
                double m0 = x[c] - (p1[0] + p2[0]) / 2, m1 = y[c] - (p1[1] + p2[1]) / 2;
                double dot = n[0] * m0 + n[1] * m1;
                if (!(dot <= 0)) {
                    n[0] = -n[0];
                    n[1] = -n[1];
                }
                normals[8 * c + 2 * k + 0] = n[0];
                normals[8 * c + 2 * k + 1] = n[1];
            }
        }
}

/* ------------------------------------------------------------------------------------------ */
/* theory                                                                                       */
/* ------------------------------------------------------------------------------------------ */

/* prim_conserve!   src/Theory/theory_macro.jl:10-33 */
void orc_prim_conserve(const double *prim, int n, double gam, double *W) {
    if (n == 3) {
        W[0] = prim[0];
        W[1] = prim[0] * prim[1];
        W[2] = 0.5 * prim[0] / prim[2] / (gam - 1.0) + 0.5 * prim[0] * (prim[1] * prim[1]);
    } else if (n == 4) {
        W[0] = prim[0];
        W[1] = prim[0] * prim[1];
        W[2] = prim[0] * prim[2];
        W[3] = 0.5 * prim[0] / prim[3] / (gam - 1.0) +
               0.5 * prim[0] * (prim[1] * prim[1] + prim[2] * prim[2]);
    } else if (n == 5) {
        W[0] = prim[0];
        W[1] = prim[0] * prim[1];
        W[2] = prim[0] * prim[2];
        W[3] = prim[0] * prim[3];
        W[4] = 0.5 * prim[0] / prim[4] / (gam - 1.0) +
               0.5 * prim[0] * (prim[1] * prim[1] + prim[2] * prim[2] + prim[3] * prim[3]);
    }
}

/* conserve_prim!   src/Theory/theory_macro.jl:185-206 */
void orc_conserve_prim(const double *W, int n, double gam, double *prim) {
    double p[5];
    if (n == 3) {
        p[0] = W[0];
        p[1] = W[1] / W[0];
        p[2] = 0.5 * W[0] / (gam - 1.0) / (W[2] - 0.5 * (W[1] * W[1]) / W[0]);
    } else if (n == 4) {
        p[0] = W[0];
        p[1] = W[1] / W[0];
        p[2] = W[2] / W[0];
        p[3] = 0.5 * W[0] / (gam - 1.0) / (W[3] - 0.5 * (W[1] * W[1] + W[2] * W[2]) / W[0]);
    } else {
        p[0] = W[0];
        p[1] = W[1] / W[0];
        p[2] = W[2] / W[0];
        p[3] = W[3] / W[0];
        p[4] = 0.5 * W[0] / (gam - 1.0) /
               (W[4] - 0.5 * (W[1] * W[1] + W[2] * W[2] + W[3] * W[3]) / W[0]);
    }
    memcpy(prim, p, sizeof(double) * n);
}

/* heat_capacity_ratio   src/Theory/theory_macro.jl:376-384 */
double orc_heat_capacity_ratio(double K, int D) {
    if (D == 1) return (K + 3.0) / (K + 1.0);
    if (D == 2) return (K + 4.0) / (K + 2.0);
    return (K + 5.0) / (K + 3.0);
}

/* sound_speed   src/Theory/theory_macro.jl:421 */
double orc_sound_speed(double lam, double gam) { return pow(0.5 * gam / lam, 0.5); }

/* ref_vhs_vis   src/Theory/theory_macro.jl:447-449 */
double orc_ref_vhs_vis(double Kn, double alpha, double omega) {
    return 5.0 * (alpha + 1.0) * (alpha + 2.0) * sqrt(PI_) /
           (4.0 * alpha * (5.0 - 2.0 * omega) * (7.0 - 2.0 * omega)) * Kn;
}

/* vhs_collision_time(rho, lambda, mu_r, omega)   src/Theory/theory_relaxation.jl:28 */
double orc_vhs_collision_time(double rho, double lam, double mu, double om) {
    return mu * 2.0 * pow(lam, 1.0 - om) / rho;
}

/* maxwellian 1V/2V/3V   src/Theory/theory_equilibrium.jl:8, :20, :32-33 */
void orc_maxwellian_1v(const double *u, int n, double rho, double U, double lam, double *M) {
    for (int i = 0; i < n; i++) {
        double c = u[i] - U;
        M[i] = rho * sqrt(lam / PI_) * exp(-lam * (c * c));
    }
}
void orc_maxwellian_2v(const double *u, const double *v, int n, double rho, double U, double V,
                       double lam, double *M) {
    for (int i = 0; i < n; i++) {
        double cu = u[i] - U, cv = v[i] - V;
        M[i] = rho * (lam / PI_) * exp(-lam * (cu * cu + cv * cv));
    }
}
void orc_maxwellian_3v(const double *u, const double *v, const double *w, int n, double rho,
                       double U, double V, double W, double lam, double *M) {
    double l3 = lam / PI_;
    for (int i = 0; i < n; i++) {
        double cu = u[i] - U, cv = v[i] - V, cw = w[i] - W;
        M[i] = rho * sqrt(l3 * l3 * l3) * exp(-lam * (cu * cu + cv * cv + cw * cw));
    }
}

/* gauss_moments(prim [, inK])   src/Theory/theory_moments.jl:10-86.  prim[end] is lambda. */
void orc_gauss_moments(const double *prim, int nprim, double K, orc_gm *g) {
    double lam = prim[nprim - 1], U = prim[1];
    g->MuL[0] = 0.5 * erfc(-sqrt(lam) * U);
    g->MuL[1] = U * g->MuL[0] + 0.5 * exp(-lam * (U * U)) / sqrt(PI_ * lam);
    g->MuR[0] = 0.5 * erfc(sqrt(lam) * U);
    g->MuR[1] = U * g->MuR[0] - 0.5 * exp(-lam * (U * U)) / sqrt(PI_ * lam);
    for (int i = 2; i <= 6; i++) {
        g->MuL[i] = U * g->MuL[i - 1] + 0.5 * (i - 1) * g->MuL[i - 2] / lam;
        g->MuR[i] = U * g->MuR[i - 1] + 0.5 * (i - 1) * g->MuR[i - 2] / lam;
    }
    for (int i = 0; i <= 6; i++) g->Mu[i] = g->MuL[i] + g->MuR[i];
    for (int i = 0; i <= 6; i++) g->Mv[i] = g->Mw[i] = 0.0;
    if (nprim >= 4) {
        double V = prim[2];
        g->Mv[0] = 1.0;
        g->Mv[1] = V;
        for (int i = 2; i <= 6; i++) g->Mv[i] = V * g->Mv[i - 1] + 0.5 * (i - 1) * g->Mv[i - 2] / lam;
    }
    if (nprim == 5) {
        double W = prim[3];
        g->Mw[0] = 1.0;
        g->Mw[1] = W;
        for (int i = 2; i <= 6; i++) g->Mw[i] = W * g->Mw[i - 1] + 0.5 * (i - 1) * g->Mw[i - 2] / lam;
    }
    g->Mxi[0] = 1.0;
    g->Mxi[1] = 0.5 * K / lam;
    g->Mxi[2] = (K * K + 2.0 * K) / (4.0 * (lam * lam));
}

/* moments_conserve(Mu, Mxi, alpha, delta)   src/Theory/theory_moments.jl:98-110 */
void orc_moments_conserve_1d(const double *Mu, const double *Mxi, int alpha, int delta, double *out) {
    out[0] = Mu[alpha] * Mxi[delta / 2];
    out[1] = Mu[alpha + 1] * Mxi[delta / 2];
    out[2] = 0.5 * (Mu[alpha + 2] * Mxi[delta / 2] + Mu[alpha] * Mxi[(delta + 2) / 2]);
}

/* moments_conserve(Mu, Mv, Mw=Mxi, ...) internal-motion branch   theory_moments.jl:123-131 */
void orc_moments_conserve_2d(const double *Mu, const double *Mv, const double *Mxi, int alpha,
                             int beta, int delta, double *out) {
    out[0] = Mu[alpha] * Mv[beta] * Mxi[delta / 2];
    out[1] = Mu[alpha + 1] * Mv[beta] * Mxi[delta / 2];
    out[2] = Mu[alpha] * Mv[beta + 1] * Mxi[delta / 2];
    out[3] = 0.5 * (Mu[alpha + 2] * Mv[beta] * Mxi[delta / 2] + Mu[alpha] * Mv[beta + 2] * Mxi[delta / 2] +
                    Mu[alpha] * Mv[beta] * Mxi[(delta + 2) / 2]);
}

/* moments_conserve(Mu, Mv, Mw, ...) 3V branch   theory_moments.jl:132-142 */
void orc_moments_conserve_3d(const double *Mu, const double *Mv, const double *Mw, int alpha,
                             int beta, int delta, double *out) {
    out[0] = Mu[alpha] * Mv[beta] * Mw[delta];
    out[1] = Mu[alpha + 1] * Mv[beta] * Mw[delta];
    out[2] = Mu[alpha] * Mv[beta + 1] * Mw[delta];
    out[3] = Mu[alpha] * Mv[beta] * Mw[delta + 1];
    out[4] = 0.5 * (Mu[alpha + 2] * Mv[beta] * Mw[delta] + Mu[alpha] * Mv[beta + 2] * Mw[delta] +
                    Mu[alpha] * Mv[beta] * Mw[delta + 2]);
}

/* moments_conserve_slope(a, Mu, Mxi, alpha)   theory_moments.jl:335-339 */
void orc_moments_conserve_slope_1d(const double *a, const double *Mu, const double *Mxi, int alpha,
                                   double *out) {
    double t0[3], t1[3], t2[3], t3[3];
    orc_moments_conserve_1d(Mu, Mxi, alpha + 0, 0, t0);
    orc_moments_conserve_1d(Mu, Mxi, alpha + 1, 0, t1);
    orc_moments_conserve_1d(Mu, Mxi, alpha + 2, 0, t2);
    orc_moments_conserve_1d(Mu, Mxi, alpha + 0, 2, t3);
    for (int m = 0; m < 3; m++)
        out[m] = a[0] * t0[m] + a[1] * t1[m] + 0.5 * a[2] * t2[m] + 0.5 * a[2] * t3[m];
}

/* moments_conserve_slope(a, Mu, Mv, Mw=Mxi, alpha, beta)   theory_moments.jl:344-359 */
void orc_moments_conserve_slope_2d(const double *a, const double *Mu, const double *Mv,
                                   const double *Mxi, int alpha, int beta, double *out) {
    double t0[4], t1[4], t2[4], t3[4], t4[4], t5[4];
    orc_moments_conserve_2d(Mu, Mv, Mxi, alpha + 0, beta + 0, 0, t0);
    orc_moments_conserve_2d(Mu, Mv, Mxi, alpha + 1, beta + 0, 0, t1);
    orc_moments_conserve_2d(Mu, Mv, Mxi, alpha + 0, beta + 1, 0, t2);
    orc_moments_conserve_2d(Mu, Mv, Mxi, alpha + 2, beta + 0, 0, t3);
    orc_moments_conserve_2d(Mu, Mv, Mxi, alpha + 0, beta + 2, 0, t4);
    orc_moments_conserve_2d(Mu, Mv, Mxi, alpha + 0, beta + 0, 2, t5);
    for (int m = 0; m < 4; m++)
        out[m] = a[0] * t0[m] + a[1] * t1[m] + a[2] * t2[m] + 0.5 * a[3] * t3[m] +
                 0.5 * a[3] * t4[m] + 0.5 * a[3] * t5[m];
}

/* pdf_slope(prim, sw, inK)   src/Theory/theory_pdf.jl:11-49 */
void orc_pdf_slope(const double *prim, const double *sw, int n, double K, double *sl) {
    if (n == 3) {
        sl[2] = 4.0 * (prim[2] * prim[2]) / (K + 1.0) / prim[0] *
                (2.0 * sw[2] - 2.0 * prim[1] * sw[1] + sw[0] * ((prim[1] * prim[1]) - 0.5 * (K + 1.0) / prim[2]));
        sl[1] = 2.0 * prim[2] / prim[0] * (sw[1] - prim[1] * sw[0]) - prim[1] * sl[2];
        sl[0] = sw[0] / prim[0] - prim[1] * sl[1] -
                0.5 * ((prim[1] * prim[1]) + 0.5 * (K + 1.0) / prim[2]) * sl[2];
    } else if (n == 4) {
        sl[3] = 4.0 * (prim[3] * prim[3]) / (K + 2.0) / prim[0] *
                (2.0 * sw[3] - 2.0 * prim[1] * sw[1] - 2.0 * prim[2] * sw[2] +
                 sw[0] * ((prim[1] * prim[1]) + (prim[2] * prim[2]) - 0.5 * (K + 2.0) / prim[3]));
        sl[2] = 2.0 * prim[3] / prim[0] * (sw[2] - prim[2] * sw[0]) - prim[2] * sl[3];
        sl[1] = 2.0 * prim[3] / prim[0] * (sw[1] - prim[1] * sw[0]) - prim[1] * sl[3];
        sl[0] = sw[0] / prim[0] - prim[1] * sl[1] - prim[2] * sl[2] -
                0.5 * ((prim[1] * prim[1]) + (prim[2] * prim[2]) + 0.5 * (K + 2.0) / prim[3]) * sl[3];
    } else {
        sl[4] = 4.0 * (prim[4] * prim[4]) / (K + 3.0) / prim[0] *
                (2.0 * sw[4] - 2.0 * prim[1] * sw[1] - 2.0 * prim[2] * sw[2] - 2.0 * prim[3] * sw[3] +
                 sw[0] * ((prim[1] * prim[1]) + (prim[2] * prim[2]) + (prim[3] * prim[3]) -
                          0.5 * (K + 3.0) / prim[4]));
        sl[3] = 2.0 * prim[4] / prim[0] * (sw[3] - prim[3] * sw[0]) - prim[3] * sl[4];
        sl[2] = 2.0 * prim[4] / prim[0] * (sw[2] - prim[2] * sw[0]) - prim[2] * sl[4];
        sl[1] = 2.0 * prim[4] / prim[0] * (sw[1] - prim[1] * sw[0]) - prim[1] * sl[4];
        sl[0] = sw[0] / prim[0] - prim[1] * sl[1] - prim[2] * sl[2] - prim[3] * sl[3] -
                0.5 * ((prim[1] * prim[1]) + (prim[2] * prim[2]) + (prim[3] * prim[3]) +
                       0.5 * (K + 3.0) / prim[4]) * sl[4];
    }
}

/* global_frame(w, cosa, sina)   src/Geometry/geo_general.jl:6-26 */
void orc_global_frame(const double *w, int n, double c, double s, double *out) {
    double G[4];
    if (n == 2) {
        G[0] = w[0] * c - w[1] * s;
        G[1] = w[0] * s + w[1] * c;
    } else {
        G[0] = w[0];
        G[1] = w[1] * c - w[2] * s;
        G[2] = w[1] * s + w[2] * c;
        G[3] = w[3];
    }
    memcpy(out, G, sizeof(double) * n);
}

/* local_frame(w, cosa, sina)   src/Geometry/geo_general.jl:130-150 */
void orc_local_frame(const double *w, int n, double c, double s, double *out) {
    double L[4];
    if (n == 2) {
        L[0] = w[0] * c + w[1] * s;
        L[1] = w[1] * c - w[0] * s;
    } else {
        L[0] = w[0];
        L[1] = w[1] * c + w[2] * s;
        L[2] = w[2] * c - w[1] * s;
        L[3] = w[3];
    }
    memcpy(out, L, sizeof(double) * n);
}

/* ------------------------------------------------------------------------------------------ */
/* KFVS   src/Flux/flux_kfvs.jl                                                                 */
/* ------------------------------------------------------------------------------------------ */

/* 1F1V   flux_kfvs.jl:238-264 (flux_conserve_1d! theory_moments.jl:389-393) */
void orc_flux_kfvs_1f1v(double *fw, double *ff, const double *fL, const double *fR, const double *u,
                        const double *w, int n, double dt, const double *sfL, const double *sfR) {
    double s0 = 0, s1 = 0, s2 = 0;
    for (int i = 0; i < n; i++) {
        double d = orc_heaviside(u[i]);
        double f = fL[i] * d + fR[i] * (1.0 - d);
        double sf = (sfL ? sfL[i] : 0.0) * d + (sfR ? sfR[i] : 0.0) * (1.0 - d);
        ff[i] = dt * u[i] * f - 0.5 * (dt * dt) * (u[i] * u[i]) * sf;
    }
    for (int i = 0; i < n; i++) {
        s0 += w[i] * ff[i];
        s1 += u[i] * w[i] * ff[i];
        s2 += (u[i] * u[i]) * w[i] * ff[i];
    }
    fw[0] = s0;
    fw[1] = s1;
    fw[2] = 0.5 * s2;
}

/* 2F1V   flux_kfvs.jl:303-340 (flux_conserve_1d! theory_moments.jl:396-400) */
void orc_flux_kfvs_2f1v(double *fw, double *fh, double *fb, const double *hL, const double *bL,
                        const double *hR, const double *bR, const double *u, const double *w, int n,
                        double dt, const double *shL, const double *sbL, const double *shR,
                        const double *sbR) {
    double s0 = 0, s1 = 0, s2 = 0, s3 = 0;
    for (int i = 0; i < n; i++) {
        double d = orc_heaviside(u[i]);
        double h = hL[i] * d + hR[i] * (1.0 - d);
        double b = bL[i] * d + bR[i] * (1.0 - d);
        double sh = (shL ? shL[i] : 0.0) * d + (shR ? shR[i] : 0.0) * (1.0 - d);
        double sb = (sbL ? sbL[i] : 0.0) * d + (sbR ? sbR[i] : 0.0) * (1.0 - d);
        fh[i] = dt * u[i] * h - 0.5 * (dt * dt) * (u[i] * u[i]) * sh;
        fb[i] = dt * u[i] * b - 0.5 * (dt * dt) * (u[i] * u[i]) * sb;
    }
    for (int i = 0; i < n; i++) {
        s0 += w[i] * fh[i];
        s1 += u[i] * w[i] * fh[i];
        s2 += (u[i] * u[i]) * w[i] * fh[i];
        s3 += w[i] * fb[i];
    }
    fw[0] = s0;
    fw[1] = s1;
    fw[2] = 0.5 * (s2 + s3);
}

/* 1F3V   flux_kfvs.jl:443-477 (flux_conserve_3d! theory_moments.jl:419-425) */
void orc_flux_kfvs_1f3v(double *fw, double *ff, const double *fL, const double *fR, const double *u,
                        const double *v, const double *wv, const double *w, int n, double dt,
                        double len, const double *sfL, const double *sfR) {
    double s[5] = {0, 0, 0, 0, 0};
    for (int i = 0; i < n; i++) {
        double d = orc_heaviside(u[i]);
        double f = fL[i] * d + fR[i] * (1.0 - d);
        double sf = (sfL ? sfL[i] : 0.0) * d + (sfR ? sfR[i] : 0.0) * (1.0 - d);
        ff[i] = (dt * u[i] * f - 0.5 * (dt * dt) * (u[i] * u[i]) * sf) * len;
    }
    for (int i = 0; i < n; i++) {
        s[0] += w[i] * ff[i];
        s[1] += u[i] * w[i] * ff[i];
        s[2] += v[i] * w[i] * ff[i];
        s[3] += wv[i] * w[i] * ff[i];
        s[4] += ((u[i] * u[i]) + (v[i] * v[i]) + (wv[i] * wv[i])) * w[i] * ff[i];
    }
    fw[0] = s[0];
    fw[1] = s[1];
    fw[2] = s[2];
    fw[3] = s[3];
    fw[4] = 0.5 * s[4];
}

/* 2F2V   flux_kfvs.jl:654-697 (flux_conserve_2d! theory_moments.jl:411-416) */
void orc_flux_kfvs_2f2v(double *fw, double *fh, double *fb, const double *hL, const double *bL,
                        const double *hR, const double *bR, const double *u, const double *v,
                        const double *w, int n, double dt, double len, const double *shL,
                        const double *sbL, const double *shR, const double *sbR) {
    double s[5] = {0, 0, 0, 0, 0};
    for (int i = 0; i < n; i++) {
        double d = orc_heaviside(u[i]);
        double h = hL[i] * d + hR[i] * (1.0 - d);
        double b = bL[i] * d + bR[i] * (1.0 - d);
        double sh = (shL ? shL[i] : 0.0) * d + (shR ? shR[i] : 0.0) * (1.0 - d);
        double sb = (sbL ? sbL[i] : 0.0) * d + (sbR ? sbR[i] : 0.0) * (1.0 - d);
        fh[i] = (dt * u[i] * h - 0.5 * (dt * dt) * (u[i] * u[i]) * sh) * len;
        fb[i] = (dt * u[i] * b - 0.5 * (dt * dt) * (u[i] * u[i]) * sb) * len;
    }
    for (int i = 0; i < n; i++) {
        s[0] += w[i] * fh[i];
        s[1] += u[i] * w[i] * fh[i];
        s[2] += v[i] * w[i] * fh[i];
        s[3] += ((u[i] * u[i]) + (v[i] * v[i])) * w[i] * fh[i];
        s[4] += w[i] * fb[i];
    }
    fw[0] = s[0];
    fw[1] = s[1];
    fw[2] = s[2];
    fw[3] = 0.5 * (s[3] + s[4]);
}

/* 2D1F2V   flux_kfvs.jl:615-647 (flux_conserve! theory_moments.jl:403-407) */
void orc_flux_kfvs_1f2v(double *fw, double *ff, const double *fL, const double *fR, const double *u,
                        const double *v, const double *w, int n, double dt, double len,
                        const double *sfL, const double *sfR) {
    double s[4] = {0, 0, 0, 0};
    for (int i = 0; i < n; i++) {
        double d = orc_heaviside(u[i]);
        double f = fL[i] * d + fR[i] * (1.0 - d);
        double sf = (sfL ? sfL[i] : 0.0) * d + (sfR ? sfR[i] : 0.0) * (1.0 - d);
        ff[i] = (dt * u[i] * f - 0.5 * (dt * dt) * (u[i] * u[i]) * sf) * len;
    }
    for (int i = 0; i < n; i++) {
        s[0] += w[i] * ff[i];
        s[1] += u[i] * w[i] * ff[i];
        s[2] += v[i] * w[i] * ff[i];
        s[3] += ((u[i] * u[i]) + (v[i] * v[i])) * w[i] * ff[i];
    }
    fw[0] = s[0];
    fw[1] = s[1];
    fw[2] = s[2];
    fw[3] = 0.5 * s[3];
}

/* ------------------------------------------------------------------------------------------ */
/* KCU   src/Flux/flux_kcu.jl                                                                   */
/* ------------------------------------------------------------------------------------------ */

/* interface equilibrium shared by all KCU/UGKS kernels: w = rhoL<MuL> + rhoR<MuR>, prim(w) */
static double kcu_tau_extra(const double *primL, const double *primR, int np, double dt) {
    double aL = primL[0] / primL[np - 1], aR = primR[0] / primR[np - 1];
    return fabs(aL - aR) / (aL + aR) * dt * 2.0;
}

/* 1D1F1V   flux_kcu.jl:241-303 */
void orc_flux_kcu_1f1v(double *fw, double *ff, const double *wL, const double *fL, const double *wR,
                       const double *fR, const double *u, const double *w, int n, double K,
                       double gam, double mu, double om, double Pr, double dt) {
    (void)Pr;
    double primL[3], primR[3], Muv1[3], Muv2[3], ww[3], prim[3], Muv[3];
    orc_gm g1, g2, g;
    orc_conserve_prim(wL, 3, gam, primL);
    orc_conserve_prim(wR, 3, gam, primR);
    orc_gauss_moments(primL, 3, K, &g1);
    orc_moments_conserve_1d(g1.MuL, g1.Mxi, 0, 0, Muv1);
    orc_gauss_moments(primR, 3, K, &g2);
    orc_moments_conserve_1d(g2.MuR, g2.Mxi, 0, 0, Muv2);
    for (int m = 0; m < 3; m++) ww[m] = primL[0] * Muv1[m] + primR[0] * Muv2[m];
    orc_conserve_prim(ww, 3, gam, prim);
    double tau = orc_vhs_collision_time(prim[0], prim[2], mu, om) + kcu_tau_extra(primL, primR, 3, dt);
    double Mt2 = tau * (1.0 - exp(-dt / tau)), Mt1 = dt - Mt2;
    orc_gauss_moments(prim, 3, K, &g);
    orc_moments_conserve_1d(g.Mu, g.Mxi, 1, 0, Muv);
    for (int m = 0; m < 3; m++) fw[m] = Mt1 * prim[0] * Muv[m];
    double s0 = 0, s1 = 0, s2 = 0;
    for (int i = 0; i < n; i++) {
        double d = orc_heaviside(u[i]);
        double f = fL[i] * d + fR[i] * (1.0 - d);
        double c = u[i] - prim[1];
        double gi = prim[0] * sqrt(prim[2] / PI_) * exp(-prim[2] * (c * c));
        s0 += w[i] * u[i] * f;
        s1 += w[i] * (u[i] * u[i]) * f;
        s2 += w[i] * (u[i] * u[i] * u[i]) * f;
        ff[i] = Mt1 * u[i] * gi + Mt2 * u[i] * f;
    }
    fw[0] += Mt2 * s0;
    fw[1] += Mt2 * s1;
    fw[2] += Mt2 * 0.5 * s2;
}

/* 1D2F1V   flux_kcu.jl:463-530 */
void orc_flux_kcu_2f1v(double *fw, double *fh, double *fb, const double *wL, const double *hL,
                       const double *bL, const double *wR, const double *hR, const double *bR,
                       const double *u, const double *w, int n, double K, double gam, double mu,
                       double om, double Pr, double dt) {
    (void)Pr;
    double primL[3], primR[3], Muv1[3], Muv2[3], ww[3], prim[3], Muv[3];
    orc_gm g1, g2, g;
    orc_conserve_prim(wL, 3, gam, primL);
    orc_conserve_prim(wR, 3, gam, primR);
    orc_gauss_moments(primL, 3, K, &g1);
    orc_moments_conserve_1d(g1.MuL, g1.Mxi, 0, 0, Muv1);
    orc_gauss_moments(primR, 3, K, &g2);
    orc_moments_conserve_1d(g2.MuR, g2.Mxi, 0, 0, Muv2);
    for (int m = 0; m < 3; m++) ww[m] = primL[0] * Muv1[m] + primR[0] * Muv2[m];
    orc_conserve_prim(ww, 3, gam, prim);
    double tau = orc_vhs_collision_time(prim[0], prim[2], mu, om) + kcu_tau_extra(primL, primR, 3, dt);
    double Mt2 = tau * (1.0 - exp(-dt / tau)), Mt1 = dt - Mt2;
    orc_gauss_moments(prim, 3, K, &g);
    orc_moments_conserve_1d(g.Mu, g.Mxi, 1, 0, Muv);
    for (int m = 0; m < 3; m++) fw[m] = Mt1 * prim[0] * Muv[m];
    double s0 = 0, s1 = 0, s2 = 0, s3 = 0;
    for (int i = 0; i < n; i++) {
        double d = orc_heaviside(u[i]);
        double h = hL[i] * d + hR[i] * (1.0 - d);
        double b = bL[i] * d + bR[i] * (1.0 - d);
        double c = u[i] - prim[1];
        double Mh = prim[0] * sqrt(prim[2] / PI_) * exp(-prim[2] * (c * c));
        double Mb = Mh * K / (2.0 * prim[2]);
        s0 += w[i] * u[i] * h;
        s1 += w[i] * (u[i] * u[i]) * h;
        s2 += w[i] * (u[i] * u[i] * u[i]) * h;
        s3 += w[i] * u[i] * b;
        fh[i] = Mt1 * u[i] * Mh + Mt2 * u[i] * h;
        fb[i] = Mt1 * u[i] * Mb + Mt2 * u[i] * b;
    }
    fw[0] += Mt2 * s0;
    fw[1] += Mt2 * s1;
    fw[2] += Mt2 * 0.5 * (s2 + s3);
}

/* 1D1F3V   flux_kcu.jl:388-456 */
void orc_flux_kcu_1f3v(double *fw, double *ff, const double *wL, const double *fL, const double *wR,
                       const double *fR, const double *u, const double *v, const double *wv,
                       const double *w, int n, double K, double gam, double mu, double om,
                       double Pr, double dt, double len) {
    (void)Pr;
    double primL[5], primR[5], Muv1[5], Muv2[5], ww[5], prim[5], Muv[5];
    orc_gm g1, g2, g;
    orc_conserve_prim(wL, 5, gam, primL);
    orc_conserve_prim(wR, 5, gam, primR);
    orc_gauss_moments(primL, 5, K, &g1);
    orc_moments_conserve_3d(g1.MuL, g1.Mv, g1.Mw, 0, 0, 0, Muv1);
    orc_gauss_moments(primR, 5, K, &g2);
    orc_moments_conserve_3d(g2.MuR, g2.Mv, g2.Mw, 0, 0, 0, Muv2);
    for (int m = 0; m < 5; m++) ww[m] = primL[0] * Muv1[m] + primR[0] * Muv2[m];
    orc_conserve_prim(ww, 5, gam, prim);
    double tau = orc_vhs_collision_time(prim[0], prim[4], mu, om) + kcu_tau_extra(primL, primR, 5, dt);
    double Mt2 = tau * (1.0 - exp(-dt / tau)), Mt1 = dt - Mt2;
    orc_gauss_moments(prim, 5, K, &g);
    orc_moments_conserve_3d(g.Mu, g.Mv, g.Mw, 1, 0, 0, Muv);
    for (int m = 0; m < 5; m++) fw[m] = Mt1 * prim[0] * Muv[m];
    double s[5] = {0, 0, 0, 0, 0};
    double l3 = prim[4] / PI_;
    for (int i = 0; i < n; i++) {
        double d = orc_heaviside(u[i]);
        double f = fL[i] * d + fR[i] * (1.0 - d);
        double cu = u[i] - prim[1], cv = v[i] - prim[2], cw = wv[i] - prim[3];
        double gi = prim[0] * sqrt(l3 * l3 * l3) * exp(-prim[4] * (cu * cu + cv * cv + cw * cw));
        s[0] += w[i] * u[i] * f;
        s[1] += w[i] * (u[i] * u[i]) * f;
        s[2] += w[i] * u[i] * v[i] * f;
        s[3] += w[i] * u[i] * wv[i] * f;
        s[4] += w[i] * u[i] * ((u[i] * u[i]) + (v[i] * v[i]) + (wv[i] * wv[i])) * f;
        ff[i] = (Mt1 * u[i] * gi + Mt2 * u[i] * f) * len;
    }
    fw[0] += Mt2 * s[0];
    fw[1] += Mt2 * s[1];
    fw[2] += Mt2 * s[2];
    fw[3] += Mt2 * s[3];
    fw[4] += Mt2 * 0.5 * s[4];
    for (int m = 0; m < 5; m++) fw[m] *= len;
}

/* 2D2F2V   flux_kcu.jl:898-967 */
void orc_flux_kcu_2f2v(double *fw, double *fh, double *fb, const double *wL, const double *hL,
                       const double *bL, const double *wR, const double *hR, const double *bR,
                       const double *u, const double *v, const double *w, int n, double K,
                       double gam, double mu, double om, double Pr, double dt, double len) {
    (void)Pr;
    double primL[4], primR[4], Muv1[4], Muv2[4], ww[4], prim[4], Muv[4];
    orc_gm g1, g2, g;
    orc_conserve_prim(wL, 4, gam, primL);
    orc_conserve_prim(wR, 4, gam, primR);
    orc_gauss_moments(primL, 4, K, &g1);
    orc_moments_conserve_2d(g1.MuL, g1.Mv, g1.Mxi, 0, 0, 0, Muv1);
    orc_gauss_moments(primR, 4, K, &g2);
    orc_moments_conserve_2d(g2.MuR, g2.Mv, g2.Mxi, 0, 0, 0, Muv2);
    for (int m = 0; m < 4; m++) ww[m] = primL[0] * Muv1[m] + primR[0] * Muv2[m];
    orc_conserve_prim(ww, 4, gam, prim);
    double tau = orc_vhs_collision_time(prim[0], prim[3], mu, om) + kcu_tau_extra(primL, primR, 4, dt);
    double Mt2 = tau * (1.0 - exp(-dt / tau)), Mt1 = dt - Mt2;
    orc_gauss_moments(prim, 4, K, &g);
    orc_moments_conserve_2d(g.Mu, g.Mv, g.Mxi, 1, 0, 0, Muv);
    for (int m = 0; m < 4; m++) fw[m] = Mt1 * prim[0] * Muv[m] * len;
    double s[5] = {0, 0, 0, 0, 0};
    for (int i = 0; i < n; i++) {
        double d = orc_heaviside(u[i]);
        double h = hL[i] * d + hR[i] * (1.0 - d);
        double b = bL[i] * d + bR[i] * (1.0 - d);
        double cu = u[i] - prim[1], cv = v[i] - prim[2];
        double H = prim[0] * (prim[3] / PI_) * exp(-prim[3] * (cu * cu + cv * cv));
        double B = H * K / (2.0 * prim[3]);
        s[0] += w[i] * u[i] * h;
        s[1] += w[i] * (u[i] * u[i]) * h;
        s[2] += w[i] * v[i] * u[i] * h;
        s[3] += w[i] * u[i] * ((u[i] * u[i]) + (v[i] * v[i])) * h;
        s[4] += w[i] * u[i] * b;
        fh[i] = (Mt1 * u[i] * H + Mt2 * u[i] * h) * len;
        fb[i] = (Mt1 * u[i] * B + Mt2 * u[i] * b) * len;
    }
    fw[0] += Mt2 * s[0] * len;
    fw[1] += Mt2 * s[1] * len;
    fw[2] += Mt2 * s[2] * len;
    fw[3] += Mt2 * 0.5 * (s[3] + s[4]) * len;
}

/* 2D1F2V   flux_kcu.jl:828-891 */
void orc_flux_kcu_1f2v(double *fw, double *ff, const double *wL, const double *fL, const double *wR,
                       const double *fR, const double *u, const double *v, const double *w, int n,
                       double K, double gam, double mu, double om, double Pr, double dt, double len) {
    (void)Pr;
    double primL[4], primR[4], Muv1[4], Muv2[4], ww[4], prim[4], Muv[4];
    orc_gm g1, g2, g;
    orc_conserve_prim(wL, 4, gam, primL);
    orc_conserve_prim(wR, 4, gam, primR);
    orc_gauss_moments(primL, 4, K, &g1);
    orc_moments_conserve_2d(g1.MuL, g1.Mv, g1.Mxi, 0, 0, 0, Muv1);
    orc_gauss_moments(primR, 4, K, &g2);
    orc_moments_conserve_2d(g2.MuR, g2.Mv, g2.Mxi, 0, 0, 0, Muv2);
    for (int m = 0; m < 4; m++) ww[m] = primL[0] * Muv1[m] + primR[0] * Muv2[m];
    orc_conserve_prim(ww, 4, gam, prim);
    double tau = orc_vhs_collision_time(prim[0], prim[3], mu, om) + kcu_tau_extra(primL, primR, 4, dt);
    double Mt2 = tau * (1.0 - exp(-dt / tau)), Mt1 = dt - Mt2;
    orc_gauss_moments(prim, 4, K, &g);
    orc_moments_conserve_2d(g.Mu, g.Mv, g.Mxi, 1, 0, 0, Muv);
    for (int m = 0; m < 4; m++) fw[m] = Mt1 * prim[0] * Muv[m] * len;
    double s[4] = {0, 0, 0, 0};
    for (int i = 0; i < n; i++) {
        double d = orc_heaviside(u[i]);
        double f = fL[i] * d + fR[i] * (1.0 - d);
        double cu = u[i] - prim[1], cv = v[i] - prim[2];
        double g0 = prim[0] * (prim[3] / PI_) * exp(-prim[3] * (cu * cu + cv * cv));
        s[0] += w[i] * u[i] * f;
        s[1] += w[i] * (u[i] * u[i]) * f;
        s[2] += w[i] * v[i] * u[i] * f;
        s[3] += w[i] * u[i] * ((u[i] * u[i]) + (v[i] * v[i])) * f;
        ff[i] = (Mt1 * u[i] * g0 + Mt2 * u[i] * f) * len;
    }
    fw[0] += Mt2 * s[0] * len;
    fw[1] += Mt2 * s[1] * len;
    fw[2] += Mt2 * s[2] * len;
    fw[3] += Mt2 * 0.5 * s[3] * len;
}

/* ------------------------------------------------------------------------------------------ */
/* UGKS   src/Flux/flux_ugks.jl                                                                 */
/* ------------------------------------------------------------------------------------------ */

/* integral time coefficients   flux_ugks.jl:66-71 */
static void ugks_mt(double tau, double dt, double *Mt) {
    Mt[3] = tau * (1.0 - exp(-dt / tau));
    Mt[4] = -tau * dt * exp(-dt / tau) + tau * Mt[3];
    Mt[0] = dt - Mt[3];
    Mt[1] = -tau * Mt[0] + Mt[4];
    Mt[2] = (dt * dt) / 2.0 - tau * Mt[0];
}

/* 1D2F1V   flux_ugks.jl:8-114 */
void orc_flux_ugks_2f1v(double *fw, double *fh, double *fb, const double *wL, const double *hL,
                        const double *bL, const double *wR, const double *hR, const double *bR,
                        const double *u, const double *w, int n, double K, double gam, double mu,
                        double om, double Pr, double dt, double dxL, double dxR, const double *shL,
                        const double *sbL, const double *shR, const double *sbR) {
    (void)Pr;
    double primL[3], primR[3], Muv1[3], Muv2[3], ww[3], prim[3], t[3];
    double aL[3], aR[3], aT[3], MauL[3], MauR[3], MauT[3], Muv[3], Mt[5];
    orc_gm g1, g2, g;
    orc_conserve_prim(wL, 3, gam, primL);
    orc_conserve_prim(wR, 3, gam, primR);
    orc_gauss_moments(primL, 3, K, &g1);
    orc_moments_conserve_1d(g1.MuL, g1.Mxi, 0, 0, Muv1);
    orc_gauss_moments(primR, 3, K, &g2);
    orc_moments_conserve_1d(g2.MuR, g2.Mxi, 0, 0, Muv2);
    for (int m = 0; m < 3; m++) ww[m] = primL[0] * Muv1[m] + primR[0] * Muv2[m];
    orc_conserve_prim(ww, 3, gam, prim);
    for (int m = 0; m < 3; m++) t[m] = (ww[m] - wL[m]) / dxL;
    orc_pdf_slope(prim, t, 3, K, aL);
    for (int m = 0; m < 3; m++) t[m] = (wR[m] - ww[m]) / dxR;
    orc_pdf_slope(prim, t, 3, K, aR);
    orc_gauss_moments(prim, 3, K, &g);
    orc_moments_conserve_slope_1d(aL, g.MuL, g.Mxi, 1, MauL);
    orc_moments_conserve_slope_1d(aR, g.MuR, g.Mxi, 1, MauR);
    for (int m = 0; m < 3; m++) t[m] = -prim[0] * (MauL[m] + MauR[m]);
    orc_pdf_slope(prim, t, 3, K, aT);
    double tau = orc_vhs_collision_time(prim[0], prim[2], mu, om);
    ugks_mt(tau, dt, Mt);
    orc_moments_conserve_1d(g.Mu, g.Mxi, 1, 0, Muv);
    orc_moments_conserve_slope_1d(aL, g.MuL, g.Mxi, 2, MauL);
    orc_moments_conserve_slope_1d(aR, g.MuR, g.Mxi, 2, MauR);
    orc_moments_conserve_slope_1d(aT, g.Mu, g.Mxi, 1, MauT);
    for (int m = 0; m < 3; m++)
        fw[m] = Mt[0] * prim[0] * Muv[m] + Mt[1] * prim[0] * (MauL[m] + MauR[m]) + Mt[2] * prim[0] * MauT[m];
    /* moments of the free-transport part: s0 = Σωuh, s1 = Σωu²sh, s2 = Σωu²h, s3 = Σωu³sh,
     * s4 = Σωu³h, s5 = Σωub, s6 = Σωu⁴sh, s7 = Σωu²sb */
    double s[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    double Mxi2 = g.Mxi[2];
    for (int i = 0; i < n; i++) {
        double d = orc_heaviside(u[i]);
        double h = hL[i] * d + hR[i] * (1.0 - d);
        double b = bL[i] * d + bR[i] * (1.0 - d);
        double sh = (shL ? shL[i] : 0.0) * d + (shR ? shR[i] : 0.0) * (1.0 - d);
        double sb = (sbL ? sbL[i] : 0.0) * d + (sbR ? sbR[i] : 0.0) * (1.0 - d);
        double c = u[i] - prim[1];
        double H = prim[0] * sqrt(prim[2] / PI_) * exp(-prim[2] * (c * c));
        double B = H * K / (2.0 * prim[2]);
        double uu = u[i] * u[i];
        s[0] += w[i] * u[i] * h;
        s[1] += w[i] * uu * sh;
        s[2] += w[i] * uu * h;
        s[3] += w[i] * (uu * u[i]) * sh;
        s[4] += w[i] * (uu * u[i]) * h;
        s[5] += w[i] * u[i] * b;
        s[6] += w[i] * (uu * uu) * sh;
        s[7] += w[i] * uu * sb;
        fh[i] = Mt[0] * u[i] * H +
                Mt[1] * uu * (aL[0] * H + aL[1] * u[i] * H + 0.5 * aL[2] * (uu * H + B)) * d +
                Mt[1] * uu * (aR[0] * H + aR[1] * u[i] * H + 0.5 * aR[2] * (uu * H + B)) * (1.0 - d) +
                Mt[2] * u[i] * (aT[0] * H + aT[1] * u[i] * H + 0.5 * aT[2] * (uu * H + B)) +
                Mt[3] * u[i] * h - Mt[4] * uu * sh;
        fb[i] = Mt[0] * u[i] * B +
                Mt[1] * uu * (aL[0] * B + aL[1] * u[i] * B + 0.5 * aL[2] * (uu * B + Mxi2 * H)) * d +
                Mt[1] * uu * (aR[0] * B + aR[1] * u[i] * B + 0.5 * aR[2] * (uu * B + Mxi2 * H)) * (1.0 - d) +
                Mt[2] * u[i] * (aT[0] * B + aT[1] * u[i] * B + 0.5 * aT[2] * (uu * B + Mxi2 * H)) +
                Mt[3] * u[i] * b - Mt[4] * uu * sb;
    }
    fw[0] += Mt[3] * s[0] - Mt[4] * s[1];
    fw[1] += Mt[3] * s[2] - Mt[4] * s[3];
    fw[2] += Mt[3] * 0.5 * (s[4] + s[5]) - Mt[4] * 0.5 * (s[6] + s[7]);
}

/* 2D2F2V   flux_ugks.jl:121-253 */
void orc_flux_ugks_2f2v(double *fw, double *fh, double *fb, const double *wL, const double *hL,
                        const double *bL, const double *wR, const double *hR, const double *bR,
                        const double *u, const double *v, const double *w, int n, double K,
                        double gam, double mu, double om, double Pr, double dt, double dxL,
                        double dxR, double len, const double *shL, const double *sbL,
                        const double *shR, const double *sbR) {
    (void)Pr;
    double primL[4], primR[4], Muv1[4], Muv2[4], ww[4], prim[4], t[4];
    double aL[4], aR[4], aT[4], MauL[4], MauR[4], MauT[4], Muv[4], Mt[5];
    orc_gm g1, g2, g;
    orc_conserve_prim(wL, 4, gam, primL);
    orc_conserve_prim(wR, 4, gam, primR);
    orc_gauss_moments(primL, 4, K, &g1);
    orc_moments_conserve_2d(g1.MuL, g1.Mv, g1.Mxi, 0, 0, 0, Muv1);
    orc_gauss_moments(primR, 4, K, &g2);
    orc_moments_conserve_2d(g2.MuR, g2.Mv, g2.Mxi, 0, 0, 0, Muv2);
    for (int m = 0; m < 4; m++) ww[m] = primL[0] * Muv1[m] + primR[0] * Muv2[m];
    orc_conserve_prim(ww, 4, gam, prim);
    for (int m = 0; m < 4; m++) t[m] = (ww[m] - wL[m]) / dxL;
    orc_pdf_slope(prim, t, 4, K, aL);
    for (int m = 0; m < 4; m++) t[m] = (wR[m] - ww[m]) / dxR;
    orc_pdf_slope(prim, t, 4, K, aR);
    orc_gauss_moments(prim, 4, K, &g);
    orc_moments_conserve_slope_2d(aL, g.MuL, g.Mv, g.Mxi, 1, 0, MauL);
    orc_moments_conserve_slope_2d(aR, g.MuR, g.Mv, g.Mxi, 1, 0, MauR);
    for (int m = 0; m < 4; m++) t[m] = -prim[0] * (MauL[m] + MauR[m]);
    orc_pdf_slope(prim, t, 4, K, aT);
    double tau = orc_vhs_collision_time(prim[0], prim[3], mu, om);
    ugks_mt(tau, dt, Mt);
    orc_moments_conserve_2d(g.Mu, g.Mv, g.Mxi, 1, 0, 0, Muv);
    orc_moments_conserve_slope_2d(aL, g.MuL, g.Mv, g.Mxi, 2, 0, MauL);
    orc_moments_conserve_slope_2d(aR, g.MuR, g.Mv, g.Mxi, 2, 0, MauR);
    orc_moments_conserve_slope_2d(aT, g.Mu, g.Mv, g.Mxi, 1, 0, MauT);
    for (int m = 0; m < 4; m++)
        fw[m] = Mt[0] * prim[0] * Muv[m] + Mt[1] * prim[0] * (MauL[m] + MauR[m]) + Mt[2] * prim[0] * MauT[m];
    double s[10];
    memset(s, 0, sizeof s);
    double Mxi2 = g.Mxi[2];
    for (int i = 0; i < n; i++) {
        double d = orc_heaviside(u[i]);
        double h = hL[i] * d + hR[i] * (1.0 - d);
        double b = bL[i] * d + bR[i] * (1.0 - d);
        double sh = (shL ? shL[i] : 0.0) * d + (shR ? shR[i] : 0.0) * (1.0 - d);
        double sb = (sbL ? sbL[i] : 0.0) * d + (sbR ? sbR[i] : 0.0) * (1.0 - d);
        double cu = u[i] - prim[1], cv = v[i] - prim[2];
        double H = prim[0] * (prim[3] / PI_) * exp(-prim[3] * (cu * cu + cv * cv));
        double B = H * K / (2.0 * prim[3]);
        double uu = u[i] * u[i], q2 = uu + v[i] * v[i];
        s[0] += w[i] * u[i] * h;
        s[1] += w[i] * uu * sh;
        s[2] += w[i] * uu * h;
        s[3] += w[i] * (uu * u[i]) * sh;
        s[4] += w[i] * v[i] * u[i] * h;
        s[5] += w[i] * v[i] * uu * sh;
        s[6] += w[i] * u[i] * q2 * h;
        s[7] += w[i] * u[i] * b;
        s[8] += w[i] * uu * q2 * sh;
        s[9] += w[i] * uu * sb;
        fh[i] = Mt[0] * u[i] * H +
                Mt[1] * uu * (aL[0] * H + aL[1] * u[i] * H + aL[2] * v[i] * H + 0.5 * aL[3] * (q2 * H + B)) * d +
                Mt[1] * uu * (aR[0] * H + aR[1] * u[i] * H + aR[2] * v[i] * H + 0.5 * aR[3] * (q2 * H + B)) * (1.0 - d) +
                Mt[2] * u[i] * (aT[0] * H + aT[1] * u[i] * H + aT[2] * v[i] * H + 0.5 * aT[3] * (q2 * H + B)) +
                Mt[3] * u[i] * h - Mt[4] * uu * sh;
        fb[i] = Mt[0] * u[i] * B +
                Mt[1] * uu * (aL[0] * B + aL[1] * u[i] * B + aL[2] * v[i] * B + 0.5 * aL[3] * (q2 * B + Mxi2 * H)) * d +
                Mt[1] * uu * (aR[0] * B + aR[1] * u[i] * B + aR[2] * v[i] * B + 0.5 * aR[3] * (q2 * B + Mxi2 * H)) * (1.0 - d) +
                Mt[2] * u[i] * (aT[0] * B + aT[1] * u[i] * B + aT[2] * v[i] * B + 0.5 * aT[3] * (q2 * B + Mxi2 * H)) +
                Mt[3] * u[i] * b - Mt[4] * uu * sb;
        fh[i] *= len;
        fb[i] *= len;
    }
    fw[0] += Mt[3] * s[0] - Mt[4] * s[1];
    fw[1] += Mt[3] * s[2] - Mt[4] * s[3];
    fw[2] += Mt[3] * s[4] - Mt[4] * s[5];
    fw[3] += Mt[3] * 0.5 * (s[6] + s[7]) - Mt[4] * 0.5 * (s[8] + s[9]);
    for (int m = 0; m < 4; m++) fw[m] *= len;
}

/* ------------------------------------------------------------------------------------------ */
/* Maxwell diffuse walls   src/Flux/flux_boundary.jl                                            */
/* ------------------------------------------------------------------------------------------ */

/* 1D2F1V   flux_boundary.jl:39-61 (bc = [rho_w, U_w, lambda_w], only U_w and lambda_w used) */
void orc_flux_boundary_maxwell_2f1v(double *fw, double *fh, double *fb, const double *bc,
                                    const double *h, const double *b, const double *u,
                                    const double *w, int n, double K, double dt, int rot) {
    double SF = 0, SG = 0;
    for (int i = 0; i < n; i++) {
        double d = orc_heaviside(u[i] * rot);
        double c = u[i] - bc[1];
        SF += w[i] * u[i] * h[i] * (1.0 - d);
        SG += w[i] * u[i] * exp(-bc[2] * (c * c)) * d;
    }
    SG *= pow(bc[2] / PI_, 0.5);
    double prim[3] = {-SF / SG, bc[1], bc[2]};
    double s0 = 0, s1 = 0, s2 = 0, s3 = 0;
    for (int i = 0; i < n; i++) {
        double d = orc_heaviside(u[i] * rot);
        double c = u[i] - prim[1];
        double H = prim[0] * sqrt(prim[2] / PI_) * exp(-prim[2] * (c * c));
        double B = H * K / (2.0 * prim[2]);
        double hW = H * d + h[i] * (1.0 - d);
        double bW = B * d + b[i] * (1.0 - d);
        s0 += w[i] * u[i] * hW;
        s1 += w[i] * (u[i] * u[i]) * hW;
        s2 += w[i] * u[i] * (hW * (u[i] * u[i]));
        s3 += w[i] * u[i] * bW;
        fh[i] = u[i] * hW * dt;
        fb[i] = u[i] * bW * dt;
    }
    fw[0] = s0 * dt;
    fw[1] = s1 * dt;
    fw[2] = (0.5 * s2 + 0.5 * s3) * dt;
}

/* 2D2F2V   flux_boundary.jl:190-220 (bc in the local frame, velocities u=vn, v=vt) */
void orc_flux_boundary_maxwell_2f2v(double *fw, double *fh, double *fb, const double *bc,
                                    const double *h, const double *b, const double *u,
                                    const double *v, const double *w, int n, double K, double dt,
                                    double len, int rot) {
    double SF = 0, SG = 0;
    for (int i = 0; i < n; i++) {
        double d = orc_heaviside(u[i] * rot);
        double cu = u[i] - bc[1], cv = v[i] - bc[2];
        SF += w[i] * u[i] * h[i] * (1.0 - d);
        SG += w[i] * u[i] * exp(-bc[3] * (cu * cu + cv * cv)) * d;
    }
    SG *= (bc[3] / PI_);
    double prim[4] = {-SF / SG, bc[1], bc[2], bc[3]};
    double s[5] = {0, 0, 0, 0, 0};
    for (int i = 0; i < n; i++) {
        double d = orc_heaviside(u[i] * rot);
        double cu = u[i] - prim[1], cv = v[i] - prim[2];
        double H = prim[0] * (prim[3] / PI_) * exp(-prim[3] * (cu * cu + cv * cv));
        double B = H * K / (2.0 * prim[3]);
        double hW = H * d + h[i] * (1.0 - d);
        double bW = B * d + b[i] * (1.0 - d);
        s[0] += w[i] * u[i] * hW;
        s[1] += w[i] * (u[i] * u[i]) * hW;
        s[2] += w[i] * v[i] * (hW * u[i]);
        s[3] += w[i] * u[i] * (hW * ((u[i] * u[i]) + (v[i] * v[i])));
        s[4] += w[i] * u[i] * bW;
        fh[i] = u[i] * hW * len * dt;
        fb[i] = u[i] * bW * len * dt;
    }
    fw[0] = s[0] * len * dt;
    fw[1] = s[1] * len * dt;
    fw[2] = s[2] * len * dt;
    fw[3] = (0.5 * s[3] + 0.5 * s[4]) * len * dt;
}

/* 2D1F2V   flux_boundary.jl:160-183 (bc in the local frame, velocities u=vn, v=vt) */
void orc_flux_boundary_maxwell_1f2v(double *fw, double *fh, const double *bc, const double *h,
                                    const double *u, const double *v, const double *w, int n,
                                    double dt, double len, int rot) {
    double SF = 0, SG = 0;
    for (int i = 0; i < n; i++) {
        double d = orc_heaviside(u[i] * rot);
        double cu = u[i] - bc[1], cv = v[i] - bc[2];
        SF += w[i] * u[i] * h[i] * (1.0 - d);
        SG += w[i] * u[i] * exp(-bc[3] * (cu * cu + cv * cv)) * d;
    }
    SG *= (bc[3] / PI_);
    double prim[4] = {-SF / SG, bc[1], bc[2], bc[3]};
    double s[4] = {0, 0, 0, 0};
    for (int i = 0; i < n; i++) {
        double d = orc_heaviside(u[i] * rot);
        double cu = u[i] - prim[1], cv = v[i] - prim[2];
        double H = prim[0] * (prim[3] / PI_) * exp(-prim[3] * (cu * cu + cv * cv));
        double hW = H * d + h[i] * (1.0 - d);
        s[0] += w[i] * u[i] * hW;
        s[1] += w[i] * (u[i] * u[i]) * hW;
        s[2] += w[i] * v[i] * (hW * u[i]);
        s[3] += w[i] * u[i] * (hW * ((u[i] * u[i]) + (v[i] * v[i])));
        fh[i] = u[i] * hW * len * dt;
    }
    fw[0] = s[0] * len * dt;
    fw[1] = s[1] * len * dt;
    fw[2] = s[2] * len * dt;
    fw[3] = 0.5 * s[3] * len * dt;
}

/* ------------------------------------------------------------------------------------------ */
/* step!   src/Solver/step_1f.jl, step_2f.jl                                                    */
/* ------------------------------------------------------------------------------------------ */

/* 1D1F1V   step_1f.jl:6-55 (heat_flux theory_moments.jl:523, shakhov theory_equilibrium.jl:178-186) */
void orc_step_1f1v(double *w, double *prim, double *f, const double *fwL, const double *ffL,
                   const double *fwR, const double *ffR, const double *u, const double *weights,
                   int n, double gam, double mu, double om, double Pr, double dx, double dt,
                   double *RES, double *AVG, int coll) {
    double w_old[3] = {w[0], w[1], w[2]};
    double *S = (double *)calloc(n, sizeof(double));
    if (coll == ORC_SHAKHOV) {
        double q = 0;
        for (int i = 0; i < n; i++) {
            double c = u[i] - prim[1];
            q += weights[i] * c * (c * c) * f[i];
        }
        q *= 0.5;
        for (int i = 0; i < n; i++) {
            double c = u[i] - prim[1];
            double Mo = prim[0] * sqrt(prim[2] / PI_) * exp(-prim[2] * (c * c));
            S[i] = 0.8 * (1.0 - Pr) * (prim[2] * prim[2]) / prim[0] * c * q *
                   (2.0 * prim[2] * (c * c) - 5.0) * Mo;
        }
    }
    for (int m = 0; m < 3; m++) w[m] += (fwL[m] - fwR[m]) / dx;
    orc_conserve_prim(w, 3, gam, prim);
    for (int m = 0; m < 3; m++) {
        RES[m] += (w[m] - w_old[m]) * (w[m] - w_old[m]);
        AVG[m] += fabs(w[m]);
    }
    double tau = orc_vhs_collision_time(prim[0], prim[2], mu, om);
    for (int i = 0; i < n; i++) {
        double c = u[i] - prim[1];
        double M = prim[0] * sqrt(prim[2] / PI_) * exp(-prim[2] * (c * c)) + S[i];
        f[i] = (f[i] + (ffL[i] - ffR[i]) / dx + dt / tau * M) / (1.0 + dt / tau);
    }
    free(S);
}

/* 1D2F1V   step_2f.jl:6-64 (heat_flux theory_moments.jl:530-531, shakhov theory_equilibrium.jl:193-206) */
void orc_step_2f1v(double *w, double *prim, double *h, double *b, const double *fwL,
                   const double *fhL, const double *fbL, const double *fwR, const double *fhR,
                   const double *fbR, const double *u, const double *weights, int n, double K,
                   double gam, double mu, double om, double Pr, double dx, double dt, double *RES,
                   double *AVG, int coll) {
    double w_old[3] = {w[0], w[1], w[2]};
    double *SH = (double *)calloc(2 * n, sizeof(double));
    double *SB = SH + n;
    if (coll == ORC_SHAKHOV) {
        double q1 = 0, q2 = 0;
        for (int i = 0; i < n; i++) {
            double c = u[i] - prim[1];
            q1 += weights[i] * c * (c * c) * h[i];
            q2 += weights[i] * c * b[i];
        }
        double q = 0.5 * (q1 + q2);
        for (int i = 0; i < n; i++) {
            double c = u[i] - prim[1];
            double MH = prim[0] * sqrt(prim[2] / PI_) * exp(-prim[2] * (c * c));
            double MB = MH * K / (2.0 * prim[2]);
            double pre = 0.8 * (1.0 - Pr) * (prim[2] * prim[2]) / prim[0] * c * q;
            SH[i] = pre * (2.0 * prim[2] * (c * c) + K - 5.0) * MH;
            SB[i] = pre * (2.0 * prim[2] * (c * c) + K - 3.0) * MB;
        }
    }
    for (int m = 0; m < 3; m++) w[m] += (fwL[m] - fwR[m]) / dx;
    orc_conserve_prim(w, 3, gam, prim);
    for (int m = 0; m < 3; m++) {
        RES[m] += (w[m] - w_old[m]) * (w[m] - w_old[m]);
        AVG[m] += fabs(w[m]);
    }
    double tau = orc_vhs_collision_time(prim[0], prim[2], mu, om);
    for (int i = 0; i < n; i++) {
        double c = u[i] - prim[1];
        double MH = prim[0] * sqrt(prim[2] / PI_) * exp(-prim[2] * (c * c));
        double MB = MH * K / (2.0 * prim[2]);
        MH += SH[i];
        MB += SB[i];
        h[i] = (h[i] + (fhL[i] - fhR[i]) / dx + dt / tau * MH) / (1.0 + dt / tau);
        b[i] = (b[i] + (fbL[i] - fbR[i]) / dx + dt / tau * MB) / (1.0 + dt / tau);
    }
    free(SH);
}

/* 1D1F3V   step_1f.jl:62-115.  The Shakhov branch there calls shakhov(..., K) with K undefined
 * (step_1f.jl:91), so KitBase throws; this returns -1 for it and steps BGK otherwise. */
int orc_step_1f3v(double *w, double *prim, double *f, const double *fwL, const double *ffL,
                  const double *fwR, const double *ffR, const double *u, const double *v,
                  const double *wv, const double *weights, int n, double gam, double mu,
                  double om, double Pr, double dx, double dt, double *RES, double *AVG, int coll) {
    (void)weights;
    (void)Pr;
    if (coll == ORC_SHAKHOV) return -1;
    double w_old[5];
    memcpy(w_old, w, sizeof w_old);
    for (int m = 0; m < 5; m++) w[m] += (fwL[m] - fwR[m]) / dx;
    orc_conserve_prim(w, 5, gam, prim);
    for (int m = 0; m < 5; m++) {
        RES[m] += (w[m] - w_old[m]) * (w[m] - w_old[m]);
        AVG[m] += fabs(w[m]);
    }
    double tau = orc_vhs_collision_time(prim[0], prim[4], mu, om);
    double l3 = prim[4] / PI_;
    for (int i = 0; i < n; i++) {
        double cu = u[i] - prim[1], cv = v[i] - prim[2], cw = wv[i] - prim[3];
        double M = prim[0] * sqrt(l3 * l3 * l3) * exp(-prim[4] * (cu * cu + cv * cv + cw * cw));
        f[i] = (f[i] + (ffL[i] - ffR[i]) / dx + dt / tau * M) / (1.0 + dt / tau);
    }
    return 0;
}

/* 2D1F2V   step_1f.jl:164-221.  The Shakhov branch there reads heat_flux(h, b, ...) with b undefined
 * (step_1f.jl:194), so KitBase throws; this returns -1 for it and steps BGK otherwise. */
int orc_step_1f2v(double *w, double *prim, double *f, const double *fwL, const double *ffL,
                  const double *fwR, const double *ffR, const double *fwD, const double *ffD,
                  const double *fwU, const double *ffU, const double *u, const double *v,
                  const double *weights, int n, double gam, double mu, double om, double Pr,
                  double ds, double dt, double *RES, double *AVG, int coll) {
    (void)weights;
    (void)Pr;
    if (coll == ORC_SHAKHOV) return -1;
    double w_old[4] = {w[0], w[1], w[2], w[3]};
    for (int m = 0; m < 4; m++) w[m] += (fwL[m] - fwR[m] + fwD[m] - fwU[m]) / ds;
    orc_conserve_prim(w, 4, gam, prim);
    for (int m = 0; m < 4; m++) {
        RES[m] += (w[m] - w_old[m]) * (w[m] - w_old[m]);
        AVG[m] += fabs(w[m]);
    }
    double tau = orc_vhs_collision_time(prim[0], prim[3], mu, om);
    for (int i = 0; i < n; i++) {
        double cu = u[i] - prim[1], cv = v[i] - prim[2];
        double M = prim[0] * (prim[3] / PI_) * exp(-prim[3] * (cu * cu + cv * cv));
        f[i] = (f[i] + (ffL[i] - ffR[i] + ffD[i] - ffU[i]) / ds + dt / tau * M) / (1.0 + dt / tau);
    }
    return 0;
}

/* 2D2F2V   step_2f.jl:179-249 (heat_flux theory_moments.jl:567-578, shakhov theory_equilibrium.jl:227-238) */
void orc_step_2f2v(double *w, double *prim, double *h, double *b, const double *fwL,
                   const double *fhL, const double *fbL, const double *fwR, const double *fhR,
                   const double *fbR, const double *fwD, const double *fhD, const double *fbD,
                   const double *fwU, const double *fhU, const double *fbU, const double *u,
                   const double *v, const double *weights, int n, double K, double gam, double mu,
                   double om, double Pr, double ds, double dt, double *RES, double *AVG, int coll) {
    double w_old[4] = {w[0], w[1], w[2], w[3]};
    double *SH = (double *)calloc(2 * n, sizeof(double));
    double *SB = SH + n;
    if (coll == ORC_SHAKHOV) {
        double q[2] = {0, 0}, qa = 0, qb = 0, qc = 0, qd = 0;
        for (int i = 0; i < n; i++) {
            double cu = u[i] - prim[1], cv = v[i] - prim[2];
            double c2 = cu * cu + cv * cv;
            qa += weights[i] * cu * c2 * h[i];
            qb += weights[i] * cu * b[i];
            qc += weights[i] * cv * c2 * h[i];
            qd += weights[i] * cv * b[i];
        }
        q[0] = 0.5 * (qa + qb);
        q[1] = 0.5 * (qc + qd);
        for (int i = 0; i < n; i++) {
            double cu = u[i] - prim[1], cv = v[i] - prim[2];
            double c2 = cu * cu + cv * cv;
            double MH = prim[0] * (prim[3] / PI_) * exp(-prim[3] * c2);
            double MB = MH * K / (2.0 * prim[3]);
            double pre = 0.8 * (1.0 - Pr) * (prim[3] * prim[3]) / prim[0] * (cu * q[0] + cv * q[1]);
            SH[i] = pre * (2.0 * prim[3] * c2 + K - 5.0) * MH;
            SB[i] = pre * (2.0 * prim[3] * c2 + K - 3.0) * MB;
        }
    }
    for (int m = 0; m < 4; m++) w[m] += (fwL[m] - fwR[m] + fwD[m] - fwU[m]) / ds;
    orc_conserve_prim(w, 4, gam, prim);
    for (int m = 0; m < 4; m++) {
        RES[m] += (w[m] - w_old[m]) * (w[m] - w_old[m]);
        AVG[m] += fabs(w[m]);
    }
    double tau = orc_vhs_collision_time(prim[0], prim[3], mu, om);
    for (int i = 0; i < n; i++) {
        double cu = u[i] - prim[1], cv = v[i] - prim[2];
        double MH = prim[0] * (prim[3] / PI_) * exp(-prim[3] * (cu * cu + cv * cv));
        double MB = MH * K / (2.0 * prim[3]);
        MH += SH[i];
        MB += SB[i];
        h[i] = (h[i] + (fhL[i] - fhR[i] + fhD[i] - fhU[i]) / ds + dt / tau * MH) / (1.0 + dt / tau);
        b[i] = (b[i] + (fbL[i] - fbR[i] + fbD[i] - fbU[i]) / ds + dt / tau * MB) / (1.0 + dt / tau);
    }
    free(SH);
}

/* ------------------------------------------------------------------------------------------ */
/* 1D whole solves                                                                             */
/* ------------------------------------------------------------------------------------------ */

static int nw_1d(const orc_solver1d *S) { return S->vdim == 3 ? 5 : 3; }

/* reconstruct!   src/Solver/solver_reconstruction.jl:56-153 (reconstruct3!  reconstruction.jl:136-151) */
int orc_reconstruct_1d(orc_solver1d *S) {
    if (S->interp_order == 1) return 0;
    const int ng = S->ng, nx = S->nx, nw = nw_1d(S), NF = S->nf * S->nv;
    const int lo = ng > 0 ? 1 : 2, hi = ng > 0 ? nx : nx - 1;
#pragma omp parallel for schedule(static)
    for (int i = lo; i <= hi; i++) {
        int k = i - 1 + ng;
        double dL = 0.5 * (S->dx[k - 1] + S->dx[k]), dR = 0.5 * (S->dx[k] + S->dx[k + 1]);
        for (int m = 0; m < nw; m++) {
            double a = S->w[(k - 1) * nw + m], c = S->w[k * nw + m], e = S->w[(k + 1) * nw + m];
            S->sw[k * nw + m] = orc_limiter(S->limiter, (c - a) / dL, (e - c) / dR);
        }
        const double *fa = S->f + (size_t)(k - 1) * NF, *fc = S->f + (size_t)k * NF, *fe = S->f + (size_t)(k + 1) * NF;
        double *sf = S->sf + (size_t)k * NF;
        for (int q = 0; q < NF; q++) sf[q] = orc_limiter(S->limiter, (fc[q] - fa[q]) / dL, (fe[q] - fc[q]) / dR);
    }
    return 0;
}

/* evolve!   src/Solver/solver_evolution.jl:39-89 with the cell-level wrappers flux_kfvs.jl:9-64,
 * flux_kcu.jl:12-102 / :196-230 and the UGKS harness (wL = w + dx/2 sw, as solver_evolution.jl:285-312);
 * then evolve_boundary! (evolve_boundary.jl:4-47) for 2F. */
int orc_evolve_1d(orc_solver1d *S, double dt) {
    const int ng = S->ng, nx = S->nx, nw = nw_1d(S), nv = S->nv, nf = S->nf, NF = nf * nv;
    const int lo = ng > 0 ? 1 : 2, hi = ng > 0 ? nx + 1 : nx;
    if (S->flux == ORC_UGKS && nf != 2) return -2; /* KitBase has no 1F UGKS kernel for 1V/3V */
    int err = 0;
#pragma omp parallel
    {
        double *L = (double *)malloc(sizeof(double) * 2 * NF);
        double *R = L + NF;
#pragma omp for schedule(static)
        for (int i = lo; i <= hi; i++) {
            int kL = i - 2 + ng, kR = i - 1 + ng;
            double dxL = 0.5 * S->dx[kL], dxR = 0.5 * S->dx[kR];
            const double *fL0 = S->f + (size_t)kL * NF, *fR0 = S->f + (size_t)kR * NF;
            const double *sL = S->sf + (size_t)kL * NF, *sR = S->sf + (size_t)kR * NF;
            double *fw = S->fw + (size_t)(i - 1) * nw, *ff = S->ff + (size_t)(i - 1) * NF;
            double wL[5], wR[5];
            for (int q = 0; q < NF; q++) {
                L[q] = fL0[q] + sL[q] * dxL;
                R[q] = fR0[q] - sR[q] * dxR;
            }
            for (int m = 0; m < nw; m++) {
                wL[m] = S->w[kL * nw + m] + S->sw[kL * nw + m] * dxL;
                wR[m] = S->w[kR * nw + m] - S->sw[kR * nw + m] * dxR;
            }
            if (S->vdim == 1 && nf == 1) {
                if (S->flux == ORC_KFVS)
                    orc_flux_kfvs_1f1v(fw, ff, L, R, S->u, S->weights, nv, dt, sL, sR);
                else
                    orc_flux_kcu_1f1v(fw, ff, wL, L, wR, R, S->u, S->weights, nv, S->K, S->gam, S->mu, S->om, S->Pr, dt);
            } else if (S->vdim == 1 && nf == 2) {
                if (S->flux == ORC_KFVS)
                    orc_flux_kfvs_2f1v(fw, ff, ff + nv, L, L + nv, R, R + nv, S->u, S->weights, nv, dt, sL, sL + nv, sR, sR + nv);
                else if (S->flux == ORC_KCU)
                    orc_flux_kcu_2f1v(fw, ff, ff + nv, wL, L, L + nv, wR, R, R + nv, S->u, S->weights, nv, S->K, S->gam, S->mu, S->om, S->Pr, dt);
                else
                    orc_flux_ugks_2f1v(fw, ff, ff + nv, wL, L, L + nv, wR, R, R + nv, S->u, S->weights, nv, S->K, S->gam, S->mu, S->om, S->Pr, dt, dxL, dxR, sL, sL + nv, sR, sR + nv);
            } else if (S->vdim == 3 && nf == 1) {
                if (S->flux == ORC_KFVS)
                    orc_flux_kfvs_1f3v(fw, ff, L, R, S->u, S->v, S->wv, S->weights, nv, dt, 1.0, sL, sR);
                else
                    orc_flux_kcu_1f3v(fw, ff, wL, L, wR, R, S->u, S->v, S->wv, S->weights, nv, S->K, S->gam, S->mu, S->om, S->Pr, dt, 1.0);
            } else {
#pragma omp atomic write
                err = -3;
            }
        }
        free(L);
    }
    if (err) return err;
    if (nf == 2) {
        if (S->bc[0] == ORC_BC_MAXWELL) {
            int k = ng; /* cell 1 */
            orc_flux_boundary_maxwell_2f1v(S->fw, S->ff, S->ff + nv, S->bc_prim[0], S->f + (size_t)k * NF,
                                           S->f + (size_t)k * NF + nv, S->u, S->weights, nv, S->K, dt, 1);
        }
        if (S->bc[1] == ORC_BC_MAXWELL) {
            int k = nx - 1 + ng; /* cell nx */
            orc_flux_boundary_maxwell_2f1v(S->fw + (size_t)nx * nw, S->ff + (size_t)nx * NF, S->ff + (size_t)nx * NF + nv,
                                           S->bc_prim[1], S->f + (size_t)k * NF, S->f + (size_t)k * NF + nv,
                                           S->u, S->weights, nv, S->K, dt, -1);
        }
    }
    return 0;
}

static int step_cell_1d(orc_solver1d *S, int i, double dt, double *RES, double *AVG, int coll) {
    const int ng = S->ng, nw = nw_1d(S), nv = S->nv, NF = S->nf * nv;
    int k = i - 1 + ng;
    double *w = S->w + (size_t)k * nw, *prim = S->prim + (size_t)k * nw, *f = S->f + (size_t)k * NF;
    const double *fwL = S->fw + (size_t)(i - 1) * nw, *fwR = S->fw + (size_t)i * nw;
    const double *ffL = S->ff + (size_t)(i - 1) * NF, *ffR = S->ff + (size_t)i * NF;
    if (S->vdim == 1 && S->nf == 1) {
        orc_step_1f1v(w, prim, f, fwL, ffL, fwR, ffR, S->u, S->weights, nv, S->gam, S->mu, S->om, S->Pr, S->dx[k], dt, RES, AVG, coll);
        return 0;
    }
    if (S->vdim == 1 && S->nf == 2) {
        orc_step_2f1v(w, prim, f, f + nv, fwL, ffL, ffL + nv, fwR, ffR, ffR + nv, S->u, S->weights, nv, S->K, S->gam, S->mu, S->om, S->Pr, S->dx[k], dt, RES, AVG, coll);
        return 0;
    }
    if (coll == ORC_FSM)
        return orc_step_1f3v_fsm(w, prim, f, fwL, ffL, fwR, ffR, S->gam, S->fsm_kn, S->fsm_nm, S->fsm_phi, S->fsm_psi,
                                 S->fsm_phipsi, S->fsm_nu, S->fsm_nv, S->fsm_nw, S->dx[k], dt, RES, AVG);
    return orc_step_1f3v(w, prim, f, fwL, ffL, fwR, ffR, S->u, S->v, S->wv, S->weights, nv, S->gam, S->mu, S->om, S->Pr, S->dx[k], dt, RES, AVG, coll);
}

/* update!   src/Solver/solver_update.jl:61-93, then update_boundary! update_boundary.jl:6-61: the
 * boundary cells are stepped unless the side is :fix, and with the default collision (BGK) because
 * update_boundary! does not forward `coll` (update_boundary.jl:33,38). */
int orc_update_1d(orc_solver1d *S, double dt, double *res) {
    const int nx = S->nx, nw = nw_1d(S);
    double sumRes[5] = {0}, sumAvg[5] = {0};
    int err = 0;
#pragma omp parallel
    {
        double r[5] = {0}, a[5] = {0};
#pragma omp for schedule(static)
        for (int i = 2; i <= nx - 1; i++) {
            int e = step_cell_1d(S, i, dt, r, a, S->coll);
            if (e) {
#pragma omp atomic write
                err = e;
            }
        }
#pragma omp critical
        for (int m = 0; m < nw; m++) {
            sumRes[m] += r[m];
            sumAvg[m] += a[m];
        }
    }
    if (err) return err;
    double rL[5] = {0}, aL[5] = {0}, rR[5] = {0}, aR[5] = {0};
    if (S->bc[0] != ORC_BC_FIX) err |= step_cell_1d(S, 1, dt, rL, aL, ORC_BGK);
    if (S->bc[1] != ORC_BC_FIX) err |= step_cell_1d(S, nx, dt, rR, aR, ORC_BGK);
    if (res)
        for (int m = 0; m < nw; m++)
            res[m] = sqrt(sumRes[m] * nx) / (sumAvg[m] + 1.e-7) +
                     sqrt((rL[m] + rR[m]) * 2) / (aL[m] + aR[m] + 1.e-7);
    return err;
}

/* the solve! loop body (solver_main.jl:32-51) repeated nsteps times with a fixed dt */
int orc_run_1d(orc_solver1d *S, double dt, int nsteps, double *res) {
    for (int it = 0; it < nsteps; it++) {
        int e = orc_reconstruct_1d(S);
        if (!e) e = orc_evolve_1d(S, dt);
        if (!e) e = orc_update_1d(S, dt, res);
        if (e) return e;
    }
    return 0;
}

/* ------------------------------------------------------------------------------------------ */
/* 2D whole solves (2F2V)                                                                      */
/* ------------------------------------------------------------------------------------------ */

#define C2(i, j) ((size_t)((i) - 1) + (size_t)S->nx * ((j) - 1))
#define XF(i, j) ((size_t)((i) - 1) + (size_t)(S->nx + 1) * ((j) - 1))
#define YF(i, j) ((size_t)((i) - 1) + (size_t)S->nx * ((j) - 1))

/* slopes: sw[cell][dir][4], sf[cell][dir][2][nv] */
static double *SW2(orc_solver2d *S, size_t c, int dir) { return S->sw + c * 8 + (size_t)dir * 4; }
static int nf_2d(const orc_solver2d *S) { return S->nf == 1 ? 1 : 2; }
static double *SF2(orc_solver2d *S, size_t c, int dir) {
    const size_t NF = (size_t)nf_2d(S) * S->nv;
    return S->sf + c * 2 * NF + (size_t)dir * NF;
}

/* reconstruct!   src/Solver/solver_reconstruction.jl:475-620 (reconstruct2! reconstruction.jl:48-50
 * on the boundary rows, reconstruct3! inside) */
int orc_reconstruct_2d(orc_solver2d *S) {
    if (S->interp_order == 1) return 0;
    const int nx = S->nx, ny = S->ny, nv = S->nv, NF = nf_2d(S) * nv;
    /* boundary x */
#pragma omp parallel for schedule(static)
    for (int j = 1; j <= ny; j++) {
        size_t c1 = C2(1, j), c2 = C2(2, j), cm = C2(nx - 1, j), cn = C2(nx, j);
        double d1 = 0.5 * (S->dx[c1] + S->dx[c2]), dn = 0.5 * (S->dx[cm] + S->dx[cn]);
        for (int m = 0; m < 4; m++) {
            SW2(S, c1, 0)[m] = (S->w[c2 * 4 + m] - S->w[c1 * 4 + m]) / d1;
            SW2(S, cn, 0)[m] = (S->w[cn * 4 + m] - S->w[cm * 4 + m]) / dn;
        }
        for (int q = 0; q < NF; q++) {
            SF2(S, c1, 0)[q] = (S->f[c2 * NF + q] - S->f[c1 * NF + q]) / d1;
            SF2(S, cn, 0)[q] = (S->f[cn * NF + q] - S->f[cm * NF + q]) / dn;
        }
    }
    /* boundary y */
#pragma omp parallel for schedule(static)
    for (int i = 1; i <= nx; i++) {
        size_t c1 = C2(i, 1), c2 = C2(i, 2), cm = C2(i, ny - 1), cn = C2(i, ny);
        double d1 = 0.5 * (S->dy[c1] + S->dy[c2]), dn = 0.5 * (S->dy[cm] + S->dy[cn]);
        for (int m = 0; m < 4; m++) {
            SW2(S, c1, 1)[m] = (S->w[c2 * 4 + m] - S->w[c1 * 4 + m]) / d1;
            SW2(S, cn, 1)[m] = (S->w[cn * 4 + m] - S->w[cm * 4 + m]) / dn;
        }
        for (int q = 0; q < NF; q++) {
            SF2(S, c1, 1)[q] = (S->f[c2 * NF + q] - S->f[c1 * NF + q]) / d1;
            SF2(S, cn, 1)[q] = (S->f[cn * NF + q] - S->f[cm * NF + q]) / dn;
        }
    }
    /* inner x */
#pragma omp parallel for schedule(static)
    for (int j = 1; j <= ny; j++)
        for (int i = 2; i <= nx - 1; i++) {
            size_t a = C2(i - 1, j), c = C2(i, j), e = C2(i + 1, j);
            double dL = 0.5 * (S->dx[a] + S->dx[c]), dR = 0.5 * (S->dx[c] + S->dx[e]);
            for (int m = 0; m < 4; m++)
                SW2(S, c, 0)[m] = orc_limiter(S->limiter, (S->w[c * 4 + m] - S->w[a * 4 + m]) / dL,
                                              (S->w[e * 4 + m] - S->w[c * 4 + m]) / dR);
            for (int q = 0; q < NF; q++)
                SF2(S, c, 0)[q] = orc_limiter(S->limiter, (S->f[c * NF + q] - S->f[a * NF + q]) / dL,
                                              (S->f[e * NF + q] - S->f[c * NF + q]) / dR);
        }
    /* inner y */
#pragma omp parallel for schedule(static)
    for (int j = 2; j <= ny - 1; j++)
        for (int i = 1; i <= nx; i++) {
            size_t a = C2(i, j - 1), c = C2(i, j), e = C2(i, j + 1);
            double dL = 0.5 * (S->dy[a] + S->dy[c]), dR = 0.5 * (S->dy[c] + S->dy[e]);
            for (int m = 0; m < 4; m++)
                SW2(S, c, 1)[m] = orc_limiter(S->limiter, (S->w[c * 4 + m] - S->w[a * 4 + m]) / dL,
                                              (S->w[e * 4 + m] - S->w[c * 4 + m]) / dR);
            for (int q = 0; q < NF; q++)
                SF2(S, c, 1)[q] = orc_limiter(S->limiter, (S->f[c * NF + q] - S->f[a * NF + q]) / dL,
                                              (S->f[e * NF + q] - S->f[c * NF + q]) / dR);
        }
    return 0;
}

/* one interior face: the 2D cell-level wrappers flux_kfvs.jl:98-135, flux_kcu.jl:149-191 and the
 * UGKS harness defined the same way as the KCU wrapper */
static void face_2d(orc_solver2d *S, size_t cL, size_t cR, int dir, double dxL, double dxR,
                    const double *n, double len, double dt, double *fw, double *fh, double *fb,
                    double *buf) {
    const int nv = S->nv, NF = nf_2d(S) * nv;
    if (nf_2d(S) == 1) {
        /* 2D1F2V: the 1F cell-level wrappers flux_kfvs.jl:65-93, flux_kcu.jl:104-147 */
        double *fL = buf, *fR = buf + nv, *vn = buf + 2 * nv, *vt = buf + 3 * nv;
        const double *sL = SF2(S, cL, dir), *sR = SF2(S, cR, dir);
        const double *f0 = S->f + cL * NF, *f1 = S->f + cR * NF;
        for (int q = 0; q < nv; q++) {
            fL[q] = f0[q] + sL[q] * dxL;
            fR[q] = f1[q] - sR[q] * dxR;
            vn[q] = S->u[q] * n[0] + S->v[q] * n[1];
            vt[q] = S->v[q] * n[0] - S->u[q] * n[1];
        }
        double fwl[4];
        if (S->flux == ORC_KFVS) {
            orc_flux_kfvs_1f2v(fwl, fh, fL, fR, vn, vt, S->weights, nv, dt, len, sL, sR);
        } else {
            double wl[4], wr[4], WL[4], WR[4];
            const double *swL = SW2(S, cL, dir), *swR = SW2(S, cR, dir);
            for (int m = 0; m < 4; m++) {
                wl[m] = S->w[cL * 4 + m] + dxL * swL[m];
                wr[m] = S->w[cR * 4 + m] - dxR * swR[m];
            }
            orc_local_frame(wl, 4, n[0], n[1], WL);
            orc_local_frame(wr, 4, n[0], n[1], WR);
            orc_flux_kcu_1f2v(fwl, fh, WL, fL, WR, fR, vn, vt, S->weights, nv, S->K, S->gam, S->mu, S->om, S->Pr, dt, len);
        }
        orc_global_frame(fwl, 4, n[0], n[1], fw);
        return;
    }
    double *hL = buf, *bL = buf + nv, *hR = buf + 2 * nv, *bR = buf + 3 * nv, *vn = buf + 4 * nv, *vt = buf + 5 * nv;
    const double *sL = SF2(S, cL, dir), *sR = SF2(S, cR, dir);
    const double *fL = S->f + cL * NF, *fR = S->f + cR * NF;
    for (int q = 0; q < nv; q++) {
        hL[q] = fL[q] + sL[q] * dxL;
        bL[q] = fL[nv + q] + sL[nv + q] * dxL;
        hR[q] = fR[q] - sR[q] * dxR;
        bR[q] = fR[nv + q] - sR[nv + q] * dxR;
        vn[q] = S->u[q] * n[0] + S->v[q] * n[1];
        vt[q] = S->v[q] * n[0] - S->u[q] * n[1];
    }
    double fwl[4];
    if (S->flux == ORC_KFVS) {
        orc_flux_kfvs_2f2v(fwl, fh, fb, hL, bL, hR, bR, vn, vt, S->weights, nv, dt, len, sL, sL + nv, sR, sR + nv);
    } else {
        double wl[4], wr[4], WL[4], WR[4];
        const double *swL = SW2(S, cL, dir), *swR = SW2(S, cR, dir);
        for (int m = 0; m < 4; m++) {
            wl[m] = S->w[cL * 4 + m] + dxL * swL[m];
            wr[m] = S->w[cR * 4 + m] - dxR * swR[m];
        }
        orc_local_frame(wl, 4, n[0], n[1], WL);
        orc_local_frame(wr, 4, n[0], n[1], WR);
        if (S->flux == ORC_KCU)
            orc_flux_kcu_2f2v(fwl, fh, fb, WL, hL, bL, WR, hR, bR, vn, vt, S->weights, nv, S->K, S->gam, S->mu, S->om, S->Pr, dt, len);
        else
            orc_flux_ugks_2f2v(fwl, fh, fb, WL, hL, bL, WR, hR, bR, vn, vt, S->weights, nv, S->K, S->gam, S->mu, S->om, S->Pr, dt, dxL, dxR, len, sL, sL + nv, sR, sR + nv);
    }
    orc_global_frame(fwl, 4, n[0], n[1], fw);
}

/* a wall face: evolve_boundary.jl:71-175 -> flux_boundary.jl:190-220 */
static void wall_2d(orc_solver2d *S, size_t c, const double *n, double len, const double *bcprim,
                    int rot, double dt, double *fw, double *fh, double *fb, double *buf) {
    const int nv = S->nv, NF = nf_2d(S) * nv;
    double *vn = buf, *vt = buf + nv;
    for (int q = 0; q < nv; q++) {
        vn[q] = S->u[q] * n[0] + S->v[q] * n[1];
        vt[q] = S->v[q] * n[0] - S->u[q] * n[1];
    }
    double bcl[4], fwl[4];
    orc_local_frame(bcprim, 4, n[0], n[1], bcl);
    if (nf_2d(S) == 1) /* the inline 2D1F walls of solver_evolution.jl:500-604 */
        orc_flux_boundary_maxwell_1f2v(fwl, fh, bcl, S->f + c * NF, vn, vt, S->weights, nv, dt, len, rot);
    else
        orc_flux_boundary_maxwell_2f2v(fwl, fh, fb, bcl, S->f + c * NF, S->f + c * NF + nv, vn, vt, S->weights, nv, S->K, dt, len, rot);
    orc_global_frame(fwl, 4, n[0], n[1], fw);
}

/* evolve!   src/Solver/solver_evolution.jl:608-679 (no ghost layers) + evolve_boundary.jl:52-178 */
int orc_evolve_2d(orc_solver2d *S, double dt) {
    const int nx = S->nx, ny = S->ny, nv = S->nv, NF = nf_2d(S) * nv;
    for (int s = 0; s < 4; s++)
        if (S->bc[s] != ORC_BC_MAXWELL && S->bc[s] != ORC_BC_FIX) return -4;
    if (nf_2d(S) == 1 && S->flux == ORC_UGKS) return -2; /* KitBase has no 2D1F2V UGKS kernel */
#pragma omp parallel
    {
        double *buf = (double *)malloc(sizeof(double) * 6 * nv);
#pragma omp for schedule(static)
        for (int j = 1; j <= ny; j++)
            for (int i = 2; i <= nx; i++) {
                size_t fi = XF(i, j), cL = C2(i - 1, j), cR = C2(i, j);
                face_2d(S, cL, cR, 0, 0.5 * S->dx[cL], 0.5 * S->dx[cR], S->xn + 2 * fi, S->xlen[fi], dt,
                        S->xfw + 4 * fi, S->xff + fi * NF, S->xff + fi * NF + nv, buf);
            }
#pragma omp for schedule(static)
        for (int j = 2; j <= ny; j++)
            for (int i = 1; i <= nx; i++) {
                size_t fi = YF(i, j), cL = C2(i, j - 1), cR = C2(i, j);
                face_2d(S, cL, cR, 1, 0.5 * S->dy[cL], 0.5 * S->dy[cR], S->yn + 2 * fi, S->ylen[fi], dt,
                        S->yfw + 4 * fi, S->yff + fi * NF, S->yff + fi * NF + nv, buf);
            }
        if (S->bc[0] == ORC_BC_MAXWELL) {
#pragma omp for schedule(static)
            for (int j = 1; j <= ny; j++) {
                size_t fi = XF(1, j);
                wall_2d(S, C2(1, j), S->xn + 2 * fi, S->xlen[fi], S->xwall + 4 * fi, 1, dt, S->xfw + 4 * fi,
                        S->xff + fi * NF, S->xff + fi * NF + nv, buf);
            }
        }
        if (S->bc[1] == ORC_BC_MAXWELL) {
#pragma omp for schedule(static)
            for (int j = 1; j <= ny; j++) {
                size_t fi = XF(nx + 1, j);
                wall_2d(S, C2(nx, j), S->xn + 2 * fi, S->xlen[fi], S->xwall + 4 * fi, -1, dt, S->xfw + 4 * fi,
                        S->xff + fi * NF, S->xff + fi * NF + nv, buf);
            }
        }
        if (S->bc[2] == ORC_BC_MAXWELL) {
#pragma omp for schedule(static)
            for (int i = 1; i <= nx; i++) {
                size_t fi = YF(i, 1);
                wall_2d(S, C2(i, 1), S->yn + 2 * fi, S->ylen[fi], S->ywall + 4 * fi, 1, dt, S->yfw + 4 * fi,
                        S->yff + fi * NF, S->yff + fi * NF + nv, buf);
            }
        }
        if (S->bc[3] == ORC_BC_MAXWELL) {
#pragma omp for schedule(static)
            for (int i = 1; i <= nx; i++) {
                size_t fi = YF(i, ny + 1);
                wall_2d(S, C2(i, ny), S->yn + 2 * fi, S->ylen[fi], S->ywall + 4 * fi, -1, dt, S->yfw + 4 * fi,
                        S->yff + fi * NF, S->yff + fi * NF + nv, buf);
            }
        }
        free(buf);
    }
    return 0;
}

static int step_cell_2d(orc_solver2d *S, int i, int j, double dt, double *RES, double *AVG) {
    const int nv = S->nv, NF = nf_2d(S) * nv;
    size_t c = C2(i, j), L = XF(i, j), R = XF(i + 1, j), D = YF(i, j), U = YF(i, j + 1);
    double *f = S->f + c * NF;
    if (nf_2d(S) == 1)
        return orc_step_1f2v(S->w + 4 * c, S->prim + 4 * c, f, S->xfw + 4 * L, S->xff + L * NF, S->xfw + 4 * R,
                             S->xff + R * NF, S->yfw + 4 * D, S->yff + D * NF, S->yfw + 4 * U, S->yff + U * NF, S->u,
                             S->v, S->weights, nv, S->gam, S->mu, S->om, S->Pr, S->dx[c] * S->dy[c], dt, RES, AVG,
                             S->coll);
    orc_step_2f2v(S->w + 4 * c, S->prim + 4 * c, f, f + nv, S->xfw + 4 * L, S->xff + L * NF,
                  S->xff + L * NF + nv, S->xfw + 4 * R, S->xff + R * NF, S->xff + R * NF + nv,
                  S->yfw + 4 * D, S->yff + D * NF, S->yff + D * NF + nv, S->yfw + 4 * U,
                  S->yff + U * NF, S->yff + U * NF + nv, S->u, S->v, S->weights, nv, S->K, S->gam,
                  S->mu, S->om, S->Pr, S->dx[c] * S->dy[c], dt, RES, AVG, S->coll);
    return 0;
}

/* update!   src/Solver/solver_update.jl:193-246, then update_boundary! update_boundary.jl:172-306
 * (the 2D version does forward `coll`) */
int orc_update_2d(orc_solver2d *S, double dt, double *res) {
    const int nx = S->nx, ny = S->ny;
    double sumRes[4] = {0}, sumAvg[4] = {0};
    if (nf_2d(S) == 1 && S->coll == ORC_SHAKHOV) return -1; /* step_1f.jl:194 */
#pragma omp parallel
    {
        double r[4] = {0}, a[4] = {0};
#pragma omp for schedule(static)
        for (int j = 2; j <= ny - 1; j++)
            for (int i = 2; i <= nx - 1; i++) step_cell_2d(S, i, j, dt, r, a);
#pragma omp critical
        for (int m = 0; m < 4; m++) {
            sumRes[m] += r[m];
            sumAvg[m] += a[m];
        }
    }
    double rb[4][4], ab[4][4];
    memset(rb, 0, sizeof rb);
    memset(ab, 0, sizeof ab);
    if (S->bc[0] != ORC_BC_FIX)
        for (int j = 1; j <= ny; j++) step_cell_2d(S, 1, j, dt, rb[0], ab[0]);
    if (S->bc[1] != ORC_BC_FIX)
        for (int j = 1; j <= ny; j++) step_cell_2d(S, nx, j, dt, rb[1], ab[1]);
    if (S->bc[2] != ORC_BC_FIX)
        for (int i = 2; i <= nx - 1; i++) step_cell_2d(S, i, 1, dt, rb[2], ab[2]);
    if (S->bc[3] != ORC_BC_FIX)
        for (int i = 2; i <= nx - 1; i++) step_cell_2d(S, i, ny, dt, rb[3], ab[3]);
    if (res)
        for (int m = 0; m < 4; m++)
            res[m] = sqrt(sumRes[m] * nx * ny) / (sumAvg[m] + 1.e-7) +
                     sqrt((rb[0][m] + rb[1][m] + rb[3][m] + rb[2][m]) * 2) /
                         (ab[0][m] + ab[1][m] + ab[3][m] + ab[2][m] + 1.e-7);
    return 0;
}

int orc_run_2d(orc_solver2d *S, double dt, int nsteps, double *res) {
    for (int it = 0; it < nsteps; it++) {
        int e = orc_reconstruct_2d(S);
        if (!e) e = orc_evolve_2d(S, dt);
        if (!e) e = orc_update_2d(S, dt, res);
        if (e) return e;
    }
    return 0;
}

int orc_num_threads(void) {
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}

void orc_set_num_threads(int n) {
#ifdef _OPENMP
    omp_set_num_threads(n);
#else
    (void)n;
#endif
}

/* ------------------------------------------------------------------------------------------ */
/* Unstructured triangle meshes                                                                 */
/* ------------------------------------------------------------------------------------------ */

/* evolve!   src/Solver/solver_evolution.jl:872-1035 at interpOrder 1: the face values
 * f + sf . (x_face - x_cell) are the cell values, the slopes staying zero (reconstruct! returns at
 * once, solver_reconstruction.jl:809-880).  KitBase's unstructured evolve! calls the 2D Maxwell wall
 * without its `rot` argument (flux_boundary.jl:160, :190 take one), which has no method; the face
 * normal of a boundary face points into the mesh, so the wall is taken with rot = 1 (incoming
 * velocities vn > 0 carry the wall Maxwellian). */
int orc_evolve_us(orc_solverus *S, double dt) {
    const int nv = S->nv, nf = S->nf, RS = nf * nv;
    int bad = 0;
#pragma omp parallel for schedule(dynamic, 16)
    for (int f = 0; f < S->nface; f++) {
        if (S->fkind[f] == 2) continue;
        double *vn = (double *)malloc(2 * nv * sizeof(double)), *vt = vn + nv;
        const double n1 = S->fn[2 * f], n2 = S->fn[2 * f + 1], len = S->flen[f];
        for (int i = 0; i < nv; i++) {
            vn[i] = S->u[i] * n1 + S->v[i] * n2;
            vt[i] = S->v[i] * n1 - S->u[i] * n2;
        }
        double fw[4];
        double *ffo = S->ff + (size_t)f * RS;
        if (S->fkind[f] == 0) {
            const double *a = S->f + (size_t)S->fc[2 * f] * RS, *b = S->f + (size_t)S->fc[2 * f + 1] * RS;
            if (nf == 1) orc_flux_kfvs_1f2v(fw, ffo, a, b, vn, vt, S->weights, nv, dt, len, NULL, NULL);
            else orc_flux_kfvs_2f2v(fw, ffo, ffo + nv, a, a + nv, b, b + nv, vn, vt, S->weights, nv, dt, len, NULL, NULL, NULL, NULL);
        } else {
            const int c = S->fc[2 * f] >= 0 ? S->fc[2 * f] : S->fc[2 * f + 1];
            if (c < 0) {
                bad = 1;
                free(vn);
                continue;
            }
            double bc[4];
            orc_local_frame(S->fwall + 4 * (size_t)f, 4, n1, n2, bc);
            const double *h = S->f + (size_t)c * RS;
            if (nf == 1) orc_flux_boundary_maxwell_1f2v(fw, ffo, bc, h, vn, vt, S->weights, nv, dt, len, 1);
            else orc_flux_boundary_maxwell_2f2v(fw, ffo, ffo + nv, bc, h, h + nv, vn, vt, S->weights, nv, S->K, dt, len, 1);
        }
        orc_global_frame(fw, 4, n1, n2, S->fw + 4 * (size_t)f);
        free(vn);
    }
    return bad ? -1 : 0;
}

/* triangle step!   src/Solver/step_1f.jl:228-285 (BGK: its Shakhov branch reads h and b, which the
 * one-distribution step does not have), src/Solver/step_2f.jl:256-326 */
static void orc_step_tri(const orc_solverus *S, int c, double dt, double *RES, double *AVG) {
    const int nv = S->nv, nf = S->nf, RS = nf * nv;
    double *w = S->w + 4 * (size_t)c, *prim = S->prim + 4 * (size_t)c, *h = S->f + (size_t)c * RS;
    double *b = nf == 2 ? h + nv : NULL;
    const double *u = S->u, *v = S->v, *wt = S->weights, K = S->K, ds = S->area[c];
    double w_old[4] = {w[0], w[1], w[2], w[3]};
    double *SH = (double *)calloc(2 * nv, sizeof(double)), *SB = SH + nv;
    if (nf == 2 && S->coll == ORC_SHAKHOV) {
        double qa = 0, qb = 0, qc = 0, qd = 0;
        for (int i = 0; i < nv; i++) {
            const double cu = u[i] - prim[1], cv = v[i] - prim[2], c2 = cu * cu + cv * cv;
            qa += wt[i] * cu * c2 * h[i];
            qb += wt[i] * cu * b[i];
            qc += wt[i] * cv * c2 * h[i];
            qd += wt[i] * cv * b[i];
        }
        const double q0 = 0.5 * (qa + qb), q1 = 0.5 * (qc + qd);
        for (int i = 0; i < nv; i++) {
            const double cu = u[i] - prim[1], cv = v[i] - prim[2], c2 = cu * cu + cv * cv;
            const double MH = prim[0] * (prim[3] / PI_) * exp(-prim[3] * c2), MB = MH * K / (2.0 * prim[3]);
            const double pre = 0.8 * (1.0 - S->Pr) * (prim[3] * prim[3]) / prim[0] * (cu * q0 + cv * q1);
            SH[i] = pre * (2.0 * prim[3] * c2 + K - 5.0) * MH;
            SB[i] = pre * (2.0 * prim[3] * c2 + K - 3.0) * MB;
        }
    }
    const int *cf = S->cf + 3 * (size_t)c;
    const double *dr = S->dirc + 3 * (size_t)c;
    for (int m = 0; m < 4; m++)
        w[m] -= (S->fw[4 * (size_t)cf[0] + m] * dr[0] + S->fw[4 * (size_t)cf[1] + m] * dr[1] + S->fw[4 * (size_t)cf[2] + m] * dr[2]) / ds;
    orc_conserve_prim(w, 4, S->gam, prim);
    for (int m = 0; m < 4; m++) {
        RES[m] += (w[m] - w_old[m]) * (w[m] - w_old[m]);
        AVG[m] += fabs(w[m]);
    }
    const double tau = orc_vhs_collision_time(prim[0], prim[3], S->mu, S->om);
    const double *F1 = S->ff + (size_t)cf[0] * RS, *F2 = S->ff + (size_t)cf[1] * RS, *F3 = S->ff + (size_t)cf[2] * RS;
    for (int i = 0; i < nv; i++) {
        const double cu = u[i] - prim[1], cv = v[i] - prim[2];
        double MH = prim[0] * (prim[3] / PI_) * exp(-prim[3] * (cu * cu + cv * cv));
        double MB = MH * K / (2.0 * prim[3]);
        MH += SH[i];
        MB += SB[i];
        h[i] = (h[i] - (F1[i] * dr[0] + F2[i] * dr[1] + F3[i] * dr[2]) / ds + dt / tau * MH) / (1.0 + dt / tau);
        if (nf == 2)
            b[i] = (b[i] - (F1[nv + i] * dr[0] + F2[nv + i] * dr[1] + F3[nv + i] * dr[2]) / ds + dt / tau * MB) / (1.0 + dt / tau);
    }
    free(SH);
}

/* update!   src/Solver/solver_update.jl:359-471, then update_boundary.jl:520-590 (cellType 3) */
int orc_update_us(orc_solverus *S, double dt, double *res) {
    if (S->nf == 1 && S->coll == ORC_SHAKHOV) return -1;
    double RES[4] = {0, 0, 0, 0}, AVG[4] = {0, 0, 0, 0};
    for (int c = 0; c < S->ncell; c++)
        if (S->ctype[c] == 0 || S->ctype[c] == 2) orc_step_tri(S, c, dt, RES, AVG);
    if (res)
        for (int m = 0; m < 4; m++) res[m] = sqrt(RES[m] * S->ncell) / (AVG[m] + 1.e-7);
    const int RS = S->nf * S->nv;
    for (int c = 0; c < S->ncell; c++) {
        if (S->ctype[c] != 3) continue;
        int id[2], k = 0;
        for (int j = 0; j < 3 && k < 2; j++)
            if (S->nb[3 * (size_t)c + j] != -1) id[k++] = S->nb[3 * (size_t)c + j];
        if (k < 2) return -1;
        for (int m = 0; m < 4; m++) S->w[4 * (size_t)c + m] = 0.5 * (S->w[4 * (size_t)id[0] + m] + S->w[4 * (size_t)id[1] + m]);
        for (int i = 0; i < RS; i++)
            S->f[(size_t)c * RS + i] = 0.5 * (S->f[(size_t)id[0] * RS + i] + S->f[(size_t)id[1] * RS + i]);
        orc_conserve_prim(S->w + 4 * (size_t)c, 4, S->gam, S->prim + 4 * (size_t)c);
    }
    return 0;
}

int orc_run_us(orc_solverus *S, double dt, int nsteps, double *res) {
    for (int t = 0; t < nsteps; t++) {
        int rc = orc_evolve_us(S, dt);
        if (!rc) rc = orc_update_us(S, dt, res);
        if (rc) return rc;
    }
    return 0;
}

/* ------------------------------------------------------------------------------------------ */
/* Fast spectral method (1D1F3V collision :fsm)                                                 */
/* ------------------------------------------------------------------------------------------ */

/* lgwt   src/Math/math_general.jl:58-98: N Gauss-Legendre nodes and weights on [a, b] by the damped
 * Newton iteration of KitBase (its derivative carries (N+1) for N, undone in the weights); nodes in
 * descending order. */
void orc_lgwt(int N, double a, double b, double *x, double *w) {
    const int N1 = N, N2 = N + 1;
    double *y = (double *)malloc(sizeof(double) * N1 * 3), *y0 = y + N1, *Lp = y0 + N1;
    double *L = (double *)malloc(sizeof(double) * N1 * N2);
    for (int i = 1; i <= N1; i++) {
        y[i - 1] = cos((2.0 * (i - 1.0) + 1.0) * 4.0 * atan(1.0) / (2.0 * (N - 1.0) + 2.0)) +
                   0.27 / N1 * sin(4.0 * atan(1.0) * (-1.0 + i * 2.0 / (N1 - 1.0)) * (N - 1.0) / N2);
        y0[i - 1] = 2.0;
    }
    for (;;) {
        double dmax = 0.0;
        for (int i = 0; i < N1; i++) dmax = fmax(dmax, fabs(y[i] - y0[i]));
        if (!(dmax > 0.0000000000001)) break;
        for (int i = 0; i < N1; i++) {
            double *Li = L + (size_t)i * N2;  /* L[i, 1..N2] */
            Li[0] = 1.0;
            Li[1] = y[i];
            for (int k = 2; k <= N1; k++) Li[k] = ((2.0 * k - 1.0) * y[i] * Li[k - 1] - (k - 1) * Li[k - 2]) / k;
            Lp[i] = N2 * (Li[N1 - 1] - y[i] * Li[N2 - 1]) / (1.0 - y[i] * y[i]);
            y0[i] = y[i];
            y[i] = y0[i] - Li[N2 - 1] / Lp[i];
        }
    }
    for (int i = 0; i < N1; i++) {
        x[i] = (a * (1.0 - y[i]) + b * (1.0 + y[i])) / 2.0;
        w[i] = (double)N2 * N2 * (b - a) / ((1.0 - y[i] * y[i]) * Lp[i] * Lp[i]) / ((double)N1 * N1);
    }
    free(L);
    free(y);
}

/* hs_boltz_kn   src/Theory/theory_fsm.jl:16-17 (Gamma(2) = 1) */
double orc_hs_boltz_kn(double mu_ref, double alpha) {
    return 64.0 * pow(sqrt(2.0), alpha) / 5.0 * tgamma((alpha + 3.0) / 2.0) * tgamma(2.0) * sqrt(PI_) * mu_ref;
}

/* kernel_mode   src/Theory/theory_fsm.jl:24-90 (the variant fsm_kernel calls, with the grid
 * spacings): phi, psi [M(M-1)][nw][nv][nu] and phipsi [nw][nv][nu], u fastest (Julia's column-major
 * phi[i, j, k, idx]).  besselj(1, .) is libm's j1. */
void orc_kernel_mode(int M, double umax, double vmax, double wmax, double du, double dv, double dw, int nu, int nv,
                     int nw, double alpha, int quad_num, double *phi, double *psi, double *phipsi) {
    const double supp = sqrt(2.0) * 2.0 * fmax(fmax(umax, vmax), wmax) / (3.0 + sqrt(2.0));
    double *fx = (double *)malloc(sizeof(double) * (nu + nv + nw)), *fy = fx + nu, *fz = fy + nv;
    /* range(-pi/d, (n/2 - 1) 2 pi / n / d; length = n): start + k step */
    for (int i = 0; i < nu; i++) fx[i] = -PI_ / du + i * (((nu / 2 - 1) * 2.0 * PI_ / nu / du - (-PI_ / du)) / (nu - 1));
    for (int j = 0; j < nv; j++) fy[j] = -PI_ / dv + j * (((nv / 2 - 1) * 2.0 * PI_ / nv / dv - (-PI_ / dv)) / (nv - 1));
    for (int k = 0; k < nw; k++) fz[k] = -PI_ / dw + k * (((nw / 2 - 1) * 2.0 * PI_ / nw / dw - (-PI_ / dw)) / (nw - 1));
    double *ab = (double *)malloc(sizeof(double) * quad_num * 2), *gw = ab + quad_num;
    orc_lgwt(quad_num, 0.0, supp, ab, gw);
    const size_t n3 = (size_t)nu * nv * nw;
    for (size_t q = 0; q < n3; q++) phipsi[q] = 0.0;
    for (int loop = 1; loop <= M - 1; loop++) {
        const double theta = PI_ / M * loop;
        for (int loop2 = 1; loop2 <= M; loop2++) {
            const double theta2 = PI_ / M * loop2;
            const int idx = (loop - 1) * M + loop2 - 1;
#pragma omp parallel for schedule(static)
            for (int k = 0; k < nw; k++)
                for (int j = 0; j < nv; j++)
                    for (int i = 0; i < nu; i++) {
                        const size_t q = (size_t)i + (size_t)nu * (j + (size_t)nv * k);
                        double s = fx[i] * sin(theta) * cos(theta2) + fy[j] * sin(theta) * sin(theta2) + fz[k] * cos(theta);
                        double int_temp = 0.0;
                        for (int id = 0; id < quad_num; id++) int_temp += 2.0 * gw[id] * cos(s * ab[id]) * pow(ab[id], alpha);
                        const double ph = int_temp * sin(theta);
                        s = fx[i] * fx[i] + fy[j] * fy[j] + fz[k] * fz[k] - s * s;
                        double ps;
                        if (s <= 0.0) {
                            ps = PI_ * supp * supp;
                        } else {
                            s = sqrt(s);
                            ps = 2.0 * PI_ * supp * j1(supp * s) / s;
                        }
                        phi[(size_t)idx * n3 + q] = ph;
                        psi[(size_t)idx * n3 + q] = ps;
                        phipsi[q] += ph * ps;
                    }
        }
    }
    free(ab);
    free(fx);
}

/* one axis of an unnormalised 3D DFT over [nw][nv][nu] (u fastest): sign -1 forward (fft), +1
 * backward (bfft), X_k = sum_n x_n exp(sign 2 pi i k n / N) */
static void dft_axis(double *re, double *im, int nu, int nv, int nw, int axis, int sign) {
    const int n = axis == 0 ? nu : (axis == 1 ? nv : nw);
    const size_t stride = axis == 0 ? 1 : (axis == 1 ? (size_t)nu : (size_t)nu * nv);
    const size_t lines = (size_t)nu * nv * nw / n;
    double *cs = (double *)malloc(sizeof(double) * n * 2), *sn = cs + n;
    for (int t = 0; t < n; t++) {
        cs[t] = cos(2.0 * PI_ * t / n);
        sn[t] = sign * sin(2.0 * PI_ * t / n);
    }
#pragma omp parallel
    {
        double *xr = (double *)malloc(sizeof(double) * n * 2), *xi = xr + n;
#pragma omp for schedule(static)
        for (size_t l = 0; l < lines; l++) {
            /* the line's base: every index with the axis coordinate zero */
            size_t base;
            if (axis == 0) base = l * nu;
            else if (axis == 1) base = (l % nu) + (l / nu) * (size_t)nu * nv;
            else base = l;
            for (int t = 0; t < n; t++) {
                xr[t] = re[base + t * stride];
                xi[t] = im[base + t * stride];
            }
            for (int kk = 0; kk < n; kk++) {
                double ar = 0.0, ai = 0.0;
                for (int t = 0; t < n; t++) {
                    const int e = (int)(((long)kk * t) % n);
                    ar += xr[t] * cs[e] - xi[t] * sn[e];
                    ai += xr[t] * sn[e] + xi[t] * cs[e];
                }
                re[base + kk * stride] = ar;
                im[base + kk * stride] = ai;
            }
        }
        free(xr);
    }
    free(cs);
}

void orc_dft3(double *re, double *im, int nu, int nv, int nw, int sign) {
    for (int a = 0; a < 3; a++) dft_axis(re, im, nu, nv, nw, a, sign);
}

/* fftshift of a [nw][nv][nu] array (each dimension rolled by n / 2) */
void orc_fftshift3(const double *in, double *out, int nu, int nv, int nw) {
    for (int k = 0; k < nw; k++)
        for (int j = 0; j < nv; j++)
            for (int i = 0; i < nu; i++) {
                const int i2 = (i + nu / 2) % nu, j2 = (j + nv / 2) % nv, k2 = (k + nw / 2) % nw;
                out[(size_t)i2 + (size_t)nu * (j2 + (size_t)nv * k2)] = in[(size_t)i + (size_t)nu * (j + (size_t)nv * k)];
            }
}

/* boltzmann_fft!   src/Theory/theory_fsm.jl:271-314: f_spec = fftshift(bfft(f) / N^3); gain
 * sum_i fft(f_spec phi_i) fft(f_spec psi_i), loss fft(f_spec phipsi) fft(f_spec);
 * Q = 4 pi^2 / Kn / M^2 Re(gain - loss) */
void orc_boltzmann_fft(double *Q, const double *f, double Kn, int M, const double *phi, const double *psi,
                       const double *phipsi, int nu, int nv, int nw) {
    const size_t n3 = (size_t)nu * nv * nw;
    double *sr = (double *)calloc(n3 * 8, sizeof(double));
    double *si = sr + n3, *ar = si + n3, *ai = ar + n3, *br = ai + n3, *bi = br + n3, *tr = bi + n3, *ti = tr + n3;
    memcpy(ar, f, n3 * sizeof(double));
    orc_dft3(ar, ai, nu, nv, nw, +1);
    const double inv = 1.0 / (double)n3;
    for (size_t q = 0; q < n3; q++) {
        ar[q] *= inv;
        ai[q] *= inv;
    }
    orc_fftshift3(ar, sr, nu, nv, nw);
    orc_fftshift3(ai, si, nu, nv, nw);
    for (size_t q = 0; q < n3; q++) tr[q] = ti[q] = 0.0;
    for (int d = 0; d < M * (M - 1); d++) {
        const double *ph = phi + (size_t)d * n3, *ps = psi + (size_t)d * n3;
        for (size_t q = 0; q < n3; q++) {
            ar[q] = sr[q] * ph[q];
            ai[q] = si[q] * ph[q];
            br[q] = sr[q] * ps[q];
            bi[q] = si[q] * ps[q];
        }
        orc_dft3(ar, ai, nu, nv, nw, -1);
        orc_dft3(br, bi, nu, nv, nw, -1);
        for (size_t q = 0; q < n3; q++) {
            tr[q] += ar[q] * br[q] - ai[q] * bi[q];
            ti[q] += ar[q] * bi[q] + ai[q] * br[q];
        }
    }
    for (size_t q = 0; q < n3; q++) {
        ar[q] = sr[q] * phipsi[q];
        ai[q] = si[q] * phipsi[q];
        br[q] = sr[q];
        bi[q] = si[q];
    }
    orc_dft3(ar, ai, nu, nv, nw, -1);
    orc_dft3(br, bi, nu, nv, nw, -1);
    for (size_t q = 0; q < n3; q++) tr[q] -= ar[q] * br[q] - ai[q] * bi[q];
    const double c = 4.0 * PI_ * PI_ / Kn / ((double)M * M);
    for (size_t q = 0; q < n3; q++) Q[q] = c * tr[q];
    free(sr);
}

/* 1D1F3V fast spectral step!   src/Solver/step_1f.jl:122-157: w += (fwL - fwR) / dx, prim, the
 * residual sums, then f += (ffL - ffR) / dx + dt Q(f) with Q from f before the update */
int orc_step_1f3v_fsm(double *w, double *prim, double *f, const double *fwL, const double *ffL, const double *fwR,
                      const double *ffR, double gam, double Kn, int M, const double *phi, const double *psi,
                      const double *phipsi, int nu, int nv, int nw, double dx, double dt, double *RES, double *AVG) {
    double w_old[5];
    memcpy(w_old, w, sizeof w_old);
    for (int m = 0; m < 5; m++) w[m] += (fwL[m] - fwR[m]) / dx;
    orc_conserve_prim(w, 5, gam, prim);
    for (int m = 0; m < 5; m++) {
        RES[m] += (w[m] - w_old[m]) * (w[m] - w_old[m]);
        AVG[m] += fabs(w[m]);
    }
    const size_t n3 = (size_t)nu * nv * nw;
    double *Q = (double *)malloc(sizeof(double) * n3);
    orc_boltzmann_fft(Q, f, Kn, M, phi, psi, phipsi, nu, nv, nw);
    for (size_t i = 0; i < n3; i++) f[i] += (ffL[i] - ffR[i]) / dx + dt * Q[i];
    free(Q);
    return 0;
}
