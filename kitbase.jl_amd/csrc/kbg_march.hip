// kbg_march.hip -- the fused 1D time step (reconstruct! + evolve! + update! in one pass) for gfx950.
//
// One wavefront owns a segment [c0, c1] of cells of one solve and marches along x in blocks of B
// cells.  Lane l holds the velocities l + 64 v (v < VPT) of every row it touches, so each f row is
// read and written with fully coalesced 256-byte wavefront accesses; f^n is read once (plus a
// two-row halo per segment) and f^{n+1} written once, the compulsory HBM traffic of the step.  The
// rows of the next block are loaded while the current block computes.
//
// Per block:
//   1. velocity phase: limited slopes and the B face fluxes per velocity, kept in registers, with
//      per-lane partial sums of the face moments (and of the Shakhov heat flux of each cell);
//   2. one reduce-scatter of all NVAL*B partial sums across the wavefront (a pairwise tree in the
//      storage precision: at most VPT + 6 roundings per moment);
//   3. lane-parallel scalars: lane j finishes face j's moments and lane t runs cell t's FP64 chain
//      (w update, conserve_prim, tau, Maxwellian and Shakhov coefficients) -- B cells in one pass
//      instead of the whole wavefront repeating each cell's scalar work;
//   4. velocity phase: the coefficients are broadcast from lane t and every lane relaxes its
//      velocities of cell t.
// KCU and UGKS face coefficients (interface state, Mt, pdf slopes) are likewise computed by one
// lane per face and broadcast.  The segment's per-cell FP64 scalars (w, prim, dx) are staged in the
// wavefront's LDS slice.
//
// Reference semantics (KitBase.jl v0.9.31):
//   slopes    solver_reconstruction.jl:56-153, reconstruct3! reconstruction.jl:136-151
//   faces     solver_evolution.jl:39-89; flux_kfvs.jl:9-64,238-340; flux_kcu.jl:12-102,241-530;
//             UGKS flux_ugks.jl:8-114 with wL = w + dx/2 sw (as solver_evolution.jl:285-312)
//   cells     solver_update.jl:61-93, update_boundary.jl:6-61; step_1f.jl:6-55, step_2f.jl:6-64
#include "kbg_device.h"
#include "kbg_types.h"

namespace kbg {

template <class T, int NF, int VPT> struct Row {
    T a[NF][VPT];
};

struct RAcc {
    double ri[3], ai[3], rb[3], ab[3];
};

// block geometry: B cells per block, NVAL partial sums per face, V = NVAL * B reduced values
template <class T, int NF, int FLUX, int VPT> struct Blk {
    static constexpr int es = sizeof(T) / 4;
    static constexpr int raw = 16 / (VPT * NF * es);  // ~ (3 B + 5) rows of VPT*NF*es registers
    static constexpr int b = raw >= 8 ? 8 : (raw >= 4 ? 4 : 2);
    static constexpr int NVAL = FLUX == UGKS ? 8 : 4;
    static constexpr int B = (b * NVAL > 64) ? 64 / NVAL : b;
    static constexpr int V = B * NVAL;
    static constexpr int QI = FLUX == UGKS ? 6 : 3;  // index of the Shakhov heat-flux partial sum
    static constexpr int NRAW = FLUX == UGKS ? 6 : 3;
};

// per-cell FP64 scalars of a segment, staged in the wavefront's LDS slice:
// w[3], prim[3], dx, 1 / (0.5 (dx[j-1] + dx[j]))
constexpr int kLdsCell = 8;

// face coefficients broadcast to the velocity phase (storage precision)
template <class T> struct FaceC {
    T A, lam, Ui, kb, Mxi2, M0, M1, M2, M3, M4, aL0, aL1, aL2, aR0, aR1, aR2, aT0, aT1, aT2;
};
// face coefficients kept by the face's own lane (FP64)
struct FaceK {
    double fw0[3], m3, m4;
};
// cell coefficients broadcast to the relaxation phase
template <class T> struct CellC {
    T An, ln, Un, rT, inv, rdx, kbn, pre, Ao, lo, Uo, kbo;
};

template <class T, int NF, int FLUX, int VPT>
__device__ __forceinline__ void march1d(const Solve1D& S, int c0, int c1, int par, int lane, bool want_res, RAcc& acc,
                                        double* __restrict__ lds) {
    using P = Blk<T, NF, FLUX, VPT>;
    constexpr int B = P::B, NVAL = P::NVAL, V = P::V, QI = P::QI, NRAW = P::NRAW;
    using R = Row<T, NF, VPT>;
    const int nv = S.nv, ng = S.ng, nx = S.nx, RS = NF * nv;
    const int in = (S.cur + par) & 1;
    const gptr<const T> fin = to_global((const T*)S.f[in]);
    const gptr<T> fout = to_global((T*)S.f[in ^ 1]);
    const gptr<const double> win = to_global((const double*)S.w[in]);
    const gptr<double> wout = to_global(S.w[in ^ 1]);
    const gptr<double> gprim = to_global(S.prim);
    const gptr<const double> gdx = to_global(S.dx);
    const gptr<const T> Uv = to_global((const T*)S.u);
    const gptr<const T> WT = to_global((const T*)S.wt);
    const double dt = S.dt, K = S.K;
    const int lk = S.limiter;
    const bool shakS = S.coll == SHAKHOV;
    const int jlo = c0 - 2, jmax = c1 + 2;  // rows the segment touches

    T u[VPT], wt[VPT];
    bool act[VPT];
#pragma unroll
    for (int v = 0; v < VPT; v++) {
        const int iv = lane + 64 * v;
        act[v] = iv < nv;
        u[v] = act[v] ? Uv[iv] : T(0);
        wt[v] = act[v] ? WT[iv] : T(0);
    }
    const T dtT = T(dt), hdt2 = T(0.5 * dt * dt);
    auto exists = [&](int i) { return i >= 1 - ng && i <= nx + ng && i <= jmax; };

    // ---- stage w, prim (Shakhov only), dx and the slope distance of rows jlo..jmax into LDS
    {
        const int ncell = jmax - jlo + 1;
        for (int x = lane; x < ncell * kLdsCell; x += 64) {
            const int cc = x / kLdsCell, comp = x - cc * kLdsCell, i = jlo + cc;
            double val = comp >= 6 ? 1.0 : 0.0;
            if (exists(i)) {
                const size_t r = (size_t)(i - 1 + ng);
                if (comp < 3) val = win[r * 3 + comp];
                else if (comp < 6) val = shakS ? gprim[r * 3 + comp - 3] : 0.0;
                else if (comp == 6) val = gdx[r];
                else if (i - 1 >= 1 - ng) val = 1.0 / (0.5 * (gdx[r - 1] + gdx[r]));
            }
            lds[x] = val;
        }
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    }
    auto L = [&](int i, int comp) -> double { return lds[(i - jlo) * kLdsCell + comp]; };

    auto load = [&](int i, R& r) {
        if (exists(i)) {
            const gptr<const T> row = fin + (size_t)(i - 1 + ng) * RS;
#pragma unroll
            for (int q = 0; q < NF; q++)
#pragma unroll
                for (int v = 0; v < VPT; v++) r.a[q][v] = act[v] ? row[q * nv + lane + 64 * v] : T(0);
        } else {
#pragma unroll
            for (int q = 0; q < NF; q++)
#pragma unroll
                for (int v = 0; v < VPT; v++) r.a[q][v] = T(0);
        }
    };
    auto zero = [&](R& r) {
#pragma unroll
        for (int q = 0; q < NF; q++)
#pragma unroll
            for (int v = 0; v < VPT; v++) r.a[q][v] = T(0);
    };
    // limited slope of cell j from rows j-1, j, j+1 (reconstruct3!)
    auto slope = [&](int j, const R& a, const R& c, const R& e, R& s) {
        if (j >= S.slope_lo && j <= S.slope_hi) {
            const T rL = T(L(j, 7)), rR = T(L(j + 1, 7));
#pragma unroll
            for (int q = 0; q < NF; q++)
#pragma unroll
                for (int v = 0; v < VPT; v++)
                    s.a[q][v] = limf<T>(lk, (c.a[q][v] - a.a[q][v]) * rL, (e.a[q][v] - c.a[q][v]) * rR);
        } else {
            zero(s);
        }
    };
    // conserved-variable slope component m of cell j (FP64, any lane)
    auto wslope = [&](int j, int m) -> double {
        if (j >= S.slope_lo && j <= S.slope_hi)
            return limiter<double>(lk, (L(j, m) - L(j - 1, m)) * L(j, 7), (L(j + 1, m) - L(j, m)) * L(j + 1, 7));
        return 0.0;
    };
    auto buffered = [&](int k) { return (k == 1 && S.faceL_buf) || (k == nx + 1 && S.faceR_buf); };

    // FP64 coefficients of face k (KCU / UGKS) from the staged w
    auto face_coef = [&](int k, FaceC<T>& C, FaceK& Kc) {
        const double dxL = 0.5 * L(k - 1, 6), dxR = 0.5 * L(k, 6);
        double wL[3], wR[3];
        for (int m = 0; m < 3; m++) {
            wL[m] = L(k - 1, m) + wslope(k - 1, m) * dxL;
            wR[m] = L(k, m) - wslope(k, m) * dxR;
        }
        if (FLUX == KCU) {
            KcuFace<3> F;
            kcu_face<3>(wL, wR, K, S.gam, S.mu, S.om, dt, F);
            C.A = T(F.prim[0] * sqrt(F.prim[2] / kPi));
            C.lam = T(F.prim[2]);
            C.Ui = T(F.prim[1]);
            C.kb = T(K / (2.0 * F.prim[2]));
            C.M0 = T(F.Mt1);
            C.M1 = T(F.Mt2);
            for (int m = 0; m < 3; m++) Kc.fw0[m] = F.fw0[m];
            Kc.m3 = F.Mt2;
            Kc.m4 = 0.0;
        } else if (FLUX == UGKS) {
            UgksFace1 F;
            ugks_face_1d(wL, wR, K, S.gam, S.mu, S.om, dt, dxL, dxR, F);
            C.A = T(F.prim[0] * sqrt(F.prim[2] / kPi));
            C.lam = T(F.prim[2]);
            C.Ui = T(F.prim[1]);
            C.kb = T(K / (2.0 * F.prim[2]));
            C.Mxi2 = T(F.Mxi2);
            C.M0 = T(F.Mt[0]);
            C.M1 = T(F.Mt[1]);
            C.M2 = T(F.Mt[2]);
            C.M3 = T(F.Mt[3]);
            C.M4 = T(F.Mt[4]);
            C.aL0 = T(F.aL[0]);
            C.aL1 = T(F.aL[1]);
            C.aL2 = T(F.aL[2]);
            C.aR0 = T(F.aR[0]);
            C.aR1 = T(F.aR[1]);
            C.aR2 = T(F.aR[2]);
            C.aT0 = T(F.aT[0]);
            C.aT1 = T(F.aT[1]);
            C.aT2 = T(F.aT[2]);
            for (int m = 0; m < 3; m++) Kc.fw0[m] = F.fw0[m];
            Kc.m3 = F.Mt[3];
            Kc.m4 = F.Mt[4];
        }
    };
    auto bcast_coef = [&](const FaceC<T>& C, int l) {
        FaceC<T> o;
        o.A = bcast(C.A, l);
        o.lam = bcast(C.lam, l);
        o.Ui = bcast(C.Ui, l);
        o.kb = bcast(C.kb, l);
        o.M0 = bcast(C.M0, l);
        o.M1 = bcast(C.M1, l);
        if (FLUX == UGKS) {
            o.Mxi2 = bcast(C.Mxi2, l);
            o.M2 = bcast(C.M2, l);
            o.M3 = bcast(C.M3, l);
            o.M4 = bcast(C.M4, l);
            o.aL0 = bcast(C.aL0, l);
            o.aL1 = bcast(C.aL1, l);
            o.aL2 = bcast(C.aL2, l);
            o.aR0 = bcast(C.aR0, l);
            o.aR1 = bcast(C.aR1, l);
            o.aR2 = bcast(C.aR2, l);
            o.aT0 = bcast(C.aT0, l);
            o.aT1 = bcast(C.aT1, l);
            o.aT2 = bcast(C.aT2, l);
        }
        return o;
    };
    // per-velocity flux of face k between rows (a, sa) and (b, sb); raw partial sums into ps[0..NRAW)
    auto face_flux = [&](int k, const R& a, const R& sa, const R& b, const R& sb, const FaceC<T>& C, R& ff, double* ps) {
        const T dxLT = T(0.5 * L(k - 1, 6)), dxRT = T(0.5 * L(k, 6));
        if (FLUX == KFVS) {
            // flux_kfvs.jl:238-264 (1F1V), :303-340 (2F1V)
            T m0 = 0, m1 = 0, m2 = 0, m3 = 0;
#pragma unroll
            for (int v = 0; v < VPT; v++) {
                const bool d = u[v] >= T(0);
                const T uu = u[v] * u[v];
#pragma unroll
                for (int q = 0; q < NF; q++) {
                    const T fu = d ? a.a[q][v] + sa.a[q][v] * dxLT : b.a[q][v] - sb.a[q][v] * dxRT;
                    const T su = d ? sa.a[q][v] : sb.a[q][v];
                    ff.a[q][v] = dtT * u[v] * fu - hdt2 * uu * su;
                }
                const T g = wt[v] * ff.a[0][v];
                m0 += g;
                m1 += u[v] * g;
                m2 += uu * g;
                if (NF == 2) m3 += wt[v] * ff.a[NF - 1][v];
            }
            ps[0] = double(m0);
            ps[1] = double(m1);
            ps[2] = double(m2) + double(m3);
        } else if (FLUX == KCU) {
            // flux_kcu.jl:241-303 (1F1V), :463-530 (2F1V)
            T s0 = 0, s1 = 0, s2 = 0, s3 = 0;
#pragma unroll
            for (int v = 0; v < VPT; v++) {
                const bool d = u[v] >= T(0);
                const T cc = u[v] - C.Ui;
                const T g = C.A * texp(-C.lam * (cc * cc));
                const T h = d ? a.a[0][v] + sa.a[0][v] * dxLT : b.a[0][v] - sb.a[0][v] * dxRT;
                s0 += wt[v] * u[v] * h;
                s1 += wt[v] * (u[v] * u[v]) * h;
                s2 += wt[v] * (u[v] * u[v] * u[v]) * h;
                ff.a[0][v] = C.M0 * u[v] * g + C.M1 * u[v] * h;
                if (NF == 2) {
                    const T bb = d ? a.a[NF - 1][v] + sa.a[NF - 1][v] * dxLT : b.a[NF - 1][v] - sb.a[NF - 1][v] * dxRT;
                    s3 += wt[v] * u[v] * bb;
                    ff.a[NF - 1][v] = C.M0 * u[v] * (g * C.kb) + C.M1 * u[v] * bb;
                }
            }
            ps[0] = double(s0);
            ps[1] = double(s1);
            ps[2] = double(s2) + double(s3);
        } else {
            // flux_ugks.jl:8-114 (1D2F1V)
            T s0 = 0, s1 = 0, s2 = 0, s3 = 0, s4 = 0, s5 = 0;
#pragma unroll
            for (int v = 0; v < VPT; v++) {
                const bool d = u[v] >= T(0);
                const T x = u[v], uu = x * x;
                const T h = d ? a.a[0][v] + sa.a[0][v] * dxLT : b.a[0][v] - sb.a[0][v] * dxRT;
                const T bb = d ? a.a[NF - 1][v] + sa.a[NF - 1][v] * dxLT : b.a[NF - 1][v] - sb.a[NF - 1][v] * dxRT;
                const T sh = d ? sa.a[0][v] : sb.a[0][v];
                const T sbv = d ? sa.a[NF - 1][v] : sb.a[NF - 1][v];
                const T cc = x - C.Ui;
                const T H = C.A * texp(-C.lam * (cc * cc));
                const T Bq = H * C.kb;
                const T wq = wt[v];
                s0 += wq * x * h;
                s1 += wq * uu * sh;
                s2 += wq * uu * h;
                s3 += wq * (uu * x) * sh;
                s4 += wq * (uu * x) * h + wq * x * bb;
                s5 += wq * (uu * uu) * sh + wq * uu * sbv;
                const T a0 = d ? C.aL0 : C.aR0, a1 = d ? C.aL1 : C.aR1, a2 = d ? C.aL2 : C.aR2;
                const T pH = a0 * H + a1 * x * H + T(0.5) * a2 * (uu * H + Bq);
                const T pTH = C.aT0 * H + C.aT1 * x * H + T(0.5) * C.aT2 * (uu * H + Bq);
                const T pB = a0 * Bq + a1 * x * Bq + T(0.5) * a2 * (uu * Bq + C.Mxi2 * H);
                const T pTB = C.aT0 * Bq + C.aT1 * x * Bq + T(0.5) * C.aT2 * (uu * Bq + C.Mxi2 * H);
                ff.a[0][v] = C.M0 * x * H + C.M1 * uu * pH + C.M2 * x * pTH + C.M3 * x * h - C.M4 * uu * sh;
                ff.a[NF - 1][v] = C.M0 * x * Bq + C.M1 * uu * pB + C.M2 * x * pTB + C.M3 * x * bb - C.M4 * uu * sbv;
            }
            ps[0] = double(s0);
            ps[1] = double(s1);
            ps[2] = double(s2);
            ps[3] = double(s3);
            ps[4] = double(s4);
            ps[5] = double(s5);
        }
    };
    // face moments from the reduced raw sums and the face's own FP64 coefficients
    auto finish_face = [&](const double* r, const FaceK& Kc, double* fw) {
        if (FLUX == KFVS) {
            fw[0] = r[0];
            fw[1] = r[1];
            fw[2] = 0.5 * r[2];
        } else if (FLUX == KCU) {
            fw[0] = Kc.fw0[0] + Kc.m3 * r[0];
            fw[1] = Kc.fw0[1] + Kc.m3 * r[1];
            fw[2] = Kc.fw0[2] + Kc.m3 * 0.5 * r[2];
        } else {
            fw[0] = Kc.fw0[0] + (Kc.m3 * r[0] - Kc.m4 * r[1]);
            fw[1] = Kc.fw0[1] + (Kc.m3 * r[2] - Kc.m4 * r[3]);
            fw[2] = Kc.fw0[2] + (Kc.m3 * 0.5 * r[4] - Kc.m4 * 0.5 * r[5]);
        }
    };
    // Shakhov heat flux partial of cell c from row fc (theory_moments.jl:523, :530-531)
    auto qpart = [&](int c, const R& fc) -> double {
        const T Uo = T(L(c, 4));
        T qp = T(0);
#pragma unroll
        for (int v = 0; v < VPT; v++) {
            const T cc = u[v] - Uo;
            qp += wt[v] * cc * (cc * cc) * fc.a[0][v];
            if (NF == 2) qp += wt[v] * cc * fc.a[NF - 1][v];
        }
        return double(qp);
    };

    // ---- prologue: face c0 (the left face of the first cell), rows c0 and c0+1, slope of c0
    // Rw[i] = row b0 + i (i = 0 .. B+1) of the current block; ffa[0] = flux of face b0,
    // ffa[t+1] = flux of face b0+1+t
    R Rw[B + 2], ffa[B + 1], sPrev;
    double fwL[3];
    {
        R fm2, fm1, sm1;
        load(c0 - 2, fm2);
        load(c0 - 1, fm1);
        load(c0, Rw[0]);
        load(c0 + 1, Rw[1]);
        slope(c0 - 1, fm2, fm1, Rw[0], sm1);
        slope(c0, fm1, Rw[0], Rw[1], sPrev);
        if (buffered(c0)) {
            const int side = c0 == 1 ? 0 : 1;
            const T* bf = (const T*)S.bff + (size_t)side * RS;
#pragma unroll
            for (int q = 0; q < NF; q++)
#pragma unroll
                for (int v = 0; v < VPT; v++) ffa[0].a[q][v] = act[v] ? bf[q * nv + lane + 64 * v] : T(0);
            for (int m = 0; m < 3; m++) fwL[m] = S.bfw[side * 3 + m];
        } else {
            FaceC<T> C;
            FaceK Kc;
            face_coef(c0, C, Kc);
            double ps[6] = {0, 0, 0, 0, 0, 0};
            face_flux(c0, fm1, sm1, Rw[0], sPrev, C, ffa[0], ps);
            wave_allreduce<NRAW>(ps);
            finish_face(ps, Kc, fwL);
        }
    }

    // ---- blocks of B cells
    R pf[B];
    static_for<B>([&](auto tc) { constexpr int t = tc; load(c0 + 2 + t, pf[t]); });
    for (int b0 = c0; b0 <= c1; b0 += B) {
        static_for<B>([&](auto tc) {
            constexpr int t = tc;
            Rw[2 + t] = pf[t];
            load(b0 + B + 2 + t, pf[t]);
        });

        // face coefficients of faces b0+1 .. b0+B, one lane each (KCU / UGKS)
        FaceC<T> Cmine;
        FaceK Kmine;
        if (FLUX != KFVS) {
            const int kf = min(b0 + 1 + min(lane, B - 1), c1 + 1);
            if (!buffered(kf)) face_coef(kf, Cmine, Kmine);
        }

        // 1. velocity phase: fluxes and partial sums (storage precision, pairwise tree below)
        T vals[V];
#pragma unroll
        for (int i = 0; i < V; i++) vals[i] = T(0);
        static_for<B>([&](auto tc) {
            constexpr int t = tc;
            const int k = b0 + 1 + t, c = b0 + t;
            if (k > c1 + 1) {
                zero(ffa[t + 1]);
                return;
            }
            R sNew;
            slope(k, Rw[t], Rw[t + 1], Rw[t + 2], sNew);
            if (buffered(k)) {
                const T* bf = (const T*)S.bff + (size_t)(k == 1 ? 0 : 1) * RS;
#pragma unroll
                for (int q = 0; q < NF; q++)
#pragma unroll
                    for (int v = 0; v < VPT; v++) ffa[t + 1].a[q][v] = act[v] ? bf[q * nv + lane + 64 * v] : T(0);
            } else {
                FaceC<T> C;
                if (FLUX != KFVS) C = bcast_coef(Cmine, t);
                double ps[6];
                face_flux(k, Rw[t], sPrev, Rw[t + 1], sNew, C, ffa[t + 1], ps);
#pragma unroll
                for (int m = 0; m < NRAW; m++) vals[NVAL * t + m] = T(ps[m]);
            }
            if (shakS && c <= c1 && c > 1 && c < nx) vals[NVAL * t + QI] = T(qpart(c, Rw[t]));
            sPrev = sNew;
        });

        // 2. one reduce-scatter of all partial sums
        const double mine = double(wave_reduce_scatter<V>(vals));

        // 3a. lane j finishes face b0+1+j
        const int jl = min(lane, B - 1);
        double fwj[3], raw[6];
#pragma unroll
        for (int m = 0; m < NRAW; m++) raw[m] = __shfl(mine, scatter_lane<V>(NVAL * jl + m), 64);
        const double qj = __shfl(mine, scatter_lane<V>(NVAL * jl + QI), 64);
        finish_face(raw, Kmine, fwj);
        {
            const int kf = b0 + 1 + jl;
            if (buffered(kf))
                for (int m = 0; m < 3; m++) fwj[m] = S.bfw[(kf == 1 ? 0 : 1) * 3 + m];
        }
        // 3b. lane t runs cell b0+t: step! (step_1f.jl:6-55, step_2f.jl:6-64)
        double fwl[3];
#pragma unroll
        for (int m = 0; m < 3; m++) {
            const double prev = __shfl(fwj[m], lane > 0 ? lane - 1 : 0, 64);
            fwl[m] = lane == 0 ? fwL[m] : prev;
        }
        CellC<T> cm;
        {
            const bool mycell = lane < B && b0 + lane <= c1;
            const int c = min(b0 + min(lane, B - 1), c1);
            const double dxcell = L(c, 6);
            double wo[3], wn[3], pn[3];
            for (int m = 0; m < 3; m++) {
                wo[m] = L(c, m);
                wn[m] = wo[m] + (fwl[m] - fwj[m]) / dxcell;
            }
            conserve_prim<3>(wn, S.gam, pn);
            const T rT = relax_ratio<T>(pn[0], pn[2], S.mu, S.om, dt);
            cm.rT = rT;
            cm.inv = T(1) / (T(1) + rT);
            cm.rdx = T(1.0 / dxcell);
            cm.An = T(pn[0] * sqrt(pn[2] / kPi));
            cm.ln = T(pn[2]);
            cm.Un = T(pn[1]);
            cm.kbn = T(K / (2.0 * pn[2]));
            if (shakS) {
                const double p0 = L(c, 3), p1 = L(c, 4), p2 = L(c, 5);
                const double q = 0.5 * qj;
                cm.pre = T(0.8 * (1.0 - S.Pr) * (p2 * p2) / p0 * q);
                cm.Ao = T(p0 * sqrt(p2 / kPi));
                cm.lo = T(p2);
                cm.Uo = T(p1);
                cm.kbo = T(K / (2.0 * p2));
            }
            if (mycell) {
                const size_t r = (size_t)(c - 1 + ng) * 3;
                for (int m = 0; m < 3; m++) {
                    wout[r + m] = wn[m];
                    gprim[r + m] = pn[m];
                }
                if (want_res) {
                    const bool bnd = c == 1 || c == nx;
                    for (int m = 0; m < 3; m++) {
                        const double d2 = (wn[m] - wo[m]) * (wn[m] - wo[m]), a = fabs(wn[m]);
                        if (bnd) {
                            acc.rb[m] += d2;
                            acc.ab[m] += a;
                        } else {
                            acc.ri[m] += d2;
                            acc.ai[m] += a;
                        }
                    }
                }
            }
        }
        // carried left face of the next block
#pragma unroll
        for (int m = 0; m < 3; m++) fwL[m] = __shfl(fwj[m], B - 1, 64);

        // 4. velocity phase: relax every cell of the block
        static_for<B>([&](auto tc) {
            constexpr int t = tc;
            const int c = b0 + t;
            if (c > c1) return;
            const bool shak = shakS && c > 1 && c < nx;
            const T An = bcast(cm.An, t), ln = bcast(cm.ln, t), Un = bcast(cm.Un, t);
            const T rT = bcast(cm.rT, t), inv = bcast(cm.inv, t), rdx = bcast(cm.rdx, t);
            const T kbn = NF == 2 ? bcast(cm.kbn, t) : T(0);
            T pre = 0, Ao = 0, lo = 0, Uo = 0, kbo = 0;
            if (shak) {
                pre = bcast(cm.pre, t);
                Ao = bcast(cm.Ao, t);
                lo = bcast(cm.lo, t);
                Uo = bcast(cm.Uo, t);
                if (NF == 2) kbo = bcast(cm.kbo, t);
            }
            const gptr<T> row = fout + (size_t)(c - 1 + ng) * RS;
#pragma unroll
            for (int v = 0; v < VPT; v++) {
                const T cc = u[v] - Un;
                T MH = An * texp(-ln * (cc * cc));
                T MB = MH * kbn;
                if (shak) {
                    const T co = u[v] - Uo, c2 = co * co;
                    const T Mo = Ao * texp(-lo * c2);
                    if (NF == 1) {
                        MH += pre * co * (T(2) * lo * c2 - T(5)) * Mo;
                    } else {
                        MH += pre * co * (T(2) * lo * c2 + T(K) - T(5)) * Mo;
                        MB += pre * co * (T(2) * lo * c2 + T(K) - T(3)) * (Mo * kbo);
                    }
                }
                if (act[v]) {
                    row[lane + 64 * v] = (Rw[t].a[0][v] + (ffa[t].a[0][v] - ffa[t + 1].a[0][v]) * rdx + rT * MH) * inv;
                    if (NF == 2)
                        row[nv + lane + 64 * v] =
                            (Rw[t].a[NF - 1][v] + (ffa[t].a[NF - 1][v] - ffa[t + 1].a[NF - 1][v]) * rdx + rT * MB) * inv;
                }
            }
        });
        // carry rows and the left-face flux into the next block
        ffa[0] = ffa[B];
        Rw[0] = Rw[B];
        Rw[1] = Rw[B + 1];
    }
}

// occupancy target: 4 wavefronts per SIMD (<= 128 VGPRs) for the FP32 kernels, 2 for FP64
template <class T, int FLUX> constexpr int waves_per_eu() { return 1; }

template <class T, int NF, int FLUX, int VPT>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(waves_per_eu<T, FLUX>()))) void k_march1d(
    const Solve1D* __restrict__ solves, const WorkItem* __restrict__ items, int nitems, int par, double* __restrict__ rpart,
    int seg) {
    extern __shared__ double kbg_lds[];
    const int lane = threadIdx.x & 63;
    double* lds = kbg_lds + (size_t)(threadIdx.x >> 6) * (seg + 4) * kLdsCell;
    const int nwaves = gridDim.x * (blockDim.x >> 6);
    // the wavefront index is uniform: say so, so that the item, its solve and all control flow
    // derived from them live in scalar registers
    const int wid = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    for (int it = blockIdx.x * (blockDim.x >> 6) + wid; it < nitems; it += nwaves) {
        const WorkItem wi = items[it];
        const Solve1D& S = solves[wi.solve];
        RAcc acc;
        for (int m = 0; m < 3; m++) acc.ri[m] = acc.ai[m] = acc.rb[m] = acc.ab[m] = 0.0;
        march1d<T, NF, FLUX, VPT>(S, wi.c0, wi.c1, par, lane, rpart != nullptr, acc, lds);
        if (rpart) {
            // each lane accumulated the cells it owned; sum over the wavefront, lane m writes entry m
            double r[12];
            for (int m = 0; m < 3; m++) {
                r[m] = acc.ri[m];
                r[3 + m] = acc.ai[m];
                r[6 + m] = acc.rb[m];
                r[9 + m] = acc.ab[m];
            }
            wave_allreduce<12>(r);
            double y = 0.0;
#pragma unroll
            for (int m = 0; m < 12; m++)
                if (lane == m) y = r[m];
            if (lane < 12) rpart[(size_t)it * 12 + lane] = y;
        }
        // the next item restages this wavefront's LDS slice
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    }
}

template <class T, int NF, int FLUX>
static hipError_t launch_vpt(int vpt, dim3 g, const Solve1D* d_solves, const WorkItem* d_items, int nitems, int par,
                             double* d_rpart, int seg, hipStream_t st) {
    dim3 b(256);
    const size_t lds = (size_t)4 * (seg + 4) * kLdsCell * sizeof(double);
    switch (vpt) {
    case 1: hipLaunchKernelGGL((k_march1d<T, NF, FLUX, 1>), g, b, lds, st, d_solves, d_items, nitems, par, d_rpart, seg); break;
    case 2: hipLaunchKernelGGL((k_march1d<T, NF, FLUX, 2>), g, b, lds, st, d_solves, d_items, nitems, par, d_rpart, seg); break;
    case 3: hipLaunchKernelGGL((k_march1d<T, NF, FLUX, 3>), g, b, lds, st, d_solves, d_items, nitems, par, d_rpart, seg); break;
    case 4: hipLaunchKernelGGL((k_march1d<T, NF, FLUX, 4>), g, b, lds, st, d_solves, d_items, nitems, par, d_rpart, seg); break;
    default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

hipError_t launch_march1d(bool fp64, int nf, int flux, int vpt, const Solve1D* d_solves, const WorkItem* d_items,
                          int nitems, int grid, int par, double* d_rpart, int seg, hipStream_t st) {
    if (nitems <= 0) return hipSuccess;
    dim3 g(grid);
#define KBG_M(T, NF, FL) return launch_vpt<T, NF, FL>(vpt, g, d_solves, d_items, nitems, par, d_rpart, seg, st)
    if (!fp64) {
        if (nf == 1 && flux == KFVS) KBG_M(float, 1, KFVS);
        if (nf == 1 && flux == KCU) KBG_M(float, 1, KCU);
        if (nf == 2 && flux == KFVS) KBG_M(float, 2, KFVS);
        if (nf == 2 && flux == KCU) KBG_M(float, 2, KCU);
        if (nf == 2 && flux == UGKS) KBG_M(float, 2, UGKS);
    } else {
        if (nf == 1 && flux == KFVS) KBG_M(double, 1, KFVS);
        if (nf == 1 && flux == KCU) KBG_M(double, 1, KCU);
        if (nf == 2 && flux == KFVS) KBG_M(double, 2, KFVS);
        if (nf == 2 && flux == KCU) KBG_M(double, 2, KCU);
        if (nf == 2 && flux == UGKS) KBG_M(double, 2, UGKS);
    }
#undef KBG_M
    return hipErrorInvalidValue;
}

}  // namespace kbg
