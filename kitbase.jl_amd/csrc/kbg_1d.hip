// kbg_1d.hip -- 1D discrete-velocity kernels for gfx950 (phase-exact path).
//
// Two paths share the physics (kbg_device.h):
//  * k_sweep1d (kbg_sweep.h): the fused time step of one-velocity-dimension solves.
//  * k_recon1d / k_evolve1d / k_wall1d / k_update1d: one workgroup per cell or face, the three
//    phases kept separate exactly as KitBase runs them, with slopes and face fluxes stored in HBM.
//    These back the phase-level entry points and handle any velocity count (1F3V included).
#include "kbg_device.h"
#include "kbg_sweep.h"
#include "kbg_types.h"

namespace kbg {

// One cell's distributions in the solve's f layout: element o = q nv + v (q = 0 h, 1 b)
template <class T> struct FRow {
    T* p;
    int nv;
    long qs, vs;
    KBG_D T& operator[](int o) const {
        const int q = o >= nv ? 1 : 0;
        return p[q * qs + (long)(o - q * nv) * vs];
    }
};
template <class T> KBG_D FRow<T> frow(const Solve1D& S, T* f, int k) { return FRow<T>{f + (long)k * S.fcs, S.nv, S.fqs, S.fvs}; }

// ------------------------------------------------------------------------------------------------
// phase-exact kernels (one workgroup of 256 threads per cell or face, 1024 for large velocity grids)
// ------------------------------------------------------------------------------------------------

// reconstruct!  solver_reconstruction.jl:56-153: slopes of the cells slope_lo..slope_hi
template <class T> __global__ __launch_bounds__(1024) void k_recon1d(const Solve1D* __restrict__ pS) {
    const Solve1D& S = *pS;
    const int i = S.slope_lo + blockIdx.x;
    const int k = i - 1 + S.ng, nw = S.nw, RS = S.nf * S.nv;
    const double dL = 0.5 * (S.dx[k - 1] + S.dx[k]), dR = 0.5 * (S.dx[k] + S.dx[k + 1]);
    const double* w = S.w[S.cur];
    if (threadIdx.x < nw) {
        const int m = threadIdx.x;
        S.sw[(size_t)k * nw + m] = limiter<double>(S.limiter, (w[(size_t)k * nw + m] - w[(size_t)(k - 1) * nw + m]) / dL,
                                                   (w[(size_t)(k + 1) * nw + m] - w[(size_t)k * nw + m]) / dR);
    }
    const T* f = (const T*)S.f[S.cur];
    const FRow<const T> fa = frow(S, f, k - 1), fc = frow(S, f, k), fe = frow(S, f, k + 1);
    T* sf = (T*)S.sf;
    const T rL = T(1.0 / dL), rR = T(1.0 / dR);
    for (int q = threadIdx.x; q < RS; q += blockDim.x) {
        const T a = fa[q], c = fc[q], e = fe[q];
        sf[(size_t)k * RS + q] = limf<T>(S.limiter, (c - a) * rL, (e - c) * rR);
    }
}

// evolve!  solver_evolution.jl:39-89 over the faces face_lo..face_hi (flux kernels flux_kfvs.jl,
// flux_kcu.jl, flux_ugks.jl; 1F1V, 2F1V and 1F3V)
template <class T, int NF, int VD, int FLUX> __global__ __launch_bounds__(1024) void k_evolve1d(const Solve1D* __restrict__ pS) {
    constexpr int NW = VD == 3 ? 5 : 3;
    __shared__ double red_s[16 * 8];
    const Solve1D& S = *pS;
    const int k = S.face_lo + blockIdx.x;  // physical face
    const int ng = S.ng, nv = S.nv, RS = NF * nv;
    const int kL = k - 2 + ng, kR = k - 1 + ng;
    const double dxL = 0.5 * S.dx[kL], dxR = 0.5 * S.dx[kR];
    const T dxLT = T(dxL), dxRT = T(dxR);
    const T* f = (const T*)S.f[S.cur];
    const T* sf = (const T*)S.sf;
    const double* W = S.w[S.cur];
    const T* U = (const T*)S.u;
    const T* V = (const T*)S.v;
    const T* Wv = (const T*)S.wv;
    const T* WT = (const T*)S.wt;
    T* ff = (T*)S.ff + (size_t)(k - 1) * RS;
    double* fw = S.fw + (size_t)(k - 1) * NW;
    const double dt = S.dt, K = S.K;
    const T dtT = T(dt), hdt2 = T(0.5 * dt * dt);
    const FRow<const T> fL = frow(S, f, kL), fR = frow(S, f, kR);
    const T* sL = sf + (size_t)kL * RS;
    const T* sR = sf + (size_t)kR * RS;
    double wL[NW], wR[NW];
    for (int m = 0; m < NW; m++) {
        wL[m] = W[(size_t)kL * NW + m] + S.sw[(size_t)kL * NW + m] * dxL;
        wR[m] = W[(size_t)kR * NW + m] - S.sw[(size_t)kR * NW + m] * dxR;
    }
    KSum<T> acc[8];
    if (FLUX == KFVS) {
        for (int iv = threadIdx.x; iv < nv; iv += blockDim.x) {
            const T x = U[iv], wq = WT[iv];
            const bool d = x >= T(0);
            T fv[NF];
            for (int q = 0; q < NF; q++) {
                const int o = q * nv + iv;
                const T fu = d ? fL[o] + sL[o] * dxLT : fR[o] - sR[o] * dxRT;
                const T su = d ? sL[o] : sR[o];
                fv[q] = dtT * x * fu - hdt2 * (x * x) * su;
                ff[o] = fv[q];
            }
            const T g = wq * fv[0];
            acc[0].add(g);
            acc[1].add(x * g);
            if (VD == 1) {
                acc[2].add((x * x) * g);
                if (NF == 2) acc[3].add(wq * fv[NF - 1]);
            } else {
                const T y = V[iv], z = Wv[iv];
                acc[2].add(y * g);
                acc[3].add(z * g);
                acc[4].add((x * x + y * y + z * z) * g);
            }
        }
        double r[5] = {acc[0].val(), acc[1].val(), acc[2].val(), acc[3].val(), acc[4].val()};
        block_allreduce<5>(r, red_s);
        if (threadIdx.x == 0) {
            if (VD == 1) {
                fw[0] = r[0];
                fw[1] = r[1];
                fw[2] = 0.5 * (r[2] + r[3]);
            } else {
                fw[0] = r[0];
                fw[1] = r[1];
                fw[2] = r[2];
                fw[3] = r[3];
                fw[4] = 0.5 * r[4];
            }
        }
    } else if (FLUX == KCU) {
        KcuFace<NW> F;
        kcu_face<NW>(wL, wR, K, S.gam, S.mu, S.om, dt, F);
        const double lamd = F.prim[NW - 1];
        const T A = VD == 1 ? T(F.prim[0] * sqrt(lamd / kPi)) : T(F.prim[0] * sqrt((lamd / kPi) * (lamd / kPi) * (lamd / kPi)));
        const T lam = T(lamd), Ui = T(F.prim[1]), Vi = T(VD == 3 ? F.prim[2] : 0.0), Wi = T(VD == 3 ? F.prim[3] : 0.0);
        const T Mt1 = T(F.Mt1), Mt2 = T(F.Mt2), kb = T(K / (2.0 * lamd));
        for (int iv = threadIdx.x; iv < nv; iv += blockDim.x) {
            const T x = U[iv], wq = WT[iv];
            const bool d = x >= T(0);
            T c2 = (x - Ui) * (x - Ui);
            T y = 0, z = 0;
            if (VD == 3) {
                y = V[iv];
                z = Wv[iv];
                c2 += (y - Vi) * (y - Vi) + (z - Wi) * (z - Wi);
            }
            const T g = A * texp(-lam * c2);
            const T h = d ? fL[iv] + sL[iv] * dxLT : fR[iv] - sR[iv] * dxRT;
            ff[iv] = Mt1 * x * g + Mt2 * x * h;
            acc[0].add(wq * x * h);
            acc[1].add(wq * (x * x) * h);
            if (VD == 1) {
                acc[2].add(wq * (x * x * x) * h);
                if (NF == 2) {
                    const int o = nv + iv;
                    const T b = d ? fL[o] + sL[o] * dxLT : fR[o] - sR[o] * dxRT;
                    ff[o] = Mt1 * x * (g * kb) + Mt2 * x * b;
                    acc[3].add(wq * x * b);
                }
            } else {
                acc[2].add(wq * x * y * h);
                acc[3].add(wq * x * z * h);
                acc[4].add(wq * x * (x * x + y * y + z * z) * h);
            }
        }
        double r[5] = {acc[0].val(), acc[1].val(), acc[2].val(), acc[3].val(), acc[4].val()};
        block_allreduce<5>(r, red_s);
        if (threadIdx.x == 0) {
            if (VD == 1) {
                fw[0] = F.fw0[0] + F.Mt2 * r[0];
                fw[1] = F.fw0[1] + F.Mt2 * r[1];
                fw[2] = F.fw0[2] + F.Mt2 * 0.5 * (r[2] + r[3]);
            } else {
                fw[0] = F.fw0[0] + F.Mt2 * r[0];
                fw[1] = F.fw0[1] + F.Mt2 * r[1];
                fw[2] = F.fw0[2] + F.Mt2 * r[2];
                fw[3] = F.fw0[3] + F.Mt2 * r[3];
                fw[4] = F.fw0[4] + F.Mt2 * 0.5 * r[4];
            }
        }
    } else {
        UgksFace1 F;
        ugks_face_1d(wL, wR, K, S.gam, S.mu, S.om, dt, dxL, dxR, F);
        const T A = T(F.prim[0] * sqrt(F.prim[2] / kPi)), lam = T(F.prim[2]), Ui = T(F.prim[1]);
        const T kb = T(K / (2.0 * F.prim[2])), Mxi2 = T(F.Mxi2);
        const T Mt0 = T(F.Mt[0]), Mt1 = T(F.Mt[1]), Mt2 = T(F.Mt[2]), Mt3 = T(F.Mt[3]), Mt4 = T(F.Mt[4]);
        for (int iv = threadIdx.x; iv < nv; iv += blockDim.x) {
            const T x = U[iv], wq = WT[iv], uu = x * x;
            const bool d = x >= T(0);
            const int o = nv + iv;
            const T h = d ? fL[iv] + sL[iv] * dxLT : fR[iv] - sR[iv] * dxRT;
            const T b = d ? fL[o] + sL[o] * dxLT : fR[o] - sR[o] * dxRT;
            const T sh = d ? sL[iv] : sR[iv];
            const T sbv = d ? sL[o] : sR[o];
            const T cc = x - Ui;
            const T H = A * texp(-lam * (cc * cc));
            const T B = H * kb;
            acc[0].add(wq * x * h);
            acc[1].add(wq * uu * sh);
            acc[2].add(wq * uu * h);
            acc[3].add(wq * (uu * x) * sh);
            acc[4].add(wq * (uu * x) * h + wq * x * b);
            acc[5].add(wq * (uu * uu) * sh + wq * uu * sbv);
            const double* aL = F.aL;
            const double* aR = F.aR;
            const double* aT = F.aT;
            const T pLH = T(aL[0]) * H + T(aL[1]) * x * H + T(0.5) * T(aL[2]) * (uu * H + B);
            const T pRH = T(aR[0]) * H + T(aR[1]) * x * H + T(0.5) * T(aR[2]) * (uu * H + B);
            const T pTH = T(aT[0]) * H + T(aT[1]) * x * H + T(0.5) * T(aT[2]) * (uu * H + B);
            const T pLB = T(aL[0]) * B + T(aL[1]) * x * B + T(0.5) * T(aL[2]) * (uu * B + Mxi2 * H);
            const T pRB = T(aR[0]) * B + T(aR[1]) * x * B + T(0.5) * T(aR[2]) * (uu * B + Mxi2 * H);
            const T pTB = T(aT[0]) * B + T(aT[1]) * x * B + T(0.5) * T(aT[2]) * (uu * B + Mxi2 * H);
            ff[iv] = Mt0 * x * H + Mt1 * uu * (d ? pLH : pRH) + Mt2 * x * pTH + Mt3 * x * h - Mt4 * uu * sh;
            ff[o] = Mt0 * x * B + Mt1 * uu * (d ? pLB : pRB) + Mt2 * x * pTB + Mt3 * x * b - Mt4 * uu * sbv;
        }
        double r[6] = {acc[0].val(), acc[1].val(), acc[2].val(), acc[3].val(), acc[4].val(), acc[5].val()};
        block_allreduce<6>(r, red_s);
        if (threadIdx.x == 0) {
            fw[0] = F.fw0[0] + (F.Mt[3] * r[0] - F.Mt[4] * r[1]);
            fw[1] = F.fw0[1] + (F.Mt[3] * r[2] - F.Mt[4] * r[3]);
            fw[2] = F.fw0[2] + (F.Mt[3] * 0.5 * r[4] - F.Mt[4] * 0.5 * r[5]);
        }
    }
}

// Maxwell diffuse wall, 1D2F1V   flux_boundary.jl:39-61 via evolve_boundary.jl:4-47.
// blockIdx.x = 0: face 1 from cell 1 (rot = 1); 1: face nx+1 from cell nx (rot = -1).
// fused != 0 writes the boundary-face buffers (reading f[cur ^ par]) instead of fw/ff.
template <class T> __global__ __launch_bounds__(1024) void k_wall1d(const Solve1D* __restrict__ pS, int fused, int par) {
    __shared__ double red_s[16 * 8];
    const Solve1D& S = *pS;
    const int side = blockIdx.x;
    if ((side == 0 && !S.wallL) || (side == 1 && !S.wallR)) return;
    const int nv = S.nv, RS = 2 * nv;
    const int cell = side == 0 ? 1 : S.nx;
    const int rot = side == 0 ? 1 : -1;
    const int k = cell - 1 + S.ng;
    const int buf = fused ? ((S.cur + par) & 1) : S.cur;
    const FRow<const T> h = frow(S, (const T*)S.f[buf], k);
    const T* U = (const T*)S.u;
    const T* WT = (const T*)S.wt;
    const double* bc = S.bcprim + side * 3;
    const double dt = S.dt, K = S.K;
    // the fused FP32 sweep stores the departure from M~(prim) (kbg_sweep.h): add it back
    const bool del = fused && S.delta;
    MaxwCoef mc{};
    if (del) {
        double p[3];
        conserve_prim<3>(S.w[buf] + (size_t)k * 3, S.gam, p);
        mc = maxw_coef(p, K);
    }
    const VChunk<float>* vtf = (const VChunk<float>*)S.vtab;
    auto hv = [&](int q, int iv) -> T {
        T x = h[q * nv + iv];
        if (del) {
            const float M = maxw_at(vtf[iv / kVChunk], mc.A, mc.nl2, mc.U, iv % kVChunk);
            x += T(q == 0 ? M : maxw_b(M, mc.kb));
        }
        return x;
    };
    double* fw = fused ? S.bfw + side * 3 : S.fw + (size_t)(side == 0 ? 0 : S.nx) * 3;
    T* ff = fused ? (T*)S.bff + (size_t)side * RS : (T*)S.ff + (size_t)(side == 0 ? 0 : S.nx) * RS;
    const T Uw = T(bc[1]), lw = T(bc[2]);
    KSum<T> a0, a1;
    for (int iv = threadIdx.x; iv < nv; iv += blockDim.x) {
        const T x = U[iv], wq = WT[iv];
        const bool d = x * T(rot) >= T(0);
        const T c = x - Uw;
        if (!d) a0.add(wq * x * hv(0, iv));
        else a1.add(wq * x * texp(-lw * (c * c)));
    }
    double r[2] = {a0.val(), a1.val()};
    block_allreduce<2>(r, red_s);
    const double SG = pow(bc[2] / kPi, 0.5) * r[1];
    const double rho = -r[0] / SG;
    const T A = T(rho * sqrt(bc[2] / kPi)), kb = T(K / (2.0 * bc[2])), dtT = T(dt);
    KSum<T> m0, m1, m2, m3;
    for (int iv = threadIdx.x; iv < nv; iv += blockDim.x) {
        const T x = U[iv], wq = WT[iv];
        const bool d = x * T(rot) >= T(0);
        const T c = x - Uw;
        const T H = A * texp(-lw * (c * c));
        const T hW = d ? H : hv(0, iv);
        const T bW = d ? H * kb : hv(1, iv);
        m0.add(wq * x * hW);
        m1.add(wq * (x * x) * hW);
        m2.add(wq * x * (hW * (x * x)));
        m3.add(wq * x * bW);
        ff[iv] = x * hW * dtT;
        ff[nv + iv] = x * bW * dtT;
    }
    double s[4] = {m0.val(), m1.val(), m2.val(), m3.val()};
    block_allreduce<4>(s, red_s);
    if (threadIdx.x == 0) {
        fw[0] = s[0] * dt;
        fw[1] = s[1] * dt;
        fw[2] = (0.5 * s[2] + 0.5 * s[3]) * dt;
    }
}

// update!  solver_update.jl:61-93 + update_boundary.jl:6-61 over the cells upd_lo..upd_hi; the
// boundary cells 1 and nx are stepped with BGK (update_boundary! does not forward `coll`).
// step kernels step_1f.jl:6-55 (1F1V), :62-115 (1F3V), step_2f.jl:6-64 (2F1V)
template <class T, int NF, int VD> __global__ __launch_bounds__(1024) void k_update1d(const Solve1D* __restrict__ pS, double* __restrict__ rcell) {
    constexpr int NW = VD == 3 ? 5 : 3;
    __shared__ double red_s[16 * 8];
    const Solve1D& S = *pS;
    const int c = S.upd_lo + blockIdx.x;
    const int ng = S.ng, nv = S.nv, RS = NF * nv, nx = S.nx;
    const int k = c - 1 + ng;
    double* w = S.w[S.cur] + (size_t)k * NW;
    double* prim = S.prim + (size_t)k * NW;
    const FRow<T> f = frow(S, (T*)S.f[S.cur], k);
    const T* ffL = (const T*)S.ff + (size_t)(c - 1) * RS;
    const T* ffR = (const T*)S.ff + (size_t)c * RS;
    const double* fwL = S.fw + (size_t)(c - 1) * NW;
    const double* fwR = S.fw + (size_t)c * NW;
    const T* U = (const T*)S.u;
    const T* V = (const T*)S.v;
    const T* Wv = (const T*)S.wv;
    const T* WT = (const T*)S.wt;
    const double K = S.K, dt = S.dt, dxc = S.dx[k];
    const bool shak = VD == 1 && S.coll == SHAKHOV && c > 1 && c < nx;
    // fast spectral step! (step_1f.jl:122-157) for the interior cells; boundary cells take BGK
    const bool fsm = VD == 3 && S.coll == FSM && c > 1 && c < nx;
    double po[NW], wo[NW];
    for (int m = 0; m < NW; m++) {
        po[m] = prim[m];
        wo[m] = w[m];
    }
    double q = 0.0;
    if (shak) {
        KSum<T> a;
        const T Uo = T(po[1]);
        for (int iv = threadIdx.x; iv < nv; iv += blockDim.x) {
            const T cc = U[iv] - Uo;
            a.add(WT[iv] * cc * (cc * cc) * f[iv]);
            if (NF == 2) a.add(WT[iv] * cc * f[nv + iv]);
        }
        double r[1] = {a.val()};
        block_allreduce<1>(r, red_s);
        q = 0.5 * r[0];
    }
    double wn[NW], pn[NW];
    for (int m = 0; m < NW; m++) wn[m] = wo[m] + (fwL[m] - fwR[m]) / dxc;
    conserve_prim<NW>(wn, S.gam, pn);
    if (threadIdx.x == 0) flag_health(S.health, prim_health<NW>(pn));
    const double lamn = pn[NW - 1];
    const double tau = vhs_tau(pn[0], lamn, S.mu, S.om);
    const double r = dt / tau;
    const T rT = T(r), inv = T(1.0 / (1.0 + r)), rdx = T(1.0 / dxc);
    const T An = VD == 1 ? T(pn[0] * sqrt(lamn / kPi)) : T(pn[0] * sqrt((lamn / kPi) * (lamn / kPi) * (lamn / kPi)));
    const T ln = T(lamn), Un = T(pn[1]), Vn = T(VD == 3 ? pn[2] : 0.0), Wn = T(VD == 3 ? pn[3] : 0.0);
    const T kbn = T(K / (2.0 * lamn));
    T pre = 0, Ao = 0, lo = 0, Uo = 0, kbo = 0;
    if (shak) {
        pre = T(0.8 * (1.0 - S.Pr) * (po[NW - 1] * po[NW - 1]) / po[0] * q);
        Ao = T(po[0] * sqrt(po[NW - 1] / kPi));
        lo = T(po[NW - 1]);
        Uo = T(po[1]);
        kbo = T(K / (2.0 * po[NW - 1]));
    }
    if (fsm) {
        const double* __restrict__ Q = S.fsmq + (size_t)(c - S.fsm_c0) * nv;
        const T dtT = T(dt);
        for (int iv = threadIdx.x; iv < nv; iv += blockDim.x) f[iv] = f[iv] + (ffL[iv] - ffR[iv]) * rdx + dtT * T(Q[iv]);
    }
    for (int iv = threadIdx.x; iv < nv && !fsm; iv += blockDim.x) {
        const T x = U[iv];
        T c2 = (x - Un) * (x - Un);
        if (VD == 3) {
            const T y = V[iv] - Vn, z = Wv[iv] - Wn;
            c2 += y * y + z * z;
        }
        T MH = An * texp(-ln * c2);
        T MB = MH * kbn;
        if (shak) {
            const T co = x - Uo, cq = co * co;
            const T Mo = Ao * texp(-lo * cq);
            if (NF == 1) {
                MH += pre * co * (T(2) * lo * cq - T(5)) * Mo;
            } else {
                MH += pre * co * (T(2) * lo * cq + T(K) - T(5)) * Mo;
                MB += pre * co * (T(2) * lo * cq + T(K) - T(3)) * (Mo * kbo);
            }
        }
        f[iv] = (f[iv] + (ffL[iv] - ffR[iv]) * rdx + rT * MH) * inv;
        if (NF == 2) f[nv + iv] = (f[nv + iv] + (ffL[nv + iv] - ffR[nv + iv]) * rdx + rT * MB) * inv;
    }
    __syncthreads();
    if (threadIdx.x < NW) {
        const int m = threadIdx.x;
        w[m] = wn[m];
        prim[m] = pn[m];
        if (rcell) {
            double* row = rcell + (size_t)blockIdx.x * 4 * NW;
            const bool bnd = (c == 1 || c == nx);
            row[(bnd ? 2 : 0) * NW + m] = (wn[m] - wo[m]) * (wn[m] - wo[m]);
            row[(bnd ? 3 : 1) * NW + m] = fabs(wn[m]);
            row[(bnd ? 0 : 2) * NW + m] = 0.0;
            row[(bnd ? 1 : 3) * NW + m] = 0.0;
        }
    }
}

// residual (solver_update.jl:86-88 and update_boundary.jl:41-44): rows of [res_in, avg_in,
// res_bnd, avg_bnd] x nw, summed in a fixed order by one workgroup.
__global__ __launch_bounds__(256) void k_resid(const double* __restrict__ rows, int nrows, int nw, int stride,
                                                double ncell, double* __restrict__ out) {
    __shared__ double part[256 * 20];
    const int ncol = 4 * nw;
    for (int col = 0; col < ncol; col++) {
        double s = 0.0;
        for (int r = threadIdx.x; r < nrows; r += 256) s += rows[(size_t)r * stride + col];
        part[threadIdx.x * 20 + col] = s;
    }
    __syncthreads();
    if (threadIdx.x < ncol) {
        double s = 0.0;
        for (int t = 0; t < 256; t++) s += part[t * 20 + threadIdx.x];
        part[threadIdx.x] = s;  // row 0 is only read below after the barrier
    }
    __syncthreads();
    if (threadIdx.x < nw) {
        const int m = threadIdx.x;
        const double rin = part[m], ain = part[nw + m], rb = part[2 * nw + m], ab = part[3 * nw + m];
        out[m] = sqrt(rin * ncell) / (ain + 1.e-7) + sqrt(rb * 2.0) / (ab + 1.e-7);
    }
}

// Workgroup size of the phase kernels: 256 threads, or 1024 for large velocity grids (1D1F3V), whose
// few cells and faces would otherwise leave each thread a long serial loop
static int phase_threads(const Solve1D& S) { return S.nv >= 2048 ? 1024 : 256; }

hipError_t launch_recon1d(bool fp64, const Solve1D& S, const Solve1D* d_S, hipStream_t st) {
    const int n = S.slope_hi - S.slope_lo + 1;
    if (n <= 0 || !S.interp2) return hipSuccess;
    if (fp64) hipLaunchKernelGGL(k_recon1d<double>, dim3(n), dim3(phase_threads(S)), 0, st, d_S);
    else hipLaunchKernelGGL(k_recon1d<float>, dim3(n), dim3(phase_threads(S)), 0, st, d_S);
    return hipGetLastError();
}

hipError_t launch_evolve1d(bool fp64, const Solve1D& S, const Solve1D* d_S, hipStream_t st) {
    const int n = S.face_hi - S.face_lo + 1;
    if (n > 0) {
        dim3 g(n), b(phase_threads(S));
#define KBG_E(T, NF, VD, FL) hipLaunchKernelGGL((k_evolve1d<T, NF, VD, FL>), g, b, 0, st, d_S)
#define KBG_EALL(T)                                                          \
    if (S.vdim == 1 && S.nf == 1 && S.flux == KFVS) KBG_E(T, 1, 1, KFVS);    \
    else if (S.vdim == 1 && S.nf == 1 && S.flux == KCU) KBG_E(T, 1, 1, KCU); \
    else if (S.vdim == 1 && S.nf == 2 && S.flux == KFVS) KBG_E(T, 2, 1, KFVS); \
    else if (S.vdim == 1 && S.nf == 2 && S.flux == KCU) KBG_E(T, 2, 1, KCU); \
    else if (S.vdim == 1 && S.nf == 2 && S.flux == UGKS) KBG_E(T, 2, 1, UGKS); \
    else if (S.vdim == 3 && S.nf == 1 && S.flux == KFVS) KBG_E(T, 1, 3, KFVS); \
    else if (S.vdim == 3 && S.nf == 1 && S.flux == KCU) KBG_E(T, 1, 3, KCU); \
    else return hipErrorInvalidValue;
        if (fp64) {
            KBG_EALL(double)
        } else {
            KBG_EALL(float)
        }
#undef KBG_EALL
#undef KBG_E
        hipError_t e = hipGetLastError();
        if (e != hipSuccess) return e;
    }
    return hipSuccess;
}

hipError_t launch_wall1d(bool fp64, const Solve1D& S, const Solve1D* d_S, int fused, int par, hipStream_t st) {
    if (!(S.wallL || S.wallR)) return hipSuccess;
    if (fp64) hipLaunchKernelGGL(k_wall1d<double>, dim3(2), dim3(phase_threads(S)), 0, st, d_S, fused, par);
    else hipLaunchKernelGGL(k_wall1d<float>, dim3(2), dim3(phase_threads(S)), 0, st, d_S, fused, par);
    return hipGetLastError();
}

hipError_t launch_update1d(bool fp64, const Solve1D& S, const Solve1D* d_S, double* d_rcell, hipStream_t st) {
    const int n = S.upd_hi - S.upd_lo + 1;
    if (n <= 0) return hipSuccess;
    dim3 g(n), b(phase_threads(S));
#define KBG_U(T, NF, VD) hipLaunchKernelGGL((k_update1d<T, NF, VD>), g, b, 0, st, d_S, d_rcell)
    if (fp64) {
        if (S.vdim == 1 && S.nf == 1) KBG_U(double, 1, 1);
        else if (S.vdim == 1 && S.nf == 2) KBG_U(double, 2, 1);
        else if (S.vdim == 3 && S.nf == 1) KBG_U(double, 1, 3);
        else return hipErrorInvalidValue;
    } else {
        if (S.vdim == 1 && S.nf == 1) KBG_U(float, 1, 1);
        else if (S.vdim == 1 && S.nf == 2) KBG_U(float, 2, 1);
        else if (S.vdim == 3 && S.nf == 1) KBG_U(float, 1, 3);
        else return hipErrorInvalidValue;
    }
#undef KBG_U
    return hipGetLastError();
}

hipError_t launch_resid1d(const Solve1D& S, const double* d_rcell, int nrows, int stride_cols, double* d_res, hipStream_t st) {
    hipLaunchKernelGGL(k_resid, dim3(1), dim3(256), 0, st, d_rcell, nrows, S.nw, stride_cols, double(S.nx), d_res);
    return hipGetLastError();
}

}  // namespace kbg
