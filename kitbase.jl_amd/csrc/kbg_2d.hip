// kbg_2d.hip -- 2D2F2V discrete-velocity kernels for gfx950 (structured quadrilateral mesh).
//
// One workgroup of 256 threads per cell or face, threads striding the velocity grid (u fastest, as
// KitBase lays out VSpace2D).  Face fluxes are computed in the face's local frame (vn, vt) with the
// face normal and length taken from the mesh (geo_struct.jl:64-175 computes them with a +1e-6 in
// the normal's length, which the host passes through unchanged), then rotated back.
#include "kbg_device.h"
#include "kbg_types.h"

namespace kbg {

// reconstruct!  solver_reconstruction.jl:475-620: one-sided differences on the boundary rows
// (reconstruct2!, reconstruction.jl:48-50), limited three-point slopes inside.
template <class T> __global__ __launch_bounds__(256) void k_recon2d(const Solve2D* __restrict__ pS) {
    const Solve2D& S = *pS;
    const int nx = S.nx, ny = S.ny, nv = S.nv, RS = 2 * nv;
    const int c = blockIdx.x, i = c % nx + 1, j = c / nx + 1;
    const T* f = (const T*)S.f;
    T* sf = (T*)S.sf;
    for (int dir = 0; dir < 2; dir++) {
        const int n = dir == 0 ? nx : ny;
        const int p = dir == 0 ? i : j;
        const int step = dir == 0 ? 1 : nx;
        const double* d = dir == 0 ? S.dx : S.dy;
        T* out = sf + (size_t)c * 2 * RS + (size_t)dir * RS;
        double* sw = S.sw + (size_t)c * 8 + dir * 4;
        if (p == 1 || p == n) {
            const int a = p == 1 ? c : c - step, e = p == 1 ? c + step : c;
            const double D = 0.5 * (d[a] + d[e]);
            if (threadIdx.x < 4) sw[threadIdx.x] = (S.w[(size_t)e * 4 + threadIdx.x] - S.w[(size_t)a * 4 + threadIdx.x]) / D;
            const T r = T(1.0 / D);
            for (int q = threadIdx.x; q < RS; q += blockDim.x) out[q] = (f[(size_t)e * RS + q] - f[(size_t)a * RS + q]) * r;
        } else {
            const int a = c - step, e = c + step;
            const double dL = 0.5 * (d[a] + d[c]), dR = 0.5 * (d[c] + d[e]);
            if (threadIdx.x < 4) {
                const int m = threadIdx.x;
                sw[m] = limiter<double>(S.limiter, (S.w[(size_t)c * 4 + m] - S.w[(size_t)a * 4 + m]) / dL,
                                        (S.w[(size_t)e * 4 + m] - S.w[(size_t)c * 4 + m]) / dR);
            }
            const T rL = T(1.0 / dL), rR = T(1.0 / dR);
            for (int q = threadIdx.x; q < RS; q += blockDim.x) {
                const T fa = f[(size_t)a * RS + q], fc = f[(size_t)c * RS + q], fe = f[(size_t)e * RS + q];
                out[q] = limf<T>(S.limiter, (fc - fa) * rL, (fe - fc) * rR);
            }
        }
    }
}

// evolve!  solver_evolution.jl:608-679: interior x-faces (i = 2..nx) then y-faces (j = 2..ny).
// Wrappers flux_kfvs.jl:98-135 and flux_kcu.jl:149-191; UGKS uses the KCU wrapper's framing with
// flux_ugks.jl:121-253.  The KCU / UGKS face chain (FP64) comes from the records of k_face2d.
template <class T, int FLUX> __global__ __launch_bounds__(256) void k_evolve2d(const Solve2D* __restrict__ pS) {
    __shared__ double red_s[4 * 10];
    const Solve2D& S = *pS;
    const int nx = S.nx, ny = S.ny, nv = S.nv, RS = 2 * nv;
    const int nxf = (nx - 1) * ny;
    int b = blockIdx.x, dir, fi, cL, cR;
    if (b < nxf) {
        dir = 0;
        const int i = b % (nx - 1) + 2, j = b / (nx - 1) + 1;
        fi = (i - 1) + (nx + 1) * (j - 1);
        cL = (i - 2) + nx * (j - 1);
        cR = cL + 1;
    } else {
        b -= nxf;
        dir = 1;
        const int i = b % nx + 1, j = b / nx + 2;
        fi = (i - 1) + nx * (j - 1);
        cL = (i - 1) + nx * (j - 2);
        cR = cL + nx;
    }
    const double* geo = (dir == 0 ? S.xgeo : S.ygeo) + (size_t)fi * 3;
    const double n1 = geo[0], n2 = geo[1], len = geo[2];
    const double* d = dir == 0 ? S.dx : S.dy;
    const double dxL = 0.5 * d[cL], dxR = 0.5 * d[cR];
    const T dxLT = T(dxL), dxRT = T(dxR);
    double* fwout = (dir == 0 ? S.xfw : S.yfw) + (size_t)fi * 4;
    T* ffh = (T*)(dir == 0 ? S.xff : S.yff) + (size_t)fi * RS;
    T* ffb = ffh + nv;
    const T* f = (const T*)S.f;
    const T* sf = (const T*)S.sf;
    const T* hL = f + (size_t)cL * RS;
    const T* bL = hL + nv;
    const T* hR = f + (size_t)cR * RS;
    const T* bR = hR + nv;
    const T* shL = sf + (size_t)cL * 2 * RS + (size_t)dir * RS;
    const T* sbL = shL + nv;
    const T* shR = sf + (size_t)cR * 2 * RS + (size_t)dir * RS;
    const T* sbR = shR + nv;
    const T* U = (const T*)S.u;
    const T* V = (const T*)S.v;
    const T* WT = (const T*)S.wt;
    const T n1T = T(n1), n2T = T(n2), lenT = T(len);
    const double dt = S.dt, K = S.K;
    const T dtT = T(dt), hdt2 = T(0.5 * dt * dt);
    double fw[4];
    if (FLUX == KFVS) {
        KSum<T> a[5];
        for (int iv = threadIdx.x; iv < nv; iv += blockDim.x) {
            const T x = U[iv] * n1T + V[iv] * n2T, y = V[iv] * n1T - U[iv] * n2T, wq = WT[iv];
            const bool dd = x >= T(0);
            const T h = dd ? hL[iv] + shL[iv] * dxLT : hR[iv] - shR[iv] * dxRT;
            const T bb = dd ? bL[iv] + sbL[iv] * dxLT : bR[iv] - sbR[iv] * dxRT;
            const T sh = dd ? shL[iv] : shR[iv];
            const T sb = dd ? sbL[iv] : sbR[iv];
            const T fh = (dtT * x * h - hdt2 * (x * x) * sh) * lenT;
            const T fb = (dtT * x * bb - hdt2 * (x * x) * sb) * lenT;
            ffh[iv] = fh;
            ffb[iv] = fb;
            a[0].add(wq * fh);
            a[1].add(x * wq * fh);
            a[2].add(y * wq * fh);
            a[3].add((x * x + y * y) * wq * fh);
            a[4].add(wq * fb);
        }
        double r[5] = {a[0].val(), a[1].val(), a[2].val(), a[3].val(), a[4].val()};
        block_allreduce256<5>(r, red_s);
        fw[0] = r[0];
        fw[1] = r[1];
        fw[2] = r[2];
        fw[3] = 0.5 * (r[3] + r[4]);
    } else {
        // the face chain was evaluated once per face by k_face2d
        const double* rec = (dir == 0 ? S.xcoef : S.ycoef) + (size_t)fi * kFaceCoef;
        if (FLUX == KCU) {
            const KcuFace<4>& F = *(const KcuFace<4>*)rec;
            const T A = T(F.prim[0] * (F.prim[3] / kPi)), lam = T(F.prim[3]), Ui = T(F.prim[1]), Vi = T(F.prim[2]);
            const T Mt1 = T(F.Mt1), Mt2 = T(F.Mt2), kb = T(K / (2.0 * F.prim[3]));
            KSum<T> a[5];
            for (int iv = threadIdx.x; iv < nv; iv += blockDim.x) {
                const T x = U[iv] * n1T + V[iv] * n2T, y = V[iv] * n1T - U[iv] * n2T, wq = WT[iv];
                const bool dd = x >= T(0);
                const T h = dd ? hL[iv] + shL[iv] * dxLT : hR[iv] - shR[iv] * dxRT;
                const T bb = dd ? bL[iv] + sbL[iv] * dxLT : bR[iv] - sbR[iv] * dxRT;
                const T cu = x - Ui, cv = y - Vi;
                const T H = A * texp(-lam * (cu * cu + cv * cv));
                const T B = H * kb;
                a[0].add(wq * x * h);
                a[1].add(wq * (x * x) * h);
                a[2].add(wq * y * x * h);
                a[3].add(wq * x * (x * x + y * y) * h);
                a[4].add(wq * x * bb);
                ffh[iv] = (Mt1 * x * H + Mt2 * x * h) * lenT;
                ffb[iv] = (Mt1 * x * B + Mt2 * x * bb) * lenT;
            }
            double r[5] = {a[0].val(), a[1].val(), a[2].val(), a[3].val(), a[4].val()};
            block_allreduce256<5>(r, red_s);
            fw[0] = F.fw0[0] * len + F.Mt2 * r[0] * len;
            fw[1] = F.fw0[1] * len + F.Mt2 * r[1] * len;
            fw[2] = F.fw0[2] * len + F.Mt2 * r[2] * len;
            fw[3] = F.fw0[3] * len + F.Mt2 * 0.5 * (r[3] + r[4]) * len;
        } else {
            const UgksFace2& F = *(const UgksFace2*)rec;
            const T A = T(F.prim[0] * (F.prim[3] / kPi)), lam = T(F.prim[3]), Ui = T(F.prim[1]), Vi = T(F.prim[2]);
            const T kb = T(K / (2.0 * F.prim[3])), Mxi2 = T(F.Mxi2);
            const T Mt0 = T(F.Mt[0]), Mt1 = T(F.Mt[1]), Mt2 = T(F.Mt[2]), Mt3 = T(F.Mt[3]), Mt4 = T(F.Mt[4]);
            T aL[4], aR[4], aT[4];
            for (int m = 0; m < 4; m++) {
                aL[m] = T(F.aL[m]);
                aR[m] = T(F.aR[m]);
                aT[m] = T(F.aT[m]);
            }
            KSum<T> a[8];
            for (int iv = threadIdx.x; iv < nv; iv += blockDim.x) {
                const T x = U[iv] * n1T + V[iv] * n2T, y = V[iv] * n1T - U[iv] * n2T, wq = WT[iv];
                const bool dd = x >= T(0);
                const T uu = x * x, q2 = uu + y * y;
                const T h = dd ? hL[iv] + shL[iv] * dxLT : hR[iv] - shR[iv] * dxRT;
                const T bb = dd ? bL[iv] + sbL[iv] * dxLT : bR[iv] - sbR[iv] * dxRT;
                const T sh = dd ? shL[iv] : shR[iv];
                const T sb = dd ? sbL[iv] : sbR[iv];
                const T cu = x - Ui, cv = y - Vi;
                const T H = A * texp(-lam * (cu * cu + cv * cv));
                const T B = H * kb;
                a[0].add(wq * x * h);
                a[1].add(wq * uu * sh);
                a[2].add(wq * uu * h);
                a[3].add(wq * (uu * x) * sh);
                a[4].add(wq * y * x * h);
                a[5].add(wq * y * uu * sh);
                a[6].add(wq * x * q2 * h + wq * x * bb);
                a[7].add(wq * uu * q2 * sh + wq * uu * sb);
                const T* s = dd ? aL : aR;
                const T pH = s[0] * H + s[1] * x * H + s[2] * y * H + T(0.5) * s[3] * (q2 * H + B);
                const T pTH = aT[0] * H + aT[1] * x * H + aT[2] * y * H + T(0.5) * aT[3] * (q2 * H + B);
                const T pB = s[0] * B + s[1] * x * B + s[2] * y * B + T(0.5) * s[3] * (q2 * B + Mxi2 * H);
                const T pTB = aT[0] * B + aT[1] * x * B + aT[2] * y * B + T(0.5) * aT[3] * (q2 * B + Mxi2 * H);
                ffh[iv] = (Mt0 * x * H + Mt1 * uu * pH + Mt2 * x * pTH + Mt3 * x * h - Mt4 * uu * sh) * lenT;
                ffb[iv] = (Mt0 * x * B + Mt1 * uu * pB + Mt2 * x * pTB + Mt3 * x * bb - Mt4 * uu * sb) * lenT;
            }
            double r[8];
            for (int m = 0; m < 8; m++) r[m] = a[m].val();
            block_allreduce256<8>(r, red_s);
            fw[0] = (F.fw0[0] + (F.Mt[3] * r[0] - F.Mt[4] * r[1])) * len;
            fw[1] = (F.fw0[1] + (F.Mt[3] * r[2] - F.Mt[4] * r[3])) * len;
            fw[2] = (F.fw0[2] + (F.Mt[3] * r[4] - F.Mt[4] * r[5])) * len;
            fw[3] = (F.fw0[3] + (F.Mt[3] * 0.5 * r[6] - F.Mt[4] * 0.5 * r[7])) * len;
        }
    }
    if (threadIdx.x == 0) {
        double g[4];
        global_frame4(fw, n1, n2, g);
        for (int m = 0; m < 4; m++) fwout[m] = g[m];
    }
}

// Maxwell diffuse walls: evolve_boundary.jl:52-178 -> flux_boundary.jl:190-220.
// Blocks: [0, ny) left, [ny, 2ny) right, [2ny, 2ny+nx) bottom, [2ny+nx, 2ny+2nx) top.
template <class T> __global__ __launch_bounds__(256) void k_wall2d(const Solve2D* __restrict__ pS) {
    __shared__ double red_s[4 * 8];
    const Solve2D& S = *pS;
    const int nx = S.nx, ny = S.ny, nv = S.nv, RS = 2 * nv;
    int b = blockIdx.x, side, fi, c, rot;
    if (b < ny) {
        side = 0; const int j = b + 1;
        fi = (nx + 1) * (j - 1); c = nx * (j - 1); rot = 1;
    } else if (b < 2 * ny) {
        side = 1; const int j = b - ny + 1;
        fi = nx + (nx + 1) * (j - 1); c = (nx - 1) + nx * (j - 1); rot = -1;
    } else if (b < 2 * ny + nx) {
        side = 2; const int i = b - 2 * ny + 1;
        fi = i - 1; c = i - 1; rot = 1;
    } else {
        side = 3; const int i = b - 2 * ny - nx + 1;
        fi = (i - 1) + nx * ny; c = (i - 1) + nx * (ny - 1); rot = -1;
    }
    if (S.bc[side] != 1) return;
    const bool xs = side < 2;
    const double* geo = (xs ? S.xgeo : S.ygeo) + (size_t)fi * 3;
    const double n1 = geo[0], n2 = geo[1], len = geo[2];
    const double* wp = (xs ? S.xwall : S.ywall) + (size_t)fi * 4;
    double bcl[4];
    local_frame4(wp, n1, n2, bcl);
    double* fwout = (xs ? S.xfw : S.yfw) + (size_t)fi * 4;
    T* ffh = (T*)(xs ? S.xff : S.yff) + (size_t)fi * RS;
    T* ffb = ffh + nv;
    const T* h = (const T*)S.f + (size_t)c * RS;
    const T* bb = h + nv;
    const T* U = (const T*)S.u;
    const T* V = (const T*)S.v;
    const T* WT = (const T*)S.wt;
    const T n1T = T(n1), n2T = T(n2), rotT = T(rot);
    const T Uw = T(bcl[1]), Vw = T(bcl[2]), lw = T(bcl[3]);
    const double dt = S.dt, K = S.K;
    KSum<T> a0, a1;
    for (int iv = threadIdx.x; iv < nv; iv += blockDim.x) {
        const T x = U[iv] * n1T + V[iv] * n2T, y = V[iv] * n1T - U[iv] * n2T;
        const bool d = x * rotT >= T(0);
        if (!d) a0.add(WT[iv] * x * h[iv]);
        else {
            const T cu = x - Uw, cv = y - Vw;
            a1.add(WT[iv] * x * texp(-lw * (cu * cu + cv * cv)));
        }
    }
    double r[2] = {a0.val(), a1.val()};
    block_allreduce256<2>(r, red_s);
    const double rho = -r[0] / ((bcl[3] / kPi) * r[1]);
    const T A = T(rho * (bcl[3] / kPi)), kb = T(K / (2.0 * bcl[3])), sc = T(len * dt);
    KSum<T> m[5];
    for (int iv = threadIdx.x; iv < nv; iv += blockDim.x) {
        const T x = U[iv] * n1T + V[iv] * n2T, y = V[iv] * n1T - U[iv] * n2T, wq = WT[iv];
        const bool d = x * rotT >= T(0);
        const T cu = x - Uw, cv = y - Vw;
        const T H = A * texp(-lw * (cu * cu + cv * cv));
        const T hW = d ? H : h[iv];
        const T bW = d ? H * kb : bb[iv];
        m[0].add(wq * x * hW);
        m[1].add(wq * (x * x) * hW);
        m[2].add(wq * y * (hW * x));
        m[3].add(wq * x * (hW * (x * x + y * y)));
        m[4].add(wq * x * bW);
        ffh[iv] = x * hW * sc;
        ffb[iv] = x * bW * sc;
    }
    double s[5] = {m[0].val(), m[1].val(), m[2].val(), m[3].val(), m[4].val()};
    block_allreduce256<5>(s, red_s);
    if (threadIdx.x == 0) {
        double fw[4] = {s[0] * len * dt, s[1] * len * dt, s[2] * len * dt, (0.5 * s[3] + 0.5 * s[4]) * len * dt}, g[4];
        global_frame4(fw, n1, n2, g);
        for (int k = 0; k < 4; k++) fwout[k] = g[k];
    }
}

// update!  solver_update.jl:193-246 + update_boundary.jl:172-306 (coll is forwarded in 2D);
// step! step_2f.jl:179-249.  One block per cell; cells that are not stepped only clear their
// residual row.
template <class T> __global__ __launch_bounds__(256) void k_update2d(const Solve2D* __restrict__ pS, double* __restrict__ rcell) {
    __shared__ double red_s[4 * 8];
    const Solve2D& S = *pS;
    const int nx = S.nx, ny = S.ny, nv = S.nv, RS = 2 * nv;
    const int c = blockIdx.x, i = c % nx + 1, j = c / nx + 1;
    const bool inner = i > 1 && i < nx && j > 1 && j < ny;
    const bool mid = i > 1 && i < nx;
    const bool stepped = inner || (i == 1 && S.bc[0] != 0) || (i == nx && S.bc[1] != 0) ||
                         (j == 1 && mid && S.bc[2] != 0) || (j == ny && mid && S.bc[3] != 0);
    if (!stepped) {
        if (rcell && threadIdx.x < 16) rcell[(size_t)c * 16 + threadIdx.x] = 0.0;
        return;
    }
    const int L = (i - 1) + (nx + 1) * (j - 1), R = L + 1;
    const int D = (i - 1) + nx * (j - 1), Up = D + nx;
    double* w = S.w + (size_t)c * 4;
    double* prim = S.prim + (size_t)c * 4;
    T* h = (T*)S.f + (size_t)c * RS;
    T* b = h + nv;
    const T* xff = (const T*)S.xff;
    const T* yff = (const T*)S.yff;
    const T* fhL = xff + (size_t)L * RS;
    const T* fhR = xff + (size_t)R * RS;
    const T* fhD = yff + (size_t)D * RS;
    const T* fhU = yff + (size_t)Up * RS;
    const double *fwL = S.xfw + (size_t)L * 4, *fwR = S.xfw + (size_t)R * 4;
    const double *fwD = S.yfw + (size_t)D * 4, *fwU = S.yfw + (size_t)Up * 4;
    const T* U = (const T*)S.u;
    const T* V = (const T*)S.v;
    const T* WT = (const T*)S.wt;
    const double K = S.K, dt = S.dt, ds = S.dx[c] * S.dy[c];
    const bool shak = S.coll == SHAKHOV;
    double po[4], wo[4];
    for (int m = 0; m < 4; m++) {
        po[m] = prim[m];
        wo[m] = w[m];
    }
    double q1 = 0, q2 = 0;
    if (shak) {
        KSum<T> a[4];
        const T Uo = T(po[1]), Vo = T(po[2]);
        for (int iv = threadIdx.x; iv < nv; iv += blockDim.x) {
            const T cu = U[iv] - Uo, cv = V[iv] - Vo, c2 = cu * cu + cv * cv, wq = WT[iv];
            a[0].add(wq * cu * c2 * h[iv]);
            a[1].add(wq * cu * b[iv]);
            a[2].add(wq * cv * c2 * h[iv]);
            a[3].add(wq * cv * b[iv]);
        }
        double r[4] = {a[0].val(), a[1].val(), a[2].val(), a[3].val()};
        block_allreduce256<4>(r, red_s);
        q1 = 0.5 * (r[0] + r[1]);
        q2 = 0.5 * (r[2] + r[3]);
    }
    double wn[4], pn[4];
    for (int m = 0; m < 4; m++) wn[m] = wo[m] + (fwL[m] - fwR[m] + fwD[m] - fwU[m]) / ds;
    conserve_prim<4>(wn, S.gam, pn);
    if (threadIdx.x == 0) flag_health(S.health, prim_health<4>(pn));
    const double tau = vhs_tau(pn[0], pn[3], S.mu, S.om);
    const double r = dt / tau;
    const T rT = T(r), inv = T(1.0 / (1.0 + r)), rds = T(1.0 / ds);
    const T An = T(pn[0] * (pn[3] / kPi)), ln = T(pn[3]), Un = T(pn[1]), Vn = T(pn[2]), kbn = T(K / (2.0 * pn[3]));
    T preu = 0, prev = 0, Ao = 0, lo = 0, Uo = 0, Vo = 0, kbo = 0;
    if (shak) {
        const double pre = 0.8 * (1.0 - S.Pr) * (po[3] * po[3]) / po[0];
        preu = T(pre * q1);
        prev = T(pre * q2);
        Ao = T(po[0] * (po[3] / kPi));
        lo = T(po[3]);
        Uo = T(po[1]);
        Vo = T(po[2]);
        kbo = T(K / (2.0 * po[3]));
    }
    for (int iv = threadIdx.x; iv < nv; iv += blockDim.x) {
        const T cu = U[iv] - Un, cv = V[iv] - Vn;
        T MH = An * texp(-ln * (cu * cu + cv * cv));
        T MB = MH * kbn;
        if (shak) {
            const T ou = U[iv] - Uo, ov = V[iv] - Vo, c2 = ou * ou + ov * ov;
            const T Mo = Ao * texp(-lo * c2);
            const T g = preu * ou + prev * ov;
            MH += g * (T(2) * lo * c2 + T(K) - T(5)) * Mo;
            MB += g * (T(2) * lo * c2 + T(K) - T(3)) * (Mo * kbo);
        }
        const int o = nv + iv;
        h[iv] = (h[iv] + (fhL[iv] - fhR[iv] + fhD[iv] - fhU[iv]) * rds + rT * MH) * inv;
        b[iv] = (b[iv] + (fhL[o] - fhR[o] + fhD[o] - fhU[o]) * rds + rT * MB) * inv;
    }
    if (threadIdx.x < 4) {
        const int m = threadIdx.x;
        w[m] = wn[m];
        prim[m] = pn[m];
        if (rcell) {
            double* row = rcell + (size_t)c * 16;
            row[(inner ? 0 : 2) * 4 + m] = (wn[m] - wo[m]) * (wn[m] - wo[m]);
            row[(inner ? 1 : 3) * 4 + m] = fabs(wn[m]);
            row[(inner ? 2 : 0) * 4 + m] = 0.0;
            row[(inner ? 3 : 1) * 4 + m] = 0.0;
        }
    }
}

__global__ void k_resid2(const double* __restrict__ rows, int nrows, double ncell, double* __restrict__ out) {
    __shared__ double part[256 * 16];
    for (int col = 0; col < 16; col++) {
        double s = 0.0;
        for (int r = threadIdx.x; r < nrows; r += 256) s += rows[(size_t)r * 16 + col];
        part[threadIdx.x * 16 + col] = s;
    }
    __syncthreads();
    __shared__ double tot[16];
    if (threadIdx.x < 16) {
        double s = 0.0;
        for (int t = 0; t < 256; t++) s += part[t * 16 + threadIdx.x];
        tot[threadIdx.x] = s;
    }
    __syncthreads();
    if (threadIdx.x < 4) {
        const int m = threadIdx.x;
        out[m] = sqrt(tot[m] * ncell) / (tot[4 + m] + 1.e-7) + sqrt(tot[8 + m] * 2.0) / (tot[12 + m] + 1.e-7);
    }
}

// timestep (solver_main.jl:240-342): tmax = max over the cells of (max(u1, |U|) + c) / dx, plus
// (max(v1, |V|) + c) / dy in 2D, with the sound speed c = (gamma / (2 lambda))^(1/2)
// (theory_macro.jl:421).  A maximum does not depend on the order it is taken in, so the result is
// the same on every run.
__global__ __launch_bounds__(256) void k_cfl(const double* __restrict__ prim, int nw, int k0, int n, const double* __restrict__ dx,
                                             const double* __restrict__ dy, double u1, double v1, double gam,
                                             double* __restrict__ out) {
    __shared__ double part[256];
    double m = 0.0;
    for (int k = threadIdx.x; k < n; k += 256) {
        const double* p = prim + (size_t)(k0 + k) * nw;
        const double sos = sqrt(0.5 * gam / p[nw - 1]);
        double t = (fmax(u1, fabs(p[1])) + sos) / dx[k0 + k];
        if (dy) t += (fmax(v1, fabs(p[2])) + sos) / dy[k0 + k];
        m = fmax(m, t);
    }
    part[threadIdx.x] = m;
    __syncthreads();
    for (int h = 128; h >= 1; h >>= 1) {
        if (threadIdx.x < h) part[threadIdx.x] = fmax(part[threadIdx.x], part[threadIdx.x + h]);
        __syncthreads();
    }
    if (threadIdx.x == 0) out[0] = part[0];
}

hipError_t launch_cfl(const double* d_prim, int nw, int k0, int n, const double* d_dx, const double* d_dy, double u1, double v1,
                      double gam, double* d_out, hipStream_t st) {
    hipLaunchKernelGGL(k_cfl, dim3(1), dim3(256), 0, st, d_prim, nw, k0, n, d_dx, d_dy, u1, v1, gam, d_out);
    return hipGetLastError();
}

hipError_t launch_recon2d(bool fp64, const Solve2D& S, const Solve2D* d_S, hipStream_t st) {
    if (!S.interp2) return hipSuccess;
    const int n = S.nx * S.ny;
    if (fp64) hipLaunchKernelGGL(k_recon2d<double>, dim3(n), dim3(256), 0, st, d_S);
    else hipLaunchKernelGGL(k_recon2d<float>, dim3(n), dim3(256), 0, st, d_S);
    return hipGetLastError();
}

hipError_t launch_evolve2d(bool fp64, const Solve2D& S, const Solve2D* d_S, hipStream_t st) {
    const int nfaces = (S.nx - 1) * S.ny + S.nx * (S.ny - 1);
    if (nfaces > 0) {
        hipError_t e0 = launch_face2d(false, S.flux, d_S, nullptr, 1, nfaces, st);
        if (e0 != hipSuccess) return e0;
        dim3 g(nfaces), b(256);
        if (fp64) {
            if (S.flux == KFVS) hipLaunchKernelGGL((k_evolve2d<double, KFVS>), g, b, 0, st, d_S);
            else if (S.flux == KCU) hipLaunchKernelGGL((k_evolve2d<double, KCU>), g, b, 0, st, d_S);
            else hipLaunchKernelGGL((k_evolve2d<double, UGKS>), g, b, 0, st, d_S);
        } else {
            if (S.flux == KFVS) hipLaunchKernelGGL((k_evolve2d<float, KFVS>), g, b, 0, st, d_S);
            else if (S.flux == KCU) hipLaunchKernelGGL((k_evolve2d<float, KCU>), g, b, 0, st, d_S);
            else hipLaunchKernelGGL((k_evolve2d<float, UGKS>), g, b, 0, st, d_S);
        }
        hipError_t e = hipGetLastError();
        if (e != hipSuccess) return e;
    }
    const int nw = 2 * (S.nx + S.ny);
    if (fp64) hipLaunchKernelGGL(k_wall2d<double>, dim3(nw), dim3(256), 0, st, d_S);
    else hipLaunchKernelGGL(k_wall2d<float>, dim3(nw), dim3(256), 0, st, d_S);
    return hipGetLastError();
}

hipError_t launch_update2d(bool fp64, const Solve2D& S, const Solve2D* d_S, double* d_rcell, hipStream_t st) {
    const int n = S.nx * S.ny;
    if (fp64) hipLaunchKernelGGL(k_update2d<double>, dim3(n), dim3(256), 0, st, d_S, d_rcell);
    else hipLaunchKernelGGL(k_update2d<float>, dim3(n), dim3(256), 0, st, d_S, d_rcell);
    return hipGetLastError();
}

hipError_t launch_resid2d(const Solve2D& S, const double* d_rcell, double* d_res, hipStream_t st) {
    hipLaunchKernelGGL(k_resid2, dim3(1), dim3(256), 0, st, d_rcell, S.nx * S.ny, double(S.nx) * double(S.ny), d_res);
    return hipGetLastError();
}

}  // namespace kbg
