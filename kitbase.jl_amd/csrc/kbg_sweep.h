// kbg_sweep.h -- the fused 1D time step (reconstruct! + evolve! + update! in one pass) for gfx950.
//
// Layout.  A swept solve stores its distributions cell-fastest: f[q][v][k] at ((q nv + v) np + k)
// for distribution q (h, b), velocity v and storage cell k, each velocity row padded to np cells
// (a multiple of 32, so every row starts on a 128-byte line).
//
// Work decomposition.  One workgroup per segment of up to 60 consecutive cells of one solve.  Its NW
// wavefronts all hold the same 64 storage cells -- lane L is cell k0 + L: the segment plus two halo
// cells on each side -- and split the velocity grid: wavefront w takes velocities
// [w VW, (w + 1) VW).  Neighbouring cells are neighbouring lanes, reached with DPP whole-wavefront
// shifts, so the stencil needs no LDS and no extra loads.
//
//   pass 1, per velocity: one coalesced load of the 64 cells' f, the limited slope (reconstruct!),
//     the upwind flux of the lane's left face (evolve!), its moment sums in FP64, and the transport
//     update G = f + (ff_left - ff_right) / dx, kept in a register;
//   the wavefronts' moment sums are added through LDS in a fixed order;
//   every lane runs its own cell's FP64 chain (step!): w update, conserve_prim, tau, equilibrium and
//     Shakhov coefficients -- 64 cells per pass;
//   pass 2, per velocity: relax G toward the new equilibrium and store f^{n+1}.
//
// f^n is read once and f^{n+1} written once (plus the 4-cell halo per segment): the compulsory HBM
// traffic of a time step.  The per-cell FP64 work is lane-parallel over the cells.
//
// Reference semantics (KitBase.jl v0.9.31):
//   slopes    solver_reconstruction.jl:56-153 (reconstruct3! reconstruction.jl:136-151, limiters
//             recon_limiter.jl:10-105)
//   faces     solver_evolution.jl:39-89; flux_kfvs.jl:238-340; flux_kcu.jl:241-303,463-530;
//             UGKS flux_ugks.jl:8-114 with the reconstructed wL = w + dx/2 sw (solver_evolution.jl:285-312)
//   walls     evolve_boundary.jl:4-47 (boundary-face buffers written by k_wall1d)
//   cells     solver_update.jl:61-93, update_boundary.jl:6-61; step_1f.jl:6-55, step_2f.jl:6-64
#pragma once

#include <algorithm>

#include "kbg_device.h"
#include "kbg_types.h"

// alternate the pack order between steps (1) or keep it (0)
#ifndef KBG_SWEEP_ZIGZAG
#define KBG_SWEEP_ZIGZAG 1
#endif
// cache policy of the f^{n+1} stores (0: default)
#ifndef KBG_SWEEP_STORE_AUX
#define KBG_SWEEP_STORE_AUX 0
#endif
#ifndef KBG_SWEEP_WPE_F32_1F
#define KBG_SWEEP_WPE_F32_1F 4
#endif

namespace kbg {

constexpr int kSweepLanes = 64;  // storage cells held by a wavefront
constexpr int kSweepOut = 60;    // cells a segment updates (lanes 2 .. 61)
constexpr int kSweepOut2 = 56;   // cells a two-step segment updates (lanes 4 .. 59)

// DPP whole-wavefront shifts: shr1(x)[L] = x[L - 1], shl1(x)[L] = x[L + 1] (0 beyond the ends)
KBG_D float shr1(float x) {
    return __builtin_bit_cast(float, __builtin_amdgcn_mov_dpp(__builtin_bit_cast(int, x), 0x138, 0xf, 0xf, true));
}
KBG_D float shl1(float x) {
    return __builtin_bit_cast(float, __builtin_amdgcn_mov_dpp(__builtin_bit_cast(int, x), 0x130, 0xf, 0xf, true));
}
KBG_D double shr1(double x) {
    const long long b = __builtin_bit_cast(long long, x);
    const int lo = __builtin_amdgcn_mov_dpp((int)b, 0x138, 0xf, 0xf, true);
    const int hi = __builtin_amdgcn_mov_dpp((int)(b >> 32), 0x138, 0xf, 0xf, true);
    return __builtin_bit_cast(double, ((long long)hi << 32) | (long long)(unsigned)lo);
}
KBG_D double shl1(double x) {
    const long long b = __builtin_bit_cast(long long, x);
    const int lo = __builtin_amdgcn_mov_dpp((int)b, 0x130, 0xf, 0xf, true);
    const int hi = __builtin_amdgcn_mov_dpp((int)(b >> 32), 0x130, 0xf, 0xf, true);
    return __builtin_bit_cast(double, ((long long)hi << 32) | (long long)(unsigned)lo);
}

// exp2 in the storage precision (the FP32 path uses the hardware exp2 directly)
KBG_D float texp2(float x) { return __builtin_amdgcn_exp2f(x); }
KBG_D double texp2(double x) { return exp2(x); }
// reciprocal in the storage precision
KBG_D float trcp(float x) { return __builtin_amdgcn_rcpf(x); }
KBG_D double trcp(double x) { return 1.0 / x; }

constexpr double kLog2e = 1.4426950408889634074;

// Per-velocity constants of a swept solve in chunks of 8 velocities (structure of arrays, so a
// chunk's values of one kind arrive in one scalar load), rebuilt on the host when dt changes.
// Velocities past nv have zero weights and up = 1: they add nothing and are never stored.
constexpr int kVChunk = 8;
template <class T> struct alignas(64) VChunk {
    T wk[5][kVChunk];         // quadrature weight x u^k, k = 0..4
    T u[kVChunk], uu[kVChunk];
    T cu[kVChunk], cs[kVChunk];  // dt u, -dt^2 u^2 / 2 (KFVS)
    T cs2[kVChunk];           // -dt^2 u^2 (KFVS with half slopes)
    T du;                     // node spacing when the chunk's nodes are equally spaced, else 0
    int upmask;               // bit j: u >= 0, the flux takes the left state (heaviside, math_general.jl:13)
};
constexpr int vchunks(int nv) { return nv / kVChunk + 2; }

// The FP32 sweep stores each distribution as its departure from M~(prim), the Maxwellian of the
// cell's primitive variables (conserve_prim of the w buffer paired with the f buffer) evaluated in
// FP32: the departure is small where f is near
// equilibrium, so the FP32 rounding of the stored state (and of the per-step update, done on the
// departure) shrinks with it.  Every kernel that reads or writes that storage evaluates M~ from the
// same FP32 coefficients with the same operations in the same order, so a stored f - M~ reads back
// against bit-identical M~.
struct MaxwCoef {
    float A, nl2, U, kb;  // rho (lambda / pi)^(1/2), -lambda log2(e), U, K / (2 lambda)
};
// in FP32 from the FP64 primitive variables (one FP32 square root and reciprocal)
KBG_D MaxwCoef maxw_coef(const double* p, double K) {
    const float lam = float(p[2]);
    return MaxwCoef{float(p[0]) * __builtin_amdgcn_sqrtf(lam * float(1.0 / kPi)), float(-p[2] * 1.4426950408889634074), float(p[1]),
                    float(K) * __builtin_amdgcn_rcpf(2.f * lam)};
}
// M~ at the 8 nodes of a velocity chunk: one exp2 per node, or on equally spaced nodes the product
// recurrence M(u + du) = M(u) R, R <- R Q with R = exp(-lambda du (2 (u - U) + du)), Q = exp(-2 lambda du^2).
// Compiled without floating-point contraction: a product left unrounded inside a fused multiply-add
// at one use and rounded at another would give the same node two values.
template <class CHK> KBG_D void maxw8(CHK& ch, float A, float nl2, float U0, float (&M)[kVChunk]) {
#pragma clang fp contract(off)
    // plain loops: the pragma reaches the statements of this body, not those of a lambda in it
    const float d = ch.du;
    if (d != 0.f) {
        const float c0 = ch.u[0] - U0;
        float m = A * __builtin_amdgcn_exp2f(nl2 * (c0 * c0));
        float R = __builtin_amdgcn_exp2f((nl2 * d) * __builtin_fmaf(2.f, c0, d));
        const float Q = __builtin_amdgcn_exp2f((2.f * nl2) * (d * d));
#pragma unroll
        for (int j = 0; j < kVChunk; j++) {
            M[j] = m;
            if (j + 1 < kVChunk) {
                m *= R;
                R *= Q;
            }
        }
    } else {
#pragma unroll
        for (int j = 0; j < kVChunk; j++) {
            const float c = ch.u[j] - U0;
            M[j] = A * __builtin_amdgcn_exp2f(nl2 * (c * c));
        }
    }
}
// the b distribution's M~ K / (2 lambda): a rounded product (no contraction into a later addition)
KBG_D float maxw_b(float M, float kb) {
#pragma clang fp contract(off)
    return M * kb;
}
// the same value at node j of the chunk alone
template <class CHK> KBG_D float maxw_at(CHK& ch, float A, float nl2, float U0, int j) {
#pragma clang fp contract(off)
    const float d = ch.du;
    if (d != 0.f) {
        const float c0 = ch.u[0] - U0;
        float m = A * __builtin_amdgcn_exp2f(nl2 * (c0 * c0));
        float R = __builtin_amdgcn_exp2f((nl2 * d) * __builtin_fmaf(2.f, c0, d));
        const float Q = __builtin_amdgcn_exp2f((2.f * nl2) * (d * d));
        for (int jj = 0; jj < j; jj++) {
            m *= R;
            R *= Q;
        }
        return m;
    }
    const float c = ch.u[j] - U0;
    return A * __builtin_amdgcn_exp2f(nl2 * (c * c));
}

// velocities per wavefront and moment-sum counts of a kernel instance
template <class T, int NF, int FLUX> struct SweepCfg {
    static constexpr int VW = sweep_vw(sizeof(T) == 8, NF);
    static constexpr int CH = kVChunk;  // velocities loaded together (one table chunk)
    static constexpr int NRAW = FLUX == UGKS ? 6 : (NF == 2 ? 4 : 3);
    static constexpr int NR = NRAW + 1;  // + the Shakhov heat-flux sum
};

// limiter family of a kernel path: the van Leer default and first order are specialised, the others
// take the generic runtime switch
enum { LIM_VANLEER = 0, LIM_GENERIC = 1, LIM_NONE = 2 };

// HALF: return half the slope (the caller doubles the factors it multiplies the slope by)
template <class T, int LIMK, bool HALF> KBG_D T sweep_limit(int kind, T sL, T sR) {
    if constexpr (LIMK == LIM_NONE) {
        return T(0);
    } else if constexpr (LIMK == LIM_VANLEER) {
        // (sgn sL + sgn sR) |sL| |sR| / (|sL| + |sR| + 1e-7) (recon_limiter.jl:17-19): zero unless
        // sL sR > 0, and then 2 |sL sR| / (|sL| + |sR| + 1e-7) with the sign of sL
        const T p = fmax(sL * sR, T(0));
        const T r = p * trcp(fabs(sL) + fabs(sR) + T(1e-7));
        return copysign(HALF ? r : T(2) * r, sL);
    } else {
        const T r = limiter<T>(kind, sL, sR);
        return HALF ? T(0.5) * r : r;
    }
}

// face coefficients of the lane's left face, in the storage precision (KCU / UGKS)
template <class T> struct SFace {
    T A, nl2, Ui, kb, Mxi2, M0, M1, M2, M3, M4, aL0, aL1, aL2, aR0, aR1, aR2, aT0, aT1, aT2;
};

// NSTEP = 2 advances the segment two time steps in one pass (KFVS, no boundary-face buffers): the
// window holds 4 halo cells per side, step 1's f^{n+1} stays in the registers that carry G, and only
// f^{n+2} is stored -- one read and one write of f per two steps.  Cells the first step does not
// update (boundary or ghost cells) keep f^n, reloaded for the second step.
template <class T, int NF, int FLUX, int LIMK, bool SHAK, int NSTEP>
KBG_D void sweep1d(const Solve1D& S, const WorkItem& wi, int par, int lane, int wave, int NW, double* lds,
                   double* rrow) {
    static_assert(NSTEP == 1 || FLUX == KFVS, "two steps per pass are built for KFVS");
    using C = SweepCfg<T, NF, FLUX>;
    constexpr int VW = C::VW, CH = C::CH, NRAW = C::NRAW, NR = C::NR;
    const int nx = S.nx, ng = S.ng, nv = S.nv, N = nx + 2 * ng;
    const long np = S.np;
    const int in = (S.cur + par) & 1;
    // f rows as raw buffers: lane offset k * sizeof(T), row offset (q nv + v) np * sizeof(T)
    const unsigned fbytes = (unsigned)((long)NF * nv * np * sizeof(T));
    const rsrc_t fin = make_rsrc(S.f[in], fbytes), fout = make_rsrc(S.f[in ^ 1], fbytes);
    const int rowb = (int)(np * sizeof(T));
    // read-only for the kernel's lifetime: the constant address space gives scalar (SMEM) loads
    // read-only for the kernel's lifetime: the constant address space gives scalar (SMEM) loads
    using VC = const __attribute__((address_space(4))) VChunk<T>;
    VC* vt = (VC*)S.vtab;
    const double dt = S.dt, K = S.K;
    const int lk = S.limiter;

    // ---- per-lane cell
    const int k0 = wi.c0 - 1 + ng - 2 * NSTEP;  // storage cell of lane 0
    const int k = k0 + lane;
    const int kc = k < 0 ? 0 : (k >= N ? N - 1 : k);  // clamped: lanes past the ends load valid memory
    const int i = k - ng + 1;                          // physical index
    const int voff = k * (int)sizeof(T);               // f lane offset (k < 0 reads 0)
    const double d0 = S.dx[kc];
    const double dm = S.dx[kc > 0 ? kc - 1 : 0], dp = S.dx[kc < N - 1 ? kc + 1 : N - 1];
    const bool slope_on = S.interp2 && i >= S.slope_lo && i <= S.slope_hi && k >= 0 && k < N;
    // zero slope distances give exactly zero slopes outside reconstruct!'s range
    const double rL64 = slope_on ? 1.0 / (0.5 * (dm + d0)) : 0.0, rR64 = slope_on ? 1.0 / (0.5 * (d0 + dp)) : 0.0;
    const T rL = T(rL64), rR = T(rR64);
    const double dxh64 = 0.5 * d0, rdx64 = 1.0 / d0;
    const T dxh = T(dxh64), rdx = T(rdx64);
    // KFVS carries half slopes (one multiply less in the limiter): the half cell width doubles
    constexpr bool HALF = FLUX == KFVS;
    const T dxs = HALF ? T(2.0 * dxh64) : dxh;
    // cells stored by this item (lanes 2..61, or 4..59 for two steps), and, for two steps, the
    // cells the first step updates (every stepped cell with two neighbours per side in the window)
    const bool upd = NSTEP == 1 ? (lane >= 2 && lane < 2 + kSweepOut && i >= wi.c0 && i <= wi.c1)
                                : (lane >= 4 && lane < 4 + kSweepOut2 && i >= wi.c0 && i <= wi.c1);
    const bool upd_mid = NSTEP == 2 && lane >= 2 && lane < 62 && i >= S.upd_lo && i <= S.upd_hi;
    // a wavefront whose lanes 2..61 (the ones the second step reads) hold cells the first step leaves
    // alone -- boundary or ghost cells -- reloads their f^n; lanes 0, 1, 62, 63 feed no stored cell
    const bool keep_any = NSTEP == 2 && __ballot(!upd_mid && lane >= 2 && lane < 62 && k >= 0 && k < N) != 0ull;
    // the lane's left face comes from the boundary-face buffers (only where a stepped cell uses it)
    const bool bufL = (i == 1 && S.faceL_buf && S.upd_lo == 1) || (i == nx + 1 && S.faceR_buf && S.upd_hi == nx);
    const int bside = i == 1 ? 0 : 1;
    const bool seg_buf = __builtin_amdgcn_readfirstlane(__ballot(bufL) != 0ull ? 1 : 0) != 0;
    // the cell's w and prim are read where they are used (after pass 1), to keep pass 1's
    // register footprint small; pass 1 needs only the old velocity for the Shakhov heat flux
    const gptr<const double> win = to_global((const double*)S.w[in]);
    const bool shak = SHAK && i > 1 && i < nx;  // boundary cells relax with BGK (update_boundary.jl)
    // The cell's primitive variables at the start of the pass, prim^n = conserve_prim(w^n), from the
    // w buffer, double-buffered like f -- not from the stored prim, which a neighbouring workgroup may
    // already have overwritten with its halo cells' prim^{n+1} (updated cells store prim =
    // conserve_prim(w), the initial state KitBase's init_fvm sets the same way)
    constexpr bool DEL = sizeof(T) == 4;
    double pq[3] = {1.0, 0.0, 1.0};
    if (SHAK || DEL) {
        double wq[3];
#pragma unroll
        for (int m = 0; m < 3; m++) wq[m] = win[(size_t)kc * 3 + m];
        conserve_prim<3>(wq, S.gam, pq);
    }
    const double Uo64 = SHAK ? pq[1] : 0.0;
    // FP32: the stored f of the lane's cell departs from M~(prim^n) (for the second step of a pass,
    // prim^{n+1} where the first step updated the cell); lanes outside the solve read 0
    float dA = 0.f, dnl = 0.f, dU = 0.f, dkb = 0.f;
    if constexpr (DEL) {
        const MaxwCoef mc = maxw_coef(pq, K);
        dA = (k >= 0 && k < N) ? mc.A : 0.f;
        dnl = mc.nl2;
        dU = mc.U;
        dkb = mc.kb;
    }

    // ---- face coefficients of the lane's left face (KCU / UGKS), FP64 per lane
    SFace<T> F{};
    double fw0[3] = {0, 0, 0}, m3 = 0.0, m4 = 0.0;
    if constexpr (FLUX != KFVS) {
        double wo[3], sw[3], wL[3], wR[3];
#pragma unroll
        for (int m = 0; m < 3; m++) wo[m] = win[(size_t)kc * 3 + m];
        const double dxL = shr1(dxh64);
#pragma unroll
        for (int m = 0; m < 3; m++) {
            const double wl = shr1(wo[m]), wr = shl1(wo[m]);
            sw[m] = LIMK == LIM_NONE ? 0.0 : limiter<double>(lk, (wo[m] - wl) * rL64, (wr - wo[m]) * rR64);
        }
#pragma unroll
        for (int m = 0; m < 3; m++) {
            wL[m] = shr1(wo[m]) + shr1(sw[m]) * dxL;
            wR[m] = wo[m] - sw[m] * dxh64;
        }
        if constexpr (FLUX == KCU) {
            KcuFace<3> Fk;
            kcu_face<3>(wL, wR, K, S.gam, S.mu, S.om, dt, Fk);
            F.A = T(Fk.prim[0] * sqrt(Fk.prim[2] / kPi));
            F.nl2 = T(-Fk.prim[2] * kLog2e);
            F.Ui = T(Fk.prim[1]);
            F.kb = T(K / (2.0 * Fk.prim[2]));
            F.M0 = T(Fk.Mt1);
            F.M1 = T(Fk.Mt2);
            for (int m = 0; m < 3; m++) fw0[m] = Fk.fw0[m];
            m3 = Fk.Mt2;
        } else {
            UgksFace1 Fu;
            ugks_face_1d(wL, wR, K, S.gam, S.mu, S.om, dt, dxL, dxh64, Fu);
            F.A = T(Fu.prim[0] * sqrt(Fu.prim[2] / kPi));
            F.nl2 = T(-Fu.prim[2] * kLog2e);
            F.Ui = T(Fu.prim[1]);
            F.kb = T(K / (2.0 * Fu.prim[2]));
            F.Mxi2 = T(Fu.Mxi2);
            F.M0 = T(Fu.Mt[0]);
            F.M1 = T(Fu.Mt[1]);
            F.M2 = T(Fu.Mt[2]);
            F.M3 = T(Fu.Mt[3]);
            F.M4 = T(Fu.Mt[4]);
            F.aL0 = T(Fu.aL[0]);
            F.aL1 = T(Fu.aL[1]);
            F.aL2 = T(Fu.aL[2]);
            F.aR0 = T(Fu.aR[0]);
            F.aR1 = T(Fu.aR[1]);
            F.aR2 = T(Fu.aR[2]);
            F.aT0 = T(Fu.aT[0]);
            F.aT1 = T(Fu.aT[1]);
            F.aT2 = T(Fu.aT[2]);
            for (int m = 0; m < 3; m++) fw0[m] = Fu.fw0[m];
            m3 = Fu.Mt[3];
            m4 = Fu.Mt[4];
        }
    }

    // ---- pass 1: slopes, face fluxes, moment sums, transport update
    const int ivw = sweep_item_vw(nv, NW);  // this item's velocities per wavefront (<= VW)
    const int vbeg = wave * ivw;
    const int vend = min(nv, vbeg + ivw);
    T G[NF][VW];
    double acc[NR];
#pragma unroll
    for (int m = 0; m < NR; m++) acc[m] = 0.0;
    const gptr<const T> bff = to_global((const T*)S.bff);
    T UoT = T(Uo64);  // old velocity of the Shakhov heat flux (step 1's new one for a second step)

    // f rows of chunk c into buffer c & 1 (the next chunk's loads are in flight while a chunk computes)
    T frb[2][NF][CH];
    auto load_chunk = [&](auto cc) {
        constexpr int c = cc;
        const int vb = vbeg + c * CH;
        if (vb >= vend) return;
        static_for<CH>([&](auto jc) {
            constexpr int j = jc;
#pragma unroll
            for (int q = 0; q < NF; q++) frb[c & 1][q][j] = bload((T*)nullptr, fin, voff, (q * nv + vb + j) * rowb);
        });
    };

    // one chunk of CH velocities; UPM: 0 all take the right state, 1 all the left, 2 mixed;
    // BUF: a lane's left face comes from the boundary-face buffers
    auto chunk1 = [&](auto cc, auto upm, auto buf, auto stc) {
        constexpr int c = cc, UPM = upm, SS = stc;
        constexpr bool BUF = buf;
        const int vb = vbeg + c * CH;
        VC& ch = vt[vb / CH];
        T(&fr)[NF][CH] = frb[c & 1];
        T part[NR];  // chunk sums in the storage precision, added to the FP64 totals below
#pragma unroll
        for (int m = 0; m < NR; m++) part[m] = T(0);
        float Md[CH];  // M~ the stored departures are taken from
        if constexpr (DEL) maxw8(ch, dA, dnl, dU, Md);
        static_for<CH>([&](auto jc) {
            constexpr int j = jc;
            const T u = ch.u[j], uu = ch.uu[j];
            const bool up = (ch.upmask >> j) & 1;
            T fq[NF], dq[NF], sq[NF], fu[NF], su[NF], ffq[NF];
#pragma unroll
            for (int q = 0; q < NF; q++) {
                T f;
                if constexpr (SS == 0) f = fr[q][j];
                else f = G[q][c * CH + j];  // the first step's f^{n+1} (departure in FP32), still in registers
                dq[q] = f;
                if constexpr (DEL) f = (q == 0 ? T(Md[j]) : T(maxw_b(Md[j], dkb))) + f;
                fq[q] = f;
                // the right difference f[L+1] - f[L] is the next lane's left one, bit for bit
                const T d = f - shr1(f);
                sq[q] = sweep_limit<T, LIMK, HALF>(lk, d * rL, shl1(d) * rR);
            }
            // upwind state and slope of the left face (flux_kfvs.jl:238-264, :303-340)
#pragma unroll
            for (int q = 0; q < NF; q++) {
                if constexpr (UPM == 1) {
                    fu[q] = shr1(fq[q] + sq[q] * dxs);
                    su[q] = shr1(sq[q]);
                } else if constexpr (UPM == 0) {
                    fu[q] = fq[q] - sq[q] * dxs;
                    su[q] = sq[q];
                } else {
                    const T l = shr1(fq[q] + sq[q] * dxs), r = fq[q] - sq[q] * dxs;
                    const T sl = shr1(sq[q]);
                    fu[q] = up ? l : r;
                    su[q] = up ? sl : sq[q];
                }
            }
            if constexpr (FLUX == KFVS) {
#pragma unroll
                for (int q = 0; q < NF; q++) ffq[q] = fu[q] * ch.cu[j] + su[q] * ch.cs2[j];
            } else if constexpr (FLUX == KCU) {
                // flux_kcu.jl:241-303 (1F1V), :463-530 (2F1V)
                const T cI = u - F.Ui;
                const T g = F.A * texp2(F.nl2 * (cI * cI));
                ffq[0] = F.M0 * u * g + F.M1 * u * fu[0];
                if constexpr (NF == 2) ffq[1] = F.M0 * u * (g * F.kb) + F.M1 * u * fu[1];
            } else {
                // flux_ugks.jl:91-111 (1D2F1V)
                const T x = u;
                const T cI = x - F.Ui;
                const T H = F.A * texp2(F.nl2 * (cI * cI));
                const T Bq = H * F.kb;
                T a0, a1, a2;
                if constexpr (UPM == 1) {
                    a0 = F.aL0, a1 = F.aL1, a2 = F.aL2;
                } else if constexpr (UPM == 0) {
                    a0 = F.aR0, a1 = F.aR1, a2 = F.aR2;
                } else {
                    a0 = up ? F.aL0 : F.aR0, a1 = up ? F.aL1 : F.aR1, a2 = up ? F.aL2 : F.aR2;
                }
                const T pH = a0 * H + a1 * x * H + T(0.5) * a2 * (uu * H + Bq);
                const T pTH = F.aT0 * H + F.aT1 * x * H + T(0.5) * F.aT2 * (uu * H + Bq);
                const T pB = a0 * Bq + a1 * x * Bq + T(0.5) * a2 * (uu * Bq + F.Mxi2 * H);
                const T pTB = F.aT0 * Bq + F.aT1 * x * Bq + T(0.5) * F.aT2 * (uu * Bq + F.Mxi2 * H);
                ffq[0] = F.M0 * x * H + F.M1 * uu * pH + F.M2 * x * pTH + F.M3 * x * fu[0] - F.M4 * uu * su[0];
                ffq[NF - 1] = F.M0 * x * Bq + F.M1 * uu * pB + F.M2 * x * pTB + F.M3 * x * fu[NF - 1] - F.M4 * uu * su[NF - 1];
            }
            if constexpr (BUF) {
                // a face with no neighbour to reconstruct from, or a wall: the boundary-face buffers
                const int v = min(vb + j, nv - 1);
#pragma unroll
                for (int q = 0; q < NF; q++) {
                    const T bv = bff[((long)bside * NF + q) * nv + v];
                    ffq[q] = bufL ? bv : ffq[q];
                }
            }
            // moment sums of the face
            if constexpr (FLUX == KFVS) {
                part[0] += ch.wk[0][j] * ffq[0];
                part[1] += ch.wk[1][j] * ffq[0];
                part[2] += ch.wk[2][j] * ffq[0];
                if constexpr (NF == 2) part[3] += ch.wk[0][j] * ffq[1];
            } else if constexpr (FLUX == KCU) {
                part[0] += ch.wk[1][j] * fu[0];
                part[1] += ch.wk[2][j] * fu[0];
                part[2] += ch.wk[3][j] * fu[0];
                if constexpr (NF == 2) part[3] += ch.wk[1][j] * fu[1];
            } else {
                part[0] += ch.wk[1][j] * fu[0];
                part[1] += ch.wk[2][j] * su[0];
                part[2] += ch.wk[2][j] * fu[0];
                part[3] += ch.wk[3][j] * su[0];
                part[4] += ch.wk[3][j] * fu[0] + ch.wk[1][j] * fu[1];
                part[5] += ch.wk[4][j] * su[0] + ch.wk[2][j] * su[1];
            }
            if constexpr (SHAK) {
                // heat flux of the cell from the old f and prim (theory_moments.jl:523, :530-531)
                const T co = u - UoT;
                const T wc = ch.wk[0][j] * co;
                part[NRAW] += (wc * (co * co)) * fq[0];
                if constexpr (NF == 2) part[NRAW] += wc * fq[1];
            }
            // transport update G = f + (ff_left - ff_right) / dx
#pragma unroll
            for (int q = 0; q < NF; q++) G[q][c * CH + j] = (DEL ? dq[q] : fq[q]) + (ffq[q] - shl1(ffq[q])) * rdx;
        });
#pragma unroll
        for (int m = 0; m < NR; m++) acc[m] += double(part[m]);
    };
    using I0 = std::integral_constant<int, 0>;
    using I1 = std::integral_constant<int, 1>;
    using I2 = std::integral_constant<int, 2>;
    using BT = std::integral_constant<bool, true>;
    using BF = std::integral_constant<bool, false>;
    double wkeep[3] = {0.0, 0.0, 0.0}, pkeep[3] = {1.0, 0.0, 1.0};  // wavefront 0: step 1's w, prim
    static_for<NSTEP>([&](auto sc) {
        constexpr int s = sc;
        constexpr bool LAST = s == NSTEP - 1;
        using SI = std::integral_constant<int, s>;
        if constexpr (s > 0) {
#pragma unroll
            for (int m = 0; m < NR; m++) acc[m] = 0.0;
        }
        if constexpr (s == 0) load_chunk(I0{});
        static_for<VW / CH>([&](auto cc) {
            constexpr int c = cc;
            const int vb = vbeg + c * CH;
            if (vb >= vend) return;
            if constexpr (s == 0 && c + 1 < VW / CH) load_chunk(std::integral_constant<int, c + 1>{});
            if (LIMK == LIM_VANLEER && !seg_buf) {
                // velocities are sorted: a chunk whose velocities agree in sign takes one side throughout
                const int m = vt[vb / CH].upmask;
                if (m == 0xff) chunk1(cc, I1{}, BF{}, SI{});
                else if (m == 0) chunk1(cc, I0{}, BF{}, SI{});
                else chunk1(cc, I2{}, BF{}, SI{});
            } else if (seg_buf) {
                chunk1(cc, I2{}, BT{}, SI{});
            } else {
                chunk1(cc, I2{}, BF{}, SI{});
            }
        });

        // ---- wavefront 0 adds the wavefronts' sums (fixed order) and runs the cells' step! chain;
        // the others take its coefficients from LDS
        T An, A1, nl2, Un, inv, rTi, kbn, pre = 0, Ao = 0, nlo2 = 0, Uo = 0, lo2 = 0, kbo = 0;
        T* coef = (T*)(lds + (size_t)NW * NR * kSweepLanes);  // [13][64]
        if (NW > 1) {
#pragma unroll
            for (int m = 0; m < NR; m++) lds[(wave * NR + m) * kSweepLanes + lane] = acc[m];
            __syncthreads();
        }
        if (wave == 0) {
            double tot[NR];
            if (NW > 1) {
#pragma unroll
                for (int m = 0; m < NR; m++) {
                    double sm = 0.0;
                    for (int w = 0; w < NW; w++) sm += lds[(w * NR + m) * kSweepLanes + lane];
                    tot[m] = sm;
                }
            } else {
#pragma unroll
                for (int m = 0; m < NR; m++) tot[m] = acc[m];
            }
            double wo[3], po[3] = {1.0, 0.0, 1.0};
            if constexpr (s == 0) {
#pragma unroll
                for (int m = 0; m < 3; m++) wo[m] = win[(size_t)kc * 3 + m];
                if (SHAK) {
#pragma unroll
                    for (int m = 0; m < 3; m++) po[m] = pq[m];
                }
            } else {
#pragma unroll
                for (int m = 0; m < 3; m++) {
                    wo[m] = wkeep[m];
                    po[m] = pkeep[m];
                }
            }

            // face moments of the lane's left face, and of its right face (the next lane's left)
            double fwl[3];
            if constexpr (FLUX == KFVS) {
                fwl[0] = tot[0];
                fwl[1] = tot[1];
                fwl[2] = NF == 2 ? 0.5 * (tot[2] + tot[3]) : 0.5 * tot[2];
            } else if constexpr (FLUX == KCU) {
                fwl[0] = fw0[0] + m3 * tot[0];
                fwl[1] = fw0[1] + m3 * tot[1];
                fwl[2] = fw0[2] + m3 * 0.5 * (NF == 2 ? tot[2] + tot[3] : tot[2]);
            } else {
                fwl[0] = fw0[0] + (m3 * tot[0] - m4 * tot[1]);
                fwl[1] = fw0[1] + (m3 * tot[2] - m4 * tot[3]);
                fwl[2] = fw0[2] + (m3 * 0.5 * tot[4] - m4 * 0.5 * tot[5]);
            }
            // pinned: otherwise the override below becomes one load through a pointer that selects
            // between bfw and the moment sums, which sends the sums to scratch memory
#pragma unroll
            for (int m = 0; m < 3; m++) asm volatile("" : "+v"(fwl[m]));
            if (bufL) {
#pragma unroll
                for (int m = 0; m < 3; m++) fwl[m] = S.bfw[bside * 3 + m];
            }
            double wn[3], pn[3];
#pragma unroll
            for (int m = 0; m < 3; m++) wn[m] = wo[m] + (fwl[m] - shl1(fwl[m])) * rdx64;

            // the cell's step! coefficients (FP64 chain, lane-parallel)
            conserve_prim<3>(wn, S.gam, pn);
            if (LAST ? upd : upd_mid) flag_health(S.health, prim_health<3>(pn));
            const T rT = relax_ratio<T>(pn[0], pn[2], S.mu, S.om, dt);
            if constexpr (DEL) {
                // FP32: M~^{n+1} from the same coefficients every kernel evaluates M~ with
                inv = __builtin_amdgcn_rcpf(1.f + rT);
                const MaxwCoef mc = maxw_coef(pn, K);
                A1 = mc.A;
                nl2 = mc.nl2;
                Un = mc.U;
                kbn = NF == 2 ? mc.kb : 0.f;
            } else {
                inv = T(1) / (T(1) + rT);
                A1 = T(pn[0] * sqrt(pn[2] / kPi));
                nl2 = T(-pn[2] * kLog2e);
                Un = T(pn[1]);
                kbn = NF == 2 ? T(K / (2.0 * pn[2])) : T(0);  // the b distribution's factor only
            }
            rTi = rT * inv;
            // the relaxation weight rT / (1 + rT) is folded into the equilibrium amplitude (and the
            // Shakhov prefactor below), one multiply per velocity less in pass 2
            An = A1 * rTi;
            if (SHAK && shak) {
                const double q = 0.5 * tot[NRAW];
                pre = T(0.8 * (1.0 - S.Pr) * (po[2] * po[2]) / po[0] * q) * rTi;
                Uo = T(po[1]);
                lo2 = T(2.0 * po[2]);
                kbo = NF == 2 ? T(K / (2.0 * po[2])) : T(0);
                if constexpr (!DEL) {
                    // the heat-flux prefactor rides on the old equilibrium's amplitude (pass 2 uses it
                    // only in their product)
                    Ao = T(po[0] * sqrt(po[2] / kPi)) * pre;
                    nlo2 = T(-po[2] * kLog2e);
                }
            }
            if constexpr (!LAST) {
                // the first of two steps: the cells it updates carry w, prim to the second
#pragma unroll
                for (int m = 0; m < 3; m++) {
                    wkeep[m] = upd_mid ? wn[m] : wo[m];
                    pkeep[m] = upd_mid ? pn[m] : po[m];
                }
            }
            if (LAST && upd) {
                const size_t r = (size_t)kc * 3;
                const gptr<double> wout = to_global(S.w[in ^ 1]);
                const gptr<double> gp = to_global(S.prim);
#pragma unroll
                for (int m = 0; m < 3; m++) {
                    wout[r + m] = wn[m];
                    gp[r + m] = pn[m];
                }
            }
            if (LAST && rrow) {
                // residual sums (solver_update.jl:86-88, update_boundary.jl:41-44): interior / boundary cells
                double r[12];
                const bool bnd = i == 1 || i == nx;
#pragma unroll
                for (int m = 0; m < 3; m++) {
                    const double d2 = upd ? (wn[m] - wo[m]) * (wn[m] - wo[m]) : 0.0, a = upd ? fabs(wn[m]) : 0.0;
                    r[m] = bnd ? 0.0 : d2;
                    r[3 + m] = bnd ? 0.0 : a;
                    r[6 + m] = bnd ? d2 : 0.0;
                    r[9 + m] = bnd ? a : 0.0;
                }
                wave_allreduce<12>(r);
                double y = 0.0;
#pragma unroll
                for (int m = 0; m < 12; m++)
                    if (lane == m) y = r[m];
                if (lane < 12) rrow[lane] = y;
            }
            if (NW > 1) {
                const T cv[13] = {An, nl2, Un, inv, rTi, kbn, pre, Ao, nlo2, Uo, lo2, kbo, A1};
#pragma unroll
                for (int m = 0; m < 13; m++) coef[m * kSweepLanes + lane] = cv[m];
            }
        }
        if (NW > 1) {
            __syncthreads();
            if (wave != 0) {
                An = coef[0 * kSweepLanes + lane];
                nl2 = coef[1 * kSweepLanes + lane];
                Un = coef[2 * kSweepLanes + lane];
                inv = coef[3 * kSweepLanes + lane];
                rTi = coef[4 * kSweepLanes + lane];
                kbn = coef[5 * kSweepLanes + lane];
                A1 = coef[12 * kSweepLanes + lane];
                if (SHAK) {
                    pre = coef[6 * kSweepLanes + lane];
                    Ao = coef[7 * kSweepLanes + lane];
                    nlo2 = coef[8 * kSweepLanes + lane];
                    Uo = coef[9 * kSweepLanes + lane];
                    lo2 = coef[10 * kSweepLanes + lane];
                    kbo = coef[11 * kSweepLanes + lane];
                }
            }
        }
        if constexpr (!LAST) UoT = Un;  // the second step's Shakhov heat flux is about prim^{n+1}

        // ---- pass 2: relaxation toward M (+ Shakhov) and the store of f^{n+1}
        // M at the chunk's nodes: one exp per node, or -- FP32 on equally spaced nodes -- the product
        // recurrence M(u + du) = M(u) R, R <- R Q with R = exp(-lam du (2 (u - U) + du)), Q = exp(-2 lam du^2),
        // restarted every chunk of 8 nodes (at most ~30 roundings against the exp per node)
        auto maxw = [&](VC& ch, T A, T nlam2, T U0, T (&M)[CH]) {
            if (sizeof(T) == 4 && ch.du != T(0)) {
                const T d = ch.du, c0 = ch.u[0] - U0;
                T m = A * texp2(nlam2 * (c0 * c0));
                T R = texp2(nlam2 * d * (T(2) * c0 + d));
                const T Q = texp2(T(2) * nlam2 * (d * d));
                static_for<CH>([&](auto jc) {
                    constexpr int j = jc;
                    M[j] = m;
                    if constexpr (j + 1 < CH) {
                        m *= R;
                        R *= Q;
                    }
                });
            } else {
                static_for<CH>([&](auto jc) {
                    constexpr int j = jc;
                    const T c = ch.u[j] - U0;
                    M[j] = A * texp2(nlam2 * (c * c));
                });
            }
        };
        static_for<VW / CH>([&](auto cc) {
            constexpr int c = cc;
            const int vb = vbeg + c * CH;
            if (vb >= vend) return;
            VC& ch = vt[vb / CH];
            T fn[NF][CH], Mn[CH], Mo[CH];
            if constexpr (DEL) {
                // departure form: d^{n+1} = f^{n+1} - M~^{n+1} = (M~^n - M~^{n+1} + G) / (1 + rT) + S rT / (1 + rT),
                // G = d^n + (ff_left - ff_right) / dx; the Shakhov term from the old M~ times its prefactor
                float Mnd[CH], Mod[CH];
                maxw8(ch, A1, nl2, Un, Mnd);
                maxw8(ch, dA, dnl, dU, Mod);
                static_for<CH>([&](auto jc) {
                    constexpr int j = jc;
                    T sh0 = T(0), sh1 = T(0);
                    if constexpr (SHAK) {
                        const T co = ch.u[j] - Uo, c2 = co * co;
                        const T t = co * (Mod[j] * pre);
                        if constexpr (NF == 1) {
                            sh0 = t * (lo2 * c2 - T(5));
                        } else {
                            sh0 = t * (lo2 * c2 + T(K) - T(5));
                            sh1 = (t * kbo) * (lo2 * c2 + T(K) - T(3));
                        }
                    }
                    fn[0][j] = ((Mod[j] - Mnd[j]) + G[0][c * CH + j]) * inv + sh0;
                    if constexpr (NF == 2) fn[1][j] = ((maxw_b(Mod[j], dkb) - maxw_b(Mnd[j], kbn)) + G[1][c * CH + j]) * inv + sh1;
                });
            } else {
            maxw(ch, An, nl2, Un, Mn);
            if constexpr (SHAK) maxw(ch, Ao, nlo2, Uo, Mo);
            static_for<CH>([&](auto jc) {
                constexpr int j = jc;
                T MH = Mn[j];
                T MB = MH * kbn;
                if constexpr (SHAK) {
                    const T co = ch.u[j] - Uo, c2 = co * co;
                    const T t = co * Mo[j];  // Mo carries the prefactor
                    if constexpr (NF == 1) {
                        MH += t * (lo2 * c2 - T(5));
                    } else {
                        MH += t * (lo2 * c2 + T(K) - T(5));
                        MB += (t * kbo) * (lo2 * c2 + T(K) - T(3));
                    }
                }
                fn[0][j] = G[0][c * CH + j] * inv + MH;  // MH, MB carry the weight rT / (1 + rT)
                if constexpr (NF == 2) fn[1][j] = G[1][c * CH + j] * inv + MB;
            });
            }
            if constexpr (LAST) {
                if (upd) {
                    if (vb + CH <= vend) {
                        static_for<CH>([&](auto jc) {
                            constexpr int j = jc;
#pragma unroll
                            for (int q = 0; q < NF; q++) bstore<KBG_SWEEP_STORE_AUX>(fn[q][j], fout, voff, (q * nv + vb + j) * rowb);
                        });
                    } else {
                        static_for<CH>([&](auto jc) {
                            constexpr int j = jc;
                            if (vb + j < vend) {
#pragma unroll
                                for (int q = 0; q < NF; q++) bstore<KBG_SWEEP_STORE_AUX>(fn[q][j], fout, voff, (q * nv + vb + j) * rowb);
                            }
                        });
                    }
                }
            } else {
                // f^{n+1} becomes the second step's input; cells the first step left alone keep f^n
                static_for<CH>([&](auto jc) {
                    constexpr int j = jc;
#pragma unroll
                    for (int q = 0; q < NF; q++) G[q][c * CH + j] = fn[q][j];
                });
                if (keep_any) {
                    static_for<CH>([&](auto jc) {
                        constexpr int j = jc;
                        if (vb + j < vend) {
#pragma unroll
                            for (int q = 0; q < NF; q++) {
                                const T f0 = bload((T*)nullptr, fin, voff, (q * nv + vb + j) * rowb);
                                G[q][c * CH + j] = upd_mid ? G[q][c * CH + j] : f0;
                            }
                        }
                    });
                }
            }
        });
        if constexpr (DEL && !LAST) {
            // the second step's departures are from M~(prim^{n+1}) where the first step updated the cell
            dA = upd_mid ? float(A1) : dA;
            dnl = upd_mid ? float(nl2) : dnl;
            dU = upd_mid ? float(Un) : dU;
            dkb = upd_mid ? (NF == 2 ? float(kbn) : dkb) : dkb;
        }
    });
}

template <class T, int NF, int FLUX, int LIMK, int NSTEP>
KBG_D void sweep_coll(const Solve1D& S, const WorkItem& wi, int par, int lane, int wave, int NW, double* lds, double* rrow) {
    if (S.coll == SHAKHOV) sweep1d<T, NF, FLUX, LIMK, true, NSTEP>(S, wi, par, lane, wave, NW, lds, rrow);
    else sweep1d<T, NF, FLUX, LIMK, false, NSTEP>(S, wi, par, lane, wave, NW, lds, rrow);
}
template <class T, int NF, int FLUX, int NSTEP>
KBG_D void sweep_lim(const Solve1D& S, const WorkItem& wi, int par, int lane, int wave, int NW, double* lds, double* rrow) {
    if (!S.interp2) sweep_coll<T, NF, FLUX, LIM_NONE, NSTEP>(S, wi, par, lane, wave, NW, lds, rrow);
    else if (S.limiter == VANLEER) sweep_coll<T, NF, FLUX, LIM_VANLEER, NSTEP>(S, wi, par, lane, wave, NW, lds, rrow);
    else sweep_coll<T, NF, FLUX, LIM_GENERIC, NSTEP>(S, wi, par, lane, wave, NW, lds, rrow);
}

// LDS of one item (doubles): the wavefronts' moment sums and wavefront 0's cell coefficients
template <class T, int NF, int FLUX> constexpr size_t sweep_lds_item(int nw) {
    return nw > 1 ? (size_t)nw * SweepCfg<T, NF, FLUX>::NR * kSweepLanes + (13 * kSweepLanes * sizeof(T) + 7) / 8 : 0;
}

// occupancy target: 4 wavefronts per SIMD (<= 128 VGPRs) for the FP32 one-distribution sweeps,
// whatever the register file allows for the others
template <class T, int NF, int FLUX> constexpr int sweep_waves_per_eu() { return sizeof(T) == 4 && NF == 1 ? KBG_SWEEP_WPE_F32_1F : 1; }

// One workgroup per pack; packs of one XCD (workgroup ids congruent mod 8) are a contiguous run,
// so neighbouring segments share that XCD's L2.  A pack's items sit side by side: wavefront w
// serves item w / nw as its wavefront w % nw.
template <class T, int NF, int FLUX>
__global__ __launch_bounds__(64 * kSweepMaxWaves) __attribute__((amdgpu_waves_per_eu(sweep_waves_per_eu<T, NF, FLUX>()))) void k_sweep1d(
    const Solve1D* __restrict__ solves, const WorkItem* __restrict__ items, const Pack* __restrict__ packs, int npacks, int par,
    double* __restrict__ rpart, int nstep) {
    extern __shared__ double kbg_sweep_lds[];
    const int G = gridDim.x, b = blockIdx.x;
    const int x = b & 7, q8 = G >> 3, r8 = G & 7;
    // each XCD walks its run of packs forwards on even steps and backwards on odd ones, so a step
    // first reads the rows the previous step wrote last (still in the memory-side cache)
    const int nx8 = q8 + (x < r8 ? 1 : 0);
    const int pk = x * q8 + (x < r8 ? x : r8) + ((KBG_SWEEP_ZIGZAG && (par & 1)) ? nx8 - 1 - (b >> 3) : (b >> 3));
    if (pk >= npacks) return;
    const Pack P = packs[pk];
    const int lane = threadIdx.x & 63;
    const int wv = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int NW = P.nw;
    const int sub = wv / NW, wave = wv - sub * NW;
    if (sub >= P.count) {
        // an unused slot still takes part in the item barriers (two per time step when nw > 1)
        if (NW > 1) {
            const int nb = (FLUX == KFVS && nstep == 2) ? 4 : 2;
            for (int b2 = 0; b2 < nb; b2++) __syncthreads();
        }
        return;
    }
    const int it = P.item0 + sub;
    const WorkItem wi = items[it];
    const Solve1D& S = solves[wi.solve];
    double* lds = kbg_sweep_lds + (size_t)sub * sweep_lds_item<T, NF, FLUX>(NW);
    double* rrow = rpart ? rpart + (size_t)it * 12 : nullptr;
    if constexpr (FLUX == KFVS) {
        if (nstep == 2) {
            sweep_lim<T, NF, FLUX, 2>(S, wi, par, lane, wave, NW, lds, rrow);
            return;
        }
    }
    sweep_lim<T, NF, FLUX, 1>(S, wi, par, lane, wave, NW, lds, rrow);
}

template <class T, int NF, int FLUX>
hipError_t launch_sweep_t(int wg_waves, const Solve1D* d_solves, const WorkItem* d_items, const Pack* d_packs, int npacks,
                          int par, double* d_rpart, int nstep, hipStream_t st) {
    // the largest LDS a pack of this launch can need: wg_waves / nw items of nw wavefronts each
    size_t lds = 0;
    for (int nw = 2; nw <= wg_waves; nw *= 2) lds = std::max(lds, (size_t)(wg_waves / nw) * sweep_lds_item<T, NF, FLUX>(nw));
    hipLaunchKernelGGL((k_sweep1d<T, NF, FLUX>), dim3(npacks), dim3(64 * wg_waves), lds * sizeof(double), st, d_solves, d_items,
                       d_packs, npacks, par, d_rpart, nstep);
    return hipGetLastError();
}

}  // namespace kbg
