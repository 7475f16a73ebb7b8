// kbg_api.hip -- the C ABI of libkbgpu.so (declared in include/kbgpu.h): contexts, device-resident
// solves, host<->device conversion, the phase-exact reconstruct/evolve/update entry points, and the
// fused single-solve and batched time steps.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <climits>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <initializer_list>
#include <map>
#include <mutex>
#include <string>
#include <thread>
#include <tuple>
#include <vector>

#include "../../include/kbgpu.h"
#include "kbg_sweep.h"
#include "kbg_types.h"

using kbg::Solve1D;
using kbg::Solve2D;
using kbg::WorkItem;

namespace {

thread_local std::string g_err;

int fail(int code, const std::string& msg) {
    g_err = msg;
    return code;
}

#define HIPCHK(expr)                                                                                   \
    do {                                                                                               \
        hipError_t e_ = (expr);                                                                        \
        if (e_ != hipSuccess) return fail(KBG_ERR_HIP, std::string(#expr) + ": " + hipGetErrorString(e_)); \
    } while (0)

constexpr int kMaxSeg = kbg::kSweepOut;  // fused 1D sweep: cells per work item

}  // namespace

struct kbg_ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    hipEvent_t ev[16] = {};
    // independent kernel groups of a batch run concurrently: group 0 on `stream`, the others on
    // these side streams (default priority), forked from and joined back into `stream`
    static constexpr int kSide = 3;
    hipStream_t side[kSide] = {};
    hipEvent_t fork = nullptr, join[kSide] = {};
    int seg_cells = kbg::kSweepOut;  // cells per sweep workgroup
    int streams = 0;                 // concurrent streams of a batch (0: all, 1 + kSide)
    // batch staging (device)
    Solve1D* d_descs = nullptr;
    size_t descs_cap = 0;
    WorkItem* d_items = nullptr;
    size_t items_cap = 0;
    double* d_rpart = nullptr;
    size_t rpart_cap = 0;
    kbg::Pack* d_packs = nullptr;
    size_t packs_cap = 0;
    double* d_res = nullptr;  // [4096][5]
    Solve2D* d_descs2 = nullptr;  // fused 2D launches: descriptors and cell / face prefix sums
    size_t descs2_cap = 0;
    int* d_pref2 = nullptr;
    size_t pref2_cap = 0;
    int sweep_steps = 2;  // time steps per fused 1D sweep launch where the solves allow two
    int sweep_vw = 0;     // most velocities per sweep wavefront (0: the kernel's width)
    // HIP graphs of a batch's steps, captured for the launch plan whose signature is gsig:
    // [p << 1 | 0] a round of two steps, [p << 1 | 1] a round of one step, where p is the buffer
    // parity of the sweep launches that take two steps per launch at the round's start (launches of
    // one step per launch start every round at parity 0)
    bool graphs = false;
    std::string gsig;
    hipGraph_t graph[4] = {};
    hipGraphExec_t gexec[4] = {};
    void drop_graphs() {
        for (int k = 0; k < 4; k++) {
            if (gexec[k]) (void)hipGraphExecDestroy(gexec[k]);
            if (graph[k]) (void)hipGraphDestroy(graph[k]);
            gexec[k] = nullptr;
            graph[k] = nullptr;
        }
        gsig.clear();
    }
    // host <-> device transfers of f: two pinned host buffers (hipHostMalloc) that worker threads
    // fill or drain while the other one's copy runs, allocated at the first large transfer
    static constexpr size_t kPinBytes = size_t(32) << 20;
    void* pin[2] = {};
    hipEvent_t pin_ev[2] = {};
};

struct kbg_solve {
    kbg_ctx* ctx = nullptr;
    kbg_desc d{};
    bool fp64 = false;
    size_t esz = 4;
    int nw = 3, N = 0, RS = 0;
    bool fused = false;   // stepped by the fused sweep (cell-fastest f)
    bool rep_full = false;  // FP32 solve stored as departures whose buffers hold full values for now
    std::vector<float> mref;  // 2D departure solves: the reference Maxwellian (h, b) as stored
    // copies of the host arrays the descriptor pointed to (the caller's arrays need only live through
    // kbg_solve_create); d's pointers are redirected to them
    std::vector<double> hu, hv, hw, hwt, hdx, hdy, hbc, hxgeo, hygeo, hxwall, hywall;
    int thr2 = 0;         // 2D: threads per workgroup of the fused step (0: phase kernels)
    int sweep_nw = 0;     // wavefronts per sweep workgroup
    double vtab_dt = -1;  // dt the velocity table was built for
    void* vtab = nullptr;
    // common device buffers
    void *u = nullptr, *v = nullptr, *wv = nullptr, *wt = nullptr;
    double *dx = nullptr, *dy = nullptr;
    double* rcell = nullptr;
    double* res = nullptr;
    // 1D
    Solve1D h1{};
    Solve1D* d1 = nullptr;
    // 2D
    Solve2D h2{};
    Solve2D* d2 = nullptr;
    // unstructured triangle solve (kbg_usolve_create; d.dim is 0)
    bool us = false;
    kbg::SolveUS hus{};
    kbg::SolveUS* dus = nullptr;
    int thru = 0;  // threads per workgroup of its face and cell kernels
    // collision :fsm (1D1F3V): kernel modes, work buffers and hipFFT plans
    bool has_fsm = false;
    kbg::FsmPlan fsm{};
    std::vector<void*> allocs;
};

namespace {

int dev_alloc(kbg_solve* s, void** p, size_t bytes) {
    if (bytes == 0) bytes = 8;
    hipError_t e = hipMalloc(p, bytes);
    if (e != hipSuccess) return fail(KBG_ERR_NOMEM, std::string("hipMalloc(") + std::to_string(bytes) + "): " + hipGetErrorString(e));
    s->allocs.push_back(*p);
    return KBG_OK;
}

// f(i0, i1) over [0, n) split across up to 8 host threads (one below 1 M elements)
template <class F> void par_for(size_t n, F&& f) {
    static const int hw = std::max(1, std::min(8, (int)std::thread::hardware_concurrency()));
    const int T = n < (size_t(1) << 20) ? 1 : hw;
    if (T == 1) {
        f(size_t(0), n);
        return;
    }
    const size_t per = (n + T - 1) / T;
    std::vector<std::thread> th;
    for (int t = 1; t < T; t++) {
        const size_t a = t * per, b = std::min(n, a + per);
        if (a < b) th.emplace_back([&f, a, b] { f(a, b); });
    }
    f(size_t(0), std::min(n, per));
    for (auto& x : th) x.join();
}

int pin_ready(kbg_ctx* c) {
    for (int b = 0; b < 2; b++) {
        if (!c->pin[b]) HIPCHK(hipHostMalloc(&c->pin[b], kbg_ctx::kPinBytes, hipHostMallocDefault));
        if (!c->pin_ev[b]) HIPCHK(hipEventCreateWithFlags(&c->pin_ev[b], hipEventDisableTiming));
    }
    return KBG_OK;
}

// dst[i] = conv(i) for i < n on the device (T the device type): chunks converted by worker threads
// into one pinned buffer while the other one's copy runs
template <class T, class Conv> int stage_up(kbg_solve* s, T* dst, size_t n, Conv conv) {
    kbg_ctx* c = s->ctx;
    int rc;
    if ((rc = pin_ready(c))) return rc;
    const size_t per = kbg_ctx::kPinBytes / sizeof(T);
    for (size_t k = 0, i0 = 0; i0 < n; k++, i0 += per) {
        const int b = int(k & 1);
        const size_t m = std::min(per, n - i0);
        if (k >= 2) HIPCHK(hipEventSynchronize(c->pin_ev[b]));  // this buffer's previous copy is done
        T* hb = static_cast<T*>(c->pin[b]);
        par_for(m, [&](size_t a, size_t e) {
            for (size_t i = a; i < e; i++) hb[i] = conv(i0 + i);
        });
        HIPCHK(hipMemcpyAsync(dst + i0, hb, m * sizeof(T), hipMemcpyHostToDevice, c->stream));
        HIPCHK(hipEventRecord(c->pin_ev[b], c->stream));
    }
    HIPCHK(hipStreamSynchronize(c->stream));
    return KBG_OK;
}

// store(i, src[i]) for i < n from the device: the next chunk's copy runs while worker threads drain
// the current one
template <class T, class Store> int stage_down(kbg_solve* s, const T* src, size_t n, Store store) {
    kbg_ctx* c = s->ctx;
    int rc;
    if ((rc = pin_ready(c))) return rc;
    const size_t per = kbg_ctx::kPinBytes / sizeof(T), nch = (n + per - 1) / per;
    auto issue = [&](size_t k) -> int {
        const int b = int(k & 1);
        const size_t i0 = k * per, m = std::min(per, n - i0);
        HIPCHK(hipMemcpyAsync(c->pin[b], src + i0, m * sizeof(T), hipMemcpyDeviceToHost, c->stream));
        HIPCHK(hipEventRecord(c->pin_ev[b], c->stream));
        return KBG_OK;
    };
    if (nch && (rc = issue(0))) return rc;
    for (size_t k = 0; k < nch; k++) {
        if (k + 1 < nch && (rc = issue(k + 1))) return rc;  // into the buffer chunk k - 1 was drained from
        const int b = int(k & 1);
        HIPCHK(hipEventSynchronize(c->pin_ev[b]));
        const size_t i0 = k * per, m = std::min(per, n - i0);
        const T* hb = static_cast<const T*>(c->pin[b]);
        par_for(m, [&](size_t a, size_t e) {
            for (size_t i = a; i < e; i++) store(i0 + i, hb[i]);
        });
    }
    return KBG_OK;
}

// copy n doubles to device as the solve's storage type
int put_vec(kbg_solve* s, void* dst, const double* src, size_t n) {
    if (s->fp64) {
        HIPCHK(hipMemcpyAsync(dst, src, n * 8, hipMemcpyHostToDevice, s->ctx->stream));
        HIPCHK(hipStreamSynchronize(s->ctx->stream));
        return KBG_OK;
    }
    return stage_up(s, static_cast<float*>(dst), n, [src](size_t i) { return float(src[i]); });
}

int get_vec(kbg_solve* s, double* dst, const void* src, size_t n) {
    if (s->fp64) {
        HIPCHK(hipMemcpyAsync(dst, src, n * 8, hipMemcpyDeviceToHost, s->ctx->stream));
        HIPCHK(hipStreamSynchronize(s->ctx->stream));
        return KBG_OK;
    }
    return stage_down(s, static_cast<const float*>(src), n, [dst](size_t i, float v) { dst[i] = double(v); });
}

int put_f64(kbg_solve* s, double* dst, const double* src, size_t n) {
    HIPCHK(hipMemcpyAsync(dst, src, n * 8, hipMemcpyHostToDevice, s->ctx->stream));
    HIPCHK(hipStreamSynchronize(s->ctx->stream));
    return KBG_OK;
}

// FP64 host values to device values of the solve's real type (FP32 when f32)
int put_real(kbg_solve* s, void* dst, const double* src, size_t n, bool f32) {
    if (!f32) return put_f64(s, (double*)dst, src, n);
    std::vector<float> tmp(src, src + n);
    HIPCHK(hipMemcpyAsync(dst, tmp.data(), n * 4, hipMemcpyHostToDevice, s->ctx->stream));
    HIPCHK(hipStreamSynchronize(s->ctx->stream));
    return KBG_OK;
}

int get_f64(kbg_solve* s, double* dst, const double* src, size_t n) {
    HIPCHK(hipMemcpyAsync(dst, src, n * 8, hipMemcpyDeviceToHost, s->ctx->stream));
    HIPCHK(hipStreamSynchronize(s->ctx->stream));
    return KBG_OK;
}

// host f [N][nf][nv] (FP64) <-> a 1D solve's device f in its own layout and precision
template <class T> void to_layout(const Solve1D& S, int N, const double* src, T* dst) {
    const int nf = S.nf, nv = S.nv;
    for (int k = 0; k < N; k++)
        for (int q = 0; q < nf; q++)
            for (int v = 0; v < nv; v++)
                dst[(size_t)k * S.fcs + (size_t)q * S.fqs + (size_t)v * S.fvs] = T(src[((size_t)k * nf + q) * nv + v]);
}
template <class T> void from_layout(const Solve1D& S, int N, const T* src, double* dst) {
    const int nf = S.nf, nv = S.nv;
    for (int k = 0; k < N; k++)
        for (int q = 0; q < nf; q++)
            for (int v = 0; v < nv; v++)
                dst[((size_t)k * nf + q) * nv + v] = double(src[(size_t)k * S.fcs + (size_t)q * S.fqs + (size_t)v * S.fvs]);
}
size_t f_elems_1d(const Solve1D& S, int N) { return S.swept ? (size_t)S.nf * S.nv * S.np : (size_t)N * S.nf * S.nv; }

int put_f1d(kbg_solve* s, void* dst, const double* src) {
    const Solve1D& S = s->h1;
    const size_t n = f_elems_1d(S, s->N);
    if (s->fp64) {
        std::vector<double> tmp(n, 0.0);
        to_layout<double>(S, s->N, src, tmp.data());
        HIPCHK(hipMemcpyAsync(dst, tmp.data(), n * 8, hipMemcpyHostToDevice, s->ctx->stream));
        HIPCHK(hipStreamSynchronize(s->ctx->stream));
    } else {
        std::vector<float> tmp(n, 0.0f);
        to_layout<float>(S, s->N, src, tmp.data());
        HIPCHK(hipMemcpyAsync(dst, tmp.data(), n * 4, hipMemcpyHostToDevice, s->ctx->stream));
        HIPCHK(hipStreamSynchronize(s->ctx->stream));
    }
    return KBG_OK;
}

int get_f1d(kbg_solve* s, double* dst, const void* src) {
    const Solve1D& S = s->h1;
    const size_t n = f_elems_1d(S, s->N);
    if (s->fp64) {
        std::vector<double> tmp(n);
        HIPCHK(hipMemcpyAsync(tmp.data(), src, n * 8, hipMemcpyDeviceToHost, s->ctx->stream));
        HIPCHK(hipStreamSynchronize(s->ctx->stream));
        from_layout<double>(S, s->N, tmp.data(), dst);
    } else {
        std::vector<float> tmp(n);
        HIPCHK(hipMemcpyAsync(tmp.data(), src, n * 4, hipMemcpyDeviceToHost, s->ctx->stream));
        HIPCHK(hipStreamSynchronize(s->ctx->stream));
        from_layout<float>(S, s->N, tmp.data(), dst);
    }
    return KBG_OK;
}

// FP32 swept solves store the departure f - M~(prim): host FP64 f -> departure in buffers `bufs`
// (prim must be on the device), through an FP64 staging copy in the swept layout
int put_f1d_delta(kbg_solve* s, const double* src, std::initializer_list<int> bufs) {
    const Solve1D& S = s->h1;
    const size_t n = f_elems_1d(S, s->N);
    std::vector<double> tmp(n, 0.0);
    to_layout<double>(S, s->N, src, tmp.data());
    double* stage = nullptr;
    HIPCHK(hipMalloc((void**)&stage, n * 8));
    hipError_t e = hipMemcpyAsync(stage, tmp.data(), n * 8, hipMemcpyHostToDevice, s->ctx->stream);
    for (int b : bufs)
        if (e == hipSuccess) e = kbg::launch_delta1d(S, s->d1, stage, 0, b, s->ctx->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(s->ctx->stream);
    (void)hipFree(stage);
    if (e != hipSuccess) return fail(KBG_ERR_HIP, std::string("departure upload: ") + hipGetErrorString(e));
    s->rep_full = false;
    return KBG_OK;
}

// f of buffer `buf` as FP64 on the host, exactly M~ + departure
int get_f1d_delta(kbg_solve* s, double* dst, int buf) {
    const Solve1D& S = s->h1;
    const size_t n = f_elems_1d(S, s->N);
    std::vector<double> tmp(n);
    double* stage = nullptr;
    HIPCHK(hipMalloc((void**)&stage, n * 8));
    hipError_t e = hipMemsetAsync(stage, 0, n * 8, s->ctx->stream);
    if (e == hipSuccess) e = kbg::launch_delta1d(S, s->d1, stage, 1, buf, s->ctx->stream);
    if (e == hipSuccess) e = hipMemcpyAsync(tmp.data(), stage, n * 8, hipMemcpyDeviceToHost, s->ctx->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(s->ctx->stream);
    (void)hipFree(stage);
    if (e != hipSuccess) return fail(KBG_ERR_HIP, std::string("departure download: ") + hipGetErrorString(e));
    from_layout<double>(S, s->N, tmp.data(), dst);
    return KBG_OK;
}

// the phase-level kernels work on f itself: convert a departure-form solve once (in place, FP32)
int to_full_rep(kbg_solve* s) {
    if (s->rep_full) return KBG_OK;
    if (s->d.dim == 1 && s->h1.delta) {
        HIPCHK(kbg::launch_delta1d(s->h1, s->d1, nullptr, 2, s->h1.cur, s->ctx->stream));
        s->rep_full = true;
    } else if (s->d.dim == 2 && s->h2.delta) {
        HIPCHK(kbg::launch_delta2d(s->h2, s->d2, 2, s->ctx->stream));
        s->rep_full = true;
    }
    return KBG_OK;
}

// The reference Maxwellian of a 2D departure solve: the Maxwellian of the mean conserved state over
// the cells (host w) with rho, T and U / sqrt(T), V / sqrt(T) rounded to a grid of 1/16 (in log2 for
// rho and T), so a state uploaded again later in the run (a checkpoint) usually gives the same
// reference; b = h K / (2 lambda); rounded to FP32 as stored; its moments in FP64
void set_mref2d(kbg_solve* s, const double* w) {
    const kbg_desc& d = s->d;
    Solve2D& S = s->h2;
    const size_t C = s->N;
    double wm[4] = {0, 0, 0, 0}, p[4];
    for (size_t c = 0; c < C; c++)
        for (int m = 0; m < 4; m++) wm[m] += w[c * 4 + m];
    for (int m = 0; m < 4; m++) wm[m] /= double(C);
    kbg::conserve_prim<4>(wm, d.gamma, p);
    {
        const double q = 16.0;
        const double T = std::exp2(std::round(q * std::log2(0.5 / p[3])) / q), c = std::sqrt(T);
        p[0] = std::exp2(std::round(q * std::log2(p[0])) / q);
        p[1] = std::round(q * p[1] / c) / q * c;
        p[2] = std::round(q * p[2] / c) / q * c;
        p[3] = 0.5 / T;
    }
    const int nv = d.nv;
    s->mref.assign(2 * (size_t)nv, 0.0f);
    const double kb = S.K / (2.0 * p[3]);
    for (int m = 0; m < 11; m++) S.mmom[m] = 0.0;
    for (int v = 0; v < nv; v++) {
        const double u = d.u[v], vv = d.v[v], wq = d.weights[v];
        const double cu = u - p[1], cv = vv - p[2];
        const float mh = float(p[0] * (p[3] / kbg::kPi) * std::exp(-p[3] * (cu * cu + cv * cv)));
        const float mb = float(double(mh) * kb);
        s->mref[v] = mh;
        s->mref[nv + v] = mb;
        const double a = wq * double(mh), b = wq * double(mb);
        S.mmom[0] += a * u;
        S.mmom[1] += a * vv;
        S.mmom[2] += a * u * u;
        S.mmom[3] += a * u * vv;
        S.mmom[4] += a * vv * vv;
        S.mmom[5] += a * u * u * u;
        S.mmom[6] += a * u * u * vv;
        S.mmom[7] += a * u * vv * vv;
        S.mmom[8] += a * vv * vv * vv;
        S.mmom[9] += b * u;
        S.mmom[10] += b * vv;
    }
}

// per-velocity constants of the sweep for time step dt, in chunks of 8 (kbg_sweep.h VChunk)
template <class T> size_t vtab_bytes(int nv) { return sizeof(kbg::VChunk<T>) * kbg::vchunks(nv); }
template <class T> void build_vtab(const kbg_desc& d, double dt, std::vector<char>& out) {
    const int nc = kbg::vchunks(d.nv);
    out.assign(vtab_bytes<T>(d.nv), 0);
    auto* ch = reinterpret_cast<kbg::VChunk<T>*>(out.data());
    for (int v = 0; v < nc * kbg::kVChunk; v++) {
        kbg::VChunk<T>& c = ch[v / kbg::kVChunk];
        const int j = v % kbg::kVChunk;
        if (v >= d.nv) {
            c.upmask |= 1 << j;
            continue;
        }
        const double u = d.u[v], w = d.weights[v];
        double p = w;
        for (int m = 0; m < 5; m++, p *= u) c.wk[m][j] = T(p);
        c.u[j] = T(u);
        c.uu[j] = T(u * u);
        c.cu[j] = T(dt * u);
        c.cs[j] = T(-0.5 * dt * dt * (u * u));
        c.cs2[j] = T(-dt * dt * (u * u));
        if (u >= 0.0) c.upmask |= 1 << j;
    }
    // equally spaced chunks (to 1e-12 relative) carry their node spacing
    for (int c = 0; c < nc; c++) {
        const int v0 = c * kbg::kVChunk, v1 = std::min(d.nv, v0 + kbg::kVChunk) - 1;
        if (v1 - v0 < 1) continue;
        const double du = (d.u[v1] - d.u[v0]) / (v1 - v0);
        bool eq = du > 0.0;
        for (int v = v0; v < v1 && eq; v++) eq = std::fabs(d.u[v + 1] - d.u[v] - du) <= 1e-12 * du;
        if (eq) ch[c].du = T(du);
    }
}

// st: the stream of the calling batch (a solve may belong to another context of the device), or
// nullptr for the solve's own context stream
int ensure_vtab(kbg_solve* s, double dt, hipStream_t st = nullptr) {
    if (!s->fused || s->vtab_dt == dt) return KBG_OK;
    std::vector<char> tab;
    if (s->fp64) build_vtab<double>(s->d, dt, tab);
    else build_vtab<float>(s->d, dt, tab);
    const hipStream_t q = st ? st : s->ctx->stream;
    HIPCHK(hipMemcpyAsync(s->vtab, tab.data(), tab.size(), hipMemcpyHostToDevice, q));
    HIPCHK(hipStreamSynchronize(q));
    s->vtab_dt = dt;
    return KBG_OK;
}

int validate(const kbg_desc* d) {
    if (!d) return fail(KBG_ERR_ARG, "desc is NULL");
    if (d->precision != KBG_FP32 && d->precision != KBG_FP64) return fail(KBG_ERR_ARG, "precision must be KBG_FP32 or KBG_FP64");
    if (d->flux < KBG_KFVS || d->flux > KBG_UGKS) return fail(KBG_ERR_ARG, "unknown flux");
    if (d->collision != KBG_BGK && d->collision != KBG_SHAKHOV && d->collision != KBG_FSM) return fail(KBG_ERR_ARG, "unknown collision");
    if (d->collision == KBG_FSM) {
        if (d->dim != 1 || d->vdim != 3 || d->nf != 1)
            return fail(KBG_ERR_UNSUPPORTED, "the fast spectral collision (:fsm) is KitBase's 1D1F3V step! (step_1f.jl:122-157)");
        if (d->fsm_nm < 2 || !(d->fsm_kn > 0.0) || !d->fsm_phi || !d->fsm_psi || !d->fsm_phipsi)
            return fail(KBG_ERR_ARG, ":fsm needs fsm_nm >= 2, fsm_kn > 0 and the kernel modes phi, psi, phipsi");
        if (d->fsm_shape[0] < 2 || d->fsm_shape[1] < 2 || d->fsm_shape[2] < 2 ||
            (long)d->fsm_shape[0] * d->fsm_shape[1] * d->fsm_shape[2] != d->nv)
            return fail(KBG_ERR_ARG, ":fsm needs fsm_shape = (nu, nv, nw) with nu nv nw = nv");
    }
    if (d->limiter < KBG_VANLEER || d->limiter > KBG_LINEAR) return fail(KBG_ERR_ARG, "unknown limiter");
    if (d->interp_order != 1 && d->interp_order != 2) return fail(KBG_ERR_ARG, "interp_order must be 1 or 2");
    if (d->nv < 1) return fail(KBG_ERR_ARG, "nv must be positive");
    if (!d->u || !d->weights || !d->dx) return fail(KBG_ERR_ARG, "u, weights and dx are required");
    if (d->dim == 1) {
        if (d->nx < 3) return fail(KBG_ERR_ARG, "1D needs nx >= 3");
        if (d->ng < 0) return fail(KBG_ERR_ARG, "ng must be >= 0");
        if (d->vdim == 1) {
            if (d->nf != 1 && d->nf != 2) return fail(KBG_ERR_ARG, "1D1V needs nf = 1 or 2");
            if (d->flux == KBG_UGKS && d->nf != 2) return fail(KBG_ERR_UNSUPPORTED, "KitBase has no 1D1F1V UGKS flux (flux_ugks.jl covers 1D2F1V)");
        } else if (d->vdim == 3) {
            if (d->nf != 1) return fail(KBG_ERR_UNSUPPORTED, "1D3V supports nf = 1 (1D1F3V)");
            if (!d->v || !d->w) return fail(KBG_ERR_ARG, "1D3V needs v and w");
            if (d->flux == KBG_UGKS) return fail(KBG_ERR_UNSUPPORTED, "KitBase has no 1D1F3V UGKS flux");
            if (d->collision == KBG_SHAKHOV)
                return fail(KBG_ERR_UNSUPPORTED, "1D1F3V Shakhov step! references an undefined K (step_1f.jl:91); only BGK is defined");
        } else {
            return fail(KBG_ERR_ARG, "1D needs vdim 1 or 3");
        }
        for (int sd = 0; sd < 2; sd++)
            if (d->bc[sd] != KBG_BC_FIX && d->bc[sd] != KBG_BC_MAXWELL) return fail(KBG_ERR_UNSUPPORTED, "1D boundaries: fix or maxwell");
        if (d->nf == 2 && (d->bc[0] == KBG_BC_MAXWELL || d->bc[1] == KBG_BC_MAXWELL) && !d->bc_prim)
            return fail(KBG_ERR_ARG, "Maxwell walls need bc_prim");
    } else if (d->dim == 2) {
        if ((d->nf != 1 && d->nf != 2) || d->vdim != 2) return fail(KBG_ERR_UNSUPPORTED, "2D supports 2F2V and 1F2V");
        if (d->nf == 1 && d->flux == KBG_UGKS) return fail(KBG_ERR_UNSUPPORTED, "KitBase has no 2D1F2V UGKS flux (flux_ugks.jl covers 2D2F2V)");
        if (d->nf == 1 && d->collision == KBG_SHAKHOV)
            return fail(KBG_ERR_UNSUPPORTED, "2D1F2V Shakhov step! references an undefined b (step_1f.jl:194); only BGK is defined");
        if (d->nf == 1 && d->K != 0.0)
            return fail(KBG_ERR_UNSUPPORTED, "2D1F2V is built for K = 0 (no internal degrees of freedom)");
        if (d->nx < 3 || d->ny < 3) return fail(KBG_ERR_ARG, "2D needs nx, ny >= 3");
        if (d->ng != 0) return fail(KBG_ERR_UNSUPPORTED, "2D ghost layers are not supported");
        if (!d->v || !d->dy || !d->xgeo || !d->ygeo) return fail(KBG_ERR_ARG, "2D needs v, dy, xgeo, ygeo");
        for (int sd = 0; sd < 4; sd++) {
            if (d->bc[sd] != KBG_BC_FIX && d->bc[sd] != KBG_BC_MAXWELL) return fail(KBG_ERR_UNSUPPORTED, "2D boundaries: fix or maxwell");
            if (d->bc[sd] == KBG_BC_MAXWELL && (!d->xwall || !d->ywall)) return fail(KBG_ERR_ARG, "Maxwell walls need xwall and ywall");
        }
    } else {
        return fail(KBG_ERR_ARG, "dim must be 1 or 2");
    }
    return KBG_OK;
}

int nw_of(const kbg_desc* d) { return d->dim == 2 ? 4 : (d->vdim == 3 ? 5 : 3); }

// the fused sweep takes one velocity dimension, up to kSweepMaxWaves wavefronts of velocities, and f buffers
// addressed with 31-bit byte offsets
bool swept_1d(const kbg_desc& d) {
    const size_t N = size_t(d.nx) + 2 * size_t(d.ng), np = (N + 31) / 32 * 32;
    const size_t bytes = np * d.nf * d.nv * (d.precision == KBG_FP64 ? 8 : 4);
    return d.dim == 1 && d.vdim == 1 && d.nv <= kbg::sweep_max_nv(d.precision == KBG_FP64, d.nf) && bytes < (size_t(1) << 31);
}

void fill1d(kbg_solve* s) {
    const kbg_desc& d = s->d;
    Solve1D& S = s->h1;
    S.nx = d.nx;
    S.ng = d.ng;
    S.nv = d.nv;
    S.nf = d.nf;
    S.vdim = d.vdim;
    S.nw = s->nw;
    S.flux = d.flux;
    S.coll = d.collision;
    S.limiter = d.limiter;
    S.interp2 = d.interp_order == 2;
    S.upd_lo = d.bc[0] == KBG_BC_FIX ? 2 : 1;
    S.upd_hi = d.bc[1] == KBG_BC_FIX ? d.nx - 1 : d.nx;
    if (S.interp2) {
        S.slope_lo = d.ng > 0 ? 1 : 2;
        S.slope_hi = d.ng > 0 ? d.nx : d.nx - 1;
    } else {
        S.slope_lo = 1;
        S.slope_hi = 0;
    }
    S.face_lo = d.ng > 0 ? 1 : 2;
    S.face_hi = d.ng > 0 ? d.nx + 1 : d.nx;
    S.wallL = d.nf == 2 && d.bc[0] == KBG_BC_MAXWELL;
    S.wallR = d.nf == 2 && d.bc[1] == KBG_BC_MAXWELL;
    S.faceL_buf = d.ng == 0 || S.wallL;
    S.faceR_buf = d.ng == 0 || S.wallR;
    S.cur = 0;
    S.K = d.K;
    S.gam = d.gamma;
    S.mu = d.mu_ref;
    S.om = d.omega;
    S.Pr = d.Pr;
    S.dt = 0.0;
    // the fused sweep takes one velocity dimension and up to kSweepMaxWaves wavefronts of velocities
    const int N = d.nx + 2 * d.ng;
    S.swept = swept_1d(d);
    S.delta = S.swept && d.precision == KBG_FP32;
    if (S.swept) {
        S.np = (N + 31) / 32 * 32;
        S.fcs = 1;
        S.fvs = S.np;
        S.fqs = (long)d.nv * S.np;
    } else {
        S.np = N;
        S.fcs = (long)d.nf * d.nv;
        S.fqs = d.nv;
        S.fvs = 1;
    }
}

void fill2d(kbg_solve* s) {
    const kbg_desc& d = s->d;
    Solve2D& S = s->h2;
    S.nx = d.nx;
    S.ny = d.ny;
    S.nv = d.nv;
    S.flux = d.flux;
    S.coll = d.collision;
    S.limiter = d.limiter;
    S.interp2 = d.interp_order == 2;
    for (int k = 0; k < 4; k++) S.bc[k] = d.bc[k];
    // 2D1F2V (K = 0) runs in the two-distribution layout with b = 0: with no internal degrees of
    // freedom every b term vanishes, which leaves KitBase's 1F2V step (step_1f.jl:164-221), KFVS / KCU
    // flux (flux_kfvs.jl:615-647, flux_kcu.jl:828-891) and wall (flux_boundary.jl:160-183)
    S.K = d.K;
    S.gam = d.gamma;
    S.mu = d.mu_ref;
    S.om = d.omega;
    S.Pr = d.Pr;
    S.dt = 0.0;
}

// the solve's device descriptor from its host copy, on st (nullptr: the solve's context stream)
int push1d(kbg_solve* s, hipStream_t st = nullptr) {
    const hipStream_t q = st ? st : s->ctx->stream;
    HIPCHK(hipMemcpyAsync(s->d1, &s->h1, sizeof(Solve1D), hipMemcpyHostToDevice, q));
    HIPCHK(hipStreamSynchronize(q));
    return KBG_OK;
}

int push2d(kbg_solve* s, hipStream_t st = nullptr) {
    const hipStream_t q = st ? st : s->ctx->stream;
    HIPCHK(hipMemcpyAsync(s->d2, &s->h2, sizeof(Solve2D), hipMemcpyHostToDevice, q));
    HIPCHK(hipStreamSynchronize(q));
    return KBG_OK;
}

int pushus(kbg_solve* s, hipStream_t st = nullptr) {
    const hipStream_t q = st ? st : s->ctx->stream;
    HIPCHK(hipMemcpyAsync(s->dus, &s->hus, sizeof(kbg::SolveUS), hipMemcpyHostToDevice, q));
    HIPCHK(hipStreamSynchronize(q));
    return KBG_OK;
}

// one fused evolve! + update! of an unstructured solve on stream st, reading f from the buffer of
// relative parity par and writing the other (rows: keep the residual rows)
int step_us(kbg_solve* s, int par, int rows, hipStream_t st) {
    HIPCHK(kbg::launch_ustep(s->fp64, s->thru, s->hus, s->dus, par, rows, st));
    HIPCHK(kbg::launch_uavg(s->fp64, s->hus, s->dus, par + 1, st));
    return KBG_OK;
}

void* us_f(const kbg_solve* s) { return s->hus.cur ? s->hus.f1 : s->hus.f; }

constexpr int kUMaxNv = 4096;  // 1024 threads x 4 velocities

// slope / face buffers of the phase-exact path, allocated on first use
int ensure_phase_buffers(kbg_solve* s) {
    int rc;
    if (s->d.dim == 1) {
        Solve1D& S = s->h1;
        if (S.sw) return KBG_OK;
        const size_t N = s->N, nf1 = size_t(s->d.nx) + 1;
        void* p;
        if ((rc = dev_alloc(s, &p, N * s->nw * 8))) return rc;
        S.sw = (double*)p;
        HIPCHK(hipMemsetAsync(S.sw, 0, N * s->nw * 8, s->ctx->stream));
        if ((rc = dev_alloc(s, &p, N * s->RS * s->esz))) return rc;
        S.sf = p;
        HIPCHK(hipMemsetAsync(S.sf, 0, N * s->RS * s->esz, s->ctx->stream));
        if ((rc = dev_alloc(s, &p, nf1 * s->nw * 8))) return rc;
        S.fw = (double*)p;
        HIPCHK(hipMemsetAsync(S.fw, 0, nf1 * s->nw * 8, s->ctx->stream));
        if ((rc = dev_alloc(s, &p, nf1 * s->RS * s->esz))) return rc;
        S.ff = p;
        HIPCHK(hipMemsetAsync(S.ff, 0, nf1 * s->RS * s->esz, s->ctx->stream));
        // the stored faces 1 and nx+1 start from the fused path's boundary-face buffers
        HIPCHK(hipMemcpyAsync(S.fw, S.bfw, s->nw * 8, hipMemcpyDeviceToDevice, s->ctx->stream));
        HIPCHK(hipMemcpyAsync(S.fw + (nf1 - 1) * s->nw, S.bfw + s->nw, s->nw * 8, hipMemcpyDeviceToDevice, s->ctx->stream));
        HIPCHK(hipMemcpyAsync(S.ff, S.bff, s->RS * s->esz, hipMemcpyDeviceToDevice, s->ctx->stream));
        HIPCHK(hipMemcpyAsync((char*)S.ff + (nf1 - 1) * s->RS * s->esz, (char*)S.bff + s->RS * s->esz, s->RS * s->esz,
                              hipMemcpyDeviceToDevice, s->ctx->stream));
        return push1d(s);
    }
    Solve2D& S = s->h2;
    if (S.sf) return KBG_OK;
    const size_t C = s->N, nv = s->d.nv;
    struct {
        void** dst;
        size_t bytes;
    } bufs[] = {
        {(void**)&S.sw, C * 8 * 8},
        {&S.sf, C * 4 * nv * s->esz},
    };
    for (auto& b : bufs) {
        void* p;
        if ((rc = dev_alloc(s, &p, b.bytes))) return rc;
        HIPCHK(hipMemsetAsync(p, 0, b.bytes, s->ctx->stream));
        *b.dst = p;
    }
    return push2d(s);
}



// boundary-face buffers follow the stored faces 1 and nx+1 after phase-level calls
int sync_boundary_faces(kbg_solve* s) {
    Solve1D& S = s->h1;
    if (!S.fw) return KBG_OK;
    const size_t nf1 = size_t(s->d.nx) + 1;
    HIPCHK(hipMemcpyAsync(S.bfw, S.fw, s->nw * 8, hipMemcpyDeviceToDevice, s->ctx->stream));
    HIPCHK(hipMemcpyAsync(S.bfw + s->nw, S.fw + (nf1 - 1) * s->nw, s->nw * 8, hipMemcpyDeviceToDevice, s->ctx->stream));
    HIPCHK(hipMemcpyAsync(S.bff, S.ff, s->RS * s->esz, hipMemcpyDeviceToDevice, s->ctx->stream));
    HIPCHK(hipMemcpyAsync((char*)S.bff + s->RS * s->esz, (char*)S.ff + (nf1 - 1) * s->RS * s->esz, s->RS * s->esz,
                          hipMemcpyDeviceToDevice, s->ctx->stream));
    return KBG_OK;
}

// segments of the stepped cells, in order along x (neighbouring segments share halo rows)
void build_items(kbg_solve* s, int seg, int solve_index, int nw, std::vector<WorkItem>& out) {
    const Solve1D& S = s->h1;
    for (int c0 = S.upd_lo; c0 <= S.upd_hi; c0 += seg) {
        WorkItem wi;
        wi.solve = solve_index;
        wi.c0 = c0;
        wi.c1 = std::min(c0 + seg - 1, S.upd_hi);
        wi.nw = nw;
        out.push_back(wi);
    }
}

int step_phase_1d(kbg_solve* s, double dt, double* residual, hipStream_t st);
int step_phase_2d(kbg_solve* s, double dt, double* residual, hipStream_t st);
int batch_step_impl(kbg_ctx* ctx, kbg_solve* const* solves, int n, const double* dts, int nsteps, double* residuals, bool noalloc);

// Staging of kbg_batch_step_noalloc: a fixed device array per GPU, held until the process ends.
// Host code never passes `scratch` itself to a kernel or a copy (in host code the name is a host-side
// stand-in, not a device address): kbg_ctx_create looks up each device's address once, with that
// device current, and every use goes through scratch_ptr(device).
constexpr size_t kNoallocScratchBytes = KBG_NOALLOC_SCRATCH_BYTES;
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
void* g_scratch_addr[kMaxDevices] = {};

void* scratch_ptr(int device) {
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    return device >= 0 && device < kMaxDevices ? g_scratch_addr[device] : nullptr;
}

}  // namespace

static __device__ __align__(256) char scratch[kNoallocScratchBytes];

namespace {

// called by kbg_ctx_create with `device` current
int setup_scratch(int device) {
    if (device < 0 || device >= kMaxDevices) return fail(KBG_ERR_ARG, "device index beyond the scratch table");
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (g_scratch_addr[device]) return KBG_OK;
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p) return fail(KBG_ERR_HIP, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
    g_scratch_addr[device] = p;
    return KBG_OK;
}

}  // namespace

extern "C" {

int kbg_version(void) { return 1 * 10000 + 0 * 100 + 0; }

const char* kbg_last_error(void) { return g_err.c_str(); }

int kbg_ctx_create(int device, kbg_ctx** out) {
    if (!out) return fail(KBG_ERR_ARG, "out is NULL");
    int n = 0;
    hipError_t e = hipGetDeviceCount(&n);
    if (e != hipSuccess || n <= 0) return fail(KBG_ERR_HIP, "no HIP device available");
    if (device < 0 || device >= n) return fail(KBG_ERR_ARG, "device index out of range");
    HIPCHK(hipSetDevice(device));
    if (int rc = setup_scratch(device)) return rc;
    kbg_ctx* c = new kbg_ctx();
    c->device = device;
    e = hipStreamCreate(&c->stream);
    for (auto& sd : c->side)
        if (e == hipSuccess) e = hipStreamCreate(&sd);
    for (auto& ev : c->ev)
        if (e == hipSuccess) e = hipEventCreate(&ev);
    if (e == hipSuccess) e = hipEventCreateWithFlags(&c->fork, hipEventDisableTiming);
    for (auto& ev : c->join)
        if (e == hipSuccess) e = hipEventCreateWithFlags(&ev, hipEventDisableTiming);
    if (e == hipSuccess) e = hipMalloc((void**)&c->d_res, 4096 * 5 * sizeof(double));
    if (e != hipSuccess) {
        kbg_ctx_destroy(c);
        return fail(KBG_ERR_HIP, std::string("kbg_ctx_create: ") + hipGetErrorString(e));
    }
    *out = c;
    return KBG_OK;
}

int kbg_ctx_destroy(kbg_ctx* c) {
    if (!c) return KBG_OK;
    if (c->stream) (void)hipStreamSynchronize(c->stream);
    for (auto& sd : c->side)
        if (sd) (void)hipStreamSynchronize(sd);
    for (auto& ev : c->ev)
        if (ev) (void)hipEventDestroy(ev);
    for (auto& ev : c->join)
        if (ev) (void)hipEventDestroy(ev);
    if (c->fork) (void)hipEventDestroy(c->fork);
    c->drop_graphs();
    (void)hipFree(c->d_descs);
    (void)hipFree(c->d_items);
    (void)hipFree(c->d_rpart);
    (void)hipFree(c->d_packs);
    (void)hipFree(c->d_res);
    (void)hipFree(c->d_descs2);
    (void)hipFree(c->d_pref2);
    for (int b = 0; b < 2; b++) {
        if (c->pin[b]) (void)hipHostFree(c->pin[b]);
        if (c->pin_ev[b]) (void)hipEventDestroy(c->pin_ev[b]);
    }
    for (auto& sd : c->side)
        if (sd) (void)hipStreamDestroy(sd);
    if (c->stream) (void)hipStreamDestroy(c->stream);
    delete c;
    return KBG_OK;
}

int kbg_ctx_sync(kbg_ctx* c) {
    if (!c) return fail(KBG_ERR_ARG, "ctx is NULL");
    HIPCHK(hipStreamSynchronize(c->stream));
    return KBG_OK;
}

int kbg_ctx_set_schedule(kbg_ctx* c, int seg_cells, int streams) {
    if (!c) return fail(KBG_ERR_ARG, "ctx is NULL");
    // a sweep workgroup's wavefronts hold 64 cells: the segment plus two halo cells on each side
    if (seg_cells < 1 || seg_cells > kMaxSeg)
        return fail(KBG_ERR_ARG, "schedule: seg_cells must be in [1, " + std::to_string(kMaxSeg) + "]");
    if (streams < 0 || streams > 1 + kbg_ctx::kSide)
        return fail(KBG_ERR_ARG, "schedule: streams must be in [0, " + std::to_string(1 + kbg_ctx::kSide) + "]");
    c->seg_cells = seg_cells;
    c->streams = streams;
    return KBG_OK;
}

int kbg_ctx_get_schedule(kbg_ctx* c, int* seg_cells, int* streams) {
    if (!c) return fail(KBG_ERR_ARG, "ctx is NULL");
    if (seg_cells) *seg_cells = c->seg_cells;
    if (streams) *streams = c->streams;
    return KBG_OK;
}

int kbg_ctx_set_graphs(kbg_ctx* ctx, int enable) {
    if (!ctx) return fail(KBG_ERR_ARG, "ctx is NULL");
    ctx->graphs = enable != 0;
    if (!ctx->graphs) {
        HIPCHK(hipStreamSynchronize(ctx->stream));
        ctx->drop_graphs();
    }
    return KBG_OK;
}

int kbg_ctx_set_sweep_steps(kbg_ctx* ctx, int steps) {
    if (!ctx) return fail(KBG_ERR_ARG, "ctx is NULL");
    if (steps != 1 && steps != 2) return fail(KBG_ERR_ARG, "sweep steps must be 1 or 2");
    ctx->sweep_steps = steps;
    return KBG_OK;
}

int kbg_ctx_set_sweep_vw(kbg_ctx* ctx, int vw) {
    if (!ctx) return fail(KBG_ERR_ARG, "ctx is NULL");
    if (vw != 0 && vw != 8 && vw != 16 && vw != 32 && vw != 64) return fail(KBG_ERR_ARG, "sweep_vw must be 0, 8, 16, 32 or 64");
    ctx->sweep_vw = vw;
    return KBG_OK;
}

int kbg_ctx_event_record(kbg_ctx* c, int slot) {
    if (!c || slot < 0 || slot >= 16) return fail(KBG_ERR_ARG, "bad event slot");
    HIPCHK(hipEventRecord(c->ev[slot], c->stream));
    return KBG_OK;
}

int kbg_ctx_event_elapsed(kbg_ctx* c, int a, int b, float* ms) {
    if (!c || a < 0 || a >= 16 || b < 0 || b >= 16 || !ms) return fail(KBG_ERR_ARG, "bad event slot");
    HIPCHK(hipEventSynchronize(c->ev[b]));
    HIPCHK(hipEventElapsedTime(ms, c->ev[a], c->ev[b]));
    return KBG_OK;
}

size_t kbg_solve_bytes(const kbg_desc* d) {
    if (!d) return 0;
    const size_t esz = d->precision == KBG_FP64 ? 8 : 4;
    const int nw = nw_of(d);
    if (d->dim == 1) {
        const size_t N = size_t(d->nx) + 2 * size_t(d->ng);
        const bool swept = swept_1d(*d);
        const size_t np = swept ? (N + 31) / 32 * 32 : N;
        const size_t vt = swept ? (esz == 8 ? vtab_bytes<double>(d->nv) : vtab_bytes<float>(d->nv)) : 0;
        size_t fsm = 0;
        if (d->collision == KBG_FSM && d->nx > 2) {
            // spectrum, the transforms' products, kernel modes (all in the storage precision), Q (FP64);
            // hipFFT's own work area not counted
            const size_t nc = size_t(d->nx) - 2, n3 = size_t(d->nv), nd = size_t(d->fsm_nm) * (d->fsm_nm - 1);
            const int cn = kbg::fsm_cube_size(d->fsm_shape[0], d->fsm_shape[1], d->fsm_shape[2]);
            const size_t work = (cn && esz == 4 ? nc * (nd + 1) * n3 * 4 : nc * (2 * nd + 2) * n3 * 2 * esz) + size_t(cn) * 2 * esz;
            fsm = nc * n3 * 2 * esz + work + (2 * nd + 1) * n3 * esz + nc * n3 * 8;
        }
        return 2 * np * d->nf * d->nv * esz + 3 * N * nw * 8 + N * 8 + 4 * size_t(d->nv) * esz + 2 * d->nf * d->nv * esz +
               N * 4 * nw * 8 + vt + fsm;
    }
    const size_t C = size_t(d->nx) * d->ny, F = size_t(d->nx + 1) * d->ny + size_t(d->nx) * (d->ny + 1);
    const size_t nv = d->nv;
    // f, face fluxes, w / prim / residual rows / dx / dy, geometry + walls + fw, velocity grid
    size_t b = C * 2 * nv * esz + F * 2 * nv * esz + C * (2 * 4 * 8 + 16 * 8 + 2 * 8) + F * (3 + 4 + 4) * 8 + 3 * nv * esz;
    if (d->flux != KBG_KFVS) b += F * kbg::kFaceCoef * 8 + (esz == 4 ? F * kbg::kFaceC32 * 4 : 0);
    if (kbg::step2d_threads(d->precision == KBG_FP64, d->nv) == 0) b += C * 8 * 8 + C * 4 * nv * esz;  // slopes
    else if (esz == 4) b += 2 * nv * 4 + F * 8;  // departure form: reference Maxwellian, per-face coefficients
    return b;
}

int kbg_solve_create(kbg_ctx* ctx, const kbg_desc* desc, kbg_solve** out) {
    if (!ctx || !out) return fail(KBG_ERR_ARG, "ctx or out is NULL");
    int rc = validate(desc);
    if (rc) return rc;
    HIPCHK(hipSetDevice(ctx->device));
    kbg_solve* s = new kbg_solve();
    s->ctx = ctx;
    s->d = *desc;
    {
        kbg_desc& dd = s->d;
        auto keep = [](std::vector<double>& v, const double*& ptr, size_t n) {
            if (!ptr) return;
            v.assign(ptr, ptr + n);
            ptr = v.data();
        };
        const size_t nv = dd.nv;
        keep(s->hu, dd.u, nv);
        keep(s->hv, dd.v, nv);
        keep(s->hw, dd.w, nv);
        keep(s->hwt, dd.weights, nv);
        if (dd.dim == 1) {
            keep(s->hdx, dd.dx, size_t(dd.nx) + 2 * size_t(dd.ng));
            keep(s->hbc, dd.bc_prim, 2 * size_t(nw_of(desc)));
            dd.dy = dd.xgeo = dd.ygeo = dd.xwall = dd.ywall = nullptr;
        } else {
            const size_t C = size_t(dd.nx) * dd.ny, FX = size_t(dd.nx + 1) * dd.ny, FY = size_t(dd.nx) * (dd.ny + 1);
            keep(s->hdx, dd.dx, C);
            keep(s->hdy, dd.dy, C);
            keep(s->hxgeo, dd.xgeo, FX * 3);
            keep(s->hygeo, dd.ygeo, FY * 3);
            keep(s->hxwall, dd.xwall, FX * 4);
            keep(s->hywall, dd.ywall, FY * 4);
            dd.bc_prim = nullptr;
        }
    }
    s->fp64 = desc->precision == KBG_FP64;
    s->esz = s->fp64 ? 8 : 4;
    s->nw = nw_of(desc);
    const kbg_desc& d = s->d;
    s->RS = d.nf * d.nv;
    auto bail = [&](int code) {
        kbg_solve_destroy(s);
        return code;
    };
    void* p;
    // velocity grid
    if ((rc = dev_alloc(s, &s->u, d.nv * s->esz)) || (rc = put_vec(s, s->u, d.u, d.nv))) return bail(rc);
    if ((rc = dev_alloc(s, &s->wt, d.nv * s->esz)) || (rc = put_vec(s, s->wt, d.weights, d.nv))) return bail(rc);
    if (d.v) {
        if ((rc = dev_alloc(s, &s->v, d.nv * s->esz)) || (rc = put_vec(s, s->v, d.v, d.nv))) return bail(rc);
    }
    if (d.w && d.dim == 1 && d.vdim == 3) {
        if ((rc = dev_alloc(s, &s->wv, d.nv * s->esz)) || (rc = put_vec(s, s->wv, d.w, d.nv))) return bail(rc);
    }
    if (d.dim == 1) {
        s->N = d.nx + 2 * d.ng;
        const size_t N = s->N;
        fill1d(s);
        Solve1D& S = s->h1;
        if ((rc = dev_alloc(s, &p, N * 8))) return bail(rc);
        s->dx = (double*)p;
        if ((rc = put_f64(s, s->dx, d.dx, N))) return bail(rc);
        const size_t fel = f_elems_1d(S, (int)N);
        for (int b = 0; b < 2; b++) {
            if ((rc = dev_alloc(s, &p, N * s->nw * 8))) return bail(rc);
            S.w[b] = (double*)p;
            HIPCHK(hipMemset(S.w[b], 0, N * s->nw * 8));
            if ((rc = dev_alloc(s, &p, fel * s->esz))) return bail(rc);
            S.f[b] = p;
            HIPCHK(hipMemset(S.f[b], 0, fel * s->esz));
        }
        if ((rc = dev_alloc(s, &p, N * s->nw * 8))) return bail(rc);
        S.prim = (double*)p;
        HIPCHK(hipMemset(S.prim, 0, N * s->nw * 8));
        if ((rc = dev_alloc(s, &p, 2 * s->nw * 8))) return bail(rc);
        S.bfw = (double*)p;
        HIPCHK(hipMemset(S.bfw, 0, 2 * s->nw * 8));
        if ((rc = dev_alloc(s, &p, 2 * s->RS * s->esz))) return bail(rc);
        S.bff = p;
        HIPCHK(hipMemset(S.bff, 0, 2 * s->RS * s->esz));
        if ((rc = dev_alloc(s, &p, 2 * s->nw * 8))) return bail(rc);
        S.bcprim = (const double*)p;
        if (d.bc_prim) {
            if ((rc = put_f64(s, (double*)p, d.bc_prim, 2 * s->nw))) return bail(rc);
        } else {
            HIPCHK(hipMemset(p, 0, 2 * s->nw * 8));
        }
        S.dx = s->dx;
        S.u = s->u;
        S.v = s->v;
        S.wv = s->wv;
        S.wt = s->wt;
        if ((rc = dev_alloc(s, &p, N * 4 * s->nw * 8))) return bail(rc);
        s->rcell = (double*)p;
        if ((rc = dev_alloc(s, &p, 8 * 8))) return bail(rc);
        s->res = (double*)p;
        if ((rc = dev_alloc(s, &p, 8))) return bail(rc);
        HIPCHK(hipMemset(p, 0, 8));
        S.health = (int*)p;
        if ((rc = dev_alloc(s, &p, sizeof(Solve1D)))) return bail(rc);
        s->d1 = (Solve1D*)p;
        s->fused = S.swept != 0;
        if (s->fused) {
            s->sweep_nw = kbg::sweep_nw(s->fp64, d.nf, d.nv);
            const size_t vb = s->fp64 ? vtab_bytes<double>(d.nv) : vtab_bytes<float>(d.nv);
            if ((rc = dev_alloc(s, &s->vtab, vb))) return bail(rc);
            HIPCHK(hipMemset(s->vtab, 0, vb));
            S.vtab = s->vtab;
            if ((rc = ensure_vtab(s, 0.0))) return bail(rc);
        }
        if (d.collision == KBG_FSM) {
            // Q for the interior cells 2 .. nx-1 (update! steps them with :fsm; boundary cells take BGK)
            kbg::FsmPlan& P = s->fsm;
            P.nu = d.fsm_shape[0];
            P.nv = d.fsm_shape[1];
            P.nw = d.fsm_shape[2];
            P.M = d.fsm_nm;
            P.kn = d.fsm_kn;
            P.c0 = 2;
            P.ncell = d.nx - 2;
            P.f32 = !s->fp64;
            P.cube_n = kbg::fsm_cube_size(P.nu, P.nv, P.nw);
            const size_t n3 = size_t(d.nv), nd = size_t(P.M) * (P.M - 1), nc = size_t(P.ncell), re = P.f32 ? 4 : 8;
            if (S.fvs != 1) return bail(fail(KBG_ERR_STATE, ":fsm needs the velocity-contiguous f layout"));
            // the FP32 LDS path keeps one FP32 product per pair instead of the transformed mode products
            const bool lds = P.cube_n && P.f32;
            const size_t work = lds ? 0 : nc * (2 * nd + 2) * n3 * 2 * re;
            const size_t part = lds ? nc * (nd + 1) * n3 * 4 : 0, twb = size_t(P.cube_n) * 2 * re;
            struct {
                void** dst;
                size_t bytes;
            } bufs[] = {{&P.spec, nc * n3 * 2 * re}, {&P.work, work},       {&P.phi, nd * n3 * re},
                        {&P.psi, nd * n3 * re},      {&P.phipsi, n3 * re},  {(void**)&P.Q, nc * n3 * 8},
                        {(void**)&P.part, part},     {&P.tw, twb}};
            for (auto& b : bufs) {
                if (!b.bytes) continue;
                if ((rc = dev_alloc(s, &p, b.bytes))) return bail(rc);
                *b.dst = p;
            }
            if ((rc = put_real(s, P.phi, d.fsm_phi, nd * n3, P.f32)) || (rc = put_real(s, P.psi, d.fsm_psi, nd * n3, P.f32)) ||
                (rc = put_real(s, P.phipsi, d.fsm_phipsi, n3, P.f32)))
                return bail(rc);
            if (P.cube_n) {
                std::vector<double> tw(2 * P.cube_n);
                for (int j = 0; j < P.cube_n; j++) {
                    tw[2 * j] = std::cos(-2.0 * kbg::kPi * j / P.cube_n);
                    tw[2 * j + 1] = std::sin(-2.0 * kbg::kPi * j / P.cube_n);
                }
                if ((rc = put_real(s, P.tw, tw.data(), tw.size(), P.f32))) return bail(rc);
            }
            HIPCHK(hipMemset(P.Q, 0, nc * n3 * 8));
            if (kbg::fsm_plan_create(P) != hipSuccess) return bail(fail(KBG_ERR_HIP, "hipfftPlanMany failed for the :fsm transforms"));
            s->has_fsm = true;
            S.fsmq = P.Q;
            S.fsm_c0 = P.c0;
        }
        // the descriptor's kernel-mode arrays were copied above; the caller's need not outlive this call
        s->d.fsm_phi = s->d.fsm_psi = s->d.fsm_phipsi = nullptr;
        if ((rc = push1d(s))) return bail(rc);
    } else {
        const size_t C = size_t(d.nx) * d.ny, FX = size_t(d.nx + 1) * d.ny, FY = size_t(d.nx) * (d.ny + 1);
        s->N = (int)C;
        fill2d(s);
        Solve2D& S = s->h2;
        double* dd;
        if ((rc = dev_alloc(s, &p, C * 8)) || (rc = put_f64(s, (double*)p, d.dx, C))) return bail(rc);
        s->dx = (double*)p;
        if ((rc = dev_alloc(s, &p, C * 8)) || (rc = put_f64(s, (double*)p, d.dy, C))) return bail(rc);
        s->dy = (double*)p;
        if ((rc = dev_alloc(s, &p, FX * 3 * 8)) || (rc = put_f64(s, (double*)p, d.xgeo, FX * 3))) return bail(rc);
        S.xgeo = (const double*)p;
        if ((rc = dev_alloc(s, &p, FY * 3 * 8)) || (rc = put_f64(s, (double*)p, d.ygeo, FY * 3))) return bail(rc);
        S.ygeo = (const double*)p;
        if ((rc = dev_alloc(s, &p, FX * 4 * 8))) return bail(rc);
        dd = (double*)p;
        if (d.xwall) {
            if ((rc = put_f64(s, dd, d.xwall, FX * 4))) return bail(rc);
        } else {
            HIPCHK(hipMemset(dd, 0, FX * 4 * 8));
        }
        S.xwall = dd;
        if ((rc = dev_alloc(s, &p, FY * 4 * 8))) return bail(rc);
        dd = (double*)p;
        if (d.ywall) {
            if ((rc = put_f64(s, dd, d.ywall, FY * 4))) return bail(rc);
        } else {
            HIPCHK(hipMemset(dd, 0, FY * 4 * 8));
        }
        S.ywall = dd;
        const bool coef = d.flux != KBG_KFVS;
        struct {
            void** dst;
            size_t bytes;
        } bufs[] = {
            {(void**)&S.w, C * 4 * 8},
            {(void**)&S.prim, C * 4 * 8},
            {&S.f, C * 2 * d.nv * s->esz},
            {(void**)&S.xfw, FX * 4 * 8},
            {&S.xff, FX * 2 * d.nv * s->esz},
            {(void**)&S.yfw, FY * 4 * 8},
            {&S.yff, FY * 2 * d.nv * s->esz},
            {(void**)&S.xcoef, coef ? FX * kbg::kFaceCoef * 8 : 0},
            {(void**)&S.ycoef, coef ? FY * kbg::kFaceCoef * 8 : 0},
            {(void**)&S.xc32, coef && !s->fp64 ? FX * kbg::kFaceC32 * 4 : 0},
            {(void**)&S.yc32, coef && !s->fp64 ? FY * kbg::kFaceC32 * 4 : 0},
            {(void**)&s->rcell, C * 16 * 8},
        };
        for (auto& b : bufs) {
            if (!b.bytes) continue;
            if ((rc = dev_alloc(s, &p, b.bytes))) return bail(rc);
            HIPCHK(hipMemset(p, 0, b.bytes));
            *b.dst = p;
        }
        S.dx = s->dx;
        S.dy = s->dy;
        S.u = s->u;
        S.v = s->v;
        S.wt = s->wt;
        S.rcell = s->rcell;
        // the fused step addresses f and the face fluxes with 32-bit byte offsets
        s->thr2 = FX * 2 * d.nv * s->esz < (1ull << 31) && FY * 2 * d.nv * s->esz < (1ull << 31) ? kbg::step2d_threads(s->fp64, d.nv) : 0;
        // FP32 solves of the fused step store departures from a reference Maxwellian (set by upload)
        S.delta = !s->fp64 && s->thr2 > 0;
        if (S.delta) {
            if ((rc = dev_alloc(s, &p, 2 * size_t(d.nv) * 4))) return bail(rc);
            HIPCHK(hipMemset(p, 0, 2 * size_t(d.nv) * 4));
            S.mref = (const float*)p;
            s->mref.assign(2 * size_t(d.nv), 0.0f);
            if ((rc = dev_alloc(s, &p, FX * 8))) return bail(rc);
            HIPCHK(hipMemset(p, 0, FX * 8));
            S.xcref = (double*)p;
            if ((rc = dev_alloc(s, &p, FY * 8))) return bail(rc);
            HIPCHK(hipMemset(p, 0, FY * 8));
            S.ycref = (double*)p;
        }
        if ((rc = dev_alloc(s, &p, 8 * 8))) return bail(rc);
        s->res = (double*)p;
        if ((rc = dev_alloc(s, &p, 8))) return bail(rc);
        HIPCHK(hipMemset(p, 0, 8));
        S.health = (int*)p;
        if ((rc = dev_alloc(s, &p, sizeof(Solve2D)))) return bail(rc);
        s->d2 = (Solve2D*)p;
        // slope buffers of the phase-level calls are allocated on first use, or now when the fused
        // step cannot take this velocity grid
        if (s->thr2 == 0) {
            if ((rc = ensure_phase_buffers(s))) return bail(rc);
        } else if ((rc = push2d(s))) {
            return bail(rc);
        }
    }
    *out = s;
    return KBG_OK;
}

size_t kbg_usolve_bytes(const kbg_udesc* d) {
    if (!d) return 0;
    const size_t esz = d->precision == KBG_FP64 ? 8 : 4, C = d->ncell > 0 ? d->ncell : 0, F = d->nface > 0 ? d->nface : 0;
    const size_t nv = d->nv > 0 ? d->nv : 0, nf = d->nf > 0 ? d->nf : 0;
    // f and face fluxes; w, prim, dmin, area, dirc, residual rows; fw, geometry, walls; index arrays;
    // velocity grid
    return (2 * C + F) * nf * nv * esz + C * (4 + 4 + 1 + 1 + 3 + 8) * 8 + F * (4 + 3 + 4) * 8 + (C * 8 + F * 3) * 4 + 3 * nv * esz +
           sizeof(kbg::SolveUS) + 16 * 8;
}

int kbg_usolve_create(kbg_ctx* ctx, const kbg_udesc* ud, kbg_solve** out) {
    if (!ctx || !ud || !out) return fail(KBG_ERR_ARG, "ctx, desc or out is NULL");
    const kbg_udesc& u = *ud;
    if (u.nf != 1 && u.nf != 2) return fail(KBG_ERR_ARG, "nf must be 1 or 2");
    if (u.nv < 1 || u.nv > kUMaxNv) return fail(KBG_ERR_UNSUPPORTED, "unstructured solves take 1..4096 velocity points");
    if (u.ncell < 1 || u.nface < 1) return fail(KBG_ERR_ARG, "ncell and nface must be positive");
    if (u.precision != KBG_FP32 && u.precision != KBG_FP64) return fail(KBG_ERR_ARG, "precision must be KBG_FP32 or KBG_FP64");
    if (u.collision != KBG_BGK && u.collision != KBG_SHAKHOV) return fail(KBG_ERR_ARG, "unknown collision");
    if (u.collision == KBG_SHAKHOV && u.nf == 1)
        return fail(KBG_ERR_UNSUPPORTED, "Shakhov on one distribution: the triangle step of KitBase reads h and b");
    if (u.interp_order != 1)
        return fail(KBG_ERR_UNSUPPORTED, "unstructured solves are built for interpOrder 1 (first order: the cells' slopes stay zero); "
                                         "the second-order reconstruct! for unstructured cells is not built");
    if (!u.u || !u.v || !u.weights || !u.face_cells || !u.face_kind || !u.face_geo || !u.cell_faces || !u.cell_dirc ||
        !u.cell_area || !u.cell_type || !u.cell_nb)
        return fail(KBG_ERR_ARG, "a required unstructured array is NULL");
    if (!(u.gamma > 1.0) || !(u.mu_ref > 0.0) || !(u.K >= 0.0)) return fail(KBG_ERR_ARG, "gas constants out of range");
    const size_t C = u.ncell, F = u.nface;
    // mesh consistency: every index the kernels follow stays inside its array
    std::vector<int> avg;
    for (size_t f = 0; f < F; f++) {
        const int a = u.face_cells[2 * f], b = u.face_cells[2 * f + 1], k = u.face_kind[f];
        if (a < -1 || a >= (int)C || b < -1 || b >= (int)C) return fail(KBG_ERR_ARG, "face " + std::to_string(f) + ": cell index out of range");
        if (k == 0 && (a < 0 || b < 0)) return fail(KBG_ERR_ARG, "face " + std::to_string(f) + ": an interior face needs two cells");
        if (k == 1 && ((a < 0) == (b < 0))) return fail(KBG_ERR_ARG, "face " + std::to_string(f) + ": a wall face needs exactly one cell");
        if (k < 0 || k > 2) return fail(KBG_ERR_ARG, "face " + std::to_string(f) + ": kind must be 0, 1 or 2");
        if (k == 1 && !u.face_wall) return fail(KBG_ERR_ARG, "wall faces need face_wall");
        if (k == 1 && !(u.face_wall[4 * f + 3] > 0.0)) return fail(KBG_ERR_ARG, "face " + std::to_string(f) + ": wall lambda must be positive");
    }
    for (size_t c = 0; c < C; c++) {
        for (int j = 0; j < 3; j++) {
            const int f = u.cell_faces[3 * c + j], n = u.cell_nb[3 * c + j];
            if (f < 0 || f >= (int)F) return fail(KBG_ERR_ARG, "cell " + std::to_string(c) + ": face index out of range");
            if (n < -1 || n >= (int)C) return fail(KBG_ERR_ARG, "cell " + std::to_string(c) + ": neighbour index out of range");
        }
        const int t = u.cell_type[c];
        if ((t == 0 || t == 2) && !(u.cell_area[c] > 0.0)) return fail(KBG_ERR_ARG, "cell " + std::to_string(c) + ": area must be positive");
        if (t == 3) {
            int k = 0;
            for (int j = 0; j < 3; j++) k += u.cell_nb[3 * c + j] != -1;
            if (k < 2) return fail(KBG_ERR_ARG, "cell " + std::to_string(c) + ": a cellType-3 cell needs two neighbours");
            avg.push_back((int)c);
        }
    }
    HIPCHK(hipSetDevice(ctx->device));
    kbg_solve* s = new kbg_solve();
    s->ctx = ctx;
    s->us = true;
    kbg_desc& d = s->d;
    d.dim = 0;
    d.nf = u.nf;
    d.vdim = 2;
    d.nv = u.nv;
    d.nx = u.ncell;
    d.ny = 1;
    d.precision = u.precision;
    d.flux = KBG_KFVS;
    d.collision = u.collision;
    d.interp_order = u.interp_order;
    d.K = u.K;
    d.gamma = u.gamma;
    d.mu_ref = u.mu_ref;
    d.omega = u.omega;
    d.Pr = u.Pr;
    s->hu.assign(u.u, u.u + u.nv);
    s->hv.assign(u.v, u.v + u.nv);
    s->hwt.assign(u.weights, u.weights + u.nv);
    d.u = s->hu.data();
    d.v = s->hv.data();
    d.weights = s->hwt.data();
    s->fp64 = u.precision == KBG_FP64;
    s->esz = s->fp64 ? 8 : 4;
    s->nw = 4;
    s->N = u.ncell;
    s->RS = u.nf * u.nv;
    s->thru = kbg::ustep_threads(u.nv);
    auto bail = [&](int code) {
        kbg_solve_destroy(s);
        return code;
    };
    int rc;
    void* p;
    kbg::SolveUS& S = s->hus;
    S.ncell = u.ncell;
    S.nface = u.nface;
    S.nv = u.nv;
    S.nf = u.nf;
    S.coll = u.collision == KBG_SHAKHOV ? kbg::SHAKHOV : kbg::BGK;
    S.navg = (int)avg.size();
    S.K = u.K;
    S.gam = u.gamma;
    S.mu = u.mu_ref;
    S.om = u.omega;
    S.Pr = u.Pr;
    S.dt = 0.0;
    S.cur = 0;
    if ((rc = dev_alloc(s, &s->u, u.nv * s->esz)) || (rc = put_vec(s, s->u, u.u, u.nv))) return bail(rc);
    if ((rc = dev_alloc(s, &s->v, u.nv * s->esz)) || (rc = put_vec(s, s->v, u.v, u.nv))) return bail(rc);
    if ((rc = dev_alloc(s, &s->wt, u.nv * s->esz)) || (rc = put_vec(s, s->wt, u.weights, u.nv))) return bail(rc);
    S.u = s->u;
    S.v = s->v;
    S.wt = s->wt;
    auto put_i = [&](const int*& dst, const int* src, size_t n) -> int {
        int r;
        if ((r = dev_alloc(s, &p, n * 4))) return r;
        if (n) {
            HIPCHK(hipMemcpyAsync(p, src, n * 4, hipMemcpyHostToDevice, s->ctx->stream));
            HIPCHK(hipStreamSynchronize(s->ctx->stream));
        }
        dst = (const int*)p;
        return KBG_OK;
    };
    auto put_d = [&](const double*& dst, const double* src, size_t n) -> int {
        int r;
        if ((r = dev_alloc(s, &p, n * 8))) return r;
        if (src) {
            if ((r = put_f64(s, (double*)p, src, n))) return r;
        } else {
            HIPCHK(hipMemset(p, 0, n * 8));
        }
        dst = (const double*)p;
        return KBG_OK;
    };
    if ((rc = put_i(S.fc, u.face_cells, 2 * F)) || (rc = put_i(S.fkind, u.face_kind, F)) || (rc = put_i(S.cf, u.cell_faces, 3 * C)) ||
        (rc = put_i(S.ctype, u.cell_type, C)) || (rc = put_i(S.nb, u.cell_nb, 3 * C)) || (rc = put_i(S.avg, avg.data(), avg.size())) ||
        (rc = put_d(S.fgeo, u.face_geo, 3 * F)) || (rc = put_d(S.fwall, u.face_wall, 4 * F)) || (rc = put_d(S.dirc, u.cell_dirc, 3 * C)) ||
        (rc = put_d(S.area, u.cell_area, C)))
        return bail(rc);
    if (u.cell_dmin) {
        const double* dm = nullptr;
        if ((rc = put_d(dm, u.cell_dmin, C))) return bail(rc);
        s->dx = const_cast<double*>(dm);
    }
    struct {
        void** dst;
        size_t bytes;
    } bufs[] = {
        {(void**)&S.w, C * 4 * 8},          {(void**)&S.prim, C * 4 * 8}, {&S.f, C * s->RS * s->esz}, {&S.f1, C * s->RS * s->esz},
        {(void**)&S.fw, F * 4 * 8},         {&S.ff, F * s->RS * s->esz},  {(void**)&s->rcell, C * 8 * 8},
        {(void**)&s->res, 8 * 8},           {(void**)&S.health, 8},
    };
    for (auto& b : bufs) {
        if ((rc = dev_alloc(s, &p, b.bytes))) return bail(rc);
        HIPCHK(hipMemset(p, 0, b.bytes));
        *b.dst = p;
    }
    S.rcell = s->rcell;
    if ((rc = dev_alloc(s, &p, sizeof(kbg::SolveUS)))) return bail(rc);
    s->dus = (kbg::SolveUS*)p;
    if ((rc = pushus(s))) return bail(rc);
    HIPCHK(hipDeviceSynchronize());
    *out = s;
    return KBG_OK;
}

int kbg_solve_destroy(kbg_solve* s) {
    if (!s) return KBG_OK;
    if (s->ctx) (void)hipStreamSynchronize(s->ctx->stream);
    kbg::fsm_plan_destroy(s->fsm);
    for (void* p : s->allocs) (void)hipFree(p);
    delete s;
    return KBG_OK;
}

int kbg_solve_health(kbg_solve* s, int* flags, int clear) {
    if (!s || !flags) return fail(KBG_ERR_ARG, "NULL argument");
    int* h = s->us ? s->hus.health : s->d.dim == 1 ? s->h1.health : s->h2.health;
    HIPCHK(hipMemcpyAsync(flags, h, sizeof(int), hipMemcpyDeviceToHost, s->ctx->stream));
    if (clear) HIPCHK(hipMemsetAsync(h, 0, sizeof(int), s->ctx->stream));
    HIPCHK(hipStreamSynchronize(s->ctx->stream));
    return KBG_OK;
}

}  // extern "C"

namespace {

// 2D1F2V host arrays [rows][1][nv] <-> the two-distribution device layout [rows][2][nv] (b = 0)
std::vector<double> widen2d(const kbg_solve* s, const double* src, size_t rows) {
    const size_t nv = s->d.nv;
    std::vector<double> out(rows * 2 * nv, 0.0);
    for (size_t r = 0; r < rows; r++) std::copy(src + r * nv, src + (r + 1) * nv, out.begin() + r * 2 * nv);
    return out;
}
void narrow2d(const kbg_solve* s, const std::vector<double>& src, double* dst, size_t rows) {
    const size_t nv = s->d.nv;
    for (size_t r = 0; r < rows; r++) std::copy(src.begin() + r * 2 * nv, src.begin() + r * 2 * nv + nv, dst + r * nv);
}

}  // namespace

extern "C" {

int kbg_upload(kbg_solve* s, const double* w, const double* prim, const double* f) {
    if (!s || !w || !prim || !f) return fail(KBG_ERR_ARG, "NULL argument");
    int rc;
    const size_t N = s->N;
    if (s->us) {
        const kbg::SolveUS& S = s->hus;
        if ((rc = put_f64(s, S.w, w, N * 4)) || (rc = put_f64(s, S.prim, prim, N * 4))) return rc;
        return put_vec(s, us_f(s), f, N * s->RS);
    }
    if (s->d.dim == 1) {
        Solve1D& S = s->h1;
        if ((rc = put_f64(s, S.prim, prim, N * s->nw))) return rc;
        for (int b = 0; b < 2; b++)
            if ((rc = put_f64(s, S.w[b], w, N * s->nw))) return rc;
        if (S.delta) {
            if ((rc = push1d(s))) return rc;
            return put_f1d_delta(s, f, {0, 1});
        }
        for (int b = 0; b < 2; b++)
            if ((rc = put_f1d(s, S.f[b], f))) return rc;
        return KBG_OK;
    }
    Solve2D& S = s->h2;
    if ((rc = put_f64(s, S.w, w, N * 4))) return rc;
    if ((rc = put_f64(s, S.prim, prim, N * 4))) return rc;
    std::vector<double> wide;
    if (s->d.nf == 1) {
        wide = widen2d(s, f, N);
        f = wide.data();
    }
    if (!S.delta) return put_vec(s, S.f, f, N * 2 * s->d.nv);
    // departures from a reference Maxwellian of this state; the face fluxes count as full values
    set_mref2d(s, w);
    const size_t RS = 2 * (size_t)s->d.nv, FX = size_t(s->d.nx + 1) * s->d.ny, FY = size_t(s->d.nx) * (s->d.ny + 1);
    const float* mr = s->mref.data();
    const double* fs = f;
    if ((rc = stage_up(s, static_cast<float*>(S.f), N * RS, [fs, mr, RS](size_t i) { return float(fs[i] - double(mr[i % RS])); })))
        return rc;
    HIPCHK(hipMemcpyAsync((void*)S.mref, s->mref.data(), RS * 4, hipMemcpyHostToDevice, s->ctx->stream));
    HIPCHK(hipMemsetAsync(S.xcref, 0, FX * 8, s->ctx->stream));
    HIPCHK(hipMemsetAsync(S.ycref, 0, FY * 8, s->ctx->stream));
    HIPCHK(hipStreamSynchronize(s->ctx->stream));
    s->rep_full = false;
    return push2d(s);
}

int kbg_download(kbg_solve* s, double* w, double* prim, double* f) {
    if (!s) return fail(KBG_ERR_ARG, "NULL solve");
    int rc;
    const size_t N = s->N;
    if (s->us) {
        const kbg::SolveUS& S = s->hus;
        if (w && (rc = get_f64(s, w, S.w, N * 4))) return rc;
        if (prim && (rc = get_f64(s, prim, S.prim, N * 4))) return rc;
        if (f && (rc = get_vec(s, f, us_f(s), N * s->RS))) return rc;
        return KBG_OK;
    }
    if (s->d.dim == 1) {
        Solve1D& S = s->h1;
        if (w && (rc = get_f64(s, w, S.w[S.cur], N * s->nw))) return rc;
        if (prim && (rc = get_f64(s, prim, S.prim, N * s->nw))) return rc;
        if (f) {
            if (S.delta && !s->rep_full) {
                if ((rc = push1d(s)) || (rc = get_f1d_delta(s, f, S.cur))) return rc;
            } else if ((rc = get_f1d(s, f, S.f[S.cur]))) {
                return rc;
            }
        }
        return KBG_OK;
    }
    Solve2D& S = s->h2;
    if (w && (rc = get_f64(s, w, S.w, N * 4))) return rc;
    if (prim && (rc = get_f64(s, prim, S.prim, N * 4))) return rc;
    if (!f) return KBG_OK;
    std::vector<double> wide;
    double* fo = f;
    if (s->d.nf == 1) {
        wide.resize(N * 2 * s->d.nv);
        fo = wide.data();
    }
    if (S.delta && !s->rep_full) {
        // departure + Mr, exactly (two FP32 values summed in FP64)
        const size_t RS = 2 * (size_t)s->d.nv;
        const float* mr = s->mref.data();
        if ((rc = stage_down(s, static_cast<const float*>(S.f), N * RS,
                             [fo, mr, RS](size_t i, float v) { fo[i] = double(v) + double(mr[i % RS]); })))
            return rc;
    } else if ((rc = get_vec(s, fo, S.f, N * 2 * s->d.nv))) {
        return rc;
    }
    if (s->d.nf == 1) narrow2d(s, wide, f, N);
    return KBG_OK;
}

int kbg_upload_slopes(kbg_solve* s, const double* sw, const double* sf) {
    if (!s) return fail(KBG_ERR_ARG, "NULL solve");
    if (s->us) return fail(KBG_ERR_UNSUPPORTED, "an unstructured solve keeps no slopes (it is built for interpOrder 1, where they are zero)");
    int rc;
    const size_t N = s->N;
    if (s->d.dim == 1) {
        if ((rc = ensure_phase_buffers(s))) return rc;
        if (sw && (rc = put_f64(s, s->h1.sw, sw, N * s->nw))) return rc;
        if (sf && (rc = put_vec(s, s->h1.sf, sf, N * s->RS))) return rc;
        return KBG_OK;
    }
    if ((rc = ensure_phase_buffers(s))) return rc;
    if (sw && (rc = put_f64(s, s->h2.sw, sw, N * 8))) return rc;
    if (sf && s->d.nf == 1) return put_vec(s, s->h2.sf, widen2d(s, sf, 2 * N).data(), N * 4 * s->d.nv);
    if (sf && (rc = put_vec(s, s->h2.sf, sf, N * 4 * s->d.nv))) return rc;
    return KBG_OK;
}

int kbg_download_slopes(kbg_solve* s, double* sw, double* sf) {
    if (!s) return fail(KBG_ERR_ARG, "NULL solve");
    if (s->us) return fail(KBG_ERR_UNSUPPORTED, "an unstructured solve keeps no slopes (it is built for interpOrder 1, where they are zero)");
    int rc;
    const size_t N = s->N;
    if (s->d.dim == 1) {
        if ((rc = ensure_phase_buffers(s))) return rc;
        if (sw && (rc = get_f64(s, sw, s->h1.sw, N * s->nw))) return rc;
        if (sf && (rc = get_vec(s, sf, s->h1.sf, N * s->RS))) return rc;
        return KBG_OK;
    }
    if ((rc = ensure_phase_buffers(s))) return rc;
    if (sw && (rc = get_f64(s, sw, s->h2.sw, N * 8))) return rc;
    if (sf && s->d.nf == 1) {
        std::vector<double> wide(N * 4 * s->d.nv);
        if ((rc = get_vec(s, wide.data(), s->h2.sf, wide.size()))) return rc;
        narrow2d(s, wide, sf, 2 * N);
        return KBG_OK;
    }
    if (sf && (rc = get_vec(s, sf, s->h2.sf, N * 4 * s->d.nv))) return rc;
    return KBG_OK;
}

int kbg_upload_faces(kbg_solve* s, const double* fw, const double* ff) {
    if (!s) return fail(KBG_ERR_ARG, "NULL solve");
    int rc;
    if (s->us) {
        const size_t F = s->hus.nface;
        if (fw && (rc = put_f64(s, s->hus.fw, fw, F * 4))) return rc;
        if (ff && (rc = put_vec(s, s->hus.ff, ff, F * s->RS))) return rc;
        return KBG_OK;
    }
    if (s->d.dim == 1) {
        if ((rc = ensure_phase_buffers(s))) return rc;
        const size_t nf1 = size_t(s->d.nx) + 1;
        if (fw && (rc = put_f64(s, s->h1.fw, fw, nf1 * s->nw))) return rc;
        if (ff && (rc = put_vec(s, s->h1.ff, ff, nf1 * s->RS))) return rc;
        return sync_boundary_faces(s);
    }
    if ((rc = ensure_phase_buffers(s))) return rc;
    const size_t FX = size_t(s->d.nx + 1) * s->d.ny, FY = size_t(s->d.nx) * (s->d.ny + 1), RS = 2 * s->d.nv;
    if (ff && s->h2.delta && !s->rep_full) {
        // uploaded face fluxes are full values: none leaves out the reference Maxwellian's part
        HIPCHK(hipMemsetAsync(s->h2.xcref, 0, FX * 8, s->ctx->stream));
        HIPCHK(hipMemsetAsync(s->h2.ycref, 0, FY * 8, s->ctx->stream));
    }
    if (fw) {
        if ((rc = put_f64(s, s->h2.xfw, fw, FX * 4))) return rc;
        if ((rc = put_f64(s, s->h2.yfw, fw + FX * 4, FY * 4))) return rc;
    }
    if (ff) {
        std::vector<double> wide;
        if (s->d.nf == 1) {
            wide = widen2d(s, ff, FX + FY);
            ff = wide.data();
        }
        if ((rc = put_vec(s, s->h2.xff, ff, FX * RS))) return rc;
        if ((rc = put_vec(s, s->h2.yff, ff + FX * RS, FY * RS))) return rc;
    }
    return KBG_OK;
}

int kbg_download_faces(kbg_solve* s, double* fw, double* ff) {
    if (!s) return fail(KBG_ERR_ARG, "NULL solve");
    int rc;
    if (s->us) {
        const size_t F = s->hus.nface;
        if (fw && (rc = get_f64(s, fw, s->hus.fw, F * 4))) return rc;
        if (ff && (rc = get_vec(s, ff, s->hus.ff, F * s->RS))) return rc;
        return KBG_OK;
    }
    if (s->d.dim == 1) {
        if ((rc = ensure_phase_buffers(s))) return rc;
        const size_t nf1 = size_t(s->d.nx) + 1;
        if (fw && (rc = get_f64(s, fw, s->h1.fw, nf1 * s->nw))) return rc;
        if (ff && (rc = get_vec(s, ff, s->h1.ff, nf1 * s->RS))) return rc;
        return KBG_OK;
    }
    if ((rc = ensure_phase_buffers(s))) return rc;
    const size_t FX = size_t(s->d.nx + 1) * s->d.ny, FY = size_t(s->d.nx) * (s->d.ny + 1), RS = 2 * s->d.nv;
    if (fw) {
        if ((rc = get_f64(s, fw, s->h2.xfw, FX * 4))) return rc;
        if ((rc = get_f64(s, fw + FX * 4, s->h2.yfw, FY * 4))) return rc;
    }
    if (ff) {
        std::vector<double> wide;
        double* fo = ff;
        if (s->d.nf == 1) {
            wide.resize((FX + FY) * RS);
            fo = wide.data();
        }
        if ((rc = get_vec(s, fo, s->h2.xff, FX * RS))) return rc;
        if ((rc = get_vec(s, fo + FX * RS, s->h2.yff, FY * RS))) return rc;
        if (s->h2.delta && !s->rep_full) {
            // the fused step's face fluxes leave out c x Mr: add it back (FP64, on the host)
            std::vector<double> c(FX + FY);
            if ((rc = get_f64(s, c.data(), s->h2.xcref, FX)) || (rc = get_f64(s, c.data() + FX, s->h2.ycref, FY))) return rc;
            const kbg_desc& d = s->d;
            for (size_t fc = 0; fc < FX + FY; fc++) {
                if (c[fc] == 0.0) continue;
                const double* g = fc < FX ? d.xgeo + fc * 3 : d.ygeo + (fc - FX) * 3;
                for (size_t k = 0; k < RS; k++) {
                    const size_t iv = k % d.nv;
                    fo[fc * RS + k] += c[fc] * (g[0] * d.u[iv] + g[1] * d.v[iv]) * double(s->mref[k]);
                }
            }
        }
        if (s->d.nf == 1) narrow2d(s, wide, ff, FX + FY);
    }
    return KBG_OK;
}

int kbg_reconstruct(kbg_solve* s) {
    if (!s) return fail(KBG_ERR_ARG, "NULL solve");
    int rc;
    if (s->us) return KBG_OK;  // interpOrder 1: reconstruct! leaves the (zero) slopes as they are
    if (s->d.dim == 1) {
        if ((rc = ensure_phase_buffers(s)) || (rc = push1d(s)) || (rc = to_full_rep(s))) return rc;
        HIPCHK(kbg::launch_recon1d(s->fp64, s->h1, s->d1, s->ctx->stream));
    } else {
        if ((rc = ensure_phase_buffers(s)) || (rc = push2d(s)) || (rc = to_full_rep(s))) return rc;
        HIPCHK(kbg::launch_recon2d(s->fp64, s->h2, s->d2, s->ctx->stream));
    }
    HIPCHK(hipStreamSynchronize(s->ctx->stream));
    return KBG_OK;
}

int kbg_evolve(kbg_solve* s, double dt) {
    if (!s) return fail(KBG_ERR_ARG, "NULL solve");
    int rc;
    if (s->us) {
        s->hus.dt = dt;
        if ((rc = pushus(s))) return rc;
        HIPCHK(kbg::launch_uface(s->fp64, s->thru, s->hus, s->dus, s->ctx->stream));
        HIPCHK(hipStreamSynchronize(s->ctx->stream));
        return KBG_OK;
    }
    if (s->d.dim == 1) {
        if ((rc = ensure_phase_buffers(s))) return rc;
        s->h1.dt = dt;
        if ((rc = push1d(s)) || (rc = to_full_rep(s))) return rc;
        HIPCHK(kbg::launch_evolve1d(s->fp64, s->h1, s->d1, s->ctx->stream));
        HIPCHK(kbg::launch_wall1d(s->fp64, s->h1, s->d1, 0, 0, s->ctx->stream));
        if ((rc = sync_boundary_faces(s))) return rc;
    } else {
        if ((rc = ensure_phase_buffers(s))) return rc;
        s->h2.dt = dt;
        if ((rc = push2d(s)) || (rc = to_full_rep(s))) return rc;
        HIPCHK(kbg::launch_evolve2d(s->fp64, s->h2, s->d2, s->ctx->stream));
    }
    HIPCHK(hipStreamSynchronize(s->ctx->stream));
    return KBG_OK;
}

int kbg_update(kbg_solve* s, double dt, double* residual) {
    if (!s) return fail(KBG_ERR_ARG, "NULL solve");
    int rc;
    if (s->us) {
        s->hus.dt = dt;
        if ((rc = pushus(s))) return rc;
        hipStream_t st = s->ctx->stream;
        HIPCHK(kbg::launch_ucell(s->fp64, s->thru, s->hus, s->dus, residual ? 1 : 0, st));
        HIPCHK(kbg::launch_uavg(s->fp64, s->hus, s->dus, 0, st));
        if (residual) {
            HIPCHK(kbg::launch_uresid(s->rcell, s->hus.ncell, s->res, st));
            HIPCHK(hipMemcpyAsync(residual, s->res, 4 * 8, hipMemcpyDeviceToHost, st));
        }
        HIPCHK(hipStreamSynchronize(st));
        return KBG_OK;
    }
    if (s->d.dim == 1) {
        if ((rc = ensure_phase_buffers(s))) return rc;
        s->h1.dt = dt;
        if ((rc = push1d(s)) || (rc = to_full_rep(s))) return rc;
        rc = step_phase_1d(s, dt, residual, s->ctx->stream);
    } else {
        if ((rc = ensure_phase_buffers(s))) return rc;
        s->h2.dt = dt;
        if ((rc = push2d(s)) || (rc = to_full_rep(s))) return rc;
        rc = step_phase_2d(s, dt, residual, s->ctx->stream);
    }
    if (rc) return rc;
    HIPCHK(hipStreamSynchronize(s->ctx->stream));
    return KBG_OK;
}

int kbg_timestep(kbg_solve* s, double cfl, double u1, double v1, double sim_time, double max_time, double* dt) {
    if (!s || !dt) return fail(KBG_ERR_ARG, "NULL argument");
    if (!(cfl > 0.0)) return fail(KBG_ERR_ARG, "cfl must be positive");
    hipStream_t st = s->ctx->stream;
    if (s->us) {
        if (!s->dx) return fail(KBG_ERR_ARG, "the unstructured solve was created without cell_dmin");
        HIPCHK(kbg::launch_ucfl(s->hus.prim, s->dx, s->hus.ncell, u1, v1, s->d.gamma, s->res, st));
    } else if (s->d.dim == 1) {
        const Solve1D& S = s->h1;
        HIPCHK(kbg::launch_cfl(S.prim, s->nw, s->d.ng, s->d.nx, s->dx, nullptr, u1, 0.0, s->d.gamma, s->res, st));
    } else {
        const Solve2D& S = s->h2;
        HIPCHK(kbg::launch_cfl(S.prim, 4, 0, s->N, s->dx, s->dy, u1, v1, s->d.gamma, s->res, st));
    }
    double tmax = 0.0;
    HIPCHK(hipMemcpyAsync(&tmax, s->res, 8, hipMemcpyDeviceToHost, st));
    HIPCHK(hipStreamSynchronize(st));
    const double d = cfl / tmax;
    *dt = d < (max_time - sim_time) ? d : max_time - sim_time;
    return KBG_OK;
}

int kbg_batch_timestep(kbg_ctx* ctx, kbg_solve* const* solves, int n, const double* cfl, const double* u1, const double* v1,
                       const double* sim_time, const double* max_time, double* dts) {
    if (!ctx || (n > 0 && (!solves || !cfl || !u1 || !dts))) return fail(KBG_ERR_ARG, "NULL argument");
    hipStream_t st = ctx->stream;
    // one reduction per solve on the context stream, each into its solve's result slot; one sync
    for (int i = 0; i < n; i++) {
        kbg_solve* s = solves[i];
        if (!s || s->ctx->device != ctx->device) return fail(KBG_ERR_ARG, "solve " + std::to_string(i) + " is NULL or on another device");
        if (!(cfl[i] > 0.0)) return fail(KBG_ERR_ARG, "cfl must be positive");
        if (s->us) {
            if (!s->dx) return fail(KBG_ERR_ARG, "solve " + std::to_string(i) + ": unstructured, created without cell_dmin");
            HIPCHK(kbg::launch_ucfl(s->hus.prim, s->dx, s->hus.ncell, u1[i], v1 ? v1[i] : 0.0, s->d.gamma, s->res, st));
        } else if (s->d.dim == 1)
            HIPCHK(kbg::launch_cfl(s->h1.prim, s->nw, s->d.ng, s->d.nx, s->dx, nullptr, u1[i], 0.0, s->d.gamma, s->res, st));
        else
            HIPCHK(kbg::launch_cfl(s->h2.prim, 4, 0, s->N, s->dx, s->dy, u1[i], v1 ? v1[i] : 0.0, s->d.gamma, s->res, st));
    }
    std::vector<double> tmax(n, 0.0);
    for (int i = 0; i < n; i++) HIPCHK(hipMemcpyAsync(&tmax[i], solves[i]->res, 8, hipMemcpyDeviceToHost, st));
    HIPCHK(hipStreamSynchronize(st));
    for (int i = 0; i < n; i++) {
        const double d = cfl[i] / tmax[i];
        const double left = (max_time ? max_time[i] : INFINITY) - (sim_time ? sim_time[i] : 0.0);
        dts[i] = d < left ? d : left;
    }
    return KBG_OK;
}

int kbg_batch_step_cfl(kbg_ctx* ctx, kbg_solve* const* solves, int n, const double* cfl, const double* u1, const double* v1,
                       const double* sim_time, const double* max_time, int nsteps, double* dts, double* residuals) {
    if (int rc = kbg_batch_timestep(ctx, solves, n, cfl, u1, v1, sim_time, max_time, dts)) return rc;
    return kbg_batch_step(ctx, solves, n, dts, nsteps, residuals);
}

int kbg_step(kbg_solve* s, double dt, int nsteps, double* residual) {
    if (!s) return fail(KBG_ERR_ARG, "NULL solve");
    if (nsteps < 0) return fail(KBG_ERR_ARG, "nsteps < 0");
    if (nsteps == 0) return KBG_OK;
    kbg_solve* one[1] = {s};
    double dts[1] = {dt};
    return kbg_batch_step(s->ctx, one, 1, dts, nsteps, residual);
}

int kbg_batch_step(kbg_ctx* ctx, kbg_solve* const* solves, int n, const double* dts, int nsteps, double* residuals) {
    return batch_step_impl(ctx, solves, n, dts, nsteps, residuals, false);
}

int kbg_batch_step_noalloc(kbg_ctx* ctx, kbg_solve* const* solves, int n, const double* dts, int nsteps, double* residuals) {
    return batch_step_impl(ctx, solves, n, dts, nsteps, residuals, true);
}

}  // extern "C"

namespace {

// kbg_batch_step and its allocation-free variant.  The batch's device staging (1D descriptors, work
// items, packs, residual partial rows; 2D descriptors and prefix sums) lives in the context's
// buffers, grown with hipMalloc when a batch outgrows them, or -- noalloc -- in this device's
// file-scope `scratch` array, carved per call.
int batch_step_impl(kbg_ctx* ctx, kbg_solve* const* solves, int n, const double* dts, int nsteps, double* residuals, bool noalloc) {
    if (!ctx || (n > 0 && (!solves || !dts))) return fail(KBG_ERR_ARG, "NULL argument");
    if (nsteps <= 0 || n <= 0) return KBG_OK;
    hipStream_t st = ctx->stream;
    int rc;
    // fused 1D solves go to the sweep; the rest are stepped phase by phase
    std::vector<int> fused, fused2, others;
    for (int i = 0; i < n; i++) {
        kbg_solve* s = solves[i];
        if (!s || s->ctx->device != ctx->device) return fail(KBG_ERR_ARG, "solve is NULL or on another device");
        if (s->d.dim == 1 && s->fused) {
            fused.push_back(i);
            if ((rc = ensure_vtab(s, dts[i], st))) return rc;
        } else if (s->d.dim == 2 && s->thr2 > 0) {
            fused2.push_back(i);
        } else {
            others.push_back(i);
        }
    }
    // fused 2D solves: face-chain, face-flux and cell launches per (precision, flux, workgroup size)
    // group and step, all on the context stream
    struct Launch2 {
        bool fp64;
        int flux, threads, n, ncells, nfaces, ninterior;
        size_t desc0, cpref0, fpref0, ipref0;
        double work;
    };
    std::vector<Launch2> launches2;
    std::vector<Solve2D> descs2;
    std::vector<int> pref2;
    {
        std::map<std::tuple<int, int, int>, std::vector<int>> groups;
        for (int i : fused2) groups[{solves[i]->fp64 ? 1 : 0, solves[i]->d.flux, solves[i]->thr2}].push_back(i);
        for (auto& kv : groups) {
            Launch2 L{std::get<0>(kv.first) == 1, std::get<1>(kv.first), std::get<2>(kv.first), (int)kv.second.size(), 0, 0, 0,
                      descs2.size(), 0, 0, 0, 0.0};
            std::vector<int> cp{0}, fp{0}, ip{0};
            for (int i : kv.second) {
                kbg_solve* s = solves[i];
                s->h2.dt = dts[i];
                descs2.push_back(s->h2);
                const long nx = s->d.nx, ny = s->d.ny;
                const long nc = nx * ny, nf = (nx + 1) * ny + nx * (ny + 1), ni = (nx - 1) * ny + nx * (ny - 1);
                if (cp.back() + nc > INT_MAX / 2 || fp.back() + nf > INT_MAX / 2) return fail(KBG_ERR_ARG, "2D batch too large for one launch");
                cp.push_back(int(cp.back() + nc));
                fp.push_back(int(fp.back() + nf));
                ip.push_back(int(ip.back() + ni));
                L.work += double(nc) * s->d.nv * 2;
            }
            L.ncells = cp.back();
            L.nfaces = fp.back();
            L.ninterior = ip.back();
            L.cpref0 = pref2.size();
            pref2.insert(pref2.end(), cp.begin(), cp.end());
            L.fpref0 = pref2.size();
            pref2.insert(pref2.end(), fp.begin(), fp.end());
            L.ipref0 = pref2.size();
            pref2.insert(pref2.end(), ip.begin(), ip.end());
            launches2.push_back(L);
        }
    }
    // two time steps per sweep launch for each fused 1D solve that allows it (KFVS, no boundary-face
    // buffers: no walls, no stepped cell next to a ghost-less boundary); not under graph replay.  The
    // choice is the solve's own, so a solve advances in the same step pairs (counted from the first
    // step of the call) in any batch: a two-step pass may round differently from two one-step passes.
    const bool two_ok = ctx->sweep_steps == 2;
    std::vector<char> two_of(n, 0);
    bool two = false;  // some launch takes two steps per round
    for (int i : fused) {
        const Solve1D& S = solves[i]->h1;
        two_of[i] = two_ok && S.flux == kbg::KFVS && !S.wallL && !S.wallR && !(S.faceL_buf && S.upd_lo == 1) &&
                    !(S.faceR_buf && S.upd_hi == S.nx);
        two = two || two_of[i];
    }
    auto seg_of = [&](int i) { return std::min(std::max(1, ctx->seg_cells), two_of[i] ? kbg::kSweepOut2 : kbg::kSweepOut); };
    // Independent units -- fused 1D solves, fused 2D groups and solves stepped by the phase kernels
    // -- are spread over the streams by work (cells x velocities x distributions), heaviest first;
    // each stream advances its own share.
    const int units = int(fused.size() + launches2.size() + others.size());
    const int nstreams = std::max(1, std::min<int>(ctx->streams > 0 ? ctx->streams : 1 + kbg_ctx::kSide, units));
    std::vector<int> stream_of(n, 0), stream2(launches2.size(), 0);
    {
        auto work = [&](int i) { return double(solves[i]->N) * solves[i]->d.nv * solves[i]->d.nf; };
        std::vector<std::pair<double, int>> order;  // (work, unit): unit >= 0 a solve, < 0 a 2D group (-1 - g)
        for (int i : fused) order.push_back({work(i), i});
        for (int i : others) order.push_back({work(i), i});
        for (size_t g = 0; g < launches2.size(); g++) order.push_back({launches2[g].work, -1 - int(g)});
        std::stable_sort(order.begin(), order.end(), [](const auto& a, const auto& b) { return a.first > b.first; });
        std::vector<double> load(nstreams, 0.0);
        for (const auto& u : order) {
            const int k = int(std::min_element(load.begin(), load.end()) - load.begin());
            if (u.second >= 0) stream_of[u.second] = k;
            else stream2[-1 - u.second] = k;
            load[k] += u.first;
        }
    }
    // solves stepped by the phase kernels: buffers and descriptors before the streams fork
    for (int i : others) {
        kbg_solve* s = solves[i];
        if (s->us) {
            s->hus.dt = dts[i];
            if ((rc = pushus(s, st))) return rc;
            continue;
        }
        if (noalloc && s->has_fsm)
            return fail(KBG_ERR_STATE, "kbg_batch_step_noalloc: a solve with the fast spectral collision runs hipFFT, which is "
                                       "not covered by the allocation-free contract");
        if (noalloc && !(s->d.dim == 1 ? s->h1.sw != nullptr : s->h2.sf != nullptr))
            return fail(KBG_ERR_STATE, "kbg_batch_step_noalloc: a solve stepped by the phase kernels has no slope and face "
                                       "buffers yet (one kbg_step, kbg_batch_step or phase-level call allocates them)");
        if ((rc = ensure_phase_buffers(s))) return rc;
        if (s->d.dim == 1) {
            s->h1.dt = dts[i];
            if ((rc = push1d(s, st))) return rc;
        } else {
            s->h2.dt = dts[i];
            if ((rc = push2d(s, st))) return rc;
        }
    }
    // descriptors, work items and packs; one launch per (stream, kernel instance)
    struct Launch {
        bool fp64;
        int nf, flux, wg_waves, stream;
        size_t pack0, npacks;
        std::vector<int> members;  // solve indices (for the wall kernels)
        bool two;                  // two time steps per launch in a round of two
        int flips;                 // sweep launches enqueued so far (the f / w buffer parity)
    };
    std::vector<Solve1D> descs;
    std::vector<int> desc_of(n, -1);
    std::vector<WorkItem> items;
    std::vector<kbg::Pack> packs;
    std::vector<Launch> launches;
    std::map<std::tuple<int, int, int, int, int>, std::vector<int>> keyed;  // (stream, fp64, nf, flux, two) -> solves
    for (int i : fused)
        keyed[{stream_of[i], solves[i]->fp64 ? 1 : 0, solves[i]->d.nf, solves[i]->d.flux, two_of[i] ? 1 : 0}].push_back(i);
    std::vector<int> launch_of(n, -1);
    for (auto& kv : keyed) {
        Launch L{std::get<1>(kv.first) == 1, std::get<2>(kv.first), std::get<3>(kv.first), 1, std::get<0>(kv.first),
                 packs.size(), 0, kv.second, std::get<4>(kv.first) == 1, 0};
        for (int i : kv.second) launch_of[i] = (int)launches.size();
        std::map<int, std::vector<int>> by_nw;  // nw -> item indices, in solve order
        for (int i : kv.second) {
            kbg_solve* s = solves[i];
            s->h1.dt = dts[i];
            desc_of[i] = (int)descs.size();
            descs.push_back(s->h1);
            const size_t i0 = items.size();
            // wavefronts per item: the kernel's width, or more when the context caps the velocities
            // per wavefront (a function of the solve and the context only, so a solve's sums keep
            // their order in any batch)
            int nw = s->sweep_nw;
            if (ctx->sweep_vw > 0)
                while (nw < kbg::kSweepMaxWaves && (s->d.nv + nw - 1) / nw > ctx->sweep_vw) nw *= 2;
            build_items(s, seg_of(i), desc_of[i], nw, items);
            for (size_t it = i0; it < items.size(); it++) by_nw[nw].push_back((int)it);
            L.wg_waves = std::max(L.wg_waves, nw);
        }
        // workgroups of at least 4 wavefronts: side-by-side items of the same nw fill them
        L.wg_waves = std::max(L.wg_waves, std::min(4, kbg::kSweepMaxWaves));
        // items are contiguous per nw bucket, so a pack is a run of consecutive items
        std::vector<WorkItem> reordered;
        for (auto& b : by_nw) {
            const int per = L.wg_waves / b.first;
            for (size_t k = 0; k < b.second.size(); k += per) {
                kbg::Pack p{(int)(items.size() + reordered.size()), (int)std::min<size_t>(per, b.second.size() - k), b.first, 0};
                for (int m = 0; m < p.count; m++) reordered.push_back(items[b.second[k + m]]);
                packs.push_back(p);
            }
        }
        // the reordered items replace this launch's items at the end of the list
        const size_t first = items.size() - reordered.size();
        for (kbg::Pack* p = packs.data() + L.pack0; p != packs.data() + packs.size(); ++p) p->item0 -= (int)reordered.size();
        std::copy(reordered.begin(), reordered.end(), items.begin() + first);
        L.npacks = packs.size() - L.pack0;
        launches.push_back(std::move(L));
    }
    // device staging of the batch: the context's buffers (grown on demand) or this device's scratch
    Solve1D* d_descs = nullptr;
    WorkItem* d_items = nullptr;
    double* d_rpart = nullptr;
    kbg::Pack* d_packs = nullptr;
    Solve2D* d_descs2 = nullptr;
    int* d_pref2 = nullptr;
    if (noalloc) {
        char* base = static_cast<char*>(scratch_ptr(ctx->device));
        if (!base) return fail(KBG_ERR_STATE, "kbg_batch_step_noalloc: the device's scratch address was not set up");
        size_t off = 0;
        auto carve = [&](size_t bytes) {
            char* p = base + off;
            off += (bytes + 255) / 256 * 256;
            return p;
        };
        d_descs = reinterpret_cast<Solve1D*>(carve(descs.size() * sizeof(Solve1D)));
        d_items = reinterpret_cast<WorkItem*>(carve(items.size() * sizeof(WorkItem)));
        d_rpart = reinterpret_cast<double*>(carve(items.size() * 12 * sizeof(double)));
        d_packs = reinterpret_cast<kbg::Pack*>(carve(packs.size() * sizeof(kbg::Pack)));
        d_descs2 = reinterpret_cast<Solve2D*>(carve(descs2.size() * sizeof(Solve2D)));
        d_pref2 = reinterpret_cast<int*>(carve(pref2.size() * sizeof(int)));
        if (off > kNoallocScratchBytes)
            return fail(KBG_ERR_ARG, "kbg_batch_step_noalloc: the batch needs " + std::to_string(off) +
                                         " bytes of staging; the allocation-free entry point holds at most " +
                                         std::to_string(kNoallocScratchBytes) + " (use kbg_batch_step or a smaller batch)");
    } else {
        if (descs.size() > ctx->descs_cap) {
            (void)hipFree(ctx->d_descs);
            ctx->d_descs = nullptr;
            ctx->descs_cap = descs.size() * 2;
            HIPCHK(hipMalloc((void**)&ctx->d_descs, ctx->descs_cap * sizeof(Solve1D)));
        }
        if (items.size() > ctx->items_cap) {
            (void)hipFree(ctx->d_items);
            (void)hipFree(ctx->d_rpart);
            ctx->d_items = nullptr;
            ctx->d_rpart = nullptr;
            ctx->items_cap = items.size() * 2;
            HIPCHK(hipMalloc((void**)&ctx->d_items, ctx->items_cap * sizeof(WorkItem)));
            HIPCHK(hipMalloc((void**)&ctx->d_rpart, ctx->items_cap * 12 * sizeof(double)));
        }
        if (packs.size() > ctx->packs_cap) {
            (void)hipFree(ctx->d_packs);
            ctx->d_packs = nullptr;
            ctx->packs_cap = packs.size() * 2;
            HIPCHK(hipMalloc((void**)&ctx->d_packs, ctx->packs_cap * sizeof(kbg::Pack)));
        }
        if (descs2.size() > ctx->descs2_cap) {
            (void)hipFree(ctx->d_descs2);
            ctx->d_descs2 = nullptr;
            ctx->descs2_cap = descs2.size() * 2;
            HIPCHK(hipMalloc((void**)&ctx->d_descs2, ctx->descs2_cap * sizeof(Solve2D)));
        }
        if (pref2.size() > ctx->pref2_cap) {
            (void)hipFree(ctx->d_pref2);
            ctx->d_pref2 = nullptr;
            ctx->pref2_cap = pref2.size() * 2;
            HIPCHK(hipMalloc((void**)&ctx->d_pref2, ctx->pref2_cap * sizeof(int)));
        }
        d_descs = ctx->d_descs;
        d_items = ctx->d_items;
        d_rpart = ctx->d_rpart;
        d_packs = ctx->d_packs;
        d_descs2 = ctx->d_descs2;
        d_pref2 = ctx->d_pref2;
    }
    if (!descs.empty()) {
        HIPCHK(hipMemcpyAsync(d_descs, descs.data(), descs.size() * sizeof(Solve1D), hipMemcpyHostToDevice, st));
        HIPCHK(hipMemcpyAsync(d_items, items.data(), items.size() * sizeof(WorkItem), hipMemcpyHostToDevice, st));
        HIPCHK(hipMemcpyAsync(d_packs, packs.data(), packs.size() * sizeof(kbg::Pack), hipMemcpyHostToDevice, st));
    }
    if (!descs2.empty()) {
        HIPCHK(hipMemcpyAsync(d_descs2, descs2.data(), descs2.size() * sizeof(Solve2D), hipMemcpyHostToDevice, st));
        HIPCHK(hipMemcpyAsync(d_pref2, pref2.data(), pref2.size() * sizeof(int), hipMemcpyHostToDevice, st));
    }
    if (!descs.empty() || !descs2.empty()) HIPCHK(hipStreamSynchronize(st));
    // fused FP32 solves left holding f by phase-level calls go back to departures (in place)
    for (int i : fused) {
        kbg_solve* s = solves[i];
        if (s->h1.delta && s->rep_full) {
            HIPCHK(kbg::launch_delta1d(s->h1, d_descs + desc_of[i], nullptr, 3, s->h1.cur, st));
            s->rep_full = false;
        }
    }
    for (int i : fused2) {
        kbg_solve* s = solves[i];
        if (s->h2.delta && s->rep_full) {
            if ((rc = push2d(s, st))) return rc;
            HIPCHK(kbg::launch_delta2d(s->h2, s->d2, 3, st));
            s->rep_full = false;
        }
    }
    // stream k > 0 is side stream k - 1; they fork from and join back into the context stream
    const int nside = nstreams - 1;
    std::vector<hipStream_t> sst(nstreams, st);
    for (int k = 0; k < nside; k++) sst[k + 1] = ctx->side[k];
    auto fork = [&]() -> int {
        if (nside > 0) {
            HIPCHK(hipEventRecord(ctx->fork, st));
            for (int k = 0; k < nside; k++) HIPCHK(hipStreamWaitEvent(ctx->side[k], ctx->fork, 0));
        }
        return KBG_OK;
    };
    auto join = [&]() -> int {
        for (int k = 0; k < nside; k++) {
            HIPCHK(hipEventRecord(ctx->join[k], ctx->side[k]));
            HIPCHK(hipStreamWaitEvent(st, ctx->join[k], 0));
        }
        return KBG_OK;
    };
    // one round of every unit on its stream: nst time steps (1 or 2; a fused 1D sweep launch whose
    // solves allow it takes both in one pass, the other units are launched nst times); each sweep
    // launch reads the f / w buffers of its own parity (launches so far), last marks the round whose
    // residual rows are kept
    // ubase: the parity of the round's first step among the call's steps (unstructured solves flip
    // their f buffer every step)
    auto enqueue_round = [&](int nst, bool last_round, int ubase) -> int {
        int e;
        for (Launch& L : launches) {
            const int per = L.two ? nst : 1;  // steps per launch
            for (int ks = 0; ks < nst; ks += per) {
                const int par = L.flips & 1;
                for (int di : L.members) {
                    const Solve1D& S = descs[desc_of[di]];
                    if (S.wallL || S.wallR)
                        HIPCHK(kbg::launch_wall1d(solves[di]->fp64, S, d_descs + desc_of[di], 1, par, sst[L.stream]));
                }
                const bool keep = last_round && residuals && ks + per == nst;
                HIPCHK(kbg::launch_sweep1d(L.fp64, L.nf, L.flux, L.wg_waves, d_descs, d_items, d_packs + L.pack0, (int)L.npacks,
                                           par, keep ? d_rpart : nullptr, per, sst[L.stream]));
                L.flips++;
            }
        }
        for (int ks = 0; ks < nst; ks++) {
        const bool last = last_round && ks == nst - 1;
        for (size_t g = 0; g < launches2.size(); g++) {
            const Launch2& L = launches2[g];
            const Solve2D* dd = d_descs2 + L.desc0;
            hipStream_t ss = sst[stream2[g]];
            HIPCHK(kbg::launch_face2d(true, L.flux, dd, d_pref2 + L.ipref0, L.n, L.ninterior, ss));
            HIPCHK(kbg::launch_flux2d(L.fp64, L.flux, L.threads, dd, d_pref2 + L.fpref0, L.n, L.nfaces, ss));
            HIPCHK(kbg::launch_cell2d(L.fp64, L.threads, dd, d_pref2 + L.cpref0, L.n, L.ncells, (last && residuals) ? 1 : 0, ss));
        }
        for (int i : others) {
            kbg_solve* s = solves[i];
            hipStream_t ss = sst[stream_of[i]];
            double* r = (last && residuals) ? residuals + 5 * (size_t)i : nullptr;
            if (s->us) {
                if ((e = step_us(s, (ubase + ks) & 1, r ? 1 : 0, ss))) return e;
                if (r) {
                    HIPCHK(kbg::launch_uresid(s->rcell, s->hus.ncell, s->res, ss));
                    HIPCHK(hipMemcpyAsync(r, s->res, 4 * 8, hipMemcpyDeviceToHost, ss));
                    HIPCHK(hipStreamSynchronize(ss));
                }
            } else if (s->d.dim == 1) {
                HIPCHK(kbg::launch_recon1d(s->fp64, s->h1, s->d1, ss));
                HIPCHK(kbg::launch_evolve1d(s->fp64, s->h1, s->d1, ss));
                HIPCHK(kbg::launch_wall1d(s->fp64, s->h1, s->d1, 0, 0, ss));
                if ((e = step_phase_1d(s, dts[i], r, ss))) return e;
            } else {
                HIPCHK(kbg::launch_recon2d(s->fp64, s->h2, s->d2, ss));
                HIPCHK(kbg::launch_evolve2d(s->fp64, s->h2, s->d2, ss));
                if ((e = step_phase_2d(s, dts[i], r, ss))) return e;
            }
        }
        }
        return KBG_OK;
    };
    // Without residuals the steps are replayed from HIP graphs: rounds of two steps and of one step
    // at both parities of the two-step launches, captured once per launch plan (the signature below: every launch's shape, stream and
    // arguments; descriptors are read from device memory, re-uploaded by every call) and reused
    // by later calls on the same batch.
    bool any_fsm = false;
    for (int i : others) any_fsm = any_fsm || solves[i]->has_fsm;
    // hipFFT calls stay out of stream capture: batches with :fsm solves launch eagerly
    const bool use_graph = ctx->graphs && !residuals && nsteps >= 2 && !noalloc && !any_fsm;
    std::string sig;
    if (use_graph) {
        auto put = [&](const void* p, size_t b) { sig.append(static_cast<const char*>(p), b); };
        auto puti = [&](long long v) { put(&v, sizeof v); };
        puti(nstreams);
        const void* bufs[] = {d_descs, d_items, d_packs, d_descs2, d_pref2};
        put(bufs, sizeof bufs);
        for (Solve1D d : descs) {
            d.cur = 0;  // the parity base flips between calls; the kernels read it from the descriptor
            put(&d, sizeof d);
        }
        put(items.data(), items.size() * sizeof(WorkItem));
        put(packs.data(), packs.size() * sizeof(kbg::Pack));
        for (const Launch& L : launches) {
            for (long long v : {(long long)L.fp64, (long long)L.nf, (long long)L.flux, (long long)L.wg_waves, (long long)L.stream,
                                (long long)L.pack0, (long long)L.npacks, (long long)L.two})
                puti(v);
            for (int di : L.members) {
                puti(di);
                puti(desc_of[di]);
                puti(solves[di]->fp64);
            }
            puti(-1);
        }
        put(descs2.data(), descs2.size() * sizeof(Solve2D));
        put(pref2.data(), pref2.size() * sizeof(int));
        for (size_t g = 0; g < launches2.size(); g++) {
            const Launch2& L = launches2[g];
            for (long long v : {(long long)L.fp64, (long long)L.flux, (long long)L.threads, (long long)L.n, (long long)L.ncells,
                                (long long)L.nfaces, (long long)L.ninterior, (long long)L.desc0, (long long)L.cpref0,
                                (long long)L.fpref0, (long long)L.ipref0, (long long)stream2[g]})
                puti(v);
        }
        for (int i : others) {
            const kbg_solve* s = solves[i];
            const void* ids[] = {s, s->d1, s->d2, s->dus};
            put(ids, sizeof ids);
            puti(stream_of[i]);
            if (s->us) {
                kbg::SolveUS h = s->hus;
                h.dt = 0.0;  // dt and the buffer base are read from the device descriptor, pushed by every call
                h.cur = 0;
                put(&h, sizeof h);
            } else if (s->d.dim == 1) {
                Solve1D h = s->h1;
                h.cur = 0;
                put(&h, sizeof h);
            } else {
                put(&s->h2, sizeof s->h2);
            }
        }
        if (sig != ctx->gsig || !ctx->gexec[0] || !ctx->gexec[1] || !ctx->gexec[2] || !ctx->gexec[3]) {
            ctx->drop_graphs();
            for (int which = 0; which < 4; which++) {
                HIPCHK(hipStreamBeginCapture(st, hipStreamCaptureModeRelaxed));
                for (Launch& L : launches) L.flips = L.two ? which >> 1 : 0;
                int e = fork();
                if (!e) e = enqueue_round((which & 1) || !two ? 1 : 2, false, 0);
                if (!e && !(which & 1) && !two) e = enqueue_round(1, false, 1);
                if (!e) e = join();
                hipGraph_t g = nullptr;
                const hipError_t ce = hipStreamEndCapture(st, &g);
                if (e || ce != hipSuccess) {
                    if (g) (void)hipGraphDestroy(g);
                    ctx->drop_graphs();
                    return e ? e : fail(KBG_ERR_HIP, std::string("hipStreamEndCapture: ") + hipGetErrorString(ce));
                }
                ctx->graph[which] = g;
                HIPCHK(hipGraphInstantiate(&ctx->gexec[which], g, nullptr, nullptr, 0));
            }
            ctx->gsig = std::move(sig);
        }
    }
    HIPCHK(hipEventRecord(ctx->ev[KBG_EV_STEPS_BEGIN], st));
    for (Launch& L : launches) L.flips = 0;
    if (use_graph) {
        const int pairs = nsteps / 2;
        for (int t = 0; t < pairs; t++) HIPCHK(hipGraphLaunch(ctx->gexec[(t & 1) << 1], st));
        if (nsteps & 1) HIPCHK(hipGraphLaunch(ctx->gexec[(pairs & 1) << 1 | 1], st));
        for (Launch& L : launches) L.flips = L.two ? pairs + (nsteps & 1) : nsteps;
    } else {
        if ((rc = fork())) return rc;
        for (int t = 0; t < nsteps;) {
            const int nst = two && nsteps - t >= 2 ? 2 : 1;
            if ((rc = enqueue_round(nst, t + nst == nsteps, t & 1))) return rc;
            t += nst;
        }
        if ((rc = join())) return rc;
    }
    HIPCHK(hipEventRecord(ctx->ev[KBG_EV_STEPS_END], st));
    // residuals of the fused solves: per-item partial rows reduced per solve
    if (residuals) {
        std::vector<int> fused_idx;
        for (int i = 0; i < n; i++)
            if (desc_of[i] >= 0) fused_idx.push_back(i);
        if (!fused_idx.empty()) {
            std::vector<double> rp(items.size() * 12);
            HIPCHK(hipMemcpyAsync(rp.data(), d_rpart, rp.size() * 8, hipMemcpyDeviceToHost, st));
            HIPCHK(hipStreamSynchronize(st));
            std::vector<double> acc(descs.size() * 12, 0.0);
            for (size_t it = 0; it < items.size(); it++)
                for (int m = 0; m < 12; m++) acc[items[it].solve * 12 + m] += rp[it * 12 + m];
            for (int i : fused_idx) {
                const double* a = acc.data() + desc_of[i] * 12;
                const double nx = solves[i]->d.nx;
                for (int m = 0; m < 3; m++)
                    residuals[5 * (size_t)i + m] = std::sqrt(a[m] * nx) / (a[3 + m] + 1.e-7) + std::sqrt(a[6 + m] * 2.0) / (a[9 + m] + 1.e-7);
            }
        }
    }
    // the buffer flip is recorded on the host; a solve's device descriptor (d1) is uploaded by
    // every entry point that launches with it, right before the launch, so it is not pushed here
    // (one synchronous copy per solve: 7.5 ms for a batch of 512)
    for (int i = 0; i < n; i++) {
        if (desc_of[i] < 0) continue;
        kbg_solve* s = solves[i];
        s->h1.cur = (s->h1.cur + launches[launch_of[i]].flips) & 1;
    }
    for (int i : others)
        if (solves[i]->us) solves[i]->hus.cur = (solves[i]->hus.cur + nsteps) & 1;
    // fused 2D solves: residuals from the last step's rows
    if (residuals) {
        for (int i : fused2) {
            kbg_solve* s = solves[i];
            HIPCHK(kbg::launch_resid2d(s->h2, s->rcell, s->res, st));
            HIPCHK(hipMemcpyAsync(residuals + 5 * (size_t)i, s->res, 4 * 8, hipMemcpyDeviceToHost, st));
        }
    }
    HIPCHK(hipStreamSynchronize(st));
    return KBG_OK;
}

// update! on the phase-exact path (kernels k_update1d / k_resid), optional residual download
int step_phase_1d(kbg_solve* s, double dt, double* residual, hipStream_t st) {
    (void)dt;
    double* rc_ = residual ? s->rcell : nullptr;
    if (s->has_fsm) {
        // Q(f^n) of the interior cells before update! changes f (step_1f.jl:146-151)
        const Solve1D& S = s->h1;
        HIPCHK(kbg::fsm_collision(s->fsm, s->fp64, S.f[S.cur], S.fcs, s->fsm.c0 - 1 + S.ng, st));
    }
    HIPCHK(kbg::launch_update1d(s->fp64, s->h1, s->d1, rc_, st));
    if (residual) {
        const int nrows = s->h1.upd_hi - s->h1.upd_lo + 1;
        HIPCHK(kbg::launch_resid1d(s->h1, s->rcell, nrows, 4 * s->nw, s->res, st));
        HIPCHK(hipMemcpyAsync(residual, s->res, s->nw * 8, hipMemcpyDeviceToHost, st));
        HIPCHK(hipStreamSynchronize(st));
    }
    return KBG_OK;
}

int step_phase_2d(kbg_solve* s, double dt, double* residual, hipStream_t st) {
    (void)dt;
    double* rc_ = residual ? s->rcell : nullptr;
    HIPCHK(kbg::launch_update2d(s->fp64, s->h2, s->d2, rc_, st));
    if (residual) {
        HIPCHK(kbg::launch_resid2d(s->h2, s->rcell, s->res, st));
        HIPCHK(hipMemcpyAsync(residual, s->res, 4 * 8, hipMemcpyDeviceToHost, st));
        HIPCHK(hipStreamSynchronize(st));
    }
    return KBG_OK;
}

}  // namespace
