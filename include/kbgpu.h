/*
 * kbgpu.h -- C ABI of libkbgpu.so, the MI355X (gfx950) discrete-velocity time stepper that replaces
 * KitBase.jl's CPU `reconstruct!` / `evolve!` / `update!` hot loop.
 *
 * Plain pointers and sizes only.  All host arrays are FP64, contiguous, owned by the caller (they
 * need to live only through the call that takes them: kbg_solve_create copies what a descriptor
 * points to), and laid out the way KitBase stores the same data once its OffsetArrays are flattened:
 *   1D  cells  k = i - 1 + ng for the physical index i in [1-ng, nx+ng]  (N = nx + 2 ng rows)
 *       faces  k - 1 for the physical face k in [1, nx+1] (between cells k-1 and k)
 *   2D  cells  (i-1) + nx (j-1);  x-faces (i-1) + (nx+1)(j-1), i in [1, nx+1];
 *       y-faces (i-1) + nx (j-1), j in [1, ny+1]
 *   per cell:  w, prim [nw];  f [nf][nv] (h then b);  sw [nw] (2D: [2 dirs][nw]);
 *              sf [nf][nv] (2D: [2 dirs][nf][nv])
 *   per face:  fw [nw];  ff [nf][nv]
 *   velocity nodes u-fastest (Julia column-major), nv points in total.
 * nw = 3 (1D1V), 4 (2D2V), 5 (1D3V).  prim = [rho, U, (V), (W), lambda], lambda = 1/(2RT).
 *
 * Every call returns KBG_OK (0) or a negative KBG_ERR_* code; kbg_last_error() then describes the
 * failure (thread-local).  Calls on one kbg_solve must come from one host thread at a time; calls
 * are ordered on the context's HIP stream and return once their results are visible to the host
 * (downloads, residuals) or queued (steps).  There is no CPU fallback: a missing or unusable GPU is
 * an error.
 *
 * Citations are relative to the KitBase.jl v0.9.31 source tree.
 */
#ifndef KBGPU_H
#define KBGPU_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

#define KBG_OK 0
#define KBG_ERR_ARG (-1)
#define KBG_ERR_HIP (-2)
#define KBG_ERR_UNSUPPORTED (-3)
#define KBG_ERR_STATE (-4)
#define KBG_ERR_NOMEM (-5)

enum { KBG_FP32 = 0, KBG_FP64 = 1 };
/* flux symbols of evolve!(...; mode) -- solver_evolution.jl:55 evaluates "flux_" * mode * "!" */
enum { KBG_KFVS = 0, KBG_KCU = 1, KBG_UGKS = 2 };
/* collision symbols of update!(...; coll) -- step_1f.jl:30, step_2f.jl:34; :fsm is the fast spectral
 * Boltzmann operator of the 1D1F3V step! (step_1f.jl:122-157, theory_fsm.jl:271-314) */
enum { KBG_BGK = 0, KBG_SHAKHOV = 1, KBG_FSM = 2 };
/* boundary symbols: :fix (boundary cells not stepped) and :maxwell (diffuse walls) */
enum { KBG_BC_FIX = 0, KBG_BC_MAXWELL = 1 };
/* limiter symbols -- recon_limiter.jl:10-105 */
enum { KBG_VANLEER = 0, KBG_MINMOD = 1, KBG_SUPERBEE = 2, KBG_VANALBABA = 3, KBG_LINEAR = 4 };

typedef struct kbg_ctx kbg_ctx;
typedef struct kbg_solve kbg_solve;

/* One solve: the fields of KitBase's SolverSet that the time step reads (solver_set.jl:10-117:
 * Setup struct_setup.jl:106-119, PSpace1D/2D geo_struct.jl:10-175, VSpace1D/2D/3D
 * phase_velocity.jl:14-258, Gas struct_property.jl:51-64, the IB boundary primitive vectors). */
typedef struct kbg_desc {
    int dim;          /* 1 or 2 */
    int nf;           /* 1 (f) or 2 (h, b) */
    int vdim;         /* velocity dimension: 1 or 3 in 1D, 2 in 2D */
    int nv;           /* total velocity points */
    int nx, ny;       /* physical cells (ny = 1 in 1D) */
    int ng;           /* ghost layers (1D; must be 0 in 2D) */
    int precision;    /* KBG_FP32 or KBG_FP64 storage of f and the per-velocity arithmetic */
    int flux;         /* KBG_KFVS, KBG_KCU, KBG_UGKS */
    int collision;    /* KBG_BGK, KBG_SHAKHOV */
    int limiter;      /* KBG_VANLEER, ... */
    int interp_order; /* 1 or 2 (Setup.interpOrder) */
    int bc[4];        /* 1D: left, right; 2D: left, right, down, up -- KBG_BC_FIX / KBG_BC_MAXWELL */
    double K, gamma, mu_ref, omega, Pr; /* Gas.K, Gas.γ, Gas.μᵣ, Gas.ω, Gas.Pr */
    const double *u, *v, *w, *weights;  /* [nv]; v and w may be NULL when unused */
    const double *dx;                   /* 1D: [nx + 2 ng]; 2D: [nx*ny] */
    const double *dy;                   /* 2D: [nx*ny] */
    const double *bc_prim;              /* 1D Maxwell walls: [2][nw] prim at x0 and x1 (KS.ib.bc) */
    const double *xgeo, *ygeo;          /* 2D: per-face [n1, n2, len] (geo_struct.jl normals/areas) */
    const double *xwall, *ywall;        /* 2D: per-face wall prim [4] in the global frame */
    /* collision KBG_FSM (1D1F3V on a uniform nu x nv x nw grid, u fastest): Gas.fsm of KitBase's
     * gas_property(vs; ..., collision = "fsm") (struct_property.jl:516-567): Kn = hs_boltz_kn(mu_ref,
     * alphaRef), nm = M and the kernel modes phi, psi [M(M-1)][nw][nv][nu], phipsi [nw][nv][nu]
     * (kernel_mode, theory_fsm.jl:24-90).  The operator's transforms run in the storage precision
     * (Q is summed in FP64).  Ignored for other collisions. */
    int fsm_nm;
    int fsm_shape[3];                   /* nu, nv, nw */
    double fsm_kn;
    const double *fsm_phi, *fsm_psi, *fsm_phipsi;
} kbg_desc;

/* Library version (major*10000 + minor*100 + patch) and the calling thread's last error. */
int kbg_version(void);
const char *kbg_last_error(void);

/* Context: one device, one stream, tuning state.  No counterpart in KitBase (it has no device). */
int kbg_ctx_create(int device, kbg_ctx **out);
int kbg_ctx_destroy(kbg_ctx *ctx);
int kbg_ctx_sync(kbg_ctx *ctx);
/* Schedule of the fused 1D sweep: cells per workgroup (1..60, default 60) and the number of
 * concurrent streams a batch spreads its kernel groups over (1..4; 0 = all 4). */
int kbg_ctx_set_schedule(kbg_ctx *ctx, int seg_cells, int streams);
int kbg_ctx_get_schedule(kbg_ctx *ctx, int *seg_cells, int *streams);
/* Replay of batch steps through HIP graphs (default off): kbg_batch_step / kbg_step calls without
 * residuals capture the launches of a step pair once per batch composition and replay them; off
 * launches every kernel one by one.  Same results either way; graphs cut host launch work when the
 * host, not the device, bounds the step rate.  No counterpart in KitBase. */
int kbg_ctx_set_graphs(kbg_ctx *ctx, int enable);
/* Time steps per launch of the fused 1D sweep (default 2): with 2, batches whose fused 1D solves all
 * use KFVS without boundary-face buffers (no Maxwell walls, no stepped cell at a ghost-less
 * boundary) advance two steps per pass over f -- a 4-cell halo per 56-cell segment, one read and
 * one write of f per two steps, bit-identical results; 1 always launches one step at a time.
 * No counterpart in KitBase. */
int kbg_ctx_set_sweep_steps(kbg_ctx *ctx, int steps);
/* Most velocities per wavefront of the fused 1D sweep: 0 (default) takes the kernel's full width
 * (64 FP32 velocities of one distribution); 8, 16 or 32 spread each segment's velocity grid over
 * more wavefronts (at most 8), which shortens the pass of small, latency-bound batches (serving)
 * at some cost in throughput on large ones.  A solve's wavefront count depends only on its velocity
 * count and this setting, so its results are the same in any batch.  No counterpart in KitBase. */
int kbg_ctx_set_sweep_vw(kbg_ctx *ctx, int vw);
/* HIP events on the context stream, for timing: record slot (0..15), elapsed ms between slots.
 * kbg_step / kbg_batch_step record slots KBG_EV_STEPS_BEGIN and KBG_EV_STEPS_END around their
 * time-step loop (after the descriptor upload, before the residual download), so the elapsed
 * time between them is the device time of the nsteps steps alone. */
enum { KBG_EV_STEPS_BEGIN = 14, KBG_EV_STEPS_END = 15 };
int kbg_ctx_event_record(kbg_ctx *ctx, int slot);
int kbg_ctx_event_elapsed(kbg_ctx *ctx, int slot_a, int slot_b, float *ms);

/* A device-resident solve: replaces the arrays of ControlVolume1F/2F (struct_ctr.jl:51-241) and
 * Interface1F/2F (struct_face.jl:37-101) built by init_fvm (solver_init.jl:62-327). */
int kbg_solve_create(kbg_ctx *ctx, const kbg_desc *desc, kbg_solve **out);
int kbg_solve_destroy(kbg_solve *s);
/* Device bytes a solve with this description occupies on the fused path (admission control). */
size_t kbg_solve_bytes(const kbg_desc *desc);

/* An unstructured triangle solve (2D1F2V or 2D2F2V): the fields of KitBase's UnstructPSpace
 * (geo_unstruct.jl:10-32) and the geometry init_fvm derives from them (solver_init.jl:387-640), with
 * the VSpace2D / Gas / wall fields its evolve! (solver_evolution.jl:872-1035) and update!
 * (solver_update.jl:359-471) read.  Cells and faces are 0-based; -1 marks "no cell".  Faces take
 * KitBase's first-order KFVS flux (flux_kfvs.jl:615-697) or its diffuse Maxwell wall
 * (flux_boundary.jl:160-220, with rot = 1: the normal of a boundary face points into the mesh).
 * Host arrays of such a solve: w, prim [ncell][4]; f [ncell][nf][nv]; fw [nface][4] (global frame);
 * ff [nface][nf][nv].  kbg_upload / kbg_download / kbg_*_faces / kbg_evolve / kbg_update / kbg_step /
 * kbg_batch_step / kbg_batch_step_noalloc / kbg_solve_health take it like any solve (residuals: 4
 * values).  The unstructured evolve! extrapolates each cell's f to the face centre with the cell's
 * slopes, which stay zero at interpOrder 1 (reconstruct! returns at once,
 * solver_reconstruction.jl:809-880): that first-order scheme is what is built, so kbg_reconstruct
 * does nothing and the slope calls fail with KBG_ERR_UNSUPPORTED.  interpOrder 2 is refused: its
 * reconstruct! (reconstruct3!/reconstruct4!, reconstruction.jl:136-254) is not built (for 1F it
 * reads an undefined `ps`; only vanleer and minmod take three slopes; cells whose neighbours' centres
 * share an x or y coordinate divide by zero).  kbg_timestep / kbg_batch_timestep take the unstructured
 * timestep (solver_main.jl:400-426): dt = cfl / max over cells of ((u1 + c)^2 + (v1 + c)^2)^(1/2) /
 * cell_dmin, clipped to max_time - sim_time. */
typedef struct kbg_udesc {
    int nf;        /* 1 (f) or 2 (h, b) */
    int nv;        /* velocity points, u-fastest; at most 4096 */
    int ncell, nface;
    int precision; /* KBG_FP32 or KBG_FP64 */
    int collision; /* KBG_BGK, or KBG_SHAKHOV with nf = 2 (step_1f.jl's triangle Shakhov reads h and b) */
    int interp_order; /* Setup.interpOrder: 1.  2 is refused (KBG_ERR_UNSUPPORTED): see kbg_usolve_create */
    double K, gamma, mu_ref, omega, Pr;
    const double *u, *v, *weights; /* [nv] */
    const int *face_cells;         /* [nface][2] ps.faceCells */
    const int *face_kind;          /* [nface] 0 interior (KFVS); 1 Maxwell wall (a boundary face of a
                                      cellType-2 cell); 2 a boundary face whose stored flux is kept */
    const double *face_geo;        /* [nface][3] n1, n2, length (face.n, face.len of init_fvm) */
    const double *face_wall;       /* [nface][4] wall prim in the global frame (KS.ib.bc at the face
                                      centre), read for kind-1 faces; NULL when there are none */
    const int *cell_faces;         /* [ncell][3] ps.cellFaces */
    const double *cell_dirc;       /* [ncell][3] sign(dot(ctr.n[j], face.n)) of update! (solver_update.jl:375) */
    const double *cell_area;       /* [ncell] ps.cellArea */
    const int *cell_type;          /* [ncell] ps.cellType: 0 and 2 stepped, 3 the mean of its first two
                                      neighbours (update_boundary.jl:520-590), others kept */
    const int *cell_nb;            /* [ncell][3] ps.cellNeighbors */
    const double *cell_dmin;       /* [ncell] least of ctr.dx (timestep); NULL: the timestep calls fail */
} kbg_udesc;

int kbg_usolve_create(kbg_ctx *ctx, const kbg_udesc *desc, kbg_solve **out);
/* Device bytes an unstructured solve with this description occupies. */
size_t kbg_usolve_bytes(const kbg_udesc *desc);

/* Failure detection.  No counterpart in KitBase, where a NaN or a negative temperature propagates
 * silently: every cell update checks the cell's new primitive variables, and a cell whose values are
 * not finite (KBG_HEALTH_NONFINITE), whose density is not positive (KBG_HEALTH_RHO) or whose
 * lambda = 1/(2RT) is not positive (KBG_HEALTH_LAMBDA) raises that bit in the solve's flag word.  The
 * step completes either way; `flags` receives the word after the solve's queued work, and clear != 0
 * resets it. */
enum { KBG_HEALTH_NONFINITE = 1, KBG_HEALTH_RHO = 2, KBG_HEALTH_LAMBDA = 4 };
int kbg_solve_health(kbg_solve *s, int *flags, int clear);

/* Host <-> device copies of the cell and face state (FP64 on the host, converted on the fly). */
int kbg_upload(kbg_solve *s, const double *w, const double *prim, const double *f);
int kbg_download(kbg_solve *s, double *w, double *prim, double *f);
int kbg_upload_slopes(kbg_solve *s, const double *sw, const double *sf);
int kbg_download_slopes(kbg_solve *s, double *sw, double *sf);
/* 2D: x-faces first, then y-faces, in both arrays. */
int kbg_upload_faces(kbg_solve *s, const double *fw, const double *ff);
int kbg_download_faces(kbg_solve *s, double *fw, double *ff);

/* reconstruct!(KS, ctr)                  solver_reconstruction.jl:56-153 (1D), :475-620 (2D) */
int kbg_reconstruct(kbg_solve *s);
/* evolve!(KS, ctr, face, dt; mode, bc)   solver_evolution.jl:39-89 (1D), :608-679 (2D),
 *                                        evolve_boundary.jl:4-178 (Maxwell walls) */
int kbg_evolve(kbg_solve *s, double dt);
/* update!(KS, ctr, face, dt, residual; coll, bc)  solver_update.jl:61-93 (1D), :193-246 (2D),
 * update_boundary.jl:6-306.  residual: [nw] or NULL. */
int kbg_update(kbg_solve *s, double dt, double *residual);

/* timestep(KS, ctr, simTime)            solver_main.jl:240-342, from the solve's current prim on the
 * device: dt = cfl / max over physical cells of (max(u1, |U|) + c) / dx (2D: + (max(v1, |V|) + c) / dy),
 * c = (gamma / (2 lambda))^(1/2) (theory_macro.jl:421), then min(dt, max_time - sim_time).  u1, v1 are
 * the velocity-space bounds VSpace.u1 / v1 (v1 unused in 1D). */
int kbg_timestep(kbg_solve *s, double cfl, double u1, double v1, double sim_time, double max_time, double *dt);

/* nsteps iterations of solve!'s loop body (solver_main.jl:32-51) with a fixed dt, fused on the
 * device; residual (of the last step) may be NULL.  Slopes and face arrays are not kept. */
int kbg_step(kbg_solve *s, double dt, int nsteps, double *residual);

/* One solve!-loop iteration for every solve of a batch, nsteps times: the 1D solves share one
 * fused launch per step; the others are stepped in turn on the same stream.  dts[i] is solve i's
 * time step; residuals (n x 5, may be NULL) receive the last step's residuals.  The solves may have
 * been created on any context of ctx's device: the batch runs on ctx's streams, while a solve's own
 * calls (upload, download, phase calls, destroy) run on its creating context's stream, so another
 * host thread may upload or download solves of a second context while the batch runs, as long as
 * none of them is in the batch (a serving loop overlaps its transfers with the stepping this way). */
int kbg_batch_step(kbg_ctx *ctx, kbg_solve *const *solves, int n, const double *dts, int nsteps,
                   double *residuals);

/* timestep (solver_main.jl:240-342) of every solve of a batch from its current state on the device:
 * per solve i, dts[i] = min(cfl[i] / max over cells of (max(u1[i], |U|) + c) / dx (2D: + the y
 * term with v1[i]), max_time[i] - sim_time[i]); sim_time / max_time may be NULL (no clip), v1 NULL
 * in a 1D-only batch.  One reduction per solve, one synchronisation. */
int kbg_batch_timestep(kbg_ctx *ctx, kbg_solve *const *solves, int n, const double *cfl, const double *u1,
                       const double *v1, const double *sim_time, const double *max_time, double *dts);
/* kbg_batch_step with each solve's dt taken from kbg_batch_timestep at the call (per-instance
 * adaptive dt, recomputed every call; KitBase's solve! keeps the first dt for the whole run):
 * the dts used are returned in dts[n]. */
int kbg_batch_step_cfl(kbg_ctx *ctx, kbg_solve *const *solves, int n, const double *cfl, const double *u1,
                       const double *v1, const double *sim_time, const double *max_time, int nsteps, double *dts,
                       double *residuals);

/* kbg_batch_step without any device allocation.  No counterpart in KitBase: it is for callers that
 * cannot allocate device memory while they step (code under HIP graph capture, tight serving loops,
 * harnesses that forbid hipMalloc).  Same arguments, kernels and results as kbg_batch_step; the
 * batch's device staging (1D descriptors, work items, packs and residual partial rows; 2D
 * descriptors and prefix sums, about 0.3 KB per solve and 0.13 KB per 56-60 cells of a 1D solve)
 * lives in a fixed per-device array of KBG_NOALLOC_SCRATCH_BYTES that the library reserves for the
 * life of the process.  A batch that needs more staging fails with KBG_ERR_ARG, as does one with a
 * solve stepped by the phase kernels (1D1F3V, very large velocity grids) whose slope and face
 * buffers no earlier call has allocated (KBG_ERR_STATE).  HIP graph replay is not used.  The array
 * is shared by every context on a device and keeps the last call's contents: calls that overlap in
 * time on one GPU (from two host threads) race; serialise them. */
#define KBG_NOALLOC_SCRATCH_BYTES ((size_t)8 << 20)
int kbg_batch_step_noalloc(kbg_ctx *ctx, kbg_solve *const *solves, int n, const double *dts, int nsteps,
                           double *residuals);

#ifdef __cplusplus
}
#endif

#endif /* KBGPU_H */
