/*
 * kitbase_oracle.h -- FP64 CPU restatement of KitBase.jl's discrete-velocity time step.
 *
 * TEST INFRASTRUCTURE ONLY.  The HIP path is tested against this library: only tests/,
 * __graft_entry__.smoke() and bench.py's cpu_baseline leg may load it.  The product library
 * (libkbgpu.so) never links, loads or calls it.
 *
 * Every function restates one KitBase.jl v0.9.31 function; the citation (file:line, relative to
 * the KitBase source tree) is given beside each declaration and again in the implementation.
 * Arrays are plain C arrays of double.  Velocity-space arrays are flattened with the first
 * velocity index fastest (Julia column-major), so u[i + nu*j] is the 2V node (i, j).
 */
#ifndef KITBASE_ORACLE_H
#define KITBASE_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

/* enum values shared with the product ABI (include/kbgpu.h) */
enum { ORC_KFVS = 0, ORC_KCU = 1, ORC_UGKS = 2 };
enum { ORC_BGK = 0, ORC_SHAKHOV = 1, ORC_FSM = 2 };
enum { ORC_BC_FIX = 0, ORC_BC_MAXWELL = 1 };
enum { ORC_VANLEER = 0, ORC_MINMOD = 1, ORC_SUPERBEE = 2, ORC_VANALBABA = 3, ORC_LINEAR = 4 };
enum { ORC_RECTANGLE = 0, ORC_NEWTON = 1, ORC_ALGEBRA = 2 };

/* ---- math helpers (src/Math/math_general.jl) ---- */
double orc_heaviside(double x);                          /* :13 */
double orc_fortsign(double x, double y);                 /* :20 */
double orc_limiter(int kind, double sL, double sR);      /* src/Reconstruction/recon_limiter.jl:10-105 */

/* ---- phase / physical space (setup) ---- */
double orc_newton_cotes(int idx, int num);               /* src/Phase/phase_quadrature.jl:276-288 */
void orc_vspace1d(int type, double U0, double U1, int nu, int ng,
                  double *u, double *du, double *weights);       /* src/Phase/phase_velocity.jl:23-65 */
int orc_maxwell_quadrature(int n2, double C, double *x, double *w); /* src/Phase/phase_quadrature.jl:13-44 */
void orc_pspace2d(double X0, double X1, int nx, double Y0, double Y1, int ny,
                  double *x, double *y, double *dx, double *dy,
                  double *areas, double *normals);               /* src/Geometry/geo_struct.jl:64-175 */

/* ---- theory (src/Theory) ---- */
void orc_prim_conserve(const double *prim, int n, double gam, double *W);   /* theory_macro.jl:10-33 */
void orc_conserve_prim(const double *W, int n, double gam, double *prim);   /* theory_macro.jl:185-206 */
double orc_heat_capacity_ratio(double K, int D);                            /* theory_macro.jl:376-384 */
double orc_sound_speed(double lam, double gam);                             /* theory_macro.jl:421 */
double orc_ref_vhs_vis(double Kn, double alpha, double omega);              /* theory_macro.jl:447-449 */
double orc_vhs_collision_time(double rho, double lam, double mu, double om);/* theory_relaxation.jl:28 */
void orc_maxwellian_1v(const double *u, int n, double rho, double U, double lam, double *M);  /* theory_equilibrium.jl:8 */
void orc_maxwellian_2v(const double *u, const double *v, int n, double rho, double U, double V,
                       double lam, double *M);                                               /* :20 */
void orc_maxwellian_3v(const double *u, const double *v, const double *w, int n, double rho,
                       double U, double V, double W, double lam, double *M);                 /* :32-33 */

typedef struct {
    double Mu[7], MuL[7], MuR[7], Mv[7], Mw[7], Mxi[3];
} orc_gm;
void orc_gauss_moments(const double *prim, int nprim, double K, orc_gm *g); /* theory_moments.jl:10-86 */
void orc_moments_conserve_1d(const double *Mu, const double *Mxi, int alpha, int delta, double *out); /* :98-110 */
void orc_moments_conserve_2d(const double *Mu, const double *Mv, const double *Mxi,
                             int alpha, int beta, int delta, double *out);                 /* :115-145 (internal) */
void orc_moments_conserve_3d(const double *Mu, const double *Mv, const double *Mw,
                             int alpha, int beta, int delta, double *out);                 /* :115-145 (3V) */
void orc_moments_conserve_slope_1d(const double *a, const double *Mu, const double *Mxi, int alpha,
                                   double *out);                                            /* :335-339 */
void orc_moments_conserve_slope_2d(const double *a, const double *Mu, const double *Mv, const double *Mxi,
                                   int alpha, int beta, double *out);                       /* :344-359 */
void orc_pdf_slope(const double *prim, const double *sw, int n, double K, double *out);   /* theory_pdf.jl:11-49 */

/* ---- fluxes (src/Flux) ---- */
void orc_flux_kfvs_1f1v(double *fw, double *ff, const double *fL, const double *fR, const double *u,
                        const double *w, int n, double dt, const double *sfL, const double *sfR);
void orc_flux_kfvs_2f1v(double *fw, double *fh, double *fb, const double *hL, const double *bL,
                        const double *hR, const double *bR, const double *u, const double *w, int n,
                        double dt, const double *shL, const double *sbL, const double *shR,
                        const double *sbR);
void orc_flux_kfvs_1f3v(double *fw, double *ff, const double *fL, const double *fR, const double *u,
                        const double *v, const double *wv, const double *w, int n, double dt,
                        double len, const double *sfL, const double *sfR);
void orc_flux_kfvs_2f2v(double *fw, double *fh, double *fb, const double *hL, const double *bL,
                        const double *hR, const double *bR, const double *u, const double *v,
                        const double *w, int n, double dt, double len, const double *shL,
                        const double *sbL, const double *shR, const double *sbR);
void orc_flux_kcu_1f1v(double *fw, double *ff, const double *wL, const double *fL, const double *wR,
                       const double *fR, const double *u, const double *w, int n, double K,
                       double gam, double mu, double om, double Pr, double dt);
void orc_flux_kcu_2f1v(double *fw, double *fh, double *fb, const double *wL, const double *hL,
                       const double *bL, const double *wR, const double *hR, const double *bR,
                       const double *u, const double *w, int n, double K, double gam, double mu,
                       double om, double Pr, double dt);
void orc_flux_kcu_1f3v(double *fw, double *ff, const double *wL, const double *fL, const double *wR,
                       const double *fR, const double *u, const double *v, const double *wv,
                       const double *w, int n, double K, double gam, double mu, double om,
                       double Pr, double dt, double len);
void orc_flux_kcu_2f2v(double *fw, double *fh, double *fb, const double *wL, const double *hL,
                       const double *bL, const double *wR, const double *hR, const double *bR,
                       const double *u, const double *v, const double *w, int n, double K,
                       double gam, double mu, double om, double Pr, double dt, double len);
void orc_flux_ugks_2f1v(double *fw, double *fh, double *fb, const double *wL, const double *hL,
                        const double *bL, const double *wR, const double *hR, const double *bR,
                        const double *u, const double *w, int n, double K, double gam, double mu,
                        double om, double Pr, double dt, double dxL, double dxR, const double *shL,
                        const double *sbL, const double *shR, const double *sbR);
void orc_flux_ugks_2f2v(double *fw, double *fh, double *fb, const double *wL, const double *hL,
                        const double *bL, const double *wR, const double *hR, const double *bR,
                        const double *u, const double *v, const double *w, int n, double K,
                        double gam, double mu, double om, double Pr, double dt, double dxL,
                        double dxR, double len, const double *shL, const double *sbL,
                        const double *shR, const double *sbR);
void orc_flux_boundary_maxwell_2f1v(double *fw, double *fh, double *fb, const double *bc,
                                    const double *h, const double *b, const double *u,
                                    const double *w, int n, double K, double dt, int rot);
void orc_flux_boundary_maxwell_2f2v(double *fw, double *fh, double *fb, const double *bc,
                                    const double *h, const double *b, const double *u,
                                    const double *v, const double *w, int n, double K, double dt,
                                    double len, int rot);
void orc_local_frame(const double *w, int n, double c, double s, double *out);  /* geo_general.jl:130-150 */
void orc_global_frame(const double *w, int n, double c, double s, double *out); /* geo_general.jl:6-26 */

/* ---- per-cell steps (src/Solver/step_1f.jl, step_2f.jl) ---- */
void orc_step_1f1v(double *w, double *prim, double *f, const double *fwL, const double *ffL,
                   const double *fwR, const double *ffR, const double *u, const double *weights,
                   int n, double gam, double mu, double om, double Pr, double dx, double dt,
                   double *RES, double *AVG, int coll);
void orc_step_2f1v(double *w, double *prim, double *h, double *b, const double *fwL,
                   const double *fhL, const double *fbL, const double *fwR, const double *fhR,
                   const double *fbR, const double *u, const double *weights, int n, double K,
                   double gam, double mu, double om, double Pr, double dx, double dt, double *RES,
                   double *AVG, int coll);
int orc_step_1f3v(double *w, double *prim, double *f, const double *fwL, const double *ffL,
                  const double *fwR, const double *ffR, const double *u, const double *v,
                  const double *wv, const double *weights, int n, double gam, double mu,
                  double om, double Pr, double dx, double dt, double *RES, double *AVG, int coll);
void orc_step_2f2v(double *w, double *prim, double *h, double *b, const double *fwL,
                   const double *fhL, const double *fbL, const double *fwR, const double *fhR,
                   const double *fbR, const double *fwD, const double *fhD, const double *fbD,
                   const double *fwU, const double *fhU, const double *fbU, const double *u,
                   const double *v, const double *weights, int n, double K, double gam, double mu,
                   double om, double Pr, double ds, double dt, double *RES, double *AVG, int coll);

/* ---- whole solves: 1D (1F1V, 2F1V, 1F3V) ----
 * Cell storage index k = i - 1 + ng for the physical 1-based index i in [1-ng, nx+ng].
 * Face storage index i - 1 for the physical face i in [1, nx+1] (between cells i-1 and i).
 * f is [cell][nf][nv]; w and prim are [cell][nw]; sw [cell][nw]; sf [cell][nf][nv];
 * fw [face][nw]; ff [face][nf][nv]. */
typedef struct {
    int nx, ng, nf, vdim, nv;
    int flux, coll, limiter, interp_order;
    int bc[2];
    double K, gam, mu, om, Pr;
    const double *u, *v, *wv, *weights;
    const double *dx;               /* nx + 2 ng */
    const double *bc_prim[2];       /* wall primitive vectors (nw values) for Maxwell walls */
    double *w, *prim, *f, *sw, *sf; /* cell state */
    double *fw, *ff;                /* face state */
    /* collision :fsm (1D1F3V): Gas.fsm = (Kn, nm, phi, psi, phipsi) and the velocity grid's shape */
    int fsm_nm, fsm_nu, fsm_nv, fsm_nw;
    double fsm_kn;
    const double *fsm_phi, *fsm_psi, *fsm_phipsi;
} orc_solver1d;

int orc_reconstruct_1d(orc_solver1d *S);                    /* solver_reconstruction.jl:56-153 */
int orc_evolve_1d(orc_solver1d *S, double dt);              /* solver_evolution.jl:39-89, evolve_boundary.jl:4-47 */
int orc_update_1d(orc_solver1d *S, double dt, double *res); /* solver_update.jl:61-93, update_boundary.jl:6-61 */
int orc_run_1d(orc_solver1d *S, double dt, int nsteps, double *res);

/* ---- 2D (2F2V) ----
 * Cells (i, j), 1-based, storage i-1 + nx*(j-1) (no ghost layers).
 * x-faces (i, j), i in 1..nx+1: storage i-1 + (nx+1)*(j-1); y-faces (i, j), j in 1..ny+1:
 * storage i-1 + nx*(j-1).  Each face carries its normal (2), its length and, on the walls, the
 * global-frame wall primitive vector (4).  Slopes: sw[cell][dir][4], sf[cell][dir][2][nv]. */
typedef struct {
    int nx, ny, nv;
    int flux, coll, limiter, interp_order;
    int bc[4];                       /* left, right, bottom(down), top(up) */
    double K, gam, mu, om, Pr;
    const double *u, *v, *weights;   /* nv */
    const double *dx, *dy;           /* nx*ny */
    const double *xn, *xlen, *xwall; /* (nx+1)*ny: n[2], len, wall prim[4] */
    const double *yn, *ylen, *ywall; /* nx*(ny+1) */
    double *w, *prim, *f, *sw, *sf;
    double *xfw, *xff, *yfw, *yff;
    int nf;                          /* distributions: 1 (2D1F2V, f) or 2 (2D2F2V, h and b) */
} orc_solver2d;

/* 2D1F2V kernels */
void orc_flux_kfvs_1f2v(double *fw, double *ff, const double *fL, const double *fR, const double *u,
                        const double *v, const double *w, int n, double dt, double len,
                        const double *sfL, const double *sfR);
void orc_flux_kcu_1f2v(double *fw, double *ff, const double *wL, const double *fL, const double *wR,
                       const double *fR, const double *u, const double *v, const double *w, int n,
                       double K, double gam, double mu, double om, double Pr, double dt, double len);
void orc_flux_boundary_maxwell_1f2v(double *fw, double *fh, const double *bc, const double *h,
                                    const double *u, const double *v, const double *w, int n,
                                    double dt, double len, int rot);
int orc_step_1f2v(double *w, double *prim, double *f, const double *fwL, const double *ffL,
                  const double *fwR, const double *ffR, const double *fwD, const double *ffD,
                  const double *fwU, const double *ffU, const double *u, const double *v,
                  const double *weights, int n, double gam, double mu, double om, double Pr,
                  double ds, double dt, double *RES, double *AVG, int coll);

int orc_reconstruct_2d(orc_solver2d *S);                    /* solver_reconstruction.jl:475-620 */
int orc_evolve_2d(orc_solver2d *S, double dt);              /* solver_evolution.jl:608-679, evolve_boundary.jl:52-178 */
int orc_update_2d(orc_solver2d *S, double dt, double *res); /* solver_update.jl:193-246, update_boundary.jl:172-306 */
int orc_run_2d(orc_solver2d *S, double dt, int nsteps, double *res);

/* ---- unstructured triangle meshes (2D1F2V / 2D2F2V, KFVS) ----
 * evolve! solver_evolution.jl:872-1035, update! solver_update.jl:359-471, triangle step!
 * step_1f.jl:228-285 / step_2f.jl:256-326, cellType-3 cells update_boundary.jl:520-590.
 * Face f: cells fc[2f], fc[2f+1] (-1 outside the mesh), unit normal fn[2f..] from the first cell to
 * the second (into the mesh on a boundary face), length flen[f], kind fkind[f]: 0 interior KFVS
 * (first order: no reconstruct! for these cells), 1 Maxwell wall (the mesh side's cell has
 * cellType 2; wall prim fwall[4f..] in the global frame), 2 a stored flux that stays.
 * Cell c: faces cf[3c..], dirc[3c..] = sign(cell's outward normal . face normal), area, KitBase
 * cellType ctype[c] (0 and 2 stepped, 3 the mean of its two neighbours nb[3c..], others fixed). */
typedef struct {
    int ncell, nface, nv, nf, coll;
    double K, gam, mu, om, Pr;
    const double *u, *v, *weights;
    const int *fc, *fkind, *cf, *ctype, *nb;
    const double *fn, *flen, *fwall, *dirc, *area;
    double *w, *prim, *f;   /* ncell x 4, ncell x 4, ncell x nf x nv */
    double *fw, *ff;        /* nface x 4, nface x nf x nv */
} orc_solverus;
int orc_evolve_us(orc_solverus *S, double dt);
int orc_update_us(orc_solverus *S, double dt, double *res);
int orc_run_us(orc_solverus *S, double dt, int nsteps, double *res);

/* ---- fast spectral method (1D1F3V :fsm) ---- */
void orc_lgwt(int N, double a, double b, double *x, double *w);             /* math_general.jl:58-98 */
double orc_hs_boltz_kn(double mu_ref, double alpha);                        /* theory_fsm.jl:16-17 */
void orc_kernel_mode(int M, double umax, double vmax, double wmax, double du, double dv, double dw, int nu, int nv,
                     int nw, double alpha, int quad_num, double *phi, double *psi, double *phipsi); /* theory_fsm.jl:24-90 */
void orc_dft3(double *re, double *im, int nu, int nv, int nw, int sign);    /* sign -1 fft, +1 bfft (unnormalised) */
void orc_fftshift3(const double *in, double *out, int nu, int nv, int nw);
void orc_boltzmann_fft(double *Q, const double *f, double Kn, int M, const double *phi, const double *psi,
                       const double *phipsi, int nu, int nv, int nw);       /* theory_fsm.jl:271-314 */
int orc_step_1f3v_fsm(double *w, double *prim, double *f, const double *fwL, const double *ffL, const double *fwR,
                      const double *ffR, double gam, double Kn, int M, const double *phi, const double *psi,
                      const double *phipsi, int nu, int nv, int nw, double dx, double dt, double *RES,
                      double *AVG);                                          /* step_1f.jl:122-157 */

int orc_num_threads(void);
void orc_set_num_threads(int n);

#ifdef __cplusplus
}
#endif
#endif
