/*
 * ImmerseFlow oracle -- TEST INFRASTRUCTURE ONLY.
 *
 * A plain-C, FP64, CPU restatement of the ImmerseFlow++ fractional-step time
 * step.  Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg
 * may load this library, and only as the checker.  The product
 * (immerseflow_amd/) never links or calls it.
 *
 * Two families of routines:
 *
 *  compat (of_compat_*): the reference kernels *as coded*, including their
 *    quirks (stale-ghost face interpolation, vf loop-continuation, ping-pong
 *    ghost parity, signed Poisson residual, test pressure BC).  Pinned by the
 *    reference's own output files results/uc.dat, vc.dat, p.dat and
 *    final_results.dat (copied under tests/golden/).
 *
 *  physics (of_phys_*): the reference discretisation completed into a full
 *    fractional step (divergence RHS, Poisson, projection, divergence check)
 *    with a sharp-interface ghost-cell immersed boundary.  The reference has
 *    no code for these parts (PPESolver.cu does not compile,
 *    AD_PPE_Correction.cu is empty, iBlankComputeKernel is a stub), so no
 *    reference output exists to compare them with; each such function is
 *    pinned by closed forms, invariants and brute force in tests/ instead
 *    (test_oracle_physics.py, test_oracle_moving.py).
 *
 * Layouts follow the reference (src/include/preSim.cu:10): cell arrays are
 * nx*ny with idx = i + j*nx, nx/ny INCLUDING the ghost ring; east-face
 * velocities uf are (nx-1)*(ny-2) with uf[i + (j-1)*(nx-1)] the east face of
 * cell (i,j); north-face velocities vf are (nx-2)*(ny-1) with
 * vf[(i-1) + j*(nx-2)] the north face of cell (i,j) (ADSolver.cu:66-74).
 */
#ifndef IMMERSEFLOW_ORACLE_H
#define IMMERSEFLOW_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- grid metric (preSim.cu:293-355) ------------------------------------ */
/* From nf face coordinates build n = nf+1 cell centres (with ghost mirrors)
 * and n cell widths (with ghost copies).  Status 0 on success. */
int of_grid_metric(int nf, const double *xf, double *xc, double *dx);

/* ---- compat mode: the reference as coded ------------------------------- */
/* Vortex initial condition (preSim.cu:53-97); uf/vf/p zeroed. */
void of_compat_init(int nx, int ny, const double *xc, const double *yc,
                    double *u, double *v, double *uf, double *vf, double *p);

/* One call of ImmerseFlow::ADsolver (ADSolver.cu:268-395).  u,v are the
 * solver-owned velc buffers (updated in place, following the pointer swaps),
 * uf/vf are persistent face arrays, iblank is the FP64 iBlank array.
 * Returns the number of Jacobi iterations executed; *res_out gets the last
 * uRes+vRes. */
int of_compat_ad_step(int nx, int ny, const double *dx, const double *dy,
                      double dt, double re, int ad_itermax,
                      const double *iblank, double *u, double *v,
                      double *uf, double *vf, double *res_out);

/* The same with the w_AD relaxation the reference writes and comments out
 * (ADSolver.cu:95,116): unew = (1.0 - w_AD) * u + w_AD * unew after each
 * Jacobi update.  w_ad = 1 is of_compat_ad_step. */
int of_compat_ad_step_w(int nx, int ny, const double *dx, const double *dy,
                        double dt, double re, int ad_itermax, double w_ad,
                        const double *iblank, double *u, double *v,
                        double *uf, double *vf, double *res_out);

/* ImmerseFlow::PPESolver as intended (PPESolver.cu:13-71,137-205):
 * set_pressure_BC, Jacobi sweeps until the signed residual sum <= 1e-6 or
 * itermax, final set_pressure_BC.  Returns sweeps executed. */
int of_compat_ppe(int nx, int ny, const double *dx, const double *dy,
                  int ppe_itermax, double *p, double *res_out);

/* ---- physics mode ------------------------------------------------------- */

/* Body kinds.  Geometry parameters are plain data computed by the caller
 * (no transcendental functions inside classification). */
enum { OF_BODY_CIRCLE = 0, OF_BODY_ELLIPSE = 1, OF_BODY_POLYGON = 2 };

typedef struct {
    int kind;
    double xc, yc;          /* centre (circle, ellipse); polygon: reference point */
    double r;               /* circle radius; polygon: bounding radius */
    double a, b;            /* ellipse semi-axes */
    double c, s;            /* ellipse cos/sin of the rotation angle */
    int nvert;              /* polygon vertex count */
    const double *vx, *vy;  /* polygon vertices (closed implicitly) */
    double ub, vb;          /* rigid-body velocity */
    double axb, ayb;        /* rigid-body acceleration (Neumann dp/dn term) */
} of_body;

/* Cell flag bits (one byte per cell). */
enum {
    OF_FLUID = 1,   /* iBlank == 1 */
    OF_GHOST = 2,   /* solid cell with >= 1 fluid 4-neighbour (interior only) */
    OF_OPEN_E = 4,  /* east face open: both cells fluid, not a domain face */
    OF_OPEN_W = 8,
    OF_OPEN_N = 16,
    OF_OPEN_S = 32
};

/* iBlank classification: flags[idx] gets OF_FLUID for fluid cells (all ghost
 * ring cells are fluid), then ghost and open-face bits.  body_id (may be NULL)
 * gets the index of the containing body or -1. */
void of_classify(int nx, int ny, const double *xc, const double *yc,
                 int nbody, const of_body *bodies, uint8_t *flags,
                 int8_t *body_id);

/* Ghost-cell list sorted by linear index; returns the count (<= cap). */
int of_ghost_list(int nx, int ny, const uint8_t *flags, int32_t *list,
                  int cap);

/* Per-ghost image-point record. */
typedef struct {
    int32_t gidx;        /* ghost cell linear index */
    int32_t body;        /* body index */
    int32_t node[4];     /* bilinear stencil cells (linear index) */
    double wd[4];        /* Dirichlet weights on fluid nodes (0 elsewhere) */
    double cd;           /* Dirichlet coefficient on the body velocity */
    double wn[4];        /* Neumann weights on fluid nodes (0 elsewhere) */
    double lnx, lny;     /* |IP-GC| * outward normal (Neumann accel term) */
    double bix, biy;     /* boundary intercept */
    double ipx, ipy;     /* image point */
} of_ghost;

/* Build image-point records for the ghost list. */
void of_ghost_weights(int nx, int ny, const double *xc, const double *yc,
                      const uint8_t *flags, int nbody, const of_body *bodies,
                      const int8_t *body_id, int ng, const int32_t *list,
                      of_ghost *out);

/* Closest point on a body boundary (the boundary intercept). */
void of_closest_point(const of_body *b, double px, double py, double *bx,
                      double *by);

/* Physics-mode parameters. */
typedef struct {
    double dt, re;
    double conv_factor;   /* 0.5 = the reference's ADSource factor */
    double ubc, vbc;      /* Dirichlet velocity on the four domain sides */
    int ad_itermax;
    double ad_tol;        /* L1 sum of |dq| (ADSolver.cu:315) */
    int ppe_solver;       /* 0 = Jacobi, 1 = red-black SOR, 2 = zebra line SOR */
    double ppe_omega;
    int ppe_itermax;
    double ppe_tol;       /* stop when dt*||b - A p||_inf <= ppe_tol */
    int check_every;      /* sweeps between Poisson residual checks */
    double ad_omega;      /* w_AD relaxation of the viscous sweep,
                             q' = (1 - w) q + w q_Jacobi (ADSolver.cu:95,116);
                             0 (unset) means w = 1, no relaxation */
} of_params;

typedef struct {
    int ad_iters;
    double ad_res;
    int ppe_sweeps;
    double ppe_res;       /* dt*||b - A p||_inf at exit */
    double div_max;       /* max |D uf| over fluid cells after projection */
} of_stats;

/* Initial state: u=ubc, v=vbc in fluid, body velocity in solid cells,
 * ghost values from the image points, domain ghosts by the BC, face
 * velocities interpolated (blanked faces = body velocity, domain faces =
 * BC), p = 0. */
void of_phys_init(int nx, int ny, const double *dx, const double *dy,
                  const uint8_t *flags, const int8_t *body_id, int nbody,
                  const of_body *bodies, int ng, const of_ghost *gh,
                  const of_params *prm, double *u, double *v, double *uf,
                  double *vf, double *p);

/* Restart from u, v, p (ghost ring included, e.g. read from the writer's
 * files): ghost-cell velocities refreshed from the image points, then the
 * faces interpolated from the cells (blanked faces at the body velocity,
 * domain faces at the BC), as at initialisation; the predecessor's restart
 * rebuilds its faces from u, v in the same way (test/UTIL_PRE_SIM.f90:130-164). */
void of_phys_restart(int nx, int ny, const double *dx, const double *dy,
                     const uint8_t *flags, const int8_t *body_id, int nbody,
                     const of_body *bodies, int ng, const of_ghost *gh,
                     const of_params *prm, double *u, double *v, double *uf,
                     double *vf, double *p);

/* One fractional step. */
void of_phys_step(int nx, int ny, const double *dx, const double *dy,
                  const uint8_t *flags, const int8_t *body_id, int nbody,
                  const of_body *bodies, int ng, const of_ghost *gh,
                  const of_params *prm, double *u, double *v, double *uf,
                  double *vf, double *p, of_stats *st);

/* Moving bodies: before a step the caller places every body at t_{n+1}
 * (position, velocity, acceleration as data).  This reclassifies all cells
 * (flags, body ids) and gives each fresh cell -- solid before, fluid now --
 * the velocity of the body that left it.  The caller then rebuilds the
 * ghost list and image-point records.  Returns the number of fresh cells. */
int of_phys_move_bodies(int nx, int ny, const double *xc, const double *yc,
                        int nbody, const of_body *bodies, uint8_t *flags,
                        int8_t *body_id, double *u, double *v);

/* Pieces of the step, exposed for closed-form tests. */
void of_phys_rhs(int nx, int ny, const double *dx, const double *dy,
                 const uint8_t *flags, double dt, const double *uf,
                 const double *vf, double *rhs);
double of_phys_residual(int nx, int ny, const double *dx, const double *dy,
                        const uint8_t *flags, const double *rhs,
                        const double *p, double *r /* may be NULL */);
int of_phys_poisson(int nx, int ny, const double *dx, const double *dy,
                    const uint8_t *flags, const double *rhs,
                    const of_params *prm, double *p, double *res_out);
double of_phys_divergence(int nx, int ny, const double *dx, const double *dy,
                          const uint8_t *flags, const double *uf,
                          const double *vf, double *div /* may be NULL */);
void of_phys_ghost_velocity(int nx, int ny, int ng, const of_ghost *gh,
                            int nbody, const of_body *bodies, double *u,
                            double *v);
void of_phys_ghost_pressure(int nx, int ny, int ng, const of_ghost *gh,
                            int nbody, const of_body *bodies, double *p);

#ifdef __cplusplus
}
#endif
#endif
