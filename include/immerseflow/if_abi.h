/*
 * ImmerseFlow-AMD C ABI -- the drop-in boundary for ImmerseFlow++'s
 * fractional-step time step on MI355X (gfx950).
 *
 * The reference (sushrutkr/ImmerseFlow) exposes its solver as the member
 * functions of `struct ImmerseFlow` (src/header/globalVariables.cuh:70-88)
 * called in a fixed order by main() (src/main.cu:75-102), plus its files:
 * inputs.txt, xgrid/ygrid face files and Tecplot .dat outputs.  Each entry
 * point below names the reference interface it replaces.
 *
 * Conventions
 *  - plain C types, caller-owned host buffers, library-owned device memory;
 *  - every call returns an if_status (0 = success); the library never calls
 *    exit() (the reference does, globalVariables.cuh:91-107); the message of
 *    the last failure on the calling thread is if_last_error();
 *  - a handle is not re-entrant: calls on one handle must not overlap; other
 *    handles may be driven from other threads;
 *  - cell arrays use the reference layout: nx*ny values with nx, ny counting
 *    the ghost ring and idx = i + j*nx (src/include/preSim.cu:10); face
 *    arrays: uf (nx-1)*(ny-2), vf (nx-2)*(ny-1) (ADSolver.cu:66-74).
 */
#ifndef IMMERSEFLOW_IF_ABI_H
#define IMMERSEFLOW_IF_ABI_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef int if_status;

enum {
    IF_OK = 0,
    IF_ERR_ARG = -1,      /* bad argument or descriptor */
    IF_ERR_IO = -2,       /* file missing / unreadable / unwritable */
    IF_ERR_PARSE = -3,    /* malformed deck or grid file */
    IF_ERR_HIP = -4,      /* HIP runtime error */
    IF_ERR_NOMEM = -5,    /* host or device allocation failed */
    IF_ERR_STATE = -6,    /* call not valid for this handle's mode/state */
    IF_ERR_NODEV = -7,    /* no usable gfx950 device */
    IF_ERR_BUDGET = -8,   /* batch admission refused: HBM budget */
    IF_ERR_DIVERGED = -9  /* non-finite residual: the case is marked failed */
};

enum { IF_MODE_PHYSICS = 0, IF_MODE_COMPAT = 1 };
/* Precision.  IF_FP64: every pass in FP64, the oracle's expressions.
 * IF_MIXED: the Poisson error sweeps in FP32 (the FP64 residual and defect
 * correction around them), everything else FP64.  IF_MIXED_V32: as
 * IF_MIXED, and the velocity stencils (convective source, viscous sweeps,
 * projection) with u, v stored in FP32 as well; faster on the velocity
 * passes, but the pressure then answers the FP32 rounding of the predicted
 * velocity (dp ~ eps32 L / dt), about 1e-4 relative on the workloads here
 * against under 1e-5 for IF_MIXED. */
enum { IF_FP64 = 0, IF_MIXED = 1, IF_MIXED_V32 = 2 };
enum { IF_PPE_JACOBI = 0, IF_PPE_RBSOR = 1, IF_PPE_LINE = 2 };
enum { IF_BODY_CIRCLE = 0, IF_BODY_ELLIPSE = 1, IF_BODY_POLYGON = 2 };
enum { IF_MOTION_NONE = 0, IF_MOTION_HEAVE = 1 };

enum {
    IF_FIELD_U = 0, IF_FIELD_V = 1, IF_FIELD_P = 2, IF_FIELD_UF = 3,
    IF_FIELD_VF = 4, IF_FIELD_IBLANK = 5, IF_FIELD_FLAGS = 6
};

#define IF_MAX_BODIES 16

/* One immersed body.  Geometry is plain data: orientation enters as
 * cos/sin computed by the caller, polygons as vertex arrays. */
typedef struct {
    int kind;               /* IF_BODY_* */
    double xc, yc;          /* centre / polygon reference point */
    double r;               /* circle radius; polygon bounding radius */
    double a, b;            /* ellipse semi-axes */
    double cth, sth;        /* ellipse rotation cos/sin */
    int nvert;              /* polygon vertices */
    const double *vx, *vy;  /* polygon vertices relative to (xc, yc) */
    int motion;             /* IF_MOTION_* */
    double amp, freq;       /* heave y = yc + amp*sin(2 pi freq t) */
} if_body;

/* The deck of src/main.cu:10-59 (CFDInput, globalVariables.cuh:15-33), nx,
 * ny including the ghost ring after main.cu:55-58, plus the physics-mode
 * extension read from keyword lines the reference parser ignores. */
typedef struct {
    int restart, restart_time;
    int nx, ny, nxf, nyf;
    double lx, ly;
    int w_ad, w_ppe, ad_itermax, ppe_itermax, ad_solver, ppe_solver;
    double error_max, tmax, dt, re, mu;
    int write_interval;
    /* extension ("Fractional step" block) */
    int mode;               /* IF_MODE_*; default compat */
    int precision;          /* IF_FP64 / IF_MIXED / IF_MIXED_V32 */
    double conv_factor;     /* default 0.5 (ADSolver.cu:66-74) */
    double ppe_omega;       /* SOR relaxation; <= 0: from the grid */
    double ppe_tol;         /* dt*||b - A p||_inf stop level */
    int check_every;        /* sweeps per Poisson check */
    double ad_omega;        /* real-valued w_AD (optional 7th value of the
                               "Mode" line); 0 when absent: the CLI then
                               uses the deck's integer w_ad */
    int nbody;
    if_body bodies[IF_MAX_BODIES];
    double poly_store[IF_MAX_BODIES][2][512]; /* storage for parsed polygons */
} if_inputs;

/* A case: grid + parameters + bodies (ImmerseFlow's allocation +
 * readGridData + initializeData, preSim.cu:147-240,242-372). */
typedef struct {
    int mode;               /* IF_MODE_* */
    int precision;          /* IF_FP64 / IF_MIXED / IF_MIXED_V32 (physics mode) */
    int nxf, nyf;           /* face counts (NX+1, NY+1) */
    const double *xf, *yf;  /* face coordinates (grid-file contents) */
    double dt, re;
    double conv_factor;     /* ADSource factor, 0.5 in the reference */
    double ubc, vbc;        /* Dirichlet velocity on the four sides */
    int ad_itermax;
    double ad_tol;          /* L1 sum of |dq| (ADSolver.cu:315), 1e-6 */
    int ppe_solver;         /* IF_PPE_JACOBI / IF_PPE_RBSOR / IF_PPE_LINE (zebra line SOR; rows of at most
                               8064 interior cells, IF_ERR_ARG beyond) */
    double ppe_omega;
    int ppe_itermax;
    double ppe_tol;
    int check_every;
    double t0;              /* start time, for moving bodies */
    int nbody;
    const if_body *bodies;
    double ad_omega;        /* w_AD: relaxation of the viscous Jacobi sweep,
                               q' = (1 - w) q + w q_J, as the reference writes
                               it and comments it out (ADSolver.cu:95,116);
                               0 means 1 (no relaxation) */
} if_case_desc;

typedef struct {
    int steps;              /* steps taken by this call */
    int ad_iters;           /* last step: Jacobi sweeps executed */
    double ad_res;          /* last step: final uRes+vRes */
    int ppe_sweeps;         /* last step: Poisson sweeps executed */
    double ppe_res;         /* last step: dt*||b - A p||_inf */
    double div_max;         /* last step: max |D uf| over fluid cells */
    long long total_ad_iters, total_ppe_sweeps; /* over the call */
    double max_div;         /* max over the call's steps */
    int failed;             /* non-finite state detected */
    int persistent_timeouts; /* steps whose persistent Poisson chunk timed out
                                (blocks not all resident) and was solved again
                                with one launch per stage */
    int weight_table;       /* last step: the red-black sweeps looked each
                               cell's weight up by its face flags, because
                               every red-black case resident has a grid with
                               one coefficient value per axis in the Poisson
                               precision; mixed: instead of loading a per-cell
                               weight, with the uniform grid's five-operation
                               update; FP64: instead of dividing on every
                               launch.  Results are the same either way. */
} if_step_stats;

typedef struct {
    int32_t gidx;           /* ghost cell, reference linear index */
    int32_t body;
    int32_t node[4];        /* bilinear stencil cells, reference indices */
    double wd[4], cd;       /* Dirichlet weights, body-velocity coefficient */
    double wn[4];           /* Neumann weights */
    double lnx, lny;        /* IP - GC (moving-body Neumann term) */
    double bix, biy, ipx, ipy;
} if_ghost;

/* Result of the launch-shape autotuner (if_case_autotune). */
typedef struct {
    int kernel, tile_y, sweeps_per_launch;  /* the fastest Poisson sweep shape */
    int check_every;        /* chosen check interval in sweeps */
    double sweep_ns;        /* its time per interior cell-sweep (all running cases) */
    double check_sweeps;    /* one convergence check, in sweeps of that shape */
    int typical_sweeps;     /* sweeps per solve the check interval was chosen for (0: none yet) */
    int candidates;         /* shapes timed */
} if_tune_result;

typedef struct if_case if_case;
typedef struct if_batch if_batch;

/* ---- device -------------------------------------------------------------
 * Replaces ImmerseFlow::CUDAQuery (preSim.cu:187-199). */
if_status if_device_query(int *ndev, char *name, size_t name_len,
                          int *cus, size_t *hbm_bytes);

/* ---- files ---------------------------------------------------------------
 * if_parse_inputs replaces readInputFile (main.cu:10-59), same grammar. */
if_status if_parse_inputs(const char *path, if_inputs *out);
/* if_read_grid replaces the file part of readGridData (preSim.cu:267-291):
 * nf "<id> <x>" lines; xf must hold nf values. */
if_status if_read_grid(const char *path, int nf, double *xf);
/* if_grid_metric replaces the metric part of readGridData
 * (preSim.cu:293-355): n = nf+1 centres (ghost mirrors) and widths. */
if_status if_grid_metric(int nf, const double *xf, double *xc, double *dx);
/* if_write_tecplot replaces write_results_to_file (postSim.cu:41-66). */
if_status if_write_tecplot(const char *path, int nx, int ny, const double *x,
                           const double *y, const double *f);
/* The writer's format at `digits` significant digits (1..17; 0 = the
 * reference's "%f").  17 digits are a lossless FP64 checkpoint.  No
 * counterpart in the reference (its writer is fixed at "%f"). */
if_status if_write_tecplot_digits(const char *path, int nx, int ny, const double *x,
                                  const double *y, const double *f, int digits);
/* Read a file in the writer's format (any digit count): the ZONE line's I/J
 * must equal nx/ny; x (nx) and y (ny) may be NULL.  Restart source, as the
 * predecessor's data files are (test/UTIL_PRE_SIM.f90:130-149). */
if_status if_read_tecplot(const char *path, int nx, int ny, double *x, double *y,
                          double *f);

/* ---- single case ---------------------------------------------------------
 * if_case_create replaces allocation + readGridData + initializeData. */
if_status if_case_create(const if_case_desc *desc, if_case **out);
/* if_case_destroy replaces freeAllocation (preSim.cu:164-179). */
void if_case_destroy(if_case *c);
/* Compat mode: one ImmerseFlow::ADsolver call (ADSolver.cu:268-395). */
if_status if_case_ad_step(if_case *c, if_step_stats *st);
/* Compat mode: ImmerseFlow::PPESolver (PPESolver.cu:137-205). */
if_status if_case_ppe_solve(if_case *c, if_step_stats *st);
/* Physics mode: nsteps full fractional steps (predictor, divergence RHS,
 * Poisson, projection, divergence check). */
if_status if_case_step(if_case *c, int nsteps, if_step_stats *st);
/* Copy a field to the host (replaces saveDataToFile's D2H copy,
 * postSim.cu:10-39); n must equal the field's reference-layout size. */
if_status if_case_get_field(if_case *c, int which, double *out, size_t n);
/* Load a field (restart); same layouts as if_case_get_field. */
if_status if_case_set_field(if_case *c, int which, const double *in, size_t n);
/* Restart (the deck's Restart flag, main.cu:27-30; predecessor
 * test/UTIL_PRE_SIM.f90:130-158): load u, v, p (nx*ny, ghosts included)
 * and the face velocities uf ((nx-1)*(ny-2)) and vf ((nx-2)*(ny-1)).  With
 * uf and vf NULL the faces are interpolated from u and v with the blanked
 * faces at the body velocity, as at initialisation and as the predecessor
 * does.  Ghost-cell velocities are then refreshed.  The case's time is the
 * t0 it was created with (create it with t0 = the restart time). */
if_status if_case_restart(if_case *c, const double *u, const double *v, const double *p,
                          const double *uf, const double *vf);
/* Grid vectors of the case (xc, yc, dx, dy: nx and ny values). */
if_status if_case_grid(if_case *c, double *xc, double *yc, double *dx,
                       double *dy);
if_status if_case_dims(if_case *c, int *nx, int *ny);
/* Ghost-cell list (sorted reference linear indices) and image-point
 * records.  *n gets the count; out/recs may be NULL to query it. */
if_status if_case_ghosts(if_case *c, int32_t *out, if_ghost *recs,
                         size_t cap, size_t *n);
/* Poisson sweep launch shape.  kernel 0: LDS-tiled red-black SOR, tile
 * height 8/16/32 cells; kernel 1: register-streaming red-black SOR (one
 * wave64 per column strip), strip height 16/32/64/128 rows; kernel 2:
 * register-tile red-black SOR (four wave64s hold a 128-column tile in
 * registers), tile height 16/32/48 rows, 1, 2 or 4 sweeps per launch;
 * kernel 3: the same with eight wave64s per tile, tile height 24/40/56
 * (2 sweeps per launch), 32/48/64 (4) or 32/48/64 (8); kernel 4: sixteen
 * wave64s per tile, tile height 48/64/80 (4 sweeps per launch) or 32/64/96
 * (8), and in mixed precision only also 96 (8), 92 (9), 88 (10), 84 (11)
 * and 80 (12; the tile computes 128 x 128 cells, stores 128 - 4m of them in
 * each direction and recomputes the rest).  Several red-black sweeps per launch is
 * temporal blocking; check_every is the number of sweeps between
 * convergence checks.  Results do not depend on the kernel, the tile
 * height or the sweeps per launch.  Defaults: kernel 4, 64 rows, 8 sweeps
 * (mixed); FP64 (until tuned): kernel 4, 4 sweeps as persistent chunks of
 * the smallest tile height of 48, 64 and 80 rows whose tiles fit on the
 * device in one launch, else kernel 3, 32 rows, 4 sweeps. */
if_status if_case_tune(if_case *c, int kernel, int tile_y, int sweeps_per_launch,
                       int check_every);
/* Launch mode of the Poisson sweeps (no counterpart in the reference,
 * whose PPESolver launches each Jacobi sweep from the host,
 * PPESolver.cu:172-189).  enable = 1 (the default): each chunk of
 * check_every sweeps is captured once into a hipGraph of kernel nodes and
 * replayed, recaptured when the resident set or the launch shape changes;
 * 0: the same kernels are launched one by one.  Results are identical. */
if_status if_case_graphs(if_case *c, int enable);
/* Persistent Poisson chunks (no counterpart in the reference).  enable = 1
 * (the default): with a register-tile shape that has a persistent form
 * (mixed: 16 waves x 96/64 rows x 8 sweeps, 16 x 48 x 4, 8 x 32 x 8 and
 * 8 x 32 x 4; FP64: 16 x 80 x 4 and 16 x 48 x 4), when every block of the
 * launch can be resident on the device at once, each chunk of check_every
 * sweeps runs as one launch whose tiles stay in registers, in stages of
 * sweeps_per_launch sweeps separated by a halo exchange with the
 * neighbouring tiles; otherwise one launch per stage.  Results are
 * identical.  A neighbour wait that times out (20 ms; the blocks were not
 * all resident) ends that solve early with a finite pressure; the step then
 * turns the form off, solves again from that pressure with one launch per
 * stage, completes as usual (converged to the same stop level) and counts
 * it in if_step_stats.persistent_timeouts. */
if_status if_case_persistent(if_case *c, int enable);
/* Monitors (the predecessor's probes, test/UTIL_PRE_SIM.f90:172-202, and a
 * body-force diagnostic with no counterpart there): after every step the
 * device records t, the force (Fx, Fy) the fluid exerts on the bodies and
 * u, v, p at each probe, into a ring of `capacity` records.  A probe is the
 * interior cell whose centre is nearest (x, y) per axis, as the
 * predecessor's minloc (:189-194).  The force sums, over the faces between
 * a fluid and a body cell, the pressure times the face length and the wall
 * shear mu (u_P - u_body) 2/d times the face length (mu = 1/Re; d the
 * cell width normal to the face).  nprobe <= 32; xy holds x0 y0 x1 y1 ...
 * Calling it again replaces the probes and empties the ring. */
if_status if_case_monitor(if_case *c, int nprobe, const double *xy, int capacity);
/* The records since the last read, oldest first, each 3 + 3*nprobe doubles
 * (t, Fx, Fy, then u, v, p per probe); at most cap_records are copied (the
 * newest), *dropped counts older ones lost to the ring or to cap_records.
 * out may be NULL to query *nrec only (the ring is then kept). */
if_status if_case_monitor_read(if_case *c, double *out, size_t cap_records, size_t *nrec,
                               size_t *dropped);
/* One monitor record of the current fields at the case's current time,
 * without a step: the initial state, the state after a restart, or fields
 * loaded with if_case_set_field (the predecessor writes its probes from
 * the current state too, test/UTIL_PRE_SIM.f90:172-202). */
if_status if_case_monitor_sample(if_case *c);
/* if_case_monitor_sample without any allocation: no counterpart in the
 * reference; for callers that cannot allocate (code under HIP graph
 * capture, tight serving loops, harnesses that forbid hipMalloc).  The same
 * kernels and the same record, with the descriptor copy, the tile table and
 * the per-tile force partials staged in a static device array of the
 * library (about 1.5 MiB per GPU, held until the process ends) instead of
 * memory allocated per call.  Cases of more than 65,536 interior tiles of
 * 64 x 16 cells (67 M cells) are refused with IF_ERR_ARG.  Calls that
 * overlap on one GPU share that array and race: serialise them. */
if_status if_case_monitor_sample_noalloc(if_case *c);
/* Autotuner (no counterpart in the reference, whose loop is fixed):
 * times every Poisson sweep launch shape of if_case_tune on the case's
 * current fields (a few launches each, on the solver's scratch buffers;
 * the solution and the loop state are left as they were; a shape that
 * would run as persistent chunks is timed as one) and keeps the
 * fastest.  The check interval is then chosen from the measured cost of a
 * convergence check c (in sweeps) and the sweep count S of the last solve:
 * checking every K sweeps costs about S c / K checks plus K / 2 sweeps of
 * overshoot, least at K = sqrt(2 S c) (rounded to whole launches, 4..1024).
 * Before the first step S is unknown and the interval is left unchanged.
 * Results do not depend on the shape; the interval changes only where the
 * solve stops (still at the same tolerance). */
if_status if_case_autotune(if_case *c, if_tune_result *out);
/* Accumulated event timing of the Poisson sweep kernel (milliseconds,
 * launches) over the steps since the last reset.  A persistent chunk
 * counts as one launch. */
if_status if_case_sweep_timing(if_case *c, int reset, double *ms,
                               long long *launches);
/* Algorithmic HBM bytes per interior cell of one Poisson sweep launch (13
 * or 25 bytes per sweep times the sweeps per launch of the last step:
 * sweeps_per_launch, or check_every for a persistent chunk) and of one
 * step at the given sweep counts (the roofline byte model). */
if_status if_case_byte_model(if_case *c, double *sweep_bytes_per_cell,
                             double *step_bytes_per_cell);

/* ---- batches and serving -------------------------------------------------
 * Many independent cases advanced together: every kernel launch covers the
 * tiles of all resident cases; per-case device flags end each case's
 * iterations independently; per-case reductions run in a fixed order, so a
 * case's result does not depend on what else is resident. */
if_status if_batch_create(int precision, size_t hbm_budget, if_batch **out);
void if_batch_destroy(if_batch *b);
/* Admit a case that should run nsteps steps; IF_ERR_BUDGET when its
 * footprint does not fit the remaining budget. *slot gets its id. */
if_status if_batch_add(if_batch *b, const if_case_desc *desc, int nsteps,
                       int *slot);
/* Advance every resident unfinished case by up to nsteps steps (one batched
 * fractional step per iteration).  *finished gets how many cases completed
 * their step count during the call. */
if_status if_batch_step(if_batch *b, int nsteps, int *finished);
/* State of a slot: 0 running, 1 finished, -1 failed, -2 empty. */
if_status if_batch_slot_state(if_batch *b, int slot, int *state,
                              int *steps_done, if_step_stats *st);
if_status if_batch_get_field(if_batch *b, int slot, int which, double *out,
                             size_t n);
/* Release a slot (frees its device memory back to the budget). */
if_status if_batch_remove(if_batch *b, int slot);
/* Resident cases, device bytes charged to the budget, and the event timing
 * of Poisson sweep launches over full convergence-check chunks:
 * milliseconds, launches, and interior cell-sweeps done in them (for the
 * achieved-bandwidth figure of batched launches). */
if_status if_batch_info(if_batch *b, int *resident, size_t *bytes_in_use,
                        double *sweep_ms, long long *sweep_launches,
                        double *cell_sweeps);
/* As if_case_tune, for every case of the batch. */
if_status if_batch_tune(if_batch *b, int kernel, int tile_y, int sweeps_per_launch,
                        int check_every);
/* As if_case_graphs, for the batch. */
if_status if_batch_graphs(if_batch *b, int enable);
/* As if_case_persistent, for the batch. */
if_status if_batch_persistent(if_batch *b, int enable);
/* As if_case_autotune, over every running case of the batch at once (the
 * batched launches are what is timed; S is the median of the cases' last
 * sweep counts). */
if_status if_batch_autotune(if_batch *b, if_tune_result *out);
/* Footprint in device bytes that if_batch_add would charge for desc. */
if_status if_case_footprint(const if_case_desc *desc, int precision,
                            size_t *bytes);

/* ---- errors ---------------------------------------------------------------*/
const char *if_last_error(void);

#ifdef __cplusplus
}
#endif
#endif
