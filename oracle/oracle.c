/*
 * ImmerseFlow oracle -- TEST INFRASTRUCTURE ONLY (see oracle.h).
 *
 * Plain C, FP64, OpenMP over rows.  Built with -ffp-contract=off so every
 * expression rounds exactly as written; the product's HIP kernels evaluate
 * the same expressions in the same order, which is what lets the parity
 * tests demand bit equality where no reduction order is involved.
 *
 * Reductions that feed control flow (the AD L1 residual) are summed row by
 * row in a fixed order so the oracle is deterministic at any thread count.
 */
#include "oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

#define IDX(i, j, nx) ((i) + (j) * (nx))

/* ======================================================================== */
/* grid metric: ImmerseFlow::readGridData, src/include/preSim.cu:293-355    */
/* ======================================================================== */
int of_grid_metric(int nf, const double *xf, double *xc, double *dx)
{
    const int n = nf + 1; /* nx = NX + 2 with nf = NX + 1 (main.cu:55-58) */
    if (nf < 2) return -1;
    /* centres, preSim.cu:294-296 */
    for (int i = 1; i < n - 1; i++) xc[i] = (xf[i - 1] + xf[i]) / 2.0;
    /* ghost mirrors, preSim.cu:304-307: the west mirror is about x = 0 */
    xc[0] = -1 * xc[1];
    xc[n - 1] = xf[nf - 1] + (xf[nf - 1] - xc[n - 2]);
    /* widths, preSim.cu:329-334, and ghost copies, :337-355 */
    for (int i = 1; i < n - 1; i++) dx[i] = xf[i] - xf[i - 1];
    dx[0] = dx[1];
    dx[n - 1] = dx[n - 2];
    return 0;
}

/* ======================================================================== */
/* compat mode                                                              */
/* ======================================================================== */

/* initializeKernel, preSim.cu:53-97 (vortex IC; fields zeroed first). */
void of_compat_init(int nx, int ny, const double *xc, const double *yc,
                    double *u, double *v, double *uf, double *vf, double *p)
{
    for (int j = 0; j < ny; j++)
        for (int i = 0; i < nx; i++) {
            const int id = IDX(i, j, nx);
            double r = sqrt(pow(xc[i] - 0.5, 2.0) + pow(yc[j] - 0.5, 2.0));
            double r0 = 0.1;
            u[id] = 1.0 - 0.25 * (yc[j] - 0.5) * exp((1.0 - pow(r / r0, 2.0)) / 2.0);
            v[id] = 0.25 * (xc[i] - 0.5) * exp((1.0 - pow(r / r0, 2.0)) / 2.0);
            p[id] = 0.0;
        }
    memset(uf, 0, sizeof(double) * (size_t)(nx - 1) * (ny - 2));
    memset(vf, 0, sizeof(double) * (size_t)(nx - 2) * (ny - 1));
}

/* Per-axis viscous coefficients of calculateADCoefficients,
 * ADSolver.cu:34-39, evaluated with the reference's operation order:
 *   coeff   = (1.0 + (dt/Re)*(X)) + (dt/Re)*(Y)
 *   dx2_p1  = (dt/Re) * (2 / (dx_i * (dx_i + dx_ip1)))  etc. */
typedef struct {
    double *xp, *xm, *yp, *ym; /* off-diagonals */
    double *px, *py;            /* aP = px[i] + py[j] */
} ad_coef;

static void ad_coef_build(int nx, int ny, const double *dx, const double *dy,
                          double dt, double re, ad_coef *c)
{
    const double k = dt / re;
    c->xp = calloc(nx, sizeof(double));
    c->xm = calloc(nx, sizeof(double));
    c->px = calloc(nx, sizeof(double));
    c->yp = calloc(ny, sizeof(double));
    c->ym = calloc(ny, sizeof(double));
    c->py = calloc(ny, sizeof(double));
    for (int i = 1; i < nx - 1; i++) {
        const double dxi = dx[i], dxp = dx[i + 1], dxm = dx[i - 1];
        c->xp[i] = k * (2 / (dxi * (dxi + dxp)));
        c->xm[i] = k * (2 / (dxi * (dxi + dxm)));
        c->px[i] = 1.0 + k * ((2 / (dxi * (dxi + dxp))) + (2 / (dxi * (dxi + dxm))));
    }
    for (int j = 1; j < ny - 1; j++) {
        const double dyj = dy[j], dyp = dy[j + 1], dym = dy[j - 1];
        c->yp[j] = k * (2 / (dyj * (dyj + dyp)));
        c->ym[j] = k * (2 / (dyj * (dyj + dym)));
        c->py[j] = k * ((2 / (dyj * (dyj + dyp))) + (2 / (dyj * (dyj + dym))));
    }
}

static void ad_coef_free(ad_coef *c)
{
    free(c->xp); free(c->xm); free(c->px);
    free(c->yp); free(c->ym); free(c->py);
}

/* Compute_velf, ADSolver.cu:162-189, including the loop-continuation quirk:
 * with the reference launch (one pass of the grid-stride loop per thread)
 * the vf loop only runs for ids in [(nx-1)(ny-2), (nx-2)(ny-1)). */
static void compat_velf(int nx, int ny, const double *dx, const double *dy,
                        const double *u, const double *v, double *uf,
                        double *vf)
{
    const int nuf = (nx - 1) * (ny - 2);
    const int nvf = (nx - 2) * (ny - 1);
    for (int id = 0; id < nuf; id++) {
        const int i = id % (nx - 1), j = id / (nx - 1);
        const int idc = j * nx + i;
        /* dx is j-independent (preSim.cu:331): dx[idc+nx] = dx_i */
        uf[id] = 1.0 / (dx[i] + dx[i + 1]) * (u[idc + nx + 1] * dx[i] + u[idc + nx] * dx[i + 1]);
    }
    for (int id = nuf; id < nvf; id++) {
        const int i = id % (nx - 2), j = id / (nx - 2);
        const int idc = j * nx + i;
        vf[id] = 1.0 / (dy[j + 1] + dy[j]) * (v[idc + nx + 1] * dy[j] + v[idc + 1] * dy[j + 1]);
    }
}

/* set_velocity_BC, ADSolver.cu:191-217.  The reference applies the four
 * side rules in place from all threads at once, so the corner cells read
 * neighbouring ghosts that other threads are writing.  The oracle fixes one
 * outcome: the order of a sequential sweep over id (i fastest), side rules in
 * source order.  Corner (0,0) then reads the pre-call value of ghost (0,1);
 * the other three corners read freshly written edge ghosts.  Edge cells read
 * interior cells only, which the kernel does not write.  The uf/vf loops of
 * the kernel (:221-263) never run (loop-continuation quirk) and are omitted. */
static void compat_bc(int nx, int ny, double *q, double qbc)
{
    for (int j = 0; j < ny; j++)
        for (int i = 0; i < nx; i++) {
            const int id = IDX(i, j, nx);
            if (i == 0) q[id] = -q[id + 1] + qbc * 2.0;
            if (j == 0) q[id] = -q[id + nx] + qbc * 2.0;
            if (i == nx - 1) q[id] = -q[id - 1] + qbc * 2.0;
            if (j == ny - 1) q[id] = -q[id - nx] + qbc * 2.0;
        }
}

/* ADSource, ADSolver.cu:46-79 (factor c; the reference uses 0.5). */
static void source_cells(int nx, int ny, const double *dx, const double *dy,
                         double dt, double c, const double *q,
                         const double *uf, const double *vf, double *s)
{
#pragma omp parallel for schedule(static)
    for (int j = 0; j < ny; j++)
        for (int i = 0; i < nx; i++) {
            const int id = IDX(i, j, nx);
            s[id] = 0.0;
            if (i > 0 && i < nx - 1 && j > 0 && j < ny - 1) {
                const double dx_i = dx[i], dx_ip1 = dx[i + 1], dx_im1 = dx[i - 1];
                const double dy_j = dy[j], dy_jp1 = dy[j + 1], dy_jm1 = dy[j - 1];
                const double ufe = uf[i + (j - 1) * (nx - 1)];
                const double ufw = uf[i - 1 + (j - 1) * (nx - 1)];
                const double vfn = vf[i - 1 + j * (nx - 2)];
                const double vfs = vf[i - 1 + (j - 1) * (nx - 2)];
                s[id] = (q[id] - (dt / dx_i) * c * (((1 / (dx_i + dx_ip1)) * (q[id + 1] * dx_i + q[id] * dx_ip1) * ufe)
                                                  - ((1 / (dx_i + dx_im1)) * (q[id] * dx_im1 + q[id - 1] * dx_i) * ufw)))
                        - (dt / dy_j) * c * (((1 / (dy_j + dy_jp1)) * (q[id + nx] * dy_j + q[id] * dy_jp1) * vfn)
                                            - ((1 / (dy_j + dy_jm1)) * (q[id] * dy_jm1 + q[id - nx] * dy_j) * vfs));
            }
        }
}

/* ADusolver_kernel / ADvsolver_kernel, ADSolver.cu:81-120: one Jacobi sweep
 * on interior cells, masked by iBlank; ghost cells of qn are not written.
 * w != 1: the relaxation the reference comments out at :95,:116,
 * qn = (1.0 - w) * q + w * qn, applied after the masked update. */
static void ad_sweep(int nx, int ny, const ad_coef *c, const double *iblank,
                     const double *s, const double *q, double *qn, double w)
{
#pragma omp parallel for schedule(static)
    for (int j = 1; j < ny - 1; j++)
        for (int i = 1; i < nx - 1; i++) {
            const int id = IDX(i, j, nx);
            double t = s[id] + c->xp[i] * q[id + 1] + c->xm[i] * q[id - 1]
                             + c->yp[j] * q[id + nx] + c->ym[j] * q[id - nx];
            double r = iblank[id] * t / (c->px[i] + c->py[j]);
            if (w != 1.0) r = (1.0 - w) * q[id] + w * r;
            qn[id] = r;
        }
}

/* Compute_{u,v}Residual_AD + Reduction: sum of |qn - q| over interior cells
 * with iBlank == 1 (ADSolver.cu:122-160, preSim.cu:376-445).  Row partials
 * summed in row order. */
static double l1_change(int nx, int ny, const double *iblank, const double *a,
                        const double *b)
{
    double *row = calloc(ny, sizeof(double));
#pragma omp parallel for schedule(static)
    for (int j = 1; j < ny - 1; j++) {
        double acc = 0.0;
        for (int i = 1; i < nx - 1; i++) {
            const int id = IDX(i, j, nx);
            if (iblank[id] == 1) acc += fabs(a[id] - b[id]);
        }
        row[j] = acc;
    }
    double s = 0.0;
    for (int j = 1; j < ny - 1; j++) s += row[j];
    free(row);
    return s;
}

/* ImmerseFlow::Reduction, preSim.cu:376-445: two launches of reduce6<256>
 * (:22-50), the first with CUDAQuery's blocksPerGrid = ceil(n/256) blocks
 * (:190-193) over the n values, the second with one block over the block
 * sums.  The summation order is restated exactly: thread tid of block b
 * starts at i = b*512 + tid and adds g[i] + g[i+256] (or g[i] alone when
 * i+256 >= n) per grid-stride step; the block then folds its 256 sums
 * pairwise, s[t] += s[t+h] for h = 128, 64, ..., 1 (warpReduce's lockstep
 * reads give the same s[0]). */
static double reduce6_block(const double *g, unsigned n, unsigned b, unsigned grid_size)
{
    double sd[256];
    for (unsigned tid = 0; tid < 256; tid++) {
        double acc = 0;
        for (unsigned i = b * 512 + tid; i < n; i += grid_size) {
            if (i + 256 < n) acc += g[i] + g[i + 256];
            else acc += g[i];
        }
        sd[tid] = acc;
    }
    for (unsigned h = 128; h >= 1; h >>= 1)
        for (unsigned tid = 0; tid < h; tid++) sd[tid] += sd[tid + h];
    return sd[0];
}

static double ref_reduction(const double *g, unsigned n)
{
    const unsigned nb = (n + 255) / 256;
    double *part = malloc(sizeof(double) * nb);
    for (unsigned b = 0; b < nb; b++) part[b] = reduce6_block(g, n, b, 512 * nb);
    const double s = reduce6_block(part, nb, 0, 512);
    free(part);
    return s;
}

/* Compute_{u,v}Residual_AD (ADSolver.cu:122-160): |a - b| on interior cells
 * with iBlank == 1, 0 elsewhere, summed by Reduction. */
static double compat_l1(int nx, int ny, const double *iblank, const double *a,
                        const double *b, double *tmp)
{
    for (int j = 0; j < ny; j++)
        for (int i = 0; i < nx; i++) {
            const int id = IDX(i, j, nx);
            tmp[id] = (i > 0 && i < nx - 1 && j > 0 && j < ny - 1 && iblank[id] == 1)
                          ? fabs(a[id] - b[id]) : 0.0;
        }
    return ref_reduction(tmp, (unsigned)(nx * ny));
}

/* ImmerseFlow::ADsolver, ADSolver.cu:268-395, as coded. */
int of_compat_ad_step(int nx, int ny, const double *dx, const double *dy,
                      double dt, double re, int ad_itermax,
                      const double *iblank, double *u, double *v,
                      double *uf, double *vf, double *res_out)
{
    return of_compat_ad_step_w(nx, ny, dx, dy, dt, re, ad_itermax, 1.0, iblank, u, v, uf, vf, res_out);
}

/* The same with the w_AD relaxation of ADSolver.cu:95,116. */
int of_compat_ad_step_w(int nx, int ny, const double *dx, const double *dy,
                        double dt, double re, int ad_itermax, double w_ad,
                        const double *iblank, double *u, double *v,
                        double *uf, double *vf, double *res_out)
{
    const size_t n = (size_t)nx * ny;
    ad_coef c;
    ad_coef_build(nx, ny, dx, dy, dt, re, &c); /* :289 */
    /* velc buffers and the per-call scratch (:275-287; fresh scratch is
     * modelled as zeros -- its ghosts are overwritten by the second
     * iteration's BC before they can reach the output) */
    double *bu[2] = {malloc(n * sizeof(double)), calloc(n, sizeof(double))};
    double *bv[2] = {malloc(n * sizeof(double)), calloc(n, sizeof(double))};
    double *sx = calloc(n, sizeof(double)), *sy = calloc(n, sizeof(double));
    double *rtmp = malloc(n * sizeof(double));
    memcpy(bu[0], u, n * sizeof(double));
    memcpy(bv[0], v, n * sizeof(double));
    int cur = 0;

    compat_velf(nx, ny, dx, dy, bu[cur], bv[cur], uf, vf);   /* :298 stale ghosts */
    compat_bc(nx, ny, bu[cur], 1.0);                          /* :304 */
    compat_bc(nx, ny, bv[cur], 0.0);
    source_cells(nx, ny, dx, dy, dt, 0.5, bu[cur], uf, vf, sx); /* :308 */
    source_cells(nx, ny, dx, dy, dt, 0.5, bv[cur], uf, vf, sy);

    double ures = 1.0, vres = 1.0;
    int iter = 0;
    while (ures + vres > pow(10.0, -6.0) && iter < ad_itermax) { /* :315 */
        compat_bc(nx, ny, bu[cur], 1.0);                          /* :317 */
        compat_bc(nx, ny, bv[cur], 0.0);
        compat_velf(nx, ny, dx, dy, bu[cur], bv[cur], uf, vf);   /* :321 */
        ad_sweep(nx, ny, &c, iblank, sx, bu[cur], bu[1 - cur], w_ad); /* :326 */
        ad_sweep(nx, ny, &c, iblank, sy, bv[cur], bv[1 - cur], w_ad); /* :331 */
        compat_bc(nx, ny, bu[cur], 1.0);                          /* :337 old buffer */
        compat_bc(nx, ny, bv[cur], 0.0);
        cur = 1 - cur;                                            /* :342-348 swap */
        ures = compat_l1(nx, ny, iblank, bu[cur], bu[1 - cur], rtmp); /* :350-366 */
        vres = compat_l1(nx, ny, iblank, bv[cur], bv[1 - cur], rtmp);
        iter += 1;
    }
    memcpy(u, bu[cur], n * sizeof(double));
    memcpy(v, bv[cur], n * sizeof(double));
    if (res_out) *res_out = ures + vres;
    free(bu[0]); free(bu[1]); free(bv[0]); free(bv[1]);
    free(sx); free(sy); free(rtmp);
    ad_coef_free(&c);
    return iter;
}

/* Poisson coefficients as intended by calculatePPECoefficients,
 * PPESolver.cu:93-99: a_E = 2/(dx_i(dx_i+dx_ip1)) ..., a_P = -(sum). */
static void ppe_axis(int n, const double *d, double *cp, double *cm)
{
    for (int i = 1; i < n - 1; i++) {
        cp[i] = (2 / (d[i] * (d[i] + d[i + 1])));
        cm[i] = (2 / (d[i] * (d[i] + d[i - 1])));
    }
    cp[0] = cm[0] = cp[n - 1] = cm[n - 1] = 0.0;
}

/* set_pressure_BC, PPESolver.cu:54-71: p = 100 on the i==0 and j==0 ghosts. */
static void compat_pbc(int nx, int ny, double *p)
{
    for (int j = 0; j < ny; j++)
        for (int i = 0; i < nx; i++) {
            const int id = IDX(i, j, nx);
            if (i == 0) p[id] = 100.0;
            if (j == 0) p[id] = 100.0;
        }
}

/* ImmerseFlow::PPESolver loop, PPESolver.cu:13-50,164-195. */
int of_compat_ppe(int nx, int ny, const double *dx, const double *dy,
                  int ppe_itermax, double *p, double *res_out)
{
    const size_t n = (size_t)nx * ny;
    double *cxp = calloc(nx, sizeof(double)), *cxm = calloc(nx, sizeof(double));
    double *cyp = calloc(ny, sizeof(double)), *cym = calloc(ny, sizeof(double));
    ppe_axis(nx, dx, cxp, cxm);
    ppe_axis(ny, dy, cyp, cym);
    double *pt = calloc(n, sizeof(double)); /* managed memory reads as zero */
    double *res_cell = calloc(n, sizeof(double));
    double *pc = p;
    compat_pbc(nx, ny, pc);
    double residual = 1.0;
    int iter = 0;
    while (residual > pow(10.0, -6.0) && iter < ppe_itermax) {
        /* jacobiIteration, :13-31 */
#pragma omp parallel for schedule(static)
        for (int j = 0; j < ny; j++)
            for (int i = 0; i < nx; i++) {
                const int id = IDX(i, j, nx);
                pt[id] = pc[id];
                if (i > 0 && i < nx - 1 && j > 0 && j < ny - 1) {
                    const double ap = -1 * ((cxp[i] + cxm[i]) + (cyp[j] + cym[j]));
                    pt[id] = -(pc[id + 1] * cxp[i] + pc[id - 1] * cxm[i]
                               + pc[id + nx] * cyp[j] + pc[id - nx] * cym[j]) / ap;
                }
            }
        double *t = pc; pc = pt; pt = t; /* :178-180 */
        /* Compute_Residual (signed A.p, no RHS; the boundary entries of the
         * managed array stay zero), :33-50, summed by Reduction, :182 */
#pragma omp parallel for schedule(static)
        for (int j = 0; j < ny; j++)
            for (int i = 0; i < nx; i++) {
                const int id = IDX(i, j, nx);
                double r = 0.0;
                if (i > 0 && i < nx - 1 && j > 0 && j < ny - 1) {
                    const double ap = -1 * ((cxp[i] + cxm[i]) + (cyp[j] + cym[j]));
                    r = pc[id] * ap + pc[id + 1] * cxp[i] + pc[id - 1] * cxm[i]
                        + pc[id + nx] * cyp[j] + pc[id - nx] * cym[j];
                }
                res_cell[id] = r;
            }
        residual = ref_reduction(res_cell, (unsigned)n);
        iter += 1;
    }
    compat_pbc(nx, ny, pc); /* :195 */
    if (pc != p) memcpy(p, pc, n * sizeof(double));
    if (res_out) *res_out = residual;
    free(pc == p ? pt : pc);
    free(cxp); free(cxm); free(cyp); free(cym); free(res_cell);
    return iter;
}

/* ======================================================================== */
/* physics mode: geometry                                                   */
/* ======================================================================== */

/* Inside tests: + - * / and compares only, FP64, no contraction. */
static int inside_body(const of_body *b, double x, double y)
{
    const double ex = x - b->xc, ey = y - b->yc;
    switch (b->kind) {
    case OF_BODY_CIRCLE:
        return ex * ex + ey * ey <= b->r * b->r;
    case OF_BODY_ELLIPSE: {
        const double X = ex * b->c + ey * b->s;
        const double Y = ey * b->c - ex * b->s;
        return (X * X) / (b->a * b->a) + (Y * Y) / (b->b * b->b) <= 1.0;
    }
    case OF_BODY_POLYGON: {
        if (!(ex * ex + ey * ey <= b->r * b->r)) return 0; /* bounding circle */
        int in = 0;
        const int n = b->nvert;
        for (int k = 0; k < n; k++) {
            const int k1 = (k + 1 == n) ? 0 : k + 1;
            const double ax = b->vx[k], ay = b->vy[k];
            const double bx = b->vx[k1], by = b->vy[k1];
            if ((ay > y) != (by > y)) {
                const double cr = (bx - ax) * (y - ay) - (by - ay) * (x - ax);
                if (by > ay) { if (cr > 0) in = !in; }
                else         { if (cr < 0) in = !in; }
            }
        }
        return in;
    }
    }
    return 0;
}

void of_classify(int nx, int ny, const double *xc, const double *yc,
                 int nbody, const of_body *bodies, uint8_t *flags,
                 int8_t *body_id)
{
    /* iBlank: 1 = fluid.  The reference's iBlankComputeKernel
     * (preSim.cu:110-136) is a stub writing 1.0; its commented-out circle
     * test (:126-130) is the intended classification. */
#pragma omp parallel for schedule(static)
    for (int j = 0; j < ny; j++)
        for (int i = 0; i < nx; i++) {
            int owner = -1;
            for (int b = 0; b < nbody && owner < 0; b++)
                if (inside_body(&bodies[b], xc[i], yc[j])) owner = b;
            flags[IDX(i, j, nx)] = (owner < 0) ? OF_FLUID : 0;
            if (body_id) body_id[IDX(i, j, nx)] = (int8_t)owner;
        }
    /* ghost cells and open faces on interior cells */
#pragma omp parallel for schedule(static)
    for (int j = 1; j < ny - 1; j++)
        for (int i = 1; i < nx - 1; i++) {
            const int id = IDX(i, j, nx);
            uint8_t f = flags[id] & OF_FLUID;
            const int fe = flags[id + 1] & OF_FLUID, fw = flags[id - 1] & OF_FLUID;
            const int fn = flags[id + nx] & OF_FLUID, fs = flags[id - nx] & OF_FLUID;
            if (f) {
                if (i + 1 <= nx - 2 && fe) f |= OF_OPEN_E;
                if (i - 1 >= 1 && fw) f |= OF_OPEN_W;
                if (j + 1 <= ny - 2 && fn) f |= OF_OPEN_N;
                if (j - 1 >= 1 && fs) f |= OF_OPEN_S;
            } else if (fe || fw || fn || fs) {
                f |= OF_GHOST;
            }
            flags[id] = f;
        }
}

int of_phys_move_bodies(int nx, int ny, const double *xc, const double *yc,
                        int nbody, const of_body *bodies, uint8_t *flags,
                        int8_t *body_id, double *u, double *v)
{
    const size_t n = (size_t)nx * ny;
    uint8_t *of = malloc(n);
    int8_t *ob = malloc(n);
    memcpy(of, flags, n);
    memcpy(ob, body_id, n);
    of_classify(nx, ny, xc, yc, nbody, bodies, flags, body_id);
    int fresh = 0;
    for (size_t id = 0; id < n; id++)
        if (!(of[id] & OF_FLUID) && (flags[id] & OF_FLUID) && ob[id] >= 0) {
            u[id] = bodies[ob[id]].ub;
            v[id] = bodies[ob[id]].vb;
            fresh++;
        }
    free(of);
    free(ob);
    return fresh;
}

int of_ghost_list(int nx, int ny, const uint8_t *flags, int32_t *list, int cap)
{
    int n = 0;
    for (int id = 0; id < nx * ny; id++)
        if (flags[id] & OF_GHOST) {
            if (n < cap) list[n] = id;
            n++;
        }
    return n;
}

/* Distance from a point to an ellipse: D. Eberly, "Distance from a Point to
 * an Ellipse, an Ellipsoid, or a Hyperellipsoid" (Geometric Tools, 2013),
 * robust bisection on F(s) = (r0 z0/(s+r0))^2 + (z1/(s+1))^2 - 1.  First
 * quadrant, e0 >= e1 > 0, y0, y1 >= 0. */
static void ellipse_closest_q1(double e0, double e1, double y0, double y1,
                               double *x0, double *x1)
{
    if (y1 > 0) {
        if (y0 > 0) {
            const double z0 = y0 / e0, z1 = y1 / e1;
            double g = z0 * z0 + z1 * z1 - 1.0;
            if (g != 0) {
                const double r0 = (e0 / e1) * (e0 / e1);
                const double n0 = r0 * z0;
                double s0 = z1 - 1.0;
                double s1 = (g < 0) ? 0.0 : sqrt(n0 * n0 + z1 * z1) - 1.0;
                double s = 0.0;
                for (int it = 0; it < 200; it++) {
                    s = (s0 + s1) / 2.0;
                    if (s == s0 || s == s1) break;
                    const double q0 = n0 / (s + r0), q1 = z1 / (s + 1.0);
                    g = q0 * q0 + q1 * q1 - 1.0;
                    if (g > 0) s0 = s;
                    else if (g < 0) s1 = s;
                    else break;
                }
                *x0 = r0 * y0 / (s + r0);
                *x1 = y1 / (s + 1.0);
            } else {
                *x0 = y0;
                *x1 = y1;
            }
        } else {
            *x0 = 0.0;
            *x1 = e1;
        }
    } else {
        const double numer0 = e0 * y0, denom0 = e0 * e0 - e1 * e1;
        if (numer0 < denom0) {
            const double xde0 = numer0 / denom0;
            *x0 = e0 * xde0;
            *x1 = e1 * sqrt(1.0 - xde0 * xde0);
        } else {
            *x0 = e0;
            *x1 = 0.0;
        }
    }
}

void of_closest_point(const of_body *b, double px, double py, double *bx,
                      double *by)
{
    const double ex = px - b->xc, ey = py - b->yc;
    switch (b->kind) {
    case OF_BODY_CIRCLE: {
        const double L = sqrt(ex * ex + ey * ey);
        if (L == 0.0) { *bx = b->xc + b->r; *by = b->yc; return; }
        *bx = b->xc + b->r * (ex / L);
        *by = b->yc + b->r * (ey / L);
        return;
    }
    case OF_BODY_ELLIPSE: {
        /* body frame */
        const double X = ex * b->c + ey * b->s;
        const double Y = ey * b->c - ex * b->s;
        const int swap = b->a < b->b;
        const double e0 = swap ? b->b : b->a, e1 = swap ? b->a : b->b;
        const double y0 = fabs(swap ? Y : X), y1 = fabs(swap ? X : Y);
        double x0, x1;
        ellipse_closest_q1(e0, e1, y0, y1, &x0, &x1);
        double cx = swap ? x1 : x0, cy = swap ? x0 : x1;
        if (X < 0) cx = -cx;
        if (Y < 0) cy = -cy;
        /* back to the world frame */
        *bx = b->xc + (cx * b->c - cy * b->s);
        *by = b->yc + (cx * b->s + cy * b->c);
        return;
    }
    case OF_BODY_POLYGON: {
        double best = -1.0, bxx = b->vx[0], byy = b->vy[0];
        const int n = b->nvert;
        for (int k = 0; k < n; k++) {
            const int k1 = (k + 1 == n) ? 0 : k + 1;
            const double ax = b->vx[k], ay = b->vy[k];
            const double sx = b->vx[k1] - ax, sy = b->vy[k1] - ay;
            const double L2 = sx * sx + sy * sy;
            double t = (L2 > 0.0) ? ((px - ax) * sx + (py - ay) * sy) / L2 : 0.0;
            if (t < 0.0) t = 0.0;
            if (t > 1.0) t = 1.0;
            const double cx = ax + t * sx, cy = ay + t * sy;
            const double d2 = (px - cx) * (px - cx) + (py - cy) * (py - cy);
            if (best < 0.0 || d2 < best) { best = d2; bxx = cx; byy = cy; }
        }
        *bx = bxx;
        *by = byy;
        return;
    }
    }
    *bx = px;
    *by = py;
}

/* Largest k in [0, n-2] with c[k] <= x (c strictly increasing). */
static int bracket(int n, const double *c, double x)
{
    int lo = 0, hi = n - 2;
    if (x <= c[0]) return 0;
    if (x >= c[n - 2]) return n - 2;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (c[mid] <= x) lo = mid; else hi = mid - 1;
    }
    return lo;
}

/* Image-point records (sharp-interface ghost-cell method, Mittal et al.,
 * J. Comput. Phys. 227 (2008) 4825-4852): boundary intercept BI = closest
 * boundary point, image point IP = 2 BI - GC, bilinear stencil around IP.
 * Stencil nodes that are fluid carry weights; the ghost cell itself is
 * eliminated implicitly; other solid nodes are taken at the body value
 * (Dirichlet) or dropped with renormalisation (Neumann). */
void of_ghost_weights(int nx, int ny, const double *xc, const double *yc,
                      const uint8_t *flags, int nbody, const of_body *bodies,
                      const int8_t *body_id, int ng, const int32_t *list,
                      of_ghost *out)
{
    (void)nbody;
#pragma omp parallel for schedule(static)
    for (int g = 0; g < ng; g++) {
        const int id = list[g];
        const int gi = id % nx, gj = id / nx;
        const int b = body_id[id] < 0 ? 0 : body_id[id];
        const double gx = xc[gi], gy = yc[gj];
        double bix, biy;
        of_closest_point(&bodies[b], gx, gy, &bix, &biy);
        const double ipx = 2.0 * bix - gx, ipy = 2.0 * biy - gy;
        const int i0 = bracket(nx, xc, ipx), j0 = bracket(ny, yc, ipy);
        const double fx = (ipx - xc[i0]) / (xc[i0 + 1] - xc[i0]);
        const double fy = (ipy - yc[j0]) / (yc[j0 + 1] - yc[j0]);
        const double w[4] = {(1.0 - fx) * (1.0 - fy), fx * (1.0 - fy),
                             (1.0 - fx) * fy, fx * fy};
        const int nd[4] = {IDX(i0, j0, nx), IDX(i0 + 1, j0, nx),
                           IDX(i0, j0 + 1, nx), IDX(i0 + 1, j0 + 1, nx)};
        double wself = 0.0, wsolid = 0.0, wfluid = 0.0;
        for (int k = 0; k < 4; k++) {
            if (nd[k] == id) wself += w[k];
            else if (flags[nd[k]] & OF_FLUID) wfluid += w[k];
            else wsolid += w[k];
        }
        of_ghost *o = &out[g];
        o->gidx = id;
        o->body = b;
        o->cd = (2.0 - wsolid) / (1.0 + wself);
        for (int k = 0; k < 4; k++) {
            const int fl = (nd[k] != id) && (flags[nd[k]] & OF_FLUID);
            o->node[k] = fl ? nd[k] : id;
            o->wd[k] = fl ? -w[k] / (1.0 + wself) : 0.0;
            o->wn[k] = (fl && wfluid > 0.0) ? w[k] / wfluid : 0.0;
        }
        o->lnx = ipx - gx;
        o->lny = ipy - gy;
        o->bix = bix; o->biy = biy;
        o->ipx = ipx; o->ipy = ipy;
    }
}

void of_phys_ghost_velocity(int nx, int ny, int ng, const of_ghost *gh,
                            int nbody, const of_body *bodies, double *u,
                            double *v)
{
    (void)nx; (void)ny; (void)nbody;
    for (int g = 0; g < ng; g++) {
        const of_ghost *o = &gh[g];
        const of_body *b = &bodies[o->body];
        double su = o->cd * b->ub, sv = o->cd * b->vb;
        for (int k = 0; k < 4; k++) {
            su = su + o->wd[k] * u[o->node[k]];
            sv = sv + o->wd[k] * v[o->node[k]];
        }
        u[o->gidx] = su;
        v[o->gidx] = sv;
    }
}

void of_phys_ghost_pressure(int nx, int ny, int ng, const of_ghost *gh,
                            int nbody, const of_body *bodies, double *p)
{
    (void)nx; (void)ny; (void)nbody;
    for (int g = 0; g < ng; g++) {
        const of_ghost *o = &gh[g];
        const of_body *b = &bodies[o->body];
        double sp = o->wn[0] * p[o->node[0]];
        for (int k = 1; k < 4; k++) sp = sp + o->wn[k] * p[o->node[k]];
        p[o->gidx] = sp + (o->lnx * b->axb + o->lny * b->ayb);
    }
}

/* ======================================================================== */
/* physics mode: the fractional step                                        */
/* ======================================================================== */

/* Domain ghosts: q_g = -q_in + 2 q_bc (ADSolver.cu:199-217) on the four
 * edges; corners are never written. */
static void domain_bc_vel(int nx, int ny, double *q, double qbc)
{
    for (int j = 1; j < ny - 1; j++) {
        q[IDX(0, j, nx)] = -q[IDX(1, j, nx)] + qbc * 2.0;
        q[IDX(nx - 1, j, nx)] = -q[IDX(nx - 2, j, nx)] + qbc * 2.0;
    }
    for (int i = 1; i < nx - 1; i++) {
        q[IDX(i, 0, nx)] = -q[IDX(i, 1, nx)] + qbc * 2.0;
        q[IDX(i, ny - 1, nx)] = -q[IDX(i, ny - 2, nx)] + qbc * 2.0;
    }
}

/* Homogeneous Neumann pressure ghosts: p_g = p_in. */
static void domain_bc_p(int nx, int ny, double *p)
{
    for (int j = 1; j < ny - 1; j++) {
        p[IDX(0, j, nx)] = p[IDX(1, j, nx)];
        p[IDX(nx - 1, j, nx)] = p[IDX(nx - 2, j, nx)];
    }
    for (int i = 1; i < nx - 1; i++) {
        p[IDX(i, 0, nx)] = p[IDX(i, 1, nx)];
        p[IDX(i, ny - 1, nx)] = p[IDX(i, ny - 2, nx)];
    }
}

/* Velocity on a blanked interior face (at least one side solid): the
 * velocity of the body owning the first solid side (classification ids). */
static double blank_vel(const int8_t *body_id, int ida, int idb,
                        const of_body *bodies, int comp)
{
    int b = body_id[ida] >= 0 ? body_id[ida] : body_id[idb];
    if (b < 0) return 0.0;
    return comp == 0 ? bodies[b].ub : bodies[b].vb;
}

/* Face velocities from cell velocities: linear interpolation of
 * Compute_velf (ADSolver.cu:173-174,184-185) on open faces, body velocity
 * on blanked faces, the Dirichlet BC on domain faces. */
static void faces_from_cells(int nx, int ny, const double *dx, const double *dy,
                             const uint8_t *flags, const int8_t *body_id,
                             const of_body *bodies, const of_params *prm,
                             const double *u, const double *v, double *uf,
                             double *vf)
{
#pragma omp parallel for schedule(static)
    for (int j = 1; j < ny - 1; j++)
        for (int i = 0; i < nx - 1; i++) {
            const int id = IDX(i, j, nx);
            double val;
            if (i == 0 || i == nx - 2) val = prm->ubc;
            else if (flags[id] & OF_OPEN_E)
                val = (1.0 / (dx[i] + dx[i + 1])) * (u[id + 1] * dx[i] + u[id] * dx[i + 1]);
            else val = blank_vel(body_id, id, id + 1, bodies, 0);
            uf[i + (j - 1) * (nx - 1)] = val;
        }
#pragma omp parallel for schedule(static)
    for (int j = 0; j < ny - 1; j++)
        for (int i = 1; i < nx - 1; i++) {
            const int id = IDX(i, j, nx);
            double val;
            if (j == 0 || j == ny - 2) val = prm->vbc;
            else if (flags[id] & OF_OPEN_N)
                val = (1.0 / (dy[j] + dy[j + 1])) * (v[id + nx] * dy[j] + v[id] * dy[j + 1]);
            else val = blank_vel(body_id, id, id + nx, bodies, 1);
            vf[(i - 1) + j * (nx - 2)] = val;
        }
}

/* Discrete divergence of the face field at interior cell (i,j). */
static inline double cell_div(int nx, int i, int j, const double *dx,
                              const double *dy, const double *uf,
                              const double *vf)
{
    const double ufe = uf[i + (j - 1) * (nx - 1)], ufw = uf[i - 1 + (j - 1) * (nx - 1)];
    const double vfn = vf[i - 1 + j * (nx - 2)], vfs = vf[i - 1 + (j - 1) * (nx - 2)];
    return (ufe - ufw) * (1.0 / dx[i]) + (vfn - vfs) * (1.0 / dy[j]);
}

void of_phys_rhs(int nx, int ny, const double *dx, const double *dy,
                 const uint8_t *flags, double dt, const double *uf,
                 const double *vf, double *rhs)
{
    const double rdt = 1.0 / dt;
#pragma omp parallel for schedule(static)
    for (int j = 0; j < ny; j++)
        for (int i = 0; i < nx; i++) {
            const int id = IDX(i, j, nx);
            rhs[id] = 0.0;
            if (i > 0 && i < nx - 1 && j > 0 && j < ny - 1 && (flags[id] & OF_FLUID))
                rhs[id] = cell_div(nx, i, j, dx, dy, uf, vf) * rdt;
        }
}

double of_phys_divergence(int nx, int ny, const double *dx, const double *dy,
                          const uint8_t *flags, const double *uf,
                          const double *vf, double *div)
{
    double *row = calloc(ny, sizeof(double));
#pragma omp parallel for schedule(static)
    for (int j = 1; j < ny - 1; j++) {
        double m = 0.0;
        for (int i = 1; i < nx - 1; i++) {
            const int id = IDX(i, j, nx);
            double d = 0.0;
            if (flags[id] & OF_FLUID) d = cell_div(nx, i, j, dx, dy, uf, vf);
            if (div) div[id] = d;
            if (!(fabs(d) <= 1.0e300)) m = INFINITY;
            else if (fabs(d) > m) m = fabs(d);
        }
        row[j] = m;
    }
    double m = 0.0;
    for (int j = 1; j < ny - 1; j++) if (row[j] > m) m = row[j];
    free(row);
    return m;
}

/* Poisson operator at fluid cell (i,j): coefficients of PPESolver.cu:93-99
 * on open faces, zero across blanked and domain faces (Neumann). */
typedef struct { double *xp, *xm, *yp, *ym; } ppe_coef;

static void ppe_coef_build(int nx, int ny, const double *dx, const double *dy,
                           ppe_coef *c)
{
    c->xp = calloc(nx, sizeof(double)); c->xm = calloc(nx, sizeof(double));
    c->yp = calloc(ny, sizeof(double)); c->ym = calloc(ny, sizeof(double));
    ppe_axis(nx, dx, c->xp, c->xm);
    ppe_axis(ny, dy, c->yp, c->ym);
}

static void ppe_coef_free(ppe_coef *c)
{
    free(c->xp); free(c->xm); free(c->yp); free(c->ym);
}

/* r = rhs - A p at one fluid cell; returns the neighbour sum in *sum and
 * the diagonal in *ap. */
static inline void ppe_cell(const ppe_coef *c, int nx, int i, int j,
                            uint8_t f, const double *p, double *ap,
                            double *sum)
{
    const int id = IDX(i, j, nx);
    const double aE = (f & OF_OPEN_E) ? c->xp[i] : 0.0;
    const double aW = (f & OF_OPEN_W) ? c->xm[i] : 0.0;
    const double aN = (f & OF_OPEN_N) ? c->yp[j] : 0.0;
    const double aS = (f & OF_OPEN_S) ? c->ym[j] : 0.0;
    *ap = -((aE + aW) + (aN + aS));
    *sum = (aE * p[id + 1] + aW * p[id - 1]) + (aN * p[id + nx] + aS * p[id - nx]);
}

static double residual_inf(const ppe_coef *c, int nx, int ny,
                           const uint8_t *flags, const double *rhs,
                           const double *p, double *r)
{
    double *row = calloc(ny, sizeof(double));
#pragma omp parallel for schedule(static)
    for (int j = 1; j < ny - 1; j++) {
        double m = 0.0;
        for (int i = 1; i < nx - 1; i++) {
            const int id = IDX(i, j, nx);
            const uint8_t f = flags[id];
            double rr = 0.0;
            if (f & OF_FLUID) {
                double ap, sum;
                ppe_cell(c, nx, i, j, f, p, &ap, &sum);
                if (ap != 0.0) rr = rhs[id] - (ap * p[id] + sum);
            }
            if (r) r[id] = rr;
            if (!(fabs(rr) <= 1.0e300)) m = INFINITY;
            else if (fabs(rr) > m) m = fabs(rr);
        }
        row[j] = m;
    }
    double m = 0.0;
    for (int j = 1; j < ny - 1; j++) if (row[j] > m) m = row[j];
    free(row);
    return m;
}

double of_phys_residual(int nx, int ny, const double *dx, const double *dy,
                        const uint8_t *flags, const double *rhs,
                        const double *p, double *r)
{
    ppe_coef c;
    ppe_coef_build(nx, ny, dx, dy, &c);
    const double m = residual_inf(&c, nx, ny, flags, rhs, p, r);
    ppe_coef_free(&c);
    return m;
}

/* Zebra line SOR, the deck's PPE_Solver 2 "Line SOR" (inputs.txt:14-15):
 * rows j odd first, then rows j even; each row's cells are solved together
 * with the rows above and below held, then relaxed, p <- p + om (x - p).
 * Row equation of a fluid cell with sa = (aE + aW) + (aN + aS) != 0:
 *     -aW x[i-1] + sa x[i] - aE x[i+1] = (aN pN + aS pS) - rhs;
 * any other cell keeps its value (an identity row, l = u = 0, b = 1,
 * d = p).
 *
 * The row's tridiagonal system (cells i = 1..n, n = nx - 2) is solved by a
 * partition (Schur-complement) method, the form a GPU runs with one lane
 * per segment:
 *   - the row is cut into S blocks of B cells, B = max(2, ceil(n / 64)),
 *     S = ceil(n / B) <= 64; the last cell of each block but the last is a
 *     separator, the block's other cells its segment (at least one cell);
 *   - within a segment, x = g + h xL + q xR, with xL, xR the separators on
 *     either side (0 where there is none): g solves the segment's system
 *     with the row's right side, h and q with -l of the first cell and -u
 *     of the last cell (the Thomas algorithm, pivots as reciprocals r);
 *   - the separators' Schur complement is tridiagonal (m = S - 1
 *     unknowns) and is solved by parallel cyclic reduction with strides
 *     1, 2, 4, ... < m.
 * Every quantity that depends only on the flags and the metric (r, the
 * negated products -l r and -u r, h, q, the reduction's multipliers and
 * final reciprocals) is formed once per sweep by line_factor_row; a row
 * solve applies them to the right side with the exact operations listed
 * there (fma is the fused multiply-add, one rounding).  The result equals
 * the Thomas algorithm's to round-off (pinned by the row equations). */
enum { LF_R, LF_NLR, LF_NCP, LF_H, LF_Q, LF_PLANES };  /* per-cell planes [B][64] */
enum { LF_PCR_LEVELS = 6, LF_PCR = 2 * LF_PCR_LEVELS + 1 };  /* alpha, gamma per level; 1 / B */

static void line_partition(int n, int *B, int *S)
{
    int b = (n + 63) / 64;
    if (b < 2) b = 2;
    *B = b;
    *S = (n + b - 1) / b;
}

/* Row coefficients of cell (i, j): l = -aW, b = sa, u = -aE for a solved
 * cell (returns 1), the identity row otherwise. */
static int line_cell(const ppe_coef *c, int nx, int i, int j, uint8_t f, double *l, double *b, double *u)
{
    *l = 0.0;
    *b = 1.0;
    *u = 0.0;
    (void)nx;
    if (!(f & OF_FLUID)) return 0;
    const double aE = (f & OF_OPEN_E) ? c->xp[i] : 0.0;
    const double aW = (f & OF_OPEN_W) ? c->xm[i] : 0.0;
    const double aN = (f & OF_OPEN_N) ? c->yp[j] : 0.0;
    const double aS = (f & OF_OPEN_S) ? c->ym[j] : 0.0;
    const double sa = (aE + aW) + (aN + aS);
    if (sa == 0.0) return 0;
    *l = -aW;
    *b = sa;
    *u = -aE;
    return 1;
}

/* Factors of row j: F[(plane * B + t) * 64 + k] for segment k's t-th cell
 * (plane LF_R at t = B - 1 holds the separator's -l, LF_NLR its -u), then
 * the reduction's LF_PCR planes of 64 at F + LF_PLANES * B * 64. */
static void line_factor_row(const ppe_coef *c, int nx, int j, const uint8_t *flags, int B, int S, double *F)
{
    const int n = nx - 2, m = S - 1;
#define LFP(pl, t, k) F[((size_t)(pl) * B + (t)) * 64 + (k)]
    for (int k = 0; k < S; k++) {
        const int a = 1 + k * B, len = k < S - 1 ? B - 1 : n - (S - 1) * B;
        double cpv = 0.0, hpv = 0.0;
        for (int t = 0; t < len; t++) {
            const int i = a + t;
            double l, b, u;
            line_cell(c, nx, i, j, flags[IDX(i, j, nx)], &l, &b, &u);
            const double den = t == 0 ? b : b - l * cpv;
            const double r = 1.0 / den, cp = u * r, nlr = -(l * r);
            const double hp = t == 0 ? nlr : nlr * hpv;
            LFP(LF_R, t, k) = r;
            LFP(LF_NLR, t, k) = nlr;
            LFP(LF_NCP, t, k) = -cp;
            LFP(LF_H, t, k) = hp;
            cpv = cp;
            hpv = hp;
        }
        LFP(LF_Q, len - 1, k) = LFP(LF_NCP, len - 1, k);
        for (int t = len - 2; t >= 0; t--) {
            LFP(LF_H, t, k) = fma(LFP(LF_NCP, t, k), LFP(LF_H, t + 1, k), LFP(LF_H, t, k));
            LFP(LF_Q, t, k) = LFP(LF_NCP, t, k) * LFP(LF_Q, t + 1, k);
        }
    }
    double A[64], Bv[64], Cv[64], nA[64], nB[64], nC[64];
    double *pcr = F + (size_t)LF_PLANES * B * 64;
    for (int q = 0; q < LF_PCR * 64; q++) pcr[q] = 0.0;
    for (int k = 0; k < m; k++) {
        const int sp = (k + 1) * B;
        double l, b, u;
        line_cell(c, nx, sp, j, flags[IDX(sp, j, nx)], &l, &b, &u);
        const double hL = LFP(LF_H, B - 2, k), qL = LFP(LF_Q, B - 2, k);
        const double hR = LFP(LF_H, 0, k + 1), qR = LFP(LF_Q, 0, k + 1);
        A[k] = l * hL;
        Cv[k] = u * qR;
        Bv[k] = (b + l * qL) + u * hR;
        LFP(LF_R, B - 1, k) = -l;
        LFP(LF_NLR, B - 1, k) = -u;
    }
    int lev = 0;
    for (int sg = 1; sg < m; sg *= 2, lev++) {
        for (int k = 0; k < m; k++) {
            const int lo = k >= sg, hi = k + sg < m;
            const double alpha = lo ? -(A[k] / Bv[k - sg]) : 0.0;
            const double gamma = hi ? -(Cv[k] / Bv[k + sg]) : 0.0;
            const double Am = lo ? A[k - sg] : 0.0, Cm = lo ? Cv[k - sg] : 0.0;
            const double Ap = hi ? A[k + sg] : 0.0, Cp = hi ? Cv[k + sg] : 0.0;
            nA[k] = alpha * Am;
            nC[k] = gamma * Cp;
            nB[k] = (Bv[k] + alpha * Cm) + gamma * Ap;
            pcr[(2 * lev) * 64 + k] = alpha;
            pcr[(2 * lev + 1) * 64 + k] = gamma;
        }
        memcpy(A, nA, sizeof(double) * m);
        memcpy(Bv, nB, sizeof(double) * m);
        memcpy(Cv, nC, sizeof(double) * m);
    }
    for (int k = 0; k < m; k++) pcr[(2 * LF_PCR_LEVELS) * 64 + k] = 1.0 / Bv[k];
#undef LFP
}

/* Solve row j with the held rows above and below and relax it (p in
 * place).  d, g, x: scratch of B * 64. */
static void line_solve_row(const ppe_coef *c, int nx, int j, const uint8_t *flags, const double *rhs, double om,
                           int B, int S, const double *F, double *p, double *g)
{
    const int n = nx - 2, m = S - 1;
#define LFP(pl, t, k) F[((size_t)(pl) * B + (t)) * 64 + (k)]
#define GV(t, k) g[(size_t)(t) * 64 + (k)]
    /* the right side d (in g), segment layout */
    for (int i = 1; i <= n; i++) {
        const int id = IDX(i, j, nx);
        const uint8_t f = flags[id];
        double d = p[id];
        if (f & OF_FLUID) {
            const double aE = (f & OF_OPEN_E) ? c->xp[i] : 0.0;
            const double aW = (f & OF_OPEN_W) ? c->xm[i] : 0.0;
            const double aN = (f & OF_OPEN_N) ? c->yp[j] : 0.0;
            const double aS = (f & OF_OPEN_S) ? c->ym[j] : 0.0;
            if ((aE + aW) + (aN + aS) != 0.0) d = (aN * p[id + nx] + aS * p[id - nx]) - rhs[id];
        }
        GV((i - 1) % B, (i - 1) / B) = d;
    }
    /* segments: forward gp = fma(-l r, gp, d r), backward g = fma(-u r, g, gp) */
    for (int k = 0; k < S; k++) {
        const int len = k < S - 1 ? B - 1 : n - (S - 1) * B;
        double gp = 0.0;
        for (int t = 0; t < len; t++) {
            const double dr = GV(t, k) * LFP(LF_R, t, k);
            gp = t == 0 ? dr : fma(LFP(LF_NLR, t, k), gp, dr);
            GV(t, k) = gp;
        }
        double gx = GV(len - 1, k);
        for (int t = len - 2; t >= 0; t--) {
            gx = fma(LFP(LF_NCP, t, k), gx, GV(t, k));
            GV(t, k) = gx;
        }
    }
    /* separators: D = fma(-u, g right, fma(-l, g left, d)), then the
     * reduction D = fma(gamma, D[k + s], fma(alpha, D[k - s], D)) (0
     * outside 0..m-1) and x = D / B as D times the reciprocal */
    double D[64], nD[64], xs[64];
    for (int k = 0; k < m; k++)
        D[k] = fma(LFP(LF_NLR, B - 1, k), GV(0, k + 1), fma(LFP(LF_R, B - 1, k), GV(B - 2, k), GV(B - 1, k)));
    const double *pcr = F + (size_t)LF_PLANES * B * 64;
    int lev = 0;
    for (int sg = 1; sg < m; sg *= 2, lev++) {
        for (int k = 0; k < m; k++) {
            const double Dm = k >= sg ? D[k - sg] : 0.0, Dp = k + sg < m ? D[k + sg] : 0.0;
            nD[k] = fma(pcr[(2 * lev + 1) * 64 + k], Dp, fma(pcr[(2 * lev) * 64 + k], Dm, D[k]));
        }
        memcpy(D, nD, sizeof(double) * m);
    }
    for (int k = 0; k < m; k++) xs[k] = D[k] * pcr[(2 * LF_PCR_LEVELS) * 64 + k];
    /* x = fma(q, xR, fma(h, xL, g)); separators take xs */
    for (int k = 0; k < S; k++) {
        const int len = k < S - 1 ? B - 1 : n - (S - 1) * B;
        const double xL = k >= 1 ? xs[k - 1] : 0.0, xR = k < m ? xs[k] : 0.0;
        for (int t = 0; t < len; t++) GV(t, k) = fma(LFP(LF_Q, t, k), xR, fma(LFP(LF_H, t, k), xL, GV(t, k)));
        if (k < m) GV(B - 1, k) = xs[k];
    }
    /* relax the solved cells only */
    for (int i = 1; i <= n; i++) {
        const int id = IDX(i, j, nx);
        const uint8_t f = flags[id];
        if (f & OF_FLUID) {
            const double aE = (f & OF_OPEN_E) ? c->xp[i] : 0.0;
            const double aW = (f & OF_OPEN_W) ? c->xm[i] : 0.0;
            const double aN = (f & OF_OPEN_N) ? c->yp[j] : 0.0;
            const double aS = (f & OF_OPEN_S) ? c->ym[j] : 0.0;
            if ((aE + aW) + (aN + aS) != 0.0) {
                const double x = GV((i - 1) % B, (i - 1) / B);
                p[id] = p[id] + om * (x - p[id]);
            }
        }
    }
#undef GV
#undef LFP
}

static void line_sor_sweep(const ppe_coef *c, int nx, int ny, const uint8_t *flags,
                           const double *rhs, double om, double *p)
{
    int B, S;
    line_partition(nx - 2, &B, &S);
    const size_t fsz = ((size_t)LF_PLANES * B + LF_PCR) * 64;
    for (int colour = 1; colour >= 0; colour--) {
#pragma omp parallel
        {
            double *F = malloc(sizeof(double) * fsz), *g = malloc(sizeof(double) * (size_t)B * 64);
#pragma omp for schedule(static)
            for (int j = 1; j < ny - 1; j++) {
                if ((j & 1) != colour) continue;
                line_factor_row(c, nx, j, flags, B, S, F);
                line_solve_row(c, nx, j, flags, rhs, om, B, S, F, p, g);
            }
            free(F);
            free(g);
        }
    }
}

/* Red-black SOR sweep (colour (i+j)&1 == 0 first), the parallel form of the
 * deck's PPE_Solver 1 "Point GS" with relaxation w-PPE (inputs.txt:14-15);
 * weighted Jacobi (jacobiIteration, PPESolver.cu:13-31, with the RHS and
 * the relaxation weight) when ppe_solver == 0.  On the all-Neumann problem
 * plain Jacobi (weight 1) does not converge: the checkerboard mode has
 * iteration eigenvalue -1. */
static void ppe_sweep(const ppe_coef *c, int nx, int ny, const uint8_t *flags,
                      const double *rhs, const of_params *prm, double *p,
                      double *tmp)
{
    const double om = prm->ppe_omega;
    if (prm->ppe_solver == 0) {
#pragma omp parallel for schedule(static)
        for (int j = 1; j < ny - 1; j++)
            for (int i = 1; i < nx - 1; i++) {
                const int id = IDX(i, j, nx);
                const uint8_t f = flags[id];
                double val = p[id];
                if (f & OF_FLUID) {
                    double ap, sum;
                    ppe_cell(c, nx, i, j, f, p, &ap, &sum);
                    if (ap != 0.0) {
                        const double gs = (rhs[id] - sum) / ap;
                        val = p[id] + om * (gs - p[id]);
                    }
                }
                tmp[id] = val;
            }
#pragma omp parallel for schedule(static)
        for (int j = 1; j < ny - 1; j++)
            for (int i = 1; i < nx - 1; i++) p[IDX(i, j, nx)] = tmp[IDX(i, j, nx)];
        return;
    }
    if (prm->ppe_solver == 2) {
        line_sor_sweep(c, nx, ny, flags, rhs, om, p);
        return;
    }
    /* p' = p + om (gs - p), gs = (rhs - sum) / ap, evaluated as
     * (1 - om) p + w (sum - rhs) with the cell's weight w = om / sa,
     * sa = -ap (the form whose weight a sweep kernel can hold per launch) */
    const double c1 = 1.0 - om;
    for (int colour = 0; colour < 2; colour++) {
#pragma omp parallel for schedule(static)
        for (int j = 1; j < ny - 1; j++)
            for (int i = 1 + ((j + 1 + colour) & 1); i < nx - 1; i += 2) {
                const int id = IDX(i, j, nx);
                const uint8_t f = flags[id];
                if (!(f & OF_FLUID)) continue;
                double ap, sum;
                ppe_cell(c, nx, i, j, f, p, &ap, &sum);
                if (ap == 0.0) continue;
                const double w = om / (-ap);
                p[id] = c1 * p[id] + w * (sum - rhs[id]);
            }
    }
}

int of_phys_poisson(int nx, int ny, const double *dx, const double *dy,
                    const uint8_t *flags, const double *rhs,
                    const of_params *prm, double *p, double *res_out)
{
    ppe_coef c;
    ppe_coef_build(nx, ny, dx, dy, &c);
    double *tmp = prm->ppe_solver == 0 ? calloc((size_t)nx * ny, sizeof(double)) : NULL;
    const int every = prm->check_every > 0 ? prm->check_every : 1;
    int sweeps = 0;
    double res = prm->dt * residual_inf(&c, nx, ny, flags, rhs, p, NULL);
    while (res > prm->ppe_tol && sweeps < prm->ppe_itermax) {
        int k = every;
        if (sweeps + k > prm->ppe_itermax) k = prm->ppe_itermax - sweeps;
        for (int s = 0; s < k; s++) ppe_sweep(&c, nx, ny, flags, rhs, prm, p, tmp);
        sweeps += k;
        res = prm->dt * residual_inf(&c, nx, ny, flags, rhs, p, NULL);
    }
    if (res_out) *res_out = res;
    free(tmp);
    ppe_coef_free(&c);
    return sweeps;
}

void of_phys_init(int nx, int ny, const double *dx, const double *dy,
                  const uint8_t *flags, const int8_t *bid, int nbody,
                  const of_body *bodies, int ng, const of_ghost *gh,
                  const of_params *prm, double *u, double *v, double *uf,
                  double *vf, double *p)
{
    /* fluid cells start at the free stream, solid cells at zero; ghost
     * cells then take their image-point values. */
    const size_t n = (size_t)nx * ny;
    for (size_t id = 0; id < n; id++) {
        u[id] = (flags[id] & OF_FLUID) ? prm->ubc : 0.0;
        v[id] = (flags[id] & OF_FLUID) ? prm->vbc : 0.0;
        p[id] = 0.0;
    }
    domain_bc_vel(nx, ny, u, prm->ubc);
    domain_bc_vel(nx, ny, v, prm->vbc);
    of_phys_ghost_velocity(nx, ny, ng, gh, nbody, bodies, u, v);
    faces_from_cells(nx, ny, dx, dy, flags, bid, bodies, prm, u, v, uf, vf);
}

void of_phys_restart(int nx, int ny, const double *dx, const double *dy,
                     const uint8_t *flags, const int8_t *bid, int nbody,
                     const of_body *bodies, int ng, const of_ghost *gh,
                     const of_params *prm, double *u, double *v, double *uf,
                     double *vf, double *p)
{
    (void)p;
    of_phys_ghost_velocity(nx, ny, ng, gh, nbody, bodies, u, v);
    faces_from_cells(nx, ny, dx, dy, flags, bid, bodies, prm, u, v, uf, vf);
}

void of_phys_step(int nx, int ny, const double *dx, const double *dy,
                  const uint8_t *flags, const int8_t *bid, int nbody,
                  const of_body *bodies, int ng, const of_ghost *gh,
                  const of_params *prm, double *u, double *v, double *uf,
                  double *vf, double *p, of_stats *st)
{
    const size_t n = (size_t)nx * ny;
    double *iblank = malloc(n * sizeof(double));
    for (size_t id = 0; id < n; id++) iblank[id] = (flags[id] & OF_FLUID) ? 1.0 : 0.0;

    ad_coef ac;
    ad_coef_build(nx, ny, dx, dy, prm->dt, prm->re, &ac);
    double *sx = calloc(n, sizeof(double)), *sy = calloc(n, sizeof(double));
    double *un = malloc(n * sizeof(double)), *vn = malloc(n * sizeof(double));
    double *rhs = calloc(n, sizeof(double));

    /* P1: explicit convective source from the projected face velocities */
    source_cells(nx, ny, dx, dy, prm->dt, prm->conv_factor, u, uf, vf, sx);
    source_cells(nx, ny, dx, dy, prm->dt, prm->conv_factor, v, uf, vf, sy);

    /* P2: backward-Euler viscous Jacobi sweeps (ADSolver.cu:315-370) */
    memcpy(un, u, n * sizeof(double));
    memcpy(vn, v, n * sizeof(double));
    double res = 1.0 + 1.0;
    int iter = 0;
    double *uc = u, *vc = v, *uo = un, *vo = vn;
    const double w_ad = prm->ad_omega > 0.0 ? prm->ad_omega : 1.0;
    while (res > prm->ad_tol && iter < prm->ad_itermax) {
        ad_sweep(nx, ny, &ac, iblank, sx, uc, uo, w_ad);
        ad_sweep(nx, ny, &ac, iblank, sy, vc, vo, w_ad);
        domain_bc_vel(nx, ny, uo, prm->ubc);
        domain_bc_vel(nx, ny, vo, prm->vbc);
        of_phys_ghost_velocity(nx, ny, ng, gh, nbody, bodies, uo, vo);
        res = l1_change(nx, ny, iblank, uo, uc) + l1_change(nx, ny, iblank, vo, vc);
        double *t = uc; uc = uo; uo = t;
        t = vc; vc = vo; vo = t;
        iter++;
    }
    if (uc != u) { memcpy(u, uc, n * sizeof(double)); memcpy(v, vc, n * sizeof(double)); }

    /* P3: intermediate face velocities and the divergence RHS */
    faces_from_cells(nx, ny, dx, dy, flags, bid, bodies, prm, u, v, uf, vf);
    of_phys_rhs(nx, ny, dx, dy, flags, prm->dt, uf, vf, rhs);

    /* P4/P5: pressure Poisson, warm-started from p^n */
    double pres;
    const int sweeps = of_phys_poisson(nx, ny, dx, dy, flags, rhs, prm, p, &pres);
    domain_bc_p(nx, ny, p);
    of_phys_ghost_pressure(nx, ny, ng, gh, nbody, bodies, p);

    /* P7: projection of faces (open faces only) and cells */
#pragma omp parallel for schedule(static)
    for (int j = 1; j < ny - 1; j++)
        for (int i = 1; i < nx - 1; i++) {
            const int id = IDX(i, j, nx);
            const uint8_t f = flags[id];
            if (f & OF_OPEN_E) {
                const double g = (p[id + 1] - p[id]) * (2.0 / (dx[i] + dx[i + 1]));
                uf[i + (j - 1) * (nx - 1)] = uf[i + (j - 1) * (nx - 1)] - prm->dt * g;
            }
            if (f & OF_OPEN_N) {
                const double g = (p[id + nx] - p[id]) * (2.0 / (dy[j] + dy[j + 1]));
                vf[(i - 1) + j * (nx - 2)] = vf[(i - 1) + j * (nx - 2)] - prm->dt * g;
            }
        }
#pragma omp parallel for schedule(static)
    for (int j = 1; j < ny - 1; j++)
        for (int i = 1; i < nx - 1; i++) {
            const int id = IDX(i, j, nx);
            const double dxi = dx[i], dxp = dx[i + 1], dxm = dx[i - 1];
            const double dyj = dy[j], dyp = dy[j + 1], dym = dy[j - 1];
            const double pe = (1 / (dxi + dxp)) * (p[id + 1] * dxi + p[id] * dxp);
            const double pw = (1 / (dxi + dxm)) * (p[id] * dxm + p[id - 1] * dxi);
            const double pn = (1 / (dyj + dyp)) * (p[id + nx] * dyj + p[id] * dyp);
            const double ps = (1 / (dyj + dym)) * (p[id] * dym + p[id - nx] * dyj);
            un[id] = iblank[id] * (u[id] - prm->dt * ((pe - pw) * (1.0 / dxi)));
            vn[id] = iblank[id] * (v[id] - prm->dt * ((pn - ps) * (1.0 / dyj)));
        }
#pragma omp parallel for schedule(static)
    for (int j = 1; j < ny - 1; j++)
        for (int i = 1; i < nx - 1; i++) {
            u[IDX(i, j, nx)] = un[IDX(i, j, nx)];
            v[IDX(i, j, nx)] = vn[IDX(i, j, nx)];
        }
    domain_bc_vel(nx, ny, u, prm->ubc);
    domain_bc_vel(nx, ny, v, prm->vbc);
    of_phys_ghost_velocity(nx, ny, ng, gh, nbody, bodies, u, v);

    /* P8: divergence check */
    const double dmax = of_phys_divergence(nx, ny, dx, dy, flags, uf, vf, NULL);
    if (st) {
        st->ad_iters = iter;
        st->ad_res = res;
        st->ppe_sweeps = sweeps;
        st->ppe_res = pres;
        st->div_max = dmax;
    }
    free(iblank); free(sx); free(sy); free(un); free(vn); free(rhs);
    ad_coef_free(&ac);
}
