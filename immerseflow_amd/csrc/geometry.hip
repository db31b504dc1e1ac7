// Immersed-boundary setup on gfx950: iBlank classification, ghost-cell list
// and image-point weights.
//
// The reference's iBlankComputeKernel (src/include/preSim.cu:110-136) is a
// stub that writes iBlank = 1.0; its commented-out circle test (:126-130)
// shows the intended per-cell classification.  The ghost-cell / image-point
// machinery follows the sharp-interface method of Mittal et al. (J. Comput.
// Phys. 227 (2008) 4825-4852).  Integer outputs (flags, ghost list) are
// exact: the geometry tests use only + - * / and compares in FP64 with no
// contraction, on parameters computed by the host.
#include "common.h"
#include "kernels.h"

namespace iff {

__device__ static bool inside_body(const BodyDev &b, double x, double y)
{
    const double ex = x - b.xc, ey = y - b.yc;
    if (b.kind == IF_BODY_CIRCLE) return ex * ex + ey * ey <= b.r * b.r;
    if (b.kind == IF_BODY_ELLIPSE) {
        const double X = ex * b.c + ey * b.s;
        const double Y = ey * b.c - ex * b.s;
        return (X * X) / (b.a * b.a) + (Y * Y) / (b.b * b.b) <= 1.0;
    }
    // polygon: bounding circle, then the crossing-number test written with
    // orientation predicates (no division)
    if (!(ex * ex + ey * ey <= b.r * b.r)) return false;
    bool in = false;
    const int n = b.nvert;
    for (int k = 0; k < n; k++) {
        const int k1 = (k + 1 == n) ? 0 : k + 1;
        const double ax = b.vx[k], ay = b.vy[k];
        const double bx = b.vx[k1], by = b.vy[k1];
        if ((ay > y) != (by > y)) {
            const double cr = (bx - ax) * (y - ay) - (by - ay) * (x - ax);
            if (by > ay) { if (cr > 0) in = !in; }
            else         { if (cr < 0) in = !in; }
        }
    }
    return in;
}

__device__ static int owner_of(const CaseDev &c, int i, int j)
{
    const double x = c.xc[i], y = c.yc[j];
    for (int b = 0; b < c.nbody; b++)
        if (inside_body(c.bodies[b], x, y)) return b;
    return -1;
}

// Flags and body ids of every cell.  fresh != 0: cells that turn from solid
// to fluid (moving bodies) take the velocity of the body that left them.
__global__ __launch_bounds__(256) void k_classify(const CaseDev *cs, int fresh)
{
    const CaseDev &c = cs[blockIdx.y];
    if (fresh && (!c.moving || !c.ctrl->active)) return;
    const int L = blockIdx.x * 256 + threadIdx.x;
    const int nx = c.nx, ny = c.ny;
    if (L >= nx * ny) return;
    const int i = L % nx, j = L / nx;
    const int id = i + j * c.P;
    const int own = owner_of(c, i, j);
    uint8_t f = (own < 0) ? F_FLUID : 0;
    if (i >= 1 && i <= nx - 2 && j >= 1 && j <= ny - 2) {
        const bool fe = owner_of(c, i + 1, j) < 0, fw = owner_of(c, i - 1, j) < 0;
        const bool fn = owner_of(c, i, j + 1) < 0, fs = owner_of(c, i, j - 1) < 0;
        if (f) {
            if (i + 1 <= nx - 2 && fe) f |= F_OPEN_E;
            if (i - 1 >= 1 && fw) f |= F_OPEN_W;
            if (j + 1 <= ny - 2 && fn) f |= F_OPEN_N;
            if (j - 1 >= 1 && fs) f |= F_OPEN_S;
        } else if (fe || fw || fn || fs) {
            f |= F_GHOST;
        }
    }
    if (fresh) {
        const uint8_t old = c.flags[id];
        const int ob = c.bodyid[id];
        if (!(old & F_FLUID) && (f & F_FLUID) && ob >= 0) {
            const int base = c.ctrl->ad_base;
            if (c.prec == IF_FP64) {
                ((double *)c.u[base])[id] = c.bodies[ob].ub;
                ((double *)c.v[base])[id] = c.bodies[ob].vb;
            } else {
                ((float *)c.u[base])[id] = (float)c.bodies[ob].ub;
                ((float *)c.v[base])[id] = (float)c.bodies[ob].vb;
            }
        }
    }
    c.flags[id] = f;
    c.bodyid[id] = (int8_t)own;
}

// Ghost list, sorted by linear cell index: per-block counts over 1024
// consecutive reference indices, one exclusive scan per case, ordered
// scatter.
constexpr int kGhostSpan = 1024;

__global__ __launch_bounds__(256) void k_ghost_count(const CaseDev *cs, int fresh)
{
    const CaseDev &c = cs[blockIdx.y];
    if (fresh && (!c.moving || !c.ctrl->active)) return;
    const int ncell = c.nx * c.ny;
    // the grid is sized for the largest case of the launch
    if ((int)blockIdx.x >= (ncell + kGhostSpan - 1) / kGhostSpan) return;
    const int L0 = blockIdx.x * kGhostSpan + threadIdx.x * 4;
    int n = 0;
    for (int k = 0; k < 4; k++) {
        const int L = L0 + k;
        if (L < ncell) {
            const int i = L % c.nx, j = L / c.nx;
            if (c.flags[i + j * c.P] & F_GHOST) n++;
        }
    }
    __shared__ int s[256];
    s[threadIdx.x] = n;
    __syncthreads();
    for (int o = 128; o > 0; o >>= 1) {
        if (threadIdx.x < o) s[threadIdx.x] += s[threadIdx.x + o];
        __syncthreads();
    }
    if (threadIdx.x == 0) c.gcount[blockIdx.x] = s[0];
}

__global__ __launch_bounds__(256) void k_ghost_scan(const CaseDev *cs, int fresh)
{
    const CaseDev &c = cs[blockIdx.x];
    if (fresh && (!c.moving || !c.ctrl->active)) return;
    const int nb = (c.nx * c.ny + kGhostSpan - 1) / kGhostSpan;
    int *cnt = c.gcount;
    // serial chunks per thread, then a scan over the 256 chunk totals
    const int per = (nb + 255) / 256;
    const int b0 = threadIdx.x * per;
    int tot = 0;
    for (int k = 0; k < per; k++)
        if (b0 + k < nb) tot += cnt[b0 + k];
    __shared__ int s[256];
    s[threadIdx.x] = tot;
    __syncthreads();
    if (threadIdx.x == 0) {
        int acc = 0;
        for (int k = 0; k < 256; k++) { const int t = s[k]; s[k] = acc; acc += t; }
        Ctrl *ct = c.ctrl;
        ct->ng_total = acc;
        if (acc > c.gcap) { ct->failed = 1; acc = c.gcap; }
        ct->ng = acc;
    }
    __syncthreads();
    int acc = s[threadIdx.x];
    for (int k = 0; k < per; k++)
        if (b0 + k < nb) { const int t = cnt[b0 + k]; cnt[b0 + k] = acc; acc += t; }
}

__global__ __launch_bounds__(256) void k_ghost_scatter(const CaseDev *cs, int fresh)
{
    const CaseDev &c = cs[blockIdx.y];
    if (fresh && (!c.moving || !c.ctrl->active)) return;
    const int ncell = c.nx * c.ny;
    if ((int)blockIdx.x >= (ncell + kGhostSpan - 1) / kGhostSpan) return;
    const int L0 = blockIdx.x * kGhostSpan + threadIdx.x * 4;
    int n = 0;
    for (int k = 0; k < 4; k++) {
        const int L = L0 + k;
        if (L < ncell) {
            const int i = L % c.nx, j = L / c.nx;
            if (c.flags[i + j * c.P] & F_GHOST) n++;
        }
    }
    __shared__ int s[256];
    s[threadIdx.x] = n;
    __syncthreads();
    if (threadIdx.x == 0) {
        int acc = 0;
        for (int k = 0; k < 256; k++) { const int t = s[k]; s[k] = acc; acc += t; }
    }
    __syncthreads();
    int pos = c.gcount[blockIdx.x] + s[threadIdx.x];
    for (int k = 0; k < 4; k++) {
        const int L = L0 + k;
        if (L < ncell) {
            const int i = L % c.nx, j = L / c.nx;
            const int id = i + j * c.P;
            if (c.flags[id] & F_GHOST) {
                if (pos < c.gcap) {
                    c.gh[pos].gidx = id;
                    c.gh[pos].body = c.bodyid[id] < 0 ? 0 : c.bodyid[id];
                }
                pos++;
            }
        }
    }
}

// Distance from a point to an ellipse: D. Eberly, "Distance from a Point to
// an Ellipse, an Ellipsoid, or a Hyperellipsoid" (Geometric Tools, 2013):
// bisection on F(s) = (r0 z0/(s+r0))^2 + (z1/(s+1))^2 - 1, first quadrant,
// e0 >= e1 > 0.
__device__ static void ellipse_closest_q1(double e0, double e1, double y0, double y1,
                                          double &x0, double &x1)
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
                x0 = r0 * y0 / (s + r0);
                x1 = y1 / (s + 1.0);
            } else {
                x0 = y0;
                x1 = y1;
            }
        } else {
            x0 = 0.0;
            x1 = e1;
        }
    } else {
        const double numer0 = e0 * y0, denom0 = e0 * e0 - e1 * e1;
        if (numer0 < denom0) {
            const double xde0 = numer0 / denom0;
            x0 = e0 * xde0;
            x1 = e1 * sqrt(1.0 - xde0 * xde0);
        } else {
            x0 = e0;
            x1 = 0.0;
        }
    }
}

// Boundary intercept: the closest point of the body boundary.
__device__ static void closest_point(const BodyDev &b, double px, double py, double &bx,
                                     double &by)
{
    const double ex = px - b.xc, ey = py - b.yc;
    if (b.kind == IF_BODY_CIRCLE) {
        const double L = sqrt(ex * ex + ey * ey);
        if (L == 0.0) { bx = b.xc + b.r; by = b.yc; return; }
        bx = b.xc + b.r * (ex / L);
        by = b.yc + b.r * (ey / L);
        return;
    }
    if (b.kind == IF_BODY_ELLIPSE) {
        const double X = ex * b.c + ey * b.s;
        const double Y = ey * b.c - ex * b.s;
        const bool swap = b.a < b.b;
        const double e0 = swap ? b.b : b.a, e1 = swap ? b.a : b.b;
        const double y0 = fabs(swap ? Y : X), y1 = fabs(swap ? X : Y);
        double x0, x1;
        ellipse_closest_q1(e0, e1, y0, y1, x0, x1);
        double cx = swap ? x1 : x0, cy = swap ? x0 : x1;
        if (X < 0) cx = -cx;
        if (Y < 0) cy = -cy;
        bx = b.xc + (cx * b.c - cy * b.s);
        by = b.yc + (cx * b.s + cy * b.c);
        return;
    }
    double best = -1.0, bxx = b.vx[0], byy = b.vy[0];
    const int n = b.nvert;
    for (int k = 0; k < n; k++) {
        const int k1 = (k + 1 == n) ? 0 : k + 1;
        const double ax = b.vx[k], ay = b.vy[k];
        const double sx = b.vx[k1] - ax, sy = b.vy[k1] - ay;
        const double L2 = sx * sx + sy * sy;
        double t = (L2 > 0.0) ? ((px - ax) * sx + (py - ay) * sy) / L2 : 0.0;
        if (t < 0.0) t = 0.0;
        if (t > 1.0) t = 1.0;
        const double cx = ax + t * sx, cy = ay + t * sy;
        const double d2 = (px - cx) * (px - cx) + (py - cy) * (py - cy);
        if (best < 0.0 || d2 < best) { best = d2; bxx = cx; byy = cy; }
    }
    bx = bxx;
    by = byy;
}

// Largest k in [0, n-2] with c[k] <= x (c strictly increasing).
__device__ static int bracket(int n, const double *c, double x)
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

// Image point IP = 2 BI - GC and its bilinear stencil.  Fluid nodes carry
// weights; the ghost cell itself is eliminated implicitly; other solid
// nodes are taken at the body value (Dirichlet) or dropped with
// renormalisation (Neumann).
__global__ __launch_bounds__(256) void k_ghost_weights(const CaseDev *cs, int fresh)
{
    const CaseDev &c = cs[blockIdx.y];
    if (fresh && (!c.moving || !c.ctrl->active)) return;
    const int g = blockIdx.x * 256 + threadIdx.x;
    if (g >= c.ctrl->ng) return;
    GhostDev &o = c.gh[g];
    const int id = o.gidx;
    const int P = c.P;
    const int gi = id % P, gj = id / P;
    const BodyDev &b = c.bodies[o.body];
    const double gx = c.xc[gi], gy = c.yc[gj];
    double bix, biy;
    closest_point(b, gx, gy, bix, biy);
    const double ipx = 2.0 * bix - gx, ipy = 2.0 * biy - gy;
    const int i0 = bracket(c.nx, c.xc, ipx), j0 = bracket(c.ny, c.yc, ipy);
    const double fx = (ipx - c.xc[i0]) / (c.xc[i0 + 1] - c.xc[i0]);
    const double fy = (ipy - c.yc[j0]) / (c.yc[j0 + 1] - c.yc[j0]);
    const double w[4] = {(1.0 - fx) * (1.0 - fy), fx * (1.0 - fy), (1.0 - fx) * fy, fx * fy};
    const int nd[4] = {i0 + j0 * P, i0 + 1 + j0 * P, i0 + (j0 + 1) * P, i0 + 1 + (j0 + 1) * P};
    double wself = 0.0, wsolid = 0.0, wfluid = 0.0;
    for (int k = 0; k < 4; k++) {
        if (nd[k] == id) wself += w[k];
        else if (c.flags[nd[k]] & F_FLUID) wfluid += w[k];
        else wsolid += w[k];
    }
    o.cd = (2.0 - wsolid) / (1.0 + wself);
    for (int k = 0; k < 4; k++) {
        const bool fl = (nd[k] != id) && (c.flags[nd[k]] & F_FLUID);
        o.node[k] = fl ? nd[k] : id;
        o.wd[k] = fl ? -w[k] / (1.0 + wself) : 0.0;
        o.wn[k] = (fl && wfluid > 0.0) ? w[k] / wfluid : 0.0;
    }
    o.lnx = ipx - gx;
    o.lny = ipy - gy;
    o.bix = bix;
    o.biy = biy;
    o.ipx = ipx;
    o.ipy = ipy;
}

void launch_classify(hipStream_t st, const CaseDev *cs, int ncase, int max_cells, int fresh)
{
    hipLaunchKernelGGL(k_classify, dim3((max_cells + 255) / 256, ncase), dim3(256), 0, st, cs,
                       fresh);
}

void launch_ghost_build(hipStream_t st, const CaseDev *cs, int ncase, int max_cells, int fresh)
{
    const int nb = (max_cells + kGhostSpan - 1) / kGhostSpan;
    hipLaunchKernelGGL(k_ghost_count, dim3(nb, ncase), dim3(256), 0, st, cs, fresh);
    hipLaunchKernelGGL(k_ghost_scan, dim3(ncase), dim3(256), 0, st, cs, fresh);
    hipLaunchKernelGGL(k_ghost_scatter, dim3(nb, ncase), dim3(256), 0, st, cs, fresh);
}

void launch_ghost_weights(hipStream_t st, const CaseDev *cs, int ncase, int max_ng, int fresh)
{
    if (max_ng <= 0) return;
    hipLaunchKernelGGL(k_ghost_weights, dim3((max_ng + 255) / 256, ncase), dim3(256), 0, st, cs,
                       fresh);
}

int ghost_span() { return kGhostSpan; }

}  // namespace iff
