// gfx950 kernels of the fractional step.
//
// Every kernel is batched: a launch covers the interior tiles of all
// resident cases, listed in a tile table (case id, tile id).  Block b is
// remapped so that consecutive tiles land on the same XCD (the dispatcher
// deals blocks round-robin over the 8 XCDs), which keeps halo re-reads of
// neighbouring tiles in one L2.
//
// Each case carries a device control block (Ctrl).  Loop kernels read the
// case's "done" flags and exit at once when the case has converged, so the
// host can enqueue sweeps speculatively and a case's sweep count is decided
// on the device exactly as the reference's host loop decides it
// (ADSolver.cu:315, PPESolver.cu:172) -- without a round trip per sweep.
//
// Expressions are evaluated in the operation order of the reference kernels
// (and of the FP64 oracle), compiled without FP contraction.
#include "common.h"
#include "kernels.h"
#include "sor.h"

namespace iff {


__device__ __forceinline__ double wave_sum(double v)
{
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}

__device__ __forceinline__ double wave_max(double v)
{
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v = fmax(v, __shfl_xor(v, o, 64));
    return v;
}

// Block (256 threads) reductions in a fixed order.
__device__ __forceinline__ double block_sum(double v)
{
    __shared__ double s[4];
    v = wave_sum(v);
    if ((threadIdx.x & 63) == 0) s[threadIdx.x >> 6] = v;
    __syncthreads();
    const double r = (s[0] + s[1]) + (s[2] + s[3]);
    __syncthreads();
    return r;
}

__device__ __forceinline__ double block_max(double v)
{
    __shared__ double s[4];
    v = wave_max(v);
    if ((threadIdx.x & 63) == 0) s[threadIdx.x >> 6] = v;
    __syncthreads();
    const double r = fmax(fmax(s[0], s[1]), fmax(s[2], s[3]));
    __syncthreads();
    return r;
}

// |r| for a max-reduction that lets a non-finite value surface as +inf.
__device__ __forceinline__ double absinf(double r)
{
    const double a = fabs(r);
    return (a <= 1.0e300) ? a : HUGE_VAL;
}

template <typename T> __device__ __forceinline__ T *fld(void *p) { return (T *)p; }

struct TileCtx {
    int cid, t, tx, ty;
};

__device__ __forceinline__ bool tile_setup(const int2 *tiles, int ntiles,
                                           int ntiles_pad, const CaseDev *cs,
                                           TileCtx &tc)
{
    const int t = tile_of_block(ntiles_pad);
    if (t >= ntiles) return false;
    const int2 tl = tiles[t];
    tc.cid = tl.x;
    tc.t = tl.y;
    const int ntx = cs[tl.x].ntx;
    tc.tx = tl.y % ntx;
    tc.ty = tl.y / ntx;
    return true;
}

// Velocity of the body owning the first solid side of a blanked face.
__device__ __forceinline__ double blank_vel(const CaseDev &c, int ida, int idb,
                                            int comp)
{
    int b = c.bodyid[ida];
    if (b < 0) b = c.bodyid[idb];
    if (b < 0) return 0.0;
    return comp == 0 ? c.bodies[b].ub : c.bodies[b].vb;
}

// ---------------------------------------------------------------------------
// step bracketing
// ---------------------------------------------------------------------------
__global__ void k_step_begin(const CaseDev *cs, int ncase)
{
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= ncase) return;
    Ctrl *c = cs[k].ctrl;
    if (c->active && c->steps_done >= c->steps_target) c->active = 0;
    c->ad_done = (cs[k].ad_itermax <= 0) ? 1 : 0;
    c->ad_iter = 0;
    c->ad_res = 0.0;
    c->ppe_done = 0;
    c->ppe_sweeps = 0;
    c->ppe_res = 0.0;
}

// ---------------------------------------------------------------------------
// P1: explicit convective source (ADSource, ADSolver.cu:46-79)
// ---------------------------------------------------------------------------
template <typename TS, typename AX>
__device__ __forceinline__ TS conv_source(const TS *q, int id, int P, const AX &ax,
                                          const AX &ay, int i, int j, TS ufe,
                                          TS ufw, TS vfn, TS vfs)
{
    const TS dx_i = ax.d[i], dx_ip1 = ax.d[i + 1], dx_im1 = ax.d[i - 1];
    const TS dy_j = ay.d[j], dy_jp1 = ay.d[j + 1], dy_jm1 = ay.d[j - 1];
    const TS qc = q[id];
    return (qc - ax.cs[i] * ((ax.ip[i] * (q[id + 1] * dx_i + qc * dx_ip1) * ufe)
                             - (ax.im[i] * (qc * dx_im1 + q[id - 1] * dx_i) * ufw)))
           - ay.cs[j] * ((ay.ip[j] * (q[id + P] * dy_j + qc * dy_jp1) * vfn)
                         - (ay.im[j] * (qc * dy_jm1 + q[id - P] * dy_j) * vfs));
}

template <typename TS>
__global__ __launch_bounds__(256) void k_source(const CaseDev *cs, const int2 *tiles,
                                                int ntiles, int ntiles_pad)
{
    TileCtx tc;
    if (!tile_setup(tiles, ntiles, ntiles_pad, cs, tc)) return;
    const CaseDev &c = cs[tc.cid];
    if (!c.ctrl->active) return;
    const int i = 1 + tc.tx * kTileX + (threadIdx.x & 63);
    if (i > c.nx - 2) return;
    const int base = c.ctrl->ad_base;
    const TS *u = fld<TS>(c.u[base]);
    const TS *v = fld<TS>(c.v[base]);
    TS *sx = fld<TS>(c.sx), *sy = fld<TS>(c.sy);
    const double *uf = c.uf[0], *vf = c.vf[0];
    const int P = c.P;
    for (int r = threadIdx.x >> 6; r < c.tile_y; r += 4) {
        const int j = 1 + tc.ty * c.tile_y + r;
        if (j > c.ny - 2) break;
        const int id = i + j * P;
        const TS ufe = (TS)uf[id], ufw = (TS)uf[id - 1];
        const TS vfn = (TS)vf[id], vfs = (TS)vf[id - P];
        if constexpr (sizeof(TS) == 8) {
            sx[id] = conv_source<TS>(u, id, P, c.x, c.y, i, j, ufe, ufw, vfn, vfs);
            sy[id] = conv_source<TS>(v, id, P, c.x, c.y, i, j, ufe, ufw, vfn, vfs);
        } else {
            sx[id] = conv_source<TS>(u, id, P, c.xf, c.yf, i, j, ufe, ufw, vfn, vfs);
            sy[id] = conv_source<TS>(v, id, P, c.xf, c.yf, i, j, ufe, ufw, vfn, vfs);
        }
    }
}

// ---------------------------------------------------------------------------
// P2: fused u+v backward-Euler viscous Jacobi sweep (ADusolver_kernel /
// ADvsolver_kernel, ADSolver.cu:81-120) with the L1 change of
// Compute_{u,v}Residual_AD (:122-160) reduced per tile, and the domain ghost
// rule of set_velocity_BC (:199-217) applied to the new iterate.
// ---------------------------------------------------------------------------
template <typename TS, typename AX>
__device__ __forceinline__ TS jacobi_cell(const TS *q, const TS *s, int id, int P,
                                          const AX &ax, const AX &ay, int i, int j,
                                          TS blank)
{
    const TS t = s[id] + ax.ap[i] * q[id + 1] + ax.am[i] * q[id - 1]
                 + ay.ap[j] * q[id + P] + ay.am[j] * q[id - P];
    return blank * t / (ax.ad[i] + ay.ad[j]);
}

template <typename TS>
__global__ __launch_bounds__(256) void k_ad_sweep(const CaseDev *cs, const int2 *tiles,
                                                  int ntiles, int ntiles_pad)
{
    TileCtx tc;
    if (!tile_setup(tiles, ntiles, ntiles_pad, cs, tc)) return;
    const CaseDev &c = cs[tc.cid];
    const Ctrl *ct = c.ctrl;
    if (!ct->active || ct->ad_done) return;
    const int base = ct->ad_base;
    const TS *u = fld<TS>(c.u[base]), *v = fld<TS>(c.v[base]);
    TS *un = fld<TS>(c.u[1 - base]), *vn = fld<TS>(c.v[1 - base]);
    const TS *sx = fld<TS>(c.sx), *sy = fld<TS>(c.sy);
    const int P = c.P, nx = c.nx, ny = c.ny;
    const TS ub2 = (TS)c.ubc * (TS)2.0, vb2 = (TS)c.vbc * (TS)2.0;
    const double w = c.ad_omega;
    const int i = 1 + tc.tx * kTileX + (threadIdx.x & 63);
    double acc = 0.0;
    if (i <= nx - 2) {
        for (int r = threadIdx.x >> 6; r < c.tile_y; r += 4) {
            const int j = 1 + tc.ty * c.tile_y + r;
            if (j > ny - 2) break;
            const int id = i + j * P;
            const bool fl = c.flags[id] & F_FLUID;
            const TS blank = fl ? (TS)1.0 : (TS)0.0;
            TS qu, qv;
            if constexpr (sizeof(TS) == 8) {
                qu = jacobi_cell<TS>(u, sx, id, P, c.x, c.y, i, j, blank);
                qv = jacobi_cell<TS>(v, sy, id, P, c.x, c.y, i, j, blank);
            } else {
                qu = jacobi_cell<TS>(u, sx, id, P, c.xf, c.yf, i, j, blank);
                qv = jacobi_cell<TS>(v, sy, id, P, c.xf, c.yf, i, j, blank);
            }
            if (w != 1.0) {  // w_AD relaxation (ADSolver.cu:95,116)
                qu = (TS)(1.0 - w) * u[id] + (TS)w * qu;
                qv = (TS)(1.0 - w) * v[id] + (TS)w * qv;
            }
            un[id] = qu;
            vn[id] = qv;
            if (fl) acc += (double)fabs(qu - u[id]) + (double)fabs(qv - v[id]);
            if (i == 1) { un[id - 1] = -qu + ub2; vn[id - 1] = -qv + vb2; }
            if (i == nx - 2) { un[id + 1] = -qu + ub2; vn[id + 1] = -qv + vb2; }
            if (j == 1) { un[id - P] = -qu + ub2; vn[id - P] = -qv + vb2; }
            if (j == ny - 2) { un[id + P] = -qu + ub2; vn[id + P] = -qv + vb2; }
        }
    }
    acc = block_sum(acc);
    if (threadIdx.x == 0) c.part[tc.t] = acc;
}

// One block per case: the stopping rule of ADSolver.cu:315 on the device.
__global__ __launch_bounds__(256) void k_ad_finalize(const CaseDev *cs, int ncase)
{
    const CaseDev &c = cs[blockIdx.x];
    Ctrl *ct = c.ctrl;
    if (!ct->active || ct->ad_done) return;
    const int nt = c.ntx * c.nty;
    double s = 0.0;
    for (int k = threadIdx.x; k < nt; k += 256) s += c.part[k];
    s = block_sum(s);
    if (threadIdx.x == 0) {
        ct->ad_res = s;
        ct->ad_iter += 1;
        ct->ad_base ^= 1;
        ct->tot_ad += 1;
        if (!(s <= 1.0e300)) ct->failed = 1;
        if (!(s > c.ad_tol) || ct->ad_iter >= c.ad_itermax) ct->ad_done = 1;
    }
}

// ---------------------------------------------------------------------------
// ghost-cell (image point) updates.  mode 0: on the new AD iterate (skipped
// once the case's AD loop is done); mode 1: on the current velocity.
// Stencil nodes are fluid cells (or the ghost itself with weight 0), which
// these kernels never write: no ordering between ghost cells is needed.
// ---------------------------------------------------------------------------
template <typename TS>
__global__ __launch_bounds__(256) void k_ghost_vel(const CaseDev *cs, int mode)
{
    const CaseDev &c = cs[blockIdx.y];
    const Ctrl *ct = c.ctrl;
    const int g = blockIdx.x * 256 + threadIdx.x;
    if (g >= ct->ng || !ct->active) return;
    int buf = ct->ad_base;
    if (mode == 0) {
        if (ct->ad_done) return;
        buf = 1 - buf;
    }
    TS *u = fld<TS>(c.u[buf]), *v = fld<TS>(c.v[buf]);
    const GhostDev &o = c.gh[g];
    const BodyDev &b = c.bodies[o.body];
    double su = o.cd * b.ub, sv = o.cd * b.vb;
#pragma unroll
    for (int k = 0; k < 4; k++) {
        su = su + o.wd[k] * (double)u[o.node[k]];
        sv = sv + o.wd[k] * (double)v[o.node[k]];
    }
    u[o.gidx] = (TS)su;
    v[o.gidx] = (TS)sv;
}

// Pressure: homogeneous Neumann on the domain ring (edges), then ghost
// cells p_GC = p_IP (+ moving-body term).
__global__ __launch_bounds__(256) void k_p_ring(const CaseDev *cs)
{
    const CaseDev &c = cs[blockIdx.y];
    if (!c.ctrl->active) return;
    double *p = c.p[c.ctrl->ppe_base];
    const int k = blockIdx.x * 256 + threadIdx.x;
    const int P = c.P, nx = c.nx, ny = c.ny;
    if (k < ny - 2) {
        const int j = 1 + k;
        p[0 + j * P] = p[1 + j * P];
        p[nx - 1 + j * P] = p[nx - 2 + j * P];
    }
    if (k < nx - 2) {
        const int i = 1 + k;
        p[i] = p[i + P];
        p[i + (ny - 1) * P] = p[i + (ny - 2) * P];
    }
}

__global__ __launch_bounds__(256) void k_ghost_p(const CaseDev *cs)
{
    const CaseDev &c = cs[blockIdx.y];
    const int g = blockIdx.x * 256 + threadIdx.x;
    if (g >= c.ctrl->ng || !c.ctrl->active) return;
    double *p = c.p[c.ctrl->ppe_base];
    const GhostDev &o = c.gh[g];
    const BodyDev &b = c.bodies[o.body];
    double sp = o.wn[0] * p[o.node[0]];
#pragma unroll
    for (int k = 1; k < 4; k++) sp = sp + o.wn[k] * p[o.node[k]];
    p[o.gidx] = sp + (o.lnx * b.axb + o.lny * b.ayb);
}

// ---------------------------------------------------------------------------
// P3: intermediate face velocities (Compute_velf interpolation,
// ADSolver.cu:173-174,184-185, on open faces; body velocity on blanked
// faces; the BC on domain faces) and the divergence RHS, in one pass.
// ---------------------------------------------------------------------------
template <typename TS>
__global__ __launch_bounds__(256) void k_faces_rhs(const CaseDev *cs, const int2 *tiles,
                                                   int ntiles, int ntiles_pad, int out_buf,
                                                   int with_rhs)
{
    TileCtx tc;
    if (!tile_setup(tiles, ntiles, ntiles_pad, cs, tc)) return;
    const CaseDev &c = cs[tc.cid];
    if (!c.ctrl->active) return;
    const int i = 1 + tc.tx * kTileX + (threadIdx.x & 63);
    const int nx = c.nx, ny = c.ny, P = c.P;
    if (i > nx - 2) return;
    const int base = c.ctrl->ad_base;
    const TS *u = fld<TS>(c.u[base]), *v = fld<TS>(c.v[base]);
    double *uf = c.uf[out_buf], *vf = c.vf[out_buf];
    const AxisD &ax = c.x, &ay = c.y;
    const double dxi = ax.d[i], dxp = ax.d[i + 1], dxm = ax.d[i - 1];
    for (int r = threadIdx.x >> 6; r < c.tile_y; r += 4) {
        const int j = 1 + tc.ty * c.tile_y + r;
        if (j > ny - 2) break;
        const int id = i + j * P;
        const uint8_t f = c.flags[id];
        const double uc = (double)u[id], vc = (double)v[id];
        double ufe, ufw, vfn, vfs;
        if (i == nx - 2) ufe = c.ubc;
        else if (f & F_OPEN_E) ufe = ax.ip[i] * ((double)u[id + 1] * dxi + uc * dxp);
        else ufe = blank_vel(c, id, id + 1, 0);
        if (i == 1) ufw = c.ubc;
        else if (f & F_OPEN_W) ufw = ax.ip[i - 1] * (uc * dxm + (double)u[id - 1] * dxi);
        else ufw = blank_vel(c, id - 1, id, 0);
        const double dyj = ay.d[j], dyp = ay.d[j + 1], dym = ay.d[j - 1];
        if (j == ny - 2) vfn = c.vbc;
        else if (f & F_OPEN_N) vfn = ay.ip[j] * ((double)v[id + P] * dyj + vc * dyp);
        else vfn = blank_vel(c, id, id + P, 1);
        if (j == 1) vfs = c.vbc;
        else if (f & F_OPEN_S) vfs = ay.ip[j - 1] * (vc * dym + (double)v[id - P] * dyj);
        else vfs = blank_vel(c, id - P, id, 1);
        uf[id] = ufe;
        vf[id] = vfn;
        if (i == 1) uf[id - 1] = ufw;
        if (j == 1) vf[id - P] = vfs;
        if (with_rhs)
            c.rhs[id] = (f & F_FLUID) ? ((ufe - ufw) * ax.rd[i] + (vfn - vfs) * ay.rd[j]) * c.rdt
                                      : 0.0;
    }
}

// ---------------------------------------------------------------------------
// P4: pressure Poisson.  Coefficients of calculatePPECoefficients
// (PPESolver.cu:93-99) on open faces, zero across blanked and domain faces.
// ---------------------------------------------------------------------------
template <typename T>
struct PpeAxis {
    const T *cp, *cm;
};

// One red-black half-sweep (colour h&1) on the LDS tile, over the region
// that is still valid after h half-sweeps: the tile plus a ring of
// R = 2M-1-h cells.  All sizes are compile-time constants.
template <typename T, int TY, int M, int h>
__device__ __forceinline__ void rb_half_sweep(T (*sp)[kTileX + 4 * M], const T (*sr)[kTileX + 4 * M],
                                              const uint8_t (*sf)[kTileX + 4 * M],
                                              const T *mxp, const T *mxm, const T *myp, const T *mym,
                                              int gpar, int cpar, T om, int uni, T ax, T ratio)
{
    constexpr int H = 2 * M;
    constexpr int R = H - 1 - h;
    constexpr int colour = h & 1;
    constexpr int rows = TY + 2 * R;
    constexpr int wid = kTileX + 2 * R;
    constexpr int pairs = (wid + 1) / 2;
    constexpr int items = rows * pairs;
#pragma unroll 2
    for (int k = threadIdx.x; k < items; k += 256) {
        const int row = k / pairs, pr = k - row * pairs;
        const int ly = H - R + row;
        int lx = H - R + 2 * pr;
        // global parity of (lx, ly) is (gpar + lx + ly) & 1
        lx += ((gpar + lx + ly) ^ colour) & 1;
        if (lx >= H + kTileX + R) continue;
        const uint8_t f = sf[ly][lx];
        if (!(f & F_FLUID)) continue;
        const T aE = (f & F_OPEN_E) ? mxp[lx] : (T)0;
        const T aW = (f & F_OPEN_W) ? mxm[lx] : (T)0;
        const T aN = (f & F_OPEN_N) ? myp[ly] : (T)0;
        const T aS = (f & F_OPEN_S) ? mym[ly] : (T)0;
        const T sa = (aE + aW) + (aN + aS);
        if (sa == (T)0) continue;
        if constexpr (sizeof(T) == 4) {
            if (uni) {
                sp[ly][lx] = sor_cell_u_sa(sp[ly][lx], sr[ly][lx], sp[ly][lx + 1], sp[ly][lx - 1], sp[ly + 1][lx],
                                           sp[ly - 1][lx], sa, om, ax, ratio);
                continue;
            }
        }
        sp[ly][lx] = sor_cell<T>(sp[ly][lx], sr[ly][lx], aE, sp[ly][lx + 1], aW, sp[ly][lx - 1], aN,
                                 sp[ly + 1][lx], aS, sp[ly - 1][lx], sa, om, cpar + lx);
    }
}

template <typename T, int TY, int M, int h>
__device__ __forceinline__ void rb_half_sweeps(T (*sp)[kTileX + 4 * M], const T (*sr)[kTileX + 4 * M],
                                               const uint8_t (*sf)[kTileX + 4 * M], const T *mxp,
                                               const T *mxm, const T *myp, const T *mym, int gpar, int cpar,
                                               T om, int nst, int uni, T ax, T ratio)
{
    if constexpr (h < 2 * M) {
        if (h < nst) {
            rb_half_sweep<T, TY, M, h>(sp, sr, sf, mxp, mxm, myp, mym, gpar, cpar, om, uni, ax, ratio);
            __syncthreads();
        }
        rb_half_sweeps<T, TY, M, h + 1>(sp, sr, sf, mxp, mxm, myp, mym, gpar, cpar, om, nst, uni, ax, ratio);
    }
}

// Red-black SOR, M full sweeps per launch on an LDS tile with a 2M halo
// (temporal blocking).  Reads the unknown from buffer src, writes dst: the
// ping-pong keeps tiles from reading halos that other tiles are writing,
// and the result equals M global red-black sweeps.
// T = double: unknown p, right side rhs.  T = float (mixed precision):
// unknown e (error), right side r32; the first launch of a chunk starts
// from e = 0.  Tiles are loaded and stored with aligned vector accesses
// (the field layout puts interior column 1 on a 256-byte boundary and the
// halo width is even).
template <typename T, int TY, int M>
__global__ __launch_bounds__(256) void k_rbsor(const CaseDev *cs, const int2 *tiles,
                                               int ntiles, int ntiles_pad, int s, int K)
{
    constexpr int H = 2 * M;
    constexpr int W = kTileX + 2 * H, HT = TY + 2 * H;
    constexpr int V = (sizeof(T) == 4 && (H % 4) == 0) ? 4 : 2;  // elements per vector access
    constexpr int WV = W / V;
    using VT = typename VecT<T, V>::type;
    using VB = typename VecB<V>::type;
    __shared__ T sp[HT][W];
    __shared__ T sr[HT][W];
    __shared__ uint8_t sf[HT][W];
    __shared__ T mxp[W], mxm[W], myp[HT], mym[HT];

    TileCtx tc;
    if (!tile_setup(tiles, ntiles, ntiles_pad, cs, tc)) return;
    const CaseDev &c = cs[tc.cid];
    const Ctrl *ct = c.ctrl;
    if (c.ppe_solver != IF_PPE_RBSOR || !ct->active || ct->ppe_done) return;
    int kexec = c.ppe_itermax - ct->ppe_sweeps;
    if (kexec > K) kexec = K;
    if (s >= kexec) return;
    int meff = kexec - s;
    if (meff > M) meff = M;
    const int L = s / M;  // launch index inside the chunk

    const int nx = c.nx, ny = c.ny, P = c.P;
    const T *src;
    T *dst;
    const T *rhs;
    const T *cxp, *cxm, *cyp, *cym;
    bool zero_src = false;
    T om;
    if constexpr (sizeof(T) == 8) {
        src = c.p[(ct->ppe_base + L) & 1];
        dst = c.p[(ct->ppe_base + L + 1) & 1];
        rhs = c.rhs;
        cxp = c.x.cp; cxm = c.x.cm; cyp = c.y.cp; cym = c.y.cm;
        om = c.omega;
    } else {
        src = c.e[L & 1];
        dst = c.e[(L + 1) & 1];
        zero_src = (L == 0);
        rhs = c.r32;
        cxp = c.xf.cp; cxm = c.xf.cm; cyp = c.yf.cp; cym = c.yf.cm;
        om = c.omega_f;
    }

    const int i0 = 1 + tc.tx * kTileX, j0 = 1 + tc.ty * TY;
    // load phase: unknown, right side and flags with halo H (vector loads
    // where the chunk lies inside the array / the interior)
#pragma unroll 2
    for (int k = threadIdx.x; k < HT * WV; k += 256) {
        const int ly = k / WV, cx = k - ly * WV;
        const int gj = j0 - H + ly, gi = i0 - H + cx * V;
        T *pd = &sp[ly][cx * V];
        T *rd = &sr[ly][cx * V];
        uint8_t *fd = &sf[ly][cx * V];
        const bool rowin = gj >= 0 && gj < ny;
        const bool rowint = gj >= 1 && gj <= ny - 2;
        const size_t o = (size_t)gj * P + gi;
        if (!zero_src && rowin && gi >= 0 && gi + V - 1 < nx) {
            const VT v = gload_vec<VT>(src + o);
#pragma unroll
            for (int q = 0; q < V; q++) pd[q] = v[q];
        } else {
#pragma unroll
            for (int q = 0; q < V; q++)
                pd[q] = (!zero_src && rowin && gi + q >= 0 && gi + q < nx) ? G(src)[o + q] : (T)0;
        }
        if (rowint && gi >= 1 && gi + V - 1 <= nx - 2) {
            const VT v = gload_vec<VT>(rhs + o);
            const VB fb = gload_vec<VB>(c.flags + o);
#pragma unroll
            for (int q = 0; q < V; q++) {
                rd[q] = v[q];
                fd[q] = fb[q];
            }
        } else {
#pragma unroll
            for (int q = 0; q < V; q++) {
                const bool in = rowint && gi + q >= 1 && gi + q <= nx - 2;
                rd[q] = in ? G(rhs)[o + q] : (T)0;
                fd[q] = in ? G(c.flags)[o + q] : (uint8_t)0;
            }
        }
    }
    if (threadIdx.x < W) {
        const int gi = i0 - H + threadIdx.x;
        const bool in = gi >= 1 && gi <= nx - 2;
        mxp[threadIdx.x] = in ? G(cxp)[gi] : (T)0;
        mxm[threadIdx.x] = in ? G(cxm)[gi] : (T)0;
    } else if (threadIdx.x >= 128 && threadIdx.x < 128 + HT) {
        const int q = threadIdx.x - 128;
        const int gj = j0 - H + q;
        const bool in = gj >= 1 && gj <= ny - 2;
        myp[q] = in ? G(cyp)[gj] : (T)0;
        mym[q] = in ? G(cym)[gj] : (T)0;
    }
    __syncthreads();

    // 2*meff half-sweeps on shrinking regions
    // parity of the tile's first cell, and of its first column (LDS column
    // lx is global column i0 - H + lx)
    const int gpar = (i0 - H) + (j0 - H);
    T uax = (T)0, uratio = (T)0;
    if constexpr (sizeof(T) == 4) {
        uax = c.uax;
        uratio = c.uratio;
    }
    rb_half_sweeps<T, TY, M, 0>(sp, sr, sf, mxp, mxm, myp, mym, gpar & 1, (i0 - H) & 1, om, 2 * meff,
                                sizeof(T) == 4 && c.uni, uax, uratio);

    // write the tile (all interior cells: non-fluid values pass through)
    constexpr int VW = 16 / sizeof(T);  // 16-byte stores; i0 is 256-byte aligned
    constexpr int RW = kTileX / VW;
    using VS = typename VecT<T, VW>::type;
#pragma unroll
    for (int k = threadIdx.x; k < RW * TY; k += 256) {
        const int ly = k / RW, cx = k - ly * RW;
        const int gi = i0 + cx * VW, gj = j0 + ly;
        if (gj > ny - 2) continue;
        const T *sv = &sp[H + ly][H + cx * VW];
        if (gi + VW - 1 <= nx - 2) {
            VS v;
#pragma unroll
            for (int q = 0; q < VW; q++) v[q] = sv[q];
            gstore_vec<VS>(dst + (size_t)gj * P + gi, v);
        } else {
#pragma unroll
            for (int q = 0; q < VW; q++)
                if (gi + q <= nx - 2) GW(dst)[(size_t)gj * P + gi + q] = sv[q];
        }
    }
}

// Weighted point Jacobi (jacobiIteration, PPESolver.cu:13-31, with the RHS
// term and the relaxation weight; weight 1 is the reference's iteration).
template <typename T>
__global__ __launch_bounds__(256) void k_jacobi(const CaseDev *cs, const int2 *tiles,
                                                int ntiles, int ntiles_pad, int s, int K)
{
    TileCtx tc;
    if (!tile_setup(tiles, ntiles, ntiles_pad, cs, tc)) return;
    const CaseDev &c = cs[tc.cid];
    const Ctrl *ct = c.ctrl;
    if (c.ppe_solver != IF_PPE_JACOBI || !ct->active || ct->ppe_done) return;
    int kexec = c.ppe_itermax - ct->ppe_sweeps;
    if (kexec > K) kexec = K;
    if (s >= kexec) return;
    const int nx = c.nx, ny = c.ny, P = c.P;
    const int i = 1 + tc.tx * kTileX + (threadIdx.x & 63);
    if (i > nx - 2) return;
    const T *src;
    T *dst;
    const T *rhs;
    PpeAxis<T> ax, ay;
    if constexpr (sizeof(T) == 8) {
        src = c.p[(ct->ppe_base + s) & 1];
        dst = c.p[(ct->ppe_base + s + 1) & 1];
        rhs = c.rhs;
        ax = {c.x.cp, c.x.cm};
        ay = {c.y.cp, c.y.cm};
    } else {
        src = c.e[s & 1];
        dst = c.e[(s + 1) & 1];
        rhs = c.r32;
        ax = {c.xf.cp, c.xf.cm};
        ay = {c.yf.cp, c.yf.cm};
    }
    const bool zero_src = (sizeof(T) == 4) && s == 0;
    const T om = (sizeof(T) == 8) ? (T)c.omega : (T)c.omega_f;
    for (int r = threadIdx.x >> 6; r < c.tile_y; r += 4) {
        const int j = 1 + tc.ty * c.tile_y + r;
        if (j > ny - 2) break;
        const int id = i + j * P;
        const uint8_t f = c.flags[id];
        T val = zero_src ? (T)0 : src[id];
        if (f & F_FLUID) {
            const T aE = (f & F_OPEN_E) ? ax.cp[i] : (T)0;
            const T aW = (f & F_OPEN_W) ? ax.cm[i] : (T)0;
            const T aN = (f & F_OPEN_N) ? ay.cp[j] : (T)0;
            const T aS = (f & F_OPEN_S) ? ay.cm[j] : (T)0;
            const T ap = -((aE + aW) + (aN + aS));
            if (ap != (T)0) {
                T sum = (T)0;
                if (!zero_src)
                    sum = (aE * src[id + 1] + aW * src[id - 1]) + (aN * src[id + P] + aS * src[id - P]);
                const T gs = (rhs[id] - sum) / ap;
                val = val + om * (gs - val);
            }
        }
        dst[id] = val;
    }
}

// Zebra line SOR (the deck's PPE_Solver 2, inputs/inputs.txt:14-15), one
// colour of rows per launch: rows j odd (colour 1) first, then rows j even.
// Each row's tridiagonal system
//     -aW x[i-1] + sa x[i] - aE x[i+1] = (aN pN + aS pS) - rhs
// (fluid cells with sa != 0; identity rows elsewhere) is solved and
// relaxed, p + om (x - p).  FP64 is the oracle's arithmetic (line_sor_sweep,
// oracle.c), to the bit.
//
// One wave per row, one lane per segment of the row's partition
// (line_partition: S <= 64 blocks of B cells, the last cell of each block
// but the last a separator): a lane runs its segment's Thomas recurrences,
// the separators' Schur complement is solved by cyclic reduction across
// the lanes (shuffles), and each segment adds its separators' share.  The
// factors depend on the flags and the metric only: k_line_factor forms
// them once per step, in the segment layout [plane][t][lane] that a
// lane-per-segment loop reads coalesced.  The sweep kernel stages the
// row's right side into LDS in that layout (coalesced row loads, a
// transposing store), runs the lanes' recurrences there, and relaxes and
// stores the row with coalesced accesses.
template <typename T> __device__ __forceinline__ T fmaT(T a, T b, T c)
{
    if constexpr (sizeof(T) == 8)
        return __builtin_fma(a, b, c);
    else
        return __builtin_fmaf(a, b, c);
}

enum { kLfR, kLfNlr, kLfNcp, kLfH, kLfQ };  // per-cell factor planes (kLinePlanes)

template <typename T> struct LineMetric {
    const T *xp, *xm, *yp, *ym;
    __device__ LineMetric(const CaseDev &c)
    {
        if constexpr (sizeof(T) == 8) {
            xp = c.x.cp; xm = c.x.cm; yp = c.y.cp; ym = c.y.cm;
        } else {
            xp = c.xf.cp; xm = c.xf.cm; yp = c.yf.cp; ym = c.yf.cm;
        }
    }
    // row coefficients of a cell (x metric xpi, xmi of its column): l = -aW,
    // b = sa, u = -aE when solved
    __device__ static bool cell(unsigned f, T xpi, T xmi, T yN, T yS, T &l, T &b, T &u)
    {
        l = (T)0;
        b = (T)1;
        u = (T)0;
        if (!(f & F_FLUID)) return false;
        const T aE = (f & F_OPEN_E) ? xpi : (T)0;
        const T aW = (f & F_OPEN_W) ? xmi : (T)0;
        const T aN = (f & F_OPEN_N) ? yN : (T)0;
        const T aS = (f & F_OPEN_S) ? yS : (T)0;
        const T sa = (aE + aW) + (aN + aS);
        if (sa == (T)0) return false;
        l = -aW;
        b = sa;
        u = -aE;
        return true;
    }
};

template <typename T> __device__ __forceinline__ T shfl_up_t(T v, int d) { return __shfl_up(v, (unsigned)d, 64); }
template <typename T> __device__ __forceinline__ T shfl_down_t(T v, int d) { return __shfl_down(v, (unsigned)d, 64); }

// The factors of every interior row of every line-solve case (once per
// step): lane k forms segment k's pivots r = 1/den, -l r, -u r and spikes
// h, q; lanes 0..m-1 (m = S - 1) the separators' -l, -u and the cyclic
// reduction's multipliers and final reciprocals.  Operation for operation
// the oracle's line_factor_row.
template <typename T>
__global__ __launch_bounds__(64) void k_line_factor(const CaseDev *cs)
{
    const CaseDev &c = cs[blockIdx.y];
    if (c.ppe_solver != IF_PPE_LINE || !c.ctrl->active) return;
    const int j = 1 + (int)blockIdx.x;
    if (j > c.ny - 2) return;
    const int B = c.lB, S = c.lS, n = c.nx - 2, m = S - 1, k = (int)threadIdx.x;
    const size_t row = (size_t)j * c.P;
    const LineMetric<T> mt(c);
    const T yN = G(mt.yp)[j], yS = G(mt.ym)[j];
    const auto fl = G(c.flags) + row;
    const auto F = GW(fld<T>(c.lf) + (size_t)j * line_row_stride(B));
    auto at = [&](int pl, int t) { return ((size_t)pl * B + t) * 64 + k; };
    const int len = k < S - 1 ? B - 1 : (k == S - 1 ? n - (S - 1) * B : 0);
    T hlast = (T)0, qlast = (T)0, h0 = (T)0, q0 = (T)0;
    if (len > 0) {
        const int a = 1 + k * B;
        T cpv = (T)0, hpv = (T)0;
        for (int t = 0; t < len; t++) {
            T l, b, u;
            mt.cell(fl[a + t], G(mt.xp)[a + t], G(mt.xm)[a + t], yN, yS, l, b, u);
            const T den = t == 0 ? b : b - l * cpv;
            const T r = (T)1 / den, cp = u * r, nlr = -(l * r);
            const T hp = t == 0 ? nlr : nlr * hpv;
            F[at(kLfR, t)] = r;
            F[at(kLfNlr, t)] = nlr;
            F[at(kLfNcp, t)] = -cp;
            cpv = cp;
            hpv = hp;
            if (t == len - 1) {
                hlast = hp;
                qlast = -cp;
            } else {
                F[at(kLfH, t)] = hp;
            }
        }
        // backward: h = fma(-u r, h', hp), q = (-u r) q' (this lane's stores
        // above are read back by this lane only)
        F[at(kLfH, len - 1)] = hlast;
        F[at(kLfQ, len - 1)] = qlast;
        T hx = hlast, qx = qlast;
        for (int t = len - 2; t >= 0; t--) {
            const T ncp = F[at(kLfNcp, t)];
            hx = fmaT(ncp, hx, F[at(kLfH, t)]);
            qx = ncp * qx;
            F[at(kLfH, t)] = hx;
            F[at(kLfQ, t)] = qx;
        }
        h0 = hx;
        q0 = qx;
    }
    // separators: the Schur complement's rows A x[k-1] + Bv x[k] + Cv x[k+1]
    const T hR = shfl_down_t(h0, 1), qR = shfl_down_t(q0, 1);
    T A = (T)0, Bv = (T)1, Cv = (T)0;
    if (k < m) {
        const int sp = (k + 1) * B;
        T l, b, u;
        mt.cell(fl[sp], G(mt.xp)[sp], G(mt.xm)[sp], yN, yS, l, b, u);
        A = l * hlast;
        Cv = u * qR;
        Bv = (b + l * qlast) + u * hR;
        F[at(kLfR, B - 1)] = -l;
        F[at(kLfNlr, B - 1)] = -u;
    }
    // cyclic reduction with strides 1, 2, 4, ... < m
    const auto pcr = F + (size_t)kLinePlanes * B * 64;
    int lev = 0;
    for (int sg = 1; sg < m; sg *= 2, lev++) {
        const T Am_ = shfl_up_t(A, sg), Cm_ = shfl_up_t(Cv, sg), Bm_ = shfl_up_t(Bv, sg);
        const T Ap_ = shfl_down_t(A, sg), Cp_ = shfl_down_t(Cv, sg), Bp_ = shfl_down_t(Bv, sg);
        const bool lo = k >= sg, hi = k + sg < m;
        T alpha = (T)0, gamma = (T)0;
        if (k < m) {
            alpha = lo ? -(A / Bm_) : (T)0;
            gamma = hi ? -(Cv / Bp_) : (T)0;
            const T Am = lo ? Am_ : (T)0, Cm = lo ? Cm_ : (T)0;
            const T Ap = hi ? Ap_ : (T)0, Cp = hi ? Cp_ : (T)0;
            const T nA = alpha * Am, nC = gamma * Cp, nB = (Bv + alpha * Cm) + gamma * Ap;
            A = nA;
            Cv = nC;
            Bv = nB;
        }
        pcr[(2 * lev) * 64 + k] = alpha;
        pcr[(2 * lev + 1) * 64 + k] = gamma;
    }
    pcr[(2 * kLinePcrLevels) * 64 + k] = k < m ? (T)1 / Bv : (T)0;
}

// One colour of one sweep: one block of kLineWaves waves per row, LDS of
// B x 65 elements (the row's values in the segment layout, padded so that
// the transposing store and the lanes' column walks are conflict-free).
// A colour has few rows (a block or two per CU), so the kernel is bound by
// the latency of its dependent steps: the row passes (all waves) load
// kLineU cells per thread at once; wave 0 runs the segments, with its
// factor loads a group of 8 cells ahead of the recurrences (two register
// buffers) and the first groups of the later passes and the reduction's
// multipliers loaded at the start.
constexpr int kLineU = 8, kLineWaves = 4;
template <typename T>
__global__ __launch_bounds__(64 * kLineWaves) void k_line_sor(const CaseDev *cs, int s, int K, int colour)
{
    const CaseDev &c = cs[blockIdx.y];
    const Ctrl *ct = c.ctrl;
    if (c.ppe_solver != IF_PPE_LINE || !ct->active || ct->ppe_done) return;
    int kexec = c.ppe_itermax - ct->ppe_sweeps;
    if (kexec > K) kexec = K;
    if (s >= kexec) return;
    const int nx = c.nx, ny = c.ny, P = c.P;
    const int j = 2 - colour + 2 * (int)blockIdx.x;
    if (j > ny - 2) return;
    const T *src;
    T *dst;
    const T *rhs;
    bool zero_src = false;
    T om;
    if constexpr (sizeof(T) == 8) {
        src = c.p[(ct->ppe_base + s) & 1];
        dst = c.p[(ct->ppe_base + s + 1) & 1];
        rhs = c.rhs;
        om = c.omega;
    } else {
        src = c.e[s & 1];
        dst = c.e[(s + 1) & 1];
        zero_src = (s == 0);
        rhs = c.r32;
        om = c.omega_f;
    }
    const bool zero_nbr = colour == 1 && zero_src;
    // rows above and below: the old values (colour 1) or this sweep's (colour 0)
    const auto gsrc = G(src), gnbr = G(colour == 1 ? src : dst), grhs = G(rhs);
    const auto gdst = GW(dst);
    const LineMetric<T> mt(c);
    const auto gxp = G(mt.xp), gxm = G(mt.xm);
    const T yN = G(mt.yp)[j], yS = G(mt.ym)[j];
    const auto gfl = G(c.flags);
    const int B = c.lB, S = c.lS, n = nx - 2, m = S - 1;
    const int tid = (int)threadIdx.x, lane = tid & 63;
    const bool w0 = tid < 64;  // the wave that runs the segments
    const int len = lane < S - 1 ? B - 1 : (lane == S - 1 ? n - (S - 1) * B : 0);
    const auto F = G((const T *)fld<T>(c.lf) + (size_t)j * line_row_stride(B));
    auto fat = [&](int pl, int t) { return F[((size_t)pl * B + t) * 64 + lane]; };
    const auto pcr = F + (size_t)kLinePlanes * B * 64;
    extern __shared__ __align__(16) unsigned char line_lds[];
    T *g = (T *)line_lds;  // [t * 65 + segment]
    // segment and position of cell i: (i - 1) / B by a multiply-high
    // (exact for i - 1 < 2^32 / B), the remainder by subtraction
    const unsigned mg = 0xffffffffu / (unsigned)B + 1u;
    auto slot = [&](int i) {
        const unsigned x = (unsigned)(i - 1), kk = __umulhi(x, mg);
        return (int)((x - kk * (unsigned)B) * 65u + kk);
    };
    const int tl = len > 0 ? len - 1 : 0;  // clamp for the lanes' loads

    // data-independent loads, issued first: the reduction's multipliers,
    // the separator's -l, -u and the first groups of the backward and
    // final passes
    T pal[kLinePcrLevels], pga[kLinePcrLevels], prb = (T)0, snl = (T)0, snu = (T)0;
    T bc0[8], fh0[8], fq0[8];
    T fra[8], fna[8], frb[8], fnb[8];
    auto ld_fwd = [&](T(&fr)[8], T(&fn)[8], int t0) {
#pragma unroll
        for (int q = 0; q < 8; q++) {
            const int t = min(t0 + q, tl);
            fr[q] = fat(kLfR, t);
            fn[q] = fat(kLfNlr, t);
        }
    };
    if (w0) {
#pragma unroll
        for (int q = 0; q < kLinePcrLevels; q++) {
            pal[q] = pcr[(2 * q) * 64 + lane];
            pga[q] = pcr[(2 * q + 1) * 64 + lane];
        }
        prb = pcr[(2 * kLinePcrLevels) * 64 + lane];
        snl = fat(kLfR, B - 1);
        snu = fat(kLfNlr, B - 1);
#pragma unroll
        for (int q = 0; q < 8; q++) {
            bc0[q] = fat(kLfNcp, max(len - 2 - q, 0));
            fh0[q] = fat(kLfH, min(q, tl));
            fq0[q] = fat(kLfQ, min(q, tl));
        }
        ld_fwd(fra, fna, 0);
    }

    // ---- the right side d, staged in the segment layout (all waves) -----
    const size_t row = (size_t)j * P;
    constexpr int kT = 64 * kLineWaves;
    for (int i0 = 1; i0 <= n; i0 += kT * kLineU) {
        unsigned fv[kLineU];
        T pv[kLineU], vn[kLineU], vs[kLineU], rh[kLineU], xa[kLineU], xb[kLineU];
#pragma unroll
        for (int q = 0; q < kLineU; q++) {
            const int i = min(i0 + kT * q + tid, n);
            const size_t id = row + i;
            fv[q] = gfl[id];
            pv[q] = gsrc[id];
            vn[q] = gnbr[id + P];
            vs[q] = gnbr[id - P];
            rh[q] = grhs[id];
            xa[q] = gxp[i];
            xb[q] = gxm[i];
        }
#pragma unroll
        for (int q = 0; q < kLineU; q++) {
            const int i = i0 + kT * q + tid;
            if (i <= n) {
                const unsigned f = fv[q];
                T d = zero_src ? (T)0 : pv[q];
                T l, b, u;
                if (LineMetric<T>::cell(f, xa[q], xb[q], yN, yS, l, b, u)) {
                    const T aN = (f & F_OPEN_N) ? yN : (T)0, aS = (f & F_OPEN_S) ? yS : (T)0;
                    const T un = zero_nbr ? (T)0 : vn[q], us = zero_nbr ? (T)0 : vs[q];
                    d = (aN * un + aS * us) - rh[q];
                }
                g[slot(i)] = d;
            }
        }
    }
    __syncthreads();

    if (w0) {
        // ---- segments: gp = fma(-l r, gp, d r) forward ------------------------
        T gp = (T)0;
        auto fwd = [&](const T(&fr)[8], const T(&fn)[8], int t0) {
            T dv[8];
#pragma unroll
            for (int q = 0; q < 8; q++) dv[q] = g[min(t0 + q, tl) * 65 + lane];
#pragma unroll
            for (int q = 0; q < 8; q++) {
                const int t = t0 + q;
                if (t < len) {
                    const T dr = dv[q] * fr[q];
                    gp = t == 0 ? dr : fmaT(fn[q], gp, dr);
                    g[t * 65 + lane] = gp;
                }
            }
        };
        for (int t0 = 0; t0 < len; t0 += 16) {
            if (t0 + 8 < len) ld_fwd(frb, fnb, t0 + 8);
            fwd(fra, fna, t0);
            if (t0 + 8 >= len) break;
            if (t0 + 16 < len) ld_fwd(fra, fna, t0 + 16);
            fwd(frb, fnb, t0 + 8);
        }
        const T glast = gp;
        // ---- backward g = fma(-u r, g, gp) ------------------------------------
        T gx = gp;
        auto ld_bwd = [&](T(&fc)[8], int t0) {
#pragma unroll
            for (int q = 0; q < 8; q++) fc[q] = fat(kLfNcp, max(t0 - q, 0));
        };
        auto bwd = [&](const T(&fc)[8], int t0) {
            T gv[8];
#pragma unroll
            for (int q = 0; q < 8; q++) gv[q] = g[max(t0 - q, 0) * 65 + lane];
#pragma unroll
            for (int q = 0; q < 8; q++) {
                const int t = t0 - q;
                if (t >= 0) {
                    gx = fmaT(fc[q], gx, gv[q]);
                    g[t * 65 + lane] = gx;
                }
            }
        };
        // fra/frb are free now: the backward pass's second buffer is fra
        for (int t0 = len - 2; t0 >= 0; t0 -= 16) {
            if (t0 - 8 >= 0) ld_bwd(fra, t0 - 8);
            bwd(bc0, t0);
            if (t0 - 8 < 0) break;
            if (t0 - 16 >= 0) ld_bwd(bc0, t0 - 16);
            bwd(fra, t0 - 8);
        }
        // ---- separators: D = fma(-u, g right, fma(-l, g left, d)); cyclic
        // reduction D = fma(gamma, D[k + s], fma(alpha, D[k - s], D)) with 0
        // outside 0..m-1; x = D times the reciprocal of the reduced diagonal
        const T gR = shfl_down_t(gx, 1);
        T D = (T)0;
        if (lane < m) D = fmaT(snu, gR, fmaT(snl, glast, g[(B - 1) * 65 + lane]));
#pragma unroll
        for (int q = 0; q < kLinePcrLevels; q++) {
            const int sg = 1 << q;
            if (sg >= m) break;
            const T Dm_ = shfl_up_t(D, sg), Dp_ = shfl_down_t(D, sg);
            const T Dm = lane >= sg ? Dm_ : (T)0, Dp = lane + sg < m ? Dp_ : (T)0;
            if (lane < m) D = fmaT(pga[q], Dp, fmaT(pal[q], Dm, D));
        }
        const T xs = lane < m ? D * prb : (T)0;
        const T xl_ = shfl_up_t(xs, 1);
        const T xL = lane >= 1 ? xl_ : (T)0, xR = lane < m ? xs : (T)0;
        // ---- x = fma(q, xR, fma(h, xL, g)); separators take xs ---------------
        auto ld_fin = [&](T(&fh)[8], T(&fq)[8], int t0) {
#pragma unroll
            for (int q = 0; q < 8; q++) {
                const int t = min(t0 + q, tl);
                fh[q] = fat(kLfH, t);
                fq[q] = fat(kLfQ, t);
            }
        };
        auto fin = [&](const T(&fh)[8], const T(&fq)[8], int t0) {
            T gv[8];
#pragma unroll
            for (int q = 0; q < 8; q++) gv[q] = g[min(t0 + q, tl) * 65 + lane];
#pragma unroll
            for (int q = 0; q < 8; q++)
                if (t0 + q < len) g[(t0 + q) * 65 + lane] = fmaT(fq[q], xR, fmaT(fh[q], xL, gv[q]));
        };
        for (int t0 = 0; t0 < len; t0 += 16) {
            if (t0 + 8 < len) ld_fin(frb, fnb, t0 + 8);
            fin(fh0, fq0, t0);
            if (t0 + 8 >= len) break;
            if (t0 + 16 < len) ld_fin(fh0, fq0, t0 + 16);
            fin(frb, fnb, t0 + 8);
        }
        if (lane < m) g[(B - 1) * 65 + lane] = xs;
    }  // wave 0
    __syncthreads();

    // ---- relax the solved cells, store the row (all waves) --------------
    for (int i0 = 1; i0 <= n; i0 += kT * kLineU) {
        unsigned fv[kLineU];
        T pv[kLineU], xa[kLineU], xb[kLineU];
#pragma unroll
        for (int q = 0; q < kLineU; q++) {
            const int i = min(i0 + kT * q + tid, n);
            fv[q] = gfl[row + i];
            pv[q] = gsrc[row + i];
            xa[q] = gxp[i];
            xb[q] = gxm[i];
        }
#pragma unroll
        for (int q = 0; q < kLineU; q++) {
            const int i = i0 + kT * q + tid;
            if (i <= n) {
                const T pold = zero_src ? (T)0 : pv[q];
                T val = pold;
                T l, b, u;
                if (LineMetric<T>::cell(fv[q], xa[q], xb[q], yN, yS, l, b, u)) val = pold + om * (g[slot(i)] - pold);
                gdst[row + i] = val;
            }
        }
    }
}

void launch_line_factor(hipStream_t st, int prec, const CaseDev *cs, int ncase, int max_ny)
{
    const dim3 grid((unsigned)(max_ny - 2), (unsigned)ncase);
    if (prec == IF_FP64)
        hipLaunchKernelGGL(k_line_factor<double>, grid, dim3(64), 0, st, cs);
    else
        hipLaunchKernelGGL(k_line_factor<float>, grid, dim3(64), 0, st, cs);
}

void launch_line_sor(hipStream_t st, int prec, const CaseDev *cs, int ncase, int max_n, int s, int K)
{
    // max_n bounds every case's row length and row count: the LDS holds the
    // largest partition among line cases (at most kLineMaxB, checked at
    // case creation), and each colour has at most (max_n - 1) / 2 rows
    int B, S;
    line_partition(max_n - 2, &B, &S);
    B = std::min(B, kLineMaxB);
    const size_t lds = (size_t)B * 65 * (prec == IF_FP64 ? 8 : 4);
    const dim3 grid((unsigned)((max_n - 1) / 2), (unsigned)ncase);
    for (int colour = 1; colour >= 0; colour--) {
        if (prec == IF_FP64)
            hipLaunchKernelGGL(k_line_sor<double>, grid, dim3(64 * kLineWaves), lds, st, cs, s, K, colour);
        else
            hipLaunchKernelGGL(k_line_sor<float>, grid, dim3(64 * kLineWaves), lds, st, cs, s, K, colour);
    }
}

// FP64 residual r = rhs - A p (max norm per tile) after a chunk of K sweeps
// done in launches of M sweeps (K = 0: the check before any sweep).
__global__ __launch_bounds__(256) void k_ppe_resid(const CaseDev *cs, const int2 *tiles,
                                                   int ntiles, int ntiles_pad, int K)
{
    TileCtx tc;
    if (!tile_setup(tiles, ntiles, ntiles_pad, cs, tc)) return;
    const CaseDev &c = cs[tc.cid];
    const Ctrl *ct = c.ctrl;
    if (!ct->active || ct->ppe_done) return;
    int kexec = c.ppe_itermax - ct->ppe_sweeps;
    if (kexec > K) kexec = K;
    const int nl = (kexec + c.ppe_m - 1) / c.ppe_m;  // launches of the chunk
    const double *p = c.p[(ct->ppe_base + nl) & 1];
    const int nx = c.nx, ny = c.ny, P = c.P;
    const int i = 1 + tc.tx * kTileX + (threadIdx.x & 63);
    double m = 0.0;
    if (i <= nx - 2) {
        const double cpx = c.x.cp[i], cmx = c.x.cm[i];
        for (int r = threadIdx.x >> 6; r < c.tile_y; r += 4) {
            const int j = 1 + tc.ty * c.tile_y + r;
            if (j > ny - 2) break;
            const int id = i + j * P;
            const uint8_t f = c.flags[id];
            if (!(f & F_FLUID)) continue;
            const double aE = (f & F_OPEN_E) ? cpx : 0.0;
            const double aW = (f & F_OPEN_W) ? cmx : 0.0;
            const double aN = (f & F_OPEN_N) ? c.y.cp[j] : 0.0;
            const double aS = (f & F_OPEN_S) ? c.y.cm[j] : 0.0;
            const double ap = -((aE + aW) + (aN + aS));
            if (ap == 0.0) continue;
            const double sum = (aE * p[id + 1] + aW * p[id - 1]) + (aN * p[id + P] + aS * p[id - P]);
            const double rr = c.rhs[id] - (ap * p[id] + sum);
            m = fmax(m, absinf(rr));
        }
    }
    m = block_max(m);
    if (threadIdx.x == 0) c.part[tc.t] = m;
}

// Mixed precision: p_new = p + e (FP64) after a chunk of FP32 error sweeps,
// then the FP64 residual of p_new, stored as FP32 for the next chunk
// (defect correction / iterative refinement).  K = 0: residual only.
__global__ __launch_bounds__(256) void k_ppe_update_resid(const CaseDev *cs, const int2 *tiles,
                                                          int ntiles, int ntiles_pad, int K)
{
    TileCtx tc;
    if (!tile_setup(tiles, ntiles, ntiles_pad, cs, tc)) return;
    const CaseDev &c = cs[tc.cid];
    const Ctrl *ct = c.ctrl;
    if (!ct->active || ct->ppe_done) return;
    int kexec = c.ppe_itermax - ct->ppe_sweeps;
    if (kexec > K) kexec = K;
    const int nl = (kexec + c.ppe_m - 1) / c.ppe_m;  // launches of the chunk
    const bool upd = kexec > 0;
    const double *p = c.p[ct->ppe_base];
    double *pn = c.p[1 - ct->ppe_base];
    const float *e = c.e[nl & 1];
    const int nx = c.nx, ny = c.ny, P = c.P;
    const int i = 1 + tc.tx * kTileX + (threadIdx.x & 63);
    double m = 0.0;
    if (i <= nx - 2) {
        const double cpx = c.x.cp[i], cmx = c.x.cm[i];
        for (int r = threadIdx.x >> 6; r < c.tile_y; r += 4) {
            const int j = 1 + tc.ty * c.tile_y + r;
            if (j > ny - 2) break;
            const int id = i + j * P;
            const uint8_t f = c.flags[id];
            double pc = p[id], pE = p[id + 1], pW = p[id - 1], pN = p[id + P], pS = p[id - P];
            if (upd) {
                pc += (double)e[id];
                pE += (double)e[id + 1];
                pW += (double)e[id - 1];
                pN += (double)e[id + P];
                pS += (double)e[id - P];
                pn[id] = pc;
            }
            float r32 = 0.0f;
            if (f & F_FLUID) {
                const double aE = (f & F_OPEN_E) ? cpx : 0.0;
                const double aW = (f & F_OPEN_W) ? cmx : 0.0;
                const double aN = (f & F_OPEN_N) ? c.y.cp[j] : 0.0;
                const double aS = (f & F_OPEN_S) ? c.y.cm[j] : 0.0;
                const double ap = -((aE + aW) + (aN + aS));
                if (ap != 0.0) {
                    const double sum = (aE * pE + aW * pW) + (aN * pN + aS * pS);
                    const double rr = c.rhs[id] - (ap * pc + sum);
                    r32 = c.uni ? (float)rr * c.uinv : (float)rr;
                    m = fmax(m, absinf(rr));
                }
            }
            c.r32[id] = r32;
        }
    }
    m = block_max(m);
    if (threadIdx.x == 0) c.part[tc.t] = m;
}

// One block per case: the Poisson stopping rule on the device.
// nl_rule: 0 = FP64 ping-pong (base advances by the launch count),
//          1 = mixed (base flips once when an update happened).
// bar (persistent chunks only): once a neighbour wait of the persistent
// sweep has timed out, the loop ends here too (ppe_done, not failed); the
// host then solves again with one launch per stage.
__global__ __launch_bounds__(256) void k_ppe_finalize(const CaseDev *cs, int ncase, int K, int mixed,
                                                      const GridBar *bar)
{
    const CaseDev &c = cs[blockIdx.x];
    Ctrl *ct = c.ctrl;
    if (!ct->active || ct->ppe_done) return;
    const int nt = c.ntx * c.nty;
    // four independent loads in flight per thread (the maximum does not
    // depend on the order)
    const double *part = c.part;
    double m0 = 0.0, m1 = 0.0, m2 = 0.0, m3 = 0.0;
    int k = threadIdx.x;
    for (; k + 768 < nt; k += 1024) {
        m0 = fmax(m0, part[k]);
        m1 = fmax(m1, part[k + 256]);
        m2 = fmax(m2, part[k + 512]);
        m3 = fmax(m3, part[k + 768]);
    }
    for (; k < nt; k += 256) m0 = fmax(m0, part[k]);
    double m = block_max(fmax(fmax(m0, m1), fmax(m2, m3)));
    if (threadIdx.x == 0) {
        int kexec = c.ppe_itermax - ct->ppe_sweeps;
        if (kexec > K) kexec = K;
        const int nl = (kexec + c.ppe_m - 1) / c.ppe_m;  // launches of the chunk
        ct->ppe_sweeps += kexec;
        ct->tot_ppe += kexec;
        if (mixed) {
            if (kexec > 0) ct->ppe_base ^= 1;
        } else {
            ct->ppe_base = (ct->ppe_base + nl) & 1;
        }
        const double res = c.dt * m;
        ct->ppe_res = res;
        if (!(res <= 1.0e300)) ct->failed = 1;
        if (!(res > c.ppe_tol) || ct->ppe_sweeps >= c.ppe_itermax || ct->failed) ct->ppe_done = 1;
        if (bar && __hip_atomic_load(&bar->err, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != 0) ct->ppe_done = 1;
    }
}

// Re-open the Poisson loop of every running case (after a persistent chunk
// timed out): the solve starts again from the pressure the stopped chunk
// left, which is finite, and the iteration converges from any start.
__global__ void k_ppe_rearm(const CaseDev *cs, int ncase)
{
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= ncase) return;
    Ctrl *c = cs[k].ctrl;
    if (!c->active || c->failed) return;
    c->ppe_done = 0;
    c->ppe_sweeps = 0;
}

// The FP32 red-black SOR weight of every interior cell, om / sa over the
// open faces (sor_weight, the expression every FP32 sweep kernel uses), 0
// for a cell the sweep keeps (not fluid, or no open face).  Once per step,
// after the flags are final: the register-tile kernel loads it instead of
// forming it from the flags on every launch.  One block per row.  On a grid
// whose coefficients are one value per axis (wtab) the weight depends on
// the flags alone: the first block also writes K = w ax of each pattern of
// flag bits 0-5 (the five-operation form's weight), with the same
// expression, for launches whose cases all have such a table.
__global__ __launch_bounds__(256) void k_weights32(const CaseDev *cs)
{
    const CaseDev &c = cs[blockIdx.y];
    if (!c.w32 || !c.ctrl->active) return;  // FP64 Poisson cases hold no FP32 weights
    if (c.wtab && blockIdx.x == 0) {
        if (threadIdx.x < 64) {
            const unsigned f = threadIdx.x;
            const float om = c.omega_f;
            float w = 0.0f;
            if (f & F_FLUID) {
                const float aE = (f & F_OPEN_E) ? G(c.xf.cp)[1] : 0.0f;
                const float aW = (f & F_OPEN_W) ? G(c.xf.cm)[2] : 0.0f;
                const float aN = (f & F_OPEN_N) ? G(c.yf.cp)[1] : 0.0f;
                const float aS = (f & F_OPEN_S) ? G(c.yf.cm)[2] : 0.0f;
                const float sa = (aE + aW) + (aN + aS);
                if (sa != 0.0f) w = sor_weight(sa, om);
            }
            GW(c.wtab)[f] = w * c.uax;  // K = w ax of the five-operation form
        }
    }
    const int j = 1 + (int)blockIdx.x;
    if (j > c.ny - 2) return;
    const size_t row = (size_t)j * c.P;
    const float yp = G(c.yf.cp)[j], ym = G(c.yf.cm)[j], om = c.omega_f;
    for (int i = 1 + (int)threadIdx.x; i <= c.nx - 2; i += (int)blockDim.x) {
        const unsigned f = G(c.flags)[row + i];
        float w = 0.0f;
        if (f & F_FLUID) {
            const float aE = (f & F_OPEN_E) ? G(c.xf.cp)[i] : 0.0f;
            const float aW = (f & F_OPEN_W) ? G(c.xf.cm)[i] : 0.0f;
            const float aN = (f & F_OPEN_N) ? yp : 0.0f;
            const float aS = (f & F_OPEN_S) ? ym : 0.0f;
            const float sa = (aE + aW) + (aN + aS);
            if (sa != 0.0f) w = sor_weight(sa, om);
        }
        GW(c.w32)[row + i] = w;
    }
}

// The FP64 weight table of a case with one value per Poisson coefficient:
// the weight of each pattern of flag bits 0-5, with the expression and
// operand order of the FP64 sweep's per-launch weights (om / sa over the
// open faces, IEEE division; 0 for a cell the sweep keeps), so a looked-up
// weight has the same bits.  One block of 64 per case.
__global__ __launch_bounds__(64) void k_weight_table64(const CaseDev *cs)
{
    const CaseDev &c = cs[blockIdx.x];
    if (!c.wtab64 || !c.ctrl->active) return;
    const unsigned f = threadIdx.x;
    double w = 0.0;
    if (f & F_FLUID) {
        const double aE = (f & F_OPEN_E) ? G(c.x.cp)[1] : 0.0;
        const double aW = (f & F_OPEN_W) ? G(c.x.cm)[2] : 0.0;
        const double aN = (f & F_OPEN_N) ? G(c.y.cp)[1] : 0.0;
        const double aS = (f & F_OPEN_S) ? G(c.y.cm)[2] : 0.0;
        const double sa = (aE + aW) + (aN + aS);
        if (sa != 0.0) w = sor_weight64(sa, c.omega);
    }
    GW(c.wtab64)[f] = w;
}

void launch_weight_table64(hipStream_t st, const CaseDev *cs, int ncase)
{
    hipLaunchKernelGGL(k_weight_table64, dim3((unsigned)ncase), dim3(64), 0, st, cs);
}

void launch_weights32(hipStream_t st, const CaseDev *cs, int ncase, int max_n)
{
    hipLaunchKernelGGL(k_weights32, dim3((unsigned)(max_n - 2), (unsigned)ncase), dim3(256), 0, st, cs);
}

void launch_ppe_rearm(hipStream_t st, const CaseDev *cs, int ncase)
{
    hipLaunchKernelGGL(k_ppe_rearm, dim3((ncase + 63) / 64), dim3(64), 0, st, cs, ncase);
}

// ---------------------------------------------------------------------------
// P7/P8: projection of faces and cells, domain ghosts, divergence max.
// ---------------------------------------------------------------------------
template <typename TS>
__global__ __launch_bounds__(256) void k_project(const CaseDev *cs, const int2 *tiles,
                                                 int ntiles, int ntiles_pad)
{
    TileCtx tc;
    if (!tile_setup(tiles, ntiles, ntiles_pad, cs, tc)) return;
    const CaseDev &c = cs[tc.cid];
    const Ctrl *ct = c.ctrl;
    if (!ct->active) return;
    const int nx = c.nx, ny = c.ny, P = c.P;
    const int i = 1 + tc.tx * kTileX + (threadIdx.x & 63);
    const double *p = c.p[ct->ppe_base];
    const double *ufs = c.uf[1], *vfs_ = c.vf[1];
    double *ufn = c.uf[0], *vfn_ = c.vf[0];
    TS *u = fld<TS>(c.u[ct->ad_base]), *v = fld<TS>(c.v[ct->ad_base]);
    const double dt = c.dt;
    const TS ub2 = (TS)c.ubc * (TS)2.0, vb2 = (TS)c.vbc * (TS)2.0;
    const AxisD &ax = c.x, &ay = c.y;
    double m = 0.0;
    if (i <= nx - 2) {
        const double dxi = ax.d[i], dxp = ax.d[i + 1], dxm = ax.d[i - 1];
        for (int r = threadIdx.x >> 6; r < c.tile_y; r += 4) {
            const int j = 1 + tc.ty * c.tile_y + r;
            if (j > ny - 2) break;
            const int id = i + j * P;
            const uint8_t f = c.flags[id];
            const double pc = p[id], pE = p[id + 1], pW = p[id - 1], pN = p[id + P], pS = p[id - P];
            // faces owned by this cell (east, north) and its west/south
            // neighbours' (recomputed for the divergence, same expression)
            double ue = ufs[id], uw = ufs[id - 1], vn = vfs_[id], vs = vfs_[id - P];
            if (f & F_OPEN_E) ue = ue - dt * ((pE - pc) * ax.g[i]);
            if (f & F_OPEN_W) uw = uw - dt * ((pc - pW) * ax.g[i - 1]);
            if (f & F_OPEN_N) vn = vn - dt * ((pN - pc) * ay.g[j]);
            if (f & F_OPEN_S) vs = vs - dt * ((pc - pS) * ay.g[j - 1]);
            ufn[id] = ue;
            vfn_[id] = vn;
            if (i == 1) ufn[id - 1] = uw;
            if (j == 1) vfn_[id - P] = vs;
            // cell velocities
            const double dyj = ay.d[j], dyp = ay.d[j + 1], dym = ay.d[j - 1];
            const double pe = ax.ip[i] * (pE * dxi + pc * dxp);
            const double pw = ax.im[i] * (pc * dxm + pW * dxi);
            const double pn = ay.ip[j] * (pN * dyj + pc * dyp);
            const double ps = ay.im[j] * (pc * dym + pS * dyj);
            const double blank = (f & F_FLUID) ? 1.0 : 0.0;
            const TS qu = (TS)(blank * ((double)u[id] - dt * ((pe - pw) * ax.rd[i])));
            const TS qv = (TS)(blank * ((double)v[id] - dt * ((pn - ps) * ay.rd[j])));
            u[id] = qu;
            v[id] = qv;
            if (i == 1) { u[id - 1] = -qu + ub2; v[id - 1] = -qv + vb2; }
            if (i == nx - 2) { u[id + 1] = -qu + ub2; v[id + 1] = -qv + vb2; }
            if (j == 1) { u[id - P] = -qu + ub2; v[id - P] = -qv + vb2; }
            if (j == ny - 2) { u[id + P] = -qu + ub2; v[id + P] = -qv + vb2; }
            if (f & F_FLUID) m = fmax(m, absinf((ue - uw) * ax.rd[i] + (vn - vs) * ay.rd[j]));
        }
    }
    m = block_max(m);
    if (threadIdx.x == 0) c.part[tc.t] = m;
}

__global__ __launch_bounds__(256) void k_step_end(const CaseDev *cs, int ncase)
{
    const CaseDev &c = cs[blockIdx.x];
    Ctrl *ct = c.ctrl;
    if (!ct->active) return;
    const int nt = c.ntx * c.nty;
    double m = 0.0;
    for (int k = threadIdx.x; k < nt; k += 256) m = fmax(m, c.part[k]);
    m = block_max(m);
    if (threadIdx.x == 0) {
        ct->div_max = m;
        if (m > ct->max_div || !(m <= 1.0e300)) ct->max_div = m;
        if (!(m <= 1.0e300)) ct->failed = 1;
        ct->steps_done += 1;
        ct->t += c.dt;
        if (ct->failed || ct->steps_done >= ct->steps_target) ct->active = 0;
    }
}

// ---------------------------------------------------------------------------
// host-side launch wrappers
// ---------------------------------------------------------------------------
static inline dim3 tile_grid(int ntiles) { return dim3((unsigned)(((ntiles + 7) / 8) * 8)); }
static inline int tile_pad(int ntiles) { return ((ntiles + 7) / 8) * 8; }

void launch_step_begin(hipStream_t st, const CaseDev *cs, int ncase)
{
    hipLaunchKernelGGL(k_step_begin, dim3((ncase + 63) / 64), dim3(64), 0, st, cs, ncase);
}

void launch_source(hipStream_t st, int prec, const CaseDev *cs, const int2 *tiles, int nt)
{
    if (prec == IF_FP64)
        hipLaunchKernelGGL(k_source<double>, tile_grid(nt), dim3(256), 0, st, cs, tiles, nt, tile_pad(nt));
    else
        hipLaunchKernelGGL(k_source<float>, tile_grid(nt), dim3(256), 0, st, cs, tiles, nt, tile_pad(nt));
}

void launch_ad_sweep(hipStream_t st, int prec, const CaseDev *cs, const int2 *tiles, int nt)
{
    if (prec == IF_FP64)
        hipLaunchKernelGGL(k_ad_sweep<double>, tile_grid(nt), dim3(256), 0, st, cs, tiles, nt, tile_pad(nt));
    else
        hipLaunchKernelGGL(k_ad_sweep<float>, tile_grid(nt), dim3(256), 0, st, cs, tiles, nt, tile_pad(nt));
}

void launch_ad_finalize(hipStream_t st, const CaseDev *cs, int ncase)
{
    hipLaunchKernelGGL(k_ad_finalize, dim3(ncase), dim3(256), 0, st, cs, ncase);
}

void launch_ghost_vel(hipStream_t st, int prec, const CaseDev *cs, int ncase, int max_ng, int mode)
{
    if (max_ng <= 0) return;
    dim3 g((max_ng + 255) / 256, ncase);
    if (prec == IF_FP64) hipLaunchKernelGGL(k_ghost_vel<double>, g, dim3(256), 0, st, cs, mode);
    else hipLaunchKernelGGL(k_ghost_vel<float>, g, dim3(256), 0, st, cs, mode);
}

void launch_pressure_bc(hipStream_t st, const CaseDev *cs, int ncase, int max_ring, int max_ng)
{
    hipLaunchKernelGGL(k_p_ring, dim3((max_ring + 255) / 256, ncase), dim3(256), 0, st, cs);
    if (max_ng > 0)
        hipLaunchKernelGGL(k_ghost_p, dim3((max_ng + 255) / 256, ncase), dim3(256), 0, st, cs);
}

void launch_faces(hipStream_t st, int prec, const CaseDev *cs, const int2 *tiles, int nt,
                  int out_buf, int with_rhs)
{
    if (prec == IF_FP64)
        hipLaunchKernelGGL(k_faces_rhs<double>, tile_grid(nt), dim3(256), 0, st, cs, tiles, nt,
                           tile_pad(nt), out_buf, with_rhs);
    else
        hipLaunchKernelGGL(k_faces_rhs<float>, tile_grid(nt), dim3(256), 0, st, cs, tiles, nt,
                           tile_pad(nt), out_buf, with_rhs);
}

// Initial state of a physics case: u = ubc, v = vbc on fluid cells (the
// ghost ring included), zero on solid cells, in both ping-pong buffers
// (their corners are never rewritten); the domain-ghost rule on the edges.
template <typename TS>
__global__ __launch_bounds__(256) void k_init_fields(const CaseDev *cs)
{
    const CaseDev &c = cs[blockIdx.y];
    const int L = blockIdx.x * 256 + threadIdx.x;
    const int nx = c.nx, ny = c.ny, P = c.P;
    if (L >= nx * ny) return;
    const int i = L % nx, j = L / nx;
    const int id = i + j * P;
    const bool fl = c.flags[id] & F_FLUID;
    TS uu = fl ? (TS)c.ubc : (TS)0, vv = fl ? (TS)c.vbc : (TS)0;
    const bool ei = (i == 0 || i == nx - 1), ej = (j == 0 || j == ny - 1);
    if (ei != ej) {  // edge ghost (not a corner): -q_in + 2 q_bc
        const int ii = (i == 0) ? 1 : (i == nx - 1 ? nx - 2 : i);
        const int jj = (j == 0) ? 1 : (j == ny - 1 ? ny - 2 : j);
        const bool fin = c.flags[ii + jj * P] & F_FLUID;
        const TS uin = fin ? (TS)c.ubc : (TS)0, vin = fin ? (TS)c.vbc : (TS)0;
        uu = -uin + (TS)c.ubc * (TS)2.0;
        vv = -vin + (TS)c.vbc * (TS)2.0;
    }
    for (int b = 0; b < 2; b++) {
        fld<TS>(c.u[b])[id] = uu;
        fld<TS>(c.v[b])[id] = vv;
    }
}

void launch_init_fields(hipStream_t st, int prec, const CaseDev *cs, int ncase, int max_cells)
{
    dim3 g((max_cells + 255) / 256, ncase);
    if (prec == IF_FP64) hipLaunchKernelGGL(k_init_fields<double>, g, dim3(256), 0, st, cs);
    else hipLaunchKernelGGL(k_init_fields<float>, g, dim3(256), 0, st, cs);
}

template <typename T, int TY, int M>
static void rbsor_one(hipStream_t st, const CaseDev *cs, const int2 *tiles, int nt, int s, int K)
{
    hipLaunchKernelGGL((k_rbsor<T, TY, M>), tile_grid(nt), dim3(256), 0, st, cs, tiles, nt,
                       tile_pad(nt), s, K);
}

bool sweep_variant_ok(int ty, int m)
{
    return (ty == 8 || ty == 16 || ty == 32) && (m == 1 || m == 2 || m == 4);
}

void launch_ppe_sweep(hipStream_t st, int prec, int solver, int ty, int m, const CaseDev *cs,
                      const int2 *tiles, int nt, int s, int K)
{
    if (solver == IF_PPE_JACOBI) {
        if (prec == IF_FP64)
            hipLaunchKernelGGL(k_jacobi<double>, tile_grid(nt), dim3(256), 0, st, cs, tiles, nt,
                               tile_pad(nt), s, K);
        else
            hipLaunchKernelGGL(k_jacobi<float>, tile_grid(nt), dim3(256), 0, st, cs, tiles, nt,
                               tile_pad(nt), s, K);
        return;
    }
#define IFF_RB(T, TY_, M_) \
    if (ty == TY_ && m == M_) { rbsor_one<T, TY_, M_>(st, cs, tiles, nt, s, K); return; }
    if (prec == IF_FP64) {
        IFF_RB(double, 8, 1) IFF_RB(double, 16, 1) IFF_RB(double, 32, 1)
        IFF_RB(double, 8, 2) IFF_RB(double, 16, 2) IFF_RB(double, 32, 2)
        IFF_RB(double, 8, 4) IFF_RB(double, 16, 4) IFF_RB(double, 32, 4)
    } else {
        IFF_RB(float, 8, 1) IFF_RB(float, 16, 1) IFF_RB(float, 32, 1)
        IFF_RB(float, 8, 2) IFF_RB(float, 16, 2) IFF_RB(float, 32, 2)
        IFF_RB(float, 8, 4) IFF_RB(float, 16, 4) IFF_RB(float, 32, 4)
    }
#undef IFF_RB
}

void launch_ppe_check(hipStream_t st, int prec, const CaseDev *cs, int ncase, const int2 *tiles,
                      int nt, int K, const GridBar *bar)
{
    if (prec == IF_FP64)
        hipLaunchKernelGGL(k_ppe_resid, tile_grid(nt), dim3(256), 0, st, cs, tiles, nt, tile_pad(nt), K);
    else
        hipLaunchKernelGGL(k_ppe_update_resid, tile_grid(nt), dim3(256), 0, st, cs, tiles, nt,
                           tile_pad(nt), K);
    hipLaunchKernelGGL(k_ppe_finalize, dim3(ncase), dim3(256), 0, st, cs, ncase, K, prec == IF_FP64 ? 0 : 1, bar);
}

void launch_project(hipStream_t st, int prec, const CaseDev *cs, const int2 *tiles, int nt)
{
    if (prec == IF_FP64)
        hipLaunchKernelGGL(k_project<double>, tile_grid(nt), dim3(256), 0, st, cs, tiles, nt, tile_pad(nt));
    else
        hipLaunchKernelGGL(k_project<float>, tile_grid(nt), dim3(256), 0, st, cs, tiles, nt, tile_pad(nt));
}

// ---------------------------------------------------------------------------
// Monitors: the force the fluid exerts on the bodies, summed over the
// blanked faces (a fluid cell's face toward a non-fluid cell), and u, v, p
// at probe cells.  On a face with the body to the east the traction is
// p (x) and mu (v_P - v_B) 2/dx (y); west: -p and the same shear; north:
// p (y) and mu (u_P - u_B) 2/dy (x); south: -p and the same shear; each
// times the face length.  p is the projection pressure, mu = 1/Re, the
// body velocity that of the neighbouring body.  Per-tile sums, then a
// per-case sum in tile order (deterministic).
// ---------------------------------------------------------------------------
// sample != 0: record the current fields of every listed case, running or
// not, at the current time (if_case_monitor_sample).  part: the per-tile
// partials (Fx, Fy per tile of the listed case); null: the case's own.
template <typename TS>
__global__ __launch_bounds__(256) void k_monitor_part(const CaseDev *cs, const int2 *tiles, int ntiles,
                                                      int ntiles_pad, int sample, double *part)
{
    TileCtx tc;
    if (!tile_setup(tiles, ntiles, ntiles_pad, cs, tc)) return;
    const CaseDev &c = cs[tc.cid];
    const Ctrl *ct = c.ctrl;
    if (!c.mon || !(ct->active || sample)) return;
    const TS *u = fld<TS>(c.u[ct->ad_base]), *v = fld<TS>(c.v[ct->ad_base]);
    const double *p = c.p[ct->ppe_base];
    const int nx = c.nx, ny = c.ny, P = c.P;
    const int i = 1 + tc.tx * kTileX + (threadIdx.x & 63);
    double fx = 0.0, fy = 0.0;
    if (i <= nx - 2) {
        const double dxi = c.x.d[i];
        for (int r = threadIdx.x >> 6; r < c.tile_y; r += 4) {
            const int j = 1 + tc.ty * c.tile_y + r;
            if (j > ny - 2) break;
            const int id = i + j * P;
            const uint8_t f = c.flags[id];
            if (!(f & F_FLUID)) continue;
            const double dyj = c.y.d[j];
            const double pc = p[id], uc = (double)u[id], vc = (double)v[id];
            if (!(c.flags[id + 1] & F_FLUID)) {
                fx += pc * dyj;
                fy += c.mu * (vc - blank_vel(c, id + 1, id + 1, 1)) * 2.0 / dxi * dyj;
            }
            if (!(c.flags[id - 1] & F_FLUID)) {
                fx -= pc * dyj;
                fy += c.mu * (vc - blank_vel(c, id - 1, id - 1, 1)) * 2.0 / dxi * dyj;
            }
            if (!(c.flags[id + P] & F_FLUID)) {
                fy += pc * dxi;
                fx += c.mu * (uc - blank_vel(c, id + P, id + P, 0)) * 2.0 / dyj * dxi;
            }
            if (!(c.flags[id - P] & F_FLUID)) {
                fy -= pc * dxi;
                fx += c.mu * (uc - blank_vel(c, id - P, id - P, 0)) * 2.0 / dyj * dxi;
            }
        }
    }
    fx = block_sum(fx);
    fy = block_sum(fy);
    if (threadIdx.x == 0) {
        double *pp = part ? part : c.mon->part;
        pp[2 * tc.t] = fx;
        pp[2 * tc.t + 1] = fy;
    }
}

template <typename TS>
__global__ void k_monitor_rec(const CaseDev *cs, int ncase, int sample, const double *part)
{
    const CaseDev &c = cs[blockIdx.x];
    const Ctrl *ct = c.ctrl;
    if (!c.mon || !(ct->active || sample) || threadIdx.x != 0) return;
    Monitor &m = *c.mon;
    const int nt = c.ntx * c.nty;
    double fx = 0.0, fy = 0.0;
    for (int k = 0; k < nt; k++) {
        fx += (part ? part : m.part)[2 * k];
        fy += (part ? part : m.part)[2 * k + 1];
    }
    const int stride = 3 + 3 * m.nprobe;
    const unsigned long long n = *m.count;
    double *rec = m.ring + (size_t)(n % (unsigned long long)m.cap) * stride;
    rec[0] = sample ? ct->t : ct->t + c.dt;  // the time this step ends at (sample: now)
    rec[1] = fx;
    rec[2] = fy;
    const TS *u = fld<TS>(c.u[ct->ad_base]), *v = fld<TS>(c.v[ct->ad_base]);
    const double *p = c.p[ct->ppe_base];
    for (int k = 0; k < m.nprobe; k++) {
        const int id = m.pi[k] + m.pj[k] * c.P;
        rec[3 + 3 * k] = (double)u[id];
        rec[4 + 3 * k] = (double)v[id];
        rec[5 + 3 * k] = p[id];
    }
    *m.count = n + 1;
}

void launch_monitor(hipStream_t st, int prec, const CaseDev *cs, int ncase, const int2 *tiles, int nt,
                    int sample, double *part)
{
    if (prec == IF_FP64) {
        hipLaunchKernelGGL(k_monitor_part<double>, tile_grid(nt), dim3(256), 0, st, cs, tiles, nt, tile_pad(nt),
                           sample, part);
        hipLaunchKernelGGL(k_monitor_rec<double>, dim3(ncase), dim3(64), 0, st, cs, ncase, sample,
                           (const double *)part);
    } else {
        hipLaunchKernelGGL(k_monitor_part<float>, tile_grid(nt), dim3(256), 0, st, cs, tiles, nt, tile_pad(nt),
                           sample, part);
        hipLaunchKernelGGL(k_monitor_rec<float>, dim3(ncase), dim3(64), 0, st, cs, ncase, sample,
                           (const double *)part);
    }
}

// Chunk timing on the device: thread 0 appends the real-time counter
// (s_memrealtime, constant rate) to buf[1 ..], buf[0] counting the stamps
// (at most cap are kept).  The stamps are kernels, so they can sit inside a
// captured sweep graph: no host event is recorded between graph launches.
__global__ void k_stamp(unsigned long long *buf, int cap)
{
    if (threadIdx.x != 0) return;
    const unsigned long long n = buf[0];
    if (n < (unsigned long long)cap) buf[1 + n] = __builtin_amdgcn_s_memrealtime();
    buf[0] = n + 1;
}

__global__ void k_stamp_reset(unsigned long long *buf)
{
    if (threadIdx.x == 0) buf[0] = 0;
}

void launch_stamp(hipStream_t st, unsigned long long *buf, int cap)
{
    hipLaunchKernelGGL(k_stamp, dim3(1), dim3(64), 0, st, buf, cap);
}

void launch_stamp_reset(hipStream_t st, unsigned long long *buf)
{
    hipLaunchKernelGGL(k_stamp_reset, dim3(1), dim3(64), 0, st, buf);
}

// The tile table of one case, tiles (0, t) for t < nt, written on the device.
__global__ void k_fill_tiles(int2 *tiles, int nt)
{
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t < nt) tiles[t] = make_int2(0, t);
}

void launch_fill_tiles(hipStream_t st, int2 *tiles, int nt)
{
    if (nt > 0) hipLaunchKernelGGL(k_fill_tiles, dim3((nt + 255) / 256), dim3(256), 0, st, tiles, nt);
}

// Launch-shape timing: reset every running case's control block to the
// saved copy (saved mirrors the control array that starts at ctrl0) and
// re-open its Poisson loop, so every timed launch sees the same state.
__global__ void k_tune_arm(const CaseDev *cs, int ncase, const Ctrl *ctrl0, const Ctrl *saved)
{
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= ncase) return;
    Ctrl *c = cs[k].ctrl;
    *c = saved[c - ctrl0];
    if (!c->active || c->failed) return;
    c->ppe_done = 0;
    c->ppe_sweeps = 0;
}

void launch_tune_arm(hipStream_t st, const CaseDev *cs, int ncase, const Ctrl *ctrl0, const Ctrl *saved)
{
    hipLaunchKernelGGL(k_tune_arm, dim3((ncase + 63) / 64), dim3(64), 0, st, cs, ncase, ctrl0, saved);
}

void launch_step_end(hipStream_t st, const CaseDev *cs, int ncase)
{
    hipLaunchKernelGGL(k_step_end, dim3(ncase), dim3(256), 0, st, cs, ncase);
}

}  // namespace iff
