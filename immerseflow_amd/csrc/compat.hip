// Compat mode: ImmerseFlow::ADsolver and ImmerseFlow::PPESolver as the
// reference codes them (src/include/ADSolver.cu:268-395,
// src/include/PPESolver.cu:13-71,137-205), quirks included, on gfx950.
// These reproduce the reference's own output files (results/uc.dat, vc.dat,
// p.dat); the performance path is physics mode (kernels.hip).
//
// Layout: the reference's unpadded nx*ny arrays (idx = i + j*nx); uf is
// (nx-1)*(ny-2), vf is (nx-2)*(ny-1).
#include <cmath>
#include <cstring>
#include <vector>

#include "compat.h"
#include "io.h"

namespace iff {

__device__ __forceinline__ double blk_sum(double v)
{
    __shared__ double s[4];
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    if ((threadIdx.x & 63) == 0) s[threadIdx.x >> 6] = v;
    __syncthreads();
    const double r = (s[0] + s[1]) + (s[2] + s[3]);
    __syncthreads();
    return r;
}

// initializeKernel, preSim.cu:53-97 (vortex IC, p = 0), and the iBlank stub
// of iBlankComputeKernel, preSim.cu:110-136 (1.0 everywhere).
__global__ void kc_init(int nx, int ny, const double *xc, const double *yc, double *u,
                        double *v, double *p, double *iblank)
{
    const int id = blockIdx.x * blockDim.x + threadIdx.x;
    if (id >= nx * ny) return;
    const int i = id % nx, j = id / nx;
    const double r = sqrt(pow(xc[i] - 0.5, 2.0) + pow(yc[j] - 0.5, 2.0));
    const double r0 = 0.1;
    u[id] = 1.0 - 0.25 * (yc[j] - 0.5) * exp((1.0 - pow(r / r0, 2.0)) / 2.0);
    v[id] = 0.25 * (xc[i] - 0.5) * exp((1.0 - pow(r / r0, 2.0)) / 2.0);
    p[id] = 0.0;
    iblank[id] = 1.0;
}

// Compute_velf, ADSolver.cu:162-189.  Under the reference launch (one
// grid-stride pass per thread) the vf loop runs only for ids in
// [(nx-1)(ny-2), (nx-2)(ny-1)), which is empty unless nx > ny.
__global__ void kc_velf(int nx, int ny, const double *dx, const double *dy, const double *u,
                        const double *v, double *uf, double *vf)
{
    const int id = blockIdx.x * blockDim.x + threadIdx.x;
    const int nuf = (nx - 1) * (ny - 2), nvf = (nx - 2) * (ny - 1);
    if (id < nuf) {
        const int i = id % (nx - 1), j = id / (nx - 1);
        const int idc = j * nx + i;
        uf[id] = 1.0 / (dx[i] + dx[i + 1]) * (u[idc + nx + 1] * dx[i] + u[idc + nx] * dx[i + 1]);
    } else if (id < nvf) {
        const int i = id % (nx - 2), j = id / (nx - 2);
        const int idc = j * nx + i;
        vf[id] = 1.0 / (dy[j + 1] + dy[j]) * (v[idc + nx + 1] * dy[j] + v[idc + 1] * dy[j + 1]);
    }
}

// set_velocity_BC, ADSolver.cu:191-217.  Edge ghosts read interior cells
// only.  The reference's corner cells race with the edge writes; this
// kernel fixes the outcome of a sequential sweep over id: corner (0,0)
// takes the pre-call ghost (0,1), the other corners take freshly written
// edge ghosts.  Each corner is written by the thread that owns the edge
// ghost it depends on, so no cell is read by one thread while another
// writes it.
__global__ void kc_bc(int nx, int ny, double *q, double qbc)
{
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    const double two = qbc * 2.0;
    if (k < ny - 2) {  // west and east edges, j = 1 + k
        const int j = 1 + k;
        const int w = j * nx, e = nx - 1 + j * nx;
        const double old_w = q[w];
        const double nw = -q[w + 1] + two;
        const double ne = -q[e - 1] + two;
        q[w] = nw;
        q[e] = ne;
        if (j == 1) q[0] = -old_w + two;                               // (0,0)
        if (j == ny - 2) q[(ny - 1) * nx] = -nw + two;                 // (0,ny-1)
        if (j == ny - 2) q[nx - 1 + (ny - 1) * nx] = -ne + two;        // (nx-1,ny-1)
    } else if (k < (ny - 2) + (nx - 2)) {  // south and north edges, i = 1 + k'
        const int i = 1 + (k - (ny - 2));
        const double ns = -q[i + nx] + two;
        q[i] = ns;
        q[i + (ny - 1) * nx] = -q[i + (ny - 2) * nx] + two;
        if (i == nx - 2) q[nx - 1] = -ns + two;                        // (nx-1,0)
    }
}

// ADSource, ADSolver.cu:46-79, for u and v.
__global__ void kc_source(int nx, int ny, const double *dx, const double *dy, double dt,
                          const double *u, const double *v, const double *uf, const double *vf,
                          double *sx, double *sy)
{
    const int id = blockIdx.x * blockDim.x + threadIdx.x;
    if (id >= nx * ny) return;
    const int i = id % nx, j = id / nx;
    double su = 0.0, sv = 0.0;
    if (i > 0 && i < nx - 1 && j > 0 && j < ny - 1) {
        const double dx_i = dx[i], dx_ip1 = dx[i + 1], dx_im1 = dx[i - 1];
        const double dy_j = dy[j], dy_jp1 = dy[j + 1], dy_jm1 = dy[j - 1];
        const double ufe = uf[i + (j - 1) * (nx - 1)], ufw = uf[i - 1 + (j - 1) * (nx - 1)];
        const double vfn = vf[i - 1 + j * (nx - 2)], vfs = vf[i - 1 + (j - 1) * (nx - 2)];
        su = (u[id] - (dt / dx_i) * 0.5 * (((1 / (dx_i + dx_ip1)) * (u[id + 1] * dx_i + u[id] * dx_ip1) * ufe)
                                           - ((1 / (dx_i + dx_im1)) * (u[id] * dx_im1 + u[id - 1] * dx_i) * ufw)))
             - (dt / dy_j) * 0.5 * (((1 / (dy_j + dy_jp1)) * (u[id + nx] * dy_j + u[id] * dy_jp1) * vfn)
                                    - ((1 / (dy_j + dy_jm1)) * (u[id] * dy_jm1 + u[id - nx] * dy_j) * vfs));
        sv = (v[id] - (dt / dx_i) * 0.5 * (((1 / (dx_i + dx_ip1)) * (v[id + 1] * dx_i + v[id] * dx_ip1) * ufe)
                                           - ((1 / (dx_i + dx_im1)) * (v[id] * dx_im1 + v[id - 1] * dx_i) * ufw)))
             - (dt / dy_j) * 0.5 * (((1 / (dy_j + dy_jp1)) * (v[id + nx] * dy_j + v[id] * dy_jp1) * vfn)
                                    - ((1 / (dy_j + dy_jm1)) * (v[id] * dy_jm1 + v[id - nx] * dy_j) * vfs));
    }
    sx[id] = su;
    sy[id] = sv;
}

// ADusolver_kernel + ADvsolver_kernel, ADSolver.cu:81-120: interior only;
// coefficients per axis (calculateADCoefficients, :34-39).  w != 1: the
// w_AD relaxation written and commented out at :95,:116.
__global__ void kc_sweep(int nx, int ny, const double *xp, const double *xm, const double *px,
                         const double *yp, const double *ym, const double *py,
                         const double *iblank, const double *sx, const double *sy,
                         const double *u, const double *v, double *un, double *vn, double w)
{
    const int id = blockIdx.x * blockDim.x + threadIdx.x;
    if (id >= nx * ny) return;
    const int i = id % nx, j = id / nx;
    if (i > 0 && i < nx - 1 && j > 0 && j < ny - 1) {
        const double ap = px[i] + py[j];
        double t = sx[id] + xp[i] * u[id + 1] + xm[i] * u[id - 1] + yp[j] * u[id + nx] + ym[j] * u[id - nx];
        double qu = iblank[id] * t / ap;
        t = sy[id] + xp[i] * v[id + 1] + xm[i] * v[id - 1] + yp[j] * v[id + nx] + ym[j] * v[id - nx];
        double qv = iblank[id] * t / ap;
        if (w != 1.0) {
            qu = (1.0 - w) * u[id] + w * qu;
            qv = (1.0 - w) * v[id] + w * qv;
        }
        un[id] = qu;
        vn[id] = qv;
    }
}

// Reduction, preSim.cu:376-445, in the reference's summation order: pass 1
// is reduce6<256> (:22-50) over n values with CUDAQuery's grid of
// ceil(n/256) blocks (:190-193): thread tid of block b starts at
// i = b*512 + tid and adds val(i) + val(i+256) (val(i) alone when i+256 >= n)
// per grid-stride step of grid_size = 512 * blocks; the block then folds its
// 256 sums pairwise, s[t] += s[t+h] for h = 128 .. 1 (warpReduce's lockstep
// reads leave the same s[0]).  Pass 2 is the same over the block sums with
// one block (grid_size 512).
template <typename F>
__device__ __forceinline__ double reduce6_block(F val, unsigned n, unsigned grid_size)
{
    __shared__ double sd[256];
    const unsigned tid = threadIdx.x;
    double acc = 0;
    for (unsigned i = blockIdx.x * 512 + tid; i < n; i += grid_size) {
        if (i + 256 < n) acc += val(i) + val(i + 256);
        else acc += val(i);
    }
    sd[tid] = acc;
    __syncthreads();
    for (unsigned h = 128; h >= 1; h >>= 1) {
        if (tid < h) sd[tid] += sd[tid + h];
        __syncthreads();
    }
    const double r = sd[0];
    __syncthreads();
    return r;
}

// Compute_{u,v}Residual_AD (ADSolver.cu:122-160): |new - old| on interior
// cells with iBlank == 1, 0 elsewhere; pass 1 of Reduction for u and v.
__global__ __launch_bounds__(256) void kc_l1(int nx, int ny, const double *iblank,
                                             const double *a, const double *b, const double *c,
                                             const double *d, double *part)
{
    const unsigned n = (unsigned)(nx * ny), gs = 512u * gridDim.x;
    auto val = [&](const double *x, const double *y) {
        return [=](unsigned id) {
            const int i = (int)id % nx, j = (int)id / nx;
            return (i > 0 && i < nx - 1 && j > 0 && j < ny - 1 && iblank[id] == 1) ? fabs(x[id] - y[id]) : 0.0;
        };
    };
    const double su = reduce6_block(val(a, b), n, gs);
    const double sv = reduce6_block(val(c, d), n, gs);
    if (threadIdx.x == 0) {
        part[2 * blockIdx.x] = su;
        part[2 * blockIdx.x + 1] = sv;
    }
}

// Pass 2 of Reduction over nb block sums (stride 2: u and v interleaved).
__global__ __launch_bounds__(256) void kc_finish(const double *part, int nb, int stride, double *out)
{
    const double s0 = reduce6_block([=](unsigned k) { return part[stride * k]; }, (unsigned)nb, 512u);
    double s1 = 0.0;
    if (stride == 2) s1 = reduce6_block([=](unsigned k) { return part[2 * k + 1]; }, (unsigned)nb, 512u);
    if (threadIdx.x == 0) {
        out[0] = s0;
        out[1] = s1;
    }
}

// jacobiIteration, PPESolver.cu:13-31 (boundary copied, no RHS).
__global__ void kc_jacobi(int nx, int ny, const double *cxp, const double *cxm, const double *cyp,
                          const double *cym, const double *p, double *pn)
{
    const int id = blockIdx.x * blockDim.x + threadIdx.x;
    if (id >= nx * ny) return;
    const int i = id % nx, j = id / nx;
    pn[id] = p[id];
    if (i > 0 && i < nx - 1 && j > 0 && j < ny - 1) {
        const double ap = -1 * ((cxp[i] + cxm[i]) + (cyp[j] + cym[j]));
        pn[id] = -(p[id + 1] * cxp[i] + p[id - 1] * cxm[i] + p[id + nx] * cyp[j] + p[id - nx] * cym[j]) / ap;
    }
}

// Compute_Residual, PPESolver.cu:33-50: signed A.p per interior cell (0 on
// the ring), pass 1 of Reduction.
__global__ __launch_bounds__(256) void kc_presid(int nx, int ny, const double *cxp,
                                                 const double *cxm, const double *cyp,
                                                 const double *cym, const double *p, double *part)
{
    const unsigned n = (unsigned)(nx * ny), gs = 512u * gridDim.x;
    const double r = reduce6_block(
        [=](unsigned id) {
            const int i = (int)id % nx, j = (int)id / nx;
            if (!(i > 0 && i < nx - 1 && j > 0 && j < ny - 1)) return 0.0;
            const double ap = -1 * ((cxp[i] + cxm[i]) + (cyp[j] + cym[j]));
            return p[id] * ap + p[id + 1] * cxp[i] + p[id - 1] * cxm[i] + p[id + nx] * cyp[j] +
                   p[id - nx] * cym[j];
        },
        n, gs);
    if (threadIdx.x == 0) part[blockIdx.x] = r;
}

// set_pressure_BC, PPESolver.cu:54-71.
__global__ void kc_pbc(int nx, int ny, double *p)
{
    const int id = blockIdx.x * blockDim.x + threadIdx.x;
    if (id >= nx * ny) return;
    const int i = id % nx, j = id / nx;
    if (i == 0 || j == 0) p[id] = 100.0;
}

// ---------------------------------------------------------------------------
CompatCase::~CompatCase()
{
    if (st) hipStreamSynchronize(st);
    if (mem) hipFree(mem);
    if (h_out) hipHostFree(h_out);
    if (st) hipStreamDestroy(st);
}

if_status CompatCase::create(const if_case_desc &d)
{
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1)
        return fail(IF_ERR_NODEV, "no HIP device available");
    if (d.nxf < 3 || d.nyf < 3 || !d.xf || !d.yf) return fail(IF_ERR_ARG, "case: bad grid");
    nx = d.nxf + 1;
    ny = d.nyf + 1;
    dt = d.dt;
    re = d.re;
    ad_itermax = d.ad_itermax;
    ppe_itermax = d.ppe_itermax;
    w_ad = d.ad_omega > 0 ? d.ad_omega : 1.0;
    xc.resize(nx);
    dx.resize(nx);
    yc.resize(ny);
    dy.resize(ny);
    grid_metric(d.nxf, d.xf, xc.data(), dx.data());
    grid_metric(d.nyf, d.yf, yc.data(), dy.data());
    IFF_HIP(hipStreamCreate(&st));
    const size_t N = (size_t)nx * ny;
    nblk = (int)((N + 255) / 256);
    // per-axis vectors: dx, dy, xc, yc, AD (xp xm px / yp ym py), PPE (cxp cxm / cyp cym)
    std::vector<double> ax(8 * nx, 0.0), ay(8 * ny, 0.0);
    const double k = dt / re;
    auto fill = [&](std::vector<double> &V, const std::vector<double> &w,
                    const std::vector<double> &c, int n, bool is_x) {
        for (int i = 0; i < n; i++) { V[i] = w[i]; V[n + i] = c[i]; }
        for (int i = 1; i < n - 1; i++) {
            const double di = w[i], dp = w[i + 1], dm = w[i - 1];
            V[2 * n + i] = k * (2 / (di * (di + dp)));
            V[3 * n + i] = k * (2 / (di * (di + dm)));
            V[4 * n + i] = is_x ? 1.0 + k * ((2 / (di * (di + dp))) + (2 / (di * (di + dm))))
                                : k * ((2 / (di * (di + dp))) + (2 / (di * (di + dm))));
            V[5 * n + i] = (2 / (di * (di + dp)));
            V[6 * n + i] = (2 / (di * (di + dm)));
        }
    };
    fill(ax, dx, xc, nx, true);
    fill(ay, dy, yc, ny, false);
    const size_t nuf = (size_t)(nx - 1) * (ny - 2), nvf = (size_t)(nx - 2) * (ny - 1);
    const size_t total = 12 * N + nuf + nvf + 8 * (size_t)(nx + ny) + 2 * (size_t)nblk + 4;
    IFF_HIP(hipMalloc(&mem, total * sizeof(double)));
    IFF_HIP(hipMemsetAsync(mem, 0, total * sizeof(double), st));
    double *m = mem;
    for (int b = 0; b < 2; b++) { u[b] = m; m += N; }
    for (int b = 0; b < 2; b++) { v[b] = m; m += N; }
    for (int b = 0; b < 2; b++) { p[b] = m; m += N; }
    sx = m; m += N;
    sy = m; m += N;
    iblank = m; m += N;
    pres = m; m += N;  // unused spare (keeps the layout simple)
    m += 2 * N;
    uf = m; m += nuf;
    vf = m; m += nvf;
    vx = m; m += 8 * nx;
    vy = m; m += 8 * ny;
    part = m; m += 2 * nblk;
    scal = m; m += 4;
    IFF_HIP(hipMemcpyAsync(vx, ax.data(), ax.size() * 8, hipMemcpyHostToDevice, st));
    IFF_HIP(hipMemcpyAsync(vy, ay.data(), ay.size() * 8, hipMemcpyHostToDevice, st));
    IFF_HIP(hipHostMalloc((void **)&h_out, 4 * sizeof(double), hipHostMallocDefault));
    hipLaunchKernelGGL(kc_init, dim3(nblk), dim3(256), 0, st, nx, ny, vx + nx, vy + ny, u[0], v[0],
                       p[0], iblank);
    IFF_HIP(hipGetLastError());
    IFF_HIP(hipStreamSynchronize(st));
    return IF_OK;
}

void CompatCase::bc(double *uu, double *vv)
{
    const int nb = (nx + ny + 255) / 256;
    hipLaunchKernelGGL(kc_bc, dim3(nb), dim3(256), 0, st, nx, ny, uu, 1.0);
    hipLaunchKernelGGL(kc_bc, dim3(nb), dim3(256), 0, st, nx, ny, vv, 0.0);
}

void CompatCase::velf(const double *uu, const double *vv)
{
    const int n = std::max((nx - 1) * (ny - 2), (nx - 2) * (ny - 1));
    hipLaunchKernelGGL(kc_velf, dim3((n + 255) / 256), dim3(256), 0, st, nx, ny, vx, vy, uu, vv, uf, vf);
}

// ImmerseFlow::ADsolver, ADSolver.cu:268-395.
if_status CompatCase::ad_step(if_step_stats *stt)
{
    const double *xp = vx + 2 * nx, *xm = vx + 3 * nx, *px = vx + 4 * nx;
    const double *yp = vy + 2 * ny, *ym = vy + 3 * ny, *py = vy + 4 * ny;
    // fresh per-call scratch (cudaMalloc'd in the reference, :275-287)
    IFF_HIP(hipMemsetAsync(u[1 - cur], 0, sizeof(double) * nx * ny, st));
    IFF_HIP(hipMemsetAsync(v[1 - cur], 0, sizeof(double) * nx * ny, st));
    velf(u[cur], v[cur]);  // :298, before the BC (stale ghosts)
    bc(u[cur], v[cur]);    // :304
    hipLaunchKernelGGL(kc_source, dim3(nblk), dim3(256), 0, st, nx, ny, vx, vy, dt, u[cur], v[cur],
                       uf, vf, sx, sy);  // :308
    double ures = 1.0, vres = 1.0;
    int iter = 0;
    while (ures + vres > std::pow(10.0, -6.0) && iter < ad_itermax) {  // :315
        bc(u[cur], v[cur]);                                             // :317
        velf(u[cur], v[cur]);                                           // :321
        hipLaunchKernelGGL(kc_sweep, dim3(nblk), dim3(256), 0, st, nx, ny, xp, xm, px, yp, ym, py,
                           iblank, sx, sy, u[cur], v[cur], u[1 - cur], v[1 - cur], w_ad);  // :326,:331
        bc(u[cur], v[cur]);                                             // :337 (old buffer)
        cur = 1 - cur;                                                  // :342-348
        hipLaunchKernelGGL(kc_l1, dim3(nblk), dim3(256), 0, st, nx, ny, iblank, u[cur], u[1 - cur],
                           v[cur], v[1 - cur], part);                   // :350-358
        hipLaunchKernelGGL(kc_finish, dim3(1), dim3(256), 0, st, part, nblk, 2, scal);
        IFF_HIP(hipMemcpyAsync(h_out, scal, 2 * sizeof(double), hipMemcpyDeviceToHost, st));
        IFF_HIP(hipStreamSynchronize(st));                              // :360-366
        ures = h_out[0];
        vres = h_out[1];
        iter += 1;
    }
    IFF_HIP(hipGetLastError());
    steps++;
    if (stt) {
        std::memset(stt, 0, sizeof(*stt));
        stt->steps = 1;
        stt->ad_iters = iter;
        stt->ad_res = ures + vres;
        stt->total_ad_iters = iter;
    }
    return IF_OK;
}

// ImmerseFlow::PPESolver loop, PPESolver.cu:164-195.
if_status CompatCase::ppe_solve(if_step_stats *stt)
{
    const double *cxp = vx + 5 * nx, *cxm = vx + 6 * nx, *cyp = vy + 5 * ny, *cym = vy + 6 * ny;
    IFF_HIP(hipMemsetAsync(p[1 - pcur], 0, sizeof(double) * nx * ny, st));
    hipLaunchKernelGGL(kc_pbc, dim3(nblk), dim3(256), 0, st, nx, ny, p[pcur]);
    double residual = 1.0;
    int iter = 0;
    while (residual > std::pow(10.0, -6.0) && iter < ppe_itermax) {
        hipLaunchKernelGGL(kc_jacobi, dim3(nblk), dim3(256), 0, st, nx, ny, cxp, cxm, cyp, cym,
                           p[pcur], p[1 - pcur]);
        pcur = 1 - pcur;
        hipLaunchKernelGGL(kc_presid, dim3(nblk), dim3(256), 0, st, nx, ny, cxp, cxm, cyp, cym,
                           p[pcur], part);
        hipLaunchKernelGGL(kc_finish, dim3(1), dim3(256), 0, st, part, nblk, 1, scal);
        IFF_HIP(hipMemcpyAsync(h_out, scal, sizeof(double), hipMemcpyDeviceToHost, st));
        IFF_HIP(hipStreamSynchronize(st));
        residual = h_out[0];
        iter += 1;
    }
    hipLaunchKernelGGL(kc_pbc, dim3(nblk), dim3(256), 0, st, nx, ny, p[pcur]);
    IFF_HIP(hipGetLastError());
    IFF_HIP(hipStreamSynchronize(st));
    if (stt) {
        std::memset(stt, 0, sizeof(*stt));
        stt->ppe_sweeps = iter;
        stt->ppe_res = residual;
        stt->total_ppe_sweeps = iter;
    }
    return IF_OK;
}

if_status CompatCase::get_field(int which, double *out, size_t n)
{
    const size_t N = (size_t)nx * ny;
    const double *src = nullptr;
    size_t need = N;
    switch (which) {
    case IF_FIELD_U: src = u[cur]; break;
    case IF_FIELD_V: src = v[cur]; break;
    case IF_FIELD_P: src = p[pcur]; break;
    case IF_FIELD_UF: src = uf; need = (size_t)(nx - 1) * (ny - 2); break;
    case IF_FIELD_VF: src = vf; need = (size_t)(nx - 2) * (ny - 1); break;
    case IF_FIELD_IBLANK: src = iblank; break;
    default: return fail(IF_ERR_ARG, "get_field: field not available in compat mode");
    }
    if (!out || n != need) return fail(IF_ERR_ARG, "get_field: size mismatch");
    IFF_HIP(hipStreamSynchronize(st));
    IFF_HIP(hipMemcpy(out, src, need * sizeof(double), hipMemcpyDeviceToHost));
    return IF_OK;
}

if_status CompatCase::set_field(int which, const double *in, size_t n)
{
    const size_t N = (size_t)nx * ny;
    double *dst = nullptr;
    size_t need = N;
    switch (which) {
    case IF_FIELD_U: dst = u[cur]; break;
    case IF_FIELD_V: dst = v[cur]; break;
    case IF_FIELD_P: dst = p[pcur]; break;
    case IF_FIELD_UF: dst = uf; need = (size_t)(nx - 1) * (ny - 2); break;
    case IF_FIELD_VF: dst = vf; need = (size_t)(nx - 2) * (ny - 1); break;
    default: return fail(IF_ERR_ARG, "set_field: field not settable");
    }
    if (!in || n != need) return fail(IF_ERR_ARG, "set_field: size mismatch");
    IFF_HIP(hipStreamSynchronize(st));
    IFF_HIP(hipMemcpy(dst, in, need * sizeof(double), hipMemcpyHostToDevice));
    return IF_OK;
}

}  // namespace iff
