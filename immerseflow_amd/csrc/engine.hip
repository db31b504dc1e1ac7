// Host runtime of the physics-mode fractional step.
//
// One Engine owns a stream and a set of resident cases ("slots"), each with
// its own device arena.  A step enqueues the fixed kernel sequence of the
// fractional step over the tiles of all running cases:
//
//   begin -> P1 source -> [P2 viscous sweep + ghost cells + finalize] x K_AD
//   -> P3 faces + RHS -> [P4 Poisson sweeps + P5 check] x chunks
//   -> pressure BC -> P7/P8 projection + divergence -> ghost cells -> end
//
// The two iterative loops end on the device (per-case flags in Ctrl).  The
// host enqueues loop units speculatively and watches asynchronous snapshots
// of the control blocks, so it stops enqueueing a few units after the last
// case converged instead of waiting for a result after every unit.
#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <mutex>
#include <string>

#include "engine.h"
#include "io.h"
#include "kernels.h"

namespace iff {

constexpr int kTileY = 16;  // base tiling (reductions, stencil passes)

// viscous iteration budgets up to this are enqueued without polling
constexpr int kAdEnqueueAll = 32;

// waves per workgroup of the register-tile sweep kinds (2, 3, 4)
static int reg_waves(int kind) { return kind == 4 ? 16 : kind == 3 ? 8 : 4; }

static size_t align256(size_t x) { return (x + 255) & ~(size_t)255; }

Engine::Engine(int prec, size_t budget)
    : prec_(prec), vprec_(prec == IF_MIXED_V32 ? IF_MIXED : IF_FP64), pprec_(prec == IF_FP64 ? IF_FP64 : IF_MIXED),
      budget_(budget)
{
    // default Poisson launch shape, by measurement on the 2048 x 1024
    // ellipse: FP32 sweeps are fastest on 16-wave tiles at 8 sweeps per
    // launch; FP64 sweeps (IEEE division per cell) on 8-wave tiles at 4,
    // or as persistent 16-wave chunks when those fill the device (refresh)
    if (prec != IF_FP64) {
        sweep_kind = 4;
        sweep_ty = 64;
        sweep_m = 8;
    }
}

Engine::~Engine()
{
    if (st_) hipStreamSynchronize(st_);
    drop_graph();
    for (auto &s : slots_)
        if (s && s->arena) hipFree(s->arena);
    for (auto &s : slots_)
        if (s && s->dev.gh) hipFree(s->dev.gh);
    for (auto &s : slots_)
        if (s && s->mon_mem) hipFree(s->mon_mem);
    if (d_cases_) hipFree(d_cases_);
    if (d_ctrl_) hipFree(d_ctrl_);
    if (d_bar_) hipFree(d_bar_);
    if (d_flags_) hipFree(d_flags_);
    if (d_tiles_base_) hipFree(d_tiles_base_);
    if (d_tiles_sweep_) hipFree(d_tiles_sweep_);
    if (h_snap_) hipHostFree(h_snap_);
    for (auto &e : ev_)
        if (e) hipEventDestroy(e);
    if (d_stamps_) hipFree(d_stamps_);
    if (h_stamps_) hipHostFree(h_stamps_);
    if (st_) hipStreamDestroy(st_);
}

if_status Engine::init()
{
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1)
        return fail(IF_ERR_NODEV, "no HIP device available");
    IFF_HIP(hipStreamCreate(&st_));  // default priority
    IFF_HIP(hipMalloc(&d_ctrl_, sizeof(Ctrl) * cap_ctrl_));
    IFF_HIP(hipMemsetAsync(d_ctrl_, 0, sizeof(Ctrl) * cap_ctrl_, st_));
    IFF_HIP(hipMalloc(&d_bar_, sizeof(GridBar)));
    // neighbour waits of the persistent sweep give up after 20 ms
    // (IFF_PERSISTENT_TIMEOUT_TICKS overrides it, for tests of the time-out
    // path); b lives until the synchronisation below
    GridBar b{};
    b.timeout = 2000000;
    if (const char *e = std::getenv("IFF_PERSISTENT_TIMEOUT_TICKS")) b.timeout = std::strtoull(e, nullptr, 10);
    IFF_HIP(hipMemcpyAsync(d_bar_, &b, sizeof(GridBar), hipMemcpyHostToDevice, st_));
    IFF_HIP(hipHostMalloc((void **)&h_snap_, sizeof(Ctrl) * cap_ctrl_ * kRing, hipHostMallocDefault));
    for (auto &e : ev_) IFF_HIP(hipEventCreateWithFlags(&e, hipEventDisableTiming));
    IFF_HIP(hipMalloc(&d_stamps_, sizeof(unsigned long long) * (1 + kStamps)));
    IFF_HIP(hipMemsetAsync(d_stamps_, 0, sizeof(unsigned long long) * (1 + kStamps), st_));
    IFF_HIP(hipHostMalloc((void **)&h_stamps_, sizeof(unsigned long long) * (1 + kStamps), hipHostMallocDefault));
    {
        int dev = 0, khz = 0;
        IFF_HIP(hipGetDevice(&dev));
        if (hipDeviceGetAttribute(&khz, hipDeviceAttributeWallClockRate, dev) == hipSuccess && khz > 0)
            stamp_hz_ = 1e3 * (double)khz;
    }
    IFF_HIP(hipStreamSynchronize(st_));
    return IF_OK;
}

int Engine::resident() const
{
    int n = 0;
    for (auto &s : slots_)
        if (s && s->state != -2) n++;
    return n;
}

// ---------------------------------------------------------------------------
// arena layout
// ---------------------------------------------------------------------------
struct ArenaPlan {
    size_t off_u[2], off_v[2], off_sx, off_sy;
    size_t off_uf[2], off_vf[2], off_p[2], off_rhs;
    size_t off_e[2], off_r32, off_w32, off_wtab;
    size_t off_lf;
    size_t off_flags, off_bodyid;
    size_t off_metd, off_xc, off_yc, off_metf;
    size_t off_bodies, off_poly, off_part, off_gcount;
    size_t total;
};

static ArenaPlan plan_arena(int prec, int nx, int ny, int P, int ntiles, int nbody,
                            size_t npoly, int gblocks, bool line)
{
    ArenaPlan a{};
    size_t o = 0;
    const size_t ts = prec == IF_MIXED_V32 ? 4 : 8;  // u, v, sx, sy
    auto fieldT = [&](size_t es) {
        const size_t lead = 256 / es - 1;
        const size_t n = (size_t)P * ny + lead + 64;
        const size_t at = o;
        o = align256(o + n * es);
        return at + lead * es;  // pre-offset pointer position
    };
    for (int b = 0; b < 2; b++) a.off_u[b] = fieldT(ts);
    for (int b = 0; b < 2; b++) a.off_v[b] = fieldT(ts);
    a.off_sx = fieldT(ts);
    a.off_sy = fieldT(ts);
    for (int b = 0; b < 2; b++) a.off_uf[b] = fieldT(8);
    for (int b = 0; b < 2; b++) a.off_vf[b] = fieldT(8);
    for (int b = 0; b < 2; b++) a.off_p[b] = fieldT(8);
    a.off_rhs = fieldT(8);
    if (prec != IF_FP64) {
        for (int b = 0; b < 2; b++) a.off_e[b] = fieldT(4);
        a.off_r32 = fieldT(4);
        a.off_w32 = fieldT(4);
    }
    a.off_wtab = o;  // 64 weights, FP32 or FP64 (the Poisson precision)
    o = align256(o + sizeof(double) * 64);
    if (line) {  // the line solve's factors, one stride per row, in the Poisson precision
        int B, S;
        line_partition(nx - 2, &B, &S);
        a.off_lf = o;
        o = align256(o + (prec == IF_FP64 ? 8 : 4) * line_row_stride(B) * (size_t)ny);
    }
    a.off_flags = fieldT(1);
    a.off_bodyid = fieldT(1);
    a.off_metd = o;
    o = align256(o + sizeof(double) * (size_t)kAxisDVecs * (nx + ny));
    a.off_xc = o;
    o = align256(o + sizeof(double) * nx);
    a.off_yc = o;
    o = align256(o + sizeof(double) * ny);
    a.off_metf = o;
    o = align256(o + sizeof(float) * (size_t)kAxisFVecs * (nx + ny));
    a.off_bodies = o;
    o = align256(o + sizeof(BodyDev) * std::max(1, nbody));
    a.off_poly = o;
    o = align256(o + sizeof(double) * std::max<size_t>(1, npoly));
    a.off_part = o;
    o = align256(o + sizeof(double) * std::max(1, ntiles));
    a.off_gcount = o;
    o = align256(o + sizeof(int) * std::max(1, gblocks));
    a.total = o;
    return a;
}

static int grid_n(const if_case_desc &d, int *nx, int *ny)
{
    *nx = d.nxf + 1;
    *ny = d.nyf + 1;
    return 0;
}

size_t Engine::footprint(const if_case_desc &d, int prec)
{
    int nx, ny;
    grid_n(d, &nx, &ny);
    const int P = pitch_for(nx);
    const int ntx = (nx - 2 + kTileX - 1) / kTileX, nty = (ny - 2 + kTileY - 1) / kTileY;
    size_t npoly = 0;
    for (int b = 0; b < d.nbody; b++)
        if (d.bodies[b].kind == IF_BODY_POLYGON) npoly += 2 * (size_t)d.bodies[b].nvert;
    const ArenaPlan a = plan_arena(prec, nx, ny, P, ntx * nty, d.nbody, npoly,
                                   (nx * ny + ghost_span() - 1) / ghost_span(), d.ppe_solver == IF_PPE_LINE);
    // ghost records: a perimeter-sized estimate for the budget
    const size_t gest = (size_t)(8 * (nx + ny));
    return a.total + gest * sizeof(GhostDev);
}

// ---------------------------------------------------------------------------
// host metric vectors (operation order of the reference kernels)
// ---------------------------------------------------------------------------
static void axis_vectors(const std::vector<double> &d, double dt, double re, double conv,
                         bool is_x, std::vector<double> &V /* kAxisDVecs*n */,
                         std::vector<float> &F /* kAxisFVecs*n */)
{
    const int n = (int)d.size();
    V.assign((size_t)kAxisDVecs * n, 0.0);
    F.assign((size_t)kAxisFVecs * n, 0.0f);
    double *vd = &V[0 * n], *ip = &V[1 * n], *im = &V[2 * n], *cs = &V[3 * n], *rd = &V[4 * n],
           *g = &V[5 * n], *cp = &V[6 * n], *cm = &V[7 * n], *ap = &V[8 * n], *am = &V[9 * n],
           *ad = &V[10 * n];
    const double k = dt / re;
    for (int i = 0; i < n; i++) {
        vd[i] = d[i];
        cs[i] = (dt / d[i]) * conv;
        rd[i] = 1.0 / d[i];
        if (i + 1 < n) {
            ip[i] = 1 / (d[i] + d[i + 1]);
            g[i] = 2.0 / (d[i] + d[i + 1]);
        }
        if (i >= 1) im[i] = 1 / (d[i] + d[i - 1]);
        if (i >= 1 && i <= n - 2) {
            const double di = d[i], dp = d[i + 1], dm = d[i - 1];
            cp[i] = (2 / (di * (di + dp)));
            cm[i] = (2 / (di * (di + dm)));
            ap[i] = k * (2 / (di * (di + dp)));
            am[i] = k * (2 / (di * (di + dm)));
            ad[i] = is_x ? 1.0 + k * ((2 / (di * (di + dp))) + (2 / (di * (di + dm))))
                         : k * ((2 / (di * (di + dp))) + (2 / (di * (di + dm))));
        }
    }
    const double *src[kAxisFVecs] = {vd, ip, im, cs, ap, am, ad, cp, cm};
    for (int v = 0; v < kAxisFVecs; v++)
        for (int i = 0; i < n; i++) F[(size_t)v * n + i] = (float)src[v][i];
}

static BodyDev body_at(const if_body &b, const double *dvx, const double *dvy, double t,
                       double dt_for_state)
{
    (void)dt_for_state;
    BodyDev o{};
    o.kind = b.kind;
    o.xc = b.xc;
    o.yc = b.yc;
    o.r = b.r;
    o.a = b.a;
    o.b = b.b;
    o.c = b.cth;
    o.s = b.sth;
    o.nvert = b.nvert;
    o.vx = dvx;
    o.vy = dvy;
    if (b.motion == IF_MOTION_HEAVE) {
        const double w = 2.0 * M_PI * b.freq;
        o.yc = b.yc + b.amp * std::sin(w * t);
        o.vb = b.amp * w * std::cos(w * t);
        o.ayb = -b.amp * w * w * std::sin(w * t);
    }
    return o;
}

// ---------------------------------------------------------------------------
// case creation: allocation + readGridData + initializeData equivalents
// ---------------------------------------------------------------------------
if_status Engine::build_slot(Slot &s, const if_case_desc &d)
{
    if (d.nxf < 3 || d.nyf < 3 || !d.xf || !d.yf)
        return fail(IF_ERR_ARG, "case: need >= 3 face coordinates per axis");
    if (!(d.dt > 0) || !(d.re > 0)) return fail(IF_ERR_ARG, "case: dt and Re must be > 0");
    if (d.nbody < 0 || d.nbody > IF_MAX_BODIES || (d.nbody > 0 && !d.bodies))
        return fail(IF_ERR_ARG, "case: bad body list");
    if (d.ppe_solver != IF_PPE_JACOBI && d.ppe_solver != IF_PPE_RBSOR && d.ppe_solver != IF_PPE_LINE)
        return fail(IF_ERR_ARG, "case: unknown Poisson solver");
    if (d.ppe_solver == IF_PPE_LINE && d.nxf - 1 > 64 * kLineMaxB)
        return fail(IF_ERR_ARG, "case: line SOR takes at most " + std::to_string(64 * kLineMaxB) +
                                    " interior cells per row");
    if (d.ad_itermax < 0 || d.ppe_itermax < 0) return fail(IF_ERR_ARG, "case: negative itermax");
    s.desc = d;
    s.prec = prec_;
    s.xf.assign(d.xf, d.xf + d.nxf);
    s.yf.assign(d.yf, d.yf + d.nyf);
    for (int i = 1; i < d.nxf; i++)
        if (!(s.xf[i] > s.xf[i - 1])) return fail(IF_ERR_ARG, "case: x faces not increasing");
    for (int i = 1; i < d.nyf; i++)
        if (!(s.yf[i] > s.yf[i - 1])) return fail(IF_ERR_ARG, "case: y faces not increasing");
    s.desc.xf = s.xf.data();
    s.desc.yf = s.yf.data();
    s.bodies.assign(d.bodies, d.bodies + d.nbody);
    s.polyx.resize(d.nbody);
    s.polyy.resize(d.nbody);
    size_t npoly = 0;
    for (int b = 0; b < d.nbody; b++) {
        const if_body &B = s.bodies[b];
        if (B.kind == IF_BODY_POLYGON) {
            if (B.nvert < 3 || !B.vx || !B.vy) return fail(IF_ERR_ARG, "case: bad polygon");
            s.polyx[b].assign(B.vx, B.vx + B.nvert);
            s.polyy[b].assign(B.vy, B.vy + B.nvert);
            npoly += 2 * (size_t)B.nvert;
        } else if (B.kind == IF_BODY_ELLIPSE) {
            if (!(B.a > 0) || !(B.b > 0)) return fail(IF_ERR_ARG, "case: bad ellipse");
        } else if (B.kind == IF_BODY_CIRCLE) {
            if (!(B.r > 0)) return fail(IF_ERR_ARG, "case: bad circle");
        } else {
            return fail(IF_ERR_ARG, "case: unknown body kind");
        }
        if (B.motion == IF_MOTION_HEAVE) s.moving = true;
    }
    s.desc.bodies = s.bodies.data();
    grid_n(d, &s.nx, &s.ny);
    s.NX = s.nx - 2;
    s.NY = s.ny - 2;
    s.P = pitch_for(s.nx);
    s.xc.resize(s.nx);
    s.dx.resize(s.nx);
    s.yc.resize(s.ny);
    s.dy.resize(s.ny);
    grid_metric(d.nxf, s.xf.data(), s.xc.data(), s.dx.data());
    grid_metric(d.nyf, s.yf.data(), s.yc.data(), s.dy.data());
    s.ntx = (s.NX + kTileX - 1) / kTileX;
    s.nty = (s.NY + kTileY - 1) / kTileY;
    s.t0 = d.t0;
    s.ghost_blocks = (s.nx * s.ny + ghost_span() - 1) / ghost_span();

    const ArenaPlan a = plan_arena(prec_, s.nx, s.ny, s.P, s.ntx * s.nty, d.nbody, npoly,
                                   s.ghost_blocks, d.ppe_solver == IF_PPE_LINE);
    if (budget_ && used_ + a.total > budget_)
        return fail(IF_ERR_BUDGET, "case footprint exceeds the remaining HBM budget");
    if (hipMalloc(&s.arena, a.total) != hipSuccess) {
        s.arena = nullptr;
        return fail(IF_ERR_NOMEM, "device allocation of " + std::to_string(a.total) + " bytes failed");
    }
    s.bytes = a.total;
    char *A = (char *)s.arena;
    IFF_HIP(hipMemsetAsync(s.arena, 0, a.total, st_));

    CaseDev &c = s.dev;
    std::memset(&c, 0, sizeof(c));
    c.nx = s.nx;
    c.ny = s.ny;
    c.P = s.P;
    c.ntx = s.ntx;
    c.nty = s.nty;
    c.tile_y = kTileY;
    c.nbody = d.nbody;
    c.moving = s.moving ? 1 : 0;
    c.prec = vprec_;
    c.ppe_solver = d.ppe_solver;
    c.ad_itermax = d.ad_itermax;
    c.ppe_itermax = d.ppe_itermax;
    c.check_every = d.check_every > 0 ? d.check_every : 10;
    c.dt = d.dt;
    c.rdt = 1.0 / d.dt;
    c.omega = d.ppe_omega;
    if (!(c.omega > 0)) {
        const int N = std::max(s.NX, s.NY);
        c.omega = 2.0 / (1.0 + std::sin(M_PI / N));
    }
    c.omega_f = (float)c.omega;
    c.ppe_tol = d.ppe_tol;
    c.ad_tol = d.ad_tol;
    c.ubc = d.ubc;
    c.vbc = d.vbc;
    c.mu = 1.0 / d.re;
    c.ad_omega = d.ad_omega > 0 ? d.ad_omega : 1.0;
    c.mon = nullptr;
    for (int b = 0; b < 2; b++) {
        c.u[b] = A + a.off_u[b];
        c.v[b] = A + a.off_v[b];
        c.uf[b] = (double *)(A + a.off_uf[b]);
        c.vf[b] = (double *)(A + a.off_vf[b]);
        c.p[b] = (double *)(A + a.off_p[b]);
        if (pprec_ == IF_MIXED) c.e[b] = (float *)(A + a.off_e[b]);
    }
    c.sx = A + a.off_sx;
    c.sy = A + a.off_sy;
    c.rhs = (double *)(A + a.off_rhs);
    if (pprec_ == IF_MIXED) {
        c.r32 = (float *)(A + a.off_r32);
        c.w32 = (float *)(A + a.off_w32);
    }
    if (d.ppe_solver == IF_PPE_LINE) {
        c.lf = A + a.off_lf;
        line_partition(s.nx - 2, &c.lB, &c.lS);
    }
    c.flags = (uint8_t *)(A + a.off_flags);
    c.bodyid = (int8_t *)(A + a.off_bodyid);
    c.part = (double *)(A + a.off_part);
    c.gcount = (int *)(A + a.off_gcount);
    c.bodies = (BodyDev *)(A + a.off_bodies);
    c.ctrl = d_ctrl_ + s.id;

    // metric vectors
    std::vector<double> VX, VY;
    std::vector<float> FX, FY;
    axis_vectors(s.dx, d.dt, d.re, d.conv_factor, true, VX, FX);
    axis_vectors(s.dy, d.dt, d.re, d.conv_factor, false, VY, FY);
    // A case whose Poisson coefficients (in the Poisson precision) are each
    // one value over the faces that can open (east and north: cells
    // 1..n-3; west and south: cells 2..n-2) has one SOR weight per pattern
    // of face flags: the register-tile sweep looks the weight up by the
    // cell's flag byte instead of loading a per-cell weight (mixed) or
    // dividing on every launch (FP64).
    auto flat = [](const auto &V, int n, int icp) {
        const auto *cp = &V[(size_t)icp * n], *cm = &V[(size_t)(icp + 1) * n];
        for (int i = 2; i <= n - 3; i++)
            if (cp[i] != cp[1]) return false;
        for (int i = 3; i <= n - 2; i++)
            if (cm[i] != cm[2]) return false;
        return true;
    };
    // mixed: the five-operation FP32 form needs one value per axis on both
    // sides of a cell (aE = aW, aN = aS), which the table form then assumes
    c.uni = pprec_ == IF_MIXED && d.ppe_solver == IF_PPE_RBSOR && flat(FX, s.nx, 7) && flat(FY, s.ny, 7) &&
            s.nx >= 4 && s.ny >= 4 && FX[(size_t)7 * s.nx + 1] == FX[(size_t)8 * s.nx + 2] &&
            FY[(size_t)7 * s.ny + 1] == FY[(size_t)8 * s.ny + 2];
    if (c.uni) {
        c.uax = FX[(size_t)7 * s.nx + 1];
        c.uratio = FY[(size_t)7 * s.ny + 1] / c.uax;
        c.uinv = 1.0f / c.uax;
    }
    c.wtab = c.uni ? (float *)(A + a.off_wtab) : nullptr;
    c.wtab64 = (pprec_ == IF_FP64 && flat(VX, s.nx, 6) && flat(VY, s.ny, 6)) ? (double *)(A + a.off_wtab) : nullptr;
    double *md = (double *)(A + a.off_metd);
    float *mf = (float *)(A + a.off_metf);
    IFF_HIP(hipMemcpyAsync(md, VX.data(), VX.size() * 8, hipMemcpyHostToDevice, st_));
    IFF_HIP(hipMemcpyAsync(md + VX.size(), VY.data(), VY.size() * 8, hipMemcpyHostToDevice, st_));
    IFF_HIP(hipMemcpyAsync(mf, FX.data(), FX.size() * 4, hipMemcpyHostToDevice, st_));
    IFF_HIP(hipMemcpyAsync(mf + FX.size(), FY.data(), FY.size() * 4, hipMemcpyHostToDevice, st_));
    auto axd = [&](AxisD &x, double *base, int n) {
        const double **p[kAxisDVecs] = {&x.d, &x.ip, &x.im, &x.cs, &x.rd, &x.g,
                                        &x.cp, &x.cm, &x.ap, &x.am, &x.ad};
        for (int v = 0; v < kAxisDVecs; v++) *p[v] = base + (size_t)v * n;
    };
    auto axf = [&](AxisF &x, float *base, int n) {
        const float **p[kAxisFVecs] = {&x.d, &x.ip, &x.im, &x.cs, &x.ap, &x.am, &x.ad, &x.cp, &x.cm};
        for (int v = 0; v < kAxisFVecs; v++) *p[v] = base + (size_t)v * n;
    };
    axd(c.x, md, s.nx);
    axd(c.y, md + VX.size(), s.ny);
    axf(c.xf, mf, s.nx);
    axf(c.yf, mf + FX.size(), s.ny);
    double *dxc = (double *)(A + a.off_xc), *dyc = (double *)(A + a.off_yc);
    IFF_HIP(hipMemcpyAsync(dxc, s.xc.data(), s.nx * 8, hipMemcpyHostToDevice, st_));
    IFF_HIP(hipMemcpyAsync(dyc, s.yc.data(), s.ny * 8, hipMemcpyHostToDevice, st_));
    c.xc = dxc;
    c.yc = dyc;

    // bodies
    double *poly = (double *)(A + a.off_poly);
    s.bodies_dev.resize(d.nbody);
    size_t po = 0;
    for (int b = 0; b < d.nbody; b++) {
        const double *dvx = nullptr, *dvy = nullptr;
        if (s.bodies[b].kind == IF_BODY_POLYGON) {
            const int n = s.bodies[b].nvert;
            IFF_HIP(hipMemcpyAsync(poly + po, s.polyx[b].data(), n * 8, hipMemcpyHostToDevice, st_));
            IFF_HIP(hipMemcpyAsync(poly + po + n, s.polyy[b].data(), n * 8, hipMemcpyHostToDevice, st_));
            dvx = poly + po;
            dvy = poly + po + n;
            po += 2 * (size_t)n;
        }
        s.bodies_dev[b] = body_at(s.bodies[b], dvx, dvy, s.t0, d.dt);
    }
    if (d.nbody > 0)
        IFF_HIP(hipMemcpyAsync(c.bodies, s.bodies_dev.data(), sizeof(BodyDev) * d.nbody,
                               hipMemcpyHostToDevice, st_));

    // control block
    Ctrl ct{};
    ct.t = s.t0;
    IFF_HIP(hipMemcpyAsync(c.ctrl, &ct, sizeof(Ctrl), hipMemcpyHostToDevice, st_));

    // classification and ghost list: count first, then size the records.
    // The scratch descriptor and tile list are freed on every return, after
    // the stream has drained (an early error return leaves work queued).
    struct Scratch {
        hipStream_t st;
        CaseDev *tmp = nullptr;
        int2 *dtl = nullptr;
        ~Scratch()
        {
            if (tmp || dtl) (void)hipStreamSynchronize(st);
            if (tmp) (void)hipFree(tmp);
            if (dtl) (void)hipFree(dtl);
        }
    } scr{st_};
    IFF_HIP(hipMalloc(&scr.tmp, sizeof(CaseDev)));
    CaseDev *tmp = scr.tmp;
    IFF_HIP(hipMemcpyAsync(tmp, &c, sizeof(CaseDev), hipMemcpyHostToDevice, st_));
    launch_classify(st_, tmp, 1, s.nx * s.ny, 0);
    launch_ghost_build(st_, tmp, 1, s.nx * s.ny, 0);
    Ctrl back{};
    if (hipMemcpyAsync(&back, c.ctrl, sizeof(Ctrl), hipMemcpyDeviceToHost, st_) != hipSuccess ||
        hipStreamSynchronize(st_) != hipSuccess) {
        return fail(IF_ERR_HIP, std::string("case setup failed: ") + hipGetErrorString(hipGetLastError()));
    }
    const int total = back.ng_total;
    s.gcap = s.moving ? std::max(256, 3 * total + 256) : total;
    s.ng = total;
    c.gcap = s.gcap;
    if (s.gcap > 0) {
        if (hipMalloc(&c.gh, sizeof(GhostDev) * (size_t)s.gcap) != hipSuccess) {
            c.gh = nullptr;
            return fail(IF_ERR_NOMEM, "ghost record allocation failed");
        }
        s.bytes += sizeof(GhostDev) * (size_t)s.gcap;
        IFF_HIP(hipMemsetAsync(c.gh, 0, sizeof(GhostDev) * (size_t)s.gcap, st_));
    }
    ct = Ctrl{};
    ct.t = s.t0;
    IFF_HIP(hipMemcpyAsync(c.ctrl, &ct, sizeof(Ctrl), hipMemcpyHostToDevice, st_));
    IFF_HIP(hipMemcpyAsync(tmp, &c, sizeof(CaseDev), hipMemcpyHostToDevice, st_));
    launch_ghost_build(st_, tmp, 1, s.nx * s.ny, 0);
    launch_ghost_weights(st_, tmp, 1, s.gcap, 0);

    // initial state (u = ubc, v = vbc in fluid; ghost values; faces; p = 0)
    launch_init_fields(st_, vprec_, tmp, 1, s.nx * s.ny);
    {
        // ghost kernels need an active case; the two control blocks are
        // separate host objects, both alive until the synchronisation
        Ctrl on = ct, off = ct;
        on.active = 1;
        on.ng = off.ng = s.ng;
        on.ng_total = off.ng_total = s.ng;
        IFF_HIP(hipMemcpyAsync(c.ctrl, &on, sizeof(Ctrl), hipMemcpyHostToDevice, st_));
        launch_ghost_vel(st_, vprec_, tmp, 1, s.gcap, 1);
        std::vector<int2> tl(s.ntx * s.nty);
        for (int t = 0; t < (int)tl.size(); t++) tl[t] = make_int2(0, t);
        IFF_HIP(hipMalloc(&scr.dtl, sizeof(int2) * tl.size()));
        IFF_HIP(hipMemcpyAsync(scr.dtl, tl.data(), sizeof(int2) * tl.size(), hipMemcpyHostToDevice, st_));
        launch_faces(st_, vprec_, tmp, scr.dtl, (int)tl.size(), 0, 0);
        IFF_HIP(hipMemcpyAsync(c.ctrl, &off, sizeof(Ctrl), hipMemcpyHostToDevice, st_));
        const hipError_t e = hipStreamSynchronize(st_);
        if (e != hipSuccess) return fail(IF_ERR_HIP, std::string("case init failed: ") + hipGetErrorString(e));
    }
    return IF_OK;
}

if_status Engine::add(const if_case_desc &d, int nsteps, int *slot)
{
    if (!st_) return fail(IF_ERR_STATE, "engine not initialised");
    if (d.mode != IF_MODE_PHYSICS) return fail(IF_ERR_ARG, "batch cases must be physics mode");
    int id = -1;
    for (int k = 0; k < (int)slots_.size(); k++)
        if (!slots_[k] || slots_[k]->state == -2) { id = k; break; }
    if (id < 0) {
        if ((int)slots_.size() >= cap_ctrl_) return fail(IF_ERR_BUDGET, "too many slots");
        id = (int)slots_.size();
        slots_.emplace_back();
    }
    auto s = std::make_unique<Slot>();
    s->id = id;
    const if_status rc = build_slot(*s, d);
    if (rc != IF_OK) {
        if (s->arena) hipFree(s->arena);
        if (s->dev.gh) hipFree(s->dev.gh);
        return rc;
    }
    used_ += s->bytes;
    s->state = 0;
    s->steps_target = std::max(0, nsteps);
    // arm the case on the device
    Ctrl ct{};
    ct.active = nsteps > 0 ? 1 : 0;
    ct.steps_target = s->steps_target;
    ct.ng = s->ng;
    ct.ng_total = s->ng;
    ct.t = s->t0;
    IFF_HIP(hipMemcpyAsync(s->dev.ctrl, &ct, sizeof(Ctrl), hipMemcpyHostToDevice, st_));
    IFF_HIP(hipStreamSynchronize(st_));
    if (nsteps <= 0) s->state = 1;
    slots_[id] = std::move(s);
    if (slot) *slot = id;
    dirty_ = true;
    return IF_OK;
}

if_status Engine::remove(int k)
{
    Slot *s = slot(k);
    if (!s || s->state == -2) return fail(IF_ERR_ARG, "remove: empty slot");
    IFF_HIP(hipStreamSynchronize(st_));
    if (s->arena) hipFree(s->arena);
    if (s->dev.gh) hipFree(s->dev.gh);
    if (s->mon_mem) hipFree(s->mon_mem);
    used_ -= s->bytes;
    Ctrl z{};
    IFF_HIP(hipMemcpyAsync(d_ctrl_ + k, &z, sizeof(Ctrl), hipMemcpyHostToDevice, st_));
    IFF_HIP(hipStreamSynchronize(st_));  // z is a local
    slots_[k].reset();
    dirty_ = true;
    return IF_OK;
}

if_status Engine::set_target(int k, int nsteps)
{
    Slot *s = slot(k);
    if (!s || s->state == -2) return fail(IF_ERR_ARG, "set_target: empty slot");
    if (s->state == -1) return fail(IF_ERR_DIVERGED, "case has failed");
    std::vector<Ctrl> all;
    if_status rc = read_ctrl_sync(all);
    if (rc != IF_OK) return rc;
    Ctrl ct = all[k];
    ct.steps_target = ct.steps_done + nsteps;
    ct.active = nsteps > 0 ? 1 : 0;
    s->steps_target = ct.steps_target;
    s->steps_host = ct.steps_done;
    IFF_HIP(hipMemcpyAsync(d_ctrl_ + k, &ct, sizeof(Ctrl), hipMemcpyHostToDevice, st_));
    IFF_HIP(hipStreamSynchronize(st_));
    s->state = nsteps > 0 ? 0 : 1;
    dirty_ = true;
    return IF_OK;
}

if_status Engine::read_ctrl_sync(std::vector<Ctrl> &out)
{
    const int n = std::max(1, (int)slots_.size());
    out.resize(n);
    IFF_HIP(hipMemcpyAsync(out.data(), d_ctrl_, sizeof(Ctrl) * n, hipMemcpyDeviceToHost, st_));
    IFF_HIP(hipStreamSynchronize(st_));
    return IF_OK;
}

if_status Engine::sync()
{
    IFF_HIP(hipStreamSynchronize(st_));
    return IF_OK;
}

// Re-upload the list of running cases and the tile tables.
void Engine::drop_graph()
{
    if (g_sweeps_) hipGraphExecDestroy(g_sweeps_);
    g_sweeps_ = nullptr;
}

// Kernel launches of one Poisson chunk (K sweeps; the red-black kernels do
// M sweeps per launch).  Every argument of every launch is fixed for a
// given resident set and shape: the kernels read the loop state (buffer
// parity, sweep counts, done flags) from the control blocks on the device,
// so a captured chunk replays unchanged.
void Engine::enqueue_sweeps(int K, int M, bool any_rb, bool any_jac, bool any_line, bool pers, bool stamp)
{
    const int n = (int)order_.size();
    if (stamp) launch_stamp(st_, d_stamps_, kStamps);
    if (any_rb && pers)  // the whole chunk in one launch
        launch_ppe_sweep_pers(st_, pprec_, reg_waves(sweep_kind), tiles_sweep_ty_, M, d_cases_, d_tiles_sweep_, nt_sweep_,
                              K, d_bar_, d_flags_, wmode_);
    for (int s = 0; s < K; s++) {
        if (any_rb && !pers && s % M == 0) {
            if (sweep_kind >= 2)
                launch_ppe_sweep_reg(st_, pprec_, reg_waves(sweep_kind), tiles_sweep_ty_, M, d_cases_, d_tiles_sweep_,
                                     nt_sweep_, s, K, wmode_);
            else if (sweep_kind == 1)
                launch_ppe_sweep_stream(st_, pprec_, tiles_sweep_ty_, M, d_cases_, d_tiles_sweep_, nt_sweep_, s, K);
            else
                launch_ppe_sweep(st_, pprec_, IF_PPE_RBSOR, tiles_sweep_ty_, M, d_cases_, d_tiles_sweep_, nt_sweep_,
                                 s, K);
        }
        if (any_jac)  // point Jacobi runs on the base tiling
            launch_ppe_sweep(st_, pprec_, IF_PPE_JACOBI, kTileY, 1, d_cases_, d_tiles_base_, nt_base_, s, K);
        if (any_line) launch_line_sor(st_, pprec_, d_cases_, n, max_ring_, s, K);
    }
    if (stamp) launch_stamp(st_, d_stamps_, kStamps);
}

// Capture the chunk into a graph (kernel nodes only) unless the cached one
// matches the shape and the resident set.  refresh() invalidates the key,
// since it re-uploads the case list and tile tables the launches point at
// (and may reallocate them).  A new capture with the cached graph's
// topology (same kernels, same count) updates the instantiated graph in
// place (hipGraphExecUpdate: new kernel arguments and launch sizes); only
// a different topology destroys it and instantiates a new one.  The stream
// is idle whenever this runs (every step ends with a synchronisation).
if_status Engine::capture_sweeps(int K, int M, bool any_rb, bool any_jac, bool any_line, bool pers, bool stamp)
{
    const int key[9] = {K, M, sweep_kind, tiles_sweep_ty_, (int)any_rb, (int)any_jac, (int)any_line, (int)pers,
                        (int)stamp};
    if (g_sweeps_ && g_valid_ && std::equal(key, key + 9, g_key_)) return IF_OK;
    hipGraph_t g = nullptr;
    IFF_HIP(hipStreamBeginCapture(st_, hipStreamCaptureModeThreadLocal));
    enqueue_sweeps(K, M, any_rb, any_jac, any_line, pers, stamp);
    hipError_t e = hipStreamEndCapture(st_, &g);
    if (e != hipSuccess) return fail(IF_ERR_HIP, std::string("sweep graph capture: ") + hipGetErrorString(e));
    bool updated = false;
    if (g_sweeps_) {
        hipGraphNode_t bad = nullptr;
        hipGraphExecUpdateResult res = hipGraphExecUpdateError;
        updated = hipGraphExecUpdate(g_sweeps_, g, &bad, &res) == hipSuccess && res == hipGraphExecUpdateSuccess;
        if (!updated) {
            (void)hipGetLastError();  // a refused update leaves no error behind
            drop_graph();
        } else {
            graph_updates++;
        }
    }
    if (!updated) {
        e = hipGraphInstantiate(&g_sweeps_, g, nullptr, nullptr, 0);
        if (e != hipSuccess) {
            (void)hipGraphDestroy(g);
            g_sweeps_ = nullptr;
            return fail(IF_ERR_HIP, std::string("sweep graph instantiate: ") + hipGetErrorString(e));
        }
        graph_instantiations++;
    }
    (void)hipGraphDestroy(g);
    std::copy(key, key + 9, g_key_);
    g_valid_ = true;
    return IF_OK;
}

// After a step that ran persistent sweeps: a neighbour wait that timed out
// (blocks not all resident) leaves the results invalid.  Reset the state,
// stop using the persistent form and report the failure.
if_status Engine::check_barrier()
{
    GridBar b{};
    IFF_HIP(hipMemcpyAsync(&b, d_bar_, sizeof(GridBar), hipMemcpyDeviceToHost, st_));
    IFF_HIP(hipStreamSynchronize(st_));
    if (b.err == 0) return IF_OK;
    use_persistent = false;
    drop_graph();
    IFF_HIP(hipMemsetAsync(&d_bar_->err, 0, sizeof(unsigned), st_));
    if (cap_flags_ > 0) IFF_HIP(hipMemsetAsync(d_flags_, 0, sizeof(unsigned) * cap_flags_, st_));
    IFF_HIP(hipStreamSynchronize(st_));
    return fail(IF_ERR_HIP, "persistent Poisson sweep: a neighbour wait timed out (blocks not co-resident); "
                            "this step's result is invalid and the persistent form is now off");
}

if_status Engine::refresh()
{
    g_valid_ = false;  // the captured launches point at the old tables
    order_.clear();
    for (auto &s : slots_)
        if (s && s->state == 0) order_.push_back(s->id);
    const int n = (int)order_.size();
    {  // the weight mode of the sweep launches (launch_ppe_sweep_reg)
        int rb = 0, tab = 0, uni = 0;
        for (int id : order_) {
            const Slot &sl = *slots_[id];
            if (sl.desc.ppe_solver != IF_PPE_RBSOR) continue;
            rb++;
            tab += (pprec_ == IF_FP64 ? (void *)sl.dev.wtab64 : (void *)sl.dev.wtab) != nullptr;
            uni += sl.dev.uni;
        }
        wmode_ = tab == rb ? 1 : uni > 0 ? 2 : 0;
    }
    // FP64 default shape, by measurement: 16-wave tiles as persistent
    // chunks of the smallest height (48, 64, 80 rows) whose tiles fit in
    // one launch, i.e. the most tiles the device holds at once (1024^2:
    // 48 rows, 3.5 us per sweep against 3.9 at 64 and 5.0 at 80; 2048 x
    // 1024: only 80 rows fit, 4.6 us against 6.1 per-stage); per-stage
    // launches of 8-wave tiles of 32 rows when none fits
    if (shape_auto && pprec_ == IF_FP64) {
        auto tiles_at = [&](int ty) {
            int nt = 0;
            for (int id : order_)
                nt += ((slots_[id]->NX + reg_out_cols(4) - 1) / reg_out_cols(4)) * ((slots_[id]->NY + ty - 1) / ty);
            return nt;
        };
        int ty = 32;
        if (use_persistent)
            for (int cand : {48, 64, 80})
                if (pers_fits(IF_FP64, 16, cand, 4, tiles_at(cand))) {
                    ty = cand;
                    break;
                }
        sweep_kind = ty == 32 ? 3 : 4;
        sweep_ty = ty;
        sweep_m = 4;
    }
    if (n > cap_cases_) {
        if (d_cases_) hipFree(d_cases_);
        cap_cases_ = std::max(n, 2 * cap_cases_);
        IFF_HIP(hipMalloc(&d_cases_, sizeof(CaseDev) * cap_cases_));
    }
    std::vector<CaseDev> cs(n);
    std::vector<int2> tb, tsw;
    max_ng_ = max_ring_ = max_cells_ = max_ad_ = max_ppe_ = 0;
    any_moving_ = false;
    any_monitor_ = false;
    const int sty = sweep_ty;
    for (int k = 0; k < n; k++) {
        Slot &s = *slots_[order_[k]];
        s.dev.ppe_m = s.dev.ppe_solver == IF_PPE_RBSOR ? std::max(1, sweep_m) : 1;
        cs[k] = s.dev;
        for (int t = 0; t < s.ntx * s.nty; t++) tb.push_back(make_int2(k, t));
        const int ntys = (s.NY + sty - 1) / sty;
        const int wo = sweep_kind == 1 ? stream_out_cols(sweep_m) : sweep_kind >= 2 ? reg_out_cols(sweep_m) : 0;
        const int ntxs = wo > 0 ? (s.NX + wo - 1) / wo : s.ntx;
        for (int t = 0; t < ntxs * ntys; t++) tsw.push_back(make_int2(k, t));
        max_ng_ = std::max(max_ng_, s.gcap);
        max_ring_ = std::max(max_ring_, std::max(s.nx, s.ny));
        max_cells_ = std::max(max_cells_, s.nx * s.ny);
        max_ad_ = std::max(max_ad_, s.dev.ad_itermax);
        max_ppe_ = std::max(max_ppe_, s.dev.ppe_itermax);
        any_moving_ = any_moving_ || s.moving;
        any_monitor_ = any_monitor_ || s.mon_mem != nullptr;
    }
    if (n > 0) IFF_HIP(hipMemcpyAsync(d_cases_, cs.data(), sizeof(CaseDev) * n, hipMemcpyHostToDevice, st_));
    auto up = [&](int2 *&d, int &cap, int &cnt, const std::vector<int2> &v) -> if_status {
        if ((int)v.size() > cap) {
            if (d) hipFree(d);
            cap = std::max((int)v.size(), 2 * cap);
            IFF_HIP(hipMalloc(&d, sizeof(int2) * cap));
        }
        cnt = (int)v.size();
        if (cnt) IFF_HIP(hipMemcpyAsync(d, v.data(), sizeof(int2) * cnt, hipMemcpyHostToDevice, st_));
        return IF_OK;
    };
    if_status rc = up(d_tiles_base_, cap_tiles_base_, nt_base_, tb);
    if (rc != IF_OK) return rc;
    rc = up(d_tiles_sweep_, cap_tiles_sweep_, nt_sweep_, tsw);
    if (rc != IF_OK) return rc;
    // stage counters of the persistent sweep: one per sweep tile, all equal
    if (nt_sweep_ > cap_flags_) {
        if (d_flags_) hipFree(d_flags_);
        d_flags_ = nullptr;
        cap_flags_ = std::max(nt_sweep_, 2 * cap_flags_);
        IFF_HIP(hipMalloc(&d_flags_, sizeof(unsigned) * cap_flags_));
    }
    if (cap_flags_ > 0) IFF_HIP(hipMemsetAsync(d_flags_, 0, sizeof(unsigned) * cap_flags_, st_));
    tiles_sweep_ty_ = sty;
    IFF_HIP(hipStreamSynchronize(st_));  // host vectors go out of scope
    dirty_ = false;
    return IF_OK;
}

if_status Engine::snapshot(int r)
{
    const int n = std::max(1, (int)slots_.size());
    IFF_HIP(hipMemcpyAsync(h_snap_ + (size_t)r * cap_ctrl_, d_ctrl_, sizeof(Ctrl) * n,
                           hipMemcpyDeviceToHost, st_));
    IFF_HIP(hipEventRecord(ev_[r], st_));
    return IF_OK;
}

bool Engine::all_done(const Ctrl *snap, int which) const
{
    for (int id : order_) {
        const Ctrl &c = snap[id];
        if (!c.active) continue;
        if (which == 0 && !c.ad_done) return false;
        if (which == 1 && !c.ppe_done) return false;
    }
    return true;
}

// Moving bodies: place every moving body at t_{n+1}, reclassify, rebuild
// ghosts and weights on the device.
if_status Engine::update_moving()
{
    for (int id : order_) {
        Slot &s = *slots_[id];
        if (!s.moving) continue;
        const double t = s.t0 + (s.steps_host + 1) * s.desc.dt;
        for (int b = 0; b < (int)s.bodies.size(); b++)
            s.bodies_dev[b] = body_at(s.bodies[b], s.bodies_dev[b].vx, s.bodies_dev[b].vy, t, s.desc.dt);
        IFF_HIP(hipMemcpyAsync(s.dev.bodies, s.bodies_dev.data(), sizeof(BodyDev) * s.bodies.size(),
                               hipMemcpyHostToDevice, st_));
    }
    const int n = (int)order_.size();
    launch_classify(st_, d_cases_, n, max_cells_, 1);
    launch_ghost_build(st_, d_cases_, n, max_cells_, 1);
    launch_ghost_weights(st_, d_cases_, n, max_ng_, 1);
    return IF_OK;
}

if_status Engine::step(int nsteps, int *finished, if_step_stats *agg)
{
    if (finished) *finished = 0;
    if (agg) std::memset(agg, 0, sizeof(*agg));
    const int M = std::max(1, sweep_m);
    for (int it = 0; it < nsteps; it++) {
        bool ran_pers = false;
        if (dirty_) {
            if_status rc = refresh();
            if (rc != IF_OK) return rc;
        }
        if (order_.empty()) break;
        const int n = (int)order_.size();
        const int K = std::max(1, check_k);
        // Poisson solvers present (each kernel skips the other solvers' cases)
        bool any_rb = false, any_jac = false, any_line = false;
        for (int id : order_) {
            const int sv = slots_[id]->dev.ppe_solver;
            any_rb = any_rb || sv == IF_PPE_RBSOR;
            any_jac = any_jac || sv == IF_PPE_JACOBI;
            any_line = any_line || sv == IF_PPE_LINE;
        }
        const int m_eff = any_rb ? M : 1;  // launches per chunk: K / m_eff (rounded up)
        if (any_moving_) {
            if_status rc = update_moving();
            if (rc != IF_OK) return rc;
        }
        launch_step_begin(st_, d_cases_, n);
        launch_source(st_, vprec_, d_cases_, d_tiles_base_, nt_base_);

        // ---- P2: viscous Jacobi sweeps, ended per case on the device
        auto spec = [&](int maxn, int which, auto &&unit) -> if_status {
            int enq = 0, obs = 0;
            bool done = false;
            while (enq < maxn && !done) {
                unit(enq);
                if_status rc = snapshot(enq % kRing);
                if (rc != IF_OK) return rc;
                enq++;
                while (obs < enq) {
                    const int r = obs % kRing;
                    if (enq - obs < spec_depth && enq < maxn) {
                        const hipError_t q = hipEventQuery(ev_[r]);
                        if (q == hipErrorNotReady) break;
                        if (q != hipSuccess) return fail(IF_ERR_HIP, hipGetErrorString(q));
                    } else {
                        IFF_HIP(hipEventSynchronize(ev_[r]));
                    }
                    if (all_done(h_snap_ + (size_t)r * cap_ctrl_, which)) { done = true; break; }
                    obs++;
                }
            }
            return IF_OK;
        };
        auto ad_unit = [&](int) {
            launch_ad_sweep(st_, vprec_, d_cases_, d_tiles_base_, nt_base_);
            launch_ghost_vel(st_, vprec_, d_cases_, n, max_ng_, 0);
            launch_ad_finalize(st_, d_cases_, n);
        };
        // A short viscous loop (the usual 25-iteration budget) is enqueued
        // whole: the kernels end each case's iterations on the device, and
        // polling snapshots would cost a copy per iteration.
        if_status rc = IF_OK;
        if (max_ad_ <= kAdEnqueueAll)
            for (int u = 0; u < max_ad_; u++) ad_unit(u);
        else
            rc = spec(max_ad_, 0, ad_unit);
        if (rc != IF_OK) return rc;

        // ---- P3: faces + RHS
        launch_faces(st_, vprec_, d_cases_, d_tiles_base_, nt_base_, 1, 1);
        if (any_line) launch_line_factor(st_, pprec_, d_cases_, n, max_ring_);
        if (pprec_ == IF_MIXED)
            launch_weights32(st_, d_cases_, n, max_ring_);
        else
            launch_weight_table64(st_, d_cases_, n);

        // ---- P4/P5: Poisson in chunks of K sweeps (unit 0 = the entry check)
        const int nunits = 1 + (max_ppe_ + K - 1) / K;
        // persistent chunk: a register-tile shape with a persistent form in
        // this precision, and the grid fits on the device
        bool pers = any_rb && use_persistent && sweep_kind >= 2 &&
                    pers_fits(pprec_, reg_waves(sweep_kind), tiles_sweep_ty_, M, nt_sweep_);
        ran_pers = ran_pers || pers;
        if (any_rb) launch_sweeps = pers ? K : M;
        if (use_graphs) {
            rc = capture_sweeps(K, M, any_rb, any_jac, any_line, pers, timing);
            if (rc != IF_OK) return rc;
        }
        if (timing) launch_stamp_reset(st_, d_stamps_);
        hipError_t glaunch = hipSuccess;
        bool reopened = false;  // the solve after a persistent time-out (no entry check)
        auto ppe_unit = [&](int u) {
            if (u == 0) {
                if (!reopened) launch_ppe_check(st_, pprec_, d_cases_, n, d_tiles_base_, nt_base_, 0);
                return;
            }
            if (use_graphs) {
                const hipError_t e = hipGraphLaunch(g_sweeps_, st_);
                if (e != hipSuccess && glaunch == hipSuccess) glaunch = e;
            } else {
                enqueue_sweeps(K, M, any_rb, any_jac, any_line, pers, timing);
            }
            launch_ppe_check(st_, pprec_, d_cases_, n, d_tiles_base_, nt_base_, K, pers ? d_bar_ : nullptr);
        };
        rc = spec(nunits, 1, ppe_unit);
        if (rc != IF_OK) return rc;
        if (glaunch != hipSuccess) return fail(IF_ERR_HIP, std::string("sweep graph launch: ") + hipGetErrorString(glaunch));
        if (pers) {
            // A neighbour wait of the persistent sweep that timed out (the
            // blocks were not all resident) ended the Poisson loop early,
            // with a finite but unconverged pressure.  Turn the persistent
            // form off, clear the flag and the stage counters, and solve
            // again from that pressure with one launch per stage; the step
            // then completes as usual.
            GridBar b{};
            IFF_HIP(hipMemcpyAsync(&b, d_bar_, sizeof(GridBar), hipMemcpyDeviceToHost, st_));
            IFF_HIP(hipStreamSynchronize(st_));
            if (b.err) {
                use_persistent = false;
                pers = false;
                if (agg) agg->persistent_timeouts++;
                launch_sweeps = M;
                persistent_timeouts++;
                IFF_HIP(hipMemsetAsync(&d_bar_->err, 0, sizeof(unsigned), st_));
                if (cap_flags_ > 0) IFF_HIP(hipMemsetAsync(d_flags_, 0, sizeof(unsigned) * cap_flags_, st_));
                launch_ppe_rearm(st_, d_cases_, n);
                if (use_graphs) {
                    rc = capture_sweeps(K, M, any_rb, any_jac, any_line, false, timing);
                    if (rc != IF_OK) return rc;
                }
                if (timing) launch_stamp_reset(st_, d_stamps_);
                reopened = true;
                rc = spec(nunits, 1, ppe_unit);
                if (rc != IF_OK) return rc;
                if (glaunch != hipSuccess)
                    return fail(IF_ERR_HIP, std::string("sweep graph launch: ") + hipGetErrorString(glaunch));
            }
        }

        // ---- pressure BC, projection, divergence, ghost cells
        launch_pressure_bc(st_, d_cases_, n, max_ring_, max_ng_);
        launch_project(st_, vprec_, d_cases_, d_tiles_base_, nt_base_);
        launch_ghost_vel(st_, vprec_, d_cases_, n, max_ng_, 1);
        if (any_monitor_) launch_monitor(st_, vprec_, d_cases_, n, d_tiles_base_, nt_base_);
        launch_step_end(st_, d_cases_, n);
        for (int id : order_) slots_[id]->steps_host++;

        if (timing)  // completes with the synchronisation below
            IFF_HIP(hipMemcpyAsync(h_stamps_, d_stamps_, sizeof(unsigned long long) * (1 + kStamps),
                                   hipMemcpyDeviceToHost, st_));
        std::vector<Ctrl> all;
        rc = read_ctrl_sync(all);
        if (rc != IF_OK) return rc;
        if (ran_pers) {
            rc = check_barrier();
            if (rc != IF_OK) return rc;
        }
        const int timed = timing ? (int)std::min<unsigned long long>(h_stamps_[0], kStamps) / 2 : 0;
        if (timed > 0) {
            int maxsw = 0;
            for (int id : order_) maxsw = std::max(maxsw, all[id].ppe_sweeps);
            const int nfull = std::min(timed, maxsw / K);
            for (int u = 0; u < nfull; u++) {
                sweep_ms += 1e3 * (double)(h_stamps_[2 + 2 * u] - h_stamps_[1 + 2 * u]) / stamp_hz_;
                sweep_launches += pers ? 1 : (K + m_eff - 1) / m_eff;
            }
            for (int id : order_)
                sweep_cells += (double)std::min(all[id].ppe_sweeps, nfull * K) *
                               ((double)slots_[id]->NX * slots_[id]->NY);
        }
        for (int id : order_) {
            Slot &s = *slots_[id];
            const Ctrl &c = all[id];
            s.last.steps = c.steps_done;
            s.last.ad_iters = c.ad_iter;
            s.last.ad_res = c.ad_res;
            s.last.ppe_sweeps = c.ppe_sweeps;
            s.last.ppe_res = c.ppe_res;
            s.last.div_max = c.div_max;
            s.last.total_ad_iters = c.tot_ad;
            s.last.total_ppe_sweeps = c.tot_ppe;
            s.last.max_div = c.max_div;
            s.last.failed = c.failed;
            s.last.weight_table = wmode_ == 1;
            if (agg) {
                agg->weight_table = wmode_ == 1;
                agg->steps = std::max(agg->steps, 1 + it);
                agg->ad_iters = c.ad_iter;
                agg->ad_res = c.ad_res;
                agg->ppe_sweeps = c.ppe_sweeps;
                agg->ppe_res = c.ppe_res;
                agg->div_max = c.div_max;
                agg->total_ad_iters += c.ad_iter;
                agg->total_ppe_sweeps += c.ppe_sweeps;
                agg->max_div = std::max(agg->max_div, c.div_max);
                agg->failed |= c.failed;
            }
            if (c.failed) {
                s.state = -1;
                dirty_ = true;
            } else if (!c.active && c.steps_done >= c.steps_target) {
                s.state = 1;
                dirty_ = true;
                if (finished) (*finished)++;
            }
        }
    }
    return IF_OK;
}

// Launch-shape autotuner.  Times the candidate Poisson sweep shapes on the
// running cases' current fields: each candidate runs a few launches that
// read the solution and write only the solver's scratch buffers (FP64:
// launch 0 of a chunk reads p[base] and writes p[base^1]; mixed: launch 1
// reads e[1] and writes e[0]); the convergence check likewise writes only
// scratch (the other p buffer, r32, tile partials).  The control blocks are
// saved before and restored after, and reset to the saved copy before each
// timed check (a check flips the case's buffer parity), so neither the
// solution nor the loop state changes.
if_status Engine::autotune(if_tune_result *out)
{
    if (dirty_) {
        if_status rc = refresh();
        if (rc != IF_OK) return rc;
    }
    if (order_.empty()) return fail(IF_ERR_STATE, "autotune: no running case");
    for (int id : order_)
        if (slots_[id]->dev.ppe_solver != IF_PPE_RBSOR)
            return fail(IF_ERR_STATE, "autotune: red-black SOR cases only");
    std::vector<Ctrl> ctl;
    if_status rc = read_ctrl_sync(ctl);
    if (rc != IF_OK) return rc;
    double cells = 0.0;
    std::vector<int> last;
    for (int id : order_) {
        if (!ctl[id].active || ctl[id].failed) continue;
        cells += (double)slots_[id]->NX * slots_[id]->NY;
        if (slots_[id]->last.ppe_sweeps > 0) last.push_back(slots_[id]->last.ppe_sweeps);
    }
    if (cells <= 0.0) return fail(IF_ERR_STATE, "autotune: no running case");
    const int n = (int)order_.size();
    const int nall = std::max(1, (int)slots_.size());

    Ctrl *saved = nullptr;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    auto cleanup = [&]() {
        if (saved) (void)hipFree(saved);
        if (e0) (void)hipEventDestroy(e0);
        if (e1) (void)hipEventDestroy(e1);
    };
    if (hipMalloc(&saved, sizeof(Ctrl) * nall) != hipSuccess ||
        hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess) {
        cleanup();
        return fail(IF_ERR_NOMEM, "autotune: scratch allocation failed");
    }
    if (hipMemcpyAsync(saved, d_ctrl_, sizeof(Ctrl) * nall, hipMemcpyDeviceToDevice, st_) != hipSuccess) {
        cleanup();
        return fail(IF_ERR_HIP, "autotune: control-block copy failed");
    }

    struct Shape {
        int kind, ty, m;
    };
    // register tiles of 16, 8 and 4 waves at 8, 4 and 2 sweeps per launch,
    // and the LDS-tiled kernel; shapes a precision lacks are skipped
    static const Shape cands[] = {{4, 64, 8}, {4, 96, 8}, {4, 92, 9}, {4, 88, 10}, {4, 32, 8},
                                  {4, 48, 4}, {4, 64, 4}, {4, 80, 4}, {3, 32, 8}, {3, 48, 8},
                                  {3, 64, 8}, {3, 32, 4}, {3, 48, 4}, {3, 24, 2}, {2, 16, 4},
                                  {0, 16, 2}};
    constexpr int kReps = 6;
    auto sweep_launch = [&](const Shape &c, int s, int K) {
        if (c.kind >= 2)
            launch_ppe_sweep_reg(st_, pprec_, reg_waves(c.kind), c.ty, c.m, d_cases_, d_tiles_sweep_, nt_sweep_,
                                 s, K, wmode_);
        else if (c.kind == 1)
            launch_ppe_sweep_stream(st_, pprec_, c.ty, c.m, d_cases_, d_tiles_sweep_, nt_sweep_, s, K);
        else
            launch_ppe_sweep(st_, pprec_, IF_PPE_RBSOR, c.ty, c.m, d_cases_, d_tiles_sweep_, nt_sweep_, s, K);
    };
    auto elapsed = [&](float *ms) -> bool {
        return hipEventSynchronize(e1) == hipSuccess && hipEventElapsedTime(ms, e0, e1) == hipSuccess;
    };
    const Shape keep{sweep_kind, sweep_ty, sweep_m};
    const bool keep_auto = shape_auto;
    // each candidate is timed over its own tile table: refresh() must not
    // replace the candidate's shape by the FP64 default while tuning
    shape_auto = false;
    Shape best = keep;
    double best_ns = 1e300;
    int ncand = 0;
    for (const Shape &c : cands) {
        if (c.kind >= 2 && !reg_variant_ok(pprec_, reg_waves(c.kind), c.ty, c.m)) continue;
        sweep_kind = c.kind;
        sweep_ty = c.ty;
        sweep_m = c.m;
        rc = refresh();
        if (rc != IF_OK) break;
        // the launch below covers exactly the candidate's tiles
        {
            const int wo = c.kind == 1 ? stream_out_cols(c.m) : c.kind >= 2 ? reg_out_cols(c.m) : 0;
            int want = 0;
            for (int id : order_) {
                const Slot &sl = *slots_[id];
                want += (wo > 0 ? (sl.NX + wo - 1) / wo : sl.ntx) * ((sl.NY + c.ty - 1) / c.ty);
            }
            if (want != nt_sweep_ || tiles_sweep_ty_ != c.ty) {
                rc = fail(IF_ERR_STATE, "autotune: tile table does not match the candidate shape");
                break;
            }
        }
        // a shape that will run as persistent chunks is timed that way, over
        // 32 stages (a typical check interval); it writes only the scratch
        // error buffers and advances the stage counters of every tile alike.
        // FP64 shapes are timed per stage: an FP64 persistent launch's
        // second stage would overwrite the solve's current pressure.
        const bool pc = pprec_ != IF_FP64 && c.kind >= 2 && use_persistent &&
                        pers_fits(pprec_, reg_waves(c.kind), c.ty, c.m, nt_sweep_);
        const int kp = 32 * c.m;
        const int s0 = pprec_ == IF_FP64 ? 0 : c.m;
        auto timed = [&]() {
            if (pc)
                launch_ppe_sweep_pers(st_, pprec_, reg_waves(c.kind), c.ty, c.m, d_cases_, d_tiles_sweep_, nt_sweep_, kp,
                                      d_bar_, d_flags_, wmode_);
            else
                sweep_launch(c, s0, 2 * c.m);
        };
        launch_tune_arm(st_, d_cases_, n, d_ctrl_, saved);
        timed();  // warm-up
        timed();
        (void)hipEventRecord(e0, st_);
        for (int r = 0; r < kReps; r++) timed();
        (void)hipEventRecord(e1, st_);
        float ms = 0.f;
        if (!elapsed(&ms)) {
            rc = fail(IF_ERR_HIP, "autotune: timing failed");
            break;
        }
        const double ns = 1e6 * (double)ms / (kReps * (double)(pc ? kp : c.m) * cells);
        ncand++;
        if (ns < best_ns) {
            best_ns = ns;
            best = c;
        }
    }
    double check_sweeps = 0.0;
    if (rc == IF_OK) {
        sweep_kind = best.kind;
        sweep_ty = best.ty;
        sweep_m = best.m;
        shape_auto = false;
        rc = refresh();
    }
    if (rc == IF_OK) {
        (void)hipEventRecord(e0, st_);
        for (int r = 0; r < kReps; r++) {
            launch_tune_arm(st_, d_cases_, n, d_ctrl_, saved);
            launch_ppe_check(st_, pprec_, d_cases_, n, d_tiles_base_, nt_base_, 2 * best.m);
        }
        (void)hipEventRecord(e1, st_);
        float ms = 0.f;
        if (!elapsed(&ms)) rc = fail(IF_ERR_HIP, "autotune: timing failed");
        else check_sweeps = 1e6 * (double)ms / kReps / (best_ns * cells);
    }
    // restore the loop state
    const bool restored =
        hipMemcpyAsync(d_ctrl_, saved, sizeof(Ctrl) * nall, hipMemcpyDeviceToDevice, st_) == hipSuccess &&
        hipStreamSynchronize(st_) == hipSuccess;
    cleanup();
    if (!restored) return fail(IF_ERR_HIP, "autotune: control-block restore failed");
    if (rc == IF_OK && pprec_ != IF_FP64) rc = check_barrier();
    if (rc != IF_OK) {
        sweep_kind = keep.kind;
        sweep_ty = keep.ty;
        sweep_m = keep.m;
        shape_auto = keep_auto;
        mark_dirty();
        return rc;
    }
    int S = 0;
    if (!last.empty()) {
        std::nth_element(last.begin(), last.begin() + last.size() / 2, last.end());
        S = last[last.size() / 2];
    }
    if (S > 0) {
        const double k = std::sqrt(2.0 * S * std::max(check_sweeps, 0.0));
        const int m = best.m;
        int K = (int)std::lround(k / m) * m;
        K = std::max(std::max(4, m), std::min(1024, K));
        check_k = K;
    }
    mark_dirty();
    if (out) {
        out->kernel = best.kind;
        out->tile_y = best.ty;
        out->sweeps_per_launch = best.m;
        out->check_every = check_k;
        out->sweep_ns = best_ns;
        out->check_sweeps = check_sweeps;
        out->typical_sweeps = S;
        out->candidates = ncand;
    }
    return IF_OK;
}

if_status Engine::slot_state(int k, int *state, int *steps_done, if_step_stats *st)
{
    Slot *s = slot(k);
    if (!s) {
        if (state) *state = -2;
        return IF_OK;
    }
    if (state) *state = s->state;
    if (steps_done) *steps_done = s->last.steps;
    if (st) *st = s->last;
    return IF_OK;
}

// Device field -> reference layout (double).
if_status Engine::get_field(int k, int which, double *out, size_t n)
{
    Slot *s = slot(k);
    if (!s || s->state == -2) return fail(IF_ERR_ARG, "get_field: empty slot");
    if (!out) return fail(IF_ERR_ARG, "get_field: null output");
    std::vector<Ctrl> all;
    if_status rc = read_ctrl_sync(all);
    if (rc != IF_OK) return rc;
    const Ctrl &ct = all[k];
    const int nx = s->nx, ny = s->ny, P = s->P;
    size_t need = (size_t)nx * ny;
    if (which == IF_FIELD_UF) need = (size_t)(nx - 1) * (ny - 2);
    if (which == IF_FIELD_VF) need = (size_t)(nx - 2) * (ny - 1);
    if (n != need) return fail(IF_ERR_ARG, "get_field: size mismatch (expected " + std::to_string(need) + ")");
    const size_t np = (size_t)P * ny;
    std::vector<double> hd;
    std::vector<float> hf;
    std::vector<uint8_t> hb;
    const CaseDev &c = s->dev;
    auto fetch_ts = [&](void *p) -> if_status {
        if (vprec_ == IF_FP64) {
            hd.resize(np);
            IFF_HIP(hipMemcpy(hd.data(), p, np * 8, hipMemcpyDeviceToHost));
        } else {
            hf.resize(np);
            IFF_HIP(hipMemcpy(hf.data(), p, np * 4, hipMemcpyDeviceToHost));
            hd.assign(hf.begin(), hf.end());
        }
        return IF_OK;
    };
    auto fetch_d = [&](const double *p) -> if_status {
        hd.resize(np);
        IFF_HIP(hipMemcpy(hd.data(), p, np * 8, hipMemcpyDeviceToHost));
        return IF_OK;
    };
    switch (which) {
    case IF_FIELD_U: rc = fetch_ts(c.u[ct.ad_base]); break;
    case IF_FIELD_V: rc = fetch_ts(c.v[ct.ad_base]); break;
    case IF_FIELD_P: rc = fetch_d(c.p[ct.ppe_base]); break;
    case IF_FIELD_UF: rc = fetch_d(c.uf[0]); break;
    case IF_FIELD_VF: rc = fetch_d(c.vf[0]); break;
    case IF_FIELD_IBLANK:
    case IF_FIELD_FLAGS:
        hb.resize(np);
        IFF_HIP(hipMemcpy(hb.data(), c.flags, np, hipMemcpyDeviceToHost));
        break;
    default: return fail(IF_ERR_ARG, "get_field: unknown field");
    }
    if (rc != IF_OK) return rc;
    if (which == IF_FIELD_UF) {
        for (int j = 1; j < ny - 1; j++)
            for (int i = 0; i < nx - 1; i++) out[i + (size_t)(j - 1) * (nx - 1)] = hd[i + (size_t)j * P];
    } else if (which == IF_FIELD_VF) {
        for (int j = 0; j < ny - 1; j++)
            for (int i = 1; i < nx - 1; i++) out[(i - 1) + (size_t)j * (nx - 2)] = hd[i + (size_t)j * P];
    } else if (which == IF_FIELD_IBLANK) {
        for (int j = 0; j < ny; j++)
            for (int i = 0; i < nx; i++) out[i + (size_t)j * nx] = (hb[i + (size_t)j * P] & F_FLUID) ? 1.0 : 0.0;
    } else if (which == IF_FIELD_FLAGS) {
        for (int j = 0; j < ny; j++)
            for (int i = 0; i < nx; i++) out[i + (size_t)j * nx] = (double)hb[i + (size_t)j * P];
    } else {
        for (int j = 0; j < ny; j++)
            for (int i = 0; i < nx; i++) out[i + (size_t)j * nx] = hd[i + (size_t)j * P];
    }
    return IF_OK;
}

if_status Engine::set_field(int k, int which, const double *in, size_t n)
{
    Slot *s = slot(k);
    if (!s || s->state == -2) return fail(IF_ERR_ARG, "set_field: empty slot");
    if (!in) return fail(IF_ERR_ARG, "set_field: null input");
    std::vector<Ctrl> all;
    if_status rc = read_ctrl_sync(all);
    if (rc != IF_OK) return rc;
    const Ctrl &ct = all[k];
    const int nx = s->nx, ny = s->ny, P = s->P;
    size_t need = (size_t)nx * ny;
    if (which == IF_FIELD_UF) need = (size_t)(nx - 1) * (ny - 2);
    if (which == IF_FIELD_VF) need = (size_t)(nx - 2) * (ny - 1);
    if (which > IF_FIELD_VF) return fail(IF_ERR_ARG, "set_field: field is not settable");
    if (n != need) return fail(IF_ERR_ARG, "set_field: size mismatch");
    const size_t np = (size_t)P * ny;
    const CaseDev &c = s->dev;
    void *dst = nullptr;
    bool ts = false;
    switch (which) {
    case IF_FIELD_U: dst = c.u[ct.ad_base]; ts = true; break;
    case IF_FIELD_V: dst = c.v[ct.ad_base]; ts = true; break;
    case IF_FIELD_P: dst = c.p[ct.ppe_base]; break;
    case IF_FIELD_UF: dst = c.uf[0]; break;
    case IF_FIELD_VF: dst = c.vf[0]; break;
    }
    std::vector<double> hd(np, 0.0);
    if (ts && vprec_ == IF_MIXED) {
        std::vector<float> hf(np);
        IFF_HIP(hipMemcpy(hf.data(), dst, np * 4, hipMemcpyDeviceToHost));
        hd.assign(hf.begin(), hf.end());
    } else {
        IFF_HIP(hipMemcpy(hd.data(), dst, np * 8, hipMemcpyDeviceToHost));
    }
    if (which == IF_FIELD_UF) {
        for (int j = 1; j < ny - 1; j++)
            for (int i = 0; i < nx - 1; i++) hd[i + (size_t)j * P] = in[i + (size_t)(j - 1) * (nx - 1)];
    } else if (which == IF_FIELD_VF) {
        for (int j = 0; j < ny - 1; j++)
            for (int i = 1; i < nx - 1; i++) hd[i + (size_t)j * P] = in[(i - 1) + (size_t)j * (nx - 2)];
    } else {
        for (int j = 0; j < ny; j++)
            for (int i = 0; i < nx; i++) hd[i + (size_t)j * P] = in[i + (size_t)j * nx];
    }
    if (ts && vprec_ == IF_MIXED) {
        std::vector<float> hf(hd.begin(), hd.end());
        IFF_HIP(hipMemcpy(dst, hf.data(), np * 4, hipMemcpyHostToDevice));
    } else {
        IFF_HIP(hipMemcpy(dst, hd.data(), np * 8, hipMemcpyHostToDevice));
    }
    return IF_OK;
}

// Monitors: probes at the interior cell nearest each (x, y) (per axis the
// first nearest centre, as the predecessor's minloc,
// test/UTIL_PRE_SIM.f90:189-194) and the body force, recorded on the
// device after every step into a ring of cap records.
if_status Engine::monitor(int k, int nprobe, const double *xy, int cap)
{
    Slot *s = slot(k);
    if (!s || s->state == -2) return fail(IF_ERR_ARG, "monitor: empty slot");
    if (nprobe < 0 || nprobe > kMaxProbes) return fail(IF_ERR_ARG, "monitor: 0..32 probes");
    if (nprobe > 0 && !xy) return fail(IF_ERR_ARG, "monitor: null probe list");
    if (cap < 1) return fail(IF_ERR_ARG, "monitor: ring capacity must be >= 1");
    Monitor m{};
    m.nprobe = nprobe;
    m.cap = cap;
    auto nearest = [](const std::vector<double> &c, double x) {
        int best = 1;
        for (int i = 2; i + 1 < (int)c.size(); i++)
            if (std::fabs(c[i] - x) < std::fabs(c[best] - x)) best = i;
        return best;
    };
    for (int q = 0; q < nprobe; q++) {
        m.pi[q] = nearest(s->xc, xy[2 * q]);
        m.pj[q] = nearest(s->yc, xy[2 * q + 1]);
    }
    const size_t ntl = (size_t)s->ntx * s->nty;
    const size_t stride = 3 + 3 * (size_t)nprobe;
    // layout: the Monitor record, then the partials, the ring and the count,
    // each on a 256-byte boundary after the previous part's end
    const size_t o_part = (sizeof(Monitor) + 255) & ~(size_t)255;
    const size_t o_ring = o_part + ((sizeof(double) * 2 * ntl + 255) & ~(size_t)255);
    const size_t o_cnt = o_ring + ((sizeof(double) * stride * cap + 255) & ~(size_t)255);
    const size_t total = o_cnt + 256;
    IFF_HIP(hipStreamSynchronize(st_));
    if (s->mon_mem) {  // replaced
        (void)hipFree(s->mon_mem);
        s->mon_mem = nullptr;
        s->dev.mon = nullptr;
        used_ -= s->mon_bytes;
        s->bytes -= s->mon_bytes;
        s->mon_bytes = 0;
    }
    void *mem = nullptr;
    if (hipMalloc(&mem, total) != hipSuccess) return fail(IF_ERR_NOMEM, "monitor: device allocation failed");
    char *b = (char *)mem;
    m.part = (double *)(b + o_part);
    m.ring = (double *)(b + o_ring);
    m.count = (unsigned long long *)(b + o_cnt);
    const unsigned long long zero = 0;
    if (hipMemcpy(b, &m, sizeof(Monitor), hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(m.count, &zero, sizeof(zero), hipMemcpyHostToDevice) != hipSuccess) {
        (void)hipFree(mem);
        return fail(IF_ERR_HIP, "monitor: set-up copy failed");
    }
    s->mon_mem = mem;
    s->mon_bytes = total;  // charged to the case
    s->bytes += total;
    used_ += total;
    s->dev.mon = (Monitor *)mem;
    dirty_ = true;
    return IF_OK;
}

// The records written since the last read, oldest first (at most cap of
// them; older ones beyond the ring's capacity were overwritten: *dropped).
if_status Engine::monitor_read(int k, double *out, size_t cap, size_t *nrec, size_t *dropped)
{
    Slot *s = slot(k);
    if (!s || s->state == -2) return fail(IF_ERR_ARG, "monitor_read: empty slot");
    if (!s->mon_mem) return fail(IF_ERR_STATE, "monitor_read: no monitors on this case");
    Monitor m{};
    IFF_HIP(hipStreamSynchronize(st_));
    IFF_HIP(hipMemcpy(&m, s->mon_mem, sizeof(Monitor), hipMemcpyDeviceToHost));
    unsigned long long cnt = 0;
    IFF_HIP(hipMemcpy(&cnt, m.count, sizeof(cnt), hipMemcpyDeviceToHost));
    const size_t stride = 3 + 3 * (size_t)m.nprobe;
    const size_t held = (size_t)std::min<unsigned long long>(cnt, (unsigned long long)m.cap);
    const size_t n = out ? std::min(held, cap) : held;
    if (nrec) *nrec = n;
    if (dropped) *dropped = (size_t)(cnt - held) + (held - n);
    if (!out) return IF_OK;
    std::vector<double> ring((size_t)m.cap * stride);
    IFF_HIP(hipMemcpy(ring.data(), m.ring, ring.size() * sizeof(double), hipMemcpyDeviceToHost));
    // oldest held record: index cnt - held; the newest n of them are returned
    const unsigned long long first = cnt - n;
    for (size_t r = 0; r < n; r++) {
        const size_t slotr = (size_t)((first + r) % (unsigned long long)m.cap);
        std::memcpy(out + r * stride, ring.data() + slotr * stride, stride * sizeof(double));
    }
    const unsigned long long zero = 0;
    IFF_HIP(hipMemcpy(m.count, &zero, sizeof(zero), hipMemcpyHostToDevice));
    return IF_OK;
}

// One monitor record of a case's current fields, at its current time,
// without a step (the initial state, a restart, or fields set by the
// caller); the case's own kernels run on a one-case table.
if_status Engine::monitor_sample(int k)
{
    Slot *s = slot(k);
    if (!s || s->state == -2) return fail(IF_ERR_ARG, "monitor_sample: empty slot");
    if (!s->mon_mem) return fail(IF_ERR_STATE, "monitor_sample: no monitors on this case");
    CaseDev *tmp = nullptr;
    int2 *dtl = nullptr;
    std::vector<int2> tl(s->ntx * s->nty);
    for (int t = 0; t < (int)tl.size(); t++) tl[t] = make_int2(0, t);
    bool ok = hipMalloc(&tmp, sizeof(CaseDev)) == hipSuccess && hipMalloc(&dtl, sizeof(int2) * tl.size()) == hipSuccess;
    if (ok) {
        ok = hipMemcpyAsync(tmp, &s->dev, sizeof(CaseDev), hipMemcpyHostToDevice, st_) == hipSuccess &&
             hipMemcpyAsync(dtl, tl.data(), sizeof(int2) * tl.size(), hipMemcpyHostToDevice, st_) == hipSuccess;
        if (ok) launch_monitor(st_, vprec_, tmp, 1, dtl, (int)tl.size(), 1);
        ok = ok && hipStreamSynchronize(st_) == hipSuccess;  // tl and the scratch live until here
    }
    if (tmp) (void)hipFree(tmp);
    if (dtl) (void)hipFree(dtl);
    if (!ok) return fail(IF_ERR_HIP, "monitor_sample: device update failed");
    return IF_OK;
}

// ---------------------------------------------------------------------------
// Allocation-free monitor sample.  Its staging -- a copy of the case
// descriptor, the one-case tile table and the per-tile force partials --
// lives in a file-scope device array, so the call makes no device (or
// host) allocation.  Each GPU holds its own copy of the array until the
// process ends; its address is resolved once per device with
// hipGetSymbolAddress and kept.  Calls that overlap on one GPU share the
// array and race; every byte a call reads it has written first, so no
// clearing is needed between calls.
// ---------------------------------------------------------------------------
constexpr int kScratchMaxTiles = 1 << 16;   // interior tiles of 64 x 16 cells: up to 67 M cells
constexpr size_t kScratchDescBytes = (sizeof(CaseDev) + 255) & ~(size_t)255;
constexpr size_t kScratchTileBytes = sizeof(int2) * (size_t)kScratchMaxTiles;
constexpr size_t kScratchBytes = kScratchDescBytes + kScratchTileBytes + 2 * sizeof(double) * (size_t)kScratchMaxTiles;
static __device__ __align__(256) unsigned char scratch[kScratchBytes];

static std::mutex g_scratch_mu;
static void *g_scratch_dev[64] = {};   // device address of scratch, per device

static if_status scratch_ptr(void **out)
{
    int dev = 0;
    IFF_HIP(hipGetDevice(&dev));
    if (dev < 0 || dev >= 64) return fail(IF_ERR_NODEV, "noalloc: device index out of range");
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (!g_scratch_dev[dev]) {
        void *p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p)
            return fail(IF_ERR_HIP, std::string("noalloc: hipGetSymbolAddress(scratch) failed: ") +
                                        hipGetErrorString(e));
        g_scratch_dev[dev] = p;
    }
    *out = g_scratch_dev[dev];
    return IF_OK;
}

if_status Engine::monitor_sample_noalloc(int k)
{
    Slot *s = slot(k);
    if (!s || s->state == -2) return fail(IF_ERR_ARG, "monitor_sample_noalloc: empty slot");
    if (!s->mon_mem) return fail(IF_ERR_STATE, "monitor_sample_noalloc: no monitors on this case");
    const int nt = s->ntx * s->nty;
    if (nt > kScratchMaxTiles)
        return fail(IF_ERR_ARG, "monitor_sample_noalloc: the case has " + std::to_string(nt) +
                                    " tiles of 64 x 16 cells; the scratch array holds " +
                                    std::to_string(kScratchMaxTiles));
    void *base = nullptr;
    if_status rc = scratch_ptr(&base);
    if (rc != IF_OK) return rc;
    char *b = (char *)base;
    CaseDev *desc = (CaseDev *)b;
    int2 *tiles = (int2 *)(b + kScratchDescBytes);
    double *part = (double *)(b + kScratchDescBytes + kScratchTileBytes);
    IFF_HIP(hipMemcpyAsync(desc, &s->dev, sizeof(CaseDev), hipMemcpyHostToDevice, st_));
    launch_fill_tiles(st_, tiles, nt);
    launch_monitor(st_, vprec_, desc, 1, tiles, nt, 1, part);
    IFF_HIP(hipGetLastError());
    IFF_HIP(hipStreamSynchronize(st_));
    return IF_OK;
}

// Restart: load u, v, p (and the faces when given), then refresh the
// ghost-cell velocities and, without given faces, interpolate the faces
// from u and v -- the kernels and order of the case's initialisation, run
// on this case alone.
if_status Engine::restart(int k, const double *u, const double *v, const double *p, const double *uf,
                          const double *vf)
{
    Slot *s = slot(k);
    if (!s || s->state == -2) return fail(IF_ERR_ARG, "restart: empty slot");
    if (!u || !v || !p) return fail(IF_ERR_ARG, "restart: u, v and p are required");
    if ((uf == nullptr) != (vf == nullptr)) return fail(IF_ERR_ARG, "restart: give both uf and vf or neither");
    const size_t ncell = (size_t)s->nx * s->ny;
    if_status rc = set_field(k, IF_FIELD_U, u, ncell);
    if (rc == IF_OK) rc = set_field(k, IF_FIELD_V, v, ncell);
    if (rc == IF_OK) rc = set_field(k, IF_FIELD_P, p, ncell);
    if (rc == IF_OK && uf) rc = set_field(k, IF_FIELD_UF, uf, (size_t)(s->nx - 1) * (s->ny - 2));
    if (rc == IF_OK && vf) rc = set_field(k, IF_FIELD_VF, vf, (size_t)(s->nx - 2) * (s->ny - 1));
    if (rc != IF_OK) return rc;
    std::vector<Ctrl> all;
    rc = read_ctrl_sync(all);
    if (rc != IF_OK) return rc;
    const Ctrl ct = all[k];
    CaseDev *tmp = nullptr;
    int2 *dtl = nullptr;
    auto done = [&]() {
        if (tmp) (void)hipFree(tmp);
        if (dtl) (void)hipFree(dtl);
    };
    std::vector<int2> tl(s->ntx * s->nty);
    for (int t = 0; t < (int)tl.size(); t++) tl[t] = make_int2(0, t);
    if (hipMalloc(&tmp, sizeof(CaseDev)) != hipSuccess || hipMalloc(&dtl, sizeof(int2) * tl.size()) != hipSuccess) {
        done();
        return fail(IF_ERR_NOMEM, "restart: scratch allocation failed");
    }
    Ctrl on = ct;
    on.active = 1;  // the ghost kernels act on active cases
    bool ok = hipMemcpyAsync(tmp, &s->dev, sizeof(CaseDev), hipMemcpyHostToDevice, st_) == hipSuccess &&
              hipMemcpyAsync(dtl, tl.data(), sizeof(int2) * tl.size(), hipMemcpyHostToDevice, st_) == hipSuccess &&
              hipMemcpyAsync(s->dev.ctrl, &on, sizeof(Ctrl), hipMemcpyHostToDevice, st_) == hipSuccess;
    if (ok) {
        launch_ghost_vel(st_, vprec_, tmp, 1, s->gcap, 1);
        if (!uf) launch_faces(st_, vprec_, tmp, dtl, (int)tl.size(), 0, 0);
        ok = hipMemcpyAsync(s->dev.ctrl, &ct, sizeof(Ctrl), hipMemcpyHostToDevice, st_) == hipSuccess &&
             hipStreamSynchronize(st_) == hipSuccess;
    }
    done();
    if (!ok) return fail(IF_ERR_HIP, "restart: device update failed");
    return IF_OK;
}

if_status Engine::ghosts(int k, int32_t *out, if_ghost *recs, size_t cap, size_t *n)
{
    Slot *s = slot(k);
    if (!s || s->state == -2) return fail(IF_ERR_ARG, "ghosts: empty slot");
    std::vector<Ctrl> all;
    if_status rc = read_ctrl_sync(all);
    if (rc != IF_OK) return rc;
    const int ng = all[k].ng;
    if (n) *n = (size_t)ng;
    if (!out && !recs) return IF_OK;
    if ((size_t)ng > cap) return fail(IF_ERR_ARG, "ghosts: capacity too small");
    std::vector<GhostDev> g(ng);
    if (ng) IFF_HIP(hipMemcpy(g.data(), s->dev.gh, sizeof(GhostDev) * ng, hipMemcpyDeviceToHost));
    const int P = s->P, nx = s->nx;
    auto ref = [&](int id) { return (id % P) + (id / P) * nx; };
    for (int q = 0; q < ng; q++) {
        if (out) out[q] = ref(g[q].gidx);
        if (recs) {
            if_ghost &r = recs[q];
            r.gidx = ref(g[q].gidx);
            r.body = g[q].body;
            for (int m = 0; m < 4; m++) {
                r.node[m] = ref(g[q].node[m]);
                r.wd[m] = g[q].wd[m];
                r.wn[m] = g[q].wn[m];
            }
            r.cd = g[q].cd;
            r.lnx = g[q].lnx;
            r.lny = g[q].lny;
            r.bix = g[q].bix;
            r.biy = g[q].biy;
            r.ipx = g[q].ipx;
            r.ipy = g[q].ipy;
        }
    }
    return IF_OK;
}

}  // namespace iff
