// Internal definitions shared by the host runtime and the gfx950 kernels.
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>
#include <string>

#include "immerseflow/if_abi.h"

namespace iff {

// ---------------------------------------------------------------------------
// errors: the library reports through status codes + a thread-local message
// (the reference exits the process instead, globalVariables.cuh:91-107)
// ---------------------------------------------------------------------------
void set_error(const std::string &msg);
if_status fail(if_status code, const std::string &msg);

#define IFF_HIP(call)                                                        \
    do {                                                                     \
        hipError_t e_ = (call);                                              \
        if (e_ != hipSuccess)                                                \
            return ::iff::fail(IF_ERR_HIP, std::string(#call " failed: ") +  \
                                               hipGetErrorString(e_));       \
    } while (0)

// ---------------------------------------------------------------------------
// cell flags (one byte per cell)
// ---------------------------------------------------------------------------
enum : uint8_t {
    F_FLUID = 1,   // iBlank == 1
    F_GHOST = 2,   // solid interior cell with a fluid 4-neighbour
    F_OPEN_E = 4,  // east face joins two fluid interior cells
    F_OPEN_W = 8,
    F_OPEN_N = 16,
    F_OPEN_S = 32
};

// ---------------------------------------------------------------------------
// HBM layout.  Every cell field of a case is an nx x ny array (ghost ring
// included) with a row pitch P that is a multiple of 64 elements.  The field
// pointer is pre-offset so that interior column i=1 of every row starts on
// a 256-byte boundary for 4- and 8-byte elements (and 64 B for bytes):
// element (i,j) is f[i + j*P].
// ---------------------------------------------------------------------------
constexpr int kPitchAlign = 64;
constexpr int kTileX = 64;   // tile width (one wave64 row)

inline int pitch_for(int nx) { return ((nx + kPitchAlign - 1) / kPitchAlign) * kPitchAlign; }
template <typename T> constexpr int lead_elems() { return 256 / (int)sizeof(T) - 1; }
template <typename T> inline size_t field_alloc_bytes(int P, int ny)
{
    size_t n = (size_t)P * ny + lead_elems<T>() + 64;
    return ((n * sizeof(T) + 255) / 256) * 256;
}

// ---------------------------------------------------------------------------
// device-side per-case records
// ---------------------------------------------------------------------------
struct BodyDev {
    int kind;
    double xc, yc, r, a, b, c, s;
    int nvert;
    const double *vx, *vy;  // absolute vertex coordinates
    double ub, vb, axb, ayb;
};

struct GhostDev {
    int32_t gidx;     // pitched cell index
    int32_t body;
    int32_t node[4];  // pitched indices (the ghost itself where weight 0)
    double wd[4], cd, wn[4], lnx, lny;
    double bix, biy, ipx, ipy;
};

// Per-axis metric vectors, length n (nx or ny), all computed on the host
// with the operation order of the reference kernels.
struct AxisD {
    const double *d;   // width
    const double *ip;  // 1/(d_i + d_{i+1})
    const double *im;  // 1/(d_i + d_{i-1})
    const double *cs;  // (dt/d_i)*conv            (ADSource)
    const double *rd;  // 1.0/d_i                   (divergence)
    const double *g;   // 2.0/(d_i + d_{i+1})       (face pressure gradient)
    const double *cp;  // 2/(d_i(d_i+d_{i+1}))      (Poisson)
    const double *cm;  // 2/(d_i(d_i+d_{i-1}))
    const double *ap;  // (dt/Re)*cp                (viscous Jacobi)
    const double *am;  // (dt/Re)*cm
    const double *ad;  // x: 1.0 + (dt/Re)*(cp+cm); y: (dt/Re)*(cp+cm)
};
constexpr int kAxisDVecs = 11;
struct AxisF {         // FP32 copies for the mixed-precision sweeps
    const float *d, *ip, *im, *cs, *ap, *am, *ad, *cp, *cm;
};
constexpr int kAxisFVecs = 9;

// Device control block of one case: loop state that lets kernels end a
// case's iterations on the device (no host round trip per sweep).
struct Ctrl {
    int active;        // case participates in launches
    int failed;        // non-finite residual seen
    int steps_done, steps_target;
    int ad_done, ad_iter, ad_base;   // u,v live in buffer ad_base
    int ppe_done, ppe_sweeps, ppe_base;
    int ng;            // live ghost-cell count (rebuilt on the device for moving bodies)
    int ng_total;      // ghost cells found by the last build (may exceed capacity)
    double ad_res, ppe_res, div_max;
    double t;          // simulated time
    long long tot_ad, tot_ppe;
    double max_div;
};

// Synchronisation state of the persistent Poisson sweep (one per engine):
// a sticky error flag set when a wait times out, and the wait limit in
// ticks of the 100 MHz real-time counter.
struct GridBar {
    unsigned err, pad;
    unsigned long long timeout;
};

// Everything a kernel needs to know about one case.
// Per-step monitors of a case (probes and body force), in device memory.
constexpr int kMaxProbes = 32;
struct Monitor {
    int nprobe, cap;                   // probes; ring capacity in records
    int pi[kMaxProbes], pj[kMaxProbes];  // probe cells (array indices)
    double *part;                      // Fx, Fy partial per interior tile
    double *ring;                      // cap records of 3 + 3*nprobe doubles
    unsigned long long *count;         // records written since the last read
};

// Line solve partition of a row of n interior cells: S = ceil(n / B) <= 64
// blocks of B = max(2, ceil(n / 64)) cells; the last cell of every block
// but the last is a separator.  A row's factors take line_row_stride(B)
// elements: five planes of B x 64 per-cell values and 13 x 64 for the
// separators' cyclic reduction.  The sweep kernel stages a row in B x 65
// LDS elements, which bounds B (and so n <= 64 kLineMaxB).
constexpr int kLinePlanes = 5, kLinePcrLevels = 6, kLinePcr = 2 * kLinePcrLevels + 1;
constexpr int kLineMaxB = 126;
__host__ __device__ inline void line_partition(int n, int *B, int *S)
{
    int b = (n + 63) / 64;
    if (b < 2) b = 2;
    *B = b;
    *S = (n + b - 1) / b;
}
__host__ __device__ inline size_t line_row_stride(int B) { return (size_t)(kLinePlanes * B + kLinePcr) * 64; }

struct CaseDev {
    int nx, ny, P;
    int ntx, nty;          // tiles of the interior (kTileX x tile_y)
    int tile_y;
    int gcap, nbody;       // ghost-record capacity, bodies
    int moving;            // bodies move: reclassify every step
    int prec;              // velocity storage: IF_FP64 (double u, v) or IF_MIXED (float)
    int ppe_solver;
    int ppe_m;             // sweeps per Poisson launch of this case's solver
    int ad_itermax, ppe_itermax, check_every;
    double dt, rdt, omega, ppe_tol, ad_tol, ubc, vbc;
    double mu;             // 1/Re
    double ad_omega;       // w_AD relaxation of the viscous sweep (1: none)
    float omega_f;
    AxisD x, y;
    AxisF xf, yf;
    const double *xc, *yc; // cell centres (ghost mirrors included)
    void *u[2], *v[2], *sx, *sy;   // double or float (prec)
    double *uf[2], *vf[2];         // face velocities: [0] = U^n, [1] = U*
    double *p[2], *rhs;
    float *e[2], *r32;             // mixed Poisson error / residual
    float *w32;                    // mixed: each cell's FP32 SOR weight om / sa (0: not updated), formed once per step
    float *wtab;                   // mixed on a grid whose FP32 coefficients are each one value: the SOR
                                   // weight of each of the 64 patterns of flag bits 0-5; null: per-cell
                                   // weights (w32)
    double *wtab64;                // FP64 Poisson on a grid whose coefficients are each one value: the same
                                   // table in FP64; null: each launch forms the weights from the flags
    int uni;                       // mixed red-black case with one FP32 coefficient value per axis on every
                                   // face that can open (aE = aW = uax, aN = aS = uax * uratio): the FP32
                                   // sweeps take the five-operation form (sor_cell_u) with K = w uax, and
                                   // r32 holds the residual times uinv = 1 / uax; only such cases get wtab
                                   // (then the K of each flag pattern)
    float uax, uratio, uinv;
    void *lf;                      // line solve factors (Poisson precision), line_row_stride(lB) per row; line cases only
    int lB, lS;                    // line solve partition of a row: S blocks of B cells
    uint8_t *flags;
    int8_t *bodyid;
    GhostDev *gh;
    BodyDev *bodies;
    Ctrl *ctrl;
    double *part;                  // one partial per interior tile
    int *gcount;                   // ghost-compaction scratch (per 1024 cells)
    Monitor *mon;                  // null: no monitors
};

}  // namespace iff
