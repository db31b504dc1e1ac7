// Register-tile red-black SOR for gfx950.
//
// A workgroup of NW (4, 8 or 16) wave64s owns a tile of 128 columns by NW*NRW rows
// (halo H = 2M on every side included): lane l of every wave holds the
// column pair (c0+2l, c0+2l+1) -- one red and one black cell of each row --
// and wave w holds rows w*NRW .. w*NRW+NRW-1, all in registers.  In a
// half-sweep each lane therefore updates exactly one cell per row, and
// which of its two cells that is depends only on the row, so the whole
// wave follows one path with every lane busy:
//   * east/west neighbours: the lane's other cell, or the adjacent lane's
//     through a DPP wave shift;
//   * north/south neighbours: the lane's own registers, except at the
//     wave's first and last row, whose neighbours come from the adjacent
//     waves through a small LDS exchange (one barrier per half-sweep).
// FP64 (the oracle's update, bit for bit): rows whose cells are all fluid
// with four open faces (nearly all of them) take a fast path whose
// coefficients are the metric itself (x: per lane, y: per row); a row
// holding any other cell evaluates every lane with its face flags.  FP32
// (the mixed path's error sweeps): every cell takes one branch-free form,
// the general one or, on a grid with one coefficient value per axis, the
// five-operation one.  Every kernel computes the same expression per case
// (sor.h), so the result does not depend on the path, the tile shape, the
// kernel or the batch a case is in.  The cells' weights come from a table
// indexed by the flag byte (uniform grids) or per cell (FP32: loaded once
// per launch; FP64: formed once per launch); WM, the launch's weight mode,
// picks the form at compile time.
//
// M sweeps per launch on the tile with its 2M halo (temporal blocking):
// after half-sweep h only the tile plus a ring of 2M-1-h cells is still
// exact, so after 2M half-sweeps the interior (TW x TY) is, and only it is
// stored.  Reads come from buffer src and writes go to dst (ping-pong), so
// tiles never read halos that other tiles are writing.  HBM traffic is one
// read of the unknown, the right side and the flags (or FP32 weights),
// plus the halo, and one write of the unknown per M sweeps.
#include "common.h"
#include "kernels.h"
#include "sor.h"

#include <algorithm>
#include <utility>

namespace iff {

namespace {

template <typename T, int NRW>
struct RegTile {
    T pa[NRW], pb[NRW];  // unknown, slots a (even column) and b (odd column)
    T ra[NRW], rb[NRW];  // right side (FP32: w r once the weights are formed)
    T cyp[NRW], cym[NRW];
    T wa[NRW], wb[NRW];  // the cell's weight om/sa, 0 for a cell the sweep does not update
};

// A wave-uniform value held in scalar registers (the row coefficients are
// the same in every lane; in SGPRs they cost no vector registers and enter
// the multiplies as scalar operands).
__device__ __forceinline__ float uniform(float x)
{
    return __int_as_float(__builtin_amdgcn_readfirstlane(__float_as_int(x)));
}
__device__ __forceinline__ double uniform(double x)
{
    const long long b = __double_as_longlong(x);
    const int lo = __builtin_amdgcn_readfirstlane((int)(b & 0xffffffffLL));
    const int hi = __builtin_amdgcn_readfirstlane((int)(b >> 32));
    return __longlong_as_double(((long long)hi << 32) | (unsigned int)lo);
}

__device__ __forceinline__ void *uniform_ptr(void *p)
{
    const unsigned long long b = (unsigned long long)p;
    const unsigned lo = __builtin_amdgcn_readfirstlane((unsigned)b);
    const unsigned hi = __builtin_amdgcn_readfirstlane((unsigned)(b >> 32));
    return (void *)(((unsigned long long)hi << 32) | lo);
}

// a if face bit F of flags f is set, else +0: the bit, sign-extended to a
// full mask, ANDed with a's bits (two VALU operations, no compare).
template <typename T, int F>
__device__ __forceinline__ T face_coef(int f, T a)
{
    const int m = __builtin_amdgcn_sbfe(f, __builtin_ctz(F), 1);
    if constexpr (sizeof(T) == 8)
        return __longlong_as_double(__double_as_longlong(a) & (long long)m);
    else
        return __int_as_float(__float_as_int(a) & m);
}

// One cell of a half-sweep, FP32 (the error equation of the mixed path).
// Every cell takes this one branch-free form: the error is zero in every
// non-fluid cell (each solve starts from zero and never updates them), and
// a face is closed exactly when it lies on the domain boundary or next to a
// non-fluid cell, so the metric coefficients (zeroed on boundary faces)
// give the same neighbour sum as the flagged ones.  The closed faces enter
// through the cell's weight w = om / sa (sa over the open faces), formed
// once per launch with wr = w r; w = 0 leaves a non-fluid cell at zero.
// (a1, p1) is the neighbour held by the adjacent lane: the lane shift
// folds into that product (v_mul_f32 with DPP).
__device__ __forceinline__ float reg_cell_w(float pc, float wr, float a1, float p1, float a2, float p2, float pN,
                                            float pS, float cyp, float cym, float w, float c1)
{
    return sor_cell_w(pc, wr, a1, p1, a2, p2, cyp, pN, cym, pS, w, c1);
}

// One cell of a half-sweep, FP64, with the cell's weight w = om / sa formed
// once per launch (w = 0: a cell the sweep does not update).  Fast path
// (FAST): a fluid cell whose faces are all open except those on the domain
// boundary; the coefficients are the metric with the boundary faces' entries
// zeroed, so both paths give the same bits for such a cell.  Flag path: the
// oracle's rule, coefficient zero on a closed face, no update of a non-fluid
// cell or one with ap = 0.
template <bool FAST>
__device__ __forceinline__ double reg_cell(int f, double pc, double r, double pE, double pW, double pN, double pS,
                                           double cxp, double cxm, double cyp, double cym, double w, double c1)
{
    if constexpr (FAST) {
        const double sum = (cxp * pE + cxm * pW) + (cyp * pN + cym * pS);
        return sor_cell_w64(pc, r, sum, w, c1);
    } else {
        const double aE = face_coef<double, F_OPEN_E>(f, cxp);
        const double aW = face_coef<double, F_OPEN_W>(f, cxm);
        const double aN = face_coef<double, F_OPEN_N>(f, cyp);
        const double aS = face_coef<double, F_OPEN_S>(f, cym);
        const double sum = (aE * pE + aW * pW) + (aN * pN + aS * pS);
        const double nv = sor_cell_w64(pc, r, sum, w, c1);
        return w != 0.0 ? nv : pc;
    }
}

// One row's cells of a half-sweep.  c1 = 1 - om; the cells' weights are in R
// (UNI, FP32: K = w ax, the five-operation form with ratio = ay / ax).
template <typename T, int NRW, int Q, bool FAST, int k, bool UNI = false>
__device__ __forceinline__ void reg_row(RegTile<T, NRW> &R, int f, T sva, T svb, T nva, T nvb, T cxpa,
                                        T cxma, T cxpb, T cxmb, T c1, T ratio = (T)0)
{
    constexpr int ks = k > 0 ? k - 1 : 0, kn = k < NRW - 1 ? k + 1 : 0;
    if constexpr ((k & 1) == Q) {
        const T pS = k == 0 ? sva : R.pa[ks];
        const T pN = k == NRW - 1 ? nva : R.pa[kn];
        const T pE = R.pb[k], pW = from_west(R.pb[k]);
        if constexpr (sizeof(T) == 4 && UNI)
            R.pa[k] = sor_cell_u(R.pa[k], R.ra[k], pW, pE, pN, pS, R.wa[k], ratio, c1);
        else if constexpr (sizeof(T) == 4)  // slot a: odd column, west neighbour in lane l-1
            R.pa[k] = reg_cell_w(R.pa[k], R.ra[k], cxma, pW, cxpa, pE, pN, pS, R.cyp[k], R.cym[k], R.wa[k], c1);
        else
            R.pa[k] = reg_cell<FAST>(f & 0xff, R.pa[k], R.ra[k], pE, pW, pN, pS, cxpa, cxma, R.cyp[k], R.cym[k],
                                     R.wa[k], c1);
    } else {
        const T pS = k == 0 ? svb : R.pb[ks];
        const T pN = k == NRW - 1 ? nvb : R.pb[kn];
        const T pE = from_east(R.pa[k]), pW = R.pa[k];
        if constexpr (sizeof(T) == 4 && UNI)
            R.pb[k] = sor_cell_u(R.pb[k], R.rb[k], pW, pE, pN, pS, R.wb[k], ratio, c1);
        else if constexpr (sizeof(T) == 4)  // slot b: even column, east neighbour in lane l+1
            R.pb[k] = reg_cell_w(R.pb[k], R.rb[k], cxpb, pE, cxmb, pW, pN, pS, R.cyp[k], R.cym[k], R.wb[k], c1);
        else
            R.pb[k] = reg_cell<FAST>(f >> 8, R.pb[k], R.rb[k], pE, pW, pN, pS, cxpb, cxmb, R.cyp[k], R.cym[k],
                                     R.wb[k], c1);
    }
}

// Half-sweep on the wave's rows.  Q: rows k with (k & 1) == Q update slot
// a, the others slot b.  sv/nv: the rows just south/north of the wave.
// ALLFAST: every row of the wave takes the fast path (no branches);
// otherwise smask marks the rows that take the flag path.
template <typename T, int NRW, int Q, bool ALLFAST>
__device__ __forceinline__ void reg_half(RegTile<T, NRW> &R, const uint16_t (*fl)[64], int lane,
                                         unsigned smask, T sva, T svb, T nva, T nvb,
                                         T cxpa, T cxma, T cxpb, T cxmb, T c1)
{
    if constexpr (ALLFAST) {
        [&]<int... K>(std::integer_sequence<int, K...>) {
            (reg_row<T, NRW, Q, true, K>(R, 0, sva, svb, nva, nvb, cxpa, cxma, cxpb, cxmb, c1), ...);
        }(std::make_integer_sequence<int, NRW>{});
    } else {
        int f[NRW];  // flags of every row, read in one batch
#pragma unroll
        for (int k = 0; k < NRW; k++) f[k] = fl[k][lane];
        [&]<int... K>(std::integer_sequence<int, K...>) {
            (([&] {
                 if ((smask >> K) & 1u)
                     reg_row<T, NRW, Q, false, K>(R, f[K], sva, svb, nva, nvb, cxpa, cxma, cxpb, cxmb, c1);
                 else
                     reg_row<T, NRW, Q, true, K>(R, 0, sva, svb, nva, nvb, cxpa, cxma, cxpb, cxmb, c1);
             }()),
             ...);
        }(std::make_integer_sequence<int, NRW>{});
    }
}

// Neighbour synchronisation of the persistent sweep.  Every tile has a
// stage counter (flags[t], advanced by the number of stages per launch by
// every tile, so all counters agree when a launch starts).  A block
// publishes "stage l done" after its border strips are written, then waits
// until the 8 neighbouring tiles of its case have published the same stage.
// Strips are written and halos read at device scope (sc1: write-through,
// and loads that miss in the per-XCD L2), so no cache flush is needed; the
// waits are bounded: past bar->timeout ticks of the 100 MHz real-time
// counter bar->err is set, every later wait returns at once and every
// block stops at its next stage boundary, so the kernel always ends; the
// host checks err after the step.
constexpr int kDev = 16;                             // cache policy: sc1 (device scope)
constexpr unsigned kRsrc3 = 0x00020000;              // buffer resource word 3 (gfx9)
// s_waitcnt operand (gfx9 encoding) waiting for vmcnt = 0 only: vmcnt
// bits 3:0 and 15:14 zero, expcnt (6:4) and lgkmcnt (11:8) at their maxima
constexpr int kWaitVmcnt0 = 0x0F70;

// wait (lanes 0-7 of wave 0, one neighbour each; nb < 0: none) until
// flags[nb] has reached target
__device__ __forceinline__ void wait_neighbours(const unsigned *flags, int nb, unsigned target, GridBar *bar)
{
    if (nb >= 0) {
        const unsigned long long t0 = __builtin_amdgcn_s_memrealtime();
        const unsigned long long tmo = bar->timeout;
        while ((int)(__hip_atomic_load(&flags[nb], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) - target) < 0) {
            if (__hip_atomic_load(&bar->err, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != 0) break;
            if (__builtin_amdgcn_s_memrealtime() - t0 > tmo) {
                __hip_atomic_store(&bar->err, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                break;
            }
            __builtin_amdgcn_s_sleep(1);
        }
    }
}

}  // namespace

// PERS: one launch runs the chunk's K sweeps as stages of M
// sweeps.  The tile stays in registers; after each stage a block writes the
// border strips (depth 2M) of its stored region, waits for its neighbouring
// tiles to have written theirs, and reloads its halo ring from them.  Stage
// L leaves its output where launch L of the chunked form does
// (FP32 e[(L + 1) & 1], FP64 p[(ppe_base + L + 1) & 1]), the last stage of
// a case the whole stored region, so
// the checks read the same buffer either way.  A block is at most one
// stage ahead of its neighbours, so the two buffers never hold data a
// neighbour still needs when they are overwritten.
//
// The second launch bound holds the FP32 per-stage 128 x 128 blocks (16
// waves x 8 rows) at 8 waves per SIMD, i.e. 64 VGPRs and two blocks per CU:
// left free, some of them take a few more registers and run at one block
// per CU, a third slower.
template <typename T, int M, int NW, int NRW, bool PERS, int WM>
__global__ __launch_bounds__(64 * NW, (sizeof(T) == 4 && !PERS && NW == 16 && NRW == 8 && WM != 2) ? 8 : 1)
void k_rbsor_reg(const CaseDev *cs, const int2 *tiles, int ntiles, int ntiles_pad, int s, int K, GridBar *bar,
                 unsigned *flags)
{
    constexpr int H = 2 * M;
    constexpr int TW = 128 - 2 * H;   // stored columns per tile
    constexpr int TY = NW * NRW - 2 * H;  // stored rows per tile
    static_assert(TY > 0 && NRW <= 32, "tile shape");
    using V2 = typename VecT<T, 2>::type;
    // [parity][1 + wave][first row, last row][lane]; slots 0 and NW + 1 stay
    // zero (the rows beyond the first and last wave), so every wave reads its
    // neighbour rows unconditionally
    __shared__ V2 xch[2][NW + 2][2][64];
    // Face flags (slot a in bits 0-7, slot b in 8-15), read by the rows that
    // take the flag path; kept in LDS rather than registers.
    constexpr int kFlw = sizeof(T) == 8 ? NW : 1;  // FP64 only
    __shared__ uint16_t sfl[kFlw][NRW][64];

    const int t = tile_of_block(ntiles_pad);
    const int nstage = (K + M - 1) / M;  // PERS: stages of the launch
    if constexpr (PERS) {
        bool run = false;
        if (t < ntiles) {
            const CaseDev &cr = cs[tiles[t].x];
            const Ctrl *cc = cr.ctrl;
            run = cr.ppe_solver == IF_PPE_RBSOR && cc->active && !cc->ppe_done && cr.ppe_itermax > cc->ppe_sweeps;
        }
        if (!run) {  // whole workgroup: nobody waits on this tile, advance its counter
            if (t < ntiles && threadIdx.x == 0) {
                const unsigned f0 = __hip_atomic_load(&flags[t], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                __hip_atomic_store(&flags[t], f0 + (unsigned)nstage, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            }
            return;
        }
    } else {
        if (t >= ntiles) return;  // whole workgroup
    }
    const int2 tl = tiles[t];
    const CaseDev &c = cs[tl.x];
    const Ctrl *ct = c.ctrl;
    if (c.ppe_solver != IF_PPE_RBSOR || !ct->active || ct->ppe_done) return;  // (PERS: never taken)
    int kexec = c.ppe_itermax - ct->ppe_sweeps;
    if (kexec > K) kexec = K;
    if (s >= kexec) return;  // (PERS: s = 0 < kexec)
    int meff = kexec - s;
    if (meff > M) meff = M;
    const int nst = 2 * meff;  // half-sweeps of this launch
    const int L = s / M;       // launch index inside the chunk

    const int nx = c.nx, ny = c.ny, P = c.P;
    const T *src;
    T *dst;
    const T *rhs;
    const T *cxp_, *cxm_, *cyp_, *cym_;
    bool zero_src = false;
    T om;
    if constexpr (sizeof(T) == 8) {
        src = c.p[(ct->ppe_base + L) & 1];
        dst = c.p[(ct->ppe_base + L + 1) & 1];
        rhs = c.rhs;
        cxp_ = c.x.cp; cxm_ = c.x.cm; cyp_ = c.y.cp; cym_ = c.y.cm;
        om = c.omega;
    } else {
        src = c.e[L & 1];
        dst = c.e[(L + 1) & 1];
        zero_src = (L == 0);
        rhs = c.r32;
        cxp_ = c.xf.cp; cxm_ = c.xf.cm; cyp_ = c.yf.cp; cym_ = c.yf.cm;
        om = c.omega_f;
    }
    const int ntxs = (nx - 2 + TW - 1) / TW;
    const int tx = tl.y % ntxs, ty = tl.y / ntxs;
    const int lane = threadIdx.x & 63;
    const int w = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int c0 = 1 + tx * TW - H;            // first tile column
    const int jt0 = 1 + ty * TY - H + w * NRW;  // first row of this wave
    const int ga = c0 + 2 * lane;               // slot a column; slot b is ga + 1
    const bool ina = ga >= 1 && ga <= nx - 2;
    const bool inb = ga + 1 >= 1 && ga + 1 <= nx - 2;
    // x coefficients of the fast path: zero on the domain boundary faces
    // and outside the interior.  A slot outside the interior has weight 0:
    // its value stays finite, is never stored and is only ever multiplied
    // by the zero coefficient of a boundary face.
    const T cxpa = (ina && ga != nx - 2) ? G(cxp_)[ga] : (T)0;
    const T cxma = (ina && ga != 1) ? G(cxm_)[ga] : (T)0;
    const T cxpb = (inb && ga + 1 != nx - 2) ? G(cxp_)[ga + 1] : (T)0;
    const T cxmb = (inb && ga + 1 != 1) ? G(cxm_)[ga + 1] : (T)0;
    // faces a fluid cell of each slot has open when it is a fast-path cell
    const int xa = (ga != nx - 2 ? F_OPEN_E : 0) | (ga != 1 ? F_OPEN_W : 0);
    const int xb = (ga + 1 != nx - 2 ? F_OPEN_E : 0) | (ga + 1 != 1 ? F_OPEN_W : 0);
    constexpr int kFaces = F_FLUID | F_OPEN_E | F_OPEN_W | F_OPEN_N | F_OPEN_S;

    // WM, the weight mode of the launch (kernels rather than runtime
    // branches, which would hold both forms' registers at once):
    //   1: every case has a weight table (wtab / wtab64): the weights are
    //      looked up by flag pattern instead of loaded per cell (FP32: w32)
    //      or formed with a division on every launch (FP64); FP32 takes the
    //      five-operation form (every such case is uni);
    //   0: per-cell weights, the general form;
    //   2 (FP32): per-cell weights, the form each case's uni flag picks (a
    //      batch mixing uniform and stretched grids).
    RegTile<T, NRW> R;
    int lf[NRW];  // FP64: flags, slot a in bits 0-7, slot b in bits 8-15
    // FP32: the cells' weights om / sa, formed once per step (k_weights32),
    // or on a grid with one value per coefficient the weight of the cell's
    // flag pattern: lane q holds the table's entry q, and each cell reads
    // its entry across the wave (ds_bpermute)
    const float *w32 = c.w32;
    static_assert(WM == 0 || WM == 1 || (WM == 2 && sizeof(T) == 4), "weight mode");
    constexpr bool wuni = WM == 1;
    // FP32 with one coefficient value per axis: the five-operation update,
    // K = w ax held in place of w
    const bool uni = sizeof(T) == 4 && (WM == 1 || (WM == 2 && c.uni));
    const T uax = sizeof(T) == 4 ? (T)c.uax : (T)0, uratio = sizeof(T) == 4 ? (T)c.uratio : (T)0;
    int wtv = 0, wtv_hi = 0;  // FP32: K = w ax; FP64: w (low and high halves)
    if constexpr (wuni) {
        if constexpr (sizeof(T) == 4) {
            wtv = __float_as_int(G(c.wtab)[lane]);
        } else {
            const double d = G(c.wtab64)[lane];
            wtv = __double2loint(d);
            wtv_hi = __double2hiint(d);
        }
    }
    // Loads.  A tile lying inside the interior (the common case) loads
    // every row with aligned 8-byte vectors and no per-lane tests.
    const bool tile_in = c0 >= 1 && c0 + 127 <= nx - 2 && jt0 >= 1 && jt0 + NRW - 1 <= ny - 2;
    if (tile_in) {
#pragma unroll
        for (int k = 0; k < NRW; k++) {
            const size_t ro = (size_t)(jt0 + k) * P;  // wave-uniform row offset
            const V2 pv = zero_src ? V2{(T)0, (T)0} : gload_vec<V2>(src + ro + ga);
            const V2 rv = gload_vec<V2>(rhs + ro + ga);
            R.pa[k] = pv[0];
            R.pb[k] = pv[1];
            R.ra[k] = rv[0];
            R.rb[k] = rv[1];
            if constexpr (sizeof(T) == 4) {
                if constexpr (wuni) {
                    const u8x2 fb = gload_vec<u8x2>(c.flags + ro + ga);
                    lf[k] = (int)fb[0] | ((int)fb[1] << 8);
                } else {
                    const V2 wv = gload_vec<V2>(w32 + ro + ga);
                    R.wa[k] = wv[0];
                    R.wb[k] = wv[1];
                    lf[k] = 0;
                }
            } else {
                const u8x2 fb = gload_vec<u8x2>(c.flags + ro + ga);
                lf[k] = (int)fb[0] | ((int)fb[1] << 8);
            }
        }
    } else {
#pragma unroll
        for (int k = 0; k < NRW; k++) {
            const int gj = jt0 + k;
            const bool rowin = gj >= 0 && gj < ny;
            const bool rowint = gj >= 1 && gj <= ny - 2;
            T a = (T)0, b = (T)0, x = (T)0, y = (T)0, wa = (T)0, wb = (T)0;
            int f = 0;
            if (rowin) {
                const size_t o = (size_t)gj * P + ga;
                if (!zero_src) {
                    if (ga >= 0 && ga + 1 < nx) {
                        const V2 v = gload_vec<V2>(src + o);
                        a = v[0];
                        b = v[1];
                    } else {
                        if (ga >= 0 && ga < nx) a = G(src)[o];
                        if (ga + 1 >= 0 && ga + 1 < nx) b = G(src)[o + 1];
                    }
                }
                if (rowint) {
                    if (ina && inb) {
                        const V2 v = gload_vec<V2>(rhs + o);
                        x = v[0];
                        y = v[1];
                        if (sizeof(T) == 4 && !wuni) {
                            const V2 wv = gload_vec<V2>(w32 + o);
                            wa = wv[0];
                            wb = wv[1];
                        } else {
                            const u8x2 fb = gload_vec<u8x2>(c.flags + o);
                            f = (int)fb[0] | ((int)fb[1] << 8);
                        }
                    } else {
                        if (ina) {
                            x = G(rhs)[o];
                            if (sizeof(T) == 4 && !wuni)
                                wa = G(w32)[o];
                            else
                                f = G(c.flags)[o];
                        }
                        if (inb) {
                            y = G(rhs)[o + 1];
                            if (sizeof(T) == 4 && !wuni)
                                wb = G(w32)[o + 1];
                            else
                                f |= (int)G(c.flags)[o + 1] << 8;
                        }
                    }
                }
            }
            R.pa[k] = a;
            R.pb[k] = b;
            R.ra[k] = x;
            R.rb[k] = y;
            if constexpr (sizeof(T) == 4) {
                R.wa[k] = wa;
                R.wb[k] = wb;
            }
            lf[k] = f;
        }
    }

    // Per-row coefficients and (FP64) the fast-path test and per-cell
    // weights.  Rows outside the interior take the fast path with zero y
    // coefficients and weight 0: their values stay finite, are never stored
    // and reach interior cells only through the zero coefficients of the
    // domain's boundary faces.  FP32 weights come from k_weights32 (0
    // outside the interior and for kept cells); wr = w r is formed here.
    unsigned smask = 0;  // FP64: rows holding a cell off the fast path
#pragma unroll
    for (int k = 0; k < NRW; k++) {
        const int gj = jt0 + k;
        const bool rowint = gj >= 1 && gj <= ny - 2;
        const T yp = (rowint && gj != ny - 2) ? G(cyp_)[gj] : (T)0;
        const T ym = (rowint && gj != 1) ? G(cym_)[gj] : (T)0;
        R.cyp[k] = uniform(yp);
        R.cym[k] = uniform(ym);
        if constexpr (sizeof(T) == 4) {
            // the five-operation form holds K = w ax in place of w; r32
            // holds r' = r / ax for such a case, so K r' stands for w r
            if constexpr (wuni) {  // a cell outside the interior has flags 0: weight 0
                R.wa[k] = __int_as_float(__builtin_amdgcn_ds_bpermute((lf[k] & 63) << 2, wtv));
                R.wb[k] = __int_as_float(__builtin_amdgcn_ds_bpermute(((lf[k] >> 8) & 63) << 2, wtv));
            } else if (uni) {
                R.wa[k] = R.wa[k] * uax;
                R.wb[k] = R.wb[k] * uax;
            }
            R.ra[k] = R.wa[k] * R.ra[k];
            R.rb[k] = R.wb[k] * R.rb[k];
        } else {
            const int f = lf[k];
            sfl[w][k][lane] = (uint16_t)f;
            bool offfast = false;
            if (rowint) {
                const int y4 = F_FLUID | (gj != ny - 2 ? F_OPEN_N : 0) | (gj != 1 ? F_OPEN_S : 0);
                const int ea = y4 | xa, eb = y4 | xb;  // the fast-path pattern of each slot
                const bool fa = !ina || (ea != F_FLUID && (f & kFaces) == ea);
                const bool fb = !inb || (eb != F_FLUID && ((f >> 8) & kFaces) == eb);
                offfast = __any(!(fa && fb));
            }
            if (offfast) smask |= 1u << k;
            // the weight om / sa over the open faces (for a fast-path cell the
            // same bits as om / (cx + cy)); 0 for a cell the sweep keeps:
            // outside the interior, not fluid, or without an open face.  With
            // a table: the entry of the cell's flag pattern, formed with the
            // same expression (a cell outside the interior has flags 0)
            const int fa = f & 0xff, fb = f >> 8;
            if constexpr (wuni) {
                auto look = [&](int q) {
                    const int i = (q & 63) << 2;
                    return __hiloint2double(__builtin_amdgcn_ds_bpermute(i, wtv_hi),
                                            __builtin_amdgcn_ds_bpermute(i, wtv));
                };
                R.wa[k] = look(fa);
                R.wb[k] = look(fb);
            } else {
                const T sa_a = (face_coef<T, F_OPEN_E>(fa, cxpa) + face_coef<T, F_OPEN_W>(fa, cxma)) +
                               (face_coef<T, F_OPEN_N>(fa, yp) + face_coef<T, F_OPEN_S>(fa, ym));
                const T sa_b = (face_coef<T, F_OPEN_E>(fb, cxpb) + face_coef<T, F_OPEN_W>(fb, cxmb)) +
                               (face_coef<T, F_OPEN_N>(fb, yp) + face_coef<T, F_OPEN_S>(fb, ym));
                R.wa[k] = (rowint && (fa & F_FLUID) && sa_a != (T)0) ? sor_weight64(sa_a, om) : (T)0;
                R.wb[k] = (rowint && (fb & F_FLUID) && sa_b != (T)0) ? sor_weight64(sa_b, om) : (T)0;
            }
        }
    }
    const T c1 = (T)1 - om;

    const int Z = (c0 + jt0) & 1;  // parity of slot a in the wave's first row
    // Half-sweep h has colour h & 1, and slot a of row k has parity Z + k,
    // so rows with (k & 1) == Q update slot a where Q = (Z ^ h) & 1.  A
    // wave of fast-path rows (FP32: every wave) runs branch-free.
    // FP32: each half-sweep first updates the rows that do not touch the
    // neighbouring waves (hiding the LDS read of the neighbour rows), then
    // the first and last row, and ends by publishing those two rows for
    // the next half-sweep.  Within a half-sweep every updated cell reads
    // only cells of the other colour, so the row order changes no result.
    if (w == 0) {
#pragma unroll
        for (int b = 0; b < 2; b++)
#pragma unroll
            for (int r = 0; r < 2; r++) {
                xch[b][0][r][lane] = V2{(T)0, (T)0};
                xch[b][NW + 1][r][lane] = V2{(T)0, (T)0};
            }
    }
    // Tile row rr holds the value of the full sweep after half-sweep h only
    // while rr > h and NW * NRW - 1 - rr > h: each half-sweep loses one
    // row at each edge of the tile.  A wave none of whose rows can still
    // be valid (all of them in the halo, which is H = 2M deep) skips the
    // update and the publish and only joins the barrier; the rows next to
    // it read its stale values only when they are no longer valid either.
    // That frees issue slots for the other workgroup on the CU; the
    // persistent form (one workgroup per CU, every wave waiting at the
    // barrier for the slowest SIMD) runs every wave.
    auto live = [&](int h) { return PERS || (w * NRW + NRW - 1 > h && NW * NRW - 1 - w * NRW > h); };
    auto half32 = [&]<int Q, bool UNI>(int h) {
        const int b = h & 1;
        if (live(h)) {
            const V2 sv = xch[b][w][1][lane], nv = xch[b][w + 2][0][lane];
            [&]<int... K>(std::integer_sequence<int, K...>) {
                (reg_row<T, NRW, Q, true, K + 1, UNI>(R, 0, (T)0, (T)0, (T)0, (T)0, cxpa, cxma, cxpb, cxmb, c1,
                                                      uratio),
                 ...);
            }(std::make_integer_sequence<int, NRW - 2>{});
            reg_row<T, NRW, Q, true, 0, UNI>(R, 0, sv[0], sv[1], nv[0], nv[1], cxpa, cxma, cxpb, cxmb, c1, uratio);
            reg_row<T, NRW, Q, true, NRW - 1, UNI>(R, 0, sv[0], sv[1], nv[0], nv[1], cxpa, cxma, cxpb, cxmb, c1,
                                                   uratio);
            xch[b ^ 1][w + 1][0][lane] = V2{R.pa[0], R.pb[0]};
            xch[b ^ 1][w + 1][1][lane] = V2{R.pa[NRW - 1], R.pb[NRW - 1]};
        }
        __syncthreads();
    };
    // nst is even: the half-sweeps run in pairs whose colour order is fixed
    // per wave, so each loop body is straight-line code
    auto pairs = [&]<bool UNI>(int n) {
        xch[0][w + 1][0][lane] = V2{R.pa[0], R.pb[0]};
        xch[0][w + 1][1][lane] = V2{R.pa[NRW - 1], R.pb[NRW - 1]};
        __syncthreads();
        if (Z == 0) {
            for (int h = 0; h < n; h += 2) {
                half32.template operator()<0, UNI>(h);
                half32.template operator()<1, UNI>(h + 1);
            }
        } else {
            for (int h = 0; h < n; h += 2) {
                half32.template operator()<1, UNI>(h);
                half32.template operator()<0, UNI>(h + 1);
            }
        }
    };
    // n half-sweeps (n even).  FP64: one loop body picks the colour order
    // per half-sweep (two copies of the unrolled pair would exceed the
    // register budget)
    auto sweeps = [&](int n) {
        if constexpr (sizeof(T) == 4) {
            if constexpr (WM == 2) {
                if (uni)
                    pairs.template operator()<true>(n);
                else
                    pairs.template operator()<false>(n);
            } else {
                pairs.template operator()<WM == 1>(n);
            }
        } else {
            for (int h = 0; h < n; h++) {
                const int b = h & 1;
                xch[b][w + 1][0][lane] = V2{R.pa[0], R.pb[0]};
                xch[b][w + 1][1][lane] = V2{R.pa[NRW - 1], R.pb[NRW - 1]};
                __syncthreads();
                const V2 sv = xch[b][w][1][lane], nv = xch[b][w + 2][0][lane];
#define IFF_HALF(Q_, F_)                                                                              \
    reg_half<T, NRW, Q_, F_>(R, sfl[kFlw > 1 ? w : 0], lane, smask, sv[0], sv[1], nv[0], nv[1], cxpa, cxma, \
                             cxpb, cxmb, c1)
                if (smask == 0) {
                    if (((Z ^ h) & 1) == 0) IFF_HALF(0, true); else IFF_HALF(1, true);
                } else {
                    if (((Z ^ h) & 1) == 0) IFF_HALF(0, false); else IFF_HALF(1, false);
                }
#undef IFF_HALF
            }
        }
    };
    // store the stored region of the tile (rows H .. H+TY-1, columns H ..
    // H+TW-1), or only its border strips of depth H (the halo of the
    // neighbouring tiles)
    auto store = [&](T *out, bool border) {
        if (lane >= H / 2 && lane < 64 - H / 2) {
            const bool lb = lane < H || lane >= 64 - H;
#pragma unroll
            for (int k = 0; k < NRW; k++) {
                const int rr = w * NRW + k;
                const int gj = jt0 + k;
                if (rr < H || rr >= H + TY || gj > ny - 2) continue;
                if (border && !lb && rr >= 2 * H && rr < TY) continue;
                T *row = out + (size_t)gj * P;  // wave-uniform
                if (ga + 1 <= nx - 2)
                    gstore_vec<V2>(row + ga, V2{R.pa[k], R.pb[k]});
                else if (ga <= nx - 2)
                    GW(row)[ga] = R.pa[k];
            }
        }
    };
    if constexpr (PERS) {
        using U2 = unsigned __attribute__((ext_vector_type(2)));
        using U4 = unsigned __attribute__((ext_vector_type(4)));
        const int nexec = (kexec + M - 1) / M;  // stages with sweeps for this case
        const unsigned f0 = __hip_atomic_load(&flags[t], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        // the neighbouring tiles of the case (lanes 0-7 of wave 0)
        int nb = -1;
        if (w == 0 && lane < 8) {
            const int d = lane < 4 ? lane : lane + 1;  // 3 x 3 without the centre
            const int ntys = (ny - 2 + TY - 1) / TY;
            const int qx = tx + d % 3 - 1, qy = ty + d / 3 - 1;
            if (qx >= 0 && qx < ntxs && qy >= 0 && qy < ntys) nb = t - tl.y + qy * ntxs + qx;
        }
        const int nrec = P * ny * (int)sizeof(T);  // bytes of a field
        // buffer offsets: the lane's column in the vector offset, the row
        // (wave-uniform) in the scalar offset; every access below has
        // ga >= 0 and 1 <= gj <= ny - 2, so both are non-negative
        const int voff = ga * (int)sizeof(T);
        // the chunk's two buffers: stage l writes buffer (l + 1) & 1
        void *buf0, *buf1;
        if constexpr (sizeof(T) == 8) {
            buf0 = c.p[ct->ppe_base & 1];
            buf1 = c.p[(ct->ppe_base + 1) & 1];
        } else {
            buf0 = c.e[0];
            buf1 = c.e[1];
        }
        // the descriptors are built from wave-uniform (readfirstlane)
        // pointers: held in scalar registers, with no per-lane loop around
        // the buffer accesses
        const __amdgpu_buffer_rsrc_t rs0 = __builtin_amdgcn_make_buffer_rsrc(uniform_ptr(buf0), 0, nrec, kRsrc3);
        const __amdgpu_buffer_rsrc_t rs1 = __builtin_amdgcn_make_buffer_rsrc(uniform_ptr(buf1), 0, nrec, kRsrc3);
        for (int l = 0; l < nexec; l++) {
            int me = kexec - l * M;
            if (me > M) me = M;
            sweeps(2 * me);
            const bool last = l + 1 == nexec;
            const __amdgpu_buffer_rsrc_t ro = ((l + 1) & 1) ? rs1 : rs0;
            // stage output at device scope: the border strips, or (last
            // stage of the case) the whole stored region
            if (lane >= H / 2 && lane < 64 - H / 2) {
                const bool lb = lane < H || lane >= 64 - H;
#pragma unroll
                for (int k = 0; k < NRW; k++) {
                    const int rr = w * NRW + k;
                    const int gj = jt0 + k;
                    if (rr < H || rr >= H + TY || gj > ny - 2) continue;
                    if (!last && !lb && rr >= 2 * H && rr < TY) continue;
                    const int srow = gj * P * (int)sizeof(T);  // uniform: scalar offset
                    if constexpr (sizeof(T) == 4) {
                        if (ga + 1 <= nx - 2)
                            __builtin_amdgcn_raw_buffer_store_b64(__builtin_bit_cast(U2, V2{R.pa[k], R.pb[k]}), ro,
                                                                  voff, srow, kDev);
                        else if (ga <= nx - 2)
                            __builtin_amdgcn_raw_buffer_store_b32(__float_as_uint(R.pa[k]), ro, voff, srow, kDev);
                    } else {
                        if (ga + 1 <= nx - 2)
                            __builtin_amdgcn_raw_buffer_store_b128(__builtin_bit_cast(U4, V2{R.pa[k], R.pb[k]}), ro,
                                                                   voff, srow, kDev);
                        else if (ga <= nx - 2)
                            __builtin_amdgcn_raw_buffer_store_b64(__builtin_bit_cast(U2, R.pa[k]), ro, voff, srow,
                                                                  kDev);
                    }
                }
            }
            if (last) break;
            // publish stage l once every wave's stores are complete, wait
            // for the neighbours' stage l
            __builtin_amdgcn_s_waitcnt(kWaitVmcnt0);
            __syncthreads();
            if (threadIdx.x == 0)
                __hip_atomic_store(&flags[t], f0 + (unsigned)(l + 1), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            wait_neighbours(flags, nb, f0 + (unsigned)(l + 1), bar);
            // after a timed-out wait anywhere in the launch, stop sweeping:
            // the error buffers keep finite values, the step reports the
            // failure and the solve resumes from the current pressure
            const int bad =
                threadIdx.x == 0 ? (int)__hip_atomic_load(&bar->err, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) : 0;
            if (__builtin_amdgcn_readfirstlane(__syncthreads_or(bad))) break;
            // halo ring: the cells of the tile outside its stored region
            // that lie in the domain's interior (device-scope loads)
            const bool lh = lane < H / 2 || lane >= 64 - H / 2;
#pragma unroll
            for (int k = 0; k < NRW; k++) {
                const int rr = w * NRW + k;
                const int gj = jt0 + k;
                if (gj < 1 || gj > ny - 2) continue;
                if (!lh && rr >= H && rr < H + TY) continue;
                const int srow = gj * P * (int)sizeof(T);
                if constexpr (sizeof(T) == 4) {
                    if (ina && inb) {
                        const V2 v =
                            __builtin_bit_cast(V2, __builtin_amdgcn_raw_buffer_load_b64(ro, voff, srow, kDev));
                        R.pa[k] = v[0];
                        R.pb[k] = v[1];
                    } else {
                        if (ina) R.pa[k] = __uint_as_float(__builtin_amdgcn_raw_buffer_load_b32(ro, voff, srow, kDev));
                        if (inb)
                            R.pb[k] =
                                __uint_as_float(__builtin_amdgcn_raw_buffer_load_b32(ro, voff + 4, srow, kDev));
                    }
                } else {
                    if (ina && inb) {
                        const V2 v =
                            __builtin_bit_cast(V2, __builtin_amdgcn_raw_buffer_load_b128(ro, voff, srow, kDev));
                        R.pa[k] = v[0];
                        R.pb[k] = v[1];
                    } else {
                        if (ina)
                            R.pa[k] = __builtin_bit_cast(T, __builtin_amdgcn_raw_buffer_load_b64(ro, voff, srow, kDev));
                        if (inb)
                            R.pb[k] =
                                __builtin_bit_cast(T, __builtin_amdgcn_raw_buffer_load_b64(ro, voff + 8, srow, kDev));
                    }
                }
            }
        }
        __syncthreads();
        if (threadIdx.x == 0)
            __hip_atomic_store(&flags[t], f0 + (unsigned)nstage, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        return;
    }
    sweeps(nst);
    store(dst, false);
}

template <typename T, int M, int NW, int NRW>
static void reg_one(hipStream_t st, const CaseDev *cs, const int2 *tiles, int nt, int s, int K, int wt)
{
    const int pad = ((nt + 7) / 8) * 8;
#define IFF_RL(WM_)                                                                                             \
    hipLaunchKernelGGL((k_rbsor_reg<T, M, NW, NRW, false, WM_>), dim3(pad), dim3(64 * NW), 0, st, cs, tiles, nt, \
                       pad, s, K, (GridBar *)nullptr, (unsigned *)nullptr)
    if (wt == 1)
        IFF_RL(1);
    else if (sizeof(T) == 4 && wt == 2)
        IFF_RL(sizeof(T) == 4 ? 2 : 0);
    else
        IFF_RL(0);
#undef IFF_RL
}

// Blocks of the persistent form that can be resident at once: resident
// blocks per CU (occupancy calculator; the smaller of the two weight forms)
// times the CUs, per shape.
template <typename T, int M, int NW, int NRW>
static int pers_capacity()
{
    static int capacity = -1;  // one device type per process
    if (capacity < 0) {
        int per_cu = 0, per_cu_t = 0, dev = 0, cus = 0;
        if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k_rbsor_reg<T, M, NW, NRW, true, 0>,
                                                         64 * NW, 0) != hipSuccess ||
            hipGetDevice(&dev) != hipSuccess ||
            hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess)
            return 0;
        if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu_t, k_rbsor_reg<T, M, NW, NRW, true, 1>, 64 * NW,
                                                         0) != hipSuccess)
            return 0;
        per_cu = std::min(per_cu, per_cu_t);
        if constexpr (sizeof(T) == 4) {
            if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu_t, k_rbsor_reg<T, M, NW, NRW, true, 2>,
                                                             64 * NW, 0) != hipSuccess)
                return 0;
            per_cu = std::min(per_cu, per_cu_t);
        }
        capacity = per_cu * cus;
    }
    return capacity;
}

template <typename T, int M, int NW, int NRW>
static void pers_one(hipStream_t st, const CaseDev *cs, const int2 *tiles, int nt, int K, GridBar *bar,
                     unsigned *flags, int wt)
{
    const int pad = ((nt + 7) / 8) * 8;
#define IFF_PLW(WM_)                                                                                           \
    hipLaunchKernelGGL((k_rbsor_reg<T, M, NW, NRW, true, WM_>), dim3(pad), dim3(64 * NW), 0, st, cs, tiles, nt, \
                       pad, 0, K, bar, flags)
    if (wt == 1)
        IFF_PLW(1);
    else if (sizeof(T) == 4 && wt == 2)
        IFF_PLW(sizeof(T) == 4 ? 2 : 0);
    else
        IFF_PLW(0);
#undef IFF_PLW
}

// the shapes with a persistent form: (waves, sweeps per stage, tile_y,
// rows per wave), FP32 and FP64
#define IFF_PERS_SHAPES32(X) \
    X(float, 16, 8, 96, 8) X(float, 16, 8, 64, 6) X(float, 16, 4, 48, 4) X(float, 8, 8, 32, 8) X(float, 8, 4, 32, 6)
#define IFF_PERS_SHAPES64(X) X(double, 16, 4, 80, 6) X(double, 16, 4, 64, 5) X(double, 16, 4, 48, 4)

int pers_blocks(int prec, int waves, int ty, int m)
{
#define IFF_PB(T_, W_, M_, TY_, NRW_) \
    if (waves == W_ && m == M_ && ty == TY_) return pers_capacity<T_, M_, W_, NRW_>();
    if (prec == IF_FP64) {
        IFF_PERS_SHAPES64(IFF_PB)
    } else {
        IFF_PERS_SHAPES32(IFF_PB)
    }
#undef IFF_PB
    return 0;
}

bool pers_fits(int prec, int waves, int ty, int m, int nt)
{
    const int pad = ((nt + 7) / 8) * 8;
#define IFF_PF(T_, W_, M_, TY_, NRW_) \
    if (waves == W_ && m == M_ && ty == TY_) return pad <= pers_capacity<T_, M_, W_, NRW_>();
    if (prec == IF_FP64) {
        IFF_PERS_SHAPES64(IFF_PF)
    } else {
        IFF_PERS_SHAPES32(IFF_PF)
    }
#undef IFF_PF
    return false;
}

void launch_ppe_sweep_pers(hipStream_t st, int prec, int waves, int ty, int m, const CaseDev *cs,
                           const int2 *tiles, int nt, int K, GridBar *bar, unsigned *flags, int wt)
{
#define IFF_PL(T_, W_, M_, TY_, NRW_)                                      \
    if (waves == W_ && m == M_ && ty == TY_) {                            \
        pers_one<T_, M_, W_, NRW_>(st, cs, tiles, nt, K, bar, flags, wt); \
        return;                                                           \
    }
    if (prec == IF_FP64) {
        IFF_PERS_SHAPES64(IFF_PL)
    } else {
        IFF_PERS_SHAPES32(IFF_PL)
    }
#undef IFF_PL
}

// Stored tile: 128 - 4M columns by tile_y rows; tile_y + 4M rows per
// workgroup, split over its waves (4, 8 or 16).  M = 8 (8- and 16-wave
// tiles) trades a wider halo for half the loads per sweep.  The FP64 tile
// of 16 waves x 8 rows does not fit the 128 registers a 1024-lane
// workgroup allows, so that shape is FP32 (mixed precision) only, as are
// its deeper forms: 9 to 12 sweeps on the same 128 x 128 cells (92, 88, 84
// and 80 stored rows and columns), fewer loads per sweep for more
// recomputed halo.
bool reg_variant_ok(int prec, int waves, int ty, int m)
{
    if (waves == 4) return (m == 1 || m == 2 || m == 4) && (ty == 16 || ty == 32 || ty == 48);
    if (waves == 8)
        return (m == 2 && (ty == 24 || ty == 40 || ty == 56)) || (m == 4 && (ty == 32 || ty == 48 || ty == 64)) ||
               (m == 8 && (ty == 32 || ty == 48 || ty == 64));
    if (waves == 16)
        return (m == 4 && (ty == 48 || ty == 64 || ty == 80)) ||
               (m == 8 && (ty == 32 || ty == 64 || (ty == 96 && prec != IF_FP64))) ||
               (prec != IF_FP64 && ((m == 9 && ty == 92) || (m == 10 && ty == 88) || (m == 11 && ty == 84) ||
                                    (m == 12 && ty == 80)));
    return false;
}

int reg_out_cols(int m) { return 128 - 4 * m; }

void launch_ppe_sweep_reg(hipStream_t st, int prec, int waves, int ty, int m, const CaseDev *cs,
                          const int2 *tiles, int nt, int s, int K, int wt)
{
#define IFF_RG(T, W_, M_, TY_)                                                            \
    if (waves == W_ && m == M_ && ty == TY_) {                                            \
        reg_one<T, M_, W_, (TY_ + 4 * M_) / W_>(st, cs, tiles, nt, s, K, wt);            \
        return;                                                                           \
    }
#define IFF_RG_ALL(T)                                                                     \
    IFF_RG(T, 4, 1, 16) IFF_RG(T, 4, 1, 32) IFF_RG(T, 4, 1, 48)                           \
    IFF_RG(T, 4, 2, 16) IFF_RG(T, 4, 2, 32) IFF_RG(T, 4, 2, 48)                           \
    IFF_RG(T, 4, 4, 16) IFF_RG(T, 4, 4, 32) IFF_RG(T, 4, 4, 48)                           \
    IFF_RG(T, 8, 2, 24) IFF_RG(T, 8, 2, 40) IFF_RG(T, 8, 2, 56)                           \
    IFF_RG(T, 8, 4, 32) IFF_RG(T, 8, 4, 48) IFF_RG(T, 8, 4, 64)                           \
    IFF_RG(T, 8, 8, 32) IFF_RG(T, 8, 8, 48) IFF_RG(T, 8, 8, 64)                           \
    IFF_RG(T, 16, 4, 48) IFF_RG(T, 16, 4, 64) IFF_RG(T, 16, 4, 80) IFF_RG(T, 16, 8, 32) IFF_RG(T, 16, 8, 64)
    if (prec == IF_FP64) {
        IFF_RG_ALL(double)
    } else {
        IFF_RG_ALL(float)
        IFF_RG(float, 16, 8, 96)
        IFF_RG(float, 16, 9, 92)
        IFF_RG(float, 16, 10, 88)
        IFF_RG(float, 16, 11, 84)
        IFF_RG(float, 16, 12, 80)
    }
#undef IFF_RG_ALL
#undef IFF_RG
}

}  // namespace iff
