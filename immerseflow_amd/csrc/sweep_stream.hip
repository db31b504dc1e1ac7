// Register-streaming red-black SOR for gfx950.
//
// One wave64 owns a column strip: lane l holds column i0 - 2M + l, and the
// wave walks down the strip one row at a time, keeping a ring of 2M+2 rows
// of the unknown, the right side and the flags in registers.  East/west
// neighbours come from the adjacent lanes through DPP wave shifts,
// north/south neighbours from the ring.  At each new row t the wave applies
// the 2M half-sweeps of M red-black SOR sweeps as a wavefront: half-sweep k
// (colour k&1) updates row t-1-k.  Every cell then sees exactly the values
// it would see in M global red-black sweeps (red then black), so results are
// bit-identical to the oracle's sweep order.  The first and last 2M lanes
// and rows are halo: their values are computed but never stored.
//
// No LDS and no barriers: each wave is independent; memory traffic is one
// read of p, rhs and flags (plus the halo) and one write of p per M sweeps.
#include "common.h"
#include "kernels.h"
#include "sor.h"

namespace iff {

template <typename T, int M, int TY>
__global__ __launch_bounds__(256) void k_rbsor_stream(const CaseDev *cs, const int2 *tiles,
                                                      int ntiles, int nblk_pad, int s, int K)
{
    constexpr int H = 2 * M;        // halo (rows above/below, lanes each side)
    constexpr int WO = 64 - 2 * H;  // output columns per wave
    constexpr int L = 2 * M + 2;    // rows alive in the ring
    const int lane = threadIdx.x & 63;
    const int t = tile_of_block(nblk_pad) * 4 + (threadIdx.x >> 6);
    if (t >= ntiles) return;  // whole wave: no barriers in this kernel
    const int2 tl = tiles[t];
    const CaseDev &c = cs[tl.x];
    const Ctrl *ct = c.ctrl;
    if (c.ppe_solver != IF_PPE_RBSOR || !ct->active || ct->ppe_done) return;
    int kexec = c.ppe_itermax - ct->ppe_sweeps;
    if (kexec > K) kexec = K;
    if (s >= kexec) return;
    int meff = kexec - s;
    if (meff > M) meff = M;
    const int nst = 2 * meff;  // half-sweeps applied by this launch
    const int Lx = s / M;

    const int nx = c.nx, ny = c.ny, P = c.P;
    const T *src;
    T *dst;
    const T *rhs;
    const T *cpx_, *cmx_, *cpy_, *cmy_;
    T om;
    bool zero_src = false;
    if constexpr (sizeof(T) == 8) {
        src = c.p[(ct->ppe_base + Lx) & 1];
        dst = c.p[(ct->ppe_base + Lx + 1) & 1];
        rhs = c.rhs;
        cpx_ = c.x.cp; cmx_ = c.x.cm; cpy_ = c.y.cp; cmy_ = c.y.cm;
        om = c.omega;
    } else {
        src = c.e[Lx & 1];
        dst = c.e[(Lx + 1) & 1];
        zero_src = (Lx == 0);
        rhs = c.r32;
        cpx_ = c.xf.cp; cmx_ = c.xf.cm; cpy_ = c.yf.cp; cmy_ = c.yf.cm;
        om = c.omega_f;
    }
    // FP32, one coefficient value per axis: the five-operation update
    const bool uni = sizeof(T) == 4 && c.uni;
    const float uax = c.uax, uratio = c.uratio;
    const int ntxs = (nx - 2 + WO - 1) / WO;
    const int tx = tl.y % ntxs, ty = tl.y / ntxs;
    const int i0 = 1 + tx * WO, j0 = 1 + ty * TY;
    const int gi = i0 - H + lane;
    const bool colin = gi >= 1 && gi <= nx - 2;
    const bool colany = gi >= 0 && gi < nx;
    const T cpx = colin ? cpx_[gi] : (T)0, cmx = colin ? cmx_[gi] : (T)0;
    const bool out_lane = lane >= H && lane < 64 - H && colin;

    T Pr[L], Rr[L], CY[L], MY[L];
    int Fr[L];
#pragma unroll
    for (int q = 0; q < L; q++) { Pr[q] = (T)0; Rr[q] = (T)0; CY[q] = (T)0; MY[q] = (T)0; Fr[q] = 0; }

    const int tstart = j0 - H, tend = j0 + TY - 1 + H;
    for (int tb = tstart; tb <= tend; tb += L) {
#pragma unroll
        for (int u = 0; u < L; u++) {
            const int tt = tb + u;
            // load row tt into slot u
            {
                T pv = (T)0, rv = (T)0, cy = (T)0, my = (T)0;
                int fv = 0;
                if (tt >= 0 && tt < ny && tt <= tend) {
                    if (colany && !zero_src) pv = src[gi + tt * P];
                    if (tt >= 1 && tt <= ny - 2) {
                        cy = cpy_[tt];
                        my = cmy_[tt];
                        if (colin) {
                            rv = rhs[gi + tt * P];
                            fv = c.flags[gi + tt * P];
                        }
                    }
                }
                Pr[u] = pv; Rr[u] = rv; Fr[u] = fv; CY[u] = cy; MY[u] = my;
            }
            // wavefront: half-sweep k updates row tt-1-k
#pragma unroll
            for (int k = 0; k < 2 * M; k++) {
                if (k < nst) {
                    const int sj = ((u - 1 - k) % L + L) % L;       // row tt-1-k
                    const int sn = ((u - k) % L + L) % L;           // row tt-k   (north)
                    const int ss = ((u - 2 - k) % L + L) % L;       // row tt-2-k (south)
                    const int j = tt - 1 - k;
                    const T pc = Pr[sj];
                    const T pE = from_east(pc), pW = from_west(pc);
                    const int f = Fr[sj];
                    if ((f & F_FLUID) && (((gi + j) & 1) == (k & 1))) {
                        const T aE = (f & F_OPEN_E) ? cpx : (T)0;
                        const T aW = (f & F_OPEN_W) ? cmx : (T)0;
                        const T aN = (f & F_OPEN_N) ? CY[sj] : (T)0;
                        const T aS = (f & F_OPEN_S) ? MY[sj] : (T)0;
                        const T sa = (aE + aW) + (aN + aS);
                        if (sa != (T)0) {
                            if (uni)
                                Pr[sj] = sor_cell_u_sa(pc, Rr[sj], pE, pW, Pr[sn], Pr[ss], sa, om, uax, uratio);
                            else
                                Pr[sj] = sor_cell<T>(pc, Rr[sj], aE, pE, aW, pW, aN, Pr[sn], aS, Pr[ss], sa, om, gi);
                        }
                    }
                }
            }
            // row tt-2M is final
            {
                const int so = ((u - 2 * M) % L + L) % L;
                const int jo = tt - 2 * M;
                if (out_lane && jo >= j0 && jo < j0 + TY && jo <= ny - 2) dst[gi + jo * P] = Pr[so];
            }
        }
    }
}

template <typename T, int M, int TY>
static void stream_one(hipStream_t st, const CaseDev *cs, const int2 *tiles, int nt, int s, int K)
{
    const int nblk = (nt + 3) / 4;
    const int pad = ((nblk + 7) / 8) * 8;
    hipLaunchKernelGGL((k_rbsor_stream<T, M, TY>), dim3(pad), dim3(256), 0, st, cs, tiles, nt, pad, s, K);
}

bool stream_variant_ok(int ty, int m)
{
    return (ty == 16 || ty == 32 || ty == 64 || ty == 128) && (m == 1 || m == 2 || m == 4);
}

int stream_out_cols(int m) { return 64 - 4 * m; }

void launch_ppe_sweep_stream(hipStream_t st, int prec, int ty, int m, const CaseDev *cs,
                             const int2 *tiles, int nt, int s, int K)
{
#define IFF_ST(T, M_, TY_) \
    if (m == M_ && ty == TY_) { stream_one<T, M_, TY_>(st, cs, tiles, nt, s, K); return; }
#define IFF_ST_ALL(T)                                                            \
    IFF_ST(T, 1, 16) IFF_ST(T, 1, 32) IFF_ST(T, 1, 64) IFF_ST(T, 1, 128)        \
    IFF_ST(T, 2, 16) IFF_ST(T, 2, 32) IFF_ST(T, 2, 64) IFF_ST(T, 2, 128)        \
    IFF_ST(T, 4, 16) IFF_ST(T, 4, 32) IFF_ST(T, 4, 64) IFF_ST(T, 4, 128)
    if (prec == IF_FP64) { IFF_ST_ALL(double) } else { IFF_ST_ALL(float) }
#undef IFF_ST_ALL
#undef IFF_ST
}

}  // namespace iff
