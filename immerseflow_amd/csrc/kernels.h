// Launch wrappers of the fractional-step kernels (kernels.hip).
#pragma once
#include "common.h"

namespace iff {

void launch_step_begin(hipStream_t st, const CaseDev *cs, int ncase);
void launch_source(hipStream_t st, int prec, const CaseDev *cs, const int2 *tiles, int nt);
void launch_ad_sweep(hipStream_t st, int prec, const CaseDev *cs, const int2 *tiles, int nt);
void launch_ad_finalize(hipStream_t st, const CaseDev *cs, int ncase);
void launch_ghost_vel(hipStream_t st, int prec, const CaseDev *cs, int ncase, int max_ng, int mode);
void launch_pressure_bc(hipStream_t st, const CaseDev *cs, int ncase, int max_ring, int max_ng);
bool sweep_variant_ok(int ty, int m);
void launch_ppe_sweep(hipStream_t st, int prec, int solver, int ty, int m, const CaseDev *cs,
                      const int2 *tiles, int nt, int s, int K);
// register-streaming red-black SOR (sweep_stream.hip)
bool stream_variant_ok(int ty, int m);
int stream_out_cols(int m);
void launch_ppe_sweep_stream(hipStream_t st, int prec, int ty, int m, const CaseDev *cs,
                             const int2 *tiles, int nt, int s, int K);
// register-tile red-black SOR (sweep_reg.hip)
bool reg_variant_ok(int prec, int waves, int ty, int m);
int reg_out_cols(int m);
// wt, the weight mode: 1 every red-black case of the tile table has a weight
// table (CaseDev::wtab, FP64: wtab64); 0 none uses the five-operation FP32
// form (CaseDev::uni); 2 (FP32) some do, some do not
void launch_ppe_sweep_reg(hipStream_t st, int prec, int waves, int ty, int m, const CaseDev *cs,
                          const int2 *tiles, int nt, int s, int K, int wt);
// The K sweeps of a chunk in one launch of the register-tile kernel, in
// stages of m sweeps with a halo exchange between neighbouring tiles after
// each (prec: the Poisson sweep precision).  pers_fits: the shape has a
// persistent form in that precision and all nt tiles' blocks can be
// resident on the device at once (a host query; call it outside stream
// capture).  flags: one stage counter per tile, zeroed whenever the tile
// table changes.
bool pers_fits(int prec, int waves, int ty, int m, int nt);
int pers_blocks(int prec, int waves, int ty, int m);  // 0: no persistent form
void launch_ppe_sweep_pers(hipStream_t st, int prec, int waves, int ty, int m, const CaseDev *cs,
                           const int2 *tiles, int nt, int K, GridBar *bar, unsigned *flags, int wt);
void launch_ppe_check(hipStream_t st, int prec, const CaseDev *cs, int ncase, const int2 *tiles,
                      int nt, int K, const GridBar *bar = nullptr);
void launch_ppe_rearm(hipStream_t st, const CaseDev *cs, int ncase);
// zebra line SOR: the per-step factors (pivots, c', 1/m) of every line
// case, then one sweep (both colours) at chunk position s
void launch_line_factor(hipStream_t st, int prec, const CaseDev *cs, int ncase, int max_ny);
void launch_weights32(hipStream_t st, const CaseDev *cs, int ncase, int max_n);
void launch_weight_table64(hipStream_t st, const CaseDev *cs, int ncase);
void launch_line_sor(hipStream_t st, int prec, const CaseDev *cs, int ncase, int max_n, int s, int K);
void launch_project(hipStream_t st, int prec, const CaseDev *cs, const int2 *tiles, int nt);
void launch_step_end(hipStream_t st, const CaseDev *cs, int ncase);
void launch_monitor(hipStream_t st, int prec, const CaseDev *cs, int ncase, const int2 *tiles, int nt,
                    int sample = 0, double *part = nullptr);
void launch_fill_tiles(hipStream_t st, int2 *tiles, int nt);
// device time stamps of the Poisson chunks (buf[0]: count, then the stamps)
void launch_stamp(hipStream_t st, unsigned long long *buf, int cap);
void launch_stamp_reset(hipStream_t st, unsigned long long *buf);
void launch_tune_arm(hipStream_t st, const CaseDev *cs, int ncase, const Ctrl *ctrl0, const Ctrl *saved);

// immersed-boundary setup (geometry.hip); fresh != 0: only moving cases,
// with fresh-cell treatment
void launch_classify(hipStream_t st, const CaseDev *cs, int ncase, int max_cells, int fresh);
void launch_ghost_build(hipStream_t st, const CaseDev *cs, int ncase, int max_cells, int fresh);
void launch_ghost_weights(hipStream_t st, const CaseDev *cs, int ncase, int max_ng, int fresh);
int ghost_span();

// initial state (kernels.hip)
void launch_init_fields(hipStream_t st, int prec, const CaseDev *cs, int ncase, int max_cells);
void launch_faces(hipStream_t st, int prec, const CaseDev *cs, const int2 *tiles, int nt,
                  int out_buf, int with_rhs);

}  // namespace iff
