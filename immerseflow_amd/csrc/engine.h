// Host runtime of the physics-mode fractional step: resident cases, batched
// launches, device-side loop control.
#pragma once

#include <memory>
#include <vector>

#include "common.h"

namespace iff {

// A case resident in HBM (one device arena holds all its arrays).
struct Slot {
    int id = -1;
    int state = -2;          // 0 running, 1 finished, -1 failed, -2 empty
    int prec = IF_FP64;
    int nx = 0, ny = 0, P = 0, NX = 0, NY = 0;
    int ntx = 0, nty = 0;    // base tiling (kTileX x kTileY)
    int gcap = 0, ng = 0;
    bool moving = false;
    int steps_host = 0;      // steps enqueued while running
    int steps_target = 0;
    double t0 = 0.0;
    if_case_desc desc{};
    std::vector<double> xf, yf, xc, yc, dx, dy;
    std::vector<if_body> bodies;
    std::vector<std::vector<double>> polyx, polyy;
    std::vector<BodyDev> bodies_dev;   // host image (device pointers inside)
    void *arena = nullptr;
    void *mon_mem = nullptr;           // monitors (Monitor + partials + ring + count)
    size_t mon_bytes = 0;
    size_t bytes = 0;
    int *ghost_counts = nullptr;       // scratch for the ghost compaction
    int ghost_blocks = 0;
    CaseDev dev{};                     // host image of the device descriptor
    if_step_stats last{};
};

class Engine {
public:
    Engine(int prec, size_t budget);
    ~Engine();
    if_status init();

    if_status add(const if_case_desc &d, int nsteps, int *slot);
    if_status remove(int slot);
    if_status step(int nsteps, int *finished, if_step_stats *agg);
    if_status slot_state(int slot, int *state, int *steps_done, if_step_stats *st);
    if_status get_field(int slot, int which, double *out, size_t n);
    if_status set_field(int slot, int which, const double *in, size_t n);
    if_status restart(int slot, const double *u, const double *v, const double *p, const double *uf,
                      const double *vf);
    if_status ghosts(int slot, int32_t *out, if_ghost *recs, size_t cap, size_t *n);
    if_status set_target(int slot, int nsteps);
    if_status autotune(if_tune_result *out);
    if_status monitor(int slot, int nprobe, const double *xy, int cap);
    if_status monitor_read(int slot, double *out, size_t cap, size_t *nrec, size_t *dropped);
    if_status monitor_sample(int slot);
    if_status monitor_sample_noalloc(int slot);
    if_status sync();

    int resident() const;
    size_t in_use() const { return used_; }
    size_t budget() const { return budget_; }
    int precision() const { return prec_; }
    int poisson_precision() const { return pprec_; }
    Slot *slot(int k) { return (k >= 0 && k < (int)slots_.size()) ? slots_[k].get() : nullptr; }

    // Poisson sweep variant and check cadence (autotuned or user-set)
    // (FP64 defaults; the constructor sets 4, 64, 8 for mixed precision)
    int sweep_kind = 3;  // 0: LDS-tiled, 1: register-streaming, 2/3/4: register-tile (4/8/16 waves)
    int sweep_ty = 32, sweep_m = 4, check_k = 10;
    bool shape_auto = true;  // FP64: the shape follows the resident set (refresh) until tuned
    bool timing = false;
    // replay each chunk of Poisson sweep launches from a captured hipGraph
    bool use_graphs = true;
    // mixed precision, register-tile shapes: run each chunk as one
    // persistent launch when all its blocks fit on the device at once
    bool use_persistent = true;
    double sweep_ms = 0.0;
    long long sweep_launches = 0;
    double sweep_cells = 0.0;  // interior cell-sweeps inside the timed chunks
    int launch_sweeps = 0;     // red-black sweeps per launch in the last step (M; K when persistent)
    int spec_depth = 3;

    static size_t footprint(const if_case_desc &d, int prec);
    void mark_dirty() { dirty_ = true; }

private:
    if_status build_slot(Slot &s, const if_case_desc &d);
    if_status refresh();      // re-upload the resident case list / tile tables
    if_status snapshot(int ring_slot);
    bool all_done(const Ctrl *snap, int which) const;
    if_status read_ctrl_sync(std::vector<Ctrl> &out);
    if_status update_moving();
    // the K sweeps of one Poisson chunk, as kernel launches on st_
    void enqueue_sweeps(int K, int M, bool any_rb, bool any_jac, bool any_line, bool pers, bool stamp);
    if_status capture_sweeps(int K, int M, bool any_rb, bool any_jac, bool any_line, bool pers, bool stamp);
    if_status check_barrier();
    void drop_graph();

    int prec_;   // IF_FP64 / IF_MIXED / IF_MIXED_V32
    int vprec_;  // velocity passes: IF_FP64, or IF_MIXED (FP32) for IF_MIXED_V32
    int pprec_;  // Poisson: IF_FP64, or IF_MIXED (FP32 error sweeps) for both mixed modes
    size_t budget_, used_ = 0;
    hipStream_t st_ = nullptr;
    std::vector<std::unique_ptr<Slot>> slots_;
    std::vector<int> order_;           // resident slot ids, launch order
    CaseDev *d_cases_ = nullptr;
    int cap_cases_ = 0;
    Ctrl *d_ctrl_ = nullptr;           // one control block per slot id
    int cap_ctrl_ = 4096;
    int2 *d_tiles_base_ = nullptr, *d_tiles_sweep_ = nullptr;
    int nt_base_ = 0, nt_sweep_ = 0, cap_tiles_base_ = 0, cap_tiles_sweep_ = 0;
    int tiles_sweep_ty_ = 0;
    int wmode_ = 0;                   // sweep weight mode: 1 every red-black case resident has a weight table,
                                       // 2 (FP32) some take the five-operation form, 0 none
    int max_ng_ = 0, max_ring_ = 0, max_cells_ = 0, max_ad_ = 0, max_ppe_ = 0;
    bool any_moving_ = false;
    bool any_monitor_ = false;
    bool dirty_ = true;
    // speculation ring
    static constexpr int kRing = 8;
    Ctrl *h_snap_ = nullptr;           // pinned, kRing * cap_ctrl_
    hipEvent_t ev_[kRing] = {};
    // device time stamps of the timed Poisson chunks: [0] count, then a
    // start and an end stamp per chunk (at most kStamps)
    static constexpr int kStamps = 2 * 64;
    unsigned long long *d_stamps_ = nullptr, *h_stamps_ = nullptr;  // 1 + kStamps each (host: pinned)
    double stamp_hz_ = 100.0e6;        // rate of the real-time counter
    // captured sweep chunk, valid for the resident set of the last
    // refresh() and the launch shape in g_key_
    hipGraphExec_t g_sweeps_ = nullptr;
    int g_key_[9] = {};
    bool g_valid_ = false;     // g_sweeps_ matches the current tables

public:
    long long graph_instantiations = 0, graph_updates = 0;
    long long persistent_timeouts = 0;  // persistent chunks re-solved per stage

private:
    GridBar *d_bar_ = nullptr;     // error flag of the persistent sweep
    unsigned *d_flags_ = nullptr;  // its per-tile stage counters
    int cap_flags_ = 0;
};

}  // namespace iff
