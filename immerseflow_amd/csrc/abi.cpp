// extern "C" entry points (include/immerseflow/if_abi.h).
#include <cstring>
#include <memory>
#include <new>
#include <string>

#include "compat.h"
#include "engine.h"
#include "io.h"
#include "kernels.h"

using namespace iff;

struct if_case {
    int mode = IF_MODE_PHYSICS;
    std::unique_ptr<Engine> eng;
    int slot = -1;
    std::unique_ptr<CompatCase> cc;
};

struct if_batch {
    std::unique_ptr<Engine> eng;
};

#define IFF_GUARD(expr)                                                   \
    try {                                                                 \
        return (expr);                                                    \
    } catch (const std::bad_alloc &) {                                    \
        return fail(IF_ERR_NOMEM, "host allocation failed");              \
    } catch (const std::exception &e) {                                   \
        return fail(IF_ERR_STATE, std::string("internal error: ") + e.what()); \
    }

extern "C" {

const char *if_last_error(void) { return last_error(); }

if_status if_device_query(int *ndev, char *name, size_t name_len, int *cus, size_t *hbm)
{
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess) n = 0;
    if (ndev) *ndev = n;
    if (n < 1) return fail(IF_ERR_NODEV, "no HIP device available");
    hipDeviceProp_t prop;
    IFF_HIP(hipGetDeviceProperties(&prop, 0));
    if (name && name_len) {
        std::strncpy(name, prop.gcnArchName, name_len - 1);
        name[name_len - 1] = 0;
    }
    if (cus) *cus = prop.multiProcessorCount;
    if (hbm) *hbm = prop.totalGlobalMem;
    return IF_OK;
}

if_status if_parse_inputs(const char *path, if_inputs *out) { IFF_GUARD(parse_inputs(path, out)); }

if_status if_read_grid(const char *path, int nf, double *xf) { IFF_GUARD(read_grid(path, nf, xf)); }

if_status if_grid_metric(int nf, const double *xf, double *xc, double *dx)
{
    IFF_GUARD(grid_metric(nf, xf, xc, dx));
}

if_status if_write_tecplot_digits(const char *path, int nx, int ny, const double *x, const double *y,
                                  const double *f, int digits)
{
    IFF_GUARD(write_tecplot_digits(path, nx, ny, x, y, f, digits));
}

if_status if_read_tecplot(const char *path, int nx, int ny, double *x, double *y, double *f)
{
    IFF_GUARD(read_tecplot(path, nx, ny, x, y, f));
}

if_status if_write_tecplot(const char *path, int nx, int ny, const double *x, const double *y,
                           const double *f)
{
    IFF_GUARD(write_tecplot(path, nx, ny, x, y, f));
}

static if_status case_create(const if_case_desc *d, if_case **out)
{
    if (!d || !out) return fail(IF_ERR_ARG, "if_case_create: null argument");
    *out = nullptr;
    auto c = std::make_unique<if_case>();
    c->mode = d->mode;
    if (d->mode == IF_MODE_COMPAT) {
        c->cc = std::make_unique<CompatCase>();
        if_status rc = c->cc->create(*d);
        if (rc != IF_OK) return rc;
    } else if (d->mode == IF_MODE_PHYSICS) {
        if (d->precision != IF_FP64 && d->precision != IF_MIXED && d->precision != IF_MIXED_V32)
            return fail(IF_ERR_ARG, "if_case_create: unknown precision");
        c->eng = std::make_unique<Engine>(d->precision, 0);
        if_status rc = c->eng->init();
        if (rc != IF_OK) return rc;
        c->eng->check_k = d->check_every > 0 ? d->check_every : 10;
        rc = c->eng->add(*d, 0, &c->slot);
        if (rc != IF_OK) return rc;
    } else {
        return fail(IF_ERR_ARG, "if_case_create: unknown mode");
    }
    *out = c.release();
    return IF_OK;
}

if_status if_case_create(const if_case_desc *d, if_case **out) { IFF_GUARD(case_create(d, out)); }

void if_case_destroy(if_case *c) { delete c; }

if_status if_case_ad_step(if_case *c, if_step_stats *st)
{
    if (!c) return fail(IF_ERR_ARG, "null case");
    if (c->mode != IF_MODE_COMPAT) return fail(IF_ERR_STATE, "if_case_ad_step: compat mode only");
    IFF_GUARD(c->cc->ad_step(st));
}

if_status if_case_ppe_solve(if_case *c, if_step_stats *st)
{
    if (!c) return fail(IF_ERR_ARG, "null case");
    if (c->mode != IF_MODE_COMPAT) return fail(IF_ERR_STATE, "if_case_ppe_solve: compat mode only");
    IFF_GUARD(c->cc->ppe_solve(st));
}

static if_status case_step(if_case *c, int nsteps, if_step_stats *st)
{
    if (!c) return fail(IF_ERR_ARG, "null case");
    if (c->mode != IF_MODE_PHYSICS) return fail(IF_ERR_STATE, "if_case_step: physics mode only");
    if (nsteps < 0) return fail(IF_ERR_ARG, "if_case_step: negative step count");
    if (st) std::memset(st, 0, sizeof(*st));
    if (nsteps == 0) return IF_OK;
    if_status rc = c->eng->set_target(c->slot, nsteps);
    if (rc != IF_OK) return rc;
    int fin = 0;
    rc = c->eng->step(nsteps, &fin, st);
    if (rc != IF_OK) return rc;
    int state = 0;
    c->eng->slot_state(c->slot, &state, nullptr, nullptr);
    if (state == -1) return fail(IF_ERR_DIVERGED, "non-finite residual: case failed");
    return IF_OK;
}

if_status if_case_step(if_case *c, int nsteps, if_step_stats *st) { IFF_GUARD(case_step(c, nsteps, st)); }

if_status if_case_get_field(if_case *c, int which, double *out, size_t n)
{
    if (!c) return fail(IF_ERR_ARG, "null case");
    if (c->mode == IF_MODE_COMPAT) IFF_GUARD(c->cc->get_field(which, out, n));
    IFF_GUARD(c->eng->get_field(c->slot, which, out, n));
}

if_status if_case_set_field(if_case *c, int which, const double *in, size_t n)
{
    if (!c) return fail(IF_ERR_ARG, "null case");
    if (c->mode == IF_MODE_COMPAT) IFF_GUARD(c->cc->set_field(which, in, n));
    IFF_GUARD(c->eng->set_field(c->slot, which, in, n));
}

if_status if_case_dims(if_case *c, int *nx, int *ny)
{
    if (!c) return fail(IF_ERR_ARG, "null case");
    if (c->mode == IF_MODE_COMPAT) {
        if (nx) *nx = c->cc->nx;
        if (ny) *ny = c->cc->ny;
    } else {
        Slot *s = c->eng->slot(c->slot);
        if (nx) *nx = s->nx;
        if (ny) *ny = s->ny;
    }
    return IF_OK;
}

if_status if_case_grid(if_case *c, double *xc, double *yc, double *dx, double *dy)
{
    if (!c) return fail(IF_ERR_ARG, "null case");
    const std::vector<double> *vxc, *vyc, *vdx, *vdy;
    if (c->mode == IF_MODE_COMPAT) {
        vxc = &c->cc->xc; vyc = &c->cc->yc; vdx = &c->cc->dx; vdy = &c->cc->dy;
    } else {
        Slot *s = c->eng->slot(c->slot);
        vxc = &s->xc; vyc = &s->yc; vdx = &s->dx; vdy = &s->dy;
    }
    if (xc) std::memcpy(xc, vxc->data(), vxc->size() * 8);
    if (yc) std::memcpy(yc, vyc->data(), vyc->size() * 8);
    if (dx) std::memcpy(dx, vdx->data(), vdx->size() * 8);
    if (dy) std::memcpy(dy, vdy->data(), vdy->size() * 8);
    return IF_OK;
}

if_status if_case_ghosts(if_case *c, int32_t *out, if_ghost *recs, size_t cap, size_t *n)
{
    if (!c) return fail(IF_ERR_ARG, "null case");
    if (c->mode != IF_MODE_PHYSICS) {
        if (n) *n = 0;
        return IF_OK;
    }
    IFF_GUARD(c->eng->ghosts(c->slot, out, recs, cap, n));
}

static if_status tune(Engine &e, int kind, int ty, int m, int k)
{
    if (kind == 0 && !sweep_variant_ok(ty, m))
        return fail(IF_ERR_ARG, "tune: LDS sweep needs tile_y 8/16/32 and 1/2/4 sweeps per launch");
    if (kind == 1 && !stream_variant_ok(ty, m))
        return fail(IF_ERR_ARG, "tune: streaming sweep needs strip height 16/32/64/128 and 1/2/4 sweeps");
    const int pr = e.poisson_precision();
    if (kind == 2 && !reg_variant_ok(pr, 4, ty, m))
        return fail(IF_ERR_ARG, "tune: register-tile sweep needs tile_y 16/32/48 and 1/2/4 sweeps");
    if (kind == 3 && !reg_variant_ok(pr, 8, ty, m))
        return fail(IF_ERR_ARG, "tune: 8-wave register-tile sweep needs 2 sweeps with tile_y 24/40/56, "
                                "4 sweeps with tile_y 32/48/64 or 8 sweeps with tile_y 32/48/64");
    if (kind == 4 && !reg_variant_ok(pr, 16, ty, m))
        return fail(IF_ERR_ARG, "tune: 16-wave register-tile sweep needs 4 sweeps with tile_y 48/64/80 or 8 "
                                "sweeps with tile_y 32/64; mixed precision also 8 sweeps with tile_y 96, 9 "
                                "with 92, 10 with 88, 11 with 84 or 12 with 80");
    if (kind < 0 || kind > 4) return fail(IF_ERR_ARG, "tune: unknown sweep kernel");
    if (k < 1) return fail(IF_ERR_ARG, "tune: check_every must be >= 1");
    e.sweep_kind = kind;
    e.sweep_ty = ty;
    e.sweep_m = m;
    e.shape_auto = false;
    e.check_k = k;
    e.mark_dirty();
    return IF_OK;
}

if_status if_case_tune(if_case *c, int kernel, int tile_y, int m, int check_every)
{
    if (!c || c->mode != IF_MODE_PHYSICS) return fail(IF_ERR_STATE, "tune: physics mode only");
    return tune(*c->eng, kernel, tile_y, m, check_every);
}

if_status if_case_graphs(if_case *c, int enable)
{
    if (!c || c->mode != IF_MODE_PHYSICS) return fail(IF_ERR_STATE, "graphs: physics mode only");
    c->eng->use_graphs = enable != 0;
    return IF_OK;
}

if_status if_case_persistent(if_case *c, int enable)
{
    if (!c || c->mode != IF_MODE_PHYSICS) return fail(IF_ERR_STATE, "persistent: physics mode only");
    c->eng->use_persistent = enable != 0;
    c->eng->mark_dirty();
    return IF_OK;
}

static if_status case_autotune(if_case *c, if_tune_result *out)
{
    // a single case is resident but idle between calls: open it for one
    // step while timing, then close it again (its state is otherwise untouched)
    if_status rc = c->eng->set_target(c->slot, 1);
    if (rc != IF_OK) return rc;
    rc = c->eng->autotune(out);
    const if_status rc2 = c->eng->set_target(c->slot, 0);
    return rc != IF_OK ? rc : rc2;
}

if_status if_case_restart(if_case *c, const double *u, const double *v, const double *p, const double *uf,
                          const double *vf)
{
    if (!c || c->mode != IF_MODE_PHYSICS) return fail(IF_ERR_STATE, "restart: physics mode only");
    IFF_GUARD(c->eng->restart(c->slot, u, v, p, uf, vf));
}

if_status if_case_monitor(if_case *c, int nprobe, const double *xy, int capacity)
{
    if (!c || c->mode != IF_MODE_PHYSICS) return fail(IF_ERR_STATE, "monitor: physics mode only");
    IFF_GUARD(c->eng->monitor(c->slot, nprobe, xy, capacity));
}

if_status if_case_monitor_read(if_case *c, double *out, size_t cap_records, size_t *nrec, size_t *dropped)
{
    if (!c || c->mode != IF_MODE_PHYSICS) return fail(IF_ERR_STATE, "monitor: physics mode only");
    IFF_GUARD(c->eng->monitor_read(c->slot, out, cap_records, nrec, dropped));
}

if_status if_case_monitor_sample(if_case *c)
{
    if (!c || c->mode != IF_MODE_PHYSICS) return fail(IF_ERR_STATE, "monitor: physics mode only");
    IFF_GUARD(c->eng->monitor_sample(c->slot));
}

if_status if_case_monitor_sample_noalloc(if_case *c)
{
    if (!c || c->mode != IF_MODE_PHYSICS) return fail(IF_ERR_STATE, "monitor: physics mode only");
    IFF_GUARD(c->eng->monitor_sample_noalloc(c->slot));
}

if_status if_case_autotune(if_case *c, if_tune_result *out)
{
    if (!c || c->mode != IF_MODE_PHYSICS) return fail(IF_ERR_STATE, "autotune: physics mode only");
    IFF_GUARD(case_autotune(c, out));
}

if_status if_batch_autotune(if_batch *b, if_tune_result *out)
{
    if (!b) return fail(IF_ERR_ARG, "null batch");
    IFF_GUARD(b->eng->autotune(out));
}

if_status if_batch_tune(if_batch *b, int kernel, int tile_y, int m, int check_every)
{
    if (!b) return fail(IF_ERR_ARG, "null batch");
    return tune(*b->eng, kernel, tile_y, m, check_every);
}

if_status if_batch_graphs(if_batch *b, int enable)
{
    if (!b) return fail(IF_ERR_ARG, "null batch");
    b->eng->use_graphs = enable != 0;
    return IF_OK;
}

if_status if_batch_persistent(if_batch *b, int enable)
{
    if (!b) return fail(IF_ERR_ARG, "null batch");
    b->eng->use_persistent = enable != 0;
    b->eng->mark_dirty();
    return IF_OK;
}

if_status if_case_sweep_timing(if_case *c, int reset, double *ms, long long *launches)
{
    if (!c || c->mode != IF_MODE_PHYSICS) return fail(IF_ERR_STATE, "sweep timing: physics mode only");
    c->eng->timing = true;
    if (ms) *ms = c->eng->sweep_ms;
    if (launches) *launches = c->eng->sweep_launches;
    if (reset) {
        c->eng->sweep_ms = 0.0;
        c->eng->sweep_launches = 0;
    }
    return IF_OK;
}

// Algorithmic bytes per interior cell (HBM streams counted once per pass):
//  Poisson sweep launch (M sweeps): FP64 M*(p 8 + rhs 8 + flags 1 + p' 8);
//                                   mixed M*(e 4 + r 4 + flags 1 + e' 4)
//  step: P1 + K_AD*P2 + P3 + sweeps*P4/M + checks*P5 + P7 with the step's
//  last counts.
if_status if_case_byte_model(if_case *c, double *sweep_bpc, double *step_bpc)
{
    if (!c || c->mode != IF_MODE_PHYSICS) return fail(IF_ERR_STATE, "byte model: physics mode only");
    Engine &e = *c->eng;
    Slot *s = e.slot(c->slot);
    const bool mixed = s->prec != IF_FP64;  // FP32 Poisson error sweeps
    // sweeps per launch: M, or the whole chunk when the last step ran
    // persistent chunks
    const int M = s->dev.ppe_solver == IF_PPE_RBSOR ? (e.launch_sweeps > 0 ? e.launch_sweeps : e.sweep_m) : 1;
    const double ts = s->prec == IF_MIXED_V32 ? 4 : 8;  // velocity fields
    const double sweep = mixed ? 13.0 : 25.0;
    if (sweep_bpc) *sweep_bpc = sweep * M;
    if (step_bpc) {
        const double p1 = 4 * ts + 16 + 1 + 2 * ts;       // u,v,uf,vf,flags -> sx,sy
        const double p2 = 4 * ts + 1 + 2 * ts;            // u,v,sx,sy,flags -> u',v'
        const double p3 = 2 * ts + 1 + 16 + 8;            // u*,v*,flags -> uf*,vf*,rhs
        const double p5 = mixed ? 8 + 4 + 8 + 1 + 8 + 4 : 8 + 8 + 1;  // check (mixed: update+resid)
        const double p7 = 8 + 16 + 2 * ts + 1 + 16 + 2 * ts;          // p,uf*,vf*,u*,v*,flags -> uf,vf,u,v
        const int K = std::max(1, e.check_k);
        const double sweeps = s->last.ppe_sweeps;
        const double checks = 1 + std::ceil(sweeps / K);
        *step_bpc = p1 + s->last.ad_iters * p2 + p3 + sweeps * sweep + checks * p5 + p7;
    }
    return IF_OK;
}

// ---- batches -----------------------------------------------------------------
static if_status batch_create(int precision, size_t budget, if_batch **out)
{
    if (!out) return fail(IF_ERR_ARG, "if_batch_create: null output");
    *out = nullptr;
    if (precision != IF_FP64 && precision != IF_MIXED && precision != IF_MIXED_V32)
        return fail(IF_ERR_ARG, "if_batch_create: unknown precision");
    auto b = std::make_unique<if_batch>();
    b->eng = std::make_unique<Engine>(precision, budget);
    b->eng->timing = true;
    if_status rc = b->eng->init();
    if (rc != IF_OK) return rc;
    *out = b.release();
    return IF_OK;
}

if_status if_batch_create(int precision, size_t budget, if_batch **out)
{
    IFF_GUARD(batch_create(precision, budget, out));
}

void if_batch_destroy(if_batch *b) { delete b; }

if_status if_batch_add(if_batch *b, const if_case_desc *d, int nsteps, int *slot)
{
    if (!b || !d) return fail(IF_ERR_ARG, "if_batch_add: null argument");
    IFF_GUARD(b->eng->add(*d, nsteps, slot));
}

if_status if_batch_step(if_batch *b, int nsteps, int *finished)
{
    if (!b) return fail(IF_ERR_ARG, "null batch");
    IFF_GUARD(b->eng->step(nsteps, finished, nullptr));
}

if_status if_batch_slot_state(if_batch *b, int slot, int *state, int *steps_done, if_step_stats *st)
{
    if (!b) return fail(IF_ERR_ARG, "null batch");
    IFF_GUARD(b->eng->slot_state(slot, state, steps_done, st));
}

if_status if_batch_get_field(if_batch *b, int slot, int which, double *out, size_t n)
{
    if (!b) return fail(IF_ERR_ARG, "null batch");
    IFF_GUARD(b->eng->get_field(slot, which, out, n));
}

if_status if_batch_remove(if_batch *b, int slot)
{
    if (!b) return fail(IF_ERR_ARG, "null batch");
    IFF_GUARD(b->eng->remove(slot));
}

if_status if_batch_info(if_batch *b, int *resident, size_t *bytes, double *ms, long long *launches,
                        double *cell_sweeps)
{
    if (!b) return fail(IF_ERR_ARG, "null batch");
    if (resident) *resident = b->eng->resident();
    if (bytes) *bytes = b->eng->in_use();
    if (ms) *ms = b->eng->sweep_ms;
    if (launches) *launches = b->eng->sweep_launches;
    if (cell_sweeps) *cell_sweeps = b->eng->sweep_cells;
    return IF_OK;
}

if_status if_case_footprint(const if_case_desc *d, int precision, size_t *bytes)
{
    if (!d || !bytes) return fail(IF_ERR_ARG, "if_case_footprint: null argument");
    if (precision != IF_FP64 && precision != IF_MIXED && precision != IF_MIXED_V32)
        return fail(IF_ERR_ARG, "if_case_footprint: unknown precision");
    *bytes = Engine::footprint(*d, precision);
    return IF_OK;
}

}  // extern "C"
