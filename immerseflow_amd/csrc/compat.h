// Compat-mode case (compat.hip): the reference's ADsolver / PPESolver as
// coded, single case, reference layout.
#pragma once
#include <vector>

#include "common.h"

namespace iff {

struct CompatCase {
    int nx = 0, ny = 0, nblk = 0;
    double dt = 0, re = 0;
    double w_ad = 1.0;  // w_AD relaxation (1: none, as the reference runs)
    int ad_itermax = 0, ppe_itermax = 0;
    int cur = 0, pcur = 0;
    long long steps = 0;
    std::vector<double> xc, yc, dx, dy;
    hipStream_t st = nullptr;
    double *mem = nullptr;
    double *u[2] = {}, *v[2] = {}, *p[2] = {};
    double *sx = nullptr, *sy = nullptr, *iblank = nullptr, *pres = nullptr;
    double *uf = nullptr, *vf = nullptr, *vx = nullptr, *vy = nullptr;
    double *part = nullptr, *scal = nullptr;
    double *h_out = nullptr;

    ~CompatCase();
    if_status create(const if_case_desc &d);
    if_status ad_step(if_step_stats *st);
    if_status ppe_solve(if_step_stats *st);
    if_status get_field(int which, double *out, size_t n);
    if_status set_field(int which, const double *in, size_t n);

private:
    void bc(double *uu, double *vv);
    void velf(const double *uu, const double *vv);
};

}  // namespace iff
