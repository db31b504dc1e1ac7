// Host file I/O (io.cpp).
#pragma once
#include "common.h"

namespace iff {
if_status parse_inputs(const char *path, if_inputs *out);
if_status read_grid(const char *path, int nf, double *xf);
if_status grid_metric(int nf, const double *xf, double *xc, double *dx);
if_status write_tecplot(const char *path, int nx, int ny, const double *x, const double *y,
                        const double *f);
if_status write_tecplot_digits(const char *path, int nx, int ny, const double *x, const double *y,
                               const double *f, int digits);
if_status read_tecplot(const char *path, int nx, int ny, double *x, double *y, double *f);
const char *last_error();
}  // namespace iff
