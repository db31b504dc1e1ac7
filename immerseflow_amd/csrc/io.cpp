// Host file I/O: the reference's deck grammar, grid files, metric and
// Tecplot writer.  No GPU work here.
#include <cmath>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <sstream>
#include <string>
#include <vector>

#include "common.h"
#include "io.h"

namespace iff {

static thread_local std::string g_last_error;

void set_error(const std::string &msg) { g_last_error = msg; }

if_status fail(if_status code, const std::string &msg)
{
    g_last_error = msg;
    return code;
}

const char *last_error() { return g_last_error.c_str(); }

// Star body of a deck line: r(t) = R (1 + sum_{k=2..6} a_k cos(k t + phi_k))
// sampled at t_m = 2 pi m / n (polygon vertices, absolute coordinates).
static void star_vertices(double xc, double yc, double R, const double *a,
                          const double *phi, int n, double *vx, double *vy,
                          double *rbound)
{
    double rmax = 0.0;
    for (int m = 0; m < n; m++) {
        const double t = 2.0 * M_PI * m / n;
        double f = 1.0;
        for (int k = 2; k <= 6; k++) f += a[k - 2] * std::cos(k * t + phi[k - 2]);
        const double r = R * f;
        vx[m] = xc + r * std::cos(t);
        vy[m] = yc + r * std::sin(t);
        if (r > rmax) rmax = r;
    }
    *rbound = rmax * 1.0001 + 1e-12;
}

// Parse one body spec line of the extension block.
static bool parse_body(const std::string &line, if_inputs *in)
{
    if (in->nbody >= IF_MAX_BODIES) return false;
    std::istringstream iss(line);
    std::string kind;
    iss >> kind;
    if_body b;
    std::memset(&b, 0, sizeof(b));
    b.cth = 1.0;
    if (kind == "circle" || kind == "heave") {
        b.kind = IF_BODY_CIRCLE;
        iss >> b.xc >> b.yc >> b.r;
        if (kind == "heave") {
            b.motion = IF_MOTION_HEAVE;
            iss >> b.amp >> b.freq;
        }
    } else if (kind == "ellipse") {
        double deg = 0.0;
        b.kind = IF_BODY_ELLIPSE;
        iss >> b.xc >> b.yc >> b.a >> b.b >> deg;
        b.cth = std::cos(deg * M_PI / 180.0);
        b.sth = std::sin(deg * M_PI / 180.0);
    } else if (kind == "star") {
        double R = 0, a[5] = {0}, phi[5] = {0};
        int n = 400;
        b.kind = IF_BODY_POLYGON;
        iss >> b.xc >> b.yc >> R;
        for (double &x : a) iss >> x;
        for (double &x : phi) iss >> x;
        iss >> n;
        if (n < 3 || n > 512) return false;
        const int k = in->nbody;
        star_vertices(b.xc, b.yc, R, a, phi, n, in->poly_store[k][0],
                      in->poly_store[k][1], &b.r);
        b.nvert = n;
        b.vx = in->poly_store[k][0];
        b.vy = in->poly_store[k][1];
    } else {
        return false;
    }
    if (iss.fail()) return false;
    in->bodies[in->nbody++] = b;
    return true;
}

// readInputFile, src/main.cu:10-59: keyword lines matched by substring, the
// values on the following line, read with stream extraction into the
// CFDInput fields; header/separator/empty lines skipped (:19).  Then
// nxf = nx+1, nx += 2 (:55-58).  Keyword blocks unknown to the reference
// (whose keywords avoid the reference's substrings) carry the physics-mode
// settings and the bodies.
if_status parse_inputs(const char *path, if_inputs *out)
{
    if (!path || !out) return fail(IF_ERR_ARG, "parse_inputs: null argument");
    std::ifstream f(path);
    if (!f.is_open()) return fail(IF_ERR_IO, std::string("Unable to open file: ") + path);
    std::memset(out, 0, sizeof(*out));
    out->mode = IF_MODE_COMPAT;
    out->precision = IF_FP64;
    out->conv_factor = 0.5;
    out->ppe_tol = 1e-8;
    out->check_every = 10;
    std::string line;
    while (std::getline(f, line)) {
        if (line.empty() || line[0] == '=' || line[0] == '_') continue;
        std::istringstream iss(line);
        if (line.find("Restart") != std::string::npos) {
            std::getline(f, line); iss.str(line); iss.clear();
            iss >> out->restart >> out->restart_time;
        } else if (line.find("nx") != std::string::npos) {
            std::getline(f, line); iss.str(line); iss.clear();
            iss >> out->nx >> out->ny;
        } else if (line.find("Lx") != std::string::npos) {
            std::getline(f, line); iss.str(line); iss.clear();
            iss >> out->lx >> out->ly;
        } else if (line.find("w-AD") != std::string::npos) {
            std::getline(f, line); iss.str(line); iss.clear();
            iss >> out->w_ad >> out->w_ppe >> out->ad_itermax >> out->ppe_itermax
                >> out->ad_solver >> out->ppe_solver;
        } else if (line.find("ErrorMax") != std::string::npos) {
            std::getline(f, line); iss.str(line); iss.clear();
            iss >> out->error_max >> out->tmax >> out->dt >> out->re >> out->mu;
        } else if (line.find("Write Interval") != std::string::npos) {
            std::getline(f, line); iss.str(line); iss.clear();
            iss >> out->write_interval;
        } else if (line.find("Mode") != std::string::npos) {
            std::getline(f, line); iss.str(line); iss.clear();
            std::string mode, prec;
            iss >> mode >> prec >> out->conv_factor >> out->ppe_omega >> out->ppe_tol
                >> out->check_every;
            double w = 0.0;  // optional real-valued w_AD
            if (iss >> w) out->ad_omega = w;
            if (mode == "physics") out->mode = IF_MODE_PHYSICS;
            else if (mode == "compat") out->mode = IF_MODE_COMPAT;
            else return fail(IF_ERR_PARSE, "unknown mode '" + mode + "'");
            if (prec == "fp64") out->precision = IF_FP64;
            else if (prec == "mixed") out->precision = IF_MIXED;
            else if (prec == "mixed_v32") out->precision = IF_MIXED_V32;
            else return fail(IF_ERR_PARSE, "unknown precision '" + prec + "'");
        } else if (line.find("Body") != std::string::npos) {
            std::getline(f, line);
            if (!parse_body(line, out))
                return fail(IF_ERR_PARSE, "bad body line: '" + line + "'");
        }
    }
    out->nxf = out->nx + 1;
    out->nyf = out->ny + 1;
    out->nx += 2;
    out->ny += 2;
    if (out->nx < 3 || out->ny < 3)
        return fail(IF_ERR_PARSE, "deck has no positive 'nx ny' entry");
    return IF_OK;
}

// The file part of readGridData (preSim.cu:267-291): nf lines "<id> <x>".
// A short file is an error here (the reference reads garbage, :275-277).
if_status read_grid(const char *path, int nf, double *xf)
{
    if (!path || !xf || nf < 2) return fail(IF_ERR_ARG, "read_grid: bad argument");
    std::ifstream f(path);
    if (!f) return fail(IF_ERR_IO, std::string("Error opening ") + path);
    for (int i = 0; i < nf; i++) {
        long id;
        if (!(f >> id >> xf[i]))
            return fail(IF_ERR_PARSE, std::string(path) + ": fewer than " +
                                          std::to_string(nf) + " grid lines");
    }
    for (int i = 1; i < nf; i++)
        if (!(xf[i] > xf[i - 1]))
            return fail(IF_ERR_PARSE, std::string(path) + ": face coordinates not increasing");
    return IF_OK;
}

// The metric part of readGridData (preSim.cu:293-355), per axis.
if_status grid_metric(int nf, const double *xf, double *xc, double *dx)
{
    if (nf < 2 || !xf || !xc || !dx) return fail(IF_ERR_ARG, "grid_metric: bad argument");
    const int n = nf + 1;
    for (int i = 1; i < n - 1; i++) xc[i] = (xf[i - 1] + xf[i]) / 2.0;
    xc[0] = -1 * xc[1];
    xc[n - 1] = xf[nf - 1] + (xf[nf - 1] - xc[n - 2]);
    for (int i = 1; i < n - 1; i++) dx[i] = xf[i] - xf[i - 1];
    dx[0] = dx[1];
    dx[n - 1] = dx[n - 2];
    return IF_OK;
}

// write_results_to_file, postSim.cu:41-66: three header lines, then
// "%f,%f,%f" rows, j outer, i inner, ghosts included.
if_status write_tecplot(const char *path, int nx, int ny, const double *x,
                        const double *y, const double *f)
{
    return write_tecplot_digits(path, nx, ny, x, y, f, 0);
}

// The same file with every value at `digits` significant digits ("%.*e");
// digits = 0 is the reference's "%f".  17 digits round-trip FP64 exactly,
// which makes the file a lossless checkpoint.
if_status write_tecplot_digits(const char *path, int nx, int ny, const double *x,
                               const double *y, const double *f, int digits)
{
    if (!path || !x || !y || !f || nx < 1 || ny < 1 || digits < 0 || digits > 17)
        return fail(IF_ERR_ARG, "write_tecplot: bad argument");
    FILE *fp = std::fopen(path, "w");
    if (!fp) return fail(IF_ERR_IO, std::string("Error opening file: ") + path);
    std::vector<char> buf(1 << 20);
    std::setvbuf(fp, buf.data(), _IOFBF, buf.size());
    std::fprintf(fp, "TITLE = \"Post Processing Tecplot\"\n");
    std::fprintf(fp, "VARIABLES = \"X\",\"Y\",\"T\"\n");
    std::fprintf(fp, "ZONE T=\"BIG ZONE\", I=%d, J=%d, DATAPACKING=POINT\n", nx, ny);
    if (digits == 0) {
        for (int j = 0; j < ny; j++)
            for (int i = 0; i < nx; i++)
                std::fprintf(fp, "%f,%f,%f\n", x[i], y[j], f[i + (size_t)j * nx]);
    } else {
        const int d = digits - 1;
        for (int j = 0; j < ny; j++)
            for (int i = 0; i < nx; i++)
                std::fprintf(fp, "%.*e,%.*e,%.*e\n", d, x[i], d, y[j], d, f[i + (size_t)j * nx]);
    }
    const bool bad = std::ferror(fp);
    std::fclose(fp);
    if (bad) return fail(IF_ERR_IO, std::string("write failed: ") + path);
    return IF_OK;
}

// Reader of the writer's format (a restart source): the three header lines,
// the ZONE line's I and J must equal nx and ny, then nx*ny rows "x,y,value"
// (j outer).  x (nx values), y (ny values) may be null.  The predecessor
// restarts from such files (test/UTIL_PRE_SIM.f90:130-149).
if_status read_tecplot(const char *path, int nx, int ny, double *x, double *y, double *f)
{
    if (!path || !f || nx < 1 || ny < 1) return fail(IF_ERR_ARG, "read_tecplot: bad argument");
    FILE *fp = std::fopen(path, "r");
    if (!fp) return fail(IF_ERR_IO, std::string("Error opening file: ") + path);
    std::vector<char> buf(1 << 20);
    std::setvbuf(fp, buf.data(), _IOFBF, buf.size());
    char line[512];
    int fi = -1, fj = -1;
    bool ok = true;
    for (int h = 0; h < 3 && ok; h++) {
        if (!std::fgets(line, sizeof(line), fp)) ok = false;
        else if (h == 2) {
            const char *pi = std::strstr(line, "I="), *pj = std::strstr(line, "J=");
            if (!pi || !pj || std::strncmp(line, "ZONE", 4) != 0) ok = false;
            else {
                fi = std::atoi(pi + 2);
                fj = std::atoi(pj + 2);
            }
        }
    }
    if (!ok || fi != nx || fj != ny) {
        std::fclose(fp);
        return fail(IF_ERR_PARSE, std::string(path) + ": header is not a " + std::to_string(nx) + " x " +
                                      std::to_string(ny) + " Tecplot POINT zone");
    }
    for (int j = 0; j < ny && ok; j++)
        for (int i = 0; i < nx && ok; i++) {
            double a, b, v;
            if (std::fscanf(fp, " %lf , %lf , %lf", &a, &b, &v) != 3) {
                ok = false;
                break;
            }
            if (x && j == 0) x[i] = a;
            if (y && i == 0) y[j] = b;
            f[i + (size_t)j * nx] = v;
        }
    std::fclose(fp);
    if (!ok) return fail(IF_ERR_PARSE, std::string(path) + ": fewer than " + std::to_string((size_t)nx * ny) +
                                           " data rows");
    return IF_OK;
}

}  // namespace iff
