// immerseflow: command-line driver with the call order and file paths of
// the reference's main() (src/main.cu:75-102), on the C ABI.
//
//   immerseflow [--deck P] [--xgrid P] [--ygrid P] [--results DIR]
//               [--steps N] [--write-every N] [--restart DIR] [--checkpoint DIR]
//
// Defaults are the reference's hard-coded paths: ../inputs/inputs.txt,
// ../inputs/xgrid.dat2, ../inputs/ygrid.dat2, ../results/.  Compat decks
// (no "Mode" block) replay ADsolver tmax times and write uc.dat/vc.dat after
// every step (ADSolver.cu:378-379); physics decks run full fractional steps
// and write uc/vc/p every Write_Interval steps and at the end.
//
// Restart (physics decks): the deck's "Restart Restart_Time" line with
// Restart = 1, or --restart DIR, loads DIR/uc.dat, vc.dat, p.dat (default
// DIR: the results directory) and, when present, the faces uf.dat, vf.dat;
// the run continues from time Restart_Time.  --checkpoint DIR also writes
// those five files at 17 significant digits at every write, a lossless
// restart point.
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "immerseflow/if_abi.h"

static int die(const char *what, if_status rc)
{
    std::fprintf(stderr, "immerseflow: %s failed (%d): %s\n", what, rc, if_last_error());
    return 1;
}

int main(int argc, char **argv)
{
    std::string deck = "../inputs/inputs.txt", xg = "../inputs/xgrid.dat2", yg = "../inputs/ygrid.dat2";
    std::string res = "../results";
    std::string restart_dir, ckpt_dir;
    long steps = -1, write_every = -1;
    for (int k = 1; k < argc; k++) {
        std::string a = argv[k];
        auto val = [&](void) -> const char * {
            if (k + 1 >= argc) { std::fprintf(stderr, "missing value for %s\n", a.c_str()); std::exit(2); }
            return argv[++k];
        };
        if (a == "--deck") deck = val();
        else if (a == "--xgrid") xg = val();
        else if (a == "--ygrid") yg = val();
        else if (a == "--results") res = val();
        else if (a == "--steps") steps = std::atol(val());
        else if (a == "--write-every") write_every = std::atol(val());
        else if (a == "--restart") restart_dir = val();
        else if (a == "--checkpoint") ckpt_dir = val();
        else {
            std::fprintf(stderr, "usage: immerseflow [--deck P] [--xgrid P] [--ygrid P] [--results DIR] "
                                 "[--steps N] [--write-every N] [--restart DIR] [--checkpoint DIR]\n");
            return 2;
        }
    }
    if_inputs *in = new if_inputs();
    if_status rc = if_parse_inputs(deck.c_str(), in);
    if (rc) return die("readInputFile", rc);
    int ndev = 0, cus = 0;
    size_t hbm = 0;
    char name[64];
    rc = if_device_query(&ndev, name, sizeof(name), &cus, &hbm);
    if (rc) return die("device query", rc);
    std::printf("device %s: %d CUs, %zu MiB\n", name, cus, hbm >> 20);
    std::printf("nx, ny, %d, %d\n", in->nx - 2, in->ny - 2);
    std::vector<double> xf(in->nxf), yf(in->nyf);
    if ((rc = if_read_grid(xg.c_str(), in->nxf, xf.data()))) return die("read x grid", rc);
    if ((rc = if_read_grid(yg.c_str(), in->nyf, yf.data()))) return die("read y grid", rc);

    if_case_desc d;
    std::memset(&d, 0, sizeof(d));
    d.mode = in->mode;
    d.precision = in->precision;
    d.nxf = in->nxf;
    d.nyf = in->nyf;
    d.xf = xf.data();
    d.yf = yf.data();
    d.dt = in->dt;
    d.re = in->re;
    d.conv_factor = in->conv_factor;
    d.ubc = 1.0;
    d.vbc = 0.0;
    d.ad_itermax = in->ad_itermax;
    d.ad_tol = 1e-6;  // ADSolver.cu:315
    d.ppe_solver = in->ppe_solver == 0 ? IF_PPE_JACOBI : in->ppe_solver == 2 ? IF_PPE_LINE : IF_PPE_RBSOR;
    d.ppe_omega = in->ppe_omega;
    d.ppe_itermax = in->ppe_itermax;
    d.ppe_tol = in->ppe_tol;
    d.check_every = in->check_every;
    d.nbody = in->nbody;
    d.bodies = in->bodies;
    // w_AD: the deck's integer (CFDInput::w_AD, globalVariables.cuh:27) or
    // the real value of the "Mode" line
    d.ad_omega = in->ad_omega > 0 ? in->ad_omega : in->w_ad > 0 ? (double)in->w_ad : 1.0;
    const bool restart = in->mode == IF_MODE_PHYSICS && (in->restart == 1 || !restart_dir.empty());
    if (restart) {
        if (restart_dir.empty()) restart_dir = res;
        d.t0 = in->restart_time;
    }
    if_case *c = nullptr;
    if ((rc = if_case_create(&d, &c))) return die("initializeData", rc);
    int nx = 0, ny = 0;
    if_case_dims(c, &nx, &ny);
    std::vector<double> xc(nx), yc(ny), f((size_t)nx * ny);
    if_case_grid(c, xc.data(), yc.data(), nullptr, nullptr);
    auto save = [&](int which, const char *file) -> int {
        if_status r = if_case_get_field(c, which, f.data(), f.size());
        if (r) return die("get field", r);
        const std::string path = res + "/" + file;
        r = if_write_tecplot(path.c_str(), nx, ny, xc.data(), yc.data(), f.data());
        if (r) return die(path.c_str(), r);
        return 0;
    };
    if (save(IF_FIELD_IBLANK, "final_results.dat")) return 1;  // preSim.cu:217
    // face coordinates of uf ((nx-1) x (ny-2)) and vf ((nx-2) x (ny-1))
    std::vector<double> xm(nx - 1), ym(ny - 1);
    for (int i = 0; i + 1 < nx; i++) xm[i] = 0.5 * (xc[i] + xc[i + 1]);
    for (int j = 0; j + 1 < ny; j++) ym[j] = 0.5 * (yc[j] + yc[j + 1]);
    if (restart) {
        std::vector<double> u(f.size()), v(f.size()), p(f.size());
        std::vector<double> uf((size_t)(nx - 1) * (ny - 2)), vf((size_t)(nx - 2) * (ny - 1));
        auto rd = [&](const char *file, int mx, int my, double *dst) -> if_status {
            const std::string path = restart_dir + "/" + file;
            return if_read_tecplot(path.c_str(), mx, my, nullptr, nullptr, dst);
        };
        if ((rc = rd("uc.dat", nx, ny, u.data())) || (rc = rd("vc.dat", nx, ny, v.data())) ||
            (rc = rd("p.dat", nx, ny, p.data())))
            return die("restart read", rc);
        const bool faces = rd("uf.dat", nx - 1, ny - 2, uf.data()) == IF_OK &&
                           rd("vf.dat", nx - 2, ny - 1, vf.data()) == IF_OK;
        rc = if_case_restart(c, u.data(), v.data(), p.data(), faces ? uf.data() : nullptr,
                             faces ? vf.data() : nullptr);
        if (rc) return die("restart", rc);
        std::printf("restart from %s at t = %d (%s)\n", restart_dir.c_str(), in->restart_time,
                    faces ? "with faces" : "faces from u, v");
    }
    auto checkpoint = [&]() -> int {
        if (ckpt_dir.empty()) return 0;
        struct Item {
            int which;
            const char *file;
            int mx, my;
            const double *x, *y;
        } items[] = {{IF_FIELD_U, "uc.dat", nx, ny, xc.data(), yc.data()},
                     {IF_FIELD_V, "vc.dat", nx, ny, xc.data(), yc.data()},
                     {IF_FIELD_P, "p.dat", nx, ny, xc.data(), yc.data()},
                     {IF_FIELD_UF, "uf.dat", nx - 1, ny - 2, xm.data(), yc.data() + 1},
                     {IF_FIELD_VF, "vf.dat", nx - 2, ny - 1, xc.data() + 1, ym.data()}};
        std::vector<double> g(f.size());
        for (const Item &it : items) {
            const size_t n = (size_t)it.mx * it.my;
            if_status r = if_case_get_field(c, it.which, g.data(), n);
            if (r) return die("get field", r);
            const std::string path = ckpt_dir + "/" + it.file;
            r = if_write_tecplot_digits(path.c_str(), it.mx, it.my, it.x, it.y, g.data(), 17);
            if (r) return die(path.c_str(), r);
        }
        return 0;
    };

    const long nsteps = steps >= 0 ? steps : (long)in->tmax;
    if (in->mode == IF_MODE_COMPAT) {
        for (long s = 0; s < nsteps; s++) {  // main.cu:93-96
            if_step_stats st;
            if ((rc = if_case_ad_step(c, &st))) return die("ADsolver", rc);
            std::printf("step %ld: %d AD iterations, residual %g\n", s + 1, st.ad_iters, st.ad_res);
            const long we = write_every > 0 ? write_every : 1;
            if ((s + 1) % we == 0 || s + 1 == nsteps) {
                if (save(IF_FIELD_U, "uc.dat") || save(IF_FIELD_V, "vc.dat")) return 1;
            }
        }
    } else {
        const long we = write_every > 0 ? write_every : (in->write_interval > 0 ? in->write_interval : nsteps);
        long done = 0;
        while (done < nsteps) {
            const long n = std::min(we, nsteps - done);
            if_step_stats st;
            if ((rc = if_case_step(c, (int)n, &st))) return die("fractional step", rc);
            done += n;
            std::printf("step %ld: AD %d, Poisson sweeps %d, max|div| %.3e\n", done, st.ad_iters,
                        st.ppe_sweeps, st.div_max);
            if (save(IF_FIELD_U, "uc.dat") || save(IF_FIELD_V, "vc.dat") || save(IF_FIELD_P, "p.dat") ||
                checkpoint())
                return 1;
        }
    }
    if (save(IF_FIELD_U, "uc.dat")) return 1;  // freeAllocation, preSim.cu:165
    if_case_destroy(c);
    delete in;
    return 0;
}
