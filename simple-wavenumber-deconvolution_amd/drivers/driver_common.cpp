// driver_common.cpp -- shared pieces of the AFor and AVSPRTM drivers.
#include "driver_common.h"

#include <chrono>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <unistd.h>

namespace drv {

void die(const char *fmt, ...)
{
    va_list ap;
    va_start(ap, fmt);
    std::vfprintf(stderr, fmt, ap);
    va_end(ap);
    std::fputc('\n', stderr);
    std::exit(1);
}

void check(int rc, const char *what)
{
    if (rc < 0) {
        char msg[1024];
        swd_last_error(msg, sizeof msg);
        die("%s failed (%d): %s", what, rc, msg);
    }
}

Options parse_options(int argc, char **argv, const char *prog)
{
    Options o;
    for (int i = 1; i < argc; i++) {
        std::string a = argv[i];
        auto next = [&]() -> const char * {
            if (i + 1 >= argc) die("%s: %s needs a value", prog, a.c_str());
            return argv[++i];
        };
        if (a == "--batch") o.batch = std::atoi(next());
        else if (a == "--device") o.device = std::atoi(next());
        else if (a == "--plain") o.path = SWD_PATH_PLAIN;
        else if (a == "--batched") o.path = SWD_PATH_BATCHED;
        else if (a == "--tile") o.tile = std::atoi(next());
        else if (a == "--updown") o.updown = true;
        else if (a == "--cig") o.cig = true;
        else if (a == "--thetan") o.thetan = std::atoi(next());
        else if (a == "--delt") o.delt = std::atoi(next());
        else if (a == "-C") o.workdir = next();
        else if (a == "-h" || a == "--help") {
            std::printf("usage: %s [-C dir] [--batch N] [--device D] [--plain | --batched] [--tile 0|16|32|64]"
                        " [--updown] [--cig [--thetan N] [--delt N]]\n", prog);
            std::exit(0);
        } else die("%s: unknown option %s", prog, a.c_str());
    }
    if (o.batch < 1) die("%s: --batch must be >= 1", prog);
    if (!o.workdir.empty() && chdir(o.workdir.c_str()) != 0) die("%s: cannot enter %s", prog, o.workdir.c_str());
    return o;
}

swd_grid Setup::grid() const
{
    swd_grid g;
    g.nx = p.nx; g.nz = p.nz; g.pml = p.pml; g.Lc = p.Lc;
    g.dx = p.dx; g.dz = p.dz; g.dt = p.dt; g.nt = p.itmax;
    return g;
}

swd_coefs Setup::coefs() const
{
    swd_coefs c;
    c.rc = rc.data();
    c.a_x = ax.data(); c.a_x_half = axh.data(); c.b_x = bx.data(); c.b_x_half = bxh.data();
    c.a_z = az.data(); c.a_z_half = azh.data(); c.b_z = bz.data(); c.b_z_half = bzh.data();
    return c;
}

Setup make_setup(const std::string &param_path)
{
    Setup s;
    check(swd_read_parameters(param_path.c_str(), &s.p), "reading parameters");
    const swd_params &p = s.p;
    s.pmlc = p.pml + p.Lc;
    s.ntx = p.nx + 2 * s.pmlc;
    s.ntz = p.nz + 2 * s.pmlc;
    s.ntp = s.ntx * s.ntz;
    s.np = p.nx * p.nz;
    s.rc.resize(p.Lc);
    check(swd_cal_xishu(p.Lc, s.rc.data()), "cal_xishu");
    double tmprc = 0;
    for (float r : s.rc) tmprc += std::fabs(r);
    std::printf("Maximum velocity for stability is %f m/s\n", p.dx / (tmprc * p.dt * std::sqrt(2.0)));   // RC:117
    // CPML over pml+Lc cells (RC:462-468); the vp_max argument is unused there
    s.ax.resize(s.ntx); s.axh.resize(s.ntx); s.bx.resize(s.ntx); s.bxh.resize(s.ntx);
    s.az.resize(s.ntz); s.azh.resize(s.ntz); s.bz.resize(s.ntz); s.bzh.resize(s.ntz);
    check(swd_cpml(s.ntx, s.ntz, s.pmlc, p.dx, p.dz, p.dt, s.ax.data(), s.axh.data(), s.bx.data(), s.bxh.data(),
                   s.az.data(), s.azh.data(), s.bz.data(), s.bzh.data()),
          "CPML coefficients");
    // shots and the two borehole receiver lines at ix = pmlc and ntx-pmlc-1 (RC:150-200)
    int rnmax = 0;
    for (int is = 0; is < p.ns; is++) {
        Setup::Shot sh;
        const int rn = swd_geometry(p.nx, p.nz, p.pml, p.Lc, p.dx, p.dz, p.sx0, p.shotdx, p.shotdep, p.offsetz, is,
                                    &sh.s_ix, &sh.s_iz, nullptr, 0);
        check(rn, "geometry");
        std::vector<int> r_iz(rn > 0 ? rn : 1);
        check(swd_geometry(p.nx, p.nz, p.pml, p.Lc, p.dx, p.dz, p.sx0, p.shotdx, p.shotdep, p.offsetz, is, &sh.s_ix,
                           &sh.s_iz, r_iz.data(), rn),
              "geometry");
        if (sh.s_ix >= s.ntx || sh.s_iz >= s.ntz) die("shot %d at (%d,%d) lies outside the grid", is + 1, sh.s_ix, sh.s_iz);
        for (int i = 0; i < rn; i++) {
            sh.rec_ix.push_back(i < rn / 2 ? s.pmlc : s.ntx - s.pmlc - 1);
            sh.rec_iz.push_back(r_iz[i]);
        }
        rnmax = rn > rnmax ? rn : rnmax;
        s.shots.push_back(std::move(sh));
    }
    std::printf("The maximum trace number for source is %d\n", rnmax);
    return s;
}

void write_floats(const std::string &path, const float *a, size_t n)
{
    FILE *fp = std::fopen(path.c_str(), "wb");
    if (!fp) die("Cannot write the file: %s", path.c_str());
    if (std::fwrite(a, sizeof(float), n, fp) != n) die("short write: %s", path.c_str());
    std::fclose(fp);
}

void read_floats(const std::string &path, float *a, size_t n)
{
    FILE *fp = std::fopen(path.c_str(), "rb");
    if (!fp) die("Cannot Open the File: %s", path.c_str());
    const size_t got = std::fread(a, sizeof(float), n, fp);
    std::fclose(fp);
    if (got != n) die("File %s holds %zu of %zu floats", path.c_str(), got, n);
}

void write_interior(const std::string &path, const Setup &s, const float *padded)
{
    std::vector<float> out(static_cast<size_t>(s.np));
    for (int ix = 0; ix < s.p.nx; ix++)
        std::memcpy(&out[static_cast<size_t>(ix) * s.p.nz], padded + static_cast<size_t>(ix + s.pmlc) * s.ntz + s.pmlc,
                    sizeof(float) * s.p.nz);
    write_floats(path, out.data(), out.size());
}

double wall_seconds()
{
    return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

}  // namespace drv
