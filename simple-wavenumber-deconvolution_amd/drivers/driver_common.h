// driver_common.h -- shared pieces of the AFor and AVSPRTM drivers.
#pragma once
#include <string>
#include <vector>

#include "swd.h"

namespace drv {

// Options added to the reference command line (none are required).
struct Options {
    int batch = 16;          // shots per GPU batch
    int device = 0;
    int path = SWD_PATH_FUSED;
    int tile = 0;            // fused tile width (0 = tuned)
    bool updown = false;     // AVSPRTM: also the up/down-separated images (swd_rtm_updown)
    bool cig = false;        // AVSPRTM: also the angle-domain common-image gathers (swd_rtm_cig)
    int thetan = 10, delt = 1;   // their bins and time averaging (RC:235)
    std::string workdir;     // run as if started in this directory
};
Options parse_options(int argc, char **argv, const char *prog);

[[noreturn]] void die(const char *fmt, ...);
void check(int rc, const char *what);   // exits with the library's last error on rc < 0

// Grid and per-shot geometry of a parameter set (RC:78-200).
struct Setup {
    swd_params p{};
    int pmlc = 0, ntx = 0, ntz = 0, ntp = 0, np = 0;
    std::vector<float> rc;
    std::vector<float> ax, axh, bx, bxh, az, azh, bz, bzh;
    struct Shot {
        int s_ix, s_iz;
        std::vector<int> rec_ix, rec_iz;
    };
    std::vector<Shot> shots;
    swd_grid grid() const;
    swd_coefs coefs() const;
};
Setup make_setup(const std::string &param_path);

void write_floats(const std::string &path, const float *a, size_t n);
void read_floats(const std::string &path, float *a, size_t n);   // dies if missing or short
void write_interior(const std::string &path, const Setup &s, const float *padded);   // [nx][nz] of [ntx][ntz]

double wall_seconds();

}  // namespace drv
