// afor.cpp -- AFor: forward modelling with kz up/down separation (the
// reference's forward/acoustic_hilbert_fftkz_HLForC.cpp main, FC:13-892).
//
// Runs in the reference's forward/ directory layout: reads ../parameter.txt
// and ../input/acc_vp.bin, writes to ../output/.  scatter_flag 0 writes the
// observed gathers of the true model; scatter_flag 1 writes the background
// gathers of the smoothed model, inserts density point scatterers and writes
// the scattered (difference) gathers.  Shots run in batches on one GPU.
#include <algorithm>
#include <cstdio>
#include <string>
#include <vector>

#include "driver_common.h"

using drv::check;

namespace {

struct Gathers {
    std::vector<float> p, up, down;
};

// Runs every shot through swd_afor in batches; out[is] holds shot is's gathers.
std::vector<Gathers> run_all(const drv::Setup &s, const drv::Options &o, const std::vector<float> &vp,
                             const std::vector<float> &rho, const std::vector<float> &rick,
                             const std::vector<float> &hrick, const std::vector<float> &kz)
{
    void *plan = nullptr;
    const swd_grid g = s.grid();
    const swd_coefs c = s.coefs();
    check(swd_plan_create(&g, o.device, &plan), "plan");
    check(swd_plan_set_coefs(plan, &c), "coefficients");
    check(swd_plan_set_model(plan, vp.data(), rho.data()), "model");
    check(swd_plan_set_path(plan, o.path), "path");
    check(swd_plan_set_tile(plan, o.tile), "tile");
    const int nt = s.p.itmax;
    std::vector<Gathers> out(s.shots.size());
    for (size_t b0 = 0; b0 < s.shots.size(); b0 += o.batch) {
        const size_t b1 = std::min(s.shots.size(), b0 + static_cast<size_t>(o.batch));
        std::vector<swd_shot> sh;
        size_t ntr = 0;
        for (size_t k = b0; k < b1; k++) {
            const auto &S = s.shots[k];
            sh.push_back(swd_shot{S.s_ix, S.s_iz, static_cast<int>(S.rec_ix.size()), S.rec_ix.data(), S.rec_iz.data()});
            ntr += S.rec_ix.size();
        }
        std::vector<float> gp(ntr * nt), gu(ntr * nt), gd(ntr * nt);
        check(swd_afor(plan, static_cast<int>(sh.size()), sh.data(), rick.data(), hrick.data(), kz.data(),
                       static_cast<int>(kz.size()), gp.data(), gu.data(), gd.data()),
              "forward modelling");
        size_t row = 0;
        for (size_t k = b0; k < b1; k++) {
            const size_t n = s.shots[k].rec_ix.size() * nt;
            out[k].p.assign(gp.begin() + row, gp.begin() + row + n);
            out[k].up.assign(gu.begin() + row, gu.begin() + row + n);
            out[k].down.assign(gd.begin() + row, gd.begin() + row + n);
            row += n;
        }
        std::printf("shots %zu..%zu done\n", b0 + 1, b1);
    }
    swd_plan_destroy(plan);
    return out;
}

void write_gathers(const std::vector<Gathers> &G, const char *tag)
{
    for (size_t is = 0; is < G.size(); is++) {
        const std::string base = "../output/" + std::to_string(is + 1) + "source_seismogram_" + tag;
        drv::write_floats(base + ".bin", G[is].p.data(), G[is].p.size());
        drv::write_floats(base + "up.bin", G[is].up.data(), G[is].up.size());
        drv::write_floats(base + "down.bin", G[is].down.data(), G[is].down.size());
    }
}

}  // namespace

int main(int argc, char **argv)
{
    const drv::Options o = drv::parse_options(argc, argv, "AFor");
    int ndev = 0;
    check(swd_get_device_count(&ndev), "device query");
    std::printf("The available Device number is %d\n", ndev);
    drv::Setup s = drv::make_setup("../parameter.txt");
    const swd_params &p = s.p;
    const int nt = p.itmax;

    // kz taper of the up/down separation (FC:354-377)
    const int nfftz = swd_nfft(s.ntz);
    std::vector<float> kz(nfftz);
    const int layerz = swd_kz_taper(kz.data(), nfftz, p.dz);
    check(layerz, "kz taper");
    std::printf("\nDamping wavenumber point No. = %d; NFFTz = %d\n\n", layerz, nfftz);

    // wavelet and its Hilbert transform (FC:382-403)
    std::vector<float> rick(nt), hrick(nt);
    check(swd_ricker(rick.data(), nt, p.f0, 1.0f / p.f0, p.dt), "ricker");
    drv::write_floats("../output/rick_second_derive.bin", rick.data(), nt);
    check(swd_hilbert1d(rick.data(), hrick.data(), nt), "hilbert");
    drv::write_floats("../output/rick_hilbert.bin", hrick.data(), nt);

    // true model (FC:409-440)
    std::vector<float> vp(s.ntp), rho(s.ntp);
    check(swd_read_model("../input/acc_vp.bin", vp.data(), rho.data(), s.ntx, s.ntz, s.pmlc), "model input");
    drv::write_interior("../output/acc_vp.bin", s, vp.data());
    drv::write_floats("../output/acc_vpful.bin", vp.data(), vp.size());
    drv::write_interior("../output/acc_rho.bin", s, rho.data());
    std::printf("The true model is done\n");

    const double t0 = drv::wall_seconds();
    if (p.scatter_flag == 0) {
        std::printf("Obtain the observed seismogram !\n");
        write_gathers(run_all(s, o, vp, rho, rick, hrick, kz), "obs");
    } else {
        // smoothed background (FC:620-633), background gathers (FC:640-684)
        check(swd_smooth_model(vp.data(), s.ntz, s.ntx, s.pmlc, p.window), "smoothing");
        drv::write_interior("../output/ini_vp.bin", s, vp.data());
        std::printf("Obtain the background scatter seismogram !\n");
        const std::vector<Gathers> bg = run_all(s, o, vp, rho, rick, hrick, kz);
        write_gathers(bg, "sobs");
        // density point scatterers (FC:697-705)
        check(swd_scatterers(rho.data(), vp.data(), s.ntx, s.ntz, s.pmlc, p.swin), "scatterers");
        drv::write_interior("../output/ini_scatter_vp.bin", s, vp.data());
        drv::write_interior("../output/rho_scatter.bin", s, rho.data());
        std::printf("Obtain the observed scatter seismogram !\n");
        std::vector<Gathers> sc = run_all(s, o, vp, rho, rick, hrick, kz);
        // scattered data = scatter-model gathers minus background gathers (FC:757-792)
        for (size_t is = 0; is < sc.size(); is++)
            for (size_t i = 0; i < sc[is].p.size(); i++) {
                sc[is].p[i] -= bg[is].p[i];
                sc[is].up[i] -= bg[is].up[i];
                sc[is].down[i] -= bg[is].down[i];
            }
        write_gathers(sc, "scatter");
    }
    std::printf("====================\n      THE END\n====================\n");
    std::printf("The cost of the run time is %f seconds\n", drv::wall_seconds() - t0);
    return 0;
}
