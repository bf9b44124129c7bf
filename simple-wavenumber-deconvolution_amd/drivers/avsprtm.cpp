// avsprtm.cpp -- AVSPRTM: VSP reverse time migration and PSF imaging (the
// reference's acoustic_HLVSPRTMC.cpp main, RC:15-883).
//
// Runs in the reference's root directory layout: reads ./parameter.txt,
// ./input/acc_vp.bin and the up-going gathers AFor wrote to ./output/
// (<n>source_seismogram_obsup.bin, or ..._scatterup.bin when scatter_flag is 1),
// migrates them with the smoothed model and writes the per-shot images and the
// summed, Laplacian-filtered, edge-zeroed migration_{pp|psf}.bin.  With
// --updown the same run also images the up/down-separated wavefields
// (swd_rtm_updown) and writes migration_{pp|psf}du.bin and ..ud.bin
// (down-going source with up-going receiver field, and the reverse),
// normalised, filtered and summed like the plain image.  With --cig [--thetan
// N] [--delt N] it also computes the angle-domain common-image gathers
// (swd_rtm_cig; thetan 10 and delt 1 by default, RC:235) and writes their
// shot sum migration_{pp|psf}cig.bin, [nx][nz][thetan] (the reference's
// commented-out migration_ppcig.bin, RC:784-786).
#include <cmath>
#include <algorithm>
#include <cstdio>
#include <string>
#include <vector>

#include "driver_common.h"

using drv::check;

int main(int argc, char **argv)
{
    const drv::Options o = drv::parse_options(argc, argv, "AVSPRTM");
    int ndev = 0;
    check(swd_get_device_count(&ndev), "device query");
    std::printf("The available Device number is %d\n", ndev);
    drv::Setup s = drv::make_setup("./parameter.txt");
    const swd_params &p = s.p;
    const int nt = p.itmax;
    const char *kind = p.scatter_flag ? "psf" : "pp";

    std::vector<float> rick(nt);
    check(swd_ricker(rick.data(), nt, p.f0, 1.0f / p.f0, p.dt), "ricker");
    drv::write_floats("./output/rick_second_derive.bin", rick.data(), nt);

    // true model, then the smoothed migration model (RC:371-457)
    std::vector<float> vp(s.ntp), rho(s.ntp);
    check(swd_read_model("./input/acc_vp.bin", vp.data(), rho.data(), s.ntx, s.ntz, s.pmlc), "model input");
    drv::write_interior("./output/acc_vp.bin", s, vp.data());
    drv::write_interior("./output/acc_rho.bin", s, rho.data());
    check(swd_smooth_model(vp.data(), s.ntz, s.ntx, s.pmlc, p.window), "smoothing");
    drv::write_interior("./output/ini_vp.bin", s, vp.data());

    void *plan = nullptr;
    const swd_grid g = s.grid();
    const swd_coefs c = s.coefs();
    check(swd_plan_create(&g, o.device, &plan), "plan");
    check(swd_plan_set_coefs(plan, &c), "coefficients");
    check(swd_plan_set_model(plan, vp.data(), rho.data()), "model");
    check(swd_plan_set_path(plan, o.path), "path");
    check(swd_plan_set_tile(plan, o.tile), "tile");
    std::printf("====================\n %s RTM BEGIN\n====================\n", p.scatter_flag ? "PSF" : "NORMAL");

    std::vector<float> tmp1(s.np, 0.f), tmp3(s.np);
    // up/down imaging: the Hilbert-transformed wavelet and kz taper (FK:1257-1324, FC:354-377)
    std::vector<float> hrick(nt), kz, accdu, accud;
    int nfftz = 0;
    if (o.updown) {
        check(swd_hilbert1d(rick.data(), hrick.data(), nt), "hilbert");
        nfftz = std::max(128, swd_nfft(p.nz));   // interior rows padded to a power of two (whole columns if they fit)
        kz.resize(nfftz);
        check(swd_kz_taper(kz.data(), nfftz, p.dz), "kz taper");
        accdu.assign(s.np, 0.f);
        accud.assign(s.np, 0.f);
    }
    std::vector<float> acccig;
    if (o.cig) acccig.assign(static_cast<size_t>(s.np) * o.thetan, 0.f);
    const double t0 = drv::wall_seconds();
    for (size_t b0 = 0; b0 < s.shots.size(); b0 += o.batch) {
        const size_t b1 = std::min(s.shots.size(), b0 + static_cast<size_t>(o.batch));
        const int nb = static_cast<int>(b1 - b0);
        std::vector<swd_shot> sh;
        std::vector<size_t> row0;
        size_t ntr = 0;
        for (size_t k = b0; k < b1; k++) {
            const auto &S = s.shots[k];
            sh.push_back(swd_shot{S.s_ix, S.s_iz, static_cast<int>(S.rec_ix.size()), S.rec_ix.data(), S.rec_iz.data()});
            row0.push_back(ntr);
            ntr += S.rec_ix.size();
        }
        // up-going data, double time integration (RC:588-637)
        std::vector<float> d(ntr * nt), adj(ntr * nt);
        for (size_t k = b0; k < b1; k++) {
            const std::string f = "./output/" + std::to_string(k + 1) + "source_seismogram_" +
                                  (p.scatter_flag ? "scatterup" : "obsup") + ".bin";
            const size_t n = s.shots[k].rec_ix.size();
            drv::read_floats(f, d.data() + row0[k - b0] * nt, n * nt);
            check(swd_adjoint_source(d.data() + row0[k - b0] * nt, adj.data() + row0[k - b0] * nt,
                                     static_cast<int>(n), nt, p.dt),
                  "adjoint source");
        }
        std::vector<float> img(static_cast<size_t>(s.ntp) * nb), ils(img.size()), ilr(img.size()), syn(ntr * nt);
        std::vector<float> du, ud, cig;
        bool need_syn = true;   // the plain RTM call records the synthetic gathers on its own
        if (o.cig) {
            cig.resize(static_cast<size_t>(s.ntp) * o.thetan * nb);
            check(swd_rtm_cig(plan, nb, sh.data(), rick.data(), adj.data(), o.thetan, o.delt, img.data(), ils.data(),
                              ilr.data(), cig.data()),
                  "RTM with angle gathers");
        }
        if (o.updown) {
            // adjoint sources' Hilbert transforms along time, trace by trace
            std::vector<float> hadj(adj.size());
            for (size_t tr = 0; tr < ntr; tr++)
                check(swd_hilbert1d(adj.data() + tr * nt, hadj.data() + tr * nt, nt), "hilbert");
            du.resize(img.size());
            ud.resize(img.size());
            // the plain image comes from the angle-gather run when there is one
            const swd_updown_images out = {du.data(), ud.data(), nullptr, nullptr, o.cig ? nullptr : img.data(),
                                           o.cig ? nullptr : ils.data(), o.cig ? nullptr : ilr.data()};
            check(swd_rtm_updown(plan, nb, sh.data(), rick.data(), hrick.data(), adj.data(), hadj.data(), kz.data(),
                                 nfftz, &out),
                  "up/down RTM");
        }
        if (!o.cig && !o.updown) {
            check(swd_rtm(plan, nb, sh.data(), rick.data(), adj.data(), img.data(), ils.data(), ilr.data(),
                          syn.data()),
                  "RTM");
            need_syn = false;
        }
        if (need_syn) {
            bool diag = false;   // the synthetic gathers are only written for every 20th shot
            for (size_t k = b0; k < b1; k++) diag |= k % 20 == 0;
            if (diag) check(swd_forward(plan, nb, sh.data(), rick.data(), syn.data()), "forward");
        }
        if (o.cig) {   // raw gathers of the interior, summed over shots: [nx][nz][thetan]
            for (int k = 0; k < nb; k++)
                for (int ix = 0; ix < p.nx; ix++)
                    for (int iz = 0; iz < p.nz; iz++) {
                        const float *src = cig.data() + (static_cast<size_t>(s.ntp) * k +
                                                         static_cast<size_t>(ix + s.pmlc) * s.ntz + iz + s.pmlc) *
                                                            o.thetan;
                        float *dst = acccig.data() + (static_cast<size_t>(ix) * p.nz + iz) * o.thetan;
                        for (int t = 0; t < o.thetan; t++) dst[t] += src[t];
                    }
        }
        for (size_t k = b0; k < b1; k++) {
            const size_t n = s.shots[k].rec_ix.size();
            const std::string id = std::to_string(k + 1);
            if (k % 20 == 0) {   // diagnostics every 20th shot (RC:640-651)
                drv::write_floats("./output/" + id + "source_seismogram_syn.bin", syn.data() + row0[k - b0] * nt, n * nt);
                drv::write_floats("./output/" + id + "source_seismogram_frms.bin", adj.data() + row0[k - b0] * nt, n * nt);
            }
            float *I = img.data() + static_cast<size_t>(s.ntp) * (k - b0);
            const float *S = ils.data() + static_cast<size_t>(s.ntp) * (k - b0);
            check(swd_image_normalize(I, S, s.ntp), "normalize");   // RC:663-679
            drv::write_interior("./output/image_" + std::string(kind) + id + ".bin", s, I);
            drv::write_interior("./output/image_sources" + id + ".bin", s, S);
            check(swd_image_laplacian(I, vp.data(), tmp3.data(), tmp1.data(), s.ntx, s.ntz, s.pmlc, p.dx, p.dz),
                  "laplacian");   // RC:722-735
            drv::write_floats("./output/fimage_" + std::string(kind) + id + ".bin", tmp3.data(), tmp3.size());
            if (o.updown) {
                float *D = du.data() + static_cast<size_t>(s.ntp) * (k - b0);
                float *U = ud.data() + static_cast<size_t>(s.ntp) * (k - b0);
                check(swd_image_normalize(D, S, s.ntp), "normalize");
                check(swd_image_normalize(U, S, s.ntp), "normalize");
                check(swd_image_laplacian(D, vp.data(), tmp3.data(), accdu.data(), s.ntx, s.ntz, s.pmlc, p.dx, p.dz),
                      "laplacian");
                check(swd_image_laplacian(U, vp.data(), tmp3.data(), accud.data(), s.ntx, s.ntz, s.pmlc, p.dx, p.dz),
                      "laplacian");
            }
        }
        std::printf("shots %zu..%zu migrated\n", b0 + 1, b1);
    }
    check(swd_edge_zero(tmp1.data(), p.nx, p.nz, p.Lc), "edge zero");   // RC:754-768
    drv::write_floats(std::string("./output/migration_") + kind + ".bin", tmp1.data(), tmp1.size());
    if (o.cig)
        drv::write_floats(std::string("./output/migration_") + kind + "cig.bin", acccig.data(), acccig.size());
    if (o.updown) {
        check(swd_edge_zero(accdu.data(), p.nx, p.nz, p.Lc), "edge zero");
        check(swd_edge_zero(accud.data(), p.nx, p.nz, p.Lc), "edge zero");
        drv::write_floats(std::string("./output/migration_") + kind + "du.bin", accdu.data(), accdu.size());
        drv::write_floats(std::string("./output/migration_") + kind + "ud.bin", accud.data(), accud.size());
    }
    swd_plan_destroy(plan);
    std::printf("====================\n      THE END\n====================\n");
    std::printf("The cost of the run time is %f seconds\n", drv::wall_seconds() - t0);
    return 0;
}
