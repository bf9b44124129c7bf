// decon.cpp -- Decon: wavenumber-domain PSF deconvolution of a migrated image
// (the reference's kkdecon_acoustic/kk_decon_interp_psf_minewindow.cpp main,
// DC:17-785), one process on one GPU instead of MPI ranks over windows.
//
// Runs in the reference's kkdecon_acoustic directory layout: reads
// ./parameter.txt (nx, nz, window sizes, epsilon and the file names), the true
// and migration velocity models, the migrated image and the PSF image
// ([nx][nz] float32), and writes under ./output/ what decon2d_mine writes:
// acc_vp.bin, true_vertical_ref.bin, ini_vp.bin, acc_ref.bin, acc_rho.bin,
// start.flag, end.flag, migration_pp.bin, migration_coef.bin,
// migration_coef.txt and the deconvolved image (the parameter file's result
// name, migration_dudecon.bin in the reference's file).
#include <cstdio>
#include <string>
#include <vector>

#include "driver_common.h"

using drv::check;

namespace {

void touch(const char *path)
{
    FILE *fp = std::fopen(path, "w");
    if (!fp) drv::die("cannot write %s", path);
    std::fclose(fp);
}

}  // namespace

int main(int argc, char **argv)
{
    const drv::Options o = drv::parse_options(argc, argv, "Decon");
    swd_decon_params p;
    swd_decon_files f;
    check(swd_decon_read_parameters("./parameter.txt", &p, &f), "parameter file");
    swd_decon_geom g;
    check(swd_decon_geometry(&p, &g), "window geometry");
    const size_t np = static_cast<size_t>(p.nx) * p.nz;
    std::printf("nx=%d nz=%d psfwin=%dx%d filterwin=%dx%d snapwin=%dx%d epsilon=%g: %d windows of %dx%d\n", p.nx, p.nz,
                p.psfwinx, p.psfwinz, p.filterwinx, p.filterwinz, p.snapwinx, p.snapwinz, p.epsilon, g.windows, g.n,
                g.n);

    // models and the reflectivity written for reference (DC:199-252)
    std::vector<float> vp_old(np), vp(np), mn(np, 0.0f), rho(np, 1000.0f);
    drv::read_floats(f.vp, vp_old.data(), np);
    drv::write_floats("./output/acc_vp.bin", vp_old.data(), np);
    for (int ix = 0; ix < p.nx; ix++)
        for (int iz = 1; iz < p.nz; iz++) {
            const size_t ip = static_cast<size_t>(ix) * p.nz + iz;
            mn[ip] = (vp_old[ip] - vp_old[ip - 1]) / (vp_old[ip] + vp_old[ip - 1]);
        }
    drv::write_floats("./output/true_vertical_ref.bin", mn.data(), np);
    drv::read_floats(f.vp0, vp.data(), np);
    drv::write_floats("./output/ini_vp.bin", vp.data(), np);
    for (size_t ip = 0; ip < np; ip++) mn[ip] = (vp_old[ip] - vp[ip]) / vp_old[ip];
    drv::write_floats("./output/acc_ref.bin", mn.data(), np);
    drv::write_floats("./output/acc_rho.bin", rho.data(), np);
    std::printf("The true model is done\n");

    std::vector<float> image(np), psf(np), dec(np), coef(np);
    drv::read_floats(f.image, image.data(), np);
    drv::read_floats(f.psf, psf.data(), np);

    void *plan = nullptr;
    check(swd_decon_plan_create(&p, o.device, &plan), "deconvolution plan");
    touch("./output/start.flag");
    std::printf("Caculate the Deconvolution with the Wavenumber-domain PSFs !\n");
    const double t0 = drv::wall_seconds();
    check(swd_decon(plan, image.data(), psf.data(), dec.data(), coef.data()), "deconvolution");
    const double t1 = drv::wall_seconds();
    swd_decon_stats st;
    check(swd_decon_get_stats(plan, &st), "stats");
    std::printf("%d windows, %d PSF cells: %.3f ms on the GPU (windows %.3f ms), %.3f s with transfers\n",
                st.windows, st.cells, st.total_ms, st.windows_ms, t1 - t0);
    swd_decon_plan_destroy(plan);
    touch("./output/end.flag");

    // outputs (DC:692-731): the image as read, the summed taper weights and the result
    drv::write_floats("./output/migration_pp.bin", image.data(), np);
    drv::write_floats("./output/migration_coef.bin", coef.data(), np);
    FILE *fp = std::fopen("./output/migration_coef.txt", "w");
    if (!fp) drv::die("cannot write ./output/migration_coef.txt");
    for (int iz = 0; iz < p.nz; iz++) {
        for (int ix = 0; ix < p.nx; ix++) std::fprintf(fp, "%f  ", coef[static_cast<size_t>(ix) * p.nz + iz]);
        std::fprintf(fp, "\r\t\n");
    }
    std::fclose(fp);
    drv::write_floats(f.decon, dec.data(), np);
    std::printf("            Deconvolution END!     \n");
    return 0;
}
