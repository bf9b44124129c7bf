"""The reference's RTM run on its own inputs, against the images it shipped.

The reference ships a parameter file (845 x 702 cells, Lc 5, 7000 steps, 210
shots, two borehole receiver lines, smoothing window 5),
the true model it names and, as the deconvolution's inputs, the migration
images of that run: migration_pp.bin (the observed data) and
migration_psf.bin (the point-scatterer run), kept in
tests/golden/shipped_rtm.npz by tests/golden/make_shipped_golden.py.  This
test runs the same pipeline through the library: AFor (FC:574-795) for the
up-going data of every shot -- with scatter_flag 0 on the true model, with
scatter_flag 1 the scatterer-model gathers minus the background's, both on
the smoothed model -- then the RTM of that data on the smoothed model, each
shot's image normalised by its source illumination, Laplacian-filtered and
summed, edges zeroed (RC:552-793).  It is the one check against output that
the reference itself produced.  The reference's code and compiler rounding
are not this build's (its images sum 210 shots of 7000 steps, with MPI
reductions), so the bar is agreement to single-precision accumulation, not
bit parity: correlation, amplitude and relative L2 of the whole image.  The
PSF image's data are differences of two runs (scatterer model minus
background), which magnify every rounding difference of the two builds, so
its bar is looser.

Scatterer window: the shipped parameter file holds 51 on the line labelled
"scatter window - 31", and the shipped PSF image is a lattice of period 31
(its autocorrelation peaks at lags 31 and 62 along x and z; with 51 the
correlation is 0.004).  The PSF run therefore uses 31."""
import json
import os

import numpy as np
import pytest

import swd

pytestmark = pytest.mark.gpu
HERE = os.path.dirname(os.path.abspath(__file__))


def shipped():
    z = np.load(os.path.join(HERE, "golden", "shipped_rtm.npz"))
    return z, json.loads(str(z["meta"]))["params"]


def migrate(z, p, kind, tmp, swin=None, batch=30):
    """migration_pp (kind 'pp') or migration_psf ('psf') of the parameter set p
    (swin: the scatterer window, else p's)."""
    nx, nz, Lc, nt = p["nx"], p["nz"], p["Lc"], p["itmax"]
    grid = swd.Grid(nx, nz, p["pml"], Lc, p["dx"], p["dz"], p["dt"], nt)
    pmlc, ntx, ntz = grid.pmlc, grid.ntx, grid.ntz
    coefs = swd.coefs_for(grid)
    f = os.path.join(tmp, "acc_vp.bin")
    z["acc_vp"].tofile(f)
    vp, rho = swd.read_model(f, ntx, ntz, pmlc)                         # RC:371-440
    vps = swd.smooth_model(vp, ntz, ntx, pmlc, p["window"])             # RC:441-457, FC:620-633
    rick = swd.ricker(nt, p["f0"], 1.0 / p["f0"], p["dt"])
    hrick = swd.hilbert1d(rick)
    kz, _ = swd.kz_taper(swd.nfft(ntz), p["dz"])
    shots = []
    for is_ in range(p["ns"]):
        s_ix, s_iz, r_iz = swd.geometry(nx, nz, p["pml"], Lc, p["dx"], p["dz"], p["sx0"], p["shotdx"], p["shotdep"],
                                        p["offsetz"], is_)
        rn = r_iz.size
        rec_ix = np.where(np.arange(rn) < rn // 2, pmlc, ntx - pmlc - 1)
        shots.append((s_ix, s_iz, rec_ix, r_iz))
    if kind == "pp":
        models = [swd.Plan(grid, coefs, vp, rho)]
    else:   # background, then density point scatterers in it (FC:697-705)
        rho_sc = swd.scatterers(rho, vps, ntx, ntz, pmlc, swin or p["swin"])
        models = [swd.Plan(grid, coefs, vps, rho), swd.Plan(grid, coefs, vps, rho_sc)]
    mig = swd.Plan(grid, coefs, vps, rho)
    acc = np.zeros(nx * nz, np.float32)
    for b0 in range(0, len(shots), batch):
        S = swd.ShotSet(shots[b0:b0 + batch])
        up = [P.afor(S, rick, hrick, kz)[1] for P in models]
        d = up[0] if kind == "pp" else up[1] - up[0]                    # FC:757-792
        img, ils, _, _ = mig.rtm(S, rick, swd.adjoint_source(d, p["dt"]))   # RC:588-637
        for k in range(S.n):
            swd.image_laplacian(swd.image_normalize(img[k], ils[k]), vps, ntx, ntz, pmlc, p["dx"], p["dz"], acc)
        print(f"{kind}: shots {b0 + 1}..{b0 + S.n} migrated", flush=True)
    for P in models + [mig]:
        P.close()
    return swd.edge_zero(acc, nx, nz, Lc).reshape(nx, nz)


def compare(got, ref, label):
    """Correlation, least-squares amplitude ratio and relative L2 of two images."""
    assert np.isfinite(got).all()
    g64, r64 = got.astype(np.float64).ravel(), ref.astype(np.float64).ravel()
    cc = float(np.corrcoef(g64, r64)[0, 1])
    scale = float(g64 @ r64 / (g64 @ g64))
    rel = float(np.linalg.norm(g64 - r64) / np.linalg.norm(r64))
    print(f"shipped {label}: correlation {cc:.10f}, amplitude ratio {scale:.8f}, relative L2 {rel:.3e}, "
          f"max |difference| / max |image| {np.abs(g64 - r64).max() / np.abs(r64).max():.3e}", flush=True)
    return cc, scale, rel


# measured (the runs are deterministic):
#   pp:  correlation 1 - 1e-10, amplitude ratio 0.9999998, relative L2 7.9e-6
#   psf: correlation 0.999987, amplitude ratio 0.99996, relative L2 5.1e-3
BARS = {"pp": (0.999999, 1e-5, 2e-5), "psf": (0.9999, 1e-3, 1e-2)}


@pytest.mark.parametrize("kind", ["pp", "psf"])
def test_rtm_of_shipped_inputs_matches_shipped_image(kind, tmp_path):
    z, p = shipped()
    print(f"case shipped RTM image {kind}: {p['nx']}x{p['nz']}, Lc {p['Lc']}, nt {p['itmax']}, {p['ns']} shots",
          flush=True)
    got = migrate(z, p, kind, str(tmp_path), swin=31)
    ref = z["migration_" + kind]
    assert np.array_equal(got == 0, ref == 0)   # the same zeroed edges (RC:754-768)
    cc, scale, rel = compare(got, ref, kind)
    c_min, a_tol, r_max = BARS[kind]
    assert cc >= c_min
    assert abs(scale - 1.0) <= a_tol
    assert rel <= r_max
