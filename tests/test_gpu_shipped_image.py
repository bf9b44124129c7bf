"""The reference's RTM run on its own inputs, against the image it shipped.

The reference ships a parameter file (845 x 702 cells, Lc 5, 7000 steps, 210
shots, two borehole receiver lines), the true model it names and, as the
deconvolution's input, the migration image of that run
(kkdecon_acoustic/input/migration_pp.bin), kept in tests/golden/shipped_rtm.npz
by tests/golden/make_shipped_golden.py.  This test runs the same pipeline
through the library: the observed data of every shot on the true model with
the kz up/down separation (AFor with scatter_flag 0, FC:574-795), and the RTM
of the up-going data on the smoothed model, each shot's image normalised by
its source illumination, Laplacian-filtered and summed, edges zeroed
(RC:552-793).  It is the one check against output that the reference itself
produced.  The reference's code and compiler rounding are not this build's
(its image sums 210 shots of 7000 steps, with MPI reductions), so the bar is
agreement to single-precision accumulation, not bit parity: correlation,
amplitude and relative L2 of the whole image."""
import json
import os

import numpy as np
import pytest

import swd

pytestmark = pytest.mark.gpu
HERE = os.path.dirname(os.path.abspath(__file__))


def test_rtm_of_shipped_inputs_matches_shipped_image(tmp_path):
    z = np.load(os.path.join(HERE, "golden", "shipped_rtm.npz"))
    p = json.loads(str(z["meta"]))["params"]
    nx, nz, Lc, nt = p["nx"], p["nz"], p["Lc"], p["itmax"]
    print(f"case shipped RTM image: {nx}x{nz}, Lc {Lc}, nt {nt}, {p['ns']} shots", flush=True)
    grid = swd.Grid(nx, nz, p["pml"], Lc, p["dx"], p["dz"], p["dt"], nt)
    pmlc, ntx, ntz = grid.pmlc, grid.ntx, grid.ntz
    coefs = swd.coefs_for(grid)
    f = str(tmp_path / "acc_vp.bin")
    z["acc_vp"].astype(np.float32).tofile(f)
    vp, rho = swd.read_model(f, ntx, ntz, pmlc)                         # RC:371-440
    vps = swd.smooth_model(vp, ntz, ntx, pmlc, p["window"])             # RC:441-457
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
    obs = swd.Plan(grid, coefs, vp, rho)
    mig = swd.Plan(grid, coefs, vps, rho)
    acc = np.zeros(nx * nz, np.float32)
    batch = 30
    for b0 in range(0, len(shots), batch):
        S = swd.ShotSet(shots[b0:b0 + batch])
        _, up, _ = obs.afor(S, rick, hrick, kz)
        adj = swd.adjoint_source(up, p["dt"])                           # RC:588-637
        img, ils, _, _ = mig.rtm(S, rick, adj)
        for k in range(S.n):
            swd.image_laplacian(swd.image_normalize(img[k], ils[k]), vps, ntx, ntz, pmlc, p["dx"], p["dz"], acc)
        print(f"shots {b0 + 1}..{b0 + S.n} migrated", flush=True)
    obs.close()
    mig.close()
    got = swd.edge_zero(acc, nx, nz, Lc).reshape(nx, nz)
    ref = z["migration_pp"]
    assert np.isfinite(got).all()
    # the same zeroed edges (RC:754-768)
    assert np.array_equal(got == 0, ref == 0)
    cc = np.corrcoef(got.ravel(), ref.ravel())[0, 1]
    g64, r64 = got.astype(np.float64).ravel(), ref.astype(np.float64).ravel()
    scale = float(g64 @ r64 / (g64 @ g64))
    rel = float(np.linalg.norm(g64 - r64) / np.linalg.norm(r64))
    print(f"shipped image: correlation {cc:.8f}, amplitude ratio {scale:.8f}, relative L2 {rel:.3e}, "
          f"max |difference| / max |image| {np.abs(g64 - r64).max() / np.abs(r64).max():.3e}", flush=True)
    # measured: correlation 1 - 1e-10, amplitude 0.9999998, relative L2 7.9e-6 (the run is deterministic)
    assert cc >= 0.999999
    assert abs(scale - 1.0) <= 1e-5
    assert rel <= 2e-5
