"""End-to-end runs of the AFor and AVSPRTM drivers in the reference's directory
layout (parameter.txt, input/acc_vp.bin, output/, forward/), checked against the
oracle pipeline: AFor gathers and up/down gathers (FC:574-795), AVSPRTM per-shot
images and the summed, Laplacian-filtered, edge-zeroed migration (RC:552-793)."""
import os
import subprocess

import numpy as np
import pytest

import _oracle as O
from helpers import rel_l2

pytestmark = pytest.mark.gpu
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BIN = os.path.join(ROOT, "simple-wavenumber-deconvolution_amd", "bin")
TOL = 1e-5

P = dict(nx=80, nz=60, pml=10, Lc=4, dx=10.0, dz=10.0, itmax=300, dt=1e-3, f0=25.0, ns=3, sx0=100.0,
         shotdx=200.0, shotdep=0.0, recdix=400.0, offsetz=300.0, swin=11, window=3)


def write_params(path, scatter_flag, **over):
    keys = ["nx", "nz", "pml", "Lc", "dx", "dz", "itmax", "dt", "f0", "ns", "sx0", "shotdx", "shotdep",
            "recdix", "offsetz", "scatter_flag", "swin", "window"]
    vals = dict(P, scatter_flag=scatter_flag, **over)
    with open(path, "w") as f:
        for k in keys:
            f.write(f"--{k}\n  {vals[k]}\n")


def setup_dir(tmp):
    os.makedirs(os.path.join(tmp, "input"))
    os.makedirs(os.path.join(tmp, "output"))
    os.makedirs(os.path.join(tmp, "forward"))
    rng = np.random.default_rng(11)
    z = np.arange(P["nz"])[None, :]
    v = 1800.0 + 12.0 * z + 0 * np.arange(P["nx"])[:, None]
    v[:, 35:] += 400.0
    v = (v * (1 + 0.02 * rng.standard_normal(v.shape))).astype(np.float32)
    v.tofile(os.path.join(tmp, "input", "acc_vp.bin"))
    return v


def run(exe, cwd, *args):
    r = subprocess.run([os.path.join(BIN, exe), *args], cwd=cwd, capture_output=True, text=True, timeout=120)
    print(r.stdout[-600:], r.stderr[-600:], flush=True)
    assert r.returncode == 0, r.stderr


def oracle_case(vint, rho_pad=None, smooth=False):
    pmlc = P["pml"] + P["Lc"]
    nx, nz = P["nx"], P["nz"]
    ntx, ntz = nx + 2 * pmlc, nz + 2 * pmlc
    vp = O.model_in_pml(np.pad(vint, pmlc), ntz, ntx, pmlc)
    if smooth:
        vp = O.smooth(vp, ntz, ntx, pmlc, P["window"])
    rho = np.full(ntx * ntz, 1000.0, np.float32) if rho_pad is None else rho_pad
    probs = []
    for is_ in range(P["ns"]):
        s_ix, s_iz, r_iz = O.geometry(nx, nz, P["pml"], P["Lc"], P["dx"], P["dz"], P["sx0"], P["shotdx"],
                                      P["shotdep"], P["offsetz"], is_)
        rn = r_iz.size
        rx = np.where(np.arange(rn) < rn // 2, pmlc, ntx - pmlc - 1)
        probs.append(O.Problem(nx, nz, P["pml"], P["Lc"], P["dx"], P["dz"], P["dt"], P["itmax"], vp, rho,
                               s_ix, s_iz, rx, r_iz, f0=P["f0"]))
    return probs, vp, rho, (ntx, ntz, pmlc)


def read(tmp, name, shape=None):
    a = np.fromfile(os.path.join(tmp, "output", name), np.float32)
    return a.reshape(shape) if shape else a


def test_afor_and_avsprtm_observed(tmp_path):
    tmp = str(tmp_path)
    v = setup_dir(tmp)
    write_params(os.path.join(tmp, "parameter.txt"), 0)
    run("AFor", os.path.join(tmp, "forward"))
    probs, vp, rho, (ntx, ntz, pmlc) = oracle_case(v)
    nfftz = O.nfft(ntz)
    kz, _ = O.kz_taper(nfftz, P["dz"])
    h = O.hilbert1d(probs[0].wavelet)
    np.testing.assert_allclose(read(tmp, "rick_hilbert.bin"), h, atol=1e-6 * np.abs(h).max())
    for is_, pr in enumerate(probs):
        g, up, dn = pr.afor(h, kz)
        n = pr.nrec
        for tag, ref in (("obs", g), ("obsup", up), ("obsdown", dn)):
            got = read(tmp, f"{is_ + 1}source_seismogram_{tag}.bin", (n, P["itmax"]))
            e = rel_l2(got, ref)
            print(f"shot {is_ + 1} {tag}: relL2 {e:.2e}")
            assert e <= TOL
    # RTM with the smoothed model of the up-going data
    run("AVSPRTM", tmp)
    probs_m, vpm, _, _ = oracle_case(v, smooth=True)
    mig = np.zeros(P["nx"] * P["nz"], np.float32)
    for is_, pr in enumerate(probs_m):
        n = pr.nrec
        up = read(tmp, f"{is_ + 1}source_seismogram_obsup.bin", (n, P["itmax"]))
        adj = O.adjoint_source(up, P["dt"])
        d, st, fs = pr.forward(strips=True, final=True)
        img, ils, ilr = pr.rtm_backward(fs, st, adj)
        img = O.image_normalize(img, ils)
        got = read(tmp, f"image_pp{is_ + 1}.bin")
        ref = img.reshape(ntx, ntz)[pmlc:ntx - pmlc, pmlc:ntz - pmlc].ravel()
        e = rel_l2(got, ref)
        print(f"shot {is_ + 1} image_pp: relL2 {e:.2e}")
        assert e <= TOL
        O.image_laplacian(img, vpm, ntx, ntz, pmlc, P["dx"], P["dz"], acc=mig)
    mig = O.edge_zero(mig, P["nx"], P["nz"], P["Lc"])
    e = rel_l2(read(tmp, "migration_pp.bin"), mig)
    print(f"migration_pp: relL2 {e:.2e}")
    assert e <= TOL


def test_avsprtm_updown(tmp_path):
    """AVSPRTM --updown: the plain migration as without the option, and the
    up/down-separated migrations (du: down-going source with up-going receiver
    field, ud: the reverse) through the same normalisation, filter, sum and
    edge zeroing, against the oracle (or_rtm_updown) pipeline."""
    tmp = str(tmp_path)
    v = setup_dir(tmp)
    write_params(os.path.join(tmp, "parameter.txt"), 0)
    run("AFor", os.path.join(tmp, "forward"))
    run("AVSPRTM", tmp, "--updown")
    probs_m, vpm, _, (ntx, ntz, pmlc) = oracle_case(v, smooth=True)
    nfftz = max(128, O.nfft(P["nz"]))
    kz, _ = O.kz_taper(nfftz, P["dz"])
    hw = O.hilbert1d(probs_m[0].wavelet)
    migs = {k: np.zeros(P["nx"] * P["nz"], np.float32) for k in ("pp", "ppdu", "ppud")}
    for is_, pr in enumerate(probs_m):
        n = pr.nrec
        up = read(tmp, f"{is_ + 1}source_seismogram_obsup.bin", (n, P["itmax"]))
        adj = O.adjoint_source(up, P["dt"])
        hadj = np.stack([O.hilbert1d(r) for r in adj])
        _, st, fs = pr.forward(strips=True, final=True)
        _, hst, hfs = pr.forward(strips=True, final=True, wavelet=hw)
        img, ils, _ = pr.rtm_backward(fs, st, adj)
        du, ud, _, _ = pr.rtm_updown(hw, (fs, st), (hfs, hst), adj, hadj, kz)
        for name, a in (("pp", img), ("ppdu", du), ("ppud", ud)):
            a = O.image_normalize(a, ils)
            O.image_laplacian(a, vpm, ntx, ntz, pmlc, P["dx"], P["dz"], acc=migs[name])
    ref = {k: O.edge_zero(m, P["nx"], P["nz"], P["Lc"]) for k, m in migs.items()}
    got = {k: read(tmp, f"migration_{k}.bin") for k in ref}
    e = rel_l2(got["pp"], ref["pp"])
    print(f"migration_pp: relL2 {e:.2e}")
    assert e <= TOL
    scale = np.linalg.norm(ref["ppdu"])
    for k in ("ppdu", "ppud"):
        e = np.linalg.norm(got[k].astype(np.float64) - ref[k]) / scale
        print(f"migration_{k}: |got - oracle| / |du| {e:.2e}, relL2 {rel_l2(got[k], ref[k]):.2e}, "
              f"norm ratio to du {np.linalg.norm(ref[k]) / scale:.3f}")
        assert e <= TOL


def test_avsprtm_cig(tmp_path):
    """AVSPRTM --cig --thetan 30 --delt 2: the plain migration as without the
    option, and migration_ppcig.bin, the shot sum of the interior angle
    gathers [nx][nz][thetan], bit-identical to the oracle's (or_rtm_cig)."""
    tmp = str(tmp_path)
    v = setup_dir(tmp)
    write_params(os.path.join(tmp, "parameter.txt"), 0)
    run("AFor", os.path.join(tmp, "forward"))
    run("AVSPRTM", tmp, "--cig", "--thetan", "30", "--delt", "2")
    probs_m, vpm, _, (ntx, ntz, pmlc) = oracle_case(v, smooth=True)
    nx, nz = P["nx"], P["nz"]
    mig = np.zeros(nx * nz, np.float32)
    cig = np.zeros((nx, nz, 30), np.float32)
    for is_, pr in enumerate(probs_m):
        n = pr.nrec
        up = read(tmp, f"{is_ + 1}source_seismogram_obsup.bin", (n, P["itmax"]))
        adj = O.adjoint_source(up, P["dt"])
        _, st, fs = pr.forward(strips=True, final=True)
        img, ils, _, c = pr.rtm_cig(fs, st, adj, 30, 2)
        O.image_laplacian(O.image_normalize(img, ils), vpm, ntx, ntz, pmlc, P["dx"], P["dz"], acc=mig)
        cig += c.reshape(ntx, ntz, 30)[pmlc:ntx - pmlc, pmlc:ntz - pmlc]
    mig = O.edge_zero(mig, nx, nz, P["Lc"])
    e = rel_l2(read(tmp, "migration_pp.bin"), mig)
    print(f"migration_pp: relL2 {e:.2e}")
    assert e <= TOL
    got = read(tmp, "migration_ppcig.bin", (nx, nz, 30))
    print(f"migration_ppcig: {np.count_nonzero(got != cig)} values differ, max |cig| {np.abs(cig).max():.3e}")
    assert np.array_equal(got, cig) and np.abs(cig).max() > 0


def test_afor_and_avsprtm_psf(tmp_path):
    tmp = str(tmp_path)
    v = setup_dir(tmp)
    write_params(os.path.join(tmp, "parameter.txt"), 1)
    run("AFor", os.path.join(tmp, "forward"))
    probs_b, vpm, rho, (ntx, ntz, pmlc) = oracle_case(v, smooth=True)
    rho_s = O.scatterers(rho, vpm, ntx, ntz, pmlc, P["swin"])
    probs_s, _, _, _ = oracle_case(v, rho_pad=rho_s, smooth=True)
    nfftz = O.nfft(ntz)
    kz, _ = O.kz_taper(nfftz, P["dz"])
    h = O.hilbert1d(probs_b[0].wavelet)
    for is_ in range(P["ns"]):
        gb = probs_b[is_].afor(h, kz)
        gs = probs_s[is_].afor(h, kz)
        n = probs_b[is_].nrec
        for k, tag in enumerate(("scatter", "scatterup", "scatterdown")):
            got = read(tmp, f"{is_ + 1}source_seismogram_{tag}.bin", (n, P["itmax"]))
            ref = gs[k] - gb[k]
            e = rel_l2(got, ref)
            print(f"shot {is_ + 1} {tag}: relL2 {e:.2e}")
            assert e <= TOL   # a difference of two gathers, relative to the (small) scattered field
    # PSF migration of the up-going scattered data with the smoothed model and rho = 1000 (RC:441, 552-793)
    run("AVSPRTM", tmp)
    mig = np.zeros(P["nx"] * P["nz"], np.float32)
    for is_, pr in enumerate(probs_b):
        n = pr.nrec
        up = read(tmp, f"{is_ + 1}source_seismogram_scatterup.bin", (n, P["itmax"]))
        adj = O.adjoint_source(up, P["dt"])
        _, st, fs = pr.forward(strips=True, final=True)
        img, ils, _ = pr.rtm_backward(fs, st, adj)
        img = O.image_normalize(img, ils)
        got = read(tmp, f"image_psf{is_ + 1}.bin")
        ref = img.reshape(ntx, ntz)[pmlc:ntx - pmlc, pmlc:ntz - pmlc].ravel()
        e = rel_l2(got, ref)
        print(f"shot {is_ + 1} image_psf: relL2 {e:.2e}")
        assert e <= TOL
        O.image_laplacian(img, vpm, ntx, ntz, pmlc, P["dx"], P["dz"], acc=mig)
    mig = O.edge_zero(mig, P["nx"], P["nz"], P["Lc"])
    psf = read(tmp, "migration_psf.bin")
    e = rel_l2(psf, mig)
    print(f"migration_psf: relL2 {e:.2e}, max |psf| {np.abs(psf).max():.3e}")
    assert e <= TOL and np.abs(psf).max() > 0


def test_decon_driver(tmp_path):
    """Decon in the kkdecon_acoustic layout on the reference-data crop
    (tests/golden): the deconvolved image and the taper weights equal the
    oracle's bit for bit, and the model by-products follow DC:199-252."""
    from test_decon_oracle import crop_inputs

    img, psf = crop_inputs()
    nx, nz = img.shape
    d = str(tmp_path)
    os.makedirs(os.path.join(d, "input"))
    os.makedirs(os.path.join(d, "output"))
    rng = np.random.default_rng(21)
    v_true = (2000.0 + 500.0 * rng.random((nx, nz))).astype(np.float32)
    v_mig = (2000.0 + 500.0 * rng.random((nx, nz))).astype(np.float32)
    files = {"vp": "./input/acc_vp.bin", "vp0": "./input/ini_vp.bin", "rho": "./input/ini_rho.bin",
             "image": "./input/migration_dupp.bin", "psf": "./input/migration_dupsf.bin",
             "decon": "./output/migration_dudecon.bin", "inv": "./output/inv_psf.bin"}
    v_true.tofile(os.path.join(d, files["vp"]))
    v_mig.tofile(os.path.join(d, files["vp0"]))
    img.tofile(os.path.join(d, files["image"]))
    psf.tofile(os.path.join(d, files["psf"]))
    vals = [nx, nz, 31, 31, 71, 71, 11, 11, "5.0e-4", files["vp"], files["vp0"], files["rho"], files["image"],
            files["psf"], files["decon"], files["inv"], 0, "2.0e-2"]
    with open(os.path.join(d, "parameter.txt"), "w") as f:
        for k, v in enumerate(vals):
            f.write(f"--entry {k}\n  {v}\n")
    print(f"case Decon driver on the {nx} x {nz} reference-data crop", flush=True)
    run("Decon", d)
    out = lambda name: np.fromfile(os.path.join(d, "output", name), np.float32).reshape(nx, nz)
    odec, ocoef = O.decon(O.decon_params(nx, nz), img, psf)
    np.testing.assert_array_equal(out("migration_dudecon.bin"), odec)
    np.testing.assert_array_equal(out("migration_coef.bin"), ocoef)
    np.testing.assert_array_equal(out("migration_pp.bin"), img)
    np.testing.assert_array_equal(out("acc_ref.bin"), (v_true - v_mig) / v_true)
    ref = np.zeros_like(v_true)
    ref[:, 1:] = (v_true[:, 1:] - v_true[:, :-1]) / (v_true[:, 1:] + v_true[:, :-1])
    np.testing.assert_array_equal(out("true_vertical_ref.bin"), ref)
    assert os.path.exists(os.path.join(d, "output", "end.flag"))
    txt = open(os.path.join(d, "output", "migration_coef.txt"), "rb").read().split(b"\r\t\n")
    assert len(txt) == nz + 1 and len(txt[0].split()) == nx


def test_pipeline_afor_avsprtm_decon(tmp_path):
    """The reference's pipeline in its directory layout: AFor and AVSPRTM for
    the observed data (migration_pp.bin), again with scatter_flag 1 on a
    7-cell scatterer lattice (migration_psf.bin), then Decon in a
    kkdecon_acoustic-style directory on those two images; its result equals
    the oracle's deconvolution of the same images bit for bit."""
    tmp = str(tmp_path)
    setup_dir(tmp)
    print("case pipeline AFor -> AVSPRTM (pp, psf) -> Decon", flush=True)
    for flag in (0, 1):
        write_params(os.path.join(tmp, "parameter.txt"), flag, swin=7)
        run("AFor", os.path.join(tmp, "forward"))
        run("AVSPRTM", tmp)
    nx, nz = P["nx"], P["nz"]
    pp = read(tmp, "migration_pp.bin", (nx, nz))
    psf = read(tmp, "migration_psf.bin", (nx, nz))
    assert np.abs(pp).max() > 0 and np.abs(psf).max() > 0
    kk = os.path.join(tmp, "kkdecon")
    os.makedirs(os.path.join(kk, "output"))
    vals = [nx, nz, 7, 7, 15, 15, 3, 3, "5.0e-4", "../input/acc_vp.bin", "../output/ini_vp.bin",
            "../input/acc_vp.bin", "../output/migration_pp.bin", "../output/migration_psf.bin",
            "./output/migration_dudecon.bin", "./output/inv_psf.bin", 0, "2.0e-2"]
    with open(os.path.join(kk, "parameter.txt"), "w") as f:
        for k, v in enumerate(vals):
            f.write(f"--entry {k}\n  {v}\n")
    run("Decon", kk)
    dec = np.fromfile(os.path.join(kk, "output", "migration_dudecon.bin"), np.float32).reshape(nx, nz)
    odec, _ = O.decon(O.DeconParams(nx, nz, 7, 7, 15, 15, 3, 3, 5e-4), pp, psf)
    np.testing.assert_array_equal(dec, odec)
    assert np.isfinite(dec).all() and np.abs(dec).max() > 0
