"""CPU tests of libswd: the C-ABI surface and the host-side setup routines.

The setup routines (parameter parsing, geometry, FD/CPML coefficients, model
smoothing, kz taper, wavelet Hilbert, adjoint-source integration, image
post-processing) must reproduce the oracle's restatement of the reference bit
for bit (the Hilbert transform, an FFT here and a direct DFT in the oracle, to
1e-6).  No GPU is touched: plan creation must fail cleanly without a device.
"""
import ctypes as C
import os
import re

import numpy as np
import pytest

import _oracle as O
import swd

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_library_exports_every_declared_symbol():
    L = swd.load()
    header = open(os.path.join(ROOT, "include", "swd.h")).read()
    names = sorted(set(re.findall(r"\bint\s+(swd_\w+)\s*\(", header)))
    assert len(names) >= 30
    missing = [n for n in names if not hasattr(L, n)]
    assert not missing, missing
    assert set(names) == set(swd.EXPORTS)


def test_errors_without_device(monkeypatch):
    if swd.device_count() > 0:
        pytest.skip("a HIP device is visible; covered by the gpu tests")
    g = swd.Grid(40, 30, 5, 4, 10.0, 10.0, 1e-3, 10)
    h = C.c_void_p()
    rc = swd.load().swd_plan_create(C.byref(g), 0, C.byref(h))
    assert rc == -2 and "device" in swd.last_error()


def test_argument_errors():
    L = swd.load()
    assert L.swd_cal_xishu(9, C.c_void_p(0)) == -1
    g = swd.Grid(4, 30, 5, 4, 10.0, 10.0, 1e-3, 10)   # nx < 2*Lc+1
    h = C.c_void_p()
    assert L.swd_plan_create(C.byref(g), 0, C.byref(h)) in (-1, -2)
    with pytest.raises(swd.SwdError):
        swd.kz_taper(100, 10.0)   # not a power of two


def test_null_plan_rejected_by_every_plan_entry_point():
    """Entry points of the later imaging modes refuse a null plan (no device needed)."""
    L = swd.load()
    null = C.c_void_p(0)
    out = swd.UpDownImages()
    s = C.c_void_p()
    assert L.swd_rtm_updown(null, 1, null, null, null, null, null, null, 128, C.byref(out)) < 0
    assert L.swd_rtm_updown_dev(null, 1, null, null, null, null, null, null, 128, C.byref(out)) < 0
    assert L.swd_updown_image_dev(null, 1, null, null, null, null, null, 128, C.byref(out)) < 0
    assert L.swd_rtm_cig(null, 1, null, null, null, 10, 1, null, null, null, null) < 0
    assert L.swd_rtm_cig_dev(null, 1, null, null, null, 10, 1, null, null, null, null) < 0
    assert L.swd_plan_get_stream(null, C.byref(s)) < 0
    assert L.swd_afor_dev(null, 1, null, null, null, null, 128, null, null, null) < 0
    assert "null plan" in swd.last_error()


PARAMS = [("--x grid number: nx", "845"), ("--z grid number: nz", "702"), ("--PML numbers: pml", "16"),
          ("--FD order: Lc", "5"), ("--x grid interval (m): dx", "10.0"), ("--z grid interval (m): dz", "10.0"),
          ("--recording time: itmax", "7000"), ("--time sample (s): dt", "1.0e-3"),
          ("--Riker wavelet dominant frequency (Hz): f0", "20.0"), ("--shot numbers: ns", "210"),
          ("--first shot (m): sx0", "50.0"), ("--shot interval X (m): shotdx", "40.0"),
          ("--the depth of shots (m): shotdep", "0.0"), ("--receivers: recdix", "5000.0"),
          ("--maximum offset", "5000.0 "), ("--scatter_flag", "1"), ("--scatter window", "51"),
          ("--velocity smooth window", "5")]


def test_read_parameters(tmp_path):
    p = tmp_path / "parameter.txt"
    p.write_text("".join(f"{a}\n  {b}\n" for a, b in PARAMS) + "\n\n###\n--velocity filename:\n  acc_vp.bin\n")
    d = swd.read_parameters(str(p))
    assert (d["nx"], d["nz"], d["pml"], d["Lc"], d["itmax"], d["ns"]) == (845, 702, 16, 5, 7000, 210)
    assert d["dt"] == np.float32(1e-3) and d["f0"] == 20.0 and d["offsetz"] == 5000.0
    assert (d["scatter_flag"], d["swin"], d["window"]) == (1, 51, 5)


def test_read_parameters_errors(tmp_path):
    with pytest.raises(swd.SwdError) as e:
        swd.read_parameters(str(tmp_path / "missing.txt"))
    assert e.value.code == -4
    p = tmp_path / "short.txt"
    p.write_text("".join(f"{a}\n  {b}\n" for a, b in PARAMS[:7]))
    with pytest.raises(swd.SwdError) as e:
        swd.read_parameters(str(p))
    assert e.value.code == -4


def test_read_model_and_short_file(tmp_path):
    nx, nz, pmlc = 12, 9, 5
    v = np.arange(nx * nz, dtype=np.float32).reshape(nx, nz) + 1500
    f = tmp_path / "acc_vp.bin"
    v.tofile(f)
    vp, rho = swd.read_model(str(f), nx + 2 * pmlc, nz + 2 * pmlc, pmlc)
    ref = O.model_in_pml(np.pad(v, pmlc), nz + 2 * pmlc, nx + 2 * pmlc, pmlc)
    np.testing.assert_array_equal(vp, ref)
    assert np.all(rho == 1000.0)
    v[:5].tofile(f)
    with pytest.raises(swd.SwdError) as e:
        swd.read_model(str(f), nx + 2 * pmlc, nz + 2 * pmlc, pmlc)
    assert e.value.code == -4


@pytest.mark.parametrize("Lc", range(1, 9))
def test_cal_xishu_matches_oracle(Lc):
    np.testing.assert_array_equal(swd.cal_xishu(Lc), O.cal_xishu(Lc))


@pytest.mark.parametrize("f0,dt,nt", [(25.0, 1e-3, 1000), (20.0, 8e-4, 4000), (20.0, 4e-4, 8000)])
def test_ricker_matches_oracle(f0, dt, nt):
    np.testing.assert_array_equal(swd.ricker(nt, f0, 1 / f0, dt), O.ricker(nt, f0, 1 / f0, dt))


@pytest.mark.parametrize("n,pmlc,h,dt", [(348, 24, 10.0, 1e-3), (933, 44, 10.0, 8e-4), (1818, 64, 5.0, 4e-4)])
def test_cpml_matches_oracle(n, pmlc, h, dt):
    a = swd.cpml(n, n - 100, pmlc, h, h, dt)
    b = O.cpml(n, n - 100, pmlc, h, h, dt)
    for k in a:
        np.testing.assert_array_equal(a[k], b[k], err_msg=k)


def test_smoothing_matches_oracle():
    rng = np.random.default_rng(0)
    v = (2000 + 500 * rng.random((90, 70))).astype(np.float32)
    for w in (3, 5, 7):
        np.testing.assert_array_equal(swd.smooth_model(v, 70, 90, 12, w), O.smooth(v, 70, 90, 12, w))


@pytest.mark.parametrize("nfftz,dz", [(256, 10.0), (1024, 10.0), (2048, 5.0)])
def test_kz_taper_matches_oracle(nfftz, dz):
    a, la = swd.kz_taper(nfftz, dz)
    b, lb = O.kz_taper(nfftz, dz)
    assert la == lb
    np.testing.assert_array_equal(a, b)


@pytest.mark.parametrize("n", [248, 590, 1129, 1000, 4000])
def test_nfft_matches_oracle(n):
    assert swd.nfft(n) == O.nfft(n)


@pytest.mark.parametrize("nt", [1000, 4000])
def test_hilbert_matches_oracle(nt):
    r = O.ricker(nt, 20.0, 0.05, 1e-3)
    a, b = swd.hilbert1d(r), O.hilbert1d(r)
    assert np.abs(a - b).max() <= 1e-6 * np.abs(b).max()


@pytest.mark.parametrize("args", [(845, 702, 16, 5, 10.0, 10.0, 50.0, 40.0, 0.0, 5000.0, 7),
                                  (300, 200, 20, 4, 10.0, 10.0, 0.0, 12.0, 20.0, 800.0, 3)])
def test_geometry_matches_oracle(args):
    a = swd.geometry(*args)
    b = O.geometry(*args)
    assert a[:2] == b[:2]
    np.testing.assert_array_equal(a[2], b[2])


def test_postprocessing_matches_oracle():
    rng = np.random.default_rng(2)
    ntx, ntz, pmlc = 60, 50, 8
    d = rng.standard_normal((7, 300)).astype(np.float32)
    np.testing.assert_array_equal(swd.adjoint_source(d, 8e-4), O.adjoint_source(d, 8e-4))
    img = rng.standard_normal(ntx * ntz).astype(np.float32)
    ill = rng.random(ntx * ntz).astype(np.float32)
    np.testing.assert_array_equal(swd.image_normalize(img, ill), O.image_normalize(img, ill))
    vp = (2000 + rng.random(ntx * ntz)).astype(np.float32)
    acc_a = np.ones((ntx - 2 * pmlc) * (ntz - 2 * pmlc), np.float32)
    acc_b = acc_a.copy()
    la = swd.image_laplacian(img, vp, ntx, ntz, pmlc, 10.0, 7.5, acc_a)
    lb = O.image_laplacian(img, vp, ntx, ntz, pmlc, 10.0, 7.5, acc_b)
    np.testing.assert_array_equal(la, lb)
    np.testing.assert_array_equal(acc_a, acc_b)
    m = rng.standard_normal(44 * 34).astype(np.float32)
    np.testing.assert_array_equal(swd.edge_zero(m, 44, 34, 4), O.edge_zero(m, 44, 34, 4))
    rho = np.full(ntx * ntz, 1000.0, np.float32)
    np.testing.assert_array_equal(swd.scatterers(rho, vp, ntx, ntz, pmlc, 11),
                                  O.scatterers(rho, vp, ntx, ntz, pmlc, 11))


def test_edge_zero_matches_the_shipped_image():
    """The reference's shipped migration image (kkdecon_acoustic/input/
    migration_pp.bin, 845 x 702, Lc 5) is zero exactly on the edges that
    swd_edge_zero clears (RC:754-768: ix <= 3, ix >= nx - Lc, iz < 7,
    iz >= nz - Lc), and nowhere inside them."""
    import json
    import os
    z = np.load(os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "shipped_rtm.npz"))
    p = json.loads(str(z["meta"]))["params"]
    nx, nz, Lc = p["nx"], p["nz"], p["Lc"]
    cleared = swd.edge_zero(np.ones(nx * nz, np.float32), nx, nz, Lc).reshape(nx, nz) == 0
    for kind in ("pp", "psf"):
        ref = z["migration_" + kind]
        assert ref.shape == (nx, nz) and ref.dtype == np.float32
        np.testing.assert_array_equal(ref == 0, cleared)


def test_shipped_psf_lattice_period_is_the_scatterer_window():
    """The shipped PSF image (migration_psf.bin) repeats with period 31 along x
    and z, the lattice swd_scatterers lays with swin 31 (FC:697-705); the
    shipped parameter file's value on that line is 51."""
    import json
    import os
    z = np.load(os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "shipped_rtm.npz"))
    p = json.loads(str(z["meta"]))["params"]
    a = z["migration_psf"].astype(np.float64)
    a -= a.mean()
    ac = np.fft.irfft2(np.abs(np.fft.rfft2(a)) ** 2, a.shape)
    lag_x = 2 + int(np.argmax(ac[2:80, 0]))
    lag_z = 2 + int(np.argmax(ac[0, 2:80]))
    assert (lag_x, lag_z) == (31, 31) and p["swin"] == 51
    pml, Lc = p["pml"], p["Lc"]
    pmlc = pml + Lc
    ntx, ntz = p["nx"] + 2 * pmlc, p["nz"] + 2 * pmlc
    rho = np.full(ntx * ntz, 1000.0, np.float32)
    vp = np.full(ntx * ntz, 2000.0, np.float32)
    r = swd.scatterers(rho, vp, ntx, ntz, pmlc, 31).reshape(ntx, ntz)
    xs, zs = np.nonzero(r != 1000.0)
    assert xs.size > 0
    assert set(np.diff(np.unique(xs))) == {31} and set(np.diff(np.unique(zs))) == {31}
