"""CPU tests of the deconvolution's host side: parameter.txt parsing in the
kkdecon format, derived geometry and output taper against the oracle, argument
checks, and plan creation failing cleanly without a device."""
import ctypes as C

import numpy as np
import pytest

import _oracle as O
import swd

LINES = [("--x grid number: nx", "845"), ("--z grid number: nz", "702"), ("--PSF x number: psfwinx", "31"),
         ("--PSF z number: psfwinz", "31"), ("--Filter x number: filterwinx", "71"),
         ("--Filter z number: filterwinz", "71"), ("--Snap x number: snapwinx", "11"),
         ("--Snap z number: snapwinz", "11"), ("--FFT epsilon", "5.0e-4"), ("--true vp", "./input/acc_vp.bin"),
         ("--migration vp", "./input/ini_vp.bin"), ("--density", "./input/ini_rho.bin"),
         ("--image", "./input/migration_dupp.bin"), ("--psf image", "./input/migration_dupsf.bin"),
         ("--result", "./output/migration_dudecon.bin"), ("--inverse psf", "./output/inv_psf.bin"),
         ("--phase shift flag", "0"), ("--reg eps", "2.0e-2")]


def test_read_decon_parameters(tmp_path):
    f = tmp_path / "parameter.txt"
    f.write_text("".join(f"{a}\n  {b}\n" for a, b in LINES) + "\n\n# trailing notes\n--SVD epsilon mu1\n  1.0e-6\n")
    p, files = swd.decon_read_parameters(str(f))
    assert (p.nx, p.nz, p.psfwinx, p.psfwinz, p.filterwinx, p.filterwinz, p.snapwinx, p.snapwinz) == \
        (845, 702, 31, 31, 71, 71, 11, 11)
    assert p.epsilon == np.float32(5e-4)
    assert files["image"] == "./input/migration_dupp.bin" and files["psf"] == "./input/migration_dupsf.bin"
    assert files["decon"] == "./output/migration_dudecon.bin" and files["shift_flag"] == 0
    short = tmp_path / "short.txt"
    short.write_text("".join(f"{a}\n  {b}\n" for a, b in LINES[:5]))
    with pytest.raises(swd.SwdError) as e:
        swd.decon_read_parameters(str(short))
    assert e.value.code == -4


@pytest.mark.parametrize("args", [(845, 702, 31, 71, 11, 5e-4), (310, 279, 31, 71, 11, 5e-4),
                                  (400, 300, 21, 61, 9, 1e-3), (500, 500, 31, 97, 13, 5e-4)])
def test_geometry_and_taper_match_oracle(args):
    nx, nz, pw, fw, sw, eps = args
    p = swd.decon_params(nx, nz, pw, fw, sw, eps)
    g = swd.decon_geometry(p)
    og = O.decon_geometry(O.decon_params(nx, nz, pw, fw, sw, eps))
    assert (g["n"], g["outwinx"], g["stepwinx"], g["stepnumx"], g["stepnumz"], g["windows"], g["nxp"], g["nzp"]) == \
        (og.fullwx, og.outwinx, og.stepwinx, og.stepnumx, og.stepnumz, og.totalnum, og.nxp, og.nzp)
    assert g["cells"] == (og.nxwpsf - 6) * (og.nzwpsf - 6)
    np.testing.assert_array_equal(swd.decon_taper(p), O.decon_coefr(O.decon_params(nx, nz, pw, fw, sw, eps)))


@pytest.mark.parametrize("p,code", [
    (swd.DeconParams(845, 702, 30, 30, 71, 71, 11, 11, 5e-4), -6),   # even PSF window
    (swd.DeconParams(845, 702, 31, 31, 71, 61, 11, 11, 5e-4), -6),   # fullwx != fullwz
    (swd.DeconParams(845, 702, 31, 31, 21, 21, 11, 11, 5e-4), -6),   # filter window < output window
    (swd.DeconParams(150, 702, 31, 31, 71, 71, 11, 11, 5e-4), -6),   # fewer than 7 PSF cells along x
    (swd.DeconParams(845, 702, 31, 31, 111, 111, 11, 11, 5e-4), -6),  # window side 141 > 127
    (swd.DeconParams(845, 0, 31, 31, 71, 71, 11, 11, 5e-4), -1),
])
def test_geometry_rejects(p, code):
    with pytest.raises(swd.SwdError) as e:
        swd.decon_geometry(p)
    assert e.value.code == code


def test_plan_without_device():
    if swd.device_count() > 0:
        pytest.skip("a HIP device is visible; covered by the gpu tests")
    h = C.c_void_p()
    rc = swd.load().swd_decon_plan_create(C.byref(swd.decon_params(845, 702)), 0, C.byref(h))
    assert rc == -2 and "device" in swd.last_error()


def test_null_arguments_fail_cleanly():
    L = swd.load()
    p = swd.decon_params(845, 702)
    assert L.swd_decon(None, None, None, None, None) == -1
    assert L.swd_decon_dev(None, None, None, None, None) == -1
    assert L.swd_decon_plan_destroy(None) == -1
    assert L.swd_decon_get_stats(None, None) == -1
    assert L.swd_decon_geometry(C.byref(p), None) == -1
    assert L.swd_decon_taper(None, None) == -1
    assert L.swd_decon_read_parameters(None, None, None) == -1
    h = C.c_void_p()
    assert L.swd_decon_plan_create(None, 0, C.byref(h)) == -1
