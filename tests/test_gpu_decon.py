"""GPU parity of the wavenumber-domain PSF deconvolution against the oracle.

The HIP kernels follow the oracle's arithmetic operation for operation (fixed
fma chains, the same float / double roundings), so the deconvolved image and
the summed taper weights must be equal bit for bit, on:
  - the reference's own data: a 310 x 279 lattice-aligned crop of
    kkdecon_acoustic/input/migration_dupp.bin and migration_dupsf.bin
    (tests/golden/), parameters of kkdecon_acoustic/parameter.txt;
  - the full 845 x 702 synthetic d1 workload (the reference's grid size);
  - a spike-lattice PSF with per-cell shifts (the closed-form case the oracle
    is pinned on);
  - a window too large for the specialised launch (filterwin 95, n 125), which
    runs the generic configuration.
"""
import numpy as np
import pytest
import torch

import _oracle as O
import swd
from swd import configs
from test_decon_oracle import _spike_lattice, crop_inputs

pytestmark = pytest.mark.gpu


def _check(name, p_args, img, psf):
    nx, nz = img.shape
    print(f"case {name}: {nx} x {nz}, psf/filter/snap {p_args}", flush=True)
    plan = swd.Decon(swd.decon_params(nx, nz, *p_args), 0)
    try:
        dec, coef = plan.run(img, psf)
        st = plan.stats()
    finally:
        plan.close()
    odec, ocoef = O.decon(O.decon_params(nx, nz, *p_args), img, psf)
    diff = np.abs(dec.astype(np.float64) - odec).max() / max(np.abs(odec).max(), 1e-30)
    print(f"  {st['windows']} windows, {st['cells']} PSF cells, n {st['n']}: GPU {st['total_ms']:.3f} ms "
          f"(windows {st['windows_ms']:.3f} ms); max rel diff vs oracle {diff:.3e}", flush=True)
    np.testing.assert_array_equal(dec, odec)
    np.testing.assert_array_equal(coef, ocoef)
    return dec


def test_reference_data_crop_matches_oracle():
    img, psf = crop_inputs()
    _check("reference data crop", (31, 71, 11, 5e-4), img, psf)


def test_full_size_synthetic_matches_oracle():
    img, psf = configs.decon_case()
    _check("d1 845 x 702", (31, 71, 11, 5e-4), img, psf)


def test_spike_lattice_matches_oracle():
    rng = np.random.default_rng(3)
    nx, nz = 7 * 31 + 10, 8 * 31 + 3
    img = rng.standard_normal((nx, nz)).astype(np.float32)
    table = rng.integers(-5, 6, size=(16, 16, 2))
    psf = _spike_lattice(nx, nz, 31, lambda i, j: tuple(table[i, j]))
    _check("spike lattice", (31, 71, 11, 5e-4), img, psf)


def test_generic_launch_configuration_matches_oracle():
    img, psf = configs.decon_case(nx=420, nz=400, seed=12)
    _check("generic launch (n 125)", (31, 95, 13, 1e-3), img, psf)


def test_unequal_axes_match_oracle():
    """psf 31 x 21, filter 71 x 81, snap 11 x 9: every x/z offset, output block
    and window step differs by axis."""
    img, psf = configs.decon_case(nx=300, nz=280, seed=14, psfwin=31, psfwin_z=21)
    nx, nz = img.shape
    print("case unequal axes 300 x 280", flush=True)
    plan = swd.Decon(swd.DeconParams(nx, nz, 31, 21, 71, 81, 11, 9, 5e-4), 0)
    try:
        dec, coef = plan.run(img, psf)
    finally:
        plan.close()
    odec, ocoef = O.decon(O.DeconParams(nx, nz, 31, 21, 71, 81, 11, 9, 5e-4), img, psf)
    np.testing.assert_array_equal(dec, odec)
    np.testing.assert_array_equal(coef, ocoef)


def test_device_pointers_and_repeat_calls():
    img, psf = crop_inputs()
    nx, nz = img.shape
    plan = swd.Decon(swd.decon_params(nx, nz), 0)
    try:
        d0, c0 = plan.run(img, psf)
        ti = torch.from_numpy(img).cuda()
        tp = torch.from_numpy(psf).cuda()
        td = torch.empty_like(ti)
        tc = torch.empty_like(ti)
        torch.cuda.synchronize()
        plan.run_dev(ti.data_ptr(), tp.data_ptr(), td.data_ptr(), tc.data_ptr())
        assert np.array_equal(td.cpu().numpy(), d0) and np.array_equal(tc.cpu().numpy(), c0)
        d1, _ = plan.run(img, psf)
        assert np.array_equal(d1, d0)
    finally:
        plan.close()
