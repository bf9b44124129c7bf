"""Pins the CPU oracle of the PSF deconvolution (kkdecon_acoustic/kk_decon_interp_psf_minewindow.cpp).

The reference ships inputs (migration_dupp.bin, migration_dupsf.bin: crops in
tests/golden/) but no deconvolved output, and FFTW is not in this image.  The
oracle is pinned by:
  - the DFT definition: its half-spectrum transform and block inverse against
    numpy's FFT;
  - a second restatement of the reference's own per-window procedure (divide
    by the float sum, full complex 2D FFT with numpy, fft_max over all bins,
    tmpfloat, full inverse) on windows of the real-data crop;
  - delta PSFs, for which the deconvolution has a closed form: a lattice of
    shifted unit spikes shifts the image back, scaled by n^2 / (pp_max (1+eps));
    per-cell shifts pin the PSF lattice clamp;
  - invariants: the output taper overlap-adds to one, scaling the image or the
    PSF by a power of two changes nothing, and the edge replication equals
    numpy's edge padding.
"""
import json
import os

import numpy as np
import pytest

import _oracle as O

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def crop_inputs():
    meta = json.load(open(os.path.join(GOLD, "decon_crop.json")))
    nx, nz = meta["crop"]
    img = np.fromfile(os.path.join(GOLD, "decon_dupp_crop.f32"), np.float32).reshape(nx, nz)
    psf = np.fromfile(os.path.join(GOLD, "decon_dupsf_crop.f32"), np.float32).reshape(nx, nz)
    return img, psf


def test_geometry_of_reference_parameters():
    """kkdecon_acoustic/parameter.txt: 845 x 702, psf 31, filter 71, snap 11."""
    g = O.decon_geometry(O.decon_params(845, 702))
    got = {k: getattr(g, k) for k in O.GEOM_FIELDS}
    want = dict(halfpsfwx=15, halfltwx=35, nxwpsf=28, nzwpsf=23, fullwx=101, fullwz=101, midpx=50,
                outwinx=29, halfoutwx=14, stepwinx=18, stepnumx=54, stepnumz=46, totalnum=2484, nxp=972, nzp=828)
    assert {k: got[k] for k in want} == want


def test_geometry_rejects_what_the_reference_rejects():
    with pytest.raises(ValueError):
        O.decon_geometry(O.DeconParams(845, 702, 30, 31, 71, 71, 11, 11, 5e-4))   # even psf window
    with pytest.raises(ValueError):
        O.decon_geometry(O.DeconParams(845, 702, 31, 31, 21, 21, 11, 11, 5e-4))   # filter < outwin


def test_expand_is_edge_padding():
    a = np.random.default_rng(1).standard_normal((37, 23)).astype(np.float32)
    got = O.decon_expand(a, 54, 36)
    np.testing.assert_array_equal(got, np.pad(a, ((0, 17), (0, 13)), mode="edge"))


@pytest.mark.parametrize("n,mx,mz,offx,offz", [(101, 71, 71, 15, 15), (101, 31, 31, 0, 0), (64, 40, 30, 12, 17),
                                               (15, 15, 15, 0, 0)])
def test_spectrum_is_the_dft(n, mx, mz, offx, offz):
    x = np.random.default_rng(n + mx).standard_normal((mx, mz))
    buf = np.zeros((n, n))
    buf[offx:offx + mx, offz:offz + mz] = x
    ref = np.fft.fft2(buf)[:, :n // 2 + 1]
    got = O.decon_spectrum(x, offx, offz, n)
    assert np.abs(got - ref).max() <= 1e-13 * np.abs(ref).max()


@pytest.mark.parametrize("n,r0,nr,c0,nc", [(101, 21, 29, 21, 29), (64, 5, 20, 30, 34), (15, 0, 15, 0, 15)])
def test_inverse_is_the_real_part_of_the_idft(n, r0, nr, c0, nc):
    d = np.fft.fft2(np.random.default_rng(n).standard_normal((n, n)))   # Hermitian
    got = O.decon_inverse(np.ascontiguousarray(d[:, :n // 2 + 1]), r0, nr, c0, nc)
    ref = (np.fft.ifft2(d) * n * n).real[r0:r0 + nr, c0:c0 + nc]
    assert np.abs(got - ref).max() <= 1e-13 * np.abs(ref).max()


def _fsum(values):
    """Sequential float32 sum of magnitudes (the reference's float accumulator)."""
    s = np.float32(0)
    for v in np.abs(np.asarray(values, np.float32)).ravel():
        s = np.float32(s + v)
    return s


def reference_order_window(p, g, imgf, psff, step):
    """The reference's per-window procedure (DC:543-674) with numpy's FFT in place of FFTW."""
    n = g.fullwx
    ix, iz = g.stepwinx * (step // g.stepnumz), g.stepwinz * (step % g.stepnumz)
    snapm = np.zeros((n, n))
    x0, z0 = ix - g.halfltwx, iz - g.halfltwz
    win = np.zeros((p.filterwinx, p.filterwinz), np.float32)
    xs, zs = max(0, -x0), max(0, -z0)
    xe, ze = min(p.filterwinx, g.nxp - x0), min(p.filterwinz, g.nzp - z0)
    win[xs:xe, zs:ze] = imgf[x0 + xs:x0 + xe, z0 + zs:z0 + ze]
    snapm[g.midpx - g.halfltwx:g.midpx - g.halfltwx + p.filterwinx,
          g.midpz - g.halfltwz:g.midpz - g.halfltwz + p.filterwinz] = win
    pp = _fsum(win)
    if pp <= 1e-15:
        pp = np.float32(1e-10)
    M = np.fft.fft2(snapm / np.float64(pp))
    cx = min(max((ix - g.halfpsfwx) // p.psfwinx if ix >= g.halfpsfwx else 0, 3), g.nxwpsf - 4)
    cz = min(max((iz - g.halfpsfwz) // p.psfwinz if iz >= g.halfpsfwz else 0, 3), g.nzwpsf - 4)
    snapsf = np.zeros((n, n))
    patch = psff[cx * p.psfwinx:(cx + 1) * p.psfwinx, cz * p.psfwinz:(cz + 1) * p.psfwinz]
    snapsf[:p.psfwinx, :p.psfwinz] = patch
    S = np.fft.fft2(snapsf / np.float64(_fsum(patch)))
    q = S.real * S.real + S.imag * S.imag
    fft_max = np.float32(q.astype(np.float32).max()) * np.float32(p.epsilon)
    t = (q + np.float64(fft_max)).astype(np.float32).astype(np.float64)
    D = ((M.real * S.real + M.imag * S.imag) + 1j * (-M.real * S.imag + M.imag * S.real)) / t
    out = (np.fft.ifft2(D) * n * n).real
    r0, c0 = g.halfltwx - g.halfoutwx, g.halfltwz - g.halfoutwz
    return out[r0:r0 + g.outwinx, c0:c0 + g.outwinz]


def test_window_matches_reference_procedure_on_real_data():
    img, psf = crop_inputs()
    p = O.decon_params(*img.shape)
    g = O.decon_geometry(p)
    imgf, psff = O.decon_expand(img, g.nxp, g.nzp), O.decon_expand(psf, g.nxp, g.nzp)
    for step in (0, 5, 100, g.totalnum // 2, g.totalnum // 2 + 7, g.totalnum - 1):
        got = O.decon_window(p, img, psf, step)
        ref = reference_order_window(p, g, imgf, psff, step)
        err = np.abs(got - ref).max() / np.abs(ref).max()
        print(f"window {step}: max rel diff vs numpy-FFT reference procedure {err:.2e}")
        assert err <= 1e-12


def test_window_matches_reference_procedure_with_unequal_axes():
    """Windows whose sizes differ along x and z (psf 31 x 21, filter 71 x 81,
    snap 11 x 9: offsets, output blocks and steps all differ by axis)."""
    from swd import configs

    img, psf = configs.decon_case(nx=300, nz=280, seed=14, psfwin=31, psfwin_z=21)
    p = O.DeconParams(300, 280, 31, 21, 71, 81, 11, 9, 5e-4)
    g = O.decon_geometry(p)
    imgf, psff = O.decon_expand(img, g.nxp, g.nzp), O.decon_expand(psf, g.nxp, g.nzp)
    for step in (0, 40, g.totalnum // 2 + 3, g.totalnum - 2):
        got = O.decon_window(p, img, psf, step)
        ref = reference_order_window(p, g, imgf, psff, step)
        assert got.shape == (g.outwinx, g.outwinz)
        err = np.abs(got - ref).max() / np.abs(ref).max()
        print(f"unequal axes, window {step}: max rel diff {err:.2e}")
        assert err <= 1e-12
    _, coef = O.decon(p, img, psf)
    assert np.abs(coef.astype(np.float64) - 1.0).max() <= 2.5e-7


def test_output_taper_overlap_adds_to_one():
    img, psf = crop_inputs()
    _, coef = O.decon(O.decon_params(*img.shape), img, psf)
    assert np.abs(coef.astype(np.float64) - 1.0).max() <= 2.5e-7
    c = O.decon_coefr(O.decon_params(845, 702))
    assert c.shape == (29, 29) and c[0].max() == 0.0 and c[14, 14] == 1.0
    np.testing.assert_array_equal(c, c.T)


def test_power_of_two_scaling_changes_nothing():
    img, psf = crop_inputs()
    p = O.decon_params(*img.shape)
    d0, c0 = O.decon(p, img, psf)
    d1, c1 = O.decon(p, 2.0 * img, 4.0 * psf)
    np.testing.assert_array_equal(d0, d1)
    np.testing.assert_array_equal(c0, c1)


def _spike_lattice(nx, nz, w, shifts):
    """PSF image: a unit spike at (h + w i + a, h + w j + b) with (a, b) = shifts(i, j)."""
    h = w // 2
    psf = np.zeros((nx, nz), np.float32)
    for i in range((nx + w - 1) // w):
        for j in range((nz + w - 1) // w):
            a, b = shifts(i, j)
            x, z = h + w * i + a, h + w * j + b
            if 0 <= x < nx and 0 <= z < nz:
                psf[x, z] = 1.0
    return psf


@pytest.mark.parametrize("varying", [False, True])
def test_spike_psf_shifts_the_image_back(varying):
    """With a unit spike at offset (a, b) from each lattice node, the Wiener
    filter is conj(S) / (1 + eps) per window: the window's output is the image
    shifted by (a, b), times n^2 / (pp_max (1+eps)) (the inverse transform is
    unnormalised, as FFTW's backward transform in the reference).  With shifts that differ per
    cell the expected output also pins which cell each window uses."""
    rng = np.random.default_rng(3)
    nx, nz, w = 7 * 31 + 10, 8 * 31 + 3, 31
    img = rng.standard_normal((nx, nz)).astype(np.float32)
    table = rng.integers(-5, 6, size=(16, 16, 2)) if varying else np.tile([2, -3], (16, 16, 1))
    psf = _spike_lattice(nx, nz, w, lambda i, j: tuple(table[i, j]))
    p = O.decon_params(nx, nz)
    g = O.decon_geometry(p)
    dec, _ = O.decon(p, img, psf)
    imgf = O.decon_expand(img, g.nxp, g.nzp)
    coefr = O.decon_coefr(p).astype(np.float64)
    tden = np.float64(np.float32(1.0 + np.float64(np.float32(np.float32(1.0) * np.float32(p.epsilon)))))
    exp = np.zeros((g.nxp, g.nzp))
    for step in range(g.totalnum):
        ix, iz = g.stepwinx * (step // g.stepnumz), g.stepwinz * (step % g.stepnumz)
        x0, z0 = ix - g.halfltwx, iz - g.halfltwz
        win = imgf[max(0, x0):min(g.nxp, x0 + p.filterwinx), max(0, z0):min(g.nzp, z0 + p.filterwinz)]
        pp = _fsum(win)
        if pp <= 1e-15:
            pp = np.float32(1e-10)
        cx = min(max((ix - 15) // 31 if ix >= 15 else 0, 3), g.nxwpsf - 4)
        cz = min(max((iz - 15) // 31 if iz >= 15 else 0, 3), g.nzwpsf - 4)
        a, b = table[cx, cz]
        for i in range(g.outwinx):
            x = ix - g.halfoutwx + i
            if not 0 <= x < g.nxp:
                continue
            for j in range(g.outwinz):
                z = iz - g.halfoutwz + j
                if not 0 <= z < g.nzp:
                    continue
                xs, zs = x + a, z + b
                v = imgf[xs, zs] if (0 <= xs < g.nxp and 0 <= zs < g.nzp) else 0.0
                exp[x, z] += coefr[i, j] * v * (g.fullwx * g.fullwz) / (np.float64(pp) * tden)
    exp = exp[:nx, :nz]
    err = np.abs(dec - exp).max() / np.abs(exp).max()
    print(f"spike-lattice PSF (varying={varying}): max rel diff {err:.2e}")
    assert err <= 2e-6
