"""Seeded workload generators (CPU)."""
import numpy as np

from swd import configs


def test_bandlimited_spectrum_and_scale():
    m = configs.bandlimited(300, 200, 10.0, 10.0, 10.0, 30.0, 1500.0, 3)
    assert m.shape == (300, 200) and m.dtype == np.float32
    assert abs(np.abs(m).max() - 1.0) < 1e-6
    F = np.abs(np.fft.rfft2(m.astype(np.float64))) ** 2
    k = np.hypot(np.fft.fftfreq(300, 10.0)[:, None], np.fft.rfftfreq(200, 10.0)[None, :])
    inside = (k >= 10.0 / 1500.0) & (k <= 30.0 / 1500.0)
    assert F[~inside].sum() < 1e-12 * F[inside].sum() + 1e-6   # float32 rounding only
    assert np.array_equal(m, configs.bandlimited(300, 200, 10.0, 10.0, 10.0, 30.0, 1500.0, 3))
    assert not np.array_equal(m, configs.bandlimited(300, 200, 10.0, 10.0, 10.0, 30.0, 1500.0, 4))


def test_perturbed_vp_scales_bulk_modulus():
    c = configs.c1(nt=10)
    g = c.grid
    m = 1e-3 * configs.bandlimited(g.nx, g.nz, g.dx, g.dz, 10.0, 30.0, 1500.0, 5)
    vp = configs.perturbed_vp(c, m)
    p = c.pmlc
    ratio = (vp[p:p + g.nx, p:p + g.nz].astype(np.float64) / c.vp[p:p + g.nx, p:p + g.nz]) ** 2 - 1.0
    assert np.max(np.abs(ratio - m)) < 3e-7
    frame = np.ones_like(vp, bool)
    frame[p:p + g.nx, p:p + g.nz] = False
    assert np.array_equal(vp[frame], c.vp[frame])


def test_c3_fields():
    c = configs.c3(nt=50)
    assert c.born_m.shape == (c.grid.nx, c.grid.nz)
    assert len(c.born_dirs) == 5
    assert c.born_probe.shape == (c.shots[0][2].size, 50)
