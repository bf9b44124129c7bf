"""Pins the CPU oracle against what the reference's data and the mathematics fix.

The reference ships no tests; its only golden data for this path is the
migration model (kkdecon_acoustic/input/ini_vp.bin) its smoothing rule made
from input/acc_vp.bin (crops in tests/golden/).  Everything else is pinned by
closed forms (FD coefficients, Ricker, adjoint-source integration, Laplacian),
invariants (CPML profile, linearity, reciprocity, dot-product test, H(cos)=sin)
and the analytic 2D Green's function of a homogeneous medium.
"""
import json
import os

import numpy as np
import pytest
import scipy.special

import _oracle as O
from helpers import rel_l2

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


# ---------------------------------------------------------------- FD coefficients
@pytest.mark.parametrize("Lc,exact", [
    (1, [1.0]),
    (2, [9 / 8, -1 / 24]),
    (4, [1225 / 1024, -245 / 3072, 49 / 5120, -5 / 7168]),
    (5, [19845 / 16384, -735 / 8192, 567 / 40960, -405 / 229376, 35 / 294912]),
])
def test_cal_xishu_closed_form(Lc, exact):
    rc = O.cal_xishu(Lc).astype(np.float64)
    np.testing.assert_allclose(rc, exact, rtol=2e-7, atol=0)


@pytest.mark.parametrize("Lc", range(1, 9))
def test_cal_xishu_consistency(Lc):
    # a staggered first derivative is exact for linear functions: sum rc_i (2i+1) = 1,
    # and for cubics: sum rc_i (2i+1)^3 = 0 (Lc >= 2)
    rc = O.cal_xishu(Lc).astype(np.float64)
    k = 2 * np.arange(Lc) + 1
    assert abs(np.sum(rc * k) - 1) < 1e-6
    if Lc >= 2:
        assert abs(np.sum(rc * k ** 3)) < 1e-5


# ---------------------------------------------------------------- Ricker
def test_ricker_closed_form():
    nt, f0, dt = 1000, 25.0, 1e-3
    r = O.ricker(nt, f0, 1 / f0, dt).astype(np.float64)
    t = np.arange(nt) * dt - 1 / f0
    a = (np.pi * f0 * t) ** 2
    np.testing.assert_allclose(r, (1 - 2 * a) * np.exp(-a), rtol=0, atol=3e-6)
    assert r[40] == 1.0
    np.testing.assert_allclose(r[40 - 20:40], r[41:61][::-1], atol=3e-6)  # float phase rounding


# ---------------------------------------------------------------- CPML
def test_cpml_invariants():
    nx, nz, pml, Lc = 120, 80, 20, 4
    pmlc = pml + Lc
    ntx, ntz = nx + 2 * pmlc, nz + 2 * pmlc
    c = O.cpml(ntx, ntz, pmlc, 10.0, 10.0, 1e-3)
    for ax, n in (("x", ntx), ("z", ntz)):
        a, ah, b, bh = c["a_" + ax], c["a_" + ax + "_half"], c["b_" + ax], c["b_" + ax + "_half"]
        assert np.all(b > 0) and np.all(b <= 1) and np.all(bh > 0) and np.all(bh <= 1)
        assert np.all(a <= 0) and np.all(ah <= 0)
        # damping only inside the CPML: a = 0 on [pmlc, n-pmlc-1], non-zero outside
        assert np.all(a[pmlc:n - pmlc] == 0) and np.all(ah[pmlc:n - pmlc - 1] == 0)
        assert np.all(a[:pmlc] < 0) and np.all(a[n - pmlc:] < 0)
        assert np.all(b[pmlc + 1:n - pmlc - 1] == 1.0)
        # mirror symmetry of the profiles
        np.testing.assert_array_equal(b, b[::-1])
        np.testing.assert_allclose(bh[:-1], bh[:-1][::-1], rtol=2e-6)
        # |a| grows towards the outer edge (d grows quadratically there)
        assert np.all(np.diff(a[:pmlc + 1]) >= 0)


def test_cpml_closed_form_value():
    # d0 = -(N+1) Vpmax ln(Rc) / (2 L), L = pmlc*dz, N = 2, Rc = 1e-5, Vpmax = 5500,
    # alpha_max = 15 pi; outermost integer point: delta = L
    ntx, ntz, pmlc, dx, dt = 200, 150, 24, 10.0, 1e-3
    c = O.cpml(ntx, ntz, pmlc, dx, dx, dt)
    L = pmlc * dx
    d0 = -3 * 5500 * np.log(1e-5) / (2 * L)
    am = 15 * np.pi
    d = d0
    al = am * 0 + 0.1 * am
    b = np.exp(-(d + al) * dt)
    a = d / (d + al) * (b - 1)
    assert abs(c["b_x"][0] - b) < 2e-6 and abs(c["a_x"][0] - a) / abs(a) < 2e-6


# ---------------------------------------------------------------- smoothing (golden)
def _crop_case(name):
    meta = json.load(open(os.path.join(GOLD, "ini_vp_crops.json")))
    Cn = meta["crop"]
    x0, z0 = meta["crops"][name]
    acc = np.fromfile(os.path.join(GOLD, f"acc_vp_{name}.f32"), np.float32).reshape(Cn, Cn)
    ini = np.fromfile(os.path.join(GOLD, f"ini_vp_{name}.f32"), np.float32).reshape(Cn, Cn)
    nx, nz = meta["shape"]
    # comparable region: cells whose 11x11 window lies in the crop or beyond a true edge,
    # and (away from the top edge) outside the first 8 crop rows the crop run keeps
    w = meta["window"]
    lo_x = 0 if x0 == 0 else w
    hi_x = Cn if x0 + Cn == nx else Cn - w
    lo_z = 0 if z0 == 0 else max(w, 8)
    hi_z = Cn if z0 + Cn == nz else Cn - w
    return acc, ini, (slice(lo_x, hi_x), slice(lo_z, hi_z)), w


@pytest.mark.parametrize("name", ["tl", "mid", "br"])
def test_smoothing_reproduces_reference_migration_model(name):
    acc, ini, sl, w = _crop_case(name)
    pmlc = 21
    v = np.pad(acc, pmlc, mode="edge")
    ntx, ntz = v.shape
    s = O.smooth(v, ntz, ntx, pmlc, w).reshape(ntx, ntz)[pmlc:-pmlc, pmlc:-pmlc]
    np.testing.assert_array_equal(s[sl], ini[sl])   # bit-exact


def test_smoothing_wrong_window_fails():
    acc, ini, sl, w = _crop_case("mid")
    v = np.pad(acc, 21, mode="edge")
    s = O.smooth(v, v.shape[1], v.shape[0], 21, 4).reshape(v.shape)[21:-21, 21:-21]
    assert rel_l2(s[sl], ini[sl]) > 1e-5


def test_smoothing_constant_and_top_rows():
    v = np.full((60, 50), 2500.0, np.float32)
    v[:, 12] = 3000.0
    s = O.smooth(v, 50, 60, 10, 5).reshape(60, 50)
    np.testing.assert_array_equal(s[:, 10:18], v[:, 10:18])   # first 8 interior rows kept
    c = O.smooth(np.full((60, 50), 2500.0, np.float32), 50, 60, 10, 5)
    assert np.allclose(c, 2500.0, rtol=1e-7)


# ---------------------------------------------------------------- kz taper, Hilbert
@pytest.mark.parametrize("nfftz", [256, 1024, 2048])
def test_kz_taper(nfftz):
    c, layerz = O.kz_taper(nfftz, 10.0)
    assert layerz == nfftz // 15
    assert c[0] == 0.0 and c[layerz] == 1.0 and np.all(c[layerz:nfftz - layerz + 1] == 1.0)
    assert np.all((c >= 0) & (c <= 1))
    k = np.arange(layerz + 1)
    np.testing.assert_allclose(c[:layerz + 1], 0.5 - 0.5 * np.cos(np.pi * k / layerz), atol=1e-7)


def test_nfft_rule():
    assert [O.nfft(n) for n in (248, 256, 257, 590, 1129, 4000)] == [256, 256, 512, 1024, 2048, 4096]


def test_hilbert_of_cosine_is_sine():
    N = 512
    n = np.arange(N)
    for k0 in (3, 17, 100):
        x = np.cos(2 * np.pi * k0 * n / N).astype(np.float32)
        h = O.hilbert1d(x)
        np.testing.assert_allclose(h, np.sin(2 * np.pi * k0 * n / N), atol=2e-6)


def test_hilbert_twice_is_minus_identity():
    rng = np.random.default_rng(1)
    N = 256
    X = np.zeros(N, complex)
    X[1:N // 2] = rng.standard_normal(N // 2 - 1) + 1j * rng.standard_normal(N // 2 - 1)
    X[N // 2 + 1:] = np.conj(X[1:N // 2][::-1])
    x = np.fft.ifft(X).real.astype(np.float32)
    hh = O.hilbert1d(O.hilbert1d(x))
    np.testing.assert_allclose(hh, -x, atol=2e-6 * np.abs(x).max())


# ---------------------------------------------------------------- host post-processing
def test_adjoint_source_closed_form():
    nt, dt, c = 200, 1e-3, 0.5
    d = np.full((2, nt), c, np.float32)
    r = O.adjoint_source(d, dt)
    k = np.arange(nt) + 1.0
    np.testing.assert_allclose(r[0], -c * dt * dt * k * (k + 1) / 2, rtol=1e-5)


def test_image_laplacian_quadratic():
    ntx, ntz, pmlc, dx, dz = 40, 30, 5, 10.0, 5.0
    ix = np.arange(ntx)[:, None].astype(np.float64)
    iz = np.arange(ntz)[None, :].astype(np.float64)
    img = (3 * ix ** 2 + 2 * iz ** 2).astype(np.float32)
    vp = np.full((ntx, ntz), 2000.0, np.float32)
    out = O.image_laplacian(img, vp, ntx, ntz, pmlc, dx, dz)
    expect = -(6 / dx ** 2 + 4 / dz ** 2) * 2000.0 ** 2
    np.testing.assert_allclose(out, expect, rtol=1e-6)


def test_image_normalize_and_edge_zero():
    illum = np.array([0, 1, 2, 4], np.float32)
    img = np.array([1, 1, 1, 1], np.float32)
    out = O.image_normalize(img, illum)
    np.testing.assert_allclose(out, 1 / (illum + 1e-3 * 4), rtol=1e-7)
    nx, nz, Lc = 20, 16, 4
    m = O.edge_zero(np.ones(nx * nz, np.float32), nx, nz, Lc).reshape(nx, nz)
    keep = np.zeros((nx, nz), bool)
    keep[4:nx - Lc, 7:nz - Lc] = True
    np.testing.assert_array_equal(m == 1, keep)


def test_geometry_reference_parameters():
    # parameter.txt of the reference: nx 845, nz 702, pml 16, Lc 5, dx dz 10,
    # sx0 50, shotdx 40, shotdep 0, offsetz 5000: rows with |iz|*dz <= 5000 m are
    # iz = 0..500, so r_n = 2*501 (two boreholes sharing the rows)
    s_ix, s_iz, r_iz = O.geometry(845, 702, 16, 5, 10.0, 10.0, 50.0, 40.0, 0.0, 5000.0, 3)
    assert (s_ix, s_iz) == (21 + 5 + 4 * 3, 21)
    assert r_iz.size == 1002 and np.array_equal(r_iz[:501], r_iz[501:])
    assert r_iz[0] == 21 and r_iz[500] == 21 + 500
    s_ix, s_iz, r_iz = O.geometry(845, 702, 16, 5, 10.0, 10.0, 50.0, 40.0, 100.0, 300.0, 0)
    assert s_iz == 31 and r_iz.size == 2 * 41 and r_iz[0] == 21


def test_scatterer_lattice():
    ntx, ntz, pmlc, swin = 120, 100, 12, 31
    vp = np.full(ntx * ntz, 2000.0, np.float32)
    rho = np.full(ntx * ntz, 1000.0, np.float32)
    r = O.scatterers(rho, vp, ntx, ntz, pmlc, swin).reshape(ntx, ntz)
    ix, iz = np.nonzero(r != 1000.0)
    assert set(ix) == {27, 58, 89} and set(iz) == {27, 58}
    assert np.allclose(r[ix, iz], 1002.0)


# ---------------------------------------------------------------- propagation
def _homog(nx=160, nz=160, pml=20, Lc=4, nt=700, v=2000.0, f0=12.0, dx=10.0, dt=1e-3, rec=None):
    pmlc = pml + Lc
    ntx, ntz = nx + 2 * pmlc, nz + 2 * pmlc
    vp = np.full((ntx, ntz), v, np.float32)
    rho = np.full((ntx, ntz), 1000.0, np.float32)
    s = (pmlc + nx // 2, pmlc + nz // 2)
    rx, rz = rec if rec is not None else ([s[0] + 40, s[0], s[0] + 30], [s[1], s[1] - 50, s[1] + 30])
    return O.Problem(nx, nz, pml, Lc, dx, dx, dt, nt, vp, rho, s[0], s[1], rx, rz, f0=f0), s


def test_forward_matches_analytic_green_function():
    """Homogeneous 2D medium: p = G * d/dt(source rate), G the 2D Green's function.

    The scheme adds w[n] to p at the source node each step, i.e. a pressure rate
    w(t)/dt over one cell, so u_tt - c^2 lap u = dx*dz*w'(t)/dt * delta(x); with
    e^{+i w t} the outgoing solution is  (-i/4) H0^(2)(w r / c) / c^2 per unit source.
    """
    P, s = _homog()
    g, _, _ = P.forward(np.float64)
    dt, dx, c, nt = 1e-3, 10.0, 2000.0, P.nt
    w = P.wavelet.astype(np.float64)
    nfft = 8 * nt
    W = np.fft.rfft(w, nfft)
    om = 2 * np.pi * np.fft.rfftfreq(nfft, dt)
    for tr in range(3):
        r = dx * np.hypot(P.rec_ix[tr] - s[0], P.rec_iz[tr] - s[1])
        Gw = np.zeros_like(W)
        Gw[1:] = (-0.25j) * scipy.special.hankel2(0, om[1:] * r / c) / c ** 2
        # source rate: dx*dz/dt * w, time derivative -> i*om
        U = Gw * W * (1j * om) * dx * dx / dt
        u = np.fft.irfft(U, nfft)[:nt]
        # p after step n holds the source summed over k <= n, i.e. the rate integrated
        # up to t_n + dt/2: compare with the analytic field half a step later
        t = np.arange(nt) * dt
        u = np.interp(t + 0.5 * dt, t, u, left=0.0)
        cc = np.dot(u, g[tr]) / np.linalg.norm(u) / np.linalg.norm(g[tr])
        amp = np.linalg.norm(g[tr]) / np.linalg.norm(u)
        print(f"trace {tr} r={r:.0f} m: correlation {cc:.6f}, amplitude ratio {amp:.4f}")
        assert cc > 0.9999 and abs(amp - 1) < 0.005


def test_forward_fp32_vs_fp64_and_linearity():
    P, _ = _homog(nt=400)
    g32, _, _ = P.forward()
    g64, _, _ = P.forward(np.float64)
    assert rel_l2(g32, g64) < 1e-5
    w = P.wavelet.copy()
    P.wavelet = (2 * w).astype(np.float32)
    g2, _, _ = P.forward()
    g2_64, _, _ = P.forward(np.float64)
    # in fp64 scaling the source by 2 scales every field exactly; in fp32 the
    # stencil's precursor ahead of the wavefront passes through the subnormal
    # range, where the relative precision it loses grows with it (~5e-7)
    np.testing.assert_array_equal(g2_64, 2 * g64)
    assert rel_l2(g2, 2 * g32) < 2e-6
    # superposition of two sources
    P.wavelet = w
    shifted = np.roll(w, 37).astype(np.float32)
    P.wavelet = (w + shifted).astype(np.float32)
    gs, _, _ = P.forward(np.float64)
    P.wavelet = shifted
    gb, _, _ = P.forward(np.float64)
    assert rel_l2(gs, g64 + gb) < 1e-7   # the summed wavelet is rounded to float


def test_cpml_absorbs():
    P, s = _homog(nx=100, nz=100, nt=1500, f0=15.0)
    _, _, fs = P.forward(np.float64, final=True)
    p_end = np.abs(fs[2]).max()
    gpk, _, _ = P.forward(np.float64)
    assert p_end < 5e-3 * np.abs(gpk).max()


def test_reciprocity_homogeneous():
    a, s = _homog(nt=500, rec=([0], [0]))
    ra = (s[0] + 35, s[1] - 20)
    Pa, _ = _homog(nt=500, rec=([ra[0]], [ra[1]]))
    ga, _, _ = Pa.forward(np.float64)
    Pb, _ = _homog(nt=500, rec=([s[0]], [s[1]]))
    Pb.shot.s_ix, Pb.shot.s_iz = ra
    gb, _, _ = Pb.forward(np.float64)
    assert rel_l2(gb, ga) < 1e-9


def test_rtm_fp32_vs_fp64():
    P, s = _homog(nx=80, nz=60, nt=400, rec=(np.full(60, 24), np.arange(24, 84)))
    d, st, fs = P.forward(strips=True, final=True)
    adj = O.adjoint_source(d, 1e-3)
    i32 = P.rtm_backward(fs, st, adj)
    d64, st64, fs64 = P.forward(np.float64, strips=True, final=True)
    i64 = P.rtm_backward(fs64, st64, adj, np.float64)
    for a, b in zip(i32, i64):
        assert rel_l2(a, b) < 1e-5
    assert np.all(i32[1] >= 0) and np.all(i32[2] >= 0)


def test_afor_up_down_separation_of_direct_waves():
    """Receivers in a borehole offset from the source: the direct wave of a shallow
    source is down-going there, that of a deep source up-going.  The separated
    gathers are unnormalised (cuFFT inverse without 1/nfftz, FK:768-769), so their
    magnitude is about nfftz times the pressure."""
    P, _ = _homog(nx=200, nz=200, nt=1300, f0=15.0, rec=(np.full(41, 24 + 150), np.arange(24 + 100, 24 + 141)))
    nfftz = O.nfft(P.ntz)
    kz, _ = O.kz_taper(nfftz, 10.0)
    h = O.hilbert1d(P.wavelet)
    win = slice(600, 1100)   # around the direct arrival (~865 steps)
    P.shot.s_ix, P.shot.s_iz = 24 + 20, 24 + 5
    g, up, dn = P.afor(h, kz, np.float64)
    e_up, e_dn = np.linalg.norm(up[:, win]), np.linalg.norm(dn[:, win])
    print(f"shallow source: up {e_up:.3e} down {e_dn:.3e}")
    assert e_dn > 50 * e_up
    assert 0.7 * nfftz < e_dn / np.linalg.norm(g[:, win]) < 1.1 * nfftz
    P.shot.s_iz = 24 + 195
    g, up, dn = P.afor(h, kz, np.float64)
    e_up, e_dn = np.linalg.norm(up[:, win]), np.linalg.norm(dn[:, win])
    print(f"deep source: up {e_up:.3e} down {e_dn:.3e}")
    assert e_up > 50 * e_dn


def test_born_dot_product_and_finite_difference():
    nx, nz, pml, Lc, nt = 50, 40, 10, 4, 300
    pmlc = pml + Lc
    ntx, ntz = nx + 2 * pmlc, nz + 2 * pmlc
    rng = np.random.default_rng(5)
    v = (2000 + 300 * rng.random((ntx, ntz))).astype(np.float32)
    rho = np.full((ntx, ntz), 1000.0, np.float32)
    rz = np.arange(pmlc, pmlc + nz, 2)
    P = O.Problem(nx, nz, pml, Lc, 10.0, 10.0, 1e-3, nt, v, rho, pmlc + 25, pmlc + 3,
                  np.full(rz.size, pmlc + 10), rz, f0=25.0)
    m = rng.standard_normal(nx * nz).astype(np.float32)
    d = rng.standard_normal((rz.size, nt)).astype(np.float32)
    Lm = P.born(m, np.float64)
    LTd = P.born_adjoint(d, np.float64)
    lhs, rhs = np.dot(Lm.ravel(), d.ravel()), np.dot(m.astype(np.float64), LTd)
    assert abs(lhs - rhs) <= 1e-12 * abs(lhs)
    # central finite difference of the forward operator along m (K -> K(1 + h m))
    h = 1e-3
    outs = []
    for sgn in (1, -1):
        vv = v.astype(np.float64).copy()
        vv[pmlc:pmlc + nx, pmlc:pmlc + nz] *= np.sqrt(1 + sgn * h * m.reshape(nx, nz))
        Q = O.Problem(nx, nz, pml, Lc, 10.0, 10.0, 1e-3, nt, vv.astype(np.float32), rho, pmlc + 25, pmlc + 3,
                      np.full(rz.size, pmlc + 10), rz, wavelet=P.wavelet, coefs=P.coefs)
        outs.append(Q.forward(np.float64)[0])
    fd = (outs[0] - outs[1]) / (2 * h)
    print(f"Born vs central FD relL2 {rel_l2(Lm, fd):.3e}")
    assert rel_l2(Lm, fd) < 2e-3
