"""Pins the oracle of the up/down-separated RTM images (or_rtm_updown).

The reference declares this imaging without running it (image_ppdu/ppud in
forward/headrtm.h:51,119, the unused wavefield_seperation_kernel FK:140-165),
so no reference output exists.  The restatement is pinned by what its
definition fixes:
  - completeness: with a taper of ones, du + ud + uu + dd is the plain RTM
    image of or_rtm_backward (the split of each analytic field is exact),
    over whole padded columns (nfftz >= ntz) or their interior rows;
  - the split point: a taper that keeps only one half of the kz axis empties
    the images of the other half exactly;
  - direction: a flat reflector under a surface source and surface receivers
    images in du (down-going source, up-going receiver field) at its depth,
    with the AFor convention whose direction the direct-wave test of the AFor
    oracle pins;
  - precision: the fp32 twin agrees with the fp64 one.
"""
import numpy as np
import pytest

import _oracle as O
from helpers import rel_l2
from swd import configs


def reflector_problem(nx=120, nz=100, nt=700, refl=40):
    """Homogeneous 2000 m/s background over a 3000 m/s half-space from row refl;
    surface source and receivers (interior row 2).  Returns the background-model
    oracle problem and the reflection data (true minus background gathers).
    The record is long enough to hold the whole reflection: a truncated one has
    a poor Hilbert transform, and its receiver field leaks into "down"."""
    vb = np.full((nx, nz), 2000.0, np.float32)
    vt = vb.copy()
    vt[:, refl:] = 3000.0
    rx = np.arange(4, nx - 4, 2)
    shots = [(nx // 2, 2, rx, np.full(rx.size, 2))]
    probs = []
    for v in (vb, vt):
        c = configs.make_case("reflector", nx, nz, 10, 4, 10.0, 10.0, 1e-3, nt, v, 25.0, shots)
        g = c.grid
        probs.append(O.Problem(g.nx, g.nz, g.pml, g.Lc, g.dx, g.dz, g.dt, g.nt, c.vp, c.rho, *c.shots[0],
                               wavelet=c.wavelet, coefs=c.coefs))
    Pb, Pt = probs
    d = (Pt.forward(np.float64)[0] - Pb.forward(np.float64)[0]).astype(np.float32)
    return Pb, d


def hilbert_rows(a):
    return np.stack([O.hilbert1d(r) for r in np.asarray(a, np.float32)])


def updown(P, adj, kz, dtype=np.float64):
    hw = O.hilbert1d(P.wavelet)
    _, st, fs = P.forward(dtype, strips=True, final=True)
    _, hst, hfs = P.forward(dtype, strips=True, final=True, wavelet=hw)
    imgs = P.rtm_updown(hw, (fs, st), (hfs, hst), adj, hilbert_rows(adj), kz, dtype)
    plain = P.rtm_backward(fs, st, adj, dtype)[0]
    return imgs, plain


def interior(P, a):
    return np.asarray(a).reshape(P.ntx, P.ntz)[P.pmlc:-P.pmlc, P.pmlc:-P.pmlc]


@pytest.fixture(scope="module")
def reflector():
    return reflector_problem()


def test_components_sum_to_plain_image(reflector):
    P, d = reflector
    nfftz = O.nfft(P.ntz)
    (du, ud, uu, dd), plain = updown(P, d, np.ones(nfftz, np.float32))
    e = rel_l2(du + ud + uu + dd, plain)
    print(f"fp64: |du+ud+uu+dd - image| / |image| = {e:.2e}")
    assert e < 1e-12
    (du, ud, uu, dd), plain = updown(P, d, np.ones(nfftz, np.float32), np.float32)
    e = rel_l2(du.astype(np.float64) + ud + uu + dd, plain)
    print(f"fp32: {e:.2e}")
    assert e < 1e-5


def test_half_axis_tapers_empty_the_other_half(reflector):
    """Bins 0..nfftz/2 are "up", the rest "down": a taper of ones on one half
    leaves only the same-half image, identical to its value under a taper of
    ones everywhere."""
    P, d = reflector
    nfftz = O.nfft(P.ntz)
    k = np.arange(nfftz)
    only_up = (k <= nfftz // 2).astype(np.float32)
    (fdu, fud, fuu, fdd), _ = updown(P, d, np.ones(nfftz, np.float32))
    (du, ud, uu, dd), _ = updown(P, d, only_up)
    assert not du.any() and not ud.any() and not dd.any()
    assert np.array_equal(uu, fuu) and np.abs(uu).max() > 0
    (du, ud, uu, dd), _ = updown(P, d, 1.0 - only_up)
    assert not du.any() and not ud.any() and not uu.any()
    assert np.array_equal(dd, fdd) and np.abs(dd).max() > 0


@pytest.mark.parametrize("taper", ["ones", "kz"])
def test_flat_reflector_images_in_down_up(reflector, taper):
    P, d = reflector
    nfftz = O.nfft(P.ntz)
    kz = np.ones(nfftz, np.float32) if taper == "ones" else O.kz_taper(nfftz, 10.0)[0]
    (du, ud, uu, dd), plain = updown(P, d, kz)
    rows = slice(32, 48)   # around the reflector (row 40)
    x = slice(20, 100)     # away from the lateral edges
    rms = {n: float(np.sqrt(np.mean(interior(P, a)[x, rows] ** 2))) for n, a in
           (("du", du), ("ud", ud), ("uu", uu), ("dd", dd), ("image", plain))}
    print(f"taper {taper}, reflector rows rms:", {k: f"{v:.3e}" for k, v in rms.items()})
    prof = np.abs(interior(P, du)[x, 10:90]).sum(0)
    assert abs(10 + int(np.argmax(prof)) - 39) <= 2   # the image peaks at the contrast
    assert rms["du"] > 100 * rms["ud"] and rms["du"] > 100 * rms["dd"] and rms["du"] > 50 * rms["uu"]
    assert abs(rms["du"] / rms["image"] - 1) < 0.01


@pytest.mark.parametrize("nfftz", [128, 256])
def test_interior_window(nfftz):
    """nz = 120 with a 14-cell frame: ntz = 148.  nfftz 256 transforms whole
    padded columns, nfftz 128 (< ntz, >= nz) the 120 interior rows.  Both are
    complete and image the reflector in du."""
    P, d = reflector_problem(nz=120)
    assert P.ntz == 148
    (du, ud, uu, dd), plain = updown(P, d, np.ones(nfftz, np.float32))
    e = rel_l2(du + ud + uu + dd, plain)
    rows, x = slice(32, 48), slice(20, 100)
    rms = {n: float(np.sqrt(np.mean(interior(P, a)[x, rows] ** 2))) for n, a in
           (("du", du), ("ud", ud), ("dd", dd), ("image", plain))}
    print(f"nfftz {nfftz}: completeness {e:.2e}, reflector rows rms {rms}")
    assert e < 1e-12
    assert rms["du"] > 100 * rms["ud"] and rms["du"] > 50 * rms["dd"]
    assert abs(rms["du"] / rms["image"] - 1) < 0.02


def test_fp32_vs_fp64_with_kz_taper():
    c = configs.small(nt=300)
    g = c.grid
    s_ix, s_iz, rx, rz = c.shots[0]
    P = O.Problem(g.nx, g.nz, g.pml, g.Lc, g.dx, g.dz, g.dt, g.nt, c.vp, c.rho, s_ix, s_iz, rx, rz,
                  wavelet=c.wavelet, coefs=c.coefs)
    d = P.forward()[0]
    adj = O.adjoint_source(d, g.dt)
    kz, _ = O.kz_taper(O.nfft(P.ntz), g.dz)
    i64, _ = updown(P, adj, kz)
    i32, _ = updown(P, adj, kz, np.float32)
    for name, a, b in zip(("du", "ud", "uu", "dd"), i32, i64):
        e = rel_l2(a, b)
        print(f"{name}: fp32 vs fp64 {e:.2e}")
        assert e < 1e-5
