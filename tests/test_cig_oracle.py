"""Pins the oracle of the angle-domain common-image gathers (or_rtm_cig).

The reference's gather kernel (sum_image_GPU_kernel_imageCIG, RK:819-905) is
never launched (RK:1149-1157 is commented out) and its gathers are never
written (RC:784-786 is commented out), so no reference output exists.  The
restatement is pinned by what its definition fixes:
  - the binning of one cell (or_cig_cells) for Poynting vectors of known
    directions: half the opening angle in degrees, the rounding of the bin,
    the Gaussian weights exp(-ii^2/18) over five bins, the [2, thetan) range,
    the delt-step averaging and its start-up, and the skip of undefined angles;
  - the plain image of the same run is or_rtm_backward's, bit for bit;
  - physics: on a flat reflector under a surface source the gathers peak at the
    incidence angle atan(offset / depth).
The receiver Poynting vector is +p v of the time-reversed field (the reference
writes -p v, which with its receiver propagation gives 90 degrees minus the
angle; the flat-reflector test shows the difference).
"""
import numpy as np
import pytest

import _oracle as O
from helpers import rel_l2
from test_updown_oracle import reflector_problem

W = np.exp(-np.arange(3) ** 2 / 18.0)   # weights of bins 0, +-1, +-2 away from the centre


def directions(alpha_deg, beta_deg, ps=1.0, pr=1.0):
    """Fields whose source Poynting vector points at alpha and receiver one at
    beta (degrees from +z towards +x): (-ps vs) and (+pr vr)."""
    a, b = np.radians(alpha_deg), np.radians(beta_deg)
    return (ps, -np.sin(a) / ps, -np.cos(a) / ps, pr, np.sin(b) / pr, np.cos(b) / pr)


def run_cells(cases, thetan, delt=1, steps=None, nt=100, dtype=np.float64):
    n = len(steps[0]) if steps else len(cases)
    ringF = np.zeros((delt, n), dtype)
    ringG = np.zeros((delt, n), dtype)
    cig = np.zeros((n, thetan), dtype)
    for k, batch in enumerate(steps or [cases]):
        f = np.array(batch, np.float64).T
        O.cig_cells(*f, nt - 2 - k, nt, delt, thetan, ringF, ringG, cig, dtype)
    return cig


def expected_bins(theta, thetan, pp=1.0):
    out = np.zeros(thetan)
    c = int(np.ceil(theta - 0.5))
    for ii in range(-2, 3):
        b = c + ii
        if 2 <= b < thetan:
            out[b] += pp * W[abs(ii)]
    return out


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
def test_binning_of_known_angles(dtype):
    thetan = 40
    cases, want = [], []
    for alpha, beta, ps, pr in ((0, 0, 1, 1), (10, 50.6, 2.0, -3.0), (-30, 14.4, 0.5, 0.25), (80, 30.2, 1, 1),
                                (0, 77.4, 1, 1), (5, 8.6, 1, 1), (0, 180, 1, 1), (45, -45, 1.5, 1.5)):
        cases.append(directions(alpha, beta, ps, pr))
        theta = abs(alpha - beta) / 2.0
        want.append(expected_bins(theta, thetan, ps * pr))
    got = run_cells(cases, thetan, dtype=dtype)
    for k, (g, w) in enumerate(zip(got, want)):
        np.testing.assert_allclose(g, w, rtol=2e-6, atol=1e-6, err_msg=f"case {k}")
    # (angles stay off the bin edges, where the float and double roundings of the
    # angle may differ) theta 0: only the +2 bin is at or above 2; theta 90
    # (opposite vectors): bins 88 .. 92, all cut at 40
    assert got[0][2] == pytest.approx(W[2], rel=1e-6) and got[0].sum() == pytest.approx(W[2], rel=1e-6)
    assert not got[6].any()


def test_delt_averaging_and_startup():
    """delt = 2: the first step (it = nt-2) only fills the rings; later steps
    bin the angle of the summed products of the last two steps."""
    thetan = 90
    same = directions(0, 0)       # F = G = 1: opening angle 0
    opposite = directions(0, 180)  # F = -1, G = 1: opening angle 180
    got = run_cells(None, thetan, delt=2, steps=[[same], [opposite]])
    # step 1 not ready; step 2: tht = (1 - 1) / 2 = 0, theta = 45
    np.testing.assert_allclose(got[0], expected_bins(45.0, thetan), rtol=1e-6, atol=1e-7)
    got1 = run_cells(None, thetan, delt=2, steps=[[same]])
    assert not got1.any()
    got3 = run_cells(None, thetan, delt=2, steps=[[same], [opposite], [opposite]])
    # step 3 averages steps 2 and 3: theta 90, bins 88, 89 (90..92 >= thetan)
    want = expected_bins(45.0, thetan) + expected_bins(90.0, thetan)
    np.testing.assert_allclose(got3[0], want, rtol=1e-6, atol=1e-7)


def test_undefined_angle_adds_nothing():
    cases = [(0.0, 0.3, 0.2, 1.0, 0.1, 0.1), (1.0, 0.0, 0.0, 1.0, 0.5, 0.5)]   # zero source vector
    assert not run_cells(cases, 20, dtype=np.float32).any()


@pytest.fixture(scope="module")
def reflector_wide():
    return reflector_problem(nx=160, nz=100, nt=900, refl=50)


def test_plain_image_unchanged(reflector_wide):
    P, d = reflector_wide
    _, st, fs = P.forward(strips=True, final=True)
    a = P.rtm_cig(fs, st, d, thetan=10)
    b = P.rtm_backward(fs, st, d)
    for x, y in zip(a[:3], b):
        assert np.array_equal(x, y)


def test_flat_reflector_gathers_peak_at_incidence(reflector_wide):
    P, d = reflector_wide
    _, st, fs = P.forward(np.float64, strips=True, final=True)
    _, _, _, cig = P.rtm_cig(fs, st, d, thetan=90, dtype=np.float64)
    pm = P.pmlc
    C = cig.reshape(P.ntx, P.ntz, 90)[pm:-pm, pm:-pm]
    sx, refl, depth = 80, 50, 48 * 10.0
    peaks = []
    for h in (0, 10, 20, 30):
        g = np.abs(C[sx + h, refl - 3:refl + 3]).sum(0)
        theta = np.degrees(np.arctan2(h * 10.0, depth))
        peaks.append(int(np.argmax(g)))
        print(f"offset {h}: incidence {theta:.1f} deg, gather peak bin {peaks[-1]}")
    assert peaks[0] <= 4
    assert abs(peaks[1] - 11.8) <= 3 and abs(peaks[2] - 22.6) <= 5
    assert peaks[0] < peaks[1] < peaks[2] < peaks[3] < 45


def test_fp32_vs_fp64(reflector_wide):
    P, d = reflector_wide
    _, st, fs = P.forward(strips=True, final=True)
    _, st64, fs64 = P.forward(np.float64, strips=True, final=True)
    c32 = P.rtm_cig(fs, st, d, thetan=60)[3]
    c64 = P.rtm_cig(fs64, st64, d, thetan=60, dtype=np.float64)[3]
    e = rel_l2(c32, c64)
    es = rel_l2(c32.sum(1), c64.sum(1))
    print(f"cig fp32 vs fp64: {e:.2e}; summed over angles: {es:.2e}")
    # an angle that sits on a bin edge can round to either bin in the two
    # precisions, moving that step's contribution to the neighbouring bin
    assert e < 2e-3 and es < 2e-4
