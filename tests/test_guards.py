"""CPU tests of the library's range guards and of the Python harness's argument checks.

* swd_plan_create refuses grids whose padded field reaches 2^30 floats (the
  kernels form element offsets in 32 bits and address them as bytes).
* swd_cfl_vmax is the stability limit the reference prints (RC:110-118):
  checked against the staggered scheme's von Neumann bound computed from its
  coefficients, and against the oracle's own time stepping, which stays bounded
  just below the limit and blows up just above it.
* swd_check_model (run by swd_plan_set_model) rejects non-positive or
  non-finite models and velocities above the limit.
* The harness refuses wavelets and trace arrays of the wrong size before any
  call into the library.
"""
import ctypes as C

import numpy as np
import pytest

import _oracle as O
import swd


def _rc_exact(Lc):
    """cal_xishu's coefficients from the closed form (products of odd numbers), in double."""
    out = []
    for m in range(1, Lc + 1):
        om = 2 * m - 1
        num = den = 1.0
        for i in range(1, Lc + 1):
            if i != m:
                oi = 2 * i - 1
                num *= oi * oi
                den *= om * om - oi * oi
        out.append((-1) ** (m + 1) * num / (abs(den) * om))
    return np.array(out)


def test_plan_create_rejects_fields_beyond_32bit_offsets():
    L = swd.load()
    h = C.c_void_p()
    g = swd.Grid(33000, 33000, 20, 4, 10.0, 10.0, 1e-3, 10)
    assert L.swd_plan_create(C.byref(g), 0, C.byref(h)) == -6
    assert "2^30" in swd.last_error()
    g = swd.Grid(1 << 20, 1000, 20, 4, 10.0, 10.0, 1e-3, 10)
    assert L.swd_plan_create(C.byref(g), 0, C.byref(h)) == -6
    if swd.device_count() == 0:
        g = swd.Grid(32000, 32000, 20, 4, 10.0, 10.0, 1e-3, 10)   # just below: fails on the device only
        assert L.swd_plan_create(C.byref(g), 0, C.byref(h)) == -2


@pytest.mark.parametrize("Lc", [1, 2, 4, 8])
@pytest.mark.parametrize("dx,dz,dt", [(10.0, 10.0, 1e-3), (5.0, 5.0, 4e-4), (10.0, 4.0, 5e-4)])
def test_cfl_limit_is_the_von_neumann_bound(Lc, dx, dz, dt):
    g = swd.Grid(40, 40, 10, Lc, dx, dz, dt, 10)
    vmax = swd.cfl_vmax(g)
    dx, dz, dt = (float(np.float32(a)) for a in (dx, dz, dt))   # the grid holds floats
    # |sin(w dt/2)| = (v dt/2) |D(kx)/dx, D(kz)/dz|, D(k h) = 2 sum c_m sin((m-1/2) k h): the scheme is
    # stable while the right side stays <= 1 for every wavenumber
    c = _rc_exact(Lc)
    kh = np.linspace(0.0, np.pi, 20001)
    D = 2.0 * np.abs(np.sin(np.outer(kh, np.arange(1, Lc + 1) - 0.5)) @ c).max()
    bound = 2.0 / (dt * D * np.hypot(1.0 / dx, 1.0 / dz))
    assert abs(vmax - bound) <= 1e-7 * bound
    if dx == dz:   # the reference's printed form, dx / (sum|rc| dt sqrt 2)
        rc = swd.cal_xishu(Lc)
        ref = dx / (np.abs(rc.astype(np.float64)).sum() * dt * np.sqrt(2.0))
        assert abs(vmax - ref) <= 1e-9 * ref


def _homogeneous_peak(v, Lc=4, nt=600):
    nx = nz = 40
    pml = 10
    pmlc = pml + Lc
    ntx, ntz = nx + 2 * pmlc, nz + 2 * pmlc
    vp = np.full((ntx, ntz), v, np.float32)
    rho = np.full((ntx, ntz), 1000.0, np.float32)
    P = O.Problem(nx, nz, pml, Lc, 10.0, 10.0, 1e-3, nt, vp, rho, pmlc + 20, pmlc + 20,
                  np.array([pmlc + 25]), np.array([pmlc + 25]), f0=20.0)
    g, _, _ = P.forward(np.float64)
    return float(np.abs(g).max()), float(np.abs(g[0, -50:]).max())


def test_oracle_turns_unstable_past_the_limit():
    vmax = swd.cfl_vmax(swd.Grid(40, 40, 10, 4, 10.0, 10.0, 1e-3, 10))
    peak, tail = _homogeneous_peak(0.99 * vmax)
    assert np.isfinite(peak) and tail < peak   # bounded, decaying into the CPML
    peak, tail = _homogeneous_peak(1.05 * vmax)
    assert not np.isfinite(tail) or tail > 1e6 * _homogeneous_peak(0.99 * vmax)[0]


def test_check_model():
    g = swd.Grid(30, 20, 5, 4, 10.0, 10.0, 1e-3, 10)
    n = g.ntx * g.ntz
    vmax = swd.cfl_vmax(g)
    rho = np.full(n, 1000.0, np.float32)
    vp = np.full(n, 0.99 * vmax, np.float32)
    swd.check_model(g, vp, rho)
    bad = vp.copy()
    bad[g.ntz * 7 + 9] = 1.01 * vmax
    with pytest.raises(swd.SwdError) as e:
        swd.check_model(g, bad, rho)
    assert e.value.code == -6 and "ix=7 iz=9" in str(e.value) and "stability limit" in str(e.value)
    for v in (0.0, -5.0, np.nan, np.inf):
        bad = vp.copy()
        bad[3] = v
        with pytest.raises(swd.SwdError) as e:
            swd.check_model(g, bad, rho)
        assert e.value.code == -1
    bad = rho.copy()
    bad[5] = 0.0
    with pytest.raises(swd.SwdError):
        swd.check_model(g, vp, bad)
    # the plan's own coefficients set the limit: halved coefficients double it
    rc = swd.cal_xishu(4) * 0.5
    swd.check_model(g, np.full(n, 1.9 * vmax, np.float32), rho, rc=rc)
    with pytest.raises(ValueError):
        swd.check_model(g, vp[:-1], rho)


def _fake_plan(nt=50):
    """A Plan object without a device handle: the argument checks run before any library call."""
    P = swd.Plan.__new__(swd.Plan)
    P.L = swd.load()
    P.grid = swd.Grid(30, 20, 5, 4, 10.0, 10.0, 1e-3, nt)
    P.h = None
    return P


def test_harness_checks_sizes_before_calling_the_library():
    P = _fake_plan()
    nt = P.grid.nt
    shots = [(10, 10, np.array([12, 13]), np.array([15, 15]))]
    w = np.zeros(nt, np.float32)
    adj = np.zeros((2, nt), np.float32)
    ntp = P.grid.nx * P.grid.nz
    calls = [
        lambda: P.forward(shots, w[:-1]),
        lambda: P.afor(shots, w, w[:-3], np.ones(128, np.float32)),
        lambda: P.rtm(shots, w, adj[:, :-1]),
        lambda: P.rtm(shots, w, adj[:1]),
        lambda: P.rtm(shots, np.zeros(nt + 1, np.float32), adj),
        lambda: P.rtm_cig(shots, w, adj[:, 1:]),
        lambda: P.rtm_updown(shots, w, w, adj, adj[:, :-2], np.ones(128, np.float32)),
        lambda: P.rtm_dev(shots, w[:10], 0, 0, 0, 0),
        lambda: P.born(shots, w, np.zeros(ntp - 1, np.float32)),
        lambda: P.born_adjoint(shots, w, adj.T.copy()),
        lambda: P.born_dev(shots, w[:-1], 0, 0),
    ]
    for k, f in enumerate(calls):
        with pytest.raises(ValueError):
            f()
        assert P.h is None, k
