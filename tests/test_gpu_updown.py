"""GPU parity of the up/down-separated RTM (swd_rtm_updown) against the oracle.

Bar: <= 1e-5 relative L2 against the fp32 oracle (or_rtm_updown_f32) for
du, ud, uu and dd; the fp64 error is printed.  The up/down images of the HIP
path pack both halves of a field into one inverse transform (the oracle runs
two), so agreement is to rounding, not bitwise.
"""
import numpy as np
import pytest

import _oracle as O
import swd
from swd import configs
from helpers import oracle_problem, rel_l2, report
from test_updown_oracle import reflector_problem, interior

pytestmark = pytest.mark.gpu
TOL = 1e-5
NAMES = ("du", "ud", "uu", "dd")


def plan_of(case, path=swd.SWD_PATH_FUSED):
    return swd.Plan(case.grid, case.coefs, case.vp, case.rho, path=path)


def oracle_updown(P, adj, hadj, kz, hw, dtype=np.float32):
    _, st, fs = P.forward(dtype, strips=True, final=True)
    _, hst, hfs = P.forward(dtype, strips=True, final=True, wavelet=hw)
    return P.rtm_updown(hw, (fs, st), (hfs, hst), adj, hadj, kz, dtype), P.rtm_backward(fs, st, adj, dtype)


@pytest.mark.parametrize("path", [swd.SWD_PATH_FUSED, swd.SWD_PATH_PLAIN])
def test_updown_small(path):
    c = configs.small(nt=300)
    print(f"case small updown path={path}", flush=True)
    P = oracle_problem(c)
    d = P.forward()[0]
    adj = O.adjoint_source(d, c.grid.dt)
    hadj = swd.hilbert_traces(adj)
    hw = swd.hilbert1d(c.wavelet)
    nfftz = swd.nfft(c.grid.ntz)
    kz, _ = swd.kz_taper(nfftz, c.grid.dz)
    g = plan_of(c, path).rtm_updown(c.shots, c.wavelet, hw, adj, hadj, kz, want_same_dir=True, want_image=True)
    (o, plain) = oracle_updown(P, adj, hadj, kz, hw)
    (o64, _) = oracle_updown(P, adj, hadj, kz, hw, np.float64)
    for k, name in enumerate(NAMES):
        e32, _ = report(f"updown {name}", g[name][0], o[k], o64[k])
        assert e32 <= TOL
    for k, name in enumerate(("image", "illum_src", "illum_rcv")):
        e32, _ = report(f"updown run {name}", g[name][0], plain[k])
        assert e32 <= TOL


def test_updown_reflector_physics_and_completeness():
    """Flat reflector, reflection data: du carries the image; with a taper of
    ones the four images add up to the plain image of the same run."""
    P, d = reflector_problem()
    print("case reflector updown", flush=True)
    g = swd.Grid(P.g.nx, P.g.nz, P.g.pml, P.g.Lc, P.g.dx, P.g.dz, P.g.dt, P.g.nt)
    plan = swd.Plan(g, P.coefs, P.vp, P.rho)
    shots = [(P.shot.s_ix, P.shot.s_iz, P.rec_ix, P.rec_iz)]
    hd = swd.hilbert_traces(d)
    hw = swd.hilbert1d(P.wavelet)
    nfftz = swd.nfft(P.ntz)
    out = plan.rtm_updown(shots, P.wavelet, hw, d, hd, np.ones(nfftz, np.float32), want_same_dir=True,
                          want_image=True)
    s = sum(out[k][0].astype(np.float64) for k in NAMES)
    e = rel_l2(s, out["image"][0])
    print(f"|du+ud+uu+dd - image| / |image| = {e:.2e}", flush=True)
    assert e <= TOL
    x, rows = slice(20, 100), slice(32, 48)
    rms = {k: float(np.sqrt(np.mean(interior(P, out[k][0])[x, rows] ** 2))) for k in NAMES}
    print("reflector rows rms:", rms, flush=True)
    assert rms["du"] > 100 * rms["ud"] and rms["du"] > 100 * rms["dd"] and rms["du"] > 50 * rms["uu"]
    # ud, uu and dd are 50-30000 times smaller than du here, sums of products
    # that cancel: their own-norm relative error measures fp32 cancellation
    # (the fp32 oracle shows the same against fp64), so their bar is taken
    # relative to du, the image they are split from.
    o, _ = oracle_updown(P, d, hd, np.ones(nfftz, np.float32), hw)
    o64, _ = oracle_updown(P, d, hd, np.ones(nfftz, np.float32), hw, np.float64)
    scale = np.linalg.norm(o[0])
    for k, name in enumerate(NAMES):
        report(f"reflector {name}", out[name][0], o[k], o64[k])
        err = np.linalg.norm(out[name][0].astype(np.float64) - o[k]) / scale
        print(f"reflector {name}: |gpu - fp32 oracle| / |du| = {err:.2e}", flush=True)
        assert err <= TOL
    assert rel_l2(out["du"][0], o[0]) <= TOL


def test_updown_batch_groups_streams_and_device_pointers():
    """Several shots with duplicate receivers: each shot against the oracle;
    any grouping and stream count, and the device-pointer entry point, give the
    same bits."""
    import torch
    c = configs.small(nx=110, nz=70, nt=220, nshots=3, dup=True)
    print("case updown batch of 3", flush=True)
    Ps = [oracle_problem(c, k) for k in range(3)]
    adj = np.concatenate([O.adjoint_source(P.forward()[0], c.grid.dt) for P in Ps])
    hadj = swd.hilbert_traces(adj)
    hw = swd.hilbert1d(c.wavelet)
    nfftz = swd.nfft(c.grid.ntz)
    kz, _ = swd.kz_taper(nfftz, c.grid.dz)
    plan = plan_of(c)
    ref = plan.rtm_updown(c.shots, c.wavelet, hw, adj, hadj, kz, want_same_dir=True)
    row = 0
    for k, P in enumerate(Ps):
        n = P.nrec
        o, _ = oracle_updown(P, adj[row:row + n], hadj[row:row + n], kz, hw)
        for q, name in enumerate(NAMES):
            e32, _ = report(f"batch shot {k} {name}", ref[name][k], o[q])
            assert e32 <= TOL
        row += n
    for gf, gb, streams in ((1, 1, 1), (2, 3, 1), (1, 2, 2), (0, 0, 3)):
        plan.set_groups(gf, gb)
        plan.set_streams(streams)
        got = plan.rtm_updown(c.shots, c.wavelet, hw, adj, hadj, kz, want_same_dir=True)
        for name in NAMES:
            assert np.array_equal(got[name], ref[name]), (gf, gb, streams, name)
    plan.set_groups(0, 0)
    plan.set_streams(2)
    dev = torch.device("cuda:0")
    ta, th = torch.from_numpy(adj).to(dev), torch.from_numpy(hadj).to(dev)
    outs = {k: torch.zeros((3, c.grid.ntx * c.grid.ntz), dtype=torch.float32, device=dev) for k in ("du", "ud")}
    plan.rtm_updown_dev(c.shots, c.wavelet, hw, ta.data_ptr(), th.data_ptr(), kz,
                        {k: v.data_ptr() for k, v in outs.items()})
    torch.cuda.synchronize()
    for k, v in outs.items():
        assert np.array_equal(v.cpu().numpy(), ref[k]), k


@pytest.mark.parametrize("Lc", [2, 8])
def test_updown_stencil_orders(Lc):
    c = configs.small(nx=90, nz=60, pml=8, Lc=Lc, nt=200)
    print(f"case updown Lc={Lc}", flush=True)
    P = oracle_problem(c)
    adj = O.adjoint_source(P.forward()[0], c.grid.dt)
    hadj = swd.hilbert_traces(adj)
    hw = swd.hilbert1d(c.wavelet)
    nfftz = swd.nfft(c.grid.ntz)
    kz, _ = swd.kz_taper(nfftz, c.grid.dz)
    g = plan_of(c).rtm_updown(c.shots, c.wavelet, hw, adj, hadj, kz, want_same_dir=True)
    o, _ = oracle_updown(P, adj, hadj, kz, hw)
    for k, name in enumerate(NAMES):
        e32, _ = report(f"Lc {Lc} updown {name}", g[name][0], o[k])
        assert e32 <= TOL


@pytest.mark.parametrize("nfftz", [1024, 512])
def test_updown_c2_grid(nfftz):
    """The c2 grid (590 x 933 padded) for 120 steps, one shot: nfftz 1024 over
    whole padded columns, 512 over the 502 interior rows ("nz padded to 512")."""
    c = configs.c2(nt=120, shot_ix=(420,))
    print(f"case c2 updown nt=120 nfftz={nfftz}", flush=True)
    P = oracle_problem(c)
    adj = O.adjoint_source(P.forward()[0], c.grid.dt)
    hadj = swd.hilbert_traces(adj)
    hw = swd.hilbert1d(c.wavelet)
    kz, _ = swd.kz_taper(nfftz, c.grid.dz)
    g = plan_of(c).rtm_updown(c.shots, c.wavelet, hw, adj, hadj, kz, want_same_dir=True)
    o, _ = oracle_updown(P, adj, hadj, kz, hw)
    for k, name in enumerate(NAMES):
        e32, _ = report(f"c2 nfftz {nfftz} updown {name}", g[name][0], o[k])
        assert e32 <= TOL


def test_updown_argument_errors():
    c = configs.small(nt=60)
    plan = plan_of(c)
    nfftz = swd.nfft(c.grid.ntz)
    ntr = c.shots[0][2].size
    adj = np.zeros((ntr, c.grid.nt), np.float32)
    for bad in (nfftz // 2, nfftz + 1, 8192):
        with pytest.raises(swd.SwdError, match="nfftz"):
            plan.rtm_updown(c.shots, c.wavelet, c.wavelet, adj, adj, np.ones(bad, np.float32))


def numpy_updown_step(src, hsrc, rcv, hrcv, kz, nfftz, ntx, ntz, pmlc):
    """One step of the up/down imaging condition in numpy fp64 (the definition
    of or_rtm_updown, written with numpy's FFT): returns du, ud, uu, dd."""
    n = nfftz
    up = np.arange(n) <= n // 2
    z0, z1 = (0, ntz) if n >= ntz else (pmlc, ntz - pmlc)   # transformed rows

    def split(a, b):
        x = np.zeros((ntx, n), complex)
        x[:, :z1 - z0] = (a.reshape(ntx, ntz) + 1j * b.reshape(ntx, ntz))[:, z0:z1]
        X = np.fft.fft(x, axis=1) * kz[None, :].astype(np.float64)
        u = np.zeros((ntx, ntz))
        d = np.zeros((ntx, ntz))
        u[:, z0:z1] = np.fft.ifft(np.where(up[None, :], X, 0), axis=1).real[:, :z1 - z0]
        d[:, z0:z1] = np.fft.ifft(np.where(up[None, :], 0, X), axis=1).real[:, :z1 - z0]
        return u, d

    su, sd = split(src, hsrc)
    ru, rd = split(rcv, hrcv)
    out = [sd * ru, su * rd, su * ru, sd * rd]
    mask = np.zeros((ntx, ntz), bool)
    mask[pmlc:ntx - pmlc, pmlc:ntz - pmlc] = True
    return [np.where(mask, o, 0.0).reshape(-1) for o in out]


@pytest.mark.parametrize("nz,pml,interior", [(80, 10, False), (160, 20, False), (380, 20, False), (560, 20, False),
                                             (1000, 20, False), (2100, 20, False), (120, 10, True),
                                             (500, 20, True), (1020, 20, True)])
def test_updown_image_step_vs_numpy(nz, pml, interior):
    """swd_updown_image_dev on random fields against numpy's fp64 FFT; two
    shots, accumulation onto non-zero images, a non-trivial taper."""
    import torch
    c = configs.small(nx=70, nz=nz, pml=pml, nt=10)
    g = c.grid
    ntx, ntz, pmlc = g.ntx, g.ntz, g.pmlc
    nfftz = swd.nfft(g.nz) if interior else swd.nfft(ntz)
    assert (nfftz < ntz) == interior
    print(f"case updown step ntz={ntz} nfftz={nfftz}", flush=True)
    kz, _ = swd.kz_taper(nfftz, g.dz)
    rng = np.random.default_rng(5)
    f = rng.standard_normal((4, 2, ntx * ntz)).astype(np.float32)
    inner = np.zeros((ntx, ntz), np.float32)
    inner[pmlc:ntx - pmlc, pmlc:ntz - pmlc] = 1
    f[0] *= inner.reshape(-1)
    f[1] *= inner.reshape(-1)
    img0 = rng.standard_normal((4, 2, ntx * ntz)).astype(np.float32) * inner.reshape(-1)
    plan = plan_of(c)
    dev = torch.device("cuda:0")
    tf = torch.from_numpy(f).to(dev)
    ti = torch.from_numpy(img0).to(dev)
    plan.updown_image_dev(2, tf[0].data_ptr(), tf[1].data_ptr(), tf[2].data_ptr(), tf[3].data_ptr(), kz,
                          {k: ti[q].data_ptr() for q, k in enumerate(NAMES)})
    plan.updown_image_dev(2, tf[0].data_ptr(), tf[1].data_ptr(), tf[2].data_ptr(), tf[3].data_ptr(), kz,
                          {k: ti[q].data_ptr() for q, k in enumerate(NAMES)})
    torch.cuda.synchronize()
    got = ti.cpu().numpy()
    for s in range(2):
        ref = numpy_updown_step(f[0, s], f[1, s], f[2, s], f[3, s], kz, nfftz, ntx, ntz, pmlc)
        for q, name in enumerate(NAMES):
            want = img0[q, s].astype(np.float64) + 2 * ref[q]
            e = rel_l2(got[q, s] - img0[q, s], want - img0[q, s])
            print(f"shot {s} {name}: relL2 of the increment vs numpy fp64 {e:.2e}", flush=True)
            assert e <= TOL
            assert not got[q, s][inner.reshape(-1) == 0].any()
