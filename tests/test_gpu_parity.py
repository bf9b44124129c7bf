"""GPU parity of the HIP path against the CPU oracle (fp32 host loops, fp64 twin).

Bar: <= 1e-5 relative L2 against the fp32 restatement of the
reference's per-cell equations; the fp64 error is printed for every result.
Tolerances are written in each test.
"""
import numpy as np
import pytest

import _oracle as O
import swd
from swd import configs
from helpers import oracle_problem, rel_l2, report

pytestmark = pytest.mark.gpu
TOL = 1e-5
PATHS = [swd.SWD_PATH_FUSED, swd.SWD_PATH_BATCHED, swd.SWD_PATH_PLAIN]


def plan_of(case, path=swd.SWD_PATH_FUSED, tile=0):
    return swd.Plan(case.grid, case.coefs, case.vp, case.rho, path=path, tile=tile)


@pytest.mark.parametrize("path", PATHS)
def test_forward_small(path):
    c = configs.small()
    print(f"case small forward path={path}", flush=True)
    g = plan_of(c, path).forward(c.shots, c.wavelet)
    P = oracle_problem(c)
    o32, _, _ = P.forward()
    o64, _, _ = P.forward(np.float64)
    e32, _ = report("forward small", g, o32, o64)
    assert e32 <= TOL


@pytest.mark.parametrize("Lc", [1, 2, 5, 8])
def test_forward_orders(Lc):
    c = configs.small(Lc=Lc, nt=250)
    print(f"case small forward Lc={Lc}", flush=True)
    g = plan_of(c).forward(c.shots, c.wavelet)
    o32, _, _ = oracle_problem(c).forward()
    e32, _ = report(f"forward Lc={Lc}", g, o32)
    assert e32 <= TOL


def test_forward_batch_with_duplicate_receivers():
    c = configs.small(nshots=3, dup=True, nt=300)
    print("case small batch of 3 shots, duplicate receivers", flush=True)
    g = plan_of(c).forward(c.shots, c.wavelet)
    row = 0
    for k in range(3):
        o32, _, _ = oracle_problem(c, k).forward()
        e32, _ = report(f"batch shot {k}", g[row:row + o32.shape[0]], o32)
        assert e32 <= TOL
        row += o32.shape[0]


def test_forward_c1():
    c = configs.c1()
    print("case c1 forward 200x300 nt=1000", flush=True)
    g = plan_of(c).forward(c.shots, c.wavelet)
    P = oracle_problem(c)
    o32, _, _ = P.forward()
    o64, _, _ = P.forward(np.float64)
    e32, e64 = report("forward c1", g, o32, o64)
    assert g.shape == (91, 1000)
    assert e32 <= TOL and e64 <= TOL


@pytest.mark.parametrize("path", PATHS)
def test_rtm_small(path):
    c = configs.small(nt=300)
    print(f"case small rtm path={path}", flush=True)
    P = oracle_problem(c)
    d, st, fs = P.forward(strips=True, final=True)
    adj = O.adjoint_source(d, c.grid.dt)
    img, ils, ilr, gs = plan_of(c, path).rtm(c.shots, c.wavelet, adj, want_gather=True)
    o = P.rtm_backward(fs, st, adj)
    d64, st64, fs64 = P.forward(np.float64, strips=True, final=True)
    o64 = P.rtm_backward(fs64, st64, adj, np.float64)
    assert rel_l2(gs, d) <= TOL
    for name, a, b, b64 in (("image", img[0], o[0], o64[0]), ("illum_src", ils[0], o[1], o64[1]),
                            ("illum_rcv", ilr[0], o[2], o64[2])):
        e32, _ = report(f"rtm {name}", a, b, b64)
        assert e32 <= TOL


@pytest.mark.parametrize("path", [swd.SWD_PATH_FUSED, swd.SWD_PATH_BATCHED])
def test_afor_small(path):
    c = configs.small(nt=300)
    print(f"case small afor path={path}", flush=True)
    nfftz = swd.nfft(c.grid.ntz)
    kz, _ = swd.kz_taper(nfftz, c.grid.dz)
    hw = swd.hilbert1d(c.wavelet)
    g, up, dn = plan_of(c, path).afor(c.shots, c.wavelet, hw, kz)
    P = oracle_problem(c)
    og, ou, od = P.afor(hw, kz)
    og64, ou64, od64 = P.afor(hw, kz, np.float64)
    for name, a, b, b64 in (("gather", g, og, og64), ("up", up, ou, ou64), ("down", dn, od, od64)):
        e32, _ = report(f"afor {name}", a, b, b64)
        assert e32 <= TOL


def test_forward_and_afor_groups_streams_device():
    """One-shot groups on two streams give the gathers of one group of the whole
    batch bit for bit, and the device-pointer AFor those of the host one."""
    import torch

    c = configs.small(nshots=3, dup=True, nt=300)
    print("case small batch of 3 shots: forward and afor in one-shot groups on two streams", flush=True)
    nfftz = swd.nfft(c.grid.ntz)
    kz, _ = swd.kz_taper(nfftz, c.grid.dz)
    hw = swd.hilbert1d(c.wavelet)
    one = plan_of(c)
    ref_g = one.forward(c.shots, c.wavelet)
    ref_a = one.afor(c.shots, c.wavelet, hw, kz)
    grp = plan_of(c)
    grp.set_groups(1, 0)
    grp.set_streams(2)
    assert np.array_equal(grp.forward(c.shots, c.wavelet), ref_g)
    for a, b in zip(grp.afor(c.shots, c.wavelet, hw, kz), ref_a):
        assert np.array_equal(a, b)
    S = swd.ShotSet(c.shots)
    d = torch.zeros((3, S.ntr, c.grid.nt), dtype=torch.float32, device="cuda:0")
    grp.afor_dev(S, c.wavelet, hw, kz, d[0].data_ptr(), d[1].data_ptr(), d[2].data_ptr())
    torch.cuda.synchronize()
    for a, b in zip(d.cpu().numpy(), ref_a):
        assert np.array_equal(a, b)
    # and the oracle, shot by shot
    row = 0
    for k in range(3):
        og, ou, od = oracle_problem(c, k).afor(hw, kz)
        n = og.shape[0]
        for name, a, b in (("gather", ref_a[0], og), ("up", ref_a[1], ou), ("down", ref_a[2], od)):
            e32, _ = report(f"afor shot {k} {name}", a[row:row + n], b)
            assert e32 <= TOL
        row += n


@pytest.mark.parametrize("Lc", [2, 8])
def test_afor_orders_in_groups_on_three_streams(Lc):
    """AFor at other FD orders, its wavefields in groups of two on three
    streams, against the oracle shot by shot."""
    c = configs.small(nshots=2, dup=True, nt=240, Lc=Lc)
    print(f"case small afor Lc={Lc}, 2 shots, groups of 2 wavefields on 3 streams", flush=True)
    nfftz = swd.nfft(c.grid.ntz)
    kz, _ = swd.kz_taper(nfftz, c.grid.dz)
    hw = swd.hilbert1d(c.wavelet)
    plan = plan_of(c)
    plan.set_groups(2, 0)
    plan.set_streams(3)
    g, up, dn = plan.afor(c.shots, c.wavelet, hw, kz)
    row = 0
    for k in range(2):
        og, ou, od = oracle_problem(c, k).afor(hw, kz)
        n = og.shape[0]
        for name, a, b in (("gather", g, og), ("up", up, ou), ("down", dn, od)):
            e32, _ = report(f"afor Lc={Lc} shot {k} {name}", a[row:row + n], b)
            assert e32 <= TOL
        row += n


def test_born_and_adjoint_small():
    c = configs.small(nt=300)
    print("case small born", flush=True)
    rng = np.random.default_rng(3)
    m = rng.standard_normal(c.grid.nx * c.grid.nz).astype(np.float32)
    dd = rng.standard_normal((c.shots[0][2].size, c.grid.nt)).astype(np.float32)
    plan = plan_of(c)
    d = plan.born(c.shots, c.wavelet, m)
    g = plan.born_adjoint(c.shots, c.wavelet, dd)
    P = oracle_problem(c)
    od = P.born(m)
    og = P.born_adjoint(dd)
    e1, _ = report("born", d, od, P.born(m, np.float64))
    e2, _ = report("born adjoint", g, og, P.born_adjoint(dd, np.float64))
    assert e1 <= TOL and e2 <= TOL
    lhs = float(np.dot(d.ravel().astype(np.float64), dd.ravel()))
    rhs = float(np.dot(m.astype(np.float64), g.astype(np.float64)))
    print(f"dot test <Lm,d>={lhs:.9e} <m,L^T d>={rhs:.9e} rel {abs(lhs - rhs) / abs(lhs):.2e}", flush=True)
    assert abs(lhs - rhs) <= 1e-5 * abs(lhs)


@pytest.mark.parametrize("path", PATHS)
def test_rtm_batch_duplicates_strip_source(path):
    """Several shots, receivers with duplicate nodes, sources on the top frame row
    (shot depth 0, as the reference's parameter.txt), every path against the oracle."""
    c = configs.small(nx=150, nz=100, nt=260, nshots=3, dup=True)
    shots = []
    for k, (s_ix, s_iz, rx, rz) in enumerate(c.shots):
        shots.append((s_ix, c.pmlc, rx, rz))
    c.shots = shots
    print(f"case rtm batch dup path={path}", flush=True)
    Ps = [oracle_problem(c, k) for k in range(3)]
    outs = [P.forward(strips=True, final=True) for P in Ps]
    adj = np.concatenate([O.adjoint_source(o[0], c.grid.dt) for o in outs])
    img, ils, ilr, gs = plan_of(c, path).rtm(c.shots, c.wavelet, adj, want_gather=True)
    row = 0
    for k, P in enumerate(Ps):
        d, st, fs = outs[k]
        n = d.shape[0]
        o = P.rtm_backward(fs, st, adj[row:row + n])
        assert rel_l2(gs[row:row + n], d) <= TOL
        for name, a, b in (("image", img[k], o[0]), ("illum_src", ils[k], o[1]), ("illum_rcv", ilr[k], o[2])):
            e32, _ = report(f"rtm shot {k} {name}", a, b)
            assert e32 <= TOL
        row += n


def test_rtm_c2_grid_two_shots():
    """The c2 grid (590 x 933 padded, 401 receivers on 201 nodes) for 160 steps, 2 shots."""
    c = configs.c2(nt=160, shot_ix=(300, 520))
    print("case c2 grid rtm nt=160", flush=True)
    Ps = [oracle_problem(c, k) for k in range(2)]
    outs = [P.forward(strips=True, final=True) for P in Ps]
    adj = np.concatenate([O.adjoint_source(o[0], c.grid.dt) for o in outs])
    for path in PATHS:
        img, ils, ilr, gs = plan_of(c, path).rtm(c.shots, c.wavelet, adj, want_gather=True)
        row = 0
        for k, P in enumerate(Ps):
            d, st, fs = outs[k]
            n = d.shape[0]
            o = P.rtm_backward(fs, st, adj[row:row + n])
            e_g, _ = report(f"c2 path {path} shot {k} gather", gs[row:row + n], d)
            e_i, _ = report(f"c2 path {path} shot {k} image", img[k], o[0])
            assert e_g <= TOL and e_i <= TOL
            row += n


@pytest.mark.parametrize("Lc", [1, 3, 4, 8])
@pytest.mark.parametrize("tile", [16, 32, 64])
def test_fused_tiles_forward_and_rtm(Lc, tile):
    """Every fused tile width and halo depth (R = 1 for Lc <= 4, R = 2 above)
    against the oracle, with a second shot whose source lies in the absorbing
    frame outside the interior (its reconstruction keeps subtracting there)."""
    c = configs.small(nx=100, nz=70, pml=8, Lc=Lc, nt=220, nshots=2)
    s_ix, s_iz, rx, rz = c.shots[1]
    c.shots[1] = (3, 5, rx, rz)
    print(f"case fused tile={tile} Lc={Lc}", flush=True)
    Ps = [oracle_problem(c, k) for k in range(2)]
    outs = [P.forward(strips=True, final=True) for P in Ps]
    adj = np.concatenate([O.adjoint_source(o[0], c.grid.dt) for o in outs])
    plan = plan_of(c, tile=tile)
    g = plan.forward(c.shots, c.wavelet)
    img, ils, ilr, gs = plan.rtm(c.shots, c.wavelet, adj, want_gather=True)
    row = 0
    for k, P in enumerate(Ps):
        d, st, fs = outs[k]
        n = d.shape[0]
        o = P.rtm_backward(fs, st, adj[row:row + n])
        e_f, _ = report(f"tile {tile} Lc {Lc} shot {k} forward", g[row:row + n], d)
        assert e_f <= TOL and rel_l2(gs[row:row + n], d) <= TOL
        for name, a, b in (("image", img[k], o[0]), ("illum_src", ils[k], o[1]), ("illum_rcv", ilr[k], o[2])):
            e32, _ = report(f"tile {tile} Lc {Lc} shot {k} {name}", a, b)
            assert e32 <= TOL
        row += n


# Lc 4 tiles of the forward and paired RTM kernels: width x rows x wavefront bound
TILE_CONFIGS = [(w, r, b) for r in (32, 64, 128) for b in (0, 5) for w in (16, 32, 64) if not (r == 128 and w == 64)]


@pytest.mark.parametrize("w,r,b", TILE_CONFIGS)
def test_tile_configs_forward_and_rtm(w, r, b):
    """Every tile height and wavefront bound the tuner may pick, against the
    oracle: forward gathers and the RTM image and illuminations, a second shot
    with its source in the absorbing frame."""
    c = configs.small(nx=100, nz=70, pml=8, Lc=4, nt=160, nshots=2)
    s_ix, s_iz, rx, rz = c.shots[1]
    c.shots[1] = (3, 5, rx, rz)
    print(f"case tile {w} x {r} rows, bound {b}", flush=True)
    Ps = [oracle_problem(c, k) for k in range(2)]
    outs = [P.forward(strips=True, final=True) for P in Ps]
    adj = np.concatenate([O.adjoint_source(o[0], c.grid.dt) for o in outs])
    plan = plan_of(c)
    plan.set_tile_config(w, r, b)
    g = plan.forward(c.shots, c.wavelet)
    st = plan.stats()
    assert (st["tile_x"], st["tile_rows"], st["tile_waves"]) == (w, r, b)
    img, ils, ilr, gs = plan.rtm(c.shots, c.wavelet, adj, want_gather=True)
    row = 0
    for k, P in enumerate(Ps):
        d, st_, fs = outs[k]
        n = d.shape[0]
        o = P.rtm_backward(fs, st_, adj[row:row + n])
        assert np.array_equal(g[row:row + n], d) and np.array_equal(gs[row:row + n], d)
        for name, a, ref in (("image", img[k], o[0]), ("illum_src", ils[k], o[1]), ("illum_rcv", ilr[k], o[2])):
            e32, _ = report(f"tile {w}x{r}/{b} shot {k} {name}", a, ref)
            assert e32 <= TOL
        row += n
    with pytest.raises(swd.SwdError):
        plan.set_tile_config(64, 128, 0)
    plan.close()


def test_tuner_picks_and_records_a_configuration(tmp_path, monkeypatch):
    """tile = 0: the tuner times every configuration on the loop's first steps,
    stores its pick in the cache file and gives the bits of a fixed tile."""
    cache = tmp_path / "tiles.txt"
    monkeypatch.setenv("SWD_TUNE_CACHE", str(cache))
    c = configs.small(nx=110, nz=90, nt=260, nshots=2)   # a grid no other test tunes (the picks persist in-process)
    print("case tuner: 110 x 90, 2 shots, nt 260", flush=True)
    ref = plan_of(c, tile=16).forward(c.shots, c.wavelet)
    plan = plan_of(c, tile=0)
    got = plan.forward(c.shots, c.wavelet)
    st = plan.stats()
    print(f"tuned tile {st['tile_x']} x {st['tile_rows']} rows, bound {st['tile_waves']}", flush=True)
    assert np.array_equal(got, ref)
    assert (st["tile_x"], st["tile_rows"], st["tile_waves"]) != (0, 0, 0)
    lines = cache.read_text().splitlines()
    print(lines, flush=True)
    assert any(" fwd " in ln and "Lc 4" in ln for ln in lines)
    plan.close()


def test_fused_autotuned_tile_matches_fixed():
    """The tuned tile width (chosen on the first steps of the loop) gives the
    same bits as a fixed width."""
    c = configs.small(nt=120, nshots=2)
    print("case fused autotune", flush=True)
    a = plan_of(c, tile=0).forward(c.shots, c.wavelet)
    b = plan_of(c, tile=32).forward(c.shots, c.wavelet)
    assert np.array_equal(a, b)


def test_fused_rtm_shot_groups_match():
    """RTM in shot groups (forward and backward grouped differently, uneven
    last group), on one stream or with groups alternating over two, gives the
    same images as one group."""
    c = configs.small(nt=150, nshots=5, dup=True)
    print("case fused rtm groups", flush=True)
    outs = [oracle_problem(c, k).forward() for k in range(5)]
    adj = np.concatenate([O.adjoint_source(o[0], c.grid.dt) for o in outs])
    plan = plan_of(c)
    ref = plan.rtm(c.shots, c.wavelet, adj, want_gather=True)
    for gf, gb, streams in ((1, 1, 1), (2, 3, 1), (5, 2, 1), (1, 1, 2), (2, 3, 2), (0, 0, 2)):
        plan.set_groups(gf, gb)
        plan.set_streams(streams)
        got = plan.rtm(c.shots, c.wavelet, adj, want_gather=True)
        for a, b in zip(got, ref):
            assert np.array_equal(a, b), (gf, gb, streams)
