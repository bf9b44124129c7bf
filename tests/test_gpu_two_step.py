"""Two time steps per launch (temporal blocking) on the fused path, against
the one-step kernel and the CPU oracle.

Bar: the gathers, RTM images and illuminations equal the one-step kernel's
bit for bit (both make the same per-cell operations), and the fp32 oracle's
within 1e-5 relative L2 (the forward gathers bit for bit).  Step counts are
odd and even so that the one-step remainder of a two-step schedule is
exercised; sources sit in the absorbing frame, on the interior frame and in
the fast interior, receivers cross tile edges.
"""
import numpy as np
import pytest

import _oracle as O
import swd
from swd import configs
from helpers import oracle_problem, rel_l2, report

pytestmark = pytest.mark.gpu
TOL = 1e-5
# configurations of the two-step kernel: columns, rows, wavefronts per workgroup
SHAPES = [(32, 32, 4), (32, 32, -4), (64, 32, -8), (64, 32, 8), (16, 64, 4), (32, 64, -8)]   # negative: split launches


def plan_of(c, steps, shape=None):
    p = swd.Plan(c.grid, c.coefs, c.vp, c.rho, path=swd.SWD_PATH_FUSED)
    p.set_steps_per_launch(steps)
    if shape:
        p.set_tile_config(*shape)
    else:
        p.set_tile_config(16, 64, 0)
    return p


def run_both(c, shape, nshots):
    """Forward and RTM with one and with two steps per launch, and the oracle's."""
    Ps = [oracle_problem(c, k) for k in range(nshots)]
    outs = [P.forward(strips=True, final=True) for P in Ps]
    adj = np.concatenate([O.adjoint_source(o[0], c.grid.dt) for o in outs])
    res = {}
    for steps in (1, 2):
        plan = plan_of(c, steps, shape if steps == 2 else None)
        g = plan.forward(c.shots, c.wavelet)
        st = plan.stats()
        assert st["steps_per_launch"] == steps
        if shape and steps == 2:
            assert (st["tile_x"], st["tile_rows"], st["tile_group_waves"]) == shape
        r = plan.rtm(c.shots, c.wavelet, adj, want_gather=True)
        st = plan.stats()
        assert st["steps_per_launch"] == steps
        res[steps] = (g,) + tuple(r)
        plan.close()
    return Ps, outs, adj, res


def check(c, shape, nshots, label):
    Ps, outs, adj, res = run_both(c, shape, nshots)
    for a, b in zip(res[1], res[2]):
        assert np.array_equal(a, b), label
    g, img, ils, ilr, gs = res[2]
    row = 0
    for k, P in enumerate(Ps):
        d, st, fs = outs[k]
        n = d.shape[0]
        o = P.rtm_backward(fs, st, adj[row:row + n])
        assert np.array_equal(g[row:row + n], d), f"{label} shot {k} forward"
        assert np.array_equal(gs[row:row + n], d), f"{label} shot {k} RTM forward"
        for name, a, ref in (("image", img[k], o[0]), ("illum_src", ils[k], o[1]), ("illum_rcv", ilr[k], o[2])):
            e32, _ = report(f"{label} shot {k} {name}", a, ref)
            assert e32 <= TOL
        row += n


@pytest.mark.parametrize("shape", SHAPES)
@pytest.mark.parametrize("nt", [160, 163])
def test_two_step_small(shape, nt):
    """A grid smaller than most tiles' margins (frame tiles everywhere), a second
    shot with its source in the absorbing frame."""
    c = configs.small(nx=100, nz=70, pml=8, Lc=4, nt=nt, nshots=2)
    s_ix, s_iz, rx, rz = c.shots[1]
    c.shots[1] = (3, 5, rx, rz)
    print(f"case two-step small {shape} nt={nt}", flush=True)
    check(c, shape, 2, f"small {shape} nt {nt}")


@pytest.mark.parametrize("shape", SHAPES)
def test_two_step_interior_sources_and_receivers(shape):
    """A grid with fast (damping- and frame-free) tiles: one source deep in the
    interior, one on the interior frame, receiver columns at tile edges."""
    c = configs.small(nx=300, nz=260, pml=10, Lc=4, nt=161, nshots=2)
    pm = c.pmlc
    rz = np.arange(pm + 2, pm + 250, 3)
    c.shots = [(pm + 150, pm + 130, np.full(rz.size, pm + 96), rz),
               (pm + 1, pm + 60, np.full(rz.size, pm + 127), rz)]
    print(f"case two-step interior {shape}", flush=True)
    check(c, shape, 2, f"interior {shape}")


def test_two_step_c2_grid():
    """The c2 grid (590 x 933 padded, 401 receivers) for 161 steps, two shots,
    one of them in the receivers' column."""
    c = configs.c2(nt=161, shot_ix=(300, 420))
    print("case two-step c2 grid nt=161", flush=True)
    check(c, (32, 32, 4), 2, "c2")


def test_two_step_tuner_and_groups():
    """steps = 0 (the tuner times one- and two-step configurations) and shot
    groups on two streams give the one-step bits."""
    c = configs.small(nx=180, nz=150, pml=10, nt=300, nshots=4, dup=True)
    print("case two-step tuner and groups", flush=True)
    outs = [oracle_problem(c, k).forward() for k in range(4)]
    adj = np.concatenate([O.adjoint_source(o[0], c.grid.dt) for o in outs])
    ref_plan = plan_of(c, 1)
    ref = ref_plan.rtm(c.shots, c.wavelet, adj, want_gather=True)
    ref_plan.close()
    plan = swd.Plan(c.grid, c.coefs, c.vp, c.rho, path=swd.SWD_PATH_FUSED)
    got = plan.rtm(c.shots, c.wavelet, adj, want_gather=True)
    print(f"tuned: {plan.stats()['steps_per_launch']} steps per launch, tile {plan.stats()['tile_x']} x "
          f"{plan.stats()['tile_rows']}", flush=True)
    for a, b in zip(got, ref):
        assert np.array_equal(a, b)
    plan.set_steps_per_launch(2)
    for gf, gb, streams in ((1, 1, 2), (3, 2, 1)):
        plan.set_groups(gf, gb)
        plan.set_streams(streams)
        got = plan.rtm(c.shots, c.wavelet, adj, want_gather=True)
        for a, b in zip(got, ref):
            assert np.array_equal(a, b), (gf, gb, streams)
    with pytest.raises(swd.SwdError):
        plan.set_steps_per_launch(3)
    plan.close()


@pytest.mark.parametrize("shape", SHAPES)
def test_two_step_frame_sources_and_receivers(shape):
    """The AVSPRTM driver test's geometry (80 x 60 interior, pml 10): three shots
    with their sources on the interior frame's top row, receivers on the left
    and right frame columns, a constant adjoint source; bit for bit against one
    step per launch.  (A register-bounded build of the 16 x 64 configuration
    that spilled gave wrong source illuminations here.)"""
    nx, nz, pml, nt = 80, 60, 10, 41
    c = configs.small(nx=nx, nz=nz, pml=pml, Lc=4, nt=nt, nshots=3)
    pm = c.pmlc
    rz = np.arange(pm, pm + nz, 2)
    rx = np.where(np.arange(rz.size) < rz.size // 2, pm, nx + pm - 1)
    c.shots = [(pm + 10 + 20 * k, pm, rx, rz) for k in range(3)]
    adj = np.full((3 * rz.size, nt), 1e-3, np.float32)
    print(f"case two-step frame geometry {shape}", flush=True)
    res = {}
    for steps in (1, 2):
        plan = plan_of(c, steps, shape if steps == 2 else None)
        res[steps] = plan.rtm(c.shots, c.wavelet, adj, want_gather=True)
        assert plan.stats()["steps_per_launch"] == steps
        plan.close()
    for name, a, b in zip(("image", "illum_src", "illum_rcv", "gather"), res[1], res[2]):
        assert np.isfinite(b).all(), name
        assert np.array_equal(a, b), name


@pytest.mark.parametrize("shape", SHAPES)
def test_two_step_afor(shape):
    """AFor (p and its Hilbert companion per shot, column capture for the kz
    split) two steps per launch: gathers and up/down gathers equal one step per
    launch bit for bit and the oracle's within 1e-5; odd nt."""
    c = configs.small(nx=150, nz=110, pml=10, Lc=4, nt=203, nshots=2, dup=True)
    print(f"case two-step afor {shape}", flush=True)
    nfftz = swd.nfft(c.grid.ntz)
    kz, _ = swd.kz_taper(nfftz, c.grid.dz)
    hw = swd.hilbert1d(c.wavelet)
    res = {}
    for steps in (1, 2):
        plan = plan_of(c, steps, shape if steps == 2 else None)
        res[steps] = plan.afor(c.shots, c.wavelet, hw, kz)
        assert plan.stats()["steps_per_launch"] == steps
        plan.close()
    for name, a, b in zip(("gather", "up", "down"), res[1], res[2]):
        assert np.array_equal(a, b), name
    row = 0
    for k in range(2):
        og, ou, od = oracle_problem(c, k).afor(hw, kz)
        n = og.shape[0]
        for name, a, ref in (("gather", res[2][0], og), ("up", res[2][1], ou), ("down", res[2][2], od)):
            e32, _ = report(f"two-step afor {shape} shot {k} {name}", a[row:row + n], ref)
            assert e32 <= TOL
        row += n


@pytest.mark.parametrize("shape", [(32, 32, 4), (64, 32, -8), (16, 64, 4)])
@pytest.mark.parametrize("nt", [161, 164])
def test_two_step_updown(shape, nt):
    """Up/down-separated RTM two steps per launch (dp/dt of both levels and the
    receiver pressure of the first level stored for the kz split after each
    level): du, ud, uu, dd and the plain image equal one step per launch bit for
    bit, with two shots in groups on two streams."""
    c = configs.small(nx=150, nz=110, pml=10, Lc=4, nt=nt, nshots=2)
    s_ix, s_iz, rx, rz = c.shots[1]
    c.shots[1] = (c.pmlc + 1, c.pmlc + 40, rx, rz)
    print(f"case two-step updown {shape} nt={nt}", flush=True)
    outs = [oracle_problem(c, k).forward()[0] for k in range(2)]
    adj = np.concatenate([O.adjoint_source(d, c.grid.dt) for d in outs])
    hadj = swd.hilbert_traces(adj)
    hw = swd.hilbert1d(c.wavelet)
    nfftz = swd.nfft(c.grid.ntz)
    kz, _ = swd.kz_taper(nfftz, c.grid.dz)
    res = {}
    for steps in (1, 2):
        plan = plan_of(c, steps, shape if steps == 2 else None)
        plan.set_groups(1, 1)
        res[steps] = plan.rtm_updown(c.shots, c.wavelet, hw, adj, hadj, kz, want_same_dir=True, want_image=True)
        assert plan.stats()["steps_per_launch"] == steps
        plan.close()
    for name in res[1]:
        assert np.isfinite(res[2][name]).all(), name
        assert np.array_equal(res[1][name], res[2][name]), name


@pytest.mark.parametrize("nt", [260, 261, 262, 263])
def test_two_step_born_forward(nt):
    """The fused Born pair two steps per launch (the background's two levels,
    then the scattered field's two in the same workgroup, the Born source of
    each level carried in registers): the gathers equal the one-step pairs'
    bit for bit at every nt mod 4, in one-shot groups on two streams, and the
    oracle's within 1e-5, with receivers sharing nodes."""
    c = configs.small(nshots=3, dup=True, nt=nt, Lc=4)
    print(f"case two-step born forward nt={nt}", flush=True)
    rng = np.random.default_rng(90 + nt)
    nint = c.grid.nx * c.grid.nz
    m = rng.standard_normal(3 * nint).astype(np.float32)
    res = {}
    for steps in (1, 2):
        plan = plan_of(c, steps)
        plan.set_groups(2, 0)
        plan.set_streams(2)
        res[steps] = plan.born(c.shots, c.wavelet, m)
        assert plan.stats()["steps_per_launch"] == steps
        plan.close()
    assert np.isfinite(res[2]).all()
    assert np.array_equal(res[1], res[2])
    row = 0
    for k in range(3):
        od = oracle_problem(c, k).born(m[k * nint:(k + 1) * nint])
        n = od.shape[0]
        e32, _ = report(f"two-step born nt={nt} shot {k}", res[2][row:row + n], od)
        assert e32 <= TOL
        row += n


@pytest.mark.parametrize("nsnap", [7, 9, 65, 87])
def test_two_step_born_adjoint_checkpoints(nsnap):
    """The Born adjoint's background passes on the fused forward (two steps per
    launch, snapshots taken from and restored into the slot that holds the
    state, the recompute's sums from both levels of a launch and from the
    batched kernels for a segment's leftover steps): the gradient equals the
    one-step fused passes' and the batched path's bit for bit, with segments of
    38, 29, 5 (one pair of launches and one leftover step) and 3 steps (no
    pair), and the oracle's within 1e-5."""
    c = configs.small(nshots=3, dup=True, nt=261, Lc=4)
    print(f"case two-step born adjoint checkpoints nsnap={nsnap}", flush=True)
    rng = np.random.default_rng(80 + nsnap)
    ntr = sum(sh[2].size for sh in c.shots)
    d = rng.standard_normal((ntr, c.grid.nt)).astype(np.float32)
    res = {}
    for steps in (1, 2):
        plan = plan_of(c, steps)
        res[steps] = plan.born_adjoint(c.shots, c.wavelet, d, nsnap=nsnap)
        plan.close()
    batched = swd.Plan(c.grid, c.coefs, c.vp, c.rho, device=0, path=swd.SWD_PATH_BATCHED)
    gb = batched.born_adjoint(c.shots, c.wavelet, d, nsnap=nsnap)
    batched.close()
    assert np.isfinite(res[2]).all()
    assert np.array_equal(res[1], res[2])
    assert np.array_equal(res[2], gb)
    og = np.zeros_like(gb, dtype=np.float64)
    row = 0
    for k in range(3):
        n = c.shots[k][2].size
        og += oracle_problem(c, k).born_adjoint(d[row:row + n]).astype(np.float64)
        row += n
    e32, _ = report(f"two-step born adjoint nsnap={nsnap}", res[2], og)
    assert e32 <= TOL


@pytest.mark.parametrize("delt", [1, 3])
@pytest.mark.parametrize("nt", [161, 162, 163, 164])
def test_two_step_angle_gathers(nt, delt):
    """Angle gathers with the backward two steps per launch: the paired
    kernel's middle-level variant stores both fields' first level and k_cig
    bins the two steps of each launch (delt 1: one k_cig2 launch, the first
    launch binning step nt-1 itself; delt 3: time-averaged over a ring).
    Gathers, image and illuminations equal one step per launch bit for bit,
    for every remainder of the step count modulo 4, two shots (one source on
    the interior frame) on two streams, and the oracle's."""
    c = configs.small(nx=150, nz=110, pml=10, Lc=4, nt=nt, nshots=2)
    s_ix, s_iz, rx, rz = c.shots[1]
    c.shots[1] = (c.pmlc + 1, c.pmlc + 40, rx, rz)
    thetan = 10 if delt == 1 else 30
    print(f"case two-step angle gathers nt={nt} delt={delt}", flush=True)
    outs = [oracle_problem(c, k).forward(strips=True, final=True) for k in range(2)]
    adj = np.concatenate([O.adjoint_source(o[0], c.grid.dt) for o in outs])
    res = {}
    for steps in (1, 2):
        plan = plan_of(c, steps)
        plan.set_groups(1, 1)
        plan.set_streams(2)
        res[steps] = plan.rtm_cig(c.shots, c.wavelet, adj, thetan, delt)
        assert plan.stats()["steps_per_launch"] == steps
        plan.close()
    for name, a, b in zip(("image", "illum_src", "illum_rcv", "cig"), res[1], res[2]):
        assert np.isfinite(b).all(), name
        assert np.array_equal(a, b), name
    assert np.abs(res[2][3]).max() > 0
    row = 0
    for k in range(2):
        d, st, fs = outs[k]
        n = d.shape[0]
        o = oracle_problem(c, k).rtm_cig(fs, st, adj[row:row + n], thetan, delt)
        assert np.array_equal(res[2][3][k], o[3]), f"shot {k} cig"
        assert np.array_equal(res[2][0][k], o[0]), f"shot {k} image"
        row += n
