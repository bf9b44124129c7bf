"""GPU parity of the RTM angle-domain common-image gathers (swd_rtm_cig)
against the oracle (or_rtm_cig_f32).

The gathers follow the fp32 oracle operation for operation (the same fused
multiply-adds, sqrt and acos in double rounded to float, the same bin rule),
so the bar is bit-identity, with the plain image of the same run bit-identical
to the oracle's as in swd_rtm.
"""
import numpy as np
import pytest

import _oracle as O
import swd
from swd import configs
from helpers import oracle_problem, rel_l2
from test_updown_oracle import reflector_problem

pytestmark = pytest.mark.gpu


def plan_of(case):
    return swd.Plan(case.grid, case.coefs, case.vp, case.rho)


def compare(name, got, want):
    diff = int(np.count_nonzero(got != want))
    print(f"{name}: {diff} of {want.size} values differ, relL2 {rel_l2(got, want):.2e}", flush=True)
    assert np.array_equal(got, want), name


@pytest.mark.parametrize("thetan,delt", [(10, 1), (60, 3)])
def test_cig_small(thetan, delt):
    c = configs.small(nt=300)
    print(f"case small cig thetan={thetan} delt={delt}", flush=True)
    P = oracle_problem(c)
    d, st, fs = P.forward(strips=True, final=True)
    adj = O.adjoint_source(d, c.grid.dt)
    img, ils, ilr, cig = plan_of(c).rtm_cig(c.shots, c.wavelet, adj, thetan, delt)
    o = P.rtm_cig(fs, st, adj, thetan, delt)
    for name, a, b in (("image", img[0], o[0]), ("illum_src", ils[0], o[1]), ("illum_rcv", ilr[0], o[2]),
                       ("cig", cig[0], o[3])):
        compare(f"small {name}", a, b)
    assert np.abs(cig).max() > 0


@pytest.mark.parametrize("Lc", [1, 2, 5, 8])
def test_cig_stencil_orders(Lc):
    """The reconstruction velocities k_cig re-forms use the Lc-deep stencil."""
    c = configs.small(nx=90, nz=60, pml=8, Lc=Lc, nt=200)
    print(f"case cig Lc={Lc}", flush=True)
    P = oracle_problem(c)
    d, st, fs = P.forward(strips=True, final=True)
    adj = O.adjoint_source(d, c.grid.dt)
    _, _, _, cig = plan_of(c).rtm_cig(c.shots, c.wavelet, adj, 45, 1)
    compare(f"Lc {Lc} cig", cig[0], P.rtm_cig(fs, st, adj, 45, 1)[3])


def test_cig_batch_groups_streams_device():
    import torch
    c = configs.small(nx=110, nz=70, nt=220, nshots=3, dup=True)
    print("case cig batch of 3", flush=True)
    Ps = [oracle_problem(c, k) for k in range(3)]
    outs = [P.forward(strips=True, final=True) for P in Ps]
    adj = np.concatenate([O.adjoint_source(o[0], c.grid.dt) for o in outs])
    plan = plan_of(c)
    ref = plan.rtm_cig(c.shots, c.wavelet, adj, 30, 2)
    row = 0
    for k, P in enumerate(Ps):
        d, st, fs = outs[k]
        n = d.shape[0]
        o = P.rtm_cig(fs, st, adj[row:row + n], 30, 2)
        compare(f"batch shot {k} image", ref[0][k], o[0])
        compare(f"batch shot {k} cig", ref[3][k], o[3])
        row += n
    for gf, gb, streams in ((1, 1, 1), (2, 3, 1), (0, 0, 3)):
        plan.set_groups(gf, gb)
        plan.set_streams(streams)
        got = plan.rtm_cig(c.shots, c.wavelet, adj, 30, 2)
        for a, b in zip(got, ref):
            assert np.array_equal(a, b), (gf, gb, streams)
    plan.set_groups(0, 0)
    plan.set_streams(2)
    dev = torch.device("cuda:0")
    ntp = c.grid.ntx * c.grid.ntz
    ta = torch.from_numpy(adj).to(dev)
    im = torch.zeros((3, 3, ntp), dtype=torch.float32, device=dev)
    tc = torch.zeros((3, ntp, 30), dtype=torch.float32, device=dev)
    plan.rtm_cig_dev(c.shots, c.wavelet, ta.data_ptr(), 30, 2, im[0].data_ptr(), im[1].data_ptr(),
                     im[2].data_ptr(), tc.data_ptr())
    torch.cuda.synchronize()
    assert np.array_equal(tc.cpu().numpy(), ref[3]) and np.array_equal(im[0].cpu().numpy(), ref[0])


def test_cig_flat_reflector_peaks_at_incidence():
    P, d = reflector_problem(nx=160, nz=100, nt=900, refl=50)
    print("case cig flat reflector", flush=True)
    g = swd.Grid(P.g.nx, P.g.nz, P.g.pml, P.g.Lc, P.g.dx, P.g.dz, P.g.dt, P.g.nt)
    plan = swd.Plan(g, P.coefs, P.vp, P.rho)
    _, _, _, cig = plan.rtm_cig([(P.shot.s_ix, P.shot.s_iz, P.rec_ix, P.rec_iz)], P.wavelet, d, 90, 1)
    pm = P.pmlc
    C = cig[0].reshape(P.ntx, P.ntz, 90)[pm:-pm, pm:-pm]
    peaks = [int(np.argmax(np.abs(C[80 + h, 47:53]).sum(0))) for h in (0, 10, 20, 30)]
    print("peaks at offsets 0, 10, 20, 30:", peaks, flush=True)
    assert peaks[0] <= 4 and abs(peaks[1] - 11.8) <= 3 and abs(peaks[2] - 22.6) <= 5
    assert peaks[0] < peaks[1] < peaks[2] < peaks[3] < 45
    _, st, fs = P.forward(strips=True, final=True)
    compare("reflector cig", cig[0], P.rtm_cig(fs, st, d, 90, 1)[3])


def test_cig_c2_grid():
    c = configs.c2(nt=120, shot_ix=(420,))
    print("case c2 cig nt=120", flush=True)
    P = oracle_problem(c)
    d, st, fs = P.forward(strips=True, final=True)
    adj = O.adjoint_source(d, c.grid.dt)
    img, _, _, cig = plan_of(c).rtm_cig(c.shots, c.wavelet, adj, 10, 1)
    o = P.rtm_cig(fs, st, adj, 10, 1)
    compare("c2 image", img[0], o[0])
    compare("c2 cig", cig[0], o[3])


def test_cig_argument_errors():
    c = configs.small(nt=40)
    plan = plan_of(c)
    adj = np.zeros((c.shots[0][2].size, c.grid.nt), np.float32)
    for thetan, delt in ((0, 1), (10, 0), (2000, 1), (10, 65)):
        with pytest.raises(swd.SwdError):
            plan.rtm_cig(c.shots, c.wavelet, adj, thetan, delt)
