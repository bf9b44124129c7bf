"""The c5 grid (1001 x 1690 interior, 5 m, 60-cell CPML, 801 receivers on 401
nodes) and memory-planned shot waves.

The full c5 run (128 shots, nt 8000) is a throughput workload; here the same
grid runs short time windows against the oracle, and the wave planner that
sizes large batches to device memory is checked to leave results unchanged.
"""
import numpy as np
import pytest

import _oracle as O
import swd
from swd import configs
from helpers import oracle_problem, report

pytestmark = pytest.mark.gpu


def test_rtm_c5_grid_vs_oracle():
    c = configs.c5(nt=240, shot_index=[0, 64])
    print("case c5 grid rtm nt=240, 2 shots", flush=True)
    Ps = [oracle_problem(c, k) for k in range(2)]
    outs = [P.forward(strips=True, final=True) for P in Ps]
    adj = np.concatenate([O.adjoint_source(o[0], c.grid.dt) for o in outs])
    plan = swd.Plan(c.grid, c.coefs, c.vp, c.rho)
    img, ils, ilr, gs = plan.rtm(c.shots, c.wavelet, adj, want_gather=True)
    row = 0
    for k, P in enumerate(Ps):
        d, st, fs = outs[k]
        n = d.shape[0]
        o = P.rtm_backward(fs, st, adj[row:row + n])
        e_g, _ = report(f"c5 shot {k} gather", gs[row:row + n], d)
        e_i, _ = report(f"c5 shot {k} image", img[k], o[0])
        e_s, _ = report(f"c5 shot {k} illum_src", ils[k], o[1])
        assert e_g <= 1e-5 and e_i <= 1e-5 and e_s <= 1e-5
        row += n
    plan.close()


def test_born_adjoint_c5_grid_vs_oracle():
    c = configs.c5(nt=160, shot_index=[10])
    print("case c5 grid born adjoint nt=160", flush=True)
    rng = np.random.default_rng(21)
    d = rng.standard_normal((c.shots[0][2].size, c.grid.nt)).astype(np.float32)
    plan = swd.Plan(c.grid, c.coefs, c.vp, c.rho)
    g = plan.born_adjoint(c.shots, c.wavelet, d)
    P = oracle_problem(c)
    e32, _ = report("c5 born adjoint", g, P.born_adjoint(d))
    assert e32 <= 1e-5
    plan.close()


def test_shot_waves_match_single_wave():
    c = configs.small(nt=150, nshots=5, dup=True)
    print("case small, 5 shots, memory-limited waves", flush=True)
    outs = [oracle_problem(c, k).forward() for k in range(5)]
    adj = np.concatenate([O.adjoint_source(o[0], c.grid.dt) for o in outs])
    rng = np.random.default_rng(5)
    m = rng.standard_normal(5 * c.grid.nx * c.grid.nz).astype(np.float32)
    plan = swd.Plan(c.grid, c.coefs, c.vp, c.rho)
    ref_f = plan.forward(c.shots, c.wavelet)
    ref_r = plan.rtm(c.shots, c.wavelet, adj, want_gather=True)
    ref_b = plan.born(c.shots, c.wavelet, m)
    ref_a = plan.born_adjoint(c.shots, c.wavelet, adj)
    assert plan.stats()["waves"] == 1
    for limit, op in ((2 << 20, "forward"), (6_500_000, "rtm"), (2_300_000, "born"), (17 << 20, "born_adjoint")):
        plan.set_memory_limit(limit)
        if op == "forward":
            assert np.array_equal(plan.forward(c.shots, c.wavelet), ref_f)
        elif op == "rtm":
            for a, b in zip(plan.rtm(c.shots, c.wavelet, adj, want_gather=True), ref_r):
                assert np.array_equal(a, b)
        elif op == "born":
            assert np.array_equal(plan.born(c.shots, c.wavelet, m), ref_b)
        else:
            assert np.array_equal(plan.born_adjoint(c.shots, c.wavelet, adj), ref_a)
        waves = plan.stats()["waves"]
        print(f"{op}: limit {limit / 1e6:.1f} MB -> {waves} waves", flush=True)
        assert waves >= 2
    plan.set_memory_limit(1000)
    with pytest.raises(swd.SwdError):
        plan.forward(c.shots, c.wavelet)
    plan.close()
