"""Uneven-size batching (c4) and captured time loops.

c4: 64 surface shots on the c1 grid with density point scatterers on a 40 m
lattice over the c1 migration model, plus 32 single-shot instances of seeded
log-uniform sizes from 128 x 128 to 502 x 845, run by swd_multi_forward in one
fused launch per time step (replayed from a captured hipGraph).
"""
import time

import numpy as np
import pytest

import _oracle as O
import swd
from swd import configs
from helpers import oracle_problem, report

pytestmark = pytest.mark.gpu


def plans_of(cases, graphs=True):
    ps = []
    for c in cases:
        p = swd.Plan(c.grid, c.coefs, c.vp, c.rho, device=0, tile=16)
        p.set_graphs(graphs)
        ps.append(p)
    return ps


def test_multi_forward_c4_matches_single_plans_and_oracle(tmp_path, monkeypatch):
    """c4 with the tile tuned (tile = 0): the pick is recorded in the cache file."""
    cache = tmp_path / "tiles.txt"
    monkeypatch.setenv("SWD_TUNE_CACHE", str(cache))
    cases = configs.c4(nt=300)
    print(f"case c4: {len(cases)} grids, {sum(len(c.shots) for c in cases)} shots, nt 300", flush=True)
    plans = plans_of(cases)
    plans[0].set_tile(0)
    t0 = time.perf_counter()
    got = swd.multi_forward(plans, [c.shots for c in cases], [c.wavelet for c in cases])
    t_multi = time.perf_counter() - t0
    st = plans[0].stats()
    print(f"c4 tile {st['tile_x']} x {st['tile_rows']} rows, bound {st['tile_waves']}, "
          f"{st['steps_per_launch']} steps per launch", flush=True)
    rec = [ln for ln in cache.read_text().splitlines() if " multi " in ln]
    print(rec, flush=True)
    assert len(rec) == 1
    picked = int(rec[0].split()[-1])
    assert picked & 0xff == 16 and 4 * ((picked >> 8) & 0xff) in (32, 64, 128) and (picked >> 16) & 0xff in (0, 5)
    if st["steps_per_launch"] == 2:
        # two steps per launch on 32 x 32-row tiles, the tuned one-step tile for the leftover steps
        assert (st["tile_x"], st["tile_rows"]) == (32, 32)
    else:
        assert picked == st["tile_x"] | (st["tile_rows"] // 4) << 8 | st["tile_waves"] << 16
    t0 = time.perf_counter()
    ref = [p.forward(c.shots, c.wavelet) for p, c in zip(plans, cases)]
    t_single = time.perf_counter() - t0
    print(f"multi-grid launch {t_multi * 1e3:.1f} ms (first call, includes capture), "
          f"one call per grid {t_single * 1e3:.1f} ms", flush=True)
    for k, (a, b) in enumerate(zip(got, ref)):
        assert np.array_equal(a, b), k
    # the smallest and the largest uneven instance against the oracle
    sizes = [c.grid.nx * c.grid.nz for c in cases[1:]]
    for k in (1 + int(np.argmin(sizes)), 1 + int(np.argmax(sizes))):
        c = cases[k]
        o32, _, _ = oracle_problem(c).forward()
        e32, _ = report(f"c4 instance {k} ({c.grid.nx}x{c.grid.nz})", got[k], o32)
        assert e32 <= 1e-5
    # first shot of the 64-shot grid
    o32, _, _ = oracle_problem(cases[0]).forward()
    e32, _ = report("c4 shot 0", got[0][:o32.shape[0]], o32)
    assert e32 <= 1e-5
    for p in plans:
        p.close()


@pytest.mark.parametrize("streams", [1, 3])
def test_multi_forward_groups_on_other_stream_counts(streams):
    """The multi-grid groups (whole instances, their count a multiple of the
    streams) give the per-grid gathers on one and on three streams."""
    cases = configs.c4(nt=120)
    print(f"case c4 nt 120 on {streams} streams", flush=True)
    plans = plans_of(cases)
    plans[0].set_streams(streams)
    got = swd.multi_forward(plans, [c.shots for c in cases], [c.wavelet for c in cases])
    again = swd.multi_forward(plans, [c.shots for c in cases], [c.wavelet for c in cases])   # graph replays
    for p in plans[1:]:
        p.set_streams(streams)
    ref = [p.forward(c.shots, c.wavelet) for p, c in zip(plans, cases)]
    for k, (a, b, r) in enumerate(zip(got, again, ref)):
        assert np.array_equal(a, r), k
        assert np.array_equal(b, r), k
    for p in plans:
        p.close()


def test_graph_replays_match_eager():
    """Several replays of a captured loop, with eager work in between, all
    equal to the eager (graphs off) result."""
    cases = configs.c4(nt=120, nshots=3, n_uneven=3)
    print("case c4 small, graph replays", flush=True)
    eager_plans = plans_of(cases, graphs=False)
    eager = swd.multi_forward(eager_plans, [c.shots for c in cases], [c.wavelet for c in cases])
    plans = plans_of(cases, graphs=True)
    for rep in range(3):
        got = swd.multi_forward(plans, [c.shots for c in cases], [c.wavelet for c in cases])
        for a, b in zip(got, eager):
            assert np.array_equal(a, b), rep
        # eager work on the same plan between replays
        plans[1].set_graphs(False)
        g1 = plans[1].forward(cases[1].shots, cases[1].wavelet)
        plans[1].set_graphs(True)
        assert np.array_equal(g1, eager[1])
    for p in plans + eager_plans:
        p.close()


def test_rtm_graph_replay_matches_eager():
    c = configs.small(nt=160, nshots=3, dup=True)
    print("case small rtm, graph replay", flush=True)
    outs = [oracle_problem(c, k).forward() for k in range(3)]
    adj = np.concatenate([O.adjoint_source(o[0], c.grid.dt) for o in outs])
    pe = swd.Plan(c.grid, c.coefs, c.vp, c.rho, tile=16)
    pe.set_graphs(False)
    ref = pe.rtm(c.shots, c.wavelet, adj, want_gather=True)
    pg = swd.Plan(c.grid, c.coefs, c.vp, c.rho, tile=16)
    for rep in range(3):
        got = pg.rtm(c.shots, c.wavelet, adj, want_gather=True)
        for a, b in zip(got, ref):
            assert np.array_equal(a, b), rep
        pg.forward(c.shots, c.wavelet)   # other captured loops in between
    pe.close()
    pg.close()


@pytest.mark.parametrize("nt", [121, 123])
def test_multi_forward_two_steps_per_launch(nt):
    """The multi-grid forward two steps per launch (k_fstep2_multi over 32 x
    32-row tiles, the leftover steps on the one-step kernel) gives the one-step
    multi-grid launch's gathers bit for bit, on two streams."""
    cases = configs.c4(nt=nt)
    print(f"case c4 nt {nt}: two steps per launch against one", flush=True)
    res = {}
    for steps in (1, 2):
        plans = plans_of(cases)
        plans[0].set_tile(0)
        plans[0].set_steps_per_launch(steps)
        res[steps] = swd.multi_forward(plans, [c.shots for c in cases], [c.wavelet for c in cases])
        assert plans[0].stats()["steps_per_launch"] == steps
        for p in plans:
            p.close()
    for k, (a, b) in enumerate(zip(res[1], res[2])):
        assert np.array_equal(a, b), k

