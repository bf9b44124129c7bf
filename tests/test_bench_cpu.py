"""CPU checks of bench.py: the one-GPU refusal and the algorithmic-byte
formulas its rooflines divide by (c2 and c5 figures computed by hand from the
per-step byte counts: state, model, CPML strips, gathers, frame strips)."""
import os
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)

import bench  # noqa: E402
import swd  # noqa: E402


def c2_grid(nt=4000):
    return swd.Grid(845, 502, 40, 4, 10.0, 10.0, 8e-4, nt)


def test_more_than_one_gpu_is_refused():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"], capture_output=True,
                       text=True, timeout=120)
    assert r.returncode == 2
    assert "one GPU" in r.stderr
    assert r.stdout.strip() == ""


def test_rtm_bytes_c2():
    g = c2_grid()
    assert (g.ntx, g.ntz) == (933, 590)
    N, Nint, P = 933 * 590, 845 * 502, 2 * 44 * 590 + 2 * 44 * 933
    nframe = 2 * 4 * (845 + 502 - 8)
    fwd, bwd = bench.rtm_bytes_per_step(g, 401, nframe)
    assert fwd == 32 * N + 16 * P + 4 * 401 + 4 * nframe
    assert bwd == 48 * Nint + 32 * N + 16 * P + 4 * nframe + 4 * 401
    assert fwd == pytest.approx(19.8e6, rel=0.01) and bwd == pytest.approx(40.2e6, rel=0.01)
    assert bench.rtm_cells(g) == 4000 * N + 3999 * (N + Nint)


def test_afor_and_updown_bytes_c2():
    g = c2_grid()
    assert bench.afor_bytes_per_step(g, 401) == pytest.approx(35.12e6, rel=0.005)
    # k_updown, interior rows (nfftz 512 < ntz): four field columns over nz rows, two images read and written
    assert bench.updown_kernel_bytes(g, 4, 2, 512) == 4 * 845 * (16 * 502 + 16 * 502)
    # whole padded columns when they fit
    assert bench.updown_kernel_bytes(g, 1, 2, 1024) == 845 * (16 * 590 + 16 * 502)


def test_cpu_baseline_host_and_native_oracle():
    """The CPU baseline's host description and its -march=native oracle build,
    which must compute what the portable oracle computes."""
    import numpy as np
    from swd import configs
    from helpers import oracle_problem

    threads, cpus = bench.host_cpu()
    assert threads >= 1 and "nproc" in cpus and "affinity" in cpus
    L, flags = bench.native_oracle()
    c = configs.small(nt=120)
    P = oracle_problem(c)
    ref = P.forward(strips=True, final=True)
    P.L = L
    got = P.forward(strips=True, final=True)
    for a, b in zip(got, ref):
        assert np.array_equal(a, b), flags
