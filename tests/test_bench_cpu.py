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
    N, Nint = 933 * 590, 845 * 502
    Px, Pz = 2 * 44 * 590, 2 * 44 * 933   # the CPML bands of x and z
    nframe = 2 * 4 * (845 + 502 - 8)
    rw = 2   # read and written
    # (arrays, cells, passes) of one shot-step: forward state, x-band memories (psx, pvx), z-band memories
    # (psz, pvz), the gather written, the strips written
    fwd_inv = [(3, N, rw), (2, Px, rw), (2, Pz, rw), (1, 401, 1), (1, nframe, 1)]
    # backward: the reconstruction's interior state, image and both illuminations, the receiver field's
    # state and memories, the strips and the adjoint source read
    bwd_inv = [(3, Nint, rw), (3, Nint, rw), (3, N, rw), (2, Px, rw), (2, Pz, rw), (1, nframe, 1), (1, 401, 1)]
    fwd, bwd, mf, mb = bench.rtm_bytes(g, 401, nframe)
    assert fwd == sum(4 * a * n * k for a, n, k in fwd_inv)
    assert bwd == sum(4 * a * n * k for a, n, k in bwd_inv)
    assert (mf, mb) == (4 * 3 * N, 4 * 4 * N)   # K, Bx, Bz (+ K0) once per launch
    # a 3-shot backward launch moves 116.1 MB, a one-shot launch 44.6 MB; a 4-shot forward launch 68.2 MB
    assert 3 * bwd + mb == pytest.approx(116.09e6, rel=1e-3) and bwd + mb == pytest.approx(44.57e6, rel=1e-3)
    assert 4 * fwd + mf == pytest.approx(68.21e6, rel=1e-3)
    assert bench.rtm_cells(g) == 4000 * N + 3999 * (N + Nint)


def test_afor_and_updown_bytes_c2():
    g = c2_grid()
    per_shot, per_launch = bench.afor_bytes(g, 401)
    assert per_shot + per_launch == pytest.approx(37.32e6, rel=0.005)
    assert per_launch == 12 * 933 * 590
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
