"""Benchmark of the RTM hot path on one MI355X.

Workload: the c2 configuration (single-shot VSP RTM at reference scale:
interior 502 x 845, 10 m, dt 0.8 ms, nt 4000, 40-cell CPML, Lc 4, 401 borehole
receivers, smoothed migration model) run as a batch of shots from surface
positions.  One step = one batched RTM pass (forward with boundary strips,
source reconstruction, receiver back-propagation, imaging) over the batch, with
the adjoint sources already in HBM.

Prints one JSON line: Gcell-updates/s (cell-update = one leapfrog step of one
wavefield at one padded node; RTM per shot = nt*N + (nt-1)*(N + N_int)),
shots/s, the plain one-shot-per-launch path's shots/s, the roofline of the
dominant kernel (the fused backward step: algorithmic bytes per launch over the
HIP-event launch time; HBM traffic per launch from the committed rocprofv3
counter summary) and a CPU baseline (the oracle's fp32 host loops with OpenMP,
on a bounded sample of the same workload).

Other workloads (--workload), each one JSON line of the same form:
  decon   the d1 PSF deconvolution (845 x 702 image, 2484 windows)
  updown  the c2 up/down-separated RTM (du, ud images)
  cig     the c2 RTM with angle-domain common-image gathers
  afor    the c2 AFor forward (p and Hilbert wavefields, kz up/down gathers)
  born    the c5 Born forward and exact adjoint (sqrt(nt) checkpoints)
  multi   the c4 uneven-grid forward (33 grids in one launch per step and group)
  c5      the c5 batched RTM + up/down RTM + Born misfit gradient (128 shots, nt 8000)
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "simple-wavenumber-deconvolution_amd"))

import numpy as np  # noqa: E402

HBM_PEAK_GBS = 8000.0


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def rtm_bytes(g, nrec, nframe):
    """Algorithmic HBM bytes of the RTM steps: (forward per shot, backward per shot, forward per launch,
    backward per launch).  Per shot and step: the wavefield state (p, vx, vz read and written), the CPML
    memories on the frame, the strips and the receiver traces; the backward adds the reconstruction's
    interior state and the image and both illuminations, read and written.  Per launch, shared by the shots
    of its group: the model, K, Bx and Bz, in the backward also K0."""
    N = g.ntx * g.ntz
    Nint = g.nx * g.nz
    P = 2 * g.pmlc * g.ntz + 2 * g.pmlc * g.ntx
    fwd = 24 * N + 16 * P + 4 * nrec + 4 * nframe
    bwd = 24 * Nint + 24 * N + 16 * P + 24 * Nint + 4 * nframe + 4 * nrec
    return fwd, bwd, 12 * N, 16 * N


def rtm_cells(g):
    N = g.ntx * g.ntz
    return g.nt * N + (g.nt - 1) * (N + g.nx * g.nz)


def build_case(nshots, nt):
    from swd import configs

    nx = 845
    xs = [int(v) for v in np.linspace(60, nx - 60, nshots)]
    return configs.c2(nt=nt, shot_ix=xs)


def host_cpu():
    """(threads to use, description of the host's CPUs): the threads OMP_NUM_THREADS
    grants (a host shared between GPUs may give each GPU's process a share of its
    CPUs), else every CPU this process may run on."""
    avail = len(os.sched_getaffinity(0))
    try:
        threads = min(avail, int(os.environ.get("OMP_NUM_THREADS", "")))
    except ValueError:
        threads = avail
    model = "unknown CPU"
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    return threads, f"{model}; nproc {os.cpu_count()}, {avail} in this process's affinity mask"


def native_oracle():
    """The oracle built with -march=native for this host (the CPU baseline only; the
    parity tests use the portable build), or the portable build if that fails."""
    import ctypes as C
    import subprocess
    import tempfile

    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import _oracle as O

    out = tempfile.mkdtemp(prefix="swd_oracle_native_")
    so = os.path.join(out, "libswd_oracle_native.so")
    cmd = ["gcc", "-O3", "-march=native", "-ffp-contract=off", "-fno-fast-math", "-fopenmp", "-fPIC", "-w",
           "-shared", "-o", so, os.path.join(ROOT, "oracle", "swd_oracle.c"), "-lm"]
    try:
        subprocess.run(cmd, check=True, capture_output=True, timeout=120)
        return O._bind(C.CDLL(so)), "gcc -O3 -march=native -ffp-contract=off"
    except (OSError, subprocess.SubprocessError):
        return O.lib(), "gcc -O3 -march=x86-64-v3 -ffp-contract=off (portable build; the native build failed)"


def cpu_baseline(case, nt_sample, runs=5):
    """The oracle's fp32 host loops (OpenMP over x) on one shot of the workload with
    nt_sample steps (forward with strips, reconstruction, receiver propagation,
    imaging): one warm-up run, then the median of `runs`."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import _oracle as O
    from helpers import oracle_problem
    import swd

    L, flags = native_oracle()
    threads, cpus = host_cpu()
    L.or_set_threads(threads)
    g = case.grid
    sub = swd.configs.Case(case.name, swd.Grid(g.nx, g.nz, g.pml, g.Lc, g.dx, g.dz, g.dt, nt_sample),
                           case.vp, case.rho, case.coefs, case.wavelet[:nt_sample], case.shots[:1], case.f0)
    P = oracle_problem(sub, 0)
    P.L = L
    times = []
    for r in range(runs + 1):
        t0 = time.perf_counter()
        d, st, fs = P.forward(strips=True, final=True)
        adj = O.adjoint_source(d, g.dt)
        P.rtm_backward(fs, st, adj)
        if r > 0:
            times.append(time.perf_counter() - t0)
    med = float(np.median(times))
    rate = rtm_cells(sub.grid) / med / 1e9
    return {"value": rate, "unit": "Gcell-updates/s", "cores": threads, "kind": "port",
            "sample": f"oracle fp32 host loops (OpenMP over x, {flags}) on 1 c2 shot with nt={nt_sample} "
                      f"(forward with strips + reconstruction + receiver propagation + imaging): median of "
                      f"{runs} runs after 1 warm-up, {med:.2f}s each; {threads} threads on {cpus}",
            "cpu": cpus, "runs_s": times}


def launch_config(st):
    """(steps per launch, tile columns, rows, waves argument) of a plan's last fused loop, as
    swd_plan_set_steps_per_launch and swd_plan_set_tile_config take them."""
    two = st["steps_per_launch"] == 2
    return (st["steps_per_launch"], st["tile_x"], st["tile_rows"], st["tile_group_waves"] if two else st["tile_waves"])


def backward_launch_ms(swd, torch, case, adj, nshots, group, streams, nt, tile, cfg=None):
    """HIP-event time per backward launch of an RTM of the first nshots shots of
    case cut to nt steps, in groups of `group` shots on `streams` streams: the
    plan's backward-phase events (on its stream) over its launches, on a second
    call (the first tunes the tile and captures the graphs), and the launches
    per step.  cfg (launch_config) pins the launch configuration."""
    g = case.grid
    sub = swd.configs.Case(case.name, swd.Grid(g.nx, g.nz, g.pml, g.Lc, g.dx, g.dz, g.dt, nt), case.vp, case.rho,
                           case.coefs, case.wavelet[:nt], case.shots[:nshots], case.f0)
    plan = swd.Plan(sub.grid, sub.coefs, sub.vp, sub.rho, device=0, tile=tile)
    if cfg is not None:
        plan.set_steps_per_launch(cfg[0])
        plan.set_tile_config(*cfg[1:])
    plan.set_groups(group, group)
    plan.set_streams(streams)
    shots = swd.ShotSet(sub.shots)
    dev = torch.device("cuda:0")
    d_adj = torch.from_numpy(np.ascontiguousarray(adj[:shots.ntr, :nt])).to(dev)
    ntp = g.ntx * g.ntz
    d_img = torch.zeros((3, nshots, ntp), dtype=torch.float32, device=dev)
    for _ in range(2):
        plan.rtm_dev(shots, sub.wavelet, d_adj.data_ptr(), d_img[0].data_ptr(), d_img[1].data_ptr(),
                     d_img[2].data_ptr())
    st = plan.stats()
    plan.close()
    del d_adj, d_img
    return st["bwd_ms"] / max(1, st["bwd_launches"]), st["bwd_launches"] / (nt - 1)


def copy_bandwidth(torch, dev, nbytes=1 << 30, reps=5):
    """Achievable HBM bandwidth on this device: a device-to-device copy of 1 GiB
    (read + write bytes over the best of a few HIP-event-timed copies)."""
    src = torch.ones(nbytes // 4, dtype=torch.float32, device=dev)
    dst = torch.empty_like(src)
    best = float("inf")
    for _ in range(reps):
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record()
        dst.copy_(src)
        e1.record()
        e1.synchronize()
        best = min(best, e0.elapsed_time(e1))
    del src, dst
    return 2 * nbytes / (best * 1e-3) / 1e9


def decon_window_flops(g):
    """FMAs x 2 of one window in k_decon_windows: the rows pass over the half
    spectrum, the columns pass over conjugate wavenumber pairs, the Wiener
    division (8 flops a bin, the division counted once) and the inverse at the
    output block."""
    n, nh, mx, mz, out = g["n"], g["n"] // 2 + 1, g["filterwin"], g["filterwin"], g["outwinx"]
    fma = mx * nh * mz * 2 + mx * nh * nh * 4 + out * nh * n * 4 + out * out * nh * 2
    return 2 * fma + 8 * n * nh


def bench_decon(a):
    """The wavenumber-domain PSF deconvolution (kkdecon) of the d1 workload:
    845 x 702 image and PSF image, psfwin 31, filterwin 71, snapwin 11, eps 5e-4
    (2484 windows of 101 x 101).  One step = one whole deconvolution (PSF cell
    spectra, window transforms, overlap-add) with image and PSF resident in HBM."""
    import torch
    import swd
    from swd import configs

    nx, nz = 845, 702
    img, psf = configs.decon_case(nx, nz)
    prm = swd.decon_params(nx, nz, **{k: configs.DECON_PARAMS[k] for k in ("psfwin", "filterwin", "snapwin", "epsilon")})
    geo = swd.decon_geometry(prm)
    plan = swd.Decon(prm, 0)
    dev = torch.device("cuda:0")
    ti, tp = torch.from_numpy(img).to(dev), torch.from_numpy(psf).to(dev)
    td, tc = torch.empty_like(ti), torch.empty_like(ti)
    torch.cuda.synchronize()

    def step():
        plan.run_dev(ti.data_ptr(), tp.data_ptr(), td.data_ptr(), tc.data_ptr())

    for i in range(a.warmup):
        step()
    torch.cuda.synchronize()
    win_ms = 0.0
    t0 = time.perf_counter()
    for i in range(a.steps):
        step()
        win_ms += plan.stats()["windows_ms"]
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t0
    st = plan.stats()
    windows = geo["windows"]
    flops = decon_window_flops({"n": geo["n"], "filterwin": prm.filterwinx, "outwinx": geo["outwinx"]}) * windows
    achieved = flops / (win_ms / a.steps * 1e-3) / 1e12
    cpu = None
    if not a.no_cpu:
        sys.path.insert(0, os.path.join(ROOT, "tests"))
        import _oracle as O

        threads, cpus = host_cpu()
        O.set_threads(threads)
        t1 = time.perf_counter()
        odec, _ = O.decon(O.decon_params(nx, nz), img, psf)
        dt = time.perf_counter() - t1
        same = bool(np.array_equal(odec, td.cpu().numpy()))
        cpu = {"value": windows / dt, "unit": "windows/s", "cores": threads, "kind": "port",
               "sample": f"oracle (direct fp64 DFTs, OpenMP over windows) on the whole d1 image, {dt:.2f}s; "
                         f"GPU output bit-identical: {same}; {threads} threads on {cpus}"}
    out = {
        "metric": "PSF-deconvolution windows/s (kkdecon, 101x101 windows)",
        "value": windows * a.steps / elapsed,
        "unit": "windows/s",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": elapsed * 1e3 / a.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (d1: depth-varying wavelet PSF lattice, dipping reflectors and diffractors, seed 11)",
        "config": {"workload": "d1 PSF deconvolution: 845x702 image, psfwin 31, filterwin 71, snapwin 11, "
                               "epsilon 5e-4, 2484 windows, 374 PSF cells",
                   "kernel_ms": {k: st[k] for k in ("cells_ms", "windows_ms", "gather_ms")}},
        "roofline": {"bound": "mfma", "achieved": achieved, "peak": 78.6, "unit": "TFLOP/s",
                     "frac": achieved / 78.6, "traffic": None,
                     "kernel": "k_decon_windows (fp64 FMAs on the vector ALUs; fp64 peak 78.6 TFLOP/s)",
                     "flops_per_launch": flops, "avg_launch_ms": win_ms / a.steps},
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)


def updown_kernel_bytes(g, nshots, nimg, nfftz):
    """Algorithmic HBM bytes of one k_updown launch: per interior column of a
    shot the four field columns are read over the transformed rows (the whole
    padded column when nfftz >= ntz, else the nz interior rows) and nimg images
    are read and written on the interior rows."""
    nzw = g.ntz if nfftz >= g.ntz else g.nz
    return nshots * g.nx * (16 * nzw + 8 * nimg * g.nz)


def bench_updown(a):
    """Up/down-separated RTM of the c2 batch: one step = forward of p and its
    time-Hilbert companion with strips, the paired backward of both, and the
    kz split + cross-correlation of every interior column after every backward
    step (du, ud images), with the adjoint sources and their Hilbert
    transforms resident in HBM."""
    import torch
    import swd

    log(f"bench: c2 up/down RTM, batch {a.batch} shots, nt {a.nt}")
    case = build_case(a.batch, a.nt)
    g = case.grid
    plan = swd.Plan(g, case.coefs, case.vp, case.rho, device=0, tile=a.tile)
    plan.set_groups(*a.groups)
    plan.set_streams(a.streams)
    shots = swd.ShotSet(case.shots)
    plan_true = swd.Plan(g, case.coefs, case.vp_true, case.rho, device=0)
    obs = plan_true.forward(shots, case.wavelet)
    plan_true.close()
    adj = swd.adjoint_source(obs, g.dt)
    t0 = time.perf_counter()
    hadj = swd.hilbert_traces(adj)
    log(f"bench: Hilbert transform of {adj.shape[0]} adjoint traces on the host in {time.perf_counter() - t0:.1f}s")
    hw = swd.hilbert1d(case.wavelet)
    # the c2 configuration's "nz padded to 512": the interior rows of each column
    nfftz = a.nfftz or max(128, swd.nfft(g.nz))
    kz, _ = swd.kz_taper(nfftz, g.dz)
    dev = torch.device("cuda:0")
    ntp = g.ntx * g.ntz
    d_adj, d_hadj = torch.from_numpy(adj).to(dev), torch.from_numpy(hadj).to(dev)
    d_img = torch.zeros((2, a.batch, ntp), dtype=torch.float32, device=dev)
    torch.cuda.synchronize()

    def step():
        plan.rtm_updown_dev(shots, case.wavelet, hw, d_adj.data_ptr(), d_hadj.data_ptr(), kz,
                            {"du": d_img[0].data_ptr(), "ud": d_img[1].data_ptr()})

    copy_gbs = copy_bandwidth(torch, dev)
    for i in range(a.warmup):
        step()
        log(f"bench: warmup {i + 1}/{a.warmup} done")
    torch.cuda.synchronize()
    fwd_ms = bwd_ms = 0.0
    t0 = time.perf_counter()
    for i in range(a.steps):
        step()
        st = plan.stats()
        fwd_ms += st["fwd_ms"]
        bwd_ms += st["bwd_ms"]
        log(f"bench: step {i + 1}/{a.steps} fwd {st['fwd_ms']:.1f} ms bwd {st['bwd_ms']:.1f} ms")
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t0
    shots_per_s = a.batch * a.steps / elapsed
    cells = 2 * rtm_cells(g) * a.batch
    du = d_img[0].cpu().numpy()

    # the plain RTM of the same batch, for the cost of the separation
    rtm = None
    if not a.no_plain:
        im = torch.zeros((3, a.batch, ntp), dtype=torch.float32, device=dev)
        plan.rtm_dev(shots, case.wavelet, d_adj.data_ptr(), im[0].data_ptr(), im[1].data_ptr(), im[2].data_ptr())
        torch.cuda.synchronize()
        t1 = time.perf_counter()
        plan.rtm_dev(shots, case.wavelet, d_adj.data_ptr(), im[0].data_ptr(), im[1].data_ptr(), im[2].data_ptr())
        torch.cuda.synchronize()
        tr = time.perf_counter() - t1
        rtm = {"shots_per_s": a.batch / tr, "updown_cost_vs_rtm": (a.batch / tr) / shots_per_s}
        del im

    # k_updown alone: one imaging step of n_k c2 shots on the plan's stream, HIP events around 40 launches
    n_k = 4
    rng = np.random.default_rng(1)
    f = torch.from_numpy(rng.standard_normal((4, n_k, ntp)).astype(np.float32)).to(dev)
    im = torch.zeros((2, n_k, ntp), dtype=torch.float32, device=dev)
    ptrs = {"du": im[0].data_ptr(), "ud": im[1].data_ptr()}
    args = (n_k, f[0].data_ptr(), f[1].data_ptr(), f[2].data_ptr(), f[3].data_ptr(), kz, ptrs)
    torch.cuda.synchronize()
    ps = torch.cuda.ExternalStream(plan.stream())
    reps = 40
    with torch.cuda.stream(ps):
        for _ in range(3):
            plan.updown_image_dev(*args)
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record()
        for _ in range(reps):
            plan.updown_image_dev(*args)
        e1.record()
    e1.synchronize()
    k_ms = e0.elapsed_time(e1) / reps
    k_bytes = updown_kernel_bytes(g, n_k, 2, nfftz)
    k_flops = n_k * g.nx * 4 * 5 * nfftz * int(np.log2(nfftz))
    achieved = k_bytes / (k_ms * 1e-3) / 1e9
    del f, im
    traffic = None
    tfile = os.path.join(ROOT, "profiles", "updown_c2_traffic.json")
    if os.path.exists(tfile) and nfftz == 512 and (g.nx, g.nz, g.pml, g.Lc) == (845, 502, 40, 4):
        with open(tfile) as fh:
            traffic = json.load(fh)["l2_miss_bytes_per_shot"]["4-shot launch"] * n_k

    out = {
        "metric": "up/down-separated RTM Gcell-updates/s (c2 batch, du and ud images)",
        "value": cells * a.steps / elapsed / 1e9,
        "unit": "Gcell-updates/s",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": elapsed * 1e3 / a.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (seeded layered model with dipping reflectors, c2 geometry)",
        "config": {"workload": "c2 batched up/down-separated VSP RTM: 502x845 interior (590x933 padded), nt "
                               f"{g.nt}, nfftz {nfftz} ({'whole padded columns' if nfftz >= g.ntz else 'interior rows'}), "
                               "p and Hilbert-companion wavefields on both sides, kz split of every interior column "
                               "after every backward step",
                   "shots_per_step": a.batch, "shots_per_s": shots_per_s, "plain_rtm": rtm,
                   "fwd_ms_per_step": fwd_ms / a.steps, "bwd_ms_per_step": bwd_ms / a.steps,
                   "du_finite": bool(np.isfinite(du).all())},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": traffic, "copy_GBs": copy_gbs,
                     "kernel": f"k_updown (kz split + imaging of one backward step, {n_k} c2 shots, "
                               f"{g.nx} columns each, nfftz {nfftz})",
                     "bytes_per_launch": k_bytes, "avg_launch_ms": k_ms,
                     "fft_gflops": k_flops / (k_ms * 1e-3) / 1e9,
                     "timing": "HIP events on the plan's stream around 40 back-to-back launches"},
        "cpu_baseline": None,
    }
    print(json.dumps(out), flush=True)


def bench_cig(a):
    """RTM of the c2 batch with angle-domain common-image gathers (thetan 10,
    delt 1 as the reference's RC:235): one step = the fused RTM plus the
    gather update after every backward step, adjoint sources resident in HBM.
    The plain RTM of the same batch runs beside it for the cost of the gathers."""
    import torch
    import swd

    thetan, delt = a.thetan, 1
    log(f"bench: c2 RTM with angle gathers (thetan {thetan}), batch {a.batch} shots, nt {a.nt}")
    case = build_case(a.batch, a.nt)
    g = case.grid
    plan = swd.Plan(g, case.coefs, case.vp, case.rho, device=0, tile=a.tile)
    plan.set_groups(*a.groups)
    plan.set_streams(a.streams)
    shots = swd.ShotSet(case.shots)
    plan_true = swd.Plan(g, case.coefs, case.vp_true, case.rho, device=0)
    obs = plan_true.forward(shots, case.wavelet)
    plan_true.close()
    adj = swd.adjoint_source(obs, g.dt)
    dev = torch.device("cuda:0")
    ntp = g.ntx * g.ntz
    d_adj = torch.from_numpy(adj).to(dev)
    d_img = torch.zeros((3, a.batch, ntp), dtype=torch.float32, device=dev)
    d_cig = torch.zeros((a.batch, ntp, thetan), dtype=torch.float32, device=dev)
    torch.cuda.synchronize()

    def step():
        plan.rtm_cig_dev(shots, case.wavelet, d_adj.data_ptr(), thetan, delt, d_img[0].data_ptr(),
                         d_img[1].data_ptr(), d_img[2].data_ptr(), d_cig.data_ptr())

    def plain():
        plan.rtm_dev(shots, case.wavelet, d_adj.data_ptr(), d_img[0].data_ptr(), d_img[1].data_ptr(),
                     d_img[2].data_ptr())

    for i in range(a.warmup):
        step()
        plain()
    torch.cuda.synchronize()
    bwd = {"cig": 0.0, "plain": 0.0}
    launches = 0
    t_cig = t_plain = 0.0
    for i in range(a.steps):
        t0 = time.perf_counter()
        step()
        torch.cuda.synchronize()
        t_cig += time.perf_counter() - t0
        st = plan.stats()
        bwd["cig"] += st["bwd_ms"]
        launches = st["bwd_launches"]
        t0 = time.perf_counter()
        plain()
        torch.cuda.synchronize()
        t_plain += time.perf_counter() - t0
        bwd["plain"] += plan.stats()["bwd_ms"]
        log(f"bench: step {i + 1}/{a.steps} backward with gathers {st['bwd_ms']:.1f} ms, plain "
            f"{plan.stats()['bwd_ms']:.1f} ms")
    cells = rtm_cells(g) * a.batch
    n_cig = launches // 2   # one k_cig after each backward k_fstep launch
    added_ms = (bwd["cig"] - bwd["plain"]) / a.steps / max(1, n_cig)
    shots_per_launch = a.batch * (g.nt - 1) / max(1, n_cig)
    nint = g.nx * g.nz
    k_bytes = 32.0 * nint * shots_per_launch   # p, vx, vz of both fields, Bx, Bz per interior cell
    achieved = k_bytes / (added_ms * 1e-3) / 1e9 if added_ms > 0 else None
    out = {
        "metric": "RTM with angle-domain CIGs, Gcell-updates/s (c2 batch)",
        "value": cells * a.steps / t_cig / 1e9,
        "unit": "Gcell-updates/s",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": t_cig * 1e3 / a.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (seeded layered model with dipping reflectors, c2 geometry)",
        "config": {"workload": f"c2 batched VSP RTM with angle gathers (thetan {thetan}, delt {delt}): 502x845 "
                               f"interior, nt {g.nt}",
                   "shots_per_step": a.batch, "plain_rtm_gcell_per_s": cells * a.steps / t_plain / 1e9,
                   "bwd_ms_with_gathers": bwd["cig"] / a.steps, "bwd_ms_plain": bwd["plain"] / a.steps,
                   "gather_cost_fraction": (t_cig - t_plain) / t_plain,
                   "cig_finite": bool(torch.isfinite(d_cig).all().item())},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS if achieved else None, "traffic": None,
                     "kernel": "k_cig (angle-gather update of one backward step, one launch per group)",
                     "bytes_per_launch": k_bytes, "avg_launch_ms": added_ms,
                     "timing": "backward-phase HIP events with gathers minus without, over the k_cig launches: "
                               "the time the gathers add, with the two streams' overlap"},
        "cpu_baseline": None,
    }
    print(json.dumps(out), flush=True)


def afor_bytes(g, nrec):
    """Algorithmic HBM bytes of an AFor step: (per shot, per launch).  Per shot:
    the p and Hilbert wavefields' state and CPML memories and the gathers (the
    kz transforms of the receiver columns add under 0.1%); per launch the model
    (K, Bx, Bz), shared by the launch's wavefields."""
    N = g.ntx * g.ntz
    P = 2 * g.pmlc * g.ntz + 2 * g.pmlc * g.ntx
    return 48 * N + 32 * P + 8 * nrec, 12 * N


def bench_afor(a):
    """AFor of the c2 batch: one step = the p and Hilbert-companion forward
    wavefields of every shot (nt steps each), the receiver-column capture and
    the kz up/down separation of the gathers (nfftz by the reference's rule,
    the padded column length rounded up to a power of two), gathers left in HBM."""
    import torch
    import swd

    log(f"bench: c2 AFor, batch {a.batch} shots, nt {a.nt}")
    case = build_case(a.batch, a.nt)
    g = case.grid
    plan = swd.Plan(g, case.coefs, case.vp_true, case.rho, device=0, tile=a.tile)
    plan.set_groups(*a.groups)
    plan.set_streams(a.streams)
    plan.set_steps_per_launch(a.steps_per_launch)
    shots = swd.ShotSet(case.shots)
    nfftz = swd.nfft(g.ntz)
    kz, _ = swd.kz_taper(nfftz, g.dz)
    hw = swd.hilbert1d(case.wavelet)
    dev = torch.device("cuda:0")
    d = torch.zeros((3, shots.ntr, g.nt), dtype=torch.float32, device=dev)
    torch.cuda.synchronize()

    def step(pl, sh, out):
        pl.afor_dev(sh, case.wavelet, hw, kz, out[0].data_ptr(), out[1].data_ptr(), out[2].data_ptr())

    copy_gbs = copy_bandwidth(torch, dev)
    for i in range(a.warmup):
        step(plan, shots, d)
        log(f"bench: warmup {i + 1}/{a.warmup} done")
    torch.cuda.synchronize()
    fwd_ms = fft_ms = 0.0
    t0 = time.perf_counter()
    for i in range(a.steps):
        step(plan, shots, d)
        st = plan.stats()
        fwd_ms += st["fwd_ms"]
        fft_ms += st["fft_ms"]
        log(f"bench: step {i + 1}/{a.steps} wavefields {st['fwd_ms']:.1f} ms, kz separation {st['fft_ms']:.1f} ms")
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t0
    N = g.ntx * g.ntz
    cells = 2 * g.nt * N * a.batch
    shots_per_s = a.batch * a.steps / elapsed
    nrec = len(case.shots[0][2])
    st = plan.stats()
    launch_ms = fwd_ms / a.steps / max(1, st["fwd_launches"])
    launches_per_step = st["fwd_launches"] / g.nt
    afs, afm = afor_bytes(g, nrec)
    bytes_per_launch = (afs * a.batch + afm * launches_per_step) / launches_per_step
    achieved = bytes_per_launch / (launch_ms * 1e-3) / 1e9
    up0 = d[1, :nrec].cpu().numpy()
    traffic = None
    # the counter figures are per wavefield-step of the one-step kernel, or of the two-step kernel
    tfile = os.path.join(ROOT, "profiles", "afor2_c2_traffic.json" if st["fwd_steps_per_launch"] == 2
                         else "afor_c2_traffic.json")
    if os.path.exists(tfile) and (g.nx, g.nz, g.pml, g.Lc) == (845, 502, 40, 4):
        with open(tfile) as f:
            traffic = json.load(f)["l2_miss_bytes_per_wavefield_step"] * 2 * a.batch / launches_per_step

    plain = None
    if not a.no_plain:
        one = swd.ShotSet(case.shots[:1])
        pp = swd.Plan(g, case.coefs, case.vp_true, case.rho, device=0, path=swd.SWD_PATH_PLAIN)
        d1 = torch.zeros((3, one.ntr, g.nt), dtype=torch.float32, device=dev)
        torch.cuda.synchronize()
        t1 = time.perf_counter()
        step(pp, one, d1)
        torch.cuda.synchronize()
        tp = time.perf_counter() - t1
        pp.close()
        rel = float(np.linalg.norm(d1[1].cpu().numpy() - up0) / max(1e-30, np.linalg.norm(up0)))
        plain = {"shots_per_s": 1.0 / tp, "gcell_per_s": 2 * g.nt * N / tp / 1e9, "up_rel_l2_vs_batched": rel}
        log(f"bench: plain path one shot {tp:.2f}s, up-going gather rel diff vs batched {rel:.2e}")
        del d1

    cpu = None
    if not a.no_cpu:
        sys.path.insert(0, os.path.join(ROOT, "tests"))
        import _oracle as O
        from helpers import oracle_problem

        threads, cpus = host_cpu()
        O.set_threads(threads)
        true_case = swd.configs.Case(case.name, g, case.vp_true, case.rho, case.coefs, case.wavelet,
                                     case.shots[:2], case.f0)
        t1 = time.perf_counter()
        for k in range(2):
            oracle_problem(true_case, k).afor(hw, kz)
        dt = time.perf_counter() - t1
        cpu = {"value": 2 * 2 * g.nt * N / dt / 1e9, "unit": "Gcell-updates/s", "cores": threads, "kind": "port",
               "sample": f"oracle fp32 host loops (OpenMP) on 2 c2 shots with nt={g.nt} (p and Hilbert "
                         f"wavefields, capture, kz separation), {dt:.1f}s; {threads} threads on {cpus}"}
        log(f"bench: cpu baseline {cpu['value']:.3f} Gcell/s in {dt:.1f}s on {threads} threads")

    out = {
        "metric": "AFor Gcell-updates/s (c2 batch, p and Hilbert wavefields, kz up/down gathers)",
        "value": cells * a.steps / elapsed / 1e9,
        "unit": "Gcell-updates/s",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": elapsed * 1e3 / a.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (seeded layered model with dipping reflectors, c2 geometry)",
        "config": {"workload": f"c2 batched AFor: 502x845 interior (590x933 padded), nt {g.nt}, Lc 4, "
                               f"{nrec} borehole receivers, nfftz {nfftz}",
                   "shots_per_step": a.batch, "shots_per_s": shots_per_s, "plain_path": plain,
                   "speedup_vs_plain_shots_per_s": (shots_per_s / plain["shots_per_s"]) if plain else None,
                   "wavefield_ms_per_step": fwd_ms / a.steps, "kz_separation_ms_per_step": fft_ms / a.steps,
                   "up_finite": bool(np.isfinite(up0).all())},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": traffic,
                     "copy_GBs": copy_gbs, "frac_of_copy": achieved / copy_gbs,
                     "kernel": (f"k_fstep2 (two forward steps per launch, {st['fwd_tile_x']} x {st['fwd_tile_rows']}-row "
                                f"tiles, of a group of p and Hilbert wavefields with capture)"
                                if st["fwd_steps_per_launch"] == 2 else
                                "k_fstep (forward step of a group of p and Hilbert wavefields with capture)"),
                     "steps_per_launch": st["fwd_steps_per_launch"],
                     "bytes_per_launch": bytes_per_launch, "avg_launch_ms": launch_ms,
                     "timing": "HIP events on the plan's stream around the wavefield loop (joining the other "
                               "streams) / launches: the interval between launch completions",
                     "traffic_source": f"rocprofv3 FETCH_SIZE x2 + WRITE_SIZE per wavefield-step, "
                                       f"profiles/{os.path.basename(tfile)}" if traffic else None},
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)


def bench_born(a):
    """Born modelling d = L m and its exact adjoint m_grad = L^T d on the c5 grid
    (1001 x 1690 interior, 5 m, dt 0.4 ms, nt 8000, 60-cell CPML, 801 receivers
    on 401 nodes): one step = the Born forward of the batch, then its adjoint
    with sqrt(nt) checkpoints, m and d resident in HBM."""
    import torch
    import swd

    from swd import configs

    nshots = a.born_shots
    log(f"bench: c5 Born forward + adjoint, {nshots} shots, nt {a.nt_born}")
    idx = [int(v) for v in np.linspace(0, 127, nshots)]
    case = configs.c5(nt=a.nt_born, shot_index=idx)
    g = case.grid
    plan = swd.Plan(g, case.coefs, case.vp, case.rho, device=0)
    plan.set_groups(*a.groups)
    plan.set_streams(a.streams)
    shots = swd.ShotSet(case.shots)
    rng = np.random.default_rng(3)
    nint = g.nx * g.nz
    dev = torch.device("cuda:0")
    d_m = torch.from_numpy(rng.standard_normal(nshots * nint).astype(np.float32)).to(dev)
    d_d = torch.zeros((shots.ntr, g.nt), dtype=torch.float32, device=dev)
    d_res = torch.from_numpy(rng.standard_normal((shots.ntr, g.nt)).astype(np.float32)).to(dev)
    d_g = torch.zeros(nint, dtype=torch.float32, device=dev)
    torch.cuda.synchronize()
    copy_gbs = copy_bandwidth(torch, dev)

    def fwd():
        plan.born_dev(shots, case.wavelet, d_m.data_ptr(), d_d.data_ptr())

    def adj():
        plan.born_adjoint_dev(shots, case.wavelet, d_res.data_ptr(), d_g.data_ptr())

    for i in range(a.warmup):
        fwd()
        adj()
        log(f"bench: warmup {i + 1}/{a.warmup} done")
    torch.cuda.synchronize()
    t_f = t_a = f_ms = 0.0
    f_launches = f_spl = 0
    for i in range(a.steps):
        t0 = time.perf_counter()
        fwd()
        torch.cuda.synchronize()
        t1 = time.perf_counter()
        st = plan.stats()
        f_ms += st["fwd_ms"]
        f_launches = st["fwd_launches"]
        f_spl = st["steps_per_launch"]
        adj()
        torch.cuda.synchronize()
        t2 = time.perf_counter()
        t_f += t1 - t0
        t_a += t2 - t1
        log(f"bench: step {i + 1}/{a.steps} Born forward {1e3 * (t1 - t0):.0f} ms, adjoint {1e3 * (t2 - t1):.0f} ms")
    st = plan.stats()
    N = g.ntx * g.ntz
    cells = 2 * g.nt * N * nshots   # per operator
    P = 2 * g.pmlc * g.ntz + 2 * g.pmlc * g.ntx
    nrec = len(case.shots[0][2])
    # per shot: background and scattered wavefields, m read, the gather; per launch: the model (K, Bx, Bz)
    shot_step_bytes = 48 * N + 32 * P + 4 * nint + 4 * nrec
    launches_per_step = f_launches / g.nt
    bytes_per_launch = (shot_step_bytes * nshots + 12 * N * launches_per_step) / launches_per_step
    launch_ms = f_ms / a.steps / max(1, f_launches)
    achieved = bytes_per_launch / (launch_ms * 1e-3) / 1e9
    traffic = None
    # the counter figures are per shot-step of the one-step pair, or of the two-step pair
    tfile = os.path.join(ROOT, "profiles", "born2_c5_traffic.json" if f_spl == 2 else "born_c5_traffic.json")
    if os.path.exists(tfile) and (g.nx, g.nz, g.pml, g.Lc) == (1690, 1001, 60, 4):
        with open(tfile) as f:
            traffic = json.load(f)["l2_miss_bytes_per_shot_step"] * nshots / launches_per_step
    grad_finite = bool(torch.isfinite(d_g).all().item())

    cpu = None
    if not a.no_cpu:
        sys.path.insert(0, os.path.join(ROOT, "tests"))
        import _oracle as O
        from helpers import oracle_problem

        threads, cpus = host_cpu()
        O.set_threads(threads)
        nts = min(g.nt, a.cpu_nt_born)
        sub = configs.Case(case.name, swd.Grid(g.nx, g.nz, g.pml, g.Lc, g.dx, g.dz, g.dt, nts), case.vp,
                               case.rho, case.coefs, case.wavelet[:nts], case.shots[:1], case.f0)
        Pq = oracle_problem(sub, 0)
        m0 = d_m[:nint].cpu().numpy()
        r0 = rng.standard_normal((nrec, nts)).astype(np.float32)
        t0 = time.perf_counter()
        Pq.born(m0)
        Pq.born_adjoint(r0)
        dt = time.perf_counter() - t0
        cpu = {"value": 2 * 2 * nts * N / dt / 1e9, "unit": "Gcell-updates/s", "cores": threads, "kind": "port",
               "sample": f"oracle fp32 host loops (OpenMP) on 1 c5 shot with nt={nts}: Born forward + adjoint "
                         f"(the oracle keeps every step's background, no recompute), {dt:.1f}s; {threads} threads "
                         f"on {cpus}"}
        log(f"bench: cpu baseline {cpu['value']:.3f} Gcell/s in {dt:.1f}s on {threads} threads")

    out = {
        "metric": "Born forward + adjoint Gcell-updates/s (c5 grid, sqrt(nt) checkpoints)",
        "value": 2 * cells * a.steps / (t_f + t_a) / 1e9,
        "unit": "Gcell-updates/s",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": (t_f + t_a) * 1e3 / a.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (seeded layered c5 model, Gaussian m = dK/K, random residual d)",
        "config": {"workload": f"c5 batched Born forward + exact adjoint: 1001x1690 interior (1129x1818 padded), "
                               f"nt {g.nt}, Lc 4, {nrec} receivers, {nshots} shots",
                   "shots_per_step": nshots, "born_fwd_gcell_per_s": cells * a.steps / t_f / 1e9,
                   "born_adj_gcell_per_s": cells * a.steps / t_a / 1e9,
                   "adj_recompute_cells_per_step": g.nt * N * nshots,
                   "born_fwd_ms": t_f * 1e3 / a.steps, "born_adj_ms": t_a * 1e3 / a.steps,
                   "ckpt_GB": st["ckpt_bytes"] / 1e9, "full_storage_GB": st["full_bytes"] / 1e9,
                   "waves": st["waves"], "grad_finite": grad_finite},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": traffic,
                     "copy_GBs": copy_gbs, "frac_of_copy": achieved / copy_gbs,
                     "kernel": ("k_fstep2 Born pair (fused Born forward, two steps per launch: the background's "
                                "two levels, then the scattered field's, of each shot of a group in one workgroup "
                                "per tile)" if f_spl == 2 else
                                "k_fstep Born pair (fused Born forward step: background and scattered field of "
                                "each shot of a group in one workgroup per tile)"),
                     "steps_per_launch": f_spl,
                     "bytes_per_launch": bytes_per_launch, "avg_launch_ms": launch_ms,
                     "launches_per_step": launches_per_step,
                     "timing": "HIP events on the plan's stream around the Born forward loop (joining the other "
                               "streams) / launches: the interval between launch completions",
                     "traffic_source": "rocprofv3 FETCH_SIZE x2 + WRITE_SIZE per shot-step, "
                                       "profiles/born_c5_8shots.md" if traffic else None},
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)


def bench_multi(a):
    """Forward modelling of the c4 batch: 64 surface shots on the c1 grid with
    point scatterers plus 32 single-shot grids of log-uniform sizes from 128x128
    to 502x845 (all Lc 4, nt 1000), every grid's shots in one fused launch per
    time step (swd_multi_forward), against one call per grid."""
    import torch
    import swd
    from swd import configs

    cases = configs.c4(nt=a.nt_multi)
    log(f"bench: c4 multi-grid forward, {len(cases)} grids, {sum(len(c.shots) for c in cases)} shots, "
        f"nt {a.nt_multi}")
    plans = [swd.Plan(c.grid, c.coefs, c.vp, c.rho, device=0) for c in cases]
    plans[0].set_steps_per_launch(a.steps_per_launch)
    sets = [swd.ShotSet(c.shots) for c in cases]
    wavs = [c.wavelet for c in cases]
    cells = sum(c.grid.ntx * c.grid.ntz * c.grid.nt * len(c.shots) for c in cases)
    bytes_per_step = 0
    for c in cases:
        g = c.grid
        P = 2 * g.pmlc * g.ntz + 2 * g.pmlc * g.ntx
        # per shot the state, CPML memories and gather; the grid's model (K, Bx, Bz) once a step
        bytes_per_step += sum(24 * g.ntx * g.ntz + 16 * P + 4 * len(sh[2]) for sh in c.shots) + 12 * g.ntx * g.ntz
    dev = torch.device("cuda:0")
    copy_gbs = copy_bandwidth(torch, dev)
    for i in range(a.warmup):
        swd.multi_forward(plans, sets, wavs)
        log(f"bench: warmup {i + 1}/{a.warmup} done")
    t_multi = dev_ms = 0.0
    for i in range(a.steps):
        t0 = time.perf_counter()
        outs = swd.multi_forward(plans, sets, wavs)
        t_multi += time.perf_counter() - t0
        dev_ms += plans[0].stats()["fwd_ms"]
    # one call per grid (each plan's own fused loop, its tile tuned on its first call)
    for pl, S, w in zip(plans, sets, wavs):
        pl.forward(S, w)
    t_each = 0.0
    rel = 0.0
    for i in range(a.steps):
        t0 = time.perf_counter()
        each = [pl.forward(S, w) for pl, S, w in zip(plans, sets, wavs)]
        t_each += time.perf_counter() - t0
    for x, y in zip(outs, each):
        rel = max(rel, float(np.linalg.norm(x - y) / max(1e-30, np.linalg.norm(y))))
    launch_ms = dev_ms / a.steps / cases[0].grid.nt
    achieved = bytes_per_step / (launch_ms * 1e-3) / 1e9

    cpu = None
    if not a.no_cpu:
        sys.path.insert(0, os.path.join(ROOT, "tests"))
        import _oracle as O
        from helpers import oracle_problem

        threads, cpus = host_cpu()
        O.set_threads(threads)
        t0 = time.perf_counter()
        ccells = 0
        for k in range(8):
            oracle_problem(cases[0], k).forward()
            ccells += cases[0].grid.ntx * cases[0].grid.ntz * cases[0].grid.nt
        for c in cases[1:]:
            oracle_problem(c, 0).forward()
            ccells += c.grid.ntx * c.grid.ntz * c.grid.nt
        dt = time.perf_counter() - t0
        cpu = {"value": ccells / dt / 1e9, "unit": "Gcell-updates/s", "cores": threads, "kind": "port",
               "sample": f"oracle fp32 host loops (OpenMP) on 8 of the c1-grid shots and the 32 uneven grids, "
                         f"nt {cases[0].grid.nt}, {dt:.1f}s; {threads} threads on {cpus}"}
        log(f"bench: cpu baseline {cpu['value']:.3f} Gcell/s in {dt:.1f}s on {threads} threads")

    out = {
        "metric": "multi-grid forward Gcell-updates/s (c4: uneven grid sizes in one launch per step)",
        "value": cells * a.steps / t_multi / 1e9,
        "unit": "Gcell-updates/s",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": t_multi * 1e3 / a.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (c4: c1 layered model with point scatterers, 32 seeded uneven grids)",
        "config": {"workload": f"c4 multi-grid forward: 64 shots on the c1 grid (200x300) + 32 single-shot grids "
                               f"from 128x128 to 502x845, Lc 4, nt {cases[0].grid.nt}",
                   "grids": len(cases), "shots": sum(len(c.shots) for c in cases),
                   "one_call_per_grid_gcell_per_s": cells * a.steps / t_each / 1e9,
                   "speedup_vs_one_call_per_grid": t_each / t_multi,
                   "gather_rel_l2_vs_one_call_per_grid": rel,
                   "timing_note": "wall clock per call, the gathers' copy to the host included"},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": None,
                     "copy_GBs": copy_gbs, "frac_of_copy": achieved / copy_gbs,
                     "kernel": ("k_fstep2_multi (two forward steps of every grid's shots in one launch, 32 x "
                                "32-row tiles)" if plans[0].stats()["steps_per_launch"] == 2 else
                                "k_fstep_multi (forward step of every grid's shots in one launch)"),
                     "bytes_per_launch": bytes_per_step, "avg_launch_ms": launch_ms,
                     "timing": "HIP events on the first plan's stream around the time loop / nt launches"},
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)


def hilbert_rows(x, chunk=2048):
    """Hilbert transform along time of every row of x [rows][nt], the
    definition of swd_hilbert1d (FK:1257-1324: nt padded to a power of two,
    -i on bins 1..NZ/2-1, +i from NZ/2, DC zeroed) with numpy's fp64 FFT in
    chunks of rows: input preparation of the c5 workload, outside the timing."""
    nt = x.shape[1]
    nz = 1 << (nt - 1).bit_length()
    mult = np.zeros(nz, np.complex128)
    mult[1:nz // 2] = -1j
    mult[nz // 2:] = 1j
    out = np.empty_like(x)
    for r0 in range(0, x.shape[0], chunk):
        X = np.fft.fft(x[r0:r0 + chunk].astype(np.float64), n=nz, axis=1)
        out[r0:r0 + chunk] = np.fft.ifft(X * mult, axis=1).real[:, :nt].astype(np.float32)
    return out


def bench_c5(a):
    """The large batched RTM + adjoint configuration (c5): 1001 x 1690 interior,
    5 m, dt 0.4 ms, nt 8000, 60-cell CPML, 128 surface shots, 801 borehole
    receivers.  One step = the batched RTM of all shots, their up/down-separated
    RTM (du, ud at nfftz 1024 over the interior rows) and one data-domain L2
    misfit gradient L^T (L m - d) through the batched Born forward and adjoint
    with sqrt(nt) checkpoints; shots run in memory-sized waves and cache-sized
    groups.  Inputs (observed data, adjoint sources and their Hilbert
    transforms, m, d) are resident in HBM before the timed region."""
    import torch
    import swd
    from swd import configs

    nshots, nt = a.c5_shots, a.nt_c5
    case = configs.c5(nt=nt, nshots=128, shot_index=[int(v) for v in np.linspace(0, 127, nshots)])
    g = case.grid
    log(f"bench: c5 RTM + up/down RTM + Born misfit gradient, {nshots} shots, nt {nt}")
    plan = swd.Plan(g, case.coefs, case.vp, case.rho, device=0, tile=a.tile)
    plan.set_streams(a.streams)
    shots = swd.ShotSet(case.shots)
    dev = torch.device("cuda:0")
    ntp, nint = g.ntx * g.ntz, g.nx * g.nz
    t0 = time.perf_counter()
    plan_true = swd.Plan(g, case.coefs, case.vp_true, case.rho, device=0)
    obs = plan_true.forward(shots, case.wavelet)
    plan_true.close()
    syn = plan.forward(shots, case.wavelet)
    log(f"bench: observed and modelled data {obs.shape} in {time.perf_counter() - t0:.1f}s")
    adj = swd.adjoint_source(obs, g.dt)
    hadj = hilbert_rows(adj)
    h0 = swd.hilbert1d(adj[0])
    log(f"bench: Hilbert transforms of {adj.shape[0]} traces, first trace vs swd_hilbert1d "
        f"{float(np.abs(hadj[0] - h0).max() / max(1e-30, np.abs(h0).max())):.1e}")
    hw = swd.hilbert1d(case.wavelet)
    nfftz = max(128, swd.nfft(g.nz))
    kz, _ = swd.kz_taper(nfftz, g.dz)
    m = (0.01 * configs.bandlimited(g.nx, g.nz, g.dx, g.dz, 10.0, 30.0, 1500.0, 3)).astype(np.float32).ravel()
    d_adj, d_hadj = torch.from_numpy(adj).to(dev), torch.from_numpy(hadj).to(dev)
    d_scat = torch.from_numpy(obs).to(dev) - torch.from_numpy(syn).to(dev)   # reflection data d
    del obs, syn, hadj
    d_m = torch.from_numpy(m).to(dev).repeat(nshots)
    d_lm = torch.zeros((shots.ntr, nt), dtype=torch.float32, device=dev)
    d_img = torch.zeros((3, nshots, ntp), dtype=torch.float32, device=dev)
    d_ud = torch.zeros((2, nshots, ntp), dtype=torch.float32, device=dev)
    d_grad = torch.zeros(nint, dtype=torch.float32, device=dev)
    torch.cuda.synchronize()
    copy_gbs = copy_bandwidth(torch, dev)

    def rtm():
        plan.rtm_dev(shots, case.wavelet, d_adj.data_ptr(), d_img[0].data_ptr(), d_img[1].data_ptr(),
                     d_img[2].data_ptr())

    def updown():
        plan.rtm_updown_dev(shots, case.wavelet, hw, d_adj.data_ptr(), d_hadj.data_ptr(), kz,
                            {"du": d_ud[0].data_ptr(), "ud": d_ud[1].data_ptr()})

    def gradient():
        plan.born_dev(shots, case.wavelet, d_m.data_ptr(), d_lm.data_ptr())
        d_lm.sub_(d_scat)   # residual L m - d
        plan.born_adjoint_dev(shots, case.wavelet, d_lm.data_ptr(), d_grad.data_ptr())

    phases = (("rtm", rtm), ("updown", updown), ("gradient", gradient))
    for i in range(a.warmup):
        for name, f in phases:
            f()
            torch.cuda.synchronize()
            log(f"bench: warmup {i + 1}/{a.warmup} {name} done")
    t = {name: 0.0 for name, _ in phases}
    stats = {}
    for i in range(a.steps):
        for name, f in phases:
            t1 = time.perf_counter()
            f()
            torch.cuda.synchronize()
            t[name] += time.perf_counter() - t1
            stats[name] = plan.stats()
            log(f"bench: step {i + 1}/{a.steps} {name} {t[name] / (i + 1):.1f}s")
    total = sum(t.values())
    N = g.ntx * g.ntz
    P = 2 * g.pmlc * g.ntz + 2 * g.pmlc * g.ntx
    nrec = len(case.shots[0][2])
    nframe = 2 * g.Lc * (g.nx + g.nz - 2 * g.Lc)
    bf, bb, mf, mb = rtm_bytes(g, nrec, nframe)
    cells = {"rtm": rtm_cells(g) * nshots, "updown": 2 * rtm_cells(g) * nshots, "gradient": 4 * nt * N * nshots}
    # algorithmic bytes per shot of each phase, with the model read once a shot and step (the cache budget
    # gives one-shot groups at c5; a launch's wavefields share the model): the RTM forward and backward; the
    # up/down RTM twice both plus the split (four field columns and two images per interior row); the Born
    # forward (background and scattered field, m read) and the adjoint (adjoint and recomputed background
    # state, the background sums read, the gradient read and written, the data read)
    ud_split = 16 * g.nx * g.nz + 16 * g.nx * g.nz
    byts = {"rtm": nt * (bf + mf) + (nt - 1) * (bb + mb),
            "updown": nt * (2 * bf + mf) + (nt - 1) * (2 * bb + mb) + (nt - 1) * ud_split,
            "gradient": nt * (48 * N + 32 * P + 4 * nint + 4 * nrec + 12 * N) +
                        nt * (2 * (24 * N + 16 * P) + 12 * nint + 4 * nrec + 12 * N)}
    per_phase = {}
    for name, _ in phases:
        sec = t[name] / a.steps
        ach = byts[name] * nshots / sec / 1e9
        per_phase[name] = {"s_per_step": sec, "gcell_per_s": cells[name] / sec / 1e9, "shots_per_s": nshots / sec,
                           "algorithmic_GBs": ach, "frac_of_peak": ach / HBM_PEAK_GBS,
                           "waves": stats[name]["waves"]}
    per_phase["gradient"]["ckpt_GB"] = stats["gradient"]["ckpt_bytes"] / 1e9
    per_phase["gradient"]["full_storage_GB"] = stats["gradient"]["full_bytes"] / 1e9
    # the dominant kernel (the paired RTM backward step, one-shot groups at c5) alone on one stream
    cfg_b = launch_config(stats["rtm"])
    # its counter traffic per shot-step, measured for two steps per launch on 32 x 32-row tiles
    c5_traffic = None
    tfile = os.path.join(ROOT, "profiles", "rtm_fused2_c5_traffic.json")
    if tuple(cfg_b) == (2, 32, 32, 4) and os.path.exists(tfile) and (g.nx, g.nz, g.Lc) == (1690, 1001, 4):
        with open(tfile) as f:
            c5_traffic = json.load(f)["l2_miss_bytes_per_shot_step"]
    nt_iso = min(a.nt_iso, a.nt_c5)
    iso_ms, iso_lps = backward_launch_ms(swd, torch, case, adj, 1, 1, 1, nt_iso, a.tile, cfg_b)
    iso_bytes = (bb + mb * iso_lps) / iso_lps   # a launch advances 1 / iso_lps steps; the model once
    iso = iso_bytes / (iso_ms * 1e-3) / 1e9
    log(f"bench: isolated 1-shot backward launch {iso_ms * 1e3:.1f} us, {iso:.0f} GB/s")
    # HBM-resident: a 4-shot group alone on one stream, a per-step working set (119 MB a shot) about twice the
    # 256 MB last-level cache
    res_n = min(4, nshots)
    res_ms, res_lps = backward_launch_ms(swd, torch, case, adj, res_n, res_n, 1, nt_iso, a.tile, cfg_b)
    res_set = 4.0 * (9 * g.nx * g.nz + 6 * g.ntx * g.ntz + 6 * (g.ntx * g.ntz - g.nx * g.nz)) * res_n
    res_bytes = (bb * res_n + mb * res_lps) / res_lps
    res = res_bytes / (res_ms * 1e-3) / 1e9
    log(f"bench: {res_n}-shot backward launch ({res_set / 2**20:.0f} MiB working set) {res_ms * 1e3:.1f} us, "
        f"{res:.0f} GB/s")
    plain = None
    if not a.no_plain:
        one = swd.ShotSet(case.shots[:1])
        plan.set_path(swd.SWD_PATH_PLAIN)
        d1 = torch.from_numpy(adj[:nrec].copy()).to(dev)
        im1 = torch.zeros((3, ntp), dtype=torch.float32, device=dev)
        torch.cuda.synchronize()
        t1 = time.perf_counter()
        plan.rtm_dev(one, case.wavelet, d1.data_ptr(), im1[0].data_ptr(), im1[1].data_ptr(), im1[2].data_ptr())
        torch.cuda.synchronize()
        tp = time.perf_counter() - t1
        plan.set_path(swd.SWD_PATH_FUSED)
        rel = float(np.linalg.norm(im1[0].cpu().numpy() - d_img[0, 0].cpu().numpy()) /
                    max(1e-30, float(np.linalg.norm(d_img[0, 0].cpu().numpy()))))
        plain = {"shots_per_s": 1.0 / tp, "image_rel_l2_vs_batched": rel}
        log(f"bench: plain path one c5 shot {tp:.2f}s, image rel diff vs batched {rel:.2e}")
    cpu = None
    if not a.no_cpu:
        sys.path.insert(0, os.path.join(ROOT, "tests"))
        import _oracle as O
        from helpers import oracle_problem

        L, flags = native_oracle()
        threads, cpus = host_cpu()
        L.or_set_threads(threads)
        nts = min(nt, a.cpu_nt_c5)
        sub = configs.Case(case.name, swd.Grid(g.nx, g.nz, g.pml, g.Lc, g.dx, g.dz, g.dt, nts), case.vp, case.rho,
                           case.coefs, case.wavelet[:nts], case.shots[:1], case.f0)
        Pq = oracle_problem(sub, 0)
        Pq.L = L
        t1 = time.perf_counter()
        d, st, fs = Pq.forward(strips=True, final=True)
        Pq.rtm_backward(fs, st, O.adjoint_source(d, g.dt))
        dt = time.perf_counter() - t1
        rate = rtm_cells(sub.grid) / dt / 1e9
        cpu = {"value": rate, "unit": "Gcell-updates/s", "cores": threads, "kind": "port",
               "sample": f"oracle fp32 host loops (OpenMP over x, {flags}) on 1 c5 shot with nt={nts} (RTM: forward "
                         f"with strips, reconstruction, receiver propagation, imaging), {dt:.1f}s; the full c5 step "
                         f"would take {rate and sum(cells.values()) / rate / 1e9 / 3600:.0f} h at that rate; "
                         f"{threads} threads on {cpus}"}
        log(f"bench: cpu baseline {rate:.3f} Gcell/s in {dt:.1f}s")
    out = {
        "metric": "c5 batched RTM + up/down RTM + Born misfit gradient, Gcell-updates/s",
        "value": sum(cells.values()) * a.steps / total / 1e9,
        "unit": "Gcell-updates/s",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": total * 1e3 / a.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (seeded layered c5 model with 8 dipping reflectors, band-limited m, c5 geometry)",
        "config": {"workload": f"c5: 1001x1690 interior (1129x1818 padded), 5 m, dt 0.4 ms, nt {nt}, Lc 4, {nshots} "
                               f"surface shots, {nrec} borehole receivers; RTM, up/down RTM (nfftz {nfftz}, interior "
                               "rows), Born forward + adjoint L2 misfit gradient with sqrt(nt) checkpoints",
                   "shots_per_step": nshots, "phases": per_phase,
                   "shots_per_s_rtm": nshots / (t["rtm"] / a.steps),
                   "plain_path": plain,
                   "speedup_vs_plain_shots_per_s": (nshots / (t["rtm"] / a.steps) / plain["shots_per_s"]) if plain
                   else None,
                   "grad_finite": bool(torch.isfinite(d_grad).all().item())},
        "roofline": {"bound": "hbm", "achieved": iso, "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": iso / HBM_PEAK_GBS,
                     "traffic": c5_traffic["backward"] / iso_lps if c5_traffic else None,
                     "traffic_note": ("FETCH_SIZE x2 + WRITE_SIZE per launch (profiles/rtm_fused2_c5_traffic.json): "
                                      "L2-miss bytes, which include last-level-cache hits, so an upper bound on HBM "
                                      "bytes") if c5_traffic else None,
                     "copy_GBs": copy_gbs, "frac_of_copy": iso / copy_gbs,
                     "kernel": (f"{'k_fstep2' if cfg_b[0] == 2 else 'k_fstep'} paired RTM backward "
                                f"({cfg_b[0]} step(s) per launch, {cfg_b[1]} x {cfg_b[2]}-row tiles), one-shot group "
                                "alone on one stream"),
                     "steps_per_launch": 1.0 / iso_lps, "rtm_backward_launch": dict(zip(
                         ("steps_per_launch", "tile_x", "tile_rows", "waves"), cfg_b)),
                     "bytes_per_launch": iso_bytes, "avg_launch_ms": iso_ms,
                     "timing": f"HIP events on the plan's stream around {nt_iso - 1} graph-replayed launches",
                     "hbm_resident": {"achieved": res, "frac": res / HBM_PEAK_GBS, "shots_per_launch": res_n,
                                      "working_set_MiB": res_set / 2**20, "bytes_per_launch": res_bytes,
                                      "traffic": (c5_traffic["backward_4shot_group"] * res_n / res_lps
                                                  if c5_traffic and res_n == 4 else None),
                                      "avg_launch_ms": res_ms,
                                      "timing": f"{res_n}-shot group alone on one stream, {nt_iso - 1} steps: a "
                                                "per-step working set about twice the 256 MB last-level cache"}},
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--batch", type=int, default=16, help="shots per RTM step")
    ap.add_argument("--nt", type=int, default=4000)
    ap.add_argument("--no-plain", action="store_true")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--cpu-nt", type=int, default=2000, help="time steps of the CPU baseline's c2 shot")
    ap.add_argument("--nt-iso", type=int, default=400,
                    help="time steps of the isolated and HBM-resident backward-launch measurements")
    ap.add_argument("--tile", type=int, default=0, help="fused tile width (0 = tuned on the first steps)")
    ap.add_argument("--steps-per-launch", type=int, default=0, choices=(0, 1, 2),
                    help="time steps per fused launch of the RTM loops (0 = the tuner picks one or two)")
    ap.add_argument("--path", choices=("fused", "batched"), default="fused")
    ap.add_argument("--groups", type=int, nargs=2, default=(0, 0), metavar=("FWD", "BWD"),
                    help="shots per fused RTM forward/backward loop (0 = sized to the cache)")
    ap.add_argument("--calibrate-copy", action="store_true",
                    help="also copy 256 MiB device-to-device once (reference for HBM byte counters)")
    ap.add_argument("--streams", type=int, default=0, choices=(0, 1, 2, 3, 4),
                    help="RTM shot groups in flight at once (swd_plan_set_streams); 0: 3 for --workload updown, "
                         "else 2 (the fastest measured)")
    ap.add_argument("--born-shots", type=int, default=8, help="shots of --workload born")
    ap.add_argument("--nt-born", type=int, default=8000, help="time steps of --workload born (c5: 8000)")
    ap.add_argument("--cpu-nt-born", type=int, default=1000, help="time steps of the born CPU sample")
    ap.add_argument("--nt-multi", type=int, default=1000, help="time steps of --workload multi (c4: 1000)")
    ap.add_argument("--c5-shots", type=int, default=128, help="shots of --workload c5 (c5: 128)")
    ap.add_argument("--nt-c5", type=int, default=8000, help="time steps of --workload c5 (c5: 8000)")
    ap.add_argument("--cpu-nt-c5", type=int, default=400, help="time steps of the c5 CPU sample")
    ap.add_argument("--workload", choices=("rtm", "decon", "updown", "cig", "afor", "born", "multi", "c5"),
                    default="rtm",
                    help="rtm: the batched RTM hot path (default); decon: the PSF deconvolution; "
                         "updown: the up/down-separated RTM; cig: RTM with angle gathers; afor: AFor forward "
                         "modelling with the kz up/down separation; born: Born forward + adjoint on the c5 grid; multi: the c4 uneven-grid forward; "
                         "c5: the large batched RTM + up/down + Born misfit gradient")
    ap.add_argument("--thetan", type=int, default=10, help="angle bins of --workload cig (RC:235: 10)")
    ap.add_argument("--nfftz", type=int, default=0,
                    help="kz transform length of --workload updown (0: nz padded to a power of two, at least 128)")
    a = ap.parse_args()
    if a.streams == 0:
        a.streams = 3 if a.workload == "updown" else 2
    if a.gpus != 1:
        print(f"bench.py runs on one GPU only (asked for --gpus {a.gpus}); the multi-GPU shot sharding is "
              "design-only in this build", file=sys.stderr)
        sys.exit(2)

    if a.workload == "decon":
        bench_decon(a)
        return
    if a.workload == "updown":
        bench_updown(a)
        return
    if a.workload == "cig":
        bench_cig(a)
        return
    if a.workload == "afor":
        bench_afor(a)
        return
    if a.workload == "born":
        bench_born(a)
        return
    if a.workload == "multi":
        bench_multi(a)
        return
    if a.workload == "c5":
        bench_c5(a)
        return

    import torch
    import swd

    log(f"bench: c2 RTM, batch {a.batch} shots, nt {a.nt}")
    case = build_case(a.batch, a.nt)
    g = case.grid
    path = swd.SWD_PATH_FUSED if a.path == "fused" else swd.SWD_PATH_BATCHED
    plan = swd.Plan(g, case.coefs, case.vp, case.rho, device=0, path=path, tile=a.tile)
    plan.set_groups(*a.groups)
    plan.set_streams(a.streams)
    plan.set_steps_per_launch(a.steps_per_launch)
    shots = swd.ShotSet(case.shots)
    # observed data on the true model, up-going integration as the adjoint source (RC:603-637)
    plan_true = swd.Plan(g, case.coefs, case.vp_true, case.rho, device=0)
    t0 = time.perf_counter()
    obs = plan_true.forward(shots, case.wavelet)
    plan_true.close()
    adj = swd.adjoint_source(obs, g.dt)
    log(f"bench: observed data {obs.shape} in {time.perf_counter() - t0:.1f}s")
    dev = torch.device("cuda:0")
    ntp = g.ntx * g.ntz
    d_adj = torch.from_numpy(adj).to(dev)
    d_img = torch.zeros((3, a.batch, ntp), dtype=torch.float32, device=dev)
    torch.cuda.synchronize()

    def step():
        plan.rtm_dev(shots, case.wavelet, d_adj.data_ptr(), d_img[0].data_ptr(), d_img[1].data_ptr(),
                     d_img[2].data_ptr())

    if a.calibrate_copy:
        src = torch.ones(64 * 1024 * 1024, dtype=torch.float32, device=dev)
        dst = torch.empty_like(src)
        torch.cuda.synchronize()
        dst.copy_(src)
        torch.cuda.synchronize()
        del src, dst
    copy_gbs = copy_bandwidth(torch, dev)
    log(f"bench: device copy bandwidth {copy_gbs:.0f} GB/s")
    for i in range(a.warmup):
        step()
        log(f"bench: warmup {i + 1}/{a.warmup} done")
    torch.cuda.synchronize()
    fwd_ms = bwd_ms = 0.0
    t0 = time.perf_counter()
    for i in range(a.steps):
        step()
        st = plan.stats()
        fwd_ms += st["fwd_ms"]
        bwd_ms += st["bwd_ms"]
        log(f"bench: step {i + 1}/{a.steps} fwd {st['fwd_ms']:.1f} ms bwd {st['bwd_ms']:.1f} ms")
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t0
    ms_per_step = elapsed * 1e3 / a.steps
    cells = rtm_cells(g) * a.batch
    gcells = cells * a.steps / elapsed / 1e9
    shots_per_s = a.batch * a.steps / elapsed
    img = d_img[0].cpu().numpy()
    finite = bool(np.isfinite(img).all())

    # the dominant kernel, the fused backward step, timed three ways with HIP events on the plan's stream
    nrec = len(case.shots[0][2])
    nframe = 2 * g.Lc * (g.nx + g.nz - 2 * g.Lc)
    bf, bb, mf, mb = rtm_bytes(g, nrec, nframe)
    st = plan.stats()
    cfg_b = launch_config(st)
    spl = st["steps_per_launch"]
    # launches per step: the groups of the batch / the steps a launch advances (a two-step loop runs its first
    # step and the remainder of its step count modulo 4 one step per launch); per launch the algorithmic bytes
    # are the per-step bytes of its shots times the steps it advances, plus the model once
    bwd_launches_per_step = st["bwd_launches"] / (g.nt - 1)
    ngroups = max(1, int(round(bwd_launches_per_step * spl)))
    group = -(-a.batch // ngroups)   # the largest backward group of the run
    # (1) overlapped: the timed region's backward phase over its launches (groups on a.streams streams)
    ov_ms = bwd_ms / a.steps / max(1, st["bwd_launches"])
    ov_bytes = (bb * a.batch + mb * bwd_launches_per_step) / bwd_launches_per_step
    ov_achieved = ov_bytes / (ov_ms * 1e-3) / 1e9
    fwd_launch_ms = fwd_ms / a.steps / max(1, st["fwd_launches"])
    fwd_launches_per_step = st["fwd_launches"] / g.nt
    fwd_achieved = (bf * a.batch + mf * fwd_launches_per_step) / fwd_launches_per_step / (fwd_launch_ms * 1e-3) / 1e9
    # (2) isolated: one group of that size alone on one stream, graph-replayed launches back to back
    nt_iso = min(a.nt_iso, a.nt)
    iso_ms, iso_lps = backward_launch_ms(swd, torch, case, adj, group, group, 1, nt_iso, a.tile, cfg_b)
    iso_bytes = (bb * group + mb * iso_lps) / iso_lps
    iso_achieved = iso_bytes / (iso_ms * 1e-3) / 1e9
    log(f"bench: isolated {group}-shot backward launch {iso_ms * 1e3:.2f} us, {iso_achieved:.0f} GB/s")
    # (3) HBM-resident: every shot of the batch in one launch on one stream, a working set twice the
    # last-level cache (32 MB a shot at c2)
    res_ms, res_lps = backward_launch_ms(swd, torch, case, adj, a.batch, a.batch, 1, nt_iso, a.tile, cfg_b)
    res_set = 4.0 * (9 * g.nx * g.nz + 6 * g.ntx * g.ntz + 6 * (g.ntx * g.ntz - g.nx * g.nz)) * a.batch
    res_bytes = (bb * a.batch + mb * res_lps) / res_lps
    res_achieved = res_bytes / (res_ms * 1e-3) / 1e9
    log(f"bench: {a.batch}-shot backward launch ({res_set / 2**20:.0f} MiB working set) {res_ms * 1e3:.1f} us, "
        f"{res_achieved:.0f} GB/s")

    plain = None
    if not a.no_plain:
        one = swd.ShotSet(case.shots[:1])
        plan.set_path(swd.SWD_PATH_PLAIN)
        rows = len(case.shots[0][2])
        d1 = torch.from_numpy(adj[:rows].copy()).to(dev)
        im1 = torch.zeros((3, ntp), dtype=torch.float32, device=dev)
        # one untimed shot (the plain path's first-call set-up), then the median of three
        runs = []
        for r in range(4):
            torch.cuda.synchronize()
            t1 = time.perf_counter()
            plan.rtm_dev(one, case.wavelet, d1.data_ptr(), im1[0].data_ptr(), im1[1].data_ptr(), im1[2].data_ptr())
            torch.cuda.synchronize()
            if r:
                runs.append(time.perf_counter() - t1)
        tp = float(np.median(runs))
        plain = {"shots_per_s": 1.0 / tp, "gcell_per_s": rtm_cells(g) / tp / 1e9,
                 "timing": "one c2 shot on the plain path: median of 3 runs after 1 untimed run"}
        rel = float(np.linalg.norm(im1[0].cpu().numpy() - img[0]) / max(1e-30, np.linalg.norm(img[0])))
        plain["image_rel_l2_vs_batched"] = rel
        plan.set_path(path)
        log(f"bench: plain path one shot {tp:.3f}s (runs {', '.join(f'{v:.3f}' for v in runs)}), "
            f"image rel diff vs batched {rel:.2e}")

    cpu = None
    if not a.no_cpu:
        cpu = cpu_baseline(case, a.cpu_nt)
        log(f"bench: cpu baseline {cpu['value']:.3f} Gcell/s ({cpu['sample']})")

    traffic = None
    tname = "rtm_fused2_c2_traffic.json" if spl == 2 else "rtm_fused_c2_traffic.json"
    tfile = os.path.join(ROOT, "profiles", tname)
    if a.path == "fused" and os.path.exists(tfile) and (g.nx, g.nz, g.pml, g.Lc) == (845, 502, 40, 4):
        with open(tfile) as f:
            t = json.load(f)
        traffic = t["l2_miss_bytes_per_shot_step"]["backward"] * group / iso_lps   # per launch
    if spl == 2:
        kname = (f"k_fstep2 paired backward (reconstruction + receiver + imaging), two steps per launch, "
                 f"{cfg_b[1]} x {cfg_b[2]}-row tiles, {abs(cfg_b[3])} wavefronts"
                 f"{' in split launches' if cfg_b[3] < 0 else ''}, {group}-shot group alone on one stream")
    else:
        kname = (f"k_fstep paired backward step (reconstruction + receiver + imaging), {group}-shot "
                 "group alone on one stream")

    out = {
        "metric": "Gcell-updates/s and batched shots/s vs plain GPU path at <=1e-5 rel L2",
        "value": gcells,
        "unit": "Gcell-updates/s",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (seeded layered model with dipping reflectors, c2 geometry)",
        "config": {"workload": "c2 batched VSP RTM: 502x845 interior (590x933 padded), nt 4000, Lc 4, "
                               "401 borehole receivers, forward+strips+reconstruction+imaging",
                   "shots_per_step": a.batch, "nt": g.nt, "padded_grid": [g.ntx, g.ntz],
                   "shots_per_s": shots_per_s,
                   "plain_path": plain,
                   "speedup_vs_plain_shots_per_s": (shots_per_s / plain["shots_per_s"]) if plain else None,
                   "fwd_ms_per_step": fwd_ms / a.steps, "bwd_ms_per_step": bwd_ms / a.steps,
                   "fwd_achieved_GBs": fwd_achieved, "image_finite": finite,
                   "backward_launch": dict(zip(("steps_per_launch", "tile_x", "tile_rows", "waves"), cfg_b)),
                   "forward_launch": {"steps_per_launch": st["fwd_steps_per_launch"], "tile_x": st["fwd_tile_x"],
                                      "tile_rows": st["fwd_tile_rows"], "waves": st["fwd_tile_group_waves"]}},
        "roofline": {"bound": "hbm", "achieved": iso_achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": iso_achieved / HBM_PEAK_GBS, "traffic": traffic,
                     "traffic_note": f"FETCH_SIZE x2 + WRITE_SIZE per launch (profiles/{tname}): "
                                     "L2-miss bytes, which include last-level-cache hits, so an upper bound on "
                                     "HBM bytes" if traffic else None,
                     "copy_GBs": copy_gbs, "frac_of_copy": iso_achieved / copy_gbs,
                     "kernel": kname if a.path == "fused" else "RTM backward step (4 sub-step kernels)",
                     "steps_per_launch": 1.0 / iso_lps,
                     "bytes_note": "algorithmic bytes of a shot-step times the steps a launch advances, the "
                                   "model once per launch (a two-step launch reads and writes the state once "
                                   "for its two steps)",
                     "bytes_per_launch": iso_bytes, "avg_launch_ms": iso_ms,
                     "timing": f"HIP events on the plan's stream around {nt_iso - 1} graph-replayed launches of one "
                               f"{group}-shot group with no other work in flight: launch duration plus launch gap",
                     "overlapped": {"achieved": ov_achieved, "frac": ov_achieved / HBM_PEAK_GBS,
                                    "bytes_per_launch": ov_bytes, "avg_launch_ms": ov_ms,
                                    "timing": f"the timed region: backward phase (HIP events on the plan's stream, "
                                              f"joining the others) over its launches, groups on {a.streams} "
                                              "streams overlapping: the interval between launch completions; "
                                              "the groups' working set is served largely by the last-level cache"},
                     "hbm_resident": {"achieved": res_achieved, "frac": res_achieved / HBM_PEAK_GBS,
                                      "shots_per_launch": a.batch, "working_set_MiB": res_set / 2**20,
                                      "bytes_per_launch": res_bytes, "avg_launch_ms": res_ms,
                                      "timing": f"{a.batch}-shot group alone on one stream, {nt_iso - 1} "
                                                "launches: a per-step working set twice the 256 MB last-level cache"},
                     "forward_overlapped_GBs": fwd_achieved},
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
