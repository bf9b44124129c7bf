"""Born operator L and its exact discrete adjoint L^T behind torch.autograd.

Workload c3: the c1 grid and model (200 x 300 interior, 10 m, 20-cell CPML,
Lc 4, 91 borehole receivers) with nt = 4000; m = dK/K a band-limited random
field (seed 3), data probe seed 4, finite-difference directions seeds 5-9.
"""
import numpy as np
import pytest
import torch

import swd
from swd import configs
from swd.autograd import born
from helpers import oracle_problem, rel_l2, report

pytestmark = pytest.mark.gpu

DEV = "cuda:0"


@pytest.fixture(scope="module")
def c3():
    print("case c3 (c1 grid, nt 4000)", flush=True)
    return configs.c3()


@pytest.fixture(scope="module")
def plan(c3):
    p = swd.Plan(c3.grid, c3.coefs, c3.vp, c3.rho, device=0)
    yield p
    p.close()


def test_born_autograd_dot_product(c3, plan):
    """<L m, d> = <m, L^T d> to 1e-5, with L^T reached through autograd."""
    m = torch.tensor(c3.born_m.ravel(), device=DEV, requires_grad=True)
    d = born(m, plan, c3.shots, c3.wavelet)
    r = torch.from_numpy(c3.born_probe).to(DEV)
    (d * r).sum().backward()
    lhs = float(np.dot(d.detach().cpu().numpy().ravel().astype(np.float64), c3.born_probe.ravel()))
    rhs = float(np.dot(c3.born_m.ravel().astype(np.float64), m.grad.cpu().numpy().astype(np.float64)))
    rel = abs(lhs - rhs) / abs(lhs)
    print(f"dot test <Lm,d> {lhs:.9e} <m,L^T d> {rhs:.9e} rel {rel:.2e}", flush=True)
    assert rel <= 1e-5


def test_born_finite_differences(c3, plan):
    """Central differences of the nonlinear forward along 5 seeded directions
    (K -> K(1 +- h dm), h = 1e-3) against L dm.  Tolerance 1e-3: rounding
    vp*sqrt(1+h dm) to float perturbs K by ~1e-7 against h = 1e-3, which the
    fp32 and fp64 oracles both show as 1e-4 .. 4e-4."""
    h = 1e-3
    fwd = swd.Plan(c3.grid, c3.coefs, c3.vp, c3.rho, device=0)
    try:
        for k, dm in enumerate(c3.born_dirs):
            fwd.set_model(configs.perturbed_vp(c3, h * dm), c3.rho)
            dp = fwd.forward(c3.shots, c3.wavelet)
            fwd.set_model(configs.perturbed_vp(c3, -h * dm), c3.rho)
            dn = fwd.forward(c3.shots, c3.wavelet)
            fd = (dp.astype(np.float64) - dn) / (2 * h)
            with torch.no_grad():
                L = born(torch.tensor(dm.ravel(), device=DEV), plan, c3.shots, c3.wavelet).cpu().numpy()
            e = rel_l2(fd, L)
            print(f"direction seed {5 + k}: relL2(FD, L dm) {e:.3e}", flush=True)
            assert e <= 1e-3
    finally:
        fwd.close()


def test_born_and_adjoint_vs_oracle(c3, plan):
    """GPU L m and L^T d against the fp32 oracle and the fp64 oracle, each <= 1e-5."""
    P = oracle_problem(c3)
    m = c3.born_m.ravel()
    d = plan.born(c3.shots, c3.wavelet, m)
    e32, e64 = report("c3 born L m", d, P.born(m), P.born(m, np.float64))
    assert e32 <= 1e-5 and e64 <= 1e-5
    g = plan.born_adjoint(c3.shots, c3.wavelet, c3.born_probe)
    e32, e64 = report("c3 born adjoint L^T d", g, P.born_adjoint(c3.born_probe),
                      P.born_adjoint(c3.born_probe, np.float64))
    assert e32 <= 1e-5 and e64 <= 1e-5


def test_born_checkpoint_memory_and_invariance(c3, plan):
    """sqrt(nt) checkpointing (64 segments of 63 steps at nt 4000) against full
    storage of the background sums; the result does not depend on the segment
    length (the recomputation repeats the same arithmetic)."""
    g0 = plan.born_adjoint(c3.shots, c3.wavelet, c3.born_probe)
    st = plan.stats()
    print(f"checkpointed adjoint: {st['ckpt_bytes'] / 2**20:.1f} MiB (snapshots + one segment), "
          f"full storage {st['full_bytes'] / 2**20:.1f} MiB, plan workspace {st['ws_bytes'] / 2**20:.1f} MiB",
          flush=True)
    assert st["ckpt_bytes"] * 4 < st["full_bytes"]
    g8 = plan.born_adjoint(c3.shots, c3.wavelet, c3.born_probe, nsnap=8)
    assert np.array_equal(g0, g8)


def test_born_autograd_host_and_device_tensors_agree(c3, plan):
    m = torch.tensor(c3.born_m.ravel(), requires_grad=True)
    d = born(m, plan, c3.shots, c3.wavelet)
    d.backward(torch.from_numpy(c3.born_probe))
    md = torch.tensor(c3.born_m.ravel(), device=DEV, requires_grad=True)
    dd = born(md, plan, c3.shots, c3.wavelet)
    dd.backward(torch.from_numpy(c3.born_probe).to(DEV))
    assert torch.equal(d, dd.cpu())
    assert torch.equal(m.grad, md.grad.cpu())


def test_born_batch_gradient_is_sum_of_shots():
    """Two shots in one call: data are the per-shot data, the gradient is the
    shot-ordered sum of the per-shot gradients, bit for bit."""
    c = configs.small(nt=300, nshots=2, dup=True)
    print("case small, 2 shots with duplicate receivers", flush=True)
    p = swd.Plan(c.grid, c.coefs, c.vp, c.rho, device=0)
    try:
        nint = c.grid.nx * c.grid.nz
        m = configs.bandlimited(c.grid.nx, c.grid.nz, c.grid.dx, c.grid.dz, 10.0, 30.0, 1500.0, 11).ravel()
        mt = torch.tensor(m, device=DEV, requires_grad=True)
        d = born(mt, p, c.shots, c.wavelet)
        r = np.random.default_rng(12).standard_normal(tuple(d.shape)).astype(np.float32)
        d.backward(torch.from_numpy(r).to(DEV))
        n0 = c.shots[0][2].size
        d0 = p.born(c.shots[:1], c.wavelet, m)
        d1 = p.born(c.shots[1:], c.wavelet, m)
        assert np.array_equal(d.detach().cpu().numpy(), np.concatenate([d0, d1]))
        g0 = p.born_adjoint(c.shots[:1], c.wavelet, r[:n0])
        g1 = p.born_adjoint(c.shots[1:], c.wavelet, r[n0:])
        assert np.array_equal(mt.grad.cpu().numpy(), (g0 + g1).astype(np.float32))
        assert m.size == nint
    finally:
        p.close()


@pytest.mark.parametrize("Lc", [2, 4, 8])
def test_fused_born_pairs_match_batched_and_oracle(Lc):
    """The fused Born forward (background and scattered field of a shot in one
    workgroup, the Born source from the background's sums in registers), run in
    one-shot groups on two streams, gives the batched path's gathers and the
    oracle's, shot by shot, with receivers sharing nodes."""
    c = configs.small(nshots=3, dup=True, nt=260, Lc=Lc)
    print(f"case small fused born pairs Lc={Lc}, 3 shots", flush=True)
    rng = np.random.default_rng(40 + Lc)
    nint = c.grid.nx * c.grid.nz
    m = rng.standard_normal(3 * nint).astype(np.float32)
    fused = swd.Plan(c.grid, c.coefs, c.vp, c.rho, device=0)
    fused.set_groups(2, 0)
    fused.set_streams(2)
    batched = swd.Plan(c.grid, c.coefs, c.vp, c.rho, device=0, path=swd.SWD_PATH_BATCHED)
    df = fused.born(c.shots, c.wavelet, m)
    db = batched.born(c.shots, c.wavelet, m)
    assert rel_l2(df, db) <= 1e-6
    row = 0
    for k in range(3):
        od = oracle_problem(c, k).born(m[k * nint:(k + 1) * nint])
        n = od.shape[0]
        e32, _ = report(f"fused born Lc={Lc} shot {k}", df[row:row + n], od)
        assert e32 <= 1e-5
        row += n
    fused.close()
    batched.close()


@pytest.mark.parametrize("Lc", [2, 4, 8])
def test_tiled_adjoint_matches_batched_and_oracle(Lc):
    """The tiled adjoint step of the fused path (two launches a step, CPML
    memories double-buffered by step parity) gives the four-kernel step's
    gradient and the oracle's, with receivers sharing nodes and several
    checkpoint segments."""
    c = configs.small(nshots=3, dup=True, nt=260, Lc=Lc)
    print(f"case small tiled born adjoint Lc={Lc}, 3 shots", flush=True)
    rng = np.random.default_rng(60 + Lc)
    ntr = sum(sh[2].size for sh in c.shots)
    d = rng.standard_normal((ntr, c.grid.nt)).astype(np.float32)
    tiled = swd.Plan(c.grid, c.coefs, c.vp, c.rho, device=0)
    batched = swd.Plan(c.grid, c.coefs, c.vp, c.rho, device=0, path=swd.SWD_PATH_BATCHED)
    gt = tiled.born_adjoint(c.shots, c.wavelet, d, nsnap=7)
    gb = batched.born_adjoint(c.shots, c.wavelet, d, nsnap=7)
    assert rel_l2(gt, gb) <= 1e-6
    og = np.zeros_like(gt, dtype=np.float64)
    row = 0
    for k in range(3):
        n = c.shots[k][2].size
        og += oracle_problem(c, k).born_adjoint(d[row:row + n]).astype(np.float64)
        row += n
    e32, _ = report(f"tiled born adjoint Lc={Lc}", gt, og)
    assert e32 <= 1e-5
    tiled.close()
    batched.close()
