"""swd_updown_image_dev_noalloc, the step of swd_updown_image_dev with its tables
in the library's fixed per-GPU array, against swd_updown_image_dev itself.

Both run the same kernel on the same tables, so the images must be identical.
Two calls in a row on one plan, the second with fewer shots and a shorter
transform, check that the array's leftover contents from the first call do
not leak into the second; the plan's workspace stays unallocated throughout.
"""
import numpy as np
import pytest

import swd
from swd import configs

pytestmark = pytest.mark.gpu
NAMES = ("du", "ud", "uu", "dd")


def _device_fields(torch, case, nshots, seed):
    g = case.grid
    rng = np.random.default_rng(seed)
    f = rng.standard_normal((4, nshots, g.ntx * g.ntz)).astype(np.float32)
    return torch.tensor(f, device="cuda:0")


def test_noalloc_step_matches_workspace_step():
    import torch

    c = configs.small(nt=10)
    ref = swd.Plan(c.grid, c.coefs, c.vp, c.rho)
    fresh = swd.Plan(c.grid, c.coefs, c.vp, c.rho)
    assert fresh.stats()["ws_bytes"] == 0
    for nshots, nfftz, seed in ((5, 256, 1), (2, 128, 2)):
        print(f"noalloc step: {nshots} shots, nfftz {nfftz}", flush=True)
        kz, _ = swd.kz_taper(nfftz, c.grid.dz)
        f = _device_fields(torch, c, nshots, seed)
        outs = {w: torch.zeros((4, nshots, c.grid.ntx * c.grid.ntz), device="cuda:0") for w in ("ref", "na")}
        ptrs = {w: {k: outs[w][i].data_ptr() for i, k in enumerate(NAMES)} for w in outs}
        args = (nshots, f[0].data_ptr(), f[1].data_ptr(), f[2].data_ptr(), f[3].data_ptr(), kz)
        ref.updown_image_dev(*args, ptrs["ref"])
        fresh.updown_image_dev_noalloc(*args, ptrs["na"])
        torch.cuda.synchronize()
        a, b = outs["ref"].cpu().numpy(), outs["na"].cpu().numpy()
        assert np.abs(a[0]).max() > 0
        for i, k in enumerate(NAMES):
            assert np.array_equal(a[i], b[i]), (nshots, nfftz, k)
        assert fresh.stats()["ws_bytes"] == 0
    # the same plan again through the workspace entry: still the same bits
    kz, _ = swd.kz_taper(128, c.grid.dz)
    f = _device_fields(torch, c, 3, 3)
    o = torch.zeros((2, 4, 3, c.grid.ntx * c.grid.ntz), device="cuda:0")
    args = (3, f[0].data_ptr(), f[1].data_ptr(), f[2].data_ptr(), f[3].data_ptr(), kz)
    fresh.updown_image_dev_noalloc(*args, {k: o[0, i].data_ptr() for i, k in enumerate(NAMES)})
    fresh.updown_image_dev(*args, {k: o[1, i].data_ptr() for i, k in enumerate(NAMES)})
    torch.cuda.synchronize()
    assert torch.equal(o[0], o[1])
    ref.close()
    fresh.close()


def test_noalloc_step_rejects_batches_beyond_its_array():
    import torch

    c = configs.small(nt=10)
    plan = swd.Plan(c.grid, c.coefs, c.vp, c.rho)
    n = swd.SWD_NOALLOC_MAX_SHOTS + 1
    f = torch.zeros((c.grid.ntx * c.grid.ntz,), device="cuda:0")
    p = f.data_ptr()
    kz, _ = swd.kz_taper(128, c.grid.dz)
    with pytest.raises(swd.SwdError) as e:
        plan.updown_image_dev_noalloc(n, p, p, p, p, kz, {"du": p, "ud": p})
    assert e.value.code == -6 and "at most 64" in str(e.value)
    with pytest.raises(swd.SwdError) as e:
        plan.updown_image_dev_noalloc(1, p, p, p, p, np.ones(8192, np.float32), {"du": p, "ud": p})
    assert e.value.code == -1 and "nfftz" in str(e.value)
    plan.close()
