"""torch.autograd.Function over libswd's Born operator.

``born(m, plan, shots, wavelet)`` returns the Born data d = L m of every shot
of the batch ([sum nrec][nt], float32) for a reflectivity m = dK/K on the
interior ([nx][nz], shared by the shots).  Its backward is the exact discrete
adjoint L^T (reversed time loop of the transposed CPML staggered scheme) with
sqrt(nt) snapshot checkpointing, so ``d.backward(r)`` leaves L^T r in m.grad.

Tensors may live on the CPU (host entry points) or on the plan's GPU
(device-pointer entry points, no host round trip).  The operator is linear in
m, so no state is saved for the backward beyond the call's configuration.
"""
import numpy as np
import torch

from . import ShotSet


def _shots(shots):
    return shots if isinstance(shots, ShotSet) else ShotSet(shots)


def _check_device(plan, t):
    if t.is_cuda and t.device.index != plan.device:
        raise ValueError(f"tensor on {t.device} but the plan runs on cuda:{plan.device}")


class BornOp(torch.autograd.Function):
    @staticmethod
    def forward(ctx, m, plan, shots, wavelet, nsnap):
        S = _shots(shots)
        g = plan.grid
        nint = g.nx * g.nz
        if m.dtype != torch.float32 or m.numel() != nint:
            raise ValueError(f"m must be float32 with nx*nz = {nint} elements, got {m.dtype} {tuple(m.shape)}")
        _check_device(plan, m)
        ctx.cfg = (plan, S, np.ascontiguousarray(wavelet, np.float32), int(nsnap), m.shape)
        mm = m.detach().reshape(1, nint).expand(S.n, nint).contiguous()
        if m.is_cuda:
            d = torch.empty((S.ntr, g.nt), dtype=torch.float32, device=m.device)
            torch.cuda.current_stream(m.device).synchronize()   # the library runs on its own stream
            plan.born_dev(S, ctx.cfg[2], mm.data_ptr(), d.data_ptr())
            return d
        return torch.from_numpy(plan.born(S, ctx.cfg[2], mm.numpy()))

    @staticmethod
    def backward(ctx, gd):
        plan, S, wavelet, nsnap, shape = ctx.cfg
        if not ctx.needs_input_grad[0]:
            return None, None, None, None, None
        g = plan.grid
        gd = gd.detach().to(torch.float32).contiguous()
        _check_device(plan, gd)
        if gd.is_cuda:
            gm = torch.empty(g.nx * g.nz, dtype=torch.float32, device=gd.device)
            torch.cuda.current_stream(gd.device).synchronize()
            plan.born_adjoint_dev(S, wavelet, gd.data_ptr(), gm.data_ptr(), nsnap)
        else:
            gm = torch.from_numpy(plan.born_adjoint(S, wavelet, gd.numpy(), nsnap))
        return gm.reshape(shape), None, None, None, None


def born(m, plan, shots, wavelet, nsnap=0):
    """d = L m with autograd support (see the module docstring)."""
    return BornOp.apply(m, plan, shots, wavelet, nsnap)
