"""Shared helpers of the parity tests."""
import numpy as np

import _oracle as O


def rel_l2(a, ref):
    a = np.asarray(a, np.float64)
    ref = np.asarray(ref, np.float64)
    n = np.linalg.norm(ref)
    return float(np.linalg.norm(a - ref) / (n if n > 0 else 1.0))


def oracle_problem(case, shot=0):
    """Oracle problem of one shot of a swd.configs.Case (same model, coefficients, wavelet)."""
    g = case.grid
    s_ix, s_iz, rx, rz = case.shots[shot]
    return O.Problem(g.nx, g.nz, g.pml, g.Lc, g.dx, g.dz, g.dt, g.nt, case.vp, case.rho, s_ix, s_iz, rx, rz,
                     wavelet=case.wavelet, coefs=case.coefs)


def report(name, gpu, o32, o64=None):
    """Prints the errors a parity test asserts on (shown with pytest -s / on failure)."""
    e32 = rel_l2(gpu, o32)
    mx = float(np.max(np.abs(np.asarray(gpu, np.float64) - np.asarray(o32, np.float64))))
    msg = f"{name}: relL2 vs fp32 oracle {e32:.3e} (max|d| {mx:.3e})"
    e64 = None
    if o64 is not None:
        e64 = rel_l2(gpu, o64)
        msg += f", vs fp64 oracle {e64:.3e}, fp32-oracle vs fp64 {rel_l2(o32, o64):.3e}"
    print(msg, flush=True)
    return e32, e64
