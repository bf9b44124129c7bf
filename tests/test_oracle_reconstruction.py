"""Pins of the oracle's RTM backward (RK:434-620, 765-817, 946-1215) that do not
restate its formulas.

The source reconstruction runs the forward leapfrog backwards from the final
state, subtracting the wavelet at it+1 (RK:478-481) and restoring the saved
interior frame of step it (RK:575-620, launched at RK:1087), from it = nt-2 down
to 0 (RK:1073).  The receiver field is overwritten by the adjoint source at its
nodes (RK:765-780), and the image cross-correlates dp/dt = rho vp^2 div v^{it+1}
with the receiver p of step it (RK:475, 811).  What fixes each of these
independently of the backward code:

* exact reconstruction: while the wavefield never reaches the interior frame the
  reverse step is the exact inverse of the forward step, so the reconstructed p
  of step it equals the forward run's p of step it to fp64 rounding;
* frame restore: on the frame the reconstructed p is the forward p of the same
  step, bit for bit, also where the wavefield crosses it;
* dp/dt: the forward run gives it as (p^{it+1} - p^{it} - delta_s wavelet[it+1]) / dt,
  and illum_s is the sum of its squares over it = 0 .. nt-2;
* hard replacement: at a receiver node the receiver p of step it is the adjoint
  source sample (the last trace wins on a shared node), so illum_r there is the
  sum of the squared samples;
* imaging: at a receiver node the image is sum_it dp/dt(it) adj(it), with dp/dt
  from the forward run.

Each deliberate mistake the oracle can be built with (OR_PLANT in
oracle/swd_oracle_prop.inc) must fail the pin that names it.
"""
import numpy as np
import pytest

import _oracle as O

NX = NZ = 90
PML, LC = 10, 4
DX, DT, F0 = 10.0, 1e-3, 25.0


def _layered(ntx, ntz, pmlc):
    vp = np.empty((ntx, ntz), np.float32)
    z = np.arange(ntz) - pmlc
    prof = np.where(z < 30, 1800.0, np.where(z < 60, 2300.0, 2700.0)).astype(np.float32)
    vp[:] = prof[None, :]
    return vp


def _frame_mask(P):
    """Interior frame of width Lc (RK:200-219): its cells as a boolean [ntx][ntz]."""
    m = np.zeros((P.ntx, P.ntz), bool)
    a, bx, bz = P.pmlc, P.ntx - P.pmlc, P.ntz - P.pmlc
    m[a:a + LC, a:bz] = True
    m[bx - LC:bx, a:bz] = True
    m[a:bx, a:a + LC] = True
    m[a:bx, bz - LC:bz] = True
    return m.reshape(-1)


def _interior_mask(P):
    m = np.zeros((P.ntx, P.ntz), bool)
    m[P.pmlc:P.ntx - P.pmlc, P.pmlc:P.ntz - P.pmlc] = True
    return m.reshape(-1)


def _problem(s_ix, s_iz, rec_ix, rec_iz, nt, L):
    pmlc = PML + LC
    ntx, ntz = NX + 2 * pmlc, NZ + 2 * pmlc
    vp = _layered(ntx, ntz, pmlc)
    rho = np.full((ntx, ntz), 1000.0, np.float32)
    P = O.Problem(NX, NZ, PML, LC, DX, DX, DT, nt, vp, rho, s_ix, s_iz, rec_ix, rec_iz, f0=F0)
    P.L = L
    return P


def _quiet_case(L):
    """Source in the middle, nt short enough that no wave reaches the frame.
    Receivers: a column beside the source, one node carrying two traces."""
    pmlc = PML + LC
    s_ix, s_iz = pmlc + 45, pmlc + 45
    rz = np.arange(s_iz - 10, s_iz + 11)
    rec_ix = np.concatenate([np.full(rz.size, s_ix + 6), [s_ix + 6]])
    rec_iz = np.concatenate([rz, [rz[3]]])
    return _problem(s_ix, s_iz, rec_ix, rec_iz, 100, L)


def _loud_case(L):
    """Source two cells inside the frame's corner: the wavefield crosses the frame."""
    pmlc = PML + LC
    return _problem(pmlc + 6, pmlc + 3, np.array([pmlc + 20]), np.array([pmlc + 20]), 200, L)


def _adj(P, seed):
    return np.random.default_rng(seed).standard_normal((P.nrec, P.nt)).astype(np.float32)


def pin_errors(L=None):
    """The error of each pin (0 when exact): a dict name -> float."""
    err = {}
    # ---- quiet frame, fp64
    P = _quiet_case(L)
    _, st, fs, ph = P.forward_hist(np.float64)
    frame = _frame_mask(P)
    inner = _interior_mask(P)
    scale = np.abs(ph).max()
    assert np.abs(ph[:, frame]).max() < 1e-11 * scale, "the quiet case reaches the frame"
    adj = _adj(P, 11)
    image, ill_s, ill_r, sh, rh, dh = P.rtm_backward_hist(fs, st, adj, np.float64)
    nt = P.nt
    its = np.arange(nt - 1)
    d_rec = np.abs(sh[its] - ph[its]).max()
    err["exact_reconstruction"] = float(d_rec / scale)
    src = P.shot.s_ix * P.ntz + P.shot.s_iz
    dt = float(np.float32(DT))                           # the float time step the oracle steps with
    dp = (ph[1:] - ph[:-1]) / dt                         # rows it = 0 .. nt-2
    dp[:, src] -= P.wavelet[1:].astype(np.float64) / dt
    dp[:, ~inner] = 0.0
    e_dpdt = np.abs(dh[its] - dp).max() / np.abs(dp).max()
    ill_ref = (dp ** 2).sum(axis=0)
    e_ill = np.linalg.norm(ill_s - ill_ref) / np.linalg.norm(ill_ref)
    err["dpdt"] = float(max(e_dpdt, e_ill))
    nodes = P.rec_ix * P.ntz + P.rec_iz
    last = {int(n): tr for tr, n in enumerate(nodes)}     # the last trace on a node wins
    e_rep = e_ilr = e_img = 0.0
    for n, tr in last.items():
        a = adj[tr].astype(np.float64)
        e_rep = max(e_rep, float(np.abs(rh[its, n] - a[its]).max() / np.abs(a).max()))
        ref = float((a[its] ** 2).sum())
        e_ilr = max(e_ilr, abs(ill_r[n] - ref) / ref)
        ref_img = float((dp[:, n] * a[its]).sum())
        e_img = max(e_img, abs(image[n] - ref_img) / np.sqrt(float((dp[:, n] ** 2).sum()) * ref))
    err["hard_replacement"] = max(e_rep, e_ilr)
    err["imaging"] = e_img
    # ---- loud frame, fp32: the frame restore is exact wherever the field is
    Q = _loud_case(L)
    _, st, fs, ph = Q.forward_hist(np.float32)
    frame = _frame_mask(Q)
    fr = ph[:, frame]
    moving = np.mean(np.any(fr[1:] != fr[:-1], axis=1))
    assert moving > 0.9, f"the frame carries too little signal ({moving:.2f} of the steps change it)"
    _, _, _, sh, _, _ = Q.rtm_backward_hist(fs, st, _adj(Q, 12), np.float32)
    its = np.arange(Q.nt - 1)
    err["frame_restore"] = float(np.abs(sh[its][:, frame] - fr[its]).max() / np.abs(fr).max())
    return err


TOL = {"exact_reconstruction": 1e-12, "dpdt": 1e-11, "hard_replacement": 1e-13, "imaging": 1e-12,
       "frame_restore": 0.0}


def test_reconstruction_pins_hold():
    err = pin_errors()
    print({k: f"{v:.2e}" for k, v in err.items()})
    for k, tol in TOL.items():
        assert err[k] <= tol, (k, err[k], tol)


@pytest.mark.parametrize("plant,pin", [
    (1, "exact_reconstruction"),   # wavelet[it] subtracted instead of wavelet[it+1]
    (2, "frame_restore"),          # strips of step it+1 restored instead of it
    (3, "hard_replacement"),       # adjoint source added instead of replacing p
    (4, "exact_reconstruction"),   # backward loop from nt-3
    (5, "dpdt"),                   # dp/dt with dt rho vp^2
])
def test_planted_mistake_fails_its_pin(plant, pin, tmp_path):
    err = pin_errors(O.build_variant(str(tmp_path), plant))
    print(plant, {k: f"{v:.2e}" for k, v in err.items()})
    assert err[pin] > 100 * max(TOL[pin], 1e-13), (plant, pin, err[pin])
