"""Seeded synthetic workloads of the FDTD/RTM path.

Each builder returns a ``Case``: grid, padded model (vp, rho as [ntx][ntz]),
FD/CPML coefficients from the library's own setup routines, wavelet and shots
(padded-grid source node and receiver lists).  Seeds are fixed so every run
(tests, bench, CPU baseline) sees the same inputs.

Workloads (interior nz x nx):
  c1  200 x 300, 10 m, dt 1 ms, nt 1000, pml 20, Lc 4: 5 flat layers 1500-3500 m/s
      times (1 + 0.03 g), g a sigma=5-cell Gaussian-smoothed N(0,1) field scaled to
      max|g| = 1 (seed 1); rho 1000; 25 Hz Ricker at (iz 2, ix 150); 91 receivers at
      ix 100, iz 10..190 step 2.
  c2  502 x 845, 10 m, dt 0.8 ms, nt 4000, pml 40, Lc 4: layered 1500-4500 m/s with
      8 random dipping reflectors (seed 2); migration model = 15x15 box average;
      20 Hz Ricker at the surface; 401 receivers at ix 420 spaced 5 m (nearest node).
  c4  point-scatterer (PSF) modelling: density scatterers on a 40 m lattice over
      the c1 migration model (15 x 15 box average of the c1 model), 64 surface
      shots every 12 m; plus 32 single-shot instances of seeded log-uniform
      sizes from 128 x 128 to 502 x 845 (seed 10), all with Lc 4 and the same nt.
  c5  1001 x 1690, 5 m, dt 0.4 ms, nt 8000, pml 60, Lc 4: the c2 kind of model
      (layered 1500-4500 m/s with 8 dipping reflectors, seed 5) on the finer
      grid, 15 x 15 box-smoothed migration model, 128 seeded surface shots
      (seed 6), 801 receivers at ix 840 spaced 2.5 m (nearest node).
  c3  the c1 grid and model with nt 4000 for the Born operator; reflectivity
      m = dK/K a band-limited random field (10-30 Hz equivalent at 1500 m/s,
      seed 3), data-space probe seed 4, finite-difference directions seeds 5-9.

PSF deconvolution workload (``decon_case``), [nx][nz] float32 images:
  d1  845 x 702 (the reference's deconvolution grid), psfwin 31, filterwin 71,
      snapwin 11, epsilon 5e-4: a PSF image with a depth-varying 2D wavelet
      (Ricker along z, Gaussian along x, peak frequency falling and width
      growing with depth) at every lattice node (15 + 31 i, 15 + 31 j), and a
      migrated image = 8 dipping reflectors + 300 point diffractors (seed 11)
      blurred by the mid-depth wavelet.
"""
import dataclasses

import numpy as np

from . import Grid, cal_xishu, coefs_for, ricker, scatterers


@dataclasses.dataclass
class Case:
    name: str
    grid: Grid
    vp: np.ndarray          # [ntx][ntz]
    rho: np.ndarray
    coefs: dict
    wavelet: np.ndarray
    shots: list             # [(s_ix, s_iz, rec_ix, rec_iz)]
    f0: float
    vp_true: np.ndarray = None

    @property
    def pmlc(self):
        return self.grid.pmlc


def _gauss_smooth(field, sigma):
    """Separable Gaussian smoothing with reflecting edges (numpy only)."""
    r = int(np.ceil(4 * sigma))
    x = np.arange(-r, r + 1)
    k = np.exp(-0.5 * (x / sigma) ** 2)
    k /= k.sum()
    out = field
    for ax in (0, 1):
        pad = [(0, 0), (0, 0)]
        pad[ax] = (r, r)
        f = np.pad(out, pad, mode="reflect")
        out = np.apply_along_axis(lambda v: np.convolve(v, k, mode="valid"), ax, f)
    return out


def pad_model(interior, pmlc):
    """Interior [nx][nz] -> padded [ntx][ntz] by edge replication (RC:1428-1468)."""
    return np.pad(np.asarray(interior, np.float32), pmlc, mode="edge")


def layered_c1(nx=300, nz=200, seed=1):
    rng = np.random.default_rng(seed)
    vel = np.array([1500, 2000, 2500, 3000, 3500], np.float64)
    layer = np.minimum(np.arange(nz) * 5 // nz, 4)
    v = np.tile(vel[layer], (nx, 1))
    g = _gauss_smooth(rng.standard_normal((nx, nz)), 5.0)
    g /= np.abs(g).max()
    return (v * (1.0 + 0.03 * g)).astype(np.float32)


def make_case(name, nx, nz, pml, Lc, dx, dz, dt, nt, vp_int, f0, shots_int, rho_val=1000.0):
    """Builds a Case from an interior model and shots in interior coordinates
    (s_ix, s_iz, rec_ix[], rec_iz[]) which are shifted by pmlc."""
    g = Grid(nx, nz, pml, Lc, dx, dz, dt, nt)
    pmlc = g.pmlc
    vp = pad_model(vp_int, pmlc)
    rho = np.full_like(vp, rho_val)
    coefs = coefs_for(g)
    wav = ricker(nt, f0, 1.0 / f0, dt)
    shots = [(s_ix + pmlc, s_iz + pmlc, np.asarray(rx, np.int32) + pmlc, np.asarray(rz, np.int32) + pmlc)
             for (s_ix, s_iz, rx, rz) in shots_int]
    return Case(name, g, vp, rho, coefs, wav, shots, f0)


def c1(nt=1000):
    vp = layered_c1()
    rz = np.arange(10, 191, 2)
    return make_case("c1", 300, 200, 20, 4, 10.0, 10.0, 1e-3, nt, vp, 25.0,
                     [(150, 2, np.full(rz.size, 100), rz)])


def bandlimited(nx, nz, dx, dz, fmin, fmax, vmin, seed):
    """Seeded random field [nx][nz] whose spatial spectrum is confined to the
    wavenumbers of fmin..fmax Hz at vmin (|k| in [fmin/vmin, fmax/vmin]
    cycles/m), scaled to max|m| = 1."""
    rng = np.random.default_rng(seed)
    W = np.fft.rfft2(rng.standard_normal((nx, nz)))
    kx = np.fft.fftfreq(nx, dx)[:, None]
    kz = np.fft.rfftfreq(nz, dz)[None, :]
    k = np.hypot(kx, kz)
    m = np.fft.irfft2(W * ((k >= fmin / vmin) & (k <= fmax / vmin)), s=(nx, nz))
    return (m / np.abs(m).max()).astype(np.float32)


def c3(nt=4000):
    """Born workload: c1 grid/model/geometry, nt steps, and the seeded fields."""
    c = c1(nt=nt)
    c.name = "c3"
    g = c.grid
    c.born_m = bandlimited(g.nx, g.nz, g.dx, g.dz, 10.0, 30.0, 1500.0, 3)
    c.born_dirs = [bandlimited(g.nx, g.nz, g.dx, g.dz, 10.0, 30.0, 1500.0, s) for s in range(5, 10)]
    ntr = c.shots[0][2].size
    c.born_probe = np.random.default_rng(4).standard_normal((ntr, nt)).astype(np.float32)
    return c


def perturbed_vp(case, m):
    """Velocity with K -> K (1 + m) on the interior (vp scaled by sqrt(1 + m), in
    double, rounded once to float); the CPML frame is unchanged."""
    vp = case.vp.astype(np.float64).copy()
    p = case.pmlc
    vp[p:p + case.grid.nx, p:p + case.grid.nz] *= np.sqrt(1.0 + np.asarray(m, np.float64).reshape(case.grid.nx, case.grid.nz))
    return vp.astype(np.float32)


def layered_c2(nx=845, nz=502, seed=2):
    rng = np.random.default_rng(seed)
    z = np.arange(nz)[None, :] * 10.0
    x = np.arange(nx)[:, None] * 10.0
    # smooth background 1500 -> 4500 m/s with depth
    v = 1500.0 + 3000.0 * (z / (nz * 10.0)) ** 0.8 + 0.0 * x
    for _ in range(8):
        z0 = rng.uniform(500, nz * 10.0 - 300)
        dip = np.tan(np.deg2rad(rng.uniform(-15, 15)))
        dv = rng.uniform(-250, 350)
        v = v + dv * (z > z0 + dip * (x - nx * 5.0))
    return np.clip(v, 1500, 4500).astype(np.float32)


def box_smooth(v, w=15):
    """w x w box average with edge replication (migration model of c2)."""
    r = w // 2
    p = np.pad(v.astype(np.float64), r, mode="edge")
    c = p.cumsum(0).cumsum(1)
    c = np.pad(c, ((1, 0), (1, 0)))
    s = c[w:, w:] - c[:-w, w:] - c[w:, :-w] + c[:-w, :-w]
    return (s / (w * w)).astype(np.float32)


def c2_receivers(nrec=401, spacing=5.0, dz=10.0):
    zr = np.arange(nrec) * spacing
    return np.floor(zr / dz + 0.5).astype(np.int32)


def c2(nt=4000, migration=True, shot_ix=(300,), nx=845, nz=502):
    vt = layered_c2(nx, nz)
    vm = box_smooth(vt) if migration else vt
    rz = c2_receivers()
    rz = rz[rz < nz]
    shots = [(s, 0, np.full(rz.size, 420 if nx > 420 else nx // 2), rz) for s in shot_ix]
    c = make_case("c2", nx, nz, 40, 4, 10.0, 10.0, 8e-4, nt, vm, 20.0, shots)
    c.vp_true = pad_model(vt, c.pmlc)
    return c


def small(nx=120, nz=80, pml=10, Lc=4, nt=400, seed=7, f0=25.0, nshots=1, dup=False):
    """Small heterogeneous case for parity tests (oracle runs in well under a second)."""
    rng = np.random.default_rng(seed)
    v = np.linspace(1800, 3200, nz)[None, :] * np.ones((nx, 1))
    v = v * (1.0 + 0.05 * _gauss_smooth(rng.standard_normal((nx, nz)), 3.0) /
             max(1e-9, np.abs(_gauss_smooth(rng.standard_normal((nx, nz)), 3.0)).max()))
    shots = []
    for k in range(nshots):
        s_ix = int(nx * (k + 1) / (nshots + 1))
        rz = np.arange(2, nz - 2, 3)
        rx = np.full(rz.size, nx // 4 + k)
        if dup:
            rz = np.concatenate([rz, rz[:5]])
            rx = np.concatenate([rx, rx[:5]])
        shots.append((s_ix, 1 + k, rx, rz))
    c = make_case("small", nx, nz, pml, Lc, 10.0, 10.0, 1e-3, nt, v.astype(np.float32), f0, shots)
    c.rho = c.rho * (1.0 + 0.1 * (np.arange(c.rho.shape[1])[None, :] % 7 == 0))   # density contrasts too
    c.rho = c.rho.astype(np.float32)
    return c


def with_scatterers(case, swin):
    """Density point scatterers (+1e-3 vp) on a lattice of swin cells (FC:697-705)."""
    g = case.grid
    case.rho = scatterers(case.rho.copy(), case.vp, g.ntx, g.ntz, g.pmlc, swin).reshape(case.rho.shape)
    return case


def c4(nt=1000, nshots=64, n_uneven=32, seed=10):
    """[64-shot c1-grid PSF case] + 32 uneven single-shot PSF cases."""
    vm = box_smooth(layered_c1(), 15)
    rz = np.arange(10, 191, 2)
    shots = [(60 + int(round(1.2 * k)), 0, np.full(rz.size, 100), rz) for k in range(nshots)]
    base = with_scatterers(make_case("c4", 300, 200, 20, 4, 10.0, 10.0, 1e-3, nt, vm, 20.0, shots), 4)
    rng = np.random.default_rng(seed)
    cases = [base]
    for k in range(n_uneven):
        nz = int(round(np.exp(rng.uniform(np.log(128), np.log(502)))))
        nx = int(round(np.exp(rng.uniform(np.log(128), np.log(845)))))
        v = (1500.0 + 1500.0 * (np.arange(nz) / nz))[None, :] * np.ones((nx, 1))
        rzk = np.arange(10, nz - 10, 2)
        sh = [(nx // 2, 0, np.full(rzk.size, nx // 3), rzk)]
        cases.append(with_scatterers(make_case(f"c4u{k}", nx, nz, 20, 4, 10.0, 10.0, 1e-3, nt, v.astype(np.float32),
                                               20.0, sh), 4))
    return cases


def layered(nx, nz, dz, seed, nrefl=8):
    """Smooth 1500 -> 4500 m/s background with nrefl seeded dipping reflectors."""
    rng = np.random.default_rng(seed)
    z = np.arange(nz)[None, :] * dz
    x = np.arange(nx)[:, None] * dz
    depth = nz * dz
    v = 1500.0 + 3000.0 * (z / depth) ** 0.8 + 0.0 * x
    for _ in range(nrefl):
        z0 = rng.uniform(0.1 * depth, 0.94 * depth)
        dip = np.tan(np.deg2rad(rng.uniform(-15, 15)))
        dv = rng.uniform(-250, 350)
        v = v + dv * (z > z0 + dip * (x - nx * dz / 2))
    return np.clip(v, 1500, 4500).astype(np.float32)


def c5(nt=8000, nshots=128, shot_index=None):
    """c5 (see the module docstring); shot_index selects some of the 128 shots."""
    nx, nz, dz = 1690, 1001, 5.0
    vt = layered(nx, nz, dz, 5)
    vm = box_smooth(vt, 15)
    rz = np.floor(np.arange(801) * 2.5 / dz + 0.5).astype(np.int32)
    rz = rz[rz < nz]
    sx = np.sort(np.random.default_rng(6).integers(60, nx - 60, size=nshots))
    idx = range(nshots) if shot_index is None else shot_index
    shots = [(int(sx[k]), 0, np.full(rz.size, 840), rz) for k in idx]
    c = make_case("c5", nx, nz, 60, 4, dz, dz, 4e-4, nt, vm, 20.0, shots)
    c.vp_true = pad_model(vt, c.pmlc)
    return c


DECON_PARAMS = dict(psfwin=31, filterwin=71, snapwin=11, epsilon=5e-4)


def _psf_wavelet(w, depth_frac, rng_tilt=0.0, wz=None):
    """w x wz wavelet centred in the patch: Ricker along z, Gaussian along x."""
    wz = w if wz is None else wz
    h, hz = w // 2, wz // 2
    ax, az = np.meshgrid(np.arange(-h, h + 1), np.arange(-hz, hz + 1), indexing="ij")
    az = az + rng_tilt * ax
    fz = 0.16 - 0.07 * depth_frac              # cycles per cell along z
    sx = 2.0 + 4.0 * depth_frac                # Gaussian width along x (cells)
    a = (np.pi * fz * az) ** 2
    return ((1.0 - 2.0 * a) * np.exp(-a) * np.exp(-0.5 * (ax / sx) ** 2)).astype(np.float32)


def decon_case(nx=845, nz=702, seed=11, psfwin=31, psfwin_z=None):
    """(image, psf) [nx][nz] float32 of the d1 kind on an nx x nz grid (PSF
    lattice spacing psfwin along x and psfwin_z, default psfwin, along z)."""
    rng = np.random.default_rng(seed)
    pwz = psfwin if psfwin_z is None else psfwin_z
    psf = np.zeros((nx, nz), np.float32)
    h, hz = psfwin // 2, pwz // 2
    for cx in range(h, nx - h, psfwin):
        for cz in range(hz, nz - hz, pwz):
            psf[cx - h:cx + h + 1, cz - hz:cz + hz + 1] = _psf_wavelet(psfwin, cz / nz, 0.1 * (cx / nx - 0.5), pwz)
    refl = np.zeros((nx, nz), np.float32)
    xs = np.arange(nx)
    for _ in range(8):
        z0 = rng.uniform(0.1, 0.9) * nz
        dip = rng.uniform(-0.15, 0.15)
        iz = np.clip(np.round(z0 + dip * (xs - nx / 2)).astype(int), 0, nz - 1)
        refl[xs, iz] += rng.uniform(-1.0, 1.0)
    px = rng.integers(0, nx, 300)
    pz = rng.integers(0, nz, 300)
    np.add.at(refl, (px, pz), rng.uniform(-2.0, 2.0, 300).astype(np.float32))
    k = _psf_wavelet(psfwin, 0.5, 0.0, pwz)
    img = np.zeros((nx + psfwin - 1, nz + pwz - 1), np.float64)
    for i in range(psfwin):
        for j in range(pwz):
            if k[i, j] != 0.0:
                img[i:i + nx, j:j + nz] += k[i, j] * refl
    image = (1e6 * img[h:h + nx, hz:hz + nz]).astype(np.float32)
    return np.ascontiguousarray(image), np.ascontiguousarray(1e5 * psf)
