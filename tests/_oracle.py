"""ctypes bindings of the CPU oracle (oracle/libswd_oracle.so).

Test infrastructure: the oracle is the checker the HIP path is compared with.
Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg load it.
"""
import ctypes as C
import os
import subprocess

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE_DIR = os.path.join(ROOT, "oracle")
LIB_PATH = os.path.join(ORACLE_DIR, "_build", "libswd_oracle.so")

f32p = np.ctypeslib.ndpointer(np.float32, flags="C_CONTIGUOUS")
f64p = np.ctypeslib.ndpointer(np.float64, flags="C_CONTIGUOUS")
i32p = np.ctypeslib.ndpointer(np.int32, flags="C_CONTIGUOUS")


class Grid(C.Structure):
    _fields_ = [("nx", C.c_int), ("nz", C.c_int), ("pml", C.c_int), ("Lc", C.c_int),
                ("dx", C.c_float), ("dz", C.c_float), ("dt", C.c_float), ("nt", C.c_int)]


class Coefs(C.Structure):
    _fields_ = [(n, C.c_void_p) for n in ("rc", "a_x", "a_x_half", "b_x", "b_x_half",
                                          "a_z", "a_z_half", "b_z", "b_z_half")]


class Shot(C.Structure):
    _fields_ = [("s_ix", C.c_int), ("s_iz", C.c_int), ("nrec", C.c_int),
                ("rec_ix", C.c_void_p), ("rec_iz", C.c_void_p)]


_lib = None


def build():
    subprocess.run(["make", "-s", "-C", ORACLE_DIR], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        _lib = _bind(C.CDLL(LIB_PATH))
    return _lib


def build_variant(out_dir, plant):
    """Builds the oracle with the deliberate mistake OR_PLANT=plant (see
    swd_oracle_prop.inc) into out_dir and returns the bound library: the
    reconstruction tests check that each such mistake fails them."""
    so = os.path.join(out_dir, f"libswd_oracle_plant{plant}.so")
    subprocess.run(["gcc", "-O3", "-march=x86-64-v3", "-ffp-contract=off", "-fno-fast-math", "-fopenmp", "-fPIC",
                    "-w", f"-DOR_PLANT={int(plant)}", "-shared", "-o", so, os.path.join(ORACLE_DIR, "swd_oracle.c"),
                    "-lm"], check=True)
    return _bind(C.CDLL(so))


def _bind(L):
    """Sets the argument and result types of the oracle's entry points on L."""
    L.or_cal_xishu.argtypes = [C.c_int, f32p]
    L.or_ricker.argtypes = [f32p, C.c_int, C.c_float, C.c_float, C.c_float]
    L.or_cpml.argtypes = [C.c_int, C.c_int, C.c_int, C.c_float, C.c_float, C.c_float] + [f32p] * 8
    L.or_model_in_pml.argtypes = [f32p, C.c_int, C.c_int, C.c_int]
    L.or_smooth.argtypes = [f32p, C.c_int, C.c_int, C.c_int, C.c_int]
    L.or_nfft.argtypes = [C.c_int]
    L.or_nfft.restype = C.c_int
    L.or_kz_taper.argtypes = [f32p, C.c_int, C.c_float]
    L.or_kz_taper.restype = C.c_int
    L.or_hilbert1d.argtypes = [f32p, f32p, C.c_int]
    L.or_geometry.argtypes = [C.c_int, C.c_int, C.c_int, C.c_int, C.c_float, C.c_float,
                              C.c_float, C.c_float, C.c_float, C.c_float, C.c_int,
                              C.POINTER(C.c_int), C.POINTER(C.c_int), C.c_void_p, C.c_int]
    L.or_geometry.restype = C.c_int
    L.or_adjoint_source.argtypes = [f32p, f32p, C.c_int, C.c_int, C.c_float]
    L.or_scatterers.argtypes = [f32p, f32p, C.c_int, C.c_int, C.c_int, C.c_int]
    L.or_image_normalize.argtypes = [f32p, f32p, C.c_int]
    L.or_image_laplacian.argtypes = [f32p, f32p, f32p, C.c_void_p, C.c_int, C.c_int, C.c_int,
                                     C.c_float, C.c_float]
    L.or_edge_zero.argtypes = [f32p, C.c_int, C.c_int, C.c_int]
    L.or_frame_size.argtypes = [C.POINTER(Grid)]
    L.or_frame_size.restype = C.c_int
    L.or_set_threads.argtypes = [C.c_int]
    for name in ("or_forward_f32", "or_forward_f64", "or_rtm_backward_f32", "or_rtm_backward_f64",
                 "or_afor_f32", "or_afor_f64", "or_born_f32", "or_born_f64",
                 "or_born_adjoint_f32", "or_born_adjoint_f64", "or_rtm_updown_f32", "or_rtm_updown_f64",
                 "or_rtm_cig_f32", "or_rtm_cig_f64", "or_forward_hist_f32", "or_forward_hist_f64",
                 "or_rtm_backward_hist_f32", "or_rtm_backward_hist_f64"):
        if hasattr(L, name):
            getattr(L, name).restype = C.c_int
    return L


def ptr(a):
    return a.ctypes.data_as(C.c_void_p) if a is not None else None


class Problem:
    """A padded-grid problem: grid, model, coefficients and one shot."""

    def __init__(self, nx, nz, pml, Lc, dx, dz, dt, nt, vp, rho, s_ix, s_iz, rec_ix, rec_iz,
                 wavelet=None, f0=None, coefs=None):
        self.g = Grid(nx, nz, pml, Lc, dx, dz, dt, nt)
        self.pmlc = pml + Lc
        self.ntx = nx + 2 * self.pmlc
        self.ntz = nz + 2 * self.pmlc
        self.ntp = self.ntx * self.ntz
        self.vp = np.ascontiguousarray(vp, np.float32).reshape(-1)
        self.rho = np.ascontiguousarray(rho, np.float32).reshape(-1)
        assert self.vp.size == self.ntp and self.rho.size == self.ntp
        if coefs is None:
            coefs = setup_coefs(self.ntx, self.ntz, self.pmlc, Lc, dx, dz, dt)
        self.coefs = coefs
        self.c = Coefs(*[ptr(coefs[k]) for k in ("rc", "a_x", "a_x_half", "b_x", "b_x_half",
                                                 "a_z", "a_z_half", "b_z", "b_z_half")])
        self.rec_ix = np.ascontiguousarray(rec_ix, np.int32)
        self.rec_iz = np.ascontiguousarray(rec_iz, np.int32)
        self.nrec = self.rec_ix.size
        self.shot = Shot(int(s_ix), int(s_iz), self.nrec, ptr(self.rec_ix), ptr(self.rec_iz))
        if wavelet is None:
            wavelet = ricker(nt, f0, 1.0 / f0, dt)
        self.wavelet = np.ascontiguousarray(wavelet, np.float32)
        self.nt = nt
        self.L = None   # a library from build_variant() in place of the oracle

    def _lib(self):
        return self.L if self.L is not None else lib()

    def frame_size(self):
        return self._lib().or_frame_size(C.byref(self.g))

    def _args(self):
        return (C.byref(self.g), ptr(self.vp), ptr(self.rho), C.byref(self.c), C.byref(self.shot))

    def forward(self, dtype=np.float32, strips=False, final=False, wavelet=None):
        L = self._lib()
        w = self.wavelet if wavelet is None else np.ascontiguousarray(wavelet, np.float32)
        gather = np.zeros((self.nrec, self.nt), dtype)
        st = np.zeros((self.nt, self.frame_size()), dtype) if strips else None
        fs = np.zeros((3, self.ntp), dtype) if final else None
        fn = L.or_forward_f32 if dtype == np.float32 else L.or_forward_f64
        rc = fn(*self._args(), ptr(w), ptr(gather), ptr(st), ptr(fs))
        assert rc == 0, rc
        return gather, st, fs

    def rtm_backward(self, final_state, strips, adj, dtype=np.float32):
        L = self._lib()
        adj = np.ascontiguousarray(adj, np.float32)
        out = [np.zeros(self.ntp, dtype) for _ in range(3)]
        fn = L.or_rtm_backward_f32 if dtype == np.float32 else L.or_rtm_backward_f64
        fs = np.ascontiguousarray(final_state, dtype)
        st = np.ascontiguousarray(strips, dtype)
        rc = fn(*self._args(), ptr(self.wavelet), ptr(fs), ptr(st), ptr(adj), *[ptr(o) for o in out])
        assert rc == 0, rc
        return out

    def forward_hist(self, dtype=np.float32):
        """(gather, strips, final_state, phist): phist[it] is p after step it, [nt][ntp]."""
        L = self._lib()
        gather = np.zeros((self.nrec, self.nt), dtype)
        st = np.zeros((self.nt, self.frame_size()), dtype)
        fs = np.zeros((3, self.ntp), dtype)
        ph = np.zeros((self.nt, self.ntp), dtype)
        fn = L.or_forward_hist_f32 if dtype == np.float32 else L.or_forward_hist_f64
        rc = fn(*self._args(), ptr(self.wavelet), ptr(gather), ptr(st), ptr(fs), ptr(ph))
        assert rc == 0, rc
        return gather, st, fs, ph

    def rtm_backward_hist(self, final_state, strips, adj, dtype=np.float32):
        """(image, illum_s, illum_r, shist, rhist, dhist): at backward step it,
        shist[it] is the reconstructed p, rhist[it] the receiver p after the hard
        replacement, dhist[it] dp/dt ([nt][ntp]; row nt-1 stays zero)."""
        L = self._lib()
        adj = np.ascontiguousarray(adj, np.float32)
        out = [np.zeros(self.ntp, dtype) for _ in range(3)]
        hist = [np.zeros((self.nt, self.ntp), dtype) for _ in range(3)]
        fn = L.or_rtm_backward_hist_f32 if dtype == np.float32 else L.or_rtm_backward_hist_f64
        fs = np.ascontiguousarray(final_state, dtype)
        st = np.ascontiguousarray(strips, dtype)
        rc = fn(*self._args(), ptr(self.wavelet), ptr(fs), ptr(st), ptr(adj), *[ptr(o) for o in out],
                *[ptr(h) for h in hist])
        assert rc == 0, rc
        return (*out, *hist)

    def rtm_cig(self, final_state, strips, adj, thetan=10, delt=1, dtype=np.float32):
        """RTM with angle gathers: image, illum_s, illum_r [ntp] and cig [ntp][thetan]."""
        L = self._lib()
        adj = np.ascontiguousarray(adj, np.float32)
        out = [np.zeros(self.ntp, dtype) for _ in range(3)]
        cig = np.zeros((self.ntp, thetan), dtype)
        fn = L.or_rtm_cig_f32 if dtype == np.float32 else L.or_rtm_cig_f64
        fs = np.ascontiguousarray(final_state, dtype)
        st = np.ascontiguousarray(strips, dtype)
        rc = fn(*self._args(), ptr(self.wavelet), ptr(fs), ptr(st), ptr(adj), int(thetan), int(delt),
                *[ptr(o) for o in out], ptr(cig))
        assert rc == 0, rc
        return (*out, cig)

    def rtm_updown(self, hwavelet, fwd, hfwd, adj, hadj, kzcoef, dtype=np.float32):
        """Up/down-separated images (du, ud, uu, dd), each [ntp].  fwd and hfwd
        are the (final_state, strips) of the forward runs with the wavelet and
        with hwavelet, in dtype."""
        L = self._lib()
        nfftz = kzcoef.size
        out = [np.zeros(self.ntp, dtype) for _ in range(4)]
        fn = L.or_rtm_updown_f32 if dtype == np.float32 else L.or_rtm_updown_f64
        arrs = [np.ascontiguousarray(a, dtype) for a in (fwd[0], fwd[1], hfwd[0], hfwd[1])]
        f32 = [np.ascontiguousarray(a, np.float32) for a in (hwavelet, adj, hadj, kzcoef)]
        rc = fn(*self._args(), ptr(self.wavelet), ptr(f32[0]), *[ptr(a) for a in arrs], ptr(f32[1]), ptr(f32[2]),
                ptr(f32[3]), nfftz, *[ptr(o) for o in out])
        assert rc == 0, rc
        return out

    def afor(self, hwavelet, kzcoef, dtype=np.float32):
        L = self._lib()
        nfftz = kzcoef.size
        outs = [np.zeros((self.nrec, self.nt), dtype) for _ in range(3)]
        fn = L.or_afor_f32 if dtype == np.float32 else L.or_afor_f64
        rc = fn(*self._args(), ptr(self.wavelet), ptr(np.ascontiguousarray(hwavelet, np.float32)),
                ptr(np.ascontiguousarray(kzcoef, np.float32)), nfftz, *[ptr(o) for o in outs])
        assert rc == 0, rc
        return outs

    def born(self, m, dtype=np.float32):
        L = self._lib()
        m = np.ascontiguousarray(m, np.float32).reshape(-1)
        d = np.zeros((self.nrec, self.nt), dtype)
        fn = L.or_born_f32 if dtype == np.float32 else L.or_born_f64
        rc = fn(*self._args(), ptr(self.wavelet), ptr(m), ptr(d))
        assert rc == 0, rc
        return d

    def born_adjoint(self, d, dtype=np.float32):
        L = self._lib()
        d = np.ascontiguousarray(d, np.float32)
        g = np.zeros(self.g.nx * self.g.nz, dtype)
        fn = L.or_born_adjoint_f32 if dtype == np.float32 else L.or_born_adjoint_f64
        rc = fn(*self._args(), ptr(self.wavelet), ptr(d), ptr(g))
        assert rc == 0, rc
        return g


def cal_xishu(Lc):
    rc = np.zeros(Lc, np.float32)
    lib().or_cal_xishu(Lc, rc)
    return rc


def ricker(nt, f0, t0, dt):
    r = np.zeros(nt, np.float32)
    lib().or_ricker(r, nt, f0, t0, dt)
    return r


def cpml(ntx, ntz, pml_arg, dx, dz, dt):
    arrs = [np.zeros(ntx, np.float32) for _ in range(4)] + [np.zeros(ntz, np.float32) for _ in range(4)]
    lib().or_cpml(ntx, ntz, pml_arg, dx, dz, dt, *arrs)
    keys = ("a_x", "a_x_half", "b_x", "b_x_half", "a_z", "a_z_half", "b_z", "b_z_half")
    return dict(zip(keys, arrs))


def setup_coefs(ntx, ntz, pmlc, Lc, dx, dz, dt):
    c = cpml(ntx, ntz, pmlc, dx, dz, dt)
    c["rc"] = cal_xishu(Lc)
    return c


def smooth(vp_padded, ntz, ntx, pmlc, window):
    v = np.ascontiguousarray(vp_padded, np.float32).reshape(-1).copy()
    lib().or_smooth(v, ntz, ntx, pmlc, window)
    return v


def model_in_pml(vp_padded, ntz, ntx, pmlc):
    v = np.ascontiguousarray(vp_padded, np.float32).reshape(-1).copy()
    lib().or_model_in_pml(v, ntz, ntx, pmlc)
    return v


def cig_cells(ps, vxs, vzs, pr, vxr, vzr, it, nt, delt, thetan, ringF, ringG, cig, dtype=np.float32):
    """One step of the angle-gather update on n cells (or_cig_cells); rings
    [delt][n] and cig [n][thetan] are updated in place (arrays of dtype)."""
    L = lib()
    fn = L.or_cig_cells_f32 if dtype == np.float32 else L.or_cig_cells_f64
    f = [np.ascontiguousarray(a, dtype) for a in (ps, vxs, vzs, pr, vxr, vzr)]
    for a in (ringF, ringG, cig):
        assert a.dtype == dtype and a.flags.c_contiguous
    fn(int(f[0].size), *[ptr(a) for a in f], int(it), int(nt), int(delt), int(thetan), ptr(ringF), ptr(ringG),
       ptr(cig))


def nfft(n):
    return lib().or_nfft(n)


def kz_taper(nfftz, dz):
    c = np.zeros(nfftz, np.float32)
    layerz = lib().or_kz_taper(c, nfftz, dz)
    return c, layerz


def hilbert1d(rick):
    rick = np.ascontiguousarray(rick, np.float32)
    h = np.zeros_like(rick)
    lib().or_hilbert1d(rick, h, rick.size)
    return h


def geometry(nx, nz, pml, Lc, dx, dz, sx0, shotdx, shotdep, offsetz, is_, max_rn=1 << 20):
    s_ix, s_iz = C.c_int(), C.c_int()
    rn = lib().or_geometry(nx, nz, pml, Lc, dx, dz, sx0, shotdx, shotdep, offsetz, is_,
                           C.byref(s_ix), C.byref(s_iz), None, 0)
    r_iz = np.zeros(max(rn, 1), np.int32)
    lib().or_geometry(nx, nz, pml, Lc, dx, dz, sx0, shotdx, shotdep, offsetz, is_,
                      C.byref(s_ix), C.byref(s_iz), ptr(r_iz), rn)
    return s_ix.value, s_iz.value, r_iz[:rn]


def adjoint_source(d, dt):
    d = np.ascontiguousarray(d, np.float32)
    out = np.zeros_like(d)
    lib().or_adjoint_source(d, out, d.shape[0], d.shape[1], dt)
    return out


def scatterers(rho, vp, ntx, ntz, pmlc, swin):
    r = np.ascontiguousarray(rho, np.float32).reshape(-1).copy()
    lib().or_scatterers(r, np.ascontiguousarray(vp, np.float32).reshape(-1), ntx, ntz, pmlc, swin)
    return r


def image_normalize(image, illum):
    im = np.ascontiguousarray(image, np.float32).reshape(-1).copy()
    lib().or_image_normalize(im, np.ascontiguousarray(illum, np.float32).reshape(-1), im.size)
    return im


def image_laplacian(image, vp, ntx, ntz, pmlc, dx, dz, acc=None):
    nx, nz = ntx - 2 * pmlc, ntz - 2 * pmlc
    out = np.zeros(nx * nz, np.float32)
    lib().or_image_laplacian(np.ascontiguousarray(image, np.float32).reshape(-1),
                             np.ascontiguousarray(vp, np.float32).reshape(-1), out,
                             ptr(acc), ntx, ntz, pmlc, dx, dz)
    return out


def edge_zero(mig, nx, nz, Lc):
    m = np.ascontiguousarray(mig, np.float32).reshape(-1).copy()
    lib().or_edge_zero(m, nx, nz, Lc)
    return m


def set_threads(n):
    lib().or_set_threads(n)


# ---------------------------------------------------------------------------
# Wavenumber-domain PSF deconvolution (kkdecon_acoustic/kk_decon_interp_psf_minewindow.cpp)

class DeconParams(C.Structure):
    _fields_ = [(n, C.c_int) for n in ("nx", "nz", "psfwinx", "psfwinz", "filterwinx", "filterwinz",
                                       "snapwinx", "snapwinz")] + [("epsilon", C.c_float)]


GEOM_FIELDS = ("halfpsfwx", "halfpsfwz", "halfltwx", "halfltwz", "nxwpsf", "nzwpsf", "fullwx", "fullwz",
               "midpx", "midpz", "outwinx", "outwinz", "halfoutwx", "halfoutwz", "stepwinx", "stepwinz",
               "stepnumx", "stepnumz", "totalnum", "nxp", "nzp")


class DeconGeom(C.Structure):
    _fields_ = [(n, C.c_int) for n in GEOM_FIELDS]


def _decon_lib():
    L = lib()
    if not getattr(L, "_decon_bound", False):
        P, G = C.POINTER(DeconParams), C.POINTER(DeconGeom)
        L.or_decon_geometry.argtypes = [P, G]
        L.or_decon_geometry.restype = C.c_int
        L.or_decon_coefr.argtypes = [P, G, f32p]
        L.or_decon_expand.argtypes = [f32p, C.c_int, C.c_int, f32p, C.c_int, C.c_int]
        L.or_twiddles.argtypes = [C.c_int, f64p, f64p]
        L.or_decon_spectrum.argtypes = [f64p, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, f64p, f64p]
        L.or_decon_inverse.argtypes = [f64p, f64p, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, f64p]
        L.or_decon_cell.argtypes = [P, G, C.c_int, C.c_int, C.POINTER(C.c_int), C.POINTER(C.c_int)]
        L.or_decon_psf.argtypes = [P, G, f32p, C.c_int, C.c_int, f64p, f64p, f32p]
        L.or_decon_window.argtypes = [P, G, f32p, f32p, C.c_int, f64p]
        L.or_decon.argtypes = [P, f32p, f32p, f32p, f32p]
        L.or_decon.restype = C.c_int
        L._decon_bound = True
    return L


def decon_params(nx, nz, psfwin=31, filterwin=71, snapwin=11, epsilon=5e-4):
    return DeconParams(nx, nz, psfwin, psfwin, filterwin, filterwin, snapwin, snapwin, epsilon)


def decon_geometry(p):
    g = DeconGeom()
    if _decon_lib().or_decon_geometry(C.byref(p), C.byref(g)) != 0:
        raise ValueError("invalid deconvolution windows")
    return g


def decon_coefr(p):
    g = decon_geometry(p)
    c = np.zeros((g.outwinx, g.outwinz), np.float32)
    _decon_lib().or_decon_coefr(C.byref(p), C.byref(g), c)
    return c


def decon_expand(a, nxp, nzp):
    a = np.ascontiguousarray(a, np.float32)
    out = np.zeros((nxp, nzp), np.float32)
    _decon_lib().or_decon_expand(a, a.shape[0], a.shape[1], out, nxp, nzp)
    return out


def twiddles(n):
    c = np.zeros(n)
    sn = np.zeros(n)
    _decon_lib().or_twiddles(n, c, sn)
    return c, sn


def decon_spectrum(x, offx, offz, n):
    """Half spectrum [n][n/2+1] of the n x n buffer holding x at (offx, offz)."""
    x = np.ascontiguousarray(x, np.float64)
    nh = n // 2 + 1
    re = np.zeros((n, nh))
    im = np.zeros((n, nh))
    _decon_lib().or_decon_spectrum(x, x.shape[0], x.shape[1], offx, offz, n, re, im)
    return re + 1j * im


def decon_inverse(d, r0, nr, c0, nc):
    """Real part of the inverse DFT of the Hermitian spectrum with half d, on a block."""
    n = d.shape[0]
    dr = np.ascontiguousarray(d.real)
    di = np.ascontiguousarray(d.imag)
    out = np.zeros((nr, nc))
    _decon_lib().or_decon_inverse(dr, di, n, r0, nr, c0, nc, out)
    return out


def decon_cell(p, ix, iz):
    g = decon_geometry(p)
    cx, cz = C.c_int(), C.c_int()
    _decon_lib().or_decon_cell(C.byref(p), C.byref(g), ix, iz, C.byref(cx), C.byref(cz))
    return cx.value, cz.value


def decon_psf(p, psf, cx, cz):
    """(S, tden): normalised PSF half spectrum of lattice cell (cx, cz) and the Wiener denominators."""
    g = decon_geometry(p)
    psff = decon_expand(psf, g.nxp, g.nzp)
    n, nh = g.fullwx, g.fullwx // 2 + 1
    sr = np.zeros((n, nh))
    si = np.zeros((n, nh))
    t = np.zeros((n, nh), np.float32)
    _decon_lib().or_decon_psf(C.byref(p), C.byref(g), psff, cx, cz, sr, si, t)
    return sr + 1j * si, t


def decon_window(p, image, psf, step):
    """Deconvolved real part of window `step` at its outwin x outwin output points."""
    g = decon_geometry(p)
    imgf = decon_expand(image, g.nxp, g.nzp)
    psff = decon_expand(psf, g.nxp, g.nzp)
    out = np.zeros((g.outwinx, g.outwinz))
    _decon_lib().or_decon_window(C.byref(p), C.byref(g), imgf, psff, int(step), out)
    return out


def decon(p, image, psf):
    """(mig_dec, mig_coef), both [nx][nz] float32."""
    image = np.ascontiguousarray(image, np.float32)
    psf = np.ascontiguousarray(psf, np.float32)
    assert image.shape == (p.nx, p.nz) and psf.shape == (p.nx, p.nz)
    dec = np.zeros((p.nx, p.nz), np.float32)
    coef = np.zeros((p.nx, p.nz), np.float32)
    if _decon_lib().or_decon(C.byref(p), image, psf, dec, coef) != 0:
        raise ValueError("invalid deconvolution windows")
    return dec, coef
