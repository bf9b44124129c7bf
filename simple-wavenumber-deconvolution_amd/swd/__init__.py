"""Python harness over libswd's C-ABI (include/swd.h).

Thin ctypes bindings: host numpy arrays in, numpy arrays out, plus device-pointer
entry points for callers that keep data in HBM (torch tensors on ``cuda``).
The compute path is the HIP library; there is no CPU fallback, and loading
fails loudly when ``lib/libswd.so`` has not been built.
"""
import ctypes as C
import os

import numpy as np

PKG_DIR = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB_PATH = os.path.join(PKG_DIR, "lib", "libswd.so")

SWD_PATH_PLAIN = 0
SWD_PATH_BATCHED = 1
SWD_PATH_FUSED = 2
SWD_NOALLOC_MAX_SHOTS = 64
ERRORS = {-1: "SWD_E_ARG", -2: "SWD_E_DEVICE", -3: "SWD_E_NOMEM", -4: "SWD_E_IO", -5: "SWD_E_STATE",
          -6: "SWD_E_RANGE"}


class SwdError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"{ERRORS.get(code, code)}: {msg}")
        self.code = code


class Grid(C.Structure):
    _fields_ = [("nx", C.c_int), ("nz", C.c_int), ("pml", C.c_int), ("Lc", C.c_int),
                ("dx", C.c_float), ("dz", C.c_float), ("dt", C.c_float), ("nt", C.c_int)]

    @property
    def pmlc(self):
        return self.pml + self.Lc

    @property
    def ntx(self):
        return self.nx + 2 * self.pmlc

    @property
    def ntz(self):
        return self.nz + 2 * self.pmlc


class Coefs(C.Structure):
    _fields_ = [(n, C.c_void_p) for n in ("rc", "a_x", "a_x_half", "b_x", "b_x_half",
                                          "a_z", "a_z_half", "b_z", "b_z_half")]


class Shot(C.Structure):
    _fields_ = [("s_ix", C.c_int), ("s_iz", C.c_int), ("nrec", C.c_int),
                ("rec_ix", C.c_void_p), ("rec_iz", C.c_void_p)]


class Params(C.Structure):
    _fields_ = [("nx", C.c_int), ("nz", C.c_int), ("pml", C.c_int), ("Lc", C.c_int),
                ("dx", C.c_float), ("dz", C.c_float), ("itmax", C.c_int), ("dt", C.c_float),
                ("f0", C.c_float), ("ns", C.c_int), ("sx0", C.c_float), ("shotdx", C.c_float),
                ("shotdep", C.c_float), ("recdix", C.c_float), ("offsetz", C.c_float),
                ("scatter_flag", C.c_int), ("swin", C.c_int), ("window", C.c_int)]


class Stats(C.Structure):
    _fields_ = [("fwd_ms", C.c_double), ("bwd_ms", C.c_double), ("fft_ms", C.c_double),
                ("fwd_launches", C.c_longlong), ("bwd_launches", C.c_longlong),
                ("fwd_cells", C.c_longlong), ("bwd_cells", C.c_longlong),
                ("ws_bytes", C.c_longlong), ("ckpt_bytes", C.c_longlong), ("full_bytes", C.c_longlong),
                ("waves", C.c_longlong), ("tile_x", C.c_int), ("tile_rows", C.c_int), ("tile_waves", C.c_int),
                ("steps_per_launch", C.c_int), ("tile_group_waves", C.c_int),
                ("fwd_tile_x", C.c_int), ("fwd_tile_rows", C.c_int), ("fwd_tile_group_waves", C.c_int),
                ("fwd_steps_per_launch", C.c_int)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


class UpDownImages(C.Structure):
    """swd_updown_images: output pointers of swd_rtm_updown (NULL = not wanted)."""
    NAMES = ("du", "ud", "uu", "dd", "image", "illum_src", "illum_rcv")
    _fields_ = [(k, C.c_void_p) for k in NAMES]


_lib = None

EXPORTS = ("swd_last_error", "swd_get_device_count", "swd_read_parameters", "swd_cal_xishu", "swd_ricker",
           "swd_cpml", "swd_model_in_pml", "swd_read_model", "swd_smooth_model", "swd_cfl_vmax", "swd_check_model",
           "swd_nfft", "swd_kz_taper",
           "swd_hilbert1d", "swd_geometry", "swd_adjoint_source", "swd_scatterers", "swd_image_normalize",
           "swd_image_laplacian", "swd_edge_zero", "swd_plan_create", "swd_plan_destroy", "swd_plan_set_coefs",
           "swd_plan_set_model", "swd_plan_set_path", "swd_plan_set_tile", "swd_plan_set_tile_config", "swd_plan_set_steps_per_launch", "swd_plan_set_streams", "swd_plan_set_groups", "swd_plan_set_graphs",
           "swd_plan_set_memory_limit", "swd_plan_get_stats", "swd_forward", "swd_afor", "swd_afor_dev",
           "swd_rtm", "swd_rtm_dev", "swd_rtm_cig", "swd_rtm_cig_dev", "swd_rtm_updown", "swd_rtm_updown_dev", "swd_updown_image_dev",
           "swd_updown_image_dev_noalloc",
           "swd_plan_get_stream", "swd_born", "swd_born_adjoint", "swd_born_dev",
           "swd_born_adjoint_dev", "swd_multi_forward", "swd_decon_read_parameters", "swd_decon_geometry",
           "swd_decon_taper", "swd_decon_plan_create", "swd_decon_plan_destroy", "swd_decon", "swd_decon_dev",
           "swd_decon_get_stats")


def load():
    """Loads lib/libswd.so (raises if it has not been built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"libswd.so not built ({LIB_PATH}); run __graft_entry__.build()")
        # PyTorch ships its own libamdhip64.so.7 and links it by another name:
        # loaded after ours it would be a second HIP runtime that finds no
        # device.  Loaded first, libswd binds to it and both share one runtime.
        try:
            import torch  # noqa: F401
        except ImportError:
            pass
        L = C.CDLL(LIB_PATH)
        for name in EXPORTS:
            getattr(L, name).restype = C.c_int
        _lib = L
    return _lib


def last_error():
    buf = C.create_string_buffer(1024)
    load().swd_last_error(buf, 1024)
    return buf.value.decode(errors="replace")


def check(rc):
    if rc < 0:
        raise SwdError(rc, last_error())
    return rc


def _p(a):
    return None if a is None else C.c_void_p(a.ctypes.data)


def f32(a):
    return np.ascontiguousarray(a, dtype=np.float32)


def _need(a, shape, name):
    """Raises ValueError unless a has the given size (and, for 2-D shapes, that shape):
    the C-ABI reads these buffers by the sizes the plan implies."""
    if a.size != int(np.prod(shape)) or (len(shape) > 1 and a.ndim > 1 and a.shape != tuple(shape)):
        raise ValueError(f"{name}: expected shape {tuple(shape)}, got {a.shape}")
    return a


def i32(a):
    return np.ascontiguousarray(a, dtype=np.int32)


# ---------------------------------------------------------------- setup (host)
def device_count():
    n = C.c_int(0)
    rc = load().swd_get_device_count(C.byref(n))
    return n.value if rc == 0 else 0


def read_parameters(path):
    p = Params()
    check(load().swd_read_parameters(path.encode(), C.byref(p)))
    return {k: getattr(p, k) for k, _ in Params._fields_}


def cal_xishu(Lc):
    rc = np.zeros(Lc, np.float32)
    check(load().swd_cal_xishu(Lc, _p(rc)))
    return rc


def ricker(nt, f0, t0, dt):
    r = np.zeros(nt, np.float32)
    check(load().swd_ricker(_p(r), nt, C.c_float(f0), C.c_float(t0), C.c_float(dt)))
    return r


def cpml(ntx, ntz, pml_arg, dx, dz, dt):
    xs = [np.zeros(ntx, np.float32) for _ in range(4)]
    zs = [np.zeros(ntz, np.float32) for _ in range(4)]
    check(load().swd_cpml(ntx, ntz, pml_arg, C.c_float(dx), C.c_float(dz), C.c_float(dt),
                          *[_p(a) for a in xs + zs]))
    keys = ("a_x", "a_x_half", "b_x", "b_x_half", "a_z", "a_z_half", "b_z", "b_z_half")
    return dict(zip(keys, xs + zs))


def coefs_for(grid):
    """FD + CPML coefficients of a grid as the reference driver computes them
    (CPML thickness pml+Lc, RC:462-468)."""
    c = cpml(grid.ntx, grid.ntz, grid.pmlc, grid.dx, grid.dz, grid.dt)
    c["rc"] = cal_xishu(grid.Lc)
    return c


def model_in_pml(vp, ntz, ntx, pml):
    v = f32(vp).reshape(-1).copy()
    check(load().swd_model_in_pml(_p(v), ntz, ntx, pml))
    return v


def read_model(path, ntx, ntz, pml):
    vp = np.zeros(ntx * ntz, np.float32)
    rho = np.zeros(ntx * ntz, np.float32)
    check(load().swd_read_model(path.encode(), _p(vp), _p(rho), ntx, ntz, pml))
    return vp, rho


def smooth_model(vp, ntz, ntx, pml, window):
    v = f32(vp).reshape(-1).copy()
    check(load().swd_smooth_model(_p(v), ntz, ntx, pml, window))
    return v


def cfl_vmax(grid, rc=None):
    """Largest stable velocity of the grid's scheme (swd_cfl_vmax)."""
    v = C.c_double(0.0)
    check(load().swd_cfl_vmax(C.byref(grid), _p(None if rc is None else f32(rc)), C.byref(v)))
    return v.value


def check_model(grid, vp, rho, rc=None):
    """The checks a plan makes on a padded model (swd_check_model); raises SwdError."""
    vp, rho = f32(vp).reshape(-1), f32(rho).reshape(-1)
    _need(vp, (grid.ntx * grid.ntz,), "vp")
    _need(rho, (grid.ntx * grid.ntz,), "rho")
    check(load().swd_check_model(C.byref(grid), _p(None if rc is None else f32(rc)), _p(vp), _p(rho)))


def nfft(n):
    return check(load().swd_nfft(n))


def kz_taper(nfftz, dz):
    c = np.zeros(nfftz, np.float32)
    layerz = check(load().swd_kz_taper(_p(c), nfftz, C.c_float(dz)))
    return c, layerz


def hilbert1d(rick):
    r = f32(rick)
    h = np.zeros_like(r)
    check(load().swd_hilbert1d(_p(r), _p(h), r.size))
    return h


def hilbert_traces(d):
    """Hilbert transform along time of every trace of d [..., nt] (swd_hilbert1d per trace)."""
    d = f32(d)
    rows = d.reshape(-1, d.shape[-1])
    return np.stack([hilbert1d(r) for r in rows]).reshape(d.shape)


def geometry(nx, nz, pml, Lc, dx, dz, sx0, shotdx, shotdep, offsetz, is_):
    L = load()
    s_ix, s_iz = C.c_int(), C.c_int()
    f = C.c_float
    rn = check(L.swd_geometry(nx, nz, pml, Lc, f(dx), f(dz), f(sx0), f(shotdx), f(shotdep), f(offsetz), is_,
                              C.byref(s_ix), C.byref(s_iz), None, 0))
    r_iz = np.zeros(max(rn, 1), np.int32)
    check(L.swd_geometry(nx, nz, pml, Lc, f(dx), f(dz), f(sx0), f(shotdx), f(shotdep), f(offsetz), is_,
                         C.byref(s_ix), C.byref(s_iz), _p(r_iz), rn))
    return s_ix.value, s_iz.value, r_iz[:rn]


def adjoint_source(d, dt):
    d = f32(d)
    out = np.zeros_like(d)
    check(load().swd_adjoint_source(_p(d), _p(out), d.shape[0], d.shape[1], C.c_float(dt)))
    return out


def scatterers(rho, vp, ntx, ntz, pmlc, swin):
    r = f32(rho).reshape(-1).copy()
    check(load().swd_scatterers(_p(r), _p(f32(vp).reshape(-1)), ntx, ntz, pmlc, swin))
    return r


def image_normalize(image, illum):
    im = f32(image).reshape(-1).copy()
    check(load().swd_image_normalize(_p(im), _p(f32(illum).reshape(-1)), im.size))
    return im


def image_laplacian(image, vp, ntx, ntz, pmlc, dx, dz, acc=None):
    out = np.zeros((ntx - 2 * pmlc) * (ntz - 2 * pmlc), np.float32)
    check(load().swd_image_laplacian(_p(f32(image).reshape(-1)), _p(f32(vp).reshape(-1)), _p(out), _p(acc),
                                     ntx, ntz, pmlc, C.c_float(dx), C.c_float(dz)))
    return out


def edge_zero(mig, nx, nz, Lc):
    m = f32(mig).reshape(-1).copy()
    check(load().swd_edge_zero(_p(m), nx, nz, Lc))
    return m


# ---------------------------------------------------------------- shots
class ShotSet:
    """A batch of shots: source nodes and receiver lists in padded-grid indices."""

    def __init__(self, shots):
        self._keep = []
        arr = (Shot * len(shots))()
        for k, (s_ix, s_iz, rec_ix, rec_iz) in enumerate(shots):
            rx, rz = i32(rec_ix).reshape(-1), i32(rec_iz).reshape(-1)
            assert rx.size == rz.size
            self._keep += [rx, rz]
            arr[k] = Shot(int(s_ix), int(s_iz), rx.size, rx.ctypes.data if rx.size else None,
                          rz.ctypes.data if rz.size else None)
        self.arr = arr
        self.n = len(shots)
        self.ntr = sum(int(i32(s[2]).size) for s in shots)


# ---------------------------------------------------------------- plan
class Plan:
    """A device plan: grid, coefficients and model resident on one GPU."""

    def __init__(self, grid, coefs, vp, rho, device=0, path=SWD_PATH_FUSED, tile=0):
        self.L = load()
        self.grid = grid
        self.device = int(device)
        h = C.c_void_p()
        check(self.L.swd_plan_create(C.byref(grid), device, C.byref(h)))
        self.h = h
        self._coefs = {k: f32(v) for k, v in coefs.items()}
        c = Coefs(*[self._coefs[k].ctypes.data for k in ("rc", "a_x", "a_x_half", "b_x", "b_x_half",
                                                          "a_z", "a_z_half", "b_z", "b_z_half")])
        check(self.L.swd_plan_set_coefs(self.h, C.byref(c)))
        self.set_model(vp, rho)
        self.set_path(path)
        self.set_tile(tile)

    def set_model(self, vp, rho):
        vp, rho = f32(vp).reshape(-1), f32(rho).reshape(-1)
        _need(vp, (self.grid.ntx * self.grid.ntz,), "vp")
        _need(rho, (self.grid.ntx * self.grid.ntz,), "rho")
        check(self.L.swd_plan_set_model(self.h, _p(vp), _p(rho)))

    def set_path(self, path):
        check(self.L.swd_plan_set_path(self.h, int(path)))

    def set_groups(self, fwd_shots=0, bwd_shots=0):
        """Shots per fused RTM forward / backward loop (0 = sized to the cache)."""
        check(self.L.swd_plan_set_groups(self.h, int(fwd_shots), int(bwd_shots)))

    def set_memory_limit(self, nbytes):
        """Device bytes a call may use (0 = 85% of the free memory); larger batches run in waves."""
        L = self.L
        L.swd_plan_set_memory_limit.argtypes = [C.c_void_p, C.c_size_t]
        check(L.swd_plan_set_memory_limit(self.h, int(nbytes)))

    def set_graphs(self, on):
        """Replay fused time loops from captured hipGraphs (default on)."""
        check(self.L.swd_plan_set_graphs(self.h, 1 if on else 0))

    def set_streams(self, streams):
        """Groups of a fused loop (RTM, forward, AFor, Born, multi-grid) in flight at once on the
        plan's streams (1 to 4); results do not change."""
        check(self.L.swd_plan_set_streams(self.h, int(streams)))

    def set_tile(self, tile):
        """Fused-path tile width (16, 32, 64) or 0 to tune it on the first steps."""
        check(self.L.swd_plan_set_tile(self.h, int(tile)))

    def set_tile_config(self, tile_x=0, tile_rows=0, waves=0):
        """Fused-path tile: columns (16, 32, 64), rows (64; at Lc 4 also 32, 128)
        and a wavefronts-per-SIMD bound (0; at Lc 4 also 5); all 0 to tune."""
        check(self.L.swd_plan_set_tile_config(self.h, int(tile_x), int(tile_rows), int(waves)))

    def set_steps_per_launch(self, steps):
        """Time steps per launch of the fused forward and paired RTM backward at Lc 4:
        1, 2 (temporal blocking) or 0 to let the tuner pick; results do not change."""
        check(self.L.swd_plan_set_steps_per_launch(self.h, int(steps)))

    def close(self):
        if self.h:
            self.L.swd_plan_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def stats(self):
        s = Stats()
        check(self.L.swd_plan_get_stats(self.h, C.byref(s)))
        return s.as_dict()

    def _wav(self, w, name="wavelet"):
        return _need(f32(w).reshape(-1), (self.grid.nt,), name)

    def _traces(self, S, a, name):
        return _need(f32(a), (S.ntr, self.grid.nt), name)

    def forward(self, shots, wavelet):
        S = shots if isinstance(shots, ShotSet) else ShotSet(shots)
        g = np.zeros((S.ntr, self.grid.nt), np.float32)
        check(self.L.swd_forward(self.h, S.n, S.arr, _p(self._wav(wavelet)), _p(g)))
        return g

    def afor(self, shots, wavelet, hwavelet, kzcoef):
        S = shots if isinstance(shots, ShotSet) else ShotSet(shots)
        kz = f32(kzcoef)
        outs = [np.zeros((S.ntr, self.grid.nt), np.float32) for _ in range(3)]
        check(self.L.swd_afor(self.h, S.n, S.arr, _p(self._wav(wavelet)), _p(self._wav(hwavelet, "hwavelet")), _p(kz),
                              kz.size, *[_p(o) for o in outs]))
        return outs

    def afor_dev(self, shots, wavelet, hwavelet, kzcoef, gather_ptr, up_ptr, down_ptr):
        """Device-pointer AFor: the three [sum nrec][nt] gathers at the given addresses."""
        S = shots if isinstance(shots, ShotSet) else ShotSet(shots)
        kz = f32(kzcoef)
        vp = C.c_void_p
        check(self.L.swd_afor_dev(self.h, S.n, S.arr, _p(self._wav(wavelet)), _p(self._wav(hwavelet, "hwavelet")),
                                  _p(kz), kz.size,
                                  vp(int(gather_ptr)), vp(int(up_ptr)), vp(int(down_ptr))))

    def rtm(self, shots, wavelet, adj, want_gather=False):
        S = shots if isinstance(shots, ShotSet) else ShotSet(shots)
        ntp = self.grid.ntx * self.grid.ntz
        outs = [np.zeros((S.n, ntp), np.float32) for _ in range(3)]
        gs = np.zeros((S.ntr, self.grid.nt), np.float32) if want_gather else None
        check(self.L.swd_rtm(self.h, S.n, S.arr, _p(self._wav(wavelet)), _p(self._traces(S, adj, "adj")),
                             *[_p(o) for o in outs], _p(gs)))
        return (*outs, gs)

    def rtm_dev(self, shots, wavelet, adj_ptr, image_ptr, illum_s_ptr, illum_r_ptr, gather_ptr=None):
        """Device-pointer RTM (pointers are integers, e.g. torch tensor data_ptr())."""
        S = shots if isinstance(shots, ShotSet) else ShotSet(shots)
        vp = lambda x: None if x is None else C.c_void_p(int(x))
        check(self.L.swd_rtm_dev(self.h, S.n, S.arr, _p(self._wav(wavelet)), vp(adj_ptr), vp(image_ptr),
                                 vp(illum_s_ptr), vp(illum_r_ptr), vp(gather_ptr)))

    def rtm_cig(self, shots, wavelet, adj, thetan=10, delt=1):
        """RTM with angle-domain common-image gathers (swd_rtm_cig): image,
        illum_src, illum_rcv [nshots][ntx*ntz] and cig [nshots][ntx*ntz][thetan]."""
        S = shots if isinstance(shots, ShotSet) else ShotSet(shots)
        ntp = self.grid.ntx * self.grid.ntz
        outs = [np.zeros((S.n, ntp), np.float32) for _ in range(3)]
        cig = np.zeros((S.n, ntp, int(thetan)), np.float32)
        check(self.L.swd_rtm_cig(self.h, S.n, S.arr, _p(self._wav(wavelet)), _p(self._traces(S, adj, "adj")),
                                 int(thetan), int(delt),
                                 *[_p(o) for o in outs], _p(cig)))
        return (*outs, cig)

    def rtm_cig_dev(self, shots, wavelet, adj_ptr, thetan, delt, image_ptr, illum_s_ptr, illum_r_ptr, cig_ptr):
        """Device-pointer RTM with angle gathers (pointers are integers)."""
        S = shots if isinstance(shots, ShotSet) else ShotSet(shots)
        vp = lambda x: C.c_void_p(int(x))
        check(self.L.swd_rtm_cig_dev(self.h, S.n, S.arr, _p(self._wav(wavelet)), vp(adj_ptr), int(thetan), int(delt),
                                     vp(image_ptr), vp(illum_s_ptr), vp(illum_r_ptr), vp(cig_ptr)))

    def rtm_updown(self, shots, wavelet, hwavelet, adj, hadj, kzcoef, want_same_dir=False, want_image=False):
        """Up/down-separated RTM images (swd_rtm_updown): returns a dict of
        [nshots][ntx*ntz] arrays du, ud, and with want_same_dir uu, dd, with
        want_image the plain image, illum_src and illum_rcv of the same run.
        hwavelet and hadj are the Hilbert transforms along time of wavelet and adj."""
        S = shots if isinstance(shots, ShotSet) else ShotSet(shots)
        ntp = self.grid.ntx * self.grid.ntz
        names = ["du", "ud"] + (["uu", "dd"] if want_same_dir else []) + (
            ["image", "illum_src", "illum_rcv"] if want_image else [])
        outs = {k: np.zeros((S.n, ntp), np.float32) for k in names}
        imgs = UpDownImages(*[C.c_void_p(outs[k].ctypes.data) if k in outs else None for k in UpDownImages.NAMES])
        kz = f32(kzcoef)
        check(self.L.swd_rtm_updown(self.h, S.n, S.arr, _p(self._wav(wavelet)), _p(self._wav(hwavelet, "hwavelet")),
                                    _p(self._traces(S, adj, "adj")), _p(self._traces(S, hadj, "hadj")), _p(kz),
                                    kz.size, C.byref(imgs)))
        return outs

    def rtm_updown_dev(self, shots, wavelet, hwavelet, adj_ptr, hadj_ptr, kzcoef, out_ptrs):
        """Device-pointer up/down RTM; out_ptrs maps output names (UpDownImages.NAMES)
        to device pointers (integers, e.g. torch tensor data_ptr())."""
        S = shots if isinstance(shots, ShotSet) else ShotSet(shots)
        imgs = UpDownImages(*[C.c_void_p(int(out_ptrs[k])) if out_ptrs.get(k) else None for k in UpDownImages.NAMES])
        kz = f32(kzcoef)
        check(self.L.swd_rtm_updown_dev(self.h, S.n, S.arr, _p(self._wav(wavelet)), _p(self._wav(hwavelet, "hwavelet")),
                                        C.c_void_p(int(adj_ptr)), C.c_void_p(int(hadj_ptr)), _p(kz), kz.size,
                                        C.byref(imgs)))

    def updown_image_dev(self, nshots, src_ptr, hsrc_ptr, rcv_ptr, hrcv_ptr, kzcoef, out_ptrs):
        """One step of the up/down imaging condition on device fields [nshots][ntx*ntz]
        (swd_updown_image_dev); queued on the plan's stream."""
        imgs = UpDownImages(*[C.c_void_p(int(out_ptrs[k])) if out_ptrs.get(k) else None for k in UpDownImages.NAMES])
        kz = f32(kzcoef)
        vp = lambda x: C.c_void_p(int(x))
        check(self.L.swd_updown_image_dev(self.h, int(nshots), vp(src_ptr), vp(hsrc_ptr), vp(rcv_ptr), vp(hrcv_ptr),
                                          _p(kz), kz.size, C.byref(imgs)))

    def updown_image_dev_noalloc(self, nshots, src_ptr, hsrc_ptr, rcv_ptr, hrcv_ptr, kzcoef, out_ptrs):
        """The same step with its tables in the library's fixed per-GPU array
        (swd_updown_image_dev_noalloc): no device allocation, at most
        SWD_NOALLOC_MAX_SHOTS shots; calls overlapping on one GPU race."""
        imgs = UpDownImages(*[C.c_void_p(int(out_ptrs[k])) if out_ptrs.get(k) else None for k in UpDownImages.NAMES])
        kz = f32(kzcoef)
        vp = lambda x: C.c_void_p(int(x))
        check(self.L.swd_updown_image_dev_noalloc(self.h, int(nshots), vp(src_ptr), vp(hsrc_ptr), vp(rcv_ptr),
                                                  vp(hrcv_ptr), _p(kz), kz.size, C.byref(imgs)))

    def stream(self):
        """The plan's hipStream_t as an integer (e.g. for torch.cuda.ExternalStream)."""
        s = C.c_void_p()
        check(self.L.swd_plan_get_stream(self.h, C.byref(s)))
        return s.value or 0

    def born(self, shots, wavelet, m):
        S = shots if isinstance(shots, ShotSet) else ShotSet(shots)
        m = _need(f32(m).reshape(-1), (S.n * self.grid.nx * self.grid.nz,), "m")
        d = np.zeros((S.ntr, self.grid.nt), np.float32)
        check(self.L.swd_born(self.h, S.n, S.arr, _p(self._wav(wavelet)), _p(m), _p(d)))
        return d

    def born_adjoint(self, shots, wavelet, d, nsnap=0):
        S = shots if isinstance(shots, ShotSet) else ShotSet(shots)
        g = np.zeros(self.grid.nx * self.grid.nz, np.float32)
        check(self.L.swd_born_adjoint(self.h, S.n, S.arr, _p(self._wav(wavelet)), _p(self._traces(S, d, "d")), _p(g),
                                      int(nsnap)))
        return g

    def born_dev(self, shots, wavelet, m_ptr, d_ptr):
        """Device-pointer Born forward: m [nshots][nx][nz] -> d [sum nrec][nt]."""
        S = shots if isinstance(shots, ShotSet) else ShotSet(shots)
        check(self.L.swd_born_dev(self.h, S.n, S.arr, _p(self._wav(wavelet)), C.c_void_p(int(m_ptr)),
                                  C.c_void_p(int(d_ptr))))

    def born_adjoint_dev(self, shots, wavelet, d_ptr, g_ptr, nsnap=0):
        """Device-pointer Born adjoint: d [sum nrec][nt] -> m_grad [nx][nz] (summed over shots)."""
        S = shots if isinstance(shots, ShotSet) else ShotSet(shots)
        check(self.L.swd_born_adjoint_dev(self.h, S.n, S.arr, _p(self._wav(wavelet)), C.c_void_p(int(d_ptr)),
                                          C.c_void_p(int(g_ptr)), int(nsnap)))


def multi_forward(plans, shotsets, wavelets):
    """Forward modelling of several plans (uneven grids, same Lc and nt) in
    one fused launch per time step; returns one gather per plan."""
    n = len(plans)
    L = load()
    sets = [S if isinstance(S, ShotSet) else ShotSet(S) for S in shotsets]
    wav = [_need(f32(w).reshape(-1), (plans[i].grid.nt,), "wavelet") for i, w in enumerate(wavelets)]
    outs = [np.zeros((S.ntr, plans[i].grid.nt), np.float32) for i, S in enumerate(sets)]
    arr_p = (C.c_void_p * n)(*[p.h.value for p in plans])
    arr_n = (C.c_int * n)(*[S.n for S in sets])
    arr_s = (C.c_void_p * n)(*[C.cast(S.arr, C.c_void_p).value for S in sets])
    arr_w = (C.c_void_p * n)(*[w.ctypes.data for w in wav])
    arr_g = (C.c_void_p * n)(*[o.ctypes.data for o in outs])
    check(L.swd_multi_forward(n, arr_p, arr_n, arr_s, arr_w, arr_g))
    return outs


# ---------------------------------------------------------------------------
# Wavenumber-domain PSF deconvolution (kkdecon_acoustic/kk_decon_interp_psf_minewindow.cpp)

class DeconParams(C.Structure):
    _fields_ = [(n, C.c_int) for n in ("nx", "nz", "psfwinx", "psfwinz", "filterwinx", "filterwinz",
                                       "snapwinx", "snapwinz")] + [("epsilon", C.c_float)]


class DeconFiles(C.Structure):
    _fields_ = [(n, C.c_char * 256) for n in ("vp", "vp0", "rho", "image", "psf", "decon", "inv_psf")] + \
        [("shift_flag", C.c_int), ("reg_eps", C.c_float)]


class DeconGeom(C.Structure):
    _fields_ = [(n, C.c_int) for n in ("n", "outwinx", "outwinz", "stepwinx", "stepwinz", "stepnumx", "stepnumz",
                                       "windows", "nxp", "nzp", "cells")]


class DeconStats(C.Structure):
    _fields_ = [(n, C.c_double) for n in ("sums_ms", "cells_ms", "windows_ms", "gather_ms", "total_ms")] + \
        [(n, C.c_int) for n in ("windows", "cells", "n")] + [("ws_bytes", C.c_longlong)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


def decon_params(nx, nz, psfwin=31, filterwin=71, snapwin=11, epsilon=5e-4):
    """Parameters of the reference's kkdecon parameter.txt (square windows by default)."""
    return DeconParams(int(nx), int(nz), psfwin, psfwin, filterwin, filterwin, snapwin, snapwin, float(epsilon))


def decon_read_parameters(path):
    """(DeconParams, dict of file names and flags) from a kkdecon parameter.txt."""
    p, f = DeconParams(), DeconFiles()
    check(load().swd_decon_read_parameters(path.encode(), C.byref(p), C.byref(f)))
    files = {k: getattr(f, k).decode() for k in ("vp", "vp0", "rho", "image", "psf", "decon", "inv_psf")}
    files.update(shift_flag=f.shift_flag, reg_eps=f.reg_eps)
    return p, files


def decon_geometry(p):
    g = DeconGeom()
    check(load().swd_decon_geometry(C.byref(p), C.byref(g)))
    return {k: getattr(g, k) for k, _ in g._fields_}


def decon_taper(p):
    g = decon_geometry(p)
    c = np.zeros((g["outwinx"], g["outwinz"]), np.float32)
    check(load().swd_decon_taper(C.byref(p), _p(c)))
    return c


class Decon:
    """A deconvolution plan: twiddle tables, taper and workspace of one
    parameter set on one GPU."""

    def __init__(self, params, device=0):
        self.L = load()
        self.params = params
        self.device = int(device)
        h = C.c_void_p()
        check(self.L.swd_decon_plan_create(C.byref(params), self.device, C.byref(h)))
        self.h = h

    def run(self, image, psf):
        """(dec, coef) [nx][nz] float32 of host arrays image, psf [nx][nz]."""
        p = self.params
        image, psf = f32(image), f32(psf)
        assert image.shape == (p.nx, p.nz) and psf.shape == (p.nx, p.nz)
        dec = np.zeros((p.nx, p.nz), np.float32)
        coef = np.zeros((p.nx, p.nz), np.float32)
        check(self.L.swd_decon(self.h, _p(image), _p(psf), _p(dec), _p(coef)))
        return dec, coef

    def run_dev(self, image_ptr, psf_ptr, dec_ptr, coef_ptr):
        """Same on device arrays of the plan's GPU (raw pointers)."""
        check(self.L.swd_decon_dev(self.h, C.c_void_p(image_ptr), C.c_void_p(psf_ptr), C.c_void_p(dec_ptr),
                                   C.c_void_p(coef_ptr)))

    def stats(self):
        s = DeconStats()
        check(self.L.swd_decon_get_stats(self.h, C.byref(s)))
        return s.as_dict()

    def close(self):
        if self.h:
            self.L.swd_decon_plan_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
