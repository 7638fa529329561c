"""ctypes binding of libgabe_amd.so (include/gabe.h) with reference-shaped names.

The reference drives one lattice through globals and the calls step(),
calcEnergy(s), adf(s), fft_stresstensor(s) and specOut() (g2header.h:127-177).
`Lattice` exposes the same operations on a device-resident lattice.  There is
no CPU fallback: a missing or unloadable library raises.
"""
import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libgabe_amd.so")

GABE_OK = 0
GABE_ERR_ARG = -1
GABE_ERR_HIP = -2
GABE_ERR_FFT = -3
GABE_ERR_NONFINITE = -4
GABE_ERR_STATE = -5
GABE_FP64 = 0
GABE_FP32 = 1
GABE_NKCLASS = 5
KERNEL_CLASSES = ("sweep_a", "finalize", "stress_fft", "gw_modes", "sweep_b")

EXPORTS = (
    "gabe_create", "gabe_destroy", "gabe_last_error", "gabe_numbins",
    "gabe_upload_state", "gabe_download_state", "gabe_download_plane", "gabe_means_vars", "gabe_means_vars_noalloc",
    "gabe_upload_gw", "gabe_download_gw", "gabe_checkpoint_size", "gabe_checkpoint_save", "gabe_checkpoint_load",
    "gabe_set_scalars", "gabe_get_scalars", "gabe_energy", "gabe_step", "gabe_sync",
    "gabe_spectra", "gabe_add_fluctuation", "gabe_effmass", "gabe_tt_source",
    "gabe_history", "gabe_kernel_times", "gabe_query", "gabe_autotune_sweep", "gabe_mode_tiles",
    "gabe_ensemble_create", "gabe_ensemble_member", "gabe_ensemble_step", "gabe_ensemble_sync",
    "gabe_ensemble_kernel_times",
    "gabe_ensemble_destroy",
    "gabe_restrict", "gabe_closure_create", "gabe_closure_destroy", "gabe_closure_params",
    "gabe_closure_set_params", "gabe_closure_scales", "gabe_closure_set_scales", "gabe_closure_sample",
    "gabe_closure_num_samples", "gabe_closure_train", "gabe_closure_loss", "gabe_closure_forward",
    "gabe_closure_gradient", "gabe_closure_attach", "gabe_closure_detach", "gabe_closure_status",
    "gabe_closure_kernel_times",
)


class GabeParams(C.Structure):
    _fields_ = [
        ("N", C.c_int), ("nflds", C.c_int), ("L", C.c_double), ("dt", C.c_double),
        ("calc_gws", C.c_int), ("expansion_type", C.c_int), ("precision", C.c_int),
        ("m2_phi", C.c_double), ("lam", C.c_double), ("g2", C.c_double), ("m2_chi", C.c_double),
        ("gw_spectrum_quirk", C.c_int), ("planes_per_block", C.c_int), ("transform_tile", C.c_int),
        ("mode_tile", C.c_int), ("stress_chunk", C.c_int), ("reserved", C.c_int * 3),
    ]


class GabeScalars(C.Structure):
    _fields_ = [
        ("a", C.c_double * 2), ("adot", C.c_double * 2), ("edkin", C.c_double * 2),
        ("edpot", C.c_double * 2), ("edgrad", C.c_double * 2), ("edrho", C.c_double * 2),
        ("t", C.c_double), ("nstep", C.c_longlong), ("adot0_valid", C.c_int), ("error", C.c_int),
    ]


class GabeError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"gabe error {code}: {msg}")
        self.code = code


_lib = None


def load():
    """Load libgabe_amd.so (raises if it is missing: there is no fallback)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(f"{LIB_PATH} is not built; run __graft_entry__.build()")
    L = C.CDLL(LIB_PATH)
    vp, dp, ip = C.c_void_p, C.POINTER(C.c_double), C.c_int
    fp = C.POINTER(C.c_float)
    sig = {
        "gabe_create": [C.POINTER(GabeParams), C.POINTER(vp)],
        "gabe_destroy": [vp],
        "gabe_numbins": [ip],
        "gabe_upload_state": [vp, dp, dp],
        "gabe_download_state": [vp, dp, dp],
        "gabe_download_plane": [vp, ip, dp, dp],
        "gabe_means_vars": [vp, dp],
        "gabe_means_vars_noalloc": [vp, ip, C.c_longlong, ip, dp, vp, dp],
        "gabe_checkpoint_size": [vp, C.POINTER(C.c_size_t)],
        "gabe_checkpoint_save": [vp, vp, C.c_size_t],
        "gabe_checkpoint_load": [vp, vp, C.c_size_t],
        "gabe_upload_gw": [vp, dp, dp],
        "gabe_download_gw": [vp, dp, dp],
        "gabe_set_scalars": [vp, C.POINTER(GabeScalars)],
        "gabe_get_scalars": [vp, C.POINTER(GabeScalars)],
        "gabe_energy": [vp, ip],
        "gabe_step": [vp, ip],
        "gabe_sync": [vp],
        "gabe_spectra": [vp, dp],
        "gabe_add_fluctuation": [vp, ip, dp, dp],
        "gabe_effmass": [vp, ip, dp],
        "gabe_tt_source": [vp, ip, dp],
        "gabe_history": [vp, dp, ip],
        "gabe_kernel_times": [vp, ip, dp],
        "gabe_query": [vp, C.POINTER(GabeParams)],
        "gabe_autotune_sweep": [C.POINTER(GabeParams), C.POINTER(C.c_int), dp],
        "gabe_mode_tiles": [ip, ip],
        "gabe_ensemble_create": [C.POINTER(GabeParams), ip, C.POINTER(vp)],
        "gabe_ensemble_member": [vp, ip, C.POINTER(vp)],
        "gabe_ensemble_step": [vp, ip],
        "gabe_ensemble_sync": [vp],
        "gabe_ensemble_kernel_times": [vp, ip, dp],
        "gabe_ensemble_destroy": [vp],
        "gabe_restrict": [vp, vp],
        "gabe_closure_create": [ip, C.c_double, C.c_uint, C.POINTER(vp)],
        "gabe_closure_destroy": [vp],
        "gabe_closure_params": [vp, fp],
        "gabe_closure_set_params": [vp, fp],
        "gabe_closure_scales": [vp, dp, dp],
        "gabe_closure_set_scales": [vp, dp, dp],
        "gabe_closure_sample": [vp, vp, vp, ip, ip],
        "gabe_closure_num_samples": [vp],
        "gabe_closure_train": [vp, ip, C.c_double, C.c_uint, dp],
        "gabe_closure_loss": [vp, ip, dp],
        "gabe_closure_forward": [vp, fp, fp],
        "gabe_closure_gradient": [vp, fp, fp, fp, dp],
        "gabe_closure_attach": [vp, vp, C.c_double],
        "gabe_closure_detach": [vp],
        "gabe_closure_status": [vp, C.POINTER(C.c_int), C.POINTER(C.c_longlong), dp],
        "gabe_closure_kernel_times": [vp, ip, dp],
    }
    for name, args in sig.items():
        fn = getattr(L, name)
        fn.argtypes = args
        fn.restype = C.c_int
    L.gabe_last_error.restype = C.c_char_p
    L.gabe_last_error.argtypes = []
    _lib = L
    return L


def _dp(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


def _check(rc):
    if rc < 0:
        raise GabeError(rc, load().gabe_last_error().decode())
    return rc


def numbins(N):
    return load().gabe_numbins(N)


def make_params(N, nflds=2, L=10.0, dt=None, calc_gws=0, expansion_type=1, precision=GABE_FP64, m2_phi=1.0,
                lam=0.0, g2=0.0, m2_chi=0.0, gw_spectrum_quirk=1, planes_per_block=0, transform_tile=0,
                mode_tile=0, stress_chunk=0):
    if dt is None:
        dt = L / N / 20.0
    return GabeParams(N, nflds, L, dt, calc_gws, expansion_type, precision, m2_phi, lam, g2, m2_chi,
                      gw_spectrum_quirk, planes_per_block, transform_tile, mode_tile, stress_chunk)


def means_vars_noalloc(dev_ptr, precision, nsites, nflds, base=None, stream=None):
    """gabe_means_vars_noalloc on fields at device address dev_ptr (nflds
    arrays of nsites doubles or floats); stream is a hipStream_t handle
    (None: the null stream).  Returns [mean, variance] per field."""
    out = np.zeros(2 * nflds)
    b = None if base is None else np.ascontiguousarray(base, dtype=np.float64)
    _check(load().gabe_means_vars_noalloc(C.c_void_p(dev_ptr), precision, nsites, nflds,
                                           None if b is None else _dp(b), C.c_void_p(stream or 0), _dp(out)))
    return out


def mode_tiles(N, precision):
    """number of tile variants of the fused i-transform + mode kernel"""
    return load().gabe_mode_tiles(N, precision)


def autotune_sweep(**kw):
    """Time the candidate sweep tilings for these parameters (and, with
    tensor modes, the j-transform widths and mode-kernel variants); later
    lattices of the same (N, nflds, precision, calc_gws) use the fastest.
    Returns (planes per workgroup, ms per step)."""
    p = make_params(**kw)
    ppb = C.c_int()
    ms = C.c_double()
    _check(load().gabe_autotune_sweep(C.byref(p), C.byref(ppb), C.byref(ms)))
    return ppb.value, ms.value


class Lattice:
    """One device-resident lattice (a gabe_ctx).

    Host arrays use the reference layouts: fields [nflds][N][N][N], tensor
    perturbations [12][N][N][N/2+1] with planes (11,12,13,22,23,33) x (Re,Im).
    """

    def __init__(self, N, nflds=2, L=10.0, dt=None, calc_gws=0, expansion_type=1,
                 precision=GABE_FP64, m2_phi=1.0, lam=0.0, g2=0.0, m2_chi=0.0,
                 gw_spectrum_quirk=1, planes_per_block=0, transform_tile=0, mode_tile=0, stress_chunk=0,
                 _handle=None):
        self.params = make_params(N, nflds, L, dt, calc_gws, expansion_type, precision, m2_phi, lam, g2, m2_chi,
                                  gw_spectrum_quirk, planes_per_block, transform_tile, mode_tile, stress_chunk)
        self.N, self.nflds, self.calc_gws = N, nflds, calc_gws
        self.M = N * N * (N // 2 + 1)
        self._owned = _handle is None
        if _handle is not None:
            self._h = _handle
            return
        h = C.c_void_p()
        _check(load().gabe_create(C.byref(self.params), C.byref(h)))
        self._h = h

    def close(self):
        if getattr(self, "_owned", True) and getattr(self, "_h", None) and self._h.value:
            load().gabe_destroy(self._h)
        self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    # -- state -------------------------------------------------------------
    def upload_state(self, field, dfield):
        f = np.ascontiguousarray(field, dtype=np.float64).reshape(self.nflds, self.N, self.N, self.N)
        d = np.ascontiguousarray(dfield, dtype=np.float64).reshape(f.shape)
        _check(load().gabe_upload_state(self._h, _dp(f), _dp(d)))

    def download_state(self):
        shp = (self.nflds, self.N, self.N, self.N)
        f = np.empty(shp)
        d = np.empty(shp)
        _check(load().gabe_download_state(self._h, _dp(f), _dp(d)))
        return f, d

    def download_plane(self, i):
        """Plane i of the state: (field, dfield), each [nflds][N][N]."""
        shp = (self.nflds, self.N, self.N)
        f = np.empty(shp)
        d = np.empty(shp)
        _check(load().gabe_download_plane(self._h, int(i), _dp(f), _dp(d)))
        return f, d

    def meansvars(self):
        """[<f0>, var f0, <f1>, var f1] as in g2output.cpp:70-121."""
        out = np.empty(2 * self.nflds)
        _check(load().gabe_means_vars(self._h, _dp(out)))
        return out

    def checkpoint(self):
        """The context's state as a bytes-like uint8 array (gabe_checkpoint_save)."""
        n = C.c_size_t()
        _check(load().gabe_checkpoint_size(self._h, C.byref(n)))
        buf = np.empty(n.value, dtype=np.uint8)
        _check(load().gabe_checkpoint_save(self._h, buf.ctypes.data, n.value))
        return buf

    def restore(self, blob):
        b = np.ascontiguousarray(np.frombuffer(blob, dtype=np.uint8))
        _check(load().gabe_checkpoint_load(self._h, b.ctypes.data, b.size))

    def upload_gw(self, h=None, l=None):
        shp = (12, self.N, self.N, self.N // 2 + 1)
        hh = None if h is None else np.ascontiguousarray(h, dtype=np.float64).reshape(shp)
        ll = None if l is None else np.ascontiguousarray(l, dtype=np.float64).reshape(shp)
        _check(load().gabe_upload_gw(self._h, None if hh is None else _dp(hh), None if ll is None else _dp(ll)))

    def download_gw(self):
        shp = (12, self.N, self.N, self.N // 2 + 1)
        h = np.empty(shp)
        l = np.empty(shp)
        _check(load().gabe_download_gw(self._h, _dp(h), _dp(l)))
        return h, l

    def set_scalars(self, a, adot, edkin=(0, 0), edpot=(0, 0), edgrad=(0, 0), edrho=(0, 0), t=0.0, nstep=0):
        s = GabeScalars()
        for k in range(2):
            s.a[k], s.adot[k] = a[k], adot[k]
            s.edkin[k], s.edpot[k], s.edgrad[k], s.edrho[k] = edkin[k], edpot[k], edgrad[k], edrho[k]
        s.t, s.nstep = t, nstep
        _check(load().gabe_set_scalars(self._h, C.byref(s)))

    def get_scalars(self):
        s = GabeScalars()
        rc = load().gabe_get_scalars(self._h, C.byref(s))
        out = {k: [getattr(s, k)[0], getattr(s, k)[1]] for k in ("a", "adot", "edkin", "edpot", "edgrad", "edrho")}
        out.update(t=s.t, nstep=s.nstep, adot0_valid=s.adot0_valid, error=s.error)
        _check(rc)
        return out

    # -- reference operations ---------------------------------------------
    def calcEnergy(self, s=0):
        _check(load().gabe_energy(self._h, s))

    def step(self, nsteps=1):
        _check(load().gabe_step(self._h, nsteps))

    def sync(self):
        _check(load().gabe_sync(self._h))

    def specOut(self):
        nb = numbins(self.N)
        ncol = 2 * self.nflds + (3 if self.calc_gws else 0)
        out = np.zeros(ncol * nb)
        _check(load().gabe_spectra(self._h, _dp(out)))
        return out.reshape(ncol, nb)

    def add_fluctuation(self, fld, fk, fkd):
        fk = np.ascontiguousarray(fk, dtype=np.complex128)
        fkd = np.ascontiguousarray(fkd, dtype=np.complex128)
        _check(load().gabe_add_fluctuation(self._h, fld, _dp(fk.view(np.float64)), _dp(fkd.view(np.float64))))

    def effMass(self, fld):
        out = C.c_double()
        _check(load().gabe_effmass(self._h, fld, C.byref(out)))
        return out.value

    def tt_source(self, s):
        out = np.empty((12, self.N, self.N, self.N // 2 + 1))
        _check(load().gabe_tt_source(self._h, s, _dp(out)))
        return out

    def history(self, max_rows=1 << 16):
        out = np.zeros((max_rows, 8))
        n = _check(load().gabe_history(self._h, _dp(out), max_rows))
        return out[:n]

    def kernel_times(self, nsteps):
        out = np.zeros(GABE_NKCLASS)
        _check(load().gabe_kernel_times(self._h, nsteps, _dp(out)))
        return dict(zip(KERNEL_CLASSES, out.tolist()))

    def query(self):
        p = GabeParams()
        _check(load().gabe_query(self._h, C.byref(p)))
        return p


class Ensemble:
    """Independent lattices stepped together (gabe_ensemble_*).  members[i] is
    a Lattice view of member i, usable with every per-lattice call."""

    def __init__(self, member_kwargs):
        arr = (GabeParams * len(member_kwargs))(*[make_params(**kw) for kw in member_kwargs])
        h = C.c_void_p()
        _check(load().gabe_ensemble_create(arr, len(member_kwargs), C.byref(h)))
        self._h = h
        self.members = []
        for i, kw in enumerate(member_kwargs):
            mh = C.c_void_p()
            _check(load().gabe_ensemble_member(self._h, i, C.byref(mh)))
            self.members.append(Lattice(_handle=mh, **kw))

    def step(self, nsteps=1):
        _check(load().gabe_ensemble_step(self._h, nsteps))

    def sync(self):
        _check(load().gabe_ensemble_sync(self._h))

    def kernel_times(self, nsteps):
        """mean ms of the batched sweep A, finalize and sweep B over nsteps steps"""
        out = np.zeros(3)
        _check(load().gabe_ensemble_kernel_times(self._h, nsteps, _dp(out)))
        return dict(zip(("sweep_a", "finalize", "sweep_b"), out.tolist()))

    def close(self):
        if getattr(self, "_h", None) and self._h.value:
            for m in self.members:
                m._h = C.c_void_p()
            load().gabe_ensemble_destroy(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
