"""ctypes binding of libgabe_host.so: the kept host program's start-up
(vacuum initial conditions drawn with libc rand(), g2init.cpp / g2model.cpp)."""
import ctypes as C
import os

from . import lattice as _lat

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libgabe_host.so")


class HostInit(C.Structure):
    _fields_ = [("f0", C.c_double * 2), ("df0", C.c_double * 2), ("rescale_B", C.c_double),
                ("randseed", C.c_uint), ("rand_init", C.c_int), ("symmetric_modes", C.c_int),
                ("spec_cut_off", C.c_double), ("spec_smooth", C.c_double)]


_lib = None


def load():
    global _lib
    if _lib is None:
        _lat.load()
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} is not built; run __graft_entry__.build()")
        L = C.CDLL(LIB_PATH)
        L.gabe_host_initialize.argtypes = [C.c_void_p, C.POINTER(_lat.GabeParams), C.POINTER(HostInit)]
        L.gabe_host_initialize.restype = C.c_int
        L.gabe_host_friedmann_residual.argtypes = [C.POINTER(C.c_double), C.c_int, C.POINTER(C.c_double)]
        L.gabe_host_friedmann_residual.restype = C.c_double
        _lib = L
    return _lib


def initialize(lat, cfg, rand_init=1):
    """Run the reference start-up sequence on a device lattice."""
    init = HostInit()
    for f in range(cfg.nflds):
        init.f0[f] = cfg.f0[f]
        init.df0[f] = cfg.df0[f]
    init.rescale_B = cfg.rescale_B
    init.randseed = cfg.randseed
    init.rand_init = rand_init
    init.symmetric_modes = 0 if cfg.field_full_rand else 1
    init.spec_cut_off = cfg.spec_cut_off
    init.spec_smooth = cfg.spec_smooth
    _lat._check(load().gabe_host_initialize(lat._h, C.byref(lat.params), C.byref(init)))


def friedmann_residual(hist):
    """R(t) = |rho_dot + 3H(rho + p)| / (3 H rho) per history row (0 at the ends)."""
    import numpy as np
    h = np.ascontiguousarray(hist, dtype=np.float64)
    out = np.zeros(h.shape[0])
    worst = load().gabe_host_friedmann_residual(h.ctypes.data_as(C.POINTER(C.c_double)), h.shape[0],
                                                out.ctypes.data_as(C.POINTER(C.c_double)))
    return worst, out


def initialize_member(lat, cfg, rand_init=1):
    """Reference start-up sequence on an existing lattice (e.g. an ensemble member)."""
    initialize(lat, cfg, rand_init)


def initialized_lattice(cfg, precision=_lat.GABE_FP32, **kw):
    lat = _lat.Lattice(precision=precision, **cfg.lattice_kwargs(), **kw)
    initialize(lat, cfg)
    return lat
