"""ctypes view of the CPU FP64 oracle (liboracle_gabe.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
CPU-baseline leg of bench.py, always as the checker.  The product never
imports it.
"""
import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "liboracle_gabe.so")


class OrcParams(C.Structure):
    _fields_ = [
        ("N", C.c_int), ("nflds", C.c_int), ("L", C.c_double), ("dt", C.c_double),
        ("calc_gws", C.c_int), ("expansion_type", C.c_int),
        ("m2_phi", C.c_double), ("lam", C.c_double), ("g2", C.c_double), ("m2_chi", C.c_double),
        ("rescale_B", C.c_double), ("nthreads", C.c_int),
    ]


class OrcState(C.Structure):
    _fields_ = [
        ("p", OrcParams), ("dx", C.c_double), ("N3", C.c_long), ("M", C.c_long),
        ("field", C.POINTER(C.c_double)), ("dfield", C.POINTER(C.c_double)),
        ("a", C.c_double * 2), ("adot", C.c_double * 2), ("edkin", C.c_double * 2),
        ("edpot", C.c_double * 2), ("edgrad", C.c_double * 2), ("edrho", C.c_double * 2),
        ("h", C.POINTER(C.c_double)), ("l", C.POINTER(C.c_double)), ("T_gw", C.POINTER(C.c_double)),
        ("t", C.c_double),
    ]


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            raise RuntimeError("oracle library not built: run `make -C oracle`")
        L = C.CDLL(_LIB_PATH)
        P = C.POINTER(OrcState)
        dp = C.POINTER(C.c_double)
        sig = {
            "orc_create": (P, [C.POINTER(OrcParams)]),
            "orc_destroy": (None, [P]),
            "orc_numbins": (C.c_int, [C.c_int]),
            "orc_laplacian": (C.c_double, [P, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int]),
            "orc_potential": (C.c_double, [P, C.c_int, C.c_int, C.c_int, C.c_int]),
            "orc_dVdf": (C.c_double, [P, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int]),
            "orc_effMass": (C.c_double, [P, C.c_int, C.c_int]),
            "orc_avgGrad": (C.c_double, [P, C.c_int]),
            "orc_avgPot": (C.c_double, [P, C.c_int]),
            "orc_avgKin": (C.c_double, [P, C.c_int]),
            "orc_calcEnergy": (None, [P, C.c_int]),
            "orc_adf": (C.c_double, [P, C.c_int]),
            "orc_ddfield": (C.c_double, [P, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int]),
            "orc_step": (None, [P]),
            "orc_gradF2": (C.c_double, [P, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int]),
            "orc_calcrho": (C.c_double, [P, C.c_int, C.c_int, C.c_int, C.c_int]),
            "orc_stresstensor": (C.c_double, [P, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int]),
            "orc_fft_stresstensor": (None, [P, C.c_int]),
            "orc_gw_update": (None, [P, C.c_int, C.c_int, C.c_double]),
            "orc_evolve_perts": (None, [P, C.c_int, C.c_int, C.c_double]),
            "orc_specout": (None, [P, dp, C.c_int]),
            "orc_srand": (None, [C.c_uint]),
            "orc_rand": (C.c_int, []),
            "orc_initfields_homog": (None, [P, dp, dp]),
            "orc_initexpansion": (None, [P]),
            "orc_randinit": (None, [P, C.c_int, C.c_double, dp, dp]),
            "orc_randinit_opt": (None, [P, C.c_int, C.c_double, C.c_int, C.c_double, C.c_double, dp, dp]),
            "orc_initfields_fluct": (None, [P]),
            "orc_meansvars": (None, [P, dp]),
            "orc_r2c3d": (None, [C.c_int, dp, dp, C.c_int]),
            "orc_c2r3d": (None, [C.c_int, dp, dp, C.c_int]),
            "orc_fft1d": (None, [C.c_int, dp, dp, C.c_int]),
        }
        for name, (res, args) in sig.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


def _dp(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


def r2c3d(x, nthreads=4):
    x = np.ascontiguousarray(x, dtype=np.float64)
    N = x.shape[0]
    out = np.empty((N, N, N // 2 + 1), dtype=np.complex128)
    lib().orc_r2c3d(N, _dp(x), _dp(out.view(np.float64)), nthreads)
    return out


def c2r3d(X, nthreads=4):
    X = np.ascontiguousarray(X, dtype=np.complex128)
    N = X.shape[0]
    out = np.empty((N, N, N), dtype=np.float64)
    lib().orc_c2r3d(N, _dp(X.view(np.float64)), _dp(out), nthreads)
    return out


def fft1d(x, sign=-1):
    x = np.ascontiguousarray(x, dtype=np.complex128)
    out = np.empty_like(x)
    lib().orc_fft1d(x.shape[0], _dp(x.view(np.float64)), _dp(out.view(np.float64)), sign)
    return out


def numbins(N):
    return lib().orc_numbins(N)


class Oracle:
    """One lattice in the oracle.  Array views alias the C state."""

    def __init__(self, N, nflds=2, L=10.0, dt=None, calc_gws=0, expansion_type=1,
                 m2_phi=1.0, lam=0.0, g2=0.0, m2_chi=0.0, rescale_B=1.0, nthreads=4):
        if dt is None:
            dt = L / N / 20.0
        self.params = OrcParams(N, nflds, L, dt, calc_gws, expansion_type,
                                m2_phi, lam, g2, m2_chi, rescale_B, nthreads)
        self._st = lib().orc_create(C.byref(self.params))
        st = self._st.contents
        self.N, self.nflds, self.M = N, nflds, st.M
        n3 = N * N * N
        self.field = np.ctypeslib.as_array(st.field, shape=(2 * nflds * n3,)).reshape(2, nflds, N, N, N)
        self.dfield = np.ctypeslib.as_array(st.dfield, shape=(2 * nflds * n3,)).reshape(2, nflds, N, N, N)
        if calc_gws:
            shp = (2, 12, N, N, N // 2 + 1)
            cnt = 2 * 12 * st.M
            self.h = np.ctypeslib.as_array(st.h, shape=(cnt,)).reshape(shp)
            self.l = np.ctypeslib.as_array(st.l, shape=(cnt,)).reshape(shp)
            self.T_gw = np.ctypeslib.as_array(st.T_gw, shape=(cnt,)).reshape(shp)
        else:
            self.h = self.l = self.T_gw = None

    def __del__(self):
        try:
            if self._st:
                lib().orc_destroy(self._st)
                self._st = None
        except Exception:
            pass

    @property
    def st(self):
        return self._st.contents

    @property
    def dx(self):
        return self.st.dx

    def scalar(self, name, s):
        return getattr(self.st, name)[s]

    def set_scalar(self, name, s, v):
        getattr(self.st, name)[s] = v

    def scalars(self):
        st = self.st
        return {k: [getattr(st, k)[0], getattr(st, k)[1]]
                for k in ("a", "adot", "edkin", "edpot", "edgrad", "edrho")}

    def step(self, n=1):
        for _ in range(n):
            lib().orc_step(self._st)

    def calcEnergy(self, s):
        lib().orc_calcEnergy(self._st, s)

    def adf(self, s):
        return lib().orc_adf(self._st, s)

    def laplacian(self, s, fld, i, j, k):
        return lib().orc_laplacian(self._st, s, fld, i, j, k)

    def stresstensor(self, aa, bb, s, i, j, k):
        return lib().orc_stresstensor(self._st, aa, bb, s, i, j, k)

    def fft_stresstensor(self, s):
        lib().orc_fft_stresstensor(self._st, s)

    def gw_update(self, s, snew, deet):
        lib().orc_gw_update(self._st, s, snew, deet)

    def evolve_perts(self, s, snew, deet):
        lib().orc_evolve_perts(self._st, s, snew, deet)

    def specout(self, gw_quirk=1):
        nb = numbins(self.N)
        ncol = (2 * self.nflds + 3) if self.params.calc_gws else 2 * self.nflds
        out = np.zeros(ncol * nb)
        lib().orc_specout(self._st, _dp(out), gw_quirk)
        return out.reshape(ncol, nb)

    def effMass(self, s, fld):
        return lib().orc_effMass(self._st, s, fld)

    def initfields_homog(self, f0, df0):
        f0 = np.asarray(f0, dtype=np.float64)
        df0 = np.asarray(df0, dtype=np.float64)
        lib().orc_initfields_homog(self._st, _dp(f0), _dp(df0))

    def initexpansion(self):
        lib().orc_initexpansion(self._st)

    def randinit(self, fld, d2vdf2, want_modes=False, full_rand=1, cut_off=0.0, smooth=0.0):
        """g2init.cpp:68-189; cut_off <= 0 leaves spec_cut_off undefined."""
        if want_modes:
            fk = np.zeros((self.N, self.N, self.N // 2 + 1), dtype=np.complex128)
            fkd = np.zeros_like(fk)
            lib().orc_randinit_opt(self._st, fld, d2vdf2, full_rand, cut_off, smooth,
                                   _dp(fk.view(np.float64)), _dp(fkd.view(np.float64)))
            return fk, fkd
        lib().orc_randinit_opt(self._st, fld, d2vdf2, full_rand, cut_off, smooth, None, None)
        return None

    def initfields_fluct(self):
        lib().orc_initfields_fluct(self._st)

    def meansvars(self):
        out = np.zeros(2 * self.nflds)
        lib().orc_meansvars(self._st, _dp(out))
        return out

    def calcrho(self, s, i, j, k):
        return lib().orc_calcrho(self._st, s, i, j, k)


def srand(seed):
    lib().orc_srand(seed)


def rand():
    return lib().orc_rand()
