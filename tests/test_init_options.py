"""The initial-condition options of g2init.cpp:111-165 (field_full_rand,
spec_cut_off, spec_smooth; g2parameters.h:72-74), CPU only.

Oracle pins: the tanh window multiplies each drawn mode by a factor that
depends on |n| alone, is exactly 1/2 on the shell |n| = N cut_off, falls
with |n|, and leaves the rand() stream where the plain draws leave it; the
symmetric modes draw nothing, are real, give an even field, carry
fkd = -H0 fk and have the power that the random modes have on average.
Host library: gabe_host_modes_opt (the kept host's mode builder) reproduces
the oracle's modes from the same draws for every option combination.
"""
import ctypes as C
import math

import numpy as np
import pytest

from gabe_amd import configs
from oracle import pyoracle as po

CUT = math.sqrt(3.0) / 8.0


def fresh_oracle(N=16):
    cfg = configs.lambda_phi4_config(N=N)
    o = po.Oracle(N, nflds=2, L=cfg.L, m2_phi=0.0, lam=cfg.lam, g2=cfg.g2, rescale_B=1.0)
    o.initfields_homog(cfg.f0, cfg.df0)
    o.initexpansion()
    return o, cfg


def nabs(N):
    i = np.array([min(a, N - a) for a in range(N)], dtype=float)
    return np.sqrt(i[:, None, None] ** 2 + i[None, :, None] ** 2 + np.arange(N // 2 + 1)[None, None, :] ** 2)


def modes(o, seed, m2, **kw):
    """randinit adds its fluctuation to the field, so the effective mass is
    taken once by the caller and passed to every call."""
    po.srand(seed)
    fk, fkd = o.randinit(0, m2, want_modes=True, **kw)
    return fk, fkd, po.rand()


def test_window_scales_each_drawn_mode():
    N = 16
    o, _ = fresh_oracle(N)
    m2 = o.effMass(0, 0)
    fk, fkd, after = modes(o, 77, m2)
    fkc, fkdc, after_c = modes(o, 77, m2, cut_off=CUT, smooth=0.5)
    assert after == after_c                      # the same number of draws
    n = nabs(N)
    live = n > 0
    ratio = (fkc[live] / fk[live])
    assert np.abs(ratio.imag).max() < 1e-12
    r = ratio.real
    # one factor per shell |n|
    for v in np.unique(np.round(n[live] ** 2)):
        sel = np.isclose(n[live] ** 2, v)
        assert np.ptp(r[sel]) < 1e-12
    # exactly 1/2 on |n| = N cut_off = sqrt(12): n = (2, 2, 2)
    assert abs(fkc[2, 2, 2] / fk[2, 2, 2] - 0.5) < 1e-12
    order = np.argsort(n[live])
    assert np.all(np.diff(r[order]) <= 1e-12)    # falls with |n|
    assert r.max() < 1 and r[n[live] > 3 * N * CUT].max() < 1e-3
    H0 = math.sqrt(8 * math.pi / 3 * o.scalar("edrho", 0))
    g, gc = fkd + H0 * fk, fkdc + H0 * fkc       # the velocity part carries the same factor
    assert np.allclose(gc[live], r * g[live], rtol=1e-12, atol=1e-30)


def test_symmetric_modes():
    N = 16
    o, _ = fresh_oracle(N)
    m2 = o.effMass(0, 0)
    po.srand(5)
    first = po.rand()
    fks, fkds, after = modes(o, 5, m2, full_rand=0, cut_off=CUT, smooth=0.5)
    assert after == first                        # no draws
    assert np.all(fks.imag == 0) and np.all(fkds.imag == 0)
    H0 = math.sqrt(8 * math.pi / 3 * o.scalar("edrho", 0))
    assert np.allclose(fkds, -H0 * fks, rtol=1e-15, atol=0)
    # the deterministic amplitude is the mean power of the drawn modes
    n = nabs(N)
    live = n > 0
    pw = []
    for seed in range(40):
        fk, _, _ = modes(o, 1000 + seed, m2, cut_off=CUT, smooth=0.5)
        pw.append(np.abs(fk[live]) ** 2 / fks[live].real ** 2)
    pw = np.concatenate(pw)
    assert abs(pw.mean() - 1.0) < 4 * pw.std() / math.sqrt(pw.size)
    # real symmetric modes: the transformed field is even, f(x) = f(-x)
    o2, cfg = fresh_oracle(N)
    before = o2.field[0, 0].copy()
    o2.randinit(0, o2.effMass(0, 0), full_rand=0, cut_off=CUT, smooth=0.5)
    df = o2.field[0, 0] - before
    idx = (-np.arange(N)) % N
    assert np.abs(df - df[np.ix_(idx, idx, idx)]).max() <= 1e-12 * np.abs(df).max()
    assert np.abs(df).max() > 0


def host_lib():
    from gabe_amd import host
    L = host.load()

    class Modes(C.Structure):
        _fields_ = [("full_rand", C.c_int), ("cut_off", C.c_double), ("smooth", C.c_double)]

    dp = C.POINTER(C.c_double)
    L.gabe_host_modes_opt.argtypes = [C.c_int, C.c_double, C.c_double, C.c_double, C.c_double,
                                      C.POINTER(Modes), C.POINTER(C.c_int), dp, dp]
    L.gabe_host_modes_opt.restype = C.c_int
    L.gabe_host_draws_per_field.argtypes = [C.c_int]
    L.gabe_host_draws_per_field.restype = C.c_long
    return L, Modes


@pytest.mark.parametrize("full_rand,cut", [(1, 0.0), (1, CUT), (0, CUT)])
def test_host_mode_builder_matches_oracle(full_rand, cut):
    N = 8
    L, Modes = host_lib()
    o, cfg = fresh_oracle(N)
    m2 = o.effMass(0, 0)
    H0 = math.sqrt(8 * math.pi / 3 * o.scalar("edrho", 0))
    fk_o, fkd_o, _ = modes(o, 31, m2, full_rand=full_rand, cut_off=cut, smooth=0.5)
    ndraw = L.gabe_host_draws_per_field(N) if full_rand else 0
    po.srand(31)
    draws = (C.c_int * max(ndraw, 1))(*[po.rand() for _ in range(ndraw)])
    fk = np.zeros((N, N, N // 2 + 1), dtype=np.complex128)
    fkd = np.zeros_like(fk)
    opt = Modes(full_rand, cut, 0.5)
    rc = L.gabe_host_modes_opt(N, cfg.L, 1.0, m2, H0, C.byref(opt), draws,
                               fk.view(np.float64).ctypes.data_as(C.POINTER(C.c_double)),
                               fkd.view(np.float64).ctypes.data_as(C.POINTER(C.c_double)))
    assert rc == 0
    assert np.array_equal(fk, fk_o) and np.array_equal(fkd, fkd_o)


def test_host_rejects_symmetric_modes_without_window():
    L, Modes = host_lib()
    N = 8
    fk = np.zeros(2 * N * N * (N // 2 + 1))
    opt = Modes(0, 0.0, 0.5)
    p = fk.ctypes.data_as(C.POINTER(C.c_double))
    assert L.gabe_host_modes_opt(N, 20.0, 1.0, 1.0, 0.1, C.byref(opt), None, p, p) < 0
