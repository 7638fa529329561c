"""Pins of the oracle's tensor-perturbation path (g2evolutionpert.cpp), spectra
(g2spectra.cpp) and vacuum initial conditions (g2init.cpp).

Stress: closed form on a plane wave.  TT source: transversality,
tracelessness, symmetry and idempotence, and equality with the projector
written as P S P - P tr(P S)/2 on numpy's transform of the oracle's stress.
Mode update: the midpoint-RK2 amplification matrix of a free wave and the
exact fixed point of a constant source.  Spectra: a single mode lands in a
single bin with the right normalisation, Parseval's relation, and the bin
weights reproduced from a full complex transform.  ICs: glibc's rand()
restated from its published additive-feedback recurrence, and the mode
variance B^2/(2 w L^3) of the vacuum formula.
"""
import math

import numpy as np
import pytest

from gabe_amd import configs
from oracle import pyoracle as po

PAIRS = [(1, 1), (1, 2), (1, 3), (2, 2), (2, 3), (3, 3)]


def test_stress_plane_wave_closed_form():
    N, n = 8, (1, 2, 3)
    o = po.Oracle(N, nflds=1, L=2.0, calc_gws=1)
    x = np.arange(N)
    th = 2 * np.pi * (n[0] * x[:, None, None] + n[1] * x[None, :, None] + n[2] * x[None, None, :]) / N
    o.field[0, 0] = np.cos(th)
    s = np.array([math.sin(2 * math.pi * a / N) for a in n])    # central difference of cos: -sin(th) s_a
    i, j, k = 1, 4, 6
    g = -np.sin(th[i, j, k]) * s
    for aa, bb in PAIRS:
        ref = g[aa - 1] * g[bb - 1] / o.dx ** 2 - (aa == bb) * np.dot(g, g) / 3 / o.dx ** 2
        assert o.stresstensor(aa, bb, 0, i, j, k) == pytest.approx(ref, abs=1e-12 / o.dx ** 2)


def gw_oracle(N=8, seed=1):
    o = po.Oracle(N, nflds=2, L=3.0, calc_gws=1)
    rng = np.random.default_rng(seed)
    o.field[0] = rng.standard_normal((2, N, N, N))
    return o


def stress_arrays(o, s=0):
    N = o.N
    S = np.zeros((6, N, N, N))
    for c, (aa, bb) in enumerate(PAIRS):
        for i in range(N):
            for j in range(N):
                for k in range(N):
                    S[c, i, j, k] = o.stresstensor(aa, bb, s, i, j, k)
    return S


def T_complex(o, s=0):
    T = o.T_gw[s]
    return T[0::2] + 1j * T[1::2]   # (11,12,13,22,23,33)


def khat(N):
    px = np.array([i if i <= N // 2 else i - N for i in range(N)], dtype=float)
    pz = np.arange(N // 2 + 1, dtype=float)
    P = np.stack(np.meshgrid(px, px, pz, indexing="ij"))
    nrm = np.sqrt((P ** 2).sum(0))
    nrm[0, 0, 0] = 1.0
    return P / nrm


def as_matrix(c6):
    m = np.empty((3, 3) + c6.shape[1:], dtype=c6.dtype)
    idx = {(0, 0): 0, (0, 1): 1, (0, 2): 2, (1, 1): 3, (1, 2): 4, (2, 2): 5}
    for (a, b), c in idx.items():
        m[a, b] = c6[c]
        m[b, a] = c6[c]
    return m


def test_tt_source_invariants_and_projector_form():
    o = gw_oracle()
    N = o.N
    o.fft_stresstensor(0)
    T = as_matrix(T_complex(o))
    n = khat(N)
    scale = np.abs(T).max()
    assert np.abs(np.einsum("i...,ij...->j...", n, T)).max() < 1e-12 * scale      # k_i T_ij = 0
    assert np.abs(T[0, 0] + T[1, 1] + T[2, 2]).max() < 1e-12 * scale              # traceless
    assert np.all(T[:, :, 0, 0, 0] == 0)                                           # zero mode
    # independent: S_hat by numpy, projector P S P - P tr(PS)/2, times dx^3
    Sh = as_matrix(np.fft.rfftn(stress_arrays(o), axes=(1, 2, 3)))
    P = np.eye(3)[:, :, None, None, None] - n[:, None] * n[None, :]
    PSP = np.einsum("ij...,jk...,kl...->il...", P, Sh, P)
    trPS = np.einsum("ij...,ji...->...", P, Sh)
    ref = (PSP - 0.5 * P * trPS) * o.dx ** 3
    ref[:, :, 0, 0, 0] = 0
    assert np.abs(T - ref).max() < 1e-12 * scale
    # idempotent: projecting the projected source changes nothing
    PTP = np.einsum("ij...,jk...,kl...->il...", P, T, P)
    trPT = np.einsum("ij...,ji...->...", P, T)
    assert np.abs(PTP - 0.5 * P * trPT - T).max() < 1e-12 * scale


def test_gw_free_wave_amplification():
    """T = 0, fixed a, adot = 0: (h, l) of a mode with k^2 = (2 pi/L)^2 |n|^2/a^2
    is mapped by the midpoint-RK2 matrix [[1 - k^2 D^2/2, D], [-k^2 D, 1 - k^2 D^2/2]]."""
    o = gw_oracle()
    N, L, a, D = o.N, 3.0, 1.7, 0.01
    rng = np.random.default_rng(2)
    o.h[0] = rng.standard_normal(o.h[0].shape)
    o.l[0] = rng.standard_normal(o.l[0].shape)
    h0, l0 = o.h[0].copy(), o.l[0].copy()
    o.T_gw[:] = 0
    for s in (0, 1):
        o.set_scalar("a", s, a)
        o.set_scalar("adot", s, 0.0)
    o.gw_update(0, 1, D / 2)
    o.gw_update(1, 0, D)
    px = np.array([i if i <= N // 2 else i - N for i in range(N)], dtype=float)
    k2 = (2 * np.pi / L / a) ** 2 * (px[:, None, None] ** 2 + px[None, :, None] ** 2
                                     + np.arange(N // 2 + 1)[None, None, :] ** 2)
    A = 1 - k2 * D * D / 2
    assert np.abs(o.h[0] - (A * h0 + D * l0)).max() < 1e-13 * np.abs(h0).max() * 10
    assert np.abs(o.l[0] - (-k2 * D * h0 + A * l0)).max() < 1e-12 * np.abs(k2 * h0).max()


def test_gw_constant_source_fixed_point():
    """h* = 16 pi T / k^2 (with adot = 0, l = 0) is stationary under both stages."""
    o = gw_oracle()
    N, L, a = o.N, 3.0, 1.0
    rng = np.random.default_rng(4)
    T = rng.standard_normal(o.T_gw[0].shape)
    T[:, 0, 0, 0] = 0.0      # the source has no k = 0 mode (g2evolutionpert.cpp:319-321)
    o.T_gw[0] = T
    o.T_gw[1] = T
    for s in (0, 1):
        o.set_scalar("a", s, a)
        o.set_scalar("adot", s, 0.0)
    px = np.array([i if i <= N // 2 else i - N for i in range(N)], dtype=float)
    k2 = (2 * np.pi / L) ** 2 * (px[:, None, None] ** 2 + px[None, :, None] ** 2
                                 + np.arange(N // 2 + 1)[None, None, :] ** 2)
    k2[0, 0, 0] = np.inf
    hstar = 16 * np.pi * T / k2
    o.h[0] = hstar
    o.l[0] = 0
    o.gw_update(0, 1, 0.005)
    o.gw_update(1, 0, 0.01)
    assert np.abs(o.h[0] - hstar).max() < 1e-14 * np.abs(hstar).max()
    assert np.abs(o.l[0]).max() < 1e-12 * np.abs(hstar).max()


def spectrum_oracle(N, L=5.0, gws=1):
    return po.Oracle(N, nflds=2, L=L, calc_gws=gws)


def test_single_mode_spectrum():
    N, L = 16, 5.0
    o = spectrum_oracle(N, L, gws=0)
    n = (2, 3, 1)
    x = np.arange(N)
    th = 2 * np.pi * (n[0] * x[:, None, None] + n[1] * x[None, :, None] + n[2] * x[None, None, :]) / N
    A = 0.3
    o.field[0, 0] = A * np.cos(th)
    sp = o.specout()
    b = int(math.floor(math.sqrt(sum(v * v for v in n))))
    P = sp[0]
    assert np.count_nonzero(P > 1e-20 * P.max()) == 1 and P[b] > 0
    # |F|^2 = (A N^3 / 2)^2 at +k and -k, averaged over the bin's weight
    count = full_bin_counts(N)[b]
    spec_norm = (L / N) ** 6 * (2 * math.pi / L) ** 3
    assert P[b] == pytest.approx(2 * (A * N ** 3 / 2) ** 2 * spec_norm / count, rel=1e-12)
    assert sp[1][b] == pytest.approx((2 * math.pi * b / L) ** 3 * P[b] / (2 * math.pi ** 2), rel=1e-14)


def full_bin_counts(N):
    """modes of the full N^3 transform per bin floor(|n|) with n in (-N/2, N/2]"""
    p = np.array([i if i <= N // 2 else i - N for i in range(N)], dtype=float)
    mag = np.sqrt(p[:, None, None] ** 2 + p[None, :, None] ** 2 + p[None, None, :] ** 2)
    return np.bincount(np.floor(mag).astype(int).ravel(), minlength=po.numbins(N))


@pytest.mark.parametrize("N", [8, 12, 16])
def test_field_spectrum_equals_full_transform_binning(N):
    """Half-spectrum with weights 2 (interior kz) and 1 (kz = 0, N/2) equals the
    plain average of |F|^2 over all N^3 modes of each shell."""
    L = 4.0
    o = spectrum_oracle(N, L, gws=0)
    rng = np.random.default_rng(N)
    o.field[0] = rng.standard_normal((2, N, N, N))
    sp = o.specout()
    p = np.array([i if i <= N // 2 else i - N for i in range(N)], dtype=float)
    mag = np.sqrt(p[:, None, None] ** 2 + p[None, :, None] ** 2 + p[None, None, :] ** 2)
    bins = np.floor(mag).astype(int).ravel()
    nb = po.numbins(N)
    spec_norm = (L / N) ** 6 * (2 * math.pi / L) ** 3
    for f in range(2):
        F2 = np.abs(np.fft.fftn(o.field[0, f])) ** 2
        s = np.bincount(bins, weights=F2.ravel(), minlength=nb)
        c = np.bincount(bins, minlength=nb)
        ref = np.where(c > 0, s * spec_norm / np.maximum(c, 1), 0)
        assert np.abs(sp[2 * f] - ref).max() < 1e-12 * ref.max()
        # Parseval: sum over bins of P * count = spec_norm * N^3 * sum x^2
        assert np.sum(sp[2 * f] * c) == pytest.approx(spec_norm * N ** 3 * np.sum(o.field[0, f] ** 2), rel=1e-12)


@pytest.mark.parametrize("quirk", [0, 1])
def test_gw_spectrum_binning(quirk):
    N, L = 8, 4.0
    o = spectrum_oracle(N, L)
    rng = np.random.default_rng(9)
    o.field[0] = rng.standard_normal((2, N, N, N))
    o.l[0] = rng.standard_normal(o.l[0].shape)
    o.set_scalar("a", 0, 1.0)
    o.calcEnergy(0)
    sp = o.specout(gw_quirk=quirk)
    lc = o.l[0][0::2] + 1j * o.l[0][1::2]
    w = np.array([1, 2, 2, 1, 2, 1])[:, None, None, None]
    E = np.sum(w * np.abs(lc) ** 2, axis=0)
    if quirk:   # g2spectra.cpp:149-164: the kz = 0 term reads the kz = N/2 plane
        E[:, :, 0] = E[:, :, N // 2]
    p = np.array([i if i <= N // 2 else i - N for i in range(N)], dtype=float)
    kz = np.arange(N // 2 + 1)
    mag = np.sqrt(p[:, None, None] ** 2 + p[None, :, None] ** 2 + kz[None, None, :] ** 2)
    wt = np.where((kz == 0) | (kz == N // 2), 1.0, 2.0)[None, None, :] * np.ones_like(mag)
    bins = np.floor(mag).astype(int).ravel()
    nb = po.numbins(N)
    s = np.bincount(bins, weights=(wt * E).ravel(), minlength=nb)
    c = np.bincount(bins, weights=wt.ravel(), minlength=nb)
    spec_norm = (L / N) ** 6 * (2 * math.pi / L) ** 3
    ref = np.where(c > 0, s * spec_norm / np.maximum(c, 1), 0)
    assert np.abs(sp[4] - ref).max() < 1e-12 * ref.max()
    H = math.sqrt(8 * math.pi / 3) * math.sqrt(o.scalar("edrho", 0))
    norm1 = 4e-5 / 100 ** 0.333 / 24 / math.pi ** 2 / L ** 3 / H ** 2
    kb = 2 * math.pi * np.arange(nb) / L
    assert sp[5].max() > 0
    assert np.abs(sp[5] - norm1 * kb ** 3 * np.where(c > 0, s / np.maximum(c, 1), 0)).max() < 1e-12 * sp[5].max()
    assert np.allclose(sp[6], kb)


def glibc_rand(seed, count):
    """glibc rand(): TYPE_3 additive feedback r[i] = r[i-3] + r[i-31] (mod 2^32),
    seeded by r[i] = 16807 r[i-1] mod (2^31 - 1), first 310 outputs discarded."""
    r = [0] * (count + 344)
    r[0] = seed if seed != 0 else 1
    for i in range(1, 31):
        r[i] = (16807 * r[i - 1]) % 2147483647
    for i in range(31, 34):
        r[i] = r[i - 31]
    for i in range(34, count + 344):
        r[i] = (r[i - 31] + r[i - 3]) & 0xFFFFFFFF
    return [x >> 1 for x in r[344:344 + count]]


def test_libc_rand_sequence_is_glibc_type3():
    po.srand(44463132)
    got = [po.rand() for _ in range(200)]
    assert got == glibc_rand(44463132, 200)


def test_vacuum_mode_variance():
    """E|f_k|^2 = B^2 / (2 w_k L^3) for the draws of g2init.cpp:117-151."""
    cfg = configs.lambda_phi4_config(N=32)
    o = po.Oracle(32, nflds=2, L=cfg.L, m2_phi=0.0, lam=cfg.lam, g2=cfg.g2, rescale_B=1.0)
    o.initfields_homog(cfg.f0, cfg.df0)
    o.initexpansion()
    po.srand(12345)
    m2 = o.effMass(0, 0)
    fk, fkd = o.randinit(0, m2, want_modes=True)
    N, L = 32, cfg.L
    i2 = np.array([min(i, N - i) ** 2 for i in range(N)], dtype=float)
    w = np.sqrt(4 * math.pi ** 2 / L ** 2 * (i2[:, None, None] + i2[None, :, None]
                                              + np.arange(N // 2 + 1)[None, None, :] ** 2) + m2)
    ratio = (np.abs(fk) ** 2 * 2 * w * L ** 3).ravel()[1:]
    assert abs(ratio.mean() - 1.0) < 4 * ratio.std() / math.sqrt(ratio.size)
    H0 = math.sqrt(8 * math.pi / 3 * o.scalar("edrho", 0))
    # fkd + H0 fk = i w (RI1 - RI2)/sqrt2: same variance as fk, uncorrelated with it
    g = (fkd + H0 * fk) / (1j * w)
    r2 = (np.abs(g) ** 2 * 2 * w * L ** 3).ravel()[1:]
    assert abs(r2.mean() - 1.0) < 4 * r2.std() / math.sqrt(r2.size)
    corr = np.mean((fk * np.conj(g)).ravel()[1:] * 2 * w.ravel()[1:] * L ** 3)
    assert abs(corr) < 4 / math.sqrt(ratio.size)
    assert fk[0, 0, 0] == 0 and fkd[0, 0, 0] == 0
