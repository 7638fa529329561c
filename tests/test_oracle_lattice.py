"""Pins of the oracle's lattice operators, energies, expansion and step
(g2functions.cpp, g2model.cpp) by closed forms and independent formulations.

Each check is chosen so that a plausible slip (a dropped term, a wrong sign,
a swapped slot or axis) changes the result: plane-wave eigenvalues of the
7-point Laplacian, Parseval's relation for the gradient energy, the
homogeneous-field reduction of the step to the scalar ODE with its
second-order convergence, and the per-mode amplification matrix of midpoint
RK2 for a free field.
"""
import math

import numpy as np
import pytest
from scipy.integrate import solve_ivp

from gabe_amd import configs
from oracle import pyoracle as po


def plane_wave(N, n, amp=1.0, phase=0.3):
    x = np.arange(N)
    th = 2 * np.pi * (n[0] * x[:, None, None] + n[1] * x[None, :, None] + n[2] * x[None, None, :]) / N
    return amp * np.cos(th + phase), th


def mu(N, dx, n):
    return 4.0 / dx ** 2 * sum(math.sin(math.pi * a / N) ** 2 for a in n)


@pytest.mark.parametrize("n", [(1, 0, 0), (0, 2, 1), (3, 1, 2), (4, 4, 4), (1, 5, 7)])
def test_laplacian_plane_wave_eigenvalue(n):
    N = 8
    o = po.Oracle(N, nflds=1, L=4.0)
    f, _ = plane_wave(N, n)
    o.field[0, 0] = f
    lap = np.array([[[o.laplacian(0, 0, i, j, k) for k in range(N)] for j in range(N)] for i in range(N)])
    assert np.abs(lap + mu(N, o.dx, n) * f).max() < 1e-12 * mu(N, o.dx, n)


@pytest.mark.parametrize("n", [(1, 0, 0), (2, 3, 1)])
def test_energies_single_mode_closed_form(n):
    N, A, B, a = 16, 0.7, 0.4, 1.3
    o = po.Oracle(N, nflds=1, L=5.0, m2_phi=2.0, lam=0.5)
    f, th = plane_wave(N, n, A)
    o.field[0, 0] = f
    o.dfield[0, 0] = B * np.sin(th)
    o.set_scalar("a", 0, a)
    o.calcEnergy(0)
    sc = o.scalars()
    assert abs(sc["edgrad"][0] - mu(N, o.dx, n) * A * A / (4 * a * a)) < 1e-13
    assert abs(sc["edkin"][0] - B * B / 4) < 1e-14
    # <V> of A cos: m2 A^2/4 + lam A^4 3/32
    assert abs(sc["edpot"][0] - (2.0 * A * A / 4 + 0.5 * 3 * A ** 4 / 32)) < 1e-14
    assert sc["edrho"][0] == pytest.approx(sc["edkin"][0] + sc["edpot"][0] + sc["edgrad"][0], rel=1e-15)
    assert o.adf(0) == pytest.approx(math.sqrt(8 * math.pi / 3 * sc["edrho"][0]) * a, rel=1e-15)


def test_gradient_energy_parseval():
    """-sum f lap f = (1/N^3) sum_k mu_k |F_k|^2 with the lattice eigenvalues."""
    N, a = 12, 0.9
    rng = np.random.default_rng(5)
    o = po.Oracle(N, nflds=2, L=3.0)
    o.field[0] = rng.standard_normal((2, N, N, N))
    o.set_scalar("a", 0, a)
    G = o.scalars()["edgrad"][0]
    o.calcEnergy(0)
    G = o.scalars()["edgrad"][0]
    k = np.fft.fftfreq(N) * N
    m = (4 / o.dx ** 2) * (np.sin(np.pi * k / N) ** 2)
    mu3 = m[:, None, None] + m[None, :, None] + m[None, None, :]
    ref = sum(np.sum(mu3 * np.abs(np.fft.fftn(o.field[0, f])) ** 2) for f in range(2)) / N ** 3
    assert G == pytest.approx(ref / (2 * N ** 3 * a * a), rel=1e-12)


def test_effmass_and_potential_closed_form():
    cfg = configs.lambda_phi4_config(N=8)
    o = po.Oracle(8, nflds=2, L=cfg.L, m2_phi=0.3, lam=cfg.lam, g2=cfg.g2, m2_chi=0.2)
    o.field[0, 0] = 0.5
    o.field[0, 1] = -0.1
    assert o.effMass(0, 0) == pytest.approx(0.3 + cfg.g2 * 0.01 + 3 * cfg.lam * 0.25, rel=1e-13)
    assert o.effMass(0, 1) == pytest.approx(cfg.g2 * 0.25 + 0.2, rel=1e-13)
    o.calcEnergy(0)
    V = 0.5 * 0.3 * 0.25 + 0.25 * cfg.lam * 0.5 ** 4 + 0.5 * cfg.g2 * 0.25 * 0.01 + 0.5 * 0.2 * 0.01
    assert o.scalars()["edpot"][0] == pytest.approx(V, rel=1e-14)


def test_fourth_order_gradient_plane_wave():
    """calcrho's gradient (g2functions.cpp:70-122) on a plane wave."""
    N, n = 16, (1, 2, 3)
    o = po.Oracle(N, nflds=1, L=2.0)
    f, th = plane_wave(N, n, phase=0.0)
    o.field[0, 0] = f
    s = [math.sin(2 * math.pi * a / N) * 4 / 3 - math.sin(4 * math.pi * a / N) / 6 for a in n]
    i, j, k = 3, 5, 7
    d = -np.sin(th[i, j, k]) * np.array(s) / o.dx
    got = po.lib().orc_gradF2(o._st, 0, 0, i, j, k)
    assert got == pytest.approx(float(np.sum(d * d)), rel=1e-12)


def test_meansvars_closed_form():
    N = 8
    o = po.Oracle(N, nflds=2, L=1.0)
    f, _ = plane_wave(N, (1, 0, 2), 0.2)
    o.field[0, 0] = 3.0 + f
    o.field[0, 1] = -1.5
    mv = o.meansvars()
    assert mv[0] == pytest.approx(3.0, rel=1e-14)
    assert mv[1] == pytest.approx(0.02, rel=1e-9)
    assert mv[2] == pytest.approx(-1.5, rel=1e-15) and abs(mv[3]) < 1e-14


def scalar_ode_rk2(cfg, nsteps):
    """Midpoint RK2 of the homogeneous system phi'' = -V'(phi) - 3 H phi',
    a' = a sqrt(8 pi rho / 3), in the stage order of g2functions.cpp:234-270."""
    def dV(p):
        return cfg.m2_phi * p + cfg.lam * p ** 3

    def rho(p, v):
        return 0.5 * v * v + 0.5 * cfg.m2_phi * p * p + 0.25 * cfg.lam * p ** 4

    p, v, a = cfg.f0[0], cfg.df0[0], 1.0
    ad = a * math.sqrt(8 * math.pi / 3 * rho(p, v))
    dt = cfg.dt
    for _ in range(nsteps):
        p1 = p + 0.5 * dt * v
        v1 = v + 0.5 * dt * (-dV(p) - 3 * ad / a * v)
        a1 = a + 0.5 * dt * ad
        ad1 = a1 * math.sqrt(8 * math.pi / 3 * rho(p1, v1))
        p, v = p + dt * v1, v + dt * (-dV(p1) - 3 * ad1 / a1 * v1)
        a = a + dt * ad1
        ad = a * math.sqrt(8 * math.pi / 3 * rho(p, v))
    return p, v, a, ad


def homogeneous_oracle(cfg):
    o = po.Oracle(cfg.N, nflds=1, L=cfg.L, dt=cfg.dt, m2_phi=cfg.m2_phi, lam=cfg.lam)
    o.initfields_homog(cfg.f0[:1], cfg.df0[:1])
    o.initexpansion()
    return o


def test_homogeneous_step_is_the_scalar_ode():
    cfg = configs.lambda_phi4_config(N=4)
    cfg.m2_phi = 0.5
    o = homogeneous_oracle(cfg)
    o.step(25)
    p, v, a, ad = scalar_ode_rk2(cfg, 25)
    assert np.abs(o.field[0, 0] - p).max() < 1e-13
    assert np.abs(o.dfield[0, 0] - v).max() < 1e-13
    assert o.scalar("a", 0) == pytest.approx(a, rel=1e-14)
    assert o.scalar("adot", 0) == pytest.approx(ad, rel=1e-13)


def test_homogeneous_step_converges_at_second_order():
    """Against an accurate ODE solution: the error shrinks 4x per halving of dt."""
    cfg = configs.lambda_phi4_config(N=4)
    T = 0.5

    def rhs(t, y):
        p, v, a = y
        rho = 0.5 * v * v + 0.25 * cfg.lam * p ** 4
        H = math.sqrt(8 * math.pi / 3 * rho)
        return [v, -cfg.lam * p ** 3 - 3 * H * v, a * H]

    ref = solve_ivp(rhs, (0, T), [cfg.f0[0], cfg.df0[0], 1.0], rtol=1e-13, atol=1e-15).y[:, -1]
    errs = []
    for nsteps in (40, 80, 160):
        cfg.dt = T / nsteps
        o = homogeneous_oracle(cfg)
        o.step(nsteps)
        errs.append(abs(o.field[0, 0, 0, 0, 0] - ref[0]) + abs(o.scalar("a", 0) - ref[2]))
    assert 3.5 < errs[0] / errs[1] < 4.5 and 3.5 < errs[1] / errs[2] < 4.5


def test_free_field_modes_follow_the_rk2_amplification_matrix():
    """No expansion, V = m^2 f^2/2: each lattice mode is an oscillator with
    w^2 = mu_k + m^2, and one midpoint-RK2 step maps (F, F') by
    [[1 - w^2 dt^2/2, dt], [-w^2 dt, 1 - w^2 dt^2/2]]."""
    N, m2, dt = 8, 0.7, 0.05
    o = po.Oracle(N, nflds=1, L=4.0, dt=dt, expansion_type=0, m2_phi=m2)
    rng = np.random.default_rng(3)
    o.field[0, 0] = rng.standard_normal((N, N, N))
    o.dfield[0, 0] = rng.standard_normal((N, N, N))
    F0, V0 = np.fft.fftn(o.field[0, 0]), np.fft.fftn(o.dfield[0, 0])
    o.set_scalar("a", 0, 1.0)
    o.set_scalar("a", 1, 1.0)
    o.step(3)
    k = np.fft.fftfreq(N) * N
    m = (4 / o.dx ** 2) * np.sin(np.pi * k / N) ** 2
    w2 = m[:, None, None] + m[None, :, None] + m[None, None, :] + m2
    A = 1 - w2 * dt * dt / 2
    F, V = F0, V0
    for _ in range(3):
        F, V = A * F + dt * V, -w2 * dt * F + A * V
    assert np.abs(np.fft.fftn(o.field[0, 0]) - F).max() < 1e-11 * np.abs(F).max()
    assert np.abs(np.fft.fftn(o.dfield[0, 0]) - V).max() < 1e-11 * np.abs(V).max()


def test_step_time_and_slots():
    cfg = configs.lambda_phi4_config(N=4)
    o = homogeneous_oracle(cfg)
    o.step(3)
    assert o.st.t == pytest.approx(3 * cfg.dt)
    # slot 1 holds the last midpoint: f1 = f0_prev + dt/2 fdot0_prev
    p, v, a, ad = scalar_ode_rk2(cfg, 2)
    assert o.field[1, 0, 0, 0, 0] == pytest.approx(p + 0.5 * cfg.dt * v, rel=1e-14)
