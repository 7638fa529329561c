"""Friedmann-constraint residual R(t) = |rho_dot + 3H(rho + p)| / (3 H rho).

The reference computes no residual (adot is defined from rho,
g2functions.cpp:188-191); this diagnostic scores both codes identically from
per-step FP64 scalars.  The residual function is pinned on closed-form FRW
histories; on the GPU the device run must be no worse than the oracle's.
"""
import math

import numpy as np
import pytest

from gabe_amd import configs, host
from tests.helpers import GABE_FP32, GABE_FP64, lattice_from_oracle, seeded_oracle


def frw_history(w_p, n=400, dt=1e-3, rho0=0.05):
    """Exact flat FRW with p = w rho: a = (1 + 3(1+w)H0 t/2)^(2/(3(1+w)))."""
    H0 = math.sqrt(8 * math.pi / 3 * rho0)
    rows = []
    for i in range(n):
        t = i * dt
        a = (1 + 1.5 * (1 + w_p) * H0 * t) ** (2 / (3 * (1 + w_p)))
        rho = rho0 * a ** (-3 * (1 + w_p))
        adot = a * math.sqrt(8 * math.pi / 3 * rho)
        # split rho into K, V, G with p = K - G/3 - V = w rho and K + V + G = rho
        G = 0.0
        K = 0.5 * (1 + w_p) * rho
        V = rho - K
        rows.append([t, a, adot, K, V, G, rho, i])
    return np.array(rows)


@pytest.mark.parametrize("w", [0.0, 1.0 / 3.0, 1.0])
def test_residual_vanishes_on_exact_frw(w):
    worst, R = host.friedmann_residual(frw_history(w))
    assert worst < 5e-5
    assert R[0] == 0 and R[-1] == 0
    # what is left is the O(dt^2) error of the central difference
    w2, _ = host.friedmann_residual(frw_history(w, n=800, dt=5e-4))
    assert 3.5 < worst / w2 < 4.5


def test_residual_detects_a_wrong_equation_of_state():
    h = frw_history(0.0)
    h[:, 4] += 0.1 * h[:, 6]     # extra potential: p no longer matches the expansion
    worst, _ = host.friedmann_residual(h)
    assert worst > 0.05


def oracle_history(o, nsteps):
    rows = []
    for n in range(nsteps):
        st = o.st
        rows.append([st.t, st.a[0], st.adot[0], st.edkin[0], st.edpot[0], st.edgrad[0], st.edrho[0], n])
        o.step(1)
    return np.array(rows)


@pytest.mark.gpu
@pytest.mark.parametrize("precision", [GABE_FP64, GABE_FP32])
def test_device_residual_no_worse_than_oracle(precision):
    cfg = configs.lambda_phi4_config(N=32, calc_gws=1)
    o = seeded_oracle(cfg, amp=2e-3)
    lat = lattice_from_oracle(o, cfg, precision)
    nsteps = 40
    ho = oracle_history(o, nsteps)
    lat.step(nsteps)
    hd = lat.history()
    assert hd.shape == (nsteps, 8)
    assert np.allclose(hd[:, 1], ho[:, 1], rtol=1e-10)
    wo, Ro = host.friedmann_residual(ho)
    wd, Rd = host.friedmann_residual(hd)
    # interior rows from the second on: the first row carries the homogeneous adot (Q4)
    wo, wd = Ro[2:-1].max(), Rd[2:-1].max()
    assert wd <= wo * (1 + 1e-6) + 1e-12
    if precision == GABE_FP64:
        assert np.abs(Rd[2:-1] - Ro[2:-1]).max() <= 1e-8 * max(wo, 1e-12) + 1e-13
