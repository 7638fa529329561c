"""Writes tests/golden/chaos_32.npz: the CPU FP64 oracle through the chaotic
phase of a 32^3 lattice with two fields and tensor modes (the lambda phi^4 +
g^2 phi^2 chi^2 model, g^2/lambda = 100, L = 20, dt = dx/20), 10000 midpoint
RK2 steps from the reference's seeded fluctuations (randseed 44463132,
g2init.cpp:68-189) drawn at the amplitude of a rescaling B = 0.3 instead of
the physical one, so that parametric resonance fragments the condensate
within reach of the CPU.

Beside the oracle's own run, an ensemble of ENS runs starts from the same
state with both fields multiplied, site by site, by 1 + 1e-7 x (independent
unit normal noise): differences of the size of FP32 rounding.  Once the run
is chaotic these runs share nothing pointwise with the oracle's, and their
scatter is the yardstick a statistical comparison is held to.

Stored (all FP64):
  spectra     [201][7][numbins]  the specOut columns (g2spectra.cpp:23-197,
                                 GW column with the reference's kz = 0
                                 reading) at steps 0, 50, ..., 10000
  history     [W][8]             t, a, adot, K, V, G, rho, step at the start of
                                 every step from WIN0 - 50 to 10000 (the rows
                                 the Friedmann residual over the chaotic
                                 window is computed from)
  planes      [5][2][32][32]     plane i = 0 of phi and chi after 100, 200,
                                 300, 400 and 500 steps (the ensemble runs
                                 reach the fluctuation's own size by step 700)
  ens_dev     [ENS][200]         each ensemble run's max |f - f_oracle| over
                                 both fields, relative to the oracle's
                                 fluctuation amplitude, every 50 steps
  ens_spec    [ENS][7][numbins]  each ensemble run's spectra averaged over the
                                 chaotic window (steps WIN0, WIN0 + 50, ...,
                                 10000)
  ens_scal    [ENS][5]           each ensemble run's window means of K, V, G,
                                 a and rho (per-step rows)
  ens_resid   [ENS][2]           each ensemble run's largest and mean
                                 Friedmann residual over the window

Run from the repository root after building (the oracle and the host
library, whose residual routine is used):
    python tests/golden/make_chaos_32.py
"""
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)

from gabe_amd import configs, host  # noqa: E402
from tests.helpers import seeded_oracle  # noqa: E402

N, STEPS, EVERY, AMP = 32, 10000, 50, 0.3
WIN0 = 7500                 # first step of the chaotic window
PLANE_STEPS = (100, 200, 300, 400, 500)
ENS, NOISE = 8, 1e-7
OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "chaos_32.npz")


def history_row(o, n):
    st = o.st
    return [st.t, st.a[0], st.adot[0], st.edkin[0], st.edpot[0], st.edgrad[0], st.edrho[0], n]


def run(member, threads, ref_fields=None):
    """One oracle run; member < 0 is the unperturbed run.  Returns its spectra
    every EVERY steps, its history rows from WIN0 - EVERY, its planes, its
    fields every EVERY steps (unperturbed run) or its deviation from
    ref_fields (ensemble runs)."""
    cfg = configs.lambda_phi4_config(N=N, calc_gws=1)
    o = seeded_oracle(cfg, amp=AMP, nthreads=threads)
    if member >= 0:
        rng = np.random.default_rng(1000 + member)
        f = o.field[0]
        f *= 1.0 + NOISE * rng.standard_normal(f.shape)
    spectra, hist, planes, fields, dev = [o.specout(gw_quirk=1)], [], [], [], []
    for n in range(STEPS):
        if n >= WIN0 - EVERY:
            hist.append(history_row(o, n))
        o.step(1)
        m = n + 1
        if m % EVERY == 0:
            spectra.append(o.specout(gw_quirk=1))
            f = o.field[0]
            if ref_fields is None:
                fields.append(f.copy())
            else:
                r = ref_fields[m // EVERY - 1]
                dev.append(max(np.abs(f[q] - r[q]).max() / np.abs(r[q] - r[q].mean()).max()
                               for q in range(2)))
            if m % 1000 == 0:
                print(f"run {member}: step {m}", flush=True)
        if m in PLANE_STEPS:
            planes.append(o.field[0][:, 0].copy())
    hist.append(history_row(o, STEPS))
    return np.array(spectra), np.array(hist), np.array(planes), fields, np.array(dev)


def window_stats(spectra, hist):
    spec = spectra[WIN0 // EVERY:].mean(axis=0)
    rows = hist[EVERY:]                      # steps WIN0 .. STEPS
    scal = [rows[:, 3].mean(), rows[:, 4].mean(), rows[:, 5].mean(), rows[:, 1].mean(), rows[:, 6].mean()]
    _, R = host.friedmann_residual(hist)
    w = R[EVERY:-1]
    return spec, np.array(scal), [float(w.max()), float(w.mean())]


def main():
    threads = min(16, os.cpu_count() or 1)
    spectra, hist, planes, fields, _ = run(-1, threads)
    ens_dev, ens_spec, ens_scal, ens_resid = [], [], [], []
    for m in range(ENS):
        sp, hi, _, _, dev = run(m, threads, ref_fields=fields)
        s, c, r = window_stats(sp, hi)
        ens_dev.append(dev)
        ens_spec.append(s)
        ens_scal.append(c)
        ens_resid.append(r)
    cfg = configs.lambda_phi4_config(N=N, calc_gws=1)
    np.savez_compressed(OUT, spectra=spectra, history=hist, planes=planes, ens_dev=np.array(ens_dev),
                        ens_spec=np.array(ens_spec), ens_scal=np.array(ens_scal),
                        ens_resid=np.array(ens_resid), N=N, steps=STEPS, every=EVERY, win0=WIN0,
                        amp=AMP, noise=NOISE, plane_steps=np.array(PLANE_STEPS), L=cfg.L, dt=cfg.dt,
                        randseed=cfg.randseed)
    print(f"wrote {OUT}")


if __name__ == "__main__":
    main()
