"""Writes tests/golden/window_64.npz: the CPU FP64 oracle over 2000 midpoint
RK2 steps of a 64^3 lattice with two fields and tensor modes (the lambda
phi^4 + g^2 phi^2 chi^2 model, g^2/lambda = 100, L = 20, dt = dx/20), from the
reference's seeded vacuum fluctuations at their physical amplitude (randseed
44463132, g2init.cpp:68-189, one-thread draw order).

Stored (all FP64):
  spectra   [41][7][numbins]  the specOut columns (g2spectra.cpp:23-197, GW
                              column with the reference's kz = 0 reading) at
                              steps 0, 50, ..., 2000
  history   [2001][8]         t, a, adot, K, V, G, rho, step at the start of
                              every step (the per-step scalars the Friedmann
                              residual is computed from) and after the last
  planes    [4][2][64][64]    plane i = 0 of phi and chi after 500, 1000,
                              1500 and 2000 steps
  twin_diff [40]              max |f - f'| / fluctuation amplitude every 50
                              steps between two runs without tensor modes
                              that start 1e-14 apart in one site (a measure of
                              how fast rounding differences grow)

Run from the repository root after `make -C oracle`:
    python tests/golden/make_window_64.py
"""
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)

from gabe_amd import configs  # noqa: E402
from tests.helpers import seeded_oracle  # noqa: E402

N, STEPS, EVERY = 64, 2000, 50
OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "window_64.npz")


def history_row(o, n):
    st = o.st
    return [st.t, st.a[0], st.adot[0], st.edkin[0], st.edpot[0], st.edgrad[0], st.edrho[0], n]


def main():
    threads = min(16, os.cpu_count() or 1)
    cfg = configs.lambda_phi4_config(N=N, calc_gws=1)
    o = seeded_oracle(cfg, nthreads=threads)
    spectra, hist, planes = [o.specout(gw_quirk=1)], [], []
    for n in range(STEPS):
        hist.append(history_row(o, n))
        o.step(1)
        if (n + 1) % EVERY == 0:
            spectra.append(o.specout(gw_quirk=1))
            print(f"step {n + 1}", flush=True)
        if (n + 1) % 500 == 0:
            planes.append(o.field[0][:, 0].copy())
    hist.append(history_row(o, STEPS))

    twin = []
    a = seeded_oracle(configs.lambda_phi4_config(N=N, calc_gws=0), nthreads=threads)
    b = seeded_oracle(configs.lambda_phi4_config(N=N, calc_gws=0), nthreads=threads)
    b.field[0][0, 3, 5, 7] *= 1 + 1e-14
    for _ in range(STEPS // 50):
        a.step(50)
        b.step(50)
        f, g = a.field[0], b.field[0]
        twin.append(float(np.abs(f - g).max() / np.abs(f[0] - f[0].mean()).max()))
    np.savez_compressed(OUT, spectra=np.array(spectra), history=np.array(hist), planes=np.array(planes),
                        twin_diff=np.array(twin), N=N, steps=STEPS, every=EVERY, L=cfg.L, dt=cfg.dt,
                        randseed=cfg.randseed)
    print(f"wrote {OUT}")


if __name__ == "__main__":
    main()
