"""The chaotic phase against the CPU FP64 oracle: 10000 steps of a 32^3
lattice with two fields and tensor modes, fluctuations drawn at the amplitude
of a rescaling B = 0.3 so that parametric resonance fragments the condensate,
the oracle's spectra every 50 steps and the statistics of an ensemble of
perturbed oracle runs committed as tests/golden/chaos_32.npz
(tests/golden/make_chaos_32.py).

The ensemble runs (fields perturbed by 1e-7 relative noise per site, FP32
rounding size) show the resonance: their distance from the oracle grows
about tenfold every 110 steps from step 100, reaches 0.24 of the fluctuation
amplitude at step 500 and the amplitude itself by step 700.  Two windows:
  early (steps 100 - 500): the runs are still apart by less than the
    fluctuation, so the device follows the oracle pointwise.  FP64 form:
    planes within 1e-8 of the fluctuation amplitude (rounding differences of
    1e-16 grown by the same factor stay near 1e-10 there).  FP32 form: each plane's
    deviation within 2.5 x the largest deviation of the ensemble at that
    step.
  chaotic (steps 7500 - 10000, 51 spectra): runs that started apart by
    rounding share nothing pointwise, so the comparison is statistical and
    the ensemble's scatter is the yardstick:
      * each spectrum bin's window mean (phi, chi, GW power) within 2.5 x the
        ensemble's largest deviation from the oracle in that bin;
      * the window means of K, V, G, a and rho within 2.5 x the ensemble's
        largest deviation from the oracle's;
      * the Friedmann residual R(t) over the window no worse than the
        oracle's: its mean at most the oracle's x (1 + 2 d_m), d_m the
        ensemble's largest relative deviation of that mean; its largest value
        at most the oracle's x (1 + 2 d) + 5 s, d the ensemble's largest
        relative excess and s the run's own step-to-step jitter of R (the rms
        second difference over sqrt 6).  R is a central difference of the
        per-step energy density: an FP64 state leaves it smooth (s ~ 1e-10,
        the oracle's and the FP64 form's alike), an FP32 state's rounding puts
        ~8e-7 of noise on it, which raises its largest values but not its
        mean.
Both forms are held to the chaotic-window bars.
"""
import os

import numpy as np
import pytest

from gabe_amd import configs, host
from tests.helpers import GABE_FP32, GABE_FP64, lattice_from_oracle, seeded_oracle

HERE = os.path.dirname(os.path.abspath(__file__))
FIXTURE = os.path.join(HERE, "golden", "chaos_32.npz")
SPEC_COLS = (0, 2, 4)     # phi power, chi power, GW power


def load_fixture():
    return np.load(FIXTURE, allow_pickle=False)


def window_means(spectra, hist, win0, every):
    """Window-mean spectra and the window means of K, V, G, a, rho from the
    per-step rows of steps win0 .. end (hist starts at win0 - every)."""
    spec = spectra[win0 // every:].mean(axis=0)
    rows = hist[every:]
    scal = np.array([rows[:, 3].mean(), rows[:, 4].mean(), rows[:, 5].mean(), rows[:, 1].mean(),
                     rows[:, 6].mean()])
    return spec, scal


def test_fixture_is_the_oracle_on_its_first_steps():
    """CPU: the committed fixture's first spectra are what the oracle gives
    now on the same seeded state, its ensemble ran apart from the oracle
    (the window is chaotic: the runs' final deviation is of the order of the
    fluctuation itself) and the windows' bookkeeping is consistent."""
    z = load_fixture()
    N = int(z["N"])
    cfg = configs.lambda_phi4_config(N=N, calc_gws=1)
    o = seeded_oracle(cfg, amp=float(z["amp"]))
    assert np.array_equal(o.specout(gw_quirk=1), z["spectra"][0])
    o.step(int(z["every"]))
    assert np.array_equal(o.specout(gw_quirk=1), z["spectra"][1])
    steps, every, win0 = int(z["steps"]), int(z["every"]), int(z["win0"])
    assert z["spectra"].shape[0] == steps // every + 1
    assert z["history"].shape[0] == steps - win0 + every + 1
    assert z["history"][0, 7] == win0 - every and z["history"][-1, 7] == steps
    dev = z["ens_dev"]
    # rounding-size starts: small early, of the fluctuation's own size in the window
    assert dev[:, 0].max() < 1e-4
    assert dev[:, win0 // every - 1:].min() > 0.1, dev[:, win0 // every - 1:].min()
    # the first stored plane is the oracle's plane after that many steps
    o.step(int(z["plane_steps"][0]) - every)
    assert np.array_equal(o.field[0][:, 0], z["planes"][0])


def _run_device(z, precision):
    N, steps, every = int(z["N"]), int(z["steps"]), int(z["every"])
    cfg = configs.lambda_phi4_config(N=N, calc_gws=1)
    o = seeded_oracle(cfg, amp=float(z["amp"]))
    lat = lattice_from_oracle(o, cfg, precision)
    plane_steps = [int(s) for s in z["plane_steps"]]
    spectra, planes = [lat.specOut()], []
    done = 0
    while done < steps:
        nxt = min(s for s in [*plane_steps, done - done % every + every] if s > done)
        lat.step(nxt - done)
        done = nxt
        if done in plane_steps:
            f, _ = lat.download_state()
            planes.append(f[:, 0].copy())
        if done % every == 0:
            spectra.append(lat.specOut())
    lat.step(1)     # its history row at start is the state after the last step
    hist = lat.history()
    lat.close()
    return np.array(spectra), np.array(planes), hist


@pytest.mark.gpu
@pytest.mark.parametrize("precision", [GABE_FP64, GABE_FP32])
def test_device_through_the_chaotic_phase(precision):
    z = load_fixture()
    steps, every, win0 = int(z["steps"]), int(z["every"]), int(z["win0"])
    fp64 = precision == GABE_FP64
    spectra, planes, hist = _run_device(z, precision)
    assert hist.shape[0] == steps + 1

    # early window: pointwise
    for p, s in enumerate(int(s) for s in z["plane_steps"]):
        ref = z["planes"][p]
        e = max(np.abs(planes[p][q] - ref[q]).max() / np.abs(ref[q] - ref[q].mean()).max()
                for q in range(2))
        bar = 1e-8 if fp64 else 2.5 * z["ens_dev"][:, s // every - 1].max()
        print(f"step {s}: plane deviation {e:.2e} (bar {bar:.2e})", flush=True)
        assert e < bar, (s, e, bar)

    # chaotic window: statistical
    hd = hist[win0 - every:]
    assert hd[0, 7] == win0 - every
    spec, scal = window_means(spectra, hd, win0, every)
    ospec, oscal = window_means(z["spectra"], z["history"], win0, every)
    worst = {}
    for c in SPEC_COLS:
        live = np.abs(ospec[c]) > 1e-9 * np.abs(ospec[c][1:]).max()
        live[0] = False
        spread = np.abs(z["ens_spec"][:, c, :] - ospec[c]).max(axis=0)
        d = np.abs(spec[c] - ospec[c])
        assert np.all(d[live] <= 2.5 * spread[live]), (c, np.nonzero(d[live] > 2.5 * spread[live]))
        worst[c] = float(np.max(d[live] / np.abs(ospec[c][live])))
    print("window-mean spectra, worst relative deviation per column:",
          {c: f"{v:.3f}" for c, v in worst.items()}, flush=True)
    sspread = np.abs(z["ens_scal"] - oscal).max(axis=0)
    assert np.all(np.abs(scal - oscal) <= 2.5 * sspread), (scal, oscal, sspread)

    _, Ro = host.friedmann_residual(z["history"])
    _, Rd = host.friedmann_residual(hd)
    ro, rd = Ro[every:-1], Rd[every:-1]
    d2 = rd[2:] - 2 * rd[1:-1] + rd[:-2]
    jitter = float(np.std(d2) / np.sqrt(6.0))
    ens_max, ens_mean = z["ens_resid"][:, 0], z["ens_resid"][:, 1]
    excess = max(0.0, float((ens_max / ro.max() - 1).max()))
    dm = float(np.abs(ens_mean / ro.mean() - 1).max())
    print(f"Friedmann residual over the window: device max {rd.max():.4e} mean {rd.mean():.4e} "
          f"jitter {jitter:.2e}; oracle max {ro.max():.4e} mean {ro.mean():.4e}; ensemble excess "
          f"{excess:.3f}, mean spread {dm:.4f}", flush=True)
    assert rd.mean() <= ro.mean() * (1 + 2 * dm)
    assert rd.max() <= ro.max() * (1 + 2 * excess) + 5 * jitter
