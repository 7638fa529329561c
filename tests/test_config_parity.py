"""Device path against the CPU FP64 oracle at the sizes of the benchmark
workloads themselves, from the reference's own seeded initial conditions at
their physical amplitude (no enlarged fluctuations):

* 256^3 with tensor modes, 2 steps, FP64 and FP32 forms: pins the kernels
  the throughput figure comes from (k rows transformed in sweep A with the
  one-row split, the 256-wide j transform, the fused i transform + modes);
* 512^3 with tensor modes (config 3's lattice), 2 steps, both forms;
* 64^3 FP64 without tensor modes over 200 steps (the plain CPU workload),
  with the FP32 form beside it;
* 256^3 FP32 against FP64 over the early window (200 steps) and over config 2's
  2000 steps with spectra every 200;
* 512^3 FP32 against FP64 over 400 steps with spectra every 50 (config 3's
  cadence; the full 4000-step run is gabe_amd.precision's);
* a 64-lattice ensemble of 32^3 ... 128^3 members (bitwise against single
  runs), and an FP64 96^3 member against the oracle;
* the learned closure's network at 128^3 against the CPU network.

Tolerances (stated per test below):
  FP64 vs oracle: fields and tensor modes within 1e-9 of their amplitude
  (1e-8 after 200 steps), a and adot within 1e-13, energies within 1e-12 of
  rho;
  FP32 form vs oracle: fields within 1e-5 of the fluctuation amplitude after
  2 steps and 1e-4 after 200, tensor modes within 1e-4, a within 1e-12.
"""
import os

import numpy as np
import pytest

from gabe_amd import configs, host
from gabe_amd.lattice import GABE_FP32, GABE_FP64, Ensemble, Lattice
from tests.helpers import fluct_err, lattice_from_oracle, rel_err, seeded_oracle

pytestmark = pytest.mark.gpu

THREADS = max(1, min(16, os.cpu_count() or 1))


def live_bins(ref):
    """bins above 1e-9 of the column's largest value outside the DC bin, which
    is left out (it holds the field's mean, or rounding for a zero-mean one)"""
    assert np.all(np.isfinite(ref))
    live = np.abs(ref) > 1e-9 * np.abs(ref[1:]).max()
    live[0] = False
    assert live.any()
    return live


def report_nonfinite(name, sp):
    bad = [(c, np.flatnonzero(~np.isfinite(sp[c])).tolist()) for c in range(sp.shape[0]) if not np.all(np.isfinite(sp[c]))]
    if bad:
        print(f"{name}: non-finite spectrum bins (column, bins): {bad}", flush=True)
    assert not bad, name


def check_scalars(got, orc, rtol_a=1e-13, rtol_e=1e-12):
    for k in ("a", "adot"):
        assert abs(got[k][0] - orc[k][0]) <= rtol_a * abs(orc[k][0]), (k, got[k][0], orc[k][0])
    scale = abs(orc["edrho"][0])
    for k in ("edkin", "edpot", "edgrad", "edrho"):
        assert abs(got[k][0] - orc[k][0]) <= rtol_e * scale, (k, got[k][0], orc[k][0])


def test_256_two_steps_match_oracle():
    """256^3, two fields + tensor modes, the reference's seeded vacuum
    fluctuations (randseed 44463132) on the slow-roll background: 2 steps of
    the FP64 form and of the FP32 form against 2 oracle steps."""
    cfg = configs.lambda_phi4_config(N=256, calc_gws=1)
    o = seeded_oracle(cfg, nthreads=THREADS)
    lat64 = lattice_from_oracle(o, cfg, GABE_FP64)
    lat32 = lattice_from_oracle(o, cfg, GABE_FP32)
    q = lat32.query()
    print(f"FP32 tiling: {q.planes_per_block} planes per sweep workgroup", flush=True)
    o.step(2)
    sc = o.scalars()
    ref_sp = o.specout(gw_quirk=1)
    results = {}
    for name, lat in (("fp64", lat64), ("fp32", lat32)):
        lat.step(2)
        sp = lat.specOut()
        report_nonfinite(name, sp)
        for col in range(ref_sp.shape[0]):
            live = live_bins(ref_sp[col])
            err = float(np.max(np.abs(sp[col][live] / ref_sp[col][live] - 1)))
            print(f"{name} spectrum column {col}: worst bin {err:.2e}", flush=True)
            assert err < (1e-8 if name == "fp64" else 1e-3), (name, col)
        f, d = lat.download_state()
        ef, ed = fluct_err(f, o.field[0]), fluct_err(d, o.dfield[0])
        del f, d
        h, l = lat.download_gw()
        eh, el = rel_err(h, o.h[0]), rel_err(l, o.l[0])
        del h, l
        results[name] = (ef, ed, eh, el, lat.get_scalars())
        print(f"{name}: field {ef:.2e} velocity {ed:.2e} h {eh:.2e} l {el:.2e}", flush=True)
        lat.close()
    ef, ed, eh, el, got = results["fp64"]
    assert ef < 1e-9 and ed < 1e-9 and eh < 1e-9 and el < 1e-9
    check_scalars(got, sc)
    ef, ed, eh, el, got = results["fp32"]
    assert ef < 1e-5 and ed < 1e-5 and eh < 1e-4 and el < 1e-4
    check_scalars(got, sc, rtol_a=1e-12, rtol_e=1e-10)


def test_512_two_steps_match_oracle():
    """Config 3's lattice: 512^3 with tensor modes, 2 steps from the seeded
    vacuum state, FP64 and FP32 forms against 2 oracle steps.  Pins the 512^3
    kernels: the FP32 form's k rows on 512-thread workgroups and its 512-wide
    mode tiles, the FP64 form's library rows and separate i pass.  Tolerances
    as at 256^3.  (About 90 s, most of it the oracle; its state at 512^3 with
    tensor modes holds about 60 GB of host memory.)"""
    cfg = configs.lambda_phi4_config(N=512, calc_gws=1)
    o = seeded_oracle(cfg, nthreads=THREADS)
    lats = {name: lattice_from_oracle(o, cfg, prec) for name, prec in (("fp64", GABE_FP64), ("fp32", GABE_FP32))}
    o.step(2)
    sc = o.scalars()
    ref_sp = o.specout(gw_quirk=1)
    for name, lat in lats.items():
        lat.step(2)
        sp = lat.specOut()
        report_nonfinite(name, sp)
        for col in range(ref_sp.shape[0]):
            live = live_bins(ref_sp[col])
            err = float(np.max(np.abs(sp[col][live] / ref_sp[col][live] - 1)))
            print(f"{name} spectrum column {col}: worst bin {err:.2e}", flush=True)
            assert err < (1e-8 if name == "fp64" else 1e-3), (name, col)
        f, d = lat.download_state()
        ef, ed = fluct_err(f, o.field[0]), fluct_err(d, o.dfield[0])
        del f, d
        h, l = lat.download_gw()
        eh, el = rel_err(h, o.h[0]), rel_err(l, o.l[0])
        del h, l
        print(f"{name}: field {ef:.2e} velocity {ed:.2e} h {eh:.2e} l {el:.2e}", flush=True)
        if name == "fp64":
            assert ef < 1e-9 and ed < 1e-9 and eh < 1e-9 and el < 1e-9
            check_scalars(lat.get_scalars(), sc)
        else:
            assert ef < 1e-5 and ed < 1e-5 and eh < 1e-4 and el < 1e-4
            check_scalars(lat.get_scalars(), sc, rtol_a=1e-12, rtol_e=1e-10)
        lat.close()


def test_64_fp64_200_steps_match_oracle():
    """The plain CPU workload: 64^3, no tensor modes, 200 steps, compared
    every 50 steps; the FP32 form against the same oracle run."""
    cfg = configs.lambda_phi4_config(N=64, calc_gws=0)
    o = seeded_oracle(cfg, nthreads=THREADS)
    lat64 = lattice_from_oracle(o, cfg, GABE_FP64)
    lat32 = lattice_from_oracle(o, cfg, GABE_FP32)
    for n in range(4):
        o.step(50)
        lat64.step(50)
        lat32.step(50)
        f, d = lat64.download_state()
        e64 = (fluct_err(f, o.field[0]), fluct_err(d, o.dfield[0]))
        f, d = lat32.download_state()
        e32 = (fluct_err(f, o.field[0]), fluct_err(d, o.dfield[0]))
        print(f"step {50 * (n + 1)}: FP64 {e64[0]:.2e} {e64[1]:.2e}  FP32 {e32[0]:.2e} {e32[1]:.2e}", flush=True)
        assert max(e64) < 1e-8, n
        assert max(e32) < 1e-4, n
        check_scalars(lat64.get_scalars(), o.scalars())
        s32 = lat32.get_scalars()
        for k in ("a", "adot"):
            assert abs(s32[k][0] - o.scalars()[k][0]) <= 1e-12 * abs(o.scalars()[k][0]), k
    sp64, sp32, ref = lat64.specOut(), lat32.specOut(), o.specout()
    for col in range(ref.shape[0]):
        scale = np.abs(ref[col]).max()
        assert np.abs(sp64[col] - ref[col]).max() <= 1e-8 * scale, col
    for col in (0, 2):   # field spectra
        live = live_bins(ref[col])
        assert np.max(np.abs(sp32[col][live] / ref[col][live] - 1)) < 1e-3, col


def test_256_fp32_tracks_fp64_over_the_early_window():
    """256^3 with tensor modes, both forms from the same device-generated
    initial state (the kept host's start-up), 200 steps: field L2 error within
    1e-5 of the fluctuation, per-bin field and GW spectra within 1e-3, rho and
    a within 1e-12, and the FP32 form's Friedmann residual no worse than the
    FP64 form's (to 1e-6 of it)."""
    cfg = configs.lambda_phi4_config(N=256, calc_gws=1)
    out = {}
    for name, prec in (("fp64", GABE_FP64), ("fp32", GABE_FP32)):
        lat = host.initialized_lattice(cfg, precision=prec)
        lat.step(200)
        out[name] = dict(state=lat.download_state(), spectra=lat.specOut(), scalars=lat.get_scalars(),
                         friedmann=host.friedmann_residual(lat.history())[0])
        lat.close()
    a, b = out["fp32"], out["fp64"]
    for k in (0, 1):
        for f in range(2):
            ref = b["state"][k][f]
            fl = ref - ref.mean()
            err = np.linalg.norm(a["state"][k][f] - ref) / np.linalg.norm(fl)
            print(f"array {k} field {f}: L2 {err:.2e}", flush=True)
            assert err < 1e-5, (k, f)
    report_nonfinite("fp64", b["spectra"])
    report_nonfinite("fp32", a["spectra"])
    for col in (0, 2, 4):
        ref, got = b["spectra"][col], a["spectra"][col]
        live = live_bins(ref)
        worst = np.max(np.abs(got[live] / ref[live] - 1))
        print(f"spectrum column {col}: worst bin {worst:.2e}", flush=True)
        assert worst < 1e-3, col
    for k in ("a", "edrho"):
        assert abs(a["scalars"][k][0] - b["scalars"][k][0]) <= 1e-12 * abs(b["scalars"][k][0]), k
    assert a["friedmann"] <= b["friedmann"] * (1 + 1e-6), (a["friedmann"], b["friedmann"])


def test_256_fp32_tracks_fp64_over_config2_2000_steps():
    """Config 2 in full: 256^3 with tensor modes, the FP32 form against the
    FP64 form from the same device-generated initial state over 2000 steps,
    spectra of both every 200 steps (gabe_amd.precision): every output's field
    and GW spectra within 1e-3 per bin, the final field L2 error within 1e-5 of
    the fluctuation, rho and a within 1e-12, and the FP32 form's Friedmann
    residual over the run no worse than the FP64 form's (to 1e-6 of it)."""
    from gabe_amd import precision
    res = precision.run(256, 2000, 200, log=lambda m: print(m, flush=True))
    assert len(res["series"]) == 10
    for row in res["series"]:
        assert row["field_max_rel"] < 1e-3 and row["gw_max_rel"] < 1e-3, row
    assert max(res["field_l2_rel"]) < 1e-5, res["field_l2_rel"]
    assert res["edrho_rel"] < 1e-12 and res["a_rel"] < 1e-12, (res["edrho_rel"], res["a_rel"])
    assert res["friedmann_max_fp32"] <= res["friedmann_max_fp64"] * (1 + 1e-6)
    print(f"worst bins: field {res['field_spectra_max_rel']:.2e}, GW {res['gw_spectrum_max_rel']:.2e}; "
          f"field L2 {max(res['field_l2_rel']):.2e}; Friedmann {res['friedmann_max_fp32']:.8e} / "
          f"{res['friedmann_max_fp64']:.8e}", flush=True)


def test_512_fp32_tracks_fp64_with_spectra_every_50_steps():
    """Config 3's lattice and output cadence, shortened: 512^3 with tensor
    modes, both forms from the same device-generated initial state, 400 steps
    with the spectra of both every 50 steps (gabe_amd.precision, the tool the
    full 4000-step run uses): every output's field and GW spectra within 1e-3
    per bin (the target for spectra against FP64), field L2 error within 1e-5
    of the fluctuation, rho and a within 1e-12, and the FP32 form's Friedmann
    residual no worse than the FP64 form's (to 1e-6 of it)."""
    from gabe_amd import precision
    res = precision.run(512, 400, 50, log=lambda m: print(m, flush=True))
    assert len(res["series"]) == 8
    for row in res["series"]:
        assert row["field_max_rel"] < 1e-3 and row["gw_max_rel"] < 1e-3, row
    assert max(res["field_l2_rel"]) < 1e-5, res["field_l2_rel"]
    assert res["edrho_rel"] < 1e-12 and res["a_rel"] < 1e-12, (res["edrho_rel"], res["a_rel"])
    assert res["friedmann_max_fp32"] <= res["friedmann_max_fp64"] * (1 + 1e-6)
    print(f"worst bins: field {res['field_spectra_max_rel']:.2e}, GW {res['gw_spectrum_max_rel']:.2e}; "
          f"{res['fp32_ms_per_step']:.2f} / {res['fp64_ms_per_step']:.2f} ms per step", flush=True)


def test_ensemble_workload_members_match_single_runs():
    """The 64-lattice ensemble (16 each of 32^3, 64^3, 96^3, 128^3, g^2/lambda
    log-uniform in [1, 200], distinct seeds, FP32 form with autotuned planes
    per workgroup): after 12 steps, one member of each size and the last
    member are bitwise equal to the same lattice stepped alone."""
    cfgs = configs.ensemble_configs()
    ens = Ensemble([dict(precision=GABE_FP32, **c.lattice_kwargs()) for c in cfgs])
    for c, m in zip(cfgs, ens.members):
        host.initialize_member(m, c)
    ens.step(12)
    ens.sync()
    for idx in (0, 17, 34, 51, 63):
        c, m = cfgs[idx], ens.members[idx]
        ppb = m.query().planes_per_block
        s = Lattice(precision=GABE_FP32, planes_per_block=ppb, **c.lattice_kwargs())
        host.initialize(s, c)
        s.step(12)
        fm, dm = m.download_state()
        fs, ds = s.download_state()
        print(f"member {idx}: {c.N}^3, {ppb} planes per workgroup", flush=True)
        assert np.array_equal(fm, fs) and np.array_equal(dm, ds), idx
        sm, ss = m.get_scalars(), s.get_scalars()
        for k in ("a", "adot", "edrho"):
            assert sm[k][0] == ss[k][0], (idx, k)
        assert np.array_equal(m.history(), s.history()), idx
        s.close()
    ens.close()


def test_ensemble_fp64_96_member_matches_oracle():
    """An FP64 96^3 member (batched with a 128^3 and a 32^3 member) against
    the oracle over 10 steps from the seeded reference state."""
    cfg = configs.lambda_phi4_config(N=96, calc_gws=0, g2_over_lambda=37.0)
    o = seeded_oracle(cfg, nthreads=THREADS)
    others = [configs.lambda_phi4_config(N=n, calc_gws=0) for n in (128, 32)]
    ens = Ensemble([dict(precision=GABE_FP64, **c.lattice_kwargs()) for c in [others[0], cfg, others[1]]])
    m = ens.members[1]
    m.upload_state(o.field[0], o.dfield[0])
    sc = o.scalars()
    m.set_scalars(sc["a"], sc["adot"], sc["edkin"], sc["edpot"], sc["edgrad"], sc["edrho"])
    for c, mem in ((others[0], ens.members[0]), (others[1], ens.members[2])):
        host.initialize_member(mem, c)
    o.step(10)
    ens.step(10)
    f, d = m.download_state()
    assert fluct_err(f, o.field[0]) < 1e-9 and fluct_err(d, o.dfield[0]) < 1e-9
    check_scalars(m.get_scalars(), o.scalars())
    ens.close()
