"""Device path vs the CPU FP64 oracle on the same seeded initial states.

Tolerances:
  FP64 device vs oracle: fields and GW modes within 1e-9 of the fluctuation
  amplitude, a/adot within 1e-13 relative, energies within 1e-12 of edrho
  (only summation order differs);
  FP32 form vs FP64 device: fields within 1e-4 of the fluctuation amplitude,
  GW modes within 1e-3 of their amplitude, a/adot within 1e-10 relative,
  over the early window tested here.
"""
import math

import os

import numpy as np
import pytest

from tests.helpers import (GABE_FP32, GABE_FP64, configs, fluct_err, lattice_from_oracle, po, rel_err,
                           seeded_oracle)

pytestmark = pytest.mark.gpu

SC_KEYS = ("a", "adot", "edkin", "edpot", "edgrad", "edrho")


def check_scalars(got, orc, rtol_a=1e-13, rtol_e=1e-12):
    for k in ("a", "adot"):
        assert abs(got[k][0] - orc[k][0]) <= rtol_a * abs(orc[k][0]), (k, got[k][0], orc[k][0])
    scale = abs(orc["edrho"][0])
    for k in ("edkin", "edpot", "edgrad", "edrho"):
        assert abs(got[k][0] - orc[k][0]) <= rtol_e * scale, (k, got[k][0], orc[k][0])


@pytest.mark.parametrize("N,gw", [(16, 1), (24, 1), (32, 1), (32, 0)])
def test_fp64_steps_match_oracle(N, gw):
    cfg = configs.lambda_phi4_config(N=N, calc_gws=gw)
    o = seeded_oracle(cfg, amp=2e-3)
    lat = lattice_from_oracle(o, cfg, GABE_FP64)
    for block in (1, 4, 15):
        o.step(block)
        lat.step(block)
        f, d = lat.download_state()
        assert fluct_err(f, o.field[0]) < 1e-9
        assert fluct_err(d, o.dfield[0]) < 1e-9
        check_scalars(lat.get_scalars(), o.scalars())
        if gw:
            h, l = lat.download_gw()
            assert rel_err(h, o.h[0]) < 1e-9
            assert rel_err(l, o.l[0]) < 1e-9


def test_fp64_reference_default_model():
    cfg = configs.reference_default_config(N=16, calc_gws=1)
    o = seeded_oracle(cfg, amp=1e-2)
    lat = lattice_from_oracle(o, cfg, GABE_FP64)
    o.step(10)
    lat.step(10)
    f, d = lat.download_state()
    assert fluct_err(f, o.field[0]) < 1e-9
    assert fluct_err(d, o.dfield[0]) < 1e-9
    check_scalars(lat.get_scalars(), o.scalars())
    h, l = lat.download_gw()
    assert rel_err(l, o.l[0]) < 1e-9


@pytest.mark.parametrize("exp_type,nflds", [(0, 2), (2, 2), (1, 1)])
def test_fp64_expansion_types_and_one_field(exp_type, nflds):
    cfg = configs.lambda_phi4_config(N=16, calc_gws=0)
    cfg.expansion_type = exp_type
    if nflds == 1:
        cfg.nflds, cfg.g2, cfg.f0, cfg.df0 = 1, 0.0, cfg.f0[:1], cfg.df0[:1]
    if exp_type == 2:
        # type 2 is user-defined: g2init.cpp:274-283 leaves a(0) unset; start it
        # as type 1 does (a = 1, adot = adf(0)), then evolve with the type-2 step
        cfg.expansion_type = 1
        o = seeded_oracle(cfg, amp=2e-3)
        o.st.p.expansion_type = 2
        cfg.expansion_type = 2
    else:
        o = seeded_oracle(cfg, amp=2e-3)
    lat = lattice_from_oracle(o, cfg, GABE_FP64)
    o.step(8)
    lat.step(8)
    f, d = lat.download_state()
    assert fluct_err(f, o.field[0]) < 1e-9
    assert fluct_err(d, o.dfield[0]) < 1e-9
    check_scalars(lat.get_scalars(), o.scalars())


@pytest.mark.parametrize("s", [0, 1])
def test_tt_source_matches_oracle(s):
    cfg = configs.lambda_phi4_config(N=16, calc_gws=1)
    o = seeded_oracle(cfg, amp=2e-3)
    lat = lattice_from_oracle(o, cfg, GABE_FP64)
    if s == 1:
        o.field[1] = o.field[0] + 0.5 * cfg.dt * o.dfield[0]
    o.fft_stresstensor(s)
    T = lat.tt_source(s)
    assert rel_err(T, o.T_gw[s]) < 1e-11


@pytest.mark.parametrize("N,quirk", [(16, 1), (16, 0), (24, 1)])
def test_spectra_match_oracle(N, quirk):
    cfg = configs.lambda_phi4_config(N=N, calc_gws=1)
    o = seeded_oracle(cfg, amp=2e-3)
    lat = lattice_from_oracle(o, cfg, GABE_FP64, gw_spectrum_quirk=quirk)
    o.step(6)
    lat.step(6)
    got = lat.specOut()
    ref = o.specout(gw_quirk=quirk)
    for col in range(ref.shape[0]):
        scale = np.abs(ref[col]).max()
        assert np.abs(got[col] - ref[col]).max() <= 1e-10 * scale, col


def test_effmass_and_fluctuation_add_match_oracle():
    cfg = configs.lambda_phi4_config(N=16, calc_gws=0)
    cfg.rescale_B = 1e-3
    o = seeded_oracle(cfg)  # reference initialisation on the CPU
    # the same modes, drawn by the oracle before any device call
    o2 = po.Oracle(cfg.N, nflds=2, L=cfg.L, dt=cfg.dt, m2_phi=cfg.m2_phi, lam=cfg.lam, g2=cfg.g2,
                   rescale_B=cfg.rescale_B)
    o2.initfields_homog(cfg.f0, cfg.df0)
    o2.initexpansion()
    po.srand(cfg.randseed)
    masses, modes = [], []
    for fld in range(2):
        masses.append(o2.effMass(0, fld))
        modes.append(o2.randinit(fld, masses[-1], want_modes=True))
    lat = lattice_from_oracle(o2, cfg, GABE_FP64)
    lat.upload_state(np.broadcast_to(np.array(cfg.f0)[:, None, None, None], o2.field[0].shape),
                     np.broadcast_to(np.array(cfg.df0)[:, None, None, None], o2.field[0].shape))
    for fld in range(2):
        m_dev = lat.effMass(fld)
        assert abs(m_dev - masses[fld]) <= 1e-13 * abs(masses[fld])
        lat.add_fluctuation(fld, *modes[fld])
    f, d = lat.download_state()
    assert fluct_err(f, o.field[0]) < 1e-11
    assert fluct_err(d, o.dfield[0]) < 1e-11


@pytest.mark.parametrize("precision,full_rand,cut", [(GABE_FP64, 1, 0.0), (GABE_FP32, 1, 0.0),
                                                    (GABE_FP64, 1, math.sqrt(3) / 8), (GABE_FP64, 0, math.sqrt(3) / 8)])
def test_host_initialize_matches_reference_init(precision, full_rand, cut):
    """gabe_host_initialize (the kept host's start-up) vs the oracle's
    g2main.cpp:109-119 sequence: same fields, energies and adot[0], for the
    default modes, the windowed modes and the symmetric modes."""
    from gabe_amd import host
    cfg = configs.lambda_phi4_config(N=24, calc_gws=1)
    cfg.rescale_B = 1e-3
    cfg.field_full_rand = full_rand
    cfg.spec_cut_off = cut
    o = seeded_oracle(cfg)
    lat = host.initialized_lattice(cfg, precision=precision)
    f, d = lat.download_state()
    tol = 1e-11 if precision == GABE_FP64 else 1e-6
    assert fluct_err(f, o.field[0]) < tol
    assert fluct_err(d, o.dfield[0]) < tol
    got, ref = lat.get_scalars(), o.scalars()
    assert abs(got["adot"][0] - ref["adot"][0]) <= 1e-13 * ref["adot"][0]   # homogeneous adot (Q4)
    # FP32 stores the fluctuation velocity in FP32: rho carries ~1e-12 of its rounding
    etol = 1e-12 if precision == GABE_FP64 else 1e-10
    assert abs(got["edrho"][0] - ref["edrho"][0]) <= etol * ref["edrho"][0]


@pytest.mark.parametrize("precision", [GABE_FP64, GABE_FP32])
def test_planes_and_means_vars(precision):
    """gabe_download_plane returns the planes of gabe_download_state;
    gabe_means_vars matches a two-pass numpy reduction of the downloaded
    state (mean to 1e-15 of the field, variance to 1e-9) and the oracle's
    g2output.cpp:70-121 formula."""
    cfg = configs.lambda_phi4_config(N=24, calc_gws=0)
    o = seeded_oracle(cfg, amp=2e-3)
    lat = lattice_from_oracle(o, cfg, precision)
    lat.step(5)
    f, d = lat.download_state()
    for i in (0, 1, 23):
        pf, pd = lat.download_plane(i)
        assert np.array_equal(pf, f[:, i]) and np.array_equal(pd, d[:, i])
    with pytest.raises(Exception):
        lat.download_plane(24)
    mv = lat.meansvars()
    o.step(5)
    ref = o.meansvars()
    for fld in range(2):
        x = f[fld].astype(np.float64)
        m = x.mean()
        v = ((x - m) ** 2).mean()
        assert abs(mv[2 * fld] - m) <= 1e-15 * max(abs(m), 1e-3)
        assert abs(mv[2 * fld + 1] - v) <= 1e-9 * v
        tol = 1e-9 if precision == GABE_FP64 else 1e-5
        assert abs(mv[2 * fld + 1] - ref[2 * fld + 1]) <= tol * ref[2 * fld + 1]


@pytest.mark.parametrize("N", [32, 48])
def test_fp32_form_tracks_fp64(N):
    cfg = configs.lambda_phi4_config(N=N, calc_gws=1)
    o = seeded_oracle(cfg)  # physical amplitude: fluctuations ~1e-6 of the mean
    l64 = lattice_from_oracle(o, cfg, GABE_FP64)
    l32 = lattice_from_oracle(o, cfg, GABE_FP32)
    l64.step(40)
    l32.step(40)
    f64, d64 = l64.download_state()
    f32, d32 = l32.download_state()
    assert fluct_err(f32, f64) < 1e-4
    assert fluct_err(d32, d64) < 1e-4
    s64, s32 = l64.get_scalars(), l32.get_scalars()
    for k in ("a", "adot"):
        assert abs(s32[k][0] - s64[k][0]) <= 1e-10 * abs(s64[k][0])
    h64, ll64 = l64.download_gw()
    h32, ll32 = l32.download_gw()
    assert rel_err(ll32, ll64) < 1e-3
    sp64, sp32 = l64.specOut(), l32.specOut()
    for col in range(sp64.shape[0]):
        # bins below 1e-9 of the column maximum are roundoff (e.g. the DC bin of
        # a zero-mean field) and carry no physics
        nz = np.abs(sp64[col]) > 1e-9 * np.abs(sp64[col]).max()
        assert np.max(np.abs(sp32[col][nz] / sp64[col][nz] - 1)) < 1e-3, col


def test_history_and_deterministic_replay():
    cfg = configs.lambda_phi4_config(N=16, calc_gws=1)
    o = seeded_oracle(cfg, amp=2e-3)
    a = lattice_from_oracle(o, cfg, GABE_FP32)
    b = lattice_from_oracle(o, cfg, GABE_FP32)
    a.step(5)
    a.get_scalars()   # forces an energy pass between steps: must not perturb the trajectory
    a.step(5)
    b.step(10)
    fa, da = a.download_state()
    fb, db = b.download_state()
    assert np.array_equal(fa, fb) and np.array_equal(da, db)
    ha, hb = a.history(), b.history()
    assert ha.shape == (10, 8)
    assert np.array_equal(ha, hb)
    assert np.all(np.diff(ha[:, 7]) == 1)


def test_gw_upload_keeps_the_tt_part():
    """Upload/download of tensor modes: a TT tensor round-trips, any other
    tensor comes back as its TT projection (the oracle's source is TT)."""
    cfg = configs.lambda_phi4_config(N=16, calc_gws=1)
    o = seeded_oracle(cfg, amp=2e-3)
    o.fft_stresstensor(0)
    tt = o.T_gw[0].copy()
    lat = lattice_from_oracle(o, cfg, GABE_FP64)
    lat.upload_gw(tt, 2.0 * tt)
    h, l = lat.download_gw()
    assert rel_err(h, tt) < 1e-13 and rel_err(l, 2.0 * tt) < 1e-13
    rng = np.random.default_rng(7)
    junk = rng.standard_normal(tt.shape)
    lat.upload_gw(tt + 0.0 * junk, None)
    # a pure-trace tensor (delta_ij f(k)) has no TT part
    trace = np.zeros_like(tt)
    for c in (0, 3, 5):
        trace[2 * c] = junk[0]
        trace[2 * c + 1] = junk[1]
    lat.upload_gw(trace, None)
    h2, _ = lat.download_gw()
    assert np.abs(h2).max() < 1e-13 * np.abs(junk).max()


@pytest.mark.parametrize("precision,N", [(GABE_FP32, 16), (GABE_FP32, 32), (GABE_FP32, 64), (GABE_FP32, 128),
                                         (GABE_FP32, 256), (GABE_FP64, 16), (GABE_FP64, 32), (GABE_FP64, 64),
                                         (GABE_FP64, 128), (GABE_FP64, 256)])
def test_fused_gw_transform_matches_separate_passes(precision, N, monkeypatch):
    """The i transform fused with the mode update gives the same tensor modes
    and fields as the separate i pass followed by the mode kernel, in both
    storage forms (FP64: the k rows come from the library transform, FP32:
    from sweep A)."""
    cfg = configs.lambda_phi4_config(N=N, calc_gws=1)
    o = seeded_oracle(cfg, amp=2e-3)
    fused = lattice_from_oracle(o, cfg, precision)
    monkeypatch.setenv("GABE_UNFUSED_GW", "1")
    sep = lattice_from_oracle(o, cfg, precision)
    monkeypatch.delenv("GABE_UNFUSED_GW")
    fused.step(12)
    sep.step(12)
    hf, lf = fused.download_gw()
    hs, ls = sep.download_gw()
    tol = 1e-5 if precision == GABE_FP32 else 1e-12
    assert rel_err(lf, ls) < tol and rel_err(hf, hs) < tol
    ff, _ = fused.download_state()
    fs, _ = sep.download_state()
    assert np.array_equal(ff, fs)   # the field path does not depend on the GW kernels


def test_fused_gw_transform_matches_separate_passes_512():
    """FP32 at 512^3: k rows transformed in sweep A (512-thread workgroups)
    and the fused i transform + modes give the tensor modes of the library
    rows with the separate i pass and mode kernel, on the reference's own
    initial conditions generated on the device.  The fields agree to FP32
    rounding rather than bitwise: a 512-thread workgroup adds its eight waves'
    energy sums before the finalize, where two 256-thread workgroups leave two
    partials, so the FP64 sums (and a) differ in the last bits."""
    from gabe_amd import host
    cfg = configs.lambda_phi4_config(N=512, calc_gws=1)
    fused = host.initialized_lattice(cfg, precision=GABE_FP32)
    os.environ["GABE_UNFUSED_GW"] = "1"
    try:
        sep = host.initialized_lattice(cfg, precision=GABE_FP32)
    finally:
        del os.environ["GABE_UNFUSED_GW"]
    fused.step(3)
    sep.step(3)
    hf, lf = fused.download_gw()
    hs, ls = sep.download_gw()
    assert rel_err(lf, ls) < 1e-6 and rel_err(hf, hs) < 1e-6
    ff, _ = fused.download_state()
    fs, _ = sep.download_state()
    for f in range(ff.shape[0]):
        fl = fs[f] - fs[f].mean()
        assert np.linalg.norm(ff[f] - fs[f]) <= 1e-6 * np.linalg.norm(fl)
    fused.close()
    sep.close()


@pytest.mark.parametrize("precision,N", [(GABE_FP32, 16), (GABE_FP32, 32), (GABE_FP32, 64), (GABE_FP32, 128),
                                         (GABE_FP32, 256), (GABE_FP64, 16), (GABE_FP64, 64), (GABE_FP64, 256)])
def test_rows_transformed_in_sweep_match_library_rows(precision, N, monkeypatch):
    """Sweep A transforming its own k-rows of the stress (one complex
    transform per pair of real rows, split into both half-spectra) gives the
    tensor modes of the library row transform, in both storage forms; the
    stress transform it feeds is checked directly through gabe_tt_source."""
    cfg = configs.lambda_phi4_config(N=N, calc_gws=1)
    o = seeded_oracle(cfg, amp=2e-3)
    fused = lattice_from_oracle(o, cfg, precision)
    monkeypatch.setenv("GABE_LIBRARY_ROWS", "1")
    lib = lattice_from_oracle(o, cfg, precision)
    monkeypatch.delenv("GABE_LIBRARY_ROWS")
    fp32 = precision == GABE_FP32
    for s in (0, 1):
        tf, tl = fused.tt_source(s), lib.tt_source(s)
        assert rel_err(tf, tl) < (2e-6 if fp32 else 1e-13), s
    fused.step(6)
    lib.step(6)
    hf, lf = fused.download_gw()
    hl, ll = lib.download_gw()
    tol = 1e-5 if fp32 else 1e-12
    assert rel_err(lf, ll) < tol and rel_err(hf, hl) < tol


@pytest.mark.parametrize("precision,N,ppb,gw", [(GABE_FP32, 16, 16, 1), (GABE_FP32, 32, 7, 0), (GABE_FP32, 48, 16, 1),
                                                (GABE_FP32, 64, 5, 1), (GABE_FP32, 96, 32, 0), (GABE_FP32, 128, 32, 1),
                                                (GABE_FP32, 256, 32, 1), (GABE_FP64, 32, 7, 1), (GABE_FP64, 64, 32, 0),
                                                (GABE_FP64, 128, 16, 1)])
def test_tiled_sweeps_match_untiled(precision, N, ppb, gw, monkeypatch):
    """Sweeps reading their planes from LDS-staged tiles (64 x 4, 32 x 8 or
    16 x 16 columns with one-site edges, two planes in flight) against the
    per-thread sweeps that load every neighbour themselves, for each tile
    shape and for plane counts per workgroup that leave partial march groups.
    With tensor modes, or in FP64, only sweep B is tiled and the bits agree;
    in FP32 without them sweep A is tiled too, its sums are grouped by tile
    rather than by row, and the energies agree to FP64 rounding, the fields to
    FP32 rounding."""
    from gabe_amd import host
    cfg = configs.lambda_phi4_config(N=N, calc_gws=gw)
    tiled = host.initialized_lattice(cfg, precision=precision, planes_per_block=ppb)
    monkeypatch.setenv("GABE_UNTILED_SWEEPS", "1")
    plain = host.initialized_lattice(cfg, precision=precision, planes_per_block=ppb)
    monkeypatch.delenv("GABE_UNTILED_SWEEPS")
    bitwise = gw or precision == GABE_FP64
    st, sp = tiled.get_scalars(), plain.get_scalars()
    if not bitwise:
        # the energies of the initial state: one tiled sweep A against one per-thread sweep A
        for k in ("edkin", "edpot", "edgrad", "edrho"):
            assert abs(st[k][0] - sp[k][0]) <= 1e-14 * abs(sp["edrho"][0]), k
    tiled.step(9)
    plain.step(9)
    ft, dt = tiled.download_state()
    fp, dp = plain.download_state()
    st, sp = tiled.get_scalars(), plain.get_scalars()
    if bitwise:
        assert np.array_equal(ft, fp) and np.array_equal(dt, dp)
        assert st["a"][0] == sp["a"][0]
    else:
        assert fluct_err(ft, fp) < 1e-5 and fluct_err(dt, dp) < 1e-5
        assert abs(st["a"][0] - sp["a"][0]) <= 1e-14 * sp["a"][0]
        assert abs(st["edrho"][0] - sp["edrho"][0]) <= 1e-12 * abs(sp["edrho"][0])
    if gw:
        ht, lt = tiled.download_gw()
        hp, lp = plain.download_gw()
        assert np.array_equal(ht, hp) and np.array_equal(lt, lp)
    tiled.close()
    plain.close()


@pytest.mark.parametrize("ppb", [5, 6, 7])
def test_plane_counts_per_workgroup_agree(ppb):
    """Sweep workgroups marching 5, 6 or 7 planes give the fields and tensor
    modes of the default tiling (1 plane at 32^3): FP32 sweep B's march
    unrolled by four runs with a 1-, 2- or 3-plane tail, and sweep A's row
    transforms end on a single plane.  The energy partials are grouped
    differently, so a(t) may differ in its last bits: agreement to FP32
    rounding."""
    cfg = configs.lambda_phi4_config(N=32, calc_gws=1)
    o = seeded_oracle(cfg, amp=2e-3)
    ref = lattice_from_oracle(o, cfg, GABE_FP32)
    lat = lattice_from_oracle(o, cfg, GABE_FP32, planes_per_block=ppb)
    assert lat.query().planes_per_block == ppb
    ref.step(8)
    lat.step(8)
    f0, d0 = ref.download_state()
    f1, d1 = lat.download_state()
    assert fluct_err(f1, f0) < 1e-6 and fluct_err(d1, d0) < 1e-6
    h0, l0 = ref.download_gw()
    h1, l1 = lat.download_gw()
    assert rel_err(l1, l0) < 1e-5 and rel_err(h1, h0) < 1e-5


def test_odd_plane_count_at_256():
    """256^3 FP32 with 7 planes per sweep workgroup: the last row-transform call
    of each workgroup holds one plane (the one-row split path with a single
    plane), checked against the default 32-plane tiling on the reference's own
    initial conditions generated on the device."""
    from gabe_amd import host
    cfg = configs.lambda_phi4_config(N=256, calc_gws=1)
    ref = host.initialized_lattice(cfg, precision=GABE_FP32)
    lat = host.initialized_lattice(cfg, precision=GABE_FP32, planes_per_block=7)
    assert lat.query().planes_per_block == 7
    ref.step(3)
    lat.step(3)
    h0, l0 = ref.download_gw()
    h1, l1 = lat.download_gw()
    assert rel_err(l1, l0) < 1e-5 and rel_err(h1, h0) < 1e-5
    f0, _ = ref.download_state()
    f1, _ = lat.download_state()
    for f in range(f0.shape[0]):
        fl = f0[f] - f0[f].mean()
        assert np.linalg.norm(f1[f] - f0[f]) <= 1e-6 * np.linalg.norm(fl)
    ref.close()
    lat.close()


@pytest.mark.parametrize("precision", [GABE_FP32, GABE_FP64])
def test_transform_and_mode_tiles_agree_at_256(precision):
    """Every tile variant of the fused i-transform + mode kernel, and the
    32-column j transform, give the tensor modes of the default tiling (the
    one checked against the oracle at 256^3) to rounding: the variants only
    regroup the same transforms and mode updates over workgroups."""
    from gabe_amd import host
    from gabe_amd.lattice import mode_tiles
    cfg = configs.lambda_phi4_config(N=256, calc_gws=1)
    ref = host.initialized_lattice(cfg, precision=precision, transform_tile=16, mode_tile=1)
    ref.step(2)
    h0, l0 = ref.download_gw()
    f0, _ = ref.download_state()
    ref.close()
    tol = 1e-6 if precision == GABE_FP32 else 1e-13
    n = mode_tiles(256, precision)
    assert n == (7 if precision == GABE_FP32 else 3)
    for mt, tt in [(m, 16) for m in range(2, n + 1)] + [(1, 32)]:
        lat = host.initialized_lattice(cfg, precision=precision, transform_tile=tt, mode_tile=mt)
        q = lat.query()
        assert (q.mode_tile, q.transform_tile) == (mt, tt)
        lat.step(2)
        h, l = lat.download_gw()
        print(f"mode tile {mt}, transform tile {tt}: h {rel_err(h, h0):.1e} l {rel_err(l, l0):.1e}", flush=True)
        assert rel_err(h, h0) < tol and rel_err(l, l0) < tol, (mt, tt)
        f, _ = lat.download_state()
        assert np.array_equal(f, f0)   # the field path does not depend on these tiles
        lat.close()


@pytest.mark.parametrize("N,precision", [(256, GABE_FP32), (256, GABE_FP64), (128, GABE_FP32), (512, GABE_FP32)])
def test_interleaved_stress_chunks_are_bitwise_equal(N, precision):
    """Sweep A and the j-transform interleaved in launches of a few plane
    blocks (each launch's spectra transformed along j on a stream of their
    own) run the same kernels on the same values as the whole-lattice
    launches: fields, velocities, h, l and the scalars after 3 steps are
    bitwise equal, for every chunk size the autotuner tries."""
    from gabe_amd import host
    cfg = configs.lambda_phi4_config(N=N, calc_gws=1)
    ppb = N // 16
    ref = host.initialized_lattice(cfg, precision=precision, planes_per_block=ppb, stress_chunk=-1)
    assert ref.query().stress_chunk == 0
    ref.step(3)
    f0, v0 = ref.download_state()
    h0, l0 = ref.download_gw()
    s0 = ref.get_scalars()
    ref.close()
    for planes in (N // 8, N // 4, N // 2):
        lat = host.initialized_lattice(cfg, precision=precision, planes_per_block=ppb, stress_chunk=planes)
        assert lat.query().stress_chunk == planes
        lat.step(3)
        f, v = lat.download_state()
        h, l = lat.download_gw()
        s = lat.get_scalars()
        lat.close()
        print(f"N={N} chunk {planes} planes: fields {np.array_equal(f, f0)}, h/l {np.array_equal(h, h0)}", flush=True)
        assert np.array_equal(f, f0) and np.array_equal(v, v0), planes
        assert np.array_equal(h, h0) and np.array_equal(l, l0), planes
        assert s == s0, planes


def test_tile_parameters_are_validated():
    cfg = configs.lambda_phi4_config(N=32, calc_gws=1)
    from gabe_amd.lattice import GabeError, Lattice
    with pytest.raises(GabeError):
        Lattice(precision=GABE_FP32, transform_tile=8, **cfg.lattice_kwargs())
    with pytest.raises(GabeError):
        Lattice(precision=GABE_FP32, mode_tile=2, **cfg.lattice_kwargs())
    with pytest.raises(GabeError):
        Lattice(precision=GABE_FP32, stress_chunk=-2, **cfg.lattice_kwargs())
