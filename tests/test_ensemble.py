"""Ensembles: every member of a batched ensemble (one launch per RK2 phase for
all members, two steps per HIP graph) must give the bits of the same lattice
stepped alone, and an FP64 member must match the oracle."""
import math

import numpy as np
import pytest

from gabe_amd import configs, host
from gabe_amd.lattice import GABE_FP32, GABE_FP64, Ensemble, Lattice, autotune_sweep
from tests.helpers import fluct_err, seeded_oracle

pytestmark = pytest.mark.gpu


def member_cfgs(sizes, seed=3):
    rng = np.random.default_rng(seed)
    out = []
    for n, N in enumerate(sizes):
        ratio = math.exp(rng.uniform(math.log(1.0), math.log(200.0)))   # g^2/lambda log-uniform in [1, 200]
        cfg = configs.lambda_phi4_config(N=N, calc_gws=0, g2_over_lambda=ratio)
        cfg.randseed = 1000 + n
        cfg.rescale_B = 1e-3
        out.append(cfg)
    return out


@pytest.mark.parametrize("nsteps,sizes", [(10, (16, 24, 32, 16)), (7, (16, 24, 32, 16)), (9, (16, 32, 48, 64))])
def test_members_match_single_lattices_bitwise(nsteps, sizes):
    """With a 24^3 member the batched sweep B is the per-thread one while the
    single 16^3 and 32^3 lattices use the LDS-tiled one; with sizes that are
    all multiples of 16 both sides are tiled."""
    cfgs = member_cfgs(list(sizes))
    ens = Ensemble([dict(precision=GABE_FP32, **c.lattice_kwargs()) for c in cfgs])
    singles = []
    for c, m in zip(cfgs, ens.members):
        host.initialize_member(m, c)
        ppb = m.query().planes_per_block
        s = Lattice(precision=GABE_FP32, planes_per_block=ppb, **c.lattice_kwargs())
        host.initialize(s, c)
        singles.append(s)
    ens.step(nsteps)
    ens.sync()
    for s in singles:
        s.step(nsteps)
    for m, s in zip(ens.members, singles):
        fm, dm = m.download_state()
        fs, ds = s.download_state()
        assert np.array_equal(fm, fs) and np.array_equal(dm, ds)
        sm, ss = m.get_scalars(), s.get_scalars()
        for k in ("a", "adot", "edrho"):
            assert sm[k][0] == ss[k][0], k
        assert np.array_equal(m.history(), s.history())


def test_fp64_member_matches_oracle():
    cfg = configs.lambda_phi4_config(N=16, calc_gws=0)
    o = seeded_oracle(cfg, amp=2e-3)
    other = configs.lambda_phi4_config(N=24, calc_gws=0)
    ens = Ensemble([dict(precision=GABE_FP64, **cfg.lattice_kwargs()),
                    dict(precision=GABE_FP64, **other.lattice_kwargs())])
    m = ens.members[0]
    m.upload_state(o.field[0], o.dfield[0])
    sc = o.scalars()
    m.set_scalars(sc["a"], sc["adot"], sc["edkin"], sc["edpot"], sc["edgrad"], sc["edrho"])
    host.initialize_member(ens.members[1], other)
    o.step(9)
    ens.step(9)
    f, d = m.download_state()
    assert fluct_err(f, o.field[0]) < 1e-9 and fluct_err(d, o.dfield[0]) < 1e-9
    assert abs(m.get_scalars()["a"][0] - o.scalars()["a"][0]) < 1e-13


def test_autotuner_choice_is_used():
    ppb, ms = autotune_sweep(N=32, nflds=2, L=20.0, precision=GABE_FP32, lam=8.5, g2=855.0)
    assert ppb in (1, 2, 4, 8, 16, 32) and ms > 0
    lat = Lattice(N=32, nflds=2, L=20.0, precision=GABE_FP32, lam=8.5, g2=855.0)
    assert lat.query().planes_per_block == ppb


def test_autotuner_picks_transform_tiles_with_tensor_modes():
    """With tensor modes the autotuner also times the j-transform widths and
    the mode-kernel variants; a later context uses its choice."""
    from gabe_amd.lattice import mode_tiles
    kw = dict(N=32, nflds=2, L=20.0, precision=GABE_FP32, lam=8.5, g2=855.0, calc_gws=1)
    ppb, ms = autotune_sweep(**kw)
    assert ms > 0
    q = Lattice(**kw).query()
    assert q.planes_per_block == ppb and q.transform_tile in (16, 32)
    assert 1 <= q.mode_tile <= mode_tiles(32, GABE_FP32)
