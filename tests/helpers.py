"""Shared set-up for the parity tests: one initial state fed to both the
oracle and the device lattice."""
import numpy as np

from gabe_amd import configs
from gabe_amd.lattice import GABE_FP32, GABE_FP64, Lattice
from oracle import pyoracle as po


def oracle_for(cfg, nthreads=4):
    return po.Oracle(cfg.N, nflds=cfg.nflds, L=cfg.L, dt=cfg.dt, calc_gws=cfg.calc_gws,
                     expansion_type=cfg.expansion_type, m2_phi=cfg.m2_phi, lam=cfg.lam, g2=cfg.g2,
                     m2_chi=cfg.m2_chi, rescale_B=cfg.rescale_B, nthreads=nthreads)


def seeded_oracle(cfg, amp=None, nthreads=4):
    """Reference initialisation (g2main.cpp:109-119): homogeneous fields,
    initexpansion, vacuum fluctuations with rand() seeded by randseed.
    amp overrides rescale_B (larger fluctuations make small lattices
    exercise the nonlinear terms)."""
    if amp is not None:
        cfg.rescale_B = amp
    o = oracle_for(cfg, nthreads)
    o.initfields_homog(cfg.f0, cfg.df0)
    o.initexpansion()
    po.srand(cfg.randseed)
    if cfg.field_full_rand and cfg.spec_cut_off <= 0:
        o.initfields_fluct()
    else:   # g2model.cpp:132-145 with the options of g2init.cpp:111-165
        for fld in range(cfg.nflds):
            o.randinit(fld, o.effMass(0, fld), full_rand=cfg.field_full_rand, cut_off=cfg.spec_cut_off,
                       smooth=cfg.spec_smooth)
        o.calcEnergy(0)
    return o


def lattice_from_oracle(o, cfg, precision=GABE_FP64, **kw):
    lat = Lattice(precision=precision, **cfg.lattice_kwargs(), **kw)
    lat.upload_state(o.field[0], o.dfield[0])
    sc = o.scalars()
    lat.set_scalars(sc["a"], sc["adot"], sc["edkin"], sc["edpot"], sc["edgrad"], sc["edrho"], t=o.st.t)
    if cfg.calc_gws:
        lat.upload_gw(o.h[0], o.l[0])
    return lat


def fluct_err(a, b):
    """max |a - b| relative to the fluctuation amplitude of b (per field)."""
    out = []
    for f in range(b.shape[0]):
        scale = np.abs(b[f] - b[f].mean()).max()
        out.append(np.abs(a[f] - b[f]).max() / scale)
    return max(out)


def rel_err(a, b):
    return np.abs(a - b).max() / np.abs(b).max()


__all__ = ["configs", "oracle_for", "seeded_oracle", "lattice_from_oracle", "fluct_err", "rel_err",
           "GABE_FP32", "GABE_FP64", "Lattice", "po"]
