"""Model and lattice configurations used by the benchmark and the tests.

Two models are provided, both instances of the polynomial family in
csrc/gabe_model.h:

* reference_default_config(): the model shipped in g2parameters.h:43-62 and
  g2model.cpp:66-87, V = m^2 phi^2/2 + g^2 phi^2 chi^2/2 in units of B = m;
* lambda_phi4_config(): V = lambda phi^4/4 + g^2 phi^2 chi^2/2 with
  lambda = 9e-14, g^2/lambda = 100, phi0 = 0.342 M_pl, in units of
  B = sqrt(lambda) phi0 (so V_pr = phi^4/(4 phi0^2) + 100 phi^2 chi^2/(2 phi0^2)).
  The homogeneous velocity is the slow-roll value -V'/(3H) with
  H = sqrt(8 pi V/3).  L = 20 and dt = dx/20.
"""
import dataclasses
import math


@dataclasses.dataclass
class ModelConfig:
    N: int
    L: float
    dt: float
    nflds: int
    m2_phi: float
    lam: float
    g2: float
    m2_chi: float
    f0: tuple
    df0: tuple
    rescale_B: float
    randseed: int = 44463132           # g2parameters.h:56
    calc_gws: int = 1
    expansion_type: int = 1
    field_full_rand: int = 1           # g2parameters.h:72
    spec_cut_off: float = 0.0          # g2parameters.h:73 (0: not defined)
    spec_smooth: float = 0.5           # g2parameters.h:74

    @property
    def dx(self):
        return self.L / self.N

    def lattice_kwargs(self):
        return dict(N=self.N, nflds=self.nflds, L=self.L, dt=self.dt, calc_gws=self.calc_gws,
                    expansion_type=self.expansion_type, m2_phi=self.m2_phi, lam=self.lam,
                    g2=self.g2, m2_chi=self.m2_chi)


def reference_default_config(N=64, L=10.0, calc_gws=1):
    """g2parameters.h:43-62: mphi=1e-6, gsq=2.5e-7, phi0=0.193, dt=dx/80."""
    mphi, gsq = 1.e-6, 2.5e-7
    return ModelConfig(N=N, L=L, dt=L / N / 80.0, nflds=2, m2_phi=1.0, lam=0.0,
                       g2=gsq / mphi / mphi, m2_chi=0.0, f0=(0.193, 0.0), df0=(-0.142231, 0.0),
                       rescale_B=mphi, calc_gws=calc_gws)


def lambda_phi4_config(N=256, L=20.0, calc_gws=1, g2_over_lambda=100.0, phi0=0.342, lam=9e-14):
    B = math.sqrt(lam) * phi0
    lam_pr = 1.0 / (phi0 * phi0)
    g2_pr = g2_over_lambda / (phi0 * phi0)
    V = 0.25 * lam_pr * phi0 ** 4
    H = math.sqrt(8.0 * math.pi / 3.0 * V)
    dphi0 = -(lam_pr * phi0 ** 3) / (3.0 * H)
    dx = L / N
    return ModelConfig(N=N, L=L, dt=dx / 20.0, nflds=2, m2_phi=0.0, lam=lam_pr, g2=g2_pr,
                       m2_chi=0.0, f0=(phi0, 0.0), df0=(dphi0, 0.0), rescale_B=B, calc_gws=calc_gws)


def ensemble_configs(sizes=(32, 64, 96, 128), per_size=16, seed=20240):
    """The batched-ensemble workload: 16 each of 32^3, 64^3, 96^3 and 128^3
    lattices with L = 20, dt = dx/20, distinct IC seeds and g^2/lambda drawn
    log-uniformly from [1, 200]; no tensor modes (the batched path steps the
    fields and the expansion only)."""
    import random
    rng = random.Random(seed)
    cfgs = []
    for N in sizes:
        for _ in range(per_size):
            ratio = math.exp(rng.uniform(0.0, math.log(200.0)))
            c = lambda_phi4_config(N=N, calc_gws=0, g2_over_lambda=ratio)
            c.randseed = 1 + len(cfgs) * 7919
            cfgs.append(c)
    return cfgs
