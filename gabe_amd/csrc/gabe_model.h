/*
 * gabe_model.h -- pointwise model functions shared by the kernels and the host.
 *
 * The reference's model plug-in (g2model.cpp:59-108) indexes global arrays by
 * (s,i,j,k); kernels need the same mathematics pointwise.  Here the potential
 * and its derivatives take the field values at one site.  The reference-shaped
 * host wrappers potential(s,i,j,k), dVdf(s,fld,i,j,k) and effMass(s,fld) are
 * built on these in the host program.
 *
 * Model (program units): V = m2_phi phi^2/2 + lambda phi^4/4
 *                          + g2 phi^2 chi^2/2 + m2_chi chi^2/2.
 * To add a model, extend gabe_model and the three functions below.
 */
#ifndef GABE_MODEL_H
#define GABE_MODEL_H

#include <cmath>

#if defined(__HIPCC__)
#define GABE_HD __host__ __device__ __forceinline__
#else
#define GABE_HD inline
#endif

struct gabe_model {
    double m2_phi, lambda, g2, m2_chi;
};

/* g2model.cpp:66-69 (term order keeps the reference default bit-identical) */
GABE_HD double gabe_V(const gabe_model &m, double phi, double chi)
{
    return 0.5 * m.m2_phi * phi * phi + 0.5 * m.g2 * phi * phi * chi * chi
         + 0.25 * m.lambda * phi * phi * phi * phi + 0.5 * m.m2_chi * chi * chi;
}

/* g2model.cpp:72-87 */
GABE_HD double gabe_dVdf(const gabe_model &m, int fld, double phi, double chi)
{
    return fld == 0 ? m.m2_phi * phi + m.g2 * chi * chi * phi + m.lambda * phi * phi * phi
                    : m.g2 * phi * phi * chi + m.m2_chi * chi;
}

/* V and both first derivatives at one point, sharing the squares of the
 * fields: the mathematics of gabe_V and gabe_dVdf in fewer operations (the
 * sweeps evaluate them at every site, twice; the results differ from the
 * separate functions in the last bits only) */
GABE_HD void gabe_V_dV(const gabe_model &m, double phi, double chi, double &V, double &dphi, double &dchi)
{
    using std::fma;
    const double p2 = phi * phi, c2 = chi * chi;
    const double a = fma(0.5 * m.g2, c2, fma(0.25 * m.lambda, p2, 0.5 * m.m2_phi));
    V = fma(p2, a, 0.5 * m.m2_chi * c2);
    dphi = phi * fma(m.g2, c2, fma(m.lambda, p2, m.m2_phi));
    dchi = chi * fma(m.g2, p2, m.m2_chi);
}

/* d2V/dfld2, whose lattice mean is effMass (g2model.cpp:89-108) */
GABE_HD double gabe_d2Vdf2(const gabe_model &m, int fld, double phi, double chi)
{
    return fld == 0 ? m.m2_phi + m.g2 * chi * chi + 3. * m.lambda * phi * phi
                    : m.g2 * phi * phi + m.m2_chi;
}

#endif
