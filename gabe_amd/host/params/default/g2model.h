/*
 * g2model.h -- the model plug-in of the kept host program: the potential of
 * the user's g2model.cpp written as the coefficients of the polynomial the
 * GPU kernels evaluate (gabe_model.h),
 *   V_pr = m2_phi phi^2/2 + lambda phi^4/4 + g2 phi^2 chi^2/2 + m2_chi chi^2/2
 * in program units (V / rescale_B^2 with fields in the units of
 * g2parameters.h).  Reads the constants of g2parameters.h.
 *
 * This is GABE's shipped model (g2model.cpp:59-87): V_pr = phi^2/2
 * + (gsq/mphi^2) phi^2 chi^2/2, i.e. m2_phi = 1 and g2 = gsq/mphi^2.
 */
const double model_m2_phi = 1.;
const double model_lambda = 0.;
const double model_g2 = gsq / mphi / mphi;
const double model_m2_chi = 0.;
