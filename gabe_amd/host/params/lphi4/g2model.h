/*
 * g2model.h -- the benchmark model's plug-in: V_pr = phi^4/(4 phi0^2)
 * + (g^2/lambda) phi^2 chi^2/(2 phi0^2), i.e. lambda = 1/phi0^2 and
 * g2 = (g^2/lambda)/phi0^2 in program units (B = sqrt(lambda) phi0).
 */
const double model_m2_phi = 0.;
const double model_lambda = lambda_pr;
const double model_g2 = g2_over_lambda / (phi0 * phi0);
const double model_m2_chi = 0.;
