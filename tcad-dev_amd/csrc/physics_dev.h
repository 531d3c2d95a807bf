// Pointwise FP64 device functions of the drift-diffusion residual.
// Built with -ffp-contract=off so every expression rounds exactly as
// the reference's NumPy expression does (no fused multiply-add); the
// only source of difference is the transcendental library (expm1, log,
// exp, pow), which is within a few ulp of glibc's.
#pragma once

#include <hip/hip_runtime.h>

namespace tcad {

// Bernoulli B(x) = x / expm1(x), pytcad/pytcad/kernels.py:88-96.
__device__ __forceinline__ double bernoulli(double x) {
  x = fmin(fmax(x, -700.0), 700.0);
  if (fabs(x) < 1e-4) return 1.0 - x / 2.0 + x * x / 12.0;
  return x / expm1(x);
}

// dB/dx, pytcad/pytcad/kernels.py:99-109.
__device__ __forceinline__ double dbernoulli(double x) {
  x = fmin(fmax(x, -700.0), 700.0);
  if (fabs(x) < 1e-4) return -0.5 + x / 6.0 - x * x * x / 180.0;
  double B = x / expm1(x);
  double em = -expm1(-x);
  if (fabs(em) < 1e-300) em = 1e-300;
  return B * (1.0 / x - 1.0 / em);
}

// Bernoulli quartet for one edge difference d: B(d), B(-d), B'(d), B'(-d),
// each rounded as its own formula above.  The clip to [-700, 700] is
// symmetric, so clip(-d) == -clip(d) and two expm1 evaluations serve all
// four terms.
struct BQuad { double Bp, Bm, dBp, dBm; };

__device__ __forceinline__ BQuad bernoulli4(double d) {
  BQuad q;
  double x = fmin(fmax(d, -700.0), 700.0);
  double xm = -x;
  if (fabs(x) < 1e-4) {
    q.Bp = 1.0 - x / 2.0 + x * x / 12.0;
    q.dBp = -0.5 + x / 6.0 - x * x * x / 180.0;
    q.Bm = 1.0 - xm / 2.0 + xm * xm / 12.0;
    q.dBm = -0.5 + xm / 6.0 - xm * xm * xm / 180.0;
    return q;
  }
  double ep = expm1(x);
  double en = expm1(xm);
  q.Bp = x / ep;
  double em = -en;
  if (fabs(em) < 1e-300) em = 1e-300;
  q.dBp = q.Bp * (1.0 / x - 1.0 / em);
  q.Bm = xm / en;
  double emm = -ep;
  if (fabs(emm) < 1e-300) emm = 1e-300;
  q.dBm = q.Bm * (1.0 / xm - 1.0 / emm);
  return q;
}

// SRH (+ Auger) recombination and exact derivatives, Boltzmann branch of
// pytcad/pytcad/materials.py:474-518, physical units.
__device__ __forceinline__ void recombination(double n, double p, double nie, double tau_n,
                                              double tau_p, double Cn, double Cp, int auger,
                                              double &R, double &dRdn, double &dRdp) {
  double ni2 = nie * nie;
  double excess = n * p - ni2;
  double den = tau_p * (n + nie) + tau_n * (p + nie);
  double den2 = den * den;
  R = excess / den;
  dRdn = (p * den - excess * tau_p) / den2;
  dRdp = (n * den - excess * tau_n) / den2;
  if (auger) {
    double C = Cn * n + Cp * p;
    R = R + C * excess;
    dRdn = dRdn + Cn * excess + C * p;
    dRdp = dRdp + Cp * excess + C * n;
  }
}

// Lombardi CVT surface mobility, pytcad/pytcad/materials.py:389-429.
__device__ __forceinline__ double mobility_cvt(double E, double mu_ct, int hole, double T) {
  E = fmax(E, 1.0);
  double B = hole ? 5.0e7 : 2.5e8;
  double delta = hole ? 2.05e14 : 5.82e14;
  double mu_ph = B / (T * pow(E, 1.0 / 3.0));
  double mu_sr = delta / (E * E);
  double inv = 1.0 / fmax(mu_ct, 1.0) + 1.0 / fmax(mu_ph, 1.0) + 1.0 / fmax(mu_sr, 1.0);
  return 1.0 / inv;
}

}  // namespace tcad
