// Hyperelastic constitutive laws of the three AceGen tet10 elements (internal
// model index 0 = skr1, 1 = skr2, 2 = skr3):
//   skr1  W = mu/2 (I1-3) - mu lnJ + lam/4 (J^2 - 1 - 2 lnJ)
//   skr2  W = mu/2 (I1-3) - mu lnJ + lam/2 (lnJ)^2
//   skr3  W = mu/2 (I1 I3^(-1/3) - 3) + K/2 (J-1)^2,   K = E / (3 (1 - 2 nu))
// (feddlib/core/AceFemAssembly/AceInterface/NeoHookQuadraticTets{,2,3}.f).
// Shared by the assembly kernels, the surrogate's sample generation and its
// host-side training targets.
#pragma once
#include <hip/hip_runtime.h>

#include <cmath>

namespace fedd {

// Exact second Piola-Kirchhoff stress S (Voigt 6, order 11,22,33,12,13,23) and
// tangent CC = dS/dE (21 entries: upper triangle of the 6x6 Voigt matrix,
// row-major, engineering shear) from C = F^T F; CC21 may be null (stress only).
__host__ __device__ inline void exact_law(int model, double Emod, double nu, const double F[3][3], double S6[6],
                                   double CC21[21]) {
  double C[3][3];
#pragma unroll
  for (int A = 0; A < 3; ++A)
#pragma unroll
    for (int B = 0; B < 3; ++B) C[A][B] = F[0][A] * F[0][B] + F[1][A] * F[1][B] + F[2][A] * F[2][B];
  const double mu = Emod / (2. * (1. + nu));
  const double lam = 2. * mu * nu / (1. - 2. * nu);
  const double I3 = C[0][0] * (C[1][1] * C[2][2] - C[1][2] * C[2][1]) -
                    C[0][1] * (C[1][0] * C[2][2] - C[1][2] * C[2][0]) +
                    C[0][2] * (C[1][0] * C[2][1] - C[1][1] * C[2][0]);
  double Ci[3][3];
  Ci[0][0] = (C[1][1] * C[2][2] - C[1][2] * C[2][1]) / I3;
  Ci[0][1] = (C[0][2] * C[2][1] - C[0][1] * C[2][2]) / I3;
  Ci[0][2] = (C[0][1] * C[1][2] - C[0][2] * C[1][1]) / I3;
  Ci[1][0] = Ci[0][1];
  Ci[1][1] = (C[0][0] * C[2][2] - C[0][2] * C[2][0]) / I3;
  Ci[1][2] = (C[0][2] * C[1][0] - C[0][0] * C[1][2]) / I3;
  Ci[2][0] = Ci[0][2];
  Ci[2][1] = Ci[1][2];
  Ci[2][2] = (C[0][0] * C[1][1] - C[0][1] * C[1][0]) / I3;
  const double J = sqrt(I3);
  const double I1 = C[0][0] + C[1][1] + C[2][2];
  double a_id, a_ci, c_cc, c_ii, c_ic = 0.0;
  if (model == 0) {
    a_id = mu;
    a_ci = lam / 2. * I3 - mu - lam / 2.;
    c_cc = lam * I3;
    c_ii = 2. * mu + lam - lam * I3;
  } else if (model == 1) {
    const double lnJ = log(J);
    a_id = mu;
    a_ci = lam * lnJ - mu;
    c_cc = lam;
    c_ii = 2. * (mu - lam * lnJ);
  } else {
    const double K = Emod / (3. * (1. - 2. * nu));
    const double f = pow(I3, -1. / 3.);
    a_id = mu * f;
    a_ci = -mu / 3. * I1 * f + K * (J - 1.) * J;
    c_cc = 2. / 9. * mu * f * I1 + K * (2. * J - 1.) * J;
    c_ii = 2. / 3. * mu * f * I1 - 2. * K * (J - 1.) * J;
    c_ic = -2. / 3. * mu * f;
  }
  const int vA[6] = {0, 1, 2, 0, 0, 1}, vB[6] = {0, 1, 2, 1, 2, 2};
#pragma unroll
  for (int I = 0; I < 6; ++I) S6[I] = (vA[I] == vB[I] ? a_id : 0.0) + a_ci * Ci[vA[I]][vB[I]];
  if (!CC21) return;  // stress only (the surrogate supplies the tangent)
  int k = 0;
#pragma unroll
  for (int I = 0; I < 6; ++I)
#pragma unroll
    for (int Jv = I; Jv < 6; ++Jv) {
      const int A = vA[I], B = vB[I], Cc = vA[Jv], D = vB[Jv];
      CC21[k++] = c_cc * Ci[A][B] * Ci[Cc][D] + c_ii * 0.5 * (Ci[A][Cc] * Ci[B][D] + Ci[A][D] * Ci[B][Cc]) +
                  c_ic * ((A == B ? Ci[Cc][D] : 0.0) + (Cc == D ? Ci[A][B] : 0.0));
    }
}

}  // namespace fedd
