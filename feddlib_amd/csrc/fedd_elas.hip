// Hyperelastic P2 tet assembly for gfx950: quadrature-point kernel + row gather.
//
// Reference: AssembleFE_{LinElas,NonLinElas,NonLinElas2}_def.hpp call the AceGen
// tet-10 elements skr1/skr2/skr3 (NeoHookQuadraticTets{,2,3}.f) and take
// element matrix = -s, rhs = p (AssembleFE_NonLinElas_def.hpp:119, 171).  With the
// AceGen isoparametric determinant detJ_ace = -det J_F (J_F in FEDDLib's
// reference frame), that is
//     element = sum_q w_q det J_F(q) K_q ,   rhs = -sum_q w_q det J_F(q) f_int,q
// where K_q = dP/dF contracted with the material gradients and f_int,q = P grad N.
// Laws (mu = E/(2(1+nu)), lam = 2 mu nu/(1-2 nu), K = E/(3(1-2 nu))):
//   model 0 (skr1, 4-pt rule):  W = mu/2(I1-3) - mu lnJ + lam/4 (J^2-1-2 lnJ)
//   model 1 (skr2, 5-pt rule):  W = mu/2(I1-3) - mu lnJ + lam/2 (lnJ)^2
//   model 2 (skr3, 5-pt rule):  W = mu/2(I1 I3^(-1/3) - 3) + K/2 (J-1)^2
//
// The quadrature-point kernel writes, per (element, point), a 52-double record:
// F (9), S (6 Voigt), the material tangent CC = dS/dE as a symmetric 6x6 Voigt
// matrix (21), w_q det J_F (1), and the inverse isoparametric Jacobian in
// barycentric-row form (12: rows dxi_a/dX for a = 1..3 and minus their sum), so
// that a relabelled element's inverse Jacobian is a selection of three rows.  The tangent
// slot is where the learned surrogate writes instead of the exact law.  The row gather
// then builds each node's 3 rows of the tangent and its 3 residual entries in LDS and
// writes them once.
//
// FP32-local guard: the FP32 error of B^T CC B at a point grows with the conditioning of
// C = F^T F (measured on strained elements: about 10 u32 kappa of the element's largest
// entry, kappa = |C|_F |adj C|_F / det C >= cond(C)); points with kappa > 1e4 (or J = 0, or
// a non-finite F) flag their element (elem_flag, FP32-local launches only); the FP32 row
// gather leaves every pair of a flagged element
// out and lists its chunk, and an FP64 pass over the listed chunks accumulates exactly
// those pairs in LDS and adds them to the rows written.
#include "fedd_internal.h"
#include "fedd_law.h"
#include "fedd_perm.h"

#include <algorithm>

#ifndef FEDD_ELAS_WAVES
#define FEDD_ELAS_WAVES 2
#endif

// kappa bound above which a point's FP32 tangent products are not trusted
#define FEDD_FP32_KAPPA_MAX 1.0e4

namespace fedd {

__constant__ int8_t c_eperm2[10][10] = FEDD_PERM_P2_INIT;

// Voigt order (11, 22, 33, 12, 13, 23); strains engineering, stresses tensorial.
__constant__ int8_t c_vA[6] = {0, 1, 2, 0, 0, 1};
__constant__ int8_t c_vB[6] = {0, 1, 2, 1, 2, 2};

// Quadrature rules (SMSUtility.f icode 18 / 19), as barycentric orbits: the rule
// is symmetric, so point m (m < 4) is the one whose barycentric coordinate
// lambda_m is the distinguished value; the 5-point rule adds the centre (m = 4).
struct ElasRule {
  int n;
  double a, b, wa;  // orbit points: lambda_m = a, others b, weight wa
  double wc;        // centre weight (5-point rule)
};
__host__ __device__ inline ElasRule elas_rule(int model) {
  if (model == 0) return {4, 0.585410196624968454, 0.138196601125010515, 0.0416666666666666667, 0.0};
  return {5, 0.5, 1.0 / 6.0, 0.075, -0.133333333333333333};
}

// Reference point of orbit point m in FEDDLib coordinates (x, y, z) = (l1, l2, l3).
__device__ inline void qp_ref(const ElasRule& R, int m, double xi[3]) {
  if (m == 4) {
    xi[0] = xi[1] = xi[2] = 0.25;
    return;
  }
  double l[4] = {R.b, R.b, R.b, R.b};
  l[m] = R.a;
  xi[0] = l[1];
  xi[1] = l[2];
  xi[2] = l[3];
}

// P2 reference gradients (Helper.cpp:545-592).
__device__ inline void p2_dphi(const double p[3], double g[10][3]) {
  const double s = 4. * (p[0] + p[1] + p[2]);
  g[0][0] = g[0][1] = g[0][2] = -3. + s;
  g[1][0] = 4. * p[0] - 1.; g[1][1] = 0.; g[1][2] = 0.;
  g[2][0] = 0.; g[2][1] = 4. * p[1] - 1.; g[2][2] = 0.;
  g[3][0] = 0.; g[3][1] = 0.; g[3][2] = 4. * p[2] - 1.;
  g[4][0] = 4. - 8. * p[0] - 4. * p[1] - 4. * p[2]; g[4][1] = -4. * p[0]; g[4][2] = -4. * p[0];
  g[5][0] = 4. * p[1]; g[5][1] = 4. * p[0]; g[5][2] = 0.;
  g[6][0] = -4. * p[1]; g[6][1] = 4. - 4. * p[0] - 8. * p[1] - 4. * p[2]; g[6][2] = -4. * p[1];
  g[7][0] = -4. * p[2]; g[7][1] = -4. * p[2]; g[7][2] = 4. - 4. * p[0] - 4. * p[1] - 8. * p[2];
  g[8][0] = 4. * p[2]; g[8][1] = 0.; g[8][2] = 4. * p[0];
  g[9][0] = 0.; g[9][1] = 4. * p[2]; g[9][2] = 4. * p[1];
}

// Physical (material) gradients of the ten basis functions at reference point xi
// for element nodes x[10][3]; returns the signed det J_F.
__device__ inline double p2_grad(const double x[10][3], const double xi[3], double G[10][3],
                                 double Ji[3][3]) {
  double g[10][3];
  p2_dphi(xi, g);
  double J[3][3] = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}};  // J[i][a] = sum_k x_k,i dphi_k/dxi_a
#pragma unroll
  for (int k = 0; k < 10; ++k)
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
      for (int a = 0; a < 3; ++a) J[i][a] += x[k][i] * g[k][a];
  const double det = J[0][0] * J[1][1] * J[2][2] + J[0][1] * J[1][2] * J[2][0] + J[0][2] * J[1][0] * J[2][1] -
                     J[2][0] * J[1][1] * J[0][2] - J[2][1] * J[1][2] * J[0][0] - J[2][2] * J[1][0] * J[0][1];
  const double id = 1.0 / det;
  // Ji[a][i] = dxi_a/dX_i
  Ji[0][0] = (J[1][1] * J[2][2] - J[1][2] * J[2][1]) * id;
  Ji[0][1] = (J[0][2] * J[2][1] - J[0][1] * J[2][2]) * id;
  Ji[0][2] = (J[0][1] * J[1][2] - J[0][2] * J[1][1]) * id;
  Ji[1][0] = (J[1][2] * J[2][0] - J[1][0] * J[2][2]) * id;
  Ji[1][1] = (J[0][0] * J[2][2] - J[0][2] * J[2][0]) * id;
  Ji[1][2] = (J[0][2] * J[1][0] - J[0][0] * J[1][2]) * id;
  Ji[2][0] = (J[1][0] * J[2][1] - J[1][1] * J[2][0]) * id;
  Ji[2][1] = (J[0][1] * J[2][0] - J[0][0] * J[2][1]) * id;
  Ji[2][2] = (J[0][0] * J[1][1] - J[0][1] * J[1][0]) * id;
#pragma unroll
  for (int k = 0; k < 10; ++k)
#pragma unroll
    for (int A = 0; A < 3; ++A) G[k][A] = g[k][0] * Ji[0][A] + g[k][1] * Ji[1][A] + g[k][2] * Ji[2][A];
  return det;
}

struct ElasArgs {
  const double* xyz;     // SoA coordinates
  int64_t npts;
  const int32_t* conn16;  // AoS connectivity
  int64_t nelem;
  const double* u;        // displacement, 3 per node
  double* qp;             // per (element, point): 52-double record (see above)
  int nq;
  int model;
  double Emod, nu;
  const double* Evec;     // per-element Young's modulus (overrides Emod) or null
  int exact_tangent;      // 1: write CC from the law; 0: leave CC (surrogate wrote it)
  // plan
  const int64_t* nptr;
  const int32_t* chunk_node;
  const int32_t* chunk_pair;
  const int32_t* pair_elem;
  const uint4* pair_meta;
  double* vals;
  double* res;
  int with_res;
  // FP32-local: chunks whose FP32 tangent holds a non-finite entry are appended to
  // fix_list (count in fix_list[nchunks]); the FP64 fix-up launch (fixup = 1)
  // re-assembles exactly those chunks.  Chunks holding pairs of flagged elements are
  // appended to flag_list (count in flag_list[nchunks], pairs added in
  // flag_list[nchunks + 1]); the FP64 launch with fixup = 2 adds those pairs.
  int32_t* fix_list;
  int32_t* flag_list;
  int32_t* elem_flag;  // FP32-local: nonzero for elements with a flagged point (else null)
  int64_t nchunks;
  int fixup;
};

// kappa = |C|_F |adj C|_F / det C for C = F^T F (>= the condition number of C).
__device__ inline double elas_kappa(const double F[3][3]) {
  double C[3][3];
#pragma unroll
  for (int A = 0; A < 3; ++A)
#pragma unroll
    for (int B = 0; B < 3; ++B) C[A][B] = F[0][A] * F[0][B] + F[1][A] * F[1][B] + F[2][A] * F[2][B];
  const double a00 = C[1][1] * C[2][2] - C[1][2] * C[1][2], a01 = C[0][2] * C[1][2] - C[0][1] * C[2][2],
               a02 = C[0][1] * C[1][2] - C[0][2] * C[1][1], a11 = C[0][0] * C[2][2] - C[0][2] * C[0][2],
               a12 = C[0][1] * C[0][2] - C[0][0] * C[1][2], a22 = C[0][0] * C[1][1] - C[0][1] * C[0][1];
  const double det = C[0][0] * a00 + C[0][1] * a01 + C[0][2] * a02;
  const double nc = C[0][0] * C[0][0] + C[1][1] * C[1][1] + C[2][2] * C[2][2] +
                    2.0 * (C[0][1] * C[0][1] + C[0][2] * C[0][2] + C[1][2] * C[1][2]);
  const double na = a00 * a00 + a11 * a11 + a22 * a22 + 2.0 * (a01 * a01 + a02 * a02 + a12 * a12);
  return sqrt(nc * na) / det;
}

// One thread per (element, quadrature point), original labelling.
__global__ __launch_bounds__(64) void k_elas_qp(ElasArgs a) {
  // One wave per 64 consecutive records: each lane stages its record in LDS, in two
  // halves of 26 doubles, and the wave writes the 64 records out as contiguous lines.
  __shared__ double stage[26 * 64];
  const int64_t nrec = a.nelem * a.nq;
  const int64_t t0 = (int64_t)blockIdx.x * 64;
  const int lane = threadIdx.x;
  const int64_t t = t0 + lane;
  const bool live = t < nrec;
  const int64_t tt = live ? t : nrec - 1;  // idle lanes recompute the last record
  const int64_t e = tt / a.nq;
  const int m = (int)(tt - e * a.nq);
  const ElasRule R = elas_rule(a.model);
  const int32_t* cn = a.conn16 + 16 * e;
  double x[10][3], uu[10][3];
#pragma unroll
  for (int k = 0; k < 10; ++k) {
    const int32_t n = cn[k];
#pragma unroll
    for (int d = 0; d < 3; ++d) {
      x[k][d] = a.xyz[d * a.npts + n];
      uu[k][d] = a.u[3 * (int64_t)n + d];
    }
  }
  double xi[3];
  qp_ref(R, m, xi);
  double G[10][3], Ji[3][3];
  const double det = p2_grad(x, xi, G, Ji);
  double F[3][3];
#pragma unroll
  for (int i = 0; i < 3; ++i)
#pragma unroll
    for (int A = 0; A < 3; ++A) {
      double v = (i == A) ? 1.0 : 0.0;
#pragma unroll
      for (int k = 0; k < 10; ++k) v += uu[k][i] * G[k][A];
      F[i][A] = v;
    }
  double S6[6], CC[21];
#pragma unroll
  for (int k = 0; k < 21; ++k) CC[k] = 0.0;  // the surrogate writes the tangent slot afterwards
  exact_law(a.model, a.Evec ? a.Evec[e] : a.Emod, a.nu, F, S6, a.exact_tangent ? CC : nullptr);
  double rec[52];
#pragma unroll
  for (int i = 0; i < 3; ++i)
#pragma unroll
    for (int A = 0; A < 3; ++A) rec[3 * i + A] = F[i][A];
#pragma unroll
  for (int I = 0; I < 6; ++I) rec[9 + I] = S6[I];
#pragma unroll
  for (int k = 0; k < 21; ++k) rec[15 + k] = CC[k];
  rec[36] = (m == 4 ? R.wc : R.wa) * det;
#pragma unroll
  for (int A = 0; A < 3; ++A) {
    rec[37 + A] = -(Ji[0][A] + Ji[1][A] + Ji[2][A]);
    rec[40 + A] = Ji[0][A];
    rec[43 + A] = Ji[1][A];
    rec[46 + A] = Ji[2][A];
  }
  rec[49] = rec[50] = rec[51] = 0.0;
  if (a.elem_flag && live && !(elas_kappa(F) <= FEDD_FP32_KAPPA_MAX))  // also J = 0, non-finite F
    atomicOr(a.elem_flag + e, 1);
  const int64_t nout = (nrec - t0 < 64 ? nrec - t0 : 64) * 52;  // doubles this wave writes
  double* out = a.qp + 52 * t0;
#pragma unroll
  for (int h = 0; h < 2; ++h) {
#pragma unroll
    for (int k = 0; k < 26; ++k) stage[lane * 26 + k] = rec[26 * h + k];
    __syncthreads();
    // record r's half h covers out[52 r + 26 h .. + 26)
    for (int j = lane; j < 26 * 64; j += 64) {
      const int r = j / 26, k = j - 26 * r;
      const int64_t idx = 52 * (int64_t)r + 26 * h + k;
      if (idx < nout) out[idx] = stage[j];
    }
    __syncthreads();
  }
  (void)live;
}

// Canonical-frame P2 reference gradients at the canonical points (compile-time).
template <int M, int MODEL>
struct DPhiQ {
  static __device__ __forceinline__ void get(double g[10][3]) {
    const ElasRule R = elas_rule(MODEL);
    double xi[3];
    if (M == 4) {
      xi[0] = xi[1] = xi[2] = 0.25;
    } else {
      double l[4] = {R.b, R.b, R.b, R.b};
      l[M] = R.a;
      xi[0] = l[1];
      xi[1] = l[2];
      xi[2] = l[3];
    }
    p2_dphi(xi, g);
  }
};

// Adds the contribution of canonical point M to node r's rows.
template <typename T, int M, int MODEL>
__device__ __forceinline__ void elas_point(const ElasArgs& a, int32_t e, const int8_t* pm, int r,
                                           const uint8_t* mm, double* __restrict__ row, int deg,
                                           double fres[3], T& chk) {
  const int morig = M < 4 ? pm[M] : 4;
  const double* rec = a.qp + 52 * ((int64_t)e * a.nq + morig);
  double dphi[10][3];
  DPhiQ<M, MODEL>::get(dphi);
  double Jd[3][3];  // canonical inverse Jacobian: rows pm[1..3] of the barycentric form
#pragma unroll
  for (int ax = 0; ax < 3; ++ax) {
    const double* jr = rec + 37 + 3 * pm[ax + 1];
#pragma unroll
    for (int A = 0; A < 3; ++A) Jd[ax][A] = jr[A];
  }
  // Physical basis gradients from the FP64 inverse Jacobian, rounded once to T: on
  // distorted (curved) elements an FP32 inverse Jacobian costs the FP32-local tangent
  // its 1e-5 bound.  Each is formed where it is used (the row's here, the columns' in
  // the column loop), so that the ten gradients are not all live at once.
  auto grad = [&](int c, T G[3]) {
#pragma unroll
    for (int A = 0; A < 3; ++A) {
      double v = 0.0;
#pragma unroll
      for (int ax = 0; ax < 3; ++ax)
        if (dphi[c][ax] != 0.0) v += dphi[c][ax] * Jd[ax][A];
      G[A] = (T)v;
    }
  };
  T g[3], g4[3];
  grad(0, g);
  grad(4, g4);
#pragma unroll
  for (int A = 0; A < 3; ++A) g[A] = r == 0 ? g[A] : g4[A];
  const double wdet = rec[36];
  double Fd[3][3];
#pragma unroll
  for (int k = 0; k < 9; ++k) Fd[k / 3][k % 3] = rec[k];
  double Sd[6];
#pragma unroll
  for (int I = 0; I < 6; ++I) Sd[I] = rec[9 + I];
  {  // residual in FP64: f_a += w det sum_A (F S)_aA g_A, with the row gradient in FP64
    double gd[3];
#pragma unroll
    for (int A = 0; A < 3; ++A) {
      const double v0 = dphi[0][0] * Jd[0][A] + dphi[0][1] * Jd[1][A] + dphi[0][2] * Jd[2][A];
      const double v4 = dphi[4][0] * Jd[0][A] + dphi[4][1] * Jd[1][A] + dphi[4][2] * Jd[2][A];
      gd[A] = r == 0 ? v0 : v4;
    }
    const double Sg0 = Sd[0] * gd[0] + Sd[3] * gd[1] + Sd[4] * gd[2];
    const double Sg1 = Sd[3] * gd[0] + Sd[1] * gd[1] + Sd[5] * gd[2];
    const double Sg2 = Sd[4] * gd[0] + Sd[5] * gd[1] + Sd[2] * gd[2];
#pragma unroll
    for (int i = 0; i < 3; ++i) fres[i] += wdet * (Fd[i][0] * Sg0 + Fd[i][1] * Sg1 + Fd[i][2] * Sg2);
  }
  T F[3][3];
#pragma unroll
  for (int i = 0; i < 3; ++i)
#pragma unroll
    for (int j = 0; j < 3; ++j) F[i][j] = (T)Fd[i][j];
  const T w = (T)wdet;
  T Bg[3][6];  // d E_I / d u_(r,a), engineering shear
#pragma unroll
  for (int i = 0; i < 3; ++i) {
    Bg[i][0] = F[i][0] * g[0];
    Bg[i][1] = F[i][1] * g[1];
    Bg[i][2] = F[i][2] * g[2];
    Bg[i][3] = F[i][0] * g[1] + F[i][1] * g[0];
    Bg[i][4] = F[i][0] * g[2] + F[i][2] * g[0];
    Bg[i][5] = F[i][1] * g[2] + F[i][2] * g[1];
  }
  T BC[3][6];  // w * Bg CC
  {
    T CC[21];
#pragma unroll
    for (int k = 0; k < 21; ++k) CC[k] = (T)rec[15 + k];
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
      for (int J = 0; J < 6; ++J) {
        T s = 0;
#pragma unroll
        for (int I = 0; I < 6; ++I) {
          const int lo = I < J ? I : J, hi = I < J ? J : I;
          s += Bg[i][I] * CC[lo * 6 - lo * (lo - 1) / 2 + (hi - lo)];
        }
        BC[i][J] = w * s;
      }
  }
  T Sg[3];
  Sg[0] = w * ((T)Sd[0] * g[0] + (T)Sd[3] * g[1] + (T)Sd[4] * g[2]);
  Sg[1] = w * ((T)Sd[3] * g[0] + (T)Sd[1] * g[1] + (T)Sd[5] * g[2]);
  Sg[2] = w * ((T)Sd[4] * g[0] + (T)Sd[5] * g[1] + (T)Sd[2] * g[2]);
#pragma unroll
  for (int c = 0; c < 10; ++c) {
    T Gc[3];
    grad(c, Gc);
    const T geo = Sg[0] * Gc[0] + Sg[1] * Gc[1] + Sg[2] * Gc[2];
    double* rc = row + 3 * mm[2 + c];
#pragma unroll
    for (int j = 0; j < 3; ++j) {
      T Bh[6];
      Bh[0] = F[j][0] * Gc[0];
      Bh[1] = F[j][1] * Gc[1];
      Bh[2] = F[j][2] * Gc[2];
      Bh[3] = F[j][0] * Gc[1] + F[j][1] * Gc[0];
      Bh[4] = F[j][0] * Gc[2] + F[j][2] * Gc[0];
      Bh[5] = F[j][1] * Gc[2] + F[j][2] * Gc[1];
#pragma unroll
      for (int i = 0; i < 3; ++i) {
        T s = (i == j) ? geo : (T)0;
#pragma unroll
        for (int J = 0; J < 6; ++J) s += BC[i][J] * Bh[J];
        chk += s;  // a non-finite entry makes the running sum non-finite
        atomicAdd(rc + i * 3 * deg + j, (double)s);
      }
    }
  }
}

template <typename T, int MODEL>
__device__ __forceinline__ void elas_pair(const ElasArgs& a, int32_t e, const uint8_t* mm,
                                          double* __restrict__ row, int deg, double fres[3], T& chk) {
  const int li = mm[0];
  const int8_t* pm = c_eperm2[li];
  const int r = li < 4 ? 0 : 4;
  elas_point<T, 0, MODEL>(a, e, pm, r, mm, row, deg, fres, chk);
  elas_point<T, 1, MODEL>(a, e, pm, r, mm, row, deg, fres, chk);
  elas_point<T, 2, MODEL>(a, e, pm, r, mm, row, deg, fres, chk);
  elas_point<T, 3, MODEL>(a, e, pm, r, mm, row, deg, fres, chk);
  if (MODEL != 0) elas_point<T, 4, MODEL>(a, e, pm, r, mm, row, deg, fres, chk);
}

// Row gather: pair (node i, element e, local li) adds node i's 3 tangent rows and
// 3 residual entries of element e into LDS; the chunk is then written once.
// FP32-local follow-up launches (FP64): fixup = 1 re-assembles the chunks listed with a
// non-finite FP32 entry, fixup = 2 adds the pairs of flagged elements to the chunks listed
// with such pairs; the workgroups walk the list (none listed: all return at once).
template <typename T>
__device__ __forceinline__ void elas_chunk(const ElasArgs& a, int64_t ch, double* lds, int2* tab, int* queue);

template <typename T>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(FEDD_ELAS_WAVES))) void k_elas_rows(ElasArgs a) {
  extern __shared__ double lds[];
  __shared__ int2 tab[256];
  __shared__ int queue[513];  // fixup = 2: flagged pairs of the chunk (+ count)
  if (a.fixup) {
    const int32_t* list = a.fixup == 1 ? a.fix_list : a.flag_list;
    const int32_t n = list[a.nchunks];
    for (int32_t k = blockIdx.x; k < n; k += gridDim.x) {
      elas_chunk<T>(a, list[k], lds, tab, queue);
      __syncthreads();  // LDS and tab are reused by the next listed chunk
    }
    return;
  }
  elas_chunk<T>(a, xcd_chunk(blockIdx.x, gridDim.x), lds, tab, queue);
}

// Pair q of the chunk: node i's 3 tangent rows (and 3 residual entries) of element e.
template <typename T>
__device__ __forceinline__ void elas_pair_at(const ElasArgs& a, int q, const int2* tab, double* lds, int aLen,
                                             int& bad) {
  const int32_t e = a.pair_elem[q];
  const uint4 mv = a.pair_meta[q];
  const uint8_t* mm = reinterpret_cast<const uint8_t*>(&mv);
  const int2 nd = tab[mm[1]];
  double fres[3] = {0.0, 0.0, 0.0};
  T chk = 0;
  if (a.model == 0)
    elas_pair<T, 0>(a, e, mm, lds + nd.x, nd.y, fres, chk);
  else
    elas_pair<T, 1>(a, e, mm, lds + nd.x, nd.y, fres, chk);
  if (sizeof(T) == 4 && !isfinite((float)chk)) bad = 1;
  if (a.with_res) {
    double* rr = lds + aLen + 3 * mm[1];
#pragma unroll
    for (int i = 0; i < 3; ++i) atomicAdd(rr + i, -fres[i]);
  }
}

template <typename T>
__device__ __forceinline__ void elas_chunk(const ElasArgs& a, int64_t ch, double* lds, int2* tab, int* queue) {
  const int n0 = a.chunk_node[ch], n1 = a.chunk_node[ch + 1];
  const int q0 = a.chunk_pair[ch], q1 = a.chunk_pair[ch + 1];
  const int64_t na0 = a.nptr[n0], na1 = a.nptr[n1];
  const int aLen = (int)(9 * (na1 - na0));
  const int rLen = a.with_res ? 3 * (n1 - n0) : 0;
  for (int t = threadIdx.x; t < n1 - n0; t += blockDim.x) {
    const int64_t ai = a.nptr[n0 + t];
    tab[t] = make_int2((int)(9 * (ai - na0)), (int)(a.nptr[n0 + t + 1] - ai));
  }
  for (int k = threadIdx.x; k < aLen + rLen; k += blockDim.x) lds[k] = 0.0;
  if (threadIdx.x == 0) queue[512] = 0;
  __syncthreads();
  int bad = 0, flagged = 0;
  if (a.fixup == 2) {
    // FP64 pairs of flagged elements only, gathered into a queue so that every lane of a
    // wave works on one (a flagged pair is rare among its chunk's pairs)
    int done = 0;
    for (int base = q0; base < q1; base += blockDim.x) {
      const int q = base + threadIdx.x;
      if (q < q1 && a.elem_flag[a.pair_elem[q]] != 0) queue[atomicAdd(queue + 512, 1)] = q;
      __syncthreads();
      const int n = queue[512];  // < 512: drained whenever it reaches 256
      if (n >= 256 || base + (int)blockDim.x >= q1) {
        for (int k = threadIdx.x; k < n; k += blockDim.x) elas_pair_at<T>(a, queue[k], tab, lds, aLen, bad);
        done += n;
        __syncthreads();
        if (threadIdx.x == 0) queue[512] = 0;
      }
      __syncthreads();
    }
    if (threadIdx.x == 0 && done) atomicAdd(a.flag_list + a.nchunks + 1, done);
  } else {
    for (int q = q0 + threadIdx.x; q < q1; q += blockDim.x) {
      if (sizeof(T) == 4 && a.elem_flag[a.pair_elem[q]] != 0) {  // flagged: left to the FP64 pass
        flagged = 1;
        continue;
      }
      elas_pair_at<T>(a, q, tab, lds, aLen, bad);
    }
  }
  if (sizeof(T) == 4) {
    if (__syncthreads_or(bad) && threadIdx.x == 0) a.fix_list[atomicAdd(a.fix_list + a.nchunks, 1)] = (int32_t)ch;
    if (__syncthreads_or(flagged) && threadIdx.x == 0)
      a.flag_list[atomicAdd(a.flag_list + a.nchunks, 1)] = (int32_t)ch;
  } else {
    __syncthreads();
  }
  double* out = a.vals + 9 * na0;
  if (a.fixup == 2) {  // add the flagged pairs to the rows the FP32 launch wrote
    for (int k = threadIdx.x; k < aLen; k += blockDim.x)
      if (lds[k] != 0.0) out[k] += lds[k];
    if (a.with_res)
      for (int k = threadIdx.x; k < rLen; k += blockDim.x)
        if (lds[aLen + k] != 0.0) a.res[3 * (int64_t)n0 + k] += lds[aLen + k];
    return;
  }
  for (int k = threadIdx.x; k < aLen; k += blockDim.x) __builtin_nontemporal_store(lds[k], out + k);
  if (a.with_res)
    for (int k = threadIdx.x; k < rLen; k += blockDim.x) a.res[3 * (int64_t)n0 + k] = lds[aLen + k];
}

__global__ void k_fill_count(int32_t* p) {
  if (threadIdx.x == 0) *p = 0;
}

__global__ void k_zero_i32(int64_t n, int32_t* p) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = 0;
}

int launch_elasticity(Ctx* c, int model, double Emod, double nu, int precision, int with_res, int tangent,
                      const double* Evec) {
  const int exact_tangent = tangent == 2 ? 0 : 1;
  const Mesh& m = c->mesh[0];
  ElasArgs a{};
  a.xyz = m.d_xyz.as<double>();
  a.npts = m.npts;
  a.conn16 = m.d_conn16.as<int32_t>();
  a.nelem = m.nelem;
  a.u = c->vec[FEDD_VEC_U].as<double>();
  a.nq = elas_rule(model).n;
  a.model = model;
  a.Emod = Emod;
  a.nu = nu;
  a.Evec = Evec;
  a.exact_tangent = exact_tangent;
  const size_t qbytes = (size_t)m.nelem * 5 * 52 * 8;
  if (c->d_qp.bytes < qbytes) {
    int rc = alloc(c, c->d_qp, qbytes);
    if (rc) return rc;
  }
  a.qp = c->d_qp.as<double>();
  c->qp_nq = a.nq;
  const GatherPlan& pl = c->vplan;
  a.nptr = c->d_nptr.as<int64_t>();
  a.chunk_node = pl.d_chunk_node.as<int32_t>();
  a.chunk_pair = pl.d_chunk_pair.as<int32_t>();
  a.pair_elem = pl.d_pair_elem.as<int32_t>();
  a.pair_meta = pl.d_pair_meta.as<uint4>();
  a.vals = c->blk[0].d_vals.as<double>();
  a.res = c->vec[FEDD_VEC_RES_U].as<double>();
  a.with_res = with_res;
  const int64_t nqp = m.nelem * a.nq;
  if (precision == FEDD_FP32_LOCAL) {
    if (c->d_elas_fix.bytes < (size_t)(pl.nchunks + 1) * 4) {
      int rc = alloc(c, c->d_elas_fix, (size_t)(pl.nchunks + 1) * 4);
      if (rc) return rc;
    }
    if (c->d_elas_flag.bytes < (size_t)(pl.nchunks + 2 + m.nelem) * 4) {
      int rc = alloc(c, c->d_elas_flag, (size_t)(pl.nchunks + 2 + m.nelem) * 4);
      if (rc) return rc;
    }
    a.fix_list = c->d_elas_fix.as<int32_t>();
    a.flag_list = c->d_elas_flag.as<int32_t>();
    a.elem_flag = a.flag_list + pl.nchunks + 2;
    // kernels, not memsets (capturable)
    k_fill_count<<<1, 64, 0, c->stream>>>(a.fix_list + pl.nchunks);
    k_zero_i32<<<(unsigned)((m.nelem + 2 + 255) / 256), 256, 0, c->stream>>>(m.nelem + 2, a.flag_list + pl.nchunks);
  }
  ktimer_begin(c);
  k_elas_qp<<<(unsigned)((nqp + 63) / 64), 64, 0, c->stream>>>(a);
  FEDD_HIP(c, hipGetLastError());
  ktimer_end(c, "elas_qp");
  if (!exact_tangent) {
    int rc = launch_mlp_tangent(c, a.nq);
    if (rc) return rc;
  }
  const size_t lds = (size_t)(pl.lds_doubles + 3 * 256) * 8;
  a.nchunks = pl.nchunks;
  ktimer_begin(c);
  if (precision == FEDD_FP32_LOCAL)
    k_elas_rows<float><<<(unsigned)pl.nchunks, 256, lds, c->stream>>>(a);
  else
    k_elas_rows<double><<<(unsigned)pl.nchunks, 256, lds, c->stream>>>(a);
  FEDD_HIP(c, hipGetLastError());
  ktimer_end(c, "elas_rows");
  if (precision == FEDD_FP32_LOCAL) {
    const unsigned gfix = (unsigned)std::min<int64_t>(pl.nchunks, c->num_cu);
    // FP64 pairs of flagged elements, added to the rows just written: one workgroup per
    // chunk, those beyond the listed count return at once.  Its cost is about that of
    // one FP64 pass over the listed chunks (each chunk's few flagged pairs take one wave
    // through the FP64 pair code): at 32^3 x 6 with grad u up to 3, where 10% of the
    // pairs are flagged, FP32-local takes 2.0 ms and FP64 1.33 ms; with no flags it costs
    // nothing
    a.fixup = 2;
    k_elas_rows<double><<<(unsigned)pl.nchunks, 256, lds, c->stream>>>(a);
    FEDD_HIP(c, hipGetLastError());
    // FP64 re-assembly of the chunks whose FP32 tangent left FP32's range (usually none);
    // it rewrites whole rows, the pairs above included
    a.fixup = 1;
    k_elas_rows<double><<<gfix, 256, lds, c->stream>>>(a);
    FEDD_HIP(c, hipGetLastError());
  }
  c->elas_fix_valid = precision == FEDD_FP32_LOCAL;
  return FEDD_OK;
}

}  // namespace fedd
