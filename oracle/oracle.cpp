// oracle.cpp -- CPU restatement of FEDDLib's element-assembly path.
//
// TEST INFRASTRUCTURE ONLY (see oracle.h): the checker for the GPU library and
// the CPU baseline timed beside it.  Citations are FEDDLib file:line.
#include "oracle.h"

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <limits>
#include <utility>
#include <thread>
#include <vector>

namespace {

typedef std::vector<double> vec_d;
typedef std::vector<vec_d> vec2_d;
typedef std::vector<vec2_d> vec3_d;

// ---------------------------------------------------------------------------
// Tpetra-style CRS: rows receive (col,val) pairs in insertion order
// (Matrix::insertGlobalValues, Matrix_def.hpp:88-92); fillComplete sorts the
// columns of each row and sums duplicates (Matrix_def.hpp:192-199).  Duplicates
// are summed in insertion order (stable sort), explicit zeros are kept.
// ---------------------------------------------------------------------------
// Checker switch (orc_set_element_mode): 1 rounds every element value to FP32 before it
// is summed in FP64 (the error an FP32 element matrix causes by itself); 2 replaces every
// value of an element block (A, B^T, B of Navier-Stokes; the whole matrix otherwise) by the
// block's largest |value| (elem_scale), so each entry sums sum_e max |K_e| over its
// elements, the scale of the rounding error of element entries computed in FP32.
static int g_elem_mode = 0;

// Mode 2: every entry of the nr x nc block at (r0, c0) of the row-major element matrix
// (leading dimension ld) becomes the block's largest |value|.
void elem_scale(double* el, int ld, int r0, int nr, int c0, int nc) {
  if (g_elem_mode != 2) return;
  double m = 0.0;
  for (int i = 0; i < nr; ++i)
    for (int j = 0; j < nc; ++j) m = std::max(m, std::fabs(el[(r0 + i) * ld + c0 + j]));
  for (int i = 0; i < nr; ++i)
    for (int j = 0; j < nc; ++j) el[(r0 + i) * ld + c0 + j] = m;
}

struct Crs {
  std::vector<std::vector<std::pair<int32_t, double>>> rows;
  bool filled = false;
  explicit Crs(int64_t n) : rows(n) {}
  void insert(int64_t row, const int32_t* cols, const double* vals, int n) {
    auto& r = rows[row];
    for (int k = 0; k < n; ++k) r.emplace_back(cols[k], g_elem_mode == 1 ? (double)(float)vals[k] : vals[k]);
  }
  void fill_complete() {
    for (auto& r : rows) {
      std::stable_sort(r.begin(), r.end(),
                       [](const std::pair<int32_t, double>& a, const std::pair<int32_t, double>& b) {
                         return a.first < b.first;
                       });
      size_t w = 0;
      for (size_t k = 0; k < r.size(); ++k) {
        if (w > 0 && r[w - 1].first == r[k].first)
          r[w - 1].second += r[k].second;
        else
          r[w++] = r[k];
      }
      r.resize(w);
    }
    filled = true;
  }
  int64_t nnz() const {
    int64_t n = 0;
    for (auto& r : rows) n += (int64_t)r.size();
    return n;
  }
};

// ---------------------------------------------------------------------------
// Helper: quadrature (Helper.cpp:1300-1517) and basis (Helper.cpp:442-743)
// ---------------------------------------------------------------------------
int quadrature(int Degree, vec2_d& QuadPts, vec_d& QuadW) {
  // degree remap for tets (Helper.cpp:1302-1305)
  if (Degree == 2) Degree = 3;
  if (Degree == 4) Degree = 5;
  QuadPts.clear();
  QuadW.clear();
  switch (Degree) {
    case 1:  // Helper.cpp:1310-1317
      QuadPts.assign(1, vec_d{0.25, 0.25, 0.25});
      QuadW.assign(1, 1 / 6.);
      break;
    case 3: {  // Helper.cpp:1319-1350
      double a = .25, b = 1. / 6., c = .5;
      QuadPts = {{a, a, a}, {b, b, b}, {b, b, c}, {b, c, b}, {c, b, b}};
      QuadW = {-2. / 15., 3. / 40., 3. / 40., 3. / 40., 3. / 40.};
      break;
    }
    case 5: {  // Helper.cpp:1423-1517
      double a = 0.25;
      double b1 = (7. + std::sqrt(15.)) / 34.;
      double b2 = (7. - std::sqrt(15.)) / 34.;
      double c1 = (13. - 3. * std::sqrt(15.)) / 34.;
      double c2 = (13. + 3. * std::sqrt(15.)) / 34.;
      double d = (5. - std::sqrt(15.)) / 20.;
      double e = (5. + std::sqrt(15.)) / 20.;
      QuadPts = {{a, a, a},    {b1, b1, b1}, {b1, b1, c1}, {b1, c1, b1}, {c1, b1, b1},
                 {b2, b2, b2}, {b2, b2, c2}, {b2, c2, b2}, {c2, b2, b2}, {d, d, e},
                 {d, e, d},    {e, d, d},    {d, e, e},    {e, d, e},    {e, e, d}};
      double P1 = (2665. - 14. * std::sqrt(15.)) / 226800.;
      double P2 = (2665. + 14. * std::sqrt(15.)) / 226800.;
      double b = 5. / 567.;
      QuadW = {8. / 405., P1, P1, P1, P1, P2, P2, P2, P2, b, b, b, b, b, b};
      break;
    }
    default:
      return -1;
  }
  return (int)QuadW.size();
}

double phi(int fe, int i, const double* p) {
  if (fe == 1) {  // Helper.cpp:660-675
    switch (i) {
      case 0: return 1. - p[0] - p[1] - p[2];
      case 1: return p[0];
      case 2: return p[1];
      default: return p[2];
    }
  }
  switch (i) {  // P2, Helper.cpp:677-710
    case 0: return (1. - p[0] - p[1] - p[2]) * (1 - 2 * p[0] - 2 * p[1] - 2 * p[2]);
    case 1: return p[0] * (2 * p[0] - 1);
    case 2: return p[1] * (2 * p[1] - 1);
    case 3: return p[2] * (2 * p[2] - 1);
    case 4: return 4 * p[0] * (1 - p[0] - p[1] - p[2]);
    case 5: return 4 * p[0] * p[1];
    case 6: return 4 * p[1] * (1 - p[0] - p[1] - p[2]);
    case 7: return 4 * p[2] * (1 - p[0] - p[1] - p[2]);
    case 8: return 4 * p[0] * p[2];
    default: return 4 * p[1] * p[2];
  }
}

void grad_phi(int fe, int i, const double* p, double* v) {
  if (fe == 1) {  // Helper.cpp:522-543
    static const double g[4][3] = {{-1., -1., -1.}, {1., 0., 0.}, {0., 1., 0.}, {0., 0., 1.}};
    v[0] = g[i][0]; v[1] = g[i][1]; v[2] = g[i][2];
    return;
  }
  switch (i) {  // P2, Helper.cpp:545-592
    case 0: v[0] = v[1] = v[2] = -3. + 4. * p[0] + 4. * p[1] + 4. * p[2]; break;
    case 1: v[0] = 4. * p[0] - 1; v[1] = 0.; v[2] = 0.; break;
    case 2: v[0] = 0.; v[1] = 4. * p[1] - 1; v[2] = 0.; break;
    case 3: v[0] = 0.; v[1] = 0.; v[2] = 4. * p[2] - 1; break;
    case 4: v[0] = 4. - 8. * p[0] - 4. * p[1] - 4. * p[2]; v[1] = -4. * p[0]; v[2] = -4. * p[0]; break;
    case 5: v[0] = 4. * p[1]; v[1] = 4. * p[0]; v[2] = 0.; break;
    case 6: v[0] = -4. * p[1]; v[1] = 4. - 4. * p[0] - 8. * p[1] - 4. * p[2]; v[2] = -4. * p[1]; break;
    case 7: v[0] = -4. * p[2]; v[1] = -4. * p[2]; v[2] = 4. - 4. * p[0] - 4. * p[1] - 8. * p[2]; break;
    case 8: v[0] = 4. * p[2]; v[1] = 0.; v[2] = 4. * p[0]; break;
    default: v[0] = 0.; v[1] = 4. * p[2]; v[2] = 4. * p[1]; break;
  }
}

int nloc(int fe) { return fe == 1 ? 4 : 10; }

// Helper::getDPhi (Helper.cpp:273-356): DPhi[w][i][d] at the quadrature points
void getDPhi(vec3_d& DPhi, vec_d& w, int fe, int deg) {
  vec2_d qp;
  quadrature(deg, qp, w);
  int n = nloc(fe);
  DPhi.assign(w.size(), vec2_d(n, vec_d(3, 0.)));
  for (size_t k = 0; k < w.size(); ++k)
    for (int i = 0; i < n; ++i) grad_phi(fe, i, qp[k].data(), DPhi[k][i].data());
}

// Helper::getPhi (Helper.cpp:846-952)
void getPhi(vec2_d& Phi, vec_d& w, int fe, int deg) {
  vec2_d qp;
  quadrature(deg, qp, w);
  int n = nloc(fe);
  Phi.assign(w.size(), vec_d(n, 0.));
  for (size_t k = 0; k < w.size(); ++k)
    for (int i = 0; i < n; ++i) Phi[k][i] = phi(fe, i, qp[k].data());
}

// Helper::determineDegree (Helper.cpp:7-23): (dim, FEType, UN degFunc)
int degree_func(int fe, int degfunc) {
  int deg = fe + degfunc;  // P1 -> 1, P2 -> 2
  if (deg == 0) deg = 1;
  return deg;
}
// Helper.cpp:25-52: (dim, FEType, int type); type 0 = Std, 1 = Grad
int degree_type(int fe, int type) {
  int deg = (fe == 1) ? (type == 0 ? 1 : 0) : (type == 0 ? 2 : 1);
  if (deg == 0) deg = 1;
  return deg;
}
// Helper.cpp:54-136
int degree_pair(int fe1, int fe2, int t1, int t2, int extra) {
  int d1 = (fe1 == 1) ? (t1 == 0 ? 1 : 0) : (t1 == 0 ? 2 : 1);
  int d2 = (fe2 == 1) ? (t2 == 0 ? 1 : 0) : (t2 == 0 ? 2 : 1);
  int deg = d1 + d2 + extra;
  if (deg == 0) deg = 1;
  return deg;
}

// ---------------------------------------------------------------------------
// SmallMatrix (SmallMatrix.hpp:317-368) and Helper::buildTransformation /
// applyBTinv (Helper.cpp:196-271)
// ---------------------------------------------------------------------------
struct M3 {
  double v[3][3];
};

double compute_det(const M3& m) {  // SmallMatrix.hpp:349-368
  return m.v[0][0] * m.v[1][1] * m.v[2][2] + m.v[0][1] * m.v[1][2] * m.v[2][0] +
         m.v[0][2] * m.v[1][0] * m.v[2][1] - m.v[2][0] * m.v[1][1] * m.v[0][2] -
         m.v[2][1] * m.v[1][2] * m.v[0][0] - m.v[2][2] * m.v[1][0] * m.v[0][1];
}

double compute_inverse(const M3& m, M3& inv) {  // SmallMatrix.hpp:317-346
  double det = compute_det(m);
  const double(*a)[3] = m.v;
  inv.v[0][0] = (a[1][1] * a[2][2] - a[1][2] * a[2][1]) / det;
  inv.v[0][1] = (a[0][2] * a[2][1] - a[0][1] * a[2][2]) / det;
  inv.v[0][2] = (a[0][1] * a[1][2] - a[0][2] * a[1][1]) / det;
  inv.v[1][0] = (a[1][2] * a[2][0] - a[1][0] * a[2][2]) / det;
  inv.v[1][1] = (a[0][0] * a[2][2] - a[0][2] * a[2][0]) / det;
  inv.v[1][2] = (a[0][2] * a[1][0] - a[0][0] * a[1][2]) / det;
  inv.v[2][0] = (a[1][0] * a[2][1] - a[1][1] * a[2][0]) / det;
  inv.v[2][1] = (a[0][1] * a[2][0] - a[0][0] * a[2][1]) / det;
  inv.v[2][2] = (a[0][0] * a[1][1] - a[0][1] * a[1][0]) / det;
  return det;
}

// B[:,j] = x_{j+1} - x_0 (Helper.cpp:196-222); xl node-major
void build_transformation(const double* xl, M3& B) {
  for (int j = 0; j < 3; ++j)
    for (int i = 0; i < 3; ++i) B.v[i][j] = xl[3 * (j + 1) + i] - xl[i];
}

// dPhiTrans[w][i][d1] = sum_d2 dPhi[w][i][d2] * Binv[d2][d1] (Helper.cpp:258-271)
void apply_btinv(const vec3_d& in, vec3_d& out, const M3& Binv) {
  out.assign(in.size(), vec2_d(in[0].size(), vec_d(3, 0.)));
  for (size_t w = 0; w < in.size(); ++w)
    for (size_t i = 0; i < in[w].size(); ++i)
      for (int d1 = 0; d1 < 3; ++d1)
        for (int d2 = 0; d2 < 3; ++d2) out[w][i][d1] += in[w][i][d2] * Binv.v[d2][d1];
}

// ---------------------------------------------------------------------------
// AssembleFENavierStokes element (AssembleFENavierStokes_def.hpp)
// Element dof order: velocity 3*i+d (i<10), then pressure 30+i (i<4).
// ---------------------------------------------------------------------------
const int NSD = 34;
typedef std::vector<vec_d> Small;  // SmallMatrix<double>, row-major

void ns_laplacian(const double* xl, Small& el) {  // :150-196
  int deg = degree_func(2, 2);  // "UN Grad = 2" selects the degFunc overload -> 4
  vec3_d dPhi;
  vec_d w;
  getDPhi(dPhi, w, 2, deg);
  M3 B, Binv;
  build_transformation(xl, B);
  double absDetB = std::fabs(compute_inverse(B, Binv));
  vec3_d dPT;
  apply_btinv(dPhi, dPT, Binv);
  for (int i = 0; i < 10; ++i) {
    double value[10] = {0};
    for (int j = 0; j < 10; ++j) {
      for (size_t q = 0; q < dPT.size(); ++q)
        for (int d = 0; d < 3; ++d) value[j] += w[q] * dPT[q][i][d] * dPT[q][j][d];
      value[j] *= absDetB;
      for (int d = 0; d < 3; ++d) el[i * 3 + d][j * 3 + d] = value[j];
    }
  }
}

void ns_advection(const double* xl, const double* sol, Small& el) {  // :228-298
  int deg = degree_func(2, 2);
  vec3_d dPhi;
  vec2_d phiv;
  vec_d w;
  getDPhi(dPhi, w, 2, deg);
  getPhi(phiv, w, 2, deg);
  M3 B, Binv;
  build_transformation(xl, B);
  double absDetB = std::fabs(compute_inverse(B, Binv));
  vec3_d dPT;
  apply_btinv(dPhi, dPT, Binv);
  size_t nq = w.size();
  vec2_d uLoc(3, vec_d(nq, -1.));
  for (size_t q = 0; q < nq; ++q)
    for (int d = 0; d < 3; ++d) {
      uLoc[d][q] = 0.;
      for (int i = 0; i < 10; ++i) uLoc[d][q] += sol[3 * i + d] * phiv[q][i];
    }
  for (int i = 0; i < 10; ++i) {
    double value[10] = {0};
    for (int j = 0; j < 10; ++j) {
      for (size_t q = 0; q < nq; ++q)
        for (int d = 0; d < 3; ++d) value[j] += w[q] * uLoc[d][q] * phiv[q][i] * dPT[q][j][d];
      value[j] *= absDetB;
    }
    for (int d = 0; d < 3; ++d)
      for (int j = 0; j < 10; ++j) el[i * 3 + d][j * 3 + d] = value[j];
  }
}

void ns_advection_in_u(const double* xl, const double* sol, Small& el) {  // :302-369
  int deg = degree_func(2, 2);
  vec3_d dPhi;
  vec2_d phiv;
  vec_d w;
  getDPhi(dPhi, w, 2, deg);
  getPhi(phiv, w, 2, deg);
  M3 B, Binv;
  build_transformation(xl, B);
  double absDetB = std::fabs(compute_inverse(B, Binv));
  vec3_d dPT;
  apply_btinv(dPhi, dPT, Binv);
  size_t nq = w.size();
  std::vector<M3> duLoc(nq);
  for (auto& m : duLoc) std::memset(m.v, 0, sizeof(m.v));
  for (size_t q = 0; q < nq; ++q)
    for (int d1 = 0; d1 < 3; ++d1)
      for (int i = 0; i < 10; ++i)
        for (int d2 = 0; d2 < 3; ++d2) duLoc[q].v[d2][d1] += sol[3 * i + d1] * dPT[q][i][d2];
  for (int i = 0; i < 10; ++i)
    for (int d1 = 0; d1 < 3; ++d1) {
      double value[30] = {0};
      for (int j = 0; j < 10; ++j)
        for (int d2 = 0; d2 < 3; ++d2) {
          for (size_t q = 0; q < nq; ++q)
            value[3 * j + d2] += w[q] * duLoc[q].v[d2][d1] * phiv[q][i] * phiv[q][j];
          value[3 * j + d2] *= absDetB;
        }
      for (int j = 0; j < 10; ++j)
        for (int d = 0; d < 3; ++d) el[i * 3 + d1][j * 3 + d] = value[j * 3 + d];
    }
}

void ns_div_and_divt(const double* xl, Small& el) {  // :374-461
  int deg = 2;  // hard-coded (:381)
  vec3_d dPhi;
  vec2_d phiv;
  vec_d w;
  getDPhi(dPhi, w, 2, deg);
  getPhi(phiv, w, 1, deg);  // pressure P1 basis
  M3 B, Binv;
  build_transformation(xl, B);
  double absDetB = std::fabs(compute_inverse(B, Binv));
  vec3_d dPT;
  apply_btinv(dPhi, dPT, Binv);
  for (int i = 0; i < 4; ++i)
    for (int j = 0; j < 10; ++j)
      for (int d = 0; d < 3; ++d) {
        double value = 0.;
        for (size_t q = 0; q < w.size(); ++q) value += w[q] * phiv[q][i] * dPT[q][j][d];
        value *= absDetB;
        el[i + 30][3 * j + d] += value;
        el[3 * j + d][30 + i] += value;
      }
}

inline void small_add(const Small& b, Small& c) {  // SmallMatrix::add (SmallMatrix.hpp:140-171)
  for (int i = 0; i < NSD; ++i)
    for (int j = 0; j < NSD; ++j) c[i][j] = c[i][j] + b[i][j];
}
inline void small_scale(Small& a, double s) {  // SmallMatrix::scale (:229-237)
  for (auto& r : a)
    for (auto& x : r) x *= s;
}
inline Small small_new() { return Small(NSD, vec_d(NSD, 0.)); }

// constantMatrix_ (built when newtonStep_ == 0, :72-91)
Small ns_constant(const double* xl, double nu, double rho) {
  Small A = small_new(), Bm = small_new(), C = small_new();
  ns_laplacian(xl, A);
  small_scale(A, nu);
  small_scale(A, rho);
  small_add(A, C);
  ns_div_and_divt(xl, Bm);
  small_scale(Bm, -1.);
  small_add(Bm, C);
  return C;
}

// ANB_ = constantMatrix_ + rho*N (:93-97 / :133-137 / :204-209)
Small ns_anb(const double* xl, const double* sol, double rho, const Small& cst) {
  Small ANB = small_new(), N = small_new();
  small_add(cst, ANB);
  ns_advection(xl, sol, N);
  small_scale(N, rho);
  small_add(N, ANB);
  return ANB;
}

void ns_element(const double* xl, const double* sol, double nu, double rho, int newton,
                const double* coeff, double* jac, double* fp, double* rhs) {
  Small cst = ns_constant(xl, nu, rho);
  Small ANB = ns_anb(xl, sol, rho, cst);
  if (jac) {  // assembleJacobian (:67-112)
    Small J = small_new();
    small_add(ANB, J);
    if (newton) {
      Small W = small_new();
      ns_advection_in_u(xl, sol, W);
      small_scale(W, rho);
      small_add(W, J);
    }
    for (int i = 0; i < NSD; ++i)
      for (int j = 0; j < NSD; ++j) jac[i * NSD + j] = J[i][j];
  }
  if (fp) {  // assembleFixedPoint (:115-146): the fixed-point matrix is ANB_
    for (int i = 0; i < NSD; ++i)
      for (int j = 0; j < NSD; ++j) fp[i * NSD + j] = ANB[i][j];
  }
  if (rhs) {  // assembleRHS (:200-224)
    for (int i = 0; i < NSD; ++i) {
      int s = i >= 30 ? 1 : 0;
      double r = 0.;
      for (int j = 0; j < NSD; ++j) {
        int t = j >= 30 ? 1 : 0;
        r += ANB[i][j] * sol[j] * coeff[2 * s + t];
      }
      rhs[i] = r;
    }
  }
}

// FE::addFeBlock (FE_def.hpp:1419-1442): dofs*numNodes rows, 3x3 node blocks
void add_fe_block(Crs& A, const double* el, int ld, const int32_t* nodes, int nn, int dofs) {
  int32_t cols[10];
  double vals[10];
  for (int i = 0; i < nn; ++i)
    for (int di = 0; di < dofs; ++di) {
      int64_t row = (int64_t)dofs * nodes[i] + di;
      for (int d = 0; d < dofs; ++d) {
        for (int j = 0; j < nn; ++j) {
          cols[j] = dofs * nodes[j] + d;
          vals[j] = el[(dofs * i + di) * ld + dofs * j + d];
        }
        A.insert(row, cols, vals, nn);
      }
    }
}

// FE::addFeBlockMatrix (FE_def.hpp:982-1062): the four NS blocks
void add_fe_block_matrix(Crs& A, Crs& BT, Crs& B, Crs& C, const double* el, const int32_t* n1,
                         const int32_t* n2) {
  add_fe_block(A, el, NSD, n1, 10, 3);  // (0,0) identical to addFeBlock (:999-1012)
  const int off = 30;
  for (int i = 0; i < 4; ++i) {  // (1,1): only |v| > 1e-13 (:1017-1031)
    for (int j = 0; j < 4; ++j) {
      double v = el[(off + i) * NSD + off + j];
      if (std::fabs(v) > 1.e-13) {
        int32_t c = n2[j];
        C.insert(n2[i], &c, &v, 1);
      }
    }
  }
  int32_t cols[10];
  double vals[10];
  for (int i = 0; i < 10; ++i)  // (0,1) (:1035-1046)
    for (int di = 0; di < 3; ++di) {
      int64_t row = 3 * (int64_t)n1[i] + di;
      for (int j = 0; j < 4; ++j) {
        vals[j] = el[(i * 3 + di) * NSD + off + j];
        cols[j] = n2[j];
      }
      BT.insert(row, cols, vals, 4);
    }
  for (int j = 0; j < 4; ++j)  // (1,0) (:1048-1059)
    for (int d = 0; d < 3; ++d) {
      for (int i = 0; i < 10; ++i) {
        vals[i] = el[(off + j) * NSD + 3 * i + d];
        cols[i] = 3 * n1[i] + d;
      }
      B.insert(n2[j], cols, vals, 10);
    }
}

// ---------------------------------------------------------------------------
// Hyperelastic tet-10 (AceGen SKR10 restated from the strain energies; see the
// header comment of orc_nh_skr for the conventions)
// ---------------------------------------------------------------------------

// AceGen node parametrisation (NeoHookQuadraticTets2.f:88-110): local node 1 is
// the xi1 vertex, 2 the xi2 vertex, 3 the xi3 vertex, 4 the origin; edge nodes
// 5(1,2) 6(2,3) 7(1,3) 8(1,4) 9(2,4) 10(3,4).  FEDDLib's local order
// (MeshStructured_def.hpp:906-987) is passed in unchanged, so FEDDLib node k
// plays AceGen node k+1.
void ace_dshape(const double* xi, double dN[10][3]) {
  double L1 = xi[0], L2 = xi[1], L3 = xi[2], L4 = 1. - xi[0] - xi[1] - xi[2];
  // N1 = L1(2L1-1) ... N4 = L4(2L4-1); N5 = 4L1L2, N6 = 4L2L3, N7 = 4L1L3,
  // N8 = 4L1L4, N9 = 4L2L4, N10 = 4L3L4.  dL4/dxi = (-1,-1,-1).
  double g[4][3] = {{1, 0, 0}, {0, 1, 0}, {0, 0, 1}, {-1, -1, -1}};
  double L[4] = {L1, L2, L3, L4};
  for (int k = 0; k < 4; ++k)
    for (int a = 0; a < 3; ++a) dN[k][a] = (4. * L[k] - 1.) * g[k][a];
  const int ed[6][2] = {{0, 1}, {1, 2}, {0, 2}, {0, 3}, {1, 3}, {2, 3}};
  for (int e = 0; e < 6; ++e) {
    int p = ed[e][0], q = ed[e][1];
    for (int a = 0; a < 3; ++a) dN[4 + e][a] = 4. * (L[p] * g[q][a] + L[q] * g[p][a]);
  }
}

// SMSIntPoints icode 18 (4-pt) and 19 (5-pt) (SMSUtility.f:344-363, 498-502)
int ace_points(int model, double gp[5][4]) {
  if (model == 0) {
    double a = 0.585410196624968454, b = 0.138196601125010515, w = 0.0416666666666666667;
    double t[4][4] = {{a, b, b, w}, {b, a, b, w}, {b, b, a, w}, {b, b, b, w}};
    for (int i = 0; i < 4; ++i)
      for (int k = 0; k < 4; ++k) gp[i][k] = t[i][k];
    return 4;
  }
  double t[5][4] = {{0.25, 0.25, 0.25, -0.133333333333333333},
                    {0.5, 1. / 6., 1. / 6., 0.075},
                    {1. / 6., 0.5, 1. / 6., 0.075},
                    {1. / 6., 1. / 6., 0.5, 0.075},
                    {1. / 6., 1. / 6., 1. / 6., 0.075}};
  for (int i = 0; i < 5; ++i)
    for (int k = 0; k < 4; ++k) gp[i][k] = t[i][k];
  return 5;
}

// Second Piola-Kirchhoff stress S and material tangent CC = dS/dE for the three
// AceGen laws, with mu = E/(2(1+nu)):
//  model 0 (skr1, NeoHookQuadraticTets.f:27-32):   W = mu/2(I1-3) - mu lnJ + lam/4(J^2-1-2lnJ)
//  model 1 (skr2, NeoHookQuadraticTets2.f:27-30):  W = mu/2(I1-3) - mu lnJ + lam/2 (lnJ)^2
//  model 2 (skr3, NeoHookQuadraticTets3.f:27-30):  W = mu/2(I1 I3^(-1/3) - 3) + K/2 (J-1)^2,
//                                                   K = E/(3(1-2nu))
// lam = 2 mu nu/(1-2nu).
void material(int model, double Emod, double nu, const double C[3][3], double S[3][3],
              double CC[3][3][3][3]) {
  double mu = Emod / (2. * (1. + nu));
  double lam = 2. * mu * nu / (1. - 2. * nu);
  double I3 = C[0][0] * (C[1][1] * C[2][2] - C[1][2] * C[2][1]) -
              C[0][1] * (C[1][0] * C[2][2] - C[1][2] * C[2][0]) +
              C[0][2] * (C[1][0] * C[2][1] - C[1][1] * C[2][0]);
  double Ci[3][3];
  Ci[0][0] = (C[1][1] * C[2][2] - C[1][2] * C[2][1]) / I3;
  Ci[0][1] = (C[0][2] * C[2][1] - C[0][1] * C[2][2]) / I3;
  Ci[0][2] = (C[0][1] * C[1][2] - C[0][2] * C[1][1]) / I3;
  Ci[1][0] = (C[1][2] * C[2][0] - C[1][0] * C[2][2]) / I3;
  Ci[1][1] = (C[0][0] * C[2][2] - C[0][2] * C[2][0]) / I3;
  Ci[1][2] = (C[0][2] * C[1][0] - C[0][0] * C[1][2]) / I3;
  Ci[2][0] = (C[1][0] * C[2][1] - C[1][1] * C[2][0]) / I3;
  Ci[2][1] = (C[0][1] * C[2][0] - C[0][0] * C[2][1]) / I3;
  Ci[2][2] = (C[0][0] * C[1][1] - C[0][1] * C[1][0]) / I3;
  double J = std::sqrt(I3);
  double I1 = C[0][0] + C[1][1] + C[2][2];
  double a_id = 0, a_ci = 0;  // S = a_id*I + a_ci*Ci
  double c_cc = 0, c_ii = 0, c_ic = 0;  // CC = c_cc Ci(x)Ci + c_ii IICi + c_ic (I(x)Ci + Ci(x)I)
  if (model == 0) {
    a_id = mu;
    a_ci = lam / 2. * I3 - mu - lam / 2.;
    c_cc = lam * I3;
    c_ii = 2. * mu + lam - lam * I3;
  } else if (model == 1) {
    double lnJ = std::log(J);
    a_id = mu;
    a_ci = lam * lnJ - mu;
    c_cc = lam;
    c_ii = 2. * (mu - lam * lnJ);
  } else {
    double K = Emod / (3. * (1. - 2. * nu));
    double f = std::pow(I3, -1. / 3.);
    a_id = mu * f;
    a_ci = -mu / 3. * I1 * f + K * (J - 1.) * J;
    c_cc = 2. / 9. * mu * f * I1 + K * (2. * J - 1.) * J;
    c_ii = 2. / 3. * mu * f * I1 - 2. * K * (J - 1.) * J;
    c_ic = -2. / 3. * mu * f;
  }
  for (int A = 0; A < 3; ++A)
    for (int B = 0; B < 3; ++B) S[A][B] = (A == B ? a_id : 0.) + a_ci * Ci[A][B];
  for (int A = 0; A < 3; ++A)
    for (int B = 0; B < 3; ++B)
      for (int Cc = 0; Cc < 3; ++Cc)
        for (int D = 0; D < 3; ++D)
          CC[A][B][Cc][D] = c_cc * Ci[A][B] * Ci[Cc][D] +
                            c_ii * 0.5 * (Ci[A][Cc] * Ci[B][D] + Ci[A][D] * Ci[B][Cc]) +
                            c_ic * ((A == B ? Ci[Cc][D] : 0.) + (Cc == D ? Ci[A][B] : 0.));
}

void nh_skr(int model, double Emod, double nu, const double* xl, const double* ul, double* s,
            double* p) {
  double gp[5][4];
  int nq = ace_points(model, gp);
  for (int q = 0; q < nq; ++q) {
    double dN[10][3];
    ace_dshape(gp[q], dN);
    // isoparametric Jacobian Jg[i][a] = sum_k x_{k,i} dN_k/dxi_a and its inverse
    double Jg[3][3] = {{0}};
    for (int k = 0; k < 10; ++k)
      for (int i = 0; i < 3; ++i)
        for (int a = 0; a < 3; ++a) Jg[i][a] += xl[3 * k + i] * dN[k][a];
    M3 Jm, Ji;
    std::memcpy(Jm.v, Jg, sizeof(Jg));
    double detJ = compute_inverse(Jm, Ji);  // signed, AceGen v(126)
    double G[10][3];  // dN/dX = dN/dxi * Jinv
    for (int k = 0; k < 10; ++k)
      for (int A = 0; A < 3; ++A) {
        G[k][A] = 0;
        for (int a = 0; a < 3; ++a) G[k][A] += dN[k][a] * Ji.v[a][A];
      }
    double F[3][3];
    for (int i = 0; i < 3; ++i)
      for (int A = 0; A < 3; ++A) {
        F[i][A] = (i == A) ? 1. : 0.;
        for (int k = 0; k < 10; ++k) F[i][A] += ul[3 * k + i] * G[k][A];
      }
    double C[3][3];
    for (int A = 0; A < 3; ++A)
      for (int B = 0; B < 3; ++B) {
        C[A][B] = 0;
        for (int i = 0; i < 3; ++i) C[A][B] += F[i][A] * F[i][B];
      }
    double S[3][3], CC[3][3][3][3];
    material(model, Emod, nu, C, S, CC);
    double P[3][3];  // P = F S
    for (int i = 0; i < 3; ++i)
      for (int A = 0; A < 3; ++A) {
        P[i][A] = 0;
        for (int B = 0; B < 3; ++B) P[i][A] += F[i][B] * S[B][A];
      }
    // AA[a][A][b][B] = delta_ab S_AB + F_aC CC_CABD F_bD
    double AA[3][3][3][3];
    for (int a = 0; a < 3; ++a)
      for (int A = 0; A < 3; ++A)
        for (int b = 0; b < 3; ++b)
          for (int B = 0; B < 3; ++B) {
            double v = (a == b) ? S[A][B] : 0.;
            for (int Cc = 0; Cc < 3; ++Cc)
              for (int D = 0; D < 3; ++D) v += F[a][Cc] * CC[Cc][A][B][D] * F[b][D];
            AA[a][A][b][B] = v;
          }
    double wq = gp[q][3];
    for (int i = 0; i < 10; ++i)
      for (int a = 0; a < 3; ++a) {
        double f = 0;
        for (int A = 0; A < 3; ++A) f += P[a][A] * G[i][A];
        p[3 * i + a] += detJ * wq * f;
      }
    for (int i = 0; i < 10; ++i)
      for (int a = 0; a < 3; ++a)
        for (int j = 0; j < 10; ++j)
          for (int b = 0; b < 3; ++b) {
            double k = 0;
            for (int A = 0; A < 3; ++A)
              for (int B = 0; B < 3; ++B) k += G[i][A] * AA[a][A][b][B] * G[j][B];
            s[(3 * j + b) * 30 + 3 * i + a] += wq * detJ * k;  // column-major s(30,30)
          }
  }
}

}  // namespace

// ===========================================================================
// C ABI
// ===========================================================================
extern "C" {

void orc_mesh_p1(int M, double* pts, int32_t* bc, int32_t* elems) {
  // MeshStructured_def.hpp:703-806 with N = 1, one rank, unit cube at the origin
  const double eps = std::numeric_limits<double>::epsilon();
  double h = 1.0 / M;
  int64_t c = 0;
  for (int t = 0; t < M + 1; ++t)
    for (int s = 0; s < M + 1; ++s)
      for (int r = 0; r < M + 1; ++r) {
        double x[3] = {r * h, s * h, t * h};
        int flag = 0;
        for (int d = 0; d < 3; ++d) {
          if (x[d] < eps && x[d] > -eps) x[d] = 0.0;
          if (x[d] > 1.0 - eps || x[d] < eps) flag = 1;
          pts[3 * c + d] = x[d];
        }
        bc[c] = flag;
        ++c;
      }
  int64_t e = 0;
  const int64_t n = M + 1;
  for (int t = 0; t < M; ++t)
    for (int s = 0; s < M; ++s)
      for (int r = 0; r < M; ++r) {
        auto id = [&](int rr, int ss, int tt) { return (int32_t)(rr + n * ss + n * n * tt); };
        int32_t T[6][4] = {
            {id(r + 1, s, t), id(r, s, t), id(r + 1, s, t + 1), id(r + 1, s + 1, t + 1)},
            {id(r, s, t + 1), id(r, s, t), id(r + 1, s, t + 1), id(r + 1, s + 1, t + 1)},
            {id(r + 1, s, t), id(r, s, t), id(r + 1, s + 1, t), id(r + 1, s + 1, t + 1)},
            {id(r, s, t), id(r, s + 1, t), id(r + 1, s + 1, t), id(r + 1, s + 1, t + 1)},
            {id(r, s, t), id(r, s + 1, t), id(r, s + 1, t + 1), id(r + 1, s + 1, t + 1)},
            {id(r, s, t), id(r, s, t + 1), id(r, s + 1, t + 1), id(r + 1, s + 1, t + 1)}};
        for (int k = 0; k < 6; ++k) {
          for (int j = 0; j < 4; ++j) elems[4 * e + j] = T[k][j];
          ++e;
        }
      }
}

void orc_mesh_p2(int M, double* pts, int32_t* bc, int32_t* elems) {
  // MeshStructured_def.hpp:808-994 with N = 1, one rank
  const double eps = std::numeric_limits<double>::epsilon();
  double h = 1.0 / M;
  const int64_t P = 2 * (M + 1) - 1;
  int64_t c = 0;
  for (int t = 0; t < P; ++t)
    for (int s = 0; s < P; ++s)
      for (int r = 0; r < P; ++r) {
        double x[3] = {r * h / 2.0, s * h / 2.0, t * h / 2.0};
        int flag = 0;
        for (int d = 0; d < 3; ++d) {
          if (x[d] < eps && x[d] > -eps) x[d] = 0.0;
          if (x[d] > 1.0 - eps || x[d] < eps) flag = 1;
          pts[3 * c + d] = x[d];
        }
        bc[c] = flag;
        ++c;
      }
  int64_t e = 0;
  const int64_t P2 = P * P;
  for (int64_t t = 0; t < M; ++t)
    for (int64_t s = 0; s < M; ++s)
      for (int64_t r = 0; r < M; ++r) {
        int64_t T[6][10] = {
            {2 * (r + 1) + 2 * P * s + 2 * P2 * t, 2 * r + 2 * P * s + 2 * P2 * t,
             2 * (r + 1) + 2 * P * s + 2 * P2 * (t + 1), 2 * (r + 1) + 2 * P * (s + 1) + 2 * P2 * (t + 1),
             2 * r + 1 + 2 * P * s + 2 * P2 * t, 2 * r + 1 + 2 * P * s + 2 * P2 * t + P2,
             2 * (r + 1) + 2 * P * s + 2 * P2 * t + P2, 2 * (r + 1) + 2 * P * s + P + 2 * P2 * t + P2,
             2 * r + 1 + 2 * P * s + P + 2 * P2 * t + P2, 2 * (r + 1) + 2 * P * s + P + 2 * P2 * (t + 1)},
            {2 * r + 2 * P * s + 2 * P2 * (t + 1), 2 * r + 2 * P * s + 2 * P2 * t,
             2 * (r + 1) + 2 * P * s + 2 * P2 * (t + 1), 2 * (r + 1) + 2 * P * (s + 1) + 2 * P2 * (t + 1),
             2 * r + 2 * P * s + 2 * P2 * t + P2, 2 * r + 1 + 2 * P * s + 2 * P2 * t + P2,
             2 * r + 1 + 2 * P * s + 2 * P2 * (t + 1), 2 * r + 1 + 2 * P * s + P + 2 * P2 * (t + 1),
             2 * r + 1 + 2 * P * s + P + 2 * P2 * t + P2, 2 * (r + 1) + 2 * P * s + P + 2 * P2 * (t + 1)},
            {2 * (r + 1) + 2 * P * s + 2 * P2 * t, 2 * r + 2 * P * s + 2 * P2 * t,
             2 * (r + 1) + 2 * P * (s + 1) + 2 * P2 * t, 2 * (r + 1) + 2 * P * (s + 1) + 2 * P2 * (t + 1),
             2 * r + 1 + 2 * P * s + 2 * P2 * t, 2 * r + 1 + 2 * P * s + P + 2 * P2 * t,
             2 * (r + 1) + 2 * P * s + P + 2 * P2 * t, 2 * (r + 1) + 2 * P * s + P + 2 * P2 * t + P2,
             2 * r + 1 + 2 * P * s + P + 2 * P2 * t + P2, 2 * (r + 1) + 2 * P * (s + 1) + 2 * P2 * t + P2},
            {2 * r + 2 * P * s + 2 * P2 * t, 2 * r + 2 * P * (s + 1) + 2 * P2 * t,
             2 * (r + 1) + 2 * P * (s + 1) + 2 * P2 * t, 2 * (r + 1) + 2 * P * (s + 1) + 2 * P2 * (t + 1),
             2 * r + 2 * P * s + P + 2 * P2 * t, 2 * r + 1 + 2 * P * (s + 1) + 2 * P2 * t,
             2 * r + 1 + 2 * P * s + P + 2 * P2 * t, 2 * r + 1 + 2 * P * s + P + 2 * P2 * t + P2,
             2 * r + 1 + 2 * P * (s + 1) + 2 * P2 * t + P2, 2 * (r + 1) + 2 * P * (s + 1) + 2 * P2 * t + P2},
            {2 * r + 2 * P * s + 2 * P2 * t, 2 * r + 2 * P * (s + 1) + 2 * P2 * t,
             2 * r + 2 * P * (s + 1) + 2 * P2 * (t + 1), 2 * (r + 1) + 2 * P * (s + 1) + 2 * P2 * (t + 1),
             2 * r + 2 * P * s + P + 2 * P2 * t, 2 * r + 2 * P * (s + 1) + 2 * P2 * t + P2,
             2 * r + 2 * P * s + P + 2 * P2 * t + P2, 2 * r + 1 + 2 * P * s + P + 2 * P2 * t + P2,
             2 * r + 1 + 2 * P * (s + 1) + 2 * P2 * t + P2, 2 * r + 1 + 2 * P * (s + 1) + 2 * P2 * (t + 1)},
            {2 * r + 2 * P * s + 2 * P2 * t, 2 * r + 2 * P * s + 2 * P2 * (t + 1),
             2 * r + 2 * P * (s + 1) + 2 * P2 * (t + 1), 2 * (r + 1) + 2 * P * (s + 1) + 2 * P2 * (t + 1),
             2 * r + 2 * P * s + 2 * P2 * t + P2, 2 * r + 2 * P * s + P + 2 * P2 * (t + 1),
             2 * r + 2 * P * s + P + 2 * P2 * t + P2, 2 * r + 1 + 2 * P * s + P + 2 * P2 * t + P2,
             2 * r + 1 + 2 * P * s + P + 2 * P2 * (t + 1), 2 * r + 1 + 2 * P * (s + 1) + 2 * P2 * (t + 1)}};
        for (int k = 0; k < 6; ++k) {
          for (int j = 0; j < 10; ++j) elems[10 * e + j] = (int32_t)T[k][j];
          ++e;
        }
      }
}

int64_t orc_correct_orientation(int npe, const double* pts, int64_t E, int32_t* el) {
  int64_t flipped = 0;
  for (int64_t T = 0; T < E; ++T) {
    int32_t* n = el + (int64_t)npe * T;
    double xl[12];
    for (int k = 0; k < 4; ++k)
      for (int d = 0; d < 3; ++d) xl[3 * k + d] = pts[3 * (int64_t)n[k] + d];
    M3 B, Bi;
    build_transformation(xl, B);
    double det = compute_inverse(B, Bi);
    if (det < 0) {  // flipElement (Mesh_def.hpp:648-726)
      ++flipped;
      if (npe == 4) {
        std::swap(n[1], n[2]);
      } else {
        int32_t o[10];
        std::memcpy(o, n, sizeof(o));
        const int perm[10] = {0, 2, 1, 3, 6, 5, 4, 7, 9, 8};
        for (int k = 0; k < 10; ++k) n[k] = o[perm[k]];
      }
    }
  }
  return flipped;
}

void* orc_crs_create(int64_t nrows) { return new Crs(nrows); }
void orc_crs_destroy(void* A) { delete static_cast<Crs*>(A); }
void orc_crs_fill_complete(void* A) { static_cast<Crs*>(A)->fill_complete(); }
int64_t orc_crs_nnz(void* A) { return static_cast<Crs*>(A)->nnz(); }
int64_t orc_crs_nrows(void* A) { return (int64_t) static_cast<Crs*>(A)->rows.size(); }
void orc_crs_export(void* Ap, int64_t* rowptr, int32_t* colind, double* vals) {
  Crs* A = static_cast<Crs*>(Ap);
  int64_t k = 0;
  rowptr[0] = 0;
  for (size_t r = 0; r < A->rows.size(); ++r) {
    for (auto& cv : A->rows[r]) {
      if (colind) colind[k] = cv.first;
      if (vals) vals[k] = cv.second;
      ++k;
    }
    rowptr[r + 1] = k;
  }
}

int orc_quadrature(int degree, double* pts, double* w) {
  vec2_d qp;
  vec_d qw;
  int n = quadrature(degree, qp, qw);
  for (int k = 0; k < n; ++k) {
    for (int d = 0; d < 3; ++d) pts[3 * k + d] = qp[k][d];
    w[k] = qw[k];
  }
  return n;
}
void orc_phi(int fe, int i, const double* p, double* value) { *value = phi(fe, i, p); }
void orc_grad_phi(int fe, int i, const double* p, double* v) { grad_phi(fe, i, p, v); }
int orc_degree_func(int fe, int degfunc) { return degree_func(fe, degfunc); }
int orc_degree_type(int fe, int type) { return degree_type(fe, type); }
int orc_degree_pair(int fe1, int fe2, int t1, int t2, int extra) {
  return degree_pair(fe1, fe2, t1, t2, extra);
}

void orc_set_element_mode(int mode) { g_elem_mode = (mode == 1 || mode == 2) ? mode : 0; }

void* orc_assembly_laplace(int fe, const double* pts, int64_t npts, const int32_t* elems, int64_t E) {
  int n = nloc(fe);
  Crs* A = new Crs(npts);
  int deg = degree_pair(fe, fe, 1, 1, 0);  // FE_def.hpp:2459
  vec3_d dPhi;
  vec_d w;
  getDPhi(dPhi, w, fe, deg);
  for (int64_t T = 0; T < E; ++T) {
    const int32_t* nd = elems + (int64_t)n * T;
    double xl[30];
    for (int k = 0; k < n; ++k)
      for (int d = 0; d < 3; ++d) xl[3 * k + d] = pts[3 * (int64_t)nd[k] + d];
    M3 B, Binv;
    build_transformation(xl, B);
    double absDetB = std::fabs(compute_inverse(B, Binv));
    vec3_d dPT;
    apply_btinv(dPhi, dPT, Binv);
    double value[10][10] = {};
    int32_t idx[10];
    for (int i = 0; i < n; ++i)
      for (int j = 0; j < n; ++j) {
        for (size_t q = 0; q < w.size(); ++q)
          for (int d = 0; d < 3; ++d) value[i][j] += w[q] * dPT[q][i][d] * dPT[q][j][d];
        value[i][j] *= absDetB;
        idx[j] = nd[j];
      }
    elem_scale(&value[0][0], 10, 0, n, 0, n);
    for (int i = 0; i < n; ++i) A->insert(nd[i], idx, value[i], n);
  }
  A->fill_complete();
  return A;
}

void* orc_assembly_laplace_assfe(int fe, int dofs, const double* pts, int64_t npts,
                                 const int32_t* elems, int64_t E) {
  int n = nloc(fe);
  Crs* A = new Crs(npts * dofs);
  int deg = degree_func(fe, 2);  // "UN Grad = 2" (AssembleFE_Laplace_def.hpp:58,65)
  std::vector<double> el((size_t)(n * dofs) * (n * dofs));
  for (int64_t T = 0; T < E; ++T) {
    const int32_t* nd = elems + (int64_t)n * T;
    double xl[30];
    for (int k = 0; k < n; ++k)
      for (int d = 0; d < 3; ++d) xl[3 * k + d] = pts[3 * (int64_t)nd[k] + d];
    vec3_d dPhi;
    vec_d w;
    getDPhi(dPhi, w, fe, deg);  // rebuilt per element (:66)
    M3 B, Binv;
    build_transformation(xl, B);
    double absDetB = std::fabs(compute_inverse(B, Binv));
    vec3_d dPT;
    apply_btinv(dPhi, dPT, Binv);
    std::fill(el.begin(), el.end(), 0.);
    int ld = n * dofs;
    for (int i = 0; i < n; ++i)
      for (int j = 0; j < n; ++j) {
        double v = 0;
        for (size_t q = 0; q < w.size(); ++q)
          for (int d = 0; d < 3; ++d) v += w[q] * dPT[q][i][d] * dPT[q][j][d];
        v *= absDetB;
        for (int d = 0; d < dofs; ++d) el[(i * dofs + d) * ld + j * dofs + d] = v;
      }
    elem_scale(el.data(), ld, 0, ld, 0, ld);
    add_fe_block(*A, el.data(), ld, nd, n, dofs);  // FE_def.hpp:1846
  }
  A->fill_complete();
  return A;
}

void orc_assembly_rhs(int fe, const double* pts, int64_t npts, const int32_t* elems, int64_t E,
                      double f, double* rhs) {
  (void)npts;
  int n = nloc(fe);
  int deg = degree_type(fe, 0);  // Std (FE_def.hpp:7327)
  vec2_d phiv;
  vec_d w;
  getPhi(phiv, w, fe, deg);
  for (int64_t T = 0; T < E; ++T) {
    const int32_t* nd = elems + (int64_t)n * T;
    double xl[30];
    for (int k = 0; k < n; ++k)
      for (int d = 0; d < 3; ++d) xl[3 * k + d] = pts[3 * (int64_t)nd[k] + d];
    M3 B;
    build_transformation(xl, B);
    double absDetB = std::fabs(compute_det(B));
    for (int i = 0; i < n; ++i) {
      double value = 0.;
      for (size_t q = 0; q < w.size(); ++q) value += w[q] * phiv[q][i] * f;
      value *= absDetB;
      rhs[nd[i]] += value;
    }
  }
}

// FE::assemblyRHS, general source (FE_def.hpp:7352-7391): the quadrature points of the
// rule determineDegree(dim, FEType, Std) mapped to each element, quadPointsTrans[w][j] =
// sum_k B[j][k] xi_w[k] + x_0[j] (:7357-7366).
int orc_rhs_quadrature_points(int fe, const double* pts, const int32_t* elems, int64_t E, double* xq) {
  const int n = nloc(fe);
  vec2_d qp;
  vec_d w;
  const int nq = quadrature(degree_type(fe, 0), qp, w);
  for (int64_t T = 0; T < E; ++T) {
    const int32_t* nd = elems + (int64_t)n * T;
    double xl[30];
    for (int k = 0; k < n; ++k)
      for (int d = 0; d < 3; ++d) xl[3 * k + d] = pts[3 * (int64_t)nd[k] + d];
    M3 B;
    build_transformation(xl, B);
    for (int q = 0; q < nq; ++q)
      for (int j = 0; j < 3; ++j) {
        double v = 0.;
        for (int k = 0; k < 3; ++k) v += B.v[j][k] * qp[q][k];
        xq[(T * nq + q) * 3 + j] = v + xl[j];
      }
  }
  return nq;
}

// ... and the load vector from the source values at those points, fq[(T nq + w) ncomp + d]:
// value = |det B| sum_w w phi_i(w) f_d(x_w), added at row node ("Scalar", ncomp = 1) or
// 3 node + d ("Vector", ncomp = 3) (:7367-7389).
void orc_assembly_rhs_points(int fe, const double* pts, int64_t npts, const int32_t* elems, int64_t E,
                             const double* fq, int ncomp, double* rhs) {
  (void)npts;
  const int n = nloc(fe);
  vec2_d phiv;
  vec_d w;
  getPhi(phiv, w, fe, degree_type(fe, 0));
  const int nq = (int)w.size();
  for (int64_t T = 0; T < E; ++T) {
    const int32_t* nd = elems + (int64_t)n * T;
    double xl[30];
    for (int k = 0; k < n; ++k)
      for (int d = 0; d < 3; ++d) xl[3 * k + d] = pts[3 * (int64_t)nd[k] + d];
    M3 B;
    build_transformation(xl, B);
    const double absDetB = std::fabs(compute_det(B));
    for (int i = 0; i < n; ++i)
      for (int d = 0; d < ncomp; ++d) {
        double value = 0.;
        for (int q = 0; q < nq; ++q) value += w[q] * phiv[q][i] * fq[(T * nq + q) * ncomp + d];
        value *= absDetB;
        rhs[(int64_t)ncomp * nd[i] + d] += value;
      }
  }
}

void orc_ns_element(const double* xl, const double* sol, double nu, double rho, int newton,
                    const double* coeff4, double* jac, double* fp, double* rhs) {
  static const double one4[4] = {1, 1, 1, 1};
  ns_element(xl, sol, nu, rho, newton, coeff4 ? coeff4 : one4, jac, fp, rhs);
}

void orc_assembly_ns(const double* pts2, int64_t np2, const int32_t* el2, const int32_t* el1,
                     int64_t E, int64_t np1, const double* u, const double* p, double nu,
                     double rho, int newton, int mode, int reassemble, const double* coeff4,
                     void** A, void** BT, void** B, void** C, double* res_u, double* res_p) {
  static const double one4[4] = {1, 1, 1, 1};
  const double* coeff = coeff4 ? coeff4 : one4;
  Crs *a = nullptr, *bt = nullptr, *b = nullptr, *c = nullptr;
  if (mode != 2) {
    a = new Crs(3 * np2);
    if (!reassemble) {
      bt = new Crs(3 * np2);
      b = new Crs(np1);
      c = new Crs(np1);
    }
  }
  std::vector<double> el(NSD * NSD);
  double rhs[NSD];
  for (int64_t T = 0; T < E; ++T) {
    const int32_t* n2 = el2 + 10 * T;
    const int32_t* n1 = el1 + 4 * T;
    double xl[30], sol[NSD];
    for (int k = 0; k < 10; ++k)
      for (int d = 0; d < 3; ++d) {
        xl[3 * k + d] = pts2[3 * (int64_t)n2[k] + d];
        sol[3 * k + d] = u[3 * (int64_t)n2[k] + d];  // FE::getSolution (FE_def.hpp:1508-1519)
      }
    for (int k = 0; k < 4; ++k) sol[30 + k] = p[n1[k]];
    if (mode == 0 || mode == 1) {
      if (mode == 0)
        ns_element(xl, sol, nu, rho, newton, coeff, el.data(), nullptr, nullptr);
      else
        ns_element(xl, sol, nu, rho, newton, coeff, nullptr, el.data(), nullptr);
      elem_scale(el.data(), NSD, 0, 30, 0, 30);
      elem_scale(el.data(), NSD, 0, 30, 30, 4);
      elem_scale(el.data(), NSD, 30, 4, 0, 30);
      if (reassemble)
        add_fe_block(*a, el.data(), NSD, n2, 10, 3);
      else
        add_fe_block_matrix(*a, *bt, *b, *c, el.data(), n2, n1);
    } else {
      ns_element(xl, sol, nu, rho, newton, coeff, nullptr, nullptr, rhs);
      for (int i = 0; i < 10; ++i)  // addFeBlockMv (FE_def.hpp:1381-1403)
        for (int d = 0; d < 3; ++d) res_u[3 * (int64_t)n2[i] + d] += rhs[3 * i + d];
      for (int i = 0; i < 4; ++i) res_p[n1[i]] += rhs[30 + i];
    }
  }
  if (a) a->fill_complete();
  if (bt) bt->fill_complete();
  if (b) b->fill_complete();
  if (c) c->fill_complete();
  if (A) *A = a;
  if (BT) *BT = bt;
  if (B) *B = b;
  if (C) *C = c;
}

double orc_time_ns_jacobian(const double* pts2, int64_t np2, const int32_t* el2, const int32_t* el1,
                            int64_t nelem, int64_t np1, const double* u, const double* p,
                            double nu, double rho) {
  auto t0 = std::chrono::steady_clock::now();
  void *A, *BT, *B, *C;
  static const double one4[4] = {1, 1, 1, 1};
  orc_assembly_ns(pts2, np2, el2, el1, nelem, np1, u, p, nu, rho, 1, 0, 0, one4, &A, &BT, &B, &C,
                  nullptr, nullptr);
  auto t1 = std::chrono::steady_clock::now();
  orc_crs_destroy(A);
  orc_crs_destroy(BT);
  orc_crs_destroy(B);
  orc_crs_destroy(C);
  return std::chrono::duration<double>(t1 - t0).count();
}

// The reference's unit of parallelism is an MPI rank assembling its own elements
// into its own rows, followed by the export-Add of globalAssemble/fillComplete.
// Here `nthreads` element partitions are assembled concurrently (each with its own
// row-wise insertion and fillComplete), then the partial matrices are merged row by
// row in parallel.  Returns the wall time of both phases.
static Crs* merge_parts(const std::vector<Crs*>& parts, int nthreads) {
  const int64_t n = (int64_t)parts[0]->rows.size();
  Crs* M = new Crs(n);
  std::vector<std::thread> th;
  for (int t = 0; t < nthreads; ++t)
    th.emplace_back([&, t]() {
      const int64_t r0 = n * t / nthreads, r1 = n * (t + 1) / nthreads;
      for (int64_t r = r0; r < r1; ++r) {
        auto& out = M->rows[r];
        for (const Crs* P : parts) out.insert(out.end(), P->rows[r].begin(), P->rows[r].end());
        std::stable_sort(out.begin(), out.end(),
                         [](const std::pair<int32_t, double>& a, const std::pair<int32_t, double>& b) {
                           return a.first < b.first;
                         });
        size_t w = 0;
        for (size_t k = 0; k < out.size(); ++k) {
          if (w > 0 && out[w - 1].first == out[k].first)
            out[w - 1].second += out[k].second;
          else
            out[w++] = out[k];
        }
        out.resize(w);
      }
    });
  for (auto& x : th) x.join();
  M->filled = true;
  return M;
}

double orc_time_ns_jacobian_mt(const double* pts2, int64_t np2, const int32_t* el2, const int32_t* el1,
                               int64_t nelem, int64_t np1, const double* u, const double* p, double nu,
                               double rho, int nthreads) {
  if (nthreads < 1) nthreads = 1;
  static const double one4[4] = {1, 1, 1, 1};
  auto t0 = std::chrono::steady_clock::now();
  std::vector<void*> parts(4 * nthreads, nullptr);
  std::vector<std::thread> th;
  for (int t = 0; t < nthreads; ++t)
    th.emplace_back([&, t]() {
      const int64_t e0 = nelem * t / nthreads, e1 = nelem * (t + 1) / nthreads;
      orc_assembly_ns(pts2, np2, el2 + 10 * e0, el1 + 4 * e0, e1 - e0, np1, u, p, nu, rho, 1, 0, 0, one4,
                      &parts[4 * t], &parts[4 * t + 1], &parts[4 * t + 2], &parts[4 * t + 3], nullptr, nullptr);
    });
  for (auto& x : th) x.join();
  std::vector<Crs*> merged;
  for (int b = 0; b < 4; ++b) {
    std::vector<Crs*> blk;
    for (int t = 0; t < nthreads; ++t) blk.push_back(static_cast<Crs*>(parts[4 * t + b]));
    merged.push_back(nthreads > 1 ? merge_parts(blk, nthreads) : nullptr);
  }
  auto t1 = std::chrono::steady_clock::now();
  for (auto* q : parts) orc_crs_destroy(q);
  for (auto* m : merged) delete m;
  return std::chrono::duration<double>(t1 - t0).count();
}

void orc_nh_skr(int model, double E, double nu, const double* xl, const double* ul, double* s,
                double* p) {
  std::memset(s, 0, sizeof(double) * 900);
  std::memset(p, 0, sizeof(double) * 30);
  nh_skr(model, E, nu, xl, ul, s, p);
}

void* orc_assembly_elasticity(int model, double Emod, double nu, const double* pts, int64_t npts,
                              const int32_t* elems, int64_t E, const double* u, double* res) {
  Crs* A = new Crs(3 * npts);
  double s[900], pv[30], el[900];
  for (int64_t T = 0; T < E; ++T) {
    const int32_t* nd = elems + 10 * T;
    double xl[30], ul[30];
    for (int k = 0; k < 10; ++k)
      for (int d = 0; d < 3; ++d) {
        xl[3 * k + d] = pts[3 * (int64_t)nd[k] + d];
        ul[3 * k + d] = u ? u[3 * (int64_t)nd[k] + d] : 0.;
      }
    orc_nh_skr(model, Emod, nu, xl, ul, s, pv);
    for (int i = 0; i < 30; ++i)
      for (int j = 0; j < 30; ++j) el[i * 30 + j] = -s[30 * j + i];  // :119
    elem_scale(el, 30, 0, 30, 0, 30);
    add_fe_block(*A, el, 30, nd, 10, 3);
    if (res && model != 0)
      for (int i = 0; i < 10; ++i)
        for (int d = 0; d < 3; ++d) res[3 * (int64_t)nd[i] + d] += pv[3 * i + d];
  }
  A->fill_complete();
  return A;
}

// ---------------------------------------------------------------------------
// Follower pressure on the 6-node surface triangle: the load FE::assemblySurfaceIntegral-
// External / assemblyNonlinearSurfaceIntegralExternal take from AceGen's
// PressureTriangle3D6 (FE_def.hpp:6905-6917, 7000-7006; the AceGenInterface library is not
// part of the reference).  Restated from the published formulation of a pressure acting
// against the outward normal of the current surface (FE_def.hpp:6928-6929): with
// x = X + u and the 6-node basis N_a(xi, eta) of the reference triangle,
//     f_a = -p int N_a (x_xi x x_eta) dxi deta          (added to the load vector)
//     Kext_(a i)(d j) = -d f_a,i / d u_d,j              (added to the tangent, A + Kext,
//                                                        NonLinElasAssFE_def.hpp:295-300)
// integrated here with Dunavant's 7-point rule (degree 5; the integrand has degree 4).
// ---------------------------------------------------------------------------
namespace {
void tri6_basis(double xi, double eta, double N[6], double dX[6], double dE[6]) {
  const double l = 1.0 - xi - eta;
  N[0] = l * (2.0 * l - 1.0);
  N[1] = xi * (2.0 * xi - 1.0);
  N[2] = eta * (2.0 * eta - 1.0);
  N[3] = 4.0 * xi * l;
  N[4] = 4.0 * xi * eta;
  N[5] = 4.0 * eta * l;
  dX[0] = 1.0 - 4.0 * l;
  dE[0] = 1.0 - 4.0 * l;
  dX[1] = 4.0 * xi - 1.0;
  dE[1] = 0.0;
  dX[2] = 0.0;
  dE[2] = 4.0 * eta - 1.0;
  dX[3] = 4.0 * (l - xi);
  dE[3] = -4.0 * xi;
  dX[4] = 4.0 * eta;
  dE[4] = 4.0 * xi;
  dX[5] = -4.0 * eta;
  dE[5] = 4.0 * (l - eta);
}

void cross3(const double a[3], const double b[3], double c[3]) {
  c[0] = a[1] * b[2] - a[2] * b[1];
  c[1] = a[2] * b[0] - a[0] * b[2];
  c[2] = a[0] * b[1] - a[1] * b[0];
}

// Face-local node lists of MeshStructured::buildSurfaces, P2 (MeshStructured_def.hpp:3746-3751)
const int kSurfP2[4][6] = {{1, 0, 2, 4, 6, 5}, {0, 3, 2, 7, 9, 6}, {1, 2, 3, 5, 9, 8}, {1, 0, 3, 4, 7, 8}};
}  // namespace

void orc_pressure_triangle(double p, const double* X, const double* u, double* f, double* K) {
  // Dunavant's degree-5 rule on the unit triangle in closed form (weights sum to 1/2):
  // centroid 9/40; orbits b = (6 -+ sqrt 15)/21, a = 1 - 2b, weights (155 -+ sqrt 15)/1200
  const double r15 = std::sqrt(15.0);
  const double b2 = (6.0 - r15) / 21.0, a2 = 1.0 - 2.0 * b2, w2 = (155.0 - r15) / 2400.0;
  const double b1 = (6.0 + r15) / 21.0, a1 = 1.0 - 2.0 * b1, w1 = (155.0 + r15) / 2400.0;
  const double w0 = 9.0 / 80.0;
  const double qp[7][3] = {{1.0 / 3, 1.0 / 3, w0}, {a1, b1, w1}, {b1, a1, w1}, {b1, b1, w1},
                           {a2, b2, w2}, {b2, a2, w2}, {b2, b2, w2}};
  for (int k = 0; k < 18; ++k) f[k] = 0.0;
  if (K)
    for (int k = 0; k < 324; ++k) K[k] = 0.0;
  for (auto& q : qp) {
    double N[6], dX[6], dE[6];
    tri6_basis(q[0], q[1], N, dX, dE);
    double xx[3] = {0, 0, 0}, xe[3] = {0, 0, 0};
    for (int b = 0; b < 6; ++b)
      for (int i = 0; i < 3; ++i) {
        const double xb = X[3 * b + i] + u[3 * b + i];
        xx[i] += dX[b] * xb;
        xe[i] += dE[b] * xb;
      }
    double nrm[3];
    cross3(xx, xe, nrm);
    for (int a = 0; a < 6; ++a)
      for (int i = 0; i < 3; ++i) f[3 * a + i] -= p * q[2] * N[a] * nrm[i];
    if (!K) continue;
    for (int d = 0; d < 6; ++d)
      for (int j = 0; j < 3; ++j) {
        double ej[3] = {0, 0, 0};
        ej[j] = 1.0;
        double t1[3], t2[3];
        cross3(ej, xe, t1);  // d(x_xi)/du_dj = dX[d] e_j
        cross3(xx, ej, t2);  // d(x_eta)/du_dj = dE[d] e_j
        for (int a = 0; a < 6; ++a)
          for (int i = 0; i < 3; ++i)
            K[(3 * a + i) * 18 + 3 * d + j] += p * q[2] * N[a] * (dX[d] * t1[i] + dE[d] * t2[i]);
      }
  }
}

// MeshStructured::buildSurfaces(3, "P2") (MeshStructured_def.hpp:3707-3852) for the box
// [x0, x0 + L] x [y0, y0 + W] x [z0, z0 + H] (box = {x0, y0, z0, L, W, H}): every face of
// every element whose midpoint lies on a side gets that side's flag (x=0: 1, y=0: 2, z=0: 3,
// z=1: 4, y=1: 5, x=1: 6, the last test that matches wins, as there), oriented so that
// (x_1 - x_0) x (x_2 - x_0) points out (flipSurface, :3857-3876).  Returns the count; faces
// (6 ids each) and flags are written when non-null.
int64_t orc_structured_surfaces(const double* pts, const int32_t* elems, int64_t E, const double* box,
                                int32_t* faces, int32_t* flags, int32_t* elem_of) {
  const double tol = 1e-12;
  int64_t nf = 0;
  for (int64_t T = 0; T < E; ++T) {
    const int32_t* nl = elems + 10 * T;
    for (int s = 0; s < 4; ++s) {
      int32_t v[6];
      for (int k = 0; k < 6; ++k) v[k] = nl[kSurfP2[s][k]];
      auto P = [&](int k, int d) { return pts[3 * (int64_t)v[k] + d]; };
      auto normal = [&](double vE[3]) {
        const double p1[3] = {P(0, 0) - P(1, 0), P(0, 1) - P(1, 1), P(0, 2) - P(1, 2)};
        const double p2[3] = {P(0, 0) - P(2, 0), P(0, 1) - P(2, 1), P(0, 2) - P(2, 2)};
        cross3(p1, p2, vE);
      };
      auto flip = [&]() {  // (0, 2, 1, 5, 4, 3)
        std::swap(v[1], v[2]);
        std::swap(v[3], v[5]);
      };
      double vE[3];
      normal(vE);
      double mid[3];
      for (int d = 0; d < 3; ++d) mid[d] = (P(0, d) + P(1, d) + P(2, d)) / 3.0;
      int flag = 10;
      if (mid[0] < box[0] + tol) {
        flag = 1;
        if (vE[0] > 0) flip();
      }
      if (mid[0] > box[0] + tol && mid[2] < box[2] + tol) {
        flag = 3;
        if (vE[2] > 0) flip();
      }
      if (mid[0] > box[0] + tol && mid[2] > box[2] + box[5] - tol) {
        if (vE[2] < 0) flip();
        flag = 4;
      }
      if (mid[0] > box[0] + tol && mid[1] < box[1] + tol) {
        if (vE[1] > 0) flip();
        flag = 2;
      }
      if (mid[0] > box[0] + tol && mid[1] > box[1] + box[4] - tol) {
        if (vE[1] < 0) flip();
        flag = 5;
      }
      if (mid[0] > box[0] + box[3] - tol && mid[1] > box[1] + tol && mid[1] < box[1] + box[4] - tol &&
          mid[2] > box[2] + tol && mid[2] < box[2] + box[5] - tol) {
        if (vE[0] < 0) flip();
        flag = 6;
      }
      if (flag == 10) continue;
      if (faces)
        for (int k = 0; k < 6; ++k) faces[6 * nf + k] = v[k];
      if (flags) flags[nf] = flag;
      if (elem_of) elem_of[nf] = (int32_t)T;
      ++nf;
    }
  }
  return nf;
}

// The surface loop of assemblySurfaceIntegralExternal / assemblyNonlinearSurfaceIntegral-
// External (FE_def.hpp:6859-6925, 6932-7045): f = 0, then for every face with
// p = pressure[flag] != 0 the element load is added to f and, when Kext is requested, its
// 18 x 18 matrix is inserted row by row (Kext->insertGlobalValues, :7032-7034).
void* orc_assembly_surface_pressure(const double* pts, int64_t npts, const int32_t* faces, const int32_t* flags,
                                    int64_t nf, const double* u, const double* pressure, int nflag, double* f,
                                    int with_tangent) {
  for (int64_t k = 0; k < 3 * npts; ++k) f[k] = 0.0;
  Crs* Kx = with_tangent ? new Crs(3 * npts) : nullptr;
  double X[18], U[18], fe[18], Ke[324];
  for (int64_t s = 0; s < nf; ++s) {
    const int fl = flags[s];
    const double p = (fl >= 0 && fl < nflag) ? pressure[fl] : 0.0;
    if (p == 0.0) continue;
    const int32_t* v = faces + 6 * s;
    for (int a = 0; a < 6; ++a)
      for (int i = 0; i < 3; ++i) {
        X[3 * a + i] = pts[3 * (int64_t)v[a] + i];
        U[3 * a + i] = u ? u[3 * (int64_t)v[a] + i] : 0.0;
      }
    orc_pressure_triangle(p, X, U, fe, Kx ? Ke : nullptr);
    for (int a = 0; a < 6; ++a)
      for (int i = 0; i < 3; ++i) f[3 * (int64_t)v[a] + i] += fe[3 * a + i];
    if (Kx) {
      int32_t cols[18];
      for (int b = 0; b < 6; ++b)
        for (int d = 0; d < 3; ++d) cols[3 * b + d] = 3 * v[b] + d;
      for (int a = 0; a < 6; ++a)
        for (int i = 0; i < 3; ++i) Kx->insert(3 * (int64_t)v[a] + i, cols, Ke + (3 * a + i) * 18, 18);
    }
  }
  if (Kx) Kx->fill_complete();
  return Kx;
}

}  // extern "C"
