// Tests of the C++ FE facade (feddlib_amd/host/FE_gpu.hpp), written like FEDDLib's
// fe_test_* programs: build a structured cube, add the domains to an FE object, call
// the FE::assembly* methods with the reference's arguments, then check every result
// against the oracle (oracle/oracle.h, the CPU restatement of the reference loops).
//
//   fe_gpu_test          all cases (needs a GPU)
//   fe_gpu_test --cpu    only the error-behaviour cases that raise before any device
//                        work (checkFE, argument checks, the element factory)
//
// Exit status 0 when every case passes; each case prints one line.
#include "FE_gpu.hpp"
#include "oracle.h"

#include <cmath>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

using namespace feddlib_amd;

static int g_fail = 0;

static void report(const char* name, bool ok, const std::string& detail = "") {
  std::printf("%s %s%s%s\n", ok ? "PASS" : "FAIL", name, detail.empty() ? "" : ": ", detail.c_str());
  std::fflush(stdout);
  if (!ok) ++g_fail;
}

// The oracle's CRS handle as a CrsMatrix.
static CrsMatrix fromOracle(void* h) {
  CrsMatrix m;
  m.numRows = orc_crs_nrows(h);
  const int64_t nnz = orc_crs_nnz(h);
  m.rowptr.resize((size_t)m.numRows + 1);
  m.colind.resize((size_t)nnz);
  m.values.resize((size_t)nnz);
  orc_crs_export(h, m.rowptr.data(), m.colind.data(), m.values.data());
  orc_crs_destroy(h);
  return m;
}

// Same graph, and |a - b| <= rtol max(|b|, 1e-3 rowmax) + atol_row rowmax per entry.
static bool sameMatrix(const CrsMatrix& a, const CrsMatrix& b, double rtol, double atol_row, std::string& why) {
  if (a.rowptr != b.rowptr || a.colind != b.colind) {
    why = "graph differs";
    return false;
  }
  for (int64_t r = 0; r + 1 < (int64_t)b.rowptr.size(); ++r) {
    double rowmax = 0.0;
    for (int64_t k = b.rowptr[r]; k < b.rowptr[r + 1]; ++k) rowmax = std::fmax(rowmax, std::fabs(b.values[k]));
    for (int64_t k = b.rowptr[r]; k < b.rowptr[r + 1]; ++k) {
      const double tol = rtol * std::fmax(std::fabs(b.values[k]), 1e-3 * rowmax) + atol_row * rowmax;
      if (!(std::fabs(a.values[k] - b.values[k]) <= tol)) {
        char buf[160];
        std::snprintf(buf, sizeof buf, "row %lld entry %lld: %.17g vs %.17g", (long long)r, (long long)k,
                      a.values[k], b.values[k]);
        why = buf;
        return false;
      }
    }
  }
  return true;
}

static double maxRelDiff(const std::vector<double>& a, const std::vector<double>& b) {
  double d = 0.0, m = 0.0;
  for (size_t i = 0; i < b.size(); ++i) {
    d = std::fmax(d, std::fabs(a[i] - b[i]));
    m = std::fmax(m, std::fabs(b[i]));
  }
  return m > 0 ? d / m : d;
}

template <class E, class F>
static bool throws(F f, const char* msg = nullptr) {
  try {
    f();
  } catch (const E& e) {
    return !msg || std::strstr(e.what(), msg) != nullptr;
  } catch (...) {
    return false;
  }
  return false;
}

static void oneFunc(double* x, double* res, double* parameters) {
  (void)x;
  (void)parameters;
  res[0] = 1.;
}
// rhsX3D of linElasUnitTest.cpp: parameters [t, force, degree] -> (force, 0, 0).
static void rhsX3D(double* x, double* res, double* parameters) {
  (void)x;
  res[0] = parameters[1];
  res[1] = 0.;
  res[2] = 0.;
}
static void rampFunc(double* x, double* res, double* parameters) {
  (void)parameters;
  res[0] = 1.0 + x[0] * x[0] - 0.5 * x[1] * x[2];
}

// ------------------------------------------------------------ CPU cases
static void errorCases() {
  FE fe;
  MatrixPtr_Type A;
  report("checkFE without domain raises logic_error",
         throws<std::logic_error>([&] { fe.checkFE(3, "P1"); }, "Use addFE(domain)"));
  report("assemblyLaplace P0 raises logic_error",
         throws<std::logic_error>([&] { fe.assemblyLaplace(3, "P0", 2, A); }, "Not implemented for P0"));
  std::vector<double> prm = {1.0};
  report("assemblyRHS null vector raises runtime_error",
         throws<std::runtime_error>([&] { fe.assemblyRHS(3, "P1", nullptr, "Scalar", oneFunc, prm); },
                                    "MultiVector in assemblyConstRHS is null."));
  report("AssembleFEFactory unknown type raises logic_error",
         throws<std::logic_error>([] { AssembleFEFactory::build("Poroelasticity"); },
                                  "No specific implementation for your request."));
  report("AssembleFEFactory knows NavierStokes", AssembleFEFactory::build("NavierStokes") == "NavierStokes");
  ParameterList pl;
  pl.sublist("Parameter").set("Viscosity", 0.5);
  report("ParameterList get default and stored value",
         pl.sublist("Parameter").get("Viscosity", 1.) == 0.5 && pl.sublist("Parameter").get("Density", 1.) == 1.0);
  report("ParameterList type mismatch raises logic_error",
         throws<std::logic_error>([&] { pl.sublist("Parameter").get("Viscosity", 1); }));
}

// ------------------------------------------------------------ GPU cases
static void laplaceCases() {
  for (int fe_ = 1; fe_ <= 2; ++fe_) {
    const std::string T = fe_ == 1 ? "P1" : "P2";
    const int M = fe_ == 1 ? 4 : 3;
    auto dom = Domain::buildStructuredCube(M, T, false);
    FE fe;
    fe.addFE(dom);
    MatrixPtr_Type A;
    fe.assemblyLaplace(3, T, 2, A);
    CrsMatrix ref = fromOracle(orc_assembly_laplace(fe_, dom->points.data(), dom->getNumPoints(),
                                                    dom->elements.data(), dom->getNumElements()));
    std::string why;
    report(("assemblyLaplace " + T).c_str(), A->fillCompleted && sameMatrix(*A, ref, 1e-12, 1e-14, why), why);

    auto b = std::make_shared<MultiVector>(dom->getNumPoints());
    std::vector<double> prm = {1.0};
    fe.assemblyRHS(3, T, b, "Scalar", oneFunc, prm);
    fe.assemblyRHS(3, T, b, "Scalar", oneFunc, prm);  // sums into b, as valuesRhs[row] += value
    std::vector<double> rb((size_t)dom->getNumPoints(), 0.0);
    orc_assembly_rhs(fe_, dom->points.data(), dom->getNumPoints(), dom->elements.data(), dom->getNumElements(), 1.0,
                     rb.data());
    for (double& v : rb) v *= 2.0;
    const double d = maxRelDiff(b->vector(), rb);
    report(("assemblyRHS Scalar accumulates " + T).c_str(), d <= 1e-13, std::to_string(d));

    BlockMatrixPtr_Type AB;
    FE fe2;
    fe2.addFE(dom);
    fe2.assemblyLaplaceAssFE(3, T, 2, 1, AB, true);
    CrsMatrix refa = fromOracle(orc_assembly_laplace_assfe(fe_, 1, dom->points.data(), dom->getNumPoints(),
                                                           dom->elements.data(), dom->getNumElements()));
    report(("assemblyLaplaceAssFE " + T).c_str(), sameMatrix(*AB->getBlock(0, 0), refa, 1e-12, 1e-14, why), why);
  }
  // vector source of linElasUnitTest.cpp on P2: (force, 0, 0) per node
  auto dom = Domain::buildStructuredCube(2, "P2", false);
  FE fe;
  fe.addFE(dom);
  const int64_t n = dom->getNumPoints();
  auto b = std::make_shared<MultiVector>(3 * n);
  std::vector<double> prm = {0.0, 0.5, 2.0};
  fe.assemblyRHS(3, "P2", b, "Vector", rhsX3D, prm);
  std::vector<double> rb((size_t)n, 0.0), want((size_t)(3 * n), 0.0);
  orc_assembly_rhs(2, dom->points.data(), n, dom->elements.data(), dom->getNumElements(), 1.0, rb.data());
  for (int64_t i = 0; i < n; ++i) want[3 * i] = 0.5 * rb[i];
  const double d = maxRelDiff(b->vector(), want);
  report("assemblyRHS Vector rhsX3D", d <= 1e-13, std::to_string(d));
  // a source varying in space, against the oracle's restated loop at its own points
  for (int fe_ = 1; fe_ <= 2; ++fe_) {
    const std::string T = fe_ == 1 ? "P1" : "P2";
    auto dv = Domain::buildStructuredCube(3, T, false);
    FE fv;
    fv.addFE(dv);
    const int64_t nv = dv->getNumPoints(), Ev = dv->getNumElements();
    auto bs = std::make_shared<MultiVector>(nv);
    fv.assemblyRHS(3, T, bs, "Scalar", rampFunc, prm);
    std::vector<double> xq((size_t)(Ev * 5 * 3)), fq((size_t)(Ev * 5)), rr((size_t)nv, 0.0);
    const int nq = orc_rhs_quadrature_points(fe_, dv->points.data(), dv->elements.data(), Ev, xq.data());
    for (int64_t k = 0; k < Ev * nq; ++k) rampFunc(&xq[3 * k], &fq[k], nullptr);
    orc_assembly_rhs_points(fe_, dv->points.data(), nv, dv->elements.data(), Ev, fq.data(), 1, rr.data());
    const double dd = maxRelDiff(bs->vector(), rr);
    report(("assemblyRHS varying source " + T).c_str(), dd <= 1e-13, std::to_string(dd));
  }
  report("assemblyRHS invalid field type raises logic_error",
         throws<std::logic_error>([&] { fe.assemblyRHS(3, "P2", b, "Tensor", rhsX3D, prm); }, "Invalid field type."));
}

static void navierStokesCases() {
  const int M = 2;
  auto d2 = Domain::buildStructuredCube(M, "P2", false);
  auto d1 = Domain::buildStructuredCube(M, "P1", false);
  const int64_t n2 = d2->getNumPoints(), n1 = d1->getNumPoints(), E = d2->getNumElements();
  auto u = std::make_shared<MultiVector>(3 * n2);
  auto p = std::make_shared<MultiVector>(n1);
  for (int64_t i = 0; i < n2; ++i) {
    const double* x = &d2->points[3 * i];
    u->getDataNonConst()[3 * i + 0] = std::sin(2.0 * x[0]) + 0.3 * x[2] * x[1];  // div u != 0: B u != 0
    u->getDataNonConst()[3 * i + 1] = std::cos(1.5 * x[0]) * x[2];
    u->getDataNonConst()[3 * i + 2] = 0.2 - x[0] * x[1];
  }
  for (int64_t i = 0; i < n1; ++i) p->getDataNonConst()[i] = d1->points[3 * i] - 0.5 * d1->points[3 * i + 2];
  const double nu = 0.01, rho = 1.0;
  auto params = std::make_shared<ParameterList>();
  params->sublist("Parameter").set("Viscosity", nu);
  params->sublist("Parameter").set("Density", rho);
  params->sublist("General").set("Linearization", "Newton");
  SmallMatrix_Type coeff(2);
  coeff[0][0] = 1.0;
  coeff[0][1] = 1.0;
  coeff[1][0] = 0.5;
  coeff[1][1] = 0.0;
  const double c4[4] = {coeff[0][0], coeff[0][1], coeff[1][0], coeff[1][1]};

  FE fe;
  fe.addFE(d2);
  fe.addFE(d1);
  BlockMatrixPtr_Type A;
  BlockMultiVectorPtr_Type res;
  fe.assemblyNavierStokes(3, "P2", "P1", 2, 3, 1, u, p, A, res, coeff, params, false, "Jacobian");
  void *hA = nullptr, *hBT = nullptr, *hB = nullptr, *hC = nullptr;
  orc_assembly_ns(d2->points.data(), n2, d2->elements.data(), d1->elements.data(), E, n1, u->getData(), p->getData(),
                  nu, rho, 1, 0, 0, c4, &hA, &hBT, &hB, &hC, nullptr, nullptr);
  CrsMatrix rA = fromOracle(hA), rBT = fromOracle(hBT), rB = fromOracle(hB), rC = fromOracle(hC);
  std::string why;
  bool ok = sameMatrix(*A->getBlock(0, 0), rA, 1e-12, 1e-14, why) && sameMatrix(*A->getBlock(0, 1), rBT, 1e-12, 1e-14, why) &&
            sameMatrix(*A->getBlock(1, 0), rB, 1e-12, 1e-14, why) && sameMatrix(*A->getBlock(1, 1), rC, 1e-12, 1e-14, why);
  report("assemblyNavierStokes Jacobian, four blocks", ok, why);
  report("advanceNewtonStep after Jacobian", fe.getNewtonStep() == 1);

  BlockMatrixPtr_Type AF;
  fe.assemblyNavierStokes(3, "P2", "P1", 2, 3, 1, u, p, AF, res, coeff, params, true, "FixedPoint");
  orc_assembly_ns(d2->points.data(), n2, d2->elements.data(), d1->elements.data(), E, n1, u->getData(), p->getData(),
                  nu, rho, 1, 1, 1, c4, &hA, &hBT, &hB, &hC, nullptr, nullptr);
  CrsMatrix rF = fromOracle(hA);
  report("assemblyNavierStokes FixedPoint, reAssemble: A(0,0) only",
         sameMatrix(*AF->getBlock(0, 0), rF, 1e-12, 1e-14, why) && !AF->blockExists(0, 1), why);

  BlockMultiVectorPtr_Type rv;
  BlockMatrixPtr_Type none;
  fe.assemblyNavierStokes(3, "P2", "P1", 2, 3, 1, u, p, none, rv, coeff, params, false, "Rhs");
  std::vector<double> ru((size_t)(3 * n2)), rp((size_t)n1);
  orc_assembly_ns(d2->points.data(), n2, d2->elements.data(), d1->elements.data(), E, n1, u->getData(), p->getData(),
                  nu, rho, 1, 2, 0, c4, &hA, &hBT, &hB, &hC, ru.data(), rp.data());
  const double du = maxRelDiff(rv->getBlock(0)->vector(), ru), dp = maxRelDiff(rv->getBlock(1)->vector(), rp);
  report("assemblyNavierStokes Rhs with coeff", du <= 1e-12 && dp <= 1e-12,
         std::to_string(du) + " " + std::to_string(dp));

  params->sublist("Material").set("Newtonian", false);
  report("GeneralizedNewtonian raises logic_error",
         throws<std::logic_error>([&] { fe.assemblyNavierStokes(3, "P2", "P1", 2, 3, 1, u, p, A, res, coeff, params,
                                                                false, "Jacobian"); }));
}

static void elasticityCases() {
  auto dom = Domain::buildStructuredCube(2, "P2", true);
  const int64_t n = dom->getNumPoints(), E = dom->getNumElements();
  auto d = std::make_shared<MultiVector>(3 * n);
  for (int64_t i = 0; i < n; ++i)
    for (int k = 0; k < 3; ++k)
      d->getDataNonConst()[3 * i + k] = 0.02 * std::sin(3.0 * dom->points[3 * i + k] + k) * dom->points[3 * i + 2];
  const double Emod = 1000.0, nu = 0.3;
  std::string why;
  for (int model = 1; model <= 2; ++model) {
    auto params = std::make_shared<ParameterList>();
    params->sublist("Parameter").set("E", Emod);
    params->sublist("Parameter").set("Poisson Ratio", nu);
    params->sublist("Parameter").set("Neo-Hooke Modell", model);
    FE fe;
    fe.addFE(dom);
    BlockMatrixPtr_Type A;
    BlockMultiVectorPtr_Type res;
    fe.assemblyNonLinearElasticity(3, "P2", 2, 3, d, A, res, params);
    std::vector<double> rr((size_t)(3 * n), 0.0);
    CrsMatrix ref = fromOracle(orc_assembly_elasticity(model, Emod, nu, dom->points.data(), n, dom->elements.data(), E,
                                                       d->getData(), rr.data()));
    const double dr = maxRelDiff(res->getBlock(0)->vector(), rr);
    const std::string name = std::string("assemblyNonLinearElasticity Neo-Hooke Modell ") + std::to_string(model);
    report(name.c_str(), sameMatrix(*A->getBlock(0, 0), ref, 1e-12, 1e-14, why) && dr <= 1e-12,
           why + " res " + std::to_string(dr));

    // per-element E equal to the constant: the same tangent
    auto eMod = std::make_shared<MultiVector>(E);
    eMod->putScalar(Emod);
    BlockMatrixPtr_Type AE;
    BlockMultiVectorPtr_Type resE;
    fe.assemblyNonLinearElasticity(3, "P2", 2, 3, d, AE, resE, params, dom, eMod);
    report((name + ", E field").c_str(), sameMatrix(*AE->getBlock(0, 0), ref, 1e-12, 1e-14, why), why);
  }
  {
    auto params = std::make_shared<ParameterList>();
    params->sublist("Parameter").set("E", Emod);
    params->sublist("Parameter").set("Poisson Ratio", nu);
    FE fe;
    fe.addFE(dom);
    BlockMatrixPtr_Type A;
    BlockMultiVectorPtr_Type res;
    fe.assemblyLinearElasticity(3, "P2", 2, 3, d, A, res, params, false, "Jacobian");
    CrsMatrix ref = fromOracle(orc_assembly_elasticity(0, Emod, nu, dom->points.data(), n, dom->elements.data(), E,
                                                       d->getData(), nullptr));
    report("assemblyLinearElasticity (skr1)", sameMatrix(*A->getBlock(0, 0), ref, 1e-12, 1e-14, why), why);

    // at zero displacement every Neo-Hookean tangent is assemblyLinElasXDim with the same E, nu
    auto zero = std::make_shared<MultiVector>(3 * n);
    BlockMatrixPtr_Type A0;
    fe.assemblyLinearElasticity(3, "P2", 2, 3, zero, A0, res, params, false, "Jacobian");
    const double mu = Emod / (2.0 * (1.0 + nu)), lam = nu * Emod / ((1.0 + nu) * (1.0 - 2.0 * nu));
    FE fx;
    fx.addFE(dom);
    MatrixPtr_Type L;
    fx.assemblyLinElasXDim(3, "P2", L, lam, mu);
    report("assemblyLinElasXDim equals the tangent at u = 0", sameMatrix(*L, *A0->getBlock(0, 0), 1e-11, 1e-13, why),
           why);
  }
  {
    // "Number Elements not the same as number assembleFE elements." (FE_def.hpp:1140-1141)
    auto p1a = Domain::buildStructuredCube(2, "P1", false);
    auto p1b = Domain::buildStructuredCube(3, "P1", false);
    FE fe;
    fe.addFE(p1a);
    BlockMatrixPtr_Type A;
    fe.assemblyLaplaceAssFE(3, "P1", 2, 1, A, true);
    fe.addFE(p1b);  // checkFE now resolves P1 to the larger domain
    report("element count change raises logic_error",
           throws<std::logic_error>([&] { fe.assemblyLaplaceAssFE(3, "P1", 2, 1, A, true); },
                                    "Number Elements not the same as number assembleFE elements."));
  }
}

int main(int argc, char** argv) {
  const bool cpu = argc > 1 && std::strcmp(argv[1], "--cpu") == 0;
  errorCases();
  if (!cpu) {
    try {
      laplaceCases();
      navierStokesCases();
      elasticityCases();
    } catch (const std::exception& e) {
      report("unexpected exception", false, e.what());
    }
  }
  std::printf("%s: %d failed\n", g_fail ? "FAILED" : "OK", g_fail);
  return g_fail ? 1 : 0;
}
