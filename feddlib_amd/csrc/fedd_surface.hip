// Follower pressure load on the P2 surface triangles, for gfx950.
//
// Reference: FE::assemblySurfaceIntegralExternal and
// FE::assemblyNonlinearSurfaceIntegralExternal (FE_def.hpp:6859-6925, 6932-7045) loop over
// the flagged surface triangles of each element, take the pressure p = func(flag), and add
// the load and tangent of AceGen's PressureTriangle3D6 element (a pressure acting against
// the outward normal of the deformed surface, :6928-6929) to the load vector and, for the
// nonlinear variant, to Kext, which NonLinElasAssFE adds to the elasticity tangent
// (A + Kext, NonLinElasAssFE_def.hpp:284-303).  The surfaces of the structured cube are
// MeshStructured::buildSurfaces(3, "P2") (MeshStructured_def.hpp:3707-3876).
//
// Formulation (exact integration, degree 4 in the reference coordinates): with x = X + u,
// y_b = x_b - x_0 and the tensor PT[a][b][c] = int N_a (dxi N_b deta N_c - dxi N_c deta N_b)
// over the reference triangle (fedd_tensors.h),
//     f_a = -p int N_a (x_xi x x_eta) = -p sum_{1<=b<c<=5} PT[a][b][c] (y_b x y_c),
//     Kext_(a i)(d j) = -d f_a,i / d u_d,j = p eps_ijl W_ad,l,  W_ad = sum_c PT[a][d][c] y_c
// (the translation invariance sum_b PT[a][b][c] = 0 lets y replace x exactly).  One thread
// per face; the 18 load entries and the 36 skew 3x3 blocks of Kext are added with
// atomics (a face touches 6 rows of 6 columns, all inside the volume graph of its tet).
#include "fedd_internal.h"
#include "fedd_tensors.h"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

namespace fedd {

namespace {
inline unsigned nblk(int64_t n, int b) { return (unsigned)((n + b - 1) / b); }

namespace ft = fedd_tensors;

struct SurfArgs {
  const double* xyz;  // SoA coordinates
  int64_t npts;
  const double* u;  // displacement, 3 per node (null: reference configuration)
  const int32_t* faces;
  const int32_t* flags;
  int64_t nface;
  double ptab[FEDD_MAX_SURFACE_FLAGS];
  int nflag;
  double* f;  // load, 3 per node (added to)
  // tangent (null vals: load only): node-level graph of the vector block A
  const int64_t* nptr;
  const int32_t* nidx;
  double* vals;
};

__global__ __launch_bounds__(64) void k_surface_pressure(SurfArgs a) {
  const int64_t s = (int64_t)blockIdx.x * 64 + threadIdx.x;
  if (s >= a.nface) return;
  const int fl = a.flags[s];
  const double p = (fl >= 0 && fl < a.nflag) ? a.ptab[fl] : 0.0;
  if (p == 0.0) return;  // FE_def.hpp:6902, 6997: faces without pressure are skipped
  int32_t v[6];
  double x[6][3];
#pragma unroll
  for (int k = 0; k < 6; ++k) {
    v[k] = a.faces[6 * s + k];
#pragma unroll
    for (int i = 0; i < 3; ++i) x[k][i] = a.xyz[i * a.npts + v[k]] + (a.u ? a.u[3 * (int64_t)v[k] + i] : 0.0);
  }
  double y[6][3];
#pragma unroll
  for (int k = 0; k < 6; ++k)
#pragma unroll
    for (int i = 0; i < 3; ++i) y[k][i] = x[k][i] - x[0][i];
  double C[6][6][3];  // y_b x y_c, b < c, b >= 1
#pragma unroll
  for (int b = 1; b < 6; ++b)
#pragma unroll
    for (int c = b + 1; c < 6; ++c) {
      C[b][c][0] = y[b][1] * y[c][2] - y[b][2] * y[c][1];
      C[b][c][1] = y[b][2] * y[c][0] - y[b][0] * y[c][2];
      C[b][c][2] = y[b][0] * y[c][1] - y[b][1] * y[c][0];
    }
#pragma unroll
  for (int r = 0; r < 6; ++r) {
    double fr[3] = {0.0, 0.0, 0.0};
#pragma unroll
    for (int b = 1; b < 6; ++b)
#pragma unroll
      for (int c = b + 1; c < 6; ++c) {
        const double t = ft::PT[r][b][c];  // compile-time after unrolling
        if (t != 0.0)
#pragma unroll
          for (int i = 0; i < 3; ++i) fr[i] += t * C[b][c][i];
      }
#pragma unroll
    for (int i = 0; i < 3; ++i) atomicAdd(a.f + 3 * (int64_t)v[r] + i, -p * fr[i]);
  }
  if (!a.vals) return;
#pragma unroll
  for (int r = 0; r < 6; ++r) {
    const int64_t base = a.nptr[v[r]];
    const int deg = (int)(a.nptr[v[r] + 1] - base);
    const int32_t* cols = a.nidx + base;
    double* row = a.vals + 9 * base;
#pragma unroll
    for (int d = 0; d < 6; ++d) {
      double w[3] = {0.0, 0.0, 0.0};
#pragma unroll
      for (int c = 1; c < 6; ++c) {
        const double t = ft::PT[r][d][c];
        if (t != 0.0)
#pragma unroll
          for (int i = 0; i < 3; ++i) w[i] += t * y[c][i];
      }
      // column slot of node v[d] in row v[r] (sorted node columns)
      int lo = 0, hi = deg - 1;
      while (lo < hi) {
        const int mid = (lo + hi) >> 1;
        if (cols[mid] < v[d])
          lo = mid + 1;
        else
          hi = mid;
      }
      double* blk = row + 3 * lo;  // entry (i, j) at blk[i * 3 deg + j]
      const int ld = 3 * deg;
      atomicAdd(blk + 0 * ld + 1, p * w[2]);
      atomicAdd(blk + 0 * ld + 2, -p * w[1]);
      atomicAdd(blk + 1 * ld + 0, -p * w[2]);
      atomicAdd(blk + 1 * ld + 2, p * w[0]);
      atomicAdd(blk + 2 * ld + 0, p * w[1]);
      atomicAdd(blk + 2 * ld + 1, -p * w[0]);
    }
  }
}

__global__ void k_zero(int64_t n, double* x) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) x[i] = 0.0;
}
}  // namespace

// Load (into vector `out`, zeroed first unless accumulate) and, with_tangent, Kext added to
// block A, for pressures scale * pressure[flag].
int surface_pressure(Ctx* c, const double* pressure, int nflag, double scale, int with_tangent, int out,
                     bool accumulate) {
  const Mesh& m = c->mesh[0];
  const int64_t n = 3 * m.npts;
  if (c->vec_n[out] != n) {
    int rc = alloc(c, c->vec[out], n * 8);
    if (rc) return rc;
    c->vec_n[out] = n;
    accumulate = false;
  }
  SurfArgs a{};
  a.xyz = m.d_xyz.as<double>();
  a.npts = m.npts;
  a.u = c->vec_n[FEDD_VEC_U] == n ? c->vec[FEDD_VEC_U].as<double>() : nullptr;
  a.faces = c->d_faces.as<int32_t>();
  a.flags = c->d_fflags.as<int32_t>();
  a.nface = c->nface;
  a.nflag = nflag;
  for (int k = 0; k < nflag; ++k) a.ptab[k] = scale * pressure[k];
  a.f = c->vec[out].as<double>();
  if (with_tangent) {
    a.nptr = c->d_nptr.as<int64_t>();
    a.nidx = c->d_nidx.as<int32_t>();
    a.vals = c->blk[0].d_vals.as<double>();
  }
  if (!accumulate) k_zero<<<nblk(n, 256), 256, 0, c->stream>>>(n, a.f);  // f->putScalar(0.)
  if (c->nface > 0) k_surface_pressure<<<nblk(c->nface, 64), 64, 0, c->stream>>>(a);
  FEDD_HIP(c, hipGetLastError());
  return FEDD_OK;
}

}  // namespace fedd

using namespace fedd;

namespace {
// Face-local node lists of the four faces of a tet-10 (MeshStructured_def.hpp:3746-3751)
// and the flip of flipSurface (:3857-3876): 0 2 1 5 4 3.
const int kFace[4][6] = {{1, 0, 2, 4, 6, 5}, {0, 3, 2, 7, 9, 6}, {1, 2, 3, 5, 9, 8}, {1, 0, 3, 4, 7, 8}};
}  // namespace

extern "C" {

int fedd_structured_surfaces(const double* pts, int64_t npts, const int32_t* elems, int64_t nelem,
                             const double* box, int32_t* faces, int32_t* flags, int64_t* nface) {
  if (!pts || !elems || !box || !nface || npts <= 0 || nelem < 0) {
    free_error() = "fedd_structured_surfaces: bad arguments";
    return FEDD_ERR_ARG;
  }
  const double tol = 1e-12;
  const double lo[3] = {box[0], box[1], box[2]};
  const double hi[3] = {box[0] + box[3], box[1] + box[4], box[2] + box[5]};
  int64_t nf = 0;
  for (int64_t e = 0; e < nelem; ++e) {
    const int32_t* nl = elems + 10 * e;
    for (int s = 0; s < 4; ++s) {
      int32_t v[6];
      for (int k = 0; k < 6; ++k) {
        v[k] = nl[kFace[s][k]];
        if (v[k] < 0 || v[k] >= npts) {
          free_error() = "fedd_structured_surfaces: node id out of range";
          return FEDD_ERR_ARG;
        }
      }
      const double* p0 = pts + 3 * (int64_t)v[0];
      const double* p1 = pts + 3 * (int64_t)v[1];
      const double* p2 = pts + 3 * (int64_t)v[2];
      double mid[3];
      for (int d = 0; d < 3; ++d) mid[d] = (p0[d] + p1[d] + p2[d]) / 3.0;
      // the side the face lies on, with the reference's precedence (the last match wins)
      int flag = 0, axis = -1;
      const bool inside_x = mid[0] > lo[0] + tol;
      if (!inside_x) flag = 1, axis = 0;
      if (inside_x && mid[2] < lo[2] + tol) flag = 3, axis = 2;
      if (inside_x && mid[2] > hi[2] - tol) flag = 4, axis = 2;
      if (inside_x && mid[1] < lo[1] + tol) flag = 2, axis = 1;
      if (inside_x && mid[1] > hi[1] - tol) flag = 5, axis = 1;
      if (mid[0] > hi[0] - tol && mid[1] > lo[1] + tol && mid[1] < hi[1] - tol && mid[2] > lo[2] + tol &&
          mid[2] < hi[2] - tol)
        flag = 6, axis = 0;
      if (flag == 0) continue;
      // orient (x1 - x0) x (x2 - x0) out of the box: its component along the side's axis
      // points outward (the sign test of the reference)
      const double a1[3] = {p1[0] - p0[0], p1[1] - p0[1], p1[2] - p0[2]};
      const double a2[3] = {p2[0] - p0[0], p2[1] - p0[1], p2[2] - p0[2]};
      const double nrm[3] = {a1[1] * a2[2] - a1[2] * a2[1], a1[2] * a2[0] - a1[0] * a2[2],
                             a1[0] * a2[1] - a1[1] * a2[0]};
      const bool low_side = flag == 1 || flag == 2 || flag == 3;
      if ((low_side && nrm[axis] > 0) || (!low_side && nrm[axis] < 0)) {
        std::swap(v[1], v[2]);
        std::swap(v[3], v[5]);
      }
      if (faces) std::memcpy(faces + 6 * nf, v, sizeof(v));
      if (flags) flags[nf] = flag;
      ++nf;
    }
  }
  *nface = nf;
  return FEDD_OK;
}

int fedd_surface_set(fedd_ctx* cc, const int32_t* faces, const int32_t* flags, int64_t nface) {
  Ctx* c = cc;
  if (!c || nface < 0 || (nface > 0 && (!faces || !flags))) return FEDD_ERR_ARG;
  const Mesh& m = c->mesh[0];
  if (m.npe != 10) return fail(c, FEDD_ERR_STATE, "surface loads need a P2 (tet-10) primary mesh");
  for (int64_t s = 0; s < nface; ++s) {
    for (int k = 0; k < 6; ++k)
      if (faces[6 * s + k] < 0 || faces[6 * s + k] >= m.npts) return fail(c, FEDD_ERR_ARG, "face node out of range");
    if (flags[s] < 0) return fail(c, FEDD_ERR_ARG, "face flags must be >= 0");
  }
  if (c->capturing) return fail(c, FEDD_ERR_STATE, "host transfers cannot be captured");
  (void)hipSetDevice(c->device);
  int rc;
  if (nface > 0) {
    if ((rc = upload(c, c->d_faces, faces, (size_t)nface * 24))) return rc;
    if ((rc = upload(c, c->d_fflags, flags, (size_t)nface * 4))) return rc;
  }
  c->nface = nface;
  FEDD_HIP(c, hipStreamSynchronize(c->stream));
  return FEDD_OK;
}

int fedd_assembly_surface_pressure(fedd_ctx* cc, const double* pressure, int nflag, int with_tangent, int out) {
  Ctx* c = cc;
  if (!c || nflag < 0 || nflag > FEDD_MAX_SURFACE_FLAGS || (nflag > 0 && !pressure) || out < 0 ||
      out >= FEDD_NUM_VECTORS)
    return FEDD_ERR_ARG;
  if (c->mesh[0].npe != 10) return fail(c, FEDD_ERR_STATE, "surface loads need a P2 (tet-10) primary mesh");
  if (out != FEDD_VEC_FEXT && out != FEDD_VEC_FPRES && out != FEDD_VEC_B && out != FEDD_VEC_RES_U)
    return fail(c, FEDD_ERR_ARG, "the load goes into a vector of 3 entries per primary node");
  if (with_tangent && c->graph_kind != FEDD_GRAPH_VECTOR)
    return fail(c, FEDD_ERR_STATE, "Kext is added to the vector graph's block A (fedd_graph_build VECTOR)");
  (void)hipSetDevice(c->device);
  return surface_pressure(c, pressure, nflag, 1.0, with_tangent, out, false);
}

int fedd_newton_pressure(fedd_ctx* cc, const double* pressure, int nflag) {
  Ctx* c = cc;
  if (!c || nflag < 0 || nflag > FEDD_MAX_SURFACE_FLAGS || (nflag > 0 && !pressure)) return FEDD_ERR_ARG;
  c->newton_pressure.assign(pressure, pressure + nflag);
  return FEDD_OK;
}

}  // extern "C"
