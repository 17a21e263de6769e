// Navier-Stokes Newton solve on gfx950: the saddle-point system [A B^T; B 0] of the
// P2-P1 Jacobian solved by restarted, right-preconditioned GMRES on the device, and the
// Newton driver around the fused Jacobian + residual assembly.
//
// Reference roles:
//   NonLinearSolver::solveNewton (NonLinearSolver_def.hpp:356-414): residual
//     ("reverse": rhs - F(u), setBCMinusVector), Criterion "Residual" (||r_k|| / ||r_0||)
//     or "Update" (||delta_k||) against relNonLinTol, MaxNonLinIts;
//   NavierStokesAssFE::reAssemble / calculateNonLinResidualVec
//     (NavierStokesAssFE_def.hpp:225-313): Jacobian and residual on the current iterate;
//   Problem::setBoundariesSystem -> BCBuilder::setSystem (BCBuilder_def.hpp:590-704);
//   LinearSolver::solve with Belos GMRES + a block preconditioner (LinearSolver_def.hpp:23-137,
//     Preconditioner_def.hpp:206-291; Teko's block preconditioners for the monolithic system).
//
// Linear solve.  Unknowns [x_u; x_p], n = 3 N2 + N1, stored contiguously.  Right
// preconditioning, so GMRES minimises the true residual of J x = b:
//   P = [A~ B^T; 0 -S~],  A~ = 3x3 block diagonal of A,  S~ = diag(B diag(A)^-1 B^T)
// (an upper block-triangular preconditioner with a SIMPLE-type Schur diagonal):
//   P^-1 [v_u; v_p] = [A~^-1 (v_u - B^T z_p); z_p],  z_p = -S~^-1 v_p.
// Arnoldi with classical Gram-Schmidt applied twice (CGS2); the dot products of one pass
// are one kernel over all basis vectors; the (k+2) scalars of an iteration reach the host
// in one copy, where the Givens rotations and the least-squares update run.
#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <vector>

#include "fedd_internal.h"

namespace fedd {

namespace {

inline unsigned nbl(int64_t n, int t) { return (unsigned)((n + t - 1) / t); }

// y = J x: rows r < nu: A_r x_u + B^T_r x_p; rows r >= nu: B_(r-nu) x_u.  16 lanes per row.
__global__ __launch_bounds__(256) void k_spmv_ns(int64_t nu, int64_t np, const int64_t* __restrict__ arp,
                                                 const int32_t* __restrict__ aci, const double* __restrict__ av,
                                                 const int64_t* __restrict__ trp, const int32_t* __restrict__ tci,
                                                 const double* __restrict__ tv, const int64_t* __restrict__ brp,
                                                 const int32_t* __restrict__ bci, const double* __restrict__ bv,
                                                 const double* __restrict__ x, double* __restrict__ y) {
  const int64_t gid = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  const int64_t row = gid >> 4;
  const int lane = (int)(gid & 15);
  double s = 0.0;
  if (row < nu) {
    for (int64_t k = arp[row] + lane; k < arp[row + 1]; k += 16) s += av[k] * x[aci[k]];
    for (int64_t k = trp[row] + lane; k < trp[row + 1]; k += 16) s += tv[k] * x[nu + tci[k]];
  } else if (row < nu + np) {
    const int64_t r = row - nu;
    for (int64_t k = brp[r] + lane; k < brp[r + 1]; k += 16) s += bv[k] * x[bci[k]];
  }
#pragma unroll
  for (int o = 8; o > 0; o >>= 1) s += __shfl_xor(s, o, 16);
  if (lane == 0 && row < nu + np) y[row] = s;
}

// Inverse of the 3x3 diagonal node block of A (and the scalar diagonal): A's rows 3i+d
// hold full node blocks with sorted columns, so the block of node i starts at the first
// column >= 3i.
__global__ __launch_bounds__(256) void k_ns_block_diag(int64_t nnode, const int64_t* __restrict__ rowptr,
                                                       const int32_t* __restrict__ colind,
                                                       const double* __restrict__ vals, double* __restrict__ binv,
                                                       double* __restrict__ diag) {
  const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (i >= nnode) return;
  double a[3][3];
#pragma unroll
  for (int d1 = 0; d1 < 3; ++d1) {
    const int64_t row = 3 * i + d1;
    int64_t lo = rowptr[row], hi = rowptr[row + 1];
    const int32_t target = (int32_t)(3 * i);
    while (lo < hi) {
      const int64_t mid = (lo + hi) >> 1;
      if (colind[mid] < target) lo = mid + 1;
      else hi = mid;
    }
#pragma unroll
    for (int d2 = 0; d2 < 3; ++d2) a[d1][d2] = vals[lo + d2];
    diag[row] = a[d1][d1];
  }
  const double det = a[0][0] * (a[1][1] * a[2][2] - a[1][2] * a[2][1]) -
                     a[0][1] * (a[1][0] * a[2][2] - a[1][2] * a[2][0]) +
                     a[0][2] * (a[1][0] * a[2][1] - a[1][1] * a[2][0]);
  const double id = det != 0.0 ? 1.0 / det : 0.0;
  double* o = binv + 9 * i;
  o[0] = (a[1][1] * a[2][2] - a[1][2] * a[2][1]) * id;
  o[1] = (a[0][2] * a[2][1] - a[0][1] * a[2][2]) * id;
  o[2] = (a[0][1] * a[1][2] - a[0][2] * a[1][1]) * id;
  o[3] = (a[1][2] * a[2][0] - a[1][0] * a[2][2]) * id;
  o[4] = (a[0][0] * a[2][2] - a[0][2] * a[2][0]) * id;
  o[5] = (a[0][2] * a[1][0] - a[0][0] * a[1][2]) * id;
  o[6] = (a[1][0] * a[2][1] - a[1][1] * a[2][0]) * id;
  o[7] = (a[0][1] * a[2][0] - a[0][0] * a[2][1]) * id;
  o[8] = (a[0][0] * a[1][1] - a[0][1] * a[1][0]) * id;
}

// Sinv_p = 1 / sum_k B_pk^2 / diag(A)_k  (1 where the sum vanishes).  16 lanes per row.
__global__ __launch_bounds__(256) void k_schur_diag(int64_t np, const int64_t* __restrict__ brp,
                                                    const int32_t* __restrict__ bci, const double* __restrict__ bv,
                                                    const double* __restrict__ diag, double* __restrict__ sinv) {
  const int64_t gid = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  const int64_t r = gid >> 4;
  const int lane = (int)(gid & 15);
  double s = 0.0;
  if (r < np)
    for (int64_t k = brp[r] + lane; k < brp[r + 1]; k += 16) {
      const double d = diag[bci[k]];
      if (d != 0.0) s += bv[k] * bv[k] / d;
    }
#pragma unroll
  for (int o = 8; o > 0; o >>= 1) s += __shfl_xor(s, o, 16);
  if (lane == 0 && r < np) sinv[r] = s != 0.0 ? 1.0 / s : 1.0;
}

// z_p = -Sinv v_p.
__global__ __launch_bounds__(256) void k_prec_p(int64_t nu, int64_t np, const double* __restrict__ sinv,
                                                const double* __restrict__ v, double* __restrict__ z) {
  const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (i < np) z[nu + i] = -sinv[i] * v[nu + i];
}

// z_u = A~^-1 (v_u - B^T z_p), one 16-lane group per velocity node.
__global__ __launch_bounds__(256) void k_prec_u(int64_t nnode, int64_t nu, const int64_t* __restrict__ trp,
                                                const int32_t* __restrict__ tci, const double* __restrict__ tv,
                                                const double* __restrict__ binv, const double* __restrict__ v,
                                                double* __restrict__ z) {
  const int64_t gid = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  const int64_t i = gid >> 4;
  const int lane = (int)(gid & 15);
  double t[3] = {0.0, 0.0, 0.0};
  if (i < nnode)
#pragma unroll
    for (int d = 0; d < 3; ++d) {
      const int64_t row = 3 * i + d;
      for (int64_t k = trp[row] + lane; k < trp[row + 1]; k += 16) t[d] += tv[k] * z[nu + tci[k]];
    }
#pragma unroll
  for (int o = 8; o > 0; o >>= 1)
#pragma unroll
    for (int d = 0; d < 3; ++d) t[d] += __shfl_xor(t[d], o, 16);
  if (lane == 0 && i < nnode) {
    double r[3];
#pragma unroll
    for (int d = 0; d < 3; ++d) r[d] = v[3 * i + d] - t[d];
    const double* B = binv + 9 * i;
#pragma unroll
    for (int d = 0; d < 3; ++d) z[3 * i + d] = B[3 * d] * r[0] + B[3 * d + 1] * r[1] + B[3 * d + 2] * r[2];
  }
}

// part[b * kp1 + j] = sum over block b's entries of V_j . w  (j < kp1).
__global__ __launch_bounds__(256) void k_multidot(const double* __restrict__ V, int64_t ld, int kp1,
                                                  const double* __restrict__ w, int64_t n, double* __restrict__ part) {
  __shared__ double red[4];
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int j = 0; j < kp1; ++j) {
    const double* v = V + j * ld;
    double s = 0.0;
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += stride) s += v[i] * w[i];
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) s += __shfl_xor(s, o, 64);
    __syncthreads();
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = s;
    __syncthreads();
    if (threadIdx.x == 0) part[blockIdx.x * (int64_t)kp1 + j] = red[0] + red[1] + red[2] + red[3];
  }
}

// out[j] = sum_b part[b * kp1 + j].
__global__ __launch_bounds__(256) void k_colsum(const double* __restrict__ part, int nb, int kp1,
                                                double* __restrict__ out) {
  __shared__ double red[256];
  for (int j = 0; j < kp1; ++j) {
    double s = 0.0;
    for (int b = threadIdx.x; b < nb; b += blockDim.x) s += part[(int64_t)b * kp1 + j];
    red[threadIdx.x] = s;
    __syncthreads();
    for (int o = 128; o > 0; o >>= 1) {
      if ((int)threadIdx.x < o) red[threadIdx.x] += red[threadIdx.x + o];
      __syncthreads();
    }
    if (threadIdx.x == 0) out[j] = red[0];
    __syncthreads();
  }
}

// w -= sum_j h[j] V_j ; part[b] = block's sum of w^2.
__global__ __launch_bounds__(256) void k_gs_update(const double* __restrict__ V, int64_t ld, int kp1,
                                                   const double* __restrict__ h, double* __restrict__ w, int64_t n,
                                                   double* __restrict__ part) {
  __shared__ double hs[128];
  __shared__ double red[4];
  for (int j = threadIdx.x; j < kp1; j += blockDim.x) hs[j] = h[j];
  __syncthreads();
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  double s = 0.0;
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += stride) {
    double x = w[i];
    for (int j = 0; j < kp1; ++j) x -= hs[j] * V[j * ld + i];
    w[i] = x;
    s += x * x;
  }
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) s += __shfl_xor(s, o, 64);
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = s;
  __syncthreads();
  if (threadIdx.x == 0) part[blockIdx.x] = red[0] + red[1] + red[2] + red[3];
}

// out = sum_j y[j] V_j
__global__ __launch_bounds__(256) void k_lincomb(const double* __restrict__ V, int64_t ld, int kp1,
                                                 const double* __restrict__ y, double* __restrict__ out, int64_t n) {
  __shared__ double ys[128];
  for (int j = threadIdx.x; j < kp1; j += blockDim.x) ys[j] = y[j];
  __syncthreads();
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
    double x = 0.0;
    for (int j = 0; j < kp1; ++j) x += ys[j] * V[j * ld + i];
    out[i] = x;
  }
}

// y = a x + b y  (x, y of length n).  b = 0: y = a x without reading y, which may be
// fresh device memory whose bits are not a finite number (0 * NaN is NaN).
__global__ __launch_bounds__(256) void k_axpby(int64_t n, double a, const double* __restrict__ x, double b,
                                               double* __restrict__ y) {
  const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (i < n) y[i] = b == 0.0 ? a * x[i] : a * x[i] + b * y[i];
}

__global__ __launch_bounds__(256) void k_zero(int64_t n, double* __restrict__ y) {
  const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (i < n) y[i] = 0.0;
}

// dst[0, na) = a, dst[na, na + nb) = b   (concatenation)
__global__ __launch_bounds__(256) void k_concat(int64_t na, const double* __restrict__ a, int64_t nb,
                                                const double* __restrict__ b, double* __restrict__ dst) {
  const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (i < na) dst[i] = a[i];
  else if (i < na + nb) dst[i] = b ? b[i - na] : 0.0;
}

// a[0, na) = src[0, na), b[0, nb) = src[na, na + nb)
__global__ __launch_bounds__(256) void k_split(int64_t na, double* __restrict__ a, int64_t nb, double* __restrict__ b,
                                               const double* __restrict__ src, double alpha, int add) {
  const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (i < na) a[i] = add ? a[i] + alpha * src[i] : alpha * src[i];
  else if (i < na + nb) b[i - na] = add ? b[i - na] + alpha * src[i] : alpha * src[i];
}

}  // namespace

// Device work space of the NS solve.
struct NsSolveWork {
  int64_t n = 0, ld = 0;
  int m = 0, grid = 0;
  DevBuf V, w, z, b, x, r, binv, diag, sinv, part, h1, h2, ydev;
};

static int ns_work(Ctx* c, NsSolveWork& W, int m) {
  const int64_t nu = c->blk[0].nrows, np = c->blk[2].nrows;
  W.n = nu + np;
  W.ld = (W.n + 63) & ~int64_t(63);
  W.m = m;
  W.grid = (int)std::min<int64_t>(nbl(W.n, 256), (int64_t)c->num_cu * 4);
  int rc;
  if ((rc = alloc(c, W.V, (size_t)(m + 1) * W.ld * 8)) || (rc = alloc(c, W.w, W.ld * 8)) ||
      (rc = alloc(c, W.z, W.ld * 8)) || (rc = alloc(c, W.b, W.ld * 8)) || (rc = alloc(c, W.x, W.ld * 8)) ||
      (rc = alloc(c, W.r, W.ld * 8)) || (rc = alloc(c, W.binv, (nu / 3) * 9 * 8)) ||
      (rc = alloc(c, W.diag, nu * 8)) || (rc = alloc(c, W.sinv, std::max<int64_t>(np, 1) * 8)) ||
      (rc = alloc(c, W.part, (size_t)W.grid * (m + 2) * 8)) || (rc = alloc(c, W.h1, (m + 2) * 8)) ||
      (rc = alloc(c, W.h2, (m + 2) * 8)) || (rc = alloc(c, W.ydev, (m + 2) * 8)))
    return rc;
  return FEDD_OK;
}

static void spmv(Ctx* c, const double* x, double* y) {
  const Block &A = c->blk[0], &T = c->blk[1], &B = c->blk[2];
  const int64_t nu = A.nrows, np = B.nrows;
  k_spmv_ns<<<nbl((nu + np) * 16, 256), 256, 0, c->stream>>>(
      nu, np, A.d_rowptr.as<int64_t>(), A.d_colind.as<int32_t>(), A.d_vals.as<double>(), T.d_rowptr.as<int64_t>(),
      T.d_colind.as<int32_t>(), T.d_vals.as<double>(), B.d_rowptr.as<int64_t>(), B.d_colind.as<int32_t>(),
      B.d_vals.as<double>(), x, y);
}

static void precond(Ctx* c, NsSolveWork& W, const double* v, double* z) {
  const Block &A = c->blk[0], &T = c->blk[1];
  const int64_t nu = A.nrows, np = c->blk[2].nrows;
  k_prec_p<<<nbl(std::max<int64_t>(np, 1), 256), 256, 0, c->stream>>>(nu, np, W.sinv.as<double>(), v, z);
  k_prec_u<<<nbl(nu / 3 * 16, 256), 256, 0, c->stream>>>(nu / 3, nu, T.d_rowptr.as<int64_t>(), T.d_colind.as<int32_t>(),
                                                         T.d_vals.as<double>(), W.binv.as<double>(), v, z);
}

// ||v||^2 through the multidot kernel (V = v itself), on the host.
static int norm2(Ctx* c, NsSolveWork& W, const double* v, double* out) {
  k_multidot<<<W.grid, 256, 0, c->stream>>>(v, W.ld, 1, v, W.n, W.part.as<double>());
  k_colsum<<<1, 256, 0, c->stream>>>(W.part.as<double>(), W.grid, 1, W.h1.as<double>());
  FEDD_HIP(c, hipMemcpyAsync(out, W.h1.p, 8, hipMemcpyDeviceToHost, c->stream));
  FEDD_HIP(c, hipStreamSynchronize(c->stream));
  return FEDD_OK;
}

// Restarted right-preconditioned GMRES(m) on the current NS blocks: J x = b, b = W.b,
// x = W.x from zero.  Returns the iterations and the final true relative residual.
static int gmres(Ctx* c, NsSolveWork& W, double rtol, int maxit, int* iters, double* relres) {
  const Block& A = c->blk[0];
  const int64_t nu = A.nrows, np = c->blk[2].nrows, n = W.n;
  const int m = W.m;
  double* V = W.V.as<double>();
  // preconditioner set-up on the current matrix
  k_ns_block_diag<<<nbl(nu / 3, 256), 256, 0, c->stream>>>(nu / 3, A.d_rowptr.as<int64_t>(), A.d_colind.as<int32_t>(),
                                                           A.d_vals.as<double>(), W.binv.as<double>(),
                                                           W.diag.as<double>());
  if (np > 0)
    k_schur_diag<<<nbl(np * 16, 256), 256, 0, c->stream>>>(np, c->blk[2].d_rowptr.as<int64_t>(),
                                                           c->blk[2].d_colind.as<int32_t>(),
                                                           c->blk[2].d_vals.as<double>(), W.diag.as<double>(),
                                                           W.sinv.as<double>());
  k_zero<<<nbl(n, 256), 256, 0, c->stream>>>(n, W.x.as<double>());
  FEDD_HIP(c, hipGetLastError());
  double bb = 0.0;
  int rc;
  if ((rc = norm2(c, W, W.b.as<double>(), &bb))) return rc;
  const double bnorm = std::sqrt(bb);
  int it = 0;
  double res = bnorm;
  if (!(bnorm > 0.0)) {
    if (iters) *iters = 0;
    if (relres) *relres = 0.0;
    return FEDD_OK;
  }
  // r = b (x = 0)
  FEDD_HIP(c, hipMemcpyAsync(W.r.p, W.b.p, n * 8, hipMemcpyDeviceToDevice, c->stream));
  double beta = bnorm;
  std::vector<double> H((size_t)(m + 1) * m), cs(m), sn(m), g(m + 1), y(m + 1), hh(m + 2), h2(m + 2);
  while (it < maxit && beta > rtol * bnorm) {
    k_axpby<<<nbl(n, 256), 256, 0, c->stream>>>(n, 1.0 / beta, W.r.as<double>(), 0.0, V);
    std::fill(g.begin(), g.end(), 0.0);
    g[0] = beta;
    int k = 0;
    for (; k < m && it < maxit; ++k) {
      precond(c, W, V + k * W.ld, W.z.as<double>());
      spmv(c, W.z.as<double>(), W.w.as<double>());
      // CGS2: two projection passes against v_0..v_k
      for (int pass = 0; pass < 2; ++pass) {
        double* hdst = pass == 0 ? W.h1.as<double>() : W.h2.as<double>();
        k_multidot<<<W.grid, 256, 0, c->stream>>>(V, W.ld, k + 1, W.w.as<double>(), n, W.part.as<double>());
        k_colsum<<<1, 256, 0, c->stream>>>(W.part.as<double>(), W.grid, k + 1, hdst);
        k_gs_update<<<W.grid, 256, 0, c->stream>>>(V, W.ld, k + 1, hdst, W.w.as<double>(), n,
                                                   W.part.as<double>() + (size_t)W.grid * (m + 1));
      }
      k_colsum<<<1, 256, 0, c->stream>>>(W.part.as<double>() + (size_t)W.grid * (m + 1), W.grid, 1,
                                         W.h2.as<double>() + m + 1);
      FEDD_HIP(c, hipGetLastError());
      FEDD_HIP(c, hipMemcpyAsync(hh.data(), W.h1.p, (k + 1) * 8, hipMemcpyDeviceToHost, c->stream));
      FEDD_HIP(c, hipMemcpyAsync(h2.data(), W.h2.p, (m + 2) * 8, hipMemcpyDeviceToHost, c->stream));
      FEDD_HIP(c, hipStreamSynchronize(c->stream));
      const double hn = std::sqrt(std::max(h2[m + 1], 0.0));
      double* col = &H[(size_t)k * (m + 1)];
      for (int j = 0; j <= k; ++j) col[j] = hh[j] + h2[j];
      col[k + 1] = hn;
      ++it;
      // Givens rotations
      for (int j = 0; j < k; ++j) {
        const double t = cs[j] * col[j] + sn[j] * col[j + 1];
        col[j + 1] = -sn[j] * col[j] + cs[j] * col[j + 1];
        col[j] = t;
      }
      const double den = std::hypot(col[k], col[k + 1]);
      cs[k] = den > 0 ? col[k] / den : 1.0;
      sn[k] = den > 0 ? col[k + 1] / den : 0.0;
      col[k] = den;
      col[k + 1] = 0.0;
      g[k + 1] = -sn[k] * g[k];
      g[k] = cs[k] * g[k];
      res = std::fabs(g[k + 1]);
      if (!(res == res)) break;
      const bool done = res <= rtol * bnorm || hn == 0.0;
      if (!done && k + 1 < m)
        k_axpby<<<nbl(n, 256), 256, 0, c->stream>>>(n, 1.0 / hn, W.w.as<double>(), 0.0, V + (k + 1) * W.ld);
      if (done) {
        ++k;
        break;
      }
    }
    // y = H^-1 g (upper triangular, k columns); x += P^-1 V y
    for (int j = k - 1; j >= 0; --j) {
      double s = g[j];
      for (int l = j + 1; l < k; ++l) s -= H[(size_t)l * (m + 1) + j] * y[l];
      const double d = H[(size_t)j * (m + 1) + j];
      y[j] = d != 0.0 ? s / d : 0.0;
    }
    FEDD_HIP(c, hipMemcpyAsync(W.ydev.p, y.data(), k * 8, hipMemcpyHostToDevice, c->stream));
    k_lincomb<<<W.grid, 256, 0, c->stream>>>(V, W.ld, k, W.ydev.as<double>(), W.w.as<double>(), n);
    precond(c, W, W.w.as<double>(), W.z.as<double>());
    k_axpby<<<nbl(n, 256), 256, 0, c->stream>>>(n, 1.0, W.z.as<double>(), 1.0, W.x.as<double>());
    // true residual r = b - J x
    spmv(c, W.x.as<double>(), W.r.as<double>());
    k_axpby<<<nbl(n, 256), 256, 0, c->stream>>>(n, 1.0, W.b.as<double>(), -1.0, W.r.as<double>());
    FEDD_HIP(c, hipGetLastError());
    double rr = 0.0;
    if ((rc = norm2(c, W, W.r.as<double>(), &rr))) return rc;
    beta = std::sqrt(rr);
    if (!(beta == beta)) break;
  }
  if (iters) *iters = it;
  if (relres) *relres = beta / bnorm;
  return FEDD_OK;
}

static int ns_check(Ctx* c) {
  if (c->graph_kind != FEDD_GRAPH_NS) return fail(c, FEDD_ERR_STATE, "needs a Navier-Stokes graph");
  return FEDD_OK;
}

}  // namespace fedd

using namespace fedd;

extern "C" {

int fedd_ns_linear_solve(fedd_ctx* cc, double rtol, int restart, int maxit, int* iters, double* relres) {
  Ctx* c = cc;
  if (!c || !(rtol > 0) || restart < 1 || restart > 120 || maxit < 1) return FEDD_ERR_ARG;
  int rc = ns_check(c);
  if (rc) return rc;
  const int64_t nu = c->blk[0].nrows, np = c->blk[2].nrows;
  if (c->vec_n[FEDD_VEC_B] != nu || c->vec_n[FEDD_VEC_B_P] != np)
    return fail(c, FEDD_ERR_STATE, "right-hand side (VEC_B, VEC_B_P) not set");
  if (c->capturing) return fail(c, FEDD_ERR_STATE, "the solve synchronises: it cannot be captured");
  (void)hipSetDevice(c->device);
  NsSolveWork W;
  if ((rc = ns_work(c, W, restart))) return rc;
  k_concat<<<nbl(W.n, 256), 256, 0, c->stream>>>(nu, c->vec[FEDD_VEC_B].as<double>(), np,
                                                 c->vec[FEDD_VEC_B_P].as<double>(), W.b.as<double>());
  if ((rc = gmres(c, W, rtol, maxit, iters, relres))) return rc;
  for (int w : {FEDD_VEC_X, FEDD_VEC_X_P}) {
    const int64_t len = w == FEDD_VEC_X ? nu : np;
    if (c->vec_n[w] != len) {
      if ((rc = alloc(c, c->vec[w], std::max<int64_t>(len, 1) * 8))) return rc;
      c->vec_n[w] = len;
    }
  }
  k_split<<<nbl(W.n, 256), 256, 0, c->stream>>>(nu, c->vec[FEDD_VEC_X].as<double>(), np,
                                                c->vec[FEDD_VEC_X_P].as<double>(), W.x.as<double>(), 1.0, 0);
  FEDD_HIP(c, hipGetLastError());
  FEDD_HIP(c, hipStreamSynchronize(c->stream));
  return FEDD_OK;
}

int fedd_newton_navier_stokes(fedd_ctx* cc, const fedd_ns_newton_params* P, fedd_ns_newton_stats* st,
                              double* history, int history_cap) {
  Ctx* c = cc;
  if (!c || !P || !st) return FEDD_ERR_ARG;
  int rc = ns_check(c);
  if (rc) return rc;
  if (!(P->rel_tol > 0) || P->max_its < 1 || (P->criterion != 0 && P->criterion != 1) || !(P->lin_rtol > 0) ||
      P->restart < 1 || P->restart > 120 || P->lin_maxit < 1 || !(P->viscosity > 0) || !(P->density >= 0))
    return fail(c, FEDD_ERR_ARG, "bad Navier-Stokes Newton parameters");
  const int64_t nu = c->blk[0].nrows, np = c->blk[2].nrows;
  if (c->vec_n[FEDD_VEC_U] != nu) return fail(c, FEDD_ERR_STATE, "velocity iterate not uploaded");
  if (c->vec_n[FEDD_VEC_P] != np) return fail(c, FEDD_ERR_STATE, "pressure iterate not uploaded");
  if (c->nbc_dofs > 0 && (c->bc_per_node != 3 || (int64_t)c->bcflag_host.size() != nu))
    return fail(c, FEDD_ERR_STATE, "the Dirichlet set was made for another graph kind or mesh: set it again");
  if (c->capturing) return fail(c, FEDD_ERR_STATE, "the solve synchronises: it cannot be captured");
  (void)hipSetDevice(c->device);
  std::memset(st, 0, sizeof(*st));
  const auto t0 = std::chrono::steady_clock::now();
  NsSolveWork W;
  if ((rc = ns_work(c, W, P->restart))) return rc;
  for (int w : {FEDD_VEC_B, FEDD_VEC_RES_U, FEDD_VEC_RES_P, FEDD_VEC_B_P}) {
    const int64_t len = (w == FEDD_VEC_B || w == FEDD_VEC_RES_U) ? nu : np;
    if (c->vec_n[w] != len) {
      if ((rc = alloc(c, c->vec[w], std::max<int64_t>(len, 1) * 8))) return rc;
      c->vec_n[w] = len;
    }
  }
  const double* fext = c->vec_n[FEDD_VEC_FEXT] == nu ? c->vec[FEDD_VEC_FEXT].as<double>() : nullptr;
  fedd_ns_params ap{};
  ap.viscosity = P->viscosity;
  ap.density = P->density;
  ap.newton = P->linearization ? 1 : 0;
  for (int k = 0; k < 4; ++k) ap.coeff[k] = 1.0;
  hipEvent_t e0, e1;
  FEDD_HIP(c, hipEventCreate(&e0));
  FEDD_HIP(c, hipEventCreate(&e1));
  double r0 = 0.0, crit = 1.0;
  int hist = 0;
  bool converged = false;
  int its = 0;
  for (;;) {
    // calculateNonLinResidualVec("reverse") and assemble("Newton") on the same iterate
    FEDD_HIP(c, hipEventRecord(e0, c->stream));
    if ((rc = launch_ns_assembly(c, &ap, ap.newton ? FEDD_NS_JACOBIAN_RHS : FEDD_NS_FIXEDPOINT_RHS, 0, FEDD_FP64)))
      break;
    FEDD_HIP(c, hipEventRecord(e1, c->stream));
    c->newton_step++;
    // r = rhs - F(u): velocity rhs = VEC_FEXT (absent: 0), pressure rhs = 0
    if (fext)
      k_axpby<<<nbl(nu, 256), 256, 0, c->stream>>>(nu, 1.0, fext, 0.0, c->vec[FEDD_VEC_B].as<double>());
    else
      k_zero<<<nbl(nu, 256), 256, 0, c->stream>>>(nu, c->vec[FEDD_VEC_B].as<double>());
    k_axpby<<<nbl(nu, 256), 256, 0, c->stream>>>(nu, -1.0, c->vec[FEDD_VEC_RES_U].as<double>(), 1.0,
                                                 c->vec[FEDD_VEC_B].as<double>());
    if (np > 0)
      k_axpby<<<nbl(np, 256), 256, 0, c->stream>>>(np, -1.0, c->vec[FEDD_VEC_RES_P].as<double>(), 0.0,
                                                   c->vec[FEDD_VEC_B_P].as<double>());
    FEDD_HIP(c, hipGetLastError());
    if ((rc = fedd_dirichlet_vector(cc, FEDD_BC_MINUS_VECTOR, FEDD_VEC_B, FEDD_VEC_U))) break;  // g - u at Dirichlet dofs
    k_concat<<<nbl(W.n, 256), 256, 0, c->stream>>>(nu, c->vec[FEDD_VEC_B].as<double>(), np,
                                                   c->vec[FEDD_VEC_B_P].as<double>(), W.b.as<double>());
    double rr = 0.0;
    if ((rc = norm2(c, W, W.b.as<double>(), &rr))) break;
    float ams = 0.f;
    FEDD_HIP(c, hipEventElapsedTime(&ams, e0, e1));
    st->assembly_ms += ams;
    const double rnorm = std::sqrt(rr);
    if (history && hist < history_cap) history[hist++] = rnorm;
    st->final_residual = rnorm;
    if (its == 0) r0 = rnorm;
    if (P->criterion == 0) {
      crit = r0 > 0 ? rnorm / r0 : 0.0;
      st->rel_residual = crit;
      if (crit < P->rel_tol || rnorm == 0.0) {
        converged = true;
        break;
      }
    }
    if (its >= P->max_its) break;
    // setBoundariesSystem, then solveAndUpdate
    if ((rc = fedd_dirichlet_set_system(cc))) break;
    int li = 0;
    double lrel = 0.0;
    FEDD_HIP(c, hipEventRecord(e0, c->stream));
    if ((rc = gmres(c, W, P->lin_rtol, P->lin_maxit, &li, &lrel))) break;
    FEDD_HIP(c, hipEventRecord(e1, c->stream));
    st->krylov_iters += li;
    st->last_lin_relres = lrel;
    k_split<<<nbl(W.n, 256), 256, 0, c->stream>>>(nu, c->vec[FEDD_VEC_U].as<double>(), np,
                                                  c->vec[FEDD_VEC_P].as<double>(), W.x.as<double>(), 1.0, 1);
    FEDD_HIP(c, hipGetLastError());
    ++its;
    st->newton_iters = its;
    double dd = 0.0;
    if ((rc = norm2(c, W, W.x.as<double>(), &dd))) break;
    FEDD_HIP(c, hipEventElapsedTime(&ams, e0, e1));
    st->solve_ms += ams;
    st->last_update = std::sqrt(dd);
    if (P->criterion == 1 && st->last_update < P->rel_tol) {
      converged = true;
      break;
    }
  }
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  if (rc) return rc;
  FEDD_HIP(c, hipStreamSynchronize(c->stream));
  st->converged = converged ? 1 : 0;
  st->history_len = hist;
  st->seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
  return FEDD_OK;
}

}  // extern "C"
