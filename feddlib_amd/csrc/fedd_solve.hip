// Newton solve of the P2 hyperelastic problem on gfx950: Dirichlet rows/columns,
// traction loads, CSR SpMV, block-Jacobi preconditioned CG with device-resident
// scalars, and the load-stepped Newton driver with exact-tangent fallback.
//
// Reference roles: NonLinearSolver::solveNewton (NonLinearSolver_def.hpp:356-414)
// drives NonLinElasAssFE (NonLinElasAssFE_def.hpp:111-152, 262-377) with
// BCBuilder::setSystem (BCBuilder_def.hpp:590-618) and Belos/Amesos2 linear solves
// (LinearSolver_def.hpp:23-137).  Here the whole loop stays on the device.
#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <mutex>
#include <string>

#include "fedd_internal.h"

namespace fedd {

static inline unsigned nb(int64_t n, int t) { return (unsigned)((n + t - 1) / t); }

#ifndef FEDD_SPMV_LANES
#define FEDD_SPMV_LANES 32  // lanes per node in the node-block SpMV (16: +8%, 8: +42% time)
#endif

// ---------------------------------------------------------------- Dirichlet
// Constrained rows and columns become identity rows/columns (symmetric
// elimination keeps the tangent SPD; the increment of a constrained dof is 0).
__global__ __launch_bounds__(256) void k_bc_matrix(int64_t nrows, const int64_t* __restrict__ rowptr,
                                                   const int32_t* __restrict__ colind, double* __restrict__ vals,
                                                   const uint8_t* __restrict__ flag) {
  const int64_t gid = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  const int64_t row = gid >> 4;
  const int lane = (int)(gid & 15);
  if (row >= nrows) return;
  const bool fr = flag[row] != 0;
  for (int64_t k = rowptr[row] + lane; k < rowptr[row + 1]; k += 16) {
    const int32_t col = colind[k];
    if (fr || flag[col]) vals[k] = (col == row) ? 1.0 : 0.0;
  }
}

// Rows whose own dof or any column dof is constrained, appended in any order (16 lanes
// per row; the list's order does not change what k_bc_matrix_rows writes).
__global__ __launch_bounds__(256) void k_bc_rows_list(int64_t nrows, const int64_t* __restrict__ rowptr,
                                                      const int32_t* __restrict__ colind,
                                                      const uint8_t* __restrict__ flag, int32_t* __restrict__ rows,
                                                      int32_t* __restrict__ count) {
  const int64_t gid = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  const int64_t row = gid >> 4;
  const int lane = (int)(gid & 15);
  int hit = 0;
  if (row < nrows) {
    hit = flag[row] != 0;
    for (int64_t k = rowptr[row] + lane; k < rowptr[row + 1] && !hit; k += 16) hit = flag[colind[k]] != 0;
  }
#pragma unroll
  for (int o = 8; o > 0; o >>= 1) hit |= __shfl_xor(hit, o, 16);
  if (row < nrows && lane == 0 && hit) rows[atomicAdd(count, 1)] = (int32_t)row;
}

// k_bc_matrix on the listed rows only.
__global__ __launch_bounds__(256) void k_bc_matrix_rows(int64_t n, const int32_t* __restrict__ rows,
                                                        const int64_t* __restrict__ rowptr,
                                                        const int32_t* __restrict__ colind, double* __restrict__ vals,
                                                        const uint8_t* __restrict__ flag) {
  const int64_t gid = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if ((gid >> 4) >= n) return;
  const int64_t row = rows[gid >> 4];
  const int lane = (int)(gid & 15);
  const bool fr = flag[row] != 0;
  for (int64_t k = rowptr[row] + lane; k < rowptr[row + 1]; k += 16) {
    const int32_t col = colind[k];
    if (fr || flag[col]) vals[k] = (col == row) ? 1.0 : 0.0;
  }
}

// BCBuilder::setLocalRowOne / setLocalRowZero (BCBuilder_def.hpp:657-704) on the listed
// rows of one CRS block: every stored entry of a constrained row becomes 0, except, in the
// diagonal block (diag = 1), the entry in the row's own column, which becomes 1.  The
// columns of other rows and the graph stay as they are.  16 lanes per row.
__global__ __launch_bounds__(256) void k_bc_set_rows(int64_t n, const int32_t* __restrict__ rows,
                                                     const int64_t* __restrict__ rowptr,
                                                     const int32_t* __restrict__ colind, double* __restrict__ vals,
                                                     int diag) {
  const int64_t gid = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if ((gid >> 4) >= n) return;
  const int32_t row = rows[gid >> 4];
  const int lane = (int)(gid & 15);
  for (int64_t k = rowptr[row] + lane; k < rowptr[row + 1]; k += 16)
    vals[k] = (diag && colind[k] == row) ? 1.0 : 0.0;
}

// Dirichlet values into a vector at the constrained dofs (BCBuilder_def.hpp:93-160, 234-300,
// 350-420): mode 0 out = g (setRHS), 1 out = g - sub (setBCMinusVector), 2 out = sub - g
// (setVectorMinusBC).
__global__ __launch_bounds__(256) void k_bc_vector(int64_t n, const int32_t* __restrict__ dofs,
                                                   const double* __restrict__ g, const double* __restrict__ sub,
                                                   double* __restrict__ out, int mode) {
  const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (i >= n) return;
  const int32_t d = dofs[i];
  const double gv = g[d];
  out[d] = mode == 0 ? gv : mode == 1 ? gv - sub[d] : sub[d] - gv;
}

__global__ void k_fill_i32(int64_t n, int32_t* __restrict__ a, int32_t v) {
  const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (i < n) a[i] = v;
}

// R = -res - lam fext - fpres, zero on constrained dofs (res = assembled element rhs p,
// fpres = follower pressure load at the load factor's pressure).
__global__ __launch_bounds__(256) void k_newton_residual(int64_t n, const double* __restrict__ res,
                                                         const double* __restrict__ fext, double lam,
                                                         const double* __restrict__ fpres,
                                                         const uint8_t* __restrict__ flag,
                                                         double* __restrict__ R, double* __restrict__ part) {
  __shared__ double red[256];
  const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  double v = 0.0;
  if (i < n) {
    v = flag[i] ? 0.0 : -res[i] - lam * (fext ? fext[i] : 0.0) - (fpres ? fpres[i] : 0.0);
    R[i] = v;
  }
  red[threadIdx.x] = v * v;
  __syncthreads();
  for (int s = 128; s > 0; s >>= 1) {
    if ((int)threadIdx.x < s) red[threadIdx.x] += red[threadIdx.x + s];
    __syncthreads();
  }
  if (threadIdx.x == 0) part[blockIdx.x] = red[0];
}

// ------------------------------------------------------------------- Krylov
// Single-block sum of n partials into out[slot]; optional CG bookkeeping:
// mode 1: beta = sum / rz (stored in scal[5]) and rz = sum.
__global__ __launch_bounds__(1024) void k_reduce(const double* __restrict__ part, int64_t n,
                                                 double* __restrict__ scal, int slot, int mode) {
  __shared__ double red[1024];
  double v = 0.0;
  for (int64_t i = threadIdx.x; i < n; i += blockDim.x) v += part[i];
  red[threadIdx.x] = v;
  __syncthreads();
  for (int s = 512; s > 0; s >>= 1) {
    if ((int)threadIdx.x < s) red[threadIdx.x] += red[threadIdx.x + s];
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    if (mode == 1) {
      scal[5] = red[0] / scal[0];
      scal[0] = red[0];
    } else {
      scal[slot] = red[0];
    }
  }
}

// Block sum of v, added with one FP64 atomic per workgroup (256 threads) to one of
// 16 sub-slots out[0..15] (spreads same-address contention; readers sum all 16).
__device__ __forceinline__ void block_atomic_sum(double v, double* out) {
  __shared__ double red[4];
  __syncthreads();  // red[] may still be read by thread 0 of a previous call
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = v;
  __syncthreads();
  if (threadIdx.x == 0) atomicAdd(out + (blockIdx.x & 15), red[0] + red[1] + red[2] + red[3]);
}

__device__ __forceinline__ double sum16(const double* p) {
  double s = 0.0;
#pragma unroll
  for (int k = 0; k < 16; ++k) s += p[k];
  return s;
}

// PCG scalars live in rotating device slots (16 sub-slots each) so that no kernel
// clears a slot that another kernel of the same iteration reads: iteration t
// accumulates p.Ap in PAP(t), r.z in RZ(t) and r.r in RR(t), slot index t mod 3;
// rz of the previous iteration is RZ(t-1) (slot 2 for t = 0, from the first update).
#define PCG_PAP(t) (32 + 16 * ((t) % 3))
#define PCG_RZ(t) (80 + 16 * ((t) % 3))
#define PCG_RR(t) (128 + 16 * ((t) % 3))
#define PCG_SCAL 176

// y = A x on the 3x3 node-block structure of the vector graph: the three rows of
// node i are contiguous (row 3i+d starts at 9 nptr[i] + 3 deg_i d) and entry q of a
// row multiplies x[3 nidx[nptr[i] + q/3] + q%3].  G lanes per node walk each row
// with unit stride (coalesced value reads); x.y is added to *dot.
// V = float: the operator's FP32 copy (products and sums stay FP64).
template <int G, typename V = double>
__global__ __launch_bounds__(256) void k_bspmv_dot(int64_t nnode, const int64_t* __restrict__ nptr,
                                                   const int32_t* __restrict__ nidx,
                                                   const V* __restrict__ vals, const double* __restrict__ x,
                                                   double* __restrict__ y, double* __restrict__ dot) {
  const int lane = threadIdx.x % G;
  const int64_t stride = (int64_t)gridDim.x * blockDim.x / G;
  const int64_t nit = (nnode + stride - 1) / stride;  // uniform trip count (shuffles below)
  double dd = 0.0;
  for (int64_t it = 0; it < nit; ++it) {
    const int64_t i = (blockIdx.x * (int64_t)blockDim.x + threadIdx.x) / G + it * stride;
    double s[3] = {0.0, 0.0, 0.0};
    if (i < nnode) {
      const int64_t k0 = nptr[i];
      const int L = 3 * (int)(nptr[i + 1] - k0);
      const V* v = vals + 9 * k0;
      const int32_t* ni = nidx + k0;
      for (int q = lane; q < L; q += G) {
        const int k = (q * 0xAAAB) >> 17;  // q / 3 for q < 2^15
        const double xv = x[3 * (int64_t)ni[k] + (q - 3 * k)];
#pragma unroll
        for (int d = 0; d < 3; ++d) s[d] += (double)v[d * L + q] * xv;
      }
    }
#pragma unroll
    for (int o = G / 2; o > 0; o >>= 1)
#pragma unroll
      for (int d = 0; d < 3; ++d) s[d] += __shfl_xor(s[d], o, G);
    if (i < nnode && lane == 0) {
#pragma unroll
      for (int d = 0; d < 3; ++d) {
        y[3 * i + d] = s[d];
        dd += s[d] * x[3 * i + d];
      }
    }
  }
  block_atomic_sum(dd, dot);
}

// x += alpha p; r -= alpha Ap; z = Binv r (3x3 node blocks); r.z and r.r added to
// the slots of iteration t.  first: z = Binv r only, into the slots of t = -1.
__global__ __launch_bounds__(256) void k_pcg_update(int64_t nnode, double* __restrict__ x, double* __restrict__ r,
                                                    const double* __restrict__ p, const double* __restrict__ Ap,
                                                    double* __restrict__ z, const double* __restrict__ binv,
                                                    double* __restrict__ scal, int t, int first) {
  double rz = 0.0, rr = 0.0;
  const double alpha = first ? 0.0 : sum16(scal + PCG_RZ(t + 2)) / sum16(scal + PCG_PAP(t));
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < nnode; i += (int64_t)gridDim.x * blockDim.x) {
    double rl[3];
    if (first) {
#pragma unroll
      for (int d = 0; d < 3; ++d) rl[d] = r[3 * i + d];
    } else {
#pragma unroll
      for (int d = 0; d < 3; ++d) {
        x[3 * i + d] += alpha * p[3 * i + d];
        rl[d] = r[3 * i + d] - alpha * Ap[3 * i + d];
        r[3 * i + d] = rl[d];
      }
    }
    const double* B = binv + 9 * i;
#pragma unroll
    for (int d = 0; d < 3; ++d) {
      const double zv = B[3 * d] * rl[0] + B[3 * d + 1] * rl[1] + B[3 * d + 2] * rl[2];
      z[3 * i + d] = zv;
      rz += zv * rl[d];
      rr += rl[d] * rl[d];
    }
  }
  const int ts = first ? 2 : t;
  block_atomic_sum(rz, scal + PCG_RZ(ts));
  block_atomic_sum(rr, scal + PCG_RR(ts));
}

// p = z + beta p, beta = RZ(t) / RZ(t-1)  (first: p = z).  Block 0 clears the
// slots of iteration t+1 (last read in iteration t-2).
__global__ __launch_bounds__(256) void k_pcg_p(int64_t n, double* __restrict__ p, const double* __restrict__ z,
                                               double* __restrict__ scal, int t, int first) {
  const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (!first && blockIdx.x == 0 && threadIdx.x < 48) {
    const int k = threadIdx.x & 15, w = threadIdx.x >> 4;
    scal[(w == 0 ? PCG_PAP(t + 1) : w == 1 ? PCG_RZ(t + 1) : PCG_RR(t + 1)) + k] = 0.0;
  }
  if (i >= n) return;
  if (first) {
    p[i] = z[i];
    return;
  }
  const double rz_old = sum16(scal + PCG_RZ(t + 2));
  p[i] = z[i] + (rz_old != 0.0 ? sum16(scal + PCG_RZ(t)) / rz_old : 0.0) * p[i];
}

// x += alpha p; r -= alpha Ap; r.r added to RR(t) (multigrid-preconditioned CG).
__global__ __launch_bounds__(256) void k_pcg_xr(int64_t n, double* __restrict__ x, double* __restrict__ r,
                                                const double* __restrict__ p, const double* __restrict__ Ap,
                                                double* __restrict__ scal, int t) {
  const double pap = sum16(scal + PCG_PAP(t));
  const double alpha = pap != 0.0 ? sum16(scal + PCG_RZ(t + 2)) / pap : 0.0;
  double rr = 0.0;
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
    x[i] += alpha * p[i];
    const double v = r[i] - alpha * Ap[i];
    r[i] = v;
    rr += v * v;
  }
  block_atomic_sum(rr, scal + PCG_RR(t));
}

// r.z added to RZ(ts); with_rr: r.r added to RR(ts) as well.
__global__ __launch_bounds__(256) void k_pcg_dot(int64_t n, const double* __restrict__ r,
                                                 const double* __restrict__ z, double* __restrict__ scal, int ts,
                                                 int with_rr) {
  double rz = 0.0, rr = 0.0;
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
    rz += r[i] * z[i];
    rr += r[i] * r[i];
  }
  block_atomic_sum(rz, scal + PCG_RZ(ts));
  if (with_rr) block_atomic_sum(rr, scal + PCG_RR(ts));
}

// Inverses of the 3x3 diagonal node blocks (block Jacobi).
__global__ __launch_bounds__(256) void k_block_diag(int64_t nnode, const int64_t* __restrict__ rowptr,
                                                    const int32_t* __restrict__ colind,
                                                    const double* __restrict__ vals, double* __restrict__ binv) {
  const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (i >= nnode) return;
  double a[3][3];
#pragma unroll
  for (int d1 = 0; d1 < 3; ++d1) {
    const int64_t row = 3 * i + d1;
    int64_t lo = rowptr[row], hi = rowptr[row + 1];
    const int32_t target = (int32_t)(3 * i);
    while (lo < hi) {  // first column >= 3i
      const int64_t mid = (lo + hi) >> 1;
      if (colind[mid] < target) lo = mid + 1;
      else hi = mid;
    }
#pragma unroll
    for (int d2 = 0; d2 < 3; ++d2) a[d1][d2] = vals[lo + d2];
  }
  const double det = a[0][0] * (a[1][1] * a[2][2] - a[1][2] * a[2][1]) -
                     a[0][1] * (a[1][0] * a[2][2] - a[1][2] * a[2][0]) +
                     a[0][2] * (a[1][0] * a[2][1] - a[1][1] * a[2][0]);
  const double id = 1.0 / det;
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

__global__ __launch_bounds__(256) void k_axpy(int64_t n, double a, const double* __restrict__ x,
                                              double* __restrict__ y) {
  const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (i < n) y[i] += a * x[i];
}

__global__ __launch_bounds__(256) void k_scale_copy(int64_t n, double a, const double* __restrict__ x,
                                                    double* __restrict__ y) {
  const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (i < n) y[i] = a * x[i];
}

__global__ __launch_bounds__(256) void k_fill(int64_t n, double* __restrict__ y, double v) {
  const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (i < n) y[i] = v;
}

static int ensure(Ctx* c, DevBuf& b, size_t bytes) {
  if (b.bytes >= bytes) return FEDD_OK;
  return alloc(c, b, bytes);
}

static int ensure_vec(Ctx* c, int w, int64_t n) {
  if (c->vec_n[w] == n && c->vec[w].bytes >= (size_t)n * 8) return FEDD_OK;
  int rc = alloc(c, c->vec[w], n * 8);
  if (rc) return rc;
  c->vec_n[w] = n;
  return FEDD_OK;
}

// The Dirichlet set must have been made for the current graph kind (dofs per node) and mesh.
static int bc_matches_graph(Ctx* c) {
  if (c->graph_kind == GK_NONE) return fail(c, FEDD_ERR_STATE, "no graph");
  const int per_node = c->graph_kind == FEDD_GRAPH_SCALAR ? 1 : 3;
  if (c->nbc_dofs > 0 && (c->bc_per_node != per_node || (int64_t)c->bcflag_host.size() != per_node * c->mesh[0].npts))
    return fail(c, FEDD_ERR_STATE, "the Dirichlet set was made for another graph kind or mesh: set it again");
  return FEDD_OK;
}

int apply_dirichlet_matrix(Ctx* c) {
  const Block& A = c->blk[0];
  if (c->nbc == 0) return FEDD_OK;
  if (int rc = bc_matches_graph(c)) return rc;
  const int64_t* rp = A.d_rowptr.as<int64_t>();
  const int32_t* ci = A.d_colind.as<int32_t>();
  const uint8_t* fl = c->d_bcflag.as<uint8_t>();
  if ((c->nbc_rows < 0 || c->bcrows_graph != A.d_colind.p) && !c->capturing) {
    // once per Dirichlet set and graph: list the rows the constraints touch
    int rc = ensure(c, c->d_bcrows, (size_t)A.nrows * 4 + 8);
    if (rc) return rc;
    int32_t* cnt = c->d_bcrows.as<int32_t>() + A.nrows;
    k_fill_i32<<<1, 64, 0, c->stream>>>(1, cnt, 0);
    k_bc_rows_list<<<nb(A.nrows * 16, 256), 256, 0, c->stream>>>(A.nrows, rp, ci, fl, c->d_bcrows.as<int32_t>(), cnt);
    FEDD_HIP(c, hipGetLastError());
    int32_t h = 0;
    FEDD_HIP(c, hipMemcpyAsync(&h, cnt, 4, hipMemcpyDeviceToHost, c->stream));
    FEDD_HIP(c, hipStreamSynchronize(c->stream));
    c->nbc_rows = h;
    c->bcrows_graph = A.d_colind.p;
  }
  if (c->nbc_rows >= 0 && c->bcrows_graph == A.d_colind.p) {
    if (c->nbc_rows > 0)
      k_bc_matrix_rows<<<nb(c->nbc_rows * 16, 256), 256, 0, c->stream>>>(c->nbc_rows, c->d_bcrows.as<int32_t>(), rp,
                                                                          ci, A.d_vals.as<double>(), fl);
  } else {
    // a capture is open and the list is not built: every row
    k_bc_matrix<<<nb(A.nrows * 16, 256), 256, 0, c->stream>>>(A.nrows, rp, ci, A.d_vals.as<double>(), fl);
  }
  FEDD_HIP(c, hipGetLastError());
  return FEDD_OK;
}

// Work space of the block-Jacobi CG: four vectors of n, the 3x3 block inverses (9 per
// node = 3n) and the PCG_SCAL reduction slots.
struct PcgWork {
  double *r, *z, *p, *Ap, *binv, *scal;
};

// Block-Jacobi preconditioned CG on block A: A x = b, x starting from 0.  Every work
// array is written before it is read (r = b, scal and x cleared by kernels), so the
// work space may hold anything on entry.  timed: kernel-timing events around the SpMV.
static int pcg_core(Ctx* c, const double* b, double* x, const PcgWork& w, double rtol, double atol, int maxit,
                    int* iters, double* relres, bool timed) {
  const Block& A = c->blk[0];
  const int64_t n = A.nrows, nnode = n / 3;
  int rc;
  double* r = w.r;
  double* z = w.z;
  double* p = w.p;
  double* Ap = w.Ap;
  double* binv = w.binv;
  double* scal = w.scal;
  const int64_t* rp = A.d_rowptr.as<int64_t>();
  const int32_t* ci = A.d_colind.as<int32_t>();
  const double* va = A.d_vals.as<double>();
  const int64_t* nptr = c->d_nptr.as<int64_t>();
  const int32_t* nidx = c->d_nidx.as<int32_t>();
  // grid-stride kernels: a bounded number of workgroups bounds the atomics per dot product
  const unsigned gmax = (unsigned)c->num_cu * 8;
  const unsigned g_nodes = std::min(nb(nnode, 256), gmax);
  const unsigned g_spmv = std::min(nb(nnode * FEDD_SPMV_LANES, 256), gmax), g_dofs = nb(n, 256);
  k_block_diag<<<g_nodes, 256, 0, c->stream>>>(nnode, rp, ci, va, binv);
  FEDD_HIP(c, hipMemcpyAsync(r, b, n * 8, hipMemcpyDeviceToDevice, c->stream));
  k_fill<<<g_dofs, 256, 0, c->stream>>>(n, x, 0.0);
  k_fill<<<1, 256, 0, c->stream>>>(PCG_SCAL, scal, 0.0);
  k_pcg_update<<<g_nodes, 256, 0, c->stream>>>(nnode, x, r, p, Ap, z, binv, scal, -1, 1);
  k_pcg_p<<<g_dofs, 256, 0, c->stream>>>(n, p, z, scal, -1, 1);
  FEDD_HIP(c, hipGetLastError());
  double h[16];
  auto read_rr = [&](int t) -> int {
    FEDD_HIP(c, hipMemcpyAsync(h, scal + PCG_RR(t), 16 * 8, hipMemcpyDeviceToHost, c->stream));
    FEDD_HIP(c, hipStreamSynchronize(c->stream));
    return FEDD_OK;
  };
  auto sum = [&]() {
    double v = 0.0;
    for (int k = 0; k < 16; ++k) v += h[k];
    return v;
  };
  if ((rc = read_rr(2))) return rc;
  const double bnorm = std::sqrt(sum());
  double rnorm = bnorm;
  int it = 0;
  const double target = std::max(rtol * bnorm, atol);
  if (bnorm > 0.0 && rnorm > target) {
    while (it < maxit) {
      const int t = it;
      if (timed) ktimer_begin(c);
      k_bspmv_dot<FEDD_SPMV_LANES><<<g_spmv, 256, 0, c->stream>>>(nnode, nptr, nidx, va, p, Ap, scal + PCG_PAP(t));
      if (timed) ktimer_end(c, "bspmv");
      k_pcg_update<<<g_nodes, 256, 0, c->stream>>>(nnode, x, r, p, Ap, z, binv, scal, t, 0);
      k_pcg_p<<<g_dofs, 256, 0, c->stream>>>(n, p, z, scal, t, 0);
      ++it;
      if ((it & 7) == 0 || it == maxit) {
        FEDD_HIP(c, hipGetLastError());
        if ((rc = read_rr(t))) return rc;
        rnorm = std::sqrt(sum());
        if (!(rnorm == rnorm)) break;  // NaN: let the caller reject the step
        if (rnorm <= target) break;
      }
    }
  }
  FEDD_HIP(c, hipGetLastError());
  FEDD_HIP(c, hipStreamSynchronize(c->stream));
  if (iters) *iters = it;
  if (relres) *relres = bnorm > 0 ? rnorm / bnorm : 0.0;
  return FEDD_OK;
}

// Preconditioned CG on block A with b = VEC_B, x = VEC_X (starting from 0), work space
// allocated in the context.
int pcg_solve(Ctx* c, double rtol, double atol, int maxit, int* iters, double* relres) {
  const int64_t n = c->blk[0].nrows, nnode = n / 3;
  int rc;
  for (int k = 0; k < 4; ++k)
    if ((rc = ensure(c, c->d_kw[k], n * 8))) return rc;
  if ((rc = ensure(c, c->d_kw[4], nnode * 9 * 8))) return rc;
  if ((rc = ensure(c, c->d_scal, PCG_SCAL * 8))) return rc;
  if ((rc = ensure_vec(c, FEDD_VEC_X, n))) return rc;
  const PcgWork w{c->d_kw[0].as<double>(), c->d_kw[1].as<double>(), c->d_kw[2].as<double>(),
                  c->d_kw[3].as<double>(), c->d_kw[4].as<double>(), c->d_scal.as<double>()};
  return pcg_core(c, c->vec[FEDD_VEC_B].as<double>(), c->vec[FEDD_VEC_X].as<double>(), w, rtol, atol, maxit,
                  iters, relres, true);
}

// ------------------------------------------------------ no-allocation solve
// Work space of fedd_linear_solve_noalloc: one file-scope device array per GPU, held
// until the process ends.  Host code never uses the array's name as an address (on the
// host it is only a stand-in for the symbol); noalloc_init looks up each device's copy
// with hipGetSymbolAddress and every launch takes that pointer.
constexpr int64_t kNoallocDoubles = FEDD_NOALLOC_SCRATCH_BYTES / 8;
static __device__ __align__(256) double scratch[kNoallocDoubles];

namespace {
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
void* g_scratch[kMaxDevices] = {};  // device address of `scratch` per device (null: not looked up)
std::string g_scratch_err[kMaxDevices];

// Segment sizes are rounded to 256 bytes (32 doubles).
inline int64_t pad32(int64_t v) { return (v + 31) & ~int64_t(31); }
constexpr int64_t kScalPad = (PCG_SCAL + 31) & ~31;
inline int64_t noalloc_doubles(int64_t n) { return kScalPad + 4 * pad32(n) + pad32(3 * n); }
}  // namespace

// Called by fedd_ctx_create with the context's device current: looks up this device's
// copy of `scratch` once per device.  A failed lookup is kept and reported by
// fedd_linear_solve_noalloc; the other entry points do not need the array.
void noalloc_init(Ctx* c) {
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if (c->device < 0 || c->device >= kMaxDevices || g_scratch[c->device]) return;
  void* ptr = nullptr;
  const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
  if (e != hipSuccess || !ptr) {
    g_scratch_err[c->device] = std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e);
    return;
  }
  g_scratch[c->device] = ptr;
}

// CG preconditioned with one geometric-multigrid V-cycle per iteration (structured
// lattice meshes); same scalar slots and convergence test as pcg_solve.
// bnorm_known > 0: ||b|| is known to the caller (Newton: ||R||), which saves a
// synchronisation.  The convergence test reads r.r of iteration t-1 from pinned host
// memory while iteration t is already queued (at most one iteration beyond the target).
// refresh: 0 keep the hierarchy, 1 rebuild it for the current A, 2 finest level only.
#ifndef FEDD_CG_FP32_SPMV
#define FEDD_CG_FP32_SPMV 1
#endif
int pcg_mg_solve(Ctx* c, double rtol, double atol, int maxit, int* iters, double* relres, int refresh,
                 double bnorm_known, int mode, int fp32_spmv) {
  const Block& A = c->blk[0];
  const int64_t n = A.nrows, nnode = n / 3;
  int rc;
  if ((rc = mg_setup(c, mode))) return rc;
  if (refresh == 1 && (rc = mg_refresh(c))) return rc;
  if (refresh == 2 && (rc = mg_refresh_fine(c))) return rc;
  for (int k = 0; k < 4; ++k)
    if ((rc = ensure(c, c->d_kw[k], n * 8))) return rc;
  if ((rc = ensure(c, c->d_scal, PCG_SCAL * 8))) return rc;
  if ((rc = ensure_vec(c, FEDD_VEC_X, n))) return rc;
  double* r = c->d_kw[0].as<double>();
  double* z = c->d_kw[1].as<double>();
  double* p = c->d_kw[2].as<double>();
  double* Ap = c->d_kw[3].as<double>();
  double* x = c->vec[FEDD_VEC_X].as<double>();
  double* scal = c->d_scal.as<double>();
  const double* va = A.d_vals.as<double>();
  const int64_t* nptr = c->d_nptr.as<int64_t>();
  const int32_t* nidx = c->d_nidx.as<int32_t>();
  const unsigned gmax = (unsigned)c->num_cu * 8;
  const unsigned g_spmv = std::min(nb(nnode * FEDD_SPMV_LANES, 256), gmax);
  const unsigned g_dofs = nb(n, 256), g_red = std::min(nb(n, 256), gmax);
  FEDD_HIP(c, hipMemcpyAsync(r, c->vec[FEDD_VEC_B].p, n * 8, hipMemcpyDeviceToDevice, c->stream));
  k_fill<<<g_dofs, 256, 0, c->stream>>>(n, x, 0.0);
  k_fill<<<1, 256, 0, c->stream>>>(PCG_SCAL, scal, 0.0);
  if ((rc = mg_vcycle(c, r, z))) return rc;
  k_pcg_dot<<<g_red, 256, 0, c->stream>>>(n, r, z, scal, 2, 1);
  k_pcg_p<<<g_dofs, 256, 0, c->stream>>>(n, p, z, scal, -1, 1);
  FEDD_HIP(c, hipGetLastError());
  if (!c->h_rr) {
    FEDD_HIP(c, hipHostMalloc((void**)&c->h_rr, 4 * 16 * 8, hipHostMallocDefault));
    for (auto& e : c->ev_rr) FEDD_HIP(c, hipEventCreateWithFlags(&e, hipEventDisableTiming));
  }
  auto post_rr = [&](int t, int slot) -> int {
    FEDD_HIP(c, hipMemcpyAsync(c->h_rr + 16 * slot, scal + PCG_RR(t), 16 * 8, hipMemcpyDeviceToHost, c->stream));
    FEDD_HIP(c, hipEventRecord(c->ev_rr[slot], c->stream));
    return FEDD_OK;
  };
  auto wait_rr = [&](int slot, double* v) -> int {
    FEDD_HIP(c, hipEventSynchronize(c->ev_rr[slot]));
    double s = 0.0;
    for (int k = 0; k < 16; ++k) s += c->h_rr[16 * slot + k];
    *v = s;
    return FEDD_OK;
  };
  double bnorm = bnorm_known;
  if (!(bnorm > 0.0)) {
    double v;
    if ((rc = post_rr(2, 3)) || (rc = wait_rr(3, &v))) return rc;
    bnorm = std::sqrt(v);
  }
  double rnorm = bnorm;
  int it = 0;
  const double target = std::max(rtol * bnorm, atol);
  // fp32_spmv (Newton steps): the CG products read the FP32 copy of A that the smoother
  // reads (just refreshed), half the bytes of the iteration's largest pass.  The Krylov
  // solve then targets A rounded to FP32, so the step's true residual carries an extra
  // O(2^-24 cond(A)) relative term on top of the forcing term: an inexact-Newton step
  // still, while Newton's residual stays exact FP64.  Plain linear solves keep FP64.
  const float* va32 = FEDD_CG_FP32_SPMV && fp32_spmv && refresh != 0 ? mg_fine_vals32(c) : nullptr;
  if (bnorm > 0.0 && rnorm > target) {
    while (it < maxit) {
      const int t = it;
      if (va32)
        k_bspmv_dot<FEDD_SPMV_LANES, float><<<g_spmv, 256, 0, c->stream>>>(nnode, nptr, nidx, va32, p, Ap,
                                                                          scal + PCG_PAP(t));
      else
        k_bspmv_dot<FEDD_SPMV_LANES><<<g_spmv, 256, 0, c->stream>>>(nnode, nptr, nidx, va, p, Ap, scal + PCG_PAP(t));
      k_pcg_xr<<<g_red, 256, 0, c->stream>>>(n, x, r, p, Ap, scal, t);
      if ((rc = mg_vcycle(c, r, z))) return rc;
      k_pcg_dot<<<g_red, 256, 0, c->stream>>>(n, r, z, scal, t, 0);
      k_pcg_p<<<g_dofs, 256, 0, c->stream>>>(n, p, z, scal, t, 0);
      FEDD_HIP(c, hipGetLastError());
      if ((rc = post_rr(t, t & 1))) return rc;
      ++it;
      if (t == 0 && it < maxit) continue;
      // test the previous iteration while this one runs (at maxit: this one)
      const int tc = it == maxit ? t : t - 1;
      double v;
      if ((rc = wait_rr(tc & 1, &v))) return rc;
      rnorm = std::sqrt(v);
      if (!(rnorm == rnorm)) break;  // NaN: let the caller reject the step
      if (rnorm <= target && tc != t) {
        // x already holds iteration t: report its residual (CG's is not monotone)
        if ((rc = wait_rr(t & 1, &v))) return rc;
        rnorm = std::sqrt(v);
        if (!(rnorm == rnorm)) break;
      }
      if (rnorm <= target) break;
    }
  }
  FEDD_HIP(c, hipGetLastError());
  if (iters) *iters = it;
  if (relres) *relres = bnorm > 0 ? rnorm / bnorm : 0.0;
  return FEDD_OK;
}

// ------------------------------------------------------------ BiCGStab
// a.b into out0 and (if c) c.d into out1, 16 sub-slots each (block_atomic_sum).
__global__ __launch_bounds__(256) void k_dots(int64_t n, const double* __restrict__ a, const double* __restrict__ b,
                                              const double* __restrict__ c, const double* __restrict__ d,
                                              double* __restrict__ out0, double* __restrict__ out1) {
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  double s0 = 0.0, s1 = 0.0;
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += stride) {
    s0 += a[i] * b[i];
    if (c) s1 += c[i] * d[i];
  }
  block_atomic_sum(s0, out0);
  if (c) block_atomic_sum(s1, out1);
}

// p = r + beta (p - omega v)
__global__ __launch_bounds__(256) void k_bicg_p(int64_t n, double* __restrict__ p, const double* __restrict__ r,
                                                const double* __restrict__ v, double beta, double omega) {
  const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (i < n) p[i] = r[i] + beta * (p[i] - omega * v[i]);
}

// z = Binv r with the 3x3 node-block inverses
__global__ __launch_bounds__(256) void k_bj_apply(int64_t nnode, const double* __restrict__ binv,
                                                  const double* __restrict__ r, double* __restrict__ z) {
  const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (i >= nnode) return;
  const double* B = binv + 9 * i;
  const double r0 = r[3 * i], r1 = r[3 * i + 1], r2 = r[3 * i + 2];
#pragma unroll
  for (int d = 0; d < 3; ++d) z[3 * i + d] = B[3 * d] * r0 + B[3 * d + 1] * r1 + B[3 * d + 2] * r2;
}

// Right-preconditioned BiCGStab (van der Vorst 1992) on block A: b = VEC_B, x = VEC_X from
// 0, for the non-symmetric Newton matrices (follower loads).  precond: 0 block Jacobi, 1
// geometric multigrid, 2 smoothed aggregation (one V-cycle per application; refresh as
// for pcg_mg_solve).  Stops at ||r|| <= max(rtol ||b||, atol); three synchronisations per
// iteration (the dot products).
int pbicgstab_solve(Ctx* c, double rtol, double atol, int maxit, int* iters, double* relres, int precond,
                    int refresh) {
  const Block& A = c->blk[0];
  const int64_t n = A.nrows, nnode = n / 3;
  int rc;
  if (precond >= 1) {
    if ((rc = mg_setup(c, precond - 1))) return rc;
    if (refresh == 1 && (rc = mg_refresh(c))) return rc;
    if (refresh == 2 && (rc = mg_refresh_fine(c))) return rc;
  }
  for (int k = 0; k < 7; ++k)
    if ((rc = ensure(c, c->d_bw[k], n * 8))) return rc;
  if ((rc = ensure(c, c->d_kw[4], nnode * 9 * 8))) return rc;
  if ((rc = ensure(c, c->d_bscal, 64 * 8))) return rc;
  if ((rc = ensure_vec(c, FEDD_VEC_X, n))) return rc;
  double *r = c->d_bw[0].as<double>(), *rh = c->d_bw[1].as<double>(), *p = c->d_bw[2].as<double>(),
         *v = c->d_bw[3].as<double>(), *ph = c->d_bw[4].as<double>(), *sh = c->d_bw[5].as<double>(),
         *t = c->d_bw[6].as<double>();
  double* x = c->vec[FEDD_VEC_X].as<double>();
  const double* b = c->vec[FEDD_VEC_B].as<double>();
  double* bs = c->d_bscal.as<double>();
  double* binv = c->d_kw[4].as<double>();
  const int64_t* nptr = c->d_nptr.as<int64_t>();
  const int32_t* nidx = c->d_nidx.as<int32_t>();
  const double* va = A.d_vals.as<double>();
  const unsigned gmax = (unsigned)c->num_cu * 8;
  const unsigned g_spmv = std::min(nb(nnode * FEDD_SPMV_LANES, 256), gmax), g_dofs = nb(n, 256);
  const unsigned g_red = std::min(nb(n, 256), gmax);
  if (precond == 0) k_block_diag<<<nb(nnode, 256), 256, 0, c->stream>>>(nnode, A.d_rowptr.as<int64_t>(),
                                                                          A.d_colind.as<int32_t>(), va, binv);
  auto apply = [&](const double* in, double* out) -> int {
    if (precond >= 1) return mg_vcycle(c, in, out);
    k_bj_apply<<<nb(nnode, 256), 256, 0, c->stream>>>(nnode, binv, in, out);
    return FEDD_OK;
  };
  auto spmv = [&](const double* in, double* out) {
    k_bspmv_dot<FEDD_SPMV_LANES><<<g_spmv, 256, 0, c->stream>>>(nnode, nptr, nidx, va, in, out, bs + 32);
  };
  double h[32];
  auto dots = [&](const double* a1, const double* b1, const double* c1, const double* d1, double* o0,
                  double* o1) -> int {
    k_fill<<<1, 64, 0, c->stream>>>(32, bs, 0.0);
    k_dots<<<g_red, 256, 0, c->stream>>>(n, a1, b1, c1, d1, bs, bs + 16);
    FEDD_HIP(c, hipGetLastError());
    FEDD_HIP(c, hipMemcpyAsync(h, bs, 32 * 8, hipMemcpyDeviceToHost, c->stream));
    FEDD_HIP(c, hipStreamSynchronize(c->stream));
    double s0 = 0.0, s1 = 0.0;
    for (int k = 0; k < 16; ++k) {
      s0 += h[k];
      s1 += h[16 + k];
    }
    *o0 = s0;
    if (o1) *o1 = s1;
    return FEDD_OK;
  };
  k_fill<<<g_dofs, 256, 0, c->stream>>>(n, x, 0.0);
  FEDD_HIP(c, hipMemcpyAsync(r, b, n * 8, hipMemcpyDeviceToDevice, c->stream));
  FEDD_HIP(c, hipMemcpyAsync(rh, b, n * 8, hipMemcpyDeviceToDevice, c->stream));
  k_fill<<<g_dofs, 256, 0, c->stream>>>(n, p, 0.0);
  k_fill<<<g_dofs, 256, 0, c->stream>>>(n, v, 0.0);
  double rho = 1.0, alpha = 1.0, omega = 1.0, rho_new = 0.0, rr = 0.0;
  if ((rc = dots(rh, r, r, r, &rho_new, &rr))) return rc;
  const double bnorm = std::sqrt(rr);
  const double target = std::max(rtol * bnorm, atol);
  double rnorm = bnorm;
  int it = 0;
  while (rnorm > target && it < maxit) {
    if (rho_new == 0.0 || omega == 0.0) break;  // breakdown: the caller sees relres
    const double beta = (rho_new / rho) * (alpha / omega);
    k_bicg_p<<<g_dofs, 256, 0, c->stream>>>(n, p, r, v, beta, omega);
    if ((rc = apply(p, ph))) return rc;
    spmv(ph, v);
    double rhv = 0.0;
    if ((rc = dots(rh, v, nullptr, nullptr, &rhv, nullptr))) return rc;
    if (rhv == 0.0) break;
    alpha = rho_new / rhv;
    k_axpy<<<g_dofs, 256, 0, c->stream>>>(n, -alpha, v, r);  // r holds s
    k_axpy<<<g_dofs, 256, 0, c->stream>>>(n, alpha, ph, x);
    if ((rc = apply(r, sh))) return rc;
    spmv(sh, t);
    double ts = 0.0, tt = 0.0;
    if ((rc = dots(t, r, t, t, &ts, &tt))) return rc;
    omega = tt > 0.0 ? ts / tt : 0.0;
    k_axpy<<<g_dofs, 256, 0, c->stream>>>(n, omega, sh, x);
    k_axpy<<<g_dofs, 256, 0, c->stream>>>(n, -omega, t, r);
    rho = rho_new;
    ++it;
    if ((rc = dots(rh, r, r, r, &rho_new, &rr))) return rc;
    rnorm = std::sqrt(rr);
    if (!(rnorm == rnorm)) break;  // NaN: let the caller reject the step
  }
  FEDD_HIP(c, hipGetLastError());
  if (iters) *iters = it;
  if (relres) *relres = bnorm > 0 ? rnorm / bnorm : 0.0;
  return FEDD_OK;
}

// ||x||_2 of a device vector of length n (synchronises the context stream).
__global__ __launch_bounds__(256) void k_sumsq(int64_t n, const double* __restrict__ x, double* __restrict__ part) {
  __shared__ double red[256];
  const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  const double v = i < n ? x[i] : 0.0;
  red[threadIdx.x] = v * v;
  __syncthreads();
  for (int s = 128; s > 0; s >>= 1) {
    if ((int)threadIdx.x < s) red[threadIdx.x] += red[threadIdx.x + s];
    __syncthreads();
  }
  if (threadIdx.x == 0) part[blockIdx.x] = red[0];
}

static int vec_norm(Ctx* c, const double* x, int64_t n, double* out) {
  const int64_t nblk = nb(n, 256);
  int rc;
  if ((rc = ensure(c, c->d_part, nblk * 8))) return rc;
  if ((rc = ensure(c, c->d_scal, PCG_SCAL * 8))) return rc;
  k_sumsq<<<nblk, 256, 0, c->stream>>>(n, x, c->d_part.as<double>());
  k_reduce<<<1, 1024, 0, c->stream>>>(c->d_part.as<double>(), nblk, c->d_scal.as<double>(), 7, 0);
  FEDD_HIP(c, hipGetLastError());
  double h = 0.0;
  FEDD_HIP(c, hipMemcpyAsync(&h, c->d_scal.as<double>() + 7, 8, hipMemcpyDeviceToHost, c->stream));
  FEDD_HIP(c, hipStreamSynchronize(c->stream));
  *out = std::sqrt(h);
  return FEDD_OK;
}

int newton_residual(Ctx* c, double lam, bool pressure, double* norm) {
  const int64_t n = 3 * c->mesh[0].npts;
  const int64_t nblk = nb(n, 256);
  int rc;
  if ((rc = ensure(c, c->d_part, nblk * 8))) return rc;
  if ((rc = ensure(c, c->d_scal, PCG_SCAL * 8))) return rc;
  if ((rc = ensure_vec(c, FEDD_VEC_B, n))) return rc;
  if ((rc = ensure(c, c->d_bcflag, n))) return rc;
  if (c->nbc == 0) FEDD_HIP(c, hipMemsetAsync(c->d_bcflag.p, 0, n, c->stream));
  const double* fext = c->vec_n[FEDD_VEC_FEXT] == n ? c->vec[FEDD_VEC_FEXT].as<double>() : nullptr;
  const double* fp = pressure && c->vec_n[FEDD_VEC_FPRES] == n ? c->vec[FEDD_VEC_FPRES].as<double>() : nullptr;
  k_newton_residual<<<nblk, 256, 0, c->stream>>>(n, c->vec[FEDD_VEC_RES_U].as<double>(), fext, lam, fp,
                                                 c->d_bcflag.as<uint8_t>(), c->vec[FEDD_VEC_B].as<double>(),
                                                 c->d_part.as<double>());
  k_reduce<<<1, 1024, 0, c->stream>>>(c->d_part.as<double>(), nblk, c->d_scal.as<double>(), 6, 0);
  FEDD_HIP(c, hipGetLastError());
  double h = 0.0;
  FEDD_HIP(c, hipMemcpyAsync(&h, c->d_scal.as<double>() + 6, 8, hipMemcpyDeviceToHost, c->stream));
  FEDD_HIP(c, hipStreamSynchronize(c->stream));
  *norm = std::sqrt(h);
  return FEDD_OK;
}

// Traction on the face {x_axis == value}: a uniform traction t on a 6-node
// triangle gives |T|/3 t to each edge node and 0 to the vertices.
int traction_load(Ctx* c, int axis, double value, const double t[3], double tol) {
  const Mesh& m = c->mesh[0];
  const int64_t n = 3 * m.npts;
  std::vector<double> f(n, 0.0);
  static const int faces[4][3] = {{1, 2, 3}, {0, 2, 3}, {0, 1, 3}, {0, 1, 2}};
  static const int edge_of[4][4] = {{-1, 4, 6, 7}, {4, -1, 5, 8}, {6, 5, -1, 9}, {7, 8, 9, -1}};
  for (int64_t e = 0; e < m.nelem; ++e) {
    const int32_t* nd = &m.elems[e * m.npe];
    for (auto& fc : faces) {
      bool on = true;
      for (int k = 0; k < 3 && on; ++k) on = std::fabs(m.pts[3 * (int64_t)nd[fc[k]] + axis] - value) <= tol;
      if (!on) continue;
      const double* a = &m.pts[3 * (int64_t)nd[fc[0]]];
      const double* b = &m.pts[3 * (int64_t)nd[fc[1]]];
      const double* cc = &m.pts[3 * (int64_t)nd[fc[2]]];
      const double u1[3] = {b[0] - a[0], b[1] - a[1], b[2] - a[2]};
      const double u2[3] = {cc[0] - a[0], cc[1] - a[1], cc[2] - a[2]};
      const double cr[3] = {u1[1] * u2[2] - u1[2] * u2[1], u1[2] * u2[0] - u1[0] * u2[2],
                            u1[0] * u2[1] - u1[1] * u2[0]};
      const double area = 0.5 * std::sqrt(cr[0] * cr[0] + cr[1] * cr[1] + cr[2] * cr[2]);
      const int pairs[3][2] = {{fc[0], fc[1]}, {fc[1], fc[2]}, {fc[0], fc[2]}};
      for (auto& pr : pairs) {
        const int32_t en = nd[edge_of[pr[0]][pr[1]]];
        for (int d = 0; d < 3; ++d) f[3 * (int64_t)en + d] += area / 3.0 * t[d];
      }
    }
  }
  int rc = upload(c, c->vec[FEDD_VEC_FEXT], f.data(), n * 8);
  if (rc) return rc;
  c->vec_n[FEDD_VEC_FEXT] = n;
  FEDD_HIP(c, hipStreamSynchronize(c->stream));
  return FEDD_OK;
}

}  // namespace fedd

using namespace fedd;

extern "C" {

int fedd_dirichlet_set(fedd_ctx* cc, const int32_t* dofs, int64_t n) {
  Ctx* c = cc;
  if (!c || (n > 0 && !dofs)) return FEDD_ERR_ARG;
  const int per_node = c->graph_kind == FEDD_GRAPH_SCALAR ? 1 : 3;
  const int64_t nd = per_node * c->mesh[0].npts;
  if (nd == 0) return fail(c, FEDD_ERR_STATE, "no primary mesh");
  std::vector<uint8_t> flag(nd, 0);
  for (int64_t k = 0; k < n; ++k) {
    if (dofs[k] < 0 || dofs[k] >= nd) return fail(c, FEDD_ERR_ARG, "Dirichlet dof out of range");
    flag[dofs[k]] = 1;
  }
  if (c->capturing) return fail(c, FEDD_ERR_STATE, "host transfers cannot be captured");
  std::vector<int32_t> list;
  for (int64_t d = 0; d < nd; ++d)
    if (flag[d]) list.push_back((int32_t)d);
  (void)hipSetDevice(c->device);
  mg_free(c);  // the multigrid masks follow the Dirichlet set (waits for queued work first)
  sync_streams(c);  // queued kernels may still read the previous flags
  int rc = upload(c, c->d_bcflag, flag.data(), nd);
  if (rc) return rc;
  if ((rc = upload(c, c->d_bcdofs, list.data(), list.size() * 4))) return rc;
  // values follow the set: homogeneous until fedd_dirichlet_values
  if (c->d_bcval.bytes != (size_t)nd * 8 && (rc = alloc(c, c->d_bcval, nd * 8))) return rc;
  FEDD_HIP(c, hipMemsetAsync(c->d_bcval.p, 0, nd * 8, c->stream));
  c->nbc = n;
  c->nbc_dofs = (int64_t)list.size();
  c->bc_per_node = per_node;
  c->nbc_rows = -1;  // the row list follows the Dirichlet set
  c->bcflag_host = flag;
  FEDD_HIP(c, hipStreamSynchronize(c->stream));
  return FEDD_OK;
}

int fedd_dirichlet_values(fedd_ctx* cc, const double* g, int64_t n) {
  Ctx* c = cc;
  if (!c || !g) return FEDD_ERR_ARG;
  int rc = bc_matches_graph(c);
  if (rc) return rc;
  if (c->bcflag_host.empty()) return fail(c, FEDD_ERR_STATE, "no Dirichlet set (fedd_dirichlet_set)");
  if (n != (int64_t)c->bcflag_host.size()) return fail(c, FEDD_ERR_ARG, "Dirichlet values: one per dof of the primary field");
  if (c->capturing) return fail(c, FEDD_ERR_STATE, "host transfers cannot be captured");
  (void)hipSetDevice(c->device);
  if ((rc = upload(c, c->d_bcval, g, n * 8))) return rc;
  FEDD_HIP(c, hipStreamSynchronize(c->stream));
  return FEDD_OK;
}

int fedd_dirichlet_set_system(fedd_ctx* cc) {
  Ctx* c = cc;
  if (!c) return FEDD_ERR_ARG;
  int rc = bc_matches_graph(c);
  if (rc) return rc;
  if (c->nbc_dofs == 0) return FEDD_OK;
  (void)hipSetDevice(c->device);
  const int32_t* rows = c->d_bcdofs.as<int32_t>();
  const unsigned g = nb(c->nbc_dofs * 16, 256);
  // diagonal block of the constrained field: identity rows (setLocalRowOne)
  const Block& A = c->blk[0];
  k_bc_set_rows<<<g, 256, 0, c->stream>>>(c->nbc_dofs, rows, A.d_rowptr.as<int64_t>(), A.d_colind.as<int32_t>(),
                                          A.d_vals.as<double>(), 1);
  if (c->graph_kind == FEDD_GRAPH_NS) {
    // off-diagonal block of the same block row: zero rows (setLocalRowZero)
    const Block& BT = c->blk[1];
    k_bc_set_rows<<<g, 256, 0, c->stream>>>(c->nbc_dofs, rows, BT.d_rowptr.as<int64_t>(),
                                            BT.d_colind.as<int32_t>(), BT.d_vals.as<double>(), 0);
  }
  FEDD_HIP(c, hipGetLastError());
  return FEDD_OK;
}

int fedd_dirichlet_vector(fedd_ctx* cc, int mode, int out, int sub) {
  Ctx* c = cc;
  if (!c || mode < FEDD_BC_RHS || mode > FEDD_BC_VECTOR_MINUS_BC || out < 0 || out >= FEDD_NUM_VECTORS)
    return FEDD_ERR_ARG;
  if (mode != FEDD_BC_RHS && (sub < 0 || sub >= FEDD_NUM_VECTORS)) return FEDD_ERR_ARG;
  int rc = bc_matches_graph(c);
  if (rc) return rc;
  const int64_t nd = (c->graph_kind == FEDD_GRAPH_SCALAR ? 1 : 3) * c->mesh[0].npts;
  if (c->vec_n[out] != nd || (mode != FEDD_BC_RHS && c->vec_n[sub] != nd))
    return fail(c, FEDD_ERR_STATE, "vector absent or not one entry per dof of the primary field");
  if (c->nbc_dofs == 0) return FEDD_OK;
  (void)hipSetDevice(c->device);
  k_bc_vector<<<nb(c->nbc_dofs, 256), 256, 0, c->stream>>>(
      c->nbc_dofs, c->d_bcdofs.as<int32_t>(), c->d_bcval.as<double>(),
      mode == FEDD_BC_RHS ? nullptr : c->vec[sub].as<double>(), c->vec[out].as<double>(), mode);
  FEDD_HIP(c, hipGetLastError());
  return FEDD_OK;
}

int fedd_traction_load(fedd_ctx* cc, int axis, double value, const double* traction) {
  Ctx* c = cc;
  if (!c || axis < 0 || axis > 2 || !traction) return FEDD_ERR_ARG;
  if (c->mesh[0].npe != 10) return fail(c, FEDD_ERR_STATE, "traction loads need a P2 primary mesh");
  (void)hipSetDevice(c->device);
  return traction_load(c, axis, value, traction, 1e-12);
}

int fedd_linear_solve(fedd_ctx* cc, double rtol, int maxit, int* iters, double* relres) {
  Ctx* c = cc;
  if (!c || !(rtol > 0) || maxit < 1) return FEDD_ERR_ARG;
  if (c->graph_kind != FEDD_GRAPH_VECTOR) return fail(c, FEDD_ERR_STATE, "needs a vector graph");
  if (c->vec_n[FEDD_VEC_B] != c->blk[0].nrows) return fail(c, FEDD_ERR_STATE, "right-hand side not set");
  (void)hipSetDevice(c->device);
  if (c->precond >= 1) return pcg_mg_solve(c, rtol, 0.0, maxit, iters, relres, 1, 0.0, c->precond - 1, 0);
  return pcg_solve(c, rtol, 0.0, maxit, iters, relres);
}

int fedd_linear_solve_noalloc(fedd_ctx* cc, const double* d_b, double* d_x, int64_t n, double rtol, int maxit,
                              int* iters, double* relres) {
  Ctx* c = cc;
  if (!c || !d_b || !d_x || !(rtol > 0) || maxit < 1) return FEDD_ERR_ARG;
  if (c->graph_kind != FEDD_GRAPH_VECTOR) return fail(c, FEDD_ERR_STATE, "needs a vector graph");
  if (c->capturing) return fail(c, FEDD_ERR_STATE, "the solve synchronises: it cannot be captured");
  if (n != c->blk[0].nrows) return fail(c, FEDD_ERR_ARG, "n must equal the rows of block A");
  if (n > FEDD_NOALLOC_MAX_DOFS || noalloc_doubles(n) > kNoallocDoubles)
    return fail(c, FEDD_ERR_LIMIT, "the no-allocation solve holds at most FEDD_NOALLOC_MAX_DOFS (599040) unknowns");
  double* base = nullptr;
  {
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (c->device >= 0 && c->device < kMaxDevices) base = static_cast<double*>(g_scratch[c->device]);
    if (!base)
      return fail(c, FEDD_ERR_HIP,
                  c->device >= 0 && c->device < kMaxDevices && !g_scratch_err[c->device].empty()
                      ? g_scratch_err[c->device]
                      : std::string("no scratch address for this device"));
  }
  (void)hipSetDevice(c->device);
  const int64_t np = pad32(n);
  PcgWork w;
  w.scal = base;
  w.r = base + kScalPad;
  w.z = w.r + np;
  w.p = w.z + np;
  w.Ap = w.p + np;
  w.binv = w.Ap + np;
  return pcg_core(c, d_b, d_x, w, rtol, 0.0, maxit, iters, relres, false);
}

int fedd_set_preconditioner(fedd_ctx* cc, int kind) {
  Ctx* c = cc;
  if (!c || kind < 0 || kind > 2) return FEDD_ERR_ARG;
  c->precond = kind;
  return FEDD_OK;
}

int fedd_apply_dirichlet_symmetric(fedd_ctx* cc) {
  Ctx* c = cc;
  if (!c) return FEDD_ERR_ARG;
  if (c->graph_kind != FEDD_GRAPH_VECTOR) return fail(c, FEDD_ERR_STATE, "needs a vector graph");
  (void)hipSetDevice(c->device);
  return apply_dirichlet_matrix(c);
}

int fedd_newton_elasticity(fedd_ctx* cc, const fedd_newton_params* P, fedd_newton_stats* st,
                           double* history, int history_cap) {
  Ctx* c = cc;
  if (!c || !P || !st) return FEDD_ERR_ARG;
  if (c->graph_kind != FEDD_GRAPH_VECTOR || c->mesh[0].npe != 10)
    return fail(c, FEDD_ERR_STATE, "needs a vector graph on a P2 (tet-10) mesh");
  if (P->load_steps < 1 || P->max_newton < 1 || !(P->abs_tol > 0) || P->preconditioner < 0 ||
      P->preconditioner > 2 || P->criterion < 0 || P->criterion > 2 || (P->criterion > 0 && !(P->rel_tol > 0)))
    return fail(c, FEDD_ERR_ARG, "bad Newton parameters");
  const int model = P->neo_hooke_model == 1 ? 1 : 2;
  const bool press = !c->newton_pressure.empty() && c->nface > 0;  // follower pressure load
  if ((P->tangent == 1 || P->collect_samples) && !c->mlp_ready)
    return fail(c, FEDD_ERR_STATE, "surrogate tangent or sample collection requested but no network");
  if (P->tangent == 1 && !mlp_matches(c, model, P->E, P->nu))
    return fail(c, FEDD_ERR_STATE, "the surrogate was built for another law or other material constants");
  (void)hipSetDevice(c->device);
  const int64_t n = 3 * c->mesh[0].npts;
  int rc;
  if ((rc = bc_matches_graph(c))) return rc;
  if ((rc = ensure_vec(c, FEDD_VEC_U, n))) return rc;
  if ((rc = ensure_vec(c, FEDD_VEC_RES_U, n))) return rc;
  std::memset(st, 0, sizeof(*st));
  auto t0 = std::chrono::steady_clock::now();
  hipEvent_t ea, eb, sa, sb;
  FEDD_HIP(c, hipEventCreate(&ea));
  FEDD_HIP(c, hipEventCreate(&eb));
  FEDD_HIP(c, hipEventCreate(&sa));
  FEDD_HIP(c, hipEventCreate(&sb));
  bool solve_pending = false;
  // the solve's span is read after the next synchronisation (no extra stall)
  auto settle_solve = [&]() -> int {
    if (!solve_pending) return FEDD_OK;
    float ms = 0.f;
    FEDD_HIP(c, hipEventElapsedTime(&ms, sa, sb));
    st->solve_ms += ms;
    solve_pending = false;
    return FEDD_OK;
  };
  DevBuf u_prev;
  if ((rc = alloc(c, u_prev, n * 8))) return rc;
  double* u = c->vec[FEDD_VEC_U].as<double>();
  int hist = 0;
  double final_norm = 0.0;
  bool all_converged = true;
  bool coarse_current = false;  // coarse levels built (mg_reuse)
  int base_iters = 0, last_iters = 0;
  for (int step = 1; step <= P->load_steps; ++step) {
    const double lam = (double)step / P->load_steps;
    bool surrogate = P->tangent == 1;
    double prev = INFINITY, norm = 0.0, r0 = 0.0;
    bool converged = false;
    // converged: ||R|| < abs_tol, or criterion "Residual": ||R|| / ||R_0|| < rel_tol
    auto done = [&](double r) {
      st->rel_residual = r0 > 0.0 ? r / r0 : 0.0;
      return r < P->abs_tol || (P->criterion == 1 && r < P->rel_tol * r0);
    };
    for (int it = 0; it <= P->max_newton; ++it) {
      FEDD_HIP(c, hipEventRecord(ea, c->stream));
      if ((rc = launch_elasticity(c, model, P->E, P->nu, P->precision, 1, surrogate ? 2 : 1))) return rc;
      if (press && (rc = surface_pressure(c, c->newton_pressure.data(), (int)c->newton_pressure.size(), lam, 1,
                                          FEDD_VEC_FPRES, false)))
        return rc;
      FEDD_HIP(c, hipEventRecord(eb, c->stream));
      if (P->collect_samples && (rc = mlp_collect_qp(c))) return rc;
      if ((rc = newton_residual(c, lam, press, &norm))) return rc;
      if ((rc = settle_solve())) return rc;
      float ms = 0.f;
      FEDD_HIP(c, hipEventElapsedTime(&ms, ea, eb));
      st->assembly_ms += ms;
      if (history && hist < history_cap) history[hist++] = norm;
      final_norm = norm;
      if (it == 0) r0 = norm;
      if (done(norm)) {
        converged = true;
        break;
      }
      if (it == P->max_newton) break;
      if (surrogate && std::isfinite(prev) && !(norm <= P->stall_ratio * prev)) {
        // The surrogate tangent stalled: continue this load step with the exact
        // tangent.  A step that increased the residual (or made it non-finite) is
        // rejected first; a merely slow one is kept.
        if (!(norm <= prev)) {
          FEDD_HIP(c, hipMemcpyAsync(u, u_prev.p, n * 8, hipMemcpyDeviceToDevice, c->stream));
          st->rejected++;
        }
        surrogate = false;
        st->fallbacks++;
        FEDD_HIP(c, hipEventRecord(ea, c->stream));
        if ((rc = launch_elasticity(c, model, P->E, P->nu, P->precision, 1, 1))) return rc;
        if (press && (rc = surface_pressure(c, c->newton_pressure.data(), (int)c->newton_pressure.size(), lam, 1,
                                            FEDD_VEC_FPRES, false)))
          return rc;
        FEDD_HIP(c, hipEventRecord(eb, c->stream));
        if ((rc = newton_residual(c, lam, press, &norm))) return rc;
        FEDD_HIP(c, hipEventElapsedTime(&ms, ea, eb));
        st->assembly_ms += ms;
        if (done(norm)) {
          converged = true;
          break;
        }
      }
      // Newton step: K du = -R  (B holds R)
      if ((rc = apply_dirichlet_matrix(c))) return rc;
      k_scale_copy<<<nb(n, 256), 256, 0, c->stream>>>(n, -1.0, c->vec[FEDD_VEC_B].as<double>(),
                                                      c->vec[FEDD_VEC_B].as<double>());
      // inexact Newton forcing term (Eisenstat-Walker), never tighter than needed
      double eta = std::isfinite(prev) ? 0.9 * (norm / prev) * (norm / prev) : 0.1;
      eta = std::min(0.1, std::max(eta, P->lin_rtol));
      eta = std::max(eta, 0.1 * P->abs_tol / std::max(norm, 1e-300));
      int iters = 0;
      double relres = 0.0;
      // mg_reuse: the coarse levels keep the operators of the tangent they were built
      // for, across Newton and load steps, until a solve needs more than 2 n0 + 2
      // iterations (n0: iterations of the solve right after the rebuild); the finest
      // level always follows A.
      const int refresh =
          P->mg_reuse && coarse_current && last_iters <= 2 * base_iters + 2 ? 2 : 1;
      FEDD_HIP(c, hipEventRecord(sa, c->stream));
      if (press)  // A + Kext is not symmetric
        rc = pbicgstab_solve(c, eta, 0.0, P->lin_maxit, &iters, &relres, P->preconditioner, refresh);
      else if (P->preconditioner >= 1)  // b = -R, ||b|| = norm
        rc = pcg_mg_solve(c, eta, 0.0, P->lin_maxit, &iters, &relres, refresh, norm, P->preconditioner - 1, 1);
      else
        rc = pcg_solve(c, eta, 0.0, P->lin_maxit, &iters, &relres);
      if (rc) return rc;
      FEDD_HIP(c, hipEventRecord(sb, c->stream));
      solve_pending = true;  // resolved after the next synchronisation
      st->krylov_iters += iters;
      if (refresh == 1) {
        base_iters = iters;
        coarse_current = true;
      }
      last_iters = iters;
      FEDD_HIP(c, hipMemcpyAsync(u_prev.p, u, n * 8, hipMemcpyDeviceToDevice, c->stream));
      k_axpy<<<nb(n, 256), 256, 0, c->stream>>>(n, 1.0, c->vec[FEDD_VEC_X].as<double>(), u);
      FEDD_HIP(c, hipGetLastError());
      st->newton_iters++;
      prev = norm;
      if (P->criterion == 2) {  // "Update": ||du||_2 (one synchronisation per update)
        if ((rc = vec_norm(c, c->vec[FEDD_VEC_X].as<double>(), n, &st->last_update))) return rc;
        if (st->last_update < P->rel_tol) {
          converged = true;
          final_norm = norm;
          break;
        }
      }
    }
    if (!converged) all_converged = false;
  }
  FEDD_HIP(c, hipStreamSynchronize(c->stream));
  auto t1 = std::chrono::steady_clock::now();
  if ((rc = settle_solve())) return rc;
  (void)hipEventDestroy(ea);
  (void)hipEventDestroy(eb);
  (void)hipEventDestroy(sa);
  (void)hipEventDestroy(sb);
  st->seconds = std::chrono::duration<double>(t1 - t0).count();
  st->final_residual = final_norm;
  st->converged = all_converged ? 1 : 0;
  st->history_len = hist;
  return FEDD_OK;
}

}  // extern "C"
