// Multigrid preconditioners for the 3-dof-per-node (elasticity) problem:
//   mode 0: geometric, on the structured unit-cube lattice of MeshStructured::buildMesh3D;
//   mode 1: smoothed aggregation, on any mesh (described before build_aggregation below).
// Both share the node-block level layout, the smoother, the coarsest solve and the
// replayed graphs.
//
// Geometric levels: the P2 space on mesh M (lattice of P0 = 2M+1 points per axis), then P1 on
// meshes M, M/2, M/4, ... (lattices P_(k+1) = (P_k + 1) / 2) while P_k - 1 is even and
// P_k >= 5.  Every prolongation is the same lattice rule: a fine node whose lattice
// coordinates are odd on the axis set S is the midpoint of the Kuhn edge e_S between
// the coarse nodes (x - e_S)/2 and (x + e_S)/2 (weights 1/2), an all-even node copies
// the coarse node x/2.  P1 on a mesh is embedded in P2 on the same mesh, and the
// Kuhn (Freudenthal) meshes are nested, so every level is a conforming subspace.
// Coarse operators are Galerkin products P^T A P on the 3x3 node blocks; Dirichlet
// dofs are masked out of P (rows and columns) and get identity rows on every level.
// Smoother: damped 3x3 block Jacobi (nu sweeps before and after); coarsest level:
// dense inverse (Gauss-Jordan on the device).  The V-cycle is symmetric, so it is a
// valid CG preconditioner.
//
// Reference role: the FROSch / MueLu preconditioners behind Belos
// (Preconditioner_def.hpp:206-291, LinearSolver_def.hpp:23-137).
#include <algorithm>
#include <atomic>
#include <cmath>
#include <vector>

#include "fedd_internal.h"

// Smoothing sweeps before and after the coarse correction, on the finest level and below
// it.  Newton to 1e-8 at 24^3x6 with coarse-level reuse (CG iterations, time):
//   fine 2 / coarse 3: 123, 0.082 s    fine 1 / coarse 3: 145, 0.077 s
//   fine 1 / coarse 2: 149, 0.074 s    fine 1 / coarse 1: 180, 0.077 s
//   fine 3 / coarse 3: 118, 0.092 s
// A fine sweep streams the whole P2 operator; the extra CG iterations cost less than the
// sweeps they save.
#ifndef FEDD_MG_NU_COARSE
#define FEDD_MG_NU_COARSE 2
#endif
#ifndef FEDD_MG_NU
#define FEDD_MG_NU 1
#endif
// Coarse-level sweeps of the aggregation hierarchy (its first coarse level holds a sixth of
// the finest level's blocks).  Same Newton solve with the aggregation hierarchy: 2 sweeps
// 333 iterations, 0.241 s; 3: 297, 0.222 s; 4: 281, 0.226 s; 6: 265, 0.243 s.
#ifndef FEDD_AMG_NU_COARSE
#define FEDD_AMG_NU_COARSE 3
#endif
#ifndef FEDD_MG_OMEGA
#define FEDD_MG_OMEGA 0.6  // block-Jacobi damping
#endif
// The smoother and residual of the finest level read an FP32 copy of the operator
// (refreshed with it): half the bytes of the HBM-bound sweeps.  The preconditioner stays
// a fixed symmetric linear operator, so CG's FP64 iteration is unaffected.
#ifndef FEDD_MG_G0
#define FEDD_MG_G0 32  // lanes per node in the finest level's sweeps
#endif
#ifndef FEDD_MG_FP32_FINE
#define FEDD_MG_FP32_FINE 1
#endif

namespace fedd {

static inline unsigned nbk(int64_t n, int t) { return (unsigned)((n + t - 1) / t); }

struct MgLevel {
  int64_t P = 0, nn = 0, nblk = 0, maxdeg = 0;
  DevBuf nptr, nidx, vals;  // node-block operator (levels >= 1)
  const int64_t* d_nptr = nullptr;
  const int32_t* d_nidx = nullptr;
  double* d_vals = nullptr;
  DevBuf binv, free, x, x2, b, r;
  DevBuf vals32;  // FP32 copy of the operator used by the smoother (finest level; wide aggregation levels)
  // aggregation hierarchy: transfer from the next finer level F to this level
  DevBuf agg;                // int32, per F node: its aggregate (nodes kSaNodes * J + q here)
  DevBuf tent;               // per F node, its 3 x kSaModes rows of the tentative prolongator
  DevBuf pptr, pidx, pval;   // smoothed prolongator, rows = F nodes, 3x3 blocks (row-major)
  DevBuf pval32, ptval32;    // FP32 copies for the V-cycle: P in pval's order, and in the transpose lists' order
  DevBuf tptr, tidx, tpos;   // its transpose: per node of this level, (F node, block index in pval)
  DevBuf aptr, aidx, aval;   // A_F P, rows = F nodes
  int64_t maxdeg_ap = 0;
  int64_t nblk_p = 0;  // blocks of P (aggregation levels)
  std::vector<int64_t> h_ptr;
  std::vector<int32_t> h_idx;
  std::vector<int32_t> h_agg;
  std::vector<double> h_tent;
  std::vector<int64_t> h_pptr, h_tptr, h_tpos, h_aptr;
  std::vector<int32_t> h_pidx, h_tidx, h_aidx;
};

struct MgHost {
  int mode = 0;  // 0: lattice (geometric), 1: aggregation (algebraic)
  std::vector<MgLevel> lv;
  DevBuf dense, inv;  // coarsest level: Gauss-Jordan ping-pong buffers (n x n); inverse in inv
  int64_t ncoarse = 0;
  // replayed HIP graphs of the V-cycle (for one (b, x) pair) and of the refresh
  hipGraphExec_t vexec = nullptr, rexec = nullptr;
  const double* vb = nullptr;
  double* vx = nullptr;
  ~MgHost() {
    if (vexec) (void)hipGraphExecDestroy(vexec);
    if (rexec) (void)hipGraphExecDestroy(rexec);
  }
};

// The replayed graphs and the level buffers may still be in use by work queued on the
// context's streams: wait for them before destroying the executables and freeing.
void mg_free(Ctx* c) {
  if (!c->mg) return;
  sync_streams(c);
  delete c->mg;
  c->mg = nullptr;
}

// ------------------------------------------------------------------ kernels
__device__ __forceinline__ void lattice(int64_t i, int64_t P, int64_t x[3]) {
  x[0] = i % P;
  x[1] = (i / P) % P;
  x[2] = i / (P * P);
}

// Node i of lattice P: parents in lattice Pc = (P + 1) / 2 and their weights.
__device__ __forceinline__ int parents(int64_t i, int64_t P, int64_t Pc, int64_t par[2], double w[2]) {
  int64_t x[3];
  lattice(i, P, x);
  const int64_t o0 = x[0] & 1, o1 = x[1] & 1, o2 = x[2] & 1;
  if (!(o0 | o1 | o2)) {
    par[0] = (x[0] >> 1) + Pc * ((x[1] >> 1) + Pc * (x[2] >> 1));
    w[0] = 1.0;
    return 1;
  }
  par[0] = ((x[0] - o0) >> 1) + Pc * (((x[1] - o1) >> 1) + Pc * ((x[2] - o2) >> 1));
  par[1] = ((x[0] + o0) >> 1) + Pc * (((x[1] + o1) >> 1) + Pc * ((x[2] + o2) >> 1));
  w[0] = w[1] = 0.5;
  return 2;
}

// Children of coarse node I (lattice Pc) in the fine lattice Pf = 2 Pc - 1.
__device__ __forceinline__ int children(int64_t I, int64_t Pc, int64_t Pf, int64_t ch[15], double w[15]) {
  int64_t X[3];
  lattice(I, Pc, X);
  int n = 0;
  ch[n] = 2 * X[0] + Pf * (2 * X[1] + Pf * 2 * X[2]);
  w[n++] = 1.0;
  for (int s = 1; s < 8; ++s)
    for (int sg = -1; sg <= 1; sg += 2) {
      const int64_t f0 = 2 * X[0] + sg * (s & 1), f1 = 2 * X[1] + sg * ((s >> 1) & 1),
                    f2 = 2 * X[2] + sg * ((s >> 2) & 1);
      if (f0 < 0 || f1 < 0 || f2 < 0 || f0 >= Pf || f1 >= Pf || f2 >= Pf) continue;
      ch[n] = f0 + Pf * (f1 + Pf * f2);
      w[n++] = 0.5;
    }
  return n;
}

__device__ __forceinline__ int find_slot(const int32_t* __restrict__ idx, int deg, int32_t J) {
  int lo = 0, hi = deg;
  while (lo < hi) {
    const int mid = (lo + hi) >> 1;
    if (idx[mid] < J) lo = mid + 1;
    else hi = mid;
  }
  return (lo < deg && idx[lo] == J) ? lo : -1;
}

// Damped block-Jacobi sweep xout = xin + omega Binv (b - A xin), G lanes per node
// (first: xin = 0, no product).  RESID: r = b - A x instead.
// VT: stored type of the operator values (float for the FP32 copy of the finest level);
// products and sums are FP64.
template <int G, int MODE, class VT>  // MODE 0: sweep, 1: first sweep, 2: residual
__global__ __launch_bounds__(256) void k_mg_sweep(int64_t nn, const int64_t* __restrict__ nptr,
                                                  const int32_t* __restrict__ nidx,
                                                  const VT* __restrict__ vals, const double* __restrict__ binv,
                                                  const double* __restrict__ b, const double* __restrict__ xin,
                                                  double* __restrict__ out, double omega) {
  const int64_t i = (blockIdx.x * (int64_t)blockDim.x + threadIdx.x) / G;
  const int lane = threadIdx.x % G;
  double s[3] = {0.0, 0.0, 0.0};
  if (i < nn && MODE != 1) {
    const int64_t k0 = nptr[i];
    const int L = 3 * (int)(nptr[i + 1] - k0);
    const VT* v = vals + 9 * k0;
    const int32_t* ni = nidx + k0;
    for (int q = lane; q < L; q += G) {
      const int k = (q * 0xAAAB) >> 17;  // q / 3 for q < 2^15
      const double xv = xin[3 * (int64_t)ni[k] + (q - 3 * k)];
#pragma unroll
      for (int d = 0; d < 3; ++d) s[d] += (double)v[d * L + q] * xv;
    }
  }
#pragma unroll
  for (int o = G / 2; o > 0; o >>= 1)
#pragma unroll
    for (int d = 0; d < 3; ++d) s[d] += __shfl_xor(s[d], o, G);
  if (i >= nn || lane != 0) return;
  double rv[3];
#pragma unroll
  for (int d = 0; d < 3; ++d) rv[d] = b[3 * i + d] - s[d];
  if (MODE == 2) {
#pragma unroll
    for (int d = 0; d < 3; ++d) out[3 * i + d] = rv[d];
    return;
  }
  const double* B = binv + 9 * i;
#pragma unroll
  for (int d = 0; d < 3; ++d) {
    const double z = B[3 * d] * rv[0] + B[3 * d + 1] * rv[1] + B[3 * d + 2] * rv[2];
    out[3 * i + d] = (MODE == 1 ? 0.0 : xin[3 * i + d]) + omega * z;
  }
}

__global__ __launch_bounds__(256) void k_mg_to_f32(int64_t n, const double* __restrict__ a, float* __restrict__ b) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
    b[i] = (float)a[i];
}

// Restriction on the lattice: 16 lanes per coarse node, lane 0 the coincident fine node
// (weight 1), lanes 1..14 the midpoints of the 14 Kuhn edges +-e_S (weight 1/2), then a
// shuffle reduction.  (One thread per coarse node leaves too few waves in flight on the
// 8x smaller coarse level.)
// With xc set, lane 0 also performs the coarse level's first smoothing sweep from zero,
// xc = omega D^-1 bc (the arithmetic of k_mg_sweep MODE 1), saving that launch.
__global__ __launch_bounds__(256) void k_mg_restrict(int64_t Pf, int64_t Pc, const double* __restrict__ rf,
                                                     const uint8_t* __restrict__ freef,
                                                     const uint8_t* __restrict__ freec, double* __restrict__ bc,
                                                     const double* __restrict__ binvc, double* __restrict__ xc,
                                                     double omega) {
  const int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  const int64_t I = t >> 4;
  const int lane = (int)(t & 15);
  const int64_t nc = Pc * Pc * Pc;
  double acc[3] = {0.0, 0.0, 0.0};
  if (I < nc && lane < 15) {
    int64_t X[3];
    lattice(I, Pc, X);
    const int s = (lane + 1) >> 1;              // axis set of the Kuhn edge (0: the node itself)
    const int sg = (lane & 1) ? 1 : -1;         // direction along it
    const int64_t f0 = 2 * X[0] + sg * (s & 1), f1 = 2 * X[1] + sg * ((s >> 1) & 1),
                  f2 = 2 * X[2] + sg * ((s >> 2) & 1);
    if (f0 >= 0 && f1 >= 0 && f2 >= 0 && f0 < Pf && f1 < Pf && f2 < Pf) {
      const int64_t f = f0 + Pf * (f1 + Pf * f2);
      const double w = lane == 0 ? 1.0 : 0.5;
#pragma unroll
      for (int d = 0; d < 3; ++d) acc[d] = freef[3 * f + d] ? w * rf[3 * f + d] : 0.0;
    }
  }
#pragma unroll
  for (int o = 8; o > 0; o >>= 1)
#pragma unroll
    for (int d = 0; d < 3; ++d) acc[d] += __shfl_xor(acc[d], o, 16);
  if (I >= nc || lane != 0) return;
  double rv[3];
#pragma unroll
  for (int d = 0; d < 3; ++d) {
    rv[d] = freec[3 * I + d] ? acc[d] : 0.0;
    bc[3 * I + d] = rv[d];
  }
  if (xc) {
    const double* B = binvc + 9 * I;
#pragma unroll
    for (int d = 0; d < 3; ++d) {
      const double z = B[3 * d] * rv[0] + B[3 * d + 1] * rv[1] + B[3 * d + 2] * rv[2];
      xc[3 * I + d] = 0.0 + omega * z;
    }
  }
}

__global__ __launch_bounds__(256) void k_mg_prolong(int64_t Pf, int64_t Pc, const double* __restrict__ xc,
                                                    const uint8_t* __restrict__ freef,
                                                    const uint8_t* __restrict__ freec, double* __restrict__ xf) {
  const int64_t f = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (f >= Pf * Pf * Pf) return;
  int64_t par[2];
  double w[2];
  const int n = parents(f, Pf, Pc, par, w);
#pragma unroll
  for (int d = 0; d < 3; ++d) {
    if (!freef[3 * f + d]) continue;
    double v = 0.0;
    for (int k = 0; k < n; ++k) v += freec[3 * par[k] + d] ? w[k] * xc[3 * par[k] + d] : 0.0;
    xf[3 * f + d] += v;
  }
}

// Frobenius norm of every 3x3 node block of a node-block operator (strength of connection).
__global__ __launch_bounds__(256) void k_amg_block_norm(int64_t nn, const int64_t* __restrict__ nptr,
                                                        const double* __restrict__ vals, float* __restrict__ out) {
  const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (i >= nn) return;
  const int64_t k0 = nptr[i];
  const int deg = (int)(nptr[i + 1] - k0);
  const double* v = vals + 9 * k0;
  for (int s = 0; s < deg; ++s) {
    double a = 0.0;
#pragma unroll
    for (int d1 = 0; d1 < 3; ++d1)
#pragma unroll
      for (int d2 = 0; d2 < 3; ++d2) {
        const double x = v[3 * deg * d1 + 3 * s + d2];
        a += x * x;
      }
    out[k0 + s] = (float)sqrt(a);
  }
}

// Near-null-space modes per aggregate: the three translations and, with 6, the three
// rigid rotations.  An aggregate J becomes kSaNodes coarse nodes kSaNodes * J + q, each
// with 3 dofs (coarse dof (J, 3q + d) is mode 3q + d), so every level keeps the 3x3
// node-block layout.
#ifndef FEDD_SA_MODES
#define FEDD_SA_MODES 6
#endif
constexpr int kSaModes = FEDD_SA_MODES;
constexpr int kSaNodes = FEDD_SA_MODES / 3;
static_assert(kSaModes == 3 || kSaModes == 6, "3 or 6 near-null-space modes");

// Smoothed aggregation.  The tentative prolongator T maps the coarse dofs of aggregate J
// to its free fine dofs: row (i, d) of T is row (i, d) of the aggregate's orthonormalised
// mode matrix (tent holds these 3 x kSaModes rows per fine node, zero for constrained
// dofs).  The prolongator is P = (I - omega D^-1 A) T, D the 3x3 block diagonal (binv
// holds its inverse).  Row i of P has a 3x3 block for every coarse node of the aggregates
// of i's neighbours; rows of constrained fine dofs are zero.
__global__ __launch_bounds__(256) void k_sa_smooth_p(int64_t nf, const int64_t* __restrict__ fptr,
                                                     const int32_t* __restrict__ fidx, const double* __restrict__ fval,
                                                     const double* __restrict__ binv, const uint8_t* __restrict__ ffree,
                                                     const int32_t* __restrict__ agg, const double* __restrict__ tent,
                                                     const int64_t* __restrict__ pptr, const int32_t* __restrict__ pidx,
                                                     double* __restrict__ pval, double omega) {
  const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (i >= nf) return;
  const int64_t p0 = pptr[i];
  const int np = (int)(pptr[i + 1] - p0);
  const int32_t* prow = pidx + p0;
  double* pv = pval + 9 * p0;
  for (int k = 0; k < 9 * np; ++k) pv[k] = 0.0;
  const int64_t f0 = fptr[i];
  const int degf = (int)(fptr[i + 1] - f0);
  const double* fv = fval + 9 * f0;
  for (int s = 0; s < degf; ++s) {  // S = A T
    const int32_t j = fidx[f0 + s];
    const double* T = tent + 3 * kSaModes * (int64_t)j;
    double a[3][3];
#pragma unroll
    for (int d1 = 0; d1 < 3; ++d1)
#pragma unroll
      for (int e = 0; e < 3; ++e) a[d1][e] = fv[3 * degf * d1 + 3 * s + e];
#pragma unroll
    for (int q = 0; q < kSaNodes; ++q) {
      const int slot = find_slot(prow, np, kSaNodes * agg[j] + q);
      if (slot < 0) continue;  // cannot happen: row i holds the aggregate of every neighbour
#pragma unroll
      for (int d1 = 0; d1 < 3; ++d1)
#pragma unroll
        for (int d2 = 0; d2 < 3; ++d2)
          pv[9 * slot + 3 * d1 + d2] += a[d1][0] * T[3 * q + d2] + a[d1][1] * T[kSaModes + 3 * q + d2] +
                                        a[d1][2] * T[2 * kSaModes + 3 * q + d2];
    }
  }
  const double* B = binv + 9 * i;
  const double* Ti = tent + 3 * kSaModes * i;
  const int32_t own = kSaNodes * agg[i];
  for (int slot = 0; slot < np; ++slot) {
    double t[3][3];
#pragma unroll
    for (int d1 = 0; d1 < 3; ++d1)
#pragma unroll
      for (int d2 = 0; d2 < 3; ++d2)
        t[d1][d2] = B[3 * d1] * pv[9 * slot + d2] + B[3 * d1 + 1] * pv[9 * slot + 3 + d2] +
                    B[3 * d1 + 2] * pv[9 * slot + 6 + d2];
    const int q = prow[slot] - own;  // this slot is coarse node q of i's own aggregate
#pragma unroll
    for (int d1 = 0; d1 < 3; ++d1)
#pragma unroll
      for (int d2 = 0; d2 < 3; ++d2) {
        double v = -omega * t[d1][d2];
        if (q >= 0 && q < kSaNodes) v += Ti[kSaModes * d1 + 3 * q + d2];
        pv[9 * slot + 3 * d1 + d2] = ffree[3 * i + d1] ? v : 0.0;
      }
  }
}

// AP = A_F P: one workgroup (64 threads) per fine row, accumulated in LDS by aggregate
// (the row's column list is staged in LDS after the accumulators for the slot searches).
__global__ __launch_bounds__(64) void k_sa_ap(int64_t nf, const int64_t* __restrict__ fptr,
                                              const int32_t* __restrict__ fidx, const double* __restrict__ fval,
                                              const int64_t* __restrict__ pptr, const int32_t* __restrict__ pidx,
                                              const double* __restrict__ pval, const int64_t* __restrict__ aptr,
                                              const int32_t* __restrict__ aidx, double* __restrict__ aval) {
  extern __shared__ double acc[];
  for (int64_t i = blockIdx.x; i < nf; i += gridDim.x) {
    const int64_t a0 = aptr[i];
    const int na = (int)(aptr[i + 1] - a0);
    int32_t* arow = reinterpret_cast<int32_t*>(acc + 9 * na);
    for (int k = threadIdx.x; k < 9 * na; k += blockDim.x) acc[k] = 0.0;
    for (int k = threadIdx.x; k < na; k += blockDim.x) arow[k] = aidx[a0 + k];
    __syncthreads();
    const int64_t f0 = fptr[i];
    const int degf = (int)(fptr[i + 1] - f0);
    const double* fv = fval + 9 * f0;
    for (int s = threadIdx.x; s < degf; s += blockDim.x) {
      const int32_t j = fidx[f0 + s];
      double A[3][3];
#pragma unroll
      for (int d1 = 0; d1 < 3; ++d1)
#pragma unroll
        for (int e = 0; e < 3; ++e) A[d1][e] = fv[3 * degf * d1 + 3 * s + e];
      for (int64_t k = pptr[j]; k < pptr[j + 1]; ++k) {
        const int slot = find_slot(arow, na, pidx[k]);
        if (slot < 0) continue;  // cannot happen: the AP graph holds every product
        const double* P = pval + 9 * k;
#pragma unroll
        for (int d1 = 0; d1 < 3; ++d1)
#pragma unroll
          for (int d2 = 0; d2 < 3; ++d2) {
            const double v = A[d1][0] * P[d2] + A[d1][1] * P[3 + d2] + A[d1][2] * P[6 + d2];
            if (v != 0.0) atomicAdd(acc + 9 * slot + 3 * d1 + d2, v);
          }
      }
    }
    __syncthreads();
    for (int k = threadIdx.x; k < 9 * na; k += blockDim.x) aval[9 * a0 + k] = acc[k];
    __syncthreads();
  }
}

// Coarse operator P^T (A P): one workgroup per coarse row I over the fine rows i with a
// block of P in column I (the transpose lists), written in the node-block layout;
// constrained coarse dofs (all-zero columns of P) get identity rows.  The row's column
// list is staged in LDS after the accumulators.
__global__ __launch_bounds__(64) void k_sa_ptap(int64_t nc, const int64_t* __restrict__ tptr,
                                                const int32_t* __restrict__ tidx, const int64_t* __restrict__ tpos,
                                                const double* __restrict__ pval, const int64_t* __restrict__ aptr,
                                                const int32_t* __restrict__ aidx, const double* __restrict__ aval,
                                                const int64_t* __restrict__ cptr, const int32_t* __restrict__ cidx,
                                                double* __restrict__ cval, const uint8_t* __restrict__ cfree) {
  extern __shared__ double acc[];
  for (int64_t I = blockIdx.x; I < nc; I += gridDim.x) {
    const int64_t c0 = cptr[I];
    const int degc = (int)(cptr[I + 1] - c0);
    int32_t* crow = reinterpret_cast<int32_t*>(acc + 9 * degc);
    for (int k = threadIdx.x; k < 9 * degc; k += blockDim.x) acc[k] = 0.0;
    for (int k = threadIdx.x; k < degc; k += blockDim.x) crow[k] = cidx[c0 + k];
    __syncthreads();
    for (int64_t q = tptr[I] + threadIdx.x; q < tptr[I + 1]; q += blockDim.x) {
      const int64_t i = tidx[q];
      const double* P = pval + 9 * tpos[q];  // block (i, I)
      for (int64_t a = aptr[i]; a < aptr[i + 1]; ++a) {
        const int slot = find_slot(crow, degc, aidx[a]);
        if (slot < 0) continue;  // cannot happen: the coarse graph holds every product
        const double* Q = aval + 9 * a;
#pragma unroll
        for (int d1 = 0; d1 < 3; ++d1)
#pragma unroll
          for (int d2 = 0; d2 < 3; ++d2) {
            const double v = P[d1] * Q[d2] + P[3 + d1] * Q[3 + d2] + P[6 + d1] * Q[6 + d2];  // (P^T Q)[d1][d2]
            if (v != 0.0) atomicAdd(acc + 3 * degc * d1 + 3 * slot + d2, v);
          }
      }
    }
    __syncthreads();
    const int self = find_slot(crow, degc, (int32_t)I);
    for (int k = threadIdx.x; k < 9 * degc; k += blockDim.x) {
      const int d1 = k / (3 * degc), r = k - 3 * degc * d1, slot = r / 3, d2 = r - 3 * slot;
      double v = acc[k];
      if (!cfree[3 * I + d1] || !cfree[3 * (int64_t)crow[slot] + d2]) v = 0.0;
      if (!cfree[3 * I + d1] && slot == self && d2 == d1) v = 1.0;
      cval[9 * c0 + k] = v;
    }
    __syncthreads();
  }
}

// FP32 copies of P for the V-cycle's transfers, one pass over the transpose lists: the
// prolongation reads P in its fine-row order, the restriction reads the same rounded
// blocks in coarse-row order (contiguous per coarse node), so restriction stays exactly
// the transpose of prolongation and the V-cycle stays symmetric.
__global__ __launch_bounds__(256) void k_sa_p32(int64_t nq, const int64_t* __restrict__ tpos,
                                                const double* __restrict__ pval, float* __restrict__ pval32,
                                                float* __restrict__ ptval32) {
  const int64_t q = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (q >= nq) return;
  const int64_t k = tpos[q];
#pragma unroll
  for (int e = 0; e < 9; ++e) {
    const float v = (float)pval[9 * k + e];
    pval32[9 * k + e] = v;
    ptval32[9 * q + e] = v;
  }
}

// b_c = P^T r: one wave per coarse node over its transpose list (its blocks contiguous
// in ptval32), shuffle reduction.
__global__ __launch_bounds__(256) void k_sa_restrict(int64_t nc, const int64_t* __restrict__ tptr,
                                                     const int32_t* __restrict__ tidx,
                                                     const float* __restrict__ ptval32,
                                                     const double* __restrict__ rf, const uint8_t* __restrict__ cfree,
                                                     double* __restrict__ bc) {
  const int64_t I = (blockIdx.x * (int64_t)blockDim.x + threadIdx.x) >> 6;
  const int lane = threadIdx.x & 63;
  double acc[3] = {0.0, 0.0, 0.0};
  if (I < nc)
    for (int64_t q = tptr[I] + lane; q < tptr[I + 1]; q += 64) {
      const int64_t i = tidx[q];
      const float* P = ptval32 + 9 * q;
      const double r0 = rf[3 * i], r1 = rf[3 * i + 1], r2 = rf[3 * i + 2];
#pragma unroll
      for (int d = 0; d < 3; ++d) acc[d] += (double)P[d] * r0 + (double)P[3 + d] * r1 + (double)P[6 + d] * r2;
    }
#pragma unroll
  for (int o = 32; o > 0; o >>= 1)
#pragma unroll
    for (int d = 0; d < 3; ++d) acc[d] += __shfl_xor(acc[d], o, 64);
  if (I >= nc || lane != 0) return;
#pragma unroll
  for (int d = 0; d < 3; ++d) bc[3 * I + d] = cfree[3 * I + d] ? acc[d] : 0.0;
}

// x_f += P x_c (rows of constrained fine dofs are zero).
template <int G>  // lanes per fine row
__global__ __launch_bounds__(256) void k_sa_prolong(int64_t nf, const int64_t* __restrict__ pptr,
                                                    const int32_t* __restrict__ pidx, const float* __restrict__ pval,
                                                    const double* __restrict__ xc, double* __restrict__ xf) {
  const int64_t i = (blockIdx.x * (int64_t)blockDim.x + threadIdx.x) / G;
  const int lane = threadIdx.x % G;
  double s[3] = {0.0, 0.0, 0.0};
  if (i < nf) {  // the row's blocks are contiguous: the G lanes stream its 9 np values
    const int64_t k0 = pptr[i];
    const int L = 9 * (int)(pptr[i + 1] - k0);
    const float* P = pval + 9 * k0;
    for (int q = lane; q < L; q += G) {
      const int k = q / 9, r = q - 9 * k, d1 = r / 3, d2 = r - 3 * d1;
      const double v = (double)P[q] * xc[3 * (int64_t)pidx[k0 + k] + d2];
      s[0] += d1 == 0 ? v : 0.0;
      s[1] += d1 == 1 ? v : 0.0;
      s[2] += d1 == 2 ? v : 0.0;
    }
  }
#pragma unroll
  for (int o = G / 2; o > 0; o >>= 1)
#pragma unroll
    for (int d = 0; d < 3; ++d) s[d] += __shfl_xor(s[d], o, G);
  if (i >= nf || lane != 0) return;
#pragma unroll
  for (int d = 0; d < 3; ++d) xf[3 * i + d] += s[d];
}

// Galerkin product on node blocks: one workgroup (64 threads) per coarse node row,
// accumulated in LDS, then written in the node-block layout; constrained coarse dofs
// get identity rows (their rows and columns are masked to zero).
__global__ __launch_bounds__(64) void k_mg_galerkin(int64_t Pf, int64_t Pc, const int64_t* __restrict__ fptr,
                                                    const int32_t* __restrict__ fidx,
                                                    const double* __restrict__ fval,
                                                    const uint8_t* __restrict__ ffree,
                                                    const int64_t* __restrict__ cptr,
                                                    const int32_t* __restrict__ cidx, double* __restrict__ cval,
                                                    const uint8_t* __restrict__ cfree) {
  extern __shared__ double acc[];
  const int64_t nc = Pc * Pc * Pc;
  for (int64_t I = blockIdx.x; I < nc; I += gridDim.x) {
    const int64_t c0 = cptr[I];
    const int degc = (int)(cptr[I + 1] - c0);
    const int32_t* crow = cidx + c0;
    for (int k = threadIdx.x; k < 9 * degc; k += blockDim.x) acc[k] = 0.0;
    __syncthreads();
    int64_t ch[15];
    double wc[15];
    const int nch = children(I, Pc, Pf, ch, wc);
    uint8_t cf[3];
#pragma unroll
    for (int d = 0; d < 3; ++d) cf[d] = cfree[3 * I + d];
    for (int a = 0; a < nch; ++a) {
      const int64_t f = ch[a];
      const int64_t f0 = fptr[f];
      const int degf = (int)(fptr[f + 1] - f0);
      uint8_t fr[3];
#pragma unroll
      for (int d = 0; d < 3; ++d) fr[d] = ffree[3 * f + d] & cf[d];
      const double* fv = fval + 9 * f0;
      for (int k = threadIdx.x; k < degf; k += blockDim.x) {
        const int32_t j = fidx[f0 + k];
        int64_t par[2];
        double wj[2];
        const int np = parents(j, Pf, Pc, par, wj);
        double blk[3][3];
#pragma unroll
        for (int d1 = 0; d1 < 3; ++d1)
#pragma unroll
          for (int d2 = 0; d2 < 3; ++d2)
            blk[d1][d2] = (fr[d1] && ffree[3 * (int64_t)j + d2]) ? wc[a] * fv[3 * degf * d1 + 3 * k + d2] : 0.0;
        for (int p = 0; p < np; ++p) {
          const int slot = find_slot(crow, degc, (int32_t)par[p]);
          if (slot < 0) continue;  // cannot happen: the coarse graph holds every coupling
#pragma unroll
          for (int d1 = 0; d1 < 3; ++d1)
#pragma unroll
            for (int d2 = 0; d2 < 3; ++d2) {
              const double v = cfree[3 * par[p] + d2] ? wj[p] * blk[d1][d2] : 0.0;
              if (v != 0.0) atomicAdd(acc + 3 * degc * d1 + 3 * slot + d2, v);
            }
        }
      }
      __syncthreads();
    }
    const int self = find_slot(crow, degc, (int32_t)I);
    for (int k = threadIdx.x; k < 9 * degc; k += blockDim.x) {
      const int d1 = k / (3 * degc), q = k - 3 * degc * d1, slot = q / 3, d2 = q - 3 * slot;
      double v = acc[k];
      if (!cf[d1] && slot == self && d2 == d1) v = 1.0;
      cval[9 * c0 + k] = v;
    }
    __syncthreads();
  }
}

__global__ __launch_bounds__(256) void k_mg_block_diag(int64_t nn, const int64_t* __restrict__ nptr,
                                                       const int32_t* __restrict__ nidx,
                                                       const double* __restrict__ vals, double* __restrict__ binv) {
  const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (i >= nn) return;
  const int64_t k0 = nptr[i];
  const int deg = (int)(nptr[i + 1] - k0);
  const int self = find_slot(nidx + k0, deg, (int32_t)i);
  double a[3][3];
#pragma unroll
  for (int d1 = 0; d1 < 3; ++d1)
#pragma unroll
    for (int d2 = 0; d2 < 3; ++d2) a[d1][d2] = self >= 0 ? vals[9 * k0 + 3 * deg * d1 + 3 * self + d2] : (d1 == d2);
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

// Dense copy of a small node-block matrix (n = 3 nn), row-major.
__global__ __launch_bounds__(256) void k_mg_densify(int64_t nn, const int64_t* __restrict__ nptr,
                                                    const int32_t* __restrict__ nidx,
                                                    const double* __restrict__ vals, double* __restrict__ D) {
  const int64_t I = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (I >= nn) return;
  const int64_t n = 3 * nn, k0 = nptr[I];
  const int deg = (int)(nptr[I + 1] - k0);
  for (int d1 = 0; d1 < 3; ++d1)
    for (int s = 0; s < deg; ++s) {
      const int64_t J = nidx[k0 + s];
      for (int d2 = 0; d2 < 3; ++d2) D[(3 * I + d1) * n + 3 * J + d2] = vals[9 * k0 + 3 * deg * d1 + 3 * s + d2];
    }
}

// One block Gauss-Jordan step of the inverse of an SPD n x n matrix (n = 3 nn), pivot
// block K = rows/columns 3k..3k+2, from the old matrix A into B; after nn steps B = A^-1.
//   P = A_KK^-1;  B_KK = P;  B_Kj = P A_Kj;  B_iK = -A_iK P;  B_ij = A_ij - A_iK P A_Kj.
__global__ __launch_bounds__(256) void k_mg_gj_step(int n, int k, const double* __restrict__ A,
                                                    double* __restrict__ B) {
  const int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (e >= (int64_t)n * n) return;
  const int i = (int)(e / n), j = (int)(e - (int64_t)i * n);
  const int K0 = 3 * k;
  double a[3][3];
#pragma unroll
  for (int r = 0; r < 3; ++r)
#pragma unroll
    for (int q = 0; q < 3; ++q) a[r][q] = A[(int64_t)(K0 + r) * n + K0 + q];
  const double det = a[0][0] * (a[1][1] * a[2][2] - a[1][2] * a[2][1]) -
                     a[0][1] * (a[1][0] * a[2][2] - a[1][2] * a[2][0]) +
                     a[0][2] * (a[1][0] * a[2][1] - a[1][1] * a[2][0]);
  const double id = 1.0 / det;
  double P[3][3];
  P[0][0] = (a[1][1] * a[2][2] - a[1][2] * a[2][1]) * id;
  P[0][1] = (a[0][2] * a[2][1] - a[0][1] * a[2][2]) * id;
  P[0][2] = (a[0][1] * a[1][2] - a[0][2] * a[1][1]) * id;
  P[1][0] = (a[1][2] * a[2][0] - a[1][0] * a[2][2]) * id;
  P[1][1] = (a[0][0] * a[2][2] - a[0][2] * a[2][0]) * id;
  P[1][2] = (a[0][2] * a[1][0] - a[0][0] * a[1][2]) * id;
  P[2][0] = (a[1][0] * a[2][1] - a[1][1] * a[2][0]) * id;
  P[2][1] = (a[0][1] * a[2][0] - a[0][0] * a[2][1]) * id;
  P[2][2] = (a[0][0] * a[1][1] - a[0][1] * a[1][0]) * id;
  const bool iK = i >= K0 && i < K0 + 3, jK = j >= K0 && j < K0 + 3;
  double v;
  if (iK && jK) {
    v = P[i - K0][j - K0];
  } else if (iK) {  // P A_Kj
    v = P[i - K0][0] * A[(int64_t)K0 * n + j] + P[i - K0][1] * A[(int64_t)(K0 + 1) * n + j] +
        P[i - K0][2] * A[(int64_t)(K0 + 2) * n + j];
  } else if (jK) {  // -A_iK P
    const double* ai = A + (int64_t)i * n + K0;
    v = -(ai[0] * P[0][j - K0] + ai[1] * P[1][j - K0] + ai[2] * P[2][j - K0]);
  } else {
    const double* ai = A + (int64_t)i * n + K0;
    double t[3];
#pragma unroll
    for (int q = 0; q < 3; ++q) t[q] = ai[0] * P[0][q] + ai[1] * P[1][q] + ai[2] * P[2][q];
    v = A[e] - (t[0] * A[(int64_t)K0 * n + j] + t[1] * A[(int64_t)(K0 + 1) * n + j] +
                t[2] * A[(int64_t)(K0 + 2) * n + j]);
  }
  B[e] = v;
}

// Inverse of a small SPD matrix (n <= FEDD_MG_LDS_N, n = 3 nn) in one workgroup: the
// symmetric sweep operator with 3x3 pivot blocks on the packed lower triangle held in
// LDS, (i, j) with i >= j at i (i + 1) / 2 + j.  Sweeping pivot block K with S = A_KK^-1:
//   A_KK <- -S;  A_iK <- A_iK S;  A_Kj <- S A_Kj;  A_ij <- A_ij - A_iK S A_Kj.
// After every block has been swept the matrix holds -A^-1 (Goodnight, "A tutorial on
// the SWEEP operator", The American Statistician 33, 1979); it is written out negated
// as the full row-major n x n matrix.
#define FEDD_MG_LDS_N 192
__global__ __launch_bounds__(1024) void k_mg_sweep_inverse(int n, const double* __restrict__ A,
                                                           double* __restrict__ inv) {
  constexpr int NP = FEDD_MG_LDS_N * (FEDD_MG_LDS_N + 1) / 2;
  constexpr int PER = (NP + 1023) / 1024;
  __shared__ double T[NP];
  __shared__ double colK[3 * FEDD_MG_LDS_N];
  __shared__ double S[9];
  const int tid = threadIdx.x;
  const int np = n * (n + 1) / 2;
  uint32_t ij[PER];
#pragma unroll
  for (int m = 0; m < PER; ++m) {
    const int e = tid + 1024 * m;
    ij[m] = 0;
    if (e < np) {
      int i = (int)((sqrt(8.0 * e + 1.0) - 1.0) * 0.5);
      while (i * (i + 1) / 2 > e) --i;
      while ((i + 1) * (i + 2) / 2 <= e) ++i;
      const int j = e - i * (i + 1) / 2;
      ij[m] = ((uint32_t)i << 16) | (uint32_t)j;
      T[e] = A[(int64_t)i * n + j];
    }
  }
  auto get = [&](int i, int j) -> double { return i >= j ? T[i * (i + 1) / 2 + j] : T[j * (j + 1) / 2 + i]; };
  for (int k = 0; k < n / 3; ++k) {
    const int K0 = 3 * k;
    __syncthreads();
    if (tid < n) {
#pragma unroll
      for (int q = 0; q < 3; ++q) colK[3 * tid + q] = get(tid, K0 + q);
    }
    if (tid == 0) {
      double a[3][3];
#pragma unroll
      for (int r = 0; r < 3; ++r)
#pragma unroll
        for (int q = 0; q < 3; ++q) a[r][q] = get(K0 + r, K0 + q);
      const double c00 = a[1][1] * a[2][2] - a[1][2] * a[2][1];
      const double c01 = a[1][2] * a[2][0] - a[1][0] * a[2][2];
      const double c02 = a[1][0] * a[2][1] - a[1][1] * a[2][0];
      const double id = 1.0 / (a[0][0] * c00 + a[0][1] * c01 + a[0][2] * c02);
      S[0] = c00 * id;
      S[1] = (a[0][2] * a[2][1] - a[0][1] * a[2][2]) * id;
      S[2] = (a[0][1] * a[1][2] - a[0][2] * a[1][1]) * id;
      S[3] = c01 * id;
      S[4] = (a[0][0] * a[2][2] - a[0][2] * a[2][0]) * id;
      S[5] = (a[0][2] * a[1][0] - a[0][0] * a[1][2]) * id;
      S[6] = c02 * id;
      S[7] = (a[0][1] * a[2][0] - a[0][0] * a[2][1]) * id;
      S[8] = (a[0][0] * a[1][1] - a[0][1] * a[1][0]) * id;
    }
    __syncthreads();
    double s[9];
#pragma unroll
    for (int r = 0; r < 9; ++r) s[r] = S[r];
#pragma unroll
    for (int m = 0; m < PER; ++m) {
      const int e = tid + 1024 * m;
      if (e >= np) continue;
      const int i = (int)(ij[m] >> 16), j = (int)(ij[m] & 0xffffu);
      const bool iK = (unsigned)(i - K0) < 3u, jK = (unsigned)(j - K0) < 3u;
      const double* ci = colK + 3 * i;
      const double* cj = colK + 3 * j;
      double v;
      if (iK && jK) {
        v = -S[3 * (i - K0) + (j - K0)];
      } else if (jK) {  // (A_iK S)_q
        const int q = j - K0;
        v = ci[0] * S[q] + ci[1] * S[3 + q] + ci[2] * S[6 + q];
      } else if (iK) {  // (S A_Kj)_r
        const int r = i - K0;
        v = S[3 * r] * cj[0] + S[3 * r + 1] * cj[1] + S[3 * r + 2] * cj[2];
      } else {
        const double t0 = ci[0] * s[0] + ci[1] * s[3] + ci[2] * s[6];
        const double t1 = ci[0] * s[1] + ci[1] * s[4] + ci[2] * s[7];
        const double t2 = ci[0] * s[2] + ci[1] * s[5] + ci[2] * s[8];
        v = T[e] - (t0 * cj[0] + t1 * cj[1] + t2 * cj[2]);
      }
      T[e] = v;
    }
  }
  __syncthreads();
#pragma unroll
  for (int m = 0; m < PER; ++m) {
    const int e = tid + 1024 * m;
    if (e >= np) continue;
    const int i = (int)(ij[m] >> 16), j = (int)(ij[m] & 0xffffu);
    inv[(int64_t)i * n + j] = -T[e];
    inv[(int64_t)j * n + i] = -T[e];
  }
}

__global__ __launch_bounds__(256) void k_mg_fill_zero(int64_t n, double* __restrict__ y) {
  const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (i < n) y[i] = 0.0;
}

__global__ __launch_bounds__(256) void k_mg_dense_mv(int64_t n, const double* __restrict__ A,
                                                     const double* __restrict__ x, double* __restrict__ y) {
  const int64_t i = blockIdx.x * (int64_t)(blockDim.x / 64) + (threadIdx.x >> 6);
  const int lane = threadIdx.x & 63;
  double s = 0.0;
  if (i < n)
    for (int64_t j = lane; j < n; j += 64) s += A[i * n + j] * x[j];
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) s += __shfl_xor(s, o, 64);
  if (i < n && lane == 0) y[i] = s;
}

// ------------------------------------------------------------------ host side
// True when the primary mesh is the P2 structured lattice of fedd_mesh_structured:
// node id = r + s P + t P^2, 6 M^3 elements (P = 2 M + 1), vertices on even lattice
// points, edge nodes at the midpoints, and every edge a Kuhn edge +-e_S (nonzero
// components of one sign), as the prolongation and the coarse stencils assume.
static bool lattice_mesh(const Mesh& m, int64_t& P) {
  if (m.npe != 10 || m.npts < 27) return false;
  P = (int64_t)std::llround(std::cbrt((double)m.npts));
  if (P * P * P != m.npts || (P & 1) == 0) return false;
  const int64_t Mc = (P - 1) / 2;
  if (m.nelem != 6 * Mc * Mc * Mc) return false;
  std::atomic<bool> ok(true);
  parallel_for(m.nelem, [&](int64_t a, int64_t b, int) {
    const int ea[6] = {0, 1, 0, 0, 1, 2}, eb[6] = {1, 2, 2, 3, 3, 3};
    auto xyz = [&](int64_t i, int64_t x[3]) {
      x[0] = i % P;
      x[1] = (i / P) % P;
      x[2] = i / (P * P);
    };
    for (int64_t e = a; e < b && ok.load(std::memory_order_relaxed); ++e) {
      const int32_t* nd = &m.elems[e * 10];
      int64_t v[4][3];
      for (int k = 0; k < 4; ++k) {
        if (nd[k] < 0 || nd[k] >= m.npts) return (void)ok.store(false);
        xyz(nd[k], v[k]);
        if ((v[k][0] | v[k][1] | v[k][2]) & 1) return (void)ok.store(false);
      }
      for (int t = 0; t < 6; ++t) {
        if (nd[4 + t] < 0 || nd[4 + t] >= m.npts) return (void)ok.store(false);
        int64_t x[3];
        xyz(nd[4 + t], x);
        int pos = 0, neg = 0;
        for (int d = 0; d < 3; ++d) {
          const int64_t dv = v[eb[t]][d] - v[ea[t]][d];
          if (2 * x[d] != v[ea[t]][d] + v[eb[t]][d] || (dv != 0 && dv != 2 && dv != -2))
            return (void)ok.store(false);
          pos += dv > 0;
          neg += dv < 0;
        }
        if (pos && neg) return (void)ok.store(false);
      }
    }
  });
  return ok.load();
}

// Greedy aggregation of a node graph (rows include the node itself): pass 1 makes every
// node whose neighbours are all unaggregated the root of a new aggregate {i} + neighbours;
// pass 2 joins each remaining node to the aggregate of its first pass-1 neighbour; pass 3
// makes singletons of the rest.  Returns the number of aggregates.
static int64_t aggregate(int64_t n, const std::vector<int64_t>& ptr, const std::vector<int32_t>& idx,
                         const std::vector<uint8_t>* strong, std::vector<int32_t>& agg) {
  auto st = [&](int64_t k) { return !strong || (*strong)[k]; };
  agg.assign(n, -1);
  int64_t nc = 0;
  for (int64_t i = 0; i < n; ++i) {
    if (agg[i] >= 0) continue;
    bool free_nb = true;
    for (int64_t k = ptr[i]; k < ptr[i + 1] && free_nb; ++k) free_nb = !st(k) || agg[idx[k]] < 0;
    if (!free_nb) continue;
    for (int64_t k = ptr[i]; k < ptr[i + 1]; ++k)
      if (st(k)) agg[idx[k]] = (int32_t)nc;
    agg[i] = (int32_t)nc++;
  }
  std::vector<int32_t> first(agg);
  for (int64_t i = 0; i < n; ++i) {
    if (agg[i] >= 0) continue;
    for (int64_t k = ptr[i]; k < ptr[i + 1]; ++k)
      if (st(k) && first[idx[k]] >= 0) {
        agg[i] = first[idx[k]];
        break;
      }
  }
  for (int64_t i = 0; i < n; ++i)
    if (agg[i] < 0) agg[i] = (int32_t)nc++;
  return nc;
}

#define FEDD_AMG_COARSEST 64  // nodes on the coarsest aggregation level (at most)
// Damping of the prolongator smoothing P = (I - omega D^-1 A) P_tent.
#ifndef FEDD_SA_OMEGA
#define FEDD_SA_OMEGA 0.67
#endif
// Levels whose average row holds at most this many blocks aggregate over strong couplings
// only.  Off: with the six rigid-body modes, aggregating the P2 level over its strong
// couplings only (threshold 64) gives aggregates of 14 nodes and a first coarse level with
// more blocks than the finest (4.2M vs 3.3M at 24^3x6; 267 CG iterations, 0.43 s); over
// all couplings, 56 nodes and 0.54M blocks (333 iterations, 0.24 s).
#ifndef FEDD_AMG_FILTER_DEG
#define FEDD_AMG_FILTER_DEG 0
#endif
// Strong coupling: ||A_ij||_F >= theta sqrt(||A_ii||_F ||A_jj||_F).
#ifndef FEDD_AMG_THETA
#define FEDD_AMG_THETA 0.1
#endif

// Device arrays of level l (l = 0 uses the context's operator) and its Dirichlet mask;
// the host copies of the level's graphs are released afterwards.
static int level_device_setup(Ctx* c, MgLevel& L, size_t l, int mode, const std::vector<uint8_t>& fr) {
  int rc;
  if (l == 0) {
    L.d_nptr = c->d_nptr.as<int64_t>();
    L.d_nidx = c->d_nidx.as<int32_t>();
    L.d_vals = c->blk[0].d_vals.as<double>();
  } else {
    if ((rc = upload(c, L.nptr, L.h_ptr.data(), L.h_ptr.size() * 8))) return rc;
    if ((rc = upload(c, L.nidx, L.h_idx.data(), L.h_idx.size() * 4))) return rc;
    if ((rc = alloc(c, L.vals, L.nblk * 9 * 8))) return rc;
    L.d_nptr = L.nptr.as<int64_t>();
    L.d_nidx = L.nidx.as<int32_t>();
    L.d_vals = L.vals.as<double>();
    if (mode == 1) {
      L.nblk_p = (int64_t)L.h_pidx.size();
      if ((rc = upload(c, L.agg, L.h_agg.data(), L.h_agg.size() * 4)) ||
          (rc = upload(c, L.tent, L.h_tent.data(), L.h_tent.size() * 8)) ||
          (rc = upload(c, L.pptr, L.h_pptr.data(), L.h_pptr.size() * 8)) ||
          (rc = upload(c, L.pidx, L.h_pidx.data(), std::max<size_t>(L.h_pidx.size(), 1) * 4)) ||
          (rc = alloc(c, L.pval, std::max<size_t>(L.h_pidx.size(), 1) * 9 * 8)) ||
          (rc = alloc(c, L.pval32, std::max<size_t>(L.h_pidx.size(), 1) * 9 * 4)) ||
          (rc = alloc(c, L.ptval32, std::max<size_t>(L.h_pidx.size(), 1) * 9 * 4)) ||
          (rc = upload(c, L.tptr, L.h_tptr.data(), L.h_tptr.size() * 8)) ||
          (rc = upload(c, L.tidx, L.h_tidx.data(), std::max<size_t>(L.h_tidx.size(), 1) * 4)) ||
          (rc = upload(c, L.tpos, L.h_tpos.data(), std::max<size_t>(L.h_tpos.size(), 1) * 8)) ||
          (rc = upload(c, L.aptr, L.h_aptr.data(), L.h_aptr.size() * 8)) ||
          (rc = upload(c, L.aidx, L.h_aidx.data(), std::max<size_t>(L.h_aidx.size(), 1) * 4)) ||
          (rc = alloc(c, L.aval, std::max<size_t>(L.h_aidx.size(), 1) * 9 * 8)))
        return rc;
    }
  }
  if ((rc = upload(c, L.free, fr.data(), fr.size()))) return rc;
  for (DevBuf* b : {&L.x, &L.x2, &L.b, &L.r})
    if ((rc = alloc(c, *b, L.nn * 3 * 8))) return rc;
  if ((rc = alloc(c, L.binv, L.nn * 9 * 8))) return rc;
  if (l == 0 && FEDD_MG_FP32_FINE && (rc = alloc(c, L.vals32, L.nblk * 9 * 4))) return rc;
  // the wide first coarse level of the aggregation hierarchy is smoothed from an FP32 copy too
  if (l > 0 && mode == 1 && L.nblk > 48 * L.nn && (rc = alloc(c, L.vals32, L.nblk * 9 * 4))) return rc;
  std::vector<int64_t>().swap(L.h_ptr);
  std::vector<int32_t>().swap(L.h_idx);
  std::vector<int32_t>().swap(L.h_agg);
  std::vector<double>().swap(L.h_tent);
  for (auto* v : {&L.h_pptr, &L.h_tptr, &L.h_tpos, &L.h_aptr}) std::vector<int64_t>().swap(*v);
  for (auto* v : {&L.h_pidx, &L.h_tidx, &L.h_aidx}) std::vector<int32_t>().swap(*v);
  return FEDD_OK;
}

// Smoothed-aggregation products for level l from level l-1 (its operator and block
// inverse current): P, A P, P^T A P.
static int sa_galerkin(Ctx* c, MgHost* g, size_t l) {
  const MgLevel& F = g->lv[l - 1];
  MgLevel& L = g->lv[l];
  k_sa_smooth_p<<<nbk(F.nn, 256), 256, 0, c->stream>>>(F.nn, F.d_nptr, F.d_nidx, F.d_vals, F.binv.as<double>(),
                                                      F.free.as<uint8_t>(), L.agg.as<int32_t>(), L.tent.as<double>(),
                                                      L.pptr.as<int64_t>(),
                                                      L.pidx.as<int32_t>(), L.pval.as<double>(), FEDD_SA_OMEGA);
  if (L.nblk_p > 0)
    k_sa_p32<<<nbk(L.nblk_p, 256), 256, 0, c->stream>>>(L.nblk_p, L.tpos.as<int64_t>(), L.pval.as<double>(),
                                                        L.pval32.as<float>(), L.ptval32.as<float>());
  const unsigned gf = (unsigned)std::min<int64_t>(F.nn, (int64_t)c->num_cu * 16);
  k_sa_ap<<<gf, 64, (size_t)L.maxdeg_ap * (9 * 8 + 4), c->stream>>>(F.nn, F.d_nptr, F.d_nidx, F.d_vals,
                                                               L.pptr.as<int64_t>(), L.pidx.as<int32_t>(),
                                                               L.pval.as<double>(), L.aptr.as<int64_t>(),
                                                               L.aidx.as<int32_t>(), L.aval.as<double>());
  const unsigned gc = (unsigned)std::min<int64_t>(L.nn, (int64_t)c->num_cu * 16);
  k_sa_ptap<<<gc, 64, (size_t)L.maxdeg * (9 * 8 + 4), c->stream>>>(
      L.nn, L.tptr.as<int64_t>(), L.tidx.as<int32_t>(), L.tpos.as<int64_t>(), L.pval.as<double>(),
      L.aptr.as<int64_t>(), L.aidx.as<int32_t>(), L.aval.as<double>(), L.d_nptr, L.d_nidx, L.d_vals,
      L.free.as<uint8_t>());
  FEDD_HIP(c, hipGetLastError());
  return FEDD_OK;
}

// Sorted, duplicate-free rows: row r holds every column that rowfn(r, emit) emits
// (columns < ncols).  Built over host threads with a stamp array per thread.
template <class RowFn>
static void union_rows(int64_t nrows, int64_t ncols, RowFn rowfn, std::vector<int64_t>& optr,
                       std::vector<int32_t>& oidx, int64_t* maxdeg) {
  const unsigned T = nthreads();
  std::vector<std::vector<int32_t>> part_idx(T);
  std::vector<std::vector<int64_t>> part_len(T);
  std::vector<int64_t> part_begin(T, nrows);
  parallel_for(nrows, [&](int64_t a, int64_t b, int t) {
    std::vector<int32_t> stamp(ncols, -1), row;
    part_begin[t] = a;
    for (int64_t r = a; r < b; ++r) {
      row.clear();
      rowfn(r, [&](int32_t col) {
        if (stamp[col] != (int32_t)r) {
          stamp[col] = (int32_t)r;
          row.push_back(col);
        }
      });
      std::sort(row.begin(), row.end());
      part_idx[t].insert(part_idx[t].end(), row.begin(), row.end());
      part_len[t].push_back((int64_t)row.size());
    }
  });
  optr.assign(nrows + 1, 0);
  oidx.clear();
  for (unsigned t = 0; t < T; ++t) {
    int64_t r = part_begin[t];
    for (int64_t l : part_len[t]) {
      optr[r + 1] = optr[r] + l;
      *maxdeg = std::max(*maxdeg, l);
      ++r;
    }
    oidx.insert(oidx.end(), part_idx[t].begin(), part_idx[t].end());
  }
}

// Tentative prolongator of one aggregation step.  nsp holds, per fine node, the 3 x
// kSaModes rows of the near-null-space modes (level 0: translations and rotations about
// the mesh centroid; below: the R factors of the step above).  Per aggregate J the rows
// of its members' free dofs are stacked into a (3 n_J) x kSaModes matrix and factored
// V = Q R by modified Gram-Schmidt with one reorthogonalisation; a column whose remainder
// falls below 1e-8 of its norm (a singleton's rotations, a fully constrained aggregate)
// is dropped: its Q column and R row are zero and its coarse dof is marked constrained.
// tent gets Q's rows, the coarse nodes kSaNodes J + q get R's rows 3q..3q+2 as their
// modes (T R = V: the coarse space reproduces the fine level's modes exactly).
static void sa_tentative(int64_t nf, int64_t nagg, const std::vector<int32_t>& agg, const std::vector<double>& nsp,
                         const std::vector<uint8_t>& ffree, std::vector<double>& tent, std::vector<double>& cnsp,
                         std::vector<uint8_t>& cfree) {
  constexpr int W = 3 * kSaModes;  // doubles per node
  std::vector<int64_t> mptr(nagg + 1, 0);
  for (int64_t i = 0; i < nf; ++i) mptr[agg[i] + 1]++;
  for (int64_t J = 0; J < nagg; ++J) mptr[J + 1] += mptr[J];
  std::vector<int32_t> mem(nf);
  {
    std::vector<int64_t> fill(mptr.begin(), mptr.end() - 1);
    for (int64_t i = 0; i < nf; ++i) mem[fill[agg[i]]++] = (int32_t)i;
  }
  tent.assign(nf * W, 0.0);
  cnsp.assign(nagg * kSaNodes * W, 0.0);
  cfree.assign(nagg * kSaNodes * 3, 0);
  parallel_for(nagg, [&](int64_t a, int64_t b, int) {
    std::vector<double> V;
    for (int64_t J = a; J < b; ++J) {
      const int64_t m0 = mptr[J];
      const int rows = (int)(3 * (mptr[J + 1] - m0));
      V.assign((size_t)rows * kSaModes, 0.0);  // column-major: V[k * rows + r]
      for (int r = 0; r < rows; ++r) {
        const int64_t i = mem[m0 + r / 3];
        const int d = r % 3;
        if (!ffree[3 * i + d]) continue;
        for (int k = 0; k < kSaModes; ++k) V[(size_t)k * rows + r] = nsp[W * i + kSaModes * d + k];
      }
      double R[kSaModes][kSaModes] = {};
      for (int k = 0; k < kSaModes; ++k) {
        double* v = V.data() + (size_t)k * rows;
        double n0 = 0.0;
        for (int r = 0; r < rows; ++r) n0 += v[r] * v[r];
        n0 = std::sqrt(n0);
        for (int pass = 0; pass < 2; ++pass)
          for (int j = 0; j < k; ++j) {
            const double* qj = V.data() + (size_t)j * rows;
            double dot = 0.0;
            for (int r = 0; r < rows; ++r) dot += qj[r] * v[r];
            R[j][k] += dot;
            for (int r = 0; r < rows; ++r) v[r] -= dot * qj[r];
          }
        double n1 = 0.0;
        for (int r = 0; r < rows; ++r) n1 += v[r] * v[r];
        n1 = std::sqrt(n1);
        if (n0 > 0.0 && n1 > 1e-8 * n0) {
          for (int r = 0; r < rows; ++r) v[r] /= n1;
          R[k][k] = n1;
        } else {
          for (int r = 0; r < rows; ++r) v[r] = 0.0;  // V_k = sum_j<k q_j R[j][k]: kept in R's column k
        }
      }
      for (int r = 0; r < rows; ++r) {
        const int64_t i = mem[m0 + r / 3];
        const int d = r % 3;
        for (int k = 0; k < kSaModes; ++k) tent[W * i + kSaModes * d + k] = V[(size_t)k * rows + r];
      }
      for (int q = 0; q < kSaNodes; ++q) {
        const int64_t I = kSaNodes * J + q;
        for (int d = 0; d < 3; ++d) {
          cfree[3 * I + d] = R[3 * q + d][3 * q + d] != 0.0;
          for (int k = 0; k < kSaModes; ++k) cnsp[W * I + kSaModes * d + k] = R[3 * q + d][k];
        }
      }
    }
  });
}

// Aggregation hierarchy from the node graph of level 0 (any mesh).
static int build_aggregation(Ctx* c, MgHost* g, std::vector<std::vector<uint8_t>>& frees) {
  const int64_t n0 = c->mesh[0].npts;
  std::vector<int64_t> ptr(n0 + 1);
  FEDD_HIP(c, hipMemcpy(ptr.data(), c->d_nptr.p, (n0 + 1) * 8, hipMemcpyDeviceToHost));
  std::vector<int32_t> idx(ptr[n0]);
  FEDD_HIP(c, hipMemcpy(idx.data(), c->d_nidx.p, idx.size() * 4, hipMemcpyDeviceToHost));
  // Strength of the couplings of the current finest-so-far level F, from its operator on
  // the device (level 0: the context's matrix; coarser: the products just formed).
  auto strength = [&](const MgLevel& F, std::vector<uint8_t>& strong) -> int {
    const int64_t nb = (int64_t)idx.size();
    DevBuf d_nrm;
    int rc = alloc(c, d_nrm, std::max<int64_t>(nb, 1) * 4);
    if (rc) return rc;
    k_amg_block_norm<<<nbk(F.nn, 256), 256, 0, c->stream>>>(F.nn, F.d_nptr, F.d_vals, d_nrm.as<float>());
    FEDD_HIP(c, hipGetLastError());
    std::vector<float> nrm(nb);
    FEDD_HIP(c, hipMemcpyAsync(nrm.data(), d_nrm.p, nb * 4, hipMemcpyDeviceToHost, c->stream));
    FEDD_HIP(c, hipStreamSynchronize(c->stream));
    std::vector<float> diag(F.nn, 0.f);
    for (int64_t i = 0; i < F.nn; ++i)
      for (int64_t k = ptr[i]; k < ptr[i + 1]; ++k)
        if (idx[k] == i) diag[i] = nrm[k];
    strong.assign(nb, 1);
    for (int64_t i = 0; i < F.nn; ++i)
      for (int64_t k = ptr[i]; k < ptr[i + 1]; ++k)
        strong[k] = idx[k] == i || nrm[k] >= (float)FEDD_AMG_THETA * std::sqrt(diag[i] * diag[idx[k]]);
    return FEDD_OK;
  };
  // near-null-space modes of level 0: translations and rotations about the centroid
  std::vector<double> nsp((size_t)n0 * 3 * kSaModes, 0.0);
  {
    const std::vector<double>& X = c->mesh[0].pts;
    double cen[3] = {0.0, 0.0, 0.0};
    for (int64_t i = 0; i < n0; ++i)
      for (int d = 0; d < 3; ++d) cen[d] += X[3 * i + d] / (double)n0;
    for (int64_t i = 0; i < n0; ++i) {
      double* B = nsp.data() + 3 * kSaModes * i;  // B[kSaModes * d + k]
      for (int d = 0; d < 3; ++d) B[kSaModes * d + d] = 1.0;
      if (kSaModes == 6) {
        const double x = X[3 * i] - cen[0], y = X[3 * i + 1] - cen[1], z = X[3 * i + 2] - cen[2];
        // rotations about the x, y and z axes: omega x (x, y, z)
        B[kSaModes * 1 + 3] = -z, B[kSaModes * 2 + 3] = y;
        B[kSaModes * 0 + 4] = z, B[kSaModes * 2 + 4] = -x;
        B[kSaModes * 0 + 5] = -y, B[kSaModes * 1 + 5] = x;
      }
    }
  }
  while (g->lv.back().nn > FEDD_AMG_COARSEST && g->lv.size() < 16) {
    const MgLevel& F = g->lv.back();
    int rc;
    // F's block inverse (for the prolongator smoothing) and the strength of its couplings
    k_mg_block_diag<<<nbk(F.nn, 256), 256, 0, c->stream>>>(F.nn, F.d_nptr, F.d_nidx, F.d_vals, F.binv.as<double>());
    FEDD_HIP(c, hipGetLastError());
    // Levels with wide rows (the smoothed products widen the coarse stencils) are
    // aggregated over all their couplings, which keeps the hierarchy short.
    std::vector<uint8_t> strong;
    const bool filter = (int64_t)idx.size() <= FEDD_AMG_FILTER_DEG * F.nn;
    if (filter && (rc = strength(F, strong))) return rc;
    std::vector<int32_t> agg;
    const int64_t nagg = aggregate(F.nn, ptr, idx, filter ? &strong : nullptr, agg);
    const int64_t nc = kSaNodes * nagg;
    if (nc >= F.nn) break;  // no coarsening left
    MgLevel C;
    C.nn = nc;
    // prolongator graph: row i holds the coarse nodes of the aggregates of i's neighbours
    int64_t md = 0;
    union_rows(F.nn, nc, [&](int64_t i, auto&& emit) {
      for (int64_t k = ptr[i]; k < ptr[i + 1]; ++k)
        for (int q = 0; q < kSaNodes; ++q) emit(kSaNodes * agg[idx[k]] + q);
    }, C.h_pptr, C.h_pidx, &md);
    // its transpose
    C.h_tptr.assign(nc + 1, 0);
    for (int32_t J : C.h_pidx) C.h_tptr[J + 1]++;
    for (int64_t I = 0; I < nc; ++I) C.h_tptr[I + 1] += C.h_tptr[I];
    C.h_tidx.resize(C.h_pidx.size());
    C.h_tpos.resize(C.h_pidx.size());
    {
      std::vector<int64_t> fill(C.h_tptr.begin(), C.h_tptr.end() - 1);
      for (int64_t i = 0; i < F.nn; ++i)
        for (int64_t k = C.h_pptr[i]; k < C.h_pptr[i + 1]; ++k) {
          const int64_t q = fill[C.h_pidx[k]]++;
          C.h_tidx[q] = (int32_t)i;
          C.h_tpos[q] = k;
        }
    }
    // A P graph: row i holds the P columns of i's neighbours
    union_rows(F.nn, nc, [&](int64_t i, auto&& emit) {
      for (int64_t k = ptr[i]; k < ptr[i + 1]; ++k) {
        const int32_t j = idx[k];
        for (int64_t q = C.h_pptr[j]; q < C.h_pptr[j + 1]; ++q) emit(C.h_pidx[q]);
      }
    }, C.h_aptr, C.h_aidx, &C.maxdeg_ap);
    // coarse graph: row I holds the A P columns of every fine row with a block in column I
    union_rows(nc, nc, [&](int64_t I, auto&& emit) {
      for (int64_t q = C.h_tptr[I]; q < C.h_tptr[I + 1]; ++q) {
        const int32_t i = C.h_tidx[q];
        for (int64_t a = C.h_aptr[i]; a < C.h_aptr[i + 1]; ++a) emit(C.h_aidx[a]);
      }
    }, C.h_ptr, C.h_idx, &C.maxdeg);
    C.nblk = C.h_ptr[nc];
    if ((9 * 8 + 4) * std::max(C.maxdeg, C.maxdeg_ap) > 65536)
      return fail(c, FEDD_ERR_LIMIT, "aggregation: a coarse row exceeds the LDS budget");
    // tentative prolongator, the coarse level's modes and its constrained dofs
    std::vector<uint8_t> cf;
    std::vector<double> cnsp;
    sa_tentative(F.nn, nagg, agg, nsp, frees.back(), C.h_tent, cnsp, cf);
    nsp = std::move(cnsp);
    frees.push_back(std::move(cf));
    C.h_agg = std::move(agg);
    ptr = C.h_ptr;
    idx = C.h_idx;
    g->lv.push_back(std::move(C));
    const size_t l = g->lv.size() - 1;
    if ((rc = level_device_setup(c, g->lv[l], l, 1, frees[l]))) return rc;
    if ((rc = sa_galerkin(c, g, l))) return rc;
  }
  if (g->lv.back().nn > 2048) return fail(c, FEDD_ERR_LIMIT, "aggregation stalled above 2048 coarsest nodes");
  return FEDD_OK;
}

static int mg_setup_levels(Ctx* c, MgHost* g, int mode, int64_t P0);

// A failed setup leaves no hierarchy behind (the next call starts over).
int mg_setup(Ctx* c, int mode) {
  if (c->mg && c->mg->mode == mode) return FEDD_OK;
  mg_free(c);
  int64_t P0 = 0;
  if (mode == 0 && !lattice_mesh(c->mesh[0], P0))
    return fail(c, FEDD_ERR_STATE, "multigrid needs the structured P2 lattice mesh of fedd_mesh_structured");
  MgHost* g = new MgHost();
  c->mg = g;
  g->mode = mode;
  const int rc = mg_setup_levels(c, g, mode, P0);
  if (rc) mg_free(c);
  return rc;
}

static int mg_setup_levels(Ctx* c, MgHost* g, int mode, int64_t P0) {
  const Mesh& m = c->mesh[0];
  // level 0: the context's operator (9 values per node block)
  {
    MgLevel L0;
    L0.P = P0;
    L0.nn = m.npts;
    L0.nblk = c->blk[0].nnz / 9;
    g->lv.push_back(std::move(L0));
  }
  std::vector<std::vector<uint8_t>> frees(1, std::vector<uint8_t>(3 * m.npts, 1));
  for (size_t k = 0; k < c->bcflag_host.size() && k < frees[0].size(); ++k) frees[0][k] = c->bcflag_host[k] ? 0 : 1;
  int rc;
  if (mode == 1) {
    if ((rc = level_device_setup(c, g->lv[0], 0, mode, frees[0]))) return rc;
    if ((rc = build_aggregation(c, g, frees))) return rc;
  } else {
    // Coarse levels: P1 lattices.  The Galerkin product of the midpoint prolongation on the
    // Kuhn lattice couples a coarse node exactly with itself and its 14 Kuhn neighbours
    // I +- e_S (the edges of the P1 Kuhn mesh), from the P2 level as from every P1 level.
    static const int kuhn[15][3] = {{0, 0, 0},  {1, 0, 0},  {-1, 0, 0},  {0, 1, 0},   {0, -1, 0},
                                    {0, 0, 1},  {0, 0, -1}, {1, 1, 0},   {-1, -1, 0}, {1, 0, 1},
                                    {-1, 0, -1}, {0, 1, 1}, {0, -1, -1}, {1, 1, 1},   {-1, -1, -1}};
    while (g->lv.back().P >= 5 && ((g->lv.back().P - 1) & 1) == 0) {
      const MgLevel& F = g->lv.back();
      MgLevel C;
      C.P = (F.P + 1) / 2;
      C.nn = C.P * C.P * C.P;
      C.h_ptr.assign(C.nn + 1, 0);
      C.h_idx.reserve(15 * C.nn);
      for (int64_t I = 0; I < C.nn; ++I) {
        const int64_t X[3] = {I % C.P, (I / C.P) % C.P, I / (C.P * C.P)};
        for (const auto& o : kuhn) {
          const int64_t Y[3] = {X[0] + o[0], X[1] + o[1], X[2] + o[2]};
          if (Y[0] < 0 || Y[1] < 0 || Y[2] < 0 || Y[0] >= C.P || Y[1] >= C.P || Y[2] >= C.P) continue;
          C.h_idx.push_back((int32_t)(Y[0] + C.P * (Y[1] + C.P * Y[2])));
        }
        C.h_ptr[I + 1] = (int64_t)C.h_idx.size();
        std::sort(C.h_idx.begin() + C.h_ptr[I], C.h_idx.end());
        C.maxdeg = std::max(C.maxdeg, C.h_ptr[I + 1] - C.h_ptr[I]);
      }
      C.nblk = C.h_ptr[C.nn];
      // coarse dof free iff the fine dof at the same lattice point is free
      const std::vector<uint8_t>& ff = frees.back();
      std::vector<uint8_t> cf(3 * C.nn);
      for (int64_t I = 0; I < C.nn; ++I) {
        const int64_t x0 = I % C.P, x1 = (I / C.P) % C.P, x2 = I / (C.P * C.P);
        const int64_t f = 2 * x0 + F.P * (2 * x1 + F.P * 2 * x2);
        for (int d = 0; d < 3; ++d) cf[3 * I + d] = ff[3 * f + d];
      }
      frees.push_back(std::move(cf));
      g->lv.push_back(std::move(C));
    }
  }
  // device arrays and Dirichlet masks (the aggregation builder has set its levels up)
  if (mode == 0)
    for (size_t l = 0; l < g->lv.size(); ++l)
      if ((rc = level_device_setup(c, g->lv[l], l, mode, frees[l]))) return rc;
  g->ncoarse = 3 * g->lv.back().nn;
  if ((rc = alloc(c, g->inv, g->ncoarse * g->ncoarse * 8))) return rc;
  if ((rc = alloc(c, g->dense, g->ncoarse * g->ncoarse * 8))) return rc;
  return FEDD_OK;
}

// Coarse operators, smoother blocks and the coarsest inverse for the current A.
static int refresh_eager(Ctx* c) {
  MgHost* g = c->mg;
  if (!g) return fail(c, FEDD_ERR_STATE, "multigrid not set up");
  for (size_t l = 0; l < g->lv.size(); ++l) {
    MgLevel& L = g->lv[l];
    if (l > 0) {
      const MgLevel& F = g->lv[l - 1];
      const unsigned grid = (unsigned)std::min<int64_t>(L.nn, (int64_t)c->num_cu * 16);
      ktimer_begin(c);
      if (g->mode == 1) {
        int rc = sa_galerkin(c, g, l);
        if (rc) return rc;
      } else
        k_mg_galerkin<<<grid, 64, (size_t)9 * L.maxdeg * 8, c->stream>>>(
            F.P, L.P, F.d_nptr, F.d_nidx, F.d_vals, F.free.as<uint8_t>(), L.d_nptr, L.d_nidx, L.d_vals,
            L.free.as<uint8_t>());
      FEDD_HIP(c, hipGetLastError());
      ktimer_end(c, l == 1 ? "mg_galerkin_1" : "mg_galerkin_coarse");
    }
    k_mg_block_diag<<<nbk(L.nn, 256), 256, 0, c->stream>>>(L.nn, L.d_nptr, L.d_nidx, L.d_vals,
                                                           L.binv.as<double>());
    if (L.vals32.p)
      k_mg_to_f32<<<(unsigned)std::min<int64_t>(nbk(L.nblk * 9, 256), (int64_t)c->num_cu * 8), 256, 0, c->stream>>>(
          L.nblk * 9, L.d_vals, L.vals32.as<float>());
    FEDD_HIP(c, hipGetLastError());
  }
  const MgLevel& Cl = g->lv.back();
  const int64_t n = g->ncoarse;
  ktimer_begin(c);
  if (n <= FEDD_MG_LDS_N) {  // one workgroup, in LDS
    double* M0 = g->dense.as<double>();
    k_mg_fill_zero<<<nbk(n * n, 256), 256, 0, c->stream>>>(n * n, M0);
    k_mg_densify<<<nbk(Cl.nn, 256), 256, 0, c->stream>>>(Cl.nn, Cl.d_nptr, Cl.d_nidx, Cl.d_vals, M0);
    k_mg_sweep_inverse<<<1, 1024, 0, c->stream>>>((int)n, M0, g->inv.as<double>());
  } else {  // one launch per pivot block; ping-pong so that the last step lands in inv
    const int64_t nsteps = n / 3;
    double* bufs[2] = {g->inv.as<double>(), g->dense.as<double>()};
    double* M0 = bufs[nsteps % 2];
    k_mg_fill_zero<<<nbk(n * n, 256), 256, 0, c->stream>>>(n * n, M0);
    k_mg_densify<<<nbk(Cl.nn, 256), 256, 0, c->stream>>>(Cl.nn, Cl.d_nptr, Cl.d_nidx, Cl.d_vals, M0);
    for (int k = 0; k < (int)nsteps; ++k)
      k_mg_gj_step<<<nbk(n * n, 256), 256, 0, c->stream>>>((int)n, k, bufs[(nsteps - k) % 2],
                                                           bufs[(nsteps - k + 1) % 2]);
  }
  FEDD_HIP(c, hipGetLastError());
  ktimer_end(c, "mg_coarsest_inverse");
  return FEDD_OK;
}

// The first sweep (from x = 0) reads no matrix row: one lane per node.
template <int G, class VT>
static void launch_sweep(hipStream_t st, unsigned gs, int mode, const MgLevel& L, const VT* vals, const double* b,
                         const double* xin, double* out, double w) {
  const double* bi = L.binv.as<double>();
  if (mode == 0)
    k_mg_sweep<G, 0, VT><<<gs, 256, 0, st>>>(L.nn, L.d_nptr, L.d_nidx, vals, bi, b, xin, out, w);
  else if (mode == 1)
    k_mg_sweep<1, 1, VT><<<nbk(L.nn, 256), 256, 0, st>>>(L.nn, L.d_nptr, L.d_nidx, vals, bi, b, xin, out, w);
  else
    k_mg_sweep<G, 2, VT><<<gs, 256, 0, st>>>(L.nn, L.d_nptr, L.d_nidx, vals, bi, b, xin, out, w);
}

// first_done: the restriction into this level already performed its first sweep
// (result in the level's spare buffer).
static void vcycle(Ctx* c, size_t l, const double* b, double* x, bool first_done = false) {
  MgHost* g = c->mg;
  MgLevel& L = g->lv[l];
  if (l + 1 == g->lv.size()) {
    k_mg_dense_mv<<<nbk(g->ncoarse, 4), 256, 0, c->stream>>>(g->ncoarse, g->inv.as<double>(), b, x);
    return;
  }
  const double w = FEDD_MG_OMEGA;
  // coarse rows of the aggregation hierarchy hold 100+ blocks: a whole wave per node
  const bool wide = l > 0 && L.nblk > 48 * L.nn;
  const unsigned gs = nbk(L.nn * (l == 0 ? FEDD_MG_G0 : wide ? 64 : 16), 256);
  auto sweep = [&](int mode, const double* xin, double* out) {
    if (l == 0 && L.vals32.p)
      launch_sweep<FEDD_MG_G0>(c->stream, gs, mode, L, L.vals32.as<float>(), b, xin, out, w);
    else if (l == 0)
      launch_sweep<FEDD_MG_G0>(c->stream, gs, mode, L, (const double*)L.d_vals, b, xin, out, w);
    else if (wide && L.vals32.p)
      launch_sweep<64>(c->stream, gs, mode, L, L.vals32.as<float>(), b, xin, out, w);
    else if (wide)
      launch_sweep<64>(c->stream, gs, mode, L, (const double*)L.d_vals, b, xin, out, w);
    else
      launch_sweep<16>(c->stream, gs, mode, L, (const double*)L.d_vals, b, xin, out, w);
  };
  // The sweeps alternate between x and the level's spare buffer.  Starting the
  // pre-smoothing in the spare buffer makes the nu post-sweeps end in x for any nu.
  double* xa = x;
  double* xb = L.x2.as<double>();
  if (!first_done) sweep(1, nullptr, xb);  // from x = 0
  double* cur = xb;
  const int nu = l == 0 ? FEDD_MG_NU : g->mode == 1 ? FEDD_AMG_NU_COARSE : FEDD_MG_NU_COARSE;
  for (int s = 1; s < nu; ++s) {
    double* nxt = cur == xa ? xb : xa;
    sweep(0, cur, nxt);
    cur = nxt;
  }
  // coarse correction
  sweep(2, cur, L.r.as<double>());
  MgLevel& C = g->lv[l + 1];
  if (g->mode == 1)
    k_sa_restrict<<<nbk(C.nn * 64, 256), 256, 0, c->stream>>>(C.nn, C.tptr.as<int64_t>(), C.tidx.as<int32_t>(),
                                                              C.ptval32.as<float>(), L.r.as<double>(),
                                                              C.free.as<uint8_t>(), C.b.as<double>());
  const bool fuse = g->mode != 1 && l + 2 < g->lv.size();  // the coarsest level is a dense solve
  if (g->mode != 1)
    k_mg_restrict<<<nbk(C.nn * 16, 256), 256, 0, c->stream>>>(
        L.P, C.P, L.r.as<double>(), L.free.as<uint8_t>(), C.free.as<uint8_t>(), C.b.as<double>(),
        fuse ? C.binv.as<double>() : nullptr, fuse ? C.x2.as<double>() : nullptr, w);
  vcycle(c, l + 1, C.b.as<double>(), C.x.as<double>(), fuse);
  if (g->mode == 1)
    k_sa_prolong<16><<<nbk(L.nn * 16, 256), 256, 0, c->stream>>>(L.nn, C.pptr.as<int64_t>(), C.pidx.as<int32_t>(),
                                                                 C.pval32.as<float>(), C.x.as<double>(), cur);
  else
    k_mg_prolong<<<nbk(L.nn, 256), 256, 0, c->stream>>>(L.P, C.P, C.x.as<double>(), L.free.as<uint8_t>(),
                                                        C.free.as<uint8_t>(), cur);
  for (int s = 0; s < nu; ++s) {
    double* nxt = cur == xa ? xb : xa;
    sweep(0, cur, nxt);
    cur = nxt;
  }
}

// Records what `body` enqueues on the context stream into *exec (the buffers it
// touches keep their addresses while the hierarchy lives).
template <class F>
static int capture(Ctx* c, hipGraphExec_t* exec, F body) {
  FEDD_HIP(c, hipStreamBeginCapture(c->stream, hipStreamCaptureModeRelaxed));
  c->capturing = true;
  const int rc = body();
  c->capturing = false;
  hipGraph_t gr = nullptr;
  const hipError_t e = hipStreamEndCapture(c->stream, &gr);
  if (rc) {
    if (gr) (void)hipGraphDestroy(gr);
    return rc;
  }
  if (e != hipSuccess || !gr) return fail(c, FEDD_ERR_HIP, std::string("hipStreamEndCapture: ") + hipGetErrorString(e));
  const hipError_t ei = hipGraphInstantiate(exec, gr, nullptr, nullptr, 0);
  (void)hipGraphDestroy(gr);
  if (ei != hipSuccess) return fail(c, FEDD_ERR_HIP, std::string("hipGraphInstantiate: ") + hipGetErrorString(ei));
  return FEDD_OK;
}

// Eager launches when kernel timing is on (per-kernel events) or a caller's own
// capture is open; otherwise one replayed graph per call.
int mg_refresh(Ctx* c) {
  MgHost* g = c->mg;
  if (!g) return fail(c, FEDD_ERR_STATE, "multigrid not set up");
  if (c->capturing || c->timing) return refresh_eager(c);
  if (!g->rexec) {
    int rc = capture(c, &g->rexec, [&]() -> int { return refresh_eager(c); });
    if (rc) return rc;
  }
  FEDD_HIP(c, hipGraphLaunch(g->rexec, c->stream));
  return FEDD_OK;
}

int mg_refresh_fine(Ctx* c) {
  MgHost* g = c->mg;
  if (!g) return fail(c, FEDD_ERR_STATE, "multigrid not set up");
  MgLevel& L = g->lv[0];
  k_mg_block_diag<<<nbk(L.nn, 256), 256, 0, c->stream>>>(L.nn, L.d_nptr, L.d_nidx, L.d_vals, L.binv.as<double>());
  if (L.vals32.p)
    k_mg_to_f32<<<(unsigned)std::min<int64_t>(nbk(L.nblk * 9, 256), (int64_t)c->num_cu * 8), 256, 0, c->stream>>>(
        L.nblk * 9, L.d_vals, L.vals32.as<float>());
  FEDD_HIP(c, hipGetLastError());
  return FEDD_OK;
}

const float* mg_fine_vals32(Ctx* c) {
  if (!c->mg || c->mg->lv.empty()) return nullptr;
  return c->mg->lv[0].vals32.as<float>();
}

// x = V-cycle(b) on the finest level (3 dofs per node).
int mg_vcycle(Ctx* c, const double* b, double* x) {
  MgHost* g = c->mg;
  if (!g) return fail(c, FEDD_ERR_STATE, "multigrid not set up");
  if (c->capturing || c->timing) {
    ktimer_begin(c);
    vcycle(c, 0, b, x);
    ktimer_end(c, "mg_vcycle");
    FEDD_HIP(c, hipGetLastError());
    return FEDD_OK;
  }
  if (!g->vexec || g->vb != b || g->vx != x) {
    if (g->vexec) {
      sync_streams(c);  // an earlier replay of this executable may still be running
      (void)hipGraphExecDestroy(g->vexec);
    }
    g->vexec = nullptr;
    int rc = capture(c, &g->vexec, [&]() -> int {
      vcycle(c, 0, b, x);
      FEDD_HIP(c, hipGetLastError());
      return FEDD_OK;
    });
    if (rc) return rc;
    g->vb = b;
    g->vx = x;
  }
  FEDD_HIP(c, hipGraphLaunch(g->vexec, c->stream));
  return FEDD_OK;
}

}  // namespace fedd
