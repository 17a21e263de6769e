// Element-assembly kernels for gfx950: CRS graph expansion, the P2-P1
// Navier-Stokes Jacobian / fixed-point / residual row gathers, and the P1/P2
// Laplace stiffness and load vector.
//
// Row gather ("owner computes"): the rows of a contiguous chunk of nodes live in
// LDS (FP64) while the chunk's (node, element) pairs add their element-matrix
// rows into them; the chunk's values are then written to HBM exactly once, in
// CRS order.  A pair relabels its element so that the row node sits at local
// position 0 (vertex) or 4 (edge); the element integrals are exact reference
// tensors (fedd_tensors.h), constant-folded into the two row kernels.
//
// Reference semantics (FEDDLib):
//   element matrix  AssembleFENavierStokes_def.hpp:67-146 (Jacobian, FixedPoint)
//   A  = nu*rho*|det| int grad phi_i . grad phi_j                     (:150-196, :81-82)
//   N  = rho*|det| int (u . grad phi_j) phi_i, diagonal 3x3 blocks     (:228-298)
//   W  = rho*|det| int phi_i phi_j du_d1/dx_d2, full 3x3 blocks        (:302-369)
//   B  = -|det| int psi_m d phi_j/dx_d, placed as B and B^T            (:374-461, :88)
//   residual = (constantMatrix + rho N) [u;p] (.) coeff, no W          (:200-224)
//   scatter  = addFeBlock / addFeBlockMatrix                           (FE_def.hpp:982-1062, 1419-1442)
#include "fedd_internal.h"
#include "fedd_perm.h"
#include "fedd_tensors.h"

#include <algorithm>
#include <type_traits>

#ifndef FEDD_ROW_WAVES
#define FEDD_ROW_WAVES 3
#endif
#ifndef FEDD_ROW_BLOCKS_PER_CU
#define FEDD_ROW_BLOCKS_PER_CU 0  // 0: one workgroup per chunk
#endif

namespace fedd {

__constant__ int8_t c_perm2[10][10] = FEDD_PERM_P2_INIT;
__constant__ int8_t c_perm1[4][4] = FEDD_PERM_P1_INIT;
// PERM rows packed as ten 4-bit fields: one load per pair instead of ten.
constexpr uint64_t perm_code(int li) {
  uint64_t v = 0;
  for (int c = 0; c < 10; ++c) v |= (uint64_t)kPermP2[li][c] << (4 * c);
  return v;
}
__constant__ uint64_t c_perm_code[10] = {perm_code(0), perm_code(1), perm_code(2), perm_code(3), perm_code(4),
                                         perm_code(5), perm_code(6), perm_code(7), perm_code(8), perm_code(9)};

namespace ft = fedd_tensors;

typedef double dbl2 __attribute__((ext_vector_type(2)));

// ---------------------------------------------------------------------------
// CRS expansion from node adjacency
// ---------------------------------------------------------------------------

// A(0,0) of a 3-dof field: rows 3i+d1 hold the 3x3 node blocks of all neighbours
// j (sorted), columns 3j+d2 (addFeBlock inserts full node blocks, FE_def.hpp:1430-1441).
__global__ void k_expand_vec3(int64_t N, const int64_t* __restrict__ nptr,
                              const int32_t* __restrict__ adj, int64_t* __restrict__ rowptr,
                              int32_t* __restrict__ colind) {
  int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (i >= N) return;
  int64_t b = nptr[i], deg = nptr[i + 1] - b;
  for (int d1 = 0; d1 < 3; ++d1) {
    int64_t r = 9 * b + 3 * d1 * deg;
    rowptr[3 * i + d1] = r;
    for (int64_t s = 0; s < deg; ++s) {
      int32_t j = adj[b + s];
      colind[r + 3 * s + 0] = 3 * j + 0;
      colind[r + 3 * s + 1] = 3 * j + 1;
      colind[r + 3 * s + 2] = 3 * j + 2;
    }
  }
  if (i == N - 1) rowptr[3 * N] = 9 * nptr[N];
}

// Scalar field: the node graph itself.
__global__ void k_expand_scalar(int64_t N, const int64_t* __restrict__ nptr,
                                const int32_t* __restrict__ adj, int64_t* __restrict__ rowptr,
                                int32_t* __restrict__ colind) {
  int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (i >= N) return;
  int64_t b = nptr[i], deg = nptr[i + 1] - b;
  rowptr[i] = b;
  for (int64_t s = 0; s < deg; ++s) colind[b + s] = adj[b + s];
  if (i == N - 1) rowptr[N] = nptr[N];
}

// B^T (0,1): rows 3i+d, columns the pressure neighbours of i (FE_def.hpp:1035-1046).
__global__ void k_expand_bt(int64_t N, const int64_t* __restrict__ pptr,
                            const int32_t* __restrict__ padj, int64_t* __restrict__ rowptr,
                            int32_t* __restrict__ colind) {
  int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (i >= N) return;
  int64_t b = pptr[i], deg = pptr[i + 1] - b;
  for (int d = 0; d < 3; ++d) {
    int64_t r = 3 * b + d * deg;
    rowptr[3 * i + d] = r;
    for (int64_t s = 0; s < deg; ++s) colind[r + s] = padj[b + s];
  }
  if (i == N - 1) rowptr[3 * N] = 3 * pptr[N];
}

// B (1,0): rows p, columns 3j+d for the velocity neighbours j (FE_def.hpp:1048-1059).
__global__ void k_expand_b(int64_t NP, const int64_t* __restrict__ bptr,
                           const int32_t* __restrict__ badj, int64_t* __restrict__ rowptr,
                           int32_t* __restrict__ colind) {
  int64_t p = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (p >= NP) return;
  int64_t b = bptr[p], deg = bptr[p + 1] - b;
  rowptr[p] = 3 * b;
  for (int64_t s = 0; s < deg; ++s) {
    int32_t j = badj[b + s];
    colind[3 * b + 3 * s + 0] = 3 * j + 0;
    colind[3 * b + 3 * s + 1] = 3 * j + 1;
    colind[3 * b + 3 * s + 2] = 3 * j + 2;
  }
  if (p == NP - 1) rowptr[NP] = 3 * bptr[NP];
}

__global__ void k_fill_i64(int64_t n, int64_t* __restrict__ a, int64_t v) {
  int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (i < n) a[i] = v;
}

static inline unsigned nblk(int64_t n, int t) { return (unsigned)((n + t - 1) / t); }

int expand_graph(Ctx* c, int kind, const int32_t* aidx, const int32_t* btidx, const int32_t* bidx) {
  const int64_t N = c->mesh[0].npts;
  const int T = 256;
  if (kind == FEDD_GRAPH_SCALAR) {
    k_expand_scalar<<<nblk(N, T), T, 0, c->stream>>>(N, c->d_nptr.as<int64_t>(), aidx,
                                                     c->blk[0].d_rowptr.as<int64_t>(),
                                                     c->blk[0].d_colind.as<int32_t>());
  } else {
    k_expand_vec3<<<nblk(N, T), T, 0, c->stream>>>(N, c->d_nptr.as<int64_t>(), aidx,
                                                   c->blk[0].d_rowptr.as<int64_t>(),
                                                   c->blk[0].d_colind.as<int32_t>());
  }
  FEDD_HIP(c, hipGetLastError());
  if (kind == FEDD_GRAPH_NS) {
    const int64_t NP = c->mesh[1].npts;
    k_expand_bt<<<nblk(N, T), T, 0, c->stream>>>(N, c->d_pptr.as<int64_t>(), btidx,
                                                 c->blk[1].d_rowptr.as<int64_t>(),
                                                 c->blk[1].d_colind.as<int32_t>());
    k_expand_b<<<nblk(NP, T), T, 0, c->stream>>>(NP, c->d_bptr.as<int64_t>(), bidx,
                                                 c->blk[2].d_rowptr.as<int64_t>(),
                                                 c->blk[2].d_colind.as<int32_t>());
    k_fill_i64<<<nblk(NP + 1, T), T, 0, c->stream>>>(NP + 1, c->blk[3].d_rowptr.as<int64_t>(), 0);
    FEDD_HIP(c, hipGetLastError());
  }
  FEDD_HIP(c, hipStreamSynchronize(c->stream));
  return FEDD_OK;
}

// ---------------------------------------------------------------------------
// Element geometry (Helper::buildTransformation + SmallMatrix::computeInverse,
// Helper.cpp:196-222, SmallMatrix.hpp:317-368), always FP64.
// ---------------------------------------------------------------------------
struct Geo {
  double Bi[3][3];
  double adet;
};

__device__ __forceinline__ void geometry(const double* __restrict__ xyz, int64_t npts,
                                         const int32_t nodes[4], Geo& g) {
  double x[4][3];
#pragma unroll
  for (int k = 0; k < 4; ++k)
#pragma unroll
    for (int d = 0; d < 3; ++d) x[k][d] = xyz[d * npts + nodes[k]];
  double a[3][3];
#pragma unroll
  for (int i = 0; i < 3; ++i)
#pragma unroll
    for (int j = 0; j < 3; ++j) a[i][j] = x[j + 1][i] - x[0][i];
  double det = a[0][0] * a[1][1] * a[2][2] + a[0][1] * a[1][2] * a[2][0] + a[0][2] * a[1][0] * a[2][1] -
               a[2][0] * a[1][1] * a[0][2] - a[2][1] * a[1][2] * a[0][0] - a[2][2] * a[1][0] * a[0][1];
  double id = 1.0 / det;
  g.Bi[0][0] = (a[1][1] * a[2][2] - a[1][2] * a[2][1]) * id;
  g.Bi[0][1] = (a[0][2] * a[2][1] - a[0][1] * a[2][2]) * id;
  g.Bi[0][2] = (a[0][1] * a[1][2] - a[0][2] * a[1][1]) * id;
  g.Bi[1][0] = (a[1][2] * a[2][0] - a[1][0] * a[2][2]) * id;
  g.Bi[1][1] = (a[0][0] * a[2][2] - a[0][2] * a[2][0]) * id;
  g.Bi[1][2] = (a[0][2] * a[1][0] - a[0][0] * a[1][2]) * id;
  g.Bi[2][0] = (a[1][0] * a[2][1] - a[1][1] * a[2][0]) * id;
  g.Bi[2][1] = (a[0][1] * a[2][0] - a[0][0] * a[2][1]) * id;
  g.Bi[2][2] = (a[0][0] * a[1][1] - a[0][1] * a[1][0]) * id;
  g.adet = fabs(det);
}

// Symmetric metric G_ab = sum_d Bi[a][d] Bi[b][d] in the pair order of KS_*.
template <typename T>
__device__ __forceinline__ void metric(const T Bi[3][3], T Gs[6]) {
  const int pa[6] = {0, 1, 2, 0, 0, 1}, pb[6] = {0, 1, 2, 1, 2, 2};
#pragma unroll
  for (int s = 0; s < 6; ++s)
    Gs[s] = Bi[pa[s]][0] * Bi[pb[s]][0] + Bi[pa[s]][1] * Bi[pb[s]][1] + Bi[pa[s]][2] * Bi[pb[s]][2];
}

template <int CLS>
struct NsTensors;
template <>
struct NsTensors<0> {
  static constexpr const double (&KS)[6][10] = ft::KS_V;
  static constexpr const double (&TT)[3][10][10] = ft::T_V;
  static constexpr const double (&L)[4][10] = ft::L_V;
  static constexpr const double (&RV)[3][4] = ft::RV_V;
  static constexpr const double (&RG)[3][3][10] = ft::RG_V;
  static constexpr int row = 0;
};
template <>
struct NsTensors<1> {
  static constexpr const double (&KS)[6][10] = ft::KS_E;
  static constexpr const double (&TT)[3][10][10] = ft::T_E;
  static constexpr const double (&L)[4][10] = ft::L_E;
  static constexpr const double (&RV)[3][4] = ft::RV_E;
  static constexpr const double (&RG)[3][3][10] = ft::RG_E;
  static constexpr int row = 4;
};

// Convection row N_c = int phi_r (uhat . grad_hat phi_c) (without |det|), uhat_ax =
// sum_d Bi[ax][d] u_d.  grad_hat phi_c is linear, so it equals its vertex interpolant
// sum_m DV[ax][c][m] lambda_m, and
//   N_c = sum_(ax,m) DV[ax][c][m] Y[ax][m],  Y[ax][m] = sum_d Bi[ax][d] Z[d][m],
//   Z[d][m] = sum_k u_kd int lambda_m phi_r phi_k = sum_k u_kd L[m][k]
// (L is the same table as the Newton term's): 171 FMA per row instead of the 276 of
// the direct contraction with T[ax][k][c] = int phi_r phi_k d_ax phi_c.
#ifndef FEDD_NS_N_VIA_L
#define FEDD_NS_N_VIA_L 1
#endif
template <typename T, int CLS>
__device__ __forceinline__ void convection_row(const T Bi[3][3], const T U[10][3], T N[10]) {
  using NT = NsTensors<CLS>;
#if FEDD_NS_N_VIA_L
  T Z[3][4];
#pragma unroll
  for (int d = 0; d < 3; ++d)
#pragma unroll
    for (int mm = 0; mm < 4; ++mm) {
      T z = 0;
#pragma unroll
      for (int k = 0; k < 10; ++k)
        if (NT::L[mm][k] != 0.0) z += U[k][d] * (T)NT::L[mm][k];
      Z[d][mm] = z;
    }
  T Y[3][4];
#pragma unroll
  for (int ax = 0; ax < 3; ++ax)
#pragma unroll
    for (int mm = 0; mm < 4; ++mm) Y[ax][mm] = Bi[ax][0] * Z[0][mm] + Bi[ax][1] * Z[1][mm] + Bi[ax][2] * Z[2][mm];
#pragma unroll
  for (int c = 0; c < 10; ++c) {
    T n = 0;
#pragma unroll
    for (int ax = 0; ax < 3; ++ax)
#pragma unroll
      for (int mm = 0; mm < 4; ++mm)
        if (ft::DV[ax][c][mm] != 0.0) n += Y[ax][mm] * (T)ft::DV[ax][c][mm];
    N[c] = n;
  }
#else
#pragma unroll
  for (int c = 0; c < 10; ++c) N[c] = 0;
#pragma unroll
  for (int k = 0; k < 10; ++k) {
    T ut[3];  // ut_ax = sum_d Bi[ax][d] u_kd
#pragma unroll
    for (int ax = 0; ax < 3; ++ax) ut[ax] = Bi[ax][0] * U[k][0] + Bi[ax][1] * U[k][1] + Bi[ax][2] * U[k][2];
#pragma unroll
    for (int c = 0; c < 10; ++c)
#pragma unroll
      for (int ax = 0; ax < 3; ++ax)
        if (NT::TT[ax][k][c] != 0.0) N[c] += ut[ax] * (T)NT::TT[ax][k][c];
  }
#endif
}

// ---------------------------------------------------------------------------
// Navier-Stokes velocity rows
// ---------------------------------------------------------------------------
struct NsArgs {
  // meshes
  const double* xyz;       // P2 coordinates SoA
  int64_t npts;
  const int32_t* conn;     // P2 connectivity SoA
  const int32_t* pconn;    // P1 connectivity SoA
  int64_t nelem;
  const double* u;         // velocity iterate, 3 per node
  const double* p;         // pressure iterate
  // node-level row pointers
  const int64_t* nptr;     // A node graph
  const int64_t* pptr;     // B^T node graph (velocity -> pressure)
  const int64_t* bptr;     // B node graph (pressure -> velocity)
  // plan
  const int32_t* chunk_node;
  const int32_t* chunk_pair;
  const int32_t* chunk_split;
  const int32_t* pair_elem;
  const uint4* pair_meta;
  // outputs
  double* valsA;
  double* valsBT;
  double* valsB;
  double* res_u;
  double* res_p;
  // parameters
  int64_t nchunks;
  double nurho, rho;
  double coeff[4];
  int newton;
  int with_bt;
};

// RES = 1 (the fused Jacobian + residual modes): the pair also adds its row of the
// residual r_(r,d) = c00 sum_c S_rc u_c,d + c01 sum_m BT_(r,d),m p_m (S = (nu rho A + rho N)
// |det|, no W; AssembleFENavierStokes_def.hpp:200-224) into rowRes[d].  The residual stays
// FP64: with T = float the S row is then computed in FP64 too and shared with the matrix.
template <typename T, int CLS, int RES>
struct VelRow {
  // Computes the canonical row of one pair and adds it into LDS.
  //   e      element, m = pair meta (loc, node-in-chunk, 10 column slots, 4 pressure slots)
  //   Binv'  inverse of the relabelled element's affine map, from its vertex coordinates
  //   u      the ten canonical nodes' velocities
  static __device__ __forceinline__ void run(const NsArgs& a, int32_t e, const uint8_t* m,
                                             double* __restrict__ rowA, int deg,
                                             double* __restrict__ rowBT, int pdeg, double* __restrict__ rowRes) {
    using NT = NsTensors<CLS>;
    using TS = typename std::conditional<RES != 0, double, T>::type;  // type of the S row
    const uint64_t code = c_perm_code[m[0]];
    int32_t nodes[10];
#pragma unroll
    for (int c = 0; c < 10; ++c) nodes[c] = a.conn[(int64_t)((code >> (4 * c)) & 15) * a.nelem + e];
    Geo g;
    geometry(a.xyz, a.npts, nodes, g);  // of the relabelled element
    T Bi[3][3];
    TS BiS[3][3];
#pragma unroll
    for (int ax = 0; ax < 3; ++ax)
#pragma unroll
      for (int d = 0; d < 3; ++d) {
        Bi[ax][d] = (T)g.Bi[ax][d];
        BiS[ax][d] = (TS)g.Bi[ax][d];
      }
    const double adet = g.adet;
    TS U[10][3];
#pragma unroll
    for (int k = 0; k < 10; ++k) {
      const double* uk = a.u + 3 * (int64_t)nodes[k];
      U[k][0] = (TS)uk[0];
      U[k][1] = (TS)uk[1];
      U[k][2] = (TS)uk[2];
    }
    // S_c = (nu rho A_rc + rho N_rc) |det|
    TS S[10];
    {
      TS Gs[6];
      metric<TS>(BiS, Gs);
#pragma unroll
      for (int c = 0; c < 10; ++c) {
        TS acc = 0;
#pragma unroll
        for (int s = 0; s < 6; ++s)
          if (NT::KS[s][c] != 0.0) acc += Gs[s] * (TS)NT::KS[s][c];
        S[c] = (TS)a.nurho * acc;
      }
    }
    {
      TS Nacc[10];
      convection_row<TS, CLS>(BiS, U, Nacc);
#pragma unroll
      for (int c = 0; c < 10; ++c) S[c] = (TS)((double)(S[c] + (TS)a.rho * Nacc[c]) * adet);
    }
    if (rowBT || RES) {
      // B^T entries of row (r, d1): -|det| sum_ax Bi[ax][d1] RV[ax][m]
      TS pm[4] = {0, 0, 0, 0};
      if (RES) {
#pragma unroll
        for (int mm = 0; mm < 4; ++mm) pm[mm] = (TS)a.p[a.pconn[(int64_t)((code >> (4 * mm)) & 15) * a.nelem + e]];
      }
      TS rp[3] = {0, 0, 0};
#pragma unroll
      for (int mm = 0; mm < 4; ++mm)
#pragma unroll
        for (int d1 = 0; d1 < 3; ++d1) {
          TS v = 0;
#pragma unroll
          for (int ax = 0; ax < 3; ++ax)
            if (NT::RV[ax][mm] != 0.0) v += BiS[ax][d1] * (TS)NT::RV[ax][mm];
          if (rowBT) atomicAdd(rowBT + d1 * pdeg + m[12 + mm], -(double)(T)v * adet);
          if (RES) rp[d1] -= v * pm[mm];
        }
      if (RES) {
        // residual row, FP64: c00 sum_c S_c u_c + c01 B^T p
#pragma unroll
        for (int d = 0; d < 3; ++d) {
          double ru = 0.0;
#pragma unroll
          for (int c = 0; c < 10; ++c) ru += (double)S[c] * (double)U[c][d];
          atomicAdd(rowRes + d, a.coeff[0] * ru + a.coeff[1] * ((double)rp[d] * adet));
        }
      }
    }
    if (a.newton) {
      // grad u is P1: du_d1/dxhat_ax = sum_m lambda_m gh[m][ax], gh[m][ax] = sum_k u_k,d1 DV[ax][k][m]
      const T rdet = (T)(a.rho * adet);
#pragma unroll
      for (int d1 = 0; d1 < 3; ++d1) {
        T gh[4][3];
#pragma unroll
        for (int mm = 0; mm < 4; ++mm)
#pragma unroll
          for (int ax = 0; ax < 3; ++ax) gh[mm][ax] = 0;
#pragma unroll
        for (int k = 0; k < 10; ++k)
#pragma unroll
          for (int mm = 0; mm < 4; ++mm)
#pragma unroll
            for (int ax = 0; ax < 3; ++ax)
              if (ft::DV[ax][k][mm] != 0.0) gh[mm][ax] += (T)U[k][d1] * (T)ft::DV[ax][k][mm];
#pragma unroll
        for (int d2 = 0; d2 < 3; ++d2) {
          T gw[4];  // rho |det| du_d1/dx_d2 at vertex m
#pragma unroll
          for (int mm = 0; mm < 4; ++mm)
            gw[mm] = rdet * (Bi[0][d2] * gh[mm][0] + Bi[1][d2] * gh[mm][1] + Bi[2][d2] * gh[mm][2]);
#pragma unroll
          for (int c = 0; c < 10; ++c) {
            T w = 0;
#pragma unroll
            for (int mm = 0; mm < 4; ++mm)
              if (NT::L[mm][c] != 0.0) w += gw[mm] * (T)NT::L[mm][c];
            double wd = (double)w;
            if (d1 == d2) wd = (double)((T)wd + (T)S[c]);
            atomicAdd(rowA + d1 * 3 * deg + 3 * m[2 + c] + d2, wd);
          }
        }
      }
    } else {
#pragma unroll
      for (int c = 0; c < 10; ++c) {
        const double v = (double)(T)S[c];
        double* r = rowA + 3 * m[2 + c];
        atomicAdd(r + 0 * 3 * deg + 0, v);
        atomicAdd(r + 1 * 3 * deg + 1, v);
        atomicAdd(r + 2 * 3 * deg + 2, v);
      }
    }
  }
};

// One workgroup per chunk (the hardware keeps several chunks per CU in
// flight, each in a different phase: zeroing, gathering, writing).  Inside a
// chunk, vertex-class pairs occupy whole waves before the edge-class pairs start
// (nv is padded to a multiple of 64), so each wave runs one specialised row kernel.
// Block barrier that orders LDS only.  __syncthreads() also waits for every
// outstanding global store of the wave; here the nontemporal stores of one chunk
// keep draining to HBM while the block gathers the next chunk.
__device__ __forceinline__ void lds_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup", "local");
  __builtin_amdgcn_s_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup", "local");
}

// Chunks [c0, c1) of the block: contiguous, and contiguous per XCD (xcd_chunk).
__device__ __forceinline__ void block_chunks(int64_t nchunks, int64_t& c0, int64_t& c1) {
  const int64_t G = gridDim.x;
  const int64_t L = xcd_chunk(blockIdx.x, G);
  const int64_t q = nchunks / G, r = nchunks % G;
  c0 = L * q + (L < r ? L : r);
  c1 = c0 + q + (L < r ? 1 : 0);
}

// Each workgroup walks its chunks in order.  Per chunk: zero the LDS rows, add
// every pair's canonical row with LDS atomics, then stream the finished rows out
// with nontemporal stores in CRS order; the stores drain while the next chunk is
// gathered.  Inside a chunk, vertex-class pairs occupy whole waves before the
// edge-class pairs start (nv is padded to a multiple of 64), so each wave runs one
// specialised row kernel.
#ifndef FEDD_ROW_WAVES_RES
#define FEDD_ROW_WAVES_RES 3
#endif
// The FP32-local Jacobian kernel fits 128 VGPRs, so its workgroups take four waves and
// four of them still share a CU (16 waves): 40^3 x 6 FP32-local step 0.437 -> 0.426 ms.
// The FP64 and fused kernels need 163-168 VGPRs: three waves per SIMD, 192 threads.
#ifndef FEDD_ROW_THREADS_F32
#define FEDD_ROW_THREADS_F32 256
#endif
#ifndef FEDD_ROW_WAVES_F32
#define FEDD_ROW_WAVES_F32 4
#endif
// Workgroup size and waves per SIMD of each row kernel.
template <typename T, int RES>
struct RowShape {
  static constexpr int threads = kNsRowThreads;
  static constexpr int waves = RES ? FEDD_ROW_WAVES_RES : FEDD_ROW_WAVES;
};
template <>
struct RowShape<float, 0> {
  static constexpr int threads = FEDD_ROW_THREADS_F32;
  static constexpr int waves = FEDD_ROW_WAVES_F32;
};
template <typename T, int RES>
__global__ __attribute__((amdgpu_flat_work_group_size(1, RowShape<T, RES>::threads),
                          amdgpu_waves_per_eu(RowShape<T, RES>::waves))) void k_ns_vel_rows(NsArgs a) {
  extern __shared__ double lds[];
  __shared__ int4 tab[kNsChunkNodes];  // per chunk node: A-row offset, deg, B^T-row offset, pdeg
  int64_t c0, c1;
  block_chunks(a.nchunks, c0, c1);
  for (int64_t ch = c0; ch < c1; ++ch) {
    const int n0 = a.chunk_node[ch], n1 = a.chunk_node[ch + 1];
    const int q0 = a.chunk_pair[ch], q1 = a.chunk_pair[ch + 1], qs = a.chunk_split[ch];
    const int64_t na0 = a.nptr[n0], na1 = a.nptr[n1];
    const int aLen = (int)(9 * (na1 - na0));
    int64_t nb0 = 0;
    int btLen = 0;
    if (a.with_bt) {
      nb0 = a.pptr[n0];
      btLen = (int)(3 * (a.pptr[n1] - nb0));
    }
    for (int t = threadIdx.x; t < n1 - n0; t += blockDim.x) {
      const int64_t ai = a.nptr[n0 + t], ai1 = a.nptr[n0 + t + 1];
      int4 v;
      v.x = (int)(9 * (ai - na0));
      v.y = (int)(ai1 - ai);
      if (a.with_bt) {
        const int64_t bi = a.pptr[n0 + t], bi1 = a.pptr[n0 + t + 1];
        v.z = aLen + (int)(3 * (bi - nb0));
        v.w = (int)(bi1 - bi);
      } else {
        v.z = 0;
        v.w = 0;
      }
      tab[t] = v;
    }
    const int resLen = RES ? 3 * (n1 - n0) : 0;  // residual rows after A and B^T
    for (int k = threadIdx.x; k < aLen + btLen + resLen; k += blockDim.x) lds[k] = 0.0;
    lds_sync();
    const int nv = qs - q0, ne = q1 - qs;
    const int nvpad = (nv + 63) & ~63;
    for (int t = threadIdx.x; t < nvpad + ne; t += blockDim.x) {
      const bool vert = t < nvpad;  // uniform across the wave
      if (vert && t >= nv) continue;
      const int q = vert ? q0 + t : qs + (t - nvpad);
      const int32_t e = a.pair_elem[q];
      if (e < 0) continue;  // padding pair
      const uint4 mv = a.pair_meta[q];
      const uint8_t* m = reinterpret_cast<const uint8_t*>(&mv);
      const int4 nd = tab[m[1]];
      double* rowA = lds + nd.x;
      double* rowBT = a.with_bt ? lds + nd.z : nullptr;
      double* rowRes = lds + aLen + btLen + 3 * m[1];
      if (vert)
        VelRow<T, 0, RES>::run(a, e, m, rowA, nd.y, rowBT, nd.w, rowRes);
      else
        VelRow<T, 1, RES>::run(a, e, m, rowA, nd.y, rowBT, nd.w, rowRes);
    }
    lds_sync();
    double* outA = a.valsA + 9 * na0;
    for (int k = threadIdx.x; k < aLen; k += blockDim.x) __builtin_nontemporal_store(lds[k], outA + k);
    if (a.with_bt) {
      double* outB = a.valsBT + 3 * nb0;
      for (int k = threadIdx.x; k < btLen; k += blockDim.x)
        __builtin_nontemporal_store(lds[aLen + k], outB + k);
    }
    if (RES) {
      double* outR = a.res_u + 3 * (int64_t)n0;
      for (int k = threadIdx.x; k < resLen; k += blockDim.x) outR[k] = lds[aLen + btLen + k];
    }
    lds_sync();  // the rows are read out before the next chunk zeroes them
  }
}

// Pressure rows of B (1,0): row p of element matrix rows 30+m (already scaled by -1).
// RES = 1: also the pressure residual r_m = c10 sum_(j,d) B_m,(j,d) u_j,d (FP64) into
// res_p; WRITEB = 0: the residual only (reAssemble, where B is not rewritten).
template <typename T, int RES, int WRITEB>
__global__ __launch_bounds__(256) void k_ns_pres_rows(NsArgs a) {
  extern __shared__ double lds[];
  const int ch = (int)xcd_chunk(blockIdx.x, gridDim.x);
  const int n0 = a.chunk_node[ch], n1 = a.chunk_node[ch + 1];
  const int q0 = a.chunk_pair[ch], q1 = a.chunk_pair[ch + 1];
  const int64_t nb0 = a.bptr[n0];
  const int len = WRITEB ? (int)(3 * (a.bptr[n1] - nb0)) : 0;
  const int resLen = RES ? n1 - n0 : 0;
  for (int k = threadIdx.x; k < len + resLen; k += blockDim.x) lds[k] = 0.0;
  __syncthreads();
  for (int q = q0 + threadIdx.x; q < q1; q += blockDim.x) {
    const int32_t e = a.pair_elem[q];
    if (e < 0) continue;  // padding pair
    const uint4 mv = a.pair_meta[q];
    const uint8_t* m = reinterpret_cast<const uint8_t*>(&mv);
    const int li = m[0];
    const int pn = n0 + m[1];
    int32_t nodes[4];
#pragma unroll
    for (int c = 0; c < 4; ++c) nodes[c] = a.conn[(int64_t)c_perm2[li][c] * a.nelem + e];
    Geo g;
    geometry(a.xyz, a.npts, nodes, g);
    double* row = lds + 3 * (a.bptr[pn] - nb0);
    double acc = 0.0;
#pragma unroll
    for (int c = 0; c < 10; ++c) {
      const double* uc = RES ? a.u + 3 * (int64_t)a.conn[(int64_t)c_perm2[li][c] * a.nelem + e] : nullptr;
#pragma unroll
      for (int d = 0; d < 3; ++d) {
        if (WRITEB) {
          T v = 0;
#pragma unroll
          for (int ax = 0; ax < 3; ++ax)
            if (ft::RP[ax][c] != 0.0) v += (T)g.Bi[ax][d] * (T)ft::RP[ax][c];
          atomicAdd(row + 3 * m[2 + c] + d, -(double)v * g.adet);
        }
        if (RES) {
          double v = 0.0;
#pragma unroll
          for (int ax = 0; ax < 3; ++ax)
            if (ft::RP[ax][c] != 0.0) v += g.Bi[ax][d] * ft::RP[ax][c];
          acc += v * uc[d];
        }
      }
    }
    if (RES) atomicAdd(lds + len + m[1], -a.coeff[2] * acc * g.adet);
  }
  __syncthreads();
  if (RES)
    for (int k = threadIdx.x; k < resLen; k += blockDim.x) a.res_p[n0 + k] = lds[len + k];
  if (!WRITEB) return;
  double* out = a.valsB + 3 * nb0;
  for (int k = threadIdx.x; k < len; k += blockDim.x) __builtin_nontemporal_store(lds[k], out + k);
}

// Coloured element scatter (FEDD_METHOD_COLOUR_SCATTER): one lane per (element, local
// node) pair of one colour.  The lane forms the pair's canonical row exactly as the row
// gather does and adds it straight into the global values with no-return FP64 atomics.
// Elements of one colour share no node, so no two lanes of a launch add into the same
// value, and a value receives its contributions colour by colour in launch order: the
// sums repeat bit for bit.  Velocity pairs [q0, qs) are vertex class, [qs, q1) edge
// class; the vertex class is padded to whole waves so that each wave runs one row kernel.
template <typename T>
__global__ __launch_bounds__(256) void k_ns_vel_scatter(NsArgs a, const int32_t* __restrict__ elem,
                                                        const int32_t* __restrict__ node,
                                                        const uint4* __restrict__ meta, int64_t q0, int64_t qs,
                                                        int64_t q1) {
  const int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  const int64_t nv = qs - q0, nvpad = (nv + 63) & ~(int64_t)63;
  const bool vert = t < nvpad;  // uniform across the wave
  if (vert && t >= nv) return;
  const int64_t q = vert ? q0 + t : qs + (t - nvpad);
  if (q >= q1) return;
  const int32_t e = elem[q];
  const int64_t g = node[q];
  const uint4 mv = meta[q];
  const uint8_t* m = reinterpret_cast<const uint8_t*>(&mv);
  const int64_t k0 = a.nptr[g];
  double* rowA = a.valsA + 9 * k0;
  const int deg = (int)(a.nptr[g + 1] - k0);
  double* rowBT = nullptr;
  int pdeg = 0;
  if (a.with_bt) {
    const int64_t b0 = a.pptr[g];
    rowBT = a.valsBT + 3 * b0;
    pdeg = (int)(a.pptr[g + 1] - b0);
  }
  if (vert)
    VelRow<T, 0, 0>::run(a, e, m, rowA, deg, rowBT, pdeg, nullptr);
  else
    VelRow<T, 1, 0>::run(a, e, m, rowA, deg, rowBT, pdeg, nullptr);
}

// Pressure rows of B by the coloured scatter: one lane per (element, pressure vertex)
// pair of one colour, the row of k_ns_pres_rows added into the global values.
template <typename T>
__global__ __launch_bounds__(256) void k_ns_pres_scatter(NsArgs a, const int32_t* __restrict__ elem,
                                                         const int32_t* __restrict__ node,
                                                         const uint4* __restrict__ meta, int64_t q0, int64_t q1) {
  const int64_t q = q0 + blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (q >= q1) return;
  const int32_t e = elem[q];
  const uint4 mv = meta[q];
  const uint8_t* m = reinterpret_cast<const uint8_t*>(&mv);
  const int li = m[0];
  int32_t nodes[4];
#pragma unroll
  for (int c = 0; c < 4; ++c) nodes[c] = a.conn[(int64_t)c_perm2[li][c] * a.nelem + e];
  Geo g;
  geometry(a.xyz, a.npts, nodes, g);
  double* row = a.valsB + 3 * a.bptr[node[q]];
#pragma unroll
  for (int c = 0; c < 10; ++c)
#pragma unroll
    for (int d = 0; d < 3; ++d) {
      T v = 0;
#pragma unroll
      for (int ax = 0; ax < 3; ++ax)
        if (ft::RP[ax][c] != 0.0) v += (T)g.Bi[ax][d] * (T)ft::RP[ax][c];
      atomicAdd(row + 3 * m[2 + c] + d, -(double)v * g.adet);
    }
}

__global__ void k_zero_f64(int64_t n, double* __restrict__ p) {
  const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (i < n) p[i] = 0.0;
}

// Residual (assembleRHS, :200-224): velocity rows r_(i,d) = c00 sum_j S_ij u_jd
// - c01 sum_m B_m,(i,d) p_m ; pressure rows r_m = -c10 sum_(j,d) B_m,(j,d) u_jd.
template <typename T, int CLS>
__device__ __forceinline__ void res_row(const NsArgs& a, int32_t e, int li, double r[3]) {
  using NT = NsTensors<CLS>;
  int32_t nodes[10];
#pragma unroll
  for (int c = 0; c < 10; ++c) nodes[c] = a.conn[(int64_t)c_perm2[li][c] * a.nelem + e];
  Geo g;
  geometry(a.xyz, a.npts, nodes, g);
  T Bi[3][3];
#pragma unroll
  for (int i = 0; i < 3; ++i)
#pragma unroll
    for (int j = 0; j < 3; ++j) Bi[i][j] = (T)g.Bi[i][j];
  T Gs[6];
  metric<T>(Bi, Gs);
  T S[10];
#pragma unroll
  for (int c = 0; c < 10; ++c) {
    T acc = 0;
#pragma unroll
    for (int s = 0; s < 6; ++s)
      if (NT::KS[s][c] != 0.0) acc += Gs[s] * (T)NT::KS[s][c];
    S[c] = (T)a.nurho * acc;
  }
  T uk[10][3];
#pragma unroll
  for (int k = 0; k < 10; ++k)
#pragma unroll
    for (int d = 0; d < 3; ++d) uk[k][d] = (T)a.u[3 * (int64_t)nodes[k] + d];
  {
    T Nr[10];
    convection_row<T, CLS>(Bi, uk, Nr);
#pragma unroll
    for (int c = 0; c < 10; ++c) S[c] += (T)a.rho * Nr[c];
  }
  T ru[3] = {0, 0, 0};
#pragma unroll
  for (int c = 0; c < 10; ++c)
#pragma unroll
    for (int d = 0; d < 3; ++d) ru[d] += S[c] * uk[c][d];
  T rp[3] = {0, 0, 0};
#pragma unroll
  for (int mm = 0; mm < 4; ++mm) {
    const T pm = (T)a.p[a.pconn[(int64_t)c_perm2[li][mm] * a.nelem + e]];
#pragma unroll
    for (int d = 0; d < 3; ++d) {
      T v = 0;
#pragma unroll
      for (int ax = 0; ax < 3; ++ax)
        if (NT::RV[ax][mm] != 0.0) v += Bi[ax][d] * (T)NT::RV[ax][mm];
      rp[d] -= v * pm;
    }
  }
#pragma unroll
  for (int d = 0; d < 3; ++d) r[d] = g.adet * ((double)((T)a.coeff[0] * ru[d]) + (double)((T)a.coeff[1] * rp[d]));
}

// Residual rows with one wave per chunk of the velocity plan, four chunks per
// workgroup.  The plan's chunks are sized for the matrix rows (about 130 pairs each
// at 40^3), so a workgroup of 256 threads per chunk leaves half its lanes idle; here a
// wave walks its chunk's vertex pairs, then its edge pairs, 64 at a time (each pass
// class-uniform), and accumulates the chunk's 3 x (nodes) residual values in its own
// LDS slice.
constexpr int kResWaves = 4;
__device__ __forceinline__ void wave_lds_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront", "local");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront", "local");
}
template <typename T>
__global__ __launch_bounds__(64 * kResWaves) void k_ns_res_vel(NsArgs a) {
  __shared__ double lds[kResWaves][3 * kNsChunkNodes];
  const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int64_t ch = kResWaves * xcd_chunk(blockIdx.x, gridDim.x) + w;
  if (ch >= a.nchunks) return;  // the whole wave; no block barrier follows
  double* acc = lds[w];
  const int n0 = a.chunk_node[ch], n1 = a.chunk_node[ch + 1];
  const int q0 = a.chunk_pair[ch], q1 = a.chunk_pair[ch + 1], qs = a.chunk_split[ch];
  const int len = 3 * (n1 - n0);
  for (int k = lane; k < len; k += 64) acc[k] = 0.0;
  wave_lds_sync();
  for (int q = q0 + lane; q < qs; q += 64) {
    const int32_t e = a.pair_elem[q];
    if (e < 0) continue;  // padding pair
    const uint4 mv = a.pair_meta[q];
    const uint8_t* m = reinterpret_cast<const uint8_t*>(&mv);
    double r[3];
    res_row<T, 0>(a, e, m[0], r);
    atomicAdd(acc + 3 * m[1] + 0, r[0]);
    atomicAdd(acc + 3 * m[1] + 1, r[1]);
    atomicAdd(acc + 3 * m[1] + 2, r[2]);
  }
  for (int q = qs + lane; q < q1; q += 64) {
    const int32_t e = a.pair_elem[q];
    if (e < 0) continue;
    const uint4 mv = a.pair_meta[q];
    const uint8_t* m = reinterpret_cast<const uint8_t*>(&mv);
    double r[3];
    res_row<T, 1>(a, e, m[0], r);
    atomicAdd(acc + 3 * m[1] + 0, r[0]);
    atomicAdd(acc + 3 * m[1] + 1, r[1]);
    atomicAdd(acc + 3 * m[1] + 2, r[2]);
  }
  wave_lds_sync();
  for (int k = lane; k < len; k += 64) a.res_u[3 * (int64_t)n0 + k] = acc[k];
}

template <typename T>
__global__ __launch_bounds__(256) void k_ns_res_pres(NsArgs a) {
  extern __shared__ double lds[];
  const int ch = (int)xcd_chunk(blockIdx.x, gridDim.x);
  const int n0 = a.chunk_node[ch], n1 = a.chunk_node[ch + 1];
  const int q0 = a.chunk_pair[ch], q1 = a.chunk_pair[ch + 1];
  const int len = n1 - n0;
  for (int k = threadIdx.x; k < len; k += blockDim.x) lds[k] = 0.0;
  __syncthreads();
  for (int q = q0 + threadIdx.x; q < q1; q += blockDim.x) {
    const int32_t e = a.pair_elem[q];
    if (e < 0) continue;  // padding pair
    const uint4 mv = a.pair_meta[q];
    const uint8_t* m = reinterpret_cast<const uint8_t*>(&mv);
    const int li = m[0];
    int32_t nodes[10];
#pragma unroll
    for (int c = 0; c < 10; ++c) nodes[c] = a.conn[(int64_t)c_perm2[li][c] * a.nelem + e];
    Geo g;
    geometry(a.xyz, a.npts, nodes, g);
    T acc = 0;
#pragma unroll
    for (int c = 0; c < 10; ++c)
#pragma unroll
      for (int d = 0; d < 3; ++d) {
        T v = 0;
#pragma unroll
        for (int ax = 0; ax < 3; ++ax)
          if (ft::RP[ax][c] != 0.0) v += (T)g.Bi[ax][d] * (T)ft::RP[ax][c];
        acc += v * (T)a.u[3 * (int64_t)nodes[c] + d];
      }
    atomicAdd(lds + m[1], -(double)((T)a.coeff[2] * acc) * g.adet);
  }
  __syncthreads();
  for (int k = threadIdx.x; k < len; k += blockDim.x) a.res_p[n0 + k] = lds[k];
}

// ---------------------------------------------------------------------------
// Scalar Laplace (FE::assemblyLaplace, FE_def.hpp:2437-2500) and load vector
// (FE::assemblyRHS, FE_def.hpp:7303-7395), P1 or P2.
// ---------------------------------------------------------------------------
struct LapArgs {
  const double* xyz;
  int64_t npts;
  const int32_t* conn;
  int64_t nelem;
  int npe;
  const int64_t* nptr;
  const int32_t* chunk_node;
  const int32_t* chunk_pair;
  const int32_t* chunk_split;
  const int32_t* pair_elem;
  const uint4* pair_meta;
  double* vals;
  double* rhs;
  double f;
};

template <typename T, int CLS>
__device__ __forceinline__ void lap_row(const LapArgs& a, int32_t e, const uint8_t* m,
                                        double* __restrict__ row) {
  const int li = m[0];
  int32_t nodes[4];
  if (a.npe == 4) {
#pragma unroll
    for (int c = 0; c < 4; ++c) nodes[c] = a.conn[(int64_t)c_perm1[li][c] * a.nelem + e];
  } else {
#pragma unroll
    for (int c = 0; c < 4; ++c) nodes[c] = a.conn[(int64_t)c_perm2[li][c] * a.nelem + e];
  }
  Geo g;
  geometry(a.xyz, a.npts, nodes, g);
  T Bi[3][3];
#pragma unroll
  for (int i = 0; i < 3; ++i)
#pragma unroll
    for (int j = 0; j < 3; ++j) Bi[i][j] = (T)g.Bi[i][j];
  T Gs[6];
  metric<T>(Bi, Gs);
  if (a.npe == 4) {
#pragma unroll
    for (int c = 0; c < 4; ++c) {
      T acc = 0;
#pragma unroll
      for (int s = 0; s < 6; ++s)
        if (ft::KS1[s][c] != 0.0) acc += Gs[s] * (T)ft::KS1[s][c];
      atomicAdd(row + m[2 + c], (double)acc * g.adet);
    }
  } else {
    using NT = NsTensors<CLS>;
#pragma unroll
    for (int c = 0; c < 10; ++c) {
      T acc = 0;
#pragma unroll
      for (int s = 0; s < 6; ++s)
        if (NT::KS[s][c] != 0.0) acc += Gs[s] * (T)NT::KS[s][c];
      atomicAdd(row + m[2 + c], (double)acc * g.adet);
    }
  }
}

template <typename T>
__global__ __launch_bounds__(256) void k_laplace_rows(LapArgs a) {
  extern __shared__ double lds[];
  const int ch = (int)xcd_chunk(blockIdx.x, gridDim.x);
  const int n0 = a.chunk_node[ch], n1 = a.chunk_node[ch + 1];
  const int q0 = a.chunk_pair[ch], q1 = a.chunk_pair[ch + 1], qs = a.chunk_split[ch];
  const int64_t na0 = a.nptr[n0];
  const int len = (int)(a.nptr[n1] - na0);
  for (int k = threadIdx.x; k < len; k += blockDim.x) lds[k] = 0.0;
  __syncthreads();
  for (int q = q0 + threadIdx.x; q < q1; q += blockDim.x) {
    const int32_t e = a.pair_elem[q];
    if (e < 0) continue;  // padding pair
    const uint4 mv = a.pair_meta[q];
    const uint8_t* m = reinterpret_cast<const uint8_t*>(&mv);
    double* row = lds + (a.nptr[n0 + m[1]] - na0);
    if (q < qs)
      lap_row<T, 0>(a, e, m, row);
    else
      lap_row<T, 1>(a, e, m, row);
  }
  __syncthreads();
  for (int k = threadIdx.x; k < len; k += blockDim.x) a.vals[na0 + k] = lds[k];
}

// FE::assemblyLinElasXDim (FE_def.hpp:4701-5004), 3D: the element entry for rows (r, a),
// columns (c, b) is int 2 mu eps(phi_r e_a) : eps(phi_c e_b) + lambda div div
//   = |det| (mu delta_ab tr M + mu M[b][a] + lambda M[a][b]),
// M[p][q] = int_ref d_p phi_r d_q phi_c after the affine map (exact reference tensors;
// the reference's 1-point P1 and 5-point P2 rules integrate these products exactly).
template <typename T, int NPE, int CLS>
__device__ __forceinline__ void linelas_row(const LapArgs& a, int32_t e, const uint8_t* m, double* __restrict__ row,
                                            int deg, T lam, T mu) {
  const int li = m[0];
  int32_t nodes[4];
#pragma unroll
  for (int c = 0; c < 4; ++c)
    nodes[c] = a.conn[(int64_t)(NPE == 4 ? c_perm1[li][c] : c_perm2[li][c]) * a.nelem + e];
  Geo g;
  geometry(a.xyz, a.npts, nodes, g);
  T Bi[3][3];
#pragma unroll
  for (int i = 0; i < 3; ++i)
#pragma unroll
    for (int j = 0; j < 3; ++j) Bi[i][j] = (T)g.Bi[i][j];
#pragma unroll
  for (int c = 0; c < NPE; ++c) {
    T t[3][3];  // t[ax][q] = sum_ay RG[ax][ay][c] Bi[ay][q]
#pragma unroll
    for (int ax = 0; ax < 3; ++ax)
#pragma unroll
      for (int q = 0; q < 3; ++q) {
        T v = 0;
#pragma unroll
        for (int ay = 0; ay < 3; ++ay) {
          const double r = NPE == 4 ? ft::RG1[ax][ay][c] : NsTensors<CLS>::RG[ax][ay][c];
          if (r != 0.0) v += (T)r * Bi[ay][q];
        }
        t[ax][q] = v;
      }
    T M[3][3];
#pragma unroll
    for (int p = 0; p < 3; ++p)
#pragma unroll
      for (int q = 0; q < 3; ++q) M[p][q] = Bi[0][p] * t[0][q] + Bi[1][p] * t[1][q] + Bi[2][p] * t[2][q];
    const T tr = M[0][0] + M[1][1] + M[2][2];
#pragma unroll
    for (int d1 = 0; d1 < 3; ++d1)
#pragma unroll
      for (int d2 = 0; d2 < 3; ++d2) {
        const T v = lam * M[d1][d2] + mu * M[d2][d1] + (d1 == d2 ? mu * tr : (T)0);
        atomicAdd(row + d1 * 3 * deg + 3 * m[2 + c] + d2, (double)v * g.adet);
      }
  }
}

template <typename T, int NPE>
__global__ __launch_bounds__(256) void k_linelas_rows(LapArgs a, double lam, double mu) {
  extern __shared__ double lds[];
  __shared__ int2 tab[256];  // per chunk node: row offset (9 per node block), degree
  const int ch = (int)xcd_chunk(blockIdx.x, gridDim.x);
  const int n0 = a.chunk_node[ch], n1 = a.chunk_node[ch + 1];
  const int q0 = a.chunk_pair[ch], q1 = a.chunk_pair[ch + 1], qs = a.chunk_split[ch];
  const int64_t na0 = a.nptr[n0];
  const int len = (int)(9 * (a.nptr[n1] - na0));
  for (int t = threadIdx.x; t < n1 - n0; t += blockDim.x) {
    const int64_t ai = a.nptr[n0 + t];
    tab[t] = make_int2((int)(9 * (ai - na0)), (int)(a.nptr[n0 + t + 1] - ai));
  }
  for (int k = threadIdx.x; k < len; k += blockDim.x) lds[k] = 0.0;
  __syncthreads();
  for (int q = q0 + threadIdx.x; q < q1; q += blockDim.x) {
    const int32_t e = a.pair_elem[q];
    if (e < 0) continue;  // padding pair
    const uint4 mv = a.pair_meta[q];
    const uint8_t* m = reinterpret_cast<const uint8_t*>(&mv);
    const int2 nd = tab[m[1]];
    if (NPE == 4)
      linelas_row<T, 4, 0>(a, e, m, lds + nd.x, nd.y, (T)lam, (T)mu);
    else if (q < qs)
      linelas_row<T, 10, 0>(a, e, m, lds + nd.x, nd.y, (T)lam, (T)mu);
    else
      linelas_row<T, 10, 1>(a, e, m, lds + nd.x, nd.y, (T)lam, (T)mu);
  }
  __syncthreads();
  for (int k = threadIdx.x; k < len; k += blockDim.x) __builtin_nontemporal_store(lds[k], a.vals + 9 * na0 + k);
}

__global__ __launch_bounds__(256) void k_rhs_rows(LapArgs a) {
  extern __shared__ double lds[];
  const int ch = (int)xcd_chunk(blockIdx.x, gridDim.x);
  const int n0 = a.chunk_node[ch], n1 = a.chunk_node[ch + 1];
  const int q0 = a.chunk_pair[ch], q1 = a.chunk_pair[ch + 1];
  const int len = n1 - n0;
  for (int k = threadIdx.x; k < len; k += blockDim.x) lds[k] = 0.0;
  __syncthreads();
  for (int q = q0 + threadIdx.x; q < q1; q += blockDim.x) {
    const int32_t e = a.pair_elem[q];
    if (e < 0) continue;  // padding pair
    const uint4 mv = a.pair_meta[q];
    const uint8_t* m = reinterpret_cast<const uint8_t*>(&mv);
    const int li = m[0];
    int32_t nodes[4];
#pragma unroll
    for (int c = 0; c < 4; ++c)
      nodes[c] = a.conn[(int64_t)(a.npe == 4 ? c_perm1[li][c] : c_perm2[li][c]) * a.nelem + e];
    Geo g;
    geometry(a.xyz, a.npts, nodes, g);
    // int phi_r over the reference tet (exact for the 1-pt P1 / 5-pt P2 rules)
    const double mv0 = a.npe == 4 ? ft::MV1[0] : (li < 4 ? ft::MV2[0] : ft::MV2[4]);
    atomicAdd(lds + m[1], mv0 * a.f * g.adet);
  }
  __syncthreads();
  for (int k = threadIdx.x; k < len; k += blockDim.x) a.rhs[n0 + k] = lds[k];
}

// FE::assemblyRHS with a general source (FE_def.hpp:7352-7391).  The rule is
// determineDegree(dim, FEType, Std): P1 -> 1 point, P2 -> degree 2 raised to 3 -> the
// 5-point rule, in the order of Helper::getQuadratureValues (Helper.cpp:1310-1350).
struct RhsRule {
  int nq;
  double xi[5][3], w[5];
};
__host__ __device__ constexpr RhsRule rhs_rule(int npe) {
  RhsRule r{};
  if (npe == 4) {
    r.nq = 1;
    r.xi[0][0] = r.xi[0][1] = r.xi[0][2] = 0.25;
    r.w[0] = 1.0 / 6.0;
  } else {
    r.nq = 5;
    const double a = 1.0 / 6.0, b = 0.5;
    const double pts[5][3] = {{0.25, 0.25, 0.25}, {a, a, a}, {a, a, b}, {a, b, a}, {b, a, a}};
    for (int q = 0; q < 5; ++q)
      for (int d = 0; d < 3; ++d) r.xi[q][d] = pts[q][d];
    r.w[0] = -2.0 / 15.0;
    for (int q = 1; q < 5; ++q) r.w[q] = 3.0 / 40.0;
  }
  return r;
}
// w_q phi_i(xi_q) of the rule (Helper::phi, Helper.cpp:660-710), node order of the mesh.
struct RhsPhiW {
  double v[10][5];
};
__host__ __device__ constexpr RhsPhiW rhs_phiw(int npe) {
  const RhsRule r = rhs_rule(npe);
  RhsPhiW t{};
  for (int q = 0; q < r.nq; ++q) {
    const double x = r.xi[q][0], y = r.xi[q][1], z = r.xi[q][2], l = 1.0 - x - y - z;
    double ph[10] = {};
    if (npe == 4) {
      ph[0] = l; ph[1] = x; ph[2] = y; ph[3] = z;
    } else {
      ph[0] = l * (2 * l - 1); ph[1] = x * (2 * x - 1); ph[2] = y * (2 * y - 1); ph[3] = z * (2 * z - 1);
      ph[4] = 4 * x * l; ph[5] = 4 * x * y; ph[6] = 4 * y * l; ph[7] = 4 * z * l; ph[8] = 4 * x * z;
      ph[9] = 4 * y * z;
    }
    for (int i = 0; i < npe; ++i) t.v[i][q] = r.w[q] * ph[i];
  }
  return t;
}
__constant__ RhsRule c_rhs_rule[2] = {rhs_rule(4), rhs_rule(10)};
__constant__ RhsPhiW c_rhs_phiw[2] = {rhs_phiw(4), rhs_phiw(10)};

// Physical points x_q = x_0 + B xi_q of every element (quadPointsTrans, :7357-7366).
__global__ __launch_bounds__(256) void k_rhs_qpoints(const double* __restrict__ xyz, int64_t npts,
                                                      const int32_t* __restrict__ conn, int64_t nelem, int npe,
                                                      double* __restrict__ xq) {
  const RhsRule& R = c_rhs_rule[npe == 4 ? 0 : 1];
  const int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (t >= nelem * R.nq) return;
  const int64_t e = t / R.nq;
  const int q = (int)(t - e * R.nq);
  double x[4][3];
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    const int32_t n = conn[(int64_t)k * nelem + e];
#pragma unroll
    for (int d = 0; d < 3; ++d) x[k][d] = xyz[d * npts + n];
  }
#pragma unroll
  for (int j = 0; j < 3; ++j) {
    double v = 0.0;
#pragma unroll
    for (int k = 0; k < 3; ++k) v += (x[k + 1][j] - x[0][j]) * R.xi[q][k];
    xq[3 * t + j] = v + x[0][j];
  }
}

// b_(i,d) += |det B| sum_q w_q phi_i(xi_q) f_d(x_q), fq[(e nq + q) ncomp + d], row gather
// over the scalar plan; rows ncomp i + d.
__global__ __launch_bounds__(256) void k_rhs_points_rows(LapArgs a, const double* __restrict__ fq, int ncomp) {
  extern __shared__ double lds[];
  const int sel = a.npe == 4 ? 0 : 1;
  const int nq = c_rhs_rule[sel].nq;
  const int ch = (int)xcd_chunk(blockIdx.x, gridDim.x);
  const int n0 = a.chunk_node[ch], n1 = a.chunk_node[ch + 1];
  const int q0 = a.chunk_pair[ch], q1 = a.chunk_pair[ch + 1];
  const int len = ncomp * (n1 - n0);
  for (int k = threadIdx.x; k < len; k += blockDim.x) lds[k] = 0.0;
  __syncthreads();
  for (int q = q0 + threadIdx.x; q < q1; q += blockDim.x) {
    const int32_t e = a.pair_elem[q];
    if (e < 0) continue;  // padding pair
    const uint4 mv = a.pair_meta[q];
    const uint8_t* m = reinterpret_cast<const uint8_t*>(&mv);
    const int li = m[0];
    int32_t nodes[4];
#pragma unroll
    for (int c = 0; c < 4; ++c) nodes[c] = a.conn[(int64_t)c * a.nelem + e];
    Geo g;
    geometry(a.xyz, a.npts, nodes, g);
    const double* pw = c_rhs_phiw[sel].v[li];
    const double* fe = fq + (int64_t)e * nq * ncomp;
    for (int d = 0; d < ncomp; ++d) {
      double v = 0.0;
      for (int k = 0; k < nq; ++k) v += pw[k] * fe[k * ncomp + d];
      atomicAdd(lds + ncomp * m[1] + d, v * g.adet);
    }
  }
  __syncthreads();
  for (int k = threadIdx.x; k < len; k += blockDim.x) a.rhs[(int64_t)ncomp * n0 + k] = lds[k];
}

// ---------------------------------------------------------------------------
// STREAM-style copy (attainable HBM bandwidth; byte-counter calibration)
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_stream_copy(int64_t n2, const dbl2* __restrict__ src,
                                                     dbl2* __restrict__ dst) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n2; i += (int64_t)gridDim.x * blockDim.x)
    __builtin_nontemporal_store(__builtin_nontemporal_load(src + i), dst + i);
}

int launch_stream_copy(Ctx* c, int64_t bytes, int iters, float* ms) {
  DevBuf a, b;
  int rc;
  const int64_t n2 = bytes / 16;
  if ((rc = alloc(c, a, n2 * 16)) || (rc = alloc(c, b, n2 * 16))) return rc;
  FEDD_HIP(c, hipMemsetAsync(a.p, 0, n2 * 16, c->stream));
  const unsigned grid = (unsigned)std::min<int64_t>((n2 + 255) / 256, (int64_t)c->num_cu * 32);
  k_stream_copy<<<grid, 256, 0, c->stream>>>(n2, a.as<dbl2>(), b.as<dbl2>());
  FEDD_HIP(c, hipGetLastError());
  hipEvent_t e0, e1;
  FEDD_HIP(c, hipEventCreate(&e0));
  FEDD_HIP(c, hipEventCreate(&e1));
  FEDD_HIP(c, hipEventRecord(e0, c->stream));
  for (int k = 0; k < iters; ++k) k_stream_copy<<<grid, 256, 0, c->stream>>>(n2, a.as<dbl2>(), b.as<dbl2>());
  FEDD_HIP(c, hipEventRecord(e1, c->stream));
  FEDD_HIP(c, hipEventSynchronize(e1));
  float t = 0.f;
  FEDD_HIP(c, hipEventElapsedTime(&t, e0, e1));
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  if (ms) *ms = t / std::max(iters, 1);
  return FEDD_OK;
}

// ---------------------------------------------------------------------------
// launchers
// ---------------------------------------------------------------------------
static NsArgs ns_args(Ctx* c, const fedd_ns_params* p, const GatherPlan& pl) {
  NsArgs a{};
  a.xyz = c->mesh[0].d_xyz.as<double>();
  a.npts = c->mesh[0].npts;
  a.conn = c->mesh[0].d_conn.as<int32_t>();
  a.pconn = c->mesh[1].d_conn.as<int32_t>();
  a.nelem = c->mesh[0].nelem;
  a.u = c->vec[FEDD_VEC_U].as<double>();
  a.p = c->vec[FEDD_VEC_P].as<double>();
  a.nptr = c->d_nptr.as<int64_t>();
  a.pptr = c->d_pptr.as<int64_t>();
  a.bptr = c->d_bptr.as<int64_t>();
  a.chunk_node = pl.d_chunk_node.as<int32_t>();
  a.chunk_pair = pl.d_chunk_pair.as<int32_t>();
  a.chunk_split = pl.d_chunk_split.as<int32_t>();
  a.pair_elem = pl.d_pair_elem.as<int32_t>();
  a.pair_meta = pl.d_pair_meta.as<uint4>();
  a.nchunks = pl.nchunks;
  a.valsA = c->blk[0].d_vals.as<double>();
  a.valsBT = c->blk[1].d_vals.as<double>();
  a.valsB = c->blk[2].d_vals.as<double>();
  a.res_u = c->vec[FEDD_VEC_RES_U].as<double>();
  a.res_p = c->vec[FEDD_VEC_RES_P].as<double>();
  a.nurho = p->viscosity * p->density;
  a.rho = p->density;
  for (int k = 0; k < 4; ++k) a.coeff[k] = p->coeff[k];
  a.newton = p->newton ? 1 : 0;
  return a;
}

// The NS residual (both rows, FP64 in both precision modes).
static int launch_ns_residual(Ctx* c, const fedd_ns_params* p) {
  NsArgs a = ns_args(c, p, c->vplan);
  ktimer_begin(c);
  k_ns_res_vel<double><<<(unsigned)((c->vplan.nchunks + kResWaves - 1) / kResWaves), 64 * kResWaves, 0,
                          c->stream>>>(a);
  FEDD_HIP(c, hipGetLastError());
  NsArgs b = ns_args(c, p, c->pplan);
  k_ns_res_pres<double><<<(unsigned)c->pplan.nchunks, 256, 256 * 8, c->stream>>>(b);
  FEDD_HIP(c, hipGetLastError());
  ktimer_end(c, "ns_residual");
  return FEDD_OK;
}

// The matrices by the coloured scatter: zero the values, then one launch per colour
// for the velocity rows (A, B^T) and, unless reAssemble, one per colour for B.
static int launch_ns_scatter(Ctx* c, const fedd_ns_params* q, int reassemble, bool f32) {
  const ScatterPlan& sp = c->scat;
  NsArgs a = ns_args(c, q, c->vplan);
  a.with_bt = reassemble ? 0 : 1;
  ktimer_begin(c);
  const int nb = reassemble ? 1 : 3;
  for (int k = 0; k < nb; ++k) {
    const int64_t n = c->blk[k].nnz;
    if (n > 0) k_zero_f64<<<nblk(n, 256), 256, 0, c->stream>>>(n, c->blk[k].d_vals.as<double>());
  }
  const int32_t* ve = sp.d_velem.as<int32_t>();
  const int32_t* vn = sp.d_vnode.as<int32_t>();
  const uint4* vm = sp.d_vmeta.as<uint4>();
  for (int k = 0; k < sp.ncolours; ++k) {
    const int64_t q0 = sp.vq[k], qs = sp.vsplit[k], q1 = sp.vq[k + 1];
    const int64_t nt = ((qs - q0 + 63) & ~(int64_t)63) + (q1 - qs);
    if (nt == 0) continue;
    if (f32)
      k_ns_vel_scatter<float><<<nblk(nt, 256), 256, 0, c->stream>>>(a, ve, vn, vm, q0, qs, q1);
    else
      k_ns_vel_scatter<double><<<nblk(nt, 256), 256, 0, c->stream>>>(a, ve, vn, vm, q0, qs, q1);
  }
  if (!reassemble)
    for (int k = 0; k < sp.ncolours; ++k) {
      const int64_t q0 = sp.pq[k], q1 = sp.pq[k + 1];
      if (q1 == q0) continue;
      if (f32)
        k_ns_pres_scatter<float><<<nblk(q1 - q0, 256), 256, 0, c->stream>>>(
            a, sp.d_pelem.as<int32_t>(), sp.d_pnode.as<int32_t>(), sp.d_pmeta.as<uint4>(), q0, q1);
      else
        k_ns_pres_scatter<double><<<nblk(q1 - q0, 256), 256, 0, c->stream>>>(
            a, sp.d_pelem.as<int32_t>(), sp.d_pnode.as<int32_t>(), sp.d_pmeta.as<uint4>(), q0, q1);
    }
  FEDD_HIP(c, hipGetLastError());
  ktimer_end(c, "ns_scatter");
  return FEDD_OK;
}

int launch_ns_assembly(Ctx* c, const fedd_ns_params* p, int mode, int reassemble, int precision) {
  const bool f32 = precision == FEDD_FP32_LOCAL;
  const bool fused = mode == FEDD_NS_JACOBIAN_RHS || mode == FEDD_NS_FIXEDPOINT_RHS;
  if (mode != FEDD_NS_RHS && c->colour_scatter && c->scat.ncolours > 0) {
    fedd_ns_params q = *p;
    if (mode == FEDD_NS_FIXEDPOINT || mode == FEDD_NS_FIXEDPOINT_RHS) q.newton = 0;  // as below
    int rc = launch_ns_scatter(c, &q, reassemble, f32);
    if (rc == FEDD_OK && fused) rc = launch_ns_residual(c, p);
    return rc;
  }
  if (mode != FEDD_NS_RHS) {
    fedd_ns_params q = *p;
    // the fixed-point matrix never holds W (:115-146)
    if (mode == FEDD_NS_FIXEDPOINT || mode == FEDD_NS_FIXEDPOINT_RHS) q.newton = 0;
    NsArgs a = ns_args(c, &q, c->vplan);
    a.with_bt = reassemble ? 0 : 1;
    // fused: 3 residual slots per chunk node after the rows (at most 256 nodes per chunk)
    const size_t lds = (size_t)c->vplan.lds_doubles * 8 + (fused ? 3 * kNsChunkNodes * 8 : 0);
    // FEDD_ROW_BLOCKS_PER_CU > 0: that many workgroups per CU walk the chunks
    // (measured slower on MI355X than one workgroup per chunk)
    const unsigned grid = (unsigned)(FEDD_ROW_BLOCKS_PER_CU > 0 ? std::min<int64_t>(c->vplan.nchunks,
                                                                                 (int64_t)c->num_cu * FEDD_ROW_BLOCKS_PER_CU)
                                                                : c->vplan.nchunks);
    // B does not depend on the iterate: its rows are assembled on the side stream,
    // overlapping the velocity rows (with the pressure residual in the fused modes).
    const bool fork = !reassemble || fused;
    if (fork) {
      NsArgs b = ns_args(c, &q, c->pplan);
      const size_t ldsb = (reassemble ? 0 : (size_t)c->pplan.lds_doubles * 8) + (fused ? 256 * 8 : 0);
      const unsigned gb = (unsigned)c->pplan.nchunks;
      FEDD_HIP(c, hipEventRecord(c->ev_fork, c->stream));
      FEDD_HIP(c, hipStreamWaitEvent(c->side, c->ev_fork, 0));
      ktimer_begin(c, c->side);
      if (!fused) {
        if (f32)
          k_ns_pres_rows<float, 0, 1><<<gb, 256, ldsb, c->side>>>(b);
        else
          k_ns_pres_rows<double, 0, 1><<<gb, 256, ldsb, c->side>>>(b);
      } else if (reassemble) {
        k_ns_pres_rows<double, 1, 0><<<gb, 256, ldsb, c->side>>>(b);
      } else {
        if (f32)
          k_ns_pres_rows<float, 1, 1><<<gb, 256, ldsb, c->side>>>(b);
        else
          k_ns_pres_rows<double, 1, 1><<<gb, 256, ldsb, c->side>>>(b);
      }
      FEDD_HIP(c, hipGetLastError());
      ktimer_end(c, "ns_pres_rows", c->side);
      FEDD_HIP(c, hipEventRecord(c->ev_join, c->side));
    }
    ktimer_begin(c);
    if (fused) {
      if (f32)
        k_ns_vel_rows<float, 1><<<grid, RowShape<float, 1>::threads, lds, c->stream>>>(a);
      else
        k_ns_vel_rows<double, 1><<<grid, RowShape<double, 1>::threads, lds, c->stream>>>(a);
    } else {
      if (f32)
        k_ns_vel_rows<float, 0><<<grid, RowShape<float, 0>::threads, lds, c->stream>>>(a);
      else
        k_ns_vel_rows<double, 0><<<grid, RowShape<double, 0>::threads, lds, c->stream>>>(a);
    }
    FEDD_HIP(c, hipGetLastError());
    ktimer_end(c, "ns_vel_rows");
    if (fork) FEDD_HIP(c, hipStreamWaitEvent(c->stream, c->ev_join, 0));
  } else {
    // The Newton residual stays FP64 in both precision modes (FP32-local applies to
    // the element matrices only).
    return launch_ns_residual(c, p);
  }
  return FEDD_OK;
}

static LapArgs lap_args(Ctx* c) {
  LapArgs a{};
  const GatherPlan& pl = c->vplan;
  a.xyz = c->mesh[0].d_xyz.as<double>();
  a.npts = c->mesh[0].npts;
  a.conn = c->mesh[0].d_conn.as<int32_t>();
  a.nelem = c->mesh[0].nelem;
  a.npe = c->mesh[0].npe;
  a.nptr = c->d_nptr.as<int64_t>();
  a.chunk_node = pl.d_chunk_node.as<int32_t>();
  a.chunk_pair = pl.d_chunk_pair.as<int32_t>();
  a.chunk_split = pl.d_chunk_split.as<int32_t>();
  a.pair_elem = pl.d_pair_elem.as<int32_t>();
  a.pair_meta = pl.d_pair_meta.as<uint4>();
  a.vals = c->blk[0].d_vals.as<double>();
  a.rhs = c->vec[FEDD_VEC_RHS].as<double>();
  return a;
}

int launch_laplace(Ctx* c, int precision) {
  LapArgs a = lap_args(c);
  const size_t lds = (size_t)c->vplan.lds_doubles * 8;
  ktimer_begin(c);
  if (precision == FEDD_FP32_LOCAL)
    k_laplace_rows<float><<<(unsigned)c->vplan.nchunks, 256, lds, c->stream>>>(a);
  else
    k_laplace_rows<double><<<(unsigned)c->vplan.nchunks, 256, lds, c->stream>>>(a);
  FEDD_HIP(c, hipGetLastError());
  ktimer_end(c, "laplace_rows");
  return FEDD_OK;
}

int launch_linelas(Ctx* c, double lam, double mu, int precision) {
  LapArgs a = lap_args(c);
  const size_t lds = (size_t)c->vplan.lds_doubles * 8;
  const bool p1 = c->mesh[0].npe == 4;
  const unsigned g = (unsigned)c->vplan.nchunks;
  ktimer_begin(c);
  if (precision == FEDD_FP32_LOCAL) {
    if (p1)
      k_linelas_rows<float, 4><<<g, 256, lds, c->stream>>>(a, lam, mu);
    else
      k_linelas_rows<float, 10><<<g, 256, lds, c->stream>>>(a, lam, mu);
  } else {
    if (p1)
      k_linelas_rows<double, 4><<<g, 256, lds, c->stream>>>(a, lam, mu);
    else
      k_linelas_rows<double, 10><<<g, 256, lds, c->stream>>>(a, lam, mu);
  }
  FEDD_HIP(c, hipGetLastError());
  ktimer_end(c, "linelas_rows");
  return FEDD_OK;
}

int launch_rhs_qpoints(Ctx* c, double* d_xq, int* nq) {
  const Mesh& m = c->mesh[0];
  *nq = m.npe == 4 ? 1 : 5;
  const int64_t n = m.nelem * *nq;
  k_rhs_qpoints<<<nblk(n, 256), 256, 0, c->stream>>>(m.d_xyz.as<double>(), m.npts, m.d_conn.as<int32_t>(), m.nelem,
                                                     m.npe, d_xq);
  FEDD_HIP(c, hipGetLastError());
  return FEDD_OK;
}

int launch_rhs_points(Ctx* c, const double* d_fq, int ncomp) {
  LapArgs a = lap_args(c);
  ktimer_begin(c);
  k_rhs_points_rows<<<(unsigned)c->vplan.nchunks, 256, (size_t)ncomp * 256 * 8, c->stream>>>(a, d_fq, ncomp);
  FEDD_HIP(c, hipGetLastError());
  ktimer_end(c, "rhs_points_rows");
  return FEDD_OK;
}

int launch_rhs(Ctx* c, double f) {
  LapArgs a = lap_args(c);
  a.f = f;
  ktimer_begin(c);
  k_rhs_rows<<<(unsigned)c->vplan.nchunks, 256, 256 * 8, c->stream>>>(a);
  FEDD_HIP(c, hipGetLastError());
  ktimer_end(c, "rhs_rows");
  return FEDD_OK;
}

}  // namespace fedd
