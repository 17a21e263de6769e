// Internal types shared by the host code and the kernels of libfeddgpu.
#pragma once
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdint>
#include <map>
#include <string>
#include <thread>
#include <vector>

#include "../../include/feddgpu.h"

namespace fedd {

// Device buffer owned by a context (hipMalloc / hipFree).
struct DevBuf {
  void* p = nullptr;
  size_t bytes = 0;
  DevBuf() = default;
  DevBuf(const DevBuf&) = delete;
  DevBuf& operator=(const DevBuf&) = delete;
  DevBuf(DevBuf&& o) noexcept : p(o.p), bytes(o.bytes) {
    o.p = nullptr;
    o.bytes = 0;
  }
  DevBuf& operator=(DevBuf&& o) noexcept {
    if (this != &o) {
      release();
      p = o.p;
      bytes = o.bytes;
      o.p = nullptr;
      o.bytes = 0;
    }
    return *this;
  }
  ~DevBuf() { release(); }
  void release() {
    if (p) (void)hipFree(p);
    p = nullptr;
    bytes = 0;
  }
  template <class T>
  T* as() const {
    return static_cast<T*>(p);
  }
};

struct Mesh {
  int64_t npts = 0, nelem = 0;
  int npe = 0;
  std::vector<double> pts;     // host, node-major
  std::vector<int32_t> elems;  // host, element-major
  DevBuf d_xyz;                // SoA: x[npts] y[npts] z[npts]
  DevBuf d_conn;               // SoA: conn[k*nelem + e]
  DevBuf d_conn16;             // AoS: conn16[e*16 + k], k < npe (one 64-byte line per element)
  bool valid() const { return npts > 0; }
};

// Node -> element incidence: entries (element, local index), ascending element id.
struct Incidence {
  std::vector<int64_t> ptr;
  std::vector<int32_t> elem;
  std::vector<uint8_t> loc;
};

// Node adjacency (sorted unique neighbour node ids per node).
struct Adjacency {
  std::vector<int64_t> ptr;
  std::vector<int32_t> idx;
  int64_t deg(int64_t i) const { return ptr[i + 1] - ptr[i]; }
};

// One CRS block: Tpetra-identical graph + values, on the device.
struct Block {
  int64_t nrows = 0, ncols = 0, nnz = 0;
  DevBuf d_rowptr;  // int64, nrows+1
  DevBuf d_colind;  // int32, nnz
  DevBuf d_vals;    // double, nnz
};

// Row-gather schedule over node chunks (see fedd_ns.hip for the layout).
// Pairs come in segments of 64 (one wave each); padding pairs have element -1.
struct GatherPlan {
  int64_t nchunks = 0, npairs = 0;  // npairs includes the padding pairs
  int lds_doubles = 0;          // max LDS doubles of any chunk
  bool ordered = true;          // every node's pairs lie in one segment (reproducible sums)
  DevBuf d_chunk_node;          // int32 [nchunks+1]: first node of each chunk
  DevBuf d_chunk_pair;          // int32 [nchunks+1]: first pair of each chunk
  DevBuf d_chunk_split;         // int32 [nchunks]: first edge-class pair of each chunk
  DevBuf d_pair_elem;           // int32 [npairs]
  DevBuf d_pair_meta;           // uint4 [npairs]: loc, node-in-chunk, 14 slots
};

// Coloured element scatter of the Navier-Stokes matrices (fedd_set_assembly_method):
// the (element, local node) pairs of every element colour, velocity pairs of a colour
// vertex class (local node < 4) first.  The 16-byte meta is the gather plans' (loc, -,
// 10 column slots in the row, 4 pressure slots); the row is named by its node.
struct ScatterPlan {
  int ncolours = 0;
  std::vector<int64_t> vq, vsplit;  // velocity pairs of colour k: [vq[k], vq[k+1]); edge class from vsplit[k]
  std::vector<int64_t> pq;          // pressure pairs of colour k: [pq[k], pq[k+1])
  DevBuf d_velem, d_vnode, d_vmeta;
  DevBuf d_pelem, d_pnode, d_pmeta;
  // host copies for fedd_scatter_indices: pairs and the node-level row pointers
  std::vector<int32_t> h_velem, h_vnode, h_pelem, h_pnode;
  std::vector<uint8_t> h_vmeta, h_pmeta;
  std::vector<int64_t> a_ptr, bt_ptr, b_ptr;
};

enum GraphKind { GK_NONE = -1 };

// Navier-Stokes velocity row gather (k_ns_vel_rows): LDS budget of a chunk's rows in
// doubles, most nodes per chunk (the size of its per-node table), threads per workgroup.
// 32 KiB of rows + a 1 KiB table: four workgroups of three waves fit a CU's 160 KiB of
// LDS and its 12 waves of 168 VGPRs (three waves per SIMD).  40^3 x 6 Jacobian at steady
// clocks, ms per step with 28 / 30 / 32 / 34 / 36 / 40 KiB: 0.49 / 0.455 / 0.458 / 0.480 /
// 0.471 / 0.53 (the chunk count against the CU count matters more than the size); the
// 64^3 fused FP32-local iteration takes 2.87-2.90 ms with 32 KiB, 2.83-2.86 with 36.
#ifndef FEDD_NS_LDS_BUDGET
#define FEDD_NS_LDS_BUDGET 4096
#endif
#ifndef FEDD_NS_CHUNK_NODES
#define FEDD_NS_CHUNK_NODES 64
#endif
#ifndef FEDD_NS_ROW_THREADS
#define FEDD_NS_ROW_THREADS 192
#endif
constexpr int kNsLdsBudget = FEDD_NS_LDS_BUDGET;
constexpr int kNsChunkNodes = FEDD_NS_CHUNK_NODES;
constexpr int kNsRowThreads = FEDD_NS_ROW_THREADS;

struct Ctx {
  int device = 0;
  int num_cu = 256;  // compute units of the device (hipDeviceProp_t::multiProcessorCount)
  hipStream_t stream = nullptr;
  // second stream for independent work that overlaps the main stream (fork/join by
  // events; default priority like the main stream)
  hipStream_t side = nullptr;
  hipEvent_t ev_fork = nullptr, ev_join = nullptr;
  std::string err;
  Mesh mesh[2];
  int graph_kind = GK_NONE;
  Block blk[4];
  // node-level structures on the device
  DevBuf d_nptr, d_pptr, d_bptr;  // int64 node-level row pointers (A, B^T, B)
  DevBuf d_nidx;                  // int32 node-level column indices of A (node blocks)
  GatherPlan vplan, pplan;        // velocity (primary) and pressure row gathers
  bool reproducible = false;      // fedd_set_reproducible: segmented plans at the next graph build
  bool colour_scatter = false;    // fedd_set_assembly_method: NS matrices by the coloured scatter
  ScatterPlan scat;               // its schedule (built with the NS graph)
  DevBuf d_emod;                  // per-element Young's modulus field (E-field elasticity overload)
  DevBuf d_qp;                    // elasticity quadrature-point records (52 doubles each)
  DevBuf d_rhsq;                  // assemblyRHS: quadrature points / source values at them
  int qp_nq = 0;                  // points per element of the records in d_qp (0: none yet)
  DevBuf d_elas_fix;              // int32: chunks the last FP32-local elasticity launch re-did in FP64, + count
  DevBuf d_elas_flag;             // int32: chunks holding pairs of flagged elements, + count, + pairs, + element flags
  bool elas_fix_valid = false;    // d_elas_fix / d_elas_flag belong to the last elasticity assembly
  // colouring
  std::vector<int64_t> colour_ptr;
  std::vector<int32_t> colour_perm;
  // vectors
  // surface triangles of fedd_surface_set (6 node ids each, FEDDLib P2 triangle order) and
  // the follower pressure per flag of fedd_newton_pressure
  DevBuf d_faces, d_fflags;
  int64_t nface = 0;
  std::vector<double> newton_pressure;
  DevBuf vec[FEDD_NUM_VECTORS];
  int64_t vec_n[FEDD_NUM_VECTORS] = {};
  // Dirichlet constraints and Krylov work space (elasticity Newton solve)
  DevBuf d_bcflag;      // uint8 per dof of the primary field
  int64_t nbc = 0;
  DevBuf d_bcdofs;      // int32: the constrained dofs, ascending (nbc_dofs of them)
  int64_t nbc_dofs = 0;
  int bc_per_node = 3;  // dofs per node the Dirichlet set was made for (1: scalar graph)
  DevBuf d_bcval;       // double per dof of the primary field: Dirichlet values (absent: 0)
  // rows of block A that hold a constrained dof in their row or columns (int32), built
  // on the device for the current Dirichlet set and graph; nbc_rows < 0: not built
  DevBuf d_bcrows;
  int64_t nbc_rows = -1;
  const void* bcrows_graph = nullptr;  // colind buffer the list was built for
  DevBuf d_kw[5];       // r, z, p, Ap and the block-Jacobi inverse (9 per node)
  DevBuf d_bw[7];       // BiCGStab: r, r_hat, p, v, p_hat, s_hat, t
  DevBuf d_bscal;       // BiCGStab dot products (16 sub-slots each)
  DevBuf d_part;        // per-block partial sums
  DevBuf d_scal;        // device scalars of the Krylov loop
  double* h_rr = nullptr;               // pinned: r.r partial sums read back by the CG loop (4 x 16)
  hipEvent_t ev_rr[4] = {};             // completion of each h_rr slot's copy
  // learned tangent surrogate (FP32 master weights + BF16 copies)
  struct MlpHost* mlp = nullptr;
  // geometric multigrid hierarchy (structured lattice meshes), rebuilt on demand
  struct MgHost* mg = nullptr;
  int precond = 0;                      // 0: block Jacobi, 1: geometric multigrid
  std::vector<uint8_t> bcflag_host;     // Dirichlet flags per dof (host copy)
  bool mlp_ready = false;
  int newton_step = 0;
  // stream capture (fedd_capture_begin / end) and the instantiated graph
  bool capturing = false;
  hipGraph_t graph = nullptr;
  hipGraphExec_t gexec = nullptr;
  // why the captured graph was destroyed (a buffer it refers to was freed or moved)
  std::string graph_dropped;
  // timing
  hipEvent_t ev[16] = {};
  bool timing = false;
  std::map<std::string, std::pair<double, int64_t>> ktime;
  struct KEv {
    std::string name;
    hipEvent_t a, b;
  };
  std::vector<KEv> kpending;       // recorded, not yet resolved
  std::vector<hipEvent_t> kpool;   // free events
  hipEvent_t kopen = nullptr;      // start event of the launch being timed
};

// Error helpers --------------------------------------------------------------
int fail(Ctx* c, int code, const std::string& msg);
// Message of the last failed context-free call on this thread (fedd_last_error(NULL)).
std::string& free_error();
#define FEDD_HIP(c, call)                                                        \
  do {                                                                           \
    hipError_t _e = (call);                                                      \
    if (_e != hipSuccess)                                                        \
      return ::fedd::fail((c), FEDD_ERR_HIP,                                     \
                          std::string(#call) + ": " + hipGetErrorString(_e));    \
  } while (0)

// Frees b (after the context's streams are idle) and allocates `bytes`.  Freeing a
// buffer destroys the caller's captured graph first (drop_user_graph).
int alloc(Ctx* c, DevBuf& b, size_t bytes);
// Waits until the context's main and side streams are idle.
void sync_streams(Ctx* c);
// Waits for the streams, then destroys the captured graph of fedd_capture_begin/end:
// its nodes hold the addresses of buffers that are about to be freed or moved, so a
// later fedd_graph_launch must fail (FEDD_ERR_STATE with `why`) instead of replaying.
void drop_user_graph(Ctx* c, const char* why);
int upload(Ctx* c, DevBuf& b, const void* host, size_t bytes);

// Kernel-timing bracket: events recorded on the context stream around a launch
// and resolved lazily (no synchronisation inside the timed region).
// Kernel timers (events around a launch) on the context stream, or on stream s.
void ktimer_begin(Ctx* c, hipStream_t s = nullptr);
void ktimer_end(Ctx* c, const char* name, hipStream_t s = nullptr);
void ktimer_resolve(Ctx* c);

// Launchers implemented in the .hip files -------------------------------------
// Expands node-level adjacency into the dof-level CRS rowptr/colind of each block.
int expand_graph(Ctx* c, int kind, const int32_t* aidx, const int32_t* btidx, const int32_t* bidx);
int launch_ns_assembly(Ctx* c, const fedd_ns_params* p, int mode, int reassemble, int precision);
int launch_laplace(Ctx* c, int precision);
int launch_rhs(Ctx* c, double f);
int launch_rhs_qpoints(Ctx* c, double* d_xq, int* nq);
int launch_rhs_points(Ctx* c, const double* d_fq, int ncomp);
int launch_linelas(Ctx* c, double lam, double mu, int precision);
int launch_stream_copy(Ctx* c, int64_t bytes, int iters, float* ms);
// tangent: 1 = exact law, 2 = learned surrogate (the network writes the tangent slot)
// Evec: device array of per-element Young's moduli (overrides Emod) or null.
int launch_elasticity(Ctx* c, int model, double Emod, double nu, int precision, int with_res, int tangent,
                      const double* Evec = nullptr);
int launch_mlp_tangent(Ctx* c, int nq);
// Follower pressure load (pressures scale * pressure[flag]) into vector `out` (zeroed first
// unless accumulate) and, with_tangent, Kext added to block A.
int surface_pressure(Ctx* c, const double* pressure, int nflag, double scale, int with_tangent, int out,
                     bool accumulate);
void mlp_free(Ctx* c);
// True when the loaded network was built for this law and these constants.
bool mlp_matches(const Ctx* c, int model, double Emod, double nu);
// Appends the quadrature-point states in d_qp to the surrogate's training set.
int mlp_collect_qp(Ctx* c);
// Geometric multigrid (fedd_mg.hip): hierarchy for the current mesh / Dirichlet set,
// coarse operators for the current A, and one V-cycle x = M^-1 b on the finest level.
// mode 0: geometric hierarchy of the structured lattice; 1: aggregation (any mesh).
int mg_setup(Ctx* c, int mode);
int mg_refresh(Ctx* c);
// Finest level only (smoother blocks and FP32 copy); coarse levels keep their operators.
int mg_refresh_fine(Ctx* c);
int mg_vcycle(Ctx* c, const double* b, double* x);
// FP32 copy of the finest operator (block A's values, same order) that the smoother reads,
// as of the last mg_refresh / mg_refresh_fine; null when the hierarchy keeps none.
const float* mg_fine_vals32(Ctx* c);
void mg_free(Ctx* c);
// Looks up this device's copy of the no-allocation solve's work array (fedd_solve.hip).
void noalloc_init(Ctx* c);

// Host worker threads (at most 16).
inline unsigned nthreads() {
  unsigned n = std::thread::hardware_concurrency();
  return std::max(1u, std::min(n, 16u));
}

// f(begin, end, thread) over contiguous ranges of [0, n); serial below 4096 items.
template <class F>
void parallel_for(int64_t n, F f) {
  unsigned T = nthreads();
  if (n < 4096 || T == 1) {
    f(0, n, 0);
    return;
  }
  std::vector<std::thread> th;
  int64_t step = (n + T - 1) / T;
  for (unsigned t = 0; t < T; ++t) {
    int64_t a = t * step, b = std::min(n, a + step);
    if (a >= b) break;
    th.emplace_back(f, a, b, (int)t);
  }
  for (auto& x : th) x.join();
}

}  // namespace fedd

namespace fedd {
// XCD-aware chunk order.  Workgroups are dispatched round-robin over the 8 XCDs
// (block b runs on XCD b % 8) and each XCD has its own L2.  Neighbouring node
// chunks share elements, so each XCD is given one contiguous range of chunks:
// XCD x takes chunks [x*q + min(x,r), ...) in order, q = G / 8, r = G % 8.
// The mapping is a bijection of [0, G) for any G.
#ifndef FEDD_XCD_REMAP
#define FEDD_XCD_REMAP 1
#endif
__device__ __forceinline__ int64_t xcd_chunk(int64_t b, int64_t G) {
#if FEDD_XCD_REMAP
  const int64_t q = G >> 3, r = G & 7, x = b & 7;
  return x * q + (x < r ? x : r) + (b >> 3);
#else
  (void)G;
  return b;
#endif
}
}  // namespace fedd

struct fedd_ctx : public fedd::Ctx {};
