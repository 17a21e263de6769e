// Host side of libfeddgpu: context, mesh upload, Tpetra-identical graph
// construction, row-gather schedules, colouring and the C ABI entry points.
#include <algorithm>
#include <array>
#include <cstdio>
#include <cstring>
#include <limits>
#include <numeric>
#include <thread>

#include "fedd_internal.h"
#include "fedd_perm.h"

namespace fedd {

int fail(Ctx* c, int code, const std::string& msg) {
  if (c) c->err = msg;
  return code;
}

void sync_streams(Ctx* c) {
  (void)hipStreamSynchronize(c->stream);
  (void)hipStreamSynchronize(c->side);
}

void drop_user_graph(Ctx* c, const char* why) {
  if (!c->gexec && !c->graph) return;
  sync_streams(c);  // a replay may still be running
  if (c->gexec) (void)hipGraphExecDestroy(c->gexec);
  if (c->graph) (void)hipGraphDestroy(c->graph);
  c->gexec = nullptr;
  c->graph = nullptr;
  c->graph_dropped = why;
}

int alloc(Ctx* c, DevBuf& b, size_t bytes) {
  if (c->capturing) return fail(c, FEDD_ERR_STATE, "device allocation while a capture is open: run the call once before capturing");
  if (b.p) {
    // kernels queued on the streams or recorded in the captured graph may read b
    drop_user_graph(c, "a device buffer the captured graph uses was reallocated (resized vector, mesh or graph): capture again");
    sync_streams(c);
  }
  b.release();
  if (bytes == 0) return FEDD_OK;
  FEDD_HIP(c, hipMalloc(&b.p, bytes));
  b.bytes = bytes;
  return FEDD_OK;
}

int upload(Ctx* c, DevBuf& b, const void* host, size_t bytes) {
  if (b.bytes != bytes) {  // same size: keep the buffer (captured graphs keep its address)
    int rc = alloc(c, b, bytes);
    if (rc) return rc;
  }
  if (bytes) FEDD_HIP(c, hipMemcpyAsync(b.p, host, bytes, hipMemcpyHostToDevice, c->stream));
  return FEDD_OK;
}

static hipEvent_t kevent(Ctx* c) {
  if (c->kpool.empty()) {
    hipEvent_t e = nullptr;
    (void)hipEventCreate(&e);
    return e;
  }
  hipEvent_t e = c->kpool.back();
  c->kpool.pop_back();
  return e;
}

void ktimer_resolve(Ctx* c) {
  for (auto& k : c->kpending) {
    float ms = 0.f;
    (void)hipEventSynchronize(k.b);
    (void)hipEventElapsedTime(&ms, k.a, k.b);
    auto& t = c->ktime[k.name];
    t.first += ms;
    t.second += 1;
    c->kpool.push_back(k.a);
    c->kpool.push_back(k.b);
  }
  c->kpending.clear();
}

void ktimer_begin(Ctx* c, hipStream_t s) {
  if (!c->timing || c->capturing) return;
  if (c->kpending.size() >= 4096) ktimer_resolve(c);
  c->kopen = kevent(c);
  (void)hipEventRecord(c->kopen, s ? s : c->stream);
}

void ktimer_end(Ctx* c, const char* name, hipStream_t s) {
  if (!c->timing || c->capturing || !c->kopen) return;
  hipEvent_t b = kevent(c);
  (void)hipEventRecord(b, s ? s : c->stream);
  c->kpending.push_back({name, c->kopen, b});
  c->kopen = nullptr;
}

namespace {

void build_incidence(const Mesh& m, Incidence& inc) {
  inc.ptr.assign(m.npts + 1, 0);
  for (int64_t e = 0; e < m.nelem; ++e)
    for (int k = 0; k < m.npe; ++k) inc.ptr[m.elems[e * m.npe + k] + 1]++;
  for (int64_t i = 0; i < m.npts; ++i) inc.ptr[i + 1] += inc.ptr[i];
  inc.elem.resize(inc.ptr[m.npts]);
  inc.loc.resize(inc.ptr[m.npts]);
  std::vector<int64_t> fill(inc.ptr.begin(), inc.ptr.end() - 1);
  for (int64_t e = 0; e < m.nelem; ++e)
    for (int k = 0; k < m.npe; ++k) {
      int32_t n = m.elems[e * m.npe + k];
      int64_t pos = fill[n]++;
      inc.elem[pos] = (int32_t)e;
      inc.loc[pos] = (uint8_t)k;
    }
}

// For every row node i of `rows`, the sorted unique nodes of mesh `cols` found in
// the elements incident to i (the elements of `cols` have the same ids as those
// of `rows`).  This is exactly the column set Tpetra holds after inserting every
// element's rows and calling fillComplete.
void build_adjacency(const Incidence& inc, int64_t nrows, const Mesh& cols, Adjacency& adj) {
  adj.ptr.assign(nrows + 1, 0);
  std::vector<int64_t> cnt(nrows, 0);
  parallel_for(nrows, [&](int64_t a, int64_t b, int) {
    std::vector<int32_t> buf;
    for (int64_t i = a; i < b; ++i) {
      buf.clear();
      for (int64_t q = inc.ptr[i]; q < inc.ptr[i + 1]; ++q) {
        const int32_t* n = &cols.elems[(int64_t)inc.elem[q] * cols.npe];
        buf.insert(buf.end(), n, n + cols.npe);
      }
      std::sort(buf.begin(), buf.end());
      cnt[i] = std::unique(buf.begin(), buf.end()) - buf.begin();
    }
  });
  for (int64_t i = 0; i < nrows; ++i) adj.ptr[i + 1] = adj.ptr[i] + cnt[i];
  adj.idx.resize(adj.ptr[nrows]);
  parallel_for(nrows, [&](int64_t a, int64_t b, int) {
    std::vector<int32_t> buf;
    for (int64_t i = a; i < b; ++i) {
      buf.clear();
      for (int64_t q = inc.ptr[i]; q < inc.ptr[i + 1]; ++q) {
        const int32_t* n = &cols.elems[(int64_t)inc.elem[q] * cols.npe];
        buf.insert(buf.end(), n, n + cols.npe);
      }
      std::sort(buf.begin(), buf.end());
      auto end = std::unique(buf.begin(), buf.end());
      std::copy(buf.begin(), end, adj.idx.begin() + adj.ptr[i]);
    }
  });
}

inline int find_slot(const Adjacency& adj, int64_t i, int32_t j) {
  const int32_t* b = adj.idx.data() + adj.ptr[i];
  const int32_t* e = adj.idx.data() + adj.ptr[i + 1];
  return (int)(std::lower_bound(b, e, j) - b);
}

// Row-gather schedule.  Nodes are cut into contiguous chunks whose rows fit the
// LDS budget.  Inside a chunk the (node, element) pairs are ordered vertex class
// first, then edge class.  Inside a class the pairs are laid out in segments of 64,
// one wave's worth: a segment holds whole nodes (the plan inserts padding pairs,
// element -1, which the kernels skip, so that no node straddles two segments; the
// vertex class is padded to a whole number of segments), and inside a segment the
// k-th pairs of its nodes follow each other, so neighbouring lanes add into different
// rows and gather neighbouring elements.  Every row of a chunk is then added to by one
// wave only, in a fixed order, so the assembled values are reproducible bit for bit.
// Slots are stored in canonical column order (fedd_perm.h).
struct PlanSpec {
  int rows_per_node;   // LDS rows per node of the main block (1 or 3)
  int comps_per_col;   // values per neighbour node per row (1 or 3)
  const Adjacency* main;
  const Adjacency* second;  // B^T-like block (may be null)
  int second_rows;          // rows per node of the second block
  const Mesh* cols_main;    // mesh whose element nodes index `main`
  const Mesh* cols_second;  // mesh whose element nodes index `second`
  int lds_budget_doubles;
  int max_nodes = 256;  // most nodes per chunk
  // false: no segments; the k-th pairs of all the chunk's nodes follow each other inside
  // a class, a node's pairs may lie in several waves (their sums vary in the last bits
  // from run to run), the vertex class is padded in the kernels only
  bool segments = true;
};

constexpr int kSegment = 64;  // pairs per wave

// Meta of the pair (row node i, element e, local index li of i in e): loc, node-in-chunk
// (left 0), the slots of the canonical columns in row i, then the second block's slots.
static void pair_meta_of(const PlanSpec& s, int li, int64_t i, int32_t e, uint8_t m[16]) {
  for (int k = 0; k < 16; ++k) m[k] = 0;
  m[0] = (uint8_t)li;
  // the relabelling acts on the column element: P2 columns use the tet-10
  // permutation (its vertex part is the tet-4 one), P1 columns the tet-4 one
  const int ncol = s.cols_main->npe;
  const int8_t* perm = ncol == 10 ? kPermP2[li] : kPermP1[li];
  for (int cpos = 0; cpos < ncol; ++cpos) {
    const int32_t j = s.cols_main->elems[(int64_t)e * ncol + perm[cpos]];
    m[2 + cpos] = (uint8_t)find_slot(*s.main, i, j);
  }
  if (s.second)
    for (int cpos = 0; cpos < 4; ++cpos) {
      const int32_t j = s.cols_second->elems[(int64_t)e * 4 + perm[cpos]];
      m[12 + cpos] = (uint8_t)find_slot(*s.second, i, j);
    }
}

int build_plan(Ctx* c, const Incidence& inc, int64_t nnodes, const Mesh& rows_mesh,
               const PlanSpec& s, GatherPlan& plan) {
  std::vector<int32_t> chunk_node(1, 0), chunk_pair(1, 0), chunk_split;
  const int64_t npairs = inc.ptr[nnodes];
  std::vector<int32_t> pair_elem;
  std::vector<uint8_t> pair_meta;
  pair_elem.reserve(npairs + npairs / 4);
  pair_meta.reserve(16 * (npairs + npairs / 4));
  int max_lds = 0;
  bool ordered = s.segments;
  std::vector<std::array<int64_t, 3>> order;  // (key, node, incidence entry)
  auto node_lds = [&](int64_t i) {
    int64_t v = (int64_t)s.rows_per_node * s.comps_per_col * s.main->deg(i);
    if (s.second) v += (int64_t)s.second_rows * s.second->deg(i);
    return v;
  };
  for (int64_t i = 0; i < nnodes; ++i) {
    if (s.main->deg(i) > 255 || (s.second && s.second->deg(i) > 255))
      return fail(c, FEDD_ERR_LIMIT, "a node has more than 255 neighbour nodes");
    if (node_lds(i) > s.lds_budget_doubles)
      return fail(c, FEDD_ERR_LIMIT, "a single row block exceeds the LDS budget");
  }
  auto pad_to = [&](int64_t q0, int64_t fill_to) {  // padding pairs up to chunk offset fill_to
    while ((int64_t)pair_elem.size() - q0 < fill_to) {
      pair_elem.push_back(-1);
      pair_meta.insert(pair_meta.end(), 16, (uint8_t)0);
    }
  };
  int64_t n0 = 0;
  auto emit = [&](int64_t i, int64_t q) {  // pair (node i, incidence entry q) of the chunk at n0
    const int32_t e = inc.elem[q];
    pair_elem.push_back(e);
    uint8_t m[16];
    pair_meta_of(s, inc.loc[q], i, e, m);
    m[1] = (uint8_t)(i - n0);
    pair_meta.insert(pair_meta.end(), m, m + 16);
  };
  while (n0 < nnodes) {
    int64_t n1 = n0, lds = 0;
    while (n1 < nnodes && n1 - n0 < s.max_nodes && lds + node_lds(n1) <= s.lds_budget_doubles &&
           (inc.ptr[n1 + 1] - inc.ptr[n0]) <= 1024) {
      lds += node_lds(n1);
      ++n1;
    }
    if (n1 == n0) n1 = n0 + 1;  // cannot happen given the checks above
    max_lds = std::max<int64_t>(max_lds, lds);
    const int64_t q0 = (int64_t)pair_elem.size();
    int32_t split = -1;
    for (int cls = 0; cls < 2; ++cls) {
      if (cls == 1) {
        // the edge class starts on a segment boundary (when the chunk has one)
        bool any = false;
        for (int64_t q = inc.ptr[n0]; q < inc.ptr[n1] && !any; ++q) any = rows_mesh.npe == 10 && inc.loc[q] >= 4;
        const int64_t used = (int64_t)pair_elem.size() - q0;
        if (any && s.segments) pad_to(q0, (used + kSegment - 1) / kSegment * kSegment);
        split = (int32_t)pair_elem.size();
      }
      // (key, node, incidence entry) of this class, in emission order
      order.clear();
      for (int64_t i = n0; i < n1; ++i) {
        int64_t g = 0;
        for (int64_t q = inc.ptr[i]; q < inc.ptr[i + 1]; ++q)
          if (((rows_mesh.npe == 10 && inc.loc[q] >= 4) ? 1 : 0) == cls)
            order.push_back({(g << 20) | (i - n0), i, q}), ++g;
      }
      if (!s.segments) {
        // the k-th pairs of all nodes follow each other, so neighbouring lanes add into
        // different rows and gather neighbouring elements
        std::sort(order.begin(), order.end());
        for (auto& o : order) emit(o[1], o[2]);
        continue;
      }
      // segments: whole nodes per segment, interleaved inside it like the order above
      size_t k = 0;
      while (k < order.size()) {
        const int64_t used = (int64_t)pair_elem.size() - q0;
        std::vector<std::array<int64_t, 3>> seg;
        int64_t fill = used % kSegment;
        while (k < order.size()) {
          size_t g = 0;  // pairs of node order[k][1]
          while (k + g < order.size() && order[k + g][1] == order[k][1]) ++g;
          if (fill > 0 && fill + (int64_t)g > kSegment) break;  // starts the next segment
          if ((int64_t)g > kSegment) ordered = false;  // spans segments: its rows race between waves
          seg.insert(seg.end(), order.begin() + k, order.begin() + k + g);
          fill += (int64_t)g;
          k += g;
          if (fill >= kSegment) break;
        }
        std::sort(seg.begin(), seg.end());
        for (auto& o : seg) emit(o[1], o[2]);
        if (k < order.size()) {
          const int64_t u = (int64_t)pair_elem.size() - q0;
          if (u % kSegment) pad_to(q0, u - u % kSegment + kSegment);
        }
      }
    }
    chunk_split.push_back(split);
    chunk_node.push_back((int32_t)n1);
    chunk_pair.push_back((int32_t)pair_elem.size());
    n0 = n1;
  }
  if (pair_elem.size() > (size_t)std::numeric_limits<int32_t>::max())
    return fail(c, FEDD_ERR_LIMIT, "the row-gather plan holds more than 2^31 pairs");
  plan.ordered = ordered;
  plan.nchunks = (int64_t)chunk_split.size();
  plan.npairs = (int64_t)pair_elem.size();
  plan.lds_doubles = max_lds;
  int rc;
  if ((rc = upload(c, plan.d_chunk_node, chunk_node.data(), chunk_node.size() * 4))) return rc;
  if ((rc = upload(c, plan.d_chunk_pair, chunk_pair.data(), chunk_pair.size() * 4))) return rc;
  if ((rc = upload(c, plan.d_chunk_split, chunk_split.data(), chunk_split.size() * 4))) return rc;
  if ((rc = upload(c, plan.d_pair_elem, pair_elem.data(), pair_elem.size() * 4))) return rc;
  if ((rc = upload(c, plan.d_pair_meta, pair_meta.data(), pair_meta.size()))) return rc;
  FEDD_HIP(c, hipStreamSynchronize(c->stream));  // the host arrays are freed on return
  return FEDD_OK;
}

// Structured node adjacency buffers kept while the device expands them.
struct GraphHost {
  Adjacency a, bt, b;
};

// Greedy first-fit colouring in ascending element order: an element takes the
// smallest colour not used by any element sharing a node with it.  Elements sorted by
// colour (ascending id inside a colour) in perm, colour k at [ptr[k], ptr[k+1]).
int colour_elements(const Mesh& m, std::vector<int64_t>& ptr, std::vector<int32_t>& perm) {
  Incidence inc;
  build_incidence(m, inc);
  std::vector<int32_t> col(m.nelem, -1);
  std::vector<int64_t> used;  // used[colour] == e marks colour as taken for element e
  int nc = 0;
  for (int64_t e = 0; e < m.nelem; ++e) {
    for (int k = 0; k < m.npe; ++k) {
      int32_t n = m.elems[e * m.npe + k];
      for (int64_t q = inc.ptr[n]; q < inc.ptr[n + 1]; ++q) {
        int32_t f = inc.elem[q];
        if (f < e && col[f] >= 0) {
          if ((int)used.size() <= col[f]) used.resize(col[f] + 1, -1);
          used[col[f]] = e;
        }
      }
    }
    int k = 0;
    while (k < (int)used.size() && used[k] == e) ++k;
    col[e] = k;
    nc = std::max(nc, k + 1);
  }
  ptr.assign(nc + 1, 0);
  for (int64_t e = 0; e < m.nelem; ++e) ptr[col[e] + 1]++;
  for (int k = 0; k < nc; ++k) ptr[k + 1] += ptr[k];
  perm.resize(m.nelem);
  std::vector<int64_t> fill(ptr.begin(), ptr.end() - 1);
  for (int64_t e = 0; e < m.nelem; ++e) perm[fill[col[e]]++] = (int32_t)e;
  return nc;
}

// Coloured scatter schedule of the Navier-Stokes graph: per colour, the velocity pairs
// (element, local node) with the vertex class first, and the pressure pairs; the meta
// of each pair is the row gathers' (vs, ps: their plan specifications).
int build_scatter(Ctx* c, const Mesh& m, const Mesh& pm, const PlanSpec& vs, const PlanSpec& ps,
                  const GraphHost& g) {
  ScatterPlan& sp = c->scat;
  sp = ScatterPlan();
  std::vector<int64_t> cptr;
  std::vector<int32_t> cperm;
  const int nc = colour_elements(m, cptr, cperm);
  uint8_t mt[16];
  for (int k = 0; k < nc; ++k) {
    sp.vq.push_back((int64_t)sp.h_velem.size());
    for (int cls = 0; cls < 2; ++cls) {
      if (cls == 1) sp.vsplit.push_back((int64_t)sp.h_velem.size());
      for (int64_t q = cptr[k]; q < cptr[k + 1]; ++q) {
        const int32_t e = cperm[q];
        for (int li = cls ? 4 : 0; li < (cls ? 10 : 4); ++li) {
          const int32_t i = m.elems[(int64_t)e * 10 + li];
          pair_meta_of(vs, li, i, e, mt);
          sp.h_velem.push_back(e);
          sp.h_vnode.push_back(i);
          sp.h_vmeta.insert(sp.h_vmeta.end(), mt, mt + 16);
        }
      }
    }
    sp.pq.push_back((int64_t)sp.h_pelem.size());
    for (int64_t q = cptr[k]; q < cptr[k + 1]; ++q) {
      const int32_t e = cperm[q];
      for (int li = 0; li < 4; ++li) {
        const int32_t pn = pm.elems[(int64_t)e * 4 + li];
        pair_meta_of(ps, li, pn, e, mt);
        sp.h_pelem.push_back(e);
        sp.h_pnode.push_back(pn);
        sp.h_pmeta.insert(sp.h_pmeta.end(), mt, mt + 16);
      }
    }
  }
  sp.vq.push_back((int64_t)sp.h_velem.size());
  sp.pq.push_back((int64_t)sp.h_pelem.size());
  sp.a_ptr = g.a.ptr;
  sp.bt_ptr = g.bt.ptr;
  sp.b_ptr = g.b.ptr;
  int rc;
  if ((rc = upload(c, sp.d_velem, sp.h_velem.data(), sp.h_velem.size() * 4)) ||
      (rc = upload(c, sp.d_vnode, sp.h_vnode.data(), sp.h_vnode.size() * 4)) ||
      (rc = upload(c, sp.d_vmeta, sp.h_vmeta.data(), sp.h_vmeta.size())) ||
      (rc = upload(c, sp.d_pelem, sp.h_pelem.data(), sp.h_pelem.size() * 4)) ||
      (rc = upload(c, sp.d_pnode, sp.h_pnode.data(), sp.h_pnode.size() * 4)) ||
      (rc = upload(c, sp.d_pmeta, sp.h_pmeta.data(), sp.h_pmeta.size())))
    return rc;
  sp.ncolours = nc;
  c->colour_ptr = std::move(cptr);  // the same colouring, for fedd_colour_export
  c->colour_perm = std::move(cperm);
  return FEDD_OK;
}

}  // namespace
}  // namespace fedd

using namespace fedd;

// ----------------------------------------------------------------------------
// Structured unit-cube mesh (MeshStructured::buildMesh3D, one rank, N = 1,
// MeshStructured_def.hpp:622-1009).  Every hex (r,s,t) is split into the same
// six tets sharing the diagonal (r,s,t)-(r+1,s+1,t+1); P2 edge nodes are the
// fine-lattice midpoints of the element edges 4:(0,1) 5:(1,2) 6:(0,2) 7:(0,3)
// 8:(1,3) 9:(2,3).
// ----------------------------------------------------------------------------
static const int kHexTets[6][4][3] = {
    {{1, 0, 0}, {0, 0, 0}, {1, 0, 1}, {1, 1, 1}}, {{0, 0, 1}, {0, 0, 0}, {1, 0, 1}, {1, 1, 1}},
    {{1, 0, 0}, {0, 0, 0}, {1, 1, 0}, {1, 1, 1}}, {{0, 0, 0}, {0, 1, 0}, {1, 1, 0}, {1, 1, 1}},
    {{0, 0, 0}, {0, 1, 0}, {0, 1, 1}, {1, 1, 1}}, {{0, 0, 0}, {0, 0, 1}, {0, 1, 1}, {1, 1, 1}}};
static const int kTetEdges[6][2] = {{0, 1}, {1, 2}, {0, 2}, {0, 3}, {1, 3}, {2, 3}};

extern "C" int fedd_mesh_structured(int M, int fe, double* pts, int32_t* bc, int32_t* elems) {
  if (M < 1 || (fe != 1 && fe != 2) || !pts || !elems) return FEDD_ERR_ARG;
  const int64_t P = fe == 1 ? M + 1 : 2 * M + 1;  // lattice points per direction
  if (P * P * P > INT32_MAX) return FEDD_ERR_LIMIT;
  const double eps = std::numeric_limits<double>::epsilon();
  const double h = 1.0 / M;
  for (int64_t t = 0; t < P; ++t)
    for (int64_t s = 0; s < P; ++s)
      for (int64_t r = 0; r < P; ++r) {
        const int64_t id = r + s * P + t * P * P;
        const int64_t q[3] = {r, s, t};
        int flag = 0;
        for (int d = 0; d < 3; ++d) {
          double x = fe == 1 ? q[d] * h : q[d] * h / 2.0;
          if (x < eps && x > -eps) x = 0.0;
          if (x > 1.0 - eps || x < eps) flag = 1;
          pts[3 * id + d] = x;
        }
        if (bc) bc[id] = flag;
      }
  const int npe = fe == 1 ? 4 : 10;
  const int sc = fe == 1 ? 1 : 2;  // vertex stride in the lattice
  int64_t e = 0;
  for (int64_t t = 0; t < M; ++t)
    for (int64_t s = 0; s < M; ++s)
      for (int64_t r = 0; r < M; ++r)
        for (int k = 0; k < 6; ++k, ++e) {
          int64_t v[4][3];
          for (int a = 0; a < 4; ++a) {
            v[a][0] = sc * (r + kHexTets[k][a][0]);
            v[a][1] = sc * (s + kHexTets[k][a][1]);
            v[a][2] = sc * (t + kHexTets[k][a][2]);
            elems[e * npe + a] = (int32_t)(v[a][0] + v[a][1] * P + v[a][2] * P * P);
          }
          if (fe == 2)
            for (int m = 0; m < 6; ++m) {
              const int64_t* p0 = v[kTetEdges[m][0]];
              const int64_t* p1 = v[kTetEdges[m][1]];
              elems[e * npe + 4 + m] = (int32_t)((p0[0] + p1[0]) / 2 + (p0[1] + p1[1]) / 2 * P +
                                                 (p0[2] + p1[2]) / 2 * P * P);
            }
        }
  return FEDD_OK;
}

// Mesh::correctElementOrientation + flipElement (Mesh_def.hpp:543-578, 648-726):
// elements with det(B) < 0 get local nodes 1 and 2 swapped (and the P2 edge nodes
// permuted accordingly).  Returns the number of flipped elements in *flipped.
extern "C" int fedd_mesh_correct_orientation(const double* pts, int64_t npts, int32_t* elems,
                                             int64_t nelem, int npe, int64_t* flipped) {
  if (!pts || !elems || (npe != 4 && npe != 10)) return FEDD_ERR_ARG;
  int64_t nf = 0;
  for (int64_t e = 0; e < nelem; ++e) {
    int32_t* n = elems + e * npe;
    double a[3][3];
    for (int j = 0; j < 3; ++j)
      for (int i = 0; i < 3; ++i) {
        if (n[j + 1] < 0 || n[j + 1] >= npts || n[0] < 0 || n[0] >= npts) return FEDD_ERR_ARG;
        a[i][j] = pts[3 * (int64_t)n[j + 1] + i] - pts[3 * (int64_t)n[0] + i];
      }
    double det = a[0][0] * a[1][1] * a[2][2] + a[0][1] * a[1][2] * a[2][0] + a[0][2] * a[1][0] * a[2][1] -
                 a[2][0] * a[1][1] * a[0][2] - a[2][1] * a[1][2] * a[0][0] - a[2][2] * a[1][0] * a[0][1];
    if (det < 0) {
      ++nf;
      std::swap(n[1], n[2]);
      if (npe == 10) {
        std::swap(n[4], n[6]);
        std::swap(n[8], n[9]);
      }
    }
  }
  if (flipped) *flipped = nf;
  return FEDD_OK;
}


// LDS budget of one row-gather chunk, in doubles.
#ifndef FEDD_LDS_BUDGET
#define FEDD_LDS_BUDGET 5120
#endif
static const int kLdsBudget = FEDD_LDS_BUDGET;
// The vector (hyperelastic) plan: k_elas_rows is held to two waves per SIMD by its
// registers (two 256-thread workgroups per CU), so its chunks can use the LDS that the
// occupancy leaves: larger chunks, fewer of them.  32^3 x 6 assembly (FP64 / FP32-local)
// with 40 / 48 / 56 / 64 KiB: 1.23-1.24 / 1.18 / 1.13 / 1.21 ms and 1.95 / 1.77 / 1.66 /
// 1.70 ms.
#ifndef FEDD_ELAS_LDS_BUDGET
#define FEDD_ELAS_LDS_BUDGET 7168
#endif
static const int kElasLdsBudget = FEDD_ELAS_LDS_BUDGET;

extern "C" {

int fedd_device_count(int* count) {
  if (!count) return FEDD_ERR_ARG;
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess) n = 0;
  *count = n;
  return FEDD_OK;
}

int fedd_ctx_create(int device, fedd_ctx** out) {
  if (!out) return FEDD_ERR_ARG;
  *out = nullptr;
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess || n == 0) return FEDD_ERR_NODEVICE;
  if (device < 0 || device >= n) return FEDD_ERR_ARG;
  if (hipSetDevice(device) != hipSuccess) return FEDD_ERR_HIP;
  fedd_ctx* c = new fedd_ctx();
  c->device = device;
  {
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, device) == hipSuccess && prop.multiProcessorCount > 0)
      c->num_cu = prop.multiProcessorCount;
  }
  if (hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess ||
      hipStreamCreateWithFlags(&c->side, hipStreamNonBlocking) != hipSuccess ||
      hipEventCreateWithFlags(&c->ev_fork, hipEventDisableTiming) != hipSuccess ||
      hipEventCreateWithFlags(&c->ev_join, hipEventDisableTiming) != hipSuccess) {
    if (c->stream) (void)hipStreamDestroy(c->stream);
    if (c->side) (void)hipStreamDestroy(c->side);
    if (c->ev_fork) (void)hipEventDestroy(c->ev_fork);
    delete c;
    return FEDD_ERR_HIP;
  }
  for (auto& e : c->ev) (void)hipEventCreate(&e);
  noalloc_init(c);
  *out = c;
  return FEDD_OK;
}

void fedd_ctx_destroy(fedd_ctx* c) {
  if (!c) return;
  (void)hipSetDevice(c->device);
  if (c->capturing) {
    hipGraph_t g = nullptr;
    (void)hipStreamEndCapture(c->stream, &g);
    if (g) (void)hipGraphDestroy(g);
  }
  (void)hipStreamSynchronize(c->stream);
  (void)hipStreamSynchronize(c->side);
  if (c->gexec) (void)hipGraphExecDestroy(c->gexec);
  if (c->graph) (void)hipGraphDestroy(c->graph);
  mlp_free(c);
  mg_free(c);
  for (auto& e : c->ev)
    if (e) (void)hipEventDestroy(e);
  for (auto& e : c->ev_rr)
    if (e) (void)hipEventDestroy(e);
  if (c->h_rr) (void)hipHostFree(c->h_rr);
  (void)hipEventDestroy(c->ev_fork);
  (void)hipEventDestroy(c->ev_join);
  (void)hipStreamDestroy(c->side);
  ktimer_resolve(c);
  for (auto e : c->kpool) (void)hipEventDestroy(e);
  (void)hipStreamDestroy(c->stream);
  delete c;
}

const char* fedd_last_error(const fedd_ctx* c) {
  if (c) return c->err.c_str();
  return fedd::free_error().empty() ? "null context" : fedd::free_error().c_str();
}

int fedd_synchronize(fedd_ctx* c) {
  if (!c) return FEDD_ERR_ARG;
  FEDD_HIP(c, hipStreamSynchronize(c->stream));
  return FEDD_OK;
}

int fedd_mesh_upload(fedd_ctx* c, int slot, const double* pts, int64_t npts,
                     const int32_t* elems, int64_t nelem, int npe) {
  if (!c) return FEDD_ERR_ARG;
  if (c->capturing) return fail(c, FEDD_ERR_STATE, "a mesh upload cannot be captured");
  drop_user_graph(c, "the mesh was replaced after the capture: capture again");
  mg_free(c);
  if (slot < 0 || slot > 1) return fail(c, FEDD_ERR_ARG, "mesh slot must be 0 or 1");
  if (!pts || !elems || npts <= 0 || nelem <= 0)
    return fail(c, FEDD_ERR_ARG, "empty mesh");
  if (npe != 4 && npe != 10) return fail(c, FEDD_ERR_ARG, "nodes_per_elem must be 4 (P1) or 10 (P2)");
  if (slot == 1 && npe != 4) return fail(c, FEDD_ERR_ARG, "the pressure mesh must be P1");
  if (npts > INT32_MAX) return fail(c, FEDD_ERR_LIMIT, "more than 2^31-1 nodes");
  for (int64_t k = 0; k < nelem * npe; ++k)
    if (elems[k] < 0 || elems[k] >= npts) return fail(c, FEDD_ERR_ARG, "element node id out of range");
  (void)hipSetDevice(c->device);
  Mesh& m = c->mesh[slot];
  m.npts = npts;
  m.nelem = nelem;
  m.npe = npe;
  m.pts.assign(pts, pts + npts * 3);
  m.elems.assign(elems, elems + nelem * npe);
  std::vector<double> soa(npts * 3);
  for (int64_t i = 0; i < npts; ++i)
    for (int d = 0; d < 3; ++d) soa[d * npts + i] = pts[3 * i + d];
  std::vector<int32_t> conn(nelem * npe);
  for (int64_t e = 0; e < nelem; ++e)
    for (int k = 0; k < npe; ++k) conn[k * nelem + e] = elems[e * npe + k];
  std::vector<int32_t> conn16(nelem * 16, 0);
  for (int64_t e = 0; e < nelem; ++e)
    for (int k = 0; k < npe; ++k) conn16[e * 16 + k] = elems[e * npe + k];
  int rc;
  if ((rc = upload(c, m.d_xyz, soa.data(), soa.size() * 8))) return rc;
  if ((rc = upload(c, m.d_conn, conn.data(), conn.size() * 4))) return rc;
  if ((rc = upload(c, m.d_conn16, conn16.data(), conn16.size() * 4))) return rc;
  FEDD_HIP(c, hipStreamSynchronize(c->stream));
  c->graph_kind = GK_NONE;
  c->colour_ptr.clear();
  c->colour_perm.clear();
  c->newton_step = 0;
  if (slot == 0) {  // the Dirichlet set and the surface triangles belong to the primary mesh
    c->nface = 0;
    c->nbc = 0;
    c->nbc_dofs = 0;
    c->nbc_rows = -1;
    c->bcflag_host.clear();
  }
  return FEDD_OK;
}

int fedd_graph_build(fedd_ctx* c, int kind) {
  if (!c) return FEDD_ERR_ARG;
  if (c->capturing) return fail(c, FEDD_ERR_STATE, "a graph build cannot be captured");
  drop_user_graph(c, "the CRS graph was rebuilt after the capture: capture again");
  mg_free(c);
  if (kind < FEDD_GRAPH_SCALAR || kind > FEDD_GRAPH_NS) return fail(c, FEDD_ERR_ARG, "unknown graph kind");
  Mesh& m = c->mesh[0];
  if (!m.valid()) return fail(c, FEDD_ERR_STATE, "no primary mesh");
  if (kind == FEDD_GRAPH_NS) {
    Mesh& pm = c->mesh[1];
    if (!pm.valid()) return fail(c, FEDD_ERR_STATE, "no pressure mesh");
    if (m.npe != 10) return fail(c, FEDD_ERR_ARG, "Navier-Stokes needs a P2 velocity mesh");
    if (pm.nelem != m.nelem)
      return fail(c, FEDD_ERR_ARG, "Number Elements not the same as number assembleFE elements.");
  }
  (void)hipSetDevice(c->device);
  c->graph_kind = GK_NONE;
  sync_streams(c);  // queued kernels may still read the old blocks
  for (auto& b : c->blk) {
    b.nrows = b.ncols = b.nnz = 0;
    b.d_rowptr.release();
    b.d_colind.release();
    b.d_vals.release();
  }
  c->scat = ScatterPlan();
  Incidence inc;
  build_incidence(m, inc);
  GraphHost g;
  build_adjacency(inc, m.npts, m, g.a);
  int rc;
  const int64_t N = m.npts;
  c->nbc_rows = -1;  // a rebuilt graph needs a new Dirichlet row list
  if (kind == FEDD_GRAPH_SCALAR) {
    c->blk[0].nrows = N;
    c->blk[0].ncols = N;
    c->blk[0].nnz = g.a.ptr[N];
  } else {
    c->blk[0].nrows = 3 * N;
    c->blk[0].ncols = 3 * N;
    c->blk[0].nnz = 9 * g.a.ptr[N];
  }
  Incidence pinc;
  if (kind == FEDD_GRAPH_NS) {
    Mesh& pm = c->mesh[1];
    build_adjacency(inc, N, pm, g.bt);
    build_incidence(pm, pinc);
    build_adjacency(pinc, pm.npts, m, g.b);
    c->blk[1].nrows = 3 * N;
    c->blk[1].ncols = pm.npts;
    c->blk[1].nnz = 3 * g.bt.ptr[N];
    c->blk[2].nrows = pm.npts;
    c->blk[2].ncols = 3 * N;
    c->blk[2].nnz = 3 * g.b.ptr[pm.npts];
    c->blk[3].nrows = pm.npts;
    c->blk[3].ncols = pm.npts;
    c->blk[3].nnz = 0;
  }
  // node-level pointers and adjacency on the device, expanded by a kernel
  if ((rc = upload(c, c->d_nptr, g.a.ptr.data(), g.a.ptr.size() * 8))) return rc;
  DevBuf d_btidx, d_bidx;
  if ((rc = upload(c, c->d_nidx, g.a.idx.data(), g.a.idx.size() * 4))) return rc;
  if (kind == FEDD_GRAPH_NS) {
    if ((rc = upload(c, c->d_pptr, g.bt.ptr.data(), g.bt.ptr.size() * 8))) return rc;
    if ((rc = upload(c, c->d_bptr, g.b.ptr.data(), g.b.ptr.size() * 8))) return rc;
    if ((rc = upload(c, d_btidx, g.bt.idx.data(), g.bt.idx.size() * 4))) return rc;
    if ((rc = upload(c, d_bidx, g.b.idx.data(), g.b.idx.size() * 4))) return rc;
  }
  int nb = kind == FEDD_GRAPH_NS ? 4 : 1;
  for (int k = 0; k < nb; ++k) {
    Block& b = c->blk[k];
    if ((rc = alloc(c, b.d_rowptr, (b.nrows + 1) * 8))) return rc;
    if ((rc = alloc(c, b.d_colind, std::max<int64_t>(b.nnz, 1) * 4))) return rc;
    if ((rc = alloc(c, b.d_vals, std::max<int64_t>(b.nnz, 1) * 8))) return rc;
    FEDD_HIP(c, hipMemsetAsync(b.d_vals.p, 0, std::max<int64_t>(b.nnz, 1) * 8, c->stream));
  }
  c->graph_kind = kind;
  rc = expand_graph(c, kind, c->d_nidx.as<int32_t>(), d_btidx.as<int32_t>(), d_bidx.as<int32_t>());
  if (rc) {
    c->graph_kind = GK_NONE;
    return rc;
  }
  // gather schedules
  PlanSpec vs{};
  vs.main = &g.a;
  vs.cols_main = &m;
  vs.lds_budget_doubles = kind == FEDD_GRAPH_VECTOR ? kElasLdsBudget : kLdsBudget;
  if (kind == FEDD_GRAPH_SCALAR) {
    vs.rows_per_node = 1;
    vs.comps_per_col = 1;
  } else {
    vs.rows_per_node = 3;
    vs.comps_per_col = 3;
  }
  // segments when reproducible sums are asked for (fedd_set_reproducible); the
  // hyperelastic row gather keeps the rank-interleaved order
  vs.segments = c->reproducible && kind != FEDD_GRAPH_VECTOR;
  if (kind == FEDD_GRAPH_NS) {
    // the velocity rows' chunks are sized for k_ns_vel_rows (its LDS table and occupancy)
    vs.lds_budget_doubles = kNsLdsBudget;
    vs.max_nodes = kNsChunkNodes;
    vs.second = &g.bt;
    vs.second_rows = 3;
    vs.cols_second = &c->mesh[1];
  }
  if ((rc = build_plan(c, inc, N, m, vs, c->vplan))) {
    c->graph_kind = GK_NONE;
    return rc;
  }
  if (kind == FEDD_GRAPH_NS) {
    PlanSpec ps{};
    ps.rows_per_node = 1;
    ps.comps_per_col = 3;
    ps.main = &g.b;
    ps.cols_main = &m;
    ps.lds_budget_doubles = kLdsBudget;
    ps.segments = c->reproducible;
    if ((rc = build_plan(c, pinc, c->mesh[1].npts, c->mesh[1], ps, c->pplan))) {
      c->graph_kind = GK_NONE;
      return rc;
    }
    if (c->colour_scatter && (rc = build_scatter(c, m, c->mesh[1], vs, ps, g))) {
      c->graph_kind = GK_NONE;
      return rc;
    }
  }
  FEDD_HIP(c, hipStreamSynchronize(c->stream));
  c->newton_step = 0;
  return FEDD_OK;
}

int fedd_set_reproducible(fedd_ctx* c, int on) {
  if (!c || (on != 0 && on != 1)) return FEDD_ERR_ARG;
  c->reproducible = on != 0;
  return FEDD_OK;
}

int fedd_set_assembly_method(fedd_ctx* c, int method) {
  if (!c || (method != FEDD_METHOD_ROW_GATHER && method != FEDD_METHOD_COLOUR_SCATTER)) return FEDD_ERR_ARG;
  c->colour_scatter = method == FEDD_METHOD_COLOUR_SCATTER;
  return FEDD_OK;
}

int fedd_scatter_indices(fedd_ctx* c, int block, int64_t* idx) {
  if (!c || !idx || block < 0 || block > 2) return FEDD_ERR_ARG;
  const ScatterPlan& sp = c->scat;
  if (c->graph_kind != FEDD_GRAPH_NS || sp.ncolours == 0)
    return fail(c, FEDD_ERR_STATE, "no coloured scatter schedule (set the method before building the NS graph)");
  if (block < 2) {
    for (size_t q = 0; q < sp.h_velem.size(); ++q) {
      const int64_t e = sp.h_velem[q], i = sp.h_vnode[q];
      const uint8_t* mt = sp.h_vmeta.data() + 16 * q;
      const int li = mt[0];
      if (block == 0)
        for (int cpos = 0; cpos < 10; ++cpos) idx[100 * e + 10 * li + kPermP2[li][cpos]] = sp.a_ptr[i] + mt[2 + cpos];
      else
        for (int cpos = 0; cpos < 4; ++cpos) idx[40 * e + 4 * li + kPermP2[li][cpos]] = sp.bt_ptr[i] + mt[12 + cpos];
    }
  } else {
    for (size_t q = 0; q < sp.h_pelem.size(); ++q) {
      const int64_t e = sp.h_pelem[q], pn = sp.h_pnode[q];
      const uint8_t* mt = sp.h_pmeta.data() + 16 * q;
      const int li = mt[0];
      for (int cpos = 0; cpos < 10; ++cpos) idx[40 * e + 10 * li + kPermP2[li][cpos]] = sp.b_ptr[pn] + mt[2 + cpos];
    }
  }
  return FEDD_OK;
}

int fedd_block_info(fedd_ctx* c, int block, int64_t* nrows, int64_t* ncols, int64_t* nnz) {
  if (!c || block < 0 || block > 3) return FEDD_ERR_ARG;
  if (c->graph_kind == GK_NONE) return fail(c, FEDD_ERR_STATE, "no graph");
  const Block& b = c->blk[block];
  if (nrows) *nrows = b.nrows;
  if (ncols) *ncols = b.ncols;
  if (nnz) *nnz = b.nnz;
  return FEDD_OK;
}

int fedd_block_graph(fedd_ctx* c, int block, int64_t* rowptr, int32_t* colind) {
  if (!c || block < 0 || block > 3) return FEDD_ERR_ARG;
  if (c->graph_kind == GK_NONE) return fail(c, FEDD_ERR_STATE, "no graph");
  const Block& b = c->blk[block];
  if (b.nrows == 0) return fail(c, FEDD_ERR_ARG, "block not present in this graph");
  (void)hipSetDevice(c->device);
  if (rowptr)
    FEDD_HIP(c, hipMemcpyAsync(rowptr, b.d_rowptr.p, (b.nrows + 1) * 8, hipMemcpyDeviceToHost, c->stream));
  if (colind && b.nnz)
    FEDD_HIP(c, hipMemcpyAsync(colind, b.d_colind.p, b.nnz * 4, hipMemcpyDeviceToHost, c->stream));
  FEDD_HIP(c, hipStreamSynchronize(c->stream));
  return FEDD_OK;
}

int fedd_block_values(fedd_ctx* c, int block, double* values) {
  if (!c || block < 0 || block > 3 || !values) return FEDD_ERR_ARG;
  if (c->graph_kind == GK_NONE) return fail(c, FEDD_ERR_STATE, "no graph");
  const Block& b = c->blk[block];
  (void)hipSetDevice(c->device);
  if (b.nnz)
    FEDD_HIP(c, hipMemcpyAsync(values, b.d_vals.p, b.nnz * 8, hipMemcpyDeviceToHost, c->stream));
  FEDD_HIP(c, hipStreamSynchronize(c->stream));
  return FEDD_OK;
}

int fedd_block_device_values(fedd_ctx* c, int block, double** d) {
  if (!c || block < 0 || block > 3 || !d) return FEDD_ERR_ARG;
  if (c->graph_kind == GK_NONE) return fail(c, FEDD_ERR_STATE, "no graph");
  *d = c->blk[block].d_vals.as<double>();
  return FEDD_OK;
}

// ------------------------------------------------------ export formats
namespace {
struct HostCrs {
  int64_t nrows = 0, ncols = 0;
  std::vector<int64_t> rp;
  std::vector<int32_t> ci;
  std::vector<double> v;
};

int host_block(fedd::Ctx* c, int block, HostCrs& h) {
  const fedd::Block& b = c->blk[block];
  h.nrows = b.nrows;
  h.ncols = b.ncols;
  h.rp.resize(b.nrows + 1);
  h.ci.resize(b.nnz);
  h.v.resize(b.nnz);
  FEDD_HIP(c, hipMemcpyAsync(h.rp.data(), b.d_rowptr.p, (b.nrows + 1) * 8, hipMemcpyDeviceToHost, c->stream));
  if (b.nnz) {
    FEDD_HIP(c, hipMemcpyAsync(h.ci.data(), b.d_colind.p, b.nnz * 4, hipMemcpyDeviceToHost, c->stream));
    FEDD_HIP(c, hipMemcpyAsync(h.v.data(), b.d_vals.p, b.nnz * 8, hipMemcpyDeviceToHost, c->stream));
  }
  FEDD_HIP(c, hipStreamSynchronize(c->stream));
  return FEDD_OK;
}

// The 2x2 NS block matrix as one CRS matrix (BlockMatrix::merge): block (i, j) at row
// offset 3 N2 i and column offset 3 N2 j; on one rank the merged rows are the block
// rows side by side, already sorted.
int merged_ns(fedd::Ctx* c, HostCrs& M) {
  HostCrs b[4];
  for (int k = 0; k < 4; ++k) {
    int rc = host_block(c, k, b[k]);
    if (rc) return rc;
  }
  const int64_t nv = b[0].nrows, np = b[2].nrows;
  M.nrows = M.ncols = nv + np;
  M.rp.assign(M.nrows + 1, 0);
  M.ci.clear();
  M.v.clear();
  M.ci.reserve(b[0].ci.size() + b[1].ci.size() + b[2].ci.size() + b[3].ci.size());
  M.v.reserve(M.ci.capacity());
  auto append = [&](const HostCrs& B, int64_t r, int64_t coff) {
    for (int64_t k = B.rp[r]; k < B.rp[r + 1]; ++k) {
      M.ci.push_back((int32_t)(B.ci[k] + coff));
      M.v.push_back(B.v[k]);
    }
  };
  for (int64_t r = 0; r < nv; ++r) {
    append(b[0], r, 0);
    append(b[1], r, nv);
    M.rp[r + 1] = (int64_t)M.ci.size();
  }
  for (int64_t r = 0; r < np; ++r) {
    append(b[2], r, 0);
    append(b[3], r, nv);
    M.rp[nv + r + 1] = (int64_t)M.ci.size();
  }
  return FEDD_OK;
}

// MatrixMarket coordinate file (the layout of Tpetra::MatrixMarket::Writer::writeSparseFile
// behind Matrix::writeMM, Matrix_def.hpp:264-277): 1-based indices, explicit zeros kept,
// values with 17 significant digits (exact round trip).
int write_mm(fedd::Ctx* c, const HostCrs& M, const char* path, const char* name) {
  FILE* f = std::fopen(path, "w");
  if (!f) return fedd::fail(c, FEDD_ERR_ARG, std::string("cannot open ") + path);
  std::fprintf(f, "%%%%MatrixMarket matrix coordinate real general\n%% %s\n", name);
  std::fprintf(f, "%lld %lld %lld\n", (long long)M.nrows, (long long)M.ncols, (long long)M.ci.size());
  for (int64_t r = 0; r < M.nrows; ++r)
    for (int64_t k = M.rp[r]; k < M.rp[r + 1]; ++k)
      std::fprintf(f, "%lld %d %.17e\n", (long long)(r + 1), M.ci[k] + 1, M.v[k]);
  const bool ok = std::fclose(f) == 0;
  return ok ? FEDD_OK : fedd::fail(c, FEDD_ERR_ARG, std::string("write failed: ") + path);
}
}  // namespace

int fedd_ns_merged_info(fedd_ctx* c, int64_t* nrows, int64_t* nnz) {
  if (!c) return FEDD_ERR_ARG;
  if (c->graph_kind != FEDD_GRAPH_NS) return fail(c, FEDD_ERR_STATE, "needs a Navier-Stokes graph");
  if (nrows) *nrows = c->blk[0].nrows + c->blk[2].nrows;
  if (nnz) *nnz = c->blk[0].nnz + c->blk[1].nnz + c->blk[2].nnz + c->blk[3].nnz;
  return FEDD_OK;
}

int fedd_ns_merged_export(fedd_ctx* c, int64_t* rowptr, int32_t* colind, double* values) {
  if (!c) return FEDD_ERR_ARG;
  if (c->graph_kind != FEDD_GRAPH_NS) return fail(c, FEDD_ERR_STATE, "needs a Navier-Stokes graph");
  (void)hipSetDevice(c->device);
  HostCrs M;
  int rc = merged_ns(c, M);
  if (rc) return rc;
  if (rowptr) std::copy(M.rp.begin(), M.rp.end(), rowptr);
  if (colind) std::copy(M.ci.begin(), M.ci.end(), colind);
  if (values) std::copy(M.v.begin(), M.v.end(), values);
  return FEDD_OK;
}

int fedd_write_matrix_market(fedd_ctx* c, int block, const char* path) {
  if (!c || !path || block < -1 || block > 3) return FEDD_ERR_ARG;
  if (c->graph_kind == GK_NONE) return fail(c, FEDD_ERR_STATE, "no graph");
  (void)hipSetDevice(c->device);
  HostCrs M;
  int rc;
  if (block < 0) {
    if (c->graph_kind != FEDD_GRAPH_NS) return fail(c, FEDD_ERR_STATE, "the merged export needs a Navier-Stokes graph");
    rc = merged_ns(c, M);
  } else {
    rc = host_block(c, block, M);
  }
  if (rc) return rc;
  return write_mm(c, M, path, "matrix");
}

int fedd_colour_build(fedd_ctx* c, int* ncolours) {
  if (!c) return FEDD_ERR_ARG;
  Mesh& m = c->mesh[0];
  if (!m.valid()) return fail(c, FEDD_ERR_STATE, "no primary mesh");
  const int nc = colour_elements(m, c->colour_ptr, c->colour_perm);
  if (ncolours) *ncolours = nc;
  return FEDD_OK;
}

int fedd_colour_export(fedd_ctx* c, int64_t* colour_ptr, int32_t* elem_perm) {
  if (!c) return FEDD_ERR_ARG;
  if (c->colour_ptr.empty()) return fail(c, FEDD_ERR_STATE, "no colouring");
  if (colour_ptr) std::copy(c->colour_ptr.begin(), c->colour_ptr.end(), colour_ptr);
  if (elem_perm) std::copy(c->colour_perm.begin(), c->colour_perm.end(), elem_perm);
  return FEDD_OK;
}

static int64_t vec_len(fedd_ctx* c, int which) {
  switch (which) {
    case FEDD_VEC_U:
    case FEDD_VEC_RES_U: return 3 * c->mesh[0].npts;
    case FEDD_VEC_P:
    case FEDD_VEC_RES_P:
    case FEDD_VEC_X_P:
    case FEDD_VEC_B_P: return c->mesh[1].npts;
    case FEDD_VEC_RHS: return c->mesh[0].npts;
    case FEDD_VEC_FEXT:
    case FEDD_VEC_FPRES:
    case FEDD_VEC_X:
    case FEDD_VEC_B: return 3 * c->mesh[0].npts;
  }
  return -1;
}

int fedd_vector_upload(fedd_ctx* c, int which, const double* host, int64_t n) {
  if (!c || which < 0 || which >= FEDD_NUM_VECTORS || !host) return FEDD_ERR_ARG;
  if (c->capturing) return fail(c, FEDD_ERR_STATE, "host transfers cannot be captured");
  int64_t want = vec_len(c, which);
  if (n != want) return fail(c, FEDD_ERR_ARG, "vector length does not match the mesh");
  (void)hipSetDevice(c->device);
  int rc = upload(c, c->vec[which], host, n * 8);
  if (rc) return rc;
  c->vec_n[which] = n;
  FEDD_HIP(c, hipStreamSynchronize(c->stream));
  return FEDD_OK;
}

// ------------------------------------------------------------ graph capture
int fedd_capture_begin(fedd_ctx* c) {
  if (!c) return FEDD_ERR_ARG;
  if (c->capturing) return fail(c, FEDD_ERR_STATE, "a capture is already open");
  (void)hipSetDevice(c->device);
  FEDD_HIP(c, hipStreamBeginCapture(c->stream, hipStreamCaptureModeRelaxed));
  c->capturing = true;
  return FEDD_OK;
}

int fedd_capture_end(fedd_ctx* c, int64_t* nodes) {
  if (!c) return FEDD_ERR_ARG;
  if (!c->capturing) return fail(c, FEDD_ERR_STATE, "no capture is open");
  (void)hipSetDevice(c->device);
  hipGraph_t g = nullptr;
  const hipError_t e = hipStreamEndCapture(c->stream, &g);
  c->capturing = false;
  if (e != hipSuccess || !g) {
    if (g) (void)hipGraphDestroy(g);
    return fail(c, FEDD_ERR_HIP, std::string("hipStreamEndCapture: ") + hipGetErrorString(e));
  }
  if (c->gexec || c->graph) {
    sync_streams(c);  // the previous graph may still be replaying
    if (c->gexec) (void)hipGraphExecDestroy(c->gexec);
    if (c->graph) (void)hipGraphDestroy(c->graph);
  }
  c->gexec = nullptr;
  c->graph = g;
  c->graph_dropped.clear();
  FEDD_HIP(c, hipGraphInstantiate(&c->gexec, g, nullptr, nullptr, 0));
  if (nodes) {
    size_t cnt = 0;
    FEDD_HIP(c, hipGraphGetNodes(g, nullptr, &cnt));
    *nodes = (int64_t)cnt;
  }
  return FEDD_OK;
}

int fedd_graph_launch(fedd_ctx* c) {
  if (!c) return FEDD_ERR_ARG;
  if (!c->gexec)
    return fail(c, FEDD_ERR_STATE, c->graph_dropped.empty() ? std::string("no captured graph") : c->graph_dropped);
  if (c->capturing) return fail(c, FEDD_ERR_STATE, "a capture is open");
  (void)hipSetDevice(c->device);
  FEDD_HIP(c, hipGraphLaunch(c->gexec, c->stream));
  return FEDD_OK;
}

int fedd_vector_download(fedd_ctx* c, int which, double* host, int64_t n) {
  if (!c || which < 0 || which >= FEDD_NUM_VECTORS || !host) return FEDD_ERR_ARG;
  if (c->vec_n[which] != n) return fail(c, FEDD_ERR_ARG, "vector length mismatch or vector absent");
  (void)hipSetDevice(c->device);
  FEDD_HIP(c, hipMemcpyAsync(host, c->vec[which].p, n * 8, hipMemcpyDeviceToHost, c->stream));
  FEDD_HIP(c, hipStreamSynchronize(c->stream));
  return FEDD_OK;
}

int fedd_vector_device_ptr(fedd_ctx* c, int which, double** d, int64_t* n) {
  if (!c || which < 0 || which >= FEDD_NUM_VECTORS || !d) return FEDD_ERR_ARG;
  *d = c->vec[which].as<double>();
  if (n) *n = c->vec_n[which];
  return FEDD_OK;
}

int fedd_assembly_laplace(fedd_ctx* c, int precision) {
  if (!c) return FEDD_ERR_ARG;
  if (c->graph_kind != FEDD_GRAPH_SCALAR) return fail(c, FEDD_ERR_STATE, "needs a scalar graph");
  if (precision != FEDD_FP64 && precision != FEDD_FP32_LOCAL) return fail(c, FEDD_ERR_ARG, "bad precision");
  (void)hipSetDevice(c->device);
  return launch_laplace(c, precision);
}

int fedd_assembly_rhs(fedd_ctx* c, double f) {
  if (!c) return FEDD_ERR_ARG;
  if (c->graph_kind != FEDD_GRAPH_SCALAR) return fail(c, FEDD_ERR_STATE, "needs a scalar graph");
  (void)hipSetDevice(c->device);
  int64_t n = c->mesh[0].npts;
  if (c->vec_n[FEDD_VEC_RHS] != n) {
    int rc = alloc(c, c->vec[FEDD_VEC_RHS], n * 8);
    if (rc) return rc;
    c->vec_n[FEDD_VEC_RHS] = n;
  }
  return launch_rhs(c, f);
}

int fedd_rhs_quadrature_points(fedd_ctx* c, double* xq, int64_t cap, int* nq) {
  if (!c || !nq) return FEDD_ERR_ARG;
  if (c->graph_kind != FEDD_GRAPH_SCALAR) return fail(c, FEDD_ERR_STATE, "needs a scalar graph");
  const Mesh& m = c->mesh[0];
  *nq = m.npe == 4 ? 1 : 5;
  if (!xq) return FEDD_OK;
  if (c->capturing) return fail(c, FEDD_ERR_STATE, "host transfers cannot be captured");
  const int64_t n = m.nelem * *nq * 3;
  if (cap < n) return fail(c, FEDD_ERR_ARG, "xq holds fewer than nelem * nq * 3 doubles");
  (void)hipSetDevice(c->device);
  if (c->d_rhsq.bytes < (size_t)n * 8) {
    int rc = alloc(c, c->d_rhsq, (size_t)n * 8);
    if (rc) return rc;
  }
  int rc = launch_rhs_qpoints(c, c->d_rhsq.as<double>(), nq);
  if (rc) return rc;
  FEDD_HIP(c, hipMemcpyAsync(xq, c->d_rhsq.p, n * 8, hipMemcpyDeviceToHost, c->stream));
  FEDD_HIP(c, hipStreamSynchronize(c->stream));
  return FEDD_OK;
}

int fedd_assembly_rhs_points(fedd_ctx* c, const double* fq, int64_t n, int ncomp) {
  if (!c || !fq) return FEDD_ERR_ARG;
  if (c->graph_kind != FEDD_GRAPH_SCALAR) return fail(c, FEDD_ERR_STATE, "needs a scalar graph");
  if (ncomp != 1 && ncomp != 3) return fail(c, FEDD_ERR_ARG, "Invalid field type.");
  if (c->capturing) return fail(c, FEDD_ERR_STATE, "host transfers cannot be captured");
  const Mesh& m = c->mesh[0];
  const int nq = m.npe == 4 ? 1 : 5;
  if (n != m.nelem * nq * ncomp)
    return fail(c, FEDD_ERR_ARG, "source values: need nelem * nq * ncomp entries");
  (void)hipSetDevice(c->device);
  int rc = upload(c, c->d_rhsq, fq, (size_t)n * 8);
  if (rc) return rc;
  const int64_t len = (int64_t)ncomp * m.npts;
  if (c->vec_n[FEDD_VEC_RHS] != len) {
    if ((rc = alloc(c, c->vec[FEDD_VEC_RHS], len * 8))) return rc;
    c->vec_n[FEDD_VEC_RHS] = len;
  }
  return launch_rhs_points(c, c->d_rhsq.as<double>(), ncomp);
}

int fedd_assembly_navier_stokes(fedd_ctx* c, const fedd_ns_params* p, int mode, int reassemble,
                                int precision) {
  if (!c || !p) return FEDD_ERR_ARG;
  if (c->graph_kind != FEDD_GRAPH_NS) return fail(c, FEDD_ERR_STATE, "needs a Navier-Stokes graph");
  if (mode < FEDD_NS_JACOBIAN || mode > FEDD_NS_FIXEDPOINT_RHS) return fail(c, FEDD_ERR_ARG, "unknown assembleMode");
  if (precision != FEDD_FP64 && precision != FEDD_FP32_LOCAL) return fail(c, FEDD_ERR_ARG, "bad precision");
  if (c->vec_n[FEDD_VEC_U] != 3 * c->mesh[0].npts) return fail(c, FEDD_ERR_STATE, "velocity iterate not uploaded");
  if (c->vec_n[FEDD_VEC_P] != c->mesh[1].npts) return fail(c, FEDD_ERR_STATE, "pressure iterate not uploaded");
  (void)hipSetDevice(c->device);
  if (mode == FEDD_NS_RHS || mode == FEDD_NS_JACOBIAN_RHS || mode == FEDD_NS_FIXEDPOINT_RHS) {
    for (int w : {FEDD_VEC_RES_U, FEDD_VEC_RES_P}) {
      int64_t n = vec_len(c, w);
      if (c->vec_n[w] != n) {
        int rc = alloc(c, c->vec[w], n * 8);
        if (rc) return rc;
        c->vec_n[w] = n;
      }
    }
  }
  int rc = launch_ns_assembly(c, p, mode, reassemble, precision);
  if (rc) return rc;
  if (mode != FEDD_NS_RHS) c->newton_step++;  // advanceNewtonStep (FE_def.hpp:1171,1196)
  return FEDD_OK;
}

static int elasticity_common(fedd_ctx* c, int model, double E, double nu, int precision, int with_res,
                             const double* Evec = nullptr, int tangent = 1) {
  if (!c) return FEDD_ERR_ARG;
  if (c->graph_kind != FEDD_GRAPH_VECTOR) return fail(c, FEDD_ERR_STATE, "needs a vector (P2) graph");
  if (c->mesh[0].npe != 10) return fail(c, FEDD_ERR_STATE, "the hyperelastic elements need a P2 (tet-10) mesh");
  if (precision != FEDD_FP64 && precision != FEDD_FP32_LOCAL) return fail(c, FEDD_ERR_ARG, "bad precision");
  if (!(E > 0) || !(nu > -1.0 && nu < 0.5)) return fail(c, FEDD_ERR_ARG, "E must be > 0 and -1 < nu < 0.5");
  if (c->vec_n[FEDD_VEC_U] != 3 * c->mesh[0].npts) return fail(c, FEDD_ERR_STATE, "displacement not uploaded");
  (void)hipSetDevice(c->device);
  if (with_res) {
    int64_t n = 3 * c->mesh[0].npts;
    if (c->vec_n[FEDD_VEC_RES_U] != n) {
      int rc = alloc(c, c->vec[FEDD_VEC_RES_U], n * 8);
      if (rc) return rc;
      c->vec_n[FEDD_VEC_RES_U] = n;
    }
  }
  return launch_elasticity(c, model, E, nu, precision, with_res, tangent, Evec);
}

int fedd_assembly_linear_elasticity(fedd_ctx* c, double E, double nu, int precision) {
  return elasticity_common(c, 0, E, nu, precision, 0);
}

int fedd_assembly_linear_elasticity_xdim(fedd_ctx* c, double lambda, double mu, int precision) {
  if (!c) return FEDD_ERR_ARG;
  if (c->graph_kind != FEDD_GRAPH_VECTOR) return fail(c, FEDD_ERR_STATE, "needs a vector graph");
  if (precision != FEDD_FP64 && precision != FEDD_FP32_LOCAL) return fail(c, FEDD_ERR_ARG, "bad precision");
  if (!(mu > 0) || !(lambda > -2.0 / 3.0 * mu)) return fail(c, FEDD_ERR_ARG, "need mu > 0 and lambda > -2 mu / 3");
  (void)hipSetDevice(c->device);
  return launch_linelas(c, lambda, mu, precision);
}

int fedd_assembly_nonlinear_elasticity(fedd_ctx* c, int neo_hooke_model, double E, double nu, int precision) {
  // "Neo-Hooke Modell" == 1 -> NonLinearElasticity (skr2), otherwise NonLinearElasticity2 (skr3)
  return elasticity_common(c, neo_hooke_model == 1 ? 1 : 2, E, nu, precision, 1);
}

int fedd_assembly_nonlinear_elasticity_efield(fedd_ctx* c, int neo_hooke_model, const double* E_elem, int64_t nelem,
                                              double nu, int precision) {
  if (!c || !E_elem) return FEDD_ERR_ARG;
  if (nelem != c->mesh[0].nelem) return fail(c, FEDD_ERR_ARG, "Number Elements not the same as the E field length");
  for (int64_t e = 0; e < nelem; ++e)
    if (!(E_elem[e] > 0)) return fail(c, FEDD_ERR_ARG, "E must be > 0 in every element");
  (void)hipSetDevice(c->device);
  int rc = upload(c, c->d_emod, E_elem, nelem * 8);
  if (rc) return rc;
  return elasticity_common(c, neo_hooke_model == 1 ? 1 : 2, E_elem[0], nu, precision, 1, c->d_emod.as<double>());
}

int fedd_assembly_nonlinear_elasticity_tangent(fedd_ctx* c, int neo_hooke_model, double E, double nu, int precision,
                                                int tangent) {
  if (!c) return FEDD_ERR_ARG;
  if (tangent != 0 && tangent != 1) return fail(c, FEDD_ERR_ARG, "tangent must be 0 (exact) or 1 (learned)");
  const int model = neo_hooke_model == 1 ? 1 : 2;
  if (tangent == 1) {
    if (!c->mlp_ready) return fail(c, FEDD_ERR_STATE, "learned tangent requested but no network");
    if (!mlp_matches(c, model, E, nu))
      return fail(c, FEDD_ERR_STATE, "the surrogate was built for another law or other material constants");
  }
  return elasticity_common(c, model, E, nu, precision, 1, nullptr, tangent == 1 ? 2 : 1);
}

int fedd_elasticity_fp64_fixups(fedd_ctx* c, int64_t* nchunks, int64_t* npairs) {
  if (!c || !nchunks) return FEDD_ERR_ARG;
  *nchunks = 0;
  if (npairs) *npairs = 0;
  if (!c->elas_fix_valid) return FEDD_OK;
  if (c->capturing) return fail(c, FEDD_ERR_STATE, "the query synchronises: it cannot be captured");
  (void)hipSetDevice(c->device);
  int32_t h[2] = {0, 0};
  FEDD_HIP(c, hipMemcpyAsync(&h[0], c->d_elas_fix.as<int32_t>() + c->vplan.nchunks, 4, hipMemcpyDeviceToHost,
                             c->stream));
  FEDD_HIP(c, hipMemcpyAsync(&h[1], c->d_elas_flag.as<int32_t>() + c->vplan.nchunks + 1, 4, hipMemcpyDeviceToHost,
                             c->stream));
  FEDD_HIP(c, hipStreamSynchronize(c->stream));
  *nchunks = h[0];
  if (npairs) *npairs = h[1];
  return FEDD_OK;
}

int fedd_newton_step(fedd_ctx* c, int* step) {
  if (!c || !step) return FEDD_ERR_ARG;
  *step = c->newton_step;
  return FEDD_OK;
}

int fedd_stream_copy(fedd_ctx* c, int64_t bytes, int iters, float* ms_per_copy) {
  if (!c || bytes < 16 || iters < 1) return FEDD_ERR_ARG;
  (void)hipSetDevice(c->device);
  return launch_stream_copy(c, bytes, iters, ms_per_copy);
}

int fedd_event_record(fedd_ctx* c, int slot) {
  if (!c || slot < 0 || slot >= 16) return FEDD_ERR_ARG;
  FEDD_HIP(c, hipEventRecord(c->ev[slot], c->stream));
  return FEDD_OK;
}

int fedd_event_elapsed(fedd_ctx* c, int a, int b, float* ms) {
  if (!c || a < 0 || a >= 16 || b < 0 || b >= 16 || !ms) return FEDD_ERR_ARG;
  FEDD_HIP(c, hipEventSynchronize(c->ev[b]));
  FEDD_HIP(c, hipEventElapsedTime(ms, c->ev[a], c->ev[b]));
  return FEDD_OK;
}

int fedd_kernel_timing(fedd_ctx* c, int enable) {
  if (!c) return FEDD_ERR_ARG;
  ktimer_resolve(c);
  c->timing = enable != 0;
  c->ktime.clear();
  return FEDD_OK;
}

int fedd_kernel_time(fedd_ctx* c, const char* name, double* total_ms, int64_t* launches) {
  if (!c || !name) return FEDD_ERR_ARG;
  ktimer_resolve(c);
  auto it = c->ktime.find(name);
  if (total_ms) *total_ms = it == c->ktime.end() ? 0.0 : it->second.first;
  if (launches) *launches = it == c->ktime.end() ? 0 : it->second.second;
  return FEDD_OK;
}

}  // extern "C"
