/*
 * feddgpu.h -- C ABI of the MI355X (gfx950) element-assembly engine.
 *
 * Drop-in for FEDDLib's element-assembly path: the FE::assembly* loops
 * (feddlib/core/FE/FE_decl.hpp:181-186, 453-459, 475-498, 560-595) with their
 * AssembleFE element plugins (feddlib/core/AceFemAssembly/AssembleFE_decl.hpp:61-299),
 * writing into CRS matrices whose graph is identical to the Tpetra graph the
 * reference builds (insertGlobalValues + fillComplete,
 * feddlib/core/LinearAlgebra/Matrix_def.hpp:88-92, 192-199).
 *
 * Conventions
 *   - Plain pointers and sizes only.  Host arrays are caller-owned and only read
 *     during the call.  Device buffers are owned by the context.
 *   - Node-major coordinates (x0 y0 z0 x1 ...), element-major connectivity in
 *     FEDDLib local order; node ids are the one-rank GIDs (repeated map = unique
 *     map), as MeshStructured::buildMesh3D produces them
 *     (feddlib/core/Mesh/MeshStructured_def.hpp:703-994).
 *   - Dof numbering dof = 3*node + d (Map::buildVecFieldMap,
 *     feddlib/core/LinearAlgebra/Map_def.hpp:95-107).
 *   - Every entry point returns FEDD_OK (0) or a negative error class; the
 *     message is in fedd_last_error().  No exception crosses the ABI.  The
 *     reference raises std::logic_error / std::runtime_error through
 *     TEUCHOS_TEST_FOR_EXCEPTION for the same conditions.
 *   - One context per device and host thread; calls are ordered on the
 *     context's HIP stream; a context is not thread-safe.
 */
#ifndef FEDDGPU_H
#define FEDDGPU_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct fedd_ctx fedd_ctx;

enum fedd_status {
  FEDD_OK = 0,
  FEDD_ERR_ARG = -1,      /* invalid argument (std::logic_error in the reference) */
  FEDD_ERR_HIP = -2,      /* HIP runtime failure */
  FEDD_ERR_STATE = -3,    /* call order violated (e.g. assembly before graph) */
  FEDD_ERR_LIMIT = -4,    /* a documented size limit is exceeded */
  FEDD_ERR_NODEVICE = -5  /* no HIP device */
};

enum fedd_mesh_slot { FEDD_MESH_PRIMARY = 0, FEDD_MESH_PRESSURE = 1 };

enum fedd_graph_kind {
  FEDD_GRAPH_SCALAR = 0,  /* one dof per node: FE::assemblyLaplace */
  FEDD_GRAPH_VECTOR = 1,  /* three dofs per node, full 3x3 node blocks: addFeBlock */
  FEDD_GRAPH_NS = 2       /* P2-P1: A(0,0), B^T(0,1), B(1,0), C(1,1): addFeBlockMatrix */
};

enum fedd_block {
  FEDD_BLOCK_A = 0,   /* scalar Laplace, vector / elasticity, or NS A(0,0) */
  FEDD_BLOCK_BT = 1,  /* NS (0,1) */
  FEDD_BLOCK_B = 2,   /* NS (1,0) */
  FEDD_BLOCK_C = 3    /* NS (1,1): empty graph for P2-P1 (|v| > 1e-13 rule) */
};

enum fedd_vector {
  FEDD_VEC_U = 0,      /* velocity / displacement iterate, 3 per primary node */
  FEDD_VEC_P = 1,      /* pressure iterate, 1 per pressure node */
  FEDD_VEC_RES_U = 2,  /* residual, velocity / displacement block */
  FEDD_VEC_RES_P = 3,  /* residual, pressure block */
  FEDD_VEC_RHS = 4,    /* load vector of FE::assemblyRHS (1 per primary node) */
  FEDD_VEC_FEXT = 5,   /* external load of the elasticity problem, 3 per node */
  FEDD_VEC_X = 6,      /* linear-solve solution / Newton increment, 3 per node */
  FEDD_VEC_B = 7,      /* linear-solve right-hand side, 3 per node */
  FEDD_VEC_X_P = 8,    /* pressure part of the NS linear-solve solution, 1 per pressure node */
  FEDD_VEC_B_P = 9,    /* pressure part of the NS linear-solve right-hand side */
  FEDD_VEC_FPRES = 10  /* follower pressure load of the elasticity problem, 3 per node */
};
#define FEDD_NUM_VECTORS 11
/* Surface flags a pressure table may cover (flags 0 .. FEDD_MAX_SURFACE_FLAGS - 1). */
#define FEDD_MAX_SURFACE_FLAGS 16

enum fedd_precision {
  FEDD_FP64 = 0,       /* everything in FP64 */
  FEDD_FP32_LOCAL = 1  /* element-matrix products in FP32; geometry (Binv, |det|), the global
                          accumulation and every residual (NS Rhs mode, elasticity) in FP64 */
};

enum fedd_ns_mode {
  FEDD_NS_JACOBIAN = 0,
  FEDD_NS_FIXEDPOINT = 1,
  FEDD_NS_RHS = 2,
  /* No single counterpart in the reference: one pass that does what "Rhs" and
   * "Jacobian" (or "FixedPoint") do on the same iterate, as the Newton cycle
   * calculateNonLinResidualVec + assemble("Newton") needs them
   * (NonLinearSolver_def.hpp:371-377, NavierStokesAssFE_def.hpp:225-313).  The residual
   * is FP64 in both precision modes. */
  FEDD_NS_JACOBIAN_RHS = 3,
  FEDD_NS_FIXEDPOINT_RHS = 4
};

/* Parameter keys of AssembleFENavierStokes (AssembleFENavierStokes_def.hpp:44-53):
 * Parameter/Viscosity, Parameter/Density, General/Linearization ("Newton" -> 1,
 * "FixedPoint" -> 0), and the 2x2 coeff of setCoeff (:57-64), row-major. */
typedef struct {
  double viscosity;
  double density;
  int newton;
  double coeff[4];
} fedd_ns_params;

/* ---------------------------------------------------------------- context */
int fedd_device_count(int* count);
/* Creates a context on `device` with its own HIP stream. */
int fedd_ctx_create(int device, fedd_ctx** ctx);
void fedd_ctx_destroy(fedd_ctx* ctx);
const char* fedd_last_error(const fedd_ctx* ctx);
int fedd_synchronize(fedd_ctx* ctx);

/* ------------------------------------------------------------------- mesh */
/* MeshStructured::buildMesh3D (MeshStructured_def.hpp:622-1009) on one rank for the
 * unit cube: fe = 1 (P1, (M+1)^3 points) or 2 (P2, (2M+1)^3 points), 6*M^3 tets.
 * pts: npts*3, bc: npts (1 on the boundary, may be NULL), elems: nelem*(4|10). */
int fedd_mesh_structured(int M, int fe, double* pts, int32_t* bc, int32_t* elems);
/* Mesh::correctElementOrientation + flipElement (Mesh_def.hpp:543-578, 648-726). */
int fedd_mesh_correct_orientation(const double* pts, int64_t npts, int32_t* elems, int64_t nelem,
                                  int nodes_per_elem, int64_t* flipped);

/* Domain::buildMesh / getPointsRepeated / getElementsC (Domain_def.hpp:210-280).
 * slot PRIMARY: the P1 (scalar Laplace) or P2 (velocity, displacement) mesh;
 * slot PRESSURE: the P1 pressure mesh of a P2-P1 pair, element T of which is the
 * pressure element of velocity element T (FE_def.hpp:1099-1101, 1157).  Replaces
 * the mesh and drops graphs built on it. */
int fedd_mesh_upload(fedd_ctx* ctx, int slot, const double* pts, int64_t npts,
                     const int32_t* elems, int64_t nelem, int nodes_per_elem);

/* ------------------------------------------------------------------ graph */
/* Builds the CRS graph(s) exactly as insertGlobalValues + fillComplete would
 * (FE_def.hpp:1419-1442 addFeBlock, :982-1062 addFeBlockMatrix, :2470-2496
 * assemblyLaplace), plus the node incidence and scatter indices the kernels use.
 * Rows never hold more than 255 distinct neighbour nodes (FEDD_ERR_LIMIT). */
int fedd_graph_build(fedd_ctx* ctx, int kind);
/* Summation order of the row gathers.  No counterpart in the reference, which sums an
 * entry's element contributions in a fixed order (Tpetra fillComplete), so its matrices
 * repeat bit for bit.  The GPU row gathers (on = 0, the default) add them in the order in
 * which a workgroup's waves reach LDS: repeated assemblies agree to the parity rules but
 * not always in the last bits.  on = 1: the Navier-Stokes velocity and pressure plans and
 * the scalar (Laplace, load vector) plans place all of a node's pairs in one wave, so
 * repeated assemblies of the same state give identical bits; the NS velocity rows then
 * take about 15% longer (40^3 x 6).  The vector-graph (elasticity) plan is not changed.
 * Applies to graphs built after the call. */
int fedd_set_reproducible(fedd_ctx* ctx, int on);
/* How the Navier-Stokes matrices are assembled.  The reference inserts each element's
 * matrix into the Tpetra CrsMatrix (FE::assemblyNavierStokes element loop,
 * FE_def.hpp:1153-1176; addFeBlockMatrix :982-1062, insertGlobalValues adding into
 * existing entries); both methods fill the same graph with values equal to the parity
 * rules.
 *   FEDD_METHOD_ROW_GATHER (default): owner-computes row gathers, every value written
 *     once from LDS.
 *   FEDD_METHOD_COLOUR_SCATTER: the elements are coloured (fedd_colour_build's greedy
 *     rule; elements of one colour share no node) and, colour after colour, each
 *     (element, local node) pair adds its row into the global values with no-return FP64
 *     atomics that never meet inside a colour.  Every value receives its contributions in
 *     colour order, so repeated assemblies agree bit for bit.  The values are zeroed first;
 *     each contribution is a read-modify-write in the L2, so this method moves several
 *     times the row gather's bytes.  The fused residual modes assemble the matrices this
 *     way and the residual as FEDD_NS_RHS does.
 * Applies to NS graphs built after the call (the scatter schedule is built with them). */
enum {
  FEDD_METHOD_ROW_GATHER = 0,
  FEDD_METHOD_COLOUR_SCATTER = 1
};
int fedd_set_assembly_method(fedd_ctx* ctx, int method);
/* Scatter indices of the coloured scatter schedule (NS graph built with
 * FEDD_METHOD_COLOUR_SCATTER; else FEDD_ERR_STATE): for every element e, the node-level
 * CRS position of each of its element-matrix blocks, local node order of the element.
 *   block 0 (A):   idx[100 e + 10 a + b] = k with node-block column nidx[k] = node b of e
 *                  in node row a (entry (3 i + d1, 3 j + d2) is value 9 k0 + 3 deg d1 +
 *                  3 (k - k0) + d2 of block A, k0 the row's first block, deg its blocks);
 *   block 1 (B^T): idx[40 e + 4 a + m], velocity node a, pressure vertex m;
 *   block 2 (B):   idx[40 e + 10 m + b], pressure vertex m, velocity node b.
 * idx holds 100 nelem (block 0) or 40 nelem (blocks 1, 2) entries. */
int fedd_scatter_indices(fedd_ctx* ctx, int block, int64_t* idx);
int fedd_block_info(fedd_ctx* ctx, int block, int64_t* nrows, int64_t* ncols, int64_t* nnz);
/* Host copies: rowptr has nrows+1 entries, colind nnz (local = global ids on one rank). */
int fedd_block_graph(fedd_ctx* ctx, int block, int64_t* rowptr, int32_t* colind);
int fedd_block_values(fedd_ctx* ctx, int block, double* values);
/* Device pointer of the value array (owned by the context, valid until the graph is rebuilt). */
int fedd_block_device_values(fedd_ctx* ctx, int block, double** d_values);

/* --------------------------------------------------------- export formats */
/* BlockMatrix::merge (BlockMatrix_def.hpp:119-163) of the NS blocks into one square CRS
 * matrix: block (i, j) at row/column offsets (0 | 3 N2); rows sorted, explicit zeros kept. */
int fedd_ns_merged_info(fedd_ctx* ctx, int64_t* nrows, int64_t* nnz);
int fedd_ns_merged_export(fedd_ctx* ctx, int64_t* rowptr, int32_t* colind, double* values);
/* Matrix::writeMM (Matrix_def.hpp:264-277): MatrixMarket coordinate real general file of
 * one block (0..3) or, with block = -1, of the merged NS matrix; 1-based indices, values
 * with 17 significant digits. */
int fedd_write_matrix_market(fedd_ctx* ctx, int block, const char* path);

/* EpetraExt-HDF5 vector reader: HDF5Import::readVariablesHDF5 (HDF5Import_def.hpp:52-70).
 * Reads the multivector group `name` (datasets Values, GlobalLength, NumVectors) of a
 * file written by HDF5Export / EpetraExt::HDF5.  On one rank the reference's read map
 * lists every global id in order, so out[k] is the value of global id k (column 0,
 * the column the reference copies into its vector).  out may be NULL to query
 * global_length / num_vectors; cap is the capacity of out.  Context-free: a missing
 * variable ("Requested varName ... not contained", HDF5Import_def.hpp:53) or an
 * unreadable file returns FEDD_ERR_ARG with the message in fedd_last_error(NULL). */
int fedd_hdf5_read_vector(const char* path, const char* name, double* out, int64_t cap,
                          int64_t* global_length, int* num_vectors);

/* ------------------------------------------------------------- colouring */
/* Element colouring of the primary mesh: elements of one colour share no node.
 * No counterpart in the reference (it assembles serially per rank).  The coloured
 * scatter (FEDD_METHOD_COLOUR_SCATTER) runs over it; it is also exported for callers
 * that scatter element-parallel themselves (e.g. a threaded CPU assembly).  colour_ptr
 * has ncolours+1 entries, elem_perm nelem entries (elements sorted by colour, ascending
 * id). */
int fedd_colour_build(fedd_ctx* ctx, int* ncolours);
int fedd_colour_export(fedd_ctx* ctx, int64_t* colour_ptr, int32_t* elem_perm);

/* ---------------------------------------------------------------- vectors */
/* Repeated-map vectors (one rank: repeated = unique).  Upload copies host data
 * to the context's device vector; download copies back (stream-synchronous). */
int fedd_vector_upload(fedd_ctx* ctx, int which, const double* host, int64_t n);
int fedd_vector_download(fedd_ctx* ctx, int which, double* host, int64_t n);
int fedd_vector_device_ptr(fedd_ctx* ctx, int which, double** d_ptr, int64_t* n);

/* --------------------------------------------------------------- assembly */
/* FE::assemblyLaplace (FE_def.hpp:2437-2500) for P1/P2 on the primary mesh,
 * rule from determineDegree(dim,FEType,FEType,Grad,Grad).  Needs GRAPH_SCALAR. */
int fedd_assembly_laplace(fedd_ctx* ctx, int precision);
/* FE::assemblyRHS (FE_def.hpp:7303-7395), scalar field, constant f: writes VEC_RHS. */
int fedd_assembly_rhs(fedd_ctx* ctx, double f);
/* FE::assemblyRHS with a general source (FE_def.hpp:7352-7391).  The reference calls
 * func at the physical quadrature points of the rule determineDegree(dim, FEType, Std)
 * (P1: the centroid, P2: the 5-point rule, Helper.cpp:1310-1350) and adds
 * |det B| sum_q w_q phi_i(q) f_d(x_q) at row node ("Scalar") or 3 node + d ("Vector").
 * fedd_rhs_quadrature_points writes those points, xq[(e*nq + q)*3 + j] (cap doubles,
 * xq may be NULL to query nq); the caller evaluates its source there and passes
 * fq[(e*nq + q)*ncomp + d], ncomp 1 or 3, to fedd_assembly_rhs_points, which writes
 * VEC_RHS with ncomp entries per node.  Needs GRAPH_SCALAR. */
int fedd_rhs_quadrature_points(fedd_ctx* ctx, double* xq, int64_t cap, int* nq);
int fedd_assembly_rhs_points(fedd_ctx* ctx, const double* fq, int64_t n, int ncomp);
/* FE::assemblyNavierStokes (FE_def.hpp:1078-1239) with AssembleFENavierStokes
 * (AssembleFENavierStokes_def.hpp:67-224).  mode: FEDD_NS_*.  reassemble = 1 fills
 * only A(0,0) (addFeBlock), 0 fills A, B^T, B (and the empty C) (addFeBlockMatrix).
 * Jacobian / FixedPoint read VEC_U (and VEC_P) and overwrite the block values;
 * Rhs writes VEC_RES_U and VEC_RES_P; the fused modes do both in one pass.  Every
 * mode but Rhs advances the Newton-step counter.  Needs GRAPH_NS. */
int fedd_assembly_navier_stokes(fedd_ctx* ctx, const fedd_ns_params* params, int mode,
                                int reassemble, int precision);
/* FE::assemblyLinearElasticity (FE_def.hpp:362-434): the "LinearElasticity"
 * AssembleFE element, i.e. AceGen skr1 (NeoHookQuadraticTets.f, 4-point rule), at
 * the displacement VEC_U; element matrix = -s (AssembleFE_LinElas_def.hpp:109-119).
 * Fills A of a GRAPH_VECTOR graph on a P2 mesh.  Parameter/E, Parameter/Poisson Ratio. */
int fedd_assembly_linear_elasticity(fedd_ctx* ctx, double E, double nu, int precision);
/* FE::assemblyLinElasXDim (FE_def.hpp:4701-5004), 3D, as LinElas::assemble calls it
 * (LinElas_def.hpp:64-84): the linear operator int 2 mu eps(u):eps(v) + lambda div u div v
 * into A of a GRAPH_VECTOR graph on a P1 or P2 mesh (1-point / 5-point rules, both exact
 * here).  Needs no displacement. */
int fedd_assembly_linear_elasticity_xdim(fedd_ctx* ctx, double lambda, double mu, int precision);
/* FE::assemblyNonLinearElasticity (FE_def.hpp:449-523): tangent into A and residual
 * (rhs = p of the element, AssembleFE_NonLinElas_def.hpp:131-173) into VEC_RES_U.
 * neo_hooke_model = Parameter/"Neo-Hooke Modell": 1 -> skr2, otherwise skr3. */
int fedd_assembly_nonlinear_elasticity(fedd_ctx* ctx, int neo_hooke_model, double E, double nu,
                                       int precision);
/* The overload with a per-element Young's modulus field eModVec
 * (FE::assemblyNonLinearElasticity(..., domain, eModVec, ...), FE_def.hpp:538-614,
 * updateParameter("E", eModVecA[T]) per element): E_elem has one entry per element. */
int fedd_assembly_nonlinear_elasticity_efield(fedd_ctx* ctx, int neo_hooke_model, const double* E_elem,
                                              int64_t nelem, double nu, int precision);
/* fedd_assembly_nonlinear_elasticity with a choice of tangent: tangent 0 the exact law
 * (the same call), 1 the learned surrogate of fedd_mlp_* at every quadrature point (the
 * residual VEC_RES_U is always the exact law's).  No counterpart in the reference, whose
 * tangent is always AceGen's.  Needs a network built (fedd_mlp_init) for this law and
 * these constants, else FEDD_ERR_STATE. */
int fedd_assembly_nonlinear_elasticity_tangent(fedd_ctx* ctx, int neo_hooke_model, double E, double nu,
                                                int precision, int tangent);
/* FP32-local elasticity guards, both applied on the device by the same call:
 *  - an element with a quadrature point where FP32 products cannot resolve the tangent
 *    (kappa = |C|_F |adj C|_F / det C > 1e4 for C = F^T F, which includes J = 0: measured
 *    FP32 error about 10 u32 kappa of the element's largest entry) has its (node, element)
 *    pairs computed in FP64 instead; *npairs (may be null) counts them;
 *  - a chunk of rows whose FP32 tangent still holds a non-finite entry (the law's tangent
 *    can leave FP32's range) is re-assembled in FP64; *nchunks counts them,
 * so no non-finite FP32 entry is written where the FP64 path is finite.  Counts of the
 * last elasticity assembly (0 for FP64 assemblies); synchronises the context stream. */
int fedd_elasticity_fp64_fixups(fedd_ctx* ctx, int64_t* nchunks, int64_t* npairs);
/* ------------------------------------------------------- surface loads */
/* MeshStructured::buildSurfaces(3, "P2") (MeshStructured_def.hpp:3707-3876) on a P2 tet
 * mesh of the box {x0, y0, z0, L, W, H}: every element face lying on a side of the box, as
 * a 6-node triangle in FEDDLib's P2 triangle order (vertices, then edges 01, 12, 02),
 * oriented so that (x1 - x0) x (x2 - x0) points out of the box, with the side's flag
 * (x = x0: 1, y = y0: 2, z = z0: 3, z = z0 + H: 4, y = y0 + W: 5, x = x0 + L: 6), in
 * element order and, within an element, face order.  faces (6 per face) and flags may be
 * null to query the count *nface.  Host code; no context. */
int fedd_structured_surfaces(const double* pts, int64_t npts, const int32_t* elems, int64_t nelem,
                             const double* box, int32_t* faces, int32_t* flags, int64_t* nface);
/* The surface triangles of the primary mesh that pressure loads act on (the surface
 * sub-elements of FiniteElement::getSubElements, FE_def.hpp:6886-6893): nface x 6 node ids
 * as above and one flag (>= 0) each.  A primary mesh upload clears them. */
int fedd_surface_set(fedd_ctx* ctx, const int32_t* faces, const int32_t* flags, int64_t nface);
/* FE::assemblySurfaceIntegralExternal (FE_def.hpp:6859-6925, with_tangent = 0) and
 * FE::assemblyNonlinearSurfaceIntegralExternal (:6932-7045, with_tangent = 1): vector `out`
 * (FEDD_VEC_FEXT, FEDD_VEC_FPRES, FEDD_VEC_B or FEDD_VEC_RES_U) is zeroed, then every surface
 * triangle with p = pressure[flag] != 0 (the reference's func(.., flag)) adds the follower
 * load of a pressure acting against the outward normal of the current surface X + VEC_U,
 *     f_a = -p int N_a (x_xi x x_eta) dxi deta,
 * the load of AceGen's PressureTriangle3D6 (:6912-6916).  with_tangent also adds
 * Kext = -df/du to block A of the vector graph (A + Kext, NonLinElasAssFE_def.hpp:295-300).
 * Flags >= nflag carry no pressure. */
int fedd_assembly_surface_pressure(fedd_ctx* ctx, const double* pressure, int nflag, int with_tangent, int out);
/* Follower pressure per flag for fedd_newton_elasticity (nflag = 0: none): ramped with the
 * load factor like the traction load, its load and Kext are re-assembled at every Newton
 * iterate, and the Newton steps are solved by BiCGStab (Kext is not symmetric) with the
 * same preconditioners. */
int fedd_newton_pressure(fedd_ctx* ctx, const double* pressure, int nflag);
/* ------------------------------------------------------- nonlinear solve */
/* Dirichlet dofs of the primary field (the dofs BCBuilder::addBC flags,
 * BCBuilder_def.hpp:590-700): one entry per constrained dof, dof = node for a scalar graph,
 * 3 node + d otherwise (the graph built when the set is made decides).  Values are zero
 * until fedd_dirichlet_values. */
int fedd_dirichlet_set(fedd_ctx* ctx, const int32_t* dofs, int64_t n);
/* Dirichlet values g, one per dof of the primary field (only the constrained ones are read). */
int fedd_dirichlet_values(fedd_ctx* ctx, const double* g, int64_t n);
/* BCBuilder::setSystem (BCBuilder_def.hpp:590-618) with setDirichletBC /
 * setLocalRowOne / setLocalRowZero (:621-704) on the block row of the primary field:
 * every constrained row of the diagonal block (A) becomes an identity row (1 in its own
 * column, every other stored entry 0) and the same row of the off-diagonal block (NS B^T)
 * a zero row.  Columns, the pressure rows (B, C) and the graph are left as they are. */
int fedd_dirichlet_set_system(fedd_ctx* ctx);
enum fedd_bc_vector_mode {
  FEDD_BC_RHS = 0,            /* out = g        BCBuilder::setRHS (BCBuilder_def.hpp:93-160) */
  FEDD_BC_MINUS_VECTOR = 1,   /* out = g - sub  BCBuilder::setBCMinusVector (:234-300) */
  FEDD_BC_VECTOR_MINUS_BC = 2 /* out = sub - g  BCBuilder::setVectorMinusBC (:350-420) */
};
/* The Dirichlet values into vector `out` (a FEDD_VEC_* of one entry per primary-field dof)
 * at the constrained dofs; `sub` is the subtracted vector of modes 1 and 2 (ignored for 0).
 * The other entries of `out` are left as they are. */
int fedd_dirichlet_vector(fedd_ctx* ctx, int mode, int out, int sub);
/* Symmetric elimination for CG (no counterpart in the reference, whose setSystem above
 * rewrites rows only): constrained rows AND columns of block A of a vector graph become
 * identity rows/columns, which keeps an SPD tangent SPD; valid for homogeneous increments
 * (the Newton solver's constrained increments are 0).  The first call after a
 * fedd_dirichlet_set or a graph build lists the rows the constraints touch (one stream
 * synchronisation); later calls rewrite only those rows.  Inside an open capture with no
 * list built, every row is scanned instead. */
int fedd_apply_dirichlet_symmetric(fedd_ctx* ctx);
/* Uniform traction t on the face {x_axis == value} of the P2 mesh into VEC_FEXT
 * (surface term of the load-stepped source, NonLinElasAssFE_def.hpp:262-325). */
int fedd_traction_load(fedd_ctx* ctx, int axis, double value, const double* traction);
/* Preconditioned CG on block A: solves A x = VEC_B into VEC_X (Belos CG behind
 * LinearSolver_def.hpp:23-137), preconditioner chosen by fedd_set_preconditioner. */
int fedd_linear_solve(fedd_ctx* ctx, double rtol, int maxit, int* iters, double* relres);
/* 0: 3x3 block Jacobi (any mesh); 1: geometric multigrid V-cycle on the structured
 * lattice (P2 -> P1 -> coarser P1, Galerkin coarse operators, damped block-Jacobi
 * smoothing, dense coarsest solve); 2: smoothed-aggregation multigrid on any mesh (greedy
 * node aggregates, prolongator (I - omega D^-1 A) P_tent with P_tent spanning the six
 * rigid-body modes per aggregate, Galerkin coarse operators, the same smoother and
 * coarsest solve).  1 and 2 stand in for the reference's FROSch/MueLu preconditioners
 * (Preconditioner_def.hpp:206-291). */
int fedd_set_preconditioner(fedd_ctx* ctx, int kind);
/* No counterpart in the reference.  For callers that cannot allocate device memory
 * (code under stream capture, tight serving loops, harnesses that forbid hipMalloc):
 * the block-Jacobi CG of fedd_linear_solve (preconditioner 0, whatever
 * fedd_set_preconditioner chose) on block A, A x = b from x = 0, with caller-owned
 * device vectors d_b (read) and d_x (written), n = rows of A.  Its work space (four
 * Krylov vectors, the 3x3 block inverses, the reduction slots) is one
 * FEDD_NOALLOC_SCRATCH_BYTES device array per GPU, held until the process ends, so
 * n <= FEDD_NOALLOC_MAX_DOFS (FEDD_ERR_LIMIT above).  It calls no hipMalloc / hipFree.
 * Every call on one GPU shares that array: calls that overlap on one GPU (two host
 * threads, or two contexts on the same device) race.  It synchronises the context
 * stream, so it cannot be captured. */
#define FEDD_NOALLOC_SCRATCH_BYTES (32ll << 20)
#define FEDD_NOALLOC_MAX_DOFS 599040
int fedd_linear_solve_noalloc(fedd_ctx* ctx, const double* d_b, double* d_x, int64_t n, double rtol, int maxit,
                              int* iters, double* relres);

typedef struct {
  int neo_hooke_model;  /* Parameter/"Neo-Hooke Modell": 1 -> skr2 law, otherwise skr3 */
  double E, nu;
  int load_steps;       /* lambda = k/load_steps, k = 1..load_steps */
  double abs_tol;       /* converged when ||F_int - lambda F_ext||_2 < abs_tol (constrained dofs 0) */
  int max_newton;       /* per load step */
  double lin_rtol;      /* floor of the inexact-Newton forcing term */
  int lin_maxit;
  int tangent;          /* 0: exact tangent, 1: learned surrogate (exact residual always) */
  int precision;        /* FEDD_FP64 or FEDD_FP32_LOCAL for the tangent */
  double stall_ratio;   /* surrogate: switch to the exact tangent for the rest of the load step
                           when ||R_k+1|| > stall_ratio ||R_k|| (the step is kept), and reject
                           the step first when ||R_k+1|| > ||R_k|| or is not finite */
  int collect_samples;  /* nonzero: append the quadrature-point states of every Newton iterate
                           to the surrogate's training set (needs fedd_mlp_init) */
  int preconditioner;   /* as fedd_set_preconditioner: 0 block Jacobi, 1 geometric multigrid
                           (structured lattice meshes of fedd_mesh_structured only),
                           2 aggregation multigrid (any mesh) */
  int mg_reuse;         /* multigrid only.  0: rebuild the hierarchy for every tangent;
                           1: keep the coarse levels built for an earlier tangent, across
                           Newton and load steps (the finest level always follows the
                           current tangent), until a solve needs more than 2 n0 + 2 CG
                           iterations, n0 those of the first solve after the rebuild.  Only
                           the preconditioner changes: every solve still meets its
                           tolerance for the current tangent. */
  int criterion;        /* Parameter/Criterion of NonLinearSolver::solveNewton
                           (NonLinearSolver_def.hpp:367-402), in addition to abs_tol:
                           0 none; 1 "Residual": a load step also ends when
                           ||R_k|| / ||R_0|| < rel_tol (R_0 its first residual); 2 "Update":
                           when ||du_k||_2 < rel_tol after an update */
  double rel_tol;       /* Parameter/relNonLinTol of criteria 1 and 2 */
} fedd_newton_params;

typedef struct {
  int newton_iters;      /* linear solves (Newton updates), rejected ones included */
  int fallbacks;         /* load steps switched from the surrogate to the exact tangent */
  int rejected;          /* surrogate steps rejected (residual grew or became non-finite) */
  int64_t krylov_iters;
  double seconds;        /* wall time of the whole solve */
  double assembly_ms;    /* tangent + residual assembly (device time) */
  double solve_ms;       /* Krylov solves (device time) */
  double final_residual;
  int converged;
  int history_len;
  double rel_residual;   /* ||R|| / ||R_0|| at the end of the last load step (logged for every criterion) */
  double last_update;    /* ||du||_2 of the last update (criterion 2 only, else 0) */
} fedd_newton_stats;

/* NonLinearSolver::solveNewton (NonLinearSolver_def.hpp:356-414) for the hyperelastic
 * problem on the device: VEC_U is the initial displacement and the solution;
 * history (may be NULL) receives ||R|| at every Newton iterate. */
int fedd_newton_elasticity(fedd_ctx* ctx, const fedd_newton_params* params, fedd_newton_stats* stats,
                           double* history, int history_cap);

/* Navier-Stokes saddle-point solve on the current NS blocks (after fedd_dirichlet_set_system):
 * [A B^T; B C] [x_u; x_p] = [VEC_B; VEC_B_P] into VEC_X, VEC_X_P, from zero, by restarted
 * right-preconditioned GMRES(restart) (Belos GMRES behind LinearSolver_def.hpp:23-137) with
 * the upper block-triangular preconditioner [A~ B^T; 0 -S~], A~ the 3x3 block diagonal of A,
 * S~ = diag(B diag(A)^-1 B^T).  iters: GMRES iterations; relres: final true residual
 * ||b - J x|| / ||b||.  restart <= 120. */
int fedd_ns_linear_solve(fedd_ctx* ctx, double rtol, int restart, int maxit, int* iters, double* relres);

typedef struct {
  double viscosity, density;  /* Parameter/Viscosity, Parameter/Density */
  int linearization;          /* General/Linearization: 1 "Newton", 0 "FixedPoint" (Picard) */
  int criterion;              /* Parameter/Criterion: 0 "Residual" (||r_k|| / ||r_0|| < rel_tol),
                                 1 "Update" (||delta_k||_2 < rel_tol) */
  double rel_tol;             /* Parameter/relNonLinTol */
  int max_its;                /* Parameter/MaxNonLinIts */
  double lin_rtol;            /* GMRES relative tolerance per Newton step */
  int restart, lin_maxit;     /* GMRES restart length (<= 120) and iteration cap per step */
} fedd_ns_newton_params;

typedef struct {
  int newton_iters;           /* linear solves (updates) */
  int64_t krylov_iters;
  double final_residual;      /* ||r|| at the last iterate */
  double rel_residual;        /* ||r|| / ||r_0|| (criterion "Residual") */
  double last_update;         /* ||delta|| of the last update */
  double last_lin_relres;     /* true relative residual of the last GMRES solve */
  int converged;
  int history_len;
  double seconds, assembly_ms, solve_ms;
} fedd_ns_newton_stats;

/* NonLinearSolver::solveNewton (NonLinearSolver_def.hpp:356-414) for NavierStokesAssFE
 * (NavierStokesAssFE_def.hpp:225-313) on the device: VEC_U, VEC_P are the initial iterate
 * and the solution.  Every iteration assembles the Jacobian (FixedPoint matrix for
 * linearization 0) and F(u) in one fused pass, forms r = rhs - F(u) (velocity rhs
 * VEC_FEXT if uploaded with 3 entries per node, else 0; pressure rhs 0), sets
 * r = g - u at the Dirichlet dofs (BCBuilder::setBCMinusVector, values of
 * fedd_dirichlet_values), tests the criterion, rewrites the Dirichlet rows
 * (BCBuilder::setSystem) and adds the GMRES solution of J delta = r.  Unlike the
 * reference, the residual of the final iterate is evaluated (and tested) after the last
 * update.  history (may be NULL) receives ||r|| at every iterate. */
int fedd_newton_navier_stokes(fedd_ctx* ctx, const fedd_ns_newton_params* params, fedd_ns_newton_stats* stats,
                              double* history, int history_cap);

/* ------------------------------------------------- learned tangent surrogate
 * None of these has a counterpart in the reference (FEDDLib always evaluates the
 * AceGen tangent).  A 9 -> 64 -> 64 -> 21 GELU MLP maps x = (F - I)/s_in to
 * y = (CC - CC0)/s_out, CC the 21 upper-triangle entries of the 6x6 Voigt tangent
 * dS/dE (order 11,22,33,12,13,23), CC0 the tangent at F = I, s_out = lambda + 2 mu,
 * s_in = 0.1.  BF16 MFMA with FP32 accumulation, FP32 master weights, Adam.
 * Parameter vector (FEDD_MLP_PARAMS floats, row-major, zero padding):
 *   W1[64][16] b1[64] W2[64][64] b2[64] W3[32][64] b3[32]
 * (inputs 9 padded to 16, outputs 21 padded to 32). */
#define FEDD_MLP_PARAMS 7328
/* New network for law neo_hooke_model (as fedd_newton_params) with Xavier-normal
 * weights from `seed`; drops any training set. */
int fedd_mlp_init(fedd_ctx* ctx, int neo_hooke_model, double E, double nu, uint64_t seed);
/* Training samples with exact-law targets: F = I + N(0, sigma^2) per entry (det F > 0.2). */
int fedd_mlp_add_samples_random(fedd_ctx* ctx, int64_t n, double sigma, uint64_t seed);
/* Training samples from the quadrature points of the last hyperelastic assembly
 * (the solver's own iterates). */
int fedd_mlp_add_samples_qp(fedd_ctx* ctx);
int fedd_mlp_num_samples(fedd_ctx* ctx, int64_t* n);
/* `steps` Adam steps (batches of `batch` from shuffled epochs, learning rate decayed
 * from lr to lr/50 on a cosine) on the device;
 * loss_history (may be NULL) gets the batch loss sum_o (y_o - t_o)^2 / batch. */
int fedd_mlp_train(fedd_ctx* ctx, int steps, int batch, float lr, uint64_t seed, float* loss_history);
int fedd_mlp_params(fedd_ctx* ctx, float* params);            /* read FEDD_MLP_PARAMS floats */
int fedd_mlp_set_params(fedd_ctx* ctx, const float* params);  /* load (checkpoint / tests) */
/* Network output y[n][21] for normalised inputs x[n][9]. */
int fedd_mlp_forward(fedd_ctx* ctx, const float* x, int64_t n, float* y);
/* Gradient of the batch loss over all n samples (as one batch) w.r.t. the parameters. */
int fedd_mlp_gradient(fedd_ctx* ctx, const float* x, const float* t, int64_t n, float* grad, float* loss);
/* Normalisation constants: cc0[21], s_out, s_in. */
int fedd_mlp_norm(fedd_ctx* ctx, float* cc0, float* s_out, double* s_in);

/* AssembleFE::advanceNewtonStep / getNewtonStep (AssembleFE_decl.hpp:146) for the
 * NS elements: the counter the reference keeps per element. */
int fedd_newton_step(fedd_ctx* ctx, int* step);

/* ---------------------------------------------------------- graph capture
 * No counterpart in the reference.  Records the work that the calls between begin
 * and end enqueue on the context stream into one HIP graph; fedd_graph_launch
 * replays it (stream-ordered, asynchronous).  Inputs are read from the context's
 * device buffers at replay time, so update them with fedd_vector_upload between
 * replays (same lengths keep the buffers in place).  Calls that allocate, transfer
 * host data or synchronise fail with FEDD_ERR_STATE (or FEDD_ERR_HIP) while a capture
 * is open: run an assembly once before capturing it.  The NS Newton-step counter
 * advances only for the captured call, not per replay.  No memset nodes are
 * recorded by the assembly calls. */
int fedd_capture_begin(fedd_ctx* ctx);
int fedd_capture_end(fedd_ctx* ctx, int64_t* nodes);  /* nodes: graph node count (may be NULL) */
int fedd_graph_launch(fedd_ctx* ctx);

/* ----------------------------------------------------------------- timing */
/* STREAM-style copy of `bytes` (read + write) repeated `iters` times on the context
 * stream: the attainable HBM bandwidth next to the 8 TB/s datasheet figure, and the
 * calibration kernel for byte counters.  No counterpart in the reference. */
int fedd_stream_copy(fedd_ctx* ctx, int64_t bytes, int iters, float* ms_per_copy);

/* HIP events recorded on the context stream.  Kernel timing, when enabled,
 * brackets every launch of the assembly kernels with events on that stream. */
int fedd_event_record(fedd_ctx* ctx, int slot);           /* slot < 16 */
int fedd_event_elapsed(fedd_ctx* ctx, int a, int b, float* ms);
int fedd_kernel_timing(fedd_ctx* ctx, int enable);
/* Accumulated time and launch count of the named kernel since enabling. */
int fedd_kernel_time(fedd_ctx* ctx, const char* name, double* total_ms, int64_t* launches);

#ifdef __cplusplus
}
#endif
#endif /* FEDDGPU_H */
