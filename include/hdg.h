/*
 * hdg.h -- C ABI of libhdg, the MI355X (gfx950) engine for GridapHybrid's
 * per-cell hybrid pipeline: cell+facet integration of the HDG local systems,
 * static condensation, deterministic skeleton assembly and backward recovery.
 *
 * Conventions
 *   - plain pointers and sizes; "device" pointers are HIP device memory owned
 *     by the caller, "host" pointers are ordinary host memory;
 *   - per-cell matrices are column-major and cell-major ([cell][col][row]),
 *     the layout of Julia's Matrix{Float64};
 *   - ids are 0-based; trace dof ids keep Gridap's convention of negative ids
 *     (-1,-2,...) for Dirichlet dofs;
 *   - every function returns an hdg_status; hdg_last_error() describes the
 *     last failure of a context;
 *   - a context is not thread-safe (one per host thread, as Gridap caches).
 */
#ifndef HDG_H
#define HDG_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef enum {
  HDG_OK = 0,
  HDG_ERR_ARG = -1,         /* invalid argument / precondition (Gridap @check) */
  HDG_ERR_HIP = -2,         /* HIP runtime error */
  HDG_ERR_ALLOC = -3,       /* device allocation failed */
  HDG_ERR_UNSUPPORTED = -4, /* size/physics combination not built */
  HDG_ERR_STATE = -5,       /* call order violated (e.g. recover before condense) */
  HDG_ERR_NO_DEVICE = -6,   /* no HIP device / library not usable */
  HDG_SINGULAR = 1          /* a local A_ii had a zero pivot (getrf info > 0) */
} hdg_status;

/* HDG_FP64: getrf-order elimination of the whole cell system in FP64.
 * HDG_MIXED: A_ii is inverted in FP32 (Gauss-Jordan, partial pivoting),
 * W = A_ii^-1 [A_ib b_i] is formed with FP32 products and refined with FP64
 * residuals B - A_ii W until the correction reaches FP64 rounding or stops
 * shrinking; S, g and the recovery are FP64.  Local systems with ni < 64
 * (where the FP64 kernels are HBM-bound and faster) or ni > 128 use the FP64
 * path under HDG_MIXED. */
typedef enum { HDG_FP64 = 0, HDG_MIXED = 1 } hdg_precision;
typedef enum { HDG_DARCY = 0, HDG_ELASTICITY = 1 } hdg_physics;
typedef enum { HDG_SPACE_Q = 0, HDG_SPACE_P = 1 } hdg_space;

/* Built-in sources (manufactured solutions).  HDG_SOURCE_QP means the source
 * values at cell quadrature points are supplied with hdg_mesh_set_source. */
typedef enum {
  HDG_SOURCE_ZERO = 0,
  HDG_SOURCE_DARCY_KAT = 1,     /* flux (1+x_a), potential -3.14 (test/DarcyHDGTests.jl:9-14) */
  HDG_SOURCE_DARCY_LINEAR = 2,  /* potential sum x_a, flux -kappa grad */
  HDG_SOURCE_DARCY_SINPROD = 3, /* potential prod sin(pi x_a), flux -kappa grad */
  HDG_SOURCE_ELAS_LINEAR = 11,  /* linear displacement (exact in the spaces) */
  HDG_SOURCE_ELAS_SIN = 12,     /* smooth trigonometric displacement */
  HDG_SOURCE_QP = 100
} hdg_source;

typedef struct hdg_ctx hdg_ctx;
typedef struct hdg_mesh hdg_mesh;
typedef struct hdg_batch hdg_batch;

/* ------------------------------------------------------------------------ */
/* Context                                                                  */
/* ------------------------------------------------------------------------ */
/* Creates a context on HIP device `device` with its own (default-priority)
 * stream.  Replaces nothing in the reference: Gridap's per-lazy-array caches
 * (src/StaticCondensationMap.jl:41-70) are the analogue. */
int hdg_ctx_create(int device, hdg_ctx** out);
int hdg_ctx_destroy(hdg_ctx* ctx);
const char* hdg_last_error(const hdg_ctx* ctx);
/* The stream all work of the context is enqueued on (a hipStream_t). */
void* hdg_ctx_stream(hdg_ctx* ctx);
int hdg_ctx_synchronize(hdg_ctx* ctx);
/* Device memory held for the context's meshes, batches and workspaces (bytes
 * live in its buffers; bytes of freed blocks it can reuse: its pool's and the
 * device's idle blocks, which every context of the device shares) and,
 * process-wide, bytes allocated outside any context's pool.  No counterpart
 * in the reference (Julia's GC owns its arrays); for callers that admit work
 * by device memory, such as a serving loop. */
int hdg_ctx_memory(hdg_ctx* ctx, int64_t* live_bytes, int64_t* pooled_bytes,
                   int64_t* unpooled_bytes);
/* Kernel timing.  Between hdg_profile_start and hdg_profile_stop every kernel
 * launch of the library (outside stream capture) is bracketed by two HIP
 * events on the stream it runs on; hdg_profile_stop synchronises and returns
 * the launch count and summed duration per kernel family (up to max entries;
 * *n receives the number of families).  Process-wide: one profile at a time. */
typedef struct {
  char name[32];
  int64_t launches;
  double total_ms;
} hdg_kernel_time;
int hdg_profile_start(hdg_ctx* ctx);
int hdg_profile_stop(hdg_ctx* ctx, int max, hdg_kernel_time* times, int* n);

/* Copies between host and device on the context stream (synchronous). */
int hdg_memcpy_d2h(hdg_ctx* ctx, void* host_dst, const void* dev_src, int64_t bytes);
int hdg_memcpy_h2d(hdg_ctx* ctx, void* dev_dst, const void* host_src, int64_t bytes);

/* ------------------------------------------------------------------------ */
/* Map-level API: batched StaticCondensationMap / BackwardStaticCondensationMap */
/* ------------------------------------------------------------------------ */
/* Field partition shared by the two maps.  field_sizes[nfields] are the
 * block sizes of the cell system (brs == bcs, src/StaticCondensationMap.jl:52);
 * interior_fields / boundary_fields are 0-based and must partition
 * 0..nfields-1 (_check_preconditions, src/StaticCondensationMap.jl:16-34). */
typedef struct {
  int nfields;
  const int32_t* field_sizes;
  int n_interior;
  const int32_t* interior_fields;
  int n_boundary;
  const int32_t* boundary_fields;
} hdg_fields;

/* StaticCondensationMap.evaluate! over ncells cells
 * (replaces src/StaticCondensationMap.jl:152-196 applied cell by cell by
 * lazy_map, src/HybridAffineFEOperators.jl:333-341).
 *   A    device [ncells][n][n] column-major, n = sum(field_sizes), fields in
 *        their natural order (the map extracts interior/boundary blocks);
 *   b    device [ncells][n];
 *   S    device [ncells][nb][nb] column-major  (A22 - A21 A11^-1 A12);
 *   g    device [ncells][nb]                   (b2 - A21 A11^-1 b1);
 *   info device [ncells] int32 (LAPACK getrf info, 0 = ok).
 * Returns HDG_SINGULAR if any cell had a zero pivot (the reference throws at
 * the @check info==0 of :180). */
int hdg_static_condensation(hdg_ctx* ctx, const hdg_fields* fields, int64_t ncells,
                            const double* A, const double* b, double* S, double* g,
                            int32_t* info);

/* hdg_static_condensation that allocates nothing: no reference counterpart.
 * For callers that cannot allocate device memory (code under HIP graph
 * capture, tight serving loops, harnesses that forbid hipMalloc).  Same
 * arguments, checks and results (to the FP64 bar of 1e-12; the kernel family
 * may differ) as hdg_static_condensation, for cell systems of at most 175 rows
 * whose kernel needs no workspace (HDG_ERR_UNSUPPORTED otherwise).  Its
 * scratch (field permutation, singular-cell counter) is a fixed 1,280-byte
 * device array per GPU, held until the process ends: calls that overlap on
 * one GPU race on it. */
int hdg_static_condensation_noalloc(hdg_ctx* ctx, const hdg_fields* fields, int64_t ncells,
                                    const double* A, const double* b, double* S, double* g,
                                    int32_t* info);

/* BackwardStaticCondensationMap.evaluate! over ncells cells
 * (replaces src/BackwardStaticCondensationMap.jl:61-102):
 *   x    device [ncells][nb] boundary (skeleton) values of each cell;
 *   xi   device [ncells][ni] recovered interior values, interior fields
 *        concatenated in the order of interior_fields (ReblockInteriorDofsMap
 *        splits them, :133-151). */
int hdg_backward_static_condensation(hdg_ctx* ctx, const hdg_fields* fields, int64_t ncells,
                                     const double* A, const double* b, const double* x,
                                     double* xi, int32_t* info);

/* ------------------------------------------------------------------------ */
/* Operator-level API: HybridAffineFEOperator + solve                        */
/* ------------------------------------------------------------------------ */
typedef struct {
  int physics;      /* hdg_physics */
  int dim;          /* 2 or 3 */
  int space;        /* hdg_space of the cell and trace spaces */
  int k_flux;       /* Darcy flux order / elasticity stress order */
  int k_pot;        /* Darcy potential order / elasticity displacement order */
  int k_trace;      /* trace order */
  int quad_degree;  /* Gauss degree of cell and facet rules (2(k+1) in the reference) */
  double tau;       /* HDG stabilisation */
  double young;     /* elasticity E */
  int source;       /* hdg_source */
  /* elasticity: 1 = the bulk-to-skeleton L2 projection P_m of the displacement
   * in the tau terms (src/GridapAPIExtensions.jl:453-819, test/P_m.jl:4-23,
   * bulk_to_skeleton_projection=true of test/LinearElasticityHDGTests.jl:
   * 124-150), 0 = P = identity.  Ignored for Darcy. */
  int projection;
} hdg_problem;

/* Mesh of affine cells (parallelograms / parallelepipeds, which includes every
 * CartesianDiscreteModel).  All arrays are host memory, copied at creation.
 *   x0[ncells*dim]           cell origin (first vertex);
 *   jac[ncells*dim*dim]      column-major Jacobian of the affine cell map;
 *   cell_facets[ncells*2dim] facet ids in Gridap's local facet order (local
 *                            facet f fixes reference direction dim-1-f/2 at
 *                            f%2; its coordinates are the other directions,
 *                            ascending);
 *   facet_dirichlet[nfacets] 1 if the facet carries Dirichlet trace dofs;
 *   coef[ncells]             kappa (Darcy) or Poisson ratio (elasticity);
 *   facet_perm[ncells*2dim]  orientation of each (cell, local facet): the code
 *                            p of the map from the cell's local facet
 *                            coordinates (s[,t]) to the facet's own ones, in
 *                            which the facet's trace dofs are numbered (the
 *                            reference's pindex, src/SkeletonArrays.jl:28-37,
 *                            src/Skeleton.jl:587-612).  2D: 0 s' = s,
 *                            1 s' = 1-s.  3D: (a,b) = (t,s) if p&1 else (s,t),
 *                            then s' = 1-a if p&2 else a, t' = 1-b if p&4 else
 *                            b.  NULL: every code 0 (a CartesianDiscreteModel).
 *                            A code under which the trace nodes are not mapped
 *                            onto themselves (P_k traces with k >= 2 on hexahedra
 *                            flipped) is HDG_ERR_UNSUPPORTED.
 * Owner signs n.n_o follow from cell_facets: the first (lowest id) cell
 * around a facet owns it (src/Skeleton.jl:135-140).  The condensed blocks of
 * hdg_cell_blocks keep each cell's local trace node order; cell_dofs maps
 * them to the facets' own numbering. */
typedef struct {
  int64_t ncells;
  int64_t nfacets;
  const double* x0;
  const double* jac;
  const int32_t* cell_facets;
  const int8_t* facet_dirichlet;
  const double* coef;
  const int8_t* facet_perm;
} hdg_mesh_desc;

typedef struct {
  int ni, nb;            /* local interior / skeleton sizes */
  int n_flux, n_pot;     /* interior field sizes */
  int n_trace_facet;     /* trace dofs per facet */
  int nq_cell;           /* cell quadrature points */
  int64_t ncells, nfacets;
  int64_t nfree;         /* free trace dofs (skeleton system size) */
  int64_t ndir;          /* Dirichlet trace dofs */
  int64_t nnz;           /* nonzeros of the skeleton matrix */
  int64_t nfree_total;   /* length of the recovered free-value vector */
} hdg_mesh_info;

/* Uploads the mesh, builds the trace dof map (Gridap numbering,
 * src/HybridAffineFEOperators.jl:388-439) and the symbolic skeleton pattern
 * (CSC, rows sorted per column), and keeps them resident. */
int hdg_mesh_create(hdg_ctx* ctx, const hdg_problem* prob, const hdg_mesh_desc* desc,
                    hdg_mesh** out);
/* Convenience: CartesianDiscreteModel(domain, cells) topology and geometry
 * built in the library (domain = [x0,x1,y0,y1(,z0,z1)]); Dirichlet on every
 * boundary facet (tags 5:8 in 2D, all boundary faces in 3D). */
int hdg_mesh_create_cartesian(hdg_ctx* ctx, const hdg_problem* prob, const int64_t* cells,
                              const double* domain, const double* coef, hdg_mesh** out);
int hdg_mesh_destroy(hdg_mesh* mesh);
/* Upper bound of the device bytes hdg_mesh_create_cartesian + a condensation
 * in `precision` would hold for this problem and partition (per-cell state,
 * skeleton pattern bound, dof maps, workspaces), computed on the host without
 * touching the device: the admission test of a serving loop.  No reference
 * counterpart (Gridap allocates lazily). */
int hdg_footprint_cartesian(const hdg_problem* prob, const int64_t* cells, int precision,
                            int64_t* device_bytes);
int hdg_mesh_get_info(const hdg_mesh* mesh, hdg_mesh_info* info);

/* Dirichlet trace values, host [ndir] (Gridap's dirichlet_values of the trial
 * skeleton space).  hdg_mesh_set_dirichlet_mms interpolates the built-in
 * manufactured trace at the facet nodes instead. */
int hdg_mesh_set_dirichlet(hdg_mesh* mesh, const double* lam_dir);
int hdg_mesh_set_dirichlet_mms(hdg_mesh* mesh);
/* The cell quadrature rule (Gauss of degree quad_degree, tensor points with
 * the first reference direction fastest): points host [nq_cell][dim] on the
 * reference cell [0,1]^dim, weights host [nq_cell].  The points at which
 * hdg_mesh_set_source expects the source (x = x0 + J xi).  Replaces Gridap's
 * CellQuadrature(trian, 2(k+1)) points (src/GridapTmpModifications.jl:13-40). */
int hdg_mesh_cell_quadrature(const hdg_mesh* mesh, double* points, double* weights);
/* Multi-field skeleton: the vector-valued trace as `nfields` scalar trace
 * fields, one per component (elasticity: nfields = dim), i.e. the reference's
 * HybridAffineFEOperator(weakform, X, Y, bulk_fields, [3, 4, ...]) with the
 * skeleton MultiFieldFESpace of _setup_fe_spaces_skeleton_system
 * (src/HybridAffineFEOperators.jl:52-63) and the condensed cell blocks
 * re-blocked per field by Scalar2ArrayBlockMap (src/Scalar2ArrayBlockMap.jl:
 * 42-69, _block_skeleton_system_contributions at
 * src/HybridAffineFEOperators.jl:343-356).  Every skeleton quantity at this
 * API then uses the multi-field consecutive numbering -- all dofs of field 1
 * (facet by facet, node order), then field 2, ...: the CSC of
 * hdg_skeleton_system, the trace of hdg_recover (input and the tail of the
 * free values), hdg_mesh_set_dirichlet / hdg_mesh_interpolate_trace and the
 * cell_dofs of hdg_cell_blocks (free ids per field block; Dirichlet ids -1,
 * -2, ... in the same field-major order).  S and g keep each cell's local
 * order.  nfields = 1 restores the single trace field.  HDG_ERR_UNSUPPORTED
 * for other counts. */
int hdg_mesh_set_trace_fields(hdg_mesh* mesh, int nfields);

/* Source values at cell quadrature points, device [ncells][nq_cell][ncomp]
 * (Darcy: f_1..f_D, g; elasticity: f_1..f_D); used when source == HDG_SOURCE_QP.
 * The pointer is borrowed until the next set or mesh destruction. */
int hdg_mesh_set_source(hdg_mesh* mesh, const double* src_qp);

/* HybridAffineFEOperator construction (src/HybridAffineFEOperators.jl:10-50):
 * fused cell integration + StaticCondensationMap for every cell, Dirichlet
 * lifting and deterministic assembly of the skeleton CSC matrix and rhs.
 * Keeps W = A_ii^-1 [A_ib b_i] resident for recovery.  *n_singular (host,
 * may be NULL) receives the number of cells with a zero pivot.  Under
 * HDG_MIXED, a mesh with any cell whose refinement did not converge is
 * condensed again in FP64 (reported by hdg_mesh_refinement_stats). */
int hdg_condense_assemble(hdg_mesh* mesh, int precision, int64_t* n_singular);

/* Refinement statistics of the last condensation (synchronises). */
typedef struct {
  int mixed;            /* the resident factors come from the mixed path */
  int fp64_fallback;    /* the mesh was redone in FP64 after a non-converged cell */
  int max_sweeps;       /* largest number of FP64 refinement sweeps of a cell */
  double mean_sweeps;
  int64_t unconverged;  /* cells whose last correction exceeded 1e-10 of W */
} hdg_refinement_stats;
int hdg_mesh_refinement_stats(hdg_mesh* mesh, hdg_refinement_stats* stats);

/* The two halves of hdg_condense_assemble, enqueued asynchronously on the
 * context stream with no host synchronisation (for timing and pipelining):
 * hdg_condense runs the fused cell integration + condensation of every cell
 * (StaticCondensationMap, src/StaticCondensationMap.jl:152-196, over
 * lazy_map at src/HybridAffineFEOperators.jl:31); hdg_assemble the Dirichlet
 * lifting and skeleton assembly (src/HybridAffineFEOperators.jl:38-46).
 * hdg_mesh_singular_count synchronises and returns the number of cells whose
 * last condensation met a zero pivot. */
int hdg_condense(hdg_mesh* mesh, int precision);
int hdg_assemble(hdg_mesh* mesh);
int hdg_mesh_singular_count(hdg_mesh* mesh, int64_t* n_singular);

/* Interpolates the built-in manufactured trace at the nodes of the free trace
 * dofs, host [nfree] (Gridap's interpolate on the skeleton space). */
int hdg_mesh_interpolate_trace(hdg_mesh* mesh, double* lam_free);

/* Device pointers of the resident skeleton system (valid until the next
 * hdg_condense_assemble or destruction): colptr int64 [nfree+1], rowval int32
 * [nnz], nzval double [nnz], rhs double [nfree]. */
int hdg_skeleton_system(hdg_mesh* mesh, const int64_t** colptr, const int32_t** rowval,
                        const double** nzval, const double** rhs);
/* Device-side skeleton solve: replaces the host `x_skel = A \ b` of
 * src/HybridAffineFEOperators.jl:74 (and src/HybridLinearSolvers.jl:45) by
 * preconditioned MINRES on the resident system of hdg_skeleton_system,
 * enqueued on the context's stream.  The skeleton matrices of both weak forms
 * are symmetric (the product reads column i of the CSC as row i) but not all
 * definite; the preconditioner is the inverse of each facet's diagonal block,
 * negated where that block's diagonal is negative.  lam_free
 * (device, nfree) receives the solution from a zero start.  Stops when
 * |b - K x| <= rtol |b| or after maxit iterations (maxit <= 0: the smaller of
 * 10 nfree and 50,000).
 * *iters and *relres (either may be NULL) report the iterations and the final
 * relative residual, recomputed from K x.  HDG_ERR_STATE when the tolerance
 * was not met (lam_free holds the last iterate). */
int hdg_skeleton_solve(hdg_mesh* mesh, double rtol, int maxit, double* lam_free, int* iters,
                       double* relres);
/* Device pointers of the per-cell condensed blocks S [ncells][nb][nb], g
 * [ncells][nb] and trace dof ids [ncells][nb]. */
int hdg_cell_blocks(hdg_mesh* mesh, const double** S, const double** g,
                    const int32_t** cell_dofs);

/* Device pointers of the resident local factors kept for recovery:
 * U [ncells][ni(ni+1)/2], the upper LU factor of A_ii in LAPACK row order
 * packed by columns (column c holds rows 0..c at c(c+1)/2 + k), and
 * Y [ncells][nb+1][ni] = L^-1 P [A_ib b_i] in LAPACK row order.  The pivot
 * sequence is getrf's (largest magnitude, first position on ties).  After a
 * mixed-precision condensation *U is NULL and Y holds W = A_ii^-1 [A_ib b_i]
 * in natural interior order.  After a flux-block condensation (family 8) U
 * and Y are the factors of the reduced system of the m = ni - n_flux
 * non-flux interior unknowns, at the start of each cell's slot (the rest of
 * the slot holds the recovery's kept blocks and the inverses of the reduced
 * factors; not part of this interface).  Cell c's slots start at c * u_stride and c * y_stride
 * (hdg_mesh_factor_layout). */
int hdg_local_factors(hdg_mesh* mesh, const double** U, const double** Y);
/* Layout of the resident factors: rows m of the factored system (ni, or
 * ni - n_flux after a flux-block condensation), the kernel family of the last
 * condensation (-1 after a mixed-precision one) and the per-cell strides (in
 * doubles) of the U and Y slots.  No reference counterpart. */
typedef struct {
  int rows;
  int family;
  int64_t u_stride;
  int64_t y_stride;
} hdg_factor_layout;
int hdg_mesh_factor_layout(hdg_mesh* mesh, hdg_factor_layout* layout);
/* Kernel family (and resident workgroups per CU, 0 = the family's default)
 * of this mesh's FP64 condensations; family -1 restores the default for the
 * class.  HDG_ERR_ARG when the family cannot condense the class.  Not for
 * batch members (the batch picks per class).  No reference counterpart. */
int hdg_mesh_set_family(hdg_mesh* mesh, int family, int grid_per_cu);

/* solve!'s recovery (src/HybridAffineFEOperators.jl:89-150 with
 * BackwardStaticCondensationMap): lam_free device [nfree] (skeleton
 * solution), free_values device [nfree_total] = [q of all cells; u of all
 * cells; lam_free], Gridap's consecutive multi-field numbering. */
int hdg_recover(hdg_mesh* mesh, const double* lam_free, double* free_values);

/* Per-cell element systems (A, b) as the fused kernel generates them,
 * device [ncells][n][n] / [ncells][n] for cells [first, first+count): a
 * debugging/parity entry point; not on the hot path. */
int hdg_cell_systems(hdg_mesh* mesh, int64_t first, int64_t count, double* A, double* b);

/* ---------------------------------------------------------------------------
 * Batches of independent meshes.  The reference builds one
 * HybridAffineFEOperator per mesh (src/HybridAffineFEOperators.jl:10-50) and
 * condenses its cells one by one; a batch does the same work for many meshes
 * at once.  Meshes whose cell systems share reference tables (same physics,
 * dimension, orders and quadrature) form a class.  The per-cell state of each
 * class is one contiguous allocation, so each class is condensed by ONE kernel
 * launch over the cells of all its meshes.  The whole sequence -- per-mesh cell
 * preparation, one condensation per class, per-mesh Dirichlet lifting and
 * skeleton assembly -- is captured once into a HIP graph and replayed by every
 * hdg_batch_launch.  Each mesh's per-cell state becomes a slice of its class
 * buffers, so the per-mesh calls (hdg_recover, hdg_cell_blocks,
 * hdg_skeleton_system, ...) keep working on it.
 *
 * A mesh belongs to at most one batch and cannot be destroyed before it.  The
 * graph records the device pointers in use at creation; hdg_mesh_set_source
 * on a member makes the next launch capture a new graph. */
enum { HDG_BATCH_AUTOTUNE = 1 };

typedef struct {
  int ni, nb;          /* local system sizes of the class */
  int physics, dim;
  int nmesh;           /* member meshes */
  int64_t ncells;      /* cells over all members */
  int family;          /* condensation kernel family in use (see below) */
  int grid_per_cu;     /* resident workgroups per CU (0 = family default) */
  double tuned_ms;     /* measured condensation time of the pick (0 if untuned) */
} hdg_batch_class;
/* Kernel families: 0 one wavefront per cell (fixed sizes), 1 8x8 grid with
 * register tiles, 3 32x16 grid, 4 blocked HBM-workspace elimination,
 * 5 blocked elimination with a single-wavefront panel and FP64 MFMA trailing
 * updates, 7 16x16 grid (2, 6 and 9 are reserved), 8 flux-block Schur
 * complement: where the flux block of A_ii is C (x) M (per-cell coefficient
 * matrix C, fixed scalar mass matrix M; both weak forms here), the flux
 * unknowns are eliminated with C^-1 (x) M^-1 in two FP64 MFMA products and
 * the reduced system is condensed by one of the families above, 10 the same
 * elimination with the Darcy flux products formed from reference-cell
 * products and each cell's Jacobian cofactors (Darcy with det J > 0; the
 * default there).  With the environment variable HDG_AUTOTUNE_LOG set, the
 * autotuner prints every candidate's time to stderr. */

/* flags: HDG_BATCH_AUTOTUNE times every kernel family and grid size able to
 * condense each class on the batch's own cells and keeps the fastest;
 * otherwise each class uses the default family for its size. */
int hdg_batch_create(hdg_ctx* ctx, int nmesh, hdg_mesh* const* meshes, int flags,
                     hdg_batch** out);
int hdg_batch_destroy(hdg_batch* batch);
int hdg_batch_get_classes(const hdg_batch* batch, int max, hdg_batch_class* classes,
                          int* nclass);
/* Enqueues one replay of the batch graph (condensation + assembly of every
 * member) on the context stream, without host synchronisation. */
int hdg_batch_launch(hdg_batch* batch, int precision);
/* hdg_batch_launch, then synchronises; *n_singular (may be NULL) receives the
 * number of cells over all members with a zero pivot. */
int hdg_batch_condense_assemble(hdg_batch* batch, int precision, int64_t* n_singular);
/* hdg_recover for every member: lam_free[m] / free_values[m] are device
 * pointers for member m (host arrays of nmesh pointers). */
int hdg_batch_recover(hdg_batch* batch, const double* const* lam_free,
                      double* const* free_values);

#ifdef __cplusplus
}
#endif
#endif /* HDG_H */
