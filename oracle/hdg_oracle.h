/*
 * hdg_oracle.h -- CPU FP64 restatement of GridapHybrid's per-cell hybrid
 * pipeline.  TEST INFRASTRUCTURE ONLY: tests/, the smoke check and the CPU
 * baseline leg of the benchmark use this library as the checker.  The
 * product library (libhdg) never links or calls it.
 *
 * Layout conventions (match Julia's Matrix{Float64}): every per-cell matrix is
 * column-major, cell-major in memory ([cell][col][row]).  Cell vectors follow
 * the multi-field order (flux q, potential u, trace lambda); the flux is
 * component-major (dof = node + nnodes*comp); lambda is facet-major (local
 * facet order x facet-local dof), as pinned by test/SumFacetMapTests.jl:109.
 */
#ifndef HDG_ORACLE_H
#define HDG_ORACLE_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* space kinds */
#define ORC_SPACE_Q 0
#define ORC_SPACE_P 1

/* built-in manufactured solutions (Darcy) */
#define ORC_MMS_ZERO 0     /* f = 0, g = 0, lambda_D = 0 */
#define ORC_MMS_KAT 1      /* test/DarcyHDGTests.jl:9-14: flux (1+x,1+y,..), p = -3.14 */
#define ORC_MMS_LINEAR 2   /* potential x+y(+z), flux -kappa grad */
#define ORC_MMS_SINPROD 3  /* potential prod sin(pi x_d), flux -kappa grad */

/* physics */
#define ORC_PHYS_DARCY 0
#define ORC_PHYS_ELASTICITY 1

typedef struct {
  int phys;          /* ORC_PHYS_* */
  int D;             /* 2 or 3 */
  int space;         /* ORC_SPACE_Q or ORC_SPACE_P (cell and trace spaces) */
  int k_flux;        /* Darcy: flux order; elasticity: stress order */
  int k_pot;         /* Darcy: potential order; elasticity: displacement order */
  int k_trace;       /* trace order */
  int quad_degree;   /* Gauss degree (2(order+1) in the reference tests) */
  double tau;        /* stabilisation */
  double E, nu;      /* elasticity only */
  int mms;           /* ORC_MMS_* or elasticity MMS id */
  int projection;    /* elasticity: 1 = bulk-to-skeleton L2 projection P_m in the
                        tau terms (test/P_m.jl:4-23), 0 = identity */
} orc_problem;

/* Sizes of the local blocks. */
int orc_sizes(const orc_problem* p, int* n_flux, int* n_pot, int* n_trace_per_facet,
              int* ni, int* nb);

/* Number of quadrature points of the cell rule and of one facet rule. */
int orc_quad_sizes(const orc_problem* p, int* nq_cell, int* nq_facet);

/* Cartesian mesh topology with Gridap numbering: cells lexicographic
 * (x fastest); facets numbered by first encounter over (cell, local facet);
 * owner of a facet = first (lowest id) cell around it.
 * Outputs (caller-allocated, sized with orc_cartesian_counts):
 *   cell_facets [ncells*2D], cell_sign [ncells*2D] (+1 owner, -1 otherwise),
 *   facet_cells [nfacets*2] (second = -1 on the boundary),
 *   facet_lpos  [nfacets*2] (local facet index in each cell),
 *   facet_boundary [nfacets]. */
int orc_cartesian_counts(int D, const int64_t* n, int64_t* ncells, int64_t* nfacets);
int orc_cartesian_topology(int D, const int64_t* n, int32_t* cell_facets, int8_t* cell_sign,
                           int32_t* facet_cells, int8_t* facet_lpos, int8_t* facet_boundary);

/* Affine geometry of Cartesian cells: x0[ncells*D], J[ncells*D*D] (col-major). */
int orc_cartesian_geometry(int D, const int64_t* n, const double* domain, double* x0, double* J);

/* Cell system (A, b) of the hybridizable weak form for ncells cells.
 * A: [ncells][(ni+nb)^2] column-major, b: [ncells][ni+nb].
 * coef: per-cell coefficient (kappa for Darcy, nu for elasticity).
 * sign: per-cell owner signs [ncells*2D] (n . n_o).
 * src (optional, may be NULL): per-cell source values at cell quadrature
 * points, [ncells][nq_cell][ncomp] (Darcy: f_1..f_D, g; elasticity: f_1..f_D);
 * when NULL the built-in MMS p->mms is used. */
int orc_cell_systems(const orc_problem* p, int64_t ncells, const double* x0, const double* J,
                     const double* coef, const int8_t* sign, const double* src,
                     double* A, double* b);

/* Per-field reference data: node coordinates of the trace space on the
 * reference facet [n_trace_per_facet*(D-1)]. */
int orc_trace_nodes(const orc_problem* p, double* nodes);

/* StaticCondensationMap (src/StaticCondensationMap.jl:152-196), op order
 * getrf -> getrs(A12) -> gemm -> getrs(b1) -> gemv, per cell.
 * A, b are the dense cell systems with interior dofs first.  Outputs S
 * [ncells][nb*nb] col-major, g [ncells][nb], info [ncells]. */
int orc_condense(int ni, int nb, int64_t ncells, const double* A, const double* b,
                 double* S, double* g, int32_t* info);

/* BackwardStaticCondensationMap (src/BackwardStaticCondensationMap.jl:61-102):
 * b1 -= A12 x (gemv), getrf(A11) again, getrs(b1). xb [ncells][nb],
 * xi [ncells][ni]. */
int orc_recover(int ni, int nb, int64_t ncells, const double* A, const double* b,
                const double* xb, double* xi, int32_t* info);

/* Cell-wise trace dof ids (Gridap convention: free >= 0 (0-based), Dirichlet
 * negative -1,-2,... in facet order).  facet_dirichlet[nfacets] marks
 * Dirichlet facets.  Output cell_dofs [ncells*nb]; returns number of free
 * dofs via *nfree and Dirichlet dofs via *ndir. */
int orc_trace_dofs(int D, int n_tf, int64_t ncells, int64_t nfacets, const int32_t* cell_facets,
                   const int8_t* facet_dirichlet, int32_t* cell_dofs, int64_t* nfree,
                   int64_t* ndir);

/* Skeleton assembly (Gridap SparseMatrixAssembler + _attach_dirichlet,
 * src/HybridAffineFEOperators.jl:38-46): CSC with rows sorted per column,
 * duplicates summed in ascending cell order.  Two-phase: count nnz, then
 * fill.  lam_dir[ndir] holds Dirichlet values. */
int64_t orc_assemble_nnz(int nb, int64_t ncells, int64_t nfree, const int32_t* cell_dofs);
int orc_assemble(int nb, int64_t ncells, int64_t nfree, const int32_t* cell_dofs,
                 const double* S, const double* g, const double* lam_dir,
                 int64_t* colptr, int32_t* rowval, double* nzval, double* rhs);

/* Cell-wise lambda (free from lam_free, Dirichlet from lam_dir). */
int orc_gather_lambda(int nb, int64_t ncells, const int32_t* cell_dofs, const double* lam_free,
                      const double* lam_dir, double* xb);

/* Dirichlet values of the trace: for each Dirichlet facet (in facet order),
 * the MMS trace evaluated at the facet nodes.  facet geometry is taken from
 * the owner cell (facet_cells[2f], facet_lpos[2f]). */
int orc_dirichlet_values(const orc_problem* p, int64_t nfacets, const int32_t* facet_cells,
                         const int8_t* facet_lpos, const int8_t* facet_dirichlet,
                         const double* x0, const double* J, double* lam_dir);

/* L2 errors of the recovered fields against the built-in MMS:
 * err[0] = ||q - q_h||, err[1] = ||u - u_h|| (Darcy).  coef = kappa. */
int orc_l2_errors(const orc_problem* p, int64_t ncells, const double* x0, const double* J,
                  const double* coef, const double* xi, double* err);

/* Dense LU solve (partial pivoting) of n x n col-major system with nrhs
 * right-hand sides, in place.  Returns info. */
int orc_dense_solve(int n, int nrhs, double* A, double* B);

/* OpenMP threads of the following calls (n <= 0 keeps the current count);
 * returns the previous count. */
int orc_set_threads(int n);

/* CPU-baseline variant that keeps the factors instead of re-factorising in
 * the recovery: per cell getrf, W = A11^-1 A12, w = A11^-1 b1, S, g as in
 * orc_condense, then x_i = w - W xb. */
int orc_condense_recover_stored(int ni, int nb, int64_t ncells, const double* A, const double* b,
                                const double* xb, double* S, double* g, double* xi,
                                int32_t* info);

/* Facet orientation (pindex, src/SkeletonArrays.jl:28-37): code p of the map
 * from a cell's local facet coordinates (s[,t]) to the facet's own ones.
 * 2D: 0 identity, 1 s' = 1-s.  3D: (a,b) = (t,s) if p&1 else (s,t);
 * s' = 1-a if p&2 else a; t' = 1-b if p&4 else b.
 * orc_facet_node_perm: perm[i] = own index of the cell-local trace node i
 * (-2 if the trace node set is not invariant under the map). */
int orc_facet_node_perm(const orc_problem* p, int code, int* perm);
/* orc_trace_dofs with per-(cell, local facet) orientation codes
 * cell_facet_perm [ncells*2D] (NULL = all identity). */
int orc_trace_dofs_oriented(const orc_problem* p, int64_t ncells, int64_t nfacets,
                            const int32_t* cell_facets, const int8_t* cell_facet_perm,
                            const int8_t* facet_dirichlet, int32_t* cell_dofs, int64_t* nfree,
                            int64_t* ndir);
/* Nodes of the scalar Lagrange space of order k on the reference d-cube
 * (nodes[i*d + a]); returns their number. */
int orc_space_nodes(int d, int k, int pspace, double* nodes);
/* orc_dirichlet_values with each facet's values in its own node order. */
int orc_dirichlet_values_oriented(const orc_problem* p, int64_t nfacets,
                                  const int32_t* facet_cells, const int8_t* facet_lpos,
                                  const int8_t* cell_facet_perm, const int8_t* facet_dirichlet,
                                  const double* x0, const double* J, double* lam_dir);

#ifdef __cplusplus
}
#endif
#endif
