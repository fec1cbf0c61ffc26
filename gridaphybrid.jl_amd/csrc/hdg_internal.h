// hdg_internal.h -- structures shared by the host runtime (hdg_api.cpp,
// hdg_fe.cpp) and the gfx950 kernels (hdg_kernels.hip).
#pragma once
#include <stdint.h>

#include <string>
#include <vector>

namespace hdg {

constexpr int kMaxD = 3;
constexpr int kMaxFacets = 6;

enum Physics { kDarcy = 0, kElasticity = 1 };

// Reference tables of one (physics, D, space, orders) class.  Scalar bases:
//   p = flux (Darcy) / stress (elasticity) scalar basis, np functions
//   q = potential (Darcy) / displacement (elasticity) scalar basis, nq
//   m = trace scalar basis on the reference facet, nm
// Exact Gauss integrals on the reference cell [0,1]^D / reference facets;
// every matrix is stored row-major [row][col].
struct TablesHost {
  int phys = 0, D = 2, nF = 4, np = 0, nq = 0, nm = 0;
  int ncp = 0, ncq = 0, ncm = 0;  // field components (Darcy: D,1,1; elasticity: nsym,D,D)
  int na = 0, nbf = 0, ntf = 0, ni = 0, nb = 0, n = 0;
  int nqc = 0;                    // cell quadrature points
  std::vector<double> Mpp;        // [np][np]
  std::vector<double> Gpq;        // [D][np][nq]  int d_e p_i q_j
  std::vector<double> Gqp;        // [D][nq][np]  int d_e q_i p_j
  std::vector<double> Fqp;        // [nF][nq][np] int_F q_i p_j
  std::vector<double> Fqq;        // [nF][nq][nq]
  std::vector<double> Fpm;        // [nF][np][nm]
  std::vector<double> Fqm;        // [nF][nq][nm]
  std::vector<double> Fmm;        // [nm][nm]
  std::vector<uint8_t> pmask;     // [np] bit F: p_i has a nonzero trace on facet F
  std::vector<uint8_t> qmask;     // [nq]
  std::vector<double> Bp;         // [nqc][np] p_i(xi_q) * w_q
  std::vector<double> Bq;         // [nqc][nq]
  std::vector<double> xq;         // [nqc][3]  reference quadrature points
  std::vector<double> wq;         // [nqc]     their weights
  std::vector<double> trace_nodes;  // [nm][D-1] trace nodes on the reference facet
  // entry program (column-major entries of the n x n matrix)
  int nalpha = 0;
  std::vector<int32_t> prog_off;  // [n*n+1]
  std::vector<uint8_t> prog_id;   // [nterms] coefficient index
  std::vector<double> prog_val;   // [nterms] reference-integral value
  // the same program with the first term of each entry inline (one load of
  // head and one of val0 per entry; most entries have at most one term):
  // head = id0 | count << 8 | (offset of the further terms in x*) << 12
  std::vector<uint32_t> prog_head;  // [n*n]
  std::vector<double> prog_val0;    // [n*n]
  std::vector<uint8_t> prog_xid;    // further terms
  std::vector<double> prog_xval;
  // the interior rows of the program as a row-major sparse matrix of terms
  // (row r: terms rptr[r]..rptr[r+1] with column, coefficient index, value)
  std::vector<int32_t> prog_rptr;   // [ni+1]
  std::vector<uint16_t> prog_rcol;
  std::vector<uint8_t> prog_rid;
  std::vector<double> prog_rval;
  std::vector<int32_t> prog_rbsplit;  // [ni] first term of row r with a boundary column
  // Leading flux block of A_ii with Kronecker structure C (x) M: entry
  // ((a,i),(b,j)) = C[a][b] M[i][j] with C[a][b] = alpha[fb_cid[6a+b]] per
  // cell (fb_cid < 0: identically zero) and M the np x np reference mass
  // matrix of the scalar flux basis.  fb_nc = 0 when the program has no such
  // block.  fb_minv = M^-1, row-major.
  int fb_nc = 0;
  int8_t fb_cid[36] = {};
  std::vector<double> fb_minv;
  // Transpose symmetry of the flux coupling (0: none): bit 0 set when
  // A_uq = s_u A_qu^T and A_bq = s_b A_qb^T entry by entry, bit 1: s_u = -1,
  // bit 2: s_b = -1.
  int fb_sym = 0;
  // Reference decomposition of the flux Schur complement for Darcy on affine
  // cells (build_flux_ref): with c_ea = detJ Ji(e, a) the cofactor entries of
  // the cell, [A_qu A_qb] row block of component a = sum_e c_ea G_e, so
  //   [A_qu A_qb]^T (C^-1 (x) M^-1) [A_qu A_qb] = (1 / alpha_0) sum_p w_p P_p,
  // w over the pairs (e, e') of sum_a c_ea c_e'a and P_(e,e) = G_e^T M^-1 G_e,
  // P_(e,e') = G_e^T M^-1 G_e' + its transpose.  fr_npair = 0 when it does not
  // apply (not Darcy, or no symmetric scalar flux block).
  int fr_npair = 0;
  std::vector<double> fr_P;  // [npair][nr][nr] row-major, nr = nq + nb
  std::vector<double> fr_G;  // [D][np][nr]  G_e
};

// Device view of the tables (pointers into one device allocation).
struct Tables {
  int phys, D, nF, np, nq, nm;
  int ncp, ncq, ncm;
  int na, nbf, ntf, ni, nb, n;
  int nqc;
  const double* Mpp;
  const double* Gpq;
  const double* Gqp;
  const double* Fqp;
  const double* Fqq;
  const double* Fpm;
  const double* Fqm;
  const double* Fmm;
  const uint8_t* pmask;
  const uint8_t* qmask;
  const double* Bp;
  const double* Bq;
  const double* xq;
  int nalpha;
  const int32_t* prog_off;
  const uint8_t* prog_id;
  const double* prog_val;
  const uint32_t* prog_head;
  const double* prog_val0;
  const uint8_t* prog_xid;
  const double* prog_xval;
  const int32_t* prog_rptr;
  const uint16_t* prog_rcol;
  const uint8_t* prog_rid;
  const double* prog_rval;
  const int32_t* prog_rbsplit;
  int fb_nc;              // flux-block components (0: no Kronecker flux block)
  int8_t fb_cid[36];      // coefficient id of C[a][b] (< 0: zero)
  const double* fb_minv;  // [np][np] inverse of the scalar flux mass matrix
  int fb_sym;             // transpose symmetry of the flux coupling (TablesHost::fb_sym)
  int fr_npair;           // reference flux Schur decomposition (TablesHost::fr_*), 0: none
  const double* fr_P;     // [npair][nr][nr]
  const double* fr_G;     // [D][np][nr]
};

// Device view of a mesh for the fused element generator.
struct MeshDev {
  int64_t ncells;
  const double* x0;        // [ncells][D]
  const double* jac;       // [ncells][D*D] column-major
  const int8_t* sign;      // [ncells][2D]  n . n_o
  const double* coef;      // [ncells] kappa or nu
  const double* src;       // [ncells][nqc][ncomp] or null (user source values)
  double* alpha;           // [ncells][72] per-cell entry-program coefficients (scratch)
  double* srcq;            // [ncells][ni] interior right-hand side b_i (scratch)
  int source;              // hdg_source
  double tau;
  double young;
};

// Outputs of the condensation kernels (any pointer may be null).
struct CondenseOut {
  double* S = nullptr;     // [ncells][nb][nb] column-major
  double* g = nullptr;     // [ncells][nb]
  double* U = nullptr;     // [ncells][ni(ni+1)/2] packed upper factor, LAPACK row order
  double* Y = nullptr;     // [ncells][nb+1][ni] L^-1 P [A_ib b_i], LAPACK row order
  double* X = nullptr;     // [ncells][ni] back-substituted rhs (backward map)
  int32_t* info = nullptr; // [ncells]
  float* Xinv = nullptr;   // [ncells][ni][ni] FP32 inverse of A_ii (mixed path, kept)
  // per-cell strides of U and Y (0: packed for this ni, i.e. ni(ni+1)/2 and
  // (nb+1) ni); the flux-block path stores the reduced factors in the
  // per-cell slots of the full interior so every family shares one layout
  int64_t u_stride = 0, y_stride = 0;
  // (grid elimination kernels) the multipliers and the pivot order, for
  // solves with other right-hand sides: L[cell * u_stride + j * ni + r] =
  // multiplier of interior row r at step j (rows pivoted later than j),
  // P[cell * u_stride + k] = the row at LAPACK position k
  double* L = nullptr;
  double* P = nullptr;
};

// Dense cell systems in HBM for the map-level API.
struct DenseIn {
  const double* A;       // [ncells][n][n] column-major, natural field order
  const double* b;       // [ncells][n]
  const int32_t* perm;   // [n] condensed index -> natural index (device)
  int n;
  int ni, nb;
};


// Facet-pair structure of the skeleton matrix (see the assembly kernels).
constexpr uint8_t kGlposDirichlet = 0x80;
struct PairTable {
  const int64_t* fptr;     // [nfree_facets+1] first nnz of facet G's columns
  const int32_t* pptr;     // [nfree_facets+1] first pair of facet G
  const int32_t* pfacet;   // [npairs] free facet id of F
  const int32_t* pcell;    // [npairs][2] contributing cells (-1 if absent)
  const uint8_t* plpos;    // [npairs][2] (lpos_G | lpos_F << 4)
  const int32_t* gcell;    // [nfree_facets][2] cells around G (-1 if absent)
  const uint8_t* glpos;    // [nfree_facets][2] local position of G in them (bits 0-3),
                           // | kGlposDirichlet when that cell has a Dirichlet trace dof
  // facet orientation (nullptr when every (cell, local facet) has code 0):
  const uint8_t* pperm;    // [npairs][2] codes (of G | of F << 4) in the contributing cell
  const uint8_t* gperm;    // [nfree_facets][2] code of G in the cells around it
  const uint8_t* iperm;    // [codes][ntf] own trace dof -> cell-local trace dof
  int64_t nff;             // free facets
  int ntf, nb;
};

}  // namespace hdg

#include <hip/hip_runtime_api.h>
namespace hdg {
// Device workspace of the blocked large-system kernel.  Owned by a context or
// a batch; it only grows, and only outside stream capture (a captured graph
// keeps the address it saw).
struct Workspace {
  double* p = nullptr;
  size_t bytes = 0;
  int pool = -1;  // the block pool it came from
};
hipError_t workspace_reserve(Workspace& w, size_t bytes);
void workspace_release(Workspace& w);

// Device memory of the library's buffers and workspaces.  hipFree waits for
// the whole device (a serving loop frees and allocates every round while
// another round runs), so each context has a pool of freed blocks: a block
// freed by the context's objects serves the context's later allocations of at
// most its size and at least half of it.  The context's buffers are used only
// by work on its one stream, so the reuse is ordered behind their last use.
// The pool in effect is the innermost PoolScope of the calling thread (none:
// plain hipMalloc / hipFree); a buffer returns to the pool it came from.
// dev_pool_idle, called once the context's stream has been synchronized,
// hands the pool's blocks to the device's idle list, which serves any
// context (no queued work uses those blocks any more).  The device's pools
// and idle list are emptied (hipFree) when an allocation would leave less
// than 8 GB of device memory free, and before retrying a failed hipMalloc.
struct PoolScope {
  explicit PoolScope(int pool);
  ~PoolScope();
  PoolScope(const PoolScope&) = delete;
  PoolScope& operator=(const PoolScope&) = delete;
  int prev;
};
int dev_pool_create(int device);
void dev_pool_destroy(int pool);
void dev_pool_idle(int pool);
hipError_t dev_alloc(void** p, size_t bytes, int* pool);
void dev_free(void* p, size_t bytes, int pool);
void dev_cache_flush(int device);
void dev_pool_stats(int pool, int64_t* live, int64_t* pooled, int64_t* unpooled);

// Kernel launchers (hdg_kernels.hip)
bool condense_supported(int n);
// fam: kernel family (-1 = default for the size; see condense_variants);
// gridmul: resident workgroups per CU (0 = the family's default)
hipError_t launch_condense_fused(const Tables& T, const MeshDev& M, const CondenseOut& out,
                                 hipStream_t st, Workspace& ws, bool prep = true, int fam = -1,
                                 int gridmul = 0);
hipError_t launch_prep_cells(const Tables& T, const MeshDev& M, hipStream_t st);
// Mixed precision (FP32 Gauss-Jordan inverse of A_ii, FP64 refinement of
// W = A_ii^-1 [A_ib b_i], FP64 S and g).  out.Y receives W (natural interior
// order, [nb+1][ni] per cell); out.U is not written.  iters (may be null) gets
// the refinement sweeps per cell, negative when not converged.
constexpr int kMaxRefinementSweeps = 10;
constexpr int kMixedMinNi = 64;  // smaller interiors: the FP64 path is faster
bool mixed_supported(int ni, int nb);
size_t mixed_workspace_bytes(int ni, int nb, int64_t ncells);
hipError_t launch_condense_mixed(const Tables& T, const MeshDev& M, const CondenseOut& out,
                                 hipStream_t st, Workspace& ws, bool prep, int maxit,
                                 int32_t* iters);
// What the kernel families need to know of a class: the local sizes and
// whether the Kronecker flux-block elimination applies (flux_na = n_flux).
struct SysShape {
  int ni = 0, nb = 0;
  int flux_na = 0;
  bool flux = false;
  bool ref = false;    // the reference decomposition applies (family 10)
};
SysShape sys_shape(const Tables& T);
SysShape sys_shape(const TablesHost& t);
// Doubles per cell of the factor slot ("U"): ni(ni+1)/2 for the elimination
// families; for classes where the flux-block path applies, room for its
// reduced factors and the interior blocks its recovery refines with:
//   [U m(m+1)/2][L m*m][P m][A_uu m*m][A_qu na*m][A_uq m*na],  m = ni - na.
size_t factor_slot(const SysShape& sh);
// Family id of the flux-block Schur complement: the flux unknowns are
// eliminated through C^-1 (x) M^-1 and the reduced system of the remaining
// ni - n_flux interior unknowns is condensed by an elimination kernel.  The
// resident factors are then those of the reduced system.
constexpr int kFamilyFluxSchur = 8;
// (family 9, a fused variant of family 8 that never won a class, is removed;
// its number stays reserved)
// Family 10: the flux elimination of family 8 with the reduced system formed
// from the reference decomposition (TablesHost::fr_*) instead of two MFMA
// products per cell (Darcy on cells with det J > 0).
constexpr int kFamilyFluxRef = 10;
// kernel families able to condense a class; returns their count
int condense_variants(const SysShape& sh, int* fams, int max);
int condense_default_family(const SysShape& sh);
// grid sizes (resident workgroups per CU) worth timing for a family
int condense_grid_options(int fam, int* gm, int max);
size_t condense_workspace_bytes(const SysShape& sh, int fam, int64_t ncells);
hipError_t launch_condense_flux(const Tables& T, const MeshDev& M, const CondenseOut& out,
                                hipStream_t st, Workspace& ws, int gridmul, bool ref);
// Recovery after a flux-block condensation (families 8 and 10): x_u =
// U^-1 (y - Y lambda) from the reduced factors, x_q = (C^-1 (x) M^-1)(b_q -
// A_qu x_u - A_qb lambda); then
// FP64 refinement against the residual of the full interior system (the
// reduced system is worse conditioned than A_ii).
hipError_t launch_recover_flux(const Tables& T, const MeshDev& M, const double* U,
                               const double* Y, int64_t usz, int64_t ysz, bool ref,
                               const int32_t* cell_dofs, const double* lam_free,
                               const double* lam_dir, double* out, int64_t nfree,
                               hipStream_t st);
constexpr int kFluxRefineSteps = 2;  // FP64 refinement steps of the flux recovery (at most)
hipError_t launch_condense_dense(const DenseIn& in, int64_t ncells, const CondenseOut& out,
                                 hipStream_t st, Workspace& ws);
hipError_t launch_backward_dense(const DenseIn& in, const double* x, int64_t ncells, double* xi,
                                 int32_t* info, hipStream_t st, Workspace& ws);
hipError_t launch_recover(const double* U, const double* Y, const int32_t* cell_dofs,
                          const double* lam_free, const double* lam_dir, double* out,
                          int64_t ncells, int ni, int nb, int na, int64_t nfree, hipStream_t st,
                          int64_t u_stride = 0, int64_t y_stride = 0);
// Recovery after a mixed-precision condensation: x0 = w - W lambda, then one
// FP64 refinement x = x0 + X (b_i - A_ib lambda - A_ii x0) with the interior
// rows of the cell system applied term by term from the per-cell
// coefficients (the residual of the combined right-hand side keeps the
// cancellation error at the level of the direct solve) and X the kept FP32
// inverse of A_ii.
hipError_t launch_recover_mixed(const Tables& T, const MeshDev& M, const double* W,
                                const float* Xinv, const int32_t* cell_dofs,
                                const double* lam_free, const double* lam_dir, double* out,
                                int64_t nfree, hipStream_t st);
hipError_t launch_symbolic(const PairTable& P, int32_t* rowval, hipStream_t st);
// device skeleton solve (symmetric CSC read as CSR; y = scale K x)
hipError_t launch_spmv_sym(const int64_t* colptr, const int32_t* rowval, const double* nzval,
                           const double* x, double* y, int64_t n, double scale, hipStream_t st);
// SPD preconditioner blocks [nblk][bs][bs]: (s B)^-1 for each diagonal block
// B with s B positive definite (s the sign of its diagonal), else
// diag(1 / |B_ii|)
hipError_t launch_block_inv(const int64_t* colptr, const int32_t* rowval, const double* nzval,
                            double* binv, int64_t nblk, int bs, hipStream_t st);
hipError_t launch_block_apply(const double* binv, const double* r, double* z, int64_t n, int bs,
                              hipStream_t st);
hipError_t launch_krylov_update(int op, int64_t n, double a, double b, double c, double d,
                                double* u0, double* u1, const double* w0, const double* w1,
                                const double* w2, hipStream_t st);
int block_jacobi_max();
// MINRES scalar state kept on the device (indices into a double array)
struct MinresScalar {
  enum {
    kBeta1, kOldb, kBeta, kDbar, kEpsln, kPhibar, kCs, kSn, kAlfa, kInvBeta, kBetaOldb,
    kAlfaBeta, kOldEps, kDelta, kInvGamma, kPhi, kFrozen, kIndefinite, kCount
  };
};
hipError_t launch_krylov_update_dev(int op, int64_t n, const double* sc, int ia, int ib, int ic,
                                    int id, double* u0, double* u1, const double* w0,
                                    const double* w1, const double* w2, hipStream_t st);
// step 0: init from dot[0] = (r1, M^-1 r1); 1: alfa = dot[0]; 2: beta^2 = dot[0]
hipError_t launch_minres_scalar(int step, const double* dot, double* sc, hipStream_t st);
size_t multidot_scratch_doubles();
// nd (1..3) dot products (a_k, b_k); the sums land in scratch[3 * 1024 + k]
hipError_t launch_multidot(int nd, int64_t n, const double* a0, const double* b0,
                           const double* a1, const double* b1, const double* a2,
                           const double* b2, double* scratch, hipStream_t st);
hipError_t launch_gather(const double* src, const int64_t* idx, double* dst, int64_t n,
                         hipStream_t st);
hipError_t launch_copy(const double* src, double* dst, int64_t n, hipStream_t st);
hipError_t launch_assemble(const PairTable& P, const double* S, const double* g,
                           const int32_t* cell_dofs, const double* lam_dir, double* nzval,
                           double* rhs, hipStream_t st);
hipError_t launch_cell_systems(const Tables& T, const MeshDev& M, int64_t first, int64_t count,
                               double* A, double* b, hipStream_t st);
// Allocation-free condensation (hdg_static_condensation_noalloc): a fixed
// device scratch array per device holds the field permutation and the
// singular-cell counter.  scratch_init (the device current) resolves its
// address once per device; scratch_address returns it (nullptr before).
constexpr int kNoAllocMaxN = 175;            // rows of the largest cell system
constexpr size_t kScratchPermOffset = 0;     // int32 [kNoAllocMaxN]
constexpr size_t kScratchCountOffset = 1024; // uint64 singular-cell counter
constexpr size_t kScratchBytes = 1280;
hipError_t scratch_init(int device);
void* scratch_address(int device);
// the kernel family that condenses (ni, nb) without a workspace, or -1
int noalloc_family(int ni, int nb);
hipError_t launch_condense_dense_family(const DenseIn& in, int64_t ncells, const CondenseOut& out,
                                        hipStream_t st, int family);
hipError_t launch_count_nonzero(const int32_t* info, int64_t n, int64_t* count_dev,
                                hipStream_t st);
hipError_t launch_count_negative(const int32_t* v, int64_t n, int64_t* count_dev,
                                 hipStream_t st);

// Kernel clock (hdg_kernels.hip): per-name totals of event-timed launches.
struct KernelTime {
  std::string name;
  int64_t launches;
  double total_ms;
};
void clock_start();
int clock_stop(std::vector<KernelTime>& out);  // returns the number of failed timings

// Host helpers (hdg_fe.cpp)
void build_flux_ref(TablesHost& t);
// proj (elasticity): the tau (w, P_m u) facet block is tau (P_m w, P_m u), i.e.
// Fqq[F] = Fqm[F] Fmm^-1 Fqm[F]^T (every facet integral scales with the same area)
int build_tables(int phys, int D, int space, int kp, int kq, int km, int quad_degree, int proj,
                 TablesHost& t, std::string& err);

}  // namespace hdg
