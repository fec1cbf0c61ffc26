// hdg_kernels.hip -- gfx950 kernels of the HDG condense / assemble / recover
// path.
//
// Condensation ("augmented elimination"): one workgroup per cell.  The cell
// system [A_ii A_ib b_i; A_bi A_bb b_b] is generated directly into registers,
// distributed 2D-cyclically over a PR x PC thread grid (row r -> thread row
// r % PR, column c -> thread column c % PC).  The first ni columns are
// eliminated with partial pivoting restricted to interior rows (the pivot
// sequence of LAPACK getf2, tracked without moving rows), so the boundary
// rows end as S = A_bb - A_bi A_ii^-1 A_ib and g = b_b - A_bi A_ii^-1 b_i and
// a back substitution over the interior rows yields W = A_ii^-1 [A_ib b_i].
// Per elimination step only the pivot row and the multiplier column travel
// through LDS; pivot search is a shuffle reduction inside one wavefront.
// Reference op sequence replaced: src/StaticCondensationMap.jl:177-195.
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <algorithm>
#include <atomic>
#include <climits>
#include <map>
#include <mutex>
#include <vector>

#include "hdg_internal.h"

namespace hdg {

// --------------------------------------------------------------------------
// Per-cell geometry
// --------------------------------------------------------------------------
struct CellGeo {
  double detJ;          // |det J|
  double Ji[9];         // J^-1, column-major D x D
  double aF[kMaxFacets];     // facet measure factor |det J| |J^-T n_ref|
  double nF[kMaxFacets][3];  // outward unit normal
  double sF[kMaxFacets];     // owner sign n . n_o
  double coef;          // kappa (Darcy) or nu (elasticity)
  double x0[3];
  double J[9];
  double K[6][6];       // elasticity: v_a : A(sigma_b) per unit mass
};

// det J of an affine cell (J column-major D x D in the first D*D entries)
__device__ inline double jac_det(const double* J, int D) {
  if (D == 2) return J[0] * J[3] - J[2] * J[1];
#define JM(r, c) J[(r) + 3 * (c)]
  return JM(0, 0) * (JM(1, 1) * JM(2, 2) - JM(2, 1) * JM(1, 2)) -
         JM(0, 1) * (JM(1, 0) * JM(2, 2) - JM(2, 0) * JM(1, 2)) +
         JM(0, 2) * (JM(1, 0) * JM(2, 1) - JM(2, 0) * JM(1, 1));
#undef JM
}

__device__ inline void cell_geometry(const MeshDev& m, int D, int64_t cell, CellGeo& g) {
  // fixed-bound loops with guards keep every local array in registers
  double J[9], Ji[9];
#pragma unroll
  for (int t = 0; t < 9; ++t) J[t] = t < D * D ? m.jac[cell * D * D + t] : 0.0;
#pragma unroll
  for (int a = 0; a < 3; ++a) g.x0[a] = a < D ? m.x0[cell * D + a] : 0.0;
#pragma unroll
  for (int t = 0; t < 9; ++t) g.J[t] = J[t];
  const double det = jac_det(J, D);
  if (D == 2) {
    Ji[0] = J[3] / det;
    Ji[3] = J[0] / det;
    Ji[2] = -J[2] / det;
    Ji[1] = -J[1] / det;
#pragma unroll
    for (int t = 4; t < 9; ++t) Ji[t] = 0.0;
  } else {
#define JM(r, c) J[(r) + 3 * (c)]
    Ji[0 + 3 * 0] = (JM(1, 1) * JM(2, 2) - JM(1, 2) * JM(2, 1)) / det;
    Ji[0 + 3 * 1] = (JM(0, 2) * JM(2, 1) - JM(0, 1) * JM(2, 2)) / det;
    Ji[0 + 3 * 2] = (JM(0, 1) * JM(1, 2) - JM(0, 2) * JM(1, 1)) / det;
    Ji[1 + 3 * 0] = (JM(1, 2) * JM(2, 0) - JM(1, 0) * JM(2, 2)) / det;
    Ji[1 + 3 * 1] = (JM(0, 0) * JM(2, 2) - JM(0, 2) * JM(2, 0)) / det;
    Ji[1 + 3 * 2] = (JM(0, 2) * JM(1, 0) - JM(0, 0) * JM(1, 2)) / det;
    Ji[2 + 3 * 0] = (JM(1, 0) * JM(2, 1) - JM(1, 1) * JM(2, 0)) / det;
    Ji[2 + 3 * 1] = (JM(0, 1) * JM(2, 0) - JM(0, 0) * JM(2, 1)) / det;
    Ji[2 + 3 * 2] = (JM(0, 0) * JM(1, 1) - JM(0, 1) * JM(1, 0)) / det;
#undef JM
  }
  g.detJ = fabs(det);
#pragma unroll
  for (int t = 0; t < 9; ++t) g.Ji[t] = Ji[t];
  // normals and facet measures read J^-1 back from LDS (dynamic indices are
  // fine there, unlike in a register array)
  for (int f = 0; f < 2 * D; ++f) {
    const int fix = D - 1 - f / 2;
    const double sgn = (f & 1) ? 1.0 : -1.0;
    double nv = 0.0;
    for (int a = 0; a < D; ++a) {
      const double t = g.Ji[fix + D * a] * sgn;  // (J^-T n_ref)_a
      g.nF[f][a] = t;
      nv += t * t;
    }
    nv = sqrt(nv);
    for (int a = 0; a < 3; ++a) g.nF[f][a] = a < D ? g.nF[f][a] / nv : 0.0;
    g.aF[f] = g.detJ * nv;
    g.sF[f] = (double)m.sign[cell * 2 * D + f];
  }
  g.coef = m.coef[cell];
}

// --------------------------------------------------------------------------
// Manufactured sources (device side of hdg_source)
// --------------------------------------------------------------------------
__device__ inline void darcy_source(int src, int D, const double* x, double kappa, double* f,
                                    double* gsrc) {
#pragma unroll
  for (int a = 0; a < 3; ++a) f[a] = 0.0;
  *gsrc = 0.0;
  if (src == 1) {  // KAT: flux (1+x_a), potential const
#pragma unroll
    for (int a = 0; a < 3; ++a) f[a] = a < D ? 1.0 + x[a] : 0.0;
    *gsrc = (double)D;
  } else if (src == 3) {  // prod sin(pi x_a), q = -kappa grad u
    double p = 1.0;
#pragma unroll
    for (int a = 0; a < 3; ++a)
      if (a < D) p *= sin(M_PI * x[a]);
    *gsrc = kappa * D * M_PI * M_PI * p;
  }
}

// x[i] for a run-time i < 3, by selects (a dynamic index would put x in
// scratch memory)
__device__ inline double sel3(const double* x, int i) {
  return i == 0 ? x[0] : (i == 1 ? x[1] : x[2]);
}

__device__ inline double elas_h(int D, const double* x, int i, int j, int k) {
  // d2 u_i / dx_j dx_k of u_i = sin(pi (x_i + x_{i+1}/2))
  const int i1 = (i + 1) % D;
  const double dj = (j == i ? 1.0 : 0.0) + (j == i1 ? 0.5 : 0.0);
  const double dk = (k == i ? 1.0 : 0.0) + (k == i1 ? 0.5 : 0.0);
  return -M_PI * M_PI * sin(M_PI * (sel3(x, i) + 0.5 * sel3(x, i1))) * dj * dk;
}

__device__ inline void elas_source(int src, int D, const double* x, double E, double nu,
                                   double* f) {
  f[0] = f[1] = f[2] = 0.0;
  if (src != 12) return;  // the linear displacement has zero body load
  const double lam = E * nu / ((1.0 + nu) * (1.0 - (D - 1) * nu));
  const double mu2 = E / (1.0 + nu);
#pragma unroll
  for (int j = 0; j < 3; ++j) {
    if (j >= D) break;
    double s = 0.0;
    for (int i = 0; i < D; ++i) {
      const double deps = 0.5 * (elas_h(D, x, i, j, i) + elas_h(D, x, j, i, i));
      double dtr = 0.0;
      if (i == j)
        for (int q = 0; q < D; ++q) dtr += elas_h(D, x, q, q, i);
      s += mu2 * deps + lam * dtr;
    }
    f[j] = s;
  }
}

// --------------------------------------------------------------------------
// Element generators.  Each provides:
//   Shared               LDS state of one cell
//   begin(cell, sh)      cooperative set-up (all threads call it, followed by
//                        a barrier in the caller)
//   entry(sh, row, col)  entry (row, col) of [A | b] in condensed order
// --------------------------------------------------------------------------
constexpr int kMaxQP = 128;

__device__ inline void sym_pair(int D, int a, int& i, int& j) {
  // upper triangle, row-major: 2D (11,12,22), 3D (11,12,13,22,23,33);
  // packed as 2-bit (i, j) pairs to stay out of scratch memory
  const unsigned code = D == 2 ? 0x540u /* (0,0)(0,1)(1,1) */ : 0xA95840u;
  i = (code >> (4 * a)) & 3u;
  j = (code >> (4 * a + 2)) & 3u;
}

constexpr int kMaxAlpha = 72;

// Per-cell preparation (runs at the start of every condensation): the
// affine pull-back coefficients alpha of the entry program (layout fixed by
// build_program in hdg_fe.cpp).  One thread per cell computes its
// coefficients into LDS; the wavefront then stores its 64 cells' rows as one
// contiguous, coalesced range.
// Per-cell stride of the right-hand-side scratch: the interior b_i [ni].
__device__ __host__ inline int64_t srcq_stride(const Tables& T) {
  return (int64_t)T.ni;
}

constexpr int kPrepCells = 64;  // cells (threads) per prep workgroup

template <int PHYS>
__global__ void __launch_bounds__(kPrepCells) prep_kernel(Tables T, MeshDev M) {
  __shared__ double s_al[kPrepCells][kMaxAlpha + 1];  // odd row stride: no bank conflicts
  const int D = T.D;
  for (int64_t cell0 = (int64_t)blockIdx.x * kPrepCells; cell0 < M.ncells;
       cell0 += (int64_t)gridDim.x * kPrepCells) {
    const int64_t cell = cell0 + threadIdx.x;
    const int cnt = (int)(M.ncells - cell0 < kPrepCells ? M.ncells - cell0 : kPrepCells);
    if (cell < M.ncells) {
      CellGeo g;
      cell_geometry(M, D, cell, g);
      double* al = s_al[threadIdx.x];
      if (PHYS == kDarcy) {
        al[0] = g.detJ / g.coef;
        for (int t = 1; t < 10; ++t) {
          const int e = (t - 1) / D, a = (t - 1) % D;
          al[t] = e < D ? -g.detJ * g.Ji[e + D * a] : 0.0;
        }
        for (int t = 10; t < 28; ++t) {
          const int F = (t - 10) / 3, a = (t - 10) % 3;
          al[t] = (F < 2 * D && a < D) ? g.aF[F] * g.nF[F][a] : 0.0;
        }
        for (int t = 28; t < kMaxAlpha; ++t) {
          const int F = t - 28;
          al[t] = F < 2 * D ? M.tau * g.sF[F] * g.aF[F] : 0.0;
        }
      } else {
        const int nsym = D * (D + 1) / 2;
        const double nu = g.coef, c1 = (1.0 + nu) / M.young, c2 = nu / M.young;
        for (int t = 0; t < 36; ++t) {
          const int a = t / 6, b = t % 6;
          double v = 0.0;
          if (a < nsym && b < nsym) {
            int ia, ja, ib, jb;
            sym_pair(D, a, ia, ja);
            sym_pair(D, b, ib, jb);
            const double dd = (a == b) ? (ia == ja ? 1.0 : 2.0) : 0.0;
            const double tra = ia == ja ? 1.0 : 0.0, trb = ib == jb ? 1.0 : 0.0;
            v = (c1 * dd - c2 * tra * trb) * g.detJ;
          }
          al[t] = v;
        }
        for (int t = 36; t < 45; ++t) {
          const int e = (t - 36) / D, d = (t - 36) % D;
          al[t] = e < D ? g.detJ * g.Ji[e + D * d] : 0.0;
        }
        for (int t = 45; t < 63; ++t) {
          const int F = (t - 45) / 3, d = (t - 45) % 3;
          al[t] = (F < 2 * D && d < D) ? g.aF[F] * g.nF[F][d] : 0.0;
        }
        for (int t = 63; t < kMaxAlpha; ++t) {
          const int F = t - 63;
          al[t] = F < 2 * D ? M.tau * g.aF[F] : 0.0;
        }
      }
    }
    __syncthreads();
    double* dst = M.alpha + cell0 * kMaxAlpha;
    for (int i = threadIdx.x; i < cnt * kMaxAlpha; i += kPrepCells)
      dst[i] = s_al[i / kMaxAlpha][i % kMaxAlpha];
    __syncthreads();
  }
}

// |det J| times the source at cell quadrature point q (components [4]): the
// values given per quadrature point (source 100) or a built-in manufactured
// source at x = x0 + J xi_q.
template <int PHYS>
__device__ inline void cell_source(const Tables& T, const MeshDev& M, int64_t cell, int q,
                                   double* v) {
  const int D = T.D;
  const int ncomp = PHYS == kDarcy ? D + 1 : D;
  double J[9], x0[3];
#pragma unroll
  for (int t = 0; t < 9; ++t) J[t] = t < D * D ? M.jac[cell * D * D + t] : 0.0;
#pragma unroll
  for (int a = 0; a < 3; ++a) x0[a] = a < D ? M.x0[cell * D + a] : 0.0;
  const double detJ = fabs(jac_det(J, D));
#pragma unroll
  for (int c = 0; c < 4; ++c) v[c] = 0.0;
  if (M.source == 100) {
    const double* p = M.src + (cell * T.nqc + q) * ncomp;
#pragma unroll
    for (int c = 0; c < 4; ++c)
      if (c < ncomp) v[c] = p[c];
  } else if (M.source != 0) {
    double x[3];
#pragma unroll
    for (int a = 0; a < 3; ++a) {
      double t = x0[a];
#pragma unroll
      for (int e = 0; e < 3; ++e)  // J[a + D e], by constant indices
        if (a < D && e < D) t += (D == 2 ? J[a + 2 * e] : J[a + 3 * e]) * T.xq[3 * q + e];
      x[a] = t;
    }
    double f[3], gs = 0.0;
    if (PHYS == kDarcy) {
      darcy_source(M.source, D, x, M.coef[cell], f, &gs);
#pragma unroll
      for (int c = 0; c < 4; ++c) v[c] = c < D ? f[c < 3 ? c : 2] : (c == D ? gs : 0.0);
    } else {
      elas_source(M.source, D, x, M.young, M.coef[cell], f);
#pragma unroll
      for (int c = 0; c < 3; ++c) v[c] = f[c];
    }
  }
#pragma unroll
  for (int c = 0; c < 4; ++c) v[c] *= detJ;
}

// Per-cell strides of the factor outputs (see CondenseOut).
__device__ inline int64_t ustride(const CondenseOut& o, int ni) {
  return o.u_stride ? o.u_stride : (int64_t)ni * (ni + 1) / 2;
}
__device__ inline int64_t ystride(const CondenseOut& o, int ni, int nb) {
  return o.y_stride ? o.y_stride : (int64_t)(nb + 1) * ni;
}

constexpr int kMaxRhsRows = 256;  // interior rows of the fused generators, at most

struct FusedShared {
  double alpha[kMaxAlpha];  // per-cell coefficients of the entry program
  double b[kMaxRhsRows];    // right-hand side b_i of the interior rows (cell_rhs_kernel)
};
using CellCoefs = FusedShared;

template <int PHYS>
struct FusedGen {
  Tables T;
  MeshDev M;

  using Shared = FusedShared;

  __device__ void begin(int64_t cell, Shared& sh) const {
    for (int t = threadIdx.x; t < kMaxAlpha; t += blockDim.x)
      sh.alpha[t] = M.alpha[cell * kMaxAlpha + t];
    const double* sq = M.srcq + cell * srcq_stride(T);
    for (int t = threadIdx.x; t < T.ni; t += blockDim.x) sh.b[t] = sq[t];
  }

  __device__ double entry(const Shared& sh, int row, int col) const {
    const int n = T.n;
    if (col < n) {
      const int e = col * n + row;
      const uint32_t h = T.prog_head[e];
      const double v0 = T.prog_val0[e];
      const int cnt = (h >> 8) & 15;
      if (cnt == 0) return 0.0;
      double acc = fma(sh.alpha[h & 255], v0, 0.0);
      const int x0 = (int)(h >> 12);
      for (int k = 0; k < cnt - 1; ++k)
        acc = fma(sh.alpha[T.prog_xid[x0 + k]], T.prog_xval[x0 + k], acc);
      return acc;
    }
    return row < T.ni ? sh.b[row] : 0.0;
  }

  // U entries at once: every head, first value and first extra term is loaded
  // before any is used, so U L2 round trips overlap.  Same terms in the same
  // order as entry(); ok[u] false gives 0.
  template <int U>
  __device__ void entries(const Shared& sh, const int* row, const int* col, const bool* ok,
                          double* out) const {
    const int n = T.n;
    uint32_t h[U];
    double v0[U];
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const bool p = ok[u] && col[u] < n;
      const int e = p ? col[u] * n + row[u] : 0;
      h[u] = p ? T.prog_head[e] : 0u;
      v0[u] = p ? T.prog_val0[e] : 0.0;
    }
    int i1[U];
    double v1[U];
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const int cnt = (h[u] >> 8) & 15, x0 = (int)(h[u] >> 12);
      i1[u] = cnt >= 2 ? (int)T.prog_xid[x0] : 0;
      v1[u] = cnt >= 2 ? T.prog_xval[x0] : 0.0;
    }
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const int cnt = (h[u] >> 8) & 15, x0 = (int)(h[u] >> 12);
      double acc = cnt ? fma(sh.alpha[h[u] & 255], v0[u], 0.0) : 0.0;
      if (cnt >= 2) acc = fma(sh.alpha[i1[u]], v1[u], acc);
      for (int k = 1; k < cnt - 1; ++k)
        acc = fma(sh.alpha[T.prog_xid[x0 + k]], T.prog_xval[x0 + k], acc);
      if (ok[u] && col[u] >= n) acc = row[u] < T.ni ? sh.b[row[u]] : 0.0;
      out[u] = ok[u] ? acc : 0.0;
    }
  }

  // right-hand side: per-cell quadrature of the source (|det J| * values at
  // the cell quadrature points, src[q][component]) against the test basis
  __device__ double rhs(const double (*src)[4], int row) const {
    const int D = T.D, np = T.np, nq = T.nq, na = T.na, ni = T.ni;
    if (row >= ni) return 0.0;
    double s = 0.0;
    if (PHYS == kDarcy) {
      if (row < na) {  // v . f
        const int a = row / np, i = row - a * np;
#pragma unroll 4
        for (int q = 0; q < T.nqc; ++q) s += T.Bp[q * np + i] * src[q][a];
      } else {  // w g
        const int i = row - na;
#pragma unroll 4
        for (int q = 0; q < T.nqc; ++q) s += T.Bq[q * nq + i] * src[q][D];
      }
      return s;
    }
    if (row < na) return 0.0;
    const int c = (row - na) / nq, j = row - na - c * nq;  // -omega . f
#pragma unroll 4
    for (int q = 0; q < T.nqc; ++q) s += T.Bq[q * nq + j] * src[q][c];
    return -s;
  }
};

// Dense cell systems from HBM (map-level StaticCondensationMap).
struct DenseShared {
  int dummy;
};
struct DenseGen {
  DenseIn in;
  using Shared = DenseShared;
  __device__ void begin(int64_t, Shared&) const {}
  __device__ double entry(const Shared&, int row, int col, int64_t cell) const {
    const int n = in.n;
    const int r = in.perm[row];
    if (col == n) return in.b[cell * n + r];
    return in.A[(cell * n + in.perm[col]) * n + r];
  }
};

// Dense backward map: rows/cols = interior only; rhs = b_i - A_ib x.
struct BackwardGen {
  DenseIn in;
  const double* x;  // [ncells][nb]
  using Shared = DenseShared;
  __device__ void begin(int64_t, Shared&) const {}
  __device__ double entry(const Shared&, int row, int col, int64_t cell) const {
    const int n = in.n, ni = in.ni, nb = in.nb;
    const int r = in.perm[row];
    if (col < ni) return in.A[(cell * n + in.perm[col]) * n + r];
    // b1 - A12 * x  (src/BackwardStaticCondensationMap.jl:91), gemv order
    double s = in.b[cell * n + r];
    for (int j = 0; j < nb; ++j) s -= in.A[(cell * n + in.perm[ni + j]) * n + r] * x[cell * nb + j];
    return s;
  }
};

template <class G>
struct GenCall {
  __device__ static double entry(const G& gen, const typename G::Shared& sh, int row, int col,
                                 int64_t cell) {
    return gen.entry(sh, row, col, cell);
  }
  // four entries (ok[u] false gives 0)
  __device__ static void entries4(const G& gen, const typename G::Shared& sh, const int* row,
                                  const int* col, const bool* ok, int64_t cell, double* v) {
#pragma unroll
    for (int u = 0; u < 4; ++u) v[u] = ok[u] ? gen.entry(sh, row[u], col[u], cell) : 0.0;
  }
};
template <int PHYS>
struct GenCall<FusedGen<PHYS>> {
  __device__ static double entry(const FusedGen<PHYS>& gen, const FusedShared& sh, int row,
                                 int col, int64_t) {
    return gen.entry(sh, row, col);
  }
  // the batched form: every load of the four entries in flight at once
  __device__ static void entries4(const FusedGen<PHYS>& gen, const FusedShared& sh,
                                  const int* row, const int* col, const bool* ok, int64_t,
                                  double* v) {
    gen.template entries<4>(sh, row, col, ok, v);
  }
};

// --------------------------------------------------------------------------
// Augmented elimination kernel
//
// Outputs (CondenseOut):
//   S, g  boundary rows after the ni elimination steps;
//   U     packed upper factor of A_ii in LAPACK row order: column c holds rows
//         0..c at U[c(c+1)/2 + k];
//   Y     L^-1 P [A_ib b_i] in LAPACK row order, [nb+1][ni] column-major;
//   X     (backward map only) the back-substituted right-hand side [ni].
// The hot loops are branch-free: inactive rows get a zero multiplier, the
// pivot row index is made wave-uniform, and every update is a plain FMA.
// --------------------------------------------------------------------------
template <int PR, int PC, int R, int CC, class Gen, bool BACKSUB>
__global__ void __launch_bounds__(PR * PC)
    condense_kernel(Gen gen, CondenseOut out, int64_t ncells, int ni, int nb) {
  constexpr int NT = PR * PC;
  constexpr int NMAX = PR * R;
  constexpr int NCMAX = PC * CC;
  static_assert(PR <= 64 && (64 % PR) == 0, "thread-grid columns must sit inside one wavefront");
  static_assert(R <= 32, "row mask is 32 bits");
  const int n = ni + nb;   // rows
  const int ncol = n + 1;  // columns incl. right-hand side
  const int tid = threadIdx.x;
  const int tr = tid % PR;
  const int tc = tid / PR;

  __shared__ double s_prow[2][NCMAX];
  __shared__ double s_lcol[2][NMAX];
  __shared__ double s_diag[NMAX];
  __shared__ int s_pos[NMAX];
  __shared__ int s_at[NMAX];
  __shared__ int s_piv[2];
  __shared__ int s_info;
  __shared__ typename Gen::Shared s_gen;
  __shared__ double s_slab[PC][NMAX];

  for (int64_t cell = blockIdx.x; cell < ncells; cell += gridDim.x) {
    gen.begin(cell, s_gen);
    for (int i = tid; i < NMAX; i += NT) {
      s_pos[i] = i;
      s_at[i] = i;
    }
    if (tid == 0) s_info = 0;
    __syncthreads();

    // Generate the cell system one slab of PC columns at a time: slab c holds
    // local column c of every thread.  The generator is expanded once (runtime
    // loop) and the register tile is filled through an unrolled select.
    double a[R][CC];
    uint32_t act = 0;
#pragma unroll
    for (int r = 0; r < R; ++r)
      if (tr + PR * r < n) act |= 1u << r;
#pragma unroll 1
    for (int c = 0; c < CC; ++c) {
      for (int t = tid; t < n * PC; t += NT) {
        const int cc = t / n, row = t - cc * n;
        const int col = PC * c + cc;
        s_slab[cc][row] = col < ncol ? GenCall<Gen>::entry(gen, s_gen, row, col, cell) : 0.0;
      }
      __syncthreads();
      double v[R];
#pragma unroll
      for (int r = 0; r < R; ++r) {
        const int row = tr + PR * r;
        v[r] = row < n ? s_slab[tc][row] : 0.0;
      }
#pragma unroll
      for (int c2 = 0; c2 < CC; ++c2)
        if (c2 == c) {
#pragma unroll
          for (int r = 0; r < R; ++r) a[r][c2] = v[r];
        }
      __syncthreads();
    }

    // ---------------- forward elimination over the ni interior pivots ----
    // One barrier per step: the owners of column j (one half-wave group)
    // pick the pivot, publish it with the multipliers and update the LAPACK
    // position bookkeeping; after the barrier every thread group fetches the
    // pivot row from its own lane p % PR by shuffle and updates its tile.
#pragma unroll 1
    for (int j = 0; j < ni; ++j) {
      const int buf = j & 1;
      const int jc = j % PC, jl = j / PC;
      if (tc == jc) {
        double colv[R];
#pragma unroll
        for (int r = 0; r < R; ++r) colv[r] = 0.0;
#pragma unroll
        for (int c = 0; c < CC; ++c)
          if (c == jl) {
#pragma unroll
            for (int r = 0; r < R; ++r) colv[r] = a[r][c];
          }
        double bav = -1.0;
        int bpos = INT_MAX;
#pragma unroll
        for (int r = 0; r < R; ++r) {
          const int row = tr + PR * r;
          const bool cand = row < ni && ((act >> r) & 1u);
          const double av = cand ? fabs(colv[r]) : -1.0;
          const int pr = s_pos[row];
          const bool better = av > bav || (av == bav && cand && pr < bpos);
          bav = better ? av : bav;
          bpos = better ? pr : bpos;
        }
        double m = bav;
#pragma unroll
        for (int off = 1; off < PR; off <<= 1) m = fmax(m, __shfl_xor(m, off, PR));
        int key = bav == m ? bpos : INT_MAX;
#pragma unroll
        for (int off = 1; off < PR; off <<= 1) key = min(key, __shfl_xor(key, off, PR));
        const int p = s_at[key];  // pivot row (smallest position among the maxima)
        const int pl = p / PR;
        double mine = 0.0;
#pragma unroll
        for (int r = 0; r < R; ++r)
          if (r == pl) mine = colv[r];
        const double pv = __shfl(mine, p % PR, PR);
        const double rinv = pv != 0.0 ? 1.0 / pv : 0.0;
#pragma unroll
        for (int r = 0; r < R; ++r) {
          const int row = tr + PR * r;
          const bool use = ((act >> r) & 1u) && row != p;
          const double lv = use ? colv[r] * rinv : 0.0;
          s_lcol[buf][row] = lv;
          if (out.L && use && row < ni) out.L[cell * out.u_stride + (int64_t)j * ni + row] = lv;
        }
        if (tr == 0) {
          s_piv[buf] = p;
          s_diag[j] = pv;
          if (pv == 0.0 && s_info == 0) s_info = j + 1;
          // LAPACK interchange of positions j and pos(p) = key
          const int q = s_at[j];
          s_at[key] = q;
          s_pos[q] = key;
          s_at[j] = p;
          s_pos[p] = j;
        }
      }
      __syncthreads();
      const int p = __builtin_amdgcn_readfirstlane(s_piv[buf]);
      const int pl = p / PR, plane = p % PR;
      if (tr == plane) {
#pragma unroll
        for (int r = 0; r < R; ++r)
          if (r == pl) {  // wave-uniform
#pragma unroll
            for (int c = 0; c < CC; ++c) s_prow[buf][tc + PC * c] = a[r][c];
          }
        act &= ~(1u << pl);
      }
      __syncthreads();
      double l[R];
#pragma unroll
      for (int r = 0; r < R; ++r) {
        const double lv = s_lcol[buf][tr + PR * r];
        l[r] = ((act >> r) & 1u) ? lv : 0.0;
      }
      double u[CC];
#pragma unroll
      for (int c = 0; c < CC; ++c) u[c] = s_prow[buf][tc + PC * c];
#pragma unroll
      for (int c = 0; c < CC; ++c) {
        if (c < jl) continue;  // wave-uniform: local column fully eliminated
#pragma unroll
        for (int r = 0; r < R; ++r) a[r][c] = fma(-l[r], u[c], a[r][c]);
      }
    }
    __syncthreads();  // last step's pivot row / multipliers fully consumed
    const int info = s_info;
    if (out.P)
      for (int k = tid; k < ni; k += NT) out.P[cell * out.u_stride + k] = (double)s_at[k];

    // ---------------- back substitution (backward map only) ---------------
    if (BACKSUB) {
#pragma unroll 1
      for (int k = ni - 1; k >= 0; --k) {
        const int buf = k & 1;
        const int p = __builtin_amdgcn_readfirstlane(s_at[k]);
        const int pl = p / PR;
        const double d = s_diag[k];
        if (tr == p % PR) {
#pragma unroll
          for (int r = 0; r < R; ++r)
            if (r == pl) {
#pragma unroll
              for (int c = 0; c < CC; ++c) {
                const int col = tc + PC * c;
                const double x = d != 0.0 ? a[r][c] / d : 0.0;
                a[r][c] = col >= ni ? x : a[r][c];
                s_prow[buf][col] = x;
              }
            }
        }
        const int kc = k % PC, kl = k / PC;
        if (tc == kc) {
#pragma unroll
          for (int c = 0; c < CC; ++c)
            if (c == kl) {
#pragma unroll
              for (int r = 0; r < R; ++r) {
                const int row = tr + PR * r;
                const bool use = row < ni && s_pos[row] < k;
                s_lcol[buf][row] = use ? a[r][c] : 0.0;
              }
            }
        }
        __syncthreads();
#pragma unroll
        for (int r = 0; r < R; ++r) {
          const double lr = s_lcol[buf][tr + PR * r];
#pragma unroll
          for (int c = 0; c < CC; ++c) {
            const int col = tc + PC * c;
            const double u = col >= ni ? s_prow[buf][col] : 0.0;
            a[r][c] = fma(-lr, u, a[r][c]);
          }
        }
      }
      __syncthreads();
    }

    // ---------------- outputs, one slab of PC columns at a time ----------
    // (registers -> LDS with constant indices, then coalesced global stores)
#pragma unroll 1
    for (int c = 0; c < CC; ++c) {
      if (PC * c + PC <= ni && out.U == nullptr) continue;  // nothing to store
#pragma unroll
      for (int c2 = 0; c2 < CC; ++c2)
        if (c2 == c) {
#pragma unroll
          for (int r = 0; r < R; ++r) s_slab[tc][tr + PR * r] = a[r][c2];
        }
      __syncthreads();
      for (int t = tid; t < n * PC; t += NT) {
        const int cc = t / n, row = t - cc * n;
        const int col = PC * c + cc;
        if (col >= ncol) continue;
        const double v = s_slab[cc][row];
        if (row >= ni) {
          if (col < ni) continue;
          if (col < n) {
            if (out.S) out.S[(cell * nb + (col - ni)) * nb + (row - ni)] = v;
          } else if (out.g) {
            out.g[cell * nb + (row - ni)] = v;
          }
        } else {
          const int k = s_pos[row];
          if (col < ni) {
            if (out.U && col >= k)
              out.U[cell * ustride(out, ni) + col * (col + 1) / 2 + k] = v;
          } else {
            if (out.Y) out.Y[cell * ystride(out, ni, nb) + (int64_t)(col - ni) * ni + k] = v;
            if (out.X && col == n) out.X[cell * ni + k] = v;
          }
        }
      }
      __syncthreads();
    }
    if (tid == 0 && out.info) out.info[cell] = info;
    __syncthreads();  // LDS reuse by the next cell
  }
}

// --------------------------------------------------------------------------
// Wave-per-cell elimination for small systems (n <= 64 rows, n+1 <= NCOL
// columns): lane i owns row i of [A | b] in registers.  The column loop is
// unrolled so every register index is a compile-time constant; the pivot row
// is broadcast with readlane (its values land in SGPRs and feed the FMAs as
// scalar operands); the LAPACK position of each row lives in a register.
// No LDS traffic and no barriers in the elimination.
// --------------------------------------------------------------------------
__device__ inline double wave_readlane_f64(double v, int lane) {
  const int lo = __builtin_amdgcn_readlane(__double2loint(v), lane);
  const int hi = __builtin_amdgcn_readlane(__double2hiint(v), lane);
  return __hiloint2double(hi, lo);
}

template <int NI, int NB, class Gen, bool BACKSUB>
__global__ void __launch_bounds__(64)
    condense_wave_kernel(Gen gen, CondenseOut out, int64_t ncells) {
  constexpr int N = NI + NB, NCOL = N + 1;
  constexpr int SW = 8;                      // slab width (columns staged per pass)
  constexpr int NS = (NCOL + SW - 1) / SW;   // slabs
  static_assert(N <= 64, "wave kernel: one row per lane");
  const int ni = NI, nb = NB, n = N;
  const int lane = threadIdx.x;
  __shared__ typename Gen::Shared s_gen;
  __shared__ double s_slab[SW * 64];

  for (int64_t cell = blockIdx.x; cell < ncells; cell += gridDim.x) {
    gen.begin(cell, s_gen);
    __syncthreads();
    // generate SW columns at a time into LDS, pick this lane's row
    double a[NCOL];
#pragma unroll 1
    for (int sl = 0; sl < NS; ++sl) {
      for (int t = lane; t < N * SW; t += 64) {
        const int cc = t / N, row = t - cc * N;
        const int col = SW * sl + cc;
        s_slab[cc * 64 + row] = col < NCOL ? GenCall<Gen>::entry(gen, s_gen, row, col, cell) : 0.0;
      }
      __syncthreads();
      double v[SW];
#pragma unroll
      for (int cc = 0; cc < SW; ++cc) v[cc] = lane < N ? s_slab[cc * 64 + lane] : 0.0;
#pragma unroll
      for (int s2 = 0; s2 < NS; ++s2)
        if (s2 == sl) {
#pragma unroll
          for (int cc = 0; cc < SW; ++cc)
            if (SW * s2 + cc < NCOL) a[SW * s2 + cc] = v[cc];
        }
      __syncthreads();
    }

    bool active = lane < N;
    int mypos = lane;        // LAPACK position of this row (interior rows)
    int info = 0;
    double diag_mine = 0.0;  // U(k,k) when this row is the pivot of step k
#pragma unroll
    for (int j = 0; j < NI; ++j) {
      const double v = a[j];
      const bool cand = active && lane < NI;
      const double av = cand ? fabs(v) : -1.0;
      double m = av;  // wave max of |v|
#pragma unroll
      for (int off = 32; off >= 1; off >>= 1) m = fmax(m, __shfl_xor(m, off, 64));
      // among the maxima, the row with the smallest LAPACK position
      int key = (cand && av == m) ? mypos : INT_MAX;
#pragma unroll
      for (int off = 32; off >= 1; off >>= 1) key = min(key, __shfl_xor(key, off, 64));
      const uint64_t hit = __ballot(cand && av == m && mypos == key);
      const int p = __builtin_amdgcn_readfirstlane(__ffsll((unsigned long long)hit) - 1);
      const double pv = wave_readlane_f64(v, p);
      if (pv == 0.0 && info == 0) info = j + 1;
      const double rinv = pv != 0.0 ? 1.0 / pv : 0.0;
      // LAPACK interchange of positions j and pos(p)
      const int pos_p = __builtin_amdgcn_readlane(mypos, p);
      if (mypos == j) mypos = pos_p;
      if (lane == p) {
        mypos = j;
        active = false;
        diag_mine = v;
      }
      const double l = active ? v * rinv : 0.0;
#pragma unroll
      for (int c = j + 1; c < NCOL; ++c) a[c] = fma(-l, wave_readlane_f64(a[c], p), a[c]);
    }

    if (BACKSUB) {
      // backward map: single rhs column at index NI (NB == 0)
#pragma unroll
      for (int k = NI - 1; k >= 0; --k) {
        const uint64_t own = __ballot(lane < NI && mypos == k);
        const int p = __builtin_amdgcn_readfirstlane(__ffsll((unsigned long long)own) - 1);
        const double d = wave_readlane_f64(diag_mine, p);
        const double x = d != 0.0 ? wave_readlane_f64(a[NCOL - 1], p) / d : 0.0;
        if (lane == p)
          a[NCOL - 1] = x;
        else if (lane < NI && mypos < k)
          a[NCOL - 1] = fma(-a[k], x, a[NCOL - 1]);
      }
    }

    // ---------------- outputs ----------------
    if (lane < N) {
      if (lane >= NI) {
        const int br = lane - NI;
#pragma unroll
        for (int c = NI; c < N; ++c)
          if (out.S) out.S[(cell * nb + (c - NI)) * nb + br] = a[c];
        if (out.g) out.g[cell * nb + br] = a[N];
      } else {
        const int k = mypos;
        const int64_t ubase = cell * ustride(out, ni);
        if (out.U) {
#pragma unroll
          for (int c = 0; c < NI; ++c)
            if (c >= k) out.U[ubase + c * (c + 1) / 2 + k] = a[c];
        }
        if (out.Y) {
#pragma unroll
          for (int c = NI; c < NCOL; ++c) out.Y[cell * ystride(out, ni, nb) + (int64_t)(c - NI) * ni + k] = a[c];
        }
        if (out.X) out.X[cell * ni + k] = a[N];
      }
    }
    (void)n;
    if (lane == 0 && out.info) out.info[cell] = info;
    __syncthreads();
  }
}

// --------------------------------------------------------------------------
// Column-per-wave elimination for medium systems (n <= 64*R rows, n+1 <= NW*CC
// columns): lane l of wave w owns rows l + 64 r and columns w + NW c of
// [A | b].  Every wave therefore holds the pivot row p in its own lane p % 64,
// so the pivot row is read with readlane (no LDS, no barrier); only the
// multiplier column crosses waves, through LDS, once per step.  The wave that
// owns column j+1 updates it first and searches the pivot of step j+1 while
// the other waves finish step j (look-ahead), so each step costs one barrier.
// --------------------------------------------------------------------------
__device__ inline int64_t dbits(double v) { return __double_as_longlong(v); }

// Wave-wide maximum of a signed 64-bit key / minimum of a 32-bit int, with
// DPP row rotations (row_ror 1,2,4,8) and the GFX9 row broadcasts
// (row_bcast15, row_bcast31); the result is read from lane 63 and is
// wave-uniform.  Lanes masked out of a DPP step keep their own value.
template <int CTRL, int ROW_MASK>
__device__ inline int64_t dpp_i64(int64_t x) {
  const int lo = (int)(uint32_t)(uint64_t)x, hi = (int)((uint64_t)x >> 32);
  const int lo2 = __builtin_amdgcn_update_dpp(lo, lo, CTRL, ROW_MASK, 0xf, false);
  const int hi2 = __builtin_amdgcn_update_dpp(hi, hi, CTRL, ROW_MASK, 0xf, false);
  return (int64_t)(((uint64_t)(uint32_t)hi2 << 32) | (uint64_t)(uint32_t)lo2);
}
__device__ inline int64_t wave_max_i64(int64_t x) {
  int64_t y;
  y = dpp_i64<0x121, 0xf>(x); x = y > x ? y : x;  // row_ror:1
  y = dpp_i64<0x122, 0xf>(x); x = y > x ? y : x;  // row_ror:2
  y = dpp_i64<0x124, 0xf>(x); x = y > x ? y : x;  // row_ror:4
  y = dpp_i64<0x128, 0xf>(x); x = y > x ? y : x;  // row_ror:8
  y = dpp_i64<0x142, 0xa>(x); x = y > x ? y : x;  // row_bcast:15
  y = dpp_i64<0x143, 0xc>(x); x = y > x ? y : x;  // row_bcast:31
  const int lo = __builtin_amdgcn_readlane((int)(uint32_t)(uint64_t)x, 63);
  const int hi = __builtin_amdgcn_readlane((int)((uint64_t)x >> 32), 63);
  return (int64_t)(((uint64_t)(uint32_t)hi << 32) | (uint64_t)(uint32_t)lo);
}
// max over the wavefront of nonnegative doubles (ordered like their bit patterns)
__device__ inline double wave_max_f64(double x) {
  return __longlong_as_double(wave_max_i64(__double_as_longlong(x)));
}
__device__ inline int wave_min_i32(int x) {
  int y;
  y = __builtin_amdgcn_update_dpp(x, x, 0x121, 0xf, 0xf, false); x = min(x, y);
  y = __builtin_amdgcn_update_dpp(x, x, 0x122, 0xf, 0xf, false); x = min(x, y);
  y = __builtin_amdgcn_update_dpp(x, x, 0x124, 0xf, 0xf, false); x = min(x, y);
  y = __builtin_amdgcn_update_dpp(x, x, 0x128, 0xf, 0xf, false); x = min(x, y);
  y = __builtin_amdgcn_update_dpp(x, x, 0x142, 0xa, 0xf, false); x = min(x, y);
  y = __builtin_amdgcn_update_dpp(x, x, 0x143, 0xc, 0xf, false); x = min(x, y);
  return __builtin_amdgcn_readlane(x, 63);
}

// --------------------------------------------------------------------------
// Blocked elimination for large systems (up to NMAX rows): the cell system
// lives in a per-workgroup HBM workspace (column-major, leading dimension
// ld); panels of KB columns are factored in LDS (partial pivoting over the
// interior rows, LAPACK position bookkeeping), the panel's pivot rows are
// brought to U form by forward substitution (U12), and the remaining active
// rows get one rank-KB update per panel from register tiles.  Each panel
// reads and writes the active trailing matrix once, so the workspace traffic
// per cell falls with KB: the 32-column variant (systems up to 416 rows)
// stages the pivot rows of U12 in LDS 128 trailing columns at a time (the
// rest wait in the workspace's pivot rows) so that the 32-column panel fits.
// --------------------------------------------------------------------------
constexpr int kLargeMaxN = 512;
constexpr int kLargeKB = 16;
constexpr int kLarge32MaxN = 416;  // 32-column panels: systems up to this many rows

template <class Gen, bool BACKSUB, int NT, int KB = kLargeKB, int NMAX = kLargeMaxN>
__global__ void __launch_bounds__(NT)
    condense_large_kernel(Gen gen, CondenseOut out, int64_t ncells, int ni, int nb,
                          double* __restrict__ work, int ld) {
  constexpr int NW = NT / 64;
  constexpr int CWU = KB == kLargeKB ? NMAX : 128;  // trailing columns of U12 staged in LDS
  const int n = ni + nb, ncol = n + 1;
  const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
  double* M = work + (int64_t)blockIdx.x * ld * (int64_t)ncol;

  __shared__ double s_P[KB][NMAX];        // panel columns (all rows)
  __shared__ double s_U[KB][CWU + 1];     // pivot rows x (a chunk of the) trailing columns
  __shared__ int s_pos[NMAX], s_at[NMAX];
  __shared__ unsigned char s_act[NMAX];
  __shared__ int s_piv[KB];
  __shared__ double s_diag[NMAX];
  __shared__ int64_t s_wkey[NW];
  __shared__ int s_wpos[NW];
  __shared__ int s_info;
  __shared__ typename Gen::Shared s_gen;

  for (int64_t cell = blockIdx.x; cell < ncells; cell += gridDim.x) {
    gen.begin(cell, s_gen);
    for (int i = tid; i < NMAX; i += NT) {
      s_pos[i] = i;
      s_at[i] = i;
      s_act[i] = i < n ? 1 : 0;
    }
    if (tid == 0) s_info = 0;
    __syncthreads();
    // generate into the workspace (coalesced, column-major), four entries per
    // thread at a time
    for (int64_t t0 = tid; t0 < (int64_t)n * ncol; t0 += 4 * NT) {
      int rows[4], cols[4];
      bool ok[4];
      double v[4];
#pragma unroll
      for (int u = 0; u < 4; ++u) {
        const int64_t t = t0 + u * NT;
        ok[u] = t < (int64_t)n * ncol;
        cols[u] = ok[u] ? (int)(t / n) : 0;
        rows[u] = ok[u] ? (int)(t - (int64_t)cols[u] * n) : 0;
      }
      GenCall<Gen>::entries4(gen, s_gen, rows, cols, ok, cell, v);
#pragma unroll
      for (int u = 0; u < 4; ++u)
        if (ok[u]) M[(int64_t)cols[u] * ld + rows[u]] = v[u];
    }
    __syncthreads();

    for (int j0 = 0; j0 < ni; j0 += KB) {
      const int kb = min(KB, ni - j0);
      const int jt = j0 + kb;  // first trailing column
      // ---- load the panel ----
      for (int t = tid; t < kb * n; t += NT) {
        const int cc = t / n, row = t - cc * n;
        s_P[cc][row] = M[(int64_t)(j0 + cc) * ld + row];
      }
      __syncthreads();
      // ---- factor the panel in LDS ----
      for (int jj = 0; jj < kb; ++jj) {
        const int j = j0 + jj;
        int64_t bkey = -1;
        int bpos = INT_MAX;
        for (int row = tid; row < ni; row += NT) {
          if (!s_act[row]) continue;
          const int64_t key = dbits(fabs(s_P[jj][row]));
          const int pr = s_pos[row];
          if (key > bkey || (key == bkey && pr < bpos)) {
            bkey = key;
            bpos = pr;
          }
        }
        const int64_t wm = wave_max_i64(bkey);
        const int wk = wave_min_i32(bkey == wm ? bpos : INT_MAX);
        if (lane == 0) {
          s_wkey[wv] = wm;
          s_wpos[wv] = wk;
        }
        __syncthreads();
        int64_t m = s_wkey[0];
        for (int q = 1; q < NW; ++q) m = s_wkey[q] > m ? s_wkey[q] : m;
        int key = INT_MAX;
        for (int q = 0; q < NW; ++q)
          if (s_wkey[q] == m) key = min(key, s_wpos[q]);
        const int p = s_at[key];
        const double pv = s_P[jj][p];
        const double rinv = pv != 0.0 ? 1.0 / pv : 0.0;
        // multipliers in place, then update the rest of the panel
        for (int row = tid; row < n; row += NT) {
          if (!s_act[row] || row == p) continue;
          const double l = s_P[jj][row] * rinv;
          s_P[jj][row] = l;
          for (int cc = jj + 1; cc < kb; ++cc) s_P[cc][row] = fma(-l, s_P[cc][p], s_P[cc][row]);
        }
        __syncthreads();
        if (tid == 0) {
          s_piv[jj] = p;
          s_diag[j] = pv;
          if (pv == 0.0 && s_info == 0) s_info = j + 1;
          s_act[p] = 0;
          const int q = s_at[j];
          s_at[key] = q;
          s_pos[q] = key;
          s_at[j] = p;
          s_pos[p] = j;
        }
        __syncthreads();
      }
      // ---- pivot rows: panel part to the outputs, trailing part -> U12 ----
      for (int t = tid; t < kb * KB; t += NT) {
        const int jj = t / KB, cc = t - jj * KB;
        if (cc >= kb || cc < jj) continue;
        const int col = j0 + cc, k = j0 + jj;
        const double v = s_P[cc][s_piv[jj]];
        if (out.U) out.U[cell * ustride(out, ni) + col * (col + 1) / 2 + k] = v;
        if (BACKSUB) M[(int64_t)col * ld + s_piv[jj]] = v;  // U row kept for back substitution
      }
      const int ntr = ncol - jt;
      const int rt = (n + 3) / 4;
      // rank-kb update of the active rows on trailing columns [c0, c0 + cw),
      // with their U12 in s_U.  A thread's 4 x 4 tile holds rows tr + rt i
      // (i < 4): consecutive lanes take consecutive rows, so the workspace
      // accesses are contiguous and the panel reads from LDS hit distinct banks.
      auto update_chunk = [&](int c0, int cw) {
        const int ct = (cw + 3) / 4;
        for (int tile = tid; tile < rt * ct; tile += NT) {
          const int tr = tile % rt, tc4 = (tile / rt) * 4;
          double acc[4][4];
#pragma unroll
          for (int i = 0; i < 4; ++i)
#pragma unroll
            for (int q = 0; q < 4; ++q) acc[i][q] = 0.0;
          bool rowok[4];
#pragma unroll
          for (int i = 0; i < 4; ++i) rowok[i] = tr + rt * i < n && s_act[tr + rt * i];
          if (!(rowok[0] | rowok[1] | rowok[2] | rowok[3])) continue;
          // the tile's entries load first, in flight during the rank-kb products
          double mt[4][4];
#pragma unroll
          for (int q = 0; q < 4; ++q) {
            const int c = jt + c0 + tc4 + q;
#pragma unroll
            for (int i = 0; i < 4; ++i)
              mt[i][q] = (tc4 + q < cw && rowok[i]) ? M[(int64_t)c * ld + tr + rt * i] : 0.0;
          }
          for (int jj = 0; jj < kb; ++jj) {
            double l[4], u[4];
#pragma unroll
            for (int i = 0; i < 4; ++i) l[i] = rowok[i] ? s_P[jj][tr + rt * i] : 0.0;
#pragma unroll
            for (int q = 0; q < 4; ++q) u[q] = tc4 + q < cw ? s_U[jj][tc4 + q] : 0.0;
#pragma unroll
            for (int i = 0; i < 4; ++i)
#pragma unroll
              for (int q = 0; q < 4; ++q) acc[i][q] = fma(l[i], u[q], acc[i][q]);
          }
#pragma unroll
          for (int q = 0; q < 4; ++q) {
            const int c = jt + c0 + tc4 + q;
            if (tc4 + q >= cw) continue;
#pragma unroll
            for (int i = 0; i < 4; ++i)
              if (rowok[i]) M[(int64_t)c * ld + tr + rt * i] = mt[i][q] - acc[i][q];
          }
        }
      };
      // the U12 value of pivot row jj, trailing column c: the outputs (and the
      // workspace for the back substitution)
      auto put_u12 = [&](int c, int jj, double u) {
        const int k = j0 + jj;
        if (c < ni) {
          if (out.U) out.U[cell * ustride(out, ni) + c * (c + 1) / 2 + k] = u;
        } else {
          if (out.Y) out.Y[cell * ystride(out, ni, nb) + (int64_t)(c - ni) * ni + k] = u;
        }
        if (BACKSUB) M[(int64_t)c * ld + s_piv[jj]] = u;
      };
      if constexpr (KB == kLargeKB) {
        // U12 = L11^-1 A12, a thread per trailing column, all in LDS at once
        for (int c = jt + tid; c < ncol; c += NT) {
          double u[KB], mv[KB];
#pragma unroll
          for (int jj = 0; jj < KB; ++jj)  // the pivot rows' entries, all loads in flight at once
            mv[jj] = jj < kb ? M[(int64_t)c * ld + s_piv[jj]] : 0.0;
#pragma unroll
          for (int jj = 0; jj < KB; ++jj) {
            if (jj < kb) {
              double v = mv[jj];
              for (int ii = 0; ii < jj; ++ii) v = fma(-s_P[ii][s_piv[jj]], u[ii], v);
              u[jj] = v;
              s_U[jj][c - jt] = v;
            }
          }
          for (int jj = 0; jj < kb; ++jj) put_u12(c, jj, u[jj]);
        }
        __syncthreads();
        update_chunk(0, ntr);
        __syncthreads();
      } else {
        // chunks of CWU trailing columns: U12 of the chunk into s_U, then its
        // update.  A wavefront owns 8 columns of the chunk; lane (column
        // lane / 8, rl = lane % 8) holds the pivot rows rl + 8 m, and row jj's
        // final value reaches the column's lanes by a shuffle (each entry
        // subtracts the rows above it in increasing order, as in the
        // thread-per-column substitution)
        constexpr int RPL = KB / 8;
        const int rl = lane & 7;
        int prow[RPL];
#pragma unroll
        for (int m = 0; m < RPL; ++m) prow[m] = rl + 8 * m < kb ? s_piv[rl + 8 * m] : 0;
        for (int c0 = 0; c0 < ntr; c0 += CWU) {
          const int cw = min(CWU, ntr - c0);
          for (int cb = wv * 8; cb < cw; cb += NW * 8) {  // wave-uniform
            const int cc = cb + (lane >> 3), c = jt + c0 + cc;
            const bool colok = cc < cw;
            double v[RPL];
#pragma unroll
            for (int m = 0; m < RPL; ++m)
              v[m] = (colok && rl + 8 * m < kb) ? M[(int64_t)c * ld + prow[m]] : 0.0;
#pragma unroll
            for (int jj = 0; jj < KB; ++jj) {
              const double u = __shfl(v[jj >> 3], (lane & ~7) | (jj & 7), 64);
#pragma unroll
              for (int m = 0; m < RPL; ++m) {
                const int j2 = rl + 8 * m;
                if (j2 > jj && j2 < kb) v[m] = fma(-s_P[jj][prow[m]], u, v[m]);
              }
            }
#pragma unroll
            for (int m = 0; m < RPL; ++m) {
              const int j2 = rl + 8 * m;
              if (!colok || j2 >= kb) continue;
              s_U[j2][cc] = v[m];
              put_u12(c, j2, v[m]);
            }
          }
          __syncthreads();
          update_chunk(c0, cw);
          __syncthreads();
        }
      }
    }

    // ---- back substitution (backward map only: single rhs column) -------
    if (BACKSUB) {
      // rhs column n lives in M at the pivot rows (U12 values were stored)
      if (tid == 0) {
        for (int k = ni - 1; k >= 0; --k) {
          const int p = s_at[k];
          double x = M[(int64_t)n * ld + p];
          for (int c = k + 1; c < ni; ++c)
            x = fma(-M[(int64_t)c * ld + p], M[(int64_t)n * ld + s_at[c]], x);
          const double d = s_diag[k];
          x = d != 0.0 ? x / d : 0.0;
          M[(int64_t)n * ld + p] = x;
          if (out.X) out.X[cell * ni + k] = x;
        }
      }
      __syncthreads();
    }
    // ---- boundary rows: S and g ----
    for (int64_t t = tid; t < (int64_t)nb * (nb + 1); t += NT) {
      const int cb = (int)(t / nb), rb = (int)(t - (int64_t)cb * nb);
      const double v = M[(int64_t)(ni + cb) * ld + ni + rb];
      if (cb < nb) {
        if (out.S) out.S[(cell * nb + cb) * nb + rb] = v;
      } else if (out.g) {
        out.g[cell * nb + rb] = v;
      }
    }
    if (tid == 0 && out.info) out.info[cell] = s_info;
    __syncthreads();
  }
}

// --------------------------------------------------------------------------
// Blocked elimination with MFMA trailing updates (64 < n <= NMAX).  The cell
// system lives row-major in a per-workgroup HBM workspace (ld = ldc).  Per
// panel of 16 interior columns: the panel is copied to LDS and factored by
// wavefront 0 alone (pivot search by DPP reductions, getrf's choice and
// LAPACK position bookkeeping, no workgroup barriers), the pivot rows are
// brought to U form over the trailing columns (U12, one column per thread),
// and the trailing matrix gets the rank-16 update C -= L21 U12 from
// v_mfma_f64_16x16x4f64 tiles (rows of eliminated pivots carry zero
// multipliers, so every tile is updated in full).
// --------------------------------------------------------------------------
typedef double dbl4 __attribute__((ext_vector_type(4)));

__device__ inline void wave_sync_lds() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

template <int NMAX, class Gen, bool BACKSUB>
__global__ void __launch_bounds__(256)
    condense_panel_kernel(Gen gen, CondenseOut out, int64_t ncells, int ni, int nb,
                          double* __restrict__ work, int ldc) {
  constexpr int NT = 256, KB = 16, LDP = NMAX + 16;
  const int n = ni + nb, ncol = n + 1;
  const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
  double* M = work + (int64_t)blockIdx.x * (int64_t)n * ldc;  // M[row * ldc + col]

  __shared__ double s_P[KB][LDP];  // panel columns (all rows), then multipliers
  __shared__ double s_U[KB][LDP];  // U12: pivot rows x trailing columns
  __shared__ int s_pos[NMAX], s_at[NMAX];
  __shared__ unsigned char s_act[NMAX];
  __shared__ int s_piv[KB];
  __shared__ double s_diag[NMAX];
  __shared__ int s_info;
  __shared__ typename Gen::Shared s_gen;

  for (int64_t cell = blockIdx.x; cell < ncells; cell += gridDim.x) {
    gen.begin(cell, s_gen);
    for (int i = tid; i < NMAX; i += NT) {
      s_pos[i] = i;
      s_at[i] = i;
      s_act[i] = i < n ? 1 : 0;
    }
    if (tid == 0) s_info = 0;
    __syncthreads();
    for (int64_t t = tid; t < (int64_t)n * ncol; t += NT) {
      const int row = (int)(t / ncol), col = (int)(t - (int64_t)row * ncol);
      M[(int64_t)row * ldc + col] = GenCall<Gen>::entry(gen, s_gen, row, col, cell);
    }
    __syncthreads();

    for (int j0 = 0; j0 < ni; j0 += KB) {
      const int kb = min(KB, ni - j0);
      const int jt = j0 + kb;  // first trailing column
      for (int t = tid; t < n * KB; t += NT) {
        const int row = t / KB, cc = t - row * KB;
        s_P[cc][row] = cc < kb ? M[(int64_t)row * ldc + j0 + cc] : 0.0;
      }
      __syncthreads();
      // ---- factor the panel: wavefront 0 only ----
      if (wv == 0) {
        for (int jj = 0; jj < kb; ++jj) {
          const int j = j0 + jj;
          int64_t bkey = -1;
          int bpos = INT_MAX;
          for (int row = lane; row < ni; row += 64) {
            if (!s_act[row]) continue;
            const int64_t key = dbits(fabs(s_P[jj][row]));
            const int pr = s_pos[row];
            if (key > bkey || (key == bkey && pr < bpos)) {
              bkey = key;
              bpos = pr;
            }
          }
          const int64_t wm = wave_max_i64(bkey);
          const int key = wave_min_i32(bkey == wm ? bpos : INT_MAX);
          const int p = s_at[key];
          const double pv = s_P[jj][p];
          wave_sync_lds();
          if (lane == 0) {
            s_piv[jj] = p;
            s_diag[j] = pv;
            if (pv == 0.0 && s_info == 0) s_info = j + 1;
            s_act[p] = 0;
            const int q = s_at[j];
            s_at[key] = q;
            s_pos[q] = key;
            s_at[j] = p;
            s_pos[p] = j;
          }
          wave_sync_lds();
          const double rinv = pv != 0.0 ? 1.0 / pv : 0.0;
          for (int row = lane; row < n; row += 64) {
            if (!s_act[row]) continue;
            const double l = s_P[jj][row] * rinv;
            s_P[jj][row] = l;
            for (int cc = jj + 1; cc < kb; ++cc) s_P[cc][row] = fma(-l, s_P[cc][p], s_P[cc][row]);
          }
          wave_sync_lds();
        }
      }
      __syncthreads();
      // ---- pivot rows: panel part to the outputs, trailing part -> U12 ----
      for (int t = tid; t < kb * KB; t += NT) {
        const int jj = t / KB, cc = t - jj * KB;
        if (cc >= kb || cc < jj) continue;
        const int col = j0 + cc, k = j0 + jj;
        const double v = s_P[cc][s_piv[jj]];
        if (out.U) out.U[cell * ustride(out, ni) + col * (col + 1) / 2 + k] = v;
        if (BACKSUB) M[(int64_t)s_piv[jj] * ldc + col] = v;
      }
      for (int c = jt + tid; c < ncol; c += NT) {
        double u[KB], mv[KB];
#pragma unroll
        for (int jj = 0; jj < KB; ++jj)  // all loads in flight before the substitution
          mv[jj] = jj < kb ? M[(int64_t)s_piv[jj] * ldc + c] : 0.0;
#pragma unroll
        for (int jj = 0; jj < KB; ++jj) {
          if (jj < kb) {
            const int pr = s_piv[jj];
            double v = mv[jj];
            for (int ii = 0; ii < jj; ++ii) v = fma(-s_P[ii][pr], u[ii], v);
            u[jj] = v;
            s_U[jj][c - jt] = v;
            const int k = j0 + jj;
            if (c < ni) {
              if (out.U) out.U[cell * ustride(out, ni) + c * (c + 1) / 2 + k] = v;
            } else if (out.Y) {
              out.Y[cell * ystride(out, ni, nb) + (int64_t)(c - ni) * ni + k] = v;
            }
            if (BACKSUB) M[(int64_t)pr * ldc + c] = v;
          }
        }
      }
      __syncthreads();
      // ---- trailing update C -= L21 U12 with FP64 MFMA tiles ----
      const int ntr = ncol - jt;
      const int rtiles = (n + 15) >> 4, ctiles = (ntr + 15) >> 4;
      const int i16 = lane & 15, q4 = lane >> 4;
      const int ntile = rtiles * ctiles;
      // groups of TG tiles per wavefront: every load of a group is in flight
      // before its MFMAs
      constexpr int TG = 4;
      for (int t0 = wv * TG; t0 < ntile; t0 += (NT / 64) * TG) {
        dbl4 acc[TG];
        bool live[TG];
#pragma unroll
        for (int g = 0; g < TG; ++g) {
          const int tile = t0 + g;
          const int r0 = (tile % rtiles) << 4, c0 = (tile / rtiles) << 4;
          const int ra = r0 + i16, cc = c0 + i16;
          live[g] = tile < ntile && __ballot(ra < n && s_act[ra]) != 0;
          const bool cok = cc < ntr;
#pragma unroll
          for (int v = 0; v < 4; ++v) {
            const int r = r0 + q4 + 4 * v;
            acc[g][v] = (live[g] && cok && r < n) ? M[(int64_t)r * ldc + jt + cc] : 0.0;
          }
        }
#pragma unroll
        for (int g = 0; g < TG; ++g) {
          if (!live[g]) continue;
          const int tile = t0 + g;
          const int r0 = (tile % rtiles) << 4, c0 = (tile / rtiles) << 4;
          const int ra = r0 + i16, cc = c0 + i16;
          const bool aok = ra < n && s_act[ra], cok = cc < ntr;
#pragma unroll
          for (int k4 = 0; k4 < KB / 4; ++k4) {
            const int k = 4 * k4 + q4;
            const double a = (aok && k < kb) ? -s_P[k][ra] : 0.0;
            const double b = (cok && k < kb) ? s_U[k][cc] : 0.0;
            acc[g] = __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, acc[g], 0, 0, 0);
          }
        }
#pragma unroll
        for (int g = 0; g < TG; ++g) {
          if (!live[g]) continue;
          const int tile = t0 + g;
          const int r0 = (tile % rtiles) << 4, c0 = (tile / rtiles) << 4;
          const int cc = c0 + i16;
          if (cc >= ntr) continue;
#pragma unroll
          for (int v = 0; v < 4; ++v) {
            const int r = r0 + q4 + 4 * v;
            if (r < n) M[(int64_t)r * ldc + jt + cc] = acc[g][v];
          }
        }
      }
      __syncthreads();
    }

    // ---- back substitution (backward map only: single rhs column) -------
    if (BACKSUB) {
      if (tid == 0) {
        for (int k = ni - 1; k >= 0; --k) {
          const int p = s_at[k];
          double x = M[(int64_t)p * ldc + n];
          for (int c = k + 1; c < ni; ++c)
            x = fma(-M[(int64_t)p * ldc + c], M[(int64_t)s_at[c] * ldc + n], x);
          const double d = s_diag[k];
          x = d != 0.0 ? x / d : 0.0;
          M[(int64_t)p * ldc + n] = x;
          if (out.X) out.X[cell * ni + k] = x;
        }
      }
      __syncthreads();
    }
    // ---- boundary rows: S and g ----
    for (int64_t t = tid; t < (int64_t)nb * (nb + 1); t += NT) {
      const int cb = (int)(t / nb), rb = (int)(t - (int64_t)cb * nb);
      const double v = M[(int64_t)(ni + rb) * ldc + ni + cb];
      if (cb < nb) {
        if (out.S) out.S[(cell * nb + cb) * nb + rb] = v;
      } else if (out.g) {
        out.g[cell * nb + rb] = v;
      }
    }
    if (tid == 0 && out.info) out.info[cell] = s_info;
    __syncthreads();
  }
}

// --------------------------------------------------------------------------
// Recovery (BackwardStaticCondensationMap with the factors kept from the
// forward pass): r = L^-1 P (b_i - A_ib lambda) = y - Y lambda, then
// U x = r.  One wavefront per cell; the result is scattered into the
// free-value vector [q of all cells; u of all cells; lambda_free].
// --------------------------------------------------------------------------
template <int KMAX>
__global__ void __launch_bounds__(256)
    recover_kernel(const double* __restrict__ U, const double* __restrict__ Y,
                   const int32_t* __restrict__ cell_dofs, const double* __restrict__ lam_free,
                   const double* __restrict__ lam_dir, double* __restrict__ out, int64_t ncells,
                   int ni, int nb, int na, int64_t usz, int64_t ysz) {
  constexpr int WPB = 4;  // wavefronts per block
  __shared__ double s_lam[WPB][256];
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  for (int64_t cell = (int64_t)blockIdx.x * WPB + w; cell < ncells;
       cell += (int64_t)gridDim.x * WPB) {
    for (int j = lane; j < nb; j += 64) {
      const int32_t d = cell_dofs[cell * nb + j];
      s_lam[w][j] = d >= 0 ? lam_free[d] : lam_dir[-(int64_t)d - 1];
    }
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    const double* Yc = Y + cell * ysz;
    double r[KMAX];
#pragma unroll
    for (int t = 0; t < KMAX; ++t) {
      const int k = lane + 64 * t;
      double s = 0.0;
      if (k < ni) {
        s = Yc[(int64_t)nb * ni + k];
        for (int j = 0; j < nb; ++j) s = fma(-Yc[(int64_t)j * ni + k], s_lam[w][j], s);
      }
      r[t] = s;
    }
    const double* Uc = U ? U + cell * usz : nullptr;
#pragma unroll 1
    for (int c = U ? ni - 1 : -1; c >= 0; --c) {
      const int64_t col0 = (int64_t)c * (c + 1) / 2;
      const double ucc = Uc[col0 + c];
      const int tsel = c >> 6, lsel = c & 63;
      double rc = 0.0;
#pragma unroll
      for (int t = 0; t < KMAX; ++t)
        if (t == tsel) rc = r[t];
      rc = __shfl(rc, lsel, 64);
      const double x = ucc != 0.0 ? rc / ucc : 0.0;
#pragma unroll
      for (int t = 0; t < KMAX; ++t) {
        const int k = lane + 64 * t;
        if (k < c) r[t] = fma(-Uc[col0 + k], x, r[t]);
        if (k == c) r[t] = x;
      }
    }
#pragma unroll
    for (int t = 0; t < KMAX; ++t) {
      const int i = lane + 64 * t;
      if (i >= ni) continue;
      if (i < na)
        out[cell * na + i] = r[t];
      else
        out[ncells * na + cell * (ni - na) + (i - na)] = r[t];
    }
    __builtin_amdgcn_wave_barrier();
  }
}

__global__ void copy_kernel(const double* __restrict__ src, double* __restrict__ dst, int64_t n) {
  for (int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; t < n;
       t += (int64_t)gridDim.x * blockDim.x)
    dst[t] = src[t];
}

// dst[i] = src[idx[i]] (renumbering of skeleton values between trace orders)
__global__ void gather_kernel(const double* __restrict__ src, const int64_t* __restrict__ idx,
                              double* __restrict__ dst, int64_t n) {
  for (int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; t < n;
       t += (int64_t)gridDim.x * blockDim.x)
    dst[t] = src[idx[t]];
}

// --------------------------------------------------------------------------
// Skeleton assembly (deterministic; every slot sums its <= 2 contributing
// cells in ascending cell order).  Facet-pair structure: for each free facet
// G (a block of ntf columns), the free facets F coupled to it in ascending
// free-facet order, each with up to two contributing cells.
// --------------------------------------------------------------------------


__global__ void symbolic_kernel(PairTable P, int32_t* __restrict__ rowval) {
  const int ntf = P.ntf;
  for (int64_t G = blockIdx.x; G < P.nff; G += gridDim.x) {
    const int p0 = P.pptr[G], np = P.pptr[G + 1] - p0;
    const int per_col = np * ntf;
    const int64_t base = P.fptr[G];
    for (int t = threadIdx.x; t < per_col * ntf; t += blockDim.x) {
      const int rem = t % per_col;
      const int pr = rem / ntf, m = rem - pr * ntf;
      rowval[base + t] = P.pfacet[p0 + pr] * ntf + m;
    }
  }
}

__global__ void assemble_kernel(PairTable P, const double* __restrict__ S,
                                double* __restrict__ nzval) {
  constexpr int kMaxPairs = 32;
  __shared__ int64_t s_off[2 * kMaxPairs];  // S offset of each contributing block, -1 if none
  const int ntf = P.ntf, nb = P.nb;
  const float inv_ntf = 1.0f / (float)ntf;
  for (int64_t G = blockIdx.x; G < P.nff; G += gridDim.x) {
    const int p0 = P.pptr[G], np = P.pptr[G + 1] - p0;
    const int per_col = np * ntf;
    const int64_t base = P.fptr[G];
    // S offset of block (G, F) in each contributing cell: row block of F,
    // column block of G (the column within it and the row are added below)
    auto block_off = [&](int k2) -> int64_t {
      const int32_t c = P.pcell[2 * p0 + k2];
      if (c < 0) return -1;
      const uint8_t lp = P.plpos[2 * p0 + k2];
      const int lg = lp & 15, lf = lp >> 4;
      return ((int64_t)c * nb + lg * ntf) * nb + lf * ntf;
    };
    const bool staged = np <= kMaxPairs;
    __syncthreads();  // the previous facet's offsets are consumed
    if (staged)
      for (int k2 = threadIdx.x; k2 < 2 * np; k2 += blockDim.x) s_off[k2] = block_off(k2);
    __syncthreads();
    const float inv_pc = 1.0f / (float)per_col;
    for (int t = threadIdx.x; t < per_col * ntf; t += blockDim.x) {
      // t = (mc * np + pr) * ntf + m; quotients from float reciprocals, then
      // corrected to the exact integer quotient
      int mc = (int)((float)t * inv_pc);
      mc += (mc + 1) * per_col <= t;
      mc -= mc * per_col > t;
      const int rem = t - mc * per_col;
      int pr = (int)((float)rem * inv_ntf);
      pr += (pr + 1) * ntf <= rem;
      pr -= pr * ntf > rem;
      const int m = rem - pr * ntf;
      double v = 0.0;
      if (!P.iperm) {
#pragma unroll
        for (int k = 0; k < 2; ++k) {
          const int64_t off = staged ? s_off[2 * pr + k] : block_off(2 * pr + k);
          if (off >= 0) v += S[off + (int64_t)mc * nb + m];
        }
      } else {
        // oriented facets: the facets' own trace dofs (mc of G, m of F) sit at
        // permuted positions of each cell's local facet blocks
#pragma unroll
        for (int k = 0; k < 2; ++k) {
          const int64_t off = staged ? s_off[2 * pr + k] : block_off(2 * pr + k);
          if (off < 0) continue;
          const uint8_t code = P.pperm[2 * (p0 + pr) + k];
          const int col = P.iperm[(code & 15) * ntf + mc], row = P.iperm[(code >> 4) * ntf + m];
          v += S[off + (int64_t)col * nb + row];
        }
      }
      nzval[base + t] = v;
    }
  }
}

// rhs: for free dof (G, m), sum over cells around G (ascending) of
// g_c[i] - sum_{j Dirichlet} S_c[i, j] lam_dir (Gridap _attach_dirichlet).
// Only cells flagged in glpos have Dirichlet dofs; the others skip the scan
// of their dof row, which otherwise is most of the kernel's traffic.
__global__ void rhs_kernel(PairTable P, const double* __restrict__ S, const double* __restrict__ g,
                           const int32_t* __restrict__ cell_dofs, const double* __restrict__ lam_dir,
                           double* __restrict__ rhs) {
  const int ntf = P.ntf, nb = P.nb;
  const int64_t total = P.nff * ntf;
  for (int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; t < total;
       t += (int64_t)gridDim.x * blockDim.x) {
    const int64_t G = t / ntf;
    const int m = (int)(t - G * ntf);
    double acc = 0.0;
    for (int k = 0; k < 2; ++k) {
      const int32_t c = P.gcell[2 * G + k];
      if (c < 0) continue;
      const int mi = P.iperm ? P.iperm[P.gperm[2 * G + k] * ntf + m] : m;
      const uint8_t gl = P.glpos[2 * G + k];
      const int i = (gl & 15) * ntf + mi;
      double v = g[(int64_t)c * nb + i];
      if (gl & kGlposDirichlet) {
        const int32_t* d = cell_dofs + (int64_t)c * nb;
        for (int j = 0; j < nb; ++j) {
          const int32_t dj = d[j];
          if (dj < 0) {
            const double ld = lam_dir[-(int64_t)dj - 1];
            if (ld != 0.0) v -= S[((int64_t)c * nb + j) * nb + i] * ld;
          }
        }
      }
      acc += v;
    }
    rhs[t] = acc;
  }
}

// Per-cell element systems (debug / parity entry point).
template <int PHYS>
__global__ void cell_systems_kernel(FusedGen<PHYS> gen, int64_t first, int64_t count, double* A,
                                    double* b) {
  __shared__ FusedShared sh;
  const int n = gen.T.n;
  for (int64_t lc = blockIdx.x; lc < count; lc += gridDim.x) {
    const int64_t cell = first + lc;
    gen.begin(cell, sh);
    __syncthreads();
    for (int t = threadIdx.x; t < n * (n + 1); t += blockDim.x) {
      const int col = t / n, row = t - col * n;
      const double v = gen.entry(sh, row, col);
      if (col < n)
        A[(lc * n + col) * n + row] = v;
      else
        b[lc * n + row] = v;
    }
    __syncthreads();
  }
}

// --------------------------------------------------------------------------
// Mixed precision.  W = A_ii^-1 [A_ib b_i] is computed from an FP32
// Gauss-Jordan inverse X of A_ii and refined in FP64:
//   W_0 = X B,   R_t = B - A_ii W_t (FP64),   W_{t+1} = W_t + X R_t (FP32
// products, FP64 update), until the correction reaches FP64 rounding or
// stops shrinking; then S = A_bb - A_bi W and g = b_b - A_bi w in FP64.
// W (natural interior order) is kept for recovery: x_i = w - W_b lambda.
// --------------------------------------------------------------------------
constexpr int kMixedMaxNi = 128;
constexpr int kRecoverSweeps = 8;  // FP64 refinement sweeps of the mixed recovery, at most
constexpr int kRecoverMaxNb = 512;
__host__ __device__ inline int cdiv(int a, int b) { return (a + b - 1) / b; }
constexpr int64_t kMixedChunkCells = 16384;  // cells per gen / inverse / refine round

// Cell systems in FP64, [cell][col][row] with n rows and n+1 columns (the last
// column is b), for cells [first, first+count).
template <int PHYS>
__global__ void __launch_bounds__(256)
    gen_system_kernel(FusedGen<PHYS> gen, int64_t first, int64_t count, double* __restrict__ A64) {
  __shared__ FusedShared sh;
  const int n = gen.T.n;
  const int per = n * (n + 1);
  for (int64_t lc = blockIdx.x; lc < count; lc += gridDim.x) {
    gen.begin(first + lc, sh);
    __syncthreads();
    double* A = A64 + lc * (int64_t)per;
    for (int t = threadIdx.x; t < per; t += blockDim.x) {
      const int col = t / n, row = t - col * n;
      A[t] = gen.entry(sh, row, col);
    }
    __syncthreads();
  }
}

// FP32 Gauss-Jordan inverse of A_ii with partial pivoting (largest magnitude,
// lowest row on ties; rows are not moved).  A 16x16 thread grid holds the
// matrix in registers: thread (tx, ty) owns rows tx + 16a and columns
// ty + 16b.  Per step, the pivot column and the pivot row go through LDS.
// Output X [cell][ni][ni] row-major with X = A_ii^-1; info[cell] = first zero
// pivot step + 1 (0 if none).
template <int RA>
__global__ void __launch_bounds__(256)
    gj_inverse_kernel(const double* __restrict__ A64, int n, int ni, int64_t count,
                      float* __restrict__ X, int32_t* __restrict__ info) {
  __shared__ float s_col[2][kMixedMaxNi];
  __shared__ float s_row[kMixedMaxNi];
  __shared__ int s_piv[kMixedMaxNi], s_inv[kMixedMaxNi];
  __shared__ unsigned char s_used[kMixedMaxNi];
  const int tx = threadIdx.x & 15, ty = threadIdx.x >> 4, lane = threadIdx.x & 63;
  for (int64_t lc = blockIdx.x; lc < count; lc += gridDim.x) {
    const double* A = A64 + lc * (int64_t)n * (n + 1);
    float m[RA][RA];
#pragma unroll
    for (int a = 0; a < RA; ++a)
#pragma unroll
      for (int b = 0; b < RA; ++b) {
        const int i = tx + 16 * a, j = ty + 16 * b;
        m[a][b] = (i < ni && j < ni) ? (float)A[(int64_t)j * n + i] : 0.f;
      }
    if (threadIdx.x < kMixedMaxNi) s_used[threadIdx.x] = 0;
    if (ty == 0) {
#pragma unroll
      for (int a = 0; a < RA; ++a)
        if (tx + 16 * a < ni) s_col[0][tx + 16 * a] = m[a][0];
    }
    __syncthreads();
    int bad = 0;
    for (int k = 0; k < ni; ++k) {
      const float* col = s_col[k & 1];
      // pivot search, repeated by every wavefront (no broadcast needed)
      int64_t best = -1;
      for (int i = lane; i < ni; i += 64) {
        if (s_used[i]) continue;
        const int64_t key =
            ((int64_t)__float_as_uint(fabsf(col[i])) << 32) | (int64_t)(0xFFFF - i);
        best = key > best ? key : best;
      }
      best = wave_max_i64(best);
      const int p = 0xFFFF - (int)(best & 0xFFFF);
      if ((best >> 32) == 0 && !bad) bad = k + 1;
      if (tx == (p & 15)) {
        const int ap = p >> 4;
#pragma unroll
        for (int b = 0; b < RA; ++b) {
          float v = 0.f;
#pragma unroll
          for (int a = 0; a < RA; ++a)
            if (a == ap) v = m[a][b];
          if (ty + 16 * b < ni) s_row[ty + 16 * b] = v;
        }
      }
      __syncthreads();
      if (threadIdx.x == 0) {
        s_piv[k] = p;
        s_inv[p] = k;
        s_used[p] = 1;
      }
      const float pv = s_row[k];
      const float inv = pv != 0.f ? 1.f / pv : 0.f;
      float c[RA], r[RA];
#pragma unroll
      for (int a = 0; a < RA; ++a) c[a] = tx + 16 * a < ni ? col[tx + 16 * a] : 0.f;
#pragma unroll
      for (int b = 0; b < RA; ++b) r[b] = ty + 16 * b < ni ? s_row[ty + 16 * b] * inv : 0.f;
#pragma unroll
      for (int a = 0; a < RA; ++a) {
        const bool prow = tx + 16 * a == p;
#pragma unroll
        for (int b = 0; b < RA; ++b) {
          const bool pcol = ty + 16 * b == k;
          // other rows: m_ij - m_ik m_pj / m_pk, column k: -m_ik / m_pk;
          // pivot row: m_pj / m_pk, pivot: 1 / m_pk
          float v = pcol ? -c[a] * inv : fmaf(-c[a], r[b], m[a][b]);
          if (prow) v = pcol ? inv : r[b];
          m[a][b] = v;
        }
      }
      // publish column k+1 for the next step
      if (k + 1 < ni && ty == ((k + 1) & 15)) {
        const int bk = (k + 1) >> 4;
#pragma unroll
        for (int a = 0; a < RA; ++a) {
          float v = 0.f;
#pragma unroll
          for (int b = 0; b < RA; ++b)
            if (b == bk) v = m[a][b];
          if (tx + 16 * a < ni) s_col[(k + 1) & 1][tx + 16 * a] = v;
        }
      }
      __syncthreads();
    }
    // m[i][c] holds (A^-1)[inv(i)][piv(c)]
    float* Xc = X + lc * (int64_t)ni * ni;
#pragma unroll
    for (int a = 0; a < RA; ++a)
#pragma unroll
      for (int b = 0; b < RA; ++b) {
        const int i = tx + 16 * a, c = ty + 16 * b;
        if (i < ni && c < ni) Xc[(int64_t)s_inv[i] * ni + s_piv[c]] = m[a][b];
      }
    if (threadIdx.x == 0 && info) info[lc] = bad;
    __syncthreads();
  }
}

// FP64 refinement of W and the FP64 Schur complement, one cell per 512-thread
// workgroup.  Threads form column groups of RP = 64 ceil(ni/64) rows (whole
// wavefronts, so a wavefront reads the same W entries: LDS broadcasts): thread
// t works on interior row t % RP and the JC contiguous columns of group
// t / RP.  W is kept row-major with ld = ng JC >= nb+1 (padding columns stay
// zero, so the inner loops need no guards), in LDS when WLDS and in the
// workspace otherwise; X and the FP32 residual are in LDS.  The final W goes
// to its output slot ([nb+1][ni] per cell).  iters[cell] (may be null) gets
// the refinement sweeps used, negated when the last correction was still
// above 1e-10 of W (not converged).
constexpr int kRefineThreads = 512;

template <int JC, bool WLDS>
__global__ void __launch_bounds__(kRefineThreads)
    mixed_refine_kernel(const double* __restrict__ A64, const float* __restrict__ X,
                        double* __restrict__ Wws, int ni, int nb, int64_t count,
                        CondenseOut out, int maxit, int32_t* __restrict__ iters) {
  constexpr int NT = kRefineThreads, NW = NT / 64;
  extern __shared__ double s_dyn[];
  const int n = ni + nb, ncol = nb + 1, ldx = ni + 1;
  const int ng = cdiv(ncol, JC), ldr = ng * JC;
  const int RP = 64 * cdiv(ni, 64);
  double* sW = s_dyn;                                                  // [ni][ldr] (WLDS)
  float* sX = reinterpret_cast<float*>(s_dyn + (WLDS ? ni * ldr : 0));  // [ni][ni+1]
  float* sR = sX + ni * ldx;                                           // [ni][ldr]
  __shared__ double s_red[2][NW];
  const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
  const int r = tid % RP, g = tid / RP, j0 = g * JC;
  const bool act = r < ni && g < ng;
  for (int64_t lc = blockIdx.x; lc < count; lc += gridDim.x) {
    const double* A = A64 + lc * (int64_t)n * (n + 1);
    const float* Xc = X + lc * (int64_t)ni * ni;
    double* W = WLDS ? sW : Wws + (size_t)blockIdx.x * ni * ldr;  // W[k][j] at k*ldr + j
    for (int t = tid; t < ni * ni; t += NT) {
      const int rr = t / ni, cc = t - rr * ni;
      sX[rr * ldx + cc] = Xc[t];
    }
    // B = [A_ib b_i] (zero in the padding columns); R_0 = B
    if (act) {
#pragma unroll
      for (int jj = 0; jj < JC; ++jj) {
        const int j = j0 + jj;
        sR[r * ldr + j] = j < ncol ? (float)A[(int64_t)(ni + j) * n + r] : 0.f;
      }
    }
    __syncthreads();
    double dprev = 0.0;
    int it = 0, conv = 1;
    for (;;) {
      // W += X R   (W_0 = X B)
      double dmax = 0.0, wmax = 0.0;
      if (act) {
        float d[JC];
#pragma unroll
        for (int jj = 0; jj < JC; ++jj) d[jj] = 0.f;
        const float* xr = sX + r * ldx;
        const float* rr = sR + j0;
        for (int k = 0; k < ni; ++k) {
          const float xv = xr[k];
#pragma unroll
          for (int jj = 0; jj < JC; ++jj) d[jj] = fmaf(xv, rr[k * ldr + jj], d[jj]);
        }
        double* wr = W + r * ldr + j0;
#pragma unroll
        for (int jj = 0; jj < JC; ++jj) {
          const double w = (it == 0 ? 0.0 : wr[jj]) + (double)d[jj];
          wr[jj] = w;
          dmax = fmax(dmax, fabs((double)d[jj]));
          wmax = fmax(wmax, fabs(w));
        }
      }
      // workgroup max of |correction| and |W|
      dmax = __builtin_bit_cast(double, wave_max_i64(__builtin_bit_cast(int64_t, dmax)));
      wmax = __builtin_bit_cast(double, wave_max_i64(__builtin_bit_cast(int64_t, wmax)));
      if (lane == 0) {
        s_red[0][wv] = dmax;
        s_red[1][wv] = wmax;
      }
      __syncthreads();  // also publishes W for the residual
      dmax = s_red[0][0];
      wmax = s_red[1][0];
#pragma unroll
      for (int q = 1; q < NW; ++q) {
        dmax = fmax(dmax, s_red[0][q]);
        wmax = fmax(wmax, s_red[1][q]);
      }
      // converged to FP64 rounding, or stalled at the FP64 floor of the
      // residual (a stall above 1e-10 of W is not accepted: transient growth of
      // the correction is possible while I - X A_ii is far from normal)
      const bool done = it >= maxit || dmax <= 0x1p-44 * wmax ||
                        (it >= 2 && dmax <= 1e-10 * wmax && dmax > 0.5 * dprev);
      if (done) {
        conv = dmax <= 1e-10 * wmax;
        break;
      }
      dprev = dmax;
      ++it;
      // R = B - A_ii W  (FP64)
      double acc[JC];
      if (act) {
#pragma unroll
        for (int jj = 0; jj < JC; ++jj) {
          const int j = j0 + jj;
          acc[jj] = j < ncol ? A[(int64_t)(ni + j) * n + r] : 0.0;
        }
        const double* wk = W + j0;
        for (int k = 0; k < ni; ++k) {
          const double a = A[(int64_t)k * n + r];
#pragma unroll
          for (int jj = 0; jj < JC; ++jj) acc[jj] = fma(-a, wk[k * ldr + jj], acc[jj]);
        }
      }
      __syncthreads();  // everyone is done with sR (and s_red) before they change
      if (act) {
#pragma unroll
        for (int jj = 0; jj < JC; ++jj) sR[r * ldr + j0 + jj] = (float)acc[jj];
      }
      __syncthreads();
    }
    // S = A_bb - A_bi W, g = b_b - A_bi w  (FP64): boundary row per thread, a
    // wavefront-uniform strided set of columns
    {
      const int RB = 64 * cdiv(nb, 64), ngb = NT / RB;
      const int rb = tid % RB, gb = tid / RB;
      if (rb < nb) {
        for (int j = gb; j < ncol; j += ngb) {
          double v = A[(int64_t)(ni + j) * n + ni + rb];
          for (int k = 0; k < ni; ++k) v = fma(-A[(int64_t)k * n + ni + rb], W[k * ldr + j], v);
          if (j < nb) {
            if (out.S) out.S[(lc * nb + j) * nb + rb] = v;
          } else {
            out.g[lc * nb + rb] = v;
          }
        }
      }
    }
    // W to its output slot, column-major [nb+1][ni]
    for (int t = tid; t < ncol * ni; t += NT) {
      const int j = t / ni, k = t - j * ni;
      out.Y[lc * (int64_t)ncol * ni + t] = W[k * ldr + j];
    }
    if (tid == 0 && iters) iters[lc] = conv ? it : -it - 1;
    __syncthreads();
  }
}

// --------------------------------------------------------------------------
// Kernel clock: while enabled, every launch outside stream capture is
// bracketed by two HIP events on its own stream, and the elapsed times are
// summed per kernel name when the clock is read.
// --------------------------------------------------------------------------
namespace {
struct ClockMark {
  const char* name;
  hipEvent_t a, b;
};
std::vector<ClockMark>* g_clock = nullptr;

struct ClockScope {
  hipStream_t st;
  ClockMark mark{nullptr, nullptr, nullptr};
  ClockScope(hipStream_t s, const char* name) : st(s) {
    if (!g_clock) return;
    hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
    if (hipStreamIsCapturing(s, &cs) != hipSuccess || cs != hipStreamCaptureStatusNone) return;
    if (hipEventCreate(&mark.a) != hipSuccess) return;
    if (hipEventCreate(&mark.b) != hipSuccess) {
      (void)hipEventDestroy(mark.a);
      return;
    }
    mark.name = name;
    (void)hipEventRecord(mark.a, s);
  }
  ~ClockScope() {
    if (!mark.name) return;
    (void)hipEventRecord(mark.b, st);
    g_clock->push_back(mark);
  }
};
}  // namespace

void clock_start() {
  if (!g_clock) g_clock = new std::vector<ClockMark>();
}

int clock_stop(std::vector<KernelTime>& out) {
  out.clear();
  if (!g_clock) return 0;
  int bad = 0;
  for (ClockMark& m : *g_clock) {
    float ms = 0.f;
    if (hipEventSynchronize(m.b) != hipSuccess || hipEventElapsedTime(&ms, m.a, m.b) != hipSuccess)
      ++bad;
    (void)hipEventDestroy(m.a);
    (void)hipEventDestroy(m.b);
    auto it = std::find_if(out.begin(), out.end(),
                           [&](const KernelTime& k) { return k.name == m.name; });
    if (it == out.end()) {
      out.push_back(KernelTime{m.name, 0, 0.0});
      it = out.end() - 1;
    }
    it->launches++;
    it->total_ms += ms;
  }
  delete g_clock;
  g_clock = nullptr;
  return bad;
}

// --------------------------------------------------------------------------
// Launchers
// --------------------------------------------------------------------------
struct Bucket {
  int pr, pc, r, cc;
};

// Thread-grid / register-tile buckets by rows n and columns n+1.
static bool pick_bucket(int n, Bucket& b) {
  const Bucket list[] = {{8, 8, 4, 4}, {8, 8, 6, 6}, {8, 8, 9, 9}, {32, 16, 6, 11}};
  for (const auto& x : list)
    if (x.pr * x.r >= n && x.pc * x.cc >= n + 1) {
      b = x;
      return true;
    }
  return false;
}

static int grid_for(int64_t ncells, int threads) {
  // enough workgroups to fill 256 CUs several times over
  int64_t target = threads >= 512 ? 256 * 3 : 256 * 16;
  if (threads <= 64) target = 256 * 40;
  return (int)(ncells < target ? ncells : target);
}

constexpr int kLargeGrid = 256 * 2;

// Bytes of the large-system kernel's workspace: one column-major (n+1)-column
// system per resident workgroup, leading dimension rounded up to even.
// wavefronts per workgroup of the blocked kernel, measured fastest per size
// wavefronts per workgroup of the blocked elimination (100 + w: 32-column panels)
static int large_default_waves(int n) {
  return n > 256 ? (n <= kLarge32MaxN ? 116 : 16) : 8;
}

static size_t large_bytes(int n, int* ld) {
  const int l = (n + 1) & ~1;
  if (ld) *ld = l;
  return (size_t)kLargeGrid * l * (size_t)(n + 1) * sizeof(double);
}

constexpr int kPanelMaxN = 512;

static int panel_nmax(int n) { return n <= 192 ? 192 : (n <= 256 ? 256 : 512); }

// Row-major workspace of the panel kernel: n rows, ld = n+1 rounded to even,
// one system per resident workgroup.
static size_t panel_bytes(int n, int grid, int* ld) {
  const int l = (n + 2) & ~1;
  if (ld) *ld = l;
  return (size_t)grid * (size_t)n * l * sizeof(double);
}

static int panel_grid(int n, int64_t ncells, int gridmul) {
  const int per = gridmul > 0 ? gridmul : (panel_nmax(n) <= 256 ? 2 : 1);
  return (int)std::min<int64_t>(ncells, 256 * per);
}

// free device memory kept beyond an allocation before cached blocks are returned
constexpr size_t kDevFreeMargin = size_t(8) << 30;

namespace {
// Per-context block pools.  A context's buffers are only ever touched by work
// on the context's one stream (its kernels, copies and graph replays), so a
// block freed into the context's pool can be handed to the context's next
// allocation at once: that allocation's work is queued behind whatever still
// used the block.  No pool is shared between contexts.
struct Pool {
  int device = 0;
  bool live = false;
  std::multimap<size_t, void*> free;  // cached blocks by the size they were requested with
  int64_t live_bytes = 0, pooled_bytes = 0;
};
// Blocks of a device that no queued work uses any more (their context's
// stream was synchronized after they were freed): any context may reuse them.
struct IdleBlocks {
  std::multimap<size_t, void*> free;
  int64_t bytes = 0;
};
struct DevCache {
  std::mutex mu;
  std::vector<Pool> pools;
  std::map<int, IdleBlocks> idle;  // by device
  int64_t unpooled_bytes = 0;
};
DevCache& dev_cache() {
  static DevCache c;
  return c;
}
thread_local int tl_pool = -1;
// hipFree of every cached block of a device; caller holds the lock.  The
// runtime hands freed memory back once the device has drained, so a flush
// that freed anything waits for the device: the free memory it reports
// afterwards is what an allocation can get.
void flush_device(DevCache& c, int device) {
  bool any = false;
  for (Pool& pl : c.pools) {
    if (pl.device != device) continue;
    for (auto& kv : pl.free) (void)hipFree(kv.second);
    any = any || !pl.free.empty();
    pl.free.clear();
    pl.pooled_bytes = 0;
  }
  IdleBlocks& ib = c.idle[device];
  for (auto& kv : ib.free) (void)hipFree(kv.second);
  any = any || !ib.free.empty();
  ib.free.clear();
  ib.bytes = 0;
  if (any) (void)hipDeviceSynchronize();
  (void)hipGetLastError();
}
}  // namespace

PoolScope::PoolScope(int pool) : prev(tl_pool) { tl_pool = pool; }
PoolScope::~PoolScope() { tl_pool = prev; }

int dev_pool_create(int device) {
  DevCache& c = dev_cache();
  std::lock_guard<std::mutex> lock(c.mu);
  c.pools.emplace_back();
  c.pools.back().device = device;
  c.pools.back().live = true;
  return (int)c.pools.size() - 1;
}

// The context's stream is idle (synchronized) when its pool goes away.
void dev_pool_destroy(int pool) {
  DevCache& c = dev_cache();
  std::lock_guard<std::mutex> lock(c.mu);
  if (pool < 0 || pool >= (int)c.pools.size()) return;
  Pool& pl = c.pools[pool];
  for (auto& kv : pl.free) (void)hipFree(kv.second);
  pl.free.clear();
  pl.pooled_bytes = 0;
  pl.live = false;
  (void)hipGetLastError();
}

void dev_pool_idle(int pool) {
  DevCache& c = dev_cache();
  std::lock_guard<std::mutex> lock(c.mu);
  if (pool < 0 || pool >= (int)c.pools.size() || !c.pools[pool].live) return;
  Pool& pl = c.pools[pool];
  IdleBlocks& ib = c.idle[pl.device];
  for (auto& kv : pl.free) ib.free.emplace(kv.first, kv.second);
  ib.bytes += pl.pooled_bytes;
  pl.free.clear();
  pl.pooled_bytes = 0;
}

void dev_pool_stats(int pool, int64_t* live, int64_t* pooled, int64_t* unpooled) {
  DevCache& c = dev_cache();
  std::lock_guard<std::mutex> lock(c.mu);
  const bool ok = pool >= 0 && pool < (int)c.pools.size();
  *live = ok ? c.pools[pool].live_bytes : 0;
  *pooled = ok ? c.pools[pool].pooled_bytes + c.idle[c.pools[pool].device].bytes : 0;
  *unpooled = c.unpooled_bytes;
}

void dev_cache_flush(int device) {
  DevCache& c = dev_cache();
  std::lock_guard<std::mutex> lock(c.mu);
  flush_device(c, device);
}

hipError_t dev_alloc(void** p, size_t bytes, int* pool_out) {
  *p = nullptr;
  const int pool = tl_pool;
  *pool_out = -1;
  if (bytes == 0) return hipSuccess;
  DevCache& c = dev_cache();
  int device = 0;
  {
    std::lock_guard<std::mutex> lock(c.mu);
    if (pool >= 0 && pool < (int)c.pools.size() && c.pools[pool].live) {
      Pool& pl = c.pools[pool];
      device = pl.device;
      auto it = pl.free.lower_bound(bytes);
      if (it != pl.free.end() && it->first / 2 <= bytes) {
        *p = it->second;
        pl.pooled_bytes -= (int64_t)it->first;
        pl.live_bytes += (int64_t)bytes;
        pl.free.erase(it);
        *pool_out = pool;
        return hipSuccess;
      }
      IdleBlocks& ib = c.idle[device];
      auto jt = ib.free.lower_bound(bytes);
      if (jt != ib.free.end() && jt->first / 2 <= bytes) {
        *p = jt->second;
        ib.bytes -= (int64_t)jt->first;
        pl.live_bytes += (int64_t)bytes;
        ib.free.erase(jt);
        *pool_out = pool;
        return hipSuccess;
      }
      // keep a margin of free device memory for the runtime and other users
      // (the runtime's own allocations, such as scratch, fail hard): the
      // device's cached blocks go back first when the allocation would eat
      // into it, and an allocation that would still eat into it is refused
      size_t fr = 0, tot = 0;
      if (hipMemGetInfo(&fr, &tot) == hipSuccess && fr < bytes + kDevFreeMargin) {
        flush_device(c, device);
        if (hipMemGetInfo(&fr, &tot) == hipSuccess && fr < bytes + kDevFreeMargin) {
          (void)hipGetLastError();
          return hipErrorOutOfMemory;
        }
      }
      (void)hipGetLastError();
    }
  }
  hipError_t e = hipMalloc(p, bytes);
  if (e != hipSuccess) {
    (void)hipGetLastError();
    {
      std::lock_guard<std::mutex> lock(c.mu);
      flush_device(c, device);
    }
    e = hipMalloc(p, bytes);
  }
  if (e != hipSuccess) {
    (void)hipGetLastError();
    *p = nullptr;
    return e;
  }
  std::lock_guard<std::mutex> lock(c.mu);
  if (pool >= 0 && pool < (int)c.pools.size() && c.pools[pool].live) {
    *pool_out = pool;
    c.pools[pool].live_bytes += (int64_t)bytes;
  } else {
    c.unpooled_bytes += (int64_t)bytes;
  }
  return hipSuccess;
}

// The block keeps the size it was requested with: a later request of at most
// that size and at least half of it, from the same pool, reuses it.
void dev_free(void* p, size_t bytes, int pool) {
  if (!p) return;
  DevCache& c = dev_cache();
  {
    std::lock_guard<std::mutex> lock(c.mu);
    if (pool >= 0 && pool < (int)c.pools.size() && c.pools[pool].live) {
      Pool& pl = c.pools[pool];
      pl.free.emplace(bytes, p);
      pl.live_bytes -= (int64_t)bytes;
      pl.pooled_bytes += (int64_t)bytes;
      return;
    }
    if (pool < 0) c.unpooled_bytes -= (int64_t)bytes;
  }
  (void)hipFree(p);
  (void)hipGetLastError();
}

hipError_t workspace_reserve(Workspace& w, size_t bytes) {
  if (bytes <= w.bytes) return hipSuccess;
  workspace_release(w);
  void* p = nullptr;
  hipError_t e = dev_alloc(&p, bytes, &w.pool);
  if (e != hipSuccess) return e;
  w.p = static_cast<double*>(p);
  w.bytes = bytes;
  return hipSuccess;
}

void workspace_release(Workspace& w) {
  dev_free(w.p, w.bytes, w.pool);
  w.p = nullptr;
  w.bytes = 0;
  w.pool = -1;
}

// Kernel families for a local system size (see condense_variants()).
enum Family {
  kFamWave = 0,
  kFamGrid8 = 1,
  // 2: a column-set-per-wavefront elimination, and 6: a register-resident
  // MFMA tile elimination, were autotuner candidates that never won; their
  // numbers stay reserved
  kFamGrid32 = 3,
  kFamLarge = 4,
  kFamPanel = 5,
  kFamGrid16 = 7,
  kFamFlux = kFamilyFluxSchur,
  kFamRef = kFamilyFluxRef,
  kFamCount = 11
};

static bool wave_class(int ni, int nb) {
  return (ni == 12 && nb == 8) || (ni == 27 && nb == 12) || (ni == 48 && nb == 16) ||
         (ni == 32 && nb == 24) || (ni == 7 && nb == 8);
}

static bool family_ok(int fam, int ni, int nb) {
  const int n = ni + nb;
  Bucket b;
  switch (fam) {
    case kFamWave: return wave_class(ni, nb);
    case kFamGrid8: return pick_bucket(n, b) && b.pr == 8;
    case kFamGrid32: return n <= 192 && n + 1 <= 176 && n > 71;
    case kFamLarge: return n > 64 && n <= kLargeMaxN;
    case kFamPanel: return n > 64 && n <= kPanelMaxN;
    case kFamGrid16: return n >= 48 && n + 1 <= 96;
  }
  return false;  // kFamFlux depends on the tables: launch_condense_fused
}

// default choice: the measured-fastest family for the size
static int default_family(int ni, int nb) {
  const int n = ni + nb;
  if (wave_class(ni, nb)) return kFamWave;
  if (n + 1 <= 72) return kFamGrid8;
  if (n > 128 && n <= 192 && n + 1 <= 176) return kFamGrid32;
  if (n <= 256) return kFamPanel;
  return kFamLarge;
}

template <class Gen, bool BACKSUB>
static hipError_t launch_condense(const Gen& gen, const CondenseOut& out, int64_t ncells, int ni,
                                  int nb, hipStream_t st, Workspace& ws, int fam = -1,
                                  int gridmul = 0) {
  if (ncells <= 0) return hipSuccess;
  if (fam < 0) fam = default_family(ni, nb);
  if (!family_ok(fam, ni, nb)) return hipErrorInvalidValue;
  auto grid_of = [&](int per_cu_default) {
    const int64_t per = gridmul > 0 ? gridmul : per_cu_default;
    const int64_t target = 256 * per;
    return (int)(ncells < target ? ncells : target);
  };
  if (fam == kFamLarge) {
    // gridmul selects the wavefronts per workgroup (one workgroup per CU: the
    // panel and pivot-row LDS take most of it): 4 (default), 8 or 16
    int ld;
    hipError_t e = workspace_reserve(ws, large_bytes(ni + nb, &ld));
    if (e != hipSuccess) return e;
    const int grid = (int)(ncells < kLargeGrid ? ncells : kLargeGrid);
    int waves = gridmul > 0 ? gridmul : large_default_waves(ni + nb);
    // 100 + waves: 32-column panels (systems up to kLarge32MaxN rows)
    const bool kb32 = waves > 100 && ni + nb <= kLarge32MaxN;
    if (waves > 100) waves -= 100;
    ClockScope clk_(st, "condense_blocked");
    if (kb32) {
      if (waves == 16)
        hipLaunchKernelGGL((condense_large_kernel<Gen, BACKSUB, 1024, 32, kLarge32MaxN>),
                           dim3(grid), dim3(1024), 0, st, gen, out, ncells, ni, nb, ws.p, ld);
      else
        hipLaunchKernelGGL((condense_large_kernel<Gen, BACKSUB, 512, 32, kLarge32MaxN>),
                           dim3(grid), dim3(512), 0, st, gen, out, ncells, ni, nb, ws.p, ld);
    } else if (waves == 16)
      hipLaunchKernelGGL((condense_large_kernel<Gen, BACKSUB, 1024>), dim3(grid), dim3(1024), 0,
                         st, gen, out, ncells, ni, nb, ws.p, ld);
    else if (waves == 8)
      hipLaunchKernelGGL((condense_large_kernel<Gen, BACKSUB, 512>), dim3(grid), dim3(512), 0, st,
                         gen, out, ncells, ni, nb, ws.p, ld);
    else
      hipLaunchKernelGGL((condense_large_kernel<Gen, BACKSUB, 256>), dim3(grid), dim3(256), 0, st,
                         gen, out, ncells, ni, nb, ws.p, ld);
    return hipGetLastError();
  }
  if (fam == kFamPanel) {
    const int n = ni + nb, grid = panel_grid(n, ncells, gridmul);
    int ld;
    hipError_t e = workspace_reserve(ws, panel_bytes(n, grid, &ld));
    if (e != hipSuccess) return e;
    const int nmax = panel_nmax(n);
#define HDG_PANEL(NM_)                                                                        \
  if (nmax == NM_) {                                                                          \
    ClockScope clk_(st, "condense_panel");                                                    \
    hipLaunchKernelGGL((condense_panel_kernel<NM_, Gen, BACKSUB>), dim3(grid), dim3(256), 0,  \
                       st, gen, out, ncells, ni, nb, ws.p, ld);                               \
    return hipGetLastError();                                                                 \
  }
    HDG_PANEL(192)
    HDG_PANEL(256)
    HDG_PANEL(512)
#undef HDG_PANEL
    return hipErrorInvalidValue;
  }
  if (fam == kFamWave) {
    const int grid = grid_of(48);
#define HDG_WAVE(NI_, NB_)                                                                  \
  if (ni == NI_ && nb == NB_) {                                                             \
    { ClockScope clk_(st, "condense_wave"); hipLaunchKernelGGL((condense_wave_kernel<NI_, NB_, Gen, BACKSUB>), dim3(grid), dim3(64), \
                       0, st, gen, out, ncells); }                                            \
    return hipGetLastError();                                                               \
  }
    HDG_WAVE(12, 8)
    HDG_WAVE(27, 12)
    HDG_WAVE(48, 16)
    HDG_WAVE(32, 24)
    HDG_WAVE(7, 8)
#undef HDG_WAVE
    return hipErrorInvalidValue;
  }
  Bucket b;
  if (fam == kFamGrid32) {
    b = Bucket{32, 16, 6, 11};
  } else if (fam == kFamGrid16) {
    b = Bucket{16, 16, 6, 6};
  } else if (!pick_bucket(ni + nb, b)) {
    return hipErrorInvalidValue;
  }
  const int grid = gridmul > 0 ? grid_of(gridmul) : grid_for(ncells, b.pr * b.pc);
#define HDG_LAUNCH(PR, PC, R, CC)                                                             \
  if (b.pr == PR && b.pc == PC && b.r == R && b.cc == CC) {                                   \
    { ClockScope clk_(st, "condense_grid"); hipLaunchKernelGGL((condense_kernel<PR, PC, R, CC, Gen, BACKSUB>), dim3(grid),             \
                       dim3(PR * PC), 0, st, gen, out, ncells, ni, nb); }                       \
    return hipGetLastError();                                                                 \
  }
  HDG_LAUNCH(8, 8, 4, 4)
  HDG_LAUNCH(8, 8, 6, 6)
  HDG_LAUNCH(8, 8, 9, 9)
  HDG_LAUNCH(32, 16, 6, 11)
  HDG_LAUNCH(16, 16, 6, 6)
#undef HDG_LAUNCH
  return hipErrorInvalidValue;
}

int condense_variants(const SysShape& sh, int* fams, int max) {
  int k = 0;
  for (int f = 0; f < kFamCount && k < max; ++f)
    if (family_ok(f, sh.ni, sh.nb) || (f == kFamFlux && sh.flux) ||
        (f == kFamRef && sh.ref))
      fams[k++] = f;
  return k;
}

int condense_grid_options(int fam, int* gm, int max) {
  // resident workgroups per CU worth timing for each family (0 = fixed grid)
  static const int wave[] = {16, 32, 48, 64}, grid8[] = {20, 40, 80},
                   g32[] = {2, 3, 4}, pan[] = {1, 2},
                   large[] = {4, 8, 16, 108, 116},
                   g16[] = {2, 4, 6}, fixed[] = {0};
  const int* v = fixed;
  int n = 1;
  switch (fam) {
    case kFamWave: v = wave, n = 4; break;
    case kFamGrid8: v = grid8, n = 3; break;
    case kFamGrid32: v = g32, n = 3; break;
    case kFamPanel: v = pan, n = 2; break;
    case kFamLarge: v = large, n = 5; break;
    case kFamGrid16: v = g16, n = 3; break;
    default: break;
  }
  int k = 0;
  for (; k < n && k < max; ++k) gm[k] = v[k];
  return k;
}

static size_t flux_workspace_bytes(int nr, int64_t ncells);

int condense_default_family(const SysShape& sh) {
  // small systems stay with the register-resident kernels (one wavefront or
  // an 8x8 grid per cell), which no two-stage scheme beats
  if ((sh.flux || sh.ref) && !wave_class(sh.ni, sh.nb) && sh.ni + sh.nb > 96)
    return sh.ref ? kFamRef : kFamFlux;
  return default_family(sh.ni, sh.nb);
}

size_t condense_workspace_bytes(const SysShape& sh, int fam, int64_t ncells) {
  const int ni = sh.ni, nb = sh.nb;
  if (fam < 0) fam = condense_default_family(sh);
  if (fam == kFamFlux || fam == kFamRef) return flux_workspace_bytes(ni - sh.flux_na + nb, ncells);
  if (fam == kFamPanel) {
    const int n = ni + nb;
    return panel_bytes(n, panel_grid(n, INT64_MAX, 0) * 2, nullptr);  // any grid multiplier <= 2x
  }
  return fam == kFamLarge ? large_bytes(ni + nb, nullptr) : 0;
}

bool condense_supported(int n) {
  Bucket b;
  return pick_bucket(n, b) || (n > 64 && n <= kLargeMaxN);
}

// Interior right-hand side b_i of every cell, one wavefront per cell: a lane
// per cell quadrature point evaluates |det J| times the source there, then a
// lane per row runs FusedGen::rhs (the quadrature of the source against the
// test basis); b_i is kept in the per-cell scratch, so the generators read one
// value per right-hand-side entry.
template <int PHYS>
__global__ void __launch_bounds__(256) cell_rhs_kernel(FusedGen<PHYS> gen) {
  constexpr int WPB = 4;
  __shared__ double s_src[WPB][kMaxQP][4];
  const Tables& T = gen.T;
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  for (int64_t cell = (int64_t)blockIdx.x * WPB + w; cell < gen.M.ncells;
       cell += (int64_t)gridDim.x * WPB) {
    double* sq = gen.M.srcq + cell * srcq_stride(T);
    for (int q = lane; q < T.nqc; q += 64) {
      double v[4];
      cell_source<PHYS>(T, gen.M, cell, q, v);
#pragma unroll
      for (int c = 0; c < 4; ++c) s_src[w][q][c] = v[c];
    }
    wave_sync_lds();
    for (int r = lane; r < T.ni; r += 64) sq[r] = gen.rhs(s_src[w], r);
    wave_sync_lds();
  }
}

static hipError_t launch_prep(const Tables& T, const MeshDev& M, hipStream_t st) {
  if (M.ncells <= 0) return hipSuccess;
  if (T.ni > kMaxRhsRows) return hipErrorInvalidValue;
  const int64_t blocks = (M.ncells + kPrepCells - 1) / kPrepCells;
  const int grid = (int)(blocks < 16384 ? blocks : 16384);
  {
    ClockScope clk_(st, "prep");
    if (T.phys == kDarcy)
      hipLaunchKernelGGL(prep_kernel<kDarcy>, dim3(grid), dim3(kPrepCells), 0, st, T, M);
    else
      hipLaunchKernelGGL(prep_kernel<kElasticity>, dim3(grid), dim3(kPrepCells), 0, st, T, M);
  }
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  const int64_t wblocks = (M.ncells + 3) / 4;
  const int wgrid = (int)(wblocks < 65536 ? wblocks : 65536);
  ClockScope clk_(st, "prep_rhs");
  if (T.phys == kDarcy)
    hipLaunchKernelGGL(cell_rhs_kernel<kDarcy>, dim3(wgrid), dim3(256), 0, st,
                       FusedGen<kDarcy>{T, M});
  else
    hipLaunchKernelGGL(cell_rhs_kernel<kElasticity>, dim3(wgrid), dim3(256), 0, st,
                       FusedGen<kElasticity>{T, M});
  return hipGetLastError();
}

hipError_t launch_prep_cells(const Tables& T, const MeshDev& M, hipStream_t st) {
  return launch_prep(T, M, st);
}

hipError_t launch_condense_fused(const Tables& T, const MeshDev& M, const CondenseOut& out,
                                 hipStream_t st, Workspace& ws, bool prep, int fam,
                                 int gridmul) {
  if (prep) {
    hipError_t e = launch_prep(T, M, st);
    if (e != hipSuccess) return e;
  }
  if (fam < 0) fam = condense_default_family(sys_shape(T));
  if (fam == kFamFlux) return launch_condense_flux(T, M, out, st, ws, gridmul, false);
  if (fam == kFamRef) return launch_condense_flux(T, M, out, st, ws, gridmul, true);
  if (T.phys == kDarcy) {
    FusedGen<kDarcy> gen{T, M};
    return launch_condense<FusedGen<kDarcy>, false>(gen, out, M.ncells, T.ni, T.nb, st, ws, fam,
                                                    gridmul);
  }
  FusedGen<kElasticity> gen{T, M};
  return launch_condense<FusedGen<kElasticity>, false>(gen, out, M.ncells, T.ni, T.nb, st, ws,
                                                       fam, gridmul);
}

hipError_t launch_gather(const double* src, const int64_t* idx, double* dst, int64_t n,
                         hipStream_t st) {
  if (n <= 0) return hipSuccess;
  const int64_t blocks = (n + 255) / 256;
  ClockScope clk_(st, "renumber");
  hipLaunchKernelGGL(gather_kernel, dim3((unsigned)(blocks < 4096 ? blocks : 4096)), dim3(256), 0,
                     st, src, idx, dst, n);
  return hipGetLastError();
}

// --------------------------------------------------------------------------
// Device skeleton solve (MINRES with a facet block-Jacobi preconditioner)
// on the assembled CSC.
// The skeleton matrices of both weak forms are symmetric (the condensed
// blocks S are), so column i of the CSC is read as row i: y = s K x is a
// gather, one wavefront per row, with a fixed reduction order.
// --------------------------------------------------------------------------
__global__ void __launch_bounds__(256) spmv_sym_kernel(const int64_t* __restrict__ colptr,
                                                       const int32_t* __restrict__ rowval,
                                                       const double* __restrict__ nzval,
                                                       const double* __restrict__ x,
                                                       double* __restrict__ y, int64_t n,
                                                       double scale) {
  const int lane = threadIdx.x & 63;
  for (int64_t i = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) >> 6; i < n;
       i += ((int64_t)gridDim.x * blockDim.x) >> 6) {
    const int64_t k0 = colptr[i], k1 = colptr[i + 1];
    double acc = 0.0;
    for (int64_t k = k0 + lane; k < k1; k += 64) acc = fma(nzval[k], x[rowval[k]], acc);
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) acc += __shfl_xor(acc, off, 64);
    if (lane == 0) y[i] = scale * acc;
  }
}

constexpr int kBlkMax = 32;  // largest diagonal block of the block-Jacobi preconditioner

// SPD preconditioner blocks: for each bs x bs diagonal block B of K (rows
// [blk bs, blk bs + bs)) whose diagonal has one sign s, the inverse of s B
// when s B is positive definite (Gauss-Jordan without pivoting on [s B | I]:
// all pivots positive); otherwise diag(1 / |B_ii|).  One wavefront per block
// in LDS, lane c owning column c.
__global__ void __launch_bounds__(128) block_inv_kernel(const int64_t* __restrict__ colptr,
                                                        const int32_t* __restrict__ rowval,
                                                        const double* __restrict__ nzval,
                                                        double* __restrict__ binv, int64_t nblk,
                                                        int bs) {
  __shared__ double s_a[2][kBlkMax][2 * kBlkMax + 1];
  __shared__ double s_d[2][kBlkMax];
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  double(*A)[2 * kBlkMax + 1] = s_a[w];
  double* dg = s_d[w];
  for (int64_t blk = (int64_t)blockIdx.x * 2 + w; blk < nblk; blk += (int64_t)gridDim.x * 2) {
    const int64_t r0 = blk * bs;
    if (lane < bs) {  // column r0 + lane of K restricted to the block's rows
      for (int i = 0; i < bs; ++i) A[i][lane] = 0.0;
      const int64_t c = r0 + lane;
      int64_t lo = colptr[c], hi = colptr[c + 1];
      while (lo < hi) {  // first row >= r0
        const int64_t mid = (lo + hi) >> 1;
        if (rowval[mid] < r0) lo = mid + 1;
        else hi = mid;
      }
      for (int64_t k = lo; k < colptr[c + 1] && rowval[k] < r0 + bs; ++k)
        A[rowval[k] - r0][lane] = nzval[k];
    } else if (lane < 2 * bs) {
      for (int i = 0; i < bs; ++i) A[i][lane] = i == lane - bs ? 1.0 : 0.0;
    }
    wave_sync_lds();
    int pos = 0, neg = 0;
    for (int i = 0; i < bs; ++i) {
      pos += A[i][i] > 0.0;
      neg += A[i][i] < 0.0;
    }
    if (lane < bs) dg[lane] = A[lane][lane];
    const double sg = pos == bs ? 1.0 : (neg == bs ? -1.0 : 0.0);
    wave_sync_lds();
    if (lane < bs)
      for (int i = 0; i < bs; ++i) A[i][lane] *= sg;
    wave_sync_lds();
    bool ok = sg != 0.0;
    for (int k = 0; k < bs && ok; ++k) {
      const double piv = A[k][k];
      if (!(piv > 0.0)) {  // s B is not positive definite
        ok = false;
        break;
      }
      wave_sync_lds();
      const double rinv = 1.0 / piv;
      if (lane < 2 * bs) A[k][lane] *= rinv;
      wave_sync_lds();
      for (int i = 0; i < bs; ++i) {
        if (i == k) continue;
        const double f = A[i][k];
        wave_sync_lds();
        if (lane < 2 * bs) A[i][lane] = fma(-f, A[k][lane], A[i][lane]);
        wave_sync_lds();
      }
    }
    if (lane >= bs && lane < 2 * bs) {
      const int j = lane - bs;
      const double dj = fabs(dg[j]);
      for (int i = 0; i < bs; ++i)
        binv[(blk * bs + i) * bs + j] =
            ok ? A[i][lane] : (i == j ? (dj > 0.0 ? 1.0 / dj : 1.0) : 0.0);
    }
    wave_sync_lds();
  }
}

// z = blockdiag(binv) r, a thread per row
__global__ void block_apply_kernel(const double* __restrict__ binv,
                                   const double* __restrict__ r, double* __restrict__ z,
                                   int64_t n, int bs) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x) {
    const int64_t r0 = i - i % bs;
    const double* b = binv + i * bs;
    double acc = 0.0;
    for (int j = 0; j < bs; ++j) acc = fma(b[j], r[r0 + j], acc);
    z[i] = acc;
  }
}

// MINRES vector updates; with sc non-null the coefficients are read from the
// device scalar state (MinresScalar indices in ia, ib, ic, id; < 0: 0):
// op 0: y = y - a u0                        (u0 = r1 or r2)
// op 1: w = (v - a w1 - b w2) c, x += d w   (u0 = w, u1 = x, w0 = v, w1 = w1, w2 = w2)
// op 2: u0 = a w0
// op 3: u0 = w0 - w1                        (the residual b - K x)
__global__ void krylov_update_kernel(int op, int64_t n, double a, double b, double c, double d,
                                     double* u0, double* u1, const double* w0,
                                     const double* w1, const double* w2, const double* sc,
                                     int ia, int ib, int ic, int id) {
  if (sc) {
    a = ia >= 0 ? sc[ia] : 0.0;
    b = ib >= 0 ? sc[ib] : 0.0;
    c = ic >= 0 ? sc[ic] : 0.0;
    d = id >= 0 ? sc[id] : 0.0;
  }
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x) {
    if (op == 0) {
      u1[i] = fma(-a, u0[i], u1[i]);
    } else if (op == 1) {
      const double wv = (w0[i] - a * w1[i] - b * w2[i]) * c;
      u0[i] = wv;
      u1[i] = fma(d, wv, u1[i]);
    } else if (op == 2) {
      u0[i] = a * w0[i];
    } else {
      u0[i] = w0[i] - w1[i];
    }
  }
}

constexpr int kDotBlocks = 1024;

// MINRES recurrences on the device (one thread): step 0 initialises from
// (r1, y) = beta1^2; step 1 takes alfa = (v, y); step 2 takes beta^2 = (r2, y)
// and advances the Givens rotation.  A zero beta (the Krylov space is
// exhausted) or a negative beta^2 (an indefinite preconditioner) freezes the
// state: every later update is then a no-op, and the host reads the flags.
__global__ void minres_scalar_kernel(int step, const double* __restrict__ dot,
                                     double* __restrict__ sc) {
  using M = MinresScalar;
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  const bool frozen = sc[M::kFrozen] != 0.0;
  if (step == 0) {
    const double b1 = dot[0] > 0.0 ? sqrt(dot[0]) : 0.0;
    sc[M::kBeta1] = b1;
    sc[M::kBeta] = b1;
    sc[M::kPhibar] = b1;
    sc[M::kOldb] = 0.0;
    sc[M::kDbar] = 0.0;
    sc[M::kEpsln] = 0.0;
    sc[M::kCs] = -1.0;
    sc[M::kSn] = 0.0;
    sc[M::kInvBeta] = b1 > 0.0 ? 1.0 / b1 : 0.0;
    sc[M::kFrozen] = b1 > 0.0 ? 0.0 : 1.0;
    sc[M::kIndefinite] = dot[0] < 0.0 ? 1.0 : 0.0;
  } else if (step == 1) {
    const double alfa = frozen ? 0.0 : dot[0];
    sc[M::kAlfa] = alfa;
    sc[M::kAlfaBeta] = frozen ? 0.0 : alfa / sc[M::kBeta];
  } else {
    if (frozen) {
      sc[M::kPhi] = 0.0;
      sc[M::kInvGamma] = 0.0;
      sc[M::kInvBeta] = 0.0;
      sc[M::kBetaOldb] = 0.0;
      sc[M::kAlfaBeta] = 0.0;
      return;
    }
    if (dot[0] < 0.0) sc[M::kIndefinite] = 1.0;
    const double alfa = sc[M::kAlfa], cs = sc[M::kCs], sn = sc[M::kSn], dbar = sc[M::kDbar];
    const double oldb = sc[M::kBeta];
    const double beta = dot[0] > 0.0 ? sqrt(dot[0]) : 0.0;
    const double oldeps = sc[M::kEpsln];
    const double delta = cs * dbar + sn * alfa;
    const double gbar = sn * dbar - cs * alfa;
    const double gamma = fmax(hypot(gbar, beta), 1e-300);
    const double cs2 = gbar / gamma, sn2 = beta / gamma;
    const double phibar = sc[M::kPhibar];
    sc[M::kOldb] = oldb;
    sc[M::kBeta] = beta;
    sc[M::kOldEps] = oldeps;
    sc[M::kDelta] = delta;
    sc[M::kEpsln] = sn * beta;
    sc[M::kDbar] = -cs * beta;
    sc[M::kCs] = cs2;
    sc[M::kSn] = sn2;
    sc[M::kPhi] = cs2 * phibar;
    sc[M::kPhibar] = sn2 * phibar;
    sc[M::kInvGamma] = 1.0 / gamma;
    sc[M::kInvBeta] = beta > 0.0 ? 1.0 / beta : 0.0;
    sc[M::kBetaOldb] = beta / oldb;
    if (!(beta > 0.0)) sc[M::kFrozen] = 1.0;
  }
}

// up to three dot products at once: per-block partial sums in a fixed order,
// then one block sums the partials (deterministic for a given n)
__global__ void __launch_bounds__(256) multidot_kernel(int nd, int64_t n, const double* a0,
                                                       const double* b0, const double* a1,
                                                       const double* b1, const double* a2,
                                                       const double* b2, double* partial) {
  __shared__ double s_red[3][256];
  double acc[3] = {0.0, 0.0, 0.0};
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x) {
    acc[0] = fma(a0[i], b0[i], acc[0]);
    if (nd > 1) acc[1] = fma(a1[i], b1[i], acc[1]);
    if (nd > 2) acc[2] = fma(a2[i], b2[i], acc[2]);
  }
#pragma unroll
  for (int d = 0; d < 3; ++d) s_red[d][threadIdx.x] = acc[d];
  __syncthreads();
  for (int h = 128; h >= 1; h >>= 1) {
    if ((int)threadIdx.x < h)
#pragma unroll
      for (int d = 0; d < 3; ++d) s_red[d][threadIdx.x] += s_red[d][threadIdx.x + h];
    __syncthreads();
  }
  if (threadIdx.x < 3) partial[threadIdx.x * kDotBlocks + blockIdx.x] = s_red[threadIdx.x][0];
}

__global__ void __launch_bounds__(256) dot_finish_kernel(const double* partial, double* out) {
  __shared__ double s_red[3][256];
#pragma unroll
  for (int d = 0; d < 3; ++d) {
    double acc = 0.0;
    for (int b = threadIdx.x; b < kDotBlocks; b += 256) acc += partial[d * kDotBlocks + b];
    s_red[d][threadIdx.x] = acc;
  }
  __syncthreads();
  for (int h = 128; h >= 1; h >>= 1) {
    if ((int)threadIdx.x < h)
#pragma unroll
      for (int d = 0; d < 3; ++d) s_red[d][threadIdx.x] += s_red[d][threadIdx.x + h];
    __syncthreads();
  }
  if (threadIdx.x < 3) out[threadIdx.x] = s_red[threadIdx.x][0];
}

static int vec_grid(int64_t n) {
  const int64_t b = (n + 255) / 256;
  return (int)(b < 8192 ? (b > 0 ? b : 1) : 8192);
}

hipError_t launch_spmv_sym(const int64_t* colptr, const int32_t* rowval, const double* nzval,
                           const double* x, double* y, int64_t n, double scale, hipStream_t st) {
  if (n <= 0) return hipSuccess;
  const int64_t blocks = (n + 3) / 4;  // four rows (wavefronts) per workgroup
  ClockScope clk_(st, "skeleton_spmv");
  hipLaunchKernelGGL(spmv_sym_kernel, dim3((unsigned)(blocks < 65536 ? blocks : 65536)), dim3(256),
                     0, st, colptr, rowval, nzval, x, y, n, scale);
  return hipGetLastError();
}

hipError_t launch_block_inv(const int64_t* colptr, const int32_t* rowval, const double* nzval,
                            double* binv, int64_t nblk, int bs, hipStream_t st) {
  if (nblk <= 0) return hipSuccess;
  if (bs < 1 || bs > kBlkMax) return hipErrorInvalidValue;
  const int64_t blocks = (nblk + 1) / 2;  // two blocks (wavefronts) per workgroup
  hipLaunchKernelGGL(block_inv_kernel, dim3((unsigned)(blocks < 65536 ? blocks : 65536)),
                     dim3(128), 0, st, colptr, rowval, nzval, binv, nblk, bs);
  return hipGetLastError();
}

hipError_t launch_block_apply(const double* binv, const double* r, double* z, int64_t n, int bs,
                              hipStream_t st) {
  if (n <= 0) return hipSuccess;
  hipLaunchKernelGGL(block_apply_kernel, dim3(vec_grid(n)), dim3(256), 0, st, binv, r, z, n, bs);
  return hipGetLastError();
}

hipError_t launch_krylov_update(int op, int64_t n, double a, double b, double c, double d,
                                double* u0, double* u1, const double* w0, const double* w1,
                                const double* w2, hipStream_t st) {
  if (n <= 0) return hipSuccess;
  hipLaunchKernelGGL(krylov_update_kernel, dim3(vec_grid(n)), dim3(256), 0, st, op, n, a, b, c,
                     d, u0, u1, w0, w1, w2, (const double*)nullptr, -1, -1, -1, -1);
  return hipGetLastError();
}

hipError_t launch_krylov_update_dev(int op, int64_t n, const double* sc, int ia, int ib, int ic,
                                    int id, double* u0, double* u1, const double* w0,
                                    const double* w1, const double* w2, hipStream_t st) {
  if (n <= 0) return hipSuccess;
  hipLaunchKernelGGL(krylov_update_kernel, dim3(vec_grid(n)), dim3(256), 0, st, op, n, 0.0, 0.0,
                     0.0, 0.0, u0, u1, w0, w1, w2, sc, ia, ib, ic, id);
  return hipGetLastError();
}

hipError_t launch_minres_scalar(int step, const double* dot, double* sc, hipStream_t st) {
  hipLaunchKernelGGL(minres_scalar_kernel, dim3(1), dim3(64), 0, st, step, dot, sc);
  return hipGetLastError();
}

int block_jacobi_max() { return kBlkMax; }

size_t multidot_scratch_doubles() { return 3 * (size_t)kDotBlocks + 3; }

hipError_t launch_multidot(int nd, int64_t n, const double* a0, const double* b0,
                           const double* a1, const double* b1, const double* a2,
                           const double* b2, double* scratch, hipStream_t st) {
  hipLaunchKernelGGL(multidot_kernel, dim3(kDotBlocks), dim3(256), 0, st, nd, n, a0, b0,
                     a1 ? a1 : a0, b1 ? b1 : b0, a2 ? a2 : a0, b2 ? b2 : b0, scratch);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL(dot_finish_kernel, dim3(1), dim3(256), 0, st, scratch,
                     scratch + 3 * kDotBlocks);
  return hipGetLastError();
}

hipError_t launch_copy(const double* src, double* dst, int64_t n, hipStream_t st) {
  if (n <= 0) return hipSuccess;
  ClockScope clk_(st, "recover_copy");
  hipLaunchKernelGGL(copy_kernel, dim3(1024), dim3(256), 0, st, src, dst, n);
  return hipGetLastError();
}

hipError_t launch_condense_dense(const DenseIn& in, int64_t ncells, const CondenseOut& out,
                                 hipStream_t st, Workspace& ws) {
  DenseGen gen{in};
  return launch_condense<DenseGen, false>(gen, out, ncells, in.ni, in.nb, st, ws);
}

// ---------------------------------------------------------------------------
// Fixed device scratch of the allocation-free entry point.  Every call on a
// device shares that device's copy; it keeps the last call's contents.
// ---------------------------------------------------------------------------
static __device__ __align__(256) unsigned char scratch[kScratchBytes];
namespace {
constexpr int kMaxScratchDevices = 64;
std::atomic<void*> g_scratch[kMaxScratchDevices];
}  // namespace

hipError_t scratch_init(int device) {
  if (device < 0 || device >= kMaxScratchDevices) return hipErrorInvalidDevice;
  if (g_scratch[device].load()) return hipSuccess;
  void* p = nullptr;
  // the array's name is a host-side stand-in: its device address is resolved
  // here, with the device current, and used everywhere else
  const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
  if (e != hipSuccess) return e;
  g_scratch[device].store(p);
  return hipSuccess;
}

void* scratch_address(int device) {
  if (device < 0 || device >= kMaxScratchDevices) return nullptr;
  return g_scratch[device].load();
}

int noalloc_family(int ni, int nb) {
  const int n = ni + nb;
  if (n > kNoAllocMaxN) return -1;
  const int fam = default_family(ni, nb);
  if (fam != kFamLarge && fam != kFamPanel) return fam;  // no workspace
  if (family_ok(kFamGrid16, ni, nb)) return kFamGrid16;
  if (family_ok(kFamGrid32, ni, nb)) return kFamGrid32;
  return -1;
}

hipError_t launch_condense_dense_family(const DenseIn& in, int64_t ncells, const CondenseOut& out,
                                        hipStream_t st, int family) {
  if (family == kFamLarge || family == kFamPanel) return hipErrorInvalidValue;
  DenseGen gen{in};
  Workspace none;  // the families above are the only ones that use it
  return launch_condense<DenseGen, false>(gen, out, ncells, in.ni, in.nb, st, none, family);
}

hipError_t launch_backward_dense(const DenseIn& in, const double* x, int64_t ncells, double* xi,
                                 int32_t* info, hipStream_t st, Workspace& ws) {
  BackwardGen gen{in, x};
  CondenseOut out{};
  out.X = xi;
  out.info = info;
  // interior-only system: nb = 0, the single rhs column is b_i - A_ib x
  return launch_condense<BackwardGen, true>(gen, out, ncells, in.ni, 0, st, ws);
}

hipError_t launch_recover(const double* U, const double* Y, const int32_t* cell_dofs,
                          const double* lam_free, const double* lam_dir, double* out,
                          int64_t ncells, int ni, int nb, int na, int64_t nfree, hipStream_t st,
                          int64_t u_stride, int64_t y_stride) {
  if (nb > 256 || ni > 256) return hipErrorInvalidValue;
  const int64_t usz = u_stride ? u_stride : (int64_t)ni * (ni + 1) / 2;
  const int64_t ysz = y_stride ? y_stride : (int64_t)(nb + 1) * ni;
  if (ncells > 0) {
    const int64_t blocks = (ncells + 3) / 4;
    const int grid = (int)(blocks < 65536 ? blocks : 65536);
    if (ni <= 64)
      {
      ClockScope clk_(st, "recover");
      hipLaunchKernelGGL(recover_kernel<1>, dim3(grid), dim3(256), 0, st, U, Y, cell_dofs,
                         lam_free, lam_dir, out, ncells, ni, nb, na, usz, ysz);
    }
    else if (ni <= 128)
      {
      ClockScope clk_(st, "recover");
      hipLaunchKernelGGL(recover_kernel<2>, dim3(grid), dim3(256), 0, st, U, Y, cell_dofs,
                         lam_free, lam_dir, out, ncells, ni, nb, na, usz, ysz);
    }
    else
      {
      ClockScope clk_(st, "recover");
      hipLaunchKernelGGL(recover_kernel<4>, dim3(grid), dim3(256), 0, st, U, Y, cell_dofs,
                         lam_free, lam_dir, out, ncells, ni, nb, na, usz, ysz);
    }
  }
  if (nfree > 0)
    {
      ClockScope clk_(st, "recover_copy");
      hipLaunchKernelGGL(copy_kernel, dim3(1024), dim3(256), 0, st, lam_free, out + ncells * ni,
                       nfree);
    }
  return hipGetLastError();
}

// One wavefront per cell; lane l works on interior rows l + 64 t.
template <int PHYS, int KMAX>
__global__ void __launch_bounds__(256)
    recover_mixed_kernel(FusedGen<PHYS> gen, const double* __restrict__ W,
                         const float* __restrict__ Xinv, const int32_t* __restrict__ cell_dofs,
                         const double* __restrict__ lam_free, const double* __restrict__ lam_dir,
                         double* __restrict__ out) {
  constexpr int WPB = 4;
  __shared__ CellCoefs s_sh[WPB];
  __shared__ double s_v[WPB][kMixedMaxNi + kRecoverMaxNb];  // [x0; lambda]
  __shared__ double s_r[WPB][kMixedMaxNi];
  const Tables& T = gen.T;
  const int ni = T.ni, nb = T.nb, na = T.na, nqc = T.nqc, ncol = nb + 1;
  const int64_t ncells = gen.M.ncells;
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  for (int64_t cell = (int64_t)blockIdx.x * WPB + w; cell < ncells;
       cell += (int64_t)gridDim.x * WPB) {
    for (int t = lane; t < kMaxAlpha; t += 64) s_sh[w].alpha[t] = gen.M.alpha[cell * kMaxAlpha + t];
    for (int t = lane; t < ni; t += 64)
      s_sh[w].b[t] = gen.M.srcq[cell * srcq_stride(T) + t];
    for (int j = lane; j < nb; j += 64) {
      const int32_t d = cell_dofs[cell * nb + j];
      s_v[w][ni + j] = d >= 0 ? lam_free[d] : lam_dir[-(int64_t)d - 1];
    }
    wave_sync_lds();
    const double* Wc = W + cell * (int64_t)ncol * ni;
    double x0[KMAX];
#pragma unroll
    for (int t = 0; t < KMAX; ++t) {
      const int r = lane + 64 * t;
      double x = 0.0;
      if (r < ni) {
        x = Wc[(int64_t)nb * ni + r];
        for (int j = 0; j < nb; ++j) x = fma(-Wc[(int64_t)j * ni + r], s_v[w][ni + j], x);
        s_v[w][r] = x;
      }
      x0[t] = x;
    }
    wave_sync_lds();
    // FP64 refinement sweeps x += X (c - A_ii x) with the FP32 inverse X,
    // with the condensation's own stopping rule: the correction is below
    // 2^-44 of x, or it has stalled at the FP64 floor of the residual (below
    // 1e-10 of x and no longer halving; each sweep contracts the error by
    // about cond(A_ii) times the FP32 rounding), or kRecoverSweeps
    const float* Xc = Xinv + cell * (int64_t)ni * ni;
    double cprev = 0.0;
    for (int sweep = 0; sweep < kRecoverSweeps; ++sweep) {
      // residual of the interior equations at [x; lambda], term by term (FP64)
#pragma unroll
      for (int t = 0; t < KMAX; ++t) {
        const int r = lane + 64 * t;
        if (r >= ni) continue;
        double res = s_sh[w].b[r];
        const int k1 = T.prog_rptr[r + 1];
        for (int k = T.prog_rptr[r]; k < k1; ++k)
          res = fma(-(s_sh[w].alpha[T.prog_rid[k]] * T.prog_rval[k]), s_v[w][T.prog_rcol[k]], res);
        s_r[w][r] = res;
      }
      wave_sync_lds();
      double cmax = 0.0, xmax = 0.0;
#pragma unroll
      for (int t = 0; t < KMAX; ++t) {
        const int r = lane + 64 * t;
        if (r >= ni) continue;
        double acc = 0.0;
        for (int k = 0; k < ni; ++k) acc = fma((double)Xc[(int64_t)r * ni + k], s_r[w][k], acc);
        x0[t] += acc;
        s_v[w][r] = x0[t];
        cmax = fmax(cmax, fabs(acc));
        xmax = fmax(xmax, fabs(x0[t]));
      }
      wave_sync_lds();
      cmax = wave_max_f64(cmax);
      xmax = wave_max_f64(xmax);
      // (uniform across the wavefront)
      if (!(cmax > 0x1p-44 * xmax)) break;
      if (sweep >= 1 && cmax <= 1e-10 * xmax && cmax > 0.5 * cprev) break;
      cprev = cmax;
    }
#pragma unroll
    for (int t = 0; t < KMAX; ++t) {
      const int r = lane + 64 * t;
      if (r >= ni) continue;
      if (r < na)
        out[cell * na + r] = x0[t];
      else
        out[ncells * na + cell * (ni - na) + (r - na)] = x0[t];
    }
    wave_sync_lds();
  }
}

hipError_t launch_recover_mixed(const Tables& T, const MeshDev& M, const double* W,
                                const float* Xinv, const int32_t* cell_dofs,
                                const double* lam_free, const double* lam_dir, double* out,
                                int64_t nfree, hipStream_t st) {
  if (T.ni > kMixedMaxNi || T.nb > kRecoverMaxNb || !W || !Xinv) return hipErrorInvalidValue;
  if (M.ncells > 0) {
    const int64_t blocks = (M.ncells + 3) / 4;
    const int grid = (int)(blocks < 65536 ? blocks : 65536);
    ClockScope clk_(st, "recover_mixed");
#define HDG_RM(P_, K_)                                                                        \
  hipLaunchKernelGGL((recover_mixed_kernel<P_, K_>), dim3(grid), dim3(256), 0, st,            \
                     FusedGen<P_>{T, M}, W, Xinv, cell_dofs, lam_free, lam_dir, out)
    if (T.phys == kDarcy) {
      if (T.ni <= 64) HDG_RM(kDarcy, 1); else HDG_RM(kDarcy, 2);
    } else {
      if (T.ni <= 64) HDG_RM(kElasticity, 1); else HDG_RM(kElasticity, 2);
    }
#undef HDG_RM
  }
  if (nfree > 0) {
    ClockScope clk_(st, "recover_copy");
    hipLaunchKernelGGL(copy_kernel, dim3(1024), dim3(256), 0, st, lam_free,
                       out + M.ncells * T.ni, nfree);
  }
  return hipGetLastError();
}

// --------------------------------------------------------------------------
// Flux-block Schur complement (family "flux_schur").  When the leading flux
// block of A_ii is A_qq = C (x) M (per-cell C, fixed scalar mass matrix M;
// see detect_flux_block in hdg_fe.cpp), the flux unknowns are eliminated by
// two dense products instead of na pivot steps:
//   Z = A_qq^-1 [A_qu A_qb b_q] = (I (x) M^-1) (C^-1 (x) I) G1   (na x nc)
//   R = [A_uu A_ub b_u; A_bu A_bb b_b] - [A_uq; A_bq] Z            (nr x nc)
// with nr = ni - na + nb and nc = nr + 1.  R is the augmented system of the
// remaining interior unknowns and the boundary; the regular elimination
// kernels condense it (ReducedGen) into S, g and the reduced factors that the
// recovery back-substitutes.  This is exact block Gaussian elimination; A_qq
// is C (x) M with M symmetric positive definite, so the flux block needs no
// pivoting.
//
// One workgroup (NW wavefronts) per cell.  Both products run on
// v_mfma_f64_16x16x4f64:
//  * G1 is generated into LDS ([k][col], odd leading dimension), mixed over
//    components by C^-1 in place (for a diagonal C, C^-1 scales the rows of
//    Z instead), and read as B operands.  The A operands of (I (x) M^-1) come
//    from M^-1 in LDS; blocks between different components are skipped
//    (zmask).
//  * Wavefront w owns column tiles w, w + NW, ... of Z and R.  Its Z tiles
//    stay in registers: the accumulator layout of the f64 MFMA (row
//    4v + lane/16 of the tile in register v) is the B-operand layout of
//    k-steps 4t..4t+3 of the second product.
//  * Rows na + 16 rr .. of the cell system ([A_uq A_uu A_ub b_u; ...]) are
//    generated tile by tile into LDS ([col][17], double buffered, aliasing
//    the dead G1 region) and shared by all wavefronts: the first na columns
//    are the A operands (negated), the rest initialises the accumulators.
// --------------------------------------------------------------------------
constexpr int kFluxWaves = 6;
constexpr int kFluxMaxNp = 64;
constexpr int kFusedMaxNb = 128;  // boundary values held in LDS by the flux recovery
constexpr int kRecMaxNa = 192;    // flux rows of the flux recovery (three per lane)
constexpr int kFluxTileLd = 17;  // leading dimension of a generated row tile

struct FluxPlan {
  int nr, nc, KS;     // reduced rows, columns (incl. rhs), k-steps of the flux block
  int ldg;            // leading dimension of G1 in LDS (odd)
  int ncols;          // columns of a row tile: 4 KS + 16 ceil(nc / 16)
  int cdiag;          // C is diagonal (structurally): C^-1 scales the rows of Z
  int sym;            // C diagonal and [A_uq; A_bq] = diag(s_u, s_b) [A_qu A_qb]^T
  int minv_off;       // offset (doubles) of M^-1 in the dynamic LDS
  uint32_t zmask[8];  // bit s of zmask[t]: block (row tile t, k-step s) of I (x) M^-1
                      // is not identically zero
};

// Inverse of the ncomp x ncomp coefficient matrix C[a][b] = alpha[cid[6a+b]]
// (Gauss-Jordan with partial pivoting by one thread, in the LDS work array
// w[6][12]).  Returns 1 when singular.
__device__ inline int flux_cinv(int ncomp, const int8_t* cid, const double* alpha, double* w,
                                double* cinv) {
  for (int r = 0; r < 6; ++r)
    for (int c = 0; c < 12; ++c) w[12 * r + c] = 0.0;
  for (int r = 0; r < ncomp; ++r) {
    for (int c = 0; c < ncomp; ++c) {
      const int id = cid[6 * r + c];
      w[12 * r + c] = id >= 0 ? alpha[id] : 0.0;
    }
    w[12 * r + 6 + r] = 1.0;
  }
  for (int c = 0; c < ncomp; ++c) {
    int p = c;
    for (int r = c + 1; r < ncomp; ++r)
      if (fabs(w[12 * r + c]) > fabs(w[12 * p + c])) p = r;
    if (w[12 * p + c] == 0.0) return 1;
    if (p != c)
      for (int k = 0; k < 12; ++k) {
        const double t = w[12 * p + k];
        w[12 * p + k] = w[12 * c + k];
        w[12 * c + k] = t;
      }
    const double d = 1.0 / w[12 * c + c];
    for (int k = 0; k < 12; ++k) w[12 * c + k] *= d;
    for (int r = 0; r < ncomp; ++r) {
      if (r == c) continue;
      const double f = w[12 * r + c];
      if (f == 0.0) continue;
      for (int k = 0; k < 12; ++k) w[12 * r + k] = fma(-f, w[12 * c + k], w[12 * r + k]);
    }
  }
  for (int r = 0; r < 6; ++r)
    for (int c = 0; c < 6; ++c) cinv[6 * r + c] = (r < ncomp && c < ncomp) ? w[12 * r + 6 + c] : 0.0;
  return 0;
}

template <int PHYS, int CT, int KSM, bool SYM>
__global__ void __launch_bounds__(64 * kFluxWaves, 3)  // two workgroups per CU
    flux_schur_kernel(FusedGen<PHYS> gen, int64_t first,
                      int64_t count, double* __restrict__ R, FluxPlan fp,
                      double* __restrict__ blocks, int64_t bstride) {
  constexpr int NW = kFluxWaves, NT = 64 * NW, LT = kFluxTileLd;
  constexpr int RTZ = (KSM + 3) / 4;  // row tiles of Z
  const Tables& T = gen.T;
  const int na = T.na, np = T.np, ncomp = T.fb_nc;
  const int nr = fp.nr, nc = fp.nc, KS = fp.KS, ldg = fp.ldg, ncols = fp.ncols;
  const int nu = T.ni - na;
  const int nk = 4 * KS;  // flux rows incl. zero padding
  const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
  const int i16 = lane & 15, q4 = lane >> 4;
  extern __shared__ double s_dyn[];  // G1 [nk][ldg], later row tiles [2][ncols][LT]; then M^-1
  __shared__ FusedShared s_gen;
  __shared__ double s_cinv[36];
  __shared__ uint16_t s_split[16 * RTZ];  // flux index -> component << 8 | scalar index
  __shared__ double s_cw[72];
  __shared__ int s_bad;
  const int tsz = ncols * LT;
  double* s_minv = s_dyn + fp.minv_off;  // [np][np]
  for (int t = tid; t < np * np; t += NT) s_minv[t] = T.fb_minv[t];
  for (int t = tid; t < 16 * RTZ; t += NT)
    s_split[t] = t < na ? (uint16_t)(((t / np) << 8) | (t % np)) : (uint16_t)0xffff;

  for (int64_t lc = blockIdx.x; lc < count; lc += gridDim.x) {
    gen.begin(first + lc, s_gen);
    __syncthreads();
    if (tid == 0) s_bad = flux_cinv(ncomp, T.fb_cid, s_gen.alpha, s_cw, s_cinv);
    double* Rc = R + lc * (int64_t)nc * nr;  // [nc][nr] column-major
    // ---- G1 = [A_qu A_qb b_q] into LDS (entries column by column), and zeros
    //      in its padding rows [na, nk) and columns [nc, ldg) ------------------
    for (int t0 = tid; t0 < na * nc; t0 += 4 * NT) {
      int rows[4], cols[4];
      bool ok[4];
      double v[4];
#pragma unroll
      for (int u = 0; u < 4; ++u) {
        const int t = t0 + u * NT, c = t / na, k = t - c * na;
        rows[u] = k;
        cols[u] = na + c;
        ok[u] = t < na * nc;
      }
      gen.template entries<4>(s_gen, rows, cols, ok, v);
#pragma unroll
      for (int u = 0; u < 4; ++u) {
        const int t = t0 + u * NT, c = t / na, k = t - c * na;
        if (t < na * nc) s_dyn[k * ldg + c] = v[u];
      }
    }
    for (int t = tid; t < (nk - na) * ldg; t += NT) s_dyn[na * ldg + t] = 0.0;
    for (int t = tid; t < na * (ldg - nc); t += NT) {
      const int k = t / (ldg - nc);
      s_dyn[k * ldg + nc + (t - k * (ldg - nc))] = 0.0;
    }
    __syncthreads();
    // interior blocks kept for the recovery: [A_uu][A_qu][A_uq], column-major
    double* Bk = blocks + lc * bstride;
    for (int t = tid; t < na * nu; t += NT) {
      const int c = t / na, k = t - c * na;
      Bk[nu * nu + t] = s_dyn[k * ldg + c];  // A_qu [u column][flux row]
    }
    __syncthreads();  // G1 read before the component mix below rewrites it
    if (s_bad) {
      // singular C: a zero system makes the reduced elimination report the cell
      for (int t = tid; t < nc * nr; t += NT) Rc[t] = 0.0;
      __syncthreads();
      continue;
    }
    // ---- G1 <- (C^-1 (x) I) G1, one (scalar index, column) pair per thread
    //      (diagonal C: applied to the rows of Z instead) ------------------
    if (!fp.cdiag)
    for (int t = tid; t < np * nc; t += NT) {
      const int i = t / nc, c = t - i * nc;
      double v[6];
#pragma unroll
      for (int b = 0; b < 6; ++b) v[b] = b < ncomp ? s_dyn[(b * np + i) * ldg + c] : 0.0;
#pragma unroll
      for (int a = 0; a < 6; ++a) {
        if (a >= ncomp) break;
        double x = 0.0;
#pragma unroll
        for (int b = 0; b < 6; ++b) x = fma(s_cinv[6 * a + b], v[b], x);
        s_dyn[(a * np + i) * ldg + c] = x;
      }
    }
    __syncthreads();

    // ---- Z = (I (x) M^-1) G1, one 16-row tile at a time --------------------
    double zf[CT][KSM];
#pragma unroll
    for (int rt = 0; rt < RTZ; ++rt) {
      dbl4 acc[CT];
#pragma unroll
      for (int ct = 0; ct < CT; ++ct) acc[ct] = dbl4{0.0, 0.0, 0.0, 0.0};
      // A operand (I (x) M^-1)[16 rt + lane%16][4 s + lane/16] from M^-1 in LDS
      const int si = s_split[16 * rt + i16];  // 0xffff past the flux rows
      const double* mrow = s_minv + (si == 0xffff ? 0 : (si & 255) * np);
#pragma unroll 2
      for (int s = 0; s < KS; ++s) {
        if (!((fp.zmask[rt] >> s) & 1u)) continue;  // wave-uniform
        const int sk = s_split[4 * s + q4];
        const double a = (si != 0xffff && (si >> 8) == (sk >> 8)) ? mrow[sk & 255] : 0.0;
        const double* brow = s_dyn + (4 * s + q4) * ldg + i16;
#pragma unroll
        for (int ct = 0; ct < CT; ++ct)
          if (16 * (w + NW * ct) < nc)  // wave-uniform: tiles past the last column stay idle
            acc[ct] = __builtin_amdgcn_mfma_f64_16x16x4f64(a, brow[16 * (w + NW * ct)], acc[ct],
                                                           0, 0, 0);
      }
#pragma unroll
      for (int v = 0; v < 4; ++v)
        if (4 * rt + v < KSM) {
          double sc = 1.0;
          if (fp.cdiag) {
            const int row = 16 * rt + q4 + 4 * v;
            sc = row < na ? s_cinv[7 * (row / np)] : 0.0;
          }
#pragma unroll
          for (int ct = 0; ct < CT; ++ct) zf[ct][4 * rt + v] = acc[ct][v] * sc;
        }
    }
    if (SYM) {
      // ---- R = R0 - [A_uq; A_bq] Z with [A_uq; A_bq] = diag(s_u, s_b) G^T,
      //      G = [A_qu A_qb] the first nr columns of G1 (left unmixed: C is
      //      diagonal).  The A operand is read from G1 and R0 is generated
      //      straight into the accumulators: each wavefront runs its column
      //      tiles on its own, without LDS tiles or barriers ---------------------
      const double su = (T.fb_sym & 2) ? -1.0 : 1.0, sb = (T.fb_sym & 4) ? -1.0 : 1.0;
      for (int t = tid; t < na * nu; t += NT) {  // A_uq [flux column][u row]
        const int j = t / nu, r = t - j * nu;
        Bk[nu * nu + na * nu + t] = su * s_dyn[j * ldg + r];
      }
      const int RTR = (nr + 15) >> 4;
#pragma unroll 1
      for (int rr = 0; rr < RTR; ++rr) {
        const int arow = 16 * rr + i16;  // this lane's row of the A operand (G^T)
        double r0[CT][4];
#pragma unroll
        for (int ct = 0; ct < CT; ++ct) {
          const int col = 16 * (w + NW * ct) + i16;
          int rows[4], cols[4];
          bool ok[4];
#pragma unroll
          for (int v = 0; v < 4; ++v) {
            const int row = 16 * rr + q4 + 4 * v;
            rows[v] = na + row;
            cols[v] = na + col;
            ok[v] = row < nr && col < nc;
          }
          gen.template entries<4>(s_gen, rows, cols, ok, r0[ct]);
        }
#pragma unroll
        for (int ct = 0; ct < CT; ++ct) {
          const int c0 = 16 * (w + NW * ct);
          if (c0 >= nc) continue;  // wave-uniform
          const int col = c0 + i16;
          // row r of R = R0 - s_r (G^T Z): accumulate -s_r R0 + G^T Z, then
          // scale by -s_r (s_r = +-1, so both steps are exact)
          dbl4 acc;
#pragma unroll
          for (int v = 0; v < 4; ++v) {
            const int row = 16 * rr + q4 + 4 * v;
            acc[v] = row < nu ? -su * r0[ct][v] : -sb * r0[ct][v];
            if (row < nu && col < nu) Bk[(int64_t)col * nu + row] = r0[ct][v];  // A_uu
          }
#pragma unroll
          for (int s = 0; s < KSM; ++s) {
            if (s < KS)
              acc = __builtin_amdgcn_mfma_f64_16x16x4f64(s_dyn[(4 * s + q4) * ldg + arow],
                                                         zf[ct][s], acc, 0, 0, 0);
          }
#pragma unroll
          for (int v = 0; v < 4; ++v) {
            const int row = 16 * rr + q4 + 4 * v;
            if (row < nr && col < nc)
              Rc[(int64_t)col * nr + row] = row < nu ? -su * acc[v] : -sb * acc[v];
          }
        }
      }
      __syncthreads();  // LDS reuse by the next cell
      continue;
    }
    __syncthreads();  // G1 fully read: its region now holds the row tiles

    // ---- R = R0 - [A_uq; A_bq] Z, one 16-row tile of R at a time ----------
    const int RTR = (nr + 15) >> 4;
    // R tiles are stored one iteration late, after the next tile's entry
    // loads: the vector-memory counter orders loads behind earlier stores
    dbl4 pend[CT];
    int pend_rr = -1;
    auto store_pending = [&]() {
#pragma unroll
      for (int ct = 0; ct < CT; ++ct) {
        const int col = 16 * (w + NW * ct) + i16;
#pragma unroll
        for (int v = 0; v < 4; ++v) {
          const int row = 16 * pend_rr + q4 + 4 * v;
          if (row < nr && col < nc) Rc[(int64_t)col * nr + row] = pend[ct][v];
        }
      }
    };
#pragma unroll 1
    for (int rr = 0; rr < RTR; ++rr) {
      double* g = s_dyn + (rr & 1) * tsz;  // [ncols][LT]: cols < nk: -A_(u,b)q, then R0
      for (int t0 = tid; t0 < 16 * ncols; t0 += 4 * NT) {
        int rows[4], cols[4];
        bool ok[4];
        double v[4];
#pragma unroll
        for (int u = 0; u < 4; ++u) {
          const int t = t0 + u * NT, col = t >> 4, row = 16 * rr + (t & 15);
          rows[u] = na + row;
          cols[u] = col < nk ? col : na + col - nk;
          ok[u] = t < 16 * ncols && row < nr && (col < na || (col >= nk && col - nk < nc));
        }
        gen.template entries<4>(s_gen, rows, cols, ok, v);
#pragma unroll
        for (int u = 0; u < 4; ++u) {
          const int t = t0 + u * NT, col = t >> 4;
          if (t < 16 * ncols) g[col * LT + (t & 15)] = col < nk ? -v[u] : v[u];
        }
      }
      __syncthreads();
      if (pend_rr >= 0) store_pending();
      if (16 * rr < nu)  // rows of the remaining interior unknowns: A_uq and A_uu
        for (int t = tid; t < 16 * (na + nu); t += NT) {
          const int col = t >> 4, i = t & 15, row = 16 * rr + i;
          if (row >= nu) continue;
          if (col < na)
            Bk[nu * nu + na * nu + (int64_t)col * nu + row] = -g[col * LT + i];
          else
            Bk[(int64_t)(col - na) * nu + row] = g[(nk + col - na) * LT + i];
        }
#pragma unroll
      for (int ct = 0; ct < CT; ++ct) {
        const int c0 = 16 * (w + NW * ct);
        pend[ct] = dbl4{0.0, 0.0, 0.0, 0.0};
        if (c0 >= nc) continue;  // wave-uniform
        const int col = c0 + i16;
        dbl4 acc;
#pragma unroll
        for (int v = 0; v < 4; ++v) acc[v] = g[(nk + col) * LT + q4 + 4 * v];
#pragma unroll
        for (int s = 0; s < KSM; ++s)
          if (s < KS)
            acc = __builtin_amdgcn_mfma_f64_16x16x4f64(g[(4 * s + q4) * LT + i16], zf[ct][s], acc,
                                                       0, 0, 0);
        pend[ct] = acc;
      }
      pend_rr = rr;
    }
    if (pend_rr >= 0) store_pending();
    __syncthreads();  // LDS reuse by the next cell
  }
}

// --------------------------------------------------------------------------
// Darcy flux elimination from the reference decomposition (family 10).  With
// C = alpha_0 I and the flux coupling [A_qu A_qb] = sum_e c_ea G_e per flux
// component a (build_flux_ref), the reduced system of flux_schur is
//   R[:, :nr] = R0 - s_r (1/alpha_0) sum_p w_p P_p,  w_(e,e') = sum_a c_ea c_e'a
//   R[:, nr]  = r0 - s_r (1/alpha_0) sum_e G_e^T (sum_a c_ea M^-1 b_q,a)
// (s_r = -1 on the potential rows, +1 on the trace rows: A_uq = -A_qu^T,
// A_bq = A_qb^T).  No per-cell product of the flux block remains.  A
// workgroup owns a band of BR rows of R and runs over the cells of the chunk;
// each thread owns up to kRefEntries entries (row, column) of the band and
// keeps, for all cells, their values of every P_p and their entry-program
// terms (R0 has at most kRefTerms terms per entry) in registers, so a cell
// costs the cell's coefficients, a few small products and one store per
// entry.  Of the blocks kept for the recovery only A_uu is written (the
// recovery applies A_qu and A_uq from the reference matrices as well).
// --------------------------------------------------------------------------
constexpr int kRefThreads = 256;
// entries of R per thread: 3 keeps the kernel at 144 VGPRs, three wavefronts
// per SIMD (5: 203 VGPRs, two).  Measured at 64^3 k=2: 9.4 ms per step
// against 11.4 with 5; at 24^3 k=3: 3.5 against 4.0
constexpr int kRefEntries = 3;
constexpr int kRefTerms = 3;    // entry-program terms per entry of R0 kept in registers
constexpr int kRefSplit = 16;   // threads per row of the rhs column's sum

struct RefPlan {
  int nr, np, na, nu, BR, nbands, npair;
  size_t lds;  // dynamic LDS bytes
};

static bool ref_plan_sizes(int D, int np, int na, int ni, int nb, int npair, RefPlan* rp) {
  if (ni > kRefThreads) return false;  // one right-hand-side value per thread
  const int nu = ni - na, nr = nu + nb;
  if (nr + 1 >= (1 << 23)) return false;  // entry columns packed above 8 bits of band row
  int BR = std::min(nr, kRefEntries * kRefThreads / (nr + 1));
  BR = std::min(BR, kRefThreads / kRefSplit);  // the rhs column: kRefSplit threads per row
  if (BR < 1) return false;
  const int nbands = (nr + BR - 1) / BR;
  BR = (nr + nbands - 1) / nbands;  // balance the bands
  // alpha, b_i, then M^-1 (prologue only) aliased with the rhs column, and
  // the band's columns of every H_e = M^-1 G_e (rows of odd length D np | 1)
  const size_t lds = (size_t)kMaxAlpha + ni + std::max<size_t>((size_t)np * np, BR) +
                     (size_t)((D * np) | 1) * BR;
  if (lds * sizeof(double) > 64 * 1024) return false;
  if (rp) {
    rp->nr = nr;
    rp->np = np;
    rp->na = na;
    rp->nu = nu;
    rp->BR = BR;
    rp->nbands = nbands;
    rp->npair = npair;
    rp->lds = lds * sizeof(double);
  }
  return true;
}

static bool ref_plan(const Tables& T, RefPlan* rp) {
  if (T.phys != kDarcy || T.fr_npair <= 0 || !T.fr_P || !T.fr_G || !T.fb_minv) return false;
  return ref_plan_sizes(T.D, T.np, T.na, T.ni, T.nb, T.fr_npair, rp);
}

template <int D>
__global__ void __launch_bounds__(kRefThreads, 2)
    flux_ref_kernel(FusedGen<kDarcy> gen, int64_t first, int64_t count, double* __restrict__ R,
                    RefPlan rp, double* __restrict__ blocks, int64_t bstride) {
  constexpr int NT = kRefThreads, NE = kRefEntries, NK = kRefTerms, NPAIR = D * (D + 1) / 2;
  constexpr int SPL = kRefSplit;
  const Tables& T = gen.T;
  const int nr = rp.nr, nc = nr + 1, np = rp.np, na = rp.na, nu = rp.nu, BR = rp.BR;
  const int ni = T.ni, n = T.n, nk = D * np;
  const int hld = nk | 1;  // odd row stride of s_H: the rows of a wavefront hit distinct banks
  const int tid = threadIdx.x;
  const int band = (int)(blockIdx.x % rp.nbands);
  const int64_t group = blockIdx.x / rp.nbands, ngroups = gridDim.x / rp.nbands;
  const int r0 = band * BR, rows = min(BR, nr - r0);
  extern __shared__ double s_dyn[];
  double* s_al = s_dyn;                      // [kMaxAlpha]
  double* s_b = s_al + kMaxAlpha;            // [ni] interior right-hand side
  double* s_minv = s_b + ni;                 // [np][np], prologue only
  double* s_xr = s_minv;                     // [BR] rhs column of the Schur term
  double* s_H = s_minv + max(np * np, BR);   // [BR][hld]: H_e[j][r0 + rl] at rl * hld + e * np + j
  // ---- per-thread entries: P values and R0's program terms, for all cells ----
  double pv[NE][NPAIR], tv[NE][NK];
  int tid8[NE];  // term coefficient ids, 8 bits each; 0xff: no term
  int ent[NE];   // (column << 8) | band row of the entry, -1: none
#pragma unroll
  for (int u = 0; u < NE; ++u) {
    const int t = tid + NT * u;
    const bool ok = t < rows * nc;
    const int c = ok ? t / rows : 0, rl = ok ? t - c * rows : 0;
    const int r = r0 + rl;
    ent[u] = ok ? (c << 8) | rl : -1;
#pragma unroll
    for (int p = 0; p < NPAIR; ++p)
      pv[u][p] = (ok && c < nr) ? T.fr_P[((size_t)p * nr + r) * nr + c] : 0.0;
    int ids = 0xffffff;
#pragma unroll
    for (int k = 0; k < NK; ++k) tv[u][k] = 0.0;
    if (ok && c < nr) {
      const int e = (na + c) * n + na + r;
      const uint32_t h = T.prog_head[e];
      const int cnt = (h >> 8) & 15, x0 = (int)(h >> 12);
      if (cnt >= 1) {
        ids = (ids & ~0xff) | (int)(h & 255);
        tv[u][0] = T.prog_val0[e];
      }
#pragma unroll
      for (int k = 1; k < NK; ++k)
        if (cnt > k) {
          ids = (ids & ~(0xff << (8 * k))) | ((int)T.prog_xid[x0 + k - 1] << (8 * k));
          tv[u][k] = T.prog_xval[x0 + k - 1];
        }
    }
    tid8[u] = ids;
  }
  // H_e = M^-1 G_e on the band's columns (M^-1 is symmetric, so the rhs
  // column's G_e^T M^-1 z_e is H_e^T z_e)
  // M^-1 staged transposed (M^-1[j][i] at i * np + j): the threads of a
  // wavefront take consecutive j, which the row-major copy put in one bank
  // when np = 64
  for (int t = tid; t < np * np; t += NT) {
    const int i = t / np, j = t - i * np;
    s_minv[t] = T.fb_minv[j * np + i];
  }
  __syncthreads();
  for (int t = tid; t < nk * BR; t += NT) {
    const int rl = t / nk, k = t - rl * nk, e = k / np, j = k - e * np;
    double h = 0.0;
    if (rl < rows)
      for (int i = 0; i < np; ++i)
        h = fma(s_minv[i * np + j], T.fr_G[((size_t)e * np + i) * nr + r0 + rl], h);
    s_H[rl * hld + k] = h;
  }
  __syncthreads();  // s_minv is reused as s_xr
  // the cell's coefficients and right-hand side are loaded one cell ahead
  // (a thread holds at most one of each: kMaxAlpha, ni <= NT)
  auto fetch = [&](int64_t lc, double& al, double& bv) {
    const int64_t cell = first + lc;
    al = (lc < count && tid < kMaxAlpha) ? gen.M.alpha[cell * kMaxAlpha + tid] : 0.0;
    bv = (lc < count && tid < ni) ? gen.M.srcq[cell * srcq_stride(T) + tid] : 0.0;
  };
  // rhs column: thread (rl, part) sums the terms k = part, part + SPL, ... of
  // sum_k H[k][rl] z[k], z_e[j] = sum_a c_ea b_q,a[j]; the SPL partial sums of
  // a row sit in consecutive lanes of one wavefront
  const int xrl = tid / SPL, xpart = tid - xrl * SPL;
  int xe0 = 0, xj0 = xpart;  // the thread's first term k = xe0 * np + xj0
  while (xj0 >= np) xj0 -= np, ++xe0;
  double nal, nbv;
  fetch(group, nal, nbv);
  for (int64_t lc = group; lc < count; lc += ngroups) {
    if (tid < kMaxAlpha) s_al[tid] = nal;
    if (tid < ni) s_b[tid] = nbv;
    fetch(lc + ngroups, nal, nbv);
    __syncthreads();
    // C = alpha_0 I, c_ea = -alpha[1 + e * D + a]; w_p = sum_a c_ea c_e'a / alpha_0
    double c[D * D];
#pragma unroll
    for (int t = 0; t < D * D; ++t) c[t] = -s_al[1 + t];
    const double a0 = s_al[0];
    const bool bad = !(a0 != 0.0);  // a singular C leaves a zero system
    const double inva = bad ? 0.0 : 1.0 / a0;
    if (xrl < rows) {
      // (c_ea read from LDS: a register array indexed by the runtime e would
      // become a chain of selects; -z is exact, so z = -sum_a alpha b)
      double x = 0.0;
      int e = xe0, j = xj0;  // k = e * np + j
      for (int k = xpart; k < nk; k += SPL) {
        double zn = 0.0;
#pragma unroll
        for (int a = 0; a < D; ++a) zn = fma(s_al[1 + e * D + a], s_b[a * np + j], zn);
        x = fma(s_H[xrl * hld + k], -zn, x);
        j += SPL;
        while (j >= np) j -= np, ++e;
      }
#pragma unroll
      for (int off = SPL / 2; off >= 1; off >>= 1) x += __shfl_xor(x, off, SPL);
      if (xpart == 0) s_xr[xrl] = x * inva;
    }  // (a row group's SPL lanes take the same branch)
    double w[NPAIR];
    {
      int p = 0;
#pragma unroll
      for (int e = 0; e < D; ++e)
#pragma unroll
        for (int e2 = e; e2 < D; ++e2, ++p) {
          double v = 0.0;
#pragma unroll
          for (int a = 0; a < D; ++a) v = fma(c[e * D + a], c[e2 * D + a], v);
          w[p] = v * inva;
        }
    }
    __syncthreads();
    double* Rc = R + lc * (int64_t)nc * nr;  // [nc][nr] column-major
    double* Bk = blocks + lc * bstride;      // [A_uu] (A_qu, A_uq: not used by the recovery)
#pragma unroll
    for (int u = 0; u < NE; ++u) {
      if (ent[u] < 0) continue;
      const int cc = ent[u] >> 8, rl = ent[u] & 255, r = r0 + rl;
      double r0v, x;
      if (cc < nr) {  // the entry program's terms, in its order
        r0v = 0.0;
#pragma unroll
        for (int k = 0; k < NK; ++k) {
          const int id = (tid8[u] >> (8 * k)) & 0xff;
          if (id != 0xff) r0v = fma(s_al[id], tv[u][k], r0v);
        }
        x = 0.0;
#pragma unroll
        for (int p = 0; p < NPAIR; ++p) x = fma(w[p], pv[u][p], x);
      } else {  // the trace rows carry no source (the cell system's rhs is [b_i; 0])
        r0v = r < nu ? s_b[na + r] : 0.0;
        x = s_xr[rl];
      }
      // s_r = -1 on the potential rows, +1 on the trace rows
      Rc[(int64_t)cc * nr + r] = bad ? 0.0 : (r < nu ? r0v + x : r0v - x);
      if (r < nu && cc < nu) Bk[(int64_t)cc * nu + r] = r0v;  // A_uu
    }
    __syncthreads();  // LDS reuse by the next cell
  }
}

static hipError_t launch_flux_ref(const Tables& T, const MeshDev& M, int64_t first, int64_t count,
                                  double* R, const RefPlan& rp, double* blocks, int64_t bstride,
                                  hipStream_t st) {
  if (count <= 0) return hipSuccess;
  // about 2048 workgroups over all bands (eight per CU, three resident); with
  // np = 64 (k = 3) the per-workgroup set-up (P values, H_e) is four times
  // dearer and 1024 do better.  Measured at 64^3 k=2: 9.36 ms per step with
  // 2048, 9.37 with 2304, 11.5 with 1536, 14.3 with 768; at 24^3 k=3: 3.07
  // with 1024, 3.16 with 2304, 3.30 with 1536, 3.82 with 768
  const int target = rp.np >= 64 ? 1024 : 2048;
  const int64_t groups = std::max<int64_t>(
      1, std::min<int64_t>(count, (target + rp.nbands - 1) / rp.nbands));
  const int grid = (int)(groups * rp.nbands);
  FusedGen<kDarcy> gen{T, M};
  ClockScope clk_(st, "flux_ref");
  if (T.D == 3)
    hipLaunchKernelGGL(flux_ref_kernel<3>, dim3(grid), dim3(kRefThreads), rp.lds, st, gen, first,
                       count, R, rp, blocks, bstride);
  else
    hipLaunchKernelGGL(flux_ref_kernel<2>, dim3(grid), dim3(kRefThreads), rp.lds, st, gen, first,
                       count, R, rp, blocks, bstride);
  return hipGetLastError();
}

// The reduced augmented system R of a chunk of cells as an element generator
// for the regular elimination kernels (rows and columns in reduced order, the
// last column is the right-hand side).
struct ReducedGen {
  const double* R;  // [count][nc][nr]
  int nr;
  using Shared = DenseShared;
  __device__ void begin(int64_t, Shared&) const {}
  __device__ double entry(const Shared&, int row, int col, int64_t cell) const {
    return R[(cell * (nr + 1) + col) * (int64_t)nr + row];
  }
};

// Elimination family for the reduced system: the grid kernels (no workspace;
// they also keep the multipliers and the pivot order for the recovery).
static int reduced_family(int ni, int nb) {
  const int n = ni + nb;
  if (n + 1 <= 72) return kFamGrid8;
  if (n + 1 <= 96) return kFamGrid16;
  if (n <= 192 && n + 1 <= 176) return kFamGrid32;
  return -1;
}

// Row-per-lane elimination of the reduced system [R_uu R_ub r_u; R_bu R_bb
// r_b] (NU interior + NB boundary rows, column-major in the workspace, as
// ReducedGen reads it) for the sizes of the flux family's 3D classes.  One
// workgroup of two wavefronts per cell; lane r of wavefront w holds row
// 64 w + r in registers, the unrolled steps giving every register index as a
// constant.  Two passes keep the registers to one part of the row:
//  1. the NU interior columns: all interior rows sit in wavefront 0, so the
//     pivot search (largest magnitude, smallest LAPACK position on ties) is a
//     DPP reduction and the LAPACK position bookkeeping lives in a register;
//     every row's multiplier and the pivot row of each step go to LDS;
//  2. the NB + 1 trailing columns, in chunks of CW: the same rank-1 updates
//     replayed with the recorded multipliers, the pivot lane publishing its
//     chunk in LDS each step.
// One barrier per step separates the pivot row's publication from the
// updates.  Every entry sees the same operations in the same order as in the
// grid kernels.  Outputs as condense_kernel: S, g, U, Y = L^-1 P [R_ub r_u],
// the multipliers L [step][row], the pivot order P and the singular-cell info.
template <int NU, int NB>
__global__ void __launch_bounds__(128)
    reduced_rowlu_kernel(const double* __restrict__ R, CondenseOut out, int64_t ncells) {
  constexpr int NR = NU + NB, NC = NR + 1, NT = NB + 1;
  constexpr int CW = 28;  // trailing columns per pass
  static_assert(NU <= 64 && NR <= 128, "interior rows in one wavefront, two wavefronts");
  __shared__ double s_l[NU][NR];  // multiplier of each row at each step
  __shared__ double s_prow[2][CW > NU ? CW : NU];
  __shared__ double s_rinv[2];
  __shared__ int s_piv[NU];
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  const int row = 64 * w + lane;
  const int64_t us = out.u_stride, ys = out.y_stride;
  for (int64_t cell = blockIdx.x; cell < ncells; cell += gridDim.x) {
    const double* Rc = R + cell * (int64_t)NC * NR;
    int pos = row;  // LAPACK position of this interior row
    int info = 0;
    // ---- pass 1: the interior columns, with pivoting -------------------------
    {
      double a[NU];
#pragma unroll
      for (int c = 0; c < NU; ++c) a[c] = row < NR ? Rc[(int64_t)c * NR + row] : 0.0;
#pragma unroll
      for (int j = 0; j < NU; ++j) {
        const int buf = j & 1;
        double l = 0.0;
        if (w == 0) {  // wave-uniform
          const bool cand = lane < NU && pos >= j;
          const int64_t key = cand ? dbits(fabs(a[j])) : (int64_t)-1;
          const int64_t m = wave_max_i64(key);
          const int ppos = wave_min_i32(key == m && cand ? pos : INT_MAX);
          const int p = __builtin_ffsll(__ballot(cand && pos == ppos)) - 1;
          const int q = __builtin_ffsll(__ballot(lane < NU && pos == j)) - 1;
          const double pv = wave_readlane_f64(a[j], p);
          const double rinv = pv != 0.0 ? 1.0 / pv : 0.0;
          if (pv == 0.0 && info == 0) info = j + 1;
          if (lane == q) pos = ppos;  // LAPACK interchange of positions j and ppos
          if (lane == p) {
            pos = j;
#pragma unroll
            for (int c = j + 1; c < NU; ++c) s_prow[buf][c] = a[c];
          }
          if (lane == 0) {
            s_rinv[buf] = rinv;
            s_piv[j] = p;
          }
          const bool use = lane < NU ? pos > j : row < NR;
          l = use ? a[j] * rinv : 0.0;
          if (lane < NU && use && out.L) out.L[cell * us + (int64_t)j * NU + lane] = l;
        }
        __syncthreads();
        if (w != 0) l = row < NR ? a[j] * s_rinv[buf] : 0.0;
        if (row < NR) s_l[j][row] = l;
#pragma unroll
        for (int c = j + 1; c < NU; ++c) a[c] = fma(-l, s_prow[buf][c], a[c]);
      }
      if (w == 0 && lane < NU && out.U) {  // U, packed by columns in LAPACK row order
#pragma unroll
        for (int c = 0; c < NU; ++c)
          if (c >= pos) out.U[cell * us + c * (c + 1) / 2 + pos] = a[c];
      }
    }
    __syncthreads();  // the last step's pivot row is read before pass 2 reuses it
    // ---- pass 2: the trailing columns, replaying the updates ---------------
#pragma unroll
    for (int c0 = 0; c0 < NT; c0 += CW) {
      double b[CW];
#pragma unroll
      for (int c = 0; c < CW; ++c)
        b[c] = (row < NR && c0 + c < NT) ? Rc[(int64_t)(NU + c0 + c) * NR + row] : 0.0;
#pragma unroll 1
      for (int j = 0; j < NU; ++j) {
        const int buf = j & 1;
        if (row == s_piv[j]) {
#pragma unroll
          for (int c = 0; c < CW; ++c) s_prow[buf][c] = b[c];
        }
        __syncthreads();
        const double l = row < NR ? s_l[j][row] : 0.0;
#pragma unroll
        for (int c = 0; c < CW; ++c) b[c] = fma(-l, s_prow[buf][c], b[c]);
      }
      if (w == 0 && lane < NU && out.Y) {
#pragma unroll
        for (int c = 0; c < CW; ++c)
          if (c0 + c < NT) out.Y[cell * ys + (int64_t)(c0 + c) * NU + pos] = b[c];
      }
      if (row >= NU && row < NR) {
        const int rb = row - NU;
#pragma unroll
        for (int c = 0; c < CW; ++c) {
          const int col = c0 + c;
          if (col < NB) {
            if (out.S) out.S[(cell * NB + col) * NB + rb] = b[c];
          } else if (col == NB && out.g) {
            out.g[cell * NB + rb] = b[c];
          }
        }
      }
      __syncthreads();  // the pivot-row buffers are reused by the next chunk
    }
    if (w == 0 && lane < NU && out.P) out.P[cell * us + pos] = (double)lane;
    if (threadIdx.x == 0 && out.info) out.info[cell] = info;
  }
}

// Register tiling of the flux kernel for a size: column tiles per wavefront
// and the k-step bound of the unrolled loops.
static bool flux_fits(int fb_nc, int na, int np, int ni, int nb, int* ct, int* ksm) {
  if (fb_nc <= 0 || np > kFluxMaxNp || na <= 0 || na >= ni) return false;
  const int nr = ni - na + nb, KS = (na + 3) / 4;
  const int c = ((nr + 1 + 15) / 16 + kFluxWaves - 1) / kFluxWaves;
  // k-step bounds of the instantiations (21: the 3D k=2 flux block exactly,
  // which keeps Z's registers within three wavefronts per SIMD)
  const int k = KS <= 8 ? 8 : KS <= 16 ? 16 : KS <= 21 ? 21 : KS <= 24 ? 24 : 0;
  if (k == 0 || c > 2 || c * k > 32) return false;
  if (reduced_family(ni - na, nb) < 0) return false;
  const int nu = ni - na;
  if (nu > 64 || na > 128 || nb > 128) return false;  // one wavefront per cell in the recovery
  const size_t ncp = 16 * (size_t)((nr + 1 + 15) / 16);
  const size_t g1 = (size_t)4 * KS * (ncp | 1), tiles = 2 * (4 * (size_t)KS + ncp) * kFluxTileLd;
  if ((std::max(g1, tiles) + (size_t)np * np) * sizeof(double) > 150 * 1024) return false;
  if (ct) *ct = c;
  if (ksm) *ksm = k;
  return true;
}

static bool flux_plan(const Tables& T, FluxPlan& fp, int* ct, int* ksm) {
  if (!T.fb_minv || !flux_fits(T.fb_nc, T.na, T.np, T.ni, T.nb, ct, ksm))
    return false;
  const int na = T.na, np = T.np;
  fp.nr = T.ni - na + T.nb;
  fp.nc = fp.nr + 1;
  fp.KS = (na + 3) / 4;
  const int ncp = 16 * ((fp.nc + 15) / 16);
  fp.ldg = ncp | 1;
  fp.ncols = 4 * fp.KS + ncp;
  fp.minv_off = (int)std::max((size_t)4 * fp.KS * fp.ldg, (size_t)2 * fp.ncols * kFluxTileLd);
  fp.cdiag = 1;
  for (int a = 0; a < T.fb_nc; ++a)
    for (int b = 0; b < T.fb_nc; ++b)
      if (a != b && T.fb_cid[6 * a + b] >= 0) fp.cdiag = 0;
  fp.sym = fp.cdiag && (T.fb_sym & 1);
  for (int t = 0; t < 8; ++t) {
    fp.zmask[t] = 0;
    for (int s = 0; s < fp.KS && s < 32; ++s) {
      bool nz = false;
      for (int i = 16 * t; i < std::min(16 * t + 16, na) && !nz; ++i)
        for (int k = 4 * s; k < std::min(4 * s + 4, na) && !nz; ++k) nz = i / np == k / np;
      if (nz) fp.zmask[t] |= 1u << s;
    }
  }
  return true;
}

static size_t flux_lds(const FluxPlan& fp, int np) {
  return ((size_t)fp.minv_off + (size_t)np * np) * sizeof(double);
}

// Regions of the flux family's per-cell factor slot (doubles from the slot's
// start): the reduced factors U (packed by columns), the multipliers L
// [step][row] and the pivot order P, then the kept interior blocks A_uu
// [u col][u row], A_qu [u col][flux row], A_uq [flux col][u row], then
// R_uu^-1 and U^-1, both [col][row].
struct FluxSlot {
  int64_t auu, aqu, auq, ri, uinv, size;
};

static FluxSlot flux_slot(int na, int nu) {
  FluxSlot o;
  o.auu = (int64_t)nu * (nu + 1) / 2 + (int64_t)nu * nu + nu;
  o.aqu = o.auu + (int64_t)nu * nu;
  o.auq = o.aqu + (int64_t)na * nu;
  o.ri = o.auq + (int64_t)na * nu;
  o.uinv = o.ri + (int64_t)nu * nu;
  o.size = o.uinv + (int64_t)nu * nu;
  return o;
}

// R_uu^-1 = U^-1 L^-1 P and U^-1 from the reduced factors, one wavefront per
// cell with a lane per column of the result held in registers (NUM >= nu).
// U (column-major) and the multipliers (permuted to positions) are staged in
// LDS as nu rows zero-padded to NUM, with 1 / diag(U) (0 for a zero pivot), so
// the unrolled substitutions test bounds once per column (each column's
// operands load together) and read them as broadcasts.  With NUM = 32 one pass does both inverses: lanes
// [0, nu) solve for the columns of R_uu^-1, lanes [32, 32 + nu) for those of
// U^-1, and the columns go out through LDS as one contiguous copy; with
// NUM = 64 each inverse takes a pass.  The recovery then needs only products
// with kept matrices.  A zero pivot of U gives a zero row, as in the
// substitution (the condensation reports such cells as singular).
constexpr int kInvWaves = 1;

template <int NUM>
__global__ void __launch_bounds__(64 * kInvWaves)
    flux_invert_kernel(double* __restrict__ slot, int64_t usz, int64_t ncells, int nu,
                       int64_t ri_off) {
  extern __shared__ double s_inv[];
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  const int nup = nu * (nu + 1) / 2, nn = nu * nu;
  double* sU = s_inv + (int64_t)w * (2 * nu * NUM + 2 * NUM);  // U [col][row], rows padded
  double* sL = sU + nu * NUM;  // multipliers [step][position], positions padded
  double* sP = sL + nu * NUM;  // row at each position (-1 past nu)
  double* sD = sP + NUM;        // 1 / U[c][c]
  for (int64_t cell = (int64_t)blockIdx.x * kInvWaves + w; cell < ncells;
       cell += (int64_t)gridDim.x * kInvWaves) {
    double* Uc = slot + cell * usz;
    const double* Lc = Uc + nup;
    // multipliers as stored into sU's region first, then permuted into sL
    for (int t = lane; t < nn; t += 64) sU[t] = Lc[t];
    for (int t = lane; t < NUM; t += 64) sP[t] = t < nu ? Lc[nn + t] : -1.0;
    wave_sync_lds();
    for (int t = lane; t < nu * NUM; t += 64) {
      const int i = t / NUM, k = t - i * NUM;
      sL[t] = k < nu ? sU[i * nu + (int)sP[k]] : 0.0;
    }
    wave_sync_lds();
    for (int t = lane; t < nu * NUM; t += 64) {
      const int c = t / NUM, k = t - c * NUM;
      sU[t] = k <= c ? Uc[c * (c + 1) / 2 + k] : 0.0;
    }
    for (int t = lane; t < NUM; t += 64) {
      const double d = t < nu ? Uc[t * (t + 1) / 2 + t] : 0.0;
      sD[t] = d != 0.0 ? 1.0 / d : 0.0;
    }
    wave_sync_lds();
    // y <- U^-1 y
    auto back_subst = [&](double (&y)[NUM]) {
#pragma unroll
      for (int c = NUM - 1; c >= 0; --c) {
        if (c < nu) {  // one block per column: its operands load together
          const double* col = sU + c * NUM;
          y[c] *= sD[c];
#pragma unroll
          for (int k = 0; k < c; ++k) y[k] = fma(-col[k], y[c], y[k]);
        }
      }
    };
    // y <- L^-1 y (positions)
    auto forward = [&](double (&y)[NUM]) {
#pragma unroll
      for (int i = 0; i < NUM - 1; ++i) {
        if (i + 1 < nu) {
          const double* li = sL + i * NUM;
#pragma unroll
          for (int k = i + 1; k < NUM; ++k) y[k] = fma(-li[k], y[i], y[k]);
        }
      }
    };
    double* out = Uc + ri_off;  // R_uu^-1 then U^-1, [col][row]
    if (NUM == 32) {
      const bool rinv = lane < 32;
      const int j = lane & 31;
      double y[NUM];
#pragma unroll
      for (int k = 0; k < NUM; ++k)
        y[k] = rinv ? ((int)sP[k] == j ? 1.0 : 0.0) : (k == j ? 1.0 : 0.0);
      if (rinv) forward(y);
      back_subst(y);
      wave_sync_lds();  // the staged factors are dead: their LDS takes the columns
      if (j < nu) {
        double* col = sU + (rinv ? 0 : nn) + j * nu;
#pragma unroll
        for (int k = 0; k < NUM; ++k)
          if (k < nu) col[k] = y[k];
      }
      wave_sync_lds();
      for (int t = lane; t < 2 * nn; t += 64) out[t] = sU[t];
    } else {
#pragma unroll 1
      for (int pass = 0; pass < 2; ++pass) {
        const int j = lane;
        double y[NUM];
#pragma unroll
        for (int k = 0; k < NUM; ++k)
          y[k] = pass == 0 ? ((int)sP[k] == j ? 1.0 : 0.0) : (k == j ? 1.0 : 0.0);
        if (pass == 0) forward(y);
        back_subst(y);
        if (j < nu) {
          double* col = out + (int64_t)pass * nn + (int64_t)j * nu;
#pragma unroll
          for (int k = 0; k < NUM; ++k)
            if (k < nu) col[k] = y[k];
        }
      }
    }
    wave_sync_lds();
  }
}

// flux_invert_kernel for 32 < nu <= 64 with the two inverses computed side by
// side: a workgroup of two wavefronts per cell shares the staged factors,
// wavefront 0 forming R_uu^-1 and wavefront 1 U^-1.  The factors are staged
// packed (U by columns as the slot keeps it, the multipliers by steps over
// the positions after the step), 34 KB per cell, so four cells fit a CU.
// Each lane does exactly what the single-wavefront kernel does for its column
// and pass, with the same operands, so the results are the same bit for bit.
constexpr int kInv2Lds = 64 * 65 / 2 + 64 * 63 / 2 + 2 * 64;  // doubles per cell

__global__ void __launch_bounds__(128)
    flux_invert2_kernel(double* __restrict__ slot, int64_t usz, int64_t ncells, int nu,
                        int64_t ri_off) {
  constexpr int NUM = 64;
  extern __shared__ double s_inv[];
  const int lane = threadIdx.x & 63, pass = threadIdx.x >> 6;
  const int nup = nu * (nu + 1) / 2, nn = nu * nu;
  double* sU = s_inv;                   // U packed by columns: (k, c), k <= c, at c(c+1)/2 + k
  double* sL = sU + NUM * (NUM + 1) / 2;  // multipliers: (i, k), k > i, at i(2 NUM - i - 1)/2 + k - i - 1
  double* sP = sL + NUM * (NUM - 1) / 2;  // row at each position (-1 past nu)
  double* sD = sP + NUM;                  // 1 / U[c][c]
  for (int64_t cell = blockIdx.x; cell < ncells; cell += gridDim.x) {
    double* Uc = slot + cell * usz;
    const double* Lc = Uc + nup;
    for (int t = threadIdx.x; t < NUM; t += 128) sP[t] = t < nu ? Lc[nn + t] : -1.0;
    for (int t = threadIdx.x; t < nup; t += 128) sU[t] = Uc[t];
    for (int t = threadIdx.x; t < NUM; t += 128) {
      const double d = t < nu ? Uc[t * (t + 1) / 2 + t] : 0.0;
      sD[t] = d != 0.0 ? 1.0 / d : 0.0;
    }
    __syncthreads();
    for (int t = threadIdx.x; t < NUM * (NUM - 1) / 2; t += 128) {
      // t -> (i, k): row i of the packed strictly-upper layout
      int i = 0, rem = t;
      while (rem >= NUM - 1 - i) rem -= NUM - 1 - i, ++i;
      const int k = i + 1 + rem;
      sL[t] = (i < nu && k < nu) ? Lc[i * nu + (int)sP[k]] : 0.0;
    }
    __syncthreads();
    double* out = Uc + ri_off;  // R_uu^-1 then U^-1, [col][row]
    const int j = lane;
    double y[NUM];
#pragma unroll
    for (int k = 0; k < NUM; ++k)
      y[k] = pass == 0 ? ((int)sP[k] == j ? 1.0 : 0.0) : (k == j ? 1.0 : 0.0);
    if (pass == 0) {  // y <- L^-1 y (positions)
#pragma unroll
      for (int i = 0; i < NUM - 1; ++i) {
        if (i + 1 < nu) {
          const double* li = sL + i * (2 * NUM - i - 1) / 2 - (i + 1);
#pragma unroll
          for (int k = i + 1; k < NUM; ++k) y[k] = fma(-li[k], y[i], y[k]);
        }
      }
    }
    // y <- U^-1 y
#pragma unroll
    for (int c = NUM - 1; c >= 0; --c) {
      if (c < nu) {
        const double* col = sU + c * (c + 1) / 2;
        y[c] *= sD[c];
#pragma unroll
        for (int k = 0; k < c; ++k) y[k] = fma(-col[k], y[c], y[k]);
      }
    }
    if (j < nu) {
      double* col = out + (int64_t)pass * nn + (int64_t)j * nu;
#pragma unroll
      for (int k = 0; k < NUM; ++k)
        if (k < nu) col[k] = y[k];
    }
    __syncthreads();  // the staged factors are reused by the next cell
  }
}

static hipError_t launch_flux_invert(double* slot, int64_t usz, int64_t ncells, int na, int nu,
                                     hipStream_t st) {
  if (ncells <= 0) return hipSuccess;
  const FluxSlot fs = flux_slot(na, nu);
  const int num = nu <= 32 ? 32 : 64;
  const size_t lds = kInvWaves * (size_t)(2 * nu * num + 2 * num) * sizeof(double);
  const int64_t blocks = (ncells + kInvWaves - 1) / kInvWaves;
  const int grid = (int)(blocks < 65536 ? blocks : 65536);
  ClockScope clk_(st, "flux_invert");
  if (nu <= 32)
    hipLaunchKernelGGL((flux_invert_kernel<32>), dim3(grid), dim3(64 * kInvWaves), lds, st,
                       slot, usz, ncells, nu, fs.ri);
  else if (nu <= 64)
    hipLaunchKernelGGL(flux_invert2_kernel, dim3((int)std::min<int64_t>(ncells, 65536)),
                       dim3(128), (size_t)kInv2Lds * sizeof(double), st, slot, usz, ncells, nu,
                       fs.ri);
  else
    return hipErrorInvalidValue;
  return hipGetLastError();
}

// Elimination of a chunk of reduced systems: the row-per-lane kernel for the
// 3D flux classes, else the grid family of the size.
static hipError_t launch_reduced(const double* R, int nr, const CondenseOut& o, int64_t count,
                                 int nu, int nb, hipStream_t st, int rfam, int gridmul) {
  // 32 row-per-lane workgroups per CU (about four times the resident eight):
  // measured at 64^3 k=2, 12.3 ms per step against 13.2 with 16, 14.3 with 8
  // and 12.4 with one cell per workgroup
  const int grid = (int)std::min<int64_t>(count, 256 * (gridmul > 0 ? gridmul : 32));
#define HDG_ROWLU(NU_, NB_)                                                                  \
  if (nu == NU_ && nb == NB_) {                                                              \
    ClockScope clk_(st, "reduced_rowlu");                                                    \
    hipLaunchKernelGGL((reduced_rowlu_kernel<NU_, NB_>), dim3(grid), dim3(128), 0, st, R, o, \
                       count);                                                               \
    return hipGetLastError();                                                                \
  }
  HDG_ROWLU(27, 54)
  HDG_ROWLU(18, 54)
#undef HDG_ROWLU
  ReducedGen rg{R, nr};
  Workspace none;
  return launch_condense<ReducedGen, false>(rg, o, count, nu, nb, st, none, rfam, gridmul);
}

// Cells per flux chunk: the reduced systems of one chunk live in the workspace.
static int64_t flux_chunk(int nr) {
  const int64_t per = (int64_t)nr * (nr + 1) * (int64_t)sizeof(double);
  const int64_t c = ((int64_t)512 << 20) / per;
  return std::max<int64_t>(1024, std::min<int64_t>(16384, c & ~(int64_t)255));
}

static size_t flux_workspace_bytes(int nr, int64_t ncells) {
  const int64_t chunk = std::min<int64_t>(std::max<int64_t>(ncells, 1), flux_chunk(nr));
  return (size_t)chunk * nr * (nr + 1) * sizeof(double);
}


template <int PHYS>
static hipError_t launch_flux_phys(const Tables& T, const MeshDev& M, const CondenseOut& out,
                                   hipStream_t st, Workspace& ws, int gridmul, bool ref) {
  FluxPlan fp;
  int ct = 0, ksm = 0;
  RefPlan rp;
  if (ref) {
    if (!sys_shape(T).ref || !ref_plan(T, &rp)) return hipErrorInvalidValue;
  } else if (!flux_plan(T, fp, &ct, &ksm)) {
    return hipErrorInvalidValue;
  }
  const int nu = T.ni - T.na, nb = T.nb, nr = nu + nb;
  const int rfam = reduced_family(nu, nb);
  if (rfam < 0) return hipErrorInvalidValue;
  const int64_t ncells = M.ncells;
  if (ncells <= 0) return hipSuccess;
  if (!out.U) return hipErrorInvalidValue;
  const int64_t chunk = std::min<int64_t>(ncells, flux_chunk(nr));
  hipError_t e = workspace_reserve(ws, flux_workspace_bytes(nr, ncells));
  if (e != hipSuccess) return e;
  const size_t lds = ref ? 0 : flux_lds(fp, T.np);
  if (lds > 160 * 1024 - sizeof(FusedShared) - 1024) return hipErrorInvalidValue;
  FusedGen<PHYS> gen{T, M};
  // reduced factors and kept blocks in the per-cell slots of the class
  const int64_t usz = out.u_stride, ysz = out.y_stride;
  if (usz < (int64_t)factor_slot(sys_shape(T)) || ysz < (int64_t)(nb + 1) * nu)
    return hipErrorInvalidValue;
  const int64_t boff = (int64_t)nu * (nu + 1) / 2 + (int64_t)nu * nu + nu;  // after U, L, P
  for (int64_t first = 0; first < ncells; first += chunk) {
    const int64_t count = std::min<int64_t>(chunk, ncells - first);
    const int grid = (int)std::min<int64_t>(count, 256 * 4);
    if (ref) {
      e = launch_flux_ref(T, M, first, count, ws.p, rp, out.U + first * usz + boff, usz, st);
      if (e != hipSuccess) return e;
    } else {
      ClockScope clk_(st, "flux_schur");
#define HDG_FLUX(CT_, KS_)                                                                   \
  if (ct == CT_ && ksm == KS_) {                                                             \
    if (PHYS == kDarcy && fp.sym)                                                            \
      hipLaunchKernelGGL((flux_schur_kernel<PHYS, CT_, KS_, PHYS == kDarcy>), dim3(grid),    \
                         dim3(64 * kFluxWaves), lds, st, gen, first, count, ws.p, fp,         \
                         out.U + first * usz + boff, usz);                                    \
    else                                                                                     \
      hipLaunchKernelGGL((flux_schur_kernel<PHYS, CT_, KS_, false>), dim3(grid),             \
                         dim3(64 * kFluxWaves), lds, st, gen, first, count, ws.p, fp,         \
                         out.U + first * usz + boff, usz);                                    \
  }
      HDG_FLUX(1, 8)
      HDG_FLUX(1, 16)
      HDG_FLUX(1, 21)
      HDG_FLUX(1, 24)
      HDG_FLUX(2, 8)
      HDG_FLUX(2, 16)
#undef HDG_FLUX
      e = hipGetLastError();
      if (e != hipSuccess) return e;
    }
    CondenseOut o;
    o.S = out.S ? out.S + first * nb * nb : nullptr;
    o.g = out.g ? out.g + first * nb : nullptr;
    o.U = out.U ? out.U + first * usz : nullptr;
    o.Y = out.Y ? out.Y + first * ysz : nullptr;
    o.u_stride = usz;
    o.y_stride = ysz;
    o.L = o.U ? o.U + (int64_t)nu * (nu + 1) / 2 : nullptr;
    o.P = o.L ? o.L + (int64_t)nu * nu : nullptr;
    o.info = out.info ? out.info + first : nullptr;
    e = launch_reduced(ws.p, nr, o, count, nu, nb, st, rfam, gridmul);
    if (e != hipSuccess) return e;
  }
  return launch_flux_invert(out.U, usz, ncells, T.na, nu, st);
}

hipError_t launch_condense_flux(const Tables& T, const MeshDev& M, const CondenseOut& out,
                                hipStream_t st, Workspace& ws, int gridmul, bool ref) {
  if (ref && (T.phys != kDarcy || T.fr_npair == 0)) return hipErrorInvalidValue;
  return T.phys == kDarcy ? launch_flux_phys<kDarcy>(T, M, out, st, ws, gridmul, ref)
                          : launch_flux_phys<kElasticity>(T, M, out, st, ws, gridmul, false);
}

size_t factor_slot(const SysShape& sh) {
  const size_t dense = (size_t)sh.ni * (sh.ni + 1) / 2;
  if (sh.flux_na <= 0) return dense;
  const size_t m = (size_t)(sh.ni - sh.flux_na), na = (size_t)sh.flux_na;
  size_t s = std::max(dense, (size_t)flux_slot((int)na, (int)m).size);
  return s;
}

// Family 10 has kernels for the sizes: the band kernel's plan, the reduced
// elimination (row-per-lane or grid), the inverses (nu <= 64) and the
// recovery (one wavefront per cell, at most kRecMaxNa flux rows).
static bool ref_fits(int phys, int D, int fb_nc, int na, int np, int ni, int nb, int npair) {
  if (phys != kDarcy || npair <= 0 || fb_nc != D || np > kFluxMaxNp || na <= 0 || na >= ni)
    return false;
  const int nu = ni - na;
  if (nu > 64 || na > kRecMaxNa || nb > kFusedMaxNb) return false;
  if (reduced_family(nu, nb) < 0) return false;
  return ref_plan_sizes(D, np, na, ni, nb, npair, nullptr);
}

SysShape sys_shape(const Tables& T) {
  SysShape sh;
  sh.ni = T.ni;
  sh.nb = T.nb;
  sh.flux = T.fb_minv && flux_fits(T.fb_nc, T.na, T.np, T.ni, T.nb, nullptr, nullptr);
  sh.ref = T.fb_minv && T.fr_P && T.fr_G &&
           ref_fits(T.phys, T.D, T.fb_nc, T.na, T.np, T.ni, T.nb, T.fr_npair);
  sh.flux_na = sh.flux || sh.ref ? T.na : 0;
  return sh;
}

SysShape sys_shape(const TablesHost& t) {
  SysShape sh;
  sh.ni = t.ni;
  sh.nb = t.nb;
  sh.flux = !t.fb_minv.empty() && flux_fits(t.fb_nc, t.na, t.np, t.ni, t.nb, nullptr, nullptr);
  sh.ref = !t.fb_minv.empty() && !t.fr_P.empty() &&
           ref_fits(t.phys, t.D, t.fb_nc, t.na, t.np, t.ni, t.nb, t.fr_npair);
  sh.flux_na = sh.flux || sh.ref ? t.na : 0;
  return sh;
}

// Flux recovery, one wavefront per cell, for both flux families.  The cell's
// slot holds R_uu^-1, the interior blocks A_uu, A_qu, A_uq and U^-1, the
// inverse of the reduced system's upper factor; Y = L^-1 P [R_ub r_u] are the
// reduced system's lower-factor rows (RecSlot gives where):
//   c   = b_i - A_ib lambda                 (boundary columns of the program)
//   x_u = U^-1 (y - Y lambda)
//   x_q = A_qq^-1 (c_q - A_qu x_u)          (A_qq^-1 = C^-1 (x) M^-1)
// then up to kFluxRefineSteps FP64 refinement steps against the residual of
// the full interior system, r = c - A_ii x, with the structured solve as the
// approximate inverse:
//   z = A_qq^-1 r_q,  d_u = R_uu^-1 (r_u - A_uq z),  d_q = z - A_qq^-1 A_qu d_u.
// The reduced system R_uu is worse conditioned than A_ii when the flux block
// is small (large kappa): the first x_u can be off by cond(R_uu) eps in the
// near-null mode of A_uq A_qq^-1 A_qu, and the refinement brings it to the
// accuracy of a direct FP64 solve of A_ii.  Each step reads every kept block
// once: A_qu x_u is carried from step to step (A_qu x_u + A_qu d_u), and
// r_u - A_uq z is formed as c_u - A_uu x_u - A_uq (x_q + z), whose rounding is
// that of the residual itself.  It stops once a correction is below 2^-48 of
// the solution.
constexpr int kRecWaves = 4;      // cells (wavefronts) per workgroup of the flux recovery

// doubles of LDS per wavefront of the flux recovery (the last na: the
// reference variant's work vector)
__host__ __device__ inline int rec_wave_lds(int na, int nu, int nb) {
  return kMaxAlpha + 72 + nb + 5 * na + (na > 72 ? na : 72) + 2 * nu;
}

struct RecSlot {
  const double* slot;  // per-cell slots, stride usz
  int64_t usz;
  int64_t ri, auu, aqu, auq;  // R_uu^-1 [col][row], A_uu, A_qu, A_uq (offsets in the slot)
  int64_t uinv;               // U^-1 [col][row]
  const double* w;     // Y: w + cell * wsz + column * wld + u row
  int64_t wsz;
  int wld;
};

// REF (family 10): A_qu and A_uq are not read from the slot but applied from
// the reference matrices G_e (potential columns) and the cell's cofactors
// c_ea: A_qu[(a,i)][c] = sum_e c_ea G_e[i][c] and A_uq = -A_qu^T (the Darcy
// coupling of build_flux_ref).  G_e is staged in LDS once per workgroup
// (g_lds), or, where it does not fit (3D k=3), read from the table, which
// every cell shares (L2-resident).
template <int PHYS, int KQ, bool REF>
__global__ void __launch_bounds__(256)
    recover_flux_kernel(FusedGen<PHYS> gen, RecSlot rs, const int32_t* __restrict__ cell_dofs,
                        const double* __restrict__ lam_free, const double* __restrict__ lam_dir,
                        double* __restrict__ out, bool g_lds) {
  constexpr int WPB = kRecWaves;
  // M^-1 and M transposed ([col][row]: lane i reads row i at stride 1, free of
  // bank conflicts for any np), (REF with g_lds: G_e [D][np][nu]), then the
  // wavefronts' vectors
  extern __shared__ double s_m[];
  const Tables& T = gen.T;
  const int ni = T.ni, nb = T.nb, na = T.na, np = T.np, nu = ni - na, nqc = T.nqc;
  const int ncomp = T.fb_nc, D = T.D;
  const int64_t ncells = gen.M.ncells;
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  const int gsz = REF && g_lds ? D * np * nu : 0;
  double* s_gu = s_m + 2 * np * np;  // REF with g_lds: G_e[i][c] at (e * np + i) * nu + c
  // G_e[i][c] at gsrc[(e * np + i) * gld + c]
  const double* gsrc = g_lds ? s_gu : T.fr_G;
  const int gld = g_lds ? nu : nu + nb;
  double* alpha = s_m + 2 * np * np + gsz + (int64_t)w * rec_wave_lds(na, nu, nb);  // [kMaxAlpha]
  double* c6 = alpha + kMaxAlpha;  // C, then C^-1 [2][36]
  double* lam = c6 + 72;           // [nb]
  double* xq = lam + nb;           // x_q [na]
  double* cq = xq + na;            // c_q [na]
  double* aq = cq + na;            // A_qu x_u [na]
  double* qv = aq + na;            // flux work vector [na]
  double* zv = qv + na;            // [max(na, 72)]; flux_cinv's scratch at a cell's start
  double* xuv = zv + (na > 72 ? na : 72);  // x_u [nu]
  double* uv = xuv + nu;           // reduced work vector [nu]
  double* tv = uv + nu;            // REF: G_e products [D][np]
  for (int t = threadIdx.x; t < np * np; t += 64 * WPB) {
    const int i = t / np, j = t - i * np;
    s_m[j * np + i] = T.fb_minv[t];
    s_m[np * np + j * np + i] = T.Mpp[t];
  }
  if (REF && g_lds)
    for (int t = threadIdx.x; t < gsz; t += 64 * WPB) {
      const int e = t / (np * nu), rem = t - e * np * nu, i = rem / nu, c = rem - i * nu;
      s_gu[t] = T.fr_G[((size_t)e * np + i) * (nu + nb) + c];
    }
  __syncthreads();
  const double* Minv = s_m;
  const double* Mm = s_m + np * np;

  // v <- (Cm (x) Mt) v over the flux rows (v in LDS): Mt = M^-1 with Cm = C^-1
  // for a solve, Mt = M with Cm = C for a product (Mt stored transposed)
  auto flux_apply = [&](double* v, const double* Mt, const double* Cm) {
    double y[KQ];
#pragma unroll
    for (int t = 0; t < KQ; ++t) {
      const int r = lane + 64 * t;
      double acc = 0.0;
      if (r < na) {
        const int a = r / np, i = r - a * np;
        const double* mc = Mt + i;
        const double* vr = v + a * np;
        for (int j = 0; j < np; ++j) acc = fma(mc[j * np], vr[j], acc);
      }
      y[t] = acc;
    }
    wave_sync_lds();
#pragma unroll
    for (int t = 0; t < KQ; ++t) {
      const int r = lane + 64 * t;
      if (r < na) v[r] = y[t];
    }
    wave_sync_lds();
#pragma unroll
    for (int t = 0; t < KQ; ++t) {
      const int r = lane + 64 * t;
      double acc = 0.0;
      if (r < na) {
        const int a = r / np, i = r - a * np;
        for (int b = 0; b < ncomp; ++b) acc = fma(Cm[6 * a + b], v[b * np + i], acc);
      }
      y[t] = acc;
    }
    wave_sync_lds();
#pragma unroll
    for (int t = 0; t < KQ; ++t) {
      const int r = lane + 64 * t;
      if (r < na) v[r] = y[t];
    }
    wave_sync_lds();
  };
  // qv = A_qu v (v over the u unknowns, in LDS)
  auto aqu_apply = [&](const double* Aqu, const double* v) {
    if (REF) {
      // t_e[i] = sum_c G_e[i][c] v[c]; qv[(a, i)] = sum_e c_ea t_e[i]
      for (int t = lane; t < D * np; t += 64) {
        const double* g = gsrc + (int64_t)t * gld;
        double acc = 0.0;
#pragma unroll 4
        for (int c = 0; c < nu; ++c) acc = fma(g[c], v[c], acc);
        tv[t] = acc;
      }
      wave_sync_lds();
#pragma unroll
      for (int t = 0; t < KQ; ++t) {
        const int r = lane + 64 * t;
        if (r < na) {
          const int a = r / np, i = r - a * np;
          double acc = 0.0;
          for (int e = 0; e < D; ++e) acc = fma(-alpha[1 + e * D + a], tv[e * np + i], acc);
          qv[r] = acc;
        }
      }
      return;
    }
#pragma unroll
    for (int t = 0; t < KQ; ++t) {
      const int r = lane + 64 * t;
      if (r < na) {
        double acc = 0.0;
#pragma unroll 4
        for (int c = 0; c < nu; ++c) acc = fma(Aqu[(int64_t)c * na + r], v[c], acc);
        qv[r] = acc;
      }
    }
  };

  for (int64_t cell = (int64_t)blockIdx.x * WPB + w; cell < ncells;
       cell += (int64_t)gridDim.x * WPB) {
    for (int t = lane; t < kMaxAlpha; t += 64) alpha[t] = gen.M.alpha[cell * kMaxAlpha + t];
    for (int j = lane; j < nb; j += 64) {
      const int32_t d = cell_dofs[cell * nb + j];
      lam[j] = d >= 0 ? lam_free[d] : lam_dir[-(int64_t)d - 1];
    }
    wave_sync_lds();
    if (lane < 36) {
      const int a = lane / 6, b = lane - 6 * (lane / 6);
      const int id = T.fb_cid[lane];
      c6[lane] = (a < ncomp && b < ncomp && id >= 0) ? alpha[id] : 0.0;
    }
    if (lane == 0 && flux_cinv(ncomp, T.fb_cid, alpha, zv, c6 + 36))
      for (int t = 0; t < 36; ++t) c6[36 + t] = 0.0;
    wave_sync_lds();
    const double* base = rs.slot + cell * rs.usz;
    const double* Ri = base + rs.ri;
    const double* Auu = base + rs.auu;
    const double* Aqu = base + rs.aqu;
    const double* Auq = base + rs.auq;
    const double* Wc = rs.w + cell * rs.wsz;
    const double* bi = gen.M.srcq + cell * srcq_stride(T);
    // ---- c = b_i - A_ib lambda and y - Y lambda -----------------------------------
    double cu = 0.0;
#pragma unroll
    for (int t = 0; t < KQ + 1; ++t) {
      const int r = lane + 64 * t;
      if (r < ni) {
        double v = bi[r];
        const int k1 = T.prog_rptr[r + 1];
        for (int k = T.prog_rbsplit[r]; k < k1; ++k)
          v = fma(-(alpha[T.prog_rid[k]] * T.prog_rval[k]), lam[T.prog_rcol[k] - ni], v);
        if (r < na) {
          cq[r] = v;
        } else {
          const int u = r - na;
          double x = Wc[(int64_t)nb * rs.wld + u];
#pragma unroll 4
          for (int j = 0; j < nb; ++j) x = fma(-Wc[(int64_t)j * rs.wld + u], lam[j], x);
          uv[u] = v;
          xuv[u] = x;
        }
      }
    }
    wave_sync_lds();
    if (lane < nu) cu = uv[lane];
    {  // x_u = U^-1 (y - Y lambda)
      const double* Ui = base + rs.uinv;
      double x = 0.0;
      if (lane < nu)
#pragma unroll 4
        for (int c = 0; c < nu; ++c) x = fma(Ui[(int64_t)c * nu + lane], xuv[c], x);
      wave_sync_lds();
      if (lane < nu) xuv[lane] = x;
      wave_sync_lds();
    }
    // ---- x_q = A_qq^-1 (c_q - A_qu x_u) -------------------------------------------
    aqu_apply(Aqu, xuv);
#pragma unroll
    for (int t = 0; t < KQ; ++t) {
      const int r = lane + 64 * t;
      if (r < na) {
        aq[r] = qv[r];
        qv[r] = cq[r] - qv[r];
      }
    }
    wave_sync_lds();
    flux_apply(qv, Minv, c6 + 36);
#pragma unroll
    for (int t = 0; t < KQ; ++t) {
      const int r = lane + 64 * t;
      if (r < na) xq[r] = qv[r];
    }
    wave_sync_lds();
    // ---- FP64 refinement against the full interior system ---------------------
#pragma unroll 1
    for (int it = 0; it < kFluxRefineSteps; ++it) {
      // r_q = c_q - (C (x) M) x_q - A_qu x_u  (into zv)
#pragma unroll
      for (int t = 0; t < KQ; ++t) {
        const int r = lane + 64 * t;
        if (r < na) zv[r] = xq[r];
      }
      wave_sync_lds();
      flux_apply(zv, Mm, c6);
#pragma unroll
      for (int t = 0; t < KQ; ++t) {
        const int r = lane + 64 * t;
        if (r < na) zv[r] = (cq[r] - zv[r]) - aq[r];
      }
      double ru = 0.0;
      if (lane < nu) {
        ru = cu;
#pragma unroll 4
        for (int c = 0; c < nu; ++c) ru = fma(-Auu[(int64_t)c * nu + lane], xuv[c], ru);
      }
      wave_sync_lds();
      flux_apply(zv, Minv, c6 + 36);  // z = A_qq^-1 r_q
#pragma unroll
      for (int t = 0; t < KQ; ++t) {
        const int r = lane + 64 * t;
        if (r < na) qv[r] = xq[r] + zv[r];
      }
      wave_sync_lds();
      // t = r_u - A_uq z = c_u - A_uu x_u - A_uq (x_q + z);  d_u = R_uu^-1 t
      if (REF) {
        // A_uq = -A_qu^T: s_e[i] = sum_a c_ea qv[(a, i)]; ru += sum_(e,i) G_e[i][u] s_e[i]
        for (int t = lane; t < D * np; t += 64) {
          const int e = t / np, i = t - e * np;
          double acc = 0.0;
          for (int a = 0; a < D; ++a) acc = fma(-alpha[1 + e * D + a], qv[a * np + i], acc);
          tv[t] = acc;
        }
        wave_sync_lds();
        if (lane < nu) {
#pragma unroll 4
          for (int t = 0; t < D * np; ++t) ru = fma(gsrc[(int64_t)t * gld + lane], tv[t], ru);
          uv[lane] = ru;
        }
      } else if (lane < nu) {
#pragma unroll 4
        for (int j = 0; j < na; ++j) ru = fma(-Auq[(int64_t)j * nu + lane], qv[j], ru);
        uv[lane] = ru;
      }
      wave_sync_lds();
      double du = 0.0;
      if (lane < nu)
#pragma unroll 4
        for (int j = 0; j < nu; ++j) du = fma(Ri[(int64_t)j * nu + lane], uv[j], du);
      wave_sync_lds();
      if (lane < nu) uv[lane] = du;
      wave_sync_lds();
      // d_q = z - A_qq^-1 (A_qu d_u);  A_qu x_u advances by A_qu d_u
      aqu_apply(Aqu, uv);
#pragma unroll
      for (int t = 0; t < KQ; ++t) {
        const int r = lane + 64 * t;
        if (r < na) aq[r] += qv[r];
      }
      wave_sync_lds();
      flux_apply(qv, Minv, c6 + 36);
      double dmax = lane < nu ? fabs(du) : 0.0, xmax = 0.0;
#pragma unroll
      for (int t = 0; t < KQ; ++t) {
        const int r = lane + 64 * t;
        if (r < na) {
          const double d = zv[r] - qv[r];
          const double x = xq[r] + d;
          xq[r] = x;
          dmax = fmax(dmax, fabs(d));
          xmax = fmax(xmax, fabs(x));
        }
      }
      if (lane < nu) {
        const double x = xuv[lane] + du;
        xuv[lane] = x;
        xmax = fmax(xmax, fabs(x));
      }
      wave_sync_lds();
      dmax = __builtin_bit_cast(double, wave_max_i64(__builtin_bit_cast(int64_t, dmax)));
      xmax = __builtin_bit_cast(double, wave_max_i64(__builtin_bit_cast(int64_t, xmax)));
      if (dmax <= 0x1p-48 * xmax) break;  // wave-uniform
    }
#pragma unroll
    for (int t = 0; t < KQ; ++t) {
      const int r = lane + 64 * t;
      if (r < na) out[cell * na + r] = xq[r];
    }
    if (lane < nu) out[ncells * na + cell * nu + lane] = xuv[lane];
    wave_sync_lds();
  }
}

hipError_t launch_recover_flux(const Tables& T, const MeshDev& M, const double* U,
                               const double* Y, int64_t usz, int64_t ysz, bool ref,
                               const int32_t* cell_dofs, const double* lam_free,
                               const double* lam_dir, double* out, int64_t nfree,
                               hipStream_t st) {
  const int na = T.na, nu = T.ni - na, nb = T.nb;
  if (na > (ref ? kRecMaxNa : kMixedMaxNi) || nu > 64 || nb > kFusedMaxNb || T.np > kFluxMaxNp)
    return hipErrorInvalidValue;
  if (ref && (T.phys != kDarcy || !T.fr_G)) return hipErrorInvalidValue;
  RecSlot rs;
  rs.slot = U;
  rs.usz = usz;
  const FluxSlot fs = flux_slot(na, nu);
  if (usz < fs.size || ysz < (int64_t)(nb + 1) * nu) return hipErrorInvalidValue;
  rs.ri = fs.ri;
  rs.auu = fs.auu;
  rs.aqu = fs.aqu;
  rs.auq = fs.auq;
  rs.uinv = fs.uinv;
  rs.w = Y;
  rs.wsz = ysz;
  rs.wld = nu;
  if (M.ncells > 0) {
    const int64_t blocks = (M.ncells + kRecWaves - 1) / kRecWaves;
    const int grid = (int)(blocks < 65536 ? blocks : 65536);
    const size_t base = (2 * (size_t)T.np * T.np + kRecWaves * (size_t)rec_wave_lds(na, nu, nb)) *
                        sizeof(double);
    const size_t gbytes = (size_t)T.D * T.np * nu * sizeof(double);
    // G_e in LDS while two workgroups still fit a CU
    const bool g_lds = ref && base + gbytes <= 64 * 1024;
    const size_t lds = base + (g_lds ? gbytes : 0);
    if (lds > 160 * 1024) return hipErrorInvalidValue;
    ClockScope clk_(st, "recover_flux");
#define HDG_RF(P_, K_, R_)                                                                     \
  hipLaunchKernelGGL((recover_flux_kernel<P_, K_, R_>), dim3(grid), dim3(64 * kRecWaves), lds, \
                     st, FusedGen<P_>{T, M}, rs, cell_dofs, lam_free, lam_dir, out, g_lds)
    if (T.phys == kDarcy) {
      if (ref) {
        if (na <= 64) HDG_RF(kDarcy, 1, true);
        else if (na <= 128) HDG_RF(kDarcy, 2, true);
        else HDG_RF(kDarcy, 3, true);
      } else {
        if (na <= 64) HDG_RF(kDarcy, 1, false); else HDG_RF(kDarcy, 2, false);
      }
    } else {
      if (na <= 64) HDG_RF(kElasticity, 1, false); else HDG_RF(kElasticity, 2, false);
    }
#undef HDG_RF
  }
  if (nfree > 0) {
    ClockScope clk_(st, "recover_copy");
    hipLaunchKernelGGL(copy_kernel, dim3(1024), dim3(256), 0, st, lam_free,
                       out + M.ncells * T.ni, nfree);
  }
  return hipGetLastError();
}

hipError_t launch_symbolic(const PairTable& P, int32_t* rowval, hipStream_t st) {
  if (P.nff > 0) {
    const int grid = (int)(P.nff < 65536 ? P.nff : 65536);
    {
      ClockScope clk_(st, "symbolic");
      hipLaunchKernelGGL(symbolic_kernel, dim3(grid), dim3(256), 0, st, P, rowval);
    }
  }
  return hipGetLastError();
}

hipError_t launch_assemble(const PairTable& P, const double* S, const double* g,
                           const int32_t* cell_dofs, const double* lam_dir, double* nzval,
                           double* rhs, hipStream_t st) {
  if (P.nff > 0) {
    const int grid = (int)(P.nff < 65536 ? P.nff : 65536);
    {
      ClockScope clk_(st, "assemble");
      // one wavefront per facet: a 3D k=2 facet has 891 entries, a 2D k=2 one
      // 63; measured at 3D 64^3 / 2D 512^2 k=2: 2.63 / 0.28 ms per step, against
      // 2.73 / 0.43 ms with 128 threads and 3.20 / 0.68 ms with 256
      hipLaunchKernelGGL(assemble_kernel, dim3(grid), dim3(64), 0, st, P, S, nzval);
    }
    const int64_t tot = P.nff * P.ntf;
    const int grid2 = (int)((tot + 255) / 256 < 65536 ? (tot + 255) / 256 : 65536);
    {
      ClockScope clk_(st, "assemble_rhs");
      hipLaunchKernelGGL(rhs_kernel, dim3(grid2), dim3(256), 0, st, P, S, g, cell_dofs, lam_dir,
                       rhs);
    }
  }
  return hipGetLastError();
}

__global__ void count_nonzero_kernel(const int32_t* __restrict__ v, int64_t n,
                                     unsigned long long* __restrict__ count) {
  unsigned long long local = 0;
  for (int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; t < n;
       t += (int64_t)gridDim.x * blockDim.x)
    local += v[t] != 0;
  if (local) atomicAdd(count, local);  // integer count: order-independent
}

__global__ void count_negative_kernel(const int32_t* __restrict__ v, int64_t n,
                                      unsigned long long* __restrict__ count) {
  unsigned long long local = 0;
  for (int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; t < n;
       t += (int64_t)gridDim.x * blockDim.x)
    local += v[t] < 0;
  if (local) atomicAdd(count, local);  // integer count: order-independent
}

__global__ void zero_count_kernel(unsigned long long* count) {
  if (threadIdx.x == 0) count[threadIdx.x] = 0;
}

hipError_t launch_count_negative(const int32_t* v, int64_t n, int64_t* count_dev,
                                 hipStream_t st) {
  {
      ClockScope clk_(st, "count");
      hipLaunchKernelGGL(zero_count_kernel, dim3(1), dim3(1), 0, st,
                     reinterpret_cast<unsigned long long*>(count_dev));
    }
  if (n > 0) {
    const int grid = (int)((n + 255) / 256 < 1024 ? (n + 255) / 256 : 1024);
    {
      ClockScope clk_(st, "count");
      hipLaunchKernelGGL(count_negative_kernel, dim3(grid), dim3(256), 0, st, v, n,
                       reinterpret_cast<unsigned long long*>(count_dev));
    }
  }
  return hipGetLastError();
}

hipError_t launch_count_nonzero(const int32_t* info, int64_t n, int64_t* count_dev,
                                hipStream_t st) {
  // a fill kernel rather than a memset node, so the sequence stays capturable
  {
      ClockScope clk_(st, "count");
      hipLaunchKernelGGL(zero_count_kernel, dim3(1), dim3(1), 0, st,
                     reinterpret_cast<unsigned long long*>(count_dev));
    }
  if (n > 0) {
    const int grid = (int)((n + 255) / 256 < 1024 ? (n + 255) / 256 : 1024);
    {
      ClockScope clk_(st, "count");
      hipLaunchKernelGGL(count_nonzero_kernel, dim3(grid), dim3(256), 0, st, info, n,
                       reinterpret_cast<unsigned long long*>(count_dev));
    }
  }
  return hipGetLastError();
}

// MFMA form of the refinement, for ni <= 128 and nb + 1 <= 64 (one cell per
// 512-thread workgroup).  Both products of a sweep are 16x16 tile GEMMs:
//   R = B - A_ii W         v_mfma_f64_16x16x4f64, A_ii from the workspace,
//                          W (FP64) from LDS;
//   W += X R               v_mfma_f32_16x16x4f32, X and R (FP32) from LDS;
// and S, g = [A_bb b_b] - A_bi W with FP64 MFMA after the last sweep.  The 16
// rows x 16 columns output tiles are dealt to the 8 wavefronts; a wavefront
// updates only its own tiles of W, so a sweep needs two barriers.
typedef float flt4 __attribute__((ext_vector_type(4)));
constexpr int kRefMaxRT = 8;   // row tiles of W (ni <= 128)
constexpr int kRefMaxCT = 4;   // column tiles of W (nb + 1 <= 64)

template <int CTM>
__global__ void __launch_bounds__(512)
    mixed_refine_mfma_kernel(const double* __restrict__ A64, const float* __restrict__ X,
                             int ni, int nb, int64_t count, CondenseOut out, int maxit,
                             int32_t* __restrict__ iters) {
  // wavefront w owns row strip w (16 rows) of W and all its CT <= CTM column
  // tiles, so each k-step loads its A_ii (or X) operand once for CT MFMAs
  constexpr int NT = 512, NW = 8;
  const int n = ni + nb, ncol = nb + 1;
  const int RT = (ni + 15) >> 4, CT = (ncol + 15) >> 4;
  const int LDX = 16 * RT + 4, LDW = 16 * CT + 2, LDR = 16 * CT + 4;
  const int KS = (ni + 3) >> 2;  // k-steps of 4
  const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
  const int i16 = lane & 15, q4 = lane >> 4;
  extern __shared__ double s_dyn[];
  double* sW = s_dyn;                                        // [16 RT][LDW]
  float* sX = reinterpret_cast<float*>(sW + 16 * RT * LDW);  // [16 RT][LDX]
  float* sR = sX + 16 * RT * LDX;                            // [16 RT][LDR]
  __shared__ double s_red[2][NW];
  const bool strip = wv < RT;
  const int r0 = wv << 4, ra = r0 + i16;

  for (int64_t lc = blockIdx.x; lc < count; lc += gridDim.x) {
    const double* A = A64 + lc * (int64_t)n * (n + 1);  // A[col * n + row]
    const float* Xc = X + lc * (int64_t)ni * ni;
    // X, W = 0 and R_0 = B into LDS (zero padding outside ni x ncol)
    for (int t = tid; t < 16 * RT * 16 * RT; t += NT) {
      const int r = t / (16 * RT), c = t - r * (16 * RT);
      sX[r * LDX + c] = (r < ni && c < ni) ? Xc[r * ni + c] : 0.f;
    }
    for (int t = tid; t < 16 * RT * 16 * CT; t += NT) {
      const int r = t / (16 * CT), c = t - r * (16 * CT);
      sW[r * LDW + c] = 0.0;
      sR[r * LDR + c] = (r < ni && c < ncol) ? (float)A[(int64_t)(ni + c) * n + r] : 0.f;
    }
    __syncthreads();
    double dprev = 0.0;
    int it = 0, conv = 1;
    for (;;) {
      // ---- W += X R (FP32 products, FP64 update) ----
      double dmax = 0.0, wmax = 0.0;
      if (strip) {
        flt4 d[CTM];
#pragma unroll
        for (int ct = 0; ct < CTM; ++ct) d[ct] = flt4{0.f, 0.f, 0.f, 0.f};
#pragma unroll 2
        for (int ks = 0; ks < KS; ++ks) {
          const int k = 4 * ks + q4;
          const float a = sX[ra * LDX + k];
#pragma unroll
          for (int ct = 0; ct < CTM; ++ct)
            if (ct < CT)
              d[ct] = __builtin_amdgcn_mfma_f32_16x16x4f32(a, sR[k * LDR + 16 * ct + i16], d[ct],
                                                          0, 0, 0);
        }
#pragma unroll
        for (int ct = 0; ct < CTM; ++ct) {
          if (ct >= CT) continue;
#pragma unroll
          for (int v = 0; v < 4; ++v) {
            const int r = r0 + 4 * q4 + v, c = 16 * ct + i16;
            if (r < ni && c < ncol) {
              const double w = sW[r * LDW + c] + (double)d[ct][v];
              sW[r * LDW + c] = w;
              dmax = fmax(dmax, fabs((double)d[ct][v]));
              wmax = fmax(wmax, fabs(w));
            }
          }
        }
      }
      dmax = __builtin_bit_cast(double, wave_max_i64(__builtin_bit_cast(int64_t, dmax)));
      wmax = __builtin_bit_cast(double, wave_max_i64(__builtin_bit_cast(int64_t, wmax)));
      if (lane == 0) {
        s_red[0][wv] = dmax;
        s_red[1][wv] = wmax;
      }
      __syncthreads();  // W complete for the residual; the maxima published
      dmax = s_red[0][0];
      wmax = s_red[1][0];
#pragma unroll
      for (int q = 1; q < NW; ++q) {
        dmax = fmax(dmax, s_red[0][q]);
        wmax = fmax(wmax, s_red[1][q]);
      }
      const bool done = it >= maxit || dmax <= 0x1p-44 * wmax ||
                        (it >= 2 && dmax <= 1e-10 * wmax && dmax > 0.5 * dprev);
      if (done) {
        conv = dmax <= 1e-10 * wmax;
        break;
      }
      dprev = dmax;
      ++it;
      // ---- R = B - A_ii W (FP64), then to LDS in FP32 ----
      dbl4 acc[CTM];
      if (strip) {
#pragma unroll
        for (int ct = 0; ct < CTM; ++ct) {
#pragma unroll
          for (int v = 0; v < 4; ++v) {
            const int r = r0 + q4 + 4 * v, c = 16 * ct + i16;
            acc[ct][v] = (ct < CT && r < ni && c < ncol) ? A[(int64_t)(ni + c) * n + r] : 0.0;
          }
        }
#pragma unroll 2
        for (int ks = 0; ks < KS; ++ks) {
          const int k = 4 * ks + q4;
          const double a = (ra < ni && k < ni) ? -A[(int64_t)k * n + ra] : 0.0;
#pragma unroll
          for (int ct = 0; ct < CTM; ++ct)
            if (ct < CT)
              acc[ct] = __builtin_amdgcn_mfma_f64_16x16x4f64(a, sW[k * LDW + 16 * ct + i16],
                                                             acc[ct], 0, 0, 0);
        }
      }
      __syncthreads();  // everyone is done reading R before it is overwritten
      if (strip) {
#pragma unroll
        for (int ct = 0; ct < CTM; ++ct) {
          if (ct >= CT) continue;
#pragma unroll
          for (int v = 0; v < 4; ++v)
            sR[(r0 + q4 + 4 * v) * LDR + 16 * ct + i16] = (float)acc[ct][v];
        }
      }
      __syncthreads();
    }
    // ---- S = A_bb - A_bi W, g = b_b - A_bi w (FP64 MFMA), row strip per wavefront ----
    const int BT = (nb + 15) >> 4;
    for (int bs = wv; bs < BT; bs += NW) {
      const int b0 = bs << 4, rb_a = b0 + i16;
      dbl4 acc[CTM];
#pragma unroll
      for (int ct = 0; ct < CTM; ++ct) {
#pragma unroll
        for (int v = 0; v < 4; ++v) {
          const int r = b0 + q4 + 4 * v, c = 16 * ct + i16;
          acc[ct][v] = (ct < CT && r < nb && c < ncol) ? A[(int64_t)(ni + c) * n + ni + r] : 0.0;
        }
      }
#pragma unroll 2
      for (int ks = 0; ks < KS; ++ks) {
        const int k = 4 * ks + q4;
        const double a = (rb_a < nb && k < ni) ? -A[(int64_t)k * n + ni + rb_a] : 0.0;
#pragma unroll
        for (int ct = 0; ct < CTM; ++ct)
          if (ct < CT)
            acc[ct] = __builtin_amdgcn_mfma_f64_16x16x4f64(a, sW[k * LDW + 16 * ct + i16],
                                                           acc[ct], 0, 0, 0);
      }
#pragma unroll
      for (int ct = 0; ct < CTM; ++ct) {
        if (ct >= CT) continue;
#pragma unroll
        for (int v = 0; v < 4; ++v) {
          const int rb = b0 + q4 + 4 * v, c = 16 * ct + i16;
          if (rb >= nb || c >= ncol) continue;
          if (c < nb) {
            if (out.S) out.S[(lc * nb + c) * nb + rb] = acc[ct][v];
          } else {
            out.g[lc * nb + rb] = acc[ct][v];
          }
        }
      }
    }
    // W to its output slot, column-major [nb+1][ni]
    for (int t = tid; t < ncol * ni; t += NT) {
      const int j = t / ni, k = t - j * ni;
      out.Y[lc * (int64_t)ncol * ni + t] = sW[k * LDW + j];
    }
    if (tid == 0 && iters) iters[lc] = conv ? it : -it - 1;
    __syncthreads();
  }
}

static size_t refine_mfma_lds(int ni, int nb) {
  const int RT = cdiv(ni, 16), CT = cdiv(nb + 1, 16);
  return (size_t)16 * RT *
         ((16 * CT + 2) * sizeof(double) + (16 * RT + 4) * sizeof(float) +
          (16 * CT + 4) * sizeof(float));
}

static bool refine_mfma_ok(int ni, int nb) {
  return ni <= 16 * kRefMaxRT && nb + 1 <= 16 * kRefMaxCT && nb >= 1 &&
         refine_mfma_lds(ni, nb) <= 150 * 1024;
}

// columns of W per thread (mixed_refine_kernel's JC): the smallest of 4, 8,
// 16, 32 whose ceil((nb+1) / JC) column groups fit in the workgroup
static int mixed_jc(int ni, int nb) {
  const int groups = kRefineThreads / (64 * cdiv(ni, 64));
  for (int jc : {4, 8, 16, 32})
    if (cdiv(nb + 1, jc) <= groups) return jc;
  return 0;
}

static int mixed_ldr(int ni, int nb) {
  const int jc = mixed_jc(ni, nb);
  return cdiv(nb + 1, jc) * jc;
}

// LDS of mixed_refine_kernel: X and the residual in FP32, W in FP64 if wlds
static size_t mixed_lds(int ni, int nb, bool wlds) {
  const size_t ldr = mixed_ldr(ni, nb);
  return ((size_t)ni * (ni + 1) + (size_t)ni * ldr) * sizeof(float) +
         (wlds ? (size_t)ni * ldr * sizeof(double) : 0);
}

constexpr size_t kMixedMaxLds = 150 * 1024;

bool mixed_supported(int ni, int nb) {
  // below kMixedMinNi the FP64 kernels are bound by HBM traffic and faster
  // than the FP32 inverse + refinement, so mixed requests use them
  return ni >= kMixedMinNi && ni <= kMixedMaxNi && nb >= 1 && nb <= 512 &&
         mixed_jc(ni, nb) > 0 && mixed_lds(ni, nb, false) <= kMixedMaxLds;
}

// cells per chunk of the mixed path and the workspace they need: the FP64
// systems and the FP32 inverses per cell, plus one W per refine workgroup
static int64_t mixed_chunk(int64_t ncells) { return std::min<int64_t>(ncells, kMixedChunkCells); }
constexpr int kRefineGrid = 256 * 2;

static size_t mixed_cell_bytes(int ni, int nb, size_t* a_off, size_t* x_off) {
  const size_t n = ni + nb;
  const size_t a = n * (n + 1) * sizeof(double), x = (size_t)ni * ni * sizeof(float);
  if (a_off) *a_off = a;
  if (x_off) *x_off = x;
  return a + x;
}

size_t mixed_workspace_bytes(int ni, int nb, int64_t ncells) {
  return (size_t)mixed_chunk(ncells) * mixed_cell_bytes(ni, nb, nullptr, nullptr) + 8 +
         (size_t)kRefineGrid * ni * mixed_ldr(ni, nb) * sizeof(double);
}

template <int PHYS>
static hipError_t launch_mixed_phys(const Tables& T, const MeshDev& M, const CondenseOut& out,
                                    hipStream_t st, Workspace& ws, int maxit, int32_t* iters) {
  const int ni = T.ni, nb = T.nb, n = ni + nb;
  const int64_t chunk = mixed_chunk(M.ncells);
  hipError_t e = workspace_reserve(ws, mixed_workspace_bytes(ni, nb, M.ncells));
  if (e != hipSuccess) return e;
  size_t abytes, xbytes;
  (void)mixed_cell_bytes(ni, nb, &abytes, &xbytes);
  char* base = reinterpret_cast<char*>(ws.p);
  double* A64 = reinterpret_cast<double*>(base);
  float* Xws = reinterpret_cast<float*>(base + (size_t)chunk * abytes);
  double* Wws = reinterpret_cast<double*>(base + (size_t)chunk * abytes +
                                          (((size_t)chunk * xbytes + 7) & ~(size_t)7));
  FusedGen<PHYS> gen{T, M};
  const int ra = cdiv(ni, 16), jc = mixed_jc(ni, nb);
  const bool wlds = mixed_lds(ni, nb, true) <= kMixedMaxLds;
  const size_t lds = mixed_lds(ni, nb, wlds);
  for (int64_t first = 0; first < M.ncells; first += chunk) {
    const int64_t count = std::min<int64_t>(chunk, M.ncells - first);
    // the FP32 inverses go straight to their per-cell slot when one is given
    float* X = out.Xinv ? out.Xinv + first * (int64_t)ni * ni : Xws;
    const int g1 = (int)std::min<int64_t>(count, 256 * 16);
    {
      ClockScope clk_(st, "mixed_gen_system");
      hipLaunchKernelGGL(gen_system_kernel<PHYS>, dim3(g1), dim3(256), 0, st, gen, first, count,
                       A64);
    }
    const int g2 = (int)std::min<int64_t>(count, 256 * 8);
    int32_t* info = out.info ? out.info + first : nullptr;
    switch (ra) {
#define HDG_GJ(R_)                                                                            \
  case R_:                                                                                    \
    { ClockScope clk_(st, "mixed_gj_inverse"); hipLaunchKernelGGL(gj_inverse_kernel<R_>, dim3(g2), dim3(256), 0, st, A64, n, ni, count, X, \
                       info); }                                                                 \
    break;
      HDG_GJ(1) HDG_GJ(2) HDG_GJ(3) HDG_GJ(4) HDG_GJ(5) HDG_GJ(6) HDG_GJ(7) HDG_GJ(8)
#undef HDG_GJ
      default: return hipErrorInvalidValue;
    }
    CondenseOut o;
    o.S = out.S ? out.S + first * nb * nb : nullptr;
    o.g = out.g + first * nb;
    o.Y = out.Y + first * (int64_t)(nb + 1) * ni;
    int32_t* it = iters ? iters + first : nullptr;
    const int g3 = (int)std::min<int64_t>(count, kRefineGrid);
    if (refine_mfma_ok(ni, nb)) {
      const int ctiles = cdiv(nb + 1, 16);
      const size_t lds3 = refine_mfma_lds(ni, nb);
#define HDG_RMF(T_)                                                                           \
  if (ctiles <= T_) {                                                                         \
    e = hipFuncSetAttribute((const void*)mixed_refine_mfma_kernel<T_>,                        \
                            hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds3);           \
    if (e != hipSuccess) return e;                                                            \
    ClockScope clk_(st, "mixed_refine");                                                      \
    hipLaunchKernelGGL(mixed_refine_mfma_kernel<T_>, dim3(g3), dim3(512), lds3, st, A64, X,   \
                       ni, nb, count, o, maxit, it);                                          \
  } else
      HDG_RMF(1) HDG_RMF(2) HDG_RMF(4) return hipErrorInvalidValue;
#undef HDG_RMF
      e = hipGetLastError();
      if (e != hipSuccess) return e;
      continue;
    }
#define HDG_REF(J_, L_)                                                                       \
  if (jc == J_ && wlds == L_) {                                                               \
    e = hipFuncSetAttribute((const void*)mixed_refine_kernel<J_, L_>,                         \
                            hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);            \
    if (e != hipSuccess) return e;                                                            \
    { ClockScope clk_(st, "mixed_refine"); hipLaunchKernelGGL((mixed_refine_kernel<J_, L_>), dim3(g3), dim3(kRefineThreads), lds, st, \
                       A64, X, Wws, ni, nb, count, o, maxit, it); }                             \
  } else
    HDG_REF(4, true) HDG_REF(8, true) HDG_REF(16, true) HDG_REF(32, true)
    HDG_REF(4, false) HDG_REF(8, false) HDG_REF(16, false) HDG_REF(32, false)
    return hipErrorInvalidValue;
#undef HDG_REF
    e = hipGetLastError();
    if (e != hipSuccess) return e;
  }
  return hipSuccess;
}

hipError_t launch_condense_mixed(const Tables& T, const MeshDev& M, const CondenseOut& out,
                                 hipStream_t st, Workspace& ws, bool prep, int maxit,
                                 int32_t* iters) {
  if (M.ncells <= 0) return hipSuccess;
  if (!mixed_supported(T.ni, T.nb) || !out.g || !out.Y) return hipErrorInvalidValue;
  if (prep) {
    hipError_t e = launch_prep(T, M, st);
    if (e != hipSuccess) return e;
  }
  if (T.phys == kDarcy) return launch_mixed_phys<kDarcy>(T, M, out, st, ws, maxit, iters);
  return launch_mixed_phys<kElasticity>(T, M, out, st, ws, maxit, iters);
}

hipError_t launch_cell_systems(const Tables& T, const MeshDev& M, int64_t first, int64_t count,
                               double* A, double* b, hipStream_t st) {
  if (count <= 0) return hipSuccess;
  hipError_t e = launch_prep(T, M, st);
  if (e != hipSuccess) return e;
  const int grid = (int)(count < 4096 ? count : 4096);
  if (T.phys == kDarcy) {
    FusedGen<kDarcy> gen{T, M};
    {
      ClockScope clk_(st, "cell_systems");
      hipLaunchKernelGGL(cell_systems_kernel<kDarcy>, dim3(grid), dim3(256), 0, st, gen, first,
                       count, A, b);
    }
  } else {
    FusedGen<kElasticity> gen{T, M};
    {
      ClockScope clk_(st, "cell_systems");
      hipLaunchKernelGGL(cell_systems_kernel<kElasticity>, dim3(grid), dim3(256), 0, st, gen, first,
                       count, A, b);
    }
  }
  return hipGetLastError();
}

}  // namespace hdg
