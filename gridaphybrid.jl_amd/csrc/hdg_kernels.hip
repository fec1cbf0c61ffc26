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

#include <climits>

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

__device__ inline void cell_geometry(const MeshDev& m, int D, int64_t cell, CellGeo& g) {
  // fixed-bound loops with guards keep every local array in registers
  double J[9], Ji[9];
#pragma unroll
  for (int t = 0; t < 9; ++t) J[t] = t < D * D ? m.jac[cell * D * D + t] : 0.0;
#pragma unroll
  for (int a = 0; a < 3; ++a) g.x0[a] = a < D ? m.x0[cell * D + a] : 0.0;
#pragma unroll
  for (int t = 0; t < 9; ++t) g.J[t] = J[t];
  double det;
  if (D == 2) {
    det = J[0] * J[3] - J[2] * J[1];
    Ji[0] = J[3] / det;
    Ji[3] = J[0] / det;
    Ji[2] = -J[2] / det;
    Ji[1] = -J[1] / det;
#pragma unroll
    for (int t = 4; t < 9; ++t) Ji[t] = 0.0;
  } else {
#define JM(r, c) J[(r) + 3 * (c)]
    det = JM(0, 0) * (JM(1, 1) * JM(2, 2) - JM(2, 1) * JM(1, 2)) -
          JM(0, 1) * (JM(1, 0) * JM(2, 2) - JM(2, 0) * JM(1, 2)) +
          JM(0, 2) * (JM(1, 0) * JM(2, 1) - JM(2, 0) * JM(1, 1));
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

__device__ inline double elas_h(int D, const double* x, int i, int j, int k) {
  // d2 u_i / dx_j dx_k of u_i = sin(pi (x_i + x_{i+1}/2))
  const int i1 = (i + 1) % D;
  const double dj = (j == i ? 1.0 : 0.0) + (j == i1 ? 0.5 : 0.0);
  const double dk = (k == i ? 1.0 : 0.0) + (k == i1 ? 0.5 : 0.0);
  return -M_PI * M_PI * sin(M_PI * (x[i] + 0.5 * x[i1])) * dj * dk;
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

struct FusedShared {
  CellGeo geo;
  double src[kMaxQP][4];  // source values at cell quadrature points
};

__device__ inline void sym_pair(int D, int a, int& i, int& j) {
  // upper triangle, row-major: 2D (11,12,22), 3D (11,12,13,22,23,33);
  // packed as 2-bit (i, j) pairs to stay out of scratch memory
  const unsigned code = D == 2 ? 0x540u /* (0,0)(0,1)(1,1) */ : 0xA95840u;
  i = (code >> (4 * a)) & 3u;
  j = (code >> (4 * a + 2)) & 3u;
}

struct FusedGen {
  Tables T;
  MeshDev M;

  using Shared = FusedShared;

  __device__ void begin(int64_t cell, Shared& sh) const {
    const int D = T.D;
    if (threadIdx.x == 0) {
      cell_geometry(M, D, cell, sh.geo);
      if (T.phys == kElasticity) {
        double nu = sh.geo.coef, c1 = (1.0 + nu) / M.young, c2 = nu / M.young;
        int nsym = D * (D + 1) / 2;
        for (int a = 0; a < nsym; ++a)
          for (int b = 0; b < nsym; ++b) {
            int ia, ja, ib, jb;
            sym_pair(D, a, ia, ja);
            sym_pair(D, b, ib, jb);
            double dd = (a == b) ? (ia == ja ? 1.0 : 2.0) : 0.0;
            double tra = ia == ja ? 1.0 : 0.0, trb = ib == jb ? 1.0 : 0.0;
            sh.geo.K[a][b] = c1 * dd - c2 * tra * trb;
          }
      }
    }
    __syncthreads();
    // source values at the cell quadrature points
    const int ncomp = T.phys == kDarcy ? D + 1 : D;
    for (int q = threadIdx.x; q < T.nqc; q += blockDim.x) {
      double v0 = 0.0, v1 = 0.0, v2 = 0.0, v3 = 0.0;
      if (M.source == 100) {
        const double* p = M.src + (cell * T.nqc + q) * ncomp;
        v0 = p[0];
        v1 = p[1];
        if (ncomp > 2) v2 = p[2];
        if (ncomp > 3) v3 = p[3];
      } else if (M.source != 0) {
        double x[3];
#pragma unroll
        for (int a = 0; a < 3; ++a) {
          double t = sh.geo.x0[a];
#pragma unroll
          for (int e = 0; e < 3; ++e)
            if (a < D && e < D) t += sh.geo.J[a + D * e] * T.xq[3 * q + e];
          x[a] = t;
        }
        double f[3], gs = 0.0;
        if (T.phys == kDarcy)
          darcy_source(M.source, D, x, sh.geo.coef, f, &gs);
        else
          elas_source(M.source, D, x, M.young, sh.geo.coef, f);
        v0 = f[0];
        v1 = f[1];
        v2 = f[2];
        if (T.phys == kDarcy) {
          // potential source g sits right after the D flux components
          if (D == 2) v2 = gs; else v3 = gs;
        }
      }
      sh.src[q][0] = v0;
      sh.src[q][1] = v1;
      sh.src[q][2] = v2;
      sh.src[q][3] = v3;
    }
  }

  __device__ double entry(const Shared& sh, int row, int col) const {
    return T.phys == kDarcy ? darcy(sh, row, col) : elas(sh, row, col);
  }

  // Darcy HDG, test/DarcyHDGTests.jl:125-135 (kappa^-1 on the mass term)
  __device__ double darcy(const Shared& sh, int row, int col) const {
    const CellGeo& g = sh.geo;
    const int D = T.D, np = T.np, nq = T.nq, nm = T.nm, na = T.na, ni = T.ni, n = T.n;
    const double tau = M.tau;
    if (row < na) {  // flux test v (component a, function i)
      const int a = row / np, i = row - a * np;
      if (col < na) {
        const int b = col / np, j = col - b * np;
        return a == b ? T.Mpp[i * np + j] * (g.detJ / g.coef) : 0.0;
      }
      if (col < ni) {  // -(div v) p
        const int j = col - na;
        double s = 0.0;
        for (int e = 0; e < D; ++e) s += g.Ji[e + D * a] * T.Gpq[(e * np + i) * nq + j];
        return -g.detJ * s;
      }
      if (col < n) {  // (v.n) lambda
        const int F = (col - ni) / nm, m = col - ni - F * nm;
        if (!((T.pmask[i] >> F) & 1)) return 0.0;
        return g.aF[F] * g.nF[F][a] * T.Fpm[(F * np + i) * nm + m];
      }
      double s = 0.0;  // rhs: v . f
      for (int q = 0; q < T.nqc; ++q) s += T.Bp[q * np + i] * sh.src[q][a];
      return s * g.detJ;
    }
    if (row < ni) {  // potential test w
      const int i = row - na;
      if (col < na) {  // -grad(w).q + w (q.n)
        const int b = col / np, j = col - b * np;
        double s = 0.0;
        for (int e = 0; e < D; ++e) s += g.Ji[e + D * b] * T.Gqp[(e * nq + i) * np + j];
        double r = -g.detJ * s;
        unsigned fm = T.qmask[i] & T.pmask[j];
        for (int F = 0; F < 2 * D; ++F)
          if ((fm >> F) & 1) r += g.aF[F] * g.nF[F][b] * T.Fqp[(F * nq + i) * np + j];
        return r;
      }
      if (col < ni) {  // tau w p (n.n_o)
        const int j = col - na;
        unsigned fm = T.qmask[i] & T.qmask[j];
        double r = 0.0;
        for (int F = 0; F < 2 * D; ++F)
          if ((fm >> F) & 1) r += tau * g.sF[F] * g.aF[F] * T.Fqq[(F * nq + i) * nq + j];
        return r;
      }
      if (col < n) {  // -tau w lambda (n.n_o)
        const int F = (col - ni) / nm, m = col - ni - F * nm;
        if (!((T.qmask[i] >> F) & 1)) return 0.0;
        return -tau * g.sF[F] * g.aF[F] * T.Fqm[(F * nq + i) * nm + m];
      }
      double s = 0.0;  // rhs: w g
      for (int q = 0; q < T.nqc; ++q) s += T.Bq[q * nq + i] * sh.src[q][D];
      return s * g.detJ;
    }
    // trace test mu on facet F
    const int F = (row - ni) / nm, m = row - ni - F * nm;
    if (col < na) {  // mu (q.n)
      const int b = col / np, j = col - b * np;
      if (!((T.pmask[j] >> F) & 1)) return 0.0;
      return g.aF[F] * g.nF[F][b] * T.Fpm[(F * np + j) * nm + m];
    }
    if (col < ni) {  // tau mu p (n.n_o)
      const int j = col - na;
      if (!((T.qmask[j] >> F) & 1)) return 0.0;
      return tau * g.sF[F] * g.aF[F] * T.Fqm[(F * nq + j) * nm + m];
    }
    if (col < n) {  // -tau mu lambda (n.n_o)
      const int G = (col - ni) / nm, m2 = col - ni - G * nm;
      if (G != F) return 0.0;
      return -tau * g.sF[F] * g.aF[F] * T.Fmm[m * nm + m2];
    }
    return 0.0;
  }

  // Elasticity HDG, test/LinearElasticityHDGTests.jl:140-150 (no projection)
  __device__ double elas(const Shared& sh, int row, int col) const {
    const CellGeo& g = sh.geo;
    const int D = T.D, np = T.np, nq = T.nq, nm = T.nm, na = T.na, ni = T.ni, n = T.n;
    const int ntf = T.ntf;
    const double tau = M.tau;
    if (row < na) {  // stress test v = p_i E_a
      const int a = row / np, i = row - a * np;
      int ia, ja;
      sym_pair(D, a, ia, ja);
      if (col < na) {
        const int b = col / np, j = col - b * np;
        return g.K[a][b] * T.Mpp[i * np + j] * g.detJ;
      }
      if (col < ni) {  // (div v) . u
        const int c = (col - na) / nq, j = col - na - c * nq;
        double s = 0.0;
        if (c == ja)
          for (int e = 0; e < D; ++e) s += g.Ji[e + D * ia] * T.Gpq[(e * np + i) * nq + j];
        if (c == ia && ia != ja)
          for (int e = 0; e < D; ++e) s += g.Ji[e + D * ja] * T.Gpq[(e * np + i) * nq + j];
        return g.detJ * s;
      }
      if (col < n) {  // -(v.n) . uhat
        const int F = (col - ni) / ntf, r = col - ni - F * ntf, c = r / nm, m = r - c * nm;
        if (!((T.pmask[i] >> F) & 1)) return 0.0;
        double vn = (c == ia ? g.nF[F][ja] : 0.0) + ((c == ja && ia != ja) ? g.nF[F][ia] : 0.0);
        return -g.aF[F] * vn * T.Fpm[(F * np + i) * nm + m];
      }
      return 0.0;
    }
    if (row < ni) {  // displacement test omega = q_j e_c
      const int c = (row - na) / nq, j = row - na - c * nq;
      if (col < na) {  // -omega . div(sigma)
        const int a = col / np, i = col - a * np;
        int ia, ja;
        sym_pair(D, a, ia, ja);
        double s = 0.0;
        if (c == ja)
          for (int e = 0; e < D; ++e) s += g.Ji[e + D * ia] * T.Gpq[(e * np + i) * nq + j];
        if (c == ia && ia != ja)
          for (int e = 0; e < D; ++e) s += g.Ji[e + D * ja] * T.Gpq[(e * np + i) * nq + j];
        return -g.detJ * s;
      }
      if (col < ni) {  // tau omega . u
        const int c2 = (col - na) / nq, j2 = col - na - c2 * nq;
        if (c2 != c) return 0.0;
        unsigned fm = T.qmask[j] & T.qmask[j2];
        double r = 0.0;
        for (int F = 0; F < 2 * D; ++F)
          if ((fm >> F) & 1) r += tau * g.aF[F] * T.Fqq[(F * nq + j) * nq + j2];
        return r;
      }
      if (col < n) {  // -tau omega . uhat
        const int F = (col - ni) / ntf, r = col - ni - F * ntf, c2 = r / nm, m = r - c2 * nm;
        if (c2 != c || !((T.qmask[j] >> F) & 1)) return 0.0;
        return -tau * g.aF[F] * T.Fqm[(F * nq + j) * nm + m];
      }
      double s = 0.0;  // rhs: -omega . f
      for (int q = 0; q < T.nqc; ++q) s += T.Bq[q * nq + j] * sh.src[q][c];
      return -s * g.detJ;
    }
    const int F = (row - ni) / ntf, rr = row - ni - F * ntf, c = rr / nm, m = rr - c * nm;
    if (col < na) {  // mu . (sigma n)
      const int a = col / np, i = col - a * np;
      int ia, ja;
      sym_pair(D, a, ia, ja);
      if (!((T.pmask[i] >> F) & 1)) return 0.0;
      double vn = (c == ia ? g.nF[F][ja] : 0.0) + ((c == ja && ia != ja) ? g.nF[F][ia] : 0.0);
      return g.aF[F] * vn * T.Fpm[(F * np + i) * nm + m];
    }
    if (col < ni) {  // -tau mu . u
      const int c2 = (col - na) / nq, j = col - na - c2 * nq;
      if (c2 != c || !((T.qmask[j] >> F) & 1)) return 0.0;
      return -tau * g.aF[F] * T.Fqm[(F * nq + j) * nm + m];
    }
    if (col < n) {  // tau mu . uhat
      const int G = (col - ni) / ntf, r2 = col - ni - G * ntf, c2 = r2 / nm, m2 = r2 - c2 * nm;
      if (G != F || c2 != c) return 0.0;
      return tau * g.aF[F] * T.Fmm[m * nm + m2];
    }
    return 0.0;
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
};
template <>
struct GenCall<FusedGen> {
  __device__ static double entry(const FusedGen& gen, const FusedShared& sh, int row, int col,
                                 int64_t) {
    return gen.entry(sh, row, col);
  }
};

// --------------------------------------------------------------------------
// Augmented elimination kernel
// --------------------------------------------------------------------------
template <int PR, int PC, int R, int CC, class Gen>
__global__ void __launch_bounds__(PR * PC)
    condense_kernel(Gen gen, CondenseOut out, int64_t ncells, int ni, int nb) {
  constexpr int NT = PR * PC;
  constexpr int NMAX = PR * R;
  constexpr int NCMAX = PC * CC;
  static_assert(PR <= 64 && (64 % PR) == 0, "thread-grid columns must sit inside one wavefront");
  static_assert(R <= 32, "row mask is 32 bits");
  const int n = ni + nb;  // rows
  const int ncol = n + 1; // columns incl. right-hand side
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
    for (int i = tid; i < ni; i += NT) {
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
    for (int j = 0; j < ni; ++j) {
      const int buf = j & 1;
      const int jc = j % PC, jl = j / PC;
      if (tc == jc) {
        double bav = -1.0, bval = 0.0;
        int bpos = INT_MAX, brow = -1;
#pragma unroll
        for (int c = 0; c < CC; ++c) {
          if (c == jl) {
#pragma unroll
            for (int r = 0; r < R; ++r) {
              const int row = tr + PR * r;
              if (row < ni && ((act >> r) & 1u)) {
                const double v = a[r][c];
                const double av = fabs(v);
                const int p = s_pos[row];
                if (av > bav || (av == bav && p < bpos)) {
                  bav = av;
                  bval = v;
                  bpos = p;
                  brow = row;
                }
              }
            }
          }
        }
#pragma unroll
        for (int off = 1; off < PR; off <<= 1) {
          const double oav = __shfl_xor(bav, off, PR);
          const double oval = __shfl_xor(bval, off, PR);
          const int opos = __shfl_xor(bpos, off, PR);
          const int orow = __shfl_xor(brow, off, PR);
          if (oav > bav || (oav == bav && opos < bpos)) {
            bav = oav;
            bval = oval;
            bpos = opos;
            brow = orow;
          }
        }
        const double rinv = bval != 0.0 ? 1.0 / bval : 0.0;
        if (tr == 0) {
          s_piv[buf] = brow;
          s_diag[j] = bval;
          if (bval == 0.0 && s_info == 0) s_info = j + 1;
        }
#pragma unroll
        for (int c = 0; c < CC; ++c) {
          if (c == jl) {
#pragma unroll
            for (int r = 0; r < R; ++r) {
              const int row = tr + PR * r;
              if (((act >> r) & 1u) && row != brow) s_lcol[buf][row] = a[r][c] * rinv;
            }
          }
        }
      }
      __syncthreads();
      const int p = s_piv[buf];
      if (tr == p % PR) {
        const int pl = p / PR;
#pragma unroll
        for (int r = 0; r < R; ++r) {
          if (r == pl) {
#pragma unroll
            for (int c = 0; c < CC; ++c) {
              const int col = tc + PC * c;
              if (col < ncol) s_prow[buf][col] = a[r][c];
            }
            act &= ~(1u << r);
          }
        }
      }
      if (tid == 0) {
        // LAPACK row interchange bookkeeping: rows at positions j and pos[p] swap
        const int pp = s_pos[p];
        const int q = s_at[j];
        s_at[pp] = q;
        s_pos[q] = pp;
        s_at[j] = p;
        s_pos[p] = j;
      }
      __syncthreads();
#pragma unroll
      for (int c = 0; c < CC; ++c) {
        if (c < jl) continue;  // whole local column already eliminated
        const int col = tc + PC * c;
        if (col <= j || col >= ncol) continue;
        const double u = s_prow[buf][col];
#pragma unroll
        for (int r = 0; r < R; ++r) {
          if ((act >> r) & 1u) {
            const int row = tr + PR * r;
            a[r][c] = fma(-s_lcol[buf][row], u, a[r][c]);
          }
        }
      }
    }

    __syncthreads();  // last step's pivot row / multipliers fully consumed

    // ---------------- back substitution on the interior right-hand sides -
    // (act now flags exactly the boundary rows; only needed when W is kept)
    const int info = s_info;
    if (out.W != nullptr) {
      for (int k = ni - 1; k >= 0; --k) {
        const int buf = k & 1;
        const int p = s_at[k];
        const double d = s_diag[k];
        if (tr == p % PR) {
          const int pl = p / PR;
#pragma unroll
          for (int r = 0; r < R; ++r) {
            if (r == pl) {
#pragma unroll
              for (int c = 0; c < CC; ++c) {
                const int col = tc + PC * c;
                if (col >= ni && col < ncol) {
                  const double x = d != 0.0 ? a[r][c] / d : 0.0;
                  a[r][c] = x;
                  s_prow[buf][col] = x;
                }
              }
            }
          }
        }
        const int kc = k % PC, kl = k / PC;
        if (tc == kc) {
#pragma unroll
          for (int c = 0; c < CC; ++c) {
            if (c == kl) {
#pragma unroll
              for (int r = 0; r < R; ++r) {
                const int row = tr + PR * r;
                if (row < ni && s_pos[row] < k) s_lcol[buf][row] = a[r][c];
              }
            }
          }
        }
        __syncthreads();
#pragma unroll
        for (int r = 0; r < R; ++r) {
          const int row = tr + PR * r;
          if (row < ni && s_pos[row] < k) {
            const double l = s_lcol[buf][row];
#pragma unroll
            for (int c = 0; c < CC; ++c) {
              const int col = tc + PC * c;
              if (col >= ni && col < ncol) a[r][c] = fma(-l, s_prow[buf][col], a[r][c]);
            }
          }
        }
      }
    }

    // ---------------- outputs ----------------
#pragma unroll
    for (int r = 0; r < R; ++r) {
      const int row = tr + PR * r;
      if (row >= n) continue;
#pragma unroll
      for (int c = 0; c < CC; ++c) {
        const int col = tc + PC * c;
        if (col < ni || col >= ncol) continue;
        const double v = a[r][c];
        if (row >= ni) {
          if (col < n) {
            if (out.S) out.S[(cell * nb + (col - ni)) * nb + (row - ni)] = v;
          } else {
            if (out.g) out.g[cell * nb + (row - ni)] = v;
          }
        } else if (out.W) {
          out.W[(cell * (nb + 1) + (col - ni)) * ni + s_pos[row]] = v;
        }
      }
    }
    if (tid == 0 && out.info) out.info[cell] = info;
    __syncthreads();  // LDS reuse by the next cell
  }
}

// --------------------------------------------------------------------------
// Recovery: x_i = w - W lambda_cell, scattered into the free-value vector
// [q of all cells; u of all cells; lambda_free].
// --------------------------------------------------------------------------
__global__ void recover_kernel(const double* __restrict__ W, const int32_t* __restrict__ cell_dofs,
                               const double* __restrict__ lam_free,
                               const double* __restrict__ lam_dir, double* __restrict__ out,
                               int64_t ncells, int ni, int nb, int na, int cells_per_block) {
  extern __shared__ double s_lam[];  // [cells_per_block][nb]
  const int64_t c0 = (int64_t)blockIdx.x * cells_per_block;
  for (int t = threadIdx.x; t < cells_per_block * nb; t += blockDim.x) {
    const int lc = t / nb, j = t - lc * nb;
    const int64_t cell = c0 + lc;
    double v = 0.0;
    if (cell < ncells) {
      const int32_t d = cell_dofs[cell * nb + j];
      v = d >= 0 ? lam_free[d] : lam_dir[-(int64_t)d - 1];
    }
    s_lam[t] = v;
  }
  __syncthreads();
  for (int t = threadIdx.x; t < cells_per_block * ni; t += blockDim.x) {
    const int lc = t / ni, i = t - lc * ni;
    const int64_t cell = c0 + lc;
    if (cell >= ncells) continue;
    const double* Wc = W + cell * (int64_t)(nb + 1) * ni;
    double s = Wc[(int64_t)nb * ni + i];
    for (int j = 0; j < nb; ++j) s = fma(-Wc[(int64_t)j * ni + i], s_lam[lc * nb + j], s);
    if (i < na)
      out[cell * na + i] = s;
    else
      out[ncells * na + cell * (ni - na) + (i - na)] = s;
  }
}

__global__ void copy_kernel(const double* __restrict__ src, double* __restrict__ dst, int64_t n) {
  for (int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; t < n;
       t += (int64_t)gridDim.x * blockDim.x)
    dst[t] = src[t];
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
  const int ntf = P.ntf, nb = P.nb;
  for (int64_t G = blockIdx.x; G < P.nff; G += gridDim.x) {
    const int p0 = P.pptr[G], np = P.pptr[G + 1] - p0;
    const int per_col = np * ntf;
    const int64_t base = P.fptr[G];
    for (int t = threadIdx.x; t < per_col * ntf; t += blockDim.x) {
      const int mc = t / per_col, rem = t - mc * per_col;
      const int pr = rem / ntf, m = rem - pr * ntf;
      double v = 0.0;
#pragma unroll
      for (int k = 0; k < 2; ++k) {
        const int32_t c = P.pcell[2 * (p0 + pr) + k];
        if (c >= 0) {
          const uint8_t lp = P.plpos[2 * (p0 + pr) + k];
          const int lg = lp & 15, lf = lp >> 4;
          v += S[((int64_t)c * nb + (lg * ntf + mc)) * nb + lf * ntf + m];
        }
      }
      nzval[base + t] = v;
    }
  }
}

// rhs: for free dof (G, m), sum over cells around G (ascending) of
// g_c[i] - sum_{j Dirichlet} S_c[i, j] lam_dir (Gridap _attach_dirichlet)
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
      const int i = P.glpos[2 * G + k] * ntf + m;
      double v = g[(int64_t)c * nb + i];
      const int32_t* d = cell_dofs + (int64_t)c * nb;
      for (int j = 0; j < nb; ++j) {
        const int32_t dj = d[j];
        if (dj < 0) {
          const double ld = lam_dir[-(int64_t)dj - 1];
          if (ld != 0.0) v -= S[((int64_t)c * nb + j) * nb + i] * ld;
        }
      }
      acc += v;
    }
    rhs[t] = acc;
  }
}

// Per-cell element systems (debug / parity entry point).
__global__ void cell_systems_kernel(FusedGen gen, int64_t first, int64_t count, double* A,
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
// Launchers
// --------------------------------------------------------------------------
struct Bucket {
  int pr, pc, r, cc;
};

// Thread-grid / register-tile buckets by rows n and columns n+1.
static bool pick_bucket(int n, Bucket& b) {
  const Bucket list[] = {{8, 8, 4, 4}, {8, 8, 6, 6}, {8, 8, 9, 9}, {32, 16, 6, 11}, {32, 32, 8, 8}};
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

template <class Gen>
static hipError_t launch_condense(const Gen& gen, const CondenseOut& out, int64_t ncells, int ni,
                                  int nb, hipStream_t st) {
  if (ncells <= 0) return hipSuccess;
  Bucket b;
  if (!pick_bucket(ni + nb, b)) return hipErrorInvalidValue;
  const int grid = grid_for(ncells, b.pr * b.pc);
#define HDG_LAUNCH(PR, PC, R, CC)                                                          \
  if (b.pr == PR && b.pc == PC && b.r == R && b.cc == CC) {                                \
    hipLaunchKernelGGL((condense_kernel<PR, PC, R, CC, Gen>), dim3(grid), dim3(PR * PC), 0, \
                       st, gen, out, ncells, ni, nb);                                      \
    return hipGetLastError();                                                              \
  }
  HDG_LAUNCH(8, 8, 4, 4)
  HDG_LAUNCH(8, 8, 6, 6)
  HDG_LAUNCH(8, 8, 9, 9)
  HDG_LAUNCH(32, 16, 6, 11)
  HDG_LAUNCH(32, 32, 8, 8)
#undef HDG_LAUNCH
  return hipErrorInvalidValue;
}

bool condense_supported(int n) {
  Bucket b;
  return pick_bucket(n, b);
}

hipError_t launch_condense_fused(const Tables& T, const MeshDev& M, const CondenseOut& out,
                                 hipStream_t st) {
  FusedGen gen{T, M};
  return launch_condense(gen, out, M.ncells, T.ni, T.nb, st);
}

hipError_t launch_condense_dense(const DenseIn& in, int64_t ncells, const CondenseOut& out,
                                 hipStream_t st) {
  DenseGen gen{in};
  return launch_condense(gen, out, ncells, in.ni, in.nb, st);
}

hipError_t launch_backward_dense(const DenseIn& in, const double* x, int64_t ncells, double* xi,
                                 int32_t* info, hipStream_t st) {
  BackwardGen gen{in, x};
  CondenseOut out{nullptr, nullptr, xi, nullptr, info};
  // interior-only system: nb = 0, the single rhs column is b_i - A_ib x
  return launch_condense(gen, out, ncells, in.ni, 0, st);
}

hipError_t launch_recover(const double* W, const int32_t* cell_dofs, const double* lam_free,
                          const double* lam_dir, double* out, int64_t ncells, int ni, int nb,
                          int na, int64_t nfree, hipStream_t st) {
  if (ncells > 0) {
    int cpb = 256 / ni;
    if (cpb < 1) cpb = 1;
    const int64_t grid = (ncells + cpb - 1) / cpb;
    const size_t shm = sizeof(double) * (size_t)cpb * nb;
    hipLaunchKernelGGL(recover_kernel, dim3((unsigned)grid), dim3(256), shm, st, W, cell_dofs,
                       lam_free, lam_dir, out, ncells, ni, nb, na, cpb);
  }
  if (nfree > 0)
    hipLaunchKernelGGL(copy_kernel, dim3(1024), dim3(256), 0, st, lam_free, out + ncells * ni,
                       nfree);
  return hipGetLastError();
}

hipError_t launch_symbolic(const PairTable& P, int32_t* rowval, hipStream_t st) {
  if (P.nff > 0) {
    const int grid = (int)(P.nff < 65536 ? P.nff : 65536);
    hipLaunchKernelGGL(symbolic_kernel, dim3(grid), dim3(256), 0, st, P, rowval);
  }
  return hipGetLastError();
}

hipError_t launch_assemble(const PairTable& P, const double* S, const double* g,
                           const int32_t* cell_dofs, const double* lam_dir, double* nzval,
                           double* rhs, hipStream_t st) {
  if (P.nff > 0) {
    const int grid = (int)(P.nff < 65536 ? P.nff : 65536);
    hipLaunchKernelGGL(assemble_kernel, dim3(grid), dim3(256), 0, st, P, S, nzval);
    const int64_t tot = P.nff * P.ntf;
    const int grid2 = (int)((tot + 255) / 256 < 65536 ? (tot + 255) / 256 : 65536);
    hipLaunchKernelGGL(rhs_kernel, dim3(grid2), dim3(256), 0, st, P, S, g, cell_dofs, lam_dir,
                       rhs);
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

hipError_t launch_count_nonzero(const int32_t* info, int64_t n, int64_t* count_dev,
                                hipStream_t st) {
  hipError_t e = hipMemsetAsync(count_dev, 0, sizeof(int64_t), st);
  if (e != hipSuccess) return e;
  if (n > 0) {
    const int grid = (int)((n + 255) / 256 < 1024 ? (n + 255) / 256 : 1024);
    hipLaunchKernelGGL(count_nonzero_kernel, dim3(grid), dim3(256), 0, st, info, n,
                       reinterpret_cast<unsigned long long*>(count_dev));
  }
  return hipGetLastError();
}

hipError_t launch_cell_systems(const Tables& T, const MeshDev& M, int64_t first, int64_t count,
                               double* A, double* b, hipStream_t st) {
  if (count <= 0) return hipSuccess;
  FusedGen gen{T, M};
  const int grid = (int)(count < 4096 ? count : 4096);
  hipLaunchKernelGGL(cell_systems_kernel, dim3(grid), dim3(256), 0, st, gen, first, count, A, b);
  return hipGetLastError();
}

}  // namespace hdg
