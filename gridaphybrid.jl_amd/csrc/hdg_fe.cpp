// hdg_fe.cpp -- reference finite-element tables for the fused cell kernels.
//
// The reference integrates every weak-form term per cell with Gauss rules of
// degree 2(order+1) (test/DarcyHDGTests.jl:57-61) through Gridap's lazy
// arrays.  For affine cells that quadrature is exact for the polynomial
// products of the bilinear forms, so it is done once here on the reference
// cell; the GPU kernels apply the per-cell affine pull-back (|det J|, J^-1,
// facet normals and measures, kappa, owner sign).  The right-hand side keeps
// its per-cell quadrature (tables Bp/Bq hold basis values times weights).
//
// Bases: Lagrangian, equispaced nodes, nodes ordered face by face as Gridap's
// n-cube polytopes enumerate them (vertices, edges, faces, interior; within a
// dimension by free-direction set, fixed coordinates first-direction
// fastest).  space Q: tensor products of 1D Lagrange polynomials; space P:
// total degree <= k with the lattice nodes of that order (nodal basis from a
// Vandermonde solve in centred monomials).
#include <algorithm>
#include <array>
#include <cmath>
#include <cstring>
#include <map>
#include <mutex>
#include <utility>

#include "hdg_internal.h"

namespace hdg {
void build_program(TablesHost& t);
int inline_program(TablesHost& t, std::string& err);
void detect_flux_block(TablesHost& t);
void detect_flux_sym(TablesHost& t);
void build_flux_ref(TablesHost& t);
namespace {

// Gauss-Legendre nodes/weights on [0,1]: Newton on P_n from the asymptotic
// guess, three-term recurrence for P_n and P_n'.
void gauss_legendre01(int n, std::vector<double>& x, std::vector<double>& w) {
  x.assign(n, 0.0);
  w.assign(n, 0.0);
  for (int i = 0; i < n; ++i) {
    double t = std::cos(M_PI * (4.0 * i + 3.0) / (4.0 * n + 2.0));
    double pn = 0.0, dpn = 1.0;
    for (int it = 0; it < 100; ++it) {
      double pm1 = 1.0, p = t;
      if (n == 0) p = 1.0;
      for (int k = 2; k <= n; ++k) {
        double pk = ((2 * k - 1) * t * p - (k - 1) * pm1) / k;
        pm1 = p;
        p = pk;
      }
      pn = p;
      dpn = n * (pm1 - t * p) / (1.0 - t * t);
      double step = pn / dpn;
      t -= step;
      if (std::fabs(step) <= 1e-16 * std::fabs(t) + 1e-300) break;
    }
    {
      double pm1 = 1.0, p = t;
      for (int k = 2; k <= n; ++k) {
        double pk = ((2 * k - 1) * t * p - (k - 1) * pm1) / k;
        pm1 = p;
        p = pk;
      }
      dpn = n * (pm1 - t * p) / (1.0 - t * t);
    }
    // t_i descending in i: map so that x is ascending
    x[i] = 0.5 * (1.0 - t);
    w[i] = 1.0 / ((1.0 - t * t) * dpn * dpn);
  }
}

struct Basis {
  int d = 0, k = 0;
  bool pspace = false;
  int n = 0;
  std::vector<int> lat;       // [n][3] lattice index of each node (nodes = lat/k)
  std::vector<double> coef;   // P space: [n][n] monomial coefficients (row: function)
  std::vector<int> mono;      // P space: [n][3] monomial exponents

  double node(int i, int a) const { return k == 0 ? 0.5 : double(lat[3 * i + a]) / k; }

  // value and reference gradient of function i at x
  void eval(const double* x, double* val, double* grad) const {
    if (k == 0) {
      for (int i = 0; i < n; ++i) {
        val[i] = 1.0;
        if (grad)
          for (int a = 0; a < 3; ++a) grad[3 * i + a] = 0.0;
      }
      return;
    }
    if (!pspace) {
      // tensor Lagrange: prod_a l_{lat_a}(x_a)
      double L[3][8], dL[3][8];
      for (int a = 0; a < d; ++a) lagrange1d(x[a], L[a], dL[a]);
      for (int i = 0; i < n; ++i) {
        double v = 1.0;
        for (int a = 0; a < d; ++a) v *= L[a][lat[3 * i + a]];
        val[i] = v;
        if (grad)
          for (int g = 0; g < 3; ++g) {
            if (g >= d) {
              grad[3 * i + g] = 0.0;
              continue;
            }
            double dv = 1.0;
            for (int a = 0; a < d; ++a) dv *= (a == g) ? dL[a][lat[3 * i + a]] : L[a][lat[3 * i + a]];
            grad[3 * i + g] = dv;
          }
      }
      return;
    }
    // P space: phi_i = sum_m coef[i][m] prod_a (x_a - 1/2)^{e_a}
    std::vector<double> mv(n), md(3 * n);
    for (int m = 0; m < n; ++m) {
      double v = 1.0;
      for (int a = 0; a < d; ++a) v *= ipow(x[a] - 0.5, mono[3 * m + a]);
      mv[m] = v;
      for (int g = 0; g < 3; ++g) {
        if (g >= d || mono[3 * m + g] == 0) {
          md[3 * m + g] = 0.0;
          continue;
        }
        double dv = 1.0;
        for (int a = 0; a < d; ++a) {
          int e = mono[3 * m + a];
          dv *= (a == g) ? e * ipow(x[a] - 0.5, e - 1) : ipow(x[a] - 0.5, e);
        }
        md[3 * m + g] = dv;
      }
    }
    for (int i = 0; i < n; ++i) {
      double v = 0.0, g0 = 0.0, g1 = 0.0, g2 = 0.0;
      for (int m = 0; m < n; ++m) {
        double c = coef[(size_t)i * n + m];
        v += c * mv[m];
        g0 += c * md[3 * m];
        g1 += c * md[3 * m + 1];
        g2 += c * md[3 * m + 2];
      }
      val[i] = v;
      if (grad) {
        grad[3 * i] = g0;
        grad[3 * i + 1] = g1;
        grad[3 * i + 2] = g2;
      }
    }
  }

  static double ipow(double b, int e) {
    double r = 1.0;
    for (int i = 0; i < e; ++i) r *= b;
    return r;
  }

  // 1D Lagrange basis on nodes j/k, j = 0..k
  void lagrange1d(double x, double* L, double* dL) const {
    for (int j = 0; j <= k; ++j) {
      double xj = double(j) / k, v = 1.0, dv = 0.0;
      for (int m = 0; m <= k; ++m) {
        if (m == j) continue;
        double xm = double(m) / k, den = xj - xm;
        dv = dv * (x - xm) / den + v / den;
        v *= (x - xm) / den;
      }
      L[j] = v;
      dL[j] = dv;
    }
  }
};

// Nodes of the d-cube in Gridap's face order (see file header).
void cube_lattice(int d, int k, bool pspace, std::vector<int>& lat) {
  lat.clear();
  if (k == 0 || d == 0) {
    lat.assign(3, 0);
    return;
  }
  for (int fd = 0; fd <= d; ++fd) {
    // subsets of size fd of {0..d-1}, lexicographic
    std::vector<std::vector<int>> subsets;
    std::vector<int> comb(fd);
    for (int i = 0; i < fd; ++i) comb[i] = i;
    while (true) {
      subsets.push_back(comb);
      int i = fd - 1;
      while (i >= 0 && comb[i] == d - fd + i) --i;
      if (i < 0) break;
      ++comb[i];
      for (int j = i + 1; j < fd; ++j) comb[j] = comb[j - 1] + 1;
    }
    for (const auto& fr : subsets) {
      bool isfree[3] = {false, false, false};
      for (int a : fr) isfree[a] = true;
      std::vector<int> fixed;
      for (int a = 0; a < d; ++a)
        if (!isfree[a]) fixed.push_back(a);
      int nfix = (int)fixed.size();
      int nown = 1;
      for (int i = 0; i < fd; ++i) nown *= (k - 1);
      for (int fv = 0; fv < (1 << nfix); ++fv)
        for (int o = 0; o < nown; ++o) {
          int a[3] = {0, 0, 0}, t = o;
          for (int i = 0; i < fd; ++i) {
            a[fr[i]] = 1 + t % (k - 1);
            t /= (k - 1);
          }
          for (int i = 0; i < nfix; ++i) a[fixed[i]] = ((fv >> i) & 1) ? k : 0;
          if (pspace && a[0] + a[1] + a[2] > k) continue;
          lat.push_back(a[0]);
          lat.push_back(a[1]);
          lat.push_back(a[2]);
        }
    }
  }
}

bool solve_dense(int n, std::vector<double>& A, std::vector<double>& B, int nrhs) {
  // row-major A[n][n], B[n][nrhs]; Gaussian elimination with partial pivoting
  for (int j = 0; j < n; ++j) {
    int p = j;
    for (int i = j + 1; i < n; ++i)
      if (std::fabs(A[(size_t)i * n + j]) > std::fabs(A[(size_t)p * n + j])) p = i;
    if (A[(size_t)p * n + j] == 0.0) return false;
    if (p != j) {
      for (int c = 0; c < n; ++c) std::swap(A[(size_t)j * n + c], A[(size_t)p * n + c]);
      for (int c = 0; c < nrhs; ++c) std::swap(B[(size_t)j * nrhs + c], B[(size_t)p * nrhs + c]);
    }
    for (int i = j + 1; i < n; ++i) {
      double l = A[(size_t)i * n + j] / A[(size_t)j * n + j];
      if (l == 0.0) continue;
      for (int c = j; c < n; ++c) A[(size_t)i * n + c] -= l * A[(size_t)j * n + c];
      for (int c = 0; c < nrhs; ++c) B[(size_t)i * nrhs + c] -= l * B[(size_t)j * nrhs + c];
    }
  }
  for (int j = n - 1; j >= 0; --j)
    for (int c = 0; c < nrhs; ++c) {
      double s = B[(size_t)j * nrhs + c];
      for (int m = j + 1; m < n; ++m) s -= A[(size_t)j * n + m] * B[(size_t)m * nrhs + c];
      B[(size_t)j * nrhs + c] = s / A[(size_t)j * n + j];
    }
  return true;
}

bool make_basis(int d, int k, bool pspace, Basis& b) {
  b.d = d;
  b.k = k;
  b.pspace = pspace && k > 0;
  cube_lattice(d, k, pspace, b.lat);
  b.n = (int)b.lat.size() / 3;
  if (!b.pspace) return true;
  // monomials of total degree <= k
  b.mono.clear();
  for (int e2 = 0; e2 <= (d > 2 ? k : 0); ++e2)
    for (int e1 = 0; e1 <= (d > 1 ? k : 0); ++e1)
      for (int e0 = 0; e0 <= k; ++e0)
        if (e0 + e1 + e2 <= k) {
          b.mono.push_back(e0);
          b.mono.push_back(e1);
          b.mono.push_back(e2);
        }
  int n = b.n;
  if ((int)b.mono.size() / 3 != n) return false;
  // V[node][mono]; coefficients: V^T C^T = I  -> solve V X = I, coef = X^T
  std::vector<double> V((size_t)n * n), X((size_t)n * n, 0.0);
  for (int i = 0; i < n; ++i) {
    for (int m = 0; m < n; ++m) {
      double v = 1.0;
      for (int a = 0; a < d; ++a) v *= Basis::ipow(b.node(i, a) - 0.5, b.mono[3 * m + a]);
      V[(size_t)i * n + m] = v;
    }
    X[(size_t)i * n + i] = 1.0;
  }
  if (!solve_dense(n, V, X, n)) return false;
  // X = V^{-1}: phi_i(x) = sum_m X[m][i] mono_m(x)
  b.coef.assign((size_t)n * n, 0.0);
  for (int i = 0; i < n; ++i)
    for (int m = 0; m < n; ++m) b.coef[(size_t)i * n + m] = X[(size_t)m * n + i];
  return true;
}

// facet f of the reference D-cube: fixes direction D-1-f/2 at f%2
void facet_point(int D, int f, const double* s, double* xi) {
  int fix = D - 1 - f / 2, j = 0;
  for (int a = 0; a < D; ++a) xi[a] = (a == fix) ? double(f % 2) : s[j++];
  for (int a = D; a < 3; ++a) xi[a] = 0.0;
}

}  // namespace

static int build_tables_uncached(int phys, int D, int space, int kp, int kq, int km,
                                 int quad_degree, int proj, TablesHost& t, std::string& err);

// The tables of a class depend only on its description, and meshes (and the
// footprint queries of a serving loop) ask for the same few classes again and
// again: they are built once per process (up to 40 ms for 3D Darcy k=3) and
// copied from then on.
int build_tables(int phys, int D, int space, int kp, int kq, int km, int quad_degree, int proj,
                 TablesHost& t, std::string& err) {
  static std::mutex mu;
  static std::map<std::array<int, 8>, TablesHost> cache;
  const std::array<int, 8> key{phys, D, space, kp, kq, km, quad_degree, proj};
  {
    std::lock_guard<std::mutex> lock(mu);
    auto it = cache.find(key);
    if (it != cache.end()) {
      t = it->second;
      return 0;
    }
  }
  const int rc = build_tables_uncached(phys, D, space, kp, kq, km, quad_degree, proj, t, err);
  if (rc == 0) {
    std::lock_guard<std::mutex> lock(mu);
    cache.emplace(key, t);
  }
  return rc;
}

static int build_tables_uncached(int phys, int D, int space, int kp, int kq, int km,
                                 int quad_degree, int proj, TablesHost& t, std::string& err) {
  if (D != 2 && D != 3) {
    err = "dimension must be 2 or 3";
    return -1;
  }
  bool ps = space == 1;
  Basis bp, bq, bm;
  if (!make_basis(D, kp, ps, bp) || !make_basis(D, kq, ps, bq) || !make_basis(D - 1, km, ps, bm)) {
    err = "could not build reference bases";
    return -1;
  }
  t = TablesHost();
  t.phys = phys;
  t.D = D;
  t.nF = 2 * D;
  t.np = bp.n;
  t.nq = bq.n;
  t.nm = bm.n;
  if (phys == kDarcy) {
    t.ncp = D;
    t.ncq = 1;
    t.ncm = 1;
  } else {
    t.ncp = D * (D + 1) / 2;
    t.ncq = D;
    t.ncm = D;
  }
  t.na = t.ncp * t.np;
  t.nbf = t.ncq * t.nq;
  t.ntf = t.ncm * t.nm;
  t.ni = t.na + t.nbf;
  t.nb = t.nF * t.ntf;
  t.n = t.ni + t.nb;

  int m1 = (quad_degree + 2) / 2;  // ceil((degree+1)/2) points per direction
  std::vector<double> x1, w1;
  gauss_legendre01(m1, x1, w1);
  int nqc = 1, nqf = 1;
  for (int a = 0; a < D; ++a) nqc *= m1;
  for (int a = 0; a < D - 1; ++a) nqf *= m1;
  t.nqc = nqc;

  const int np = t.np, nq = t.nq, nm = t.nm;
  t.Mpp.assign((size_t)np * np, 0.0);
  t.Gpq.assign((size_t)D * np * nq, 0.0);
  t.Gqp.assign((size_t)D * nq * np, 0.0);
  t.Bp.assign((size_t)nqc * np, 0.0);
  t.Bq.assign((size_t)nqc * nq, 0.0);
  t.xq.assign((size_t)nqc * 3, 0.0);
  t.wq.assign((size_t)nqc, 0.0);
  std::vector<double> vp(np), gp(3 * np), vq(nq), gq(3 * nq), vm(nm);
  for (int q = 0; q < nqc; ++q) {
    double xi[3] = {0, 0, 0}, w = 1.0;
    int r = q;
    for (int a = 0; a < D; ++a) {
      xi[a] = x1[r % m1];
      w *= w1[r % m1];
      r /= m1;
    }
    for (int a = 0; a < 3; ++a) t.xq[3 * q + a] = xi[a];
    t.wq[q] = w;
    bp.eval(xi, vp.data(), gp.data());
    bq.eval(xi, vq.data(), gq.data());
    for (int i = 0; i < np; ++i) {
      t.Bp[(size_t)q * np + i] = vp[i] * w;
      for (int j = 0; j < np; ++j) t.Mpp[(size_t)i * np + j] += vp[i] * vp[j] * w;
      for (int e = 0; e < D; ++e)
        for (int j = 0; j < nq; ++j) t.Gpq[((size_t)e * np + i) * nq + j] += gp[3 * i + e] * vq[j] * w;
    }
    for (int i = 0; i < nq; ++i) {
      t.Bq[(size_t)q * nq + i] = vq[i] * w;
      for (int e = 0; e < D; ++e)
        for (int j = 0; j < np; ++j) t.Gqp[((size_t)e * nq + i) * np + j] += gq[3 * i + e] * vp[j] * w;
    }
  }
  const int nF = t.nF;
  t.Fqp.assign((size_t)nF * nq * np, 0.0);
  t.Fqq.assign((size_t)nF * nq * nq, 0.0);
  t.Fpm.assign((size_t)nF * np * nm, 0.0);
  t.Fqm.assign((size_t)nF * nq * nm, 0.0);
  t.Fmm.assign((size_t)nm * nm, 0.0);
  t.pmask.assign(np, 0);
  t.qmask.assign(nq, 0);
  std::vector<double> pp2((size_t)nF * np, 0.0), qq2((size_t)nF * nq, 0.0);
  for (int f = 0; f < nF; ++f)
    for (int q = 0; q < nqf; ++q) {
      double s[3] = {0, 0, 0}, w = 1.0;
      int r = q;
      for (int a = 0; a < D - 1; ++a) {
        s[a] = x1[r % m1];
        w *= w1[r % m1];
        r /= m1;
      }
      double xi[3];
      facet_point(D, f, s, xi);
      bp.eval(xi, vp.data(), nullptr);
      bq.eval(xi, vq.data(), nullptr);
      bm.eval(s, vm.data(), nullptr);
      for (int i = 0; i < nq; ++i) {
        qq2[(size_t)f * nq + i] += vq[i] * vq[i] * w;
        for (int j = 0; j < np; ++j) t.Fqp[((size_t)f * nq + i) * np + j] += vq[i] * vp[j] * w;
        for (int j = 0; j < nq; ++j) t.Fqq[((size_t)f * nq + i) * nq + j] += vq[i] * vq[j] * w;
        for (int j = 0; j < nm; ++j) t.Fqm[((size_t)f * nq + i) * nm + j] += vq[i] * vm[j] * w;
      }
      for (int i = 0; i < np; ++i) {
        pp2[(size_t)f * np + i] += vp[i] * vp[i] * w;
        for (int j = 0; j < nm; ++j) t.Fpm[((size_t)f * np + i) * nm + j] += vp[i] * vm[j] * w;
      }
      if (f == 0)
        for (int i = 0; i < nm; ++i)
          for (int j = 0; j < nm; ++j) t.Fmm[(size_t)i * nm + j] += vm[i] * vm[j] * w;
    }
  if (proj && phys != kDarcy) {
    // bulk-to-skeleton L2 projection P_m (src/GridapAPIExtensions.jl:453-500:
    // dofs = A \ B with A = (mu, mu')_F, B = (mu, u)_F): the displacement's
    // facet trace is replaced by its projection onto the trace space, so
    // tau (w, P_m u)_F = tau (P_m w, P_m u)_F = tau B^T A^-1 B.  The other tau
    // terms are unchanged: (mu, P_m u)_F = (mu, u)_F for mu in the trace space.
    std::vector<double> Acopy = t.Fmm, I((size_t)nm * nm, 0.0);
    for (int i = 0; i < nm; ++i) I[(size_t)i * nm + i] = 1.0;
    if (!solve_dense(nm, Acopy, I, nm)) {
      err = "singular trace mass matrix";
      return -1;
    }
    // I now holds Fmm^-1 (row-major)
    for (int f = 0; f < nF; ++f)
      for (int i = 0; i < nq; ++i)
        for (int j = 0; j < nq; ++j) {
          double v = 0.0;
          for (int m = 0; m < nm; ++m) {
            double w = 0.0;
            for (int m2 = 0; m2 < nm; ++m2)
              w += I[(size_t)m * nm + m2] * t.Fqm[((size_t)f * nq + j) * nm + m2];
            v += t.Fqm[((size_t)f * nq + i) * nm + m] * w;
          }
          t.Fqq[((size_t)f * nq + i) * nq + j] = v;
        }
  }
  // support masks: bit f set when the trace of the function on facet f is not
  // identically zero (its facet L2 norm is positive)
  for (int f = 0; f < nF; ++f) {
    for (int i = 0; i < np; ++i)
      if (pp2[(size_t)f * np + i] > 1e-300) t.pmask[i] |= (uint8_t)(1u << f);
    for (int i = 0; i < nq; ++i)
      if (qq2[(size_t)f * nq + i] > 1e-300) t.qmask[i] |= (uint8_t)(1u << f);
  }
  t.trace_nodes.assign((size_t)nm * (D - 1 > 0 ? D - 1 : 1), 0.0);
  for (int i = 0; i < nm; ++i)
    for (int a = 0; a < D - 1; ++a) t.trace_nodes[(size_t)i * (D - 1) + a] = bm.node(i, a);
  build_program(t);
  detect_flux_block(t);
  detect_flux_sym(t);
  build_flux_ref(t);
  return inline_program(t, err);
}

// ---------------------------------------------------------------------------
// Entry programs.  Every entry (row, col) of the condensed cell matrix
// [A_ii A_ib; A_bi A_bb] is a short linear combination of per-cell
// coefficients alpha[id] (affine pull-back factors, kappa, owner signs) with
// reference-integral values; the program lists, entry by entry in
// column-major order, the (id, value) terms with a nonzero value.
// ---------------------------------------------------------------------------
namespace {
struct Prog {
  std::vector<int32_t>& off;
  std::vector<uint8_t>& id;
  std::vector<double>& val;
  void add(int cid, double v) {
    if (v != 0.0) {
      id.push_back((uint8_t)cid);
      val.push_back(v);
    }
  }
};
inline void sym_ij(int D, int a, int& i, int& j) {
  int t = 0;
  for (int x = 0; x < D; ++x)
    for (int y = x; y < D; ++y) {
      if (t == a) {
        i = x;
        j = y;
        return;
      }
      ++t;
    }
  i = j = 0;
}
}  // namespace

void build_program(TablesHost& t) {
  const int D = t.D, np = t.np, nq = t.nq, nm = t.nm, na = t.na, ni = t.ni, n = t.n;
  const int nF = t.nF;
  t.prog_off.assign(1, 0);
  t.prog_id.clear();
  t.prog_val.clear();
  Prog P{t.prog_off, t.prog_id, t.prog_val};
  auto Mpp = [&](int i, int j) { return t.Mpp[(size_t)i * np + j]; };
  auto Gpq = [&](int e, int i, int j) { return t.Gpq[((size_t)e * np + i) * nq + j]; };
  auto Fqq = [&](int F, int i, int j) { return t.Fqq[((size_t)F * nq + i) * nq + j]; };
  auto Fpm = [&](int F, int i, int m) { return t.Fpm[((size_t)F * np + i) * nm + m]; };
  auto Fqm = [&](int F, int i, int m) { return t.Fqm[((size_t)F * nq + i) * nm + m]; };
  auto Fmm = [&](int m, int k) { return t.Fmm[(size_t)m * nm + k]; };
  auto pm = [&](int i, int F) { return (t.pmask[i] >> F) & 1; };
  auto qm = [&](int i, int F) { return (t.qmask[i] >> F) & 1; };
  if (t.phys == kDarcy) {
    // alpha: [0] detJ/kappa; [1 + e*D + a] -detJ Ji(e,a); [kAn + 3F + a] aF nF_a;
    //        [kTs + F] tau sF aF
    const int kJ = 1, kAn = 10, kTs = 28;
    t.nalpha = 34;
    for (int col = 0; col < n; ++col)
      for (int row = 0; row < n; ++row) {
        if (row < na) {
          const int a = row / np, i = row % np;
          if (col < na) {
            const int b = col / np, j = col % np;
            if (a == b) P.add(0, Mpp(i, j));
          } else if (col < ni) {
            const int j = col - na;
            for (int e = 0; e < D; ++e) P.add(kJ + e * D + a, Gpq(e, i, j));
          } else {
            const int F = (col - ni) / nm, m = (col - ni) % nm;
            if (pm(i, F)) P.add(kAn + 3 * F + a, Fpm(F, i, m));
          }
        } else if (row < ni) {
          const int i = row - na;
          if (col < na) {
            // -(grad w, q) + <w, q.n> (test/DarcyHDGTests.jl:125-129) integrated
            // by parts on the affine cell: (w, div q), i.e. minus the transpose
            // of the (div v, u) block.  With the exact quadrature of the tables
            // the two forms agree to rounding, and the flux coupling is then
            // exactly transpose-symmetric in the program (detect_flux_sym).
            const int b = col / np, j = col % np;
            for (int e = 0; e < D; ++e) P.add(kJ + e * D + b, -Gpq(e, j, i));
          } else if (col < ni) {
            const int j = col - na;
            for (int F = 0; F < nF; ++F)
              if (qm(i, F) && qm(j, F)) P.add(kTs + F, Fqq(F, i, j));
          } else {
            const int F = (col - ni) / nm, m = (col - ni) % nm;
            if (qm(i, F)) P.add(kTs + F, -Fqm(F, i, m));
          }
        } else {
          const int F = (row - ni) / nm, m = (row - ni) % nm;
          if (col < na) {
            const int b = col / np, j = col % np;
            if (pm(j, F)) P.add(kAn + 3 * F + b, Fpm(F, j, m));
          } else if (col < ni) {
            const int j = col - na;
            if (qm(j, F)) P.add(kTs + F, Fqm(F, j, m));
          } else {
            const int G = (col - ni) / nm, m2 = (col - ni) % nm;
            if (G == F) P.add(kTs + F, -Fmm(m, m2));
          }
        }
        t.prog_off.push_back((int32_t)t.prog_id.size());
      }
  } else {
    // alpha: [kK + a*6 + b] K_ab detJ; [kJ + e*D + d] detJ Ji(e,d);
    //        [kAn + 3F + d] aF nF_d; [kT + F] tau aF
    const int kK = 0, kJ = 36, kAn = 45, kT = 63;
    t.nalpha = 69;
    const int ntf = t.ntf;
    for (int col = 0; col < n; ++col)
      for (int row = 0; row < n; ++row) {
        if (row < na) {
          const int a = row / np, i = row % np;
          int ia, ja;
          sym_ij(D, a, ia, ja);
          if (col < na) {
            const int b = col / np, j = col % np;
            P.add(kK + a * 6 + b, Mpp(i, j));
          } else if (col < ni) {
            const int c = (col - na) / nq, j = (col - na) % nq;
            if (c == ja)
              for (int e = 0; e < D; ++e) P.add(kJ + e * D + ia, Gpq(e, i, j));
            if (c == ia && ia != ja)
              for (int e = 0; e < D; ++e) P.add(kJ + e * D + ja, Gpq(e, i, j));
          } else {
            const int F = (col - ni) / ntf, r = (col - ni) % ntf, c = r / nm, m = r % nm;
            if (pm(i, F)) {
              if (c == ia) P.add(kAn + 3 * F + ja, -Fpm(F, i, m));
              if (c == ja && ia != ja) P.add(kAn + 3 * F + ia, -Fpm(F, i, m));
            }
          }
        } else if (row < ni) {
          const int c = (row - na) / nq, j = (row - na) % nq;
          if (col < na) {
            const int a = col / np, i = col % np;
            int ia, ja;
            sym_ij(D, a, ia, ja);
            if (c == ja)
              for (int e = 0; e < D; ++e) P.add(kJ + e * D + ia, -Gpq(e, i, j));
            if (c == ia && ia != ja)
              for (int e = 0; e < D; ++e) P.add(kJ + e * D + ja, -Gpq(e, i, j));
          } else if (col < ni) {
            const int c2 = (col - na) / nq, j2 = (col - na) % nq;
            if (c2 == c)
              for (int F = 0; F < nF; ++F)
                if (qm(j, F) && qm(j2, F)) P.add(kT + F, Fqq(F, j, j2));
          } else {
            const int F = (col - ni) / ntf, r = (col - ni) % ntf, c2 = r / nm, m = r % nm;
            if (c2 == c && qm(j, F)) P.add(kT + F, -Fqm(F, j, m));
          }
        } else {
          const int F = (row - ni) / ntf, rr = (row - ni) % ntf, c = rr / nm, m = rr % nm;
          if (col < na) {
            const int a = col / np, i = col % np;
            int ia, ja;
            sym_ij(D, a, ia, ja);
            if (pm(i, F)) {
              if (c == ia) P.add(kAn + 3 * F + ja, Fpm(F, i, m));
              if (c == ja && ia != ja) P.add(kAn + 3 * F + ia, Fpm(F, i, m));
            }
          } else if (col < ni) {
            const int c2 = (col - na) / nq, j = (col - na) % nq;
            if (c2 == c && qm(j, F)) P.add(kT + F, -Fqm(F, j, m));
          } else {
            const int G = (col - ni) / ntf, r2 = (col - ni) % ntf, c2 = r2 / nm, m2 = r2 % nm;
            if (G == F && c2 == c) P.add(kT + F, Fmm(m, m2));
          }
        }
        t.prog_off.push_back((int32_t)t.prog_id.size());
      }
  }
}

// Inline-first-term form of the entry program (prog_head / prog_val0 /
// prog_xid / prog_xval), same terms in the same order.
int inline_program(TablesHost& t, std::string& err) {
  const size_t ne = t.prog_off.empty() ? 0 : t.prog_off.size() - 1;
  t.prog_head.assign(ne, 0u);
  t.prog_val0.assign(ne, 0.0);
  t.prog_xid.clear();
  t.prog_xval.clear();
  for (size_t e = 0; e < ne; ++e) {
    const int k0 = t.prog_off[e], k1 = t.prog_off[e + 1], cnt = k1 - k0;
    if (cnt > 15) {
      err = "entry program term count above 15";
      return 1;
    }
    const size_t x = t.prog_xid.size();
    if (x >= (1u << 20)) {
      err = "entry program too long";
      return 1;
    }
    if (cnt > 0) {
      t.prog_head[e] = (uint32_t)t.prog_id[k0] | ((uint32_t)cnt << 8) | ((uint32_t)x << 12);
      t.prog_val0[e] = t.prog_val[k0];
      for (int k = k0 + 1; k < k1; ++k) {
        t.prog_xid.push_back(t.prog_id[k]);
        t.prog_xval.push_back(t.prog_val[k]);
      }
    }
  }
  // row-major terms of the interior rows (entries are column-major: e = col n + row)
  const int n = t.n, ni = t.ni;
  t.prog_rptr.assign((size_t)ni + 1, 0);
  for (int col = 0; col < n; ++col)
    for (int row = 0; row < ni; ++row) {
      const int e = col * n + row;
      t.prog_rptr[row + 1] += t.prog_off[e + 1] - t.prog_off[e];
    }
  for (int r = 0; r < ni; ++r) t.prog_rptr[r + 1] += t.prog_rptr[r];
  const size_t nt = (size_t)t.prog_rptr[ni];
  t.prog_rcol.assign(nt, 0);
  t.prog_rid.assign(nt, 0);
  t.prog_rval.assign(nt, 0.0);
  std::vector<int32_t> fill(t.prog_rptr.begin(), t.prog_rptr.end() - 1);
  for (int col = 0; col < n; ++col)
    for (int row = 0; row < ni; ++row) {
      const int e = col * n + row;
      for (int k = t.prog_off[e]; k < t.prog_off[e + 1]; ++k) {
        const int32_t at = fill[row]++;
        t.prog_rcol[at] = (uint16_t)col;
        t.prog_rid[at] = t.prog_id[k];
        t.prog_rval[at] = t.prog_val[k];
      }
    }
  // terms are in column order within a row: the boundary columns are a suffix
  t.prog_rbsplit.assign((size_t)ni, 0);
  for (int r = 0; r < ni; ++r) {
    int k = t.prog_rptr[r];
    while (k < t.prog_rptr[r + 1] && t.prog_rcol[k] < ni) ++k;
    t.prog_rbsplit[r] = k;
  }
  return 0;
}

}  // namespace hdg

namespace hdg {
// ---------------------------------------------------------------------------
// Kronecker flux block.  Scans the entry program of the leading na x na block
// of A_ii: every entry must be either empty or one term whose value is the
// scalar mass-matrix entry M[i][j] (Mpp) and whose coefficient id depends on
// the component pair (a, b) only.  Then A_qq = C (x) M per cell with
// C[a][b] = alpha[cid(a,b)], and A_qq^-1 = C^-1 (x) M^-1.  M^-1 is computed
// once here: Gauss-Jordan with partial pivoting in long double, then two
// steps of iterative refinement of each column.
// ---------------------------------------------------------------------------
void detect_flux_block(TablesHost& t) {
  t.fb_nc = 0;
  t.fb_minv.clear();
  const int np = t.np, na = t.na, n = t.n;
  if (np <= 0 || na <= 0 || na % np != 0 || na >= t.ni) return;
  const int nc = na / np;
  if (nc > 6) return;
  int cid[36];
  for (int& c : cid) c = -2;  // -2: not seen yet
  for (int col = 0; col < na; ++col)
    for (int row = 0; row < na; ++row) {
      const int a = row / np, i = row % np, b = col / np, j = col % np;
      const double m = t.Mpp[(size_t)i * np + j];
      const size_t e = (size_t)col * n + row;
      const int k0 = t.prog_off[e], k1 = t.prog_off[e + 1];
      int& c = cid[a * 6 + b];
      if (k1 == k0) {
        if (m != 0.0) {
          if (c >= 0) return;
          c = -1;
        }
        continue;
      }
      if (k1 - k0 != 1 || t.prog_val[k0] != m) return;
      const int id = t.prog_id[k0];
      if (c == -1 || (c >= 0 && c != id)) return;
      c = id;
    }
  std::vector<long double> A((size_t)np * np), X((size_t)np * np, 0.0L);
  for (int i = 0; i < np * np; ++i) A[i] = t.Mpp[i];
  for (int i = 0; i < np; ++i) X[(size_t)i * np + i] = 1.0L;
  std::vector<long double> W = A;
  for (int c = 0; c < np; ++c) {
    int p = c;
    for (int r = c + 1; r < np; ++r)
      if (std::fabs(W[(size_t)r * np + c]) > std::fabs(W[(size_t)p * np + c])) p = r;
    if (W[(size_t)p * np + c] == 0.0L) return;
    if (p != c)
      for (int k = 0; k < np; ++k) {
        std::swap(W[(size_t)p * np + k], W[(size_t)c * np + k]);
        std::swap(X[(size_t)p * np + k], X[(size_t)c * np + k]);
      }
    const long double d = W[(size_t)c * np + c];
    for (int k = 0; k < np; ++k) {
      W[(size_t)c * np + k] /= d;
      X[(size_t)c * np + k] /= d;
    }
    for (int r = 0; r < np; ++r) {
      if (r == c) continue;
      const long double f = W[(size_t)r * np + c];
      if (f == 0.0L) continue;
      for (int k = 0; k < np; ++k) {
        W[(size_t)r * np + k] -= f * W[(size_t)c * np + k];
        X[(size_t)r * np + k] -= f * X[(size_t)c * np + k];
      }
    }
  }
  // X <- X + X (I - A X), twice
  for (int it = 0; it < 2; ++it) {
    std::vector<long double> R((size_t)np * np), D((size_t)np * np);
    for (int i = 0; i < np; ++i)
      for (int j = 0; j < np; ++j) {
        long double s = i == j ? 1.0L : 0.0L;
        for (int k = 0; k < np; ++k) s -= A[(size_t)i * np + k] * X[(size_t)k * np + j];
        R[(size_t)i * np + j] = s;
      }
    for (int i = 0; i < np; ++i)
      for (int j = 0; j < np; ++j) {
        long double s = 0.0L;
        for (int k = 0; k < np; ++k) s += X[(size_t)i * np + k] * R[(size_t)k * np + j];
        D[(size_t)i * np + j] = s;
      }
    for (int i = 0; i < np * np; ++i) X[i] += D[i];
  }
  t.fb_minv.resize((size_t)np * np);
  for (int i = 0; i < np * np; ++i) t.fb_minv[i] = (double)X[i];
  for (int k = 0; k < 36; ++k) t.fb_cid[k] = (int8_t)(cid[k] >= 0 ? cid[k] : -1);
  t.fb_nc = nc;
}

// Transpose symmetry of the flux coupling: for every flux unknown k and
// every other unknown r, entry (r, k) equals s entry (k, r), coefficient by
// coefficient (the values of each coefficient index summed over the entry's
// terms agree to a relative 1e-14), with one sign s_u for the remaining
// interior rows and one s_b for the boundary rows.  Then [A_uq; A_bq] can be
// read from the generated [A_qu A_qb] instead of being generated again.
void detect_flux_sym(TablesHost& t) {
  t.fb_sym = 0;
  if (t.fb_nc <= 0) return;
  const int na = t.na, n = t.n, ni = t.ni;
  std::vector<double> va(256, 0.0), vb(256, 0.0);
  std::vector<int> ids;
  int su = 0, sb = 0;
  for (int r = na; r < n; ++r)
    for (int k = 0; k < na; ++k) {
      const size_t e1 = (size_t)k * n + r, e2 = (size_t)r * n + k;  // (r, k) and (k, r)
      ids.clear();
      for (int q = t.prog_off[e1]; q < t.prog_off[e1 + 1]; ++q) {
        va[t.prog_id[q]] += t.prog_val[q];
        ids.push_back(t.prog_id[q]);
      }
      for (int q = t.prog_off[e2]; q < t.prog_off[e2 + 1]; ++q) {
        vb[t.prog_id[q]] += t.prog_val[q];
        ids.push_back(t.prog_id[q]);
      }
      int& s = r < ni ? su : sb;
      bool ok = true;
      for (int id : ids) {
        const double a = va[id], b = vb[id];
        if (a == 0.0 && b == 0.0) continue;
        if (a == 0.0 || b == 0.0) {
          ok = false;
          break;
        }
        if (s == 0) s = (a > 0.0) == (b > 0.0) ? 1 : -1;
        if (std::fabs(a - s * b) > 1e-14 * std::max(std::fabs(a), std::fabs(b))) {
          ok = false;
          break;
        }
      }
      for (int id : ids) va[id] = vb[id] = 0.0;
      if (!ok) return;
    }
  t.fb_sym = 1 | (su < 0 ? 2 : 0) | (sb < 0 ? 4 : 0);
}

// Reference decomposition of the Darcy flux Schur complement (family 10).
// The program's flux rows (a, i) of [A_qu A_qb] are
//   A_qu: sum_e alpha[1 + e D + a] Gpq(e, i, j),  alpha[1 + e D + a] = -c_ea,
//   A_qb: alpha[10 + 3F + a] Fpm(F, i, m),        alpha[10 + 3F + a] = aF nF_a,
// and on an affine cell with det J > 0 the facet area times normal is the
// cofactor image of the reference normal, aF nF_a = sum_e c_ea nhat_Fe.  So the
// row block of component a is sum_e c_ea G_e with the reference matrices
//   G_e[i][j] = -Gpq(e, i, j) (potential columns),
//   G_e[i][(F, m)] = nhat_Fe Fpm(F, i, m) (trace columns),
// and, C being (detJ/kappa) I = alpha_0 I, the flux Schur term
//   [A_qu A_qb]^T (C^-1 (x) M^-1) [A_qu A_qb]
//     = (1/alpha_0) sum_(e<=e') (sum_a c_ea c_e'a) P_(e,e').
// Computed once here in long double.
void build_flux_ref(TablesHost& t) {
  t.fr_npair = 0;
  t.fr_P.clear();
  t.fr_G.clear();
  const int D = t.D, np = t.np, nq = t.nq, nm = t.nm, nF = t.nF;
  if (t.phys != kDarcy || t.fb_nc != D || (t.fb_sym & 1) == 0 || t.fb_minv.empty()) return;
  // s_u = -1 (the integrated-by-parts u-q block), s_b = +1
  if ((t.fb_sym & 6) != 2) return;
  for (int a = 0; a < D; ++a)
    for (int b = 0; b < D; ++b)
      if (t.fb_cid[6 * a + b] != (a == b ? 0 : -1)) return;  // C = alpha_0 I
  const int nr = nq + nF * nm;
  // the kernel keeps at most three program terms per entry of R0 in registers
  for (int c = t.na; c < t.n; ++c)
    for (int r = t.na; r < t.n; ++r) {
      const size_t e = (size_t)c * t.n + r;
      if (t.prog_off[e + 1] - t.prog_off[e] > 3) return;
    }
  std::vector<long double> G((size_t)D * np * nr, 0.0L);
  for (int e = 0; e < D; ++e)
    for (int i = 0; i < np; ++i) {
      for (int j = 0; j < nq; ++j)
        G[((size_t)e * np + i) * nr + j] = -(long double)t.Gpq[((size_t)e * np + i) * nq + j];
      for (int F = 0; F < nF; ++F) {
        const int fix = D - 1 - F / 2;
        if (fix != e) continue;
        const long double nh = (F % 2) ? 1.0L : -1.0L;
        for (int m = 0; m < nm; ++m)
          G[((size_t)e * np + i) * nr + nq + F * nm + m] =
              nh * (long double)t.Fpm[((size_t)F * np + i) * nm + m];
      }
    }
  // H_e = M^-1 G_e
  std::vector<long double> H((size_t)D * np * nr, 0.0L);
  for (int e = 0; e < D; ++e)
    for (int i = 0; i < np; ++i)
      for (int c = 0; c < nr; ++c) {
        long double v = 0.0L;
        for (int k = 0; k < np; ++k)
          v += (long double)t.fb_minv[(size_t)i * np + k] * G[((size_t)e * np + k) * nr + c];
        H[((size_t)e * np + i) * nr + c] = v;
      }
  const int npair = D * (D + 1) / 2;
  t.fr_P.assign((size_t)npair * nr * nr, 0.0);
  int p = 0;
  for (int e = 0; e < D; ++e)
    for (int e2 = e; e2 < D; ++e2, ++p)
      for (int r = 0; r < nr; ++r)
        for (int c = 0; c < nr; ++c) {
          long double v = 0.0L;
          for (int i = 0; i < np; ++i) {
            v += G[((size_t)e * np + i) * nr + r] * H[((size_t)e2 * np + i) * nr + c];
            if (e2 != e) v += G[((size_t)e2 * np + i) * nr + r] * H[((size_t)e * np + i) * nr + c];
          }
          t.fr_P[((size_t)p * nr + r) * nr + c] = (double)v;
        }
  t.fr_G.resize(G.size());
  for (size_t k = 0; k < G.size(); ++k) t.fr_G[k] = (double)G[k];
  t.fr_npair = npair;
}

}  // namespace hdg
