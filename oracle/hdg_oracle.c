/*
 * hdg_oracle.c -- CPU FP64 restatement of GridapHybrid's per-cell hybrid
 * pipeline.  TEST INFRASTRUCTURE ONLY (the checker): used by tests/, the smoke
 * check and the CPU-baseline leg of bench.py; never linked into the product.
 *
 * What it restates, and from where:
 *   - cell + facet quadrature of the hybridizable weak forms, facet sums in
 *     the order ((f1+f2)+f3)+... (src/SumFacetsMap.jl:19-30), bulk added
 *     before skeleton (src/HybridAffineFEOperators.jl:271-274);
 *     Darcy HDG weak form: test/DarcyHDGTests.jl:125-135;
 *     elasticity HDG weak form: test/LinearElasticityHDGTests.jl:140-150.
 *   - owner sign n.n_o (src/Skeleton.jl:549-562, src/SkeletonArrays.jl:219-229)
 *   - StaticCondensationMap (src/StaticCondensationMap.jl:152-196)
 *   - BackwardStaticCondensationMap (src/BackwardStaticCondensationMap.jl:61-102)
 *   - skeleton dof ids (src/HybridAffineFEOperators.jl:388-439) and assembly
 *     with Dirichlet lifting (src/HybridAffineFEOperators.jl:38-46).
 * Unpinned Gridap-internal conventions (node order, facet enumeration) follow
 * the best-knowledge statement of Gridap 0.18; see the tests for what pins
 * each function.
 */
#include "hdg_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

#ifndef M_PI
#define M_PI 3.14159265358979323846
#endif

#define MAXN 512   /* max basis functions of one scalar space */
#define MAXQ 1024  /* max quadrature points */

/* ------------------------------------------------------------------------ */
/* Gauss-Legendre on [0,1] (Newton iteration on P_n).                        */
/* ------------------------------------------------------------------------ */
static void gauss01(int n, double* x, double* w) {
  for (int i = 0; i < n; i++) {
    double z = cos(M_PI * (i + 0.75) / (n + 0.5));
    double p0 = 1.0, p1 = z, dp = 1.0;
    for (int it = 0; it < 200; it++) {
      p0 = 1.0;
      p1 = z;
      for (int m = 2; m <= n; m++) {
        double p2 = ((2.0 * m - 1.0) * z * p1 - (m - 1.0) * p0) / m;
        p0 = p1;
        p1 = p2;
      }
      if (n == 1) {
        p1 = z;
        p0 = 1.0;
      }
      dp = n * (z * p1 - p0) / (z * z - 1.0);
      double dz = p1 / dp;
      z -= dz;
      if (fabs(dz) < 1e-17) break;
    }
    p0 = 1.0;
    p1 = z;
    for (int m = 2; m <= n; m++) {
      double p2 = ((2.0 * m - 1.0) * z * p1 - (m - 1.0) * p0) / m;
      p0 = p1;
      p1 = p2;
    }
    dp = n * (z * p1 - p0) / (z * z - 1.0);
    /* ascending order on [0,1] */
    x[i] = 0.5 * (1.0 - z);
    w[i] = 1.0 / ((1.0 - z * z) * dp * dp);
  }
}

/* ------------------------------------------------------------------------ */
/* Scalar nodal spaces on the reference d-cube via a Legendre-product        */
/* Vandermonde (basis phi_i = sum_m C[m][i] L_{a_m}(x)).                     */
/* ------------------------------------------------------------------------ */
typedef struct {
  int d;          /* dimension of the reference cube (0..3) */
  int n;          /* number of basis functions */
  int exps[MAXN][3];
  double nodes[MAXN][3];
  double* C;      /* n x n, C[m*n + i] */
} space_t;

static void legendre01(int m, double x, double* L, double* dL) {
  /* L[0..m], dL[0..m] of shifted Legendre on [0,1] */
  double t = 2.0 * x - 1.0;
  L[0] = 1.0;
  dL[0] = 0.0;
  if (m >= 1) {
    L[1] = t;
    dL[1] = 2.0;
  }
  for (int j = 1; j < m; j++) {
    L[j + 1] = ((2.0 * j + 1.0) * t * L[j] - j * L[j - 1]) / (j + 1.0);
    dL[j + 1] = ((2.0 * j + 1.0) * (2.0 * L[j] + t * dL[j]) - j * dL[j - 1]) / (j + 1.0);
  }
}

/* dense LU with partial pivoting, col-major, n x n, solve nrhs in place */
int orc_dense_solve(int n, int nrhs, double* A, double* B) {
  int info = 0;
  int* piv = (int*)malloc(sizeof(int) * (n > 0 ? n : 1));
  for (int j = 0; j < n; j++) {
    int p = j;
    double amax = fabs(A[j + (size_t)j * n]);
    for (int i = j + 1; i < n; i++) {
      double v = fabs(A[i + (size_t)j * n]);
      if (v > amax) {
        amax = v;
        p = i;
      }
    }
    piv[j] = p;
    if (A[p + (size_t)j * n] == 0.0) {
      if (!info) info = j + 1;
      continue;
    }
    if (p != j)
      for (int c = 0; c < n; c++) {
        double t = A[j + (size_t)c * n];
        A[j + (size_t)c * n] = A[p + (size_t)c * n];
        A[p + (size_t)c * n] = t;
      }
    double r = 1.0 / A[j + (size_t)j * n];
    for (int i = j + 1; i < n; i++) A[i + (size_t)j * n] *= r;
    for (int c = j + 1; c < n; c++) {
      double u = A[j + (size_t)c * n];
      if (u != 0.0)
        for (int i = j + 1; i < n; i++) A[i + (size_t)c * n] -= A[i + (size_t)j * n] * u;
    }
  }
  if (!info) {
    for (int r = 0; r < nrhs; r++) {
      double* x = B + (size_t)r * n;
      for (int j = 0; j < n; j++)
        if (piv[j] != j) {
          double t = x[j];
          x[j] = x[piv[j]];
          x[piv[j]] = t;
        }
      for (int j = 0; j < n; j++)
        for (int i = j + 1; i < n; i++) x[i] -= A[i + (size_t)j * n] * x[j];
      for (int j = n - 1; j >= 0; j--) {
        x[j] /= A[j + (size_t)j * n];
        for (int i = 0; i < j; i++) x[i] -= A[i + (size_t)j * n] * x[j];
      }
    }
  }
  free(piv);
  return info;
}

/* Node order on the d-cube (Gridap n-cube faces: by dimension; within a
 * dimension, free-direction sets in lexicographic combination order; fixed
 * coordinates lexicographic with the first fixed direction fastest; the own
 * nodes of each face lexicographic in its free coordinates). */
static int cube_nodes(int d, int k, int pspace, double nodes[][3], int exps[][3]) {
  int n = 0;
  if (d == 0) {
    nodes[0][0] = nodes[0][1] = nodes[0][2] = 0.0;
    return 1;
  }
  if (k == 0) {
    for (int a = 0; a < 3; a++) nodes[0][a] = a < d ? 0.5 : 0.0;
    exps[0][0] = exps[0][1] = exps[0][2] = 0;
    return 1;
  }
  for (int fd = 0; fd <= d; fd++) {
    /* combinations of fd free directions out of d, lexicographic */
    int comb[3];
    for (int i = 0; i < fd; i++) comb[i] = i;
    for (;;) {
      int isfree[3] = {0, 0, 0};
      for (int i = 0; i < fd; i++) isfree[comb[i]] = 1;
      int fixed[3], nfixed = 0;
      for (int a = 0; a < d; a++)
        if (!isfree[a]) fixed[nfixed++] = a;
      for (int fv = 0; fv < (1 << nfixed); fv++) {
        /* own nodes: free coords in 1..k-1 */
        int nown = 1;
        for (int i = 0; i < fd; i++) nown *= (k - 1);
        for (int o = 0; o < nown; o++) {
          int a[3] = {0, 0, 0};
          int t = o;
          for (int i = 0; i < fd; i++) {
            a[comb[i]] = 1 + t % (k - 1);
            t /= (k - 1);
          }
          for (int i = 0; i < nfixed; i++) a[fixed[i]] = ((fv >> i) & 1) ? k : 0;
          int sum = 0;
          for (int q = 0; q < d; q++) sum += a[q];
          if (pspace && sum > k) continue;
          for (int q = 0; q < 3; q++) nodes[n][q] = q < d ? (double)a[q] / k : 0.0;
          n++;
        }
      }
      /* next combination */
      int i = fd - 1;
      while (i >= 0 && comb[i] == d - fd + i) i--;
      if (i < 0) break;
      comb[i]++;
      for (int j = i + 1; j < fd; j++) comb[j] = comb[j - 1] + 1;
    }
  }
  /* polynomial exponents */
  int m = 0;
  int kk = k;
  for (int a2 = 0; a2 <= (d > 2 ? kk : 0); a2++)
    for (int a1 = 0; a1 <= (d > 1 ? kk : 0); a1++)
      for (int a0 = 0; a0 <= kk; a0++) {
        int s = a0 + a1 + a2;
        int mx = a0 > a1 ? a0 : a1;
        mx = mx > a2 ? mx : a2;
        if (pspace ? (s > k) : (mx > k)) continue;
        exps[m][0] = a0;
        exps[m][1] = a1;
        exps[m][2] = a2;
        m++;
      }
  return n == m ? n : -1;
}

static int space_init(space_t* s, int d, int k, int pspace) {
  memset(s, 0, sizeof(*s));
  s->d = d;
  int n = cube_nodes(d, k, pspace, s->nodes, s->exps);
  if (n <= 0) return -1;
  s->n = n;
  s->C = (double*)malloc(sizeof(double) * n * n);
  double* V = (double*)malloc(sizeof(double) * n * n);
  double L[3][16], dL[3][16];
  for (int j = 0; j < n; j++) {
    for (int a = 0; a < 3; a++) legendre01(k, a < d ? s->nodes[j][a] : 0.0, L[a], dL[a]);
    for (int m = 0; m < n; m++) {
      double v = 1.0;
      for (int a = 0; a < d; a++) v *= L[a][s->exps[m][a]];
      V[j + (size_t)m * n] = v; /* row j (node), col m (poly) */
    }
  }
  /* C = V^{-1}: solve V C = I */
  for (int i = 0; i < n * n; i++) s->C[i] = 0.0;
  for (int i = 0; i < n; i++) s->C[i + (size_t)i * n] = 1.0;
  int info = orc_dense_solve(n, n, V, s->C);
  free(V);
  return info ? -2 : 0;
}

static void space_free(space_t* s) { free(s->C); s->C = NULL; }

/* values and reference gradients of all basis functions at x */
static void space_eval(const space_t* s, const double* x, double* val, double* grad /*[n][3]*/) {
  int kmax = 0;
  for (int m = 0; m < s->n; m++)
    for (int a = 0; a < 3; a++)
      if (s->exps[m][a] > kmax) kmax = s->exps[m][a];
  double L[3][16], dL[3][16];
  for (int a = 0; a < 3; a++) legendre01(kmax, a < s->d ? x[a] : 0.0, L[a], dL[a]);
  double P[MAXN], dP[MAXN][3];
  for (int m = 0; m < s->n; m++) {
    double v = 1.0;
    for (int a = 0; a < s->d; a++) v *= L[a][s->exps[m][a]];
    P[m] = v;
    for (int g = 0; g < 3; g++) {
      double dv = g < s->d ? 1.0 : 0.0;
      if (g < s->d)
        for (int a = 0; a < s->d; a++) dv *= (a == g) ? dL[a][s->exps[m][a]] : L[a][s->exps[m][a]];
      dP[m][g] = dv;
    }
  }
  for (int i = 0; i < s->n; i++) {
    double v = 0.0, g0 = 0.0, g1 = 0.0, g2 = 0.0;
    for (int m = 0; m < s->n; m++) {
      double c = s->C[m + (size_t)i * s->n];
      v += c * P[m];
      g0 += c * dP[m][0];
      g1 += c * dP[m][1];
      g2 += c * dP[m][2];
    }
    val[i] = v;
    if (grad) {
      grad[3 * i + 0] = g0;
      grad[3 * i + 1] = g1;
      grad[3 * i + 2] = g2;
    }
  }
}

/* ------------------------------------------------------------------------ */
/* Reference cube facets (Gridap order): facet f fixes direction D-1-f/2 at  */
/* value f%2; its parametrisation uses the remaining directions ascending.   */
/* ------------------------------------------------------------------------ */
static void facet_to_cell(int D, int f, const double* s, double* xi) {
  int fixdir = D - 1 - f / 2;
  int j = 0;
  for (int a = 0; a < D; a++) {
    if (a == fixdir)
      xi[a] = (double)(f % 2);
    else
      xi[a] = s[j++];
  }
}
static void facet_ref_normal(int D, int f, double* nref) {
  int fixdir = D - 1 - f / 2;
  for (int a = 0; a < D; a++) nref[a] = 0.0;
  nref[fixdir] = (f % 2) ? 1.0 : -1.0;
}

/* small dense helpers for D x D (col-major) */
static double det3(int D, const double* J) {
  if (D == 2) return J[0] * J[3] - J[2] * J[1];
  return J[0] * (J[4] * J[8] - J[7] * J[5]) - J[3] * (J[1] * J[8] - J[7] * J[2]) +
         J[6] * (J[1] * J[5] - J[4] * J[2]);
}
static void inv3(int D, const double* J, double* Ji) {
  double d = det3(D, J);
  if (D == 2) {
    Ji[0] = J[3] / d;
    Ji[3] = J[0] / d;
    Ji[2] = -J[2] / d;
    Ji[1] = -J[1] / d;
    return;
  }
  /* J col-major: J[r + 3c] */
#define Jm(r, c) J[(r) + 3 * (c)]
  Ji[0 + 3 * 0] = (Jm(1, 1) * Jm(2, 2) - Jm(1, 2) * Jm(2, 1)) / d;
  Ji[0 + 3 * 1] = (Jm(0, 2) * Jm(2, 1) - Jm(0, 1) * Jm(2, 2)) / d;
  Ji[0 + 3 * 2] = (Jm(0, 1) * Jm(1, 2) - Jm(0, 2) * Jm(1, 1)) / d;
  Ji[1 + 3 * 0] = (Jm(1, 2) * Jm(2, 0) - Jm(1, 0) * Jm(2, 2)) / d;
  Ji[1 + 3 * 1] = (Jm(0, 0) * Jm(2, 2) - Jm(0, 2) * Jm(2, 0)) / d;
  Ji[1 + 3 * 2] = (Jm(0, 2) * Jm(1, 0) - Jm(0, 0) * Jm(1, 2)) / d;
  Ji[2 + 3 * 0] = (Jm(1, 0) * Jm(2, 1) - Jm(1, 1) * Jm(2, 0)) / d;
  Ji[2 + 3 * 1] = (Jm(0, 1) * Jm(2, 0) - Jm(0, 0) * Jm(2, 1)) / d;
  Ji[2 + 3 * 2] = (Jm(0, 0) * Jm(1, 1) - Jm(0, 1) * Jm(1, 0)) / d;
#undef Jm
}

/* ------------------------------------------------------------------------ */
/* Problem sizes                                                             */
/* ------------------------------------------------------------------------ */
static int ncomb(int n, int k) {
  if (k < 0 || k > n) return 0;
  long r = 1;
  for (int i = 1; i <= k; i++) r = r * (n - k + i) / i;
  return (int)r;
}
static int scalar_dim(int d, int k, int pspace) {
  if (d == 0) return 1;
  if (pspace) return ncomb(k + d, d);
  int n = 1;
  for (int a = 0; a < d; a++) n *= (k + 1);
  return n;
}

int orc_sizes(const orc_problem* p, int* n_flux, int* n_pot, int* n_tf, int* ni, int* nb) {
  int D = p->D;
  int ps = p->space == ORC_SPACE_P;
  int nF = 2 * D;
  int a, b, t;
  if (p->phys == ORC_PHYS_DARCY) {
    a = D * scalar_dim(D, p->k_flux, ps);
    b = scalar_dim(D, p->k_pot, ps);
    t = scalar_dim(D - 1, p->k_trace, ps);
  } else {
    int nsym = D * (D + 1) / 2;
    a = nsym * scalar_dim(D, p->k_flux, ps);
    b = D * scalar_dim(D, p->k_pot, ps);
    t = D * scalar_dim(D - 1, p->k_trace, ps);
  }
  if (n_flux) *n_flux = a;
  if (n_pot) *n_pot = b;
  if (n_tf) *n_tf = t;
  if (ni) *ni = a + b;
  if (nb) *nb = nF * t;
  return 0;
}

static int npts1(const orc_problem* p) { return (p->quad_degree + 2) / 2; } /* ceil((d+1)/2) */

int orc_quad_sizes(const orc_problem* p, int* nq_cell, int* nq_facet) {
  int m = npts1(p);
  int c = 1, f = 1;
  for (int a = 0; a < p->D; a++) c *= m;
  for (int a = 0; a < p->D - 1; a++) f *= m;
  if (nq_cell) *nq_cell = c;
  if (nq_facet) *nq_facet = f;
  return 0;
}

/* ------------------------------------------------------------------------ */
/* Cartesian topology                                                        */
/* ------------------------------------------------------------------------ */
int orc_cartesian_counts(int D, const int64_t* n, int64_t* ncells, int64_t* nfacets) {
  int64_t nc = 1;
  for (int a = 0; a < D; a++) nc *= n[a];
  int64_t nf = 0;
  for (int a = 0; a < D; a++) {
    int64_t t = n[a] + 1;
    for (int b = 0; b < D; b++)
      if (b != a) t *= n[b];
    nf += t;
  }
  *ncells = nc;
  *nfacets = nf;
  return 0;
}

int orc_cartesian_topology(int D, const int64_t* n, int32_t* cell_facets, int8_t* cell_sign,
                           int32_t* facet_cells, int8_t* facet_lpos, int8_t* facet_boundary) {
  int64_t nc, nf;
  orc_cartesian_counts(D, n, &nc, &nf);
  int nF = 2 * D;
  int64_t stride[3] = {1, n[0], D > 1 ? n[0] * n[1] : 0};
  int64_t next = 0;
  for (int64_t c = 0; c < nc; c++) {
    int64_t idx[3] = {0, 0, 0};
    int64_t t = c;
    for (int a = 0; a < D; a++) {
      idx[a] = t % n[a];
      t /= n[a];
    }
    for (int f = 0; f < nF; f++) {
      int dir = D - 1 - f / 2;
      int side = f % 2;
      int64_t nb_cell = -1;
      if (side == 0 && idx[dir] > 0) nb_cell = c - stride[dir];
      if (side == 1 && idx[dir] < n[dir] - 1) nb_cell = c + stride[dir];
      int64_t fid;
      if (nb_cell >= 0 && nb_cell < c) {
        /* already numbered by the neighbour through its opposite facet */
        fid = cell_facets[nb_cell * nF + (f ^ 1)];
        facet_cells[2 * fid + 1] = (int32_t)c;
        facet_lpos[2 * fid + 1] = (int8_t)f;
        cell_sign[c * nF + f] = -1;
      } else {
        fid = next++;
        facet_cells[2 * fid] = (int32_t)c;
        facet_cells[2 * fid + 1] = -1;
        facet_lpos[2 * fid] = (int8_t)f;
        facet_lpos[2 * fid + 1] = -1;
        facet_boundary[fid] = nb_cell < 0 ? 1 : 0;
        cell_sign[c * nF + f] = 1;
      }
      cell_facets[c * nF + f] = (int32_t)fid;
    }
  }
  return next == nf ? 0 : -1;
}

int orc_cartesian_geometry(int D, const int64_t* n, const double* domain, double* x0, double* J) {
  int64_t nc, nf;
  orc_cartesian_counts(D, n, &nc, &nf);
  double h[3];
  for (int a = 0; a < D; a++) h[a] = (domain[2 * a + 1] - domain[2 * a]) / n[a];
  for (int64_t c = 0; c < nc; c++) {
    int64_t t = c;
    for (int a = 0; a < D; a++) {
      int64_t i = t % n[a];
      t /= n[a];
      x0[c * D + a] = domain[2 * a] + i * h[a];
    }
    for (int r = 0; r < D * D; r++) J[c * D * D + r] = 0.0;
    for (int a = 0; a < D; a++) J[c * D * D + a + D * a] = h[a];
  }
  return 0;
}

/* ------------------------------------------------------------------------ */
/* Manufactured solutions                                                    */
/* ------------------------------------------------------------------------ */
/* Darcy: returns potential, flux q (with kappa) and sources f (D), g. */
static void darcy_mms(int mms, int D, const double* x, double kappa, double* u, double* q,
                      double* f, double* g) {
  for (int a = 0; a < D; a++) q[a] = f[a] = 0.0;
  *u = 0.0;
  *g = 0.0;
  if (mms == ORC_MMS_KAT) {
    /* flux (1+x_a), potential p = -3.14: f = flux + grad p, g = div flux */
    for (int a = 0; a < D; a++) {
      q[a] = 1.0 + x[a];
      f[a] = q[a];
    }
    *u = -3.14;
    *g = (double)D;
  } else if (mms == ORC_MMS_LINEAR) {
    double s = 0.0;
    for (int a = 0; a < D; a++) s += x[a];
    *u = s;
    for (int a = 0; a < D; a++) q[a] = -kappa;
    /* f = q/kappa + grad u = 0, g = div q = 0 */
  } else if (mms == ORC_MMS_SINPROD) {
    double sv[3], cv[3];
    for (int a = 0; a < D; a++) {
      sv[a] = sin(M_PI * x[a]);
      cv[a] = cos(M_PI * x[a]);
    }
    double p = 1.0;
    for (int a = 0; a < D; a++) p *= sv[a];
    *u = p;
    for (int a = 0; a < D; a++) {
      double gr = M_PI * cv[a];
      for (int b = 0; b < D; b++)
        if (b != a) gr *= sv[b];
      q[a] = -kappa * gr;
    }
    *g = kappa * D * M_PI * M_PI * p;
  }
}

/* Elasticity: displacement u (D), stress sigma (D x D), body load f = div
 * sigma (D).  mms 1: linear displacement (exact in the spaces), mms 2:
 * smooth trigonometric displacement. */
static void elas_mms(int mms, int D, const double* x, double E, double nu, double* u,
                     double* sig, double* f) {
  double gu[9] = {0}; /* gu[i + D*j] = d u_i / d x_j */
  double hu[27] = {0}; /* second derivatives hu[i + D*(j + D*k)] */
  for (int a = 0; a < D; a++) u[a] = f[a] = 0.0;
  if (mms == 1) {
    /* u_i = sum_j (1 + i + 2j)/(D+1) x_j  (non-symmetric gradient) */
    for (int i = 0; i < D; i++)
      for (int j = 0; j < D; j++) {
        double c = (1.0 + i + 2.0 * j) / (D + 1.0);
        u[i] += c * x[j];
        gu[i + D * j] = c;
      }
  } else if (mms == 2) {
    /* u_i = sin(pi (x_i + 0.5 x_{i+1})) */
    for (int i = 0; i < D; i++) {
      int i1 = (i + 1) % D;
      double arg = M_PI * (x[i] + 0.5 * x[i1]);
      u[i] = sin(arg);
      double c = M_PI * cos(arg), s = -M_PI * M_PI * sin(arg);
      gu[i + D * i] += c;
      gu[i + D * i1] += 0.5 * c;
      double dd[3] = {0, 0, 0};
      dd[i] += 1.0;
      dd[i1] += 0.5;
      for (int j = 0; j < D; j++)
        for (int k = 0; k < D; k++) hu[i + D * (j + D * k)] = s * dd[j] * dd[k];
    }
  }
  /* sigma = A^{-1}(eps) = E/(1+nu) (eps + nu/(1-(D-1)nu) tr(eps) I)  -- inverse of
   * A(s) = (1+nu)/E s - nu/E tr(s) I in D dimensions */
  double lam = E * nu / ((1.0 + nu) * (1.0 - (D - 1) * nu));
  double mu2 = E / (1.0 + nu);
  double tr = 0.0;
  for (int a = 0; a < D; a++) tr += gu[a + D * a];
  for (int i = 0; i < D; i++)
    for (int j = 0; j < D; j++) {
      double eps = 0.5 * (gu[i + D * j] + gu[j + D * i]);
      sig[i + D * j] = mu2 * eps + (i == j ? lam * tr : 0.0);
    }
  /* f_j = sum_i d sigma_ij / dx_i */
  for (int j = 0; j < D; j++) {
    double s = 0.0;
    for (int i = 0; i < D; i++) {
      /* d/dx_i of sigma_ij */
      double deps = 0.5 * (hu[i + D * (j + D * i)] + hu[j + D * (i + D * i)]);
      double dtr = 0.0;
      for (int a = 0; a < D; a++) dtr += hu[a + D * (a + D * i)];
      s += mu2 * deps + (i == j ? lam * dtr : 0.0);
    }
    f[j] = s;
  }
}

/* symmetric tensor component c (upper triangle, row-major) -> (i,j) */
static void sym_ij(int D, int c, int* i, int* j) {
  int t = 0;
  *i = 0;
  *j = 0;
  for (int a = 0; a < D; a++)
    for (int b = a; b < D; b++) {
      if (t == c) {
        *i = a;
        *j = b;
        return;
      }
      t++;
    }
}

/* ------------------------------------------------------------------------ */
/* Cell systems                                                              */
/* ------------------------------------------------------------------------ */
typedef struct {
  int D, nF, ni, nb, n, na, nbf, ntf; /* na: flux/stress dofs; nbf: pot/disp dofs */
  int nphi, npsi, nmu;                /* scalar basis sizes */
  int nqc, nqf;
  double qc[MAXQ][3], wc[MAXQ];
  double qf[64][3], wf[64];
  /* tables at cell quadrature points */
  double *phi_c, *dphi_c, *psi_c, *dpsi_c; /* [q][i], [q][i][3] */
  /* tables at facet quadrature points, per facet */
  double *phi_f, *psi_f, *mu_f; /* [f][q][i] */
} ctx_t;

static int ctx_init(const orc_problem* p, ctx_t* c, space_t* sphi, space_t* spsi, space_t* smu) {
  int D = p->D;
  int ps = p->space == ORC_SPACE_P;
  memset(c, 0, sizeof(*c));
  c->D = D;
  c->nF = 2 * D;
  if (space_init(sphi, D, p->k_flux, ps)) return -1;
  if (space_init(spsi, D, p->k_pot, ps)) return -1;
  if (space_init(smu, D - 1, p->k_trace, ps)) return -1;
  c->nphi = sphi->n;
  c->npsi = spsi->n;
  c->nmu = smu->n;
  orc_sizes(p, &c->na, &c->nbf, &c->ntf, &c->ni, &c->nb);
  c->n = c->ni + c->nb;
  int m = npts1(p);
  double x1[32], w1[32];
  gauss01(m, x1, w1);
  c->nqc = 1;
  for (int a = 0; a < D; a++) c->nqc *= m;
  c->nqf = c->nqc / m;
  for (int q = 0; q < c->nqc; q++) {
    int t = q;
    c->wc[q] = 1.0;
    for (int a = 0; a < 3; a++) c->qc[q][a] = 0.0;
    for (int a = 0; a < D; a++) {
      c->qc[q][a] = x1[t % m];
      c->wc[q] *= w1[t % m];
      t /= m;
    }
  }
  for (int q = 0; q < c->nqf; q++) {
    int t = q;
    c->wf[q] = 1.0;
    for (int a = 0; a < 3; a++) c->qf[q][a] = 0.0;
    for (int a = 0; a < D - 1; a++) {
      c->qf[q][a] = x1[t % m];
      c->wf[q] *= w1[t % m];
      t /= m;
    }
  }
  c->phi_c = (double*)malloc(sizeof(double) * c->nqc * c->nphi);
  c->dphi_c = (double*)malloc(sizeof(double) * c->nqc * c->nphi * 3);
  c->psi_c = (double*)malloc(sizeof(double) * c->nqc * c->npsi);
  c->dpsi_c = (double*)malloc(sizeof(double) * c->nqc * c->npsi * 3);
  for (int q = 0; q < c->nqc; q++) {
    space_eval(sphi, c->qc[q], c->phi_c + q * c->nphi, c->dphi_c + q * c->nphi * 3);
    space_eval(spsi, c->qc[q], c->psi_c + q * c->npsi, c->dpsi_c + q * c->npsi * 3);
  }
  c->phi_f = (double*)malloc(sizeof(double) * c->nF * c->nqf * c->nphi);
  c->psi_f = (double*)malloc(sizeof(double) * c->nF * c->nqf * c->npsi);
  c->mu_f = (double*)malloc(sizeof(double) * c->nF * c->nqf * c->nmu);
  for (int f = 0; f < c->nF; f++)
    for (int q = 0; q < c->nqf; q++) {
      double xi[3] = {0, 0, 0};
      facet_to_cell(D, f, c->qf[q], xi);
      space_eval(sphi, xi, c->phi_f + (f * c->nqf + q) * c->nphi, NULL);
      space_eval(spsi, xi, c->psi_f + (f * c->nqf + q) * c->npsi, NULL);
      space_eval(smu, c->qf[q], c->mu_f + (f * c->nqf + q) * c->nmu, NULL);
    }
  return 0;
}

static void ctx_free(ctx_t* c) {
  free(c->phi_c);
  free(c->dphi_c);
  free(c->psi_c);
  free(c->dpsi_c);
  free(c->phi_f);
  free(c->psi_f);
  free(c->mu_f);
}

/* physical facet data: outward unit normal n and area factor (ds = fac * dshat) */
static void facet_geom(int D, int f, const double* Ji, double detJ, double* n, double* fac) {
  double nref[3];
  facet_ref_normal(D, f, nref);
  double v[3] = {0, 0, 0};
  /* J^{-T} nref : v_a = sum_b Ji[b + D a] nref_b */
  for (int a = 0; a < D; a++)
    for (int b = 0; b < D; b++) v[a] += Ji[b + D * a] * nref[b];
  double nv = 0.0;
  for (int a = 0; a < D; a++) nv += v[a] * v[a];
  nv = sqrt(nv);
  for (int a = 0; a < D; a++) n[a] = v[a] / nv;
  *fac = fabs(detJ) * nv;
}

#define AIDX(i, j) ((size_t)(i) + (size_t)(j) * (size_t)n)

/* Darcy HDG cell system, test/DarcyHDGTests.jl:125-135 (kappa^{-1} on the
 * mass term; the reference has kappa = 1). */
static void darcy_cell(const orc_problem* p, const ctx_t* c, const double* x0, const double* J,
                       double kappa, const int8_t* sign, const double* src, double* A, double* b,
                       double* skel) {
  int D = c->D, n = c->n, np = c->nphi, nq = c->npsi, nm = c->nmu;
  int off_u = c->na, off_l = c->ni;
  double tau = p->tau;
  double Ji[9], detJ = det3(D, J);
  inv3(D, J, Ji);
  double adet = fabs(detJ);
  memset(A, 0, sizeof(double) * n * n);
  memset(b, 0, sizeof(double) * n);
  memset(skel, 0, sizeof(double) * n * n);
  /* ---- bulk: vh.uh - (div vh) ph - grad(qh).uh, and l = vh.f + qh g ---- */
  for (int q = 0; q < c->nqc; q++) {
    const double* phi = c->phi_c + q * np;
    const double* dphi = c->dphi_c + q * np * 3;
    const double* psi = c->psi_c + q * nq;
    const double* dpsi = c->dpsi_c + q * nq * 3;
    double dV = c->wc[q] * adet;
    /* physical gradients: d/dx_a = sum_e Ji[e + D a] d/dxi_e */
    double gphi[MAXN][3], gpsi[MAXN][3];
    for (int i = 0; i < np; i++)
      for (int a = 0; a < D; a++) {
        double s = 0.0;
        for (int e = 0; e < D; e++) s += Ji[e + D * a] * dphi[3 * i + e];
        gphi[i][a] = s;
      }
    for (int i = 0; i < nq; i++)
      for (int a = 0; a < D; a++) {
        double s = 0.0;
        for (int e = 0; e < D; e++) s += Ji[e + D * a] * dpsi[3 * i + e];
        gpsi[i][a] = s;
      }
    /* q-q mass */
    for (int a = 0; a < D; a++)
      for (int j = 0; j < np; j++)
        for (int i = 0; i < np; i++)
          A[AIDX(a * np + i, a * np + j)] += (phi[i] * phi[j] / kappa) * dV;
    /* q-u: -(div v) p */
    for (int j = 0; j < nq; j++)
      for (int a = 0; a < D; a++)
        for (int i = 0; i < np; i++) A[AIDX(a * np + i, off_u + j)] += (-gphi[i][a] * psi[j]) * dV;
    /* u-q: -grad(w).q */
    for (int a = 0; a < D; a++)
      for (int j = 0; j < np; j++)
        for (int i = 0; i < nq; i++) A[AIDX(off_u + i, a * np + j)] += (-gpsi[i][a] * phi[j]) * dV;
    /* rhs */
    double xq[3] = {0, 0, 0};
    for (int a = 0; a < D; a++) {
      xq[a] = x0[a];
      for (int e = 0; e < D; e++) xq[a] += J[a + D * e] * c->qc[q][e];
    }
    double f[3], g;
    if (src) {
      for (int a = 0; a < D; a++) f[a] = src[q * (D + 1) + a];
      g = src[q * (D + 1) + D];
    } else {
      double uu, qq[3];
      darcy_mms(p->mms, D, xq, kappa, &uu, qq, f, &g);
    }
    for (int a = 0; a < D; a++)
      for (int i = 0; i < np; i++) b[a * np + i] += (phi[i] * f[a]) * dV;
    for (int i = 0; i < nq; i++) b[off_u + i] += (psi[i] * g) * dV;
  }
  /* ---- skeleton terms, summed facet by facet ---- */
  for (int f = 0; f < c->nF; f++) {
    double nrm[3], fac;
    facet_geom(D, f, Ji, detJ, nrm, &fac);
    double s = (double)sign[f];
    int lo = off_l + f * nm;
    for (int q = 0; q < c->nqf; q++) {
      const double* phi = c->phi_f + (f * c->nqf + q) * np;
      const double* psi = c->psi_f + (f * c->nqf + q) * nq;
      const double* mu = c->mu_f + (f * c->nqf + q) * nm;
      double dS = c->wf[q] * fac;
      /* (vh.n) lh */
      for (int m = 0; m < nm; m++)
        for (int a = 0; a < D; a++)
          for (int i = 0; i < np; i++) skel[AIDX(a * np + i, lo + m)] += (phi[i] * nrm[a] * mu[m]) * dS;
      /* qh (uh.n) */
      for (int a = 0; a < D; a++)
        for (int j = 0; j < np; j++)
          for (int i = 0; i < nq; i++) skel[AIDX(off_u + i, a * np + j)] += (psi[i] * phi[j] * nrm[a]) * dS;
      /* tau qh ph (n.no) */
      for (int j = 0; j < nq; j++)
        for (int i = 0; i < nq; i++) skel[AIDX(off_u + i, off_u + j)] += (tau * psi[i] * psi[j] * s) * dS;
      /* - tau qh lh (n.no) */
      for (int m = 0; m < nm; m++)
        for (int i = 0; i < nq; i++) skel[AIDX(off_u + i, lo + m)] += (-tau * psi[i] * mu[m] * s) * dS;
      /* mh (uh.n) */
      for (int a = 0; a < D; a++)
        for (int j = 0; j < np; j++)
          for (int m = 0; m < nm; m++) skel[AIDX(lo + m, a * np + j)] += (mu[m] * phi[j] * nrm[a]) * dS;
      /* tau mh ph (n.no) */
      for (int j = 0; j < nq; j++)
        for (int m = 0; m < nm; m++) skel[AIDX(lo + m, off_u + j)] += (tau * mu[m] * psi[j] * s) * dS;
      /* - tau mh lh (n.no) */
      for (int m2 = 0; m2 < nm; m2++)
        for (int m = 0; m < nm; m++) skel[AIDX(lo + m, lo + m2)] += (-tau * mu[m] * mu[m2] * s) * dS;
    }
  }
  for (size_t t = 0; t < (size_t)n * n; t++) A[t] += skel[t];
}

/* Elasticity HDG cell system, test/LinearElasticityHDGTests.jl:140-150 with
 * P = identity (no bulk-to-skeleton projection) and A(s) from :15-17. */
static void elas_cell(const orc_problem* p, const ctx_t* c, const double* x0, const double* J,
                      double nu, const double* src, double* A, double* b, double* skel) {
  int D = c->D, n = c->n, np = c->nphi, nq = c->npsi, nm = c->nmu;
  int nsym = D * (D + 1) / 2;
  int off_u = c->na, off_l = c->ni;
  double tau = p->tau, E = p->E;
  double Ji[9], detJ = det3(D, J);
  inv3(D, J, Ji);
  double adet = fabs(detJ);
  memset(A, 0, sizeof(double) * n * n);
  memset(b, 0, sizeof(double) * n);
  memset(skel, 0, sizeof(double) * n * n);
  double c1 = (1.0 + nu) / E, c2 = nu / E;
  /* basis tensors of the symmetric components */
  double Et[6][9];
  for (int cc = 0; cc < nsym; cc++) {
    int i, j;
    sym_ij(D, cc, &i, &j);
    for (int t = 0; t < 9; t++) Et[cc][t] = 0.0;
    Et[cc][i + D * j] = 1.0;
    Et[cc][j + D * i] = 1.0;
  }
  for (int q = 0; q < c->nqc; q++) {
    const double* phi = c->phi_c + q * np;
    const double* dphi = c->dphi_c + q * np * 3;
    const double* psi = c->psi_c + q * nq;
    double dV = c->wc[q] * adet;
    double gphi[MAXN][3];
    for (int i = 0; i < np; i++)
      for (int a = 0; a < D; a++) {
        double s = 0.0;
        for (int e = 0; e < D; e++) s += Ji[e + D * a] * dphi[3 * i + e];
        gphi[i][a] = s;
      }
    /* v : A(sigma) */
    for (int cb = 0; cb < nsym; cb++)
      for (int ca = 0; ca < nsym; ca++) {
        double trb = 0.0, dd = 0.0, tra = 0.0;
        for (int a = 0; a < D; a++) {
          trb += Et[cb][a + D * a];
          tra += Et[ca][a + D * a];
        }
        for (int t = 0; t < D * D; t++) dd += Et[ca][t] * Et[cb][t];
        double kab = c1 * dd - c2 * trb * tra; /* v_a : A(sigma_b) */
        if (kab == 0.0) continue;
        for (int j = 0; j < np; j++)
          for (int i = 0; i < np; i++) A[AIDX(ca * np + i, cb * np + j)] += (kab * phi[i] * phi[j]) * dV;
      }
    /* (div v).u and -omega.(div sigma); (div v)_j = sum_i d_i v_ij */
    for (int ca = 0; ca < nsym; ca++)
      for (int i = 0; i < np; i++) {
        double dv[3] = {0, 0, 0};
        for (int jj = 0; jj < D; jj++)
          for (int ii = 0; ii < D; ii++) dv[jj] += Et[ca][ii + D * jj] * gphi[i][ii];
        for (int comp = 0; comp < D; comp++) {
          if (dv[comp] == 0.0) continue;
          for (int j = 0; j < nq; j++) {
            A[AIDX(ca * np + i, off_u + comp * nq + j)] += (dv[comp] * psi[j]) * dV;
            A[AIDX(off_u + comp * nq + j, ca * np + i)] += (-psi[j] * dv[comp]) * dV;
          }
        }
      }
    double xq[3] = {0, 0, 0};
    for (int a = 0; a < D; a++) {
      xq[a] = x0[a];
      for (int e = 0; e < D; e++) xq[a] += J[a + D * e] * c->qc[q][e];
    }
    double f[3];
    if (src) {
      for (int a = 0; a < D; a++) f[a] = src[q * D + a];
    } else {
      double uu[3], sg[9];
      elas_mms(p->mms, D, xq, E, nu, uu, sg, f);
    }
    for (int comp = 0; comp < D; comp++)
      for (int j = 0; j < nq; j++) b[off_u + comp * nq + j] += (-psi[j] * f[comp]) * dV;
  }
  for (int fc = 0; fc < c->nF; fc++) {
    double nrm[3], fac;
    facet_geom(D, fc, Ji, detJ, nrm, &fac);
    int lo = off_l + fc * (D * nm);
    for (int q = 0; q < c->nqf; q++) {
      const double* phi = c->phi_f + (fc * c->nqf + q) * np;
      const double* psi = c->psi_f + (fc * c->nqf + q) * nq;
      const double* mu = c->mu_f + (fc * c->nqf + q) * nm;
      double dS = c->wf[q] * fac;
      for (int ca = 0; ca < nsym; ca++) {
        double vn[3] = {0, 0, 0}; /* (E_ca . n)_comp */
        for (int ii = 0; ii < D; ii++)
          for (int jj = 0; jj < D; jj++) vn[ii] += Et[ca][ii + D * jj] * nrm[jj];
        for (int comp = 0; comp < D; comp++) {
          if (vn[comp] == 0.0) continue;
          for (int m = 0; m < nm; m++)
            for (int i = 0; i < np; i++) {
              /* - (v.n).uhat */
              skel[AIDX(ca * np + i, lo + comp * nm + m)] += (-phi[i] * vn[comp] * mu[m]) * dS;
              /* mu.(sigma.n) */
              skel[AIDX(lo + comp * nm + m, ca * np + i)] += (mu[m] * phi[i] * vn[comp]) * dS;
            }
        }
      }
      for (int comp = 0; comp < D; comp++) {
        /* tau omega.u (with P_m: added after the facet's quadrature loop) */
        if (!p->projection)
          for (int j = 0; j < nq; j++)
            for (int i = 0; i < nq; i++)
              skel[AIDX(off_u + comp * nq + i, off_u + comp * nq + j)] += (tau * psi[i] * psi[j]) * dS;
        /* - tau omega.uhat */
        for (int m = 0; m < nm; m++)
          for (int i = 0; i < nq; i++)
            skel[AIDX(off_u + comp * nq + i, lo + comp * nm + m)] += (-tau * psi[i] * mu[m]) * dS;
        /* - tau mu.u */
        for (int j = 0; j < nq; j++)
          for (int m = 0; m < nm; m++)
            skel[AIDX(lo + comp * nm + m, off_u + comp * nq + j)] += (-tau * mu[m] * psi[j]) * dS;
        /* tau mu.uhat */
        for (int m2 = 0; m2 < nm; m2++)
          for (int m = 0; m < nm; m++)
            skel[AIDX(lo + comp * nm + m, lo + comp * nm + m2)] += (tau * mu[m] * mu[m2]) * dS;
      }
    }
    if (p->projection) {
      /* P_m (test/P_m.jl:4-23, src/GridapAPIExtensions.jl:453-500): the facet
       * trace of u is replaced by its L2 projection onto the trace space,
       * dofs = A \ B with A = (mu, mu')_F and B = (mu, u)_F, so
       * tau (omega, P_m u)_F = tau B^T A^-1 B per displacement component. */
      double* Bf = (double*)calloc((size_t)nm * nq, sizeof(double));
      double* Af = (double*)calloc((size_t)nm * nm, sizeof(double));
      double* X = (double*)malloc(sizeof(double) * (size_t)nm * nq);
      for (int q = 0; q < c->nqf; q++) {
        const double* psi = c->psi_f + (fc * c->nqf + q) * nq;
        const double* mu = c->mu_f + (fc * c->nqf + q) * nm;
        double dS = c->wf[q] * fac;
        for (int j = 0; j < nq; j++)
          for (int m = 0; m < nm; m++) Bf[m + (size_t)j * nm] += mu[m] * psi[j] * dS;
        for (int m2 = 0; m2 < nm; m2++)
          for (int m = 0; m < nm; m++) Af[m + (size_t)m2 * nm] += mu[m] * mu[m2] * dS;
      }
      memcpy(X, Bf, sizeof(double) * (size_t)nm * nq);
      orc_dense_solve(nm, nq, Af, X); /* X = A^-1 B */
      for (int comp = 0; comp < D; comp++)
        for (int j = 0; j < nq; j++)
          for (int i = 0; i < nq; i++) {
            double v = 0.0;
            for (int m = 0; m < nm; m++) v += Bf[m + (size_t)i * nm] * X[m + (size_t)j * nm];
            skel[AIDX(off_u + comp * nq + i, off_u + comp * nq + j)] += tau * v;
          }
      free(Bf);
      free(Af);
      free(X);
    }
  }
  for (size_t t = 0; t < (size_t)n * n; t++) A[t] += skel[t];
}

int orc_cell_systems(const orc_problem* p, int64_t ncells, const double* x0, const double* J,
                     const double* coef, const int8_t* sign, const double* src, double* A,
                     double* b) {
  space_t sphi, spsi, smu;
  ctx_t c;
  if (ctx_init(p, &c, &sphi, &spsi, &smu)) return -1;
  int D = p->D, n = c.n, nF = c.nF;
  int srcw = p->phys == ORC_PHYS_DARCY ? D + 1 : D;
  int rc = 0;
#pragma omp parallel
  {
    double* skel = (double*)malloc(sizeof(double) * n * n);
#pragma omp for schedule(static)
    for (int64_t cell = 0; cell < ncells; cell++) {
      const double* s = src ? src + (size_t)cell * c.nqc * srcw : NULL;
      if (p->phys == ORC_PHYS_DARCY)
        darcy_cell(p, &c, x0 + cell * D, J + cell * D * D, coef[cell], sign + cell * nF, s,
                   A + (size_t)cell * n * n, b + (size_t)cell * n, skel);
      else
        elas_cell(p, &c, x0 + cell * D, J + cell * D * D, coef[cell], s, A + (size_t)cell * n * n,
                  b + (size_t)cell * n, skel);
    }
    free(skel);
  }
  ctx_free(&c);
  space_free(&sphi);
  space_free(&spsi);
  space_free(&smu);
  return rc;
}

int orc_trace_nodes(const orc_problem* p, double* nodes) {
  space_t s;
  if (space_init(&s, p->D - 1, p->k_trace, p->space == ORC_SPACE_P)) return -1;
  for (int i = 0; i < s.n; i++)
    for (int a = 0; a < p->D - 1; a++) nodes[i * (p->D - 1) + a] = s.nodes[i][a];
  space_free(&s);
  return 0;
}

/* ------------------------------------------------------------------------ */
/* LAPACK-semantics kernels (column-major), written from the published       */
/* algorithms: dgetf2 (Golub & Van Loan Alg. 3.4.1), dgetrs('N'), netlib      */
/* dgemm/dgemv loop order.                                                   */
/* ------------------------------------------------------------------------ */
static int lu_getf2(int n, double* A, int lda, int* ipiv) {
  int info = 0;
  for (int j = 0; j < n; j++) {
    int p = j;
    double amax = fabs(A[j + (size_t)j * lda]);
    for (int i = j + 1; i < n; i++) {
      double v = fabs(A[i + (size_t)j * lda]);
      if (v > amax) {
        amax = v;
        p = i;
      }
    }
    ipiv[j] = p;
    if (A[p + (size_t)j * lda] != 0.0) {
      if (p != j)
        for (int c = 0; c < n; c++) {
          double t = A[j + (size_t)c * lda];
          A[j + (size_t)c * lda] = A[p + (size_t)c * lda];
          A[p + (size_t)c * lda] = t;
        }
      double r = 1.0 / A[j + (size_t)j * lda];
      for (int i = j + 1; i < n; i++) A[i + (size_t)j * lda] *= r;
    } else if (!info) {
      info = j + 1;
    }
    for (int c = j + 1; c < n; c++) {
      double u = A[j + (size_t)c * lda];
      if (u != 0.0)
        for (int i = j + 1; i < n; i++) A[i + (size_t)c * lda] -= A[i + (size_t)j * lda] * u;
    }
  }
  return info;
}

static void lu_getrs(int n, int nrhs, const double* A, int lda, const int* ipiv, double* B,
                     int ldb) {
  for (int r = 0; r < nrhs; r++) {
    double* x = B + (size_t)r * ldb;
    for (int j = 0; j < n; j++)
      if (ipiv[j] != j) {
        double t = x[j];
        x[j] = x[ipiv[j]];
        x[ipiv[j]] = t;
      }
    for (int j = 0; j < n; j++) {
      double xj = x[j];
      if (xj != 0.0)
        for (int i = j + 1; i < n; i++) x[i] -= xj * A[i + (size_t)j * lda];
    }
    for (int j = n - 1; j >= 0; j--) {
      if (x[j] != 0.0) {
        x[j] /= A[j + (size_t)j * lda];
        double xj = x[j];
        for (int i = 0; i < j; i++) x[i] -= xj * A[i + (size_t)j * lda];
      }
    }
  }
}

/* C(m x n) += alpha A(m x k) B(k x n), beta = 1 */
static void mm_acc(int m, int n, int k, double alpha, const double* A, int lda, const double* B,
                   int ldb, double* C, int ldc) {
  for (int j = 0; j < n; j++)
    for (int l = 0; l < k; l++) {
      double t = alpha * B[l + (size_t)j * ldb];
      if (t != 0.0)
        for (int i = 0; i < m; i++) C[i + (size_t)j * ldc] += t * A[i + (size_t)l * lda];
    }
}

int orc_condense(int ni, int nb, int64_t ncells, const double* A, const double* b, double* S,
                 double* g, int32_t* info) {
  int n = ni + nb;
  int bad = 0;
#pragma omp parallel reduction(+ : bad)
  {
    double* A11 = (double*)malloc(sizeof(double) * ni * ni);
    double* A12 = (double*)malloc(sizeof(double) * ni * nb);
    double* A21 = (double*)malloc(sizeof(double) * nb * ni);
    double* b1 = (double*)malloc(sizeof(double) * ni);
    int* ipiv = (int*)malloc(sizeof(int) * ni);
#pragma omp for schedule(static)
    for (int64_t c = 0; c < ncells; c++) {
      const double* Ac = A + (size_t)c * n * n;
      const double* bc = b + (size_t)c * n;
      double* Sc = S + (size_t)c * nb * nb;
      double* gc = g + (size_t)c * nb;
      /* densify the four blocks (src/StaticCondensationMap.jl:170-175) */
      for (int j = 0; j < ni; j++)
        for (int i = 0; i < ni; i++) A11[i + (size_t)j * ni] = Ac[i + (size_t)j * n];
      for (int j = 0; j < nb; j++)
        for (int i = 0; i < ni; i++) A12[i + (size_t)j * ni] = Ac[i + (size_t)(ni + j) * n];
      for (int j = 0; j < ni; j++)
        for (int i = 0; i < nb; i++) A21[i + (size_t)j * nb] = Ac[ni + i + (size_t)j * n];
      for (int j = 0; j < nb; j++)
        for (int i = 0; i < nb; i++) Sc[i + (size_t)j * nb] = Ac[ni + i + (size_t)(ni + j) * n];
      for (int i = 0; i < ni; i++) b1[i] = bc[i];
      for (int i = 0; i < nb; i++) gc[i] = bc[ni + i];
      int inf = lu_getf2(ni, A11, ni, ipiv);      /* :179 */
      info[c] = inf;
      if (inf) {
        bad++;
        continue;
      }
      lu_getrs(ni, nb, A11, ni, ipiv, A12, ni);     /* :183 */
      mm_acc(nb, nb, ni, -1.0, A21, nb, A12, ni, Sc, nb); /* :186 */
      lu_getrs(ni, 1, A11, ni, ipiv, b1, ni);       /* :189 */
      mm_acc(nb, 1, ni, -1.0, A21, nb, b1, ni, gc, nb); /* :192 */
    }
    free(A11);
    free(A12);
    free(A21);
    free(b1);
    free(ipiv);
  }
  return bad;
}

int orc_recover(int ni, int nb, int64_t ncells, const double* A, const double* b,
                const double* xb, double* xi, int32_t* info) {
  int n = ni + nb;
  int bad = 0;
#pragma omp parallel reduction(+ : bad)
  {
    double* A11 = (double*)malloc(sizeof(double) * ni * ni);
    int* ipiv = (int*)malloc(sizeof(int) * ni);
#pragma omp for schedule(static)
    for (int64_t c = 0; c < ncells; c++) {
      const double* Ac = A + (size_t)c * n * n;
      const double* bc = b + (size_t)c * n;
      double* x = xi + (size_t)c * ni;
      for (int i = 0; i < ni; i++) x[i] = bc[i];
      /* b1 -= A12 * x  (src/BackwardStaticCondensationMap.jl:91) */
      mm_acc(ni, 1, nb, -1.0, Ac + (size_t)ni * n, n, xb + (size_t)c * nb, nb, x, ni);
      for (int j = 0; j < ni; j++)
        for (int i = 0; i < ni; i++) A11[i + (size_t)j * ni] = Ac[i + (size_t)j * n];
      int inf = lu_getf2(ni, A11, ni, ipiv); /* :95 */
      info[c] = inf;
      if (inf) {
        bad++;
        continue;
      }
      lu_getrs(ni, 1, A11, ni, ipiv, x, ni); /* :99 */
    }
    free(A11);
    free(ipiv);
  }
  return bad;
}

/* ------------------------------------------------------------------------ */
/* DOF ids, assembly, recovery gather                                        */
/* ------------------------------------------------------------------------ */
int orc_trace_dofs(int D, int n_tf, int64_t ncells, int64_t nfacets, const int32_t* cell_facets,
                   const int8_t* facet_dirichlet, int32_t* cell_dofs, int64_t* nfree,
                   int64_t* ndir) {
  int nF = 2 * D;
  int64_t* first = (int64_t*)malloc(sizeof(int64_t) * nfacets);
  int64_t fr = 0, dr = 0;
  /* L2 facet space: facet by facet, dofs consecutive; Dirichlet facets
   * numbered -1,-2,... in facet order */
  for (int64_t f = 0; f < nfacets; f++) {
    if (facet_dirichlet[f]) {
      first[f] = -(dr + 1);
      dr += n_tf;
    } else {
      first[f] = fr;
      fr += n_tf;
    }
  }
  for (int64_t c = 0; c < ncells; c++)
    for (int lf = 0; lf < nF; lf++) {
      int64_t f = cell_facets[c * nF + lf];
      for (int m = 0; m < n_tf; m++) {
        int64_t id = first[f] >= 0 ? first[f] + m : first[f] - m;
        cell_dofs[(c * nF + lf) * n_tf + m] = (int32_t)id;
      }
    }
  free(first);
  *nfree = fr;
  *ndir = dr;
  return 0;
}

static int cmp_i32(const void* a, const void* b) {
  int32_t x = *(const int32_t*)a, y = *(const int32_t*)b;
  return (x > y) - (x < y);
}

/* column lists: for each free column dof, the set of free rows coupled to it */
static int64_t build_pattern(int nb, int64_t ncells, int64_t nfree, const int32_t* cell_dofs,
                             int64_t* colptr, int32_t* rowval) {
  /* count with duplicates per column, then sort + unique */
  int64_t* cnt = (int64_t*)calloc(nfree + 1, sizeof(int64_t));
  for (int64_t c = 0; c < ncells; c++) {
    const int32_t* d = cell_dofs + c * nb;
    int nfr = 0;
    for (int i = 0; i < nb; i++) nfr += d[i] >= 0;
    for (int j = 0; j < nb; j++)
      if (d[j] >= 0) cnt[d[j] + 1] += nfr;
  }
  for (int64_t j = 0; j < nfree; j++) cnt[j + 1] += cnt[j];
  int32_t* rows = (int32_t*)malloc(sizeof(int32_t) * (cnt[nfree] > 0 ? cnt[nfree] : 1));
  int64_t* pos = (int64_t*)malloc(sizeof(int64_t) * (nfree > 0 ? nfree : 1));
  for (int64_t j = 0; j < nfree; j++) pos[j] = cnt[j];
  for (int64_t c = 0; c < ncells; c++) {
    const int32_t* d = cell_dofs + c * nb;
    for (int j = 0; j < nb; j++)
      if (d[j] >= 0)
        for (int i = 0; i < nb; i++)
          if (d[i] >= 0) rows[pos[d[j]]++] = d[i];
  }
  int64_t nnz = 0;
  for (int64_t j = 0; j < nfree; j++) {
    int64_t a = cnt[j], e = cnt[j + 1];
    qsort(rows + a, e - a, sizeof(int32_t), cmp_i32);
    if (colptr) colptr[j] = nnz;
    for (int64_t t = a; t < e; t++) {
      if (t > a && rows[t] == rows[t - 1]) continue;
      if (rowval) rowval[nnz] = rows[t];
      nnz++;
    }
  }
  if (colptr) colptr[nfree] = nnz;
  free(rows);
  free(pos);
  free(cnt);
  return nnz;
}

int64_t orc_assemble_nnz(int nb, int64_t ncells, int64_t nfree, const int32_t* cell_dofs) {
  return build_pattern(nb, ncells, nfree, cell_dofs, NULL, NULL);
}

static int64_t find_row(const int64_t* colptr, const int32_t* rowval, int64_t col, int32_t row) {
  int64_t lo = colptr[col], hi = colptr[col + 1] - 1;
  while (lo <= hi) {
    int64_t mid = (lo + hi) / 2;
    if (rowval[mid] == row) return mid;
    if (rowval[mid] < row)
      lo = mid + 1;
    else
      hi = mid - 1;
  }
  return -1;
}

int orc_assemble(int nb, int64_t ncells, int64_t nfree, const int32_t* cell_dofs, const double* S,
                 const double* g, const double* lam_dir, int64_t* colptr, int32_t* rowval,
                 double* nzval, double* rhs) {
  int64_t nnz = build_pattern(nb, ncells, nfree, cell_dofs, colptr, rowval);
  for (int64_t t = 0; t < nnz; t++) nzval[t] = 0.0;
  for (int64_t i = 0; i < nfree; i++) rhs[i] = 0.0;
  double* vec = (double*)malloc(sizeof(double) * nb);
  for (int64_t c = 0; c < ncells; c++) {
    const int32_t* d = cell_dofs + c * nb;
    const double* Sc = S + (size_t)c * nb * nb;
    /* _attach_dirichlet: vec = g - S * uhd_cell (uhd: Dirichlet values,
     * zero on free dofs) */
    for (int i = 0; i < nb; i++) vec[i] = g[(size_t)c * nb + i];
    for (int j = 0; j < nb; j++)
      if (d[j] < 0) {
        double v = lam_dir[-d[j] - 1];
        if (v != 0.0)
          for (int i = 0; i < nb; i++) vec[i] -= Sc[i + (size_t)j * nb] * v;
      }
    for (int j = 0; j < nb; j++) {
      if (d[j] < 0) continue;
      for (int i = 0; i < nb; i++) {
        if (d[i] < 0) continue;
        int64_t s = find_row(colptr, rowval, d[j], d[i]);
        nzval[s] += Sc[i + (size_t)j * nb];
      }
    }
    for (int i = 0; i < nb; i++)
      if (d[i] >= 0) rhs[d[i]] += vec[i];
  }
  free(vec);
  return 0;
}

int orc_gather_lambda(int nb, int64_t ncells, const int32_t* cell_dofs, const double* lam_free,
                      const double* lam_dir, double* xb) {
  for (int64_t t = 0; t < ncells * nb; t++) {
    int32_t d = cell_dofs[t];
    xb[t] = d >= 0 ? lam_free[d] : lam_dir[-d - 1];
  }
  return 0;
}

int orc_dirichlet_values(const orc_problem* p, int64_t nfacets, const int32_t* facet_cells,
                         const int8_t* facet_lpos, const int8_t* facet_dirichlet,
                         const double* x0, const double* J, double* lam_dir) {
  int D = p->D;
  space_t s;
  if (space_init(&s, D - 1, p->k_trace, p->space == ORC_SPACE_P)) return -1;
  int64_t k = 0;
  for (int64_t f = 0; f < nfacets; f++) {
    if (!facet_dirichlet[f]) continue;
    int64_t c = facet_cells[2 * f];
    int lf = facet_lpos[2 * f];
    for (int m = 0; m < s.n; m++) {
      double xi[3] = {0, 0, 0}, x[3] = {0, 0, 0};
      facet_to_cell(D, lf, s.nodes[m], xi);
      for (int a = 0; a < D; a++) {
        x[a] = x0[c * D + a];
        for (int e = 0; e < D; e++) x[a] += J[c * D * D + a + D * e] * xi[e];
      }
      if (p->phys == ORC_PHYS_DARCY) {
        double u, q[3], ff[3], g;
        darcy_mms(p->mms, D, x, 1.0, &u, q, ff, &g);
        lam_dir[k++] = u;
      } else {
        /* vector trace, component-major within the facet */
        double u[3], sg[9], ff[3];
        elas_mms(p->mms, D, x, p->E, p->nu, u, sg, ff);
        for (int comp = 0; comp < D; comp++) lam_dir[k + comp * s.n + m] = u[comp];
      }
    }
    if (p->phys != ORC_PHYS_DARCY) k += (int64_t)D * s.n;
  }
  space_free(&s);
  return 0;
}

int orc_l2_errors(const orc_problem* p, int64_t ncells, const double* x0, const double* J,
                  const double* coef, const double* xi, double* err) {
  space_t sphi, spsi, smu;
  ctx_t c;
  if (ctx_init(p, &c, &sphi, &spsi, &smu)) return -1;
  int D = p->D;
  int nsym = D * (D + 1) / 2;
  double e0 = 0.0, e1 = 0.0;
  for (int64_t cell = 0; cell < ncells; cell++) {
    const double* x = xi + (size_t)cell * c.ni;
    double adet = fabs(det3(D, J + cell * D * D));
    for (int q = 0; q < c.nqc; q++) {
      const double* phi = c.phi_c + q * c.nphi;
      const double* psi = c.psi_c + q * c.npsi;
      double xq[3] = {0, 0, 0};
      for (int a = 0; a < D; a++) {
        xq[a] = x0[cell * D + a];
        for (int e = 0; e < D; e++) xq[a] += J[cell * D * D + a + D * e] * c.qc[q][e];
      }
      double dV = c.wc[q] * adet;
      if (p->phys == ORC_PHYS_DARCY) {
        double u, qq[3], ff[3], g;
        darcy_mms(p->mms, D, xq, coef[cell], &u, qq, ff, &g);
        for (int a = 0; a < D; a++) {
          double v = 0.0;
          for (int i = 0; i < c.nphi; i++) v += x[a * c.nphi + i] * phi[i];
          e0 += (v - qq[a]) * (v - qq[a]) * dV;
        }
        double v = 0.0;
        for (int i = 0; i < c.npsi; i++) v += x[c.na + i] * psi[i];
        e1 += (v - u) * (v - u) * dV;
      } else {
        double u[3], sg[9], ff[3];
        elas_mms(p->mms, D, xq, p->E, coef[cell], u, sg, ff);
        for (int cc = 0; cc < nsym; cc++) {
          int i, j;
          sym_ij(D, cc, &i, &j);
          double v = 0.0;
          for (int t = 0; t < c.nphi; t++) v += x[cc * c.nphi + t] * phi[t];
          double w = (i == j) ? 1.0 : 2.0; /* sigma : sigma counts off-diagonals twice */
          e0 += w * (v - sg[i + D * j]) * (v - sg[i + D * j]) * dV;
        }
        for (int comp = 0; comp < D; comp++) {
          double v = 0.0;
          for (int t = 0; t < c.npsi; t++) v += x[c.na + comp * c.npsi + t] * psi[t];
          e1 += (v - u[comp]) * (v - u[comp]) * dV;
        }
      }
    }
  }
  err[0] = sqrt(e0);
  err[1] = sqrt(e1);
  ctx_free(&c);
  space_free(&sphi);
  space_free(&spsi);
  space_free(&smu);
  return 0;
}

/* ------------------------------------------------------------------------ */
/* CPU-baseline variants                                                      */
/* ------------------------------------------------------------------------ */
#ifdef _OPENMP
#include <omp.h>
#endif

int orc_set_threads(int n) {
#ifdef _OPENMP
  int prev = omp_get_max_threads();
  if (n > 0) omp_set_num_threads(n);
  return prev;
#else
  (void)n;
  return 1;
#endif
}

/* Condensation with the factors kept for the recovery (the alternative to
 * the reference's re-factorisation at src/BackwardStaticCondensationMap.jl:95):
 * per cell getrf, W = A11^-1 A12, w = A11^-1 b1 (getrs, :183 and :189), S and
 * g as in orc_condense, then x_i = w - W x_b (one gemv) for the given trace
 * values xb.  Outputs as orc_condense and orc_recover. */
int orc_condense_recover_stored(int ni, int nb, int64_t ncells, const double* A, const double* b,
                                const double* xb, double* S, double* g, double* xi,
                                int32_t* info) {
  int n = ni + nb;
  int bad = 0;
#pragma omp parallel reduction(+ : bad)
  {
    double* A11 = (double*)malloc(sizeof(double) * ni * ni);
    double* W = (double*)malloc(sizeof(double) * ni * nb);
    double* A21 = (double*)malloc(sizeof(double) * nb * ni);
    int* ipiv = (int*)malloc(sizeof(int) * ni);
#pragma omp for schedule(static)
    for (int64_t c = 0; c < ncells; c++) {
      const double* Ac = A + (size_t)c * n * n;
      const double* bc = b + (size_t)c * n;
      double* Sc = S + (size_t)c * nb * nb;
      double* gc = g + (size_t)c * nb;
      double* x = xi + (size_t)c * ni;
      for (int j = 0; j < ni; j++)
        for (int i = 0; i < ni; i++) A11[i + (size_t)j * ni] = Ac[i + (size_t)j * n];
      for (int j = 0; j < nb; j++)
        for (int i = 0; i < ni; i++) W[i + (size_t)j * ni] = Ac[i + (size_t)(ni + j) * n];
      for (int j = 0; j < ni; j++)
        for (int i = 0; i < nb; i++) A21[i + (size_t)j * nb] = Ac[ni + i + (size_t)j * n];
      for (int j = 0; j < nb; j++)
        for (int i = 0; i < nb; i++) Sc[i + (size_t)j * nb] = Ac[ni + i + (size_t)(ni + j) * n];
      for (int i = 0; i < ni; i++) x[i] = bc[i];
      for (int i = 0; i < nb; i++) gc[i] = bc[ni + i];
      int inf = lu_getf2(ni, A11, ni, ipiv);
      info[c] = inf;
      if (inf) {
        bad++;
        continue;
      }
      lu_getrs(ni, nb, A11, ni, ipiv, W, ni);
      mm_acc(nb, nb, ni, -1.0, A21, nb, W, ni, Sc, nb);
      lu_getrs(ni, 1, A11, ni, ipiv, x, ni);
      mm_acc(nb, 1, ni, -1.0, A21, nb, x, ni, gc, nb);
      /* recovery from the kept W and w: x = w - W xb */
      mm_acc(ni, 1, nb, -1.0, W, ni, xb + (size_t)c * nb, nb, x, ni);
    }
    free(A11);
    free(W);
    free(A21);
    free(ipiv);
  }
  return bad;
}

/* ------------------------------------------------------------------------ */
/* Facet orientation (src/SkeletonArrays.jl:28-37 pindex, src/Skeleton.jl:    */
/* 587-612): each (cell, local facet) carries a code p of the map from the   */
/* cell's local facet coordinates to the facet's own coordinates.  2D: p=0   */
/* s'=s, p=1 s'=1-s.  3D: (a,b) = (t,s) if p&1 else (s,t); s' = 1-a if p&2,   */
/* t' = 1-b if p&4.  The trace dofs of a facet are numbered in its own frame, */
/* so node i of the cell's local facet is own node perm_p(i).                */
/* ------------------------------------------------------------------------ */
static void orient_coords(int D, int p, const double* y, double* yo) {
  if (D == 2) {
    yo[0] = (p & 1) ? 1.0 - y[0] : y[0];
    return;
  }
  double a = (p & 1) ? y[1] : y[0], b = (p & 1) ? y[0] : y[1];
  yo[0] = (p & 2) ? 1.0 - a : a;
  yo[1] = (p & 4) ? 1.0 - b : b;
}

int orc_facet_node_perm(const orc_problem* p, int code, int* perm) {
  int D = p->D;
  if (code < 0 || code >= (D == 2 ? 2 : 8)) return -1;
  space_t s;
  if (space_init(&s, D - 1, p->k_trace, p->space == ORC_SPACE_P)) return -1;
  int rc = 0;
  for (int i = 0; i < s.n && !rc; i++) {
    double yo[2] = {0, 0};
    orient_coords(D, code, s.nodes[i], yo);
    int found = -1;
    for (int j = 0; j < s.n; j++) {
      double d = 0.0;
      for (int a = 0; a < D - 1; a++) d += fabs(s.nodes[j][a] - yo[a]);
      if (d < 1e-12) found = j;
    }
    if (found < 0) rc = -2; /* the node set is not invariant under this map */
    perm[i] = found;
  }
  space_free(&s);
  return rc;
}

int orc_trace_dofs_oriented(const orc_problem* p, int64_t ncells, int64_t nfacets,
                            const int32_t* cell_facets, const int8_t* cell_facet_perm,
                            const int8_t* facet_dirichlet, int32_t* cell_dofs, int64_t* nfree,
                            int64_t* ndir) {
  int D = p->D, nF = 2 * D;
  int n_flux, n_pot, n_tf, ni, nb;
  orc_sizes(p, &n_flux, &n_pot, &n_tf, &ni, &nb);
  int ncomp = p->phys == ORC_PHYS_DARCY ? 1 : D;
  int nm = n_tf / ncomp;
  int* perm = (int*)malloc(sizeof(int) * 8 * nm);
  for (int code = 0; code < (D == 2 ? 2 : 8); code++)
    if (orc_facet_node_perm(p, code, perm + code * nm))
      for (int i = 0; i < nm; i++) perm[code * nm + i] = -1;
  int rc = orc_trace_dofs(D, n_tf, ncells, nfacets, cell_facets, facet_dirichlet, cell_dofs, nfree,
                          ndir);
  for (int64_t c = 0; c < ncells && !rc; c++)
    for (int lf = 0; lf < nF && !rc; lf++) {
      int code = cell_facet_perm ? cell_facet_perm[c * nF + lf] : 0;
      if (!code) continue;
      if (perm[code * nm] < 0) {
        rc = -2;
        break;
      }
      int32_t* d = cell_dofs + (c * nF + lf) * n_tf;
      int32_t own[64];
      for (int m = 0; m < n_tf; m++) own[m] = d[m]; /* own order: dof = comp*nm + node */
      for (int comp = 0; comp < ncomp; comp++)
        for (int i = 0; i < nm; i++) d[comp * nm + i] = own[comp * nm + perm[code * nm + i]];
    }
  free(perm);
  return rc;
}

/* Dirichlet values in each facet's own node order: own node j is node
 * perm^-1(j) of the owner cell's local facet. */
int orc_dirichlet_values_oriented(const orc_problem* p, int64_t nfacets,
                                  const int32_t* facet_cells, const int8_t* facet_lpos,
                                  const int8_t* cell_facet_perm, const int8_t* facet_dirichlet,
                                  const double* x0, const double* J, double* lam_dir) {
  int D = p->D, nF = 2 * D;
  space_t s;
  if (space_init(&s, D - 1, p->k_trace, p->space == ORC_SPACE_P)) return -1;
  int nm = s.n;
  int* perm = (int*)malloc(sizeof(int) * 8 * nm);
  int* inv = (int*)malloc(sizeof(int) * nm);
  for (int code = 0; code < (D == 2 ? 2 : 8); code++)
    if (orc_facet_node_perm(p, code, perm + code * nm))
      for (int i = 0; i < nm; i++) perm[code * nm + i] = -1;
  int64_t k = 0;
  int rc = 0;
  for (int64_t f = 0; f < nfacets && !rc; f++) {
    if (!facet_dirichlet[f]) continue;
    int64_t c = facet_cells[2 * f];
    int lf = facet_lpos[2 * f];
    int code = cell_facet_perm ? cell_facet_perm[c * nF + lf] : 0;
    if (perm[code * nm] < 0) {
      rc = -2;
      break;
    }
    for (int i = 0; i < nm; i++) inv[perm[code * nm + i]] = i;
    for (int j = 0; j < nm; j++) {
      double xi[3] = {0, 0, 0}, x[3] = {0, 0, 0};
      facet_to_cell(D, lf, s.nodes[inv[j]], xi);
      for (int a = 0; a < D; a++) {
        x[a] = x0[c * D + a];
        for (int e = 0; e < D; e++) x[a] += J[c * D * D + a + D * e] * xi[e];
      }
      if (p->phys == ORC_PHYS_DARCY) {
        double u, q[3], ff[3], g;
        darcy_mms(p->mms, D, x, 1.0, &u, q, ff, &g);
        lam_dir[k++] = u;
      } else {
        double u[3], sg[9], ff[3];
        elas_mms(p->mms, D, x, p->E, p->nu, u, sg, ff);
        for (int comp = 0; comp < D; comp++) lam_dir[k + comp * nm + j] = u[comp];
      }
    }
    if (p->phys != ORC_PHYS_DARCY) k += (int64_t)D * nm;
  }
  free(perm);
  free(inv);
  space_free(&s);
  return rc;
}

/* Nodes of the scalar Lagrange space of order k on the reference d-cube, in
 * the oracle's node order: nodes[i*d + a]. */
int orc_space_nodes(int d, int k, int pspace, double* nodes) {
  space_t s;
  if (space_init(&s, d, k, pspace)) return -1;
  for (int i = 0; i < s.n; i++)
    for (int a = 0; a < d; a++) nodes[i * d + a] = s.nodes[i][a];
  int n = s.n;
  space_free(&s);
  return n;
}
