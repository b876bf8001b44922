// Host-side FP64 maths of the SEM library: GLL quadrature, Lagrange basis,
// P1 (bilinear/trilinear) geometry, hull test and inverse map.
//
// Semantics follow the reference: TensorQuad.cpp:50-174 and
// Hexahedra.cpp:74-130, :493-577 (GLL tables, D_ij = l'_j(xi_i));
// QuadP1.cpp:8-92 / HexP1.cpp:8-240 (hull test including points on the
// boundary, Newton inverse map from the element centre with absolute
// tolerance 1e-6 and at most 11 updates).  GLL nodes are computed here by
// Newton iteration on (1-x^2) P'_N (they agree with the reference's
// 15-digit literals to 1e-15).
#include <cmath>
#include <cstring>

#include "sem_internal.h"

namespace sem {

static void legendre(int N, double x, double *p, double *dp) {
  double p0 = 1.0, p1 = x;
  if (N == 0) { *p = 1; *dp = 0; return; }
  for (int k = 2; k <= N; ++k) {
    double p2 = ((2.0 * k - 1.0) * x * p1 - (k - 1.0) * p0) / k;
    p0 = p1;
    p1 = p2;
  }
  *p = p1;
  *dp = N * (x * p1 - p0) / (x * x - 1.0);
}

Gll make_gll(int order) {
  Gll g;
  g.order = order;
  g.n = order + 1;
  const int N = order;
  g.xi[0] = -1.0;
  g.xi[N] = 1.0;
  for (int i = 1; i < N; ++i) {
    double x = -std::cos(M_PI * i / N);
    for (int it = 0; it < 100; ++it) {
      double p, dp;
      legendre(N, x, &p, &dp);
      double d2p = (2.0 * x * dp - N * (N + 1.0) * p) / (1.0 - x * x);
      double dx = dp / d2p;
      x -= dx;
      if (std::fabs(dx) < 1e-17) break;
    }
    g.xi[i] = x;
  }
  // symmetrise exactly
  for (int i = 0; i < g.n / 2; ++i) {
    double a = 0.5 * (g.xi[g.n - 1 - i] - g.xi[i]);
    g.xi[i] = -a;
    g.xi[g.n - 1 - i] = a;
  }
  if (g.n % 2) g.xi[g.n / 2] = 0.0;
  for (int i = 0; i <= N; ++i) {
    double p = 1.0, dp;
    if (i != 0 && i != N) legendre(N, g.xi[i], &p, &dp);
    g.w[i] = 2.0 / (N * (N + 1.0) * p * p);
  }
  for (int i = 0; i < g.n; ++i) lagrange_deriv(g, g.xi[i], g.D + i * g.n);
  return g;
}

void lagrange(const Gll &g, double x, double *l) {
  for (int j = 0; j < g.n; ++j) {
    double v = 1.0;
    for (int m = 0; m < g.n; ++m)
      if (m != j) v *= (x - g.xi[m]) / (g.xi[j] - g.xi[m]);
    l[j] = v;
  }
}

void lagrange_deriv(const Gll &g, double x, double *dl) {
  for (int j = 0; j < g.n; ++j) {
    double den = 1.0;
    for (int m = 0; m < g.n; ++m)
      if (m != j) den *= g.xi[j] - g.xi[m];
    double s = 0.0;
    for (int k = 0; k < g.n; ++k) {
      if (k == j) continue;
      double pr = 1.0;
      for (int m = 0; m < g.n; ++m)
        if (m != j && m != k) pr *= x - g.xi[m];
      s += pr;
    }
    dl[j] = s / den;
  }
}

// Reference-coordinate signs of the vertices (QuadP1.h:98-108, HexP1.cpp:242-298).
static const double kQuadRef[4][2] = {{-1, -1}, {1, -1}, {1, 1}, {-1, 1}};
static const double kHexRef[8][3] = {{-1, -1, -1}, {-1, 1, -1}, {1, 1, -1}, {1, -1, -1},
                                     {-1, -1, 1},  {1, -1, 1},  {1, 1, 1},  {-1, 1, 1}};

void p1_weights(int dim, const double *ref, double *w) {
  if (dim == 2) {
    for (int v = 0; v < 4; ++v)
      w[v] = 0.25 * (1 + ref[0] * kQuadRef[v][0]) * (1 + ref[1] * kQuadRef[v][1]);
  } else {
    for (int v = 0; v < 8; ++v)
      w[v] = 0.125 * (1 + ref[0] * kHexRef[v][0]) * (1 + ref[1] * kHexRef[v][1]) *
             (1 + ref[2] * kHexRef[v][2]);
  }
}

double jacobian(int dim, const double *vtx, const double *ref, double *J) {
  const int nv = dim == 2 ? 4 : 8;
  for (int i = 0; i < dim * dim; ++i) J[i] = 0.0;
  for (int v = 0; v < nv; ++v) {
    double d[3];
    if (dim == 2) {
      d[0] = 0.25 * kQuadRef[v][0] * (1 + ref[1] * kQuadRef[v][1]);
      d[1] = 0.25 * kQuadRef[v][1] * (1 + ref[0] * kQuadRef[v][0]);
    } else {
      const double *s = kHexRef[v];
      d[0] = 0.125 * s[0] * (1 + ref[1] * s[1]) * (1 + ref[2] * s[2]);
      d[1] = 0.125 * s[1] * (1 + ref[0] * s[0]) * (1 + ref[2] * s[2]);
      d[2] = 0.125 * s[2] * (1 + ref[0] * s[0]) * (1 + ref[1] * s[1]);
    }
    for (int a = 0; a < dim; ++a)
      for (int b = 0; b < dim; ++b) J[a * dim + b] += d[a] * vtx[v * dim + b];
  }
  if (dim == 2) return J[0] * J[3] - J[1] * J[2];
  return J[0] * (J[4] * J[8] - J[5] * J[7]) - J[1] * (J[3] * J[8] - J[5] * J[6]) +
         J[2] * (J[3] * J[7] - J[4] * J[6]);
}

double invert(int dim, const double *J, double *inv) {
  if (dim == 2) {
    double d = J[0] * J[3] - J[1] * J[2];
    inv[0] = J[3] / d;
    inv[1] = -J[1] / d;
    inv[2] = -J[2] / d;
    inv[3] = J[0] / d;
    return d;
  }
  double c00 = J[4] * J[8] - J[5] * J[7], c01 = J[5] * J[6] - J[3] * J[8],
         c02 = J[3] * J[7] - J[4] * J[6];
  double d = J[0] * c00 + J[1] * c01 + J[2] * c02;
  inv[0] = c00 / d;
  inv[1] = (J[2] * J[7] - J[1] * J[8]) / d;
  inv[2] = (J[1] * J[5] - J[2] * J[4]) / d;
  inv[3] = c01 / d;
  inv[4] = (J[0] * J[8] - J[2] * J[6]) / d;
  inv[5] = (J[2] * J[3] - J[0] * J[5]) / d;
  inv[6] = c02 / d;
  inv[7] = (J[1] * J[6] - J[0] * J[7]) / d;
  inv[8] = (J[0] * J[4] - J[1] * J[3]) / d;
  return d;
}

int check_hull(int dim, const double *pt, const double *vtx) {
  double J[9], zero[3] = {0, 0, 0};
  if (jacobian(dim, vtx, zero, J) <= 0) return -1;
  if (dim == 2) {
    int neg = 0, pos = 0;
    for (int i = 0; i < 4; ++i) {
      const double *p0 = vtx + 2 * i, *p1 = vtx + 2 * ((i + 1) % 4);
      double c = (p1[0] - p0[0]) * (pt[1] - p0[1]) - (p1[1] - p0[1]) * (pt[0] - p0[0]);
      if (c < 0) ++neg;
      else if (c > 0) ++pos;
      else { ++neg; ++pos; }
    }
    return (neg == 4 || pos == 4) ? 1 : 0;
  }
  // face corners a, b, d of each side, outward normal (b-a) x (d-a)
  static const int side[6][3] = {{0, 1, 3}, {4, 5, 7}, {0, 3, 4}, {2, 1, 6}, {0, 4, 1}, {2, 6, 3}};
  for (int f = 0; f < 6; ++f) {
    const double *a = vtx + 3 * side[f][0], *b = vtx + 3 * side[f][1], *d = vtx + 3 * side[f][2];
    double l1[3] = {b[0] - a[0], b[1] - a[1], b[2] - a[2]};
    double l2[3] = {d[0] - a[0], d[1] - a[1], d[2] - a[2]};
    double nrm[3] = {l1[1] * l2[2] - l1[2] * l2[1], l1[2] * l2[0] - l1[0] * l2[2],
                     l1[0] * l2[1] - l1[1] * l2[0]};
    double dot = nrm[0] * (pt[0] - a[0]) + nrm[1] * (pt[1] - a[1]) + nrm[2] * (pt[2] - a[2]);
    if (dot > 0) return 0;
  }
  return 1;
}

int inverse_map(int dim, const double *pt, const double *vtx, double *ref) {
  const int nv = dim == 2 ? 4 : 8;
  double r[3] = {0, 0, 0};
  for (int it = 0;; ++it) {
    double w[8], X[3] = {0, 0, 0};
    p1_weights(dim, r, w);
    for (int v = 0; v < nv; ++v)
      for (int a = 0; a < dim; ++a) X[a] += w[v] * vtx[v * dim + a];
    double f[3];
    bool conv = true;
    for (int a = 0; a < dim; ++a) {
      f[a] = pt[a] - X[a];
      if (!(std::fabs(f[a]) < 1e-6)) conv = false;
    }
    if (conv) {
      for (int a = 0; a < dim; ++a) ref[a] = r[a];
      return 0;
    }
    double J[9], inv[9];
    jacobian(dim, vtx, r, J);
    invert(dim, J, inv);
    // dX/dref = J^T, so the Newton step is J^{-T} f
    for (int a = 0; a < dim; ++a) {
      double s = 0;
      for (int b = 0; b < dim; ++b) s += inv[b * dim + a] * f[b];
      r[a] += s;
    }
    if (it > 10) return -1;
  }
}

}  // namespace sem
