/* GLL quadrature and Lagrange basis for the FP64 oracle (test infrastructure).
 *
 * Follows: TensorQuad.cpp:50-174 (GllPointsForOrder, GllIntegrationWeights-
 * ForOrder, setupGradientOperator, interpolateLagrangePolynomials) and
 * Hexahedra.cpp:74-130, :493-577, which read the sympy tables in
 * Quad/Autogen/order<N>_square.c and Hex/Autogen/order<N>_hex.c.  The
 * abscissae and weights are the reference's literals (gll_table.h); the
 * Lagrange values and derivatives are the product form through those
 * literal nodes, which is the polynomial the autogen code expands.
 */
#include <math.h>
#include <string.h>
#include "oracle.h"
#include "gll_table.h"

int orc_gll(int order, double *xi, double *w) {
  if (order < 1 || order > ORC_GLL_MAX_ORDER) return -1;
  for (int i = 0; i <= order; i++) {
    xi[i] = orc_gll_xi_lit[order - 1][i];
    w[i] = orc_gll_w_lit[order - 1][i];
  }
  return 0;
}

/* Legendre P_N(x) and its derivative by the three-term recurrence. */
static void legendre(int N, double x, double *p, double *dp) {
  double p0 = 1.0, p1 = x;
  if (N == 0) { *p = 1.0; *dp = 0.0; return; }
  for (int k = 2; k <= N; k++) {
    double p2 = ((2.0 * k - 1.0) * x * p1 - (k - 1.0) * p0) / k;
    p0 = p1; p1 = p2;
  }
  *p = p1;
  /* P'_N = N (x P_N - P_{N-1}) / (x^2 - 1) away from the end points. */
  *dp = N * (x * p1 - p0) / (x * x - 1.0);
}

/* GLL nodes are -1, 1 and the roots of P'_N; weights 2/(N(N+1) P_N(x)^2)
 * (Canuto et al., Spectral Methods, eq. 2.3.12).  Newton on P'_N using
 * P''_N = (2x P'_N - N(N+1) P_N)/(1-x^2). */
int orc_gll_exact(int order, double *xi, double *w) {
  int N = order;
  if (N < 1) return -1;
  xi[0] = -1.0; xi[N] = 1.0;
  for (int i = 1; i < N; i++) {
    double x = -cos(M_PI * i / N);
    for (int it = 0; it < 100; it++) {
      double p, dp;
      legendre(N, x, &p, &dp);
      double d2p = (2.0 * x * dp - N * (N + 1.0) * p) / (1.0 - x * x);
      double dx = dp / d2p;
      x -= dx;
      if (fabs(dx) < 1e-17) break;
    }
    xi[i] = x;
  }
  for (int i = 0; i <= N; i++) {
    double p, dp;
    if (i == 0 || i == N) p = (i == 0 && (N % 2)) ? -1.0 : 1.0;
    else legendre(N, xi[i], &p, &dp);
    w[i] = 2.0 / (N * (N + 1.0) * p * p);
  }
  return 0;
}

void orc_lagrange(int n, const double *xi, double x, double *l) {
  for (int j = 0; j < n; j++) {
    double v = 1.0;
    for (int m = 0; m < n; m++)
      if (m != j) v *= (x - xi[m]) / (xi[j] - xi[m]);
    l[j] = v;
  }
}

/* d/dx of the product form: sum over the dropped factor. */
void orc_lagrange_deriv(int n, const double *xi, double x, double *dl) {
  for (int j = 0; j < n; j++) {
    double denom = 1.0;
    for (int m = 0; m < n; m++)
      if (m != j) denom *= (xi[j] - xi[m]);
    double s = 0.0;
    for (int k = 0; k < n; k++) {
      if (k == j) continue;
      double prod = 1.0;
      for (int m = 0; m < n; m++)
        if (m != j && m != k) prod *= (x - xi[m]);
      s += prod;
    }
    dl[j] = s / denom;
  }
}

int orc_deriv_matrix(int order, double *D) {
  double xi[ORC_GLL_MAX_ORDER + 1], w[ORC_GLL_MAX_ORDER + 1];
  if (orc_gll(order, xi, w)) return -1;
  int n = order + 1;
  for (int i = 0; i < n; i++) orc_lagrange_deriv(n, xi, xi[i], D + i * n);
  return 0;
}
