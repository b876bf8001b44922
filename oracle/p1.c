/* Bilinear quad / trilinear hex geometry for the FP64 oracle (test
 * infrastructure).
 *
 * Follows QuadP1.h:20-108, QuadP1.cpp:8-92, :176-216 and HexP1.cpp:8-354.
 * Vertex orders: quad v0(-1,-1) v1(+1,-1) v2(+1,+1) v3(-1,+1)
 * (QuadP1.h:98-108); hex v0(-1,-1,-1) v1(-1,+1,-1) v2(+1,+1,-1) v3(+1,-1,-1)
 * v4(-1,-1,+1) v5(+1,-1,+1) v6(+1,+1,+1) v7(-1,+1,+1) (HexP1.cpp:242-298).
 * J has rows d/dr, d/ds(, d/dt) and columns x, y(, z) (QuadP1.h:85-86,
 * HexP1.cpp:231-235); the element is valid when det J > 0.
 */
#include <math.h>
#include "oracle.h"

static const double quad_ref[4][2] = {{-1, -1}, {1, -1}, {1, 1}, {-1, 1}};
static const double hex_ref[8][3] = {{-1, -1, -1}, {-1, 1, -1}, {1, 1, -1}, {1, -1, -1},
                                     {-1, -1, 1},  {1, -1, 1},  {1, 1, 1},  {-1, 1, 1}};

/* QuadP1.h:79-90.  The reference builds row d/dr from dn*dr(r) and row d/ds
 * from dn*ds(s) (QuadP1.h:20-27), i.e. each row is evaluated at the other
 * coordinate than the one a bilinear map's derivative depends on.  For a
 * parallelogram J is constant and this is exact; it is restated as is. */
int orc_quad_invjac(double r, double s, const double *vtx, double *det, double *inv) {
  const double dr[4] = {-(1 - r) / 4.0, (1 - r) / 4.0, (1 + r) / 4.0, -(1 + r) / 4.0};
  const double ds[4] = {(1 - s) * -1.0 / 4.0, (1 + s) * -1.0 / 4.0, (1 + s) * 1.0 / 4.0,
                        (1 - s) * 1.0 / 4.0};
  double J[2][2] = {{0, 0}, {0, 0}};
  for (int v = 0; v < 4; v++) {
    J[0][0] += dr[v] * vtx[2 * v + 0];
    J[0][1] += dr[v] * vtx[2 * v + 1];
    J[1][0] += ds[v] * vtx[2 * v + 0];
    J[1][1] += ds[v] * vtx[2 * v + 1];
  }
  double d = J[0][0] * J[1][1] - J[0][1] * J[1][0];
  *det = d;
  if (inv) {
    inv[0] = J[1][1] / d;
    inv[1] = -J[0][1] / d;
    inv[2] = -J[1][0] / d;
    inv[3] = J[0][0] / d;
  }
  return d > 0 ? 0 : -1;
}

int orc_quad_invjac_consistent(double r, double s, const double *vtx, double *det, double *inv) {
  const double dr[4] = {-(1 - s) / 4.0, (1 - s) / 4.0, (1 + s) / 4.0, -(1 + s) / 4.0};
  const double ds[4] = {-(1 - r) / 4.0, -(1 + r) / 4.0, (1 + r) / 4.0, (1 - r) / 4.0};
  double J[2][2] = {{0, 0}, {0, 0}};
  for (int v = 0; v < 4; v++) {
    J[0][0] += dr[v] * vtx[2 * v + 0];
    J[0][1] += dr[v] * vtx[2 * v + 1];
    J[1][0] += ds[v] * vtx[2 * v + 0];
    J[1][1] += ds[v] * vtx[2 * v + 1];
  }
  double d = J[0][0] * J[1][1] - J[0][1] * J[1][0];
  *det = d;
  if (inv) {
    inv[0] = J[1][1] / d;
    inv[1] = -J[0][1] / d;
    inv[2] = -J[1][0] / d;
    inv[3] = J[0][0] / d;
  }
  return d > 0 ? 0 : -1;
}

/* QuadP1.h:98-108 */
void orc_quad_interp(double r, double s, double *w4) {
  w4[0] = +0.25 * r * s - 0.25 * r - 0.25 * s + 0.25;
  w4[1] = -0.25 * r * s + 0.25 * r - 0.25 * s + 0.25;
  w4[2] = +0.25 * r * s + 0.25 * r + 0.25 * s + 0.25;
  w4[3] = -0.25 * r * s - 0.25 * r + 0.25 * s + 0.25;
}

/* QuadP1.cpp:8-39: inside or on the boundary when the point is on the same
 * side of all four edges (zero cross products count for both sides).
 * Returns 1 inside, 0 outside, -1 when det J(0,0) <= 0 (the reference
 * throws "Poorly behaved TensorQuad"). */
int orc_quad_check_hull(double x, double y, const double *vtx) {
  double det;
  orc_quad_invjac(0, 0, vtx, &det, 0);
  if (det <= 0) return -1;
  int nneg = 0, npos = 0;
  for (int i = 0; i < 4; i++) {
    const double *p0 = vtx + 2 * i, *p1 = vtx + 2 * ((i + 1) % 4);
    double vx = p1[0] - p0[0], vy = p1[1] - p0[1];
    double px = x - p0[0], py = y - p0[1];
    double c = vx * py - vy * px;
    if (c < 0) nneg++;
    else if (c > 0) npos++;
    else { npos++; nneg++; }
  }
  return (nneg == 4 || npos == 4) ? 1 : 0;
}

static void quad_map(double r, double s, const double *vtx, double *x, double *J) {
  double w[4];
  orc_quad_interp(r, s, w);
  x[0] = x[1] = 0;
  J[0] = J[1] = J[2] = J[3] = 0; /* [[dx/dr, dx/ds], [dy/dr, dy/ds]] */
  for (int v = 0; v < 4; v++) {
    double rv = quad_ref[v][0], sv = quad_ref[v][1];
    double dndr = rv * (1 + s * sv) / 4.0, dnds = sv * (1 + r * rv) / 4.0;
    x[0] += w[v] * vtx[2 * v];
    x[1] += w[v] * vtx[2 * v + 1];
    J[0] += dndr * vtx[2 * v];
    J[1] += dnds * vtx[2 * v];
    J[2] += dndr * vtx[2 * v + 1];
    J[3] += dnds * vtx[2 * v + 1];
  }
}

/* QuadP1.cpp:41-92: Newton from (0,0) with absolute tolerance 1e-6 on the
 * physical residual; gives up after the 11th update. */
int orc_quad_inverse_map(double x, double y, const double *vtx, double *rs) {
  double r = 0, s = 0;
  for (int it = 0;; it++) {
    double X[2], J[4];
    quad_map(r, s, vtx, X, J);
    double fx = x - X[0], fy = y - X[1];
    if (fabs(fx) < 1e-6 && fabs(fy) < 1e-6) { rs[0] = r; rs[1] = s; return 0; }
    double d = J[0] * J[3] - J[1] * J[2];
    r += (J[3] * fx - J[1] * fy) / d;
    s += (-J[2] * fx + J[0] * fy) / d;
    if (it > 10) return -1;
  }
}

/* QuadP1.cpp:176-216 */
void orc_quad_nodal_points(int order, const double *vtx, double *xy) {
  double xi[16], w[16];
  orc_gll(order, xi, w);
  int n = order + 1;
  for (int j = 0; j < n; j++)
    for (int i = 0; i < n; i++) {
      double r = (xi[i] + 1) / 2, s = (xi[j] + 1) / 2;
      int p = i + n * j;
      for (int c = 0; c < 2; c++) {
        double v1 = vtx[c], v2 = vtx[2 + c], v3 = vtx[4 + c], v4 = vtx[6 + c];
        xy[2 * p + c] = v1 + (v2 - v1) * r + (v4 + (v3 - v4) * r - v1 - (v2 - v1) * r) * s;
      }
    }
}

static void hex_jac(double r, double s, double t, const double *vtx, double J[3][3]) {
  for (int a = 0; a < 3; a++)
    for (int b = 0; b < 3; b++) J[a][b] = 0;
  for (int v = 0; v < 8; v++) {
    double rv = hex_ref[v][0], sv = hex_ref[v][1], tv = hex_ref[v][2];
    double d[3] = {rv * (1 + s * sv) * (1 + t * tv) / 8.0, sv * (1 + r * rv) * (1 + t * tv) / 8.0,
                   tv * (1 + r * rv) * (1 + s * sv) / 8.0};
    for (int a = 0; a < 3; a++)
      for (int b = 0; b < 3; b++) J[a][b] += d[a] * vtx[3 * v + b];
  }
}

static double inv3(const double J[3][3], double *inv) {
  double c00 = J[1][1] * J[2][2] - J[1][2] * J[2][1];
  double c01 = J[1][2] * J[2][0] - J[1][0] * J[2][2];
  double c02 = J[1][0] * J[2][1] - J[1][1] * J[2][0];
  double det = J[0][0] * c00 + J[0][1] * c01 + J[0][2] * c02;
  if (inv) {
    inv[0] = c00 / det;
    inv[1] = (J[0][2] * J[2][1] - J[0][1] * J[2][2]) / det;
    inv[2] = (J[0][1] * J[1][2] - J[0][2] * J[1][1]) / det;
    inv[3] = c01 / det;
    inv[4] = (J[0][0] * J[2][2] - J[0][2] * J[2][0]) / det;
    inv[5] = (J[0][2] * J[1][0] - J[0][0] * J[1][2]) / det;
    inv[6] = c02 / det;
    inv[7] = (J[0][1] * J[2][0] - J[0][0] * J[2][1]) / det;
    inv[8] = (J[0][0] * J[1][1] - J[0][1] * J[1][0]) / det;
  }
  return det;
}

/* HexP1.cpp:189-240 (trilinear map derivative, det, inverse). */
int orc_hex_invjac(double r, double s, double t, const double *vtx, double *det, double *inv) {
  double J[3][3];
  hex_jac(r, s, t, vtx, J);
  *det = inv3(J, inv);
  return *det > 0 ? 0 : -1;
}

/* HexP1.cpp:242-298 */
void orc_hex_interp(double r, double s, double t, double *w8) {
  for (int v = 0; v < 8; v++)
    w8[v] = (1 + r * hex_ref[v][0]) * (1 + s * hex_ref[v][1]) * (1 + t * hex_ref[v][2]) / 8.0;
}

/* HexP1.cpp:8-84: outside when the point is strictly on the outer side of
 * any face plane (normals from (b-a)x(d-a) of the listed face corners).
 * -1 when det J(0,0,0) <= 0 ("Poorly behaved Hexahedra"). */
int orc_hex_check_hull(double x, double y, double z, const double *vtx) {
  static const int side[6][4] = {{0, 1, 2, 3}, {4, 5, 6, 7}, {0, 3, 5, 4},
                                 {2, 1, 7, 6}, {0, 4, 7, 1}, {2, 6, 5, 3}};
  double det;
  orc_hex_invjac(0, 0, 0, vtx, &det, 0);
  if (det <= 0) return -1;
  double p[3] = {x, y, z};
  for (int f = 0; f < 6; f++) {
    const double *a = vtx + 3 * side[f][0], *b = vtx + 3 * side[f][1], *d = vtx + 3 * side[f][3];
    double l1[3] = {b[0] - a[0], b[1] - a[1], b[2] - a[2]};
    double l2[3] = {d[0] - a[0], d[1] - a[1], d[2] - a[2]};
    double nrm[3] = {l1[1] * l2[2] - l1[2] * l2[1], l1[2] * l2[0] - l1[0] * l2[2],
                     l1[0] * l2[1] - l1[1] * l2[0]};
    double dot = nrm[0] * (p[0] - a[0]) + nrm[1] * (p[1] - a[1]) + nrm[2] * (p[2] - a[2]);
    if (dot > 0) return 0;
  }
  return 1;
}

/* HexP1.cpp:135-187: Newton from (0,0,0), update J^{-T} residual, tolerance
 * 1e-6 absolute, at most 11 updates. */
int orc_hex_inverse_map(double x, double y, double z, const double *vtx, double *rst) {
  double r = 0, s = 0, t = 0;
  for (int it = 0;; it++) {
    double w[8], X[3] = {0, 0, 0};
    orc_hex_interp(r, s, t, w);
    for (int v = 0; v < 8; v++)
      for (int c = 0; c < 3; c++) X[c] += w[v] * vtx[3 * v + c];
    double f[3] = {x - X[0], y - X[1], z - X[2]};
    if (fabs(f[0]) < 1e-6 && fabs(f[1]) < 1e-6 && fabs(f[2]) < 1e-6) {
      rst[0] = r; rst[1] = s; rst[2] = t;
      return 0;
    }
    double det, inv[9];
    orc_hex_invjac(r, s, t, vtx, &det, inv);
    /* solution += inv^T f */
    r += inv[0] * f[0] + inv[3] * f[1] + inv[6] * f[2];
    s += inv[1] * f[0] + inv[4] * f[1] + inv[7] * f[2];
    t += inv[2] * f[0] + inv[5] * f[1] + inv[8] * f[2];
    if (it > 10) return -1;
  }
}

/* HexP1.cpp:312-354 */
void orc_hex_nodal_points(int order, const double *vtx, double *xyz) {
  double xi[16], w[16];
  orc_gll(order, xi, w);
  int n = order + 1;
  for (int k = 0; k < n; k++)
    for (int j = 0; j < n; j++)
      for (int i = 0; i < n; i++) {
        double wv[8];
        orc_hex_interp(xi[i], xi[j], xi[k], wv);
        int p = i + n * (j + n * k);
        for (int c = 0; c < 3; c++) {
          double acc = 0;
          for (int v = 0; v < 8; v++) acc += wv[v] * vtx[3 * v + c];
          xyz[3 * p + c] = acc;
        }
      }
}
