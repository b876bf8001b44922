/* Tensor-element operators and physics for the FP64 oracle (test
 * infrastructure).
 *
 * Follows TensorQuad.cpp:176-210 (computeGradient), :286-348 (delta
 * coefficients, ParAtIntPts, applyTestAndIntegrate), :350-397
 * (applyGradTestAndIntegrate); Hexahedra.cpp:466-490, :579-754; physics
 * Scalar.cpp:34-88, Elastic2D.cpp:29-127, Elastic3D.cpp:33-133.  The
 * structure (Jacobian re-evaluated at every GLL point on every call, the
 * quad's detJ carried from computeGradient into applyGradTestAndIntegrate,
 * J^-1 applied at the output node for quads and J^-T per point for hexes)
 * is the reference's, so this is also the CPU-baseline port of that path.
 */
#include <math.h>
#include <string.h>
#include "oracle.h"
#include "oracle_internal.h"

int orc_tables(orc_tab *T, int dim, int order) {
  if (dim != 2 && dim != 3) return -1;
  if (order < 1 || order > (dim == 2 ? 10 : 7)) return -1;
  T->dim = dim; T->order = order; T->n = order + 1;
  T->nd = dim == 2 ? T->n * T->n : T->n * T->n * T->n;
  T->quad_ref = 1;
  orc_gll(order, T->xi, T->w);
  orc_deriv_matrix(order, T->D);
  return 0;
}

/* ---------------- quads ---------------- */

/* The quad Jacobian the element's mode takes (see orc_tab.quad_ref). */
static int quad_jac(const orc_tab *T, double r, double s, const double *vtx, double *det,
                    double *inv) {
  return T->quad_ref ? orc_quad_invjac(r, s, vtx, det, inv)
                     : orc_quad_invjac_consistent(r, s, vtx, det, inv);
}

/* TensorQuad.cpp:176-210; writes detJ per point (TensorQuad.cpp:194: the
 * gradient leaves mDetJac for applyGradTestAndIntegrate, :371-374). */
static void quad_gradient(const orc_tab *T, const double *vtx, const double *u, double *grad,
                          double *detJ) {
  int n = T->n;
  for (int s = 0; s < n; s++)
    for (int r = 0; r < n; r++) {
      int idx = r + s * n;
      double inv[4];
      quad_jac(T, T->xi[r], T->xi[s], vtx, &detJ[idx], inv);
      double g0 = 0, g1 = 0;
      for (int i = 0; i < n; i++) {
        g0 += T->D[r * n + i] * u[i + s * n];
        g1 += T->D[s * n + i] * u[r + i * n];
      }
      grad[2 * idx + 0] = inv[0] * g0 + inv[1] * g1;
      grad[2 * idx + 1] = inv[2] * g0 + inv[3] * g1;
    }
}

/* TensorQuad.cpp:350-397; f is n^2 x 2 row-major.  In the consistent mode
 * the flux takes J^-T at each quadrature point first and the output node
 * takes no transform (the hexes' order, Hexahedra.cpp:688-754). */
static void quad_grad_test_integrate(const orc_tab *T, const double *vtx, const double *detJ,
                                     const double *f, double *out) {
  int n = T->n;
  const double *w = T->w, *D = T->D;
  if (!T->quad_ref) {
    double fr[2 * ORC_NDMAX];
    for (int s = 0; s < n; s++)
      for (int r = 0; r < n; r++) {
        int idx = r + s * n;
        double det, inv[4];
        orc_quad_invjac_consistent(T->xi[r], T->xi[s], vtx, &det, inv);
        fr[2 * idx + 0] = inv[0] * f[2 * idx] + inv[2] * f[2 * idx + 1];
        fr[2 * idx + 1] = inv[1] * f[2 * idx] + inv[3] * f[2 * idx + 1];
      }
    for (int s = 0; s < n; s++)
      for (int r = 0; r < n; r++) {
        double d0 = 0, d1 = 0;
        for (int i = 0; i < n; i++) {
          int ri = i + s * n, si = r + i * n;
          d0 += detJ[ri] * fr[2 * ri + 0] * D[i * n + r] * w[i];
          d1 += detJ[si] * fr[2 * si + 1] * D[i * n + s] * w[i];
        }
        out[r + s * n] = d0 * w[s] + d1 * w[r];
      }
    return;
  }
  for (int s = 0; s < n; s++)
    for (int r = 0; r < n; r++) {
      double fx0 = 0, fx1 = 0, fy0 = 0, fy1 = 0;
      for (int i = 0; i < n; i++) {
        int ri = i + s * n, si = r + i * n;
        fx0 += detJ[ri] * f[2 * ri + 0] * D[i * n + r] * w[i];
        fx1 += detJ[si] * f[2 * si + 0] * D[i * n + s] * w[i];
        fy0 += detJ[ri] * f[2 * ri + 1] * D[i * n + r] * w[i];
        fy1 += detJ[si] * f[2 * si + 1] * D[i * n + s] * w[i];
      }
      fx0 *= w[s]; fx1 *= w[r]; fy0 *= w[s]; fy1 *= w[r];
      double det, inv[4];
      quad_jac(T, T->xi[r], T->xi[s], vtx, &det, inv);
      out[r + s * n] = (inv[0] * fx0 + inv[1] * fx1) + (inv[2] * fy0 + inv[3] * fy1);
    }
}

/* ---------------- hexes ---------------- */

/* Hexahedra.cpp:579-632 */
static void hex_gradient(const orc_tab *T, const double *vtx, const double *u, double *grad) {
  int n = T->n, nn = n * n;
  for (int t = 0; t < n; t++)
    for (int s = 0; s < n; s++)
      for (int r = 0; r < n; r++) {
        int idx = r + s * n + t * nn;
        double det, inv[9];
        orc_hex_invjac(T->xi[r], T->xi[s], T->xi[t], vtx, &det, inv);
        double g0 = 0, g1 = 0, g2 = 0;
        for (int i = 0; i < n; i++) {
          g0 += T->D[r * n + i] * u[i + s * n + t * nn];
          g1 += T->D[s * n + i] * u[r + i * n + t * nn];
          g2 += T->D[t * n + i] * u[r + s * n + i * nn];
        }
        for (int a = 0; a < 3; a++)
          grad[3 * idx + a] = inv[3 * a] * g0 + inv[3 * a + 1] * g1 + inv[3 * a + 2] * g2;
      }
}

/* Hexahedra.cpp:688-754; f is n^3 x 3 row-major; scratch holds 4 n^3. */
static void hex_grad_test_integrate(const orc_tab *T, const double *vtx, const double *f,
                                    double *out, double *scratch) {
  int n = T->n, nn = n * n;
  double *detJ = scratch, *fxyz = scratch + T->nd;
  for (int t = 0; t < n; t++)
    for (int s = 0; s < n; s++)
      for (int r = 0; r < n; r++) {
        int idx = r + s * n + t * nn;
        double inv[9];
        orc_hex_invjac(T->xi[r], T->xi[s], T->xi[t], vtx, &detJ[idx], inv);
        const double *fi = f + 3 * idx;
        for (int a = 0; a < 3; a++)
          fxyz[3 * idx + a] = inv[a] * fi[0] + inv[3 + a] * fi[1] + inv[6 + a] * fi[2];
      }
  const double *w = T->w, *D = T->D;
  for (int t = 0; t < n; t++)
    for (int s = 0; s < n; s++)
      for (int r = 0; r < n; r++) {
        double dr = 0, ds = 0, dt = 0;
        for (int i = 0; i < n; i++) {
          int ri = i + s * n + t * nn, si = r + i * n + t * nn, ti = r + s * n + i * nn;
          dr += detJ[ri] * fxyz[3 * ri + 0] * D[i * n + r] * w[i];
          ds += detJ[si] * fxyz[3 * si + 1] * D[i * n + s] * w[i];
          dt += detJ[ti] * fxyz[3 * ti + 2] * D[i * n + t] * w[i];
        }
        dr *= w[s] * w[t];
        ds *= w[r] * w[t];
        dt *= w[r] * w[s];
        out[r + s * n + t * nn] = dr + ds + dt;
      }
}

/* ---------------- shared element helpers ---------------- */

/* ParAtIntPts: TensorQuad.cpp:308-323, Hexahedra.cpp:634-656.
 * vpar: nv values of one parameter at the element vertices. */
void orc_par_at_pts(const orc_tab *T, const double *vpar, int stride, double *out) {
  int n = T->n;
  if (T->dim == 2) {
    for (int s = 0; s < n; s++)
      for (int r = 0; r < n; r++) {
        double w4[4];
        orc_quad_interp(T->xi[r], T->xi[s], w4);
        double v = 0;
        for (int k = 0; k < 4; k++) v += w4[k] * vpar[k * stride];
        out[r + s * n] = v;
      }
  } else {
    for (int t = 0; t < n; t++)
      for (int s = 0; s < n; s++)
        for (int r = 0; r < n; r++) {
          double w8[8];
          orc_hex_interp(T->xi[r], T->xi[s], T->xi[t], w8);
          double v = 0;
          for (int k = 0; k < 8; k++) v += w8[k] * vpar[k * stride];
          out[r + n * (s + n * t)] = v;
        }
  }
}

/* applyTestAndIntegrate: TensorQuad.cpp:325-348, Hexahedra.cpp:658-686. */
void orc_test_integrate(const orc_tab *T, const double *vtx, const double *f, double *out) {
  int n = T->n;
  if (T->dim == 2) {
    for (int s = 0; s < n; s++)
      for (int r = 0; r < n; r++) {
        double det;
        quad_jac(T, T->xi[r], T->xi[s], vtx, &det, 0);
        int idx = r + s * n;
        out[idx] = f[idx] * det * T->w[r] * T->w[s];
      }
  } else {
    for (int t = 0; t < n; t++)
      for (int s = 0; s < n; s++)
        for (int r = 0; r < n; r++) {
          double det;
          orc_hex_invjac(T->xi[r], T->xi[s], T->xi[t], vtx, &det, 0);
          int idx = r + n * (s + n * t);
          out[idx] = f[idx] * det * T->w[r] * T->w[s] * T->w[t];
        }
  }
}

/* Tensor Lagrange values at a reference point (interpolate_order<N>_*). */
void orc_tensor_lagrange(const orc_tab *T, const double *ref, double *l) {
  int n = T->n;
  double lr[16], ls[16], lt[16];
  orc_lagrange(n, T->xi, ref[0], lr);
  orc_lagrange(n, T->xi, ref[1], ls);
  if (T->dim == 3) orc_lagrange(n, T->xi, ref[2], lt);
  for (int p = 0; p < T->nd; p++) {
    int i = p % n, j = (p / n) % n, k = p / (n * n);
    l[p] = lr[i] * ls[j] * (T->dim == 3 ? lt[k] : 1.0);
  }
}

/* getDeltaFunctionCoefficients: TensorQuad.cpp:286-306, Hexahedra.cpp:466-490. */
void orc_delta_coeffs(const orc_tab *T, const double *vtx, const double *ref, double *out) {
  int n = T->n;
  orc_tensor_lagrange(T, ref, out);
  for (int p = 0; p < T->nd; p++) {
    int i = p % n, j = (p / n) % n, k = p / (n * n);
    double det;
    if (T->dim == 2) quad_jac(T, T->xi[i], T->xi[j], vtx, &det, 0);
    else orc_hex_invjac(T->xi[i], T->xi[j], T->xi[k], vtx, &det, 0);
    out[p] /= (T->w[i] * T->w[j] * (T->dim == 3 ? T->w[k] : 1.0) * det);
  }
}

/* Physical gradient of every basis function at a reference point:
 * dphi[p*d + a] = d phi_p / d x_a (J^-1 grad_r, HexP1.cpp:189).  Used by the
 * moment-tensor source, which has no reference counterpart: quads take the
 * consistent bilinear Jacobian in either mode. */
void orc_basis_gradient_at(const orc_tab *T, const double *vtx, const double *ref,
                           double *dphi) {
  int n = T->n, d = T->dim;
  double l[3][16], dl[3][16], det, inv[9];
  for (int a = 0; a < d; a++) {
    orc_lagrange(n, T->xi, ref[a], l[a]);
    orc_lagrange_deriv(n, T->xi, ref[a], dl[a]);
  }
  if (d == 2) orc_quad_invjac_consistent(ref[0], ref[1], vtx, &det, inv);
  else orc_hex_invjac(ref[0], ref[1], ref[2], vtx, &det, inv);
  for (int p = 0; p < T->nd; p++) {
    int i = p % n, j = (p / n) % n, k = p / (n * n);
    double gr[3];
    if (d == 2) {
      gr[0] = dl[0][i] * l[1][j];
      gr[1] = l[0][i] * dl[1][j];
    } else {
      gr[0] = dl[0][i] * l[1][j] * l[2][k];
      gr[1] = l[0][i] * dl[1][j] * l[2][k];
      gr[2] = l[0][i] * l[1][j] * dl[2][k];
    }
    for (int a = 0; a < d; a++) {
      double v = 0;
      for (int b = 0; b < d; b++) v += inv[a * d + b] * gr[b];
      dphi[p * d + a] = v;
    }
  }
}

/* ---------------- physics stiffness ---------------- */

/* Elastic3D.cpp:33-54 precompute: cc[q*10 + {rho,c11,c12,c13,c22,c23,c33,c44,c55,c66}]. */
void orc_e3d_constants(const orc_tab *T, const double *vpar, double *cc) {
  double rho[ORC_NDMAX], vpv[ORC_NDMAX], vph[ORC_NDMAX], vsv[ORC_NDMAX], vsh[ORC_NDMAX],
      eta[ORC_NDMAX];
  orc_par_at_pts(T, vpar + 0, 6, rho);
  orc_par_at_pts(T, vpar + 1, 6, vpv);
  orc_par_at_pts(T, vpar + 2, 6, vph);
  orc_par_at_pts(T, vpar + 3, 6, vsv);
  orc_par_at_pts(T, vpar + 4, 6, vsh);
  orc_par_at_pts(T, vpar + 5, 6, eta);
  for (int q = 0; q < T->nd; q++) {
    double c11 = rho[q] * pow(vph[q], 2), c22 = rho[q] * pow(vph[q], 2);
    double c33 = rho[q] * pow(vpv[q], 2), c44 = rho[q] * pow(vsv[q], 2);
    double c55 = rho[q] * pow(vsv[q], 2), c66 = rho[q] * pow(vsh[q], 2);
    double *o = cc + 10 * q;
    o[0] = rho[q]; o[1] = c11; o[2] = c11 - 2 * c66; o[3] = eta[q] * (c11 - 2 * c44);
    o[4] = c22; o[5] = eta[q] * (c11 - 2 * c44); o[6] = c33; o[7] = c44; o[8] = c55; o[9] = c66;
  }
}

/* Element stiffness for one element.  u, ku are column-major n^d x c.
 * vpar: vertex parameters (nv x npar); cc: Elastic3D constants or NULL. */
void orc_stiffness(const orc_tab *T, int physics, const double *vtx, const double *vpar,
                   const double *cc, const double *u, double *ku, double *work) {
  int nd = T->nd;
  double *detJ = work;              /* nd   */
  double *g0 = work + nd;           /* 3 nd */
  double *g1 = work + 4 * nd;       /* 3 nd */
  double *g2 = work + 7 * nd;       /* 3 nd */
  double *flux = work + 10 * nd;    /* 3 nd */
  double *par = work + 13 * nd;     /* 4 nd */
  double *hs = work + 17 * nd;      /* 4 nd (hex scratch) */
  if (T->dim == 2) {
    if (physics == ORC_ACOUSTIC) {
      /* Scalar.cpp:42-71 */
      quad_gradient(T, vtx, u, g0, detJ);
      orc_par_at_pts(T, vpar, 1, par);
      for (int q = 0; q < nd; q++) {
        double vp2 = pow(par[q], 2);
        flux[2 * q] = vp2 * g0[2 * q];
        flux[2 * q + 1] = vp2 * g0[2 * q + 1];
      }
      quad_grad_test_integrate(T, vtx, detJ, flux, ku);
    } else {
      /* Elastic2D.cpp:53-107: strain (ux,x ux,y uy,x uy,y); mc13 = mc23 = 0 */
      quad_gradient(T, vtx, u, g0, detJ);
      quad_gradient(T, vtx, u + nd, g1, detJ);
      double *c11 = par, *c13 = par + nd, *c33 = par + 2 * nd, *c55 = par + 3 * nd;
      orc_par_at_pts(T, vpar + 1, 5, c11);
      orc_par_at_pts(T, vpar + 2, 5, c13);
      orc_par_at_pts(T, vpar + 3, 5, c33);
      orc_par_at_pts(T, vpar + 4, 5, c55);
      double *sxx = g2, *syy = g2 + nd, *sxy = g2 + 2 * nd;
      for (int q = 0; q < nd; q++) {
        double exx = g0[2 * q], uxy = g0[2 * q + 1], uyx = g1[2 * q], eyy = g1[2 * q + 1];
        double sh = uxy + uyx;
        sxx[q] = c11[q] * exx + c13[q] * eyy + 0.0 * sh;
        syy[q] = c13[q] * exx + c33[q] * eyy + 0.0 * sh;
        sxy[q] = 0.0 * exx + 0.0 * eyy + c55[q] * sh;
      }
      for (int q = 0; q < nd; q++) { flux[2 * q] = sxx[q]; flux[2 * q + 1] = sxy[q]; }
      quad_grad_test_integrate(T, vtx, detJ, flux, ku);
      for (int q = 0; q < nd; q++) { flux[2 * q] = sxy[q]; flux[2 * q + 1] = syy[q]; }
      quad_grad_test_integrate(T, vtx, detJ, flux, ku + nd);
    }
  } else {
    if (physics == ORC_ACOUSTIC) {
      hex_gradient(T, vtx, u, g0);
      orc_par_at_pts(T, vpar, 1, par);
      for (int q = 0; q < nd; q++) {
        double vp2 = pow(par[q], 2);
        for (int a = 0; a < 3; a++) flux[3 * q + a] = vp2 * g0[3 * q + a];
      }
      hex_grad_test_integrate(T, vtx, flux, ku, hs);
    } else {
      /* Elastic3D.cpp:70-118 */
      hex_gradient(T, vtx, u, g0);
      hex_gradient(T, vtx, u + nd, g1);
      hex_gradient(T, vtx, u + 2 * nd, g2);
      for (int q = 0; q < nd; q++) {
        const double *c = cc + 10 * q;
        double e0 = g0[3 * q], e1 = g1[3 * q + 1], e2 = g2[3 * q + 2];
        double e3 = g1[3 * q + 2] + g2[3 * q + 1];
        double e4 = g0[3 * q + 2] + g2[3 * q + 0];
        double e5 = g0[3 * q + 1] + g1[3 * q + 0];
        double s0 = c[1] * e0 + c[2] * e1 + c[3] * e2;
        double s1 = c[2] * e0 + c[4] * e1 + c[5] * e2;
        double s2 = c[3] * e0 + c[5] * e1 + c[6] * e2;
        double s3 = c[7] * e3, s4 = c[8] * e4, s5 = c[9] * e5;
        /* stress kept in par[] (6 nd needed: par+hs region is 8 nd) */
        double *S = par;
        S[6 * q + 0] = s0; S[6 * q + 1] = s1; S[6 * q + 2] = s2;
        S[6 * q + 3] = s3; S[6 * q + 4] = s4; S[6 * q + 5] = s5;
      }
      double *S = par;
      double *hs2 = work + 23 * nd; /* 4 nd */
      static const int rows[3][3] = {{0, 5, 4}, {5, 1, 3}, {4, 3, 2}};
      for (int comp = 0; comp < 3; comp++) {
        for (int q = 0; q < nd; q++)
          for (int a = 0; a < 3; a++) flux[3 * q + a] = S[6 * q + rows[comp][a]];
        hex_grad_test_integrate(T, vtx, flux, ku + comp * nd, hs2);
      }
    }
  }
}

/* ---------------- single-element public wrappers ---------------- */

int orc_elem_gradient(int dim, int order, const double *vtx, const double *u, double *grad) {
  orc_tab T;
  if (orc_tables(&T, dim, order)) return -1;
  if (dim == 2) {
    double detJ[ORC_NDMAX];
    quad_gradient(&T, vtx, u, grad, detJ);
  } else {
    hex_gradient(&T, vtx, u, grad);
  }
  return 0;
}

int orc_elem_grad_test_integrate(int dim, int order, const double *vtx, const double *f,
                                 double *out) {
  orc_tab T;
  if (orc_tables(&T, dim, order)) return -1;
  if (dim == 2) {
    /* detJ as computeGradient leaves it (TensorQuad.cpp:194, :371-374) */
    double detJ[ORC_NDMAX], g[2 * ORC_NDMAX], z[ORC_NDMAX];
    memset(z, 0, sizeof z);
    quad_gradient(&T, vtx, z, g, detJ);
    quad_grad_test_integrate(&T, vtx, detJ, f, out);
  } else {
    double scratch[4 * ORC_NDMAX];
    hex_grad_test_integrate(&T, vtx, f, out, scratch);
  }
  return 0;
}

int orc_elem_test_integrate(int dim, int order, const double *vtx, const double *f, double *out) {
  orc_tab T;
  if (orc_tables(&T, dim, order)) return -1;
  orc_test_integrate(&T, vtx, f, out);
  return 0;
}

int orc_elem_delta_coeffs(int dim, int order, const double *vtx, const double *ref,
                          double *out) {
  orc_tab T;
  if (orc_tables(&T, dim, order)) return -1;
  orc_delta_coeffs(&T, vtx, ref, out);
  return 0;
}

int orc_elem_par_at_int_pts(int dim, int order, const double *vtx_par, double *out) {
  orc_tab T;
  if (orc_tables(&T, dim, order)) return -1;
  orc_par_at_pts(&T, vtx_par, 1, out);
  return 0;
}

int orc_elem_stiffness(int dim, int order, int physics, const double *vtx, const double *par,
                       const double *u, double *ku) {
  orc_tab T;
  if (orc_tables(&T, dim, order)) return -1;
  static __thread double work[ORC_WORK * ORC_NDMAX];
  double cc[10 * ORC_NDMAX];
  if (dim == 3 && physics == ORC_ELASTIC) orc_e3d_constants(&T, par, cc);
  orc_stiffness(&T, physics, vtx, par, cc, u, ku, work);
  return 0;
}
