/* Internal declarations shared by the oracle's translation units. */
#pragma once
#include "oracle.h"

#define ORC_NDMAX 1331 /* (10+1)^3 upper bound on nodes per element */
#define ORC_WORK 28    /* scratch slots of ORC_NDMAX doubles for orc_stiffness */

typedef struct {
  int dim, order, n, nd;
  /* quads: 1 = the reference's operator (QuadP1.h:20-27 Jacobian rows with
   * each derivative evaluated at the other coordinate; J^-1 applied at the
   * output node, TensorQuad.cpp:383-390), 0 = the consistent bilinear
   * Jacobian with J^-T at the quadrature points (as the hexes do) */
  int quad_ref;
  double xi[16], w[16], D[16 * 16];
} orc_tab;

int orc_tables(orc_tab *T, int dim, int order);
void orc_par_at_pts(const orc_tab *T, const double *vpar, int stride, double *out);
void orc_test_integrate(const orc_tab *T, const double *vtx, const double *f, double *out);
void orc_tensor_lagrange(const orc_tab *T, const double *ref, double *l);
void orc_delta_coeffs(const orc_tab *T, const double *vtx, const double *ref, double *out);
void orc_basis_gradient_at(const orc_tab *T, const double *vtx, const double *ref, double *dphi);
void orc_e3d_constants(const orc_tab *T, const double *vpar, double *cc);
void orc_stiffness(const orc_tab *T, int physics, const double *vtx, const double *vpar,
                   const double *cc, const double *u, double *ku, double *work);
