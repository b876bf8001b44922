/*
 * FP64 CPU oracle for the spectral-element time-stepping path of Salvus
 * (tensor quads/hexes; Scalar, Elastic2D, Elastic3D physics).
 *
 * TEST INFRASTRUCTURE ONLY.  Nothing in the product links, loads or calls
 * this library; only tests/, the smoke check and the CPU-baseline leg of
 * bench.py use it, as the checker.  It restates the reference's algorithm in
 * plain C99 (no PETSc, no Eigen); each function cites the reference
 * file:line it follows (paths relative to the reference repository root).
 *
 * Conventions
 *   order N, n = N+1 GLL points per direction, d = 2|3, c = components.
 *   Element-local node p = i + n*j (+ n*n*k), i along r (TensorQuad.cpp:187,
 *   Hexahedra.cpp:591).  Vertex coordinates per element: nv x d row-major
 *   (nv = 4 quad, 8 hex) in the reference's vertex order (QuadP1.h:98-108,
 *   HexP1.cpp:123-133).  Global fields are component-planar: f[comp*G + g].
 *   Material: per element, per vertex, per parameter: par[(e*nv + v)*npar+q].
 *   Parameter order: Scalar {VP}; Elastic2D {RHO,C11,C13,C33,C55};
 *   Elastic3D {RHO,VPV,VPH,VSV,VSH,ETA}.
 */
#pragma once
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum { ORC_ACOUSTIC = 0, ORC_ELASTIC = 1 };

/* Reference-behaviour switches (bitmask).
 * ORC_QUIRK_NEWMARK3D  : Order2Newmark.cpp:93-96 iterates {ax,ay,ax,a}; for
 *                        Elastic2D and Elastic3D the x component is advanced
 *                        twice per step, y once and z never.  Off = every
 *                        component advanced once.
 * ORC_QUIRK_E3D_SOURCE : Elastic3D.cpp:126-133 skips applyTestAndIntegrate,
 *                        so its nodal load is l_i(xs)/(w_i detJ_i) f(t).
 *                        Off = l_i(xs) f(t) as Scalar/Elastic2D give.
 * ORC_QUIRK_QUAD_JACOBIAN : quads take the reference's operator: the
 *                        Jacobian of QuadP1.h:20-27 (row d/dr built from
 *                        dN/dr evaluated at r, row d/ds from dN/ds at s;
 *                        exact for parallelograms only), J^-1 applied at the
 *                        output node (TensorQuad.cpp:383-390) and that det J
 *                        in mass and source integration.  Off = consistent
 *                        bilinear Jacobian, J^-T at the quadrature points.
 */
enum { ORC_QUIRK_NEWMARK3D = 1, ORC_QUIRK_E3D_SOURCE = 2, ORC_QUIRK_QUAD_JACOBIAN = 4 };

/* ---------------- GLL quadrature (a19) ---------------- */
/* Reference literals (Quad/Autogen/order<N>_square.c); 0 on success. */
int orc_gll(int order, double *xi, double *w);
/* Closed-form GLL nodes/weights by Newton on (1-x^2)P'_N, for pinning. */
int orc_gll_exact(int order, double *xi, double *w);
/* Lagrange basis through the n nodes xi, and its derivative, at x. */
void orc_lagrange(int n, const double *xi, double x, double *l);
void orc_lagrange_deriv(int n, const double *xi, double x, double *dl);
/* D[i*n + j] = l'_j(xi_i) (TensorQuad.cpp:106-142 grad.row(i)). */
int orc_deriv_matrix(int order, double *D);

/* ---------------- P1 geometry (a9, a10, a21, a24) ---------------- */
int orc_quad_invjac(double r, double s, const double *vtx, double *det, double *inv);
/* Consistent bilinear Jacobian of the P1 map (QuadP1.h:98-108 shape
 * functions differentiated): row d/dr uses dN/dr(s), row d/ds dN/ds(r). */
int orc_quad_invjac_consistent(double r, double s, const double *vtx, double *det, double *inv);
void orc_quad_interp(double r, double s, double *w4);
int orc_quad_check_hull(double x, double y, const double *vtx);
int orc_quad_inverse_map(double x, double y, const double *vtx, double *rs);
void orc_quad_nodal_points(int order, const double *vtx, double *xy);
int orc_hex_invjac(double r, double s, double t, const double *vtx, double *det, double *inv);
void orc_hex_interp(double r, double s, double t, double *w8);
int orc_hex_check_hull(double x, double y, double z, const double *vtx);
int orc_hex_inverse_map(double x, double y, double z, const double *vtx, double *rst);
void orc_hex_nodal_points(int order, const double *vtx, double *xyz);

/* ---------------- element operators (a5-a8, a10, a15, a16) ---------------- */
/* Single-element forms of the reference's tensor-element methods, on one
 * element with vertices vtx.  grad is n^d x d row-major. */
int orc_elem_gradient(int dim, int order, const double *vtx, const double *u, double *grad);
int orc_elem_grad_test_integrate(int dim, int order, const double *vtx, const double *f,
                                 double *out);
int orc_elem_test_integrate(int dim, int order, const double *vtx, const double *f,
                            double *out);
int orc_elem_delta_coeffs(int dim, int order, const double *vtx, const double *ref,
                          double *out);
int orc_elem_par_at_int_pts(int dim, int order, const double *vtx_par, double *out);
/* Element stiffness K_e u (n^d x c, column-major like the reference's
 * RealMat: out[comp*n^d + p]) for one element with npar vertex parameters. */
int orc_elem_stiffness(int dim, int order, int physics, const double *vtx,
                       const double *par, const double *u, double *ku);

/* ---------------- plex-lite: DMPlex numbering + gather map (a2, a3, a20) ---------------- */
typedef struct orc_plex orc_plex;
orc_plex *orc_plex_new(int dim, int order, int64_t E, int64_t V, const int32_t *cells);
void orc_plex_free(orc_plex *P);
/* counts[0..5] = {E, V, faces (3D, else 0), edges, points, global DOFs} */
void orc_plex_counts(const orc_plex *P, int64_t *counts);
/* conn[e*n^d + p] = global DOF of element-local node p (int32). */
void orc_plex_conn(const orc_plex *P, int32_t *conn);
/* Cone of each cell in points (E x 4 edges in 2D, E x 6 faces in 3D). */
void orc_plex_cell_cones(const orc_plex *P, int32_t *cones);
/* Cone (vertex points) of every edge (NEd x 2) and face (NF x 4). */
void orc_plex_edge_cones(const orc_plex *P, int32_t *cones);
void orc_plex_face_cones(const orc_plex *P, int32_t *cones);
/* Section offset of every point (npoints). */
void orc_plex_offsets(const orc_plex *P, int64_t *off);

/* Greedy element colouring in element order over vertex adjacency:
 * colour[e] = smallest colour unused by elements sharing a vertex with e.
 * Returns the number of colours. */
int orc_color_elements(int dim, int64_t E, int64_t V, const int32_t *cells, int32_t *colour);

/* ---------------- whole-problem driver (a1, a4, a11-a18, a21-a23) ---------------- */
typedef struct orc_sim orc_sim;
orc_sim *orc_sim_new(int dim, int order, int physics, int64_t E, int64_t V,
                     const int32_t *cells, const double *coords, const double *par,
                     int quirks);
void orc_sim_free(orc_sim *S);
const char *orc_last_error(void);
int64_t orc_sim_nglobal(const orc_sim *S);
int orc_sim_ncomp(const orc_sim *S);
void orc_sim_conn(const orc_sim *S, int32_t *conn);
/* Assembled inverse mass (G) after initialisation. */
void orc_sim_inverse_mass(const orc_sim *S, double *mi);
/* Global node coordinates (G x d) from element-local nodal points. */
void orc_sim_node_coords(const orc_sim *S, double *xyz);
/* Homogeneous Dirichlet: zero stiffness rows at masked global nodes; sources,
 * u and v are untouched (HomogeneousDirichlet.cpp:71-78). */
int orc_sim_set_dirichlet(orc_sim *S, const uint8_t *node_mask);
/* Ricker point source with direction dir[c] (Ricker.cpp:27-33). */
int orc_sim_add_source(orc_sim *S, const double *loc, double amp, double f0, double t0,
                       const double *dir);
/* Moment-tensor Ricker source (build-added; f_i = -M:grad(phi_i)(xs) R(t)). */
int orc_sim_add_moment_source(orc_sim *S, const double *loc, const double *M, double amp,
                              double f0, double t0);
int orc_sim_add_receiver(orc_sim *S, const double *loc);
/* Owning element and reference coordinates of source/receiver k. */
int orc_sim_source_info(const orc_sim *S, int k, int64_t *elem, double *ref);
int orc_sim_receiver_info(const orc_sim *S, int k, int64_t *elem, double *ref);
/* Nodal load of source k per unit source-time function (n^d x c column-major). */
int orc_sim_source_load(const orc_sim *S, int k, double *load);
/* Lagrange weights of receiver k (n^d). */
int orc_sim_receiver_weights(const orc_sim *S, int k, double *l);
/* which: 0 = u, 1 = v, 2 = a_ (previous acceleration); vals: c x G planar. */
int orc_sim_set_field(orc_sim *S, int which, const double *vals);
int orc_sim_get_field(const orc_sim *S, int which, double *vals);
/* Advance nsteps; seis (nsteps x R x c) gets receiver samples after each
 * step (may be NULL).  nthreads > 1 switches the element loop to a
 * colour-scheduled OpenMP loop (same maths, different summation order). */
int orc_sim_run(orc_sim *S, double dt, int64_t nsteps, double *seis, int nthreads);
double orc_sim_time(const orc_sim *S);
/* Assembled K u (c x G planar, Dirichlet rows zeroed), no source/mass. */
int orc_sim_apply_stiffness(orc_sim *S, const double *u, double *ku);

/* Step count and adjusted dt of the reference's loop (Options.cpp:84-97 +
 * Main.cpp:45): returns the number of iterations of while(t < duration). */
int64_t orc_reference_step_count(double duration, double dt_requested, double *dt_out);
double orc_ricker(double t, double amp, double f0, double t0);

#ifdef __cplusplus
}
#endif
