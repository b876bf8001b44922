/* Whole-problem FP64 driver for the oracle (test infrastructure): element
 * set-up, diagonal mass, source/receiver location, the per-step assembly
 * loop and the Newmark update.
 *
 * Follows Problem.cpp:34-142 (initializeElements: first element in local
 * order whose closed hull holds a source/receiver owns it), :143-215
 * (assembleIntoGlobalDof: gather, K u, source, a = f - Ku, scatter-add in
 * element order), Order2Newmark.cpp:32-66 (mass assembly + reciprocal),
 * :68-83 (a <- mi a), :85-115 (v += dt/2 (a + a_); u += dt v + dt^2/2 a;
 * a_ = a; t += dt), HomogeneousDirichlet.cpp:71-78 (zero stiffness rows),
 * Main.cpp:45-66 (step loop), Ricker.cpp:27-33.  Receiver sampling is
 * defined here (the reference locates receivers but never records them): s = sum_p l_p(xr) u_p
 * in the owning element, taken after each step's update.
 */
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif
#include "oracle.h"
#include "oracle_internal.h"

static __thread char g_err[512];
const char *orc_last_error(void) { return g_err; }
static void set_err(const char *m) { snprintf(g_err, sizeof g_err, "%s", m); }

typedef struct {
  int64_t elem;
  double ref[3];
  double *load; /* nd x c column-major nodal weights (multiplied by R(t)) */
  double amp, f0, t0;
} orc_src;

typedef struct {
  int64_t elem;
  double ref[3];
  double *l; /* nd */
} orc_rec;

struct orc_sim {
  orc_tab T;
  int physics, c, nv, npar, quirks;
  int64_t E, V, G;
  double *vtx;   /* E x nv x d */
  double *vpar;  /* E x nv x npar */
  double *cc;    /* Elastic3D constants E x nd x 10 */
  int32_t *conn; /* E x nd */
  int32_t *colour;
  int ncolour;
  int64_t *corder, *cstart; /* elements sorted by colour */
  double *mi;     /* G */
  uint8_t *dmask; /* G or NULL */
  double *u, *v, *a, *a_; /* c x G */
  double time;
  orc_src *src; int nsrc;
  orc_rec *rec; int nrec;
  double bbmin[3], bbmax[3];
  double *ebb; /* E x 2d element bounding boxes */
};

static int npar_of(int dim, int physics) {
  if (physics == ORC_ACOUSTIC) return 1;
  return dim == 2 ? 5 : 6;
}

void orc_sim_free(orc_sim *S) {
  if (!S) return;
  free(S->vtx); free(S->vpar); free(S->cc); free(S->conn); free(S->colour);
  free(S->corder); free(S->cstart); free(S->mi); free(S->dmask);
  free(S->u); free(S->v); free(S->a); free(S->a_); free(S->ebb);
  for (int i = 0; i < S->nsrc; i++) free(S->src[i].load);
  for (int i = 0; i < S->nrec; i++) free(S->rec[i].l);
  free(S->src); free(S->rec); free(S);
}

orc_sim *orc_sim_new(int dim, int order, int physics, int64_t E, int64_t V,
                     const int32_t *cells, const double *coords, const double *par,
                     int quirks) {
  orc_sim *S = (orc_sim *)calloc(1, sizeof(orc_sim));
  if (orc_tables(&S->T, dim, order)) { set_err("unsupported dim/order"); free(S); return NULL; }
  if (physics != ORC_ACOUSTIC && physics != ORC_ELASTIC) { set_err("bad physics"); free(S); return NULL; }
  S->physics = physics; S->quirks = quirks;
  S->T.quad_ref = (quirks & ORC_QUIRK_QUAD_JACOBIAN) ? 1 : 0;
  S->c = physics == ORC_ACOUSTIC ? 1 : dim;
  S->nv = dim == 2 ? 4 : 8;
  S->npar = npar_of(dim, physics);
  S->E = E; S->V = V;
  int nd = S->T.nd, nv = S->nv;
  orc_plex *P = orc_plex_new(dim, order, E, V, cells);
  int64_t cnt[6];
  orc_plex_counts(P, cnt);
  S->G = cnt[5];
  S->conn = (int32_t *)malloc(sizeof(int32_t) * E * nd);
  orc_plex_conn(P, S->conn);
  orc_plex_free(P);
  S->vtx = (double *)malloc(sizeof(double) * E * nv * dim);
  S->ebb = (double *)malloc(sizeof(double) * E * 2 * dim);
  for (int64_t e = 0; e < E; e++) {
    for (int k = 0; k < nv; k++)
      for (int a = 0; a < dim; a++) S->vtx[(e * nv + k) * dim + a] = coords[(int64_t)cells[e * nv + k] * dim + a];
    for (int a = 0; a < dim; a++) {
      double lo = 1e300, hi = -1e300;
      for (int k = 0; k < nv; k++) {
        double x = S->vtx[(e * nv + k) * dim + a];
        if (x < lo) lo = x;
        if (x > hi) hi = x;
      }
      S->ebb[e * 2 * dim + a] = lo; S->ebb[e * 2 * dim + dim + a] = hi;
    }
    double det;
    const double *vx = S->vtx + e * nv * dim;
    int bad = dim == 2 ? orc_quad_invjac(0, 0, vx, &det, 0) : orc_hex_invjac(0, 0, 0, vx, &det, 0);
    if (bad) { set_err("element with det J <= 0"); orc_sim_free(S); return NULL; }
  }
  S->vpar = (double *)malloc(sizeof(double) * E * nv * S->npar);
  memcpy(S->vpar, par, sizeof(double) * E * nv * S->npar);
  if (dim == 3 && physics == ORC_ELASTIC) {
    S->cc = (double *)malloc(sizeof(double) * E * nd * 10);
    for (int64_t e = 0; e < E; e++) orc_e3d_constants(&S->T, S->vpar + e * nv * 6, S->cc + e * nd * 10);
  }
  /* colouring for the threaded loop */
  S->colour = (int32_t *)malloc(sizeof(int32_t) * E);
  S->ncolour = orc_color_elements(dim, E, V, cells, S->colour);
  S->cstart = (int64_t *)calloc(S->ncolour + 1, sizeof(int64_t));
  S->corder = (int64_t *)malloc(sizeof(int64_t) * E);
  for (int64_t e = 0; e < E; e++) S->cstart[S->colour[e] + 1]++;
  for (int k = 0; k < S->ncolour; k++) S->cstart[k + 1] += S->cstart[k];
  int64_t *fill = (int64_t *)malloc(sizeof(int64_t) * (S->ncolour + 1));
  memcpy(fill, S->cstart, sizeof(int64_t) * (S->ncolour + 1));
  for (int64_t e = 0; e < E; e++) S->corder[fill[S->colour[e]]++] = e;
  free(fill);
  /* mass: Order2Newmark.cpp:42-52 */
  S->mi = (double *)calloc(S->G, sizeof(double));
  double *one = (double *)malloc(sizeof(double) * nd), *rho = (double *)malloc(sizeof(double) * nd),
         *me = (double *)malloc(sizeof(double) * nd);
  for (int q = 0; q < nd; q++) one[q] = 1.0;
  for (int64_t e = 0; e < E; e++) {
    const double *vx = S->vtx + e * nv * dim;
    if (physics == ORC_ACOUSTIC) {
      orc_test_integrate(&S->T, vx, one, me); /* Scalar.cpp:34-39 (the acoustic mass has no rho) */
    } else {
      orc_par_at_pts(&S->T, S->vpar + e * nv * S->npar, S->npar, rho); /* RHO is parameter 0 */
      orc_test_integrate(&S->T, vx, rho, me);
    }
    for (int q = 0; q < nd; q++) S->mi[S->conn[e * nd + q]] += me[q];
  }
  for (int64_t g = 0; g < S->G; g++) S->mi[g] = 1.0 / S->mi[g];
  free(one); free(rho); free(me);
  size_t fsz = sizeof(double) * S->c * S->G;
  S->u = (double *)calloc(1, fsz); S->v = (double *)calloc(1, fsz);
  S->a = (double *)calloc(1, fsz); S->a_ = (double *)calloc(1, fsz);
  return S;
}

int64_t orc_sim_nglobal(const orc_sim *S) { return S->G; }
int orc_sim_ncomp(const orc_sim *S) { return S->c; }
void orc_sim_conn(const orc_sim *S, int32_t *conn) {
  memcpy(conn, S->conn, sizeof(int32_t) * S->E * S->T.nd);
}
void orc_sim_inverse_mass(const orc_sim *S, double *mi) { memcpy(mi, S->mi, sizeof(double) * S->G); }
double orc_sim_time(const orc_sim *S) { return S->time; }

void orc_sim_node_coords(const orc_sim *S, double *xyz) {
  int d = S->T.dim, nd = S->T.nd;
  double *pts = (double *)malloc(sizeof(double) * nd * d);
  for (int64_t e = 0; e < S->E; e++) {
    const double *vx = S->vtx + e * S->nv * d;
    if (d == 2) orc_quad_nodal_points(S->T.order, vx, pts);
    else orc_hex_nodal_points(S->T.order, vx, pts);
    for (int q = 0; q < nd; q++)
      for (int a = 0; a < d; a++) xyz[(int64_t)S->conn[e * nd + q] * d + a] = pts[q * d + a];
  }
  free(pts);
}

int orc_sim_set_dirichlet(orc_sim *S, const uint8_t *mask) {
  if (!S->dmask) S->dmask = (uint8_t *)malloc(S->G);
  memcpy(S->dmask, mask, S->G);
  return 0;
}

/* Problem.cpp:57-113 + QuadP1/HexP1 checkHull and inverseCoordinateTransform. */
static int locate(orc_sim *S, const double *loc, int64_t *elem, double *ref) {
  int d = S->T.dim;
  for (int64_t e = 0; e < S->E; e++) {
    const double *bb = S->ebb + e * 2 * d;
    int out = 0;
    for (int a = 0; a < d; a++)
      if (loc[a] < bb[a] || loc[a] > bb[d + a]) out = 1;
    if (out) continue; /* a convex element's hull lies inside its bounding box */
    const double *vx = S->vtx + e * S->nv * d;
    int in = d == 2 ? orc_quad_check_hull(loc[0], loc[1], vx)
                    : orc_hex_check_hull(loc[0], loc[1], loc[2], vx);
    if (in < 0) { set_err("poorly behaved element"); return -1; }
    if (in) {
      int rc = d == 2 ? orc_quad_inverse_map(loc[0], loc[1], vx, ref)
                      : orc_hex_inverse_map(loc[0], loc[1], loc[2], vx, ref);
      if (rc) { set_err("inverseCoordinateTransform failed to converge"); return -1; }
      *elem = e;
      return 0;
    }
  }
  set_err("point outside the mesh");
  return -1;
}

static int add_source_common(orc_sim *S, const double *loc, orc_src *s) {
  if (locate(S, loc, &s->elem, s->ref)) return -1;
  S->src = (orc_src *)realloc(S->src, sizeof(orc_src) * (S->nsrc + 1));
  S->src[S->nsrc++] = *s;
  return S->nsrc - 1;
}

/* Ricker point source: nodal load per unit R(t) for each component. */
int orc_sim_add_source(orc_sim *S, const double *loc, double amp, double f0, double t0,
                       const double *dir) {
  orc_src s;
  memset(&s, 0, sizeof s);
  s.amp = amp; s.f0 = f0; s.t0 = t0;
  if (locate(S, loc, &s.elem, s.ref)) return -1;
  int nd = S->T.nd, d = S->T.dim;
  const double *vx = S->vtx + s.elem * S->nv * d;
  double *dc = (double *)malloc(sizeof(double) * nd), *tmp = (double *)malloc(sizeof(double) * nd);
  orc_delta_coeffs(&S->T, vx, s.ref, dc);
  s.load = (double *)malloc(sizeof(double) * nd * S->c);
  for (int comp = 0; comp < S->c; comp++) {
    for (int q = 0; q < nd; q++) tmp[q] = dc[q] * dir[comp];
    int e3d = (d == 3 && S->physics == ORC_ELASTIC);
    if (e3d && (S->quirks & ORC_QUIRK_E3D_SOURCE)) {
      memcpy(s.load + comp * nd, tmp, sizeof(double) * nd); /* Elastic3D.cpp:126-133 */
    } else {
      orc_test_integrate(&S->T, vx, tmp, s.load + comp * nd); /* Scalar.cpp:79-88 */
    }
  }
  free(dc); free(tmp);
  S->src = (orc_src *)realloc(S->src, sizeof(orc_src) * (S->nsrc + 1));
  S->src[S->nsrc++] = s;
  return S->nsrc - 1;
}

/* Moment-tensor source (defined here, Komatitsch & Tromp 1999):
 * f_{p,i} = -sum_j M_ij dphi_p/dx_j (xs) R(t). */
int orc_sim_add_moment_source(orc_sim *S, const double *loc, const double *M, double amp,
                              double f0, double t0) {
  orc_src s;
  memset(&s, 0, sizeof s);
  s.amp = amp; s.f0 = f0; s.t0 = t0;
  if (S->physics != ORC_ELASTIC) { set_err("moment tensor needs elastic physics"); return -1; }
  if (locate(S, loc, &s.elem, s.ref)) return -1;
  int nd = S->T.nd, d = S->T.dim;
  const double *vx = S->vtx + s.elem * S->nv * d;
  double *dphi = (double *)malloc(sizeof(double) * nd * d);
  orc_basis_gradient_at(&S->T, vx, s.ref, dphi);
  s.load = (double *)malloc(sizeof(double) * nd * S->c);
  for (int i = 0; i < d; i++)
    for (int q = 0; q < nd; q++) {
      double v = 0;
      for (int j = 0; j < d; j++) v += M[i * d + j] * dphi[q * d + j];
      s.load[i * nd + q] = -v;
    }
  free(dphi);
  (void)add_source_common;
  S->src = (orc_src *)realloc(S->src, sizeof(orc_src) * (S->nsrc + 1));
  S->src[S->nsrc++] = s;
  return S->nsrc - 1;
}

int orc_sim_add_receiver(orc_sim *S, const double *loc) {
  orc_rec r;
  memset(&r, 0, sizeof r);
  if (locate(S, loc, &r.elem, r.ref)) return -1;
  r.l = (double *)malloc(sizeof(double) * S->T.nd);
  orc_tensor_lagrange(&S->T, r.ref, r.l);
  S->rec = (orc_rec *)realloc(S->rec, sizeof(orc_rec) * (S->nrec + 1));
  S->rec[S->nrec++] = r;
  return S->nrec - 1;
}

int orc_sim_source_info(const orc_sim *S, int k, int64_t *elem, double *ref) {
  if (k < 0 || k >= S->nsrc) return -1;
  *elem = S->src[k].elem;
  memcpy(ref, S->src[k].ref, sizeof(double) * S->T.dim);
  return 0;
}
int orc_sim_receiver_info(const orc_sim *S, int k, int64_t *elem, double *ref) {
  if (k < 0 || k >= S->nrec) return -1;
  *elem = S->rec[k].elem;
  memcpy(ref, S->rec[k].ref, sizeof(double) * S->T.dim);
  return 0;
}

int orc_sim_source_load(const orc_sim *S, int k, double *load) {
  if (k < 0 || k >= S->nsrc) return -1;
  memcpy(load, S->src[k].load, sizeof(double) * S->T.nd * S->c);
  return 0;
}
int orc_sim_receiver_weights(const orc_sim *S, int k, double *l) {
  if (k < 0 || k >= S->nrec) return -1;
  memcpy(l, S->rec[k].l, sizeof(double) * S->T.nd);
  return 0;
}

static double *field_ptr(const orc_sim *S, int which) {
  return which == 0 ? S->u : which == 1 ? S->v : which == 2 ? S->a_ : which == 3 ? S->a : NULL;
}
int orc_sim_set_field(orc_sim *S, int which, const double *vals) {
  double *f = field_ptr(S, which);
  if (!f) return -1;
  memcpy(f, vals, sizeof(double) * S->c * S->G);
  return 0;
}
int orc_sim_get_field(const orc_sim *S, int which, double *vals) {
  double *f = field_ptr(S, which);
  if (!f) return -1;
  memcpy(vals, f, sizeof(double) * S->c * S->G);
  return 0;
}

double orc_ricker(double t, double amp, double f0, double t0) {
  const double factor = M_PI * M_PI * f0 * f0 * (t - t0) * (t - t0);
  return amp * ((1 - 2 * factor) * exp(-1 * factor));
}

/* One element's contribution a_loc = f - K u (Problem.cpp:183-200). */
static void element_accel(orc_sim *S, int64_t e, double t, const double *u, int with_source,
                          double *uloc, double *aloc, double *work) {
  int nd = S->T.nd, c = S->c, d = S->T.dim;
  const int32_t *cn = S->conn + e * nd;
  for (int comp = 0; comp < c; comp++)
    for (int q = 0; q < nd; q++) uloc[comp * nd + q] = u[(int64_t)comp * S->G + cn[q]];
  const double *cc = S->cc ? S->cc + e * nd * 10 : NULL;
  orc_stiffness(&S->T, S->physics, S->vtx + e * S->nv * d, S->vpar + e * S->nv * S->npar, cc,
                uloc, aloc, work);
  if (S->dmask) /* HomogeneousDirichlet.cpp:71-78 */
    for (int comp = 0; comp < c; comp++)
      for (int q = 0; q < nd; q++)
        if (S->dmask[cn[q]]) aloc[comp * nd + q] = 0.0;
  for (int k = 0; k < c * nd; k++) aloc[k] = -aloc[k];
  if (with_source)
    for (int s = 0; s < S->nsrc; s++) {
      if (S->src[s].elem != e) continue;
      double R = orc_ricker(t, S->src[s].amp, S->src[s].f0, S->src[s].t0);
      for (int k = 0; k < c * nd; k++) aloc[k] = S->src[s].load[k] * R + aloc[k];
    }
}

static void assemble(orc_sim *S, double t, const double *u, double *a, int with_source,
                     int nthreads) {
  int nd = S->T.nd, c = S->c;
  memset(a, 0, sizeof(double) * c * S->G);
  if (nthreads <= 1) {
    double *uloc = (double *)malloc(sizeof(double) * 3 * nd), *aloc = (double *)malloc(sizeof(double) * 3 * nd);
    double *work = (double *)malloc(sizeof(double) * ORC_WORK * ORC_NDMAX);
    for (int64_t e = 0; e < S->E; e++) {
      element_accel(S, e, t, u, with_source, uloc, aloc, work);
      const int32_t *cn = S->conn + e * nd;
      for (int comp = 0; comp < c; comp++)
        for (int q = 0; q < nd; q++) a[(int64_t)comp * S->G + cn[q]] += aloc[comp * nd + q];
    }
    free(uloc); free(aloc); free(work);
    return;
  }
#ifdef _OPENMP
  for (int col = 0; col < S->ncolour; col++) {
#pragma omp parallel num_threads(nthreads)
    {
      double *uloc = (double *)malloc(sizeof(double) * 3 * nd), *aloc = (double *)malloc(sizeof(double) * 3 * nd);
      double *work = (double *)malloc(sizeof(double) * ORC_WORK * ORC_NDMAX);
#pragma omp for schedule(static)
      for (int64_t i = S->cstart[col]; i < S->cstart[col + 1]; i++) {
        int64_t e = S->corder[i];
        element_accel(S, e, t, u, with_source, uloc, aloc, work);
        const int32_t *cn = S->conn + e * nd;
        for (int comp = 0; comp < c; comp++)
          for (int q = 0; q < nd; q++) a[(int64_t)comp * S->G + cn[q]] += aloc[comp * nd + q];
      }
      free(uloc); free(aloc); free(work);
    }
  }
#else
  (void)nthreads;
  assemble(S, t, u, a, with_source, 1);
#endif
}

int orc_sim_apply_stiffness(orc_sim *S, const double *u, double *ku) {
  assemble(S, 0.0, u, ku, 0, 1);
  for (int64_t k = 0; k < (int64_t)S->c * S->G; k++) ku[k] = -ku[k];
  return 0;
}

int orc_sim_run(orc_sim *S, double dt, int64_t nsteps, double *seis, int nthreads) {
  int c = S->c, nd = S->T.nd;
  int64_t G = S->G;
  const double acl = (1.0 / 2.0) * dt, dsp = (1.0 / 2.0) * (dt * dt);
  for (int64_t step = 0; step < nsteps; step++) {
    assemble(S, S->time, S->u, S->a, 1, nthreads);
    /* Order2Newmark.cpp:68-83 */
    for (int comp = 0; comp < c; comp++)
      for (int64_t g = 0; g < G; g++) S->a[comp * G + g] = S->mi[g] * S->a[comp * G + g];
    /* Order2Newmark.cpp:85-115 */
    /* With the quirk, the reference's literal field list {ax, ay, ax, a}
     * (Order2Newmark.cpp:93-96), each entry taken when the field exists:
     * elastic meshes (2-D and 3-D) advance x twice, y once and z never;
     * acoustic meshes advance "a" once.  Without it every component once. */
    int ncomp_list = c;
    int list[4] = {0, 1, 2, -1};
    if (c >= 2 && (S->quirks & ORC_QUIRK_NEWMARK3D)) {
      const int lit[3] = {0, 1, 0};
      memcpy(list, lit, sizeof lit);
      ncomp_list = 3;
    }
    for (int li = 0; li < ncomp_list; li++) {
      int comp = list[li];
      double *u = S->u + comp * G, *v = S->v + comp * G, *a = S->a + comp * G, *ap = S->a_ + comp * G;
      for (int64_t g = 0; g < G; g++) {
        v[g] += acl * a[g] + acl * ap[g];
        u[g] += dt * v[g] + dsp * a[g];
        ap[g] = a[g];
      }
    }
    S->time += dt;
    if (seis) {
      for (int r = 0; r < S->nrec; r++) {
        const int32_t *cn = S->conn + S->rec[r].elem * nd;
        for (int comp = 0; comp < c; comp++) {
          double acc = 0;
          for (int q = 0; q < nd; q++) acc += S->rec[r].l[q] * S->u[comp * G + cn[q]];
          seis[(step * S->nrec + r) * c + comp] = acc;
        }
      }
    }
  }
  return 0;
}

int64_t orc_reference_step_count(double duration, double dt_requested, double *dt_out) {
  int64_t nts = (int64_t)ceil(duration / dt_requested);
  double dt = duration / (double)nts;
  if (dt_out) *dt_out = dt;
  double t = 0;
  int64_t n = 0;
  while (t < duration) { t += dt; n++; }
  return n;
}
