/* plex-lite: a PETSc-free restatement of the DMPlex numbering, section and
 * spectral closure that the reference's gather/scatter rides on (test
 * infrastructure).
 *
 * Follows Mesh.cpp:31-51 (DMPlexCreateFromCellList with interpolation),
 * Mesh.cpp:361-452 (DMPlexCreateSection with 1 DOF per vertex, N-1 per edge,
 * (N-1)^2 per face/2D cell, (N-1)^3 per 3D cell; FixOrientation; spectral
 * closure) and Problem.cpp:302-320 (closure gather with identity ClsMap).
 * PETSc itself (branch maint + knepley/fix-plex-spectral-order,
 * conf/install-deps.sh:13-17) is not present; its published interpolation
 * algorithm is restated:
 *   points are cells [0,E), vertices [E,E+V), then (3D) faces, then edges;
 *   faces/edges are numbered in creation order, walking cells in order and
 *   each cell's raw faces (hex: bottom 0123, top 4567, front 0354, back
 *   2176, right 3265, left 0471) or raw edges (quad: 01,12,23,30); in 3D the
 *   edges are created from the faces in face order (face edges c0c1, c1c2,
 *   c2c3, c3c0).  A new face/edge keeps the creating cone's vertex order.
 *   Section offsets follow point order.
 * DOFs inside an edge run from its cone[0] to cone[1]; DOFs inside a face
 * are (i,j) with i along cone[0]->cone[1], j along cone[0]->cone[3]; cell
 * interior DOFs are tensor-lexicographic.  The gather map of a cell node is
 * found geometrically in index space, which is what FixOrientation's
 * orientation permutations (Mesh.cpp:90-278) achieve: every element sees a
 * shared node at the same global DOF.
 */
#include <stdlib.h>
#include <string.h>
#include "oracle.h"

struct orc_plex {
  int dim, order;
  int64_t E, V, NF, NEd, npoints, nglobal;
  int32_t *cells;     /* E x nv (copy) */
  int32_t *cell_cone; /* E x (4|6) point ids */
  int32_t *face_vtx;  /* NF x 4 vertex indices */
  int32_t *edge_vtx;  /* NEd x 2 vertex indices */
  int64_t *offset;    /* npoints */
  int32_t *conn;      /* E x n^d */
};

/* ---- open-addressing hash on 128-bit keys ---- */
typedef struct {
  uint64_t *k0, *k1;
  int32_t *val;
  int64_t cap;
} hmap;

static void hm_init(hmap *h, int64_t n) {
  int64_t cap = 16;
  while (cap < 2 * n + 16) cap <<= 1;
  h->cap = cap;
  h->k0 = (uint64_t *)malloc(sizeof(uint64_t) * cap);
  h->k1 = (uint64_t *)malloc(sizeof(uint64_t) * cap);
  h->val = (int32_t *)malloc(sizeof(int32_t) * cap);
  for (int64_t i = 0; i < cap; i++) h->val[i] = -1;
}
static void hm_free(hmap *h) { free(h->k0); free(h->k1); free(h->val); }
static uint64_t mix(uint64_t x) {
  x ^= x >> 33; x *= 0xff51afd7ed558ccdULL; x ^= x >> 33; x *= 0xc4ceb9fe1a85ec53ULL;
  x ^= x >> 33;
  return x;
}
/* returns existing value or inserts v and returns -1 */
static int32_t hm_get_or_put(hmap *h, uint64_t a, uint64_t b, int32_t v) {
  int64_t m = h->cap - 1, i = (int64_t)(mix(a * 0x9e3779b97f4a7c15ULL ^ b) & (uint64_t)m);
  for (;;) {
    if (h->val[i] < 0) { h->k0[i] = a; h->k1[i] = b; h->val[i] = v; return -1; }
    if (h->k0[i] == a && h->k1[i] == b) return h->val[i];
    i = (i + 1) & m;
  }
}
static int32_t hm_get(const hmap *h, uint64_t a, uint64_t b) {
  int64_t m = h->cap - 1, i = (int64_t)(mix(a * 0x9e3779b97f4a7c15ULL ^ b) & (uint64_t)m);
  for (;;) {
    if (h->val[i] < 0) return -1;
    if (h->k0[i] == a && h->k1[i] == b) return h->val[i];
    i = (i + 1) & m;
  }
}
static void edge_key(int32_t a, int32_t b, uint64_t *k0, uint64_t *k1) {
  uint32_t lo = (uint32_t)(a < b ? a : b), hi = (uint32_t)(a < b ? b : a);
  *k0 = ((uint64_t)lo << 32) | hi;
  *k1 = 0;
}
static void face_key(const int32_t *v, uint64_t *k0, uint64_t *k1) {
  uint32_t s[4] = {(uint32_t)v[0], (uint32_t)v[1], (uint32_t)v[2], (uint32_t)v[3]};
  for (int i = 1; i < 4; i++)
    for (int j = i; j > 0 && s[j - 1] > s[j]; j--) { uint32_t t = s[j]; s[j] = s[j - 1]; s[j - 1] = t; }
  *k0 = ((uint64_t)s[0] << 32) | s[1];
  *k1 = ((uint64_t)s[2] << 32) | s[3];
}

static const int hex_faces[6][4] = {{0, 1, 2, 3}, {4, 5, 6, 7}, {0, 3, 5, 4},
                                    {2, 1, 7, 6}, {3, 2, 6, 5}, {0, 4, 7, 1}};
/* hex vertex v -> index-space corner (0 or 1 per axis r,s,t) */
static const int hex_corner[8][3] = {{0, 0, 0}, {0, 1, 0}, {1, 1, 0}, {1, 0, 0},
                                     {0, 0, 1}, {1, 0, 1}, {1, 1, 1}, {0, 1, 1}};
static const int quad_corner[4][2] = {{0, 0}, {1, 0}, {1, 1}, {0, 1}};

static int local_vertex(const int32_t *cv, int nv, int32_t g) {
  for (int v = 0; v < nv; v++)
    if (cv[v] == g) return v;
  return -1;
}

orc_plex *orc_plex_new(int dim, int order, int64_t E, int64_t V, const int32_t *cells) {
  if ((dim != 2 && dim != 3) || order < 1) return NULL;
  orc_plex *P = (orc_plex *)calloc(1, sizeof(orc_plex));
  int nv = dim == 2 ? 4 : 8, N = order, n = N + 1;
  P->dim = dim; P->order = order; P->E = E; P->V = V;
  P->cells = (int32_t *)malloc(sizeof(int32_t) * E * nv);
  memcpy(P->cells, cells, sizeof(int32_t) * E * nv);
  hmap he;
  hm_init(&he, dim == 2 ? 2 * E + 8 : 3 * E + 8);
  if (dim == 2) {
    P->cell_cone = (int32_t *)malloc(sizeof(int32_t) * E * 4);
    P->edge_vtx = (int32_t *)malloc(sizeof(int32_t) * (E * 4) * 2);
    int64_t ne = 0;
    for (int64_t c = 0; c < E; c++) {
      const int32_t *cv = cells + c * 4;
      for (int k = 0; k < 4; k++) {
        int32_t a = cv[k], b = cv[(k + 1) % 4];
        uint64_t k0, k1;
        edge_key(a, b, &k0, &k1);
        int32_t id = hm_get_or_put(&he, k0, k1, (int32_t)ne);
        if (id < 0) { id = (int32_t)ne; P->edge_vtx[2 * ne] = a; P->edge_vtx[2 * ne + 1] = b; ne++; }
        P->cell_cone[c * 4 + k] = (int32_t)(E + V + id);
      }
    }
    P->NEd = ne; P->NF = 0;
  } else {
    hmap hf;
    hm_init(&hf, 3 * E + 8);
    P->cell_cone = (int32_t *)malloc(sizeof(int32_t) * E * 6);
    P->face_vtx = (int32_t *)malloc(sizeof(int32_t) * (E * 6) * 4);
    int64_t nf = 0;
    for (int64_t c = 0; c < E; c++) {
      const int32_t *cv = cells + c * 8;
      for (int f = 0; f < 6; f++) {
        int32_t fv[4];
        for (int q = 0; q < 4; q++) fv[q] = cv[hex_faces[f][q]];
        uint64_t k0, k1;
        face_key(fv, &k0, &k1);
        int32_t id = hm_get_or_put(&hf, k0, k1, (int32_t)nf);
        if (id < 0) { id = (int32_t)nf; memcpy(P->face_vtx + 4 * nf, fv, sizeof fv); nf++; }
        P->cell_cone[c * 6 + f] = (int32_t)(E + V + id);
      }
    }
    hm_free(&hf);
    P->NF = nf;
    P->edge_vtx = (int32_t *)malloc(sizeof(int32_t) * (nf * 4) * 2);
    int64_t ne = 0;
    for (int64_t f = 0; f < nf; f++) {
      const int32_t *fv = P->face_vtx + 4 * f;
      for (int k = 0; k < 4; k++) {
        int32_t a = fv[k], b = fv[(k + 1) % 4];
        uint64_t k0, k1;
        edge_key(a, b, &k0, &k1);
        if (hm_get_or_put(&he, k0, k1, (int32_t)ne) < 0) {
          P->edge_vtx[2 * ne] = a; P->edge_vtx[2 * ne + 1] = b; ne++;
        }
      }
    }
    P->NEd = ne;
  }
  P->npoints = E + V + P->NF + P->NEd;
  /* section offsets in point order */
  P->offset = (int64_t *)malloc(sizeof(int64_t) * P->npoints);
  int64_t off = 0, p = 0;
  int64_t dcell = dim == 2 ? (int64_t)(N - 1) * (N - 1) : (int64_t)(N - 1) * (N - 1) * (N - 1);
  for (int64_t c = 0; c < E; c++, p++) { P->offset[p] = off; off += dcell; }
  for (int64_t v = 0; v < V; v++, p++) { P->offset[p] = off; off += 1; }
  for (int64_t f = 0; f < P->NF; f++, p++) { P->offset[p] = off; off += (int64_t)(N - 1) * (N - 1); }
  for (int64_t e = 0; e < P->NEd; e++, p++) { P->offset[p] = off; off += N - 1; }
  P->nglobal = off;

  /* closure gather map */
  int64_t nd = dim == 2 ? n * n : n * n * n;
  P->conn = (int32_t *)malloc(sizeof(int32_t) * E * nd);
  int64_t ebase = E + V + P->NF; /* first edge point */
  for (int64_t c = 0; c < E; c++) {
    const int32_t *cv = cells + c * nv;
    for (int64_t q = 0; q < nd; q++) {
      int Pc[3] = {(int)(q % n), (int)((q / n) % n), dim == 3 ? (int)(q / (n * n)) : 0};
      int onb[3], nb = 0;
      for (int a = 0; a < dim; a++) { onb[a] = (Pc[a] == 0 || Pc[a] == N); nb += onb[a]; }
      int64_t g;
      if (nb == 0) {
        int64_t dof = dim == 2 ? (Pc[0] - 1) + (int64_t)(N - 1) * (Pc[1] - 1)
                               : (Pc[0] - 1) + (int64_t)(N - 1) * ((Pc[1] - 1) + (int64_t)(N - 1) * (Pc[2] - 1));
        g = P->offset[c] + dof;
      } else if (nb == dim) {
        int lv = -1;
        for (int v = 0; v < nv; v++) {
          int ok = 1;
          for (int a = 0; a < dim; a++) {
            int cc = dim == 2 ? quad_corner[v][a] : hex_corner[v][a];
            if (cc * N != Pc[a]) ok = 0;
          }
          if (ok) { lv = v; break; }
        }
        g = P->offset[E + cv[lv]];
      } else if (nb == dim - 1) {
        /* edge node: free axis fa; endpoints are P with Pc[fa] = 0 and N */
        int fa = 0;
        for (int a = 0; a < dim; a++) if (!onb[a]) fa = a;
        int la = -1, lb = -1;
        for (int v = 0; v < nv; v++) {
          int ok0 = 1, ok1 = 1;
          for (int a = 0; a < dim; a++) {
            int cc = (dim == 2 ? quad_corner[v][a] : hex_corner[v][a]) * N;
            int want0 = a == fa ? 0 : Pc[a], want1 = a == fa ? N : Pc[a];
            if (cc != want0) ok0 = 0;
            if (cc != want1) ok1 = 0;
          }
          if (ok0) la = v;
          if (ok1) lb = v;
        }
        uint64_t k0, k1;
        edge_key(cv[la], cv[lb], &k0, &k1);
        int32_t eid = hm_get(&he, k0, k1);
        int32_t ea = P->edge_vtx[2 * eid];
        int t = (ea == cv[la]) ? Pc[fa] : N - Pc[fa];
        g = P->offset[ebase + eid] + (t - 1);
      } else {
        /* 3D face node: fixed axis xa at Pc[xa] in {0,N} */
        int xa = 0;
        for (int a = 0; a < 3; a++) if (onb[a]) xa = a;
        int side = Pc[xa] == N;
        int fidx;
        if (xa == 2) fidx = side ? 1 : 0;
        else if (xa == 1) fidx = side ? 3 : 2;
        else fidx = side ? 4 : 5;
        int32_t fpt = P->cell_cone[c * 6 + fidx];
        const int32_t *fv = P->face_vtx + 4 * (fpt - (E + V));
        int l0 = local_vertex(cv, 8, fv[0]), l1 = local_vertex(cv, 8, fv[1]),
            l3 = local_vertex(cv, 8, fv[3]);
        int ia = 0, ja = 0;
        for (int a = 0; a < 3; a++) {
          if (hex_corner[l1][a] != hex_corner[l0][a]) ia = a;
          if (hex_corner[l3][a] != hex_corner[l0][a]) ja = a;
        }
        int ii = hex_corner[l0][ia] ? N - Pc[ia] : Pc[ia];
        int jj = hex_corner[l0][ja] ? N - Pc[ja] : Pc[ja];
        g = P->offset[fpt] + (ii - 1) + (int64_t)(N - 1) * (jj - 1);
      }
      P->conn[c * nd + q] = (int32_t)g;
    }
  }
  hm_free(&he);
  return P;
}

void orc_plex_free(orc_plex *P) {
  if (!P) return;
  free(P->cells); free(P->cell_cone); free(P->face_vtx); free(P->edge_vtx);
  free(P->offset); free(P->conn); free(P);
}

void orc_plex_counts(const orc_plex *P, int64_t *counts) {
  counts[0] = P->E; counts[1] = P->V; counts[2] = P->NF; counts[3] = P->NEd;
  counts[4] = P->npoints; counts[5] = P->nglobal;
}
void orc_plex_conn(const orc_plex *P, int32_t *conn) {
  int n = P->order + 1;
  int64_t nd = P->dim == 2 ? n * n : n * n * n;
  memcpy(conn, P->conn, sizeof(int32_t) * P->E * nd);
}
void orc_plex_cell_cones(const orc_plex *P, int32_t *cones) {
  memcpy(cones, P->cell_cone, sizeof(int32_t) * P->E * (P->dim == 2 ? 4 : 6));
}
void orc_plex_edge_cones(const orc_plex *P, int32_t *cones) {
  for (int64_t i = 0; i < 2 * P->NEd; i++) cones[i] = (int32_t)(P->edge_vtx[i] + P->E);
}
void orc_plex_face_cones(const orc_plex *P, int32_t *cones) {
  for (int64_t i = 0; i < 4 * P->NF; i++) cones[i] = (int32_t)(P->face_vtx[i] + P->E);
}
void orc_plex_offsets(const orc_plex *P, int64_t *off) {
  memcpy(off, P->offset, sizeof(int64_t) * P->npoints);
}

/* Greedy colouring (defined here; the reference has no colouring). */
int orc_color_elements(int dim, int64_t E, int64_t V, const int32_t *cells, int32_t *colour) {
  int nv = dim == 2 ? 4 : 8;
  int64_t *start = (int64_t *)calloc(V + 1, sizeof(int64_t));
  for (int64_t i = 0; i < E * nv; i++) start[cells[i] + 1]++;
  for (int64_t v = 0; v < V; v++) start[v + 1] += start[v];
  int64_t *fill = (int64_t *)malloc(sizeof(int64_t) * (V + 1));
  memcpy(fill, start, sizeof(int64_t) * (V + 1));
  int64_t *adj = (int64_t *)malloc(sizeof(int64_t) * E * nv);
  for (int64_t e = 0; e < E; e++)
    for (int k = 0; k < nv; k++) adj[fill[cells[e * nv + k]]++] = e;
  int ncol = 0;
  unsigned char used[1024];
  for (int64_t e = 0; e < E; e++) {
    memset(used, 0, sizeof used);
    for (int k = 0; k < nv; k++) {
      int32_t v = cells[e * nv + k];
      for (int64_t q = start[v]; q < start[v + 1]; q++) {
        int64_t o = adj[q];
        if (o < e) used[colour[o]] = 1;
      }
    }
    int c = 0;
    while (used[c]) c++;
    colour[e] = c;
    if (c + 1 > ncol) ncol = c + 1;
  }
  free(start); free(fill); free(adj);
  return ncol;
}
