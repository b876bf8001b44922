// C ABI of the MI355X SEM path: meshes, shots, batches and the two step
// paths (plain separate kernels; fused colour kernels in hipGraph blocks).
//
// Host-side responsibilities mirror the reference's set-up code:
//   sem_mesh_create     Mesh::read + setupGlobalDof (Mesh.cpp:31-51, :361-452),
//                       attachVertexCoordinates / attachMaterialProperties,
//                       Order2Newmark::initializeGlobalDofs mass (Order2Newmark.cpp:32-66)
//   sem_shot_add_*      Problem::initializeElements source/receiver attach
//                       (Problem.cpp:34-142; QuadP1/HexP1 checkHull + Newton)
//   sem_batch_run       the time loop (Main.cpp:45-66)
#include <hip/hip_runtime.h>

#include <algorithm>
#include <array>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <map>
#include <mutex>
#include <numeric>
#include <string>
#include <tuple>
#include <vector>

#include "batch.h"

namespace sem {

static thread_local std::string g_error;
void set_error(const std::string &msg) { g_error = msg; }
int hip_check(hipError_t e, const char *what) {
  if (e == hipSuccess) return 0;
  set_error(std::string(what) + ": " + hipGetErrorString(e));
  return -2;
}

}  // namespace sem

using namespace sem;

// ====================================================================== helpers
namespace sem {
int set_device(int device) {
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess || n == 0) {
    set_error("no HIP device visible");
    return -4;
  }
  if (device < 0 || device >= n) {
    set_error("device index out of range");
    return -1;
  }
  HIPCHK(hipSetDevice(device));
  SEM_TRY(init_device_tables(device));
  return init_tile_tables(device);
}
}  // namespace sem

static bool is_affine(int dim, const double *v) {
  // All bilinear/trilinear cross terms of the P1 map vanish.
  double scale = 0;
  const int nv = dim == 2 ? 4 : 8;
  for (int i = 0; i < nv * dim; ++i) scale = std::max(scale, std::fabs(v[i]));
  const double tol = 1e-12 * (scale + 1.0);
  if (dim == 2) {
    for (int a = 0; a < 2; ++a)
      if (std::fabs(v[0 * 2 + a] - v[1 * 2 + a] + v[2 * 2 + a] - v[3 * 2 + a]) > tol) return false;
    return true;
  }
  static const double s[8][3] = {{-1, -1, -1}, {-1, 1, -1}, {1, 1, -1}, {1, -1, -1},
                                 {-1, -1, 1},  {1, -1, 1},  {1, 1, 1},  {-1, 1, 1}};
  for (int a = 0; a < 3; ++a) {
    double rs = 0, rt = 0, st = 0, rst = 0;
    for (int k = 0; k < 8; ++k) {
      double x = v[k * 3 + a];
      rs += s[k][0] * s[k][1] * x;
      rt += s[k][0] * s[k][2] * x;
      st += s[k][1] * s[k][2] * x;
      rst += s[k][0] * s[k][1] * s[k][2] * x;
    }
    if (std::fabs(rs) > tol || std::fabs(rt) > tol || std::fabs(st) > tol || std::fabs(rst) > tol)
      return false;
  }
  return true;
}

static int upload_conn(sem_mesh *m) {
  // sorted order, with first/last/dirichlet flags
  const int64_t E = m->E, nd = m->nd;
  std::vector<int32_t> cmin(m->G, 1 << 30), cmax(m->G, -1);
  for (int64_t e = 0; e < E; ++e) {
    int32_t c = m->colour[e];
    for (int64_t p = 0; p < nd; ++p) {
      int32_t g = m->conn[e * nd + p];
      cmin[g] = std::min(cmin[g], c);
      cmax[g] = std::max(cmax[g], c);
    }
  }
  std::vector<uint32_t> words(E * nd);
  for (int64_t sp = 0; sp < E; ++sp) {
    int64_t e = m->perm[sp];
    int32_t c = m->colour[e];
    for (int64_t p = 0; p < nd; ++p) {
      int32_t g = m->conn[e * nd + p];
      uint32_t w = (uint32_t)g;
      if (cmin[g] == c) w |= kFlagFirst;
      if (cmax[g] == c) w |= kFlagLast;
      if (!m->dirichlet.empty() && m->dirichlet[g]) w |= kFlagDirichlet;
      words[sp * nd + p] = w;
    }
  }
  if (!m->d_conn) SEM_TRY(dalloc(&m->d_conn, E * nd, &m->bytes));
  HIPCHK(hipMemcpyAsync(m->d_conn, words.data(), words.size() * 4, hipMemcpyHostToDevice, m->stream));
  HIPCHK(hipStreamSynchronize(m->stream));
  return 0;
}

int sem::ensure_pointwise(sem_mesh *m) {
  if (m->d_geo_p) return 0;
  const int gm = pointwise_mode(m);
  const int ngp = point_geo_count(m->dim, m->physics, gm);
  SEM_TRY(dalloc(&m->d_geo_p, (size_t)m->E * ngp * m->nd, &m->bytes));
  SEM_TRY(launch_point_geometry(m->dim, m->order, m->physics, gm, m->E, m->d_vtx, m->d_vpar,
                                ref_npar(m->dim, m->physics), m->d_geo_p, nullptr, m->stream));
  HIPCHK(hipStreamSynchronize(m->stream));
  return 0;
}

// Tile shape: the first compiled shape of the preference list for this
// dimension/physics, or exactly SEM_TILE=x,y[,z] (no tiles if that shape has
// no kernel for the mesh's order).  Returns false when no shape applies.
static bool pick_tile(int dim, int order, int physics, int *T) {
  const char *env = getenv("SEM_TILE");
  if (env && *env) {
    int a = 0, b = 0, c = 1;
    if (sscanf(env, "%d,%d,%d", &a, &b, &c) < 2 || a <= 0 || b <= 0 || c <= 0) return false;
    T[0] = a; T[1] = b; T[2] = dim == 3 ? c : 1;
    return tile_kernel_info(dim, order, physics, T).lds_bytes > 0;
  }
  static const int pref2[][3] = {{16, 8, 1}, {8, 8, 1}};
  static const int pref3a[][3] = {{4, 4, 2}, {4, 2, 2}};
  static const int pref3e[][3] = {{4, 2, 2}, {2, 2, 2}};
  const int(*list)[3] = dim == 2 ? pref2 : (physics == SEM_ACOUSTIC ? pref3a : pref3e);
  const int cnt = 2;
  for (int i = 0; i < cnt; ++i) {
    T[0] = list[i][0]; T[1] = list[i][1]; T[2] = list[i][2];
    const TileKernelInfo k = tile_kernel_info(dim, order, physics, T);
    if (k.lds_bytes > 0 && k.lds_bytes <= 160 * 1024) return true;
  }
  return false;
}

static int upload_tile_words(sem_mesh *m) {
  const TileSet &ts = m->tiles;
  if (!m->d_tile_words) SEM_TRY(dalloc(&m->d_tile_words, ts.words.size(), &m->tile_bytes));
  if (!m->d_tile_slot) SEM_TRY(dalloc(&m->d_tile_slot, ts.slot_of.size(), &m->tile_bytes));
  if (!m->d_tile_words_lat) SEM_TRY(dalloc(&m->d_tile_words_lat, ts.words_lat.size(), &m->tile_bytes));
  HIPCHK(hipMemcpyAsync(m->d_tile_words_lat, ts.words_lat.data(), ts.words_lat.size() * 4, hipMemcpyHostToDevice, m->stream));
  HIPCHK(hipMemcpyAsync(m->d_tile_words, ts.words.data(), ts.words.size() * 4, hipMemcpyHostToDevice, m->stream));
  HIPCHK(hipMemcpyAsync(m->d_tile_slot, ts.slot_of.data(), ts.slot_of.size() * 2, hipMemcpyHostToDevice, m->stream));
  HIPCHK(hipStreamSynchronize(m->stream));
  return 0;
}

void sem::free_tiles(sem_mesh *m) {
  dfree(m->d_tile_info);
  dfree(m->d_tile_words);
  dfree(m->d_tile_slot);
  dfree(m->d_tile_words_lat);
  dfree(m->d_telem);
  dfree(m->d_mi_int);
  m->bytes -= std::min(m->bytes, m->tile_bytes);
  m->tile_bytes = 0;
  m->tiles = TileSet();
}

// Builds and uploads the tile path of an affine, logically Cartesian mesh,
// with the given tile shape or (shape null) the first of the preference list.
int sem::setup_tiles(sem_mesh *m, const int *shape) {
  free_tiles(m);
  const char *off = getenv("SEM_NO_TILES");
  if (!m->affine || (off && off[0] == '1')) return 0;
  int T[3];
  if (shape) {
    T[0] = shape[0]; T[1] = shape[1]; T[2] = m->dim == 3 ? shape[2] : 1;
    if (tile_kernel_info(m->dim, m->order, m->physics, T).lds_bytes == 0) {
      set_error("tile shape not compiled for this order and physics");
      return -1;
    }
  } else if (!pick_tile(m->dim, m->order, m->physics, T)) {
    return 0;
  }
  if (build_tiles(m, T[0], T[1], T[2], m->tiles) != 0) {
    m->tiles = TileSet();
    return 0;  // not eligible: the element kernel is used
  }
  const int rc = upload_tiles(m);
  m->bytes += m->tile_bytes;
  return rc;
}

int sem::upload_tiles(sem_mesh *m) {
  TileSet &ts = m->tiles;
  const int dim = m->dim, nv = m->nv, npar = ref_npar(dim, m->physics), kp = kernel_npar(dim, m->physics);
  // padded tile element order (padding repeats element 0; never read)
  const int64_t E = (int64_t)ts.elem_order.size();
  SEM_TRY(dalloc(&m->d_tile_info, ts.info.size(), &m->tile_bytes));
  HIPCHK(hipMemcpyAsync(m->d_tile_info, ts.info.data(), ts.info.size() * sizeof(TileInfo), hipMemcpyHostToDevice, m->stream));
  SEM_TRY(upload_tile_words(m));
  // FP64 vertices in tile element order -> affine factors on the device
  std::vector<double> tv(E * nv * dim);
  std::vector<float> tp(E * nv * kp);
  for (int64_t i = 0; i < E; ++i) {
    const int64_t e = std::max<int32_t>(ts.elem_order[i], 0);
    std::copy_n(&m->vtx[e * nv * dim], nv * dim, &tv[i * nv * dim]);
    for (int k = 0; k < nv; ++k) {
      const double *src = &m->vpar[(e * nv + k) * npar];
      float *dst = &tp[(i * nv + k) * kp];
      if (m->physics == SEM_ACOUSTIC) dst[0] = (float)src[0];
      else if (dim == 2) for (int q = 0; q < 4; ++q) dst[q] = (float)src[1 + q];
      else for (int q = 0; q < 6; ++q) dst[q] = (float)src[q];
    }
  }
  // affine factors (FP64 on the device, FP32 out), then one array of
  // [affine factors | vertex material] per element for the tile kernel
  const int ng = affine_geo_count(dim, m->physics), ned = ng + nv * kp;
  double *d_tv = nullptr;
  float *d_tg = nullptr;
  std::vector<float> tg(E * ng), te(E * ned);
  SEM_TRY(dalloc(&d_tv, tv.size()));
  int rc = hip_check(hipMemcpyAsync(d_tv, tv.data(), tv.size() * 8, hipMemcpyHostToDevice, m->stream), "upload");
  if (!rc) rc = dalloc(&d_tg, tg.size());
  if (!rc) rc = launch_affine_geometry(dim, m->physics, E, d_tv, d_tg, m->stream);
  if (!rc) rc = hip_check(hipMemcpyAsync(tg.data(), d_tg, tg.size() * 4, hipMemcpyDeviceToHost, m->stream), "download");
  if (!rc) rc = hip_check(hipStreamSynchronize(m->stream), "sync");
  if (!rc) {
    // off-diagonal entries: J^-1 row-major (elastic), packed symmetric G (acoustic)
    static const int off_e2[] = {1, 2}, off_e3[] = {1, 2, 3, 5, 6, 7}, off_a2[] = {1}, off_a3[] = {1, 2, 4};
    const int *off = m->physics == SEM_ACOUSTIC ? (dim == 2 ? off_a2 : off_a3) : (dim == 2 ? off_e2 : off_e3);
    const int noff = m->physics == SEM_ACOUSTIC ? (dim == 2 ? 1 : 3) : (dim == 2 ? 2 : 6);
    bool diag = true;
    for (int64_t i = 0; i < E && diag; ++i)
      for (int k = 0; k < noff; ++k) diag = diag && tg[i * ng + off[k]] == 0.f;
    m->tile_diag = diag;
    for (int64_t i = 0; i < E; ++i) {
      std::copy_n(&tg[i * ng], ng, &te[i * ned]);
      std::copy_n(&tp[i * nv * kp], nv * kp, &te[i * ned + ng]);
    }
    rc = dalloc(&m->d_telem, te.size(), &m->tile_bytes);
  }
  if (!rc) rc = hip_check(hipMemcpyAsync(m->d_telem, te.data(), te.size() * 4, hipMemcpyHostToDevice, m->stream), "upload");
  dfree(d_tg);
  // inverse mass in the internal numbering
  int32_t *d_to = nullptr;
  if (!rc) rc = dalloc(&d_to, (size_t)m->G);
  if (!rc) rc = hip_check(hipMemcpyAsync(d_to, ts.to_internal.data(), m->G * 4, hipMemcpyHostToDevice, m->stream), "upload");
  if (!rc) rc = dalloc(&m->d_mi_int, (size_t)m->G, &m->tile_bytes);
  if (!rc) rc = launch_permute_f32(m->G, d_to, m->d_mi, m->d_mi_int, m->stream);
  if (!rc) rc = hip_check(hipStreamSynchronize(m->stream), "sync");
  dfree(d_tv);
  dfree(d_to);
  return rc;
}

static void mesh_free(sem_mesh *m) {
  if (!m) return;
  (void)hipSetDevice(m->device);
  if (m->stream) (void)hipStreamSynchronize(m->stream);
  dfree(m->d_conn);
  dfree(m->d_geo_e);
  dfree(m->d_par_e);
  dfree(m->d_geo_p);
  dfree(m->d_mi);
  dfree(m->d_vtx);
  dfree(m->d_vpar);
  dfree(m->d_tile_info);
  dfree(m->d_tile_words);
  dfree(m->d_tile_slot);
  dfree(m->d_tile_words_lat);
  dfree(m->d_telem);
  dfree(m->d_mi_int);
  if (m->stream) (void)hipStreamDestroy(m->stream);
  delete m;
}

// ====================================================================== C ABI
extern "C" {

const char *sem_last_error(void) { return g_error.c_str(); }

int sem_device_count(int *count) {
  int n = 0;
  hipError_t e = hipGetDeviceCount(&n);
  *count = e == hipSuccess ? n : 0;
  return 0;
}

int64_t sem_reference_step_count(double duration, double dt, double *dt_out) {
  int64_t m = (int64_t)std::ceil(duration / dt);
  double h = duration / (double)m;
  if (dt_out) *dt_out = h;
  double t = 0;
  int64_t count = 0;
  while (t < duration) {
    t += h;
    ++count;
  }
  return count;
}

int sem_topology_build(int dim, int order, int64_t E, int64_t V, const int32_t *cells,
                       int32_t *conn, int32_t *colour, int64_t *info) {
  if (order < 1 || order > (dim == 2 ? kMaxOrder2D : kMaxOrder3D)) {
    set_error("Polynomial order " + std::to_string(order) + " not supported");
    return -1;
  }
  Topology T;
  SEM_TRY(build_topology(dim, order, E, V, cells, T));
  std::vector<int32_t> col;
  SEM_TRY(colour_elements(dim, E, V, cells, col));
  if (conn) std::copy(T.conn.begin(), T.conn.end(), conn);
  if (colour) std::copy(col.begin(), col.end(), colour);
  if (info) {
    info[0] = T.G;
    info[1] = E ? 1 + *std::max_element(col.begin(), col.end()) : 0;
    info[2] = T.NF;
    info[3] = T.NEd;
  }
  return 0;
}

int sem_mesh_create(sem_mesh **out, int device, int dim, int order, int physics, int64_t E,
                    int64_t V, const int32_t *cells, const double *coords, const double *vpar,
                    int quirks) {
  *out = nullptr;
  if (dim != 2 && dim != 3) { set_error("dim must be 2 or 3"); return -1; }
  if (physics != SEM_ACOUSTIC && physics != SEM_ELASTIC) { set_error("physics"); return -1; }
  if (order < 1 || order > (dim == 2 ? kMaxOrder2D : kMaxOrder3D)) {
    set_error("Polynomial order " + std::to_string(order) + " not supported");
    return -1;
  }
  if (E <= 0 || V <= 0 || !cells || !coords || !vpar) { set_error("empty mesh"); return -1; }
  SEM_TRY(set_device(device));
  auto *m = new sem_mesh();
  m->device = device;
  m->dim = dim;
  m->order = order;
  m->n = order + 1;
  m->nd = dim == 2 ? m->n * m->n : m->n * m->n * m->n;
  m->physics = physics;
  m->c = physics == SEM_ACOUSTIC ? 1 : dim;
  m->nv = dim == 2 ? 4 : 8;
  m->quirks = quirks;
  m->E = E;
  m->V = V;
  const int nv = m->nv, npar = ref_npar(dim, physics);
  int rc = 0;
  auto fail = [&](int r) { mesh_free(m); return r; };
  Topology T;
  if ((rc = build_topology(dim, order, E, V, cells, T))) return fail(rc);
  m->G = T.G;
  m->conn = std::move(T.conn);
  if ((rc = colour_elements(dim, E, V, cells, m->colour))) return fail(rc);
  m->ncol = 1 + *std::max_element(m->colour.begin(), m->colour.end());
  // sorted order: by colour, then element index
  m->col_start.assign(m->ncol + 1, 0);
  for (int64_t e = 0; e < E; ++e) ++m->col_start[m->colour[e] + 1];
  for (int c = 0; c < m->ncol; ++c) m->col_start[c + 1] += m->col_start[c];
  m->perm.resize(E);
  m->inv_perm.resize(E);
  {
    std::vector<int64_t> cur(m->col_start.begin(), m->col_start.end() - 1);
    for (int64_t e = 0; e < E; ++e) {
      int64_t sp = cur[m->colour[e]]++;
      m->perm[sp] = (int32_t)e;
      m->inv_perm[e] = (int32_t)sp;
    }
  }
  // host geometry (original order), validity and affinity
  m->vtx.resize(E * nv * dim);
  m->bbox.resize(E * 2 * dim);
  m->vpar.assign(vpar, vpar + E * nv * npar);
  bool affine = true;
  for (int64_t e = 0; e < E; ++e) {
    double *vx = &m->vtx[e * nv * dim];
    for (int k = 0; k < nv; ++k)
      for (int a = 0; a < dim; ++a) vx[k * dim + a] = coords[(int64_t)cells[e * nv + k] * dim + a];
    double J[9], ref[3] = {0, 0, 0};
    if (jacobian(dim, vx, ref, J) <= 0) {
      set_error("element " + std::to_string(e) + " has det J <= 0");
      return fail(-1);
    }
    for (int a = 0; a < dim; ++a) {
      double lo = vx[a], hi = vx[a];
      for (int k = 1; k < nv; ++k) {
        lo = std::min(lo, vx[k * dim + a]);
        hi = std::max(hi, vx[k * dim + a]);
      }
      m->bbox[e * 2 * dim + a] = lo;
      m->bbox[e * 2 * dim + dim + a] = hi;
    }
    if (affine && !is_affine(dim, vx)) affine = false;
  }
  m->affine = affine;
  // device upload (sorted order)
  if ((rc = hip_check(hipStreamCreateWithFlags(&m->stream, hipStreamNonBlocking), "stream"))) return fail(rc);
  if ((rc = upload_conn(m))) return fail(rc);
  {
    std::vector<double> sv(E * nv * dim), sp(E * nv * npar);
    for (int64_t s = 0; s < E; ++s) {
      int64_t e = m->perm[s];
      std::copy_n(&m->vtx[e * nv * dim], nv * dim, &sv[s * nv * dim]);
      std::copy_n(&m->vpar[e * nv * npar], nv * npar, &sp[s * nv * npar]);
    }
    if ((rc = dalloc(&m->d_vtx, sv.size(), &m->bytes))) return fail(rc);
    if ((rc = dalloc(&m->d_vpar, sp.size(), &m->bytes))) return fail(rc);
    if ((rc = hip_check(hipMemcpyAsync(m->d_vtx, sv.data(), sv.size() * 8, hipMemcpyHostToDevice, m->stream), "upload"))) return fail(rc);
    if ((rc = hip_check(hipMemcpyAsync(m->d_vpar, sp.data(), sp.size() * 8, hipMemcpyHostToDevice, m->stream), "upload"))) return fail(rc);
    // kernel material set (FP32 vertex values)
    const int kp = kernel_npar(dim, physics);
    std::vector<float> pe(E * nv * kp);
    for (int64_t s = 0; s < E; ++s)
      for (int k = 0; k < nv; ++k) {
        const double *src = &sp[(s * nv + k) * npar];
        float *dst = &pe[(s * nv + k) * kp];
        if (physics == SEM_ACOUSTIC) dst[0] = (float)src[0];
        else if (dim == 2) for (int q = 0; q < 4; ++q) dst[q] = (float)src[1 + q];  // C11 C13 C33 C55
        else for (int q = 0; q < 6; ++q) dst[q] = (float)src[q];
      }
    if ((rc = dalloc(&m->d_par_e, pe.size(), &m->bytes))) return fail(rc);
    if ((rc = hip_check(hipMemcpyAsync(m->d_par_e, pe.data(), pe.size() * 4, hipMemcpyHostToDevice, m->stream), "upload"))) return fail(rc);
    if ((rc = hip_check(hipStreamSynchronize(m->stream), "sync"))) return fail(rc);
  }
  if (affine) {
    if ((rc = dalloc(&m->d_geo_e, (size_t)E * affine_geo_count(dim, physics), &m->bytes))) return fail(rc);
    if ((rc = launch_affine_geometry(dim, physics, E, m->d_vtx, m->d_geo_e, m->stream))) return fail(rc);
  }
  // FP64 mass: element contributions, colour-ordered assembly, reciprocal
  {
    double *me = nullptr, *M = nullptr;
    if ((rc = dalloc(&me, (size_t)E * m->nd))) return fail(rc);
    if ((rc = dalloc(&M, (size_t)m->G))) { dfree(me); return fail(rc); }
    rc = launch_point_geometry(dim, order, physics, pointwise_mode(m), E, m->d_vtx, m->d_vpar, npar,
                               nullptr, me, m->stream);
    if (!rc) rc = launch_fill_f64(m->G, M, 0.0, m->stream);
    for (int c = 0; c < m->ncol && !rc; ++c)
      rc = launch_mass_scatter(m->col_start[c], m->col_start[c + 1] - m->col_start[c], m->nd,
                               m->d_conn, me, M, m->stream);
    if (!rc) rc = dalloc(&m->d_mi, (size_t)m->G, &m->bytes);
    if (!rc) rc = launch_reciprocal(m->G, M, m->d_mi, m->stream);
    if (!rc) rc = hip_check(hipStreamSynchronize(m->stream), "mass");
    dfree(me);
    dfree(M);
    if (rc) return fail(rc);
  }
  if (!affine && (rc = ensure_pointwise(m))) return fail(rc);
  if ((rc = setup_tiles(m, nullptr))) return fail(rc);
  *out = m;
  return 0;
}

int sem_mesh_destroy(sem_mesh *m) {
  mesh_free(m);
  return 0;
}

int sem_mesh_info(const sem_mesh *m, int64_t *info) {
  info[0] = m->E; info[1] = m->G; info[2] = m->ncol; info[3] = m->c;
  info[4] = m->nd; info[5] = m->affine; info[6] = (int64_t)m->bytes; info[7] = m->dim;
  info[8] = m->tiles.valid ? m->tiles.ntiles : 0;
  info[9] = m->tiles.valid ? m->tiles.T[0] * 10000 + m->tiles.T[1] * 100 + m->tiles.T[2] : 0;
  // tile path traffic counts: lattice slots loaded, slots on the mesh,
  // accumulator reads (slots not first) and writes (slots not last)
  int64_t valid = 0, accr = 0, accw = 0;
  for (uint32_t w : m->tiles.words) {
    if (w == kSlotEmpty) continue;
    ++valid;
    accr += (w & kFlagFirst) ? 0 : 1;
    accw += (w & kFlagLast) ? 0 : 1;
  }
  info[10] = m->tiles.valid ? m->tiles.total_slots : 0;
  info[11] = valid;
  info[12] = accr;
  info[13] = accw;
  info[14] = m->tiles.valid ? m->tiles.elems_per_tile : 0;
  info[15] = m->tiles.valid && m->tile_diag;
  return 0;
}

int sem_mesh_get_conn(const sem_mesh *m, int32_t *conn) {
  std::copy(m->conn.begin(), m->conn.end(), conn);
  return 0;
}

int sem_mesh_get_colours(const sem_mesh *m, int32_t *colour) {
  std::copy(m->colour.begin(), m->colour.end(), colour);
  return 0;
}

int sem_mesh_get_inverse_mass(const sem_mesh *m, float *mi) {
  HIPCHK(hipSetDevice(m->device));
  HIPCHK(hipMemcpy(mi, m->d_mi, m->G * sizeof(float), hipMemcpyDeviceToHost));
  return 0;
}

int sem_mesh_get_node_coords(const sem_mesh *m, double *xyz) {
  Gll g = make_gll(m->order);
  const int dim = m->dim, n = m->n;
  for (int64_t e = 0; e < m->E; ++e) {
    const double *vx = &m->vtx[e * m->nv * dim];
    for (int p = 0; p < m->nd; ++p) {
      double ref[3] = {g.xi[p % n], g.xi[(p / n) % n], dim == 3 ? g.xi[p / (n * n)] : 0.0};
      double w[8];
      p1_weights(dim, ref, w);
      int32_t gid = m->conn[e * m->nd + p];
      for (int a = 0; a < dim; ++a) {
        double s = 0;
        for (int k = 0; k < m->nv; ++k) s += w[k] * vx[k * dim + a];
        xyz[(int64_t)gid * dim + a] = s;
      }
    }
  }
  return 0;
}

int sem_mesh_set_dirichlet(sem_mesh *m, const uint8_t *mask) {
  HIPCHK(hipSetDevice(m->device));
  m->dirichlet.assign(mask, mask + m->G);
  SEM_TRY(upload_conn(m));
  if (m->tiles.valid) {
    const uint32_t flags = kFlagFirst | kFlagLast;
    TileSet &ts = m->tiles;
    std::vector<uint8_t> dint(m->G, 0);
    for (int64_t g = 0; g < m->G; ++g) dint[ts.to_internal[g]] = m->dirichlet[g];
    for (auto *ws : {&ts.words, &ts.words_lat})
      for (auto &w : *ws) {
        if (w == kSlotEmpty) continue;
        uint32_t node = w & kNodeMask;
        w = node | (w & flags) | (dint[node] ? kFlagDirichlet : 0u);
      }
    SEM_TRY(upload_tile_words(m));
  }
  return 0;
}

int sem_mesh_check_colouring(sem_mesh *m, int64_t *violations) {
  HIPCHK(hipSetDevice(m->device));
  int32_t *hits = nullptr;
  unsigned long long *viol = nullptr;
  SEM_TRY(dalloc(&hits, (size_t)m->G));
  int rc = dalloc(&viol, 1);
  if (!rc) rc = launch_fill_i32(m->G, hits, 0, m->stream);
  if (!rc) rc = hip_check(hipMemsetAsync(viol, 0, sizeof(unsigned long long), m->stream), "memset");
  for (int c = 0; c < m->ncol && !rc; ++c) {
    rc = launch_colour_touch(m->col_start[c], m->col_start[c + 1] - m->col_start[c], m->nd,
                             m->d_conn, hits, m->stream);
    if (!rc) rc = launch_colour_count(m->G, hits, viol, m->stream);
  }
  unsigned long long h = 0;
  if (!rc) rc = hip_check(hipMemcpyAsync(&h, viol, sizeof h, hipMemcpyDeviceToHost, m->stream), "copy");
  if (!rc) rc = hip_check(hipStreamSynchronize(m->stream), "sync");
  dfree(hits);
  dfree(viol);
  *violations = (int64_t)h;
  return rc;
}

}  // extern "C"

// ====================================================================== sources/receivers (host, FP64)
int sem::locate(const sem_mesh *m, const double *pt, int64_t *elem, double *ref) {
  const int dim = m->dim;
  for (int64_t e = 0; e < m->E; ++e) {
    const double *bb = &m->bbox[e * 2 * dim];
    bool out = false;
    for (int a = 0; a < dim; ++a) out = out || pt[a] < bb[a] || pt[a] > bb[dim + a];
    if (out) continue;  // a convex element lies inside its bounding box
    const double *vx = &m->vtx[e * m->nv * dim];
    int in = check_hull(dim, pt, vx);
    if (in < 0) { set_error("poorly behaved element"); return -1; }
    if (in) {
      if (inverse_map(dim, pt, vx, ref)) {
        set_error("inverseCoordinateTransform failed to converge");
        return -1;
      }
      *elem = e;
      return 0;
    }
  }
  set_error("point outside the mesh: source or receiver not added");
  return -1;
}

// Nodal load per unit source-time function, nd x c column-major (FP64).
int sem::source_load(const sem_mesh *m, const SrcSpec &s, int64_t *elem, std::vector<double> &load) {
  double ref[3] = {0, 0, 0};
  SEM_TRY(locate(m, s.loc, elem, ref));
  const int dim = m->dim, n = m->n, nd = m->nd, c = m->c;
  const double *vx = &m->vtx[*elem * m->nv * dim];
  Gll g = make_gll(m->order);
  double l[3][kMaxN], dl[3][kMaxN];
  for (int a = 0; a < dim; ++a) {
    lagrange(g, ref[a], l[a]);
    lagrange_deriv(g, ref[a], dl[a]);
  }
  load.assign((size_t)nd * c, 0.0);
  if (s.kind == 0 || s.kind == 2) {
    const bool e3d_quirk = dim == 3 && m->physics == SEM_ELASTIC && (m->quirks & SEM_QUIRK_E3D_SOURCE);
    for (int p = 0; p < nd; ++p) {
      int P[3] = {p % n, (p / n) % n, dim == 3 ? p / (n * n) : 0};
      double lp = 1.0, wq = 1.0;
      for (int a = 0; a < dim; ++a) {
        lp *= l[a][P[a]];
        wq *= g.w[P[a]];
      }
      // getDeltaFunctionCoefficients divides by w detJ at the node; the
      // test-and-integrate of Scalar/Elastic2D multiplies it back; Elastic3D's
      // source term skips it (Elastic3D.cpp:126-133).
      double r3[3] = {g.xi[P[0]], g.xi[P[1]], dim == 3 ? g.xi[P[2]] : 0.0}, J[9];
      double det = jacobian(dim, vx, r3, J);
      double delta = lp / (wq * det);
      double val = e3d_quirk ? delta : delta * det * wq;
      for (int k = 0; k < c; ++k) load[(size_t)k * nd + p] = val * s.dir[k];
    }
  } else {
    if (m->physics != SEM_ELASTIC) { set_error("moment tensor needs elastic physics"); return -1; }
    double J[9], inv[9];
    jacobian(dim, vx, ref, J);
    invert(dim, J, inv);
    for (int p = 0; p < nd; ++p) {
      int P[3] = {p % n, (p / n) % n, dim == 3 ? p / (n * n) : 0};
      double gr[3];
      for (int a = 0; a < dim; ++a) {
        double v = dl[a][P[a]];
        for (int b = 0; b < dim; ++b)
          if (b != a) v *= l[b][P[b]];
        gr[a] = v;
      }
      double gx[3];
      for (int a = 0; a < dim; ++a) {
        double v = 0;
        for (int b = 0; b < dim; ++b) v += inv[a * dim + b] * gr[b];
        gx[a] = v;
      }
      for (int i = 0; i < dim; ++i) {
        double v = 0;
        for (int j = 0; j < dim; ++j) v += s.M[i * dim + j] * gx[j];
        load[(size_t)i * nd + p] = -v;
      }
    }
  }
  return 0;
}

double sem::ricker(double t, double amp, double f0, double t0) {
  const double factor = M_PI * M_PI * f0 * f0 * (t - t0) * (t - t0);
  return amp * ((1 - 2 * factor) * std::exp(-1 * factor));
}

