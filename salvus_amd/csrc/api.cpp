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

#include "sem_internal.h"

namespace sem {

static thread_local std::string g_error;
void set_error(const std::string &msg) { g_error = msg; }
int hip_check(hipError_t e, const char *what) {
  if (e == hipSuccess) return 0;
  set_error(std::string(what) + ": " + hipGetErrorString(e));
  return -2;
}

template <typename T>
static int dalloc(T **p, size_t count, size_t *bytes = nullptr) {
  *p = nullptr;
  if (count == 0) return 0;
  hipError_t e = hipMalloc((void **)p, count * sizeof(T));
  if (e != hipSuccess) {
    set_error(std::string("device allocation of ") + std::to_string(count * sizeof(T)) +
              " bytes failed: " + hipGetErrorString(e));
    *p = nullptr;
    return -3;
  }
  if (bytes) *bytes += count * sizeof(T);
  return 0;
}
template <typename T>
static void dfree(T *&p) {
  if (p) (void)hipFree((void *)p);
  p = nullptr;
}

struct SrcSpec {
  int kind;  // 0 force, 1 moment tensor
  double loc[3], amp, f0, t0, dir[3], M[9];
};

struct Shot {
  sem_mesh *mesh = nullptr;
  int64_t nsteps = 0;
  double dt = 0;
  std::vector<SrcSpec> srcs;
  std::vector<std::array<double, 3>> recs;
  std::vector<double> init[3];
  int group = -1, slot = 0;
  int64_t seis_off = 0;
};

struct Group {
  sem_mesh *mesh = nullptr;
  std::vector<int> shots;
  int S = 1;
  int64_t nsteps = 0;
  double dt = 0;
  float *fields = nullptr;  // 6 arrays of c*G*S: u v ap ku kv acc
  SrcDev *d_src = nullptr;
  float *d_stf = nullptr, *d_load = nullptr;
  int nsrc = 0;
  bool tiled = false;  // tile kernel; fields in the mesh's tile-major numbering
};

struct LaunchSet {
  int dim, order, physics, nd, c, colour;
  bool affine;
  bool tile;
  int nitems;
  int T[3];  // tile shape (tile kernel)
  WorkItem *d_items;
};

struct RecSet {
  int C = 0, nrec = 0;
  RecDev *d_recs = nullptr;
  int32_t *d_group = nullptr;
};

}  // namespace sem

using namespace sem;

struct sem_batch {
  int device = 0;
  hipStream_t stream = nullptr;
  std::vector<Shot> shots;
  std::vector<Group> groups;
  bool prepared = false;
  int path = SEM_PATH_FUSED, epb = 0, spg = 0, block_steps = 0;
  GroupDev *d_groups = nullptr;
  std::vector<GroupDev> h_groups;
  int64_t *d_steps = nullptr;
  std::vector<LaunchSet> launches;
  std::vector<RecSet> recsets;
  float *d_rec_w = nullptr;
  uint32_t *d_rec_nodes = nullptr;
  float *d_seis = nullptr;
  int64_t seis_total = 0;
  float *d_uloc = nullptr, *d_aloc = nullptr;  // plain path scratch
  hipGraphExec_t graph = nullptr;
  hipGraph_t graph_def = nullptr;
  int64_t steps_done = 0;
  bool timing = false;
  double kernel_ms = 0;
  int64_t kernel_launches = 0;
  std::vector<hipEvent_t> ev;
};

// ====================================================================== helpers
static int set_device(int device) {
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

static int ensure_pointwise(sem_mesh *m) {
  if (m->d_geo_p) return 0;
  const int ngp = point_geo_count(m->dim, m->physics);
  SEM_TRY(dalloc(&m->d_geo_p, (size_t)m->E * ngp * m->nd, &m->bytes));
  SEM_TRY(launch_point_geometry(m->dim, m->order, m->physics, m->E, m->d_vtx, m->d_vpar,
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
  static const int pref3a[][3] = {{4, 4, 2}, {4, 2, 2}, {2, 2, 2}};
  static const int pref3e[][3] = {{2, 2, 2}, {4, 2, 2}};
  const int(*list)[3] = dim == 2 ? pref2 : (physics == SEM_ACOUSTIC ? pref3a : pref3e);
  const int cnt = dim == 2 ? 2 : (physics == SEM_ACOUSTIC ? 3 : 2);
  for (int i = 0; i < cnt; ++i) {
    T[0] = list[i][0]; T[1] = list[i][1]; T[2] = list[i][2];
    const TileKernelInfo k = tile_kernel_info(dim, order, physics, T);
    if (k.lds_bytes > 0 && k.lds_bytes <= 64 * 1024) return true;
  }
  return false;
}

static int upload_tile_words(sem_mesh *m) {
  const TileSet &ts = m->tiles;
  if (!m->d_tile_words) SEM_TRY(dalloc(&m->d_tile_words, ts.words.size(), &m->bytes));
  HIPCHK(hipMemcpyAsync(m->d_tile_words, ts.words.data(), ts.words.size() * 4, hipMemcpyHostToDevice, m->stream));
  HIPCHK(hipStreamSynchronize(m->stream));
  return 0;
}

// Builds and uploads the tile path of an affine, logically Cartesian mesh.
static int setup_tiles(sem_mesh *m) {
  const char *off = getenv("SEM_NO_TILES");
  if (!m->affine || (off && off[0] == '1')) return 0;
  int T[3];
  if (!pick_tile(m->dim, m->order, m->physics, T)) return 0;
  if (build_tiles(m, T[0], T[1], T[2], m->tiles) != 0) {
    m->tiles = TileSet();
    return 0;  // not eligible: the element kernel is used
  }
  TileSet &ts = m->tiles;
  const int dim = m->dim, nv = m->nv, npar = ref_npar(dim, m->physics), kp = kernel_npar(dim, m->physics);
  // padded tile element order (padding repeats element 0; never read)
  const int64_t E = (int64_t)ts.elem_order.size();
  SEM_TRY(dalloc(&m->d_tile_info, ts.info.size(), &m->bytes));
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
  double *d_tv = nullptr;
  SEM_TRY(dalloc(&d_tv, tv.size()));
  int rc = hip_check(hipMemcpyAsync(d_tv, tv.data(), tv.size() * 8, hipMemcpyHostToDevice, m->stream), "upload");
  if (!rc) rc = dalloc(&m->d_tgeo_e, (size_t)E * affine_geo_count(dim, m->physics), &m->bytes);
  if (!rc) rc = launch_affine_geometry(dim, m->physics, E, d_tv, m->d_tgeo_e, m->stream);
  if (!rc) rc = dalloc(&m->d_tpar_e, tp.size(), &m->bytes);
  if (!rc) rc = hip_check(hipMemcpyAsync(m->d_tpar_e, tp.data(), tp.size() * 4, hipMemcpyHostToDevice, m->stream), "upload");
  // inverse mass in the internal numbering
  int32_t *d_to = nullptr;
  if (!rc) rc = dalloc(&d_to, (size_t)m->G);
  if (!rc) rc = hip_check(hipMemcpyAsync(d_to, ts.to_internal.data(), m->G * 4, hipMemcpyHostToDevice, m->stream), "upload");
  if (!rc) rc = dalloc(&m->d_mi_int, (size_t)m->G, &m->bytes);
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
  dfree(m->d_tgeo_e);
  dfree(m->d_tpar_e);
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
    rc = launch_point_geometry(dim, order, physics, E, m->d_vtx, m->d_vpar, npar, nullptr, me, m->stream);
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
  if ((rc = setup_tiles(m))) return fail(rc);
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
    for (auto &w : ts.words) {
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
static int locate(const sem_mesh *m, const double *pt, int64_t *elem, double *ref) {
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
static int source_load(const sem_mesh *m, const SrcSpec &s, int64_t *elem, std::vector<double> &load) {
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
  if (s.kind == 0) {
    const bool e3d_quirk = dim == 3 && m->physics == SEM_ELASTIC && (m->quirks & SEM_QUIRK_E3D_SOURCE);
    for (int p = 0; p < nd; ++p) {
      int P[3] = {p % n, (p / n) % n, dim == 3 ? p / (n * n) : 0};
      double lp = 1.0, wq = 1.0;
      for (int a = 0; a < dim; ++a) {
        lp *= l[a][P[a]];
        wq *= g.w[P[a]];
      }
      // getDeltaFunctionCoefficients divides by w detJ at the node; the
      // test-and-integrate of Scalar/Elastic2D multiplies it back (Q7).
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

static double ricker(double t, double amp, double f0, double t0) {
  const double factor = M_PI * M_PI * f0 * f0 * (t - t0) * (t - t0);
  return amp * ((1 - 2 * factor) * std::exp(-1 * factor));
}

// ====================================================================== batch
static void batch_release(sem_batch *b) {
  if (b->graph) (void)hipGraphExecDestroy(b->graph);
  if (b->graph_def) (void)hipGraphDestroy(b->graph_def);
  b->graph = nullptr;
  b->graph_def = nullptr;
  for (auto &g : b->groups) {
    dfree(g.fields);
    dfree(g.d_src);
    dfree(g.d_stf);
    dfree(g.d_load);
  }
  b->groups.clear();
  for (auto &l : b->launches) dfree(l.d_items);
  b->launches.clear();
  for (auto &r : b->recsets) {
    dfree(r.d_recs);
    dfree(r.d_group);
  }
  b->recsets.clear();
  dfree(b->d_groups);
  dfree(b->d_steps);
  dfree(b->d_rec_w);
  dfree(b->d_rec_nodes);
  dfree(b->d_seis);
  dfree(b->d_uloc);
  dfree(b->d_aloc);
  for (auto e : b->ev) (void)hipEventDestroy(e);
  b->ev.clear();
  b->prepared = false;
}

extern "C" {

int sem_batch_create(sem_batch **out, int device) {
  *out = nullptr;
  SEM_TRY(set_device(device));
  auto *b = new sem_batch();
  b->device = device;
  if (hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking) != hipSuccess) {
    delete b;
    set_error("stream creation failed");
    return -2;
  }
  *out = b;
  return 0;
}

int sem_batch_destroy(sem_batch *b) {
  if (!b) return 0;
  (void)hipSetDevice(b->device);
  (void)hipStreamSynchronize(b->stream);
  batch_release(b);
  (void)hipStreamDestroy(b->stream);
  delete b;
  return 0;
}

int sem_batch_add_shot(sem_batch *b, sem_mesh *m, int64_t nsteps, double dt, int *shot) {
  if (b->prepared) { set_error("batch already prepared"); return -1; }
  if (m->device != b->device) { set_error("mesh and batch on different devices"); return -1; }
  if (nsteps < 0 || !(dt > 0)) { set_error("nsteps >= 0 and dt > 0 required"); return -1; }
  Shot s;
  s.mesh = m;
  s.nsteps = nsteps;
  s.dt = dt;
  b->shots.push_back(s);
  *shot = (int)b->shots.size() - 1;
  return 0;
}

static Shot *get_shot(sem_batch *b, int shot) {
  if (shot < 0 || shot >= (int)b->shots.size()) {
    set_error("shot index out of range");
    return nullptr;
  }
  return &b->shots[shot];
}

int sem_shot_add_source(sem_batch *b, int shot, const double *loc, double amp, double f0,
                        double t0, const double *dir) {
  Shot *s = get_shot(b, shot);
  if (!s) return -1;
  if (b->prepared) { set_error("batch already prepared"); return -1; }
  SrcSpec sp{};
  sp.kind = 0;
  for (int a = 0; a < s->mesh->dim; ++a) sp.loc[a] = loc[a];
  sp.amp = amp; sp.f0 = f0; sp.t0 = t0;
  for (int k = 0; k < s->mesh->c; ++k) sp.dir[k] = dir ? dir[k] : (k == 0 ? 1.0 : 0.0);
  int64_t e;
  std::vector<double> load;
  SEM_TRY(source_load(s->mesh, sp, &e, load));  // validates placement now
  s->srcs.push_back(sp);
  return 0;
}

int sem_shot_add_moment_source(sem_batch *b, int shot, const double *loc, const double *M,
                               double amp, double f0, double t0) {
  Shot *s = get_shot(b, shot);
  if (!s) return -1;
  if (b->prepared) { set_error("batch already prepared"); return -1; }
  SrcSpec sp{};
  sp.kind = 1;
  const int d = s->mesh->dim;
  for (int a = 0; a < d; ++a) sp.loc[a] = loc[a];
  for (int i = 0; i < d * d; ++i) sp.M[i] = M[i];
  sp.amp = amp; sp.f0 = f0; sp.t0 = t0;
  int64_t e;
  std::vector<double> load;
  SEM_TRY(source_load(s->mesh, sp, &e, load));
  s->srcs.push_back(sp);
  return 0;
}

int sem_shot_add_receiver(sem_batch *b, int shot, const double *loc) {
  Shot *s = get_shot(b, shot);
  if (!s) return -1;
  if (b->prepared) { set_error("batch already prepared"); return -1; }
  std::array<double, 3> r{0, 0, 0};
  for (int a = 0; a < s->mesh->dim; ++a) r[a] = loc[a];
  int64_t e;
  double ref[3];
  SEM_TRY(locate(s->mesh, r.data(), &e, ref));
  s->recs.push_back(r);
  return 0;
}

// Position of DMPlex DOF entry i (= comp*G + g) in a group's field arrays.
static inline int64_t field_pos(const Group &g, int64_t i) {
  if (!g.tiled) return i;
  const sem_mesh *m = g.mesh;
  const int64_t c = i / m->G, node = i - c * m->G;
  return c * m->G + m->tiles.to_internal[node];
}

static float *field_base(sem_batch *b, const Shot &s, int which) {
  Group &g = b->groups[s.group];
  const size_t len = (size_t)s.mesh->c * s.mesh->G * g.S;
  return g.fields + (size_t)which * len;
}

int sem_shot_set_field(sem_batch *b, int shot, int which, const double *vals) {
  Shot *s = get_shot(b, shot);
  if (!s) return -1;
  if (which < 0 || which > 2) { set_error("field selector"); return -1; }
  const size_t len = (size_t)s->mesh->c * s->mesh->G;
  if (!b->prepared) {
    s->init[which].assign(vals, vals + len);
    return 0;
  }
  // live update (S-strided)
  HIPCHK(hipSetDevice(b->device));
  Group &g = b->groups[s->group];
  std::vector<float> buf((size_t)len * g.S);
  float *base = field_base(b, *s, which);
  HIPCHK(hipMemcpyAsync(buf.data(), base, buf.size() * 4, hipMemcpyDeviceToHost, b->stream));
  HIPCHK(hipStreamSynchronize(b->stream));
  for (size_t i = 0; i < len; ++i) buf[field_pos(g, i) * g.S + s->slot] = (float)vals[i];
  HIPCHK(hipMemcpyAsync(base, buf.data(), buf.size() * 4, hipMemcpyHostToDevice, b->stream));
  HIPCHK(hipStreamSynchronize(b->stream));
  return 0;
}

int sem_shot_get_field(sem_batch *b, int shot, int which, double *vals) {
  Shot *s = get_shot(b, shot);
  if (!s) return -1;
  if (which < 0 || which > 2) { set_error("field selector"); return -1; }
  const size_t len = (size_t)s->mesh->c * s->mesh->G;
  if (!b->prepared) {
    for (size_t i = 0; i < len; ++i) vals[i] = s->init[which].empty() ? 0.0 : s->init[which][i];
    return 0;
  }
  HIPCHK(hipSetDevice(b->device));
  Group &g = b->groups[s->group];
  std::vector<float> buf(len * g.S);
  HIPCHK(hipMemcpyAsync(buf.data(), field_base(b, *s, which), buf.size() * 4, hipMemcpyDeviceToHost, b->stream));
  HIPCHK(hipStreamSynchronize(b->stream));
  for (size_t i = 0; i < len; ++i) vals[i] = buf[field_pos(g, i) * g.S + s->slot];
  return 0;
}

}  // extern "C"

// Default launch shape: elements per block so a block has ~128-256 threads.
static int default_epb(int nd) {
  int best = 1;
  for (int epb = 1; epb * nd <= 512; ++epb)
    if (block_threads(nd, epb) <= 256) best = epb;
  return best;
}

// Fused-path stiffness-only timing of a launch shape (no state changes
// besides the scratch accumulator).
static double time_shape(sem_batch *b, int epb, int reps);

static int build_plan(sem_batch *b) {
  // work items per (kernel key, colour); key = (tile?, dim, order, physics, affine, tile shape)
  std::map<std::tuple<int, int, int, int, int, int>, std::vector<int>> keys;
  for (int gi = 0; gi < (int)b->groups.size(); ++gi) {
    sem_mesh *m = b->groups[gi].mesh;
    const bool t = b->groups[gi].tiled;
    const int shape = t ? m->tiles.T[0] * 10000 + m->tiles.T[1] * 100 + m->tiles.T[2] : 0;
    keys[std::make_tuple((int)t, m->dim, m->order, m->physics, (int)m->affine, shape)].push_back(gi);
  }
  for (auto &l : b->launches) dfree(l.d_items);
  b->launches.clear();
  for (auto &kv : keys) {
    const bool tile = std::get<0>(kv.first) != 0;
    int maxcol = 0;
    for (int gi : kv.second) {
      sem_mesh *m = b->groups[gi].mesh;
      maxcol = std::max(maxcol, tile ? (1 << m->dim) : m->ncol);
    }
    for (int c = 0; c < maxcol; ++c) {
      std::vector<WorkItem> items;
      for (int gi : kv.second) {
        sem_mesh *m = b->groups[gi].mesh;
        if (tile) {
          const TileSet &ts = m->tiles;
          if (c >= (int)ts.colour_start.size() - 1) continue;
          for (int64_t k = ts.colour_start[c]; k < ts.colour_start[c + 1]; ++k) {
            const int32_t t = ts.by_colour[k];
            const TileInfo &ti = ts.info[t];
            WorkItem w;
            w.group = gi;
            w.first = t;
            w.count = ti.ex * ti.ey * ti.ez;
            w.pad = 0;
            items.push_back(w);
          }
        } else {
          if (c >= m->ncol) continue;
          for (int64_t f = m->col_start[c]; f < m->col_start[c + 1]; f += b->epb) {
            WorkItem w;
            w.group = gi;
            w.first = (int32_t)f;
            w.count = (int32_t)std::min<int64_t>(b->epb, m->col_start[c + 1] - f);
            w.pad = 0;
            items.push_back(w);
          }
        }
      }
      if (items.empty()) continue;
      LaunchSet L;
      L.tile = tile;
      L.dim = std::get<1>(kv.first);
      L.order = std::get<2>(kv.first);
      L.physics = std::get<3>(kv.first);
      L.affine = std::get<4>(kv.first) != 0;
      L.nd = L.dim == 2 ? (L.order + 1) * (L.order + 1) : (L.order + 1) * (L.order + 1) * (L.order + 1);
      L.c = L.physics == SEM_ACOUSTIC ? 1 : L.dim;
      L.colour = c;
      L.nitems = (int)items.size();
      const int shape = std::get<5>(kv.first);
      L.T[0] = shape / 10000;
      L.T[1] = (shape / 100) % 100;
      L.T[2] = shape % 100;
      SEM_TRY(dalloc(&L.d_items, items.size()));
      HIPCHK(hipMemcpyAsync(L.d_items, items.data(), items.size() * sizeof(WorkItem), hipMemcpyHostToDevice, b->stream));
      b->launches.push_back(L);
    }
  }
  HIPCHK(hipStreamSynchronize(b->stream));
  return 0;
}

static int enqueue_fused_step(sem_batch *b, int koff, int ku_only) {
  for (auto &L : b->launches) {
    FusedArgs a;
    a.groups = b->d_groups;
    a.items = L.d_items;
    a.nitems = L.nitems;
    a.epb = b->epb;
    a.koff = koff;
    a.ku_only = ku_only;
    auto go = [&]() {
      return L.tile ? launch_tile(L.dim, L.order, L.physics, L.T, L.nitems, a, b->stream)
                    : launch_fused(L.dim, L.order, L.physics, L.affine, L.nitems, b->epb, L.nd, L.c, a, b->stream);
    };
    if (b->timing && !ku_only) {
      hipEvent_t e0, e1;
      HIPCHK(hipEventCreate(&e0));
      HIPCHK(hipEventCreate(&e1));
      b->ev.push_back(e0);
      b->ev.push_back(e1);
      HIPCHK(hipEventRecord(e0, b->stream));
      SEM_TRY(go());
      HIPCHK(hipEventRecord(e1, b->stream));
    } else {
      SEM_TRY(go());
    }
  }
  if (ku_only) return 0;
  for (auto &R : b->recsets) {
    RecArgs ra;
    ra.groups = b->d_groups;
    ra.rec_group = R.d_group;
    ra.recs = R.d_recs;
    ra.nodes = b->d_rec_nodes;
    ra.weights = b->d_rec_w;
    ra.out = b->d_seis;
    ra.nrec = R.nrec;
    ra.nd = 0;
    ra.koff = koff;
    ra.pad = 0;
    // nd is uniform within a receiver set (grouped by C and nd)
    ra.nd = R.C >> 8;
    SEM_TRY(launch_receivers(ra, R.C & 0xff, b->stream));
  }
  return 0;
}

static int enqueue_plain_step(sem_batch *b) {
  for (int gi = 0; gi < (int)b->groups.size(); ++gi) {
    Group &g = b->groups[gi];
    GroupDev &h = b->h_groups[gi];
    sem_mesh *m = g.mesh;
    SEM_TRY(launch_plain_gather(m->E, m->nd, m->c, m->G, m->d_conn, h.u, b->d_uloc, b->stream));
    PlainArgs a;
    a.conn = m->d_conn;
    a.geo_p = m->d_geo_p;
    a.uloc = b->d_uloc;
    a.aloc = b->d_aloc;
    a.src = g.d_src;
    a.stf = g.d_stf;
    a.load = g.d_load;
    a.step = b->d_steps + gi;
    a.nsrc = g.nsrc;
    a.epb = default_epb(m->nd);
    a.E = m->E;
    a.koff = 0;
    a.pad = 0;
    if (b->timing) {
      hipEvent_t e0, e1;
      HIPCHK(hipEventCreate(&e0));
      HIPCHK(hipEventCreate(&e1));
      b->ev.push_back(e0);
      b->ev.push_back(e1);
      HIPCHK(hipEventRecord(e0, b->stream));
      SEM_TRY(launch_plain_stiffness(m->dim, m->order, m->physics, m->E, m->nd, m->c, a, b->stream));
      HIPCHK(hipEventRecord(e1, b->stream));
    } else {
      SEM_TRY(launch_plain_stiffness(m->dim, m->order, m->physics, m->E, m->nd, m->c, a, b->stream));
    }
    SEM_TRY(launch_fill_f32((int64_t)m->c * m->G, h.acc, 0.f, b->stream));
    for (int c = 0; c < m->ncol; ++c)
      SEM_TRY(launch_plain_scatter(m->col_start[c], m->col_start[c + 1] - m->col_start[c], m->nd, m->c,
                                   m->G, m->d_conn, b->d_aloc, h.acc, b->stream));
    SEM_TRY(launch_plain_update(m->G, m->c, m->quirks, h.dt, h.acl, h.dsp, m->d_mi, h.acc, h.u, h.v,
                                h.ap, h.ku, h.kv, b->stream));
  }
  for (auto &R : b->recsets) {
    RecArgs ra;
    ra.groups = b->d_groups;
    ra.rec_group = R.d_group;
    ra.recs = R.d_recs;
    ra.nodes = b->d_rec_nodes;
    ra.weights = b->d_rec_w;
    ra.out = b->d_seis;
    ra.nrec = R.nrec;
    ra.nd = R.C >> 8;
    ra.koff = 0;
    ra.pad = 0;
    SEM_TRY(launch_receivers(ra, R.C & 0xff, b->stream));
  }
  SEM_TRY(launch_advance((int)b->groups.size(), b->d_groups, b->d_steps, 1, b->stream));
  return 0;
}

static int capture_graph(sem_batch *b) {
  if (b->graph) return 0;
  HIPCHK(hipStreamBeginCapture(b->stream, hipStreamCaptureModeThreadLocal));
  int rc = 0;
  for (int k = 0; k < b->block_steps && !rc; ++k) rc = enqueue_fused_step(b, k, 0);
  if (!rc) rc = launch_advance((int)b->groups.size(), b->d_groups, b->d_steps, b->block_steps, b->stream);
  hipGraph_t gdef = nullptr;
  hipError_t e = hipStreamEndCapture(b->stream, &gdef);
  if (rc) {
    if (gdef) (void)hipGraphDestroy(gdef);
    return rc;
  }
  HIPCHK(e);
  b->graph_def = gdef;
  HIPCHK(hipGraphInstantiate(&b->graph, gdef, nullptr, nullptr, 0));
  return 0;
}

extern "C" {

int sem_batch_prepare(sem_batch *b, int path, int epb, int spg) {
  if (b->prepared) { set_error("batch already prepared"); return -1; }
  if (b->shots.empty()) { set_error("batch has no shots"); return -1; }
  if (path != SEM_PATH_PLAIN && path != SEM_PATH_FUSED) { set_error("path"); return -1; }
  HIPCHK(hipSetDevice(b->device));
  b->path = path;
  int rc = 0;
  // ---- grouping: shots sharing mesh, nsteps and dt (fused path only)
  const int max_s = path == SEM_PATH_PLAIN ? 1 : (spg > 0 ? spg : 8);
  std::map<std::tuple<sem_mesh *, int64_t, double>, std::vector<int>> buckets;
  for (int i = 0; i < (int)b->shots.size(); ++i)
    buckets[std::make_tuple(b->shots[i].mesh, b->shots[i].nsteps, b->shots[i].dt)].push_back(i);
  for (auto &kv : buckets) {
    auto &list = kv.second;
    for (size_t k = 0; k < list.size(); k += max_s) {
      Group g;
      g.mesh = std::get<0>(kv.first);
      g.nsteps = std::get<1>(kv.first);
      g.dt = std::get<2>(kv.first);
      for (size_t j = k; j < std::min(list.size(), k + (size_t)max_s); ++j) g.shots.push_back(list[j]);
      g.S = (int)g.shots.size();
      b->groups.push_back(g);
    }
  }
  auto fail = [&](int r) { batch_release(b); return r; };
  // ---- per-group device state
  std::vector<float> rec_w;
  std::vector<uint32_t> rec_nodes;
  std::vector<std::vector<std::pair<RecDev, int>>> recs_by_key;  // per (C, nd) key
  std::map<int, int> rec_key;
  int64_t seis = 0;
  for (int gi = 0; gi < (int)b->groups.size(); ++gi) {
    Group &g = b->groups[gi];
    sem_mesh *m = g.mesh;
    // tile kernel: 32-bit field offsets and at most kTileMaxSrc sources per tile
    g.tiled = path == SEM_PATH_FUSED && m->tiles.valid &&
              (uint64_t)m->c * (uint64_t)m->G * (uint64_t)g.S < (1ull << 32);
    if (g.tiled) {
      std::map<int64_t, int> per_tile;
      for (int sl = 0; sl < g.S && g.tiled; ++sl)
        for (auto &sp : b->shots[g.shots[sl]].srcs) {
          int64_t e;
          std::vector<double> ld;
          if ((rc = source_load(m, sp, &e, ld))) return fail(rc);
          if (++per_tile[m->tiles.elem_pos[e] / m->tiles.elems_per_tile] > kTileMaxSrc) {
            g.tiled = false;
            break;
          }
        }
    }
    const size_t len = (size_t)m->c * m->G * g.S;
    if ((rc = dalloc(&g.fields, 6 * len))) return fail(rc);
    std::vector<float> buf(len);
    for (int which = 0; which < 6; ++which) {
      std::fill(buf.begin(), buf.end(), 0.f);
      if (which < 3)
        for (int sl = 0; sl < g.S; ++sl) {
          const Shot &s = b->shots[g.shots[sl]];
          if (s.init[which].empty()) continue;
          for (size_t i = 0; i < (size_t)m->c * m->G; ++i) buf[field_pos(g, i) * g.S + sl] = (float)s.init[which][i];
        }
      if ((rc = hip_check(hipMemcpyAsync(g.fields + which * len, buf.data(), len * 4, hipMemcpyHostToDevice, b->stream), "upload"))) return fail(rc);
    }
    if ((rc = hip_check(hipStreamSynchronize(b->stream), "sync"))) return fail(rc);
    // sources: STF tables in FP64 at the accumulated times t_n (Main.cpp:45,
    // Order2Newmark.cpp:112), stored FP32
    std::vector<SrcDev> src;
    std::vector<float> stf, load;
    for (int sl = 0; sl < g.S; ++sl) {
      Shot &s = b->shots[g.shots[sl]];
      s.group = gi;
      s.slot = sl;
      for (auto &sp : s.srcs) {
        int64_t e;
        std::vector<double> ld;
        if ((rc = source_load(m, sp, &e, ld))) return fail(rc);
        SrcDev d;
        d.elem = g.tiled ? m->tiles.elem_pos[e] : m->inv_perm[e];
        d.stf_off = (int32_t)stf.size();
        d.load_off = (int32_t)load.size();
        d.shot = sl;
        double t = 0.0;
        for (int64_t k = 0; k < g.nsteps; ++k) {
          stf.push_back((float)ricker(t, sp.amp, sp.f0, sp.t0));
          t += g.dt;
        }
        for (double v : ld) load.push_back((float)v);
        src.push_back(d);
      }
      // receivers
      const int key = (m->c) | (m->nd << 8);
      if (!rec_key.count(key)) {
        rec_key[key] = (int)recs_by_key.size();
        recs_by_key.emplace_back();
      }
      s.seis_off = seis;
      Gll gl = make_gll(m->order);
      for (size_t r = 0; r < s.recs.size(); ++r) {
        int64_t e;
        double ref[3];
        if ((rc = locate(m, s.recs[r].data(), &e, ref))) return fail(rc);
        RecDev d;
        d.nodes_off = (int32_t)rec_nodes.size();
        for (int p = 0; p < m->nd; ++p) {
          int32_t gid = m->conn[e * m->nd + p];
          rec_nodes.push_back((uint32_t)(g.tiled ? m->tiles.to_internal[gid] : gid));
        }
        d.w_off = (int32_t)rec_w.size();
        d.shot = sl;
        d.stride = (int32_t)(s.recs.size() * m->c);
        d.out_off = seis + (int64_t)r * m->c;
        double l[3][kMaxN];
        for (int a = 0; a < m->dim; ++a) lagrange(gl, ref[a], l[a]);
        for (int p = 0; p < m->nd; ++p) {
          int P[3] = {p % m->n, (p / m->n) % m->n, m->dim == 3 ? p / (m->n * m->n) : 0};
          double v = 1;
          for (int a = 0; a < m->dim; ++a) v *= l[a][P[a]];
          rec_w.push_back((float)v);
        }
        recs_by_key[rec_key[key]].push_back({d, gi});
      }
      seis += (int64_t)s.recs.size() * m->c * s.nsteps;
    }
    g.nsrc = (int)src.size();
    if ((rc = dalloc(&g.d_src, src.size()))) return fail(rc);
    if ((rc = dalloc(&g.d_stf, stf.size()))) return fail(rc);
    if ((rc = dalloc(&g.d_load, load.size()))) return fail(rc);
    if (!src.empty()) {
      if ((rc = hip_check(hipMemcpyAsync(g.d_src, src.data(), src.size() * sizeof(SrcDev), hipMemcpyHostToDevice, b->stream), "upload"))) return fail(rc);
      if ((rc = hip_check(hipMemcpyAsync(g.d_stf, stf.data(), stf.size() * 4, hipMemcpyHostToDevice, b->stream), "upload"))) return fail(rc);
      if ((rc = hip_check(hipMemcpyAsync(g.d_load, load.data(), load.size() * 4, hipMemcpyHostToDevice, b->stream), "upload"))) return fail(rc);
    }
    if (path == SEM_PATH_PLAIN || (!m->affine && !g.tiled))
      if ((rc = ensure_pointwise(m))) return fail(rc);
    GroupDev h{};
    h.conn = m->d_conn;
    if (g.tiled) {
      h.tiles = m->d_tile_info;
      h.tile_words = m->d_tile_words;
      h.tgeo_e = m->d_tgeo_e;
      h.tpar_e = m->d_tpar_e;
    }
    h.geo_e = m->d_geo_e;
    h.par_e = m->d_par_e;
    h.geo_p = m->d_geo_p;
    h.mi = g.tiled ? m->d_mi_int : m->d_mi;
    h.u = g.fields;
    h.v = g.fields + len;
    h.ap = g.fields + 2 * len;
    h.ku = g.fields + 3 * len;
    h.kv = g.fields + 4 * len;
    h.acc = g.fields + 5 * len;
    h.src = g.d_src;
    h.stf = g.d_stf;
    h.load = g.d_load;
    h.nsrc = g.nsrc;
    h.S = g.S;
    h.G = m->G;
    h.nsteps = g.nsteps;
    h.dt = (float)g.dt;
    h.acl = (float)((1.0 / 2.0) * g.dt);   // Order2Newmark.cpp:89
    h.dsp = (float)((1.0 / 2.0) * (g.dt * g.dt));  // Order2Newmark.cpp:90
    h.quirks = m->quirks;
    b->h_groups.push_back(h);
  }
  if ((rc = dalloc(&b->d_steps, b->groups.size()))) return fail(rc);
  for (size_t gi = 0; gi < b->groups.size(); ++gi) b->h_groups[gi].step = b->d_steps + gi;
  if ((rc = dalloc(&b->d_groups, b->groups.size()))) return fail(rc);
  if ((rc = hip_check(hipMemcpyAsync(b->d_groups, b->h_groups.data(), b->h_groups.size() * sizeof(GroupDev), hipMemcpyHostToDevice, b->stream), "upload"))) return fail(rc);
  {
    std::vector<int64_t> z(b->groups.size(), 0);
    if ((rc = hip_check(hipMemcpyAsync(b->d_steps, z.data(), z.size() * 8, hipMemcpyHostToDevice, b->stream), "upload"))) return fail(rc);
  }
  // receivers
  b->seis_total = seis;
  if ((rc = dalloc(&b->d_seis, (size_t)std::max<int64_t>(seis, 1)))) return fail(rc);
  if ((rc = launch_fill_f32(std::max<int64_t>(seis, 1), b->d_seis, 0.f, b->stream))) return fail(rc);
  if ((rc = dalloc(&b->d_rec_w, rec_w.size()))) return fail(rc);
  if ((rc = dalloc(&b->d_rec_nodes, rec_nodes.size()))) return fail(rc);
  if (!rec_nodes.empty())
    if ((rc = hip_check(hipMemcpyAsync(b->d_rec_nodes, rec_nodes.data(), rec_nodes.size() * 4, hipMemcpyHostToDevice, b->stream), "upload"))) return fail(rc);
  if (!rec_w.empty())
    if ((rc = hip_check(hipMemcpyAsync(b->d_rec_w, rec_w.data(), rec_w.size() * 4, hipMemcpyHostToDevice, b->stream), "upload"))) return fail(rc);
  for (auto &kv : rec_key) {
    auto &lst = recs_by_key[kv.second];
    if (lst.empty()) continue;
    RecSet R;
    R.C = kv.first;  // low byte: components, bits 8+: nodes per element
    R.nrec = (int)lst.size();
    std::vector<RecDev> rd;
    std::vector<int32_t> rg;
    for (auto &pr : lst) {
      rd.push_back(pr.first);
      rg.push_back(pr.second);
    }
    if ((rc = dalloc(&R.d_recs, rd.size()))) return fail(rc);
    if ((rc = dalloc(&R.d_group, rg.size()))) return fail(rc);
    if ((rc = hip_check(hipMemcpyAsync(R.d_recs, rd.data(), rd.size() * sizeof(RecDev), hipMemcpyHostToDevice, b->stream), "upload"))) return fail(rc);
    if ((rc = hip_check(hipMemcpyAsync(R.d_group, rg.data(), rg.size() * 4, hipMemcpyHostToDevice, b->stream), "upload"))) return fail(rc);
    b->recsets.push_back(R);
  }
  if (path == SEM_PATH_PLAIN) {
    size_t mx = 0;
    for (auto &g : b->groups) mx = std::max(mx, (size_t)g.mesh->E * g.mesh->nd * g.mesh->c);
    if ((rc = dalloc(&b->d_uloc, mx))) return fail(rc);
    if ((rc = dalloc(&b->d_aloc, mx))) return fail(rc);
    b->epb = 0;
  } else {
    int nd_max = 0;
    for (auto &g : b->groups) nd_max = std::max(nd_max, g.mesh->nd);
    b->epb = epb > 0 ? epb : default_epb(nd_max);
    if (block_threads(nd_max, b->epb) > 512) { set_error("elements per block too large"); return fail(-1); }
    if ((rc = build_plan(b))) return fail(rc);
    if (epb <= 0) {
      // autotune elements per block on the stiffness-only pass
      double best = 1e300;
      int best_epb = b->epb;
      // one candidate per block size: the most elements that fit in bt threads
      for (int bt = 64; bt <= 512; bt += 64) {
        int cand = bt / nd_max;
        if (cand < 1) continue;
        b->epb = cand;
        if ((rc = build_plan(b))) return fail(rc);
        double t = time_shape(b, cand, 3);
        if (t < 0) return fail(-2);
        if (t < best * 0.98) { best = t; best_epb = cand; }
      }
      b->epb = best_epb;
      if ((rc = build_plan(b))) return fail(rc);
    }
    int64_t min_steps = INT64_MAX;
    for (auto &g : b->groups) min_steps = std::min(min_steps, std::max<int64_t>(g.nsteps, 1));
    b->block_steps = (int)std::min<int64_t>(std::max<int64_t>(1, 4096 / std::max<int>(1, (int)b->launches.size())), 50);
    b->block_steps = std::max(1, std::min<int>(b->block_steps, (int)min_steps));
  }
  b->spg = max_s;
  if ((rc = hip_check(hipStreamSynchronize(b->stream), "sync"))) return fail(rc);
  b->prepared = true;
  return 0;
}

int sem_batch_run(sem_batch *b, int64_t nsteps) {
  if (!b->prepared) { set_error("batch not prepared"); return -1; }
  HIPCHK(hipSetDevice(b->device));
  for (auto e : b->ev) (void)hipEventDestroy(e);
  b->ev.clear();
  int64_t left = nsteps;
  if (b->path == SEM_PATH_PLAIN) {
    for (int64_t k = 0; k < left; ++k) SEM_TRY(enqueue_plain_step(b));
  } else {
    const char *tg = getenv("SEM_TIME_KERNELS");
    b->timing = tg && tg[0] == '1';
    const char *ng = getenv("SEM_NO_GRAPH");
    const bool graphs = !(ng && ng[0] == '1') && !b->timing;
    while (left > 0) {
      if (graphs && left >= b->block_steps) {
        SEM_TRY(capture_graph(b));
        HIPCHK(hipGraphLaunch(b->graph, b->stream));
        left -= b->block_steps;
      } else {
        int k = (int)std::min<int64_t>(left, b->block_steps);
        for (int j = 0; j < k; ++j) SEM_TRY(enqueue_fused_step(b, j, 0));
        SEM_TRY(launch_advance((int)b->groups.size(), b->d_groups, b->d_steps, k, b->stream));
        left -= k;
      }
    }
  }
  HIPCHK(hipStreamSynchronize(b->stream));
  b->steps_done += nsteps;
  if (!b->ev.empty()) {
    double ms = 0;
    for (size_t i = 0; i + 1 < b->ev.size(); i += 2) {
      float t = 0;
      HIPCHK(hipEventElapsedTime(&t, b->ev[i], b->ev[i + 1]));
      ms += t;
    }
    b->kernel_ms = ms;
    b->kernel_launches = (int64_t)b->ev.size() / 2;
  }
  b->timing = false;
  return 0;
}

int sem_shot_progress(sem_batch *b, int shot, int64_t *steps, double *time) {
  Shot *s = get_shot(b, shot);
  if (!s) return -1;
  int64_t st = 0;
  if (b->prepared) {
    HIPCHK(hipSetDevice(b->device));
    HIPCHK(hipMemcpy(&st, b->d_steps + s->group, 8, hipMemcpyDeviceToHost));
  }
  *steps = st;
  double t = 0;
  for (int64_t k = 0; k < st; ++k) t += s->dt;
  *time = t;
  return 0;
}

int sem_shot_get_seismograms(sem_batch *b, int shot, float *out) {
  Shot *s = get_shot(b, shot);
  if (!s) return -1;
  if (!b->prepared) { set_error("batch not prepared"); return -1; }
  const size_t n = s->recs.size() * s->mesh->c * s->nsteps;
  if (n == 0) return 0;
  HIPCHK(hipSetDevice(b->device));
  HIPCHK(hipMemcpy(out, b->d_seis + s->seis_off, n * 4, hipMemcpyDeviceToHost));
  return 0;
}

int sem_batch_config(const sem_batch *b, int64_t *cfg) {
  cfg[0] = b->epb;
  cfg[1] = b->spg;
  cfg[2] = b->block_steps;
  cfg[3] = (int64_t)b->groups.size();
  return 0;
}

int sem_batch_kernel_time(const sem_batch *b, double *ms, int64_t *launches) {
  *ms = b->kernel_ms;
  *launches = b->kernel_launches;
  return 0;
}

int sem_mesh_apply_stiffness(sem_mesh *m, int path, const float *u, float *ku) {
  HIPCHK(hipSetDevice(m->device));
  sem_batch *b = nullptr;
  SEM_TRY(sem_batch_create(&b, m->device));
  int shot = 0;
  int rc = sem_batch_add_shot(b, m, 1, 1.0, &shot);
  if (!rc) rc = sem_batch_prepare(b, path, path == SEM_PATH_FUSED ? default_epb(m->nd) : 0, 1);
  const size_t len = (size_t)m->c * m->G;
  std::vector<float> uin(u, u + len);
  if (!rc && b->groups[0].tiled)
    for (size_t i = 0; i < len; ++i) uin[field_pos(b->groups[0], i)] = u[i];
  if (!rc) rc = hip_check(hipMemcpyAsync(b->h_groups[0].u, uin.data(), len * 4, hipMemcpyHostToDevice, b->stream), "upload");
  if (!rc) {
    if (path == SEM_PATH_FUSED) {
      rc = enqueue_fused_step(b, 0, 1);
    } else {
      GroupDev &h = b->h_groups[0];
      rc = launch_plain_gather(m->E, m->nd, m->c, m->G, m->d_conn, h.u, b->d_uloc, b->stream);
      PlainArgs a{};
      a.conn = m->d_conn;
      a.geo_p = m->d_geo_p;
      a.uloc = b->d_uloc;
      a.aloc = b->d_aloc;
      a.epb = default_epb(m->nd);
      a.E = m->E;
      if (!rc) rc = launch_plain_stiffness(m->dim, m->order, m->physics, m->E, m->nd, m->c, a, b->stream);
      if (!rc) rc = launch_fill_f32((int64_t)len, h.acc, 0.f, b->stream);
      for (int c = 0; c < m->ncol && !rc; ++c)
        rc = launch_plain_scatter(m->col_start[c], m->col_start[c + 1] - m->col_start[c], m->nd, m->c,
                                  m->G, m->d_conn, b->d_aloc, h.acc, b->stream);
    }
  }
  std::vector<float> acc(len);
  if (!rc) rc = hip_check(hipMemcpyAsync(acc.data(), b->h_groups[0].acc, len * 4, hipMemcpyDeviceToHost, b->stream), "copy");
  if (!rc) rc = hip_check(hipStreamSynchronize(b->stream), "sync");
  if (!rc)
    for (size_t i = 0; i < len; ++i) ku[i] = -acc[field_pos(b->groups[0], i)];
  sem_batch_destroy(b);
  return rc;
}

}  // extern "C"

static double time_shape(sem_batch *b, int epb, int reps) {
  (void)epb;
  hipEvent_t e0, e1;
  if (hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess) return -1;
  if (enqueue_fused_step(b, 0, 1)) return -1;  // warm
  (void)hipEventRecord(e0, b->stream);
  for (int r = 0; r < reps; ++r)
    if (enqueue_fused_step(b, 0, 1)) return -1;
  (void)hipEventRecord(e1, b->stream);
  (void)hipEventSynchronize(e1);
  float ms = 0;
  (void)hipEventElapsedTime(&ms, e0, e1);
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  return ms;
}
