// 2-D fluid-solid coupling of two meshes cut from one quad mesh: the fluid
// (acoustic) and the solid (Elastic2D) elements.  The reference holds both
// field sets on every node and assembles one mass over all elements
// (Order2Newmark.cpp:7-66); fluid elements touch only fluid fields, solid
// elements only solid fields, so the problem is two single-physics problems
// joined by
//   - one inverse mass at the nodes on both meshes: 1 / (M_fluid + M_solid);
//   - coupling loads of the coupled elements (a vertex on an element of the
//     other physics, Mesh.cpp:505-521), edge integrals over their edges
//     (TensorQuad.cpp:474-515: edge nodes from vertex i to i+1, weights
//     w_j |edge| / 2, outward normal of the element):
//       fluid, every edge with a neighbouring element of either physics
//         (Mesh.cpp:480-503, ElasticToAcoustic.cpp:27-41):
//         s_f += (v_solid . n)  (0 at edge nodes that are not solid nodes)
//       solid, every edge whose neighbour is fluid
//         (AcousticToElastic2D.cpp:49-60):
//         s_x = s_y += -rho_0 v_fluid, rho_0 the neighbour's mean vertex RHO
//     with the velocities after the previous step.
// The loads are evaluated on the device before each step's colour launches
// (k_couple: one thread per target node summing its terms in a fixed order)
// into an external-load array that the step kernels add to the node's sum at
// its first colour.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <map>
#include <unordered_map>

#include "batch.h"

struct sem_coupling {
  sem_mesh *fluid = nullptr, *solid = nullptr;
  std::vector<float> mi_f, mi_s;  // combined inverse mass, DMPlex order
  // fluid targets (DMPlex ids): terms (solid node, half w_j n_x, half w_j n_y)
  std::vector<int32_t> f_tgt, f_ptr, f_src;
  std::vector<float> f_wx, f_wy;
  // solid targets: terms (fluid node, -rho_0 half w_j)
  std::vector<int32_t> s_tgt, s_ptr, s_src;
  std::vector<float> s_coef;
  int64_t shared = 0;
};

namespace sem {
namespace {

// First node and stride of local edge i (vertex i to vertex i+1) of a quad.
void edge_nodes(int i, int n, int *first, int *stride) {
  switch (i) {
    case 0: *first = 0; *stride = 1; break;
    case 1: *first = n - 1; *stride = n; break;
    case 2: *first = n * n - 1; *stride = -1; break;
    default: *first = n * (n - 1); *stride = -n; break;
  }
}

struct Key2 {
  uint64_t x, y;
  bool operator<(const Key2 &o) const { return x != o.x ? x < o.x : y < o.y; }
  bool operator==(const Key2 &o) const { return x == o.x && y == o.y; }
};
Key2 vkey(const double *p) {
  Key2 k;
  std::memcpy(&k.x, &p[0], 8);
  std::memcpy(&k.y, &p[1], 8);
  return k;
}

// Node coordinates of a mesh (DMPlex order): the GLL nodes of its elements.
std::vector<double> node_xy(const sem_mesh *m) {
  std::vector<double> xy((size_t)m->G * 2);
  sem_mesh_get_node_coords(m, xy.data());
  return xy;
}

// Spatial hash of points within tol of each other.
struct PointIndex {
  double tol;
  std::unordered_map<int64_t, std::vector<int32_t>> cells;
  const std::vector<double> *pts = nullptr;
  static int64_t key(int64_t i, int64_t j) { return (i << 32) ^ (j & 0xffffffffLL); }
  void build(const std::vector<double> &p, double t) {
    tol = t;
    pts = &p;
    for (int32_t k = 0; k < (int32_t)(p.size() / 2); ++k)
      cells[key((int64_t)std::floor(p[2 * k] / tol), (int64_t)std::floor(p[2 * k + 1] / tol))].push_back(k);
  }
  int32_t find(const double *q) const {
    const int64_t i0 = (int64_t)std::floor(q[0] / tol), j0 = (int64_t)std::floor(q[1] / tol);
    for (int64_t i = i0 - 1; i <= i0 + 1; ++i)
      for (int64_t j = j0 - 1; j <= j0 + 1; ++j) {
        auto it = cells.find(key(i, j));
        if (it == cells.end()) continue;
        for (int32_t k : it->second)
          if (std::fabs((*pts)[2 * k] - q[0]) <= tol && std::fabs((*pts)[2 * k + 1] - q[1]) <= tol) return k;
      }
    return -1;
  }
};

// CSR of terms grouped by target in first-appearance order of the targets,
// terms of a target in insertion order.
template <class Term>
void to_csr(const std::vector<std::pair<int32_t, Term>> &terms, std::vector<int32_t> &tgt,
            std::vector<int32_t> &ptr, std::vector<Term> &out) {
  std::map<int32_t, int32_t> slot;
  std::vector<std::vector<Term>> lists;
  for (auto &t : terms) {
    auto it = slot.find(t.first);
    if (it == slot.end()) {
      it = slot.emplace(t.first, (int32_t)lists.size()).first;
      tgt.push_back(t.first);
      lists.emplace_back();
    }
    lists[it->second].push_back(t.second);
  }
  ptr.assign(1, 0);
  for (auto &l : lists) {
    out.insert(out.end(), l.begin(), l.end());
    ptr.push_back((int32_t)out.size());
  }
}

__global__ void k_couple(CoupleDev c) {
  const int32_t t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t < c.nf) {
    float s = 0.f;
    for (int32_t k = c.f_ptr[t]; k < c.f_ptr[t + 1]; ++k) {
      const int32_t i = c.f_src[k];
      const float vx = c.ws[i] - c.acls * c.aps[i];
      const float vy = c.ws[i + c.Gs] - c.acls * c.aps[i + c.Gs];
      s = fmaf(c.f_wx[k], vx, s);
      s = fmaf(c.f_wy[k], vy, s);
    }
    c.extf[c.f_tgt[t]] = s;
  } else if (t < c.nf + c.ns) {
    const int32_t r = t - c.nf;
    float s = 0.f;
    for (int32_t k = c.s_ptr[r]; k < c.s_ptr[r + 1]; ++k) {
      const int32_t i = c.s_src[k];
      s = fmaf(c.s_coef[k], c.wf[i] - c.aclf * c.apf[i], s);
    }
    c.exts[c.s_tgt[r]] = s;
    c.exts[c.s_tgt[r] + c.Gs] = s;
  }
}

}  // namespace

const std::vector<float> &coupled_mass(const CouplePair &p, const sem_mesh *m) {
  return m == p.cp->fluid ? p.cp->mi_f : p.cp->mi_s;
}

void free_couples(sem_batch *b) {
  for (auto &p : b->couples)
    if (p.dev) {
      (void)hipFree(p.dev);
      p.dev = nullptr;
    }
}

int setup_couples(sem_batch *b) {
  free_couples(b);
  for (auto &p : b->couples) {
    const Shot &fs = b->shots[p.fluid_shot], &ss = b->shots[p.solid_shot];
    if (fs.group < 0 || ss.group < 0) { set_error("coupled shots are not set up"); return -1; }
    const Group &gf = b->groups[fs.group], &gs = b->groups[ss.group];
    const GroupDev &hf = b->h_groups[fs.group], &hs = b->h_groups[ss.group];
    const sem_coupling *cp = p.cp;
    const int64_t Gs = cp->solid->G;
    // field indices in the two groups' numbering (one shot per group)
    std::vector<int32_t> f_tgt(cp->f_tgt.size()), f_src(cp->f_src.size());
    std::vector<int32_t> s_tgt(cp->s_tgt.size()), s_src(cp->s_src.size());
    for (size_t i = 0; i < f_tgt.size(); ++i) f_tgt[i] = (int32_t)field_pos(gf, cp->f_tgt[i]);
    for (size_t i = 0; i < f_src.size(); ++i) f_src[i] = (int32_t)field_pos(gs, cp->f_src[i]);
    for (size_t i = 0; i < s_tgt.size(); ++i) s_tgt[i] = (int32_t)field_pos(gs, cp->s_tgt[i]);
    for (size_t i = 0; i < s_src.size(); ++i) s_src[i] = (int32_t)field_pos(gf, cp->s_src[i]);
    // one device block: int arrays, then float arrays
    const size_t ni = f_tgt.size() + cp->f_ptr.size() + f_src.size() + s_tgt.size() + cp->s_ptr.size() +
                      s_src.size();
    const size_t nf = cp->f_wx.size() + cp->f_wy.size() + cp->s_coef.size();
    std::vector<char> host((ni + nf) * 4 + 16);
    char *dev = nullptr;
    HIPCHK(hipMalloc((void **)&dev, host.size()));
    p.dev = dev;
    size_t off = 0;
    auto put = [&](const void *src, size_t n) {
      std::memcpy(host.data() + off, src, n * 4);
      const char *d = dev + off;
      off += n * 4;
      return d;
    };
    CoupleDev &c = p.h;
    c.nf = (int32_t)f_tgt.size();
    c.ns = (int32_t)s_tgt.size();
    c.f_tgt = (const int32_t *)put(f_tgt.data(), f_tgt.size());
    c.f_ptr = (const int32_t *)put(cp->f_ptr.data(), cp->f_ptr.size());
    c.f_src = (const int32_t *)put(f_src.data(), f_src.size());
    c.s_tgt = (const int32_t *)put(s_tgt.data(), s_tgt.size());
    c.s_ptr = (const int32_t *)put(cp->s_ptr.data(), cp->s_ptr.size());
    c.s_src = (const int32_t *)put(s_src.data(), s_src.size());
    c.f_wx = (const float *)put(cp->f_wx.data(), cp->f_wx.size());
    c.f_wy = (const float *)put(cp->f_wy.data(), cp->f_wy.size());
    c.s_coef = (const float *)put(cp->s_coef.data(), cp->s_coef.size());
    c.Gs = Gs;
    c.wf = hf.w;
    c.apf = hf.ap;
    c.ws = hs.w;
    c.aps = hs.ap;
    c.extf = gf.ext;
    c.exts = gs.ext;
    c.aclf = hf.acl;
    c.acls = hs.acl;
    HIPCHK(hipMemcpyAsync(dev, host.data(), off, hipMemcpyHostToDevice, b->stream));
    HIPCHK(hipStreamSynchronize(b->stream));
  }
  return 0;
}

int enqueue_couples(sem_batch *b) {
  for (auto &p : b->couples) {
    const int n = p.h.nf + p.h.ns;
    if (n <= 0) continue;
    hipLaunchKernelGGL(k_couple, dim3((n + 255) / 256), dim3(256), 0, b->stream, p.h);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) { set_error(std::string("coupling: ") + hipGetErrorString(e)); return -2; }
  }
  return 0;
}

}  // namespace sem

using namespace sem;

extern "C" {

int sem_coupling_create(sem_coupling **out, sem_mesh *fluid, sem_mesh *solid, const double *rho) {
  *out = nullptr;
  if (!fluid || !solid || !rho) { set_error("null argument"); return -1; }
  if (fluid->dim != 2 || solid->dim != 2 || fluid->physics != SEM_ACOUSTIC ||
      solid->physics != SEM_ELASTIC || fluid->order != solid->order || fluid->device != solid->device) {
    set_error("coupling needs a 2-D acoustic and a 2-D elastic mesh of one order on one device");
    return -1;
  }
  auto *cp = new sem_coupling();
  cp->fluid = fluid;
  cp->solid = solid;
  const int n = fluid->n, N = fluid->order;
  const Gll gl = make_gll(N);
  // ---- nodes on both meshes, and the combined inverse mass
  const std::vector<double> xf = node_xy(fluid), xs = node_xy(solid);
  double ext = 0;
  for (double v : xf) ext = std::max(ext, std::fabs(v));
  for (double v : xs) ext = std::max(ext, std::fabs(v));
  PointIndex idx;
  idx.build(xs, 1e-9 * std::max(ext, 1.0));
  std::vector<int32_t> f2s(fluid->G, -1), s2f(solid->G, -1);
  for (int64_t g = 0; g < fluid->G; ++g) {
    const int32_t k = idx.find(&xf[2 * g]);
    if (k >= 0) {
      f2s[g] = k;
      s2f[k] = (int32_t)g;
      ++cp->shared;
    }
  }
  cp->mi_f.resize(fluid->G);
  cp->mi_s.resize(solid->G);
  int rc = sem_mesh_get_inverse_mass(fluid, cp->mi_f.data());
  if (!rc) rc = sem_mesh_get_inverse_mass(solid, cp->mi_s.data());
  if (rc) { delete cp; return rc; }
  for (int64_t g = 0; g < fluid->G; ++g) {
    const int32_t k = f2s[g];
    if (k < 0) continue;
    const double m = 1.0 / (double)cp->mi_f[g] + 1.0 / (double)cp->mi_s[k];
    cp->mi_f[g] = cp->mi_s[k] = (float)(1.0 / m);
  }
  // ---- element edges: owners by vertex pair, vertex sets of each physics
  std::map<std::pair<Key2, Key2>, std::vector<std::pair<int, int64_t>>> owners;  // (0 fluid / 1 solid, elem)
  std::map<Key2, int> vset;  // bit 0: on a fluid element, bit 1: on a solid element
  const sem_mesh *ms[2] = {fluid, solid};
  for (int ph = 0; ph < 2; ++ph)
    for (int64_t e = 0; e < ms[ph]->E; ++e) {
      const double *v = &ms[ph]->vtx[e * 8];
      for (int i = 0; i < 4; ++i) {
        Key2 a = vkey(v + 2 * i), b = vkey(v + 2 * ((i + 1) % 4));
        if (b < a) std::swap(a, b);
        owners[{a, b}].push_back({ph, e});
        vset[vkey(v + 2 * i)] |= 1 << ph;
      }
    }
  auto coupled = [&](int ph, int64_t e) {
    const double *v = &ms[ph]->vtx[e * 8];
    for (int i = 0; i < 4; ++i)
      if (vset[vkey(v + 2 * i)] & (1 << (1 - ph))) return true;
    return false;
  };
  struct FTerm { int32_t src; float wx, wy; };
  struct STerm { int32_t src; float coef; };
  std::vector<std::pair<int32_t, FTerm>> fterms;
  std::vector<std::pair<int32_t, STerm>> sterms;
  for (int ph = 0; ph < 2; ++ph) {
    const sem_mesh *m = ms[ph];
    for (int64_t e = 0; e < m->E; ++e) {
      if (!coupled(ph, e)) continue;
      const double *v = &m->vtx[e * 8];
      for (int i = 0; i < 4; ++i) {
        const double *p0 = v + 2 * i, *p1 = v + 2 * ((i + 1) % 4);
        Key2 a = vkey(p0), b = vkey(p1);
        if (b < a) std::swap(a, b);
        const double dx = p1[0] - p0[0], dy = p1[1] - p0[1], len = std::hypot(dx, dy);
        const double half = len / 2.0, nx = dy / len, ny = -dx / len;  // outward (counter-clockwise cells)
        int first, stride;
        edge_nodes(i, n, &first, &stride);
        for (auto &o : owners[{a, b}]) {
          if (o.first == ph && o.second == e) continue;
          if (ph == 1 && o.first != 0) continue;  // the solid couples to fluid neighbours only
          double rho0 = 0;
          if (ph == 1) {
            for (int k = 0; k < 4; ++k) rho0 += rho[o.second * 4 + k];
            rho0 /= 4.0;
          }
          for (int j = 0; j < n; ++j) {
            const int q = first + j * stride;
            const int32_t g = m->conn[e * m->nd + q];
            if (ph == 0) {
              if (f2s[g] >= 0)
                fterms.push_back({g, FTerm{f2s[g], (float)(half * gl.w[j] * nx), (float)(half * gl.w[j] * ny)}});
            } else if (s2f[g] >= 0) {
              sterms.push_back({g, STerm{s2f[g], (float)(-rho0 * half * gl.w[j])}});
            }
          }
        }
      }
    }
  }
  std::vector<FTerm> fl;
  std::vector<STerm> sl;
  to_csr(fterms, cp->f_tgt, cp->f_ptr, fl);
  to_csr(sterms, cp->s_tgt, cp->s_ptr, sl);
  for (auto &t : fl) { cp->f_src.push_back(t.src); cp->f_wx.push_back(t.wx); cp->f_wy.push_back(t.wy); }
  for (auto &t : sl) { cp->s_src.push_back(t.src); cp->s_coef.push_back(t.coef); }
  *out = cp;
  return 0;
}

int sem_coupling_destroy(sem_coupling *cp) {
  delete cp;
  return 0;
}

int sem_coupling_info(const sem_coupling *cp, int64_t *info) {
  info[0] = cp->shared;
  info[1] = (int64_t)cp->f_tgt.size();
  info[2] = (int64_t)cp->s_tgt.size();
  info[3] = (int64_t)(cp->f_src.size() + cp->s_src.size());
  return 0;
}

int sem_batch_couple(sem_batch *b, sem_coupling *cp, int fshot, int sshot) {
  if (!b || !cp) { set_error("null argument"); return -1; }
  if (b->prepared) { set_error("couple shots before sem_batch_prepare"); return -1; }
  if (fshot < 0 || sshot < 0 || fshot >= (int)b->shots.size() || sshot >= (int)b->shots.size() ||
      fshot == sshot) {
    set_error("shot index out of range");
    return -1;
  }
  Shot &f = b->shots[fshot], &s = b->shots[sshot];
  if (f.mesh != cp->fluid || s.mesh != cp->solid) {
    set_error("the shots are not on the coupling's fluid and solid meshes");
    return -1;
  }
  if (f.nsteps != s.nsteps || f.dt != s.dt) { set_error("coupled shots need the same dt and step count"); return -1; }
  if (f.coupled >= 0 || s.coupled >= 0) { set_error("shot already coupled"); return -1; }
  CouplePair p;
  p.cp = cp;
  p.fluid_shot = fshot;
  p.solid_shot = sshot;
  f.coupled = s.coupled = (int)b->couples.size();
  b->couples.push_back(p);
  return 0;
}

}  // extern "C"
