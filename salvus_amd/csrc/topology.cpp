// DMPlex-compatible topology and gather map without PETSc, plus the element
// colouring that schedules the race-free scatter.
//
// Restates what Mesh::read(dim, ...) (Mesh.cpp:31-51: DMPlexCreateFromCellList
// with interpolation) and Mesh::setupGlobalDof (Mesh.cpp:361-452: section
// with 1 / N-1 / (N-1)^2 / (N-1)^3 DOFs per vertex / edge / face / cell,
// FixOrientation, DMPlexCreateSpectralClosurePermutation) produce:
//  * points: cells [0,E), vertices [E,E+V), faces (3D), edges; faces and
//    edges numbered in order of creation while walking cells (3D edges:
//    walking faces) and their raw faces/edges in PETSc's fixed order;
//  * DOF offsets in point order;
//  * closure gather: a node shared by several cells maps to one DOF whose
//    position inside its edge/face follows the edge/face's own cone.
#include <algorithm>
#include <cstring>

#include "sem_internal.h"

namespace sem {

namespace {

// Raw faces of a hex (bottom, top, front, back, right, left) as local vertices.
const int kHexFace[6][4] = {{0, 1, 2, 3}, {4, 5, 6, 7}, {0, 3, 5, 4},
                            {2, 1, 7, 6}, {3, 2, 6, 5}, {0, 4, 7, 1}};
// Local vertex -> corner in index space (0 or 1 along r, s, t).
const int kHexCorner[8][3] = {{0, 0, 0}, {0, 1, 0}, {1, 1, 0}, {1, 0, 0},
                              {0, 0, 1}, {1, 0, 1}, {1, 1, 1}, {0, 1, 1}};
const int kQuadCorner[4][2] = {{0, 0}, {1, 0}, {1, 1}, {0, 1}};

// Open-addressing table keyed by a sorted vertex tuple (up to 4 ids).
class TupleTable {
 public:
  explicit TupleTable(int64_t expected) {
    cap_ = 64;
    while (cap_ < 2 * expected) cap_ <<= 1;
    hi_.assign(cap_, 0);
    lo_.assign(cap_, 0);
    id_.assign(cap_, -1);
  }
  // Returns the existing id, or stores `fresh` and returns -1.
  int32_t find_or_add(uint64_t hi, uint64_t lo, int32_t fresh) {
    uint64_t h = hash(hi, lo);
    for (uint64_t i = h & (cap_ - 1);; i = (i + 1) & (cap_ - 1)) {
      if (id_[i] < 0) {
        hi_[i] = hi;
        lo_[i] = lo;
        id_[i] = fresh;
        return -1;
      }
      if (hi_[i] == hi && lo_[i] == lo) return id_[i];
    }
  }
  int32_t find(uint64_t hi, uint64_t lo) const {
    uint64_t h = hash(hi, lo);
    for (uint64_t i = h & (cap_ - 1);; i = (i + 1) & (cap_ - 1)) {
      if (id_[i] < 0) return -1;
      if (hi_[i] == hi && lo_[i] == lo) return id_[i];
    }
  }

 private:
  static uint64_t hash(uint64_t a, uint64_t b) {
    uint64_t x = a ^ (b * 0xbf58476d1ce4e5b9ULL);
    x ^= x >> 31;
    x *= 0x94d049bb133111ebULL;
    x ^= x >> 29;
    return x;
  }
  uint64_t cap_;
  std::vector<uint64_t> hi_, lo_;
  std::vector<int32_t> id_;
};

inline void pair_key(int32_t a, int32_t b, uint64_t &hi, uint64_t &lo) {
  uint64_t x = (uint32_t)std::min(a, b), y = (uint32_t)std::max(a, b);
  hi = (x << 32) | y;
  lo = 0;
}
inline void quad_key(const int32_t *v, uint64_t &hi, uint64_t &lo) {
  uint32_t s[4] = {(uint32_t)v[0], (uint32_t)v[1], (uint32_t)v[2], (uint32_t)v[3]};
  std::sort(s, s + 4);
  hi = ((uint64_t)s[0] << 32) | s[1];
  lo = ((uint64_t)s[2] << 32) | s[3];
}

}  // namespace

int build_topology(int dim, int order, int64_t E, int64_t V, const int32_t *cells, Topology &T) {
  if (dim != 2 && dim != 3) { set_error("dim must be 2 or 3"); return -1; }
  const int N = order, n = N + 1, nv = dim == 2 ? 4 : 8;
  const int64_t nd = dim == 2 ? n * n : n * n * n;
  for (int64_t i = 0; i < E * nv; ++i)
    if (cells[i] < 0 || cells[i] >= V) { set_error("cell references a vertex out of range"); return -1; }
  T.dim = dim;
  T.order = order;
  T.E = E;
  T.V = V;

  std::vector<int32_t> cell_entity;  // 2D: E x 4 edge ids; 3D: E x 6 face ids
  std::vector<int32_t> face_vtx;     // NF x 4 (3D)
  std::vector<int32_t> face_edge;    // NF x 4 (3D): edge id of raw edge k
  std::vector<int32_t> edge_vtx;     // NEd x 2
  TupleTable etab(dim == 2 ? 2 * E + 16 : 3 * E + 16);

  if (dim == 2) {
    cell_entity.resize(E * 4);
    edge_vtx.reserve(2 * (2 * E + 8));
    for (int64_t c = 0; c < E; ++c) {
      const int32_t *cv = cells + c * 4;
      for (int k = 0; k < 4; ++k) {
        uint64_t hi, lo;
        pair_key(cv[k], cv[(k + 1) & 3], hi, lo);
        int32_t fresh = (int32_t)(edge_vtx.size() / 2);
        int32_t id = etab.find_or_add(hi, lo, fresh);
        if (id < 0) {
          id = fresh;
          edge_vtx.push_back(cv[k]);
          edge_vtx.push_back(cv[(k + 1) & 3]);
        }
        cell_entity[c * 4 + k] = id;
      }
    }
    T.NF = 0;
  } else {
    TupleTable ftab(3 * E + 16);
    cell_entity.resize(E * 6);
    face_vtx.reserve(4 * (3 * E + 8));
    for (int64_t c = 0; c < E; ++c) {
      const int32_t *cv = cells + c * 8;
      for (int f = 0; f < 6; ++f) {
        int32_t fv[4] = {cv[kHexFace[f][0]], cv[kHexFace[f][1]], cv[kHexFace[f][2]],
                         cv[kHexFace[f][3]]};
        uint64_t hi, lo;
        quad_key(fv, hi, lo);
        int32_t fresh = (int32_t)(face_vtx.size() / 4);
        int32_t id = ftab.find_or_add(hi, lo, fresh);
        if (id < 0) {
          id = fresh;
          face_vtx.insert(face_vtx.end(), fv, fv + 4);
        }
        cell_entity[c * 6 + f] = id;
      }
    }
    const int64_t NF = (int64_t)face_vtx.size() / 4;
    T.NF = NF;
    face_edge.resize(NF * 4);
    edge_vtx.reserve(2 * (3 * E + 8));
    for (int64_t f = 0; f < NF; ++f) {
      const int32_t *fv = &face_vtx[f * 4];
      for (int k = 0; k < 4; ++k) {
        uint64_t hi, lo;
        pair_key(fv[k], fv[(k + 1) & 3], hi, lo);
        int32_t fresh = (int32_t)(edge_vtx.size() / 2);
        int32_t id = etab.find_or_add(hi, lo, fresh);
        if (id < 0) {
          id = fresh;
          edge_vtx.push_back(fv[k]);
          edge_vtx.push_back(fv[(k + 1) & 3]);
        }
        face_edge[f * 4 + k] = id;
      }
    }
  }
  T.NEd = (int64_t)edge_vtx.size() / 2;

  // DOF offsets in point order: cells, vertices, faces, edges.
  const int64_t dc = dim == 2 ? (int64_t)(N - 1) * (N - 1) : (int64_t)(N - 1) * (N - 1) * (N - 1);
  const int64_t off_vertex = E * dc;
  const int64_t off_face = off_vertex + V;
  const int64_t off_edge = off_face + T.NF * (int64_t)(N - 1) * (N - 1);
  T.G = off_edge + T.NEd * (N - 1);
  if (T.G >= (int64_t)kNodeMask) { set_error("mesh too large: more than 2^29 nodes"); return -1; }

  // Per local node: entity kind and local entity, computed once per order.
  struct NodeClass {
    int kind;    // 0 interior, 1 vertex, 2 edge, 3 face
    int ent;     // local vertex / face index, or edge endpoint la
    int lb;      // edge endpoint lb
    int P[3];    // index-space position
  };
  std::vector<NodeClass> cls(nd);
  for (int64_t q = 0; q < nd; ++q) {
    NodeClass &k = cls[q];
    k.P[0] = (int)(q % n);
    k.P[1] = (int)((q / n) % n);
    k.P[2] = dim == 3 ? (int)(q / (n * n)) : 0;
    int onb = 0;
    for (int a = 0; a < dim; ++a) onb += (k.P[a] == 0 || k.P[a] == N);
    auto corner = [&](int v, int a) { return (dim == 2 ? kQuadCorner[v][a] : kHexCorner[v][a]) * N; };
    if (onb == 0) {
      k.kind = 0;
    } else if (onb == dim) {
      k.kind = 1;
      for (int v = 0; v < nv; ++v) {
        bool ok = true;
        for (int a = 0; a < dim; ++a) ok = ok && corner(v, a) == k.P[a];
        if (ok) k.ent = v;
      }
    } else if (onb == dim - 1) {
      k.kind = 2;
      int fa = 0;
      for (int a = 0; a < dim; ++a)
        if (k.P[a] != 0 && k.P[a] != N) fa = a;
      for (int v = 0; v < nv; ++v) {
        bool at0 = true, atN = true;
        for (int a = 0; a < dim; ++a) {
          int want0 = a == fa ? 0 : k.P[a], wantN = a == fa ? N : k.P[a];
          at0 = at0 && corner(v, a) == want0;
          atN = atN && corner(v, a) == wantN;
        }
        if (at0) k.ent = v;
        if (atN) k.lb = v;
      }
    } else {
      k.kind = 3;
      int xa = 0;
      for (int a = 0; a < 3; ++a)
        if (k.P[a] == 0 || k.P[a] == N) xa = a;
      bool hi = k.P[xa] == N;
      k.ent = xa == 2 ? (hi ? 1 : 0) : xa == 1 ? (hi ? 3 : 2) : (hi ? 4 : 5);
    }
  }

  T.conn.assign(E * nd, 0);
  for (int64_t c = 0; c < E; ++c) {
    const int32_t *cv = cells + c * nv;
    int32_t *out = &T.conn[c * nd];
    auto local_of = [&](int32_t g) {
      for (int v = 0; v < nv; ++v)
        if (cv[v] == g) return v;
      return -1;
    };
    for (int64_t q = 0; q < nd; ++q) {
      const NodeClass &k = cls[q];
      int64_t g = 0;
      switch (k.kind) {
        case 0:
          g = c * dc + (k.P[0] - 1) + (int64_t)(N - 1) * ((k.P[1] - 1) + (dim == 3 ? (int64_t)(N - 1) * (k.P[2] - 1) : 0));
          break;
        case 1:
          g = off_vertex + cv[k.ent];
          break;
        case 2: {
          int32_t ga = cv[k.ent], gb = cv[k.lb];
          int32_t eid = -1;
          if (dim == 2) {
            for (int j = 0; j < 4; ++j) {
              int32_t e = cell_entity[c * 4 + j];
              int32_t a = edge_vtx[2 * e], b = edge_vtx[2 * e + 1];
              if ((a == ga && b == gb) || (a == gb && b == ga)) { eid = e; break; }
            }
          } else {
            for (int f = 0; f < 6 && eid < 0; ++f) {
              int32_t fid = cell_entity[c * 6 + f];
              for (int j = 0; j < 4; ++j) {
                int32_t e = face_edge[fid * 4 + j];
                int32_t a = edge_vtx[2 * e], b = edge_vtx[2 * e + 1];
                if ((a == ga && b == gb) || (a == gb && b == ga)) { eid = e; break; }
              }
            }
          }
          int fa = 0;
          for (int a = 0; a < dim; ++a)
            if (k.P[a] != 0 && k.P[a] != N) fa = a;
          int t = edge_vtx[2 * eid] == ga ? k.P[fa] : N - k.P[fa];
          g = off_edge + (int64_t)eid * (N - 1) + (t - 1);
          break;
        }
        default: {
          int32_t fid = cell_entity[c * 6 + k.ent];
          const int32_t *fv = &face_vtx[fid * 4];
          int l0 = local_of(fv[0]), l1 = local_of(fv[1]), l3 = local_of(fv[3]);
          int ia = 0, ja = 0;
          for (int a = 0; a < 3; ++a) {
            if (kHexCorner[l1][a] != kHexCorner[l0][a]) ia = a;
            if (kHexCorner[l3][a] != kHexCorner[l0][a]) ja = a;
          }
          int ii = kHexCorner[l0][ia] ? N - k.P[ia] : k.P[ia];
          int jj = kHexCorner[l0][ja] ? N - k.P[ja] : k.P[ja];
          g = off_face + (int64_t)fid * (N - 1) * (N - 1) + (ii - 1) + (int64_t)(N - 1) * (jj - 1);
        }
      }
      out[q] = (int32_t)g;
    }
  }
  return 0;
}

int colour_elements(int dim, int64_t E, int64_t V, const int32_t *cells,
                    std::vector<int32_t> &colour) {
  const int nv = dim == 2 ? 4 : 8;
  std::vector<int64_t> head(V + 1, 0);
  for (int64_t i = 0; i < E * nv; ++i) ++head[cells[i] + 1];
  for (int64_t v = 0; v < V; ++v) head[v + 1] += head[v];
  std::vector<int64_t> cursor(head.begin(), head.end() - 1);
  std::vector<int32_t> elems(E * nv);
  for (int64_t e = 0; e < E; ++e)
    for (int k = 0; k < nv; ++k) elems[cursor[cells[e * nv + k]]++] = (int32_t)e;
  colour.assign(E, -1);
  std::vector<int64_t> seen(256, -1);  // seen[colour] = last element that saw it
  for (int64_t e = 0; e < E; ++e) {
    for (int k = 0; k < nv; ++k) {
      int32_t v = cells[e * nv + k];
      for (int64_t q = head[v]; q < head[v + 1]; ++q) {
        int32_t o = elems[q];
        if (o < e) seen[colour[o]] = e;
      }
    }
    int c = 0;
    while (c < 256 && seen[c] == e) ++c;
    if (c == 256) { set_error("more than 256 colours"); return -1; }
    colour[e] = c;
  }
  return 0;
}

}  // namespace sem
