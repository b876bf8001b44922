// Tiling of logically Cartesian meshes for the tile kernel.
//
// A tile is a block of up to Tx x Ty (x Tz) elements of the element
// lattice.  Its nodes form an (ex N + 1) x (ey N + 1) (x (ez N + 1)) node
// lattice that the tile kernel stages in LDS; element (a,b,c) of the tile,
// node (i,j,k) sits at lattice slot (aN+i) + nx ((bN+j) + ny (cN+k)).  The
// slot -> global DOF table is filled from the DMPlex-compatible gather map
// and checked for consistency, so the tile path is used only when every
// element's reference axes follow the lattice axes.
//
// Tiles are coloured by lattice parity (2^dim colours; tiles of one colour
// share no node).  Nodes are renumbered tile-major (each tile's new nodes in
// slot order) so that a tile's gather and update touch contiguous memory.
#include <algorithm>
#include <cmath>
#include <cstring>

#include "sem_internal.h"

namespace sem {

int build_tiles(sem_mesh *m, int Tx, int Ty, int Tz, TileSet &ts) {
  ts = TileSet();
  const int dim = m->dim, N = m->order, n = m->n, nv = m->nv;
  const int64_t E = m->E;
  if (dim == 2) Tz = 1;
  // ---- element lattice from centroids and a common element size
  double lo[3] = {1e300, 1e300, 1e300}, h[3] = {0, 0, 0};
  for (int a = 0; a < dim; ++a) h[a] = m->bbox[dim + a] - m->bbox[a];
  for (int64_t e = 0; e < E; ++e)
    for (int a = 0; a < dim; ++a) {
      lo[a] = std::min(lo[a], m->bbox[e * 2 * dim + a]);
      double he = m->bbox[e * 2 * dim + dim + a] - m->bbox[e * 2 * dim + a];
      if (std::fabs(he - h[a]) > 1e-9 * std::max(1.0, std::fabs(h[a]))) return 1;  // not uniform
    }
  int64_t nel[3] = {1, 1, 1};
  std::vector<int64_t> lat(E * 3, 0);
  for (int64_t e = 0; e < E; ++e)
    for (int a = 0; a < dim; ++a) {
      double c = 0.5 * (m->bbox[e * 2 * dim + a] + m->bbox[e * 2 * dim + dim + a]);
      double f = (c - lo[a]) / h[a] - 0.5;
      int64_t k = (int64_t)std::llround(f);
      if (std::fabs(f - k) > 1e-6 || k < 0) return 1;
      lat[e * 3 + a] = k;
      nel[a] = std::max(nel[a], k + 1);
    }
  if (nel[0] * nel[1] * nel[2] != E) return 1;
  std::vector<int32_t> at(E, -1);  // lattice cell -> element
  for (int64_t e = 0; e < E; ++e) {
    int64_t id = lat[e * 3] + nel[0] * (lat[e * 3 + 1] + nel[1] * lat[e * 3 + 2]);
    if (at[id] >= 0) return 1;
    at[id] = (int32_t)e;
  }
  const int T[3] = {Tx, Ty, Tz};
  int64_t nt[3];
  for (int a = 0; a < 3; ++a) nt[a] = (nel[a] + T[a] - 1) / T[a];
  const int64_t ntiles = nt[0] * nt[1] * nt[2];
  ts.ntiles = ntiles;
  ts.info.resize(ntiles);
  ts.colour_start.assign((1 << dim) + 1, 0);
  // ---- per tile: element list (lattice order), slot -> PETSc node map
  std::vector<int32_t> slot_node;  // all tiles concatenated (PETSc ids)
  int64_t elem_off = 0, node_off = 0;
  ts.elem_order.reserve(E);
  for (int64_t tz = 0; tz < nt[2]; ++tz)
    for (int64_t ty = 0; ty < nt[1]; ++ty)
      for (int64_t tx = 0; tx < nt[0]; ++tx) {
        int64_t t = tx + nt[0] * (ty + nt[1] * tz);
        TileInfo &ti = ts.info[t];
        int ex = (int)std::min<int64_t>(T[0], nel[0] - tx * T[0]);
        int ey = (int)std::min<int64_t>(T[1], nel[1] - ty * T[1]);
        int ez = dim == 3 ? (int)std::min<int64_t>(T[2], nel[2] - tz * T[2]) : 1;
        ti.ex = ex; ti.ey = ey; ti.ez = ez;
        ti.elem_off = (int32_t)elem_off;
        ti.node_off = node_off;
        ti.colour = (int32_t)((tx & 1) | ((ty & 1) << 1) | (dim == 3 ? ((tz & 1) << 2) : 0));
        const int nx = ex * N + 1, ny = ey * N + 1, nz = dim == 3 ? ez * N + 1 : 1;
        const int64_t ntn = (int64_t)nx * ny * nz;
        slot_node.resize(node_off + ntn, -1);
        for (int c = 0; c < ez; ++c)
          for (int b = 0; b < ey; ++b)
            for (int a = 0; a < ex; ++a) {
              int64_t cell = (tx * T[0] + a) + nel[0] * ((ty * T[1] + b) + nel[1] * (tz * T[2] + c));
              int32_t e = at[cell];
              ts.elem_order.push_back(e);
              for (int p = 0; p < m->nd; ++p) {
                int i = p % n, j = (p / n) % n, k = dim == 3 ? p / (n * n) : 0;
                int64_t s = (a * N + i) + (int64_t)nx * ((b * N + j) + (int64_t)ny * (c * N + k));
                int32_t g = m->conn[(int64_t)e * m->nd + p];
                int32_t &dst = slot_node[node_off + s];
                if (dst < 0) dst = g;
                else if (dst != g) return 1;  // reference axes do not follow the lattice
              }
            }
        elem_off += (int64_t)ex * ey * ez;
        node_off += ntn;
      }
  ts.total_slots = node_off;
  for (int64_t s = 0; s < node_off; ++s)
    if (slot_node[s] < 0) return 1;
  (void)nv;
  // ---- internal numbering: tile-major, first appearance
  ts.to_internal.assign(m->G, -1);
  int64_t next = 0;
  for (int64_t t = 0; t < ntiles; ++t) {
    const TileInfo &ti = ts.info[t];
    int64_t ntn = (t + 1 < ntiles ? ts.info[t + 1].node_off : node_off) - ti.node_off;
    for (int64_t s = 0; s < ntn; ++s) {
      int32_t g = slot_node[ti.node_off + s];
      if (ts.to_internal[g] < 0) ts.to_internal[g] = (int32_t)next++;
    }
  }
  if (next != m->G) return 1;
  // ---- per node: min/max tile colour -> first/last flags per slot
  std::vector<int8_t> cmin(m->G, 127), cmax(m->G, -1);
  for (int64_t t = 0; t < ntiles; ++t) {
    const TileInfo &ti = ts.info[t];
    int64_t end = t + 1 < ntiles ? ts.info[t + 1].node_off : node_off;
    for (int64_t s = ti.node_off; s < end; ++s) {
      int32_t g = slot_node[s];
      cmin[g] = std::min<int8_t>(cmin[g], (int8_t)ti.colour);
      cmax[g] = std::max<int8_t>(cmax[g], (int8_t)ti.colour);
    }
  }
  ts.words.resize(node_off);
  for (int64_t t = 0; t < ntiles; ++t) {
    const TileInfo &ti = ts.info[t];
    int64_t end = t + 1 < ntiles ? ts.info[t + 1].node_off : node_off;
    for (int64_t s = ti.node_off; s < end; ++s) {
      int32_t g = slot_node[s];
      uint32_t w = (uint32_t)ts.to_internal[g];
      if (cmin[g] == ti.colour) w |= kFlagFirst;
      if (cmax[g] == ti.colour) w |= kFlagLast;
      if (!m->dirichlet.empty() && m->dirichlet[g]) w |= kFlagDirichlet;
      ts.words[s] = w;
    }
  }
  // ---- colour order of tiles
  for (int64_t t = 0; t < ntiles; ++t) ++ts.colour_start[ts.info[t].colour + 1];
  for (int c = 0; c < (1 << dim); ++c) ts.colour_start[c + 1] += ts.colour_start[c];
  ts.by_colour.resize(ntiles);
  {
    std::vector<int64_t> cur(ts.colour_start.begin(), ts.colour_start.end() - 1);
    for (int64_t t = 0; t < ntiles; ++t) ts.by_colour[cur[ts.info[t].colour]++] = (int32_t)t;
  }
  ts.elem_pos.assign(E, -1);
  for (int64_t i = 0; i < E; ++i) ts.elem_pos[ts.elem_order[i]] = (int32_t)i;
  ts.T[0] = Tx; ts.T[1] = Ty; ts.T[2] = Tz;
  ts.valid = true;
  return 0;
}

}  // namespace sem
