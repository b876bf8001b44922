// Tiling of logically Cartesian meshes for the tile kernel.
//
// A tile is a box of Tx x Ty (x Tz) elements of the element lattice.  Its
// nodes form an (Tx N + 1) x (Ty N + 1) (x (Tz N + 1)) node lattice that the
// tile kernel stages in LDS; element (a,b,c) of the tile, node (i,j,k) sits
// at lattice slot (aN+i) + Lx ((bN+j) + Ly (cN+k)).  Tiles at the mesh edge
// keep the full box: their missing elements and slots are padding.  The
// slot -> global DOF table is filled from the DMPlex-compatible gather map
// and checked for consistency, so the tile path is used only when every
// element's reference axes follow the lattice axes.
//
// Tiles are coloured by lattice parity (2^dim colours; tiles of one colour
// share no node).  The words of a tile (node | flags) are stored in the
// order of their nodes, each with its lattice slot (slot_of), and in
// lattice order (words_lat, for the high-occupancy acoustic variant).  Nodes are renumbered tile-major by the tile that finishes
// them (the last colour holding the node): each tile's nodes in slot order,
// its inside first and then one run per lower/upper face, edge and corner, so
// that a tile's Newmark state and each piece its neighbours share are
// contiguous.
#include <algorithm>
#include <cmath>
#include <cstring>

#include "sem_internal.h"

namespace sem {

int build_tiles(sem_mesh *m, int Tx, int Ty, int Tz, TileSet &ts) {
  ts = TileSet();
  const int dim = m->dim, N = m->order, n = m->n;
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
  const int LX = Tx * N + 1, LY = Ty * N + 1, LZ = dim == 3 ? Tz * N + 1 : 1;
  const int64_t NTL = (int64_t)LX * LY * LZ;
  const int EPT = Tx * Ty * Tz;
  int64_t nt[3];
  for (int a = 0; a < 3; ++a) nt[a] = (nel[a] + T[a] - 1) / T[a];
  const int64_t ntiles = nt[0] * nt[1] * nt[2];
  if (ntiles * EPT > INT32_MAX || ntiles * NTL > ((int64_t)1 << 40)) return 1;
  ts.ntiles = ntiles;
  ts.slots_per_tile = NTL;
  ts.elems_per_tile = EPT;
  ts.total_slots = ntiles * NTL;
  ts.info.resize(ntiles);
  ts.colour_start.assign((1 << dim) + 1, 0);
  ts.elem_order.assign(ntiles * EPT, -1);
  // ---- per tile: element list (lattice order), slot -> PETSc node map
  std::vector<int32_t> slot_node(ntiles * NTL, -1);  // PETSc ids, -1 = padding
  for (int64_t tz = 0; tz < nt[2]; ++tz)
    for (int64_t ty = 0; ty < nt[1]; ++ty)
      for (int64_t tx = 0; tx < nt[0]; ++tx) {
        const int64_t t = tx + nt[0] * (ty + nt[1] * tz);
        TileInfo &ti = ts.info[t];
        ti.ex = (int)std::min<int64_t>(T[0], nel[0] - tx * T[0]);
        ti.ey = (int)std::min<int64_t>(T[1], nel[1] - ty * T[1]);
        ti.ez = dim == 3 ? (int)std::min<int64_t>(T[2], nel[2] - tz * T[2]) : 1;
        ti.elem_off = (int32_t)(t * EPT);
        ti.pad = 0;
        ti.node_off = t * NTL;
        ti.colour = (int32_t)((tx & 1) | ((ty & 1) << 1) | (dim == 3 ? ((tz & 1) << 2) : 0));
        for (int c = 0; c < ti.ez; ++c)
          for (int b = 0; b < ti.ey; ++b)
            for (int a = 0; a < ti.ex; ++a) {
              const int64_t cell = (tx * T[0] + a) + nel[0] * ((ty * T[1] + b) + nel[1] * (tz * T[2] + c));
              const int32_t e = at[cell];
              ts.elem_order[t * EPT + a + Tx * (b + Ty * c)] = e;
              for (int p = 0; p < m->nd; ++p) {
                const int i = p % n, j = (p / n) % n, k = dim == 3 ? p / (n * n) : 0;
                const int64_t s = (a * N + i) + (int64_t)LX * ((b * N + j) + (int64_t)LY * (c * N + k));
                const int32_t g = m->conn[(int64_t)e * m->nd + p];
                int32_t &dst = slot_node[t * NTL + s];
                if (dst < 0) dst = g;
                else if (dst != g) return 1;  // reference axes do not follow the lattice
              }
            }
      }
  // ---- per node: min/max colour of the tiles holding it (first/last flags)
  std::vector<int8_t> cmin(m->G, 127), cmax(m->G, -1);
  for (int64_t t = 0; t < ntiles; ++t)
    for (int64_t s = t * NTL; s < (t + 1) * NTL; ++s) {
      const int32_t g = slot_node[s];
      if (g < 0) continue;
      cmin[g] = std::min<int8_t>(cmin[g], (int8_t)ts.info[t].colour);
      cmax[g] = std::max<int8_t>(cmax[g], (int8_t)ts.info[t].colour);
    }
  // ---- internal numbering, tile-major: each node belongs to the tile that
  // finishes it (the last colour holding it), so a tile's Newmark reads and
  // writes are one contiguous range; in it the nodes come in runs by position
  // (inside, then each lower/upper face, edge and corner), so every piece a
  // neighbouring tile gathers or accumulates is contiguous too.
  ts.to_internal.assign(m->G, -1);
  int64_t next = 0;
  // Per tile the eligible slots are bucketed by category (a counting sort in
  // slot order), then numbered bucket by bucket.
  const int ncat = dim == 3 ? 27 : 9;
  std::vector<int32_t> slot_cat(NTL), order(NTL);
  std::vector<int64_t> start(ncat + 1);
  for (int64_t t = 0; t < ntiles; ++t) {
    const TileInfo &ti = ts.info[t];
    const int fx = ti.ex * N, fy = ti.ey * N, fz = ti.ez * N;  // last lattice index
    std::fill(start.begin(), start.end(), 0);
    for (int64_t s = 0; s < NTL; ++s) {
      const int32_t g = slot_node[t * NTL + s];
      slot_cat[s] = -1;
      if (g < 0 || ts.to_internal[g] >= 0 || cmax[g] != ti.colour) continue;
      const int x = (int)(s % LX), y = (int)((s / LX) % LY), z = (int)(s / ((int64_t)LX * LY));
      const int px = x == 0 ? 1 : x == fx ? 2 : 0, py = y == 0 ? 1 : y == fy ? 2 : 0;
      const int pz = dim == 3 ? (z == 0 ? 1 : z == fz ? 2 : 0) : 0;
      const int c = px + 3 * py + 9 * pz;
      slot_cat[s] = c;
      ++start[c + 1];
    }
    for (int c = 0; c < ncat; ++c) start[c + 1] += start[c];
    const int64_t cnt = start[ncat];
    for (int64_t s = 0; s < NTL; ++s)
      if (slot_cat[s] >= 0) order[start[slot_cat[s]]++] = (int32_t)s;
    for (int64_t i = 0; i < cnt; ++i) ts.to_internal[slot_node[t * NTL + order[i]]] = (int32_t)next++;
  }
  if (next != m->G) return 1;
  // ---- per tile, the slots in the order of their internal node (padding
  // last): the kernel's threads take the slots in this order, so the lanes
  // of a wave touch consecutive nodes (one run of the tile's own nodes, then
  // the runs shared with its neighbours) and every HBM access of the gather
  // and the finish is coalesced; slot_of gives each position's lattice slot.
  if (NTL > 65535) return 1;
  ts.words.assign(ntiles * NTL, kSlotEmpty);
  ts.words_lat.assign(ntiles * NTL, kSlotEmpty);
  ts.slot_of.assign(ntiles * NTL, 0);
  std::vector<std::pair<int64_t, int32_t>> key(NTL);
  for (int64_t t = 0; t < ntiles; ++t) {
    for (int64_t s = 0; s < NTL; ++s) {
      const int32_t g = slot_node[t * NTL + s];
      key[s] = {g < 0 ? (int64_t)1 << 40 : (int64_t)ts.to_internal[g], (int32_t)s};
    }
    std::sort(key.begin(), key.end());
    for (int64_t k = 0; k < NTL; ++k) {
      const int32_t s = key[k].second;
      ts.slot_of[t * NTL + k] = (uint16_t)s;
      const int32_t g = slot_node[t * NTL + s];
      if (g < 0) continue;
      uint32_t w = (uint32_t)ts.to_internal[g];
      if (cmin[g] == ts.info[t].colour) w |= kFlagFirst;
      if (cmax[g] == ts.info[t].colour) w |= kFlagLast;
      if (!m->dirichlet.empty() && m->dirichlet[g]) w |= kFlagDirichlet;
      ts.words[t * NTL + k] = w;
      ts.words_lat[t * NTL + s] = w;
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
  for (int64_t i = 0; i < ntiles * EPT; ++i)
    if (ts.elem_order[i] >= 0) ts.elem_pos[ts.elem_order[i]] = (int32_t)i;
  ts.T[0] = Tx; ts.T[1] = Ty; ts.T[2] = Tz;
  ts.valid = true;
  return 0;
}

}  // namespace sem
