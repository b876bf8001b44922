// Batches: shots advanced together through the plain path (one shot per
// launch: gather / stiffness / colour scatter / update) or the batched path
// (tile or element colour kernels over every group of the batch, in
// hipGraph-captured step blocks).  The time loop is Main.cpp:45-66; the
// update is Order2Newmark.cpp:68-115.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <map>
#include <set>
#include <tuple>

#include "batch.h"

using namespace sem;

namespace sem {

int64_t field_pos(const Group &g, int64_t i) {
  if (!g.tiled) return i;
  const sem_mesh *m = g.mesh;
  const int64_t c = i / m->G, node = i - c * m->G;
  return c * m->G + m->tiles.to_internal[node];
}

void free_group(Group &g) {
  dfree(g.fields);
  dfree(g.d_src);
  dfree(g.d_stf);
  dfree(g.d_load);
  dfree(g.d_step);
  dfree(g.d_seis);
  dfree(g.ext);
  dfree(g.mi_own);
  g.recs.clear();
  g.rec_nodes.clear();
  g.rec_w.clear();
  g.bytes = 0;
}

void drop_graph(sem_batch *b) {
  if (b->graph) (void)hipGraphExecDestroy(b->graph);
  if (b->graph_def) (void)hipGraphDestroy(b->graph_def);
  if (b->tail) (void)hipGraphExecDestroy(b->tail);
  if (b->tail_def) (void)hipGraphDestroy(b->tail_def);
  b->graph = nullptr;
  b->graph_def = nullptr;
  b->tail = nullptr;
  b->tail_def = nullptr;
  b->tail_steps = 0;
}

// Default launch shape: elements per block so a block has ~128-256 threads.
int default_epb(int nd) {
  int best = 1;
  for (int epb = 1; epb * nd <= 512; ++epb)
    if (block_threads(nd, epb) <= 256) best = epb;
  return best;
}

int setup_group(sem_batch *b, int gi) {
  Group &g = b->groups[gi];
  sem_mesh *m = g.mesh;
  const int path = b->path;
  int rc = 0;
  // tile kernel: 32-bit field offsets and at most kTileMaxSrc sources per tile
  g.tiled = path == SEM_PATH_FUSED && m->tiles.valid &&
            (uint64_t)m->c * (uint64_t)m->G * (uint64_t)g.S < (1ull << 32);
  if (g.tiled) {
    std::map<int64_t, int> per_tile;
    for (int sl = 0; sl < g.S && g.tiled; ++sl)
      for (auto &sp : b->shots[g.shots[sl]].srcs) {
        int64_t e;
        std::vector<double> ld;
        SEM_TRY(source_load(m, sp, &e, ld));
        int &cnt = per_tile[m->tiles.elem_pos[e] / m->tiles.elems_per_tile];
        cnt += sp.kind == 2 ? m->c : 1;  // a file source is one source per component
        if (cnt > kTileMaxSrc) {
          g.tiled = false;
          break;
        }
      }
  }
  const size_t len = (size_t)m->c * m->G * g.S;
  SEM_TRY(dalloc(&g.fields, 5 * len, &g.bytes));
  std::vector<float> buf;
  // u, w, a_, kc, acc; kc = 0 is two bf16 zeros; w = v + dt/2 a_ from the
  // shot's initial v and a_
  const double half_dt = 0.5 * g.dt;
  for (int which = 0; which < 5; ++which) {
    bool any = false;
    for (int sl = 0; sl < g.S && which < 3; ++sl) {
      const Shot &s = b->shots[g.shots[sl]];
      any |= !s.init[which].empty() || (which == 1 && !s.init[2].empty());
    }
    if (!any) {  // zero start: a fill kernel, no host buffer
      SEM_TRY(launch_fill_f32((int64_t)len, g.fields + which * len, 0.f, b->stream));
      continue;
    }
    buf.assign(len, 0.f);
    for (int sl = 0; sl < g.S; ++sl) {
      const Shot &s = b->shots[g.shots[sl]];
      const size_t cg = (size_t)m->c * m->G;
      const std::vector<double> &v0 = s.init[1], &a0 = s.init[2];
      if (which == 1) {
        if (v0.empty() && a0.empty()) continue;
        for (size_t i = 0; i < cg; ++i)
          buf[sl * cg + field_pos(g, i)] =
              (float)((v0.empty() ? 0.0 : v0[i]) + (a0.empty() ? 0.0 : half_dt * a0[i]));
        continue;
      }
      if (s.init[which].empty()) continue;
      for (size_t i = 0; i < cg; ++i) buf[sl * cg + field_pos(g, i)] = (float)s.init[which][i];
    }
    HIPCHK(hipMemcpyAsync(g.fields + which * len, buf.data(), len * 4, hipMemcpyHostToDevice, b->stream));
    HIPCHK(hipStreamSynchronize(b->stream));
  }
  // step counter
  SEM_TRY(dalloc(&g.d_step, 1, &g.bytes));
  {
    const int64_t zero = 0;
    HIPCHK(hipMemcpyAsync(g.d_step, &zero, 8, hipMemcpyHostToDevice, b->stream));
    HIPCHK(hipStreamSynchronize(b->stream));
  }
  // sources: STF tables in FP64 at the accumulated times t_n (Main.cpp:45,
  // Order2Newmark.cpp:112), stored FP32
  std::vector<SrcDev> src;
  std::vector<float> stf, load;
  std::vector<int64_t> rec_out;  // seismogram offsets of g.recs
  int64_t seis = 0;
  g.rec_key = m->c | (m->nd << 8);
  const Gll gl = make_gll(m->order);
  for (int sl = 0; sl < g.S; ++sl) {
    Shot &s = b->shots[g.shots[sl]];
    s.group = gi;
    s.slot = sl;
    for (auto &sp0 : s.srcs) {
      // a file source is one unit-direction source per component with that
      // component's tabulated series as its source-time function
      const int parts = sp0.kind == 2 ? m->c : 1;
      for (int k = 0; k < parts; ++k) {
        SrcSpec sp = sp0;
        if (sp0.kind == 2) {
          sp.kind = 0;
          for (int q = 0; q < 3; ++q) sp.dir[q] = q == k ? 1.0 : 0.0;
        }
        int64_t e;
        std::vector<double> ld;
        SEM_TRY(source_load(m, sp, &e, ld));
        SrcDev d;
        d.elem = g.tiled ? m->tiles.elem_pos[e] : m->inv_perm[e];
        d.stf_off = (int32_t)stf.size();
        d.load_off = (int32_t)load.size();
        d.shot = sl;
        if (sp0.kind == 2) {
          // (a shot shorter than its group's step count: zeros past its end)
          for (int64_t n = 0; n < g.nsteps; ++n)
            stf.push_back(n < sp0.nser ? (float)sp0.series[k * sp0.nser + n] : 0.f);
        } else {
          double t = 0.0;
          for (int64_t n = 0; n < g.nsteps; ++n) {
            stf.push_back((float)ricker(t, sp.amp, sp.f0, sp.t0));
            t += g.dt;
          }
        }
        for (double v : ld) load.push_back((float)v);
        src.push_back(d);
      }
    }
    // receivers: Lagrange weights of the owning element at the receiver
    s.seis_off = seis;
    for (size_t r = 0; r < s.recs.size(); ++r) {
      int64_t e;
      double ref[3];
      SEM_TRY(locate(m, s.recs[r].data(), &e, ref));
      RecDev d;
      d.nodes_off = (int32_t)g.rec_nodes.size();
      for (int p = 0; p < m->nd; ++p) {
        const int32_t gid = m->conn[e * m->nd + p];
        g.rec_nodes.push_back((uint32_t)(g.tiled ? m->tiles.to_internal[gid] : gid));
      }
      d.w_off = (int32_t)g.rec_w.size();
      d.shot = sl;
      d.stride = (int32_t)(s.recs.size() * m->c);
      d.out = nullptr;
      rec_out.push_back(seis + (int64_t)r * m->c);
      double l[3][kMaxN];
      for (int a = 0; a < m->dim; ++a) lagrange(gl, ref[a], l[a]);
      for (int p = 0; p < m->nd; ++p) {
        const int P[3] = {p % m->n, (p / m->n) % m->n, m->dim == 3 ? p / (m->n * m->n) : 0};
        double v = 1;
        for (int a = 0; a < m->dim; ++a) v *= l[a][P[a]];
        g.rec_w.push_back((float)v);
      }
      g.recs.push_back(d);
    }
    // rows for every step of the group (a server groups shots whose step
    // counts differ a little; a shot's own rows are the first s.nsteps)
    seis += (int64_t)s.recs.size() * m->c * g.nsteps;
  }
  g.seis_len = seis;
  SEM_TRY(dalloc(&g.d_seis, (size_t)std::max<int64_t>(seis, 1), &g.bytes));
  SEM_TRY(launch_fill_f32(std::max<int64_t>(seis, 1), g.d_seis, 0.f, b->stream));
  for (size_t r = 0; r < g.recs.size(); ++r) g.recs[r].out = g.d_seis + rec_out[r];
  g.nsrc = (int)src.size();
  SEM_TRY(dalloc(&g.d_src, src.size(), &g.bytes));
  SEM_TRY(dalloc(&g.d_stf, stf.size(), &g.bytes));
  SEM_TRY(dalloc(&g.d_load, load.size(), &g.bytes));
  if (!src.empty()) {
    HIPCHK(hipMemcpyAsync(g.d_src, src.data(), src.size() * sizeof(SrcDev), hipMemcpyHostToDevice, b->stream));
    HIPCHK(hipMemcpyAsync(g.d_stf, stf.data(), stf.size() * 4, hipMemcpyHostToDevice, b->stream));
    HIPCHK(hipMemcpyAsync(g.d_load, load.data(), load.size() * 4, hipMemcpyHostToDevice, b->stream));
  }
  if (path == SEM_PATH_PLAIN || (!m->affine && !g.tiled)) SEM_TRY(ensure_pointwise(m));
  GroupDev h{};
  h.conn = m->d_conn;
  if (g.tiled) {
    h.tiles = m->d_tile_info;
    h.tile_words = m->d_tile_words;
    h.tile_slot = m->d_tile_slot;
    h.tile_words_lat = m->d_tile_words_lat;
    h.telem = m->d_telem;
  }
  h.geo_e = m->d_geo_e;
  h.par_e = m->d_par_e;
  h.geo_p = m->d_geo_p;
  h.mi = g.tiled ? m->d_mi_int : m->d_mi;
  h.u = g.fields;
  h.w = g.fields + len;
  h.ap = g.fields + 2 * len;
  h.kc = reinterpret_cast<uint32_t *>(g.fields + 3 * len);
  h.acc = g.fields + 4 * len;
  h.src = g.d_src;
  h.stf = g.d_stf;
  h.load = g.d_load;
  h.nsrc = g.nsrc;
  h.S = g.S;
  h.G = m->G;
  h.nsteps = g.nsteps;
  h.step = g.d_step;
  h.dt = (float)g.dt;
  h.acl = (float)((1.0 / 2.0) * g.dt);           // Order2Newmark.cpp:89
  h.quirks = m->quirks;
  // a coupled shot (a group of its own): external load and combined mass
  {
    const Shot &s0 = b->shots[g.shots[0]];
    if (s0.coupled >= 0) {
      const std::vector<float> &mc = coupled_mass(b->couples[s0.coupled], m);
      std::vector<float> mo(m->G);
      for (int64_t i = 0; i < m->G; ++i) mo[field_pos(g, i)] = mc[i];
      SEM_TRY(dalloc(&g.mi_own, (size_t)m->G, &g.bytes));
      SEM_TRY(dalloc(&g.ext, len, &g.bytes));
      SEM_TRY(launch_fill_f32((int64_t)len, g.ext, 0.f, b->stream));
      HIPCHK(hipMemcpyAsync(g.mi_own, mo.data(), mo.size() * 4, hipMemcpyHostToDevice, b->stream));
      HIPCHK(hipStreamSynchronize(b->stream));
      h.mi = g.mi_own;
      h.ext = g.ext;
      h.ap_always = 1;
    }
  }
  if ((int)b->h_groups.size() <= gi) b->h_groups.resize(gi + 1);
  b->h_groups[gi] = h;
  HIPCHK(hipStreamSynchronize(b->stream));
  return rc;
}

static void release_plan(sem_batch *b) {
  for (auto &l : b->launches) dfree(l.d_items);
  b->launches.clear();
  for (auto &r : b->recsets) {
    dfree(r.d_recs);
    dfree(r.d_group);
    dfree(r.d_nodes);
    dfree(r.d_w);
  }
  b->recsets.clear();
}

static void batch_release(sem_batch *b) {
  drop_graph(b);
  free_couples(b);
  for (auto &g : b->groups) free_group(g);
  b->groups.clear();
  b->h_groups.clear();
  release_plan(b);
  dfree(b->d_groups);
  b->d_groups_cap = 0;
  dfree(b->d_uloc);
  dfree(b->d_aloc);
  b->uloc_cap = 0;
  for (auto e : b->ev) (void)hipEventDestroy(e);
  b->ev.clear();
  b->ev_launches = 0;
  b->prepared = false;
}

// Work items per (kernel key, colour): tiles of tiled groups, EPB-element
// slices of the element kernel's colour ranges otherwise.
static int build_plan(sem_batch *b) {
  // key = (lane, tile? | diagonal geometry << 1, dim, order, physics,
  // geometry mode, tile shape); SEM_TILE_GENERAL=1 keeps the general tile
  // kernel
  const char *gen = getenv("SEM_TILE_GENERAL");
  const bool general = gen && gen[0] == '1';
  // lanes: SEM_LANES=k (1..4) splits the groups over k streams (not with
  // coupled pairs, whose two groups exchange loads every step)
  const char *ln = getenv("SEM_LANES");
  const int ngroups = (int)b->groups.size();
  std::vector<int> lane(ngroups, 0);
  b->lanes = ln && atoi(ln) > 0 ? std::min(atoi(ln), kMaxLanes) : 1;
  if (!b->couples.empty()) b->lanes = 1;
  b->lanes = std::min(b->lanes, std::max(1, ngroups));
  if (b->lanes > 1) {
    // largest first onto the lightest lane (work: elements x shots)
    std::vector<int> order(ngroups);
    for (int gi = 0; gi < ngroups; ++gi) order[gi] = gi;
    auto work = [&](int gi) { return (double)b->groups[gi].mesh->E * b->groups[gi].S; };
    std::stable_sort(order.begin(), order.end(), [&](int x, int y) { return work(x) > work(y); });
    std::vector<double> load(b->lanes, 0.0);
    for (int gi : order) {
      const int l = (int)(std::min_element(load.begin(), load.end()) - load.begin());
      lane[gi] = l;
      load[l] += work(gi);
    }
    while ((int)b->lane_streams.size() < b->lanes - 1) {
      hipStream_t st;
      if (hipStreamCreateWithFlags(&st, hipStreamNonBlocking) != hipSuccess) {
        set_error("stream creation failed");
        return -2;
      }
      b->lane_streams.push_back(st);
    }
    while (b->lane_ev.size() < (size_t)kMaxBlockSteps * kMaxLanes) {
      hipEvent_t e;
      HIPCHK(hipEventCreateWithFlags(&e, hipEventDisableTiming));
      b->lane_ev.push_back(e);
    }
  }
  std::map<std::tuple<int, int, int, int, int, int, int>, std::vector<int>> keys;
  for (int gi = 0; gi < ngroups; ++gi) {
    sem_mesh *m = b->groups[gi].mesh;
    const bool t = b->groups[gi].tiled;
    const int shape = t ? m->tiles.T[0] * 10000 + m->tiles.T[1] * 100 + m->tiles.T[2] : 0;
    const int tk = t ? 1 + 2 * (int)(m->tile_diag && !general) : 0;
    keys[std::make_tuple(lane[gi], tk, m->dim, m->order, m->physics, geo_mode(m), shape)].push_back(gi);
  }
  for (auto &l : b->launches) dfree(l.d_items);
  b->launches.clear();
  for (auto &kv : keys) {
    const bool tile = std::get<1>(kv.first) != 0;
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
      L.diag = (std::get<1>(kv.first) & 2) != 0;
      L.dim = std::get<2>(kv.first);
      L.order = std::get<3>(kv.first);
      L.physics = std::get<4>(kv.first);
      L.gmode = std::get<5>(kv.first);
      L.nd = L.dim == 2 ? (L.order + 1) * (L.order + 1) : (L.order + 1) * (L.order + 1) * (L.order + 1);
      L.c = L.physics == SEM_ACOUSTIC ? 1 : L.dim;
      L.colour = c;
      L.nitems = (int)items.size();
      const int shape = std::get<6>(kv.first);
      L.T[0] = shape / 10000;
      L.T[1] = (shape / 100) % 100;
      L.T[2] = shape % 100;
      L.grid = 0;
      L.lane = std::get<0>(kv.first);
      if (tile) {
        int per_cu = 1;
        SEM_TRY(tile_kernel_prepare(L.dim, L.order, L.physics, L.T, &per_cu));
        L.grid = L.nitems;  // one workgroup per tile
      }
      SEM_TRY(dalloc(&L.d_items, items.size()));
      HIPCHK(hipMemcpyAsync(L.d_items, items.data(), items.size() * sizeof(WorkItem), hipMemcpyHostToDevice, b->stream));
      b->launches.push_back(L);
    }
  }
  // colour launches of one step: launches of the same kernel and colour in
  // different lanes run side by side and count once
  std::set<std::tuple<int, int, int, int, int, int, int, int>> per_step;
  for (const auto &L : b->launches)
    per_step.insert(std::make_tuple((int)L.tile, (int)L.diag, L.dim, L.order, L.physics, L.gmode,
                                    L.T[0] * 10000 + L.T[1] * 100 + L.T[2], L.colour));
  b->step_launches = (int)per_step.size();
  HIPCHK(hipStreamSynchronize(b->stream));
  return 0;
}

int finalize_groups(sem_batch *b, int epb_hint) {
  drop_graph(b);
  release_plan(b);
  SEM_TRY(setup_couples(b));
  // group table
  const size_t ng = b->groups.size();
  if (ng > b->d_groups_cap) {
    dfree(b->d_groups);
    b->d_groups_cap = std::max<size_t>(std::max<size_t>(ng, 2 * b->d_groups_cap), 16);
    SEM_TRY(dalloc(&b->d_groups, b->d_groups_cap));
  }
  if (ng)
    HIPCHK(hipMemcpyAsync(b->d_groups, b->h_groups.data(), ng * sizeof(GroupDev), hipMemcpyHostToDevice, b->stream));
  // receiver sets, one per (components, nodes per element)
  std::map<int, std::vector<int>> by_key;
  for (int gi = 0; gi < (int)ng; ++gi)
    if (!b->groups[gi].recs.empty()) by_key[b->groups[gi].rec_key].push_back(gi);
  for (auto &kv : by_key) {
    std::vector<RecDev> rd;
    std::vector<int32_t> rg;
    std::vector<uint32_t> nodes;
    std::vector<float> w;
    for (int gi : kv.second) {
      const Group &g = b->groups[gi];
      const int32_t noff = (int32_t)nodes.size(), woff = (int32_t)w.size();
      for (RecDev d : g.recs) {
        d.nodes_off += noff;
        d.w_off += woff;
        rd.push_back(d);
        rg.push_back(gi);
      }
      nodes.insert(nodes.end(), g.rec_nodes.begin(), g.rec_nodes.end());
      w.insert(w.end(), g.rec_w.begin(), g.rec_w.end());
    }
    RecSet R;
    R.C = kv.first;
    R.nrec = (int)rd.size();
    SEM_TRY(dalloc(&R.d_recs, rd.size()));
    SEM_TRY(dalloc(&R.d_group, rg.size()));
    SEM_TRY(dalloc(&R.d_nodes, nodes.size()));
    SEM_TRY(dalloc(&R.d_w, w.size()));
    HIPCHK(hipMemcpyAsync(R.d_recs, rd.data(), rd.size() * sizeof(RecDev), hipMemcpyHostToDevice, b->stream));
    HIPCHK(hipMemcpyAsync(R.d_group, rg.data(), rg.size() * 4, hipMemcpyHostToDevice, b->stream));
    HIPCHK(hipMemcpyAsync(R.d_nodes, nodes.data(), nodes.size() * 4, hipMemcpyHostToDevice, b->stream));
    HIPCHK(hipMemcpyAsync(R.d_w, w.data(), w.size() * 4, hipMemcpyHostToDevice, b->stream));
    HIPCHK(hipStreamSynchronize(b->stream));
    b->recsets.push_back(R);
  }
  // launch plan
  if (b->path == SEM_PATH_PLAIN) {
    size_t mx = 0;
    for (auto &g : b->groups) mx = std::max(mx, (size_t)g.mesh->E * g.mesh->nd * g.mesh->c);
    if (mx > b->uloc_cap) {
      dfree(b->d_uloc);
      dfree(b->d_aloc);
      b->uloc_cap = 0;
      SEM_TRY(dalloc(&b->d_uloc, mx));
      SEM_TRY(dalloc(&b->d_aloc, mx));
      b->uloc_cap = mx;
    }
    b->epb = 0;
  } else {
    int nd_max = 1;
    for (auto &g : b->groups) nd_max = std::max(nd_max, g.mesh->nd);
    b->epb = epb_hint > 0 ? epb_hint : default_epb(nd_max);
    if (block_threads(nd_max, b->epb) > 512) { set_error("elements per block too large"); return -1; }
    SEM_TRY(build_plan(b));
    int64_t min_steps = INT64_MAX;
    for (auto &g : b->groups) min_steps = std::min(min_steps, std::max<int64_t>(g.nsteps, 1));
    if (b->groups.empty()) min_steps = 1;
    b->block_steps = (int)std::min<int64_t>(std::max<int64_t>(1, 4096 / std::max<int>(1, (int)b->launches.size())), kMaxBlockSteps);
    b->block_steps = std::max(1, std::min<int>(b->block_steps, (int)min_steps));
    // SEM_GRAPH_STEPS=k: steps per captured graph block (1..50)
    const char *gs = getenv("SEM_GRAPH_STEPS");
    if (gs && atoi(gs) > 0) b->block_steps = std::max(1, std::min(atoi(gs), kMaxBlockSteps));
  }
  HIPCHK(hipStreamSynchronize(b->stream));
  return 0;
}

static int enqueue_receivers(sem_batch *b, int koff) {
  for (auto &R : b->recsets) {
    RecArgs ra;
    ra.groups = b->d_groups;
    ra.rec_group = R.d_group;
    ra.recs = R.d_recs;
    ra.nodes = R.d_nodes;
    ra.weights = R.d_w;
    ra.nrec = R.nrec;
    ra.nd = R.C >> 8;
    ra.koff = koff;
    ra.pad = 0;
    SEM_TRY(launch_receivers(ra, R.C & 0xff, b->stream));
  }
  return 0;
}

int enqueue_fused_step(sem_batch *b, int koff, int ku_only, int last) {
  if (!ku_only) SEM_TRY(enqueue_couples(b));  // coupling loads from the velocities after the last step
  // timing: one event pair around all colour launches of the step (a pair
  // per launch would add its own overhead to short launches)
  const bool timed = b->timing && !ku_only && !b->launches.empty();
  if (timed) {
    hipEvent_t e0, e1;
    HIPCHK(hipEventCreate(&e0));
    HIPCHK(hipEventCreate(&e1));
    b->ev.push_back(e0);
    b->ev.push_back(e1);
    b->ev_launches += (int64_t)b->step_launches;
    HIPCHK(hipEventRecord(e0, b->stream));
  }
  // lanes: fork from the batch stream, launch, join back (events of this
  // step's slot; a captured block holds at most kMaxBlockSteps steps)
  const int nl = b->lanes;
  hipEvent_t *lev = nl > 1 ? &b->lane_ev[(size_t)(koff % kMaxBlockSteps) * kMaxLanes] : nullptr;
  if (nl > 1) {
    HIPCHK(hipEventRecord(lev[0], b->stream));
    for (int l = 1; l < nl; ++l) HIPCHK(hipStreamWaitEvent(b->lane_streams[l - 1], lev[0], 0));
  }
  for (auto &L : b->launches) {
    hipStream_t st = L.lane > 0 && L.lane < nl ? b->lane_streams[L.lane - 1] : b->stream;
    FusedArgs a;
    a.groups = b->d_groups;
    a.items = L.d_items;
    a.nitems = L.nitems;
    a.epb = b->epb;
    a.koff = koff;
    a.ku_only = ku_only;
    a.ap_koff = last;
    a.pad = 0;
    auto go = [&]() {
      return L.tile ? launch_tile(L.dim, L.order, L.physics, L.T, L.diag, L.grid, a, st)
                    : launch_fused(L.dim, L.order, L.physics, L.gmode, L.nitems, b->epb, L.nd, L.c, a, st);
    };
    SEM_TRY(go());
  }
  for (int l = 1; l < nl; ++l) {
    HIPCHK(hipEventRecord(lev[l], b->lane_streams[l - 1]));
    HIPCHK(hipStreamWaitEvent(b->stream, lev[l], 0));
  }
  if (timed) HIPCHK(hipEventRecord(b->ev.back(), b->stream));
  if (ku_only) return 0;
  return enqueue_receivers(b, koff);
}

int enqueue_plain_step(sem_batch *b) {
  SEM_TRY(enqueue_couples(b));
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
    a.step = g.d_step;
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
      ++b->ev_launches;
      HIPCHK(hipEventRecord(e0, b->stream));
      SEM_TRY(launch_plain_stiffness(m->dim, m->order, m->physics, pointwise_mode(m), m->E, m->nd, m->c, a, b->stream));
      HIPCHK(hipEventRecord(e1, b->stream));
    } else {
      SEM_TRY(launch_plain_stiffness(m->dim, m->order, m->physics, pointwise_mode(m), m->E, m->nd, m->c, a, b->stream));
    }
    SEM_TRY(launch_fill_f32((int64_t)m->c * m->G, h.acc, 0.f, b->stream));
    for (int c = 0; c < m->ncol; ++c)
      SEM_TRY(launch_plain_scatter(m->col_start[c], m->col_start[c + 1] - m->col_start[c], m->nd, m->c,
                                   m->G, m->d_conn, b->d_aloc, h.acc, b->stream));
    SEM_TRY(launch_plain_update(m->G, m->c, m->quirks, h.dt, h.mi, h.acc, h.ext, h.u, h.w, h.ap,
                                h.kc, b->stream));
  }
  SEM_TRY(enqueue_receivers(b, 0));
  SEM_TRY(launch_advance((int)b->groups.size(), b->d_groups, 1, b->stream));
  return 0;
}

int capture_graph(sem_batch *b, int steps) {
  const bool full = steps == b->block_steps;
  if (full && b->graph) return 0;
  if (!full && b->tail && b->tail_steps == steps) return 0;
  hipGraphExec_t &exec = full ? b->graph : b->tail;
  hipGraph_t &def = full ? b->graph_def : b->tail_def;
  if (exec) (void)hipGraphExecDestroy(exec);
  if (def) (void)hipGraphDestroy(def);
  exec = nullptr;
  def = nullptr;
  HIPCHK(hipStreamBeginCapture(b->stream, hipStreamCaptureModeThreadLocal));
  int rc = 0;
  for (int k = 0; k < steps && !rc; ++k) rc = enqueue_fused_step(b, k, 0, steps - 1);
  if (!rc) rc = launch_advance((int)b->groups.size(), b->d_groups, steps, b->stream);
  hipGraph_t gdef = nullptr;
  hipError_t e = hipStreamEndCapture(b->stream, &gdef);
  if (rc) {
    if (gdef) (void)hipGraphDestroy(gdef);
    return rc;
  }
  HIPCHK(e);
  def = gdef;
  HIPCHK(hipGraphInstantiate(&exec, gdef, nullptr, nullptr, 0));
  if (!full) b->tail_steps = steps;
  return 0;
}

}  // namespace sem

// Full batched-path steps (state advances; step counters do not), each timed
// with HIP events after one warm step; returns the fastest step's
// milliseconds (the least disturbed sample).
static double time_steps(sem_batch *b, int reps) {
  std::vector<hipEvent_t> ev(reps + 1, nullptr);
  auto done = [&](double r) {
    for (auto e : ev)
      if (e) (void)hipEventDestroy(e);
    return r;
  };
  for (auto &e : ev)
    if (hipEventCreate(&e) != hipSuccess) return done(-1);
  if (enqueue_fused_step(b, 0, 0, 0)) return done(-1);
  (void)hipEventRecord(ev[0], b->stream);
  for (int r = 0; r < reps; ++r) {
    if (enqueue_fused_step(b, 0, 0, 0)) return done(-1);
    (void)hipEventRecord(ev[r + 1], b->stream);
  }
  if (hipEventSynchronize(ev[reps]) != hipSuccess) return done(-1);
  double best = 1e300;
  for (int r = 0; r < reps; ++r) {
    float ms = 0;
    if (hipEventElapsedTime(&ms, ev[r], ev[r + 1]) != hipSuccess) return done(-1);
    best = std::min(best, (double)ms);
  }
  return done(best);
}

// key: mesh, nsteps, dt and (coupled shots, which form groups of their own)
// the shot index
using Bucket = std::map<std::tuple<sem_mesh *, int64_t, double, int>, std::vector<int>>;
static Bucket buckets_of(const sem_batch *b) {
  Bucket bk;
  for (int i = 0; i < (int)b->shots.size(); ++i)
    bk[std::make_tuple(b->shots[i].mesh, b->shots[i].nsteps, b->shots[i].dt,
                       b->shots[i].coupled >= 0 ? i : -1)].push_back(i);
  return bk;
}

static int largest_bucket(const sem_batch *b) {
  size_t mx = 0;
  for (auto &kv : buckets_of(b)) mx = std::max(mx, kv.second.size());
  return (int)mx;
}

// Groups of up to max_s shots sharing mesh, nsteps and dt.
static void form_groups(sem_batch *b, int max_s) {
  b->groups.clear();
  for (auto &kv : buckets_of(b)) {
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
}

static Shot *get_shot(sem_batch *b, int shot) {
  if (shot < 0 || shot >= (int)b->shots.size()) {
    set_error("shot index out of range");
    return nullptr;
  }
  return &b->shots[shot];
}

static float *field_base(sem_batch *b, const Shot &s, int which) {
  Group &g = b->groups[s.group];
  const size_t len = (size_t)s.mesh->c * s.mesh->G * g.S;
  return g.fields + (size_t)which * len;
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
  for (auto e : b->lane_ev) (void)hipEventDestroy(e);
  for (auto st : b->lane_streams) (void)hipStreamDestroy(st);
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

int sem_shot_add_file_source(sem_batch *b, int shot, const double *loc, int ncomp,
                             int64_t nsamples, const double *stf) {
  Shot *s = get_shot(b, shot);
  if (!s) return -1;
  if (b->prepared) { set_error("batch already prepared"); return -1; }
  if (ncomp != s->mesh->c) { set_error("file source needs one series per field component"); return -1; }
  if (nsamples < s->nsteps || !stf) { set_error("file source needs a sample for every step"); return -1; }
  SrcSpec sp{};
  sp.kind = 2;
  for (int a = 0; a < s->mesh->dim; ++a) sp.loc[a] = loc[a];
  sp.dir[0] = 1.0;
  int64_t e;
  std::vector<double> load;
  SEM_TRY(source_load(s->mesh, sp, &e, load));  // validates placement now
  sp.nser = nsamples;
  sp.series.assign(stf, stf + (size_t)ncomp * nsamples);
  s->srcs.push_back(std::move(sp));
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

int sem_shot_set_field(sem_batch *b, int shot, int which, const double *vals) {
  Shot *s = get_shot(b, shot);
  if (!s) return -1;
  if (which < 0 || which > 2) { set_error("field selector"); return -1; }
  const size_t len = (size_t)s->mesh->c * s->mesh->G;
  if (!b->prepared) {
    s->init[which].assign(vals, vals + len);
    return 0;
  }
  // live update of this shot's slice; v and a_ live in the state as
  // w = v + dt/2 a_ and a_, so setting one keeps the other
  HIPCHK(hipSetDevice(b->device));
  Group &g = b->groups[s->group];
  const double half_dt = 0.5 * g.dt;
  std::vector<float> buf(len), wcur, acur;
  auto fetch = [&](int w, std::vector<float> &out) {
    out.resize(len);
    HIPCHK(hipMemcpyAsync(out.data(), field_base(b, *s, w) + (size_t)s->slot * len, len * 4,
                          hipMemcpyDeviceToHost, b->stream));
    HIPCHK(hipStreamSynchronize(b->stream));
    return 0;
  };
  if (which >= 1) {
    SEM_TRY(fetch(1, wcur));
    SEM_TRY(fetch(2, acur));
  }
  auto put = [&](int w) {
    HIPCHK(hipMemcpyAsync(field_base(b, *s, w) + (size_t)s->slot * len, buf.data(), len * 4,
                          hipMemcpyHostToDevice, b->stream));
    HIPCHK(hipStreamSynchronize(b->stream));
    return 0;
  };
  if (which == 0) {
    for (size_t i = 0; i < len; ++i) buf[field_pos(g, i)] = (float)vals[i];
    return put(0);
  }
  if (which == 1) {  // w = v_new + dt/2 a_
    for (size_t i = 0; i < len; ++i) {
      const size_t k = field_pos(g, i);
      buf[k] = (float)(vals[i] + half_dt * acur[k]);
    }
    return put(1);
  }
  // a_ new: w = (w - dt/2 a_old) + dt/2 a_new keeps v
  for (size_t i = 0; i < len; ++i) {
    const size_t k = field_pos(g, i);
    buf[k] = (float)((double)wcur[k] - half_dt * acur[k] + half_dt * vals[i]);
  }
  SEM_TRY(put(1));
  for (size_t i = 0; i < len; ++i) buf[field_pos(g, i)] = (float)vals[i];
  return put(2);
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
  std::vector<float> buf(len), ap;
  HIPCHK(hipMemcpyAsync(buf.data(), field_base(b, *s, which) + (size_t)s->slot * len, len * 4,
                        hipMemcpyDeviceToHost, b->stream));
  if (which == 1) {  // v = w - dt/2 a_
    ap.resize(len);
    HIPCHK(hipMemcpyAsync(ap.data(), field_base(b, *s, 2) + (size_t)s->slot * len, len * 4,
                          hipMemcpyDeviceToHost, b->stream));
  }
  HIPCHK(hipStreamSynchronize(b->stream));
  const double half_dt = 0.5 * g.dt;
  for (size_t i = 0; i < len; ++i) {
    const size_t k = field_pos(g, i);
    vals[i] = which == 1 ? (double)buf[k] - half_dt * (double)ap[k] : (double)buf[k];
  }
  return 0;
}

int sem_batch_prepare(sem_batch *b, int path, int epb, int spg) {
  if (b->prepared) { set_error("batch already prepared"); return -1; }
  if (b->shots.empty()) { set_error("batch has no shots"); return -1; }
  if (path != SEM_PATH_PLAIN && path != SEM_PATH_FUSED) { set_error("path"); return -1; }
  HIPCHK(hipSetDevice(b->device));
  b->path = path;
  auto fail = [&](int r) { batch_release(b); return r; };
  int rc = 0;
  // ---- shots per group: given, 1 on the plain path, or autotuned: every
  // power of two up to the largest bucket is set up and timed over full
  // steps (the state is rebuilt from the shot specifications afterwards)
  int max_s = path == SEM_PATH_PLAIN ? 1 : spg;
  if (max_s <= 0) {
    const int biggest = largest_bucket(b);
    max_s = std::min(biggest, 8);
    if (biggest > 1) {
      double best = 1e300;
      for (int cand = 1; cand <= std::min(biggest, 8); cand *= 2) {
        form_groups(b, cand);
        b->h_groups.resize(b->groups.size());
        for (int gi = 0; gi < (int)b->groups.size() && !rc; ++gi) rc = setup_group(b, gi);
        if (!rc) rc = finalize_groups(b, epb);
        const double t = rc ? -1 : time_steps(b, 3);
        batch_release(b);
        if (rc) return rc;
        if (t < 0) return fail(-2);
        if (t < best * 0.99) {
          best = t;
          max_s = cand;
        }
      }
      b->path = path;
    }
  }
  // ---- grouping: shots sharing mesh, nsteps and dt
  form_groups(b, max_s);
  b->h_groups.resize(b->groups.size());
  for (int gi = 0; gi < (int)b->groups.size(); ++gi)
    if ((rc = setup_group(b, gi))) return fail(rc);
  if ((rc = finalize_groups(b, epb))) return fail(rc);
  if (path == SEM_PATH_FUSED && epb <= 0) {
    bool any_elem = false;
    int nd_max = 1;
    for (auto &g : b->groups) {
      any_elem |= !g.tiled;
      nd_max = std::max(nd_max, g.mesh->nd);
    }
    if (any_elem) {
      // autotune elements per block of the element kernel over full steps
      // (gather, operator and finish: the stiffness pass alone ranks them
      // differently), one candidate per distinct block size; the timed steps
      // advance the state, so it is rebuilt from the shot specifications
      double best = 1e300;
      int best_epb = b->epb, last = 0;
      for (int bt = 64; bt <= 512; bt += 64) {
        const int cand = bt / nd_max;
        if (cand < 1 || cand == last) continue;
        last = cand;
        b->epb = cand;
        if ((rc = build_plan(b))) return fail(rc);
        const double t = time_steps(b, 3);
        if (t < 0) return fail(-2);
        if (t < best * 0.99) {
          best = t;
          best_epb = cand;
        }
      }
      batch_release(b);
      b->path = path;
      form_groups(b, max_s);
      b->h_groups.resize(b->groups.size());
      for (int gi = 0; gi < (int)b->groups.size(); ++gi)
        if ((rc = setup_group(b, gi))) return fail(rc);
      if ((rc = finalize_groups(b, best_epb))) return fail(rc);
    }
  }
  b->spg = max_s;
  // the full-block graph is captured here (capture runs nothing), so the
  // first run's whole blocks are already replays
  const char *ng = getenv("SEM_NO_GRAPH");
  if (path == SEM_PATH_FUSED && !(ng && ng[0] == '1') && (rc = capture_graph(b, b->block_steps)))
    return fail(rc);
  if ((rc = hip_check(hipStreamSynchronize(b->stream), "sync"))) return fail(rc);
  b->prepared = true;
  return 0;
}

int sem_batch_run(sem_batch *b, int64_t nsteps) {
  if (!b->prepared) { set_error("batch not prepared"); return -1; }
  HIPCHK(hipSetDevice(b->device));
  for (auto e : b->ev) (void)hipEventDestroy(e);
  b->ev.clear();
  b->ev_launches = 0;
  int64_t left = nsteps;
  if (b->path == SEM_PATH_PLAIN) {
    const char *tg = getenv("SEM_TIME_KERNELS");
    b->timing = tg && tg[0] == '1';
    for (int64_t k = 0; k < left; ++k) SEM_TRY(enqueue_plain_step(b));
  } else {
    const char *tg = getenv("SEM_TIME_KERNELS");
    b->timing = tg && tg[0] == '1';
    const char *ng = getenv("SEM_NO_GRAPH");
    const bool graphs = !(ng && ng[0] == '1') && !b->timing;
    while (left > 0) {
      if (graphs) {  // full blocks, then one (cached) tail block
        const int k = (int)std::min<int64_t>(left, b->block_steps);
        SEM_TRY(capture_graph(b, k));
        HIPCHK(hipGraphLaunch(k == b->block_steps ? b->graph : b->tail, b->stream));
        left -= k;
      } else {
        const int k = (int)std::min<int64_t>(left, b->block_steps);
        for (int j = 0; j < k; ++j) SEM_TRY(enqueue_fused_step(b, j, 0, k - 1));
        SEM_TRY(launch_advance((int)b->groups.size(), b->d_groups, k, b->stream));
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
    b->kernel_launches = b->ev_launches;
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
    HIPCHK(hipMemcpy(&st, b->groups[s->group].d_step, 8, hipMemcpyDeviceToHost));
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
  HIPCHK(hipMemcpy(out, b->groups[s->group].d_seis + s->seis_off, n * 4, hipMemcpyDeviceToHost));
  return 0;
}

int sem_batch_config(const sem_batch *b, int64_t *cfg) {
  cfg[0] = b->epb;
  cfg[1] = b->spg;
  cfg[2] = b->block_steps;
  cfg[3] = (int64_t)b->groups.size();
  int64_t tiled = 0;
  for (const auto &g : b->groups) tiled += g.tiled ? 1 : 0;
  cfg[4] = tiled;
  return 0;
}

int sem_batch_lanes(const sem_batch *b, int *lanes) {
  *lanes = b->prepared ? b->lanes : 0;
  return 0;
}

int sem_batch_kernel_time(const sem_batch *b, double *ms, int64_t *launches) {
  *ms = b->kernel_ms;
  *launches = b->kernel_launches;
  return 0;
}

int sem_mesh_autotune(sem_mesh *m, int reps, int shots, int64_t *shape) {
  *shape = 0;
  HIPCHK(hipSetDevice(m->device));
  if (!m->tiles.valid) return 0;  // element kernel: nothing to choose per mesh
  if (shots < 1 || shots > 64) { set_error("autotune shots must be 1..64"); return -1; }
  reps = std::max(reps, 1);
  int shapes[3 * 16];
  const int ns = tile_kernel_shapes(m->dim, m->order, m->physics, shapes, 16);
  double best = 1e300;
  int bi = -1;
  for (int i = 0; i < ns; ++i) {
    SEM_TRY(setup_tiles(m, &shapes[3 * i]));
    if (!m->tiles.valid) continue;
    sem_batch *b = nullptr;
    SEM_TRY(sem_batch_create(&b, m->device));
    int rc = 0;
    for (int s = 0; s < shots && !rc; ++s) {
      int k = 0;
      rc = sem_batch_add_shot(b, m, 2 * reps + 4, 1.0, &k);
    }
    if (!rc) rc = sem_batch_prepare(b, SEM_PATH_FUSED, default_epb(m->nd), shots);
    const double t = rc ? -1 : time_steps(b, 2 * reps);
    sem_batch_destroy(b);
    if (rc) return rc;
    if (t < 0) { set_error("autotune timing failed"); return -2; }
    if (t < best * 0.98) {
      best = t;
      bi = i;
    }
  }
  SEM_TRY(setup_tiles(m, bi >= 0 ? &shapes[3 * bi] : nullptr));
  if (m->tiles.valid) *shape = m->tiles.T[0] * 10000 + m->tiles.T[1] * 100 + m->tiles.T[2];
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
      rc = enqueue_fused_step(b, 0, 1, 0);
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
      if (!rc) rc = launch_plain_stiffness(m->dim, m->order, m->physics, pointwise_mode(m), m->E, m->nd, m->c, a, b->stream);
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
