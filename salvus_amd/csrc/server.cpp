// Serving loop: shot requests on device-resident meshes with continuous
// batching.  Admitted requests become groups of a long-lived batch; all
// active groups advance together in the same colour launches (one persistent
// tile launch per colour and kernel shape covers the tiles of every active
// group), in blocks of K steps.  Between blocks, finished requests are
// retired (seismograms copied out; a group's device memory is released when
// its last request is done) and requests that have arrived are admitted in
// arrival order while the HBM budget and the active-request cap allow.
// Requests keep their wavefields on the device from admission to completion.
//
// Multi-shot groups: the request at the head of the queue is admitted
// together with later arrived requests on the same mesh and dt whose step
// counts lie within a spread of the group's largest (at most shots_per_group
// shots): they share one tile launch per tile and colour, which reads the
// tile's mesh data once for all of them.  The group steps to its largest
// count; a shorter request's seismograms are complete (and copied out) at
// its own count, its fields then keep stepping with the group unobserved.
//
// A shot's arithmetic does not depend on which other shots share its
// launches or its group (every shot owns its fields and sources, and the
// kernels sum in a fixed order), so a request's result is the same alone and
// in any mix.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <deque>

#include "batch.h"

namespace sem {

struct Request {
  Shot shot;
  double arrival = 0, t_admit = -1, t_done = -1;
  int state = 0;  // 0 pending, 1 active, 2 done, 3 failed, 4 cancelled
  int64_t bytes = 0, steps = 0;
  int group_shots = 0;  // shots of the group it was admitted in
  std::vector<float> seis;
};

}  // namespace sem

using namespace sem;

struct sem_server {
  sem_batch *b = nullptr;
  int path = SEM_PATH_FUSED;
  int64_t budget = 0, used = 0;
  int max_active = 0, block_steps = 32;
  int group_cap = 4;       // shots per group (fused path)
  double spread = 0.1;     // largest step-count gap within a group, of its largest
  std::vector<Request> reqs;
  std::deque<int> pending;  // arrival order
  std::vector<std::vector<int>> active;  // requests of each batch group, in slot order
  int nactive = 0;          // requests in state 1
  std::chrono::steady_clock::time_point t0;
  int64_t done = 0;
};

static double clock_of(const sem_server *s) {
  return std::chrono::duration<double>(std::chrono::steady_clock::now() - s->t0).count();
}

// Device bytes a request holds while active in a group of `nsteps` steps.
static int64_t request_bytes(const sem_server *s, const Shot &sh, int64_t nsteps) {
  const sem_mesh *m = sh.mesh;
  int64_t b = 5 * 4 * (int64_t)m->c * m->G + 8;
  b += 4 * (int64_t)sh.recs.size() * m->c * nsteps;
  b += (int64_t)sh.srcs.size() * (4 * nsteps + 4 * (int64_t)m->nd * m->c + (int64_t)sizeof(SrcDev));
  if (s->path == SEM_PATH_PLAIN) b += 2 * 4 * m->E * (int64_t)m->nd * m->c;  // gather/scatter scratch
  return b;
}

// Removes active group gi (its request's device memory and its shot's host
// data) and renumbers the groups behind it.
static void drop_active(sem_server *s, int gi) {
  sem_batch *b = s->b;
  Group &g = b->groups[gi];
  for (int k : g.shots) b->shots[k] = Shot{};  // the shot's sources, receivers, fields
  free_group(g);
  b->groups.erase(b->groups.begin() + gi);
  b->h_groups.erase(b->h_groups.begin() + gi);
  s->active.erase(s->active.begin() + gi);
  for (int gj = gi; gj < (int)b->groups.size(); ++gj)
    for (int k : b->groups[gj].shots) b->shots[k].group = gj;
  // trailing released shots are not referenced any more
  while (!b->shots.empty() && !b->shots.back().mesh) b->shots.pop_back();
}

// Releases group gi and the bytes of its requests.
static void release_group(sem_server *s, int gi) {
  for (int id : s->active[gi]) s->used -= s->reqs[id].bytes;
  drop_active(s, gi);
}

// Copies out the seismograms of every request that has reached its step
// count; drops groups none of whose requests is still active.
static bool retire(sem_server *s, double now) {
  sem_batch *b = s->b;
  bool dropped = false;
  for (int gi = (int)s->active.size() - 1; gi >= 0; --gi) {
    Group &g = b->groups[gi];
    bool live = false;
    for (size_t k = 0; k < s->active[gi].size(); ++k) {
      Request &r = s->reqs[s->active[gi][k]];
      if (r.state == 1 && r.steps >= r.shot.nsteps) {
        const Shot &sh = b->shots[g.shots[k]];
        const size_t n = sh.recs.size() * (size_t)g.mesh->c * (size_t)r.shot.nsteps;
        r.seis.resize(n);
        const bool ok = n == 0 || hipMemcpy(r.seis.data(), g.d_seis + sh.seis_off, n * 4,
                                            hipMemcpyDeviceToHost) == hipSuccess;
        r.state = ok ? 2 : 3;
        r.t_done = now;
        --s->nactive;
        ++s->done;
      }
      live = live || r.state == 1;
    }
    if (!live) {
      release_group(s, gi);
      dropped = true;
    }
  }
  return dropped;
}

extern "C" {

int sem_server_create(sem_server **out, int device, int path, int64_t hbm_budget, int max_active,
                      int block_steps) {
  *out = nullptr;
  if (path != SEM_PATH_PLAIN && path != SEM_PATH_FUSED) { set_error("path"); return -1; }
  sem_batch *b = nullptr;
  SEM_TRY(sem_batch_create(&b, device));
  auto *s = new sem_server();
  s->b = b;
  s->path = path;
  b->path = path;
  s->max_active = max_active;
  s->block_steps = block_steps > 0 ? block_steps : 32;
  if (hbm_budget > 0) {
    s->budget = hbm_budget;
  } else {
    size_t fr = 0, tot = 0;
    if (hipMemGetInfo(&fr, &tot) != hipSuccess) {
      sem_batch_destroy(b);
      delete s;
      set_error("hipMemGetInfo failed");
      return -2;
    }
    const int64_t reserve = (int64_t)4 << 30;
    s->budget = std::max<int64_t>((int64_t)fr - reserve, (int64_t)fr / 2);
  }
  s->t0 = std::chrono::steady_clock::now();
  *out = s;
  return 0;
}

int sem_server_destroy(sem_server *s) {
  if (!s) return 0;
  sem_batch_destroy(s->b);
  delete s;
  return 0;
}

int sem_server_clock(const sem_server *s, double *t) {
  *t = clock_of(s);
  return 0;
}

int sem_server_submit(sem_server *s, sem_mesh *m, int64_t nsteps, double dt, double arrival, int nsrc,
                      const double *src, int nrec, const double *rec, int *request) {
  if (m->device != s->b->device) { set_error("mesh and server on different devices"); return -1; }
  if (nsteps < 1 || !(dt > 0)) { set_error("nsteps >= 1 and dt > 0 required"); return -1; }
  Request r;
  r.shot.mesh = m;
  r.shot.nsteps = nsteps;
  r.shot.dt = dt;
  r.arrival = arrival;
  for (int k = 0; k < nsrc; ++k) {
    const double *p = src + (size_t)k * SEM_SOURCE_DOUBLES;
    SrcSpec sp{};
    sp.kind = (int)p[0];
    if (sp.kind != 0 && sp.kind != 1) { set_error("source kind"); return -1; }
    for (int a = 0; a < 3; ++a) sp.loc[a] = a < m->dim ? p[1 + a] : 0.0;
    sp.amp = p[4];
    sp.f0 = p[5];
    sp.t0 = p[6];
    for (int a = 0; a < 3; ++a) sp.dir[a] = a < m->c ? p[7 + a] : 0.0;
    for (int i = 0; i < 9; ++i) sp.M[i] = 0.0;
    for (int i = 0; i < m->dim; ++i)
      for (int j = 0; j < m->dim; ++j) sp.M[i * m->dim + j] = p[10 + 3 * i + j];
    int64_t e;
    std::vector<double> load;
    SEM_TRY(source_load(m, sp, &e, load));
    r.shot.srcs.push_back(sp);
  }
  for (int k = 0; k < nrec; ++k) {
    std::array<double, 3> x{0, 0, 0};
    for (int a = 0; a < m->dim; ++a) x[a] = rec[3 * k + a];
    int64_t e;
    double ref[3];
    SEM_TRY(locate(m, x.data(), &e, ref));
    r.shot.recs.push_back(x);
  }
  r.bytes = request_bytes(s, r.shot, nsteps);
  if (r.bytes > s->budget) { set_error("request larger than the server's HBM budget"); return -1; }
  const int id = (int)s->reqs.size();
  s->reqs.push_back(std::move(r));
  // keep the queue in arrival order (stable for equal arrivals)
  auto it = s->pending.end();
  while (it != s->pending.begin() && s->reqs[*(it - 1)].arrival > arrival) --it;
  s->pending.insert(it, id);
  *request = id;
  return 0;
}

int sem_server_pump(sem_server *s, double *status) {
  sem_batch *b = s->b;
  HIPCHK(hipSetDevice(b->device));
  const double now = clock_of(s);
  // ---- admission in arrival order
  bool changed = false;
  while (!s->pending.empty()) {
    const int id = s->pending.front();
    Request &r = s->reqs[id];
    if (r.arrival > now) break;
    if (s->max_active > 0 && s->nactive >= s->max_active) break;
    // the head and later arrived requests on its mesh and dt whose step
    // counts stay within the spread (fused path)
    std::vector<int> mem = {id};
    int64_t nmax = r.shot.nsteps, nmin = r.shot.nsteps;
    if (s->path == SEM_PATH_FUSED)
      for (auto it = s->pending.begin() + 1; it != s->pending.end() && (int)mem.size() < s->group_cap; ++it) {
        const Request &q = s->reqs[*it];
        if (q.arrival > now) break;  // the queue is in arrival order
        if (s->max_active > 0 && s->nactive + (int)mem.size() >= s->max_active) break;
        if (q.shot.mesh != r.shot.mesh || q.shot.dt != r.shot.dt) continue;
        const int64_t hi = std::max(nmax, q.shot.nsteps), lo = std::min(nmin, q.shot.nsteps);
        if ((double)(hi - lo) > s->spread * (double)hi) continue;
        mem.push_back(*it);
        nmax = hi;
        nmin = lo;
      }
    auto bytes_of = [&](const std::vector<int> &ids, int64_t steps) {
      int64_t t = 0;
      for (int k : ids) t += request_bytes(s, s->reqs[k].shot, steps);
      return t;
    };
    if (mem.size() > 1 && s->used + bytes_of(mem, nmax) > s->budget) {
      mem.resize(1);  // the head alone, at its own step count
      nmax = r.shot.nsteps;
    }
    if (s->used + bytes_of(mem, nmax) > s->budget) break;
    Group g;
    g.mesh = r.shot.mesh;
    g.nsteps = nmax;
    g.dt = r.shot.dt;
    g.S = (int)mem.size();
    for (int k : mem) {
      b->shots.push_back(s->reqs[k].shot);
      g.shots.push_back((int)b->shots.size() - 1);
    }
    b->groups.push_back(g);
    const int gi = (int)b->groups.size() - 1;
    const int rc = setup_group(b, gi);
    if (rc) {
      free_group(b->groups[gi]);
      b->groups.pop_back();
      b->shots.resize(b->shots.size() - mem.size());  // admitted again later from the requests
      if ((int)b->h_groups.size() > gi) b->h_groups.resize(gi);
      if (rc == -3) break;  // device memory short: retry after retirements
      return rc;
    }
    for (int k : mem) {
      Request &q = s->reqs[k];
      q.state = 1;
      q.t_admit = now;
      q.bytes = request_bytes(s, q.shot, nmax);
      q.group_shots = (int)mem.size();
      s->used += q.bytes;
      ++s->nactive;
      s->pending.erase(std::find(s->pending.begin(), s->pending.end(), k));
    }
    s->active.push_back(mem);
    changed = true;
  }
  if (changed) {
    SEM_TRY(finalize_groups(b, 0));
    b->prepared = true;
  }
  // ---- one block of steps for every active request
  int64_t k = 0;
  if (!s->active.empty()) {
    b->block_steps = s->block_steps;
    k = s->block_steps;
    SEM_TRY(sem_batch_run(b, k));
    for (auto &grp : s->active)
      for (int id : grp) {
        Request &r = s->reqs[id];
        r.steps = std::min<int64_t>(r.shot.nsteps, r.steps + k);
      }
    if (retire(s, clock_of(s))) {
      SEM_TRY(finalize_groups(b, 0));
      b->prepared = true;
    }
  }
  status[0] = (double)s->nactive;
  status[1] = (double)s->pending.size();
  status[2] = (double)s->done;
  status[3] = (double)k;
  status[4] = s->pending.empty() ? -1.0 : s->reqs[s->pending.front()].arrival;
  status[5] = (double)s->used;
  return 0;
}

int sem_server_cancel(sem_server *s, int request) {
  if (request < 0 || request >= (int)s->reqs.size()) { set_error("request index"); return -1; }
  Request &r = s->reqs[request];
  if (r.state == 0) {
    for (auto it = s->pending.begin(); it != s->pending.end(); ++it)
      if (*it == request) {
        s->pending.erase(it);
        break;
      }
  } else if (r.state == 1) {
    HIPCHK(hipSetDevice(s->b->device));
    HIPCHK(hipStreamSynchronize(s->b->stream));
    r.state = 4;
    --s->nactive;
    // its group is released at once when no other request in it is active;
    // otherwise the shot steps on with its group, unobserved
    for (int gi = 0; gi < (int)s->active.size(); ++gi) {
      const auto &grp = s->active[gi];
      if (std::find(grp.begin(), grp.end(), request) == grp.end()) continue;
      bool live = false;
      for (int id : grp) live = live || s->reqs[id].state == 1;
      if (!live) {
        release_group(s, gi);
        SEM_TRY(finalize_groups(s->b, 0));
        s->b->prepared = true;
      }
      break;
    }
  } else {
    set_error("request already finished or cancelled");
    return -1;
  }
  r.state = 4;
  r.t_done = clock_of(s);
  return 0;
}

int sem_server_set_grouping(sem_server *s, int shots_per_group, double step_spread) {
  if (!s || shots_per_group < 1 || !(step_spread >= 0.0 && step_spread < 1.0)) {
    set_error("shots_per_group >= 1 and 0 <= step_spread < 1 required");
    return -1;
  }
  s->group_cap = shots_per_group;
  s->spread = step_spread;
  return 0;
}

int sem_server_request_group(const sem_server *s, int request, int *shots) {
  if (request < 0 || request >= (int)s->reqs.size()) { set_error("request index"); return -1; }
  *shots = s->reqs[request].group_shots;
  return 0;
}

int sem_server_request(const sem_server *s, int request, double *info) {
  if (request < 0 || request >= (int)s->reqs.size()) { set_error("request index"); return -1; }
  const Request &r = s->reqs[request];
  info[0] = r.state;
  info[1] = r.arrival;
  info[2] = r.t_admit;
  info[3] = r.t_done;
  info[4] = (double)r.steps;
  info[5] = (double)r.bytes;
  return 0;
}

int sem_server_result(const sem_server *s, int request, float *seis) {
  if (request < 0 || request >= (int)s->reqs.size()) { set_error("request index"); return -1; }
  const Request &r = s->reqs[request];
  if (r.state != 2) { set_error("request not finished"); return -1; }
  std::copy(r.seis.begin(), r.seis.end(), seis);
  return 0;
}

}  // extern "C"
