// Internal host structures of batches (shots advanced together) and of the
// serving loop that keeps a changing set of shots on the device.
#pragma once

#include <array>
#include <string>
#include <vector>

#include "sem_internal.h"

namespace sem {

constexpr int kMaxBlockSteps = 50;  // steps per captured graph block
constexpr int kMaxLanes = 4;        // streams one step's launches may use

template <typename T>
inline int dalloc(T **p, size_t count, size_t *bytes = nullptr) {
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
inline void dfree(T *&p) {
  if (p) (void)hipFree((void *)p);
  p = nullptr;
}

struct SrcSpec {
  int kind;  // 0 force (Ricker), 1 moment tensor (Ricker), 2 file (tabulated)
  double loc[3], amp, f0, t0, dir[3], M[9];
  // kind 2: component k's force at step n is series[k * nser + n]
  // (SourceHdf5.cpp:31-36: the rows of a column-major nsteps x c matrix)
  int64_t nser = 0;
  std::vector<double> series;
};

struct Shot {
  sem_mesh *mesh = nullptr;
  int64_t nsteps = 0;
  double dt = 0;
  std::vector<SrcSpec> srcs;
  std::vector<std::array<double, 3>> recs;
  std::vector<double> init[3];
  int group = -1, slot = 0;
  int64_t seis_off = 0;  // within the group's seismogram buffer
  int coupled = -1;      // index of its coupling pair in the batch, or -1
};

// Shots sharing a mesh, nsteps and dt, advanced in lockstep; fields are
// shot-major: field[(slot*c + comp)*G + node].
struct Group {
  sem_mesh *mesh = nullptr;
  std::vector<int> shots;
  int S = 1;
  int64_t nsteps = 0;
  double dt = 0;
  float *fields = nullptr;  // 5 arrays of c*G*S: u v ap kc acc
  SrcDev *d_src = nullptr;
  float *d_stf = nullptr, *d_load = nullptr;
  int nsrc = 0;
  bool tiled = false;       // tile kernel; fields in the mesh's tile-major numbering
  int64_t *d_step = nullptr;  // device step counter
  float *d_seis = nullptr;    // seismograms of the group's shots
  int64_t seis_len = 0;
  // receivers (offsets local to the group; RecDev::out is absolute)
  std::vector<RecDev> recs;
  std::vector<uint32_t> rec_nodes;
  std::vector<float> rec_w;
  int rec_key = 0;  // components | nodes per element << 8
  float *ext = nullptr;    // coupled shots: external load added at the first colour
  float *mi_own = nullptr; // coupled shots: the combined inverse mass (group order)
  size_t bytes = 0;  // device bytes held
};

// One coupled fluid-solid pair of shots of a batch: the coupling terms in
// the field numbering of the two groups (each coupled shot is a group of its
// own), evaluated before every step's colour launches.
struct CoupleDev {
  int32_t nf, ns;             // fluid and solid target nodes
  const int32_t *f_tgt;       // fluid field index of each fluid target
  const int32_t *f_ptr;       // nf + 1 term offsets
  const int32_t *f_src;       // solid x field index of each term (y: + G_s)
  const float *f_wx, *f_wy;   // half length x GLL weight x outward normal
  const int32_t *s_tgt;       // solid x field index of each solid target (y: + G_s)
  const int32_t *s_ptr;
  const int32_t *s_src;       // fluid field index of each term
  const float *s_coef;        // -rho_0 x half length x GLL weight
  int64_t Gs;                 // solid nodes (y component offset)
  const float *wf, *apf;      // fluid state w, a_
  const float *ws, *aps;      // solid state w, a_ (x then y)
  float *extf, *exts;         // external loads of the two shots
  float aclf, acls;           // dt/2 of the two shots
};

struct CouplePair {
  sem_coupling *cp = nullptr;
  int fluid_shot = -1, solid_shot = -1;
  void *dev = nullptr;        // device arrays of the terms
  CoupleDev h{};
};

struct LaunchSet {
  int dim, order, physics, nd, c, colour;
  int gmode;  // geometry mode of the element kernel (kGeo*)
  bool tile;
  bool diag;  // tile kernel: diagonal-geometry variant
  int nitems;
  int T[3];  // tile shape (tile kernel)
  int grid;  // tile kernel: workgroups (one per tile)
  int lane;  // stream lane (0: the batch stream, 1: the second stream)
  WorkItem *d_items;
};

struct RecSet {
  int C = 0, nrec = 0;  // C: components | nodes per element << 8
  RecDev *d_recs = nullptr;
  int32_t *d_group = nullptr;
  uint32_t *d_nodes = nullptr;
  float *d_w = nullptr;
};

}  // namespace sem

struct sem_batch {
  int device = 0;
  hipStream_t stream = nullptr;
  std::vector<sem::Shot> shots;
  std::vector<sem::Group> groups;
  bool prepared = false;
  int path = SEM_PATH_FUSED, epb = 0, spg = 0, block_steps = 0;
  sem::GroupDev *d_groups = nullptr;
  size_t d_groups_cap = 0;
  std::vector<sem::GroupDev> h_groups;
  std::vector<sem::LaunchSet> launches;
  std::vector<sem::RecSet> recsets;
  float *d_uloc = nullptr, *d_aloc = nullptr;  // plain path scratch
  size_t uloc_cap = 0;
  hipGraphExec_t graph = nullptr;   // block_steps steps
  hipGraph_t graph_def = nullptr;
  hipGraphExec_t tail = nullptr;    // tail_steps (< block_steps) steps
  hipGraph_t tail_def = nullptr;
  int tail_steps = 0;
  int64_t steps_done = 0;
  bool timing = false;
  double kernel_ms = 0;
  int64_t kernel_launches = 0;
  std::vector<hipEvent_t> ev;  // timing: event pairs around timed kernels
  int64_t ev_launches = 0;       // kernel launches the event pairs cover
  std::vector<sem::CouplePair> couples;  // coupled fluid-solid shot pairs
  // Lanes: the groups split into parts of about equal work, each part's
  // colour launches on its own stream (lane 0: `stream`), forked from and
  // joined back into `stream` once per step.
  int lanes = 1;
  int step_launches = 0;                  // colour launches per step, lanes side by side counted once
  std::vector<hipStream_t> lane_streams;  // lanes 1.. (created on demand)
  std::vector<hipEvent_t> lane_ev;        // per step of a block: fork, then a join per lane
};

namespace sem {

// api.cpp
int set_device(int device);  // selects the device and uploads its tables once
int ensure_pointwise(sem_mesh *m);
int locate(const sem_mesh *m, const double *pt, int64_t *elem, double *ref);
int source_load(const sem_mesh *m, const SrcSpec &s, int64_t *elem, std::vector<double> &load);
double ricker(double t, double amp, double f0, double t0);  // Ricker.cpp:27-33
// Tile path of a mesh with the given shape (null: preference list); frees
// any previous tiling first.
int setup_tiles(sem_mesh *m, const int *shape);
int upload_tiles(sem_mesh *m);
void free_tiles(sem_mesh *m);
// batch.cpp
// Position of DMPlex DOF entry i (= comp*G + g) in a group's field arrays.
int64_t field_pos(const Group &g, int64_t i);
// Device state of b->groups[gi] from its shots (fields, sources, receivers);
// fills b->h_groups[gi].
int setup_group(sem_batch *b, int gi);
void free_group(Group &g);
// Uploads the group table, receivers and launch plan of the current groups.
int finalize_groups(sem_batch *b, int epb_hint);
void drop_graph(sem_batch *b);
// One fused step at offset koff of a block whose last offset is `last`.
int enqueue_fused_step(sem_batch *b, int koff, int ku_only, int last);
// Default elements per block of the element kernel for nd nodes per element.
int default_epb(int nd);
int enqueue_plain_step(sem_batch *b);
// Captures (once) the graph of k steps: k == block_steps, or one cached
// shorter tail block.
int capture_graph(sem_batch *b, int k);
// Coupling pairs (coupling.cpp): device terms of every pair after the groups
// are set up, the coupling kernel of every pair before a step's launches,
// and their release.
int setup_couples(sem_batch *b);
// The pair's combined inverse mass (DMPlex order) of one of its meshes.
const std::vector<float> &coupled_mass(const CouplePair &p, const sem_mesh *m);
int enqueue_couples(sem_batch *b);
void free_couples(sem_batch *b);

}  // namespace sem
