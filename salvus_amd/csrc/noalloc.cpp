// Allocation-free stiffness application: sem_mesh_apply_stiffness_noalloc.
//
// sem_mesh_apply_stiffness builds a one-shot batch (group descriptor, step
// counter, work items, fields) in freshly allocated device memory.  This
// variant takes device input and output buffers from the caller and keeps
// everything else -- the group descriptor, the step counter and one colour's
// work items -- in a file-scope device array, so callers that may not call
// hipMalloc (graph capture, tight serving loops) can use it.  It runs the
// element-colour kernel (k_fused, accumulate-only mode) on the DMPlex-ordered
// fields directly, one launch per element colour, the same kernel, elements
// per block and launch order as the allocating entry point takes on meshes
// without tiles.  Every kernel here writes the scratch itself (no memset, no
// host copy), so the calls can be captured.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <mutex>

#include "batch.h"
#include "sem_internal.h"

namespace sem {
namespace {

// Scratch: [0, 256) group descriptor, [256, 512) step counter, then work
// items of one colour.
constexpr size_t kScratchBytes = size_t(4) << 20;
constexpr size_t kItemsOffset = 512;
constexpr int64_t kMaxItems = (int64_t)((kScratchBytes - kItemsOffset) / sizeof(WorkItem));
static_assert(sizeof(GroupDev) <= 256, "group descriptor exceeds its scratch slot");

static __device__ __align__(256) unsigned char scratch[kScratchBytes];

// Per-device address of `scratch` (hipGetSymbolAddress with that device
// current); the array's host-side name is not a device address.
std::mutex g_mu;
void *g_ptr[64] = {nullptr};

int scratch_ptr(int device, unsigned char **out) {
  std::lock_guard<std::mutex> lk(g_mu);
  if (device < 0 || device >= 64) { set_error("device index"); return -1; }
  if (!g_ptr[device]) {
    void *p = nullptr;
    hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p) {
      set_error(std::string("scratch address: ") + hipGetErrorString(e));
      return -2;
    }
    g_ptr[device] = p;
  }
  *out = static_cast<unsigned char *>(g_ptr[device]);
  return 0;
}

__global__ void k_put_group(GroupDev g, GroupDev *dst, int64_t *step) {
  if (threadIdx.x == 0 && blockIdx.x == 0) {
    *dst = g;
    *step = 0;
  }
}

__global__ void k_colour_items(int64_t first, int64_t count, int epb, WorkItem *items) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int64_t f = first + i * epb;
  if (i * epb >= count) return;
  WorkItem w;
  w.group = 0;
  w.first = (int32_t)f;
  w.count = (int32_t)min((int64_t)epb, first + count - f);
  w.pad = 0;
  items[i] = w;
}

__global__ void k_negate(int64_t n, float *x) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) x[i] = -x[i];
}

int launched(const char *what) {
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) {
    set_error(std::string(what) + ": " + hipGetErrorString(e));
    return -2;
  }
  return 0;
}

}  // namespace
}  // namespace sem

using namespace sem;

extern "C" int sem_mesh_apply_stiffness_noalloc(sem_mesh *m, const float *u, float *ku, void *stream) {
  if (!m || !u || !ku) { set_error("null argument"); return -1; }
  const int epb = default_epb(m->nd);
  int64_t worst = 0;
  for (int c = 0; c < m->ncol; ++c)
    worst = std::max<int64_t>(worst, (m->col_start[c + 1] - m->col_start[c] + epb - 1) / epb);
  if (worst > kMaxItems) {
    set_error("mesh too large for the allocation-free stiffness: a colour needs " +
              std::to_string(worst) + " work items, the scratch holds " + std::to_string(kMaxItems) +
              " (at most " + std::to_string(kMaxItems * epb) + " elements per colour)");
    return -1;
  }
  HIPCHK(hipSetDevice(m->device));
  unsigned char *base = nullptr;
  SEM_TRY(scratch_ptr(m->device, &base));
  GroupDev *d_group = reinterpret_cast<GroupDev *>(base);
  int64_t *d_step = reinterpret_cast<int64_t *>(base + 256);
  WorkItem *d_items = reinterpret_cast<WorkItem *>(base + kItemsOffset);
  hipStream_t st = stream ? static_cast<hipStream_t>(stream) : m->stream;

  GroupDev h{};
  h.conn = m->d_conn;
  h.geo_e = m->d_geo_e;
  h.par_e = m->d_par_e;
  h.geo_p = m->d_geo_p;
  h.mi = m->d_mi;
  h.u = const_cast<float *>(u);  // read only: the accumulate-only mode never updates u
  h.acc = ku;
  h.S = 1;
  h.G = m->G;
  h.nsteps = 1;
  h.step = d_step;
  h.quirks = m->quirks;
  hipLaunchKernelGGL(k_put_group, dim3(1), dim3(64), 0, st, h, d_group, d_step);
  SEM_TRY(launched("stiffness descriptor"));
  FusedArgs a;
  a.groups = d_group;
  a.items = d_items;
  a.epb = epb;
  a.koff = 0;
  a.ku_only = 1;
  a.ap_koff = 0;
  a.pad = 0;
  for (int c = 0; c < m->ncol; ++c) {
    const int64_t first = m->col_start[c], count = m->col_start[c + 1] - first;
    if (count <= 0) continue;
    const int64_t nit = (count + epb - 1) / epb;
    hipLaunchKernelGGL(k_colour_items, dim3((unsigned)((nit + 255) / 256)), dim3(256), 0, st, first, count,
                       epb, d_items);
    SEM_TRY(launched("stiffness work items"));
    a.nitems = (int32_t)nit;
    SEM_TRY(launch_fused(m->dim, m->order, m->physics, geo_mode(m), a.nitems, epb, m->nd, m->c, a, st));
  }
  const int64_t len = (int64_t)m->c * m->G;
  hipLaunchKernelGGL(k_negate, dim3((unsigned)((len + 255) / 256)), dim3(256), 0, st, len, ku);
  SEM_TRY(launched("stiffness sign"));
  if (!stream) HIPCHK(hipStreamSynchronize(st));
  return 0;
}
