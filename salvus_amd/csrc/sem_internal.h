// Internal declarations of the MI355X SEM library (host + device).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <memory>
#include <string>
#include <vector>

#include "../../include/salvus_sem.h"
#include "tile_layout.h"

namespace sem {

// ---------------------------------------------------------------- limits
constexpr int kMaxOrder3D = 7;   // Hexahedra::mMaxOrder with the order-7 table
constexpr int kMaxOrder2D = 10;  // TensorQuad::mMaxOrder
constexpr int kMaxN = 11;

// Gather-map words carry three flags above the node index (G < 2^29).
constexpr uint32_t kNodeMask = (1u << 29) - 1;
constexpr uint32_t kFlagDirichlet = 1u << 29;  // zero the stiffness row only
constexpr uint32_t kFlagFirst = 1u << 30;      // first colour to touch the node: store
constexpr uint32_t kFlagLast = 1u << 31;       // last colour: finalise (mass + Newmark)
// Tile lattice slot outside the mesh (partial tile at the mesh edge).
constexpr uint32_t kSlotEmpty = 0xffffffffu;

// ---------------------------------------------------------------- errors
void set_error(const std::string &msg);
#define SEM_TRY(expr)                                   \
  do {                                                  \
    int _rc = (expr);                                   \
    if (_rc) return _rc;                                \
  } while (0)
int hip_check(hipError_t e, const char *what);
#define HIPCHK(x) SEM_TRY(::sem::hip_check((x), #x))

// ---------------------------------------------------------------- host maths (FP64)
struct Gll {
  int order = 0, n = 0;
  double xi[kMaxN], w[kMaxN], D[kMaxN * kMaxN];  // D[i*n+j] = l'_j(xi_i)
};
Gll make_gll(int order);
void lagrange(const Gll &g, double x, double *l);
void lagrange_deriv(const Gll &g, double x, double *dl);

// J rows d/dr, d/ds(, d/dt); returns det.
double jacobian(int dim, const double *vtx, const double *ref, double *J);
double invert(int dim, const double *J, double *inv);
void p1_weights(int dim, const double *ref, double *w);
// 1 inside/on hull, 0 outside, -1 degenerate element
int check_hull(int dim, const double *pt, const double *vtx);
int inverse_map(int dim, const double *pt, const double *vtx, double *ref);

// ---------------------------------------------------------------- mesh topology
struct Topology {
  int dim = 0, order = 0;
  int64_t E = 0, V = 0, NF = 0, NEd = 0, G = 0;
  std::vector<int32_t> conn;  // E x nd (original element order)
};
int build_topology(int dim, int order, int64_t E, int64_t V, const int32_t *cells, Topology &T);
int colour_elements(int dim, int64_t E, int64_t V, const int32_t *cells,
                    std::vector<int32_t> &colour);

// ---------------------------------------------------------------- device-side descriptors
// Global-memory view of a device pointer.  Pointers that kernels read out of
// structs in memory are generic, and generic (flat) accesses also count on
// the LDS wait counter, so every wait for an LDS read would also wait for all
// outstanding global loads; kernels access HBM through these views instead.
// (The host pass parses kernel bodies too; there the view is the plain pointer.)
#if defined(__HIP_DEVICE_COMPILE__)
template <class T>
using gptr = __attribute__((address_space(1))) T *;
#else
template <class T>
using gptr = T *;
#endif
template <class T>
__device__ __forceinline__ gptr<T> glb(T *p) {
  return (gptr<T>)p;
}

// Kahan compensations of u and v are stored as the two bf16 halves of one
// word (round to nearest even).  A compensation is the rounding error of the
// last update, at most half an ulp of the sum, so its leading 8 bits keep the
// compensated u and v within about one FP32 ulp over thousands of steps while
// the state moves 8 bytes less per component and node.
__device__ __forceinline__ uint32_t bf16_bits(float x) {
  const uint32_t b = __float_as_uint(x);
  return (b + 0x7fffu + ((b >> 16) & 1u)) >> 16;
}
__device__ __forceinline__ uint32_t pack_comp(float cu, float cv) {
  return bf16_bits(cu) | (bf16_bits(cv) << 16);
}
__device__ __forceinline__ float comp_u(uint32_t w) { return __uint_as_float(w << 16); }
__device__ __forceinline__ float comp_v(uint32_t w) { return __uint_as_float(w & 0xffff0000u); }

// Per-element factor counts.
// Affine geometry (one Jacobian per element): acoustic Gsym (dim*(dim+1)/2);
// elastic Jinv (dim^2) + detJ (1).
// Pointwise geometry (per GLL node): acoustic w*vp^2*Gsym; elastic 2D
// Jinv(4)+w detJ(1)+C11,C13,C33,C55(4); elastic 3D Jinv(9)+w detJ(1)+9 c's.
__host__ __device__ inline int affine_geo_count(int dim, int physics) {
  return physics == SEM_ACOUSTIC ? dim * (dim + 1) / 2 : dim * dim + 1;
}
// Geometry modes of the element kernels: pointwise factors, per-element
// affine factors, or pointwise factors for the reference's quad operator
// (SEM_QUIRK_QUAD_JACOBIAN on non-affine quads: J^-1 and w det J of the
// QuadP1.h:20-27 Jacobian per node, J^-1 applied at the output node).
enum { kGeoPoint = 0, kGeoAffine = 1, kGeoQuadRef = 2 };
// With kGeoQuadRef the acoustic factors are Jinv(4) + w det J vp^2 (1).
__host__ __device__ inline int point_geo_count(int dim, int physics, int gmode = kGeoPoint) {
  if (physics == SEM_ACOUSTIC) return gmode == kGeoQuadRef ? 5 : dim * (dim + 1) / 2;
  return dim == 2 ? 9 : 19;
}
// Vertex material the fused kernels interpolate on the fly (stiffness only):
// acoustic {VP}; Elastic2D {C11,C13,C33,C55}; Elastic3D {RHO,VPV,VPH,VSV,VSH,ETA}.
__host__ __device__ inline int kernel_npar(int dim, int physics) {
  if (physics == SEM_ACOUSTIC) return 1;
  return dim == 2 ? 4 : 6;
}
__host__ __device__ inline int ref_npar(int dim, int physics) {
  if (physics == SEM_ACOUSTIC) return 1;
  return dim == 2 ? 5 : 6;
}

struct SrcDev {
  int32_t elem;     // sorted element position
  int32_t stf_off;  // offset of this source's STF samples
  int32_t load_off; // offset of nd*c load weights
  int32_t shot;     // shot slot within the group
};

struct RecDev {
  int32_t nodes_off; // offset of the owning element's nd node ids (field numbering)
  int32_t w_off;     // offset of nd Lagrange weights
  int32_t shot;      // shot slot within the group
  int32_t stride;    // floats between consecutive steps (R_shot * c)
  float *out;        // this receiver's step-0, component-0 sample
};

// One tile of a logically Cartesian mesh (see tiles.cpp).
struct TileInfo {
  int32_t ex, ey, ez;  // elements per axis in this tile
  int32_t colour;      // lattice parity colour
  int32_t elem_off;    // first element in tile order
  int32_t pad;
  int64_t node_off;    // first node slot
};

// One group = shots sharing a mesh, fields shot-major:
// field[(s*c + comp)*G + g] (each shot's field contiguous).
struct GroupDev {
  const uint32_t *conn;   // E x nd (sorted order, flags)
  const float *geo_e;     // affine: E x ng
  const float *par_e;     // E x nv x kpar
  const float *geo_p;     // pointwise: E x ngp x nd
  const float *mi;        // G
  // Newmark state: u, w = v + dt/2 a_ (what the update reads), a_ (written
  // for read-out only: v = w - dt/2 a_)
  float *u, *w, *ap;
  uint32_t *kc;           // Kahan compensations of u and w (bf16 pair per word)
  float *acc;
  const SrcDev *src;
  const float *stf;
  const float *load;
  int32_t nsrc;
  int32_t S;
  int64_t G;
  int64_t nsteps;
  int64_t *step;          // device step counter of this group
  float dt, acl;           // acl = dt/2 (Order2Newmark.cpp:89), host conversions of w
  int32_t quirks;
  int32_t pad;
  // tile kernel (structured meshes): fields in tile-major node numbering
  const TileInfo *tiles;
  const uint32_t *tile_words;   // per tile, in node order (padding last)
  const uint16_t *tile_slot;    // lattice slot of each tile word
  const uint32_t *tile_words_lat;  // per tile, in lattice-slot order
  const float *telem;     // per element in tile order: affine factors, then vertex material
  // coupled shots (fluid-solid pairs): external load added to a node's sum
  // at its first colour, null otherwise; a_ stored every step (the coupling
  // reads the velocities v = w - dt/2 a_ after each step)
  const float *ext;
  int32_t ap_always;
  int32_t pad2;
};

// Work item of a colour launch: EPB elements of one group (element kernel),
// or one tile of one group (tile kernel: `first` = tile id).
struct WorkItem {
  int32_t group;
  int32_t first;  // first sorted element position / tile id
  int32_t count;  // elements in this item
  int32_t pad;
};

// Every tile has the full Tx x Ty (x Tz) element box and its whole node
// lattice: elements and slots beyond a partial tile at the mesh edge are
// padding (element order -1, slot word kSlotEmpty).
struct TileSet {
  bool valid = false;
  int T[3] = {0, 0, 0};
  int64_t ntiles = 0, total_slots = 0;
  int64_t slots_per_tile = 0;        // (Tx N + 1)(Ty N + 1)(Tz N + 1)
  int elems_per_tile = 0;            // Tx Ty Tz
  std::vector<TileInfo> info;
  std::vector<uint32_t> words;       // per tile, in node order: internal node | flags, or kSlotEmpty
  std::vector<uint16_t> slot_of;     // lattice slot of each word
  std::vector<uint32_t> words_lat;   // the same words in lattice-slot order
  std::vector<int32_t> to_internal;  // DMPlex DOF -> internal node
  std::vector<int32_t> elem_order;   // tile-order position -> element (-1: padding)
  std::vector<int32_t> elem_pos;     // element -> tile-order position
  std::vector<int64_t> colour_start; // tiles by colour
  std::vector<int32_t> by_colour;
};

// Kernel argument blocks (passed by value; recorded by graph capture).
struct FusedArgs {
  const GroupDev *groups;
  const WorkItem *items;
  int32_t nitems;
  int32_t epb;
  int32_t koff;      // step offset within the captured block
  int32_t ku_only;   // 1: accumulate -K u into acc on every colour (no update)
  int32_t ap_koff;   // a_ (read out only) is stored at this step offset of the
                     // block -- its last -- and at a shot's last step
  int32_t pad;
};

struct PlainArgs {
  const uint32_t *conn;
  const float *geo_p;
  const float *uloc;
  float *aloc;
  const SrcDev *src;
  const float *stf;
  const float *load;
  const int64_t *step;
  int32_t nsrc;
  int32_t epb;
  int64_t E;
  int32_t koff;
  int32_t pad;
};

struct RecArgs {
  const GroupDev *groups;
  const int32_t *rec_group;  // group of each receiver
  const RecDev *recs;
  const uint32_t *nodes;
  const float *weights;
  int32_t nrec;
  int32_t nd;
  int32_t koff;
  int32_t pad;
};

// ---------------------------------------------------------------- launchers (kernels.hip)
int init_device_tables(int device);
int launch_affine_geometry(int dim, int physics, int64_t E, const double *vtx, float *geo,
                           hipStream_t st);
int launch_point_geometry(int dim, int order, int physics, int gmode, int64_t E, const double *vtx,
                          const double *vpar, int npar, float *geo_p, double *me, hipStream_t st);
int launch_mass_scatter(int64_t first, int64_t count, int nd, const uint32_t *conn,
                        const double *me, double *M, hipStream_t st);
int launch_reciprocal(int64_t G, const double *M, float *mi, hipStream_t st);
int launch_fill_f64(int64_t n, double *p, double v, hipStream_t st);
int launch_fill_f32(int64_t n, float *p, float v, hipStream_t st);
int launch_fill_i32(int64_t n, int32_t *p, int32_t v, hipStream_t st);
int launch_colour_touch(int64_t first, int64_t count, int nd, const uint32_t *conn, int32_t *hits,
                        hipStream_t st);
int launch_colour_count(int64_t G, int32_t *hits, unsigned long long *viol, hipStream_t st);
// Threads per block for EPB elements of nd nodes (multiple of 64).
inline int block_threads(int nd, int epb) { return ((nd * epb + 63) / 64) * 64; }
inline size_t fused_lds_bytes(int dim, int nd, int c, int epb) {
  return sizeof(float) * (size_t)epb * c * nd * (1 + dim);
}
int launch_fused(int dim, int order, int physics, int gmode, int nitems, int epb, int nd, int c,
                 const FusedArgs &a, hipStream_t st);
int launch_permute_f32(int64_t n, const int32_t *to, const float *src, float *dst, hipStream_t st);
int launch_plain_gather(int64_t E, int nd, int C, int64_t G, const uint32_t *conn, const float *u,
                        float *uloc, hipStream_t st);
int launch_plain_stiffness(int dim, int order, int physics, int gmode, int64_t E, int nd, int c,
                           const PlainArgs &a, hipStream_t st);
int launch_plain_scatter(int64_t first, int64_t count, int nd, int C, int64_t G,
                         const uint32_t *conn, const float *aloc, float *acc, hipStream_t st);
// Newmark update in the (u, w = v + dt/2 a_) form: w += dt a, u += dt w.
int launch_plain_update(int64_t G, int C, int quirks, float dt, const float *mi, const float *acc,
                        const float *ext, float *u, float *w, float *ap, uint32_t *kc, hipStream_t st);
int launch_receivers(const RecArgs &a, int C, hipStream_t st);
// Advances every group's step counter by k (clamped at its nsteps).
int launch_advance(int ngroups, const GroupDev *groups, int64_t k, hipStream_t st);

// ---------------------------------------------------------------- tile kernel (tile_kernels.hip)
int init_tile_tables(int device);
struct TileKernelInfo {
  size_t lds_bytes;  // 0: no kernel compiled for this order and tile shape
  int threads;
};
TileKernelInfo tile_kernel_info(int dim, int order, int physics, const int *T);
// Compiled tile shapes for (dim, order, physics): up to max triples; returns the count.
int tile_kernel_shapes(int dim, int order, int physics, int *shapes, int max);
// Readies the kernel of this shape (LDS limit) and returns its workgroups per CU.
int tile_kernel_prepare(int dim, int order, int physics, const int *T, int *per_cu);
// One workgroup per tile: grid = a.nitems.
int launch_tile(int dim, int order, int physics, const int *T, bool diag, int grid, const FusedArgs &a,
                hipStream_t st);

}  // namespace sem

// ---------------------------------------------------------------- public structs
struct sem_mesh {
  int device = 0, dim = 0, order = 0, n = 0, nd = 0, physics = 0, c = 0, nv = 0, quirks = 0;
  int64_t E = 0, V = 0, G = 0;
  int ncol = 0;
  bool affine = false;
  std::vector<int64_t> col_start;  // ncol+1 positions in sorted order
  std::vector<int32_t> perm;       // sorted position -> original element
  std::vector<int32_t> inv_perm;   // original element -> sorted position
  std::vector<int32_t> conn;       // original order E x nd (plain)
  std::vector<int32_t> colour;     // original order
  std::vector<double> vtx;         // original order E x nv x dim
  std::vector<double> vpar;        // original order E x nv x npar
  std::vector<double> bbox;        // original order E x 2 dim
  std::vector<uint8_t> dirichlet;  // G (empty = none)
  // device
  hipStream_t stream = nullptr;
  uint32_t *d_conn = nullptr;      // sorted E x nd with flags
  float *d_geo_e = nullptr;        // affine per-element factors
  float *d_par_e = nullptr;        // per-element vertex material (kernel set)
  float *d_geo_p = nullptr;        // pointwise factors (lazy)
  float *d_mi = nullptr;
  double *d_vtx = nullptr;         // sorted E x nv x dim (FP64)
  double *d_vpar = nullptr;        // sorted E x nv x npar (FP64)
  // tile path (logically Cartesian affine meshes)
  sem::TileSet tiles;
  sem::TileInfo *d_tile_info = nullptr;
  uint32_t *d_tile_words = nullptr;
  uint16_t *d_tile_slot = nullptr;
  uint32_t *d_tile_words_lat = nullptr;
  float *d_telem = nullptr, *d_mi_int = nullptr;
  bool tile_diag = false;  // every element's affine factors diagonal (axis-aligned lattice)
  size_t bytes = 0;       // device bytes held (tile path included)
  size_t tile_bytes = 0;  // of which the tile path
};

namespace sem {
struct Shot;
struct Group;
}  // namespace sem

struct sem_batch;

namespace sem {
// Geometry mode of a mesh's element kernels (kGeo*).
inline int geo_mode(const sem_mesh *m) {
  if (m->affine) return kGeoAffine;
  return (m->dim == 2 && (m->quirks & SEM_QUIRK_QUAD_JACOBIAN)) ? kGeoQuadRef : kGeoPoint;
}
// Mode of the pointwise factors (plain path, mass): the reference quad
// operator where it differs from the consistent one, else kGeoPoint.
inline int pointwise_mode(const sem_mesh *m) {
  return geo_mode(m) == kGeoQuadRef ? kGeoQuadRef : kGeoPoint;
}
// Tiles a logically Cartesian mesh; 0 = tiled, 1 = not eligible.
int build_tiles(sem_mesh *m, int Tx, int Ty, int Tz, TileSet &ts);
}  // namespace sem
