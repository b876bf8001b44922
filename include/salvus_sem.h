/*
 * salvus_sem.h -- C ABI of the MI355X spectral-element time-stepping path
 * (tensor quads/hexes; Scalar, Elastic2D, Elastic3D physics).
 *
 * This is the drop-in boundary between the reference's C++ Problem/Newmark
 * loop and the device.  Plain pointers and sizes only; every entry point
 * returns 0 on success and a negative status otherwise, with the message in
 * sem_last_error() (the reference throws std::runtime_error instead,
 * Main.cpp:70-74).  The caller owns host buffers; the library owns device
 * buffers.  A mesh or batch is bound to one device and one HIP stream and is
 * driven from one host thread at a time.
 *
 * Conventions (same as the reference):
 *   order N, n = N+1, nd = n^dim nodes per element, node p = i + n j (+ n^2 k)
 *   with i along r (TensorQuad.cpp:187, Hexahedra.cpp:591);
 *   cells in the vertex order of QuadP1.h:98-108 / HexP1.cpp:123-133;
 *   vertex material per element: par[(e*nv + v)*npar + q] with parameters
 *   Scalar {VP}, Elastic2D {RHO,C11,C13,C33,C55},
 *   Elastic3D {RHO,VPV,VPH,VSV,VSH,ETA};
 *   global fields are component-planar: f[comp*G + g] in the DMPlex DOF
 *   numbering (cells, vertices, faces, edges; Mesh.cpp:361-452).
 */
#ifndef SALVUS_SEM_H
#define SALVUS_SEM_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct sem_mesh sem_mesh;
typedef struct sem_batch sem_batch;
typedef struct sem_server sem_server;
typedef struct sem_coupling sem_coupling;

enum { SEM_ACOUSTIC = 0, SEM_ELASTIC = 1 };

/* Reference-behaviour switches:
 * SEM_QUIRK_E3D_SOURCE keeps Elastic3D's source scaling (its source term
 *   skips test-and-integrate, Elastic3D.cpp:126-133);
 * SEM_QUIRK_NEWMARK3D follows the field list {ax, ay, ax, a} of
 *   Order2Newmark.cpp:93-96: on elastic meshes (2-D and 3-D) x advances twice
 *   per step, y once and z never (the name keeps its first use);
 * SEM_QUIRK_QUAD_JACOBIAN gives non-parallelogram quads the reference's
 *   operator: the Jacobian of QuadP1.h:20-27 (row d/dr from dN/dr evaluated
 *   at r, row d/ds from dN/ds at s), J^-1 applied at the output node
 *   (TensorQuad.cpp:383-390) and that det J in the mass.  Parallelograms and
 *   hexes are unaffected (both Jacobians are one there).  Off: the consistent
 *   bilinear Jacobian with J^-T at the quadrature points. */
enum { SEM_QUIRK_NEWMARK3D = 1, SEM_QUIRK_E3D_SOURCE = 2, SEM_QUIRK_QUAD_JACOBIAN = 4 };
#define SEM_DEFAULT_QUIRKS (SEM_QUIRK_E3D_SOURCE | SEM_QUIRK_QUAD_JACOBIAN)

/* Execution paths of a batch. */
enum {
  SEM_PATH_PLAIN = 0, /* one shot per launch: gather, stiffness, colour scatter, update */
  SEM_PATH_FUSED = 1  /* fused colour kernels over all shots, hipGraph step blocks */
};

/* Field selectors: displacement, velocity, previous acceleration (a_). */
enum { SEM_FIELD_U = 0, SEM_FIELD_V = 1, SEM_FIELD_A_PREV = 2 };

/* Thread-local text of the last error (replaces the reference's exception
 * messages, Main.cpp:70-74, Logging.cpp:9-12). */
const char *sem_last_error(void);
/* Number of visible HIP devices. */
int sem_device_count(int *count);

/* Host-only gather map and colouring (no device needed): the closure
 * numbering of Mesh::setupGlobalDof (Mesh.cpp:361-452) and the element
 * colouring used by the scatter.  conn: E x nd (or NULL), colour: E (or
 * NULL); info[0..3] = {G unique nodes, colours, faces, edges}. */
int sem_topology_build(int dim, int order, int64_t num_cells, int64_t num_verts,
                       const int32_t *cells, int32_t *conn, int32_t *colour, int64_t *info);

/* ------------------------------------------------------------------ mesh
 * Replaces Mesh::read(dim, numCells, numVerts, numVertsPerElem, cells,
 * coords) (Mesh.cpp:31-51), Mesh::setupGlobalDof (Mesh.cpp:361-452),
 * Element::attachVertexCoordinates / attachMaterialProperties
 * (TensorQuad.cpp:212-251, Hexahedra.cpp:132-162, :414-425,
 * Elastic3D.cpp:33-54) and Order2Newmark::initializeGlobalDofs' mass
 * assembly (Order2Newmark.cpp:32-66).  Builds the gather map, element
 * colouring and FP64 geometry/mass on the device. */
int sem_mesh_create(sem_mesh **mesh, int device, int dim, int order, int physics,
                    int64_t num_cells, int64_t num_verts, const int32_t *cells,
                    const double *coords, const double *vertex_par, int quirks);
int sem_mesh_destroy(sem_mesh *mesh);
/* info[0..15] = {E, G (unique nodes), ncolours, components, nodes/element,
 *                affine (1/0), device bytes held, dim, tiles (0 = element
 *                kernel), tile shape Tx*10000 + Ty*100 + Tz, tile lattice
 *                slots, slots on the mesh, accumulator reads and writes per
 *                shot-step, elements per tile (padded), diagonal
 *                geometry (1: every element's J^-1 is diagonal and the tile
 *                kernel skips the off-diagonal terms)} */
int sem_mesh_info(const sem_mesh *mesh, int64_t *info);
/* Tile-shape autotuner (elements per workgroup of the tile kernel): for
 * every compiled tile shape of the mesh's order and physics, times 2 x `reps`
 * full batched steps of one group of `shots` shots (the batch size the mesh
 * will serve; the best shape depends on it) one by one and keeps the shape
 * with the fastest step; *shape = Tx*10000 + Ty*100 + Tz (0: the mesh takes
 * the element kernel).  No reference counterpart.  Call before the mesh
 * joins a batch. */
int sem_mesh_autotune(sem_mesh *mesh, int reps, int shots, int64_t *shape);
/* Gather map conn[e*nd + p] = global DOF, element order as given
 * (DMPlexVecGetClosure + ClsMap, Problem.cpp:302-320). */
int sem_mesh_get_conn(const sem_mesh *mesh, int32_t *conn);
/* Element colours (no two elements of one colour share a node). */
int sem_mesh_get_colours(const sem_mesh *mesh, int32_t *colour);
/* Assembled inverse diagonal mass, FP32 (Order2Newmark.cpp:52). */
int sem_mesh_get_inverse_mass(const sem_mesh *mesh, float *mi);
/* Global node coordinates (G x dim, FP64). */
int sem_mesh_get_node_coords(const sem_mesh *mesh, double *xyz);
/* Homogeneous Dirichlet: zero stiffness rows at masked nodes
 * (HomogeneousDirichlet.cpp:71-78). */
int sem_mesh_set_dirichlet(sem_mesh *mesh, const uint8_t *node_mask);
/* One assembled stiffness application ku = K u (c x G, FP32) on the device
 * through the given path's kernels (Problem.cpp:176-208 without source and
 * mass); u and ku are host buffers. */
int sem_mesh_apply_stiffness(sem_mesh *mesh, int path, const float *u, float *ku);
/* Allocation-free variant of sem_mesh_apply_stiffness (no counterpart in
 * the reference; for callers that cannot allocate device memory: code under
 * HIP graph capture, tight serving loops).  u and ku are DEVICE buffers of
 * c x G floats in the DMPlex DOF order; ku = K u.  It allocates nothing: its
 * group descriptor, step counter and work items live in a 4 MiB device array
 * of the library, which each GPU holds until the process ends; a mesh whose
 * largest element colour needs more work items than that array holds
 * (262,112 items of up to 256 threads' worth of elements: at order 4 in 3-D
 * two elements each, 524,224 elements per colour) is rejected with an
 * error.  Runs the element-colour kernel (also on meshes with tiles)
 * on `stream` (a hipStream_t; NULL = the mesh's stream, and the call then
 * waits for the result).  Calls that overlap on one GPU share that array and
 * race: serialise them. */
int sem_mesh_apply_stiffness_noalloc(sem_mesh *mesh, const float *u, float *ku, void *stream);
/* Race check on the device: returns in *violations the number of
 * (colour, node) pairs touched by more than one element of that colour. */
int sem_mesh_check_colouring(sem_mesh *mesh, int64_t *violations);

/* ----------------------------------------------------------------- shots
 * A batch holds shots (sources + receivers + wavefields) on one or more
 * meshes and advances them together.  Replaces Problem::initializeElements'
 * source/receiver attachment (Problem.cpp:34-142), the time loop
 * (Main.cpp:45-66) with assembleIntoGlobalDof / applyInverseMassMatrix /
 * takeTimeStep (Problem.cpp:143-215, Order2Newmark.cpp:68-115). */
int sem_batch_create(sem_batch **batch, int device);
int sem_batch_destroy(sem_batch *batch);
/* Adds a shot on `mesh` that will advance nsteps steps of dt; *shot gets
 * its index in the batch. */
int sem_batch_add_shot(sem_batch *batch, sem_mesh *mesh, int64_t nsteps, double dt, int *shot);
/* Ricker point source (Ricker.cpp:27-33) at loc, owned by the first element
 * whose closed hull holds it (Problem.cpp:57-113); dir has c entries. */
int sem_shot_add_source(sem_batch *batch, int shot, const double *loc, double amp, double f0,
                        double t0, const double *dir);
/* Tabulated point source (SourceHdf5.cpp:12-36 with the physics'
 * computeSourceTerm, Scalar.cpp:79-88, Elastic2D.cpp:115-127,
 * Elastic3D.cpp:126-133): at step n component k of the force is
 * stf[k * nsamples + n], i.e. row n of the column-major nsamples x c matrix
 * SourceHdf5::fire reads (time_idx = n); ncomp must equal the mesh's c and
 * nsamples must cover the shot's steps. */
int sem_shot_add_file_source(sem_batch *batch, int shot, const double *loc, int ncomp,
                             int64_t nsamples, const double *stf);
/* Moment-tensor Ricker source (no reference counterpart):
 * f_p = -M grad(phi_p)(loc) R(t); M is dim x dim row-major. */
int sem_shot_add_moment_source(sem_batch *batch, int shot, const double *loc, const double *M,
                               double amp, double f0, double t0);
/* Receiver sampling u at loc after every step (the reference locates
 * receivers but never records, Receiver.cpp:48-59). */
int sem_shot_add_receiver(sem_batch *batch, int shot, const double *loc);
/* Initial/current fields (c x G, FP64 host values stored FP32). */
int sem_shot_set_field(sem_batch *batch, int shot, int which, const double *vals);
int sem_shot_get_field(sem_batch *batch, int shot, int which, double *vals);
/* Allocates device state and records the step plan.  path: SEM_PATH_*;
 * elems_per_block / shots_per_group: 0 = autotuned (each candidate timed
 * over full steps; the timed steps' state is discarded and rebuilt from the
 * shot specifications, which include initial fields set before this call).
 * No reference counterpart (the reference runs one problem per process). */
int sem_batch_prepare(sem_batch *batch, int path, int elems_per_block, int shots_per_group);
/* Advances every unfinished shot by up to nsteps steps. */
int sem_batch_run(sem_batch *batch, int64_t nsteps);
/* Steps taken so far by a shot and its simulated time. */
int sem_shot_progress(sem_batch *batch, int shot, int64_t *steps, double *time);
/* Receiver samples nsteps x R x c (FP32), sampled after each step. */
int sem_shot_get_seismograms(sem_batch *batch, int shot, float *out);

/* ------------------------------------------------------------------ files
 * HDF5 (libhdf5 >= 1.10 opened at run time: SEM_HDF5_LIB, else libhdf5.so,
 * libhdf5.so.103 or /opt/conda/lib; host only, no GPU). */
/* ReceiverHdf5::write (ReceiverHdf5.cpp:31-87): per receiver r a group
 * /names[r] holding one FP32 dataset /names[r]/fields[k] of nsamples values
 * for each of ncomp fields; data is nsamples x nrec x ncomp (the layout of
 * sem_shot_get_seismograms).  append = 0 creates (truncates) the file,
 * 1 adds to it (a receiver name already present is an error).  Names hold
 * no '/'. */
int sem_receivers_write_h5(const char *path, int append, int nrec, const char *const *names,
                           int ncomp, const char *const *fields, int64_t nsamples,
                           const float *data);
/* SourceHdf5::loadData (SourceHdf5.cpp:40-50): reads dataset "<name>/data"
 * as nsteps x ncomp doubles in the order stored (the reference's column-major
 * matrix: component k's series is out[k * nsteps + n]); the dataset must hold
 * exactly nsteps * ncomp values. */
int sem_source_read_h5(const char *path, const char *name, int64_t nsteps, int ncomp, double *out);
/* The mesh part of the movie file, written before any frame as the
 * reference DMViews the mesh into its viewer (Problem.cpp:133-137): creates
 * (truncates) `path` with /geometry/vertices [num_verts][dim] FP64 and
 * /topology/cells [num_cells][2^dim] int32 (attribute cell_dim), the names
 * of PETSc's DMPlex HDF5 view, and /geometry/nodes [num_nodes][dim] FP64,
 * the coordinates of the global DOFs whose values the frames hold
 * (sem_mesh_get_node_coords).  Frames are then appended with
 * sem_movie_write_h5(path, append = 1, ...). */
int sem_movie_write_mesh_h5(const char *path, int dim, int64_t num_verts, const double *coords,
                            int64_t num_cells, const int32_t *cells, int64_t num_nodes,
                            const double *node_xyz);
/* Problem::saveSolution (Problem.cpp:217-237; the driver calls it after
 * every step, Main.cpp:55-62, with the time after the step): appends one
 * movie frame to `path` (append = 0 creates / truncates it): field k's n
 * values (values[k * n + i], global DOF order) to the extendible FP64
 * dataset /fields/<names[k]> [frames][n], and `time` to /time [frames].
 * Every dataset of a frame must hold the same number of frames before it is
 * appended (else -1 and nothing is written).  Names hold no '/'. */
int sem_movie_write_h5(const char *path, int append, int nfields, const char *const *names,
                       int64_t n, const double *values, double time);
/* ------------------------------------------------------------ coupling
 * 2-D fluid-solid coupling (AcousticToElastic2D.cpp:17-60,
 * ElasticToAcoustic.cpp:17-41; coupled elements and edge neighbours,
 * Mesh.cpp:480-521; edge quadrature, TensorQuad.cpp:400-515).  The fluid
 * (acoustic) and the solid (Elastic2D) elements are two meshes whose cells
 * come from one quad mesh (shared vertices have identical coordinates).
 * The coupling holds:
 *  - the combined inverse mass of nodes on both meshes (the reference
 *    assembles one mass over all elements, Order2Newmark.cpp:32-66);
 *  - the fluid load: over every edge of a fluid element with a vertex on a
 *    solid element that has a neighbouring element (either physics), the
 *    edge integral of the solid velocity . the element's outward normal
 *    (ElasticToAcoustic.cpp:27-41; 0 where the edge node is not a solid
 *    node);
 *  - the solid load: over every edge of a solid element with a vertex on a
 *    fluid element whose neighbour is fluid, -rho_0 times the edge integral
 *    of the fluid velocity, added to both the x and the y row
 *    (AcousticToElastic2D.cpp:49-60), rho_0 the mean of the neighbour's
 *    vertex RHO (fluid_vertex_rho: fluid cells x 4).
 * Velocities are those after the previous step.  The meshes themselves are
 * not changed. */
int sem_coupling_create(sem_coupling **coupling, sem_mesh *fluid, sem_mesh *solid,
                        const double *fluid_vertex_rho);
int sem_coupling_destroy(sem_coupling *coupling);
/* info[0..3] = {nodes on both meshes, fluid target nodes, solid target nodes,
 *               coupling terms} */
int sem_coupling_info(const sem_coupling *coupling, int64_t *info);
/* Couples two shots of a batch before sem_batch_prepare: fluid_shot on the
 * coupling's fluid mesh, solid_shot on its solid mesh, with the same dt and
 * step count; both advance as one coupled problem (each becomes a group of
 * its own, a_ is stored every step). */
int sem_batch_couple(sem_batch *batch, sem_coupling *coupling, int fluid_shot, int solid_shot);

/* Launch configuration chosen: cfg[0..4] = {elems/block, shots/group,
 * graph block steps, groups, groups on the tile kernel (the others take the
 * element kernel)}. */
int sem_batch_config(const sem_batch *batch, int64_t *cfg);
/* Streams (lanes) the batched path's launches use: the groups split into
 * parts of about equal work, each part's colour launches on its own stream,
 * forked from and joined into the batch stream once per step; results are
 * the same bit for bit with any count.  1 by default; SEM_LANES=k (1..4)
 * chooses (at most one per group; 1 with coupled pairs); 0 before
 * sem_batch_prepare.  No counterpart in the reference (one problem
 * per process). */
int sem_batch_lanes(const sem_batch *batch, int *lanes);
/* Timing of the dominant kernel during the last sem_batch_run when the
 * environment sets SEM_TIME_KERNELS=1 (steps then run without graphs): total
 * ms and number of launches, from HIP events on the batch stream (batched
 * path: one pair around each step's colour launches of the tile or element
 * kernel, the launches of one colour in different lanes, which run side by
 * side, counted once; plain path: one pair around each stiffness launch). */
int sem_batch_kernel_time(const sem_batch *batch, double *ms, int64_t *launches);

/* Reference time-loop semantics (Options.cpp:84-97, Main.cpp:45): number
 * of iterations of while (t < duration) and the adjusted dt. */
int64_t sem_reference_step_count(double duration, double dt, double *dt_out);

/* ---------------------------------------------------------------- serving
 * Shot requests on device-resident meshes with continuous batching.  No
 * counterpart in the reference (which runs one problem per process,
 * Main.cpp:12-75); this is the request loop the batched path serves.  A
 * request is one shot: its wavefields live on the device from admission to
 * completion; all active requests advance together, K steps per pump, in the
 * same colour launches; between blocks finished requests are retired and
 * arrived requests admitted in arrival order while the HBM budget (bytes) and
 * the active cap (0 = none, counted in requests) allow.  On the fused path
 * the request at the head of the queue is admitted with later arrived
 * requests on the same mesh and dt whose step counts lie within a spread of
 * the largest (sem_server_set_grouping; default up to 4 shots, 10 %): they
 * form one multi-shot group that steps to its largest count, and each
 * request's seismograms are copied out at its own count.  Times are seconds
 * on the server's clock (sem_server_clock; 0 at creation).  path
 * SEM_PATH_PLAIN with max_active 1 is the one-shot-per-launch baseline (one
 * shot per group). */
#define SEM_SOURCE_DOUBLES 19 /* kind (0 force, 1 moment), x, y, z, amp, f0, t0,
                                 dir[3], M[3][3] row-major */
int sem_server_create(sem_server **server, int device, int path, int64_t hbm_budget,
                      int max_active, int block_steps);
int sem_server_destroy(sem_server *server);
int sem_server_clock(const sem_server *server, double *t);
/* sources: nsrc x SEM_SOURCE_DOUBLES; receivers: nrec x 3 coordinates */
int sem_server_submit(sem_server *server, sem_mesh *mesh, int64_t nsteps, double dt,
                      double arrival, int nsrc, const double *sources, int nrec,
                      const double *receivers, int *request);
/* status[0..5] = {active, pending, finished, steps advanced, next arrival
 *                 (-1: none), device bytes held by active requests} */
int sem_server_pump(sem_server *server, double *status);
/* Cancels a request (no reference counterpart): a pending request leaves
 * the queue; an active one is retired at once and its partial results
 * dropped, its group's device memory released unless other active requests
 * share the group (then the shot steps on with it, unobserved, until the
 * group ends).  Finished or cancelled requests are an error.  The other
 * requests' results do not change (each request's arithmetic is its own). */
int sem_server_cancel(sem_server *server, int request);
/* Multi-shot groups of the fused path: at most shots_per_group requests per
 * group (1: every request alone), step counts within step_spread (0 <= s <
 * 1) of the group's largest.  Applies to later admissions. */
int sem_server_set_grouping(sem_server *server, int shots_per_group, double step_spread);
/* *shots = shots of the group the request was admitted in (0: not admitted) */
int sem_server_request_group(const sem_server *server, int request, int *shots);
/* info[0..5] = {state (0 pending, 1 active, 2 done, 3 failed, 4 cancelled),
 *               arrival, admission, completion, steps done, device bytes} */
int sem_server_request(const sem_server *server, int request, double *info);
/* seismograms of a finished request: nsteps x receivers x components */
int sem_server_result(const sem_server *server, int request, float *seis);

#ifdef __cplusplus
}
#endif

#endif /* SALVUS_SEM_H */
