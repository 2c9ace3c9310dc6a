/*
 * cmcx.h -- C ABI of the MI355X-native ADI time step (drop-in for the
 * AdiSolver3D path of samech-git/cmc-fluid-solver).
 *
 * Plain C: opaque handle, pointers and sizes, int status codes, no exceptions
 * and no torch types across the boundary.  One handle is driven by one host
 * thread at a time; all device work of a handle is issued on its stream.
 *
 * Reference interfaces each entry point replaces (paths are inside the
 * reference tree):
 *   cmcx_create           AdiSolver3D::Init                 src/FluidSolver3D/AdiSolver3D.cpp:166-268
 *                         + Grid3D node data (SetType/SetData) src/FluidSolver3D/Grid3D.h:120-121
 *   cmcx_set_options      AdiSolver3D::SetOptionsGPU        src/FluidSolver3D/AdiSolver3D.cpp:160-164
 *   cmcx_build_segments   AdiSolver3D::CreateSegments       src/FluidSolver3D/AdiSolver3D.cpp:553-562
 *   cmcx_get_segments     h_listX/Y/Z (Segment3D lists)     src/FluidSolver3D/AdiSolver3D.h:78
 *   cmcx_get_segment_nodes  d_node_listX/Y/Z (NodesBoundary3D) src/FluidSolver3D/Grid3D.cpp:129-146
 *   cmcx_update_boundaries AdiSolver3D::UpdateBoundaries    src/FluidSolver3D/AdiSolver3D.cpp:286-304
 *   cmcx_step             AdiSolver3D::TimeStep             src/FluidSolver3D/AdiSolver3D.cpp:306-391
 *   cmcx_div_error        TimeLayer3D::EvalDivError         src/FluidSolver3D/TimeLayer3D.h:595-641
 *   cmcx_get_layer        Solver3D::GetLayer                src/FluidSolver3D/Solver3D.cpp:21-25
 *   cmcx_get_fields       (current state u,v,w,p,T; the reference keeps it in the
 *                          protected `cur` layer, src/FluidSolver3D/Solver3D.h:46)
 *   cmcx_set_fields       Solver2D::SetLayer analogue       src/FluidSolver2D/Solver2D.cpp:37-46
 *   cmcx_set_boundary     Grid3D::SetData                   src/FluidSolver3D/Grid3D.h:121
 *   cmcx_params_*         FluidParams constructors          src/Common/Geometry.h:545-561
 *   cmcx_destroy          AdiSolver3D::~AdiSolver3D         src/FluidSolver3D/AdiSolver3D.cpp:153-158
 *   cmcx_geometry_load    Grid3D::LoadFromFile (Load3DShape, Grid2D::LoadFromFile)
 *                         src/FluidSolver3D/Grid3D.cpp:373-436, 488-514; src/FluidSolver2D/Grid2D.cpp:272-356
 *   cmcx_geometry_build   Grid3D::Prepare (Build, Prepare2D) src/FluidSolver3D/Grid3D.cpp:574-666, 864-981
 *   cmcx_coards_write_header  OutputNetCDF3D_header        src/Common/IO.h:136-297
 *   cmcx_coards_write_layer   OutputNetCDF3D_layer         src/Common/IO.h:355-388
 *
 * 2D ADI (the reference's CPU-only FluidSolver2D path):
 *   cmcx2d_create/update_boundaries/step/get_fields/destroy
 *                         AdiSolver2D::Init/TimeStep, Solver2D::UpdateBoundaries
 *                         src/FluidSolver2D/AdiSolver2D.cpp:21-50, 279-323; Solver2D.cpp:48-63
 */
#ifndef CMCX_H
#define CMCX_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* status codes */
#define CMCX_OK 0
#define CMCX_DIVERGED 1          /* divergence guard fired (mean |div| > 0.01); state not advanced */
#define CMCX_ERR_ARG (-1)        /* bad argument (null pointer, size, enum) */
#define CMCX_ERR_MASK (-2)       /* mask not well formed: NODE_IN on a box face, or NODE_OUT next to NODE_IN */
#define CMCX_ERR_HIP (-3)        /* HIP runtime error (message in cmcx_last_error) */
#define CMCX_ERR_STATE (-4)      /* call out of order (e.g. step before build_segments) */
#define CMCX_ERR_UNSUPPORTED (-5)
#define CMCX_ERR_TOO_LARGE (-6)
#define CMCX_ERR_PICARD (-7)     /* 2D: Picard loop exceeded 100 iterations or error > 1.0 */

/* precision of the fields (the reference's FTYPE) */
#define CMCX_F32 0
#define CMCX_F64 1

/* node types and boundary kinds, same numbering as the reference's enums
   (src/Common/Geometry.h:31-41) */
#define CMCX_NODE_IN 0
#define CMCX_NODE_OUT 1
#define CMCX_NODE_BOUND 2
#define CMCX_NODE_VALVE 3
#define CMCX_BC_NOSLIP 0
#define CMCX_BC_FREE 1

/* axes (DirType X, Y, Z) */
#define CMCX_X 0
#define CMCX_Y 1
#define CMCX_Z 2

/* sweep implementations (cmcx_set_options `path`) */
#define CMCX_PATH_PLAIN 0   /* one thread per (segment, field), one launch per axis */
#define CMCX_PATH_FUSED 1   /* one wavefront per grid line, LDS-staged tiles, one launch per axis */

/* Picard merge order (cmcx_set_options `merge_order`) */
#define CMCX_MERGE_CPU 0        /* the reference's CPU backend (AdiSolver3D.cpp:591-604, 347-357) */
#define CMCX_MERGE_GPU_ORDER 1  /* the reference's CUDA backend (AdiSolver3D.cu:337-364): per field, in order */

typedef struct cmcx_solver cmcx_solver;

/* Per-node input, x-major with z fastest: node (i,j,k) is element (i*dimy + j)*dimz + k.
   Structure-of-arrays form of the reference's Node array (Grid3D.h:73-88). */
typedef struct cmcx_domain {
    int32_t dims[3];          /* dimx, dimy, dimz (each >= 3) */
    double h[3];              /* dx, dy, dz */
    double base_t;            /* Grid3D::baseT */
    const uint8_t* type;      /* NodeType per node */
    const uint8_t* bc_vel;    /* BCtype for u, v, w per node */
    const uint8_t* bc_temp;   /* BCtype for T per node */
    const double* v[3];       /* node velocity (Node.v), used on BOUND/VALVE nodes */
    const double* T;          /* node temperature (Node.T) */
} cmcx_domain;

/* FluidParams (Geometry.h:538-562); rounded to the field type inside. */
typedef struct cmcx_params {
    double v_T, v_vis, t_vis, t_phi;
} cmcx_params;

const char* cmcx_version(void);
/* message of the last failing call on this host thread */
const char* cmcx_last_error(void);

void cmcx_params_normalized(double Re, double Pr, double lambda, cmcx_params* out);
void cmcx_params_dimensional(double vis, double rho, double R, double k, double cv,
                             cmcx_params* out);

/* Allocates the layers on the current HIP device and uploads the domain.
   State starts as the node values (TimeLayer3D(grid) -> CopyFromGrid).
   stream: a hipStream_t, or NULL for the default stream. */
int cmcx_create(const cmcx_domain* dom, const cmcx_params* params, int dtype, void* stream,
                cmcx_solver** out);
/* Several independent grids (each with its own mask, boundary data and
   parameters) advanced together: every fused sweep covers all grids in one
   launch.  The per-grid calls below (segments, fields, layer output, boundary
   data, div_error) address the grid chosen with cmcx_select (0 at creation);
   cmcx_step / cmcx_run advance all grids, and the divergence guard fires when
   any grid exceeds the threshold (no grid is advanced then).  No reference
   counterpart (the reference runs one grid per process). */
int cmcx_create_batch(const cmcx_domain* doms, const cmcx_params* params, int count, int dtype,
                      void* stream, cmcx_solver** out);
int cmcx_batch_size(cmcx_solver* s, int* count);
int cmcx_select(cmcx_solver* s, int index);
/* Divergence errors of every grid from the last evaluation (count doubles). */
int cmcx_grid_errors(cmcx_solver* s, double* errs);
int cmcx_destroy(cmcx_solver* s);

int cmcx_set_options(cmcx_solver* s, int path, int merge_order);
/* Z-block reordering of the X and Y segment lists (the reference's nblockZ,
   Grid3D.cpp:59-62, 124-125).  Affects the list order only; 1 = none. */
int cmcx_set_nblockz(cmcx_solver* s, int nblockz);

/* Builds the X, Y, Z segment lists on the device (count, scan, emit). */
int cmcx_build_segments(cmcx_solver* s);
int cmcx_num_segments(cmcx_solver* s, int axis, int64_t* n);
/* 7 int32 per segment: posx, posy, posz, endx, endy, endz, size; in the
   reference's list order.  cap = capacity in segments. */
int cmcx_get_segments(cmcx_solver* s, int axis, int32_t* out, int64_t cap);
/* Endpoint node records (NodesBoundary3D: first, last).  Per segment:
   flags[6] = {first.type, first.bc_vel, first.bc_temp, last.type, last.bc_vel, last.bc_temp},
   vals[8]  = {first.v.x, first.v.y, first.v.z, first.T, last.v.x, last.v.y, last.v.z, last.T}. */
int cmcx_get_segment_nodes(cmcx_solver* s, int axis, uint8_t* flags, double* vals, int64_t cap);

/* Writes node velocity/temperature into the state on NODE_BOUND and NODE_VALVE nodes. */
int cmcx_update_boundaries(cmcx_solver* s);

/* One ADI time step: num_global Picard iterations of Z, Y, X sweeps with
   num_local sweeps each.  compute_err != 0 evaluates the divergence of the new
   state into *err (may be NULL).  Returns CMCX_DIVERGED (state not advanced)
   when the error exceeds 0.01, as the reference throws before its swap. */
int cmcx_step(cmcx_solver* s, double dt, int num_global, int num_local, int compute_err,
              double* err);
/* Several steps with update_boundaries before each (the reference's time loop,
   FluidSolver3D.cpp:226-242): error evaluated every 10th step and on the last. */
int cmcx_run(cmcx_solver* s, double dt, int steps, int num_global, int num_local, double* err);

/* Replays each time step as a HIP graph (captured on first use per dt,
   num_global, num_local and state buffer; cmcx_run's graph includes the
   boundary fill).  Options, variants and segments invalidate the graphs.
   0 = direct launches.  No reference counterpart. */
int cmcx_set_graph(cmcx_solver* s, int enable);

/* Divergence error of the current state (selected grid). */
int cmcx_div_error(cmcx_solver* s, double* err);

/* Current state to host arrays of the field type (float for CMCX_F32, double for
   CMCX_F64), N elements each; any pointer may be NULL.  p = v_T * T.
   Selected grid of a batch. */
int cmcx_get_fields(cmcx_solver* s, void* u, void* v, void* w, void* p, void* T);
/* Same into device buffers (N elements each), ordered on the handle's stream. */
int cmcx_get_fields_device(cmcx_solver* s, void* u, void* v, void* w, void* p, void* T);
/* Replace the current state (initial conditions, checkpoint restore). */
int cmcx_set_fields(cmcx_solver* s, const void* u, const void* v, const void* w, const void* T);
/* Device pointers of the current state fields {u, v, w, T} (valid until the next step).
   The solver takes the state as rewritten at this call (the next boundary fill and
   step set-up then visit every BOUND/VALVE and non-IN node): write through the
   pointers before the next cmcx_update_boundaries, cmcx_step or cmcx_run. */
int cmcx_state_device_ptrs(cmcx_solver* s, void** ptrs4);

/* Solver3D::GetLayer: the reference output view.  Sets NODE_OUT nodes of the
   layer it reads to 99999 and resamples it to (ox, oy, oz) by nearest index
   (0 = full size).  Like the reference, it reads the buffer that holds the
   state before the last step.  vel: 3 values per output node (field type),
   T: double per output node. */
int cmcx_get_layer(cmcx_solver* s, void* vel, double* T, int ox, int oy, int oz);

/* Grid3D::SetData for every node at once (new boundary data; the mask is fixed). */
int cmcx_set_boundary(cmcx_solver* s, const uint8_t* bc_vel, const uint8_t* bc_temp,
                      const double* vx, const double* vy, const double* vz, const double* T);

/* Kernel timing with HIP events on the handle's stream (0 = off). */
int cmcx_profile(cmcx_solver* s, int enable);
/* Accumulated time of the sweep kernels since the last reset: total ms and launches. */
int cmcx_profile_read(cmcx_solver* s, double* sweep_ms, int64_t* sweep_launches, int reset);
/* Autotuner (no reference counterpart): times the fused path's kernel variants
   of the X, Y and Z sweeps on this domain (writing scratch layers only; the state is
   untouched), keeps the fastest per axis and caches the choice process-wide for
   the grid shape and field type.  choice (may be NULL) receives 3 ints. */
int cmcx_autotune(cmcx_solver* s, int reps, int* choice);
/* Forces the fused path's kernel variant of one axis.  X (0) and Y (1) sweeps:
   0 = 128-byte z rows per work-group (one work-group per CU), 1 = 64-byte rows
   (two work-groups per CU), 2 = 32-byte rows (four per CU); 3, 4, 5 = the same
   with half the rows per lane (lines of at most 128 nodes); 6, 7, 8, 9 =
   variants 0, 1, 2, 3 keeping the Picard layer in LDS between the elimination
   and the merge (variants that do not fit the grid return
   CMCX_ERR_UNSUPPORTED).  Z sweep (2), 0-7: bit 0 = registers capped for more
   waves per SIMD, bit 1 = the Picard layer parked in LDS until the merge,
   bit 2 = each z line spread over 64 lanes instead of 32.  No reference
   counterpart (tuning knob). */
int cmcx_set_tile_variant(cmcx_solver* s, int axis, int variant);
/* FluidParams of the selected grid (e.g. a new viscosity between rollouts). */
int cmcx_set_params(cmcx_solver* s, const cmcx_params* params);

/* ---- reverse mode: gradients of a rollout (no reference counterpart) -------
   cmcx_adjoint_forward runs `steps` steps (UpdateBoundaries + TimeStep with
   num_local = 1, no divergence evaluation) from the current state, keeping a
   device checkpoint of the state every `checkpoint_every` steps; the final
   state becomes the current state.  cmcx_adjoint_backward then takes
   seed[4] = dL/d(final u, v, w, T) (host arrays of the field type, N each,
   NULL = 0; only NODE_IN nodes are read) and returns, for the recorded
   rollout, gbar[4] = dL/d(node values v, T) (non-zero on BOUND/VALVE nodes),
   sbar[4] = dL/d(initial state) on NODE_IN nodes and pbar[4] = dL/d(v_T, v_vis,
   t_vis, t_phi).  Each checkpoint interval is recomputed once with every
   sweep's layers kept on the device, then reversed step by step (transposed
   line solves in reverse sweep order and the adjoint of the Picard merges);
   an interval's recompute runs on a second stream while the interval after it
   is reversed (two tape buffers).
   Any array pointer may be NULL.  One grid, CPU merge order, axes of at most
   256 nodes, no BC_FREE node shared by two segments (else
   CMCX_ERR_UNSUPPORTED).  The current state is unchanged by the backward pass.
   cmcx_adjoint_release frees the checkpoints and the recording.  The recording
   belongs to the parameters, node values, BC kinds and options it was made
   with: cmcx_set_params, cmcx_set_boundary and cmcx_set_options (with another
   path or merge order) clear it, and cmcx_adjoint_backward then returns
   CMCX_ERR_STATE.  cmcx_adjoint_generation gives the current recording's id
   (it changes with every recording and every clearing; 0 = none), so that a
   caller holding several rollouts can tell whether the one it wants a gradient
   for is still the recorded one. */
int cmcx_adjoint_forward(cmcx_solver* s, double dt, int steps, int num_global, int checkpoint_every);
int cmcx_adjoint_backward(cmcx_solver* s, const void* const* seed, void* const* gbar,
                          void* const* sbar, double* pbar);
/* The same with seed, gbar and sbar in device memory (N elements of the field
   type each, on the handle's device): no host transfer of the fields. */
int cmcx_adjoint_backward_device(cmcx_solver* s, const void* const* seed, void* const* gbar,
                                 void* const* sbar, double* pbar);
int cmcx_adjoint_release(cmcx_solver* s);
int cmcx_adjoint_generation(cmcx_solver* s, uint64_t* generation);

/* EvalDivError of fields already in device memory, allocating nothing: the
   variant of cmcx_div_error for callers that cannot allocate device memory
   (code under HIP graph capture, tight serving loops).  No counterpart in the
   reference.  u, v, w: device arrays of the field type (dtype), type: device
   array of node types, x-major with z fastest, dims[3] each >= 3 with fewer
   than 2^31 nodes in all (else CMCX_ERR_TOO_LARGE), h[3] the spacings.  The
   reduction's partial sums live in a fixed device array of the library
   (2 + 4096 doubles, held by each GPU the library runs on until the process
   ends), whatever the grid size; calls that overlap on one GPU race on it.
   Work is ordered on `stream` (NULL = default stream); *err is written after
   the stream has finished the call's work. */
int cmcx_div_error_fields(const void* u, const void* v, const void* w, const uint8_t* type,
                          const int32_t* dims, const double* h, int dtype, void* stream,
                          double* err);

/* Algorithmic HBM bytes of one time step: G*L*(16*w+1)*(S_X+S_Y+S_Z). */
int cmcx_step_bytes(cmcx_solver* s, int num_global, int num_local, double* bytes);
int cmcx_synchronize(cmcx_solver* s);

/* ---- geometry ingestion (the reference's Shape3D / Shape2D input formats) ----
   A triangle-mesh file (in_fmt Shape3D: frames of vertices with velocities,
   then triangles) or a polyline file (in_fmt Shape2D: frames of polylines,
   each "Passive" or "Motion" with a velocity, extruded `depth` deep with the
   bottom raised by `depth_var`) becomes the node arrays of a 3D grid on the
   GPU: rasterised walls, flood-filled outside, fluid inside, the reference's
   BC kinds and node values.  Coordinates in the file are millimetres; spacing
   and depth in metres; align rounds every dimension up to a multiple of 32.
   The arrays feed cmcx_domain (h = spacing, base_t). */
#define CMCX_SHAPE3D 3
#define CMCX_SHAPE2D 2
typedef struct cmcx_geometry cmcx_geometry;
int cmcx_geometry_load(int kind, const char* path, const double* spacing, double base_t, double depth,
                       double depth_var, int align, cmcx_geometry** out);
/* Grid dimensions, the output box (metres; an extruded shape spans z in
   [-depth, 0]), the number of frames and the cycle length (s).  Any pointer may
   be NULL. */
int cmcx_geometry_info(cmcx_geometry* g, int32_t* dims, double* bbox_min, double* bbox_max, int* nframes,
                       double* cycle);
/* Builds the grid of the frame at `time` (frames blended as the reference's
   Prepare) into host arrays of dims[0]*dims[1]*dims[2] elements (any may be NULL). */
int cmcx_geometry_build(cmcx_geometry* g, double time, uint8_t* type, uint8_t* bc_vel, uint8_t* bc_temp,
                        double* vx, double* vy, double* vz, double* T);
/* Grid2D::Prepare(time) (src/FluidSolver2D/Grid2D.cpp:109-266, 456-480) of a
   polyline input: the 2D solvers' grid, dims[0]*dims[1] cells, y fastest (cell
   type, BC kind = BC_NOSLIP as with the reference's NoSlip setting, velocity,
   temperature).  A CMCX_SHAPE2D load with depth = 0 makes only this grid
   (dims[2] = 0; cmcx_geometry_build then fails). */
int cmcx_geometry_build2d(cmcx_geometry* g, double time, uint8_t* type, uint8_t* bc, double* vx,
                          double* vy, double* T);
int cmcx_geometry_destroy(cmcx_geometry* g);

/* ---- COARDS result files (the reference's netCDF output, IO.h:136-388) -----
   Written as netCDF classic with 64-bit offsets (no netCDF library needed),
   with the reference's data model: dimensions x, y, z, t (unlimited); float
   x(x), y(y), z(z), double time(t); the requested variables among u, v, w, T
   as double (t, x, y, z) and d as float (x, y); missing_value 99999 and the
   reference's attributes.  The header announces time/timestep records
   (time = i * timestep), their field slices holding netCDF's fill value until a
   layer is written; writing a layer past the last record appends records. */
typedef struct cmcx_coards_header {
    double bbox_min[3], bbox_max[3];  /* physical extent (x and y axis values: min + i * (max-min)/n) */
    double timestep;                  /* output interval (dt * out_time_steps) */
    double time;                      /* final time */
    int32_t outdims[3];               /* output grid (GetLayer's ox, oy, oz) */
    const char* vars;                 /* output variables, e.g. "u v w T" (d = depth) */
    int xy_degree_units;              /* x, y in degrees (bathymetry input) instead of metres */
    const float* depth;               /* d at the output resolution, x-major (outdims[0]*outdims[1]), or NULL */
} cmcx_coards_header;
int cmcx_coards_write_header(const char* path, const cmcx_coards_header* header);
/* One output layer (GetLayer's arrays: 3 velocity values per node of the field
   type vel_dtype, double T) at record time_index. */
int cmcx_coards_write_layer(const char* path, const void* vel, int vel_dtype, const double* T,
                            int time_index, int dimx, int dimy, int dimz, const char* vars);

/* ---- 2D ADI (the reference's FluidSolver2D path) -------------------------- */
typedef struct cmcx2d_solver cmcx2d_solver;
typedef struct cmcx2d_domain {
    int32_t dims[2];          /* dimx, dimy; node (i,j) is element i*dimy + j */
    double h[2];
    double start_t;           /* Grid2D::startT */
    const uint8_t* type;      /* NodeType */
    const uint8_t* bc;        /* CondData2D.type (BCtype) */
    const double* v[2];
    const double* T;
} cmcx2d_domain;

int cmcx2d_create(const cmcx2d_domain* dom, const cmcx_params* params, int dtype, void* stream,
                  cmcx2d_solver** out);
/* The reference's 2D solver choice (Config.h:25, FluidSolver2D.cpp:65-70):
   CMCX2D_ADI (AdiSolver2D; what cmcx2d_create makes) or CMCX2D_STABLE
   (StableSolver2D.cpp:21-234: explicit advection-diffusion of u and v, then a
   projection whose Poisson equation is relaxed cell by cell, boundary cells
   then inner cells in (i, j) order, until the largest relative change of a
   sweep is below 1e-2; T is carried, not solved).  The Stable solver needs
   every NODE_IN cell off the grid's edge and at most 512 rows (dimx); a
   boundary cell on the edge reads q = 0 beyond the grid (the reference reads
   outside its arrays there). */
#define CMCX2D_ADI 1
#define CMCX2D_STABLE 2
int cmcx2d_create_method(const cmcx2d_domain* dom, const cmcx_params* params, int method, int dtype,
                         void* stream, cmcx2d_solver** out);
int cmcx2d_update_boundaries(cmcx2d_solver* s);
/* Returns CMCX_ERR_PICARD where the reference exits (AdiSolver2D.cpp:303-313),
   with err and iters of the iteration that exited and the state as the step
   found it (as the reference leaves it).  iters: Picard iterations
   done (may be NULL).  The iterations below num_global are queued without a
   host read between them; the host waits once for their errors. */
int cmcx2d_step(cmcx2d_solver* s, double dt, int num_global, int num_local, double* err,
                int* iters);
int cmcx2d_get_fields(cmcx2d_solver* s, void* u, void* v, void* p, void* T);
int cmcx2d_destroy(cmcx2d_solver* s);
/* Grid2D::Prepare(t) between steps (FluidSolver2D.cpp:131-133): new cell
   types, BC kinds and node values for UpdateBoundaries and the next step (the
   reference rebuilds its line lists every step, so the mask may change). */
int cmcx2d_set_grid(cmcx2d_solver* s, const uint8_t* type, const uint8_t* bc, const double* vx,
                    const double* vy, const double* T);
/* Stable solver: relaxation tolerance (the reference's POISSON_ERR_THRESHOLD
   1e-2, StableSolver2D.h:23) and a cap on sweeps per project (the reference has
   none; default 1e6); a step that reaches the cap returns CMCX_ERR_PICARD. */
int cmcx2d_set_poisson(cmcx2d_solver* s, double tol, int64_t max_sweeps);
/* Stable solver: the relaxation sweeps of each project of the last step
   (count = how many projects; up to max written to out). */
int cmcx2d_poisson_sweeps(cmcx2d_solver* s, int32_t* out, int max, int* count);

#ifdef __cplusplus
}
#endif

#endif /* CMCX_H */
