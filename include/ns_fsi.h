/*
 * ns_fsi.h — C-ABI of the MI355X-native explicit compressible Navier–Stokes time step.
 *
 * One handle = one simulation case on one GPU, one HIP stream. All device memory is owned by the
 * handle; host pointers are borrowed only for the duration of a call. Every entry point returns an
 * ns_status; a handle keeps the text of its last error (ns_last_error). No exception crosses this
 * boundary. Handles are not thread-safe; calls on one handle must be serialised by the caller.
 *
 * The reference (Frederik-Kris/3D-NS-FSI) has no FFI: the path sits behind C++ member calls. Each
 * entry point below names the reference call it replaces (paths relative to the reference root).
 * Field arrays crossing this boundary are dense NI*NJ*NK arrays in x-fastest order,
 * index = (k*NJ + j)*NI + i; the reference stores k fastest, (i*NJ + j)*NK + k (src/Array3D.h:42).
 */
#ifndef NS_FSI_H_
#define NS_FSI_H_

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define NS_FSI_ABI_VERSION 2

typedef enum ns_status {
    NS_OK = 0,
    NS_ERR_CONFIG = 1,       /* ConfigFile parse / type / value error (src/ConfigSettings.cpp:31-163) */
    NS_ERR_GEOMETRY = 2,     /* body/image-point invariant violated (src/Boundary.cpp:724, :36) */
    NS_ERR_NONFINITE = 3,    /* non-finite conserved value found (src/Mesh.h:48-61) */
    NS_ERR_HIP = 4,          /* HIP runtime error, or no gfx950 device */
    NS_ERR_ARGUMENT = 5,     /* null handle, bad pointer or out-of-range argument */
    NS_ERR_STATE = 6,        /* call out of order (e.g. step before init) */
    NS_ERR_UNSUPPORTED = 7   /* configuration the device path does not implement */
} ns_status;

typedef enum ns_bc_type {
    NS_BC_INLET = 0, NS_BC_OUTLET = 1, NS_BC_PERIODIC = 2, NS_BC_SYMMETRY = 3
} ns_bc_type;

/* Face order used by ns_config.bc[]. */
typedef enum ns_face { NS_XMIN = 0, NS_XMAX = 1, NS_YMIN = 2, NS_YMAX = 3, NS_ZMIN = 4, NS_ZMAX = 5 } ns_face;

typedef enum ns_body_type { NS_BODY_NONE = 0, NS_BODY_CYLINDER = 1, NS_BODY_SPHERE = 2 } ns_body_type;

typedef enum ns_ic_type {
    NS_IC_STAGNATION = 0,          /* src/Solver.cpp:31-51 (the reference default) */
    NS_IC_UNIFORM_OBLIQUE = 1,     /* src/Solver.cpp:54-75 */
    NS_IC_FREESTREAM_GAUSSIAN = 2, /* u = M_0 + N(0, a^2), v = N(0,a^2), w = N(0,a^2), p' = T' = 0 */
    NS_IC_FREESTREAM_UNIFORM = 3   /* u = M_0 + U[-a,a], v = U[-a,a], w = U[-a,a], p' = T' = 0 */
} ns_ic_type;

typedef enum ns_precision { NS_FP64 = 0, NS_FP32 = 1 } ns_precision;

typedef enum ns_field {
    NS_RHO = 0, NS_RHO_U = 1, NS_RHO_V = 2, NS_RHO_W = 3, NS_RHO_E = 4,
    NS_U = 5, NS_V = 6, NS_W = 7, NS_P = 8, NS_T = 9, NS_MU = 10, NS_KAPPA = 11
} ns_field;

/* Node classes; values equal the reference enum order (src/Node.h:27-34). */
typedef enum ns_node_type {
    NS_FLUID_INTERIOR = 0, NS_FLUID_EDGE = 1, NS_FLUID_GHOST = 2, NS_SOLID_INACTIVE = 3, NS_SOLID_GHOST = 4
} ns_node_type;

#define NS_MAX_FILTER 16
#define NS_MAX_BODIES 8

/* One immersed body (the reference's CylinderBody / SphereBody, src/Boundary.h:255-346). */
typedef struct ns_body {
    int32_t type;       /* ns_body_type */
    int32_t axis;       /* 0,1,2 = x,y,z (cylinder only) */
    double centre[3];
    double radius;
} ns_body;

/* Every ConfigFile key of the reference (ConfigFile:22-125, src/ConfigSettings.h:39-74) plus the
 * extension keys for what the reference hard-codes (src/Mesh.cpp:33-57, src/Solver.cpp:26,
 * src/Boundary.cpp:267). Fill with ns_config_default(), then override. */
typedef struct ns_config {
    int32_t NI, NJ, NK;
    int32_t stop_criterion;            /* 0 = "timesteps", 1 = "end_time" */
    uint64_t stop_time_level;
    double t_end;
    double conv_stability_limit;
    double visc_stability_limit;
    int32_t n_filter_intervals;
    uint32_t filter_intervals[NS_MAX_FILTER];
    int32_t n_filter_change_times;
    double filter_change_times[NS_MAX_FILTER];
    double status_report_interval;
    int32_t save_density, save_momentum, save_energy, save_velocity;
    int32_t save_pressure, save_temperature, save_viscosity, save_thermal_cond;
    int32_t save_ic, save_final, save_intervals;
    double save_period, save_intervals_start_time, save_intervals_end_time;
    int32_t conv_history;              /* 0 = "none", 1 = "density", 2 = "all" */
    double gamma, Pr, R, Re, sutherlands_C2, M_0, T_0, L_x, L_y, L_z;
    /* derived: machine-precision buffer eps*(L_x+L_y+L_z) (src/ConfigSettings.cpp:160-163) */
    double machine_precision_buffer;
    /* ---- extension keys (defaults reproduce the reference's hard-coded layout) ---- */
    int32_t bc[6];                     /* ns_bc_type per ns_face */
    int32_t body_type;                 /* ns_body_type */
    int32_t body_axis;                 /* 0,1,2 = x,y,z (cylinder only) */
    double body_centre[3];
    double body_radius;
    int32_t ic_type;                   /* ns_ic_type */
    uint64_t ic_seed;
    double ic_amplitude;
    double inlet_ramp_time;            /* reference: 10 (src/Boundary.cpp:267) */
    int32_t precision;                 /* ns_precision */
    int32_t kahan;                     /* FP32 only: compensated stage combine */
    /* Further immersed bodies after the one above, classified and applied in this order as the reference's
     * vector of bodies (src/Mesh.cpp:52-57, src/Mesh.h:91-92). Their bounding boxes (radius + 3 nodes) must be
     * disjoint from each other's. */
    int32_t n_extra_bodies;
    int32_t pad_bodies;
    ns_body extra_bodies[NS_MAX_BODIES - 1];
} ns_config;

typedef struct ns_handle ns_handle;

/* Defaults: the reference's shipped ConfigFile and its hard-coded geometry. */
void ns_config_default(ns_config* cfg);

/* Parse libconfig-syntax text (the reference ConfigFile format) into cfg. Replaces
 * ConfigSettings::loadSettings (src/ConfigSettings.cpp:18-28). On NS_ERR_CONFIG, errbuf (if not
 * null) receives the message. */
ns_status ns_config_parse(const char* text, ns_config* cfg, char* errbuf, size_t errbuf_len);

/* Replaces Solver::Solver / Mesh::Mesh (src/Solver.cpp:13-18, src/Mesh.cpp:10-21). Allocates the
 * device fields. Fails with NS_ERR_HIP when no gfx950 device is present. */
ns_status ns_create(const ns_config* cfg, ns_handle** out);

/* Replaces Solver::initialize (src/Solver.cpp:22-28): boundaries, node classification (device
 * kernel + compaction), ghost geometry, initial condition, first dt. */
ns_status ns_init(ns_handle* h);

/* Replaces the IC upload of applyStagnation_IC (src/Solver.cpp:31-51) with an arbitrary field:
 * U5 holds rho', rho*u, rho*v, rho*w, rho*E' as 5 consecutive x-fastest NI*NJ*NK arrays. Primitive
 * and transport fields are derived everywhere; dt is recomputed. Call after ns_init, before steps. */
ns_status ns_set_initial_field(ns_handle* h, const double* U5);

/* Replaces nsteps calls of Solver::marchTimeStep(t, timeLevel) (src/Solver.cpp:130-167). t, level
 * and dt (any may be null) receive the values after the last step; dt is the size of the next step
 * (public member Solver::dt, src/Solver.h:31). */
ns_status ns_step(ns_handle* h, int32_t nsteps, double* t, uint64_t* level, double* dt);

/* Runs Simulation::run's loop (src/Simulation.cpp:19-52) until the stopping criterion holds. */
ns_status ns_run(ns_handle* h, double* t, uint64_t* level);

/* Current t, timeLevel and dt without stepping. */
ns_status ns_get_time(ns_handle* h, double* t, uint64_t* level, double* dt);

/* Copy one field (ns_field) to host, x-fastest NI*NJ*NK doubles. The conserved fields are the
 * state Mesh::conservedVariables; primitives and transport are the values the reference keeps in
 * Mesh::primitiveVariables / transportProperties (src/Mesh.h:82-85). */
ns_status ns_get_field(ns_handle* h, int32_t which, double* out);

/* Node classes (ns_node_type), x-fastest NI*NJ*NK bytes (Mesh::nodeType, src/Mesh.h:88). */
ns_status ns_get_node_types(ns_handle* h, uint8_t* out);

/* Number of solid ghost nodes and filter nodes of the immersed body (src/Boundary.h:200-210). */
ns_status ns_get_counts(ns_handle* h, int64_t* n_ghosts, int64_t* n_filter, int64_t* n_interior);

/* Ghost nodes (GhostNode, src/Node.h:15-22), sorted by reference index (i*NJ+j)*NK+k:
 * ijk[3*n] node indices, bi[3*n] body intercept, ip[3*n] image point, box[3*n] lower corner of the
 * 8-node image-point box (src/SmallVectors.h:108-127). Any pointer may be null. */
ns_status ns_get_ghosts(ns_handle* h, int32_t* ijk, double* bi, double* ip, int32_t* box);

/* Filter nodes (src/Boundary.cpp:1010-1012, :1099-1101), sorted by reference index, ijk[3*n]. */
ns_status ns_get_filter_nodes(ns_handle* h, int32_t* ijk);

/* Norm-of-change history (Solver::getConvergenceHistory, src/Solver.h:27; src/Mesh.cpp:171-178):
 * var 0..4 = rho..rho_E; copies min(n, available) values, *count receives the number available. */
ns_status ns_norm_history(ns_handle* h, int32_t var, double* out, int64_t n, int64_t* count);

/* Force on the immersed body after each step taken while the run-time option "body_force" is 1 (no
 * reference counterpart: the reference computes no forces). Definition: the discrete momentum balance of
 * the time step. Summed over all FluidInterior nodes, the RK4 right-hand side of the momentum equations
 * telescopes into face fluxes across the outer boundary and across the faces between FluidInterior and
 * ghost nodes; the latter are the momentum per unit time the body gives the fluid. The force on the body
 * is minus that sum times dx*dy*dz, evaluated in FP64 on the state after the step's last boundary
 * application, in the nondimensional units of the state (density rho_0, velocity c_0, length 1).
 * Drag coefficient: C_D = Fx / (0.5 * M_0^2 * A_ref) with A_ref = pi r^2 (sphere) or 2 r times the span
 * of FluidInterior planes along the cylinder axis ((N_axis - 2) * h_axis).
 * out receives min(n, available) entries of 4 doubles (timeLevel after the step, Fx, Fy, Fz); *count
 * the number available. */
ns_status ns_force_history(ns_handle* h, double* out, int64_t n, int64_t* count);

/* Prescribed step sizes (no reference counterpart): dts[L] is the size of the step that leaves time level
 * L; the stability limit is still scanned but not used. n must exceed the last time level stepped to.
 * Null / n = 0 returns to the stability limit. Used to differentiate runs at a frozen step sequence. */
ns_status ns_set_dt_schedule(ns_handle* h, const double* dts, int64_t n);

/* Inlet velocity profile (extension; the reference's inlet is uniform): the inlet's normal velocity at a
 * node of the inlet plane is min(1, t / inletRampTime) * M_0 * q[a + n_a * b], where (a, b) are the node's
 * two in-plane indices in (i, j, k) order and n_a the extent of the first (x inlet: q[j + NJ * k]).
 * n must equal the plane's node count. Null restores the uniform inlet. */
ns_status ns_set_inflow_profile(ns_handle* h, const double* q, int64_t n);

/* Discrete adjoint of the drag objective (no reference counterpart). From the initial state (time level 0,
 * no periodic density filter, an immersed body), runs nsteps steps in the case's precision and evaluates
 *   J = 1/(n1 - n0 + 1) * sum_{L=n0..n1} C_D(U^L),  C_D = Fx / (0.5 M_0^2 A_ref)
 * (Fx and A_ref as in ns_force_history), then the hand-written reverse sweep (transposed right-hand side,
 * combine and boundary application of every stage; checkpoints every `checkpoint` steps with
 * recomputation) at the step sizes of that run held fixed. The reverse sweep is FP64 in both precisions:
 * for an fp32 case it is the adjoint of the FP64 step linearised at the FP32 run's stage states (forces,
 * J and step sizes reduced in FP64 either way). Outputs (any may be null):
 *   J; grad_u0 = dJ/d(initial conserved fields), 5 x-fastest arrays as ns_get_field;
 *   grad_inflow = dJ/d(inlet profile factors q), the inlet plane's size (see ns_set_inflow_profile);
 *   grad_re = dJ/dRe; dts = the nsteps + 1 step sizes by time level (pass them to ns_set_dt_schedule to
 *   evaluate J at the same frozen steps, e.g. for finite differences).
 * The handle is left at time level nsteps with its previous step-size rule. */
ns_status ns_adjoint_drag(ns_handle* h, int32_t nsteps, int32_t n0, int32_t n1, int32_t checkpoint, double* J,
                          double* grad_u0, double* grad_inflow, double* grad_re, double* dts);

/* FP64 totals over FluidInterior nodes of the current state: sum of each conserved variable times
 * dx*dy*dz (mass, momenta, energy). No reference counterpart; conservation monitor. */
ns_status ns_totals(ns_handle* h, double out[5]);

/* Grid spacing and origin offset (Mesh::dx/dy/dz, positionOffset; src/Mesh.cpp:44-49). */
ns_status ns_get_geometry(ns_handle* h, double spacing[3], double offset[3]);

/* Measurement hook (no reference counterpart): runs nsteps time steps with HIP events recorded on the
 * handle's stream around each kernel group and returns the mean duration in milliseconds of one launch
 * of each group: ms[0] the RK4 stage kernel, ms[1] the primitive-variable pass (0 when fused),
 * ms[2] the boundary kernels of one stage, ms[3] the end-of-step reductions; launches[g] = launches per
 * step. The steps advance the solution like ns_step. */
ns_status ns_profile(ns_handle* h, int32_t nsteps, double ms[4], int32_t launches[4]);

/* Run-time option (no reference counterpart):
 *   "stage_kernel" 1 (default) fused stage kernel, 0 the unfused pair (primitive pass + per-node RHS);
 *   "tile_ty"      output rows per block of the fused stage (4, 6, 8 or 14; default 8);
 *   "tile_kc"      z-planes per block of the fused stage (default 64);
 *   "autotune"     after ns_init: time every (tile_ty, tile_kc) pair (tile_kc from 12, 16, 20, 24, 32,
 *                  40, 48, 64, 96, 128) on this case with `value` launches
 *                  each (the state is not changed) and keep the fastest;
 *   "fused_scans"  1 (default): the end-of-step scans (norms, dt maxima, finite check) run inside the
 *                  last fused stage kernel plus a small boundary pass; 0: one separate pass over the grid
 *                  (the unfused stage kernel always uses the separate pass);
 *   "body_force"   1: record the force on the body after every step (ns_force_history), 0 (default);
 *   "graph"        1: ns_step replays each time step as one captured HIP graph (no per-kernel launch
 *                  cost), 0 (default): launches kernel by kernel. Same results either way.
 *   "global_access" FP32 fused stage: 1 addresses the fields by per-field global pointers (the form
 *                  used for states beyond 4 GiB, which one buffer resource cannot span), 0 (default)
 *                  by buffer resources where the state fits. Same results either way.
 *   "fused_outlet" 1 (default): an outlet face at x = NI-1 is applied by the fused stage kernel itself
 *                  when a row's last x-segment (62 nodes per segment) holds at least three nodes and the
 *                  face is independent of the faces applied before it; 0 leaves it to the edge-face pass.
 *                  Same results either way.
 * Results do not depend on the tile shape (the per-node arithmetic is the same). */
ns_status ns_set_option(ns_handle* h, const char* name, int64_t value);

/* Checkpoint and restart (no reference counterpart: the reference writes output snapshots only). ns_save writes
 * to the file `path` everything a later step reads: the state buffers (including the previous step's values the
 * reference keeps in its "Old" arrays), the stored primitive fields, the pre-filter density and energy of the
 * immersed-filter nodes, the FP32 compensation carry, the initial primitives of untouched solid nodes, t, dt,
 * timeLevel and the norm and force histories. ns_restore reads such a file into an initialised handle created
 * with the same grid, boundary types, body, initial-state generator and physical parameters (stopping, output
 * and stability settings may differ; NS_ERR_ARGUMENT otherwise): a restarted run continues bit for bit as the
 * uninterrupted one would. Run-time options, tile, prescribed step sizes and the inlet profile are not part of
 * a checkpoint. */
ns_status ns_save(ns_handle* h, const char* path);
ns_status ns_restore(ns_handle* h, const char* path);

/* Tile shape in use by the fused stage (after autotuning, the chosen one). */
ns_status ns_get_tile(ns_handle* h, int32_t* ty, int32_t* kc);

/* Last error text of this handle (never null). */
const char* ns_last_error(ns_handle* h);

/* Releases all device memory. Null is accepted. */
void ns_destroy(ns_handle* h);

/* ---- Batched cases (no reference counterpart: the reference runs one case per process).
 * Many independent cases (each its own ns_config: grid size, body, Mach, Re, initial field, stability
 * limits) advance together on one GPU; every kernel of the time step runs once for all cases, each case
 * keeps its own dt / t / timeLevel (src/Solver.cpp:80-126 per case), and a step replays as one captured
 * HIP graph. Results of each case equal those of ns_step on that case alone with the same tile shape.
 * Restrictions: precision fp32, the same kahan setting in every case, no periodic density filter
 * (NS_ERR_UNSUPPORTED otherwise). */
typedef struct ns_batch ns_batch;

/* Creates the cases (like ns_create for each config). On error *out still receives a handle that holds
 * the error text; destroy it with ns_batch_destroy. */
ns_status ns_batch_create(const ns_config* cfgs, int32_t n, ns_batch** out);
/* Solver::initialize of every case (classification, ghost geometry, initial field, first dt). */
ns_status ns_batch_init(ns_batch* b);
/* nsteps time steps of every case (each at its own dt). NS_ERR_NONFINITE if any case holds a
 * non-finite value afterwards. */
ns_status ns_batch_step(ns_batch* b, int32_t nsteps);
/* Time, time level and next dt of case c. */
ns_status ns_batch_get_time(ns_batch* b, int32_t c, double* t, uint64_t* level, double* dt);
/* Field `which` of case c (ids and layout as ns_get_field). */
ns_status ns_batch_get_field(ns_batch* b, int32_t c, int32_t which, double* out);
/* Norm-of-change history of case c (as ns_norm_history). */
ns_status ns_batch_norm_history(ns_batch* b, int32_t c, int32_t var, double* out, int64_t n, int64_t* count);
/* Options: "graph" (1 default), "tile_ty" / "tile_kc" (tile shape of the batched stage, shared by all
 * cases), "autotune" (time every tile shape on this batch with `value` launches each and keep the
 * fastest; states unchanged), "persistent" (0 default: one block per tile; 1: one block per resident
 * slot, each XCD's blocks pulling tiles from their own range and then from the others'; same results). */
ns_status ns_batch_set_option(ns_batch* b, const char* name, int64_t value);
ns_status ns_batch_get_tile(ns_batch* b, int32_t* ty, int32_t* kc);
const char* ns_batch_last_error(ns_batch* b);
void ns_batch_destroy(ns_batch* b);

/* ---- Allocation-free entry point.
 * No counterpart in the reference: it is for callers that cannot allocate device memory (code under HIP graph
 * capture, tight serving loops, harnesses that forbid hipMalloc). Same result as ns_totals, bit for bit, on
 * fields the caller holds in device memory: fields = 5 consecutive dense x-fastest NI*NJ*NK arrays (rho',
 * rho*u, rho*v, rho*w, rho*E') of double (precision NS_FP64) or float (NS_FP32); node_types = NI*NJ*NK bytes
 * holding ns_node_type values, the layout ns_get_node_types writes, in device memory; spacing = dx, dy, dz;
 * out = 5 doubles of device memory receiving the sums over NS_FLUID_INTERIOR nodes times dx*dy*dz.
 * stream: a hipStream_t (null: the default stream); the call only enqueues work there and returns.
 * It allocates nothing: its scratch is a fixed 256 x 5 doubles held by the library on each GPU for the life of
 * the process, shared by every call on that GPU, so calls that overlap on one GPU race. Grids beyond 2^36
 * nodes are refused with NS_ERR_ARGUMENT; the text of the last error is ns_noalloc_last_error(). */
ns_status ns_totals_noalloc(const void* fields, int32_t precision, const uint8_t* node_types, int32_t NI, int32_t NJ,
                            int32_t NK, const double spacing[3], double* out, void* stream);
const char* ns_noalloc_last_error(void);

/* Library ABI version, NS_FSI_ABI_VERSION. */
int32_t ns_abi_version(void);

#ifdef __cplusplus
}
#endif

#endif /* NS_FSI_H_ */
