/*
 * erod.h — C ABI of the MI355X-native Cosserat-rod block timestep.
 *
 * A batch is one device-resident, ghost-padded multi-rod block: the
 * reference's blocks::Block<CosseratRod> (elastica/Systems/CosseratRods/Block.hpp)
 * with every per-variable array moved into HBM.  Rod i starts at the same
 * slot s_i in the node, element and Voronoi arrays, s_{i+1} = s_i + n_i + 2
 * (Block.hpp:401-443); each field is component-major (component c of slot k
 * at c * slots + k; a director is 9 lines, line 3p + r holding Q(p, r)).
 *
 * All calls return EROD_OK (0) or a negative erod_status.  A batch owns a
 * HIP stream; calls on one batch are ordered on that stream and are not
 * thread-safe.  (A fused launch whose plan has several tile groups forks
 * some of them onto a second stream of the batch and joins it back before
 * the call returns, so the order holds.)  Host pointers are plain host
 * memory.
 */
#ifndef EROD_H
#define EROD_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef enum {
  EROD_OK = 0,
  EROD_ERR_INVALID = -1,      /* bad argument (size, index, field, n_elem < 2) */
  EROD_ERR_HIP = -2,          /* a HIP runtime call failed; see erod_last_error */
  EROD_ERR_UNSUPPORTED = -3,  /* the requested path cannot run this batch */
  EROD_ERR_NO_DEVICE = -4,    /* no gfx950 device visible */
  EROD_ERR_NONFINITE = -5     /* a non-finite state value was detected */
} erod_status;

/* Field ids: the reference's tags (elastica/Systems/CosseratRods/Tags.hpp and
 * the component variable lists).  Rigidity, inertia and inverse inertia are
 * stored as their diagonals (3 lines): the reference's 3x3 tensors are
 * diagonal and its matvec with them only adds exact zeros
 * (WithDiagonalLinearHyperElasticModel.hpp:358, WithRodKinematics.hpp:358-364). */
typedef enum {
  /* nodes (n + 1 per rod) */
  EROD_POSITION = 0, EROD_VELOCITY, EROD_ACCELERATION, EROD_MASS, EROD_INV_MASS,
  EROD_INTERNAL_LOADS, EROD_EXTERNAL_LOADS, EROD_FORCE_DAMPING_RATE,
  /* elements (n per rod) */
  EROD_DIRECTOR, EROD_ANGULAR_VELOCITY, EROD_ANGULAR_ACCELERATION,
  EROD_REFERENCE_ELEMENT_LENGTH, EROD_ELEMENT_LENGTH, EROD_ELEMENT_DILATATION,
  EROD_TANGENT, EROD_ELEMENT_DIMENSION, EROD_ELEMENT_VOLUME,
  EROD_SHEAR_STRETCH_RIGIDITY, EROD_MASS_SECOND_MOMENT_OF_INERTIA,
  EROD_INV_MASS_SECOND_MOMENT_OF_INERTIA, EROD_SHEAR_STRETCH_STRAIN,
  EROD_REFERENCE_SHEAR_STRETCH_STRAIN, EROD_INTERNAL_STRESS, EROD_INTERNAL_TORQUES,
  EROD_EXTERNAL_TORQUES, EROD_TORQUE_DAMPING_RATE, EROD_DUMMY_ELEMENT_VECTOR,
  EROD_DUMMY_ELEMENT_VECTOR2,
  /* voronoi (n - 1 per rod) */
  EROD_REFERENCE_VORONOI_LENGTH, EROD_VORONOI_LENGTH, EROD_VORONOI_DILATATION,
  EROD_CURVATURE, EROD_REFERENCE_CURVATURE, EROD_BENDING_TWIST_RIGIDITY,
  EROD_INTERNAL_COUPLE, EROD_DUMMY_VORONOI_VECTOR,
  EROD_N_FIELDS
} erod_field;

typedef enum { EROD_NODE = 0, EROD_ELEMENT = 1, EROD_VORONOI = 2 } erod_stagger;

typedef enum {
  EROD_FP64 = 0,       /* everything in FP64 */
  EROD_FP32_BULK = 1   /* FP64 state, integrator and exponential map; FP32 strains/stresses/loads */
} erod_precision;

typedef enum { EROD_RAMP_NONE = 0, EROD_RAMP_LINEAR = 1, EROD_RAMP_SIN = 2 } erod_ramp;

typedef struct erod_batch erod_batch;

/* ---- lifetime and layout ------------------------------------------------ */

/* Replaces Block::emplace for n_rods rods at once (Block.hpp:568-611): sizes
 * the ghost-padded arrays, fills every ghost slot (Block.hpp:815-849) and
 * zeroes the rest.  `damped` selects CosseratRod (1) or
 * CosseratRodWithoutDamping (0) (Aliases/CosseratRod*.hpp). */
int erod_batch_create(const uint64_t* n_elems, uint64_t n_rods, int damped, erod_batch** out);
int erod_batch_destroy(erod_batch* b);
/* Device ordinal the batch lives on. */
int erod_batch_device(const erod_batch* b, int* device);
/* Total element slots (ghosts included); nodes = slots + 1, voronoi = slots - 1. */
int erod_batch_slots(const erod_batch* b, uint64_t* element_slots);
/* Number of rods (slots), and, when `capacity` is not null, each rod's slot
 * capacity in elements (n_rods entries; Block::emplace's n of each system,
 * Block.hpp:568-611). */
int erod_batch_rods(const erod_batch* b, uint64_t* n_rods, uint64_t* capacity);
/* First slot of a rod (Block::get_start_of_slice, Block.hpp:216-227). */
int erod_rod_start(const erod_batch* b, uint64_t rod, uint64_t* start);
/* Lines and stagger of a field. */
int erod_field_shape(int field, int* ncomp, int* stagger);

/* ---- data movement ------------------------------------------------------- */

/* Whole-block copies: host holds ncomp lines of the stagger's slot count. */
int erod_upload_field(erod_batch* b, int field, const double* host);
int erod_download_field(erod_batch* b, int field, double* host);
/* One rod's slice (blocks::get<tag>(slice)): host holds ncomp lines of the
 * rod's n + 1 / n / n - 1 values. */
int erod_upload_rod(erod_batch* b, int field, uint64_t rod, const double* host);
int erod_download_rod(erod_batch* b, int field, uint64_t rod, double* host);
/* Rewrite every ghost slot (fill_ghosts, Block.hpp:815-849). */
int erod_fill_ghosts(erod_batch* b);

/* ---- the three per-block calls of the stepper (op-by-op FP64 path) ------- */

/* PositionVerlet::KinematicSubStep (PositionVerlet.hpp:60-69): x += (dt/2) v,
 * Q <- exp(-(dt/2)[w]x) Q over every slot. */
int erod_kinematic_update(erod_batch* b, double dt);
/* The same update with any coefficient: x += c v, Q <- exp(-c [w]x) Q, that
 * is sys.kinematic_states() += c * sys.kinematic_rates(t) for the
 * coefficient a symplectic algorithm forms (PositionVerlet.hpp:60-69 uses
 * c = dt / 2, VelocityVerlet.hpp:74 c = dt). */
int erod_kinematic_update_by(erod_batch* b, double c);
/* Every slot of a field (ghosts included) set to `value`: the block-wide
 * assignment of a scalar to a variable, as in external_loads(system) = 0.0
 * (SymplecticSteppers.hpp:178-186; Access.hpp:229-247). */
int erod_fill_field(erod_batch* b, int field, double value);
/* ExternalLoads = ExternalTorques = 0, then compute_internal_dynamics
 * (SymplecticSteppers.hpp:178-188, ElasticityInterface.hpp:360-365). */
int erod_compute_internal_dynamics(erod_batch* b, double time);
/* PositionVerlet::DynamicSubStep with update_acceleration
 * (PositionVerlet.hpp:71-77, WithRodKinematics.hpp:683-714). */
int erod_dynamic_update(erod_batch* b, double dt);

/* Device-side hook descriptors that replace per-step host lambdas.
 * Clamp: value constraint after the kinematic half step (position of a node,
 * director of an element), rate constraint after the dynamic
 * step (velocity and angular velocity zeroed); the constraint's per-index
 * constrain_primitives / constrain_rates (ConstrainPrimitives /
 * ConstrainRates, elastica/Simulator/Interface/Constraints/ExecutionPolicy.hpp:591-593,
 * elastica/Simulator/Constraints/Constraint.hpp:290-313).  index runs from 0
 * to n (the rod's active length) and holds node `index` with element
 * min(index, n - 1): index n clamps the rod's last end as index 0 clamps its
 * first.  A rod takes up to two clamps at different indices (a clamp at an
 * index that has one replaces it; clamps at n - 1 and n, which would both
 * hold element n - 1, are refused); a third returns EROD_ERR_UNSUPPORTED, an
 * index past n EROD_ERR_INVALID. */
int erod_add_clamp(erod_batch* b, uint64_t rod, uint64_t index, const double x[3], const double Q[9]);
/* ExternalLoads(:, last node) = F * ramp(t) before the dynamic step
 * (examples/timoshenko/TimoshenkoCaseStudy.cpp:128-141; ramps
 * elastica/Utilities/Rampup.hpp:95-150). */
int erod_add_end_force(erod_batch* b, uint64_t rod, const double F[3], int ramp, double start, double interval);
/* ExternalLoads += g (x) Mass (GravityField.hpp:239-246). */
int erod_set_gravity(erod_batch* b, const double g[3]);
/* Per-node plane contact: a node whose height n . x above the plane
 * n . x = z0 is below offset gets a linear spring-dashpot normal force
 * max(0, k delta + c v_n) along n and friction min(ct |v_t|, mu f_N) against
 * its slip in the plane (LinearSpringDashpot.cpp:49-71 applied per node).
 * The normal n is (0, 0, 1) unless erod_set_plane_normal sets it. */
int erod_set_plane_contact(erod_batch* b, double k, double c, double mu, double ct, double z0, double offset);
/* Unit normal of the contact plane (PlanePrimitive's normal, distance along it
 * as in CapsulePlane.hpp:73-80); scaled to unit length, zero or non-finite
 * normals are refused. */
int erod_set_plane_normal(erod_batch* b, const double n[3]);
int erod_apply_post_kinematic(erod_batch* b, double time);
int erod_apply_pre_dynamic(erod_batch* b, double time);
int erod_apply_post_dynamic(erod_batch* b, double time);

/* n full position-Verlet steps through the op-by-op kernels
 * (SymplecticSteppers.hpp:133-217); *time advances by n dt and every rod
 * clock (erod_get_rod_time) is set to it. */
int erod_step(erod_batch* b, double* time, double dt, uint64_t n);

/* ---- fused / persistent path ---------------------------------------------- */

/* Per-rod timestep for the fused path (rods are independent; default: 0 = unset). */
int erod_set_rod_dt(erod_batch* b, const double* dt_per_rod);
/* Per-rod clock (seconds), read back after fused steps. */
int erod_get_rod_time(erod_batch* b, double* time_per_rod);
int erod_set_rod_time(erod_batch* b, const double* time_per_rod);
/* Tile / launch shape of the persistent kernels: a workgroup owns whole rods
 * totalling at most lanes_per_tile elements (64 ... 1024), each thread holding
 * elements_per_thread of them (1, 2 or 4); 0 leaves a value free.
 * steps_per_launch: steps a workgroup advances with its rods' state held on
 * chip (0 = 100).  The rods of a batch share one shape.  Without a request or
 * an autotune, a batch takes the smallest capacity that holds its longest rod
 * (one element per thread up to 256 lanes, two from 512); rods longer than
 * the tile run as pieces (erod_set_fused_pieces). */
int erod_set_fused_config(erod_batch* b, int lanes_per_tile, int elements_per_thread, int steps_per_launch);
int erod_get_fused_config(const erod_batch* b, int* lanes_per_tile, int* elements_per_thread, int* steps_per_launch);
/* A second group of tiles: a rod of at most lanes_per_tile elements that finds
 * no room beside longer rods in the main tiles runs in tiles of this smaller
 * shape (elements_per_thread 0 = either), launched after the main ones; 0, 0
 * turns it off.  For fused FP64 on ragged batches, whose long-rod tiles spill
 * registers while shorter tiles do not.  No counterpart in the reference (its
 * blocks are sized by one policy, BlockingPolicy.hpp:291-321).  The getter
 * reports the group's shape and tile count of the current plan (0 if none). */
int erod_set_fused_split(erod_batch* b, int lanes_per_tile, int elements_per_thread);
int erod_get_fused_split(const erod_batch* b, int* lanes_per_tile, int* elements_per_thread, uint64_t* tiles);
/* Rods longer than the main tile, or longer than lanes_per_tile when it is
 * not 0, run as pieces in tiles of that shape (0: the main shape;
 * elements_per_thread 0 = either): each piece owns a part of its rod and
 * recomputes `halo` elements (0 = 16, at least 2) of each neighbouring piece
 * from the state they had when the launch began, so that pieces need no
 * exchange within a launch; a launch over pieces then takes at most halo / 2
 * steps, and the pieces' group is launched as often as needed.  Results do
 * not depend on the cut (bitwise).  For rods longer than any tile (the
 * reference's blocks emplace rods of any length, Block.hpp:568-611) and for
 * fused FP64, whose long-rod tiles spill registers.  The getter reports the
 * pieces' shape, halo, tiles and pieces in the current plan (0 if none). */
int erod_set_fused_pieces(erod_batch* b, int lanes_per_tile, int elements_per_thread, int halo);
int erod_get_fused_pieces(const erod_batch* b, int* lanes_per_tile, int* elements_per_thread, int* halo,
                          uint64_t* tiles, uint64_t* pieces);
/* Times every feasible tile shape on `trial_steps` steps from a snapshot of
 * the state (restored afterwards) and keeps the fastest; reports its ms/step.
 * No counterpart in the reference (its blocks are sized by a fixed policy,
 * elastica/Simulator/Configuration/ModuleConfiguration/BlockingPolicy.hpp:291-321). */
int erod_autotune_fused(erod_batch* b, int precision, int trial_steps, double* best_ms_per_step);
/* n position-Verlet steps with the fused persistent kernels at the given
 * precision.  Keeps x, v, Q, w (and the rod clocks) current; derived fields
 * are refreshed on request by erod_compute_internal_dynamics. */
int erod_step_fused(erod_batch* b, int precision, uint64_t n);
/* Same, replaying a captured hipGraph of `chunk` steps n / chunk times, then
 * launching the n % chunk steps left directly (in launches of
 * steps_per_launch, the last one shorter).  A chunk whose graph is a single
 * kernel (one launch, one tile group) goes out as that kernel's launch
 * instead of a replay: same work, less device time. */
int erod_step_fused_graph(erod_batch* b, int precision, uint64_t n, uint64_t chunk);
/* Capture the hipGraph of one chunk of `chunk` fused steps now (as
 * erod_step_fused_graph would on its first replay), without launching it, so
 * that a later erod_step_fused_graph of that chunk starts replaying at once. */
int erod_capture_fused_graph(erod_batch* b, int precision, uint64_t chunk);
/* Re-orthonormalise every director (one Newton-Schulz polar step, FP64). */
int erod_reorthonormalize(erod_batch* b);
/* Schedule the same step on the fused paths (erod_step_fused,
 * erod_step_fused_graph, and inside their captured graphs): after the launch
 * during which the batch's fused step count, counted from this call, reaches
 * a multiple of `every`, the directors of every rod with steps left (no step
 * budget, or one not yet spent) are re-orthonormalised; rods whose budget is
 * spent keep their state.  every = 0 (the default) schedules nothing.  A
 * graph chunk that is a multiple of `every` replays one captured schedule;
 * other chunks capture again whenever a replay starts at another phase.
 * No reference counterpart: its directors are never re-orthonormalised. */
int erod_set_reortho_interval(erod_batch* b, uint64_t every);
/* The interval and the steps taken since the last scheduled step. */
int erod_get_reortho_interval(const erod_batch* b, uint64_t* every, uint64_t* phase);
/* Average duration (ms) of one launch of the fused step kernel over the last
 * erod_step_fused / erod_step_fused_graph call: HIP events recorded on the
 * batch stream before the first and after the last launch, divided by the
 * number of launches (reported in *launches).  The gaps between back-to-back
 * launches are included. */
int erod_last_kernel_ms(const erod_batch* b, double* ms, uint64_t* launches);

/* ---- continuous batching: slots, budgets, rod copies ------------------------ */

/* Active length of rod slot `rod`: 0 (an empty slot) or 2 ... its capacity (the
 * n it was created with).  Per-rod transfers and the fused path use the active
 * length; the op-by-op path needs every rod at full capacity
 * (EROD_ERR_UNSUPPORTED otherwise).  No counterpart in the reference, whose
 * blocks are rebuilt to change their rods (Block.hpp:568-611); for callers that
 * reuse device slots for new rods. */
int erod_set_rod_active_length(erod_batch* b, uint64_t rod, uint64_t n);
int erod_get_rod_active_length(const erod_batch* b, uint64_t rod, uint64_t* n);
/* Step budget of a rod on the fused path: a fused call of n steps advances the
 * rod by min(n, budget) steps and spends them; a rod with no budget left keeps
 * its state.  UINT64_MAX (the default) is unlimited.  No counterpart in the
 * reference (one clock for all systems); for callers that mix rods with
 * different horizons in one batch. */
int erod_set_rod_step_budget(erod_batch* b, uint64_t rod, uint64_t steps);
int erod_get_rod_step_budget(const erod_batch* b, uint64_t rod, uint64_t* steps);
/* Copy whole rods on the device: every field, the rod clock, hooks, damping
 * rates and dt of src_rods[i] go into slot dst_rods[i] of another batch on the
 * same device; the slot's active length becomes the source rod's and its budget
 * unlimited.  No counterpart in the reference; for admission of requests into
 * a pool of slots. */
int erod_copy_rods(erod_batch* dst, const uint64_t* dst_rods, erod_batch* src, const uint64_t* src_rods, uint64_t count);

/* ---- serving: continuous batching of requests ------------------------------- */

/* A server owns a pool batch of empty rod slots in size classes (n_classes
 * capacities and slot counts) and serves requests made of rods of `library`
 * (which must outlive it).  The pool's HBM (all fields of all slots) must fit
 * `hbm_budget_bytes`.  Every tick admits queued requests in order (all rods
 * of a request at once, each into the smallest free slot that holds it; the
 * earliest request that does not fit reserves the classes it needs, and later
 * ones are admitted past it into slots of other classes only), advances every
 * slot with budget left by up to `steps_per_tick` fused steps at `precision`,
 * waits for the device, and reports finished requests.  No counterpart in
 * the reference; its nearest analogue is the first-fit blocking of rods at
 * emplace time (BlockingPolicy.hpp:202-222). */
typedef struct erod_server erod_server;
int erod_server_create(erod_batch* library, const uint64_t* class_capacity, const uint64_t* class_slots, int n_classes,
                       int precision, int steps_per_tick, double hbm_budget_bytes, erod_server** out);
/* Same, admitting against the HBM budget: the pool starts with a few slots
 * of each class and, when the earliest waiting request finds no room, grows
 * (each class it lacks slots of doubles, or grows by what it lacks, or by
 * what it lacks alone when doubling would pass the budget) while the pool
 * stays within `hbm_budget_bytes`.  Growing rebuilds the pool batch: the rods
 * in it are copied bit for bit under the same slot ids, so results do not
 * change, and a pool batch taken from erod_server_pool before a tick may be
 * gone after it.  During a growth the old and the new pool are both held. */
int erod_server_create_budgeted(erod_batch* library, const uint64_t* class_capacity, int n_classes, int precision,
                                int steps_per_tick, double hbm_budget_bytes, erod_server** out);
int erod_server_destroy(erod_server* s);
/* The pool batch (for its tile configuration).  The pool steps in the
 * library's environment (gravity, contact), taken anew at every tick. */
int erod_server_pool(erod_server* s, erod_batch** pool);
/* Queue request `id`: the given library rods, advanced `steps` steps each with
 * their own dt.  Rejected if it could not fit the empty pool. */
int erod_server_submit(erod_server* s, uint64_t id, const uint64_t* library_rods, uint64_t n_rods, uint64_t steps);
/* One tick; up to `capacity` finished request ids go to done_ids (the rest are
 * reported by later ticks). */
int erod_server_tick(erod_server* s, uint64_t* done_ids, uint64_t capacity, uint64_t* n_done);
/* Field of rod `rod` (index within the request) of a finished request. */
int erod_server_read(erod_server* s, uint64_t id, uint64_t rod, int field, double* host);
/* Free a finished request's slots. */
int erod_server_release(erod_server* s, uint64_t id);
int erod_server_status(const erod_server* s, uint64_t* queued, uint64_t* running, uint64_t* active_elements,
                       uint64_t* pool_bytes);
/* Totals since creation: ticks that launched, and element-steps they advanced
 * (each rod counts min(steps per tick, its budget left) steps times its active
 * length), so that throughput over a window is a difference of two readings. */
int erod_server_stats(const erod_server* s, uint64_t* ticks, uint64_t* element_steps);
/* Slots in the pool, pool rebuilds so far (growths), and the HBM budget. */
int erod_server_pool_info(const erod_server* s, uint64_t* slots, uint64_t* growths, double* hbm_budget_bytes);

/* ---- checkpoints ------------------------------------------------------------ */

/* Write everything the batch holds (all fields, rod clocks, rod records with
 * dt, damping and hooks, active lengths, step budgets, environment) to one
 * binary file; erod_load creates a batch from it that continues bit for bit.
 * Replaces the block archive of elastica/Systems/Block/Serialize.hpp:37-62 and
 * elastica/Systems/CosseratRods/Serialize.hpp:41-72 (variables + ghost node
 * buffer).  The layout is documented in elasticapp_amd/csrc/erod_io.hip. */
int erod_save(erod_batch* b, const char* path);
int erod_load(const char* path, erod_batch** out);

/* ---- synchronisation and errors ------------------------------------------ */
int erod_synchronize(erod_batch* b);
/* Flags a non-finite value in x, v, Q or w of any rod (0/1 per rod). */
int erod_check_finite(erod_batch* b, uint8_t* rod_flags);
/* The same check without allocating device memory: no counterpart in the
 * reference; for callers that cannot allocate (code under HIP graph capture,
 * tight serving loops).  The per-rod flags live in a 1 MiB device array the
 * library holds on each GPU it runs on until the process ends, so a batch
 * may have at most 1,048,576 rods (EROD_ERR_INVALID beyond).  Calls that
 * overlap on one GPU share that array and race. */
int erod_check_finite_noalloc(erod_batch* b, uint8_t* rod_flags);
/* Message of the last failed call in this thread. */
const char* erod_last_error(void);

#ifdef __cplusplus
}
#endif

#endif /* EROD_H */
