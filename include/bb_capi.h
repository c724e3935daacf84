/*
 * bb_capi.h — C-ABI of the MI355X-native rigid-body step (beat-box drop-in).
 *
 * Plain C: POD structs, opaque handles, status-code returns, no exceptions
 * cross this boundary, no HIP/torch types in any signature (streams are passed
 * as `void*`, NULL = the library's own stream).
 *
 * Each entry point replaces one piece of the reference's C++ surface
 * (Jaisiero/beat-box).  The citation next to each declaration names the
 * reference function whose meaning, argument semantics and failure behaviour it
 * keeps.  The reference exposes no C ABI of its own; its callers are
 * SceneManager::load_scene and RendererManager::render.
 *
 * Thread safety: one handle per thread, or callers serialise.  There is no
 * global mutable state apart from per-device kernel modules.
 * Memory: the library owns device memory; the caller owns every host buffer.
 */
#ifndef BB_CAPI_H
#define BB_CAPI_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---------------------------------------------------------------- status */
typedef enum bb_status {
  BB_OK = 0,
  BB_ERR_INVALID = 1,   /* bad argument / malformed descriptor              */
  BB_ERR_CAPACITY = 2,  /* body / world capacity exceeded (add-body refusal)*/
  BB_ERR_DEVICE = 3,    /* HIP runtime error                                */
  BB_ERR_NO_DEVICE = 4, /* no GPU visible: there is no CPU fallback         */
  BB_ERR_OVERFLOW = 5,  /* device-side buffer overflow flagged during a step*/
  BB_ERR_STATE = 6,     /* call not valid in the current state              */
  BB_ERR_BUDGET = 7     /* serving: request does not fit the HBM budget     */
} bb_status;

/* --------------------------------------------------------- enums / flags */
/* SimSolverType, include/shared.inl:210-216.  PGS (plain sequential impulse)
 * exists in the reference only on the per-island path; with graph colouring on
 * (the default) "PGS" runs as PGS_SOFT, which is what this library does. */
typedef enum bb_solver {
  BB_SOLVER_PGS = 0,
  BB_SOLVER_PGS_SOFT = 1,
  BB_SOLVER_AVBD = 2
} bb_solver;

/* SimFlag, include/shared.inl:150-161 (same bit values). */
enum {
  BB_SIM_ACCUM_IMPULSE = 1u << 0,
  BB_SIM_FRICTION = 1u << 1,
  BB_SIM_WARM_STARTING = 1u << 2,
  BB_SIM_ADVECTION = 1u << 3,
  BB_SIM_DEBUG_INFO = 1u << 4,
  BB_SIM_USE_GRAPH_COLORING = 1u << 5,
  BB_SIM_DEBUG_GRAPH_COLORS = 1u << 6,
  BB_SIM_SLEEPING_ENABLED = 1u << 7
};
/* Default set, src/rigid_body_manager.hpp:140-147. */
#define BB_SIM_DEFAULT_FLAGS                                                   \
  (BB_SIM_ADVECTION | BB_SIM_DEBUG_INFO | BB_SIM_FRICTION |                    \
   BB_SIM_ACCUM_IMPULSE | BB_SIM_WARM_STARTING | BB_SIM_USE_GRAPH_COLORING |   \
   BB_SIM_SLEEPING_ENABLED)

/* RigidBodyFlag, include/shared.inl:91-97 (same bit values). */
enum {
  BB_BODY_NONE = 0,
  BB_BODY_COLLIDING = 1u << 0,
  BB_BODY_DYNAMIC = 1u << 1,
  BB_BODY_GRAVITY = 1u << 2,
  BB_BODY_SLEEPING = 1u << 3
};

typedef enum bb_shape { BB_SHAPE_BOX = 0, BB_SHAPE_SPHERE = 1, BB_SHAPE_VOXEL = 2 } bb_shape;

/* A voxel body's shape word: BB_SHAPE_VOXEL in the low byte, the 1-based
 * voxel shape id (RigidBody::shape_index, include/shared.inl:335-476) above. */
#define BB_SHAPE_VOXEL_OF(id) ((uint32_t)BB_SHAPE_VOXEL | ((uint32_t)(id) << 8))
#define BB_SHAPE_KIND(shape) ((uint32_t)(shape)&0xFFu)
#define BB_SHAPE_VOXEL_ID(shape) ((uint32_t)(shape) >> 8)

/* Voxel shape pools, include/shared.inl:772-774. */
#define BB_MAX_VOXEL_SHAPES 64u
#define BB_MAX_VOXEL_OCC_WORDS 16384u
#define BB_MAX_VOXEL_SURFACE 16384u

/* A concave voxel shape (VoxelShape, include/shared.inl:938-952): occupancy
 * bits over a dims grid, bit x + y*dims[0] + z*dims[0]*dims[1] of the words
 * (the solid(x,y,z) predicate of SceneManager::build_voxel_shape,
 * src/scene_manager.hpp:478-566). */
typedef struct bb_voxel_shape_desc {
  uint32_t dims[3];           /* cells per axis, 1..255                     */
  float voxel_size;           /* cell edge, > 0                             */
  float density;              /* mass per unit volume, > 0                  */
  const uint32_t* occupancy;  /* ceil(dims product / 32) words              */
} bb_voxel_shape_desc;

/* What build_voxel_shape returns (VoxelShapeBuild): the body origin is the
 * centre of mass; bodies of this shape take mass, inverse inertia and local
 * bounds from here (push_voxel_body, src/scene_manager.hpp:568-572). */
typedef struct bb_voxel_shape_info {
  uint32_t shape_id;      /* 1-based: body shape = BB_SHAPE_VOXEL_OF(id)     */
  uint32_t solid_count;   /* solid cells                                     */
  uint32_t surface_count; /* cells with an empty face neighbour              */
  float mass;
  float inv_inertia[9];   /* row-major, about the centre of mass             */
  float minimum[3];       /* grid bounds in the body frame                   */
  float maximum[3];
} bb_voxel_shape_info;

/* Reference-compatibility switches ("parity mode").  Each reproduces one
 * documented behaviour of the reference so results can be compared with it. */
enum {
  /* inverse inertia built from inv_mass where mass is expected
   * (src/scene_manager.hpp:861 -> include/math.hpp:106-130)              */
  BB_QUIRK_INERTIA_FROM_INV_MASS = 1u << 0,
  /* world inverse inertia as R^T I^-1 R (include/shared.inl:367-375,
   * src/shaders/extensions.slang:184-188)                                */
  BB_QUIRK_TRANSPOSED_WORLD_INERTIA = 1u << 1,
  /* advect/integrate dispatched over ceil(A/32)*32 threads but indexing rows
   * < N (src/rigid_body_manager.cpp advect/IP tasks use the ACTIVE dispatch) */
  BB_QUIRK_ACTIVE_DISPATCH_ROWS = 1u << 2
};
#define BB_QUIRKS_PARITY                                                       \
  (BB_QUIRK_INERTIA_FROM_INV_MASS | BB_QUIRK_TRANSPOSED_WORLD_INERTIA |        \
   BB_QUIRK_ACTIVE_DISPATCH_ROWS)

/* ------------------------------------------------------------ descriptors */
/* World descriptor.  Defaults (bb_world_desc_default) follow the reference:
 * dt = TIME_STEP 0.01667f, gravity = -GRAVITY, iteration_count = 5,
 * solver = AVBD, flags = default set (include/defines.hpp:354-364,
 * src/rigid_body_manager.hpp:140-147, 258-261). */
typedef struct bb_world_desc {
  uint32_t capacity_bodies; /* 0 = grow on demand                          */
  uint32_t solver;          /* bb_solver                                    */
  uint32_t iterations;      /* PGS iteration_count (create() argument)      */
  uint32_t flags;           /* BB_SIM_* bitmask                             */
  float dt;
  float gravity; /* signed y acceleration, reference uses -9.81             */
  /* 30 = reference keys: 30-bit Morton over the fixed box [-64,64]^3
   * (src/shaders/simulation/RB_sim.slang:99-101); 63 = 21 bits/axis over
   * the AABB of this world's body centres (extension, no body cap). */
  uint32_t morton_bits;
  uint32_t quirks; /* BB_QUIRK_* bitmask (BB_QUIRKS_PARITY for parity)     */
} bb_world_desc;

/* One body to add.  Boxes use local bounds minimum/maximum (reference
 * RigidBody::minimum/maximum); spheres use `radius` (extension: the reference
 * has no spheres).  mass == 0 means immovable, as in load_scene.  A voxel
 * body (shape = BB_SHAPE_VOXEL_OF(id)) takes mass, inverse inertia and bounds
 * from its shape and ignores minimum, maximum, radius and mass; without
 * BB_BODY_DYNAMIC it is immovable (mass 0). */
typedef struct bb_body_desc {
  uint32_t shape; /* bb_shape                                             */
  uint32_t flags; /* BB_BODY_* (DYNAMIC, GRAVITY)                         */
  float position[3];
  float rotation[4]; /* quaternion x, y, z, w                              */
  float minimum[3];
  float maximum[3];
  float radius;
  float mass;
  float velocity[3];
  float omega[3];
  float restitution; /* stored, ignored like the reference (solvers.slang:49-56) */
  float friction;
} bb_body_desc;

/* Readback of one body. */
typedef struct bb_body_state {
  uint32_t id;
  uint32_t flags;
  uint32_t island_index; /* MAX = not in an island                         */
  uint32_t row;          /* current Morton-sorted row                      */
  float position[3];
  float rotation[4];
  float velocity[3];
  float omega[3];
} bb_body_state;

typedef enum bb_order { BB_ORDER_BY_ID = 0, BB_ORDER_BY_ROW = 1 } bb_order;

/* SimConfig-equivalent per-step counters (include/shared.inl:571-666). */
typedef struct bb_step_stats {
  uint32_t rigid_body_count;            /* N                                */
  uint32_t active_rigid_body_count;     /* A                                */
  uint32_t broad_phase_collision_count; /* P                                */
  uint32_t collision_count;             /* M (manifolds)                    */
  uint32_t manifold_node_count;         /* 2M                               */
  uint32_t contact_count;               /* C                                */
  uint32_t island_count;
  uint32_t contact_island_count;
  uint32_t sleeping_count;
  uint32_t graph_color_count;
  uint32_t graph_color_violations;
  uint32_t graph_color_overflow;
  uint32_t avbd_color_count;
  uint32_t avbd_violations;
  uint32_t avbd_stick_count;
  uint32_t dbg_fresh;  /* contacts with warm_start == 0                   */
  uint32_t dbg_pen;    /* deepest penetration this step, integer mm       */
  uint32_t dbg_id_sum; /* sum of ids over rows (permutation checksum)     */
  uint32_t dbg_maxv;   /* max |v| over dynamic rows, integer mm/s         */
  uint32_t overflow;   /* nonzero: a device-side capacity was exceeded    */
  uint64_t step_index;
} bb_step_stats;

/* Debug taps: canonical intermediate outputs for the oracle comparison. */
typedef enum bb_tap {
  BB_TAP_MORTON = 0,      /* u32 (30-bit) or u64 (63-bit) per pre-sort row */
  BB_TAP_SORTED = 1,      /* {u64 code, u32 index, u32 pad} in sorted order */
  BB_TAP_LBVH = 2,        /* bb_lbvh_node[2N-1]                             */
  BB_TAP_PAIRS = 3,       /* {u32 a, u32 b} broad-phase pairs, (a,b) order  */
  BB_TAP_MANIFOLDS = 4,   /* bb_manifold_tap[M] after the solve             */
  BB_TAP_ISLANDS = 5,     /* u32 per row: island root (min active index)    */
  BB_TAP_EDGE_COLORS = 6, /* u32 per manifold                               */
  BB_TAP_BODY_COLORS = 7, /* u32 per row (AVBD)                             */
  BB_TAP_ROWS = 8         /* bb_body_state per row                          */
} bb_tap;

typedef struct bb_lbvh_node { /* LBVHNode, include/shared.inl:867-897     */
  float minimum[3];
  float maximum[3];
  int32_t left;
  int32_t right;
  uint32_t primitive_idx;
  uint32_t pad;
} bb_lbvh_node;

typedef struct bb_contact_tap {
  float position[3];
  float penetration;
  float normal_impulse;
  float tangent_impulse[2];
  float mass_coefficient;
  float tangent_mass[2];
  float anchor_a[3];
  float anchor_b[3];
  uint32_t stick;
  uint32_t fp[4];
  uint32_t warm_start;
} bb_contact_tap;

typedef struct bb_manifold_tap {
  uint32_t obb1_index;
  uint32_t obb2_index;
  int32_t key;
  int32_t error;
  int32_t contact_count;
  uint32_t color;
  float normal[3];
  float tangent[2][3];
  float anchor_a[3];
  float anchor_b[3];
  bb_contact_tap contacts[8];
} bb_manifold_tap;

/* ------------------------------------------------------- single world API */
typedef struct bb_world bb_world;

/* Fills *desc with the reference defaults. */
void bb_world_desc_default(bb_world_desc* desc);

/* RigidBodyManager::create(name, renderer, gui, iterations = 5),
 * src/rigid_body_manager.hpp:17 / src/rigid_body_manager.cpp:121-1488. */
bb_status bb_world_create(const bb_world_desc* desc, bb_world** out);

/* SceneManager::load_scene body rules + AccelerationStructureManager::
 * build_accel_structs/build_AS (src/scene_manager.hpp:850-887,
 * src/acceleration_structure_manager.cpp:224-441).  Ids are assigned
 * sequentially from 0; inverse mass/inertia follow the load_scene rules.
 * Fails with BB_ERR_CAPACITY (nothing added) when the world is full, as
 * build_accel_structs returns false when N > MAX_RIGID_BODY_COUNT. */
bb_status bb_world_add_bodies(bb_world* w, const bb_body_desc* bodies,
                              uint32_t n, uint32_t* out_ids);

/* SceneManager::build_voxel_shape (src/scene_manager.hpp:478-566): appends a
 * voxel shape to the world's pools (occupancy bits, surface-cell list, shape
 * record) and returns its id and mass properties.  BB_ERR_INVALID for bad
 * dims, sizes or an empty grid; BB_ERR_CAPACITY past the pool limits (the
 * reference prints an error and overruns). */
bb_status bb_world_add_voxel_shape(bb_world* w, const bb_voxel_shape_desc* desc,
                                   bb_voxel_shape_info* out);

/* RigidBodyManager::set_sim_flags (ORs the bits in), clear_sim_flags, and
 * get_sim_flags (src/rigid_body_manager.hpp:35-53).  Like the reference, a
 * flag change takes effect on the next step and keeps warm-start state. */
bb_status bb_world_set_flags(bb_world* w, uint32_t flags);
bb_status bb_world_clear_flags(bb_world* w, uint32_t flags);
uint32_t bb_world_get_flags(const bb_world* w);
/* RigidBodyManager::set_sim_type (src/rigid_body_manager.hpp:55-60).  As in
 * the reference (set_sim_type marks the sim dirty and update_sim() rewrites
 * SimConfig, zeroing the previous manifold count), the next step finds no old
 * manifolds to warm start from. */
bb_status bb_world_set_solver(bb_world* w, uint32_t solver);
/* create(..., iterations) argument (src/rigid_body_manager.hpp:17). */
bb_status bb_world_set_iterations(bb_world* w, uint32_t iterations);
/* RigidBodyManager::skip_warm_starting_once (src/rigid_body_manager.hpp:29-31):
 * the next step runs with WARM_STARTING cleared. */
bb_status bb_world_skip_warm_start_once(bb_world* w);

/* RigidBodyManager::simulate() (src/rigid_body_manager.cpp:1606-1622), nsteps
 * times, on the library stream (`stream` is reserved).  The reference's fixed
 * pair capacity becomes a growing one: after each step the call reads the
 * step's pair-overflow bit (one 4-byte read and a stream synchronisation per
 * step), and a step that overflowed is run again, from the state it started
 * from, with the world's pair capacity doubled (at most 16 times, else
 * BB_ERR_CAPACITY).  The result equals a step with enough capacity from the
 * start, bit for bit.  A world stepped by a workgroup team whose workgroups
 * did not all meet at a barrier (overflow bit 1) is stepped again from the
 * same state by one workgroup, with the same result, and stays on that path
 * (BB_ERR_DEVICE if that fails too).  Returns when the steps are done. */
bb_status bb_world_step(bb_world* w, uint32_t nsteps, void* stream);
/* device.wait_idle() after a step (src/renderer_manager.cpp:254-257). */
bb_status bb_world_sync(bb_world* w);


/* Body readback.  BY_ID goes through the id -> row map (rigid_body_entries,
 * src/shaders/simulation/RB_sim.slang:374-381); BY_ROW returns Morton rows.
 * The reference has no host body readback (its consumer is the TLAS update,
 * src/shaders/AS_mngr.slang:19-27); this is that consumer's read path. */
bb_status bb_world_read_bodies(bb_world* w, bb_body_state* out, uint32_t n,
                               bb_order order);
/* RigidBodyManager::read_back_sim_config + get_sim_config_reference
 * (src/rigid_body_manager.cpp:1681-1693, :79-82): counters of the last step. */
bb_status bb_world_read_stats(bb_world* w, bb_step_stats* out);
/* Canonical intermediate outputs of the last step (for parity tests). */
bb_status bb_world_debug_tap(bb_world* w, bb_tap kind, void* dst, size_t cap,
                             size_t* written);
/* Body count of the world. */
uint32_t bb_world_body_count(const bb_world* w);
/* Mechanical energy of the world's current state (see bb_batch_energy). */
bb_status bb_world_energy(bb_world* w, double* out);
/* Render hand-off: replaces entry_update_acceleration_structures
 * (src/shaders/AS_mngr.slang:5-36) with RigidBody::get_instance_transform
 * (include/shared.inl:377-419), which fill the TLAS instance buffer per
 * body id.  Writes, for body id i, 12 floats at dst + 12*i: the 3x4
 * row-major affine transform [R | t] of the current pose (the
 * VkTransformMatrixKHR layout), R by Quaternion::to_matrix
 * (include/math.hpp:292-309).  n = capacity of dst in bodies (at least the
 * world's body count, else BB_ERR_CAPACITY).  With dst_on_device != 0, dst
 * is device memory and the write is queued on the world's stream after the
 * steps (bb_world_sync waits for it); otherwise dst is host memory and the
 * call returns when it is filled. */
bb_status bb_world_instance_transforms(bb_world* w, float* dst, uint32_t n,
                                       int dst_on_device);
/* RigidBodyManager::get_lbvh_node_buffer (src/rigid_body_manager.cpp:
 * 1601-1604): the BVH the last step built, as the 2N-1 records of the
 * reference's LBVHNode layout (bb_lbvh_node, 40 bytes: min, max, left and
 * right child, primitive index = the leaf's row or 0xFFFFFFFF for an inner
 * node, pad), node 0 the root.  The reference's renderer builds its BLAS from
 * this buffer as procedural AABB geometry with a 40-byte stride
 * (AccelerationStructureManager::update, src/acceleration_structure_manager.
 * cpp:490-535).  *count receives 2N-1 (0 before the first step); with dst ==
 * NULL nothing else happens.  cap = capacity of dst in nodes
 * (BB_ERR_CAPACITY if short).  With dst_on_device != 0 the records are
 * written by a kernel queued on the world's stream (bb_world_sync waits for
 * it); otherwise dst is host memory, filled when the call returns. */
bb_status bb_world_lbvh_nodes(bb_world* w, bb_lbvh_node* dst, uint32_t cap,
                              uint32_t* count, int dst_on_device);
/* Checkpoint of a world for a later bb_world_restore, in this process or
 * another.  The reference keeps a paused simulation's buffers only in memory
 * (StatusManager::switch_simulating, src/status_manager.hpp:133-152, with the
 * current -> previous copies of AccelerationStructureManager::
 * record_update_AS_buffers_tasks, src/acceleration_structure_manager.cpp:
 * 813-948); this blob holds everything the world's next step reads: the
 * newest body rows in row order, the id -> row map, the pair list with its
 * per-row offsets, the manifold slots that warm start the next step, the
 * counters, the step index and the world and body descriptors.  With
 * dst == NULL, *size receives the bytes needed; otherwise cap must be at
 * least that (BB_ERR_CAPACITY).  The world's voxel shapes travel in the
 * blob and are rebuilt, with the same ids, on restore. */
bb_status bb_world_snapshot(bb_world* w, void* dst, size_t cap, size_t* size);
/* Creates a world from a checkpoint.  warm_start != 0: the world continues
 * bit-identically to the one that wrote the blob.  warm_start == 0: like the
 * reference's resume after a pause (skip_warm_starting_once,
 * src/status_manager.hpp:140-146), the next step starts without warm start.
 * BB_ERR_INVALID for a blob that is not a checkpoint of this version. */
bb_status bb_world_restore(const void* src, size_t size, int warm_start,
                           bb_world** out);
/* RigidBodyManager::destroy (src/rigid_body_manager.cpp:1550-1585). */
void bb_world_destroy(bb_world* w);

/* ------------------------------------------------ allocation-free stepping */
/* No counterpart in the reference.  For callers that cannot allocate device
 * memory (code next to HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc): what bb_world_create, bb_world_add_bodies (bodies[0..n)),
 * bb_world_step(nsteps), bb_world_read_bodies(out, out_n, BB_ORDER_BY_ID) and
 * bb_world_read_stats(stats, may be NULL) give for a fresh world, in one call
 * that allocates no device memory.  The world's device arrays are cut from
 * one static device array of BB_NOALLOC_SCRATCH_BYTES, which every GPU the
 * library runs on holds until the process ends; its pair capacity is the
 * largest of 16, 8, 4, 2 or 1 times bb_world_step's starting capacity that
 * fits there.  Limits: n from 1 to BB_NOALLOC_MAX_BODIES (BB_ERR_CAPACITY
 * beyond, with a message on stderr), boxes and spheres only (voxel bodies:
 * BB_ERR_INVALID).  Where bb_world_step would grow the pair capacity and
 * replay a step, this call keeps the pairs that fit and returns
 * BB_ERR_OVERFLOW, with out and stats filled.  Work is queued on `stream`
 * (a hipStream_t; NULL = the null stream) and the call returns when out is
 * filled.  The array keeps the last call's contents and each call clears the
 * part it uses.  All calls on one GPU share that GPU's array: calls that
 * overlapped on one GPU would race, so the library runs them one at a time
 * within a process (a lock per GPU). */
#define BB_NOALLOC_MAX_BODIES 2048u
#define BB_NOALLOC_SCRATCH_BYTES (64ull << 20)
bb_status bb_world_step_noalloc(const bb_world_desc* desc, const bb_body_desc* bodies, uint32_t n,
                                uint32_t nsteps, bb_body_state* out, uint32_t out_n, bb_step_stats* stats,
                                void* stream);

/* ------------------------------------------------------------- batch API */
/* Many independent worlds of uneven size stepped together: segmented sorts,
 * BVHs and colour sweeps, one workgroup per world.  No reference counterpart
 * (the reference steps one world). */
typedef struct bb_batch bb_batch;

bb_status bb_batch_create(uint32_t capacity_worlds, uint32_t capacity_bodies,
                          bb_batch** out);
/* Adds a world with its bodies; returns its index in *out_world. */
bb_status bb_batch_add_world(bb_batch* b, const bb_world_desc* desc,
                             const bb_body_desc* bodies, uint32_t n,
                             uint32_t* out_world);
/* bb_world_add_voxel_shape for a batch: the pools are shared by the batch's
 * worlds, so a voxel body of any world may use any shape added before it. */
bb_status bb_batch_add_voxel_shape(bb_batch* b, const bb_voxel_shape_desc* desc,
                                   bb_voxel_shape_info* out);
/* Steps every world nsteps times.  With use_graph != 0 the step sequence is
 * captured once as a hipGraph and replayed.  Asynchronous: a world whose
 * pairs overflow its capacity keeps the pairs that fit (the reference's early
 * return on a full buffer, src/shaders/simulation/LBVH.slang:248) and flags
 * the step: bit 0 of its `overflow` counter for that step, and a sticky mark
 * that stays until bb_batch_overflow collects it, whichever later steps do.
 * Bit 1 means a workgroup team did not meet at a barrier and left the step
 * unfinished.  The replay-and-grow of bb_world_step is for single worlds. */
bb_status bb_batch_step(bb_batch* b, uint32_t nsteps, void* stream,
                        int use_graph);
/* Waits for the batch's steps.  BB_ERR_OVERFLOW while any world carries an
 * uncollected overflow mark (the results are there, but such a world dropped
 * pairs or left a step unfinished). */
bb_status bb_batch_sync(bb_batch* b);
/* Collects the overflow marks: the indices of the worlds whose steps flagged
 * an overflow since the last call (at most cap of them written to worlds,
 * *count receives how many there are), then clears every mark. */
bb_status bb_batch_overflow(bb_batch* b, uint32_t* worlds, uint32_t cap,
                            uint32_t* count);
bb_status bb_batch_read_world(bb_batch* b, uint32_t world, bb_body_state* out,
                              uint32_t n, bb_order order);
bb_status bb_batch_world_stats(bb_batch* b, uint32_t world,
                               bb_step_stats* out);
bb_status bb_batch_debug_tap(bb_batch* b, uint32_t world, bb_tap kind,
                             void* dst, size_t cap, size_t* written);
uint32_t bb_batch_world_count(const bb_batch* b);
uint32_t bb_batch_body_count(const bb_batch* b);
/* Stage kernels of one step, in launch order (bb_batch_profile indices). */
enum {
  BB_KERNEL_BUILD = 0,  /* Morton keys, radix sort, LBVH, reorder          */
  BB_KERNEL_BROAD = 1,  /* BVH traversal, pair lists                       */
  BB_KERNEL_NARROW = 2, /* contacts, warm start, union-find                */
  BB_KERNEL_POST = 3,   /* advect, islands, sleeping, edge colouring       */
  BB_KERNEL_PGS = 4,    /* PGS_SOFT colour sweeps + integration            */
  BB_KERNEL_AVBD = 5,   /* AVBD colouring and sweeps                       */
  BB_KERNEL_FINISH = 6, /* end-of-step counters                            */
  BB_KERNEL_COUNT = 7
};
/* Batch totals of one step's counters (bb_batch_profile `counts` fields). */
enum {
  BB_COUNT_BODIES = 0,     /* sum N                                        */
  BB_COUNT_ACTIVE = 1,     /* sum A                                        */
  BB_COUNT_PAIRS = 2,      /* sum P                                        */
  BB_COUNT_MANIFOLDS = 3,  /* sum M                                        */
  BB_COUNT_CONTACTS = 4,   /* sum C                                        */
  BB_COUNT_COLOR_WORK = 5, /* sum over worlds of edge colours used x M     */
  BB_COUNT_MAX_PEN_MM = 6, /* max penetration over worlds (dbg_pen, mm)    */
  BB_COUNT_SLEEPING = 7,   /* sum sleeping bodies                          */
  BB_COUNT_FIELDS = 8
};
/* Measurement helper (no reference counterpart; the reference prints the
 * wait-bracketed step time, src/renderer_manager.cpp:254-260).  Runs nsteps
 * eager steps with hipEvents recorded on the library stream around every
 * stage kernel.  step_ms[nsteps] receives each step's duration,
 * kernel_ms[BB_KERNEL_COUNT] the summed duration of each stage over the steps
 * (a world stepped by a workgroup team runs all stages in one kernel; its
 * time is split by stage with the team's own clock marks, from the slowest
 * team),
 * counts[nsteps * BB_COUNT_FIELDS] each step's batch totals.  Any output may
 * be NULL. */
bb_status bb_batch_profile(bb_batch* b, uint32_t nsteps, float* step_ms, float* kernel_ms,
                           uint64_t* counts);
/* Parts of BB_KERNEL_BUILD (bb_batch_build_split indices): the Morton keys
 * (RB_sim.slang:75-111), the LSD radix sort (LBVH.slang:34-152), the Karras
 * hierarchy (RB_sim.slang:206-278), the bounds pass up to its last team-wide
 * barrier (RB_sim.slang:282-349), the lead workgroup's climb of the top
 * nodes above it, then the reorder (RB_sim.slang:351-386). */
enum {
  BB_BUILD_KEYS = 0,
  BB_BUILD_SORT = 1,
  BB_BUILD_HIERARCHY = 2,
  BB_BUILD_BOUNDS = 3,
  BB_BUILD_CLIMB = 4,   /* the lead workgroup's climb of the top nodes */
  BB_BUILD_REORDER = 5,
  BB_BUILD_COUNT = 6
};
/* Measurement helper (no reference counterpart): the build stage of the
 * slowest workgroup team, split into the parts above by the team's own clock
 * marks, summed over the steps of the last bb_batch_profile call
 * (build_ms[BB_BUILD_COUNT]; zeros when no world of the batch runs as a
 * team). */
bb_status bb_batch_build_split(const bb_batch* b, float* build_ms);
/* The AVBD stage of the slowest workgroup team, the same way (no reference
 * counterpart): adjacency, colouring and prepare (avbd.slang:70-128,
 * 812-981); warm start, support depth and the grouping by (colour, layer)
 * (avbd.slang:389-473, 667-724); the main sweeps with their dual updates
 * (avbd.slang:475-658, 726-808); the finalize and the post-stabilisation
 * (avbd.slang:130-202).  avbd_ms[BB_AVBD_COUNT]. */
enum { BB_AVBD_COLOUR = 0, BB_AVBD_GROUP = 1, BB_AVBD_SWEEPS = 2, BB_AVBD_POST = 3, BB_AVBD_COUNT = 4 };
bb_status bb_batch_avbd_split(const bb_batch* b, float* avbd_ms);
/* Mechanical energy of each world's current state, in double (no reference
 * counterpart): sum over dynamic bodies of m|v|^2/2 + w.I.w/2 - m*g*y with
 * the physical inertia of the shape.  per_world[n], n >= world count. */
bb_status bb_batch_energy(bb_batch* b, double* per_world, uint32_t n);
/* bb_world_lbvh_nodes for one world of a batch. */
bb_status bb_batch_lbvh_nodes(bb_batch* b, uint32_t world, bb_lbvh_node* dst,
                              uint32_t cap, uint32_t* count, int dst_on_device);
/* bb_world_instance_transforms for one world of a batch. */
bb_status bb_batch_instance_transforms(bb_batch* b, uint32_t world, float* dst,
                                       uint32_t n, int dst_on_device);
/* Device bytes held by the batch. */
uint64_t bb_batch_device_bytes(const bb_batch* b);

/* Block size used by the per-world kernels (autotuner hook); 0 = automatic. */
bb_status bb_batch_set_block_size(bb_batch* b, uint32_t threads);

/* Launch parameters (no reference counterpart).  None of them changes any
 * result: every world's trajectory is bit-identical for every setting. */
typedef enum bb_param {
  BB_PARAM_BLOCK = 0,        /* threads per world workgroup: 0 = automatic,
                                128, 256, 512 or 1024                       */
  BB_PARAM_PGS_LDS_ROWS = 1, /* rows per world whose velocities and world
                                inverse inertia the PGS sweeps keep in LDS
                                (80 B each): 0..2048                        */
  BB_PARAM_TEAM_BODIES = 2,  /* worlds with at least this many bodies are
                                stepped by a team of workgroups (k_team);
                                0 = every world by one workgroup            */
  BB_PARAM_LANE_BODIES = 3,  /* worlds with at least this many bodies (and
                                below the team threshold) run the per-world
                                kernels in a second lane, concurrently with
                                the rest; 0 = one lane.  Default 384        */
  BB_PARAM_LANE_BLOCK = 4,   /* threads per workgroup in that lane: 256, 512
                                or 1024.  Default 512                       */
  BB_PARAM_LANE_LDS_ROWS = 5, /* BB_PARAM_PGS_LDS_ROWS of that lane: 0..2048.
                                Default 1,024                               */
  BB_PARAM_TEAM_SPIN = 6,    /* spins of a team barrier wait before the team
                                gives the step up (overflow bit 1); 0 = the
                                default, about a second.  Small values force
                                that failure path (tests): bb_world_step then
                                replays the step by one workgroup, a batch
                                flags the world                             */
  BB_PARAM_PAIR_SCALE = 7,   /* pair slots of every world, as a multiple
                                (1..16, default 1) of the starting capacity
                                8n + 64 (six times that with voxel bodies):
                                room a batch world cannot grow into during
                                bb_batch_step.  Set it before the first step:
                                a later change lays the batch out anew, like
                                adding a world, and the next step starts
                                without warm start                          */
  BB_PARAM_SORT_BLOCKS = 8,  /* workgroups of a workgroup team that run its
                                radix sort (the rest wait at the barrier
                                after it): 1..2048, 0 = the default 32.
                                Fewer members mean cheaper sort barriers and
                                longer chunks each; the sort is stable, so
                                the order is the same for every value      */
  BB_PARAM_FUSED = 9,        /* 1: the per-world stages of a step run as one
                                launch per lane (k_fused: a world's
                                workgroup goes from stage to stage without
                                waiting for the lane's slowest world); 0
                                (default): one launch per stage.  Profiling
                                (bb_batch_profile) always times the stages  */
  BB_PARAM_AVBD_LANES = 10,  /* the AVBD solver tile of the per-world
                                kernels: lanes that assemble one body's 6x6
                                system together in the primal sweeps (1, 2,
                                4 or 8; 0 = the default 4).  The partial
                                sums are added in manifold order whatever
                                the value, so results do not change        */
  BB_PARAM_PGS_PREFETCH = 11, /* the PGS colour-sweep tile: 1 = each thread
                                fetches its row of the next colour while it
                                solves this colour's row (the next colour's
                                rows are written by nobody else meanwhile);
                                0 (default) = fetch where solved            */
  BB_PARAM_WALK_BUDGET = 12, /* broad phase: tree nodes one thread walks for
                                its body before the body is handed to its
                                whole workgroup (up to 64 bodies per world;
                                a ground under a pile overlaps thousands);
                                0 = the default 128, 0xFFFFFFFF = never.
                                The pair list is the same for every value  */
  BB_PARAM_TEAM_BLOCKS = 13, /* workgroups of the k_team launch at most
                                (1..2048); never more than the device holds
                                resident at once.  0 = half of that (the
                                default).  More workgroups cost more per
                                team barrier; results do not change        */
  BB_PARAM_PGS_BLOCK = 14    /* threads per workgroup of the colour sweeps
                                (k_pgs) of the worlds below the second
                                lane's threshold: 64, 128, 256, 512 or
                                1024; 0 (default) = BB_PARAM_BLOCK.  A small
                                world has a few dozen rows per colour: one
                                wave per world makes each colour's barrier
                                a wave's and lets more worlds share a CU.
                                Its LDS rows are BB_PARAM_PGS_LDS_ROWS     */
} bb_param;
bb_status bb_batch_set_param(bb_batch* b, bb_param p, uint32_t value);
/* Effective value (BB_PARAM_BLOCK resolves 0 to the automatic choice). */
uint32_t bb_batch_get_param(const bb_batch* b, bb_param p);
/* bb_batch_set_param for a single world (no reference counterpart). */
bb_status bb_world_set_param(bb_world* w, bb_param p, uint32_t value);
/* Autotuner: times `reps` replays of the next step for every candidate,
 * each replay restarting from the same state, and keeps the fastest: first
 * the grid of workgroup size x PGS LDS rows, then, one at a time from the
 * best so far, one launch per step or per stage (BB_PARAM_FUSED), the PGS
 * sweep tile (BB_PARAM_PGS_PREFETCH, when a world runs PGS), the AVBD
 * solver tile (BB_PARAM_AVBD_LANES, when a world runs AVBD in the per-world
 * kernels), the radix sort's sub-team (BB_PARAM_SORT_BLOCKS, when a world
 * runs as a workgroup team), the colour sweeps' workgroup and LDS rows
 * (BB_PARAM_PGS_BLOCK) and the team launch's size (BB_PARAM_TEAM_BLOCKS:
 * half or all of the resident workgroups).  The broad phase's walk budget
 * keeps its value: an early step cannot show how long the walks grow once
 * piles settle.  No candidate changes any result.  The
 * simulation state afterwards is exactly the state before the call.
 * *best_ms (may be NULL) receives the chosen step time. */
bb_status bb_batch_autotune(bb_batch* b, uint32_t reps, float* best_ms);
/* Checkpoint of a whole batch for a later bb_batch_restore: its voxel shapes
 * and pair scale, then per world what bb_world_snapshot keeps (descriptors,
 * newest rows, pairs, warm-start manifolds, counters) and its pair scale.
 * The reference pauses one world in memory (src/status_manager.hpp:133-152,
 * src/acceleration_structure_manager.cpp:813-948); no reference counterpart
 * for a batch.  Launch parameters are not part of the blob.  With
 * dst == NULL, *size receives the bytes needed; otherwise cap must be at least
 * that (BB_ERR_CAPACITY).  Overflow marks not yet collected are not kept. */
bb_status bb_batch_snapshot(bb_batch* b, void* dst, size_t cap, size_t* size);
/* Creates a batch from a bb_batch_snapshot blob, its worlds in the same order.
 * warm_start != 0: every world continues bit-identically to the batch that
 * wrote the blob; warm_start == 0: each world's next step starts without warm
 * start (src/status_manager.hpp:140-146).  Every record is checked on the
 * host first; BB_ERR_INVALID for a blob that is not a batch checkpoint of
 * this version. */
bb_status bb_batch_restore(const void* src, size_t size, int warm_start,
                           bb_batch** out);
void bb_batch_destroy(bb_batch* b);

/* ----------------------------------------------------------- serving loop */
/* Continuous batching of independent worlds (no reference counterpart; the
 * reference steps one scene).  Requests are queued on submit and admitted at
 * step boundaries, first come first served within their pool, while the
 * resident worlds' device state fits the HBM budget (and, optionally, a
 * resident-body cap).  The worlds of a pool advance one step per pool step in
 * one launch sequence; pools (split_bodies) step independently.  A world that
 * has run its requested number of steps is retired: its final state is
 * copied back and its device memory and slot are reused.  Each world's
 * trajectory is bit-identical to stepping it alone (bb_world_*), wherever it
 * sits in the device arrays and whatever else is resident. */
typedef struct bb_server bb_server;
#define BB_SERVER_MAX_POOLS 4

typedef struct bb_server_desc {
  uint64_t hbm_budget_bytes;    /* device bytes for resident state (admission) */
  uint32_t max_worlds;          /* resident world slots                       */
  uint32_t max_bodies;          /* body rows; 0 = as many as the budget holds */
  uint32_t max_resident_bodies; /* 0 = no cap; else admit while the resident
                                   bodies stay within it (latency bound)      */
  uint32_t block;               /* threads per world workgroup of the first
                                   pool (the only one without a split):
                                   128, 256, 512 or 1024, 0 = 256; the other
                                   pools run 512                             */
  uint32_t team_bodies;         /* worlds this large run as workgroup teams;
                                   0 = 8192, 0xFFFFFFFF = never              */
  uint32_t split_bodies;        /* 0: one pool, every resident world steps in
                                   lockstep.  Else pools by body count, each
                                   stepping at its own cadence on its own
                                   stream: [0, s) with s = split_bodies
                                   (workgroups of `block` threads), [s, 4s)
                                   and [4s, team_bodies) when 4s is below
                                   team_bodies (else [s, team_bodies)),
                                   both with 512-thread workgroups, and
                                   [team_bodies, ...) (teams); at most
                                   BB_SERVER_MAX_POOLS pools, max_worlds
                                   slots each                                */
} bb_server_desc;

typedef enum bb_request_state {
  BB_REQ_QUEUED = 0,
  BB_REQ_RUNNING = 1,
  BB_REQ_DONE = 2,
  BB_REQ_UNKNOWN = 3,
  /* done, but a step overflowed even after the request was run again from
   * its first step 17 times (pair capacity doubled, or one workgroup in place
   * of a team): the bodies are the last run's and may have dropped pairs   */
  BB_REQ_FAILED = 4
} bb_request_state;

typedef struct bb_server_report {
  uint64_t submitted, admitted, retired, queued;
  uint64_t resident_worlds, resident_bodies;
  uint64_t bytes_in_use, bytes_budget, bytes_allocated;
  uint64_t server_steps, world_steps, body_steps;
  double device_ms;                  /* first step start .. last completed step */
  double step_ms_p50, step_ms_p99;   /* per world-step latency: completion of
                                        step s - completion of step s-1 (or the
                                        world's admission for s = 1)          */
  double queue_ms_p50, queue_ms_p99; /* submit -> admission, host clock       */
  uint64_t reruns;                   /* requests run again from their first
                                        step after an overflow             */
  /* per pool (pools = 1 without a split): the smallest body count it takes,
   * its world-steps and their latency percentiles, as step_ms_p50/p99 */
  uint32_t pools, pad;
  uint32_t pool_min_bodies[BB_SERVER_MAX_POOLS];
  uint64_t pool_world_steps[BB_SERVER_MAX_POOLS];
  double pool_step_ms_p50[BB_SERVER_MAX_POOLS];
  double pool_step_ms_p99[BB_SERVER_MAX_POOLS];
} bb_server_report;

bb_status bb_server_create(const bb_server_desc* desc, bb_server** out);
/* Queues a world (bodies follow the load_scene rules, as bb_world_add_bodies)
 * that runs `steps` steps.  BB_ERR_BUDGET when the world alone exceeds the
 * budget or the capacities; *ticket identifies the request. */
/* bb_world_add_voxel_shape for the server: requests submitted afterwards may
 * hold voxel bodies of this shape (BB_SHAPE_VOXEL_OF(out->shape_id)); such a
 * world takes six pair slots per default pair slot.  The server's steps in
 * flight complete first.  No reference counterpart (the reference builds its
 * shapes per scene, src/scene_manager.hpp:478-572). */
bb_status bb_server_add_voxel_shape(bb_server* s, const bb_voxel_shape_desc* desc, bb_voxel_shape_info* out);
bb_status bb_server_submit(bb_server* s, const bb_world_desc* desc, const bb_body_desc* bodies, uint32_t n,
                           uint32_t steps, uint64_t* ticket);
/* Runs nsteps server rounds.  In a round every pool with queued or resident
 * worlds and fewer than two steps in flight takes one step (admission, one
 * step of each of its resident worlds, retirement); a pool whose two steps
 * are still running is passed over, so each pool keeps its own cadence, and
 * when no pool can step the call waits until one can.  With one pool a
 * round is one lockstep step of every resident world.  Launches are
 * asynchronous. */
bb_status bb_server_step(bb_server* s, uint32_t nsteps);
/* Steps until every submitted request is done. */
bb_status bb_server_drain(bb_server* s);
/* State of a request; when done, its final bodies (by id, n >= body count)
 * and last-step counters are copied to out / stats (either may be NULL) and
 * the result is released.  *state receives a bb_request_state.  A request
 * whose world overflowed in any step (pairs past its capacity, or a team
 * that left a step unfinished) is run again from its first step with twice
 * the pair capacity, or by one workgroup, before it is reported: its result
 * equals the world stepped alone by bb_world_step, which grows the same way. */
bb_status bb_server_poll(bb_server* s, uint64_t ticket, uint32_t* state, bb_body_state* out, uint32_t n,
                         bb_step_stats* stats);
/* Checkpoint of the serving loop (no reference counterpart): waits for
 * the steps in flight, then writes every resident world (its request and
 * the state its next step reads: rows, pairs, warm-start manifolds,
 * counters, steps done), every queued request and every result not yet
 * polled.  dst = NULL: *size receives the blob size only.  The server
 * does not step between the two calls, so the size holds.  Returns
 * BB_ERR_CAPACITY when cap < *size. */
bb_status bb_server_checkpoint(bb_server* s, void* dst, size_t cap, size_t* size);
/* A new server (from desc, which may differ from the one that wrote the
 * blob) holding the blob's worlds: resident worlds are admitted first and
 * continue from their state, bit for bit as in the server that wrote the
 * checkpoint; queued requests follow in order; results can be polled; new
 * tickets continue after the blob's.  The blob is checked on the host
 * (sizes, descriptors, every index the step uses) before anything reaches
 * the device: BB_ERR_INVALID if it fails, BB_ERR_BUDGET if a world does not
 * fit desc's budget. */
bb_status bb_server_restore(const bb_server_desc* desc, const void* src, size_t size, bb_server** out);
bb_status bb_server_report_get(bb_server* s, bb_server_report* out);
void bb_server_destroy(bb_server* s);

/* Measurement helper (no reference counterpart): copies `bytes` between two
 * device buffers `reps` times with a grid-stride float4 copy kernel
 * (k_copy_calibrate) and returns the mean kernel time in *ms.  Used to
 * calibrate HBM byte counters and as the practical bandwidth ceiling. */
bb_status bb_copy_bandwidth(uint64_t bytes, uint32_t reps, float* ms);

/* Library version string. */
const char* bb_version(void);

#ifdef __cplusplus
} /* extern "C" */
#endif

#endif /* BB_CAPI_H */
