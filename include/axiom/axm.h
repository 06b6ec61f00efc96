/*
 * axm.h — C ABI of the MI355X rigid-body collision & contact pipeline.
 *
 * Plain C, plain pointers and sizes: no C++ or torch types cross this
 * boundary.  It is the drop-in for the stages the reference reserves but never
 * implements (src/collision and src/dynamics hold only .gitkeep; both modules
 * are commented out in src/CMakeLists.txt:23-27).  Each entry point below cites
 * the reference interface or convention it stands in for.
 *
 * Conventions
 *   - Status codes are the numeric values of axiom::core::ErrorCode
 *     (include/axiom/core/error_code.hpp:21-57); 0 = Success.
 *   - Y-up, gravity (0,-9.81,0) by default (include/axiom/gui/physics_panel.hpp:35).
 *   - Quaternions are x,y,z,w (include/axiom/math/quat.hpp:29-34).
 *   - Shapes: 0 sphere (r), 1 box (hx,hy,hz), 2 capsule (r, half-height along
 *     local +Y), following gui::ShapeType numbering
 *     (include/axiom/gui/body_inspector.hpp:20-27); static half-space planes are
 *     world colliders (debug::ShapeType::Plane, include/axiom/debug/physics_debug_draw.hpp:87-94).
 *   - Contacts: normal from A to B, depth > 0 penetrating, position = witness
 *     point on A, witness on B = position - normal*depth
 *     (src/debug/physics_debug_draw.cpp:111-115).
 *   - Host pointers are borrowed for the duration of a call.  Calls on one
 *     handle are not thread-safe; each handle owns its own HIP stream.
 *     Steps run asynchronously until axm_world_sync or an axm_*_get_* call.
 */
#ifndef AXIOM_AXM_H
#define AXIOM_AXM_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef int32_t axm_status;

/* axiom::core::ErrorCode values (include/axiom/core/error_code.hpp:21-57). */
enum {
    AXM_SUCCESS = 0,
    AXM_OUT_OF_MEMORY = 200,
    AXM_NULL_POINTER = 202,
    AXM_INVALID_SHAPE = 300,
    AXM_GJK_FAILED_TO_CONVERGE = 301,
    AXM_EPA_FAILED_TO_CONVERGE = 302,
    AXM_INVALID_RIGID_BODY = 400,
    AXM_CONSTRAINT_SOLVER_FAILURE = 401,
    AXM_BUFFER_ALLOCATION_FAILED = 502,
    AXM_GPU_INVALID_OPERATION = 503,
    AXM_GPU_TIMEOUT = 504,
    AXM_GPU_OPERATION_FAILED = 505,
    AXM_INVALID_PARAMETER = 600,
    AXM_OUT_OF_RANGE = 601
};

enum { AXM_SHAPE_SPHERE = 0, AXM_SHAPE_BOX = 1, AXM_SHAPE_CAPSULE = 2, AXM_SHAPE_PLANE = 3 };

#define AXM_PLANE_BIT 0x80000000u /* pair partner index flag: static plane k = AXM_PLANE_BIT | k */
#define AXM_MAX_PLANES 8

enum { AXM_WARM_NONE = 0, AXM_WARM_CACHE = 1, AXM_WARM_MLP = 2 };

/* Step flags. */
enum {
    AXM_STEP_TIMINGS = 1u << 0,  /* fill the per-stage millisecond fields (adds event syncs) */
    AXM_STEP_RESIDUAL = 1u << 1, /* compute the residual after the last sweep */
    AXM_STEP_SYNC = 1u << 2,     /* block until the step is complete */
    AXM_STEP_KEEP_STAGES = 1u << 3 /* keep the solver's input velocities for axm_world_get_pre_solve_velocities */
};

/* World description.  Stands in for gui::PhysicsWorldConfig
 * (include/axiom/gui/physics_panel.hpp:34-43) plus the solver constants the
 * reference leaves unspecified.  The leading fields up to `sleep_time`
 * have the same layout as the CPU oracle's parameter block. */
typedef struct axm_world_desc {
    float gravity[3];
    float friction;        /* Coulomb mu (pyramid) */
    float restitution;     /* e, applied when the approach speed exceeds rest_threshold */
    float rest_threshold;  /* m/s */
    float baumgarte;       /* position-error feedback factor */
    float slop;            /* m of allowed penetration */
    float max_bias;        /* m/s cap on the position-error bias */
    float margin;          /* m: AABB inflation and speculative contact distance */
    float lin_damp;        /* per-step multiplicative damping (body_inspector_example.cpp:53-55) */
    float ang_damp;
    float cell_size;       /* m: uniform-grid cell; 0 = automatic (max body diameter) */
    float origin[3];       /* grid origin; cells clamp to [0,1023] per axis */
    uint32_t num_planes;
    float planes[AXM_MAX_PLANES][4]; /* n.xyz, d: solid where dot(n,x) < d */
    uint32_t warm_start;   /* AXM_WARM_* */
    uint32_t record_residual; /* 1: residual after every sweep (diagnostic) */
    float softness;        /* relative compliance eps of every contact row: a row solves v + eps*k*lambda = target
                              (k: the row's inverse effective mass).  0 = rigid rows; > 0 makes each step's
                              contact problem strictly convex, so the sweeps converge to any residual */
    uint32_t allow_sleep;  /* PhysicsWorldConfig::allowSleep: an island whose bodies all rest (|v| <
                              sleep_linear, |w| < sleep_angular) for sleep_time seconds falls asleep and
                              acts as static until an awake body touches it (physics_panel.hpp:39-42) */
    float sleep_linear;    /* PhysicsWorldConfig::sleepLinearThreshold (m/s) */
    float sleep_angular;   /* PhysicsWorldConfig::sleepAngularThreshold (rad/s) */
    float sleep_time;      /* PhysicsWorldConfig::sleepTimeThreshold (s) */
    uint32_t island_solve; /* 1: every island of at most 4096 manifolds (bodies joined by contacts; static
                              bodies and planes join nothing) is solved on its own and, with residual_tol,
                              stops at its own residual; stats.iterations reports the most sweeps any island
                              ran.  Without a tolerance the results are identical either way.  Not applied
                              with record_residual or with 2^23 bodies or more */
    /* capacities (0 = sized from the body count) */
    uint32_t max_bodies;
    uint32_t max_pairs;
    uint32_t max_manifolds;
    uint32_t auto_grow;    /* 1: grow pair/manifold buffers on overflow instead of failing */
    float residual_tol;    /* > 0: stop sweeping once the residual <= tol (checked every residual_every sweeps,
                              inside the sweep launch; stats.iterations reports the sweeps run) */
    uint32_t residual_every;
    uint32_t max_iterations; /* cap when residual_tol > 0 */
    int32_t device;        /* HIP device ordinal */
    uint32_t solver_mode;  /* 0: one persistent sweep kernel ordered by per-body counters (when <= 64 colours);
                              1: one launch per colour per sweep */
    uint32_t strict;       /* 1: a step (or run) that completes with a GJK or EPA failure returns
                              GJKFailedToConverge / EPAFailedToConverge, and one whose sweeps stop at
                              max_iterations above residual_tol returns ConstraintSolverFailure; the
                              results are kept (the state has advanced).  0: such steps succeed and
                              report it in their statistics only */
} axm_world_desc;

/* Host-side body arrays (structure of arrays).  Field set follows
 * gui::RigidBodyData (include/axiom/gui/body_inspector.hpp:58-90): position,
 * rotation, linear / angular velocity, mass (as inverse mass, 0 = static),
 * diagonal inertia (as its inverse, body frame), shape type and extents. */
typedef struct axm_body_soa {
    float* pos;          /* 3n */
    float* quat;         /* 4n, x y z w */
    float* lin_vel;      /* 3n */
    float* ang_vel;      /* 3n */
    float* inv_mass;     /* n */
    float* inv_inertia;  /* 3n */
    uint32_t* shape_type;/* n */
    float* shape_params; /* 3n */
    uint32_t* world;     /* n or NULL; batched worlds: world index per body (bodies of one world contiguous) */
    /* Optional per-body extras (NULL: the world description's values), following
     * gui::RigidBodyData's material, damping, type and filter fields
     * (include/axiom/gui/body_inspector.hpp:29-42, 76-87): */
    float* material;     /* 4n or NULL: friction, restitution, linear damping, angular damping.  A contact uses
                            mu = sqrt(mu_A mu_B) (mu_A when equal) and e = max(e_A, e_B); a plane has the
                            world's friction and restitution */
    uint32_t* flags;     /* n or NULL: bit 0 kinematic (BodyType::Kinematic: inverse mass 0, moved by its own
                            velocity, pushes dynamic bodies and ignores static ones) */
    uint32_t* filter;    /* 3n or NULL: FilterInfo categoryBits, maskBits, groupIndex (as int32).  Two bodies
                            collide iff they share a nonzero group that is positive, or, otherwise, each one's
                            category meets the other's mask; planes collide with every dynamic body */
} axm_body_soa;

/* Per-step statistics.  Mirrors gui::PhysicsWorldStats
 * (include/axiom/gui/physics_panel.hpp:11-31) and adds the solver's own counts. */
typedef struct axm_step_stats {
    uint32_t pairs;
    uint32_t manifolds;
    uint32_t contacts;      /* contact points (PhysicsWorldStats::contactPointCount) */
    uint32_t colours;
    uint32_t colour_rounds;
    uint32_t iterations;    /* sweeps actually run */
    float residual;         /* velocity-space fixed-point residual after the last sweep */
    float residual_first;   /* after the first sweep (when recorded) */
    uint32_t gjk_fail;
    uint32_t epa_fail;
    uint32_t warm_hits;
    uint32_t overflow;      /* bit 0 pairs, bit 1 manifolds */
    float broadphase_ms;    /* PhysicsWorldStats::broadphaseTime */
    float narrowphase_ms;   /* PhysicsWorldStats::narrowphaseTime */
    float solver_ms;        /* PhysicsWorldStats::solverTime */
    float integration_ms;   /* PhysicsWorldStats::integrationTime */
    float total_ms;         /* PhysicsWorldStats::totalStepTime */
    uint32_t islands;       /* PhysicsWorldStats::islandCount: islands of dynamic bodies, awake and asleep */
    uint32_t sleeping;      /* PhysicsWorldStats::sleepingBodies */
} axm_step_stats;

/* Kernel-level timing of the solver sweep kernel, accumulated over steps run
 * with profiling enabled (HIP events around every launch on the world's stream). */
typedef struct axm_kernel_profile {
    uint64_t steps;
    uint64_t pgs_launches;
    double pgs_ms;        /* summed kernel durations */
    double pgs_bytes;     /* summed algorithmic bytes of those launches */
    uint64_t sweeps;
    /* the bundle kernel of the islands solved on their own (island_solve) */
    uint64_t bundle_launches;
    double bundle_ms;     /* summed kernel durations */
    uint64_t bundle_point_sweeps; /* contact points x sweeps of its islands (240 B each) */
} axm_kernel_profile;

typedef struct axm_world_s* axm_world;

/* Number of visible HIP devices (0 when none or when the runtime is unavailable). */
int32_t axm_device_count(void);

/* Mirrors errorCodeToString (src/core/error_code.cpp:5-65). */
const char* axm_status_string(axm_status s);

/* Defaults for a world description: PhysicsWorldConfig defaults
 * (include/axiom/gui/physics_panel.hpp:35-42) plus solver constants. */
void axm_world_desc_default(axm_world_desc* desc);

/* Factory, as VkContext::create (include/axiom/gpu/vk_instance.hpp:34): returns
 * BufferAllocationFailed / GPU_OPERATION_FAILED on failure. */
axm_status axm_world_create(const axm_world_desc* desc, axm_world* out);
axm_status axm_world_destroy(axm_world w);
/* Replaces the solver settings of a live world (gravity, friction, restitution and the other
 * contact constants, softness, sleep settings, damping, planes, warm start, residual tolerance;
 * as gui::PhysicsWorldConfig changes between frames, physics_panel.hpp:34-43).  Capacities, the
 * grid, auto_grow and the device keep their values; captured step graphs are rebuilt. */
axm_status axm_world_configure(axm_world w, const axm_world_desc* desc);
/* The world's current description. */
axm_status axm_world_get_desc(axm_world w, axm_world_desc* out);

/* Uploads n bodies (copies in).  Invalid bodies (NaN state, negative mass,
 * unknown shape) return InvalidRigidBody / InvalidShape and leave the world unchanged. */
axm_status axm_world_set_bodies(axm_world w, const axm_body_soa* bodies, uint32_t n);
/* Copies out pos, quat, lin_vel, ang_vel (other fields ignored; NULL fields skipped). */
axm_status axm_world_get_bodies(axm_world w, axm_body_soa* out);
axm_status axm_world_body_count(axm_world w, uint32_t* n);

/* One step of PhysicsWorld::step(dt) (examples/profiling_example.cpp:132-165):
 * broadphase -> narrowphase -> solver (iters sweeps) -> integration. */
axm_status axm_world_step(axm_world w, float dt, uint32_t iters, uint32_t flags, axm_step_stats* stats);
/* `steps` steps; with graph = 1 the step is captured once into a hipGraph and
 * replayed (not with record_residual; residual-stopped steps replay too: the
 * sweep launch checks the residual itself). */
axm_status axm_world_run(axm_world w, float dt, uint32_t iters, uint32_t steps, int graph, axm_step_stats* last);
axm_status axm_world_sync(axm_world w);
/* Islands and sleep state after the last step (the reference's SleepInfo and island
 * statistics: include/axiom/gui/body_inspector.hpp:45-51, physics_panel.hpp:23): per body
 * the island label (the smallest body index of its island; 0x7FFFFFFF for static bodies and
 * bodies that have not been in a step yet), whether it is asleep, and its rest time in
 * seconds (SleepInfo::sleepTime).  NULL outputs are skipped. */
axm_status axm_world_get_islands(axm_world w, uint32_t* label, uint32_t* asleep, float* rest_time);

/* Totals over the steps run since the last reset (counted on the device, so graph-replayed
 * steps need no host round trip): steps, steps whose sweeps stopped because the residual met
 * residual_tol, sweeps, contact points x sweeps (the per-step contact count times that step's
 * sweeps), and the most sweeps one step ran.  Synchronises the world's stream; reset = 1
 * clears the totals after reading them.  Benchmark hook, no reference counterpart. */
typedef struct axm_world_totals {
    uint64_t steps;
    uint64_t steps_at_tol;
    uint64_t sweeps;
    uint64_t contact_sweeps;
    uint32_t max_sweeps;
    uint32_t pad;
    uint64_t local_contact_sweeps;  /* the part of contact_sweeps in islands solved on their own */
} axm_world_totals;
axm_status axm_world_get_totals(axm_world w, axm_world_totals* out, int reset);

/* World snapshot: the description, every body's state (including sleep state, material,
 * flags and filters) and the warm-start cache of the last step, in a checksummed binary
 * file (magic "AXMWORLD", version 1).  axm_world_load replaces the world's description
 * (keeping its device), bodies and cache with the file's: stepping on from a loaded
 * snapshot gives bit-identical results to stepping on from where it was saved.  A file
 * that does not match returns InvalidParameter and leaves the world as it was.  Stand-in
 * for the reference's save/load of binary caches (PipelineCache::save/load,
 * include/axiom/gpu/vk_compute_pipeline.hpp:263-270). */
axm_status axm_world_save(axm_world w, const char* path);
axm_status axm_world_load(axm_world w, const char* path);

/* Enables / disables per-launch event timing of the sweep kernel; `out`
 * (optional) receives the accumulated profile, which is then reset. */
axm_status axm_world_profile(axm_world w, int enable, axm_kernel_profile* out);

/* Launch shape of the persistent sweep kernel: workgroups of one wave (0: one per CU;
 * at most what can be resident at once), KiB of LDS per workgroup for the staged
 * Jacobian rows and impulses of its slots (0: the device maximum, up to 160), and
 * policy bits (bit 1, default on: poll one 16-byte hint before loading the
 * velocities; bit 2: back off longer while a wave makes no progress; bits 4-6: poll
 * back-off level 0..7, default 1).  Results
 * do not depend on it; throughput does (tuning hook, no reference counterpart). */
axm_status axm_world_set_sweep_config(axm_world w, uint32_t groups, uint32_t lds_kib, uint32_t policy);

/* The sweep launch shape chosen by axm_world_autotune. */
typedef struct axm_sweep_config {
    uint32_t groups, lds_kib, policy;
    double sweep_ms; /* its average sweep-kernel time over the trial steps */
} axm_sweep_config;
/* Autotuner of the persistent sweep: times `steps` graph-replayed steps (after one
 * untimed capture step) for each candidate launch shape (one, two and four
 * workgroups per CU, LDS split between them), keeps the fastest and reports it.
 * The world advances by (steps + 1) x candidates steps; results do not depend on the
 * shape (no reference counterpart: the batched-world tuning hook). */
axm_status axm_world_autotune(axm_world w, float dt, uint32_t iters, uint32_t steps, axm_sweep_config* out);

/* Stage outputs of the last step (collision::ContactPoint / debug::DebugContactPoint,
 * include/axiom/debug/physics_debug_draw.hpp:128-132).  Undersized buffers return
 * OutOfRange with *count set to the required size. */
axm_status axm_world_get_pairs(axm_world w, uint32_t* a, uint32_t* b, uint32_t cap, uint32_t* count);
axm_status axm_world_get_manifolds(axm_world w, uint32_t* a, uint32_t* b, uint32_t* npts, float* pos /*4x3 per*/,
                                   float* normal /*4x3*/, float* depth /*4*/, uint32_t* feature /*4*/, uint32_t cap,
                                   uint32_t* count);
/* colour per manifold (pair order), solve order, colour count. */
axm_status axm_world_get_colours(axm_world w, uint32_t* colour, uint32_t* solve_order, uint32_t cap, uint32_t* count,
                                 uint32_t* num_colours);
/* impulses per manifold (pair order) x 4 points x (normal, t1, t2): initial (warm start) and final. */
axm_status axm_world_get_impulses(axm_world w, float* initial, float* final_, uint32_t cap, uint32_t* count);
/* friction bound mu * B of every contact point of the last step (manifold order x 4 points):
 * each friction row is clamped to |lambda_t| <= mu * B, with B fixed when the step starts: the
 * point's normal impulse of the previous step when the warm-start cache holds its feature,
 * else m_N (target - v_n)+, the impulse that stops its approach on its own. */
axm_status axm_world_get_friction_bounds(axm_world w, float* bound, uint32_t cap, uint32_t* count);
/* residual after each sweep of the last step (needs record_residual). */
axm_status axm_world_get_residuals(axm_world w, float* out, uint32_t cap, uint32_t* count);
/* velocities entering the solver of the last step (after gravity). */
axm_status axm_world_get_pre_solve_velocities(axm_world w, float* lin_vel, float* ang_vel);

/* Stand-alone broadphase on the current state (no time advance). */
axm_status axm_broadphase(axm_world w, uint32_t* count);
/* Stand-alone broadphase + narrowphase on the current state. */
axm_status axm_collide(axm_world w, uint32_t* manifold_count);

/* Batched pairwise narrowphase: n shape pairs in, <= 4 contacts per pair out.
 * type_b == AXM_SHAPE_PLANE: prm_b holds the plane normal and pos_b[0] its offset.
 * Stand-in for the narrowphase dispatch the reference names (CLAUDE.md:88). */
axm_status axm_collide_pairs(int32_t device, uint32_t n, const uint32_t* type_a, const float* prm_a, const float* pos_a,
                             const float* quat_a, const uint32_t* type_b, const float* prm_b, const float* pos_b,
                             const float* quat_b, float margin, uint32_t* npts, float* pos, float* normal, float* depth,
                             uint32_t* feature);
/* axm_collide_pairs without allocating: the same kernel and results, with the device
 * copies of the inputs, the outputs and the per-pair status staged in a fixed buffer the
 * library holds on each device until the process ends (3.5 MiB: at most 16384 pairs per
 * call; more returns OutOfRange).  It calls no hipMalloc / hipFree, for callers that
 * cannot allocate (tight stepping loops, harnesses that forbid device allocation).  Calls
 * that overlap on one device share the buffer and race: serialise them.  No counterpart in
 * the reference. */
axm_status axm_collide_pairs_noalloc(int32_t device, uint32_t n, const uint32_t* type_a, const float* prm_a,
                                     const float* pos_a, const float* quat_a, const uint32_t* type_b,
                                     const float* prm_b, const float* pos_b, const float* quat_b, float margin,
                                     uint32_t* npts, float* pos, float* normal, float* depth, uint32_t* feature);
/* Batched GJK (+ EPA when the cores overlap) on shape cores.  out: 10 floats per
 * pair (distance or depth, pA[3], pB[3], n[3]); status per pair: 0 separated,
 * 1 penetrating, 301 / 302 on failure (GJKFailedToConverge / EPAFailedToConverge,
 * include/axiom/core/error_code.hpp:37-38). */
/* The same with the iteration caps of GJK and EPA given (0: the defaults, 32 and 48; at most
 * 4096): a pair that needs more returns 301 / 302, as in the world step's narrowphase. */
axm_status axm_gjk_epa_pairs_capped(int32_t device, uint32_t n, const uint32_t* type_a, const float* prm_a,
                                    const float* pos_a, const float* quat_a, const uint32_t* type_b,
                                    const float* prm_b, const float* pos_b, const float* quat_b,
                                    uint32_t gjk_max_iters, uint32_t epa_max_iters, float* out, int32_t* status);
axm_status axm_gjk_epa_pairs(int32_t device, uint32_t n, const uint32_t* type_a, const float* prm_a, const float* pos_a,
                             const float* quat_a, const uint32_t* type_b, const float* prm_b, const float* pos_b,
                             const float* quat_b, float* out, int32_t* status);

/* ---------------------------------------------------------------- learned warm start
 * A 24 -> 128 -> 128 -> 3 ReLU network (bf16 weights / activations, fp32
 * accumulation, MFMA kernels) predicting, for each contact point, the change of
 * its impulses (normal, tangent 1, tangent 2) from the previous step's impulse
 * of the same feature, in units of u = mN * 9.81 m/s^2 * dt, from 24 per-point
 * inputs (normal, depth, contact-frame relative velocity and lever arms, mass
 * ratios, bias, the previous impulse).  With desc.warm_start == AXM_WARM_MLP and
 * a network attached, the solver starts from previous + prediction (projected
 * onto lambda_n >= 0 and the friction pyramid) and iterates as usual, so the
 * converged answer does not depend on it; a zero network is the cached warm start.
 * The reference has no learned component (no counterpart); these entry points
 * serve the warm-start path of axm_world_step. */
typedef struct axm_mlp_s* axm_mlp;
#define AXM_MLP_INPUTS 24
#define AXM_MLP_OUTPUTS 3
#define AXM_MLP_PARAMS 20099 /* W1[128][24], b1[128], W2[128][128], b2[128], W3[3][128], b3[3] */
axm_status axm_mlp_create(int32_t device, axm_mlp* out);
axm_status axm_mlp_destroy(axm_mlp m);
/* He-uniform weights from a seed, zero biases, Adam state cleared. */
axm_status axm_mlp_init(axm_mlp m, uint64_t seed);
axm_status axm_mlp_set_params(axm_mlp m, const float* params);
axm_status axm_mlp_get_params(axm_mlp m, float* params);
/* Y[n][3] = network(X[n][24]) (host arrays). */
axm_status axm_mlp_forward(axm_mlp m, const float* X, uint32_t n, float* Y);
/* Mean-squared-error loss over n rows and its gradient (host arrays; grad may be NULL). */
axm_status axm_mlp_loss_grad(axm_mlp m, const float* X, const float* Y, uint32_t n, float* grad, double* loss);
/* One bias-corrected Adam step (Kingma & Ba 2015) with the given gradient (NULL: the last
 * axm_mlp_loss_grad gradient). */
axm_status axm_mlp_adam_step(axm_mlp m, const float* grad, float lr, float beta1, float beta2, float eps);
/* Minibatch Adam (beta 0.9 / 0.999, eps 1e-8) over n rows in the given order, `epochs` passes;
 * epoch_loss (optional) receives each pass's mean batch loss. */
axm_status axm_mlp_train(axm_mlp m, const float* X, const float* Y, uint32_t n, uint32_t epochs, uint32_t batch,
                         float lr, float* epoch_loss);
/* Network file: parameters, Adam moments and step count in a checksummed binary file
 * (magic "AXMMLP01", version 1).  Loading restores the network and its optimiser
 * state exactly; a file that does not match returns InvalidParameter.  Stand-in for the
 * reference's binary cache files (PipelineCache::save/load,
 * include/axiom/gpu/vk_compute_pipeline.hpp:263-270). */
axm_status axm_mlp_save(axm_mlp m, const char* path);
axm_status axm_mlp_load(axm_mlp m, const char* path);
/* Attaches a network to a world (same device; NULL detaches).  The caller keeps it alive. */
axm_status axm_world_set_mlp(axm_world w, axm_mlp m);
/* Keeps every step's per-point network inputs for axm_world_get_features. */
axm_status axm_world_set_feature_logging(axm_world w, int enable);
/* Network inputs X[cap][4][24] and targets Y[cap][4][3] ((final - previous impulse) / u) of
 * the last step, manifold order (pair order) x point; rows past a manifold's point count are zero. */
axm_status axm_world_get_features(axm_world w, float* X, float* Y, uint32_t cap, uint32_t* count);

/* Synthetic scenes (host only, seeded DeterministicRNG, include/axiom/math/random.hpp:19-68).
 * config: 0 mixed 1k, 1 dense pile, 2 sparse 1M, 3 batched (n_worlds worlds).
 * n = body count (0 = the config's default).  Call with out->pos == NULL to query
 * the body count and the recommended world description. */
axm_status axm_scene_generate(uint32_t config, uint64_t seed, uint32_t n, uint32_t n_worlds, axm_body_soa* out,
                              uint32_t* count, axm_world_desc* desc);
/* Inverse mass and inverse body-frame principal inertia of a solid shape of the
 * given density (kg/m^3): sphere (r), box (half extents), capsule (r, half-height,
 * axis +Y).  Host only.  Returns AXM_INVALID_SHAPE for an unknown type or a
 * non-positive size.  Replaces the mass/inertia fields a caller fills in
 * gui::RigidBodyData (include/axiom/gui/body_inspector.hpp:58-90). */
axm_status axm_mass_properties(uint32_t type, const float* params, float density, float* inv_mass,
                               float* inv_inertia);
/* Raw DeterministicRNG stream (u32) and nextFloat(min,max) stream for a seed. */
void axm_rng_stream(uint64_t seed, uint32_t n, uint32_t* u32_out, float lo, float hi, float* f_out);

#ifdef __cplusplus
}
#endif

#endif /* AXIOM_AXM_H */
