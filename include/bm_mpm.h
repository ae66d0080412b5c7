/* bm_mpm.h -- C ABI of the MI355X MLS-MPM soil substep (libbm_mpm.so).
 *
 * One handle = one soil scene resident in HBM: particle state (FP32 SoA),
 * a dense block-linear node grid over a fixed domain, and the scratch the
 * kernels need.  Every entry point returns an int status (BM_OK = 0) and
 * leaves a message for bm_last_error() on failure; no C++ exception crosses
 * the boundary.  Work is queued on the handle's HIP stream; entry points that
 * hand data back to host memory synchronise that stream first.
 *
 * The entry points replace the reference's C++ API for this path (paths
 * relative to the reference root):
 *   basements::mpm::MPMSolver           include/basements/physics/mpm/mpm_solver.h:15-95
 *   MPMRigidCoupler::apply (M0)         include/basements/physics/coupling/mpm_rigid_coupler.h:212-214
 *   MPMRigidCouplerPIC::apply (M1)      include/basements/physics/coupling/mpm_rigid_coupler_pic.h:88-90
 *   TimeIntegrator::integrate           include/basements/physics/dynamics/integrator.h:19-57
 * The C style (POD structs, int status, version check, last-error string)
 * follows the reference's plugin ABI, include/basements/engine/plugins/plugin_capi.h.
 *
 * A handle is not thread-safe; distinct handles are independent.
 */
#ifndef BM_MPM_H
#define BM_MPM_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define BM_MPM_ABI_VERSION 1

enum bm_status {
    BM_OK = 0,
    BM_ERR_INVALID_ARG = 1,
    BM_ERR_OOM = 2,
    BM_ERR_HIP = 3,
    BM_ERR_NONFINITE = 4,
    BM_ERR_NOT_SUPPORTED = 5,
    BM_ERR_OUT_OF_DOMAIN = 6, /* a particle's 3x3x3 stencil left the grid domain */
    BM_ERR_STATE = 7          /* phase called out of order (e.g. g2p before p2g) */
};

/* Substep pipeline.
 *   BM_MODE_REFERENCE: the reference's phase order -- P2G (stress from an SVD
 *     of F), grid update, couplers, G2P with the plasticity SVD -- with the
 *     reference's operation order and FP32 Jacobi; force kept apart from
 *     momentum on the grid.  mpm_solver.h:71-87.
 *   BM_MODE_FAST: the same physics with G2P of substep n and the particle
 *     half of P2G of substep n+1 in one pass over the particles, then a
 *     per-cell deposit into the next grid; the Kirchhoff stress of n+1 comes
 *     from the plasticity SVD of n (F_new = U diag(sigma') V^T is its own
 *     SVD), the stress force is folded into the node momentum, particles are
 *     kept binned by grid block, and the SVD has an FP64 fallback. */
enum bm_mode { BM_MODE_REFERENCE = 0, BM_MODE_FAST = 1 };

/* Coupler method ids; the numbering is the reference's M0/M1
 * (research/mpm_rigid_coupling/bench_compare_all.cpp:60-63). */
/* M0 Dirichlet (mpm_rigid_coupler.cpp:58-143), M1 PIC-rigid
 * (mpm_rigid_coupler_pic.h:88-184), and the research harness's M4.5 ladder
 * (couplers/m4_5_hybrid.h:104-313): M4.5a = M1 with the body mass spread over
 * the inside nodes by cell-overlap fraction (two passes, sum = body mass),
 * M4.5b = a + Coulomb tangent blend, M4.5 = b + quadratic B-spline weight. */
enum bm_coupler {
    BM_COUPLER_M0_DIRICHLET = 0,
    BM_COUPLER_M1_PIC_RIGID = 1,
    BM_COUPLER_M45A_OVERLAP = 2,
    BM_COUPLER_M45B_FRICTION = 3,
    BM_COUPLER_M45_BSPLINE = 4
};

enum bm_shape { BM_SHAPE_BOX = 0, BM_SHAPE_CYLINDER = 1 };

typedef struct bm_mpm_solver* bm_mpm_handle;

typedef struct {
    float dx;                /* node spacing [m]; MPMSolver::initialize(dx) */
    int32_t grid_origin[3];  /* node index of the domain's lowest corner */
    int32_t grid_dims[3];    /* nodes per axis; rounded up to a multiple of 4 */
    int64_t max_particles;   /* capacity of the particle arrays */
    int32_t mode;            /* enum bm_mode */
    int32_t sort_every;      /* BM_MODE_FAST: re-bin particles every k substeps (0 = default 32) */
    void* stream;            /* hipStream_t to run on; NULL = the handle creates its own */
} bm_mpm_desc;

/* Host-side particle state, SoA, FP32.  Matrices are row-major 3x3 per
 * particle (F[9*i + 3*r + c]).  Field meaning: struct Particle,
 * include/basements/physics/mpm/particle.h:9-23 (volume = V0). */
typedef struct {
    float* x;  /* 3n */
    float* v;  /* 3n */
    float* F;  /* 9n */
    float* C;  /* 9n */
    float* m;  /* n  */
    float* V0; /* n  */
} bm_particles;

/* Device pointers of the handle's own particle arrays: component-major SoA,
 * component c of particle i at base[c * capacity + i]. */
typedef struct {
    float* x;  /* 3 arrays */
    float* v;  /* 3 arrays */
    float* F;  /* 9 arrays, row-major component order */
    float* C;  /* 9 arrays */
    float* m;
    float* V0;
    int64_t capacity;
} bm_particles_device;

/* RigidColliderState, mpm_rigid_coupler.h:72-84.  quat = (w, x, y, z);
 * CYLINDER: half_extents = (radius, half_length, radius), axis = local Y. */
typedef struct {
    float center[3];
    float quat[4];
    float half_extents[3];
    float lin_vel[3];
    float ang_vel[3];
    float friction;
    float mass;
    int32_t shape; /* enum bm_shape */
} bm_collider;

/* CouplingReaction, mpm_rigid_coupler.h:201-205. */
typedef struct {
    float force[3];
    float torque[3];
    int32_t nodes_inside;
} bm_reaction;

/* ---------------------------------------------------------------- library */
int bm_abi_version(void);
const char* bm_last_error(void);

/* ---------------------------------------------------------------- lifetime */
/* MPMSolver() + initialize(dx) (mpm_solver.h:17-21). */
int bm_mpm_create(const bm_mpm_desc* desc, bm_mpm_handle* out);
int bm_mpm_destroy(bm_mpm_handle h);
int bm_mpm_synchronize(bm_mpm_handle h);

/* ---------------------------------------------------------------- parameters */
/* set_material (mpm_solver.h:42-51): rescales V0 of existing particles by
 * rho_old/rho_new when the density changes. */
int bm_mpm_set_material(bm_mpm_handle h, float young_modulus, float poisson_ratio, float bulk_density);
int bm_mpm_set_gravity(bm_mpm_handle h, float gx, float gy, float gz);   /* mpm_solver.h:52 */
int bm_mpm_set_plastic_alpha(bm_mpm_handle h, float alpha);               /* mpm_solver.h:53 */
int bm_mpm_set_floor_friction(bm_mpm_handle h, float mu_floor);           /* mpm_solver.h:54 */
/* Optional slip walls on the six domain faces, `thickness` nodes deep: the
 * outward normal velocity of nodes in that layer is zeroed.  Off (0) by
 * default; the reference has only the floor (mpm_solver.h:263-288). */
int bm_mpm_set_walls(bm_mpm_handle h, int32_t thickness);

/* ---------------------------------------------------------------- particles */
/* add_particle for n particles (mpm_solver.h:23-34): F = I, C = 0,
 * V0 = m / rho (1 when rho <= 1e-12).  pos, vel: 3n floats; mass: n floats. */
int bm_mpm_add_particles(bm_mpm_handle h, int64_t n, const float* pos, const float* vel, const float* mass);
/* Replace the whole particle state (get_particles_mutable(), mpm_solver.h:56-59). */
int bm_mpm_upload_particles(bm_mpm_handle h, int64_t n, const bm_particles* host);
/* Copy the state back in the caller's (insertion) order (get_particles(), mpm_solver.h:89). */
int bm_mpm_download_particles(bm_mpm_handle h, bm_particles* host);
int bm_mpm_num_particles(bm_mpm_handle h, int64_t* n);
/* Device-resident access: the caller writes the arrays directly (already in
 * HBM), then declares the count with bm_mpm_commit_device_particles. */
int bm_mpm_particles_device(bm_mpm_handle h, bm_particles_device* out);
int bm_mpm_commit_device_particles(bm_mpm_handle h, int64_t n);

/* ---------------------------------------------------------------- substep */
/* MPMSolver::step(dt) (mpm_solver.h:62-87): clear, P2G, grid update, G2P. */
int bm_mpm_step(bm_mpm_handle h, float dt);
/* k substeps back to back (one call, no host round trip). */
int bm_mpm_step_n(bm_mpm_handle h, float dt, int32_t k);
/* The phases of step(), for callers that act on the grid in between -- the
 * couplers' contract is "between update_grid() and g2p"
 * (mpm_rigid_coupler.h:209-211).  The reference keeps these private. */
int bm_mpm_p2g(bm_mpm_handle h, float dt);
int bm_mpm_grid_update(bm_mpm_handle h, float dt);
int bm_mpm_g2p(bm_mpm_handle h, float dt);

/* ---------------------------------------------------------------- couplers */
/* M0 / M1 on the current grid (after bm_mpm_grid_update, before bm_mpm_g2p).
 * Synchronous: the reaction is written to host memory.  dt <= 0 gives a zero
 * reaction, as in the reference (mpm_rigid_coupler.cpp:62, _pic.h:92). */
int bm_mpm_couple(bm_mpm_handle h, int32_t method, const bm_collider* body, float dt, bm_reaction* out);

/* bm_mpm_couple without allocation or host synchronisation, for callers that
 * cannot allocate (code under HIP graph capture, tight serving loops): the
 * coupler's partial reactions live in a fixed 7,232-byte device array of the
 * library (one per GPU, held until the process ends), and the reaction is
 * written to `reaction` in the caller's DEVICE memory, on the handle's stream;
 * the call returns without waiting.  The grid is updated in place as by
 * bm_mpm_couple.  Calls that overlap on one GPU share that array and race:
 * order them on one stream.  No counterpart in the reference (its couplers,
 * mpm_rigid_coupler.h:212-214, return the reaction by value). */
int bm_mpm_couple_noalloc(bm_mpm_handle h, int32_t method, const bm_collider* body, float dt, bm_reaction* reaction);

/* ---------------------------------------------------------------- device body */
/* The rigid body of the coupled loop, kept in device memory so a coupled
 * substep needs no host round trip.  Body fields follow RigidBody
 * (rigid_body.h:23-60); the collider fields complete RigidColliderState
 * (mpm_rigid_coupler.h:72-84), whose pose and velocities come from the body. */
/* BM_BODY_STATIC: STATIC or sleeping, integration skipped (integrator.h:25);
 * BM_BODY_DYNAMIC: TimeIntegrator::integrate;
 * BM_BODY_KINEMATIC: moves at its own linear / angular velocity whatever
 *   the reaction (a dragged tool; no counterpart in the reference). */
enum bm_body_motion { BM_BODY_STATIC = 0, BM_BODY_DYNAMIC = 1, BM_BODY_KINEMATIC = 2 };

typedef struct {
    float position[3];
    float orientation[4];     /* quaternion (w, x, y, z) */
    float linear_velocity[3];
    float angular_velocity[3];
    float force[3];           /* accumulators, cleared by every integration */
    float torque[3];
    float inv_mass;
    float inv_inertia[9];     /* inv_inertia_tensor, body space, row-major */
    float linear_damping;
    float angular_damping;
    int32_t integrate;        /* enum bm_body_motion */
    float applied_force[3];   /* added to force at the start of each coupled substep (gravity * mass) */
    float applied_torque[3];  /* added to torque likewise (a drive torque) */
    float half_extents[3];    /* collider shape, as bm_collider */
    float friction;
    float mass;               /* RigidColliderState::mass */
    int32_t shape;
} bm_body;

/* Set / read the device body.  get also returns the reaction of the last
 * coupled substep (reaction may be NULL); both synchronize the stream. */
int bm_mpm_set_body(bm_mpm_handle h, const bm_body* body);
int bm_mpm_get_body(bm_mpm_handle h, bm_body* body, bm_reaction* reaction);
/* One coupling of the device body, the body half of the reference's coupled
 * loop (bench_compare_all.cpp:642-663): force += applied_force, torque +=
 * applied_torque; reaction = coupler(grid, collider of the body, dt); force +=
 * reaction.force, torque += reaction.torque; TimeIntegrator::integrate
 * (integrator.h:19-57).  Between bm_mpm_grid_update and bm_mpm_g2p.  out may
 * be NULL: then nothing is copied to the host and the call does not wait. */
int bm_mpm_couple_body(bm_mpm_handle h, int32_t method, float dt, bm_reaction* out);
/* k coupled substeps (p2g, grid update, couple_body, g2p) with no host
 * round trip: the order of mpm_rigid_coupler.h:209-211. */
int bm_mpm_step_coupled(bm_mpm_handle h, int32_t method, float dt, int32_t k);

/* ---------------------------------------------------------------- grid access */
/* Dense copy of the node box [org, org + dims) (get_grid(), mpm_solver.h:90).
 * Any output may be NULL.  mass_q: the int64 fixed-point node mass in units
 * of 2^mass_quantum_exp kg; vel: momentum after P2G, velocity after the grid
 * update (the reference keeps both in GridNode::velocity); force: the P2G
 * stress force (BM_MODE_REFERENCE only); alloc: block ever touched by P2G. */
int bm_mpm_download_grid(bm_mpm_handle h, const int32_t org[3], const int32_t dims[3], int64_t* mass_q,
                         float* mass, float* vel, float* force, uint8_t* active, uint8_t* alloc);
/* Seed nodes of the box directly (the reference's coupler tests seed
 * GridNode velocity/mass/active, tests/test_mpm_rigid_coupler.cpp:26-92).
 * Leaves the grid in the post-update phase. */
int bm_mpm_upload_grid(bm_mpm_handle h, const int32_t org[3], const int32_t dims[3], const float* mass,
                       const float* vel, const uint8_t* active);
/* The research harness's grid step of its grid-only coupler loops
 * (bench_compare_all.cpp:117-126 box drop, :666-675 rover wheel): on every
 * active node with mass > 0 of the current (post-update) grid, v_y += dvy,
 * then v_y = 0 if negative.  dvy is the caller's -g * dt. */
int bm_mpm_grid_settle(bm_mpm_handle h, float dvy);
/* Sum of the fixed-point node mass of the last P2G (exactly the sum of the
 * particles' quanta when every particle was inside the domain). */
int bm_mpm_grid_mass_total(bm_mpm_handle h, int64_t* mass_q_total);
int bm_mpm_mass_quantum_exp(bm_mpm_handle h, int32_t* e);
/* Number of blocks touched by the last P2G, and of 3x3 SVDs that took the
 * FP64 fallback since creation. */
int bm_mpm_stats(bm_mpm_handle h, int64_t* active_blocks, int64_t* fp64_fallbacks);

/* ---------------------------------------------------------------- tuning */
/* BM_MODE_FAST launch parameters (what the autotuner searches):
 * unit_particles: particles per workgroup unit (16..1024, default 640);
 * sort_every: substeps between particle re-binnings (>= 1, default 32);
 * variant: bit 2 = the two-kernel form (per-particle G2P writing deposit
 * records, then a per-block P2G accumulating each cell's particles in
 * registers; always set -- the single-kernel fused form was retired and a
 * variant without bit 2 returns BM_ERR_NOT_SUPPORTED); bit 3 = the per-cell
 * deposit in two passes over each cell's particles (fewer registers,
 * particle data read twice); bit 4 = the G2P per particle over the flat
 * particle range, bit 5 = the G2P over the re-binning units with each unit's
 * grid nodes staged in LDS (not both; with neither the staged form runs
 * when the scene has at least 6144 units).  Default 4.
 * Values <= 0 / < 0 keep the current setting.  No effect in
 * BM_MODE_REFERENCE. */
int bm_mpm_set_tuning(bm_mpm_handle h, int32_t unit_particles, int32_t sort_every, int32_t variant);

/* ---------------------------------------------------------------- timing */
/* Per-phase device time from HIP events recorded on the handle's stream
 * around every kernel launch while timing is on (the reference times only
 * wall-clock per coupler step, research/mpm_rigid_coupling/metrics.h:122-132). */
enum bm_phase {
    BM_PHASE_CLEAR = 0,   /* zero the blocks of the previous substep */
    BM_PHASE_P2G = 1,     /* particle-to-grid scatter */
    BM_PHASE_UPDATE = 2,  /* grid update (gravity, floor, walls) */
    BM_PHASE_COUPLE = 3,  /* M0 / M1 coupler and its reaction reduction */
    BM_PHASE_G2P = 4,     /* grid-to-particle gather + plasticity */
    BM_PHASE_G2P_P2G = 5, /* the retired fused G2P(n) + P2G(n+1) kernel (no longer recorded) */
    BM_PHASE_SORT = 6,    /* particle re-binning (BM_MODE_FAST) */
    BM_PHASE_OTHER = 7,
    BM_NUM_PHASES = 8
};
/* on = 1: every phase (event pairs recorded between the launches, which
 * adds their own gaps to the stream); on = 2: only the two-kernel fast
 * form's G2P and P2G kernels, timed by events that their own dispatches
 * record (nothing between launches); on = 3: as 2, the G2P (particle pass)
 * kernel only; 0: off.  Events recorded by a dispatch still lengthen it:
 * level 2 adds about 10 us per wheel-bed substep of 0.32 ms. */
int bm_mpm_set_timing(bm_mpm_handle h, int32_t on);
/* Synchronises, then returns the accumulated milliseconds and launch counts
 * per phase (arrays of BM_NUM_PHASES); reset != 0 clears the accumulators. */
int bm_mpm_read_timing(bm_mpm_handle h, int32_t reset, double* ms, int64_t* launches);

/* ---------------------------------------------------------------- batches */
/* Many independent scenes advanced together (no counterpart in the
 * reference, whose solver holds one scene).  The scenes share device arrays:
 * particles of scene s occupy one contiguous range, its node box one range of
 * grid blocks, and every kernel of a substep runs once for all scenes, so
 * small scenes fill the GPU together.  Each scene has its own dx, box,
 * material, mass quantum and optionally a device body coupled by M0 or M1.
 * The physics is BM_MODE_FAST's, scene by scene.  Scenes can be added and
 * removed between steps (the remaining scenes are packed; their state is
 * kept).  With graphs on, the substeps are captured once per batch layout
 * and replayed (the launches of 2 x sort_every substeps per replay). */
typedef struct bm_batch* bm_batch_handle;

typedef struct {
    int64_t max_particles;  /* particles of all scenes together */
    int64_t max_blocks;     /* 4^3-node grid blocks of all scene boxes together */
    int32_t max_scenes;
    void* stream;           /* hipStream_t; NULL = the batch creates its own */
} bm_batch_desc;

typedef struct {
    float dx;
    int32_t grid_origin[3];  /* the scene's node box, as bm_mpm_desc */
    int32_t grid_dims[3];
    float young_modulus, poisson_ratio, bulk_density;
    float plastic_alpha;
    float floor_friction;
    float gravity[3];
    int32_t walls;           /* as bm_mpm_set_walls */
    int32_t coupler;         /* BM_COUPLER_* for the scene's body; ignored without a body */
} bm_scene_desc;

int bm_batch_create(const bm_batch_desc* desc, bm_batch_handle* out);
int bm_batch_destroy(bm_batch_handle b);
int bm_batch_synchronize(bm_batch_handle b);
/* Add a scene with n particles (full state, as bm_mpm_upload_particles) and
 * an optional device body (NULL: none).  BM_ERR_OOM when the batch lacks
 * particle, block or scene room -- the caller's admission signal. */
int bm_batch_add_scene(bm_batch_handle b, const bm_scene_desc* desc, int64_t n, const bm_particles* host,
                       const bm_body* body, int32_t* scene_id);
int bm_batch_remove_scene(bm_batch_handle b, int32_t scene_id);
/* k coupled substeps of every scene (p2g, grid update, the scene's body
 * coupling, g2p). */
int bm_batch_step(bm_batch_handle b, float dt, int32_t k);
/* Errors are kept per scene: a scene whose particles left its box (or whose
 * body outgrew the coupler's scan bound) fails its own download / get_body
 * with the error, the other scenes step on.  bm_batch_synchronize reports the
 * first failed scene. */
int bm_batch_download_scene(bm_batch_handle b, int32_t scene_id, bm_particles* host);
int bm_batch_get_body(bm_batch_handle b, int32_t scene_id, bm_body* body, bm_reaction* reaction);
/* The scene's error status: BM_OK, or the code its download would return. */
int bm_batch_scene_error(bm_batch_handle b, int32_t scene_id, int32_t* code);
/* Scenes, particles and grid blocks in use. */
int bm_batch_usage(bm_batch_handle b, int32_t* scenes, int64_t* particles, int64_t* blocks);
/* Launch parameters: unit_particles and sort_every as bm_mpm_set_tuning;
 * deposit_groups = workgroups of the cell deposit, which walk the units the
 * device counted (default 4096).  Values <= 0 keep the current setting. */
int bm_batch_set_tuning(bm_batch_handle b, int32_t unit_particles, int32_t sort_every, int32_t deposit_groups);
/* Time each candidate (unit 512 / 640 x sort_every 8 / 16 / 32 x deposit
 * groups 2048 / 4096 / 8192) over at least `substeps` substeps (a whole
 * number of re-binning cycles), keep the fastest and report it.  The scenes'
 * state is restored afterwards; needs device memory for one copy of the
 * particles. */
int bm_batch_autotune(bm_batch_handle b, float dt, int32_t substeps, int32_t* unit_particles, int32_t* sort_every,
                      int32_t* deposit_groups);
/* Capture the substeps in a HIP graph and replay it (on != 0). */
int bm_batch_set_graph(bm_batch_handle b, int32_t on);
/* Device bytes a batch of these capacities allocates (linear in each). */
int bm_batch_footprint(int64_t particles, int64_t blocks, int32_t scenes, int64_t* bytes);

/* ---------------------------------------------------------------- serving */
/* A request loop over one batch (no counterpart in the reference, which has
 * no serving).  A request is a scene, its particles, an optional body and a
 * substep count, with an arrival time.  Requests whose arrival has come are
 * admitted while the batch has particle, block and scene room and the HBM
 * budget allows (a later request may pass an older one that does not fit,
 * at most max_bypass times), advance together with everything else in the
 * batch, and are retired one by one when their own substeps are done --
 * continuous batching: admissions and retirements happen between chunks of
 * at most max_chunk substeps, never overshooting a request's count.  Every
 * request runs with the server's dt.  The caller keeps a request's input
 * buffers until it is admitted and its output buffers until it is done.
 * A scene that raises a device error (a particle leaving its box) retires as
 * BM_REQ_FAILED with that status; the others run on. */
typedef struct bm_server* bm_server_handle;

typedef struct {
    bm_batch_desc batch;  /* the arenas, allocated once */
    int64_t hbm_budget;   /* bytes the running requests may hold (bm_batch_footprint of each
                             request's particles, blocks and scene); 0 = the arenas' */
    float dt;
    int32_t max_chunk;    /* most substeps between scheduling points (0 = 32) */
    int32_t max_bypass;   /* admissions past the oldest waiting request (0 = 8) */
} bm_server_desc;

typedef struct {
    bm_scene_desc scene;
    int64_t n;
    const bm_particles* input;  /* read at admission */
    const bm_body* body;        /* NULL = none; copied at submission */
    bm_particles* output;       /* written at retirement; NULL = not copied back */
    int32_t substeps;
    double arrival;             /* seconds on the server clock; < 0 = now */
} bm_request;

enum bm_request_state { BM_REQ_QUEUED = 0, BM_REQ_RUNNING = 1, BM_REQ_DONE = 2, BM_REQ_FAILED = 3, BM_REQ_CANCELLED = 4 };

typedef struct {
    int32_t state;              /* bm_request_state */
    int32_t status;             /* BM_OK, or the error of a failed request */
    int32_t substeps_done;
    int32_t pad;
    double arrival, admitted, finished;  /* server clock, seconds; -1 = not yet */
    bm_body body;               /* the body at retirement (requests with a body) */
    bm_reaction reaction;
} bm_request_info;

int bm_server_create(const bm_server_desc* desc, bm_server_handle* out);
int bm_server_destroy(bm_server_handle s);
/* Seconds since the server was created (the clock of arrivals and timings). */
int bm_server_clock(bm_server_handle s, double* seconds);
/* The server's batch, for tuning (bm_batch_set_tuning / _autotune / _set_graph).
 * Waits for the chunk in flight; the batch is the caller's until the next poll. */
int bm_server_batch(bm_server_handle s, bm_batch_handle* batch);
/* Queue a request.  BM_ERR_OOM when it could not fit even an empty server. */
int bm_server_submit(bm_server_handle s, const bm_request* req, int64_t* request_id);
/* One scheduling round: the chunk launched by the previous round ends and
 * the requests it finished leave the batch; admissions; the next chunk is
 * launched and not waited for -- with overlap on (the default) a launcher
 * thread of the server enqueues its kernels -- and while it runs, the
 * finished requests' outputs are written and arrived requests are staged
 * onto the device for their admission.  Sleeps until the next arrival when
 * nothing is running and nothing is left to report.  Writes up to
 * `capacity` ids of requests finished (done or failed) since the last
 * report; the rest are reported by later calls.  A request's output is
 * written when its id is reported; `finished` in its info is the time it
 * left the batch.  substeps_done counts launched substeps. */
int bm_server_poll(bm_server_handle s, int64_t* done_ids, int32_t capacity, int32_t* n_done);
/* Overlap of staging with stepping (on by default): a copy stream and a
 * staging area sized for the batch's particle capacity (108 bytes per
 * particle, in device memory and in pinned host memory, outside hbm_budget).
 * Off: each poll waits for its chunk, and admissions and retirements use the
 * batch's synchronous paths.  BM_ERR_OOM when the staging area could not be
 * allocated at creation. */
int bm_server_set_overlap(bm_server_handle s, int32_t on);
int bm_server_info(bm_server_handle s, int64_t request_id, bm_request_info* info);
/* Drop a queued or running request (its output is not written).  Waits for
 * the chunk in flight. */
int bm_server_cancel(bm_server_handle s, int64_t request_id);
int bm_server_load(bm_server_handle s, int32_t* queued, int32_t* running, int64_t* bytes_in_use, int64_t* budget);
/* Where the time went: seconds[4] = admitting, stepping (launch until the
 * device is idle, including the host staging done meanwhile), retiring,
 * sleeping for arrivals; counts[2] = chunks and substeps run.  Either may be
 * NULL. */
int bm_server_stats(bm_server_handle s, double* seconds, int64_t* counts);

/* ---------------------------------------------------------------- adjoint */
/* Differentiable rollout (no counterpart in the reference, whose autodiff
 * tape records nothing: mpm_autodiff_tape.h:63-92).  The forward substep is
 * the reference algorithm (P2G with Hencky-Kirchhoff stress, grid update
 * with gravity and the floor rule, the M0 clamp of a kinematic box tool,
 * G2P with the Drucker-Prager return map) in FP32 or FP64 on a dense node
 * box; the backward pass is its hand-written adjoint, recomputing each run
 * of checkpoint_every substeps from the state saved at its start.  The loss
 * is the mean over particles of |x(T) - x(0) - target|^2. */
typedef struct bm_diff* bm_diff_handle;

typedef struct {
    float dx;
    int32_t grid_origin[3];
    int32_t grid_dims[3];
    int64_t max_particles;
    int32_t precision;         /* 32 or 64 */
    int32_t checkpoint_every;  /* substeps between saved states (0 = 16) */
    int32_t reference_svd;     /* 1: the reference's SVD stopping rule (for comparing forwards; the
                                  adjoint needs the default, tighter rule) */
    int32_t rollouts;          /* rollouts run side by side, each with its own parameters and
                                  the same particles and target (0 = 1; at most
                                  BM_DIFF_MAX_ROLLOUTS): every launch covers all of them */
    void* stream;              /* hipStream_t; NULL = the handle creates its own */
} bm_diff_desc;

#define BM_DIFF_MAX_ROLLOUTS 4096

typedef struct {
    double young_modulus, poisson_ratio;
    double friction_deg;       /* Drucker-Prager alpha = sqrt(2/3) 2 sin(phi) / (3 - sin(phi)) */
    double floor_friction;
    double gravity[3];
    int32_t has_tool;          /* a kinematic box tool coupled by M0 */
    int32_t pad;
    double tool_center[3];     /* at substep 0; it moves with tool_velocity */
    double tool_half_extents[3];
    double tool_quat[4];       /* (w, x, y, z) */
    double tool_velocity[3];
    double tool_friction;
} bm_diff_params;

typedef struct {
    double tool_velocity[3];   /* d loss / d parameter */
    double friction_deg;
    double young_modulus;
    double alpha, mu, lambda;  /* the same through the solver's own constants */
} bm_diff_grad;

int bm_diff_create(const bm_diff_desc* desc, bm_diff_handle* out);
int bm_diff_destroy(bm_diff_handle h);
/* Parameters of every rollout. */
int bm_diff_set_params(bm_diff_handle h, const bm_diff_params* params);
/* Parameters of rollout r alone (a parameter sweep or a population of
 * candidates); a forward needs every rollout's. */
int bm_diff_set_rollout_params(bm_diff_handle h, int32_t r, const bm_diff_params* params);
/* Device memory the forward may spend keeping every substep's grid (m, q and
 * the updated velocity: 7 values per node and substep) for the backward,
 * which then recomputes only the particle states of each checkpoint segment
 * instead of every segment's grids too.  A forward whose grids exceed it
 * keeps none.  Default 4 GiB; 0 = never.  No counterpart in the reference
 * (it has no adjoint); takes effect from the next forward. */
int bm_diff_set_grid_budget(bm_diff_handle h, int64_t bytes);
/* Initial particle state and the target displacement (3n floats; NULL = 0). */
int bm_diff_upload(bm_diff_handle h, int64_t n, const bm_particles* p, const float* target_displacement);
/* Run `substeps` substeps of every rollout, saving a state every
 * checkpoint_every; the loss of rollout 0. */
int bm_diff_forward(bm_diff_handle h, float dt, int32_t substeps, double* loss);
/* The same; losses[r] = rollout r's loss (one per rollout, or NULL). */
int bm_diff_forward_batch(bm_diff_handle h, float dt, int32_t substeps, double* losses);
/* Gradient of the last forward's loss of rollout 0; may be called again (the
 * same gradient) until the next forward or parameter change.  The backward
 * runs every rollout. */
int bm_diff_backward(bm_diff_handle h, bm_diff_grad* grad);
/* The same; grads[r] = rollout r's gradient (one per rollout). */
int bm_diff_backward_batch(bm_diff_handle h, bm_diff_grad* grads);
/* Final state of the last forward's rollout 0 (x, v, F, C; m and V0 are not
 * written). */
int bm_diff_download(bm_diff_handle h, bm_particles* p);
/* The same for rollout r. */
int bm_diff_download_rollout(bm_diff_handle h, int32_t r, bm_particles* p);
/* The rollout's two per-particle SVD maps and their hand-written
 * vector-Jacobian products on their own, for checking the adjoint against an
 * independent differentiation: map 0 = the Drucker-Prager return map
 * F -> U h(sigma) V^T (mpm_solver.h:114-166), map 1 = the first
 * Piola-Kirchhoff stress P(F) = U diag(tau_k / sigma_k) V^T of the
 * Hencky-Kirchhoff model (mpm_solver.h:186-217).  n row-major 3x3 matrices
 * F and cotangents Gbar in; G = map(F) and Fbar = Gbar . dG/dF out (9n doubles
 * each).  params supplies E, nu and the friction angle; precision 32 or 64 is
 * the arithmetic of the kernel.  No counterpart in the reference (its tape
 * records nothing). */
int bm_diff_map_vjp(int32_t map, int64_t n, const double* F, const double* Gbar, const bm_diff_params* params,
                    int32_t precision, double* G, double* Fbar);

#ifdef __cplusplus
}
#endif

#endif /* BM_MPM_H */
