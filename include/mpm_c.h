/*
 * mpm_c.h — C ABI of the MI355X (gfx950) MPM snow substep library
 * (libmpm_hip.so).  Plain pointers, sizes and status codes; no C++ or torch
 * types cross it.  It replaces the simulation-step surface of the reference's
 * MaterialPointMethod::LagrangeEulerView
 * (realtime-deformations/material_point_method.hpp:174-200), which
 * include/mpm/material_point_method.hpp re-exposes on top of these calls.
 *
 * Conventions
 *  - Every function returns MPM_OK (0) or an MPM_ERR_* code; the message of
 *    the last failure is mpm_last_error(ctx).  No exception crosses the ABI.
 *  - Host arrays are caller-owned and copied synchronously; the context owns
 *    all device memory, one HIP stream and its HIP graphs, and keeps no caller
 *    pointer.  A context is single-device and externally serialised.
 *  - Vectors are 3 floats per particle (x, y, z); matrices are 9 floats per
 *    particle in glm's column-major order (element (row r, col c) at 3c + r),
 *    the layout of the reference's glm::mat3 members.
 *  - Grid arrays are dense over the nodes, index i*J*K + j*K + k
 *    (realtime-deformations/material_point_method.hpp:129-131).
 */
#ifndef MPM_C_H
#define MPM_C_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct mpm_ctx mpm_ctx;

enum {
  MPM_OK = 0,
  MPM_ERR_ARG = 1,       /* invalid argument                                   */
  MPM_ERR_HIP = 2,       /* HIP runtime failure (message has the HIP error)    */
  MPM_ERR_BOUNDS = 3,    /* a particle's 5^3 stencil leaves the grid           */
  MPM_ERR_NONFINITE = 4, /* the SVD met a non-finite F (reference: SVD Error),
                            or a particle's velocity or stress input is not
                            finite on the fused path                           */
  MPM_ERR_STATE = 5,     /* call out of order (e.g. no particles set)          */
  MPM_ERR_NOMEM = 6,     /* device allocation failed                           */
  MPM_ERR_CFL = 7        /* warning, reported by mpm_scene_status only: on the
                            fused path a particle moved more than one cell in a
                            substep (|v| dt > h), or the elastic wave of its
                            hardened moduli crosses more than two cells
                            (dt sqrt((lambda + 2 mu) V0 / m) > 2h)              */
};

/* F_E recomposition after the singular-value clamp
 * (realtime-deformations/material_point_method.cpp:323-326). */
enum {
  MPM_COMPOSE_LITERAL = 0, /* the reference's U^T S V of the SVD of F_hat^T    */
  MPM_COMPOSE_TEXTBOOK = 1 /* V S U^T, the physically correct recomposition    */
};

/* Grid boundary handling (bit mask). */
enum {
  MPM_BC_COLLIDERS = 1, /* SDF box colliders, material_point_method.cpp:264-304 */
  MPM_BC_DOMAIN = 2     /* sticky floor band j<3, separating walls elsewhere    */
};

/* Stages of one substep, in the reference caller's order
 * (realtime-deformations/main.cpp:192-218). */
enum {
  MPM_STAGE_P2G = 0,     /* rasterizeParticlesToGrid            cpp:94-129  */
  MPM_STAGE_VOLUMES = 1, /* computeParticleVolumesAndDensities  cpp:131-142 */
  MPM_STAGE_FORCES = 2,  /* computeExplicitGridForces           cpp:235-254 */
  MPM_STAGE_GRID = 3,    /* gridVelocitiesUpdate                cpp:256-262 */
  MPM_STAGE_COLLIDE = 4, /* gridBasedCollisions                 cpp:298-304 */
  MPM_STAGE_FUPDATE = 5, /* updateDeformationGradient           cpp:306-330 */
  MPM_STAGE_G2P = 6,     /* updateParticleVelocities            cpp:332-342 */
  MPM_STAGE_ADVECT = 7   /* updateParticlePositions             cpp:344-350 */
};

typedef struct mpm_params {
  int32_t dims[3];   /* grid nodes per axis: the ctor's max_i, max_j, max_k     */
  int32_t compose;   /* MPM_COMPOSE_*                                           */
  int32_t bc_mode;   /* MPM_BC_* mask                                           */
  int32_t flags;     /* reserved, 0                                             */
  float h;           /* grid spacing: the reference's global WeightCalculator::h*/
  float E;           /* Young's modulus when no per-particle E0 is given (1.4e5)*/
  float nu;          /* Poisson ratio (0.2)                                     */
  float xi;          /* hardening coefficient (10)                              */
  double theta_c;    /* clamp 1 - theta_c (2.5e-2)                              */
  double theta_s;    /* clamp 1 + theta_s (5e-3 in the reference)              */
  float gravity[3];  /* (0, -9.8, 0)                                            */
  float svd_gap_fp64;/* TEXTBOOK only: FP64 re-solve when the smallest gap
                        between singular values is below this times the
                        largest; 0 disables                                    */
  int32_t resort_every; /* fused path: re-bin particles every this many
                        substeps (0: 16).  Results do not depend on it: a
                        particle that left its block's halo margin takes a
                        global-memory path until the next sort.              */
} mpm_params;

/* Oriented box collider: the pose the reference reads from Object3D
 * (realtime-deformations/mesh.hpp:11-24) and MeshCollider::velocity
 * (realtime-deformations/material_point_method.hpp:74-93). */
typedef struct mpm_collider {
  float translation[3];
  float rotation[4]; /* quaternion w, x, y, z                                   */
  float scale[3];    /* half extents                                            */
  float velocity[3];
} mpm_collider;

/* Fill p with the reference's constants for a grid of dims and spacing h. */
void mpm_default_params(mpm_params* p, int32_t ni, int32_t nj, int32_t nk, float h);

/* LagrangeEulerView(max_i, max_j, max_k, particlesNum)
 * (realtime-deformations/material_point_method.cpp:144-154).
 * On failure *out is either NULL (invalid arguments) or a context that only
 * carries the reason (mpm_last_error) and must still go to mpm_destroy; the
 * same holds for mpm_create_batch and mpm_create_from_snapshot. */
int mpm_create(const mpm_params* p, int32_t n_particles, mpm_ctx** out);
int mpm_destroy(mpm_ctx* ctx);
const char* mpm_last_error(const mpm_ctx* ctx);
int32_t mpm_num_particles(const mpm_ctx* ctx);

/* Particle positions, velocities, masses and optional per-particle Young's
 * moduli (NULL: params E); F_E = F_P = I, B = 0, V0 = 0, as
 * initializeParticles leaves them (material_point_method.cpp:18-63).  For a
 * batch the arrays hold every scene's particles, scene after scene. */
int mpm_set_particles(mpm_ctx* ctx, const float* x, const float* v, const float* m, const float* E0);
/* Overwrite any of B, F_E, F_P (9 floats each) and V0; NULL leaves a field. */
int mpm_set_state(mpm_ctx* ctx, const float* B, const float* FE, const float* FP, const float* V0);
/* The same colliders for every scene of the context. */
int mpm_set_colliders(mpm_ctx* ctx, const mpm_collider* c, int32_t k);

/* rasterizeParticlesToGrid + computeParticleVolumesAndDensities
 * (realtime-deformations/main.cpp:53-54), on the device for every scene:
 * binned mass scatter, node masses, density gathered with the particle's
 * weights, V0 = m / density (material_point_method.cpp:131-142). */
int mpm_init_volumes(mpm_ctx* ctx);

/* One legacy stage (MPM_STAGE_*), exactly the reference method's semantics. */
int mpm_stage(mpm_ctx* ctx, int32_t stage, float dt);

/* nsteps substeps in the reference's order on the fused, binned path
 * (main.cpp:192-218 repeated).  Asynchronous on the context stream; errors
 * raised by the device are reported by the next synchronising call. */
int mpm_substeps(mpm_ctx* ctx, float dt, int32_t nsteps);
int mpm_sync(mpm_ctx* ctx);
/* Replay runs of substeps through captured HIP graphs (default 1). */
int mpm_use_graphs(mpm_ctx* ctx, int32_t enable);

/* Reproducible mode (default 0).  The reference sums its grid serially in
 * particle order (material_point_method.cpp:99-117), so a run repeats
 * exactly; the device sums with atomics, whose order varies from run to
 * run.  In this mode every P2G contribution of mpm_substeps and
 * mpm_init_volumes is rounded to a per-scene quantum first: 2^(e-40) for
 * mass and 2^(e-38) for momentum, where 2^e is the power of two just above
 * the scene's largest particle mass (taken again whenever masses are set).
 * The grid's FP64 sums of such multiples are exact while a node stays below
 * 2^13 m_ref of mass and 2^15 m_ref m/s of momentum, so they do not depend
 * on the order of the atomics: results repeat bit for bit from run to run,
 * for any resort_every and block-pass width, and a scene gives the same
 * bits alone as inside a batch.  A contribution of 2^-40 m_ref or less is
 * rounded (to zero below half that), so nodes reached only by stencil
 * weights below about 1e-12 lose their mass.  The fused pass costs more
 * (two FP64 adds per contribution).  Not covered: mpm_stage (FP32 force
 * atomics) and the reverse mode's recording passes. */
int mpm_set_reproducible(mpm_ctx* ctx, int32_t enable);
int mpm_get_reproducible(mpm_ctx* ctx, int32_t* enabled);

/* Readback into host arrays (NULL skips a field).  getParticles()
 * (material_point_method.hpp:181-183). */
int mpm_get_particles(mpm_ctx* ctx, float* x, float* v, float* B, float* FE, float* FP, float* V0);
/* Dense grid of scene 0: mass, velocity (3), force (3, stage path only), active flag. */
int mpm_get_grid(mpm_ctx* ctx, float* m, float* vel, float* force, uint8_t* active);

/* ---- batched scenes (no reference counterpart: the reference simulates one
 * scene through globals).  One context steps any number of independent
 * scenes together: each scene has its own grid (dims, h), material,
 * gravity, boundary mode and colliders; one binning sort and one launch per
 * stage cover all of them.  params[0] sets compose and resort_every for the
 * batch (compose must agree).  Particles are indexed scene after scene;
 * mpm_num_particles is the total.  The per-stage calls (mpm_stage) need a
 * single-scene context. ---------------------------------------------- */
int mpm_create_batch(const mpm_params* params, const int32_t* n_particles, int32_t nscenes, mpm_ctx** out);
int32_t mpm_num_scenes(const mpm_ctx* ctx);
int mpm_scene_set_particles(mpm_ctx* ctx, int32_t scene, const float* x, const float* v, const float* m,
                            const float* E0);
int mpm_scene_set_state(mpm_ctx* ctx, int32_t scene, const float* B, const float* FE, const float* FP,
                        const float* V0);
int mpm_scene_set_colliders(mpm_ctx* ctx, int32_t scene, const mpm_collider* c, int32_t k);
int mpm_scene_get_particles(mpm_ctx* ctx, int32_t scene, float* x, float* v, float* B, float* FE, float* FP,
                            float* V0);
int mpm_scene_get_grid(mpm_ctx* ctx, int32_t scene, float* m, float* vel, float* force, uint8_t* active);
/* Synchronises and returns, per scene, MPM_OK or the MPM_ERR_* its particles
 * raised since the last status read (a non-finite scene does not stop the
 * others); clears the flags.  Precedence: MPM_ERR_NONFINITE (a non-finite
 * velocity, stress input or F on the fused path, or the SVD's rejection),
 * then MPM_ERR_BOUNDS, then the MPM_ERR_CFL warning.  Other calls report the
 * first two as errors of the context; the CFL warning stays pending until
 * this call reads it.  The flags are raised on the device each substep
 * (an atomic OR only by a failing particle) and cost no synchronisation. */
int mpm_scene_status(mpm_ctx* ctx, int32_t* status);

/* ---- device scene generators.  Seeded particle placement written on the
 * GPU into a scene's particle slots (caller order = generation order), with
 * F_E = F_P = I, B = 0, V0 = 0 as after mpm_set_particles; follow with
 * mpm_init_volumes.  Counter-based randomness (Philox4x64-10 keyed by seed):
 * the same seed gives the same particles on any launch shape.
 *   MPM_GEN_LATTICE replaces initializeParticles
 *     (realtime-deformations/material_point_method.cpp:18-63): eight
 *     quarter-offset candidates per cell of the cube of half width
 *     int(radius / h) cells around ivec3(centre / h), jittered uniformly in a
 *     ball of 1/4 cell (the reference draws it from rand(), utils.h:114-127),
 *     kept within `radius` (the reference's width) of `centre`, filled from the
 *     back of the array.  *candidates returns the kept count; fewer than the
 *     scene's particles is MPM_ERR_ARG (the reference leaves those particles
 *     unplaced).
 *   MPM_GEN_BALL (no reference counterpart): n points uniform in a ball.
 *   MPM_GEN_CELLS (no reference counterpart): a box of cells with
 *     U{ppc_lo..ppc_hi} particles each, adjusted by +-1 steps in a seeded cell
 *     order to exactly the scene's count, uniform inside each cell.
 * E0 is uniform in [E_lo, E_hi) (E_lo == E_hi: constant).  A particle whose
 * stencil leaves the grid is MPM_ERR_BOUNDS and leaves the scene unset. ---- */
enum { MPM_GEN_BALL = 0, MPM_GEN_CELLS = 1, MPM_GEN_LATTICE = 2 };
typedef struct mpm_generator {
  int32_t kind;
  int32_t reserved;
  uint64_t seed;
  float centre[3];     /* BALL: centre; LATTICE: origin */
  float radius;        /* BALL: radius; LATTICE: width */
  int32_t cell_lo[3];  /* CELLS: first cell (i, j, k) of the box */
  int32_t cells[3];    /* CELLS: box extent in cells */
  int32_t ppc_lo, ppc_hi;
  float velocity[3];
  float mass;          /* per particle */
  float E_lo, E_hi;
} mpm_generator;
int mpm_scene_generate(mpm_ctx* ctx, int32_t scene, const mpm_generator* g, int64_t* candidates);
int mpm_generate(mpm_ctx* ctx, const mpm_generator* g, int64_t* candidates);

/* ---- readback that allocates nothing (no counterpart in the reference).
 * For callers that cannot allocate device memory: code under HIP graph
 * capture, tight serving loops, harnesses that forbid hipMalloc.  The same
 * fields as mpm_scene_get_particles (caller order, NULL skips), equal to them
 * bit for bit, without any device allocation: each field is staged through a
 * file-scope device array of 9 * MPM_NOALLOC_MAX_PARTICLES floats (9 MiB),
 * which every GPU the library runs on holds until the process ends.  A scene
 * with more particles than MPM_NOALLOC_MAX_PARTICLES is MPM_ERR_ARG.  Calls
 * on one GPU share that array, each leaving its last field there: calls that
 * overlap on one GPU (from several threads or contexts) race. */
#define MPM_NOALLOC_MAX_PARTICLES 262144
int mpm_scene_get_particles_noalloc(mpm_ctx* ctx, int32_t scene, float* x, float* v, float* B, float* FE,
                                    float* FP, float* V0);

/* ---- readback and snapshots ---------------------------------------------
 * m and E0 of one scene (caller order; NULL skips). */
int mpm_scene_get_material(mpm_ctx* ctx, int32_t scene, float* m, float* E0);
/* Positions only, all particles in caller order (n x 3): the viewer's read
 * of getParticles() each frame (realtime-deformations/main.cpp:257-271).
 * Goes through a context-owned device staging buffer and pinned host mirror
 * (allocated on first use, 24 bytes per particle).  mpm_get_positions copies
 * into x and synchronises.  mpm_positions_begin enqueues the gather on the
 * context stream and the copy on a second stream, so it overlaps substeps
 * issued after it; mpm_positions_wait blocks until the copy lands and points
 * *x at the pinned mirror, valid until the next mpm_positions_begin. */
int mpm_get_positions(mpm_ctx* ctx, float* x);
int mpm_positions_begin(mpm_ctx* ctx);
int mpm_positions_wait(mpm_ctx* ctx, const float** x);
/* Snapshot file: every scene's parameters, colliders and full particle state
 * (x, v, B, F_E, F_P, m, V0, E0 in caller order) plus a caller step counter;
 * the layout is documented in mpm_io.hip and read by mpm_amd.snapshot.
 * mpm_snapshot_load needs a context with the same scenes (particle counts and
 * parameters) and replaces its state, clearing the status flags;
 * mpm_create_from_snapshot creates that context.  Resume = create from the
 * snapshot and continue with mpm_substeps. */
int mpm_snapshot_save(mpm_ctx* ctx, const char* path, int64_t step);
int mpm_snapshot_load(mpm_ctx* ctx, const char* path, int64_t* step);
int mpm_create_from_snapshot(const char* path, mpm_ctx** out, int64_t* step);

/* ---- launch tuning (no reference counterpart).  resort_every: substeps
 * between re-binnings; block_grid: workgroups of the persistent block pass.
 * mpm_autotune times each candidate (resort_every in {8, 16, 32, 64} x
 * block_grid in {1024, 2048, 4096}, capped at the block count) over at least
 * `steps` substeps, rounded up to whole re-bin cycles and at least 128, from
 * the current state; it restores that state, keeps the fastest and returns
 * its time per substep.  After either call the next substep re-bins, so a
 * run of whole re-bin cycles starts there. */
int mpm_set_tuning(mpm_ctx* ctx, int32_t resort_every, int32_t block_grid);
int mpm_get_tuning(mpm_ctx* ctx, int32_t* resort_every, int32_t* block_grid);
int mpm_autotune(mpm_ctx* ctx, float dt, int32_t steps, double* best_ms_per_substep);

/* ---- reverse mode (no reference counterpart: the reference has no
 * gradients).  mpm_adjoint_forward runs nsteps substeps from the current
 * state, keeping a checkpoint every `checkpoint_every` substeps (device
 * memory: 132 bytes per particle per checkpoint; checkpoint_every = 1
 * instead keeps every substep's record, 132 bytes per particle plus room for
 * 16 bytes per grid node and 4 per grid block per substep, and the backward
 * re-runs nothing; checkpoint_every = 0 chooses: 1 when those records fit in
 * half of the free device memory, else 20).  mpm_adjoint_interval reports
 * the interval the last forward used (0 before any).  Checkpoints, records and the backward's workspaces stay
 * allocated for the next call of the same shape, which reuses them; records
 * are released by a checkpoint_every = 1 call of another length, everything
 * by mpm_destroy.
 * mpm_adjoint_backward takes
 * dL/dx and dL/dv of the final particles (n x 3, caller order; NULL = 0) and
 * returns dL/dx0, dL/dv0 (n x 3) and dL/dE0 (n) of the state the forward
 * started from (NULL skips), recomputing each segment from its checkpoint; it
 * leaves the context at that initial state.  TEXTBOOK composition and domain
 * walls only (MPM_ERR_STATE otherwise); V0 is held fixed, so dL/dx0 leaves
 * out V0's dependence on the initial positions. */
int mpm_adjoint_forward(mpm_ctx* ctx, float dt, int32_t nsteps, int32_t checkpoint_every);
int mpm_adjoint_interval(mpm_ctx* ctx, int32_t* checkpoint_every);
int mpm_adjoint_backward(mpm_ctx* ctx, const float* gx, const float* gv, float* gx0, float* gv0, float* gE0);

/* ---- measurement (no reference counterpart: the reference times stages
 * with millisecond std::chrono stamps, realtime-deformations/utils.h:138-154,
 * main.cpp:192-225) ----------------------------------------------------- */
enum {
  MPM_KERNEL_P2G = 0,   /* first substep of a call: stress + F update + P2G   */
  MPM_KERNEL_GRID = 1,  /* grid velocity update + boundary conditions         */
  MPM_KERNEL_G2P = 2,   /* end of a call: G2P + advection                     */
  MPM_KERNEL_BIN = 3,   /* binning: keys, radix sort, reorder                 */
  MPM_KERNEL_FUSED = 4, /* other substeps: G2P + advection + stress + F
                           update + P2G in one pass                           */
  MPM_NKERNELS = 8
};
/* Starts a profiled region: a HIP event on the context stream marks its
 * start.  With per_kernel != 0 every fused-path launch is also bracketed by
 * events (and launched directly, not through the substep HIP graph);
 * mpm_profile_end synchronises and returns, per MPM_KERNEL_* id, the summed
 * kernel milliseconds and launch counts, and the region's wall time. */
int mpm_profile_begin(mpm_ctx* ctx, int32_t per_kernel);
int mpm_profile_end(mpm_ctx* ctx, double* kernel_ms, int32_t* launches, double* total_ms);
/* Number of grid nodes with non-zero mass after the last grid update. */
int mpm_count_active(mpm_ctx* ctx, int64_t* active_nodes);
/* A plain streaming copy of `bytes` bytes (a multiple of 16) between two
 * device buffers on the context stream: the known-size kernel that scales
 * the profiler's HBM byte counters. */
int mpm_copy_probe(mpm_ctx* ctx, int64_t bytes);

/* Binning check (no reference counterpart): bins the current particles as
 * the fused path does and returns, for n particles in their current device
 * order, the sorted keys (block * 64 + cell-in-block), the stable sort
 * permutation (sorted position -> current slot) and, after the round-robin
 * reorder, the caller index held by every slot.  Any output may be NULL. */
int mpm_debug_bin(mpm_ctx* ctx, uint32_t* sorted_keys, uint32_t* perm, int32_t* slot_caller_index);

/* Home-block list of the last binning (no reference counterpart), in the
 * order the block passes claim it: per block (grid block over all scenes,
 * first slot, end slot, scene), 4 ints each.  *count receives the number of
 * blocks; at most max_blocks are written (blocks may be NULL to query the
 * count). */
int mpm_debug_blocks(mpm_ctx* ctx, int32_t* blocks, int32_t max_blocks, int32_t* count);

#ifdef __cplusplus
}
#endif
#endif /* MPM_C_H */
