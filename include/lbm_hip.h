/*
 * lbm_hip.h — C ABI of liblbm_hip.so, the MI355X (gfx950) time step of a
 * lattice-Boltzmann engine for squirmers (LBMengine.jl semantics).
 *
 * The reference has no FFI: its only seam is the backend directory chosen
 * at module load (src/LBMengine.jl:12-24), where every backend defines the
 * same generic functions over its own `LBMmodel`.  The entry points below
 * are what a third backend directory would `ccall` in place of those
 * functions; each one cites the reference function it replaces.
 *
 * Conventions
 *  - Every function returns int: 0 (LBM_OK) on success, a negative LBM_E*
 *    code on failure; lbm_last_error() then returns a thread-local message.
 *    No C++ exception or abort crosses this boundary.
 *  - Arrays use the reference's Julia layout: column-major, x fastest, the
 *    velocity (or vector-component) index slowest; q is 0-based here where
 *    the reference is 1-based.  Host buffers are caller-owned and copied
 *    synchronously; device memory is owned by the handle.
 *  - A handle is used from one host thread at a time.
 */
#ifndef LBM_HIP_H
#define LBM_HIP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define LBM_OK 0
#define LBM_EINVAL -1     /* invalid argument or configuration            */
#define LBM_ENOMEM -2     /* device or host allocation failed             */
#define LBM_EDEVICE -3    /* HIP runtime error                            */
#define LBM_ESTATE -4     /* operation not valid in the current state     */
#define LBM_ERANGE -5     /* a particle left the region the step supports */

#define LBM_COLLISION_BGK 0 /* :bgk  gpu/fluids.jl:69-71  */
#define LBM_COLLISION_TRT 1 /* :trt  gpu/fluids.jl:72-88  */

#define LBM_STORAGE_F64 0   /* populations f_i stored in FP64                  */
#define LBM_STORAGE_F32 1   /* deviations f_i - w_i rho0 stored in FP32        */
#define LBM_STORAGE_F16 2   /* deviations scaled by 2^14/w_i stored in FP16    */
#define LBM_STORAGE_F16C 3  /* deviations over w_i, (f_i - w_i rho0) / w_i, in a
                               16-bit float of 1 sign, 4 exponent and 11
                               mantissa bits, |value| < 1/8 (clamped beyond) */

#define LBM_SEMANTICS_GPU 0 /* follow src/gpu/ (Rodrigues rotation of e, ...)  */
#define LBM_SEMANTICS_CPU 1 /* follow src/cpu/ (swimming direction not rotated) */

#define LBM_FORCE_NONE 0
#define LBM_FORCE_UNIFORM 1 /* constant Guo body force, zeroed on walls         */
#define LBM_FORCE_FIELD 2   /* per-node Guo body force (lbm_set_force_field)   */

#define LBM_PARTICLE_BEAD 0     /* addBead!      gpu/initMethods.jl:7-89   */
#define LBM_PARTICLE_SQUIRMER 1 /* addSquirmer!  gpu/initMethods.jl:91-222 */

/* Particles per lattice.  Up to 8 (the particles a box block works with),
 * every kernel of a run applies the previous one's moves itself; with more,
 * a small move kernel follows each step kernel, and the run fails with
 * LBM_ERANGE if more than 8 particles come within reach of one another
 * (their link boxes, grown by two nodes, overlap).  The adjoint, the boxes
 * of a batch and lbm_run_noalloc take at most 8. */
#define LBM_MAX_PARTICLES 64

/* lbm_config.flags */
#define LBM_FLAG_INCOMPRESSIBLE 1 /* isFluidCompressible = false: the equilibrium
                                     w rho + w rho0 (...) (gpu/fluids.jl:51), with
                                     rho0 = lbm_config.rho0 the uniform initial
                                     density                                   */

/* Resolved model parameters: the output of modelInit's unit and scheme
 * resolution (gpu/initMethods.jl:440-702), done by the host front end. */
typedef struct lbm_config {
  int32_t dims;          /* 2 or 3                                        */
  int32_t q;             /* 9 (D2Q9, dims 2); 15, 19 or 27 (D3Q15, D3Q19,
                            D3Q27, dims 3; gpu/misc.jl:195-272)           */
  int64_t n[3];          /* nodes per dimension; n[2] = 1 when dims == 2  */
  int32_t collision;     /* LBM_COLLISION_*                               */
  int32_t storage;       /* LBM_STORAGE_*                                 */
  int32_t semantics;     /* LBM_SEMANTICS_*                               */
  int32_t walled;        /* bit k: dimension k has a wall layer at index 0
                            and n[k]-1 (walledDimensions, misc.jl:7-31)   */
  int32_t force;         /* LBM_FORCE_*                                   */
  int32_t flags;         /* LBM_FLAG_*                                    */
  double force_uniform[3]; /* used when force == LBM_FORCE_UNIFORM        */
  double dt;             /* timeStep                                      */
  double dx;             /* latticeParameter (the GPU backend forces dx == dt) */
  double tau;            /* relaxationTime (BGK) or relaxationTimePlus (TRT) */
  double tau_minus;      /* relaxationTimeMinus (TRT only)                */
  double rho0;           /* shift of the deviation storage (FP32/FP16)    */
  const double *coords[3]; /* node coordinates along each dimension, length
                              n[k]; the values of Julia's range(xlims..., step=dx) */
} lbm_config;

/* One rigid spherical particle, parameters resolved by addSquirmer!/addBead!
 * (mass from the origin-centred sphere, B1 from swimmingSpeed, ...). */
typedef struct lbm_particle {
  int32_t kind;            /* LBM_PARTICLE_*                              */
  int32_t couple_forces;   /* coupleForces                                */
  int32_t couple_torques;  /* coupleTorques                               */
  int32_t reserved0;
  double radius;
  double position[3];
  double velocity[3];
  double angular_velocity[3]; /* in 2D only [0] is used (scalar omega)   */
  double swimming_direction[3];
  double B1, B2;
  double inverse_mass;
  double inverse_moment_of_inertia;
} lbm_particle;

typedef struct lbm_handle lbm_handle;

/* Thread-local text of the last error on this thread ("" if none). */
const char *lbm_last_error(void);

/* Library version string. */
const char *lbm_version(void);

/* Number of visible HIP devices (0 when none; never an error). */
int lbm_device_count(int32_t *count);

/* Replaces the device allocation half of modelInit (gpu/initMethods.jl:440-702):
 * allocates the populations for cfg and the step's working state.  The
 * populations are undefined until lbm_init_uniform or lbm_set_populations. */
int lbm_create(const lbm_config *cfg, lbm_handle **out);

int lbm_destroy(lbm_handle *h);

/* findInitialConditions for uniform rho and u (gpu/initMethods.jl:417-438,
 * 657-668): f_i = feq_i(rho, u - dt F / (2 rho)), with rho = 0 on walls. */
int lbm_init_uniform(lbm_handle *h, double rho, const double *u);

/* Per-node Guo force density, host FP64 (N..., dims); zeroed on walls as
 * modelInit does (gpu/initMethods.jl:643-647).  Requires LBM_FORCE_FIELD. */
int lbm_set_force_field(lbm_handle *h, const double *force);

/* The wall region of modelInit's `walls` function (gpu/initMethods.jl:
 * 591-605): wall[n] != 0 marks node n (N..., host, Julia order) a wall
 * node, in addition to the layers of the walled dimensions.  Wall nodes hold
 * zero populations and zero rho and u; the fluid next to them bounces back
 * halfway (bounceBackPrep, gpu/aux.jl:187-198).  Call before the populations
 * are set and before particles are added. */
int lbm_set_walls(lbm_handle *h, const uint8_t *wall);

/* solidNodeVelocity of modelInit's `movingWalls` function (gpu/initMethods.jl:
 * 607-626): u (N..., dims), host FP64.  Nodes with a non-zero velocity join
 * the wall region; a population that bounces back off a moving wall node
 * x - c_q takes 2 w_q / c_s^2 rho(x) (u_w(x - c_q) . c_q) more
 * (gpu/fluids.jl:168-174).  Call before the populations are set and before
 * particles are added. */
int lbm_set_wall_velocity(lbm_handle *h, const double *u);

/* model.distributions <- f / f <- model.distributions, host FP64 (N..., Q). */
int lbm_set_populations(lbm_handle *h, const double *f);
int lbm_get_populations(lbm_handle *h, double *f);

/* hydroVariablesUpdate! (gpu/fluids.jl:7-37) evaluated on the current
 * populations: rho (N...) and u (N..., dims), host FP64.  Either pointer
 * may be NULL. */
int lbm_get_hydro(lbm_handle *h, double *rho, double *u);

/* addSquirmer!/addBead! after parameter resolution, including the initial
 * moveParticles!(; initialSetup = true) (gpu/initMethods.jl:91-222,
 * gpu/particles.jl:7-92).  *id receives the 0-based particle index. */
int lbm_add_particle(lbm_handle *h, const lbm_particle *p, int32_t *id);

/* Current state of particle id (position, velocity, angular velocity,
 * swimming direction; the other fields echo what was added). */
int lbm_get_particle(lbm_handle *h, int32_t id, lbm_particle *out);

/* Momentum-exchange force and torque of the last step, before the
 * integrator consumed them (forceInput/torqueInput, gpu/fluids.jl:193-204). */
int lbm_get_particle_load(lbm_handle *h, int32_t id, double force[3],
                          double torque[3]);

/* ---- particle-particle interactions (molecularDynamicsTick!,
 * gpu/particles.jl:100-197) ----
 * model.particleInteractions with each pair list expanded, in the order
 * the reference applies them: every term adds to forceInput from the
 * positions before the tick's moves, then every particle moves. */
#define LBM_INTERACTION_LINEAR 0        /* addLinearBond!    gpu/initMethods.jl:229-245, particles.jl:110-121 */
#define LBM_INTERACTION_BISTABLE 1      /* addBistableBond!  gpu/initMethods.jl:282-308, particles.jl:123-134 */
#define LBM_INTERACTION_POLAR 2         /* addPolarBond!     gpu/initMethods.jl:247-273, particles.jl:136-158 */
#define LBM_INTERACTION_DIPOLE 3        /* addDipoles!, one pair   gpu/initMethods.jl:316-359, particles.jl:160-177 */
#define LBM_INTERACTION_LENNARD_JONES 4 /* addLennardJones!/addWCA!, one pair  gpu/initMethods.jl:361-414, particles.jl:179-197 */

#define LBM_MAX_INTERACTIONS 256

typedef struct lbm_interaction {
  int32_t kind;      /* LBM_INTERACTION_*                                   */
  int32_t ids[3];    /* 0-based particle indices id1, id2 (and id3, POLAR)   */
  double params[4];  /* LINEAR: equilibriumDisplacement, stiffness
                        BISTABLE: trapRadius, a, b
                        POLAR: equilibriumAngle, stiffness
                        DIPOLE: dipoleConstant, scalar-field flag (1: B is
                                the number field[0], 2D), schedule flag
                                (1: B changes with time and is read, tick
                                by tick, from lbm_set_field_schedule)
                        LENNARD_JONES: epsilon, sigma, cutoff             */
  double field[3];   /* DIPOLE: the magnetic field B (when not scheduled)   */
} lbm_interaction;

/* Replaces the handle's interaction terms with terms[0..n) (n = 0 clears
 * them).  GPU semantics only, as the CPU backend has no interactions. */
int lbm_set_interactions(lbm_handle *h, const lbm_interaction *terms, int32_t n);

/* A magnetic field that is a function of time (addDipoles! with
 * magneticField::Function, gpu/initMethods.jl:337-346, evaluated as
 * magneticField(model.time) at gpu/particles.jl:161): B[i][0..3) is the field
 * of tick first_tick + i, i < count, that is the field at the model time
 * that tick ends at (the time molecularDynamicsTick! sees, after tick!
 * advanced it).  Dipole terms with the schedule flag read it; a tick whose
 * field is zero adds no dipole force (`norm(B) ≈ 0`).  lbm_run and
 * lbm_batch_run refuse (LBM_ESTATE) a run with such terms that reaches a
 * tick outside the schedule.  Replaces the previous schedule; count = 0
 * clears it.  Waits for the handle's stream. */
int lbm_set_field_schedule(lbm_handle *h, const double *B, int64_t first_tick, int64_t count);

/* ---- per-tick particle log (saveData's particlesTrj.csv,
 * gpu/fluids.jl:252-259, gpu/aux.jl:319-362) ----
 * With capacity > 0 every tick's particle move also writes each particle's
 * position, velocity and angular velocity after the move into a device
 * ring of `capacity` ticks; 0 turns the log off and frees the ring.  Ticks
 * run before the log was on are not logged. */
int lbm_set_trajectory_log(lbm_handle *h, int64_t capacity);

/* Logged states of particle id for ticks first .. first + count - 1, as
 * out[count][9] = (position[3], velocity[3], angular velocity[3]) with
 * unused components 0.  Every tick must be one the ring still holds: later
 * than the particle's addition and than the log's start, at most the
 * current tick, and within the last `capacity` ticks.  Waits for the
 * handle's stream. */
int lbm_get_trajectory(lbm_handle *h, int32_t id, int64_t first, int64_t count, double *out);

/* writeTrajectories (gpu/aux.jl:271-317) as text: fluidTrj_<tick>.csv with
 * columns tick, time, id_x, id_y[, id_z], coordinate_*, massDensity,
 * fluidVelocity_*, f1..fQ, one row per node in column-major order; id_* are
 * 1-based indices.  rho is (N...), u (N..., dims) and f (N..., Q), column
 * major, FP64 host arrays.  Host code only: no handle or device.  Floats
 * are written as the shortest decimal that reads back to the same double. */
int lbm_write_fluid_csv(const char *path, int32_t dims, const int64_t n[3], int64_t tick, double time,
                        const double *const coords[3], const double *rho, const double *u, int32_t q,
                        const double *f);

/* writeTensor (gpu/aux.jl:364-400) as text: <name>.csv with columns tick,
 * time, id_*, coordinate_*, then component_ab for the dims x dims tensor t
 * in column-major order (component_xx, component_yx, ...); t is
 * (N..., dims, dims), column major, FP64 host.  Host code only. */
int lbm_write_tensor_csv(const char *path, int32_t dims, const int64_t n[3], int64_t tick, double time,
                         const double *const coords[3], const double *t);

/* viscousStressTensor (cpu/fluids.jl:507-520) of the current populations:
 * sigma_ab = -(1 - dt / (2 tau)) sum_i c_ia c_ib (f_i - feq_i(rho, u)), with
 * rho and u of hydroVariablesUpdate! and the integer velocities c_i;
 * evaluated on the device in FP64 whatever the storage.  sigma is
 * (N..., dims, dims), column major, FP64 host. */
int lbm_viscous_stress(lbm_handle *h, double *sigma);

/* Health check of the current populations (no counterpart in the
 * reference, which checks nothing while it runs): *nonfinite receives the
 * number of NaN or infinite populations of fluid nodes, *saturated the
 * number held at the storage's largest magnitude (F16C's clamp: a
 * deviation beyond |f - w rho0| >= w/8, outside the storage's range; 0 for
 * the other storages).  One read of the populations; waits for the stream.
 * Either pointer may be NULL. */
int lbm_check_populations(lbm_handle *h, int64_t *nonfinite, int64_t *saturated);

/* Boundary links of particle id in the reference's findall order over its
 * streamingInvasionRegions (gpu/aux.jl:187-198, gpu/fluids.jl:181): link
 * (node, j) means node is fluid and node + c_j is solid.  Writes at most cap
 * entries; *count receives the total. */
int lbm_get_links(lbm_handle *h, int32_t id, int64_t *node, int32_t *dir,
                  int64_t cap, int64_t *count);

/* LBMpropagate!(model; ticks = nsteps) (gpu/fluids.jl:228-264): nsteps calls
 * of tick! (gpu/fluids.jl:132-226).  Asynchronous on the handle's stream:
 * one collide kernel, nsteps - 1 fused gather+collide kernels and one gather
 * kernel (plus one small kernel applying the last particle move), so the
 * populations before and after are the reference's distributions. */
int lbm_run(lbm_handle *h, int64_t nsteps);

/* lbm_run without any allocation, for callers that cannot allocate device
 * memory at call time (code under HIP graph capture, tight serving loops,
 * harnesses that forbid hipMalloc).  No counterpart in the reference.
 * The same kernels as lbm_run, launched directly (no graph is captured, no
 * timing event recorded), with the particles' momentum-exchange partials and
 * link-slot values in a fixed 8 MiB device array that the library holds on
 * each device it runs on until the process ends.  Returns LBM_EINVAL, before
 * launching anything, when the particles of this block shape need more.
 * Every call on one device shares that array, so calls that overlap on one
 * device (several host threads, or handles on different streams) race.
 * The result is bit-identical to lbm_run's. */
int lbm_run_noalloc(lbm_handle *h, int64_t nsteps);

/* Wait for the handle's stream; reports a particle that left the supported
 * region (LBM_ERANGE) or a HIP error. */
int lbm_sync(lbm_handle *h);

/* model.tick and model.time. */
int lbm_get_tick(lbm_handle *h, int64_t *tick, double *time);

/* State snapshots, to resume a run: lbm_save_state writes the populations in
 * the storage's own encoding, every particle's state record (kinematics, the
 * node-velocity snapshot, the move count) and the tick and time into `buf`
 * (lbm_state_size bytes); lbm_load_state puts them into a handle created
 * with the same configuration and particles, which then runs on bit for bit
 * as the saved one would have (LBM_EINVAL for another lattice, storage or
 * particle set; the per-tick log then refuses the ticks before the restored
 * one).  No counterpart in the reference, which cannot resume a run. */
int lbm_state_size(lbm_handle *h, int64_t *bytes);
int lbm_save_state(lbm_handle *h, void *buf, int64_t bytes);
int lbm_load_state(lbm_handle *h, const void *buf, int64_t bytes);

/* Block shape of the step kernel (threads along x, y, z; product 64..512, a
 * multiple of 64), normally chosen by the autotuner; 0 resets to the
 * default. */
int lbm_set_block(lbm_handle *h, int32_t bx, int32_t by, int32_t bz);

/* Current block shape of the step kernel. */
int lbm_get_block(lbm_handle *h, int32_t *bx, int32_t *by, int32_t *bz);

/* Order of the step kernel's tiles over its grid (normally chosen by the
 * autotuner; lbm_set_block resets it to 1): 1 gives each of the 8 XCDs one
 * contiguous range of tiles, 0 the plain order.  No counterpart in the
 * reference. */
int lbm_set_tile_order(lbm_handle *h, int32_t xcd_remap);
int lbm_get_tile_order(lbm_handle *h, int32_t *xcd_remap);

/* Occupancy variant of the gathering kernels (no reference counterpart; a
 * tuning knob like the block shape): 0 (default) builds them for the most
 * waves per SIMD the tiles use (FP32/FP16: 8, FP64: 4), capping registers so
 * that the particles' box path spills; 1 for half the waves and twice the
 * registers, which pays on lattices up to about 10^6 nodes with particles,
 * where the box path's dependent chain sets the kernel time, and for FP64
 * with a force, whose collision spills under the default cap.  lbm_autotune
 * times both. */
int lbm_set_occupancy(lbm_handle *h, int32_t low);
int lbm_get_occupancy(lbm_handle *h, int32_t *low);

/* Autotuner: times the fused step kernel over candidate block shapes, tile
 * orders and both occupancy builds, `reps` kernels each, then the best few
 * again three times with more kernels, and keeps the lowest median; returns
 * it with its time per launch.  The trials run as a run does (a
 * collide-only kernel, then fused kernels that move the particles) on a
 * copy of the populations and records, so the state is untouched.
 * No counterpart in the reference (CUDA.jl launches 8x8x8 blocks,
 * gpu/aux.jl:12-21). */
int lbm_autotune(lbm_handle *h, int32_t reps, int32_t best[3], double *best_ms);

/* lbm_run replays captured hipGraphs of `ticks_per_graph` fused kernels
 * (even) and launches the remainder directly; 0 launches every kernel
 * directly.  Default 32.  The result is the same either way. */
int lbm_set_graph(lbm_handle *h, int32_t ticks_per_graph);

/* Measurement (no counterpart in the reference, which has no timing):
 * with timing on, lbm_run records HIP events around every fused step kernel
 * on the handle's stream, waits for the run and keeps the averages. */
int lbm_set_timing(lbm_handle *h, int on);

/* Average device time per step kernel over the last timed lbm_run, in ms
 * (0 if no timed run happened). */
int lbm_last_run_ms(lbm_handle *h, double *ms);

/* Averages of the last timed lbm_run: fused step kernel time and whole run
 * time per tick, in ms.  Either pointer may be NULL. */
int lbm_last_run_times(lbm_handle *h, double *kernel_ms, double *step_ms);

/* Measurement: average time (ms) of a plain device-to-device copy of `bytes`
 * bytes (a multiple of 16), over `reps` copies — the achievable HBM
 * bandwidth beside the step, and a kernel of known traffic for scaling
 * profiler byte counters.  No counterpart in the reference. */
int lbm_measure_copy(int64_t bytes, int32_t reps, double *ms);

/* Measurement: average time (ms) of the D3Q19 step's data movement alone on
 * a periodic nx x ny x nz lattice of the given storage, over `reps` launches:
 * every node pulls its 19 populations f_q(x - c_q) from the row-blocked
 * layout and stores them unchanged, one node per lane, as the step kernels'
 * tiles do, with no collision — the ceiling of the step's access pattern.
 * No counterpart in the reference. */
int lbm_measure_pull(int32_t storage, int32_t nx, int32_t ny, int32_t nz, int32_t reps, double *ms);

/* ------------------------------------------------------------ batches
 * A batch advances up to 256 independent boxes (handles made by lbm_create,
 * of one lattice, storage, collision and force kind; any sizes, parameters,
 * walls and particles) together: each step kernel of a run is one launch
 * over every box, and the run's fused kernels replay from captured
 * hipGraphs.  Each box ends exactly where lbm_run(box, nsteps) would leave
 * it (same kernels, same arithmetic, bit for bit).
 * No counterpart in the reference, which advances one LBMmodel per
 * LBMpropagate! call (gpu/fluids.jl:228-264); a sweep there is a loop over
 * models. */
typedef struct lbm_batch lbm_batch;

/* The batch borrows the handles (they must outlive it) and gives every box
 * its block shape (default: box 0's, if it fits every box). */
int lbm_batch_create(lbm_handle *const *boxes, int32_t n, lbm_batch **out);
int lbm_batch_destroy(lbm_batch *b);

/* LBMpropagate!(box; ticks = nsteps) for every box of the batch, as one
 * sequence of batched kernels on the batch's stream.  Each box's own stream
 * waits for it, so later calls on a box see the result. */
int lbm_batch_run(lbm_batch *b, int64_t nsteps);

/* Wait for the batch; reports LBM_ERANGE for a particle that left its box's
 * supported region, as lbm_sync does. */
int lbm_batch_sync(lbm_batch *b);

/* Block shape shared by the boxes (the rules of lbm_set_block) and the
 * order of the tiles (the rules of lbm_set_tile_order). */
int lbm_batch_set_block(lbm_batch *b, int32_t bx, int32_t by, int32_t bz);
int lbm_batch_get_block(lbm_batch *b, int32_t *bx, int32_t *by, int32_t *bz);
/* Row tiles instead of one shared shape: blocks of `threads` threads (64..512,
 * a multiple of 64), and every box its own tile shape of that many threads
 * whose x extent holds a whole row of the box (the least divisor of
 * `threads` of at least 16 that does, or `threads`), so that a row's x
 * neighbours lie in lines its own tile reads.  get returns 0 while the
 * boxes share one shape (lbm_batch_set_block). */
int lbm_batch_set_row_tiles(lbm_batch *b, int32_t threads);
int lbm_batch_get_row_tiles(lbm_batch *b, int32_t *threads);
int lbm_batch_set_tile_order(lbm_batch *b, int32_t xcd_remap);
int lbm_batch_get_tile_order(lbm_batch *b, int32_t *xcd_remap);
int lbm_batch_set_occupancy(lbm_batch *b, int32_t low);  /* as lbm_set_occupancy */
int lbm_batch_get_occupancy(lbm_batch *b, int32_t *low);

/* Autotuner of the batched fused kernel over the block shapes that fit every
 * box, row tiles of 64..512 threads, both tile orders and both occupancy
 * builds, ranked as lbm_autotune ranks (the boxes' state is untouched).
 * best: the block shape, or {threads, 0, 0} when row tiles are fastest. */
int lbm_batch_autotune(lbm_batch *b, int32_t reps, int32_t best[3], double *best_ms);

/* Ticks per captured graph (0 or even; default 32), timing and the averages
 * of the last timed run, as for one handle. */
int lbm_batch_set_graph(lbm_batch *b, int32_t ticks_per_graph);
int lbm_batch_set_timing(lbm_batch *b, int on);
int lbm_batch_last_run_times(lbm_batch *b, double *kernel_ms, double *step_ms);

/* ------------------------------------------------------------ adjoint
 * Gradient of a swimming-speed loss by the discrete adjoint of the step
 * (D3Q19, GPU semantics; any storage, collision and force kind):
 *
 *   L = (mean over ticks T-window+1..T of |V_swimmer| - target_speed)^2
 *       + force_penalty * sum F^2
 *
 * with F the model's Guo force field (zero at wall nodes, where the step
 * zeroes it: gpu/initMethods.jl:636-654; dL/dF is 0 there)
 * from the handle's current state over `ticks` ticks, with respect to every
 * particle's B1 and B2 and the Guo force field.  The forward run saves the
 * populations and particle records every `checkpoint_every` ticks; the
 * reverse sweep recomputes each segment, keeping its states, and runs the
 * adjoint tick kernel backwards through it.  The link set is a
 * piecewise-constant function of the positions and carries no derivative.
 * No counterpart in the reference (whose step is not differentiable:
 * gpu/fluids.jl:132-226 runs on Float64 arrays in place). */
typedef struct lbm_adjoint_config {
  int64_t ticks;            /* forward ticks T                                   */
  int64_t checkpoint_every; /* K: states saved every K ticks                     */
  int64_t window;           /* the speed is averaged over the last `window` ticks */
  int32_t particle;         /* id of the swimmer                                 */
  int32_t reserved;         /* 0                                                 */
  double target_speed;
  double force_penalty;
} lbm_adjoint_config;

/* Runs the forward ticks (the handle ends ticks ahead, as after lbm_run) and
 * the reverse sweep.  Outputs (each may be NULL): *loss; dB1, dB2 (one per
 * particle, in id order); dF ((N, 3), Julia layout; requires a force field).
 * Device memory: about ticks/K + 2.5 K + 4 population buffers (two segment
 * sets of every tick's post-collision populations and moments). */
int lbm_adjoint_run(lbm_handle *h, const lbm_adjoint_config *cfg, double *loss, double *dB1, double *dB2,
                    double *dF);

/* Wall times (ms) of the last lbm_adjoint_run: forward, reverse (with the
 * recomputation). */
int lbm_adjoint_last_times(lbm_handle *h, double *forward_ms, double *reverse_ms);

#ifdef __cplusplus
}
#endif

#endif /* LBM_HIP_H */
