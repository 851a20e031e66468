/*
 * ss_api.h -- C ABI of the MI355X-native per-timestep hot loop of
 * stochasticsampling (active-suspension Smoluchowski / Stokes solver).
 *
 * Every entry point takes plain pointers and sizes.  Host buffers are owned by
 * the caller; device memory is owned by the library.  Return codes: SS_OK (0)
 * on success, a negative SS_ERR_* on failure; ss_last_error() then returns a
 * thread-local message.  One handle per host thread; handles are independent
 * and the library takes no internal locks.
 *
 * Reference citations name files of clamydo/stochasticsampling v2.1.1.
 */
#ifndef SS_API_H
#define SS_API_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define SS_API_VERSION 4

/* Return codes. */
#define SS_OK 0
#define SS_ERR_INVALID (-1)     /* bad argument or parameter combination */
#define SS_ERR_DEVICE (-2)      /* HIP runtime error or no GPU           */
#define SS_ERR_STATE (-3)       /* call out of order (e.g. step before init) */
#define SS_ERR_UNSUPPORTED (-4) /* size or mode outside what is built     */
#define SS_ERR_ALLOC (-5)       /* device or host allocation failed       */

/* Arithmetic of particle states and FFTs. */
#define SS_FP64 0 /* plain FP64 throughout                                  */
#define SS_FP32 1 /* FP32 states with compensated positions, FP32 FFT       */

/* Source of the per-particle Gaussian / Rayleigh increments. */
#define SS_NOISE_PHILOX 0   /* counter-based Philox4x32-10 keyed by (seed, step, particle id) */
#define SS_NOISE_EXTERNAL 1 /* caller-supplied RandomVector buffer, see ss_set_noise */
#define SS_NOISE_REFERENCE 2 /* the reference's own stream: noise_streams Pcg32 generators
                                seeded alike by seed_from_u64(seed), chunk c of N/T + 1
                                particles drawn from generator c (rand_distr ziggurat
                                normals, rand uniforms), src/bin/simulation/mod.rs:111-128,
                                259-286 */

/* Binning strategy for particle -> (x,y,z,phi,theta) histogram. */
#define SS_BIN_AUTO 0      /* SORTED for FP32, PLAIN_TILED for FP64 (each falls back to
                              HISTOGRAM where it does not fit)                        */
#define SS_BIN_HISTOGRAM 1 /* plain: full 5-D u32 histogram by global atomics           */
#define SS_BIN_SORTED 2    /* particles grouped by spatial tile; binning fused into the
                              update with LDS-privatised fixed-point moments          */
#define SS_BIN_PLAIN_TILED 3 /* plain path (separate binning, solve and update) with the
                              particles kept in tile order: per-tile LDS counts stored
                              with plain stores, a counting sort every resort_interval
                              steps after the update                                  */

/* FFT variant used by the spectral solve. */
#define SS_FFT_AUTO 0  /* autotuned among the variants that pass the accuracy gate */
#define SS_FFT_RADIX 1 /* LDS-staged mixed-radix Stockham                          */
#define SS_FFT_DFT 2   /* dense DFT matrix product on MFMA (3-part split bf16) */

/* Flags. */
#define SS_FLAG_QUASI2D 1u            /* z = 0, theta = pi/2 after every step (feature "quasi2d") */
#define SS_FLAG_MAGNETIC_ALWAYS 2u    /* run the magnetic solve even when its coefficients are 0  */
#define SS_FLAG_KEEP_DISTRIBUTION 4u  /* keep the full count grid of every step's binning         */
#define SS_FLAG_NO_GRAPH 8u           /* enqueue every step eagerly instead of replaying captured
                                         hipGraphs of whole re-sort periods                       */
#define SS_FLAG_AXIS_PASSES 16u       /* tiled path: run the spectral stage as separate kernels
                                         (convert, one pass per FFT axis, solve, pack) instead of
                                         the three fused plane/pencil kernels (same results)      */
#define SS_FLAG_PAIRED_SPECTRAL 32u   /* fused spectral stage with two real fields per complex
                                         transform (pencils solved with their mirrors) instead of
                                         one transform per component: half the transforms, the same
                                         fields to rounding; slower at 64^3 on MI355X             */

/* Size of the simulation box (lib.rs BoxSize). */
typedef struct ss_box {
    double x, y, z;
} ss_box;

/* Size of the discrete grid (lib.rs GridSize). */
typedef struct ss_grid {
    int32_t x, y, z, phi, theta;
} ss_grid;

/*
 * Everything Simulation::new reads (src/bin/simulation/mod.rs:91-160): the
 * SimulationSettings fields used by the step (settings/mod.rs:108-118) and all
 * Parameters fields (settings/mod.rs:51-67), plus the build's own knobs.
 */
typedef struct ss_params {
    ss_box box_size;
    ss_grid grid_size;
    uint64_t number_of_particles;
    double timestep;
    uint64_t seed;
    /* Parameters */
    double magnetic_drag;
    double shape;
    double hydro_screening;
    double volume_exclusion;
    double magnetic_reorientation;
    double diffusion_translational;
    double diffusion_rotational;
    double stress_active;
    double stress_magnetic;
    double magnetic_dipole_dipole;
    /* Direction of the external field; (0,0,0) selects the reference's y axis
       (modifiers.rs:108-110, stress.rs:121-141). */
    double field_direction[3];
    /* Build knobs. */
    int32_t precision;   /* SS_FP64 | SS_FP32          */
    int32_t noise_mode;  /* SS_NOISE_*                 */
    int32_t binning;     /* SS_BIN_*                   */
    int32_t fft_variant; /* SS_FFT_*                   */
    uint32_t flags;      /* SS_FLAG_*                  */
    int32_t device;      /* HIP device ordinal          */
    /* Tiled path (SS_BIN_SORTED): tile extent in cells per axis (0: chosen
       automatically) and steps between re-sorts of the particles by tile
       (0: 16).  SS_BIN_PLAIN_TILED uses resort_interval alone. */
    int32_t tile[3];
    int32_t resort_interval;
    /* SS_NOISE_REFERENCE: number of generator streams, the reference run's
       RAYON_NUM_THREADS (mod.rs:113-128); required (>= 1) in that mode. */
    int32_t noise_streams;
} ss_params;

/* Fills p with the reference's defaults: zero optional parameters, field along y,
   FP64, Philox noise, automatic binning and FFT, device 0. */
void ss_params_default(ss_params* p);

/* Thread-local message of the last failing call ("" if none). */
const char* ss_last_error(void);

/* SS_API_VERSION of the loaded library. */
int ss_version(void);

/* Number of visible GPUs (0 without a GPU; never an error in that case). */
int ss_device_count(int* count);

/* ---------------------------------------------------------------------------
 * Simulation handle -- replaces struct Simulation (src/bin/simulation/mod.rs:60-67).
 * ------------------------------------------------------------------------- */
typedef struct ss_sim ss_sim;

/* Simulation::new (src/bin/simulation/mod.rs:91-160).  All planning and device
   allocation happens here; ss_step allocates nothing. */
int ss_create(const ss_params* params, ss_sim** out);

/* Drop for Simulation (src/bin/simulation/mod.rs:377-381). */
int ss_destroy(ss_sim* sim);

/* Simulation::init (src/bin/simulation/mod.rs:163-196).  particles is N x 5
   doubles in serde order [x, y, z, phi, theta] (particle.rs:363-395).  Every
   particle is sanitised like Particle::new (particle.rs:186-201) and the
   distribution is sampled once.  n must equal number_of_particles. */
int ss_init(ss_sim* sim, const double* particles, uint64_t n);

/* Replaces the noise draw of do_timestep (src/bin/simulation/mod.rs:259-286)
   for noise_mode SS_NOISE_EXTERNAL: rv is N x 5 doubles per particle, in the
   order of RandomVector (mod.rs:51-57): [x, y, z, axis_angle, rotate_angle],
   already scaled like the reference (rotate_angle = dr*sqrt(-2 ln(1-U))).
   Every following step uses this buffer until it is replaced. */
int ss_set_noise(ss_sim* sim, const double* rv, uint64_t n);

/* Simulation::do_timestep / Iterator::next (src/bin/simulation/mod.rs:251-393),
   nsteps times.  timestep_out (may be NULL) receives the counter afterwards. */
int ss_step(ss_sim* sim, uint64_t nsteps, uint64_t* timestep_out);

/* Simulation::get_timestep (src/bin/simulation/mod.rs:246-248). */
int ss_get_timestep(const ss_sim* sim, uint64_t* timestep);

/* Simulation::get_particles / get_particles_head (mod.rs:221-228): the first
   n particles (n <= N) as n x 5 doubles, in the order they were given to init. */
int ss_get_particles(ss_sim* sim, double* out, uint64_t n);

/* Simulation::get_distribution (mod.rs:231-233): the distribution sampled at
   the start of the last step (or by init).  counts (may be NULL) receives the
   integer bin counts [x][y][z][phi][theta]; dist (may be NULL) the normalised
   values exactly as Distribution::sample_from times the box volume
   (distribution/mod.rs:164-184, mod.rs:253-257).  With SS_BIN_SORTED this
   needs SS_FLAG_KEEP_DISTRIBUTION. */
int ss_get_distribution(ss_sim* sim, uint32_t* counts, double* dist);

/* Simulation::get_flow_field (mod.rs:236-238) and the gradient the step used:
   u is [3][nx][ny][nz], grad_u (may be NULL) is [3][3][nx][ny][nz] with
   grad_u[a][b] = d_a u_b (spectral_solver.rs:192-229), real parts. */
int ss_get_flow_field(ss_sim* sim, double* u, double* grad_u);

/* Simulation::get_magnetic_field (mod.rs:241-243) and its gradient
   (magnetic_solver.rs:139-169), real parts, same layouts as above. */
int ss_get_magnetic_field(ss_sim* sim, double* b, double* grad_b);

/* Simulation::get_snapshot (mod.rs:210-217).  particles: N x 5; compensation
   (may be NULL): N x 3 low-order position parts of the FP32 path (zeros in
   FP64).  The Philox noise state is (seed, timestep); SS_NOISE_REFERENCE
   adds the generators of ss_get_rng_state. */
int ss_get_snapshot(ss_sim* sim, double* particles, double* compensation, uint64_t* timestep);

/* Simulation::resume (mod.rs:199-207): init from the snapshot, restore the
   compensation (may be NULL) and the timestep counter. */
int ss_resume(ss_sim* sim, const double* particles, const double* compensation, uint64_t n,
              uint64_t timestep);

/* Snapshot::rng_state (mod.rs:80-86, 199-217) for SS_NOISE_REFERENCE: the
   noise_streams generators as (state, increment) pairs, the serde fields of
   rand_pcg's Lcg64Xsh32.  ss_get_rng_state writes 2 * noise_streams values
   (count >= noise_streams, else SS_ERR_INVALID) and returns noise_streams;
   ss_set_rng_state replaces the first min(count, noise_streams) generators,
   as Simulation::resume's zip does (call it after ss_resume).  Other noise
   modes have no generator state: get returns 0, set accepts count 0 only. */
int ss_get_rng_state(ss_sim* sim, uint64_t* state_increment, int32_t count);
int ss_set_rng_state(ss_sim* sim, const uint64_t* state_increment, int32_t count);

/* The RandomVector array of the last step (N x 5 doubles: x, y, z,
   axis_angle, rotate_angle) for SS_NOISE_EXTERNAL and SS_NOISE_REFERENCE
   (mod.rs:73, 259-286); SS_ERR_STATE before the first step. */
int ss_get_noise(ss_sim* sim, double* rv, uint64_t n);

/* Per-stage device time of the last ss_profile_steps call, measured with HIP
   events on the simulation's stream.  Stage names: "bin", "fft_forward",
   "solve", "fft_inverse", "pack", "noise" (SS_NOISE_REFERENCE draws),
   "update", "sort".  Returns the number of
   stages written (at most max_stages); ms_per_step[i] is the mean per step. */
int ss_profile_steps(ss_sim* sim, uint64_t nsteps, double* ms_per_step, char* names,
                     int name_stride, int max_stages);

/* Starts a new re-sort period: the next step's update writes the particles
   in tile order (offsets from the counts of the last binning), as every
   resort_interval-th step does.  SS_BIN_PLAIN_TILED sorts the particles
   at once (no effect when nothing moved since the last sort).  No effect on
   SS_BIN_HISTOGRAM or when a re-sort is already due; results do not depend on
   it (both tiled binnings are exact).  No reference counterpart (the reference keeps particles unsorted).
   SS_ERR_STATE for a member of an ensemble. */
int ss_resort(ss_sim* sim);

/* Captures the graphs that the next ss_step(sim, nsteps) will replay (whole
   re-sort periods, and the runs of steps between period boundaries that
   ss_step otherwise captures only when a run of the same phase and length
   is asked for a second time), without running anything or changing the
   simulation's state: the set-up of a run of steps, as an FFT plan is the
   set-up of its transforms.  Optional; the tiled path only (no effect with
   the kept distribution, SS_NOISE_REFERENCE or elsewhere).  No reference
   counterpart.  SS_ERR_STATE before ss_init or for a member of an ensemble. */
int ss_prepare_steps(ss_sim* sim, uint64_t nsteps);

/* Keep (on != 0) or stop keeping the full count grid of every step's
   binning, as SS_FLAG_KEEP_DISTRIBUTION does from ss_create, so that
   ss_get_distribution (Simulation::get_distribution, mod.rs:231-233) can be
   read after every step on the tiled path too.  Turning it on allocates the
   two count grids (kept until ss_destroy) and bins the current particles at
   once, so ss_get_distribution is valid immediately (the distribution the
   next step samples) and after every following step; while it is on, steps
   are enqueued eagerly (no captured graphs) and the update adds the count
   atomics.  No effect on the plain path, which always keeps the grid.
   SS_ERR_UNSUPPORTED for a member of an ensemble. */
int ss_keep_distribution(ss_sim* sim, int on);

/* Global diagnostics of the current particles, reduced in FP64 (exact
   64-bit fixed-point sums, independent of the particle order), 15 doubles:
   out[0..2]  mean orientation vector n over the particles;
   out[3..5]  mean bin-centre orientation = the sum over cells of the mean
              polarisation field of tools/DataStreamer.py:177-196;
   out[6..14] mean stress [3][3] = (1/Ns) sum over cells of average_stress
              (stress.rs:60-96) for the current binning.
   No reference counterpart inside the step (the reference computes these in
   Python post-processing). */
int ss_get_diagnostics(ss_sim* sim, double* out);

/* Wait for all work of the handle. */
int ss_synchronize(ss_sim* sim);

/* Execution plan chosen at ss_create (no reference counterpart; for tests and
   reports). */
typedef struct ss_info {
    int32_t sorted;          /* tiled binning path (particles grouped by tile) */
    int32_t fused_spectral;  /* spectral stage as the three fused kernels: 1 one
                                transform per component, 2 two fields per transform */
    int32_t tile[3];         /* tile extent in cells (tiled path)              */
    int32_t ntiles;
    int32_t resort_interval; /* steps between re-sorts (tiled path)            */
    int32_t pencils;         /* (y, z) pencils per workgroup (fused spectral)  */
    int32_t graphs;          /* whole periods replayed from captured hipGraphs */
    int32_t update_groups;   /* workgroups of the tile update (one per tile)   */
    int32_t plain_tiled;     /* plain path with tile-ordered particles
                                (SS_BIN_PLAIN_TILED); tile, ntiles and
                                resort_interval describe its tiles       */
} ss_info;
int ss_get_info(const ss_sim* sim, ss_info* out);

/* ---------------------------------------------------------------------------
 * Batched ensemble -- no reference counterpart (the reference runs one system
 * per process).  For sweeps of many independent systems that differ in grid,
 * box, particle count and parameters.  Members share precision, noise mode,
 * moment layout (magnetic coefficients zero or not, magnetic stress zero or
 * not) and the re-sort interval, and always use the tiled path.  Every step
 * runs the members' spectral solves as parallel branches and one persistent
 * update launch over all members' tiles; whole re-sort periods replay from a
 * captured hipGraph.
 * ------------------------------------------------------------------------- */
typedef struct ss_ensemble ss_ensemble;

/* params: nsys entries, one per member (1 <= nsys <= 256).  SS_ERR_UNSUPPORTED
 * when the largest member tile window plus 4 (nsys + 1) bytes exceed the LDS
 * of a workgroup (160 KiB on gfx950). */
int ss_ensemble_create(const ss_params* params, int nsys, ss_ensemble** out);
int ss_ensemble_destroy(ss_ensemble* ens);

/* Borrowed handle of member `sys` for ss_init / ss_set_noise / getters
   (ss_step, ss_profile_steps and ss_resort return SS_ERR_STATE on it; do not
   ss_destroy it).  All members must be initialised together
   before ss_ensemble_step. */
int ss_ensemble_member(ss_ensemble* ens, int sys, ss_sim** out);

/* nsteps steps of every member; timestep_out (may be NULL) receives the
   members' common step counter. */
int ss_ensemble_step(ss_ensemble* ens, uint64_t nsteps, uint64_t* timestep_out);

/* nsteps eager ensemble steps timed with HIP events on the ensemble's
   stream, like ss_profile_steps: stages "spectral" (all members' spectral
   stages, from the fork to the join of the side streams), "update" (the
   persistent update launch over every member's tiles) and "sort" (the
   offsets of a due re-sort).  Returns the number of stages written. */
int ss_ensemble_profile_steps(ss_ensemble* ens, uint64_t nsteps, double* ms_per_step, char* names,
                              int name_stride, int max_stages);

/* ---------------------------------------------------------------------------
 * Initial conditions (no reference counterpart for the perturbed state; the
 * uniform part mirrors Particle::create_isotropic, particle.rs:224-256).
 * Positions uniform in the box; cos(theta) drawn by exact inverse CDF from
 * 1 + amplitude*sin(2 pi x / Lx)*cos(theta); phi uniform.  Uniforms come
 * from Philox4x32-10 with key (seed) and counter (index, 0xC0FFEE, 0, 0).
 * ------------------------------------------------------------------------- */
int ss_initial_condition(const ss_params* params, double amplitude, uint64_t seed,
                         double* particles, uint64_t n);

/* Particle::create_isotropic (src/particle.rs:224-255) and
   Particle::create_homogeneous (src/particle.rs:257-297), drawn on `device`
   from the reference's stream: Pcg64Mcg::seed_from_u64(seed), five
   Uniform(0, 1) values per particle in order.  particles: n x 5 doubles on
   the host.  create_homogeneous returns SS_ERR_INVALID with the reference's
   panic message for kappa == 0 or an angle that is NaN. */
int ss_create_isotropic(uint64_t n, const ss_box* box, uint64_t seed, int32_t device, double* particles);
int ss_create_homogeneous(uint64_t n, double kappa, const ss_box* box, uint64_t seed, int32_t device,
                          double* particles);

/* DensityGradient::get_gradient + get_real_gradient
   (src/distribution/density_gradient.rs:22-110): the density of a
   normalised distribution ([x][y][z][phi][theta] doubles, like ss_flow_solve)
   summed over the angles times dtheta dphi, its gradient by FFT; grad: 3 x
   Ns real parts.  Not used by the step (the reference comments it out). */
int ss_density_gradient(const ss_params* params, const double* dist, double* grad);

/* interpolate_vector_field (src/mesh/interpolate.rs:11-123): trilinear
   interpolation of a field [3][nx][ny][nz] (cell-centred, periodic) at n
   positions (n x 3) into out (n x 3), in FP64.  The reference takes the z
   offset from position.y; this entry point takes it from position.z.  Not
   used by the step. */
int ss_interpolate_vector_field(const ss_params* params, const double* field, const double* positions, uint64_t n,
                                double* out);

/* ---------------------------------------------------------------------------
 * FFT plan -- replaces FFTPlan (crates/fftw3/src/fft.rs:33-244): unnormalised,
 * in place, forward sign -1, backward +1, batched over `batch` contiguous
 * [nx][ny][nz] complex arrays (interleaved re, im).
 * ------------------------------------------------------------------------- */
typedef struct ss_fft_plan ss_fft_plan;

/* FFTPlan::new_c2c_inplace_3d (fft.rs:156-178).  direction: -1 forward, +1
   backward. */
int ss_fft_plan_create(int nx, int ny, int nz, int direction, int batch, int precision,
                       int variant, int device, ss_fft_plan** out);

/* FFTPlan::reexecute3d (fft.rs:226-231) on device memory, enqueued on the given
   hipStream_t (NULL: the plan's own stream; the call then waits for it). */
int ss_fft_execute(ss_fft_plan* plan, void* device_data, void* hip_stream);

/* Same on host memory: copies in, executes, copies out. */
int ss_fft_execute_host(ss_fft_plan* plan, void* host_data);

/* Largest plan batch (nx*ny*nz*batch complex values, in bytes) that
   ss_fft_execute_host_noalloc stages: 64 MiB. */
#define SS_FFT_NOALLOC_MAX_BYTES (64u << 20)

/* No counterpart in the reference.  ss_fft_execute_host without a device
   allocation, for callers that cannot allocate (tight loops, harnesses that
   forbid hipMalloc): the plan's batch is staged in a buffer the library holds
   on each GPU until the process ends (SS_FFT_NOALLOC_MAX_BYTES; a larger
   batch returns SS_ERR_UNSUPPORTED).  All calls on one GPU share that buffer:
   calls that overlap on one GPU (other host threads) race. */
int ss_fft_execute_host_noalloc(ss_fft_plan* plan, void* host_data);

/* Drop for FFTPlan (fft.rs:235-244). */
int ss_fft_plan_destroy(ss_fft_plan* plan);

/* ---------------------------------------------------------------------------
 * Solver entry points for unit-level parity.
 * ------------------------------------------------------------------------- */

/* Reusable solver handles -- SpectralSolver::new (spectral_solver.rs:33-68)
   and MagneticSolver::new (magnetic_solver.rs:32-62): plan and allocate once
   for params' grid, box, stress prefactors, shape and precision; every solve
   then reuses them (no allocation per call).  One handle per host thread. */
typedef struct ss_solver ss_solver;
#define SS_SOLVER_FLOW 0
#define SS_SOLVER_MAGNETIC 1
int ss_solver_create(const ss_params* params, int kind, ss_solver** out);
int ss_solver_destroy(ss_solver* solver);

/* SpectralSolver::mean_flow_field(hydro_screening, dist) + get_real_flow_field
   (spectral_solver.rs:233-258) on a distribution [nx][ny][nz][phi][theta]
   (already multiplied by the box volume, like the caller in
   src/bin/simulation/mod.rs:253-257): real parts of u [3][..] and grad_u
   [3][3][..] (may be NULL).  SS_SOLVER_FLOW handles only. */
int ss_solver_mean_flow_field(ss_solver* solver, double hydro_screening, const double* dist, double* u,
                              double* grad_u);

/* MagneticSolver::mean_magnetic_field (magnetic_solver.rs:139-161): real parts
   of B [3][..] and grad B [3][3][..] (may be NULL).  SS_SOLVER_MAGNETIC
   handles only. */
int ss_solver_mean_magnetic_field(ss_solver* solver, const double* dist, double* b, double* grad_b);

/* SpectralSolver::new + mean_flow_field (spectral_solver.rs:33-68,233-254) on a
   given distribution dist [nx][ny][nz][phi][theta] (already multiplied by the
   box volume, like the caller in src/bin/simulation/mod.rs:253-257).  Uses
   params' grid, box, stress prefactors, shape, hydro_screening and precision.
   Outputs real parts: u [3][nx][ny][nz], grad_u [3][3][nx][ny][nz] (may be NULL). */
int ss_flow_solve(const ss_params* params, const double* dist, double* u, double* grad_u);

/* MagneticSolver::mean_magnetic_field (magnetic_solver.rs:139-161) on a given
   distribution; outputs real parts of B [3][..] and grad B [3][3][..]. */
int ss_magnetic_solve(const ss_params* params, const double* dist, double* b, double* grad_b);

/* Number of doubles per particle in the fields argument of ss_debug_update. */
#define SS_DEBUG_FIELDS 34

/* The device particle update -- LangevinBuilder chain of
   src/bin/simulation/mod.rs:323-363 with modifiers.rs:39-231 -- applied to n
   independent particles with explicitly given cell values.  fields per
   particle: u[3], W[3][3], E[3][3], rho, B[3] (before the magnetic_dipole_dipole
   factor), gradB[3][3] (row-major, [a][b]); rv per particle: RandomVector
   [x,y,z,axis_angle,rotate_angle].  particles / out: n x 5.  Computed in
   params->precision. */
int ss_debug_update(const ss_params* params, uint64_t n, const double* particles,
                    const double* fields, const double* rv, double* out);

/* The device binning rule -- Distribution::coord_to_grid
   (distribution/mod.rs:86-141) -- for n particles (n x 5, converted to
   params->precision first); out receives n x 5 int32 grid coordinates. */
int ss_debug_coord_to_grid(const ss_params* params, uint64_t n, const double* particles,
                           int32_t* out);

#ifdef __cplusplus
}
#endif

#endif /* SS_API_H */
