/*
 * ss_oracle.h -- FP64 CPU restatement of stochasticsampling's per-timestep loop.
 *
 * TEST INFRASTRUCTURE ONLY.  Only tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py may load this library, and only as the checker
 * or as the timed CPU baseline.  The product (libss) never links or calls it.
 */
#ifndef SS_ORACLE_H
#define SS_ORACLE_H

#include <stdint.h>

#include "../include/ss_api.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct osim osim;

/* Simulation::new / init / do_timestep / getters (src/bin/simulation/mod.rs). */
osim* osim_create(const ss_params* p);
void osim_destroy(osim* s);
int osim_init(osim* s, const double* particles, uint64_t n);
int osim_set_noise(osim* s, const double* rv, uint64_t n);
int osim_step(osim* s, uint64_t nsteps);
uint64_t osim_get_timestep(const osim* s);
int osim_get_particles(const osim* s, double* out, uint64_t n);
int osim_get_counts(const osim* s, uint32_t* counts);
int osim_get_distribution(const osim* s, double* dist);
int osim_get_flow_field(const osim* s, double* u, double* grad_u);
int osim_get_magnetic_field(const osim* s, double* b, double* grad_b);
/* Last step's noise (N x 5), whichever source produced it. */
int osim_get_noise(const osim* s, double* rv);
/* SS_NOISE_REFERENCE streams as (state, increment) pairs; get returns the
   stream count (-1 if count is too small). */
int osim_get_rng_state(const osim* s, uint64_t* st, int count);
int osim_set_rng_state(osim* s, const uint64_t* st, int count);
int osim_set_timestep(osim* s, uint64_t t);
/* Number of OpenMP threads the parallel parts use (0: OpenMP default). */
void osim_set_threads(int n);

/* Pure functions. */
double o_modulo(double f, double m);
void o_ang_pbc(double phi, double theta, double* out2);
void o_particle_new(const double in[5], const ss_box* bs, double out[5]);
void o_orientation_from_vector(const double v[3], double out2[2]);
void o_coord_to_grid(const ss_params* p, const double part[5], int32_t out[5]);
void o_coord_to_grid_f32(const ss_params* p, const float part[5], int32_t out[5]);
void o_k_sampling(int n, double L, double* out);
void o_inverse_norm_squared(const ss_params* p, double* out);
void o_stress_kernel(const ss_params* p, double* out);
void o_orientation_kernel(const ss_params* p, double* out);
void o_linspace(double a, double b, int n, double* out);
void o_fft1d(int n, double* data, int sign);
void o_fft3d(int nx, int ny, int nz, double* data, int sign);
void o_histogram(const ss_params* p, const double* particles, uint64_t n, uint32_t* counts);
void o_sample_from(const ss_params* p, const uint32_t* counts, uint64_t n, double* dist);
int o_mean_flow_field(const ss_params* p, double screening, const double* dist, double* u,
                      double* grad_u);
int o_mean_magnetic_field(const ss_params* p, const double* dist, double* b, double* grad_b);
/* SpectralSolver::solve_flow_field (deprecated formulation, screening 0):
   u [3][Ns] real parts. */
int o_solve_flow_field(const ss_params* p, const double* dist, double* u);
void o_update_one(const ss_params* p, const double part[5], const double fields[34],
                  const double rv[5], double out[5]);
void o_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]);
/* DensityGradient (real parts, 3 x Ns) and trilinear interpolation. */
void o_density_gradient(const ss_params* p, const double* dist, double* grad);
double o_trilinear(const double rel[3], const double v[8]);
void o_neighbour_values(const double* f, int nx, int ny, int nz, const double pos[3], const double gw[3],
                        double v[8]);
void o_interpolate_vector_field(const double* field, int nx, int ny, int nz, const double pos[3],
                                const double gw[3], double out[3]);
void o_noise(uint64_t seed, uint64_t step, uint64_t id, double dr, double out[5]);
/* The reference's random streams (rand_pcg 0.2.0, rand_core 0.5.0, rand 0.7.0,
   rand_distr 0.2.1). */
typedef struct o_pcg32 {
    uint64_t state, inc;
} o_pcg32;
typedef struct o_mcg128 {
    uint64_t lo, hi;
} o_mcg128;
void o_pcg32_new(uint64_t state, uint64_t stream, o_pcg32* r);
void o_pcg32_from_seed(const uint8_t seed[16], o_pcg32* r);
void o_pcg32_seed_from_u64(uint64_t s, o_pcg32* r);
uint32_t o_pcg32_next_u32(o_pcg32* r);
uint64_t o_pcg32_next_u64(o_pcg32* r);
void o_mcg128_new(uint64_t lo, uint64_t hi, o_mcg128* r);
void o_mcg128_seed_from_u64(uint64_t s, o_mcg128* r);
uint64_t o_mcg128_next_u64(o_mcg128* r);
double o_uniform01(uint64_t bits);
double o_gen_f64(uint64_t bits);
double o_open01(uint64_t bits);
void o_zig_tables(double* x, double* f);
double o_standard_normal(o_pcg32* r);
void o_ref_noise(o_pcg32* rngs, int t, uint64_t n, double dr, double* rv);
void o_create_isotropic(uint64_t n, const ss_box* bs, uint64_t seed, double* out);
int o_create_homogeneous(uint64_t n, double kappa, const ss_box* bs, uint64_t seed, double* out);

#ifdef __cplusplus
}
#endif

#endif
