/*
 * ss_oracle.c -- FP64 CPU restatement of stochasticsampling's per-timestep
 * loop (Simulation::do_timestep, src/bin/simulation/mod.rs:251-374).
 *
 * TEST INFRASTRUCTURE ONLY: the checker for the HIP library and the timed
 * "reference-equivalent CPU" baseline.  Nothing in the product links it.
 *
 * It follows the reference's structure where that structure is the cost the
 * baseline measures: serial histogram and serial stress contraction with the
 * outer loop over the nine components, the magnetic solve every step, full
 * complex-to-complex 3-D transforms (36 per step), complex W/E arrays, and a
 * parallel per-particle update (rayon there, OpenMP here).
 *
 * Third-party arithmetic restated from published definitions:
 *   FFTW 3.3 (unpinned version)  -> own mixed-radix Stockham DFT, unnormalised,
 *                                   sign -1 forward, +1 backward.
 *   quaternion 0.3.1             -> Hamilton product rotation q (0,v) q*.
 *   ndarray 0.12.1 dot           -> 8-way unrolled accumulation for contiguous
 *                                   1-D dots (numeric_util::unrolled_dot).
 *   num-complex 0.2.3 division   -> 1/z computed as re/(re*re) for real z.
 *   rand_pcg 0.2.0 / rand_core 0.5.0 / rand 0.7.0 / rand_distr 0.2.1
 *                                -> PCG32 and MCG128 streams, seed_from_u64,
 *                                   the f64 uniform conversions and the
 *                                   256-layer ziggurat normal, for the
 *                                   reference's own noise (SS_NOISE_REFERENCE)
 *                                   and initial conditions.  The default noise
 *                                   is Philox4x32-10 (Salmon et al., SC'11),
 *                                   identical in the HIP library.
 */
#include "ss_oracle.h"

#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define O_PI 3.14159265358979323846
static const double PI_ = 3.141592653589793;          /* std::f64::consts::PI */
static const double TWOPI_ = 2.0 * 3.141592653589793; /* consts.rs:8         */

/* ------------------------------------------------------------------------ */
/* Scalars: particle.rs:26-44, 106-139                                       */
/* ------------------------------------------------------------------------ */

/* modulo (particle.rs:26-35): Rust's % on floats is fmod. */
double o_modulo(double f, double m) {
    double r = fmod(f, m);
    return r < 0.0 ? r + fabs(m) : r;
}

/* ang_pbc (particle.rs:37-44). */
void o_ang_pbc(double phi, double theta, double* out2) {
    double t = o_modulo(theta, TWOPI_);
    if (t > PI_) {
        out2[0] = o_modulo(phi + PI_, TWOPI_);
        out2[1] = TWOPI_ - t;
    } else {
        out2[0] = o_modulo(phi, TWOPI_);
        out2[1] = t;
    }
}

/* Particle::new (particle.rs:186-201): Position::new + Orientation::new, then
   pbc once more. */
void o_particle_new(const double in[5], const ss_box* bs, double out[5]) {
    double a[2];
    out[0] = o_modulo(in[0], bs->x);
    out[1] = o_modulo(in[1], bs->y);
    out[2] = o_modulo(in[2], bs->z);
    o_ang_pbc(in[3], in[4], a);
    out[0] = o_modulo(out[0], bs->x);
    out[1] = o_modulo(out[1], bs->y);
    out[2] = o_modulo(out[2], bs->z);
    o_ang_pbc(a[0], a[1], out + 3);
}

/* Orientation::from_vector (particle.rs:106-121). */
void o_orientation_from_vector(const double v[3], double out2[2]) {
    double rxy = sqrt(v[0] * v[0] + v[1] * v[1]);
    out2[0] = atan2(v[1], v[0]);
    out2[1] = PI_ / 2.0 - atan2(v[2], rxy);
}

/* ------------------------------------------------------------------------ */
/* Mesh: grid_width.rs:23-31, fft_helper.rs:26-100                           */
/* ------------------------------------------------------------------------ */
typedef struct {
    double x, y, z, phi, theta;
} ogw;

static ogw grid_width(const ss_params* p) {
    ogw g;
    g.x = p->box_size.x / (double)p->grid_size.x;
    g.y = p->box_size.y / (double)p->grid_size.y;
    g.z = p->box_size.z / (double)p->grid_size.z;
    g.phi = TWOPI_ / (double)p->grid_size.phi;
    g.theta = PI_ / (double)p->grid_size.theta;
    return g;
}

/* get_k_sampling (fft_helper.rs:26-51): FFTW order, Nyquist stored negative. */
void o_k_sampling(int n, double L, double* out) {
    int a = n / 2;
    int b = (n % 2 == 0) ? n / 2 : n / 2 + 1;
    double step = TWOPI_ / L;
    for (int i = 0; i < b; ++i) out[i] = (double)i * step;
    for (int i = b; i < n; ++i) out[i] = (double)(i - n) * step;
    (void)a;
}

/* get_inverse_norm_squared (fft_helper.rs:79-88): complex 1/v evaluated by
   num-complex as v.re / (v.re*v.re) for a real v; k=0 set to 0. */
void o_inverse_norm_squared(const ss_params* p, double* out) {
    int nx = p->grid_size.x, ny = p->grid_size.y, nz = p->grid_size.z;
    double* kx = malloc(sizeof(double) * nx);
    double* ky = malloc(sizeof(double) * ny);
    double* kz = malloc(sizeof(double) * nz);
    o_k_sampling(nx, p->box_size.x, kx);
    o_k_sampling(ny, p->box_size.y, ky);
    o_k_sampling(nz, p->box_size.z, kz);
    for (int i = 0; i < nx; ++i)
        for (int j = 0; j < ny; ++j)
            for (int k = 0; k < nz; ++k) {
                double c = (kx[i] * kx[i] + ky[j] * ky[j]) + kz[k] * kz[k];
                out[((size_t)i * ny + j) * nz + k] = c / (c * c);
            }
    out[0] = 0.0;
    free(kx);
    free(ky);
    free(kz);
}

/* get_inverse_norm (fft_helper.rs:93-100): 1/sqrt(v.re), k=0 set to 0. */
static void inverse_norm(const ss_params* p, double* out) {
    int nx = p->grid_size.x, ny = p->grid_size.y, nz = p->grid_size.z;
    double kx[4096], ky[4096], kz[4096];
    o_k_sampling(nx, p->box_size.x, kx);
    o_k_sampling(ny, p->box_size.y, ky);
    o_k_sampling(nz, p->box_size.z, kz);
    for (int i = 0; i < nx; ++i)
        for (int j = 0; j < ny; ++j)
            for (int k = 0; k < nz; ++k) {
                double c = (kx[i] * kx[i] + ky[j] * ky[j]) + kz[k] * kz[k];
                out[((size_t)i * ny + j) * nz + k] = 1.0 / sqrt(c);
            }
    out[0] = 0.0;
}

/* ------------------------------------------------------------------------ */
/* Kernels on angular bin centres: stress.rs:29-56,103-159, director.rs:69-92 */
/* ------------------------------------------------------------------------ */

/* ndarray 0.12 Array::linspace: start + step*i, step = (b-a)/(n-1). */
void o_linspace(double a, double b, int n, double* out) {
    double step = n > 1 ? (b - a) / ((double)n - 1.0) : 0.0;
    for (int i = 0; i < n; ++i) out[i] = a + step * (double)i;
}

static void stress_active(double phi, double theta, double s[9]) {
    double cp = cos(phi), sp = sin(phi), ct = cos(theta), st = sin(theta);
    s[0] = -(1. / 3.) + cp * cp * st * st;
    s[1] = cp * st * st * sp;
    s[2] = ct * cp * st;
    s[3] = cp * st * st * sp;
    s[4] = -(1. / 3.) + st * st * sp * sp;
    s[5] = ct * st * sp;
    s[6] = ct * cp * st;
    s[7] = ct * st * sp;
    s[8] = -(1. / 3.) + ct * ct;
}

/* stress_magnetic for a field along d: 0.5 (n d^T - d n^T); with d = y it is
   the reference's table (stress.rs:125-141). */
static void stress_magnetic(double phi, double theta, const double d[3], double s[9]) {
    double n[3] = {sin(theta) * cos(phi), sin(theta) * sin(phi), cos(theta)};
    if (d[0] == 0.0 && d[1] == 1.0 && d[2] == 0.0) {
        double cp = cos(phi), st = sin(theta), ct = cos(theta);
        s[0] = 0.;
        s[1] = cp * st;
        s[2] = 0.;
        s[3] = -cp * st;
        s[4] = 0.;
        s[5] = -ct;
        s[6] = 0.;
        s[7] = ct;
        s[8] = 0.;
    } else {
        for (int i = 0; i < 3; ++i)
            for (int j = 0; j < 3; ++j) s[3 * i + j] = n[i] * d[j] - d[i] * n[j];
    }
    for (int i = 0; i < 9; ++i) s[i] *= 0.5;
}

static void stress_magnetic_rods(double phi, double theta, double s[9]) {
    double ct = cos(theta), st = sin(theta), cp = cos(phi), sp = sin(phi);
    s[0] = 2. * ct * (cp * cp) * (st * st);
    s[1] = ct * (st * st) * cos(2. * phi);
    s[2] = cos(2. * theta) * cp * st;
    s[3] = ct * (st * st) * cos(2. * phi);
    s[4] = st * sin(2. * theta) * (sp * sp);
    s[5] = cos(2. * theta) * st * sp;
    s[6] = cos(2. * theta) * cp * st;
    s[7] = cos(2. * theta) * st * sp;
    s[8] = -2. * ct * (st * st);
    for (int i = 0; i < 9; ++i) s[i] *= 0.5;
}

static void field_dir(const ss_params* p, double d[3]) {
    d[0] = p->field_direction[0];
    d[1] = p->field_direction[1];
    d[2] = p->field_direction[2];
    if (d[0] == 0.0 && d[1] == 0.0 && d[2] == 0.0) d[1] = 1.0;
}

/* stress_kernel (stress.rs:29-56) with the closure of
   src/bin/simulation/mod.rs:100-104.  out: [3][3][nphi][ntheta]. */
void o_stress_kernel(const ss_params* p, double* out) {
    int np = p->grid_size.phi, nt = p->grid_size.theta;
    ogw gw = grid_width(p);
    double* ap = malloc(sizeof(double) * np);
    double* at = malloc(sizeof(double) * nt);
    o_linspace(0. + gw.phi / 2., TWOPI_ - gw.phi / 2., np, ap);
    o_linspace(0. + gw.theta / 2., PI_ - gw.theta / 2., nt, at);
    double d[3];
    field_dir(p, d);
    for (int i = 0; i < np; ++i)
        for (int j = 0; j < nt; ++j) {
            double sa[9], sm[9], sr[9];
            stress_active(ap[i], at[j], sa);
            stress_magnetic(ap[i], at[j], d, sm);
            stress_magnetic_rods(ap[i], at[j], sr);
            for (int q = 0; q < 9; ++q) {
                double v = p->stress_active * sa[q] + p->stress_magnetic * sm[q] + p->shape * sr[q];
                out[((size_t)q * np + i) * nt + j] = v;
            }
        }
    free(ap);
    free(at);
}

/* orientation_kernel (director.rs:69-92). out: [3][nphi][ntheta]. */
void o_orientation_kernel(const ss_params* p, double* out) {
    int np = p->grid_size.phi, nt = p->grid_size.theta;
    ogw gw = grid_width(p);
    double* ap = malloc(sizeof(double) * np);
    double* at = malloc(sizeof(double) * nt);
    o_linspace(0. + gw.phi / 2., TWOPI_ - gw.phi / 2., np, ap);
    o_linspace(0. + gw.theta / 2., PI_ - gw.theta / 2., nt, at);
    for (int i = 0; i < np; ++i)
        for (int j = 0; j < nt; ++j) {
            double o[2];
            o_ang_pbc(ap[i], at[j], o); /* Orientation::new applies ang_pbc */
            double st = sin(o[1]), ct = cos(o[1]), cp = cos(o[0]), sp = sin(o[0]);
            out[((size_t)0 * np + i) * nt + j] = st * cp;
            out[((size_t)1 * np + i) * nt + j] = st * sp;
            out[((size_t)2 * np + i) * nt + j] = ct;
        }
    free(ap);
    free(at);
}

/* ndarray numeric_util::unrolled_dot for contiguous slices. */
static double unrolled_dot(const double* x, const double* y, size_t len) {
    double sum = 0.0;
    double p0 = 0, p1 = 0, p2 = 0, p3 = 0, p4 = 0, p5 = 0, p6 = 0, p7 = 0;
    while (len >= 8) {
        p0 = p0 + x[0] * y[0];
        p1 = p1 + x[1] * y[1];
        p2 = p2 + x[2] * y[2];
        p3 = p3 + x[3] * y[3];
        p4 = p4 + x[4] * y[4];
        p5 = p5 + x[5] * y[5];
        p6 = p6 + x[6] * y[6];
        p7 = p7 + x[7] * y[7];
        x += 8;
        y += 8;
        len -= 8;
    }
    sum = sum + (p0 + p4);
    sum = sum + (p1 + p5);
    sum = sum + (p2 + p6);
    sum = sum + (p3 + p7);
    for (size_t i = 0; i < len; ++i) sum = sum + x[i] * y[i];
    return sum;
}

/* ------------------------------------------------------------------------ */
/* Binning: distribution/mod.rs:86-184                                       */
/* ------------------------------------------------------------------------ */
static int32_t clamp_idx(double q, double g, int n) {
    int64_t i = (int64_t)floor(q / g);
    if (i == n) i -= 1;
    return (int32_t)i;
}

/* coord_to_grid (distribution/mod.rs:86-141). */
void o_coord_to_grid(const ss_params* p, const double part[5], int32_t out[5]) {
    ogw gw = grid_width(p);
    out[0] = clamp_idx(part[0], gw.x, p->grid_size.x);
    out[1] = clamp_idx(part[1], gw.y, p->grid_size.y);
    out[2] = clamp_idx(part[2], gw.z, p->grid_size.z);
    out[3] = clamp_idx(part[3], gw.phi, p->grid_size.phi);
    out[4] = clamp_idx(part[4], gw.theta, p->grid_size.theta);
}

/* The same rule evaluated in binary32: widths rounded from FP64, the quotient
   correctly rounded in FP32.  This is the binning of the FP32 path. */
void o_coord_to_grid_f32(const ss_params* p, const float part[5], int32_t out[5]) {
    ogw gw = grid_width(p);
    float g[5] = {(float)gw.x, (float)gw.y, (float)gw.z, (float)gw.phi, (float)gw.theta};
    int n[5] = {p->grid_size.x, p->grid_size.y, p->grid_size.z, p->grid_size.phi,
                p->grid_size.theta};
    for (int a = 0; a < 5; ++a) {
        volatile float q = part[a] / g[a];
        int64_t i = (int64_t)floorf(q);
        if (i == n[a]) i -= 1;
        out[a] = (int32_t)i;
    }
}

static size_t bin_index(const ss_params* p, const int32_t c[5]) {
    return ((((size_t)c[0] * p->grid_size.y + c[1]) * p->grid_size.z + c[2]) * p->grid_size.phi +
            c[3]) *
               p->grid_size.theta +
           c[4];
}

/* histogram_from (distribution/mod.rs:146-160): serial. */
void o_histogram(const ss_params* p, const double* particles, uint64_t n, uint32_t* counts) {
    size_t nb = (size_t)p->grid_size.x * p->grid_size.y * p->grid_size.z * p->grid_size.phi *
                p->grid_size.theta;
    memset(counts, 0, nb * sizeof(uint32_t));
    for (uint64_t i = 0; i < n; ++i) {
        int32_t c[5];
        o_coord_to_grid(p, particles + 5 * i, c);
        counts[bin_index(p, c)] += 1;
    }
}

/* sample_from (distribution/mod.rs:164-184) followed by the volume rescale of
   src/bin/simulation/mod.rs:253-257. */
void o_sample_from(const ss_params* p, const uint32_t* counts, uint64_t n, double* dist) {
    ogw gw = grid_width(p);
    size_t nb = (size_t)p->grid_size.x * p->grid_size.y * p->grid_size.z * p->grid_size.phi *
                p->grid_size.theta;
    double den = gw.x * gw.y * gw.z * gw.phi * gw.theta * (double)n;
    double vol = p->box_size.x * p->box_size.y * p->box_size.z;
    for (size_t i = 0; i < nb; ++i) dist[i] = (double)counts[i] / den;
    for (size_t i = 0; i < nb; ++i) dist[i] *= vol;
}

/* ------------------------------------------------------------------------ */
/* FFT (FFTW's transform, restated): unnormalised c2c, interleaved complex    */
/* ------------------------------------------------------------------------ */
static int factorize(int n, int* f) {
    int k = 0, m = n;
    while (m % 4 == 0) { f[k++] = 4; m /= 4; }
    while (m % 2 == 0) { f[k++] = 2; m /= 2; }
    while (m % 3 == 0) { f[k++] = 3; m /= 3; }
    while (m % 5 == 0) { f[k++] = 5; m /= 5; }
    for (int q = 7; m > 1; q += 2)
        while (m % q == 0) { f[k++] = q; m /= q; }
    return k;
}

/* Stockham autosort, one line; tw[2t] + i tw[2t+1] = exp(sign 2 pi i t / n).
   Returns the buffer (a or b) that holds the result. */
static double* stockham_line(int n, double* a, double* b, const double* tw, const int* f, int nf) {
    int Ns = 1;
    for (int s = 0; s < nf; ++s) {
        int R = f[s];
        int m = n / R;
        for (int j = 0; j < m; ++j) {
            int k = j % Ns;
            double vr[64], vi[64], wr[64], wi[64];
            for (int r = 0; r < R; ++r) {
                double xr = a[2 * (j + r * m)], xi = a[2 * (j + r * m) + 1];
                int t = (int)(((long)r * k * (n / (Ns * R))) % n);
                double cr = tw[2 * t], ci = tw[2 * t + 1];
                vr[r] = xr * cr - xi * ci;
                vi[r] = xr * ci + xi * cr;
            }
            for (int q = 0; q < R; ++q) {
                double sr = 0.0, si = 0.0;
                for (int r = 0; r < R; ++r) {
                    int t = (int)(((long)r * q * m) % n);
                    double cr = tw[2 * t], ci = tw[2 * t + 1];
                    sr += vr[r] * cr - vi[r] * ci;
                    si += vr[r] * ci + vi[r] * cr;
                }
                wr[q] = sr;
                wi[q] = si;
            }
            int idxD = (j / Ns) * Ns * R + k;
            for (int r = 0; r < R; ++r) {
                b[2 * (idxD + r * Ns)] = wr[r];
                b[2 * (idxD + r * Ns) + 1] = wi[r];
            }
        }
        double* t = a;
        a = b;
        b = t;
        Ns *= R;
    }
    return a;
}

static double* twiddles(int n, int sign) {
    double* tw = malloc(sizeof(double) * 2 * n);
    for (int t = 0; t < n; ++t) {
        double ang = 2.0 * O_PI * (double)t / (double)n;
        tw[2 * t] = cos(ang);
        tw[2 * t + 1] = sign * sin(ang);
    }
    return tw;
}

void o_fft1d(int n, double* data, int sign) {
    int f[64];
    int nf = factorize(n, f);
    double* tw = twiddles(n, sign);
    double* tmp = malloc(sizeof(double) * 2 * n);
    double* a = malloc(sizeof(double) * 2 * n);
    memcpy(a, data, sizeof(double) * 2 * n);
    const double* res = stockham_line(n, a, tmp, tw, f, nf);
    memcpy(data, res, sizeof(double) * 2 * n);
    free(a);
    free(tmp);
    free(tw);
}

static void fft_axis(double* data, int n, size_t count_outer, size_t stride_outer, size_t count_inner,
                     size_t stride_el, int sign) {
    int f[64];
    int nf = factorize(n, f);
    double* tw = twiddles(n, sign);
    double* a = malloc(sizeof(double) * 2 * n);
    double* b = malloc(sizeof(double) * 2 * n);
    for (size_t o = 0; o < count_outer; ++o)
        for (size_t l = 0; l < count_inner; ++l) {
            double* base = data + 2 * (o * stride_outer + l);
            for (int i = 0; i < n; ++i) {
                a[2 * i] = base[2 * i * stride_el];
                a[2 * i + 1] = base[2 * i * stride_el + 1];
            }
            const double* res = stockham_line(n, a, b, tw, f, nf);
            for (int i = 0; i < n; ++i) {
                base[2 * i * stride_el] = res[2 * i];
                base[2 * i * stride_el + 1] = res[2 * i + 1];
            }
        }
    free(a);
    free(b);
    free(tw);
}

/* In-place 3-D c2c over [nx][ny][nz]; sign -1 forward, +1 backward. */
void o_fft3d(int nx, int ny, int nz, double* data, int sign) {
    size_t nyz = (size_t)ny * nz;
    /* z: lines (x,y) contiguous of length nz */
    fft_axis(data, nz, (size_t)nx * ny, nz, 1, 1, sign);
    /* y: for each x, lines over z (stride 1), elements stride nz */
    fft_axis(data, ny, nx, nyz, nz, nz, sign);
    /* x: lines over (y,z) (stride 1), elements stride ny*nz */
    fft_axis(data, nx, 1, 0, nyz, nyz, sign);
}

/* ------------------------------------------------------------------------ */
/* Spectral solves: spectral_solver.rs:133-254, magnetic_solver.rs:71-161    */
/* ------------------------------------------------------------------------ */
typedef struct {
    int nx, ny, nz, np, nt;
    size_t ns, na;
    double *kx, *ky, *kz; /* 1-D samplings */
    double* ik2;          /* [Ns] */
    double* kinv;         /* [Ns] */
    double* skern;        /* [9][Na] */
    double* okern;        /* [3][Na] */
    double measure;
} osolver;

static void solver_init(osolver* s, const ss_params* p) {
    s->nx = p->grid_size.x;
    s->ny = p->grid_size.y;
    s->nz = p->grid_size.z;
    s->np = p->grid_size.phi;
    s->nt = p->grid_size.theta;
    s->ns = (size_t)s->nx * s->ny * s->nz;
    s->na = (size_t)s->np * s->nt;
    s->kx = malloc(sizeof(double) * s->nx);
    s->ky = malloc(sizeof(double) * s->ny);
    s->kz = malloc(sizeof(double) * s->nz);
    o_k_sampling(s->nx, p->box_size.x, s->kx);
    o_k_sampling(s->ny, p->box_size.y, s->ky);
    o_k_sampling(s->nz, p->box_size.z, s->kz);
    s->ik2 = malloc(sizeof(double) * s->ns);
    s->kinv = malloc(sizeof(double) * s->ns);
    o_inverse_norm_squared(p, s->ik2);
    inverse_norm(p, s->kinv);
    s->skern = malloc(sizeof(double) * 9 * s->na);
    s->okern = malloc(sizeof(double) * 3 * s->na);
    o_stress_kernel(p, s->skern);
    o_orientation_kernel(p, s->okern);
    ogw gw = grid_width(p);
    s->measure = gw.phi * gw.theta;
}

static void solver_free(osolver* s) {
    free(s->kx);
    free(s->ky);
    free(s->kz);
    free(s->ik2);
    free(s->kinv);
    free(s->skern);
    free(s->okern);
}

static inline void kvec(const osolver* s, size_t idx, double k[3]) {
    size_t iz = idx % s->nz, iy = (idx / s->nz) % s->ny, ix = idx / ((size_t)s->ny * s->nz);
    k[0] = s->kx[ix];
    k[1] = s->ky[iy];
    k[2] = s->kz[iz];
}

/* update_gradient (spectral_solver.rs:192-229 / magnetic_solver.rs:98-135):
   G_ab = i k_a v_b, then nine backward transforms.  v: [3][Ns], g: [9][Ns]. */
static void gradient_and_inverse(const osolver* s, const double* v, double* g) {
#pragma omp parallel for schedule(static)
    for (long long idx = 0; idx < (long long)s->ns; ++idx) {
        double k[3];
        kvec(s, (size_t)idx, k);
        for (int a = 0; a < 3; ++a)
            for (int b = 0; b < 3; ++b) {
                double vr = v[2 * ((size_t)b * s->ns + idx)], vi = v[2 * ((size_t)b * s->ns + idx) + 1];
                double er = k[a] * vr, ei = k[a] * vi; /* k (im 0) times v */
                g[2 * ((size_t)(3 * a + b) * s->ns + idx)] = -ei; /* times i */
                g[2 * ((size_t)(3 * a + b) * s->ns + idx) + 1] = er;
            }
    }
#pragma omp parallel for schedule(dynamic, 1)
    for (int c = 0; c < 9; ++c) o_fft3d(s->nx, s->ny, s->nz, g + 2 * (size_t)c * s->ns, +1);
}

/* fft_mean_flow_field + update_gradient + backward of u
   (spectral_solver.rs:133-254).  sigma: [9][Ns] complex work; u: [3][Ns]
   complex out; g: [9][Ns] complex out. */
static void flow_core(const osolver* s, double screening, const double* dist, double* sigma,
                      double* u, double* g) {
    /* average_stress (stress.rs:60-96): outer loop over components, serial */
    for (int q = 0; q < 9; ++q)
        for (size_t c = 0; c < s->ns; ++c) {
            sigma[2 * ((size_t)q * s->ns + c)] =
                unrolled_dot(s->skern + (size_t)q * s->na, dist + c * s->na, s->na) * s->measure;
            sigma[2 * ((size_t)q * s->ns + c) + 1] = 0.0;
        }
#pragma omp parallel for schedule(dynamic, 1)
    for (int q = 0; q < 9; ++q) o_fft3d(s->nx, s->ny, s->nz, sigma + 2 * (size_t)q * s->ns, -1);

    double norm = (double)s->ns;
#pragma omp parallel for schedule(static)
    for (long long idx = 0; idx < (long long)s->ns; ++idx) {
        double k[3], kn[3];
        kvec(s, (size_t)idx, k);
        for (int a = 0; a < 3; ++a) kn[a] = k[a] * s->kinv[idx];
        double sr[3], si[3];
        for (int i = 0; i < 3; ++i) {
            double ar = 0.0, ai = 0.0;
            for (int j = 0; j < 3; ++j) {
                size_t off = 2 * ((size_t)(3 * i + j) * s->ns + idx);
                ar = ar + sigma[off] * k[j];
                ai = ai + sigma[off + 1] * k[j];
            }
            sr[i] = ar;
            si[i] = ai;
        }
        double kr = 0.0, ki = 0.0;
        for (int i = 0; i < 3; ++i) {
            kr = kr + kn[i] * sr[i];
            ki = ki + kn[i] * si[i];
        }
        double fac = (s->ik2[idx] + screening) / norm;
        for (int i = 0; i < 3; ++i) {
            double ar = (sr[i] - kn[i] * kr) * fac;
            double ai = (si[i] - kn[i] * ki) * fac;
            u[2 * ((size_t)i * s->ns + idx)] = -ai; /* times i */
            u[2 * ((size_t)i * s->ns + idx) + 1] = ar;
        }
    }
    gradient_and_inverse(s, u, g);
#pragma omp parallel for schedule(dynamic, 1)
    for (int c = 0; c < 3; ++c) o_fft3d(s->nx, s->ny, s->nz, u + 2 * (size_t)c * s->ns, +1);
}

/* DirectorField::from_distribution + fft_mean_magnetic_field + update_gradient
   + backward of b (director.rs:36-61, magnetic_solver.rs:71-161). */
static void magnetic_core(const osolver* s, const double* dist, double* b, double* g) {
#pragma omp parallel for schedule(static)
    for (long long c = 0; c < (long long)s->ns; ++c)
        for (int q = 0; q < 3; ++q) {
            b[2 * ((size_t)q * s->ns + c)] =
                unrolled_dot(s->okern + (size_t)q * s->na, dist + (size_t)c * s->na, s->na) *
                s->measure;
            b[2 * ((size_t)q * s->ns + c) + 1] = 0.0;
        }
#pragma omp parallel for schedule(dynamic, 1)
    for (int q = 0; q < 3; ++q) o_fft3d(s->nx, s->ny, s->nz, b + 2 * (size_t)q * s->ns, -1);
    double norm = (double)s->ns;
#pragma omp parallel for schedule(static)
    for (long long idx = 0; idx < (long long)s->ns; ++idx) {
        double k[3], kn[3];
        kvec(s, (size_t)idx, k);
        for (int a = 0; a < 3; ++a) kn[a] = k[a] * s->kinv[idx];
        double kr = 0.0, ki = 0.0;
        for (int q = 0; q < 3; ++q) {
            kr = kr + kn[q] * b[2 * ((size_t)q * s->ns + idx)];
            ki = ki + kn[q] * b[2 * ((size_t)q * s->ns + idx) + 1];
        }
        for (int q = 0; q < 3; ++q) {
            size_t off = 2 * ((size_t)q * s->ns + idx);
            b[off] = (b[off] - kn[q] * kr) / norm;
            b[off + 1] = (b[off + 1] - kn[q] * ki) / norm;
        }
    }
    gradient_and_inverse(s, b, g);
#pragma omp parallel for schedule(dynamic, 1)
    for (int q = 0; q < 3; ++q) o_fft3d(s->nx, s->ny, s->nz, b + 2 * (size_t)q * s->ns, +1);
}

static void real_parts(const double* c, double* r, size_t n) {
    for (size_t i = 0; i < n; ++i) r[i] = c[2 * i];
}

int o_mean_flow_field(const ss_params* p, double screening, const double* dist, double* u,
                      double* grad_u) {
    osolver s;
    solver_init(&s, p);
    double* sig = malloc(sizeof(double) * 2 * 9 * s.ns);
    double* uc = malloc(sizeof(double) * 2 * 3 * s.ns);
    double* gc = malloc(sizeof(double) * 2 * 9 * s.ns);
    if (!sig || !uc || !gc) return -1;
    flow_core(&s, screening, dist, sig, uc, gc);
    if (u) real_parts(uc, u, 3 * s.ns);
    if (grad_u) real_parts(gc, grad_u, 9 * s.ns);
    free(sig);
    free(uc);
    free(gc);
    solver_free(&s);
    return 0;
}

/* SpectralSolver::solve_flow_field (spectral_solver.rs:85-131), the
   deprecated second formulation of the same solve (no screening):
   sigma_k_i = i (sum_j sigma^_ij k_j) / k^2, u^ = sigma_k - k (k . sigma_k) / k^2
   with the unnormalised k mesh and 1/k^2 (k = 0 -> 0), backward transforms,
   then u = Re / N.  Sums over a component axis run in order (ndarray's
   sum_axis over an outer axis adds the sub-views one after another).
   u: [3][Ns] real parts. */
int o_solve_flow_field(const ss_params* p, const double* dist, double* u) {
    osolver s;
    solver_init(&s, p);
    double* sig = malloc(sizeof(double) * 2 * 9 * s.ns);
    double* uc = malloc(sizeof(double) * 2 * 3 * s.ns);
    if (!sig || !uc) {
        free(sig);
        free(uc);
        solver_free(&s);
        return -1;
    }
    for (int q = 0; q < 9; ++q)
        for (size_t c = 0; c < s.ns; ++c) {
            sig[2 * ((size_t)q * s.ns + c)] =
                unrolled_dot(s.skern + (size_t)q * s.na, dist + c * s.na, s.na) * s.measure;
            sig[2 * ((size_t)q * s.ns + c) + 1] = 0.0;
        }
#pragma omp parallel for schedule(dynamic, 1)
    for (int q = 0; q < 9; ++q) o_fft3d(s.nx, s.ny, s.nz, sig + 2 * (size_t)q * s.ns, -1);
#pragma omp parallel for schedule(static)
    for (long long idx = 0; idx < (long long)s.ns; ++idx) {
        double k[3];
        kvec(&s, (size_t)idx, k);
        double ik2 = s.ik2[idx];
        double skr[3], ski[3];
        for (int i = 0; i < 3; ++i) {
            double ar = 0.0, ai = 0.0;
            for (int j = 0; j < 3; ++j) {
                size_t off = 2 * ((size_t)(3 * i + j) * s.ns + idx);
                ar = ar + sig[off] * k[j];
                ai = ai + sig[off + 1] * k[j];
            }
            ar = ar * ik2;
            ai = ai * ik2;
            skr[i] = -ai; /* times i */
            ski[i] = ar;
        }
        double kr = 0.0, ki = 0.0;
        for (int i = 0; i < 3; ++i) {
            kr = kr + k[i] * skr[i];
            ki = ki + k[i] * ski[i];
        }
        kr = kr * ik2;
        ki = ki * ik2;
        for (int i = 0; i < 3; ++i) {
            uc[2 * ((size_t)i * s.ns + idx)] = skr[i] - k[i] * kr;
            uc[2 * ((size_t)i * s.ns + idx) + 1] = ski[i] - k[i] * ki;
        }
    }
#pragma omp parallel for schedule(dynamic, 1)
    for (int c = 0; c < 3; ++c) o_fft3d(s.nx, s.ny, s.nz, uc + 2 * (size_t)c * s.ns, +1);
    double norm = (double)s.ns;
    for (size_t i = 0; i < 3 * s.ns; ++i) u[i] = uc[2 * i] / norm;
    free(sig);
    free(uc);
    solver_free(&s);
    return 0;
}

int o_mean_magnetic_field(const ss_params* p, const double* dist, double* b, double* grad_b) {
    osolver s;
    solver_init(&s, p);
    double* bc = malloc(sizeof(double) * 2 * 3 * s.ns);
    double* gc = malloc(sizeof(double) * 2 * 9 * s.ns);
    if (!bc || !gc) return -1;
    magnetic_core(&s, dist, bc, gc);
    if (b) real_parts(bc, b, 3 * s.ns);
    if (grad_b) real_parts(gc, grad_b, 9 * s.ns);
    free(bc);
    free(gc);
    solver_free(&s);
    return 0;
}

/* ------------------------------------------------------------------------ */
/* Physics the live step leaves unused: DensityGradient                      */
/* (distribution/density_gradient.rs:22-110) and the trilinear interpolation  */
/* of mesh/interpolate.rs:11-123                                             */
/* ------------------------------------------------------------------------ */
/* ndarray numeric_util::unrolled_sum for contiguous slices (ArrayBase::sum). */
static double unrolled_sum(const double* x, size_t len) {
    double sum = 0.0;
    double p0 = 0, p1 = 0, p2 = 0, p3 = 0, p4 = 0, p5 = 0, p6 = 0, p7 = 0;
    while (len >= 8) {
        p0 = p0 + x[0];
        p1 = p1 + x[1];
        p2 = p2 + x[2];
        p3 = p3 + x[3];
        p4 = p4 + x[4];
        p5 = p5 + x[5];
        p6 = p6 + x[6];
        p7 = p7 + x[7];
        x += 8;
        len -= 8;
    }
    sum = sum + (p0 + p4);
    sum = sum + (p1 + p5);
    sum = sum + (p2 + p6);
    sum = sum + (p3 + p7);
    for (size_t i = 0; i < len; ++i) sum = sum + x[i];
    return sum;
}

/* DensityGradient::get_real_gradient after update_gradient: density per
   cell (sum over the angles times dtheta dphi), forward FFT, times k (i / N),
   three inverse FFTs, real parts.  grad: 3 x Ns. */
void o_density_gradient(const ss_params* p, const double* dist, double* grad) {
    const int nx = p->grid_size.x, ny = p->grid_size.y, nz = p->grid_size.z;
    const size_t ns = (size_t)nx * ny * nz, na = (size_t)p->grid_size.phi * p->grid_size.theta;
    ogw gw = grid_width(p);
    const double dthph = gw.theta * gw.phi;
    double* rho = calloc(2 * ns, sizeof(double));
    double* g = calloc(2 * 3 * ns, sizeof(double));
    double *kx = malloc(sizeof(double) * nx), *ky = malloc(sizeof(double) * ny), *kz = malloc(sizeof(double) * nz);
    o_k_sampling(nx, p->box_size.x, kx);
    o_k_sampling(ny, p->box_size.y, ky);
    o_k_sampling(nz, p->box_size.z, kz);
    for (size_t c = 0; c < ns; ++c) {
        /* Complex::from(sum) * dthph: (re * dthph, 0 * dthph) */
        rho[2 * c] = unrolled_sum(dist + c * na, na) * dthph;
        rho[2 * c + 1] = 0.0 * dthph;
    }
    o_fft3d(nx, ny, nz, rho, -1);
    /* norm = Complex(0, 1) / n; e = (k * d) * norm in num-complex arithmetic,
       k = (k_a, 0) */
    const double nre = 0.0 / (double)ns, nim = 1.0 / (double)ns;
    for (int ix = 0; ix < nx; ++ix)
        for (int iy = 0; iy < ny; ++iy)
            for (int iz = 0; iz < nz; ++iz) {
                const size_t c = ((size_t)ix * ny + iy) * nz + iz;
                const double k[3] = {kx[ix], ky[iy], kz[iz]};
                const double dre = rho[2 * c], dim = rho[2 * c + 1];
                for (int a = 0; a < 3; ++a) {
                    const double kre = k[a], kim = 0.0;
                    const double e1re = kre * dre - kim * dim, e1im = kre * dim + kim * dre;
                    g[2 * ((size_t)a * ns + c)] = e1re * nre - e1im * nim;
                    g[2 * ((size_t)a * ns + c) + 1] = e1re * nim + e1im * nre;
                }
            }
    for (int a = 0; a < 3; ++a) o_fft3d(nx, ny, nz, g + 2 * (size_t)a * ns, +1);
    for (size_t q = 0; q < 3 * ns; ++q) grad[q] = g[2 * q];
    free(rho);
    free(g);
    free(kx);
    free(ky);
    free(kz);
}

/* lerp 0.2.0: a (1 - t) + b t. */
static inline double o_lerp(double a, double b, double t) { return a * (1.0 - t) + b * t; }

/* trilinear_interpolation (interpolate.rs:105-123): v[i][j][k] at i, j, k in
   {0, 1} along x, y, z; x first, then y, then z. */
double o_trilinear(const double rel[3], const double v[8]) {
    const double c00 = o_lerp(v[0], v[4], rel[0]);
    const double c01 = o_lerp(v[1], v[5], rel[0]);
    const double c10 = o_lerp(v[2], v[6], rel[0]);
    const double c11 = o_lerp(v[3], v[7], rel[0]);
    const double c0 = o_lerp(c00, c10, rel[1]);
    const double c1 = o_lerp(c01, c11, rel[1]);
    return o_lerp(c0, c1, rel[2]);
}

/* get_neighbouring_cell_values (interpolate.rs:61-103): the 2x2x2 cells
   around the position, from floor(q / g - 1/2), wrapped periodically. */
void o_neighbour_values(const double* f, int nx, int ny, int nz, const double pos[3], const double gw[3],
                        double v[8]) {
    const int n[3] = {nx, ny, nz};
    int idx[3];
    for (int a = 0; a < 3; ++a) idx[a] = (int)floor(pos[a] / gw[a] - 0.5);
    for (int i = 0; i < 2; ++i)
        for (int j = 0; j < 2; ++j)
            for (int k = 0; k < 2; ++k) {
                const int r[3] = {idx[0] + i, idx[1] + j, idx[2] + k};
                int w[3];
                for (int a = 0; a < 3; ++a) w[a] = (r[a] + n[a]) % n[a];
                v[4 * i + 2 * j + k] = f[((size_t)w[0] * ny + w[1]) * nz + w[2]];
            }
}

/* interpolate_vector_field (interpolate.rs:11-39) for a field [3][nx][ny][nz]:
   the fractional offsets (q / g + 1/2) mod 1 -- the reference takes z's from
   position.y, restated here from position.z. */
void o_interpolate_vector_field(const double* field, int nx, int ny, int nz, const double pos[3],
                                const double gw[3], double out[3]) {
    double rel[3];
    for (int a = 0; a < 3; ++a) rel[a] = fmod(pos[a] / gw[a] + 0.5, 1.0);
    const size_t ns = (size_t)nx * ny * nz;
    for (int c = 0; c < 3; ++c) {
        double v[8];
        o_neighbour_values(field + (size_t)c * ns, nx, ny, nz, pos, gw, v);
        out[c] = o_trilinear(rel, v);
    }
}

/* ------------------------------------------------------------------------ */
/* Noise: Philox4x32-10 (Salmon et al. 2011) -> the RandomVector of          */
/* src/bin/simulation/mod.rs:51-57,259-286,383-385                           */
/* ------------------------------------------------------------------------ */
void o_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]) {
    uint32_t c0 = ctr[0], c1 = ctr[1], c2 = ctr[2], c3 = ctr[3];
    uint32_t k0 = key[0], k1 = key[1];
    for (int r = 0; r < 10; ++r) {
        if (r > 0) {
            k0 += 0x9E3779B9u;
            k1 += 0xBB67AE85u;
        }
        uint64_t p0 = (uint64_t)0xD2511F53u * c0;
        uint64_t p1 = (uint64_t)0xCD9E8D57u * c2;
        uint32_t hi0 = (uint32_t)(p0 >> 32), lo0 = (uint32_t)p0;
        uint32_t hi1 = (uint32_t)(p1 >> 32), lo1 = (uint32_t)p1;
        c0 = hi1 ^ c1 ^ k0;
        c1 = lo1;
        c2 = hi0 ^ c3 ^ k1;
        c3 = lo0;
    }
    out[0] = c0;
    out[1] = c1;
    out[2] = c2;
    out[3] = c3;
}

/* Twelve words per particle and step: counter (j, id, step lo, step hi),
   j = 0,1,2; key (seed lo, seed hi).  Six 21-bit fields come from the words
   w0..w3 of call 0 (the top 21 bits of each word; then the low 11 bits of
   w0|w1 and of w2|w3, 11 + 10 bits); uniform k is (m_k + 1/2) 2^-52 with m_k
   the field followed by the top 31 bits of word 4+k (calls 1 and 2).  The
   uniforms are exactly representable and strictly inside (0, 1). */
void o_noise(uint64_t seed, uint64_t step, uint64_t id, double dr, double out[5]) {
    uint32_t key[2] = {(uint32_t)seed, (uint32_t)(seed >> 32)};
    uint32_t w[12];
    for (uint32_t j = 0; j < 3; ++j) {
        uint32_t ctr[4] = {j, (uint32_t)id, (uint32_t)step, (uint32_t)(step >> 32)};
        o_philox4x32_10(ctr, key, w + 4 * j);
    }
    uint32_t f[6] = {w[0] >> 11, w[1] >> 11, w[2] >> 11, w[3] >> 11,
                     ((w[0] & 0x7FFu) << 10) | ((w[1] & 0x7FFu) >> 1),
                     ((w[2] & 0x7FFu) << 10) | ((w[3] & 0x7FFu) >> 1)};
    double u[6];
    for (int k = 0; k < 6; ++k) {
        uint64_t m = ((uint64_t)f[k] << 31) | (uint64_t)(w[4 + k] >> 1);
        u[k] = ((double)m + 0.5) * (1.0 / 4503599627370496.0);
    }
    double r0 = sqrt(-2.0 * log(u[0]));
    double r1 = sqrt(-2.0 * log(u[2]));
    out[0] = r0 * cos(TWOPI_ * u[1]);
    out[1] = r0 * sin(TWOPI_ * u[1]);
    out[2] = r1 * cos(TWOPI_ * u[3]);
    out[3] = TWOPI_ * u[4];
    out[4] = dr * sqrt(-2.0 * log(1.0 - u[5])); /* rayleigh_pdf (mod.rs:383-385) */
}

/* ------------------------------------------------------------------------ */
/* The reference's own random streams, restated from the published           */
/* algorithms of its pinned dependencies (Cargo.lock):                        */
/*   rand_pcg 0.2.0   Lcg64Xsh32 (= Pcg32, PCG XSH-RR 64/32) and Mcg128Xsl64  */
/*                    (= Pcg64Mcg, MCG XSL-RR 128/64), O'Neill 2014;          */
/*   rand_core 0.5.0  seed_from_u64 (PCG32 output words fill the seed);       */
/*   rand 0.7.0       Uniform<f64> in [0, 1), Standard and Open01 for f64;    */
/*   rand_distr 0.2.1 StandardNormal: 256-layer ziggurat (Marsaglia & Tsang   */
/*                    2000, Doornik 2005 layout) with its R, V constants;     */
/* used by Simulation::new / do_timestep (src/bin/simulation/mod.rs:111-128,  */
/* 259-286) and Particle::create_isotropic / create_homogeneous               */
/* (src/particle.rs:224-297).                                                 */
/* ------------------------------------------------------------------------ */
static inline void qmul(const double a0, const double* a, const double b0, const double* b,
                        double* r0, double* r);

static const uint64_t PCG_MUL = 6364136223846793005ull;
static const uint64_t SEED_INC = 11634580027462260723ull; /* rand_core seed_from_u64 */

static inline uint32_t pcg_xsh_rr(uint64_t s) {
    uint32_t x = (uint32_t)(((s >> 18) ^ s) >> 27);
    uint32_t r = (uint32_t)(s >> 59);
    return (x >> r) | (x << ((32u - r) & 31u));
}

/* Lcg64Xsh32::from_state_incr: state += increment, then one step. */
static void pcg32_from_state_incr(uint64_t state, uint64_t inc, o_pcg32* r) {
    r->inc = inc;
    r->state = (state + inc) * PCG_MUL + inc;
}

void o_pcg32_new(uint64_t state, uint64_t stream, o_pcg32* r) {
    pcg32_from_state_incr(state, (stream << 1) | 1u, r);
}

/* from_seed: two little-endian u64, the increment forced odd. */
void o_pcg32_from_seed(const uint8_t seed[16], o_pcg32* r) {
    uint64_t w[2] = {0, 0};
    for (int k = 0; k < 16; ++k) w[k / 8] |= (uint64_t)seed[k] << (8 * (k % 8));
    pcg32_from_state_incr(w[0], w[1] | 1u, r);
}

/* rand_core 0.5 SeedableRng::seed_from_u64: each 4-byte word of the seed is
   a PCG32 output of a private LCG advanced before the output. */
static void seed_bytes_from_u64(uint64_t state, uint8_t* seed, int nbytes) {
    for (int k = 0; k < nbytes; k += 4) {
        state = state * PCG_MUL + SEED_INC;
        uint32_t x = pcg_xsh_rr(state);
        for (int b = 0; b < 4 && k + b < nbytes; ++b) seed[k + b] = (uint8_t)(x >> (8 * b));
    }
}

void o_pcg32_seed_from_u64(uint64_t s, o_pcg32* r) {
    uint8_t seed[16];
    seed_bytes_from_u64(s, seed, 16);
    o_pcg32_from_seed(seed, r);
}

uint32_t o_pcg32_next_u32(o_pcg32* r) {
    uint64_t old = r->state;
    r->state = old * PCG_MUL + r->inc;
    return pcg_xsh_rr(old);
}

/* rand_core impls::next_u64_via_u32: low word first. */
uint64_t o_pcg32_next_u64(o_pcg32* r) {
    uint64_t lo = o_pcg32_next_u32(r);
    uint64_t hi = o_pcg32_next_u32(r);
    return (hi << 32) | lo;
}

typedef unsigned __int128 u128;
static const u128 MCG_MUL = ((u128)0x2360ED051FC65DA4ull << 64) | 0x4385DF649FCCF645ull;

/* Mcg128Xsl64::new: the state forced odd. */
void o_mcg128_new(uint64_t lo, uint64_t hi, o_mcg128* r) {
    r->lo = lo | 1u;
    r->hi = hi;
}

void o_mcg128_seed_from_u64(uint64_t s, o_mcg128* r) {
    uint8_t seed[16];
    seed_bytes_from_u64(s, seed, 16);
    uint64_t w[2] = {0, 0};
    for (int k = 0; k < 16; ++k) w[k / 8] |= (uint64_t)seed[k] << (8 * (k % 8));
    o_mcg128_new(w[0], w[1], r);
}

/* One multiplication, then XSL-RR of the new state. */
uint64_t o_mcg128_next_u64(o_mcg128* r) {
    u128 s = (((u128)r->hi << 64) | r->lo) * MCG_MUL;
    r->lo = (uint64_t)s;
    r->hi = (uint64_t)(s >> 64);
    uint64_t x = r->hi ^ r->lo;
    unsigned rot = (unsigned)(r->hi >> 58);
    return (x >> rot) | (x << ((64u - rot) & 63u));
}

/* Bits 12..63 as the mantissa of a double with exponent e: [2^e, 2^(e+1)). */
static inline double mant_exp(uint64_t bits, int e) {
    uint64_t v = (bits >> 12) | ((uint64_t)(1023 + e) << 52);
    double d;
    memcpy(&d, &v, 8);
    return d;
}

/* rand 0.7 UniformFloat::sample for Uniform::new(0, 1): scale 1, low 0. */
double o_uniform01(uint64_t bits) { return (mant_exp(bits, 0) - 1.0) * 1.0 + 0.0; }
/* Standard for f64: the top 53 bits times 2^-53. */
double o_gen_f64(uint64_t bits) { return (double)(bits >> 11) * (1.0 / 9007199254740992.0); }
/* Open01 for f64: [1, 2) minus (1 - eps/2). */
double o_open01(uint64_t bits) { return mant_exp(bits, 0) - (1.0 - 2.220446049250313e-16 / 2.0); }

/* Ziggurat tables of rand_distr's normal: x_0 = V / f(R), x_1 = R,
   x_i = f^-1(V / x_(i-1) + f(x_(i-1))), x_256 = 0, f_i = f(x_i), with
   f(x) = exp(-x^2/2).  The crate's tables are these numbers printed with 18
   decimals; printing and reading back reproduces them. */
static double ZX[257], ZF[257];
static int zig_ready = 0;
void o_zig_tables(double* x, double* f) {
    if (!zig_ready) {
        const double R = 3.6541528853610088, V = 0.00492867323399;
        double xv[257], fv[257];
        xv[0] = V / exp(-R * R / 2.0);
        xv[1] = R;
        for (int i = 2; i < 256; ++i) {
            double last = xv[i - 1];
            xv[i] = sqrt(-2.0 * log(V / last + exp(-last * last / 2.0)));
        }
        xv[256] = 0.0;
        for (int i = 0; i < 257; ++i) fv[i] = exp(-xv[i] * xv[i] / 2.0);
        char buf[64];
        for (int i = 0; i < 257; ++i) {
            snprintf(buf, sizeof buf, "%.18f", xv[i]);
            ZX[i] = strtod(buf, NULL);
            snprintf(buf, sizeof buf, "%.18f", fv[i]);
            ZF[i] = strtod(buf, NULL);
        }
        zig_ready = 1;
    }
    if (x) memcpy(x, ZX, sizeof ZX);
    if (f) memcpy(f, ZF, sizeof ZF);
}

/* StandardNormal for f64 (rand_distr 0.2.1): symmetric ziggurat on one u64
   per try (layer from the low 8 bits, u in [-1, 1) from the top 52), tail
   from pairs of Open01 draws, wedge test against exp(-x^2/2). */
double o_standard_normal(o_pcg32* r) {
    o_zig_tables(NULL, NULL);
    const double R = ZX[1];
    for (;;) {
        uint64_t bits = o_pcg32_next_u64(r);
        int i = (int)(bits & 0xff);
        double u = mant_exp(bits, 1) - 3.0;
        double x = u * ZX[i];
        if (fabs(x) < ZX[i + 1]) return x;
        if (i == 0) {
            double xx = 1.0, yy = 0.0;
            while (-2.0 * yy < xx * xx) {
                double x_ = o_open01(o_pcg32_next_u64(r));
                double y_ = o_open01(o_pcg32_next_u64(r));
                xx = log(x_) / R;
                yy = log(y_);
            }
            return u < 0.0 ? xx - R : R - xx;
        }
        double g = o_gen_f64(o_pcg32_next_u64(r));
        if (ZF[i + 1] + (ZF[i] - ZF[i + 1]) * g < exp(-x * x / 2.0)) return x;
    }
}

/* The noise draw of do_timestep (mod.rs:259-286): chunks of N / T + 1
   particles, chunk c drawn in order from stream c; per particle three
   normals, then axis angle 2 pi U and rotate angle rayleigh_pdf(dr, U)
   (mod.rs:383-385). */
void o_ref_noise(o_pcg32* rngs, int t, uint64_t n, double dr, double* rv) {
    uint64_t cs = n / (uint64_t)t + 1;
#pragma omp parallel for schedule(static, 1)
    for (int c = 0; c < t; ++c) {
        uint64_t b = (uint64_t)c * cs;
        if (b >= n) continue;
        uint64_t e = b + cs < n ? b + cs : n;
        o_pcg32* r = rngs + c;
        for (uint64_t i = b; i < e; ++i) {
            double* o = rv + 5 * i;
            o[0] = o_standard_normal(r);
            o[1] = o_standard_normal(r);
            o[2] = o_standard_normal(r);
            o[3] = TWOPI_ * o_uniform01(o_pcg32_next_u64(r));
            o[4] = dr * sqrt(-2. * log(1. - o_uniform01(o_pcg32_next_u64(r))));
        }
    }
}

/* Particle::create_isotropic (particle.rs:224-255): Pcg64Mcg seeded from
   the u64 seed, five uniforms per particle: box * U for the position,
   2 pi U, pdf_sin(2 U) = acos(1 - 2 U); then Particle::new. */
void o_create_isotropic(uint64_t n, const ss_box* bs, uint64_t seed, double* out) {
    o_mcg128 r;
    o_mcg128_seed_from_u64(seed, &r);
    for (uint64_t i = 0; i < n; ++i) {
        double in[5];
        in[0] = bs->x * o_uniform01(o_mcg128_next_u64(&r));
        in[1] = bs->y * o_uniform01(o_mcg128_next_u64(&r));
        in[2] = bs->z * o_uniform01(o_mcg128_next_u64(&r));
        in[3] = TWOPI_ * o_uniform01(o_mcg128_next_u64(&r));
        in[4] = acos(1. - 2. * o_uniform01(o_mcg128_next_u64(&r)));
        o_particle_new(in, bs, out + 5 * i);
    }
}

/* Particle::create_homogeneous (particle.rs:257-297): polar angle from
   pdf_homogeneous_fixpoint (particle.rs:347-360), Particle::new, then the
   orientation vector rotated by the quaternion axis_angle([1, 0, 0], -pi/2)
   (quaternion 0.3.1: (cos h, axis sin h), h = angle / 2; q (0, v) q*) and
   written back by Orientation::from_vector_mut without a further wrap.
   Returns -1 (the reference panics) for kappa == 0 or a NaN angle. */
int o_create_homogeneous(uint64_t n, double kappa, const ss_box* bs, uint64_t seed, double* out) {
    if (kappa == 0.0) return -1;
    o_mcg128 r;
    o_mcg128_seed_from_u64(seed, &r);
    const double half = -(PI_ / 2.) / 2.;
    const double sh = sin(half), q0 = cos(half);
    const double qv[3] = {1. * sh, 0. * sh, 0. * sh};
    const double qc[3] = {-qv[0], -qv[1], -qv[2]};
    for (uint64_t i = 0; i < n; ++i) {
        double in[5];
        in[0] = bs->x * o_uniform01(o_mcg128_next_u64(&r));
        in[1] = bs->y * o_uniform01(o_mcg128_next_u64(&r));
        in[2] = bs->z * o_uniform01(o_mcg128_next_u64(&r));
        in[3] = TWOPI_ * o_uniform01(o_mcg128_next_u64(&r));
        double x = o_uniform01(o_mcg128_next_u64(&r));
        in[4] = acos(log(exp(kappa) - 2. * x * sinh(kappa)) / kappa);
        if (isnan(in[4])) return -1;
        double* p = out + 5 * i;
        o_particle_new(in, bs, p);
        double cp = cos(p[3]), sp = sin(p[3]), ct = cos(p[4]), st = sin(p[4]);
        double v[3] = {st * cp, st * sp, ct};
        double t0, tv[3], r0, w[3];
        qmul(q0, qv, 0.0, v, &t0, tv);
        qmul(t0, tv, q0, qc, &r0, w);
        o_orientation_from_vector(w, p + 3);
    }
    return 0;
}

/* ------------------------------------------------------------------------ */
/* Particle update: LangevinBuilder chain of src/bin/simulation/mod.rs:323-363 */
/* with the modifiers of langevin_builder/modifiers.rs                       */
/* ------------------------------------------------------------------------ */
static inline double dot3(const double* a, const double* b) {
    return ((0.0 + a[0] * b[0]) + a[1] * b[1]) + a[2] * b[2];
}
static inline void matvec(const double* m, const double* v, double* out) {
    for (int a = 0; a < 3; ++a) out[a] = dot3(m + 3 * a, v);
}
static inline void qmul(const double a0, const double* a, const double b0, const double* b,
                        double* r0, double* r) {
    /* quaternion 0.3.1 mul: (a0 b0 - a.b, b*a0 + a*b0 + a x b) */
    *r0 = a0 * b0 - (a[0] * b[0] + a[1] * b[1] + a[2] * b[2]);
    double c[3] = {a[1] * b[2] - a[2] * b[1], a[2] * b[0] - a[0] * b[2], a[0] * b[1] - a[1] * b[0]};
    for (int i = 0; i < 3; ++i) r[i] = (b[i] * a0 + a[i] * b0) + c[i];
}

/* fields: u[3] W[9] E[9] rho B[3] gradB[9] (B before the mu_dd factor). */
void o_update_one(const ss_params* p, const double part[5], const double fields[34],
                  const double rv[5], double out[5]) {
    const double* u = fields;
    const double* W = fields + 3;
    const double* E = fields + 12;
    double rho = fields[21];
    const double* Bc = fields + 22;
    const double* GB = fields + 25;
    double dt = p->timestep;

    double cp = cos(part[3]), sp = sin(part[3]), ct = cos(part[4]), st = sin(part[4]);
    double pos[3] = {part[0], part[1], part[2]};
    double n[3] = {st * cp, st * sp, ct};
    double dp[3] = {0.0, 0.0, 0.0}, dn[3] = {0.0, 0.0, 0.0};
    double f[3], b[3];

    /* self_propulsion (modifiers.rs:39-45) */
    for (int c = 0; c < 3; ++c) { dp[c] = dp[c] + n[c]; dn[c] = dn[c] + 0.0; }
    /* convection (49-55) */
    for (int c = 0; c < 3; ++c) { dp[c] = dp[c] + u[c]; dn[c] = dn[c] + 0.0; }
    /* magnetic_dipole_dipole_force (59-69) + mean_force (magnetic_interaction/mod.rs:13-17) */
    matvec(GB, n, f);
    for (int c = 0; c < 3; ++c) { dp[c] = dp[c] + f[c] * p->magnetic_drag; dn[c] = dn[c] + 0.0; }
    /* external_field_alignment (103-118) */
    {
        double d[3];
        field_dir(p, d);
        for (int c = 0; c < 3; ++c) b[c] = d[c] * p->magnetic_reorientation;
        if (p->field_direction[0] == 0.0 && p->field_direction[1] == 0.0 &&
            p->field_direction[2] == 0.0) {
            b[0] = 0.0;
            b[1] = p->magnetic_reorientation;
            b[2] = 0.0;
        }
        double nb = dot3(n, b);
        for (int c = 0; c < 3; ++c) b[c] -= n[c] * nb;
        for (int c = 0; c < 3; ++c) { dp[c] = dp[c] + 0.0; dn[c] = dn[c] + b[c]; }
    }
    /* magnetic_dipole_dipole_rotation (173-186), b = Re B(cell) * mu_dd */
    {
        for (int c = 0; c < 3; ++c) b[c] = Bc[c] * p->magnetic_dipole_dipole;
        double nb = dot3(n, b);
        for (int c = 0; c < 3; ++c) b[c] -= n[c] * nb;
        for (int c = 0; c < 3; ++c) { dp[c] = dp[c] + 0.0; dn[c] = dn[c] + b[c]; }
    }
    /* jeffrey_vorticity (123-142) */
    matvec(W, n, f);
    for (int c = 0; c < 3; ++c) { dp[c] = dp[c] + 0.0; dn[c] = dn[c] + f[c] * (-1.0); }
    /* jeffrey_strain (147-169) */
    {
        matvec(E, n, f);
        for (int c = 0; c < 3; ++c) f[c] *= p->shape;
        double fn = dot3(f, n);
        for (int c = 0; c < 3; ++c) f[c] -= n[c] * fn;
        for (int c = 0; c < 3; ++c) { dp[c] = dp[c] + 0.0; dn[c] = dn[c] + f[c]; }
    }
    /* step (langevin_builder/mod.rs:114-119) */
    for (int c = 0; c < 3; ++c) { dp[c] *= dt; dn[c] *= dt; }
    /* translational_diffusion (89-99) with the std of mod.rs:346 */
    {
        double diff = sqrt(2. * dt * (p->diffusion_translational + p->volume_exclusion * rho));
        for (int c = 0; c < 3; ++c) { dp[c] = dp[c] + rv[c] * diff; dn[c] = dn[c] + 0.0; }
    }
    /* rotational_diffusion (189-231) */
    {
        double ca = cos(rv[3]), sa = sin(rv[3]);
        double ax[3] = {cp * ct * sa - ca * sp, ca * cp + ct * sa * sp, -sa * st};
        double half = rv[4] / 2.0;
        double q0 = cos(half), sh = sin(half);
        double qv[3] = {ax[0] * sh, ax[1] * sh, ax[2] * sh};
        double t0, tv[3], r0, rv3[3];
        qmul(q0, qv, 0.0, n, &t0, tv);
        double qc[3] = {-qv[0], -qv[1], -qv[2]};
        qmul(t0, tv, q0, qc, &r0, rv3);
        for (int c = 0; c < 3; ++c) { dp[c] = dp[c] + 0.0; dn[c] = dn[c] + (rv3[c] - n[c]); }
    }
    /* finalize (langevin_builder/mod.rs:121-125) */
    double np_[3], nv[3], ang[2];
    for (int c = 0; c < 3; ++c) { np_[c] = pos[c] + dp[c]; nv[c] = n[c] + dn[c]; }
    o_orientation_from_vector(nv, ang);
    out[0] = o_modulo(np_[0], p->box_size.x);
    out[1] = o_modulo(np_[1], p->box_size.y);
    out[2] = o_modulo(np_[2], p->box_size.z);
    o_ang_pbc(ang[0], ang[1], out + 3);
    if (p->flags & SS_FLAG_QUASI2D) {
        out[2] = 0.0;
        out[4] = PI_ / 2.;
    }
}

/* ------------------------------------------------------------------------ */
/* Simulation: src/bin/simulation/mod.rs                                     */
/* ------------------------------------------------------------------------ */
struct osim {
    ss_params p;
    osolver s;
    uint64_t n;
    double* particles; /* N x 5 */
    double* noise;     /* N x 5 */
    int have_noise;
    uint32_t* counts;  /* [Ns][Na] */
    double* dist;      /* [Ns][Na] */
    double* sigma;     /* 9 complex */
    double* u;         /* 3 complex */
    double* g;         /* 9 complex */
    double* b;         /* 3 complex */
    double* gb;        /* 9 complex */
    uint64_t timestep;
    int initialised;
    o_pcg32* rngs; /* SS_NOISE_REFERENCE: one stream per chunk */
    int nrng;
};

void osim_set_threads(int n) {
#ifdef _OPENMP
    if (n > 0) omp_set_num_threads(n);
#else
    (void)n;
#endif
}

osim* osim_create(const ss_params* p) {
    osim* s = calloc(1, sizeof(osim));
    if (!s) return NULL;
    s->p = *p;
    solver_init(&s->s, p);
    s->n = p->number_of_particles;
    size_t nb = s->s.ns * s->s.na;
    s->particles = malloc(sizeof(double) * 5 * s->n);
    s->noise = calloc(5 * s->n, sizeof(double));
    s->counts = calloc(nb, sizeof(uint32_t));
    s->dist = calloc(nb, sizeof(double));
    s->sigma = calloc(2 * 9 * s->s.ns, sizeof(double));
    s->u = calloc(2 * 3 * s->s.ns, sizeof(double));
    s->g = calloc(2 * 9 * s->s.ns, sizeof(double));
    s->b = calloc(2 * 3 * s->s.ns, sizeof(double));
    s->gb = calloc(2 * 9 * s->s.ns, sizeof(double));
    if (p->noise_mode == SS_NOISE_REFERENCE) {
        /* Simulation::new (mod.rs:111-128): T streams, all seeded alike */
        s->nrng = p->noise_streams;
        s->rngs = malloc(sizeof(o_pcg32) * (s->nrng > 0 ? s->nrng : 1));
        for (int c = 0; c < s->nrng && s->rngs; ++c) o_pcg32_seed_from_u64(p->seed, s->rngs + c);
    }
    if (!s->particles || !s->noise || !s->counts || !s->dist || !s->sigma || !s->u || !s->g ||
        !s->b || !s->gb || (p->noise_mode == SS_NOISE_REFERENCE && (s->nrng < 1 || !s->rngs))) {
        osim_destroy(s);
        return NULL;
    }
    return s;
}

void osim_destroy(osim* s) {
    if (!s) return;
    free(s->particles);
    free(s->noise);
    free(s->counts);
    free(s->dist);
    free(s->sigma);
    free(s->u);
    free(s->g);
    free(s->b);
    free(s->gb);
    free(s->rngs);
    solver_free(&s->s);
    free(s);
}

static void sample(osim* s) {
    o_histogram(&s->p, s->particles, s->n, s->counts);
    o_sample_from(&s->p, s->counts, s->n, s->dist);
}

int osim_init(osim* s, const double* particles, uint64_t n) {
    if (n != s->n) return -1;
    for (uint64_t i = 0; i < n; ++i) o_particle_new(particles + 5 * i, &s->p.box_size, s->particles + 5 * i);
    sample(s);
    s->initialised = 1;
    return 0;
}

int osim_set_noise(osim* s, const double* rv, uint64_t n) {
    if (n != s->n) return -1;
    memcpy(s->noise, rv, sizeof(double) * 5 * n);
    s->have_noise = 1;
    return 0;
}

static int magnetic_needed(const ss_params* p) {
    return (p->flags & SS_FLAG_MAGNETIC_ALWAYS) || p->magnetic_dipole_dipole != 0.0 ||
           p->magnetic_drag != 0.0;
}

static void one_step(osim* s) {
    const ss_params* p = &s->p;
    const osolver* sv = &s->s;
    sample(s);
    double dr = sqrt(2. * p->diffusion_rotational * p->timestep);
    if (p->noise_mode == SS_NOISE_PHILOX) {
#pragma omp parallel for schedule(static)
        for (long long i = 0; i < (long long)s->n; ++i)
            o_noise(p->seed, s->timestep, (uint64_t)i, dr, s->noise + 5 * i);
    } else if (p->noise_mode == SS_NOISE_REFERENCE) {
        o_ref_noise(s->rngs, s->nrng, s->n, dr, s->noise);
    }
    if (magnetic_needed(p)) magnetic_core(sv, s->dist, s->b, s->gb);
    flow_core(sv, p->hydro_screening, s->dist, s->sigma, s->u, s->g);

    /* W and E as full complex arrays (src/bin/simulation/mod.rs:302-308) */
    size_t ns = sv->ns;
    double* W = malloc(sizeof(double) * 2 * 9 * ns);
    double* E = malloc(sizeof(double) * 2 * 9 * ns);
    for (int a = 0; a < 3; ++a)
        for (int bb = 0; bb < 3; ++bb)
            for (size_t c = 0; c < ns; ++c)
                for (int z = 0; z < 2; ++z) {
                    double gab = s->g[2 * ((size_t)(3 * a + bb) * ns + c) + z];
                    double gba = s->g[2 * ((size_t)(3 * bb + a) * ns + c) + z];
                    W[2 * ((size_t)(3 * a + bb) * ns + c) + z] = (gab - gba) * 0.5;
                    E[2 * ((size_t)(3 * a + bb) * ns + c) + z] = (gab + gba) * 0.5;
                }
    /* density (mod.rs:312-316), serial */
    ogw gw = grid_width(p);
    double dthph = gw.theta * gw.phi;
    double* dens = malloc(sizeof(double) * ns);
    for (size_t c = 0; c < ns; ++c) {
        double acc = 0.0;
        for (size_t a = 0; a < sv->na; ++a) acc = acc + s->dist[c * sv->na + a];
        dens[c] = acc * dthph;
    }
    int nx = sv->nx, ny = sv->ny, nz = sv->nz;
#pragma omp parallel for schedule(static)
    for (long long i = 0; i < (long long)s->n; ++i) {
        double* pp = s->particles + 5 * i;
        /* get_cell_index (mesh/mod.rs:44-66) */
        int64_t ix = (int64_t)floor(pp[0] / gw.x), iy = (int64_t)floor(pp[1] / gw.y),
                iz = (int64_t)floor(pp[2] / gw.z);
        if (ix == nx) ix -= 1;
        if (iy == ny) iy -= 1;
        if (iz == nz) iz -= 1;
        size_t c = ((size_t)ix * ny + iy) * nz + iz;
        double fields[34];
        for (int q = 0; q < 3; ++q) fields[q] = s->u[2 * ((size_t)q * ns + c)];
        for (int q = 0; q < 9; ++q) fields[3 + q] = W[2 * ((size_t)q * ns + c)];
        for (int q = 0; q < 9; ++q) fields[12 + q] = E[2 * ((size_t)q * ns + c)];
        fields[21] = dens[c];
        for (int q = 0; q < 3; ++q) fields[22 + q] = s->b[2 * ((size_t)q * ns + c)];
        for (int q = 0; q < 9; ++q) fields[25 + q] = s->gb[2 * ((size_t)q * ns + c)];
        double out[5];
        o_update_one(p, pp, fields, s->noise + 5 * i, out);
        memcpy(pp, out, sizeof(out));
    }
    free(W);
    free(E);
    free(dens);
    s->timestep += 1;
}

int osim_step(osim* s, uint64_t nsteps) {
    if (!s->initialised) return -3;
    if (s->p.noise_mode == SS_NOISE_EXTERNAL && !s->have_noise) return -3;
    for (uint64_t k = 0; k < nsteps; ++k) one_step(s);
    return 0;
}

uint64_t osim_get_timestep(const osim* s) { return s->timestep; }

/* Snapshot::rng_state (mod.rs:80-86, 199-217): (state, increment) per
   stream.  Setting copies min(count, T) streams, like the reference's zip. */
int osim_get_rng_state(const osim* s, uint64_t* st, int count) {
    if (count < s->nrng) return -1;
    for (int c = 0; c < s->nrng; ++c) {
        st[2 * c] = s->rngs[c].state;
        st[2 * c + 1] = s->rngs[c].inc;
    }
    return s->nrng;
}

int osim_set_rng_state(osim* s, const uint64_t* st, int count) {
    for (int c = 0; c < s->nrng && c < count; ++c) {
        s->rngs[c].state = st[2 * c];
        s->rngs[c].inc = st[2 * c + 1];
    }
    return 0;
}

int osim_set_timestep(osim* s, uint64_t t) {
    s->timestep = t;
    return 0;
}

int osim_get_particles(const osim* s, double* out, uint64_t n) {
    if (n > s->n) return -1;
    memcpy(out, s->particles, sizeof(double) * 5 * n);
    return 0;
}

int osim_get_counts(const osim* s, uint32_t* counts) {
    memcpy(counts, s->counts, sizeof(uint32_t) * s->s.ns * s->s.na);
    return 0;
}

int osim_get_distribution(const osim* s, double* dist) {
    memcpy(dist, s->dist, sizeof(double) * s->s.ns * s->s.na);
    return 0;
}

int osim_get_flow_field(const osim* s, double* u, double* grad_u) {
    if (u) real_parts(s->u, u, 3 * s->s.ns);
    if (grad_u) real_parts(s->g, grad_u, 9 * s->s.ns);
    return 0;
}

int osim_get_magnetic_field(const osim* s, double* b, double* grad_b) {
    if (b) real_parts(s->b, b, 3 * s->s.ns);
    if (grad_b) real_parts(s->gb, grad_b, 9 * s->s.ns);
    return 0;
}

int osim_get_noise(const osim* s, double* rv) {
    memcpy(rv, s->noise, sizeof(double) * 5 * s->n);
    return 0;
}
