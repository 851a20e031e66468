// ss_device.h -- device-side building blocks shared by the kernels of libss:
// Philox4x32-10 noise, the periodic wraps, the binning rule and the particle
// update (the LangevinBuilder chain of src/bin/simulation/mod.rs:323-363).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace ss {

template <typename T>
struct cplx {
    T x, y;
};

template <typename T>
struct Math;

template <>
struct Math<double> {
    static constexpr double PI = 3.141592653589793;
    static constexpr double TWOPI = 2.0 * 3.141592653589793;
    __device__ static inline double sin(double a) { return ::sin(a); }
    __device__ static inline double cos(double a) { return ::cos(a); }
    __device__ static inline void sincos(double a, double* s, double* c) { ::sincos(a, s, c); }
    __device__ static inline double atan2(double y, double x) { return ::atan2(y, x); }
    __device__ static inline double sqrt(double a) { return ::sqrt(a); }
    __device__ static inline double log(double a) { return ::log(a); }
    __device__ static inline double floor(double a) { return ::floor(a); }
    __device__ static inline double fmod(double a, double b) { return ::fmod(a, b); }
    __device__ static inline double fabs(double a) { return ::fabs(a); }
    __device__ static inline void sincos_angle(double a, double* s, double* c);
    __device__ static inline double atan2_angle(double y, double x) { return ::atan2(y, x); }
};

// sin and cos of an angle in [-pi/4, 2 pi + pi/4] (particle angles, axis
// angles) in binary32: quadrant k = rint(x 2/pi), r = x - k pi/2 with pi/2
// split in three binary32 parts (Cody and Waite), and the Taylor polynomials
// of sin (degree 9) and cos (degree 10) on |r| <= pi/4, whose truncation
// errors (1.7e-9, 1.1e-10) are far below an ulp; measured within 0.6 ulp of
// 1 on [0, 2 pi].
__device__ inline void sincos_angle_f32(float x, float* s, float* c) {
#pragma clang fp contract(on)
    const float k = __builtin_rintf(x * 0.636619772367581343f);
    float r = __builtin_fmaf(-k, 1.57079637050628662109375f, x);
    r = __builtin_fmaf(-k, -4.37113882867379290e-08f, r);
    r = __builtin_fmaf(-k, -1.71512451000588187e-15f, r);
    const float z = r * r;
    float ps = __builtin_fmaf(z, 2.75573192239858907e-06f, -1.98412698412698413e-04f);
    ps = __builtin_fmaf(z, ps, 8.33333333333333333e-03f);
    ps = __builtin_fmaf(z, ps, -1.66666666666666667e-01f);
    const float sr = __builtin_fmaf(r * z, ps, r);
    float pc = __builtin_fmaf(z, -2.75573192239858907e-07f, 2.48015873015873016e-05f);
    pc = __builtin_fmaf(z, pc, -1.38888888888888889e-03f);
    pc = __builtin_fmaf(z, pc, 4.16666666666666667e-02f);
    pc = __builtin_fmaf(z, pc, -0.5f);
    const float cr = __builtin_fmaf(z, pc, 1.0f);
    const int q = (int)k;
    const float a = (q & 1) ? cr : sr, b = (q & 1) ? sr : cr;
    *s = (q & 2) ? -a : a;
    *c = ((q + 1) & 2) ? -b : b;
}

// atan2(y, x) in binary32 with C's signed-zero cases: a = min/max of |x|, |y|
// (reciprocal), atan(a) for a > tan(pi/8) as pi/4 + atan((a - 1)/(a + 1)),
// and atan(t) = t + t^3 P(t^2) on |t| <= tan(pi/8) with P of degree 4 fitted
// by least squares on Chebyshev nodes (relative error 2.6e-9); within 1.3
// ulp of the result on finite inputs.
__device__ inline float atan2_angle_f32(float y, float x) {
#pragma clang fp contract(on)
    const float ax = __builtin_fabsf(x), ay = __builtin_fabsf(y);
    const float mx = __builtin_fmaxf(ax, ay), mn = __builtin_fminf(ax, ay);
    const float a = mx > 0.0f ? mn * __builtin_amdgcn_rcpf(mx) : 0.0f;
    const bool big = a > 0.414213562373095049f;
    const float t = big ? (a - 1.0f) * __builtin_amdgcn_rcpf(a + 1.0f) : a;
    const float z = t * t;
    float p = __builtin_fmaf(z, -0.06459424011434889f, 0.10746316861293508f);
    p = __builtin_fmaf(z, p, -0.14264234924234587f);
    p = __builtin_fmaf(z, p, 0.19999550188924334f);
    p = __builtin_fmaf(z, p, -0.3333333181487349f);
    float r = __builtin_fmaf(t * z, p, t);
    if (big) r = r + 0.785398163397448310f;
    if (ay > ax) r = 1.57079632679489662f - r;
    if (__builtin_signbit(x)) r = 3.14159265358979324f - r;
    return __builtin_signbit(y) ? -r : r;
}

// sin and cos of an angle in binary64, for the angles of the path (particle
// angles, axis angles, half rotation angles, 2 pi u of the noise; |x| up to
// a few times 2 pi): quadrant k = rint(x 2/pi) and r = x - k pi/2 as a
// double-length pair rh + rl, with pi/2 split in three parts (Cody and
// Waite; the first two of 31 and 32 significant bits, so that k P1 and k P2
// are exact for |k| < 2^21, the second difference taken exactly by
// Knuth's two-sum); then the Taylor polynomials of sin (to r^17) and cos (to
// r^16) on |r| <= pi/4, whose truncation errors (8e-20, 3e-18) are far below
// an ulp, with the first-order term of rl added to each.  Within 1 ulp of
// the exact values.  Beyond |x| = 2^20 (never reached by the path) the
// library.
__host__ __device__ inline void sincos_angle_f64(double x, double* s, double* c) {
#pragma clang fp contract(on)
    if (!(__builtin_fabs(x) < 1048576.0)) {
        ::sincos(x, s, c);
        return;
    }
    const double k = __builtin_rint(x * 0.6366197723675814);
    const double r1 = __builtin_fma(-k, 1.5707963267341256, x);  // exact
    const double w = k * 6.077100506303966e-11;                    // exact
    const double rh = r1 - w;
    const double bb = rh - r1;
    const double rl = ((r1 - (rh - bb)) - (w + bb)) - k * 2.0222662487959506e-21;
    const double z = rh * rh;
    double ps = __builtin_fma(z, 2.8114572543455206e-15, -7.647163731819816e-13);
    ps = __builtin_fma(z, ps, 1.6059043836821613e-10);
    ps = __builtin_fma(z, ps, -2.505210838544172e-08);
    ps = __builtin_fma(z, ps, 2.7557319223985893e-06);
    ps = __builtin_fma(z, ps, -0.0001984126984126984);
    ps = __builtin_fma(z, ps, 0.008333333333333333);
    ps = __builtin_fma(z, ps, -0.16666666666666666);
    // sin(rh + rl) = sin rh + rl cos rh; cos rh = 1 - z/2 to the order kept
    const double sr = rh + __builtin_fma(rh * z, ps, __builtin_fma(-0.5 * z, rl, rl));
    double pc = __builtin_fma(z, 4.779477332387385e-14, -1.1470745597729725e-11);
    pc = __builtin_fma(z, pc, 2.08767569878681e-09);
    pc = __builtin_fma(z, pc, -2.755731922398589e-07);
    pc = __builtin_fma(z, pc, 2.48015873015873e-05);
    pc = __builtin_fma(z, pc, -0.001388888888888889);
    pc = __builtin_fma(z, pc, 0.041666666666666664);
    // cos(rh + rl) = cos rh - rl sin rh; 1 - z/2 split so that the large
    // part is exact: hz = z/2 rounded, its rounding kept in the tail
    const double hz = 0.5 * z;
    const double one_hz = 1.0 - hz;
    const double tail = ((1.0 - one_hz) - hz) + (__builtin_fma(rh, rh, -z) * -0.5);
    const double cr = one_hz + (__builtin_fma(z * z, pc, tail) - rl * rh);
    const int q = (int)(long long)k;
    const double a = (q & 1) ? cr : sr, b = (q & 1) ? sr : cr;
    *s = x == 0.0 ? x : ((q & 2) ? -a : a);
    *c = ((q + 1) & 2) ? -b : b;
}

// ln u for u in (0, 1] (the uniforms of the noise, at least 2^-53): u = 2^e m
// with m = 1 + f in [sqrt(1/2), sqrt(2)) (f exact), s = f/(2 + f) in
// [-0.172, 0.172], ln m = 2 atanh s = 2s + s Q(s^2), Q = 2s^2/3 + 2s^4/5 +
// ... to s^20 (truncation 2e-19), and with 2s = f - s f, ln m = f - s (f -
// Q): the exact f carries the result and the rounded s enters only a
// correction a sixth its size.  ln 2 split in two parts (the first of 42
// bits, so that e ln2_hi is exact).  Within 1 ulp.  Other arguments: the
// library.
__host__ __device__ inline double log_unit_f64(double u) {
#pragma clang fp contract(on)
    if (!(u >= 2.2250738585072014e-308 && u <= 1.0)) return ::log(u);
    const unsigned long long bits = __builtin_bit_cast(unsigned long long, u);
    int e = (int)((bits >> 52) & 0x7FFull) - 1022;
    double m = __builtin_bit_cast(double, (bits & 0x000FFFFFFFFFFFFFull) | 0x3FE0000000000000ull);
    if (m < 0.7071067811865476) {
        m = m * 2.0;
        e -= 1;
    }
    const double f = m - 1.0;
    const double s = f / (2.0 + f);
    const double z = s * s;
    double p = __builtin_fma(z, 0.09523809523809523, 0.10526315789473684);
    p = __builtin_fma(z, p, 0.11764705882352941);
    p = __builtin_fma(z, p, 0.13333333333333333);
    p = __builtin_fma(z, p, 0.15384615384615385);
    p = __builtin_fma(z, p, 0.18181818181818182);
    p = __builtin_fma(z, p, 0.2222222222222222);
    p = __builtin_fma(z, p, 0.2857142857142857);
    p = __builtin_fma(z, p, 0.4);
    p = __builtin_fma(z, p, 0.6666666666666666);
    const double Q = z * p;
    const double de = (double)e;
    const double lm = f - (s * (f - Q) - de * 5.497923018708371e-14);
    return __builtin_fma(de, 0.6931471805598903, lm);
}

__device__ inline void Math<double>::sincos_angle(double a, double* s, double* c) { sincos_angle_f64(a, s, c); }

template <>
struct Math<float> {
    static constexpr float PI = 3.14159265358979323846f;
    static constexpr float TWOPI = 6.28318530717958647692f;
    __device__ static inline float sin(float a) { return ::sinf(a); }
    __device__ static inline float cos(float a) { return ::cosf(a); }
    __device__ static inline void sincos(float a, float* s, float* c) { ::sincosf(a, s, c); }
    __device__ static inline float atan2(float y, float x) { return ::atan2f(y, x); }
    __device__ static inline float sqrt(float a) { return ::sqrtf(a); }
    __device__ static inline float log(float a) { return ::logf(a); }
    __device__ static inline float floor(float a) { return ::floorf(a); }
    __device__ static inline float fmod(float a, float b) { return ::fmodf(a, b); }
    __device__ static inline float fabs(float a) { return ::fabsf(a); }
    __device__ static inline void sincos_angle(float a, float* s, float* c) { sincos_angle_f32(a, s, c); }
    __device__ static inline float atan2_angle(float y, float x) { return atan2_angle_f32(y, x); }
};

// ---------------------------------------------------------------------------
// Philox4x32-10 (Salmon, Moraes, Dror, Shaw, SC'11).
// ---------------------------------------------------------------------------

// a ^ b ^ k; on the device one gfx950 v_bitop3_b32 with truth table 0x96
// (odd parity of the three inputs).  k must be uniform across the wavefront
// (the Philox key: the seed and round constants), so that it can be the
// instruction's scalar operand; the compiler otherwise splits the three-way
// xor into two two-way ones.  (The host pass, which draws the initial
// conditions, uses the plain expression.)
__host__ __device__ inline uint32_t xor3_key(uint32_t a, uint32_t b, uint32_t k) {
#if defined(__HIP_DEVICE_COMPILE__)
    uint32_t r;
    asm("v_bitop3_b32 %0, %1, %2, %3 bitop3:0x96" : "=v"(r) : "v"(a), "v"(b), "s"(k));
    return r;
#else
    return (a ^ b) ^ k;
#endif
}

__host__ __device__ inline void philox4x32_10(uint32_t c0, uint32_t c1, uint32_t c2, uint32_t c3, uint32_t k0,
                                     uint32_t k1, uint32_t out[4]) {
#pragma unroll
    for (int r = 0; r < 10; ++r) {
        if (r > 0) {
            k0 += 0x9E3779B9u;
            k1 += 0xBB67AE85u;
        }
        uint64_t p0 = (uint64_t)0xD2511F53u * c0;
        uint64_t p1 = (uint64_t)0xCD9E8D57u * c2;
        uint32_t n0 = xor3_key((uint32_t)(p1 >> 32), c1, k0);
        uint32_t n2 = xor3_key((uint32_t)(p0 >> 32), c3, k1);
        c1 = (uint32_t)p1;
        c3 = (uint32_t)p0;
        c0 = n0;
        c2 = n2;
    }
    out[0] = c0;
    out[1] = c1;
    out[2] = c2;
    out[3] = c3;
}

// Noise of one particle and step.  sa, ca: sine and cosine of the rotation
// axis angle, so that generators that draw the angle as a fraction of a turn
// can evaluate them directly.
template <typename T>
struct Noise {
    T x, y, z;  // standard normals
    T sa, ca;   // sin and cos of axis_angle
    T rot;      // rotate_angle
};

// Philox layout, per particle and step: counter (j, id, step lo, step hi),
// key (seed lo, seed hi).  Call j = 0 gives four words w0..w3; six 21-bit
// fields f0..f5 are the top 21 bits of w0..w3 and, for f4 and f5, the low 11
// bits of w0|w1 and w2|w3 (11 + 10 bits).  31 more bits from the top of the
// words of calls j = 1, 2 (in order) extend each f_k to a 52-bit integer m_k,
// and u_k = (m_k + 1/2) 2^-52, strictly inside (0, 1) (smallest 2^-53, so the
// normals reach 8.6 sigma).  Both precisions draw these same uniforms: FP64
// exactly, FP32 rounded to binary32 (within one ulp); 1 - u5 is the uniform
// of the complemented integer, exact in both.  The transforms differ only in
// the transcendentals (libm in FP64, the hardware's FP32 log2, sin and cos of
// turns in FP32), so the FP32 increments equal the FP64 ones to about 1e-6
// relative.  Transform: Box-Muller normals x, y from (u0, u1), z from (u2,
// u3), axis angle 2 pi u4, rotate angle dr sqrt(-2 ln(1 - u5)) -- the
// RandomVector of src/bin/simulation/mod.rs:51-57,259-286,383-385.
__host__ __device__ inline void philox_fields(const uint32_t w[4], uint32_t f[6]) {
    f[0] = w[0] >> 11;
    f[1] = w[1] >> 11;
    f[2] = w[2] >> 11;
    f[3] = w[3] >> 11;
    f[4] = ((w[0] & 0x7FFu) << 10) | ((w[1] & 0x7FFu) >> 1);
    f[5] = ((w[2] & 0x7FFu) << 10) | ((w[3] & 0x7FFu) >> 1);
}

__device__ inline void philox_noise(uint64_t seed, uint64_t step, uint32_t id, double dr, Noise<double>& nz) {
    uint32_t w[12];
    const uint32_t k0 = (uint32_t)seed, k1 = (uint32_t)(seed >> 32);
    const uint32_t s0 = (uint32_t)step, s1 = (uint32_t)(step >> 32);
    philox4x32_10(0u, id, s0, s1, k0, k1, w);
    philox4x32_10(1u, id, s0, s1, k0, k1, w + 4);
    philox4x32_10(2u, id, s0, s1, k0, k1, w + 8);
    uint32_t f[6];
    philox_fields(w, f);
    double u[6];
#pragma unroll
    for (int k = 0; k < 6; ++k) {
        const uint64_t m = ((uint64_t)f[k] << 31) | (uint64_t)(w[4 + k] >> 1);
        u[k] = ((double)m + 0.5) * (1.0 / 4503599627370496.0);
    }
    const double TWOPI = Math<double>::TWOPI;
    const double r0 = ::sqrt(-2.0 * log_unit_f64(u[0]));
    const double r1 = ::sqrt(-2.0 * log_unit_f64(u[2]));
    double sn1, cs1, sn3, cs3;
    sincos_angle_f64(TWOPI * u[1], &sn1, &cs1);
    sincos_angle_f64(TWOPI * u[3], &sn3, &cs3);
    nz.x = r0 * cs1;
    nz.y = r0 * sn1;
    nz.z = r1 * cs3;
    sincos_angle_f64(TWOPI * u[4], &nz.sa, &nz.ca);
    nz.rot = dr * ::sqrt(-2.0 * log_unit_f64(1.0 - u[5]));
}

// (m + 1/2) 2^-52 for m = f 2^31 + t (f: 21 bits, t: 31 bits) in binary32:
// f 2^-21 is exact, the tail is rounded once before the fused sum, so the
// result is within one ulp of the exact value.
__device__ inline float uniform52_f32(uint32_t f, uint32_t t) {
    return __builtin_fmaf((float)t + 0.5f, 2.220446049250313080847e-16f, (float)f * 4.76837158203125e-07f);
}

__device__ inline void philox_noise(uint64_t seed, uint64_t step, uint32_t id, float dr, Noise<float>& nz) {
    uint32_t w[12];
    const uint32_t k0 = (uint32_t)seed, k1 = (uint32_t)(seed >> 32);
    const uint32_t s0 = (uint32_t)step, s1 = (uint32_t)(step >> 32);
    philox4x32_10(0u, id, s0, s1, k0, k1, w);
    philox4x32_10(1u, id, s0, s1, k0, k1, w + 4);
    philox4x32_10(2u, id, s0, s1, k0, k1, w + 8);
    uint32_t f[6];
    philox_fields(w, f);
    float u[5];
#pragma unroll
    for (int k = 0; k < 5; ++k) u[k] = uniform52_f32(f[k], w[4 + k] >> 1);
    // 1 - u5: the complement of the 52-bit integer
    const float u5c = uniform52_f32(~f[5] & 0x1FFFFFu, ~(w[9] >> 1) & 0x7FFFFFFFu);
    // hardware transcendentals: v_log_f32 is log2, v_sin/v_cos take turns
    const float m2ln2 = -1.38629436111989061883f;  // -2 ln 2
    const float r0 = __builtin_amdgcn_sqrtf(m2ln2 * __builtin_amdgcn_logf(u[0]));
    const float r1 = __builtin_amdgcn_sqrtf(m2ln2 * __builtin_amdgcn_logf(u[2]));
    nz.x = r0 * __builtin_amdgcn_cosf(u[1]);
    nz.y = r0 * __builtin_amdgcn_sinf(u[1]);
    nz.z = r1 * __builtin_amdgcn_cosf(u[3]);
    nz.sa = __builtin_amdgcn_sinf(u[4]);
    nz.ca = __builtin_amdgcn_cosf(u[4]);
    nz.rot = dr * __builtin_amdgcn_sqrtf(m2ln2 * __builtin_amdgcn_logf(u5c));
}

// Caller-supplied RandomVector [x, y, z, axis_angle, rotate_angle].
template <typename T>
__device__ inline void noise_from_vector(const T rv[5], Noise<T>& nz) {
    nz.x = rv[0];
    nz.y = rv[1];
    nz.z = rv[2];
    Math<T>::sincos_angle(rv[3], &nz.sa, &nz.ca);
    nz.rot = rv[4];
}

// ---------------------------------------------------------------------------
// Periodic wraps: modulo and ang_pbc (particle.rs:26-44).
// ---------------------------------------------------------------------------
template <typename T>
__device__ inline T modulo(T f, T m) {
    T r = Math<T>::fmod(f, m);
    return r < T(0) ? r + Math<T>::fabs(m) : r;
}

// modulo for |f| < 2|m| without the division inside fmod; the same value as
// modulo() (fmod is exact, and so is f - m for m <= f < 2m).
template <typename T>
__device__ inline T modulo_near(T f, T m) {
    if (f >= T(0) && f < m) return f;
    if (f >= m) {
        const T r = f - m;
        return r < m ? r : modulo(f, m);
    }
    if (f < T(0)) {
        if (f >= -m) return f + m;
        return modulo(f, m);
    }
    return f;
}

// ang_pbc of angles just produced by Orientation::from_vector: theta =
// pi/2 - atan2(z, rxy) lies in [0, pi] and phi = atan2(y, x) in [-pi, pi], so
// ang_pbc (particle.rs:37-44) reduces to wrapping phi once.
template <typename T>
__device__ inline void ang_pbc_from_vector(T phi, T theta, T& ophi, T& otheta) {
    ophi = phi < T(0) ? phi + Math<T>::TWOPI : phi;
    otheta = theta;
}

template <typename T>
__device__ inline void ang_pbc(T phi, T theta, T& ophi, T& otheta) {
    const T PI = Math<T>::PI, TWOPI = Math<T>::TWOPI;
    T t = modulo(theta, TWOPI);
    if (t > PI) {
        ophi = modulo(phi + PI, TWOPI);
        otheta = TWOPI - t;
    } else {
        ophi = modulo(phi, TWOPI);
        otheta = t;
    }
}

// Binning rule of Distribution::coord_to_grid / get_cell_index
// (distribution/mod.rs:112-137, mesh/mod.rs:49-63): floor(q / g), an index equal
// to n folds to n-1.  The lower clamp only guards memory; pbc keeps q >= 0.
template <typename T>
__device__ inline int grid_index(T q, T g, int n) {
    T f = Math<T>::floor(q / g);
    int i = (int)f;
    if (f >= (T)n) i = n - 1;
    if (i == n) i = n - 1;
    if (f < T(0)) i = 0;
    return i;
}

// The same index without the division: t = q * rg with rg = fl(1/g) is within
// n (2u + u^2) of q/g for q in [0, n g] (u the unit roundoff), and fl(q/g)
// within n u of it, so floor(t) = floor(fl(q/g)) unless t lies within
// tol >= 3 n u of an integer; there the division decides.
// fr = fract(t) = t - floor(t) is exact for t >= 0.  half = fl(1/2 - tol):
// |fl(fr - 1/2)| < half implies tol <= fr <= 1 - tol (rounding is monotonic
// and odd, so fr < tol gives fl(fr - 1/2) <= -half and fr > 1 - tol gives
// fl(fr - 1/2) >= half).
template <typename T>
__device__ inline int grid_index_r(T q, T g, T rg, T half, int n) {
    const T t = q * rg;
    T f = Math<T>::floor(t);
    T fr;
    if constexpr (sizeof(T) == 4)
        fr = __builtin_amdgcn_fractf(t);
    else
        fr = __builtin_amdgcn_fract(t);
    if (!(Math<T>::fabs(fr - T(0.5)) < half)) f = Math<T>::floor(q / g);
    // clamp to [0, n - 1] before the conversion (an index equal to n folds to
    // n - 1)
    if constexpr (sizeof(T) == 4)
        f = __builtin_amdgcn_fmed3f(f, 0.0f, (float)(n - 1));
    else
        f = __builtin_fmin(__builtin_fmax(f, 0.0), (double)(n - 1));
    return (int)f;
}

// ---------------------------------------------------------------------------
// Field record at one cell, as the update reads it.
// ---------------------------------------------------------------------------
enum : int {
    REC_U = 0,   // u[3]
    REC_W = 3,   // W01, W02, W12 (W antisymmetric)
    REC_E = 8,   // E00, E01, E02, E11, E12, E22 (E symmetric)
    REC_RHO = 14,
    REC_B = 15,  // Re B * magnetic_dipole_dipole
    REC_GB = 18, // grad B [a][b]
    REC_SMALL = 8,
    REC_FULL = 32,
};

template <typename T>
struct UpdateParams;

// Whether any term that FULL = false leaves out is on.
template <typename T>
__host__ __device__ inline bool full_physics(const UpdateParams<T>& P);

template <typename T>
struct UpdateParams {
    T dt;
    T drag;            // magnetic_drag
    T shape;
    T vol_excl;        // volume_exclusion
    T kappa;           // magnetic_reorientation
    T dir[3];          // external field direction
    int default_dir;   // 1: reference's y axis, built as [0, kappa, 0]
    T d_trans;         // diffusion.translational
    T dr;              // sqrt(2 D_r dt)
    T box[3];
    int use_e, use_rho, use_b, use_gb;
    int quasi2d;
};

template <typename T>
__host__ __device__ inline bool full_physics(const UpdateParams<T>& P) {
    return P.use_gb || P.kappa != T(0) || P.use_b || P.use_e || P.use_rho || P.quasi2d;
}

// One particle: the LangevinBuilder chain (src/bin/simulation/mod.rs:348-363,
// modifiers.rs).  rec: field record of the particle's cell.  rv: RandomVector.
// Outputs the position increment dp (to be added with pbc) and the new
// orientation angles after ang_pbc.
// sin and cos of the half rotation angle: a short Taylor polynomial in FP32
// (|x| < 0.5, error below 1e-8), the library functions otherwise; in FP64
// the reduced-range sin and cos above.
template <typename T>
__device__ inline void half_angle(T x, T* s, T* c) {
#pragma clang fp contract(on)
    if constexpr (sizeof(T) == 8) {
        sincos_angle_f64(x, s, c);
        return;
    }
    if constexpr (sizeof(T) == 4) {
        if (x < 0.5f && x > -0.5f) {
            const float x2 = x * x;
            *s = x * (1.0f + x2 * (-1.0f / 6.0f + x2 * (1.0f / 120.0f - x2 * (1.0f / 5040.0f))));
            *c = 1.0f + x2 * (-0.5f + x2 * (1.0f / 24.0f + x2 * (-1.0f / 720.0f + x2 * (1.0f / 40320.0f))));
            return;
        }
    }
    Math<T>::sincos(x, s, c);
}

// FULL = false compiles out the terms that need a field or option the
// parameters leave off (magnetic force and rotation, external field, strain,
// density-dependent diffusion, quasi-2D); the caller guarantees they are off.
template <typename T, bool FULL = true>
__device__ inline void langevin(const UpdateParams<T>& P, T phi, T theta, const T* rec, const Noise<T>& rv,
                                T dp[3], T& ophi, T& otheta) {
    // contractions fixed by the source, so that both instantiations (and the
    // ensemble kernel) round alike
#pragma clang fp contract(on)
    T sp, cp, st, ct;
    Math<T>::sincos_angle(phi, &sp, &cp);
    Math<T>::sincos_angle(theta, &st, &ct);
    const T n0 = st * cp, n1 = st * sp, n2 = ct;
    // self_propulsion + convection (modifiers.rs:39-55)
    T p0 = n0 + rec[REC_U + 0];
    T p1 = n1 + rec[REC_U + 1];
    T p2 = n2 + rec[REC_U + 2];
    T o0 = T(0), o1 = T(0), o2 = T(0);
    // magnetic_dipole_dipole_force (modifiers.rs:59-69, magnetic_interaction/mod.rs:13-17)
    if (FULL && P.use_gb) {
        const T* g = rec + REC_GB;
        T f0 = (g[0] * n0 + g[1] * n1) + g[2] * n2;
        T f1 = (g[3] * n0 + g[4] * n1) + g[5] * n2;
        T f2 = (g[6] * n0 + g[7] * n1) + g[8] * n2;
        p0 = p0 + f0 * P.drag;
        p1 = p1 + f1 * P.drag;
        p2 = p2 + f2 * P.drag;
    }
    // external_field_alignment (modifiers.rs:103-118)
    if (FULL && P.kappa != T(0)) {
        T b0, b1, b2;
        if (P.default_dir) {
            b0 = T(0);
            b1 = P.kappa;
            b2 = T(0);
        } else {
            b0 = P.dir[0] * P.kappa;
            b1 = P.dir[1] * P.kappa;
            b2 = P.dir[2] * P.kappa;
        }
        T nb = (n0 * b0 + n1 * b1) + n2 * b2;
        o0 = o0 + (b0 - n0 * nb);
        o1 = o1 + (b1 - n1 * nb);
        o2 = o2 + (b2 - n2 * nb);
    }
    // magnetic_dipole_dipole_rotation (modifiers.rs:173-186)
    if (FULL && P.use_b) {
        const T* b = rec + REC_B;
        T nb = (n0 * b[0] + n1 * b[1]) + n2 * b[2];
        o0 = o0 + (b[0] - n0 * nb);
        o1 = o1 + (b[1] - n1 * nb);
        o2 = o2 + (b[2] - n2 * nb);
    }
    // jeffrey_vorticity (modifiers.rs:123-142): -W n with W antisymmetric
    {
        const T w01 = rec[REC_W + 0], w02 = rec[REC_W + 1], w12 = rec[REC_W + 2];
        T f0 = w01 * n1 + w02 * n2;
        T f1 = -w01 * n0 + w12 * n2;
        T f2 = -w02 * n0 - w12 * n1;
        o0 = o0 - f0;
        o1 = o1 - f1;
        o2 = o2 - f2;
    }
    // jeffrey_strain (modifiers.rs:147-169)
    if (FULL && P.use_e) {
        const T* e = rec + REC_E;
        T f0 = ((e[0] * n0 + e[1] * n1) + e[2] * n2) * P.shape;
        T f1 = ((e[1] * n0 + e[3] * n1) + e[4] * n2) * P.shape;
        T f2 = ((e[2] * n0 + e[4] * n1) + e[5] * n2) * P.shape;
        T fn = (f0 * n0 + f1 * n1) + f2 * n2;
        o0 = o0 + (f0 - n0 * fn);
        o1 = o1 + (f1 - n1 * fn);
        o2 = o2 + (f2 - n2 * fn);
    }
    // step (langevin_builder/mod.rs:114-119)
    p0 *= P.dt;
    p1 *= P.dt;
    p2 *= P.dt;
    o0 *= P.dt;
    o1 *= P.dt;
    o2 *= P.dt;
    // translational_diffusion (modifiers.rs:89-99), std of mod.rs:346
    {
        T diff = (FULL && P.use_rho) ? Math<T>::sqrt(T(2) * P.dt * (P.d_trans + P.vol_excl * rec[REC_RHO]))
                           : Math<T>::sqrt(T(2) * P.dt * P.d_trans);
        p0 = p0 + rv.x * diff;
        p1 = p1 + rv.y * diff;
        p2 = p2 + rv.z * diff;
    }
    // rotational_diffusion (modifiers.rs:189-231): q (0,n) q*, quaternion 0.3.1
    {
        const T sa = rv.sa, ca = rv.ca;
        const T ax0 = cp * ct * sa - ca * sp;
        const T ax1 = ca * cp + ct * sa * sp;
        const T ax2 = -sa * st;
        T sh, q0;
        half_angle<T>(rv.rot / T(2), &sh, &q0);
        const T q1 = ax0 * sh, q2 = ax1 * sh, q3 = ax2 * sh;
        // t = q * (0, n); the Hamilton product's terms with the zero scalar
        // part of (0, n) are left out (they add zeros)
        const T t0 = -(q1 * n0 + q2 * n1 + q3 * n2);
        const T t1 = n0 * q0 + (q2 * n2 - q3 * n1);
        const T t2 = n1 * q0 + (q3 * n0 - q1 * n2);
        const T t3 = n2 * q0 + (q1 * n1 - q2 * n0);
        // r = t * conj(q), vector part
        const T c1 = -q1, c2 = -q2, c3 = -q3;
        const T r1 = (c1 * t0 + t1 * q0) + (t2 * c3 - t3 * c2);
        const T r2 = (c2 * t0 + t2 * q0) + (t3 * c1 - t1 * c3);
        const T r3 = (c3 * t0 + t3 * q0) + (t1 * c2 - t2 * c1);
        o0 = o0 + (r1 - n0);
        o1 = o1 + (r2 - n1);
        o2 = o2 + (r3 - n2);
    }
    // finalize (langevin_builder/mod.rs:121-125, particle.rs:106-121)
    const T v0 = n0 + o0, v1 = n1 + o1, v2 = n2 + o2;
    const T rxy = Math<T>::sqrt(v0 * v0 + v1 * v1);
    const T nphi = Math<T>::atan2_angle(v1, v0);
    const T ntheta = Math<T>::PI / T(2) - Math<T>::atan2_angle(v2, rxy);
    ang_pbc_from_vector(nphi, ntheta, ophi, otheta);
    if (FULL && P.quasi2d) otheta = Math<T>::PI / T(2);
    dp[0] = p0;
    dp[1] = p1;
    dp[2] = p2;
}

// Particle::new (particle.rs:186-201): modulo and ang_pbc, twice, in FP64.
__device__ inline void sanitize(const double* in, const double* box, double* out) {
    double a0, a1, b0, b1;
    out[0] = modulo(in[0], box[0]);
    out[1] = modulo(in[1], box[1]);
    out[2] = modulo(in[2], box[2]);
    ang_pbc(in[3], in[4], a0, a1);
    out[0] = modulo(out[0], box[0]);
    out[1] = modulo(out[1], box[1]);
    out[2] = modulo(out[2], box[2]);
    ang_pbc(a0, a1, b0, b1);
    out[3] = b0;
    out[4] = b1;
}

}  // namespace ss
