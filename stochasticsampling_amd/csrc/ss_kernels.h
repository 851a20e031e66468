// ss_kernels.h -- the per-step kernels of the hot loop (binning, moment
// contraction, k-space solve, field packing, particle update, counting sort).
// Included by ss_sim.hip only.
#pragma once

#include "ss_device.h"

namespace ss {

constexpr uint32_t INVALID_KEY = 0xFFFFFFFFu;

// Geometry and scaling constants of one system, in the working precision.
template <typename T>
struct Geo {
    int nx, ny, nz, np, nt;
    int ns, na;
    T gx, gy, gz, gphi, gtheta;  // GridWidth (grid_width.rs:23-31)
    T box[3];
    // per axis (x, y, z, phi, theta): fl(1 / width) and fl(1/2 - tol) with
    // the tolerance tol = 4 u (n + 1) of grid_index_r
    T rg[5], tol[5];
};

// Distribution::coord_to_grid per axis (distribution/mod.rs:112-137).
template <typename T>
__device__ inline int bin_x(const Geo<T>& g, T q) { return grid_index_r(q, g.gx, g.rg[0], g.tol[0], g.nx); }
template <typename T>
__device__ inline int bin_y(const Geo<T>& g, T q) { return grid_index_r(q, g.gy, g.rg[1], g.tol[1], g.ny); }
template <typename T>
__device__ inline int bin_z(const Geo<T>& g, T q) { return grid_index_r(q, g.gz, g.rg[2], g.tol[2], g.nz); }
template <typename T>
__device__ inline int bin_phi(const Geo<T>& g, T q) { return grid_index_r(q, g.gphi, g.rg[3], g.tol[3], g.np); }
template <typename T>
__device__ inline int bin_theta(const Geo<T>& g, T q) {
    return grid_index_r(q, g.gtheta, g.rg[4], g.tol[4], g.nt);
}

template <typename T>
__device__ inline int cell_of(const Geo<T>& g, T x, T y, T z) {
    return (bin_x(g, x) * g.ny + bin_y(g, y)) * g.nz + bin_z(g, z);
}

template <typename T>
__device__ inline int angle_of(const Geo<T>& g, T phi, T theta) {
    return bin_phi(g, phi) * g.nt + bin_theta(g, theta);
}

// Wave64 run detection on a key that is mostly constant across neighbouring
// lanes (particles sorted by cell).  A run is a maximal stretch of adjacent
// lanes with equal keys.
struct Run {
    int head;     // lane of this lane's run head
    int len;      // length of the run (valid on head lanes)
    bool is_head;
};

__device__ inline Run wave_run(uint32_t key) {
    const int lane = __lane_id();
    uint32_t prev = __shfl_up(key, 1);
    bool head = (lane == 0) || (prev != key);
    unsigned long long heads = __ballot(head);
    unsigned long long le = (lane == 63) ? ~0ull : ((2ull << lane) - 1ull);
    unsigned long long above = heads & ~le;
    Run r;
    r.is_head = head;
    r.len = (above ? __builtin_ctzll(above) : 64) - lane;
    r.head = 63 - __builtin_clzll(heads & le);
    return r;
}

// ---------------------------------------------------------------------------
// Binning, plain path: full 5-D count grid by global atomics
// (Distribution::histogram_from, distribution/mod.rs:146-160).
// ---------------------------------------------------------------------------
template <typename T>
__global__ void __launch_bounds__(256)
    k_bin_hist(const T* __restrict__ x, const T* __restrict__ y, const T* __restrict__ z,
               const T* __restrict__ phi, const T* __restrict__ theta, uint64_t n, Geo<T> g,
               uint32_t* __restrict__ hist) {
    uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    int c = cell_of(g, x[i], y[i], z[i]);
    int a = angle_of(g, phi[i], theta[i]);
    atomicAdd(hist + (size_t)c * g.na + a, 1u);
}

// Contraction of the angular counts of each cell against the moment kernels
// (average_stress, stress.rs:60-96; DirectorField, director.rs:36-61; density,
// src/bin/simulation/mod.rs:312-316).  kern: [nm][na]; out: moments m (times
// scale) as complex with zero imaginary part.  CONTRACT_LANES lanes per cell
// read the cell's row in 16-byte pieces side by side (a wavefront reads eight
// rows' 128-byte stretches, whole cache lines, where one thread per cell
// touched 64 lines per load), sum their pieces in order and combine by a
// fixed butterfly, so the result does not depend on the launch.  The kernels
// sit in LDS.
constexpr int CONTRACT_LANES = 8;
constexpr int CONTRACT_GROUPS = 4;

template <typename T, int NM>
__global__ void __launch_bounds__(256)
    k_contract_hist(const uint32_t* __restrict__ hist, Geo<T> g, const T* __restrict__ kern_g, T scale,
                    T rho_scale, cplx<T>* __restrict__ spec, T* __restrict__ rho) {
    extern __shared__ __attribute__((aligned(16))) unsigned char smem_raw[];
    T* kern = reinterpret_cast<T*>(smem_raw);
    const int na = g.na;
    for (int t = threadIdx.x; t < NM * na; t += blockDim.x) kern[t] = kern_g[t];
    __syncthreads();
    const int part = threadIdx.x % CONTRACT_LANES;
    // CONTRACT_GROUPS groups of 256 / CONTRACT_LANES cells per workgroup (the
    // kernels are loaded once per workgroup)
    for (int grp = 0; grp < CONTRACT_GROUPS; ++grp) {
        const int cell = ((blockIdx.x * CONTRACT_GROUPS + grp) * (int)blockDim.x + (int)threadIdx.x) / CONTRACT_LANES;
        const bool live = cell < g.ns;
        T acc[NM];
        uint32_t tot = 0;
#pragma unroll
        for (int q = 0; q < NM; ++q) acc[q] = T(0);
        auto add = [&](int a, uint32_t n) {
            const T c = (T)n;
            tot += n;
#pragma unroll
            for (int q = 0; q < NM; ++q) acc[q] += kern[q * na + a] * c;
        };
        if (live) {
            const uint32_t* cnt = hist + (size_t)cell * na;
            if ((na & 3) == 0) {
                const uint4* c4 = reinterpret_cast<const uint4*>(cnt);
                for (int a = 4 * part; a < na; a += 4 * CONTRACT_LANES) {
                    const uint4 v = c4[a >> 2];
                    add(a, v.x);
                    add(a + 1, v.y);
                    add(a + 2, v.z);
                    add(a + 3, v.w);
                }
            } else {
                for (int a = part; a < na; a += CONTRACT_LANES) add(a, cnt[a]);
            }
        }
        for (int off = CONTRACT_LANES / 2; off > 0; off >>= 1) {
            tot += __shfl_xor(tot, off);
#pragma unroll
            for (int q = 0; q < NM; ++q) acc[q] += __shfl_xor(acc[q], off);
        }
        if (live && part == 0) {
#pragma unroll
            for (int q = 0; q < NM; ++q) spec[(size_t)q * g.ns + cell] = cplx<T>{acc[q] * scale, T(0)};
            if (rho) rho[cell] = (T)tot * rho_scale;
        }
    }
}

// ---------------------------------------------------------------------------
// k-space solve: fft_mean_flow_field + update_gradient
// (spectral_solver.rs:133-229) and fft_mean_magnetic_field + update_gradient
// (magnetic_solver.rs:71-135), one thread per wave vector.
//
// Input spec_in: sigma-hat in `nsig` components (6 symmetric or 9), then
// P-hat (3) when `mag`.  Output layout:
//   hot  (full == 0): u(3), W01 W02 W12 (3), [E 6], [b 3], [gradB 9]
//   full (full == 1): u(3), G(9), [b 3, gradB 9]
// W and E are formed in k-space: the inverse transform is linear, so
// IFFT((G - G^T)/2) equals (IFFT(G) - IFFT(G)^T)/2 of
// src/bin/simulation/mod.rs:302-308.
// ---------------------------------------------------------------------------
struct SolveFlags {
    int nsig;
    int mag;      // P-hat present
    int full;
    int want_e, want_b, want_gb;
    int zz_trace;  // flow_point: sigma_zz from the trace (the component is not binned)
};

// Stokes solve at one wave vector (spectral_solver.rs:133-189): v = sigma k,
// v -= kn (kn . v), u = i v (1/k^2 + screening) / N.  sig(q) returns stress
// component q: the 6 of a symmetric sigma (xx xy xz yy yz zz) or all 9.
// zz_trace (6 components): sigma_zz = -(sigma_xx + sigma_yy), sig(5) not
// read.  Every stress kernel of the path is traceless (the active nn - I/3,
// the antisymmetric magnetic kernel, the rods kernel of stress.rs:144-159),
// so the tiled path sums and transforms 5 of the 6 components.
template <typename T, typename Sig>
__device__ inline void flow_point(Sig sig, int nsig, const T k[3], T ki, T ik2, T screening, T norm, T ur[3],
                                  T ui[3], bool zz_trace = false) {
    // contractions fixed by the source: the solve kernel and the fused pencil
    // kernels inline this with different surroundings and must round alike
#pragma clang fp contract(on)
    const T kn[3] = {k[0] * ki, k[1] * ki, k[2] * ki};
    T sr[3], si[3];
#pragma unroll
    for (int i = 0; i < 3; ++i) {
        T ar = T(0), ai = T(0);
#pragma unroll
        for (int j = 0; j < 3; ++j) {
            const int q = nsig == 6 ? (i <= j ? 3 * i + j - (i * (i + 1)) / 2 : 3 * j + i - (j * (j + 1)) / 2)
                                    : 3 * i + j;
            cplx<T> s;
            if (zz_trace && nsig == 6 && q == 5) {
                const cplx<T> a = sig(0), b = sig(3);
                s = cplx<T>{-(a.x + b.x), -(a.y + b.y)};
            } else {
                s = sig(q);
            }
            ar = ar + s.x * k[j];
            ai = ai + s.y * k[j];
        }
        sr[i] = ar;
        si[i] = ai;
    }
    T kr = (kn[0] * sr[0] + kn[1] * sr[1]) + kn[2] * sr[2];
    T kim = (kn[0] * si[0] + kn[1] * si[1]) + kn[2] * si[2];
    const T fac = (ik2 + screening) / norm;
#pragma unroll
    for (int i = 0; i < 3; ++i) {
        T ar = (sr[i] - kn[i] * kr) * fac;
        T ai = (si[i] - kn[i] * kim) * fac;
        ur[i] = -ai;
        ui[i] = ar;
    }
}

// Magnetic solve at one wave vector (magnetic_solver.rs:88-94): b = (P -
// kn (kn . P)) / N, the k = 0 mode kept.  pq(q) returns P-hat component q.
// Shared by the solve kernel and the fused pencil kernel, which must round
// alike.
template <typename T, typename Pq>
__device__ inline void mag_point(Pq pq, const T kn[3], T norm, T br[3], T bi[3]) {
#pragma clang fp contract(on)
    T kr = T(0), kim = T(0);
#pragma unroll
    for (int q = 0; q < 3; ++q) {
        const cplx<T> v = pq(q);
        kr = kr + kn[q] * v.x;
        kim = kim + kn[q] * v.y;
    }
#pragma unroll
    for (int q = 0; q < 3; ++q) {
        const cplx<T> v = pq(q);
        br[q] = (v.x - kn[q] * kr) / norm;
        bi[q] = (v.y - kn[q] * kim) / norm;
    }
}

// grad b_ab = i k_a b_b
template <typename T>
__device__ inline cplx<T> grad_b(const T k[3], const T br[3], const T bi[3], int a, int b) {
    return cplx<T>{-(k[a] * bi[b]), k[a] * br[b]};
}

// G_ab = i k_a u_b (spectral_solver.rs:192-229)
template <typename T>
__device__ inline cplx<T> grad_u(const T k[3], const T ur[3], const T ui[3], int a, int b) {
    return cplx<T>{-(k[a] * ui[b]), k[a] * ur[b]};
}

// W = (G - G^T) / 2 as W01, W02, W12 (w = 0..2) and E = (G + G^T) / 2 as E00,
// E01, E02, E11, E12, E22 (w = 0..5) (src/bin/simulation/mod.rs:302-308).
template <typename T>
__device__ inline cplx<T> vort_part(const T k[3], const T ur[3], const T ui[3], int w) {
#pragma clang fp contract(on)
    const int a = w == 2 ? 1 : 0, b = w == 0 ? 1 : 2;
    const cplx<T> gab = grad_u(k, ur, ui, a, b), gba = grad_u(k, ur, ui, b, a);
    return cplx<T>{(gab.x - gba.x) * T(0.5), (gab.y - gba.y) * T(0.5)};
}
template <typename T>
__device__ inline cplx<T> strain_part(const T k[3], const T ur[3], const T ui[3], int w) {
#pragma clang fp contract(on)
    const int a = w < 3 ? 0 : (w < 5 ? 1 : 2);
    const int b = w < 3 ? w : (w < 5 ? w - 2 : 2);
    const cplx<T> gab = grad_u(k, ur, ui, a, b), gba = grad_u(k, ur, ui, b, a);
    return cplx<T>{(gab.x + gba.x) * T(0.5), (gab.y + gba.y) * T(0.5)};
}

template <typename T>
__global__ void __launch_bounds__(256)
    k_solve(const cplx<T>* __restrict__ in, cplx<T>* __restrict__ out, Geo<T> g,
            const T* __restrict__ kx, const T* __restrict__ ky, const T* __restrict__ kz,
            const T* __restrict__ ik2, const T* __restrict__ kinv, T screening, T norm, SolveFlags f) {
    const int idx = blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= g.ns) return;
    const int ns = g.ns;
    const int iz = idx % g.nz, iy = (idx / g.nz) % g.ny, ix = idx / (g.ny * g.nz);
    const T k[3] = {kx[ix], ky[iy], kz[iz]};
    const T ki = kinv[idx];
    const T kn[3] = {k[0] * ki, k[1] * ki, k[2] * ki};
    T ur[3], ui[3];
    flow_point<T>([&](int q) { return in[(size_t)q * ns + idx]; }, f.nsig, k, ki, ik2[idx], screening, norm, ur,
                  ui, f.zz_trace != 0);
    int o = 0;
#pragma unroll
    for (int i = 0; i < 3; ++i) out[(size_t)(o + i) * ns + idx] = cplx<T>{ur[i], ui[i]};
    o += 3;
    if (f.full) {
        for (int a = 0; a < 3; ++a)
            for (int b = 0; b < 3; ++b) out[(size_t)(o + 3 * a + b) * ns + idx] = grad_u(k, ur, ui, a, b);
        o += 9;
    } else {
#pragma unroll
        for (int w = 0; w < 3; ++w) out[(size_t)(o + w) * ns + idx] = vort_part(k, ur, ui, w);
        o += 3;
        if (f.want_e) {
#pragma unroll
            for (int w = 0; w < 6; ++w) out[(size_t)(o + w) * ns + idx] = strain_part(k, ur, ui, w);
            o += 6;
        }
    }
    if (f.mag && (f.full || f.want_b || f.want_gb)) {
        T br[3], bi[3];
        const cplx<T>* P = in + (size_t)f.nsig * ns;
        mag_point<T>([&](int q) { return P[(size_t)q * ns + idx]; }, kn, norm, br, bi);
        if (f.full || f.want_b) {
#pragma unroll
            for (int q = 0; q < 3; ++q) out[(size_t)(o + q) * ns + idx] = cplx<T>{br[q], bi[q]};
            o += 3;
        }
        if (f.full || f.want_gb) {
            for (int a = 0; a < 3; ++a)
                for (int b = 0; b < 3; ++b) out[(size_t)(o + 3 * a + b) * ns + idx] = grad_b(k, br, bi, a, b);
            o += 9;
        }
    }
}

// Real parts of the inverse-transformed fields -> the per-cell record read by
// the update (REC_* offsets), B pre-multiplied by magnetic_dipole_dipole like
// src/bin/simulation/mod.rs:337-338.
template <typename T>
__global__ void __launch_bounds__(256)
    k_pack(const cplx<T>* __restrict__ out, const T* __restrict__ rho, int ns, int rs, SolveFlags f,
           T mu_dd, T* __restrict__ rec) {
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= ns) return;
    T* r = rec + (size_t)c * rs;
    int o = 0;
    for (int q = 0; q < 6; ++q) r[q] = out[(size_t)q * ns + c].x;
    o = 6;
    if (f.want_e) {
        for (int q = 0; q < 6; ++q) r[REC_E + q] = out[(size_t)(o + q) * ns + c].x;
        o += 6;
    }
    if (rho) r[REC_RHO] = rho[c];
    if (f.want_b) {
        for (int q = 0; q < 3; ++q) r[REC_B + q] = out[(size_t)(o + q) * ns + c].x * mu_dd;
        o += 3;
    }
    if (f.want_gb) {
        for (int q = 0; q < 9; ++q) r[REC_GB + q] = out[(size_t)(o + q) * ns + c].x;
        o += 9;
    }
}

// ---------------------------------------------------------------------------
// Particle state (structure of arrays).  lo* hold the low-order parts of the
// FP32 positions (true position = x + lo, lo_decode), nullptr in FP64.
// ---------------------------------------------------------------------------
typedef int16_t lo_t;

template <typename T>
struct State {
    T *x, *y, *z, *phi, *theta;
    lo_t *lx, *ly, *lz;
    uint32_t* id;  // original index (sorted path); nullptr: slot == index
};

// The low-order part of an FP32 position hi >= 0 is stored as a 16-bit
// integer k: the part is k u(hi), u(hi) = 2^(e - 39) for hi in [2^e, 2^(e+1)),
// i.e. 2^-16 of hi's ulp (e clamped below at -10: positions near zero keep an
// absolute step of 2^-49).  A normalized part, |lo| <= ulp(hi) / 2 =
// 2^15 u(hi), is held to u(hi) / 2: 40 significant bits for the pair, the
// precision the compensated sum needs over millions of steps, in 6 bytes per
// particle instead of 12.
__device__ inline float lo_unit(float hi) {
    const uint32_t ex = (__float_as_uint(hi) >> 23) & 0xFFu;  // biased exponent of |hi|
    return __uint_as_float(((ex < 117u ? 117u : ex) - 39u) << 23);
}
__device__ inline float lo_decode(float hi, lo_t k) { return (float)k * lo_unit(hi); }
__device__ inline lo_t lo_encode(float hi, float lo) {
    // lo / u(hi), exact: the reciprocal of a power of two
    const uint32_t ex = (__float_as_uint(hi) >> 23) & 0xFFu;
    const float inv = __uint_as_float((293u - (ex < 117u ? 117u : ex)) << 23);
    const float k = __builtin_rintf(lo * inv);
    return (lo_t)(int)__builtin_fminf(__builtin_fmaxf(k, -32768.0f), 32767.0f);
}
// The same from a double part (the exact residual p - hi of an imported
// FP64 position), rounded once.
__device__ inline lo_t lo_encode(float hi, double lo) {
    const double k = __builtin_rint(lo / (double)lo_unit(hi));
    return (lo_t)(int)__builtin_fmin(__builtin_fmax(k, -32768.0), 32767.0);
}

template <typename T>
__device__ inline T add_wrap(T x, T d, T L, float* lo) {
    // FP64: Position::pbc after pos + delta (particle.rs:72-76)
    return modulo_near(x + d, L);
}

template <>
__device__ inline float add_wrap<float>(float x, float d, float L, float* lo) {
    // Compensated (Kahan / Fast2Sum) increment, then the periodic wrap with the
    // rounding error of the wrap folded into the low part.
    float y = d + *lo;
    float t = x + y;
    float bp = t - x;
    float l = (x - (t - bp)) + (y - bp);  // TwoSum: t + l == x + y exactly
    if (!(t > 0.0f && t < L)) {  // on or beyond an edge of the box (rare)
        // wrap on the value t + l, so that it lands in [0, L); t itself may equal L
        if (t > L || (t == L && l >= 0.0f)) {
            t = t - L;  // exact (Sterbenz) for t < 2L
        } else if (t < 0.0f || (t == 0.0f && l < 0.0f)) {
            float s = t + L;
            l = l + (t - (s - L));
            t = s;
        }
        if (!(t >= 0.0f && t <= L)) {  // increment of a box length or more
            t = modulo(t, L);
            l = 0.0f;
        }
        // the wrap can leave |l| above half an ulp of the new t: renormalize
        // (Fast2Sum, |t| >= |l|), as the stored part must be (lo_encode)
        const float s = t + l;
        l = l - (s - t);
        t = s;
    }
    *lo = l;
    return t;
}

// Element i of a particle array.  A 32-bit slot (the tiled path, whose host
// side keeps n * 8 bytes below 2^32) addresses through a 32-bit byte offset,
// so that loads and stores take the array's scalar base plus one vector offset
// instead of a 64-bit vector address per array.
template <typename T>
__device__ inline T& at(T* base, uint64_t i) {
    return base[i];
}
template <typename T>
__device__ inline T& at(T* base, uint32_t i) {
    return *reinterpret_cast<T*>(reinterpret_cast<char*>(base) + i * (uint32_t)sizeof(T));
}

// One particle's state as loaded from the arrays (the low-order position
// parts only on the FP32 path).
template <typename T>
struct Loaded {
    T x, y, z, phi, theta;
    float lx, ly, lz;
};

template <typename T, typename I>
__device__ inline void load_particle(const State<T>& s, I i, Loaded<T>& q) {
    q.x = at(s.x, i);
    q.y = at(s.y, i);
    q.z = at(s.z, i);
    q.phi = at(s.phi, i);
    q.theta = at(s.theta, i);
    if constexpr (sizeof(T) == 4) {
        q.lx = lo_decode(q.x, at(s.lx, i));
        q.ly = lo_decode(q.y, at(s.ly, i));
        q.lz = lo_decode(q.z, at(s.lz, i));
    } else {
        q.lx = q.ly = q.lz = 0.0f;
    }
}

// Update of one loaded particle with its cell record (global memory or
// registers): run the LangevinBuilder chain, move, wrap, store at slot i.
template <typename T, bool FULL = true, typename I>
__device__ inline void move_loaded(const State<T>& s, I i, const Loaded<T>& q, const UpdateParams<T>& P,
                                   const T* rec, const Noise<T>& rv, T& nx, T& ny, T& nz, T& ophi, T& otheta) {
    T dp[3];
    langevin<T, FULL>(P, q.phi, q.theta, rec, rv, dp, ophi, otheta);
    if constexpr (sizeof(T) == 4) {
        float lx = q.lx, ly = q.ly, lz = q.lz;
        nx = add_wrap<T>(q.x, dp[0], P.box[0], &lx);
        ny = add_wrap<T>(q.y, dp[1], P.box[1], &ly);
        nz = add_wrap<T>(q.z, dp[2], P.box[2], &lz);
        if (FULL && P.quasi2d) {
            nz = T(0);
            lz = 0.0f;
        }
        at(s.lx, i) = lo_encode(nx, lx);
        at(s.ly, i) = lo_encode(ny, ly);
        at(s.lz, i) = lo_encode(nz, lz);
    } else {
        nx = add_wrap<T>(q.x, dp[0], P.box[0], nullptr);
        ny = add_wrap<T>(q.y, dp[1], P.box[1], nullptr);
        nz = add_wrap<T>(q.z, dp[2], P.box[2], nullptr);
        if (FULL && P.quasi2d) nz = T(0);
    }
    at(s.x, i) = nx;
    at(s.y, i) = ny;
    at(s.z, i) = nz;
    at(s.phi, i) = ophi;
    at(s.theta, i) = otheta;
}

// FP32 cell record into registers with 16-byte loads: the first 8 values
// (u, W) always, the rest when the record is the full one.
__device__ inline void gather_record(const float* __restrict__ src, int rs, float r[REC_FULL]) {
    const float4* v = reinterpret_cast<const float4*>(src);
    const int nv = rs == REC_FULL ? REC_FULL / 4 : REC_SMALL / 4;
#pragma unroll
    for (int k = 0; k < REC_FULL / 4; ++k) {
        float4 t = make_float4(0.f, 0.f, 0.f, 0.f);
        if (k < nv) t = v[k];
        r[4 * k + 0] = t.x;
        r[4 * k + 1] = t.y;
        r[4 * k + 2] = t.z;
        r[4 * k + 3] = t.w;
    }
}

// Update of every particle, plain path (particles in their original order).
template <typename T, bool FULL = true>
__device__ inline void move_particle(const State<T>& s, uint64_t i, const Geo<T>& g, const UpdateParams<T>& P,
                                     const T* __restrict__ rec, int rs, const Noise<T>& rv, T& nx, T& ny, T& nz,
                                     T& ophi, T& otheta) {
    Loaded<T> q;
    load_particle(s, i, q);
    const int c = cell_of(g, q.x, q.y, q.z);
    move_loaded<T, FULL>(s, i, q, P, rec + (size_t)c * rs, rv, nx, ny, nz, ophi, otheta);
}

// Noise sources of the update: Philox drawn in the kernel; caller-supplied
// RandomVectors in five component arrays indexed by particle id; or records
// of NOISE_REC values (the first five used) at each particle's storage slot,
// which the reference-stream generator writes so that the update reads them
// in order, 32 (FP32) or 64 (FP64) bytes per particle.
enum { NZ_PHILOX = 0, NZ_EXT = 1, NZ_REC = 2 };
constexpr int NOISE_REC = 8;

template <typename T>
struct NoiseRec {
    T v[5];
};

__device__ inline NoiseRec<float> load_noise_rec(const float* __restrict__ noise, uint32_t id) {
    const float4* p = reinterpret_cast<const float4*>(noise + (size_t)id * NOISE_REC);
    const float4 a = p[0], b = p[1];
    return NoiseRec<float>{{a.x, a.y, a.z, a.w, b.x}};
}

__device__ inline NoiseRec<double> load_noise_rec(const double* __restrict__ noise, uint32_t id) {
    const double2* p = reinterpret_cast<const double2*>(noise + (size_t)id * NOISE_REC);
    const double2 a = p[0], b = p[1], c = p[2];
    return NoiseRec<double>{{a.x, a.y, b.x, b.y, c.x}};
}

template <typename T, int NZ>
__device__ inline void draw_noise(const UpdateParams<T>& P, uint64_t seed, uint64_t step, uint32_t id,
                                  uint64_t n, const T* __restrict__ noise, Noise<T>& rv) {
    if constexpr (NZ == NZ_EXT) {
        T v[5];
#pragma unroll
        for (int k = 0; k < 5; ++k) v[k] = noise[(size_t)k * n + id];
        noise_from_vector<T>(v, rv);
    } else if constexpr (NZ == NZ_REC) {
        const NoiseRec<T> r = load_noise_rec(noise, id);
        noise_from_vector<T>(r.v, rv);
    } else {
        philox_noise(seed, step, id, P.dr, rv);
    }
}

template <typename T, int NZ, bool FULL = true>
__global__ void __launch_bounds__(256)
    k_update_plain(State<T> s, uint64_t n, Geo<T> g, UpdateParams<T> P, const T* __restrict__ rec, int rs,
                   uint64_t seed, const uint64_t* __restrict__ step, uint32_t step_off,
                   const T* __restrict__ noise) {
    const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    Noise<T> rv;
    // Philox and caller-supplied noise belong to the particle (its index at
    // ss_init), the reference stream's records to the storage slot
    const uint32_t key = (NZ == NZ_REC || !s.id) ? (uint32_t)i : s.id[i];
    draw_noise<T, NZ>(P, seed, *step + step_off, key, n, noise, rv);
    T nx, ny, nz, ophi, otheta;
    move_particle<T, FULL>(s, i, g, P, rec, rs, rv, nx, ny, nz, ophi, otheta);
}

// Global diagnostics in FP64 (mean orientation of the particles, mean
// bin-centre orientation = spatial sum of the polarisation field of
// tools/DataStreamer.py:177-196, mean stress = (1/Ns) sum over cells of
// average_stress).  All are particle averages; each thread converts its
// values to 64-bit fixed point, so the sums are exact and independent of the
// particle order.
constexpr int DIAG_N = 15;

// A grid-stride loop: each thread sums its particles' fixed-point values in
// registers, the workgroup combines its four wavefronts in LDS, and one
// device atomic per workgroup and value adds the result (one atomic per
// wavefront and value serialised on the 15 addresses: 25 ms at 2^24).
template <typename T>
__global__ void __launch_bounds__(256)
    k_diag(State<T> s, uint64_t n, Geo<T> g, const double* __restrict__ table /*[12][na]*/, double scale,
           unsigned long long* __restrict__ acc) {
    __shared__ long long part[DIAG_N][4];
    long long v[DIAG_N];
#pragma unroll
    for (int q = 0; q < DIAG_N; ++q) v[q] = 0;
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x) {
        const double ph = (double)s.phi[i], th = (double)s.theta[i];
        double sp, cp, st, ct;
        ::sincos(ph, &sp, &cp);
        ::sincos(th, &st, &ct);
        v[0] += llrint(st * cp * scale);
        v[1] += llrint(st * sp * scale);
        v[2] += llrint(ct * scale);
        const int a = angle_of(g, s.phi[i], s.theta[i]);
#pragma unroll
        for (int q = 0; q < 3; ++q) v[3 + q] += llrint(table[(size_t)(9 + q) * g.na + a] * scale);
#pragma unroll
        for (int q = 0; q < 9; ++q) v[6 + q] += llrint(table[(size_t)q * g.na + a] * scale);
    }
    const int wave = threadIdx.x >> 6;
#pragma unroll
    for (int q = 0; q < DIAG_N; ++q) {
        long long x = v[q];
        for (int off = 32; off > 0; off >>= 1) x += __shfl_xor(x, off);
        if (__lane_id() == 0) part[q][wave] = x;
    }
    __syncthreads();
    if (threadIdx.x < DIAG_N) {
        const long long x = part[threadIdx.x][0] + part[threadIdx.x][1] + part[threadIdx.x][2] + part[threadIdx.x][3];
        atomicAdd(acc + threadIdx.x, (unsigned long long)x);
    }
}

// Fill and counter kernels (graph-capturable replacements of memset).
__global__ void __launch_bounds__(256) k_fill_u32(uint32_t* __restrict__ p, size_t n, uint32_t v) {
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
        p[i] = v;
}

__global__ void k_add_step(uint64_t* step, uint32_t k) {
    if (threadIdx.x == 0 && blockIdx.x == 0) *step += k;
}

// Exclusive scan of the per-cell counts, two kernels: chunk sums, then chunk
// scans offset by the sum of the preceding chunks.  The second kernel also
// clears the counts and the scatter cursors for the next step.
constexpr int SCAN_CHUNK = 1024;

__global__ void __launch_bounds__(256) k_scan_reduce(const uint32_t* __restrict__ cnt, int ns,
                                                     uint32_t* __restrict__ part) {
    __shared__ uint32_t red[256];
    const int base = blockIdx.x * SCAN_CHUNK;
    uint32_t acc = 0;
    for (int t = threadIdx.x; t < SCAN_CHUNK; t += 256)
        if (base + t < ns) acc += cnt[base + t];
    red[threadIdx.x] = acc;
    __syncthreads();
    for (int o = 128; o > 0; o >>= 1) {
        if (threadIdx.x < o) red[threadIdx.x] += red[threadIdx.x + o];
        __syncthreads();
    }
    if (threadIdx.x == 0) part[blockIdx.x] = red[0];
}

__global__ void __launch_bounds__(256)
    k_scan_apply(uint32_t* __restrict__ cnt, int ns, const uint32_t* __restrict__ part, int nparts,
                 uint32_t* __restrict__ off, uint32_t* __restrict__ cursor) {
    __shared__ uint32_t red[256];
    __shared__ uint32_t sc[256];
    // prefix of the preceding chunks
    uint32_t acc = 0;
    for (int t = threadIdx.x; t < (int)blockIdx.x; t += 256) acc += part[t];
    red[threadIdx.x] = acc;
    __syncthreads();
    for (int o = 128; o > 0; o >>= 1) {
        if (threadIdx.x < o) red[threadIdx.x] += red[threadIdx.x + o];
        __syncthreads();
    }
    const uint32_t prefix = red[0];
    // each thread owns 4 consecutive entries
    const int base = blockIdx.x * SCAN_CHUNK + threadIdx.x * 4;
    uint32_t v[4];
    uint32_t tsum = 0;
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        v[k] = (base + k < ns) ? cnt[base + k] : 0u;
        tsum += v[k];
    }
    sc[threadIdx.x] = tsum;
    __syncthreads();
    for (int o = 1; o < 256; o <<= 1) {
        uint32_t add = threadIdx.x >= (unsigned)o ? sc[threadIdx.x - o] : 0u;
        __syncthreads();
        sc[threadIdx.x] += add;
        __syncthreads();
    }
    uint32_t run = prefix + sc[threadIdx.x] - tsum;
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        if (base + k < ns) {
            off[base + k] = run;
            cnt[base + k] = 0u;
            cursor[base + k] = 0u;
        }
        run += v[k];
    }
    if (blockIdx.x == nparts - 1 && threadIdx.x == 255) off[ns] = run;
}

// ---------------------------------------------------------------------------
// Import / export between caller AoS doubles and the device state.
// ---------------------------------------------------------------------------
template <typename T>
__global__ void __launch_bounds__(256)
    k_import(const double* __restrict__ aos, const double* __restrict__ comp, uint64_t n, State<T> s,
             double bx, double by, double bz, int do_sanitize) {
    const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    double p[5];
    const double box[3] = {bx, by, bz};
    if (do_sanitize)
        sanitize(aos + 5 * i, box, p);
    else
        for (int k = 0; k < 5; ++k) p[k] = aos[5 * i + k];
    T* dst[5] = {s.x, s.y, s.z, s.phi, s.theta};
    for (int k = 0; k < 5; ++k) dst[k][i] = (T)p[k];
    if (s.lx) {
        lo_t* lo[3] = {s.lx, s.ly, s.lz};
        for (int k = 0; k < 3; ++k) {
            // hi = fl(p) and p - hi (exact) is a normalized part; a given
            // compensation word (a snapshot's, which lo_decode made) encodes
            // back to the same integer
            const float hi = (float)p[k];
            lo[k][i] = lo_encode(hi, comp ? comp[3 * i + k] : p[k] - (double)hi);
        }
    }
    if (s.id) s.id[i] = (uint32_t)i;
}

template <typename T>
__global__ void __launch_bounds__(256)
    k_export(State<T> s, uint64_t n, double* __restrict__ aos, double* __restrict__ comp, double box0,
             double box1, double box2) {
    const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const uint64_t j = s.id ? s.id[i] : i;
    const T* src[5] = {s.x, s.y, s.z, s.phi, s.theta};
    for (int k = 0; k < 5; ++k) aos[5 * j + k] = (double)src[k][i];
    if (s.lx) {
        const lo_t* lo[3] = {s.lx, s.ly, s.lz};
        if (comp) {
            for (int k = 0; k < 3; ++k) comp[3 * j + k] = (double)lo_decode((float)src[k][i], lo[k][i]);
        } else {
            const double L[3] = {box0, box1, box2};
            for (int k = 0; k < 3; ++k) {
                double v = (double)src[k][i] + (double)lo_decode((float)src[k][i], lo[k][i]);
                if (v < 0.0) v += L[k];
                if (v >= L[k]) v -= L[k];
                aos[5 * j + k] = v;
            }
        }
    } else if (comp) {
        for (int k = 0; k < 3; ++k) comp[3 * j + k] = 0.0;
    }
}

// stride: distance between a particle's components (n: component arrays,
// 1: records); rec: values per record (1 for component arrays); inv (may be
// null): slot of each particle id.
template <typename T>
__global__ void __launch_bounds__(256) k_export_noise(const T* __restrict__ noise, uint64_t n, uint64_t stride,
                                                      int rec, const uint32_t* __restrict__ inv,
                                                      double* __restrict__ aos) {
    const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const uint64_t slot = inv ? inv[i] : i;
    for (int k = 0; k < 5; ++k) aos[5 * i + k] = (double)noise[(size_t)k * stride + slot * (uint64_t)rec];
}

__global__ void __launch_bounds__(256) k_inverse_perm(const uint32_t* __restrict__ id, uint64_t n,
                                                      uint32_t* __restrict__ inv) {
    const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) inv[id[i]] = (uint32_t)i;
}

template <typename T>
__global__ void __launch_bounds__(256) k_import_noise(const double* __restrict__ aos, uint64_t n,
                                                      T* __restrict__ noise) {
    const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    for (int k = 0; k < 5; ++k) noise[(size_t)k * n + i] = (T)aos[5 * i + k];
}

template <typename T>
__global__ void __launch_bounds__(256) k_real_parts(const cplx<T>* __restrict__ c, size_t cnt,
                                                    double* __restrict__ out) {
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < cnt) out[i] = (double)c[i].x;
}

// DensityGradient (distribution/density_gradient.rs:73-110), not part of the
// step: the density of each cell as ndarray sums a contiguous row (eight
// partial sums, numeric_util::unrolled_sum) times dtheta dphi.
__global__ void __launch_bounds__(256) k_density_rows(const double* __restrict__ dist, uint64_t ns, int na,
                                                      double dthph, double* __restrict__ rho) {
    const uint64_t c = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= ns) return;
    const double* x = dist + c * (uint64_t)na;
    double p[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    int len = na, o = 0;
    while (len >= 8) {
        for (int k = 0; k < 8; ++k) p[k] = p[k] + x[o + k];
        o += 8;
        len -= 8;
    }
    double sum = 0.0;
    sum = sum + (p[0] + p[4]);
    sum = sum + (p[1] + p[5]);
    sum = sum + (p[2] + p[6]);
    sum = sum + (p[3] + p[7]);
    for (int i = 0; i < len; ++i) sum = sum + x[o + i];
    rho[c] = sum * dthph;
}

template <typename T>
__global__ void __launch_bounds__(256) k_density_spec(const double* __restrict__ rho, uint64_t ns, T dthph,
                                                      cplx<T>* __restrict__ spec) {
    const uint64_t c = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (c < ns) spec[c] = cplx<T>{(T)rho[c], T(0) * dthph};
}

// g_a = (k_a d) (i / N) in complex arithmetic with k = (k_a, 0), as the
// reference evaluates it (no fused multiply-adds).
template <typename T>
__global__ void __launch_bounds__(256) k_density_grad(const cplx<T>* __restrict__ d, Geo<T> g,
                                                      const T* __restrict__ kx, const T* __restrict__ ky,
                                                      const T* __restrict__ kz, T nre, T nim,
                                                      cplx<T>* __restrict__ out) {
#pragma clang fp contract(off)
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= g.ns) return;
    const int iz = c % g.nz, iy = (c / g.nz) % g.ny, ix = c / (g.nz * g.ny);
    const T k[3] = {kx[ix], ky[iy], kz[iz]};
    const T dre = d[c].x, dim = d[c].y;
    for (int a = 0; a < 3; ++a) {
        const T kre = k[a], kim = T(0);
        const T e1re = kre * dre - kim * dim, e1im = kre * dim + kim * dre;
        out[(size_t)a * g.ns + c] = cplx<T>{e1re * nre - e1im * nim, e1re * nim + e1im * nre};
    }
}

// interpolate_vector_field (mesh/interpolate.rs:11-123) of a field
// [3][nx][ny][nz] at n positions: the 2x2x2 cells around floor(q / g - 1/2)
// (periodic), trilinear with lerp(a, b, t) = a (1 - t) + b t along x, y, z.
// The reference takes z's fractional offset from position.y; here from z.
__global__ void __launch_bounds__(256) k_interpolate(const double* __restrict__ f, int nx, int ny, int nz,
                                                     double gx, double gy, double gz,
                                                     const double* __restrict__ pos, uint64_t n,
                                                     double* __restrict__ out) {
#pragma clang fp contract(off)
    const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const double q[3] = {pos[3 * i], pos[3 * i + 1], pos[3 * i + 2]};
    const double gw[3] = {gx, gy, gz};
    const int dims[3] = {nx, ny, nz};
    double rel[3];
    int w0[3], w1[3];
    for (int a = 0; a < 3; ++a) {
        rel[a] = fmod(q[a] / gw[a] + 0.5, 1.0);
        const int idx = (int)floor(q[a] / gw[a] - 0.5);
        w0[a] = (idx + dims[a]) % dims[a];
        w1[a] = (idx + 1 + dims[a]) % dims[a];
    }
    const size_t ns = (size_t)nx * ny * nz;
    auto lerp = [](double a, double b, double t) { return a * (1.0 - t) + b * t; };
    for (int comp = 0; comp < 3; ++comp) {
        const double* fc = f + (size_t)comp * ns;
        double v[8];
        for (int a = 0; a < 2; ++a)
            for (int b = 0; b < 2; ++b)
                for (int c = 0; c < 2; ++c)
                    v[4 * a + 2 * b + c] =
                        fc[((size_t)(a ? w1[0] : w0[0]) * ny + (b ? w1[1] : w0[1])) * nz + (c ? w1[2] : w0[2])];
        const double c00 = lerp(v[0], v[4], rel[0]), c01 = lerp(v[1], v[5], rel[0]);
        const double c10 = lerp(v[2], v[6], rel[0]), c11 = lerp(v[3], v[7], rel[0]);
        const double c0 = lerp(c00, c10, rel[1]), c1 = lerp(c01, c11, rel[1]);
        out[3 * i + comp] = lerp(c0, c1, rel[2]);
    }
}

// Contraction of a caller-given distribution (values, not counts) for the
// solver entry points ss_flow_solve / ss_magnetic_solve.
template <typename T, int NM>
__global__ void __launch_bounds__(256)
    k_contract_dense(const double* __restrict__ dist, Geo<T> g, const T* __restrict__ kern_g, T measure,
                     cplx<T>* __restrict__ spec) {
    const int wave = threadIdx.x >> 6;
    const int cell = blockIdx.x * 4 + wave;
    if (cell >= g.ns) return;
    const int lane = __lane_id();
    T acc[NM];
    for (int q = 0; q < NM; ++q) acc[q] = T(0);
    for (int a = lane; a < g.na; a += 64) {
        const T f = (T)dist[(size_t)cell * g.na + a];
        for (int q = 0; q < NM; ++q) acc[q] += kern_g[q * g.na + a] * f;
    }
    for (int off = 32; off > 0; off >>= 1)
        for (int q = 0; q < NM; ++q) acc[q] += __shfl_xor(acc[q], off);
    if (lane == 0)
        for (int q = 0; q < NM; ++q) spec[(size_t)q * g.ns + cell] = cplx<T>{acc[q] * measure, T(0)};
}

template <typename T>
__global__ void k_coord_to_grid(const double* __restrict__ aos, uint64_t n, Geo<T> g,
                                int32_t* __restrict__ out) {
    const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const T* dummy = nullptr;
    (void)dummy;
    out[5 * i + 0] = bin_x(g, (T)aos[5 * i + 0]);
    out[5 * i + 1] = bin_y(g, (T)aos[5 * i + 1]);
    out[5 * i + 2] = bin_z(g, (T)aos[5 * i + 2]);
    out[5 * i + 3] = bin_phi(g, (T)aos[5 * i + 3]);
    out[5 * i + 4] = bin_theta(g, (T)aos[5 * i + 4]);
}

// Debug update of n independent particles with explicit per-particle fields
// (u[3] W[9] E[9] rho B[3] gradB[9]) and noise.
template <typename T>
__global__ void k_debug_update(const double* __restrict__ parts, const double* __restrict__ fields,
                               const double* __restrict__ rvs, uint64_t n, UpdateParams<T> P, T mu_dd,
                               double* __restrict__ out) {
    const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const double* f = fields + 34 * i;
    T rec[REC_FULL];
    for (int q = 0; q < REC_FULL; ++q) rec[q] = T(0);
    for (int q = 0; q < 3; ++q) rec[REC_U + q] = (T)f[q];
    rec[REC_W + 0] = (T)f[3 + 1];
    rec[REC_W + 1] = (T)f[3 + 2];
    rec[REC_W + 2] = (T)f[3 + 5];
    rec[REC_E + 0] = (T)f[12 + 0];
    rec[REC_E + 1] = (T)f[12 + 1];
    rec[REC_E + 2] = (T)f[12 + 2];
    rec[REC_E + 3] = (T)f[12 + 4];
    rec[REC_E + 4] = (T)f[12 + 5];
    rec[REC_E + 5] = (T)f[12 + 8];
    rec[REC_RHO] = (T)f[21];
    for (int q = 0; q < 3; ++q) rec[REC_B + q] = (T)f[22 + q] * mu_dd;
    for (int q = 0; q < 9; ++q) rec[REC_GB + q] = (T)f[25 + q];
    T v[5];
    for (int k = 0; k < 5; ++k) v[k] = (T)rvs[5 * i + k];
    Noise<T> rv;
    noise_from_vector<T>(v, rv);
    const double* p = parts + 5 * i;
    T dp[3], ophi, otheta;
    langevin<T>(P, (T)p[3], (T)p[4], rec, rv, dp, ophi, otheta);
    out[5 * i + 0] = (double)modulo((T)p[0] + dp[0], P.box[0]);
    out[5 * i + 1] = (double)modulo((T)p[1] + dp[1], P.box[1]);
    out[5 * i + 2] = P.quasi2d ? 0.0 : (double)modulo((T)p[2] + dp[2], P.box[2]);
    out[5 * i + 3] = (double)ophi;
    out[5 * i + 4] = (double)otheta;
}

}  // namespace ss
