// ss_spectral.h -- the spectral part of one step on the tiled path in three
// kernels instead of convert + 3 forward axis passes + solve + 3 inverse
// passes + pack:
//
//   k_spec_planes_fwd  per (x plane, moment component): fixed-point moments ->
//                      working precision, z and y transforms in LDS
//   k_spec_pencils     per group of (y, z) pencils: x transform of every
//                      stress component, the Stokes solve, inverse x
//                      transform of u and W (and E)
//   k_spec_planes_inv  per (x plane, field component): inverse y and z
//                      transforms in LDS, real part into the cell records
//
// The transforms are the Stockham stages of ss_fft_device.h in the same axis
// order as the axis-pass path (forward z, y, x; backward x, y, z) and the
// solve is flow_point, so both paths give identical results.  Where all three
// axes have one length P of 16, 32, 48 or 64 points the stages are
// instantiated for it (sizes fixed at compile time; planes transformed in
// place).  The *_batched kernels run the same bodies for several systems in
// one launch (ensembles); the *2 kernels are the paired variant (two real
// fields per complex transform).
#pragma once
#include "ss_fft_device.h"
#include "ss_kernels.h"

namespace ss {

// Workgroup sizes: the plane kernels keep one plane per workgroup, the
// pencil kernel a group of pencils; more threads per workgroup give each
// workgroup's LDS-bound stages more waves to overlap.
constexpr int SPEC_PLANE_THREADS = 512;
constexpr int SPEC_PENCIL_THREADS = 512;
constexpr int SPEC_MAX_PENCILS = 16;  // pencils (or pencil pairs) per workgroup

// Fields of the pencil output and of the inverse planes after u (3) and W
// (3), present as flagged, in this order (the order of the solve kernel's
// hot outputs): E (6), b (3), grad b (9).
constexpr int SPEC_OUT_E = 1, SPEC_OUT_B = 2, SPEC_OUT_GB = 4;

// A strided loop over t < n whose global loads are issued U iterations at a
// time, the stores after them: a loop that consumes each load before the next
// one waits for the memory latency once per iteration.
template <int NT, int U = 8, typename Load, typename Store>
__device__ inline void batched_loop(int n, Load load, Store store) {
    constexpr int st = NT;
    for (int b = threadIdx.x; b < n; b += U * st) {
        decltype(load(0)) v[U];
#pragma unroll
        for (int u = 0; u < U; ++u)
            if (b + u * st < n) v[u] = load(b + u * st);
#pragma unroll
        for (int u = 0; u < U; ++u)
            if (b + u * st < n) store(b + u * st, v[u]);
    }
}

// The transform of nlines lines of one axis: with P > 0 every axis has the
// length P and the stages run with their sizes fixed at compile time (NL > 0:
// nlines = NL as well); the arithmetic is the same.
template <typename T, bool LINES_FAST, int NT, int P, int NL>
__device__ inline cplx<T>* spec_lines(cplx<T>* a, cplx<T>* b, const cplx<T>* tw, const SpecAxis& ax, int nlines,
                                      int ls, int es, int sign) {
    if constexpr (P > 0)
        return fft_lines_ct<T, LINES_FAST, NT, P, NL>(a, b, tw, nlines, ls, es, sign);
    else
        return fft_lines<T, LINES_FAST, NT>(a, b, tw, ax, nlines, ls, es, sign);
}

// The count component of one plane: main + inflow, inflow cleared, the
// density into rho and the cell records when wanted.
template <typename T, int NT>
__device__ inline void spec_count_plane(const long long* __restrict__ mc, long long* __restrict__ ic, int np,
                                        double rho_scale, T* __restrict__ rho, T* __restrict__ rec, size_t base,
                                        int rs) {
    if (!rho) {  // the counts are not needed: only clear the inflow
        for (int c = threadIdx.x; c < np; c += NT) ic[c] = 0;
        return;
    }
    batched_loop<NT>(
        np, [&](int c) { return mc[c] + ic[c]; },
        [&](int c, long long v) {
            ic[c] = 0;
            const T r = (T)((double)v * rho_scale);
            rho[base + c] = r;
            rec[(base + c) * rs + REC_RHO] = r;
        });
}

// LDS of one plane kernel: twiddles of y and z, two plane buffers.
// With sizes fixed at compile time (P > 0) the transforms run in place in one
// plane buffer.
inline size_t spec_plane_lds(int ny, int nz, size_t esz, bool inplace) {
    return esz * 2 * ((size_t)ny + nz + (inplace ? 1 : 2) * (size_t)ny * nz);
}

// The two transforms of a plane held at b0 (b1: second buffer, run-time
// sizes only): forward z then y, or inverse y then z.  Returns the buffer
// with the result.
template <typename T, int P>
__device__ inline cplx<T>* spec_plane_fwd(cplx<T>* b0, cplx<T>* b1, const cplx<T>* ty, const cplx<T>* tz,
                                          const SpecAxis& ay, const SpecAxis& az, int ny, int nz) {
    // z lines: ny contiguous lines; y lines: nz lines interleaved at stride nz
    if constexpr (P > 0) {
        fft_lines_ct_inplace<T, false, SPEC_PLANE_THREADS, P, P>(b0, tz, nz, 1, -1);
        fft_lines_ct_inplace<T, true, SPEC_PLANE_THREADS, P, P>(b0, ty, 1, nz, -1);
        return b0;
    } else {
        cplx<T>* r1 = spec_lines<T, false, SPEC_PLANE_THREADS, 0, 0>(b0, b1, tz, az, ny, nz, 1, -1);
        return spec_lines<T, true, SPEC_PLANE_THREADS, 0, 0>(r1, r1 == b0 ? b1 : b0, ty, ay, nz, 1, nz, -1);
    }
}

template <typename T, int P>
__device__ inline cplx<T>* spec_plane_inv(cplx<T>* b0, cplx<T>* b1, const cplx<T>* ty, const cplx<T>* tz,
                                          const SpecAxis& ay, const SpecAxis& az, int ny, int nz) {
    if constexpr (P > 0) {
        fft_lines_ct_inplace<T, true, SPEC_PLANE_THREADS, P, P>(b0, ty, 1, nz, +1);
        fft_lines_ct_inplace<T, false, SPEC_PLANE_THREADS, P, P>(b0, tz, nz, 1, +1);
        return b0;
    } else {
        cplx<T>* r1 = spec_lines<T, true, SPEC_PLANE_THREADS, 0, 0>(b0, b1, ty, ay, nz, 1, nz, +1);
        return spec_lines<T, false, SPEC_PLANE_THREADS, 0, 0>(r1, r1 == b0 ? b1 : b0, tz, az, ny, nz, 1, +1);
    }
}

// Forward: component q < NM of plane x; the blocks q == NM convert the count
// component (density, when wanted) and transform nothing, so that no block
// carries both.  Clears the inflow entries it reads.
// bid: the block's index in the launch of one system.
template <typename T, int NM, int P>
__device__ __forceinline__ void spec_planes_fwd_body(const long long* __restrict__ main, long long* __restrict__ inflow, int nx,
                                            double scale, double rho_scale, const SpecAxis& ay, const SpecAxis& az,
                                            const cplx<T>* __restrict__ twy, const cplx<T>* __restrict__ twz,
                                            cplx<T>* __restrict__ spec, T* __restrict__ rho, T* __restrict__ rec,
                                            int rs, int bid, int zz_trace) {
    extern __shared__ __attribute__((aligned(16))) unsigned char smem_raw[];
    const int ny = P ? P : ay.n, nz = P ? P : az.n, np = ny * nz;
    cplx<T>* ty = reinterpret_cast<cplx<T>*>(smem_raw);
    cplx<T>* tz = ty + ny;
    cplx<T>* b0 = tz + nz;
    cplx<T>* b1 = b0 + np;
    const int x = bid % nx, q = bid / nx;
    const size_t ns = (size_t)nx * np, base = (size_t)x * np;
    if (NM == 6 && zz_trace && q == 5) return;  // sigma_zz comes from the trace (flow_point)
    if (q == NM) {
        spec_count_plane<T, SPEC_PLANE_THREADS>(main + (size_t)NM * ns + base, inflow + (size_t)NM * ns + base, np,
                                                rho_scale, rho, rec, base, rs);
        return;
    }
    for (int t = threadIdx.x; t < ny; t += SPEC_PLANE_THREADS) ty[t] = twy[t];
    for (int t = threadIdx.x; t < nz; t += SPEC_PLANE_THREADS) tz[t] = twz[t];
    const long long* mq = main + (size_t)q * ns + base;
    long long* iq = inflow + (size_t)q * ns + base;
    batched_loop<SPEC_PLANE_THREADS>(
        np, [&](int c) { return mq[c] + iq[c]; },
        [&](int c, long long v) {
            iq[c] = 0;
            b0[c] = cplx<T>{(T)((double)v * scale), T(0)};
        });
    __syncthreads();
    cplx<T>* r2 = spec_plane_fwd<T, P>(b0, b1, ty, tz, ay, az, ny, nz);
    cplx<T>* out = spec + (size_t)q * ns + base;
    for (int c = threadIdx.x; c < np; c += SPEC_PLANE_THREADS) out[c] = r2[c];
}

template <typename T, int NM, int P>
__global__ void __launch_bounds__(SPEC_PLANE_THREADS)
    k_spec_planes_fwd(const long long* __restrict__ main, long long* __restrict__ inflow, int nx, double scale,
                      double rho_scale, SpecAxis ay, SpecAxis az, const cplx<T>* __restrict__ twy,
                      const cplx<T>* __restrict__ twz, cplx<T>* __restrict__ spec, T* __restrict__ rho,
                      T* __restrict__ rec, int rs, int zz_trace) {
    spec_planes_fwd_body<T, NM, P>(main, inflow, nx, scale, rho_scale, ay, az, twy, twz, spec, rho, rec, rs,
                                   (int)blockIdx.x, zz_trace);
}

// Wave vector, 1/|k| and 1/k^2 of bin (x, pencil p = y nz + z) from the
// spectral solver's tables.
template <typename T>
struct KPoint {
    T k[3] = {T(0), T(0), T(0)};
    T ki = T(0), ik2 = T(0);
    __device__ static KPoint load(const T* kx, const T* ky, const T* kz, const T* kinv, const T* ik2t, int x, int p,
                                  int np, int nz) {
        KPoint r;
        const size_t idx = (size_t)x * np + p;
        r.k[0] = kx[x];
        r.k[1] = ky[p / nz];
        r.k[2] = kz[p % nz];
        r.ki = kinv[idx];
        r.ik2 = ik2t[idx];
        return r;
    }
};

// LDS of the pencil kernel: x twiddles and two buffers of nl lines of
// nx + 1 points (padded against bank conflicts).
inline size_t spec_pencil_lds(int nx, int nl, size_t esz) {
    return esz * 2 * ((size_t)nx + 2 * (size_t)nl * (nx + 1));
}

// Pencils p0 .. p0+pb-1 (p = y nz + z) of the nsig stress components and,
// with nmag = 3, the polarisation P-hat after them: forward x transform, flow
// solve (and magnetic solve), inverse x transform of the nout outputs: u, W
// and the fields `outs` flags (SPEC_OUT_*).
// bid, groups: the block's index and the number of blocks in the launch of
// one system.  zz_trace and nmag are not used together.
template <typename T, int P>
__device__ __forceinline__ void spec_pencils_body(const cplx<T>* __restrict__ in, cplx<T>* __restrict__ out, int np, int nz,
                                         int pb_max, int nsig, int nout, const SpecAxis& ax,
                                         const cplx<T>* __restrict__ twx, const T* __restrict__ kx,
                                         const T* __restrict__ ky, const T* __restrict__ kz,
                                         const T* __restrict__ ik2, const T* __restrict__ kinv, T screening, T norm,
                                         int bid, int groups, int zz_trace, int nmag, int outs) {
    extern __shared__ __attribute__((aligned(16))) unsigned char smem_raw[];
    const int nx = P ? P : ax.n, ls = nx + 1;
    const int nin = nsig + nmag;
    const int nl = (nin > nout ? nin : nout) * pb_max;
    cplx<T>* tw = reinterpret_cast<cplx<T>*>(smem_raw);
    cplx<T>* A = tw + nx;
    cplx<T>* B = A + (size_t)nl * ls;
    const size_t ns = (size_t)nx * np;
    // workgroups are dispatched round-robin over the XCDs: give each XCD a
    // contiguous run of pencil groups, so that the two groups sharing a
    // cache line of every x row are read through the same L2
    const int g = groups % 8 == 0 ? (bid % 8) * (groups / 8) + bid / 8 : bid;
    const int p0 = g * pb_max;
    const int pb = min(pb_max, np - p0);
    for (int t = threadIdx.x; t < nx; t += SPEC_PENCIL_THREADS) tw[t] = twx[t];
    // line (c, p) at A[(c pb_max + p) ls]; lanes over p for coalesced loads
    // four loads at a time: eight take the kernel past 128 VGPRs (one
    // workgroup per CU) where all axes have compile-time lengths; with
    // zz_trace the last (zz) component is neither loaded nor transformed
    const int nload = zz_trace && nsig == 6 ? 5 : nin;
    batched_loop<SPEC_PENCIL_THREADS, 4>(
        nload * nx * pb_max,
        [&](int f) {
            const int p = f % pb_max, xx = (f / pb_max) % nx, c = f / (pb_max * nx);
            return p < pb ? in[(size_t)c * ns + (size_t)xx * np + p0 + p] : cplx<T>{T(0), T(0)};
        },
        [&](int f, cplx<T> v) {
            const int p = f % pb_max, xx = (f / pb_max) % nx, c = f / (pb_max * nx);
            A[(size_t)(c * pb_max + p) * ls + xx] = v;
        });
    __syncthreads();
    cplx<T>* S = spec_lines<T, false, SPEC_PENCIL_THREADS, P, 0>(A, B, tw, ax, nload * pb_max, ls, 1, -1);
    cplx<T>* O = S == A ? B : A;
    for (int f = threadIdx.x; f < nx * pb_max; f += SPEC_PENCIL_THREADS) {
        const int p = f % pb_max, xx = f / pb_max;
        if (p >= pb) {
            for (int o = 0; o < nout; ++o) O[(size_t)(o * pb_max + p) * ls + xx] = cplx<T>{T(0), T(0)};
            continue;
        }
        const KPoint<T> kp = KPoint<T>::load(kx, ky, kz, kinv, ik2, xx, p0 + p, np, nz);
        const T* k = kp.k;
        T ur[3], ui[3];
        flow_point<T>([&](int q) { return S[(size_t)(q * pb_max + p) * ls + xx]; }, nsig, k, kp.ki, kp.ik2, screening,
                      norm, ur, ui, zz_trace != 0);
#pragma unroll
        for (int i = 0; i < 3; ++i) O[(size_t)(i * pb_max + p) * ls + xx] = cplx<T>{ur[i], ui[i]};
#pragma unroll
        for (int w = 0; w < 3; ++w) O[(size_t)((3 + w) * pb_max + p) * ls + xx] = vort_part(k, ur, ui, w);
        int o = 6;
        if (outs & SPEC_OUT_E) {
#pragma unroll
            for (int w = 0; w < 6; ++w) O[(size_t)((o + w) * pb_max + p) * ls + xx] = strain_part(k, ur, ui, w);
            o += 6;
        }
        if (nmag) {
            const T kn[3] = {k[0] * kp.ki, k[1] * kp.ki, k[2] * kp.ki};
            T br[3], bi[3];
            mag_point<T>([&](int q) { return S[(size_t)((nsig + q) * pb_max + p) * ls + xx]; }, kn, norm, br, bi);
            if (outs & SPEC_OUT_B) {
#pragma unroll
                for (int q = 0; q < 3; ++q) O[(size_t)((o + q) * pb_max + p) * ls + xx] = cplx<T>{br[q], bi[q]};
                o += 3;
            }
            if (outs & SPEC_OUT_GB) {
#pragma unroll
                for (int a = 0; a < 3; ++a)
#pragma unroll
                    for (int b = 0; b < 3; ++b)
                        O[(size_t)((o + 3 * a + b) * pb_max + p) * ls + xx] = grad_b(k, br, bi, a, b);
            }
        }
    }
    __syncthreads();
    cplx<T>* R = spec_lines<T, false, SPEC_PENCIL_THREADS, P, 0>(O, S, tw, ax, nout * pb_max, ls, 1, +1);
    for (int f = threadIdx.x; f < nout * nx * pb_max; f += SPEC_PENCIL_THREADS) {
        const int p = f % pb_max, xx = (f / pb_max) % nx, o = f / (pb_max * nx);
        if (p < pb) out[(size_t)o * ns + (size_t)xx * np + p0 + p] = R[(size_t)(o * pb_max + p) * ls + xx];
    }
}

template <typename T, int P>
__global__ void __launch_bounds__(SPEC_PENCIL_THREADS)
    k_spec_pencils(const cplx<T>* __restrict__ in, cplx<T>* __restrict__ out, int np, int nz, int pb_max,
                   int nsig, int nout, SpecAxis ax, const cplx<T>* __restrict__ twx, const T* __restrict__ kx,
                   const T* __restrict__ ky, const T* __restrict__ kz, const T* __restrict__ ik2,
                   const T* __restrict__ kinv, T screening, T norm, int zz_trace, int nmag, int outs) {
    spec_pencils_body<T, P>(in, out, np, nz, pb_max, nsig, nout, ax, twx, kx, ky, kz, ik2, kinv, screening, norm,
                            (int)blockIdx.x, (int)gridDim.x, zz_trace, nmag, outs);
}

// Inverse: field component o of plane x; the real part goes to record slot
// o (u, W), then to the slots of E, b (times mu_dd) and grad b, those `outs`
// flags, in that order (the solve kernel and k_pack's layout).
template <typename T, int P>
__device__ __forceinline__ void spec_planes_inv_body(const cplx<T>* __restrict__ spec, int nx, const SpecAxis& ay,
                                            const SpecAxis& az, const cplx<T>* __restrict__ twy,
                                            const cplx<T>* __restrict__ twz, T* __restrict__ rec, int rs, int bid,
                                            int outs, T mu_dd) {
    extern __shared__ __attribute__((aligned(16))) unsigned char smem_raw[];
    const int ny = P ? P : ay.n, nz = P ? P : az.n, np = ny * nz;
    cplx<T>* ty = reinterpret_cast<cplx<T>*>(smem_raw);
    cplx<T>* tz = ty + ny;
    cplx<T>* b0 = tz + nz;
    cplx<T>* b1 = b0 + np;
    const int x = bid % nx, o = bid / nx;
    const size_t ns = (size_t)nx * np, base = (size_t)x * np;
    for (int t = threadIdx.x; t < ny; t += SPEC_PLANE_THREADS) ty[t] = twy[t];
    for (int t = threadIdx.x; t < nz; t += SPEC_PLANE_THREADS) tz[t] = twz[t];
    const cplx<T>* src = spec + (size_t)o * ns + base;
    batched_loop<SPEC_PLANE_THREADS>(np, [&](int c) { return src[c]; }, [&](int c, cplx<T> v) { b0[c] = v; });
    __syncthreads();
    cplx<T>* r2 = spec_plane_inv<T, P>(b0, b1, ty, tz, ay, az, ny, nz);
    int slot = o, r = o - 6;
    bool times_mu = false;
    if (o >= 6) {
        if ((outs & SPEC_OUT_E) && r < 6) {
            slot = REC_E + r;
        } else {
            if (outs & SPEC_OUT_E) r -= 6;
            if ((outs & SPEC_OUT_B) && r < 3) {
                slot = REC_B + r;
                times_mu = true;
            } else {
                if (outs & SPEC_OUT_B) r -= 3;
                slot = REC_GB + r;
            }
        }
    }
    if (times_mu)
        for (int c = threadIdx.x; c < np; c += SPEC_PLANE_THREADS) rec[(base + c) * rs + slot] = r2[c].x * mu_dd;
    else
        for (int c = threadIdx.x; c < np; c += SPEC_PLANE_THREADS) rec[(base + c) * rs + slot] = r2[c].x;
}

template <typename T, int P>
__global__ void __launch_bounds__(SPEC_PLANE_THREADS)
    k_spec_planes_inv(const cplx<T>* __restrict__ spec, int nx, SpecAxis ay, SpecAxis az,
                      const cplx<T>* __restrict__ twy, const cplx<T>* __restrict__ twz, T* __restrict__ rec,
                      int rs, int outs, T mu_dd) {
    spec_planes_inv_body<T, P>(spec, nx, ay, az, twy, twz, rec, rs, (int)blockIdx.x, outs, mu_dd);
}

// ---------------------------------------------------------------------------
// Batched launches: the spectral stages of several systems (an ensemble's
// members of one axis-length class, without the magnetic solve) in one
// launch of each kernel.  first[m]
// is the first block of system m (first[n] the total), so every block runs
// the body of its own system with its index within that system.
// ---------------------------------------------------------------------------
template <typename T>
struct SpecJob {
    const long long* main;
    long long* inflow;
    cplx<T>* spec_in;
    cplx<T>* spec_out;
    T* rho;  // nullptr: no density
    T* rec;
    const cplx<T> *twx, *twy, *twz;
    const T *kx, *ky, *kz, *ik2, *kinv;
    double scale, rho_scale;
    T screening, norm;
    SpecAxis ax, ay, az;
    int nx, np, nz, rs, pb, nsig, nout;
    int zz_trace;  // sigma_zz from the trace (flow_point)
};

// The system a block belongs to: the last m with first[m] <= b.
__device__ inline int spec_job_of(const int* __restrict__ first, int n, int b) {
    int lo = 0, hi = n - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (first[mid] <= b) lo = mid; else hi = mid - 1;
    }
    return __builtin_amdgcn_readfirstlane(lo);  // uniform: the job stays in scalar registers
}

template <typename T, int NM, int P>
__global__ void __launch_bounds__(SPEC_PLANE_THREADS)
    k_spec_planes_fwd_batched(const SpecJob<T>* __restrict__ jobs, const int* __restrict__ first, int n) {
    const int m = spec_job_of(first, n, (int)blockIdx.x);
    const SpecJob<T>& a = jobs[m];
    spec_planes_fwd_body<T, NM, P>(a.main, a.inflow, a.nx, a.scale, a.rho_scale, a.ay, a.az, a.twy, a.twz,
                                   a.spec_in, a.rho, a.rec, a.rs, (int)blockIdx.x - first[m], a.zz_trace);
}

template <typename T, int P>
__global__ void __launch_bounds__(SPEC_PENCIL_THREADS)
    k_spec_pencils_batched(const SpecJob<T>* __restrict__ jobs, const int* __restrict__ first, int n) {
    const int m = spec_job_of(first, n, (int)blockIdx.x);
    const SpecJob<T>& a = jobs[m];
    spec_pencils_body<T, P>(a.spec_in, a.spec_out, a.np, a.nz, a.pb, a.nsig, a.nout, a.ax, a.twx, a.kx, a.ky, a.kz,
                            a.ik2, a.kinv, a.screening, a.norm, (int)blockIdx.x - first[m], first[m + 1] - first[m],
                            a.zz_trace, 0, a.nout == 12 ? SPEC_OUT_E : 0);
}

template <typename T, int P>
__global__ void __launch_bounds__(SPEC_PLANE_THREADS)
    k_spec_planes_inv_batched(const SpecJob<T>* __restrict__ jobs, const int* __restrict__ first, int n) {
    const int m = spec_job_of(first, n, (int)blockIdx.x);
    const SpecJob<T>& a = jobs[m];
    spec_planes_inv_body<T, P>(a.spec_out, a.nx, a.ay, a.az, a.twy, a.twz, a.rec, a.rs, (int)blockIdx.x - first[m],
                               a.nout == 12 ? SPEC_OUT_E : 0, T(1));
}

// ---------------------------------------------------------------------------
// Paired variant: two real fields per complex transform.
//
// The moments and the fields are real, so two of them share one complex
// transform (a + i b).  The pencil kernel takes a (y, z) pencil together with
// its mirror (-y, -z), where the transforms of a and b separate:
// A(k) = (Z(k) + conj Z(-k)) / 2, B(k) = (Z(k) - conj Z(-k)) / 2i.  The solve
// runs at k and at -k (with the reference's k table, whose Nyquist entries
// are their own mirror), and the Hermitian part H(k) = (X(k) + conj X(-k)) / 2
// of each output goes into the inverse transform: its inverse is real and
// equals the real part of the inverse of X, which is what the reference
// keeps.  Off the Nyquist planes X(-k) = conj X(k) exactly, so H = X there.
// Half the transforms of the per-component path, the same results up to
// rounding.
// ---------------------------------------------------------------------------

// Forward: packed component c = moments 2c + i 2c+1 of plane x (2c+1 may be
// absent); the blocks c == (NM + 1) / 2 convert the count component only.
template <typename T, int NM, int P>
__global__ void __launch_bounds__(SPEC_PLANE_THREADS)
    k_spec_planes_fwd2(const long long* __restrict__ main, long long* __restrict__ inflow, int nx, double scale,
                       double rho_scale, SpecAxis ay, SpecAxis az, const cplx<T>* __restrict__ twy,
                       const cplx<T>* __restrict__ twz, cplx<T>* __restrict__ spec, T* __restrict__ rho,
                       T* __restrict__ rec, int rs) {
    extern __shared__ __attribute__((aligned(16))) unsigned char smem_raw[];
    const int ny = P ? P : ay.n, nz = P ? P : az.n, np = ny * nz;
    cplx<T>* ty = reinterpret_cast<cplx<T>*>(smem_raw);
    cplx<T>* tz = ty + ny;
    cplx<T>* b0 = tz + nz;
    cplx<T>* b1 = b0 + np;
    const int x = blockIdx.x % nx, c = blockIdx.x / nx;
    const int q0 = 2 * c, q1 = 2 * c + 1;
    const size_t ns = (size_t)nx * np, base = (size_t)x * np;
    if (c == (NM + 1) / 2) {
        spec_count_plane<T, SPEC_PLANE_THREADS>(main + (size_t)NM * ns + base, inflow + (size_t)NM * ns + base, np,
                                                rho_scale, rho, rec, base, rs);
        return;
    }
    for (int t = threadIdx.x; t < ny; t += SPEC_PLANE_THREADS) ty[t] = twy[t];
    for (int t = threadIdx.x; t < nz; t += SPEC_PLANE_THREADS) tz[t] = twz[t];
    const long long* m0 = main + (size_t)q0 * ns + base;
    long long* i0 = inflow + (size_t)q0 * ns + base;
    const long long* m1 = main + (size_t)q1 * ns + base;
    long long* i1 = inflow + (size_t)q1 * ns + base;
    if (q1 < NM)
        batched_loop<SPEC_PLANE_THREADS>(
            np, [&](int t) { return longlong2{m0[t] + i0[t], m1[t] + i1[t]}; },
            [&](int t, longlong2 v) {
                i0[t] = 0;
                i1[t] = 0;
                b0[t] = cplx<T>{(T)((double)v.x * scale), (T)((double)v.y * scale)};
            });
    else
        batched_loop<SPEC_PLANE_THREADS>(
            np, [&](int t) { return m0[t] + i0[t]; },
            [&](int t, long long v) {
                i0[t] = 0;
                b0[t] = cplx<T>{(T)((double)v * scale), T(0)};
            });
    __syncthreads();
    cplx<T>* r2 = spec_plane_fwd<T, P>(b0, b1, ty, tz, ay, az, ny, nz);
    cplx<T>* out = spec + (size_t)c * ns + base;
    for (int t = threadIdx.x; t < np; t += SPEC_PLANE_THREADS) out[t] = r2[t];
}

// (a - conj b) / 2i
template <typename T>
__device__ inline cplx<T> half_diff_over_i(cplx<T> a, cplx<T> b) {
    return cplx<T>{(a.y + b.y) * T(0.5), (b.x - a.x) * T(0.5)};
}

// Pencil pairs g*pb .. g*pb+pb-1 of `pairs` (pencil p, mirror m; p == m for
// pencils that are their own mirror).  npk packed stress components in, nout
// outputs (u, W [, E]) out as nout/2 packed fields.  Line (c, s, w) of pair s,
// pencil w (0: p, 1: m) sits at A[((c pb + s) 2 + w) (nx + 1)].
template <typename T, int P>
__global__ void __launch_bounds__(SPEC_PENCIL_THREADS)
    k_spec_pencils2(const cplx<T>* __restrict__ in, cplx<T>* __restrict__ out, const uint2* __restrict__ pairs,
                    int npairs, int np, int nz, int ny, int pb_max, int nsig, int nout, SpecAxis ax,
                    const cplx<T>* __restrict__ twx, const T* __restrict__ kx, const T* __restrict__ ky,
                    const T* __restrict__ kz, const T* __restrict__ ik2, const T* __restrict__ kinv, T screening,
                    T norm) {
    extern __shared__ __attribute__((aligned(16))) unsigned char smem_raw[];
    const int nx = P ? P : ax.n, ls = nx + 1;
    const int npk = (nsig + 1) / 2, nop = nout / 2;
    const int nl = 2 * (npk > nop ? npk : nop) * pb_max;
    cplx<T>* tw = reinterpret_cast<cplx<T>*>(smem_raw);
    cplx<T>* A = tw + nx;
    cplx<T>* B = A + (size_t)nl * ls;
    const size_t ns = (size_t)nx * np;
    // XCD-aware order, as in k_spec_pencils
    const int groups = gridDim.x;
    const int g = groups % 8 == 0 ? (blockIdx.x % 8) * (groups / 8) + blockIdx.x / 8 : blockIdx.x;
    const int s0 = g * pb_max;
    const int pb = min(pb_max, npairs - s0);
    __shared__ uint2 prs[SPEC_MAX_PENCILS];
    for (int t = threadIdx.x; t < nx; t += SPEC_PENCIL_THREADS) tw[t] = twx[t];
    if (threadIdx.x < pb) prs[threadIdx.x] = pairs[s0 + threadIdx.x];
    __syncthreads();
    // lanes over (w, s) for loads of neighbouring pencils
    batched_loop<SPEC_PENCIL_THREADS>(
        npk * nx * 2 * pb_max,
        [&](int f) {
            const int ws = f % (2 * pb_max), xx = (f / (2 * pb_max)) % nx, c = f / (2 * pb_max * nx);
            const int s = ws >> 1, w = ws & 1;
            cplx<T> v = cplx<T>{T(0), T(0)};
            if (s < pb) v = in[(size_t)c * ns + (size_t)xx * np + (w ? prs[s].y : prs[s].x)];
            return v;
        },
        [&](int f, cplx<T> v) {
            const int ws = f % (2 * pb_max), xx = (f / (2 * pb_max)) % nx, c = f / (2 * pb_max * nx);
            A[(size_t)((c * pb_max + (ws >> 1)) * 2 + (ws & 1)) * ls + xx] = v;
        });
    __syncthreads();
    cplx<T>* S = spec_lines<T, false, SPEC_PENCIL_THREADS, P, 0>(A, B, tw, ax, npk * pb_max * 2, ls, 1, -1);
    cplx<T>* O = S == A ? B : A;
    // one thread per mirror pair of bins: (p, i) and (m, -i)
    for (int f = threadIdx.x; f < nx * pb_max; f += SPEC_PENCIL_THREADS) {
        const int s = f % pb_max, i = f / pb_max;
        const int im = i == 0 ? 0 : nx - i;
        if (s >= pb) {
            for (int d = 0; d < nop; ++d) {
                O[(size_t)((d * pb_max + s) * 2 + 0) * ls + i] = cplx<T>{T(0), T(0)};
                O[(size_t)((d * pb_max + s) * 2 + 1) * ls + i] = cplx<T>{T(0), T(0)};
            }
            continue;
        }
        const uint2 pr = prs[s];
        const int p = (int)pr.x, m = (int)pr.y;
        if (p == m && i > im) continue;  // its own mirror: the pair (i, -i) once
        // unpacked stress spectra at (p, i); at (m, -i) they are the conjugates
        cplx<T> sg[9];
        for (int c = 0; c < npk; ++c) {
            const cplx<T> zp = S[(size_t)((c * pb_max + s) * 2 + 0) * ls + i];
            const cplx<T> zm = S[(size_t)((c * pb_max + s) * 2 + 1) * ls + im];
            sg[2 * c] = cplx<T>{(zp.x + zm.x) * T(0.5), (zp.y - zm.y) * T(0.5)};
            if (2 * c + 1 < nsig) sg[2 * c + 1] = half_diff_over_i(zp, zm);
        }
        const KPoint<T> q1 = KPoint<T>::load(kx, ky, kz, kinv, ik2, i, p, np, nz);
        const KPoint<T> q2 = KPoint<T>::load(kx, ky, kz, kinv, ik2, im, m, np, nz);
        const T* k1 = q1.k;
        const T* k2 = q2.k;
        T u1r[3], u1i[3], u2r[3], u2i[3];
        flow_point<T>([&](int q) { return sg[q]; }, nsig, k1, q1.ki, q1.ik2, screening, norm, u1r, u1i);
        flow_point<T>([&](int q) { return cplx<T>{sg[q].x, -sg[q].y}; }, nsig, k2, q2.ki, q2.ik2, screening, norm,
                      u2r, u2i);
        // outputs at k and at -k, then their Hermitian part
        cplx<T> h[12];
        for (int o = 0; o < nout; ++o) {
            cplx<T> a, b;
            if (o < 3) {
                a = cplx<T>{u1r[o], u1i[o]};
                b = cplx<T>{u2r[o], u2i[o]};
            } else if (o < 6) {
                a = vort_part(k1, u1r, u1i, o - 3);
                b = vort_part(k2, u2r, u2i, o - 3);
            } else {
                a = strain_part(k1, u1r, u1i, o - 6);
                b = strain_part(k2, u2r, u2i, o - 6);
            }
            h[o] = cplx<T>{(a.x + b.x) * T(0.5), (a.y - b.y) * T(0.5)};
        }
        // packed outputs 2d + i (2d+1) at (p, i) and, with H(-k) = conj H(k), at (m, -i)
        for (int d = 0; d < nop; ++d) {
            const cplx<T> h0 = h[2 * d], h1 = h[2 * d + 1];
            O[(size_t)((d * pb_max + s) * 2 + 0) * ls + i] = cplx<T>{h0.x - h1.y, h0.y + h1.x};
            O[(size_t)((d * pb_max + s) * 2 + 1) * ls + im] = cplx<T>{h0.x + h1.y, h1.x - h0.y};
            if (p == m) {  // the same pencil: (p, -i) too
                O[(size_t)((d * pb_max + s) * 2 + 0) * ls + im] = cplx<T>{h0.x + h1.y, h1.x - h0.y};
                O[(size_t)((d * pb_max + s) * 2 + 1) * ls + i] = cplx<T>{h0.x - h1.y, h0.y + h1.x};
            }
        }
    }
    __syncthreads();
    cplx<T>* R = spec_lines<T, false, SPEC_PENCIL_THREADS, P, 0>(O, S, tw, ax, nop * pb_max * 2, ls, 1, +1);
    for (int f = threadIdx.x; f < nop * nx * 2 * pb_max; f += SPEC_PENCIL_THREADS) {
        const int ws = f % (2 * pb_max), xx = (f / (2 * pb_max)) % nx, d = f / (2 * pb_max * nx);
        const int s = ws >> 1, w = ws & 1;
        if (s >= pb) continue;
        const uint2 pr = prs[s];
        if (w == 1 && pr.x == pr.y) continue;
        out[(size_t)d * ns + (size_t)xx * np + (w ? pr.y : pr.x)] = R[(size_t)((d * pb_max + s) * 2 + w) * ls + xx];
    }
}

// Inverse: packed field d (outputs 2d + i 2d+1) of plane x; the real part is
// output 2d, the imaginary part output 2d+1.
template <typename T, int P>
__global__ void __launch_bounds__(SPEC_PLANE_THREADS)
    k_spec_planes_inv2(const cplx<T>* __restrict__ spec, int nx, SpecAxis ay, SpecAxis az,
                       const cplx<T>* __restrict__ twy, const cplx<T>* __restrict__ twz, T* __restrict__ rec,
                       int rs) {
    extern __shared__ __attribute__((aligned(16))) unsigned char smem_raw[];
    const int ny = P ? P : ay.n, nz = P ? P : az.n, np = ny * nz;
    cplx<T>* ty = reinterpret_cast<cplx<T>*>(smem_raw);
    cplx<T>* tz = ty + ny;
    cplx<T>* b0 = tz + nz;
    cplx<T>* b1 = b0 + np;
    const int x = blockIdx.x % nx, d = blockIdx.x / nx;
    const size_t ns = (size_t)nx * np, base = (size_t)x * np;
    for (int t = threadIdx.x; t < ny; t += SPEC_PLANE_THREADS) ty[t] = twy[t];
    for (int t = threadIdx.x; t < nz; t += SPEC_PLANE_THREADS) tz[t] = twz[t];
    const cplx<T>* src = spec + (size_t)d * ns + base;
    batched_loop<SPEC_PLANE_THREADS>(np, [&](int t) { return src[t]; }, [&](int t, cplx<T> v) { b0[t] = v; });
    __syncthreads();
    cplx<T>* r2 = spec_plane_inv<T, P>(b0, b1, ty, tz, ay, az, ny, nz);
    const int o0 = 2 * d, o1 = 2 * d + 1;
    const int sl0 = o0 < 6 ? o0 : REC_E + o0 - 6, sl1 = o1 < 6 ? o1 : REC_E + o1 - 6;
    for (int t = threadIdx.x; t < np; t += SPEC_PLANE_THREADS) {
        rec[(base + t) * rs + sl0] = r2[t].x;
        rec[(base + t) * rs + sl1] = r2[t].y;
    }
}

// Getter path: packed z,y-transformed moments -> the per-component layout
// (plane x keeps its own Hermitian symmetry in (y, z)).
template <typename T>
__global__ void __launch_bounds__(256) k_spec_unpack(const cplx<T>* __restrict__ packed, int nsig, int nx, int ny,
                                                     int nz, cplx<T>* __restrict__ out) {
    const size_t np = (size_t)ny * nz, ns = (size_t)nx * np;
    const size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= ns) return;
    const size_t x = t / np, p = t % np;
    const int y = (int)(p / nz), z = (int)(p % nz);
    const size_t pm = (size_t)((ny - y) % ny) * nz + (size_t)((nz - z) % nz);
    for (int c = 0; 2 * c < nsig; ++c) {
        const cplx<T> zp = packed[(size_t)c * ns + x * np + p];
        const cplx<T> zm = packed[(size_t)c * ns + x * np + pm];
        out[(size_t)(2 * c) * ns + t] = cplx<T>{(zp.x + zm.x) * T(0.5), (zp.y - zm.y) * T(0.5)};
        if (2 * c + 1 < nsig) out[(size_t)(2 * c + 1) * ns + t] = half_diff_over_i(zp, zm);
    }
}

}  // namespace ss
