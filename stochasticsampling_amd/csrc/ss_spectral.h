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
// solve is flow_point, so both paths give identical results.
#pragma once
#include "ss_fft_device.h"
#include "ss_kernels.h"

namespace ss {

// Workgroup sizes: the plane kernels keep one plane per workgroup, the
// pencil kernel a group of pencils; more threads per workgroup give each
// workgroup's LDS-bound stages more waves to overlap.
constexpr int SPEC_PLANE_THREADS = 512;
constexpr int SPEC_PENCIL_THREADS = 512;

// LDS of one plane kernel: twiddles of y and z, two plane buffers.
inline size_t spec_plane_lds(int ny, int nz, size_t esz) { return esz * 2 * ((size_t)ny + nz + 2 * (size_t)ny * nz); }

// Forward: component q of plane x.  Blocks with q == 0 also convert the count
// component (density, when wanted).  Clears the inflow entries it reads.
template <typename T, int NM>
__global__ void __launch_bounds__(SPEC_PLANE_THREADS)
    k_spec_planes_fwd(const long long* __restrict__ main, long long* __restrict__ inflow, int nx, double scale,
                      double rho_scale, SpecAxis ay, SpecAxis az, const cplx<T>* __restrict__ twy,
                      const cplx<T>* __restrict__ twz, cplx<T>* __restrict__ spec, T* __restrict__ rho,
                      T* __restrict__ rec, int rs) {
    extern __shared__ __attribute__((aligned(16))) unsigned char smem_raw[];
    const int ny = ay.n, nz = az.n, np = ny * nz;
    cplx<T>* ty = reinterpret_cast<cplx<T>*>(smem_raw);
    cplx<T>* tz = ty + ny;
    cplx<T>* b0 = tz + nz;
    cplx<T>* b1 = b0 + np;
    const int x = blockIdx.x % nx, q = blockIdx.x / nx;
    const size_t ns = (size_t)nx * np, base = (size_t)x * np;
    for (int t = threadIdx.x; t < ny; t += blockDim.x) ty[t] = twy[t];
    for (int t = threadIdx.x; t < nz; t += blockDim.x) tz[t] = twz[t];
    const long long* mq = main + (size_t)q * ns + base;
    long long* iq = inflow + (size_t)q * ns + base;
    for (int c = threadIdx.x; c < np; c += blockDim.x) {
        const long long v = mq[c] + iq[c];
        iq[c] = 0;
        b0[c] = cplx<T>{(T)((double)v * scale), T(0)};
    }
    if (q == 0) {
        const long long* mc = main + (size_t)NM * ns + base;
        long long* ic = inflow + (size_t)NM * ns + base;
        for (int c = threadIdx.x; c < np; c += blockDim.x) {
            const long long v = mc[c] + ic[c];
            ic[c] = 0;
            if (rho) {
                const T r = (T)((double)v * rho_scale);
                rho[base + c] = r;
                rec[(base + c) * rs + REC_RHO] = r;
            }
        }
    }
    __syncthreads();
    // z lines: ny contiguous lines; y lines: nz lines interleaved at stride nz
    cplx<T>* r1 = fft_lines<T, false>(b0, b1, tz, az, ny, nz, 1, -1);
    cplx<T>* r2 = fft_lines<T, true>(r1, r1 == b0 ? b1 : b0, ty, ay, nz, 1, nz, -1);
    cplx<T>* out = spec + (size_t)q * ns + base;
    for (int c = threadIdx.x; c < np; c += blockDim.x) out[c] = r2[c];
}

// LDS of the pencil kernel: x twiddles and two buffers of nl lines of
// nx + 1 points (padded against bank conflicts).
inline size_t spec_pencil_lds(int nx, int nl, size_t esz) {
    return esz * 2 * ((size_t)nx + 2 * (size_t)nl * (nx + 1));
}

// Pencils p0 .. p0+pb-1 (p = y nz + z) of the nsig stress components: forward
// x transform, flow solve, inverse x transform of the nout = 6 (u, W) or 12
// (u, W, E) outputs.
template <typename T>
__global__ void __launch_bounds__(SPEC_PENCIL_THREADS)
    k_spec_pencils(const cplx<T>* __restrict__ in, cplx<T>* __restrict__ out, int np, int nz, int pb_max,
                   int nsig, int nout, SpecAxis ax, const cplx<T>* __restrict__ twx, const T* __restrict__ kx,
                   const T* __restrict__ ky, const T* __restrict__ kz, const T* __restrict__ ik2,
                   const T* __restrict__ kinv, T screening, T norm) {
    extern __shared__ __attribute__((aligned(16))) unsigned char smem_raw[];
    const int nx = ax.n, ls = nx + 1;
    const int nl = (nsig > nout ? nsig : nout) * pb_max;
    cplx<T>* tw = reinterpret_cast<cplx<T>*>(smem_raw);
    cplx<T>* A = tw + nx;
    cplx<T>* B = A + (size_t)nl * ls;
    const size_t ns = (size_t)nx * np;
    // workgroups are dispatched round-robin over the XCDs: give each XCD a
    // contiguous run of pencil groups, so that the two groups sharing a
    // cache line of every x row are read through the same L2
    const int groups = gridDim.x;
    const int g = groups % 8 == 0 ? (blockIdx.x % 8) * (groups / 8) + blockIdx.x / 8 : blockIdx.x;
    const int p0 = g * pb_max;
    const int pb = min(pb_max, np - p0);
    for (int t = threadIdx.x; t < nx; t += blockDim.x) tw[t] = twx[t];
    // line (c, p) at A[(c pb_max + p) ls]; lanes over p for coalesced loads
    for (int f = threadIdx.x; f < nsig * nx * pb_max; f += blockDim.x) {
        const int p = f % pb_max, xx = (f / pb_max) % nx, c = f / (pb_max * nx);
        A[(size_t)(c * pb_max + p) * ls + xx] =
            p < pb ? in[(size_t)c * ns + (size_t)xx * np + p0 + p] : cplx<T>{T(0), T(0)};
    }
    __syncthreads();
    cplx<T>* S = fft_lines<T, false>(A, B, tw, ax, nsig * pb_max, ls, 1, -1);
    cplx<T>* O = S == A ? B : A;
    for (int f = threadIdx.x; f < nx * pb_max; f += blockDim.x) {
        const int p = f % pb_max, xx = f / pb_max;
        if (p >= pb) {
            for (int o = 0; o < nout; ++o) O[(size_t)(o * pb_max + p) * ls + xx] = cplx<T>{T(0), T(0)};
            continue;
        }
        const int pp = p0 + p;
        const size_t idx = (size_t)xx * np + pp;
        const T k[3] = {kx[xx], ky[pp / nz], kz[pp % nz]};
        T ur[3], ui[3];
        flow_point<T>([&](int q) { return S[(size_t)(q * pb_max + p) * ls + xx]; }, nsig, k, kinv[idx], ik2[idx],
                      screening, norm, ur, ui);
#pragma unroll
        for (int i = 0; i < 3; ++i) O[(size_t)(i * pb_max + p) * ls + xx] = cplx<T>{ur[i], ui[i]};
#pragma unroll
        for (int w = 0; w < 3; ++w) O[(size_t)((3 + w) * pb_max + p) * ls + xx] = vort_part(k, ur, ui, w);
        if (nout == 12) {
#pragma unroll
            for (int w = 0; w < 6; ++w) O[(size_t)((6 + w) * pb_max + p) * ls + xx] = strain_part(k, ur, ui, w);
        }
    }
    __syncthreads();
    cplx<T>* R = fft_lines<T, false>(O, S, tw, ax, nout * pb_max, ls, 1, +1);
    for (int f = threadIdx.x; f < nout * nx * pb_max; f += blockDim.x) {
        const int p = f % pb_max, xx = (f / pb_max) % nx, o = f / (pb_max * nx);
        if (p < pb) out[(size_t)o * ns + (size_t)xx * np + p0 + p] = R[(size_t)(o * pb_max + p) * ls + xx];
    }
}

// Inverse: field component o of plane x; the real part goes to record slot
// o (u, W) or REC_E + o - 6 (E).
template <typename T>
__global__ void __launch_bounds__(SPEC_PLANE_THREADS)
    k_spec_planes_inv(const cplx<T>* __restrict__ spec, int nx, SpecAxis ay, SpecAxis az,
                      const cplx<T>* __restrict__ twy, const cplx<T>* __restrict__ twz, T* __restrict__ rec,
                      int rs) {
    extern __shared__ __attribute__((aligned(16))) unsigned char smem_raw[];
    const int ny = ay.n, nz = az.n, np = ny * nz;
    cplx<T>* ty = reinterpret_cast<cplx<T>*>(smem_raw);
    cplx<T>* tz = ty + ny;
    cplx<T>* b0 = tz + nz;
    cplx<T>* b1 = b0 + np;
    const int x = blockIdx.x % nx, o = blockIdx.x / nx;
    const size_t ns = (size_t)nx * np, base = (size_t)x * np;
    for (int t = threadIdx.x; t < ny; t += blockDim.x) ty[t] = twy[t];
    for (int t = threadIdx.x; t < nz; t += blockDim.x) tz[t] = twz[t];
    const cplx<T>* src = spec + (size_t)o * ns + base;
    for (int c = threadIdx.x; c < np; c += blockDim.x) b0[c] = src[c];
    __syncthreads();
    cplx<T>* r1 = fft_lines<T, true>(b0, b1, ty, ay, nz, 1, nz, +1);
    cplx<T>* r2 = fft_lines<T, false>(r1, r1 == b0 ? b1 : b0, tz, az, ny, nz, 1, +1);
    const int slot = o < 6 ? o : REC_E + o - 6;
    for (int c = threadIdx.x; c < np; c += blockDim.x) rec[(base + c) * rs + slot] = r2[c].x;
}

}  // namespace ss
