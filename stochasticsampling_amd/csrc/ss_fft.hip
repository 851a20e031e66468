// ss_fft.hip -- LDS-staged mixed-radix Stockham FFT for gfx950, batched 3-D
// c2c in place (the transform FFTW performs for the reference,
// crates/fftw3/src/fft.rs:156-231: unnormalised, sign -1 forward).
//
// A 3-D transform is three passes, one per axis.  Each workgroup stages a tile
// of TL lines (n <= 1024 points each) in LDS, loading them so that
// neighbouring lanes touch neighbouring addresses (contiguous lines for the
// z axis, TL adjacent lines per element for the strided x and y axes), runs
// the radix stages between two LDS buffers and writes the tile back.
#include <math.h>

#include <cstring>

#include "ss_fft_device.h"
#include "ss_internal.h"

namespace ss {

int factorize(int n, int* f) {
    int k = 0, m = n;
    while (m % 4 == 0) { f[k++] = 4; m /= 4; }
    while (m % 2 == 0) { f[k++] = 2; m /= 2; }
    while (m % 3 == 0) { f[k++] = 3; m /= 3; }
    while (m % 5 == 0) { f[k++] = 5; m /= 5; }
    for (int q = 7; m > 1; q += 2)
        while (m % q == 0) {
            f[k++] = q;
            m /= q;
        }
    return k;
}

struct AxisDesc {
    int n, nf;
    int fac[16];
    int tl;            // lines per tile
    int lines;         // lines per outer index
    int tiles;         // tiles per outer index
    long long outer_stride;
    long long line_stride;
    long long elem_stride;
    int contiguous;    // lines are contiguous runs of n elements
};

// Generic radix for odd primes above 5 (test sizes such as 7, 11, 13).
template <typename T>
__device__ inline void dft_generic(cplx<T>* v, int R, const cplx<T>* tw, int n, int sign) {
    cplx<T> w[64];
    const int step = n / R;
    for (int q = 0; q < R; ++q) {
        T sr = T(0), si = T(0);
        for (int r = 0; r < R; ++r) {
            cplx<T> c = tw[((r * q) % R) * step];
            if (sign > 0) c.y = -c.y;
            sr += v[r].x * c.x - v[r].y * c.y;
            si += v[r].x * c.y + v[r].y * c.x;
        }
        w[q] = {sr, si};
    }
    for (int q = 0; q < R; ++q) v[q] = w[q];
}

template <typename T>
__device__ inline void stage_generic(const cplx<T>* src, cplx<T>* dst, const cplx<T>* tw, int n,
                                     int Ns, int R, int nlines, int sign) {
    const int m = n / R;
    const int tstep = n / (Ns * R);
    for (int b = threadIdx.x; b < nlines * m; b += blockDim.x) {
        const int l = b / m, j = b - l * m;
        const int k = j % Ns;
        const cplx<T>* s = src + l * n;
        cplx<T> v[64];
        for (int r = 0; r < R; ++r) {
            cplx<T> x = s[j + r * m];
            if (r > 0 && k > 0) {
                cplx<T> c = tw[(r * k * tstep) % n];
                if (sign > 0) c.y = -c.y;
                x = cmul(x, c);
            }
            v[r] = x;
        }
        dft_generic<T>(v, R, tw, n, sign);
        const int idxD = (j / Ns) * Ns * R + k;
        cplx<T>* d = dst + l * n;
        for (int r = 0; r < R; ++r) d[idxD + r * Ns] = v[r];
    }
}

template <typename T, bool GENERIC>
__global__ void __launch_bounds__(256) k_fft_axis(cplx<T>* __restrict__ data, AxisDesc d,
                                                  const cplx<T>* __restrict__ twg, int sign) {
    extern __shared__ __attribute__((aligned(16))) unsigned char smem_raw[];
    cplx<T>* tw = reinterpret_cast<cplx<T>*>(smem_raw);
    cplx<T>* buf0 = tw + d.n;
    cplx<T>* buf1 = buf0 + d.tl * d.n;

    const int outer = blockIdx.x / d.tiles;
    const int tile = blockIdx.x - outer * d.tiles;
    const int l0 = tile * d.tl;
    const int nl = min(d.tl, d.lines - l0);
    cplx<T>* base = data + (long long)outer * d.outer_stride + (long long)l0 * d.line_stride;

    for (int t = threadIdx.x; t < d.n; t += blockDim.x) tw[t] = twg[t];
    const int total = d.tl * d.n;
    for (int f = threadIdx.x; f < total; f += blockDim.x) {
        int l, i;
        if (d.contiguous) {
            l = f / d.n;
            i = f - l * d.n;
        } else {
            i = f / d.tl;
            l = f - i * d.tl;
        }
        if (l < nl) buf0[l * d.n + i] = base[(long long)l * d.line_stride + (long long)i * d.elem_stride];
    }
    __syncthreads();

    cplx<T>* src = buf0;
    cplx<T>* dst = buf1;
    int Ns = 1;
    for (int s = 0; s < d.nf; ++s) {
        const int R = d.fac[s];
        switch (R) {
            case 4: stage<T, 4, false>(src, dst, tw, d.n, Ns, nl, d.n, 1, sign); break;
            case 2: stage<T, 2, false>(src, dst, tw, d.n, Ns, nl, d.n, 1, sign); break;
            case 3: stage<T, 3, false>(src, dst, tw, d.n, Ns, nl, d.n, 1, sign); break;
            case 5: stage<T, 5, false>(src, dst, tw, d.n, Ns, nl, d.n, 1, sign); break;
            default:
                if constexpr (GENERIC) stage_generic<T>(src, dst, tw, d.n, Ns, R, nl, sign);
                break;
        }
        __syncthreads();
        cplx<T>* t = src;
        src = dst;
        dst = t;
        Ns *= R;
    }

    for (int f = threadIdx.x; f < total; f += blockDim.x) {
        int l, i;
        if (d.contiguous) {
            l = f / d.n;
            i = f - l * d.n;
        } else {
            i = f / d.tl;
            l = f - i * d.tl;
        }
        if (l < nl) base[(long long)l * d.line_stride + (long long)i * d.elem_stride] = src[l * d.n + i];
    }
}

int Fft3d::init(int nx_, int ny_, int nz_, int precision_) {
    nx = nx_;
    ny = ny_;
    nz = nz_;
    precision = precision_;
    int dims[3] = {nx, ny, nz};
    for (int a = 0; a < 3; ++a) {
        int n = dims[a];
        if (n < 1 || n > 1024) return fail(SS_ERR_UNSUPPORTED, "FFT axis length must be in [1, 1024]");
        ax[a].n = n;
        ax[a].nf = factorize(n, ax[a].fac);
        for (int s = 0; s < ax[a].nf; ++s)
            if (ax[a].fac[s] > 64) return fail(SS_ERR_UNSUPPORTED, "FFT axis has a prime factor above 64");
        const size_t esz = precision == SS_FP64 ? 16 : 8;
        std::vector<unsigned char> host(esz * n);
        for (int t = 0; t < n; ++t) {
            double ang = 2.0 * M_PI * (double)t / (double)n;
            double c = cos(ang), s = -sin(ang);
            if (precision == SS_FP64) {
                double* h = reinterpret_cast<double*>(host.data());
                h[2 * t] = c;
                h[2 * t + 1] = s;
            } else {
                float* h = reinterpret_cast<float*>(host.data());
                h[2 * t] = (float)c;
                h[2 * t + 1] = (float)s;
            }
        }
        SS_HIP_CHECK(hipMalloc(&ax[a].tw, esz * n));
        SS_HIP_CHECK(hipMemcpy(ax[a].tw, host.data(), esz * n, hipMemcpyHostToDevice));
    }
    return SS_OK;
}

void Fft3d::release() {
    for (int a = 0; a < 3; ++a)
        if (ax[a].tw) {
            (void)hipFree(ax[a].tw);
            ax[a].tw = nullptr;
        }
}

template <typename T>
static int run_axis(cplx<T>* data, const FftAxis& fa, int axis, int nx, int ny, int nz, int batch,
                    int sign, hipStream_t stream) {
    AxisDesc d;
    std::memset(&d, 0, sizeof(d));
    d.n = fa.n;
    d.nf = fa.nf;
    for (int s = 0; s < fa.nf; ++s) d.fac[s] = fa.fac[s];
    const long long nyz = (long long)ny * nz;
    const long long vol = (long long)nx * nyz;
    if (d.n == 1) return SS_OK;
    bool generic = false;
    for (int s = 0; s < fa.nf; ++s)
        if (fa.fac[s] > 5) generic = true;
    if (axis == 2) {
        d.contiguous = 1;
        d.tl = d.n >= 1024 ? 1 : 1024 / d.n;
        d.lines = (int)((long long)batch * nx * ny);
        d.outer_stride = 0;
        d.line_stride = nz;
        d.elem_stride = 1;
        d.tiles = (d.lines + d.tl - 1) / d.tl;
        dim3 grid(d.tiles);
        size_t lds = sizeof(cplx<T>) * (d.n + 2 * d.tl * d.n);
        if (generic)
            hipLaunchKernelGGL((k_fft_axis<T, true>), grid, dim3(256), lds, stream, data, d,
                               reinterpret_cast<const cplx<T>*>(fa.tw), sign);
        else
            hipLaunchKernelGGL((k_fft_axis<T, false>), grid, dim3(256), lds, stream, data, d,
                               reinterpret_cast<const cplx<T>*>(fa.tw), sign);
    } else {
        d.contiguous = 0;
        int tl = 16;
        while (tl > 1 && tl * d.n > 1024) tl /= 2;
        d.tl = tl;
        if (axis == 1) {
            d.lines = nz;
            d.outer_stride = nyz;
            d.elem_stride = nz;
        } else {
            d.lines = (int)nyz;
            d.outer_stride = vol;
            d.elem_stride = nyz;
        }
        d.line_stride = 1;
        d.tiles = (d.lines + d.tl - 1) / d.tl;
        long long outers = axis == 1 ? (long long)batch * nx : (long long)batch;
        dim3 grid((unsigned)(outers * d.tiles));
        size_t lds = sizeof(cplx<T>) * (d.n + 2 * d.tl * d.n);
        if (generic)
            hipLaunchKernelGGL((k_fft_axis<T, true>), grid, dim3(256), lds, stream, data, d,
                               reinterpret_cast<const cplx<T>*>(fa.tw), sign);
        else
            hipLaunchKernelGGL((k_fft_axis<T, false>), grid, dim3(256), lds, stream, data, d,
                               reinterpret_cast<const cplx<T>*>(fa.tw), sign);
    }
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return fail(SS_ERR_DEVICE, std::string("fft launch: ") + hipGetErrorString(e));
    return SS_OK;
}

int Fft3d::run_axis(void* data, int batch, int sign, int axis, hipStream_t stream) const {
    if (precision == SS_FP64)
        return ss::run_axis<double>(reinterpret_cast<cplx<double>*>(data), ax[axis], axis, nx, ny, nz, batch, sign,
                                    stream);
    return ss::run_axis<float>(reinterpret_cast<cplx<float>*>(data), ax[axis], axis, nx, ny, nz, batch, sign,
                               stream);
}

int Fft3d::run(void* data, int batch, int sign, hipStream_t stream) const {
    for (int k = 0; k < 3; ++k) {
        const int axis = sign < 0 ? 2 - k : k;
        const int rc = run_axis(data, batch, sign, axis, stream);
        if (rc != SS_OK) return rc;
    }
    return SS_OK;
}

bool Fft3d::small_radices() const {
    for (int a = 0; a < 3; ++a)
        for (int s = 0; s < ax[a].nf; ++s)
            if (ax[a].fac[s] > 5) return false;
    return true;
}

}  // namespace ss

// ---------------------------------------------------------------------------
// C ABI: FFTPlan replacement (crates/fftw3/src/fft.rs:33-244).
// ---------------------------------------------------------------------------
struct ss_fft_plan {
    ss::Fft3d fft;
    int direction = -1;
    int batch = 1;
    int device = 0;
    hipStream_t stream = nullptr;
};

extern "C" int ss_fft_plan_create(int nx, int ny, int nz, int direction, int batch, int precision,
                                  int variant, int device, ss_fft_plan** out) {
    if (!out) return ss::fail(SS_ERR_INVALID, "out is NULL");
    *out = nullptr;
    if (direction != -1 && direction != 1) return ss::fail(SS_ERR_INVALID, "direction must be -1 or +1");
    if (batch < 1) return ss::fail(SS_ERR_INVALID, "batch must be >= 1");
    if (precision != SS_FP64 && precision != SS_FP32) return ss::fail(SS_ERR_INVALID, "bad precision");
    if (variant != SS_FFT_AUTO && variant != SS_FFT_RADIX)
        return ss::fail(SS_ERR_UNSUPPORTED, "only the radix FFT variant is built");
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) return ss::fail(SS_ERR_DEVICE, "no GPU");
    if (device < 0 || device >= ndev) return ss::fail(SS_ERR_INVALID, "bad device ordinal");
    SS_HIP_CHECK(hipSetDevice(device));
    ss_fft_plan* p = new ss_fft_plan();
    p->direction = direction;
    p->batch = batch;
    p->device = device;
    int rc = p->fft.init(nx, ny, nz, precision);
    if (rc != SS_OK) {
        p->fft.release();
        delete p;
        return rc;
    }
    if (hipStreamCreate(&p->stream) != hipSuccess) {
        p->fft.release();
        delete p;
        return ss::fail(SS_ERR_DEVICE, "hipStreamCreate failed");
    }
    *out = p;
    return SS_OK;
}

extern "C" int ss_fft_execute(ss_fft_plan* p, void* device_data, void* hip_stream) {
    if (!p || !device_data) return ss::fail(SS_ERR_INVALID, "NULL argument");
    SS_HIP_CHECK(hipSetDevice(p->device));
    hipStream_t s = hip_stream ? reinterpret_cast<hipStream_t>(hip_stream) : p->stream;
    int rc = p->fft.run(device_data, p->batch, p->direction, s);
    if (rc != SS_OK) return rc;
    if (!hip_stream) SS_HIP_CHECK(hipStreamSynchronize(s));
    return SS_OK;
}

extern "C" int ss_fft_execute_host(ss_fft_plan* p, void* host_data) {
    if (!p || !host_data) return ss::fail(SS_ERR_INVALID, "NULL argument");
    SS_HIP_CHECK(hipSetDevice(p->device));
    size_t bytes = (size_t)p->fft.nx * p->fft.ny * p->fft.nz * p->batch *
                   (p->fft.precision == SS_FP64 ? 16 : 8);
    void* d = nullptr;
    SS_HIP_CHECK(hipMalloc(&d, bytes));
    int rc = SS_OK;
    if (hipMemcpyAsync(d, host_data, bytes, hipMemcpyHostToDevice, p->stream) != hipSuccess)
        rc = ss::fail(SS_ERR_DEVICE, "copy in failed");
    if (rc == SS_OK) rc = p->fft.run(d, p->batch, p->direction, p->stream);
    if (rc == SS_OK && hipMemcpyAsync(host_data, d, bytes, hipMemcpyDeviceToHost, p->stream) != hipSuccess)
        rc = ss::fail(SS_ERR_DEVICE, "copy out failed");
    if (hipStreamSynchronize(p->stream) != hipSuccess && rc == SS_OK)
        rc = ss::fail(SS_ERR_DEVICE, "FFT execution failed");
    (void)hipFree(d);
    return rc;
}

extern "C" int ss_fft_plan_destroy(ss_fft_plan* p) {
    if (!p) return SS_OK;
    (void)hipSetDevice(p->device);
    if (p->stream) (void)hipStreamDestroy(p->stream);
    p->fft.release();
    delete p;
    return SS_OK;
}
