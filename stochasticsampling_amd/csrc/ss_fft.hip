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

#include <atomic>
#include <cstring>
#include <string>

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
            case 4: stage<T, 4, false, 256>(src, dst, tw, d.n, Ns, nl, d.n, 1, sign); break;
            case 2: stage<T, 2, false, 256>(src, dst, tw, d.n, Ns, nl, d.n, 1, sign); break;
            case 3: stage<T, 3, false, 256>(src, dst, tw, d.n, Ns, nl, d.n, 1, sign); break;
            case 5: stage<T, 5, false, 256>(src, dst, tw, d.n, Ns, nl, d.n, 1, sign); break;
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

// ---------------------------------------------------------------------------
// Dense-DFT variant for axes of at most 64 points: each line is a complex
// vector x of n points, stored as the real 2n-vector [Re x; Im x], and the
// transform is the real matrix product [Re X; Im X] = M [Re x; Im x] with
// M = [[C, -s S], [s S, C]], C = cos(2 pi jk / n), S = sin(2 pi jk / n).  It
// runs on the matrix cores with gfx950's v_mfma_f32_32x32x16_bf16 (16-deep
// products).  Each factor is split into three bf16 parts, v = h + m + l with
// h = bf16(v), m = bf16(v - h), l = bf16(v - h - m) (24 significant bits),
// and the six products whose parts weigh 2^-16 or more (hh, hm, mh, hl, lh,
// mm) are accumulated in FP32, so each product keeps about 24 bits, the
// accuracy of the radix FFT; three passes (hh, hm, mh over two-part splits)
// gave 8e-6 relative in u at 64^3 in the same time, the axis passes being
// bound by their memory traffic.  Lane layout: A[i][k] and B[k][j] hold
// i (or j) = lane % 32, k = 8 (lane / 32) + t, t = 0..7; D[i][j] in register v
// holds j = lane % 32, i = 8 (v / 4) + 4 (lane / 32) + v % 4.
// ---------------------------------------------------------------------------
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x16 __attribute__((ext_vector_type(16)));

constexpr int DFT_LINES = 32;  // lines (matrix columns) per workgroup

__device__ inline short bf16_rne(float x) {
    uint32_t b = __float_as_uint(x);
    b += 0x7FFFu + ((b >> 16) & 1u);
    return (short)(b >> 16);
}
__device__ inline float bf16_value(short h) { return __uint_as_float(((uint32_t)(uint16_t)h) << 16); }
__device__ inline __bf16 bf16_of_bits(short h) { return __builtin_bit_cast(__bf16, h); }

// One tile of DFT_LINES lines; four wavefronts, wavefront w computes output
// rows 32 w .. 32 w + 31 (rows <= 128).
__global__ void __launch_bounds__(256) k_dft_axis(cplx<float>* __restrict__ data, AxisDesc d,
                                                  const bf16x8* __restrict__ mh, const bf16x8* __restrict__ mm,
                                                  const bf16x8* __restrict__ ml,
                                                  int rows, int kdim) {
    extern __shared__ __attribute__((aligned(16))) unsigned char smem_raw[];
    float* xs = reinterpret_cast<float*>(smem_raw);  // [DFT_LINES][kdim + 4]
    const int n = d.n, ldx = kdim + 4;
    const int outer = blockIdx.x / d.tiles;
    const int tile = blockIdx.x - outer * d.tiles;
    const int l0 = tile * DFT_LINES;
    const int nl = min(DFT_LINES, d.lines - l0);
    cplx<float>* base = data + (long long)outer * d.outer_stride + (long long)l0 * d.line_stride;
    for (int f = threadIdx.x; f < DFT_LINES * n; f += blockDim.x) {
        int l, i;
        if (d.contiguous) {
            l = f / n;
            i = f - l * n;
        } else {
            i = f / DFT_LINES;
            l = f - i * DFT_LINES;
        }
        cplx<float> v = {0.f, 0.f};
        if (l < nl) v = base[(long long)l * d.line_stride + (long long)i * d.elem_stride];
        xs[l * ldx + i] = v.x;
        xs[l * ldx + n + i] = v.y;
    }
    for (int f = threadIdx.x; f < DFT_LINES * (kdim - 2 * n); f += blockDim.x) {
        const int l = f / (kdim - 2 * n), e = 2 * n + f % (kdim - 2 * n);
        xs[l * ldx + e] = 0.f;
    }
    __syncthreads();
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int ksteps = kdim / 16;
    f32x16 acc = {};
    const bool active = wave * 32 < rows;
    if (active) {
        for (int ks = 0; ks < ksteps; ++ks) {
            const int fr = (wave * ksteps + ks) * 64 + lane;
            const bf16x8 ah = mh[fr], am = mm[fr], al = ml[fr];
            const float* xp = xs + (lane & 31) * ldx + ks * 16 + 8 * (lane >> 5);
            bf16x8 bh, bm, bl;
#pragma unroll
            for (int t = 0; t < 8; ++t) {
                const float x = xp[t];
                const short h = bf16_rne(x);
                const float r = x - bf16_value(h);
                const short m = bf16_rne(r);
                bh[t] = bf16_of_bits(h);
                bm[t] = bf16_of_bits(m);
                bl[t] = bf16_of_bits(bf16_rne(r - bf16_value(m)));
            }
            // smallest terms first
            f32x16 lo = {};
            lo = __builtin_amdgcn_mfma_f32_32x32x16_bf16(am, bm, lo, 0, 0, 0);
            lo = __builtin_amdgcn_mfma_f32_32x32x16_bf16(al, bh, lo, 0, 0, 0);
            lo = __builtin_amdgcn_mfma_f32_32x32x16_bf16(ah, bl, lo, 0, 0, 0);
            lo = __builtin_amdgcn_mfma_f32_32x32x16_bf16(am, bh, lo, 0, 0, 0);
            lo = __builtin_amdgcn_mfma_f32_32x32x16_bf16(ah, bm, lo, 0, 0, 0);
            acc += lo;
            acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(ah, bh, acc, 0, 0, 0);
        }
    }
    __syncthreads();  // every wavefront has read its inputs
    if (active) {
#pragma unroll
        for (int v = 0; v < 16; ++v) {
            const int i = wave * 32 + (v / 4) * 8 + (lane >> 5) * 4 + (v % 4);
            if (i < 2 * n) xs[(lane & 31) * ldx + i] = acc[v];
        }
    }
    __syncthreads();
    for (int f = threadIdx.x; f < DFT_LINES * n; f += blockDim.x) {
        int l, i;
        if (d.contiguous) {
            l = f / n;
            i = f - l * n;
        } else {
            i = f / DFT_LINES;
            l = f - i * DFT_LINES;
        }
        if (l < nl)
            base[(long long)l * d.line_stride + (long long)i * d.elem_stride] =
                cplx<float>{xs[l * ldx + i], xs[l * ldx + n + i]};
    }
}

static uint16_t bf16_bits_rne(float x) {
    uint32_t b;
    std::memcpy(&b, &x, 4);
    b += 0x7FFFu + ((b >> 16) & 1u);
    return (uint16_t)(b >> 16);
}
static double bf16_double(uint16_t h) {
    const uint32_t b = (uint32_t)h << 16;
    float f;
    std::memcpy(&f, &b, 4);
    return (double)f;
}

// DFT matrices of one axis for both signs, split into bf16 high and low parts
// in fragment order [row tile][k step][lane][8].
static int build_dft(FftAxis& a) {
    const int n = a.n;
    a.dft_rows = (2 * n + 31) / 32 * 32;
    a.dft_k = (2 * n + 15) / 16 * 16;
    const int rt = a.dft_rows / 32, ks = a.dft_k / 16;
    const size_t cnt = (size_t)rt * ks * 64 * 8;
    for (int sg = 0; sg < 2; ++sg) {
        const double s = sg == 0 ? -1.0 : 1.0;
        std::vector<uint16_t> hi(cnt), mid(cnt), lo(cnt);
        auto m = [&](int i, int k) -> double {
            if (i >= 2 * n || k >= 2 * n) return 0.0;
            const int e = i % n, j = k % n;
            const double ang = 2.0 * M_PI * (double)((long long)e * j % n) / (double)n;
            const bool im_out = i >= n, im_in = k >= n;
            if (im_out == im_in) return cos(ang);
            return im_out ? s * sin(ang) : -s * sin(ang);
        };
        for (int r = 0; r < rt; ++r)
            for (int q = 0; q < ks; ++q)
                for (int l = 0; l < 64; ++l)
                    for (int t = 0; t < 8; ++t) {
                        const double v = m(r * 32 + l % 32, q * 16 + 8 * (l / 32) + t);
                        const uint16_t h = bf16_bits_rne((float)v);
                        const double rest = v - bf16_double(h);
                        const uint16_t md = bf16_bits_rne((float)rest);
                        const size_t idx = (((size_t)r * ks + q) * 64 + l) * 8 + t;
                        hi[idx] = h;
                        mid[idx] = md;
                        lo[idx] = bf16_bits_rne((float)(rest - bf16_double(md)));
                    }
        const std::vector<uint16_t>* parts[3] = {&hi, &mid, &lo};
        for (int h = 0; h < 3; ++h) {
            SS_HIP_CHECK(hipMalloc(&a.dft[sg][h], cnt * 2));
            SS_HIP_CHECK(hipMemcpy(a.dft[sg][h], parts[h]->data(), cnt * 2, hipMemcpyHostToDevice));
        }
    }
    return SS_OK;
}

int Fft3d::init(int nx_, int ny_, int nz_, int precision_, int variant_) {
    nx = nx_;
    ny = ny_;
    nz = nz_;
    precision = precision_;
    variant = variant_ == SS_FFT_DFT ? SS_FFT_DFT : SS_FFT_RADIX;
    if (variant == SS_FFT_DFT && precision != SS_FP32)
        return fail(SS_ERR_UNSUPPORTED, "the dense-DFT FFT variant is FP32 only");
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
        // axes of 2..64 points run as dense DFTs in the DFT variant, longer
        // ones keep the radix stages
        if (variant == SS_FFT_DFT && n >= 2 && n <= 64) {
            const int rc = build_dft(ax[a]);
            if (rc != SS_OK) return rc;
        }
    }
    return SS_OK;
}

void Fft3d::release() {
    for (int a = 0; a < 3; ++a) {
        if (ax[a].tw) {
            (void)hipFree(ax[a].tw);
            ax[a].tw = nullptr;
        }
        for (int sg = 0; sg < 2; ++sg)
            for (int h = 0; h < 3; ++h)
                if (ax[a].dft[sg][h]) {
                    (void)hipFree(ax[a].dft[sg][h]);
                    ax[a].dft[sg][h] = nullptr;
                }
    }
}

// One axis as dense DFTs on the matrix cores (FP32, 2..64 points).
static int run_dft_axis(cplx<float>* data, const FftAxis& fa, int axis, int nx, int ny, int nz, int batch,
                        int sign, hipStream_t stream) {
    AxisDesc d;
    std::memset(&d, 0, sizeof(d));
    d.n = fa.n;
    d.tl = DFT_LINES;
    const long long nyz = (long long)ny * nz;
    long long outers = 1;
    if (axis == 2) {
        d.contiguous = 1;
        d.lines = (int)((long long)batch * nx * ny);
        d.line_stride = nz;
        d.elem_stride = 1;
    } else if (axis == 1) {
        d.lines = nz;
        d.outer_stride = nyz;
        d.line_stride = 1;
        d.elem_stride = nz;
        outers = (long long)batch * nx;
    } else {
        d.lines = (int)nyz;
        d.outer_stride = (long long)nx * nyz;
        d.line_stride = 1;
        d.elem_stride = nyz;
        outers = batch;
    }
    d.tiles = (d.lines + DFT_LINES - 1) / DFT_LINES;
    const int sg = sign < 0 ? 0 : 1;
    const size_t lds = sizeof(float) * DFT_LINES * (fa.dft_k + 4);
    hipLaunchKernelGGL(k_dft_axis, dim3((unsigned)(outers * d.tiles)), dim3(256), lds, stream, data, d,
                       reinterpret_cast<const bf16x8*>(fa.dft[sg][0]), reinterpret_cast<const bf16x8*>(fa.dft[sg][1]),
                       reinterpret_cast<const bf16x8*>(fa.dft[sg][2]),
                       fa.dft_rows, fa.dft_k);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return fail(SS_ERR_DEVICE, std::string("dft launch: ") + hipGetErrorString(e));
    return SS_OK;
}

template <typename T>
static int run_axis(cplx<T>* data, const FftAxis& fa, int axis, int nx, int ny, int nz, int batch,
                    int sign, hipStream_t stream) {
    if constexpr (sizeof(T) == 4) {
        if (fa.dft[0][0]) return run_dft_axis(data, fa, axis, nx, ny, nz, batch, sign, stream);
    }
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
    if (variant != SS_FFT_AUTO && variant != SS_FFT_RADIX && variant != SS_FFT_DFT)
        return ss::fail(SS_ERR_INVALID, "bad FFT variant");
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) return ss::fail(SS_ERR_DEVICE, "no GPU");
    if (device < 0 || device >= ndev) return ss::fail(SS_ERR_INVALID, "bad device ordinal");
    SS_HIP_CHECK(hipSetDevice(device));
    ss_fft_plan* p = new ss_fft_plan();
    p->direction = direction;
    p->batch = batch;
    p->device = device;
    int rc = p->fft.init(nx, ny, nz, precision, variant);
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

// Staging buffer of ss_fft_execute_host_noalloc, one per GPU for the life of
// the process.  In host code its name stands for no device address: host code
// only passes the address hipGetSymbolAddress returns for the current device
// (kept per device in noalloc_ptr).
static __device__ __align__(256) unsigned char scratch[SS_FFT_NOALLOC_MAX_BYTES];

namespace {
constexpr int NOALLOC_MAX_DEVICES = 64;
std::atomic<void*> noalloc_ptr[NOALLOC_MAX_DEVICES];

// Device address of `scratch` on `device` (which must be current), looked up
// on the first call for that device.  Two threads that race here store the
// same address.
int noalloc_scratch(int device, void** out) {
    if (device < 0 || device >= NOALLOC_MAX_DEVICES)
        return ss::fail(SS_ERR_UNSUPPORTED, "device ordinal beyond the staging-buffer table");
    void* ptr = noalloc_ptr[device].load(std::memory_order_acquire);
    if (!ptr) {
        const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !ptr)
            return ss::fail(SS_ERR_DEVICE, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
        noalloc_ptr[device].store(ptr, std::memory_order_release);
    }
    *out = ptr;
    return SS_OK;
}
}  // namespace

extern "C" int ss_fft_execute_host_noalloc(ss_fft_plan* p, void* host_data) {
    if (!p || !host_data) return ss::fail(SS_ERR_INVALID, "NULL argument");
    const size_t bytes = (size_t)p->fft.nx * p->fft.ny * p->fft.nz * p->batch *
                         (p->fft.precision == SS_FP64 ? 16 : 8);
    if (bytes > (size_t)SS_FFT_NOALLOC_MAX_BYTES)
        return ss::fail(SS_ERR_UNSUPPORTED, "the plan's batch of " + std::to_string(bytes) +
                                                " bytes exceeds the staging buffer of " +
                                                std::to_string((size_t)SS_FFT_NOALLOC_MAX_BYTES) +
                                                " bytes; use ss_fft_execute_host");
    SS_HIP_CHECK(hipSetDevice(p->device));
    void* d = nullptr;
    int rc = noalloc_scratch(p->device, &d);
    if (rc != SS_OK) return rc;
    // the copy in overwrites every staged byte, so nothing is cleared first
    if (hipMemcpyAsync(d, host_data, bytes, hipMemcpyHostToDevice, p->stream) != hipSuccess)
        rc = ss::fail(SS_ERR_DEVICE, "copy in failed");
    if (rc == SS_OK) rc = p->fft.run(d, p->batch, p->direction, p->stream);
    if (rc == SS_OK && hipMemcpyAsync(host_data, d, bytes, hipMemcpyDeviceToHost, p->stream) != hipSuccess)
        rc = ss::fail(SS_ERR_DEVICE, "copy out failed");
    if (hipStreamSynchronize(p->stream) != hipSuccess && rc == SS_OK)
        rc = ss::fail(SS_ERR_DEVICE, "FFT execution failed");
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
