// ss_internal.h -- host-side declarations shared by the translation units of
// libss (not part of the C ABI).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <string>
#include <vector>

#include "../../include/ss_api.h"
#include "ss_device.h"

namespace ss {

// Thread-local error message behind ss_last_error().
void set_error(const std::string& msg);
int fail(int code, const std::string& msg);

#define SS_HIP_CHECK(expr)                                                             \
    do {                                                                               \
        hipError_t e_ = (expr);                                                        \
        if (e_ != hipSuccess)                                                          \
            return ::ss::fail(SS_ERR_DEVICE, std::string(#expr) + ": " + hipGetErrorString(e_)); \
    } while (0)

// ---------------------------------------------------------------------------
// Batched 3-D c2c FFT over contiguous [nx][ny][nz] complex arrays.
// ---------------------------------------------------------------------------
struct FftAxis {
    int n = 0;
    int nf = 0;
    int fac[16] = {0};
    void* tw = nullptr;  // n complex twiddles exp(-2 pi i t / n) in the plan precision
    // dense-DFT variant: the real 2n x 2n DFT matrix of each sign split into
    // three bf16 parts (high, middle, low), in MFMA fragment order
    // ([sign][part])
    void* dft[2][3] = {{nullptr, nullptr, nullptr}, {nullptr, nullptr, nullptr}};
    int dft_rows = 0;  // 2n rounded up to 32 (output rows)
    int dft_k = 0;     // 2n rounded up to 16 (reduction length)
};

struct Fft3d {
    int nx = 0, ny = 0, nz = 0;
    int precision = SS_FP64;
    int variant = SS_FFT_RADIX;  // SS_FFT_RADIX or SS_FFT_DFT (FP32, axes of at most 64 points)
    FftAxis ax[3];
    int init(int nx, int ny, int nz, int precision, int variant = SS_FFT_RADIX);
    void release();
    // In-place, unnormalised; sign -1 forward, +1 backward.  The forward
    // transform runs the axes z, y, x, the backward one x, y, z.
    int run(void* data, int batch, int sign, hipStream_t stream) const;
    // One axis (0 = x, 1 = y, 2 = z) of the same transform.
    int run_axis(void* data, int batch, int sign, int axis, hipStream_t stream) const;
    // True when every axis factors into radices 2, 3, 4 and 5.
    bool small_radices() const;
};

int factorize(int n, int* fac);

// ---------------------------------------------------------------------------
// SS_NOISE_REFERENCE: the reference's Pcg32 streams decoded on the GPU
// (ss_refrng.hip).
// ---------------------------------------------------------------------------
struct RefNoise;
int refnoise_create(RefNoise** out, uint64_t n, int streams, uint64_t seed);
void refnoise_destroy(RefNoise* r);
int refnoise_streams(const RefNoise* r);
// Enqueue one step's RandomVectors into noise: a record of 8 values (the
// first five used; double or float) at slot inv[id] of each particle id
// (inv null: at slot id).
hipError_t refnoise_enqueue(RefNoise* r, void* noise, bool fp64, double dr, const uint32_t* inv,
                            hipStream_t stream);
int refnoise_get(RefNoise* r, uint64_t* state_increment, int count, hipStream_t stream);
int refnoise_set(RefNoise* r, const uint64_t* state_increment, int count, hipStream_t stream);

}  // namespace ss
