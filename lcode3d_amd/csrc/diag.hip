// On-device plasma-noise diagnostic zn (lcode.py:980-989): the mean of
// |ro - G_sigma * ro| over the grid, divided by 4.23045376e-04, where
// G_sigma is a separable Gaussian blur of sigma = 0.25 / grid_step_size
// cells with scipy.ndimage.gaussian_filter's conventions (taps
// exp(-k^2 / (2 sigma^2)) normalised to sum 1, |k| <= int(4 sigma + 0.5),
// half-sample 'reflect' boundaries; axis 0 first, then axis 1).  Two
// kernels: the axis-0 blur into a scratch grid, then the axis-1 blur fused
// with |ro - blurred| and a block sum; one FP64 atomic per wave.

#include "common.hpp"

namespace lcode {

namespace {

constexpr int kDiagBlock = 256;

// Index of the half-sample symmetric extension of [0, n).
__device__ __forceinline__ int reflect_index(int i, int n) {
    const int p = 2 * n;
    int m = i % p;
    if (m < 0) m += p;
    return m < n ? m : p - 1 - m;
}

__global__ void __launch_bounds__(kDiagBlock)
blur_axis0_kernel(const double* ro, const double* w, int r, int N, double* out) {
    const int e = blockIdx.x * kDiagBlock + threadIdx.x;
    if (e >= N * N) return;
    const int i = e / N, j = e - i * N;
    double s = 0.;
    for (int k = -r; k <= r; ++k)
        s += w[k + r] * ro[(size_t)reflect_index(i + k, N) * N + j];
    out[e] = s;
}

__global__ void __launch_bounds__(kDiagBlock)
blur_axis1_absdiff_kernel(const double* ro, const double* tmp, const double* w,
                          int r, int N, double* sum) {
    const int e = blockIdx.x * kDiagBlock + threadIdx.x;
    double v = 0.;
    if (e < N * N) {
        const int i = e / N, j = e - i * N;
        const double* row = tmp + (size_t)i * N;
        double s = 0.;
        for (int k = -r; k <= r; ++k) s += w[k + r] * row[reflect_index(j + k, N)];
        v = fabs(ro[e] - s);
    }
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
    if ((threadIdx.x & 63) == 0) atomicAdd(sum, v);
}

__global__ void zero_one(double* p) { *p = 0.; }

__global__ void zn_finish(const double* sum, int N, double* out) {
    *out = *sum / ((double)N * N) / 4.23045376e-04;   // lcode.py:987
}

}  // namespace

int launch_diag_zn(const double* ro, const double* w, int r, int N, double* tmp,
                   double* sum, double* out, hipStream_t s, Profiler* prof) {
    const int blocks = (N * N + kDiagBlock - 1) / kDiagBlock;
    {
        ProfScope ps(prof, "diag_zn", s);
        zero_one<<<1, 1, 0, s>>>(sum);
        blur_axis0_kernel<<<blocks, kDiagBlock, 0, s>>>(ro, w, r, N, tmp);
        blur_axis1_absdiff_kernel<<<blocks, kDiagBlock, 0, s>>>(ro, tmp, w, r, N, sum);
        zn_finish<<<1, 1, 0, s>>>(sum, N, out);
    }
    LCODE_HIP(hipGetLastError());
    return LCODE_OK;
}

}  // namespace lcode
