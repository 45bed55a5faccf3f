// vs_dft.hip -- FP64 discrete-time Fourier transform of probe waveforms at
// arbitrary frequencies (the spectrum behind the transfer function and the
// sub-bin formant refinement).
//
// X(f) = sum_n x[n] w[n] exp(-i 2 pi f n / fs).  One work-group per
// (frequency, signal): each lane starts its phasor from an exact sincos at its
// first sample and advances it by a rotation of 256 samples per stride, so the
// accumulated phasor error stays at ~(n / 256) ulp; partial sums are reduced in
// double through LDS.

#include "vs_common.h"

#include <mutex>

namespace vs {
namespace {

constexpr int kDftThreads = 256;

__global__ void __launch_bounds__(kDftThreads)
k_dtft(const double* __restrict__ x, long long stride_sig, int n, double inv_fs, int hann,
       const double* __restrict__ freqs, int nf, double* __restrict__ re, double* __restrict__ im) {
    const int fi = blockIdx.x, sig = blockIdx.y;
    const double* xs = x + (long long)sig * stride_sig;
    const double w = -2.0 * M_PI * freqs[fi] * inv_fs;     // radians per sample
    double s0, c0, sB, cB;
    sincos(w * threadIdx.x, &s0, &c0);
    sincos(w * kDftThreads, &sB, &cB);
    double ar = 0.0, ai = 0.0, pr = c0, pi = s0;
    const double hw = 2.0 * M_PI / (double)(n > 1 ? n - 1 : 1);
    for (int k = threadIdx.x; k < n; k += kDftThreads) {
        double v = xs[k];
        if (hann) v *= 0.5 - 0.5 * cos(hw * k);
        ar += v * pr;
        ai += v * pi;
        const double nr = pr * cB - pi * sB;
        pi = pr * sB + pi * cB;
        pr = nr;
    }
    __shared__ double sr[kDftThreads], si[kDftThreads];
    sr[threadIdx.x] = ar;
    si[threadIdx.x] = ai;
    __syncthreads();
    for (int s = kDftThreads / 2; s > 0; s >>= 1) {
        if (threadIdx.x < s) {
            sr[threadIdx.x] += sr[threadIdx.x + s];
            si[threadIdx.x] += si[threadIdx.x + s];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        re[(long long)sig * nf + fi] = sr[0];
        im[(long long)sig * nf + fi] = si[0];
    }
}

}  // namespace

hipError_t launch_dtft(const double* d_x, long long stride, int nsig, int n, double fs, int hann,
                       const double* d_f, int nf, double* d_re, double* d_im, hipStream_t st) {
    hipLaunchKernelGGL(k_dtft, dim3(nf, nsig), dim3(kDftThreads), 0, st, d_x, stride, n, 1.0 / fs, hann,
                       d_f, nf, d_re, d_im);
    return hipGetLastError();
}

// ---------------------------------------------------------------------------
// Allocation-free transform: the device copies of the samples, frequencies
// and results live in `scratch`, a file-scope device array (one per GPU,
// held until the process ends), instead of buffers allocated per call.
// Layout in doubles: samples [0, kNoAllocMaxSamples), frequencies, real
// parts and imaginary parts kNoAllocMaxFreqs each after it.  The host never
// uses the array's name as an address: its device address comes from
// hipGetSymbolAddress, once per device with that device current, and is kept
// per device.  Calls on one GPU share the array; dtft_noalloc holds a
// per-device lock for the whole call, and calls from other processes or
// through other copies of the library on the same GPU would race.  Nothing in
// the array has to start at zero: every call overwrites the parts it reads.
// ---------------------------------------------------------------------------
namespace {
constexpr size_t kScratchDoubles = (size_t)kNoAllocMaxSamples + 3 * (size_t)kNoAllocMaxFreqs;
__device__ __align__(256) double scratch[kScratchDoubles];

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu[kMaxDevices];
double* g_scratch_ptr[kMaxDevices] = {};
}  // namespace

hipError_t dtft_noalloc(const double* x, int n, double fs, int hann, const double* freqs, int nf, double* re,
                        double* im) {
    if (n < 1 || n > kNoAllocMaxSamples || nf < 1 || nf > kNoAllocMaxFreqs) return hipErrorInvalidValue;
    int dev = 0;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess) return e;
    if (dev < 0 || dev >= kMaxDevices) return hipErrorInvalidDevice;
    std::lock_guard<std::mutex> lock(g_scratch_mu[dev]);
    if (!g_scratch_ptr[dev]) {
        void* p = nullptr;
        e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess) return e;
        g_scratch_ptr[dev] = static_cast<double*>(p);
    }
    double* const base = g_scratch_ptr[dev];
    double* const d_x = base;
    double* const d_f = base + kNoAllocMaxSamples;
    double* const d_re = d_f + kNoAllocMaxFreqs;
    double* const d_im = d_re + kNoAllocMaxFreqs;
    hipStream_t st = nullptr;
    e = hipMemcpyAsync(d_x, x, sizeof(double) * n, hipMemcpyHostToDevice, st);
    if (e == hipSuccess) e = hipMemcpyAsync(d_f, freqs, sizeof(double) * nf, hipMemcpyHostToDevice, st);
    if (e == hipSuccess) e = launch_dtft(d_x, 0, 1, n, fs, hann, d_f, nf, d_re, d_im, st);
    if (e == hipSuccess) e = hipMemcpyAsync(re, d_re, sizeof(double) * nf, hipMemcpyDeviceToHost, st);
    if (e == hipSuccess) e = hipMemcpyAsync(im, d_im, sizeof(double) * nf, hipMemcpyDeviceToHost, st);
    if (e == hipSuccess) e = hipStreamSynchronize(st);
    return e;
}

}  // namespace vs
