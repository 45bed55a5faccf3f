// vs_audio.hip -- audio export of probe traces (results/generate_audio.py:9-42):
// normalise by max |p|, clip to [-1, 1], and for fs != 44.1 kHz a
// second-order Butterworth low-pass at 5050 Hz followed by decimation by
// srate_mul; float32 samples at 44.1 kHz.
//
// One work-group per tube: the block reduces max |p| over the probe, then one
// lane runs the recursive filter in the transposed direct form II order of
// scipy.signal.lfilter (y = z0 + b0 x; z0 = z1 + x b1 - y a1; z1 = x b2 - y a2),
// every operation rounded separately.

#include "vs_common.h"

#include <cmath>

namespace vs {
namespace {

constexpr int kAudioThreads = 256;

// np.clip semantics: NaN stays NaN
__device__ inline double clip1(double a) { return a < -1.0 ? -1.0 : (a > 1.0 ? 1.0 : a); }

__global__ __launch_bounds__(kAudioThreads) void k_audio(const double* __restrict__ probe,
                                                         const AudioJob* __restrict__ jobs,
                                                         float* __restrict__ out) {
    const AudioJob J = jobs[blockIdx.x];
    const double* x = probe + J.in_off;
    const int n = J.n;
    const int decim = J.decim;
    float* y = out + J.out_off;
    __shared__ double red[kAudioThreads];
    double m = 0.0;
    for (int i = threadIdx.x; i < n; i += blockDim.x) m = fmax(m, fabs(x[i]));
    red[threadIdx.x] = m;
    __syncthreads();
    for (int s = blockDim.x / 2; s > 0; s >>= 1) {
        if (threadIdx.x < s) red[threadIdx.x] = fmax(red[threadIdx.x], red[threadIdx.x + s]);
        __syncthreads();
    }
    const double peak = red[0];
    if (threadIdx.x != 0) return;
    if (decim == 1) {
        for (int i = 0; i < n; ++i) y[i] = (float)clip1(x[i] / peak);
        return;
    }
    const double b0 = J.flt.b[0], b1 = J.flt.b[1], b2 = J.flt.b[2], a1 = J.flt.a[1], a2 = J.flt.a[2];
    double z0 = 0.0, z1 = 0.0;
    int k = 0;
    for (int i = 0; i < n; ++i) {
        const double a = clip1(x[i] / peak);
        const double v = z0 + b0 * a;
        z0 = z1 + a * b1 - v * a1;
        z1 = a * b2 - v * a2;
        if (i % decim == 0) y[k++] = (float)v;
    }
}

}  // namespace

AudioFilter butter2_lowpass(double fc, double fs) {
    // Bilinear transform of the analogue prototype 1 / (s^2 + sqrt(2) s + 1)
    // with the cutoff pre-warped: K = tan(pi fc / fs).
    const double K = std::tan(M_PI * fc / fs);
    const double K2 = K * K;
    const double r2 = std::sqrt(2.0);
    const double den = 1.0 + r2 * K + K2;
    AudioFilter f;
    f.b[0] = K2 / den;
    f.b[1] = 2.0 * K2 / den;
    f.b[2] = K2 / den;
    f.a[0] = 1.0;
    f.a[1] = 2.0 * (K2 - 1.0) / den;
    f.a[2] = (1.0 - r2 * K + K2) / den;
    return f;
}

hipError_t launch_audio(const double* d_probe, const AudioJob* d_jobs, int njobs, float* d_out, hipStream_t st) {
    hipLaunchKernelGGL(k_audio, dim3(njobs), dim3(kAudioThreads), 0, st, d_probe, d_jobs, d_out);
    return hipGetLastError();
}

}  // namespace vs
