// vs_measure.hip -- LDS read bandwidth of the device, measured with a
// streaming kernel, the peak against which the on-chip step kernel's LDS
// traffic is rated.
//
// Every work-group (1024 threads) fills a 32 KiB LDS region, then each lane
// reads 8 independent W-word vectors per iteration at lane-consecutive
// addresses (conflict-free: lane l of a read touches words [l W, l W + W)),
// all 8 at distinct addresses, stepping through the region; the words are
// summed so no read is dead.
// The grid holds 8 work-groups per CU so every CU streams for the whole
// launch.  Bytes = groups * 1024 * iters * 8 * 4 W.

#include "vs_common.h"

namespace vs {
namespace {

constexpr int kLdsWords = 8192;          // 32 KiB
constexpr int kLdsThreads = 1024;

template <int W>
__global__ __launch_bounds__(kLdsThreads) void k_lds_read(int iters, float* __restrict__ out) {
    typedef float V __attribute__((ext_vector_type(W)));
    __shared__ __align__(16) float buf[kLdsWords];
    for (int i = threadIdx.x; i < kLdsWords; i += kLdsThreads) buf[i] = (float)(i & 7);
    __syncthreads();
    const V* vb = reinterpret_cast<const V*>(buf);
    constexpr int NV = kLdsWords / W;    // vectors in the region (a power of two)
    V acc[8];
#pragma unroll
    for (int u = 0; u < 8; ++u) acc[u] = V(0.f);
    // the 8 reads of an iteration sit NV / 8 vectors apart, so no two of them
    // (nor of consecutive iterations) read the same address and none can be
    // merged by the compiler
    int base = threadIdx.x;
    for (int it = 0; it < iters; ++it) {
#pragma unroll
        for (int u = 0; u < 8; ++u) acc[u] += vb[(base + u * (NV / 8)) & (NV - 1)];
        base = (base + 64) & (NV - 1);
    }
    float s = 0.f;
#pragma unroll
    for (int u = 0; u < 8; ++u)
#pragma unroll
        for (int w = 0; w < W; ++w) s += acc[u][w];
    out[blockIdx.x * kLdsThreads + threadIdx.x] = s;
}

}  // namespace

// Launches the read kernel of width W (words) on `st`, timed with HIP events.
hipError_t measure_lds_read(int width_words, int groups, int iters, float* d_out, hipStream_t st, float* ms) {
    hipEvent_t a, b;
    hipError_t e = hipEventCreate(&a);
    if (e != hipSuccess) return e;
    e = hipEventCreate(&b);
    if (e != hipSuccess) { (void)hipEventDestroy(a); return e; }
    e = hipEventRecord(a, st);
    if (e == hipSuccess) {
        switch (width_words) {
        case 1: hipLaunchKernelGGL(k_lds_read<1>, dim3(groups), dim3(kLdsThreads), 0, st, iters, d_out); break;
        case 2: hipLaunchKernelGGL(k_lds_read<2>, dim3(groups), dim3(kLdsThreads), 0, st, iters, d_out); break;
        default: hipLaunchKernelGGL(k_lds_read<4>, dim3(groups), dim3(kLdsThreads), 0, st, iters, d_out); break;
        }
        e = hipGetLastError();
    }
    if (e == hipSuccess) e = hipEventRecord(b, st);
    if (e == hipSuccess) e = hipEventSynchronize(b);
    if (e == hipSuccess) e = hipEventElapsedTime(ms, a, b);
    (void)hipEventDestroy(a);
    (void)hipEventDestroy(b);
    return e;
}

}  // namespace vs
