// mathcheck.hip — test library: compares the engine's correctly rounding float
// routines (exact_math.h) with the device library's double routines rounded to
// float, on every one of the 2^32 float bit patterns.  Test infrastructure only.
#include <hip/hip_runtime.h>

#include <cstdint>

#include "../../ppi-auto_amd/csrc/exact_math.h"

using namespace ppi::dev;

namespace {

constexpr int kSamples = 64;

template <int FN> __device__ __forceinline__ float fast(float x) {
    if constexpr (FN == 0) return exact_exp(x);
    else if constexpr (FN == 1) return exact_log(x);
    else if constexpr (FN == 2) return exact_sin(x);
    else if constexpr (FN == 3) return exact_cos(x);
    else return exact_sqrt(x);
}
template <int FN> __device__ __forceinline__ float ref(float x) {
    if constexpr (FN == 0) return exp_of_double(x);
    else if constexpr (FN == 1) return log_of_double(x);
    else if constexpr (FN == 2) return sin_of_double(x);
    else if constexpr (FN == 3) return cos_of_double(x);
    else return sqrtf(x);
}

// Grid-stride sweep over bit patterns [start, start + count); every thread exits
// when its index passes the end.
template <int FN>
__global__ void sweep(uint64_t start, uint64_t count, unsigned long long* bad, uint32_t* nsamp, uint32_t* samples) {
    stage_exact_tables();
    __syncthreads();
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    unsigned long long nbad = 0;
    for (uint64_t i = start + (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < start + count; i += stride) {
        const float x = __uint_as_float((uint32_t)i);
        const float a = fast<FN>(x), b = ref<FN>(x);
        const bool same = __float_as_uint(a) == __float_as_uint(b) || (a != a && b != b);
        if (!same) {
            ++nbad;
            const uint32_t s = atomicAdd(nsamp, 1u);
            if (s < (uint32_t)kSamples) samples[s] = (uint32_t)i;
        }
    }
    if (nbad) atomicAdd(bad, nbad);
}

}  // namespace

// fn: 0 exp, 1 log, 2 sin, 3 cos, 4 sqrt (against sqrtf).  Sweeps all 2^32 inputs; returns 0 on success
// (mismatch count in *mismatches, up to 64 mismatching inputs in samples[]), or
// a HIP error code.
extern "C" int ppi_mathcheck(int fn, unsigned long long* mismatches, uint32_t* samples, uint32_t* nsamples,
                             float* ms) {
    unsigned long long* d_bad = nullptr;
    uint32_t* d_ns = nullptr;
    uint32_t* d_s = nullptr;
    hipError_t e = hipMalloc(&d_bad, sizeof(unsigned long long));
    if (!e) e = hipMalloc(&d_ns, sizeof(uint32_t));
    if (!e) e = hipMalloc(&d_s, kSamples * sizeof(uint32_t));
    if (!e) e = hipMemset(d_bad, 0, sizeof(unsigned long long));
    if (!e) e = hipMemset(d_ns, 0, sizeof(uint32_t));
    hipEvent_t t0 = nullptr, t1 = nullptr;
    if (!e) e = hipEventCreate(&t0);
    if (!e) e = hipEventCreate(&t1);
    if (!e) {
        const dim3 grid(2048), block(256);
        const uint64_t total = 1ull << 32;
        (void)hipEventRecord(t0, 0);
        switch (fn) {
        case 0: hipLaunchKernelGGL(sweep<0>, grid, block, 0, 0, 0ull, total, d_bad, d_ns, d_s); break;
        case 1: hipLaunchKernelGGL(sweep<1>, grid, block, 0, 0, 0ull, total, d_bad, d_ns, d_s); break;
        case 2: hipLaunchKernelGGL(sweep<2>, grid, block, 0, 0, 0ull, total, d_bad, d_ns, d_s); break;
        case 3: hipLaunchKernelGGL(sweep<3>, grid, block, 0, 0, 0ull, total, d_bad, d_ns, d_s); break;
        case 4: hipLaunchKernelGGL(sweep<4>, grid, block, 0, 0, 0ull, total, d_bad, d_ns, d_s); break;
        default: e = hipErrorInvalidValue;
        }
        if (!e) e = hipGetLastError();
        (void)hipEventRecord(t1, 0);
        if (!e) e = hipEventSynchronize(t1);
        if (!e) e = hipEventElapsedTime(ms, t0, t1);
    }
    if (!e) e = hipMemcpy(mismatches, d_bad, sizeof(unsigned long long), hipMemcpyDeviceToHost);
    if (!e) e = hipMemcpy(nsamples, d_ns, sizeof(uint32_t), hipMemcpyDeviceToHost);
    if (!e) e = hipMemcpy(samples, d_s, kSamples * sizeof(uint32_t), hipMemcpyDeviceToHost);
    if (t0) (void)hipEventDestroy(t0);
    if (t1) (void)hipEventDestroy(t1);
    (void)hipFree(d_bad);
    (void)hipFree(d_ns);
    (void)hipFree(d_s);
    return (int)e;
}
