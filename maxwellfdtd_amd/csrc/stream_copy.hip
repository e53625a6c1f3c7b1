// Plain device copy used as the HBM reference: achievable copy bandwidth on this GPU
// and the scale of the FETCH_SIZE / WRITE_SIZE counters for a known byte count.
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <vector>

#include "engine.hpp"

namespace mfd {

__global__ __launch_bounds__(256) void k_copy(const float4 *__restrict__ src,
                                              float4 *__restrict__ dst, long long n4) {
    const long long stride = (long long)gridDim.x * blockDim.x;
    for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n4; i += stride)
        dst[i] = src[i];
}

// 4 float4 per lane in flight over block-contiguous chunks, nontemporal both ways
__global__ __launch_bounds__(256) void k_copy_fast(const float4 *__restrict__ src,
                                                   float4 *__restrict__ dst, long long n4) {
    typedef float v4 __attribute__((ext_vector_type(4)));
    typedef __attribute__((address_space(1))) v4 gv4;
    const long long chunk = 4LL * blockDim.x;
    for (long long base = blockIdx.x * chunk; base < n4; base += (long long)gridDim.x * chunk) {
        v4 r[4];
#pragma unroll
        for (int u = 0; u < 4; u++) {
            const long long i = base + u * blockDim.x + threadIdx.x;
            if (i < n4) r[u] = __builtin_nontemporal_load(&((const gv4 *)src)[i]);
        }
#pragma unroll
        for (int u = 0; u < 4; u++) {
            const long long i = base + u * blockDim.x + threadIdx.x;
            if (i < n4) __builtin_nontemporal_store(r[u], &((gv4 *)dst)[i]);
        }
    }
}

static int32_t time_copy(int64_t bytes, int32_t iters, double *ms_per_iter, bool fast);

}  // namespace mfd

extern "C" int32_t mfd_stream_copy(int64_t bytes, int32_t iters, double *ms_per_iter) {
    return mfd::time_copy(bytes, iters, ms_per_iter, false);
}

extern "C" int32_t mfd_stream_copy_fast(int64_t bytes, int32_t iters, double *ms_per_iter) {
    return mfd::time_copy(bytes, iters, ms_per_iter, true);
}

static int32_t mfd::time_copy(int64_t bytes, int32_t iters, double *ms_per_iter, bool fast) {
    using namespace mfd;
    if (bytes < 1024 || iters < 1 || !ms_per_iter) { set_error("bad arguments"); return MFD_ERR_ARG; }
    const long long n4 = bytes / 16;
    float4 *a = nullptr, *b = nullptr;
    MFD_HIP(hipMalloc(&a, n4 * 16));
    if (hipMalloc(&b, n4 * 16) != hipSuccess) { hipFree(a); set_error("out of memory"); return MFD_ERR_NOMEM; }
    hipMemset(a, 0, n4 * 16);
    hipMemset(b, 0, n4 * 16);
    int dev = 0, cus = 256;
    hipGetDevice(&dev);
    hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev);
    const int grid = cus * (fast ? 16 : 8);
    auto launch = [&] {
        if (fast) hipLaunchKernelGGL(k_copy_fast, dim3(grid), dim3(256), 0, 0, a, b, n4);
        else hipLaunchKernelGGL(k_copy, dim3(grid), dim3(256), 0, 0, a, b, n4);
    };
    hipEvent_t e0, e1;
    hipEventCreate(&e0);
    hipEventCreate(&e1);
    launch();
    hipEventRecord(e0, 0);
    for (int i = 0; i < iters; i++) launch();
    hipEventRecord(e1, 0);
    hipEventSynchronize(e1);
    float ms = 0;
    hipEventElapsedTime(&ms, e0, e1);
    *ms_per_iter = ms / iters;
    hipEventDestroy(e0);
    hipEventDestroy(e1);
    hipFree(a);
    hipFree(b);
    return MFD_OK;
}

// Multi-stream variant: NIN arrays read, NOUT arrays written per element, as the
// half-step kernels do; used to measure how HBM throughput depends on the number of
// concurrent streams.
namespace mfd {
template <int NIN, int NOUT>
__global__ __launch_bounds__(256) void k_multi(float4 *const *__restrict__ in,
                                               float4 *const *__restrict__ out, long long n4) {
    const long long stride = (long long)gridDim.x * blockDim.x;
    const float4 *ip[NIN];
    float4 *op[NOUT];
#pragma unroll
    for (int k = 0; k < NIN; k++) ip[k] = in[k];
#pragma unroll
    for (int k = 0; k < NOUT; k++) op[k] = out[k];
    for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n4; i += stride) {
        typedef float v4 __attribute__((ext_vector_type(4)));
        v4 s = {0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
        for (int k = 0; k < NIN; k++) s += ((const __attribute__((address_space(1))) v4 *)ip[k])[i];
#pragma unroll
        for (int k = 0; k < NOUT; k++) ((__attribute__((address_space(1))) v4 *)op[k])[i] = s;
    }
}
}  // namespace mfd

extern "C" int32_t mfd_stream_multi(int64_t bytes_per_array, int32_t nin, int32_t nout,
                                    int32_t blocks_per_cu, int32_t iters, double *ms_per_iter) {
    using namespace mfd;
    if (nin < 1 || nin > 6 || nout < 1 || nout > 6 || nin != nout || iters < 1 ||
        blocks_per_cu < 1 || bytes_per_array < 1024) {
        set_error("need nin == nout in 1..6, blocks_per_cu >= 1, bytes >= 1024");
        return MFD_ERR_ARG;
    }
    const long long n4 = bytes_per_array / 16;
    std::vector<float4 *> bufs(nin + nout);
    for (auto &b : bufs) {
        if (hipMalloc(&b, n4 * 16) != hipSuccess) { set_error("oom"); return MFD_ERR_NOMEM; }
        hipMemset(b, 0, n4 * 16);
    }
    float4 **d_in, **d_out;
    MFD_HIP(hipMalloc(&d_in, nin * sizeof(float4 *)));
    MFD_HIP(hipMalloc(&d_out, nout * sizeof(float4 *)));
    MFD_HIP(hipMemcpy(d_in, bufs.data(), nin * sizeof(float4 *), hipMemcpyHostToDevice));
    MFD_HIP(hipMemcpy(d_out, bufs.data() + nin, nout * sizeof(float4 *), hipMemcpyHostToDevice));
    int dev = 0, cus = 256;
    hipGetDevice(&dev);
    hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev);
    const int grid = cus * blocks_per_cu;
    auto launch = [&]() {
        switch (nin) {
            case 1: hipLaunchKernelGGL((k_multi<1, 1>), dim3(grid), dim3(256), 0, 0, d_in, d_out, n4); break;
            case 2: hipLaunchKernelGGL((k_multi<2, 2>), dim3(grid), dim3(256), 0, 0, d_in, d_out, n4); break;
            case 3: hipLaunchKernelGGL((k_multi<3, 3>), dim3(grid), dim3(256), 0, 0, d_in, d_out, n4); break;
            case 4: hipLaunchKernelGGL((k_multi<4, 4>), dim3(grid), dim3(256), 0, 0, d_in, d_out, n4); break;
            case 5: hipLaunchKernelGGL((k_multi<5, 5>), dim3(grid), dim3(256), 0, 0, d_in, d_out, n4); break;
            default: hipLaunchKernelGGL((k_multi<6, 6>), dim3(grid), dim3(256), 0, 0, d_in, d_out, n4); break;
        }
    };
    hipEvent_t e0, e1;
    hipEventCreate(&e0);
    hipEventCreate(&e1);
    launch();
    hipEventRecord(e0, 0);
    for (int i = 0; i < iters; i++) launch();
    hipEventRecord(e1, 0);
    hipEventSynchronize(e1);
    float ms = 0;
    hipEventElapsedTime(&ms, e0, e1);
    *ms_per_iter = ms / iters;
    hipEventDestroy(e0);
    hipEventDestroy(e1);
    for (auto b : bufs) hipFree(b);
    hipFree(d_in);
    hipFree(d_out);
    return MFD_OK;
}
