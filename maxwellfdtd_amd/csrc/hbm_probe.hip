// HBM bandwidth probes, built as their own library (libmfd_hbm_probe.so) beside the
// engine: a plain device copy (the achievable copy bandwidth on this GPU, and the scale
// of the FETCH_SIZE / WRITE_SIZE counters for a known byte count), the fastest copy form
// measured here, and a many-stream variant.  Nothing in the engine calls them.
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <string>
#include <vector>

#include "hbm_probe.h"

namespace mfd_probe {

enum { MFD_OK = 0, MFD_ERR_ARG = -1, MFD_ERR_NOMEM = -2, MFD_ERR_HIP = -3 };

static thread_local std::string g_err;
static void set_error(const std::string &s) { g_err = s; }

#define MFD_HIP(x)                                                           \
    do {                                                                     \
        const hipError_t e_ = (x);                                           \
        if (e_ != hipSuccess) {                                              \
            set_error(std::string(#x) + ": " + hipGetErrorString(e_));       \
            return MFD_ERR_HIP;                                              \
        }                                                                    \
    } while (0)

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

// device buffers and events of one probe, released on every return path
struct ProbeRes {
    std::vector<void *> mem;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    ~ProbeRes() {
        if (e0) (void)hipEventDestroy(e0);
        if (e1) (void)hipEventDestroy(e1);
        for (void *p : mem) (void)hipFree(p);
    }
    template <class T>
    int alloc(T **p, size_t bytes) {
        if (hipMalloc(p, bytes) != hipSuccess) { set_error("out of memory"); return MFD_ERR_NOMEM; }
        mem.push_back(*p);
        return MFD_OK;
    }
};

#define PROBE_TRY(x)                                     \
    do {                                                 \
        const int st_ = (x);                             \
        if (st_ != MFD_OK) return st_;                   \
    } while (0)

// one warm-up launch, then `iters` timed launches between two events on the null stream
template <class F>
static int32_t time_launches(ProbeRes &r, int32_t iters, F launch, double *ms_per_iter) {
    MFD_HIP(hipEventCreate(&r.e0));
    MFD_HIP(hipEventCreate(&r.e1));
    launch();
    MFD_HIP(hipGetLastError());
    MFD_HIP(hipEventRecord(r.e0, 0));
    for (int i = 0; i < iters; i++) launch();
    MFD_HIP(hipEventRecord(r.e1, 0));
    MFD_HIP(hipEventSynchronize(r.e1));
    float ms = 0;
    MFD_HIP(hipEventElapsedTime(&ms, r.e0, r.e1));
    *ms_per_iter = ms / iters;
    return MFD_OK;
}

static int32_t cu_count_of_device(int *cus) {
    int dev = 0;
    MFD_HIP(hipGetDevice(&dev));
    MFD_HIP(hipDeviceGetAttribute(cus, hipDeviceAttributeMultiprocessorCount, dev));
    return MFD_OK;
}

static int32_t time_copy(int64_t bytes, int32_t iters, double *ms_per_iter, bool fast) {
    if (bytes < 1024 || iters < 1 || !ms_per_iter) { set_error("bad arguments"); return MFD_ERR_ARG; }
    const long long n4 = bytes / 16;
    ProbeRes r;
    float4 *a = nullptr, *b = nullptr;
    PROBE_TRY(r.alloc(&a, n4 * 16));
    PROBE_TRY(r.alloc(&b, n4 * 16));
    MFD_HIP(hipMemset(a, 0, n4 * 16));
    MFD_HIP(hipMemset(b, 0, n4 * 16));
    int cus = 0;
    PROBE_TRY(cu_count_of_device(&cus));
    const int grid = cus * (fast ? 16 : 8);
    return time_launches(r, iters, [&] {
        if (fast) hipLaunchKernelGGL(k_copy_fast, dim3(grid), dim3(256), 0, 0, a, b, n4);
        else hipLaunchKernelGGL(k_copy, dim3(grid), dim3(256), 0, 0, a, b, n4);
    }, ms_per_iter);
}

// Multi-stream variant: NIN arrays read, NOUT arrays written per element, as the
// half-step kernels do; used to measure how HBM throughput depends on the number of
// concurrent streams.
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
static int32_t time_multi(int64_t bytes_per_array, int32_t nin, int32_t nout, int32_t blocks_per_cu,
                          int32_t iters, double *ms_per_iter) {
    if (nin < 1 || nin > 6 || nout < 1 || nout > 6 || nin != nout || iters < 1 ||
        blocks_per_cu < 1 || bytes_per_array < 1024) {
        set_error("need nin == nout in 1..6, blocks_per_cu >= 1, bytes >= 1024");
        return MFD_ERR_ARG;
    }
    if (!ms_per_iter) { set_error("bad arguments"); return MFD_ERR_ARG; }
    const long long n4 = bytes_per_array / 16;
    ProbeRes r;
    std::vector<float4 *> bufs(nin + nout);
    for (auto &b : bufs) {
        PROBE_TRY(r.alloc(&b, n4 * 16));
        MFD_HIP(hipMemset(b, 0, n4 * 16));
    }
    float4 **d_in, **d_out;
    PROBE_TRY(r.alloc(&d_in, nin * sizeof(float4 *)));
    PROBE_TRY(r.alloc(&d_out, nout * sizeof(float4 *)));
    MFD_HIP(hipMemcpy(d_in, bufs.data(), nin * sizeof(float4 *), hipMemcpyHostToDevice));
    MFD_HIP(hipMemcpy(d_out, bufs.data() + nin, nout * sizeof(float4 *), hipMemcpyHostToDevice));
    int cus = 0;
    PROBE_TRY(cu_count_of_device(&cus));
    const int grid = cus * blocks_per_cu;
    return time_launches(r, iters, [&]() {
        switch (nin) {
            case 1: hipLaunchKernelGGL((k_multi<1, 1>), dim3(grid), dim3(256), 0, 0, d_in, d_out, n4); break;
            case 2: hipLaunchKernelGGL((k_multi<2, 2>), dim3(grid), dim3(256), 0, 0, d_in, d_out, n4); break;
            case 3: hipLaunchKernelGGL((k_multi<3, 3>), dim3(grid), dim3(256), 0, 0, d_in, d_out, n4); break;
            case 4: hipLaunchKernelGGL((k_multi<4, 4>), dim3(grid), dim3(256), 0, 0, d_in, d_out, n4); break;
            case 5: hipLaunchKernelGGL((k_multi<5, 5>), dim3(grid), dim3(256), 0, 0, d_in, d_out, n4); break;
            default: hipLaunchKernelGGL((k_multi<6, 6>), dim3(grid), dim3(256), 0, 0, d_in, d_out, n4); break;
        }
    }, ms_per_iter);
}

}  // namespace mfd_probe

extern "C" {

const char *mfd_probe_last_error(void) { return mfd_probe::g_err.c_str(); }

int32_t mfd_probe_copy(int64_t bytes, int32_t iters, double *ms_per_iter) {
    return mfd_probe::time_copy(bytes, iters, ms_per_iter, false);
}

int32_t mfd_probe_copy_fast(int64_t bytes, int32_t iters, double *ms_per_iter) {
    return mfd_probe::time_copy(bytes, iters, ms_per_iter, true);
}

int32_t mfd_probe_multi(int64_t bytes_per_array, int32_t nin, int32_t nout, int32_t blocks_per_cu,
                        int32_t iters, double *ms_per_iter) {
    return mfd_probe::time_multi(bytes_per_array, nin, nout, blocks_per_cu, iters, ms_per_iter);
}

}  // extern "C"
