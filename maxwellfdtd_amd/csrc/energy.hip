// Total electromagnetic field energy of a model: W_E = 1/2 sum eps0 eps |E|^2 dV and
// W_H = 1/2 sum mu0 mu |H|^2 dV over every cell, from the stored E^n and H^(n-1/2).
// The reference computes no energy; this is the result check that the field snapshots
// are judged by beside their relative L2 error.  eps0*eps and mu0*mu of a cell come
// from its material's update coefficients: eps0*eps = 1/Me, mu0*mu = dt/Mh
// (src/materials.jl:393-400; vacuum src/models.jl:111-114).  dV is dx*dy*dz in 3D,
// dx*dz in 2D (per metre along y) and dz in 1D (per square metre).
//
// Two launches: each workgroup reduces a grid-stride share of the cells into one FP64
// pair (W_E, W_H) in a partial-sum buffer, then one workgroup sums the partials in a
// fixed order.  The result depends only on the cell count, not on the stream or call.
#include <hip/hip_runtime.h>

#include <mutex>
#include <string>

#include "engine.hpp"

namespace mfd {

constexpr int ENERGY_THREADS = 256;
constexpr int ENERGY_MAX_BLOCKS = 2048;   // partial sums per call: 2048 x 16 B = 32 KiB

struct EnergyArgs {
    const float *E[3];
    const float *H[3];
    const double *E64[3];     // an FP64 model's fields (E, H null)
    const double *H64[3];
    const unsigned char *id;
    const DevMat *mats;
    const DevMat64 *mats64;   // an FP64 model's coefficients (else null)
    long long N;
    int nE, nH;           // components present (3D 3/3, 2D E 2 / H 1, 1D 1/1)
    double dt, dV;
};

__device__ inline double2 block_sum2(double2 v, double2 *sh) {
    const int t = threadIdx.x;
    sh[t] = v;
    __syncthreads();
    for (int o = blockDim.x / 2; o > 0; o >>= 1) {
        if (t < o) { sh[t].x += sh[t + o].x; sh[t].y += sh[t + o].y; }
        __syncthreads();
    }
    return sh[0];
}

__global__ __launch_bounds__(ENERGY_THREADS) void k_energy_partial(EnergyArgs a, double2 *__restrict__ part) {
    __shared__ double2 sh[ENERGY_THREADS];
    double we = 0.0, wh = 0.0;
    const long long stride = (long long)gridDim.x * blockDim.x;
    for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < a.N; i += stride) {
        const int g = a.id[i] & ID_MASK;
        double e2 = 0.0, h2 = 0.0;
        for (int c = 0; c < a.nE; c++) { const double v = a.E[c] ? a.E[c][i] : a.E64[c][i]; e2 += v * v; }
        for (int c = 0; c < a.nH; c++) { const double v = a.H[c] ? a.H[c][i] : a.H64[c][i]; h2 += v * v; }
        if (a.mats64) {
            we += e2 / a.mats64[g].Me;
            wh += a.dt / a.mats64[g].Mh * h2;
        } else {
            we += (double)a.mats[g].iMe * e2;
            wh += a.dt / (double)a.mats[g].Mh * h2;
        }
    }
    const double2 s = block_sum2(make_double2(we, wh), sh);
    if (threadIdx.x == 0) part[blockIdx.x] = s;
}

__global__ __launch_bounds__(ENERGY_THREADS) void k_energy_final(const double2 *__restrict__ part, int nb,
                                                                 double scale, double *__restrict__ out) {
    __shared__ double2 sh[ENERGY_THREADS];
    double2 v = make_double2(0.0, 0.0);
    for (int b = threadIdx.x; b < nb; b += blockDim.x) { v.x += part[b].x; v.y += part[b].y; }
    const double2 s = block_sum2(v, sh);
    if (threadIdx.x == 0) { out[0] = s.x * scale; out[1] = s.y * scale; }
}

// Launch shape of one call: the number of workgroups (partial sums) for N cells.
int energy_blocks(long long N) {
    long long nb = (N + ENERGY_THREADS * 8 - 1) / (ENERGY_THREADS * 8);
    if (nb < 1) nb = 1;
    if (nb > ENERGY_MAX_BLOCKS) nb = ENERGY_MAX_BLOCKS;
    return (int)nb;
}

static EnergyArgs energy_args(const Model &m) {
    EnergyArgs a{};
    const DevModel &d = m.dm;
    if (m.d.ndims == 3) {
        a.nE = a.nH = 3;
        for (int c = 0; c < 3; c++) { a.E[c] = d.F[EX + c]; a.H[c] = d.F[HX + c]; }
        a.dV = m.d.dx * m.d.dy * m.d.dz;
    } else if (m.d.ndims == 2) {
        a.nE = 2; a.nH = 1;
        a.E[0] = d.F[EX]; a.E[1] = d.F[EZ]; a.H[0] = d.F[HY];
        a.dV = m.d.dx * m.d.dz;
    } else {
        a.nE = a.nH = 1;
        a.E[0] = d.F[EX]; a.H[0] = d.F[HY];
        a.E64[0] = d.F64[EX]; a.H64[0] = d.F64[HY];
        a.dV = m.d.dz;
    }
    a.id = d.id;
    a.mats = d.mats;
    a.mats64 = d.f64 ? d.mats64 : nullptr;
    a.N = m.N;
    a.dt = m.d.dt;
    return a;
}

// The two launches of one energy evaluation, partial sums in `part` (energy_blocks(N)
// entries), result (W_E, W_H) in out[0..1] (device memory), on stream s.
void launch_energy(const Model &m, double2 *part, double *out, hipStream_t s) {
    const EnergyArgs a = energy_args(m);
    const int nb = energy_blocks(a.N);
    hipLaunchKernelGGL(k_energy_partial, dim3(nb), dim3(ENERGY_THREADS), 0, s, a, part);
    hipLaunchKernelGGL(k_energy_final, dim3(1), dim3(ENERGY_THREADS), 0, s, (const double2 *)part, nb,
                       0.5 * a.dV, out);
}

// ---- allocation-free variant ----------------------------------------------------------
// Partial sums of mfd_model_get_energy_noalloc: one file-scope device array, of which
// every GPU the library runs on holds its own copy until the process ends.  Each call's
// first launch writes every slot its second launch reads, so nothing has to start at
// zero; calls that overlap on one GPU (two streams, two threads) share the array and
// race.  In host code the array's name is only a stand-in for the symbol: its device
// address, per device, comes from hipGetSymbolAddress and is what kernels receive.
static __device__ __align__(256) double2 scratch[ENERGY_MAX_BLOCKS];

constexpr int SCRATCH_MAX_DEVICES = 64;
static std::mutex g_scratch_mu;
static double2 *g_scratch[SCRATCH_MAX_DEVICES];   // device address of `scratch`, per device

// The device address of this device's `scratch`, looked up once per device (with that
// device current) on the first call there.
static int scratch_address(double2 **out) {
    int dev = 0;
    MFD_HIP(hipGetDevice(&dev));
    if (dev < 0 || dev >= SCRATCH_MAX_DEVICES) {
        set_error("energy scratch: device index beyond the 64 the library keeps addresses for");
        return MFD_ERR_ARG;
    }
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (!g_scratch[dev]) {
        void *p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p) {
            set_error(std::string("energy scratch: hipGetSymbolAddress: ") + hipGetErrorString(e));
            return MFD_ERR_HIP;
        }
        g_scratch[dev] = (double2 *)p;
    }
    *out = g_scratch[dev];
    return MFD_OK;
}

}  // namespace mfd

using namespace mfd;

extern "C" {

int32_t mfd_model_get_energy(mfd_model *h, double *out) {
    if (!h || !out) { set_error("null argument"); return MFD_ERR_ARG; }
    Model &m = h->m;
    if (!m.uploaded) { out[0] = out[1] = 0.0; return MFD_OK; }
    const int nb = energy_blocks(m.N);
    double2 *buf = nullptr;     // nb partial sums, then the result pair
    MFD_HIP(hipMalloc(&buf, (nb + 1) * sizeof(double2)));
    launch_energy(m, buf, (double *)(buf + nb), m.io);
    hipError_t e = hipGetLastError();
    double res[2] = {0.0, 0.0};
    if (e == hipSuccess) e = copy_sync(res, buf + nb, sizeof(res), hipMemcpyDeviceToHost, m.io);
    (void)hipFree(buf);
    if (e != hipSuccess) { set_error(std::string("energy: ") + hipGetErrorString(e)); return MFD_ERR_HIP; }
    out[0] = res[0];
    out[1] = res[1];
    return MFD_OK;
}

// Same sums as mfd_model_get_energy with no allocation: the partial sums go to the
// file-scope `scratch` array, the result (W_E, W_H) to the caller's device memory
// d_out[0..1], asynchronously on `stream` (null: the model's copy stream).
int32_t mfd_model_get_energy_noalloc(mfd_model *h, double *d_out, void *stream) {
    if (!h || !d_out) { set_error("null argument"); return MFD_ERR_ARG; }
    Model &m = h->m;
    if (!m.uploaded) {
        set_error("energy without allocation needs an uploaded model (mfd_model_upload or a step)");
        return MFD_ERR_STATE;
    }
    const int nb = energy_blocks(m.N);
    if (nb > ENERGY_MAX_BLOCKS) {
        set_error("energy without allocation: " + std::to_string(nb) + " partial sums exceed the " +
                  std::to_string(ENERGY_MAX_BLOCKS) + " of the device scratch array");
        return MFD_ERR_ARG;
    }
    double2 *part = nullptr;
    const int st = scratch_address(&part);
    if (st != MFD_OK) return st;
    hipStream_t s = stream ? (hipStream_t)stream : m.io;
    launch_energy(m, part, d_out, s);
    MFD_HIP(hipGetLastError());
    return MFD_OK;
}

}  // extern "C"
