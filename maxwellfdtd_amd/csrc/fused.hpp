// Device helpers shared by the fused-step kernels (3D and 2D): byte-offset global
// accesses, lane shifts, CPML slab lookup and the boundary rules of a neighbour outside
// the grid.
#pragma once
#include "common.hpp"

namespace mfd {
namespace fused {

typedef const __attribute__((address_space(1))) char gcC;
typedef __attribute__((address_space(1))) char gC;

// global accesses at a byte offset from a base pointer
template <class T>
__device__ __forceinline__ T ldb(const void *base, unsigned off) {
    return *(const __attribute__((address_space(1))) T *)((gcC *)base + off);
}
template <class T>
__device__ __forceinline__ void stb(void *base, unsigned off, T v) {
    *(__attribute__((address_space(1))) T *)((gC *)base + off) = v;
}
__device__ __forceinline__ float ldnt(const void *base, unsigned off) {
    return __builtin_nontemporal_load((const __attribute__((address_space(1))) float *)((gcC *)base + off));
}
__device__ __forceinline__ void stnt(void *base, unsigned off, float v) {
    __builtin_nontemporal_store(v, (__attribute__((address_space(1))) float *)((gC *)base + off));
}

__device__ __forceinline__ float rdlane(float v, int l) {
    return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), l));
}
// value of lane + 1 / lane - 1 (DPP wave shifts; the edge lane's result is not used)
__device__ __forceinline__ float from_next(float v) {
    return __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), 0x130, 0xf, 0xf, false));
}
__device__ __forceinline__ float from_prev(float v) {
    return __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), 0x138, 0xf, 0xf, false));
}

// profile values of one CPML slab at one position: B, A and 1/K (src/pml.jl:68-81)
struct Prof {
    float B, A, Ki;
};

// CPML slab of axis a at coordinate c of that axis: face index (-1 none) and profile index
__device__ __forceinline__ int slab_of(const DevModel &m, int a, int c, int &p) {
    const DevLayer &L0 = m.L[2 * a], &L1 = m.L[2 * a + 1];
    if (L0.present && c <= L0.ib) { p = c; return 2 * a; }
    if (L1.present && c >= L1.ib) { p = c - L1.ib; return 2 * a + 1; }
    p = 0;
    return -1;
}

__device__ __forceinline__ Prof prof_of(const DevModel &m, int f, int p) {
    Prof r{0.0f, 0.0f, 1.0f};
    if (f >= 0) {
        const DevLayer &L = m.L[f];
        r.B = G(L.B)[p];
        r.A = G(L.A)[p];
        r.Ki = G(L.Ki)[p];
    }
    return r;
}

// boundary rule for a neighbour outside the grid: 0 the loaded (periodic) value, 1 zero
// (Dirichlet), 2 the cell's own value (Neumann), as in src/models.jl:755-802,920-973
__device__ __forceinline__ float bcsel(int mode, float loaded, float own) {
    return mode == 0 ? loaded : mode == 1 ? 0.0f : own;
}
__device__ __forceinline__ int bcmode(int bc) { return bc == BC_PERIODIC ? 0 : bc == BC_DIRICHLET ? 1 : 2; }

}  // namespace fused
}  // namespace mfd
