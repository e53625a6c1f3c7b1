// E of the cells that carry ADE polarisation or plasma state.
//
// The E half-step kernels update D everywhere and E everywhere except in cells of
// stateful materials (Debye/Drude/Lorentz poles, plasma); this pass finishes those
// cells: one lane per state slot, so the state arrays (slot-contiguous) are read and
// written coalesced and the main kernels keep the register budget of the simple
// materials.  The per-cell statements are update_E_kernel!'s (src/models.jl:1043-1154,
// 2D :646-713, 1D :381-434) through material_e, then the non-finite guard on Ex and the
// integral variables (src/outputs.jl:182-191,318-327,475-485) of the cell.
#include "common.hpp"
#include "material.hpp"

namespace mfd {

template <int NC, bool INTEG>
__global__ __launch_bounds__(256) void k_e_state(const DevModel *__restrict__ models,
                                                 const int4 *__restrict__ chunks, int nchunks, int pp) {
    const int4 ch = chunks[blockIdx.x];
    const DevModel &m = models[ch.x];
    const int it = *m.it;
    if (it >= m.Nt) return;
    // the material record through the global address space: its scalars are uniform
    // over the block and load as such, with no flat access in the per-lane work
    const __attribute__((address_space(1))) DevMat &mt =
        *(const __attribute__((address_space(1))) DevMat *)(m.mats + ch.y);
    const int s = ch.z + threadIdx.x;
    bool active = s < ch.w;
    float e0 = 0.0f;
    if (active) {
        const int i = G(mt.cells)[s];
        // component order of the 1D/2D/3D fields: D and E of (x), (x, z), (x, y, z)
        constexpr int nd = NC;
        const int cD[3] = {DX, nd == 2 ? DZ : DY, DZ};
        const int cE[3] = {EX, nd == 2 ? EZ : EY, EZ};
        float d[NC], e[NC];
#pragma unroll
        for (int c = 0; c < NC; c++) {
            d[c] = G((const float *)m.F[cD[c]])[i];
            e[c] = G(fld_ptr(m, cE[c], pp))[i];
        }
        material_e<NC>(mt, s, it & 1, d, e, m.dt, m.dtf);
#pragma unroll
        for (int c = 0; c < NC; c++) G((float *)fld_ptr(m, cE[c], pp))[i] = e[c];
        e0 = e[0];
        if (INTEG && !(G(m.id)[i] & ID_TOUCHED)) {
            const float dt = m.dtf;
            if (NC == 3) {
                const float hx = G(fld_ptr(m, HX, pp))[i], hy = G(fld_ptr(m, HY, pp))[i];
                const float hz = G(fld_ptr(m, HZ, pp))[i];
                const float ex = e[0], ey = e[1 % NC], ez = e[2 % NC];
                if (m.Sa) {
                    const float a = ey * hz - ez * hy, b = ez * hx - ex * hz, c = ex * hy - ey * hx;
                    G(m.Sa)[i] += sa_sqrt(a * a + b * b + c * c) * dt;
                }
                if (m.E2) G(m.E2)[i] += (ex * ex + ey * ey + ez * ez) * dt;
            } else if (NC == 2) {
                const float hy = G(fld_ptr(m, HY, pp))[i];
                const float ex = e[0], ez = e[1 % NC];
                if (m.Sa) {
                    const float a = -ez * hy, b = ex * hy;
                    G(m.Sa)[i] += sa_sqrt(a * a + b * b) * dt;
                }
                if (m.E2) G(m.E2)[i] += (ex * ex + ez * ez) * dt;
            } else {
                const float hy = G(fld_ptr(m, HY, pp))[i];
                const float ex = e[0];
                if (m.Sa) {
                    const float a = ex * hy;
                    G(m.Sa)[i] += sa_sqrt(a * a) * dt;
                }
                if (m.E2) G(m.E2)[i] += ex * ex * dt;
            }
        }
    }
    check_finite(e0, m.flag, active);
}

void launch_e_state(const DevModel *m, const int4 *chunks, int nchunks, int ndims, bool integ,
                    hipStream_t s, bool pp) {
    if (nchunks <= 0) return;
#define ES_(NC, IN) hipLaunchKernelGGL((k_e_state<NC, IN>), dim3(nchunks), dim3(256), 0, s, m, chunks, nchunks, pp ? 1 : 0)
    if (ndims == 3) { if (integ) ES_(3, true); else ES_(3, false); }
    else if (ndims == 2) { if (integ) ES_(2, true); else ES_(2, false); }
    else { if (integ) ES_(1, true); else ES_(1, false); }
#undef ES_
}

}  // namespace mfd
