// 1D (z) Yee half-steps with CPML and materials (update_H_kernel! /
// update_E_kernel! 1D, src/models.jl:248-444): fields Hy, Dx, Ex.  One thread per
// cell; a tile is 256 consecutive z cells.
#include "common.hpp"
#include "material.hpp"

namespace mfd {

template <int MODE>
__global__ __launch_bounds__(256) void k_h1d(const DevModel *__restrict__ models,
                                             const Tile *__restrict__ tiles, int ntiles) {
    for (int t = blockIdx.x; t < ntiles; t += gridDim.x) {
        const Tile tl = tiles[t];
        if (tl.x < 0) continue;
        const DevModel &m = models[tl.x];
        if (*m.it >= m.Nt) continue;
        const int iz = tile_z0(tl) + threadIdx.x;
        const int Nz = m.Nz;
        if (iz >= tile_z1(tl)) continue;
        const float *__restrict__ Ex = m.F[EX];
        float *__restrict__ Hy = m.F[HY];
        const int bcz = m.bcs[5];
        float Mh = m.mats[0].Mh;
        if (MODE != MODE_VAC) Mh = m.mats[m.id[iz] & ID_MASK].Mh;
        const float ex = Ex[iz];
        float ex_zp;
        if (iz + 1 < Nz) ex_zp = Ex[iz + 1];
        else if (bcz == BC_PERIODIC) ex_zp = Ex[0];
        else if (bcz == BC_DIRICHLET) ex_zp = 0.0f;
        else ex_zp = ex;
        float dExz = (ex_zp - ex) * m.idz;
        float hy = Hy[iz];
        {
            const DevLayer &l = m.L[4];
            if (l.present && iz <= l.ib) {
                const int p = iz;
                const float a = l.B[p] * l.psiE0[p] + l.A[p] * dExz;
                l.psiE0[p] = a;
                hy -= Mh * a;
                dExz = dExz / l.K[p];
            }
        }
        {
            const DevLayer &l = m.L[5];
            if (l.present && iz >= l.ib) {
                const int p = iz - l.ib;
                const float a = l.B[p] * l.psiE0[p] + l.A[p] * dExz;
                l.psiE0[p] = a;
                hy -= Mh * a;
                dExz = dExz / l.K[p];
            }
        }
        hy -= Mh * (0.0f + dExz);
        Hy[iz] = hy;
    }
}

template <int MODE, bool INTEG>
__global__ __launch_bounds__(256) void k_e1d(const DevModel *__restrict__ models,
                                             const Tile *__restrict__ tiles, int ntiles) {
    for (int t = blockIdx.x; t < ntiles; t += gridDim.x) {
        const Tile tl = tiles[t];
        if (tl.x < 0) continue;
        const DevModel &m = models[tl.x];
        if (*m.it >= m.Nt) continue;
        const int iz = tile_z0(tl) + threadIdx.x;
        const int Nz = m.Nz;
        if (iz >= tile_z1(tl)) continue;
        const float *__restrict__ Hy = m.F[HY];
        float *__restrict__ Dx = m.F[DX];
        float *__restrict__ Ex = m.F[EX];
        const int bcz = m.bcs[4];
        int idb = 0;
        if (MODE != MODE_VAC || INTEG) idb = m.id[iz];
        const DevMat &mt = m.mats[MODE != MODE_VAC ? (idb & ID_MASK) : 0];
        const float Md1 = mt.Md1, Md2 = mt.Md2;
        const float hy = Hy[iz];
        float hy_zm;
        if (iz > 0) hy_zm = Hy[iz - 1];
        else if (bcz == BC_PERIODIC) hy_zm = Hy[Nz - 1];
        else if (bcz == BC_DIRICHLET) hy_zm = 0.0f;
        else hy_zm = hy;
        float dHyz = (hy - hy_zm) * m.idz;
        float d[1] = {Dx[iz]};
        {
            const DevLayer &l = m.L[4];
            if (l.present && iz <= l.ib) {
                const int p = iz;
                const float a = l.B[p] * l.psiH1[p] + l.A[p] * dHyz;
                l.psiH1[p] = a;
                d[0] -= Md2 * a;
                dHyz = dHyz / l.K[p];
            }
        }
        {
            const DevLayer &l = m.L[5];
            if (l.present && iz >= l.ib) {
                const int p = iz - l.ib;
                const float a = l.B[p] * l.psiH1[p] + l.A[p] * dHyz;
                l.psiH1[p] = a;
                d[0] -= Md2 * a;
                dHyz = dHyz / l.K[p];
            }
        }
        d[0] = Md1 * d[0] + Md2 * (0.0f - dHyz);
        Dx[iz] = d[0];
        float e[1] = {0.0f};
        if (MODE == MODE_FULL && (mt.flags & (MAT_DISP | MAT_PLASMA))) {
            e[0] = Ex[iz];
            material_e<1>(mt, m.slot[iz], d, e, m.dt, m.dtf);
        } else if (MODE == MODE_FULL && (mt.flags & MAT_KERR)) {
            material_e<1>(mt, 0, d, e, m.dt, m.dtf);
        } else {
            e[0] = d[0] * mt.Me;
        }
        Ex[iz] = e[0];
        check_finite(e[0], m.flag, true);
        if (INTEG && !(idb & ID_TOUCHED)) {   // src/outputs.jl:186-189
            const float dt = m.dtf;
            if (m.Sa) {
                const float a = e[0] * hy;
                m.Sa[iz] += sa_sqrt(a * a) * dt;
            }
            if (m.E2) m.E2[iz] += e[0] * e[0] * dt;
        }
    }
}

void launch_h1d(const DevModel *m, const Tile *tiles, int ntiles, int grid, int mode,
                hipStream_t s) {
    if (mode == MODE_VAC) hipLaunchKernelGGL(k_h1d<MODE_VAC>, dim3(grid), dim3(TX1), 0, s, m, tiles, ntiles);
    else hipLaunchKernelGGL(k_h1d<MODE_SIMPLE>, dim3(grid), dim3(TX1), 0, s, m, tiles, ntiles);
}

void launch_e1d(const DevModel *m, const Tile *tiles, int ntiles, int grid, int mode, bool integ,
                hipStream_t s) {
#define E1(MD, IN) hipLaunchKernelGGL((k_e1d<MD, IN>), dim3(grid), dim3(TX1), 0, s, m, tiles, ntiles)
    if (mode == MODE_VAC) { if (integ) E1(MODE_VAC, true); else E1(MODE_VAC, false); }
    else if (mode == MODE_SIMPLE) { if (integ) E1(MODE_SIMPLE, true); else E1(MODE_SIMPLE, false); }
    else { if (integ) E1(MODE_FULL, true); else E1(MODE_FULL, false); }
#undef E1
}

}  // namespace mfd
