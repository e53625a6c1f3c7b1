// 2D (x, z) Yee half-steps with CPML and materials (update_H_kernel! /
// update_E_kernel! 2D, src/models.jl:450-726): fields Hy, Dx, Dz, Ex, Ez.
// A tile is 256 consecutive x cells (four 64-wide waves) and a z range; each thread
// marches up z keeping the z-neighbour in registers.
#include "common.hpp"
#include "material.hpp"

namespace mfd {

template <int MODE>
__global__ __launch_bounds__(256) void k_h2d(const DevModel *__restrict__ models,
                                             const Tile *__restrict__ tiles, int ntiles) {
    for (int t = blockIdx.x; t < ntiles; t += gridDim.x) {
        const Tile tl = tiles[t];
        if (tl.x < 0) continue;
        const DevModel &m = models[tl.x];
        if (*m.it >= m.Nt) continue;
        const int ix = tl.y + threadIdx.x;
        const int z0 = tile_z0(tl), z1 = tile_z1(tl);
        const int Nx = m.Nx, Nz = m.Nz;
        if (ix >= Nx) continue;
        const long long sz = m.sz;
        const float *__restrict__ Ex = m.F[EX];
        const float *__restrict__ Ez = m.F[EZ];
        float *__restrict__ Hy = m.F[HY];
        const float idx = m.idx, idz = m.idz;
        const bool xlast = ix == Nx - 1;
        const long long xoff = xlast ? -(long long)(Nx - 1) : 1;
        const int bcx = m.bcs[1], bcz = m.bcs[5];
        const float Mh0 = m.mats[0].Mh;
        float ex = Ex[ix + sz * z0];
        for (int iz = z0; iz < z1; iz++) {
            const long long i = ix + sz * iz;
            const float ez = Ez[i];
            float ex_zp;
            if (iz + 1 < Nz) ex_zp = Ex[i + sz];
            else if (bcz == BC_PERIODIC) ex_zp = Ex[ix];
            else if (bcz == BC_DIRICHLET) ex_zp = 0.0f;
            else ex_zp = ex;
            float ez_xp;
            if (!xlast || bcx == BC_PERIODIC) ez_xp = Ez[i + xoff];
            else if (bcx == BC_DIRICHLET) ez_xp = 0.0f;
            else ez_xp = ez;
            float Mh = Mh0;
            if (MODE != MODE_VAC) Mh = m.mats[m.id[i] & ID_MASK].Mh;
            float dExz = (ex_zp - ex) * idz;
            float dEzx = (ez_xp - ez) * idx;
            float hy = Hy[i];
            {
                const DevLayer &l = m.L[0];
                if (l.present && ix <= l.ib) {
                    const int p = ix;
                    const long long j = p + (long long)l.nb * iz;
                    const float a = l.B[p] * l.psiE1[j] + l.A[p] * dEzx;
                    l.psiE1[j] = a;
                    hy += Mh * a;
                    dEzx = dEzx / l.K[p];
                }
            }
            {
                const DevLayer &l = m.L[1];
                if (l.present && ix >= l.ib) {
                    const int p = ix - l.ib;
                    const long long j = p + (long long)l.nb * iz;
                    const float a = l.B[p] * l.psiE1[j] + l.A[p] * dEzx;
                    l.psiE1[j] = a;
                    hy += Mh * a;
                    dEzx = dEzx / l.K[p];
                }
            }
            {
                const DevLayer &l = m.L[4];
                if (l.present && iz <= l.ib) {
                    const int p = iz;
                    const long long j = ix + (long long)Nx * p;
                    const float a = l.B[p] * l.psiE0[j] + l.A[p] * dExz;
                    l.psiE0[j] = a;
                    hy -= Mh * a;
                    dExz = dExz / l.K[p];
                }
            }
            {
                const DevLayer &l = m.L[5];
                if (l.present && iz >= l.ib) {
                    const int p = iz - l.ib;
                    const long long j = ix + (long long)Nx * p;
                    const float a = l.B[p] * l.psiE0[j] + l.A[p] * dExz;
                    l.psiE0[j] = a;
                    hy -= Mh * a;
                    dExz = dExz / l.K[p];
                }
            }
            hy -= Mh * (dExz - dEzx);
            Hy[i] = hy;
            ex = ex_zp;
        }
    }
}

template <int MODE, bool INTEG>
__global__ __launch_bounds__(256) void k_e2d(const DevModel *__restrict__ models,
                                             const Tile *__restrict__ tiles, int ntiles) {
    for (int t = blockIdx.x; t < ntiles; t += gridDim.x) {
        const Tile tl = tiles[t];
        if (tl.x < 0) continue;
        const DevModel &m = models[tl.x];
        if (*m.it >= m.Nt) continue;
        const int ix = tl.y + threadIdx.x;
        const int z0 = tile_z0(tl), z1 = tile_z1(tl);
        const int Nx = m.Nx, Nz = m.Nz;
        if (ix >= Nx) continue;
        const long long sz = m.sz;
        const float *__restrict__ Hy = m.F[HY];
        float *__restrict__ Dx = m.F[DX];
        float *__restrict__ Dz = m.F[DZ];
        float *__restrict__ Ex = m.F[EX];
        float *__restrict__ Ez = m.F[EZ];
        const float idx = m.idx, idz = m.idz;
        const bool xfirst = ix == 0;
        const long long xoff = xfirst ? (long long)(Nx - 1) : -1;
        const int bcx = m.bcs[0], bcz = m.bcs[4];
        const DevMat &vac = m.mats[0];
        float hy_zm = 0.0f;
        bool zm_own = false;
        if (z0 > 0) hy_zm = Hy[ix + sz * (z0 - 1)];
        else if (bcz == BC_PERIODIC) hy_zm = Hy[ix + sz * (Nz - 1)];
        else if (bcz == BC_NEUMANN) zm_own = true;
        for (int iz = z0; iz < z1; iz++) {
            const long long i = ix + sz * iz;
            const float hy = Hy[i];
            if (zm_own) { hy_zm = hy; zm_own = false; }
            float hy_xm;
            if (!xfirst || bcx == BC_PERIODIC) hy_xm = Hy[i + xoff];
            else if (bcx == BC_DIRICHLET) hy_xm = 0.0f;
            else hy_xm = hy;
            int idb = 0;
            if (MODE != MODE_VAC || INTEG) idb = m.id[i];
            const DevMat &mt = MODE != MODE_VAC ? m.mats[idb & ID_MASK] : vac;
            const float Md1 = mt.Md1, Md2 = mt.Md2;
            float dHyx = (hy - hy_xm) * idx;
            float dHyz = (hy - hy_zm) * idz;
            float d[2] = {Dx[i], Dz[i]};
            {
                const DevLayer &l = m.L[0];
                if (l.present && ix <= l.ib) {
                    const int p = ix;
                    const long long j = p + (long long)l.nb * iz;
                    const float a = l.B[p] * l.psiH0[j] + l.A[p] * dHyx;
                    l.psiH0[j] = a;
                    d[1] += Md2 * a;
                    dHyx = dHyx / l.K[p];
                }
            }
            {
                const DevLayer &l = m.L[1];
                if (l.present && ix >= l.ib) {
                    const int p = ix - l.ib;
                    const long long j = p + (long long)l.nb * iz;
                    const float a = l.B[p] * l.psiH0[j] + l.A[p] * dHyx;
                    l.psiH0[j] = a;
                    d[1] += Md2 * a;
                    dHyx = dHyx / l.K[p];
                }
            }
            {
                const DevLayer &l = m.L[4];
                if (l.present && iz <= l.ib) {
                    const int p = iz;
                    const long long j = ix + (long long)Nx * p;
                    const float a = l.B[p] * l.psiH1[j] + l.A[p] * dHyz;
                    l.psiH1[j] = a;
                    d[0] -= Md2 * a;
                    dHyz = dHyz / l.K[p];
                }
            }
            {
                const DevLayer &l = m.L[5];
                if (l.present && iz >= l.ib) {
                    const int p = iz - l.ib;
                    const long long j = ix + (long long)Nx * p;
                    const float a = l.B[p] * l.psiH1[j] + l.A[p] * dHyz;
                    l.psiH1[j] = a;
                    d[0] -= Md2 * a;
                    dHyz = dHyz / l.K[p];
                }
            }
            d[0] = Md1 * d[0] + Md2 * (0.0f - dHyz);
            d[1] = Md1 * d[1] + Md2 * (dHyx - 0.0f);
            Dx[i] = d[0]; Dz[i] = d[1];
            float e[2] = {0.0f, 0.0f};
            // cells of ADE / plasma materials: E is finished by the state pass (k_e_state)
            const bool defer = MODE == MODE_FULL && (mt.flags & (MAT_DISP | MAT_PLASMA));
            if (MODE == MODE_FULL && !defer && (mt.flags & MAT_KERR)) {
                material_e<2>(mt, 0, d, e, m.dt, m.dtf);
            } else {
                e[0] = d[0] * mt.Me; e[1] = d[1] * mt.Me;
            }
            if (!defer) { Ex[i] = e[0]; Ez[i] = e[1]; }
            check_finite(e[0], m.flag, !defer);
            if (INTEG && !defer && !(idb & ID_TOUCHED)) {   // src/outputs.jl:322-325
                const float dt = m.dtf;
                if (m.Sa) {
                    const float a = -e[1] * hy, b = e[0] * hy;
                    m.Sa[i] += sqrtf(a * a + b * b) * dt;
                }
                if (m.E2) m.E2[i] += (e[0] * e[0] + e[1] * e[1]) * dt;
            }
            hy_zm = hy;
        }
    }
}

void launch_h2d(const DevModel *m, const Tile *tiles, int ntiles, int grid, int mode,
                hipStream_t s) {
    if (mode == MODE_VAC) hipLaunchKernelGGL(k_h2d<MODE_VAC>, dim3(grid), dim3(TX2), 0, s, m, tiles, ntiles);
    else hipLaunchKernelGGL(k_h2d<MODE_SIMPLE>, dim3(grid), dim3(TX2), 0, s, m, tiles, ntiles);
}

void launch_e2d(const DevModel *m, const Tile *tiles, int ntiles, int grid, int mode, bool integ,
                hipStream_t s) {
#define E2(MD, IN) hipLaunchKernelGGL((k_e2d<MD, IN>), dim3(grid), dim3(TX2), 0, s, m, tiles, ntiles)
    if (mode == MODE_VAC) { if (integ) E2(MODE_VAC, true); else E2(MODE_VAC, false); }
    else if (mode == MODE_SIMPLE) { if (integ) E2(MODE_SIMPLE, true); else E2(MODE_SIMPLE, false); }
    else { if (integ) E2(MODE_FULL, true); else E2(MODE_FULL, false); }
#undef E2
}

}  // namespace mfd
