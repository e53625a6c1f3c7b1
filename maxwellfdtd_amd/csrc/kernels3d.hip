// 3D Yee H and E half-steps fused with the CPML psi recursions and the per-cell
// material update (update_H_kernel! / update_E_kernel! 3D, src/models.jl:735-1175).
//
// Work decomposition: a tile is 64 (x) x 4 (y) columns and a z range; each of the 256
// threads owns one (x, y) column and marches up z, keeping the z-neighbour plane it
// already loaded in registers (E(z+1) becomes E(z) of the next plane in the H half,
// H(z) becomes H(z-1) in the E half).  x and y neighbours come from the L1/L2 lines the
// neighbouring lanes and waves of the same workgroup load.  A wave is one 64-wide x row,
// so every field load and store is a fully coalesced 256-byte row.  CPML work sits in
// branches that only the slab cells take.
#include "common.hpp"
#include "material.hpp"

namespace mfd {

template <int MODE>
__global__ __launch_bounds__(256) void k_h3d(const DevModel *__restrict__ models,
                                             const Tile *__restrict__ tiles, int ntiles) {
    for (int t = blockIdx.x; t < ntiles; t += gridDim.x) {
        const Tile tl = tiles[t];
        if (tl.x < 0) continue;
        const DevModel &m = models[tl.x];
        if (*m.it >= m.Nt) continue;
        const int ix = tl.y + threadIdx.x, iy = tl.z + threadIdx.y;
        const int z0 = tile_z0(tl), z1 = tile_z1(tl);
        const int Nx = m.Nx, Ny = m.Ny, Nz = m.Nz;
        if (ix >= Nx || iy >= Ny) continue;
        const long long sy = m.sy, sz = m.sz;
        const float *__restrict__ Ex = m.F[EX];
        const float *__restrict__ Ey = m.F[EY];
        const float *__restrict__ Ez = m.F[EZ];
        float *__restrict__ Hx = m.F[HX];
        float *__restrict__ Hy = m.F[HY];
        float *__restrict__ Hz = m.F[HZ];
        const float idx = m.idx, idy = m.idy, idz = m.idz;
        const long long col = ix + sy * iy;
        // x+1 / y+1 neighbour offsets with the high-face boundary conditions
        const bool xlast = ix == Nx - 1, ylast = iy == Ny - 1;
        const long long xoff = xlast ? -(long long)(Nx - 1) : 1;
        const long long yoff = ylast ? -(long long)(Ny - 1) * sy : sy;
        const int bcx = m.bcs[1], bcy = m.bcs[3], bcz = m.bcs[5];
        const float Mh0 = m.mats[0].Mh;
        float ex = Ex[col + sz * z0], ey = Ey[col + sz * z0];
        for (int iz = z0; iz < z1; iz++) {
            const long long i = col + sz * iz;
            const float ez = Ez[i];
            float ex_zp, ey_zp;
            if (iz + 1 < Nz) {
                ex_zp = Ex[i + sz];
                ey_zp = Ey[i + sz];
            } else if (bcz == BC_PERIODIC) {
                ex_zp = Ex[col];
                ey_zp = Ey[col];
            } else if (bcz == BC_DIRICHLET) {
                ex_zp = 0.0f; ey_zp = 0.0f;
            } else {
                ex_zp = ex; ey_zp = ey;
            }
            float ey_xp, ez_xp, ex_yp, ez_yp;
            if (!xlast || bcx == BC_PERIODIC) { ey_xp = Ey[i + xoff]; ez_xp = Ez[i + xoff]; }
            else if (bcx == BC_DIRICHLET) { ey_xp = 0.0f; ez_xp = 0.0f; }
            else { ey_xp = ey; ez_xp = ez; }
            if (!ylast || bcy == BC_PERIODIC) { ex_yp = Ex[i + yoff]; ez_yp = Ez[i + yoff]; }
            else if (bcy == BC_DIRICHLET) { ex_yp = 0.0f; ez_yp = 0.0f; }
            else { ex_yp = ex; ez_yp = ez; }
            float Mh = Mh0;
            if (MODE != MODE_VAC) Mh = m.mats[m.id[i] & ID_MASK].Mh;
            float dExy = (ex_yp - ex) * idy;
            float dExz = (ex_zp - ex) * idz;
            float dEyx = (ey_xp - ey) * idx;
            float dEyz = (ey_zp - ey) * idz;
            float dEzx = (ez_xp - ez) * idx;
            float dEzy = (ez_yp - ez) * idy;
            float hx = Hx[i], hy = Hy[i], hz = Hz[i];
            // CPML slabs (src/models.jl:805-871), in the reference's order
            {
                const DevLayer &l = m.L[0];
                if (l.present && ix <= l.ib) {
                    const int p = ix;
                    const long long j = p + (long long)l.nb * (iy + (long long)Ny * iz);
                    const v2f q = l.pE[j];
                    const float a0 = l.B[p] * q.x + l.A[p] * dEyx;
                    const float a1 = l.B[p] * q.y + l.A[p] * dEzx;
                    l.pE[j] = mk2(a0, a1);
                    hy += Mh * a1; hz -= Mh * a0;
                    dEyx = dEyx / l.K[p]; dEzx = dEzx / l.K[p];
                }
            }
            {
                const DevLayer &l = m.L[1];
                if (l.present && ix >= l.ib) {
                    const int p = ix - l.ib;
                    const long long j = p + (long long)l.nb * (iy + (long long)Ny * iz);
                    const v2f q = l.pE[j];
                    const float a0 = l.B[p] * q.x + l.A[p] * dEyx;
                    const float a1 = l.B[p] * q.y + l.A[p] * dEzx;
                    l.pE[j] = mk2(a0, a1);
                    hy += Mh * a1; hz -= Mh * a0;
                    dEyx = dEyx / l.K[p]; dEzx = dEzx / l.K[p];
                }
            }
            {
                const DevLayer &l = m.L[2];
                if (l.present && iy <= l.ib) {
                    const int p = iy;
                    const long long j = ix + (long long)Nx * (p + (long long)l.nb * iz);
                    const v2f q = l.pE[j];
                    const float a0 = l.B[p] * q.x + l.A[p] * dExy;
                    const float a1 = l.B[p] * q.y + l.A[p] * dEzy;
                    l.pE[j] = mk2(a0, a1);
                    hx -= Mh * a1; hz += Mh * a0;
                    dExy = dExy / l.K[p]; dEzy = dEzy / l.K[p];
                }
            }
            {
                const DevLayer &l = m.L[3];
                if (l.present && iy >= l.ib) {
                    const int p = iy - l.ib;
                    const long long j = ix + (long long)Nx * (p + (long long)l.nb * iz);
                    const v2f q = l.pE[j];
                    const float a0 = l.B[p] * q.x + l.A[p] * dExy;
                    const float a1 = l.B[p] * q.y + l.A[p] * dEzy;
                    l.pE[j] = mk2(a0, a1);
                    hx -= Mh * a1; hz += Mh * a0;
                    dExy = dExy / l.K[p]; dEzy = dEzy / l.K[p];
                }
            }
            {
                const DevLayer &l = m.L[4];
                if (l.present && iz <= l.ib) {
                    const int p = iz;
                    const long long j = col + sz * p;
                    const v2f q = l.pE[j];
                    const float a0 = l.B[p] * q.x + l.A[p] * dExz;
                    const float a1 = l.B[p] * q.y + l.A[p] * dEyz;
                    l.pE[j] = mk2(a0, a1);
                    hx += Mh * a1; hy -= Mh * a0;
                    dExz = dExz / l.K[p]; dEyz = dEyz / l.K[p];
                }
            }
            {
                const DevLayer &l = m.L[5];
                if (l.present && iz >= l.ib) {
                    const int p = iz - l.ib;
                    const long long j = col + sz * p;
                    const v2f q = l.pE[j];
                    const float a0 = l.B[p] * q.x + l.A[p] * dExz;
                    const float a1 = l.B[p] * q.y + l.A[p] * dEyz;
                    l.pE[j] = mk2(a0, a1);
                    hx += Mh * a1; hy -= Mh * a0;
                    dExz = dExz / l.K[p]; dEyz = dEyz / l.K[p];
                }
            }
            hx -= Mh * (dEzy - dEyz);
            hy -= Mh * (dExz - dEzx);
            hz -= Mh * (dEyx - dExy);
            Hx[i] = hx; Hy[i] = hy; Hz[i] = hz;
            ex = ex_zp; ey = ey_zp;
        }
    }
}

template <int MODE, bool INTEG>
__global__ __launch_bounds__(256) void k_e3d(const DevModel *__restrict__ models,
                                             const Tile *__restrict__ tiles, int ntiles) {
    for (int t = blockIdx.x; t < ntiles; t += gridDim.x) {
        const Tile tl = tiles[t];
        if (tl.x < 0) continue;
        if (tl.x < 0) continue;
        const DevModel &m = models[tl.x];
        if (*m.it >= m.Nt) continue;
        const int ix = tl.y + threadIdx.x, iy = tl.z + threadIdx.y;
        const int z0 = tile_z0(tl), z1 = tile_z1(tl);
        const int Nx = m.Nx, Ny = m.Ny, Nz = m.Nz;
        if (ix >= Nx || iy >= Ny) continue;
        const long long sy = m.sy, sz = m.sz;
        const float *__restrict__ Hx = m.F[HX];
        const float *__restrict__ Hy = m.F[HY];
        const float *__restrict__ Hz = m.F[HZ];
        float *__restrict__ Dx = m.F[DX];
        float *__restrict__ Dy = m.F[DY];
        float *__restrict__ Dz = m.F[DZ];
        float *__restrict__ Ex = m.F[EX];
        float *__restrict__ Ey = m.F[EY];
        float *__restrict__ Ez = m.F[EZ];
        const float idx = m.idx, idy = m.idy, idz = m.idz;
        const long long col = ix + sy * iy;
        const bool xfirst = ix == 0, yfirst = iy == 0;
        const long long xoff = xfirst ? (long long)(Nx - 1) : -1;
        const long long yoff = yfirst ? (long long)(Ny - 1) * sy : -sy;
        const int bcx = m.bcs[0], bcy = m.bcs[2], bcz = m.bcs[4];
        const DevMat &vac = m.mats[0];
        // H(z0-1) for the first plane
        float hx_zm = 0.0f, hy_zm = 0.0f;
        bool zm_own = false;
        if (z0 > 0) {
            hx_zm = Hx[col + sz * (z0 - 1)];
            hy_zm = Hy[col + sz * (z0 - 1)];
        } else if (bcz == BC_PERIODIC) {
            hx_zm = Hx[col + sz * (Nz - 1)];
            hy_zm = Hy[col + sz * (Nz - 1)];
        } else if (bcz == BC_NEUMANN) {
            zm_own = true;
        }
        for (int iz = z0; iz < z1; iz++) {
            const long long i = col + sz * iz;
            const float hx = Hx[i], hy = Hy[i], hz = Hz[i];
            if (zm_own) { hx_zm = hx; hy_zm = hy; zm_own = false; }
            float hy_xm, hz_xm, hx_ym, hz_ym;
            if (!xfirst || bcx == BC_PERIODIC) { hy_xm = Hy[i + xoff]; hz_xm = Hz[i + xoff]; }
            else if (bcx == BC_DIRICHLET) { hy_xm = 0.0f; hz_xm = 0.0f; }
            else { hy_xm = hy; hz_xm = hz; }
            if (!yfirst || bcy == BC_PERIODIC) { hx_ym = Hx[i + yoff]; hz_ym = Hz[i + yoff]; }
            else if (bcy == BC_DIRICHLET) { hx_ym = 0.0f; hz_ym = 0.0f; }
            else { hx_ym = hx; hz_ym = hz; }
            int idb = 0;
            if (MODE != MODE_VAC || INTEG) idb = m.id[i];
            const int g = idb & ID_MASK;
            const DevMat &mt = MODE != MODE_VAC ? m.mats[g] : vac;
            const float Md1 = mt.Md1, Md2 = mt.Md2;
            float dHxy = (hx - hx_ym) * idy;
            float dHxz = (hx - hx_zm) * idz;
            float dHyx = (hy - hy_xm) * idx;
            float dHyz = (hy - hy_zm) * idz;
            float dHzx = (hz - hz_xm) * idx;
            float dHzy = (hz - hz_ym) * idy;
            float d[3] = {Dx[i], Dy[i], Dz[i]};
            // CPML slabs (src/models.jl:976-1035)
            {
                const DevLayer &l = m.L[0];
                if (l.present && ix <= l.ib) {
                    const int p = ix;
                    const long long j = p + (long long)l.nb * (iy + (long long)Ny * iz);
                    const v2f q = l.pH[j];
                    const float a0 = l.B[p] * q.x + l.A[p] * dHyx;
                    const float a1 = l.B[p] * q.y + l.A[p] * dHzx;
                    l.pH[j] = mk2(a0, a1);
                    d[1] -= Md2 * a1; d[2] += Md2 * a0;
                    dHyx = dHyx / l.K[p]; dHzx = dHzx / l.K[p];
                }
            }
            {
                const DevLayer &l = m.L[1];
                if (l.present && ix >= l.ib) {
                    const int p = ix - l.ib;
                    const long long j = p + (long long)l.nb * (iy + (long long)Ny * iz);
                    const v2f q = l.pH[j];
                    const float a0 = l.B[p] * q.x + l.A[p] * dHyx;
                    const float a1 = l.B[p] * q.y + l.A[p] * dHzx;
                    l.pH[j] = mk2(a0, a1);
                    d[1] -= Md2 * a1; d[2] += Md2 * a0;
                    dHyx = dHyx / l.K[p]; dHzx = dHzx / l.K[p];
                }
            }
            {
                const DevLayer &l = m.L[2];
                if (l.present && iy <= l.ib) {
                    const int p = iy;
                    const long long j = ix + (long long)Nx * (p + (long long)l.nb * iz);
                    const v2f q = l.pH[j];
                    const float a0 = l.B[p] * q.x + l.A[p] * dHxy;
                    const float a1 = l.B[p] * q.y + l.A[p] * dHzy;
                    l.pH[j] = mk2(a0, a1);
                    d[0] += Md2 * a1; d[2] -= Md2 * a0;
                    dHxy = dHxy / l.K[p]; dHzy = dHzy / l.K[p];
                }
            }
            {
                const DevLayer &l = m.L[3];
                if (l.present && iy >= l.ib) {
                    const int p = iy - l.ib;
                    const long long j = ix + (long long)Nx * (p + (long long)l.nb * iz);
                    const v2f q = l.pH[j];
                    const float a0 = l.B[p] * q.x + l.A[p] * dHxy;
                    const float a1 = l.B[p] * q.y + l.A[p] * dHzy;
                    l.pH[j] = mk2(a0, a1);
                    d[0] += Md2 * a1; d[2] -= Md2 * a0;
                    dHxy = dHxy / l.K[p]; dHzy = dHzy / l.K[p];
                }
            }
            {
                const DevLayer &l = m.L[4];
                if (l.present && iz <= l.ib) {
                    const int p = iz;
                    const long long j = col + sz * p;
                    const v2f q = l.pH[j];
                    const float a0 = l.B[p] * q.x + l.A[p] * dHxz;
                    const float a1 = l.B[p] * q.y + l.A[p] * dHyz;
                    l.pH[j] = mk2(a0, a1);
                    d[0] -= Md2 * a1; d[1] += Md2 * a0;
                    dHxz = dHxz / l.K[p]; dHyz = dHyz / l.K[p];
                }
            }
            {
                const DevLayer &l = m.L[5];
                if (l.present && iz >= l.ib) {
                    const int p = iz - l.ib;
                    const long long j = col + sz * p;
                    const v2f q = l.pH[j];
                    const float a0 = l.B[p] * q.x + l.A[p] * dHxz;
                    const float a1 = l.B[p] * q.y + l.A[p] * dHyz;
                    l.pH[j] = mk2(a0, a1);
                    d[0] -= Md2 * a1; d[1] += Md2 * a0;
                    dHxz = dHxz / l.K[p]; dHyz = dHyz / l.K[p];
                }
            }
            d[0] = Md1 * d[0] + Md2 * (dHzy - dHyz);
            d[1] = Md1 * d[1] + Md2 * (dHxz - dHzx);
            d[2] = Md1 * d[2] + Md2 * (dHyx - dHxy);
            Dx[i] = d[0]; Dy[i] = d[1]; Dz[i] = d[2];
            float e[3] = {0.0f, 0.0f, 0.0f};
            if (MODE == MODE_FULL && (mt.flags & (MAT_DISP | MAT_PLASMA))) {
                e[0] = Ex[i]; e[1] = Ey[i]; e[2] = Ez[i];
                material_e<3>(mt, m.slot[i], *m.it & 1, d, e, m.dt, m.dtf);
            } else if (MODE == MODE_FULL && (mt.flags & MAT_KERR)) {
                material_e<3>(mt, 0, 0, d, e, m.dt, m.dtf);
            } else {
                const float Me = mt.Me;
                e[0] = d[0] * Me; e[1] = d[1] * Me; e[2] = d[2] * Me;
            }
            Ex[i] = e[0]; Ey[i] = e[1]; Ez[i] = e[2];
            check_finite(e[0], m.flag, true);
            // integral variables of the step (src/outputs.jl:480-483); cells that a source
            // rewrites afterwards are accumulated by the post-source fixup instead
            if (INTEG && !(idb & ID_TOUCHED)) {
                const float dt = m.dtf;
                if (m.Sa) {
                    const float a = e[1] * hz - e[2] * hy, b = e[2] * hx - e[0] * hz;
                    const float c = e[0] * hy - e[1] * hx;
                    m.Sa[i] += sa_sqrt(a * a + b * b + c * c) * dt;
                }
                if (m.E2) m.E2[i] += (e[0] * e[0] + e[1] * e[1] + e[2] * e[2]) * dt;
            }
            hx_zm = hx; hy_zm = hy;
        }
    }
}

void launch_h3d(const DevModel *m, const Tile *tiles, int ntiles, int grid, int mode,
                hipStream_t s) {
    dim3 block(TX3, TY3);
    if (mode == MODE_VAC) hipLaunchKernelGGL(k_h3d<MODE_VAC>, dim3(grid), block, 0, s, m, tiles, ntiles);
    else hipLaunchKernelGGL(k_h3d<MODE_SIMPLE>, dim3(grid), block, 0, s, m, tiles, ntiles);
}

void launch_e3d(const DevModel *m, const Tile *tiles, int ntiles, int grid, int mode, bool integ,
                hipStream_t s) {
    dim3 block(TX3, TY3);
#define E3(MD, IN) hipLaunchKernelGGL((k_e3d<MD, IN>), dim3(grid), block, 0, s, m, tiles, ntiles)
    if (mode == MODE_VAC) { if (integ) E3(MODE_VAC, true); else E3(MODE_VAC, false); }
    else if (mode == MODE_SIMPLE) { if (integ) E3(MODE_SIMPLE, true); else E3(MODE_SIMPLE, false); }
    else { if (integ) E3(MODE_FULL, true); else E3(MODE_FULL, false); }
#undef E3
}

}  // namespace mfd
