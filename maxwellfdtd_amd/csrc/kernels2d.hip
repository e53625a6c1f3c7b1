// 2D (x, z) Yee half-steps with CPML and materials (update_H_kernel! /
// update_E_kernel! 2D, src/models.jl:450-726): fields Hy, Dx, Dz, Ex, Ez.
//
// A tile is 64 * TW consecutive x cells (TW waves side by side) and a z range; each lane
// owns one x column and marches up z.  The loads of plane k+1 are issued before plane k
// is computed and stored (two planes in flight per wave), the in-wave x neighbour comes
// through a cross-lane shuffle (edge lanes load theirs), material coefficients come
// from LDS tables, and all field accesses use the global address space with 32-bit
// cell offsets.  Cells of ADE/plasma materials get their E from the state pass.
#include "common.hpp"
#include "material.hpp"

namespace mfd {

typedef __attribute__((address_space(1))) float gf1;
typedef __attribute__((address_space(1))) const float gcf1;
typedef const __attribute__((address_space(1))) unsigned char gcu1;

// CPML slab update of one derivative: psi = B psi + A d, d /= K (src/models.jl:463-520),
// the division as a product with 1/K rounded once on the host
__device__ __forceinline__ float slab1(const DevLayer &l, int p, unsigned j, float *psi, float &d) {
    const float a = G(l.B)[p] * G(psi)[j] + G(l.A)[p] * d;
    G(psi)[j] = a;
    d = d * G(l.Ki)[p];
    return a;
}

struct H2P {
    float ez, exn, hy, ezE;
    unsigned char g;
};

template <int MODE, int TW>
__global__ __launch_bounds__(64 * TW) void k_h2d(const DevModel *__restrict__ models,
                                                 const Tile *__restrict__ tiles, int ntiles) {
    __shared__ float mh_tab[128];
    const int lane = threadIdx.x, w = threadIdx.y;
    for (int t = blockIdx.x; t < ntiles; t += gridDim.x) {
        const Tile tl = tiles[t];
        if (tl.x < 0) continue;
        const DevModel &m = models[tl.x];
        if (*m.it >= m.Nt) continue;
        const int nm = MODE == MODE_VAC ? 0 : m.nmat;
        for (int k = lane + 64 * w; k <= nm; k += 64 * TW) mh_tab[k] = m.mats[k].Mh;
        __syncthreads();
        const int Nx = m.Nx, Nz = m.Nz;
        const int ix = tl.y + 64 * w + lane;
        if (ix < Nx) {
            const int z0 = tile_z0(tl), z1 = tile_z1(tl);
            gcf1 *Ex = G((const float *)m.F[EX]), *Ez = G((const float *)m.F[EZ]);
            gf1 *Hy = G(m.F[HY]);
            gcu1 *id = G(m.id);
            const unsigned sz = (unsigned)m.sz;
            const float idx = m.idx, idz = m.idz;
            const int bcx = m.bcs[1], bcz = m.bcs[5];
            const bool edge = lane == 63 || ix + 1 >= Nx;
            const bool eload = edge && (ix + 1 < Nx || bcx == BC_PERIODIC);
            const unsigned xoff = ix + 1 < Nx ? 1u : (unsigned)(1 - Nx);
            const DevLayer &L0 = m.L[0], &L1 = m.L[1], &L4 = m.L[4], &L5 = m.L[5];
            const bool xs0 = L0.present && ix <= L0.ib, xs1 = L1.present && ix >= L1.ib;
            auto load = [&](int iz, H2P &P) {
                const unsigned i = ix + sz * iz;
                P.ez = Ez[i];
                if (iz + 1 < Nz) P.exn = Ex[i + sz];
                else if (bcz == BC_PERIODIC) P.exn = Ex[ix];
                P.hy = Hy[i];
                if (eload) P.ezE = Ez[i + xoff];
                if (MODE != MODE_VAC) P.g = id[i];
            };
            float ex = Ex[ix + sz * z0];
            auto compute = [&](int iz, H2P &P) {
                const unsigned i = ix + sz * iz;
                if (iz + 1 >= Nz && bcz != BC_PERIODIC) P.exn = bcz == BC_DIRICHLET ? 0.0f : ex;
                float ezx = __shfl_down(P.ez, 1);
                if (edge) {
                    if (eload) ezx = P.ezE;
                    else if (bcx == BC_DIRICHLET) ezx = 0.0f;
                    else ezx = P.ez;
                }
                const float Mh = MODE != MODE_VAC ? mh_tab[P.g & ID_MASK] : mh_tab[0];
                float dExz = (P.exn - ex) * idz;
                float dEzx = (ezx - P.ez) * idx;
                float hy = P.hy;
                if (xs0) hy += Mh * slab1(L0, ix, ix + (unsigned)L0.nb * iz, L0.psiE1, dEzx);
                if (xs1) {
                    const int p = ix - L1.ib;
                    hy += Mh * slab1(L1, p, p + (unsigned)L1.nb * iz, L1.psiE1, dEzx);
                }
                if (L4.present && iz <= L4.ib) hy -= Mh * slab1(L4, iz, ix + (unsigned)Nx * iz, L4.psiE0, dExz);
                if (L5.present && iz >= L5.ib) {
                    const int p = iz - L5.ib;
                    hy -= Mh * slab1(L5, p, ix + (unsigned)Nx * p, L5.psiE0, dExz);
                }
                hy -= Mh * (dExz - dEzx);
                Hy[i] = hy;
                ex = P.exn;
            };
            H2P A, B;
            load(z0, A);
            for (int iz = z0; iz < z1; iz += 2) {
                if (iz + 1 < z1) load(iz + 1, B);
                compute(iz, A);
                if (iz + 1 >= z1) break;
                if (iz + 2 < z1) load(iz + 2, A);
                compute(iz + 1, B);
            }
        }
        __syncthreads();
    }
}

struct E2P {
    float hy, d0, d1, hyE, sa, e2;
    unsigned char g;
};

template <int MODE, bool INTEG, int TW>
__global__ __launch_bounds__(64 * TW) void k_e2d(const DevModel *__restrict__ models,
                                                 const Tile *__restrict__ tiles, int ntiles) {
    // (Md1, Md2, Me, kind): kind 0 plain, 1 Kerr only, 2 ADE/plasma state; cw = Me Md2
    __shared__ float4 ctab[128];
    __shared__ float2 ktab[128];
    __shared__ float cw[128];
    const int lane = threadIdx.x, w = threadIdx.y;
    for (int t = blockIdx.x; t < ntiles; t += gridDim.x) {
        const Tile tl = tiles[t];
        if (tl.x < 0) continue;
        const DevModel &m = models[tl.x];
        if (*m.it >= m.Nt) continue;
        const int nm = MODE == MODE_VAC ? 0 : m.nmat;
        for (int k = lane + 64 * w; k <= nm; k += 64 * TW) {
            const DevMat &mt = m.mats[k];
            const float kind = (mt.flags & (MAT_DISP | MAT_PLASMA)) ? 2.0f : (mt.flags & MAT_KERR) ? 1.0f : 0.0f;
            ctab[k] = make_float4(mt.Md1, mt.Md2, mt.Me, kind);
            cw[k] = mt.MeMd2;
            if (MODE == MODE_FULL) ktab[k] = make_float2(mt.K2, mt.K3);
        }
        __syncthreads();
        const int Nx = m.Nx, Nz = m.Nz;
        const int ix = tl.y + 64 * w + lane;
        if (ix < Nx) {
            const int z0 = tile_z0(tl), z1 = tile_z1(tl);
            gcf1 *Hy = G((const float *)m.F[HY]);
            gf1 *Dx = G(m.F[DX]), *Dz = G(m.F[DZ]), *Ex = G(m.F[EX]), *Ez = G(m.F[EZ]);
            gf1 *Sa = G(m.Sa), *E2 = G(m.E2);
            const bool doSa = INTEG && m.Sa, doE2 = INTEG && m.E2;
            gcu1 *id = G(m.id);
            const unsigned sz = (unsigned)m.sz;
            const float idx = m.idx, idz = m.idz, dtf = m.dtf;
            const int bcx = m.bcs[0], bcz = m.bcs[4];
            const bool edge = lane == 0 || ix == 0;
            const bool eload = edge && (ix > 0 || bcx == BC_PERIODIC);
            const unsigned xoff = ix > 0 ? (unsigned)-1 : (unsigned)(Nx - 1);
            const DevLayer &L0 = m.L[0], &L1 = m.L[1], &L4 = m.L[4], &L5 = m.L[5];
            const bool xs0 = L0.present && ix <= L0.ib, xs1 = L1.present && ix >= L1.ib;
            // without stored D, a plain cell that no source writes reads E and takes the
            // update in E form, E' = Md1 (E + Me Md2 psi) + Me Md2 curl H (Me times the D
            // form); Kerr, ADE and source cells keep their D, loaded once the id is known
            const bool dless = m.dless;
            gcf1 *Dsx = G((const float *)m.F[DX]), *Dsz = G((const float *)m.F[DZ]);
            gcf1 *Esx = G((const float *)m.F[EX]), *Esz = G((const float *)m.F[EZ]);
            auto load = [&](int iz, E2P &P) {
                const unsigned i = ix + sz * iz;
                P.hy = Hy[i];
                P.d0 = (dless ? Esx : Dsx)[i]; P.d1 = (dless ? Esz : Dsz)[i];
                if (eload) P.hyE = Hy[i + xoff];
                if (MODE != MODE_VAC || INTEG || dless) P.g = id[i];
                if (doSa) P.sa = Sa[i];
                if (doE2) P.e2 = E2[i];
            };
            float hym = 0.0f;       // Hy(z-1)
            bool zm_own = false;
            if (z0 > 0) hym = Hy[ix + sz * (z0 - 1)];
            else if (bcz == BC_PERIODIC) hym = Hy[ix + sz * (Nz - 1)];
            else if (bcz == BC_NEUMANN) zm_own = true;
            auto compute = [&](int iz, E2P &P) {
                const unsigned i = ix + sz * iz;
                if (zm_own) { hym = P.hy; zm_own = false; }
                float hyx = __shfl_up(P.hy, 1);
                if (edge) {
                    if (eload) hyx = P.hyE;
                    else if (bcx == BC_DIRICHLET) hyx = 0.0f;
                    else hyx = P.hy;
                }
                const int g = P.g & ID_MASK;
                const float4 cf = MODE != MODE_VAC ? ctab[g] : ctab[0];
                const bool dc = !dless || (P.g & ID_TOUCHED) || (MODE == MODE_FULL && cf.w != 0.0f);
                float d0 = P.d0, d1 = P.d1;
                if (dless && dc) { d0 = Dsx[i]; d1 = Dsz[i]; }
                const float Md1 = cf.x, Md2 = dc ? cf.y : (MODE != MODE_VAC ? cw[g] : cw[0]);
                float dHyx = (P.hy - hyx) * idx;
                float dHyz = (P.hy - hym) * idz;
                if (xs0) d1 += Md2 * slab1(L0, ix, ix + (unsigned)L0.nb * iz, L0.psiH0, dHyx);
                if (xs1) {
                    const int p = ix - L1.ib;
                    d1 += Md2 * slab1(L1, p, p + (unsigned)L1.nb * iz, L1.psiH0, dHyx);
                }
                if (L4.present && iz <= L4.ib) d0 -= Md2 * slab1(L4, iz, ix + (unsigned)Nx * iz, L4.psiH1, dHyz);
                if (L5.present && iz >= L5.ib) {
                    const int p = iz - L5.ib;
                    d0 -= Md2 * slab1(L5, p, ix + (unsigned)Nx * p, L5.psiH1, dHyz);
                }
                d0 = Md1 * d0 + Md2 * (0.0f - dHyz);
                d1 = Md1 * d1 + Md2 * (dHyx - 0.0f);
                float e0 = dc ? d0 * cf.z : d0, e1 = dc ? d1 * cf.z : d1;
                bool defer = false;
                if (MODE == MODE_FULL && cf.w != 0.0f) {
                    if (cf.w > 1.5f) {
                        defer = true;
                    } else {   // Kerr Pade, src/models.jl:699-704
                        const float2 kk = ktab[g];
                        e0 = (1.0f + kk.x * d0 + 2.0f * kk.y * (d0 * d0)) /
                             (1.0f + 2.0f * kk.x * d0 + 3.0f * kk.y * (d0 * d0)) * d0 * cf.z;
                        e1 = (1.0f + kk.x * d1 + 2.0f * kk.y * (d1 * d1)) /
                             (1.0f + 2.0f * kk.x * d1 + 3.0f * kk.y * (d1 * d1)) * d1 * cf.z;
                    }
                }
                if (dc) { Dx[i] = d0; Dz[i] = d1; }
                if (!defer) { Ex[i] = e0; Ez[i] = e1; }
                check_finite(e0, m.flag, !defer);
                if (INTEG && !defer && !(P.g & ID_TOUCHED)) {   // src/outputs.jl:322-325
                    if (doSa) {
                        const float a = -e1 * P.hy, b = e0 * P.hy;
                        Sa[i] = P.sa + sa_sqrt(a * a + b * b) * dtf;
                    }
                    if (doE2) E2[i] = P.e2 + (e0 * e0 + e1 * e1) * dtf;
                }
                hym = P.hy;
            };
            E2P A, B;
            load(z0, A);
            for (int iz = z0; iz < z1; iz += 2) {
                if (iz + 1 < z1) load(iz + 1, B);
                compute(iz, A);
                if (iz + 1 >= z1) break;
                if (iz + 2 < z1) load(iz + 2, A);
                compute(iz + 1, B);
            }
        }
        __syncthreads();
    }
}

template <int TW>
static void lh2(const DevModel *m, const Tile *tiles, int ntiles, int grid, int mode, hipStream_t s) {
    const dim3 block(64, TW);
    if (mode == MODE_VAC) hipLaunchKernelGGL((k_h2d<MODE_VAC, TW>), dim3(grid), block, 0, s, m, tiles, ntiles);
    else hipLaunchKernelGGL((k_h2d<MODE_SIMPLE, TW>), dim3(grid), block, 0, s, m, tiles, ntiles);
}

template <int TW>
static void le2(const DevModel *m, const Tile *tiles, int ntiles, int grid, int mode, bool integ,
                hipStream_t s) {
    const dim3 block(64, TW);
#define E2_(MD, IN) hipLaunchKernelGGL((k_e2d<MD, IN, TW>), dim3(grid), block, 0, s, m, tiles, ntiles)
    if (mode == MODE_VAC) { if (integ) E2_(MODE_VAC, true); else E2_(MODE_VAC, false); }
    else if (mode == MODE_SIMPLE) { if (integ) E2_(MODE_SIMPLE, true); else E2_(MODE_SIMPLE, false); }
    else { if (integ) E2_(MODE_FULL, true); else E2_(MODE_FULL, false); }
#undef E2_
}

void launch_h2d(const DevModel *m, const Tile *tiles, int ntiles, int grid, int mode, int tw,
                hipStream_t s) {
    if (tw == 1) lh2<1>(m, tiles, ntiles, grid, mode, s);
    else if (tw == 2) lh2<2>(m, tiles, ntiles, grid, mode, s);
    else if (tw == 4) lh2<4>(m, tiles, ntiles, grid, mode, s);
    else lh2<8>(m, tiles, ntiles, grid, mode, s);
}

void launch_e2d(const DevModel *m, const Tile *tiles, int ntiles, int grid, int mode, bool integ,
                int tw, hipStream_t s) {
    if (tw == 1) le2<1>(m, tiles, ntiles, grid, mode, integ, s);
    else if (tw == 2) le2<2>(m, tiles, ntiles, grid, mode, integ, s);
    else if (tw == 4) le2<4>(m, tiles, ntiles, grid, mode, integ, s);
    else le2<8>(m, tiles, ntiles, grid, mode, integ, s);
}

}  // namespace mfd
