// 3D Yee H and E half-steps, pipelined single-cell-per-lane form
// (update_H_kernel! / update_E_kernel! 3D, src/models.jl:735-1175).
//
// A tile is 64 x-cells by TY rows by a z range; each lane owns one (x, y) column and
// marches up z.  Every load a plane needs (own fields, the y-neighbour row, the z+1
// plane of the H half, material ids, integral accumulators, and the edge lane's x
// neighbour) is issued one plane ahead of that plane's compute and stores, so two planes
// of loads are in flight per wave while registers stay few enough for a high wave count.
// In-wave x neighbours come through a cross-lane shuffle.  Material coefficients come
// from an LDS table.  All field pointers are accessed in the global address space.
//
// PML = false instantiations serve tiles that no CPML layer reaches: without the six
// slab branches the kernels need about 50 fewer VGPRs (E: 118 -> 69), so a SIMD holds
// more waves and more loads are in flight.
#include "common.hpp"
#include "material.hpp"
#include "tfsf.hpp"

#include <algorithm>

namespace mfd {

typedef __attribute__((address_space(1))) float gf1;
typedef __attribute__((address_space(1))) const float gcf1;
typedef const __attribute__((address_space(1))) unsigned char gcu1;

struct LayP {
    int present, ib, nb;
    const float *Ki, *A, *B;     // 1/K, A, B
    v2f *p;                      // psi pairs
};
__device__ __forceinline__ LayP layp(const DevLayer &l, bool h) {
    LayP r;
    r.present = l.present; r.ib = l.ib; r.nb = l.nb;
    r.Ki = l.Ki; r.A = l.A; r.B = l.B;
    r.p = h ? l.pE : l.pH;
    return r;
}

// one CPML slab update of a cell: psi = B psi + A d for the derivative pair, returns psi
__device__ __forceinline__ void slab2(const LayP &L, int p, unsigned j, float &d0, float &d1,
                                      float &a0, float &a1) {
    const float B = G(L.B)[p], A = G(L.A)[p], Ki = G(L.Ki)[p];
    const v2f q = G(L.p)[j];
    a0 = B * q.x + A * d0;
    a1 = B * q.y + A * d1;
    G(L.p)[j] = mk2(a0, a1);
    d0 = d0 * Ki;      // d / K, with 1/K rounded once on the host
    d1 = d1 * Ki;
}

// the same update with the old psi pair and the profile values already in registers
__device__ __forceinline__ void slab2v(const LayP &L, unsigned j, float psi0, float psi1, float B,
                                       float A, float Ki, float &d0, float &d1, float &a0, float &a1) {
    a0 = B * psi0 + A * d0;
    a1 = B * psi1 + A * d1;
    G(L.p)[j] = mk2(a0, a1);
    d0 = d0 * Ki;
    d1 = d1 * Ki;
}

// ---------------------------------- H half ---------------------------------------------
struct HP {
    float ez, exn, eyn, exy, ezy, hx, hy, hz, eyE, ezE;
    unsigned char g;
};

template <int MODE, int TY, bool PML>
__global__ __launch_bounds__(64 * TY) __attribute__((amdgpu_waves_per_eu(5, 8))) void k_h3d_p(const DevModel *__restrict__ models,
                                                   const Tile *__restrict__ tiles, int ntiles) {
    __shared__ float mh_tab[128];
    const int lane = threadIdx.x, ty = threadIdx.y;
    for (int t = blockIdx.x; t < ntiles; t += gridDim.x) {
        const Tile tl = tiles[t];
        if (tl.x < 0) continue;
        const DevModel &m = models[tl.x];
        if (*m.it >= m.Nt) continue;
        const int nm = MODE == MODE_VAC ? 0 : m.nmat;
        for (int k = lane + 64 * ty; k <= nm; k += 64 * TY) mh_tab[k] = m.mats[k].Mh;
        __syncthreads();
        const int Nx = m.Nx, Ny = m.Ny, Nz = m.Nz;
        const int ix = tl.y + lane, iy = tl.z + ty;
        if (ix < Nx && iy < Ny) {
            const int z0 = tile_z0(tl), z1 = tile_z1(tl);
            gcf1 *Ex = G((const float *)m.F[EX]), *Ey = G((const float *)m.F[EY]);
            gcf1 *Ez = G((const float *)m.F[EZ]);
            gf1 *Hx = G(m.F[HX]), *Hy = G(m.F[HY]), *Hz = G(m.F[HZ]);
            gcu1 *id = G(m.id);
            const unsigned sz = (unsigned)m.sz, col = (unsigned)(ix + m.sy * iy);
            const float idx = m.idx, idy = m.idy, idz = m.idz;
            const int bcx = m.bcs[1], bcy = m.bcs[3], bcz = m.bcs[5];
            const bool yin = iy + 1 < Ny;
            const bool yload = yin || bcy == BC_PERIODIC;
            const unsigned yoff = yin ? (unsigned)m.sy : (unsigned)(-(long long)(Ny - 1) * m.sy);
            const bool edge = lane == 63 || ix + 1 >= Nx;
            const bool eload = edge && (ix + 1 < Nx || bcx == BC_PERIODIC);
            const unsigned xoff = ix + 1 < Nx ? 1u : (unsigned)(1 - Nx);
            LayP L[6];
#pragma unroll
            for (int f = 0; f < 6; f++) {
                L[f] = layp(m.L[f], true);
                if (!PML) L[f].present = 0;
            }
            const bool xs0 = L[0].present && ix <= L[0].ib, xs1 = L[1].present && ix >= L[1].ib;
            const bool ys0 = L[2].present && iy <= L[2].ib, ys1 = L[3].present && iy >= L[3].ib;
            auto load = [&](int iz, HP &P) {
                const unsigned i = col + sz * iz;
                P.ez = Ez[i];
                if (iz + 1 < Nz) { P.exn = Ex[i + sz]; P.eyn = Ey[i + sz]; }
                else if (bcz == BC_PERIODIC) { P.exn = Ex[col]; P.eyn = Ey[col]; }
                if (yload) { P.exy = Ex[i + yoff]; P.ezy = Ez[i + yoff]; }
                P.hx = Hx[i]; P.hy = Hy[i]; P.hz = Hz[i];
                if (eload) { P.eyE = Ey[i + xoff]; P.ezE = Ez[i + xoff]; }
                if (MODE != MODE_VAC) P.g = id[i];
            };
            float ex = Ex[col + sz * z0], ey = Ey[col + sz * z0];
            // x-slab CPML as in the E half: profile values once per tile, each plane's psi
            // pair issued before the next plane's prefetch
            float xKi = 0.0f, xA = 0.0f, xB = 0.0f;
            if (xs0) { xKi = G(L[0].Ki)[ix]; xA = G(L[0].A)[ix]; xB = G(L[0].B)[ix]; }
            else if (xs1) {
                const int p = ix - L[1].ib;
                xKi = G(L[1].Ki)[p]; xA = G(L[1].A)[p]; xB = G(L[1].B)[p];
            }
            struct SP { float x0, x1; };
            auto slab_pre = [&](int iz, SP &S) {
                if (xs0) {
                    const unsigned j = ix + (unsigned)L[0].nb * (iy + (unsigned)Ny * iz);
                    { const v2f q = G((const v2f *)L[0].p)[j]; S.x0 = q.x; S.x1 = q.y; }
                } else if (xs1) {
                    const int p = ix - L[1].ib;
                    const unsigned j = p + (unsigned)L[1].nb * (iy + (unsigned)Ny * iz);
                    { const v2f q = G((const v2f *)L[1].p)[j]; S.x0 = q.x; S.x1 = q.y; }
                }
            };
            auto compute = [&](int iz, HP &P, const SP &S) {
                const unsigned i = col + sz * iz;
                if (iz + 1 >= Nz && bcz != BC_PERIODIC) {
                    P.exn = bcz == BC_DIRICHLET ? 0.0f : ex;
                    P.eyn = bcz == BC_DIRICHLET ? 0.0f : ey;
                }
                if (!yload) {
                    P.exy = bcy == BC_DIRICHLET ? 0.0f : ex;
                    P.ezy = bcy == BC_DIRICHLET ? 0.0f : P.ez;
                }
                float eyx = __shfl_down(ey, 1), ezx = __shfl_down(P.ez, 1);
                if (edge) {
                    if (eload) { eyx = P.eyE; ezx = P.ezE; }
                    else if (bcx == BC_DIRICHLET) { eyx = 0.0f; ezx = 0.0f; }
                    else { eyx = ey; ezx = P.ez; }
                }
                const float Mh = MODE != MODE_VAC ? mh_tab[P.g & ID_MASK] : mh_tab[0];
                float dExy = (P.exy - ex) * idy;
                float dExz = (P.exn - ex) * idz;
                float dEyx = (eyx - ey) * idx;
                float dEyz = (P.eyn - ey) * idz;
                float dEzx = (ezx - P.ez) * idx;
                float dEzy = (P.ezy - P.ez) * idy;
                float hx = P.hx, hy = P.hy, hz = P.hz;
                float a0, a1;
                if (xs0) {   // src/models.jl:805-815
                    slab2v(L[0], ix + (unsigned)L[0].nb * (iy + (unsigned)Ny * iz), S.x0, S.x1, xB, xA, xKi,
                           dEyx, dEzx, a0, a1);
                    hy += Mh * a1; hz -= Mh * a0;
                }
                if (xs1) {   // src/models.jl:816-826
                    const int p = ix - L[1].ib;
                    const unsigned j = p + (unsigned)L[1].nb * (iy + (unsigned)Ny * iz);
                    if (!xs0) slab2v(L[1], j, S.x0, S.x1, xB, xA, xKi, dEyx, dEzx, a0, a1);
                    else slab2(L[1], p, j, dEyx, dEzx, a0, a1);     // both x layers reach the cell
                    hy += Mh * a1; hz -= Mh * a0;
                }
                if (ys0) {   // src/models.jl:827-837
                    slab2(L[2], iy, ix + (unsigned)Nx * (iy + (unsigned)L[2].nb * iz), dExy, dEzy, a0, a1);
                    hx -= Mh * a1; hz += Mh * a0;
                }
                if (ys1) {   // src/models.jl:838-848
                    const int p = iy - L[3].ib;
                    slab2(L[3], p, ix + (unsigned)Nx * (p + (unsigned)L[3].nb * iz), dExy, dEzy, a0, a1);
                    hx -= Mh * a1; hz += Mh * a0;
                }
                if (L[4].present && iz <= L[4].ib) {   // src/models.jl:849-859
                    slab2(L[4], iz, col + sz * iz, dExz, dEyz, a0, a1);
                    hx += Mh * a1; hy -= Mh * a0;
                }
                if (L[5].present && iz >= L[5].ib) {   // src/models.jl:860-871
                    const int p = iz - L[5].ib;
                    slab2(L[5], p, col + sz * p, dExz, dEyz, a0, a1);
                    hx += Mh * a1; hy -= Mh * a0;
                }
                hx -= Mh * (dEzy - dEyz);
                hy -= Mh * (dExz - dEzx);
                hz -= Mh * (dEyx - dExy);
                Hx[i] = hx; Hy[i] = hy; Hz[i] = hz;
                ex = P.exn; ey = P.eyn;
            };
            HP A, B;
            load(z0, A);
            for (int iz = z0; iz < z1; iz += 2) {
                SP S0 = {0.0f, 0.0f};
                slab_pre(iz, S0);
                if (iz + 1 < z1) load(iz + 1, B);
                compute(iz, A, S0);
                if (iz + 1 >= z1) break;
                SP S1 = {0.0f, 0.0f};
                slab_pre(iz + 1, S1);
                if (iz + 2 < z1) load(iz + 2, A);
                compute(iz + 1, B, S1);
            }
        }
        __syncthreads();
    }
}

// ---------------------------------- E half ---------------------------------------------
struct EP {
    float hx, hy, hz, hxy, hzy, d0, d1, d2, sa, e2, hyE, hzE;
    unsigned char g;
};

template <int MODE, bool INTEG, int TY, bool PML>
__global__ __launch_bounds__(64 * TY) void k_e3d_p(const DevModel *__restrict__ models,
                                                   const Tile *__restrict__ tiles, int ntiles) {
    // (Md1, Md2, Me, kind): kind 0 plain, 1 Kerr only, 2 ADE/plasma state (E of such
    // cells is finished by the state pass, k_e_state)
    __shared__ float4 ctab[128];
    __shared__ float2 ktab[128];    // Kerr Pade (K2, K3)
    __shared__ float cw[128];       // Me*Md2 (the E-form update without D), full mode
    // incident values of the TFSF x faces for this tile's rows and planes:
    // [side][H_u, H_v, E_u, E_v][row][plane - z0]
    __shared__ float xinc[2][4][TY][XFACE_ZC];
    const int lane = threadIdx.x, ty = threadIdx.y;
    for (int t = blockIdx.x; t < ntiles; t += gridDim.x) {
        const Tile tl = tiles[t];
        if (tl.x < 0) continue;
        const DevModel &m = models[tl.x];
        if (*m.it >= m.Nt) continue;
        const int nm = MODE == MODE_VAC ? 0 : m.nmat;
        for (int k = lane + 64 * ty; k <= nm; k += 64 * TY) {
            const DevMat &mt = m.mats[k];
            const float kind = (mt.flags & (MAT_DISP | MAT_PLASMA)) ? 2.0f : (mt.flags & MAT_KERR) ? 1.0f : 0.0f;
            // w: the material kind with ADE/Kerr/plasma, else Me*Md2 for the runs without D
            ctab[k] = make_float4(mt.Md1, mt.Md2, mt.Me, MODE == MODE_FULL ? kind : mt.MeMd2);
            if (MODE == MODE_FULL) { ktab[k] = make_float2(mt.K2, mt.K3); cw[k] = mt.MeMd2; }
        }
        // TFSF x faces finished in this kernel: the source's scalars come into registers
        // once per tile and the incident values of the tile's x-face cells into LDS, so
        // the z march issues no dependent or flat load for them
        int xb[6] = {0, -2, 0, -1, 0, -1};
        float xcm = 0.0f, xce = 0.0f, xcd = 0.0f;
        bool xl[4] = {false, false, false, false};
        bool xtile = false;
        if (m.xtfsf) {
            const __attribute__((address_space(1))) DevSource &S =
                *(const __attribute__((address_space(1))) DevSource *)m.xtfsf;
#pragma unroll
            for (int q = 0; q < 6; q++) xb[q] = S.box[q];
            xcm = S.cm[0]; xce = S.ce[0]; xcd = S.cd[0];
#pragma unroll
            for (int kk = 0; kk < 4; kk++) xl[kk] = tfsf_live(S, 0, kk);
            const int x0 = tl.y, zt0 = tile_z0(tl), nzt = tile_z1(tl) - zt0;
            xtile = ((xb[0] - 1 >= x0 && xb[0] - 1 < x0 + 64) || (xb[0] >= x0 && xb[0] < x0 + 64) ||
                     (xb[1] >= x0 && xb[1] < x0 + 64)) &&
                    tl.z <= xb[3] && tl.z + TY - 1 >= xb[2] && zt0 <= xb[5] && zt0 + nzt - 1 >= xb[4];
            if (xtile) {
                const int it0 = *m.it, Nzm = m.Nz;
                for (int k = lane + 64 * ty; k < 8 * TY * nzt; k += 64 * TY) {
                    const int zz = k % nzt, r = (k / nzt) % TY, kk = (k / (nzt * TY)) % 4, side = k / (nzt * TY * 4);
                    const int u = tl.z + r, v = zt0 + zz;
                    const bool in = (kk == 0 || kk == 3) ? (u >= xb[2] && u <= xb[3] && v >= xb[4] && v <= xb[5] - 1)
                                                         : (u >= xb[2] && u <= xb[3] - 1 && v >= xb[4] && v <= xb[5]);
                    xinc[side][kk][r][zz] = in && xl[kk] ? tfsf_inc(S, Nzm, 0, side, kk, u, v, it0) : 0.0f;
                }
            }
        }
        __syncthreads();
        const int Nx = m.Nx, Ny = m.Ny, Nz = m.Nz;
        const int ix = tl.y + lane, iy = tl.z + ty;
        if (ix < Nx && iy < Ny) {
            const int z0 = tile_z0(tl), z1 = tile_z1(tl);
            gcf1 *Hx = G((const float *)m.F[HX]), *Hy = G((const float *)m.F[HY]);
            gcf1 *Hz = G((const float *)m.F[HZ]);
            gf1 *Dx = G(m.F[DX]), *Dy = G(m.F[DY]), *Dz = G(m.F[DZ]);
            gf1 *Ex = G(m.F[EX]), *Ey = G(m.F[EY]), *Ez = G(m.F[EZ]);
            gf1 *Sa = G(m.Sa), *E2 = G(m.E2);
            const bool doSa = INTEG && m.Sa, doE2 = INTEG && m.E2;
            gcu1 *id = G(m.id);
            const unsigned sz = (unsigned)m.sz, col = (unsigned)(ix + m.sy * iy);
            const float idx = m.idx, idy = m.idy, idz = m.idz, dtf = m.dtf;
            const int bcx = m.bcs[0], bcy = m.bcs[2], bcz = m.bcs[4];
            const bool yin = iy > 0;
            const bool yload = yin || bcy == BC_PERIODIC;
            const unsigned yoff = yin ? (unsigned)(-m.sy) : (unsigned)((Ny - 1) * m.sy);
            const bool edge = lane == 0 || ix == 0;
            const bool eload = edge && (ix > 0 || bcx == BC_PERIODIC);
            const unsigned xoff = ix > 0 ? (unsigned)-1 : (unsigned)(Nx - 1);
            LayP L[6];
#pragma unroll
            for (int f = 0; f < 6; f++) {
                L[f] = layp(m.L[f], false);
                if (!PML) L[f].present = 0;
            }
            const bool xs0 = L[0].present && ix <= L[0].ib, xs1 = L[1].present && ix >= L[1].ib;
            const bool ys0 = L[2].present && iy <= L[2].ib, ys1 = L[3].present && iy >= L[3].ib;
            // lanes on a TFSF x-face column (planes N1-1, N1, N2 of the x faces)
            const bool xcol = xtile && (ix == xb[0] - 1 || ix == xb[0] || ix == xb[1]) && iy >= xb[2] &&
                              iy <= xb[3];
            const int xside = ix == xb[1] ? 1 : 0;
            // without stored D, a plain cell reads E and takes the update in E form,
            // E' = Md1 (E + Me Md2 psi) + Me Md2 curl H, which is Me times the D form.
            // Cells that keep their D: the TFSF face cells, whose D the face corrections
            // change apart from E (known from the plane, so D is loaded ahead), and Kerr,
            // ADE, plasma and point-source cells (known from the id: D loaded on use)
            const bool dless = m.dless;
            auto dcell = [&](int iz) { return !dless || tfsf_touched_box(xb, ix, iy, iz); };
            gcf1 *Dsx = G((const float *)m.F[DX]), *Dsy = G((const float *)m.F[DY]);
            gcf1 *Dsz = G((const float *)m.F[DZ]);
            gcf1 *Esx = G((const float *)m.F[EX]), *Esy = G((const float *)m.F[EY]);
            gcf1 *Esz = G((const float *)m.F[EZ]);
            auto load = [&](int iz, EP &P) {
                const unsigned i = col + sz * iz;
                P.hx = Hx[i]; P.hy = Hy[i]; P.hz = Hz[i];
                if (yload) { P.hxy = Hx[i + yoff]; P.hzy = Hz[i + yoff]; }
                const bool dc = dcell(iz);
                P.d0 = (dc ? Dsx : Esx)[i]; P.d1 = (dc ? Dsy : Esy)[i]; P.d2 = (dc ? Dsz : Esz)[i];
                if (eload) { P.hyE = Hy[i + xoff]; P.hzE = Hz[i + xoff]; }
                if (MODE != MODE_VAC || INTEG || dless) P.g = id[i];
                // the integrals are read and written once per step and not touched by the
                // H half: streamed past the caches (nontemporal), leaving them to H and E
                if (doSa) P.sa = __builtin_nontemporal_load(&Sa[i]);
                if (doE2) P.e2 = __builtin_nontemporal_load(&E2[i]);
            };
            float hxm = 0.0f, hym = 0.0f;     // H(z-1)
            bool zm_own = false;
            int nonfin = 0;                   // 1 NaN, 2 Inf seen in this lane's Ex
            if (z0 > 0) { hxm = Hx[col + sz * (z0 - 1)]; hym = Hy[col + sz * (z0 - 1)]; }
            else if (bcz == BC_PERIODIC) { hxm = Hx[col + sz * (Nz - 1)]; hym = Hy[col + sz * (Nz - 1)]; }
            else if (bcz == BC_NEUMANN) zm_own = true;
            // x- and y-slab CPML: the profile values depend only on the lane (x) or the row
            // (y) and are loaded once.  Each plane's x and y psi pairs are issued before the
            // next plane's prefetch (slab_pre in the march below): vector loads complete in
            // order, so a wait for a psi value issued after the prefetch would drain it.
            float xKi = 0.0f, xA = 0.0f, xB = 0.0f, yKi = 0.0f, yA = 0.0f, yB = 0.0f;
            if (xs0) { xKi = G(L[0].Ki)[ix]; xA = G(L[0].A)[ix]; xB = G(L[0].B)[ix]; }
            else if (xs1) {
                const int p = ix - L[1].ib;
                xKi = G(L[1].Ki)[p]; xA = G(L[1].A)[p]; xB = G(L[1].B)[p];
            }
            if (ys0) { yKi = G(L[2].Ki)[iy]; yA = G(L[2].A)[iy]; yB = G(L[2].B)[iy]; }
            else if (ys1) {
                const int p = iy - L[3].ib;
                yKi = G(L[3].Ki)[p]; yA = G(L[3].A)[p]; yB = G(L[3].B)[p];
            }
            struct SP { float x0, x1, y0, y1, z0, z1, zKi, zA, zB; };
            auto zs0 = [&](int iz) { return L[4].present && iz <= L[4].ib; };
            auto zs1 = [&](int iz) { return L[5].present && iz >= L[5].ib; };
            auto slab_pre = [&](int iz, SP &S) {
                // z slabs: the plane's psi pair and its profile values (uniform per plane)
                if (zs0(iz)) {
                    const unsigned j = col + sz * iz;
                    { const v2f q = G((const v2f *)L[4].p)[j]; S.z0 = q.x; S.z1 = q.y; }
                    S.zKi = G(L[4].Ki)[iz]; S.zA = G(L[4].A)[iz]; S.zB = G(L[4].B)[iz];
                } else if (zs1(iz)) {
                    const int p = iz - L[5].ib;
                    const unsigned j = col + sz * p;
                    { const v2f q = G((const v2f *)L[5].p)[j]; S.z0 = q.x; S.z1 = q.y; }
                    S.zKi = G(L[5].Ki)[p]; S.zA = G(L[5].A)[p]; S.zB = G(L[5].B)[p];
                }
                if (xs0) {
                    const unsigned j = ix + (unsigned)L[0].nb * (iy + (unsigned)Ny * iz);
                    { const v2f q = G((const v2f *)L[0].p)[j]; S.x0 = q.x; S.x1 = q.y; }
                } else if (xs1) {
                    const int p = ix - L[1].ib;
                    const unsigned j = p + (unsigned)L[1].nb * (iy + (unsigned)Ny * iz);
                    { const v2f q = G((const v2f *)L[1].p)[j]; S.x0 = q.x; S.x1 = q.y; }
                }
                if (ys0) {
                    const unsigned j = ix + (unsigned)Nx * (iy + (unsigned)L[2].nb * iz);
                    { const v2f q = G((const v2f *)L[2].p)[j]; S.y0 = q.x; S.y1 = q.y; }
                } else if (ys1) {
                    const int p = iy - L[3].ib;
                    const unsigned j = ix + (unsigned)Nx * (p + (unsigned)L[3].nb * iz);
                    { const v2f q = G((const v2f *)L[3].p)[j]; S.y0 = q.x; S.y1 = q.y; }
                }
            };
            auto compute = [&](int iz, EP &P, const SP &S) {
                const unsigned i = col + sz * iz;
                if (zm_own) { hxm = P.hx; hym = P.hy; zm_own = false; }
                if (!yload) {
                    P.hxy = bcy == BC_DIRICHLET ? 0.0f : P.hx;
                    P.hzy = bcy == BC_DIRICHLET ? 0.0f : P.hz;
                }
                float hyx = __shfl_up(P.hy, 1), hzx = __shfl_up(P.hz, 1);
                if (edge) {
                    if (eload) { hyx = P.hyE; hzx = P.hzE; }
                    else if (bcx == BC_DIRICHLET) { hyx = 0.0f; hzx = 0.0f; }
                    else { hyx = P.hy; hzx = P.hz; }
                }
                const int g = P.g & ID_MASK;
                const float4 cf = MODE != MODE_VAC ? ctab[g] : ctab[0];
                bool dc = dcell(iz);
                if (dless && !dc && ((P.g & ID_TOUCHED) || (MODE == MODE_FULL && cf.w != 0.0f))) {
                    dc = true;
                    P.d0 = Dsx[i]; P.d1 = Dsy[i]; P.d2 = Dsz[i];
                }
                const float Md1 = cf.x, Md2 = dc ? cf.y : (MODE == MODE_FULL ? cw[g] : cf.w);
                float dHxy = (P.hx - P.hxy) * idy;
                float dHxz = (P.hx - hxm) * idz;
                float dHyx = (P.hy - hyx) * idx;
                float dHyz = (P.hy - hym) * idz;
                float dHzx = (P.hz - hzx) * idx;
                float dHzy = (P.hz - P.hzy) * idy;
                float d0 = P.d0, d1 = P.d1, d2 = P.d2;
                float a0, a1;
                if (xs0) {   // src/models.jl:976-986
                    slab2v(L[0], ix + (unsigned)L[0].nb * (iy + (unsigned)Ny * iz), S.x0, S.x1, xB, xA, xKi,
                           dHyx, dHzx, a0, a1);
                    d1 -= Md2 * a1; d2 += Md2 * a0;
                }
                if (xs1) {   // src/models.jl:987-997
                    const int p = ix - L[1].ib;
                    const unsigned j = p + (unsigned)L[1].nb * (iy + (unsigned)Ny * iz);
                    if (!xs0) slab2v(L[1], j, S.x0, S.x1, xB, xA, xKi, dHyx, dHzx, a0, a1);
                    else slab2(L[1], p, j, dHyx, dHzx, a0, a1);     // both x layers reach the cell
                    d1 -= Md2 * a1; d2 += Md2 * a0;
                }
                if (ys0) {   // src/models.jl:998-1006
                    slab2v(L[2], ix + (unsigned)Nx * (iy + (unsigned)L[2].nb * iz), S.y0, S.y1, yB, yA, yKi,
                           dHxy, dHzy, a0, a1);
                    d0 += Md2 * a1; d2 -= Md2 * a0;
                }
                if (ys1) {   // src/models.jl:1007-1015
                    const int p = iy - L[3].ib;
                    const unsigned j = ix + (unsigned)Nx * (p + (unsigned)L[3].nb * iz);
                    if (!ys0) slab2v(L[3], j, S.y0, S.y1, yB, yA, yKi, dHxy, dHzy, a0, a1);
                    else slab2(L[3], p, j, dHxy, dHzy, a0, a1);     // both y layers reach the cell
                    d0 += Md2 * a1; d2 -= Md2 * a0;
                }
                if (zs0(iz)) {   // src/models.jl:1016-1025
                    slab2v(L[4], col + sz * iz, S.z0, S.z1, S.zB, S.zA, S.zKi, dHxz, dHyz, a0, a1);
                    d0 -= Md2 * a1; d1 += Md2 * a0;
                }
                if (zs1(iz)) {   // src/models.jl:1026-1035
                    const int p = iz - L[5].ib;
                    if (!zs0(iz)) slab2v(L[5], col + sz * p, S.z0, S.z1, S.zB, S.zA, S.zKi, dHxz, dHyz, a0, a1);
                    else slab2(L[5], p, col + sz * p, dHxz, dHyz, a0, a1);     // both z layers reach the cell
                    d0 -= Md2 * a1; d1 += Md2 * a0;
                }
                d0 = Md1 * d0 + Md2 * (dHzy - dHyz);
                d1 = Md1 * d1 + Md2 * (dHxz - dHzx);
                d2 = Md1 * d2 + Md2 * (dHyx - dHxy);
                float e[3];
                if (dc) { e[0] = d0 * cf.z; e[1] = d1 * cf.z; e[2] = d2 * cf.z; }
                else { e[0] = d0; e[1] = d1; e[2] = d2; }       // already E
                bool defer = false;
                if (MODE == MODE_FULL && cf.w != 0.0f) {
                    if (cf.w > 1.5f) {
                        defer = true;
                    } else {   // Kerr Pade, src/models.jl:1140-1149
                        const float2 kk = ktab[g];
                        const float dd[3] = {d0, d1, d2};
#pragma unroll
                        for (int c = 0; c < 3; c++) {
                            const float u = dd[c];
                            e[c] = (1.0f + kk.x * u + 2.0f * kk.y * (u * u)) /
                                   (1.0f + 2.0f * kk.x * u + 3.0f * kk.y * (u * u)) * u * cf.z;
                        }
                    }
                }
                // x-face TFSF epilogue: the corrections add_source!(::TFSFSource3D) makes on
                // an x face (src/sources.jl:725-745) after this update, in face order; the
                // H correction of the H-out cells enters only the integrals here (the face
                // kernel writes it, after every neighbour has read the uncorrected H)
                float hyI = P.hy, hzI = P.hz;
                bool xh = false;
                if (xcol && iz >= xb[4] && iz <= xb[5]) {
                    xh = tfsf_xhandled_box(xb, ix, iy, iz, defer);
                    if (xh) {
                        const int u = iy, v = iz, side = xside, zz = iz - z0;
                        const int U2 = xb[3], V2 = xb[5];
                        const int nH = side ? xb[1] : xb[0] - 1, nE = side ? xb[1] : xb[0];
                        const float s0 = side ? 1.0f : -1.0f;
                        if (ix == nH) {
                            if (u <= U2 && v <= V2 - 1 && xl[3])
                                hyI = __builtin_fmaf(s0 * xcm, xinc[side][3][ty][zz], hyI);
                            if (u <= U2 - 1 && v <= V2 && xl[2])
                                hzI = __builtin_fmaf(-s0 * xcm, xinc[side][2][ty][zz], hzI);
                        }
                        if (ix == nE) {
                            if (u <= U2 - 1 && v <= V2 && xl[1]) {
                                const float iv = xinc[side][1][ty][zz];
                                e[1] = __builtin_fmaf(-s0 * xce, iv, e[1]);
                                d1 = __builtin_fmaf(-s0 * xcd, iv, d1);
                            }
                            if (u <= U2 && v <= V2 - 1 && xl[0]) {
                                const float iv = xinc[side][0][ty][zz];
                                e[2] = __builtin_fmaf(s0 * xce, iv, e[2]);
                                d2 = __builtin_fmaf(s0 * xcd, iv, d2);
                            }
                        }
                    }
                }
                if (dc) { Dx[i] = d0; Dy[i] = d1; Dz[i] = d2; }
                if (!defer) { Ex[i] = e[0]; Ey[i] = e[1]; Ez[i] = e[2]; }
                // non-finite guard on Ex: collected per lane, reported once per tile
                if (!defer && !isfinite(e[0])) nonfin |= isnan(e[0]) ? 1 : 2;
                if (INTEG && !defer && (!(P.g & ID_TOUCHED) || xh)) {   // src/outputs.jl:480-483
                    if (doSa) {
                        const float a = e[1] * hzI - e[2] * hyI, b = e[2] * P.hx - e[0] * hzI;
                        const float cc = e[0] * hyI - e[1] * P.hx;
                        __builtin_nontemporal_store(P.sa + sa_sqrt(a * a + b * b + cc * cc) * dtf, &Sa[i]);
                    }
                    if (doE2) __builtin_nontemporal_store(P.e2 + (e[0] * e[0] + e[1] * e[1] + e[2] * e[2]) * dtf, &E2[i]);
                }
                hxm = P.hx; hym = P.hy;
            };
            EP A, B;
            load(z0, A);
            for (int iz = z0; iz < z1; iz += 2) {
                SP S0 = {0.0f, 0.0f, 0.0f, 0.0f, 0.0f, 0.0f, 0.0f, 0.0f, 0.0f};
                slab_pre(iz, S0);
                if (iz + 1 < z1) load(iz + 1, B);
                compute(iz, A, S0);
                if (iz + 1 >= z1) break;
                SP S1 = {0.0f, 0.0f, 0.0f, 0.0f, 0.0f, 0.0f, 0.0f, 0.0f, 0.0f};
                slab_pre(iz + 1, S1);
                if (iz + 2 < z1) load(iz + 2, A);
                compute(iz + 1, B, S1);
            }
            if (__any(nonfin != 0) && nonfin) atomicOr(m.flag, nonfin);
        }
        __syncthreads();
    }
}

template <int TY, bool PML>
static void lh(const DevModel *m, const Tile *tiles, int ntiles, int grid, int mode, hipStream_t s) {
    dim3 block(64, TY);
    if (mode == MODE_VAC) hipLaunchKernelGGL((k_h3d_p<MODE_VAC, TY, PML>), dim3(grid), block, 0, s, m, tiles, ntiles);
    else hipLaunchKernelGGL((k_h3d_p<MODE_SIMPLE, TY, PML>), dim3(grid), block, 0, s, m, tiles, ntiles);
}

template <int TY, bool PML>
static void le(const DevModel *m, const Tile *tiles, int ntiles, int grid, int mode, bool integ,
               hipStream_t s) {
    dim3 block(64, TY);
#define EP_(MD, IN) hipLaunchKernelGGL((k_e3d_p<MD, IN, TY, PML>), dim3(grid), block, 0, s, m, tiles, ntiles)
    if (mode == MODE_VAC) { if (integ) EP_(MODE_VAC, true); else EP_(MODE_VAC, false); }
    else if (mode == MODE_SIMPLE) { if (integ) EP_(MODE_SIMPLE, true); else EP_(MODE_SIMPLE, false); }
    else { if (integ) EP_(MODE_FULL, true); else EP_(MODE_FULL, false); }
#undef EP_
}

void launch_h3d_p(const DevModel *m, const Tile *tiles, int ntiles, int grid, int mode, int ty,
                  bool pml, hipStream_t s) {
    if (ntiles <= 0) return;
#define LH_(T) (pml ? lh<T, true>(m, tiles, ntiles, grid, mode, s) : lh<T, false>(m, tiles, ntiles, grid, mode, s))
    if (ty == 1) LH_(1);
    else if (ty == 2) LH_(2);
    else if (ty == 4) LH_(4);
    else LH_(8);
#undef LH_
}

// D = E/Me (the E kernel's own rebuild) for every plain cell of a model that ran
// without stored D, so that D reads back and later runs start from a full state.
__global__ __launch_bounds__(256) void k_d_from_e(const DevModel m) {
    int xb[6] = {0, -2, 0, -1, 0, -1};
    if (m.xtfsf)
        for (int q = 0; q < 6; q++) xb[q] = m.xtfsf->box[q];
    const unsigned N = (unsigned)m.N;
    for (unsigned i = blockIdx.x * blockDim.x + threadIdx.x; i < N; i += gridDim.x * blockDim.x) {
        const int ix = (int)(i % (unsigned)m.Nx), iy = (int)((i / (unsigned)m.Nx) % (unsigned)m.Ny);
        const int iz = (int)(i / ((unsigned)m.Nx * (unsigned)m.Ny));
        if (tfsf_touched_box(xb, ix, iy, iz)) continue;
        const unsigned char g = G(m.id)[i];
        if (g & ID_TOUCHED) continue;
        const DevMat &mt = m.mats[g & ID_MASK];
        if (mt.flags & (MAT_DISP | MAT_PLASMA | MAT_KERR)) continue;
        const float iMe = mt.iMe;
        G(m.F[DX])[i] = G((const float *)m.F[EX])[i] * iMe;
        G(m.F[DY])[i] = G((const float *)m.F[EY])[i] * iMe;
        G(m.F[DZ])[i] = G((const float *)m.F[EZ])[i] * iMe;
    }
}

// 2D: the cells that keep D are those with Kerr, ADE or plasma state and those a source
// writes (the id's touched bit); every other cell gets D = E/Me
__global__ __launch_bounds__(256) void k_d_from_e_2d(const DevModel m) {
    const unsigned N = (unsigned)m.N;
    for (unsigned i = blockIdx.x * blockDim.x + threadIdx.x; i < N; i += gridDim.x * blockDim.x) {
        const unsigned char g = G(m.id)[i];
        if (g & ID_TOUCHED) continue;
        const DevMat &mt = m.mats[g & ID_MASK];
        if (mt.flags & (MAT_DISP | MAT_PLASMA | MAT_KERR)) continue;
        const float iMe = mt.iMe;
        G(m.F[DX])[i] = G((const float *)m.F[EX])[i] * iMe;
        G(m.F[DZ])[i] = G((const float *)m.F[EZ])[i] * iMe;
    }
}

void launch_d_from_e(const DevModel &m, hipStream_t s) {
    const long long nb = std::min<long long>((m.N + 255) / 256, 8192);
    if (m.ndims == 2) hipLaunchKernelGGL(k_d_from_e_2d, dim3((unsigned)nb), dim3(256), 0, s, m);
    else hipLaunchKernelGGL(k_d_from_e, dim3((unsigned)nb), dim3(256), 0, s, m);
}

void launch_e3d_p(const DevModel *m, const Tile *tiles, int ntiles, int grid, int mode, bool integ,
                  int ty, bool pml, hipStream_t s) {
    if (ntiles <= 0) return;
#define LE_(T) (pml ? le<T, true>(m, tiles, ntiles, grid, mode, integ, s) \
                    : le<T, false>(m, tiles, ntiles, grid, mode, integ, s))
    if (ty == 1) LE_(1);
    else if (ty == 2) LE_(2);
    else if (ty == 4) LE_(4);
    else LE_(8);
#undef LE_
}

}  // namespace mfd
