// 3D Yee H and E half-steps, vector form for grids with Nx % 4 == 0
// (update_H_kernel! / update_E_kernel! 3D, src/models.jl:735-1175).
//
// Work decomposition: a tile is 256 x-cells by TY rows by a z range.  Each wave owns one
// row segment of 256 cells, 4 consecutive cells per lane, so every field access is one
// 16-byte load or store per lane (1 KiB per wave instruction).  The x neighbour that
// lies in the next (H half) or previous (E half) lane comes through a cross-lane
// shuffle; only the wave's edge lane reads it from memory.  The y neighbour row is read
// from L2 (the row belongs to the neighbouring wave or tile on the same XCD), and the z
// neighbour plane stays in registers while the column marches up z.  Per-material
// coefficients sit in an LDS table; CPML slab work runs in branches only slab cells take,
// with y and z slab psi accessed as float4.  Arithmetic per cell is the same statement
// sequence as the scalar kernels.
#include "common.hpp"
#include "material.hpp"

namespace mfd {

typedef __attribute__((address_space(1))) float gf;
typedef __attribute__((address_space(1))) const float gcf;
typedef float v4f __attribute__((ext_vector_type(4)));
typedef unsigned char v4u8 __attribute__((ext_vector_type(4)));

__device__ __forceinline__ void ld4(gcf *p, long long i, float (&v)[4]) {
    const v4f q = *(const __attribute__((address_space(1))) v4f *)(p + i);
    v[0] = q.x; v[1] = q.y; v[2] = q.z; v[3] = q.w;
}
__device__ __forceinline__ void st4(gf *p, long long i, const float (&v)[4]) {
    v4f q;
    q.x = v[0]; q.y = v[1]; q.z = v[2]; q.w = v[3];
    *(__attribute__((address_space(1))) v4f *)(p + i) = q;
}
__device__ __forceinline__ void ldid(const __attribute__((address_space(1))) unsigned char *id,
                                     long long i, unsigned char (&g)[4]) {
    const v4u8 q = *(const __attribute__((address_space(1))) v4u8 *)(id + i);
    g[0] = q.x; g[1] = q.y; g[2] = q.z; g[3] = q.w;
}
__device__ __forceinline__ void cp4(const float (&a)[4], float (&b)[4]) {
#pragma unroll
    for (int k = 0; k < 4; k++) b[k] = a[k];
}
__device__ __forceinline__ void zero4(float (&b)[4]) {
#pragma unroll
    for (int k = 0; k < 4; k++) b[k] = 0.0f;
}

// the CPML layer fields a tile needs, hoisted into registers before any store
struct Lay {
    int present, ib, nb;
    const float *K, *A, *B;
    float *p0, *p1;
};
__device__ __forceinline__ Lay lay(const DevLayer &l, bool h) {
    Lay r;
    r.present = l.present; r.ib = l.ib; r.nb = l.nb;
    r.K = l.K; r.A = l.A; r.B = l.B;
    r.p0 = h ? l.psiE0 : l.psiH0;
    r.p1 = h ? l.psiE1 : l.psiH1;
    return r;
}

// psi = B psi + A d for one x-slab cell (scalar), returns the new psi values
__device__ __forceinline__ void psi_x(const Lay &L, int p, long long j, float d0, float d1,
                                      float &a0, float &a1) {
    const float B = G(L.B)[p], A = G(L.A)[p];
    a0 = B * G(L.p0)[j] + A * d0;
    a1 = B * G(L.p1)[j] + A * d1;
    G(L.p0)[j] = a0;
    G(L.p1)[j] = a1;
}

// ================================ H half-step ========================================
// Everything one plane needs from memory; loaded one plane ahead of the compute.
struct HPlane {
    float ez[4], exn[4], eyn[4], exy[4], ezy[4], hx[4], hy[4], hz[4];
    float eyE, ezE;            // x+1 neighbour of the lane's last cell (edge lane only)
    unsigned char g[4];
};

struct HCtx {
    gcf *Ex, *Ey, *Ez;
    gf *Hx, *Hy, *Hz;
    const __attribute__((address_space(1))) unsigned char *id;
    long long col, sz, yoff, xoff;
    int Nz, bcz;
    bool yload, edge, eload;
};

template <int MODE>
__device__ __forceinline__ void h_load(const HCtx &c, int iz, HPlane &P) {
    const long long i = c.col + c.sz * iz;
    ld4(c.Ez, i, P.ez);
    if (iz + 1 < c.Nz) { ld4(c.Ex, i + c.sz, P.exn); ld4(c.Ey, i + c.sz, P.eyn); }
    else if (c.bcz == BC_PERIODIC) { ld4(c.Ex, c.col, P.exn); ld4(c.Ey, c.col, P.eyn); }
    if (c.yload) { ld4(c.Ex, i + c.yoff, P.exy); ld4(c.Ez, i + c.yoff, P.ezy); }
    ld4(c.Hx, i, P.hx);
    ld4(c.Hy, i, P.hy);
    ld4(c.Hz, i, P.hz);
    if (c.eload) { P.eyE = c.Ey[i + c.xoff]; P.ezE = c.Ez[i + c.xoff]; }
    if (MODE != MODE_VAC) ldid(c.id, i, P.g);
}

template <int MODE>
__device__ __forceinline__ void h_compute(const DevModel &m, const HCtx &c, int iz, HPlane &P,
                                          float (&ex)[4], float (&ey)[4], int xi, int iy,
                                          const float *mh_tab, const Lay (&L)[6], bool xs0,
                                          bool xs1, bool ys0, bool ys1) {
    const int Ny = m.Ny, Nx = m.Nx;
    const long long i = c.col + c.sz * iz;
    const float idx = m.idx, idy = m.idy, idz = m.idz;
    if (iz + 1 >= c.Nz && c.bcz != BC_PERIODIC) {
        if (c.bcz == BC_DIRICHLET) { zero4(P.exn); zero4(P.eyn); }
        else { cp4(ex, P.exn); cp4(ey, P.eyn); }
    }
    if (!c.yload) {
        if (m.bcs[3] == BC_DIRICHLET) { zero4(P.exy); zero4(P.ezy); }
        else { cp4(ex, P.exy); cp4(P.ez, P.ezy); }
    }
    float ey_n = __shfl_down(ey[0], 1), ez_n = __shfl_down(P.ez[0], 1);
    if (c.edge) {
        if (c.eload) { ey_n = P.eyE; ez_n = P.ezE; }
        else if (m.bcs[1] == BC_DIRICHLET) { ey_n = 0.0f; ez_n = 0.0f; }
        else { ey_n = ey[3]; ez_n = P.ez[3]; }
    }
    float Mh[4];
#pragma unroll
    for (int k = 0; k < 4; k++) Mh[k] = MODE != MODE_VAC ? mh_tab[P.g[k] & ID_MASK] : mh_tab[0];
    float dExy[4], dExz[4], dEyx[4], dEyz[4], dEzx[4], dEzy[4];
    float (&hx)[4] = P.hx, (&hy)[4] = P.hy, (&hz)[4] = P.hz;
#pragma unroll
    for (int k = 0; k < 4; k++) {
        const float eyx1 = k < 3 ? ey[k + 1] : ey_n;
        const float ezx1 = k < 3 ? P.ez[k + 1] : ez_n;
        dExy[k] = (P.exy[k] - ex[k]) * idy;
        dExz[k] = (P.exn[k] - ex[k]) * idz;
        dEyx[k] = (eyx1 - ey[k]) * idx;
        dEyz[k] = (P.eyn[k] - ey[k]) * idz;
        dEzx[k] = (ezx1 - P.ez[k]) * idx;
        dEzy[k] = (P.ezy[k] - P.ez[k]) * idy;
    }
    // CPML, x slabs (per cell), src/models.jl:805-826
    if (xs0) {
#pragma unroll
        for (int k = 0; k < 4; k++) {
            const int ix = xi + k;
            if (ix <= L[0].ib) {
                const long long j = ix + (long long)L[0].nb * (iy + (long long)Ny * iz);
                float a0, a1;
                psi_x(L[0], ix, j, dEyx[k], dEzx[k], a0, a1);
                hy[k] += Mh[k] * a1; hz[k] -= Mh[k] * a0;
                const float K = G(L[0].K)[ix];
                dEyx[k] = dEyx[k] / K; dEzx[k] = dEzx[k] / K;
            }
        }
    }
    if (xs1) {
#pragma unroll
        for (int k = 0; k < 4; k++) {
            const int ix = xi + k;
            if (ix >= L[1].ib) {
                const int p = ix - L[1].ib;
                const long long j = p + (long long)L[1].nb * (iy + (long long)Ny * iz);
                float a0, a1;
                psi_x(L[1], p, j, dEyx[k], dEzx[k], a0, a1);
                hy[k] += Mh[k] * a1; hz[k] -= Mh[k] * a0;
                const float K = G(L[1].K)[p];
                dEyx[k] = dEyx[k] / K; dEzx[k] = dEzx[k] / K;
            }
        }
    }
    // y slabs (whole row), src/models.jl:827-848
    for (int side = 0; side < 2; side++) {
        if (side == 0 ? !ys0 : !ys1) continue;
        const Lay &Ls = L[2 + side];
        const int p = side == 0 ? iy : iy - Ls.ib;
        const long long j = xi + (long long)Nx * (p + (long long)Ls.nb * iz);
        const float B = G(Ls.B)[p], A = G(Ls.A)[p], K = G(Ls.K)[p];
        float a0[4], a1[4];
        ld4(G((const float *)Ls.p0), j, a0);
        ld4(G((const float *)Ls.p1), j, a1);
#pragma unroll
        for (int k = 0; k < 4; k++) {
            a0[k] = B * a0[k] + A * dExy[k];
            a1[k] = B * a1[k] + A * dEzy[k];
            hx[k] -= Mh[k] * a1[k]; hz[k] += Mh[k] * a0[k];
            dExy[k] = dExy[k] / K; dEzy[k] = dEzy[k] / K;
        }
        st4(G(Ls.p0), j, a0);
        st4(G(Ls.p1), j, a1);
    }
    // z slabs (whole plane), src/models.jl:849-871
    for (int side = 0; side < 2; side++) {
        const Lay &Ls = L[4 + side];
        if (!Ls.present) continue;
        if (side == 0 ? iz > Ls.ib : iz < Ls.ib) continue;
        const int p = side == 0 ? iz : iz - Ls.ib;
        const long long j = c.col + c.sz * p;
        const float B = G(Ls.B)[p], A = G(Ls.A)[p], K = G(Ls.K)[p];
        float a0[4], a1[4];
        ld4(G((const float *)Ls.p0), j, a0);
        ld4(G((const float *)Ls.p1), j, a1);
#pragma unroll
        for (int k = 0; k < 4; k++) {
            a0[k] = B * a0[k] + A * dExz[k];
            a1[k] = B * a1[k] + A * dEyz[k];
            hx[k] += Mh[k] * a1[k]; hy[k] -= Mh[k] * a0[k];
            dExz[k] = dExz[k] / K; dEyz[k] = dEyz[k] / K;
        }
        st4(G(Ls.p0), j, a0);
        st4(G(Ls.p1), j, a1);
    }
#pragma unroll
    for (int k = 0; k < 4; k++) {
        hx[k] -= Mh[k] * (dEzy[k] - dEyz[k]);
        hy[k] -= Mh[k] * (dExz[k] - dEzx[k]);
        hz[k] -= Mh[k] * (dEyx[k] - dExy[k]);
    }
    st4(c.Hx, i, hx);
    st4(c.Hy, i, hy);
    st4(c.Hz, i, hz);
    cp4(P.exn, ex);
    cp4(P.eyn, ey);
}

template <int MODE>
__device__ __forceinline__ void h3d_v4_tile(const DevModel &m, const Tile &tl, int lane, int ty,
                                            const float *mh_tab) {
    const int Nx = m.Nx, Ny = m.Ny, Nz = m.Nz;
    const int xi = tl.y + 4 * lane, iy = tl.z + ty;
    if (xi >= Nx || iy >= Ny) return;
    const int z0 = tile_z0(tl), z1 = tile_z1(tl);
    HCtx c;
    c.Ex = G((const float *)m.F[EX]); c.Ey = G((const float *)m.F[EY]); c.Ez = G((const float *)m.F[EZ]);
    c.Hx = G(m.F[HX]); c.Hy = G(m.F[HY]); c.Hz = G(m.F[HZ]);
    c.id = G(m.id);
    c.sz = m.sz;
    c.col = xi + m.sy * iy;
    c.Nz = Nz;
    c.bcz = m.bcs[5];
    const bool yin = iy + 1 < Ny;
    c.yload = yin || m.bcs[3] == BC_PERIODIC;
    c.yoff = yin ? m.sy : -(long long)(Ny - 1) * m.sy;
    c.edge = lane == 63 || xi + 4 >= Nx;
    c.eload = c.edge && (xi + 4 < Nx || m.bcs[1] == BC_PERIODIC);
    c.xoff = xi + 4 < Nx ? 4 : 4 - Nx;
    Lay L[6];
#pragma unroll
    for (int f = 0; f < 6; f++) L[f] = lay(m.L[f], true);
    const bool xs0 = L[0].present && xi <= L[0].ib, xs1 = L[1].present && xi + 3 >= L[1].ib;
    const bool ys0 = L[2].present && iy <= L[2].ib, ys1 = L[3].present && iy >= L[3].ib;
    float ex[4], ey[4];
    ld4(c.Ex, c.col + c.sz * z0, ex);
    ld4(c.Ey, c.col + c.sz * z0, ey);
    HPlane A, B;
    h_load<MODE>(c, z0, A);
    for (int iz = z0; iz < z1; iz += 2) {
        if (iz + 1 < z1) h_load<MODE>(c, iz + 1, B);
        h_compute<MODE>(m, c, iz, A, ex, ey, xi, iy, mh_tab, L, xs0, xs1, ys0, ys1);
        if (iz + 1 >= z1) break;
        if (iz + 2 < z1) h_load<MODE>(c, iz + 2, A);
        h_compute<MODE>(m, c, iz + 1, B, ex, ey, xi, iy, mh_tab, L, xs0, xs1, ys0, ys1);
    }
}

template <int MODE, int TY>
__global__ __launch_bounds__(64 * TY) void k_h3d_v4(const DevModel *__restrict__ models,
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
        h3d_v4_tile<MODE>(m, tl, lane, ty, mh_tab);
        __syncthreads();
    }
}

// ================================ E half-step ========================================
struct EPlane {
    float hx[4], hy[4], hz[4], hxy[4], hzy[4], d0[4], d1[4], d2[4], sa[4], e2[4];
    float hyE, hzE;            // x-1 neighbour of the lane's first cell (edge lane only)
    unsigned char g[4];
};

struct ECtx {
    gcf *Hx, *Hy, *Hz;
    gf *Dx, *Dy, *Dz, *Ex, *Ey, *Ez, *Sa, *E2;
    const __attribute__((address_space(1))) unsigned char *id;
    long long col, sz, yoff, xoff;
    bool yload, edge, eload, doSa, doE2;
};

template <int MODE, bool INTEG>
__device__ __forceinline__ void e_load(const ECtx &c, int iz, EPlane &P) {
    const long long i = c.col + c.sz * iz;
    ld4(c.Hx, i, P.hx);
    ld4(c.Hy, i, P.hy);
    ld4(c.Hz, i, P.hz);
    if (c.yload) { ld4(c.Hx, i + c.yoff, P.hxy); ld4(c.Hz, i + c.yoff, P.hzy); }
    ld4(G((const float *)c.Dx), i, P.d0);
    ld4(G((const float *)c.Dy), i, P.d1);
    ld4(G((const float *)c.Dz), i, P.d2);
    if (c.eload) { P.hyE = c.Hy[i + c.xoff]; P.hzE = c.Hz[i + c.xoff]; }
    if (MODE != MODE_VAC || INTEG) ldid(c.id, i, P.g);
    if (INTEG && c.doSa) ld4(G((const float *)c.Sa), i, P.sa);
    if (INTEG && c.doE2) ld4(G((const float *)c.E2), i, P.e2);
}

template <int MODE, bool INTEG>
__device__ __forceinline__ void e_compute(const DevModel &m, const ECtx &c, int iz, EPlane &P,
                                          float (&hxm)[4], float (&hym)[4], int xi, int iy,
                                          const float4 *ctab, const Lay (&L)[6], bool xs0,
                                          bool xs1, bool ys0, bool ys1) {
    const int Ny = m.Ny, Nx = m.Nx;
    const long long i = c.col + c.sz * iz;
    const float idx = m.idx, idy = m.idy, idz = m.idz, dtf = m.dtf;
    if (!c.yload) {
        if (m.bcs[2] == BC_DIRICHLET) { zero4(P.hxy); zero4(P.hzy); }
        else { cp4(P.hx, P.hxy); cp4(P.hz, P.hzy); }
    }
    float hy_p = __shfl_up(P.hy[3], 1), hz_p = __shfl_up(P.hz[3], 1);
    if (c.edge) {
        if (c.eload) { hy_p = P.hyE; hz_p = P.hzE; }
        else if (m.bcs[0] == BC_DIRICHLET) { hy_p = 0.0f; hz_p = 0.0f; }
        else { hy_p = P.hy[0]; hz_p = P.hz[0]; }
    }
    float4 cf[4];
#pragma unroll
    for (int k = 0; k < 4; k++) cf[k] = MODE != MODE_VAC ? ctab[P.g[k] & ID_MASK] : ctab[0];
    float dHxy[4], dHxz[4], dHyx[4], dHyz[4], dHzx[4], dHzy[4];
    float (&d0)[4] = P.d0, (&d1)[4] = P.d1, (&d2)[4] = P.d2;
#pragma unroll
    for (int k = 0; k < 4; k++) {
        const float hyx0 = k > 0 ? P.hy[k - 1] : hy_p;
        const float hzx0 = k > 0 ? P.hz[k - 1] : hz_p;
        dHxy[k] = (P.hx[k] - P.hxy[k]) * idy;
        dHxz[k] = (P.hx[k] - hxm[k]) * idz;
        dHyx[k] = (P.hy[k] - hyx0) * idx;
        dHyz[k] = (P.hy[k] - hym[k]) * idz;
        dHzx[k] = (P.hz[k] - hzx0) * idx;
        dHzy[k] = (P.hz[k] - P.hzy[k]) * idy;
    }
    // CPML x slabs (per cell), src/models.jl:976-997
    if (xs0) {
#pragma unroll
        for (int k = 0; k < 4; k++) {
            const int ix = xi + k;
            if (ix <= L[0].ib) {
                const long long j = ix + (long long)L[0].nb * (iy + (long long)Ny * iz);
                float a0, a1;
                psi_x(L[0], ix, j, dHyx[k], dHzx[k], a0, a1);
                d1[k] -= cf[k].y * a1; d2[k] += cf[k].y * a0;
                const float K = G(L[0].K)[ix];
                dHyx[k] = dHyx[k] / K; dHzx[k] = dHzx[k] / K;
            }
        }
    }
    if (xs1) {
#pragma unroll
        for (int k = 0; k < 4; k++) {
            const int ix = xi + k;
            if (ix >= L[1].ib) {
                const int p = ix - L[1].ib;
                const long long j = p + (long long)L[1].nb * (iy + (long long)Ny * iz);
                float a0, a1;
                psi_x(L[1], p, j, dHyx[k], dHzx[k], a0, a1);
                d1[k] -= cf[k].y * a1; d2[k] += cf[k].y * a0;
                const float K = G(L[1].K)[p];
                dHyx[k] = dHyx[k] / K; dHzx[k] = dHzx[k] / K;
            }
        }
    }
    for (int side = 0; side < 2; side++) {     // y slabs, src/models.jl:998-1015
        if (side == 0 ? !ys0 : !ys1) continue;
        const Lay &Ls = L[2 + side];
        const int p = side == 0 ? iy : iy - Ls.ib;
        const long long j = xi + (long long)Nx * (p + (long long)Ls.nb * iz);
        const float B = G(Ls.B)[p], A = G(Ls.A)[p], K = G(Ls.K)[p];
        float a0[4], a1[4];
        ld4(G((const float *)Ls.p0), j, a0);
        ld4(G((const float *)Ls.p1), j, a1);
#pragma unroll
        for (int k = 0; k < 4; k++) {
            a0[k] = B * a0[k] + A * dHxy[k];
            a1[k] = B * a1[k] + A * dHzy[k];
            d0[k] += cf[k].y * a1[k]; d2[k] -= cf[k].y * a0[k];
            dHxy[k] = dHxy[k] / K; dHzy[k] = dHzy[k] / K;
        }
        st4(G(Ls.p0), j, a0);
        st4(G(Ls.p1), j, a1);
    }
    for (int side = 0; side < 2; side++) {     // z slabs, src/models.jl:1016-1035
        const Lay &Ls = L[4 + side];
        if (!Ls.present) continue;
        if (side == 0 ? iz > Ls.ib : iz < Ls.ib) continue;
        const int p = side == 0 ? iz : iz - Ls.ib;
        const long long j = c.col + c.sz * p;
        const float B = G(Ls.B)[p], A = G(Ls.A)[p], K = G(Ls.K)[p];
        float a0[4], a1[4];
        ld4(G((const float *)Ls.p0), j, a0);
        ld4(G((const float *)Ls.p1), j, a1);
#pragma unroll
        for (int k = 0; k < 4; k++) {
            a0[k] = B * a0[k] + A * dHxz[k];
            a1[k] = B * a1[k] + A * dHyz[k];
            d0[k] -= cf[k].y * a1[k]; d1[k] += cf[k].y * a0[k];
            dHxz[k] = dHxz[k] / K; dHyz[k] = dHyz[k] / K;
        }
        st4(G(Ls.p0), j, a0);
        st4(G(Ls.p1), j, a1);
    }
    float e0[4], e1[4], e2[4];
#pragma unroll
    for (int k = 0; k < 4; k++) {
        d0[k] = cf[k].x * d0[k] + cf[k].y * (dHzy[k] - dHyz[k]);
        d1[k] = cf[k].x * d1[k] + cf[k].y * (dHxz[k] - dHzx[k]);
        d2[k] = cf[k].x * d2[k] + cf[k].y * (dHyx[k] - dHxy[k]);
        e0[k] = d0[k] * cf[k].z;
        e1[k] = d1[k] * cf[k].z;
        e2[k] = d2[k] * cf[k].z;
    }
    if (MODE == MODE_FULL) {
        // cells whose material carries ADE / Kerr / plasma terms (flag in ctab .w)
#pragma unroll
        for (int k = 0; k < 4; k++) {
            if (cf[k].w != 0.0f) {
                const DevMat &mt = m.mats[P.g[k] & ID_MASK];
                const float dd[3] = {d0[k], d1[k], d2[k]};
                float ee[3] = {0.0f, 0.0f, 0.0f};
                int s = 0;
                if (mt.flags & (MAT_DISP | MAT_PLASMA)) {
                    ee[0] = c.Ex[i + k]; ee[1] = c.Ey[i + k]; ee[2] = c.Ez[i + k];
                    s = G(m.slot)[i + k];
                }
                material_e<3>(mt, s, dd, ee, m.dt, dtf);
                e0[k] = ee[0]; e1[k] = ee[1]; e2[k] = ee[2];
            }
        }
    }
    st4(c.Dx, i, d0);
    st4(c.Dy, i, d1);
    st4(c.Dz, i, d2);
    st4(c.Ex, i, e0);
    st4(c.Ey, i, e1);
    st4(c.Ez, i, e2);
    bool bad = false;
#pragma unroll
    for (int k = 0; k < 4; k++) bad |= !isfinite(e0[k]);
    if (__any(bad) && bad) {
        int f = 0;
#pragma unroll
        for (int k = 0; k < 4; k++) f |= !isfinite(e0[k]) ? (isnan(e0[k]) ? 1 : 2) : 0;
        atomicOr(m.flag, f);
    }
    if (INTEG) {   // src/outputs.jl:480-483, skipped for cells a source rewrites
        if (c.doSa) {
#pragma unroll
            for (int k = 0; k < 4; k++) {
                if (P.g[k] & ID_TOUCHED) continue;
                const float a = e1[k] * P.hz[k] - e2[k] * P.hy[k];
                const float b = e2[k] * P.hx[k] - e0[k] * P.hz[k];
                const float cc = e0[k] * P.hy[k] - e1[k] * P.hx[k];
                P.sa[k] += sa_sqrt(a * a + b * b + cc * cc) * dtf;
            }
            st4(c.Sa, i, P.sa);
        }
        if (c.doE2) {
#pragma unroll
            for (int k = 0; k < 4; k++) {
                if (P.g[k] & ID_TOUCHED) continue;
                P.e2[k] += (e0[k] * e0[k] + e1[k] * e1[k] + e2[k] * e2[k]) * dtf;
            }
            st4(c.E2, i, P.e2);
        }
    }
    cp4(P.hx, hxm);
    cp4(P.hy, hym);
}

template <int MODE, bool INTEG>
__device__ __forceinline__ void e3d_v4_tile(const DevModel &m, const Tile &tl, int lane, int ty,
                                            const float4 *ctab) {
    const int Nx = m.Nx, Ny = m.Ny, Nz = m.Nz;
    const int xi = tl.y + 4 * lane, iy = tl.z + ty;
    if (xi >= Nx || iy >= Ny) return;
    const int z0 = tile_z0(tl), z1 = tile_z1(tl);
    ECtx c;
    c.Hx = G((const float *)m.F[HX]); c.Hy = G((const float *)m.F[HY]); c.Hz = G((const float *)m.F[HZ]);
    c.Dx = G(m.F[DX]); c.Dy = G(m.F[DY]); c.Dz = G(m.F[DZ]);
    c.Ex = G(m.F[EX]); c.Ey = G(m.F[EY]); c.Ez = G(m.F[EZ]);
    c.Sa = G(m.Sa); c.E2 = G(m.E2);
    c.doSa = INTEG && m.Sa; c.doE2 = INTEG && m.E2;
    c.id = G(m.id);
    c.sz = m.sz;
    c.col = xi + m.sy * iy;
    const bool yin = iy > 0;
    c.yload = yin || m.bcs[2] == BC_PERIODIC;
    c.yoff = yin ? -m.sy : (long long)(Ny - 1) * m.sy;
    c.edge = lane == 0 || xi == 0;
    c.eload = c.edge && (xi > 0 || m.bcs[0] == BC_PERIODIC);
    c.xoff = xi > 0 ? -1 : Nx - 1;
    Lay L[6];
#pragma unroll
    for (int f = 0; f < 6; f++) L[f] = lay(m.L[f], false);
    const bool xs0 = L[0].present && xi <= L[0].ib, xs1 = L[1].present && xi + 3 >= L[1].ib;
    const bool ys0 = L[2].present && iy <= L[2].ib, ys1 = L[3].present && iy >= L[3].ib;
    const int bcz = m.bcs[4];
    float hxm[4], hym[4];      // H(z-1)
    bool zm_own = false;
    if (z0 > 0) { ld4(c.Hx, c.col + c.sz * (z0 - 1), hxm); ld4(c.Hy, c.col + c.sz * (z0 - 1), hym); }
    else if (bcz == BC_PERIODIC) { ld4(c.Hx, c.col + c.sz * (Nz - 1), hxm); ld4(c.Hy, c.col + c.sz * (Nz - 1), hym); }
    else if (bcz == BC_DIRICHLET) { zero4(hxm); zero4(hym); }
    else zm_own = true;
    EPlane A, B;
    e_load<MODE, INTEG>(c, z0, A);
    if (zm_own) { cp4(A.hx, hxm); cp4(A.hy, hym); }
    for (int iz = z0; iz < z1; iz += 2) {
        if (iz + 1 < z1) e_load<MODE, INTEG>(c, iz + 1, B);
        e_compute<MODE, INTEG>(m, c, iz, A, hxm, hym, xi, iy, ctab, L, xs0, xs1, ys0, ys1);
        if (iz + 1 >= z1) break;
        if (iz + 2 < z1) e_load<MODE, INTEG>(c, iz + 2, A);
        e_compute<MODE, INTEG>(m, c, iz + 1, B, hxm, hym, xi, iy, ctab, L, xs0, xs1, ys0, ys1);
    }
}

template <int MODE, bool INTEG, int TY>
__global__ __launch_bounds__(64 * TY) void k_e3d_v4(const DevModel *__restrict__ models,
                                                    const Tile *__restrict__ tiles, int ntiles) {
    __shared__ float4 ctab[128];    // (Md1, Md2, Me, has ADE/Kerr/plasma)
    const int lane = threadIdx.x, ty = threadIdx.y;
    for (int t = blockIdx.x; t < ntiles; t += gridDim.x) {
        const Tile tl = tiles[t];
        if (tl.x < 0) continue;
        const DevModel &m = models[tl.x];
        if (*m.it >= m.Nt) continue;
        const int nm = MODE == MODE_VAC ? 0 : m.nmat;
        for (int k = lane + 64 * ty; k <= nm; k += 64 * TY) {
            const DevMat &mt = m.mats[k];
            ctab[k] = make_float4(mt.Md1, mt.Md2, mt.Me, mt.flags ? 1.0f : 0.0f);
        }
        __syncthreads();
        e3d_v4_tile<MODE, INTEG>(m, tl, lane, ty, ctab);
        __syncthreads();
    }
}

template <int TY>
static void launch_h_ty(const DevModel *m, const Tile *tiles, int ntiles, int grid, int mode,
                        hipStream_t s) {
    dim3 block(64, TY);
    if (mode == MODE_VAC) hipLaunchKernelGGL((k_h3d_v4<MODE_VAC, TY>), dim3(grid), block, 0, s, m, tiles, ntiles);
    else hipLaunchKernelGGL((k_h3d_v4<MODE_SIMPLE, TY>), dim3(grid), block, 0, s, m, tiles, ntiles);
}

template <int TY>
static void launch_e_ty(const DevModel *m, const Tile *tiles, int ntiles, int grid, int mode,
                        bool integ, hipStream_t s) {
    dim3 block(64, TY);
#define EV(MD, IN) hipLaunchKernelGGL((k_e3d_v4<MD, IN, TY>), dim3(grid), block, 0, s, m, tiles, ntiles)
    if (mode == MODE_VAC) { if (integ) EV(MODE_VAC, true); else EV(MODE_VAC, false); }
    else if (mode == MODE_SIMPLE) { if (integ) EV(MODE_SIMPLE, true); else EV(MODE_SIMPLE, false); }
    else { if (integ) EV(MODE_FULL, true); else EV(MODE_FULL, false); }
#undef EV
}

void launch_h3d_v4(const DevModel *m, const Tile *tiles, int ntiles, int grid, int mode, int ty,
                   hipStream_t s) {
    if (ty == 1) launch_h_ty<1>(m, tiles, ntiles, grid, mode, s);
    else if (ty == 2) launch_h_ty<2>(m, tiles, ntiles, grid, mode, s);
    else if (ty == 4) launch_h_ty<4>(m, tiles, ntiles, grid, mode, s);
    else launch_h_ty<8>(m, tiles, ntiles, grid, mode, s);
}

void launch_e3d_v4(const DevModel *m, const Tile *tiles, int ntiles, int grid, int mode,
                   bool integ, int ty, hipStream_t s) {
    if (ty == 1) launch_e_ty<1>(m, tiles, ntiles, grid, mode, integ, s);
    else if (ty == 2) launch_e_ty<2>(m, tiles, ntiles, grid, mode, integ, s);
    else if (ty == 4) launch_e_ty<4>(m, tiles, ntiles, grid, mode, integ, s);
    else launch_e_ty<8>(m, tiles, ntiles, grid, mode, integ, s);
}

}  // namespace mfd
