// 2D (x, z) Yee step in one z march: update_H_kernel! and update_E_kernel! 2D
// (src/models.jl:450-726) fused into one launch, the 2D counterpart of the 3D fused step.
//
// Fields Hy, Dx, Dz, Ex, Ez.  Each step reads the previous step's Hy, Ex, Ez from one
// buffer set and writes the new ones into the other.  A tile is 64 x columns (one per
// lane) by a z range; each plane (z row) is done in one pass: the new Hy of the row, then
// its new E, which needs the new Hy of the row below (kept from the previous plane) and
// of the column left of lane 0.  That column belongs to another workgroup, so the tile
// computes its Hy again from the old buffers; a first pass over plane z0 - 1 (Hy only)
// gives the row below the tile.  The multiply-adds are written out and the file is built
// with -ffp-contract=off, so recomputed values are the owner's bits and the results do not
// depend on the tiling.
//
// The H-half CPML psi (x faces: Ezx in psiE1, z faces: Exz in psiE0) is ping-ponged like
// the fields, its other buffer in psiEb; the E-half psi, D, the integrals and the material
// state stay in place.  Cells with ADE / plasma state keep Eⁿ in the new buffer for the
// state pass.  All addressing is a scalar base plus a 32-bit byte offset, and x neighbours
// move between lanes by DPP.
#include "common.hpp"
#include "fused.hpp"
#include "material.hpp"

namespace mfd {
namespace {

using namespace fused;

// Hy of one cell (src/models.jl:455-530): its Ex and Ez, Ez of its +x neighbour (ezx),
// Ex of its +z neighbour (exz), its old Hy.  Slab psi (x: Ezx, z: Exz) in qx / qz,
// corrections before the division by K, x before z as in the reference.
__device__ __forceinline__ float h_cell2(float Mh, float idx, float idz, float ex, float ez, float ezx, float exz,
                                         float hy, bool bx, const Prof &px, bool bz, const Prof &pz, float &qx,
                                         float &qz) {
    float dExz = (exz - ex) * idz;
    float dEzx = (ezx - ez) * idx;
    if (bx) {
        qx = __builtin_fmaf(px.B, qx, px.A * dEzx);
        hy = __builtin_fmaf(Mh, qx, hy);
        dEzx = dEzx * px.Ki;
    }
    if (bz) {
        qz = __builtin_fmaf(pz.B, qz, pz.A * dExz);
        hy = __builtin_fmaf(-Mh, qz, hy);
        dExz = dExz * pz.Ki;
    }
    return __builtin_fmaf(-Mh, dExz - dEzx, hy);
}

// Dx, Dz of one cell (or Ex, Ez in E form for a cell without stored D), src/models.jl
// :590-660: its new Hy, the new Hy of its -x (hyx) and -z (hym) neighbours.  Slab psi
// (x: Hyx, z: Hyz) in qx / qz.
__device__ __forceinline__ void e_cell2(float Md1, float Md2, float idx, float idz, float hy, float hyx, float hym,
                                        float &d0, float &d1, bool bx, const Prof &px, bool bz, const Prof &pz,
                                        float &qx, float &qz) {
    float dHyx = (hy - hyx) * idx;
    float dHyz = (hy - hym) * idz;
    if (bx) {
        qx = __builtin_fmaf(px.B, qx, px.A * dHyx);
        d1 = __builtin_fmaf(Md2, qx, d1);
        dHyx = dHyx * px.Ki;
    }
    if (bz) {
        qz = __builtin_fmaf(pz.B, qz, pz.A * dHyz);
        d0 = __builtin_fmaf(-Md2, qz, d0);
        dHyz = dHyz * pz.Ki;
    }
    d0 = __builtin_fmaf(Md1, d0, Md2 * (0.0f - dHyz));
    d1 = __builtin_fmaf(Md1, d1, Md2 * dHyx);
}

// Sa and E2 of a 2D cell advanced by one step (src/outputs.jl:322-325)
__device__ __forceinline__ float sa_step2(float sa, float ex, float ez, float hy, float dt) {
    const float a = -ez * hy, b = ex * hy;
    return __builtin_fmaf(sa_sqrt(__builtin_fmaf(a, a, b * b)), dt, sa);
}
__device__ __forceinline__ float e2_step2(float e2, float ex, float ez, float dt) {
    return __builtin_fmaf(__builtin_fmaf(ex, ex, ez * ez), dt, e2);
}

}  // namespace

// INTEG: 0 no integral variables; 3 Sa and E2 in every model; 4 each model's own.
// PM: 0 no CPML layer in any model of the launch, 1 any.
template <int MODE, int INTEG, int PM>
__global__ __launch_bounds__(64) void k_step2d_f(const DevModel *__restrict__ models, const Tile *__restrict__ tiles,
                                                 int ntiles) {
    __shared__ float4 ctab[128];
    __shared__ float2 ktab[128];    // Kerr Pade (K2, K3)
    __shared__ float cw[128];       // Me*Md2 (the E-form update without D), full mode
    const int lane = threadIdx.x;
    for (int t = blockIdx.x; t < ntiles; t += gridDim.x) {
        const Tile tl = tiles[t];
        if (tl.x < 0) continue;
        const DevModel &m = models[tl.x];
        if (*m.it >= m.Nt) continue;
        const int nm = MODE == MODE_VAC ? 0 : m.nmat;
        for (int k = lane; k <= nm; k += 64) {
            const DevMat &mt = m.mats[k];
            const float kind = (mt.flags & (MAT_DISP | MAT_PLASMA)) ? 2.0f : (mt.flags & MAT_KERR) ? 1.0f : 0.0f;
            ctab[k] = make_float4(mt.Md1, mt.Md2, mt.Me, MODE == MODE_FULL ? kind : mt.MeMd2);
            if (MODE == MODE_FULL) { ktab[k] = make_float2(mt.K2, mt.K3); cw[k] = mt.MeMd2; }
        }
        __syncthreads();

        const int Nx = m.Nx, Nz = m.Nz;
        const int x0 = tl.y, z0 = tile_z0(tl), z1 = tile_z1(tl);
        const int ix = x0 + lane;
        const bool xin = ix < Nx;
        const int ixc = xin ? ix : Nx - 1;
        const int lx = min(63, Nx - 1 - x0);            // last lane inside the grid
        const unsigned szB = 4u * (unsigned)m.sz;
        const float idx = m.idx, idz = m.idz, dtf = m.dtf;
        const float Mh = m.mats[0].Mh;                  // every material has mu = 1
        const float *Ex = m.F[EX], *Ez = m.F[EZ], *Hy = m.F[HY];         // previous step
        float *Exw = m.Falt[EX], *Ezw = m.Falt[EZ], *Hyw = m.Falt[HY];   // this step
        float *Dx = m.F[DX], *Dz = m.F[DZ];
        float *Sa = m.Sa, *E2 = m.E2;
        const bool doSa = INTEG == 3 || (INTEG == 4 && Sa), doE2 = INTEG == 3 || (INTEG == 4 && E2);
        const unsigned char *id = m.id;
        const bool dless = m.dless;
        const bool useg = MODE != MODE_VAC || INTEG || dless;
        // column left of lane 0 (its new Hy feeds Ez of lane 0): mode 0 computed from the
        // old buffers, 1 zero (Dirichlet), 2 lane 0's own value (Neumann); xhn: the rule for
        // that column's own +x neighbour (0: the tile's first column)
        int xh, xhm, xhn;
        if (x0 > 0) { xh = x0 - 1; xhm = 0; xhn = 0; }
        else if (m.bcs[0] == BC_PERIODIC) { xh = Nx - 1; xhm = 0; xhn = bcmode(m.bcs[1]); }
        else { xh = 0; xhm = bcmode(m.bcs[0]); xhn = 0; }
        // column right of the last lane (Ez for its Hy)
        int xe, xem;
        if (x0 + lx + 1 < Nx) { xe = x0 + lx + 1; xem = 0; }
        else { xem = bcmode(m.bcs[1]); xe = xem == 0 ? 0 : Nx - 1; }
        // first plane of the march: z0 - 1 (Hy only) when E of plane z0 needs Hy below it
        int pb, zbm;
        if (z0 > 0) { pb = z0 - 1; zbm = 0; }
        else if (m.bcs[4] == BC_PERIODIC) { pb = -1; zbm = 0; }
        else { pb = 0; zbm = bcmode(m.bcs[4]); }
        const int zhm = bcmode(m.bcs[5]);   // +z neighbour of plane Nz-1
        auto wz = [&](int q) { return (unsigned)(q < 0 ? q + Nz : (q >= Nz ? q - Nz : q)); };
        const unsigned colB = 4u * (unsigned)ixc, xhB = 4u * (unsigned)xh, xeB = 4u * (unsigned)xe;

        // CPML x slabs of this lane and of the left column, with their profile values
        int fx = -1, px = 0, fxh = -1, pxh = 0;
        Prof xP{0.0f, 0.0f, 1.0f}, xhP = xP;
        if (PM) {
            fx = slab_of(m, 0, ixc, px);
            fxh = slab_of(m, 0, xh, pxh);
            xP = prof_of(m, fx, px);
            xhP = prof_of(m, fxh, pxh);
        }
        // psi: x faces' arrays hold [face 0: p + nb0 * z][face 1: p + nb1 * z]; z faces'
        // [face 4: x + Nx * p][face 5: x + Nx * p]; each pair of faces in one allocation
        const DevLayer &LX = m.L[0].present ? m.L[0] : m.L[1];
        const float *XE = LX.psiE1, *XH = LX.psiH0;
        float *XEw = LX.psiEb, *XHw = LX.psiH0;
        const unsigned nbx0 = (unsigned)m.L[0].nb, nbx1 = (unsigned)m.L[1].nb;
        const unsigned x1B = m.L[0].present ? 4u * nbx0 * (unsigned)Nz : 0u;
        auto xo = [&](int f, int p, int iz) {
            const unsigned nb = f == 0 ? nbx0 : nbx1;
            return (f == 1 ? x1B : 0u) + 4u * ((unsigned)p + nb * (unsigned)iz);
        };
        const DevLayer &LZ = m.L[4].present ? m.L[4] : m.L[5];
        const float *ZE = LZ.psiE0, *ZH = LZ.psiH1;
        float *ZEw = LZ.psiEb, *ZHw = LZ.psiH1;
        const unsigned z1B = m.L[4].present ? 4u * (unsigned)Nx * (unsigned)m.L[4].nb : 0u;
        auto zof = [&](int f, int p, unsigned cB) {   // cB: byte offset of the column
            return (f == 5 ? z1B : 0u) + cB + 4u * (unsigned)Nx * (unsigned)p;
        };

        // Loads of a plane, in two groups.  The unconditional group (U) of plane p + 1 is
        // issued while plane p is computed, so its latency overlaps that work: the row's
        // Ez, Hy and integrals, Ex and the id of the plane above, the halo columns.  The
        // group that depends on the cells (stored D, CPML psi) is issued at the start of
        // the plane, before the next U group, so that waiting for it leaves the U group in
        // flight (vector loads complete in order).  Values a plane does not load (psi
        // outside the slabs, D of plain cells, the prologue plane's integrals, a halo the
        // boundary rule replaces) stay unset: every use sits behind the condition of its
        // load.
        struct U2 {
            float exn, ez, hy, hexn, cez, hez, hhy, sa, e2;
            int gidn;
        };
        auto loadU = [&](int p, U2 &u) {
            const int zg = p < 0 ? p + Nz : p;
            const unsigned zo = szB * (unsigned)zg, zn = szB * wz(p + 1), o = colB + zo;
            u.exn = ldb<float>(Ex, colB + zn);
            u.gidn = useg ? (int)ldb<unsigned char>(id, (colB + zn) >> 2) : 0;
            u.ez = ldb<float>(Ez, o);
            u.hy = ldb<float>(Hy, o);
            u.hexn = ldb<float>(Ex, xhB + zn);
            u.cez = ldb<float>(Ez, xeB + zo);
            u.hez = ldb<float>(Ez, xhB + zo);
            u.hhy = ldb<float>(Hy, xhB + zo);
            if (doSa) u.sa = ldnt(Sa, o);
            if (doE2) u.e2 = ldnt(E2, o);
        };

        // carried from one plane to the next: Ex and the id of the row (loaded with the
        // plane below as its +z neighbour), Ex of the left column, the new Hy below
        float ex, hex, hym = 0.0f;
        int gid = 0;
        {
            const unsigned zo = szB * wz(pb);
            ex = ldb<float>(Ex, colB + zo);
            gid = useg ? (int)ldb<unsigned char>(id, (colB + zo) >> 2) : 0;
            hex = ldb<float>(Ex, xhB + zo);
        }
        bool zm_own = zbm == 2;
        int nonfin = 0;

        auto plane = [&](int p, const U2 &u, U2 &un) {
            const int zg = p < 0 ? p + Nz : p;
            const int ztm = zg == Nz - 1 ? zhm : 0;   // rule for the +z neighbour of this plane
            const bool pro = p < z0;                   // the plane below the tile: Hy only
            const unsigned zo = szB * (unsigned)zg;
            const unsigned o = colB + zo;
            int fz = -1, pz = 0;
            Prof zP{0.0f, 0.0f, 1.0f};
            if (PM) {
                fz = slab_of(m, 2, zg, pz);
                zP = prof_of(m, fz, pz);
            }
            const bool bz = PM && fz >= 0, bx = PM && fx >= 0, bxh = PM && fxh >= 0;
            // ---- the cell-dependent loads of this plane, then the U group of the next
            float d0s, d1s, qhx, qhz, qex, qez, qxh, qzh;
            bool d = false;
            if (!pro) {
                d = !dless;
                if (useg && dless) {
                    d = (gid & ID_TOUCHED) || (MODE == MODE_FULL && ctab[gid & ID_MASK].w != 0.0f);
                }
                if (d) { d0s = ldb<float>(Dx, o); d1s = ldb<float>(Dz, o); }
            }
            if (bx) {
                const unsigned j = xo(fx, px, zg);
                qhx = ldb<float>(XE, j);
                if (!pro) qex = ldb<float>(XH, j);
            }
            if (bz) {
                const unsigned j = zof(fz, pz, colB);
                qhz = ldb<float>(ZE, j);
                if (!pro) qez = ldb<float>(ZH, j);
            }
            if (!pro && xhm == 0) {
                if (bxh) qxh = ldb<float>(XE, xo(fxh, pxh, zg));
                if (bz) qzh = ldb<float>(ZE, zof(fz, pz, xhB));
            }
            if (p + 1 < z1) loadU(p + 1, un);

            // ---- new Hy of the row (src/models.jl:455-530)
            float ezx = from_next(u.ez);
            if (lane == lx) ezx = bcsel(xem, u.cez, u.ez);
            const float exz = ztm == 0 ? u.exn : bcsel(ztm, 0.0f, ex);
            const float hn = h_cell2(Mh, idx, idz, ex, u.ez, ezx, exz, u.hy, bx, xP, bz, zP, qhx, qhz);
            if (PM && !pro && xin) {
                if (bx) stb<float>(XEw, xo(fx, px, zg), qhx);
                if (bz) stb<float>(ZEw, zof(fz, pz, colB), qhz);
            }
            if (!pro) {
                // ---- new Hy of the column left of lane 0 (recomputed; its tile stores it)
                float hyh = 0.0f;
                if (xhm == 0) {
                    const float ezxh = bcsel(xhn, rdlane(u.ez, 0), u.hez);
                    const float exzh = ztm == 0 ? u.hexn : bcsel(ztm, 0.0f, hex);
                    hyh = h_cell2(Mh, idx, idz, hex, u.hez, ezxh, exzh, u.hhy, bxh, xhP, bz, zP, qxh, qzh);
                }
                // ---- new E of the row (src/models.jl:590-725)
                if (zm_own) hym = hn;
                float hyx = from_prev(hn);
                if (lane == 0) hyx = bcsel(xhm, hyh, hn);
                const int g = gid & ID_MASK;
                const float4 cf = MODE != MODE_VAC ? ctab[g] : ctab[0];
                const float Md1 = cf.x, Md2 = d ? cf.y : (MODE == MODE_FULL ? cw[g] : cf.w);
                const float exo = ex, ezo = u.ez;
                float d0 = d ? d0s : exo;
                float d1 = d ? d1s : ezo;
                e_cell2(Md1, Md2, idx, idz, hn, hyx, hym, d0, d1, bx, xP, bz, zP, qex, qez);
                if (PM && xin) {
                    if (bx) stb<float>(XHw, xo(fx, px, zg), qex);
                    if (bz) stb<float>(ZHw, zof(fz, pz, colB), qez);
                }
                float e0, e1;
                if (d) { e0 = d0 * cf.z; e1 = d1 * cf.z; }
                else { e0 = d0; e1 = d1; }       // already E
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
                if (xin) {
                    stb<float>(Hyw, o, hn);
                    if (d) { stb<float>(Dx, o, d0); stb<float>(Dz, o, d1); }
                    // a cell whose E the state pass finishes keeps Eⁿ, which that pass reads
                    if (defer) { e0 = exo; e1 = ezo; }
                    stb<float>(Exw, o, e0);
                    stb<float>(Ezw, o, e1);
                    if (!defer && !isfinite(e0)) nonfin |= isnan(e0) ? 1 : 2;
                    if (INTEG && !defer && !(gid & ID_TOUCHED)) {   // src/outputs.jl:322-325
                        if (doSa) stnt(Sa, o, sa_step2(u.sa, e0, e1, hn, dtf));
                        if (doE2) stnt(E2, o, e2_step2(u.e2, e0, e1, dtf));
                    }
                }
                zm_own = false;
            }
            hym = hn;
            ex = u.exn;
            gid = u.gidn;
            hex = u.hexn;
        };

        // planes in pairs, the two U groups alternating between A and B
        U2 A, B;
        loadU(pb, A);
        for (int p = pb; p < z1; p += 2) {
            plane(p, A, B);
            if (p + 1 >= z1) break;
            plane(p + 1, B, A);
        }
        if (__any(nonfin != 0) && nonfin) atomicOr(m.flag, nonfin);
        // the LDS tables are rewritten by the next tile
        __syncthreads();
    }
}

template <int PM>
static void lf2(DevModel *m, const Tile *tiles, int ntiles, int grid, int mode, int integ, hipStream_t s) {
#define SF_(MD, IN) hipLaunchKernelGGL((k_step2d_f<MD, IN, PM>), dim3(grid), dim3(64), 0, s, m, tiles, ntiles)
#define SI_(MD) do { if (integ == 3) SF_(MD, 3); else if (integ) SF_(MD, 4); else SF_(MD, 0); } while (0)
    if (mode == MODE_VAC) SI_(MODE_VAC);
    else if (mode == MODE_SIMPLE) SI_(MODE_SIMPLE);
    else SI_(MODE_FULL);
#undef SI_
#undef SF_
}

// pml: some model of the launch has a CPML layer.  integ: 0 none, 3 Sa and E2 in every
// model, else each model's own
void launch_step2d_f(DevModel *m, const Tile *tiles, int ntiles, int grid, int mode, int integ, bool pml,
                     hipStream_t s) {
    if (ntiles <= 0) return;
    if (pml) lf2<1>(m, tiles, ntiles, grid, mode, integ, s);
    else lf2<0>(m, tiles, ntiles, grid, mode, integ, s);
}

}  // namespace mfd
