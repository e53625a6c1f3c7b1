// 3D Yee step in one z march: update_H_kernel! and update_E_kernel! 3D
// (src/models.jl:735-1175) fused into one launch.
//
// Each step reads the previous step's E and H from one buffer set and writes the new E
// and H into the other (ping-pong), so a step reads and writes each field once instead
// of twice.  A tile is 64 x-columns (one per lane) by R rows (every lane holds all R
// rows in registers) by a z range.  Each plane is done in one pass: the new H of the
// tile's cells at that plane, then their new E, which needs the new H of the plane
// below (kept from the previous plane) and of the row below and the column to the left.
// Those two belong to other workgroups running at the same time, so the tile computes
// them again from the old buffers: the row below on every lane, the column to the left
// on R lanes, and the plane below the tile in a first pass over plane z0-1.  That
// recomputation runs the same code as the owning tile (fused multiply-adds written out,
// the file built with -ffp-contract=off so the compiler fuses nothing else), so the
// results do not depend on the tiling.
//
// The H-half CPML psi arrays are ping-ponged like the fields (the halo cells read the
// old psi); the E-half psi, D, the integrals and the material state stay in place.
// A one-thread-per-model launch right after this one swaps the buffer pointers in the
// model descriptors, so every later launch of the step sees the new fields where it
// always looks (a per-tile atomic count instead, to let the last tile swap, serialises
// thousands of same-address atomics per step).
//
// Every load of a plane is issued at once (with Ex, Ey and the id of the plane above,
// which this plane's H and the next plane's stored-D choice need), then the plane is
// computed: a wave keeps R rows of loads in flight and few registers, so a SIMD holds
// many waves.  All addressing is a scalar base plus a 32-bit byte offset (models below
// 2^30 cells), and x neighbours move between lanes by DPP.
#include "common.hpp"
#include "material.hpp"
#include "tfsf.hpp"
#include "fused.hpp"

#include <algorithm>


namespace mfd {
namespace {

using namespace fused;

struct Hv {
    float x, y, z;
};

// H half-step of one cell (src/models.jl:755-876): its E, the E of its +x, +y, +z
// neighbours (eyx/ezx, exy/ezy, exz/eyz) and its old H.  Present slabs get their psi
// pairs updated in q (x: Eyx, Ezx; y: Exy, Ezy; z: Exz, Eyz), corrections before the
// division by K, in the x, y, z order of the reference.
__device__ __forceinline__ Hv h_cell(float Mh, float idx, float idy, float idz, float ex, float ey,
                                     float ez, float eyx, float ezx, float exy, float ezy, float exz,
                                     float eyz, float hx, float hy, float hz, bool bx, const Prof &px,
                                     bool by, const Prof &py, bool bz, const Prof &pz, float (&q)[6]) {
    float dExy = (exy - ex) * idy;
    float dExz = (exz - ex) * idz;
    float dEyx = (eyx - ey) * idx;
    float dEyz = (eyz - ey) * idz;
    float dEzx = (ezx - ez) * idx;
    float dEzy = (ezy - ez) * idy;
    if (bx) {   // src/models.jl:805-826
        q[0] = __builtin_fmaf(px.B, q[0], px.A * dEyx);
        q[1] = __builtin_fmaf(px.B, q[1], px.A * dEzx);
        hy = __builtin_fmaf(Mh, q[1], hy);
        hz = __builtin_fmaf(-Mh, q[0], hz);
        dEyx = dEyx * px.Ki;
        dEzx = dEzx * px.Ki;
    }
    if (by) {   // src/models.jl:827-848
        q[2] = __builtin_fmaf(py.B, q[2], py.A * dExy);
        q[3] = __builtin_fmaf(py.B, q[3], py.A * dEzy);
        hx = __builtin_fmaf(-Mh, q[3], hx);
        hz = __builtin_fmaf(Mh, q[2], hz);
        dExy = dExy * py.Ki;
        dEzy = dEzy * py.Ki;
    }
    if (bz) {   // src/models.jl:849-871
        q[4] = __builtin_fmaf(pz.B, q[4], pz.A * dExz);
        q[5] = __builtin_fmaf(pz.B, q[5], pz.A * dEyz);
        hx = __builtin_fmaf(Mh, q[5], hx);
        hy = __builtin_fmaf(-Mh, q[4], hy);
        dExz = dExz * pz.Ki;
        dEyz = dEyz * pz.Ki;
    }
    Hv r;   // src/models.jl:874-876
    r.x = __builtin_fmaf(-Mh, dEzy - dEyz, hx);
    r.y = __builtin_fmaf(-Mh, dExz - dEzx, hy);
    r.z = __builtin_fmaf(-Mh, dEyx - dExy, hz);
    return r;
}

// D (or, for a cell without stored D, E in E form) of one cell in the E half-step
// (src/models.jl:920-1040): its new H, the new H of its -y (hxy/hzy), -x (hyx/hzx) and
// -z (hxm/hym) neighbours.  Slab corrections before the Md1 scaling, as in the reference;
// psi pairs in q (x: Hyx, Hzx; y: Hxy, Hzy; z: Hxz, Hyz).
__device__ __forceinline__ void e_cell(float Md1, float Md2, float idx, float idy, float idz, float hx,
                                       float hy, float hz, float hxy, float hzy, float hyx, float hzx,
                                       float hxm, float hym, float &d0, float &d1, float &d2, bool bx,
                                       const Prof &px, bool by, const Prof &py, bool bz, const Prof &pz,
                                       float (&q)[6]) {
    float dHxy = (hx - hxy) * idy;
    float dHxz = (hx - hxm) * idz;
    float dHyx = (hy - hyx) * idx;
    float dHyz = (hy - hym) * idz;
    float dHzx = (hz - hzx) * idx;
    float dHzy = (hz - hzy) * idy;
    if (bx) {   // src/models.jl:976-997
        q[0] = __builtin_fmaf(px.B, q[0], px.A * dHyx);
        q[1] = __builtin_fmaf(px.B, q[1], px.A * dHzx);
        d1 = __builtin_fmaf(-Md2, q[1], d1);
        d2 = __builtin_fmaf(Md2, q[0], d2);
        dHyx = dHyx * px.Ki;
        dHzx = dHzx * px.Ki;
    }
    if (by) {   // src/models.jl:998-1015
        q[2] = __builtin_fmaf(py.B, q[2], py.A * dHxy);
        q[3] = __builtin_fmaf(py.B, q[3], py.A * dHzy);
        d0 = __builtin_fmaf(Md2, q[3], d0);
        d2 = __builtin_fmaf(-Md2, q[2], d2);
        dHxy = dHxy * py.Ki;
        dHzy = dHzy * py.Ki;
    }
    if (bz) {   // src/models.jl:1016-1035
        q[4] = __builtin_fmaf(pz.B, q[4], pz.A * dHxz);
        q[5] = __builtin_fmaf(pz.B, q[5], pz.A * dHyz);
        d0 = __builtin_fmaf(-Md2, q[5], d0);
        d1 = __builtin_fmaf(Md2, q[4], d1);
        dHxz = dHxz * pz.Ki;
        dHyz = dHyz * pz.Ki;
    }
    d0 = __builtin_fmaf(Md1, d0, Md2 * (dHzy - dHyz));   // src/models.jl:1038-1040
    d1 = __builtin_fmaf(Md1, d1, Md2 * (dHxz - dHzx));
    d2 = __builtin_fmaf(Md1, d2, Md2 * (dHyx - dHxy));
}

}  // namespace

// INTEG: 0 no integral variables; 3 Sa and E2 in every model of the launch; 4 each
// model's own (Sa and / or E2, read from its descriptor)
template <int MODE, int INTEG, int R, int PM>
__global__ __launch_bounds__(64) void k_step3d_f(const DevModel *__restrict__ models, const Tile *__restrict__ tiles,
                                                 int ntiles) {
    __shared__ float4 ctab[128];
    __shared__ float2 ktab[128];    // Kerr Pade (K2, K3)
    __shared__ float cw[128];       // Me*Md2 (the E-form update without D), full mode
    // incident values of the TFSF x faces for this tile's rows and planes:
    // [side][H_u, H_v, E_u, E_v][row][plane - z0]
    __shared__ float xinc[2][4][R][FUSED_XFACE_ZC];
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
        const int Nx = m.Nx, Ny = m.Ny, Nz = m.Nz;
        const int x0 = tl.y, y0 = tl.z, z0 = tile_z0(tl), z1 = tile_z1(tl);
        // TFSF x faces finished here, as in the two-pass E kernel: the incident values of
        // the tile's x-face cells come into LDS once per tile
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
            const int nzt = z1 - z0;
            xtile = ((xb[0] - 1 >= x0 && xb[0] - 1 < x0 + 64) || (xb[0] >= x0 && xb[0] < x0 + 64) ||
                     (xb[1] >= x0 && xb[1] < x0 + 64)) &&
                    y0 <= xb[3] && y0 + R - 1 >= xb[2] && z0 <= xb[5] && z1 - 1 >= xb[4];
            if (xtile) {
                const int it0 = *m.it;
                for (int k = lane; k < 8 * R * nzt; k += 64) {
                    const int zz = k % nzt, r = (k / nzt) % R, kk = (k / (nzt * R)) % 4, side = k / (nzt * R * 4);
                    const int u = y0 + r, v = z0 + zz;
                    const bool in = (kk == 0 || kk == 3) ? (u >= xb[2] && u <= xb[3] && v >= xb[4] && v <= xb[5] - 1)
                                                         : (u >= xb[2] && u <= xb[3] - 1 && v >= xb[4] && v <= xb[5]);
                    xinc[side][kk][r][zz] = in && xl[kk] ? tfsf_inc(S, Nz, 0, side, kk, u, v, it0) : 0.0f;
                }
            }
        }
        __syncthreads();

        const int ix = x0 + lane;
        const bool xin = ix < Nx;
        const int ixc = xin ? ix : Nx - 1;
        const int lx = min(63, Nx - 1 - x0);            // last lane inside the grid
        const int nr = min(R, Ny - y0);                 // rows inside the grid
        const unsigned syB = 4u * (unsigned)m.sy, szB = 4u * (unsigned)m.sz;
        const float idx = m.idx, idy = m.idy, idz = m.idz, dtf = m.dtf;
        const float Mh = m.mats[0].Mh;                  // every material has mu = 1
        const float *Ex = m.F[EX], *Ey = m.F[EY], *Ez = m.F[EZ];          // previous step
        const float *Hx = m.F[HX], *Hy = m.F[HY], *Hz = m.F[HZ];
        float *Exw = m.Falt[EX], *Eyw = m.Falt[EY], *Ezw = m.Falt[EZ];    // this step
        float *Hxw = m.Falt[HX], *Hyw = m.Falt[HY], *Hzw = m.Falt[HZ];
        float *Dx = m.F[DX], *Dy = m.F[DY], *Dz = m.F[DZ];
        float *Sa = m.Sa, *E2 = m.E2;
        const bool doSa = INTEG == 3 || (INTEG == 4 && Sa), doE2 = INTEG == 3 || (INTEG == 4 && E2);
        const unsigned char *id = m.id;
        const bool dless = m.dless;
        const bool useg = MODE != MODE_VAC || INTEG || dless;
        // row below the tile (its new Hx, Hz feed E of row 0): mode 0 computed from the old
        // buffers, 1 zero (Dirichlet), 2 row 0's own value (Neumann); ybn: the rule for
        // that row's own +y neighbour (0: the tile's row 0)
        int yb, ybm, ybn;
        if (y0 > 0) { yb = y0 - 1; ybm = 0; ybn = 0; }
        else if (m.bcs[2] == BC_PERIODIC) { yb = Ny - 1; ybm = 0; ybn = bcmode(m.bcs[3]); }
        else { yb = 0; ybm = bcmode(m.bcs[2]); ybn = 0; }
        // row above the last row (Ex, Ez for its H)
        int yt, ytm;
        if (y0 + nr < Ny) { yt = y0 + nr; ytm = 0; }
        else { ytm = bcmode(m.bcs[3]); yt = ytm == 0 ? 0 : Ny - 1; }
        // column left of lane 0 (its new Hy, Hz feed E of lane 0); xhn: the rule for that
        // column's own +x neighbour (0: the tile's first column)
        int xh, xhm, xhn;
        if (x0 > 0) { xh = x0 - 1; xhm = 0; xhn = 0; }
        else if (m.bcs[0] == BC_PERIODIC) { xh = Nx - 1; xhm = 0; xhn = bcmode(m.bcs[1]); }
        else { xh = 0; xhm = bcmode(m.bcs[0]); xhn = 0; }
        // column right of the last lane (Ey, Ez for its H)
        int xe, xem;
        if (x0 + lx + 1 < Nx) { xe = x0 + lx + 1; xem = 0; }
        else { xem = bcmode(m.bcs[1]); xe = xem == 0 ? 0 : Nx - 1; }
        // first plane of the march: z0 - 1 (H only) when E of plane z0 needs H below it
        int pb, zbm;
        if (z0 > 0) { pb = z0 - 1; zbm = 0; }
        else if (m.bcs[4] == BC_PERIODIC) { pb = -1; zbm = 0; }
        else { pb = 0; zbm = bcmode(m.bcs[4]); }
        const int zhm = bcmode(m.bcs[5]);   // +z neighbour of plane Nz-1

        const unsigned colB = 4u * (unsigned)ixc;
        unsigned rowB[R];   // byte offset of each row's cell in a plane
#pragma unroll
        for (int r = 0; r < R; r++) rowB[r] = colB + syB * (unsigned)min(y0 + r, Ny - 1);
        const unsigned bB = colB + syB * (unsigned)yb, tB = colB + syB * (unsigned)yt;
        // lane-distributed halo columns: left column lane t = row t (lane nr: the row
        // above), right column lane 0 = the row below, lane t = row t - 1
        const int xhrow = lane < nr ? y0 + lane : yt;
        const unsigned xhB = 4u * (unsigned)xh + syB * (unsigned)xhrow;
        const int erow = lane == 0 ? yb : y0 + min(lane - 1, nr - 1);
        const unsigned eB = 4u * (unsigned)xe + syB * (unsigned)erow;
        auto wz = [&](int q) { return (unsigned)(q < 0 ? q + Nz : (q >= Nz ? q - Nz : q)); };

        // cells on a TFSF face plane (they keep D): lane, row and plane parts of the test
        const bool xf_l = ix == xb[0] - 1 || ix == xb[0] || ix == xb[1], xi_l = ix >= xb[0] && ix <= xb[1];
        bool yf_r[R], yi_r[R];
#pragma unroll
        for (int r = 0; r < R; r++) {
            const int iy = y0 + r;
            yf_r[r] = (iy == xb[2] - 1 || iy == xb[2] || iy == xb[3]);
            yi_r[r] = iy >= xb[2] && iy <= xb[3];
        }

        // CPML slabs fixed over the march, with their profile values: x of this lane and
        // of the left column, y of the rows, of the row below and of the left column's
        // lane row (lane t: row t)
        int fx = -1, px = 0, fxh = -1, pxh = 0, fyb = -1, pyb = 0, fyh = -1, pyh = 0;
        int fy[R], py[R];
        Prof xP{0.0f, 0.0f, 1.0f}, xhP = xP, ybP = xP, yhP = xP, yP[R];
#pragma unroll
        for (int r = 0; r < R; r++) { fy[r] = -1; py[r] = 0; yP[r] = xP; }
        if (PM) {
            fx = slab_of(m, 0, ixc, px);
            fxh = slab_of(m, 0, xh, pxh);
            fyb = slab_of(m, 1, yb, pyb);
            fyh = slab_of(m, 1, min(xhrow, Ny - 1), pyh);
            xP = prof_of(m, fx, px);
            xhP = prof_of(m, fxh, pxh);
            ybP = prof_of(m, fyb, pyb);
            yhP = prof_of(m, fyh, pyh);
#pragma unroll
            for (int r = 0; r < R; r++) {
                fy[r] = slab_of(m, 1, min(y0 + r, Ny - 1), py[r]);
                yP[r] = prof_of(m, fy[r], py[r]);
            }
        }
        const int hrow = min(xhrow, Ny - 1);
        // x-slab psi arrays through one base (the lower face's, or the upper face's when
        // the lower face has no layer): the upper face's array follows the lower face's in
        // the same allocation, so a lane of either face needs only a byte offset
        const DevLayer &LX = m.L[0].present ? m.L[0] : m.L[1];
        const v2f *XE = LX.pE, *XH = LX.pH;
        v2f *XEw = LX.pEb, *XHw = LX.pH;
        const unsigned nbx0 = (unsigned)m.L[0].nb, nbx1 = (unsigned)m.L[1].nb;
        const unsigned x1B = m.L[0].present ? 8u * nbx0 * (unsigned)Ny * (unsigned)Nz : 0u;
        auto xo = [&](int f, int p, int iy, int iz) {
            const unsigned nb = f == 0 ? nbx0 : nbx1;
            return (f == 1 ? x1B : 0u) + 8u * ((unsigned)p + nb * ((unsigned)iy + (unsigned)Ny * (unsigned)iz));
        };
        // the same for the y and z slabs
        const DevLayer &LY = m.L[2].present ? m.L[2] : m.L[3];
        const v2f *YE = LY.pE, *YH = LY.pH;
        v2f *YEw = LY.pEb, *YHw = LY.pH;
        const unsigned nby0 = (unsigned)m.L[2].nb, nby1 = (unsigned)m.L[3].nb;
        const unsigned y1B = m.L[2].present ? 8u * (unsigned)Nx * nby0 * (unsigned)Nz : 0u;
        auto yo = [&](int f, int p, int ix_, int iz) {
            const unsigned nb = f == 2 ? nby0 : nby1;
            return (f == 3 ? y1B : 0u) + 8u * ((unsigned)ix_ + (unsigned)Nx * ((unsigned)p + nb * (unsigned)iz));
        };
        const DevLayer &LZ = m.L[4].present ? m.L[4] : m.L[5];
        const v2f *ZE = LZ.pE, *ZH = LZ.pH;
        v2f *ZEw = LZ.pEb, *ZHw = LZ.pH;
        const unsigned z1B = m.L[4].present ? 4u * (unsigned)m.L[4].nb * (unsigned)m.sz : 0u;
        auto zof = [&](int f, int p, unsigned cellB) {   // cellB: byte offset of the cell in its plane
            return 2u * ((f == 5 ? z1B : 0u) + cellB + szB * (unsigned)p);
        };
        // carried from one plane to the next: Ex, Ey and the id of the rows (loaded with
        // the plane below as its +z neighbour), Ey of the row below, Ex of the left column
        float ex[R], ey[R], bey, hex;
        int gid[R];
        {
            const unsigned zo = szB * wz(pb);
#pragma unroll
            for (int r = 0; r < R; r++) {
                ex[r] = ldb<float>(Ex, rowB[r] + zo);
                ey[r] = ldb<float>(Ey, rowB[r] + zo);
                gid[r] = useg ? (int)ldb<unsigned char>(id, (rowB[r] + zo) >> 2) : 0;
            }
            bey = ldb<float>(Ey, bB + zo);
            hex = ldb<float>(Ex, xhB + zo);
        }
        float hxm[R], hym[R];     // new H of the plane below
        bool zm_own = false;
#pragma unroll
        for (int r = 0; r < R; r++) { hxm[r] = 0.0f; hym[r] = 0.0f; }
        if (zbm == 2) zm_own = true;
        int nonfin = 0;
        // x-face TFSF columns of this lane; xhs: their H corrections stored here
        const bool xcol = xtile && xf_l;
        const bool xhs = m.xface_h != 0;
        const bool nts = m.ntst != 0;
        const int xside = ix == xb[1] ? 1 : 0;

        for (int p = pb; p < z1; p++) {
            const int zg = p < 0 ? p + Nz : p;
            const int ztm = zg == Nz - 1 ? zhm : 0;   // rule for the +z neighbour of this plane
            const bool pro = p < z0;                   // the plane below the tile: H only
            const unsigned zo = szB * (unsigned)zg, zn = szB * wz(p + 1);
            int fz = -1, pz = 0;
            Prof zP{0.0f, 0.0f, 1.0f};
            if (PM & 4) {
                fz = slab_of(m, 2, zg, pz);
                zP = prof_of(m, fz, pz);
            }
            const bool bz = (PM & 4) && fz >= 0, bx = (PM & 1) && fx >= 0;
            const bool zf_p = zg == xb[4] - 1 || zg == xb[4] || zg == xb[5], zi_p = zg >= xb[4] && zg <= xb[5];
            // ---- every load of the plane, issued together: fields, integrals, stored D,
            // psi pairs, the halo rows and columns.  Values a plane does not load (psi
            // outside the slabs, D of plain cells, integrals of the prologue plane, a
            // halo the boundary rule replaces) are left unset: every use sits behind the
            // condition of its load
            float exn[R], eyn[R], ez[R], hx[R], hy[R], hz[R], sa[R], e2[R], dd[R][3];
            float hq[R][6], eq[R][6];
            int gidn[R];
            bool dc[R];
#pragma unroll
            for (int r = 0; r < R; r++) {
                const unsigned o = rowB[r] + zo, on = rowB[r] + zn;
                const int iyc = min(y0 + r, Ny - 1);
                exn[r] = ldb<float>(Ex, on);
                eyn[r] = ldb<float>(Ey, on);
                gidn[r] = useg ? (int)ldb<unsigned char>(id, on >> 2) : 0;
                ez[r] = ldb<float>(Ez, o);
                hx[r] = ldb<float>(Hx, o);
                hy[r] = ldb<float>(Hy, o);
                hz[r] = ldb<float>(Hz, o);
                // cells that keep D: on a TFSF face plane and, from the id, cells a source
                // writes and Kerr / ADE / plasma cells
                bool d = false;
                if (!pro) {
                    d = !dless || (xf_l && yi_r[r] && zi_p) || (yf_r[r] && xi_l && zi_p) || (zf_p && xi_l && yi_r[r]);
                    if (useg && dless && !d) {
                        const int g = gid[r];
                        d = (g & ID_TOUCHED) || (MODE == MODE_FULL && ctab[g & ID_MASK].w != 0.0f);
                    }
                    if (d) { dd[r][0] = ldb<float>(Dx, o); dd[r][1] = ldb<float>(Dy, o); dd[r][2] = ldb<float>(Dz, o); }
                    if (doSa) sa[r] = ldnt(Sa, o);
                    if (doE2) e2[r] = ldnt(E2, o);
                }
                dc[r] = d;
                if (bx) {
                    const unsigned j = xo(fx, px, iyc, zg);
                    const v2f q = ldb<v2f>(XE, j);
                    hq[r][0] = q.x;
                    hq[r][1] = q.y;
                    if (!pro) { const v2f w = ldb<v2f>(XH, j); eq[r][0] = w.x; eq[r][1] = w.y; }
                }
                if ((PM & 2) && fy[r] >= 0) {
                    const unsigned j = yo(fy[r], py[r], ixc, zg);
                    const v2f q = ldb<v2f>(YE, j);
                    hq[r][2] = q.x;
                    hq[r][3] = q.y;
                    if (!pro) { const v2f w = ldb<v2f>(YH, j); eq[r][2] = w.x; eq[r][3] = w.y; }
                }
                if (bz) {
                    const unsigned j = zof(fz, pz, rowB[r]);
                    const v2f q = ldb<v2f>(ZE, j);
                    hq[r][4] = q.x;
                    hq[r][5] = q.y;
                    if (!pro) { const v2f w = ldb<v2f>(ZH, j); eq[r][4] = w.x; eq[r][5] = w.y; }
                }
            }
            const float beyn = ldb<float>(Ey, bB + zn), hexn = ldb<float>(Ex, xhB + zn);
            const float tex = ldb<float>(Ex, tB + zo), tez = ldb<float>(Ez, tB + zo);
            const float cey = ldb<float>(Ey, eB + zo), cez = ldb<float>(Ez, eB + zo);
            // the row below: Ex, Ez, old Hx, Hz and the psi values its Hx, Hz need
            float bexv, bez, bhx, bhz, bq[6];
            const bool bby = (PM & 2) && fyb >= 0;
            if (!pro && ybm == 0) {
                bexv = ldb<float>(Ex, bB + zo); bez = ldb<float>(Ez, bB + zo);
                bhx = ldb<float>(Hx, bB + zo); bhz = ldb<float>(Hz, bB + zo);
                if (bx) bq[0] = ldb<float>(XE, xo(fx, px, yb, zg));
                if (bby) {
                    const unsigned j = yo(fyb, pyb, ixc, zg);
                    const v2f q = ldb<v2f>(YE, j);
                    bq[2] = q.x;
                    bq[3] = q.y;
                }
                if (bz) bq[5] = ldb<float>(ZE, zof(fz, pz, bB) + 4u);
            }
            // the left column (lane t: row t): Ey, Ez, old Hy, Hz and the psi values its Hy,
            // Hz need
            float hey, hez, hhy, hhz, xq[6];
            const bool xbx = (PM & 1) && fxh >= 0, xby = (PM & 2) && fyh >= 0;
            if (!pro && xhm == 0) {
                hey = ldb<float>(Ey, xhB + zo); hez = ldb<float>(Ez, xhB + zo);
                hhy = ldb<float>(Hy, xhB + zo); hhz = ldb<float>(Hz, xhB + zo);
                if (PM && lane < nr) {
                        if (xbx) {
                        const unsigned j = xo(fxh, pxh, hrow, zg);
                        const v2f q = ldb<v2f>(XE, j);
                        xq[0] = q.x;
                        xq[1] = q.y;
                    }
                    if (xby) xq[2] = ldb<float>(YE, yo(fyh, pyh, xh, zg));
                    if (bz) xq[4] = ldb<float>(ZE, zof(fz, pz, 4u * (unsigned)xh + syB * (unsigned)hrow));
                }
            }

            // ---- new H of the tile's rows (src/models.jl:735-889)
            Hv Hn[R];
#pragma unroll
            for (int r = 0; r < R; r++) {
                float eyx = from_next(ey[r]), ezx = from_next(ez[r]);
                const float cy = rdlane(cey, r + 1), cz = rdlane(cez, r + 1);
                if (lane == lx) {
                    eyx = bcsel(xem, cy, ey[r]);
                    ezx = bcsel(xem, cz, ez[r]);
                }
                float exy, ezy;
                if (r + 1 < R && r + 1 < nr) { exy = ex[r + 1]; ezy = ez[r + 1]; }
                else { exy = bcsel(ytm, tex, ex[r]); ezy = bcsel(ytm, tez, ez[r]); }
                const float exz = ztm == 0 ? exn[r] : bcsel(ztm, 0.0f, ex[r]);
                const float eyz = ztm == 0 ? eyn[r] : bcsel(ztm, 0.0f, ey[r]);
                const bool by = (PM & 2) && fy[r] >= 0;
                Hn[r] = h_cell(Mh, idx, idy, idz, ex[r], ey[r], ez[r], eyx, ezx, exy, ezy, exz, eyz, hx[r], hy[r],
                               hz[r], bx, xP, by, yP[r], bz, zP, hq[r]);
                if (PM && !pro && xin && r < nr) {
                    const int iyc = min(y0 + r, Ny - 1);
                    if (bx) {
                        const unsigned j = xo(fx, px, iyc, zg);
                        stb<v2f>(XEw, j, mk2(hq[r][0], hq[r][1]));
                    }
                    if (by) {
                        const unsigned j = yo(fy[r], py[r], ixc, zg);
                        stb<v2f>(YEw, j, mk2(hq[r][2], hq[r][3]));
                    }
                    if (bz) {
                        const unsigned j = zof(fz, pz, rowB[r]);
                        stb<v2f>(ZEw, j, mk2(hq[r][4], hq[r][5]));
                    }
                }
            }
            if (!pro) {
                // ---- new Hx, Hz of the row below (recomputed; the other tile stores them)
                float hxb = 0.0f, hzb = 0.0f;
                if (ybm == 0) {
                    float eyx = from_next(bey);
                    const float cy = rdlane(cey, 0);
                    if (lane == lx) eyx = bcsel(xem, cy, bey);
                    const float exy = bcsel(ybn, ex[0], bexv), ezy = bcsel(ybn, ez[0], bez);
                    const float eyz = ztm == 0 ? beyn : bcsel(ztm, 0.0f, bey);
                    const Hv hb = h_cell(Mh, idx, idy, idz, bexv, bey, bez, eyx, bez, exy, ezy, bexv, eyz, bhx,
                                         0.0f, bhz, bx, xP, bby, ybP, bz, zP, bq);
                    hxb = hb.x;
                    hzb = hb.z;
                }
                // ---- new Hy, Hz of the column left of lane 0, on lanes 0..nr-1 (lane t: row t)
                float hyh = 0.0f, hzh = 0.0f;
                if (xhm == 0) {
                    float ezr = 0.0f, eyr = 0.0f;   // E of the tile's first column, row t
#pragma unroll
                    for (int r = 0; r < R; r++) {
                        const float vz = rdlane(ez[r], 0), vy = rdlane(ey[r], 0);
                        if (lane == r) { ezr = vz; eyr = vy; }
                    }
                    ezr = bcsel(xhn, ezr, hez);
                    eyr = bcsel(xhn, eyr, hey);
                    float exy = from_next(hex);
                    if (lane == nr - 1) exy = bcsel(ytm, exy, hex);
                    const float exz = ztm == 0 ? hexn : bcsel(ztm, 0.0f, hex);
                    const Hv hh = h_cell(Mh, idx, idy, idz, hex, hey, hez, eyr, ezr, exy, hez, exz, hey, 0.0f, hhy,
                                         hhz, xbx, xhP, xby, yhP, bz, zP, xq);
                    hyh = hh.y;
                    hzh = hh.z;
                }
                // ---- new E of the tile's rows (src/models.jl:895-1175)
#pragma unroll
                for (int r = 0; r < R; r++) {
                    const int iy = y0 + r;
                    const bool act = xin && r < nr;
                    const unsigned o = rowB[r] + zo;
                    if (zm_own) { hxm[r] = Hn[r].x; hym[r] = Hn[r].y; }
                    float hxy, hzy;
                    if (r > 0) { hxy = Hn[r - 1].x; hzy = Hn[r - 1].z; }
                    else { hxy = bcsel(ybm, hxb, Hn[0].x); hzy = bcsel(ybm, hzb, Hn[0].z); }
                    float hyx = from_prev(Hn[r].y), hzx = from_prev(Hn[r].z);
                    const float vy = rdlane(hyh, r), vz = rdlane(hzh, r);
                    if (lane == 0) {
                        hyx = bcsel(xhm, vy, Hn[r].y);
                        hzx = bcsel(xhm, vz, Hn[r].z);
                    }
                    const int gg = gid[r];
                    const int g = gg & ID_MASK;
                    const float4 cf = MODE != MODE_VAC ? ctab[g] : ctab[0];
                    const bool dcr = dc[r];
                    const float Md1 = cf.x, Md2 = dcr ? cf.y : (MODE == MODE_FULL ? cw[g] : cf.w);
                    float d0 = dcr ? dd[r][0] : ex[r];
                    float d1 = dcr ? dd[r][1] : ey[r];
                    float d2 = dcr ? dd[r][2] : ez[r];
                    const bool by = (PM & 2) && fy[r] >= 0;
                    e_cell(Md1, Md2, idx, idy, idz, Hn[r].x, Hn[r].y, Hn[r].z, hxy, hzy, hyx, hzx, hxm[r], hym[r],
                           d0, d1, d2, bx, xP, by, yP[r], bz, zP, eq[r]);
                    if (PM && act) {
                        const int iyc = min(iy, Ny - 1);
                        if (bx) {
                            const unsigned j = xo(fx, px, iyc, zg);
                            stb<v2f>(XHw, j, mk2(eq[r][0], eq[r][1]));
                        }
                        if (by) {
                            const unsigned j = yo(fy[r], py[r], ixc, zg);
                            stb<v2f>(YHw, j, mk2(eq[r][2], eq[r][3]));
                        }
                        if (bz) {
                            const unsigned j = zof(fz, pz, rowB[r]);
                            stb<v2f>(ZHw, j, mk2(eq[r][4], eq[r][5]));
                        }
                    }
                    float e[3];
                    if (dcr) { e[0] = d0 * cf.z; e[1] = d1 * cf.z; e[2] = d2 * cf.z; }
                    else { e[0] = d0; e[1] = d1; e[2] = d2; }       // already E
                    bool defer = false;
                    if (MODE == MODE_FULL && cf.w != 0.0f) {
                        if (cf.w > 1.5f) {
                            defer = true;
                        } else {   // Kerr Pade, src/models.jl:1140-1149
                            const float2 kk = ktab[g];
                            const float du[3] = {d0, d1, d2};
#pragma unroll
                            for (int c = 0; c < 3; c++) {
                                const float u = du[c];
                                e[c] = (1.0f + kk.x * u + 2.0f * kk.y * (u * u)) /
                                       (1.0f + 2.0f * kk.x * u + 3.0f * kk.y * (u * u)) * u * cf.z;
                            }
                        }
                    }
                    // x-face TFSF epilogue (src/sources.jl:725-745); the H correction of the
                    // H-out cells enters the integrals here, and the stored H too when
                    // m.xface_h is set (else the face kernel writes it)
                    float hyI = Hn[r].y, hzI = Hn[r].z;
                    bool xh2 = false;
                    if (xcol && iy >= xb[2] && iy <= xb[3] && zg >= xb[4] && zg <= xb[5]) {
                        xh2 = tfsf_xhandled_box(xb, ix, iy, zg, defer);
                        if (xh2) {
                            const int u = iy, v = zg, side = xside, zz = zg - z0;
                            const int U2 = xb[3], V2 = xb[5];
                            const int nH = side ? xb[1] : xb[0] - 1, nE = side ? xb[1] : xb[0];
                            const float s0 = side ? 1.0f : -1.0f;
                            if (ix == nH) {
                                if (u <= U2 && v <= V2 - 1 && xl[3]) hyI = __builtin_fmaf(s0 * xcm, xinc[side][3][r][zz], hyI);
                                if (u <= U2 - 1 && v <= V2 && xl[2]) hzI = __builtin_fmaf(-s0 * xcm, xinc[side][2][r][zz], hzI);
                            }
                            if (ix == nE) {
                                if (u <= U2 - 1 && v <= V2 && xl[1]) {
                                    const float iv = xinc[side][1][r][zz];
                                    e[1] = __builtin_fmaf(-s0 * xce, iv, e[1]);
                                    d1 = __builtin_fmaf(-s0 * xcd, iv, d1);
                                }
                                if (u <= U2 && v <= V2 - 1 && xl[0]) {
                                    const float iv = xinc[side][0][r][zz];
                                    e[2] = __builtin_fmaf(s0 * xce, iv, e[2]);
                                    d2 = __builtin_fmaf(s0 * xcd, iv, d2);
                                }
                            }
                        }
                    }
                    if (act) {
                        if (dcr) { stb<float>(Dx, o, d0); stb<float>(Dy, o, d1); stb<float>(Dz, o, d2); }
                        // a cell whose E the state pass finishes keeps E^n, which that pass reads
                        if (defer) { e[0] = ex[r]; e[1] = ey[r]; e[2] = ez[r]; }
                        const float h1 = xhs ? hyI : Hn[r].y, h2 = xhs ? hzI : Hn[r].z;
                        if (nts) {   // the new fields are read by the next step only
                            // the empty asm statements keep the two branches' stores apart:
                            // hoisted or sunk into one, they would lose the non-temporal hint
                            asm volatile("" ::: "memory");
                            stnt(Hxw, o, Hn[r].x); stnt(Hyw, o, h1); stnt(Hzw, o, h2);
                            stnt(Exw, o, e[0]); stnt(Eyw, o, e[1]); stnt(Ezw, o, e[2]);
                            asm volatile("" ::: "memory");
                        } else {
                            stb<float>(Hxw, o, Hn[r].x); stb<float>(Hyw, o, h1); stb<float>(Hzw, o, h2);
                            stb<float>(Exw, o, e[0]); stb<float>(Eyw, o, e[1]); stb<float>(Ezw, o, e[2]);
                        }
                        if (!defer && !isfinite(e[0])) nonfin |= isnan(e[0]) ? 1 : 2;
                        if (INTEG && !defer && (!(gg & ID_TOUCHED) || xh2)) {   // src/outputs.jl:480-483
                            if (doSa) stnt(Sa, o, sa_step3(sa[r], e[0], e[1], e[2], Hn[r].x, hyI, hzI, dtf));
                            if (doE2) stnt(E2, o, e2_step3(e2[r], e[0], e[1], e[2], dtf));
                        }
                    }
                }
                zm_own = false;
            }
#pragma unroll
            for (int r = 0; r < R; r++) {
                hxm[r] = Hn[r].x;
                hym[r] = Hn[r].y;
                ex[r] = exn[r];
                ey[r] = eyn[r];
                gid[r] = gidn[r];
            }
            bey = beyn;
            hex = hexn;
        }
        if (__any(nonfin != 0) && nonfin) atomicOr(m.flag, nonfin);
        // the LDS tables are rewritten by the next tile
        __syncthreads();
    }
}

// Swaps the H/E buffers and the H-half psi arrays of every model the fused launch just
// stepped (one thread per model; models that had finished did not step).
__global__ void k_swap_pp(DevModel *models, int nmodels) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= nmodels) return;
    DevModel &m = models[k];
    if (*m.it >= m.Nt) return;
    const int cs[6] = {HX, HY, HZ, EX, EY, EZ};
    for (int c = 0; c < 6; c++) {
        float *a = m.F[cs[c]];
        m.F[cs[c]] = m.Falt[cs[c]];
        m.Falt[cs[c]] = a;
    }
    for (int f = 0; f < 6; f++) {
        DevLayer &L = m.L[f];
        if (m.ndims == 3) {
            v2f *a = L.pE;
            L.pE = L.pEb;
            L.pEb = a;
        } else {
            float *&e = f < 2 ? L.psiE1 : L.psiE0;
            float *a = e;
            e = L.psiEb;
            L.psiEb = a;
        }
    }
}

// k_finish and k_swap_pp in one launch, for steps in which nothing reads the fields
// after the fused launch: the buffer swap, the step counter and the DFT block position
__global__ void k_finish_swap(DevModel *models, int nmodels) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= nmodels) return;
    DevModel &m = models[k];
    const int it = *m.it;
    if (it >= m.Nt) return;
    const int cs[6] = {HX, HY, HZ, EX, EY, EZ};
    for (int c = 0; c < 6; c++) {
        float *a = m.F[cs[c]];
        m.F[cs[c]] = m.Falt[cs[c]];
        m.Falt[cs[c]] = a;
    }
    for (int f = 0; f < 6; f++) {
        DevLayer &L = m.L[f];
        if (m.ndims == 3) {
            v2f *a = L.pE;
            L.pE = L.pEb;
            L.pEb = a;
        } else {
            float *&e = f < 2 ? L.psiE1 : L.psiE0;
            float *a = e;
            e = L.psiEb;
            L.psiEb = a;
        }
    }
    *m.it = it + 1;
    if (it + 1 - *m.dftdone == m.dft_block) *m.dftdone = it + 1;
}

template <int R, int PM>
static void lf(DevModel *m, const Tile *tiles, int ntiles, int grid, int mode, int integ, hipStream_t s) {
#define SF_(MD, IN) hipLaunchKernelGGL((k_step3d_f<MD, IN, R, PM>), dim3(grid), dim3(64), 0, s, m, tiles, ntiles)
#define SI_(MD) do { if (integ == 3) SF_(MD, 3); else if (integ) SF_(MD, 4); else SF_(MD, 0); } while (0)
    if (mode == MODE_VAC) SI_(MODE_VAC);
    else if (mode == MODE_SIMPLE) SI_(MODE_SIMPLE);
    else SI_(MODE_FULL);
#undef SI_
#undef SF_
}

void launch_swap_pp(DevModel *m, int nmodels, hipStream_t s) {
    hipLaunchKernelGGL(k_swap_pp, dim3((nmodels + 63) / 64), dim3(64), 0, s, m, nmodels);
}

void launch_finish_swap(DevModel *m, int nmodels, hipStream_t s) {
    hipLaunchKernelGGL(k_finish_swap, dim3((nmodels + 63) / 64), dim3(64), 0, s, m, nmodels);
}

// pmask: the CPML axes the tiles' cells (and recomputed halo cells) may lie in; the
// kernel is built for 0 (none), 1 (x only) or 7 (any).  integ: 0 none, 3 Sa and E2 in
// every model, else each model's own
void launch_step3d_f(DevModel *m, const Tile *tiles, int ntiles, int grid, int mode, int integ, int rows,
                     int pmask, hipStream_t s) {
    if (ntiles <= 0) return;
#define LF_(RR) (pmask == 0 ? lf<RR, 0>(m, tiles, ntiles, grid, mode, integ, s)   \
                 : pmask == 1 ? lf<RR, 1>(m, tiles, ntiles, grid, mode, integ, s) \
                              : lf<RR, 7>(m, tiles, ntiles, grid, mode, integ, s))
    if (rows == 2) LF_(2);
    else LF_(4);
#undef LF_
}

}  // namespace mfd
