// 1D (z) Yee half-steps with CPML and materials (update_H_kernel! /
// update_E_kernel! 1D, src/models.jl:248-444): fields Hy, Dx, Ex.  One thread per
// cell; a tile is 256 consecutive z cells.  Instantiated for the FP32 state and for
// models that keep their state in FP64.
#include "common.hpp"
#include "material.hpp"

namespace mfd {

// T = float: the FP32 state; T = double: a model that keeps its state in FP64 (its CPML
// layer arrays then hold doubles too).
template <int MODE, class T>
__global__ __launch_bounds__(256) void k_h1d(const DevModel *__restrict__ models,
                                             const Tile *__restrict__ tiles, int ntiles) {
    typedef Prec<T> P;
    for (int t = blockIdx.x; t < ntiles; t += gridDim.x) {
        const Tile tl = tiles[t];
        if (tl.x < 0) continue;
        const DevModel &m = models[tl.x];
        if (*m.it >= m.Nt) continue;
        const int iz = tile_z0(tl) + threadIdx.x;
        const int Nz = m.Nz;
        if (iz >= tile_z1(tl)) continue;
        const T *__restrict__ Ex = P::field(m, EX);
        T *__restrict__ Hy = P::field(m, HY);
        const int bcz = m.bcs[5];
        T Mh = P::mats(m)[0].Mh;
        if (MODE != MODE_VAC) Mh = P::mats(m)[m.id[iz] & ID_MASK].Mh;
        const T ex = Ex[iz];
        T ex_zp;
        if (iz + 1 < Nz) ex_zp = Ex[iz + 1];
        else if (bcz == BC_PERIODIC) ex_zp = Ex[0];
        else if (bcz == BC_DIRICHLET) ex_zp = (T)0;
        else ex_zp = ex;
        T dExz = (ex_zp - ex) * P::idz(m);
        T hy = Hy[iz];
        {
            const DevLayer &l = m.L[4];
            if (l.present && iz <= l.ib) {
                const int p = iz;
                T *psi = (T *)l.psiE0;
                const T a = ((const T *)l.B)[p] * psi[p] + ((const T *)l.A)[p] * dExz;
                psi[p] = a;
                hy -= Mh * a;
                dExz = dExz / ((const T *)l.K)[p];
            }
        }
        {
            const DevLayer &l = m.L[5];
            if (l.present && iz >= l.ib) {
                const int p = iz - l.ib;
                T *psi = (T *)l.psiE0;
                const T a = ((const T *)l.B)[p] * psi[p] + ((const T *)l.A)[p] * dExz;
                psi[p] = a;
                hy -= Mh * a;
                dExz = dExz / ((const T *)l.K)[p];
            }
        }
        hy -= Mh * ((T)0 + dExz);
        Hy[iz] = hy;
    }
}

template <int MODE, bool INTEG, class T>
__global__ __launch_bounds__(256) void k_e1d(const DevModel *__restrict__ models,
                                             const Tile *__restrict__ tiles, int ntiles) {
    typedef Prec<T> P;
    for (int t = blockIdx.x; t < ntiles; t += gridDim.x) {
        const Tile tl = tiles[t];
        if (tl.x < 0) continue;
        const DevModel &m = models[tl.x];
        if (*m.it >= m.Nt) continue;
        const int iz = tile_z0(tl) + threadIdx.x;
        const int Nz = m.Nz;
        if (iz >= tile_z1(tl)) continue;
        const T *__restrict__ Hy = P::field(m, HY);
        T *__restrict__ Dx = P::field(m, DX);
        T *__restrict__ Ex = P::field(m, EX);
        const int bcz = m.bcs[4];
        int idb = 0;
        if (MODE != MODE_VAC || INTEG) idb = m.id[iz];
        const typename P::Mat &mt = P::mats(m)[MODE != MODE_VAC ? (idb & ID_MASK) : 0];
        const T Md1 = mt.Md1, Md2 = mt.Md2;
        const T hy = Hy[iz];
        T hy_zm;
        if (iz > 0) hy_zm = Hy[iz - 1];
        else if (bcz == BC_PERIODIC) hy_zm = Hy[Nz - 1];
        else if (bcz == BC_DIRICHLET) hy_zm = (T)0;
        else hy_zm = hy;
        T dHyz = (hy - hy_zm) * P::idz(m);
        T d[1] = {Dx[iz]};
        {
            const DevLayer &l = m.L[4];
            if (l.present && iz <= l.ib) {
                const int p = iz;
                T *psi = (T *)l.psiH1;
                const T a = ((const T *)l.B)[p] * psi[p] + ((const T *)l.A)[p] * dHyz;
                psi[p] = a;
                d[0] -= Md2 * a;
                dHyz = dHyz / ((const T *)l.K)[p];
            }
        }
        {
            const DevLayer &l = m.L[5];
            if (l.present && iz >= l.ib) {
                const int p = iz - l.ib;
                T *psi = (T *)l.psiH1;
                const T a = ((const T *)l.B)[p] * psi[p] + ((const T *)l.A)[p] * dHyz;
                psi[p] = a;
                d[0] -= Md2 * a;
                dHyz = dHyz / ((const T *)l.K)[p];
            }
        }
        d[0] = Md1 * d[0] + Md2 * ((T)0 - dHyz);
        Dx[iz] = d[0];
        T e[1] = {(T)0};
        if (MODE == MODE_FULL && (mt.flags & (MAT_DISP | MAT_PLASMA))) {
            e[0] = Ex[iz];
            material_e<1>(mt, m.slot[iz], *m.it & 1, d, e, m.dt, P::dt(m));
        } else if (MODE == MODE_FULL && (mt.flags & MAT_KERR)) {
            material_e<1>(mt, 0, 0, d, e, m.dt, P::dt(m));
        } else {
            e[0] = d[0] * (T)mt.Me;
        }
        Ex[iz] = e[0];
        check_finite(e[0], m.flag, true);
        if (INTEG && !(idb & ID_TOUCHED)) {   // src/outputs.jl:186-189
            const T dt = P::dt(m);
            T *Sa = P::Sa(m), *E2 = P::E2(m);
            if (Sa) {
                const T a = e[0] * hy;
                Sa[iz] += sqrt_of(a * a) * dt;
            }
            if (E2) E2[iz] += e[0] * e[0] * dt;
        }
    }
}

void launch_h1d(const DevModel *m, const Tile *tiles, int ntiles, int grid, int mode, bool f64,
                hipStream_t s) {
#define H1(MD, T) hipLaunchKernelGGL((k_h1d<MD, T>), dim3(grid), dim3(TX1), 0, s, m, tiles, ntiles)
    if (f64) { if (mode == MODE_VAC) H1(MODE_VAC, double); else H1(MODE_SIMPLE, double); }
    else { if (mode == MODE_VAC) H1(MODE_VAC, float); else H1(MODE_SIMPLE, float); }
#undef H1
}

void launch_e1d(const DevModel *m, const Tile *tiles, int ntiles, int grid, int mode, bool integ,
                bool f64, hipStream_t s) {
#define E1(MD, IN, T) hipLaunchKernelGGL((k_e1d<MD, IN, T>), dim3(grid), dim3(TX1), 0, s, m, tiles, ntiles)
#define E1P(T)                                                                              \
    if (mode == MODE_VAC) { if (integ) E1(MODE_VAC, true, T); else E1(MODE_VAC, false, T); } \
    else if (mode == MODE_SIMPLE) { if (integ) E1(MODE_SIMPLE, true, T); else E1(MODE_SIMPLE, false, T); } \
    else { if (integ) E1(MODE_FULL, true, T); else E1(MODE_FULL, false, T); }
    if (f64) { E1P(double) } else { E1P(float) }
#undef E1P
#undef E1
}

}  // namespace mfd
