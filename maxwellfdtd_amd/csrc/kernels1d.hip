// 1D (z) Yee half-steps with CPML and materials (update_H_kernel! /
// update_E_kernel! 1D, src/models.jl:248-444): fields Hy, Dx, Ex.  One thread per
// cell; a tile is 256 consecutive z cells.  Instantiated for the FP32 state and for
// models that keep their state in FP64.
#include "common.hpp"
#include "material.hpp"
#include "step_dev.hpp"

namespace mfd {

// T = float: the FP32 state; T = double: a model that keeps its state in FP64 (its CPML
// layer arrays then hold doubles too).

// H half-step of cell iz (update_H_kernel! 1D, src/models.jl:248-307); hmat: Mh from the
// cell's material (some material has mu != 1), else the vacuum value
template <class T>
__device__ __forceinline__ void h1d_cell(const DevModel &m, int iz, bool hmat) {
    typedef Prec<T> P;
    const int Nz = m.Nz;
    const gp<const T> Ex = G((const T *)P::field(m, EX));
    const gp<T> Hy = G(P::field(m, HY));
    const int bcz = m.bcs[5];
    const gp<const typename P::Mat> mats = G(P::mats(m));
    T Mh = mats[0].Mh;
    if (hmat) Mh = mats[G(m.id)[iz] & ID_MASK].Mh;
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
            const gp<T> psi = G((T *)l.psiE0);
            const T a = G((const T *)l.B)[p] * psi[p] + G((const T *)l.A)[p] * dExz;
            psi[p] = a;
            hy -= Mh * a;
            dExz = dExz / G((const T *)l.K)[p];
        }
    }
    {
        const DevLayer &l = m.L[5];
        if (l.present && iz >= l.ib) {
            const int p = iz - l.ib;
            const gp<T> psi = G((T *)l.psiE0);
            const T a = G((const T *)l.B)[p] * psi[p] + G((const T *)l.A)[p] * dExz;
            psi[p] = a;
            hy -= Mh * a;
            dExz = dExz / G((const T *)l.K)[p];
        }
    }
    hy -= Mh * ((T)0 + dExz);
    Hy[iz] = hy;
}

// E half-step of cell iz at step it (update_E_kernel! 1D, src/models.jl:310-444): D with
// CPML, then E through the material, the non-finite guard and the integrals
template <int MODE, bool INTEG, class T>
__device__ __forceinline__ void e1d_cell(const DevModel &m, int iz, int it) {
    typedef Prec<T> P;
    const int Nz = m.Nz;
    const gp<const T> Hy = G((const T *)P::field(m, HY));
    const gp<T> Dx = G(P::field(m, DX));
    const gp<T> Ex = G(P::field(m, EX));
    const int bcz = m.bcs[4];
    int idb = 0;
    if (MODE != MODE_VAC || INTEG) idb = G(m.id)[iz];
    const __attribute__((address_space(1))) typename P::Mat &mt = G(P::mats(m))[MODE != MODE_VAC ? (idb & ID_MASK) : 0];
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
            const gp<T> psi = G((T *)l.psiH1);
            const T a = G((const T *)l.B)[p] * psi[p] + G((const T *)l.A)[p] * dHyz;
            psi[p] = a;
            d[0] -= Md2 * a;
            dHyz = dHyz / G((const T *)l.K)[p];
        }
    }
    {
        const DevLayer &l = m.L[5];
        if (l.present && iz >= l.ib) {
            const int p = iz - l.ib;
            const gp<T> psi = G((T *)l.psiH1);
            const T a = G((const T *)l.B)[p] * psi[p] + G((const T *)l.A)[p] * dHyz;
            psi[p] = a;
            d[0] -= Md2 * a;
            dHyz = dHyz / G((const T *)l.K)[p];
        }
    }
    d[0] = Md1 * d[0] + Md2 * ((T)0 - dHyz);
    Dx[iz] = d[0];
    T e[1] = {(T)0};
    if (MODE == MODE_FULL && (mt.flags & (MAT_DISP | MAT_PLASMA))) {
        e[0] = Ex[iz];
        material_e<1>(mt, G(m.slot)[iz], it & 1, d, e, m.dt, P::dt(m));
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
            G(Sa)[iz] += sqrt_of(a * a) * dt;
        }
        if (E2) G(E2)[iz] += e[0] * e[0] * dt;
    }
}

template <int MODE, class T>
__global__ __launch_bounds__(256) void k_h1d(const DevModel *__restrict__ models,
                                             const Tile *__restrict__ tiles, int ntiles) {
    for (int t = blockIdx.x; t < ntiles; t += gridDim.x) {
        const Tile tl = tiles[t];
        if (tl.x < 0) continue;
        const DevModel &m = models[tl.x];
        if (*m.it >= m.Nt) continue;
        const int iz = tile_z0(tl) + threadIdx.x;
        if (iz >= tile_z1(tl)) continue;
        h1d_cell<T>(m, iz, MODE != MODE_VAC);
    }
}

template <int MODE, bool INTEG, class T>
__global__ __launch_bounds__(256) void k_e1d(const DevModel *__restrict__ models,
                                             const Tile *__restrict__ tiles, int ntiles) {
    for (int t = blockIdx.x; t < ntiles; t += gridDim.x) {
        const Tile tl = tiles[t];
        if (tl.x < 0) continue;
        const DevModel &m = models[tl.x];
        if (*m.it >= m.Nt) continue;
        const int iz = tile_z0(tl) + threadIdx.x;
        if (iz >= tile_z1(tl)) continue;
        e1d_cell<MODE, INTEG, T>(m, iz, *m.it);
    }
}

// ---- whole runs: one workgroup per model steps it through nsteps steps (or to its Nt)
// in one launch, in the order of the per-step launches: H half, E half, the point
// sources in slot order, the monitors and the integrals of the cells a source wrote, a
// due DFT block, the step counter.  Workgroup barriers separate the pieces, and the
// model's few arrays stay in L2 between them.  Launched once per step, a model of a few
// thousand cells spends its time on the five launches of each step instead.
constexpr int RUN1D_THREADS = 1024;

template <int MODE, bool INTEG, class T>
__global__ __launch_bounds__(RUN1D_THREADS) void k_run1d(const DevModel *__restrict__ models, int nsteps,
                                                         int hmat) {
    __shared__ double2 ph[KMAX * WCH];
    __shared__ double sh[RUN1D_THREADS];
    const DevModel &m = models[blockIdx.x];
    const int tid = threadIdx.x, bs = blockDim.x;
    // this workgroup is the only writer of the model's counters: they stay in registers,
    // the same in every thread, so every loop bound below is uniform over the workgroup
    int it = *m.it, done = *m.dftdone;
    const int Nz = m.Nz, Nt = m.Nt;
    for (int s = 0; s < nsteps && it < Nt; s++) {
        for (int iz = tid; iz < Nz; iz += bs) h1d_cell<T>(m, iz, hmat != 0);
        __syncthreads();
        for (int iz = tid; iz < Nz; iz += bs) e1d_cell<MODE, INTEG, T>(m, iz, it);
        __syncthreads();
        for (int q = 0; q < m.nsrc; q++) {
            const __attribute__((address_space(1))) DevSource &S = G(m.src)[q];
            if (S.kind != SRC_TFSF) {
                for (int k0 = 0; k0 < S.n; k0 += bs) {
                    const int k = k0 + tid;
                    bool wroteEx = false;
                    float v = 0.0f;
                    if (k < S.n) v = point_source_entry(m, S, it, k, wroteEx);
                    check_finite(v, m.flag, wroteEx);
                }
            }
            __syncthreads();
        }
        // the monitors and the integrals of the cells a source wrote read the same fields
        // and write different arrays: one piece
        if (m.nmon) monitors_model(m, it, 0, tid, bs, true, sh);
        if (INTEG && m.int_flags) integrals_touched_model(m, 1, tid, bs);
        __syncthreads();
        if (m.nmon && it + 1 - done == m.dft_block) {
            dft_flush_model(m, done, it + 1, tid, bs, true, ph, sh);
            __syncthreads();
        }
        it++;
        if (it - done == m.dft_block) done = it;
        if (tid == 0) {
            *m.it = it;
            *m.dftdone = done;
        }
    }
}

void launch_run1d(const DevModel *m, int nmodels, int nsteps, int mode, bool integ, bool f64, bool hmat,
                  hipStream_t s) {
#define R1(MD, IN, T) \
    hipLaunchKernelGGL((k_run1d<MD, IN, T>), dim3(nmodels), dim3(RUN1D_THREADS), 0, s, m, nsteps, hmat ? 1 : 0)
#define R1P(T)                                                                              \
    if (mode == MODE_VAC) { if (integ) R1(MODE_VAC, true, T); else R1(MODE_VAC, false, T); } \
    else if (mode == MODE_SIMPLE) { if (integ) R1(MODE_SIMPLE, true, T); else R1(MODE_SIMPLE, false, T); } \
    else { if (integ) R1(MODE_FULL, true, T); else R1(MODE_FULL, false, T); }
    if (f64) { R1P(double) } else { R1P(float) }
#undef R1P
#undef R1
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
