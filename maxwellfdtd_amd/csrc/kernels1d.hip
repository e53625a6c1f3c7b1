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

// ---- whole runs with the cells in registers: models of up to REG1D_CELLS cells with
// simple materials, CPML layers of up to REG1D_NB cells that do not meet and at most
// REG1D_SRC point-source entries.  Thread t of 512 holds cells t, t + 512, ..., t + 3584
// (fields, CPML psi, integrals, material index); the z neighbours of a half-step come
// through an LDS copy of the other field; the material and CPML coefficients sit in LDS;
// the waveform values of a step are loaded at its start and used after its E half.  The
// fields go to memory only where a monitor reads them and at the end.  The arithmetic is
// h1d_cell's and e1d_cell's, term for term.
constexpr int REG1D_THREADS = 512, REG1D_CPT = 8, REG1D_CELLS = REG1D_THREADS * REG1D_CPT, REG1D_NB = 64,
              REG1D_SRC = 8;

template <int MODE, bool INTEG, class T>
__global__ __launch_bounds__(REG1D_THREADS) void k_reg1d(const DevModel *__restrict__ models, int nsteps,
                                                         int hmat) {
    typedef Prec<T> P;
    __shared__ double2 ph[KMAX * WCH];
    __shared__ double sh[REG1D_THREADS];
    __shared__ T xch[REG1D_CELLS];
    __shared__ T cMh[64], cMd1[64], cMd2[64], cMe[64];
    __shared__ T lB[2][REG1D_NB], lA[2][REG1D_NB], lK[2][REG1D_NB];
    const DevModel &m = models[blockIdx.x];
    const int tid = threadIdx.x, bs = blockDim.x;
    int it = *m.it, done = *m.dftdone;
    const int Nz = m.Nz, Nt = m.Nt;
    const gp<const typename P::Mat> mats = G(P::mats(m));
    for (int k = tid; k <= m.nmat; k += bs) {
        cMh[k] = mats[k].Mh;
        cMd1[k] = mats[k].Md1;
        cMd2[k] = mats[k].Md2;
        cMe[k] = mats[k].Me;
    }
    const DevLayer &L0 = m.L[4], &L1 = m.L[5];
    const int ib0 = L0.present ? L0.ib : -1, ib1 = L1.present ? L1.ib : Nz;
    for (int k = tid; k < REG1D_NB; k += bs) {
        if (k <= ib0) { lB[0][k] = G((const T *)L0.B)[k]; lA[0][k] = G((const T *)L0.A)[k]; lK[0][k] = G((const T *)L0.K)[k]; }
        if (ib1 + k < Nz) { lB[1][k] = G((const T *)L1.B)[k]; lA[1][k] = G((const T *)L1.A)[k]; lK[1][k] = G((const T *)L1.K)[k]; }
    }
    const gp<T> Hy = G(P::field(m, HY)), Dx = G(P::field(m, DX)), Ex = G(P::field(m, EX));
    const gp<T> Sa = G(P::Sa(m)), E2 = G(P::E2(m));
    T hy[REG1D_CPT], dx[REG1D_CPT], ex[REG1D_CPT], sa[REG1D_CPT], e2[REG1D_CPT];
    int g[REG1D_CPT];
    // CPML psi (H half, E half) of the cell the thread holds in each layer: the left layer
    // lies in the first cell block, the right one in the last (jl)
    const int jl = (Nz - 1) / bs;
    T peL = (T)0, phL = (T)0, peR = (T)0, phR = (T)0;
#pragma unroll
    for (int j = 0; j < REG1D_CPT; j++) {
        const int iz = j * bs + tid;
        hy[j] = dx[j] = ex[j] = sa[j] = e2[j] = (T)0;
        g[j] = 0;
        if (iz < Nz) {
            hy[j] = Hy[iz]; dx[j] = Dx[iz]; ex[j] = Ex[iz];
            g[j] = G(m.id)[iz];
            if (j == 0 && iz <= ib0) { peL = G((T *)L0.psiE0)[iz]; phL = G((T *)L0.psiH1)[iz]; }
            if (j == jl && iz >= ib1) { peR = G((T *)L1.psiE0)[iz - ib1]; phR = G((T *)L1.psiH1)[iz - ib1]; }
            if (INTEG && Sa) sa[j] = Sa[iz];
            if (INTEG && E2) e2[j] = E2[iz];
        }
    }
    const T idz = P::idz(m), dtT = P::dt(m);
    const int bcl = m.bcs[4], bcr = m.bcs[5];
    // the point-source entries in slot order: slot and entry of each (at most REG1D_SRC)
    __shared__ int es_q[REG1D_SRC], es_k[REG1D_SRC], es_n;
    // the monitor entries, when every monitor samples points (no sums) and there are at
    // most REG1D_SRC of them: the thread that holds a point writes its samples from its
    // registers (else the fields go to memory for monitors_model every step)
    __shared__ int em_q[REG1D_SRC], em_k[REG1D_SRC], em_n, dft_pts;
    if (tid == 0) {
        int n = 0;
        for (int q = 0; q < m.nsrc; q++)
            for (int k = 0; k < G(m.src)[q].n && n < REG1D_SRC; k++, n++) { es_q[n] = q; es_k[n] = k; }
        es_n = n;
        n = 0;
        for (int q = 0; q < m.nmon && n >= 0; q++) {
            const __attribute__((address_space(1))) DevMon &M = G(m.mon)[q];
            if (M.issum || n + M.n > REG1D_SRC) { n = -1; break; }
            for (int k = 0; k < M.n; k++, n++) { em_q[n] = q; em_k[n] = k; }
        }
        em_n = n;
        // the DFT blocks of point spectral monitors one thread per point and frequency
        int pts = 1;
        for (int q = 0; q < m.nmon; q++) {
            const __attribute__((address_space(1))) DevMon &M = G(m.mon)[q];
            if (M.spectral && (M.issum || m.dft_block * M.nw > KMAX * WCH)) pts = 0;
        }
        dft_pts = pts;
    }
    __syncthreads();
    const int nent = es_n, nment = em_n;
    const bool dftp = dft_pts != 0;
    for (int s = 0; s < nsteps && it < Nt; s++) {
        // this step's waveform values, used after the E half
        T wv[REG1D_SRC];
#pragma unroll
        for (int e = 0; e < REG1D_SRC; e++) {
            wv[e] = (T)0;
            if (e < nent) {
                const __attribute__((address_space(1))) DevSource &S = G(m.src)[es_q[e]];
                const size_t w = (size_t)it * S.n + es_k[e];
                wv[e] = m.f64 ? (T)G(S.wave64)[w] : (T)G(S.wave)[w];
            }
        }
        // ---- H half (h1d_cell)
#pragma unroll
        for (int j = 0; j < REG1D_CPT; j++) xch[j * bs + tid] = ex[j];
        __syncthreads();
#pragma unroll
        for (int j = 0; j < REG1D_CPT; j++) {
            const int iz = j * bs + tid;
            if (iz >= Nz) continue;
            const T Mh = hmat ? cMh[g[j] & ID_MASK] : cMh[0];
            const T e0 = ex[j];
            T ex_zp;
            if (iz + 1 < Nz) ex_zp = xch[iz + 1];
            else if (bcr == BC_PERIODIC) ex_zp = xch[0];
            else if (bcr == BC_DIRICHLET) ex_zp = (T)0;
            else ex_zp = e0;
            T dExz = (ex_zp - e0) * idz;
            T h = hy[j];
            if (j == 0 && iz <= ib0) {
                const T a2 = lB[0][iz] * peL + lA[0][iz] * dExz;
                peL = a2;
                h -= Mh * a2;
                dExz = dExz / lK[0][iz];
            }
            if (j == jl && iz >= ib1) {
                const int p = iz - ib1;
                const T a2 = lB[1][p] * peR + lA[1][p] * dExz;
                peR = a2;
                h -= Mh * a2;
                dExz = dExz / lK[1][p];
            }
            h -= Mh * ((T)0 + dExz);
            hy[j] = h;
        }
        __syncthreads();
        // ---- E half (e1d_cell)
#pragma unroll
        for (int j = 0; j < REG1D_CPT; j++) xch[j * bs + tid] = hy[j];
        __syncthreads();
#pragma unroll
        for (int j = 0; j < REG1D_CPT; j++) {
            const int iz = j * bs + tid;
            const bool act = iz < Nz;
            T e = (T)0;
            if (act) {
                const int gm = MODE != MODE_VAC ? (g[j] & ID_MASK) : 0;
                const T Md1 = cMd1[gm], Md2 = cMd2[gm];
                const T h = hy[j];
                T hy_zm;
                if (iz > 0) hy_zm = xch[iz - 1];
                else if (bcl == BC_PERIODIC) hy_zm = xch[Nz - 1];
                else if (bcl == BC_DIRICHLET) hy_zm = (T)0;
                else hy_zm = h;
                T dHyz = (h - hy_zm) * idz;
                T d = dx[j];
                if (j == 0 && iz <= ib0) {
                    const T a2 = lB[0][iz] * phL + lA[0][iz] * dHyz;
                    phL = a2;
                    d -= Md2 * a2;
                    dHyz = dHyz / lK[0][iz];
                }
                if (j == jl && iz >= ib1) {
                    const int p = iz - ib1;
                    const T a2 = lB[1][p] * phR + lA[1][p] * dHyz;
                    phR = a2;
                    d -= Md2 * a2;
                    dHyz = dHyz / lK[1][p];
                }
                d = Md1 * d + Md2 * ((T)0 - dHyz);
                dx[j] = d;
                e = d * cMe[gm];
                ex[j] = e;
                if (INTEG && !(g[j] & ID_TOUCHED)) {   // src/outputs.jl:186-189
                    if (Sa) {
                        const T a2 = e * h;
                        sa[j] += sqrt_of(a2 * a2) * dtT;
                    }
                    if (E2) e2[j] += e * e * dtT;
                }
            }
            check_finite(e, m.flag, act);
        }
        // ---- point sources, in slot order, by the thread that holds the cell
#pragma unroll
        for (int e = 0; e < REG1D_SRC; e++) {
            if (e >= nent) break;
            const __attribute__((address_space(1))) DevSource &S = G(m.src)[es_q[e]];
            const int c = G(S.cells)[es_k[e]];
            bool wrote = false;
            T v = (T)0;
            if (c % bs == tid) {
                const int jc = c / bs;
#pragma unroll
                for (int j = 0; j < REG1D_CPT; j++) {
                    if (j != jc) continue;
                    const int comp = S.comp;
                    const T f = comp == HY ? hy[j] : comp == DX ? dx[j] : ex[j];
                    if (m.f64) v = (S.kind == SRC_SOFT) ? f + (T)S.coeff64 * wv[e] : wv[e];
                    else v = (S.kind == SRC_SOFT) ? f + (T)S.coeff * wv[e] : wv[e];
                    if (comp == HY) hy[j] = v;
                    else if (comp == DX) dx[j] = v;
                    else ex[j] = v;
                }
                wrote = S.comp == EX;
            }
            check_finite((float)v, m.flag, wrote);
        }
        // integrals of the cells a source wrote (integrals_touched_model), now final
        if (INTEG) {
#pragma unroll
            for (int j = 0; j < REG1D_CPT; j++) {
                const int iz = j * bs + tid;
                if (iz < Nz && (g[j] & ID_TOUCHED)) {
                    const T a2 = ex[j] * hy[j];
                    if (Sa) sa[j] += sqrt_of(a2 * a2) * dtT;
                    if (E2) e2[j] += ex[j] * ex[j] * dtT;
                }
            }
        }
        // ---- monitors: from the registers of the thread that holds each point, or from
        // memory (monitors_model)
        if (m.nmon && nment >= 0) {
#pragma unroll
            for (int e = 0; e < REG1D_SRC; e++) {
                if (e >= nment) break;
                const __attribute__((address_space(1))) DevMon &M = G(m.mon)[em_q[e]];
                const int k = em_k[e], c = G(M.cells)[k];
                if (c % bs != tid) continue;
                const int jc = c / bs;
                for (int ci = 0; ci < M.ncomp; ci++) {
                    const int comp = M.comps[ci];
                    T v = (T)0;
#pragma unroll
                    for (int j = 0; j < REG1D_CPT; j++)
                        if (j == jc) v = comp == HY ? hy[j] : comp == DX ? dx[j] : ex[j];
                    if (!M.spectral) G(M.rec)[((size_t)it * M.ncomp + ci) * M.n + k] = (float)v;
                    else G(M.samp)[((size_t)(it % m.dft_block) * M.ncomp + ci) * M.n + k] = (float)v;
                }
            }
        } else if (m.nmon) {
#pragma unroll
            for (int j = 0; j < REG1D_CPT; j++) {
                const int iz = j * bs + tid;
                if (iz < Nz) { Hy[iz] = hy[j]; Dx[iz] = dx[j]; Ex[iz] = ex[j]; }
            }
            __syncthreads();
            monitors_model(m, it, 0, tid, bs, true, sh);
        }
        if (m.nmon && it + 1 - done == m.dft_block) {
            __syncthreads();
            if (dftp) dft_flush_points(m, done, it + 1, ph);
            else dft_flush_model<4>(m, done, it + 1, tid, bs, true, ph, sh);
        }
        it++;
        if (it - done == m.dft_block) done = it;
        if (tid == 0) {
            *m.it = it;
            *m.dftdone = done;
        }
        __syncthreads();   // the E half's reads of xch end before the next step writes it
    }
#pragma unroll
    for (int j = 0; j < REG1D_CPT; j++) {
        const int iz = j * bs + tid;
        if (iz >= Nz) continue;
        Hy[iz] = hy[j]; Dx[iz] = dx[j]; Ex[iz] = ex[j];
        if (j == 0 && iz <= ib0) { G((T *)L0.psiE0)[iz] = peL; G((T *)L0.psiH1)[iz] = phL; }
        if (j == jl && iz >= ib1) { G((T *)L1.psiE0)[iz - ib1] = peR; G((T *)L1.psiH1)[iz - ib1] = phR; }
        if (INTEG && Sa) Sa[iz] = sa[j];
        if (INTEG && E2) E2[iz] = e2[j];
    }
}

void launch_reg1d(const DevModel *m, int nmodels, int nsteps, int mode, bool integ, bool f64, bool hmat,
                  hipStream_t s) {
#define G1(MD, IN, T) \
    hipLaunchKernelGGL((k_reg1d<MD, IN, T>), dim3(nmodels), dim3(REG1D_THREADS), 0, s, m, nsteps, hmat ? 1 : 0)
#define G1P(T)                                                                              \
    if (mode == MODE_VAC) { if (integ) G1(MODE_VAC, true, T); else G1(MODE_VAC, false, T); } \
    else { if (integ) G1(MODE_SIMPLE, true, T); else G1(MODE_SIMPLE, false, T); }
    if (f64) { G1P(double) } else { G1P(float) }
#undef G1P
#undef G1
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
