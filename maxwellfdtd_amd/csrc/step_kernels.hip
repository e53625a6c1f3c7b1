// Per-step kernels around the H/E half-steps: soft/hard sources, TFSF (auxiliary 1D
// run + face corrections), field/spectral monitors, the time-blocked DFT, integral
// variables and the device step counter.  Every kernel takes the DevModel array of a
// batch; blockIdx.y (or blockIdx.x for one-block-per-model kernels) selects the model,
// and a model that finished its Nt steps is skipped.
#include "common.hpp"
#include "material.hpp"
#include "tfsf.hpp"
#include "step_dev.hpp"

#include <algorithm>

namespace mfd {

// ---- soft / hard sources: add_source_kernel! (src/sources.jl:77-119,190-232) ---------
__global__ __launch_bounds__(256) void k_point_sources(const DevModel *__restrict__ models,
                                                       int slot) {
    const DevModel &m = models[blockIdx.y];
    if (slot >= m.nsrc) return;
    const DevSource &S = m.src[slot];
    if (S.kind == SRC_TFSF) return;
    const int it0 = *m.it;
    if (it0 >= m.Nt) return;
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    bool wroteEx = false;
    float v = 0.0f;
    if (k < S.n) v = point_source_entry(m, S, it0, k, wroteEx);
    check_finite(v, m.flag, wroteEx);
}

void launch_point_sources(const DevModel *m, int nmodels, int slot, int maxn, hipStream_t s) {
    const int nb = (maxn + 255) / 256;
    hipLaunchKernelGGL(k_point_sources, dim3(nb, nmodels), dim3(256), 0, s, m, slot);
}

// ---- TFSF auxiliary run --------------------------------------------------------------
// One step of the x/y-uniform record run in FP64 (the transverse differences of the 3D
// update vanish, so update_H_kernel!/update_E_kernel! 3D reduce to these z recursions,
// src/models.jl:735-1162): H half, E half, then the aux soft sources.
__device__ void aux_advance(const DevSource &S, int Nz, int Nt, int bcl, int bcr, int it0) {
    double *Hx = S.aux, *Hy = S.aux + Nz, *Dx = S.aux + 2 * Nz, *Dy = S.aux + 3 * Nz;
    double *Ex = S.aux + 4 * Nz, *Ey = S.aux + 5 * Nz;
    const int nbm = S.aux_nbmax;
    const double Mh = S.Mh0, Me = S.Me0, Md1 = S.Md10, Md2 = S.Md20, dz = 1.0 / S.idz;
    for (int iz = threadIdx.x; iz < Nz; iz += blockDim.x) {
        double ex1, ey1;
        if (iz + 1 < Nz) { ex1 = Ex[iz + 1]; ey1 = Ey[iz + 1]; }
        else if (bcr == BC_PERIODIC) { ex1 = Ex[0]; ey1 = Ey[0]; }
        else if (bcr == BC_DIRICHLET) { ex1 = 0.0; ey1 = 0.0; }
        else { ex1 = Ex[iz]; ey1 = Ey[iz]; }
        double dExz = (ex1 - Ex[iz]) / dz;
        double dEyz = (ey1 - Ey[iz]) / dz;
        double hx = Hx[iz], hy = Hy[iz];
        for (int f = 0; f < 2; f++) {
            if (!S.aux_present[f]) continue;
            const bool in = f == 0 ? iz <= S.aux_ib[0] : iz >= S.aux_ib[1];
            if (!in) continue;
            const int p = f == 0 ? iz : iz - S.aux_ib[1];
            double *pEx = S.auxpsi + (size_t)(f * 4 + 0) * nbm;
            double *pEy = S.auxpsi + (size_t)(f * 4 + 1) * nbm;
            const double A = S.auxA[f][p], B = S.auxB[f][p], K = S.auxK[f][p];
            pEx[p] = B * pEx[p] + A * dExz;
            pEy[p] = B * pEy[p] + A * dEyz;
            hx += Mh * pEy[p];
            hy -= Mh * pEx[p];
            dExz = dExz / K;
            dEyz = dEyz / K;
        }
        hx -= Mh * (0.0 - dEyz);
        hy -= Mh * (dExz - 0.0);
        Hx[iz] = hx;
        Hy[iz] = hy;
    }
    __syncthreads();
    for (int iz = threadIdx.x; iz < Nz; iz += blockDim.x) {
        double hx0, hy0;
        if (iz > 0) { hx0 = Hx[iz - 1]; hy0 = Hy[iz - 1]; }
        else if (bcl == BC_PERIODIC) { hx0 = Hx[Nz - 1]; hy0 = Hy[Nz - 1]; }
        else if (bcl == BC_DIRICHLET) { hx0 = 0.0; hy0 = 0.0; }
        else { hx0 = Hx[iz]; hy0 = Hy[iz]; }
        double dHxz = (Hx[iz] - hx0) / dz;
        double dHyz = (Hy[iz] - hy0) / dz;
        double dx = Dx[iz], dy = Dy[iz];
        for (int f = 0; f < 2; f++) {
            if (!S.aux_present[f]) continue;
            const bool in = f == 0 ? iz <= S.aux_ib[0] : iz >= S.aux_ib[1];
            if (!in) continue;
            const int p = f == 0 ? iz : iz - S.aux_ib[1];
            double *pHx = S.auxpsi + (size_t)(f * 4 + 2) * nbm;
            double *pHy = S.auxpsi + (size_t)(f * 4 + 3) * nbm;
            const double A = S.auxA[f][p], B = S.auxB[f][p], K = S.auxK[f][p];
            pHx[p] = B * pHx[p] + A * dHxz;
            pHy[p] = B * pHy[p] + A * dHyz;
            dx -= Md2 * pHy[p];
            dy += Md2 * pHx[p];
            dHxz = dHxz / K;
            dHyz = dHyz / K;
        }
        dx = Md1 * dx + Md2 * (0.0 - dHyz);
        dy = Md1 * dy + Md2 * (dHxz - 0.0);
        Dx[iz] = dx;
        Dy[iz] = dy;
        Ex[iz] = dx * Me;
        Ey[iz] = dy * Me;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        for (int k = 0; k < S.aux_n; k++) {
            const int iz = S.aux_iz[k];
            const double v = S.aux_coeff[k] * S.aux_wave[(size_t)k * Nt + it0];
            double *dst;
            switch (S.aux_comp[k]) {
                case HX: dst = Hx; break;
                case HY: dst = Hy; break;
                case DX: dst = Dx; break;
                case DY: dst = Dy; break;
                case EX: dst = Ex; break;
                default: dst = Ey; break;
            }
            dst[iz] = dst[iz] + v;
        }
    }
    __syncthreads();
}

__global__ __launch_bounds__(256) void k_tfsf_aux(const DevModel *__restrict__ models,
                                                  int slot) {
    const DevModel &m = models[blockIdx.x];
    if (slot >= m.nsrc) return;
    const DevSource &S = m.src[slot];
    if (S.kind != SRC_TFSF || S.mode != 1 || S.auxtab) return;
    const int it0 = *m.it;
    if (it0 >= m.Nt) return;
    aux_advance(S, m.Nz, m.Nt, m.bcs[4], m.bcs[5], it0);
}

// The whole auxiliary run at once: the incident field does not depend on the model's
// fields, so its Hx, Hy, Ex, Ey after every step are tabulated (FP32, the precision the
// face corrections use) into tab[Nt][4][Nz] before the first step.
__global__ __launch_bounds__(256) void k_tfsf_aux_table(const DevSource *__restrict__ src,
                                                        int Nz, int Nt, int bcl, int bcr,
                                                        float *__restrict__ tab) {
    const DevSource &S = *src;
    for (int it0 = 0; it0 < Nt; it0++) {
        aux_advance(S, Nz, Nt, bcl, bcr, it0);
        float *row = tab + (size_t)it0 * 4 * Nz;
        for (int iz = threadIdx.x; iz < Nz; iz += blockDim.x) {
            row[iz] = (float)S.aux[iz];
            row[Nz + iz] = (float)S.aux[Nz + iz];
            row[2 * Nz + iz] = (float)S.aux[4 * Nz + iz];
            row[3 * Nz + iz] = (float)S.aux[5 * Nz + iz];
        }
        __syncthreads();
    }
}

void launch_tfsf_aux_table(const DevSource *S, int Nz, int Nt, int bcl, int bcr, float *tab,
                           hipStream_t s) {
    hipLaunchKernelGGL(k_tfsf_aux_table, dim3(1), dim3(256), 0, s, S, Nz, Nt, bcl, bcr, tab);
}

// ---- TFSF face corrections: add_source!(::TFSFSource3D) (src/sources.jl:713-776) -------
// Face f = 2*axis + side.  (u, v) are the tangential axes (x face: y,z; y face: x,z;
// z face: x,y); the face tables are H_u, H_v, E_u, E_v.  1D and 2D models use faces
// 4,5 (resp. 0,1,4,5) with degenerate tangential ranges.

// Both faces of one axis run in one launch (they write disjoint planes); the three
// axes run in three launches in the reference's face order.
__global__ __launch_bounds__(256) void k_tfsf_axis(const DevModel *__restrict__ models,
                                                   int slot, int axis) {
    const DevModel &m = models[blockIdx.y];
    if (slot >= m.nsrc) return;
    const DevSource &S = m.src[slot];
    if (S.kind != SRC_TFSF) return;
    const int it0 = *m.it;
    if (it0 >= m.Nt) return;
    const int ua = axis == 0 ? 1 : 0, va = axis == 2 ? 1 : 2;
    const int U1 = S.box[2 * ua], U2 = S.box[2 * ua + 1];
    const int V1 = S.box[2 * va], V2 = S.box[2 * va + 1];
    const int N1 = S.box[2 * axis], N2 = S.box[2 * axis + 1];
    const int nu = U2 - U1 + 1, nv = V2 - V1 + 1;
    int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= 2 * nu * nv) return;
    const int side = e >= nu * nv ? 1 : 0;
    e -= side * nu * nv;
    const int f = 2 * axis + side;
    const int u = U1 + e % nu, v = V1 + e / nu;
    const int nH = side ? N2 : N1 - 1, nE = side ? N2 : N1;
    const float sg = side ? -1.0f : 1.0f;
    const float s0 = (axis == 1 ? 1.0f : -1.0f) * sg;
    int cHu, cHv, cEu, cEv, cDu, cDv;
    if (axis == 0) { cHu = HY; cHv = HZ; cEu = EY; cEv = EZ; cDu = DY; cDv = DZ; }
    else if (axis == 1) { cHu = HX; cHv = HZ; cEu = EX; cEv = EZ; cDu = DX; cDv = DZ; }
    else { cHu = HX; cHv = HY; cEu = EX; cEv = EY; cDu = DX; cDv = DY; }
    const int tabc[4] = {cHu, cHv, cEu, cEv};
    const int Nu = U2 - U1, Nv = V2 - V1;
    auto inc = [&](int k) -> float {
        if (S.mode == 1) {
            int z;
            if (axis == 2) z = side ? N2 : (k < 2 ? N1 - 1 : N1);
            else z = v;
            return aux_value(S, m.Nz, tabc[k], z, it0);
        }
        const float *T = S.tab[4 * f + k];
        if (!T) return 0.0f;
        const int na = (k == 0 || k == 3) ? Nu + 1 : Nu;
        const int nb = (k == 0 || k == 3) ? Nv : Nv + 1;
        return T[(u - U1) + (size_t)na * ((v - V1) + (size_t)nb * it0)];
    };
    auto cell = [&](int n) -> long long {
        int ix, iy, iz;
        if (axis == 0) { ix = n; iy = u; iz = v; }
        else if (axis == 1) { ix = u; iy = n; iz = v; }
        else { ix = u; iy = v; iz = n; }
        return ix + m.sy * iy + m.sz * iz;
    };
    const float cm = S.cm[axis], ce = S.ce[axis], cd = S.cd[axis];
    // In the auxiliary (normal-incidence) mode the Hz and Ez incident fields are
    // identically zero; their corrections add exact zeros and are skipped.
    auto live = [&](int k) { return S.mode == 0 || (tabc[k] != HZ && tabc[k] != EZ); };
    // an FP64 model: incident values and coefficients in FP64, added to the FP64 state
    auto inc64 = [&](int k) -> double {
        if (S.mode == 1) {
            int z;
            if (axis == 2) z = side ? N2 : (k < 2 ? N1 - 1 : N1);
            else z = v;
            const int comp = tabc[k];
            const int slot = comp == HX ? 0 : comp == HY ? 1 : comp == EX ? 4 : comp == EY ? 5 : -1;
            return slot < 0 ? 0.0 : S.aux[(size_t)slot * m.Nz + z];
        }
        const double *T = S.tab64[4 * f + k];
        if (!T) return 0.0;
        const int na = (k == 0 || k == 3) ? Nu + 1 : Nu;
        const int nb = (k == 0 || k == 3) ? Nv : Nv + 1;
        return T[(u - U1) + (size_t)na * ((v - V1) + (size_t)nb * it0)];
    };
    if (m.f64) {
        const double cm64 = S.cm64[axis], ce64 = S.ce64[axis], cd64 = S.cd64[axis];
        const double s064 = s0;
        if (u <= U2 && v <= V2 - 1 && live(3)) m.F64[cHu][cell(nH)] += (s064 * cm64) * inc64(3);
        if (u <= U2 - 1 && v <= V2 && live(2)) m.F64[cHv][cell(nH)] += (-s064 * cm64) * inc64(2);
        if (u <= U2 - 1 && v <= V2 && live(1)) {
            const double iv = inc64(1);
            const long long c = cell(nE);
            m.F64[cEu][c] += (-s064 * ce64) * iv;
            m.F64[cDu][c] += (-s064 * cd64) * iv;
        }
        if (u <= U2 && v <= V2 - 1 && live(0)) {
            const double iv = inc64(0);
            const long long c = cell(nE);
            m.F64[cEv][c] += (s064 * ce64) * iv;
            m.F64[cDv][c] += (s064 * cd64) * iv;
        }
        return;
    }
    if (u <= U2 && v <= V2 - 1 && live(3)) {     // H_u += (s0) Cm/d * inc E_v
        const float iv = inc(3);
        float *F = m.F[cHu];
        const long long c = cell(nH);
        F[c] = __builtin_fmaf(s0 * cm, iv, F[c]);
    }
    if (u <= U2 - 1 && v <= V2 && live(2)) {     // H_v += (-s0) Cm/d * inc E_u
        const float iv = inc(2);
        float *F = m.F[cHv];
        const long long c = cell(nH);
        F[c] = __builtin_fmaf(-s0 * cm, iv, F[c]);
    }
    if (u <= U2 - 1 && v <= V2 && live(1)) {     // E_u, D_u += (-s0) Ce/d * inc H_v
        const float iv = inc(1);
        const long long c = cell(nE);
        float *E = m.F[cEu];
        float *D = m.F[cDu];
        E[c] = __builtin_fmaf(-s0 * ce, iv, E[c]);
        D[c] = __builtin_fmaf(-s0 * cd, iv, D[c]);
    }
    if (u <= U2 && v <= V2 - 1 && live(0)) {     // E_v, D_v += (s0) Ce/d * inc H_u
        const float iv = inc(0);
        const long long c = cell(nE);
        float *E = m.F[cEv];
        float *D = m.F[cDv];
        E[c] = __builtin_fmaf(s0 * ce, iv, E[c]);
        D[c] = __builtin_fmaf(s0 * cd, iv, D[c]);
    }
}

// ---- TFSF corrections and integral variables of the touched cells, 3D ---------------
// One lane per cell of the model's touched list (every cell a TFSF face correction can
// write).  The lane applies every correction of source `slot` that lands on its cell, in
// the reference's face order (x, y, z; lower side, upper side) and with the statements
// of k_tfsf_axis, then accumulates the cell's Sa/E2 (which the E kernel skipped) and
// checks Ex.  Used when the TFSF source is a model's only source.  When the model's E
// kernel finishes x-face cells itself (m.xtfsf set), such a cell only gets its pending
// H correction here.
__global__ __launch_bounds__(256) void k_tfsf_gather(const DevModel *__restrict__ models,
                                                     int slot, int integ, int pp) {
    const DevModel &m = models[blockIdx.y];
    const int it0 = *m.it;
    if (it0 >= m.Nt) return;
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    bool active = k < m.ntouched;
    float exf = 0.0f;
    if (active) {
        // the source record through the global address space: its fields are uniform
        // and load as scalars
        const __attribute__((address_space(1))) DevSource &S =
            *(const __attribute__((address_space(1))) DevSource *)(m.src + slot);
        const int tk = G(m.touched)[k];
        const unsigned i = (unsigned)(tk & TOUCHED_CELL);
        const unsigned qz = udiv_magic(i, m.divxy), rz = i - qz * m.divxy.w;
        const unsigned qy = udiv_magic(rz, m.divx);
        const int c3[3] = {(int)(rz - qy * m.divx.w), (int)qy, (int)qz};
        bool xh = false;
        if (m.xtfsf) xh = tfsf_xhandled(S, c3[0], c3[1], c3[2], (tk & TOUCHED_STATEFUL) != 0);
        // pp: a fused step ran and its buffer swap comes after this launch, so this
        // step's H and E are in Falt (D stays in place)
        float *fp[9];
#pragma unroll
        for (int c = 0; c < 9; c++) fp[c] = pp && (c < DX || c > DZ) ? m.Falt[c] : m.F[c];
        float f[9];
        if (xh) {
            f[HY] = G((const float *)fp[HY])[i];
            f[HZ] = G((const float *)fp[HZ])[i];
        } else {
#pragma unroll
            for (int c = 0; c < 9; c++) f[c] = G((const float *)fp[c])[i];
        }
        const unsigned dirty = tfsf_cell_fix(S, m.Nz, it0, c3, xh, f);
#pragma unroll
        for (int c = 0; c < 9; c++)
            if (dirty & (1u << c)) G(fp[c])[i] = f[c];
        if (xh) {
            active = false;     // integrals and the Ex check were done by the E kernel
        } else {
            if (integ) {   // src/outputs.jl:475-485
                const float dt = m.dtf;
                if (m.Sa) G(m.Sa)[i] = sa_step3(G(m.Sa)[i], f[EX], f[EY], f[EZ], f[HX], f[HY], f[HZ], dt);
                if (m.E2) G(m.E2)[i] = e2_step3(G(m.E2)[i], f[EX], f[EY], f[EZ], dt);
            }
            exf = f[EX];
        }
    }
    check_finite(exf, m.flag, active);
}

void launch_tfsf_gather(const DevModel *m, int nmodels, int slot, int maxn, bool integ,
                        hipStream_t s, bool pp) {
    const int nb = std::max(1, (maxn + 255) / 256);
    hipLaunchKernelGGL(k_tfsf_gather, dim3(nb, nmodels), dim3(256), 0, s, m, slot, integ ? 1 : 0, pp ? 1 : 0);
}

void launch_tfsf_aux(const DevModel *m, int nmodels, int slot, hipStream_t s) {
    hipLaunchKernelGGL(k_tfsf_aux, dim3(nmodels), dim3(256), 0, s, m, slot);
}

void launch_tfsf(const DevModel *m, int nmodels, int slot, int ndims, int maxface,
                 hipStream_t s, bool aux) {
    if (aux) launch_tfsf_aux(m, nmodels, slot, s);
    const int nb = (2 * maxface + 255) / 256;
    for (int axis = 0; axis < 3; axis++) {
        if (ndims == 1 && axis < 2) continue;
        if (ndims == 2 && axis == 1) continue;
        hipLaunchKernelGGL(k_tfsf_axis, dim3(nb, nmodels), dim3(256), 0, s, m, slot, axis);
    }
}

// ---- monitors: update_monitor! (src/monitors.jl:62-66,110-116,164-176,350-359) --------
__global__ __launch_bounds__(256) void k_monitors(const DevModel *__restrict__ models, int pp) {
    __shared__ double sh[256];
    const DevModel &m = models[blockIdx.y];
    const int it0 = *m.it;
    if (it0 >= m.Nt) return;
    monitors_model(m, it0, pp, blockIdx.x * blockDim.x + threadIdx.x, gridDim.x * blockDim.x, blockIdx.x == 0,
                   sh);
}

void launch_monitors(const DevModel *m, int nmodels, int maxn, hipStream_t s, bool pp) {
    int nb = (maxn + 255) / 256;
    if (nb < 1) nb = 1;
    if (nb > 1024) nb = 1024;
    hipLaunchKernelGGL(k_monitors, dim3(nb, nmodels), dim3(256), 0, s, m, pp ? 1 : 0);
}

// ---- time-blocked DFT (src/dft.jl:27-51): S += sum_k F_k * 2 exp(-i w (t_k - tshift)) dt
// The buffered samples of steps dftdone .. itn-1 are contracted against an FP64 phase
// table staged in LDS.  Each model keeps its own dftdone on the device, so models of one
// run need not be aligned (a batch of uneven Nt, a server's staggered arrivals):
//  - in the step (force = 0, after the monitors and before the step counter advances): a
//    model whose block is complete with this step's sample (it + 1 - dftdone ==
//    dft_block) is contracted; k_finish then moves its dftdone to it + 1;
//  - at read-back (force = 1): every model's pending partial block (it - dftdone steps),
//    after which k_dft_commit moves dftdone to it.
// A model that has finished (it >= Nt) is never contracted in the step.
__global__ __launch_bounds__(256) void k_dft_flush(const DevModel *__restrict__ models, int force) {
    __shared__ double2 ph[KMAX * WCH];
    __shared__ double sh[256];
    const DevModel &m = models[blockIdx.y];
    if (m.nmon == 0) return;
    const int it0 = *m.it;
    const int base = *m.dftdone;
    int itn;
    if (force) {
        itn = it0;
    } else {
        if (it0 >= m.Nt) return;
        itn = it0 + 1;
        if (itn - base != m.dft_block) return;
    }
    dft_flush_model(m, base, itn, blockIdx.x * blockDim.x + threadIdx.x, gridDim.x * blockDim.x, blockIdx.x == 0,
                    ph, sh);
}

void launch_dft_flush(const DevModel *m, int nmodels, int maxn, bool force, hipStream_t s) {
    int nb = (maxn + 255) / 256;
    if (nb < 1) nb = 1;
    if (nb > 2048) nb = 2048;
    hipLaunchKernelGGL(k_dft_flush, dim3(nb, nmodels), dim3(256), 0, s, m, force ? 1 : 0);
}

// After a forced flush: every model's samples up to its step counter are contracted.
__global__ void k_dft_commit(const DevModel *__restrict__ models, int n) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k < n) *models[k].dftdone = *models[k].it;
}

void launch_dft_commit(const DevModel *m, int nmodels, hipStream_t s) {
    hipLaunchKernelGGL(k_dft_commit, dim3((nmodels + 255) / 256), dim3(256), 0, s, m, nmodels);
}

// ---- integral variables of source-written cells: update_integral_variables!
// (src/outputs.jl:182-191,318-327,475-485) for the cells the E kernels skipped, after
// every source of the step has been applied.
__global__ __launch_bounds__(256) void k_integrals_touched(const DevModel *__restrict__ models,
                                                           int ndims) {
    const DevModel &m = models[blockIdx.y];
    if (*m.it >= m.Nt || m.int_flags == 0) return;
    integrals_touched_model(m, ndims, blockIdx.x * blockDim.x + threadIdx.x, gridDim.x * blockDim.x);
}

void launch_integrals_touched(const DevModel *m, int nmodels, int maxn, int ndims, hipStream_t s) {
    int nb = (maxn + 255) / 256;
    if (nb < 1) nb = 1;
    if (nb > 4096) nb = 4096;
    hipLaunchKernelGGL(k_integrals_touched, dim3(nb, nmodels), dim3(256), 0, s, m, ndims);
}

// ---- step counter ----------------------------------------------------------------------
// Advances every unfinished model's step counter; a model whose DFT block was completed
// (and contracted by k_dft_flush) in this step moves its dftdone along.
__global__ void k_finish(const DevModel *__restrict__ models, int n) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k < n) {
        const DevModel &m = models[k];
        const int it = *m.it;
        if (it < m.Nt) {
            *m.it = it + 1;
            if (it + 1 - *m.dftdone == m.dft_block) *m.dftdone = it + 1;
        }
    }
}

// Non-finite flags of every model of a run gathered into one array for one copy.
__global__ void k_collect_flags(const DevModel *__restrict__ models, int n, int *__restrict__ out) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k < n) out[k] = *models[k].flag;
}

void launch_collect_flags(const DevModel *m, int nmodels, int *out, hipStream_t s) {
    hipLaunchKernelGGL(k_collect_flags, dim3((nmodels + 255) / 256), dim3(256), 0, s, m, nmodels, out);
}

void launch_finish(const DevModel *m, int nmodels, hipStream_t s) {
    hipLaunchKernelGGL(k_finish, dim3((nmodels + 255) / 256), dim3(256), 0, s, m, nmodels);
}

}  // namespace mfd
