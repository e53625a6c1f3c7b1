// Device bodies of the per-step work around the half-steps, shared by the kernels that
// launch once per step (step_kernels.hip: any model, one launch per piece of work) and
// the whole-run 1D kernel (kernels1d.hip: one workgroup steps one model through many
// steps).  Each body covers entries gid, gid + stride, ... of its work; `lead` marks the
// one workgroup that does the per-model reductions.  Bodies that reduce over the
// workgroup (block_sum) must be reached by every thread of it.
#pragma once
#include "common.hpp"

namespace mfd {

// ---- soft / hard sources: add_source_kernel! (src/sources.jl:77-119,190-232) ---------
// Entry k of point source S at step it0: writes the component and returns its new value
// (wroteEx: the component is Ex, which the non-finite guard watches).
template <class SR>
__device__ __forceinline__ float point_source_entry(const DevModel &m, SR &S, int it0, int k, bool &wroteEx) {
    const int c = G(S.cells)[k];
    const float w = G(S.wave)[(size_t)it0 * S.n + k];
    float v;
    if (m.f64) {            // the state is FP64: the waveform and the addition too
        const gp<double> F = G(m.F64[S.comp]);
        const double w64 = G(S.wave64)[(size_t)it0 * S.n + k];
        const double v64 = (S.kind == SRC_SOFT) ? F[c] + S.coeff64 * w64 : w64;
        F[c] = v64;
        v = (float)v64;
    } else {
        const gp<float> F = G(m.F[S.comp]);
        v = (S.kind == SRC_SOFT) ? F[c] + S.coeff * w : w;
        F[c] = v;
    }
    wroteEx = S.comp == EX;
    return v;
}

// ---- monitors: update_monitor! (src/monitors.jl:62-66,110-116,164-176,350-359) --------
// Sum over the workgroup (sh: one double per thread).
__device__ inline double block_sum(double v, double *sh) {
    const int t = threadIdx.x;
    sh[t] = v;
    __syncthreads();
    for (int o = blockDim.x / 2; o > 0; o >>= 1) {
        if (t < o) sh[t] += sh[t + o];
        __syncthreads();
    }
    const double r = sh[0];
    __syncthreads();
    return r;
}

// The samples of step it0 of every monitor of m.
__device__ inline void monitors_model(const DevModel &m, int it0, int pp, int gid, int stride, bool lead,
                                      double *sh) {
    for (int q = 0; q < m.nmon; q++) {
        // the record through the global address space: its fields load as scalars
        const __attribute__((address_space(1))) DevMon &M = G(m.mon)[q];
        if (!M.spectral && M.issum) {
            if (!lead) continue;
            for (int c = 0; c < M.ncomp; c++) {
                double acc = 0.0;
                for (int k = threadIdx.x; k < M.n; k += blockDim.x)
                    acc += fld_get64(m, M.comps[c], G(M.cells)[k], pp);
                const double tot = block_sum(acc, sh);
                if (threadIdx.x == 0) G(M.sum)[(size_t)it0 * M.ncomp + c] = tot;
            }
        } else if (!M.spectral) {
            for (int k = gid; k < M.n; k += stride)
                for (int c = 0; c < M.ncomp; c++)
                    G(M.rec)[((size_t)it0 * M.ncomp + c) * M.n + k] = fld_get(m, M.comps[c], G(M.cells)[k], pp);
        } else {
            const int kk = it0 % m.dft_block;
            for (int k = gid; k < M.n; k += stride)
                for (int c = 0; c < M.ncomp; c++)
                    G(M.samp)[((size_t)kk * M.ncomp + c) * M.n + k] = fld_get(m, M.comps[c], G(M.cells)[k], pp);
        }
    }
}

// ---- time-blocked DFT (src/dft.jl:27-51): S += sum_k F_k * 2 exp(-i w (t_k - tshift)) dt
// The buffered samples of steps base .. itn-1 are contracted against an FP64 phase table
// staged in LDS (ph: KMAX * WCH entries).
constexpr int WCH = 16;     // frequencies per pass: 32 FP64 accumulators per lane
constexpr int KMAX = 64;

// WC: frequencies per pass (the per-frequency sums do not depend on it; fewer passes
// need more accumulator registers)
template <int WC = WCH>
__device__ inline void dft_flush_model(const DevModel &m, int base, int itn, int gid, int stride, bool lead,
                                       double2 *ph, double *sh) {
    const int nk = itn - base;
    // the sample ring holds dft_block <= KMAX steps
    if (base < 0 || nk <= 0 || nk > m.dft_block || nk > KMAX) return;
    const double dt = m.dt;
    for (int q = 0; q < m.nmon; q++) {
        const __attribute__((address_space(1))) DevMon &M = G(m.mon)[q];
        if (!M.spectral) continue;
        for (int w0 = 0; w0 < M.nw; w0 += WC) {
            const int nwc = min(WC, M.nw - w0);
            __syncthreads();
            for (int e = threadIdx.x; e < nk * nwc; e += blockDim.x) {
                const int k = e / nwc, w = e % nwc;
                const double t = (double)(base + k) * dt;
                const double th = G(M.w)[w0 + w] * (t - M.tshift);
                double sn, cs;
                sincos(th, &sn, &cs);
                ph[k * WC + w] = make_double2(2.0 * cs * dt, -2.0 * sn * dt);
            }
            __syncthreads();
            if (!M.issum) {
                for (int j = gid; j < M.n; j += stride) {
                    for (int c = 0; c < M.ncomp; c++) {
                        double re[WC], im[WC];
                        for (int w = 0; w < WC; w++) { re[w] = 0.0; im[w] = 0.0; }
                        for (int k = 0; k < nk; k++) {
                            const int kk = (base + k) % m.dft_block;
                            const double F = (double)G(M.samp)[((size_t)kk * M.ncomp + c) * M.n + j];
#pragma unroll
                            for (int w = 0; w < WC; w++) {
                                const double2 p = ph[k * WC + w];
                                re[w] += F * p.x;
                                im[w] += F * p.y;
                            }
                        }
                        for (int w = 0; w < nwc; w++) {
                            double2 *Sp = M.S + ((size_t)c * M.nw + (w0 + w)) * M.n + j;
                            double2 sv = *Sp;
                            sv.x += re[w];
                            sv.y += im[w];
                            *Sp = sv;
                        }
                    }
                }
            } else if (lead) {
                for (int c = 0; c < M.ncomp; c++) {
                    for (int k = 0; k < nk; k++) {
                        const int kk = (base + k) % m.dft_block;
                        double acc = 0.0;
                        for (int j = threadIdx.x; j < M.n; j += blockDim.x)
                            acc += (double)G(M.samp)[((size_t)kk * M.ncomp + c) * M.n + j];
                        const double Fs = block_sum(acc, sh);
                        if ((int)threadIdx.x < nwc) {
                            const double2 p = ph[k * WC + threadIdx.x];
                            double2 *Sp = M.S + (size_t)c * M.nw + w0 + threadIdx.x;
                            double2 sv = *Sp;
                            sv.x += Fs * p.x;
                            sv.y += Fs * p.y;
                            *Sp = sv;
                        }
                        __syncthreads();
                    }
                }
            }
        }
    }
}

// The same block for point (not summed) spectral monitors with nk * nw <= KMAX * WCH,
// one thread per (point, component, frequency) and the phase table of every (step,
// frequency) in one pass: the sums, term for term, of dft_flush_model.
__device__ inline void dft_flush_points(const DevModel &m, int base, int itn, double2 *ph) {
    const int nk = itn - base;
    if (base < 0 || nk <= 0 || nk > m.dft_block || nk > KMAX) return;
    const double dt = m.dt;
    for (int q = 0; q < m.nmon; q++) {
        const __attribute__((address_space(1))) DevMon &M = G(m.mon)[q];
        if (!M.spectral) continue;
        const int nw = M.nw;
        __syncthreads();
        for (int e = threadIdx.x; e < nk * nw; e += blockDim.x) {
            const int k = e / nw, w = e % nw;
            const double t = (double)(base + k) * dt;
            const double th = G(M.w)[w] * (t - M.tshift);
            double sn, cs;
            sincos(th, &sn, &cs);
            ph[k * nw + w] = make_double2(2.0 * cs * dt, -2.0 * sn * dt);
        }
        __syncthreads();
        for (int e = threadIdx.x; e < M.n * M.ncomp * nw; e += blockDim.x) {
            const int w = e % nw, c = (e / nw) % M.ncomp, j = e / (nw * M.ncomp);
            double re = 0.0, im = 0.0;
            for (int k = 0; k < nk; k++) {
                const int kk = (base + k) % m.dft_block;
                const double F = (double)G(M.samp)[((size_t)kk * M.ncomp + c) * M.n + j];
                const double2 p = ph[k * nw + w];
                re += F * p.x;
                im += F * p.y;
            }
            double2 *Sp = M.S + ((size_t)c * nw + w) * M.n + j;
            double2 sv = *Sp;
            sv.x += re;
            sv.y += im;
            *Sp = sv;
        }
    }
}

// ---- integral variables of source-written cells: update_integral_variables!
// (src/outputs.jl:182-191,318-327,475-485) for the cells the E kernels skipped, after
// every source of the step has been applied.
__device__ inline void integrals_touched_model(const DevModel &m, int ndims, int gid, int stride) {
    const float dt = m.dtf;
    const gp<float> Sa = G(m.Sa), E2 = G(m.E2);
    for (int k = gid; k < m.ntouched; k += stride) {
        const long long i = G(m.touched)[k] & TOUCHED_CELL;
        if (ndims == 3) {
            const float hx = G(m.F[HX])[i], hy = G(m.F[HY])[i], hz = G(m.F[HZ])[i];
            const float ex = G(m.F[EX])[i], ey = G(m.F[EY])[i], ez = G(m.F[EZ])[i];
            if (Sa) {
                const float a = ey * hz - ez * hy, b = ez * hx - ex * hz, c = ex * hy - ey * hx;
                Sa[i] += sa_sqrt(a * a + b * b + c * c) * dt;
            }
            if (E2) E2[i] += (ex * ex + ey * ey + ez * ez) * dt;
        } else if (ndims == 2) {
            const float hy = G(m.F[HY])[i], ex = G(m.F[EX])[i], ez = G(m.F[EZ])[i];
            if (Sa) {
                const float a = -ez * hy, b = ex * hy;
                Sa[i] += sa_sqrt(a * a + b * b) * dt;
            }
            if (E2) E2[i] += (ex * ex + ez * ez) * dt;
        } else if (m.f64) {
            const double hy = G(m.F64[HY])[i], ex = G(m.F64[EX])[i];
            if (m.Sa64) {
                const double a = ex * hy;
                G(m.Sa64)[i] += sqrt(a * a) * m.dt;
            }
            if (m.E2_64) G(m.E2_64)[i] += ex * ex * m.dt;
        } else {
            const float hy = G(m.F[HY])[i], ex = G(m.F[EX])[i];
            if (Sa) {
                const float a = ex * hy;
                Sa[i] += sa_sqrt(a * a) * dt;
            }
            if (E2) E2[i] += ex * ex * dt;
        }
    }
}

}  // namespace mfd
