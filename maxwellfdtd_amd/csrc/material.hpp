// Per-cell material update of the E half-step, shared by the 1D/2D/3D kernels:
// ADE poles, plasma (FP64 density and rate), Kerr Pade and E = (D - sum P) * Me.
// Statement order follows update_E_kernel! 3D (src/models.jl:1043-1154).
#pragma once
#include "common.hpp"

namespace mfd {

constexpr double DBL_EPS_ = 2.220446049250313e-16;   // eps(Float64), src/models.jl:1106

// ionrate(I) in FP64: power law sigma*I^K, or the log-log table of IonizationRate
// (examples/5_nonlinear_propagation/IonizationRate.jl:30-52).
template <class MR>
__device__ inline double ionrate(MR &mt, double I) {
    if (!(I > 0.0)) return 0.0;
    if (mt.rate_kind == 1) return mt.rate_c * pow(I, mt.rate_k);
    if (mt.rate_kind == 2) {
        const gp<const double> x = G(mt.rate_x), y = G(mt.rate_y);
        const int n = mt.rate_n;
        const double xl = log10(I);
        int i;
        if (xl <= x[0]) i = 0;
        else if (xl >= x[n - 1]) i = n - 2;
        else {
            int lo = 0, hi = n - 1;
            while (lo < hi) {
                int mid = (lo + hi) >> 1;
                if (x[mid] >= xl) hi = mid; else lo = mid + 1;
            }
            i = lo - 1;
        }
        const double dydx = (y[i + 1] - y[i]) / (x[i + 1] - x[i]);
        return exp10(y[i] + dydx * (xl - x[i]));
    }
    return 0.0;
}

// d[c]: updated D; e[c]: on entry E^n (read only when the material has ADE/plasma
// state), on exit E^{n+1}.  s is the cell's state slot within its material; par the
// model's completed steps mod 2: half par of the P and Pp arrays holds step n, the other
// half step n - 1, which this update overwrites with step n + 1 (a Debye pole, B = 0,
// reads no step n - 1).  mt is a
// DevMat (T = float) or a DevMat64 (T = double), generic or global-address-space; dtT
// is dt in T.  Whatever T, the pole and plasma-current recursions run in FP64 on FP64
// state and coefficients, and D - sum P is formed in FP64: a Drude pole (and the plasma
// current) has a root at exactly 1, whose FP32 rounding of the coefficients (A + B off 1
// by ~1e-7) and of the state the other root, 1 - gamma dt, amplifies ~1/(gamma dt) times
// into a drift of P and D over thousands of steps.  The plasma density and rate are
// FP64 as well.
template <int NC, class T, class MR>
__device__ __forceinline__ void material_e(MR &mt, int s, int par, const T (&d)[NC], T (&e)[NC],
                                           double dt, T dtT) {
    const int flags = mt.flags;
    double sumP[NC];
#pragma unroll
    for (int c = 0; c < NC; c++) sumP[c] = 0.0;
    if (flags & MAT_DISP) {
        const int nq = mt.nq, cnt = mt.cnt;
        const size_t half = (size_t)nq * NC * cnt;
        const gp<double> cur = G(mt.P) + (par ? half : 0), old = G(mt.P) + (par ? 0 : half);
        for (int q = 0; q < nq; q++) {
            const double A = G(mt.Aq)[q], B = G(mt.Bq)[q], Cc = G(mt.Cq)[q];
#pragma unroll
            for (int c = 0; c < NC; c++) {
                // pole state is read and written once per step: streamed past the caches
                const size_t j = (size_t)(q * NC + c) * cnt + s;
                const double p = __builtin_nontemporal_load(cur + j);
                const double pm = B != 0.0 ? __builtin_nontemporal_load(old + j) : 0.0;
                const double pn = A * p + B * pm + Cc * (double)e[c];
                __builtin_nontemporal_store(pn, old + j);
                sumP[c] += pn;
            }
        }
    }
    if (flags & MAT_PLASMA) {
        const int cnt = mt.cnt;
        double E2 = 0.0;
#pragma unroll
        for (int c = 0; c < NC; c++) E2 += (double)e[c] * (double)e[c];
        const double II = mt.ksi * E2;
        gp<v2d> prd = (gp<v2d>)(mt.rhod + s);
        v2d rd = __builtin_nontemporal_load(prd);
        const double cp = mt.Cprho0 * rd.x;
        double J[NC];
        const size_t half = (size_t)NC * cnt;
        const gp<double> cur = G(mt.Pp) + (par ? half : 0), old = G(mt.Pp) + (par ? 0 : half);
#pragma unroll
        for (int c = 0; c < NC; c++) {
            const size_t j = (size_t)c * cnt + s;
            const double p = __builtin_nontemporal_load(cur + j), pm = __builtin_nontemporal_load(old + j);
            const double pn = (double)mt.Ap * p + (double)mt.Bp * pm + cp * (double)e[c];
            __builtin_nontemporal_store(pn, old + j);
            sumP[c] += pn;
            J[c] = (pn - pm) / (2.0 * dt);
        }
        const double invE2 = (E2 >= DBL_EPS_) ? 1.0 / E2 : 0.0;
        const double ca = mt.Marho0 * rd.y * invE2;
#pragma unroll
        for (int c = 0; c < NC; c++) {
            gp<double> pa = G(mt.Pa) + (size_t)c * cnt + s;
            const double v = __builtin_nontemporal_load(pa) + ca * (double)e[c];
            __builtin_nontemporal_store(v, pa);
            sumP[c] += v;
        }
        const double R1 = ionrate(mt, II);
        const double R2 = mt.Rava * E2;
        double rho;
        if (R2 == 0.0) {
            rho = 1.0 - (1.0 - rd.x) * exp(-R1 * dt);
        } else {
            const double R12 = R1 - R2;
            rho = R1 / R12 - (R1 / R12 - rd.x) * exp(-R12 * dt);
        }
        rd.x = rho;
        rd.y = R1 * (1.0 - rho);
        __builtin_nontemporal_store(rd, prd);
        T je = (T)0;
#pragma unroll
        for (int c = 0; c < NC; c++) je += (T)J[c] * e[c];
        __builtin_nontemporal_store(__builtin_nontemporal_load(&G(mt.JE)[s]) + je * dtT, &G(mt.JE)[s]);
    }
    const double Me = (double)mt.Me;
    if (flags & MAT_KERR) {
        const T K2 = (T)mt.K2, K3 = (T)mt.K3, MeT = (T)mt.Me;
#pragma unroll
        for (int c = 0; c < NC; c++) {
            const T u = (T)((double)d[c] - sumP[c]);
            e[c] = ((T)1 + K2 * u + (T)2 * K3 * (u * u)) /
                   ((T)1 + (T)2 * K2 * u + (T)3 * K3 * (u * u)) * u * MeT;
        }
    } else {
#pragma unroll
        for (int c = 0; c < NC; c++) e[c] = (T)(((double)d[c] - sumP[c]) * Me);
    }
}

// Non-finite guard on Ex (src/models.jl:202-209): one atomic per wave that saw one.
template <class T>
__device__ __forceinline__ void check_finite(T ex, int *flag, bool active) {
    const bool bad = active && !isfinite(ex);
    if (__any(bad)) {
        if (bad) atomicOr(flag, isnan(ex) ? 1 : 2);
    }
}

__device__ __forceinline__ float bcval(int bc, float periodic, float own) {
    return bc == BC_PERIODIC ? periodic : (bc == BC_DIRICHLET ? 0.0f : own);
}

}  // namespace mfd
