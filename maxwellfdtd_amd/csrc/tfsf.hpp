// TFSF face-correction helpers shared by the face kernels and the E half-step epilogue
// (add_source!(::TFSFSource3D), src/sources.jl:713-776).
//
// Face f = 2*axis + side.  (u, v) are the tangential axes (x face: y, z; y face: x, z;
// z face: x, y); the replay tables per face are H_u, H_v, E_u, E_v.  On the lower side
// the H correction lands on plane N1-1 and the E/D correction on N1; on the upper side
// both land on plane N2.
#pragma once
#include "common.hpp"

namespace mfd {

// incident value of the aux run (tabulated or live) at plane z, component comp; S may
// be a generic or a global-address-space reference
template <class SR>
__device__ inline float aux_value(SR &S, int Nz, int comp, int z, int it0) {
    const float *tab = S.auxtab;
    if (tab) {
        const int k = comp == HX ? 0 : comp == HY ? 1 : comp == EX ? 2 : 3;
        return G(tab)[((size_t)it0 * 4 + k) * Nz + z];
    }
    const double *aux = S.aux;
    switch (comp) {
        case HX: return (float)G(aux)[z];
        case HY: return (float)G(aux)[Nz + z];
        case EX: return (float)G(aux)[4 * Nz + z];
        case EY: return (float)G(aux)[5 * Nz + z];
        default: return 0.0f;
    }
}

// the four face components (H_u, H_v, E_u, E_v) of an axis
__device__ __forceinline__ int tfsf_comp(int axis, int kk) {
    if (axis == 0) return kk == 0 ? HY : kk == 1 ? HZ : kk == 2 ? EY : EZ;
    if (axis == 1) return kk == 0 ? HX : kk == 1 ? HZ : kk == 2 ? EX : EZ;
    return kk == 0 ? HX : kk == 1 ? HY : kk == 2 ? EX : EY;
}

// the aux run carries no Hz/Ez: their corrections add exact zeros and are skipped
__device__ __forceinline__ bool tfsf_live_mode(int mode, int axis, int kk) {
    const int c = tfsf_comp(axis, kk);
    return mode == 0 || (c != HZ && c != EZ);
}
template <class SR>
__device__ __forceinline__ bool tfsf_live(SR &S, int axis, int kk) {
    return tfsf_live_mode(S.mode, axis, kk);
}

// incident table value kk (0 H_u, 1 H_v, 2 E_u, 3 E_v) of face (axis, side) at (u, v)
template <class SR>
__device__ inline float tfsf_inc(SR &S, int Nz, int axis, int side, int kk, int u, int v,
                                 int it0) {
    const int ua = axis == 0 ? 1 : 0, va = axis == 2 ? 1 : 2;
    const int U1 = S.box[2 * ua], U2 = S.box[2 * ua + 1];
    const int V1 = S.box[2 * va], V2 = S.box[2 * va + 1];
    const int N1 = S.box[2 * axis], N2 = S.box[2 * axis + 1];
    if (S.mode == 1) {
        int z;
        if (axis == 2) z = side ? N2 : (kk < 2 ? N1 - 1 : N1);
        else z = v;
        return aux_value(S, Nz, tfsf_comp(axis, kk), z, it0);
    }
    const float *T = S.tab[4 * (2 * axis + side) + kk];
    if (!T) return 0.0f;
    const int Nu = U2 - U1, Nv = V2 - V1;
    const int na = (kk == 0 || kk == 3) ? Nu + 1 : Nu;
    const int nb = (kk == 0 || kk == 3) ? Nv : Nv + 1;
    return G(T)[(u - U1) + (size_t)na * ((v - V1) + (size_t)nb * it0)];
}

// Whether an x-face cell is finished by the E half-step epilogue: it lies on an x face
// plane (N1-1, N1 or N2) inside the tangential ranges, on no y or z face plane, and its
// material has no ADE/plasma state (whose E comes later, from the state pass).  Such a
// cell gets its E/D corrections and its Sa/E2 (with the H correction folded in) in the E
// kernel; the face kernel only writes its H correction.
__device__ __forceinline__ bool tfsf_xhandled_box(const int (&b)[6], int ix, int iy, int iz,
                                                  bool stateful) {
    const int X1 = b[0], X2 = b[1], Y1 = b[2], Y2 = b[3], Z1 = b[4], Z2 = b[5];
    if (!((ix == X1 - 1 || ix == X1 || ix == X2) && iy >= Y1 && iy <= Y2 && iz >= Z1 && iz <= Z2))
        return false;
    const bool on_y = (iy == Y1 - 1 || iy == Y1 || iy == Y2) && ix >= X1 && ix <= X2;
    const bool on_z = (iz == Z1 - 1 || iz == Z1 || iz == Z2) && ix >= X1 && ix <= X2;
    return !on_y && !on_z && !stateful;
}
// Whether a cell lies on a TFSF face plane inside the tangential ranges: the cells a
// face correction may write (the touched set of the host's mark_touched)
__device__ __forceinline__ bool tfsf_touched_box(const int (&b)[6], int ix, int iy, int iz) {
    const int X1 = b[0], X2 = b[1], Y1 = b[2], Y2 = b[3], Z1 = b[4], Z2 = b[5];
    const bool inx = ix >= X1 && ix <= X2, iny = iy >= Y1 && iy <= Y2, inz = iz >= Z1 && iz <= Z2;
    const bool fx = (ix == X1 - 1 || ix == X1 || ix == X2) && iny && inz;
    const bool fy = (iy == Y1 - 1 || iy == Y1 || iy == Y2) && inx && inz;
    const bool fz = (iz == Z1 - 1 || iz == Z1 || iz == Z2) && inx && iny;
    return fx || fy || fz;
}

// Every correction that lands on cell c3, applied to its fields f[] in the reference's face
// order (x, y, z; lower side, upper side; src/sources.jl:725-774).  b is the box, co the
// factors (cm, ce, cd) per axis, live(axis, kk) whether a table carries values and
// inc(axis, side, kk, u, v) the incident value.  With xh (an x-face cell whose E, D and
// integrals the E half-step epilogue finished) only the x faces' H corrections.  Returns
// the mask of the components written.
template <class LIVE, class INC>
__device__ __forceinline__ unsigned cell_fix(const int (&b)[6], const float (&co)[3][3], LIVE live, INC inc,
                                             const int (&c3)[3], bool xh, float (&f)[9]) {
    unsigned dirty = 0;
#pragma unroll
    for (int axis = 0; axis < 3; axis++) {
        const int ua = axis == 0 ? 1 : 0, va = axis == 2 ? 1 : 2;
        const int U1 = b[2 * ua], U2 = b[2 * ua + 1];
        const int V1 = b[2 * va], V2 = b[2 * va + 1];
        const int N1 = b[2 * axis], N2 = b[2 * axis + 1];
        const int u = c3[ua], v = c3[va], n = c3[axis];
        if (u < U1 || u > U2 || v < V1 || v > V2) continue;
        if (xh && axis > 0) continue;
        constexpr int cH[3][2] = {{HY, HZ}, {HX, HZ}, {HX, HY}};
        constexpr int cE[3][2] = {{EY, EZ}, {EX, EZ}, {EX, EY}};
        constexpr int cD[3][2] = {{DY, DZ}, {DX, DZ}, {DX, DY}};
        const int cHu = cH[axis][0], cHv = cH[axis][1], cEu = cE[axis][0], cEv = cE[axis][1];
        const int cDu = cD[axis][0], cDv = cD[axis][1];
        const float cm = co[axis][0], ce = co[axis][1], cd = co[axis][2];
#pragma unroll
        for (int side = 0; side < 2; side++) {
            const int nH = side ? N2 : N1 - 1, nE = side ? N2 : N1;
            if (n != nH && n != nE) continue;
            const float sg = side ? -1.0f : 1.0f;
            const float s0 = (axis == 1 ? 1.0f : -1.0f) * sg;
            if (n == nH) {
                if (u <= U2 && v <= V2 - 1 && live(axis, 3)) {
                    f[cHu] = __builtin_fmaf(s0 * cm, inc(axis, side, 3, u, v), f[cHu]);
                    dirty |= 1u << cHu;
                }
                if (u <= U2 - 1 && v <= V2 && live(axis, 2)) {
                    f[cHv] = __builtin_fmaf(-s0 * cm, inc(axis, side, 2, u, v), f[cHv]);
                    dirty |= 1u << cHv;
                }
            }
            if (n == nE && !xh) {
                if (u <= U2 - 1 && v <= V2 && live(axis, 1)) {
                    const float iv = inc(axis, side, 1, u, v);
                    f[cEu] = __builtin_fmaf(-s0 * ce, iv, f[cEu]);
                    f[cDu] = __builtin_fmaf(-s0 * cd, iv, f[cDu]);
                    dirty |= (1u << cEu) | (1u << cDu);
                }
                if (u <= U2 && v <= V2 - 1 && live(axis, 0)) {
                    const float iv = inc(axis, side, 0, u, v);
                    f[cEv] = __builtin_fmaf(s0 * ce, iv, f[cEv]);
                    f[cDv] = __builtin_fmaf(s0 * cd, iv, f[cDv]);
                    dirty |= (1u << cEv) | (1u << cDv);
                }
            }
        }
    }
    return dirty;
}

// cell_fix with the tables, factors and box of source S
template <class SR>
__device__ __forceinline__ unsigned tfsf_cell_fix(SR &S, int Nz, int it0, const int (&c3)[3], bool xh, float (&f)[9]) {
    const int b[6] = {S.box[0], S.box[1], S.box[2], S.box[3], S.box[4], S.box[5]};
    const float co[3][3] = {{S.cm[0], S.ce[0], S.cd[0]}, {S.cm[1], S.ce[1], S.cd[1]}, {S.cm[2], S.ce[2], S.cd[2]}};
    return cell_fix(
        b, co, [&](int axis, int kk) { return tfsf_live(S, axis, kk); },
        [&](int axis, int side, int kk, int u, int v) { return tfsf_inc(S, Nz, axis, side, kk, u, v, it0); }, c3, xh,
        f);
}

template <class SR>
__device__ __forceinline__ bool tfsf_xhandled(SR &S, int ix, int iy, int iz, bool stateful) {
    const int b[6] = {S.box[0], S.box[1], S.box[2], S.box[3], S.box[4], S.box[5]};
    return tfsf_xhandled_box(b, ix, iy, iz, stateful);
}

}  // namespace mfd
