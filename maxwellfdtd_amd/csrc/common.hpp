// Device-side data layout of the MI355X MaxwellFDTD engine.
//
// One DevModel describes one solve resident in HBM.  All kernels take an array of
// DevModel plus a packed tile work-list, so a single solve and a batch of uneven solves
// run through the same launches.  Field arrays are SoA FP32, x fastest (the Julia
// column-major order of Field1D/2D/3D, src/fields.jl:7-90); 1D uses Nx = Ny = 1 and
// 2D uses Ny = 1.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace mfd {

// Global-address-space views of device pointers.  Pointers that kernels read out of the
// DevModel structs are generic (flat) to the compiler; casting them to address space 1
// makes every access a global_load/global_store instead of a flat access (which would
// also count against the scalar-load counter and serialise with it).
template <class T>
using gp = __attribute__((address_space(1))) T *;
template <class T>
__device__ __forceinline__ gp<T> G(T *p) { return (gp<T>)p; }
template <class T>
__device__ __forceinline__ gp<const T> G(const T *p) { return (gp<const T>)p; }

// |E x H| for the Sa integral (src/outputs.jl:182-191,318-327,475-485): the hardware
// square root (v_sqrt_f32, within 1 ulp) rather than the correctly rounded library
// sequence, which costs about 15 more instructions per cell and step
__device__ __forceinline__ float sa_sqrt(float x) { return __builtin_amdgcn_sqrtf(x); }
// Sa and E2 of a 3D cell advanced by one step (src/outputs.jl:475-485), the multiply-adds
// written out so that the fused step and the TFSF face launch round them alike whatever
// each file's contraction setting
__device__ __forceinline__ float sa_step3(float sa, float ex, float ey, float ez, float hx, float hy, float hz,
                                          float dt) {
    const float a = __builtin_fmaf(ey, hz, -(ez * hy));
    const float b = __builtin_fmaf(ez, hx, -(ex * hz));
    const float c = __builtin_fmaf(ex, hy, -(ey * hx));
    return __builtin_fmaf(sa_sqrt(__builtin_fmaf(a, a, __builtin_fmaf(b, b, c * c))), dt, sa);
}
__device__ __forceinline__ float e2_step3(float e2, float ex, float ey, float ez, float dt) {
    return __builtin_fmaf(__builtin_fmaf(ex, ex, __builtin_fmaf(ey, ey, ez * ez)), dt, e2);
}
// square root in the precision of x: the same hardware form for FP32, correctly rounded
// for FP64
__device__ __forceinline__ float sqrt_of(float x) { return sa_sqrt(x); }
__device__ __forceinline__ double sqrt_of(double x) { return sqrt(x); }

// Two-element vectors usable through address-space-1 pointers (HIP's float2/double2
// classes only bind to generic references).
typedef float v2f __attribute__((ext_vector_type(2)));
// Unsigned division by an invariant d as a multiply-high and two shifts (the round-up
// method of Granlund and Montgomery, exact for every 32-bit n): {m, s1, s2, d}
inline uint4 udiv_magic_of(unsigned d) {
    unsigned l = 0;
    while ((1ull << l) < d) l++;                       // ceil(log2 d)
    const unsigned long long m = ((1ull << 32) * ((1ull << l) - d)) / d + 1;
    uint4 r;
    r.x = (unsigned)m;
    r.y = l < 1 ? l : 1;
    r.z = l > 1 ? l - 1 : 0;
    r.w = d;
    return r;
}
__device__ __forceinline__ unsigned udiv_magic(unsigned n, const uint4 &g) {
    const unsigned t = __umulhi(g.x, n);
    return (t + ((n - t) >> g.y)) >> g.z;
}

__host__ __device__ inline v2f mk2(float a, float b) {
    v2f v;
    v.x = a;
    v.y = b;
    return v;
}
typedef double v2d __attribute__((ext_vector_type(2)));

enum { HX = 0, HY, HZ, DX, DY, DZ, EX, EY, EZ };
enum { BC_PERIODIC = 1, BC_DIRICHLET = 2, BC_NEUMANN = 3 };
enum { MAT_DISP = 1, MAT_KERR = 2, MAT_PLASMA = 4 };
enum { SRC_SOFT = 0, SRC_HARD = 1, SRC_TFSF = 2 };

// One CPML face layer: profiles from LeftPMLLayer/RightPMLLayer (src/pml.jl:57-127)
// and the psi arrays of that face (src/pml.jl:221-304), FP32.  In an FP32 model the two
// faces of an axis share each psi allocation: the upper face's array starts right after
// the lower face's.
struct DevLayer {
    int present;
    int ib;        // 0-based boundary index: left layer covers i <= ib, right i >= ib
    int nb;
    int pad;
    const float *K, *A, *B;
    const float *Ki;        // 1/K (rounded once from FP64), for the pipelined 3D kernels
    // 1D / 2D models (and FP64 models, whose arrays hold doubles): one array per psi
    float *psiE0, *psiE1;   // H half-step psi (x: Eyx,Ezx  y: Exy,Ezy  z: Exz,Eyz)
    float *psiH0, *psiH1;   // E half-step psi (x: Hyx,Hzx  y: Hxy,Hzy  z: Hxz,Hyz)
    // 3D FP32 models: the two psi of a cell interleaved, (psi0, psi1) per entry, so one
    // 8-byte access and one base pointer serve both
    v2f *pE, *pH;           // H and E half-step psi pairs
    v2f *pEb;               // fused 3D step: the other buffer of the H half-step pairs
    float *psiEb;           // fused 2D step: the other buffer of the face's H half-step
                            // psi (psiE1 on x faces, psiE0 on z faces)
};

// Material coefficient table entry (index 0 = vacuum) with the material's compact
// state: only the cells of the material hold ADE / plasma state, addressed by the
// per-cell slot index (the reference keeps full-grid arrays, src/materials.jl:424-474).
struct DevMat {
    float Mh, Me, Md1, Md2;     // src/materials.jl:393-400 (vacuum: src/models.jl:111-114)
    float iMe;                  // 1/Me = EPS0*eps: D of a plain cell from its E
    float MeMd2;                // Me*Md2: the E-form update of a plain cell without D
    int flags;                  // MAT_DISP | MAT_KERR | MAT_PLASMA
    int nq;                     // ADE poles
    float K2, K3;               // Kerr Pade: Mk2*Me^2, Mk3*Me^3 (src/models.jl:1140-1149)
    // ADE poles and plasma current: coefficients and state in FP64 (material_e)
    const double *Aq, *Bq, *Cq; // ade_coefficients (src/materials.jl:509-577)
    double *P;                  // [2][nq][nc][cnt]: half (it & 1) P^n, the other P^{n-1}
    int cnt;                    // state slots (cells of this material)
    int rate_kind, rate_n, pad;
    double Rava, ksi, Cprho0, Marho0, rate_c, rate_k, rho0;
    double Ap, Bp;
    const double *rate_x, *rate_y;
    double2 *rhod;              // [cnt]: (rho, drho), FP64
    double *Pp;                 // [2][nc][cnt]: half (it & 1) Pp^n, the other Pp^{n-1}
    double *Pa;                 // [nc][cnt]
    float *JE;                  // [cnt]
    const int *cells;           // [cnt]: linear cell index of each state slot (stateful materials)
};

// FP64 twin of DevMat for models that keep their state in FP64 (1D models with
// MFD_OPT_F64): the same recursions with double coefficients and double state.
struct DevMat64 {
    double Mh, Me, Md1, Md2;
    int flags, nq;
    double K2, K3;
    const double *Aq, *Bq, *Cq;
    double *P;                  // [2][nq][nc][cnt], as DevMat's
    int cnt;
    int rate_kind, rate_n, pad;
    double Rava, ksi, Cprho0, Marho0, rate_c, rate_k, rho0;
    double Ap, Bp;
    const double *rate_x, *rate_y;
    double2 *rhod;              // [cnt]: (rho, drho)
    double *Pp;                 // [2][nc][cnt], as DevMat's
    double *Pa;                 // [nc][cnt]
    double *JE;                 // [cnt]
};

// Soft/hard point source or TFSF source of one model.
struct DevSource {
    int kind;                   // SRC_SOFT / SRC_HARD / SRC_TFSF
    int comp;
    int n;
    float coeff;
    double coeff64;
    const int *cells;           // 0-based linear cell indices [n]
    const float *wave;          // [Nt][n]
    // TFSF: 3D face scheme (1D/2D map onto faces 4,5 / 0,1,4,5, see engine.cpp)
    int box[6];                 // 0-based X1, X2, Y1, Y2, Z1, Z2
    int mode;                   // 0 tables, 1 aux 1D
    int aux_n;
    const float *tab[24];       // replay tables per face: H_u, H_v, E_u, E_v
    float cm[3], ce[3], cd[3];  // Cm/d, Ce/d, EPS0*Ce/d per axis (src/sources.jl:725)
    double cm64[3], ce64[3], cd64[3];   // the same in FP64, for FP64 models
    // auxiliary 1D run (FP64): fields [6][Nz] = Hx, Hy, Dx, Dy, Ex, Ey
    double *aux;
    double *auxpsi;             // [2 faces][4][nbmax]: dEx/dz, dEy/dz, dHx/dz, dHy/dz
    int aux_nbmax;
    int aux_ib[2], aux_nb[2], aux_present[2];
    const double *auxK[2], *auxA[2], *auxB[2];
    const int *aux_comp, *aux_iz;
    const double *aux_coeff, *aux_wave;   // [aux_n][Nt]
    double Mh0, Me0, Md10, Md20, idz;
    const float *auxtab;        // tabulated aux run [Nt][4][Nz]: Hx, Hy, Ex, Ey (or null)
    // FP64 models: the waveform table and the replay tables in FP64 (the aux run is
    // then stepped live, in FP64, instead of tabulated)
    const double *wave64;
    const double *tab64[24];
};

// Field / spectral monitor of one model.
struct DevMon {
    int spectral, issum, n, nw;
    int ncomp;
    int comps[6];
    const int *cells;
    float *rec;                 // field monitor, per-point record [Nt][ncomp][n]
    double *sum;                // field monitor with sum=true [Nt][ncomp]
    float *samp;                // spectral: time block of samples [K][ncomp][n]
    double2 *S;                 // spectral accumulators [ncomp][nw][n] (or [ncomp][nw])
    const double *w;            // [nw]
    double tshift;
};

struct DevModel {
    int ndims, Nx, Ny, Nz;
    long long N, sy, sz;        // cells, strides in y and z
    float idx, idy, idz, dtf;   // 1/dx, 1/dy, 1/dz, dt (FP32)
    double dt;
    int bcs[6];
    int Nt;
    int int_flags;              // 1 Sa, 2 E2
    float *F[9];
    const unsigned char *id;    // material id per cell (bit 7: cell written by a source)
    const int *slot;            // state slot per cell, nullptr if no stateful material
    const DevMat *mats;         // [nmat + 1]
    int nmat;
    DevLayer L[6];
    float *Sa, *E2;
    int *it;                    // device step counter: completed steps
    int *flag;                  // non-finite Ex: 1 NaN, 2 Inf
    int *dftdone;               // steps already contracted into the DFT accumulators
    int nsrc, nmon;
    const DevSource *src;
    const DevMon *mon;
    int dft_block;              // DFT time-block length
    const int *touched;         // cells written by a source (integral fixup list); bit 31
                                // (TOUCHED_STATEFUL): the cell's material has ADE/plasma state
    int ntouched;
    const DevSource *xtfsf;     // TFSF source whose x faces the E kernel finishes, or null
    int xface_h;                // fused 3D step: it also stores the H corrections of the
                                // x-face cells it finishes (touched then lists the others)
    // cell index -> coordinates without integer division: n / Nx and n / (Nx Ny) as
    // udiv_magic(n, divx) and udiv_magic(n, divxy)
    uint4 divx, divxy;
    int dless;                  // plain cells keep no D: D = E/Me is rebuilt on demand
    int ntst;                   // fused 3D step: new E and H stored non-temporally (a lone
                                // model whose two field sets exceed the Infinity Cache)
    // FP64 state (1D models with MFD_OPT_F64): the fields are F64[] (F[] is null), the
    // material table is mats64 (mats keeps the FP32 table for the id lookups of the
    // shared kernels), and every CPML layer array (K, A, B, psi) holds doubles
    int f64;
    double *F64[9];
    const DevMat64 *mats64;
    double idz64;               // 1/dz in FP64
    double *Sa64, *E2_64;       // the integral variables of an FP64 model (Sa, E2 null)
    // fused 3D step: the other buffer of H and E (Falt[HX..HZ], Falt[EX..EZ]); each step
    // reads F and writes Falt, and a launch after it swaps the two (and the H-half psi
    // pairs of the layers)
    float *Falt[9];
};

// The field arrays and material table of a model in precision T (float: F, mats;
// double: F64, mats64).
template <class T> struct Prec;
template <> struct Prec<float> {
    typedef DevMat Mat;
    static __device__ __forceinline__ float *Sa(const DevModel &m) { return m.Sa; }
    static __device__ __forceinline__ float *E2(const DevModel &m) { return m.E2; }
    static __device__ __forceinline__ float *field(const DevModel &m, int c) { return m.F[c]; }
    static __device__ __forceinline__ const DevMat *mats(const DevModel &m) { return m.mats; }
    static __device__ __forceinline__ float idz(const DevModel &m) { return m.idz; }
    static __device__ __forceinline__ float dt(const DevModel &m) { return m.dtf; }
};
template <> struct Prec<double> {
    typedef DevMat64 Mat;
    static __device__ __forceinline__ double *Sa(const DevModel &m) { return m.Sa64; }
    static __device__ __forceinline__ double *E2(const DevModel &m) { return m.E2_64; }
    static __device__ __forceinline__ double *field(const DevModel &m, int c) { return m.F64[c]; }
    static __device__ __forceinline__ const DevMat64 *mats(const DevModel &m) { return m.mats64; }
    static __device__ __forceinline__ double idz(const DevModel &m) { return m.idz64; }
    static __device__ __forceinline__ double dt(const DevModel &m) { return m.dt; }
};

// Field access of the kernels shared by FP32 and FP64 models (sources, TFSF faces,
// monitors, integrals): values cross as FP32 inputs or outputs; additions to the state
// happen in the model's own precision.
// pp: a fused step ran and the buffer swap comes later in the step, so this step's H
// and E are in Falt (D stays in place)
__device__ __forceinline__ const float *fld_ptr(const DevModel &m, int c, int pp) {
    return pp && (c < DX || c > DZ) ? m.Falt[c] : m.F[c];
}
__device__ __forceinline__ float fld_get(const DevModel &m, int c, long long i, int pp = 0) {
    return m.f64 ? (float)G(m.F64[c])[i] : G(fld_ptr(m, c, pp))[i];
}
__device__ __forceinline__ double fld_get64(const DevModel &m, int c, long long i, int pp = 0) {
    return m.f64 ? G(m.F64[c])[i] : (double)G(fld_ptr(m, c, pp))[i];
}

constexpr unsigned char ID_MASK = 0x7f, ID_TOUCHED = 0x80;
constexpr int TOUCHED_STATEFUL = (int)0x80000000u, TOUCHED_CELL = 0x7fffffff;
// E-kernel material modes: vacuum only, coefficient table only, full (ADE/Kerr/plasma)
enum { MODE_VAC = 0, MODE_SIMPLE = 1, MODE_FULL = 2 };

// Tile of the work-list: model index, x0, y0, (z0 | z1 << 16).
typedef int4 Tile;

__host__ __device__ inline int tile_z0(const Tile &t) { return t.w & 0xffff; }
__host__ __device__ inline int tile_z1(const Tile &t) { return (t.w >> 16) & 0xffff; }

// tile geometry per dimensionality
constexpr int TX3 = 64, TY3 = 4;    // 3D scalar kernels: 64 x 4 columns, marching z
constexpr int TX2 = 64;             // 2D: 64 x-columns per wave, TW waves per tile, marching z
constexpr int TX1 = 256;            // 1D: 256 cells
constexpr int XFACE_ZC = 64;        // longest tile z range with the E kernel's x-face epilogue
constexpr int FUSED_XFACE_ZC = 32;  // the same for the fused step (its LDS holds R rows)

// launchers (defined in the .hip files)
void launch_h3d(const DevModel *m, const Tile *tiles, int ntiles, int grid, int mode, hipStream_t s);
void launch_e3d(const DevModel *m, const Tile *tiles, int ntiles, int grid, int mode, bool integ, hipStream_t s);
// pml = false: tiles that no CPML layer reaches (the kernels leave out the slab updates)
void launch_h3d_p(const DevModel *m, const Tile *tiles, int ntiles, int grid, int mode, int ty, bool pml,
                  hipStream_t s);
void launch_e3d_p(const DevModel *m, const Tile *tiles, int ntiles, int grid, int mode, bool integ, int ty,
                  bool pml, hipStream_t s);
// fused 3D step (one z march per step over ping-pong H/E buffers), rows = 2 or 4;
// pmask: CPML axes the tiles reach (0 none, 1 x only, 7 any)
void launch_step3d_f(DevModel *m, const Tile *tiles, int ntiles, int grid, int mode, int integ, int rows,
                     int pmask, hipStream_t s);
void launch_swap_pp(DevModel *m, int nmodels, hipStream_t s);
void launch_finish_swap(DevModel *m, int nmodels, hipStream_t s);
void launch_step2d_f(DevModel *m, const Tile *tiles, int ntiles, int grid, int mode, int integ, bool pml,
                     hipStream_t s);
void launch_h2d(const DevModel *m, const Tile *tiles, int ntiles, int grid, int mode, int tw, hipStream_t s);
void launch_e2d(const DevModel *m, const Tile *tiles, int ntiles, int grid, int mode, bool integ, int tw, hipStream_t s);
void launch_h1d(const DevModel *m, const Tile *tiles, int ntiles, int grid, int mode, bool f64,
                hipStream_t s);
void launch_e1d(const DevModel *m, const Tile *tiles, int ntiles, int grid, int mode, bool integ,
                bool f64, hipStream_t s);
// whole runs of 1D models without TFSF sources: one workgroup per model, nsteps steps
// (H, E, point sources, monitors and DFT, touched-cell integrals, the step counter)
void launch_run1d(const DevModel *m, int nmodels, int nsteps, int mode, bool integ, bool f64, bool hmat,
                  hipStream_t s);
// the same with each model's cells in registers (models the host found small enough)
void launch_reg1d(const DevModel *m, int nmodels, int nsteps, int mode, bool integ, bool f64, bool hmat,
                  hipStream_t s);
void launch_integrals_touched(const DevModel *m, int nmodels, int maxn, int ndims, hipStream_t s);
void launch_point_sources(const DevModel *m, int nmodels, int slot, int maxn, hipStream_t s);
void launch_tfsf(const DevModel *m, int nmodels, int slot, int ndims, int maxface, hipStream_t s,
                 bool aux);
void launch_tfsf_aux(const DevModel *m, int nmodels, int slot, hipStream_t s);
void launch_tfsf_gather(const DevModel *m, int nmodels, int slot, int maxn, bool integ,
                        hipStream_t s, bool pp = false);
void launch_tfsf_aux_table(const DevSource *S, int Nz, int Nt, int bcl, int bcr, float *tab,
                           hipStream_t s);
void launch_monitors(const DevModel *m, int nmodels, int maxn, hipStream_t s, bool pp = false);
// force = false: models whose DFT block completes with this step's sample (launched
// after the monitors, before launch_finish); force = true: every model's pending partial
// block, to be followed by launch_dft_commit
void launch_dft_flush(const DevModel *m, int nmodels, int maxn, bool force, hipStream_t s);
void launch_dft_commit(const DevModel *m, int nmodels, hipStream_t s);
void launch_finish(const DevModel *m, int nmodels, hipStream_t s);
void launch_collect_flags(const DevModel *m, int nmodels, int *out, hipStream_t s);
// D = E/Me written back into the D arrays of a model that ran without stored D
void launch_d_from_e(const DevModel &m, hipStream_t s);
// E of the cells of stateful (ADE / plasma) materials, after the E half-step kernel
// updated their D: one work item per state slot, chunks {model, material, s0, s1}.
void launch_e_state(const DevModel *m, const int4 *chunks, int nchunks, int ndims, bool integ,
                    hipStream_t s, bool pp = false);

}  // namespace mfd
