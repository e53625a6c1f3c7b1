// Host engine: model upload to HBM, the per-step launch sequence, hipGraph replay and
// the C ABI of include/maxwell_fdtd.h.
#include "engine.hpp"

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <mutex>

namespace mfd {

static thread_local std::string g_error;
void set_error(const std::string &msg) { g_error = msg; }

Arena::~Arena() { release(); }
void Arena::release() {
    for (void *p : ptrs) (void)hipFree(p);
    ptrs.clear();
    bytes = 0;
}

bool Model::comp_present(int c) const {
    if (d.ndims == 3) return true;
    if (d.ndims == 2) return c == HY || c == DX || c == DZ || c == EX || c == EZ;
    return c == HY || c == DX || c == EX;
}

static int monitor_comps(int ndims, int *out) {
    static const int c3[6] = {HX, HY, HZ, EX, EY, EZ};
    static const int c2[3] = {HY, EX, EZ};
    static const int c1[2] = {HY, EX};
    const int *c = ndims == 3 ? c3 : ndims == 2 ? c2 : c1;
    const int n = ndims == 3 ? 6 : ndims == 2 ? 3 : 2;
    for (int i = 0; i < n; i++) out[i] = c[i];
    return n;
}

static int psi_len(const Model &m, int face, int nb) {
    const int Nx = m.d.Nx, Ny = m.d.Ny, Nz = m.d.Nz;
    if (face < 2) return nb * Ny * Nz;
    if (face < 4) return Nx * nb * Nz;
    return Nx * Ny * nb;
}

int64_t Model::footprint() const {
    int64_t b = 0;
    int nf = 0;
    for (int c = 0; c < 9; c++) nf += comp_present(c);
    b += (int64_t)nf * N * (f64 ? 8 : 4);
    b += N;
    bool stateful = false;
    std::vector<long long> cnt(mats.size() + 1, 0);
    if (has_geom)
        for (long long i = 0; i < N; i++) cnt[geom[i]]++;
    for (size_t k = 0; k < mats.size(); k++) {
        const auto &md = mats[k].d;
        const long long c = cnt[k + 1];
        if (md.nq > 0 || md.isplasma) { stateful = true; b += c * 4; }
        b += (int64_t)md.nq * ncomp() * c * 16;
        if (md.isplasma) b += c * (16 + ncomp() * 24 + (f64 ? 8 : 4));
    }
    if (stateful) b += N * 4;
    for (int f = 0; f < 6; f++)
        if (L[f].present) b += (int64_t)4 * psi_len(*this, f, L[f].nb) * 4 + 4 * L[f].nb * 4;
    for (const auto &s : srcs) {
        if (s.type == 0) b += (int64_t)s.cells.size() * (4 + 4LL * d.Nt);
        else {
            for (const auto &t : s.tables) b += t.size() * 4;
            b += 6LL * d.Nz * 8 + (int64_t)s.aux_wave.size() * 8;
        }
    }
    int cm[6];
    const int nc = monitor_comps(d.ndims, cm);
    for (const auto &mo : mons) {
        const int64_t n = mo.cells.size();
        if (!mo.spectral) b += mo.issum ? (int64_t)d.Nt * nc * 8 : (int64_t)d.Nt * nc * n * 4;
        else b += (int64_t)dft_block * nc * n * 4 + (int64_t)nc * mo.w.size() * (mo.issum ? 1 : n) * 16;
    }
    if (int_flags & MFD_INT_SA) b += N * 4;
    if (int_flags & MFD_INT_E2) b += N * 4;
    // the fused step's second H/E buffers and H-half psi arrays
    if (d.ndims == 3 && !f64) {
        b += 6 * N * 4;
        for (int f = 0; f < 6; f++)
            if (L[f].present) b += (int64_t)2 * psi_len(*this, f, L[f].nb) * 4;
    }
    if (d.ndims == 2 && !f64) {
        b += 3 * N * 4;
        for (int f = 0; f < 6; f++)
            if (L[f].present) b += (int64_t)psi_len(*this, f, L[f].nb) * 4;
    }
    return b;
}

// Every host<->device copy of the engine runs on a non-blocking stream and waits for
// it: nothing touches the legacy stream, so copies from one thread never depend on a
// stream another thread (a server worker) is capturing into a graph.
hipError_t copy_sync(void *dst, const void *src, size_t n, hipMemcpyKind kind, hipStream_t s) {
    hipError_t e = hipMemcpyAsync(dst, src, n, kind, s);
    if (e != hipSuccess) return e;
    return hipStreamSynchronize(s);
}

template <class T>
static hipError_t upload_vec(Arena &a, T **dst, const std::vector<T> &v, hipStream_t s) {
    hipError_t e = a.alloc(dst, v.size());
    if (e != hipSuccess) return e;
    if (!v.empty()) e = copy_sync(*dst, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice, s);
    return e;
}

static std::vector<float> to_f32(const std::vector<double> &v) {
    std::vector<float> r(v.size());
    for (size_t i = 0; i < v.size(); i++) r[i] = (float)v[i];
    return r;
}

// 0-based face box of the 3D face scheme; 1D and 2D boxes are degenerate in the absent
// tangential axes (see the TFSF kernels).
static void tfsf_face_box(const mfd_model_desc &d, const int *box, int *b) {
    if (d.ndims == 3) {
        for (int i = 0; i < 6; i++) b[i] = box[i] - 1;
    } else if (d.ndims == 2) {
        b[0] = box[0] - 1; b[1] = box[1] - 1; b[2] = 0; b[3] = 0;
        b[4] = box[4] - 1; b[5] = box[5] - 1;
    } else {
        b[0] = 0; b[1] = 1; b[2] = 0; b[3] = 0;
        b[4] = box[4] - 1; b[5] = box[5] - 1;
    }
}

// Cells a source writes after the E half-step: their integral variables are accumulated
// after all sources (the E kernels skip them).
static void mark_touched(const Model &m, std::vector<unsigned char> &mark) {
    const auto &d = m.d;
    for (const auto &s : m.srcs) {
        if (s.type == 0) {
            for (int c : s.cells) mark[c] = 1;
            continue;
        }
        int b[6];
        tfsf_face_box(d, s.box, b);
        for (int axis = 0; axis < 3; axis++) {
            if ((d.ndims == 1 && axis < 2) || (d.ndims == 2 && axis == 1)) continue;
            const int ua = axis == 0 ? 1 : 0, va = axis == 2 ? 1 : 2;
            const int planes[4] = {b[2 * axis] - 1, b[2 * axis], b[2 * axis + 1], b[2 * axis + 1]};
            for (int n : planes)
                for (int u = b[2 * ua]; u <= b[2 * ua + 1]; u++)
                    for (int v = b[2 * va]; v <= b[2 * va + 1]; v++) {
                        int c3[3];
                        c3[axis] = n; c3[ua] = u; c3[va] = v;
                        if (c3[0] < 0 || c3[0] >= d.Nx || c3[1] < 0 || c3[1] >= d.Ny ||
                            c3[2] < 0 || c3[2] >= d.Nz)
                            continue;
                        mark[c3[0] + (long long)d.Nx * (c3[1] + (long long)d.Ny * c3[2])] = 1;
                    }
        }
    }
}

#define UP(call)                                                                       \
    do {                                                                               \
        hipError_t e_ = (call);                                                        \
        if (e_ != hipSuccess) {                                                        \
            set_error(std::string("upload: ") + #call + ": " + hipGetErrorString(e_)); \
            arena.release();                                                           \
            return e_ == hipErrorOutOfMemory ? MFD_ERR_NOMEM : MFD_ERR_HIP;            \
        }                                                                              \
    } while (0)

int Model::ensure_io() {
    if (!io) MFD_HIP(hipStreamCreateWithFlags(&io, hipStreamNonBlocking));
    return MFD_OK;
}

int Model::upload() {
    if (uploaded) return MFD_OK;
    {
        int st = ensure_io();
        if (st != MFD_OK) return st;
    }
    DevModel m{};
    m.ndims = d.ndims;
    m.Nx = d.Nx; m.Ny = d.Ny; m.Nz = d.Nz;
    m.N = N;
    m.sy = d.Nx;
    m.sz = (long long)d.Nx * d.Ny;
    m.idx = (float)(1.0 / d.dx);
    m.idy = (float)(1.0 / d.dy);
    m.idz = (float)(1.0 / d.dz);
    m.dtf = (float)d.dt;
    m.dt = d.dt;
    for (int i = 0; i < 6; i++) m.bcs[i] = d.bcs[i];
    m.Nt = d.Nt;
    m.int_flags = int_flags;
    m.divx = udiv_magic_of((unsigned)d.Nx);
    m.divxy = udiv_magic_of((unsigned)d.Nx * (unsigned)d.Ny);
    // per-cell FP32 arrays (fields, Sa, E2) come from one pool; array k starts k*stagger
    // bytes past a 256-byte boundary so that equal cell indices of different arrays do
    // not all land on the same HBM channel and bank
    {
        // an FP64 model keeps its fields and integral variables in FP64 (8 B per cell)
        void **slots[11];
        size_t esz[11];
        int nslot = 0;
        for (int c = 0; c < 9; c++) {
            m.F[c] = nullptr;
            m.F64[c] = nullptr;
            if (!comp_present(c)) continue;
            slots[nslot] = f64 ? (void **)&m.F64[c] : (void **)&m.F[c];
            esz[nslot++] = f64 ? sizeof(double) : sizeof(float);
        }
        const size_t isz = f64 ? sizeof(double) : sizeof(float);
        if (int_flags & MFD_INT_SA) { slots[nslot] = f64 ? (void **)&m.Sa64 : (void **)&m.Sa; esz[nslot++] = isz; }
        if (int_flags & MFD_INT_E2) { slots[nslot] = f64 ? (void **)&m.E2_64 : (void **)&m.E2; esz[nslot++] = isz; }
        const size_t st = ((size_t)std::max(0, stagger) + 255) / 256 * 256;
        size_t total = 0;
        for (int k = 0; k < nslot; k++) total += ((size_t)N * esz[k] + 255) / 256 * 256 + st;
        char *pool;
        UP(arena.alloc(&pool, total));
        UP(hipMemsetAsync(pool, 0, total, io));
        size_t off = 0;
        for (int k = 0; k < nslot; k++) {
            *slots[k] = pool + off;
            off += ((size_t)N * esz[k] + 255) / 256 * 256 + st;
        }
    }
    m.f64 = f64 ? 1 : 0;
    m.idz64 = 1.0 / d.dz;
    // materials: id per cell (uint8) and compact state slots
    const int nmat = (int)mats.size();
    std::vector<long long> cnt(nmat + 1, 0);
    bool stateful = false;
    for (const auto &hm : mats)
        if (hm.d.nq > 0 || hm.d.isplasma) stateful = true;
    has_mats = has_geom;
    for (long long i = 0; i < N; i++)
        if (geom[i] > nmat) {
            set_error("geometry id " + std::to_string(geom[i]) + " has no material");
            return MFD_ERR_ARG;
        }
    {
        // the id array always exists (all zeros without materials) so that a batch
        // mixing models with and without materials can read it for every model
        std::vector<unsigned char> ids(N);
        std::vector<int> slot(stateful ? N : 0);
        std::vector<unsigned char> mark(N, 0);
        mark_touched(*this, mark);
        touched.clear();
        state_cells.assign(nmat + 1, std::vector<int>());
        for (long long i = 0; i < N; i++) {
            const int g = geom[i];
            ids[i] = (unsigned char)(g | (mark[i] ? ID_TOUCHED : 0));
            if (mark[i]) touched.push_back((int)i);
            if (stateful) slot[i] = (int)cnt[g];
            if (g > 0 && (mats[g - 1].d.nq > 0 || mats[g - 1].d.isplasma)) state_cells[g].push_back((int)i);
            cnt[g]++;
        }
        unsigned char *did;
        UP(upload_vec(arena, &did, ids, io));
        m.id = did;
        // the device list carries, in bit 31, whether the cell's material has ADE or
        // plasma state, so that the face kernel needs no id and material lookups
        std::vector<int> tdev(touched);
        for (int &c : tdev) {
            const int g = geom[c];
            if (g > 0 && (mats[g - 1].d.nq > 0 || mats[g - 1].d.isplasma)) c |= TOUCHED_STATEFUL;
        }
        int *dt_;
        UP(upload_vec(arena, &dt_, tdev, io));
        m.touched = dt_;
        m.ntouched = (int)touched.size();
        if (stateful) {
            int *ds;
            UP(upload_vec(arena, &ds, slot, io));
            m.slot = ds;
        }
    }
    const int nc = ncomp();
    hmats.assign(nmat + 1, DevMat{});
    {
        DevMat &v = hmats[0];
        v.Mh = (float)d.Mh; v.Me = (float)d.Me; v.Md1 = (float)d.Md1; v.Md2 = (float)d.Md2;
        v.iMe = (float)(1.0 / d.Me);
        v.MeMd2 = (float)(d.Me * d.Md2);
    }
    for (int k = 0; k < nmat; k++) {
        const HostMat &hm = mats[k];
        DevMat &v = hmats[k + 1];
        const long long c = cnt[k + 1];
        v.Mh = (float)hm.d.Mh; v.Me = (float)hm.d.Me;
        v.Md1 = (float)hm.d.Md1; v.Md2 = (float)hm.d.Md2;
        v.iMe = (float)(1.0 / hm.d.Me);
        v.MeMd2 = (float)(hm.d.Me * hm.d.Md2);
        v.flags = (hm.d.nq > 0 ? MAT_DISP : 0) | (hm.d.iskerr ? MAT_KERR : 0) |
                  (hm.d.isplasma ? MAT_PLASMA : 0);
        v.nq = hm.d.nq;
        v.cnt = (int)c;
        v.K2 = (float)(hm.d.Mk2 * hm.d.Me * hm.d.Me);
        v.K3 = (float)(hm.d.Mk3 * hm.d.Me * hm.d.Me * hm.d.Me);
        if (!state_cells[k + 1].empty()) {
            int *cp;
            UP(upload_vec(arena, &cp, state_cells[k + 1], io));
            v.cells = cp;
        }
        if (hm.d.nq > 0) {
            double *p;
            UP(upload_vec(arena, &p, hm.Aq, io)); v.Aq = p;
            UP(upload_vec(arena, &p, hm.Bq, io)); v.Bq = p;
            UP(upload_vec(arena, &p, hm.Cq, io)); v.Cq = p;
            UP(arena.alloc(&v.P, (size_t)2 * hm.d.nq * nc * c));
            UP(hipMemsetAsync(v.P, 0, (size_t)2 * hm.d.nq * nc * c * sizeof(double), io));
        }
        if (hm.d.isplasma) {
            v.Rava = hm.d.Rava; v.ksi = hm.d.ksi; v.rho0 = hm.d.rho0;
            v.Cprho0 = hm.d.Cp * hm.d.rho0;
            v.Marho0 = hm.d.Ma * hm.d.rho0;
            v.Ap = hm.d.Ap; v.Bp = hm.d.Bp;
            v.rate_kind = hm.d.rate_kind; v.rate_c = hm.d.rate_c; v.rate_k = hm.d.rate_k;
            v.rate_n = (int)hm.rx.size();
            double *dp;
            UP(upload_vec(arena, &dp, hm.rx, io)); v.rate_x = dp;
            UP(upload_vec(arena, &dp, hm.ry, io)); v.rate_y = dp;
            UP(arena.alloc(&v.rhod, c));
            UP(hipMemsetAsync(v.rhod, 0, c * sizeof(double2), io));
            UP(arena.alloc(&v.Pp, (size_t)2 * nc * c));
            UP(hipMemsetAsync(v.Pp, 0, (size_t)2 * nc * c * sizeof(double), io));
            UP(arena.alloc(&v.Pa, (size_t)nc * c));
            UP(hipMemsetAsync(v.Pa, 0, (size_t)nc * c * sizeof(double), io));
            UP(arena.alloc(&v.JE, c));
            UP(hipMemsetAsync(v.JE, 0, c * sizeof(float), io));
        }
    }
    if (f64) {
        // the FP64 table: coefficients from the host's FP64 values, state in FP64 (the
        // FP32 table above keeps the flags and counts; its state arrays are not used)
        hmats64.assign(nmat + 1, DevMat64{});
        DevMat64 &v0 = hmats64[0];
        v0.Mh = d.Mh; v0.Me = d.Me; v0.Md1 = d.Md1; v0.Md2 = d.Md2;
        for (int k = 0; k < nmat; k++) {
            const HostMat &hm = mats[k];
            DevMat64 &v = hmats64[k + 1];
            const long long c = cnt[k + 1];
            v.Mh = hm.d.Mh; v.Me = hm.d.Me; v.Md1 = hm.d.Md1; v.Md2 = hm.d.Md2;
            v.flags = hmats[k + 1].flags;
            v.nq = hm.d.nq;
            v.cnt = (int)c;
            v.K2 = hm.d.Mk2 * hm.d.Me * hm.d.Me;
            v.K3 = hm.d.Mk3 * hm.d.Me * hm.d.Me * hm.d.Me;
            if (hm.d.nq > 0) {
                double *p;
                UP(upload_vec(arena, &p, hm.Aq, io)); v.Aq = p;
                UP(upload_vec(arena, &p, hm.Bq, io)); v.Bq = p;
                UP(upload_vec(arena, &p, hm.Cq, io)); v.Cq = p;
                UP(arena.alloc(&v.P, (size_t)2 * hm.d.nq * nc * c));
                UP(hipMemsetAsync(v.P, 0, (size_t)2 * hm.d.nq * nc * c * sizeof(double), io));
            }
            if (hm.d.isplasma) {
                const DevMat &f = hmats[k + 1];
                v.Rava = f.Rava; v.ksi = f.ksi; v.rho0 = f.rho0; v.Cprho0 = f.Cprho0;
                v.Marho0 = f.Marho0; v.Ap = hm.d.Ap; v.Bp = hm.d.Bp;
                v.rate_kind = f.rate_kind; v.rate_c = f.rate_c; v.rate_k = f.rate_k;
                v.rate_n = f.rate_n; v.rate_x = f.rate_x; v.rate_y = f.rate_y;
                v.rhod = f.rhod;
                UP(arena.alloc(&v.Pp, (size_t)2 * nc * c));
                UP(hipMemsetAsync(v.Pp, 0, (size_t)2 * nc * c * sizeof(double), io));
                UP(arena.alloc(&v.Pa, (size_t)nc * c));
                UP(hipMemsetAsync(v.Pa, 0, (size_t)nc * c * sizeof(double), io));
                UP(arena.alloc(&v.JE, c));
                UP(hipMemsetAsync(v.JE, 0, c * sizeof(double), io));
            }
        }
        DevMat64 *dm64;
        UP(upload_vec(arena, &dm64, hmats64, io));
        m.mats64 = dm64;
    }
    DevMat *dmats;
    UP(upload_vec(arena, &dmats, hmats, io));
    m.mats = dmats;
    m.nmat = nmat;
    // CPML layers
    for (int f = 0; f < 6; f++) {
        DevLayer &l = m.L[f];
        l = DevLayer{};
        if (!L[f].present) continue;
        l.present = 1;
        l.ib = L[f].ib - 1;
        l.nb = L[f].nb;
        const size_t n = psi_len(*this, f, l.nb);
        float **dst[4] = {&l.psiE0, &l.psiE1, &l.psiH0, &l.psiH1};
        if (f64) {
            // an FP64 model's layer arrays hold doubles behind the same pointers
            double *p;
            UP(upload_vec(arena, &p, L[f].K, io)); l.K = (const float *)p;
            UP(upload_vec(arena, &p, L[f].A, io)); l.A = (const float *)p;
            UP(upload_vec(arena, &p, L[f].B, io)); l.B = (const float *)p;
            for (int k = 0; k < 4; k++) {
                double *q;
                UP(arena.alloc(&q, n));
                UP(hipMemsetAsync(q, 0, n * sizeof(double), io));
                *dst[k] = (float *)q;
            }
            continue;
        }
        float *p;
        UP(upload_vec(arena, &p, to_f32(L[f].K), io)); l.K = p;
        std::vector<float> ki(L[f].K.size());
        for (size_t q = 0; q < ki.size(); q++) ki[q] = (float)(1.0 / L[f].K[q]);
        UP(upload_vec(arena, &p, ki, io)); l.Ki = p;
        UP(upload_vec(arena, &p, to_f32(L[f].A), io)); l.A = p;
        UP(upload_vec(arena, &p, to_f32(L[f].B), io)); l.B = p;
    }
    // FP32 psi arrays: the two faces of an axis share one allocation, the upper face's
    // array right after the lower face's, so a kernel can address either from one base;
    // 3D models keep the two psi of a cell interleaved (pE, pH)
    if (!f64)
        for (int a = 0; a < 3; a++) {
            DevLayer &l0 = m.L[2 * a], &l1 = m.L[2 * a + 1];
            const size_t n0 = l0.present ? psi_len(*this, 2 * a, l0.nb) : 0;
            const size_t n1 = l1.present ? psi_len(*this, 2 * a + 1, l1.nb) : 0;
            if (n0 + n1 == 0) continue;
            if (d.ndims == 3) {
                v2f **d0[2] = {&l0.pE, &l0.pH};
                v2f **d1[2] = {&l1.pE, &l1.pH};
                for (int k = 0; k < 2; k++) {
                    v2f *q;
                    UP(arena.alloc(&q, n0 + n1));
                    UP(hipMemsetAsync(q, 0, (n0 + n1) * sizeof(v2f), io));
                    if (l0.present) *d0[k] = q;
                    if (l1.present) *d1[k] = q + n0;
                }
                continue;
            }
            float **d0[4] = {&l0.psiE0, &l0.psiE1, &l0.psiH0, &l0.psiH1};
            float **d1[4] = {&l1.psiE0, &l1.psiE1, &l1.psiH0, &l1.psiH1};
            for (int k = 0; k < 4; k++) {
                float *q;
                UP(arena.alloc(&q, n0 + n1));
                UP(hipMemsetAsync(q, 0, (n0 + n1) * sizeof(float), io));
                if (l0.present) *d0[k] = q;
                if (l1.present) *d1[k] = q + n0;
            }
        }
    // sources
    hsrc.assign(srcs.size(), DevSource{});
    maxsrc_n = 1;
    maxface = 1;
    has_tfsf = false;
    for (size_t s = 0; s < srcs.size(); s++) {
        const HostSource &hs = srcs[s];
        DevSource &v = hsrc[s];
        if (hs.type == 0) {
            v.kind = hs.kind == MFD_SRC_HARD ? SRC_HARD : SRC_SOFT;
            v.comp = hs.comp;
            v.n = (int)hs.cells.size();
            v.coeff = (float)hs.coeff;
            int *ip;
            UP(upload_vec(arena, &ip, hs.cells, io)); v.cells = ip;
            float *fp;
            UP(upload_vec(arena, &fp, to_f32(hs.wave), io)); v.wave = fp;
            v.coeff64 = hs.coeff;
            if (f64) {
                double *dp;
                UP(upload_vec(arena, &dp, hs.wave, io)); v.wave64 = dp;
            }
            maxsrc_n = std::max(maxsrc_n, v.n);
            continue;
        }
        has_tfsf = true;
        v.kind = SRC_TFSF;
        v.mode = hs.mode == MFD_TFSF_AUX1D ? 1 : 0;
        // face scheme of the 3D replay; 1D and 2D degenerate tangential ranges
        int b[6];
        tfsf_face_box(d, hs.box, b);
        for (int i = 0; i < 6; i++) v.box[i] = b[i];
        for (int f = 0; f < 6; f++) {
            const int axis = f >> 1;
            const int ua = axis == 0 ? 1 : 0, va = axis == 2 ? 1 : 2;
            maxface = std::max(maxface, (b[2 * ua + 1] - b[2 * ua] + 1) * (b[2 * va + 1] - b[2 * va] + 1));
        }
        const double ds[3] = {d.dx, d.dy, d.dz};
        for (int a = 0; a < 3; a++) {
            const double Cm = d.dt / MFD_MU0, Ce = d.dt / MFD_EPS0;
            v.cm[a] = (float)(Cm / ds[a]);
            v.ce[a] = (float)(Ce / ds[a]);
            v.cd[a] = (float)(MFD_EPS0 * Ce / ds[a]);
            v.cm64[a] = Cm / ds[a];
            v.ce64[a] = Ce / ds[a];
            v.cd64[a] = MFD_EPS0 * Ce / ds[a];
        }
        if (v.mode == 0) {
            int slotmap[24];
            for (int k = 0; k < 24; k++) slotmap[k] = -1;
            if (d.ndims == 3) {
                for (int k = 0; k < 24; k++) slotmap[k] = k;
            } else if (d.ndims == 2) {
                // incHy_1z, incEz_1z, incHy_2z, incEz_2z, incHy_x1, incEx_x1, incHy_x2, incEx_x2
                slotmap[0] = 0; slotmap[3] = 1; slotmap[4] = 2; slotmap[7] = 3;
                slotmap[17] = 4; slotmap[18] = 5; slotmap[21] = 6; slotmap[22] = 7;
            } else {
                // incHy_1, incEx_1, incHy_2, incEx_2
                slotmap[17] = 0; slotmap[18] = 1; slotmap[21] = 2; slotmap[22] = 3;
            }
            for (int k = 0; k < 24; k++) {
                v.tab[k] = nullptr;
                if (slotmap[k] < 0) continue;
                float *fp;
                UP(upload_vec(arena, &fp, to_f32(hs.tables[slotmap[k]]), io));
                v.tab[k] = fp;
                if (f64) {
                    double *dp;
                    UP(upload_vec(arena, &dp, hs.tables[slotmap[k]], io));
                    v.tab64[k] = dp;
                }
            }
        } else {
            const int Nz = d.Nz;
            UP(arena.alloc(&v.aux, 6 * (size_t)Nz));
            UP(hipMemsetAsync(v.aux, 0, 6 * (size_t)Nz * sizeof(double), io));
            int nbmax = 1;
            for (int f = 0; f < 2; f++) {
                const HostLayer &hl = L[4 + f];
                v.aux_present[f] = hl.present ? 1 : 0;
                v.aux_ib[f] = hl.ib - 1;
                v.aux_nb[f] = hl.nb;
                if (hl.present) {
                    nbmax = std::max(nbmax, hl.nb);
                    double *dp;
                    UP(upload_vec(arena, &dp, hl.K, io)); v.auxK[f] = dp;
                    UP(upload_vec(arena, &dp, hl.A, io)); v.auxA[f] = dp;
                    UP(upload_vec(arena, &dp, hl.B, io)); v.auxB[f] = dp;
                }
            }
            v.aux_nbmax = nbmax;
            UP(arena.alloc(&v.auxpsi, 8 * (size_t)nbmax));
            UP(hipMemsetAsync(v.auxpsi, 0, 8 * (size_t)nbmax * sizeof(double), io));
            v.aux_n = (int)hs.aux_comp.size();
            int *ip;
            UP(upload_vec(arena, &ip, hs.aux_comp, io)); v.aux_comp = ip;
            UP(upload_vec(arena, &ip, hs.aux_iz, io)); v.aux_iz = ip;
            double *dp;
            UP(upload_vec(arena, &dp, hs.aux_coeff, io)); v.aux_coeff = dp;
            UP(upload_vec(arena, &dp, hs.aux_wave, io)); v.aux_wave = dp;
            v.Mh0 = d.Mh; v.Me0 = d.Me; v.Md10 = d.Md1; v.Md20 = d.Md2;
            v.idz = 1.0 / d.dz;
            const size_t tab_n = (size_t)d.Nt * 4 * Nz;
            // an FP64 model steps the aux run live instead (its table would be FP32)
            if (!f64 && tab_n * sizeof(float) <= ((size_t)1 << 30)) {
                float *tab;
                UP(arena.alloc(&tab, tab_n));
                v.auxtab = tab;
            }
        }
    }
    DevSource *dsrc;
    UP(upload_vec(arena, &dsrc, hsrc, io));
    m.src = dsrc;
    m.nsrc = (int)srcs.size();
    for (size_t k = 0; k < hsrc.size(); k++)
        if (hsrc[k].auxtab) {
            launch_tfsf_aux_table(dsrc + k, d.Nz, d.Nt, d.bcs[4], d.bcs[5], (float *)hsrc[k].auxtab, io);
            UP(hipGetLastError());
        }
    // monitors
    int comps[6];
    ncomp_mon = monitor_comps(d.ndims, comps);
    hmon.assign(mons.size(), DevMon{});
    maxmon_n = 1;
    has_spectral = false;
    for (size_t q = 0; q < mons.size(); q++) {
        const HostMon &hm = mons[q];
        DevMon &v = hmon[q];
        v.spectral = hm.spectral;
        v.issum = hm.issum;
        v.n = (int)hm.cells.size();
        v.ncomp = ncomp_mon;
        for (int c = 0; c < 6; c++) v.comps[c] = c < ncomp_mon ? comps[c] : 0;
        int *ip;
        UP(upload_vec(arena, &ip, hm.cells, io)); v.cells = ip;
        maxmon_n = std::max(maxmon_n, v.n);
        if (!hm.spectral) {
            if (hm.issum) {
                UP(arena.alloc(&v.sum, (size_t)d.Nt * ncomp_mon));
                UP(hipMemsetAsync(v.sum, 0, (size_t)d.Nt * ncomp_mon * sizeof(double), io));
            } else {
                UP(arena.alloc(&v.rec, (size_t)d.Nt * ncomp_mon * v.n));
                UP(hipMemsetAsync(v.rec, 0, (size_t)d.Nt * ncomp_mon * v.n * sizeof(float), io));
            }
        } else {
            has_spectral = true;
            v.nw = (int)hm.w.size();
            v.tshift = hm.tshift;
            double *dp;
            UP(upload_vec(arena, &dp, hm.w, io)); v.w = dp;
            UP(arena.alloc(&v.samp, (size_t)dft_block * ncomp_mon * v.n));
            const size_t ns = (size_t)ncomp_mon * v.nw * (hm.issum ? 1 : v.n);
            UP(arena.alloc(&v.S, ns));
            UP(hipMemsetAsync(v.S, 0, ns * sizeof(double2), io));
        }
    }
    DevMon *dmon;
    UP(upload_vec(arena, &dmon, hmon, io));
    m.mon = dmon;
    m.nmon = (int)mons.size();
    m.dft_block = dft_block;

    int *ctr;
    UP(arena.alloc(&ctr, 4));
    UP(hipMemsetAsync(ctr, 0, 4 * sizeof(int), io));
    m.it = ctr;
    m.flag = ctr + 1;
    m.dftdone = ctr + 2;
    // the fills above run on the model's copy stream; runners use other streams, so
    // finish them before any kernel can read this model
    UP(hipStreamSynchronize(io));
    dm = m;
    it = 0;
    dft_done = 0;
    pp_alloc = false;
    pp_parity = 0;
    status = MFD_OK;
    uploaded = true;
    return MFD_OK;
}

// The second H/E buffers and H-half psi arrays of the fused step (3D or 2D), allocated
// the first time a runner steps the model in that form.  Each step writes every cell of
// them, so their initial contents never reach a result; they start zeroed all the same.
int Model::ensure_pp() {
    if (pp_alloc) return MFD_OK;
    const int cs[6] = {HX, HY, HZ, EX, EY, EZ};
    for (int c : cs) {
        if (!comp_present(c)) continue;
        MFD_HIP(arena.alloc(&dm.Falt[c], (size_t)N));
        MFD_HIP(hipMemsetAsync(dm.Falt[c], 0, (size_t)N * sizeof(float), io));
    }
    // paired per axis like the psi arrays of upload()
    for (int a = 0; a < 3; a++) {
        DevLayer &l0 = dm.L[2 * a], &l1 = dm.L[2 * a + 1];
        const size_t n0 = l0.present ? psi_len(*this, 2 * a, l0.nb) : 0;
        const size_t n1 = l1.present ? psi_len(*this, 2 * a + 1, l1.nb) : 0;
        if (n0 + n1 == 0) continue;
        if (d.ndims == 3) {
            v2f *q;
            MFD_HIP(arena.alloc(&q, n0 + n1));
            MFD_HIP(hipMemsetAsync(q, 0, (n0 + n1) * sizeof(v2f), io));
            if (l0.present) l0.pEb = q;
            if (l1.present) l1.pEb = q + n0;
        } else {
            float *q;
            MFD_HIP(arena.alloc(&q, n0 + n1));
            MFD_HIP(hipMemsetAsync(q, 0, (n0 + n1) * sizeof(float), io));
            if (l0.present) l0.psiEb = q;
            if (l1.present) l1.psiEb = q + n0;
        }
    }
    MFD_HIP(hipStreamSynchronize(io));
    pp_alloc = true;
    return MFD_OK;
}

// The touched list without the x-face cells that a fused step finishes, H corrections
// included (tfsf_xhandled_box: on an x face plane, on no y or z face plane, no ADE or
// plasma state); the cells keep the stateful bit of the full list.
int Model::ensure_touched_yz() {
    if (n_touched_yz >= 0) return MFD_OK;
    const int *b = hsrc[0].box;
    std::vector<int> keep;
    for (int c : touched) {
        const int ix = (int)(c % d.Nx), iy = (int)((c / d.Nx) % d.Ny), iz = (int)(c / ((long long)d.Nx * d.Ny));
        const int g = geom[c];
        const bool st = g > 0 && (mats[g - 1].d.nq > 0 || mats[g - 1].d.isplasma);
        const bool onx = (ix == b[0] - 1 || ix == b[0] || ix == b[1]) && iy >= b[2] && iy <= b[3] && iz >= b[4] &&
                         iz <= b[5];
        const bool ony = (iy == b[2] - 1 || iy == b[2] || iy == b[3]) && ix >= b[0] && ix <= b[1];
        const bool onz = (iz == b[4] - 1 || iz == b[4] || iz == b[5]) && ix >= b[0] && ix <= b[1];
        if (onx && !ony && !onz && !st) continue;
        keep.push_back(st ? (c | TOUCHED_STATEFUL) : c);
    }
    int *p = nullptr;
    MFD_HIP(arena.alloc(&p, keep.size()));
    if (!keep.empty()) MFD_HIP(copy_sync(p, keep.data(), keep.size() * sizeof(int), hipMemcpyHostToDevice, io));
    d_touched_yz = p;
    n_touched_yz = (int)keep.size();
    return MFD_OK;
}

// Host mirror of the buffer swap made on the device after a fused step (k_swap_pp or
// k_finish_swap).
void Model::swap_pp() {
    const int cs[6] = {HX, HY, HZ, EX, EY, EZ};
    for (int c : cs) std::swap(dm.F[c], dm.Falt[c]);
    for (int f = 0; f < 6; f++) {
        DevLayer &L = dm.L[f];
        if (d.ndims == 3) std::swap(L.pE, L.pEb);
        else std::swap(f < 2 ? L.psiE1 : L.psiE0, L.psiEb);
    }
    pp_parity ^= 1;
}

// D of the plain cells after runs without stored D: rebuilt as E/Me in place.
int Model::materialize_d() {
    if (!d_stale || !uploaded) return MFD_OK;
    DevModel t = dm;
    t.xtfsf = d_tfsf ? dm.src : nullptr;
    launch_d_from_e(t, io);
    MFD_HIP(hipGetLastError());
    MFD_HIP(hipStreamSynchronize(io));
    d_stale = false;
    return MFD_OK;
}

// Device state is complete whenever no run is in flight: every Runner::run ends with a
// synchronisation of its stream (check_flags), whichever runner stepped the model.
int Model::sync() { return MFD_OK; }

void Model::release() {
    if (io) (void)hipStreamSynchronize(io);
    arena.release();
    if (io) { (void)hipStreamDestroy(io); io = nullptr; }
    uploaded = false;
}

int Model::read_field(int comp, double *out) {
    if (comp < 0 || comp > 8 || !comp_present(comp)) {
        set_error("component not present in this model");
        return MFD_ERR_ARG;
    }
    if (!uploaded) {
        std::fill(out, out + N, 0.0);
        return MFD_OK;
    }
    int st = sync();
    if (st == MFD_OK && (comp == DX || comp == DY || comp == DZ)) st = materialize_d();
    if (st != MFD_OK) return st;
    if (f64) {
        MFD_HIP(copy_sync(out, dm.F64[comp], N * sizeof(double), hipMemcpyDeviceToHost, io));
        return MFD_OK;
    }
    std::vector<float> tmp(N);
    MFD_HIP(copy_sync(tmp.data(), dm.F[comp], N * sizeof(float), hipMemcpyDeviceToHost, io));
    for (long long i = 0; i < N; i++) out[i] = tmp[i];
    return MFD_OK;
}

int Model::write_field(int comp, const double *in) {
    if (comp < 0 || comp > 8 || !comp_present(comp)) {
        set_error("component not present in this model");
        return MFD_ERR_ARG;
    }
    int st = upload();
    if (st != MFD_OK) return st;
    st = sync();
    if (st == MFD_OK) st = materialize_d();
    if (st != MFD_OK) return st;
    user_set = true;        // a state the caller wrote runs with stored D from now on
    if (f64) {
        MFD_HIP(copy_sync(dm.F64[comp], in, N * sizeof(double), hipMemcpyHostToDevice, io));
        return MFD_OK;
    }
    std::vector<float> tmp(N);
    for (long long i = 0; i < N; i++) tmp[i] = (float)in[i];
    MFD_HIP(copy_sync(dm.F[comp], tmp.data(), N * sizeof(float), hipMemcpyHostToDevice, io));
    return MFD_OK;
}

// ---- Runner --------------------------------------------------------------------------
Runner::~Runner() { release(); }

void Runner::release() {
    // work in flight on the runner's stream (a graph replay, the last step's kernels)
    // reads the graph, the descriptors, the tile lists and the flags: wait for it before
    // any of them is destroyed or freed, the graph first
    if (stream) (void)hipStreamSynchronize(stream);
    stream = nullptr;
    if (gexec) { (void)hipGraphExecDestroy(gexec); gexec = nullptr; }
    gsteps = 0;
    gmask = 0;
    if (d_models) { (void)hipFree(d_models); d_models = nullptr; }
    if (d_tiles) { (void)hipFree(d_tiles); d_tiles = nullptr; d_tiles_e = nullptr; }
    if (d_chunks) { (void)hipFree(d_chunks); d_chunks = nullptr; }
    if (d_flags) { (void)hipFree(d_flags); d_flags = nullptr; }
    if (h_flags) { (void)hipHostFree(h_flags); h_flags = nullptr; }
    nchunks = 0;
    for (auto e : ev) (void)hipEventDestroy(e);
    ev.clear();
    if (side) { (void)hipStreamSynchronize(side); (void)hipStreamDestroy(side); side = nullptr; }
    if (fork_ev) { (void)hipEventDestroy(fork_ev); fork_ev = nullptr; }
    if (join_ev) { (void)hipEventDestroy(join_ev); join_ev = nullptr; }
}

static void make_tiles(const Model &m, int idx, int zchunk, int ty, int kform, std::vector<Tile> &out) {
    const int Nx = m.d.Nx, Ny = m.d.Ny, Nz = m.d.Nz;
    if (m.d.ndims == 3 && kform > 0) {
        const int tx = 64;
        for (int z0 = 0; z0 < Nz; z0 += zchunk)
            for (int y0 = 0; y0 < Ny; y0 += ty)
                for (int x0 = 0; x0 < Nx; x0 += tx)
                    out.push_back(make_int4(idx, x0, y0, z0 | (std::min(Nz, z0 + zchunk) << 16)));
    } else if (m.d.ndims == 3) {
        for (int z0 = 0; z0 < Nz; z0 += zchunk)
            for (int y0 = 0; y0 < Ny; y0 += TY3)
                for (int x0 = 0; x0 < Nx; x0 += TX3)
                    out.push_back(make_int4(idx, x0, y0, z0 | (std::min(Nz, z0 + zchunk) << 16)));
    } else if (m.d.ndims == 2) {
        for (int z0 = 0; z0 < Nz; z0 += zchunk)
            for (int x0 = 0; x0 < Nx; x0 += TX2 * ty)
                out.push_back(make_int4(idx, x0, 0, z0 | (std::min(Nz, z0 + zchunk) << 16)));
    } else {
        for (int z0 = 0; z0 < Nz; z0 += TX1)
            out.push_back(make_int4(idx, 0, 0, z0 | (std::min(Nz, z0 + TX1) << 16)));
    }
}

int Runner::init(const std::vector<Model *> &ms, const RunOptions &o, hipStream_t s) {
    release();
    models = ms;
    opt = o;
    stream = s;
    if (ms.empty()) { set_error("no models"); return MFD_ERR_ARG; }
    ndims = ms[0]->d.ndims;
    int dev = 0;
    MFD_HIP(hipGetDevice(&dev));
    MFD_HIP(hipDeviceGetAttribute(&cu_count, hipDeviceAttributeMultiprocessorCount, dev));
    hd.clear();
    std::vector<Tile> tiles, tiles_h;
    any_mats = any_full = any_int = any_mon = any_spectral = h_mats = any_pml = false;
    maxtouched = 1;
    maxsrc = 0;
    maxsrc_n = maxmon_n = maxface = 1;
    dft_block = ms[0]->dft_block;
    long long maxN = 1;
    kform = ndims == 3 ? opt.vec : 0;
    for (Model *m : ms) {
        if (kform >= 2 && m->N >= (1LL << 30)) kform = 0;   // 32-bit cell offsets
    }
    f64 = ms[0]->f64;
    for (size_t k = 0; k < ms.size(); k++) {
        Model *m = ms[k];
        if (m->d.ndims != ndims) { set_error("a run mixes 1D/2D/3D models"); return MFD_ERR_ARG; }
        if (m->f64 != f64) { set_error("a run mixes FP32 and FP64 models"); return MFD_ERR_ARG; }
        int st = m->upload();
        if (st != MFD_OK) return st;
        any_mats |= m->has_mats;
        for (int f = 0; f < 6; f++) any_pml |= m->L[f].present;
        for (const auto &hm : m->mats) {
            any_full |= hm.d.nq > 0 || hm.d.iskerr || hm.d.isplasma;
            // the H half needs the material id only where mu differs from vacuum
            h_mats |= m->has_mats && (float)hm.d.Mh != (float)m->d.Mh;
        }
        maxtouched = std::max(maxtouched, (int)m->touched.size());
        any_int |= m->int_flags != 0;
        any_mon |= !m->mons.empty();
        any_spectral |= m->has_spectral;
        maxsrc = std::max(maxsrc, (int)m->srcs.size());
        maxsrc_n = std::max(maxsrc_n, m->maxsrc_n);
        maxmon_n = std::max(maxmon_n, m->maxmon_n);
        maxface = std::max(maxface, m->maxface);
        maxN = std::max(maxN, m->N);
        if (m->dft_block != dft_block) { set_error("DFT block mismatch"); return MFD_ERR_ARG; }
    }
    // the fused step (kform 3) covers FP32 3D models with mu = 1 everywhere, at least two
    // cells per axis and no cell inside both CPML layers of one axis; otherwise the
    // pipelined two-pass form runs
    if (kform == 3) {
        bool ok = ndims == 3 && !f64 && !h_mats;
        for (Model *m : ms) {
            ok &= m->d.Nx >= 2 && m->d.Ny >= 2 && m->d.Nz >= 2;
            for (int a = 0; a < 3; a++)
                if (m->L[2 * a].present && m->L[2 * a + 1].present) ok &= m->L[2 * a + 1].ib > m->L[2 * a].ib;
        }
        if (!ok) kform = 2;
    }
    fused = kform == 3;
    // the fused 2D step (MFD_OPT_FUSED_2D): FP32 2D models with mu = 1 everywhere, at least
    // two cells along x and z and no cell inside both CPML layers of one axis; otherwise the
    // two-pass 2D kernels run
    fused2 = ndims == 2 && opt.fused_2d && !f64 && !h_mats;
    if (fused2)
        for (Model *m : ms) {
            fused2 &= m->d.Nx >= 2 && m->d.Nz >= 2 && m->N < (1LL << 30);
            for (int a = 0; a < 3; a += 2)
                if (m->L[2 * a].present && m->L[2 * a + 1].present) fused2 &= m->L[2 * a + 1].ib > m->L[2 * a].ib;
        }
    pp_step = fused || fused2;
    // the fused step's integral variables: none, Sa and E2 in every model, or per model
    integ_f = 0;
    for (Model *m : ms)
        if (m->int_flags) integ_f = 3;
    for (Model *m : ms)
        if (integ_f && (m->int_flags & 3) != 3) integ_f = 4;
    frows = (opt.ty == 2) ? 2 : 4;
    fzc = opt.zchunk;
    if (fused && !opt.shape_set) {
        // 10-plane tiles of 2 rows (3 waves per SIMD) while they fill two rounds of
        // waves; with fewer cells, 4-plane tiles of 4 rows (2 waves per SIMD) while they
        // fill one round, else 4-plane tiles of 2 rows.  Measured: 10 x 2 beats 16 x 4 by
        // 2 % at 256^3 and by 9 % on the 384x384x192 dispersive model; at 64^3-160^3 the
        // short tiles run 1.1-3x faster than 16 x 4
        auto count = [&](int zc, int rows) {
            long long t = 0;
            for (Model *m : ms)
                t += (long long)((m->d.Nx + 63) / 64) * ((m->d.Ny + rows - 1) / rows) * ((m->d.Nz + zc - 1) / zc);
            return t;
        };
        const long long slots2 = (long long)cu_count * 4 * 2, slots3 = (long long)cu_count * 4 * 3;
        if (count(10, 2) >= 2 * slots3) { fzc = 10; frows = 2; }
        else if (count(4, 4) >= slots2) { fzc = 4; frows = 4; }
        else { fzc = 4; frows = 2; }
    }
    for (size_t k = 0; k < ms.size(); k++) {
        Model *m = ms[k];
        if (pp_step) {
            int st = m->ensure_pp();
            if (st != MFD_OK) return st;
        }
        hd.push_back(m->dm);
        // the fused step's new fields are read by the next step only.  For a lone model
        // whose two field sets (48 B per cell) outgrow the 256 MB Infinity Cache, storing
        // them past the caches leaves L2 to the old buffers' rows that neighbour tiles
        // re-read as halo: 256^3 331 -> 320 us, 192^3 156 -> 148 us, the dispersive
        // 384x384x192 step 537 -> 518 us.  At 128^3 (100 MB) it gained nothing, and in
        // the served batches (many models per launch) it was 3 % slower
        hd.back().ntst = fused && ms.size() == 1 && (double)m->d.Nx * m->d.Ny * m->d.Nz * 48.0 > 256.0 * (1 << 20);
        make_tiles(*m, (int)k, fused ? fzc : opt.zchunk, fused ? frows : fused2 ? 1 : opt.ty, kform, tiles);
        if (ndims == 3 && kform == 2 && opt.zchunk_h > 0) make_tiles(*m, (int)k, opt.zchunk_h, opt.ty, kform, tiles_h);
    }
    pp_seen.assign(ms.size(), 0);
    for (size_t k = 0; k < ms.size(); k++) pp_seen[k] = ms[k]->pp_parity;
    // cells of stateful materials get their E from the state pass on the pipelined 3D
    // kernels and the 2D kernels; the other forms update them inline
    split_state = (ndims == 3 && kform >= 2) || ndims == 2;
    std::vector<int4> chunks;
    if (split_state)
        for (size_t k = 0; k < ms.size(); k++)
            for (size_t j = 1; j < ms[k]->state_cells.size(); j++) {
                const int cnt = (int)ms[k]->state_cells[j].size();
                for (int s0 = 0; s0 < cnt; s0 += 256)
                    chunks.push_back(make_int4((int)k, (int)j, s0, std::min(cnt, s0 + 256)));
            }
    nchunks = (int)chunks.size();
    if (nchunks) {
        MFD_HIP(hipMalloc(&d_chunks, chunks.size() * sizeof(int4)));
        MFD_HIP(copy_sync(d_chunks, chunks.data(), chunks.size() * sizeof(int4), hipMemcpyHostToDevice, stream));
    }
    // sources: a model whose only source is a TFSF source gets its face corrections and
    // its touched cells' integrals from one gather launch (3D)
    aux_needed = false;
    tfsf_gather = ndims == 3;
    for (Model *m : ms) {
        tfsf_gather &= m->srcs.size() == 1 && m->srcs[0].type == 1;
        for (const DevSource &v : m->hsrc) aux_needed |= v.kind == SRC_TFSF && v.mode == 1 && !v.auxtab;
    }
    // the pipelined 3D E kernel finishes TFSF x-face cells itself
    const bool xface = tfsf_gather && kform >= 2 && (fused ? fzc <= FUSED_XFACE_ZC : opt.zchunk <= XFACE_ZC);
    if (xface)
        for (size_t k = 0; k < ms.size(); k++) hd[k].xtfsf = ms[k]->dm.src;
    // the fused step stores the x faces' H corrections itself, and the gather skips
    // those cells: their sectors lie 4 * Nx bytes apart, most of the gather's traffic
    xface_h = xface && fused && opt.tfsf_fold && !aux_needed;
    if (xface_h) {
        maxtouched = 1;
        for (size_t k = 0; k < ms.size(); k++) {
            int st = ms[k]->ensure_touched_yz();
            if (st != MFD_OK) return st;
            hd[k].xface_h = 1;
            hd[k].touched = ms[k]->d_touched_yz;
            hd[k].ntouched = ms[k]->n_touched_yz;
            maxtouched = std::max(maxtouched, ms[k]->n_touched_yz);
        }
    }
    // the launches after the fused one (state pass, TFSF gather, monitors) read the new
    // buffers where it left them, and k_finish swaps; point sources and the touched-cell
    // integrals still look in F, so with them the swap comes right after the step
    finish_swap = pp_step && (tfsf_gather || (maxsrc == 0 && !any_int));
    dless_models.clear();
    // plain cells without stored D, for a state the caller has not written, on the
    // pipelined 3D kernels and the 2D kernels: D stays stored where the id says Kerr,
    // ADE, plasma or source cell (and, in 3D, on the TFSF faces the x-face path finishes)
    for (size_t k = 0; k < ms.size(); k++) {
        Model *m = ms[k];
        const bool ok = opt.dless && !m->user_set && ((ndims == 3 && kform >= 2) || ndims == 2);
        if (!ok) {
            int st = m->materialize_d();
            if (st != MFD_OK) return st;
            continue;
        }
        hd[k].dless = 1;
        m->d_tfsf = hd[k].xtfsf != nullptr;
        dless_models.push_back(m);
    }
    src_kind_mask.assign(maxsrc, 0);
    for (Model *m : ms)
        for (size_t s2 = 0; s2 < m->srcs.size(); s2++)
            src_kind_mask[s2] |= m->srcs[s2].type == 0 ? 1 : 2;
    // 1D models of up to 64 Ki cells without TFSF sources: one workgroup per model runs
    // whole step ranges (the TFSF path keeps its per-step launches)
    run1d = ndims == 1 && opt.run1d != 0;
    for (Model *m : ms) {
        run1d &= m->d.Nz <= (1 << 16);
        for (const auto &sv : m->srcs) run1d &= sv.type == 0;
    }
    // the register form: up to 4096 cells, no Kerr / ADE / plasma material, fewer than 64
    // materials, CPML layers of up to 64 cells inside the first and the last block of 512
    // cells, at most 8 point-source entries
    reg1d = run1d && opt.run1d == 1 && !any_full;
    for (size_t k = 0; k < ms.size() && reg1d; k++) {
        const DevModel &dm = hd[k];
        const int Nz = dm.Nz, jl = (Nz - 1) / 512;
        reg1d &= Nz <= 4096 && dm.nmat < 64;
        if (dm.L[4].present) reg1d &= dm.L[4].ib < 64;
        if (dm.L[5].present) reg1d &= Nz - dm.L[5].ib <= 64 && dm.L[5].ib / 512 == jl && dm.L[5].ib > 0;
        int nent = 0;
        for (const DevSource &v : ms[k]->hsrc) nent += v.n;
        reg1d &= nent <= 8;
    }
    int_grid = (int)std::min<long long>((maxN + 255) / 256, (long long)cu_count * 16);
    MFD_HIP(hipMalloc(&d_models, hd.size() * sizeof(DevModel)));
    MFD_HIP(hipMalloc(&d_flags, hd.size() * sizeof(int)));
    MFD_HIP(hipHostMalloc(&h_flags, hd.size() * sizeof(int)));
    MFD_HIP(copy_sync(d_models, hd.data(), hd.size() * sizeof(DevModel), hipMemcpyHostToDevice, stream));
    // pipelined 3D: tiles that no CPML layer reaches go to the slab-free kernels
    // the fused step always splits them (the CPML instantiation needs far more registers)
    // and runs the two launches side by side
    const bool split = ndims == 3 && ((kform == 2 && opt.pml_split > 0) || (fused && any_pml && opt.pml_split > 0));
    if (split && opt.pml_split == 2) {
        MFD_HIP(hipStreamCreateWithFlags(&side, hipStreamNonBlocking));
        MFD_HIP(hipEventCreateWithFlags(&fork_ev, hipEventDisableTiming));
        MFD_HIP(hipEventCreateWithFlags(&join_ev, hipEventDisableTiming));
    }
    auto classify = [&](const std::vector<Tile> &all, std::vector<Tile> &tl_pml, std::vector<Tile> &tl_int) {
        for (const Tile &t : all) {
            bool in = !split;
            if (split) {
                const DevModel &dm = hd[t.x];
                const int x1 = std::min(dm.Nx, t.y + 64) - 1, y1 = std::min(dm.Ny, t.z + (fused ? frows : opt.ty)) - 1;
                const int z0 = tile_z0(t), z1 = tile_z1(t) - 1;
                const DevLayer *L = dm.L;
                in = (L[0].present && t.y <= L[0].ib) || (L[1].present && x1 >= L[1].ib) ||
                     (L[2].present && t.z <= L[2].ib) || (L[3].present && y1 >= L[3].ib) ||
                     (L[4].present && z0 <= L[4].ib) || (L[5].present && z1 >= L[5].ib);
                if (fused) {
                    // the fused step also recomputes H one cell below the tile in x, y and z
                    // (wrapping round under periodic faces): those cells count too
                    auto halo_in = [&](int a, int c0, int n) {
                        const int h = c0 > 0 ? c0 - 1 : (dm.bcs[2 * a] == BC_PERIODIC ? n - 1 : -1);
                        if (h < 0) return false;
                        return (L[2 * a].present && h <= L[2 * a].ib) || (L[2 * a + 1].present && h >= L[2 * a + 1].ib);
                    };
                    in = in || halo_in(0, t.y, dm.Nx) || halo_in(1, t.z, dm.Ny) || halo_in(2, z0, dm.Nz);
                }
            }
            (in ? tl_pml : tl_int).push_back(t);
        }
    };
    // the H half may march its own z chunk (MFD_OPT_ZCHUNK_H); both lists cover the grid
    // in the same z-major order, so the snake order still pairs E tiles with the H
    // tiles finished last
    std::vector<Tile> tl_pml, tl_int, th_pml, th_int;
    classify(tiles, tl_pml, tl_int);
    classify(tiles_h.empty() ? tiles : tiles_h, th_pml, th_int);
    // XCD-aware order: workgroups are dispatched round-robin over the 8 XCDs, so block b
    // runs on XCD b % 8.  Give every XCD one contiguous chunk of each tile list, so tiles
    // that share halo rows and planes share that XCD's L2.  The E half-step walks each
    // XCD's chunk backwards (and the lists in the opposite order), so the E kernel starts
    // on the tiles the H kernel finished last (their H is still in the Infinity Cache),
    // and the next H kernel, walking forwards, starts where E finished.
    const int nxcd = 8;
    auto order = [&](const std::vector<Tile> &in, bool backwards, std::vector<Tile> &out,
                     bool chunked = true) {
        if (in.empty()) return;
        if (!chunked) {
            for (size_t b = 0; b < in.size(); b++) out.push_back(in[backwards ? in.size() - 1 - b : b]);
            return;
        }
        // K = 0: one contiguous chunk per XCD; K > 0: blocks of K contiguous tiles dealt
        // to the XCDs in turn (XCD x runs blocks x, x + 8, ...).  At 256^3, 64-tile
        // blocks take E from 258 to 251 us and the dispersive step from 28.9 to 30.5
        // Gcell/s.  A lone 2D solve runs 20 % slower with them and keeps the chunks; a
        // batch of 2D solves takes the blocks.
        const int K = (ndims == 3 || models.size() > 1) ? opt.xcd_block : 0;
        const int per = K > 0 ? ((int)in.size() + nxcd * K - 1) / (nxcd * K) * K
                              : ((int)in.size() + nxcd - 1) / nxcd;
        for (int b = 0; b < per * nxcd; b++) {
            int j = b / nxcd;
            if (backwards) j = per - 1 - j;
            const int x = b % nxcd;
            const int src = K > 0 ? (j / K) * nxcd * K + x * K + j % K : x * per + j;
            out.push_back(src < (int)in.size() ? in[src] : make_int4(-1, 0, 0, 0));
        }
    };
    // the H list is dealt to the XCDs like the E list (MFD_OPT_XCD_H = 0: 3D H tiles in
    // plain grid order, which beats one chunk per XCD but not 64-tile blocks)
    std::vector<Tile> perm, rev;
    const bool hchunk = opt.xcd_h || ndims != 3;
    if (fused) {
        // fused step: three lists by the CPML axes a tile reaches, counting the cells one
        // below it in x, y and z whose H it recomputes (wrapping round under periodic
        // faces): x only (the x-slab columns, half the tiles of a 64-column tiling), any
        // other mix, none.  Each list gets a kernel built for those axes.
        std::vector<Tile> cl[3];
        for (const Tile &t : tiles) {
            const DevModel &dm = hd[t.x];
            if (!split) { cl[1].push_back(t); continue; }   // one launch built for any axis
            const int c0[3] = {t.y, t.z, tile_z0(t)};
            const int c1[3] = {std::min(dm.Nx, t.y + 64) - 1, std::min(dm.Ny, t.z + frows) - 1, tile_z1(t) - 1};
            const int n[3] = {dm.Nx, dm.Ny, dm.Nz};
            int mask = 0;
            for (int a = 0; a < 3; a++) {
                const DevLayer &l0 = dm.L[2 * a], &l1 = dm.L[2 * a + 1];
                const int h = c0[a] > 0 ? c0[a] - 1 : (dm.bcs[2 * a] == BC_PERIODIC ? n[a] - 1 : c0[a]);
                auto hits = [&](int lo, int hi) {
                    return (l0.present && lo <= l0.ib) || (l1.present && hi >= l1.ib);
                };
                if (hits(c0[a], c1[a]) || hits(h, h)) mask |= 1 << a;
            }
            cl[mask == 0 ? 2 : mask == 1 ? 0 : 1].push_back(t);
        }
        for (int c = 0; c < 3; c++) {
            const size_t b = perm.size();
            order(cl[c], false, perm, true);
            nfc[c] = (int)(perm.size() - b);
        }
        ntiles_pml_h = nfc[0] + nfc[1];
    } else {
        order(th_pml, false, perm, hchunk);
        ntiles_pml_h = (int)perm.size();
        order(th_int, false, perm, hchunk);
    }
    order(tl_int, opt.snake, rev);
    ntiles_int_e = (int)rev.size();
    order(tl_pml, opt.snake, rev);
    ntiles_h = (int)perm.size();
    ntiles = (int)rev.size();
    MFD_HIP(hipMalloc(&d_tiles, (perm.size() + rev.size()) * sizeof(Tile)));
    MFD_HIP(copy_sync(d_tiles, perm.data(), perm.size() * sizeof(Tile), hipMemcpyHostToDevice, stream));
    MFD_HIP(copy_sync(d_tiles + perm.size(), rev.data(), rev.size() * sizeof(Tile), hipMemcpyHostToDevice, stream));
    d_tiles_e = d_tiles + perm.size();
    return MFD_OK;
}

int Runner::launch_step(bool timing, bool dft_flush) {
    const int nm = (int)models.size();
    int grid = ntiles;
    if (opt.persistent > 0) grid = std::min(ntiles, cu_count * opt.persistent);
    // pipelined 3D: H list = ntiles_pml_h CPML tiles then the rest (ntiles_h in all);
    // E list = ntiles_int_e CPML-free tiles then the rest (ntiles in all); without the
    // CPML split every tile counts as a CPML tile
    const int nph = ntiles_pml_h, nih = ntiles_h - ntiles_pml_h;
    const int nie = ntiles_int_e, npe = ntiles - ntiles_int_e;
    auto grid_of = [&](int n) { return opt.persistent > 0 ? std::min(n, cu_count * opt.persistent) : n; };
    // the H list (d_tiles, ntiles_h entries) and the E list (d_tiles_e, ntiles) differ in
    // length when only one of them is padded to whole XCD chunks
    const int grid_h = grid_of(ntiles_h);
    if (timing && ev.empty()) {
        ev.resize(4);
        for (auto &e : ev) MFD_HIP(hipEventCreate(&e));
    }
    if (timing) MFD_HIP(hipEventRecord(ev[0], stream));
    const int mode = any_full ? MODE_FULL : any_mats ? MODE_SIMPLE : MODE_VAC;
    const int hmode = h_mats ? MODE_SIMPLE : MODE_VAC;
    if (fused) {
        // x-slab tiles on the side stream; CPML-free tiles, then the other CPML tiles, on
        // the runner's stream, side by side; then the buffer swap
        hipStream_t sp = side ? side : stream;
        const int n1 = nfc[0], n7 = nfc[1], n0 = nfc[2];
        if (side) { MFD_HIP(hipEventRecord(fork_ev, stream)); MFD_HIP(hipStreamWaitEvent(side, fork_ev, 0)); }
        launch_step3d_f(d_models, d_tiles, n1, grid_of(n1), mode, integ_f, frows, 1, sp);
        launch_step3d_f(d_models, d_tiles + n1 + n7, n0, grid_of(n0), mode, integ_f, frows, 0, stream);
        launch_step3d_f(d_models, d_tiles + n1, n7, grid_of(n7), mode, integ_f, frows, 7, stream);
        if (side) { MFD_HIP(hipEventRecord(join_ev, side)); MFD_HIP(hipStreamWaitEvent(stream, join_ev, 0)); }
        if (!finish_swap) launch_swap_pp(d_models, nm, stream);
    }
    else if (ndims == 3 && kform == 2) {
        // side by side (MFD_OPT_PML_SPLIT = 2): the CPML tiles run on the side stream,
        // forked after the previous work on the runner's stream and joined before the next
        hipStream_t sp = side ? side : stream;
        if (side) { MFD_HIP(hipEventRecord(fork_ev, stream)); MFD_HIP(hipStreamWaitEvent(side, fork_ev, 0)); }
        launch_h3d_p(d_models, d_tiles, nph, grid_of(nph), hmode, opt.ty, true, sp);
        launch_h3d_p(d_models, d_tiles + nph, nih, grid_of(nih), hmode, opt.ty, false, stream);
        if (side) { MFD_HIP(hipEventRecord(join_ev, side)); MFD_HIP(hipStreamWaitEvent(stream, join_ev, 0)); }
    }
    else if (ndims == 3) launch_h3d(d_models, d_tiles, ntiles_h, grid_h, hmode, stream);
    else if (fused2) {
        launch_step2d_f(d_models, d_tiles, ntiles_h, grid_h, mode, integ_f, any_pml, stream);
        if (!finish_swap) launch_swap_pp(d_models, nm, stream);
    }
    else if (ndims == 2) launch_h2d(d_models, d_tiles, ntiles_h, grid_h, hmode, opt.ty, stream);
    else launch_h1d(d_models, d_tiles, ntiles_h, grid_h, hmode, f64, stream);
    if (timing) { MFD_HIP(hipEventRecord(ev[1], stream)); MFD_HIP(hipEventRecord(ev[2], stream)); }
    if (fused) {
        // the fused launch did the E half too
    } else if (ndims == 3 && kform == 2) {
        hipStream_t sp = side ? side : stream;
        if (side) { MFD_HIP(hipEventRecord(fork_ev, stream)); MFD_HIP(hipStreamWaitEvent(side, fork_ev, 0)); }
        launch_e3d_p(d_models, d_tiles_e, nie, grid_of(nie), mode, any_int, opt.ty, false, stream);
        launch_e3d_p(d_models, d_tiles_e + nie, npe, grid_of(npe), mode, any_int, opt.ty, true, sp);
        if (side) { MFD_HIP(hipEventRecord(join_ev, side)); MFD_HIP(hipStreamWaitEvent(stream, join_ev, 0)); }
    }
    else if (ndims == 3) launch_e3d(d_models, d_tiles_e, ntiles, grid, mode, any_int, stream);
    else if (fused2) {
        // the fused launch did the E half too
    }
    else if (ndims == 2) launch_e2d(d_models, d_tiles_e, ntiles, grid, mode, any_int, opt.ty, stream);
    else launch_e1d(d_models, d_tiles_e, ntiles, grid, mode, any_int, f64, stream);
    if (split_state && nchunks) launch_e_state(d_models, d_chunks, nchunks, ndims, any_int, stream, finish_swap);
    if (timing) MFD_HIP(hipEventRecord(ev[3], stream));
    if (tfsf_gather) {
        if (aux_needed) launch_tfsf_aux(d_models, nm, 0, stream);
        launch_tfsf_gather(d_models, nm, 0, maxtouched, any_int, stream, finish_swap);
    } else {
        for (int s = 0; s < maxsrc; s++) {
            if (src_kind_mask[s] & 1) launch_point_sources(d_models, nm, s, maxsrc_n, stream);
            if (src_kind_mask[s] & 2) launch_tfsf(d_models, nm, s, ndims, maxface, stream, aux_needed);
        }
    }
    if (any_mon) launch_monitors(d_models, nm, maxmon_n, stream, finish_swap);
    if (dft_flush) launch_dft_flush(d_models, nm, maxmon_n, false, stream);
    if (any_int && !tfsf_gather) launch_integrals_touched(d_models, nm, maxtouched, ndims, stream);
    if (finish_swap) launch_finish_swap(d_models, nm, stream);
    else launch_finish(d_models, nm, stream);
    MFD_HIP(hipGetLastError());
    if (timing) {
        MFD_HIP(hipEventSynchronize(ev[3]));
        float a = 0, b = 0;
        MFD_HIP(hipEventElapsedTime(&a, ev[0], ev[1]));
        MFD_HIP(hipEventElapsedTime(&b, ev[2], ev[3]));
        t_h_ms += a; t_e_ms += b; n_h++; n_e++;
    }
    return MFD_OK;
}

// Host mirror of the device step bookkeeping (k_finish): the step counter, the DFT block
// position of a model with spectral monitors and, for the fused step, the buffer swap.
void Runner::advance(size_t k) {
    Model *m = models[k];
    if (m->it >= m->d.Nt) return;
    m->it++;
    if (m->it - m->dft_done == dft_block) m->dft_done = m->it;
    if (pp_step) {
        m->swap_pp();
        pp_seen[k] = m->pp_parity;
    }
}

// A model's buffers may have been swapped by another runner since this one uploaded its
// descriptor (a model stepped alone, then in a batch, then alone again): take the
// current pointers before launching.
int Runner::refresh() {
    for (size_t k = 0; k < models.size(); k++) {
        Model *m = models[k];
        if (pp_seen[k] == m->pp_parity) continue;
        DevModel &h = hd[k];
        for (int c = 0; c < 9; c++) { h.F[c] = m->dm.F[c]; h.Falt[c] = m->dm.Falt[c]; }
        for (int f = 0; f < 6; f++) {
            h.L[f].pE = m->dm.L[f].pE;
            h.L[f].pEb = m->dm.L[f].pEb;
            h.L[f].psiE0 = m->dm.L[f].psiE0;
            h.L[f].psiE1 = m->dm.L[f].psiE1;
            h.L[f].psiEb = m->dm.L[f].psiEb;
        }
        MFD_HIP(copy_sync(d_models + k, &h, sizeof(DevModel), hipMemcpyHostToDevice, stream));
        pp_seen[k] = m->pp_parity;
    }
    return MFD_OK;
}

// Steps k (0-based) of the next `steps` steps at which some model's DFT block completes:
// k_dft_flush is launched there.  With steps a multiple of the block length, the same
// steps are due on every replay until a model finishes or a read-back flushes a partial
// block.
uint64_t Runner::dft_mask(int steps) const {
    uint64_t mask = 0;
    if (!any_spectral) return 0;
    for (const Model *m : models) {
        if (!m->has_spectral) continue;
        int it = m->it, done = m->dft_done;
        for (int k = 0; k < steps && k < 64 && it < m->d.Nt; k++) {
            if (it + 1 - done == dft_block) { mask |= 1ull << k; done = it + 1; }
            it++;
        }
    }
    return mask;
}

int Runner::build_graph(int steps, uint64_t mask) {
    if (gexec) {
        MFD_HIP(hipStreamSynchronize(stream));
        (void)hipGraphExecDestroy(gexec);
        gexec = nullptr;
    }
    hipGraph_t g;
    MFD_HIP(hipStreamBeginCapture(stream, hipStreamCaptureModeThreadLocal));
    int st = MFD_OK;
    for (int k = 0; k < steps && st == MFD_OK; k++) st = launch_step(false, (mask >> k) & 1);
    hipError_t e = hipStreamEndCapture(stream, &g);
    if (st != MFD_OK) {
        if (e == hipSuccess) (void)hipGraphDestroy(g);
        return st;
    }
    if (e != hipSuccess) { set_error(std::string("graph capture: ") + hipGetErrorString(e)); return MFD_ERR_HIP; }
    e = hipGraphInstantiate(&gexec, g, nullptr, nullptr, 0);
    (void)hipGraphDestroy(g);
    if (e != hipSuccess) { set_error(std::string("graph instantiate: ") + hipGetErrorString(e)); return MFD_ERR_HIP; }
    // the first replay then finds the graph on the device already
    e = hipGraphUpload(gexec, stream);
    if (e != hipSuccess) { set_error(std::string("graph upload: ") + hipGetErrorString(e)); return MFD_ERR_HIP; }
    MFD_HIP(hipStreamSynchronize(stream));
    gsteps = steps;
    gmask = mask;
    return MFD_OK;
}

int Runner::graph_len() const {
    int G = opt.graph_steps;
    if (any_spectral && G > 0) G = dft_block;   // due steps then repeat on every replay
    return std::min(G, 64);
}

// Builds the step graph ahead of the first run that replays it, so that no capture or
// instantiation happens inside a caller's timed steps.
int Runner::prepare() {
    const int G = graph_len();
    if (G <= 0 || opt.kernel_timing) return MFD_OK;
    const uint64_t need = dft_mask(G);
    if (gexec && gsteps == G && (need & ~gmask) == 0) return MFD_OK;
    return build_graph(G, need);
}

// nsteps steps of every model; models that already finished are skipped on the device.
int Runner::run(int nsteps) {
    {
        int st = refresh();
        if (st != MFD_OK) return st;
    }
    for (Model *m : dless_models) m->d_stale = true;
    int done = 0;
    const bool timing = opt.kernel_timing;
    if (run1d && !timing && nsteps > 0) {
        const int mode = any_full ? MODE_FULL : any_mats ? MODE_SIMPLE : MODE_VAC;
        if (reg1d) launch_reg1d(d_models, (int)models.size(), nsteps, mode, any_int, f64, h_mats, stream);
        else launch_run1d(d_models, (int)models.size(), nsteps, mode, any_int, f64, h_mats, stream);
        MFD_HIP(hipGetLastError());
        for (int k = 0; k < nsteps; k++)
            for (size_t q = 0; q < models.size(); q++) advance(q);
        return check_flags();
    }
    const int G = graph_len();
    while (done < nsteps) {
        if (G > 0 && !timing && nsteps - done >= G) {
            int st = prepare();
            if (st != MFD_OK) return st;
            MFD_HIP(hipGraphLaunch(gexec, stream));
            for (int k = 0; k < G; k++)
                for (size_t q = 0; q < models.size(); q++) advance(q);
            done += G;
        } else {
            int st = launch_step(timing, (dft_mask(1) & 1) != 0);
            if (st != MFD_OK) return st;
            for (size_t q = 0; q < models.size(); q++) advance(q);
            done += 1;
        }
    }
    return check_flags();
}

// Contracts every model's partial DFT block (read-back of spectral monitors).
int Runner::flush_dft_partial() {
    if (!any_spectral) return MFD_OK;
    bool any = false;
    for (Model *m : models) any |= m->has_spectral && m->it > m->dft_done;
    if (!any) return MFD_OK;
    const int nm = (int)models.size();
    launch_dft_flush(d_models, nm, maxmon_n, true, stream);
    launch_dft_commit(d_models, nm, stream);
    MFD_HIP(hipGetLastError());
    MFD_HIP(hipStreamSynchronize(stream));
    for (Model *m : models) m->dft_done = m->it;
    return MFD_OK;
}

// Reads every model's non-finite flag; each model keeps its own status, and the first
// failing model's status is returned.
int Runner::check_flags() {
    const int n = (int)models.size();
    launch_collect_flags(d_models, n, d_flags, stream);
    MFD_HIP(hipMemcpyAsync(h_flags, d_flags, n * sizeof(int), hipMemcpyDeviceToHost, stream));
    MFD_HIP(hipStreamSynchronize(stream));
    int first = MFD_OK;
    for (int k = 0; k < n; k++) {
        Model *m = models[k];
        const int f = h_flags[k];
        if (f & 1) m->status = MFD_ERR_NAN;
        else if (f & 2) m->status = MFD_ERR_INF;
        if (m->status != MFD_OK && first == MFD_OK) first = m->status;
    }
    if (first == MFD_ERR_NAN) set_error("Something went wrong. I found NaN field values.");
    if (first == MFD_ERR_INF) set_error("Something went wrong. I found Inf field values.");
    return first;
}

}  // namespace mfd

namespace mfd {
// Launch options shared by single models, batches and the server.  `m` (optional) is
// the model whose pre-upload settings (DFT block, stagger) are changed.
int set_option(RunOptions &o, Model *m, int opt, int value) {
    switch (opt) {
        case MFD_OPT_GRAPH_STEPS: o.graph_steps = std::max(0, value); break;
        case MFD_OPT_ZCHUNK:
            if (value < 1 || value > 65535) { set_error("bad z chunk"); return MFD_ERR_ARG; }
            o.zchunk = value;
            o.shape_set = true;
            break;
        case MFD_OPT_ZCHUNK_H:
            if (value < 0 || value > 65535) { set_error("bad H z chunk"); return MFD_ERR_ARG; }
            o.zchunk_h = value;
            break;
        case MFD_OPT_PERSISTENT: o.persistent = std::max(0, value); break;
        case MFD_OPT_KERNEL_TIMING: o.kernel_timing = value != 0; break;
        case MFD_OPT_DFT_BLOCK:
            if (!m || m->uploaded || value < 1 || value > 64) { set_error("DFT block must be 1..64, set on a model before upload"); return MFD_ERR_ARG; }
            m->dft_block = value;
            break;
        case MFD_OPT_ROWS:
            if (value != 1 && value != 2 && value != 4 && value != 8) { set_error("rows per workgroup must be 1, 2, 4 or 8"); return MFD_ERR_ARG; }
            o.ty = value;
            o.shape_set = true;
            break;
        case MFD_OPT_KERNEL_FORM:
            if (value != 0 && value != 2 && value != 3) {
                set_error("3D kernel form must be 0 (one cell per thread), 2 (pipelined two-pass z march) or 3 (fused step)");
                return MFD_ERR_ARG;
            }
            o.vec = value;
            break;
        case MFD_OPT_SNAKE: o.snake = value != 0; break;
        case MFD_OPT_DLESS: o.dless = value != 0; break;
        case MFD_OPT_PML_SPLIT:
            if (value < 0 || value > 2) { set_error("CPML split must be 0, 1 or 2"); return MFD_ERR_ARG; }
            o.pml_split = value;
            break;
        case MFD_OPT_XCD_H: o.xcd_h = value != 0; break;
        case MFD_OPT_XCD_BLOCK:
            if (value < 0) { set_error("XCD block must be >= 0"); return MFD_ERR_ARG; }
            o.xcd_block = value;
            break;
        case MFD_OPT_F64:
            if (!m || m->uploaded || m->d.ndims != 1 || (value != 0 && value != 1)) {
                set_error("FP64 state: 1D models only, 0 or 1, set before upload");
                return MFD_ERR_ARG;
            }
            m->f64 = value != 0;
            break;
        case MFD_OPT_TFSF_FOLD: o.tfsf_fold = value != 0; break;
        case MFD_OPT_FUSED_2D: o.fused_2d = value != 0; break;
        case MFD_OPT_RUN1D:
            if (value < 0 || value > 2) { set_error("run1d must be 0, 1 or 2"); return MFD_ERR_ARG; }
            o.run1d = value;
            break;
        case MFD_OPT_STAGGER:
            if (!m || m->uploaded || value < 0) { set_error("stagger must be >= 0, set on a model before upload"); return MFD_ERR_ARG; }
            m->stagger = value;
            break;
        default: set_error("unknown option"); return MFD_ERR_ARG;
    }
    return MFD_OK;
}
}  // namespace mfd

// =======================================================================================
// C ABI
// =======================================================================================
using namespace mfd;


static int check_device() {
    static int ok = -1;
    static std::mutex mu;
    std::lock_guard<std::mutex> lk(mu);
    if (ok >= 0) return ok;
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess || n < 1) {
        set_error("no HIP device visible");
        ok = MFD_ERR_NODEV;
        return ok;
    }
    hipDeviceProp_t p;
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess || hipGetDeviceProperties(&p, dev) != hipSuccess) {
        set_error("cannot read the HIP device's properties");
        ok = MFD_ERR_NODEV;
        return ok;
    }
    p.gcnArchName[sizeof(p.gcnArchName) - 1] = '\0';
    if (std::strncmp(p.gcnArchName, "gfx950", 6) != 0) {
        set_error(std::string("device is not gfx950: ") + p.gcnArchName);
        ok = MFD_ERR_NODEV;
        return ok;
    }
    ok = MFD_OK;
    return ok;
}

extern "C" {

int32_t mfd_version(void) { return MFD_ABI_VERSION; }
const char *mfd_last_error(void) { return g_error.c_str(); }

int32_t mfd_device_info(int64_t *free_bytes, int64_t *total_bytes, int32_t *cu_count, char *name,
                        int32_t name_len) {
    int st = check_device();
    if (st != MFD_OK) return st;
    size_t fr = 0, tot = 0;
    MFD_HIP(hipMemGetInfo(&fr, &tot));
    int dev = 0;
    MFD_HIP(hipGetDevice(&dev));
    hipDeviceProp_t p;
    MFD_HIP(hipGetDeviceProperties(&p, dev));
    if (free_bytes) *free_bytes = (int64_t)fr;
    if (total_bytes) *total_bytes = (int64_t)tot;
    if (cu_count) *cu_count = p.multiProcessorCount;
    if (name && name_len > 0) {
        std::snprintf(name, name_len, "%s (%s)", p.name, p.gcnArchName);
    }
    return MFD_OK;
}

int32_t mfd_model_create(const mfd_model_desc *desc, mfd_model **out) {
    if (!desc || !out) { set_error("null argument"); return MFD_ERR_ARG; }
    if (desc->ndims < 1 || desc->ndims > 3 || desc->Nz < 2 || desc->Nx < 1 || desc->Ny < 1 ||
        (desc->ndims == 1 && (desc->Nx != 1 || desc->Ny != 1)) ||
        (desc->ndims == 2 && (desc->Ny != 1 || desc->Nx < 2)) ||
        (desc->ndims == 3 && (desc->Nx < 2 || desc->Ny < 2)) || desc->Nt < 1 ||
        desc->Nz > 65535 || !(desc->dt > 0)) {
        set_error("bad model descriptor");
        return MFD_ERR_ARG;
    }
    // the kernels and the cell lists index cells with 32-bit integers
    if ((long long)desc->Nx * desc->Ny * desc->Nz > 2147483647LL) {
        set_error("bad model descriptor: more than 2^31 - 1 cells");
        return MFD_ERR_ARG;
    }
    for (int i = 0; i < 6; i++)
        if (desc->bcs[i] < 1 || desc->bcs[i] > 3) {
            set_error("Wrong 'bc' value. It can be either ':periodic', ':dirichlet', or ':neumann'.");
            return MFD_ERR_ARG;
        }
    int st = check_device();
    if (st != MFD_OK) return st;
    mfd_model *h = new mfd_model();
    h->m.d = *desc;
    h->m.N = (long long)desc->Nx * desc->Ny * desc->Nz;
    h->m.geom.assign(h->m.N, 0);
    *out = h;
    return MFD_OK;
}

#define NEED_HOST(h)                                                                    \
    do {                                                                                \
        if (!(h)) { set_error("null model"); return MFD_ERR_ARG; }                      \
        if ((h)->m.uploaded) { set_error("model already uploaded"); return MFD_ERR_STATE; } \
    } while (0)

int32_t mfd_model_set_geometry(mfd_model *h, const int32_t *ids) {
    NEED_HOST(h);
    h->m.has_geom = false;
    for (long long i = 0; i < h->m.N; i++) {
        if (ids[i] < 0 || ids[i] > 127) { set_error("geometry ids must be in 0..127"); return MFD_ERR_ARG; }
        h->m.geom[i] = ids[i];
        if (ids[i] > 0) h->m.has_geom = true;
    }
    return MFD_OK;
}

int32_t mfd_model_add_material(mfd_model *h, const mfd_material_desc *mat) {
    NEED_HOST(h);
    if (!mat || mat->nq < 0) { set_error("bad material"); return MFD_ERR_ARG; }
    if (h->m.mats.size() >= 127) { set_error("at most 127 materials per model"); return MFD_ERR_ARG; }
    HostMat hm;
    hm.d = *mat;
    hm.Aq.assign(mat->Aq, mat->Aq + mat->nq);
    hm.Bq.assign(mat->Bq, mat->Bq + mat->nq);
    hm.Cq.assign(mat->Cq, mat->Cq + mat->nq);
    if (mat->isplasma) {
        if (mat->rate_kind == MFD_RATE_LOGLOG) {
            if (mat->rate_n < 2) { set_error("loglog rate table needs >= 2 points"); return MFD_ERR_ARG; }
            hm.rx.assign(mat->rate_x, mat->rate_x + mat->rate_n);
            hm.ry.assign(mat->rate_y, mat->rate_y + mat->rate_n);
        } else if (mat->rate_kind != MFD_RATE_POWER) {
            set_error("plasma material needs a power-law or log-log ionization rate");
            return MFD_ERR_ARG;
        }
    }
    hm.d.Aq = hm.d.Bq = hm.d.Cq = nullptr;
    hm.d.rate_x = hm.d.rate_y = nullptr;
    h->m.mats.push_back(std::move(hm));
    return MFD_OK;
}

int32_t mfd_model_set_pml(mfd_model *h, int32_t face, int32_t ib, int32_t nb, const double *K,
                          const double *A, const double *B) {
    NEED_HOST(h);
    const auto &d = h->m.d;
    const int n = face < 2 ? d.Nx : face < 4 ? d.Ny : d.Nz;
    if (face < 0 || face > 5 || nb < 1 || ib < 1 || ib > n ||
        (face % 2 == 0 && nb != ib) || (face % 2 == 1 && nb != n - ib + 1)) {
        set_error("bad pml layer");
        return MFD_ERR_ARG;
    }
    if ((d.ndims == 1 && face < 4) || (d.ndims == 2 && (face == 2 || face == 3))) {
        set_error("pml face not present for this dimensionality");
        return MFD_ERR_ARG;
    }
    HostLayer &l = h->m.L[face];
    l.present = true;
    l.ib = ib;
    l.nb = nb;
    l.K.assign(K, K + nb);
    l.A.assign(A, A + nb);
    l.B.assign(B, B + nb);
    return MFD_OK;
}

int32_t mfd_model_add_source(mfd_model *h, int32_t kind, int32_t comp, double coeff, int32_t n,
                             const int32_t *cells, const double *waveform) {
    NEED_HOST(h);
    if (n < 1) { set_error("I did not find any grid points which satisfy your source geometry."); return MFD_ERR_ARG; }
    if (comp < 0 || comp > 8 || !h->m.comp_present(comp) || (kind != MFD_SRC_SOFT && kind != MFD_SRC_HARD)) {
        set_error("bad source component or kind");
        return MFD_ERR_ARG;
    }
    HostSource s;
    s.type = 0;
    s.kind = kind;
    s.comp = comp;
    s.coeff = coeff;
    s.cells.assign(cells, cells + n);
    for (int c : s.cells)
        if (c < 0 || c >= h->m.N) { set_error("source cell out of range"); return MFD_ERR_ARG; }
    s.wave.assign(waveform, waveform + (size_t)n * h->m.d.Nt);
    h->m.srcs.push_back(std::move(s));
    return MFD_OK;
}

static size_t tfsf_len(const mfd_model_desc &d, const int *b, int k) {
    const size_t Nt = d.Nt;
    const int Nxi = b[1] - b[0], Nyi = b[3] - b[2], Nzi = b[5] - b[4];
    if (d.ndims == 1) return Nt;
    if (d.ndims == 2) return (size_t)(k < 4 ? Nzi : Nxi) * Nt;
    const int face = k / 4, j = k % 4;
    const bool big = j == 0 || j == 3;
    size_t a, c;
    if (face < 2) { a = big ? Nyi + 1 : Nyi; c = big ? Nzi : Nzi + 1; }
    else if (face < 4) { a = big ? Nxi + 1 : Nxi; c = big ? Nzi : Nzi + 1; }
    else { a = big ? Nxi + 1 : Nxi; c = big ? Nyi : Nyi + 1; }
    return a * c * Nt;
}

int32_t mfd_model_add_tfsf(mfd_model *h, const mfd_tfsf_desc *t) {
    NEED_HOST(h);
    const auto &d = h->m.d;
    if (!t) { set_error("null tfsf"); return MFD_ERR_ARG; }
    const int Ns[3] = {d.Nx, d.Ny, d.Nz};
    for (int a = 0; a < 3; a++) {
        if (d.ndims == 1 && a < 2) continue;
        if (d.ndims == 2 && a == 1) continue;
        const int i1 = t->box[2 * a], i2 = t->box[2 * a + 1];
        if (!(i1 >= 2 && i1 < i2 && i2 <= Ns[a])) {
            set_error("TFSF box must lie strictly inside the grid: 2 <= i1 < i2 <= N");
            return MFD_ERR_ARG;
        }
    }
    HostSource s;
    s.type = 1;
    for (int i = 0; i < 6; i++) s.box[i] = t->box[i];
    s.mode = t->mode;
    if (t->mode == MFD_TFSF_TABLES) {
        const int need = d.ndims == 1 ? 4 : d.ndims == 2 ? 8 : 24;
        if (t->ntables != need || !t->tables) { set_error("wrong number of TFSF tables"); return MFD_ERR_ARG; }
        for (int k = 0; k < need; k++) {
            const size_t len = tfsf_len(d, t->box, k);
            s.tables.emplace_back(t->tables[k], t->tables[k] + len);
        }
    } else if (t->mode == MFD_TFSF_AUX1D) {
        if (t->aux_n < 1) { set_error("plane-wave TFSF needs an auxiliary source"); return MFD_ERR_ARG; }
        for (int k = 0; k < t->aux_n; k++) {
            const int c = t->aux_comp[k];
            if (!(c == MFD_EX || c == MFD_EY || c == MFD_DX || c == MFD_DY || c == MFD_HX || c == MFD_HY) ||
                t->aux_iz[k] < 0 || t->aux_iz[k] >= d.Nz) {
                set_error("bad auxiliary TFSF source");
                return MFD_ERR_ARG;
            }
        }
        s.aux_comp.assign(t->aux_comp, t->aux_comp + t->aux_n);
        s.aux_iz.assign(t->aux_iz, t->aux_iz + t->aux_n);
        s.aux_coeff.assign(t->aux_coeff, t->aux_coeff + t->aux_n);
        s.aux_wave.assign(t->aux_wave, t->aux_wave + (size_t)t->aux_n * d.Nt);
    } else {
        set_error("bad TFSF mode");
        return MFD_ERR_ARG;
    }
    h->m.srcs.push_back(std::move(s));
    return MFD_OK;
}

int32_t mfd_model_add_field_monitor(mfd_model *h, int32_t n, const int32_t *cells, int32_t issum) {
    NEED_HOST(h);
    if (n < 1) { set_error("I did not find any grid points which satisfy your monitor geometry."); return MFD_ERR_ARG; }
    HostMon m;
    m.spectral = false;
    m.issum = issum != 0;
    m.cells.assign(cells, cells + n);
    for (int c : m.cells)
        if (c < 0 || c >= h->m.N) { set_error("monitor cell out of range"); return MFD_ERR_ARG; }
    h->m.mons.push_back(std::move(m));
    return MFD_OK;
}

int32_t mfd_model_add_spectral_monitor(mfd_model *h, int32_t n, const int32_t *cells, int32_t nw,
                                       const double *w, double tshift, int32_t issum) {
    NEED_HOST(h);
    if (n < 1) { set_error("I did not find any grid points which satisfy your monitor geometry."); return MFD_ERR_ARG; }
    if (nw < 1) { set_error("spectral monitor needs at least one frequency"); return MFD_ERR_ARG; }
    HostMon m;
    m.spectral = true;
    m.issum = issum != 0;
    m.cells.assign(cells, cells + n);
    for (int c : m.cells)
        if (c < 0 || c >= h->m.N) { set_error("monitor cell out of range"); return MFD_ERR_ARG; }
    m.w.assign(w, w + nw);
    m.tshift = tshift;
    h->m.mons.push_back(std::move(m));
    return MFD_OK;
}

int32_t mfd_model_enable_integrals(mfd_model *h, int32_t flags) {
    NEED_HOST(h);
    h->m.int_flags = flags & (MFD_INT_SA | MFD_INT_E2);
    return MFD_OK;
}

int32_t mfd_model_footprint(mfd_model *h, int64_t *bytes) {
    if (!h || !bytes) { set_error("null argument"); return MFD_ERR_ARG; }
    *bytes = h->m.uploaded ? (int64_t)h->m.arena.bytes : h->m.footprint();
    return MFD_OK;
}

int32_t mfd_model_upload(mfd_model *h) {
    if (!h) { set_error("null model"); return MFD_ERR_ARG; }
    return h->m.upload();
}

static int ensure_runner(mfd_model *h) {
    if (h->runner_ready) return MFD_OK;
    if (!h->stream) MFD_HIP(hipStreamCreateWithFlags(&h->stream, hipStreamNonBlocking));
    int st = h->runner.init({&h->m}, h->opt, h->stream);
    if (st != MFD_OK) return st;
    h->runner_ready = true;
    return MFD_OK;
}

int32_t mfd_model_step(mfd_model *h, int32_t nsteps) {
    if (!h || nsteps < 0) { set_error("bad arguments"); return MFD_ERR_ARG; }
    if (h->owner) { set_error("model belongs to a batch or server"); return MFD_ERR_STATE; }
    if (h->m.it + nsteps > h->m.d.Nt) { set_error("model already ran all Nt steps"); return MFD_ERR_STATE; }
    int st = ensure_runner(h);
    if (st == MFD_OK) st = h->runner.prepare();
    if (st != MFD_OK) return st;
    return h->runner.run(nsteps);
}

int32_t mfd_model_get_status(mfd_model *h, int32_t *status) {
    if (!h || !status) { set_error("null argument"); return MFD_ERR_ARG; }
    *status = h->m.status;
    return MFD_OK;
}

int32_t mfd_model_get_it(mfd_model *h, int32_t *it) {
    if (!h || !it) { set_error("null argument"); return MFD_ERR_ARG; }
    *it = h->m.it;
    return MFD_OK;
}

int32_t mfd_model_get_field(mfd_model *h, int32_t comp, double *out) {
    if (!h || !out) { set_error("null argument"); return MFD_ERR_ARG; }
    return h->m.read_field(comp, out);
}

int32_t mfd_model_set_field(mfd_model *h, int32_t comp, const double *in) {
    if (!h || !in) { set_error("null argument"); return MFD_ERR_ARG; }
    if (h->owner) { set_error("model belongs to a batch or server"); return MFD_ERR_STATE; }
    const int st = h->m.write_field(comp, in);
    if (st == MFD_OK && h->runner_ready) {   // the launch plan depends on user-written state
        h->runner.release();
        h->runner_ready = false;
    }
    return st;
}

int32_t mfd_model_get_monitor(mfd_model *h, int32_t imon, int32_t comp, double *out) {
    if (!h || !out || imon < 0 || imon >= (int)h->m.mons.size()) { set_error("bad monitor"); return MFD_ERR_ARG; }
    Model &m = h->m;
    const HostMon &hm = m.mons[imon];
    int comps[6];
    const int nc = monitor_comps(m.d.ndims, comps);
    int ci = -1;
    for (int c = 0; c < nc; c++) if (comps[c] == comp) ci = c;
    if (ci < 0) { set_error("component not recorded by monitors of this model"); return MFD_ERR_ARG; }
    const size_t n = hm.cells.size(), Nt = m.d.Nt;
    if (!m.uploaded) {
        const size_t len = hm.spectral ? 2 * hm.w.size() * (hm.issum ? 1 : n) : (hm.issum ? Nt : n * Nt);
        std::fill(out, out + len, 0.0);
        return MFD_OK;
    }
    int st = ensure_runner(h);
    if (st == MFD_OK) st = m.sync();
    if (st == MFD_OK) st = h->runner.flush_dft_partial();
    if (st != MFD_OK) return st;
    const DevMon &dmn = m.hmon[imon];
    if (!hm.spectral) {
        if (hm.issum) {
            std::vector<double> tmp(Nt * nc);
            MFD_HIP(copy_sync(tmp.data(), dmn.sum, tmp.size() * sizeof(double), hipMemcpyDeviceToHost, m.io));
            for (size_t t = 0; t < Nt; t++) out[t] = tmp[t * nc + ci];
        } else {
            std::vector<float> tmp(Nt * nc * n);
            MFD_HIP(copy_sync(tmp.data(), dmn.rec, tmp.size() * sizeof(float), hipMemcpyDeviceToHost, m.io));
            for (size_t t = 0; t < Nt; t++)
                for (size_t k = 0; k < n; k++) out[k + n * t] = tmp[(t * nc + ci) * n + k];
        }
    } else {
        const size_t nw = hm.w.size(), np = hm.issum ? 1 : n;
        std::vector<double2> tmp(nc * nw * np);
        MFD_HIP(copy_sync(tmp.data(), dmn.S, tmp.size() * sizeof(double2), hipMemcpyDeviceToHost, m.io));
        for (size_t w = 0; w < nw; w++)
            for (size_t k = 0; k < np; k++) {
                const double2 v = tmp[((size_t)ci * nw + w) * np + k];
                out[2 * (k + np * w)] = v.x;
                out[2 * (k + np * w) + 1] = v.y;
            }
    }
    return MFD_OK;
}

int32_t mfd_model_get_output(mfd_model *h, int32_t which, double *out) {
    if (!h || !out) { set_error("null argument"); return MFD_ERR_ARG; }
    Model &m = h->m;
    const long long N = m.N;
    if (!m.uploaded) { std::fill(out, out + N, 0.0); return MFD_OK; }
    int st = m.sync();
    if (st != MFD_OK) return st;
    if ((which == MFD_OUT_SA || which == MFD_OUT_E2) && m.f64) {
        const double *src = which == MFD_OUT_SA ? m.dm.Sa64 : m.dm.E2_64;
        if (!src) { set_error("integral not enabled"); return MFD_ERR_STATE; }
        MFD_HIP(copy_sync(out, src, N * sizeof(double), hipMemcpyDeviceToHost, m.io));
        return MFD_OK;
    }
    if (which == MFD_OUT_SA || which == MFD_OUT_E2) {
        const float *src = which == MFD_OUT_SA ? m.dm.Sa : m.dm.E2;
        if (!src) { set_error("integral not enabled"); return MFD_ERR_STATE; }
        std::vector<float> tmp(N);
        MFD_HIP(copy_sync(tmp.data(), src, N * sizeof(float), hipMemcpyDeviceToHost, m.io));
        for (long long i = 0; i < N; i++) out[i] = tmp[i];
        return MFD_OK;
    }
    if (which != MFD_OUT_JE && which != MFD_OUT_RHO) { set_error("bad output"); return MFD_ERR_ARG; }
    std::fill(out, out + N, 0.0);
    for (size_t k = 0; k < m.mats.size(); k++) {
        if (!m.mats[k].d.isplasma) continue;
        const DevMat &v = m.hmats[k + 1];
        const int cnt = v.cnt;
        if (cnt == 0) continue;
        std::vector<double> vals(cnt);
        if (which == MFD_OUT_JE && m.f64) {
            MFD_HIP(copy_sync(vals.data(), m.hmats64[k + 1].JE, cnt * sizeof(double),
                              hipMemcpyDeviceToHost, m.io));
        } else if (which == MFD_OUT_JE) {
            std::vector<float> tmp(cnt);
            MFD_HIP(copy_sync(tmp.data(), v.JE, cnt * sizeof(float), hipMemcpyDeviceToHost, m.io));
            for (int i = 0; i < cnt; i++) vals[i] = tmp[i];
        } else {
            std::vector<double2> tmp(cnt);
            MFD_HIP(copy_sync(tmp.data(), v.rhod, cnt * sizeof(double2), hipMemcpyDeviceToHost, m.io));
            for (int i = 0; i < cnt; i++) vals[i] = tmp[i].x * m.mats[k].d.rho0;
        }
        int s = 0;
        for (long long i = 0; i < N; i++)
            if (m.geom[i] == (int)k + 1) out[i] += vals[s++];
    }
    return MFD_OK;
}

int32_t mfd_model_destroy(mfd_model *h) {
    if (!h) return MFD_OK;
    if (h->owner) { set_error("model belongs to a batch or server"); return MFD_ERR_STATE; }
    if (h->stream) (void)hipStreamSynchronize(h->stream);
    h->runner.release();
    h->m.release();
    if (h->stream) (void)hipStreamDestroy(h->stream);
    delete h;
    return MFD_OK;
}

// Engine options (not part of the reference interface): 0 graph steps per replay,
// 1 z planes per tile, 2 persistent workgroups per CU (0 = one per tile),
// 3 kernel timing (hipEvent pairs around the H/E launches), 4 DFT time-block length,
// 5 rows per workgroup of the 3D tiled kernels, 6 3D kernel form (0 scalar, 1 vector,
// 2 pipelined), 7 per-array stagger of the field pool (bytes).
int32_t mfd_model_set_options(mfd_model *h, int32_t opt, int32_t value) {
    if (!h) { set_error("null model"); return MFD_ERR_ARG; }
    int st = set_option(h->opt, &h->m, opt, value);
    if (st != MFD_OK) return st;
    if (h->runner_ready) {
        h->runner.release();
        h->runner_ready = false;
    }
    return MFD_OK;
}

// Accumulated hipEvent timing of the H (which = 0) or E (which = 1) launches.
int32_t mfd_model_timing(mfd_model *h, int32_t which, double *total_ms, int32_t *count) {
    if (!h || !total_ms || !count) { set_error("null argument"); return MFD_ERR_ARG; }
    *total_ms = which == 0 ? h->runner.t_h_ms : h->runner.t_e_ms;
    *count = (int32_t)(which == 0 ? h->runner.n_h : h->runner.n_e);
    if (which == 2) { h->runner.t_h_ms = h->runner.t_e_ms = 0; h->runner.n_h = h->runner.n_e = 0; }
    return MFD_OK;
}

}  // extern "C"
