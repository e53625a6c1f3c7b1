/*
 * fdtd_oracle.c -- CPU restatement of MaxwellFDTD's per-time-step field update.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the checker the GPU engine is compared with;
 * it is used by tests/, by __graft_entry__.smoke() and by bench.py's cpu_baseline leg,
 * and never by the product path.
 *
 * It follows the reference statement by statement, with 0-based indices:
 *   update_H_kernel!/update_E_kernel! 1D  src/models.jl:248-444
 *   update_H_kernel!/update_E_kernel! 2D  src/models.jl:450-726
 *   update_H_kernel!/update_E_kernel! 3D  src/models.jl:735-1175
 *   step! order H -> E -> sources         src/models.jl:125-132
 *   solve! loop (NaN/Inf guard, monitors, integrals)   src/models.jl:199-234
 *   soft / hard sources                   src/sources.jl:77-131,190-244
 *   TFSF replay                           src/sources.jl:320-346,469-497,713-776
 *   field / spectral monitors, DFT        src/monitors.jl:62-66,110-116,164-176,350-359,
 *                                         src/dft.jl:27-51
 *   integral variables                    src/outputs.jl:182-191,318-327,475-485
 * ADE, plasma and Kerr state is kept per material over the full grid, as
 * MaterialStruct3D does (src/materials.jl:424-426,464-474).
 *
 * Compiled twice: -DOFD_REAL=double (the FP64 reference run) and -DOFD_REAL=float
 * (the reference's FP32 GPU arithmetic, used to measure the intrinsic FP32 gap).
 * The exported symbols mirror include/maxwell_fdtd.h with the prefix ofd_.
 *
 * How each part is pinned (tests/): the reference's own propagation and loss tests;
 * the exact one-cell-per-step translation at CN = 1; ADE coefficients against chi(w);
 * Kerr against the cubic; plasma density against its closed form; the plasma current
 * against the Drude response and the multiphoton loss against its energy balance;
 * Dirichlet/Neumann faces against mirror-image periodic extensions in 1D/2D/3D; the
 * CPML update (kappa >= 1, both layers) against a quadrature of the CPML memory
 * kernel; the log-log rate table against the reference's SiO2_keldysh_1.03um.tf rows;
 * TFSF replay against the auxiliary-run identity; DFT against brute-force sums.
 * Parity unpinned: nothing in this file; mismatched-grid TFSF interpolation and the
 * HDF5 layout live outside it.
 */
#include <complex.h>
#include <float.h>
#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#include "../include/maxwell_fdtd.h"

#ifndef OFD_REAL
#define OFD_REAL double
#endif
typedef OFD_REAL real;

#define REAL_IS_FLOAT (sizeof(real) == sizeof(float))
/* eps(one(T)) of the plasma E2 threshold (src/models.jl:1106) */
#define REAL_EPS (REAL_IS_FLOAT ? (real)FLT_EPSILON : (real)DBL_EPSILON)

static char g_err[512];
static void set_err(const char *s) { snprintf(g_err, sizeof g_err, "%s", s); }
const char *ofd_last_error(void) { return g_err; }
int32_t ofd_version(void) { return MFD_ABI_VERSION; }
int32_t ofd_real_size(void) { return (int32_t)sizeof(real); }

int32_t ofd_set_threads(int32_t n) {
#ifdef _OPENMP
    if (n > 0) omp_set_num_threads(n);
    return omp_get_max_threads();
#else
    (void)n;
    return 1;
#endif
}

typedef struct {
    int ib, nb;          /* ib 0-based index of the layer boundary, nb profile length */
    int present;
    real *K, *A, *B;
    real *psiE[2];       /* H-half psi (derivatives of E): order of src/pml.jl:221-249 */
    real *psiH[2];       /* E-half psi (derivatives of H)                              */
} layer_t;

typedef struct {
    real Mh, Me, Md1, Md2;
    int nq;
    real *Aq, *Bq, *Cq;
    real *P[3], *P1[3], *P2[3];          /* Px, oldPx1, oldPx2 ... [nq, N]             */
    int iskerr;
    real Mk2, Mk3;
    int isplasma;
    real Rava, ksi, rho0, Ap, Bp, Cp, Ma;
    int rate_kind;
    double rate_c, rate_k;
    int rate_n;
    double *rate_x, *rate_y;
    real *rho, *drho, *Pp[3], *Pp1[3], *Pp2[3], *Pa[3];
} mat_t;

typedef struct {
    int kind, comp, n;
    real coeff;
    int64_t *cells;
    real *wave;          /* [Nt, n] */
} src_t;

typedef struct {
    int box[6];          /* 1-based ix1,ix2,iy1,iy2,iz1,iz2 */
    int mode;
    int ntab;
    real *tab[24];
    /* aux 1D run */
    int aux_n;
    int *aux_comp, *aux_iz;
    real *aux_coeff;
    real *aux_wave;      /* [aux_n, Nt] */
    real *aHx, *aHy, *aDx, *aDy, *aEx, *aEy;
    real *apsiE[2][2], *apsiH[2][2];    /* [face][x-pol, y-pol] */
} tfsf_t;

typedef struct {
    int spectral, issum, n, nw;
    int64_t *cells;
    double *data[9];     /* field: [n, Nt] or [Nt]; spectral: complex [n, nw] or [nw] */
    double *w;
    double tshift;
} mon_t;

typedef struct ofd_model {
    int status;                 /* MFD_OK, or the non-finite status of the last step */
    int ndims, Nx, Ny, Nz;
    int64_t N;
    real dx, dy, dz, dt;
    double dt64;
    int Nt, it;
    int bcs[6];
    real Mh, Me, Md1, Md2;
    real *F[9];
    int32_t *geom;
    int isgeom;
    layer_t L[6];
    int nmat;
    mat_t mats[256];
    int nsrc;
    int src_kind[64];    /* 0 point source, 1 tfsf */
    src_t src[64];
    tfsf_t tfsf[64];
    int nmon;
    mon_t mon[64];
    int int_flags;
    real *Sa, *E2, *JE;
} ofd_model;

static void *xcalloc(size_t n, size_t s) {
    void *p = calloc(n ? n : 1, s);
    if (!p) { fprintf(stderr, "oracle: out of memory\n"); abort(); }
    return p;
}

static int comp_present(const ofd_model *m, int c) {
    if (m->ndims == 3) return 1;
    if (m->ndims == 2) return c == MFD_HY || c == MFD_DX || c == MFD_DZ || c == MFD_EX || c == MFD_EZ;
    return c == MFD_HY || c == MFD_DX || c == MFD_EX;
}

int32_t ofd_model_create(const mfd_model_desc *d, ofd_model **out) {
    if (!d || !out || d->ndims < 1 || d->ndims > 3 || d->Nx < 1 || d->Ny < 1 || d->Nz < 2 ||
        (d->ndims == 1 && (d->Nx != 1 || d->Ny != 1)) || (d->ndims == 2 && d->Ny != 1) ||
        d->Nt < 1 || !(d->dt > 0)) {
        set_err("bad model descriptor");
        return MFD_ERR_ARG;
    }
    for (int i = 0; i < 6; i++)
        if (d->bcs[i] < 1 || d->bcs[i] > 3) {
            set_err("Wrong 'bc' value. It can be either ':periodic', ':dirichlet', or ':neumann'.");
            return MFD_ERR_ARG;
        }
    ofd_model *m = xcalloc(1, sizeof *m);
    m->ndims = d->ndims; m->Nx = d->Nx; m->Ny = d->Ny; m->Nz = d->Nz;
    m->N = (int64_t)d->Nx * d->Ny * d->Nz;
    m->dx = (real)d->dx; m->dy = (real)d->dy; m->dz = (real)d->dz;
    m->dt = (real)d->dt; m->dt64 = d->dt;
    m->Nt = d->Nt; m->it = 0;
    for (int i = 0; i < 6; i++) m->bcs[i] = d->bcs[i];
    m->Mh = (real)d->Mh; m->Me = (real)d->Me; m->Md1 = (real)d->Md1; m->Md2 = (real)d->Md2;
    for (int c = 0; c < 9; c++)
        if (comp_present(m, c)) m->F[c] = xcalloc((size_t)m->N, sizeof(real));
    m->geom = xcalloc((size_t)m->N, sizeof(int32_t));
    for (int f = 0; f < 6; f++) {
        m->L[f].present = 0;
        m->L[f].ib = (f % 2 == 0) ? -1 : (f < 2 ? m->Nx : f < 4 ? m->Ny : m->Nz);
    }
    *out = m;
    return MFD_OK;
}

int32_t ofd_model_set_geometry(ofd_model *m, const int32_t *ids) {
    m->isgeom = 0;
    for (int64_t i = 0; i < m->N; i++) {
        m->geom[i] = ids[i];
        if (ids[i] > 0) m->isgeom = 1;
    }
    return MFD_OK;
}

int32_t ofd_model_add_material(ofd_model *m, const mfd_material_desc *d) {
    if (m->nmat >= 255) { set_err("too many materials"); return MFD_ERR_ARG; }
    mat_t *t = &m->mats[++m->nmat];
    t->Mh = (real)d->Mh; t->Me = (real)d->Me; t->Md1 = (real)d->Md1; t->Md2 = (real)d->Md2;
    t->nq = d->nq;
    if (t->nq > 0) {
        t->Aq = xcalloc(t->nq, sizeof(real)); t->Bq = xcalloc(t->nq, sizeof(real));
        t->Cq = xcalloc(t->nq, sizeof(real));
        for (int q = 0; q < t->nq; q++) {
            t->Aq[q] = (real)d->Aq[q]; t->Bq[q] = (real)d->Bq[q]; t->Cq[q] = (real)d->Cq[q];
        }
        for (int c = 0; c < 3; c++) {
            t->P[c] = xcalloc((size_t)t->nq * m->N, sizeof(real));
            t->P1[c] = xcalloc((size_t)t->nq * m->N, sizeof(real));
            t->P2[c] = xcalloc((size_t)t->nq * m->N, sizeof(real));
        }
    }
    t->iskerr = d->iskerr; t->Mk2 = (real)d->Mk2; t->Mk3 = (real)d->Mk3;
    t->isplasma = d->isplasma;
    if (t->isplasma) {
        t->Rava = (real)d->Rava; t->ksi = (real)d->ksi; t->rho0 = (real)d->rho0;
        t->Ap = (real)d->Ap; t->Bp = (real)d->Bp; t->Cp = (real)d->Cp; t->Ma = (real)d->Ma;
        t->rate_kind = d->rate_kind; t->rate_c = d->rate_c; t->rate_k = d->rate_k;
        t->rate_n = d->rate_n;
        if (t->rate_kind == MFD_RATE_LOGLOG) {
            if (d->rate_n < 2) { set_err("loglog rate table needs >= 2 points"); return MFD_ERR_ARG; }
            t->rate_x = xcalloc(d->rate_n, sizeof(double));
            t->rate_y = xcalloc(d->rate_n, sizeof(double));
            memcpy(t->rate_x, d->rate_x, d->rate_n * sizeof(double));
            memcpy(t->rate_y, d->rate_y, d->rate_n * sizeof(double));
        }
        t->rho = xcalloc((size_t)m->N, sizeof(real));
        t->drho = xcalloc((size_t)m->N, sizeof(real));
        for (int c = 0; c < 3; c++) {
            t->Pp[c] = xcalloc((size_t)m->N, sizeof(real));
            t->Pp1[c] = xcalloc((size_t)m->N, sizeof(real));
            t->Pp2[c] = xcalloc((size_t)m->N, sizeof(real));
            t->Pa[c] = xcalloc((size_t)m->N, sizeof(real));
        }
        if (!m->JE) m->JE = xcalloc((size_t)m->N, sizeof(real));
    }
    return MFD_OK;
}

int32_t ofd_model_set_pml(ofd_model *m, int32_t face, int32_t ib, int32_t nb,
                          const double *K, const double *A, const double *B) {
    if (face < 0 || face > 5 || nb < 1) { set_err("bad pml face"); return MFD_ERR_ARG; }
    layer_t *L = &m->L[face];
    L->present = 1;
    L->ib = ib - 1;
    L->nb = nb;
    L->K = xcalloc(nb, sizeof(real)); L->A = xcalloc(nb, sizeof(real)); L->B = xcalloc(nb, sizeof(real));
    for (int i = 0; i < nb; i++) { L->K[i] = (real)K[i]; L->A[i] = (real)A[i]; L->B[i] = (real)B[i]; }
    size_t sz;
    if (face < 2) sz = (size_t)nb * m->Ny * m->Nz;
    else if (face < 4) sz = (size_t)m->Nx * nb * m->Nz;
    else sz = (size_t)m->Nx * m->Ny * nb;
    for (int k = 0; k < 2; k++) {
        L->psiE[k] = xcalloc(sz, sizeof(real));
        L->psiH[k] = xcalloc(sz, sizeof(real));
    }
    return MFD_OK;
}

int32_t ofd_model_add_source(ofd_model *m, int32_t kind, int32_t comp, double coeff,
                             int32_t n, const int32_t *cells, const double *wave) {
    if (n < 1) { set_err("I did not find any grid points which satisfy your source geometry."); return MFD_ERR_ARG; }
    if (comp < 0 || comp > 8 || !comp_present(m, comp)) { set_err("bad source component"); return MFD_ERR_ARG; }
    int s = m->nsrc++;
    m->src_kind[s] = 0;
    src_t *S = &m->src[s];
    S->kind = kind; S->comp = comp; S->coeff = (real)coeff; S->n = n;
    S->cells = xcalloc(n, sizeof(int64_t));
    for (int i = 0; i < n; i++) S->cells[i] = cells[i];
    S->wave = xcalloc((size_t)n * m->Nt, sizeof(real));
    for (size_t i = 0; i < (size_t)n * m->Nt; i++) S->wave[i] = (real)wave[i];
    return MFD_OK;
}

static size_t tfsf_table_len(const ofd_model *m, const int *b, int k) {
    int Nxi = b[1] - b[0], Nyi = b[3] - b[2], Nzi = b[5] - b[4];
    size_t Nt = (size_t)m->Nt;
    if (m->ndims == 1) return Nt;
    if (m->ndims == 2) return (size_t)(k < 4 ? Nzi : Nxi) * Nt;
    /* 3D: the 24 arrays of TFSFSource3D, shapes from prepare_tfsf_record
       (src/sources.jl:779-833)                                                        */
    int face = k / 4, j = k % 4;
    size_t a, c;
    if (face < 2) {           /* Hy, Hz, Ey, Ez */
        a = (j == 0 || j == 3) ? Nyi + 1 : Nyi;
        c = (j == 0 || j == 3) ? Nzi : Nzi + 1;
    } else if (face < 4) {    /* Hx, Hz, Ex, Ez */
        a = (j == 0 || j == 3) ? Nxi + 1 : Nxi;
        c = (j == 0 || j == 3) ? Nzi : Nzi + 1;
    } else {                  /* Hx, Hy, Ex, Ey */
        a = (j == 0 || j == 3) ? Nxi + 1 : Nxi;
        c = (j == 0 || j == 3) ? Nyi : Nyi + 1;
    }
    return a * c * Nt;
}

int32_t ofd_model_add_tfsf(ofd_model *m, const mfd_tfsf_desc *d) {
    int s = m->nsrc++;
    m->src_kind[s] = 1;
    tfsf_t *T = &m->tfsf[s];
    memcpy(T->box, d->box, sizeof T->box);
    T->mode = d->mode;
    if (d->mode == MFD_TFSF_TABLES) {
        int need = m->ndims == 1 ? 4 : m->ndims == 2 ? 8 : 24;
        if (d->ntables != need) { set_err("wrong number of TFSF tables"); return MFD_ERR_ARG; }
        T->ntab = need;
        for (int k = 0; k < need; k++) {
            size_t len = tfsf_table_len(m, T->box, k);
            T->tab[k] = xcalloc(len, sizeof(real));
            for (size_t i = 0; i < len; i++) T->tab[k][i] = (real)d->tables[k][i];
        }
    } else {
        T->aux_n = d->aux_n;
        T->aux_comp = xcalloc(d->aux_n, sizeof(int));
        T->aux_iz = xcalloc(d->aux_n, sizeof(int));
        T->aux_coeff = xcalloc(d->aux_n, sizeof(real));
        T->aux_wave = xcalloc((size_t)d->aux_n * m->Nt, sizeof(real));
        for (int k = 0; k < d->aux_n; k++) {
            T->aux_comp[k] = d->aux_comp[k];
            T->aux_iz[k] = d->aux_iz[k];
            T->aux_coeff[k] = (real)d->aux_coeff[k];
            for (int i = 0; i < m->Nt; i++)
                T->aux_wave[(size_t)k * m->Nt + i] = (real)d->aux_wave[(size_t)k * m->Nt + i];
        }
        int Nz = m->Nz;
        T->aHx = xcalloc(Nz, sizeof(real)); T->aHy = xcalloc(Nz, sizeof(real));
        T->aDx = xcalloc(Nz, sizeof(real)); T->aDy = xcalloc(Nz, sizeof(real));
        T->aEx = xcalloc(Nz, sizeof(real)); T->aEy = xcalloc(Nz, sizeof(real));
        for (int f = 0; f < 2; f++)
            for (int p = 0; p < 2; p++) {
                int nb = m->L[4 + f].present ? m->L[4 + f].nb : 1;
                T->apsiE[f][p] = xcalloc(nb, sizeof(real));
                T->apsiH[f][p] = xcalloc(nb, sizeof(real));
            }
    }
    return MFD_OK;
}

static int ncomp_mon(const ofd_model *m) { return m->ndims == 3 ? 6 : m->ndims == 2 ? 3 : 2; }
static int mon_comp(const ofd_model *m, int j) {
    static const int c3[6] = {MFD_HX, MFD_HY, MFD_HZ, MFD_EX, MFD_EY, MFD_EZ};
    static const int c2[3] = {MFD_HY, MFD_EX, MFD_EZ};
    static const int c1[2] = {MFD_HY, MFD_EX};
    return m->ndims == 3 ? c3[j] : m->ndims == 2 ? c2[j] : c1[j];
}

int32_t ofd_model_add_field_monitor(ofd_model *m, int32_t n, const int32_t *cells, int32_t issum) {
    if (n < 1) { set_err("I did not find any grid points which satisfy your monitor geometry."); return MFD_ERR_ARG; }
    mon_t *M = &m->mon[m->nmon++];
    M->spectral = 0; M->issum = issum; M->n = n;
    M->cells = xcalloc(n, sizeof(int64_t));
    for (int i = 0; i < n; i++) M->cells[i] = cells[i];
    for (int j = 0; j < ncomp_mon(m); j++)
        M->data[mon_comp(m, j)] = xcalloc(issum ? (size_t)m->Nt : (size_t)n * m->Nt, sizeof(double));
    return MFD_OK;
}

int32_t ofd_model_add_spectral_monitor(ofd_model *m, int32_t n, const int32_t *cells, int32_t nw,
                                       const double *w, double tshift, int32_t issum) {
    if (n < 1) { set_err("I did not find any grid points which satisfy your monitor geometry."); return MFD_ERR_ARG; }
    mon_t *M = &m->mon[m->nmon++];
    M->spectral = 1; M->issum = issum; M->n = n; M->nw = nw; M->tshift = tshift;
    M->cells = xcalloc(n, sizeof(int64_t));
    for (int i = 0; i < n; i++) M->cells[i] = cells[i];
    M->w = xcalloc(nw, sizeof(double));
    memcpy(M->w, w, nw * sizeof(double));
    for (int j = 0; j < ncomp_mon(m); j++)
        M->data[mon_comp(m, j)] = xcalloc(2 * (issum ? (size_t)nw : (size_t)n * nw), sizeof(double));
    return MFD_OK;
}

int32_t ofd_model_enable_integrals(ofd_model *m, int32_t flags) {
    m->int_flags = flags;
    if ((flags & MFD_INT_SA) && !m->Sa) m->Sa = xcalloc((size_t)m->N, sizeof(real));
    if ((flags & MFD_INT_E2) && !m->E2) m->E2 = xcalloc((size_t)m->N, sizeof(real));
    return MFD_OK;
}

int32_t ofd_model_footprint(ofd_model *m, int64_t *bytes) { (void)m; *bytes = 0; return MFD_OK; }
int32_t ofd_model_upload(ofd_model *m) { (void)m; return MFD_OK; }

/* ------------------------------------------------------------------------------------ */
/* ionrate(I): power law or IonizationRate log-log table                                */
/* (examples/5_nonlinear_propagation/IonizationRate.jl:30-52)                           */
static double ionrate(const mat_t *t, double inten) {
    if (t->rate_kind == MFD_RATE_POWER) {
        if (inten <= 0) return 0.0;
        return t->rate_c * pow(inten, t->rate_k);
    }
    if (t->rate_kind == MFD_RATE_LOGLOG) {
        if (inten <= 0) return 0.0;
        double xl = log10(inten);
        const double *x = t->rate_x, *y = t->rate_y;
        int n = t->rate_n, i;
        if (xl <= x[0]) i = 0;
        else if (xl >= x[n - 1]) i = n - 2;
        else {
            /* searchsortedfirst(x, xl) - 1 (1-based) -> first index with x[i] >= xl, minus 1 */
            int lo = 0, hi = n - 1;
            while (lo < hi) { int mid = (lo + hi) / 2; if (x[mid] >= xl) hi = mid; else lo = mid + 1; }
            i = lo - 1;
        }
        double dydx = (y[i + 1] - y[i]) / (x[i + 1] - x[i]);
        return pow(10.0, y[i] + dydx * (xl - x[i]));
    }
    return 0.0;
}

/* Per-cell material update of the E half-step after the D update: ADE poles, plasma,
   Kerr, E = (D - sum P) * Me.  src/models.jl:1043-1154 (3D); :651-714 (2D); :362-430
   (1D).  nc components given by comp list (indices into P arrays 0..2 = x,y,z).        */
static void material_e(ofd_model *m, int64_t i, int g, int nc, const int *cc,
                       real *Dv[], real *Ev[]) {
    real sumP[3] = {0, 0, 0};
    real Me;
    int isdisp = 0, iskerr = 0, isplasma = 0;
    mat_t *t = NULL;
    if (g > 0) {
        t = &m->mats[g];
        Me = t->Me;
        isdisp = t->nq > 0; iskerr = t->iskerr; isplasma = t->isplasma;
    } else {
        Me = m->Me;
    }
    if (isdisp) {
        int nq = t->nq;
        for (int q = 0; q < nq; q++) {
            for (int k = 0; k < nc; k++) {
                int c = cc[k];
                int64_t j = q + (int64_t)nq * i;
                t->P2[c][j] = t->P1[c][j];
                t->P1[c][j] = t->P[c][j];
                t->P[c][j] = t->Aq[q] * t->P[c][j] + t->Bq[q] * t->P2[c][j] + t->Cq[q] * Ev[k][i];
            }
            for (int k = 0; k < nc; k++) sumP[k] += t->P[cc[k]][q + (int64_t)nq * i];
        }
    }
    if (isplasma) {
        real E2 = 0;
        for (int k = 0; k < nc; k++) E2 += Ev[k][i] * Ev[k][i];
        real II = t->ksi * E2;
        for (int k = 0; k < nc; k++) {
            int c = cc[k];
            t->Pp2[c][i] = t->Pp1[c][i];
            t->Pp1[c][i] = t->Pp[c][i];
            t->Pp[c][i] = t->Ap * t->Pp[c][i] + t->Bp * t->Pp2[c][i] + t->Cp * t->rho[i] * t->rho0 * Ev[k][i];
        }
        for (int k = 0; k < nc; k++) sumP[k] += t->Pp[cc[k]][i];
        real invE2 = (E2 >= REAL_EPS) ? (real)1 / E2 : (real)0;
        for (int k = 0; k < nc; k++) {
            int c = cc[k];
            t->Pa[c][i] += t->Ma * t->drho[i] * t->rho0 * Ev[k][i] * invE2;
        }
        for (int k = 0; k < nc; k++) sumP[k] += t->Pa[cc[k]][i];
        real R1 = (real)ionrate(t, (double)II);
        real R2 = t->Rava * E2;
        if (R2 == 0) {
            t->rho[i] = 1 - (1 - t->rho[i]) * (real)exp(-(double)(R1 * m->dt));
        } else {
            real R12 = R1 - R2;
            t->rho[i] = R1 / R12 * 1 - (R1 / R12 * 1 - t->rho[i]) * (real)exp(-(double)(R12 * m->dt));
        }
        t->drho[i] = R1 * (1 - t->rho[i]);
        real JE = 0;
        for (int k = 0; k < nc; k++) {
            int c = cc[k];
            real J = (t->Pp[c][i] - t->Pp2[c][i]) / (2 * m->dt);
            JE += J * Ev[k][i];
        }
        m->JE[i] += JE * m->dt;
    }
    for (int k = 0; k < nc; k++) {
        real u = Dv[k][i] - sumP[k];
        if (iskerr) {
            real Mk2 = t->Mk2, Mk3 = t->Mk3;
            Ev[k][i] = (1 + 1 * Mk2 * (Me * Me) * u + 2 * Mk3 * (Me * Me * Me) * (u * u)) /
                       (1 + 2 * Mk2 * (Me * Me) * u + 3 * Mk3 * (Me * Me * Me) * (u * u)) * u * Me;
        } else {
            Ev[k][i] = u * Me;
        }
    }
}

static inline real bcval(int bc, real periodic, real own) {
    return bc == MFD_BC_PERIODIC ? periodic : bc == MFD_BC_DIRICHLET ? (real)0 : own;
}

/* ---------------- 3D: src/models.jl:735-1175 ---------------------------------------- */
static void update_H_3d(ofd_model *m) {
    const int Nx = m->Nx, Ny = m->Ny, Nz = m->Nz;
    real *Hx = m->F[MFD_HX], *Hy = m->F[MFD_HY], *Hz = m->F[MFD_HZ];
    const real *Ex = m->F[MFD_EX], *Ey = m->F[MFD_EY], *Ez = m->F[MFD_EZ];
    layer_t *L = m->L;
#pragma omp parallel for schedule(static)
    for (int iz = 0; iz < Nz; iz++)
        for (int iy = 0; iy < Ny; iy++)
            for (int ix = 0; ix < Nx; ix++) {
                int64_t i = ix + (int64_t)Nx * (iy + (int64_t)Ny * iz);
                int g = m->geom[i];
                real Mh = g > 0 ? m->mats[g].Mh : m->Mh;
                real Ey_ixp1, Ez_ixp1, Ex_iyp1, Ez_iyp1, Ex_izp1, Ey_izp1;
                if (ix == Nx - 1) {
                    int64_t j = 0 + (int64_t)Nx * (iy + (int64_t)Ny * iz);
                    Ey_ixp1 = bcval(m->bcs[1], Ey[j], Ey[i]);
                    Ez_ixp1 = bcval(m->bcs[1], Ez[j], Ez[i]);
                } else { Ey_ixp1 = Ey[i + 1]; Ez_ixp1 = Ez[i + 1]; }
                if (iy == Ny - 1) {
                    int64_t j = ix + (int64_t)Nx * (0 + (int64_t)Ny * iz);
                    Ex_iyp1 = bcval(m->bcs[3], Ex[j], Ex[i]);
                    Ez_iyp1 = bcval(m->bcs[3], Ez[j], Ez[i]);
                } else { Ex_iyp1 = Ex[i + Nx]; Ez_iyp1 = Ez[i + Nx]; }
                if (iz == Nz - 1) {
                    int64_t j = ix + (int64_t)Nx * iy;
                    Ex_izp1 = bcval(m->bcs[5], Ex[j], Ex[i]);
                    Ey_izp1 = bcval(m->bcs[5], Ey[j], Ey[i]);
                } else { Ex_izp1 = Ex[i + (int64_t)Nx * Ny]; Ey_izp1 = Ey[i + (int64_t)Nx * Ny]; }
                real dExy = (Ex_iyp1 - Ex[i]) / m->dy;
                real dExz = (Ex_izp1 - Ex[i]) / m->dz;
                real dEyx = (Ey_ixp1 - Ey[i]) / m->dx;
                real dEyz = (Ey_izp1 - Ey[i]) / m->dz;
                real dEzx = (Ez_ixp1 - Ez[i]) / m->dx;
                real dEzy = (Ez_iyp1 - Ez[i]) / m->dy;
                for (int f = 0; f < 2; f++) {       /* x layers */
                    layer_t *l = &L[f];
                    if (!l->present) continue;
                    if (f == 0 ? ix <= l->ib : ix >= l->ib) {
                        int p = f == 0 ? ix : ix - l->ib;
                        int64_t j = p + (int64_t)l->nb * (iy + (int64_t)Ny * iz);
                        l->psiE[0][j] = l->B[p] * l->psiE[0][j] + l->A[p] * dEyx;   /* psiEyx */
                        l->psiE[1][j] = l->B[p] * l->psiE[1][j] + l->A[p] * dEzx;   /* psiEzx */
                        Hy[i] += Mh * l->psiE[1][j];
                        Hz[i] -= Mh * l->psiE[0][j];
                        dEyx = dEyx / l->K[p];
                        dEzx = dEzx / l->K[p];
                    }
                }
                for (int f = 2; f < 4; f++) {       /* y layers */
                    layer_t *l = &L[f];
                    if (!l->present) continue;
                    if (f == 2 ? iy <= l->ib : iy >= l->ib) {
                        int p = f == 2 ? iy : iy - l->ib;
                        int64_t j = ix + (int64_t)Nx * (p + (int64_t)l->nb * iz);
                        l->psiE[0][j] = l->B[p] * l->psiE[0][j] + l->A[p] * dExy;   /* psiExy */
                        l->psiE[1][j] = l->B[p] * l->psiE[1][j] + l->A[p] * dEzy;   /* psiEzy */
                        Hx[i] -= Mh * l->psiE[1][j];
                        Hz[i] += Mh * l->psiE[0][j];
                        dExy = dExy / l->K[p];
                        dEzy = dEzy / l->K[p];
                    }
                }
                for (int f = 4; f < 6; f++) {       /* z layers */
                    layer_t *l = &L[f];
                    if (!l->present) continue;
                    if (f == 4 ? iz <= l->ib : iz >= l->ib) {
                        int p = f == 4 ? iz : iz - l->ib;
                        int64_t j = ix + (int64_t)Nx * (iy + (int64_t)Ny * p);
                        l->psiE[0][j] = l->B[p] * l->psiE[0][j] + l->A[p] * dExz;   /* psiExz */
                        l->psiE[1][j] = l->B[p] * l->psiE[1][j] + l->A[p] * dEyz;   /* psiEyz */
                        Hx[i] += Mh * l->psiE[1][j];
                        Hy[i] -= Mh * l->psiE[0][j];
                        dExz = dExz / l->K[p];
                        dEyz = dEyz / l->K[p];
                    }
                }
                Hx[i] -= Mh * (dEzy - dEyz);
                Hy[i] -= Mh * (dExz - dEzx);
                Hz[i] -= Mh * (dEyx - dExy);
            }
}

static void update_E_3d(ofd_model *m) {
    const int Nx = m->Nx, Ny = m->Ny, Nz = m->Nz;
    const real *Hx = m->F[MFD_HX], *Hy = m->F[MFD_HY], *Hz = m->F[MFD_HZ];
    real *Dx = m->F[MFD_DX], *Dy = m->F[MFD_DY], *Dz = m->F[MFD_DZ];
    real *Dv[3] = {Dx, Dy, Dz};
    real *Ev[3] = {m->F[MFD_EX], m->F[MFD_EY], m->F[MFD_EZ]};
    static const int cc[3] = {0, 1, 2};
    layer_t *L = m->L;
#pragma omp parallel for schedule(static)
    for (int iz = 0; iz < Nz; iz++)
        for (int iy = 0; iy < Ny; iy++)
            for (int ix = 0; ix < Nx; ix++) {
                int64_t i = ix + (int64_t)Nx * (iy + (int64_t)Ny * iz);
                int g = m->geom[i];
                real Md1, Md2;
                if (g > 0) { Md1 = m->mats[g].Md1; Md2 = m->mats[g].Md2; }
                else { Md1 = m->Md1; Md2 = m->Md2; }
                real Hy_ixm1, Hz_ixm1, Hx_iym1, Hz_iym1, Hx_izm1, Hy_izm1;
                if (ix == 0) {
                    int64_t j = (Nx - 1) + (int64_t)Nx * (iy + (int64_t)Ny * iz);
                    Hy_ixm1 = bcval(m->bcs[0], Hy[j], Hy[i]);
                    Hz_ixm1 = bcval(m->bcs[0], Hz[j], Hz[i]);
                } else { Hy_ixm1 = Hy[i - 1]; Hz_ixm1 = Hz[i - 1]; }
                if (iy == 0) {
                    int64_t j = ix + (int64_t)Nx * ((Ny - 1) + (int64_t)Ny * iz);
                    Hx_iym1 = bcval(m->bcs[2], Hx[j], Hx[i]);
                    Hz_iym1 = bcval(m->bcs[2], Hz[j], Hz[i]);
                } else { Hx_iym1 = Hx[i - Nx]; Hz_iym1 = Hz[i - Nx]; }
                if (iz == 0) {
                    int64_t j = ix + (int64_t)Nx * (iy + (int64_t)Ny * (Nz - 1));
                    Hx_izm1 = bcval(m->bcs[4], Hx[j], Hx[i]);
                    Hy_izm1 = bcval(m->bcs[4], Hy[j], Hy[i]);
                } else { Hx_izm1 = Hx[i - (int64_t)Nx * Ny]; Hy_izm1 = Hy[i - (int64_t)Nx * Ny]; }
                real dHxy = (Hx[i] - Hx_iym1) / m->dy;
                real dHxz = (Hx[i] - Hx_izm1) / m->dz;
                real dHyx = (Hy[i] - Hy_ixm1) / m->dx;
                real dHyz = (Hy[i] - Hy_izm1) / m->dz;
                real dHzx = (Hz[i] - Hz_ixm1) / m->dx;
                real dHzy = (Hz[i] - Hz_iym1) / m->dy;
                for (int f = 0; f < 2; f++) {
                    layer_t *l = &L[f];
                    if (!l->present) continue;
                    if (f == 0 ? ix <= l->ib : ix >= l->ib) {
                        int p = f == 0 ? ix : ix - l->ib;
                        int64_t j = p + (int64_t)l->nb * (iy + (int64_t)Ny * iz);
                        l->psiH[0][j] = l->B[p] * l->psiH[0][j] + l->A[p] * dHyx;   /* psiHyx */
                        l->psiH[1][j] = l->B[p] * l->psiH[1][j] + l->A[p] * dHzx;   /* psiHzx */
                        Dy[i] -= Md2 * l->psiH[1][j];
                        Dz[i] += Md2 * l->psiH[0][j];
                        dHyx = dHyx / l->K[p];
                        dHzx = dHzx / l->K[p];
                    }
                }
                for (int f = 2; f < 4; f++) {
                    layer_t *l = &L[f];
                    if (!l->present) continue;
                    if (f == 2 ? iy <= l->ib : iy >= l->ib) {
                        int p = f == 2 ? iy : iy - l->ib;
                        int64_t j = ix + (int64_t)Nx * (p + (int64_t)l->nb * iz);
                        l->psiH[0][j] = l->B[p] * l->psiH[0][j] + l->A[p] * dHxy;   /* psiHxy */
                        l->psiH[1][j] = l->B[p] * l->psiH[1][j] + l->A[p] * dHzy;   /* psiHzy */
                        Dx[i] += Md2 * l->psiH[1][j];
                        Dz[i] -= Md2 * l->psiH[0][j];
                        dHxy = dHxy / l->K[p];
                        dHzy = dHzy / l->K[p];
                    }
                }
                for (int f = 4; f < 6; f++) {
                    layer_t *l = &L[f];
                    if (!l->present) continue;
                    if (f == 4 ? iz <= l->ib : iz >= l->ib) {
                        int p = f == 4 ? iz : iz - l->ib;
                        int64_t j = ix + (int64_t)Nx * (iy + (int64_t)Ny * p);
                        l->psiH[0][j] = l->B[p] * l->psiH[0][j] + l->A[p] * dHxz;   /* psiHxz */
                        l->psiH[1][j] = l->B[p] * l->psiH[1][j] + l->A[p] * dHyz;   /* psiHyz */
                        Dx[i] -= Md2 * l->psiH[1][j];
                        Dy[i] += Md2 * l->psiH[0][j];
                        dHxz = dHxz / l->K[p];
                        dHyz = dHyz / l->K[p];
                    }
                }
                Dx[i] = Md1 * Dx[i] + Md2 * (dHzy - dHyz);
                Dy[i] = Md1 * Dy[i] + Md2 * (dHxz - dHzx);
                Dz[i] = Md1 * Dz[i] + Md2 * (dHyx - dHxy);
                material_e(m, i, g, 3, cc, Dv, Ev);
            }
}

/* ---------------- 2D (x,z): src/models.jl:450-726 ----------------------------------- */
static void update_H_2d(ofd_model *m) {
    const int Nx = m->Nx, Nz = m->Nz;
    real *Hy = m->F[MFD_HY];
    const real *Ex = m->F[MFD_EX], *Ez = m->F[MFD_EZ];
    layer_t *L = m->L;
#pragma omp parallel for schedule(static)
    for (int iz = 0; iz < Nz; iz++)
        for (int ix = 0; ix < Nx; ix++) {
            int64_t i = ix + (int64_t)Nx * iz;
            int g = m->geom[i];
            real Mh = g > 0 ? m->mats[g].Mh : m->Mh;
            real Ez_ixp1, Ex_izp1;
            if (ix == Nx - 1) Ez_ixp1 = bcval(m->bcs[1], Ez[(int64_t)Nx * iz], Ez[i]);
            else Ez_ixp1 = Ez[i + 1];
            if (iz == Nz - 1) Ex_izp1 = bcval(m->bcs[5], Ex[ix], Ex[i]);
            else Ex_izp1 = Ex[i + Nx];
            real dExz = (Ex_izp1 - Ex[i]) / m->dz;
            real dEzx = (Ez_ixp1 - Ez[i]) / m->dx;
            for (int f = 0; f < 2; f++) {
                layer_t *l = &L[f];
                if (!l->present) continue;
                if (f == 0 ? ix <= l->ib : ix >= l->ib) {
                    int p = f == 0 ? ix : ix - l->ib;
                    int64_t j = p + (int64_t)l->nb * iz;
                    l->psiE[1][j] = l->B[p] * l->psiE[1][j] + l->A[p] * dEzx;   /* psiEzx */
                    Hy[i] += Mh * l->psiE[1][j];
                    dEzx = dEzx / l->K[p];
                }
            }
            for (int f = 4; f < 6; f++) {
                layer_t *l = &L[f];
                if (!l->present) continue;
                if (f == 4 ? iz <= l->ib : iz >= l->ib) {
                    int p = f == 4 ? iz : iz - l->ib;
                    int64_t j = ix + (int64_t)Nx * p;
                    l->psiE[0][j] = l->B[p] * l->psiE[0][j] + l->A[p] * dExz;   /* psiExz */
                    Hy[i] -= Mh * l->psiE[0][j];
                    dExz = dExz / l->K[p];
                }
            }
            Hy[i] -= Mh * (dExz - dEzx);
        }
}

static void update_E_2d(ofd_model *m) {
    const int Nx = m->Nx, Nz = m->Nz;
    const real *Hy = m->F[MFD_HY];
    real *Dx = m->F[MFD_DX], *Dz = m->F[MFD_DZ];
    real *Dv[2] = {Dx, Dz};
    real *Ev[2] = {m->F[MFD_EX], m->F[MFD_EZ]};
    static const int cc[2] = {0, 2};
    layer_t *L = m->L;
#pragma omp parallel for schedule(static)
    for (int iz = 0; iz < Nz; iz++)
        for (int ix = 0; ix < Nx; ix++) {
            int64_t i = ix + (int64_t)Nx * iz;
            int g = m->geom[i];
            real Md1, Md2;
            if (g > 0) { Md1 = m->mats[g].Md1; Md2 = m->mats[g].Md2; }
            else { Md1 = m->Md1; Md2 = m->Md2; }
            real Hy_ixm1, Hy_izm1;
            if (ix == 0) Hy_ixm1 = bcval(m->bcs[0], Hy[(Nx - 1) + (int64_t)Nx * iz], Hy[i]);
            else Hy_ixm1 = Hy[i - 1];
            if (iz == 0) Hy_izm1 = bcval(m->bcs[4], Hy[ix + (int64_t)Nx * (Nz - 1)], Hy[i]);
            else Hy_izm1 = Hy[i - Nx];
            real dHyx = (Hy[i] - Hy_ixm1) / m->dx;
            real dHyz = (Hy[i] - Hy_izm1) / m->dz;
            for (int f = 0; f < 2; f++) {
                layer_t *l = &L[f];
                if (!l->present) continue;
                if (f == 0 ? ix <= l->ib : ix >= l->ib) {
                    int p = f == 0 ? ix : ix - l->ib;
                    int64_t j = p + (int64_t)l->nb * iz;
                    l->psiH[0][j] = l->B[p] * l->psiH[0][j] + l->A[p] * dHyx;   /* psiHyx */
                    Dz[i] += Md2 * l->psiH[0][j];
                    dHyx = dHyx / l->K[p];
                }
            }
            for (int f = 4; f < 6; f++) {
                layer_t *l = &L[f];
                if (!l->present) continue;
                if (f == 4 ? iz <= l->ib : iz >= l->ib) {
                    int p = f == 4 ? iz : iz - l->ib;
                    int64_t j = ix + (int64_t)Nx * p;
                    l->psiH[1][j] = l->B[p] * l->psiH[1][j] + l->A[p] * dHyz;   /* psiHyz */
                    Dx[i] -= Md2 * l->psiH[1][j];
                    dHyz = dHyz / l->K[p];
                }
            }
            Dx[i] = Md1 * Dx[i] + Md2 * (0 - dHyz);
            Dz[i] = Md1 * Dz[i] + Md2 * (dHyx - 0);
            material_e(m, i, g, 2, cc, Dv, Ev);
        }
}

/* ---------------- 1D (z): src/models.jl:248-444 ------------------------------------- */
static void update_H_1d(ofd_model *m) {
    const int Nz = m->Nz;
    real *Hy = m->F[MFD_HY];
    const real *Ex = m->F[MFD_EX];
    layer_t *L = m->L;
    for (int iz = 0; iz < Nz; iz++) {
        int g = m->geom[iz];
        real Mh = g > 0 ? m->mats[g].Mh : m->Mh;
        real Ex_izp1 = (iz == Nz - 1) ? bcval(m->bcs[5], Ex[0], Ex[iz]) : Ex[iz + 1];
        real dExz = (Ex_izp1 - Ex[iz]) / m->dz;
        for (int f = 4; f < 6; f++) {
            layer_t *l = &L[f];
            if (!l->present) continue;
            if (f == 4 ? iz <= l->ib : iz >= l->ib) {
                int p = f == 4 ? iz : iz - l->ib;
                l->psiE[0][p] = l->B[p] * l->psiE[0][p] + l->A[p] * dExz;     /* psiExz */
                Hy[iz] -= Mh * l->psiE[0][p];
                dExz = dExz / l->K[p];
            }
        }
        Hy[iz] -= Mh * (0 + dExz);
    }
}

static void update_E_1d(ofd_model *m) {
    const int Nz = m->Nz;
    const real *Hy = m->F[MFD_HY];
    real *Dx = m->F[MFD_DX];
    real *Dv[1] = {Dx};
    real *Ev[1] = {m->F[MFD_EX]};
    static const int cc[1] = {0};
    layer_t *L = m->L;
    for (int iz = 0; iz < Nz; iz++) {
        int g = m->geom[iz];
        real Md1, Md2;
        if (g > 0) { Md1 = m->mats[g].Md1; Md2 = m->mats[g].Md2; }
        else { Md1 = m->Md1; Md2 = m->Md2; }
        real Hy_izm1 = (iz == 0) ? bcval(m->bcs[4], Hy[Nz - 1], Hy[iz]) : Hy[iz - 1];
        real dHyz = (Hy[iz] - Hy_izm1) / m->dz;
        for (int f = 4; f < 6; f++) {
            layer_t *l = &L[f];
            if (!l->present) continue;
            if (f == 4 ? iz <= l->ib : iz >= l->ib) {
                int p = f == 4 ? iz : iz - l->ib;
                l->psiH[1][p] = l->B[p] * l->psiH[1][p] + l->A[p] * dHyz;     /* psiHyz */
                Dx[iz] -= Md2 * l->psiH[1][p];
                dHyz = dHyz / l->K[p];
            }
        }
        Dx[iz] = Md1 * Dx[iz] + Md2 * (0 - dHyz);
        material_e(m, iz, g, 1, cc, Dv, Ev);
    }
}

/* ---------------- TFSF auxiliary 1D run ---------------------------------------------- */
/* The x/y-uniform 3D run of the record pass reduces to these 1D recursions because the
   transverse differences are exactly zero (src/models.jl:735-1162 with d/dx = d/dy = 0):
   x-polarised pair (Hy, Dx, Ex) follows update_H/E_kernel! 1D (src/models.jl:248-444);
   the y-polarised pair (Hx, Dy, Ey) follows the same 3D statements with signs
   Hx -= Mh*(0 - dEyz), Hx += Mh*psiEyz, Dy = Md1*Dy + Md2*(dHxz - 0), Dy += Md2*psiHxz. */
static void tfsf_aux_step(ofd_model *m, tfsf_t *T, int it1) {
    const int Nz = m->Nz;
    const real Mh = m->Mh, Me = m->Me, Md1 = m->Md1, Md2 = m->Md2;
    for (int iz = 0; iz < Nz; iz++) {
        real Ex1 = (iz == Nz - 1) ? bcval(m->bcs[5], T->aEx[0], T->aEx[iz]) : T->aEx[iz + 1];
        real Ey1 = (iz == Nz - 1) ? bcval(m->bcs[5], T->aEy[0], T->aEy[iz]) : T->aEy[iz + 1];
        real dExz = (Ex1 - T->aEx[iz]) / m->dz;
        real dEyz = (Ey1 - T->aEy[iz]) / m->dz;
        for (int f = 4; f < 6; f++) {
            layer_t *l = &m->L[f];
            if (!l->present) continue;
            if (f == 4 ? iz <= l->ib : iz >= l->ib) {
                int p = f == 4 ? iz : iz - l->ib;
                real *pEx = T->apsiE[f - 4][0], *pEy = T->apsiE[f - 4][1];
                pEx[p] = l->B[p] * pEx[p] + l->A[p] * dExz;
                pEy[p] = l->B[p] * pEy[p] + l->A[p] * dEyz;
                T->aHx[iz] += Mh * pEy[p];
                T->aHy[iz] -= Mh * pEx[p];
                dExz = dExz / l->K[p];
                dEyz = dEyz / l->K[p];
            }
        }
        T->aHx[iz] -= Mh * (0 - dEyz);
        T->aHy[iz] -= Mh * (dExz - 0);
    }
    for (int iz = 0; iz < Nz; iz++) {
        real Hx0 = (iz == 0) ? bcval(m->bcs[4], T->aHx[Nz - 1], T->aHx[iz]) : T->aHx[iz - 1];
        real Hy0 = (iz == 0) ? bcval(m->bcs[4], T->aHy[Nz - 1], T->aHy[iz]) : T->aHy[iz - 1];
        real dHxz = (T->aHx[iz] - Hx0) / m->dz;
        real dHyz = (T->aHy[iz] - Hy0) / m->dz;
        for (int f = 4; f < 6; f++) {
            layer_t *l = &m->L[f];
            if (!l->present) continue;
            if (f == 4 ? iz <= l->ib : iz >= l->ib) {
                int p = f == 4 ? iz : iz - l->ib;
                real *pHx = T->apsiH[f - 4][0], *pHy = T->apsiH[f - 4][1];
                pHx[p] = l->B[p] * pHx[p] + l->A[p] * dHxz;
                pHy[p] = l->B[p] * pHy[p] + l->A[p] * dHyz;
                T->aDx[iz] -= Md2 * pHy[p];
                T->aDy[iz] += Md2 * pHx[p];
                dHxz = dHxz / l->K[p];
                dHyz = dHyz / l->K[p];
            }
        }
        T->aDx[iz] = Md1 * T->aDx[iz] + Md2 * (0 - dHyz);
        T->aDy[iz] = Md1 * T->aDy[iz] + Md2 * (dHxz - 0);
        T->aEx[iz] = T->aDx[iz] * Me;
        T->aEy[iz] = T->aDy[iz] * Me;
    }
    for (int k = 0; k < T->aux_n; k++) {
        int iz = T->aux_iz[k];
        real v = T->aux_coeff[k] * T->aux_wave[(size_t)k * m->Nt + (it1 - 1)];
        real *dst = NULL;
        switch (T->aux_comp[k]) {
            case MFD_EX: dst = T->aEx; break;
            case MFD_EY: dst = T->aEy; break;
            case MFD_DX: dst = T->aDx; break;
            case MFD_DY: dst = T->aDy; break;
            case MFD_HX: dst = T->aHx; break;
            case MFD_HY: dst = T->aHy; break;
        }
        if (dst) dst[iz] = dst[iz] + v;
    }
}

/* incident value lookup: table replay or auxiliary 1D run (uniform in x, y)           */
typedef struct {
    ofd_model *m;
    tfsf_t *T;
    int it0;   /* 0-based step index */
} inc_ctx;

/* 3D table k, element (a, b) of the face (column-major, time slowest)                 */
static real inc3(inc_ctx *c, int k, int a, int b, int zidx) {
    tfsf_t *T = c->T;
    if (T->mode == MFD_TFSF_AUX1D) {
        /* component of table k: face order {Hy,Hz,Ey,Ez},{Hx,Hz,Ex,Ez},{Hx,Hy,Ex,Ey}    */
        static const int comp[24] = {
            MFD_HY, MFD_HZ, MFD_EY, MFD_EZ, MFD_HY, MFD_HZ, MFD_EY, MFD_EZ,
            MFD_HX, MFD_HZ, MFD_EX, MFD_EZ, MFD_HX, MFD_HZ, MFD_EX, MFD_EZ,
            MFD_HX, MFD_HY, MFD_EX, MFD_EY, MFD_HX, MFD_HY, MFD_EX, MFD_EY};
        switch (comp[k]) {
            case MFD_HX: return T->aHx[zidx];
            case MFD_HY: return T->aHy[zidx];
            case MFD_EX: return T->aEx[zidx];
            case MFD_EY: return T->aEy[zidx];
            default: return 0;
        }
    }
    const int *b_ = T->box;
    int Nxi = b_[1] - b_[0], Nyi = b_[3] - b_[2], Nzi = b_[5] - b_[4];
    int face = k / 4, j = k % 4;
    int na, nbb;
    if (face < 2) { na = (j == 0 || j == 3) ? Nyi + 1 : Nyi; nbb = (j == 0 || j == 3) ? Nzi : Nzi + 1; }
    else if (face < 4) { na = (j == 0 || j == 3) ? Nxi + 1 : Nxi; nbb = (j == 0 || j == 3) ? Nzi : Nzi + 1; }
    else { na = (j == 0 || j == 3) ? Nxi + 1 : Nxi; nbb = (j == 0 || j == 3) ? Nyi : Nyi + 1; }
    return T->tab[k][a + (size_t)na * (b + (size_t)nbb * c->it0)];
}


/* F[c] element at 0-based (ix, iy, iz) */
#define AT(c, ix, iy, iz) m->F[c][(ix) + (int64_t)m->Nx * ((iy) + (int64_t)m->Ny * (iz))]

/* One broadcast statement of add_source!(::TFSFSource3D) (src/sources.jl:728-773):
   F[comp][face] += coef * inc_k[:, :, it].  normal 0: x = n0, (a, b) = (y, z);
   normal 1: y = n0, (a, b) = (x, z); normal 2: z = n0, (a, b) = (x, y).
   zrec is the z plane the table was recorded at when the normal is z.               */
static void tf3(ofd_model *m, inc_ctx *c, int comp, int normal, int n0, int a0, int a1,
                int b0, int b1, real coef, int k, int zrec) {
    for (int b = b0; b <= b1; b++)
        for (int a = a0; a <= a1; a++) {
            int ix, iy, iz;
            if (normal == 0) { ix = n0; iy = a; iz = b; }
            else if (normal == 1) { ix = a; iy = n0; iz = b; }
            else { ix = a; iy = b; iz = n0; }
            int zidx = normal == 2 ? zrec : iz;
            real v = inc3(c, k, a - a0, b - b0, zidx);
            AT(comp, ix, iy, iz) = AT(comp, ix, iy, iz) + coef * v;
        }
}

static void tfsf_apply_3d(ofd_model *m, tfsf_t *T, int it0) {
    inc_ctx c = {m, T, it0};
    const int X1 = T->box[0] - 1, X2 = T->box[1] - 1, Y1 = T->box[2] - 1, Y2 = T->box[3] - 1;
    const int Z1 = T->box[4] - 1, Z2 = T->box[5] - 1;
    const double Cm = m->dt64 / MFD_MU0, Ce = m->dt64 / MFD_EPS0;
    const real cmx = (real)(Cm / m->dx), cmy = (real)(Cm / m->dy), cmz = (real)(Cm / m->dz);
    const real cex = (real)(Ce / m->dx), cey = (real)(Ce / m->dy), cez = (real)(Ce / m->dz);
    const real cdx = (real)(MFD_EPS0 * Ce / m->dx), cdy = (real)(MFD_EPS0 * Ce / m->dy);
    const real cdz = (real)(MFD_EPS0 * Ce / m->dz);
    /* 1yz: x = x1 */
    tf3(m, &c, MFD_HY, 0, X1 - 1, Y1, Y2, Z1, Z2 - 1, -cmx, 3, 0);
    tf3(m, &c, MFD_HZ, 0, X1 - 1, Y1, Y2 - 1, Z1, Z2, cmx, 2, 0);
    tf3(m, &c, MFD_EY, 0, X1, Y1, Y2 - 1, Z1, Z2, cex, 1, 0);
    tf3(m, &c, MFD_EZ, 0, X1, Y1, Y2, Z1, Z2 - 1, -cex, 0, 0);
    tf3(m, &c, MFD_DY, 0, X1, Y1, Y2 - 1, Z1, Z2, cdx, 1, 0);
    tf3(m, &c, MFD_DZ, 0, X1, Y1, Y2, Z1, Z2 - 1, -cdx, 0, 0);
    /* 2yz: x = x2 */
    tf3(m, &c, MFD_HY, 0, X2, Y1, Y2, Z1, Z2 - 1, cmx, 7, 0);
    tf3(m, &c, MFD_HZ, 0, X2, Y1, Y2 - 1, Z1, Z2, -cmx, 6, 0);
    tf3(m, &c, MFD_EY, 0, X2, Y1, Y2 - 1, Z1, Z2, -cex, 5, 0);
    tf3(m, &c, MFD_EZ, 0, X2, Y1, Y2, Z1, Z2 - 1, cex, 4, 0);
    tf3(m, &c, MFD_DY, 0, X2, Y1, Y2 - 1, Z1, Z2, -cdx, 5, 0);
    tf3(m, &c, MFD_DZ, 0, X2, Y1, Y2, Z1, Z2 - 1, cdx, 4, 0);
    /* x1z: y = y1 */
    tf3(m, &c, MFD_HX, 1, Y1 - 1, X1, X2, Z1, Z2 - 1, cmy, 11, 0);
    tf3(m, &c, MFD_HZ, 1, Y1 - 1, X1, X2 - 1, Z1, Z2, -cmy, 10, 0);
    tf3(m, &c, MFD_EX, 1, Y1, X1, X2 - 1, Z1, Z2, -cey, 9, 0);
    tf3(m, &c, MFD_EZ, 1, Y1, X1, X2, Z1, Z2 - 1, cey, 8, 0);
    tf3(m, &c, MFD_DX, 1, Y1, X1, X2 - 1, Z1, Z2, -cdy, 9, 0);
    tf3(m, &c, MFD_DZ, 1, Y1, X1, X2, Z1, Z2 - 1, cdy, 8, 0);
    /* x2z: y = y2 */
    tf3(m, &c, MFD_HX, 1, Y2, X1, X2, Z1, Z2 - 1, -cmy, 15, 0);
    tf3(m, &c, MFD_HZ, 1, Y2, X1, X2 - 1, Z1, Z2, cmy, 14, 0);
    tf3(m, &c, MFD_EX, 1, Y2, X1, X2 - 1, Z1, Z2, cey, 13, 0);
    tf3(m, &c, MFD_EZ, 1, Y2, X1, X2, Z1, Z2 - 1, -cey, 12, 0);
    tf3(m, &c, MFD_DX, 1, Y2, X1, X2 - 1, Z1, Z2, cdy, 13, 0);
    tf3(m, &c, MFD_DZ, 1, Y2, X1, X2, Z1, Z2 - 1, -cdy, 12, 0);
    /* xy1: z = z1 (H tables recorded at z1-1, E tables at z1) */
    tf3(m, &c, MFD_HX, 2, Z1 - 1, X1, X2, Y1, Y2 - 1, -cmz, 19, Z1);
    tf3(m, &c, MFD_HY, 2, Z1 - 1, X1, X2 - 1, Y1, Y2, cmz, 18, Z1);
    tf3(m, &c, MFD_EX, 2, Z1, X1, X2 - 1, Y1, Y2, cez, 17, Z1 - 1);
    tf3(m, &c, MFD_EY, 2, Z1, X1, X2, Y1, Y2 - 1, -cez, 16, Z1 - 1);
    tf3(m, &c, MFD_DX, 2, Z1, X1, X2 - 1, Y1, Y2, cdz, 17, Z1 - 1);
    tf3(m, &c, MFD_DY, 2, Z1, X1, X2, Y1, Y2 - 1, -cdz, 16, Z1 - 1);
    /* xy2: z = z2 (all tables recorded at z2) */
    tf3(m, &c, MFD_HX, 2, Z2, X1, X2, Y1, Y2 - 1, cmz, 23, Z2);
    tf3(m, &c, MFD_HY, 2, Z2, X1, X2 - 1, Y1, Y2, -cmz, 22, Z2);
    tf3(m, &c, MFD_EX, 2, Z2, X1, X2 - 1, Y1, Y2, -cez, 21, Z2);
    tf3(m, &c, MFD_EY, 2, Z2, X1, X2, Y1, Y2 - 1, cez, 20, Z2);
    tf3(m, &c, MFD_DX, 2, Z2, X1, X2 - 1, Y1, Y2, -cdz, 21, Z2);
    tf3(m, &c, MFD_DY, 2, Z2, X1, X2, Y1, Y2 - 1, cdz, 20, Z2);
}

/* add_source!(::TFSFSource2D) (src/sources.jl:469-497).  Tables: incHy_1z, incEz_1z,
   incHy_2z, incEz_2z ([Nzi, Nt]), incHy_x1, incEx_x1, incHy_x2, incEx_x2 ([Nxi, Nt]). */
static real inc2(tfsf_t *T, const ofd_model *m, int k, int a, int it0, int zrec) {
    if (T->mode == MFD_TFSF_AUX1D) {
        static const int comp[8] = {MFD_HY, MFD_EZ, MFD_HY, MFD_EZ, MFD_HY, MFD_EX, MFD_HY, MFD_EX};
        if (comp[k] == MFD_HY) return T->aHy[zrec];
        if (comp[k] == MFD_EX) return T->aEx[zrec];
        return 0;
    }
    int n = k < 4 ? T->box[5] - T->box[4] : T->box[1] - T->box[0];
    (void)m;
    return T->tab[k][a + (size_t)n * it0];
}

static void tfsf_apply_2d(ofd_model *m, tfsf_t *T, int it0) {
    const int X1 = T->box[0] - 1, X2 = T->box[1] - 1, Z1 = T->box[4] - 1, Z2 = T->box[5] - 1;
    const double dt = m->dt64;
    const real hmx = (real)(dt / (MFD_MU0 * m->dx)), hex = (real)(dt / (MFD_EPS0 * m->dx));
    const real hdx = (real)(dt / m->dx);
    const real hmz = (real)(dt / (MFD_MU0 * m->dz)), hez = (real)(dt / (MFD_EPS0 * m->dz));
    const real hdz = (real)(dt / m->dz);
    for (int iz = Z1; iz <= Z2 - 1; iz++) {
        int a = iz - Z1;
        AT(MFD_HY, X1 - 1, 0, iz) -= hmx * inc2(T, m, 1, a, it0, iz);
        AT(MFD_EZ, X1, 0, iz) -= hex * inc2(T, m, 0, a, it0, iz);
        AT(MFD_DZ, X1, 0, iz) -= hdx * inc2(T, m, 0, a, it0, iz);
    }
    for (int iz = Z1; iz <= Z2 - 1; iz++) {
        int a = iz - Z1;
        AT(MFD_HY, X2, 0, iz) += hmx * inc2(T, m, 3, a, it0, iz);
        AT(MFD_EZ, X2, 0, iz) += hex * inc2(T, m, 2, a, it0, iz);
        AT(MFD_DZ, X2, 0, iz) += hdx * inc2(T, m, 2, a, it0, iz);
    }
    for (int ix = X1; ix <= X2 - 1; ix++) {
        int a = ix - X1;
        AT(MFD_HY, ix, 0, Z1 - 1) += hmz * inc2(T, m, 5, a, it0, Z1);
        AT(MFD_EX, ix, 0, Z1) += hez * inc2(T, m, 4, a, it0, Z1 - 1);
        AT(MFD_DX, ix, 0, Z1) += hdz * inc2(T, m, 4, a, it0, Z1 - 1);
    }
    for (int ix = X1; ix <= X2 - 1; ix++) {
        int a = ix - X1;
        AT(MFD_HY, ix, 0, Z2) -= hmz * inc2(T, m, 7, a, it0, Z2);
        AT(MFD_EX, ix, 0, Z2) -= hez * inc2(T, m, 6, a, it0, Z2);
        AT(MFD_DX, ix, 0, Z2) -= hdz * inc2(T, m, 6, a, it0, Z2);
    }
}

/* add_source!(::TFSFSource1D) (src/sources.jl:320-346).  Tables incHy_1, incEx_1,
   incHy_2, incEx_2 ([Nt]).                                                             */
static void tfsf_apply_1d(ofd_model *m, tfsf_t *T, int it0) {
    const int Z1 = T->box[4] - 1, Z2 = T->box[5] - 1;
    real iHy1, iEx1, iHy2, iEx2;
    if (T->mode == MFD_TFSF_AUX1D) {
        iHy1 = T->aHy[Z1 - 1]; iEx1 = T->aEx[Z1]; iHy2 = T->aHy[Z2]; iEx2 = T->aEx[Z2];
    } else {
        iHy1 = T->tab[0][it0]; iEx1 = T->tab[1][it0]; iHy2 = T->tab[2][it0]; iEx2 = T->tab[3][it0];
    }
    const double dt = m->dt64;
    const real hm = (real)(dt / (MFD_MU0 * m->dz)), he = (real)(dt / (MFD_EPS0 * m->dz));
    const real hd = (real)(dt / m->dz);
    real *Hy = m->F[MFD_HY], *Ex = m->F[MFD_EX], *Dx = m->F[MFD_DX];
    Hy[Z1 - 1] += hm * iEx1;
    Ex[Z1] += he * iHy1;
    Dx[Z1] += hd * iHy1;
    Hy[Z2] -= hm * iEx2;
    Ex[Z2] -= he * iHy2;
    Dx[Z2] -= hd * iHy2;
}

/* add_source_kernel! soft/hard (src/sources.jl:77-119,190-232)                          */
static void point_source(ofd_model *m, src_t *S, int it0) {
    real *F = m->F[S->comp];
    const real *w = S->wave + (size_t)it0 * S->n;
    for (int k = 0; k < S->n; k++) {
        int64_t i = S->cells[k];
        if (S->kind == MFD_SRC_SOFT) F[i] = F[i] + S->coeff * w[k];
        else F[i] = w[k];
    }
}

/* update_monitor! (src/monitors.jl:62-66,110-116,164-176,350-359), DFT (src/dft.jl:36-51) */
static void monitors(ofd_model *m, int it0) {
    const int nc = ncomp_mon(m);
    const double t = (double)it0 * m->dt64;
    for (int q = 0; q < m->nmon; q++) {
        mon_t *M = &m->mon[q];
        for (int j = 0; j < nc; j++) {
            int comp = mon_comp(m, j);
            const real *F = m->F[comp];
            double *D = M->data[comp];
            if (!M->spectral) {
                if (M->issum) {
                    double s = 0;
                    for (int k = 0; k < M->n; k++) s += (double)F[M->cells[k]];
                    D[it0] = s;
                } else {
                    for (int k = 0; k < M->n; k++) D[k + (size_t)M->n * it0] = (double)F[M->cells[k]];
                }
            } else {
                for (int iw = 0; iw < M->nw; iw++) {
                    double complex e = cexp(-I * M->w[iw] * (t - M->tshift));
                    if (M->issum) {
                        double s = 0;
                        for (int k = 0; k < M->n; k++) s += (double)F[M->cells[k]];
                        double complex v = 2 * (s * e) * m->dt64;
                        D[2 * iw] += creal(v);
                        D[2 * iw + 1] += cimag(v);
                    } else {
                        double complex tmp = 2 * e * m->dt64;
                        for (int k = 0; k < M->n; k++) {
                            double complex v = (double)F[M->cells[k]] * tmp;
                            size_t o = 2 * (k + (size_t)M->n * iw);
                            D[o] += creal(v);
                            D[o + 1] += cimag(v);
                        }
                    }
                }
            }
        }
    }
}

/* update_integral_variables! (src/outputs.jl:182-191,318-327,475-485)                  */
static void integrals(ofd_model *m) {
    const int64_t N = m->N;
    const real dt = m->dt;
    if (m->ndims == 3) {
        const real *Hx = m->F[MFD_HX], *Hy = m->F[MFD_HY], *Hz = m->F[MFD_HZ];
        const real *Ex = m->F[MFD_EX], *Ey = m->F[MFD_EY], *Ez = m->F[MFD_EZ];
#pragma omp parallel for schedule(static)
        for (int64_t i = 0; i < N; i++) {
            if (m->Sa) {
                real a = Ey[i] * Hz[i] - Ez[i] * Hy[i];
                real b = Ez[i] * Hx[i] - Ex[i] * Hz[i];
                real c = Ex[i] * Hy[i] - Ey[i] * Hx[i];
                m->Sa[i] += (real)sqrt((double)(a * a + b * b + c * c)) * dt;
            }
            if (m->E2) m->E2[i] += (Ex[i] * Ex[i] + Ey[i] * Ey[i] + Ez[i] * Ez[i]) * dt;
        }
    } else if (m->ndims == 2) {
        const real *Hy = m->F[MFD_HY], *Ex = m->F[MFD_EX], *Ez = m->F[MFD_EZ];
        for (int64_t i = 0; i < N; i++) {
            if (m->Sa) {
                real a = -Ez[i] * Hy[i], b = Ex[i] * Hy[i];
                m->Sa[i] += (real)sqrt((double)(a * a + b * b)) * dt;
            }
            if (m->E2) m->E2[i] += (Ex[i] * Ex[i] + Ez[i] * Ez[i]) * dt;
        }
    } else {
        const real *Hy = m->F[MFD_HY], *Ex = m->F[MFD_EX];
        for (int64_t i = 0; i < N; i++) {
            if (m->Sa) {
                real a = Ex[i] * Hy[i];
                m->Sa[i] += (real)sqrt((double)(a * a)) * dt;
            }
            if (m->E2) m->E2[i] += Ex[i] * Ex[i] * dt;
        }
    }
}

/* one iteration of the solve! loop (src/models.jl:199-234)                              */
int32_t ofd_model_step(ofd_model *m, int32_t nsteps) {
    for (int s = 0; s < nsteps; s++) {
        if (m->it >= m->Nt) { set_err("model already ran all Nt steps"); return MFD_ERR_STATE; }
        const int it0 = m->it;
        if (m->ndims == 3) { update_H_3d(m); update_E_3d(m); }
        else if (m->ndims == 2) { update_H_2d(m); update_E_2d(m); }
        else { update_H_1d(m); update_E_1d(m); }
        for (int k = 0; k < m->nsrc; k++) {
            if (m->src_kind[k] == 0) {
                point_source(m, &m->src[k], it0);
            } else {
                tfsf_t *T = &m->tfsf[k];
                if (T->mode == MFD_TFSF_AUX1D) tfsf_aux_step(m, T, it0 + 1);
                if (m->ndims == 3) tfsf_apply_3d(m, T, it0);
                else if (m->ndims == 2) tfsf_apply_2d(m, T, it0);
                else tfsf_apply_1d(m, T, it0);
            }
        }
        m->it = it0 + 1;
        const real *Ex = m->F[MFD_EX];
        int nan = 0, inf = 0;
        for (int64_t i = 0; i < m->N; i++) {
            if (isnan(Ex[i])) nan = 1;
            else if (isinf(Ex[i])) inf = 1;
        }
        if (nan) { set_err("Something went wrong. I found NaN field values."); m->status = MFD_ERR_NAN; return MFD_ERR_NAN; }
        if (inf) { set_err("Something went wrong. I found Inf field values."); m->status = MFD_ERR_INF; return MFD_ERR_INF; }
        monitors(m, it0);
        if (m->int_flags) integrals(m);
    }
    return MFD_OK;
}

int32_t ofd_model_get_it(ofd_model *m, int32_t *it) { *it = m->it; return MFD_OK; }
int32_t ofd_model_get_status(ofd_model *m, int32_t *status) { *status = m->status; return MFD_OK; }

int32_t ofd_model_get_field(ofd_model *m, int32_t comp, double *out) {
    if (comp < 0 || comp > 8 || !m->F[comp]) { set_err("bad component"); return MFD_ERR_ARG; }
    for (int64_t i = 0; i < m->N; i++) out[i] = (double)m->F[comp][i];
    return MFD_OK;
}

int32_t ofd_model_set_field(ofd_model *m, int32_t comp, const double *in) {
    if (comp < 0 || comp > 8 || !m->F[comp]) { set_err("bad component"); return MFD_ERR_ARG; }
    for (int64_t i = 0; i < m->N; i++) m->F[comp][i] = (real)in[i];
    return MFD_OK;
}

int32_t ofd_model_get_monitor(ofd_model *m, int32_t imon, int32_t comp, double *out) {
    if (imon < 0 || imon >= m->nmon || comp < 0 || comp > 8 || !m->mon[imon].data[comp]) {
        set_err("bad monitor or component");
        return MFD_ERR_ARG;
    }
    mon_t *M = &m->mon[imon];
    size_t len = M->spectral ? 2 * (M->issum ? (size_t)M->nw : (size_t)M->n * M->nw)
                             : (M->issum ? (size_t)m->Nt : (size_t)M->n * m->Nt);
    memcpy(out, M->data[comp], len * sizeof(double));
    return MFD_OK;
}

int32_t ofd_model_get_output(ofd_model *m, int32_t which, double *out) {
    const real *src = NULL;
    if (which == MFD_OUT_SA) src = m->Sa;
    else if (which == MFD_OUT_E2) src = m->E2;
    else if (which == MFD_OUT_JE) src = m->JE;
    else if (which == MFD_OUT_RHO) {
        for (int64_t i = 0; i < m->N; i++) out[i] = 0;
        for (int g = 1; g <= m->nmat; g++)
            if (m->mats[g].isplasma)
                for (int64_t i = 0; i < m->N; i++)
                    out[i] += (double)m->mats[g].rho[i] * (double)m->mats[g].rho0;
        return MFD_OK;
    }
    if (!src) { set_err("output not enabled"); return MFD_ERR_STATE; }
    for (int64_t i = 0; i < m->N; i++) out[i] = (double)src[i];
    return MFD_OK;
}

int32_t ofd_model_destroy(ofd_model *m) {
    if (!m) return MFD_OK;
    for (int c = 0; c < 9; c++) free(m->F[c]);
    free(m->geom);
    for (int f = 0; f < 6; f++) {
        layer_t *L = &m->L[f];
        free(L->K); free(L->A); free(L->B);
        for (int k = 0; k < 2; k++) { free(L->psiE[k]); free(L->psiH[k]); }
    }
    for (int g = 1; g <= m->nmat; g++) {
        mat_t *t = &m->mats[g];
        free(t->Aq); free(t->Bq); free(t->Cq); free(t->rate_x); free(t->rate_y);
        free(t->rho); free(t->drho);
        for (int c = 0; c < 3; c++) {
            free(t->P[c]); free(t->P1[c]); free(t->P2[c]);
            free(t->Pp[c]); free(t->Pp1[c]); free(t->Pp2[c]); free(t->Pa[c]);
        }
    }
    for (int k = 0; k < m->nsrc; k++) {
        if (m->src_kind[k] == 0) { free(m->src[k].cells); free(m->src[k].wave); }
        else {
            tfsf_t *T = &m->tfsf[k];
            for (int j = 0; j < T->ntab; j++) free(T->tab[j]);
            free(T->aux_comp); free(T->aux_iz); free(T->aux_coeff); free(T->aux_wave);
            free(T->aHx); free(T->aHy); free(T->aDx); free(T->aDy); free(T->aEx); free(T->aEy);
            for (int f = 0; f < 2; f++)
                for (int p = 0; p < 2; p++) { free(T->apsiE[f][p]); free(T->apsiH[f][p]); }
        }
    }
    for (int q = 0; q < m->nmon; q++) {
        free(m->mon[q].cells); free(m->mon[q].w);
        for (int c = 0; c < 9; c++) free(m->mon[q].data[c]);
    }
    free(m->Sa); free(m->E2); free(m->JE);
    free(m);
    return MFD_OK;
}
