/*
 * maxwell_fdtd.h -- C ABI of the MI355X-native MaxwellFDTD per-time-step engine.
 *
 * This library stands in for the backend that MaxwellFDTD selects with
 * `solve!(model; backend=...)` (src/backends.jl:1-41, src/models.jl:166-183).
 * The caller keeps building its Model on the host in FP64 (src/models.jl:57-122,
 * src/pml.jl:57-304, src/materials.jl:64-483, src/sources.jl:52-73,175-186) and hands
 * the precomputed coefficients to this library through plain pointers and sizes.
 * The library copies every array it is given; no host pointer is kept after a call.
 *
 * Conventions
 *  - Every entry point returns int32 status: MFD_OK (0) or a negative MFD_ERR_* code.
 *    mfd_last_error() returns a thread-local message for the last failure.
 *  - Arrays are column-major with x fastest (Julia layout): cell (ix,iy,iz), 0-based,
 *    has linear index ix + Nx*(iy + Ny*iz).  1D grids use Nx = Ny = 1 (axis z),
 *    2D grids use Ny = 1 (axes x,z), exactly as Field1D/Field2D (src/fields.jl:7-46).
 *  - Field state, coefficients, CPML psi and ADE states are FP32 on the device
 *    (FP64 for 1D models with MFD_OPT_F64).  DFT accumulators, the plasma density /
 *    rate integration and the TFSF auxiliary 1D grid are FP64.  Time is
 *    t[it] = (it-1)*dt in FP64.
 *  - A handle is not thread-safe; distinct handles may be used from distinct threads.
 */
#ifndef MAXWELL_FDTD_H
#define MAXWELL_FDTD_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define MFD_ABI_VERSION 1

/* CODATA 2018 constants used by the TFSF face corrections (Cm = dt/MU0, Ce = dt/EPS0,
 * src/sources.jl:725), identical to PhysicalConstants.CODATA2018 (src/MaxwellFDTD.jl:17-23). */
#define MFD_C0 299792458.0
#define MFD_EPS0 8.8541878128e-12
#define MFD_MU0 1.25663706212e-6

/* ---- status codes ---------------------------------------------------------------- */
enum {
    MFD_OK = 0,
    MFD_ERR_ARG = -1,        /* bad argument (sizes, indices, enum values)            */
    MFD_ERR_STATE = -2,      /* call not valid in the handle's current state          */
    MFD_ERR_HIP = -3,        /* HIP runtime error                                     */
    MFD_ERR_NOMEM = -4,      /* device allocation failed / over budget                */
    MFD_ERR_NAN = -5,        /* "Something went wrong. I found NaN field values."     */
    MFD_ERR_INF = -6,        /* "Something went wrong. I found Inf field values."     */
    MFD_ERR_NODEV = -7       /* no usable gfx950 device                               */
};

/* ---- field components: field order of Field3D (src/fields.jl:66-78) --------------- */
enum {
    MFD_HX = 0, MFD_HY = 1, MFD_HZ = 2,
    MFD_DX = 3, MFD_DY = 4, MFD_DZ = 5,
    MFD_EX = 6, MFD_EY = 7, MFD_EZ = 8
};

/* ---- boundary conditions: bc2int (src/models.jl:1199-1216) ------------------------- */
enum { MFD_BC_PERIODIC = 1, MFD_BC_DIRICHLET = 2, MFD_BC_NEUMANN = 3 };

/* ---- CPML faces: xlayer1, xlayer2, ylayer1, ylayer2, zlayer1, zlayer2 (src/pml.jl) -- */
enum { MFD_FACE_XL = 0, MFD_FACE_XR, MFD_FACE_YL, MFD_FACE_YR, MFD_FACE_ZL, MFD_FACE_ZR };

/* ---- sources ----------------------------------------------------------------------- */
enum { MFD_SRC_SOFT = 0, MFD_SRC_HARD = 1 };
enum { MFD_TFSF_TABLES = 0, MFD_TFSF_AUX1D = 1 };

/* ---- ionization rate of a plasma material ----------------------------------------- */
enum { MFD_RATE_NONE = 0, MFD_RATE_POWER = 1, MFD_RATE_LOGLOG = 2 };

/* ---- integral outputs (src/outputs.jl:182-191,318-327,475-485,46-65) --------------- */
enum { MFD_INT_SA = 1, MFD_INT_E2 = 2 };
enum { MFD_OUT_SA = 0, MFD_OUT_E2 = 1, MFD_OUT_JE = 2, MFD_OUT_RHO = 3 };

/* Model geometry and vacuum coefficients.
 * Replaces Model(grid, source; tmax, CN, bc, pml, material) -> (Nt, dt, bcs, Mh, Me, Md1,
 * Md2) (src/models.jl:57-122) and time_step (src/models.jl:1181-1196).               */
typedef struct mfd_model_desc {
    int32_t ndims;            /* 1, 2 or 3                                            */
    int32_t Nx, Ny, Nz;       /* 1D: Nx = Ny = 1; 2D: Ny = 1                           */
    double dx, dy, dz;        /* grid steps (unused axes ignored)                      */
    double dt;                /* time step                                             */
    int32_t Nt;               /* number of time steps (length(t))                      */
    int32_t bcs[6];           /* x-left, x-right, y-left, y-right, z-left, z-right      */
    double Mh, Me, Md1, Md2;  /* vacuum update coefficients (src/models.jl:111-114)    */
} mfd_model_desc;

/* One material: MaterialStruct1D/2D/3D coefficients (src/materials.jl:115-483).
 * Materials are numbered 1,2,... in the order they are added; geometry id 0 is vacuum.  */
typedef struct mfd_material_desc {
    double Mh, Me, Md1, Md2;              /* src/materials.jl:393-400                    */
    int32_t nq;                           /* number of ADE poles (0 = no dispersion)     */
    const double *Aq, *Bq, *Cq;           /* ade_coefficients (src/materials.jl:509-577) */
    int32_t iskerr;
    double Mk2, Mk3;                      /* EPS0*chi2, EPS0*chi3 (src/materials.jl:430-437) */
    int32_t isplasma;
    double Rava, ksi, rho0, Ap, Bp, Cp, Ma;   /* src/materials.jl:64-69,450-470           */
    int32_t rate_kind;                    /* MFD_RATE_*                                  */
    double rate_c, rate_k;                /* POWER: R(I) = rate_c * I^rate_k             */
    int32_t rate_n;                       /* LOGLOG: n points, uniformly spaced log10(I) */
    const double *rate_x, *rate_y;        /* log10(I), log10(R)                          */
} mfd_material_desc;

/* TFSF source (src/sources.jl:250-877).  box = ix1,ix2,iy1,iy2,iz1,iz2, 1-based as in
 * TFSFSource3D (2D uses ix and iz, 1D uses iz only).
 *  MFD_TFSF_TABLES: replay arrays in the field order of TFSFSource1D/2D/3D, each laid out
 *    as the reference's [face dims..., Nt] array (column-major, time slowest): 4 arrays
 *    in 1D, 8 in 2D, 24 in 3D.
 *  MFD_TFSF_AUX1D: the incident field is produced by an x/y-uniform auxiliary run on the
 *    model's own z grid, z-CPML and z boundary conditions, in vacuum, driven by soft
 *    sources aux_comp[k] (MFD_EX/EY/DX/DY/HX/HY) at plane aux_iz[k] (0-based) with
 *    coefficient aux_coeff[k] and waveform aux_wave[k*Nt + it-1].  This is the run that
 *    prepare_tfsf_record/write_tfsf_record would record for a normal-incidence plane
 *    wave (src/sources.jl:779-877), computed in FP64.                                  */
typedef struct mfd_tfsf_desc {
    int32_t box[6];
    int32_t mode;
    int32_t ntables;
    const double *const *tables;
    int32_t aux_n;
    const int32_t *aux_comp;
    const int32_t *aux_iz;
    const double *aux_coeff;
    const double *aux_wave;
} mfd_tfsf_desc;

/* ---- library / device -------------------------------------------------------------- */
int32_t mfd_version(void);
const char *mfd_last_error(void);
/* Free/total HBM bytes, CU count and the device name of the current device.           */
int32_t mfd_device_info(int64_t *free_bytes, int64_t *total_bytes, int32_t *cu_count,
                        char *name, int32_t name_len);

/* ---- model construction ------------------------------------------------------------ */
typedef struct mfd_model mfd_model;

/* Model(...) (src/models.jl:57-122).  Host-side only until the first step/upload.
 * Limits: Nz <= 65535 and Nx*Ny*Nz <= 2^31 - 1 cells (cell indices are 32-bit);
 * MFD_ERR_ARG ("bad model descriptor") beyond them.                                  */
int32_t mfd_model_create(const mfd_model_desc *desc, mfd_model **out);
/* geometry Int array, 0 = vacuum, k = k-th material (src/models.jl:85-98).          */
int32_t mfd_model_set_geometry(mfd_model *m, const int32_t *ids);
/* MaterialStruct(material, grid, dt) (src/materials.jl:389-483).                     */
int32_t mfd_model_add_material(mfd_model *m, const mfd_material_desc *mat);
/* LeftPMLLayer/RightPMLLayer (src/pml.jl:57-127): ib is 1-based, nb = length(K).     */
int32_t mfd_model_set_pml(mfd_model *m, int32_t face, int32_t ib, int32_t nb,
                          const double *K, const double *A, const double *B);
/* SoftSourceStruct / HardSourceStruct + add_source_kernel! (src/sources.jl:52-244).
 * cells are 0-based linear indices; waveform[(it-1)*n + k] = w(x_k, t[it]).            */
int32_t mfd_model_add_source(mfd_model *m, int32_t kind, int32_t comp, double coeff,
                             int32_t n, const int32_t *cells, const double *waveform);
/* SourceStruct(::TFSFSource) + add_source!(::TFSFSource*) (src/sources.jl:284-776).   */
int32_t mfd_model_add_tfsf(mfd_model *m, const mfd_tfsf_desc *tfsf);
/* FieldMonitor: Monitor/update_monitor! (src/monitors.jl:43-192).                    */
int32_t mfd_model_add_field_monitor(mfd_model *m, int32_t n, const int32_t *cells,
                                    int32_t issum);
/* SpectralMonitor + DFT functor (src/monitors.jl:237-359, src/dft.jl:1-51): explicit
 * angular frequencies w[nw] and tshift; S += F*2*exp(-i*w*(t-tshift))*dt in FP64.    */
int32_t mfd_model_add_spectral_monitor(mfd_model *m, int32_t n, const int32_t *cells,
                                       int32_t nw, const double *w, double tshift,
                                       int32_t issum);
/* update_integral_variables! (src/outputs.jl:475-485): flags = MFD_INT_SA|MFD_INT_E2. */
int32_t mfd_model_enable_integrals(mfd_model *m, int32_t flags);

/* ---- running ----------------------------------------------------------------------- */
/* Device bytes the model needs once uploaded.                                         */
int32_t mfd_model_footprint(mfd_model *m, int64_t *bytes);
/* Allocate device state and copy coefficients (done implicitly by the first step).    */
int32_t mfd_model_upload(mfd_model *m);
/* Runs reference steps it+1 .. it+nsteps of the solve! loop (src/models.jl:199-234):
 * step! (H, E, sources), then monitors and integrals.  Returns MFD_ERR_NAN/MFD_ERR_INF
 * if Ex became non-finite (src/models.jl:202-209).                                     */
int32_t mfd_model_step(mfd_model *m, int32_t nsteps);
/* Number of completed steps.                                                          */
int32_t mfd_model_get_it(mfd_model *m, int32_t *it);
/* MFD_OK, or MFD_ERR_NAN / MFD_ERR_INF once the model's Ex became non-finite (a model
 * in that state is not stepped further by batches or the server).                    */
int32_t mfd_model_get_status(mfd_model *m, int32_t *status);

/* ---- results ----------------------------------------------------------------------- */
/* Copy one field component (Nx*Ny*Nz values) to host FP64.                            */
int32_t mfd_model_get_field(mfd_model *m, int32_t comp, double *out);
/* Overwrite one field component from host FP64 (initial conditions).                 */
int32_t mfd_model_set_field(mfd_model *m, int32_t comp, const double *in);
/* Monitor data.  Field monitor: FP64 [n x Nt] (point fastest) or [Nt] if issum.
 * Spectral monitor: complex FP64 interleaved (re,im) [n x nw] or [nw] if issum.       */
int32_t mfd_model_get_monitor(mfd_model *m, int32_t imon, int32_t comp, double *out);
/* MFD_OUT_SA / MFD_OUT_E2 / MFD_OUT_JE / MFD_OUT_RHO (sum of rho*rho0 over materials),
 * Nx*Ny*Nz FP64 values (src/outputs.jl:46-65).                                         */
int32_t mfd_model_get_output(mfd_model *m, int32_t which, double *out);
/* Total field energy of the current state: out[0] = W_E = 1/2 sum eps0*eps*|E|^2 dV,
 * out[1] = W_H = 1/2 sum mu0*mu*|H|^2 dV, from E^n and H^(n-1/2), summed in FP64 on the
 * device (dV = dx*dy*dz in 3D, dx*dz in 2D, dz in 1D), with eps and mu of the cell's
 * material (eps0*eps = 1/Me, mu0*mu = dt/Mh, src/materials.jl:393-400).  No
 * counterpart in the reference, which computes no energy: it is the energy check that
 * field snapshots are judged by.  Allocates a small partial-sum buffer per call.       */
int32_t mfd_model_get_energy(mfd_model *m, double *out);
/* mfd_model_get_energy without any allocation, for callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops, harnesses that forbid
 * hipMalloc).  No counterpart in the reference.  The model must already be uploaded.
 * The partial sums go to a device array inside the library (2048 FP64 pairs, 32 KiB per
 * GPU, held until the process ends; any model size fits); the result pair (W_E, W_H)
 * is written to the caller's device memory d_out[0..1], asynchronously on `stream` (a
 * hipStream_t; NULL: the model's own copy stream).  Calls that overlap on one GPU share
 * that array and race: serialise them.                                                 */
int32_t mfd_model_get_energy_noalloc(mfd_model *m, double *d_out, void *stream);
int32_t mfd_model_destroy(mfd_model *m);

/* ---- engine options and instrumentation (no counterpart in the reference: its backends
 * have no tunables; these select launch shapes of this engine) ----------------------- */
enum {
    MFD_OPT_GRAPH_STEPS = 0,   /* steps captured per hipGraph replay (0 = eager launches) */
    MFD_OPT_ZCHUNK = 1,        /* z planes marched by one 3D/2D tile (1..65535); unless   */
                               /* ZCHUNK or ROWS is set, the fused 3D step picks 16x4,    */
                               /* 4x4 or 4x2 tiles from the tile count                    */
    MFD_OPT_PERSISTENT = 2,    /* workgroups per CU of the persistent form (0 = off)       */
    MFD_OPT_KERNEL_TIMING = 3, /* hipEvent pairs around every H and E launch (eager)       */
    MFD_OPT_DFT_BLOCK = 4,     /* time steps per DFT flush (1..64, before upload)          */
    MFD_OPT_ROWS = 5,          /* y rows per 3D workgroup (1, 2, 4, 8); fused form: rows  */
                               /* per lane (2 or 4)                                       */
    MFD_OPT_KERNEL_FORM = 6,   /* 3D kernel form: 0 one cell per thread, 2 two-pass z      */
                               /* march, 3 fused H+E step over two buffer sets (default;  */
                               /* FP32 3D models with mu = 1 everywhere, else form 2)      */
    MFD_OPT_STAGGER = 7,       /* byte offset between field arrays (before upload)         */
    MFD_OPT_QUANTUM = 8,       /* server: steps per scheduling round (default 64)          */
    MFD_OPT_FILL_MCELLS = 9,   /* server: admit while active cells < value * 2^20 (16)     */
    MFD_OPT_SNAKE = 10,        /* E half walks the tile list backwards (default 1)         */
    MFD_OPT_DLESS = 11,        /* plain cells keep no D (rebuilt as E/Me on read; default 1) */
    MFD_OPT_PML_SPLIT = 12,    /* CPML-free tiles in slab-free kernels: 0 off, 1 after the */
                               /* CPML tiles, 2 beside them on a second stream; the fused  */
                               /* form splits into x-slab-only, other CPML and free tiles */
    MFD_OPT_SCHED = 13,        /* server: admit 0 in arrival order, 1 least work first (1) */
    MFD_OPT_ZCHUNK_H = 14,     /* pipelined 3D H half z planes per tile (8; 0 = ZCHUNK)    */
    MFD_OPT_XCD_H = 15,        /* 3D H tiles dealt to the XCDs like E's (1) or grid order  */
    MFD_OPT_XCD_BLOCK = 16,    /* 3D tiles per block dealt to the XCDs (64; 0: one per XCD) */
    MFD_OPT_F64 = 17,          /* 1D models: 1 keeps fields, CPML, coefficients, ADE /
                                  plasma state, source waveforms, TFSF tables and Sa/E2
                                  in FP64 (the reference's CPU{Float64} run; set before
                                  upload); monitor samples are stored in FP32           */
    MFD_OPT_TFSF_FOLD = 18,    /* fused 3D step with a TFSF source alone: 1 (default)
                                  stores the x-face H corrections in the step launch, so
                                  the face launch after it walks only the y/z-face cells;
                                  0 leaves them to the face launch; results are bitwise
                                  equal                                                  */
    MFD_OPT_FUSED_2D = 19,     /* 2D models: 1 runs the fused step (one z march per step
                                  over two Hy/Ex/Ez buffer sets), 0 (default) the two-pass
                                  kernels; results agree to rounding                    */
    MFD_OPT_RUN1D = 20         /* 1D models (up to 65536 cells, no TFSF source): 1
                                  (default) runs each mfd_model_step / mfd_batch_run call
                                  in one launch, one workgroup per model, with the cells
                                  in registers where the model allows it; 2 the same
                                  without the register form; 0 launches the pieces of
                                  every step separately; results are bitwise equal     */
};
/* Changing an option rebuilds the launch plan; field state is kept.                   */
int32_t mfd_model_set_options(mfd_model *m, int32_t opt, int32_t value);
/* Accumulated hipEvent time of the H (which = 0) or E (which = 1) half-step launches
 * since the last reset (which = 2 resets), measured on the model's stream.             */
int32_t mfd_model_timing(mfd_model *m, int32_t which, double *total_ms, int32_t *count);

/* ---- batches (no counterpart in the reference, which runs one model per solve!,
 * src/models.jl:166-242) ------------------------------------------------------------
 * A batch advances many built models of one dimensionality together: one H and one E
 * launch per half-step over the packed tile list of all unfinished models, replayed
 * from a hipGraph.  Models keep their own Nt; a model that reaches it drops out of the
 * tile list.  Every model's results are those of stepping it alone.  The caller keeps
 * ownership of the models: read them with mfd_model_get_* once the batch has run, and
 * destroy the batch before the models.                                                */
typedef struct mfd_batch mfd_batch;
int32_t mfd_batch_create(mfd_batch **out);
int32_t mfd_batch_add(mfd_batch *b, mfd_model *m);
/* MFD_OPT_GRAPH_STEPS, ZCHUNK, PERSISTENT, KERNEL_TIMING, ROWS, KERNEL_FORM.          */
int32_t mfd_batch_set_options(mfd_batch *b, int32_t opt, int32_t value);
/* Advance every unfinished model by nsteps (nsteps <= 0: until every model is done).  */
int32_t mfd_batch_run(mfd_batch *b, int32_t nsteps);
/* Time trial_steps real steps under each candidate launch shape (z-chunk, rows per
 * workgroup, persistent workgroups per CU) and keep the fastest.  The trial steps are
 * part of the solves: results do not depend on the launch shape.                      */
int32_t mfd_batch_autotune(mfd_batch *b, int32_t trial_steps, int32_t *zchunk, int32_t *rows,
                           int32_t *persistent, double *ms_per_step);
int32_t mfd_batch_active(mfd_batch *b, int32_t *nactive);
int32_t mfd_batch_destroy(mfd_batch *b);

/* ---- serving loop (no counterpart in the reference) ----------------------------------
 * A server owns a worker thread and a stream.  mfd_server_submit queues a built,
 * not-yet-stepped model; the worker admits queued models (least work, cells x Nt, first;
 * or in arrival order with MFD_OPT_SCHED = 0) while their device footprint fits the HBM
 * budget (and free HBM) and the active set holds fewer than MFD_OPT_FILL_MCELLS * 2^20
 * cells (enough to fill the GPU), uploads them, and
 * advances the active set together in rounds of MFD_OPT_QUANTUM steps, so requests that
 * arrive while others run join between rounds.  A model that finishes leaves the tile list before it
 * is reported done; its device state stays until the caller destroys the model.
 * Latency is submit -> done on the host clock.                                         */
enum { MFD_REQ_QUEUED = 0, MFD_REQ_RUNNING = 1, MFD_REQ_DONE = 2, MFD_REQ_FAILED = 3 };
typedef struct mfd_server_stats {
    int32_t submitted, queued, active, done, failed;
    double p50_ms, p99_ms, mean_ms;   /* nearest-rank percentiles over finished requests */
    double cell_updates;              /* sum of cells * Nt over finished requests        */
    double busy_s;                    /* first admitted submit -> last completion          */
    double gcells_per_s;              /* cell_updates / busy_s / 1e9                     */
    int64_t hbm_in_use;               /* footprint of admitted, unfinished requests      */
} mfd_server_stats;
typedef struct mfd_server mfd_server;
/* hbm_budget <= 0: free HBM minus 2 GiB.  max_active bounds the active set.           */
int32_t mfd_server_create(int64_t hbm_budget, int32_t max_active, mfd_server **out);
/* MFD_OPT_QUANTUM, FILL_MCELLS, SCHED, GRAPH_STEPS, ZCHUNK, PERSISTENT, ROWS, KERNEL_FORM;
 * before the first submit.                                                            */
int32_t mfd_server_set_options(mfd_server *s, int32_t opt, int32_t value);
int32_t mfd_server_submit(mfd_server *s, mfd_model *m, int64_t *ticket);
/* state = MFD_REQ_*; status = the request's MFD_OK / MFD_ERR_* once finished.          */
int32_t mfd_server_poll(mfd_server *s, int64_t ticket, int32_t *state, int32_t *status,
                        double *latency_ms);
/* Block until the ticket (ticket <= 0: every request) finished, or timeout_ms passed
 * (timeout_ms < 0: no limit).                                                         */
int32_t mfd_server_wait(mfd_server *s, int64_t ticket, int32_t timeout_ms, int32_t *state);
int32_t mfd_server_get_stats(mfd_server *s, mfd_server_stats *out);
/* Finishes every submitted request, then stops the worker.                            */
int32_t mfd_server_destroy(mfd_server *s);

#ifdef __cplusplus
}
#endif

#endif /* MAXWELL_FDTD_H */
