// Host-side engine objects behind the C ABI: a Model owns its host description and,
// once uploaded, its device state; a Runner advances a set of device-resident models
// (one model, a batch, or the server's active set) through the shared launches.
#pragma once
#include <hip/hip_runtime.h>

#include <memory>
#include <string>
#include <vector>

#include "../../include/maxwell_fdtd.h"
#include "common.hpp"

namespace mfd {

void set_error(const std::string &msg);
hipError_t copy_sync(void *dst, const void *src, size_t n, hipMemcpyKind kind, hipStream_t s);

#define MFD_HIP(call)                                                                  \
    do {                                                                               \
        hipError_t e_ = (call);                                                        \
        if (e_ != hipSuccess) {                                                        \
            ::mfd::set_error(std::string(#call) + ": " + hipGetErrorString(e_));       \
            return MFD_ERR_HIP;                                                        \
        }                                                                              \
    } while (0)

struct HostLayer {
    bool present = false;
    int ib = 0, nb = 0;
    std::vector<double> K, A, B;
};

struct HostMat {
    mfd_material_desc d{};
    std::vector<double> Aq, Bq, Cq, rx, ry;
};

struct HostSource {
    int type = 0;   // 0 point, 1 tfsf
    int kind = 0, comp = 0;
    double coeff = 0;
    std::vector<int> cells;
    std::vector<double> wave;
    int box[6] = {0, 0, 0, 0, 0, 0};
    int mode = 0;
    std::vector<std::vector<double>> tables;
    std::vector<int> aux_comp, aux_iz;
    std::vector<double> aux_coeff, aux_wave;
};

struct HostMon {
    bool spectral = false, issum = false;
    std::vector<int> cells;
    std::vector<double> w;
    double tshift = 0;
};

// device allocations of one model
struct Arena {
    std::vector<void *> ptrs;
    size_t bytes = 0;
    ~Arena();
    template <class T>
    hipError_t alloc(T **p, size_t n) {
        void *q = nullptr;
        size_t b = n * sizeof(T);
        if (b == 0) b = 16;
        hipError_t e = hipMalloc(&q, b);
        if (e != hipSuccess) return e;
        ptrs.push_back(q);
        bytes += b;
        *p = (T *)q;
        return hipSuccess;
    }
    void release();
};

struct Model {
    mfd_model_desc d{};
    long long N = 0;
    std::vector<int32_t> geom;
    bool has_geom = false;
    std::vector<HostMat> mats;
    HostLayer L[6];
    std::vector<HostSource> srcs;
    std::vector<HostMon> mons;
    int int_flags = 0;

    // device state
    bool uploaded = false;
    Arena arena;
    DevModel dm{};              // host copy of the device descriptor
    std::vector<DevMat> hmats;  // host copies of the material table
    std::vector<DevSource> hsrc;
    std::vector<DevMon> hmon;
    int it = 0;                 // completed steps (host mirror of *dm.it)
    int dft_done = 0;           // steps already contracted into the DFT accumulators
    int dft_block = 32;
    int stagger = 0;            // byte offset added per per-cell array in the field pool
    bool f64 = false;           // 1D: state kept in FP64 (MFD_OPT_F64)
    std::vector<DevMat64> hmats64;
    int ncomp_mon = 0;
    int maxsrc_n = 0, maxmon_n = 0, maxface = 0;
    std::vector<int> touched;   // cells written by sources (sorted)
    std::vector<std::vector<int>> state_cells;   // per material: cells holding ADE/plasma state
    bool has_mats = false, has_tfsf = false, has_spectral = false;
    // the touched cells less the x-face cells a fused step finishes with their H
    // corrections (the face launch's list then), built on first use
    int *d_touched_yz = nullptr;
    int n_touched_yz = -1;

    int status = 0;             // MFD_OK or the model's non-finite status
    bool user_set = false;      // the caller wrote a field component
    bool d_stale = false;       // D of plain cells not stored by the last runs
    bool d_tfsf = false;        // ... whose stored-D cells are the TFSF face cells
    hipStream_t io = nullptr;   // non-blocking stream for this model's copies
    // fused 3D step: the other H/E buffers and H-half psi arrays exist (pp_alloc), and
    // pp_parity counts the swaps mod 2 (dm always holds the current pointers)
    bool pp_alloc = false;
    int pp_parity = 0;

    int ncomp() const { return d.ndims == 3 ? 3 : d.ndims == 2 ? 2 : 1; }
    bool comp_present(int c) const;
    int64_t footprint() const;
    int ensure_io();
    int upload();
    int ensure_pp();
    int ensure_touched_yz();
    void swap_pp();
    int sync();
    int materialize_d();
    void release();
    int read_field(int comp, double *out);
    int write_field(int comp, const double *in);
};

struct RunOptions {
    int graph_steps = 16;       // steps captured per hipGraph replay (0 = eager launches)
    int zchunk = 16;            // z planes per tile (3D/2D)
    int zchunk_h = 8;           // pipelined 3D H half: its own z planes per tile (0 = zchunk)
    int ty = 4;                 // rows per workgroup of the 3D vector kernels (1, 2, 4, 8)
    int vec = 3;                // 3D kernel form: 0 one cell per thread, 2 pipelined two-pass,
                                // 3 fused step (falls back to 2 where it does not apply)
    int persistent = 0;         // 0: one workgroup per tile; >0: that many workgroups per CU
    bool kernel_timing = false; // hipEvent pairs around the H/E launches (eager mode)
    bool snake = true;          // E half walks each XCD's tile chunk backwards
    bool dless = true;          // plain cells keep no D where the run allows it
    // pipelined 3D: tiles outside every CPML layer get launches of their own, built
    // without the slab code (fewer registers, more waves).  0: one launch per half-step
    // (waves of CPML tiles, heavy in ALU work, share CUs with the memory-bound interior
    // waves); 1: the two launches one after the other; 2: the two launches side by side,
    // the CPML tiles' on a second stream forked from and joined to the runner's stream
    int pml_split = 0;
    bool xcd_h = true;          // 3D H tile list dealt to the XCDs like E's (else grid order)
    int xcd_block = 64;         // 3D: tiles per block dealt to the XCDs (0: one chunk per XCD)
    bool tfsf_fold = true;      // fused 3D step: x-face H corrections inside the launch
    bool fused_2d = false;      // 2D models: the fused step instead of the two-pass kernels
    int run1d = 1;              // 1D models without TFSF sources: whole step ranges in one
                                // launch, one workgroup per model; 1: the cells in registers
                                // where the model allows it (k_reg1d), else k_run1d; 2:
                                // k_run1d; 0: launches per step
    bool shape_set = false;     // ZCHUNK or ROWS set by the caller (else the fused step
                                // picks its tile shape from the tile count)
};

// Advances a fixed set of uploaded models together.
struct Runner {
    std::vector<Model *> models;
    DevModel *d_models = nullptr;
    Tile *d_tiles = nullptr;    // H-half tile order
    Tile *d_tiles_e = nullptr;  // E-half tile order
    int ntiles = 0;
    // pipelined 3D: d_tiles (ntiles_h) holds the ntiles_pml_h tiles some CPML layer
    // reaches, then the rest; d_tiles_e (ntiles) holds the ntiles_int_e CPML-free
    // tiles, then the layer tiles (without the CPML split, all tiles are layer tiles)
    int ntiles_h = 0, ntiles_pml_h = 0, ntiles_int_e = 0;
    int ndims = 3;
    bool any_mats = false, any_full = false, any_int = false, any_mon = false, any_spectral = false;
    bool h_mats = false;        // some material has mu != 1: the H half reads the id
    bool any_pml = false;       // some model has a CPML layer
    std::vector<Model *> dless_models;   // models this runner steps without stored D
    int maxtouched = 1;
    int kform = 0;              // 3D kernel form in use
    bool fused = false;         // kform 3: one launch per step over ping-pong H/E buffers
    bool fused2 = false;        // the same for 2D models
    bool pp_step = false;       // fused || fused2: the models' buffers swap every step
    int frows = 2;              // fused step: rows per lane
    int fzc = 16;               // fused step: z planes per tile
    int integ_f = 0;            // fused step: 0 no integrals, 3 Sa and E2 everywhere, 4 per model
    int nfc[3] = {0, 0, 0};     // fused step: tiles reaching x slabs only, other slabs, none
    std::vector<DevModel> hd;   // host copies of d_models
    std::vector<int> pp_seen;   // per model: the buffer parity d_models holds
    bool f64 = false;           // the models keep their state in FP64 (1D)
    bool run1d = false;         // run() steps the 1D models with k_run1d
    bool reg1d = false;         // ... with k_reg1d
    bool split_state = false;   // stateful cells' E computed by the state pass
    bool tfsf_gather = false;   // TFSF corrections + touched integrals in one launch
    bool xface_h = false;       // the fused step stores the x faces' H corrections too;
                                // the gather walks the other touched cells only
    bool finish_swap = false;   // fused step: nothing reads the fields between the step
                                // launch and k_finish, which then swaps the buffers too
    bool aux_needed = false;    // some TFSF aux run is stepped on the device each step
    int4 *d_chunks = nullptr;   // state-pass work list {model, material, s0, s1}
    int *d_flags = nullptr;     // per-model non-finite flags, gathered
    int *h_flags = nullptr;     // pinned host copy
    int nchunks = 0;
    int maxsrc = 0;
    std::vector<int> src_kind_mask;   // per slot: 1 point, 2 tfsf
    int maxsrc_n = 1, maxmon_n = 1, maxface = 1, int_grid = 1;
    int dft_block = 32;
    RunOptions opt;
    hipStream_t stream = nullptr;
    hipGraphExec_t gexec = nullptr;
    int gsteps = 0;
    uint64_t gmask = 0;         // steps of the graph that launch the DFT flush
    int cu_count = 256;
    std::vector<hipEvent_t> ev;
    hipStream_t side = nullptr;           // MFD_OPT_PML_SPLIT = 2: the CPML tiles' stream
    hipEvent_t fork_ev = nullptr, join_ev = nullptr;
    double t_h_ms = 0, t_e_ms = 0;
    long long n_h = 0, n_e = 0;

    ~Runner();
    int init(const std::vector<Model *> &ms, const RunOptions &o, hipStream_t s);
    void release();
    int launch_step(bool timing, bool dft_flush);   // one step of every model, on `stream`
    void advance(size_t k);           // host mirror of one step of model k
    int refresh();                    // d_models takes buffer swaps made by other runners
    uint64_t dft_mask(int steps) const;
    int graph_len() const;
    int build_graph(int steps, uint64_t mask);
    int prepare();                    // build the step graph before the first replay
    int run(int nsteps);              // nsteps steps (models that finish are skipped)
    int flush_dft_partial();
    int check_flags();
    int collect_timing();
};

int set_option(RunOptions &o, Model *m, int opt, int value);

}  // namespace mfd

// A model handle of the C ABI.  `owner` is set while a batch or a server holds it.
struct mfd_model {
    mfd::Model m;
    mfd::Runner runner;
    bool runner_ready = false;
    mfd::RunOptions opt;
    hipStream_t stream = nullptr;
    void *owner = nullptr;
};
