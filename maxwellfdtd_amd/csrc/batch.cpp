// Batched runs, the launch-shape autotuner and the serving loop.
//
// A batch advances many independent solves (uneven sizes, uneven step counts) through
// the same launches: one H and one E launch per half-step cover the packed tile list of
// every active solve, the step loop is replayed from a hipGraph, and solves that reach
// their Nt drop out of the tile list.  The server runs the same machinery on a thread
// of its own: requests queue up, are admitted against an HBM budget, join the active
// set between scheduling quanta, and leave it as soon as they finish.  The reference
// runs one model per solve! call (src/models.jl:166-242); none of this has a
// counterpart there, and every solve's arithmetic is the single-solve arithmetic.
#include <algorithm>
#include <cmath>
#include <chrono>
#include <condition_variable>
#include <deque>
#include <mutex>
#include <thread>

#include "engine.hpp"

using namespace mfd;

namespace {

double now_s() {
    using namespace std::chrono;
    return duration<double>(steady_clock::now().time_since_epoch()).count();
}

int remaining(const Model &m) { return m.d.Nt - m.it; }

// Candidate launch shapes, fastest measured one kept.  Results do not depend on the
// shape: every cell's arithmetic is the same whatever the tiling.
std::vector<RunOptions> candidates(int ndims, const RunOptions &base) {
    std::vector<RunOptions> out;
    if (ndims == 3) {
        for (int zc : {8, 16, 32})
            for (int ty : {2, 4, 8}) {
                RunOptions o = base;
                o.zchunk = zc;
                o.ty = ty;
                o.persistent = 0;
                out.push_back(o);
            }
    } else if (ndims == 2) {
        for (int zc : {8, 16, 32, 64})
            for (int tw : {1, 2, 4}) {
                RunOptions o = base;
                o.zchunk = zc;
                o.ty = tw;          // waves (64 x-columns each) per 2D tile
                o.persistent = 0;
                out.push_back(o);
            }
    } else {
        for (int p : {0, 2, 8}) {
            RunOptions o = base;
            o.persistent = p;
            out.push_back(o);
        }
    }
    return out;
}

}  // namespace

// =======================================================================================
// batch
// =======================================================================================
struct mfd_batch {
    std::vector<mfd_model *> models;
    std::vector<Model *> active;      // models in the runner's tile list
    Runner runner;
    bool ready = false;
    RunOptions opt;
    hipStream_t stream = nullptr;
    double tuned_ms = 0;
};

static int batch_refresh(mfd_batch *b) {
    std::vector<Model *> act;
    for (mfd_model *h : b->models)
        if (remaining(h->m) > 0 && h->m.status == MFD_OK) act.push_back(&h->m);
    if (b->ready && act == b->active) return MFD_OK;
    b->active = act;
    b->ready = false;
    if (act.empty()) return MFD_OK;
    if (!b->stream) MFD_HIP(hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking));
    int st = b->runner.init(act, b->opt, b->stream);
    if (st != MFD_OK) return st;
    b->ready = true;
    return MFD_OK;
}

// Runs up to n steps of the active set, repacking the tile list when solves finish.
// Solves that finish inside a chunk are skipped on the device for the rest of it.
static int batch_advance(mfd_batch *b, long long n) {
    long long left = n;
    const int G = std::max(1, b->opt.graph_steps);
    while (left > 0) {
        int st = batch_refresh(b);
        if (st != MFD_OK) return st;
        if (b->active.empty()) break;
        int rmin = INT32_MAX;
        for (Model *m : b->active) rmin = std::min(rmin, remaining(*m));
        // round up to whole graphs: a solve that finishes inside the chunk idles on
        // the device for at most G*4 - 1 steps before the repack
        const long long chunk = std::min<long long>(left, std::max<long long>(rmin, 4LL * G));
        st = b->runner.run((int)chunk);
        if (st != MFD_OK && st != MFD_ERR_NAN && st != MFD_ERR_INF) return st;
        left -= chunk;
    }
    for (mfd_model *h : b->models)
        if (h->m.status != MFD_OK) return h->m.status;
    return MFD_OK;
}

extern "C" {

int32_t mfd_batch_create(mfd_batch **out) {
    if (!out) { set_error("null argument"); return MFD_ERR_ARG; }
    *out = new mfd_batch();
    return MFD_OK;
}

int32_t mfd_batch_add(mfd_batch *b, mfd_model *h) {
    if (!b || !h) { set_error("null argument"); return MFD_ERR_ARG; }
    if (h->owner) { set_error("model already belongs to a batch or server"); return MFD_ERR_STATE; }
    if (!b->models.empty()) {
        const Model &f = b->models[0]->m;
        if (f.d.ndims != h->m.d.ndims) { set_error("a batch holds models of one dimensionality"); return MFD_ERR_ARG; }
        if (f.dft_block != h->m.dft_block) { set_error("a batch holds models of one DFT block length"); return MFD_ERR_ARG; }
    }
    h->owner = b;
    b->models.push_back(h);
    b->ready = false;
    return MFD_OK;
}

int32_t mfd_batch_set_options(mfd_batch *b, int32_t opt, int32_t value) {
    if (!b) { set_error("null batch"); return MFD_ERR_ARG; }
    if (opt == MFD_OPT_DFT_BLOCK || opt == MFD_OPT_STAGGER || opt == MFD_OPT_F64) { set_error("per-model option; set it on the model"); return MFD_ERR_ARG; }
    int st = set_option(b->opt, nullptr, opt, value);
    if (st != MFD_OK) return st;
    b->ready = false;
    return MFD_OK;
}

int32_t mfd_batch_run(mfd_batch *b, int32_t nsteps) {
    if (!b) { set_error("null batch"); return MFD_ERR_ARG; }
    const long long n = nsteps > 0 ? nsteps : (long long)INT32_MAX;
    return batch_advance(b, n);
}

int32_t mfd_batch_autotune(mfd_batch *b, int32_t trial_steps, int32_t *zchunk, int32_t *ty,
                           int32_t *persistent, double *ms_per_step) {
    if (!b || trial_steps < 1) { set_error("bad arguments"); return MFD_ERR_ARG; }
    if (b->models.empty()) { set_error("empty batch"); return MFD_ERR_STATE; }
    const int ndims = b->models[0]->m.d.ndims;
    const RunOptions base = b->opt;
    RunOptions best = base;
    double best_ms = 1e300;
    const int G = std::max(1, base.graph_steps);
    for (const RunOptions &o : candidates(ndims, base)) {
        b->opt = o;
        b->ready = false;
        int st = batch_advance(b, G);                 // builds the graph
        if (st != MFD_OK) return st;
        if (b->active.empty()) break;
        const double t0 = now_s();
        st = batch_advance(b, trial_steps);
        if (st != MFD_OK) return st;
        const double ms = (now_s() - t0) * 1e3 / trial_steps;
        if (ms < best_ms) { best_ms = ms; best = o; }
    }
    b->opt = best;
    b->ready = false;
    b->tuned_ms = best_ms;
    if (zchunk) *zchunk = best.zchunk;
    if (ty) *ty = best.ty;
    if (persistent) *persistent = best.persistent;
    if (ms_per_step) *ms_per_step = best_ms;
    return MFD_OK;
}

int32_t mfd_batch_active(mfd_batch *b, int32_t *nactive) {
    if (!b || !nactive) { set_error("null argument"); return MFD_ERR_ARG; }
    int n = 0;
    for (mfd_model *h : b->models) n += remaining(h->m) > 0 && h->m.status == MFD_OK;
    *nactive = n;
    return MFD_OK;
}

int32_t mfd_batch_destroy(mfd_batch *b) {
    if (!b) return MFD_OK;
    if (b->stream) (void)hipStreamSynchronize(b->stream);
    b->runner.release();
    for (mfd_model *h : b->models) h->owner = nullptr;
    if (b->stream) (void)hipStreamDestroy(b->stream);
    delete b;
    return MFD_OK;
}

}  // extern "C"

// =======================================================================================
// server
// =======================================================================================
namespace {

struct Request {
    int64_t ticket = 0;
    mfd_model *h = nullptr;
    int state = MFD_REQ_QUEUED;
    int state_before_finish = MFD_REQ_QUEUED;
    int status = MFD_OK;
    int64_t fp = 0;
    double t_submit = 0, t_admit = 0, t_done = 0;
    std::string err;
};

}  // namespace

struct mfd_server {
    int64_t budget = 0;
    int max_active = 64;
    int quantum = 64;
    long long fill_cells = 16LL << 20;   // admit while the active set is below this size
    int sched = 1;                       // 0: arrival order, 1: least work (cells x steps) first
    long long active_cells = 0;
    int ndims = 0;
    RunOptions opt;
    int device = 0;

    std::mutex mu;
    std::condition_variable cv_work, cv_done;
    std::deque<Request *> queue;
    std::vector<Request *> all;
    bool stop = false;
    int64_t in_use = 0;
    int nactive = 0;

    // worker-only state
    std::vector<Request *> active;
    Runner runner;
    hipStream_t stream = nullptr;
    std::thread worker;

    // statistics
    std::vector<double> latencies_ms;
    double cell_updates = 0, t_first = -1, t_last = 0;
    int nfailed = 0;

    void loop();
    void finish(Request *r, int state, int status, const std::string &err);
};

// Hands a request back to the caller.  From here on the caller may destroy its model at
// any time, so the request forgets the handle: nothing of the server touches it again
// (mfd_server_destroy included).
void mfd_server::finish(Request *r, int state, int status, const std::string &err) {
    std::lock_guard<std::mutex> lk(mu);
    r->state_before_finish = r->state;
    r->status = status;
    r->err = err;
    r->t_done = now_s();
    in_use -= r->fp;
    if (r->state_before_finish == MFD_REQ_RUNNING) active_cells -= r->h->m.N;
    if (state == MFD_REQ_DONE) {
        latencies_ms.push_back((r->t_done - r->t_submit) * 1e3);
        cell_updates += (double)r->h->m.N * r->h->m.d.Nt;
        t_last = r->t_done;
    } else {
        nfailed++;
    }
    r->h->owner = nullptr;
    r->h = nullptr;
    r->state = state;
    cv_done.notify_all();
}

void mfd_server::loop() {
    const hipError_t dev_err = hipSetDevice(device);
    bool changed = false;
    for (;;) {
        std::vector<Request *> admitted;
        {
            std::unique_lock<std::mutex> lk(mu);
            if (active.empty())
                cv_work.wait(lk, [&] { return stop || !queue.empty(); });
            if (stop && queue.empty() && active.empty()) return;
            size_t free_b = 0, total_b = 0;
            const hipError_t me = dev_err != hipSuccess ? dev_err : hipMemGetInfo(&free_b, &total_b);
            if (me != hipSuccess) {
                // the device cannot be used or queried: the queued requests fail with
                // the HIP error instead of waiting for memory that is never reported
                const std::string e = std::string("server device: ") + hipGetErrorString(me);
                while (!queue.empty()) {
                    Request *r = queue.front();
                    queue.pop_front();
                    in_use += r->fp;       // finish() subtracts it
                    lk.unlock();
                    finish(r, MFD_REQ_FAILED, MFD_ERR_HIP, e);
                    lk.lock();
                }
            }
            const int64_t reserve = (int64_t)1 << 30;
            while (!queue.empty() && (int)(active.size() + admitted.size()) < max_active) {
                // the next request: the first to arrive, or the one with the least work
                // (a queued request has not started, so its work is cells x Nt)
                auto pick = queue.begin();
                if (sched == 1)
                    for (auto it = queue.begin(); it != queue.end(); ++it)
                        if ((double)(*it)->h->m.N * (*it)->h->m.d.Nt < (double)(*pick)->h->m.N * (*pick)->h->m.d.Nt)
                            pick = it;
                Request *r = *pick;
                if (r->fp > budget) {
                    queue.erase(pick);
                    in_use += r->fp;       // finish() subtracts it
                    lk.unlock();
                    finish(r, MFD_REQ_FAILED, MFD_ERR_NOMEM, "request exceeds the server's HBM budget");
                    lk.lock();
                    continue;
                }
                if (r->fp + reserve > (int64_t)free_b && active.empty() && admitted.empty()) {
                    // nothing of this server holds device memory, so waiting frees none:
                    // the request does not fit what the device has free
                    queue.erase(pick);
                    in_use += r->fp;       // finish() subtracts it
                    lk.unlock();
                    finish(r, MFD_REQ_FAILED, MFD_ERR_NOMEM,
                           "request does not fit the free device memory (footprint + 1 GiB reserve)");
                    lk.lock();
                    continue;
                }
                if (in_use + r->fp > budget || r->fp + reserve > (int64_t)free_b) break;
                // enough cells in flight to fill the GPU: later requests wait, so that
                // requests finish close to arrival order instead of sharing the GPU
                if (!(active.empty() && admitted.empty()) && active_cells + r->h->m.N > fill_cells) break;
                queue.erase(pick);
                in_use += r->fp;
                active_cells += r->h->m.N;
                free_b -= (size_t)std::min<int64_t>((int64_t)free_b, r->fp);
                r->state = MFD_REQ_RUNNING;
                if (t_first < 0) t_first = r->t_submit;
                admitted.push_back(r);
            }
        }
        for (Request *r : admitted) {
            int st = r->h->m.upload();
            r->t_admit = now_s();
            if (st != MFD_OK) { finish(r, MFD_REQ_FAILED, st, mfd_last_error()); continue; }
            active.push_back(r);
            changed = true;
        }
        {
            std::lock_guard<std::mutex> lk(mu);
            nactive = (int)active.size();
        }
        if (active.empty()) {
            // queued work that does not fit yet: wait for the caller to free memory
            std::this_thread::sleep_for(std::chrono::milliseconds(1));
            continue;
        }
        if (changed) {
            std::vector<Model *> ms;
            for (Request *r : active) ms.push_back(&r->h->m);
            int st = runner.init(ms, opt, stream);
            if (st != MFD_OK) {
                const std::string e = mfd_last_error();
                for (Request *r : active) finish(r, MFD_REQ_FAILED, st, e);
                active.clear();
                continue;
            }
            changed = false;
        }
        int rmin = INT32_MAX;
        for (Request *r : active) rmin = std::min(rmin, remaining(r->h->m));
        const int chunk = std::max(1, std::min(quantum, rmin));
        int st = runner.run(chunk);
        if (st != MFD_OK && st != MFD_ERR_NAN && st != MFD_ERR_INF) {
            const std::string e = mfd_last_error();
            for (Request *r : active) finish(r, MFD_REQ_FAILED, st, e);
            active.clear();
            continue;
        }
        // retire finished (or failed) solves; the tile list is repacked before the
        // next chunk, so a finished solve is out of every launch before it is reported
        std::vector<Request *> keep, done;
        for (Request *r : active) {
            Model &m = r->h->m;
            if (m.status != MFD_OK || remaining(m) <= 0) done.push_back(r);
            else keep.push_back(r);
        }
        if (!done.empty()) {
            active = keep;
            changed = true;
            runner.release();
            {
                std::lock_guard<std::mutex> lk(mu);
                nactive = (int)active.size();
            }
            for (Request *r : done) {
                Model &m = r->h->m;
                m.sync();
                if (m.status != MFD_OK)
                    finish(r, MFD_REQ_FAILED, m.status,
                           m.status == MFD_ERR_NAN ? "Something went wrong. I found NaN field values."
                                                   : "Something went wrong. I found Inf field values.");
                else
                    finish(r, MFD_REQ_DONE, MFD_OK, "");
            }
        }
    }
}

extern "C" {

int32_t mfd_server_create(int64_t hbm_budget, int32_t max_active, mfd_server **out) {
    if (!out || max_active < 1) { set_error("bad arguments"); return MFD_ERR_ARG; }
    int dev = 0;
    MFD_HIP(hipGetDevice(&dev));
    size_t free_b = 0, total_b = 0;
    MFD_HIP(hipMemGetInfo(&free_b, &total_b));
    mfd_server *s = new mfd_server();
    s->device = dev;
    s->budget = hbm_budget > 0 ? hbm_budget : (int64_t)free_b - ((int64_t)2 << 30);
    s->max_active = max_active;
    if (hipStreamCreateWithFlags(&s->stream, hipStreamNonBlocking) != hipSuccess) {
        delete s;
        set_error("stream creation failed");
        return MFD_ERR_HIP;
    }
    s->worker = std::thread([s] { s->loop(); });
    *out = s;
    return MFD_OK;
}

int32_t mfd_server_set_options(mfd_server *s, int32_t opt, int32_t value) {
    if (!s) { set_error("null server"); return MFD_ERR_ARG; }
    std::lock_guard<std::mutex> lk(s->mu);
    if (!s->all.empty()) { set_error("set server options before the first submit"); return MFD_ERR_STATE; }
    if (opt == MFD_OPT_QUANTUM) {
        if (value < 1) { set_error("quantum must be >= 1"); return MFD_ERR_ARG; }
        s->quantum = value;
        return MFD_OK;
    }
    if (opt == MFD_OPT_SCHED) {
        if (value != 0 && value != 1) { set_error("schedule must be 0 (arrival order) or 1 (least work first)"); return MFD_ERR_ARG; }
        s->sched = value;
        return MFD_OK;
    }
    if (opt == MFD_OPT_FILL_MCELLS) {
        if (value < 1) { set_error("fill must be >= 1 Mi-cells"); return MFD_ERR_ARG; }
        s->fill_cells = (long long)value << 20;
        return MFD_OK;
    }
    if (opt == MFD_OPT_DFT_BLOCK || opt == MFD_OPT_STAGGER || opt == MFD_OPT_KERNEL_TIMING ||
        opt == MFD_OPT_F64) {
        set_error("not a server option");
        return MFD_ERR_ARG;
    }
    return set_option(s->opt, nullptr, opt, value);
}

int32_t mfd_server_submit(mfd_server *s, mfd_model *h, int64_t *ticket) {
    if (!s || !h || !ticket) { set_error("null argument"); return MFD_ERR_ARG; }
    if (h->owner) { set_error("model already belongs to a batch or server"); return MFD_ERR_STATE; }
    if (h->m.it != 0 || h->runner_ready) { set_error("submit a model that has not been stepped"); return MFD_ERR_STATE; }
    std::lock_guard<std::mutex> lk(s->mu);
    if (s->stop) { set_error("server is shutting down"); return MFD_ERR_STATE; }
    if (s->ndims == 0) s->ndims = h->m.d.ndims;
    if (h->m.d.ndims != s->ndims) { set_error("a server runs models of one dimensionality"); return MFD_ERR_ARG; }
    Request *r = new Request();
    r->ticket = (int64_t)s->all.size() + 1;
    r->h = h;
    r->fp = h->m.footprint();
    r->t_submit = now_s();
    h->owner = s;
    s->all.push_back(r);
    s->queue.push_back(r);
    *ticket = r->ticket;
    s->cv_work.notify_one();
    return MFD_OK;
}

int32_t mfd_server_poll(mfd_server *s, int64_t ticket, int32_t *state, int32_t *status,
                        double *latency_ms) {
    if (!s || !state) { set_error("null argument"); return MFD_ERR_ARG; }
    std::lock_guard<std::mutex> lk(s->mu);
    if (ticket < 1 || ticket > (int64_t)s->all.size()) { set_error("unknown ticket"); return MFD_ERR_ARG; }
    const Request *r = s->all[ticket - 1];
    *state = r->state;
    if (status) *status = r->status;
    if (latency_ms) *latency_ms = r->state >= MFD_REQ_DONE ? (r->t_done - r->t_submit) * 1e3 : 0.0;
    return MFD_OK;
}

int32_t mfd_server_wait(mfd_server *s, int64_t ticket, int32_t timeout_ms, int32_t *state) {
    if (!s || !state) { set_error("null argument"); return MFD_ERR_ARG; }
    std::unique_lock<std::mutex> lk(s->mu);
    auto ready = [&] {
        if (ticket > 0) return s->all[ticket - 1]->state >= MFD_REQ_DONE;
        for (const Request *r : s->all) if (r->state < MFD_REQ_DONE) return false;
        return true;
    };
    if (ticket > (int64_t)s->all.size()) { set_error("unknown ticket"); return MFD_ERR_ARG; }
    if (timeout_ms < 0) s->cv_done.wait(lk, ready);
    else s->cv_done.wait_for(lk, std::chrono::milliseconds(timeout_ms), ready);
    *state = ticket > 0 ? s->all[ticket - 1]->state : (ready() ? MFD_REQ_DONE : MFD_REQ_RUNNING);
    return MFD_OK;
}

int32_t mfd_server_get_stats(mfd_server *s, mfd_server_stats *out) {
    if (!s || !out) { set_error("null argument"); return MFD_ERR_ARG; }
    std::lock_guard<std::mutex> lk(s->mu);
    mfd_server_stats st{};
    st.submitted = (int32_t)s->all.size();
    st.queued = (int32_t)s->queue.size();
    st.active = s->nactive;
    st.done = (int32_t)s->latencies_ms.size();
    st.failed = s->nfailed;
    std::vector<double> l = s->latencies_ms;
    std::sort(l.begin(), l.end());
    auto pct = [&](double p) -> double {
        if (l.empty()) return 0.0;
        // nearest-rank percentile
        size_t k = (size_t)std::ceil(p / 100.0 * l.size());
        k = std::min(std::max<size_t>(k, 1), l.size());
        return l[k - 1];
    };
    st.p50_ms = pct(50);
    st.p99_ms = pct(99);
    double sum = 0;
    for (double v : l) sum += v;
    st.mean_ms = l.empty() ? 0.0 : sum / l.size();
    st.cell_updates = s->cell_updates;
    st.busy_s = s->t_first >= 0 && s->t_last > s->t_first ? s->t_last - s->t_first : 0.0;
    st.gcells_per_s = st.busy_s > 0 ? st.cell_updates / st.busy_s / 1e9 : 0.0;
    st.hbm_in_use = s->in_use;
    *out = st;
    return MFD_OK;
}

int32_t mfd_server_destroy(mfd_server *s) {
    if (!s) return MFD_OK;
    {
        std::lock_guard<std::mutex> lk(s->mu);
        s->stop = true;
        s->cv_work.notify_all();
    }
    if (s->worker.joinable()) s->worker.join();
    s->runner.release();
    if (s->stream) (void)hipStreamDestroy(s->stream);
    for (Request *r : s->all) {
        if (r->h && r->h->owner == s) r->h->owner = nullptr;
        delete r;
    }
    delete s;
    return MFD_OK;
}

}  // extern "C"
