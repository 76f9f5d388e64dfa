// server.cpp -- solve-request serving loop over device-resident gauge fields.
//
// Callers register lattices (links uploaded once and kept in HBM, clover
// field-strength blocks built once) and submit requests (lattice, kappa,
// tolerance, source, destination).  Each request solves D x = b with the
// even-odd mixed-precision CG driver (the reference's PreconditionedBiCGStab
// entry point, source/inverters/PreconditionedBiCGStab.cpp:43-91) to
// |b - D x| <= tol |b|, through the stopping rule |r|^2 < (0.1 tol |b|)^2 on
// the even-odd system.
//
// Continuous batching: a worker thread runs a session on one lattice, a
// batched solve whose lanes are requests (engine.h).  Between chunks of
// iterations the session returns the requests whose lanes have stopped and
// admits queued requests for its lattice -- any kappa and tolerance -- into
// the free lanes, widening the batch up to max_batch lanes.  A queued request
// joins a running session on its lattice that has room; otherwise an idle
// worker starts a new session for it.  Every worker has one HIP stream
// (default priority) for all its lattices.  Device state per (worker,
// lattice) -- context, engine with its lanes' operator data and fields -- is
// created on first use and charged against an HBM budget; a worker that would
// exceed it first releases the least recently used idle state (of any
// worker), and otherwise waits until another worker's session ends.  A
// request gives the same bits alone and in any mix (engine.h).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <thread>
#include <unordered_map>
#include <vector>

#include "../../include/lym.h"
#include "engine.h"
#include "lym_common.h"

namespace lym_internal {
int fail(int code, const char* fmt, ...);
int op_ready(lym_op* op, bool force_blocks);
}  // namespace lym_internal
using lym_internal::fail;

namespace {

using clk = std::chrono::steady_clock;

struct Request {
    uint64_t id;
    int lattice;
    double kappa, tol;
    const double* src;
    double* out;
    clk::time_point t_arrival, t_start, t_done;
    int status = LYM_REQ_PENDING;
    int error = 0;
    std::string msg;
    lym_solve_info info{};
    unsigned batch = 0;
    int worker = -1;
};

struct Lattice {
    int dims[4];
    int rep, bc, storage;
    double csw;
    lym_ctx* base = nullptr;
    lym_gauge* gauge = nullptr;
};

// device state of one worker for one lattice
struct WState {
    lym_ctx* ctx = nullptr;
    lym_internal::LaneEngine* eng = nullptr;
    size_t charge = 0;
    uint64_t last_use = 0;
    bool in_use = false;
};

void free_state(WState& s) {
    if (s.eng) lym_internal::engine_destroy(s.eng);
    if (s.ctx) lym_ctx_destroy(s.ctx);
    s.eng = nullptr;
    s.ctx = nullptr;
}

double seconds(clk::time_point a, clk::time_point b) { return std::chrono::duration<double>(b - a).count(); }

}  // namespace

struct lym_server {
    int device = 0;
    unsigned nworkers = 1, max_batch = 16;
    int sloppy = LYM_PREC_HALF;
    unsigned max_steps = 20000;
    size_t budget = 0, used = 0;
    std::vector<std::unique_ptr<Lattice>> lattices;
    std::vector<std::map<int, WState>> states;  // per worker
    std::vector<hipStream_t> streams;           // per worker
    std::vector<int> sess_lattice;              // per worker: lattice of its running session, or -1
    std::vector<int> sess_room;                 // per worker: requests its session can still take
    std::mutex m;
    std::condition_variable cv_work, cv_done, cv_mem;
    std::deque<Request*> queue;
    std::unordered_map<uint64_t, std::unique_ptr<Request>> reqs;
    std::vector<std::thread> threads;
    unsigned busy = 0;
    uint64_t next_id = 1, tick = 0;
    bool stop = false;
};

namespace {

int make_state(lym_server* S, int w, int li, WState& st) {
    const Lattice& L = *S->lattices[li];
    int rc = lym_internal::ctx_create_on_stream(L.dims, L.rep, L.bc, S->streams[w], &st.ctx);
    if (!rc) rc = lym_internal::engine_create(st.ctx, L.gauge, L.csw, S->sloppy, S->max_steps, (int)S->max_batch,
                                              &st.eng);
    if (rc) free_state(st);
    return rc;
}

// the oldest queued request that no running session on its lattice will take
// (sessions absorb queued requests of their lattice up to their room);
// called with S->m held
std::deque<Request*>::iterator startable(lym_server* S) {
    std::map<int, int> room, ahead;
    for (size_t w = 0; w < S->sess_lattice.size(); ++w)
        if (S->sess_lattice[w] >= 0) room[S->sess_lattice[w]] += S->sess_room[w];
    for (auto it = S->queue.begin(); it != S->queue.end(); ++it) {
        const int li = (*it)->lattice;
        if (ahead[li]++ >= room[li]) return it;
    }
    return S->queue.end();
}

// up to n queued requests of lattice li, oldest first; called with S->m held
std::vector<Request*> take(lym_server* S, int li, int n) {
    std::vector<Request*> b;
    for (auto it = S->queue.begin(); it != S->queue.end() && (int)b.size() < n;) {
        if ((*it)->lattice == li) {
            b.push_back(*it);
            it = S->queue.erase(it);
        } else {
            ++it;
        }
    }
    return b;
}

// device state for (w, li), created within the budget; called with S->m held.
// States of other workers may be released here, but only idle ones: a worker
// holds a pointer only to the state it is using.
WState* admit(lym_server* S, int w, int li, std::unique_lock<std::mutex>& lk, int* rc) {
    auto& mine = S->states[w];
    auto it = mine.find(li);
    if (it != mine.end()) {
        it->second.last_use = ++S->tick;
        it->second.in_use = true;
        return &it->second;
    }
    const Lattice& L = *S->lattices[li];
    const size_t need = lym_internal::engine_bytes(L.dims, L.rep, L.storage, S->sloppy, (int)S->max_batch);
    if (need > S->budget) {  // never fits, whatever is released: fail now rather than wait
        *rc = fail(LYM_ERR_NOMEM, "lattice state of %zu bytes exceeds the HBM budget of %zu", need, S->budget);
        return nullptr;
    }
    for (;;) {
        while (S->used + need > S->budget) {
            int vw = -1, vl = -1;
            uint64_t oldest = UINT64_MAX;
            for (int k = 0; k < (int)S->states.size(); ++k)
                for (auto& kv : S->states[k])
                    if (!kv.second.in_use && kv.second.last_use < oldest) {
                        oldest = kv.second.last_use;
                        vw = k;
                        vl = kv.first;
                    }
            if (vw < 0) break;
            WState dead = S->states[vw][vl];
            S->states[vw].erase(vl);
            S->used -= dead.charge;
            lk.unlock();
            free_state(dead);
            lk.lock();
        }
        if (S->used + need <= S->budget) break;
        if (S->busy <= 1) {  // nothing in use elsewhere that could become idle
            *rc = fail(LYM_ERR_NOMEM, "lattice state of %zu bytes exceeds the HBM budget of %zu", need, S->budget);
            return nullptr;
        }
        S->cv_mem.wait(lk);
    }
    S->used += need;
    lk.unlock();
    WState st;
    *rc = make_state(S, w, li, st);
    lk.lock();
    if (*rc) {
        S->used -= need;
        S->cv_mem.notify_all();
        return nullptr;
    }
    st.charge = need;
    st.last_use = ++S->tick;
    st.in_use = true;
    mine[li] = st;
    return &mine[li];
}

void finish_one(Request* r, int rc, const std::string& msg, const lym_solve_info* info, unsigned batch) {
    r->t_done = clk::now();
    r->status = rc ? LYM_REQ_FAILED : LYM_REQ_DONE;
    r->error = rc;
    if (rc) r->msg = msg;
    if (info) r->info = *info;
    r->info.seconds = seconds(r->t_start, r->t_done);
    if (batch) r->batch = batch;
}

// One session: the batched solve on lattice li, fed with queued requests of
// li between chunks until no lane is busy and none is queued.
// Returns the device error that ended it, if any (its state is then unusable).
int session(lym_server* S, int w, int li, WState* st, std::vector<Request*> batch, std::unique_lock<std::mutex>& lk) {
    using lym_internal::LaneRequest;
    using lym_internal::LaneResult;
    lym_internal::LaneEngine* e = st->eng;
    for (;;) {
        // admit (lock released while the device works)
        const auto t0 = clk::now();
        for (Request* r : batch) {
            r->status = LYM_REQ_RUNNING;
            r->t_start = t0;
            r->worker = w;
        }
        lk.unlock();
        std::vector<LaneRequest> lr;
        for (Request* r : batch) lr.push_back(LaneRequest{r->src, r->out, r->kappa, r->tol, r});
        std::vector<LaneResult> failed, done;
        int rc = lym_internal::engine_admit(e, (int)lr.size(), lr.data(), &failed);
        if (!rc) rc = lym_internal::engine_run_chunk(e);
        if (!rc) rc = lym_internal::engine_harvest(e, &done);
        std::string msg = rc ? lym_last_error() : "";
        lk.lock();
        for (auto& f : failed) finish_one((Request*)f.tag, f.error, "clover block not invertible at this kappa",
                                          nullptr, 1);
        for (auto& d : done) finish_one((Request*)d.tag, LYM_OK, "", &d.info, d.max_lanes);
        if (rc) {
            // a device error ends the session: every request it holds fails
            for (Request* r : batch)
                if (r->status == LYM_REQ_RUNNING) finish_one(r, rc, msg, nullptr, 0);
            for (auto& kv : S->reqs)
                if (kv.second->status == LYM_REQ_RUNNING && kv.second->worker == w)
                    finish_one(kv.second.get(), rc, msg, nullptr, 0);
            S->cv_done.notify_all();
            S->sess_lattice[w] = -1;
            S->sess_room[w] = 0;
            S->cv_work.notify_all();
            return rc;
        }
        if (!failed.empty() || !done.empty()) S->cv_done.notify_all();
        const int room = lym_internal::engine_room(e);
        batch = take(S, li, room);
        if (batch.empty() && lym_internal::engine_busy(e) == 0) break;
        S->sess_room[w] = room - (int)batch.size();
    }
    S->sess_lattice[w] = -1;
    S->sess_room[w] = 0;
    S->cv_work.notify_all();  // queued requests of li no longer have a session to join
    return LYM_OK;
}

void worker(lym_server* S, int w) {
    (void)hipSetDevice(S->device);
    std::unique_lock<std::mutex> lk(S->m);
    for (;;) {
        std::deque<Request*>::iterator it;
        S->cv_work.wait(lk, [&] { return (S->stop && S->queue.empty()) || (it = startable(S)) != S->queue.end(); });
        if (S->stop && S->queue.empty()) break;
        const int li = (*it)->lattice;
        ++S->busy;
        int rc = LYM_OK;
        WState* st = admit(S, w, li, lk, &rc);
        if (!st) {
            // the state cannot be created: the oldest request of li fails
            std::vector<Request*> one = take(S, li, 1);
            for (Request* r : one) {
                r->t_start = clk::now();
                r->worker = w;
                finish_one(r, rc, lym_last_error(), nullptr, 0);
            }
            --S->busy;
            S->cv_done.notify_all();
            S->cv_mem.notify_all();
            continue;
        }
        std::vector<Request*> batch = take(S, li, (int)S->max_batch);
        S->sess_lattice[w] = li;
        S->sess_room[w] = (int)S->max_batch - (int)batch.size();
        const int src = batch.empty() ? LYM_OK : session(S, w, li, st, std::move(batch), lk);
        S->sess_lattice[w] = -1;
        S->sess_room[w] = 0;
        if (src) {
            // a device error left the engine's lanes undefined: release its state
            WState dead = *st;
            S->states[w].erase(li);
            S->used -= dead.charge;
            lk.unlock();
            free_state(dead);
            lk.lock();
        } else {
            st->in_use = false;
        }
        --S->busy;
        S->cv_done.notify_all();
        S->cv_mem.notify_all();
        S->cv_work.notify_all();
    }
}

}  // namespace

extern "C" {

int lym_server_create(unsigned workers, size_t hbm_budget_bytes, unsigned max_batch, lym_server** out) {
    if (!out) return fail(LYM_ERR_INVALID, "null argument");
    *out = nullptr;
    if (workers == 0 || workers > 16) return fail(LYM_ERR_INVALID, "workers must be 1..16");
    if (max_batch == 0 || max_batch > 16) return fail(LYM_ERR_INVALID, "max_batch must be 1..16");
    int dev = 0, ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) return fail(LYM_ERR_DEVICE, "no HIP device available");
    (void)hipGetDevice(&dev);
    size_t freeb = 0, total = 0;
    if (hipMemGetInfo(&freeb, &total) != hipSuccess) return fail(LYM_ERR_DEVICE, "hipMemGetInfo");
    auto* S = new lym_server();
    S->device = dev;
    S->nworkers = workers;
    S->max_batch = max_batch;
    // default budget: 80% of the free HBM
    S->budget = hbm_budget_bytes ? std::min(hbm_budget_bytes, freeb) : (size_t)(0.8 * (double)freeb);
    S->states.resize(workers);
    S->sess_lattice.assign(workers, -1);
    S->sess_room.assign(workers, 0);
    for (unsigned w = 0; w < workers; ++w) {
        hipStream_t st = nullptr;
        if (hipStreamCreateWithFlags(&st, hipStreamNonBlocking) != hipSuccess) {
            for (hipStream_t x : S->streams) (void)hipStreamDestroy(x);
            delete S;
            return fail(LYM_ERR_DEVICE, "hipStreamCreate");
        }
        S->streams.push_back(st);
    }
    for (unsigned w = 0; w < workers; ++w) S->threads.emplace_back(worker, S, (int)w);
    *out = S;
    return LYM_OK;
}

void lym_server_destroy(lym_server* S) {
    if (!S) return;
    {
        std::lock_guard<std::mutex> lk(S->m);
        S->stop = true;
    }
    S->cv_work.notify_all();
    for (auto& t : S->threads) t.join();
    for (auto& mine : S->states)
        for (auto& kv : mine) free_state(kv.second);
    for (hipStream_t st : S->streams) (void)hipStreamDestroy(st);
    for (auto& L : S->lattices) {
        if (L->gauge) lym_gauge_destroy(L->gauge);
        if (L->base) lym_ctx_destroy(L->base);
    }
    delete S;
}

int lym_server_set_solver(lym_server* S, int sloppy_prec, unsigned max_steps) {
    if (!S) return fail(LYM_ERR_INVALID, "null server");
    if (sloppy_prec != LYM_PREC_SINGLE && sloppy_prec != LYM_PREC_HALF)
        return fail(LYM_ERR_INVALID, "sloppy precision must be LYM_PREC_SINGLE or LYM_PREC_HALF");
    std::lock_guard<std::mutex> lk(S->m);
    if (!S->reqs.empty()) return fail(LYM_ERR_INVALID, "set the solver before the first request");
    S->sloppy = sloppy_prec;
    if (max_steps) S->max_steps = max_steps;
    return LYM_OK;
}

int lym_server_add_lattice(lym_server* S, const int dims[4], int rep, int bc_t, const double* links,
                           int links_have_bc_folded, int storage, double csw, int* lattice_id) {
    if (!S || !dims || !links || !lattice_id) return fail(LYM_ERR_INVALID, "null argument");
    if ((dims[0] | dims[1] | dims[2] | dims[3]) % 2)
        return fail(LYM_ERR_UNSUPPORTED, "served lattices need all extents even (the engine runs the even-odd driver)");
    (void)hipSetDevice(S->device);
    auto L = std::make_unique<Lattice>();
    std::memcpy(L->dims, dims, sizeof(L->dims));
    L->rep = rep;
    L->bc = bc_t;
    L->storage = storage;
    L->csw = csw;
    int rc = lym_ctx_create(dims, rep, bc_t, &L->base);
    if (!rc) rc = lym_gauge_create(L->base, links, links_have_bc_folded, storage, &L->gauge);
    // Build the gauge field's clover blocks once, before workers share it (a
    // lazy build by the first worker would race with the others).  Requests
    // run the even-odd driver, which needs the blocks whatever csw is (csw = 0
    // gives unit diagonal blocks); the even-odd operator always builds them.
    if (!rc) {
        lym_op* op = nullptr;
        lym_spinor *a = nullptr, *b = nullptr;
        rc = lym_op_create(L->base, LYM_OP_EO_CLOVER, 1, 0.125, csw, 1, &op);
        if (!rc) rc = lym_op_set_gauge(op, L->gauge);
        if (!rc) rc = lym_spinor_create(L->base, &a);
        if (!rc) rc = lym_spinor_create(L->base, &b);
        if (!rc) rc = lym_op_multiply(op, b, a);
        if (!rc) rc = lym_ctx_synchronize(L->base);
        if (a) lym_spinor_destroy(a);
        if (b) lym_spinor_destroy(b);
        if (op) lym_op_destroy(op);
    }
    if (rc) {
        if (L->gauge) lym_gauge_destroy(L->gauge);
        if (L->base) lym_ctx_destroy(L->base);
        return rc;
    }
    std::lock_guard<std::mutex> lk(S->m);
    S->lattices.push_back(std::move(L));
    *lattice_id = (int)S->lattices.size() - 1;
    return LYM_OK;
}

int lym_server_submit(lym_server* S, int lattice_id, double kappa, double tolerance, const double* source,
                      double* solution, uint64_t* request_id) {
    if (!S || !source || !solution || !request_id) return fail(LYM_ERR_INVALID, "null argument");
    if (!(tolerance > 0)) return fail(LYM_ERR_INVALID, "tolerance must be positive");
    if (!(kappa > 0)) return fail(LYM_ERR_INVALID, "kappa must be positive");
    auto r = std::make_unique<Request>();
    r->t_arrival = clk::now();
    r->lattice = lattice_id;
    r->kappa = kappa;
    r->tol = tolerance;
    r->src = source;
    r->out = solution;
    {
        std::lock_guard<std::mutex> lk(S->m);
        if (lattice_id < 0 || lattice_id >= (int)S->lattices.size())
            return fail(LYM_ERR_INVALID, "unknown lattice %d", lattice_id);
        r->id = S->next_id++;
        *request_id = r->id;
        S->queue.push_back(r.get());
        S->reqs[r->id] = std::move(r);
    }
    S->cv_work.notify_one();
    return LYM_OK;
}

static void fill_info(const Request& r, lym_request_info* info) {
    std::memset(info, 0, sizeof(*info));
    info->status = r.status;
    info->error = r.error;
    info->batch_size = r.batch;
    info->worker = r.worker;
    info->solve = r.info;
    if (r.status == LYM_REQ_DONE || r.status == LYM_REQ_FAILED) {
        info->queue_seconds = seconds(r.t_arrival, r.t_start);
        info->service_seconds = seconds(r.t_start, r.t_done);
        info->latency_seconds = seconds(r.t_arrival, r.t_done);
    }
}

int lym_server_wait(lym_server* S, uint64_t request_id, double timeout_seconds, lym_request_info* info) {
    if (!S || !info) return fail(LYM_ERR_INVALID, "null argument");
    std::unique_lock<std::mutex> lk(S->m);
    auto it = S->reqs.find(request_id);
    if (it == S->reqs.end()) return fail(LYM_ERR_INVALID, "unknown request %llu", (unsigned long long)request_id);
    Request* r = it->second.get();
    auto done = [&] { return r->status == LYM_REQ_DONE || r->status == LYM_REQ_FAILED; };
    if (timeout_seconds < 0) {
        S->cv_done.wait(lk, done);
    } else if (!S->cv_done.wait_for(lk, std::chrono::duration<double>(timeout_seconds), done)) {
        fill_info(*r, info);  // still queued or running: info->status says which
        return LYM_OK;
    }
    fill_info(*r, info);
    if (r->status == LYM_REQ_FAILED) return fail(r->error, "request %llu: %s", (unsigned long long)request_id,
                                                 r->msg.c_str());
    return LYM_OK;
}

int lym_server_release(lym_server* S, uint64_t request_id) {
    if (!S) return fail(LYM_ERR_INVALID, "null server");
    std::lock_guard<std::mutex> lk(S->m);
    auto it = S->reqs.find(request_id);
    if (it == S->reqs.end()) return fail(LYM_ERR_INVALID, "unknown request");
    if (it->second->status != LYM_REQ_DONE && it->second->status != LYM_REQ_FAILED)
        return fail(LYM_ERR_INVALID, "request %llu still in flight", (unsigned long long)request_id);
    S->reqs.erase(it);
    return LYM_OK;
}

int lym_server_drain(lym_server* S) {
    if (!S) return fail(LYM_ERR_INVALID, "null server");
    std::unique_lock<std::mutex> lk(S->m);
    S->cv_done.wait(lk, [&] { return S->queue.empty() && S->busy == 0; });
    return LYM_OK;
}

int lym_server_memory(lym_server* S, size_t* budget, size_t* in_use) {
    if (!S) return fail(LYM_ERR_INVALID, "null server");
    std::lock_guard<std::mutex> lk(S->m);
    if (budget) *budget = S->budget;
    if (in_use) *in_use = S->used;
    return LYM_OK;
}

}  // extern "C"
