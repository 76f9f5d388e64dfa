// solvers.cpp -- CG, BiCGStab and the even-odd BiCGStab driver on the device.
//
// ConjugateGradient::solve        source/inverters/ConjugateGradient.cpp:10-74
// BiConjugateGradient::solve      source/inverters/BiConjugateGradient.cpp:321-507
// PreconditionedBiCGStab::solve   source/inverters/PreconditionedBiCGStab.cpp:33-100
//
// The host queues iterations in chunks on the context stream and reads the
// device state block once per chunk; kernels of iterations past the stopping
// point return immediately (see blas.hip), so the result is the one the
// reference's loop stops at.  Even-odd operators are solved on the odd half
// only: on the even half the reference's Q^2 is the identity, its residual
// starts and stays zero there, and the even part of the solution is the even
// part of the source.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdint>
#include <cstdarg>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <algorithm>
#include <map>
#include <mutex>
#include <set>
#include <string>

#include "../../include/lym.h"
#include "engine.h"
#include "lym_kernels.h"

using lym::zd;

// provided by api.cpp
namespace lym_internal {
int fail(int code, const char* fmt, ...);
int op_ready(lym_op* op, bool force_blocks);
void apply_schur(lym_op* op, zd* out_o, const zd* in_o, int gamma5, const int* skip);
void apply_full(lym_op* op, zd* out, const zd* in, int gamma5, const int* skip);
void apply_full_add(lym_op* op, zd* out, const zd* in, const zd* v2, zd alpha, int gamma5, const int* skip);
int link_kind_of(const lym_gauge* g);
int op_sloppy(lym_op* op, int prec, bool full_rows = false);
// the sloppy applies can also write per-workgroup partial sums of |out|^2
// (return value: how many; 0 when norm_partials is null or the launch has more
// than norm_max workgroups)
// upd: the fused mixed-CG update on the last launch (the return value is then
// its partial count, 0 when the launch could not take it)
int apply_full_sloppy(lym_op* op, int prec, void* out, float* out_sc, const void* in, const float* in_sc, int gamma5,
                      const int* skip, double* norm_partials = nullptr, int norm_max = 0,
                      const lym::UpdArgs* upd = nullptr);
int apply_schur_sloppy(lym_op* op, int prec, void* out_o, float* out_sc, const void* in_o, const float* in_sc,
                       void* tmp_e, float* tmp_sc, int gamma5, const int* skip, double* norm_partials = nullptr,
                       int norm_max = 0, const lym::UpdArgs* upd = nullptr);
// batch_links: the batched solver's low-precision links (all rows where the
// single-RHS copy keeps rows 0, 1)
// mfma: the matrix-core form of the FP16 stencil where it applies (SU(3)
// links with all rows, 2, 4 or 8 RHS); elsewhere the VALU form runs
int apply_full_nr(lym_op* op, int prec, int nr, int threads, void* out, float* out_sc, const void* in,
                  const float* in_sc, int gamma5, const int* skip, const int* lane_flags = nullptr,
                  int lane_skip_value = 0, bool batch_links = false, double* norm_partials = nullptr,
                  int norm_max = 0, int mfma = 0, const lym::UpdArgs* upd = nullptr,
                  const lym::BFuse* bf = nullptr);
int apply_schur_nr(lym_op* op, int prec, int nr, int threads, void* out_o, float* out_sc, const void* in_o,
                   const float* in_sc, void* tmp_e, float* tmp_sc, int gamma5, const int* skip,
                   const int* lane_flags = nullptr, int lane_skip_value = 0, bool batch_links = false,
                   double* norm_partials = nullptr, int norm_max = 0, int mfma = 0,
                   const lym::UpdArgs* upd = nullptr, const lym::BFuse* bf = nullptr);
void eo_source_nr(lym_op* op, int nr, const zd* b_e, const zd* b_o, zd* te, zd* src_o);
void eo_reconstruct_nr(lym_op* op, int nr, const zd* b_e, const zd* x_o, zd* x_e);
}  // namespace lym_internal

using lym_internal::fail;

struct lym_solver {
    lym_ctx* ctx;
    int kind;
    double precision;
    unsigned max_steps;
    int even_odd;
    zd* w[8];
    void* state;
    void* host_state;
    // mixed precision
    int sloppy_prec;
    double delta;
    int mixed_prec_alloc;   // precision the sloppy buffers below were sized for
    void* sv[4];            // p, Ap, Q^2 intermediate, Schur intermediate (full-size sloppy fields)
    float* svsc[4];         // their per-site scales [2][stride]
    lym::cplx<float>* xs;   // FP32 x_s, r_s [2][12][stride]
    lym::cplx<float>* rs;
    double* upd_part;       // |r_s|^2 partials of the fused update (NORM_FOLD_MAX)
    // multi-RHS batches
    int rhs_block;          // RHS per batch (0: autotuned)
    int tuned_nr, tuned_threads;
    int matrix_cores;       // the batched FP16 stencil may run on the matrix cores
    int tuned_mfma;         // the form the last batched call chose
    // batches whose right-hand sides have their own kappa (lanes of a serving
    // engine on this solver's context), kept for the next call with the same
    // gauge field, csw, precision, width and step limit
    lym_internal::LaneEngine* eng;
    const lym_gauge* eng_gauge;
    double eng_csw;
    int eng_prec, eng_w;
    unsigned eng_steps;
    int bnr, bdom, bprec;   // nr, domain parities and sloppy precision the batch buffers hold
    zd* bw[10];             // FP64 nr-fields
    void* bsv[4];           // sloppy nr-fields p, Ap, Q^2 intermediate, Schur intermediate
    float* bsvsc[4];
    lym::cplx<float>* bxs;
    lym::cplx<float>* brs;
    double* bsite;          // one value per (site, RHS): the fused steps' terms
    double* bpartials;
    void* bstate;
    void* bhost;
    zd** bptr_dev;
    zd** bptr_host;
    zd* bzero;              // zero source / dummy destination for padded lanes
    // plain BiCGStab on A + shift (the reference's shifted overload)
    double shift_re, shift_im;
    // captured iteration chunks (see queue_iterations), with the id of the
    // operator each one launches
    uint64_t graph_key[16];
    hipGraphExec_t graph_exec[16];
    uint64_t graph_used[16], graph_tick;
    uint64_t graph_op[16];
};

namespace {

#define HIPCHECK(expr)                                                                             \
    do {                                                                                           \
        hipError_t e_ = (expr);                                                                    \
        if (e_ != hipSuccess) return fail(LYM_ERR_DEVICE, "%s: %s", #expr, hipGetErrorString(e_)); \
    } while (0)

inline long half_n(const lym_ctx* c) { return 12L * c->g.stride; }
inline long full_n(const lym_ctx* c) { return 24L * c->g.stride; }
inline zd* par(zd* f, const lym_ctx* c, int p) { return f + (size_t)p * 12 * c->g.stride; }
inline const zd* par(const zd* f, const lym_ctx* c, int p) { return f + (size_t)p * 12 * c->g.stride; }

// A linear map for a Krylov solver: the operator to the power `power` on full
// fields, or the Schur complement (gamma5 flag given) to `power` on odd halves.
struct LinOp {
    lym_op* op;
    bool half;
    int gamma5;
    unsigned power;
    zd* scratch;  // one vector of the map's domain
    zd shift{0.0, 0.0};  // full-lattice maps only: apply A + shift (fused into the last stencil launch)
    bool shifted() const { return shift.re != 0.0 || shift.im != 0.0; }
    void last(zd* out, const zd* in, const zd* v, const int* skip) const {
        if (shifted()) lym_internal::apply_full_add(op, out, in, v, shift, gamma5, skip);
        else lym_internal::apply_full(op, out, in, gamma5, skip);
    }
    void apply(zd* out, const zd* in, const int* skip) const {
        if (power == 1) {
            if (half) lym_internal::apply_schur(op, out, in, gamma5, skip);
            else last(out, in, in, skip);
        } else {
            if (half) {
                lym_internal::apply_schur(op, scratch, in, gamma5, skip);
                lym_internal::apply_schur(op, out, scratch, gamma5, skip);
            } else {
                lym_internal::apply_full(op, scratch, in, gamma5, skip);
                last(out, scratch, in, skip);
            }
        }
    }
};

int ensure_work(lym_solver* s) {
    if (s->w[0]) return LYM_OK;
    const size_t bytes = sizeof(zd) * full_n(s->ctx);
    for (int i = 0; i < 8; ++i) {
        if (hipMalloc((void**)&s->w[i], bytes) != hipSuccess) return fail(LYM_ERR_NOMEM, "solver work vectors");
        if (hipMemsetAsync(s->w[i], 0, bytes, s->ctx->stream) != hipSuccess) return fail(LYM_ERR_DEVICE, "hipMemsetAsync");
    }
    if (hipMalloc(&s->state, 256) != hipSuccess) return fail(LYM_ERR_NOMEM, "solver state");
    if (hipHostMalloc(&s->host_state, 256, hipHostMallocDefault) != hipSuccess)
        return fail(LYM_ERR_DEVICE, "hipHostMalloc");
    return LYM_OK;
}

// iterations to queue before the next look at the state
unsigned next_chunk(double r0, double r_now, unsigned done_iters, double target, unsigned left) {
    unsigned c = 8;
    if (done_iters > 0 && r_now > 0 && r0 > r_now && target > 0) {
        double rate = std::log(r_now / r0) / done_iters;  // < 0
        double need = std::log(target / r_now) / rate;
        if (need > 0) c = (unsigned)std::ceil(need * 0.9) + 1;
    }
    if (c < 2) c = 2;
    if (c > 64) c = 64;
    if (c > left) c = left;
    return c;
}

// ---------------------------------------------------------------- iteration graphs
// An iteration of every solver here is a fixed sequence of kernel launches
// whose arguments do not change during a solve, and the host looks at the
// state only between chunks of iterations.  A chunk of GRAPH_ITERS iterations
// is therefore captured once into a HIP graph (kernel nodes only) and replayed;
// a launch-bound solve (small lattice, low precision) then costs one graph
// launch per chunk instead of ~16 kernel launches per iteration.  The capture
// is keyed by everything its launches read from the host (pointers, kappa,
// csw, flags, sizes); solvers keep the last four.  LYM_GRAPHS=0 queues the
// kernels directly.
constexpr unsigned GRAPH_ITERS = 8;
// Mixed-precision solvers launch their reliable-update kernels (accumulate,
// FP64 apply, true residual) only every REL_EVERY-th iteration; an update
// that falls due in between waits for that slot while the sloppy iteration
// goes on.  This removes 5-7 mostly-idle launches from three of four
// iterations.  GRAPH_ITERS is a multiple of REL_EVERY, so captured chunks and
// directly queued iterations place the slots identically.
#ifndef LYM_REL_EVERY
#define LYM_REL_EVERY 4
#endif
constexpr unsigned REL_EVERY = LYM_REL_EVERY;
static_assert(GRAPH_ITERS % REL_EVERY == 0, "slots must line up with captured chunks");
inline bool rel_slot(unsigned it) { return it % REL_EVERY == REL_EVERY - 1; }

struct Key {
    uint64_t h = 1469598103934665603ull;
    void bytes(const void* p, size_t n) {
        const unsigned char* c = static_cast<const unsigned char*>(p);
        for (size_t i = 0; i < n; ++i) h = (h ^ c[i]) * 1099511628211ull;
    }
    template <class T>
    Key& operator<<(const T& v) {
        bytes(&v, sizeof(v));
        return *this;
    }
};

// what the operator's launches capture (kappa is read from device memory, so a
// kappa change keeps the captures valid; csw enters only through the blocks).
// The operator's id, never reused, keeps a capture from matching another
// operator that happens to be allocated at the same addresses.
void key_op(Key& k, const lym_op* op) {
    k << op->uid << op << op->ctx << op->kind << op->power << op->dev_kappa << (op->csw != 0.0) << op->gamma5 << op->gauge << op->gauge->U
      << op->gauge->compression << op->A << op->Ainv << op->tmp << op->tmp2 << op->sloppy_prec << op->Us << op->Usf << op->As
      << op->As_sc << op->Ainvs << op->Ainvs_sc << op->ctx->g.stride << op->ctx->g.Vh;
}

bool graphs_enabled() {
    const char* e = std::getenv("LYM_GRAPHS");
    return !(e && e[0] == '0');
}

constexpr int GRAPH_SLOTS = 16;

// Live solvers, so that destroying an operator can drop every capture that
// launches it (no capture outlives the buffers it reads); the mutex also
// guards each solver's capture cache.
std::mutex& graph_mutex() {
    static std::mutex m;
    return m;
}
std::set<lym_solver*>& live_solvers() {
    static std::set<lym_solver*> s;
    return s;
}

void drop_graphs(lym_solver* s) {
    std::lock_guard<std::mutex> lk(graph_mutex());
    for (int i = 0; i < GRAPH_SLOTS; ++i) {
        if (s->graph_exec[i]) (void)hipGraphExecDestroy(s->graph_exec[i]);
        s->graph_exec[i] = nullptr;
        s->graph_key[i] = 0;
        s->graph_op[i] = 0;
    }
}

// queue k iterations of body(iteration index) (at least k: whole chunks when
// replaying a graph; iterations past the solver's stop are no-ops by its
// device state).  When replaying, the queued count is a multiple of
// GRAPH_ITERS, so the index modulo GRAPH_ITERS is the same in both paths.
template <class F>
int queue_iterations(lym_solver* s, uint64_t key, uint64_t op_uid, unsigned k, F&& body, unsigned* queued) {
    hipStream_t st = s->ctx->stream;
    if (!graphs_enabled()) {
        for (unsigned i = 0; i < k; ++i) body(*queued + i);
        *queued += k;
        return LYM_OK;
    }
    std::unique_lock<std::mutex> lk(graph_mutex());
    int slot = -1;
    for (int i = 0; i < GRAPH_SLOTS; ++i)
        if (s->graph_exec[i] && s->graph_key[i] == key) slot = i;
    if (slot < 0) {
        // a free slot, else the least recently used one (its last replay
        // ended before the host read the state that led here)
        for (int i = 0; i < GRAPH_SLOTS && slot < 0; ++i)
            if (!s->graph_exec[i]) slot = i;
        if (slot < 0) {
            slot = 0;
            for (int i = 1; i < GRAPH_SLOTS; ++i)
                if (s->graph_used[i] < s->graph_used[slot]) slot = i;
        }
        if (s->graph_exec[slot]) (void)hipGraphExecDestroy(s->graph_exec[slot]);
        s->graph_exec[slot] = nullptr;
        s->graph_key[slot] = 0;
        HIPCHECK(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
        for (unsigned i = 0; i < GRAPH_ITERS; ++i) body(i);
        hipGraph_t graph = nullptr;
        hipError_t e = hipStreamEndCapture(st, &graph);
        if (e != hipSuccess || !graph) return fail(LYM_ERR_DEVICE, "graph capture: %s", hipGetErrorString(e));
        e = hipGraphInstantiate(&s->graph_exec[slot], graph, nullptr, nullptr, 0);
        (void)hipGraphDestroy(graph);
        if (e != hipSuccess) {
            s->graph_exec[slot] = nullptr;
            return fail(LYM_ERR_DEVICE, "graph instantiate: %s", hipGetErrorString(e));
        }
        s->graph_key[slot] = key;
        s->graph_op[slot] = op_uid;
    }
    s->graph_used[slot] = ++s->graph_tick;
    hipGraphExec_t exec = s->graph_exec[slot];
    lk.unlock();
    const unsigned n = (k + GRAPH_ITERS - 1) / GRAPH_ITERS;
    for (unsigned i = 0; i < n; ++i) HIPCHECK(hipGraphLaunch(exec, st));
    *queued += n * GRAPH_ITERS;
    return LYM_OK;
}

double read_norm2(lym_ctx* c, const zd* v, long n) {
    lym::launch_norm2(v, n, c->red_partials, c->red_result, c->stream);
    double out = 0;
    (void)hipMemcpyAsync(c->host_pinned, c->red_result, sizeof(double), hipMemcpyDeviceToHost, c->stream);
    (void)hipStreamSynchronize(c->stream);
    out = c->host_pinned[0];
    return out;
}

// CG on L (ConjugateGradient.cpp:10-74): vectors of length n
int run_cg(lym_solver* s, const LinOp& L, const zd* b, zd* xout, const zd* x0, long n, lym_solve_info* info) {
    lym_ctx* c = s->ctx;
    hipStream_t st = c->stream;
    zd *x = s->w[0], *r = s->w[1], *p = s->w[2], *Ap = s->w[3];
    auto* S = reinterpret_cast<lym::CGState*>(s->state);
    auto* H = reinterpret_cast<lym::CGState*>(s->host_state);
    lym::launch_copy(x, x0 ? x0 : b, n, st);
    L.apply(Ap, x, nullptr);
    lym::launch_sub(r, b, Ap, n, st);
    lym::launch_copy(p, r, n, st);
    lym::launch_cg_init(S, r, n, c->red_partials, s->precision, (int)s->max_steps, st);
    HIPCHECK(hipMemcpyAsync(H, S, sizeof(lym::CGState), hipMemcpyDeviceToHost, st));
    HIPCHECK(hipStreamSynchronize(st));
    const double r0 = H->rr;
    unsigned queued = 0;
    Key key;
    key << 1 << s << n << b;
    key_op(key, L.op);
    key << L.half << L.gamma5 << L.power << L.scratch;
    auto body = [&](unsigned) {
        L.apply(Ap, p, &S->done);
        lym::launch_cg_alpha(S, p, Ap, n, c->red_partials, st);
        lym::launch_cg_update(S, x, r, p, Ap, n, c->red_partials, st);
        lym::launch_cg_p(S, p, r, n, c->red_partials, st);
    };
    while (!H->done && queued <= s->max_steps) {
        unsigned k = next_chunk(r0, H->rr, (unsigned)H->step, s->precision, s->max_steps + 1 - queued);
        int rc = queue_iterations(s, key.h, L.op->uid, k, body, &queued);
        if (rc) return rc;
        HIPCHECK(hipGetLastError());
        HIPCHECK(hipMemcpyAsync(H, S, sizeof(lym::CGState), hipMemcpyDeviceToHost, st));
        HIPCHECK(hipStreamSynchronize(st));
    }
    lym::launch_copy(xout, x, n, st);
    HIPCHECK(hipGetLastError());
    info->converged = H->converged;
    info->steps = H->converged ? (unsigned)H->step : s->max_steps;
    info->last_error = H->rr_new;
    info->applies = 1 + (H->converged ? (unsigned)H->step + 1 : s->max_steps);
    return LYM_OK;
}

// BiCGStab on L (BiConjugateGradient.cpp:321-507); extra_norm is added to
// |b|^2 for the zero-source test when b is the odd half of a longer vector.
int run_bicgstab(lym_solver* s, const LinOp& L, const zd* b, zd* xout, const zd* x0, long n, double extra_norm,
                 lym_solve_info* info) {
    lym_ctx* c = s->ctx;
    hipStream_t st = c->stream;
    zd *x = s->w[0], *r = s->w[1], *rh = s->w[2], *p = s->w[3], *nu = s->w[4], *sv = s->w[5], *t = s->w[6];
    auto* S = reinterpret_cast<lym::BiCGState*>(s->state);
    auto* H = reinterpret_cast<lym::BiCGState*>(s->host_state);
    if (x0) {
        lym::launch_copy(x, x0, n, st);
    } else if (read_norm2(c, b, n) + extra_norm > s->precision) {
        lym::launch_copy(x, b, n, st);
    } else {
        // the reference draws a time-seeded random start here; a fixed seed is used instead
        // (one class on a lattice with an odd extent: its second half stays zero)
        lym::launch_fill_uniform(x, std::min((int)(n / half_n(c)), 1 + c->g.tsh), c->g.stride, c->g.Vh, 0x5eedull,
                                 st);
    }
    L.apply(t, x, nullptr);
    lym::launch_bicg_setup(r, rh, p, nu, b, t, n, st);
    lym::launch_bicg_init(S, s->precision, (int)s->max_steps, st);
    HIPCHECK(hipMemcpyAsync(H, S, sizeof(lym::BiCGState), hipMemcpyDeviceToHost, st));
    HIPCHECK(hipStreamSynchronize(st));
    const double r0 = read_norm2(c, r, n);
    double rnow = r0;
    unsigned queued = 0;
    Key key;
    key << 2 << s << n << b;
    key_op(key, L.op);
    key << L.half << L.gamma5 << L.power << L.scratch << L.shift.re << L.shift.im;
    auto body = [&](unsigned) {
        lym::launch_bicg_rho(S, rh, r, n, c->red_partials, st);
        lym::launch_bicg_p(S, p, r, nu, n, c->red_partials, st);
        L.apply(nu, p, &S->done);
        lym::launch_bicg_alpha(S, rh, nu, n, c->red_partials, st);
        lym::launch_bicg_s(S, sv, r, nu, n, c->red_partials, st);
        L.apply(t, sv, &S->done);
        lym::launch_bicg_omega(S, t, sv, n, c->red_partials, st);
        lym::launch_bicg_update(S, x, r, p, sv, t, b, n, c->red_partials, st);
    };
    while (!H->done && queued <= s->max_steps) {
        unsigned k = next_chunk(r0, rnow, (unsigned)H->step, s->precision, s->max_steps + 1 - queued);
        int rc = queue_iterations(s, key.h, L.op->uid, k, body, &queued);
        if (rc) return rc;
        HIPCHECK(hipGetLastError());
        HIPCHECK(hipMemcpyAsync(H, S, sizeof(lym::BiCGState), hipMemcpyDeviceToHost, st));
        HIPCHECK(hipStreamSynchronize(st));
        rnow = H->norm;
    }
    lym::launch_copy(xout, x, n, st);
    HIPCHECK(hipGetLastError());
    info->converged = H->converged;
    unsigned iters = H->done ? (unsigned)H->step + (H->converged ? 1 : 0) : s->max_steps;
    info->steps = H->converged ? (unsigned)H->step : (H->breakdown ? (unsigned)H->step : s->max_steps);
    info->last_error = H->norm;
    info->applies = 1 + 2 * iters;
    return LYM_OK;
}

int ensure_mixed_work(lym_solver* s, int prec) {
    if (s->mixed_prec_alloc == prec) return LYM_OK;
    lym_ctx* c = s->ctx;
    for (int i = 0; i < 4; ++i) {
        if (s->sv[i]) (void)hipFree(s->sv[i]);
        if (s->svsc[i]) (void)hipFree(s->svsc[i]);
        s->sv[i] = nullptr;
        s->svsc[i] = nullptr;
    }
    if (s->xs) (void)hipFree(s->xs);
    if (s->rs) (void)hipFree(s->rs);
    s->xs = s->rs = nullptr;
    const size_t S = c->g.stride;
    const size_t es = prec == 1 ? 8 : 4;
    for (int i = 0; i < 4; ++i) {
        if (hipMalloc(&s->sv[i], 24 * S * es) != hipSuccess) return fail(LYM_ERR_NOMEM, "sloppy vectors");
        if (hipMalloc((void**)&s->svsc[i], 2 * S * sizeof(float)) != hipSuccess) return fail(LYM_ERR_NOMEM, "scales");
        if (hipMemsetAsync(s->sv[i], 0, 24 * S * es, c->stream) != hipSuccess ||
            hipMemsetAsync(s->svsc[i], 0, 2 * S * sizeof(float), c->stream) != hipSuccess)
            return fail(LYM_ERR_DEVICE, "hipMemsetAsync");
    }
    if (hipMalloc((void**)&s->xs, 24 * S * 8) != hipSuccess || hipMalloc((void**)&s->rs, 24 * S * 8) != hipSuccess)
        return fail(LYM_ERR_NOMEM, "FP32 vectors");
    if (!s->upd_part && hipMalloc((void**)&s->upd_part, sizeof(double) * lym::NORM_FOLD_MAX) != hipSuccess)
        return fail(LYM_ERR_NOMEM, "update partials");
    if (hipMemsetAsync(s->xs, 0, 24 * S * 8, c->stream) != hipSuccess ||
        hipMemsetAsync(s->rs, 0, 24 * S * 8, c->stream) != hipSuccess)
        return fail(LYM_ERR_DEVICE, "hipMemsetAsync");
    s->mixed_prec_alloc = prec;
    return LYM_OK;
}

// the sloppy counterpart of a LinOp
struct SloppyOp {
    lym_op* op;
    bool half;
    int gamma5;
    unsigned power;
    int prec;
    void* mid;     // Q^2 intermediate
    float* mid_sc;
    void* tmp_e;   // Schur intermediate
    float* tmp_sc;
    // out = Op in.  With power 2 and norm_partials set, the first apply also
    // writes per-workgroup partials of |Q in|^2 = <in, Q^2 in> (Q hermitian:
    // gamma5 set); returns how many (0: not asked for, or too many workgroups).
    // With upd also set (and the norm partials written), the second apply's
    // last launch runs the fused mixed-CG update instead of storing out;
    // *nupd receives its partial count (0: not fused, out was written).
    int apply(void* out, float* out_sc, const void* in, const float* in_sc, const int* skip,
              double* norm_partials = nullptr, lym::UpdArgs* upd = nullptr, int* nupd = nullptr) const {
        auto one = [&](void* o, float* osc, const void* i, const float* isc, double* np, const lym::UpdArgs* u) {
            if (half)
                return lym_internal::apply_schur_sloppy(op, prec, o, osc, i, isc, tmp_e, tmp_sc, gamma5, skip, np,
                                                        lym::NORM_FOLD_MAX, u);
            return lym_internal::apply_full_sloppy(op, prec, o, osc, i, isc, gamma5, skip, np, lym::NORM_FOLD_MAX, u);
        };
        if (nupd) *nupd = 0;
        if (power == 1) {
            one(out, out_sc, in, in_sc, nullptr, nullptr);
            return 0;
        }
        const int n = one(mid, mid_sc, in, in_sc, gamma5 ? norm_partials : nullptr, nullptr);
        if (upd && n > 0) {
            upd->dot = norm_partials;
            upd->ndot = n;
            const int m = one(out, out_sc, mid, mid_sc, nullptr, upd);
            if (nupd) *nupd = m;
        } else {
            one(out, out_sc, mid, mid_sc, nullptr, nullptr);
        }
        return n;
    }
};

// Mixed-precision CG on L with reliable updates (see mixed.hip); vectors of
// nparity parities
int run_mixed_cg(lym_solver* s, const LinOp& L, const SloppyOp& SL, const zd* b, zd* xout, const zd* x0, int nparity,
                 lym_solve_info* info) {
    lym_ctx* c = s->ctx;
    hipStream_t st = c->stream;
    const long n = (long)nparity * 12 * c->g.stride;
    const long nsites = (long)nparity * c->g.stride;
    const int prec = SL.prec;
    zd *x = s->w[0], *r = s->w[1], *Ax = s->w[3];
    void *p = s->sv[0], *Ap = s->sv[1];
    float *p_sc = s->svsc[0], *Ap_sc = s->svsc[1];
    auto* S = reinterpret_cast<lym::MCGState*>(s->state);
    auto* H = reinterpret_cast<lym::MCGState*>(s->host_state);
    lym::launch_copy(x, x0 ? x0 : b, n, st);
    L.apply(Ax, x, nullptr);
    lym::launch_sub(r, b, Ax, n, st);
    lym::launch_mcg_init(prec, S, r, s->rs, s->xs, p, p_sc, nsites, c->g.stride, c->red_partials, s->precision,
                         s->delta, (int)s->max_steps, st);
    HIPCHECK(hipMemcpyAsync(H, S, sizeof(lym::MCGState), hipMemcpyDeviceToHost, st));
    HIPCHECK(hipStreamSynchronize(st));
    const double r0 = H->rr_s;
    unsigned queued = 0;
    Key key;
    key << 3 << s << nsites << b << prec << s->xs << s->rs << p << Ap << p_sc << Ap_sc;
    key_op(key, L.op);
    // LYM_FUSED_NORM=0 keeps the separate <p, Ap> launch (for comparison)
    const char* fn_env = std::getenv("LYM_FUSED_NORM");
    const bool fused_norm = !(fn_env && fn_env[0] == '0');
    key << L.half << L.gamma5 << L.power << L.scratch << SL.half << SL.gamma5 << SL.power << SL.mid << SL.mid_sc
        << SL.tmp_e << SL.tmp_sc << fused_norm;
    // <p, Ap> for A = Q^2: |Q p|^2 summed by the stencil launch that writes Q p
    // (its partials in red_partials[4 RED_BLOCKS, 4 RED_BLOCKS + NORM_FOLD_MAX)),
    // else a separate dot over p and Ap
    static_assert(4 * lym::RED_BLOCKS + lym::NORM_FOLD_MAX <= 8 * lym::RED_BLOCKS, "partials buffer");
    double* norm_part = c->red_partials + 4 * lym::RED_BLOCKS;
    // the update x_s += alpha p, r_s -= alpha Ap fused into the last stencil
    // launch of A p (LYM_FUSED_UPDATE=0 keeps the separate update launch)
    const char* fu_env = std::getenv("LYM_FUSED_UPDATE");
    const bool fused_upd = fused_norm && !(fu_env && fu_env[0] == '0');
    key << fused_upd;
    lym::UpdArgs upd{S, nullptr, 0, p, p_sc, s->xs, s->rs, s->upd_part};
    auto body = [&](unsigned it) {
        int nupd = 0;
        const int nn = SL.apply(Ap, Ap_sc, p, p_sc, &S->done, fused_norm ? norm_part : nullptr,
                                fused_upd ? &upd : nullptr, &nupd);
        if (nupd == 0) {
            if (nn == 0) lym::launch_mcg_alpha(prec, S, p, p_sc, Ap, Ap_sc, nsites, c->g.stride, c->red_partials, st);
            lym::launch_mcg_update(prec, S, s->xs, s->rs, p, p_sc, Ap, Ap_sc, nsites, c->g.stride, c->red_partials,
                                   nn ? norm_part : c->red_partials, nn ? nn : lym::RED_BLOCKS, st);
        }
        lym::launch_mcg_p(prec, S, p, p_sc, s->rs, nsites, c->g.stride, c->red_partials, st,
                          nupd ? s->upd_part : nullptr, nupd);
        if (rel_slot(it)) {
            lym::launch_rel_accum(S, x, s->xs, nsites, c->g.stride, st);
            L.apply(Ax, x, &S->skip_rel);
            lym::launch_rel_resid(S, r, s->rs, b, Ax, nsites, c->g.stride, c->red_partials, st);
        }
    };
    while (!H->done && queued <= s->max_steps) {
        unsigned k = next_chunk(r0, H->rr_s, (unsigned)H->step, s->precision, s->max_steps + 1 - queued);
        int rc = queue_iterations(s, key.h, L.op->uid, k, body, &queued);
        if (rc) return rc;
        HIPCHECK(hipGetLastError());
        HIPCHECK(hipMemcpyAsync(H, S, sizeof(lym::MCGState), hipMemcpyDeviceToHost, st));
        HIPCHECK(hipStreamSynchronize(st));
    }
    if (!H->converged) lym::launch_accum_xs(x, s->xs, nsites, c->g.stride, st);
    lym::launch_copy(xout, x, n, st);
    HIPCHECK(hipGetLastError());
    info->converged = H->converged;
    info->steps = H->converged ? (unsigned)H->step : s->max_steps;
    info->last_error = H->rr_true;
    info->applies = 1 + (unsigned)H->step + (unsigned)H->nrel;
    info->reliable_updates = (unsigned)H->nrel;
    return LYM_OK;
}

void free_batch(lym_solver* s) {
    for (int i = 0; i < 10; ++i) {
        if (s->bw[i]) (void)hipFree(s->bw[i]);
        s->bw[i] = nullptr;
    }
    for (int i = 0; i < 4; ++i) {
        if (s->bsv[i]) (void)hipFree(s->bsv[i]);
        if (s->bsvsc[i]) (void)hipFree(s->bsvsc[i]);
        s->bsv[i] = nullptr;
        s->bsvsc[i] = nullptr;
    }
    if (s->bxs) (void)hipFree(s->bxs);
    if (s->brs) (void)hipFree(s->brs);
    if (s->bsite) (void)hipFree(s->bsite);
    s->bxs = s->brs = nullptr;
    s->bsite = nullptr;
    s->bnr = 0;
    s->bdom = 0;
}

int dzalloc(lym_ctx* c, void** p, size_t bytes) {
    if (hipMalloc(p, bytes) != hipSuccess) {
        *p = nullptr;
        return fail(LYM_ERR_NOMEM, "batch buffer of %zu bytes", bytes);
    }
    if (hipMemsetAsync(*p, 0, bytes, c->stream) != hipSuccess) return fail(LYM_ERR_DEVICE, "hipMemsetAsync");
    return LYM_OK;
}

// buffers for batches of nr RHS over `dom` parities (1: even-odd, 2: full)
int ensure_batch(lym_solver* s, int nr, int dom, int prec) {
    lym_ctx* c = s->ctx;
    int rc;
    if (!s->bpartials) {
        const size_t nch = (2 * (size_t)c->g.stride + lym::BATCH_CHUNK - 1) / lym::BATCH_CHUNK;
        if ((rc = dzalloc(c, (void**)&s->bpartials, sizeof(double) * lym::BMAX * nch))) return rc;
        if ((rc = dzalloc(c, &s->bstate, sizeof(lym::BState)))) return rc;
        if (hipHostMalloc(&s->bhost, sizeof(lym::BState), hipHostMallocDefault) != hipSuccess)
            return fail(LYM_ERR_DEVICE, "hipHostMalloc");
        if ((rc = dzalloc(c, (void**)&s->bptr_dev, sizeof(zd*) * lym::BMAX))) return rc;
        if (hipHostMalloc((void**)&s->bptr_host, sizeof(zd*) * lym::BMAX, hipHostMallocDefault) != hipSuccess)
            return fail(LYM_ERR_DEVICE, "hipHostMalloc");
        if ((rc = dzalloc(c, (void**)&s->bzero, sizeof(zd) * full_n(c)))) return rc;
    }
    if (s->bnr >= nr && s->bdom >= dom && s->bprec == prec) return LYM_OK;
    const size_t S = c->g.stride;
    // one component of a batched field is read through 32-bit buffer offsets
    if (S * (size_t)nr * sizeof(zd) >= (size_t(1) << 31))
        return fail(LYM_ERR_INVALID, "%d right-hand sides per batch exceed the 2 GiB component stream on this lattice", nr);
    free_batch(s);
    const size_t fullb = 24 * S * nr * sizeof(zd), domb = 12 * S * dom * nr * sizeof(zd);
    for (int i = 0; i < 10; ++i)
        if ((rc = dzalloc(c, (void**)&s->bw[i], i < 2 ? fullb : domb))) return rc;
    const size_t es = prec == 1 ? 8 : 4;
    for (int i = 0; i < 4; ++i) {
        if ((rc = dzalloc(c, &s->bsv[i], 12 * S * dom * nr * es))) return rc;
        if ((rc = dzalloc(c, (void**)&s->bsvsc[i], S * dom * nr * sizeof(float)))) return rc;
    }
    if ((rc = dzalloc(c, (void**)&s->bxs, 12 * S * dom * nr * 8))) return rc;
    if ((rc = dzalloc(c, (void**)&s->brs, 12 * S * dom * nr * 8))) return rc;
    // zero at the padding sites (past the parity's sites), which no launch writes
    if ((rc = dzalloc(c, (void**)&s->bsite, S * dom * nr * sizeof(double)))) return rc;
    s->bnr = nr;
    s->bdom = dom;
    s->bprec = prec;
    return LYM_OK;
}

// Q^power on nr-fields at storage precision prec (FP64 uses the batch's FP64 scratch)
struct BatchOp {
    lym_op* op;
    bool half;
    int gamma5;
    unsigned power;
    int nr, threads;
    void* mid;
    float* mid_sc;
    void* tmp;
    float* tmp_sc;
    int mfma = 0;  // matrix-core form of the sloppy FP16 stencil
    // lane_flags/lane_skip_value: RHS lanes whose flag equals the value do nothing
    void apply(int prec, void* out, float* out_sc, const void* in, const float* in_sc, const int* skip,
               const int* lane_flags = nullptr, int lane_skip_value = 0) const {
        auto one = [&](void* o, float* osc, const void* i, const float* isc) {
            if (half)
                lym_internal::apply_schur_nr(op, prec, nr, threads, o, osc, i, isc, tmp, tmp_sc, gamma5, skip,
                                             lane_flags, lane_skip_value, true, nullptr, 0, prec ? mfma : 0);
            else
                lym_internal::apply_full_nr(op, prec, nr, threads, o, osc, i, isc, gamma5, skip, lane_flags,
                                            lane_skip_value, true, nullptr, 0, prec ? mfma : 0);
        };
        if (power == 1) {
            one(out, out_sc, in, in_sc);
        } else {
            one(mid, mid_sc, in, in_sc);
            one(out, out_sc, mid, mid_sc);
        }
    }
};

// Autotuner: RHS per batch and workgroup size, by timing the sloppy Q apply
// per right-hand side; cached per lattice shape, link kind and precision.
struct TuneKey {
    int dims[4], lk, prec, half, mfma_allowed;
    bool operator<(const TuneKey& o) const { return std::memcmp(this, &o, sizeof(*this)) < 0; }
};
// cached choice: (nr, threads, largest nr the choice was tuned over)
// per candidate batch width: the fastest workgroup size and the time of one
// sloppy apply of a full batch with it (0: not measured)
constexpr int N_TUNE = 5;
constexpr int TUNE_NR[N_TUNE] = {4, 6, 8, 12, 16};
struct TuneVal {
    int maxnr;
    int threads[N_TUNE];
    int mfma[N_TUNE];  // the matrix-core form was the faster one
    double ms[N_TUNE];
};

// The batch width for a call of nrhs RHS: least total apply time over the
// batches it makes (full batches of nr, and the last batch of the remainder
// padded to the next supported width, at the time of the smallest measured
// width that holds it), among the widths <= maxnr.
void choose_batch(const TuneVal& tv, unsigned nrhs, int maxnr, int* nr_out, int* thr_out, int* mfma_out) {
    double best = 1e300;
    for (int i = 0; i < N_TUNE; ++i) {
        const int nr = TUNE_NR[i];
        if (nr > maxnr || tv.ms[i] <= 0) continue;
        const unsigned full = nrhs / (unsigned)nr, rem = nrhs % (unsigned)nr;
        double cost = full * tv.ms[i];
        if (rem) {
            int j = 0;
            while (j < N_TUNE && (TUNE_NR[j] < (int)rem || tv.ms[j] <= 0)) ++j;
            cost += j < N_TUNE ? tv.ms[j] : tv.ms[i];
        }
        if (nrhs == 0) cost = tv.ms[i] / nr;
        if (cost < best * 0.98) {
            best = cost;
            *nr_out = nr;
            *thr_out = tv.threads[i];
            *mfma_out = tv.mfma[i];
        }
    }
}
std::map<TuneKey, TuneVal>& tune_cache() {
    static std::map<TuneKey, TuneVal> m;
    return m;
}
std::mutex& tune_mutex() {
    static std::mutex m;
    return m;
}

// mfma_ok: the matrix-core form applies to this operator and precision and the
// solver allows it; it is then timed beside the VALU form for the widths it has
int autotune_batch(lym_solver* s, lym_op* op, bool half, int prec, unsigned nrhs, bool mfma_ok, int* nr_out,
                   int* thr_out, int* mfma_out) {
    lym_ctx* c = s->ctx;
    TuneKey key;
    std::memset(&key, 0, sizeof(key));
    for (int i = 0; i < 4; ++i) key.dims[i] = c->dims[i];
    key.lk = lym_internal::link_kind_of(op->gauge);
    key.prec = prec;
    key.half = half;
    key.mfma_allowed = mfma_ok;
    *mfma_out = 0;
    int maxnr = 4;
    for (int k : TUNE_NR)
        if ((unsigned)k <= std::max(nrhs, 4u)) maxnr = k;
    {
        std::lock_guard<std::mutex> lk(tune_mutex());
        // reuse a choice tuned over at least the candidates this call allows;
        // a choice tuned over fewer RHS is retuned
        auto it = tune_cache().find(key);
        if (it != tune_cache().end() && it->second.maxnr >= maxnr) {
            *nr_out = 4;
            *thr_out = 256;
            choose_batch(it->second, nrhs, maxnr, nr_out, thr_out, mfma_out);
            return LYM_OK;
        }
    }
    int rc = ensure_batch(s, maxnr, half ? 1 : 2, prec);
    if (rc) return rc;
    TuneVal tv;
    tv.maxnr = maxnr;
    for (int i = 0; i < N_TUNE; ++i) {
        const int nr = TUNE_NR[i];
        tv.ms[i] = 0;
        tv.threads[i] = 256;
        tv.mfma[i] = 0;
        if (nr > maxnr) continue;
        double best = 1e300;
        for (int form = 0; form < 2; ++form) {
            if (form == 1 && !(mfma_ok && (nr == 4 || nr == 8))) continue;
            for (int thr : {64, 128, 256}) {
                if (thr < nr) continue;
                BatchOp B{op, half, 1, 1, nr, thr, s->bsv[2], s->bsvsc[2], s->bsv[3], s->bsvsc[3]};
                B.mfma = form;
                B.apply(prec, s->bsv[1], s->bsvsc[1], s->bsv[0], s->bsvsc[0], nullptr);  // warm
                HIPCHECK(hipEventRecord(c->ev_start, c->stream));
                for (int k = 0; k < 3; ++k) B.apply(prec, s->bsv[1], s->bsvsc[1], s->bsv[0], s->bsvsc[0], nullptr);
                HIPCHECK(hipEventRecord(c->ev_stop, c->stream));
                HIPCHECK(hipEventSynchronize(c->ev_stop));
                float ms = 0;
                (void)hipEventElapsedTime(&ms, c->ev_start, c->ev_stop);
                if (ms / 3.0 < best * 0.98) {
                    best = ms / 3.0;
                    tv.ms[i] = best;
                    tv.threads[i] = thr;
                    tv.mfma[i] = form;
                }
            }
        }
    }
    {
        std::lock_guard<std::mutex> lk(tune_mutex());
        tune_cache()[key] = tv;
    }
    *nr_out = 4;
    *thr_out = 256;
    choose_batch(tv, nrhs, maxnr, nr_out, thr_out, mfma_out);
    return LYM_OK;
}

int next_supported(int n) {
    for (int k : {1, 2, 3, 4, 6, 8, 12, 16})
        if (k >= n) return k;
    return 16;
}

// one batch of nb <= nr RHS (lanes nb..nr-1 padded with zero sources)
int run_batch(lym_solver* s, lym_op* op, bool eo, bool half, int nr, int thr, int mfma, int nb,
              const lym_spinor* const* sources, lym_spinor* const* solutions, const double* precs,
              lym_solve_info* infos) {
    lym_ctx* c = s->ctx;
    hipStream_t st = c->stream;
    const int prec = s->sloppy_prec;
    const long S = c->g.stride;
    const int dom = (eo || half) ? 1 : 2;
    const long nfull = 24 * S * nr, ndom = 12 * S * dom * nr;
    zd *B = s->bw[0], *X = s->bw[1], *x = s->bw[2], *r = s->bw[3], *Ax = s->bw[4], *mid = s->bw[5],
       *tmp = s->bw[6], *rhs = s->bw[7], *src = s->bw[8], *te = s->bw[9];
    // padding slots of this width's layout to zero (the buffers may have held
    // another width; lattices whose parity fills its stride have none)
    {
        const int vh = c->g.Vh, G = prec == 1 ? 6 : 3;
        for (int i = 0; i < 10; ++i) lym::launch_zero_padding(16, s->bw[i], i < 2 ? 24 : 12 * dom, (int)S, vh, nr, st);
        for (int i = 0; i < 4; ++i) {
            lym::launch_zero_padding(16, s->bsv[i], G * dom, (int)S, vh, nr, st);
            lym::launch_zero_padding(4, s->bsvsc[i], dom, (int)S, vh, nr, st);
        }
        lym::launch_zero_padding(8, s->bxs, 12 * dom, (int)S, vh, nr, st);
        lym::launch_zero_padding(8, s->brs, 12 * dom, (int)S, vh, nr, st);
        lym::launch_zero_padding(8, s->bsite, dom, (int)S, vh, nr, st);
    }
    // pack sources
    for (int q = 0; q < nr; ++q) s->bptr_host[q] = q < nb ? reinterpret_cast<zd*>(sources[q]->d) : s->bzero;
    HIPCHECK(hipMemcpyAsync(s->bptr_dev, s->bptr_host, sizeof(zd*) * nr, hipMemcpyHostToDevice, st));
    lym::launch_pack(s->bptr_dev, nr, B, full_n(c), st);
    const zd* b;
    if (eo) {
        const zd* b_e = B;
        const zd* b_o = B + 12 * S * nr;
        lym_internal::eo_source_nr(op, nr, b_e, b_o, te, src);
        lym::launch_scale_g5(src, 1, (int)(S * nr), st);
        lym_internal::apply_schur_nr(op, 0, nr, thr, rhs, nullptr, src, nullptr, tmp, nullptr, 1, nullptr);
        b = rhs;
    } else {
        b = half ? B + 12 * S * nr : B;
    }
    // the EO driver solves Qhat^2 with gamma5 = 1; otherwise the operator as configured
    const int g5 = eo ? 1 : op->gamma5;
    const unsigned pw = eo ? 2u : op->power;
    BatchOp H64{op, dom == 1, g5, pw, nr, thr, mid, nullptr, tmp, nullptr};
    BatchOp HS{op, dom == 1, g5, pw, nr, thr, s->bsv[2], s->bsvsc[2], s->bsv[3], s->bsvsc[3]};
    HS.mfma = mfma;
    // FP64 start: x = b, r = b - A x
    lym::launch_copy(x, b, ndom, st);
    H64.apply(0, Ax, nullptr, x, nullptr, nullptr);
    lym::launch_sub(r, b, Ax, ndom, st);
    auto* Sdev = reinterpret_cast<lym::BState*>(s->bstate);
    auto* Sh = reinterpret_cast<lym::BState*>(s->bhost);
    std::memset(Sh, 0, sizeof(lym::BState));
    Sh->delta2 = s->delta * s->delta;
    for (int q = 0; q < nr; ++q) Sh->eps[q] = (q < nb && precs) ? precs[q] : s->precision;
    HIPCHECK(hipMemcpyAsync(Sdev, Sh, sizeof(lym::BState), hipMemcpyHostToDevice, st));
    lym::BcgArgs a;
    a.x = x;
    a.r = r;
    a.b = b;
    a.Ax = Ax;
    a.xs = s->bxs;
    a.rs = s->brs;
    a.p = s->bsv[0];
    a.p_sc = s->bsvsc[0];
    a.Ap = s->bsv[1];
    a.Ap_sc = s->bsvsc[1];
    a.nsites = (long)dom * S;
    a.stride = (int)S;
    a.max_steps = (int)s->max_steps;
    a.partials = s->bpartials;
    a.bsite = s->bsite;
    a.vh = c->g.Vh;
    if (!lym::launch_bcg_step(0, prec, nr, Sdev, a, st)) return fail(LYM_ERR_UNSUPPORTED, "batch of %d RHS", nr);
    HIPCHECK(hipMemcpyAsync(Sh, Sdev, sizeof(lym::BState), hipMemcpyDeviceToHost, st));
    HIPCHECK(hipStreamSynchronize(st));
    unsigned queued = 0;
    // steps are queued in chunks; those queued after the last RHS stopped exit at once
    Key key;
    key << 4 << s << nr << thr << prec << eo << half << g5 << pw << x << r << b << Ax << mid << tmp << a.xs << a.rs
        << a.p << a.Ap << a.p_sc << a.Ap_sc << a.nsites << HS.mid << HS.tmp << HS.mid_sc << HS.tmp_sc;
    key_op(key, op);
    // The even-odd sloppy Qhat^2 p with the CG's two sums fused into its
    // stencil launches (LYM_FUSED_BATCH=0 keeps the separate dot and update
    // launches): the launch writing Qhat p also leaves |(Qhat p)_site|^2 per
    // (site, RHS), so <p, A p> = |Qhat p|^2 sums over the fixed chunks; the
    // last launch updates x_s and r_s from its unrounded FP32 output instead of
    // storing A p and leaves |r_s|^2 per (site, RHS).
    const char* fb_env = std::getenv("LYM_FUSED_BATCH");
    const bool fused = prec >= 1 && dom == 1 && g5 == 1 && pw == 2 && !(fb_env && fb_env[0] == '0');
    key << fused << s->bsite;
    const lym::BFuse bf1{1, s->bsite, nullptr, nullptr, nullptr, nullptr, nullptr};
    const lym::BFuse bf2{2, s->bsite, &Sdev->alpha[0], a.p, a.p_sc, a.xs, a.rs};
    auto body = [&](unsigned it) {
        // finished RHS skip the sloppy apply; only RHS with a due update take the FP64 one
        if (fused) {
            lym_internal::apply_schur_nr(op, prec, nr, thr, HS.mid, HS.mid_sc, a.p, a.p_sc, HS.tmp, HS.tmp_sc, 1,
                                         &Sdev->all_done, Sdev->done, 1, true, nullptr, 0, mfma, nullptr, &bf1);
            lym::launch_bcg_step(7, prec, nr, Sdev, a, st);
            lym_internal::apply_schur_nr(op, prec, nr, thr, nullptr, nullptr, HS.mid, HS.mid_sc, HS.tmp, HS.tmp_sc,
                                         1, &Sdev->all_done, Sdev->done, 1, true, nullptr, 0, mfma, nullptr, &bf2);
            lym::launch_bcg_step(8, prec, nr, Sdev, a, st);
        } else {
            HS.apply(prec, a.Ap, a.Ap_sc, a.p, a.p_sc, &Sdev->all_done, Sdev->done, 1);
            lym::launch_bcg_step(1, prec, nr, Sdev, a, st);
            lym::launch_bcg_step(2, prec, nr, Sdev, a, st);
        }
        lym::launch_bcg_step(5, prec, nr, Sdev, a, st);
        if (rel_slot(it)) {
            lym::launch_bcg_step(3, prec, nr, Sdev, a, st);
            H64.apply(0, Ax, nullptr, x, nullptr, &Sdev->skip_rel, Sdev->rel, 0);
            lym::launch_bcg_step(4, prec, nr, Sdev, a, st);
        }
    };
    while (!Sh->all_done && queued <= s->max_steps) {
        unsigned k = queued == 0 ? 8 : 16;
        if (k > s->max_steps + 1 - queued) k = s->max_steps + 1 - queued;
        int rc = queue_iterations(s, key.h, op->uid, k, body, &queued);
        if (rc) return rc;
        HIPCHECK(hipGetLastError());
        HIPCHECK(hipMemcpyAsync(Sh, Sdev, sizeof(lym::BState), hipMemcpyDeviceToHost, st));
        HIPCHECK(hipStreamSynchronize(st));
    }
    lym::launch_bcg_step(6, prec, nr, Sdev, a, st);
    // solutions
    if (eo) {
        lym::launch_copy(X + 12 * S * nr, x, 12 * S * nr, st);
        lym_internal::eo_reconstruct_nr(op, nr, B, X + 12 * S * nr, X);
    } else if (half) {
        lym::launch_copy(X, B, 12 * S * nr, st);
        lym::launch_copy(X + 12 * S * nr, x, 12 * S * nr, st);
    } else {
        lym::launch_copy(X, x, nfull, st);
    }
    for (int q = 0; q < nr; ++q) s->bptr_host[q] = q < nb ? reinterpret_cast<zd*>(solutions[q]->d) : s->bzero;
    HIPCHECK(hipMemcpyAsync(s->bptr_dev, s->bptr_host, sizeof(zd*) * nr, hipMemcpyHostToDevice, st));
    lym::launch_unpack(X, nr, s->bptr_dev, full_n(c), st);
    HIPCHECK(hipMemcpyAsync(Sh, Sdev, sizeof(lym::BState), hipMemcpyDeviceToHost, st));
    HIPCHECK(hipStreamSynchronize(st));
    HIPCHECK(hipMemsetAsync(s->bzero, 0, sizeof(zd) * full_n(c), st));
    for (int q = 0; q < nb; ++q) {
        lym_solve_info& f = infos[q];
        f.converged = Sh->conv[q];
        f.steps = Sh->conv[q] ? (unsigned)Sh->step[q] : s->max_steps;
        f.last_error = Sh->rr_true[q];
        f.reliable_updates = (unsigned)Sh->nrel[q];
        f.applies = 1 + (unsigned)Sh->step[q] + (unsigned)Sh->nrel[q];
    }
    return LYM_OK;
}

}  // namespace

namespace lym_internal {
int solver_kind(const lym_solver* s) { return s->kind; }
int solver_even_odd(const lym_solver* s) { return s->even_odd; }

// Drops every captured iteration chunk that launches this operator (called
// when it is destroyed, after its stream has drained): no capture outlives
// the buffers it reads.
void forget_op_captures(const lym_op* op) {
    std::lock_guard<std::mutex> lk(graph_mutex());
    for (lym_solver* s : live_solvers())
        for (int i = 0; i < GRAPH_SLOTS; ++i)
            if (s->graph_exec[i] && s->graph_op[i] == op->uid) {
                (void)hipGraphExecDestroy(s->graph_exec[i]);
                s->graph_exec[i] = nullptr;
                s->graph_key[i] = 0;
                s->graph_op[i] = 0;
            }
}
}  // namespace lym_internal

extern "C" {

int lym_solver_solve_batch_kappas(lym_solver* s, lym_op* op, unsigned nrhs, const double* kappas,
                                  const lym_spinor* const* sources, lym_spinor* const* solutions,
                                  const double* precisions, lym_solve_info* infos) {
    if (!s || !op || (nrhs && (!kappas || !sources || !solutions || !infos))) return fail(LYM_ERR_INVALID, "null argument");
    lym_ctx* c = s->ctx;
    if (s->kind != LYM_SOLVER_PRECOND_MIXED_CG)
        return fail(LYM_ERR_UNSUPPORTED, "batches over several kappa need the even-odd mixed-precision CG solver");
    if (op->ctx != c) return fail(LYM_ERR_INVALID, "operator of another context");
    if (op->kind != LYM_OP_CLOVER || op->power != 1)
        return fail(LYM_ERR_UNSUPPORTED, "batches over several kappa run the even-odd driver of a clover operator (power 1)");
    if (!c->g.tsh) return fail(LYM_ERR_UNSUPPORTED, "batches over several kappa need all lattice extents even");
    for (unsigned q = 0; q < nrhs; ++q) {
        if (!sources[q] || !solutions[q] || sources[q]->ctx != c || solutions[q]->ctx != c)
            return fail(LYM_ERR_INVALID, "spinor %u of another context", q);
        if (!std::isfinite(kappas[q])) return fail(LYM_ERR_INVALID, "kappa %u is not a finite number", q);
    }
    // lanes per batch: 12 from 12 right-hand sides on (measured on 32^3x64
    // point sources, three kappas: as fast as 8 for 12 RHS, 7 % faster for 24,
    // 9 % for 36; 16 was 11 % slower than 8), else 8
    const int w = s->rhs_block ? s->rhs_block : (nrhs >= 12 ? 12 : 8);
    if (w > lym::BMAX) return fail(LYM_ERR_INVALID, "at most %d right-hand sides per batch", lym::BMAX);
    int rc = lym_internal::op_ready(op, true);  // the gauge field's clover field strength
    if (rc) return rc;
    if (!s->eng || s->eng_gauge != op->gauge || s->eng_csw != op->csw || s->eng_prec != s->sloppy_prec ||
        s->eng_w != w || s->eng_steps != s->max_steps) {
        if (s->eng) lym_internal::engine_destroy(s->eng);
        s->eng = nullptr;
        rc = lym_internal::engine_create(c, op->gauge, op->csw, s->sloppy_prec, s->max_steps, w, &s->eng);
        if (rc) return rc;
        s->eng_gauge = op->gauge;
        s->eng_csw = op->csw;
        s->eng_prec = s->sloppy_prec;
        s->eng_w = w;
        s->eng_steps = s->max_steps;
    }
    lym_internal::LaneEngine* e = s->eng;
    // right-hand sides enter free lanes in order of kappa (stable), so that the
    // lanes in flight together mostly share one clover block set (lanes of one
    // kappa read one set; a mix of kappas reads one set per kappa, measured
    // 1.6x slower for 36 RHS whose kappas alternate); a result does not depend
    // on its lane, its neighbours or the order of admission
    std::vector<unsigned> order(nrhs);
    for (unsigned q = 0; q < nrhs; ++q) order[q] = q;
    std::stable_sort(order.begin(), order.end(), [kappas](unsigned a, unsigned b) { return kappas[a] < kappas[b]; });
    HIPCHECK(hipEventRecord(c->ev_start, c->stream));
    // right-hand sides enter free lanes and leave when they stop; every chunk
    // of iterations advances every busy lane
    unsigned next = 0, finished = 0;
    const unsigned max_chunks = s->max_steps / 8 + 8;
    unsigned idle = 0;
    while (finished < nrhs) {
        const unsigned n = std::min<unsigned>((unsigned)lym_internal::engine_room(e), nrhs - next);
        if (n > 0) {
            std::vector<lym_internal::LaneRequest> reqs(n);
            for (unsigned i = 0; i < n; ++i) {
                const unsigned q = order[next + i];
                lym_internal::LaneRequest& r = reqs[i];
                r.src = nullptr;
                r.dst = nullptr;
                r.kappa = kappas[q];
                r.tol = 0;
                r.tag = reinterpret_cast<void*>((uintptr_t)q);
                r.src_dev = sources[q];
                r.dst_dev = solutions[q];
                r.eps = precisions ? precisions[q] : s->precision;
            }
            std::vector<lym_internal::LaneResult> failed;
            if ((rc = lym_internal::engine_admit(e, (int)n, reqs.data(), &failed))) return rc;
            for (const auto& f : failed) {
                const unsigned q = (unsigned)(uintptr_t)f.tag;
                std::memset(&infos[q], 0, sizeof(lym_solve_info));
                ++finished;
            }
            next += n;
        }
        if (lym_internal::engine_busy(e) == 0) continue;
        if ((rc = lym_internal::engine_run_chunk(e))) return rc;
        std::vector<lym_internal::LaneResult> done;
        if ((rc = lym_internal::engine_harvest(e, &done))) return rc;
        for (const auto& d : done) {
            const unsigned q = (unsigned)(uintptr_t)d.tag;
            infos[q] = d.info;
            ++finished;
        }
        idle = done.empty() ? idle + 1 : 0;
        if (idle > max_chunks) return fail(LYM_ERR_DEVICE, "batched solve made no progress");
    }
    HIPCHECK(hipEventRecord(c->ev_stop, c->stream));
    HIPCHECK(hipEventSynchronize(c->ev_stop));
    float ms = 0;
    (void)hipEventElapsedTime(&ms, c->ev_start, c->ev_stop);
    for (unsigned q = 0; q < nrhs; ++q) infos[q].seconds = ms * 1e-3;
    return LYM_OK;
}

int lym_solver_set_matrix_cores(lym_solver* s, int enable) {
    if (!s) return fail(LYM_ERR_INVALID, "null solver");
    s->matrix_cores = enable ? 1 : 0;
    return LYM_OK;
}

int lym_solver_get_matrix_cores(const lym_solver* s, int* enabled, int* used) {
    if (!s) return fail(LYM_ERR_INVALID, "null solver");
    if (enabled) *enabled = s->matrix_cores;
    if (used) *used = s->tuned_mfma;
    return LYM_OK;
}

int lym_solver_set_rhs_block(lym_solver* s, unsigned nrhs_per_batch) {
    if (!s) return fail(LYM_ERR_INVALID, "null solver");
    if (nrhs_per_batch != 0 && !lym::batch_nr_supported((int)nrhs_per_batch))
        return fail(LYM_ERR_INVALID, "RHS per batch must be 0 (autotune) or one of 1,2,3,4,6,8,12,16");
    s->rhs_block = (int)nrhs_per_batch;
    return LYM_OK;
}

int lym_solver_get_rhs_block(lym_solver* s, unsigned* nrhs, unsigned* threads) {
    if (!s) return fail(LYM_ERR_INVALID, "null solver");
    if (nrhs) *nrhs = (unsigned)(s->rhs_block ? s->rhs_block : s->tuned_nr);
    if (threads) *threads = (unsigned)s->tuned_threads;
    return LYM_OK;
}

int lym_solver_solve_batch(lym_solver* s, lym_op* op, unsigned nrhs, const lym_spinor* const* sources,
                           lym_spinor* const* solutions, const double* precisions, lym_solve_info* infos) {
    if (!s || !op || (nrhs && (!sources || !solutions || !infos))) return fail(LYM_ERR_INVALID, "null argument");
    lym_ctx* c = s->ctx;
    if (s->kind != LYM_SOLVER_MIXED_CG && s->kind != LYM_SOLVER_PRECOND_MIXED_CG)
        return fail(LYM_ERR_UNSUPPORTED, "batched solves need a mixed-precision CG solver");
    for (unsigned q = 0; q < nrhs; ++q)
        if (!sources[q] || !solutions[q] || sources[q]->ctx != c || solutions[q]->ctx != c)
            return fail(LYM_ERR_INVALID, "spinor %u of another context", q);
    if (op->ctx != c) return fail(LYM_ERR_INVALID, "operator of another context");
    if (!c->g.tsh) return fail(LYM_ERR_UNSUPPORTED, "batched solves need all lattice extents even");
    const bool eo = s->kind == LYM_SOLVER_PRECOND_MIXED_CG && s->even_odd && op->power == 1 &&
                    (op->kind == LYM_OP_CLOVER || op->kind == LYM_OP_EO_CLOVER) && c->dims[0] % 2 == 0 &&
                    c->dims[1] % 2 == 0 && c->dims[2] % 2 == 0 && c->dims[3] % 2 == 0;
    const bool half = op->kind == LYM_OP_EO_CLOVER;
    int rc = lym_internal::op_ready(op, eo);
    if (!rc) rc = lym_internal::op_sloppy(op, s->sloppy_prec, true);
    if (rc) return rc;
    // the matrix-core form: FP16 SU(3) with all three link rows in the batch copy
    const bool mfma_ok = s->matrix_cores && s->sloppy_prec == LYM_PREC_HALF && c->rep == LYM_REP_SU3_FUND &&
                         op->Usf != nullptr;
    int nr = s->rhs_block, thr = 0, mfma = mfma_ok ? 1 : 0;
    if (nr == 0) {
        rc = autotune_batch(s, op, eo || half, s->sloppy_prec, nrhs, mfma_ok, &nr, &thr, &mfma);
        if (rc) return rc;
        s->tuned_nr = nr;
        s->tuned_threads = thr;
    }
    s->tuned_mfma = mfma;
    HIPCHECK(hipEventRecord(c->ev_start, c->stream));
    for (unsigned q0 = 0; q0 < nrhs; q0 += nr) {
        int nb = (int)std::min<unsigned>(nr, nrhs - q0);
        int nrb = next_supported(nb);
        rc = ensure_batch(s, nrb, (eo || half) ? 1 : 2, s->sloppy_prec);
        if (!rc) rc = run_batch(s, op, eo, half, nrb, thr, mfma, nb, sources + q0, solutions + q0,
                                precisions ? precisions + q0 : nullptr, infos + q0);
        if (rc) return rc;
    }
    HIPCHECK(hipEventRecord(c->ev_stop, c->stream));
    HIPCHECK(hipEventSynchronize(c->ev_stop));
    float ms = 0;
    (void)hipEventElapsedTime(&ms, c->ev_start, c->ev_stop);
    for (unsigned q = 0; q < nrhs; ++q) infos[q].seconds = ms * 1e-3;
    return LYM_OK;
}

int lym_solver_create(lym_ctx* c, int kind, lym_solver** out) {
    if (!c || !out) return fail(LYM_ERR_INVALID, "null argument");
    *out = nullptr;
    if (kind < LYM_SOLVER_CG || kind > LYM_SOLVER_PRECOND_MIXED_CG) return fail(LYM_ERR_INVALID, "unknown solver %d", kind);
    lym_solver* s = new lym_solver();
    std::memset(s, 0, sizeof(*s));
    s->ctx = c;
    s->kind = kind;
    // defaults: ConjugateGradient.cpp:6 and Solver.h:10
    const bool cg = kind == LYM_SOLVER_CG || kind == LYM_SOLVER_MIXED_CG || kind == LYM_SOLVER_PRECOND_CG ||
                    kind == LYM_SOLVER_PRECOND_MIXED_CG;
    s->precision = cg ? 1e-11 : 1e-10;
    s->max_steps = cg ? 3000 : 1000;
    s->even_odd = 1;
    s->sloppy_prec = LYM_PREC_HALF;
    s->delta = 0.05;
    {
        std::lock_guard<std::mutex> lk(graph_mutex());
        live_solvers().insert(s);
    }
    *out = s;
    return LYM_OK;
}

void lym_solver_destroy(lym_solver* s) {
    if (!s) return;
    if (s->ctx && s->ctx->stream) (void)hipStreamSynchronize(s->ctx->stream);
    for (int i = 0; i < 8; ++i)
        if (s->w[i]) (void)hipFree(s->w[i]);
    if (s->state) (void)hipFree(s->state);
    for (int i = 0; i < 4; ++i) {
        if (s->sv[i]) (void)hipFree(s->sv[i]);
        if (s->svsc[i]) (void)hipFree(s->svsc[i]);
    }
    if (s->xs) (void)hipFree(s->xs);
    if (s->rs) (void)hipFree(s->rs);
    if (s->upd_part) (void)hipFree(s->upd_part);
    if (s->host_state) (void)hipHostFree(s->host_state);
    free_batch(s);
    if (s->eng) lym_internal::engine_destroy(s->eng);
    drop_graphs(s);
    {
        std::lock_guard<std::mutex> lk(graph_mutex());
        live_solvers().erase(s);
    }
    if (s->bpartials) (void)hipFree(s->bpartials);
    if (s->bstate) (void)hipFree(s->bstate);
    if (s->bhost) (void)hipHostFree(s->bhost);
    if (s->bptr_dev) (void)hipFree(s->bptr_dev);
    if (s->bptr_host) (void)hipHostFree(s->bptr_host);
    if (s->bzero) (void)hipFree(s->bzero);
    delete s;
}

int lym_solver_set_precision(lym_solver* s, double precision) {
    if (!s) return fail(LYM_ERR_INVALID, "null solver");
    s->precision = precision;
    return LYM_OK;
}

int lym_solver_set_max_steps(lym_solver* s, unsigned max_steps) {
    if (!s) return fail(LYM_ERR_INVALID, "null solver");
    s->max_steps = max_steps;
    return LYM_OK;
}

int lym_solver_set_sloppy_precision(lym_solver* s, int prec) {
    if (!s) return fail(LYM_ERR_INVALID, "null solver");
    if (prec != LYM_PREC_SINGLE && prec != LYM_PREC_HALF) return fail(LYM_ERR_INVALID, "sloppy precision %d", prec);
    s->sloppy_prec = prec;
    return LYM_OK;
}

int lym_solver_set_reliable_delta(lym_solver* s, double delta) {
    if (!s) return fail(LYM_ERR_INVALID, "null solver");
    if (!(delta > 0 && delta < 1)) return fail(LYM_ERR_INVALID, "reliable delta must lie in (0, 1)");
    s->delta = delta;
    return LYM_OK;
}

int lym_solver_set_shift(lym_solver* s, double re, double im) {
    if (!s) return fail(LYM_ERR_INVALID, "null solver");
    s->shift_re = re;
    s->shift_im = im;
    return LYM_OK;
}

int lym_solver_set_even_odd(lym_solver* s, int use) {
    if (!s) return fail(LYM_ERR_INVALID, "null solver");
    s->even_odd = use ? 1 : 0;
    return LYM_OK;
}

int lym_solver_solve(lym_solver* s, lym_op* op, const lym_spinor* source, lym_spinor* solution,
                     const lym_spinor* guess, lym_solve_info* info) {
    if (!s || !op || !source || !solution) return fail(LYM_ERR_INVALID, "null argument");
    lym_ctx* c = s->ctx;
    if (op->ctx != c || source->ctx != c || solution->ctx != c || (guess && guess->ctx != c))
        return fail(LYM_ERR_INVALID, "objects of different contexts");
    lym_solve_info local;
    if (!info) info = &local;
    std::memset(info, 0, sizeof(*info));
    const bool precond_kind = s->kind == LYM_SOLVER_PRECOND_BICGSTAB || s->kind == LYM_SOLVER_PRECOND_CG ||
                              s->kind == LYM_SOLVER_PRECOND_MIXED_CG;
    const bool mixed = s->kind == LYM_SOLVER_MIXED_CG || s->kind == LYM_SOLVER_PRECOND_MIXED_CG;
    bool eo_driver = precond_kind && s->even_odd && op->power == 1 &&
                     (op->kind == LYM_OP_CLOVER || op->kind == LYM_OP_EO_CLOVER) && c->dims[0] % 2 == 0 &&
                     c->dims[1] % 2 == 0 && c->dims[2] % 2 == 0 && c->dims[3] % 2 == 0;
    if ((s->shift_re != 0.0 || s->shift_im != 0.0) &&
        (s->kind != LYM_SOLVER_BICGSTAB || op->kind == LYM_OP_EO_CLOVER))
        return fail(LYM_ERR_UNSUPPORTED, "a shift is supported by plain BiCGStab on full-lattice operators only "
                                         "(the even-odd operator has no multiplyAdd)");
    if (mixed && !c->g.tsh)
        return fail(LYM_ERR_UNSUPPORTED, "mixed-precision solves need all lattice extents even (FP64 CG and "
                                         "BiCGStab run on lattices with an odd extent)");
    int rc = lym_internal::op_ready(op, eo_driver);
    if (rc) return rc;
    rc = ensure_work(s);
    if (rc) return rc;
    if (mixed) {
        rc = lym_internal::op_sloppy(op, s->sloppy_prec);
        if (!rc) rc = ensure_mixed_work(s, s->sloppy_prec);
        if (rc) return rc;
    }
    hipStream_t st = c->stream;
    HIPCHECK(hipEventRecord(c->ev_start, st));
    const zd* b = reinterpret_cast<const zd*>(source->d);
    zd* x = reinterpret_cast<zd*>(solution->d);
    const zd* x0 = guess ? reinterpret_cast<const zd*>(guess->d) : nullptr;
    const bool half = op->kind == LYM_OP_EO_CLOVER;
    zd* scratch = s->w[7];

    if (eo_driver) {
        // PreconditionedBiCGStab.cpp:56-86 with the operator's kappa and csw
        lym::DslashParams a;
        zd* te = s->w[7];
        zd* rhs = s->w[7] + half_n(c);
        zd* src = s->w[1];
        zd* xo = par(x, c, 1);
        lym::launch_clover_apply(c->g, reinterpret_cast<const zd*>(op->Ainv), par(b, c, 0), te, 0, st);
        std::memset(&a, 0, sizeof(a));
        a.g = c->g;
        a.links = op->gauge->U;
        a.kappa = op->kappa;
        a.hop_sign = -1.0;
        a.parity = 1;
        a.in[0] = te;
        a.diag_in[1] = par(b, c, 1);
        a.out[1] = src;
        lym::launch_dslash(a, lym_internal::link_kind_of(op->gauge), 1, false, st);
        lym::launch_scale_g5(src, 1, c->g.stride, st);
        lym_internal::apply_schur(op, rhs, src, 1, nullptr);
        // inner solve on Qhat^2 with gamma5 = 1 (the driver's own EO operators);
        // it uses w[0..6], the Q^2 intermediate lives in the operator's scratch
        LinOp L{op, true, 1, 2, reinterpret_cast<zd*>(op->tmp)};
        const zd* g0 = x0 ? par(x0, c, 1) : nullptr;
        if (s->kind == LYM_SOLVER_PRECOND_BICGSTAB) {
            double be = read_norm2(c, par(b, c, 0), half_n(c));
            rc = run_bicgstab(s, L, rhs, xo, g0, half_n(c), be, info);
        } else if (s->kind == LYM_SOLVER_PRECOND_CG) {
            rc = run_cg(s, L, rhs, xo, g0, half_n(c), info);
        } else {
            SloppyOp SL{op, true, 1, 2, s->sloppy_prec, s->sv[2], s->svsc[2], s->sv[3], s->svsc[3]};
            rc = run_mixed_cg(s, L, SL, rhs, xo, g0, 1, info);
        }
        if (rc) return rc;
        std::memset(&a, 0, sizeof(a));
        a.g = c->g;
        a.links = op->gauge->U;
        a.Ainv = op->Ainv;
        a.kappa = op->kappa;
        a.hop_sign = -1.0;
        a.parity = 0;
        a.in[1] = xo;
        a.diag_in[0] = par(b, c, 0);
        a.out[0] = par(x, c, 0);
        lym::launch_dslash(a, lym_internal::link_kind_of(op->gauge), 1, true, st);
    } else {
        LinOp L{op, half, op->gamma5, op->power, half ? reinterpret_cast<zd*>(op->tmp) : scratch};
        long n = half ? half_n(c) : full_n(c);
        const zd* bb = half ? par(b, c, 1) : b;
        zd* xx = half ? par(x, c, 1) : x;
        const zd* g0 = x0 ? (half ? par(x0, c, 1) : x0) : nullptr;
        if (mixed) {
            SloppyOp SL{op, half, op->gamma5, op->power, s->sloppy_prec, s->sv[2], s->svsc[2], s->sv[3], s->svsc[3]};
            rc = run_mixed_cg(s, L, SL, bb, xx, g0, half ? 1 : 2, info);
        } else if (s->kind == LYM_SOLVER_CG || s->kind == LYM_SOLVER_PRECOND_CG) {
            rc = run_cg(s, L, bb, xx, g0, n, info);
        } else {
            L.shift = lym::mk(s->shift_re, s->shift_im);
            rc = run_bicgstab(s, L, bb, xx, g0, n, half ? read_norm2(c, b, half_n(c)) : 0.0, info);
        }
        if (rc) return rc;
        if (half) HIPCHECK(hipMemcpyAsync(x, b, sizeof(zd) * half_n(c), hipMemcpyDeviceToDevice, st));
    }
    HIPCHECK(hipGetLastError());
    HIPCHECK(hipEventRecord(c->ev_stop, st));
    HIPCHECK(hipEventSynchronize(c->ev_stop));
    float ms = 0;
    (void)hipEventElapsedTime(&ms, c->ev_start, c->ev_stop);
    info->seconds = ms * 1e-3;
    return LYM_OK;
}

}  // extern "C"
