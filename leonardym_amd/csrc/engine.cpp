// engine.cpp -- continuous batching of solve requests (see engine.h).
//
// Per lane (slot) q the engine holds the operator data of that request's
// kappa: kappa[q] in device memory, the FP64 clover blocks A_q = 1 + kappa_q
// csw gamma5 B and their even-site inverses, and their low-precision copies;
// the stencil reads them through per-lane strides (DslashParams::lane_ops).
// The links and the kappa-independent blocks B are the lattice's, shared.
//
// Fields are multi-RHS nr-fields (batch.hip) of the current width w, lane
// fastest, stored in buffers sized for wmax lanes.  Growing the batch
// relayouts the fields that carry a lane's state from one iteration to the
// next (sources, right-hand side, x, r, x_s, r_s, p) to the new width; a lane
// keeps its index, so its operator data never moves.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <vector>

#include "engine.h"
#include "lym_kernels.h"

using lym::zd;

namespace lym_internal {
int fail(int code, const char* fmt, ...);
int link_kind_of(const lym_gauge* g);
}  // namespace lym_internal

namespace lym_internal {

namespace {

constexpr int CHUNK = 8;      // iterations per chunk: requests join and leave between chunks
constexpr int REL_EVERY = 4;  // reliable-update slots, as in the solvers' batched CG
static_assert(CHUNK % REL_EVERY == 0, "a lane admitted at a chunk boundary meets the slots at its own iterations");
inline bool rel_slot(int it) { return it % REL_EVERY == REL_EVERY - 1; }

int next_width(int n) {
    for (int k : {1, 2, 3, 4, 6, 8, 12, 16})
        if (k >= n) return k;
    return 16;
}

#define ECHECK(expr)                                                                               \
    do {                                                                                           \
        hipError_t e_ = (expr);                                                                    \
        if (e_ != hipSuccess) return fail(LYM_ERR_DEVICE, "%s: %s", #expr, hipGetErrorString(e_)); \
    } while (0)

struct Lane {
    bool busy = false;
    LaneRequest req{};
    unsigned max_lanes = 0;
};

}  // namespace

struct LaneEngine {
    lym_ctx* ctx = nullptr;
    const lym_gauge* g = nullptr;
    double csw = 0;
    int prec = LYM_PREC_HALF, wmax = 1, width = 0;
    int wcap = 1;    // lanes the buffers hold: the smallest batch width >= wmax
    unsigned max_steps = 20000;
    long S = 0;      // component stride
    size_t es = 4;   // bytes of one sloppy complex component
    int lk64 = 0, lks = 0;  // link kinds of the FP64 and the low-precision links
    std::vector<void*> allocs;
    // operator data per slot
    double* kappa = nullptr;
    zd *A = nullptr, *Ainv = nullptr;
    void *As = nullptr, *Ainvs = nullptr, *Us = nullptr;
    float *As_sc = nullptr, *Ainvs_sc = nullptr;
    // nr-fields (FP64: B full, X full, the rest odd halves)
    zd *B = nullptr, *X = nullptr, *x = nullptr, *r = nullptr, *Ax = nullptr, *mid = nullptr, *tmp = nullptr,
       *rhs = nullptr, *src = nullptr, *te = nullptr;
    void* sv[4] = {};  // sloppy p, Ap, Q^2 intermediate, even Schur intermediate
    float* svsc[4] = {};
    lym::cplx<float> *xs = nullptr, *rs = nullptr;
    double* partials = nullptr;
    double* bsite = nullptr;    // the fused steps' per-(site, lane) terms
    bool fused = true;          // the batched solver's fused steps (LYM_FUSED_BATCH=0: separate launches)
    lym::BState* Sd = nullptr;  // device state
    lym::BState* Sh = nullptr;  // pinned host copy
    int* mask_d = nullptr;      // per-lane mask of the admission / harvest launches
    int* mask_h = nullptr;
    int* flag_d = nullptr;      // clover inverse: [0] singular blocks, [1] small pivots
    int* flag_h = nullptr;
    int* lset_d = nullptr;      // the block set of every lane (read by the stencil)
    int* lset_h = nullptr;
    // block set s holds the clover blocks, their even-site inverses and their
    // low-precision copies for kappa set_kappa[s]; lanes of one kappa share a
    // set, so a kappa already in a set is not built again
    double set_kappa[lym::BMAX] = {};
    bool set_valid[lym::BMAX] = {};
    int set_refs[lym::BMAX] = {};
    unsigned long set_used[lym::BMAX] = {};
    unsigned long set_tick = 0;
    zd** ptr_d = nullptr;
    zd** ptr_h = nullptr;
    std::vector<lym_spinor*> lsrc, lsol;  // single-RHS transfer spinors per lane
    Lane lanes[lym::BMAX];
    // captured chunks, one per width
    int graph_w[8] = {};
    hipGraphExec_t graph[8] = {};
};

namespace {

int ealloc(LaneEngine* e, void** p, size_t bytes) {
    if (hipMalloc(p, bytes) != hipSuccess) {
        *p = nullptr;
        return fail(LYM_ERR_NOMEM, "serving engine: hipMalloc of %zu bytes", bytes);
    }
    e->allocs.push_back(*p);
    if (hipMemsetAsync(*p, 0, bytes, e->ctx->stream) != hipSuccess) return fail(LYM_ERR_DEVICE, "hipMemsetAsync");
    return LYM_OK;
}

// sloppy link bytes per component and the layout the engine keeps
bool full_rows(const lym_gauge* g) {
    const int ncomp = g->real_links ? g->compression : g->compression / 2;
    return !(g->rows_rebuild && ncomp == 6);
}

// the per-lane operator data of one launch at storage precision prec
void lane_ops(LaneEngine* e, lym::DslashParams& a, int prec) {
    a.prec = prec;
    a.lane_ops = 1;
    a.lane_set = e->lset_d;
    a.kappa_dev = e->kappa;
    const long S = e->S;
    if (prec == 0) {
        a.links = e->g->U;
        a.A = e->A;
        a.Ainv = e->Ainv;
        a.A_lane_bytes = 72 * S * (long)sizeof(zd);
        a.Ainv_lane_bytes = 36 * S * (long)sizeof(zd);
    } else {
        a.links = e->Us;
        a.A = e->As;
        a.A_sc = e->As_sc;
        a.Ainv = e->Ainvs;
        a.Ainv_sc = e->Ainvs_sc;
        a.A_lane_bytes = 72 * S * (long)e->es;
        a.Ainv_lane_bytes = 36 * S * (long)e->es;
        a.Asc_lane = 2 * S;
        a.Ainvsc_lane = S;
    }
}

lym::DslashParams base(LaneEngine* e, int prec) {
    lym::DslashParams a;
    std::memset(&a, 0, sizeof(a));
    a.g = e->ctx->g;
    a.nr = e->width;
    a.hop_sign = 1.0;
    lane_ops(e, a, prec);
    return a;
}

// Schur complement with gamma5 (the even-odd driver's Qhat) on odd-half
// nr-fields; lanes whose flag equals skip_value do nothing.  bmode: the
// batched solver's fused steps on the second launch (1: also |Qhat in|^2 per
// (site, lane) into e->bsite; 2: x_s += alpha p, r_s -= alpha Qhat in with
// alpha per lane from the state, |r_s|^2 per (site, lane), no output).
void schur(LaneEngine* e, int prec, void* out, float* out_sc, const void* in, const float* in_sc, void* tmp_e,
           float* tmp_sc, const int* skip, const int* flags, int skip_value, int bmode = 0) {
    const int lk = prec == 0 ? e->lk64 : e->lks;
    lym::DslashParams a = base(e, prec);
    a.skip = skip;
    a.lane_flags = flags;
    a.lane_skip_value = skip_value;
    lym::DslashParams b = a;
    a.parity = 0;
    a.in[1] = in;
    a.in_sc[1] = in_sc;
    a.out[0] = tmp_e;
    a.out_sc[0] = tmp_sc;
    lym::launch_dslash(a, lk, 0, true, e->ctx->stream);
    b.parity = 1;
    b.in[0] = tmp_e;
    b.in_sc[0] = tmp_sc;
    b.diag_in[1] = in;
    b.diag_sc[1] = in_sc;
    b.out[1] = out;
    b.out_sc[1] = out_sc;
    b.hop_sign = -1.0;
    b.gamma5 = 1;
    if (bmode) {
        b.bmode = bmode;
        b.bsite[1] = e->bsite;
        if (bmode == 2) {
            b.balpha = &e->Sd->alpha[0];
            b.upd_p[1] = e->sv[0];
            b.upd_p_sc[1] = e->svsc[0];
            b.upd_xs[1] = e->xs;
            b.upd_rs[1] = e->rs;
        }
    }
    lym::launch_dslash(b, lk, e->csw != 0.0 ? 2 : 1, false, e->ctx->stream);
}

// Qhat^2 at precision prec: out = Qhat (Qhat in)
void schur2(LaneEngine* e, int prec, void* out, float* out_sc, const void* in, const float* in_sc, const int* skip,
            const int* flags, int skip_value) {
    if (prec == 0) {
        schur(e, 0, e->mid, nullptr, in, nullptr, e->tmp, nullptr, skip, flags, skip_value);
        schur(e, 0, out, nullptr, e->mid, nullptr, e->tmp, nullptr, skip, flags, skip_value);
    } else {
        schur(e, prec, e->sv[2], e->svsc[2], in, in_sc, e->sv[3], e->svsc[3], skip, flags, skip_value);
        schur(e, prec, out, out_sc, e->sv[2], e->svsc[2], e->sv[3], e->svsc[3], skip, flags, skip_value);
    }
}

lym::BcgArgs bcg_args(LaneEngine* e, const int* mask) {
    lym::BcgArgs a;
    a.x = e->x;
    a.r = e->r;
    a.b = e->rhs;
    a.Ax = e->Ax;
    a.xs = e->xs;
    a.rs = e->rs;
    a.p = e->sv[0];
    a.p_sc = e->svsc[0];
    a.Ap = e->sv[1];
    a.Ap_sc = e->svsc[1];
    a.nsites = e->S;
    a.stride = (int)e->S;
    a.max_steps = (int)e->max_steps;
    a.partials = e->partials;
    a.mask = mask;
    a.bsite = e->bsite;
    a.vh = e->ctx->g.Vh;
    return a;
}

int read_state(LaneEngine* e) {
    ECHECK(hipMemcpyAsync(e->Sh, e->Sd, sizeof(lym::BState), hipMemcpyDeviceToHost, e->ctx->stream));
    ECHECK(hipStreamSynchronize(e->ctx->stream));
    return LYM_OK;
}

int write_state(LaneEngine* e) {
    ECHECK(hipMemcpyAsync(e->Sd, e->Sh, sizeof(lym::BState), hipMemcpyHostToDevice, e->ctx->stream));
    return LYM_OK;
}

// Relayout of every field that carries lane state from width e->width to w
// (X, a scratch field of wmax lanes, holds each new layout before it is copied back).
int set_width(LaneEngine* e, int w) {
    const int w0 = e->width;
    if (w == w0) return LYM_OK;
    hipStream_t st = e->ctx->stream;
    const long S = e->S;
    const int G = e->prec == LYM_PREC_SINGLE ? 6 : 3;  // 16-byte vectors per site of a packed sloppy spinor
    struct F {
        void* p;
        int esize;
        long rows;
    } fields[] = {
        {e->B, 16, 24 * S},  {e->rhs, 16, 12 * S}, {e->x, 16, 12 * S},  {e->r, 16, 12 * S},
        {e->xs, 8, 12 * S},  {e->rs, 8, 12 * S},   {e->sv[0], 16, G * S}, {e->svsc[0], 4, S},
    };
    for (const F& f : fields) {
        if (w0 > 0) {
            lym::launch_relayout(f.esize, f.p, e->X, f.rows, w0, w, st);
            ECHECK(hipMemcpyAsync(f.p, e->X, (size_t)f.rows * w * f.esize, hipMemcpyDeviceToDevice, st));
        }
    }
    ECHECK(hipGetLastError());
    // lanes beyond the old width hold no request
    for (int q = w0; q < w; ++q) {
        e->Sh->done[q] = 1;
        e->Sh->conv[q] = 0;
        e->Sh->rel[q] = 0;
    }
    e->width = w;
    // padding slots of the new layout to zero (no launch writes them)
    const int vh = e->ctx->g.Vh;
    zd* full[] = {e->B, e->X};
    zd* half[] = {e->x, e->r, e->Ax, e->mid, e->tmp, e->rhs, e->src, e->te};
    for (zd* f : full) lym::launch_zero_padding(16, f, 24, (int)S, vh, w, st);
    for (zd* f : half) lym::launch_zero_padding(16, f, 12, (int)S, vh, w, st);
    for (int i = 0; i < 4; ++i) {
        lym::launch_zero_padding(16, e->sv[i], G, (int)S, vh, w, st);
        lym::launch_zero_padding(4, e->svsc[i], 1, (int)S, vh, w, st);
    }
    lym::launch_zero_padding(8, e->xs, 12, (int)S, vh, w, st);
    lym::launch_zero_padding(8, e->rs, 12, (int)S, vh, w, st);
    lym::launch_zero_padding(8, e->bsite, 1, (int)S, vh, w, st);
    ECHECK(hipGetLastError());
    return write_state(e);
}

void iteration(LaneEngine* e, int it) {
    hipStream_t st = e->ctx->stream;
    lym::BState* Sd = e->Sd;
    lym::BcgArgs a = bcg_args(e, nullptr);
    const int nr = e->width, prec = e->prec;
    // lanes that are done skip the sloppy apply; only lanes with a due update take the FP64 one
    if (e->fused && prec >= 1) {
        // <p, A p> = |Qhat p|^2 and the update fused into the stencil launches
        // (as the batched solver's even-odd batches do)
        schur(e, prec, e->sv[2], e->svsc[2], a.p, a.p_sc, e->sv[3], e->svsc[3], &Sd->all_done, Sd->done, 1, 1);
        lym::launch_bcg_step(7, prec, nr, Sd, a, st);
        schur(e, prec, nullptr, nullptr, e->sv[2], e->svsc[2], e->sv[3], e->svsc[3], &Sd->all_done, Sd->done, 1, 2);
        lym::launch_bcg_step(8, prec, nr, Sd, a, st);
    } else {
        schur2(e, prec, a.Ap, a.Ap_sc, a.p, a.p_sc, &Sd->all_done, Sd->done, 1);
        lym::launch_bcg_step(1, prec, nr, Sd, a, st);
        lym::launch_bcg_step(2, prec, nr, Sd, a, st);
    }
    lym::launch_bcg_step(5, prec, nr, Sd, a, st);
    if (rel_slot(it)) {
        lym::launch_bcg_step(3, prec, nr, Sd, a, st);
        schur2(e, 0, e->Ax, nullptr, e->x, nullptr, &Sd->skip_rel, Sd->rel, 0);
        lym::launch_bcg_step(4, prec, nr, Sd, a, st);
    }
}

bool graphs_enabled() {
    const char* v = std::getenv("LYM_GRAPHS");
    return !(v && v[0] == '0');
}

}  // namespace

size_t engine_bytes(const int dims[4], int rep, int storage, int prec, int wmax) {
    const double V = (double)dims[0] * dims[1] * dims[2] * dims[3];
    const double S = std::ceil(V / 2 / 64) * 64;
    const double es = prec == LYM_PREC_SINGLE ? 8 : 4;
    const double le = rep == LYM_REP_SU2_ADJ ? es / 2 : es;
    (void)storage;
    const double W = next_width(wmax);
    double b = 0;
    b += 2 * 24 * S * W * 16 + 8 * 12 * S * W * 16;  // FP64 nr-fields
    b += 2 * 12 * S * W * 8 + S * W * 8;             // x_s, r_s, the fused steps' terms
    b += 4 * (12 * S * W * es + S * W * 4);          // sloppy nr-fields
    b += W * (72 * S * 16 + 36 * S * 16 + 72 * S * es + 36 * S * es + 3 * S * 4);  // operator data per lane
    b += 72 * S * le;                                // low-precision links
    b += 2 * W * 24 * S * 16 + 12 * V * 16;          // transfer spinors, context staging
    return (size_t)(b * 1.03) + (32u << 20);
}

int ctx_create_on_stream(const int dims[4], int rep, int bc_t, void* stream, lym_ctx** out) {
    int rc = lym_ctx_create(dims, rep, bc_t, out);
    if (rc) return rc;
    lym_ctx* c = *out;
    (void)hipStreamSynchronize(c->stream);
    (void)hipStreamDestroy(c->stream);
    c->stream = reinterpret_cast<hipStream_t>(stream);
    c->owns_stream = 0;
    return LYM_OK;
}

int engine_create(lym_ctx* ctx, const lym_gauge* g, double csw, int prec, unsigned max_steps, int wmax,
                  LaneEngine** out) {
    if (!ctx || !g || !out) return fail(LYM_ERR_INVALID, "null argument");
    *out = nullptr;
    if (prec != LYM_PREC_SINGLE && prec != LYM_PREC_HALF) return fail(LYM_ERR_INVALID, "sloppy precision");
    if (wmax < 1 || wmax > lym::BMAX) return fail(LYM_ERR_INVALID, "lanes per batch must be 1..%d", lym::BMAX);
    for (int i = 0; i < 4; ++i)
        if (ctx->dims[i] % 2) return fail(LYM_ERR_UNSUPPORTED, "the serving engine needs even lattice extents");
    if (!g->B) return fail(LYM_ERR_INVALID, "the lattice's clover blocks are not built");
    LaneEngine* e = new LaneEngine();
    e->ctx = ctx;
    e->g = g;
    e->csw = csw;
    e->prec = prec;
    e->wmax = wmax;
    e->max_steps = max_steps;
    e->S = ctx->g.stride;
    e->es = prec == LYM_PREC_SINGLE ? 8 : 4;
    e->lk64 = link_kind_of(g);
    e->wcap = next_width(wmax);
    const long S = e->S;
    const size_t W = (size_t)e->wcap;
    int rc = LYM_OK;
    auto A = [&](void** p, size_t bytes) {
        if (!rc) rc = ealloc(e, p, bytes);
    };
    A((void**)&e->kappa, sizeof(double) * lym::BMAX);
    A((void**)&e->A, sizeof(zd) * 72 * S * W);
    A((void**)&e->Ainv, sizeof(zd) * 36 * S * W);
    A(&e->As, 72 * S * W * e->es);
    A((void**)&e->As_sc, sizeof(float) * 2 * S * W);
    A(&e->Ainvs, 36 * S * W * e->es);
    A((void**)&e->Ainvs_sc, sizeof(float) * S * W);
    // low-precision links: all three rows when the FP64 field has them
    // (every lane would otherwise rebuild row 2 of the same link)
    const size_t le = g->real_links ? e->es / 2 : e->es;
    A(&e->Us, 72 * S * le);
    for (zd** f : {&e->B, &e->X}) A((void**)f, sizeof(zd) * 24 * S * W);
    for (zd** f : {&e->x, &e->r, &e->Ax, &e->mid, &e->tmp, &e->rhs, &e->src, &e->te})
        A((void**)f, sizeof(zd) * 12 * S * W);
    A((void**)&e->xs, sizeof(lym::cplx<float>) * 12 * S * W);
    A((void**)&e->rs, sizeof(lym::cplx<float>) * 12 * S * W);
    for (int i = 0; i < 4; ++i) {
        A(&e->sv[i], 12 * S * W * e->es);
        A((void**)&e->svsc[i], sizeof(float) * S * W);
    }
    const size_t nch = ((size_t)S + lym::BATCH_CHUNK - 1) / lym::BATCH_CHUNK;
    A((void**)&e->partials, sizeof(double) * lym::BMAX * nch);
    A((void**)&e->bsite, sizeof(double) * S * W);
    {
        const char* fb = std::getenv("LYM_FUSED_BATCH");
        e->fused = !(fb && fb[0] == '0');
    }
    A((void**)&e->Sd, sizeof(lym::BState));
    A((void**)&e->mask_d, sizeof(int) * lym::BMAX);
    A((void**)&e->flag_d, sizeof(int) * 2);
    A((void**)&e->lset_d, sizeof(int) * lym::BMAX);
    A((void**)&e->ptr_d, sizeof(zd*) * lym::BMAX);
    if (!rc && (hipHostMalloc((void**)&e->Sh, sizeof(lym::BState), hipHostMallocDefault) != hipSuccess ||
                hipHostMalloc((void**)&e->mask_h, sizeof(int) * lym::BMAX, hipHostMallocDefault) != hipSuccess ||
                hipHostMalloc((void**)&e->flag_h, sizeof(int) * 2, hipHostMallocDefault) != hipSuccess ||
                hipHostMalloc((void**)&e->ptr_h, sizeof(zd*) * lym::BMAX, hipHostMallocDefault) != hipSuccess ||
                hipHostMalloc((void**)&e->lset_h, sizeof(int) * lym::BMAX, hipHostMallocDefault) != hipSuccess))
        rc = fail(LYM_ERR_DEVICE, "hipHostMalloc");
    for (int q = 0; !rc && q < wmax; ++q) {
        lym_spinor *a = nullptr, *b = nullptr;
        rc = lym_spinor_create(ctx, &a);
        if (!rc) e->lsrc.push_back(a);
        if (!rc) rc = lym_spinor_create(ctx, &b);
        if (!rc) e->lsol.push_back(b);
    }
    if (!rc) {
        // the context's transfer staging, allocated now rather than in the first request
        std::vector<double> host((size_t)ctx->V * 24);
        rc = lym_spinor_download(e->lsrc[0], host.data());
    }
    if (!rc) {
        const int ncomp = g->real_links ? g->compression : g->compression / 2;
        if (full_rows(g)) {
            lym::launch_convert_links(prec, g->U, e->Us, 2L * 4 * ncomp * S, g->real_links, ctx->stream);
            e->lks = e->lk64;
        } else {
            // 12/6-real storage: rows 0 and 1 packed, row 2 rebuilt in the stencil
            lym::launch_pack_links(prec, g->U, ncomp, g->real_links, e->Us, (int)S, ctx->stream);
            e->lks = g->real_links ? 3 : 1;
        }
        if (hipGetLastError() != hipSuccess) rc = fail(LYM_ERR_DEVICE, "link conversion");
    }
    if (!rc) {
        std::memset(e->Sh, 0, sizeof(lym::BState));
        for (int q = 0; q < lym::BMAX; ++q) e->Sh->done[q] = 1;
        for (int q = 0; q < lym::BMAX; ++q) e->lset_h[q] = q < e->wcap ? q : 0;
        if (hipMemcpyAsync(e->lset_d, e->lset_h, sizeof(int) * lym::BMAX, hipMemcpyHostToDevice, ctx->stream) !=
            hipSuccess)
            rc = fail(LYM_ERR_DEVICE, "lane block sets");
        e->Sh->all_done = 1;
        e->Sh->skip_rel = 1;
        e->Sh->max_steps = (int)max_steps;
        rc = write_state(e);
    }
    if (!rc && hipStreamSynchronize(ctx->stream) != hipSuccess) rc = fail(LYM_ERR_DEVICE, "engine set-up");
    if (rc) {
        engine_destroy(e);
        return rc;
    }
    *out = e;
    return LYM_OK;
}

void engine_destroy(LaneEngine* e) {
    if (!e) return;
    if (e->ctx && e->ctx->stream) (void)hipStreamSynchronize(e->ctx->stream);
    for (int i = 0; i < 8; ++i)
        if (e->graph[i]) (void)hipGraphExecDestroy(e->graph[i]);
    for (auto* s : e->lsrc) lym_spinor_destroy(s);
    for (auto* s : e->lsol) lym_spinor_destroy(s);
    for (void* p : e->allocs) (void)hipFree(p);
    if (e->Sh) (void)hipHostFree(e->Sh);
    if (e->mask_h) (void)hipHostFree(e->mask_h);
    if (e->flag_h) (void)hipHostFree(e->flag_h);
    if (e->ptr_h) (void)hipHostFree(e->ptr_h);
    if (e->lset_h) (void)hipHostFree(e->lset_h);
    delete e;
}

int engine_busy(const LaneEngine* e) {
    int n = 0;
    for (int q = 0; q < e->wmax; ++q) n += e->lanes[q].busy;
    return n;
}

int engine_room(const LaneEngine* e) { return e->wmax - engine_busy(e); }

// block set q: A = 1 + kappa csw gamma5 B, its even-site inverse and the
// low-precision copies
static int build_set(LaneEngine* e, int q, double kappa) {
    lym_ctx* c = e->ctx;
    hipStream_t st = c->stream;
    const long S = e->S;
    zd* Aq = e->A + 72 * S * q;
    zd* Iq = e->Ainv + 36 * S * q;
    lym::launch_clover_A(c->g, reinterpret_cast<const zd*>(e->g->B), kappa * e->csw, Aq, st);
    // a unit block for every lane stays exact at csw = 0 (B is then ignored)
    ECHECK(hipMemsetAsync(e->flag_d, 0, 2 * sizeof(int), st));
    lym::launch_clover_inv_ldl(c->g, Aq, Iq, e->flag_d + 1, st);
    ECHECK(hipMemcpyAsync(e->flag_h, e->flag_d, 2 * sizeof(int), hipMemcpyDeviceToHost, st));
    ECHECK(hipStreamSynchronize(st));
    if (e->flag_h[1]) {
        lym::launch_clover_inv(c->g, Aq, Iq, e->flag_d, st);
        ECHECK(hipMemcpyAsync(e->flag_h, e->flag_d, sizeof(int), hipMemcpyDeviceToHost, st));
        ECHECK(hipStreamSynchronize(st));
    }
    if (e->flag_h[0]) return fail(LYM_ERR_SINGULAR, "%d clover blocks are singular at kappa %g", e->flag_h[0], kappa);
    lym::launch_convert_blocks(e->prec, Aq, (char*)e->As + 72 * S * e->es * q, e->As_sc + 2 * S * q, 2, (int)S, st);
    lym::launch_convert_blocks(e->prec, Iq, (char*)e->Ainvs + 36 * S * e->es * q, e->Ainvs_sc + S * q, 1, (int)S, st);
    ECHECK(hipGetLastError());
    return LYM_OK;
}

int engine_admit(LaneEngine* e, int n, const LaneRequest* reqs, std::vector<LaneResult>* failed) {
    if (n <= 0) return LYM_OK;
    if (n > engine_room(e)) return fail(LYM_ERR_INVALID, "no room for %d requests", n);
    lym_ctx* c = e->ctx;
    hipStream_t st = c->stream;
    const long S = e->S;
    // lanes: the lowest free ones; widen the batch to hold them
    std::vector<int> lane_of;
    int hi = 0;
    for (int q = 0; q < e->wmax; ++q)
        if (e->lanes[q].busy) hi = q + 1;
    for (int q = 0; q < e->wmax && (int)lane_of.size() < n; ++q)
        if (!e->lanes[q].busy) lane_of.push_back(q);
    for (int q : lane_of) hi = std::max(hi, q + 1);
    int rc = set_width(e, std::max(e->width, next_width(hi)));
    if (rc) return rc;
    const int w = e->width;
    // per request: source upload, |b|^2, the lane's operator data
    std::vector<double> eps(n);
    std::memset(e->mask_h, 0, sizeof(int) * lym::BMAX);
    int admitted = 0;
    for (int i = 0; i < n; ++i) {
        const int q = lane_of[i];
        double b2 = 0;
        rc = reqs[i].src_dev ? lym_spinor_copy(e->lsrc[q], reqs[i].src_dev) : lym_spinor_upload(e->lsrc[q], reqs[i].src);
        if (!rc) rc = lym_spinor_squared_norm(e->lsrc[q], &b2);
        if (rc) return rc;
        // the lane's kappa, and its block set: one already built for this
        // kappa, else a set no lane uses (an empty one first, else the least
        // recently used); there are wcap sets and at most wmax <= wcap lanes
        lym::launch_set_double(e->kappa + q, reqs[i].kappa, st);
        int bs = -1;
        for (int s = 0; s < e->wcap && bs < 0; ++s)
            if (e->set_valid[s] && e->set_kappa[s] == reqs[i].kappa) bs = s;
        if (bs < 0) {
            for (int s = 0; s < e->wcap && bs < 0; ++s)
                if (!e->set_valid[s] && e->set_refs[s] == 0) bs = s;
            for (int s = 0; s < e->wcap; ++s)
                if (e->set_refs[s] == 0 && (bs < 0 || (e->set_valid[bs] && e->set_used[s] < e->set_used[bs])))
                    bs = s;
            if (bs < 0) return fail(LYM_ERR_DEVICE, "serving engine: no free block set");
            e->set_valid[bs] = false;
            const int orc = build_set(e, bs, reqs[i].kappa);
            if (orc == LYM_ERR_SINGULAR) {
                LaneResult f{};
                f.tag = reqs[i].tag;
                f.error = orc;
                failed->push_back(f);
                continue;
            }
            if (orc) return orc;
            e->set_valid[bs] = true;
            e->set_kappa[bs] = reqs[i].kappa;
        }
        ++e->set_refs[bs];
        e->set_used[bs] = ++e->set_tick;
        e->lset_h[q] = bs;
        const double t = 0.1 * reqs[i].tol;
        eps[i] = reqs[i].eps > 0 ? reqs[i].eps : t * t * b2;
        e->lanes[q].busy = true;
        e->lanes[q].req = reqs[i];
        e->lanes[q].max_lanes = 0;
        e->mask_h[q] = 1;
        e->Sh->eps[q] = eps[i];
        ++admitted;
    }
    if (!admitted) return LYM_OK;
    e->Sh->delta2 = 0.05 * 0.05;
    e->Sh->max_steps = (int)e->max_steps;
    rc = write_state(e);
    if (rc) return rc;
    ECHECK(hipMemcpyAsync(e->mask_d, e->mask_h, sizeof(int) * lym::BMAX, hipMemcpyHostToDevice, st));
    ECHECK(hipMemcpyAsync(e->lset_d, e->lset_h, sizeof(int) * lym::BMAX, hipMemcpyHostToDevice, st));
    // pack the admitted sources into their lanes of B
    for (int q = 0; q < w; ++q) e->ptr_h[q] = e->mask_h[q] ? reinterpret_cast<zd*>(e->lsrc[q]->d) : nullptr;
    ECHECK(hipMemcpyAsync(e->ptr_d, e->ptr_h, sizeof(zd*) * w, hipMemcpyHostToDevice, st));
    lym::launch_pack(e->ptr_d, w, e->B, 24 * S, st);
    // the even-odd driver's right-hand side (PreconditionedBiCGStab.cpp:56-70):
    // src_o = b_o - D_oe R_ee^-1 b_e, rhs = Qhat gamma5 src_o
    const zd* b_e = e->B;
    const zd* b_o = e->B + 12 * S * w;
    lym::launch_clover_apply_lanes(c->g, e->Ainv, 36 * S, e->lset_d, e->mask_d, b_e, e->te, w, st);
    lym::DslashParams a = base(e, 0);
    a.hop_sign = -1.0;
    a.parity = 1;
    a.in[0] = e->te;
    a.diag_in[1] = b_o;
    a.out[1] = e->src;
    a.lane_flags = e->mask_d;
    a.lane_skip_value = 0;
    lym::launch_dslash(a, e->lk64, 1, false, st);
    lym::launch_scale_g5(e->src, 1, (int)(S * w), st);  // lanes outside the mask hold nothing in use
    schur(e, 0, e->rhs, nullptr, e->src, nullptr, e->tmp, nullptr, nullptr, e->mask_d, 0);
    // FP64 start: x = rhs, r = rhs - Qhat^2 x, then the lanes' CG state
    lym::launch_lane_copy(16, e->x, e->rhs, 12 * S, w, e->mask_d, st);
    schur2(e, 0, e->Ax, nullptr, e->x, nullptr, nullptr, e->mask_d, 0);
    lym::launch_lane_sub(e->r, e->rhs, e->Ax, 12 * S, w, e->mask_d, st);
    lym::BcgArgs ba = bcg_args(e, e->mask_d);
    if (!lym::launch_bcg_step(0, e->prec, w, e->Sd, ba, st)) return fail(LYM_ERR_UNSUPPORTED, "width %d", w);
    ECHECK(hipGetLastError());
    return read_state(e);
}

int engine_run_chunk(LaneEngine* e) {
    if (e->width == 0 || e->Sh->all_done) return LYM_OK;
    // the most requests in flight together, per lane
    const int busy = engine_busy(e);
    for (int q = 0; q < e->wmax; ++q)
        if (e->lanes[q].busy) e->lanes[q].max_lanes = std::max(e->lanes[q].max_lanes, (unsigned)busy);
    hipStream_t st = e->ctx->stream;
    if (!graphs_enabled()) {
        for (int it = 0; it < CHUNK; ++it) iteration(e, it);
    } else {
        int slot = -1;
        for (int i = 0; i < 8; ++i)
            if (e->graph[i] && e->graph_w[i] == e->width) slot = i;
        if (slot < 0) {
            for (int i = 0; i < 8 && slot < 0; ++i)
                if (!e->graph[i]) slot = i;
            if (slot < 0) return fail(LYM_ERR_DEVICE, "graph cache full");
            ECHECK(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
            for (int it = 0; it < CHUNK; ++it) iteration(e, it);
            hipGraph_t gr = nullptr;
            hipError_t err = hipStreamEndCapture(st, &gr);
            if (err != hipSuccess || !gr) return fail(LYM_ERR_DEVICE, "graph capture: %s", hipGetErrorString(err));
            err = hipGraphInstantiate(&e->graph[slot], gr, nullptr, nullptr, 0);
            (void)hipGraphDestroy(gr);
            if (err != hipSuccess) {
                e->graph[slot] = nullptr;
                return fail(LYM_ERR_DEVICE, "graph instantiate: %s", hipGetErrorString(err));
            }
            e->graph_w[slot] = e->width;
        }
        ECHECK(hipGraphLaunch(e->graph[slot], st));
    }
    ECHECK(hipGetLastError());
    return read_state(e);
}

int engine_harvest(LaneEngine* e, std::vector<LaneResult>* done) {
    lym_ctx* c = e->ctx;
    hipStream_t st = c->stream;
    const long S = e->S;
    const int w = e->width;
    std::memset(e->mask_h, 0, sizeof(int) * lym::BMAX);
    int n = 0;
    for (int q = 0; q < w; ++q)
        if (e->lanes[q].busy && e->Sh->done[q]) {
            e->mask_h[q] = 1;
            ++n;
        }
    if (!n) return LYM_OK;
    ECHECK(hipMemcpyAsync(e->mask_d, e->mask_h, sizeof(int) * lym::BMAX, hipMemcpyHostToDevice, st));
    // lanes that stopped at max steps between reliable updates: x += x_s
    lym::BcgArgs ba = bcg_args(e, e->mask_d);
    lym::launch_bcg_step(6, e->prec, w, e->Sd, ba, st);
    // x_o, and x_e = R_ee^-1 (b_e - D_eo x_o) (PreconditionedBiCGStab.cpp:80-86)
    zd* X_e = e->X;
    zd* X_o = e->X + 12 * S * w;
    lym::launch_lane_copy(16, X_o, e->x, 12 * S, w, e->mask_d, st);
    lym::DslashParams a = base(e, 0);
    a.hop_sign = -1.0;
    a.parity = 0;
    a.in[1] = X_o;
    a.diag_in[0] = e->B;
    a.out[0] = X_e;
    a.lane_flags = e->mask_d;
    a.lane_skip_value = 0;
    lym::launch_dslash(a, e->lk64, 1, true, st);
    for (int q = 0; q < w; ++q) e->ptr_h[q] = e->mask_h[q] ? reinterpret_cast<zd*>(e->lsol[q]->d) : nullptr;
    ECHECK(hipMemcpyAsync(e->ptr_d, e->ptr_h, sizeof(zd*) * w, hipMemcpyHostToDevice, st));
    lym::launch_unpack(e->X, w, e->ptr_d, 24 * S, st);
    ECHECK(hipGetLastError());
    int hi = 0;
    for (int q = 0; q < w; ++q) {
        if (!e->mask_h[q]) {
            if (e->lanes[q].busy) hi = q + 1;
            continue;
        }
        Lane& L = e->lanes[q];
        int rc = L.req.dst_dev ? lym_spinor_copy(L.req.dst_dev, e->lsol[q]) : lym_spinor_download(e->lsol[q], L.req.dst);
        if (rc) return rc;
        LaneResult res{};
        res.tag = L.req.tag;
        res.error = LYM_OK;
        res.max_lanes = std::max(L.max_lanes, 1u);
        const lym::BState* H = e->Sh;
        res.info.converged = H->conv[q];
        res.info.steps = H->conv[q] ? (unsigned)H->step[q] : e->max_steps;
        res.info.last_error = H->rr_true[q];
        res.info.reliable_updates = (unsigned)H->nrel[q];
        res.info.applies = 1 + (unsigned)H->step[q] + (unsigned)H->nrel[q];
        done->push_back(res);
        L.busy = false;
        --e->set_refs[e->lset_h[q]];
    }
    // narrow the batch when its busy lanes fit a smaller width
    if (hi == 0) {
        e->width = 0;  // empty: the next admission starts a fresh layout
        return LYM_OK;
    }
    if (next_width(hi) < w) return set_width(e, next_width(hi));
    return LYM_OK;
}

}  // namespace lym_internal
