// api.cpp -- C ABI of the MI355X fermion-solve library (declared in include/lym.h).
//
// Objects mirror the reference's: a context holds the lattice geometry and one
// HIP stream (default priority); a gauge holds the device links shared by any
// number of operators; an operator holds kappa, csw, gamma5, power and its
// clover blocks; a solver holds its work vectors and a device state block.
#include <hip/hip_runtime.h>

#include <atomic>
#include <chrono>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>

#include "../../include/lym.h"
#include "lym_kernels.h"

using lym::zd;

namespace lym_internal {

void forget_op_captures(const lym_op* op);  // solvers.cpp

thread_local std::string g_err;

int fail(int code, const char* fmt, ...) {
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof(buf), fmt, ap);
    va_end(ap);
    g_err = buf;
    return code;
}

}  // namespace lym_internal

using lym_internal::fail;
using lym_internal::g_err;

namespace {

#define HIPCHECK(expr)                                                                      \
    do {                                                                                    \
        hipError_t e_ = (expr);                                                             \
        if (e_ != hipSuccess) return fail(LYM_ERR_DEVICE, "%s: %s", #expr, hipGetErrorString(e_)); \
    } while (0)

#define CHECK_LAUNCH()                                                                        \
    do {                                                                                      \
        hipError_t e_ = hipGetLastError();                                                    \
        if (e_ != hipSuccess) return fail(LYM_ERR_DEVICE, "kernel launch: %s", hipGetErrorString(e_)); \
    } while (0)

inline long half_n(const lym_ctx* c) { return 12L * c->g.stride; }
inline long full_n(const lym_ctx* c) { return 24L * c->g.stride; }

// Allocates and zeroes on the context's stream: every later use of the buffer
// is on that stream too, so the zero fill is ordered before it (a plain
// hipMemset runs on the null stream, which a non-blocking stream does not
// wait for).
int dmalloc(lym_ctx* c, void** p, size_t bytes) {
    hipError_t e = hipMalloc(p, bytes);
    if (e != hipSuccess) {
        *p = nullptr;
        return fail(LYM_ERR_NOMEM, "hipMalloc(%zu bytes): %s", bytes, hipGetErrorString(e));
    }
    e = hipMemsetAsync(*p, 0, bytes, c->stream);
    if (e != hipSuccess) return fail(LYM_ERR_DEVICE, "hipMemsetAsync: %s", hipGetErrorString(e));
    return LYM_OK;
}

int link_kind(const lym_gauge* g) {
    if (g->real_links) return g->compression == 6 ? 3 : 2;
    return g->compression == 12 ? 1 : 0;
}

int ensure_staging(lym_ctx* c, size_t bytes) {
    if (c->staging_bytes >= bytes) return LYM_OK;
    if (c->staging) (void)hipFree(c->staging);
    c->staging = nullptr;
    c->staging_bytes = 0;
    int rc = dmalloc(c, &c->staging, bytes);
    if (rc) return rc;
    c->staging_bytes = bytes;
    return LYM_OK;
}

// ---------------------------------------------------------------- operator internals
bool op_has_clover(const lym_op* op) {
    return (op->kind == LYM_OP_CLOVER || op->kind == LYM_OP_EO_CLOVER) && op->csw != 0.0;
}

// builds B (gauge), A and Ainv (op) when stale; force_blocks also for csw = 0
int op_prepare(lym_op* op, bool force_blocks) {
    lym_ctx* c = op->ctx;
    if (!op->gauge) return fail(LYM_ERR_INVALID, "operator has no gauge field (call lym_op_set_gauge)");
    bool need_blocks = force_blocks || op_has_clover(op) || op->kind == LYM_OP_EO_CLOVER;
    if (!need_blocks || (!op->diag_dirty && op->A)) return LYM_OK;
    lym_gauge* g = const_cast<lym_gauge*>(op->gauge);
    const size_t blk = sizeof(zd) * 36 * (size_t)c->g.stride;
    if (!g->B) {
        int rc = dmalloc(c, (void**)&g->B, 2 * blk);
        if (rc) return rc;
        lym::launch_clover_B(c->g, g->U, link_kind(g), reinterpret_cast<zd*>(g->B), c->stream);
        CHECK_LAUNCH();
    }
    if (!op->A) {
        int rc = dmalloc(c, (void**)&op->A, 2 * blk);
        if (rc) return rc;
    }
    if (!op->Ainv) {
        int rc = dmalloc(c, (void**)&op->Ainv, blk);
        if (rc) return rc;
    }
    lym::launch_clover_A(c->g, reinterpret_cast<const zd*>(g->B), op->kappa * op->csw, reinterpret_cast<zd*>(op->A),
                         c->stream);
    CHECK_LAUNCH();
    // unpivoted LDL^H inverse first; the pivoted Gauss-Jordan redoes the parity
    // if any block had a small pivot, and reports singular blocks
    int* flag = reinterpret_cast<int*>(c->red_result) + 32;
    HIPCHECK(hipMemsetAsync(flag, 0, 2 * sizeof(int), c->stream));
    lym::launch_clover_inv_ldl(c->g, reinterpret_cast<const zd*>(op->A), reinterpret_cast<zd*>(op->Ainv), flag + 1,
                               c->stream);
    CHECK_LAUNCH();
    int hflag[2] = {0, 0};
    HIPCHECK(hipMemcpyAsync(hflag, flag, 2 * sizeof(int), hipMemcpyDeviceToHost, c->stream));
    HIPCHECK(hipStreamSynchronize(c->stream));
    // LYM_CLOVER_INV_PIVOTED=1 always takes the pivoted path (lets tests cover it)
    const char* force = std::getenv("LYM_CLOVER_INV_PIVOTED");
    if (hflag[1] || (force && force[0] == '1')) {
        lym::launch_clover_inv(c->g, reinterpret_cast<const zd*>(op->A), reinterpret_cast<zd*>(op->Ainv), flag,
                               c->stream);
        CHECK_LAUNCH();
        HIPCHECK(hipMemcpyAsync(hflag, flag, sizeof(int), hipMemcpyDeviceToHost, c->stream));
        HIPCHECK(hipStreamSynchronize(c->stream));
    }
    if (hflag[0]) return fail(LYM_ERR_SINGULAR, "%d clover blocks are singular", hflag[0]);
    op->diag_dirty = 0;
    op->sloppy_prec = 0;
    return LYM_OK;
}

void free_sloppy(lym_op* op) {
    void* ptrs[6] = {op->Us, op->Usf, op->As, op->As_sc, op->Ainvs, op->Ainvs_sc};
    for (void* q : ptrs)
        if (q) (void)hipFree(q);
    op->Us = op->Usf = op->As = op->Ainvs = nullptr;
    op->As_sc = op->Ainvs_sc = nullptr;
    op->sloppy_prec = 0;
    op->sloppy_alloc = 0;
}

// all three rows of 18-real (9-real) links in low precision, for multi-RHS
// applies: there the NR lanes of a site would each rebuild row 2 of the same
// link, and the batched stencil is bound by arithmetic, not bandwidth
int op_sloppy_full_rows(lym_op* op, int prec) {
    lym_ctx* c = op->ctx;
    const lym_gauge* g = op->gauge;
    const int ncomp = g->real_links ? g->compression : g->compression / 2;
    if (!g->rows_rebuild || ncomp != 9 || op->Usf) return LYM_OK;  // Us already has every row, or Usf is current
    const long S = c->g.stride;
    const size_t le = g->real_links ? (prec == 1 ? 4 : 2) : (prec == 1 ? 8 : 4);
    int rc = dmalloc(c, &op->Usf, 2L * 4 * 9 * S * le);
    if (rc) return rc;
    lym::launch_convert_links(prec, g->U, op->Usf, 2L * 4 * 9 * S, g->real_links, c->stream);
    CHECK_LAUNCH();
    return LYM_OK;
}

// low-precision copies of links and blocks for a mixed-precision solve
// (full_rows: also the full-row links of multi-RHS applies)
int op_sloppy_data(lym_op* op, int prec, bool full_rows) {
    lym_ctx* c = op->ctx;
    if (prec != 1 && prec != 2) return fail(LYM_ERR_INVALID, "sloppy precision must be 1 (FP32) or 2 (FP16)");
    if (op->sloppy_prec == prec) return full_rows ? op_sloppy_full_rows(op, prec) : LYM_OK;
    // buffers stay allocated across kappa/csw/gauge changes (a free would
    // synchronise the device); only a change of precision reallocates
    if (op->sloppy_alloc != prec) free_sloppy(op);
    const lym_gauge* g = op->gauge;
    const long S = c->g.stride;
    const int ncomp = g->real_links ? g->compression : g->compression / 2;
    const long nl = 2L * 4 * ncomp * S;
    const size_t le = g->real_links ? (prec == 1 ? 4 : 2) : (prec == 1 ? 8 : 4);
    const size_t se = prec == 1 ? 8 : 4;  // spinor/block element
    int rc = LYM_OK;
    if (!op->Us) rc = dmalloc(c, &op->Us, 2L * 4 * 9 * S * le);  // room for 18-real links
    if (rc) return rc;
    op->sloppy_alloc = prec;
    if (g->rows_rebuild) {
        // rows 0 and 1 in 16/8/4-byte vectors, row 2 rebuilt in the stencil
        lym::launch_pack_links(prec, g->U, ncomp, g->real_links, op->Us, (int)S, c->stream);
        op->sloppy_lk = g->real_links ? 3 : 1;
    } else {
        lym::launch_convert_links(prec, g->U, op->Us, nl, g->real_links, c->stream);
        op->sloppy_lk = link_kind(g);
    }
    if (op->Usf) {
        if (g->rows_rebuild && ncomp == 9) {
            lym::launch_convert_links(prec, g->U, op->Usf, nl, g->real_links, c->stream);
        } else {
            // the new gauge field's Us already serves batches (compressed
            // storage, or rows not rebuilt): a stale full-row copy must not be used
            (void)hipFree(op->Usf);
            op->Usf = nullptr;
        }
    }
    if (op->A) {
        if (!op->As) rc = dmalloc(c, &op->As, 2L * 36 * S * se);
        if (!rc && !op->As_sc) rc = dmalloc(c, (void**)&op->As_sc, 2L * S * sizeof(float));
        if (!rc && !op->Ainvs) rc = dmalloc(c, &op->Ainvs, 36L * S * se);
        if (!rc && !op->Ainvs_sc) rc = dmalloc(c, (void**)&op->Ainvs_sc, S * sizeof(float));
        if (rc) return rc;
        lym::launch_convert_blocks(prec, reinterpret_cast<const zd*>(op->A), op->As, op->As_sc, 2, (int)S, c->stream);
        lym::launch_convert_blocks(prec, reinterpret_cast<const zd*>(op->Ainv), op->Ainvs, op->Ainvs_sc, 1, (int)S,
                                   c->stream);
    }
    CHECK_LAUNCH();
    op->sloppy_prec = prec;
    return full_rows ? op_sloppy_full_rows(op, prec) : LYM_OK;
}

lym::DslashParams base_params(const lym_op* op) {
    lym::DslashParams a;
    std::memset(&a, 0, sizeof(a));
    a.g = op->ctx->g;
    a.links = op->gauge->U;
    a.A = reinterpret_cast<const zd*>(op->A);
    a.Ainv = reinterpret_cast<const zd*>(op->Ainv);
    a.kappa = op->kappa;
    a.kappa_dev = op->dev_kappa;
    a.hop_sign = 1.0;
    a.alpha = lym::mk(0.0, 0.0);
    return a;
}

inline zd* par(zd* f, const lym_ctx* c, int p) { return f + (size_t)p * 12 * c->g.stride; }
inline const zd* par(const zd* f, const lym_ctx* c, int p) { return f + (size_t)p * 12 * c->g.stride; }

// full-lattice single power: out = Op in (+ alpha v2); Wilson or clover
void apply_full1(lym_op* op, zd* out, const zd* in, const zd* v2, zd alpha, int gamma5, const int* skip) {
    lym_ctx* c = op->ctx;
    lym::DslashParams a = base_params(op);
    for (int p = 0; p < 2; ++p) {
        a.in[p] = par(in, c, p);
        a.diag_in[p] = par(in, c, p);
        a.out[p] = par(out, c, p);
        a.v2[p] = v2 ? par(v2, c, p) : nullptr;
    }
    a.alpha = alpha;
    a.gamma5 = gamma5;
    a.parity = c->g.tsh ? -1 : 0;  // both parities, or the one class of a lattice with an odd extent
    a.skip = skip;
    lym::launch_dslash(a, link_kind(op->gauge), op_has_clover(op) ? 2 : 1, false, c->stream);
}

// Schur complement on odd half fields: out_o = gamma5^g (A_o in_o - D_oe Ainv_e D_eo in_o)
void apply_schur1(lym_op* op, zd* out_o, const zd* in_o, int gamma5, const int* skip) {
    lym_ctx* c = op->ctx;
    int lk = link_kind(op->gauge);
    lym::DslashParams a = base_params(op);
    a.skip = skip;
    a.parity = 0;
    a.in[1] = in_o;
    a.out[0] = reinterpret_cast<zd*>(op->tmp2);
    lym::launch_dslash(a, lk, 0, true, c->stream);
    lym::DslashParams b = base_params(op);
    b.skip = skip;
    b.parity = 1;
    b.in[0] = reinterpret_cast<zd*>(op->tmp2);
    b.diag_in[1] = in_o;
    b.out[1] = out_o;
    b.hop_sign = -1.0;
    b.gamma5 = gamma5;
    lym::launch_dslash(b, lk, op->csw != 0.0 ? 2 : 1, false, c->stream);
}

int ensure_op_scratch(lym_op* op) {
    lym_ctx* c = op->ctx;
    if (!op->tmp) {
        int rc = dmalloc(c, (void**)&op->tmp, sizeof(zd) * full_n(c));
        if (rc) return rc;
    }
    if (!op->tmp2) {
        int rc = dmalloc(c, (void**)&op->tmp2, sizeof(zd) * half_n(c));
        if (rc) return rc;
    }
    return LYM_OK;
}

// Op^power applied to a vector of the operator's natural domain:
// full fields for Wilson/clover, odd half fields for the even-odd operator.
void apply_power(lym_op* op, zd* out, const zd* in, const int* skip) {
    if (op->kind == LYM_OP_EO_CLOVER) {
        if (op->power == 1) {
            apply_schur1(op, out, in, op->gamma5, skip);
        } else {
            zd* t = reinterpret_cast<zd*>(op->tmp);
            apply_schur1(op, t, in, op->gamma5, skip);
            apply_schur1(op, out, t, op->gamma5, skip);
        }
        return;
    }
    if (op->power == 1) {
        apply_full1(op, out, in, nullptr, lym::mk(0.0, 0.0), op->gamma5, skip);
    } else {
        zd* t = reinterpret_cast<zd*>(op->tmp);
        apply_full1(op, t, in, nullptr, lym::mk(0.0, 0.0), op->gamma5, skip);
        apply_full1(op, out, t, nullptr, lym::mk(0.0, 0.0), op->gamma5, skip);
    }
}

int check_op_ready(lym_op* op) {
    if (!op) return fail(LYM_ERR_INVALID, "null operator");
    int rc = op_prepare(op, false);
    if (rc) return rc;
    return ensure_op_scratch(op);
}

}  // namespace

namespace lym_internal {
int op_ready(lym_op* op, bool force_blocks) {
    if (!op) return fail(LYM_ERR_INVALID, "null operator");
    int rc = op_prepare(op, force_blocks);
    if (rc) return rc;
    return ensure_op_scratch(op);
}
void apply_schur(lym_op* op, zd* out_o, const zd* in_o, int gamma5, const int* skip) {
    apply_schur1(op, out_o, in_o, gamma5, skip);
}
void apply_full(lym_op* op, zd* out, const zd* in, int gamma5, const int* skip) {
    apply_full1(op, out, in, nullptr, lym::mk(0.0, 0.0), gamma5, skip);
}
void apply_full_add(lym_op* op, zd* out, const zd* in, const zd* v2, zd alpha, int gamma5, const int* skip) {
    apply_full1(op, out, in, v2, alpha, gamma5, skip);
}
int link_kind_of(const lym_gauge* g) { return link_kind(g); }

int op_sloppy(lym_op* op, int prec, bool full_rows) { return op_sloppy_data(op, prec, full_rows); }

static size_t selem(int prec) { return prec == 0 ? 16 : (prec == 1 ? 8 : 4); }

// operator data of a storage precision: 0 FP64 (the operator's own), 1/2 the
// sloppy copies (batch_links: the full-row links of the batched solver when
// there are such, for every batch width, so an RHS gives the same bits in a
// batch of any size); returns the link kind
static int set_op_data(lym::DslashParams& a, const lym_op* op, int prec, bool batch_links) {
    a.prec = prec;
    if (prec == 0) {
        a.links = op->gauge->U;
        a.A = op->A;
        a.Ainv = op->Ainv;
        return link_kind(op->gauge);
    }
    const bool full = batch_links && op->Usf;
    a.links = full ? op->Usf : op->Us;
    a.A = op->As;
    a.A_sc = op->As_sc;
    a.Ainv = op->Ainvs;
    a.Ainv_sc = op->Ainvs_sc;
    return full ? link_kind(op->gauge) : op->sloppy_lk;
}

// the fused mixed-CG update on a launch: nparity 1 (an odd-half vector, the
// Schur complement's last launch, odd parity) or 2 (full lattice)
static void set_upd(lym::DslashParams& a, const lym::UpdArgs& u, const lym_ctx* c, int prec, int nparity) {
    const size_t S = c->g.stride;
    a.upd_state = u.state;
    a.upd_dot = u.dot;
    a.upd_ndot = u.ndot;
    a.upd_rpart = u.rpart;
    for (int q = 0; q < 2; ++q) {
        // parity q of the launch is vector parity q (full) or 0 (odd half)
        const int vq = nparity == 2 ? q : 0;
        if (nparity == 1 && q == 0) continue;
        a.upd_p[q] = (const char*)u.p + vq * 12 * S * selem(prec);
        a.upd_p_sc[q] = u.p_sc ? u.p_sc + vq * S : nullptr;
        a.upd_xs[q] = (char*)u.xs + vq * 12 * S * sizeof(lym::cplx<float>);
        a.upd_rs[q] = (char*)u.rs + vq * 12 * S * sizeof(lym::cplx<float>);
    }
}

// the batched solver's fused steps on a launch: vectors of nparity parities
// (1: odd-half multi-RHS fields on the odd-parity launch, 2: full lattice)
static void set_bfuse(lym::DslashParams& a, const lym::BFuse& f, const lym_ctx* c, int prec, int nr, int nparity) {
    const size_t S = c->g.stride;
    a.bmode = f.mode;
    a.balpha = f.balpha;
    for (int q = 0; q < 2; ++q) {
        const int vq = nparity == 2 ? q : 0;
        if (nparity == 1 && q == 0) continue;
        a.bsite[q] = f.bsite + vq * S * nr;
        if (f.mode == 2) {
            a.upd_p[q] = (const char*)f.p + vq * 12 * S * nr * selem(prec);
            a.upd_p_sc[q] = f.p_sc ? f.p_sc + vq * S * nr : nullptr;
            a.upd_xs[q] = (char*)f.xs + vq * 12 * S * nr * sizeof(lym::cplx<float>);
            a.upd_rs[q] = (char*)f.rs + vq * 12 * S * nr * sizeof(lym::cplx<float>);
        }
    }
}

// full lattice, nr right-hand sides, storage precision prec: out = Op in
int apply_full_nr(lym_op* op, int prec, int nr, int threads, void* out, float* out_sc, const void* in,
                  const float* in_sc, int gamma5, const int* skip, const int* lane_flags, int lane_skip_value,
                  bool batch_links, double* norm_partials, int norm_max, int mfma, const lym::UpdArgs* upd,
                  const lym::BFuse* bf) {
    lym_ctx* c = op->ctx;
    const size_t po = 12 * (size_t)c->g.stride * nr * selem(prec);
    const size_t so = (size_t)c->g.stride * nr;
    lym::DslashParams a = base_params(op);
    const int lk = set_op_data(a, op, prec, batch_links);
    a.allow_split = !batch_links;  // batches keep one form for every width (same bits)
    a.nr = nr;
    a.threads = threads;
    for (int p = 0; p < 2; ++p) {
        a.in[p] = (const char*)in + p * po;
        a.diag_in[p] = a.in[p];
        a.out[p] = (char*)out + p * po;
        if (prec == 2) {
            a.in_sc[p] = in_sc + p * so;
            a.diag_sc[p] = a.in_sc[p];
            a.out_sc[p] = out_sc + p * so;
        }
    }
    a.gamma5 = gamma5;
    a.parity = c->g.tsh ? -1 : 0;
    a.skip = skip;
    a.lane_flags = lane_flags;
    a.lane_skip_value = lane_skip_value;
    a.norm_partials = norm_partials;
    a.norm_max = norm_max;
    a.mfma = mfma;
    if (upd) set_upd(a, *upd, c, prec, 2);
    if (bf) set_bfuse(a, *bf, c, prec, nr, 2);
    return lym::launch_dslash(a, lk, op_has_clover(op) ? 2 : 1, false, c->stream);
}

// Schur complement on odd-half nr-fields, storage precision prec, explicit even intermediate
int apply_schur_nr(lym_op* op, int prec, int nr, int threads, void* out_o, float* out_sc, const void* in_o,
                   const float* in_sc, void* tmp_e, float* tmp_sc, int gamma5, const int* skip,
                   const int* lane_flags, int lane_skip_value, bool batch_links, double* norm_partials,
                   int norm_max, int mfma, const lym::UpdArgs* upd, const lym::BFuse* bf) {
    lym_ctx* c = op->ctx;
    lym::DslashParams a = base_params(op);
    const int lk = set_op_data(a, op, prec, batch_links);
    a.allow_split = !batch_links;  // batches keep one form for every width (same bits)
    a.nr = nr;
    a.threads = threads;
    a.mfma = mfma;
    a.skip = skip;
    a.lane_flags = lane_flags;
    a.lane_skip_value = lane_skip_value;
    lym::DslashParams b = a;
    a.parity = 0;
    a.in[1] = in_o;
    a.in_sc[1] = in_sc;
    a.out[0] = tmp_e;
    a.out_sc[0] = tmp_sc;
    lym::launch_dslash(a, lk, 0, true, c->stream);
    b.parity = 1;
    b.in[0] = tmp_e;
    b.in_sc[0] = tmp_sc;
    b.diag_in[1] = in_o;
    b.diag_sc[1] = in_sc;
    b.out[1] = out_o;
    b.out_sc[1] = out_sc;
    b.hop_sign = -1.0;
    b.gamma5 = gamma5;
    b.norm_partials = norm_partials;
    b.norm_max = norm_max;
    if (upd) set_upd(b, *upd, c, prec, 1);
    if (bf) set_bfuse(b, *bf, c, prec, nr, 1);
    return lym::launch_dslash(b, lk, op->csw != 0.0 ? 2 : 1, false, c->stream);
}

int apply_full_sloppy(lym_op* op, int prec, void* out, float* out_sc, const void* in, const float* in_sc, int gamma5,
                      const int* skip, double* norm_partials, int norm_max, const lym::UpdArgs* upd = nullptr) {
    return apply_full_nr(op, prec, 1, 0, out, out_sc, in, in_sc, gamma5, skip, nullptr, 0, false, norm_partials,
                         norm_max, 0, upd, nullptr);
}

int apply_schur_sloppy(lym_op* op, int prec, void* out_o, float* out_sc, const void* in_o, const float* in_sc,
                       void* tmp_e, float* tmp_sc, int gamma5, const int* skip, double* norm_partials, int norm_max,
                       const lym::UpdArgs* upd = nullptr) {
    return apply_schur_nr(op, prec, 1, 0, out_o, out_sc, in_o, in_sc, tmp_e, tmp_sc, gamma5, skip, nullptr, 0, false,
                          norm_partials, norm_max, 0, upd, nullptr);
}

// the EO driver's source and reconstruction steps on nr-fields (FP64)
// src_o = b_o - D_oe Ainv b_e  (te: even scratch)
void eo_source_nr(lym_op* op, int nr, const zd* b_e, const zd* b_o, zd* te, zd* src_o) {
    lym_ctx* c = op->ctx;
    lym::launch_clover_apply_nr(c->g, reinterpret_cast<const zd*>(op->Ainv), b_e, te, nr, c->stream);
    lym::DslashParams a = base_params(op);
    a.nr = nr;
    a.hop_sign = -1.0;
    a.parity = 1;
    a.in[0] = te;
    a.diag_in[1] = b_o;
    a.out[1] = src_o;
    lym::launch_dslash(a, link_kind(op->gauge), 1, false, c->stream);
}
// x_e = Ainv (b_e - D_eo x_o)
void eo_reconstruct_nr(lym_op* op, int nr, const zd* b_e, const zd* x_o, zd* x_e) {
    lym_ctx* c = op->ctx;
    lym::DslashParams a = base_params(op);
    a.nr = nr;
    a.hop_sign = -1.0;
    a.parity = 0;
    a.in[1] = x_o;
    a.diag_in[0] = b_e;
    a.out[0] = x_e;
    lym::launch_dslash(a, link_kind(op->gauge), 1, true, c->stream);
}
}  // namespace lym_internal

// ====================================================================== C ABI
extern "C" {

const char* lym_last_error(void) { return g_err.c_str(); }

int lym_device_count(int* n) {
    int k = 0;
    hipError_t e = hipGetDeviceCount(&k);
    if (e != hipSuccess) {
        *n = 0;
        return fail(LYM_ERR_DEVICE, "hipGetDeviceCount: %s", hipGetErrorString(e));
    }
    *n = k;
    return LYM_OK;
}

int lym_host_alloc(size_t bytes, void** out) {
    if (!out) return fail(LYM_ERR_INVALID, "null argument");
    *out = nullptr;
    if (bytes == 0) return LYM_OK;
    if (hipHostMalloc(out, bytes, hipHostMallocDefault) != hipSuccess) {
        *out = nullptr;
        return fail(LYM_ERR_NOMEM, "hipHostMalloc of %zu bytes", bytes);
    }
    return LYM_OK;
}

void lym_host_free(void* p) {
    if (p) (void)hipHostFree(p);
}

int lym_ctx_create(const int dims[4], int rep, int bc_t, lym_ctx** out) {
    if (!dims || !out) return fail(LYM_ERR_INVALID, "null argument");
    *out = nullptr;
    for (int i = 0; i < 4; ++i)
        if (dims[i] < 2) return fail(LYM_ERR_INVALID, "lattice extent %d must be >= 2", dims[i]);
    if (rep != LYM_REP_SU3_FUND && rep != LYM_REP_SU2_ADJ) return fail(LYM_ERR_INVALID, "unknown representation %d", rep);
    if (bc_t != 1 && bc_t != -1) return fail(LYM_ERR_INVALID, "bc_t must be +1 or -1");
    long V = (long)dims[0] * dims[1] * dims[2] * dims[3];
    // all extents even: the even-odd layout; an odd extent: one site class
    // (only the full-lattice FP64 operators and solvers run there)
    const int tsh = (dims[0] | dims[1] | dims[2] | dims[3]) % 2 == 0 ? 1 : 0;
    // a field component is read through 32-bit buffer offsets below 2 GiB
    if ((V >> tsh) > (1L << 27) - 64) return fail(LYM_ERR_INVALID, "lattice too large (at most 2^27 sites per class)");
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) return fail(LYM_ERR_DEVICE, "no HIP device available");
    lym_ctx* c = new lym_ctx();
    std::memset(c, 0, sizeof(*c));
    for (int i = 0; i < 4; ++i) {
        c->dims[i] = dims[i];
        c->g.L[i] = dims[i];
    }
    c->V = V;
    c->g.tsh = tsh;
    c->g.Lth = dims[3] >> tsh;
    c->g.Lxh = dims[0] >> tsh;
    c->g.Vh = (int)(V >> tsh);
    c->g.stride = (int)(((V >> tsh) + 63) / 64 * 64);
    c->g.bc_t = bc_t;
    c->rep = rep;
    (void)hipGetDevice(&c->device);
    int rc = LYM_OK;
    if (hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess) rc = fail(LYM_ERR_DEVICE, "stream");
    c->owns_stream = 1;
    if (!rc) rc = dmalloc(c, (void**)&c->red_partials, sizeof(double) * 8 * lym::RED_BLOCKS);
    if (!rc) rc = dmalloc(c, (void**)&c->red_result, 256);
    if (!rc && hipHostMalloc((void**)&c->host_pinned, 4096, hipHostMallocDefault) != hipSuccess)
        rc = fail(LYM_ERR_DEVICE, "hipHostMalloc");
    if (!rc && (hipEventCreate(&c->ev_start) != hipSuccess || hipEventCreate(&c->ev_stop) != hipSuccess))
        rc = fail(LYM_ERR_DEVICE, "hipEventCreate");
    if (rc) {
        lym_ctx_destroy(c);
        return rc;
    }
    *out = c;
    return LYM_OK;
}

void lym_ctx_destroy(lym_ctx* c) {
    if (!c) return;
    if (c->stream) (void)hipStreamSynchronize(c->stream);
    if (c->staging) (void)hipFree(c->staging);
    if (c->red_partials) (void)hipFree(c->red_partials);
    if (c->red_result) (void)hipFree(c->red_result);
    if (c->host_pinned) (void)hipHostFree(c->host_pinned);
    if (c->ev_start) (void)hipEventDestroy(c->ev_start);
    if (c->ev_stop) (void)hipEventDestroy(c->ev_stop);
    if (c->stream && c->owns_stream) (void)hipStreamDestroy(c->stream);
    delete c;
}

int lym_ctx_synchronize(lym_ctx* c) {
    if (!c) return fail(LYM_ERR_INVALID, "null context");
    HIPCHECK(hipStreamSynchronize(c->stream));
    return LYM_OK;
}

void* lym_ctx_stream(lym_ctx* c) { return c ? (void*)c->stream : nullptr; }

int lym_ctx_timer_start(lym_ctx* c) {
    if (!c) return fail(LYM_ERR_INVALID, "null context");
    HIPCHECK(hipEventRecord(c->ev_start, c->stream));
    return LYM_OK;
}

int lym_ctx_timer_stop(lym_ctx* c, float* ms) {
    if (!c || !ms) return fail(LYM_ERR_INVALID, "null argument");
    HIPCHECK(hipEventRecord(c->ev_stop, c->stream));
    HIPCHECK(hipEventSynchronize(c->ev_stop));
    HIPCHECK(hipEventElapsedTime(ms, c->ev_start, c->ev_stop));
    return LYM_OK;
}

// ---------------------------------------------------------------- gauge
int lym_gauge_create(lym_ctx* c, const double* links, int folded, int storage, lym_gauge** out) {
    if (!c || !links || !out) return fail(LYM_ERR_INVALID, "null argument");
    *out = nullptr;
    if (storage != LYM_LINKS_FULL && storage != LYM_LINKS_RECONSTRUCT) return fail(LYM_ERR_INVALID, "storage");
    lym_gauge* g = new lym_gauge();
    std::memset(g, 0, sizeof(*g));
    g->ctx = c;
    g->real_links = c->rep == LYM_REP_SU2_ADJ;
    g->compression = g->real_links ? (storage ? 6 : 9) : (storage ? 12 : 18);
    const size_t elem = g->real_links ? sizeof(double) : sizeof(zd);
    const int ncomp = g->real_links ? g->compression : g->compression / 2;
    const size_t host_bytes = (size_t)c->V * 4 * 9 * elem;
    int rc = dmalloc(c, &g->U, (size_t)2 * 4 * ncomp * c->g.stride * elem);
    if (!rc) rc = ensure_staging(c, host_bytes);
    if (rc) {
        lym_gauge_destroy(g);
        return rc;
    }
    hipError_t e = hipMemcpyAsync(c->staging, links, host_bytes, hipMemcpyHostToDevice, c->stream);
    if (e != hipSuccess) {
        lym_gauge_destroy(g);
        return fail(LYM_ERR_DEVICE, "upload links: %s", hipGetErrorString(e));
    }
    lym::launch_links_to_eo(c->g, c->staging, g->real_links, ncomp, folded, g->U, c->stream);
    // 12/6-real storage holds rows 0, 1 by construction; 18/9-real links are
    // checked once here (low-precision copies then keep rows 0, 1 only)
    g->rows_rebuild = ncomp == 6;
    static_assert(sizeof(double) * lym::RED_BLOCKS <= 4096, "pinned readback holds the partials");
    if (ncomp == 9) lym::launch_link_row2_err(g->U, g->real_links, c->g.stride, c->red_partials, c->stream);
    e = ncomp == 9 ? hipMemcpyAsync(c->host_pinned, c->red_partials, sizeof(double) * lym::RED_BLOCKS,
                                    hipMemcpyDeviceToHost, c->stream)
                   : hipSuccess;
    if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
    if (e != hipSuccess) {
        lym_gauge_destroy(g);
        return fail(LYM_ERR_DEVICE, "links layout: %s", hipGetErrorString(e));
    }
    if (ncomp == 9) {
        const double* m = reinterpret_cast<const double*>(c->host_pinned);
        double worst = 0;
        for (int k = 0; k < lym::RED_BLOCKS; ++k) worst = std::fmax(worst, m[k]);
        g->rows_rebuild = worst < 1e-6;
    }
    *out = g;
    return LYM_OK;
}

void lym_gauge_destroy(lym_gauge* g) {
    if (!g) return;
    if (g->ctx && g->ctx->stream) (void)hipStreamSynchronize(g->ctx->stream);
    if (g->U) (void)hipFree(g->U);
    if (g->B) (void)hipFree(g->B);
    delete g;
}

// ---------------------------------------------------------------- spinors
int lym_spinor_create(lym_ctx* c, lym_spinor** out) {
    if (!c || !out) return fail(LYM_ERR_INVALID, "null argument");
    *out = nullptr;
    lym_spinor* s = new lym_spinor();
    s->ctx = c;
    s->d = nullptr;
    int rc = dmalloc(c, (void**)&s->d, sizeof(zd) * full_n(c));
    if (rc) {
        delete s;
        return rc;
    }
    *out = s;
    return LYM_OK;
}

void lym_spinor_destroy(lym_spinor* s) {
    if (!s) return;
    if (s->ctx && s->ctx->stream) (void)hipStreamSynchronize(s->ctx->stream);
    if (s->d) (void)hipFree(s->d);
    delete s;
}

int lym_spinor_upload(lym_spinor* s, const double* host) {
    if (!s || !host) return fail(LYM_ERR_INVALID, "null argument");
    lym_ctx* c = s->ctx;
    size_t bytes = sizeof(zd) * 12 * (size_t)c->V;
    int rc = ensure_staging(c, bytes);
    if (rc) return rc;
    HIPCHECK(hipMemcpyAsync(c->staging, host, bytes, hipMemcpyHostToDevice, c->stream));
    lym::launch_spinor_to_eo(c->g, reinterpret_cast<const zd*>(c->staging), reinterpret_cast<zd*>(s->d), c->stream);
    CHECK_LAUNCH();
    HIPCHECK(hipStreamSynchronize(c->stream));
    return LYM_OK;
}

int lym_spinor_download(const lym_spinor* s, double* host) {
    if (!s || !host) return fail(LYM_ERR_INVALID, "null argument");
    lym_ctx* c = s->ctx;
    size_t bytes = sizeof(zd) * 12 * (size_t)c->V;
    int rc = ensure_staging(c, bytes);
    if (rc) return rc;
    lym::launch_spinor_from_eo(c->g, reinterpret_cast<const zd*>(s->d), reinterpret_cast<zd*>(c->staging), c->stream);
    CHECK_LAUNCH();
    HIPCHECK(hipMemcpyAsync(host, c->staging, bytes, hipMemcpyDeviceToHost, c->stream));
    HIPCHECK(hipStreamSynchronize(c->stream));
    return LYM_OK;
}

int lym_spinor_copy(lym_spinor* dst, const lym_spinor* src) {
    if (!dst || !src || dst->ctx != src->ctx) return fail(LYM_ERR_INVALID, "spinors of different contexts");
    lym::launch_copy(dst->d ? reinterpret_cast<zd*>(dst->d) : nullptr, reinterpret_cast<const zd*>(src->d),
                     full_n(dst->ctx), dst->ctx->stream);
    CHECK_LAUNCH();
    return LYM_OK;
}

int lym_spinor_set_zero(lym_spinor* s) {
    if (!s) return fail(LYM_ERR_INVALID, "null spinor");
    HIPCHECK(hipMemsetAsync(s->d, 0, sizeof(zd) * full_n(s->ctx), s->ctx->stream));
    return LYM_OK;
}

int lym_spinor_gamma5(lym_spinor* s) {
    if (!s) return fail(LYM_ERR_INVALID, "null spinor");
    lym::launch_scale_g5(reinterpret_cast<zd*>(s->d), 2, s->ctx->g.stride, s->ctx->stream);
    CHECK_LAUNCH();
    return LYM_OK;
}

static int read_result(lym_ctx* c, int n, double* out) {
    HIPCHECK(hipMemcpyAsync(c->host_pinned, c->red_result, sizeof(double) * n, hipMemcpyDeviceToHost, c->stream));
    HIPCHECK(hipStreamSynchronize(c->stream));
    for (int i = 0; i < n; ++i) out[i] = c->host_pinned[i];
    return LYM_OK;
}

int lym_spinor_squared_norm(const lym_spinor* s, double* out) {
    if (!s || !out) return fail(LYM_ERR_INVALID, "null argument");
    lym_ctx* c = s->ctx;
    lym::launch_norm2(reinterpret_cast<const zd*>(s->d), full_n(c), c->red_partials, c->red_result, c->stream);
    CHECK_LAUNCH();
    return read_result(c, 1, out);
}

int lym_spinor_dot(const lym_spinor* a, const lym_spinor* b, double* re, double* im) {
    if (!a || !b || !re || !im || a->ctx != b->ctx) return fail(LYM_ERR_INVALID, "bad arguments");
    lym_ctx* c = a->ctx;
    lym::launch_dot(reinterpret_cast<const zd*>(a->d), reinterpret_cast<const zd*>(b->d), full_n(c), c->red_partials,
                    c->red_result, c->stream);
    CHECK_LAUNCH();
    double r[2];
    int rc = read_result(c, 2, r);
    *re = r[0];
    *im = r[1];
    return rc;
}

int lym_spinor_gamma5dot(const lym_spinor* a, const lym_spinor* b, double* re, double* im) {
    if (!a || !b || !re || !im || a->ctx != b->ctx) return fail(LYM_ERR_INVALID, "bad arguments");
    lym_ctx* c = a->ctx;
    lym::launch_gamma5dot(reinterpret_cast<const zd*>(a->d), reinterpret_cast<const zd*>(b->d), full_n(c), 2,
                          c->g.stride, c->red_partials, c->red_result, c->stream);
    CHECK_LAUNCH();
    double r[2];
    int rc = read_result(c, 2, r);
    *re = r[0];
    *im = r[1];
    return rc;
}

int lym_spinor_difference_norm(const lym_spinor* a, const lym_spinor* b, double* out) {
    if (!a || !b || !out || a->ctx != b->ctx) return fail(LYM_ERR_INVALID, "bad arguments");
    lym_ctx* c = a->ctx;
    lym::launch_diffnorm(reinterpret_cast<const zd*>(a->d), reinterpret_cast<const zd*>(b->d), full_n(c),
                         c->red_partials, c->red_result, c->stream);
    CHECK_LAUNCH();
    return read_result(c, 1, out);
}

int lym_spinor_axpy(lym_spinor* y, double are, double aim, const lym_spinor* x) {
    if (!y || !x || y->ctx != x->ctx) return fail(LYM_ERR_INVALID, "bad arguments");
    lym::launch_axpy(reinterpret_cast<zd*>(y->d), lym::mk(are, aim), reinterpret_cast<const zd*>(x->d),
                     full_n(y->ctx), y->ctx->stream);
    CHECK_LAUNCH();
    return LYM_OK;
}

// ---------------------------------------------------------------- operators
int lym_op_create(lym_ctx* c, int kind, unsigned power, double kappa, double csw, int gamma5, lym_op** out) {
    if (!c || !out) return fail(LYM_ERR_INVALID, "null argument");
    *out = nullptr;
    if (kind < LYM_OP_WILSON || kind > LYM_OP_EO_CLOVER) return fail(LYM_ERR_INVALID, "unknown operator kind %d", kind);
    if (power != 1 && power != 2) return fail(LYM_ERR_INVALID, "Power of the Dirac Wilson Operator not supported: %u", power);
    if (kind == LYM_OP_EO_CLOVER)
        for (int i = 0; i < 4; ++i)
            if (c->dims[i] % 2) return fail(LYM_ERR_INVALID, "even-odd operator needs even lattice extents");
    lym_op* op = new lym_op();
    std::memset(op, 0, sizeof(*op));
    op->ctx = c;
    static std::atomic<uint64_t> next_uid{1};
    op->uid = next_uid++;
    op->kind = kind;
    op->power = power;
    op->kappa = kappa;
    op->csw = (kind == LYM_OP_CLOVER || kind == LYM_OP_EO_CLOVER) ? csw : 0.0;
    op->gamma5 = gamma5 ? 1 : 0;
    op->diag_dirty = 1;
    int rc = dmalloc(c, (void**)&op->dev_kappa, 256);
    if (rc) {
        delete op;
        return rc;
    }
    lym::launch_set_double(op->dev_kappa, kappa, c->stream);
    *out = op;
    return LYM_OK;
}

int lym_op_create_by_name(lym_ctx* c, const char* name, unsigned power, double kappa, double csw, int gamma5,
                          lym_op** out) {
    if (!name) return fail(LYM_ERR_INVALID, "null name");
    int kind;
    if (!std::strcmp(name, "DiracWilson")) kind = LYM_OP_WILSON;
    else if (!std::strcmp(name, "BasicDiracWilson")) kind = LYM_OP_BASIC_WILSON;
    else if (!std::strcmp(name, "Improved")) kind = LYM_OP_CLOVER;
    else if (!std::strcmp(name, "EvenOddImproved")) kind = LYM_OP_EO_CLOVER;
    else return fail(LYM_ERR_UNSUPPORTED, "Dirac Wilson Operator %s not supported!", name);
    return lym_op_create(c, kind, power, kappa, csw, gamma5, out);
}

void lym_op_destroy(lym_op* op) {
    if (!op) return;
    if (op->ctx && op->ctx->stream) (void)hipStreamSynchronize(op->ctx->stream);
    lym_internal::forget_op_captures(op);
    free_sloppy(op);
    if (op->A) (void)hipFree(op->A);
    if (op->Ainv) (void)hipFree(op->Ainv);
    if (op->tmp) (void)hipFree(op->tmp);
    if (op->tmp2) (void)hipFree(op->tmp2);
    if (op->dev_kappa) (void)hipFree(op->dev_kappa);
    for (int i = 0; i < 4; ++i) {
        if (op->sp[i]) (void)hipFree(op->sp[i]);
        if (op->sp_sc[i]) (void)hipFree(op->sp_sc[i]);
    }
    delete op;
}

int lym_op_set_gauge(lym_op* op, const lym_gauge* g) {
    if (!op || !g) return fail(LYM_ERR_INVALID, "null argument");
    const lym_ctx *a = op->ctx, *b = g->ctx;
    if (a != b && (a->device != b->device || a->rep != b->rep || a->g.bc_t != b->g.bc_t ||
                   std::memcmp(a->dims, b->dims, sizeof(a->dims)) != 0))
        return fail(LYM_ERR_INVALID, "gauge field of a context with another geometry or device");
    op->gauge = g;
    op->diag_dirty = 1;
    op->sloppy_prec = 0;
    return LYM_OK;
}

int lym_op_set_kappa(lym_op* op, double kappa) {
    if (!op) return fail(LYM_ERR_INVALID, "null operator");
    if (op->kappa != kappa) {
        op->diag_dirty = 1;
        lym::launch_set_double(op->dev_kappa, kappa, op->ctx->stream);
    }
    op->kappa = kappa;
    return LYM_OK;
}

int lym_op_set_csw(lym_op* op, double csw) {
    if (!op) return fail(LYM_ERR_INVALID, "null operator");
    if (op->kind != LYM_OP_CLOVER && op->kind != LYM_OP_EO_CLOVER)
        return fail(LYM_ERR_INVALID, "csw only applies to clover operators");
    if (op->csw != csw) op->diag_dirty = 1;
    op->csw = csw;
    return LYM_OK;
}

int lym_op_set_gamma5(lym_op* op, int gamma5) {
    if (!op) return fail(LYM_ERR_INVALID, "null operator");
    op->gamma5 = gamma5 ? 1 : 0;
    return LYM_OK;
}

int lym_op_get_params(const lym_op* op, double* kappa, double* csw, int* gamma5, int* kind, unsigned* power) {
    if (!op) return fail(LYM_ERR_INVALID, "null operator");
    if (kappa) *kappa = op->kappa;
    if (csw) *csw = op->csw;
    if (gamma5) *gamma5 = op->gamma5;
    if (kind) *kind = op->kind;
    if (power) *power = op->power;
    return LYM_OK;
}

int lym_op_multiply(lym_op* op, lym_spinor* out, const lym_spinor* in) {
    int rc = check_op_ready(op);
    if (rc) return rc;
    if (!out || !in || out->ctx != op->ctx || in->ctx != op->ctx) return fail(LYM_ERR_INVALID, "bad spinors");
    if (out == in) return fail(LYM_ERR_INVALID, "multiply needs distinct input and output");
    lym_ctx* c = op->ctx;
    zd* o = reinterpret_cast<zd*>(out->d);
    const zd* i = reinterpret_cast<const zd*>(in->d);
    if (op->kind == LYM_OP_EO_CLOVER) {
        // even sites copied through, odd sites get the Schur complement
        HIPCHECK(hipMemcpyAsync(o, i, sizeof(zd) * half_n(c), hipMemcpyDeviceToDevice, c->stream));
        apply_power(op, par(o, c, 1), par(i, c, 1), nullptr);
    } else {
        apply_power(op, o, i, nullptr);
    }
    CHECK_LAUNCH();
    return LYM_OK;
}

int lym_op_multiply_sloppy(lym_op* op, int prec, lym_spinor* out, const lym_spinor* in) {
    int rc = check_op_ready(op);
    if (rc) return rc;
    if (!out || !in || out->ctx != op->ctx || in->ctx != op->ctx) return fail(LYM_ERR_INVALID, "bad spinors");
    if (out == in) return fail(LYM_ERR_INVALID, "multiply needs distinct input and output");
    if (prec != LYM_PREC_SINGLE && prec != LYM_PREC_HALF)
        return fail(LYM_ERR_INVALID, "storage precision must be LYM_PREC_SINGLE or LYM_PREC_HALF");
    if (!op->ctx->g.tsh) return fail(LYM_ERR_UNSUPPORTED, "low-precision applies need all lattice extents even");
    rc = op_sloppy_data(op, prec, false);
    if (rc) return rc;
    lym_ctx* c = op->ctx;
    const size_t S = c->g.stride;
    for (int i = 0; i < 4; ++i) {
        if (!op->sp[i] && (rc = dmalloc(c, &op->sp[i], 24 * S * sizeof(lym::cplx<float>)))) return rc;
        if (!op->sp_sc[i] && (rc = dmalloc(c, (void**)&op->sp_sc[i], 2 * S * sizeof(float)))) return rc;
    }
    const size_t es = prec == LYM_PREC_SINGLE ? 8 : 4;  // bytes per complex component
    auto po = [&](void* f, int p) { return (void*)((char*)f + p * 12 * S * es); };
    auto ps = [&](float* f, int p) { return f + p * S; };
    zd* o = reinterpret_cast<zd*>(out->d);
    const zd* i = reinterpret_cast<const zd*>(in->d);
    lym::launch_spinor_to_sloppy(prec, i, op->sp[0], op->sp_sc[0], 2 * (long)S, (int)S, c->stream);
    if (op->kind == LYM_OP_EO_CLOVER) {
        // even sites copied through (FP64), odd sites: the Schur complement on
        // the odd half in the storage precision
        void* mid = op->power == 2 ? op->sp[2] : op->sp[1];
        float* mid_sc = op->power == 2 ? op->sp_sc[2] : op->sp_sc[1];
        lym_internal::apply_schur_sloppy(op, prec, po(mid, 1), ps(mid_sc, 1), po(op->sp[0], 1), ps(op->sp_sc[0], 1), op->sp[3],
                           op->sp_sc[3], op->gamma5, nullptr, nullptr, 0);
        if (op->power == 2)
            lym_internal::apply_schur_sloppy(op, prec, po(op->sp[1], 1), ps(op->sp_sc[1], 1), po(mid, 1), ps(mid_sc, 1),
                               op->sp[3], op->sp_sc[3], op->gamma5, nullptr, nullptr, 0);
        lym::launch_spinor_from_sloppy(prec, po(op->sp[1], 1), ps(op->sp_sc[1], 1), par(o, c, 1), (long)S, (int)S,
                                       c->stream);
        HIPCHECK(hipMemcpyAsync(o, i, sizeof(zd) * half_n(c), hipMemcpyDeviceToDevice, c->stream));
    } else {
        void* mid = op->power == 2 ? op->sp[2] : op->sp[1];
        float* mid_sc = op->power == 2 ? op->sp_sc[2] : op->sp_sc[1];
        lym_internal::apply_full_sloppy(op, prec, mid, mid_sc, op->sp[0], op->sp_sc[0], op->gamma5, nullptr, nullptr, 0);
        if (op->power == 2) lym_internal::apply_full_sloppy(op, prec, op->sp[1], op->sp_sc[1], mid, mid_sc, op->gamma5, nullptr, nullptr, 0);
        lym::launch_spinor_from_sloppy(prec, op->sp[1], op->sp_sc[1], o, 2 * (long)S, (int)S, c->stream);
    }
    CHECK_LAUNCH();
    return LYM_OK;
}

int lym_op_multiply_sloppy_batch(lym_op* op, int prec, int matrix_cores, unsigned nrhs, lym_spinor* const* outs,
                                 const lym_spinor* const* ins) {
    int rc = check_op_ready(op);
    if (rc) return rc;
    if (prec != LYM_PREC_SINGLE && prec != LYM_PREC_HALF)
        return fail(LYM_ERR_INVALID, "storage precision must be LYM_PREC_SINGLE or LYM_PREC_HALF");
    if (!op->ctx->g.tsh) return fail(LYM_ERR_UNSUPPORTED, "low-precision applies need all lattice extents even");
    if (!outs || !ins || !lym::batch_nr_supported((int)nrhs))
        return fail(LYM_ERR_INVALID, "nrhs must be one of 1,2,3,4,6,8,12,16 with output and input arrays");
    lym_ctx* c = op->ctx;
    for (unsigned q = 0; q < nrhs; ++q) {
        if (!outs[q] || !ins[q] || outs[q]->ctx != c || ins[q]->ctx != c) return fail(LYM_ERR_INVALID, "bad spinors");
        for (unsigned k = 0; k < nrhs; ++k)
            if ((const lym_spinor*)outs[q] == ins[k]) return fail(LYM_ERR_INVALID, "outputs must differ from inputs");
    }
    const int nr = (int)nrhs;
    if (matrix_cores && !(prec == LYM_PREC_HALF && c->rep == LYM_REP_SU3_FUND && (nr == 2 || nr == 4 || nr == 8)))
        return fail(LYM_ERR_UNSUPPORTED, "the matrix-core stencil is FP16 SU(3) with 2, 4 or 8 right-hand sides");
    if ((rc = op_sloppy_data(op, prec, true))) return rc;
    if (matrix_cores && (op->Usf ? link_kind(op->gauge) : op->sloppy_lk) != 0)
        return fail(LYM_ERR_UNSUPPORTED, "the matrix-core stencil reads SU(3) links with all three rows");
    const size_t S = c->g.stride;
    const size_t es = prec == LYM_PREC_SINGLE ? 8 : 4;
    std::vector<void*> bufs;
    auto alloc = [&](void** p, size_t bytes) {
        if (!rc && !(rc = dmalloc(c, p, bytes))) bufs.push_back(*p);
    };
    zd *B = nullptr, *X = nullptr;
    zd** ptr_dev = nullptr;
    void *sin = nullptr, *sout = nullptr, *smid = nullptr, *stmp = nullptr;
    float *sin_sc = nullptr, *sout_sc = nullptr, *smid_sc = nullptr, *stmp_sc = nullptr;
    alloc((void**)&B, sizeof(zd) * 24 * S * nr);
    alloc((void**)&X, sizeof(zd) * 24 * S * nr);
    alloc((void**)&ptr_dev, sizeof(zd*) * nr);
    for (void** f : {&sin, &sout, &smid, &stmp}) alloc(f, 24 * S * nr * es);
    for (float** f : {&sin_sc, &sout_sc, &smid_sc, &stmp_sc}) alloc((void**)f, 2 * S * nr * sizeof(float));
    if (!rc) {
        std::vector<zd*> hp(nr);
        for (int q = 0; q < nr; ++q) hp[q] = reinterpret_cast<zd*>(ins[q]->d);
        HIPCHECK(hipMemcpyAsync(ptr_dev, hp.data(), sizeof(zd*) * nr, hipMemcpyHostToDevice, c->stream));
        lym::launch_pack(ptr_dev, nr, B, full_n(c), c->stream);
        lym::launch_spinor_to_sloppy(prec, B, sin, sin_sc, 2 * (long)S * nr, (int)(S * nr), c->stream);
        const size_t po = 12 * S * nr * es, so = S * nr;  // parity offsets (bytes of the field, scales)
        auto pf = [&](void* f, int p) { return (void*)((char*)f + p * po); };
        auto pscl = [&](float* f, int p) { return f + p * so; };
        void* res = op->power == 2 ? sout : smid;
        float* res_sc = op->power == 2 ? sout_sc : smid_sc;
        if (op->kind == LYM_OP_EO_CLOVER) {
            lym_internal::apply_schur_nr(op, prec, nr, 0, pf(smid, 1), pscl(smid_sc, 1), pf(sin, 1), pscl(sin_sc, 1),
                                         stmp, stmp_sc, op->gamma5, nullptr, nullptr, 0, true, nullptr, 0, matrix_cores, nullptr, nullptr);
            if (op->power == 2)
                lym_internal::apply_schur_nr(op, prec, nr, 0, pf(sout, 1), pscl(sout_sc, 1), pf(smid, 1),
                                             pscl(smid_sc, 1), stmp, stmp_sc, op->gamma5, nullptr, nullptr, 0, true,
                                             nullptr, 0, matrix_cores, nullptr, nullptr);
            // even sites copied through in FP64, odd sites from the sloppy result
            lym::launch_spinor_from_sloppy(prec, pf(res, 1), pscl(res_sc, 1), X + 12 * S * nr, (long)S * nr,
                                           (int)(S * nr), c->stream);
            HIPCHECK(hipMemcpyAsync(X, B, sizeof(zd) * 12 * S * nr, hipMemcpyDeviceToDevice, c->stream));
        } else {
            lym_internal::apply_full_nr(op, prec, nr, 0, smid, smid_sc, sin, sin_sc, op->gamma5, nullptr, nullptr, 0,
                                        true, nullptr, 0, matrix_cores, nullptr, nullptr);
            if (op->power == 2)
                lym_internal::apply_full_nr(op, prec, nr, 0, sout, sout_sc, smid, smid_sc, op->gamma5, nullptr,
                                            nullptr, 0, true, nullptr, 0, matrix_cores, nullptr, nullptr);
            lym::launch_spinor_from_sloppy(prec, res, res_sc, X, 2 * (long)S * nr, (int)(S * nr), c->stream);
        }
        for (int q = 0; q < nr; ++q) hp[q] = reinterpret_cast<zd*>(outs[q]->d);
        HIPCHECK(hipMemcpyAsync(ptr_dev, hp.data(), sizeof(zd*) * nr, hipMemcpyHostToDevice, c->stream));
        lym::launch_unpack(X, nr, ptr_dev, full_n(c), c->stream);
        hipError_t e = hipGetLastError();
        if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
        if (e != hipSuccess) rc = fail(LYM_ERR_DEVICE, "batched sloppy apply: %s", hipGetErrorString(e));
    }
    (void)hipStreamSynchronize(c->stream);
    for (void* p : bufs) (void)hipFree(p);
    return rc;
}

int lym_op_multiply_add(lym_op* op, lym_spinor* out, const lym_spinor* v1, const lym_spinor* v2, double are,
                        double aim) {
    int rc = check_op_ready(op);
    if (rc) return rc;
    if (!out || !v1 || !v2) return fail(LYM_ERR_INVALID, "null spinor");
    if (out == v1 || out == v2) return fail(LYM_ERR_INVALID, "multiplyAdd output must not alias an input");
    lym_ctx* c = op->ctx;
    zd* o = reinterpret_cast<zd*>(out->d);
    const zd* a = reinterpret_cast<const zd*>(v1->d);
    const zd* b = reinterpret_cast<const zd*>(v2->d);
    zd al = lym::mk(are, aim);
    if (op->kind == LYM_OP_EO_CLOVER) {
        if (op->power == 1)
            return fail(LYM_ERR_UNSUPPORTED, "EvenOddImprovedDiracWilsonOperator::multiplyAdd not implemented");
        // SquareEvenOdd multiplyAdd: Q^2 v1 + alpha v2 on all sites (even part of Q^2 is the identity)
        HIPCHECK(hipMemcpyAsync(o, a, sizeof(zd) * half_n(c), hipMemcpyDeviceToDevice, c->stream));
        apply_power(op, par(o, c, 1), par(a, c, 1), nullptr);
        lym::launch_axpy(o, al, b, full_n(c), c->stream);
    } else if (op->power == 1) {
        apply_full1(op, o, a, b, al, op->gamma5, nullptr);
    } else {
        zd* t = reinterpret_cast<zd*>(op->tmp);
        apply_full1(op, t, a, nullptr, lym::mk(0.0, 0.0), op->gamma5, nullptr);
        apply_full1(op, o, t, b, al, op->gamma5, nullptr);
    }
    CHECK_LAUNCH();
    return LYM_OK;
}

}  // extern "C"
