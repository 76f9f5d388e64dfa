// lym_kernels.h -- kernel parameter blocks and host-side launchers.
#pragma once

#include "lym_common.h"

namespace lym {

typedef cplx<double> zd;

// Parameters of one stencil launch.  Field pointers are per parity; the
// storage element type follows `prec` (0 FP64, 1 FP32, 2 FP16 + per-site
// scales); multi-RHS fields hold `nr` right-hand sides, rhs fastest.
struct DslashParams {
    Geom g;
    int prec;              // storage precision of spinors, links and clover blocks
    int nr;                // right-hand sides per field
    int spb;               // sites per workgroup (set by the launcher)
    const void* in[2];     // per parity: the field read at neighbours is in[1 - p]
    const float* in_sc[2];
    const void* diag_in[2];  // per parity: the field under the site-diagonal term
    const float* diag_sc[2];
    void* out[2];
    float* out_sc[2];
    const void* v2[2];     // optional axpy operand
    const float* v2_sc[2];
    const void* links;
    const void* A;         // [2][36][stride] clover blocks of D
    const float* A_sc;     // [2][stride] (FP16)
    const void* Ainv;      // [36][stride] inverse blocks (even sites)
    const float* Ainv_sc;  // [stride] (FP16)
    double kappa;
    const double* kappa_dev;  // when set, kappa is read from device memory instead
    double hop_sign;
    zd alpha;
    int gamma5;
    int parity;            // 0 or 1: that parity only; -1: both
    int nblocks_per_parity;
    const int* skip;       // optional device flag: nonzero -> the launch does nothing (solver done)
    const int* lane_flags; // optional per-RHS device flags: lane r does nothing when
    int lane_skip_value;   //   lane_flags[r] == lane_skip_value (an RHS that is done or needs no update)
    int threads;           // workgroup size (0: LYM_DSLASH_THREADS); a workgroup holds threads / nr sites
    int xslice_blocks;     // set by the launcher: site blocks per x-slice and parity when an x-slice
                           // is a whole number of blocks, else 0
    int allow_split;       // the caller allows the two-threads-per-site form (low precision)
    int mfma;              // the caller allows the matrix-core form (FP16 SU(3) multi-RHS, nr 2, 4 or 8)
    int split;             // set by the launcher: threads per half of that form, or 0
    double* norm_partials; // optional (FP32/FP16 only): workgroup b writes the sum of |out|^2 over
    int norm_max;          //   its sites to norm_partials[b], when the launch has at most norm_max workgroups
    // Per-lane operators (multi-RHS launches of the serving engine): lane r uses
    // kappa_dev[r] and the clover blocks at A + r * A_lane_bytes (scales at
    // A_sc + r * Asc_lane), Ainv + r * Ainv_lane_bytes (Ainv_sc + r * Ainvsc_lane)
    int lane_ops;
    long A_lane_bytes, Ainv_lane_bytes, Asc_lane, Ainvsc_lane;
    const int* lane_set;  // lane_ops: lane r reads the block set lane_set[r] (NULL: set r)
    // Fused update of the mixed-precision CG (single RHS, FP32/FP16 storage; the
    // last launch of A p): instead of storing its output v = (A p)_x, the
    // launch updates x_s += alpha p and r_s -= alpha v at every site, with
    // alpha = rr_s / (the sum of the upd_ndot <p, Ap> partials at upd_dot) and
    // rr_s from the solver state upd_state (an MCGState); workgroup b writes the
    // sum of |r_s|^2 over its sites to upd_rpart[b], workgroup 0 stores alpha
    // and rr_prev = rr_s in the state.  p: storage precision; x_s, r_s: complex
    // float [12][stride] per parity.
    void* upd_state;
    const double* upd_dot;
    int upd_ndot;
    const void* upd_p[2];
    const float* upd_p_sc[2];
    void* upd_xs[2];
    void* upd_rs[2];
    double* upd_rpart;
    // The same two steps for a batch, one value per (site, RHS) so that every
    // RHS sums its terms over the batched solver's fixed site chunks:
    // bmode 1 (the launch that writes Q p): also bsite[p][site * nr + r] =
    // |(Q p)_site|^2 of lane r; bmode 2 (the last launch of Q^2 p): no output
    // store; x_s += alpha_r p and r_s -= alpha_r v (alpha from balpha[r]; p,
    // x_s, r_s as upd_p / upd_xs / upd_rs, multi-RHS fields) and bsite = the
    // site's |r_s|^2 of lane r.
    int bmode;
    double* bsite[2];
    const double* balpha;
};

// Arguments of the fused mixed-CG update (DslashParams::upd_*) for one vector
// of nparity parities: p in storage precision with its scales, x_s and r_s
// complex float [nparity][12][stride]; dot: the <p, Ap> partials (count set by
// the caller once known); rpart: NORM_FOLD_MAX doubles for the |r_s|^2 partials
struct UpdArgs {
    void* state;
    const double* dot;
    int ndot;
    const void* p;
    const float* p_sc;
    void* xs;
    void* rs;
    double* rpart;
};

// link_kind: 0 complex 18, 1 complex 12, 2 real 9, 3 real 6; diag: 0 none, 1 identity, 2 clover.
// Returns the number of output-norm partials written (0: none were asked for or they did not fit).
int launch_dslash(DslashParams a, int link_kind, int diag, bool inv, hipStream_t st);

// ---------------------------------------------------------------- layout (layout.hip)
void launch_spinor_to_eo(const Geom& g, const zd* aos, zd* eo, hipStream_t st);
void launch_spinor_from_eo(const Geom& g, const zd* eo, zd* aos, hipStream_t st);
// links_aos: complex [V][4][3x3 col-major] (real_links = 0) or real (real_links = 1)
void launch_links_to_eo(const Geom& g, const void* links_aos, int real_links, int ncomp, int unfold_bc, void* eo,
                        hipStream_t st);

// ---------------------------------------------------------------- clover (clover.hip)
void launch_clover_B(const Geom& g, const void* links, int link_kind, zd* B, hipStream_t st);
void launch_clover_A(const Geom& g, const zd* B, double kc, zd* A, hipStream_t st);
// Ainv on parity 0 from A; returns via *flag (device int) nonzero when a block is singular
void launch_clover_inv(const Geom& g, const zd* A, zd* Ainv, int* flag, hipStream_t st);  // pivoted; flag: singular count
void launch_clover_inv_ldl(const Geom& g, const zd* A, zd* Ainv, int* fallback, hipStream_t st);  // unpivoted
void launch_clover_apply(const Geom& g, const zd* blocks, const zd* in, zd* out, int gamma5, hipStream_t st);

// ---------------------------------------------------------------- BLAS-1 (blas.hip)
constexpr int RED_BLOCKS = 512;

// reductions over n complex elements; results land in `result` (device) as doubles
void launch_norm2(const zd* a, long n, double* partials, double* result, hipStream_t st);
void launch_dot(const zd* a, const zd* b, long n, double* partials, double* result, hipStream_t st);
void launch_gamma5dot(const zd* a, const zd* b, long half_stride_n, int nparity, int stride, double* partials,
                      double* result, hipStream_t st);
void launch_diffnorm(const zd* a, const zd* b, long n, double* partials, double* result, hipStream_t st);

void launch_axpy(zd* y, zd a, const zd* x, long n, hipStream_t st);         // y += a x
void launch_xpay(zd* y, const zd* x, zd a, long n, hipStream_t st);         // y = x + a y
void launch_scale_g5(zd* y, int nparity, int stride, hipStream_t st);       // lower spins negated
void launch_copy(zd* y, const zd* x, long n, hipStream_t st);
void launch_set_double(double* p, double v, hipStream_t st);                // *p = v, ordered on st
void launch_sub(zd* out, const zd* a, const zd* b, long n, hipStream_t st); // out = a - b

// ---------------------------------------------------------------- solver state machines
struct CGState {
    int done, converged, step, max_steps;
    double rr, rr_new, alpha_re, alpha_im, beta, eps;
    double rr_prev;  // rr at the start of the iteration (set by the update, read by the p update)
};

struct BiCGState {
    int done, converged, step, max_steps, breakdown, copy_b;
    double rho_re, rho_im, rhon_re, rhon_im;
    double alpha_re, alpha_im, omega_re, omega_im, beta_re, beta_im;
    double norm, precision;
};

void launch_cg_init(CGState* s, const zd* r, long n, double* partials, double eps, int max_steps, hipStream_t st);
// one CG iteration after Ap = A p: <p, Ap> partials (slots 0, 1 of partials),
// the update folding them into alpha (|r|^2 partials into slot 2), the p update
// folding those into beta and the stopping rule
void launch_cg_alpha(CGState* s, const zd* p, const zd* Ap, long n, double* partials, hipStream_t st);
void launch_cg_update(CGState* s, zd* x, zd* r, const zd* p, const zd* Ap, long n, double* partials,
                      hipStream_t st);
void launch_cg_p(CGState* s, zd* p, const zd* r, long n, double* partials, hipStream_t st);

void launch_bicg_init(BiCGState* s, double precision, int max_steps, hipStream_t st);
// r = b - Ax, rh = b + Ax, p = nu = 0
void launch_bicg_setup(zd* r, zd* rh, zd* p, zd* nu, const zd* b, const zd* Ax, long n, hipStream_t st);
void launch_bicg_rho(BiCGState* s, const zd* rh, const zd* r, long n, double* partials, hipStream_t st);
// the p, s and x/r updates fold the partials of the dot launched before them
// (rho, alpha, omega) into their scalars
void launch_bicg_p(BiCGState* s, zd* p, const zd* r, const zd* nu, long n, const double* partials, hipStream_t st);
void launch_bicg_alpha(BiCGState* s, const zd* rh, const zd* nu, long n, double* partials, hipStream_t st);
void launch_bicg_s(BiCGState* s, zd* sv, const zd* r, const zd* nu, long n, const double* partials, hipStream_t st);
void launch_bicg_omega(BiCGState* s, const zd* t, const zd* sv, long n, double* partials, hipStream_t st);
void launch_bicg_update(BiCGState* s, zd* x, zd* r, const zd* p, const zd* sv, const zd* t, const zd* b, long n,
                        double* partials, hipStream_t st);

}  // namespace lym

namespace lym {
struct MCGState {
    int done, converged, step, max_steps, rel, skip_rel, nrel, pad;
    double rr_s, rr_true, maxrr, alpha, beta, eps, delta2;
    double rr_prev;  // rr_s at the start of the iteration (set by the update, read by the p update)
};

// low-precision copies of operator data (mixed.hip); prec 1 FP32, 2 FP16
void launch_convert_links(int prec, const void* src, void* dst, long n, int real_links, hipStream_t st);
void launch_convert_blocks(int prec, const zd* src, void* dst, float* sc, int nparity, int stride, hipStream_t st);
// rows 0 and 1 of FP64 links [2][4][nc][stride] into the packed low-precision
// layout of link kinds 1 / 3 (row 2 rebuilt in the stencil)
void launch_pack_links(int prec, const void* src, int nc, int real_links, void* dst, int stride, hipStream_t st);
// per-workgroup maxima (RED_BLOCKS of them) of the deviation of row 2 of
// 18-real / 9-real FP64 links from the row rebuilt out of rows 0 and 1
void launch_link_row2_err(const void* src, int real_links, int stride, double* partials, hipStream_t st);
// FP64 spinor field <-> sloppy field of storage precision prec (FP16: with its
// per-site scales), nsites = nparity * stride sites
void launch_spinor_to_sloppy(int prec, const zd* src, void* dst, float* dst_sc, long nsites, int stride,
                             hipStream_t st);
void launch_spinor_from_sloppy(int prec, const void* src, const float* src_sc, zd* dst, long nsites, int stride,
                               hipStream_t st);
// mixed-precision CG with reliable updates; vectors span nsites = nparity * stride sites
void launch_mcg_init(int prec, MCGState* s, const zd* r, cplx<float>* rs, cplx<float>* xs, void* p, float* p_sc,
                     long nsites, int stride, double* partials, double eps, double delta, int max_steps, hipStream_t st);
// <p, Ap> partials into partials[0, RED_BLOCKS); launch_mcg_update folds ndot
// partials of <p, Ap> (those, or the stencil's output-norm partials: for
// A = Q^2 with Q hermitian, <p, Ap> = |Q p|^2) into alpha and writes |r_s|^2
// partials to partials[RED_BLOCKS, 2 RED_BLOCKS)
constexpr int NORM_FOLD_MAX = 2048;  // most partials the update folds
void launch_mcg_alpha(int prec, MCGState* s, void* p, float* p_sc, void* Ap, float* Ap_sc, long nsites, int stride,
                      double* partials, hipStream_t st);
void launch_mcg_update(int prec, MCGState* s, cplx<float>* xs, cplx<float>* rs, void* p, float* p_sc, void* Ap,
                       float* Ap_sc, long nsites, int stride, double* partials, const double* dot_partials, int ndot,
                       hipStream_t st);
void launch_rel_accum(MCGState* s, zd* x, cplx<float>* xs, long nsites, int stride, hipStream_t st);
void launch_accum_xs(zd* x, const cplx<float>* xs, long nsites, int stride, hipStream_t st);
void launch_rel_resid(MCGState* s, zd* r, cplx<float>* rs, const zd* b, const zd* Ax, long nsites, int stride,
                      double* partials, hipStream_t st);
// p = r_s + beta p after launch_mcg_update, with beta and the state step taken
// from the update's |r_s|^2 partials (nupd > 0: the nupd partials at upd_part
// of an update fused into the stencil)
void launch_mcg_p(int prec, MCGState* s, void* p, float* p_sc, const cplx<float>* rs, long nsites, int stride,
                  const double* partials, hipStream_t st, const double* upd_part = nullptr, int nupd = 0);

// ---------------------------------------------------------------- multi-RHS mixed CG (batch.hip)
constexpr int BMAX = 16;
struct BState {
    int all_done, any_rel, skip_rel, max_steps;
    int done[BMAX], conv[BMAX], step[BMAX], rel[BMAX], nrel[BMAX];
    double delta2;
    double eps[BMAX], rr_s[BMAX], rr_true[BMAX], maxrr[BMAX], alpha[BMAX], beta[BMAX];
};
constexpr int BATCH_CHUNK = 64;  // sites per fixed summation chunk of the batched solver
struct BcgArgs {
    zd* x;
    zd* r;
    const zd* b;
    const zd* Ax;
    cplx<float>* xs;
    cplx<float>* rs;
    void* p;
    float* p_sc;
    void* Ap;
    float* Ap_sc;
    long nsites;
    int stride;
    int max_steps;
    double* partials;  // BMAX * ceil(nsites / BATCH_CHUNK) doubles
    const int* mask = nullptr;  // steps 0 (init) and 6 (final accumulate): only lanes with mask[q] != 0
    const double* bsite = nullptr;  // steps 7, 8: one value per (site, RHS), site-major
    int vh = 0;                     // steps 7, 8: the real sites of each parity (of stride)
};
// the fused batch steps on one stencil launch (DslashParams::bmode)
struct BFuse {
    int mode;
    double* bsite;
    const double* balpha;
    const void* p;
    const float* p_sc;
    void* xs;
    void* rs;
};
// which: 0 init, 1 alpha, 2 update, 3 reliable accumulate, 4 reliable residual, 5 p update, 6 final accumulate,
// 7 alpha and 8 the update's sums from the per-(site, RHS) values at bsite (a stencil's bmode 1 / 2)
bool launch_bcg_step(int which, int prec, int nr, BState* s, const BcgArgs& a, hipStream_t st);
bool batch_nr_supported(int nr);
void launch_pack(const zd* const* srcs_dev, int nr, zd* multi, long n, hipStream_t st);
void launch_unpack(const zd* multi, int nr, zd* const* dsts_dev, long n, hipStream_t st);
// lanes of nr-fields (rows x nr elements of esize bytes, lane fastest): copy or
// subtract on the lanes with mask[q] != 0; relayout to another width (a lane
// keeps its index; new lanes are zero)
void launch_lane_copy(int esize, void* dst, const void* src, long rows, int nr, const int* mask, hipStream_t st);
void launch_lane_sub(zd* out, const zd* a, const zd* b, long rows, int nr, const int* mask, hipStream_t st);
void launch_relayout(int esize, const void* in, void* out, long rows, int w_old, int w_new, hipStream_t st);
// zero the padding slots (sites vh..stride-1) of an nr-field of `groups` groups
// of stride * nr elements of esize (16, 8 or 4) bytes.  No launch writes them,
// and a buffer reused at another width would otherwise hold, there, values of
// real sites of the old layout, which per-RHS sums and copies read.
void launch_zero_padding(int esize, void* buf, long groups, int stride, int vh, int nr, hipStream_t st);
void launch_clover_apply_nr(const Geom& g, const zd* blocks, const zd* in, zd* out, int nr, hipStream_t st);
// the same with lane r's blocks at blocks + r * lane_elems, only for lanes with mask[r] != 0
// lane r applies block set lane_set[r] (lane_set NULL: set r)
void launch_clover_apply_lanes(const Geom& g, const zd* blocks, long lane_elems, const int* lane_set, const int* mask,
                               const zd* in,
                               zd* out, int nr, hipStream_t st);

// x = uniform [0,1) complex on real sites, zero on padding (BiCGStab start for a zero source)
void launch_fill_uniform(zd* x, int nparity, int stride, int Vh, uint64_t seed, hipStream_t st);
}  // namespace lym
