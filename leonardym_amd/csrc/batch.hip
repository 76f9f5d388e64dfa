// batch.hip -- multi-RHS mixed-precision CG with per-RHS reliable updates.
//
// NR right-hand sides share every operator apply (rhs-fastest fields, so one
// link load serves all NR lanes of a site) while each keeps its own CG
// recursion: alpha, beta, |r|^2, its reliable-update decisions, its tolerance
// and its stopping point.  A converged RHS is frozen.  Every per-RHS sum runs
// over the same site partition and in the same order as the single-RHS kernels
// of mixed.hip (thread per site, 512 x 256 grid-stride, fixed trees), and the
// per-site arithmetic is the same, so an RHS gives bit-identical results alone
// and in any batch.
#include "lym_device.h"

namespace lym {

// Every per-RHS sum runs over fixed chunks of BC consecutive sites: a chunk's
// sum is an LDS tree over its BC sites, and the final sum an LDS tree over
// per-thread strided sums of the chunk partials, with a fixed FT-thread
// block.  Neither tree depends on NR, so an RHS sums its terms in the same
// order in a batch of any size.
constexpr int BC = BATCH_CHUNK;  // sites per chunk
constexpr int FT = 1024;   // threads of the final-sum block

// chunk sum of v (one value per (site, lane) thread; block = BC sites x NR lanes)
template <int NR>
__device__ __forceinline__ void chunk_store(double v, double* partials, long nchunks) {
    __shared__ double red[NR * BC];
    const int t = threadIdx.x, sl = t / NR, q = t - sl * NR;
    red[q * BC + sl] = v;
    __syncthreads();
#pragma unroll
    for (int off = BC / 2; off > 0; off >>= 1) {
        if (sl < off) red[q * BC + sl] += red[q * BC + sl + off];
        __syncthreads();
    }
    if (sl == 0) partials[(long)q * nchunks + blockIdx.x] = red[q * BC];
}

// final per-RHS sums (valid in thread 0), one FT-thread block
template <int NR>
__device__ __forceinline__ void final_sums(const double* partials, long nchunks, double out[NR]) {
    __shared__ double sh[FT];
    for (int q = 0; q < NR; ++q) {
        double v = 0;
        for (long c = threadIdx.x; c < nchunks; c += FT) v += partials[(long)q * nchunks + c];
        sh[threadIdx.x] = v;
        __syncthreads();
#pragma unroll
        for (int off = FT / 2; off > 0; off >>= 1) {
            if ((int)threadIdx.x < off) sh[threadIdx.x] += sh[threadIdx.x + off];
            __syncthreads();
        }
        if (threadIdx.x == 0) out[q] = sh[0];
        __syncthreads();
    }
}

#define BSITE_LOOP(j, n) for (long j = (long)blockIdx.x * blockDim.x + threadIdx.x; j < (n); j += (long)gridDim.x * blockDim.x)

// (site, lane) of this thread in a chunked launch; false past the field
template <int NR>
__device__ __forceinline__ bool chunk_pair(long nsites, long& j, int& q) {
    const int t = threadIdx.x, sl = t / NR;
    q = t - sl * NR;
    j = (long)blockIdx.x * BC + sl;
    return j < nsites;
}

#define PAIR_LOOP(j, q, NR, nsites)                                                                        \
    for (long t_ = (long)blockIdx.x * blockDim.x + threadIdx.x, j, q; t_ < (nsites) * (NR) &&              \
                                                                  ((j = t_ / (NR)), (q = t_ - j * (NR)), 1); \
         t_ += (long)gridDim.x * blockDim.x)

// element (parity p, component k, site i, rhs r) of an nr-field: ((p*12 + k)*stride + i)*nr + r
__device__ __forceinline__ long bidx(int stride, int nr, long j, int k, int r) {
    const long p = j / stride, i = j % stride;
    return ((p * 12 + k) * stride + i) * nr + r;
}

// parity base of a low-precision nr-field and the lane element of (site j, rhs r)
template <int PR, int NR>
__device__ __forceinline__ long bs_parity_off(int stride, long j) {
    return (j / stride) * 12L * stride * NR;
}

template <int PR, int NR>
__device__ __forceinline__ void bs_load(const typename Prec<PR>::S* d, const float* sc, int stride, long j, int r,
                                        cplx<float> s[12]) {
    typename Prec<PR>::S e[12];
    load_packed<PR>(d + bs_parity_off<PR, NR>(stride, j), (size_t)stride * NR, (unsigned)((j % stride) * NR + r), e);
#pragma unroll
    for (int k = 0; k < 12; ++k) {
        cplx<typename Prec<PR>::T> z = to_c<PR>(e[k]);
        s[k] = mk((float)z.re, (float)z.im);
    }
    if constexpr (PR == 2) {
        const float f = sc[j * NR + r];
#pragma unroll
        for (int k = 0; k < 12; ++k) s[k] = f * s[k];
    }
}

template <int PR, int NR>
__device__ __forceinline__ void bs_store(typename Prec<PR>::S* d, float* sc, int stride, long j, int r,
                                         const cplx<float> s[12]) {
    if constexpr (PR == 2) {
        float m = 0.f;
#pragma unroll
        for (int k = 0; k < 12; ++k) m = fmaxf(m, fmaxf(fabsf(s[k].re), fabsf(s[k].im)));
        const float inv = fp16_site_inv(m);
        typename Prec<PR>::S e[12];
#pragma unroll
        for (int k = 0; k < 12; ++k) e[k] = from_c<PR>(inv * s[k]);
        store_packed<PR, false>(d + bs_parity_off<PR, NR>(stride, j), (size_t)stride * NR,
                                (unsigned)((j % stride) * NR + r), e);
        sc[j * NR + r] = m;
    } else {
        typename Prec<PR>::S e[12];
#pragma unroll
        for (int k = 0; k < 12; ++k) e[k] = from_c<PR>(s[k]);
        store_packed<PR, false>(d + bs_parity_off<PR, NR>(stride, j), (size_t)stride * NR,
                                (unsigned)((j % stride) * NR + r), e);
    }
}

template <class E, int NR>
__device__ __forceinline__ void bg_load(const E* d, int stride, long j, int r, E s[12]) {
#pragma unroll
    for (int k = 0; k < 12; ++k) s[k] = d[bidx(stride, NR, j, k, r)];
}
template <class E, int NR>
__device__ __forceinline__ void bg_store(E* d, int stride, long j, int r, const E s[12]) {
#pragma unroll
    for (int k = 0; k < 12; ++k) d[bidx(stride, NR, j, k, r)] = s[k];
}

// ---------------------------------------------------------------- kernels
// Sums: one thread per (site, RHS) pair, RHS fastest (coalesced like the
// fields), chunk_store / final_sums.  Element-wise kernels: PAIR_LOOP.
// mask (may be null: every lane): only the lanes with mask[q] != 0 start
template <int PR, int NR>
__global__ __launch_bounds__(BC * NR) void k_bcg_init(const zd* r, cplx<float>* rs, cplx<float>* xs,
                                                      typename Prec<PR>::S* p, float* p_sc, long nsites, int stride,
                                                      double* partials, long nchunks, const int* mask) {
    long j;
    int q;
    double v = 0;
    if (chunk_pair<NR>(nsites, j, q) && (!mask || mask[q])) {
        zd a[12];
        cplx<float> f[12], z[12];
        bg_load<zd, NR>(r, stride, j, q, a);
#pragma unroll
        for (int k = 0; k < 12; ++k) {
            v += a[k].re * a[k].re + a[k].im * a[k].im;
            f[k] = mk((float)a[k].re, (float)a[k].im);
            z[k] = mk(0.f, 0.f);
        }
        bg_store<cplx<float>, NR>(rs, stride, j, q, f);
        bg_store<cplx<float>, NR>(xs, stride, j, q, z);
        bs_store<PR, NR>(p, p_sc, stride, j, q, f);
    }
    chunk_store<NR>(v, partials, nchunks);
}

template <int NR>
__global__ __launch_bounds__(FT) void k_bcg_init_final(BState* s, const double* partials, long nchunks,
                                                       int max_steps, const int* mask) {
    double out[NR];
    final_sums<NR>(partials, nchunks, out);
    if (threadIdx.x == 0) {
        int all = 1;
        for (int q = 0; q < NR; ++q) {
            if (mask && !mask[q]) {  // a lane that keeps its state
                all = all && s->done[q];
                continue;
            }
            s->rr_s[q] = s->rr_true[q] = s->maxrr[q] = out[q];
            s->step[q] = 0;
            s->nrel[q] = 0;
            s->rel[q] = 0;
            s->alpha[q] = s->beta[q] = 0;
            // a zero source has the solution x = x0 = 0: nothing to iterate
            s->conv[q] = out[q] == 0.0 ? 1 : 0;
            s->done[q] = (max_steps <= 0 || out[q] == 0.0) ? 1 : 0;
            all = all && s->done[q];
        }
        s->max_steps = max_steps;
        s->all_done = all;
        // a lane that keeps its state may hold an update due at the next slot
        int any = 0;
        for (int q = 0; q < NR; ++q) any = any || (s->rel[q] && !s->done[q]);
        s->any_rel = any;
        s->skip_rel = any ? 0 : 1;
    }
}

template <int PR, int NR>
__global__ __launch_bounds__(BC * NR) void k_bcg_dot(const BState* s, const typename Prec<PR>::S* p,
                                                     const float* p_sc, const typename Prec<PR>::S* Ap,
                                                     const float* Ap_sc, long nsites, int stride, double* partials,
                                                     long nchunks) {
    if (s->all_done) return;
    long j;
    int q;
    double v = 0;
    if (chunk_pair<NR>(nsites, j, q) && !s->done[q]) {
        cplx<float> a[12], b[12];
        bs_load<PR, NR>(p, p_sc, stride, j, q, a);
        bs_load<PR, NR>(Ap, Ap_sc, stride, j, q, b);
#pragma unroll
        for (int k = 0; k < 12; ++k) v += (double)a[k].re * b[k].re + (double)a[k].im * b[k].im;
    }
    chunk_store<NR>(v, partials, nchunks);
}

// chunk sums of one value per (site, RHS) written by a fused stencil launch
// (bsite[j * NR + q]); lanes that are done contribute zero, as in k_bcg_dot.
// Only the vh real sites of each parity count: the stencil never writes the
// padding slots up to the stride, which may hold values a wider batch left.
template <int NR>
__global__ __launch_bounds__(BC * NR) void k_bcg_chunksum(const BState* s, const double* bsite, long nsites,
                                                          int stride, int vh, double* partials, long nchunks) {
    if (s->all_done) return;
    long j;
    int q;
    double v = 0;
    if (chunk_pair<NR>(nsites, j, q) && !s->done[q] && j % stride < vh) v = bsite[j * NR + q];
    chunk_store<NR>(v, partials, nchunks);
}

template <int NR>
__global__ __launch_bounds__(FT) void k_bcg_alpha_final(BState* s, const double* partials, long nchunks) {
    if (s->all_done) return;
    double out[NR];
    final_sums<NR>(partials, nchunks, out);
    if (threadIdx.x == 0)
        for (int q = 0; q < NR; ++q)
            if (!s->done[q]) s->alpha[q] = s->rr_s[q] / out[q];
}

template <int PR, int NR>
__global__ __launch_bounds__(BC * NR) void k_bcg_update(const BState* s, cplx<float>* xs, cplx<float>* rs,
                                                        const typename Prec<PR>::S* p, const float* p_sc,
                                                        const typename Prec<PR>::S* Ap, const float* Ap_sc,
                                                        long nsites, int stride, double* partials, long nchunks) {
    if (s->all_done) return;
    long j;
    int q;
    double v = 0;
    if (chunk_pair<NR>(nsites, j, q) && !s->done[q]) {
        const float al = (float)s->alpha[q];
        cplx<float> x[12], r[12], a[12], b[12];
        bg_load<cplx<float>, NR>(xs, stride, j, q, x);
        bg_load<cplx<float>, NR>(rs, stride, j, q, r);
        bs_load<PR, NR>(p, p_sc, stride, j, q, a);
        bs_load<PR, NR>(Ap, Ap_sc, stride, j, q, b);
#pragma unroll
        for (int k = 0; k < 12; ++k) {
            x[k] = x[k] + al * a[k];
            r[k] = r[k] - al * b[k];
            v += (double)r[k].re * r[k].re + (double)r[k].im * r[k].im;
        }
        bg_store<cplx<float>, NR>(xs, stride, j, q, x);
        bg_store<cplx<float>, NR>(rs, stride, j, q, r);
    }
    chunk_store<NR>(v, partials, nchunks);
}

template <int NR>
__global__ __launch_bounds__(FT) void k_bcg_update_final(BState* s, const double* partials, long nchunks) {
    if (s->all_done) return;
    double out[NR];
    final_sums<NR>(partials, nchunks, out);
    if (threadIdx.x == 0) {
        int any = 0;
        for (int q = 0; q < NR; ++q) {
            if (s->done[q]) continue;
            const double rn = out[q];
            if (rn > s->maxrr[q]) s->maxrr[q] = rn;
            // due reliable updates run at the next reliable-update slot
            if (rn < s->delta2 * s->maxrr[q] || rn < s->eps[q]) s->rel[q] = 1;
            s->beta[q] = rn / s->rr_s[q];
            s->rr_s[q] = rn;
            s->step[q] += 1;
            if (s->step[q] >= s->max_steps) s->done[q] = 1;
        }
        for (int q = 0; q < NR; ++q) any = any || (s->rel[q] && !s->done[q]);
        s->any_rel = any;
        s->skip_rel = any ? 0 : 1;
        int all = 1;
        for (int q = 0; q < NR; ++q) all = all && s->done[q];
        s->all_done = all;
    }
}

// reliable update, part 1, for the RHS with a due update: x += x_s, x_s = 0
template <int NR>
__global__ void k_brel_accum(const BState* s, zd* x, cplx<float>* xs, long nsites, int stride) {
    if (s->skip_rel) return;
    PAIR_LOOP(j, q, NR, nsites) {
        if (!s->rel[q]) continue;
        zd a[12];
        cplx<float> f[12];
        bg_load<zd, NR>(x, stride, j, (int)q, a);
        bg_load<cplx<float>, NR>(xs, stride, j, (int)q, f);
#pragma unroll
        for (int k = 0; k < 12; ++k) {
            a[k] = a[k] + mk((double)f[k].re, (double)f[k].im);
            f[k] = mk(0.f, 0.f);
        }
        bg_store<zd, NR>(x, stride, j, (int)q, a);
        bg_store<cplx<float>, NR>(xs, stride, j, (int)q, f);
    }
}

// reliable update, part 2: r = b - Ax (FP64), r_s = r, |r|^2
template <int NR>
__global__ __launch_bounds__(BC * NR) void k_brel_resid(const BState* s, zd* r, cplx<float>* rs, const zd* b,
                                                        const zd* Ax, long nsites, int stride, double* partials,
                                                        long nchunks) {
    if (s->skip_rel) return;
    long j;
    int q;
    double v = 0;
    if (chunk_pair<NR>(nsites, j, q) && s->rel[q]) {
        zd bb[12], ax[12];
        cplx<float> f[12];
        bg_load<zd, NR>(b, stride, j, q, bb);
        bg_load<zd, NR>(Ax, stride, j, q, ax);
#pragma unroll
        for (int k = 0; k < 12; ++k) {
            bb[k] = bb[k] - ax[k];
            v += bb[k].re * bb[k].re + bb[k].im * bb[k].im;
            f[k] = mk((float)bb[k].re, (float)bb[k].im);
        }
        bg_store<zd, NR>(r, stride, j, q, bb);
        bg_store<cplx<float>, NR>(rs, stride, j, q, f);
    }
    chunk_store<NR>(v, partials, nchunks);
}

template <int NR>
__global__ __launch_bounds__(FT) void k_brel_final(BState* s, const double* partials, long nchunks) {
    if (s->skip_rel) return;
    double out[NR];
    final_sums<NR>(partials, nchunks, out);
    if (threadIdx.x == 0) {
        for (int q = 0; q < NR; ++q) {
            if (!s->rel[q]) continue;
            const double rt = out[q];
            s->rr_true[q] = rt;
            s->nrel[q] += 1;
            s->rel[q] = 0;
            if (rt < s->eps[q]) {
                s->done[q] = 1;
                s->conv[q] = 1;
                continue;
            }
            s->rr_s[q] = rt;
            s->maxrr[q] = rt;
        }
        s->any_rel = 0;
        s->skip_rel = 1;
        int all = 1;
        for (int q = 0; q < NR; ++q) all = all && s->done[q];
        s->all_done = all;
    }
}

// p = r_s + beta p for the RHS still running
template <int PR, int NR>
__global__ void k_bcg_p(const BState* s, typename Prec<PR>::S* p, float* p_sc, const cplx<float>* rs, long nsites,
                        int stride) {
    if (s->all_done) return;
    PAIR_LOOP(j, q, NR, nsites) {
        if (s->done[q]) continue;
        const float be = (float)s->beta[q];
        cplx<float> a[12], r[12];
        bs_load<PR, NR>(p, p_sc, stride, j, (int)q, a);
        bg_load<cplx<float>, NR>(rs, stride, j, (int)q, r);
#pragma unroll
        for (int k = 0; k < 12; ++k) a[k] = r[k] + be * a[k];
        bs_store<PR, NR>(p, p_sc, stride, j, (int)q, a);
    }
}

// x += x_s for the RHS that stopped between reliable updates (mask may be null: every lane)
template <int NR>
__global__ void k_baccum(const BState* s, zd* x, const cplx<float>* xs, long nsites, int stride, const int* mask) {
    PAIR_LOOP(j, q, NR, nsites) {
        if (s->conv[q] || (mask && !mask[q])) continue;
        zd a[12];
        cplx<float> f[12];
        bg_load<zd, NR>(x, stride, j, (int)q, a);
        bg_load<cplx<float>, NR>(xs, stride, j, (int)q, f);
#pragma unroll
        for (int k = 0; k < 12; ++k) a[k] = a[k] + mk((double)f[k].re, (double)f[k].im);
        bg_store<zd, NR>(x, stride, j, (int)q, a);
    }
}

// pack / unpack single full fields <-> an nr-field (a null pointer leaves
// its lane alone); also multi-RHS clover block apply
__global__ void k_pack(const zd* const* srcs, int nr, zd* multi, long n) {
    BSITE_LOOP(i, n) {
        for (int r = 0; r < nr; ++r)
            if (srcs[r]) multi[i * nr + r] = srcs[r][i];
    }
}
__global__ void k_unpack(const zd* multi, int nr, zd* const* dsts, long n) {
    BSITE_LOOP(i, n) {
        for (int r = 0; r < nr; ++r)
            if (dsts[r]) dsts[r][i] = multi[i * nr + r];
    }
}

// ---------------------------------------------------------------- lanes of an nr-field
// An nr-field is rows x nr elements, lane fastest.  These act on the lanes
// with mask[q] != 0 only, or move every lane to another width.
template <class E>
__global__ void k_lane_copy(E* dst, const E* src, long rows, int nr, const int* mask) {
    BSITE_LOOP(t, rows * nr) {
        if (mask[t % nr]) dst[t] = src[t];
    }
}
__global__ void k_lane_sub(zd* out, const zd* a, const zd* b, long rows, int nr, const int* mask) {
    BSITE_LOOP(t, rows * nr) {
        if (mask[t % nr]) out[t] = a[t] - b[t];
    }
}
// out (width w_new) from in (width w_old): lane j keeps its index, lanes
// j >= w_old are zero
template <class E>
__global__ void k_relayout(const E* in, E* out, long rows, int w_old, int w_new) {
    BSITE_LOOP(t, rows * w_new) {
        const long i = t / w_new;
        const int j = (int)(t - i * w_new);
        E e;
        if (j < w_old) e = in[i * w_old + j];
        else __builtin_memset(&e, 0, sizeof(E));
        out[t] = e;
    }
}

__global__ void k_clover_apply_nr(Geom g, const zd* __restrict__ blocks, const zd* __restrict__ in,
                                  zd* __restrict__ out, int nr) {
    long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (long)g.Vh * nr) return;
    const int idx = (int)(t / nr), r = (int)(t % nr);
    zd s[12], o[12];
#pragma unroll
    for (int k = 0; k < 12; ++k) s[k] = in[((long)k * g.stride + idx) * nr + r];
    herm6(blocks, g.stride, idx, s, o);
    herm6(blocks + (size_t)18 * g.stride, g.stride, idx, s + 6, o + 6);
#pragma unroll
    for (int k = 0; k < 12; ++k) out[((long)k * g.stride + idx) * nr + r] = o[k];
}

__global__ void k_clover_apply_lanes(Geom g, const zd* __restrict__ blocks, long lane_elems, const int* lane_set,
                                     const int* mask, const zd* __restrict__ in, zd* __restrict__ out, int nr) {
    long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (long)g.Vh * nr) return;
    const int idx = (int)(t / nr), r = (int)(t % nr);
    if (!mask[r]) return;
    const zd* B = blocks + (lane_set ? lane_set[r] : r) * lane_elems;
    zd s[12], o[12];
#pragma unroll
    for (int k = 0; k < 12; ++k) s[k] = in[((long)k * g.stride + idx) * nr + r];
    herm6(B, g.stride, idx, s, o);
    herm6(B + (size_t)18 * g.stride, g.stride, idx, s + 6, o + 6);
#pragma unroll
    for (int k = 0; k < 12; ++k) out[((long)k * g.stride + idx) * nr + r] = o[k];
}

// ---------------------------------------------------------------- launchers
static inline int bsite_blocks(long n) {
    long b = (n + 255) / 256;
    return (int)(b < 4096 ? b : 4096);
}

static inline int pair_blocks(long pairs) {
    long b = (pairs + 255) / 256;
    return (int)(b < 8192 ? b : 8192);
}

template <int PR, int NR>
static void bcg_step_impl(int which, BState* s, const BcgArgs& a, hipStream_t st) {
    typedef typename Prec<PR>::S S;
    S* p = reinterpret_cast<S*>(a.p);
    S* Ap = reinterpret_cast<S*>(a.Ap);
    const long nch = (a.nsites + BC - 1) / BC;
    const dim3 cg((unsigned)nch), cb(BC * NR), fb(FT);
    const int eb = pair_blocks(a.nsites * NR);
    switch (which) {
        case 0:
            hipLaunchKernelGGL((k_bcg_init<PR, NR>), cg, cb, 0, st, a.r, a.rs, a.xs, p, a.p_sc, a.nsites, a.stride,
                               a.partials, nch, a.mask);
            hipLaunchKernelGGL((k_bcg_init_final<NR>), dim3(1), fb, 0, st, s, a.partials, nch, a.max_steps, a.mask);
            break;
        case 1:
            hipLaunchKernelGGL((k_bcg_dot<PR, NR>), cg, cb, 0, st, s, p, a.p_sc, Ap, a.Ap_sc, a.nsites, a.stride,
                               a.partials, nch);
            hipLaunchKernelGGL((k_bcg_alpha_final<NR>), dim3(1), fb, 0, st, s, a.partials, nch);
            break;
        case 2:
            hipLaunchKernelGGL((k_bcg_update<PR, NR>), cg, cb, 0, st, s, a.xs, a.rs, p, a.p_sc, Ap, a.Ap_sc, a.nsites,
                               a.stride, a.partials, nch);
            hipLaunchKernelGGL((k_bcg_update_final<NR>), dim3(1), fb, 0, st, s, a.partials, nch);
            break;
        case 3:
            hipLaunchKernelGGL((k_brel_accum<NR>), dim3(eb), dim3(256), 0, st, s, a.x, a.xs, a.nsites, a.stride);
            break;
        case 4:
            hipLaunchKernelGGL((k_brel_resid<NR>), cg, cb, 0, st, s, a.r, a.rs, a.b, a.Ax, a.nsites, a.stride,
                               a.partials, nch);
            hipLaunchKernelGGL((k_brel_final<NR>), dim3(1), fb, 0, st, s, a.partials, nch);
            break;
        case 5:
            hipLaunchKernelGGL((k_bcg_p<PR, NR>), dim3(eb), dim3(256), 0, st, s, p, a.p_sc, a.rs, a.nsites, a.stride);
            break;
        case 7:
            hipLaunchKernelGGL((k_bcg_chunksum<NR>), cg, cb, 0, st, s, a.bsite, a.nsites, a.stride, a.vh, a.partials,
                               nch);
            hipLaunchKernelGGL((k_bcg_alpha_final<NR>), dim3(1), fb, 0, st, s, a.partials, nch);
            break;
        case 8:
            hipLaunchKernelGGL((k_bcg_chunksum<NR>), cg, cb, 0, st, s, a.bsite, a.nsites, a.stride, a.vh, a.partials,
                               nch);
            hipLaunchKernelGGL((k_bcg_update_final<NR>), dim3(1), fb, 0, st, s, a.partials, nch);
            break;
        default:
            hipLaunchKernelGGL((k_baccum<NR>), dim3(eb), dim3(256), 0, st, s, a.x, a.xs, a.nsites, a.stride, a.mask);
            break;
    }
}

template <int PR>
static bool bcg_step_pr(int which, int nr, BState* s, const BcgArgs& a, hipStream_t st) {
    switch (nr) {
        case 1: bcg_step_impl<PR, 1>(which, s, a, st); return true;
        case 2: bcg_step_impl<PR, 2>(which, s, a, st); return true;
        case 3: bcg_step_impl<PR, 3>(which, s, a, st); return true;
        case 4: bcg_step_impl<PR, 4>(which, s, a, st); return true;
        case 6: bcg_step_impl<PR, 6>(which, s, a, st); return true;
        case 8: bcg_step_impl<PR, 8>(which, s, a, st); return true;
        case 12: bcg_step_impl<PR, 12>(which, s, a, st); return true;
        case 16: bcg_step_impl<PR, 16>(which, s, a, st); return true;
        default: return false;
    }
}

bool launch_bcg_step(int which, int prec, int nr, BState* s, const BcgArgs& a, hipStream_t st) {
    return prec == 1 ? bcg_step_pr<1>(which, nr, s, a, st) : bcg_step_pr<2>(which, nr, s, a, st);
}

bool batch_nr_supported(int nr) { return nr == 1 || nr == 2 || nr == 3 || nr == 4 || nr == 6 || nr == 8 || nr == 12 || nr == 16; }

void launch_pack(const zd* const* srcs_dev, int nr, zd* multi, long n, hipStream_t st) {
    hipLaunchKernelGGL(k_pack, dim3(bsite_blocks(n)), dim3(256), 0, st, srcs_dev, nr, multi, n);
}
void launch_unpack(const zd* multi, int nr, zd* const* dsts_dev, long n, hipStream_t st) {
    hipLaunchKernelGGL(k_unpack, dim3(bsite_blocks(n)), dim3(256), 0, st, multi, nr, dsts_dev, n);
}

void launch_lane_copy(int esize, void* dst, const void* src, long rows, int nr, const int* mask, hipStream_t st) {
    const dim3 g(bsite_blocks(rows * nr)), b(256);
    switch (esize) {
        case 16: hipLaunchKernelGGL(k_lane_copy<zd>, g, b, 0, st, (zd*)dst, (const zd*)src, rows, nr, mask); break;
        case 8:
            hipLaunchKernelGGL(k_lane_copy<cplx<float>>, g, b, 0, st, (cplx<float>*)dst, (const cplx<float>*)src,
                               rows, nr, mask);
            break;
        default:
            hipLaunchKernelGGL(k_lane_copy<float>, g, b, 0, st, (float*)dst, (const float*)src, rows, nr, mask);
            break;
    }
}
void launch_lane_sub(zd* out, const zd* a, const zd* b, long rows, int nr, const int* mask, hipStream_t st) {
    hipLaunchKernelGGL(k_lane_sub, dim3(bsite_blocks(rows * nr)), dim3(256), 0, st, out, a, b, rows, nr, mask);
}
void launch_relayout(int esize, const void* in, void* out, long rows, int w_old, int w_new, hipStream_t st) {
    const dim3 g(bsite_blocks(rows * w_new)), b(256);
    switch (esize) {
        case 16:
            hipLaunchKernelGGL(k_relayout<V16>, g, b, 0, st, (const V16*)in, (V16*)out, rows, w_old, w_new);
            break;
        case 8:
            hipLaunchKernelGGL(k_relayout<cplx<float>>, g, b, 0, st, (const cplx<float>*)in, (cplx<float>*)out, rows,
                               w_old, w_new);
            break;
        default:
            hipLaunchKernelGGL(k_relayout<float>, g, b, 0, st, (const float*)in, (float*)out, rows, w_old, w_new);
            break;
    }
}
// padding slots of an nr-field: sites vh..stride-1 of each of `groups` groups
// of stride * nr elements (lane fastest)
template <class E>
__global__ void k_zero_padding(E* buf, long groups, int stride, int vh, int nr) {
    const long per = (long)(stride - vh) * nr;
    const E z{};
    for (long t = (long)blockIdx.x * blockDim.x + threadIdx.x; t < groups * per; t += (long)gridDim.x * blockDim.x) {
        const long g = t / per;
        buf[(g * stride + vh) * nr + (t - g * per)] = z;
    }
}
void launch_zero_padding(int esize, void* buf, long groups, int stride, int vh, int nr, hipStream_t st) {
    if (!buf || stride <= vh) return;
    const dim3 g(bsite_blocks(groups * (stride - vh) * nr)), b(256);
    switch (esize) {
        case 16: hipLaunchKernelGGL(k_zero_padding<V16>, g, b, 0, st, (V16*)buf, groups, stride, vh, nr); break;
        case 8:
            hipLaunchKernelGGL(k_zero_padding<cplx<float>>, g, b, 0, st, (cplx<float>*)buf, groups, stride, vh, nr);
            break;
        default: hipLaunchKernelGGL(k_zero_padding<float>, g, b, 0, st, (float*)buf, groups, stride, vh, nr); break;
    }
}
void launch_clover_apply_lanes(const Geom& g, const zd* blocks, long lane_elems, const int* lane_set, const int* mask,
                               const zd* in, zd* out, int nr, hipStream_t st) {
    long n = (long)g.Vh * nr;
    hipLaunchKernelGGL(k_clover_apply_lanes, dim3((n + 255) / 256), dim3(256), 0, st, g, blocks, lane_elems, lane_set,
                       mask, in, out, nr);
}
void launch_clover_apply_nr(const Geom& g, const zd* blocks, const zd* in, zd* out, int nr, hipStream_t st) {
    long n = (long)g.Vh * nr;
    hipLaunchKernelGGL(k_clover_apply_nr, dim3((n + 255) / 256), dim3(256), 0, st, g, blocks, in, out, nr);
}

}  // namespace lym
