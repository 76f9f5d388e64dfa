// mixed.hip -- low-precision copies of the operator data and the kernels of the
// mixed-precision CG with reliable updates.
//
// The sloppy operator computes in FP32 on FP32 or FP16 storage (FP16 spinors
// and clover blocks carry one FP32 scale per site; links are unitary or
// orthogonal and are stored unscaled).  The Krylov recursion keeps x_s, r_s in
// FP32 and p, Ap in the sloppy storage; every dot product is accumulated in
// FP64.  When the sloppy residual has dropped by `delta` since the last reliable
// update (or looks converged), x += x_s in FP64 and the true residual
// r = b - A x is recomputed with the FP64 operator; convergence is declared
// only on that true residual, so the FP64 stopping rule |r|^2 < eps of the
// reference CG (source/inverters/ConjugateGradient.cpp:49-50) holds exactly.
// The search direction p is kept across updates (reliable updates in the sense
// of Sleijpen and van der Vorst, 1996).
#include "lym_device.h"

namespace lym {

constexpr int MR_THREADS = 256;

template <int NV>
__device__ __forceinline__ void m_block_store(double v[NV], double* partials) {
    __shared__ double sh[NV][MR_THREADS / 64];
#pragma unroll
    for (int off = 32; off > 0; off >>= 1)
#pragma unroll
        for (int k = 0; k < NV; ++k) v[k] += __shfl_xor(v[k], off, 64);
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    if (lane == 0)
#pragma unroll
        for (int k = 0; k < NV; ++k) sh[k][w] = v[k];
    __syncthreads();
    if (threadIdx.x == 0)
#pragma unroll
        for (int k = 0; k < NV; ++k) {
            double s = 0;
#pragma unroll
            for (int j = 0; j < MR_THREADS / 64; ++j) s += sh[k][j];
            partials[k * RED_BLOCKS + blockIdx.x] = s;
        }
}

template <int NV>
__device__ __forceinline__ void m_final_sum(const double* partials, double out[NV]) {
    __shared__ double sh[NV][RED_BLOCKS / 64];
    double v[NV];
#pragma unroll
    for (int k = 0; k < NV; ++k) v[k] = partials[k * RED_BLOCKS + threadIdx.x];
#pragma unroll
    for (int off = 32; off > 0; off >>= 1)
#pragma unroll
        for (int k = 0; k < NV; ++k) v[k] += __shfl_xor(v[k], off, 64);
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    if (lane == 0)
#pragma unroll
        for (int k = 0; k < NV; ++k) sh[k][w] = v[k];
    __syncthreads();
    if (threadIdx.x == 0)
#pragma unroll
        for (int k = 0; k < NV; ++k) {
            double s = 0;
#pragma unroll
            for (int j = 0; j < RED_BLOCKS / 64; ++j) s += sh[k][j];
            out[k] = s;
        }
}

// The same total as m_final_sum, folded by a workgroup of MR_THREADS threads:
// every workgroup that calls it gets the identical value.  Split in two so a
// kernel can issue the partial loads, then its first streaming loads, and only
// then wait for the partials: m_partials_load gives each thread its share (in
// a fixed order), m_fold_block folds the shares over the workgroup (result in
// every thread).
constexpr int PART_PER_THREAD = RED_BLOCKS / MR_THREADS;
static_assert(RED_BLOCKS % MR_THREADS == 0, "partials per thread");
__device__ __forceinline__ void m_partials_load(const double* partials, double pl[PART_PER_THREAD]) {
#pragma unroll
    for (int i = 0; i < PART_PER_THREAD; ++i) pl[i] = partials[threadIdx.x + i * MR_THREADS];
}
__device__ __forceinline__ double m_fold_block(const double pl[PART_PER_THREAD]) {
    __shared__ double sh[MR_THREADS / 64];
    double v = 0;
#pragma unroll
    for (int i = 0; i < PART_PER_THREAD; ++i) v += pl[i];
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
    if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = v;
    __syncthreads();
    double t = 0;
#pragma unroll
    for (int w = 0; w < MR_THREADS / 64; ++w) t += sh[w];
    __syncthreads();
    return t;
}

// The same for n <= NORM_FOLD_MAX partials (the stencil's per-workgroup
// output norms: one per workgroup of its launch); slots past n count as zero,
// so n = RED_BLOCKS gives m_fold_block's total.
constexpr int NORM_PER_THREAD = NORM_FOLD_MAX / MR_THREADS;
static_assert(NORM_FOLD_MAX % MR_THREADS == 0, "norm partials per thread");
__device__ __forceinline__ void m_partials_load_n(const double* partials, int n, double pl[NORM_PER_THREAD]) {
#pragma unroll
    for (int i = 0; i < NORM_PER_THREAD; ++i) {
        const int k = (int)threadIdx.x + i * MR_THREADS;
        pl[i] = k < n ? partials[k] : 0.0;
    }
}
__device__ __forceinline__ double m_fold_block_n(const double pl[NORM_PER_THREAD]) {
    __shared__ double sh[MR_THREADS / 64];
    double v = 0;
#pragma unroll
    for (int i = 0; i < NORM_PER_THREAD; ++i) v += pl[i];
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
    if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = v;
    __syncthreads();
    double t = 0;
#pragma unroll
    for (int w = 0; w < MR_THREADS / 64; ++w) t += sh[w];
    __syncthreads();
    return t;
}

#define SITE_LOOP(j, n) for (long j = (long)blockIdx.x * blockDim.x + threadIdx.x; j < (n); j += (long)gridDim.x * blockDim.x)

// A sloppy spinor vector over nsites = nparity * stride sites: site j lives in
// parity j / stride at index j % stride.
template <int PR>
struct SVec {
    typename Prec<PR>::S* d;  // [nparity][12][stride]
    float* sc;                // [nparity][stride] (FP16)
};

template <int PR>
__device__ __forceinline__ void sv_load(const SVec<PR>& v, int stride, long j, cplx<float> s[12]) {
    const long p = j / stride, i = j % stride;
    typename Prec<PR>::S e[12];
    load_packed<PR>(v.d + p * 12 * stride, (size_t)stride, (unsigned)i, e);
#pragma unroll
    for (int k = 0; k < 12; ++k) {
        cplx<typename Prec<PR>::T> z = to_c<PR>(e[k]);
        s[k] = mk((float)z.re, (float)z.im);
    }
    if constexpr (PR == 2) {
        const float sc = v.sc[j];
#pragma unroll
        for (int k = 0; k < 12; ++k) s[k] = sc * s[k];
    }
}

template <int PR>
__device__ __forceinline__ void sv_store(const SVec<PR>& v, int stride, long j, const cplx<float> s[12]) {
    const long p = j / stride, i = j % stride;
    typename Prec<PR>::S* b = v.d + p * 12 * stride;
    typename Prec<PR>::S e[12];
    if constexpr (PR == 2) {
        float m = 0.f;
#pragma unroll
        for (int k = 0; k < 12; ++k) m = fmaxf(m, fmaxf(fabsf(s[k].re), fabsf(s[k].im)));
        const float inv = fp16_site_inv(m);
#pragma unroll
        for (int k = 0; k < 12; ++k) e[k] = from_c<PR>(inv * s[k]);
        store_packed<PR, false>(b, (size_t)stride, (unsigned)i, e);
        v.sc[j] = m;
    } else {
#pragma unroll
        for (int k = 0; k < 12; ++k) e[k] = from_c<PR>(s[k]);
        store_packed<PR, false>(b, (size_t)stride, (unsigned)i, e);
    }
}

__device__ __forceinline__ void f_load(const cplx<float>* v, int stride, long j, cplx<float> s[12]) {
    const long p = j / stride, i = j % stride;
#pragma unroll
    for (int k = 0; k < 12; ++k) s[k] = v[p * 12 * stride + (long)k * stride + i];
}
__device__ __forceinline__ void f_store(cplx<float>* v, int stride, long j, const cplx<float> s[12]) {
    const long p = j / stride, i = j % stride;
#pragma unroll
    for (int k = 0; k < 12; ++k) v[p * 12 * stride + (long)k * stride + i] = s[k];
}
__device__ __forceinline__ void d_load(const zd* v, int stride, long j, zd s[12]) {
    const long p = j / stride, i = j % stride;
#pragma unroll
    for (int k = 0; k < 12; ++k) s[k] = v[p * 12 * stride + (long)k * stride + i];
}
__device__ __forceinline__ void d_store(zd* v, int stride, long j, const zd s[12]) {
    const long p = j / stride, i = j % stride;
#pragma unroll
    for (int k = 0; k < 12; ++k) v[p * 12 * stride + (long)k * stride + i] = s[k];
}

// ---------------------------------------------------------------- operator data conversion
template <int PR>
__global__ void k_convert_links(const void* src, void* dst, long n, int real_links) {
    SITE_LOOP(i, n) {
        if (real_links) {
            reinterpret_cast<typename Prec<PR>::R*>(dst)[i] =
                (typename Prec<PR>::R)reinterpret_cast<const double*>(src)[i];
        } else {
            zd z = reinterpret_cast<const zd*>(src)[i];
            reinterpret_cast<typename Prec<PR>::S*>(dst)[i] =
                from_c<PR>(mk((typename Prec<PR>::T)z.re, (typename Prec<PR>::T)z.im));
        }
    }
}

// FP64 links [2][4][nc][stride] (nc = 9 or 6; entries 0..5 are rows 0 and 1)
// -> rows 0, 1 in the packed low-precision layout (lym_device.h)
template <int PR, bool REAL>
__global__ void k_pack_links(const void* src, int nc, void* dst, int stride) {
    typedef PackedLink<PR, REAL> P;
    typedef typename P::E E;
    const long n = 8L * stride;
    SITE_LOOP(j, n) {
        const long plane = j / stride, i = j % stride;
        E e[6];
#pragma unroll
        for (int k = 0; k < 6; ++k) {
            const long o = (plane * nc + k) * stride + i;
            if constexpr (REAL) e[k] = (E)reinterpret_cast<const double*>(src)[o];
            else {
                const zd z = reinterpret_cast<const zd*>(src)[o];
                e[k] = from_c<PR>(mk((typename Prec<PR>::T)z.re, (typename Prec<PR>::T)z.im));
            }
        }
        unsigned char buf[P::BYTES];
        __builtin_memcpy(buf, e, P::BYTES);
        char* gb = reinterpret_cast<char*>(dst);
        int off = 0;
        for (int g = 0; g < P::N16; ++g) {
            __builtin_memcpy(gb + (size_t)j * 16, buf + off, 16);
            off += 16;
            gb += (size_t)n * 16;
        }
        if constexpr (P::H8) {
            __builtin_memcpy(gb + (size_t)j * 8, buf + off, 8);
            off += 8;
            gb += (size_t)n * 8;
        }
        if constexpr (P::H4) __builtin_memcpy(gb + (size_t)j * 4, buf + off, 4);
    }
}

// largest |U_2j - rebuilt U_2j| over the 18-real (9-real) FP64 links: row 2
// rebuilt as conj(row0 x row1) (complex) or row0 x row1 (real); one maximum
// per workgroup into partials[blockIdx.x]
__global__ __launch_bounds__(MR_THREADS) void k_link_row2_err(const void* src, int real_links, int stride,
                                                              double* partials) {
    double m = 0;
    const long n = 8L * stride;
    SITE_LOOP(j, n) {
        const long plane = j / stride, i = j % stride;
        if (real_links) {
            const double* u = reinterpret_cast<const double*>(src) + plane * 9 * stride + i;
            double a[9];
            for (int k = 0; k < 9; ++k) a[k] = u[(long)k * stride];
            const double r6 = a[1] * a[5] - a[2] * a[4], r7 = a[2] * a[3] - a[0] * a[5], r8 = a[0] * a[4] - a[1] * a[3];
            m = fmax(m, fmax(fabs(a[6] - r6), fmax(fabs(a[7] - r7), fabs(a[8] - r8))));
        } else {
            const zd* u = reinterpret_cast<const zd*>(src) + plane * 9 * stride + i;
            zd a[9];
            for (int k = 0; k < 9; ++k) a[k] = u[(long)k * stride];
            const zd r[3] = {conj(a[1] * a[5] - a[2] * a[4]), conj(a[2] * a[3] - a[0] * a[5]),
                             conj(a[0] * a[4] - a[1] * a[3])};
            for (int k = 0; k < 3; ++k) m = fmax(m, fmax(fabs(a[6 + k].re - r[k].re), fabs(a[6 + k].im - r[k].im)));
        }
    }
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) m = fmax(m, __shfl_xor(m, off, 64));
    __shared__ double sh[MR_THREADS / 64];
    if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = m;
    __syncthreads();
    if (threadIdx.x == 0) {
        double t = 0;
        for (int w = 0; w < MR_THREADS / 64; ++w) t = fmax(t, sh[w]);
        partials[blockIdx.x] = t;
    }
}

// blocks [nparity][36][stride] -> PR storage with one scale per site
// blocks [nparity][36][stride] (FP64) -> the packed low-precision layout of
// lym_device.h (planes 2 p + block), FP16 with one scale per site
template <int PR>
__global__ void k_convert_blocks(const zd* src, void* dst, float* sc, int nparity, int stride) {
    typedef PackedBlock<PR> P;
    const long ngroup = 2L * nparity * stride;
    SITE_LOOP(j, (long)nparity * stride) {
        const long p = j / stride, i = j % stride;
        const zd* s = src + p * 36 * stride + i;
        float m = 1.f;
        if constexpr (PR == 2) {
            m = 0.f;
            for (int k = 0; k < 36; ++k) {
                zd z = s[(long)k * stride];
                m = fmaxf(m, fmaxf(fabsf((float)z.re), fabsf((float)z.im)));
            }
        }
        const float inv = fp16_site_inv(m);
        if constexpr (PR == 2) sc[j] = m;
        for (int blk = 0; blk < 2; ++blk) {
            typename P::E e[18];
            for (int k = 0; k < 18; ++k) {
                zd z = s[(long)(18 * blk + k) * stride];
                e[k] = from_c<PR>(mk((typename Prec<PR>::T)(inv * z.re), (typename Prec<PR>::T)(inv * z.im)));
            }
            unsigned char buf[P::BYTES];
            __builtin_memcpy(buf, e, P::BYTES);
            const long el = (2 * p + blk) * stride + i;
            char* gb = reinterpret_cast<char*>(dst);
            int off = 0;
            for (int g = 0; g < P::N16; ++g) {
                __builtin_memcpy(gb + el * 16, buf + off, 16);
                off += 16;
                gb += ngroup * 16;
            }
            if constexpr (P::H8) __builtin_memcpy(gb + el * 8, buf + off, 8);
        }
    }
}

static inline int site_blocks(long n) {
    long b = (n + 255) / 256;
    return (int)(b < 4096 ? b : 4096);
}

void launch_convert_links(int prec, const void* src, void* dst, long n, int real_links, hipStream_t st) {
    if (prec == 1) hipLaunchKernelGGL(k_convert_links<1>, dim3(site_blocks(n)), dim3(256), 0, st, src, dst, n, real_links);
    else hipLaunchKernelGGL(k_convert_links<2>, dim3(site_blocks(n)), dim3(256), 0, st, src, dst, n, real_links);
}

void launch_pack_links(int prec, const void* src, int nc, int real_links, void* dst, int stride, hipStream_t st) {
    const long n = 8L * stride;
    if (prec == 1) {
        if (real_links) hipLaunchKernelGGL((k_pack_links<1, true>), dim3(site_blocks(n)), dim3(256), 0, st, src, nc, dst, stride);
        else hipLaunchKernelGGL((k_pack_links<1, false>), dim3(site_blocks(n)), dim3(256), 0, st, src, nc, dst, stride);
    } else {
        if (real_links) hipLaunchKernelGGL((k_pack_links<2, true>), dim3(site_blocks(n)), dim3(256), 0, st, src, nc, dst, stride);
        else hipLaunchKernelGGL((k_pack_links<2, false>), dim3(site_blocks(n)), dim3(256), 0, st, src, nc, dst, stride);
    }
}

void launch_link_row2_err(const void* src, int real_links, int stride, double* partials, hipStream_t st) {
    hipLaunchKernelGGL(k_link_row2_err, dim3(RED_BLOCKS), dim3(MR_THREADS), 0, st, src, real_links, stride, partials);
}

void launch_convert_blocks(int prec, const zd* src, void* dst, float* sc, int nparity, int stride, hipStream_t st) {
    long n = (long)nparity * stride;
    if (prec == 1)
        hipLaunchKernelGGL(k_convert_blocks<1>, dim3(site_blocks(n)), dim3(256), 0, st, src, dst, sc, nparity, stride);
    else
        hipLaunchKernelGGL(k_convert_blocks<2>, dim3(site_blocks(n)), dim3(256), 0, st, src, dst, sc, nparity, stride);
}

// ---------------------------------------------------------------- spinor precision conversion
// FP64 field -> sloppy field (rounded to the storage precision, FP16 with its
// per-site scale) and back, over nsites = nparity * stride sites
template <int PR>
__global__ void k_spinor_to_sloppy(const zd* src, SVec<PR> dst, long nsites, int stride) {
    SITE_LOOP(j, nsites) {
        zd a[12];
        cplx<float> f[12];
        d_load(src, stride, j, a);
#pragma unroll
        for (int k = 0; k < 12; ++k) f[k] = mk((float)a[k].re, (float)a[k].im);
        sv_store<PR>(dst, stride, j, f);
    }
}
template <int PR>
__global__ void k_spinor_from_sloppy(SVec<PR> src, zd* dst, long nsites, int stride) {
    SITE_LOOP(j, nsites) {
        cplx<float> f[12];
        zd a[12];
        sv_load<PR>(src, stride, j, f);
#pragma unroll
        for (int k = 0; k < 12; ++k) a[k] = mk((double)f[k].re, (double)f[k].im);
        d_store(dst, stride, j, a);
    }
}

void launch_spinor_to_sloppy(int prec, const zd* src, void* dst, float* dst_sc, long nsites, int stride,
                             hipStream_t st) {
    if (prec == 1)
        hipLaunchKernelGGL(k_spinor_to_sloppy<1>, dim3(site_blocks(nsites)), dim3(256), 0, st, src,
                           SVec<1>{reinterpret_cast<cplx<float>*>(dst), nullptr}, nsites, stride);
    else
        hipLaunchKernelGGL(k_spinor_to_sloppy<2>, dim3(site_blocks(nsites)), dim3(256), 0, st, src,
                           SVec<2>{reinterpret_cast<h2_t*>(dst), dst_sc}, nsites, stride);
}
void launch_spinor_from_sloppy(int prec, const void* src, const float* src_sc, zd* dst, long nsites, int stride,
                               hipStream_t st) {
    if (prec == 1)
        hipLaunchKernelGGL(k_spinor_from_sloppy<1>, dim3(site_blocks(nsites)), dim3(256), 0, st,
                           SVec<1>{reinterpret_cast<cplx<float>*>(const_cast<void*>(src)), nullptr}, dst, nsites,
                           stride);
    else
        hipLaunchKernelGGL(k_spinor_from_sloppy<2>, dim3(site_blocks(nsites)), dim3(256), 0, st,
                           SVec<2>{reinterpret_cast<h2_t*>(const_cast<void*>(src)), const_cast<float*>(src_sc)},
                           dst, nsites, stride);
}

// ---------------------------------------------------------------- mixed CG
// r_s = r (FP32), p = r (sloppy), x_s = 0, partial |r|^2
template <int PR>
__global__ __launch_bounds__(MR_THREADS) void k_mcg_init(const zd* r, cplx<float>* rs, cplx<float>* xs, SVec<PR> p,
                                                         long nsites, int stride, double* partials) {
    double v[1] = {0};
    SITE_LOOP(j, nsites) {
        zd a[12];
        cplx<float> f[12], z[12];
        d_load(r, stride, j, a);
#pragma unroll
        for (int k = 0; k < 12; ++k) {
            v[0] += a[k].re * a[k].re + a[k].im * a[k].im;
            f[k] = mk((float)a[k].re, (float)a[k].im);
            z[k] = mk(0.f, 0.f);
        }
        f_store(rs, stride, j, f);
        f_store(xs, stride, j, z);
        sv_store<PR>(p, stride, j, f);
    }
    m_block_store<1>(v, partials);
}

__global__ __launch_bounds__(RED_BLOCKS) void k_mcg_init_final(MCGState* s, const double* partials, double eps,
                                                               double delta, int max_steps) {
    double out[1];
    m_final_sum<1>(partials, out);
    if (threadIdx.x == 0) {
        s->rr_s = out[0];
        s->rr_true = out[0];
        s->maxrr = out[0];
        s->eps = eps;
        s->delta2 = delta * delta;
        s->max_steps = max_steps;
        s->step = 0;
        s->nrel = 0;
        s->rel = 0;
        s->skip_rel = 1;
        // a zero initial residual (zero source, or a guess that solves exactly)
        // is converged at step 0: the first alpha would be 0/0
        s->converged = out[0] == 0.0 ? 1 : 0;
        s->done = (max_steps <= 0 || out[0] == 0.0) ? 1 : 0;
        s->alpha = s->beta = 0;
    }
}

// <p, Ap> accumulated in FP64
template <int PR>
__global__ __launch_bounds__(MR_THREADS) void k_mcg_dot(const MCGState* s, SVec<PR> p, SVec<PR> Ap, long nsites,
                                                        int stride, double* partials) {
    if (s->done) return;
    double v[1] = {0};
    SITE_LOOP(j, nsites) {
        cplx<float> a[12], b[12];
        sv_load<PR>(p, stride, j, a);
        sv_load<PR>(Ap, stride, j, b);
#pragma unroll
        for (int k = 0; k < 12; ++k) v[0] += (double)a[k].re * b[k].re + (double)a[k].im * b[k].im;
    }
    m_block_store<1>(v, partials);
}

// alpha = rr_s / <p, Ap>, then x_s += alpha p, r_s -= alpha Ap, partial |r_s|^2.
// Every workgroup folds the ndot <p, Ap> partials itself (in one fixed order,
// so all get the same alpha) instead of a separate one-workgroup kernel.
template <int PR>
__global__ __launch_bounds__(MR_THREADS) void k_mcg_update(MCGState* s, cplx<float>* xs, cplx<float>* rs,
                                                           SVec<PR> p, SVec<PR> Ap, long nsites, int stride,
                                                           const double* dot_partials, int ndot, double* partials) {
    if (s->done) return;
    double pl[NORM_PER_THREAD];
    m_partials_load_n(dot_partials, ndot, pl);
    // the first site's vectors load while the partials are folded
    long j = (long)blockIdx.x * blockDim.x + threadIdx.x;
    const long jstep = (long)gridDim.x * blockDim.x;
    cplx<float> x[12], r[12], a[12], b[12];
    if (j < nsites) {
        f_load(xs, stride, j, x);
        f_load(rs, stride, j, r);
        sv_load<PR>(p, stride, j, a);
        sv_load<PR>(Ap, stride, j, b);
    }
    const double alpha = s->rr_s / m_fold_block_n(pl);
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        s->alpha = alpha;
        s->rr_prev = s->rr_s;  // the p update of this iteration divides by it
    }
    const float al = (float)alpha;
    double v[1] = {0};
    while (j < nsites) {
#pragma unroll
        for (int k = 0; k < 12; ++k) {
            x[k] = x[k] + al * a[k];
            r[k] = r[k] - al * b[k];
            v[0] += (double)r[k].re * r[k].re + (double)r[k].im * r[k].im;
        }
        f_store(xs, stride, j, x);
        f_store(rs, stride, j, r);
        j += jstep;
        if (j < nsites) {
            f_load(xs, stride, j, x);
            f_load(rs, stride, j, r);
            sv_load<PR>(p, stride, j, a);
            sv_load<PR>(Ap, stride, j, b);
        }
    }
    m_block_store<1>(v, partials);
}

// reliable update, part 1: x += x_s, x_s = 0
__global__ void k_rel_accum(const MCGState* s, zd* x, cplx<float>* xs, long nsites, int stride) {
    if (s->skip_rel || s->done) return;
    SITE_LOOP(j, nsites) {
        zd a[12];
        cplx<float> f[12];
        d_load(x, stride, j, a);
        f_load(xs, stride, j, f);
#pragma unroll
        for (int k = 0; k < 12; ++k) {
            a[k] = a[k] + mk((double)f[k].re, (double)f[k].im);
            f[k] = mk(0.f, 0.f);
        }
        d_store(x, stride, j, a);
        f_store(xs, stride, j, f);
    }
}

// x += x_s unconditionally (end of a solve that stopped between reliable updates)
__global__ void k_accum_xs(zd* x, const cplx<float>* xs, long nsites, int stride) {
    SITE_LOOP(j, nsites) {
        zd a[12];
        cplx<float> f[12];
        d_load(x, stride, j, a);
        f_load(xs, stride, j, f);
#pragma unroll
        for (int k = 0; k < 12; ++k) a[k] = a[k] + mk((double)f[k].re, (double)f[k].im);
        d_store(x, stride, j, a);
    }
}

// reliable update, part 2: r = b - Ax (FP64), r_s = r, partial |r|^2
__global__ __launch_bounds__(MR_THREADS) void k_rel_resid(const MCGState* s, zd* r, cplx<float>* rs, const zd* b,
                                                          const zd* Ax, long nsites, int stride, double* partials) {
    if (s->skip_rel || s->done) return;
    double v[1] = {0};
    SITE_LOOP(j, nsites) {
        zd bb[12], ax[12];
        cplx<float> f[12];
        d_load(b, stride, j, bb);
        d_load(Ax, stride, j, ax);
#pragma unroll
        for (int k = 0; k < 12; ++k) {
            bb[k] = bb[k] - ax[k];
            v[0] += bb[k].re * bb[k].re + bb[k].im * bb[k].im;
            f[k] = mk((float)bb[k].re, (float)bb[k].im);
        }
        d_store(r, stride, j, bb);
        f_store(rs, stride, j, f);
    }
    m_block_store<1>(v, partials);
}

__global__ __launch_bounds__(RED_BLOCKS) void k_rel_final(MCGState* s, const double* partials) {
    if (s->skip_rel || s->done) return;
    double out[1];
    m_final_sum<1>(partials, out);
    if (threadIdx.x == 0) {
        const double rt = out[0];
        s->rr_true = rt;
        s->nrel += 1;
        s->rel = 0;
        s->skip_rel = 1;
        if (rt < s->eps) {
            s->done = 1;
            s->converged = 1;
            return;
        }
        // r_s now holds the true residual; p (already updated) is kept
        s->rr_s = rt;
        s->maxrr = rt;
    }
}

// beta = |r_s|^2 / rr_prev (every workgroup folds the |r_s|^2 partials of the
// update itself), p = r_s + beta p; workgroup 0 also advances the state: the
// step count, the largest residual since the last reliable update, and the
// reliable-update flag (the update itself runs at the next reliable-update
// slot, the iteration completes in sloppy precision).  Only workgroup 0 writes
// the state and no workgroup reads what it writes, except `done` at the start
// (read once per workgroup: a solve stopped here by max_steps may leave p
// partly updated, which nothing reads).
template <int PR>
__global__ __launch_bounds__(MR_THREADS) void k_mcg_p(MCGState* s, SVec<PR> p, const cplx<float>* rs, long nsites,
                                                      int stride, const double* rpart, int nrpart) {
    __shared__ int stop;
    if (threadIdx.x == 0) stop = s->done;
    __syncthreads();
    if (stop) return;
    // nrpart partials (the fused update's, one per stencil workgroup), else RED_BLOCKS
    constexpr int NP = NORM_PER_THREAD > PART_PER_THREAD ? NORM_PER_THREAD : PART_PER_THREAD;
    double pl[NP];
    if (nrpart > 0) {
        m_partials_load_n(rpart, nrpart, pl);
    } else {
        m_partials_load(rpart, pl);
#pragma unroll
        for (int i = PART_PER_THREAD; i < NP; ++i) pl[i] = 0.0;
    }
    // the first site's vectors load while the partials are folded
    long j = (long)blockIdx.x * blockDim.x + threadIdx.x;
    const long jstep = (long)gridDim.x * blockDim.x;
    cplx<float> a[12], r[12];
    if (j < nsites) {
        sv_load<PR>(p, stride, j, a);
        f_load(rs, stride, j, r);
    }
    const double rn = m_fold_block_n(pl);
    const double beta = rn / s->rr_prev;
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        if (rn > s->maxrr) s->maxrr = rn;
        if (rn < s->delta2 * s->maxrr || rn < s->eps) {
            s->rel = 1;
            s->skip_rel = 0;
        }
        s->beta = beta;
        s->rr_s = rn;
        s->step += 1;
        if (s->step >= s->max_steps) s->done = 1;
    }
    const float be = (float)beta;
    while (j < nsites) {
#pragma unroll
        for (int k = 0; k < 12; ++k) a[k] = r[k] + be * a[k];
        sv_store<PR>(p, stride, j, a);
        j += jstep;
        if (j < nsites) {
            sv_load<PR>(p, stride, j, a);
            f_load(rs, stride, j, r);
        }
    }
}

// ---------------------------------------------------------------- launchers
template <int PR>
static SVec<PR> sv(void* d, float* sc) {
    return SVec<PR>{reinterpret_cast<typename Prec<PR>::S*>(d), sc};
}

void launch_mcg_init(int prec, MCGState* s, const zd* r, cplx<float>* rs, cplx<float>* xs, void* p, float* p_sc,
                     long nsites, int stride, double* partials, double eps, double delta, int max_steps,
                     hipStream_t st) {
    if (prec == 1)
        hipLaunchKernelGGL(k_mcg_init<1>, dim3(RED_BLOCKS), dim3(MR_THREADS), 0, st, r, rs, xs, sv<1>(p, p_sc), nsites,
                           stride, partials);
    else
        hipLaunchKernelGGL(k_mcg_init<2>, dim3(RED_BLOCKS), dim3(MR_THREADS), 0, st, r, rs, xs, sv<2>(p, p_sc), nsites,
                           stride, partials);
    hipLaunchKernelGGL(k_mcg_init_final, dim3(1), dim3(RED_BLOCKS), 0, st, s, partials, eps, delta, max_steps);
}

void launch_mcg_alpha(int prec, MCGState* s, void* p, float* p_sc, void* Ap, float* Ap_sc, long nsites, int stride,
                      double* partials, hipStream_t st) {
    if (prec == 1)
        hipLaunchKernelGGL(k_mcg_dot<1>, dim3(RED_BLOCKS), dim3(MR_THREADS), 0, st, s, sv<1>(p, p_sc), sv<1>(Ap, Ap_sc),
                           nsites, stride, partials);
    else
        hipLaunchKernelGGL(k_mcg_dot<2>, dim3(RED_BLOCKS), dim3(MR_THREADS), 0, st, s, sv<2>(p, p_sc), sv<2>(Ap, Ap_sc),
                           nsites, stride, partials);
}

void launch_mcg_update(int prec, MCGState* s, cplx<float>* xs, cplx<float>* rs, void* p, float* p_sc, void* Ap,
                       float* Ap_sc, long nsites, int stride, double* partials, const double* dot_partials, int ndot,
                       hipStream_t st) {
    // |r_s|^2 partials into the second RED_BLOCKS slots
    double* rpart = partials + RED_BLOCKS;
    if (prec == 1)
        hipLaunchKernelGGL(k_mcg_update<1>, dim3(RED_BLOCKS), dim3(MR_THREADS), 0, st, s, xs, rs, sv<1>(p, p_sc),
                           sv<1>(Ap, Ap_sc), nsites, stride, dot_partials, ndot, rpart);
    else
        hipLaunchKernelGGL(k_mcg_update<2>, dim3(RED_BLOCKS), dim3(MR_THREADS), 0, st, s, xs, rs, sv<2>(p, p_sc),
                           sv<2>(Ap, Ap_sc), nsites, stride, dot_partials, ndot, rpart);
}

void launch_rel_accum(MCGState* s, zd* x, cplx<float>* xs, long nsites, int stride, hipStream_t st) {
    hipLaunchKernelGGL(k_rel_accum, dim3(site_blocks(nsites)), dim3(256), 0, st, s, x, xs, nsites, stride);
}

void launch_accum_xs(zd* x, const cplx<float>* xs, long nsites, int stride, hipStream_t st) {
    hipLaunchKernelGGL(k_accum_xs, dim3(site_blocks(nsites)), dim3(256), 0, st, x, xs, nsites, stride);
}

void launch_rel_resid(MCGState* s, zd* r, cplx<float>* rs, const zd* b, const zd* Ax, long nsites, int stride,
                      double* partials, hipStream_t st) {
    hipLaunchKernelGGL(k_rel_resid, dim3(RED_BLOCKS), dim3(MR_THREADS), 0, st, s, r, rs, b, Ax, nsites, stride, partials);
    hipLaunchKernelGGL(k_rel_final, dim3(1), dim3(RED_BLOCKS), 0, st, s, partials);
}

void launch_mcg_p(int prec, MCGState* s, void* p, float* p_sc, const cplx<float>* rs, long nsites, int stride,
                  const double* partials, hipStream_t st, const double* upd_part, int nupd) {
    // |r_s|^2 partials: launch_mcg_update's, or nupd of the fused update
    const double* rpart = nupd > 0 ? upd_part : partials + RED_BLOCKS;
    if (prec == 1)
        hipLaunchKernelGGL(k_mcg_p<1>, dim3(RED_BLOCKS), dim3(MR_THREADS), 0, st, s, sv<1>(p, p_sc), rs, nsites, stride,
                           rpart, nupd);
    else
        hipLaunchKernelGGL(k_mcg_p<2>, dim3(RED_BLOCKS), dim3(MR_THREADS), 0, st, s, sv<2>(p, p_sc), rs, nsites, stride,
                           rpart, nupd);
}

}  // namespace lym
