// blas.hip -- spinor BLAS-1 and the device-resident CG / BiCGStab state machines.
//
// Reductions (AlgebraUtils::squaredNorm, dot, gamma5dot, differenceNorm,
// source/algebra_utils/AlgebraUtils.h:17-27,182-254) run on a fixed grid of
// RED_BLOCKS workgroups: each wavefront folds its lanes with xor-shuffles,
// the workgroup folds its waves through LDS and writes one partial per
// workgroup; a single-workgroup second pass folds the partials in a fixed
// tree.  The order of every addition is therefore fixed, so a result does not
// depend on what else runs on the device (the reference sums in long double;
// here every sum is FP64).
//
// The solver kernels keep all scalars (alpha, beta, rho, omega, |r|^2) in a
// device state block, so an iteration needs no host round trip; every kernel
// returns at once when the state says `done`, which makes the iteration
// count and stopping point those of the reference loops
// (source/inverters/ConjugateGradient.cpp:33-67, BiConjugateGradient.cpp:364-501)
// however many iterations the host has queued.
#include "lym_kernels.h"

namespace lym {

constexpr int RB_THREADS = 256;

template <int NV>
__device__ __forceinline__ void wave_sum(double v[NV]) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1)
#pragma unroll
        for (int k = 0; k < NV; ++k) v[k] += __shfl_xor(v[k], off, 64);
}

// block of RB_THREADS threads -> partials[k * RED_BLOCKS + blockIdx.x]
template <int NV>
__device__ __forceinline__ void block_store(double v[NV], double* partials) {
    __shared__ double sh[NV][RB_THREADS / 64];
    wave_sum<NV>(v);
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    if (lane == 0)
#pragma unroll
        for (int k = 0; k < NV; ++k) sh[k][w] = v[k];
    __syncthreads();
    if (threadIdx.x == 0) {
#pragma unroll
        for (int k = 0; k < NV; ++k) {
            double s = 0;
#pragma unroll
            for (int j = 0; j < RB_THREADS / 64; ++j) s += sh[k][j];
            partials[k * RED_BLOCKS + blockIdx.x] = s;
        }
    }
}

// one block of RED_BLOCKS threads folds the partials; thread 0 returns totals
template <int NV>
__device__ __forceinline__ void final_sum(const double* partials, double out[NV]) {
    __shared__ double sh[NV][RED_BLOCKS / 64];
    double v[NV];
#pragma unroll
    for (int k = 0; k < NV; ++k) v[k] = partials[k * RED_BLOCKS + threadIdx.x];
    wave_sum<NV>(v);
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

// The total of final_sum, folded by every workgroup of RB_THREADS threads that
// needs it (the same fixed order in each, so all get the same value): the
// partials are loaded first (fold_load), the caller may then issue its own
// loads, and fold_block folds across the workgroup (result in every thread).
constexpr int FOLD_PER_THREAD = RED_BLOCKS / RB_THREADS;
static_assert(RED_BLOCKS % RB_THREADS == 0, "partials per thread");
template <int NV>
__device__ __forceinline__ void fold_load(const double* partials, double pl[NV][FOLD_PER_THREAD]) {
#pragma unroll
    for (int k = 0; k < NV; ++k)
#pragma unroll
        for (int i = 0; i < FOLD_PER_THREAD; ++i) pl[k][i] = partials[k * RED_BLOCKS + threadIdx.x + i * RB_THREADS];
}
template <int NV>
__device__ __forceinline__ void fold_block(const double pl[NV][FOLD_PER_THREAD], double out[NV]) {
    __shared__ double sh[NV][RB_THREADS / 64];
    double v[NV];
#pragma unroll
    for (int k = 0; k < NV; ++k) {
        v[k] = 0;
#pragma unroll
        for (int i = 0; i < FOLD_PER_THREAD; ++i) v[k] += pl[k][i];
    }
    wave_sum<NV>(v);
    if ((threadIdx.x & 63) == 0)
#pragma unroll
        for (int k = 0; k < NV; ++k) sh[k][threadIdx.x >> 6] = v[k];
    __syncthreads();
#pragma unroll
    for (int k = 0; k < NV; ++k) {
        double t = 0;
#pragma unroll
        for (int j = 0; j < RB_THREADS / 64; ++j) t += sh[k][j];
        out[k] = t;
    }
    __syncthreads();
}

#define GRID_LOOP(i, n) for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < (n); i += (long)gridDim.x * blockDim.x)

// ---------------------------------------------------------------- plain reductions
__global__ __launch_bounds__(RB_THREADS) void k_norm2(const zd* __restrict__ a, long n, double* partials) {
    double v[1] = {0};
    GRID_LOOP(i, n) {
        zd x = a[i];
        v[0] += x.re * x.re + x.im * x.im;
    }
    block_store<1>(v, partials);
}

__global__ __launch_bounds__(RB_THREADS) void k_dot(const zd* __restrict__ a, const zd* __restrict__ b, long n,
                                                    double* partials) {
    double v[2] = {0, 0};
    GRID_LOOP(i, n) {
        zd x = a[i], y = b[i];
        v[0] += x.re * y.re + x.im * y.im;
        v[1] += x.re * y.im - x.im * y.re;
    }
    block_store<2>(v, partials);
}

__global__ __launch_bounds__(RB_THREADS) void k_gamma5dot(const zd* __restrict__ a, const zd* __restrict__ b, long n,
                                                          int stride, double* partials) {
    double v[2] = {0, 0};
    GRID_LOOP(i, n) {
        zd x = a[i], y = b[i];
        double s = ((i / stride) % 12) < 6 ? 1.0 : -1.0;
        v[0] += s * (x.re * y.re + x.im * y.im);
        v[1] += s * (x.re * y.im - x.im * y.re);
    }
    block_store<2>(v, partials);
}

__global__ __launch_bounds__(RB_THREADS) void k_diffnorm(const zd* __restrict__ a, const zd* __restrict__ b, long n,
                                                         double* partials) {
    double v[1] = {0};
    GRID_LOOP(i, n) {
        zd d = a[i] - b[i];
        v[0] += d.re * d.re + d.im * d.im;
    }
    block_store<1>(v, partials);
}

template <int NV>
__global__ __launch_bounds__(RED_BLOCKS) void k_final(const double* partials, double* result) {
    double out[NV];
    final_sum<NV>(partials, out);
    if (threadIdx.x == 0)
        for (int k = 0; k < NV; ++k) result[k] = out[k];
}

void launch_norm2(const zd* a, long n, double* partials, double* result, hipStream_t st) {
    hipLaunchKernelGGL(k_norm2, dim3(RED_BLOCKS), dim3(RB_THREADS), 0, st, a, n, partials);
    hipLaunchKernelGGL(k_final<1>, dim3(1), dim3(RED_BLOCKS), 0, st, partials, result);
}
void launch_dot(const zd* a, const zd* b, long n, double* partials, double* result, hipStream_t st) {
    hipLaunchKernelGGL(k_dot, dim3(RED_BLOCKS), dim3(RB_THREADS), 0, st, a, b, n, partials);
    hipLaunchKernelGGL(k_final<2>, dim3(1), dim3(RED_BLOCKS), 0, st, partials, result);
}
void launch_gamma5dot(const zd* a, const zd* b, long n, int nparity, int stride, double* partials, double* result,
                      hipStream_t st) {
    (void)nparity;
    hipLaunchKernelGGL(k_gamma5dot, dim3(RED_BLOCKS), dim3(RB_THREADS), 0, st, a, b, n, stride, partials);
    hipLaunchKernelGGL(k_final<2>, dim3(1), dim3(RED_BLOCKS), 0, st, partials, result);
}
void launch_diffnorm(const zd* a, const zd* b, long n, double* partials, double* result, hipStream_t st) {
    hipLaunchKernelGGL(k_diffnorm, dim3(RED_BLOCKS), dim3(RB_THREADS), 0, st, a, b, n, partials);
    hipLaunchKernelGGL(k_final<1>, dim3(1), dim3(RED_BLOCKS), 0, st, partials, result);
}

// ---------------------------------------------------------------- elementwise
// one element per thread (grid-stride only beyond 2^28 elements): on the
// MI355X a copy with one element per thread streams at 6.2 TB/s, the same
// copy as a grid-stride loop over 8192 workgroups at 5.1 TB/s
static inline int ew_blocks(long n) {
    long b = (n + 255) / 256;
    return (int)(b < (1L << 20) ? b : (1L << 20));
}

__global__ void k_axpy(zd* __restrict__ y, zd a, const zd* __restrict__ x, long n) {
    GRID_LOOP(i, n) { y[i] = y[i] + a * x[i]; }
}
__global__ void k_xpay(zd* __restrict__ y, const zd* __restrict__ x, zd a, long n) {
    GRID_LOOP(i, n) { y[i] = x[i] + a * y[i]; }
}
__global__ void k_g5(zd* __restrict__ y, long n, int stride) {
    GRID_LOOP(i, n) {
        if (((i / stride) % 12) >= 6) y[i] = -y[i];
    }
}
__global__ void k_copy(zd* __restrict__ y, const zd* __restrict__ x, long n) {
    GRID_LOOP(i, n) { y[i] = x[i]; }
}
__global__ void k_sub(zd* __restrict__ o, const zd* __restrict__ a, const zd* __restrict__ b, long n) {
    GRID_LOOP(i, n) { o[i] = a[i] - b[i]; }
}

void launch_axpy(zd* y, zd a, const zd* x, long n, hipStream_t st) {
    hipLaunchKernelGGL(k_axpy, dim3(ew_blocks(n)), dim3(256), 0, st, y, a, x, n);
}
void launch_xpay(zd* y, const zd* x, zd a, long n, hipStream_t st) {
    hipLaunchKernelGGL(k_xpay, dim3(ew_blocks(n)), dim3(256), 0, st, y, x, a, n);
}
void launch_scale_g5(zd* y, int nparity, int stride, hipStream_t st) {
    long n = (long)nparity * 12 * stride;
    hipLaunchKernelGGL(k_g5, dim3(ew_blocks(n)), dim3(256), 0, st, y, n, stride);
}
__global__ void k_set_double(double* p, double v) { *p = v; }
void launch_set_double(double* p, double v, hipStream_t st) {
    hipLaunchKernelGGL(k_set_double, dim3(1), dim3(1), 0, st, p, v);
}

void launch_copy(zd* y, const zd* x, long n, hipStream_t st) {
    hipLaunchKernelGGL(k_copy, dim3(ew_blocks(n)), dim3(256), 0, st, y, x, n);
}
void launch_sub(zd* o, const zd* a, const zd* b, long n, hipStream_t st) {
    hipLaunchKernelGGL(k_sub, dim3(ew_blocks(n)), dim3(256), 0, st, o, a, b, n);
}

// ---------------------------------------------------------------- CG
__global__ __launch_bounds__(RED_BLOCKS) void k_cg_init_final(CGState* s, const double* partials, double eps,
                                                              int max_steps) {
    double out[1];
    final_sum<1>(partials, out);
    if (threadIdx.x == 0) {
        s->rr = out[0];
        s->rr_new = out[0];
        s->eps = eps;
        s->max_steps = max_steps;
        s->step = 0;
        s->converged = 0;
        s->done = max_steps <= 0 ? 1 : 0;
        s->alpha_re = s->alpha_im = s->beta = 0;
    }
}

void launch_cg_init(CGState* s, const zd* r, long n, double* partials, double eps, int max_steps, hipStream_t st) {
    hipLaunchKernelGGL(k_norm2, dim3(RED_BLOCKS), dim3(RB_THREADS), 0, st, r, n, partials);
    hipLaunchKernelGGL(k_cg_init_final, dim3(1), dim3(RED_BLOCKS), 0, st, s, partials, eps, max_steps);
}

__global__ __launch_bounds__(RB_THREADS) void k_cg_dot(const CGState* s, const zd* __restrict__ p,
                                                       const zd* __restrict__ Ap, long n, double* partials) {
    if (s->done) return;
    double v[2] = {0, 0};
    GRID_LOOP(i, n) {
        zd x = p[i], y = Ap[i];
        v[0] += x.re * y.re + x.im * y.im;
        v[1] += x.re * y.im - x.im * y.re;
    }
    block_store<2>(v, partials);
}

void launch_cg_alpha(CGState* s, const zd* p, const zd* Ap, long n, double* partials, hipStream_t st) {
    hipLaunchKernelGGL(k_cg_dot, dim3(RED_BLOCKS), dim3(RB_THREADS), 0, st, s, p, Ap, n, partials);
}

// alpha = |r|^2 / <p, A p> (complex, ConjugateGradient.cpp:36), folded by
// every workgroup from the <p, Ap> partials; then x += alpha p, r -= alpha Ap,
// partial |r|^2.  Workgroup 0 records alpha and the |r|^2 the p update
// divides by.
__global__ __launch_bounds__(RB_THREADS) void k_cg_update(CGState* s, zd* __restrict__ x, zd* __restrict__ r,
                                                          const zd* __restrict__ p, const zd* __restrict__ Ap, long n,
                                                          const double* dot_partials, double* partials) {
    if (s->done) return;
    double pl[2][FOLD_PER_THREAD];
    fold_load<2>(dot_partials, pl);
    long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    const long istep = (long)gridDim.x * blockDim.x;
    zd xv, rv, pv, av;
    if (i < n) {
        xv = x[i];
        rv = r[i];
        pv = p[i];
        av = Ap[i];
    }
    double d[2];
    fold_block<2>(pl, d);
    const double den = d[0] * d[0] + d[1] * d[1];
    const zd al = mk(s->rr * d[0] / den, -s->rr * d[1] / den);
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        s->alpha_re = al.re;
        s->alpha_im = al.im;
        s->rr_prev = s->rr;
    }
    double v[1] = {0};
    while (i < n) {
        x[i] = xv + al * pv;
        const zd rn = rv - al * av;
        r[i] = rn;
        v[0] += rn.re * rn.re + rn.im * rn.im;
        i += istep;
        if (i < n) {
            xv = x[i];
            rv = r[i];
            pv = p[i];
            av = Ap[i];
        }
    }
    block_store<1>(v, partials);
}

void launch_cg_update(CGState* s, zd* x, zd* r, const zd* p, const zd* Ap, long n, double* partials,
                      hipStream_t st) {
    // <p, Ap> partials in slots 0 and 1 (launch_cg_alpha), |r|^2 in slot 2
    hipLaunchKernelGGL(k_cg_update, dim3(RED_BLOCKS), dim3(RB_THREADS), 0, st, s, x, r, p, Ap, n, partials,
                       partials + 2 * RED_BLOCKS);
}

// |r|^2 folded by every workgroup from the update's partials; workgroup 0
// advances the state (ConjugateGradient.cpp:49-50: stop at |r|^2 < eps, else
// beta = |r|^2 / |r_prev|^2, step + 1); then p = r + beta p.  Only workgroup 0
// writes the state and no workgroup reads what it writes except `done` at the
// start (read once per workgroup: after a stop decided here p may be left
// partly updated, which nothing reads).
__global__ __launch_bounds__(RB_THREADS) void k_cg_p(CGState* s, zd* __restrict__ p, const zd* __restrict__ r, long n,
                                                     const double* rpart) {
    __shared__ int stop;
    if (threadIdx.x == 0) stop = s->done;
    __syncthreads();
    if (stop) return;
    double pl[1][FOLD_PER_THREAD];
    fold_load<1>(rpart, pl);
    long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    const long istep = (long)gridDim.x * blockDim.x;
    zd pv, rv;
    if (i < n) {
        pv = p[i];
        rv = r[i];
    }
    double rn[1];
    fold_block<1>(pl, rn);
    const double b = rn[0] / s->rr_prev;
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        s->rr_new = rn[0];
        if (rn[0] < s->eps) {
            s->done = 1;
            s->converged = 1;
        } else {
            s->beta = b;
            s->rr = rn[0];
            s->step += 1;
            if (s->step >= s->max_steps) s->done = 1;
        }
    }
    while (i < n) {
        p[i] = rv + b * pv;
        i += istep;
        if (i < n) {
            pv = p[i];
            rv = r[i];
        }
    }
}

void launch_cg_p(CGState* s, zd* p, const zd* r, long n, double* partials, hipStream_t st) {
    hipLaunchKernelGGL(k_cg_p, dim3(RED_BLOCKS), dim3(RB_THREADS), 0, st, s, p, r, n, partials + 2 * RED_BLOCKS);
}

// ---------------------------------------------------------------- BiCGStab
__device__ __forceinline__ zd zdiv(zd a, zd b) {
    double den = b.re * b.re + b.im * b.im;
    return mk((a.re * b.re + a.im * b.im) / den, (a.im * b.re - a.re * b.im) / den);
}

__global__ void k_bicg_init(BiCGState* s, double precision, int max_steps) {
    s->rho_re = 1; s->rho_im = 0;
    s->alpha_re = 1; s->alpha_im = 0;
    s->omega_re = 1; s->omega_im = 0;
    s->rhon_re = s->rhon_im = s->beta_re = s->beta_im = 0;
    s->norm = 0;
    s->precision = precision;
    s->max_steps = max_steps;
    s->step = 0;
    s->converged = 0;
    s->breakdown = 0;
    s->copy_b = 0;
    s->done = max_steps <= 0 ? 1 : 0;
}

void launch_bicg_init(BiCGState* s, double precision, int max_steps, hipStream_t st) {
    hipLaunchKernelGGL(k_bicg_init, dim3(1), dim3(1), 0, st, s, precision, max_steps);
}

__global__ void k_bicg_setup(zd* __restrict__ r, zd* __restrict__ rh, zd* __restrict__ p, zd* __restrict__ nu,
                             const zd* __restrict__ b, const zd* __restrict__ Ax, long n) {
    GRID_LOOP(i, n) {
        zd bb = b[i], ax = Ax[i];
        r[i] = bb - ax;
        rh[i] = bb + ax;
        p[i] = mk(0.0, 0.0);
        nu[i] = mk(0.0, 0.0);
    }
}

void launch_bicg_setup(zd* r, zd* rh, zd* p, zd* nu, const zd* b, const zd* Ax, long n, hipStream_t st) {
    hipLaunchKernelGGL(k_bicg_setup, dim3(ew_blocks(n)), dim3(256), 0, st, r, rh, p, nu, b, Ax, n);
}

template <class S>
__global__ __launch_bounds__(RB_THREADS) void k_state_dot(const S* s, const zd* __restrict__ a,
                                                          const zd* __restrict__ b, long n, double* partials) {
    if (s->done) return;
    double v[2] = {0, 0};
    GRID_LOOP(i, n) {
        zd x = a[i], y = b[i];
        v[0] += x.re * y.re + x.im * y.im;
        v[1] += x.re * y.im - x.im * y.re;
    }
    block_store<2>(v, partials);
}

void launch_bicg_rho(BiCGState* s, const zd* rh, const zd* r, long n, double* partials, hipStream_t st) {
    hipLaunchKernelGGL(k_state_dot<BiCGState>, dim3(RED_BLOCKS), dim3(RB_THREADS), 0, st, s, rh, r, n, partials);
}

// Each of the p, s and x/r updates below folds the partial sums of the dot
// product before it in every workgroup (a fixed order, so all get the same
// value); workgroup 0 writes the state, and no workgroup reads what it
// writes except `done` at the start (read once per workgroup: after a stop
// decided there a vector may be left partly updated, which nothing reads).
#define BICG_STOP_CHECK()                        \
    __shared__ int stop_;                        \
    if (threadIdx.x == 0) stop_ = s->done;      \
    __syncthreads();                             \
    if (stop_) return

// rho_next = <rh, r>; breakdown if zero; beta = (rho_next / rho)(alpha / omega);
// p = r + beta (p - omega nu)
__global__ __launch_bounds__(RB_THREADS) void k_bicg_p(BiCGState* s, zd* __restrict__ p, const zd* __restrict__ r,
                                                       const zd* __restrict__ nu, long n, const double* partials) {
    BICG_STOP_CHECK();
    double pl[2][FOLD_PER_THREAD];
    fold_load<2>(partials, pl);
    double out[2];
    fold_block<2>(pl, out);
    const zd rn = mk(out[0], out[1]);
    const bool brk = rn.re * rn.re + rn.im * rn.im == 0.0;
    const zd beta = zdiv(rn, mk(s->rho_re, s->rho_im)) * zdiv(mk(s->alpha_re, s->alpha_im), mk(s->omega_re, s->omega_im));
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        s->rhon_re = rn.re;
        s->rhon_im = rn.im;
        if (brk) {
            s->done = 1;
            s->breakdown = 1;
            s->converged = 0;
        } else {
            s->beta_re = beta.re;
            s->beta_im = beta.im;
        }
    }
    if (brk) return;
    const zd om = mk(s->omega_re, s->omega_im);
    GRID_LOOP(i, n) { p[i] = r[i] + beta * (p[i] - om * nu[i]); }
}

void launch_bicg_p(BiCGState* s, zd* p, const zd* r, const zd* nu, long n, const double* partials, hipStream_t st) {
    hipLaunchKernelGGL(k_bicg_p, dim3(RED_BLOCKS), dim3(RB_THREADS), 0, st, s, p, r, nu, n, partials);
}

void launch_bicg_alpha(BiCGState* s, const zd* rh, const zd* nu, long n, double* partials, hipStream_t st) {
    hipLaunchKernelGGL(k_state_dot<BiCGState>, dim3(RED_BLOCKS), dim3(RB_THREADS), 0, st, s, rh, nu, n, partials);
}

// alpha = rho_next / <rh, nu>; s = r - alpha nu
__global__ __launch_bounds__(RB_THREADS) void k_bicg_s(BiCGState* s, zd* __restrict__ sv, const zd* __restrict__ r,
                                                       const zd* __restrict__ nu, long n, const double* partials) {
    BICG_STOP_CHECK();
    double pl[2][FOLD_PER_THREAD];
    fold_load<2>(partials, pl);
    double out[2];
    fold_block<2>(pl, out);
    const zd al = zdiv(mk(s->rhon_re, s->rhon_im), mk(out[0], out[1]));
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        s->alpha_re = al.re;
        s->alpha_im = al.im;
    }
    GRID_LOOP(i, n) { sv[i] = r[i] - al * nu[i]; }
}

void launch_bicg_s(BiCGState* s, zd* sv, const zd* r, const zd* nu, long n, const double* partials, hipStream_t st) {
    hipLaunchKernelGGL(k_bicg_s, dim3(RED_BLOCKS), dim3(RB_THREADS), 0, st, s, sv, r, nu, n, partials);
}

__global__ __launch_bounds__(RB_THREADS) void k_bicg_ts(const BiCGState* s, const zd* __restrict__ t,
                                                        const zd* __restrict__ sv, long n, double* partials) {
    if (s->done) return;
    double v[4] = {0, 0, 0, 0};
    GRID_LOOP(i, n) {
        zd x = t[i], y = sv[i];
        v[0] += x.re * y.re + x.im * y.im;
        v[1] += x.re * y.im - x.im * y.re;
        v[2] += x.re * x.re + x.im * x.im;
    }
    block_store<4>(v, partials);
}

void launch_bicg_omega(BiCGState* s, const zd* t, const zd* sv, long n, double* partials, hipStream_t st) {
    hipLaunchKernelGGL(k_bicg_ts, dim3(RED_BLOCKS), dim3(RB_THREADS), 0, st, s, t, sv, n, partials);
}

// omega = <t,s>/<t,t>; <t,t> == 0 -> solution = source (BiConjugateGradient.cpp:449-458);
// x += alpha p + omega s, r = s - omega t, partial |r|^2 (into rpart)
__global__ __launch_bounds__(RB_THREADS) void k_bicg_update(BiCGState* s, zd* __restrict__ x, zd* __restrict__ r,
                                                            const zd* __restrict__ p, const zd* __restrict__ sv,
                                                            const zd* __restrict__ t, const zd* __restrict__ b,
                                                            long n, const double* partials, double* rpart) {
    BICG_STOP_CHECK();
    double pl[4][FOLD_PER_THREAD];
    fold_load<4>(partials, pl);
    double out[4];
    fold_block<4>(pl, out);
    if (out[2] == 0.0) {
        if (blockIdx.x == 0 && threadIdx.x == 0) {
            s->copy_b = 1;
            s->done = 1;
            s->converged = 1;
        }
        GRID_LOOP(i, n) { x[i] = b[i]; }
        return;
    }
    const zd om = mk(out[0] / out[2], out[1] / out[2]);
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        s->omega_re = om.re;
        s->omega_im = om.im;
    }
    const zd al = mk(s->alpha_re, s->alpha_im);
    double v[1] = {0};
    GRID_LOOP(i, n) {
        zd sv_i = sv[i];
        x[i] = x[i] + al * p[i] + om * sv_i;
        zd rn = sv_i - om * t[i];
        r[i] = rn;
        v[0] += rn.re * rn.re + rn.im * rn.im;
    }
    block_store<1>(v, rpart);
}

__global__ __launch_bounds__(RED_BLOCKS) void k_bicg_update_final(BiCGState* s, const double* partials) {
    if (s->done) return;
    double out[1];
    final_sum<1>(partials, out);
    if (threadIdx.x == 0) {
        s->norm = out[0];
        if (out[0] < s->precision) {
            s->done = 1;
            s->converged = 1;
            return;
        }
        s->rho_re = s->rhon_re;
        s->rho_im = s->rhon_im;
        s->step += 1;
        if (s->step >= s->max_steps) s->done = 1;
    }
}

void launch_bicg_update(BiCGState* s, zd* x, zd* r, const zd* p, const zd* sv, const zd* t, const zd* b, long n,
                        double* partials, hipStream_t st) {
    // <t,s>, <t,t> partials (launch_bicg_omega) in slots 0..3, |r|^2 in slot 4
    double* rpart = partials + 4 * RED_BLOCKS;
    hipLaunchKernelGGL(k_bicg_update, dim3(RED_BLOCKS), dim3(RB_THREADS), 0, st, s, x, r, p, sv, t, b, n, partials,
                       rpart);
    hipLaunchKernelGGL(k_bicg_update_final, dim3(1), dim3(RED_BLOCKS), 0, st, s, rpart);
}

}  // namespace lym

namespace lym {

__device__ __forceinline__ uint64_t mix64(uint64_t x) {
    x += 0x9E3779B97F4A7C15ull;
    x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ull;
    x = (x ^ (x >> 27)) * 0x94D049BB133111EBull;
    return x ^ (x >> 31);
}

__global__ void k_fill_uniform(zd* x, long n, int stride, int Vh, uint64_t seed) {
    GRID_LOOP(i, n) {
        if ((int)(i % stride) >= Vh) {
            x[i] = mk(0.0, 0.0);
        } else {
            uint64_t a = mix64(seed ^ mix64(2 * (uint64_t)i)), b = mix64(seed ^ mix64(2 * (uint64_t)i + 1));
            x[i] = mk((a >> 11) * 0x1.0p-53, (b >> 11) * 0x1.0p-53);
        }
    }
}

void launch_fill_uniform(zd* x, int nparity, int stride, int Vh, uint64_t seed, hipStream_t st) {
    long n = (long)nparity * 12 * stride;
    hipLaunchKernelGGL(k_fill_uniform, dim3(ew_blocks(n)), dim3(256), 0, st, x, n, stride, Vh, seed);
}

}  // namespace lym
