// stout.hip -- stout smearing of the fundamental gauge links on the device:
// StoutSmearing::smearing (source/utils/StoutSmearing.cpp:12-20, smearLink
// :55-62), the step Environment::synchronize applies stout_smearing_levels
// times before it builds the fermion links (source/Environment.cpp:22-47).
//
// Per link: the staple sum of WilsonGaugeAction::staple
// (source/actions/WilsonGaugeAction.cpp:9-20)
//   S = sum_{nu != mu} U_nu(x+mu) U_mu(x+nu)^+ U_nu(x)^+ + U_nu(x+mu-nu)^+ U_mu(x-nu)^+ U_nu(x-nu),
// Omega = rho S^+ U_mu(x)^+, X = (Omega - Omega^+)/2 minus its trace / N, and
// U' = exp(X) U_mu(x), with exp in closed form: the Cayley-Hamilton form of
// ExponentialMap::exp for SU(3) (source/utils/ExpMap.h:100-140; Morningstar and
// Peardon, Phys. Rev. D 69 (2004) 054501) and the rotation form for SU(2)
// (:142-155).  Deliberate differences: an exponent that is exactly zero (cold
// links, rho = 0) gives the identity where the reference divides 0 by 0, and
// the ratio under acos is clamped to 1 against rounding.
//
// Device layout: complex double [4 mu][N*N entries, column-major][V], sites
// in the t-fastest lexicographic order; one thread per link (site, mu).
#include <cstring>

#include "../../include/lym.h"
#include "lym_kernels.h"

namespace lym_internal {
int fail(int code, const char* fmt, ...);
}
using lym_internal::fail;

namespace lym {
namespace {

template <int N>
struct Mat {
    zd a[N * N];  // (i, j) at i + N j
    __device__ zd& operator()(int i, int j) { return a[i + N * j]; }
    __device__ const zd& operator()(int i, int j) const { return a[i + N * j]; }
};

struct Lex {
    int L[4];
    long V;
};

__device__ __forceinline__ long lex_of(const Lex& g, const int c[4]) {
    return (((long)c[0] * g.L[1] + c[1]) * g.L[2] + c[2]) * g.L[3] + c[3];
}

template <int N>
__device__ __forceinline__ Mat<N> ld(const zd* U, long V, long s, int mu) {
    Mat<N> m;
#pragma unroll
    for (int k = 0; k < N * N; ++k) m.a[k] = U[((long)mu * N * N + k) * V + s];
    return m;
}

// A B, A^+ B, A B^+
template <int N>
__device__ __forceinline__ Mat<N> mul(const Mat<N>& A, const Mat<N>& B) {
    Mat<N> C;
#pragma unroll
    for (int i = 0; i < N; ++i)
#pragma unroll
        for (int j = 0; j < N; ++j) {
            zd s = A(i, 0) * B(0, j);
#pragma unroll
            for (int k = 1; k < N; ++k) cfma(s, A(i, k), B(k, j));
            C(i, j) = s;
        }
    return C;
}
template <int N>
__device__ __forceinline__ Mat<N> mul_dn(const Mat<N>& A, const Mat<N>& B) {  // A^+ B
    Mat<N> C;
#pragma unroll
    for (int i = 0; i < N; ++i)
#pragma unroll
        for (int j = 0; j < N; ++j) {
            zd s = cmulc(A(0, i), B(0, j));
#pragma unroll
            for (int k = 1; k < N; ++k) cfmac(s, A(k, i), B(k, j));
            C(i, j) = s;
        }
    return C;
}
template <int N>
__device__ __forceinline__ Mat<N> mul_nd(const Mat<N>& A, const Mat<N>& B) {  // A B^+
    Mat<N> C;
#pragma unroll
    for (int i = 0; i < N; ++i)
#pragma unroll
        for (int j = 0; j < N; ++j) {
            zd s = A(i, 0) * conj(B(j, 0));
#pragma unroll
            for (int k = 1; k < N; ++k) cfma(s, A(i, k), conj(B(j, k)));
            C(i, j) = s;
        }
    return C;
}

// exp of an anti-hermitian traceless X
__device__ Mat<2> expmap(const Mat<2>& X) {
    Mat<2> R;
    const double a = X(0, 0).im, b = X(0, 1).re, c = X(0, 1).im;
    const double r = sqrt(a * a + b * b + c * c);
    if (r == 0.0) {
        R(0, 0) = R(1, 1) = mk(1.0, 0.0);
        R(0, 1) = R(1, 0) = mk(0.0, 0.0);
        return R;
    }
    // exp(i r n.sigma) = cos r + i sin r (n.sigma) with X = i (c sx + b sy + a sz)
    double sn, cs;
    sincos(r, &sn, &cs);
    const double s = sn / r;
    R(0, 0) = mk(cs, a * s);
    R(0, 1) = mk(b * s, c * s);
    R(1, 0) = mk(-b * s, c * s);
    R(1, 1) = mk(cs, -a * s);
    return R;
}

__device__ Mat<3> expmap(const Mat<3>& X) {
    // Q = -i X (hermitian, traceless); exp(i Q) = f0 + f1 Q + f2 Q^2 with
    // c0 = det Q, c1 = tr(Q^2) / 2
    Mat<3> Q;
#pragma unroll
    for (int k = 0; k < 9; ++k) Q.a[k] = mk(X.a[k].im, -X.a[k].re);
    const Mat<3> Q2 = mul(Q, Q);
    const zd det = Q(0, 0) * (Q(1, 1) * Q(2, 2) - Q(1, 2) * Q(2, 1)) - Q(0, 1) * (Q(1, 0) * Q(2, 2) - Q(1, 2) * Q(2, 0)) +
                   Q(0, 2) * (Q(1, 0) * Q(2, 1) - Q(1, 1) * Q(2, 0));
    double c0 = det.re;
    const double c1 = 0.5 * (Q2(0, 0).re + Q2(1, 1).re + Q2(2, 2).re);
    Mat<3> R;
    if (c1 == 0.0) {
#pragma unroll
        for (int i = 0; i < 3; ++i)
#pragma unroll
            for (int j = 0; j < 3; ++j) R(i, j) = mk(i == j ? 1.0 : 0.0, 0.0);
        return R;
    }
    const bool neg = c0 < 0;
    if (neg) c0 = -c0;
    const double c0max = 2.0 * (c1 / 3.0) * sqrt(c1 / 3.0);
    const double theta = acos(fmin(1.0, c0 / c0max));
    const double u = sqrt(c1 / 3.0) * cos(theta / 3.0);
    const double w = sqrt(c1) * sin(theta / 3.0);
    const double uu = u * u, ww = w * w;
    double s2u, c2u, su, cu, sw, cw;
    sincos(2.0 * u, &s2u, &c2u);
    sincos(u, &su, &cu);
    sincos(w, &sw, &cw);
    // xi0(w) = sin(w) / w, its series below |w| = 0.05
    const double xi = fabs(w) < 0.05 ? 1.0 - ww / 6.0 * (1.0 - ww / 20.0 * (1.0 - ww / 42.0)) : sw / w;
    const zd e2 = mk(c2u, s2u), e1 = mk(cu, -su);
    // h0 = (u^2 - w^2) e^{2iu} + e^{-iu} (8 u^2 cos w + 2 i u (3 u^2 + w^2) xi)
    // h1 = 2 u e^{2iu} - e^{-iu} (2 u cos w - i (3 u^2 - w^2) xi)
    // h2 = e^{2iu} - e^{-iu} (cos w + 3 i u xi)
    const zd h0 = (uu - ww) * e2 + e1 * mk(8.0 * uu * cw, 2.0 * u * (3.0 * uu + ww) * xi);
    const zd h1 = (2.0 * u) * e2 - e1 * mk(2.0 * u * cw, -(3.0 * uu - ww) * xi);
    const zd h2 = e2 - e1 * mk(cw, 3.0 * u * xi);
    const double inv = 1.0 / (9.0 * uu - ww);
    zd f0 = inv * h0, f1 = inv * h1, f2 = inv * h2;
    if (neg) {  // exp(iQ) for det Q < 0 from f_j(-c0) = (-1)^j conj(f_j(c0))
        f0 = conj(f0);
        f1 = -conj(f1);
        f2 = conj(f2);
    }
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int j = 0; j < 3; ++j) {
            zd v = f1 * Q(i, j) + f2 * Q2(i, j);
            if (i == j) v = v + f0;
            R(i, j) = v;
        }
    return R;
}

template <int N>
__global__ __launch_bounds__(128) void k_stout(Lex g, double rho, const zd* __restrict__ in, zd* __restrict__ out) {
    const long j = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= 4 * g.V) return;
    const int mu = (int)(j / g.V);
    const long s = j - (long)mu * g.V;
    int c[4];
    {
        long r = s;
        c[3] = (int)(r % g.L[3]);
        r /= g.L[3];
        c[2] = (int)(r % g.L[2]);
        r /= g.L[2];
        c[1] = (int)(r % g.L[1]);
        c[0] = (int)(r / g.L[1]);
    }
    int cu[4] = {c[0], c[1], c[2], c[3]};
    cu[mu] = c[mu] + 1 == g.L[mu] ? 0 : c[mu] + 1;
    const long smu = lex_of(g, cu);
    Mat<N> st;
#pragma unroll
    for (int k = 0; k < N * N; ++k) st.a[k] = mk(0.0, 0.0);
    for (int nu = 0; nu < 4; ++nu) {
        if (nu == mu) continue;
        int cn[4] = {c[0], c[1], c[2], c[3]}, cd[4] = {c[0], c[1], c[2], c[3]};
        cn[nu] = c[nu] + 1 == g.L[nu] ? 0 : c[nu] + 1;
        cd[nu] = c[nu] == 0 ? g.L[nu] - 1 : c[nu] - 1;
        int cmd[4] = {cd[0], cd[1], cd[2], cd[3]};
        cmd[mu] = cd[mu] + 1 == g.L[mu] ? 0 : cd[mu] + 1;
        const long snu = lex_of(g, cn), sdn = lex_of(g, cd), smudn = lex_of(g, cmd);
        // U_nu(x+mu) U_mu(x+nu)^+ U_nu(x)^+
        const Mat<N> up = mul_nd(mul_nd(ld<N>(in, g.V, smu, nu), ld<N>(in, g.V, snu, mu)), ld<N>(in, g.V, s, nu));
        // U_nu(x+mu-nu)^+ U_mu(x-nu)^+ U_nu(x-nu) = (U_mu(x-nu) U_nu(x+mu-nu))^+ U_nu(x-nu)
        const Mat<N> dn = mul_dn(mul(ld<N>(in, g.V, sdn, mu), ld<N>(in, g.V, smudn, nu)), ld<N>(in, g.V, sdn, nu));
#pragma unroll
        for (int k = 0; k < N * N; ++k) st.a[k] = st.a[k] + up.a[k] + dn.a[k];
    }
    const Mat<N> U = ld<N>(in, g.V, s, mu);
    // Omega = rho S^+ U^+ = rho (U S)^+; X = (Omega - Omega^+) / 2 - tr / N
    const Mat<N> US = mul(U, st);
    Mat<N> X;
    zd tr = mk(0.0, 0.0);
#pragma unroll
    for (int i = 0; i < N; ++i)
#pragma unroll
        for (int k = 0; k < N; ++k) {
            // Omega(i,k) = rho conj(US(k,i)), Omega^+(i,k) = rho US(i,k)
            const zd v = (0.5 * rho) * (conj(US(k, i)) - US(i, k));
            X(i, k) = v;
            if (i == k) tr = tr + v;
        }
#pragma unroll
    for (int i = 0; i < N; ++i) X(i, i) = X(i, i) - (1.0 / N) * tr;
    const Mat<N> E = mul(expmap(X), U);
#pragma unroll
    for (int k = 0; k < N * N; ++k) out[((long)mu * N * N + k) * g.V + s] = E.a[k];
}

// host AoS [V][4][N*N] <-> device [4][N*N][V]
template <bool TO_DEV>
__global__ void k_gauge_transpose(long V, int nn, const zd* __restrict__ src, zd* __restrict__ dst) {
    const long j = (long)blockIdx.x * blockDim.x + threadIdx.x;  // (mu, k, s) of the device layout
    if (j >= 4L * nn * V) return;
    const long s = j % V;
    const long mk_ = j / V;  // mu * nn + k
    const long a = s * 4 * nn + mk_;  // AoS index
    if (TO_DEV) dst[j] = src[a];
    else dst[a] = src[j];
}

}  // namespace
}  // namespace lym

using lym::zd;

extern "C" int lym_stout_smear(const int dims[4], int ncolors, double rho, int levels, const double* in,
                               double* out) {
    if (!dims || !in || !out) return fail(LYM_ERR_INVALID, "null argument");
    if (ncolors != 2 && ncolors != 3) return fail(LYM_ERR_INVALID, "ncolors must be 2 or 3 (got %d)", ncolors);
    if (levels < 0) return fail(LYM_ERR_INVALID, "levels must be >= 0");
    lym::Lex g;
    g.V = 1;
    for (int m = 0; m < 4; ++m) {
        if (dims[m] < 2) return fail(LYM_ERR_INVALID, "lattice extent %d must be >= 2", dims[m]);
        g.L[m] = dims[m];
        g.V *= dims[m];
    }
    const int nn = ncolors * ncolors;
    const size_t n = 4UL * nn * (size_t)g.V, bytes = n * sizeof(zd);
    if (levels == 0) {
        if (out != in) std::memmove(out, in, bytes);
        return LYM_OK;
    }
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) return fail(LYM_ERR_DEVICE, "no HIP device available");
    hipStream_t st = nullptr;
    zd *a = nullptr, *b = nullptr, *stage = nullptr;
    int rc = LYM_OK;
    auto chk = [&](hipError_t e, const char* what) {
        if (e != hipSuccess && rc == LYM_OK) rc = fail(LYM_ERR_DEVICE, "%s: %s", what, hipGetErrorString(e));
        return rc == LYM_OK;
    };
    if (chk(hipStreamCreateWithFlags(&st, hipStreamNonBlocking), "stream") &&
        chk(hipMalloc((void**)&a, bytes), "hipMalloc") && chk(hipMalloc((void**)&b, bytes), "hipMalloc") &&
        chk(hipMalloc((void**)&stage, bytes), "hipMalloc") &&
        chk(hipMemcpyAsync(stage, in, bytes, hipMemcpyHostToDevice, st), "upload")) {
        const unsigned tb = 256, grid_t = (unsigned)((n + tb - 1) / tb);
        hipLaunchKernelGGL(lym::k_gauge_transpose<true>, dim3(grid_t), dim3(tb), 0, st, g.V, nn, stage, a);
        const unsigned grid_s = (unsigned)((4 * g.V + 127) / 128);
        for (int l = 0; l < levels; ++l) {
            if (ncolors == 3) hipLaunchKernelGGL(lym::k_stout<3>, dim3(grid_s), dim3(128), 0, st, g, rho, a, b);
            else hipLaunchKernelGGL(lym::k_stout<2>, dim3(grid_s), dim3(128), 0, st, g, rho, a, b);
            std::swap(a, b);
        }
        hipLaunchKernelGGL(lym::k_gauge_transpose<false>, dim3(grid_t), dim3(tb), 0, st, g.V, nn, a, stage);
        if (chk(hipGetLastError(), "kernel launch") &&
            chk(hipMemcpyAsync(out, stage, bytes, hipMemcpyDeviceToHost, st), "download"))
            chk(hipStreamSynchronize(st), "synchronize");
    }
    if (a) (void)hipFree(a);
    if (b) (void)hipFree(b);
    if (stage) (void)hipFree(stage);
    if (st) (void)hipStreamDestroy(st);
    return rc;
}
