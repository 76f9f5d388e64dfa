// correlators.hip -- connected meson contractions of point-source propagators
// (source/correlators/MesonCorrelator.cpp:92-149).
//
// The propagator is 12 FP64 spinor fields S_{c,alpha} = D^-1 e_{alpha,c}
// (source spin alpha, colour c; the field's component 3 mu + d is sink spin
// mu, colour d).  Per site, summed over the colours c, d:
//   pion   += sum_{alpha,mu} |S_{c,alpha}[mu,d]|^2
//   scalar += Re sum_{alpha,beta,mu,nu} G[mu][nu] G[alpha][beta]
//                  conj(S_{c,alpha}[mu,d]) S_{c,beta}[nu,d],          G = gamma5 gamma_3
//   ps     += Re sum_{alpha,beta} g[alpha][beta] ( - sum_mu conj(S_{c,alpha}[mu,d]) S_{c,beta}[mu,d]
//                                                  + sum_mu conj(S_{c,mu}[alpha,d]) S_{c,mu}[beta,d] ),  g = gamma_3
// (gamma matrices of source/utils/Gamma.cpp:8-9), then each is summed over
// the sites of every timeslice in a fixed order.
#include "lym_device.h"

namespace lym {

// gamma5 gamma_3 and gamma_3 in the chiral basis (both real)
__device__ constexpr double G53[4][4] = {{0, 0, 1, 0}, {0, 0, 0, 1}, {-1, 0, 0, 0}, {0, -1, 0, 0}};
__device__ constexpr double G3[4][4] = {{0, 0, 1, 0}, {0, 0, 0, 1}, {1, 0, 0, 0}, {0, 1, 0, 0}};

// one thread per site of both parities (j = parity * stride + index): the
// three contractions into out[q * 2 * stride + j]
__global__ void k_meson_site(Geom g, const zd* const* prop, double* out) {
    const long j = (long)blockIdx.x * blockDim.x + threadIdx.x;
    const long n = 2L * g.stride;
    if (j >= n) return;
    const int p = (int)(j / g.stride), i = (int)(j % g.stride);
    if (i >= g.Vh) {
        out[j] = out[n + j] = out[2 * n + j] = 0.0;
        return;
    }
    double pion = 0, scalar = 0, ps = 0;
    for (int c = 0; c < 3; ++c) {
#pragma unroll
        for (int d = 0; d < 3; ++d) {
            // A[alpha][mu] = S_{c,alpha}[mu, d]
            zd A[4][4];
#pragma unroll
            for (int al = 0; al < 4; ++al) {
                const zd* f = prop[c * 4 + al] + (size_t)p * 12 * g.stride;
#pragma unroll
                for (int mu = 0; mu < 4; ++mu) A[al][mu] = f[(size_t)(3 * mu + d) * g.stride + i];
            }
#pragma unroll
            for (int al = 0; al < 4; ++al)
#pragma unroll
                for (int mu = 0; mu < 4; ++mu) pion += A[al][mu].re * A[al][mu].re + A[al][mu].im * A[al][mu].im;
#pragma unroll
            for (int al = 0; al < 4; ++al)
#pragma unroll
                for (int be = 0; be < 4; ++be) {
                    if (G53[al][be] != 0) {
                        double s = 0;
#pragma unroll
                        for (int mu = 0; mu < 4; ++mu)
#pragma unroll
                            for (int nu = 0; nu < 4; ++nu)
                                if (G53[mu][nu] != 0)
                                    s += G53[mu][nu] * (A[al][mu].re * A[be][nu].re + A[al][mu].im * A[be][nu].im);
                        scalar += G53[al][be] * s;
                    }
                    if (G3[al][be] != 0) {
                        double s = 0;
#pragma unroll
                        for (int mu = 0; mu < 4; ++mu) {
                            s -= A[al][mu].re * A[be][mu].re + A[al][mu].im * A[be][mu].im;
                            s += A[mu][al].re * A[mu][be].re + A[mu][al].im * A[mu][be].im;
                        }
                        ps += G3[al][be] * s;
                    }
                }
        }
    }
    out[j] = pion;
    out[n + j] = scalar;
    out[2 * n + j] = ps;
}

// block (t, q): the sum of contraction q over the sites of timeslice t, in a
// fixed order (strided per-thread sums, then an LDS tree)
__global__ __launch_bounds__(256) void k_meson_timeslice(Geom g, const double* site_vals, double* res) {
    __shared__ double sh[256];
    const int t = blockIdx.x, q = blockIdx.y;
    const long n = 2L * g.stride;
    const double* v = site_vals + q * n;
    const long nr = (long)g.L[0] * g.L[1] * g.L[2];
    double s = 0;
    for (long r = threadIdx.x; r < nr; r += blockDim.x) {
        // r = (x Ly + y) Lz + z; lexicographic site r Lt + t, class (x + y + z + t) & tsh
        const int z = (int)(r % g.L[2]);
        const long r2 = r / g.L[2];
        const int y = (int)(r2 % g.L[1]), x = (int)(r2 / g.L[1]);
        const int par = (x + y + z + t) & g.tsh;
        s += v[(long)par * g.stride + r * g.Lth + (t >> g.tsh)];
    }
    sh[threadIdx.x] = s;
    __syncthreads();
    for (int off = 128; off > 0; off >>= 1) {
        if ((int)threadIdx.x < off) sh[threadIdx.x] += sh[threadIdx.x + off];
        __syncthreads();
    }
    if (threadIdx.x == 0) res[q * g.L[3] + t] = sh[0];
}

__global__ void k_set_one(zd* f, long idx) { f[idx] = mk(1.0, 0.0); }

void launch_meson_contractions(const Geom& g, const zd* const* prop_dev, double* site_vals, double* res,
                               hipStream_t st) {
    const long n = 2L * g.stride;
    hipLaunchKernelGGL(k_meson_site, dim3((unsigned)((n + 127) / 128)), dim3(128), 0, st, g, prop_dev, site_vals);
    hipLaunchKernelGGL(k_meson_timeslice, dim3(g.L[3], 3), dim3(256), 0, st, g, site_vals, res);
}

void launch_set_one(zd* f, long idx, hipStream_t st) { hipLaunchKernelGGL(k_set_one, dim3(1), dim3(1), 0, st, f, idx); }

}  // namespace lym
