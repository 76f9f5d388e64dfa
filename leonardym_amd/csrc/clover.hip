// clover.hip -- clover field strength, site-diagonal blocks and their inverses.
//
// F_k = (Q_k - Q_k^dagger)/8 from the four clover leaves of each plane
// k = (01, 02, 03, 12, 13, 23) (source/dirac_operators/ImprovedDiracWilsonOperator.cpp:502-523),
// no trace subtraction.  The kappa-independent part B of the clover term is
// stored as two packed hermitian 6x6 blocks per site (the chiral combination of
// ImprovedDiracWilsonOperator.cpp:200-243); the operator's diagonal blocks are
// A = 1 + kappa csw gamma5 B, and R_ee^-1 is their blockwise inverse
// (EvenOddImprovedDiracWilsonOperator.cpp:285-317 inverts the full 12x12 instead).
// Links are the unsigned ones: each leaf that crosses the time boundary holds
// two time links, so the antiperiodic sign cancels in F.
#include "lym_device.h"

namespace lym {

struct M3 {
    zd a[9];  // row-major
};

__device__ __forceinline__ M3 m3mul(const M3& A, const M3& B) {
    M3 C;
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int j = 0; j < 3; ++j) {
            zd s = A.a[3 * i] * B.a[j];
            cfma(s, A.a[3 * i + 1], B.a[3 + j]);
            cfma(s, A.a[3 * i + 2], B.a[6 + j]);
            C.a[3 * i + j] = s;
        }
    return C;
}

__device__ __forceinline__ M3 m3dag(const M3& A) {
    M3 C;
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int j = 0; j < 3; ++j) C.a[3 * i + j] = conj(A.a[3 * j + i]);
    return C;
}

template <int LK>
__device__ __forceinline__ M3 get_link(const Geom& g, const void* links, const int c[4], int mu) {
    int p = site_parity(g, c);
    int idx = eo_index(g, c[0], c[1], c[2], c[3]);
    LinkMat<LK, double> L;
    load_link<0, LK>(links, g.stride, p, mu, idx, L);
    M3 m;
#pragma unroll
    for (int k = 0; k < 9; ++k) {
        if constexpr (LK <= 1) m.a[k] = L.m[k];
        else m.a[k] = mk(L.m[k], 0.0);
    }
    return m;
}

__device__ __forceinline__ void shift(const Geom& g, const int c[4], int mu, int d, int o[4]) {
    for (int k = 0; k < 4; ++k) o[k] = c[k];
    int v = c[mu] + d;
    o[mu] = v < 0 ? v + g.L[mu] : (v >= g.L[mu] ? v - g.L[mu] : v);
}

// F for plane (m, n) at coordinates c
template <int LK>
__device__ M3 field_strength(const Geom& g, const void* links, const int c[4], int m, int n) {
    int xm[4], xn[4], xmn[4], xpm[4], xpn[4], xnpm[4], xmpn[4];
    shift(g, c, m, -1, xm);
    shift(g, c, n, -1, xn);
    shift(g, xm, n, -1, xmn);
    shift(g, c, m, +1, xpm);
    shift(g, c, n, +1, xpn);
    shift(g, xn, m, +1, xnpm);
    shift(g, xm, n, +1, xmpn);
    // each leaf loads its own four links (the repeats hit in cache): fewer
    // link matrices live at once
    M3 Q = m3mul(m3mul(m3dag(get_link<LK>(g, links, xm, m)), m3dag(get_link<LK>(g, links, xmn, n))),
                 m3mul(get_link<LK>(g, links, xmn, m), get_link<LK>(g, links, xn, n)));
    {
        M3 l = m3mul(m3mul(m3dag(get_link<LK>(g, links, xn, n)), get_link<LK>(g, links, xn, m)),
                     m3mul(get_link<LK>(g, links, xnpm, n), m3dag(get_link<LK>(g, links, c, m))));
        for (int k = 0; k < 9; ++k) Q.a[k] = Q.a[k] + l.a[k];
    }
    {
        M3 l = m3mul(m3mul(get_link<LK>(g, links, c, m), get_link<LK>(g, links, xpm, n)),
                     m3mul(m3dag(get_link<LK>(g, links, xpn, m)), m3dag(get_link<LK>(g, links, c, n))));
        for (int k = 0; k < 9; ++k) Q.a[k] = Q.a[k] + l.a[k];
    }
    {
        M3 l = m3mul(m3mul(get_link<LK>(g, links, c, n), m3dag(get_link<LK>(g, links, xmpn, m))),
                     m3mul(m3dag(get_link<LK>(g, links, xm, n)), get_link<LK>(g, links, xm, m)));
        for (int k = 0; k < 9; ++k) Q.a[k] = Q.a[k] + l.a[k];
    }
    M3 F;
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int j = 0; j < 3; ++j) F.a[3 * i + j] = 0.125 * (Q.a[3 * i + j] - conj(Q.a[3 * j + i]));
    return F;
}

__device__ __forceinline__ zd ti(zd z) { return mulph<1>(z); }

// store one hermitian 6x6 (full, row-major h[6][6]) packed
__device__ __forceinline__ void store_herm6(zd* base, int stride, int idx, const zd h[36]) {
    for (int k = 0; k < 3; ++k) base[(size_t)k * stride + idx] = mk(h[(2 * k) * 7].re, h[(2 * k + 1) * 7].re);
    int k = 3;
    for (int i = 1; i < 6; ++i)
        for (int j = 0; j < i; ++j) {
            base[(size_t)k * stride + idx] = h[6 * i + j];
            ++k;
        }
}

__device__ __forceinline__ void load_herm6(const zd* base, int stride, int idx, zd h[36]) {
    for (int k = 0; k < 3; ++k) {
        zd d = base[(size_t)k * stride + idx];
        h[(2 * k) * 7] = mk(d.re, 0.0);
        h[(2 * k + 1) * 7] = mk(d.im, 0.0);
    }
    int k = 3;
    for (int i = 1; i < 6; ++i)
        for (int j = 0; j < i; ++j) {
            zd a = base[(size_t)k * stride + idx];
            h[6 * i + j] = a;
            h[6 * j + i] = conj(a);
            ++k;
        }
}

// the diagonal and lower entries of a 3x3 spin block: (0,0) (1,1) (2,2) (1,0) (2,0) (2,1)
__device__ constexpr int DI[6] = {0, 1, 2, 1, 2, 2}, DJ[6] = {0, 1, 2, 0, 0, 1};

// one plane's F into the spin blocks of the upper (du, X) and lower (dl, Y) block
template <int PL>
__device__ __forceinline__ void acc_plane(const M3& F, zd du[6], zd dl[6], zd X[9], zd Y[9]) {
    if constexpr (PL == 0 || PL == 5) {
#pragma unroll
        for (int k = 0; k < 6; ++k) {
            const zd f = ti(F.a[3 * DI[k] + DJ[k]]);
            du[k] = PL == 0 ? du[k] - f : du[k] + f;
            dl[k] = dl[k] + f;
        }
    } else {
#pragma unroll
        for (int k = 0; k < 9; ++k) {
            const zd f = F.a[k];
            if constexpr (PL == 1) {
                X[k] = X[k] - f;
                Y[k] = Y[k] + f;
            } else if constexpr (PL == 4) {
                X[k] = X[k] - f;
                Y[k] = Y[k] - f;
            } else if constexpr (PL == 2) {
                X[k] = X[k] + ti(f);
                Y[k] = Y[k] + ti(f);
            } else {
                X[k] = X[k] - ti(f);
                Y[k] = Y[k] + ti(f);
            }
        }
    }
}

// the packed hermitian 6x6 of spin blocks (0,0) = d, (1,1) = -d, (1,0) = X:
// diagonal pairs, then the lower entries row by row (the store_herm6 order)
__device__ __forceinline__ void store_spin_blocks(zd* base, int stride, int idx, const zd d[6], const zd X[9]) {
    base[idx] = mk(d[0].re, d[1].re);
    base[(size_t)stride + idx] = mk(d[2].re, -d[0].re);
    base[(size_t)2 * stride + idx] = mk(-d[1].re, -d[2].re);
    const zd e[15] = {d[3],  d[4], d[5],                                     // rows 1, 2
                      X[0],  X[1], X[2],                                     // row 3
                      X[3],  X[4], X[5], -d[3],                              // row 4
                      X[6],  X[7], X[8], -d[4], -d[5]};                      // row 5
#pragma unroll
    for (int k = 0; k < 15; ++k) base[(size_t)(3 + k) * stride + idx] = e[k];
}

// One thread per site.  Each XCD takes a contiguous range of site blocks with
// the two parities of a block interleaved, so the links a site's leaves share
// with its neighbours (both parities) are fetched once into that XCD's L2;
// dealt round-robin, neighbouring blocks ran on different XCDs and every XCD
// fetched them again.
template <int LK>
__global__ __launch_bounds__(128) void k_clover_B(Geom g, const void* links, zd* B) {
    const int lb = xcd_block((int)blockIdx.x, (int)gridDim.x);
    const int p = g.tsh ? (lb & 1) : 0;
    const int idx = (g.tsh ? lb >> 1 : lb) * 128 + (int)threadIdx.x;
    if (idx >= g.Vh) return;
    int c[4];
    eo_coords(g, p, idx, c);
    // The two 6x6 blocks (spins 0,1 and 2,3; rows 3 s + colour) as the 3x3
    // spin blocks they are made of: upper (0,0) = i(F5 - F0), (1,1) = -(0,0),
    // (1,0) = -F1 + iF2 - iF3 - F4; lower (0,0) = i(F0 + F5), (1,1) = -(0,0),
    // (1,0) = F1 + iF2 + iF3 - F4 (planes 0..5 = 01 02 03 12 13 23), summed in
    // plane order.  The (0,1) blocks are the adjoints of (1,0) and are not stored.
    zd du[6], dl[6], X[9], Y[9];
#pragma unroll
    for (int k = 0; k < 6; ++k) du[k] = dl[k] = mk(0.0, 0.0);
#pragma unroll
    for (int k = 0; k < 9; ++k) X[k] = Y[k] = mk(0.0, 0.0);
    acc_plane<0>(field_strength<LK>(g, links, c, 0, 1), du, dl, X, Y);
    acc_plane<1>(field_strength<LK>(g, links, c, 0, 2), du, dl, X, Y);
    acc_plane<2>(field_strength<LK>(g, links, c, 0, 3), du, dl, X, Y);
    acc_plane<3>(field_strength<LK>(g, links, c, 1, 2), du, dl, X, Y);
    acc_plane<4>(field_strength<LK>(g, links, c, 1, 3), du, dl, X, Y);
    acc_plane<5>(field_strength<LK>(g, links, c, 2, 3), du, dl, X, Y);
    zd* base = B + (size_t)p * 36 * g.stride;
    store_spin_blocks(base, g.stride, idx, du, X);
    store_spin_blocks(base + (size_t)18 * g.stride, g.stride, idx, dl, Y);
}

// A = 1 + kc gamma5 B: upper block 1 + kc B_u, lower block 1 - kc B_l
__global__ void k_clover_A(Geom g, const zd* __restrict__ B, double kc, zd* __restrict__ A) {
    long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    long n = 2L * 36 * g.stride;
    if (i >= n) return;
    int comp = (int)((i / g.stride) % 36);
    double s = comp < 18 ? kc : -kc;
    int k = comp % 18;
    zd b = B[i];
    zd a = mk(s * b.re, s * b.im);
    if (k < 3) {
        // diagonal pairs
        a.re += 1.0;
        a.im += 1.0;
    }
    A[i] = a;
}

// Inverse of a hermitian 6x6 block through A = L D L^H (L unit lower
// triangular, D real) without pivoting: every index is static, so the whole
// factorisation lives in registers.  Returns false when a pivot is below
// 1e-12 of the largest diagonal entry (the caller then redoes the parity with
// the pivoted kernel below).
__device__ __forceinline__ bool herm6_inverse_ldl(const zd M[36], zd R[36]) {
    zd L[36];
    double d[6], di[6];
    double scale = 0.0;
#pragma unroll
    for (int i = 0; i < 6; ++i) scale = fmax(scale, fabs(M[7 * i].re));
    bool ok = true;
#pragma unroll
    for (int j = 0; j < 6; ++j) {
        double dj = M[7 * j].re;
#pragma unroll
        for (int k = 0; k < j; ++k) dj -= (L[6 * j + k].re * L[6 * j + k].re + L[6 * j + k].im * L[6 * j + k].im) * d[k];
        d[j] = dj;
        ok = ok && fabs(dj) > 1e-12 * scale;
        di[j] = 1.0 / dj;
#pragma unroll
        for (int i = j + 1; i < 6; ++i) {
            zd s = M[6 * i + j];
#pragma unroll
            for (int k = 0; k < j; ++k) s = s - L[6 * i + k] * (d[k] * conj(L[6 * j + k]));
            L[6 * i + j] = di[j] * s;
        }
    }
#pragma unroll
    for (int c = 0; c < 6; ++c) {
        zd y[6], x[6];
#pragma unroll
        for (int i = 0; i < 6; ++i) {
            zd v = mk(i == c ? 1.0 : 0.0, 0.0);
#pragma unroll
            for (int k = 0; k < i; ++k) v = v - L[6 * i + k] * y[k];
            y[i] = v;
        }
#pragma unroll
        for (int i = 5; i >= 0; --i) {
            zd v = di[i] * y[i];
#pragma unroll
            for (int k = i + 1; k < 6; ++k) v = v - conj(L[6 * k + i]) * x[k];
            x[i] = v;
        }
#pragma unroll
        for (int i = 0; i < 6; ++i) R[6 * i + c] = x[i];
    }
    return ok;
}

__global__ __launch_bounds__(128) void k_clover_inv_ldl(Geom g, const zd* __restrict__ A, zd* __restrict__ Ainv,
                                                       int* fallback) {
    const int idx = blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= g.Vh) return;
#pragma unroll 1
    for (int blk = 0; blk < 2; ++blk) {
        zd M[36], R[36];
        load_herm6(A + (size_t)18 * blk * g.stride, g.stride, idx, M);
        if (!herm6_inverse_ldl(M, R)) atomicAdd(fallback, 1);
        store_herm6(Ainv + (size_t)18 * blk * g.stride, g.stride, idx, R);
    }
}

// blockwise inverse on parity-0 sites: Gauss-Jordan with partial pivoting
__global__ __launch_bounds__(64) void k_clover_inv(Geom g, const zd* __restrict__ A, zd* __restrict__ Ainv,
                                                   int* flag) {
    int idx = blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= g.Vh) return;
    for (int blk = 0; blk < 2; ++blk) {
        zd M[36], R[36];
        load_herm6(A + (size_t)18 * blk * g.stride, g.stride, idx, M);
        for (int k = 0; k < 36; ++k) R[k] = mk(k % 7 == 0 ? 1.0 : 0.0, 0.0);
        bool ok = true;
        for (int col = 0; col < 6; ++col) {
            int piv = col;
            double best = M[6 * col + col].re * M[6 * col + col].re + M[6 * col + col].im * M[6 * col + col].im;
            for (int r = col + 1; r < 6; ++r) {
                double v = M[6 * r + col].re * M[6 * r + col].re + M[6 * r + col].im * M[6 * r + col].im;
                if (v > best) { best = v; piv = r; }
            }
            if (best == 0.0) { ok = false; break; }
            if (piv != col)
                for (int j = 0; j < 6; ++j) {
                    zd t = M[6 * col + j]; M[6 * col + j] = M[6 * piv + j]; M[6 * piv + j] = t;
                    t = R[6 * col + j]; R[6 * col + j] = R[6 * piv + j]; R[6 * piv + j] = t;
                }
            zd d = M[6 * col + col];
            double den = d.re * d.re + d.im * d.im;
            zd inv = mk(d.re / den, -d.im / den);
            for (int j = 0; j < 6; ++j) { M[6 * col + j] = inv * M[6 * col + j]; R[6 * col + j] = inv * R[6 * col + j]; }
            for (int r = 0; r < 6; ++r) {
                if (r == col) continue;
                zd f = M[6 * r + col];
                for (int j = 0; j < 6; ++j) {
                    M[6 * r + j] = M[6 * r + j] - f * M[6 * col + j];
                    R[6 * r + j] = R[6 * r + j] - f * R[6 * col + j];
                }
            }
        }
        if (!ok) atomicAdd(flag, 1);
        store_herm6(Ainv + (size_t)18 * blk * g.stride, g.stride, idx, R);
    }
}

// out = blocks * in on one parity (half fields), optional gamma5
__global__ void k_clover_apply(Geom g, const zd* __restrict__ blocks, const zd* __restrict__ in, zd* __restrict__ out,
                               int gamma5) {
    int idx = blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= g.Vh) return;
    zd s[12], t[12];
    load_spinor(in, g.stride, idx, s);
    herm6(blocks, g.stride, idx, s, t);
    herm6(blocks + (size_t)18 * g.stride, g.stride, idx, s + 6, t + 6);
    for (int k = 0; k < 12; ++k) {
        zd v = (gamma5 && k >= 6) ? -t[k] : t[k];
        out[(size_t)k * g.stride + idx] = v;
    }
}

void launch_clover_B(const Geom& g, const void* links, int link_kind, zd* B, hipStream_t st) {
    int nb = (1 + g.tsh) * ((g.Vh + 127) / 128);
    switch (link_kind) {
        case 0: hipLaunchKernelGGL(k_clover_B<0>, dim3(nb), dim3(128), 0, st, g, links, B); break;
        case 1: hipLaunchKernelGGL(k_clover_B<1>, dim3(nb), dim3(128), 0, st, g, links, B); break;
        case 2: hipLaunchKernelGGL(k_clover_B<2>, dim3(nb), dim3(128), 0, st, g, links, B); break;
        default: hipLaunchKernelGGL(k_clover_B<3>, dim3(nb), dim3(128), 0, st, g, links, B); break;
    }
}

void launch_clover_A(const Geom& g, const zd* B, double kc, zd* A, hipStream_t st) {
    long n = 2L * 36 * g.stride;
    hipLaunchKernelGGL(k_clover_A, dim3((n + 255) / 256), dim3(256), 0, st, g, B, kc, A);
}

void launch_clover_inv(const Geom& g, const zd* A, zd* Ainv, int* flag, hipStream_t st) {
    hipLaunchKernelGGL(k_clover_inv, dim3((g.Vh + 63) / 64), dim3(64), 0, st, g, A, Ainv, flag);
}

void launch_clover_inv_ldl(const Geom& g, const zd* A, zd* Ainv, int* fallback, hipStream_t st) {
    hipLaunchKernelGGL(k_clover_inv_ldl, dim3((g.Vh + 127) / 128), dim3(128), 0, st, g, A, Ainv, fallback);
}

void launch_clover_apply(const Geom& g, const zd* blocks, const zd* in, zd* out, int gamma5, hipStream_t st) {
    hipLaunchKernelGGL(k_clover_apply, dim3((g.Vh + 255) / 256), dim3(256), 0, st, g, blocks, in, out, gamma5);
}

}  // namespace lym
