// lym_device.h -- device-side link loaders, colour matvecs and spinor loads
// shared by the stencil and the clover builder.
#pragma once

#include "lym_kernels.h"

#ifndef LYM_NT_LINKS
#define LYM_NT_LINKS 0
#endif

namespace lym {

// Non-temporal 16-byte access for data touched once per launch (clover blocks,
// axpy operands, outputs): it streams past L2 instead of evicting the
// neighbour spinors and links that other sites of the launch reuse.
typedef double nt_d2 __attribute__((ext_vector_type(2)));
__device__ __forceinline__ cplx<double> ld_nt(const cplx<double>* p) {
    nt_d2 v = __builtin_nontemporal_load(reinterpret_cast<const nt_d2*>(p));
    return {v.x, v.y};
}
__device__ __forceinline__ void st_nt(cplx<double>* p, cplx<double> z) {
    nt_d2 v = {z.re, z.im};
    __builtin_nontemporal_store(v, reinterpret_cast<nt_d2*>(p));
}

// ---------------------------------------------------------------- link loaders
// LK: 0 complex 18-real, 1 complex 12-real (row 2 = conj(row0 x row1)),
//     2 real 9 (adjoint), 3 real 6 (row 2 = row0 x row1)
template <int LK>
struct LinkMat;
template <>
struct LinkMat<0> { cplx<double> m[9]; };
template <>
struct LinkMat<1> { cplx<double> m[9]; };
template <>
struct LinkMat<2> { double m[9]; };
template <>
struct LinkMat<3> { double m[9]; };

template <int LK>
__device__ __forceinline__ void load_link(const void* base, int stride, int p, int mu, int idx, LinkMat<LK>& L) {
    if constexpr (LK == 0 || LK == 1) {
        constexpr int NC = (LK == 0) ? 9 : 6;
        const cplx<double>* b = reinterpret_cast<const cplx<double>*>(base) + (size_t)((p * 4 + mu) * NC) * stride + idx;
#pragma unroll
        for (int k = 0; k < NC; ++k) {
#if LYM_NT_LINKS
            L.m[k] = ld_nt(b + (size_t)k * stride);
#else
            L.m[k] = b[(size_t)k * stride];
#endif
        }
        if constexpr (LK == 1) {
            // row 2 = conj(row0 x row1)
            L.m[6] = conj(L.m[1] * L.m[5] - L.m[2] * L.m[4]);
            L.m[7] = conj(L.m[2] * L.m[3] - L.m[0] * L.m[5]);
            L.m[8] = conj(L.m[0] * L.m[4] - L.m[1] * L.m[3]);
        }
    } else {
        constexpr int NC = (LK == 2) ? 9 : 6;
        const double* b = reinterpret_cast<const double*>(base) + (size_t)((p * 4 + mu) * NC) * stride + idx;
#pragma unroll
        for (int k = 0; k < NC; ++k) L.m[k] = b[(size_t)k * stride];
        if constexpr (LK == 3) {
            L.m[6] = L.m[1] * L.m[5] - L.m[2] * L.m[4];
            L.m[7] = L.m[2] * L.m[3] - L.m[0] * L.m[5];
            L.m[8] = L.m[0] * L.m[4] - L.m[1] * L.m[3];
        }
    }
}

template <int LK>
__device__ __forceinline__ void matvec(const LinkMat<LK>& L, const cplx<double> h[3], cplx<double> o[3]) {
#pragma unroll
    for (int i = 0; i < 3; ++i) {
        if constexpr (LK <= 1) {
            cplx<double> s = L.m[3 * i] * h[0];
            cfma(s, L.m[3 * i + 1], h[1]);
            cfma(s, L.m[3 * i + 2], h[2]);
            o[i] = s;
        } else {
            o[i] = mk(L.m[3 * i] * h[0].re + L.m[3 * i + 1] * h[1].re + L.m[3 * i + 2] * h[2].re,
                      L.m[3 * i] * h[0].im + L.m[3 * i + 1] * h[1].im + L.m[3 * i + 2] * h[2].im);
        }
    }
}

template <int LK>
__device__ __forceinline__ void matvec_dag(const LinkMat<LK>& L, const cplx<double> h[3], cplx<double> o[3]) {
#pragma unroll
    for (int i = 0; i < 3; ++i) {
        if constexpr (LK <= 1) {
            cplx<double> s = cmulc(L.m[i], h[0]);
            cfmac(s, L.m[3 + i], h[1]);
            cfmac(s, L.m[6 + i], h[2]);
            o[i] = s;
        } else {
            o[i] = mk(L.m[i] * h[0].re + L.m[3 + i] * h[1].re + L.m[6 + i] * h[2].re,
                      L.m[i] * h[0].im + L.m[3 + i] * h[1].im + L.m[6 + i] * h[2].im);
        }
    }
}

__device__ __forceinline__ void load_spinor(const cplx<double>* f, int stride, int idx, cplx<double> s[12]) {
#pragma unroll
    for (int k = 0; k < 12; ++k) s[k] = f[(size_t)k * stride + idx];
}

// o = A v for one packed hermitian 6x6 block (base points at the block's first component)
__device__ __forceinline__ void herm6(const cplx<double>* A, int stride, int idx, const cplx<double> v[6],
                                      cplx<double> o[6]) {
    double d[6];
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        cplx<double> dd = ld_nt(A + (size_t)k * stride + idx);
        d[2 * k] = dd.re;
        d[2 * k + 1] = dd.im;
    }
#pragma unroll
    for (int i = 0; i < 6; ++i) o[i] = d[i] * v[i];
    int k = 3;
#pragma unroll
    for (int i = 1; i < 6; ++i) {
#pragma unroll
        for (int j = 0; j < i; ++j) {
            cplx<double> a = ld_nt(A + (size_t)k * stride + idx);
            cfma(o[i], a, v[j]);
            cfmac(o[j], a, v[i]);
            ++k;
        }
    }
}

}  // namespace lym
