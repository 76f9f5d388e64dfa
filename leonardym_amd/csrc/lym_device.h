// lym_device.h -- device-side field access shared by the stencil, the clover
// builder and the solver kernels, generic over the storage precision and the
// number of right-hand sides.
//
// Storage precisions (PR):
//   0  FP64: spinor double2, links double2 / double, clover double2
//   1  FP32: float2 / float, computed in FP32
//   2  FP16: _Float16 pairs, computed in FP32; spinors and clover blocks carry
//      one FP32 scale per site (and right-hand side) = the largest |real|
//      component, so the stored values lie in [-1, 1]; links are unitary
//      (or orthogonal), their entries already lie in [-1, 1] and need no scale.
// Multi-RHS fields store the right-hand side fastest:
//   element (component k, site s, rhs r) at (k * stride + s) * NR + r,
// so the NR lanes that handle one site read one link (a single broadcast
// transaction) and NR consecutive spinor elements.
#pragma once

#include <type_traits>

#include "lym_kernels.h"

#ifndef LYM_NT_LINKS
#define LYM_NT_LINKS 0
#endif

namespace lym {

typedef _Float16 h2_t __attribute__((ext_vector_type(2)));

template <int PR>
struct Prec;
template <>
struct Prec<0> {
    typedef double T;
    typedef cplx<double> S;  // complex storage element
    typedef double R;        // real storage element
};
template <>
struct Prec<1> {
    typedef float T;
    typedef cplx<float> S;
    typedef float R;
};
template <>
struct Prec<2> {
    typedef float T;
    typedef h2_t S;
    typedef _Float16 R;
};

// Non-temporal access for data touched once per launch (clover blocks, axpy
// operands, outputs): it streams past L2 instead of evicting the neighbour
// spinors and links that other sites of the launch reuse.
template <class E>
__device__ __forceinline__ E ld_nt_raw(const E* p) {
    if constexpr (sizeof(E) == 16) {
        typedef double v2 __attribute__((ext_vector_type(2)));
        v2 v = __builtin_nontemporal_load(reinterpret_cast<const v2*>(p));
        E e;
        __builtin_memcpy(&e, &v, 16);
        return e;
    } else if constexpr (sizeof(E) == 8) {
        typedef float v2 __attribute__((ext_vector_type(2)));
        v2 v = __builtin_nontemporal_load(reinterpret_cast<const v2*>(p));
        E e;
        __builtin_memcpy(&e, &v, 8);
        return e;
    } else {
        static_assert(sizeof(E) == 4, "element size");
        unsigned v = __builtin_nontemporal_load(reinterpret_cast<const unsigned*>(p));
        E e;
        __builtin_memcpy(&e, &v, 4);
        return e;
    }
}
template <class E>
__device__ __forceinline__ void st_nt_raw(E* p, E e) {
    if constexpr (sizeof(E) == 16) {
        typedef double v2 __attribute__((ext_vector_type(2)));
        v2 v;
        __builtin_memcpy(&v, &e, 16);
        __builtin_nontemporal_store(v, reinterpret_cast<v2*>(p));
    } else if constexpr (sizeof(E) == 8) {
        typedef float v2 __attribute__((ext_vector_type(2)));
        v2 v;
        __builtin_memcpy(&v, &e, 8);
        __builtin_nontemporal_store(v, reinterpret_cast<v2*>(p));
    } else {
        unsigned v;
        __builtin_memcpy(&v, &e, 4);
        __builtin_nontemporal_store(v, reinterpret_cast<unsigned*>(p));
    }
}

__device__ __forceinline__ cplx<double> ld_nt(const cplx<double>* p) { return ld_nt_raw(p); }
__device__ __forceinline__ void st_nt(cplx<double>* p, cplx<double> z) { st_nt_raw(p, z); }

// storage element -> compute complex
template <int PR>
__device__ __forceinline__ cplx<typename Prec<PR>::T> to_c(typename Prec<PR>::S s) {
    if constexpr (PR == 2) return {(float)s.x, (float)s.y};
    else return s;
}
template <int PR>
__device__ __forceinline__ typename Prec<PR>::S from_c(cplx<typename Prec<PR>::T> z) {
    if constexpr (PR == 2) {
        h2_t h;
        h.x = (_Float16)z.re;
        h.y = (_Float16)z.im;
        return h;
    } else {
        return z;
    }
}

// Field access through buffer instructions: a wave-uniform 128-bit resource
// (base address in SGPRs) per component stream plus one 32-bit per-lane byte
// offset shared by all streams of a site index.  Plain pointer arithmetic
// leaves a 64-bit VGPR address per stream (v_lshl_add_u64 for each of the ~230
// loads of a stencil site), which costs registers and occupancy.  Each
// component stream is < 2 GiB.  LYM_BUFFER_LOADS=0 keeps plain loads.
#ifndef LYM_BUFFER_LOADS
#define LYM_BUFFER_LOADS 1
#endif

template <class E>
__device__ __forceinline__ const E* lane_at(const E* ubase, unsigned lane_elems) {
    return reinterpret_cast<const E*>(reinterpret_cast<const char*>(ubase) + lane_elems * (unsigned)sizeof(E));
}
template <class E>
__device__ __forceinline__ E* lane_at(E* ubase, unsigned lane_elems) {
    return reinterpret_cast<E*>(reinterpret_cast<char*>(ubase) + lane_elems * (unsigned)sizeof(E));
}

#if LYM_BUFFER_LOADS
// raw buffer resource for gfx9-family (gfx950): stride 0, 2 GiB range
__device__ __forceinline__ __amdgpu_buffer_rsrc_t rsrc_of(const void* base) {
    return __builtin_amdgcn_make_buffer_rsrc(const_cast<void*>(base), (short)0, 0x7fffffff, 0x00020000);
}
// cache policy bits of the buffer instructions: 2 = non-temporal
template <class E, int AUX = 0>
__device__ __forceinline__ E buf_ld(const E* ubase, unsigned lane_elems) {
    const __amdgpu_buffer_rsrc_t r = rsrc_of(ubase);
    const unsigned off = lane_elems * (unsigned)sizeof(E);
    E e;
    if constexpr (sizeof(E) == 16) {
        auto v = __builtin_amdgcn_raw_buffer_load_b128(r, off, 0, AUX);
        __builtin_memcpy(&e, &v, 16);
    } else if constexpr (sizeof(E) == 8) {
        auto v = __builtin_amdgcn_raw_buffer_load_b64(r, off, 0, AUX);
        __builtin_memcpy(&e, &v, 8);
    } else if constexpr (sizeof(E) == 4) {
        auto v = __builtin_amdgcn_raw_buffer_load_b32(r, off, 0, AUX);
        __builtin_memcpy(&e, &v, 4);
    } else {
        static_assert(sizeof(E) == 2, "element size");
        auto v = __builtin_amdgcn_raw_buffer_load_b16(r, off, 0, AUX);
        __builtin_memcpy(&e, &v, 2);
    }
    return e;
}
template <class E, int AUX = 0>
__device__ __forceinline__ void buf_st(E* ubase, unsigned lane_elems, E e) {
    const __amdgpu_buffer_rsrc_t r = rsrc_of(ubase);
    const unsigned off = lane_elems * (unsigned)sizeof(E);
    if constexpr (sizeof(E) == 16) {
        typedef unsigned v4u __attribute__((ext_vector_type(4)));
        v4u v;
        __builtin_memcpy(&v, &e, 16);
        __builtin_amdgcn_raw_buffer_store_b128(v, r, off, 0, AUX);
    } else if constexpr (sizeof(E) == 8) {
        typedef unsigned v2u __attribute__((ext_vector_type(2)));
        v2u v;
        __builtin_memcpy(&v, &e, 8);
        __builtin_amdgcn_raw_buffer_store_b64(v, r, off, 0, AUX);
    } else {
        static_assert(sizeof(E) == 4, "element size");
        unsigned v;
        __builtin_memcpy(&v, &e, 4);
        __builtin_amdgcn_raw_buffer_store_b32(v, r, off, 0, AUX);
    }
}
#endif

// element (ubase)[lane] through the chosen path; NT = non-temporal
template <class E, bool NT = false>
__device__ __forceinline__ E fld(const E* ubase, unsigned lane) {
#if LYM_BUFFER_LOADS
    return buf_ld<E, NT ? 2 : 0>(ubase, lane);
#else
    if constexpr (NT && sizeof(E) >= 4) return ld_nt_raw(lane_at(ubase, lane));
    else return *lane_at(ubase, lane);
#endif
}
template <class E, bool NT = true>
__device__ __forceinline__ void fst(E* ubase, unsigned lane, E e) {
#if LYM_BUFFER_LOADS
    buf_st<E, NT ? 2 : 0>(ubase, lane, e);
#else
    if constexpr (NT && sizeof(E) >= 4) st_nt_raw(lane_at(ubase, lane), e);
    else *lane_at(ubase, lane) = e;
#endif
}

// ---------------------------------------------------------------- packed low-precision spinors
// FP32 and FP16 spinor fields store 16-byte vectors of consecutive
// spin-colour components: per parity [G][stride * nr] vectors, G = 6 (two
// float2) or 3 (four half2), lane element (site * nr + rhs).  A lane reads a
// site's spinor with G 16-byte loads instead of 12 8- or 4-byte ones; the
// bytes per site are those of the unpacked [12][stride * nr] layout, so the
// buffers keep their sizes.  FP64 spinors stay [12][stride * nr] double2.
struct alignas(16) V16 {
    unsigned w[4];
};
struct alignas(8) V8 {
    unsigned w[2];
};
template <int PR>
struct Packed {
    static constexpr int PER = 16 / (int)sizeof(typename Prec<PR>::S);  // components per vector
    static constexpr int G = 12 / PER;
};

// parity base pbase (element type S), group stride ks = stride * nr vectors, lane o
template <int PR, bool NT = false>
__device__ __forceinline__ void load_packed(const typename Prec<PR>::S* pbase, size_t ks, unsigned o,
                                            typename Prec<PR>::S e[12]) {
    constexpr int PER = Packed<PR>::PER;
    const V16* b = reinterpret_cast<const V16*>(pbase);
#pragma unroll
    for (int g = 0; g < Packed<PR>::G; ++g) {
        V16 v = fld<V16, NT>(b + g * ks, o);
        __builtin_memcpy(e + g * PER, &v, 16);
    }
}
template <int PR, bool NT = true>
__device__ __forceinline__ void store_packed(typename Prec<PR>::S* pbase, size_t ks, unsigned o,
                                             const typename Prec<PR>::S e[12]) {
    constexpr int PER = Packed<PR>::PER;
    V16* b = reinterpret_cast<V16*>(pbase);
#pragma unroll
    for (int g = 0; g < Packed<PR>::G; ++g) {
        V16 v;
        __builtin_memcpy(&v, e + g * PER, 16);
        fst<V16, NT>(b + g * ks, o, v);
    }
}

// ---------------------------------------------------------------- spinors
// One parity of a (multi-RHS) spinor field.
template <int PR>
struct SpinorRef {
    const typename Prec<PR>::S* base;
    const float* scale;  // PR == 2 only: [stride][NR]
};
template <int PR>
struct SpinorOut {
    typename Prec<PR>::S* base;
    float* scale;
};

template <int PR, bool NT = false>
__device__ __forceinline__ void load_spinor(SpinorRef<PR> f, int stride, int nr, int site, int r,
                                            cplx<typename Prec<PR>::T> s[12]) {
    const unsigned o = (unsigned)site * (unsigned)nr + (unsigned)r;
    const size_t ks = (size_t)stride * nr;
    if constexpr (PR == 0) {
#pragma unroll
        for (int k = 0; k < 12; ++k) s[k] = to_c<PR>(fld<typename Prec<PR>::S, NT>(f.base + k * ks, o));
    } else {
        typename Prec<PR>::S e[12];
        load_packed<PR, NT>(f.base, ks, o, e);
#pragma unroll
        for (int k = 0; k < 12; ++k) s[k] = to_c<PR>(e[k]);
    }
    if constexpr (PR == 2) {
        const float sc = fld<float>(f.scale, o);
#pragma unroll
        for (int k = 0; k < 12; ++k) s[k] = sc * s[k];
    }
}

// The same without the FP16 per-site scale: sc receives it (1 for FP64/FP32),
// so a caller can apply it to a product of fewer components (the stencil
// scales the two projected colour vectors after the link multiply: 12 real
// multiplies per hop instead of 24).
template <int PR>
__device__ __forceinline__ void load_spinor_unscaled(SpinorRef<PR> f, int stride, int nr, int site, int r,
                                                     cplx<typename Prec<PR>::T> s[12], typename Prec<PR>::T& sc) {
    const unsigned o = (unsigned)site * (unsigned)nr + (unsigned)r;
    const size_t ks = (size_t)stride * nr;
    if constexpr (PR == 0) {
#pragma unroll
        for (int k = 0; k < 12; ++k) s[k] = fld<typename Prec<PR>::S>(f.base + k * ks, o);
    } else {
        typename Prec<PR>::S e[12];
        load_packed<PR>(f.base, ks, o, e);
#pragma unroll
        for (int k = 0; k < 12; ++k) s[k] = to_c<PR>(e[k]);
    }
    if constexpr (PR == 2) sc = fld<float>(f.scale, o);
    else sc = (typename Prec<PR>::T)1;
}

// Reciprocal of a per-site FP16 scale m (the site's largest |component|).  A
// site whose m is below the smallest normal float is stored as zero: 1/m of a
// subnormal overflows to inf, and inf * 0 = NaN would poison the vector.  Such
// sites appear where a localised vector (a point source after a few hops)
// decays below 1e-38; they are negligible in FP32 arithmetic anyway.
__device__ __forceinline__ float fp16_site_inv(float& m) {
    if (!(m >= 1.17549435e-38f)) {
        m = 0.f;
        return 0.f;
    }
    return 1.0f / m;
}

template <int PR>
__device__ __forceinline__ void store_spinor(SpinorOut<PR> f, int stride, int nr, int site, int r,
                                             const cplx<typename Prec<PR>::T> v[12]) {
    const unsigned o = (unsigned)site * (unsigned)nr + (unsigned)r;
    const size_t ks = (size_t)stride * nr;
    if constexpr (PR == 2) {
        float m = 0.f;
#pragma unroll
        for (int k = 0; k < 12; ++k) m = fmaxf(m, fmaxf(fabsf(v[k].re), fabsf(v[k].im)));
        const float inv = fp16_site_inv(m);
        typename Prec<PR>::S e[12];
#pragma unroll
        for (int k = 0; k < 12; ++k) e[k] = from_c<PR>(inv * v[k]);
        store_packed<PR>(f.base, ks, o, e);
        fst<float, false>(f.scale, o, m);
    } else if constexpr (PR == 1) {
        typename Prec<PR>::S e[12];
#pragma unroll
        for (int k = 0; k < 12; ++k) e[k] = from_c<PR>(v[k]);
        store_packed<PR>(f.base, ks, o, e);
    } else {
#pragma unroll
        for (int k = 0; k < 12; ++k) fst(f.base + k * ks, o, v[k]);
    }
}

// ---------------------------------------------------------------- link loaders
// LK: 0 complex 18-real, 1 complex 12-real (row 2 = conj(row0 x row1)),
//     2 real 9 (adjoint), 3 real 6 (row 2 = row0 x row1)
template <int LK, class T>
struct LinkMat;
template <class T>
struct LinkMat<0, T> { cplx<T> m[9]; };
template <class T>
struct LinkMat<1, T> { cplx<T> m[9]; };
template <class T>
struct LinkMat<2, T> { T m[9]; };
template <class T>
struct LinkMat<3, T> { T m[9]; };

// ---------------------------------------------------------------- one chirality of a low-precision spinor
// Components 6h .. 6h+5 (spins 2h, 2h+1) of a packed FP32/FP16 spinor, for
// the two-threads-per-site stencil; h must be wave-uniform.  FP32 vectors hold
// components (2g, 2g+1): a chirality is three whole vectors.  FP16 vectors
// hold (4g .. 4g+3): chirality 0 is vector 0 and the low half of vector 1,
// chirality 1 the high half of vector 1 and vector 2.  No scale applied.
template <int PR, bool NT = false>
__device__ __forceinline__ void load_half_lp(const typename Prec<PR>::S* pbase, size_t ks, unsigned o, int h,
                                             typename Prec<PR>::S e[6]) {
    static_assert(PR >= 1, "low precision only");
    const V16* b = reinterpret_cast<const V16*>(pbase);
    if constexpr (PR == 1) {
#pragma unroll
        for (int g = 0; g < 3; ++g) {
            V16 v = fld<V16, NT>(b + (3 * h + g) * ks, o);
            __builtin_memcpy(e + 2 * g, &v, 16);
        }
    } else {
        const V8* b8 = reinterpret_cast<const V8*>(b + ks);  // vector 1 as 8-byte halves
        if (h == 0) {
            V16 v = fld<V16, NT>(b, o);
            V8 w = fld<V8, NT>(b8, 2 * o);
            __builtin_memcpy(e, &v, 16);
            __builtin_memcpy(e + 4, &w, 8);
        } else {
            V8 w = fld<V8, NT>(b8, 2 * o + 1);
            V16 v = fld<V16, NT>(b + 2 * ks, o);
            __builtin_memcpy(e, &w, 8);
            __builtin_memcpy(e + 2, &v, 16);
        }
    }
}
template <int PR>
__device__ __forceinline__ void store_half_lp(typename Prec<PR>::S* pbase, size_t ks, unsigned o, int h,
                                              const typename Prec<PR>::S e[6]) {
    static_assert(PR >= 1, "low precision only");
    V16* b = reinterpret_cast<V16*>(pbase);
    if constexpr (PR == 1) {
#pragma unroll
        for (int g = 0; g < 3; ++g) {
            V16 v;
            __builtin_memcpy(&v, e + 2 * g, 16);
            fst<V16>(b + (3 * h + g) * ks, o, v);
        }
    } else {
        V8* b8 = reinterpret_cast<V8*>(b + ks);
        V16 v;
        V8 w;
        if (h == 0) {
            __builtin_memcpy(&v, e, 16);
            __builtin_memcpy(&w, e + 4, 8);
            fst<V16>(b, o, v);
            fst<V8>(b8, 2 * o, w);
        } else {
            __builtin_memcpy(&w, e, 8);
            __builtin_memcpy(&v, e + 2, 16);
            fst<V8>(b8, 2 * o + 1, w);
            fst<V16>(b + 2 * ks, o, v);
        }
    }
}

// ---------------------------------------------------------------- packed low-precision links
// FP32/FP16 links of kinds 1 and 3 hold rows 0 and 1 only (6 entries; row 2 is
// rebuilt: conj(row0 x row1) for SU(3), row0 x row1 for the real adjoint
// links), packed into 16-, 8- and 4-byte vectors.  Group g of the field is an
// array [2 parity][4 mu][stride] of vectors of its size; the groups follow
// one another.  Bytes per link: FP32 complex 48 (16+16+16), FP16 complex 24
// (16+8), FP32 real 24 (16+8), FP16 real 12 (8+4).
template <int PR, bool REAL>
struct PackedLink {
    static_assert(PR >= 1, "packed links are low-precision only");
    typedef typename std::conditional<REAL, typename Prec<PR>::R, typename Prec<PR>::S>::type E;
    static constexpr int BYTES = 6 * (int)sizeof(E);
    static constexpr int N16 = BYTES / 16;
    static constexpr bool H8 = (BYTES - 16 * N16) >= 8;
    static constexpr bool H4 = (BYTES - 16 * N16) % 8 == 4;
};

template <int PR, bool REAL>
__device__ __forceinline__ void load_link_rows(const void* base, int stride, int p, int mu, int idx,
                                               typename PackedLink<PR, REAL>::E e[6]) {
    typedef PackedLink<PR, REAL> P;
    const char* gb = reinterpret_cast<const char*>(base);
    const size_t plane = (size_t)(p * 4 + mu) * stride;
    unsigned char buf[P::BYTES];
    int off = 0;
#pragma unroll
    for (int g = 0; g < P::N16; ++g) {
        V16 v = fld<V16, LYM_NT_LINKS != 0>(reinterpret_cast<const V16*>(gb) + plane, (unsigned)idx);
        __builtin_memcpy(buf + off, &v, 16);
        off += 16;
        gb += (size_t)8 * stride * 16;
    }
    if constexpr (P::H8) {
        V8 v = fld<V8, LYM_NT_LINKS != 0>(reinterpret_cast<const V8*>(gb) + plane, (unsigned)idx);
        __builtin_memcpy(buf + off, &v, 8);
        off += 8;
        gb += (size_t)8 * stride * 8;
    }
    if constexpr (P::H4) {
        unsigned v = fld<unsigned, LYM_NT_LINKS != 0>(reinterpret_cast<const unsigned*>(gb) + plane, (unsigned)idx);
        __builtin_memcpy(buf + off, &v, 4);
    }
    __builtin_memcpy(e, buf, P::BYTES);
}

template <int PR, int LK>
__device__ __forceinline__ void load_link(const void* base, int stride, int p, int mu, int idx,
                                          LinkMat<LK, typename Prec<PR>::T>& L) {
    typedef typename Prec<PR>::T T;
    if constexpr (PR >= 1 && (LK == 1 || LK == 3)) {
        constexpr bool REAL = LK == 3;
        typename PackedLink<PR, REAL>::E e[6];
        load_link_rows<PR, REAL>(base, stride, p, mu, idx, e);
#pragma unroll
        for (int k = 0; k < 6; ++k) {
            if constexpr (REAL) L.m[k] = (T)e[k];
            else L.m[k] = to_c<PR>(e[k]);
        }
        if constexpr (REAL) {
            L.m[6] = L.m[1] * L.m[5] - L.m[2] * L.m[4];
            L.m[7] = L.m[2] * L.m[3] - L.m[0] * L.m[5];
            L.m[8] = L.m[0] * L.m[4] - L.m[1] * L.m[3];
        } else {
            L.m[6] = conj(L.m[1] * L.m[5] - L.m[2] * L.m[4]);
            L.m[7] = conj(L.m[2] * L.m[3] - L.m[0] * L.m[5]);
            L.m[8] = conj(L.m[0] * L.m[4] - L.m[1] * L.m[3]);
        }
    } else if constexpr (LK == 0 || LK == 1) {
        constexpr int NC = (LK == 0) ? 9 : 6;
        const typename Prec<PR>::S* b =
            reinterpret_cast<const typename Prec<PR>::S*>(base) + (size_t)((p * 4 + mu) * NC) * stride;
#pragma unroll
        for (int k = 0; k < NC; ++k) {
            L.m[k] = to_c<PR>(fld<typename Prec<PR>::S, LYM_NT_LINKS != 0>(b + (size_t)k * stride, (unsigned)idx));
        }
        if constexpr (LK == 1) {
            L.m[6] = conj(L.m[1] * L.m[5] - L.m[2] * L.m[4]);
            L.m[7] = conj(L.m[2] * L.m[3] - L.m[0] * L.m[5]);
            L.m[8] = conj(L.m[0] * L.m[4] - L.m[1] * L.m[3]);
        }
    } else {
        constexpr int NC = (LK == 2) ? 9 : 6;
        const typename Prec<PR>::R* b =
            reinterpret_cast<const typename Prec<PR>::R*>(base) + (size_t)((p * 4 + mu) * NC) * stride;
#pragma unroll
        for (int k = 0; k < NC; ++k) L.m[k] = (T)fld<typename Prec<PR>::R, LYM_NT_LINKS != 0>(b + (size_t)k * stride, (unsigned)idx);
        if constexpr (LK == 3) {
            L.m[6] = L.m[1] * L.m[5] - L.m[2] * L.m[4];
            L.m[7] = L.m[2] * L.m[3] - L.m[0] * L.m[5];
            L.m[8] = L.m[0] * L.m[4] - L.m[1] * L.m[3];
        }
    }
}

template <int LK, class T>
__device__ __forceinline__ void matvec(const LinkMat<LK, T>& L, const cplx<T> h[3], cplx<T> o[3]) {
#pragma unroll
    for (int i = 0; i < 3; ++i) {
        if constexpr (LK <= 1) {
            cplx<T> s = L.m[3 * i] * h[0];
            cfma(s, L.m[3 * i + 1], h[1]);
            cfma(s, L.m[3 * i + 2], h[2]);
            o[i] = s;
        } else {
            o[i] = mk(L.m[3 * i] * h[0].re + L.m[3 * i + 1] * h[1].re + L.m[3 * i + 2] * h[2].re,
                      L.m[3 * i] * h[0].im + L.m[3 * i + 1] * h[1].im + L.m[3 * i + 2] * h[2].im);
        }
    }
}

template <int LK, class T>
__device__ __forceinline__ void matvec_dag(const LinkMat<LK, T>& L, const cplx<T> h[3], cplx<T> o[3]) {
#pragma unroll
    for (int i = 0; i < 3; ++i) {
        if constexpr (LK <= 1) {
            cplx<T> s = cmulc(L.m[i], h[0]);
            cfmac(s, L.m[3 + i], h[1]);
            cfmac(s, L.m[6 + i], h[2]);
            o[i] = s;
        } else {
            o[i] = mk(L.m[i] * h[0].re + L.m[3 + i] * h[1].re + L.m[6 + i] * h[2].re,
                      L.m[i] * h[0].im + L.m[3 + i] * h[1].im + L.m[6 + i] * h[2].im);
        }
    }
}

// FP64 single-RHS spinor load (layout.hip, clover.hip)
__device__ __forceinline__ void load_spinor(const cplx<double>* f, int stride, int idx, cplx<double> s[12]) {
#pragma unroll
    for (int k = 0; k < 12; ++k) s[k] = f[(size_t)k * stride + idx];
}

// ---------------------------------------------------------------- clover blocks
// o = A v for one packed hermitian 6x6 block; A points at the block's first
// component of the parity, sc is the per-site scale (PR == 2) or 1.
template <int PR>
__device__ __forceinline__ void herm6(const typename Prec<PR>::S* A, int stride, int idx,
                                      typename Prec<PR>::T sc, const cplx<typename Prec<PR>::T> v[6],
                                      cplx<typename Prec<PR>::T> o[6]) {
    typedef typename Prec<PR>::T T;
    T d[6];
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        cplx<T> dd = to_c<PR>(fld<typename Prec<PR>::S, true>(A + (size_t)k * stride, (unsigned)idx));
        d[2 * k] = sc * dd.re;
        d[2 * k + 1] = sc * dd.im;
    }
#pragma unroll
    for (int i = 0; i < 6; ++i) o[i] = d[i] * v[i];
    int k = 3;
#pragma unroll
    for (int i = 1; i < 6; ++i) {
#pragma unroll
        for (int j = 0; j < i; ++j) {
            cplx<T> a = sc * to_c<PR>(fld<typename Prec<PR>::S, true>(A + (size_t)k * stride, (unsigned)idx));
            cfma(o[i], a, v[j]);
            cfmac(o[j], a, v[i]);
            ++k;
        }
    }
}

// The same product from the block's 18 entries already in registers (scaled).
template <class T>
__device__ __forceinline__ void herm6_mul(const cplx<T> a[18], const cplx<T> v[6], cplx<T> o[6]) {
    T d[6];
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        d[2 * k] = a[k].re;
        d[2 * k + 1] = a[k].im;
    }
#pragma unroll
    for (int i = 0; i < 6; ++i) o[i] = d[i] * v[i];
    int k = 3;
#pragma unroll
    for (int i = 1; i < 6; ++i) {
#pragma unroll
        for (int j = 0; j < i; ++j) {
            cfma(o[i], a[k], v[j]);
            cfmac(o[j], a[k], v[i]);
            ++k;
        }
    }
}

// Low-precision clover blocks: the 18 entries of each hermitian 6x6 block
// (3 diagonal pairs, 15 off-diagonal) packed into 16- and 8-byte vectors.
// Group g is an array [nplanes][stride] of vectors (plane = 2 parity + block
// for the blocks of D, block for R^-1), the groups one after the other.
// FP32: 144 B = 9 x 16; FP16: 72 B = 4 x 16 + 8.
template <int PR>
struct PackedBlock {
    static_assert(PR >= 1, "packed blocks are low-precision only");
    typedef typename Prec<PR>::S E;
    static constexpr int BYTES = 18 * (int)sizeof(E);
    static constexpr int N16 = BYTES / 16;
    static constexpr bool H8 = BYTES - 16 * N16 >= 8;
};

// o = sc * A v for the block at plane_off = plane * stride of a packed field
// whose group arrays hold ngroup vectors each
template <int PR>
__device__ __forceinline__ void herm6_packed(const void* base, size_t ngroup, size_t plane_off, int idx,
                                             typename Prec<PR>::T sc, const cplx<typename Prec<PR>::T> v[6],
                                             cplx<typename Prec<PR>::T> o[6]) {
    typedef PackedBlock<PR> P;
    typedef typename Prec<PR>::T T;
    const char* gb = reinterpret_cast<const char*>(base);
    unsigned char buf[P::BYTES];
    int off = 0;
#pragma unroll
    for (int g = 0; g < P::N16; ++g) {
        V16 w = fld<V16, true>(reinterpret_cast<const V16*>(gb) + plane_off, (unsigned)idx);
        __builtin_memcpy(buf + off, &w, 16);
        off += 16;
        gb += ngroup * 16;
    }
    if constexpr (P::H8) {
        V8 w = fld<V8, true>(reinterpret_cast<const V8*>(gb) + plane_off, (unsigned)idx);
        __builtin_memcpy(buf + off, &w, 8);
    }
    typename P::E e[18];
    __builtin_memcpy(e, buf, P::BYTES);
    cplx<T> a[18];
#pragma unroll
    for (int k = 0; k < 18; ++k) a[k] = to_c<PR>(e[k]);
    herm6_mul<T>(a, v, o);
    if constexpr (PR == 2) {
        // the block's scale on the 6 outputs instead of its 18 entries
#pragma unroll
        for (int i = 0; i < 6; ++i) o[i] = sc * o[i];
    }
}

// FP64 form used by the clover builder and the EO driver
__device__ __forceinline__ void herm6(const cplx<double>* A, int stride, int idx, const cplx<double> v[6],
                                      cplx<double> o[6]) {
    herm6<0>(A, stride, idx, 1.0, v, o);
}

}  // namespace lym
