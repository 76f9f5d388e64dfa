// lym_common.h -- shared device/host definitions of the MI355X fermion-solve library.
//
// Device data layout (HBM):
//   Sites are split by parity p = (x+y+z+t) & 1.  Within a parity the index is
//   eo = lex >> 1 with lex = ((x*Ly + y)*Lz + z)*Lt + t, which is a bijection
//   onto [0, V/2) when Lt is even (every t-row pairs one even with one odd site).
//   A field component is a contiguous array of `stride` elements (stride = V/2
//   rounded up to 64), so lane i of a wavefront touches element i: every load
//   and store of the stencil and of BLAS-1 is a fully coalesced 8- or 16-byte
//   stream.
//     spinor (full) : double2 [2 parity][12 = spin*3 + colour][stride]
//     spinor (half) : double2 [12][stride]
//     links 18-real : double2 [2][4 mu][9 = 3*row + col][stride]
//     links 12-real : double2 [2][4][6 = rows 0,1][stride], row 2 = conj(row0 x row1)
//     adjoint links : double  [2][4][9 or 6][stride] (real SO(3), row 2 = row0 x row1)
//     clover blocks : double2 [2][36][stride]: two hermitian 6x6 blocks (spins 0,1
//                     and 2,3), each packed as 3 double2 of diagonal pairs then 15
//                     complex lower-triangle entries, k = i(i-1)/2 + j for i > j.
//   Links are stored without the fermion boundary sign; the kernels apply the
//   sign bc_t to the time hops that cross t = Lt-1 <-> 0, which keeps the
//   12-real reconstruction exact for the sign-flipped links.
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>

namespace lym {

// Device site order.  0: t fastest (the reference's lexicographic order; the
// even-odd split needs Lt even).  1: t slowest, x fastest (needs Lx even): the
// contiguous site range each XCD works on is then a slab of whole t-slices,
// the longest extent of the usual T > L lattices, so a smaller share of each
// slab's neighbours lies on other XCDs.
#ifndef LYM_T_SLOW
#define LYM_T_SLOW 0
#endif

// Two site classes (tsh = 1): the even-odd layout, eo = lex >> 1, parity
// (x + y + z + t) & 1, every neighbour in the other parity -- lattices whose
// four extents are all even.  One class (tsh = 0): a lattice with an odd
// extent has no such two-colouring under periodic wrapping, so all sites form
// "parity 0" in lexicographic order (eo = lex) and a hop stays in class 0.
// The same index arithmetic covers both: parity & tsh, lex >> tsh, p ^ tsh.
struct Geom {
    int L[4];    // Lx, Ly, Lz, Lt
    int Lth;     // Lt >> tsh
    int Lxh;     // Lx >> tsh
    int Vh;      // sites per class
    int stride;  // component stride (Vh rounded up to 64)
    int bc_t;    // +1 periodic, -1 antiperiodic in time
    int tsh;     // 1: even-odd layout; 0: one class (a lattice with an odd extent)
};

// ---------------------------------------------------------------- complex
template <class T>
struct alignas(2 * sizeof(T)) cplx {
    T re, im;
};

template <class T>
__host__ __device__ __forceinline__ cplx<T> mk(T r, T i) { return {r, i}; }
template <class T>
__host__ __device__ __forceinline__ cplx<T> operator+(cplx<T> a, cplx<T> b) { return {a.re + b.re, a.im + b.im}; }
template <class T>
__host__ __device__ __forceinline__ cplx<T> operator-(cplx<T> a, cplx<T> b) { return {a.re - b.re, a.im - b.im}; }
template <class T>
__host__ __device__ __forceinline__ cplx<T> operator-(cplx<T> a) { return {-a.re, -a.im}; }
template <class T>
__host__ __device__ __forceinline__ cplx<T> operator*(cplx<T> a, cplx<T> b) {
    return {a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re};
}
template <class T>
__host__ __device__ __forceinline__ cplx<T> operator*(T s, cplx<T> a) { return {s * a.re, s * a.im}; }
// conj(a) * b
template <class T>
__host__ __device__ __forceinline__ cplx<T> cmulc(cplx<T> a, cplx<T> b) {
    return {a.re * b.re + a.im * b.im, a.re * b.im - a.im * b.re};
}
template <class T>
__host__ __device__ __forceinline__ cplx<T> conj(cplx<T> a) { return {a.re, -a.im}; }
// one fused multiply-add (a single rounding)
__host__ __device__ __forceinline__ float fmadd(float a, float b, float c) { return __builtin_fmaf(a, b, c); }
__host__ __device__ __forceinline__ double fmadd(double a, double b, double c) { return __builtin_fma(a, b, c); }
// a += b * c, as two fused multiply-adds per real component: four
// instructions per complex product instead of six (a multiply, a fused
// multiply-add and an add per component), and one rounding fewer
template <class T>
__host__ __device__ __forceinline__ void cfma(cplx<T>& a, cplx<T> b, cplx<T> c) {
    a.re = fmadd(b.re, c.re, a.re);
    a.re = fmadd(-b.im, c.im, a.re);
    a.im = fmadd(b.re, c.im, a.im);
    a.im = fmadd(b.im, c.re, a.im);
}
// a += conj(b) * c, likewise
template <class T>
__host__ __device__ __forceinline__ void cfmac(cplx<T>& a, cplx<T> b, cplx<T> c) {
    a.re = fmadd(b.re, c.re, a.re);
    a.re = fmadd(b.im, c.im, a.re);
    a.im = fmadd(b.re, c.im, a.im);
    a.im = fmadd(-b.im, c.re, a.im);
}

// multiply by the phase i^PH (PH = 0: 1, 1: i, 2: -1, 3: -i)
template <int PH, class T>
__host__ __device__ __forceinline__ cplx<T> mulph(cplx<T> z) {
    if constexpr ((PH & 3) == 0) return z;
    else if constexpr ((PH & 3) == 1) return {-z.im, z.re};
    else if constexpr ((PH & 3) == 2) return {-z.re, -z.im};
    else return {z.im, -z.re};
}

// ---------------------------------------------------------------- geometry
#if LYM_T_SLOW
// device order t slowest, x fastest: dev = ((t*Lz + z)*Ly + y)*Lx + x, eo = dev >> 1
__device__ __forceinline__ void eo_coords(const Geom& g, int p, int idx, int c[4]) {
    int xh = idx % g.Lxh;
    int r = idx / g.Lxh;
    c[1] = r % g.L[1];
    r /= g.L[1];
    c[2] = r % g.L[2];
    c[3] = r / g.L[2];
    c[0] = (xh << g.tsh) + ((c[1] + c[2] + c[3] + p) & g.tsh);
}

__device__ __forceinline__ int eo_index(const Geom& g, int x, int y, int z, int t) {
    return ((t * g.L[2] + z) * g.L[1] + y) * g.Lxh + (x >> g.tsh);
}
#else
// device order t fastest (the reference's lexicographic order), eo = lex >> 1
__device__ __forceinline__ void eo_coords(const Geom& g, int p, int idx, int c[4]) {
    int th = idx % g.Lth;
    int r = idx / g.Lth;
    c[2] = r % g.L[2];
    r /= g.L[2];
    c[1] = r % g.L[1];
    c[0] = r / g.L[1];
    c[3] = (th << g.tsh) + ((c[0] + c[1] + c[2] + p) & g.tsh);
}

__device__ __forceinline__ int eo_index(const Geom& g, int x, int y, int z, int t) {
    return ((x * g.L[1] + y) * g.L[2] + z) * g.Lth + (t >> g.tsh);
}

// class of the site at coordinates c
__device__ __forceinline__ int site_parity(const Geom& g, const int c[4]) { return (c[0] + c[1] + c[2] + c[3]) & g.tsh; }
#endif

// XCD-aware block order: MI355X deals workgroups round-robin over its 8 XCDs;
// remapping gives each XCD a contiguous range of sites, so neighbour spinors
// and links fetched by one XCD are reused from that XCD's own L2.
__device__ __forceinline__ int xcd_block(int b, int nb) {
    const int nx = 8;
    int q = nb / nx, r = nb % nx;
    int xcd = b % nx, k = b / nx;
    // XCDs 0..r-1 own q+1 blocks, the rest q
    return (xcd < r ? xcd * (q + 1) : r * (q + 1) + (xcd - r) * q) + k;
}

}  // namespace lym

// ---------------------------------------------------------------- host objects
struct lym_ctx {
    lym::Geom g;
    int dims[4];
    long V;
    int rep;  // 0 SU(3) fundamental, 1 SU(2) adjoint
    int device;
    hipStream_t stream;
    int owns_stream;        // 0: the stream belongs to the caller (serving workers share one per thread)
    void* staging;          // device staging for host-layout transfers
    size_t staging_bytes;
    double* red_partials;   // reduction partials
    double* red_result;     // device reduction results
    double* host_pinned;    // pinned host readback
    hipEvent_t ev_start, ev_stop;
};

struct lym_gauge {
    lym_ctx* ctx;
    int compression;   // 18 or 12 (9 or 6 for real adjoint links)
    int real_links;    // 1 for SU(2) adjoint
    void* U;           // links, see layout above
    double2* B;        // kappa-independent clover blocks (hermitian, packed), built lazily
    int rows_rebuild;  // 1: row 2 of every link is the one rebuilt from rows 0 and 1 (SU(3) /
                       // SO(3) links), so the low-precision copies keep rows 0 and 1 only
};

struct lym_spinor {
    lym_ctx* ctx;
    double2* d;        // [2][12][stride]
};

struct lym_op {
    lym_ctx* ctx;
    uint64_t uid;      // unique for the process's lifetime (keys of captured solver iterations)
    int kind;          // LYM_OP_*
    unsigned power;
    double kappa, csw;
    double* dev_kappa;  // kappa in device memory, read by the stencil (captured launches stay valid across kappa changes)
    int gamma5;
    const lym_gauge* gauge;
    double2* A;        // [2][36][stride] site-diagonal blocks of D (1 + kappa csw gamma5 B)
    double2* Ainv;     // [36][stride] inverse blocks on even sites
    int diag_dirty;
    double2* tmp;      // full-spinor scratch for squared operators
    double2* tmp2;     // half-spinor scratch for the Schur operator
    // low-precision copies for mixed-precision solvers (sloppy_prec 1 FP32, 2 FP16; 0 none)
    int sloppy_prec;
    int sloppy_alloc;  // precision the low-precision buffers are allocated for (kept across kappa changes)
    int sloppy_lk;     // link kind of the low-precision links
    void* Usf;         // all rows of the links in low precision (multi-RHS applies), when Us holds rows 0, 1
    void* Us;          // links
    void* As;          // [2][36][stride] blocks
    float* As_sc;      // [2][stride]
    void* Ainvs;       // [36][stride]
    float* Ainvs_sc;   // [stride]
    // spinor fields of lym_op_multiply_sloppy (in, out, Q^2 intermediate, even
    // Schur intermediate), sized for FP32 storage; allocated on first use
    void* sp[4];
    float* sp_sc[4];   // [2][stride] FP16 scales
};
