// dslash.hip -- Wilson / clover hopping stencil for gfx950, FP64.
//
// One thread per lattice site of one parity.  For every direction mu the
// neighbour spinors x+mu and x-mu (other parity) are loaded as 12 coalesced
// complex streams, spin-projected to two colour vectors with (1 -+ gamma_mu),
// multiplied by U_mu(x) or U_mu(x-mu)^dagger and accumulated with the
// gamma_mu reconstruction, as in the reference's per-site loops
// (source/dirac_operators/DiracWilsonOperator.cpp:21-192,
//  BasicDiracWilsonOperator.cpp:16-93, EvenOddImprovedDiracWilsonOperator.cpp:204-281).
//
// The epilogue is fused: v = diag(psi_x) + hop_sign * (-kappa H psi),
// optionally v = Ainv v, gamma5, and v += alpha v2.  diag is nothing, the
// identity (Wilson) or the packed clover block A = 1 + kappa csw gamma5 B
// (ImprovedDiracWilsonOperator.cpp:200-243), so a single kernel covers
//   full-lattice Q / D and multiplyAdd        diag = A or 1, hop_sign = +1
//   EO  tmp_e = R_ee^-1 D_eo psi_o            diag = none,   INV
//   EO  out_o = R_oo psi_o - D_oe tmp_e       diag = A,      hop_sign = -1
//   EO driver source b_o - D_oe R^-1 b_e and even reconstruction
//   R_ee^-1 (b_e - D_eo x_o)                  diag = 1, hop_sign = -1 [, INV]
// (PreconditionedBiCGStab.cpp:56-86).
#include "lym_device.h"

// Launch shape of the stencil (swept on the device; defaults are the measured best)
#ifndef LYM_DSLASH_THREADS
#define LYM_DSLASH_THREADS 256
#endif
// both parities of a block range run on the same XCD (links and spinors of a
// region are then fetched into one L2 for the even and the odd sites)
#ifndef LYM_GRID_PER_CU
#define LYM_GRID_PER_CU 0
#endif
#ifdef LYM_WAVES_PER_EU
#define LYM_OCC __attribute__((amdgpu_waves_per_eu(LYM_WAVES_PER_EU, LYM_WAVES_PER_EU)))
#else
#define LYM_OCC
#endif

namespace lym {

template <int MU>
struct Dir;
// (1 - gamma_mu) keeps h_a = psi_a + c_a psi_{P_a} for a = 0, 1; c = i^C.
template <>
struct Dir<0> { static constexpr int P0 = 3, C0 = 1, P1 = 2, C1 = 1; };
template <>
struct Dir<1> { static constexpr int P0 = 3, C0 = 0, P1 = 2, C1 = 2; };
template <>
struct Dir<2> { static constexpr int P0 = 2, C0 = 1, P1 = 3, C1 = 3; };
template <>
struct Dir<3> { static constexpr int P0 = 2, C0 = 2, P1 = 3, C1 = 2; };

// forward:  h_a = psi_a + c_a psi_P, acc_a += U h_a, acc_P += conj(c_a) U h_a
// backward: h_a = psi_a - c_a psi_P, acc_a += U^+ h_a, acc_P -= conj(c_a) U^+ h_a
template <int MU, bool FWD, int LK>
__device__ __forceinline__ void hop_dir(cplx<double> acc[12], const cplx<double> psi[12], const LinkMat<LK>& L,
                                        bool flip) {
    using D = Dir<MU>;
    constexpr int PH0 = FWD ? D::C0 : D::C0 + 2;
    constexpr int PH1 = FWD ? D::C1 : D::C1 + 2;
    constexpr int Q0 = FWD ? 4 - D::C0 : 6 - D::C0;
    constexpr int Q1 = FWD ? 4 - D::C1 : 6 - D::C1;
    cplx<double> h0[3], h1[3], r0[3], r1[3];
#pragma unroll
    for (int c = 0; c < 3; ++c) {
        h0[c] = psi[c] + mulph<PH0>(psi[3 * D::P0 + c]);
        h1[c] = psi[3 + c] + mulph<PH1>(psi[3 * D::P1 + c]);
        if (flip) {
            h0[c] = -h0[c];
            h1[c] = -h1[c];
        }
    }
    if constexpr (FWD) {
        matvec<LK>(L, h0, r0);
        matvec<LK>(L, h1, r1);
    } else {
        matvec_dag<LK>(L, h0, r0);
        matvec_dag<LK>(L, h1, r1);
    }
#pragma unroll
    for (int c = 0; c < 3; ++c) {
        acc[c] = acc[c] + r0[c];
        acc[3 + c] = acc[3 + c] + r1[c];
        acc[3 * D::P0 + c] = acc[3 * D::P0 + c] + mulph<Q0>(r0[c]);
        acc[3 * D::P1 + c] = acc[3 * D::P1 + c] + mulph<Q1>(r1[c]);
    }
}

// Empty asm that reads and writes the accumulator and clobbers memory: the
// hop of one direction must finish before the next direction's loads are
// issued, which bounds the live registers to one hop (without it every load of
// all eight hops is hoisted ahead of the arithmetic and the kernel spills).
__device__ __forceinline__ void pin(cplx<double> acc[12]) {
#pragma unroll
    for (int k = 0; k < 12; ++k) asm volatile("" : "+v"(acc[k].re), "+v"(acc[k].im));
    asm volatile("" ::: "memory");
}

// H psi at site (p, idx) with coordinates c, reading the other parity `src`.
template <int LK>
__device__ __forceinline__ void hopping(const Geom& g, const void* links, const cplx<double>* src, int p, int idx,
                                        const int c[4], cplx<double> acc[12]) {
    const int S = g.stride;
    const int q = p ^ 1;
#pragma unroll
    for (int k = 0; k < 12; ++k) acc[k] = mk(0.0, 0.0);
    cplx<double> psi[12];
    LinkMat<LK> L;
    // mu = 0 (x)
    {
        int xu = c[0] + 1 == g.L[0] ? 0 : c[0] + 1;
        int xd = c[0] == 0 ? g.L[0] - 1 : c[0] - 1;
        int iu = eo_index(g, xu, c[1], c[2], c[3]);
        int id = eo_index(g, xd, c[1], c[2], c[3]);
        load_spinor(src, S, iu, psi);
        load_link<LK>(links, S, p, 0, idx, L);
        hop_dir<0, true, LK>(acc, psi, L, false);
        pin(acc);
        load_spinor(src, S, id, psi);
        load_link<LK>(links, S, q, 0, id, L);
        hop_dir<0, false, LK>(acc, psi, L, false);
        pin(acc);
    }
    {
        int yu = c[1] + 1 == g.L[1] ? 0 : c[1] + 1;
        int yd = c[1] == 0 ? g.L[1] - 1 : c[1] - 1;
        int iu = eo_index(g, c[0], yu, c[2], c[3]);
        int id = eo_index(g, c[0], yd, c[2], c[3]);
        load_spinor(src, S, iu, psi);
        load_link<LK>(links, S, p, 1, idx, L);
        hop_dir<1, true, LK>(acc, psi, L, false);
        pin(acc);
        load_spinor(src, S, id, psi);
        load_link<LK>(links, S, q, 1, id, L);
        hop_dir<1, false, LK>(acc, psi, L, false);
        pin(acc);
    }
    {
        int zu = c[2] + 1 == g.L[2] ? 0 : c[2] + 1;
        int zd = c[2] == 0 ? g.L[2] - 1 : c[2] - 1;
        int iu = eo_index(g, c[0], c[1], zu, c[3]);
        int id = eo_index(g, c[0], c[1], zd, c[3]);
        load_spinor(src, S, iu, psi);
        load_link<LK>(links, S, p, 2, idx, L);
        hop_dir<2, true, LK>(acc, psi, L, false);
        pin(acc);
        load_spinor(src, S, id, psi);
        load_link<LK>(links, S, q, 2, id, L);
        hop_dir<2, false, LK>(acc, psi, L, false);
        pin(acc);
    }
    {
        const bool last = c[3] + 1 == g.L[3];
        const bool first = c[3] == 0;
        int tu = last ? 0 : c[3] + 1;
        int td = first ? g.L[3] - 1 : c[3] - 1;
        int iu = eo_index(g, c[0], c[1], c[2], tu);
        int id = eo_index(g, c[0], c[1], c[2], td);
        const bool anti = g.bc_t < 0;
        load_spinor(src, S, iu, psi);
        load_link<LK>(links, S, p, 3, idx, L);
        hop_dir<3, true, LK>(acc, psi, L, anti && last);
        pin(acc);
        load_spinor(src, S, id, psi);
        load_link<LK>(links, S, q, 3, id, L);
        hop_dir<3, false, LK>(acc, psi, L, anti && first);
        pin(acc);
    }
}

template <int LK, int DIAG, bool INV>
__device__ __forceinline__ void dslash_site(const DslashParams& a, int p, int idx) {
    const int S = a.g.stride;
    int c[4];
    eo_coords(a.g, p, idx, c);
    cplx<double> acc[12];
    hopping<LK>(a.g, a.links, a.in[p ^ 1], p, idx, c, acc);
    const double hk = -a.kappa * a.hop_sign;
    cplx<double> v[12];
    if constexpr (DIAG == 0) {
#pragma unroll
        for (int k = 0; k < 12; ++k) v[k] = hk * acc[k];
    } else {
        cplx<double> s[12];
        load_spinor(a.diag_in[p], S, idx, s);
        if constexpr (DIAG == 2) {
            cplx<double> t[12];
            const cplx<double>* Ab = a.A + (size_t)p * 36 * S;
            herm6(Ab, S, idx, s, t);
            herm6(Ab + (size_t)18 * S, S, idx, s + 6, t + 6);
#pragma unroll
            for (int k = 0; k < 12; ++k) v[k] = t[k] + hk * acc[k];
        } else {
#pragma unroll
            for (int k = 0; k < 12; ++k) v[k] = s[k] + hk * acc[k];
        }
    }
    if constexpr (INV) {
        cplx<double> t[12];
        herm6(a.Ainv, S, idx, v, t);
        herm6(a.Ainv + (size_t)18 * S, S, idx, v + 6, t + 6);
#pragma unroll
        for (int k = 0; k < 12; ++k) v[k] = t[k];
    }
    if (a.gamma5) {
#pragma unroll
        for (int k = 6; k < 12; ++k) v[k] = -v[k];
    }
    if (a.v2[p]) {
        const cplx<double>* w = a.v2[p];
#pragma unroll
        for (int k = 0; k < 12; ++k) v[k] = v[k] + a.alpha * ld_nt(w + (size_t)k * S + idx);
    }
    cplx<double>* o = a.out[p];
#pragma unroll
    for (int k = 0; k < 12; ++k) st_nt(o + (size_t)k * S + idx, v[k]);
}

// Grid: gridDim.x workgroups dealt round-robin over the 8 XCDs.  XCD x owns a
// contiguous range of logical site blocks (x-major lattice order, the two
// parities of a block range interleaved) and its workgroups sweep that range
// in order, so at any time an XCD works on one compact slab and the
// neighbour spinors and links of that slab are reused from its L2.
template <int LK, int DIAG, bool INV>
__global__ __launch_bounds__(LYM_DSLASH_THREADS) LYM_OCC void k_dslash(DslashParams a) {
    if (a.skip && *a.skip) return;
    const int nbt = a.nblocks_per_parity * (a.parity >= 0 ? 1 : 2);
    const int nx = 8;
    const int xcd = blockIdx.x % nx, k = blockIdx.x / nx;
    const int nphys = ((int)gridDim.x - xcd + nx - 1) / nx;  // workgroups on this XCD
    const int q = nbt / nx, r = nbt % nx;
    const int lo = xcd < r ? xcd * (q + 1) : r * (q + 1) + (xcd - r) * q;
    const int hi = lo + (xcd < r ? q + 1 : q);
    for (int lb = lo + k; lb < hi; lb += nphys) {
        int p, bi;
        if (a.parity >= 0) {
            p = a.parity;
            bi = lb;
        } else {
            p = lb & 1;
            bi = lb >> 1;
        }
        int idx = bi * LYM_DSLASH_THREADS + threadIdx.x;
        if (idx < a.g.Vh) dslash_site<LK, DIAG, INV>(a, p, idx);
    }
}

template <int LK, int DIAG, bool INV>
static void launch_one(const DslashParams& a, int nblocks, hipStream_t st) {
    hipLaunchKernelGGL((k_dslash<LK, DIAG, INV>), dim3(nblocks), dim3(LYM_DSLASH_THREADS), 0, st, a);
}

template <int LK>
static void launch_lk(const DslashParams& a, int diag, bool inv, int nblocks, hipStream_t st) {
    if (diag == 0) {
        if (inv) launch_one<LK, 0, true>(a, nblocks, st);
        else launch_one<LK, 0, false>(a, nblocks, st);
    } else if (diag == 1) {
        if (inv) launch_one<LK, 1, true>(a, nblocks, st);
        else launch_one<LK, 1, false>(a, nblocks, st);
    } else {
        launch_one<LK, 2, false>(a, nblocks, st);
    }
}

void launch_dslash(DslashParams a, int link_kind, int diag, bool inv, hipStream_t st) {
    a.nblocks_per_parity = (a.g.Vh + LYM_DSLASH_THREADS - 1) / LYM_DSLASH_THREADS;
    int nblocks = a.nblocks_per_parity * (a.parity >= 0 ? 1 : 2);
    // workgroups in flight: LYM_GRID_PER_CU per CU of the 256 CUs, sweeping
    if (LYM_GRID_PER_CU > 0 && nblocks > 256 * LYM_GRID_PER_CU) nblocks = 256 * LYM_GRID_PER_CU;
    switch (link_kind) {
        case 0: launch_lk<0>(a, diag, inv, nblocks, st); break;
        case 1: launch_lk<1>(a, diag, inv, nblocks, st); break;
        case 2: launch_lk<2>(a, diag, inv, nblocks, st); break;
        default: launch_lk<3>(a, diag, inv, nblocks, st); break;
    }
}

}  // namespace lym
