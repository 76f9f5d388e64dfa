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
// -1: automatic -- one sweeping workgroup per CU for launches of at most
// LYM_SWEEP_MAX_BLOCKS workgroups (small lattices, where it was ~4 % faster),
// all workgroups at once above that (large lattices, where sweeping was slower)
#ifndef LYM_GRID_PER_CU
#define LYM_GRID_PER_CU -1
#endif
#ifndef LYM_SWEEP_MAX_BLOCKS
#define LYM_SWEEP_MAX_BLOCKS 1024
#endif
// 1: each XCD works through its own contiguous range of site blocks;
// 0: all XCDs advance together through the lattice in block order
#ifndef LYM_XCD_MAP
#define LYM_XCD_MAP 1
#endif
// 1: within its block range an XCD takes the x-slices as the fastest index of
// its sweep (block j of every x-slice of the range, then block j + 1, ...), so
// the workgroups in flight together hold x-adjacent planes and the neighbour
// spinors and backward links of the x hops are read while still in L2;
// 0: plain block order within the range (one whole x-slice after the other).
// Measured on the 16^3x32 FP64 clover apply: L2 misses 277 -> 236-250 MB per
// apply, 49.5 -> 47.2 us; even-odd applies and solves gain alike.
#ifndef LYM_XSWEEP
#define LYM_XSWEEP 1
#endif
// 1: single-RHS FP32/FP16 applies of at most LYM_SPLIT_MAX_SITES sites per
// parity run two threads per site (k_dslash_split).  Measured: mixed CG on
// 16^3x32 (65536 sites) 5 % faster, on 24^3x48 (331776) 4 % slower
#ifndef LYM_SPLIT_LOWP
#define LYM_SPLIT_LOWP 1
#endif
#ifndef LYM_SPLIT_MAX_SITES
#define LYM_SPLIT_MAX_SITES 131072
#endif
// 1: XCDs 2i and 2i+1 share the x-slices of both their ranges, each taking half
// of every slice's site blocks, so an XCD's domain is 4 x-slices by half the
// (y, z) blocks instead of 2 whole x-slices: 3/4 of the halo spinors and
// links.  Measured on the 16^3x32 FP64 clover apply: 252 -> 221 MB read,
// 47.1 -> 44.6 us; even-odd launches 68.4 -> 65.8 us.
#ifndef LYM_XCD_PAIRS
#define LYM_XCD_PAIRS 1
#endif
// Full-lattice applies: odd-parity sites take each direction's backward hop
// before its forward hop, so the two reads of a link U_mu(y) -- by y (forward)
// and by y + mu (backward), sites of opposite parity -- happen at the same
// step of the kernel instead of one step apart.
#ifndef LYM_ALT_ORDER
#define LYM_ALT_ORDER 0
#endif
// Hops whose loads may be in flight together: a register barrier (pin) after
// every hop (1), every direction pair (2) or every two directions (4).
// More hops in flight raise memory-level parallelism per wave at the cost of
// registers, which pays when the lattice gives few waves per SIMD anyway.
// FP64: 2 (about 237 VGPRs, 2 waves per SIMD) -- the single-parity launches of
// the even-odd operator ran 7 % faster than with 1 (120 VGPRs), full applies
// alike; 4 (380 registers, 1 wave per SIMD) was slower.
#ifndef LYM_PIN_HOPS
#define LYM_PIN_HOPS 2
#endif
// the same for FP32/FP16 storage, whose hops need far fewer registers
// (8: no barrier between hops at all).  4 (125-150 VGPRs for FP16): the
// mixed-precision even-odd CG ran 5-14 % faster than with 1; 8 slowed the
// FP32 stencil down.
#ifndef LYM_PIN_HOPS_LOWP
#define LYM_PIN_HOPS_LOWP 4
#endif
// 1: single-RHS FP64 launches of 256-site blocks stage the other parity's
// spinors of their site block in LDS (k_dslash_stg); the z and t hops whose
// neighbour falls inside the block read it from there instead of from L2.
// Measured: 16^3x32 full apply 3 % faster, even-odd launches 2 %, 32^3x64
// full apply 5 % (same bits: only where the operands come from changes).
// Filling the copy with direct global-to-LDS loads instead of through
// registers, and a second copy filled for the next block while the current
// one computes, were measured too: the first the same, the second 15 %
// slower on 16^3x32.
#ifndef LYM_STAGE_ZT
#define LYM_STAGE_ZT 1
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
// SCALED: psi is an unscaled FP16 spinor and sc its site scale, applied (with
// the boundary sign) to the two colour vectors after the link multiply
template <int MU, bool FWD, int LK, class T, bool SCALED = false>
__device__ __forceinline__ void hop_dir(cplx<T> acc[12], const cplx<T> psi[12], const LinkMat<LK, T>& L, bool flip,
                                        T sc = (T)1) {
    using D = Dir<MU>;
    constexpr int PH0 = FWD ? D::C0 : D::C0 + 2;
    constexpr int PH1 = FWD ? D::C1 : D::C1 + 2;
    constexpr int Q0 = FWD ? 4 - D::C0 : 6 - D::C0;
    constexpr int Q1 = FWD ? 4 - D::C1 : 6 - D::C1;
    cplx<T> h0[3], h1[3], r0[3], r1[3];
    if constexpr (SCALED) {
        // FP16 operands: each component of h = psi_a + c psi_P as a fused
        // multiply-add by a unit the compiler cannot see through, so that the
        // conversions of both FP16 operands fold into one mixed-precision
        // instruction instead of two conversions and an add.  The value is
        // the same: the product by one is exact and the sum rounds once.
        T one = (T)1;
        asm volatile("" : "+v"(one));
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            const cplx<T> m0 = mulph<PH0>(psi[3 * D::P0 + c]), m1 = mulph<PH1>(psi[3 * D::P1 + c]);
            h0[c] = mk(fmadd(m0.re, one, psi[c].re), fmadd(m0.im, one, psi[c].im));
            h1[c] = mk(fmadd(m1.re, one, psi[3 + c].re), fmadd(m1.im, one, psi[3 + c].im));
        }
    } else {
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            h0[c] = psi[c] + mulph<PH0>(psi[3 * D::P0 + c]);
            h1[c] = psi[3 + c] + mulph<PH1>(psi[3 * D::P1 + c]);
            if (flip) {
                h0[c] = -h0[c];
                h1[c] = -h1[c];
            }
        }
    }
    if constexpr (FWD) {
        matvec<LK, T>(L, h0, r0);
        matvec<LK, T>(L, h1, r1);
    } else {
        matvec_dag<LK, T>(L, h0, r0);
        matvec_dag<LK, T>(L, h1, r1);
    }
    if constexpr (SCALED) {
        const T f = flip ? -sc : sc;
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            r0[c] = f * r0[c];
            r1[c] = f * r1[c];
        }
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
template <class T>
__device__ __forceinline__ void pin(cplx<T> acc[12]) {
#pragma unroll
    for (int k = 0; k < 12; ++k) asm volatile("" : "+v"(acc[k].re), "+v"(acc[k].im));
    asm volatile("" ::: "memory");
}

// the neighbour spinor at idx: from the LDS copy of the other parity's site
// block [lbase, lbase + LYM_DSLASH_THREADS) when lds is set and idx is inside,
// else from the field
template <int MU, int PR>
__device__ __forceinline__ void load_nb(const SpinorRef<PR>& src, int stride, int nr, int idx, int r,
                                        cplx<typename Prec<PR>::T> psi[12], typename Prec<PR>::T& sc,
                                        const cplx<double>* lds, int lbase) {
    if constexpr (PR == 0 && MU >= 2) {
        if (lds) {
            const unsigned j = (unsigned)(idx - lbase);
            if (j < (unsigned)LYM_DSLASH_THREADS) {
#pragma unroll
                for (int k = 0; k < 12; ++k) psi[k] = lds[k * LYM_DSLASH_THREADS + j];
                sc = 1.0;
                return;
            }
        }
    }
    load_spinor_unscaled<PR>(src, stride, nr, idx, r, psi, sc);
}

// one hop pair (forward and backward, or backward first when BWD_FIRST) in
// direction MU for site (p, site), rhs r; lds, lbase: as for load_nb
template <int MU, int PR, int LK, bool BWD_FIRST = false>
__device__ __forceinline__ void hop_pair(const DslashParams& a, int p, int site, int r, const int c[4],
                                         cplx<typename Prec<PR>::T> acc[12], const cplx<double>* lds = nullptr,
                                         int lbase = 0) {
    typedef typename Prec<PR>::T T;
    const Geom& g = a.g;
    const int q = p ^ g.tsh;
    int cu[4] = {c[0], c[1], c[2], c[3]}, cd[4] = {c[0], c[1], c[2], c[3]};
    cu[MU] = c[MU] + 1 == g.L[MU] ? 0 : c[MU] + 1;
    cd[MU] = c[MU] == 0 ? g.L[MU] - 1 : c[MU] - 1;
    const int iu = eo_index(g, cu[0], cu[1], cu[2], cu[3]);
    const int id = eo_index(g, cd[0], cd[1], cd[2], cd[3]);
    const bool anti = MU == 3 && g.bc_t < 0;
    const SpinorRef<PR> src{reinterpret_cast<const typename Prec<PR>::S*>(a.in[q]), a.in_sc[q]};
    constexpr bool SC = PR == 2;
    cplx<T> psi[12];
    T sc;
    LinkMat<LK, T> L;
    constexpr int PH = PR == 0 ? LYM_PIN_HOPS : LYM_PIN_HOPS_LOWP;
    constexpr bool pin_mid = PH == 1;
    constexpr bool pin_end = PH <= 2 || (PH == 4 && (MU & 1));
    if constexpr (!BWD_FIRST) {
        load_nb<MU, PR>(src, g.stride, a.nr, iu, r, psi, sc, lds, lbase);
        load_link<PR, LK>(a.links, g.stride, p, MU, site, L);
        hop_dir<MU, true, LK, T, SC>(acc, psi, L, anti && c[3] == g.L[3] - 1, sc);
        if constexpr (pin_mid) pin(acc);
    }
    {
        cplx<T> psi2[12];
        T sc2;
        LinkMat<LK, T> L2;
        load_nb<MU, PR>(src, g.stride, a.nr, id, r, psi2, sc2, lds, lbase);
        load_link<PR, LK>(a.links, g.stride, q, MU, id, L2);
        hop_dir<MU, false, LK, T, SC>(acc, psi2, L2, anti && c[3] == 0, sc2);
    }
    if constexpr (BWD_FIRST) {
        if constexpr (pin_mid) pin(acc);
        load_nb<MU, PR>(src, g.stride, a.nr, iu, r, psi, sc, lds, lbase);
        load_link<PR, LK>(a.links, g.stride, p, MU, site, L);
        hop_dir<MU, true, LK, T, SC>(acc, psi, L, anti && c[3] == g.L[3] - 1, sc);
    }
    if constexpr (pin_end) pin(acc);
}

template <int PR, int DIAG, bool INV, bool NORM, bool LO, bool UPD = false, int BM = 0>
__device__ __forceinline__ double dslash_epilogue(const DslashParams& a, int p, int site, int r,
                                                  const cplx<typename Prec<PR>::T> acc[12], float ual = 0.f);

// NORM (low precision only): returns |v|^2 of the site's output as computed,
// before rounding to the storage precision; else 0.  A template parameter, so
// kernels without it compile as if the sum did not exist.  LO: lane r has its
// own kappa and clover blocks (a.lane_ops); also a template parameter.  UPD:
// the fused mixed-CG update with step ual (a.upd_*): returns the site's
// |r_s|^2 after it, and stores no output.
template <int PR, int LK, int DIAG, bool INV, bool NORM = false, bool LO = false, bool UPD = false, int BM = 0>
__device__ __forceinline__ double dslash_site(const DslashParams& a, int p, int site, int r, float ual = 0.f,
                                              const cplx<double>* lds = nullptr, int lbase = 0) {
    typedef typename Prec<PR>::T T;
    int c[4];
    eo_coords(a.g, p, site, c);
    cplx<T> acc[12];
#pragma unroll
    for (int k = 0; k < 12; ++k) acc[k] = mk((T)0, (T)0);
    if (LYM_ALT_ORDER && a.parity < 0 && p == 1) {
        hop_pair<0, PR, LK, true>(a, p, site, r, c, acc);
        hop_pair<1, PR, LK, true>(a, p, site, r, c, acc);
        hop_pair<2, PR, LK, true>(a, p, site, r, c, acc);
        hop_pair<3, PR, LK, true>(a, p, site, r, c, acc);
    } else {
        hop_pair<0, PR, LK>(a, p, site, r, c, acc);
        hop_pair<1, PR, LK>(a, p, site, r, c, acc);
        hop_pair<2, PR, LK>(a, p, site, r, c, acc, lds, lbase);
        hop_pair<3, PR, LK>(a, p, site, r, c, acc, lds, lbase);
    }
    return dslash_epilogue<PR, DIAG, INV, NORM, LO, UPD, BM>(a, p, site, r, acc, ual);
}

// The site's output from its hopping sum acc: v = diag(psi_x) + hop_sign
// (-kappa acc), then R^-1, gamma5, the axpy and the store (the value
// returned is as for dslash_site)
template <int PR, int DIAG, bool INV, bool NORM, bool LO, bool UPD, int BM>
__device__ __forceinline__ double dslash_epilogue(const DslashParams& a, int p, int site, int r,
                                                  const cplx<typename Prec<PR>::T> acc[12], float ual) {
    typedef typename Prec<PR>::T T;
    typedef typename Prec<PR>::S S;
    const int ST = a.g.stride;
    const double kap = LO ? a.kappa_dev[r] : (a.kappa_dev ? *a.kappa_dev : a.kappa);
    const T hk = (T)(-kap * a.hop_sign);
    // this lane's clover blocks
    const void* Ab_ = a.A;
    const void* Ai_ = a.Ainv;
    const float* Asc_ = a.A_sc;
    const float* Aisc_ = a.Ainv_sc;
    if constexpr (LO) {
        // lanes of one kappa share a block set
        const int bs = a.lane_set ? a.lane_set[r] : r;
        Ab_ = reinterpret_cast<const char*>(a.A) + bs * a.A_lane_bytes;
        Ai_ = reinterpret_cast<const char*>(a.Ainv) + bs * a.Ainv_lane_bytes;
        if constexpr (PR == 2) {
            Asc_ = a.A_sc + bs * a.Asc_lane;
            Aisc_ = a.Ainv_sc + bs * a.Ainvsc_lane;
        }
    }
    cplx<T> v[12];
    if constexpr (DIAG == 0) {
#pragma unroll
        for (int k = 0; k < 12; ++k) v[k] = hk * acc[k];
    } else if constexpr (DIAG == 2 && PR >= 1) {
        // A (sc_psi psi) = (sc_A sc_psi) (A psi): both FP16 scales on the 12 outputs
        cplx<T> s[12], t[12];
        T scp;
        load_spinor_unscaled<PR>(SpinorRef<PR>{reinterpret_cast<const S*>(a.diag_in[p]), a.diag_sc[p]}, ST, a.nr,
                                 site, r, s, scp);
        const T sc = PR == 2 ? (T)Asc_[(size_t)p * ST + site] * scp : (T)1;
        herm6_packed<PR>(Ab_, 4 * (size_t)ST, (size_t)(2 * p) * ST, site, sc, s, t);
        herm6_packed<PR>(Ab_, 4 * (size_t)ST, (size_t)(2 * p + 1) * ST, site, sc, s + 6, t + 6);
#pragma unroll
        for (int k = 0; k < 12; ++k) v[k] = t[k] + hk * acc[k];
    } else {
        cplx<T> s[12];
        load_spinor<PR>(SpinorRef<PR>{reinterpret_cast<const S*>(a.diag_in[p]), a.diag_sc[p]}, ST, a.nr, site, r, s);
        if constexpr (DIAG == 2) {
            cplx<T> t[12];
            const S* Ab = reinterpret_cast<const S*>(Ab_) + (size_t)p * 36 * ST;
            herm6<PR>(Ab, ST, site, (T)1, s, t);
            herm6<PR>(Ab + (size_t)18 * ST, ST, site, (T)1, s + 6, t + 6);
#pragma unroll
            for (int k = 0; k < 12; ++k) v[k] = t[k] + hk * acc[k];
        } else {
#pragma unroll
            for (int k = 0; k < 12; ++k) v[k] = s[k] + hk * acc[k];
        }
    }
    if constexpr (INV) {
        cplx<T> t[12];
        if constexpr (PR == 0) {
            const S* Ai = reinterpret_cast<const S*>(Ai_);
            herm6<PR>(Ai, ST, site, (T)1, v, t);
            herm6<PR>(Ai + (size_t)18 * ST, ST, site, (T)1, v + 6, t + 6);
        } else {
            const T sc = PR == 2 ? (T)Aisc_[site] : (T)1;
            herm6_packed<PR>(Ai_, 2 * (size_t)ST, 0, site, sc, v, t);
            herm6_packed<PR>(Ai_, 2 * (size_t)ST, (size_t)ST, site, sc, v + 6, t + 6);
        }
#pragma unroll
        for (int k = 0; k < 12; ++k) v[k] = t[k];
    }
    if (a.gamma5) {
#pragma unroll
        for (int k = 6; k < 12; ++k) v[k] = -v[k];
    }
    if (a.v2[p]) {
        cplx<T> w[12];
        load_spinor<PR, true>(SpinorRef<PR>{reinterpret_cast<const S*>(a.v2[p]), a.v2_sc[p]}, ST, a.nr, site, r, w);
        const cplx<T> al = mk((T)a.alpha.re, (T)a.alpha.im);
#pragma unroll
        for (int k = 0; k < 12; ++k) v[k] = v[k] + al * w[k];
    }
    if constexpr (BM == 2) {
        // batch lane r: x_s += alpha_r p, r_s -= alpha_r v (v in FP32, never
        // rounded to storage), the site's |r_s|^2 to bsite
        static_assert(PR >= 1, "the fused update is the mixed-precision solver's");
        const float al = (float)a.balpha[r];
        cplx<T> pv[12];
        load_spinor<PR>(SpinorRef<PR>{reinterpret_cast<const S*>(a.upd_p[p]), a.upd_p_sc[p]}, ST, a.nr, site, r, pv);
        cplx<float>* xs = reinterpret_cast<cplx<float>*>(a.upd_xs[p]);
        cplx<float>* rs = reinterpret_cast<cplx<float>*>(a.upd_rs[p]);
        const size_t o = (size_t)site * a.nr + r, ks = (size_t)ST * a.nr;
        double n2 = 0;
#pragma unroll
        for (int k = 0; k < 12; ++k) {
            cplx<float> x = xs[k * ks + o];
            cplx<float> rr = rs[k * ks + o];
            x = x + al * pv[k];
            rr = rr - al * v[k];
            xs[k * ks + o] = x;
            rs[k * ks + o] = rr;
            n2 += (double)rr.re * rr.re + (double)rr.im * rr.im;
        }
        a.bsite[p][o] = n2;
        return 0;
    }
    if constexpr (UPD) {
        // x_s += alpha p, r_s -= alpha v (v in FP32, never rounded to storage)
        static_assert(PR >= 1, "the fused update is the mixed-precision solver's");
        cplx<T> pv[12];
        load_spinor<PR>(SpinorRef<PR>{reinterpret_cast<const S*>(a.upd_p[p]), a.upd_p_sc[p]}, ST, 1, site, 0, pv);
        cplx<float>* xs = reinterpret_cast<cplx<float>*>(a.upd_xs[p]);
        cplx<float>* rs = reinterpret_cast<cplx<float>*>(a.upd_rs[p]);
        double n2 = 0;
#pragma unroll
        for (int k = 0; k < 12; ++k) {
            cplx<float> x = fld<cplx<float>>(xs + (size_t)k * ST, (unsigned)site);
            cplx<float> rr = fld<cplx<float>>(rs + (size_t)k * ST, (unsigned)site);
            x = x + ual * pv[k];
            rr = rr - ual * v[k];
            fst<cplx<float>, false>(xs + (size_t)k * ST, (unsigned)site, x);
            fst<cplx<float>, false>(rs + (size_t)k * ST, (unsigned)site, rr);
            n2 += (double)rr.re * rr.re + (double)rr.im * rr.im;
        }
        return n2;
    }
    store_spinor<PR>(SpinorOut<PR>{reinterpret_cast<S*>(a.out[p]), a.out_sc[p]}, ST, a.nr, site, r, v);
    if constexpr (BM == 1) {  // lane r's |v|^2 at this site: the 24 squares in FP32
        float n2 = 0.f;
#pragma unroll
        for (int k = 0; k < 12; ++k) n2 += v[k].re * v[k].re + v[k].im * v[k].im;
        a.bsite[p][(size_t)site * a.nr + r] = (double)n2;
    }
    if constexpr (NORM) {  // the site's 24 squares in FP32, the sum over sites in FP64
        float n2 = 0.f;
#pragma unroll
        for (int k = 0; k < 12; ++k) n2 += v[k].re * v[k].re + v[k].im * v[k].im;
        return (double)n2;
    }
    return 0;
}

// Sum of v over the workgroup (blockDim.x a multiple of 64, at most 1024), in
// one fixed order; the total is written by thread 0 to out[blockIdx.x].  Every
// thread of the workgroup must call it.
__device__ __forceinline__ void block_sum_store(double v, double* out) {
    __shared__ double sh[16];
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
    if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = v;
    __syncthreads();
    if (threadIdx.x == 0) {
        double t = 0;
        for (int w = 0; w < (int)(blockDim.x >> 6); ++w) t += sh[w];
        out[blockIdx.x] = t;
    }
}

// Grid: gridDim.x workgroups dealt round-robin over the 8 XCDs.  XCD x owns a
// contiguous range of logical site blocks (x-major lattice order, the two
// parities of a block range interleaved) and its workgroups sweep that range
// in order, so at any time an XCD works on one compact slab and the
// neighbour spinors and links of that slab are reused from its L2.  A
// workgroup holds `spb` sites x `nr` right-hand sides, rhs fastest.
// (The mapping functions take the workgroup index and grid size as arguments so
// that the host can evaluate the same mapping: lym_launch_plan below.)
struct BlockSweep {
    int lo, hi, k, nphys;
};
__host__ __device__ __forceinline__ BlockSweep block_sweep_at(const DslashParams& a, int bx, int gdim) {
    const int nbt = a.nblocks_per_parity * (a.parity >= 0 ? 1 : 2);
    BlockSweep b;
#if LYM_XCD_MAP
    const int nx = 8;
    const int xcd = bx % nx;
    b.k = bx / nx;
    b.nphys = (gdim - xcd + nx - 1) / nx;  // workgroups on this XCD
    const int q = nbt / nx, rr = nbt % nx;
    b.lo = xcd < rr ? xcd * (q + 1) : rr * (q + 1) + (xcd - rr) * q;
    b.hi = b.lo + (xcd < rr ? q + 1 : q);
#else
    // plain order: consecutive workgroups (dealt round-robin over the XCDs) take
    // consecutive site blocks, so all XCDs advance one front through the lattice
    // and the neighbour slices they re-read are recent in the shared Infinity Cache
    b.k = bx;
    b.nphys = gdim;
    b.lo = 0;
    b.hi = nbt;
#endif
    return b;
}

// Which renumbering block_of_at applied (reported by lym_launch_plan)
enum { MAP_PLAIN = 0, MAP_XSLICE = 1, MAP_XCD_PAIRS = 2 };

// logical block lb of the sweep of workgroup bx -> parity p and site block bi
// of that parity; returns the MAP_* branch taken
__host__ __device__ __forceinline__ int block_of_at(const DslashParams& a, const BlockSweep& b, int lb, int bx, int& p,
                                                    int& bi) {
    (void)bx;
    if (a.parity >= 0) {
        p = a.parity;
        bi = lb;
    } else {
        p = lb & 1;
        bi = lb >> 1;
    }
#if LYM_XSWEEP && LYM_XCD_MAP
    {
        // the range holds whole x-slices: renumber its blocks x-slice fastest
        const int W = a.xslice_blocks;
        const int b0 = a.parity >= 0 ? b.lo : b.lo >> 1;
        const int D = a.parity >= 0 ? b.hi - b.lo : (b.hi - b.lo) >> 1;
        const bool whole = a.parity >= 0 || ((b.lo | b.hi) & 1) == 0;
#if LYM_XCD_PAIRS
        // XCDs 2i and 2i+1 share the x-slices of both their ranges, each taking
        // half of every slice's blocks (a 2-D domain: shorter x halo)
        const int nbt = a.nblocks_per_parity * (a.parity >= 0 ? 1 : 2);
        if (W > 0 && whole && W % 2 == 0 && nbt % 16 == 0 && b0 % W == 0 && D % W == 0) {
            const int x1 = (bx % 8) & 1;
            const int b0p = b0 - x1 * D, nxs2 = 2 * D / W, Wh = W / 2;
            const int u = bi - b0, j = u / nxs2;
            bi = b0p + (u - j * nxs2) * W + x1 * Wh + j;
            return MAP_XCD_PAIRS;
        }
#endif
        if (W > 0 && whole && b0 % W == 0 && D % W == 0 && D > W) {
            const int u = bi - b0, nxs = D / W;
            const int j = u / nxs;
            bi = b0 + (u - j * nxs) * W + j;
            return MAP_XSLICE;
        }
    }
#endif
    return MAP_PLAIN;
}

__device__ __forceinline__ BlockSweep block_sweep(const DslashParams& a) {
    return block_sweep_at(a, (int)blockIdx.x, (int)gridDim.x);
}
__device__ __forceinline__ void block_of(const DslashParams& a, const BlockSweep& b, int lb, int& p, int& bi) {
    (void)block_of_at(a, b, lb, (int)blockIdx.x, p, bi);
}

// The fused update's step: alpha = rr_s / (sum of the n <p, Ap> partials),
// folded by every workgroup in one fixed order (so all get the same alpha);
// workgroup 0 stores alpha and rr_prev = rr_s in the state.  Every thread of
// the workgroup must call it.
__device__ __forceinline__ float upd_alpha(const DslashParams& a) {
    __shared__ double sh[16];
    double v = 0;
    for (int k = (int)threadIdx.x; k < a.upd_ndot; k += (int)blockDim.x) v += a.upd_dot[k];
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
    if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = v;
    __syncthreads();
    double t = 0;
    for (int w = 0; w < (int)(blockDim.x >> 6); ++w) t += sh[w];
    MCGState* st = reinterpret_cast<MCGState*>(a.upd_state);
    const double rr = st->rr_s;
    const double alpha = rr / t;
    __syncthreads();  // every thread has read rr_s before workgroup 0 moves it to rr_prev
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        st->alpha = alpha;
        st->rr_prev = rr;
    }
    return (float)alpha;
}

// NORM: also the per-workgroup output norms (a.norm_partials); LO: per-lane
// operators; UPD: the fused mixed-CG update instead of the output store
template <int PR, int LK, int DIAG, bool INV, bool NORM = false, bool LO = false, bool UPD = false, int BM = 0>
__global__ __launch_bounds__(LYM_DSLASH_THREADS) LYM_OCC void k_dslash(DslashParams a) {
    static_assert(!NORM || PR >= 1, "output norms are the low-precision stencil's");
    static_assert(!(NORM && UPD), "one per-workgroup sum per launch");
    if (a.skip && *a.skip) return;
    const BlockSweep b = block_sweep(a);
    const int ls = (int)threadIdx.x / a.nr, r = (int)threadIdx.x - ls * a.nr;
    if constexpr (UPD) {
        // one RHS, no lane flags: every thread holds a site slot and reaches the sums
        const float ual = upd_alpha(a);
        double n2 = 0;
        for (int lb = b.lo + b.k; lb < b.hi; lb += b.nphys) {
            int p, bi;
            block_of(a, b, lb, p, bi);
            const int site = bi * a.spb + (int)threadIdx.x;
            if (site < a.g.Vh) n2 += dslash_site<PR, LK, DIAG, INV, false, false, true>(a, p, site, 0, ual);
        }
        block_sum_store(n2, a.upd_rpart);
    } else if constexpr (!NORM) {
        if (ls >= a.spb) return;
        if (a.lane_flags && a.lane_flags[r] == a.lane_skip_value) return;
        for (int lb = b.lo + b.k; lb < b.hi; lb += b.nphys) {
            int p, bi;
            block_of(a, b, lb, p, bi);
            const int site = bi * a.spb + ls;
            if (site < a.g.Vh) dslash_site<PR, LK, DIAG, INV, false, LO, false, BM>(a, p, site, r);
        }
    } else {
        // one RHS, no lane flags (the launcher's condition for output norms):
        // every thread holds a site slot, and every thread reaches the sum
        double n2 = 0;
        for (int lb = b.lo + b.k; lb < b.hi; lb += b.nphys) {
            int p, bi;
            block_of(a, b, lb, p, bi);
            const int site = bi * a.spb + (int)threadIdx.x;
            if (site < a.g.Vh) n2 += dslash_site<PR, LK, DIAG, INV, true>(a, p, site, 0);
        }
        block_sum_store(n2, a.norm_partials);
    }
}

#if LYM_STAGE_ZT
// FP64, one RHS, one thread per site (spb == LYM_DSLASH_THREADS): each site
// block first copies the other parity's spinors of the same block index --
// with t fastest, the z and t neighbours of most of the block's sites -- into
// LDS, coalesced, once; those hops then read LDS instead of issuing four
// reads of every such spinor to L2.
template <int LK, int DIAG, bool INV>
__global__ __launch_bounds__(LYM_DSLASH_THREADS) LYM_OCC void k_dslash_stg(DslashParams a) {
    __shared__ cplx<double> stg[12 * LYM_DSLASH_THREADS];
    if (a.skip && *a.skip) return;
    const BlockSweep b = block_sweep(a);
    const int t = (int)threadIdx.x;
    for (int lb = b.lo + b.k; lb < b.hi; lb += b.nphys) {
        int p, bi;
        block_of(a, b, lb, p, bi);
        const int base = bi * LYM_DSLASH_THREADS;
        __syncthreads();  // the previous block's reads of stg are done
        if (base + t < a.g.Vh) {
            const zd* src = reinterpret_cast<const zd*>(a.in[p ^ a.g.tsh]);
#pragma unroll
            for (int k = 0; k < 12; ++k)
                stg[k * LYM_DSLASH_THREADS + t] = fld<zd>(src + (size_t)k * a.g.stride, (unsigned)(base + t));
        }
        __syncthreads();
        const int site = base + t;
        if (site < a.g.Vh) dslash_site<0, LK, DIAG, INV>(a, p, site, 0, 0.f, stg, base);
    }
}

#endif

// ---------------------------------------------------------------- two threads per site (FP32/FP16)
// Split form of the single-RHS low-precision stencil: a workgroup holds its
// sites twice.  The waves of the first half take the x and y hops of every
// site, those of the second half the z and t hops; the halves swap, through
// LDS, the partial sums of the chirality the other half finishes, and each
// half applies the diagonal, R^-1, gamma5, the axpy and the store to one
// chirality (spins 0,1 or 2,3: the clover blocks are chiral, so the epilogue
// splits exactly).  An FP16 site's scale is the largest component of both
// halves, also swapped through LDS.  Twice the waves per site block hide the
// load latency where one thread per site leaves one wave per SIMD.
template <int PR, int LK, int DIAG, bool INV>
__device__ __forceinline__ void epilogue_half_lp(const DslashParams& a, int p, int site, int r, int h,
                                                 const cplx<float> u[6], cplx<float> v[6]) {
    typedef float T;
    typedef typename Prec<PR>::S S;
    const int ST = a.g.stride;
    const size_t ks = (size_t)ST * a.nr;
    const unsigned o = (unsigned)site * (unsigned)a.nr + (unsigned)r;
    const double kap = a.kappa_dev ? *a.kappa_dev : a.kappa;
    const T hk = (T)(-kap * a.hop_sign);
    if constexpr (DIAG == 0) {
#pragma unroll
        for (int k = 0; k < 6; ++k) v[k] = hk * u[k];
    } else {
        S e[6];
        load_half_lp<PR>(reinterpret_cast<const S*>(a.diag_in[p]), ks, o, h, e);
        const T scp = PR == 2 ? fld<float>(a.diag_sc[p], o) : (T)1;
        cplx<T> s6[6];
#pragma unroll
        for (int k = 0; k < 6; ++k) s6[k] = to_c<PR>(e[k]);
        if constexpr (DIAG == 2) {
            const T sc = PR == 2 ? (T)a.A_sc[(size_t)p * ST + site] * scp : (T)1;
            cplx<T> t[6];
            herm6_packed<PR>(a.A, 4 * (size_t)ST, (size_t)(2 * p + h) * ST, site, sc, s6, t);
#pragma unroll
            for (int k = 0; k < 6; ++k) v[k] = t[k] + hk * u[k];
        } else {
#pragma unroll
            for (int k = 0; k < 6; ++k) v[k] = scp * s6[k] + hk * u[k];
        }
    }
    if constexpr (INV) {
        const T sc = PR == 2 ? (T)a.Ainv_sc[site] : (T)1;
        cplx<T> t[6];
        herm6_packed<PR>(a.Ainv, 2 * (size_t)ST, (size_t)h * ST, site, sc, v, t);
#pragma unroll
        for (int k = 0; k < 6; ++k) v[k] = t[k];
    }
    if (a.gamma5 && h) {
#pragma unroll
        for (int k = 0; k < 6; ++k) v[k] = -v[k];
    }
    if (a.v2[p]) {
        S e[6];
        load_half_lp<PR, true>(reinterpret_cast<const S*>(a.v2[p]), ks, o, h, e);
        const T sc2 = PR == 2 ? fld<float>(a.v2_sc[p], o) : (T)1;
        const cplx<T> al = mk((T)a.alpha.re, (T)a.alpha.im);
#pragma unroll
        for (int k = 0; k < 6; ++k) v[k] = v[k] + al * (sc2 * to_c<PR>(e[k]));
    }
}

// blockDim.x = 2 * TH, TH a multiple of 64 (the launcher's thread count), so
// every wave lies in one half.  Every thread reaches both barriers of every
// block of the sweep (invalid lanes carry zeros and store nothing).
template <int PR, int LK, int DIAG, bool INV, bool NORM = false, bool UPD = false>
__global__ __launch_bounds__(2 * LYM_DSLASH_THREADS) LYM_OCC void k_dslash_split(DslashParams a) {
    static_assert(PR >= 1, "the split form is the low-precision stencil's");
    static_assert(!(NORM && UPD), "one per-workgroup sum per launch");
    typedef float T;
    typedef typename Prec<PR>::S S;
    __shared__ cplx<T> xch[2][6][LYM_DSLASH_THREADS];
    __shared__ float mx[2][LYM_DSLASH_THREADS];
    if (a.skip && *a.skip) return;
    const BlockSweep b = block_sweep(a);
    const int TH = (int)blockDim.x >> 1;
    const int h = __builtin_amdgcn_readfirstlane((int)threadIdx.x >= TH ? 1 : 0);
    const int tl = (int)threadIdx.x - h * TH;
    const int ls = tl / a.nr, r = tl - ls * a.nr;
    const bool lane_ok = ls < a.spb && !(a.lane_flags && a.lane_flags[r] == a.lane_skip_value);
    const size_t ks = (size_t)a.g.stride * a.nr;
    [[maybe_unused]] double n2 = 0;
    [[maybe_unused]] float ual = 0.f;
    if constexpr (UPD) ual = upd_alpha(a);
    for (int lb = b.lo + b.k; lb < b.hi; lb += b.nphys) {
        int p, bi;
        block_of(a, b, lb, p, bi);
        const int site = bi * a.spb + ls;
        const bool ok = lane_ok && site < a.g.Vh;
        cplx<T> acc[12];
#pragma unroll
        for (int k = 0; k < 12; ++k) acc[k] = mk((T)0, (T)0);
        if (ok) {
            int c[4];
            eo_coords(a.g, p, site, c);
            if (h == 0) {
                hop_pair<0, PR, LK>(a, p, site, r, c, acc);
                hop_pair<1, PR, LK>(a, p, site, r, c, acc);
            } else {
                hop_pair<2, PR, LK>(a, p, site, r, c, acc);
                hop_pair<3, PR, LK>(a, p, site, r, c, acc);
            }
        }
        // (constant indices only: a select between acc[k] and acc[6 + k] would
        // become a dynamic index and put the accumulator in scratch memory)
        if (h == 0) {
#pragma unroll
            for (int k = 0; k < 6; ++k) xch[0][k][tl] = acc[6 + k];
        } else {
#pragma unroll
            for (int k = 0; k < 6; ++k) xch[1][k][tl] = acc[k];
        }
        __syncthreads();
        cplx<T> u[6], v[6];
        if (h == 0) {
#pragma unroll
            for (int k = 0; k < 6; ++k) u[k] = acc[k] + xch[1][k][tl];
        } else {
#pragma unroll
            for (int k = 0; k < 6; ++k) u[k] = acc[6 + k] + xch[0][k][tl];
        }
#pragma unroll
        for (int k = 0; k < 6; ++k) v[k] = mk((T)0, (T)0);
        if (ok) epilogue_half_lp<PR, LK, DIAG, INV>(a, p, site, r, h, u, v);
        if constexpr (NORM) {  // invalid lanes hold zeros
            float h2 = 0.f;
#pragma unroll
            for (int k = 0; k < 6; ++k) h2 += v[k].re * v[k].re + v[k].im * v[k].im;
            n2 += (double)h2;
        }
        const unsigned o = (unsigned)site * (unsigned)a.nr + (unsigned)r;
        S e[6];
        if constexpr (UPD) {
            // x_s += alpha p, r_s -= alpha v on this half's components 6 h + k
            if (ok) {
                load_half_lp<PR>(reinterpret_cast<const S*>(a.upd_p[p]), ks, o, h, e);
                const T psc = PR == 2 ? fld<float>(a.upd_p_sc[p], o) : (T)1;
                const int ST = a.g.stride;
                cplx<float>* xs = reinterpret_cast<cplx<float>*>(a.upd_xs[p]) + (size_t)(6 * h) * ST;
                cplx<float>* rs = reinterpret_cast<cplx<float>*>(a.upd_rs[p]) + (size_t)(6 * h) * ST;
#pragma unroll
                for (int k = 0; k < 6; ++k) {
                    cplx<float> x = fld<cplx<float>>(xs + (size_t)k * ST, (unsigned)site);
                    cplx<float> rr = fld<cplx<float>>(rs + (size_t)k * ST, (unsigned)site);
                    x = x + ual * (psc * to_c<PR>(e[k]));
                    rr = rr - ual * v[k];
                    fst<cplx<float>, false>(xs + (size_t)k * ST, (unsigned)site, x);
                    fst<cplx<float>, false>(rs + (size_t)k * ST, (unsigned)site, rr);
                    n2 += (double)rr.re * rr.re + (double)rr.im * rr.im;
                }
            }
            __syncthreads();  // every xch read of this block is done before the next block writes it
        } else if constexpr (PR == 2) {
            float m = 0.f;
#pragma unroll
            for (int k = 0; k < 6; ++k) m = fmaxf(m, fmaxf(fabsf(v[k].re), fabsf(v[k].im)));
            mx[h][tl] = m;
            __syncthreads();  // also: every xch read of this block is done
            m = fmaxf(mx[0][tl], mx[1][tl]);
            const float inv = fp16_site_inv(m);
#pragma unroll
            for (int k = 0; k < 6; ++k) e[k] = from_c<PR>(inv * v[k]);
            if (ok) {
                store_half_lp<PR>(reinterpret_cast<S*>(a.out[p]), ks, o, h, e);
                if (h == 0) fst<float, false>(a.out_sc[p], o, m);
            }
        } else {
            __syncthreads();  // every xch read of this block is done before the next block writes it
#pragma unroll
            for (int k = 0; k < 6; ++k) e[k] = from_c<PR>(v[k]);
            if (ok) store_half_lp<PR>(reinterpret_cast<S*>(a.out[p]), ks, o, h, e);
        }
    }
    if constexpr (NORM) block_sum_store(n2, a.norm_partials);  // every thread reaches it
    if constexpr (UPD) block_sum_store(n2, a.upd_rpart);
}

// ---------------------------------------------------------------- multi-RHS FP16 stencil on the matrix cores
// k_dslash_mfma: the hopping term of a batch of nr = 2, 4 or 8 right-hand
// sides per site (FP16 storage, SU(3) links with all three rows) with the link
// product on the matrix cores, one v_mfma_f32_16x16x16_f16 per hop and site:
//   C (16 x 16) = A (16 x 16) . B (16 x 16), FP16 inputs, FP32 products and sums
//   B column n = 2 r + h (RHS r, spin half h = 0, 1); B rows k 0-5: the
//     neighbour spinor's spin h, colour k / 2, (re, im); k 6-11: its partner
//     spin P_h of the direction (negated for h = 1 when c_1 = -c_0); 12-15 zero
//   A rows m 0-5: the direct output (spin h), colour m / 2, (re, im); m 6-11:
//     the partner output (spin P_h); 12-15 zero.  A is the real form of the
//     3x3 blocks  forward: direct [U | c U], partner [c* U | U];
//                 backward: direct [U+ | -c U+], partner [-c* U+ | U+]
//     with c = i^C0 of the direction,
// so the projection (1 -+ gamma_mu), the link multiply and the
// reconstruction all happen inside the MFMA on the stored FP16 values, with
// no rounding before the FP32 sums.  The neighbour's FP16 scale (one per site
// and RHS, i.e. per column), the boundary sign and the sign of partner rows
// for h = 1 multiply C in FP32.  Lane l holds A row l % 16 and B column
// l % 16 at k = 4 (l / 16) + 0..3, and C column l % 16, rows 4 (l / 16) +
// 0..3 (checked on the device).  A wave works through its sites one at a
// time (the site, its neighbours and the link bases are wave-uniform); the
// hopping sums go through LDS to one thread per (site, RHS), which runs the
// epilogue of the VALU stencil (diagonal, R^-1, gamma5, axpy, FP16 store).
typedef _Float16 hv4 __attribute__((ext_vector_type(4)));
typedef float fv4 __attribute__((ext_vector_type(4)));

// z i^p on FP16 values (exact: a swap and sign changes)
__device__ __forceinline__ h2_t mul_ip(h2_t z, int p) {
    const bool sw = (p & 1) != 0;
    _Float16 x = sw ? z.y : z.x, y = sw ? z.x : z.y;
    if (p == 1 || p == 2) x = -x;
    if (p >= 2) y = -y;
    h2_t o;
    o.x = x;
    o.y = y;
    return o;
}

constexpr int MFMA_LDS_STRIDE = 26;  // floats per (site, RHS) slot: 24 used, padded against bank conflicts

// lane constants of the MFMA stencil
struct MfmaLane {
    int rr, hh;         // B/C column: RHS and spin half
    bool bval;          // column rr < nr and k-group < 3: this lane loads B
    bool aval;          // A row < 12 and k-group < 3: this lane loads A
    int arow, aore;     // A row: 0 direct, 1 partner; 0 real row, 1 imaginary row
    bool partP[2];      // A/B k slot s: colour block of the partner spin (else of spin h)
    unsigned lofs_f[2], lofs_b[2];  // link entry of slot s (in units of the link component stride)
    int kgrp;           // k-group l / 16
};

template <int MU>
__device__ __forceinline__ int mfma_phase(const MfmaLane& L, int s, bool fwd) {
    constexpr int C0 = Dir<MU>::C0;
    if (L.arow == 0 && L.partP[s]) return fwd ? C0 : (C0 + 2) & 3;
    if (L.arow == 1 && !L.partP[s]) return fwd ? (4 - C0) & 3 : (6 - C0) & 3;
    return 0;
}

// Operands of one hop (direction MU, forward or backward) of the wave's
// current site: A (the link block), B (the neighbour spinor's components) and
// the neighbour's FP16 scale.  Branch-free: lanes without an operand load a
// valid address and select zero, so the loads of all eight hops of a site can
// be issued before the first MFMA waits on them.  lsite: the link's site
// index on parity lp; nb: the neighbour spinor's index on parity q
template <int MU, bool FWD>
__device__ __forceinline__ void mfma_load(const DslashParams& a, const MfmaLane& L, int lp, int lsite, int q, int nb,
                                          const unsigned sofs[2], bool use_b, hv4& Av, hv4& Bv, float& sc) {
    const int ST = a.g.stride;
    const h2_t* lb = reinterpret_cast<const h2_t*>(a.links) + (size_t)((lp * 4 + MU) * 9) * ST + lsite;
    const h2_t* sb = reinterpret_cast<const h2_t*>(a.in[q]) + (size_t)nb * a.nr * 4;
    constexpr bool SIGN1 = Dir<MU>::C1 != Dir<MU>::C0;  // c_1 = -c_0
    const _Float16 hz = (_Float16)0;
#pragma unroll
    for (int s = 0; s < 2; ++s) {
        h2_t z = fld<h2_t>(lb, L.aval ? (FWD ? L.lofs_f[s] : L.lofs_b[s]) : 0u);
        if (!FWD) z.y = -z.y;
        z = mul_ip(z, mfma_phase<MU>(L, s, FWD));
        Av[2 * s] = L.aval ? (L.aore ? z.y : z.x) : hz;
        Av[2 * s + 1] = L.aval ? (L.aore ? z.x : (_Float16)(-z.y)) : hz;
        h2_t v = fld<h2_t>(sb, use_b ? sofs[s] : 0u);
        if (SIGN1 && L.hh && L.partP[s]) {
            v.x = -v.x;
            v.y = -v.y;
        }
        Bv[2 * s] = use_b ? v.x : hz;
        Bv[2 * s + 1] = use_b ? v.y : hz;
    }
    const float f = fld<float>(a.in_sc[q] + (size_t)nb * a.nr, use_b ? (unsigned)L.rr : 0u);
    sc = use_b ? f : 0.f;
}

// G += f * (A . B) for one hop: f = the neighbour's scale, with the boundary
// sign and, for partner rows of columns h = 1 where c_1 = -c_0, a minus sign
template <int MU>
__device__ __forceinline__ void mfma_acc(const MfmaLane& L, const hv4& Av, const hv4& Bv, float sc, bool flip,
                                         fv4& G) {
    constexpr bool SIGN1 = Dir<MU>::C1 != Dir<MU>::C0;
    const fv4 zero = {0.f, 0.f, 0.f, 0.f};
    const fv4 Cv = __builtin_amdgcn_mfma_f32_16x16x16f16(Av, Bv, zero, 0, 0, 0);
    const float f = flip ? -sc : sc;
#pragma unroll
    for (int t = 0; t < 4; ++t) {
        // C row 4 k + t is a partner row for k = 1, t >= 2 and for k = 2
        const bool prow = L.kgrp == 2 || (L.kgrp == 1 && t >= 2);
        G[t] += ((SIGN1 && L.hh && prow) ? -f : f) * Cv[t];
    }
}

template <int DIAG, bool INV, bool LO>
__global__ __launch_bounds__(LYM_DSLASH_THREADS) void k_dslash_mfma(DslashParams a) {
    __shared__ float accs[LYM_DSLASH_THREADS * MFMA_LDS_STRIDE];
    if (a.skip && *a.skip) return;
    const BlockSweep b = block_sweep(a);
    const int nr = a.nr, spb = a.spb;
    const int lane = (int)threadIdx.x & 63, wv = (int)threadIdx.x >> 6;
    const int spw = 64 / nr;  // sites per wave (nr divides 64)
    MfmaLane L;
    {
        const int n = lane & 15;
        L.kgrp = lane >> 4;
        L.rr = n >> 1;
        L.hh = n & 1;
        L.bval = L.rr < nr && L.kgrp < 3;
        L.arow = n < 6 ? 0 : 1;
        L.aore = n & 1;
        L.aval = n < 12 && L.kgrp < 3;
        const int ai = (n % 6) >> 1;
#pragma unroll
        for (int s = 0; s < 2; ++s) {
            const int idx = 2 * L.kgrp + s;  // 0..5: (spin h, colour 0..2), then (partner, colour 0..2)
            const int j = idx % 3;
            L.partP[s] = idx >= 3;
            L.lofs_f[s] = (unsigned)((3 * ai + j) * a.g.stride);
            L.lofs_b[s] = (unsigned)((3 * j + ai) * a.g.stride);
        }
    }
    // B offsets (h2 elements from the neighbour's base) per direction and slot
    unsigned sofs[4][2];
    {
        const size_t ks = (size_t)a.g.stride * nr;
        const int P0[4] = {Dir<0>::P0, Dir<1>::P0, Dir<2>::P0, Dir<3>::P0};
        const int P1[4] = {Dir<0>::P1, Dir<1>::P1, Dir<2>::P1, Dir<3>::P1};
#pragma unroll
        for (int mu = 0; mu < 4; ++mu)
#pragma unroll
            for (int s = 0; s < 2; ++s) {
                const int idx = 2 * L.kgrp + s;
                const int j = idx % 3;
                const int spin = idx < 3 ? L.hh : (L.hh ? P1[mu] : P0[mu]);
                const int comp = 3 * spin + j;  // 16-byte vector comp / 4, slot comp % 4
                sofs[mu][s] = L.kgrp < 3 ? (unsigned)(((comp >> 2) * ks + L.rr) * 4 + (comp & 3)) : 0u;
            }
    }
    const bool anti = a.g.bc_t < 0;
    const bool skip_col = L.rr < nr && a.lane_flags && a.lane_flags[L.rr] == a.lane_skip_value;
    for (int lb = b.lo + b.k; lb < b.hi; lb += b.nphys) {
        int p, bi;
        block_of(a, b, lb, p, bi);
        const int q = p ^ a.g.tsh;
        for (int sl = 0; sl < spw; ++sl) {
            const int ls = wv * spw + sl;
            const int site = __builtin_amdgcn_readfirstlane(bi * spb + ls);
            if (ls >= spb || site >= a.g.Vh) break;
            int c[4];
            eo_coords(a.g, p, site, c);
            const bool use_b = L.bval && !skip_col;
            // neighbour indices (wave-uniform) and the operands of all eight hops
            hv4 Ah[8], Bh[8];
            float sch[8];
            bool fl[8];
#define LYM_MFMA_LOAD(MU)                                                                                        \
    {                                                                                                            \
        int cu[4] = {c[0], c[1], c[2], c[3]}, cd[4] = {c[0], c[1], c[2], c[3]};                                  \
        cu[MU] = c[MU] + 1 == a.g.L[MU] ? 0 : c[MU] + 1;                                                         \
        cd[MU] = c[MU] == 0 ? a.g.L[MU] - 1 : c[MU] - 1;                                                         \
        const int iu = __builtin_amdgcn_readfirstlane(eo_index(a.g, cu[0], cu[1], cu[2], cu[3]));                \
        const int id = __builtin_amdgcn_readfirstlane(eo_index(a.g, cd[0], cd[1], cd[2], cd[3]));                \
        mfma_load<MU, true>(a, L, p, site, q, iu, sofs[MU], use_b, Ah[2 * MU], Bh[2 * MU], sch[2 * MU]);         \
        mfma_load<MU, false>(a, L, q, id, q, id, sofs[MU], use_b, Ah[2 * MU + 1], Bh[2 * MU + 1], sch[2 * MU + 1]); \
        fl[2 * MU] = MU == 3 && anti && c[3] == a.g.L[3] - 1;                                                    \
        fl[2 * MU + 1] = MU == 3 && anti && c[3] == 0;                                                           \
    }
            LYM_MFMA_LOAD(0)
            LYM_MFMA_LOAD(1)
            LYM_MFMA_LOAD(2)
            LYM_MFMA_LOAD(3)
#undef LYM_MFMA_LOAD
            fv4 G0 = {0.f, 0.f, 0.f, 0.f}, G1 = G0;
            mfma_acc<0>(L, Ah[0], Bh[0], sch[0], fl[0], G0);
            mfma_acc<0>(L, Ah[1], Bh[1], sch[1], fl[1], G0);
            mfma_acc<1>(L, Ah[2], Bh[2], sch[2], fl[2], G0);
            mfma_acc<1>(L, Ah[3], Bh[3], sch[3], fl[3], G0);
            mfma_acc<2>(L, Ah[4], Bh[4], sch[4], fl[4], G1);
            mfma_acc<2>(L, Ah[5], Bh[5], sch[5], fl[5], G1);
            mfma_acc<3>(L, Ah[6], Bh[6], sch[6], fl[6], G1);
            mfma_acc<3>(L, Ah[7], Bh[7], sch[7], fl[7], G1);
            // directions 0, 1 send the partner half of column h to spin 3 - h,
            // directions 2, 3 to spin 2 + h: the lane of column h = 0 collects
            // spin 2 (its G1, its neighbour's G0), the lane of h = 1 spin 3
            float other[4];
#pragma unroll
            for (int t = 0; t < 4; ++t) other[t] = __shfl_xor(G0[t], 1, 64);
            if (L.bval) {
                float* slot = accs + (ls * nr + L.rr) * MFMA_LDS_STRIDE;
#pragma unroll
                for (int t = 0; t < 4; t += 2) {
                    const int m = 4 * L.kgrp + t;  // C rows m (re), m + 1 (im)
                    float re, im;
                    int comp;
                    if (m < 6) {
                        re = G0[t] + G1[t];
                        im = G0[t + 1] + G1[t + 1];
                        comp = 3 * L.hh + (m >> 1);
                    } else {
                        re = G1[t] + other[t];
                        im = G1[t + 1] + other[t + 1];
                        comp = 3 * (2 + L.hh) + ((m - 6) >> 1);
                    }
                    slot[2 * comp] = re;
                    slot[2 * comp + 1] = im;
                }
            }
        }
        __syncthreads();
        {
            const int t = (int)threadIdx.x;
            const int ls = t / nr, r = t - ls * nr;
            const int site = bi * spb + ls;
            if (ls < spb && site < a.g.Vh && !(a.lane_flags && a.lane_flags[r] == a.lane_skip_value)) {
                cplx<float> acc[12];
                const float* slot = accs + t * MFMA_LDS_STRIDE;
#pragma unroll
                for (int k = 0; k < 12; ++k) acc[k] = mk(slot[2 * k], slot[2 * k + 1]);
                (void)dslash_epilogue<2, DIAG, INV, false, LO>(a, p, site, r, acc);
            }
        }
        __syncthreads();
    }
}

template <int PR, int LK, int DIAG, bool INV>
static void launch_one(const DslashParams& a, int nblocks, hipStream_t st) {
    if constexpr (PR == 2 && LK == 0) {
        // matrix-core form: FP16 SU(3) batches of 2, 4 or 8 RHS in whole waves
        if (a.mfma && !a.bmode && (a.nr == 2 || a.nr == 4 || a.nr == 8) && (a.spb * a.nr) % 64 == 0) {
            if (a.lane_ops)
                hipLaunchKernelGGL((k_dslash_mfma<DIAG, INV, true>), dim3(nblocks), dim3(a.spb * a.nr), 0, st, a);
            else
                hipLaunchKernelGGL((k_dslash_mfma<DIAG, INV, false>), dim3(nblocks), dim3(a.spb * a.nr), 0, st, a);
            return;
        }
    }
    if (a.lane_ops) {  // per-lane operators: the serving engine's batches (never split, no output norms)
        if constexpr (PR >= 1 && DIAG >= 1 && !INV) {
            // with the batched solver's fused steps
            if (a.bmode == 1) {
                hipLaunchKernelGGL((k_dslash<PR, LK, DIAG, INV, false, true, false, 1>), dim3(nblocks),
                                   dim3(a.spb * a.nr), 0, st, a);
                return;
            }
            if (a.bmode == 2) {
                hipLaunchKernelGGL((k_dslash<PR, LK, DIAG, INV, false, true, false, 2>), dim3(nblocks),
                                   dim3(a.spb * a.nr), 0, st, a);
                return;
            }
        }
        hipLaunchKernelGGL((k_dslash<PR, LK, DIAG, INV, false, true>), dim3(nblocks), dim3(a.spb * a.nr), 0, st, a);
        return;
    }
    if constexpr (PR >= 1 && DIAG >= 1 && !INV) {
        // the batched solver's fused steps (one thread per (site, RHS), never split)
        if (a.bmode == 1) {
            hipLaunchKernelGGL((k_dslash<PR, LK, DIAG, INV, false, false, false, 1>), dim3(nblocks),
                               dim3(a.spb * a.nr), 0, st, a);
            return;
        }
        if (a.bmode == 2) {
            hipLaunchKernelGGL((k_dslash<PR, LK, DIAG, INV, false, false, false, 2>), dim3(nblocks),
                               dim3(a.spb * a.nr), 0, st, a);
            return;
        }
    }
    // the fused mixed-CG update and output norms: asked for only on launches
    // with a diagonal and no R^-1 (the last launch of a Schur apply, a
    // full-lattice apply)
    if constexpr (PR >= 1 && DIAG >= 1 && !INV) {
        if (a.upd_state) {
            if (a.split)
                hipLaunchKernelGGL((k_dslash_split<PR, LK, DIAG, INV, false, true>), dim3(nblocks), dim3(2 * a.split),
                                   0, st, a);
            else
                hipLaunchKernelGGL((k_dslash<PR, LK, DIAG, INV, false, false, true>), dim3(nblocks),
                                   dim3(a.spb * a.nr), 0, st, a);
            return;
        }
        if (a.norm_partials) {
            if (a.split)
                hipLaunchKernelGGL((k_dslash_split<PR, LK, DIAG, INV, true>), dim3(nblocks), dim3(2 * a.split), 0, st,
                                   a);
            else
                hipLaunchKernelGGL((k_dslash<PR, LK, DIAG, INV, true>), dim3(nblocks), dim3(a.spb * a.nr), 0, st, a);
            return;
        }
    }
    if constexpr (PR >= 1) {
        if (a.split) {
            hipLaunchKernelGGL((k_dslash_split<PR, LK, DIAG, INV>), dim3(nblocks), dim3(2 * a.split), 0, st, a);
            return;
        }
    }
#if LYM_STAGE_ZT
    if constexpr (PR == 0) {
        if (a.nr == 1 && a.spb == LYM_DSLASH_THREADS && !a.lane_flags) {
            hipLaunchKernelGGL((k_dslash_stg<LK, DIAG, INV>), dim3(nblocks), dim3(LYM_DSLASH_THREADS), 0, st, a);
            return;
        }
    }
#endif
    hipLaunchKernelGGL((k_dslash<PR, LK, DIAG, INV>), dim3(nblocks), dim3(a.spb * a.nr), 0, st, a);
}

template <int PR, int LK>
static void launch_lk(const DslashParams& a, int diag, bool inv, int nblocks, hipStream_t st) {
    if (diag == 0) {
        if (inv) launch_one<PR, LK, 0, true>(a, nblocks, st);
        else launch_one<PR, LK, 0, false>(a, nblocks, st);
    } else if (diag == 1) {
        if (inv) launch_one<PR, LK, 1, true>(a, nblocks, st);
        else launch_one<PR, LK, 1, false>(a, nblocks, st);
    } else {
        launch_one<PR, LK, 2, false>(a, nblocks, st);
    }
}

template <int PR>
static void launch_pr(const DslashParams& a, int link_kind, int diag, bool inv, int nblocks, hipStream_t st) {
    switch (link_kind) {
        case 0: launch_lk<PR, 0>(a, diag, inv, nblocks, st); break;
        case 1: launch_lk<PR, 1>(a, diag, inv, nblocks, st); break;
        case 2: launch_lk<PR, 2>(a, diag, inv, nblocks, st); break;
        default: launch_lk<PR, 3>(a, diag, inv, nblocks, st); break;
    }
}

// Launch geometry: sites per workgroup, the split form, site blocks per parity
// and per x-slice; returns the number of workgroups.
static int plan_dslash(DslashParams& a) {
    if (a.nr < 1) a.nr = 1;
    const int threads = (a.threads > 0 && a.threads <= LYM_DSLASH_THREADS) ? a.threads : LYM_DSLASH_THREADS;
    a.spb = threads / a.nr;
    if (a.spb < 1) a.spb = 1;
    // two threads per site: low-precision single-RHS applies that allow it
    a.split = (LYM_SPLIT_LOWP && a.allow_split && !a.lane_ops && !a.bmode && a.prec >= 1 && threads % 64 == 0 &&
               (long)a.g.Vh * a.nr <= LYM_SPLIT_MAX_SITES)
                  ? threads
                  : 0;
    a.nblocks_per_parity = (a.g.Vh + a.spb - 1) / a.spb;
    {
        const int xs = a.g.Vh / a.g.L[0];  // sites of one parity per x-slice
        a.xslice_blocks = (xs % a.spb == 0) ? xs / a.spb : 0;
    }
    int nblocks = a.nblocks_per_parity * (a.parity >= 0 ? 1 : 2);
    // workgroups in flight: LYM_GRID_PER_CU per CU of the 256 CUs, sweeping
    const int gpc = LYM_GRID_PER_CU >= 0 ? LYM_GRID_PER_CU : (nblocks <= LYM_SWEEP_MAX_BLOCKS ? 1 : 0);
    if (gpc > 0 && nblocks > 256 * gpc) nblocks = 256 * gpc;
    return nblocks;
}

int launch_dslash(DslashParams a, int link_kind, int diag, bool inv, hipStream_t st) {
    const int threads = (a.threads > 0 && a.threads <= LYM_DSLASH_THREADS) ? a.threads : LYM_DSLASH_THREADS;
    const int nblocks = plan_dslash(a);
    // output norm partials: low precision only, one per workgroup, and only
    // when the caller's buffer holds one per workgroup of this launch
    // (the workgroup sum shuffles over whole 64-lane waves)
    if (a.prec == 0 || diag == 0 || inv || a.nr != 1 || a.lane_flags || threads % 64 != 0 ||
        nblocks > a.norm_max)
        a.norm_partials = nullptr;
    // the fused update: under the same conditions (the caller checks the return
    // value and runs the separate update when it is 0)
    if (a.prec == 0 || diag == 0 || inv || a.nr != 1 || a.lane_flags || a.lane_ops || threads % 64 != 0 ||
        nblocks > a.norm_max || a.norm_partials)
        a.upd_state = nullptr;
    switch (a.prec) {
        case 0: launch_pr<0>(a, link_kind, diag, inv, nblocks, st); break;
        case 1: launch_pr<1>(a, link_kind, diag, inv, nblocks, st); break;
        default: launch_pr<2>(a, link_kind, diag, inv, nblocks, st); break;
    }
    return (a.norm_partials || a.upd_state) ? nblocks : 0;
}

}  // namespace lym

// ---------------------------------------------------------------- launch plan (host)
#include <cstring>

#include "../../include/lym.h"

namespace lym_internal {
int fail(int code, const char* fmt, ...);
}

extern "C" int lym_launch_plan(const int dims[4], int parity, int nr, int threads, int lowp, int* grid, int* spb,
                               int* nbpp, int* xslice_blocks, int* visits, long visits_len, int* branches) {
    using lym_internal::fail;
    if (!dims || !grid || !spb || !nbpp || !xslice_blocks) return fail(LYM_ERR_INVALID, "null argument");
    for (int i = 0; i < 4; ++i)
        if (dims[i] < 2) return fail(LYM_ERR_INVALID, "lattice extent %d must be >= 2", dims[i]);
    // the context's layout: even-odd when all extents are even, else one class
    // (whose launches are parity 0)
    const int tsh = (dims[0] | dims[1] | dims[2] | dims[3]) % 2 == 0 ? 1 : 0;
    if (parity < -1 || parity > 1 || nr < 1 || nr > 16 || threads < 0 || threads > LYM_DSLASH_THREADS ||
        (!tsh && parity != 0))
        return fail(LYM_ERR_INVALID, "bad launch arguments");
    const long V = (long)dims[0] * dims[1] * dims[2] * dims[3];
    if ((V >> tsh) > (1L << 27) - 64) return fail(LYM_ERR_INVALID, "lattice too large");
    lym::DslashParams a;
    std::memset(&a, 0, sizeof(a));
    for (int i = 0; i < 4; ++i) a.g.L[i] = dims[i];
    a.g.tsh = tsh;
    a.g.Lth = dims[3] >> tsh;
    a.g.Lxh = dims[0] >> tsh;
    a.g.Vh = (int)(V >> tsh);
    a.g.stride = (int)(((V >> tsh) + 63) / 64 * 64);
    a.g.bc_t = -1;
    a.nr = nr;
    a.threads = threads;
    a.parity = parity;
    a.prec = lowp ? 2 : 0;
    a.allow_split = lowp ? 1 : 0;
    const int gdim = lym::plan_dslash(a);
    *grid = gdim;
    *spb = a.spb;
    *nbpp = a.nblocks_per_parity;
    *xslice_blocks = a.xslice_blocks;
    if (visits && visits_len < 2L * a.nblocks_per_parity) return fail(LYM_ERR_INVALID, "visits array too short");
    if (visits) std::memset(visits, 0, sizeof(int) * 2 * (size_t)a.nblocks_per_parity);
    if (branches) branches[0] = branches[1] = branches[2] = 0;
    for (int bx = 0; bx < gdim; ++bx) {
        const lym::BlockSweep b = lym::block_sweep_at(a, bx, gdim);
        for (int lb = b.lo + b.k; lb < b.hi; lb += b.nphys) {
            int p = 0, bi = 0;
            const int br = lym::block_of_at(a, b, lb, bx, p, bi);
            if (p < 0 || p > 1 || bi < 0 || bi >= a.nblocks_per_parity)
                return fail(LYM_ERR_INVALID, "workgroup %d maps block %d to (%d, %d), outside the lattice", bx, lb, p,
                            bi);
            if (visits) ++visits[p * a.nblocks_per_parity + bi];
            if (branches) ++branches[br];
        }
    }
    return LYM_OK;
}
