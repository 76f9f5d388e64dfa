// force.hip -- the two-flavour Wilson fermion force on the device:
// TwoFlavorFermionAction::force / updateForce
// (source/actions/TwoFlavorFermionAction.cpp:17-43), the derivative of
// DiracWilsonFermionForce (source/hmc_forces/DiracWilsonFermionForce.cpp:10-54)
// projected on the generators by FermionForce::force
// (source/hmc_forces/FermionForce.cpp:33-58).
//
// One thread per link (lexicographic site, mu).  For the solutions
// X = Q^-2 phi, Y = Q^-1 phi, with W the fermion link:
//   D = sum_alpha A(x)_alpha^+ ... written as outer products
//     D += t_alpha A(x)_alpha^+            t = -kappa g5(1 - g_mu) B(x+mu)
//     D -= t'_alpha (W A(x+mu)_alpha)^+    t' = -kappa g5(1 + g_mu) W^+ B(x)
//   over (A, B) = (X, Y) and (Y, X).  The second term is quadratic in W, so
//   D does not depend on the boundary sign of W.
// and the force F = i sum_a Im tr(D T^f_a L) T^g_a with L the link of the
// fermion representation without the boundary sign (the device stores links
// without it): for SU(3) fundamental fermions, by completeness of the
// generators (tr T_a T_b = delta_ab / 2),  F = (M - M^+ - tr(M - M^+)/3) / 4
// with M = L D; for SU(2) adjoint fermions (T^adj_a)_{bc} = -i eps_{abc},
// F = i sum_a c_a sigma_a / 2 with c_a = -Re sum_{bc} eps_{abc} (L D)_{cb}.
#include <algorithm>
#include <cstring>
#include <mutex>

#include "../../include/lym.h"
#include "lym_device.h"

namespace lym_internal {
int fail(int code, const char* fmt, ...);
int op_ready(lym_op* op, bool force_blocks);
int solver_kind(const lym_solver* s);
}  // namespace lym_internal
using lym_internal::fail;

namespace lym {
namespace {

// gamma5 (1 - s gamma_mu) applied to spinor b at spin row alpha:
// sum_beta [g5 (1 - s g_mu)]_{alpha beta} b_beta, with the chiral gammas
// gamma_mu[a][P_a] = i^C (projection partners of dslash.hip's Dir<mu>)
template <int MU>
struct GDir;
// gamma_mu nonzero entries: row a -> column P[a], value i^C[a]
template <>
struct GDir<0> { static constexpr int P[4] = {3, 2, 1, 0}; static constexpr int C[4] = {3, 3, 1, 1}; };
template <>
struct GDir<1> { static constexpr int P[4] = {3, 2, 1, 0}; static constexpr int C[4] = {2, 0, 0, 2}; };
template <>
struct GDir<2> { static constexpr int P[4] = {2, 3, 0, 1}; static constexpr int C[4] = {3, 1, 1, 3}; };
template <>
struct GDir<3> { static constexpr int P[4] = {2, 3, 0, 1}; static constexpr int C[4] = {0, 0, 0, 0}; };

__device__ __forceinline__ zd phase(int c, zd z) {
    switch (c & 3) {
        case 0: return z;
        case 1: return mk(-z.im, z.re);
        case 2: return mk(-z.re, -z.im);
        default: return mk(z.im, -z.re);
    }
}

// out[alpha][c] = -kappa sum_beta [g5 (1 + sgn gamma_mu)]_{alpha beta} b[beta][c]
template <int MU>
__device__ __forceinline__ void proj(const zd b[12], double sgn, double kappa, zd out[12]) {
#pragma unroll
    for (int al = 0; al < 4; ++al) {
        const double s5 = al < 2 ? -kappa : kappa;  // -kappa * gamma5 row sign
        const int P = GDir<MU>::P[al], C = GDir<MU>::C[al];
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            const zd g = phase(C, b[3 * P + c]);
            out[3 * al + c] = s5 * (b[3 * al + c] + sgn * g);
        }
    }
}

struct ForceParams {
    Geom g;
    const zd* X;  // [2][12][stride]
    const zd* Y;
    const void* links;
    int link_kind;
    double kappa;
    long s0, ns;  // the lexicographic sites [s0, s0 + ns) of this launch
    zd* out;      // [ns][4][N*N] for those sites, column-major per link
};

__device__ __forceinline__ void load_sp(const zd* f, const Geom& g, int p, int i, zd s[12]) {
#pragma unroll
    for (int k = 0; k < 12; ++k) s[k] = f[((size_t)p * 12 + k) * g.stride + i];
}

template <int MU, int LK>
__device__ void force_link(const ForceParams& a, const int c[4], long lex) {
    const Geom& g = a.g;
    const int p = site_parity(g, c);
    const int i = eo_index(g, c[0], c[1], c[2], c[3]);
    int cu[4] = {c[0], c[1], c[2], c[3]};
    cu[MU] = c[MU] + 1 == g.L[MU] ? 0 : c[MU] + 1;
    const int iu = eo_index(g, cu[0], cu[1], cu[2], cu[3]);
    LinkMat<LK, double> L;
    load_link<0, LK>(a.links, g.stride, p, MU, i, L);
    // complex link entries W(r, s) = L.m[3 r + s]
    zd W[9];
#pragma unroll
    for (int k = 0; k < 9; ++k) {
        if constexpr (LK <= 1) W[k] = L.m[k];
        else W[k] = mk((double)L.m[k], 0.0);
    }
    zd D[9];
#pragma unroll
    for (int k = 0; k < 9; ++k) D[k] = mk(0.0, 0.0);
#pragma unroll
    for (int ord = 0; ord < 2; ++ord) {
        const zd* Af = ord == 0 ? a.X : a.Y;
        const zd* Bf = ord == 0 ? a.Y : a.X;
        zd A0[12], Au[12], B0[12], Bu[12];
        load_sp(Af, g, p, i, A0);
        load_sp(Af, g, p ^ g.tsh, iu, Au);
        load_sp(Bf, g, p, i, B0);
        load_sp(Bf, g, p ^ g.tsh, iu, Bu);
        // t = -kappa g5 (1 - g_mu) B(x+mu)
        zd t[12];
        proj<MU>(Bu, -1.0, a.kappa, t);
        // W^+ B(x), then t' = -kappa g5 (1 + g_mu) W^+ B(x)
        zd wb[12];
#pragma unroll
        for (int be = 0; be < 4; ++be)
#pragma unroll
            for (int r = 0; r < 3; ++r) {
                zd s = cmulc(W[r], B0[3 * be]);
                cfmac(s, W[3 + r], B0[3 * be + 1]);
                cfmac(s, W[6 + r], B0[3 * be + 2]);
                wb[3 * be + r] = s;
            }
        zd tq[12];
        proj<MU>(wb, 1.0, a.kappa, tq);
        // W A(x+mu)
        zd wa[12];
#pragma unroll
        for (int al = 0; al < 4; ++al)
#pragma unroll
            for (int r = 0; r < 3; ++r) {
                zd s = W[3 * r] * Au[3 * al];
                cfma(s, W[3 * r + 1], Au[3 * al + 1]);
                cfma(s, W[3 * r + 2], Au[3 * al + 2]);
                wa[3 * al + r] = s;
            }
        // D(r, s) += sum_alpha t_r conj(A0_s) - t'_r conj(wa_s)
#pragma unroll
        for (int al = 0; al < 4; ++al)
#pragma unroll
            for (int r = 0; r < 3; ++r)
#pragma unroll
                for (int s = 0; s < 3; ++s) {
                    zd& d = D[3 * r + s];
                    d = d + t[3 * al + r] * conj(A0[3 * al + s]) - tq[3 * al + r] * conj(wa[3 * al + s]);
                }
    }
    // M = L D with L the link without boundary sign (= W as stored)
    zd M[9];
#pragma unroll
    for (int r = 0; r < 3; ++r)
#pragma unroll
        for (int s = 0; s < 3; ++s) {
            zd v = W[3 * r] * D[s];
            cfma(v, W[3 * r + 1], D[3 + s]);
            cfma(v, W[3 * r + 2], D[6 + s]);
            M[3 * r + s] = v;
        }
    zd* o = a.out + ((size_t)(lex - a.s0) * 4 + MU) * (LK <= 1 ? 9 : 4);
    if constexpr (LK <= 1) {
        // F = (A - tr(A)/3) / 4, A = M - M^+
        zd A[9];
#pragma unroll
        for (int r = 0; r < 3; ++r)
#pragma unroll
            for (int s = 0; s < 3; ++s) A[3 * r + s] = M[3 * r + s] - conj(M[3 * s + r]);
        const zd tr = (1.0 / 3.0) * (A[0] + A[4] + A[8]);
#pragma unroll
        for (int r = 0; r < 3; ++r)
#pragma unroll
            for (int s = 0; s < 3; ++s) {
                zd v = A[3 * r + s];
                if (r == s) v = v - tr;
                o[r + 3 * s] = 0.25 * v;  // column-major
            }
    } else {
        // c_a = -Re sum_{bc} eps_{abc} M(c, b):  eps_{abc} M_cb for (a,b,c) cyclic minus anticyclic
        const double c0 = -(M[3 * 2 + 1].re - M[3 * 1 + 2].re);
        const double c1 = -(M[3 * 0 + 2].re - M[3 * 2 + 0].re);
        const double c2 = -(M[3 * 1 + 0].re - M[3 * 0 + 1].re);
        // F = i (c0 sx + c1 sy + c2 sz) / 2
        o[0] = mk(0.0, 0.5 * c2);             // (0,0)
        o[1] = mk(-0.5 * c1, 0.5 * c0);       // (1,0): i (c0 + i c1) / 2
        o[2] = mk(0.5 * c1, 0.5 * c0);        // (0,1): i (c0 - i c1) / 2
        o[3] = mk(0.0, -0.5 * c2);            // (1,1)
    }
}

template <int LK>
__global__ __launch_bounds__(128) void k_fermion_force(ForceParams a) {
    const long j = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= 4 * a.ns) return;
    const int mu = (int)(j / a.ns);
    const long lex = a.s0 + (j - (long)mu * a.ns);
    int c[4];
    {
        long r = lex;
        c[3] = (int)(r % a.g.L[3]);
        r /= a.g.L[3];
        c[2] = (int)(r % a.g.L[2]);
        r /= a.g.L[2];
        c[1] = (int)(r % a.g.L[1]);
        c[0] = (int)(r / a.g.L[1]);
    }
    switch (mu) {
        case 0: force_link<0, LK>(a, c, lex); break;
        case 1: force_link<1, LK>(a, c, lex); break;
        case 2: force_link<2, LK>(a, c, lex); break;
        default: force_link<3, LK>(a, c, lex); break;
    }
}

}  // namespace
}  // namespace lym

using lym::zd;

namespace {

// checks of lym_fermion_force's arguments; fills everything of the launch
// parameters but the site range and the output
int force_params(lym_op* op, const lym_spinor* X, const lym_spinor* Y, const double* force, lym::ForceParams& a,
                 int& nn) {
    if (!op || !X || !Y || !force) return fail(LYM_ERR_INVALID, "null argument");
    lym_ctx* c = op->ctx;
    if (X->ctx != c || Y->ctx != c) return fail(LYM_ERR_INVALID, "objects of different contexts");
    if (op->kind != LYM_OP_WILSON && op->kind != LYM_OP_BASIC_WILSON)
        return fail(LYM_ERR_UNSUPPORTED, "the fermion force is DiracWilsonFermionForce's: Wilson operators only");
    if (!op->gauge) return fail(LYM_ERR_INVALID, "operator has no gauge field (call lym_op_set_gauge)");
    const int N = c->rep == LYM_REP_SU3_FUND ? 3 : 2;
    nn = N * N;
    a.g = c->g;
    a.X = reinterpret_cast<const zd*>(X->d);
    a.Y = reinterpret_cast<const zd*>(Y->d);
    a.links = op->gauge->U;
    a.kappa = op->kappa;
    const lym_gauge* gg = op->gauge;
    a.link_kind = gg->real_links ? (gg->compression == 6 ? 3 : 2) : (gg->compression == 12 ? 1 : 0);
    a.s0 = 0;
    a.ns = 0;
    a.out = nullptr;
    return LYM_OK;
}

void launch_force(const lym::ForceParams& a, hipStream_t st) {
    const unsigned grid = (unsigned)((4 * a.ns + 127) / 128);
    switch (a.link_kind) {
        case 0: hipLaunchKernelGGL(lym::k_fermion_force<0>, dim3(grid), dim3(128), 0, st, a); break;
        case 1: hipLaunchKernelGGL(lym::k_fermion_force<1>, dim3(grid), dim3(128), 0, st, a); break;
        case 2: hipLaunchKernelGGL(lym::k_fermion_force<2>, dim3(grid), dim3(128), 0, st, a); break;
        default: hipLaunchKernelGGL(lym::k_fermion_force<3>, dim3(grid), dim3(128), 0, st, a); break;
    }
}

}  // namespace

extern "C" int lym_fermion_force(lym_op* op, const lym_spinor* X, const lym_spinor* Y, double* force) {
    lym::ForceParams a;
    int nn = 0;
    int rc = force_params(op, X, Y, force, a, nn);
    if (rc) return rc;
    lym_ctx* c = op->ctx;
    const long V = c->V;
    const size_t bytes = sizeof(zd) * (size_t)V * 4 * nn;
    zd* dev = nullptr;
    if (hipMalloc((void**)&dev, bytes) != hipSuccess) return fail(LYM_ERR_NOMEM, "hipMalloc(%zu bytes)", bytes);
    a.ns = V;
    a.out = dev;
    launch_force(a, c->stream);
    hipError_t e = hipGetLastError();
    if (e == hipSuccess) e = hipMemcpyAsync(force, dev, bytes, hipMemcpyDeviceToHost, c->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
    (void)hipFree(dev);
    if (e != hipSuccess) return fail(LYM_ERR_DEVICE, "fermion force: %s", hipGetErrorString(e));
    return LYM_OK;
}

// ---------------------------------------------------------------- without allocation
// The force of lym_fermion_force_noalloc goes through this fixed device array
// in chunks of whole sites (LYM_FORCE_SCRATCH_BYTES: 16 MiB, 29127 SU(3) or
// 65536 SU(2) sites per chunk).  Every GPU the library runs on holds its own
// copy until the process ends.  In host code `scratch` names a host-side
// stand-in, never a device address: its device address on each GPU comes from
// hipGetSymbolAddress, once per device, and only that address is passed to
// kernels and copies.
#define LYM_FORCE_SCRATCH_BYTES (16u << 20)
static __device__ __align__(256) zd scratch[LYM_FORCE_SCRATCH_BYTES / sizeof(zd)];

namespace {

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
zd* g_scratch_ptr[kMaxDevices] = {};

// the device address of `scratch` on GPU `dev`, looked up with that GPU current
int scratch_on(int dev, zd** out) {
    if (dev < 0 || dev >= kMaxDevices) return fail(LYM_ERR_DEVICE, "device %d beyond the %d supported", dev, kMaxDevices);
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (!g_scratch_ptr[dev]) {
        int cur = -1;
        hipError_t e = hipGetDevice(&cur);
        if (e == hipSuccess && cur != dev) e = hipSetDevice(dev);
        void* p = nullptr;
        if (e == hipSuccess) e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (cur >= 0 && cur != dev) (void)hipSetDevice(cur);
        if (e != hipSuccess || !p)
            return fail(LYM_ERR_DEVICE, "hipGetSymbolAddress(scratch) on device %d: %s", dev, hipGetErrorString(e));
        g_scratch_ptr[dev] = static_cast<zd*>(p);
    }
    *out = g_scratch_ptr[dev];
    return LYM_OK;
}

}  // namespace

extern "C" int lym_fermion_force_noalloc(lym_op* op, const lym_spinor* X, const lym_spinor* Y, double* force) {
    lym::ForceParams a;
    int nn = 0;
    int rc = force_params(op, X, Y, force, a, nn);
    if (rc) return rc;
    lym_ctx* c = op->ctx;
    zd* buf = nullptr;
    if ((rc = scratch_on(c->device, &buf))) return rc;
    const long per_site = 4L * nn;  // complex entries per site
    const long chunk = (long)(LYM_FORCE_SCRATCH_BYTES / sizeof(zd)) / per_site;
    a.out = buf;
    hipError_t e = hipSuccess;
    for (long s0 = 0; s0 < c->V && e == hipSuccess; s0 += chunk) {
        a.s0 = s0;
        a.ns = std::min(chunk, c->V - s0);
        if (a.ns * per_site * (long)sizeof(zd) > (long)LYM_FORCE_SCRATCH_BYTES)
            return fail(LYM_ERR_INVALID, "fermion force chunk of %ld sites exceeds the %u-byte scratch", a.ns,
                        LYM_FORCE_SCRATCH_BYTES);
        launch_force(a, c->stream);
        e = hipGetLastError();
        // stream order keeps the next chunk's kernel behind this copy
        if (e == hipSuccess)
            e = hipMemcpyAsync(force + 2 * s0 * per_site, buf, sizeof(zd) * a.ns * per_site, hipMemcpyDeviceToHost,
                               c->stream);
    }
    if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
    if (e != hipSuccess) return fail(LYM_ERR_DEVICE, "fermion force: %s", hipGetErrorString(e));
    return LYM_OK;
}

extern "C" int lym_two_flavor_force(lym_solver* s, lym_op* op, const lym_spinor* phi, double* force,
                                    lym_solve_info* infos) {
    if (!s || !op || !phi || !force) return fail(LYM_ERR_INVALID, "null argument");
    const int kind = lym_internal::solver_kind(s);
    if (kind != LYM_SOLVER_BICGSTAB && kind != LYM_SOLVER_PRECOND_BICGSTAB)
        return fail(LYM_ERR_UNSUPPORTED, "updateForce solves Q Y = phi and Q X = Y with BiCGStab");
    if (op->power != 1 || !op->gamma5)
        return fail(LYM_ERR_INVALID, "the operator must be Q = gamma5 D (power 1, gamma5 set)");
    lym_solve_info local[2];
    if (!infos) infos = local;
    lym_spinor *X = nullptr, *Y = nullptr;
    int rc = lym_spinor_create(op->ctx, &Y);
    if (!rc) rc = lym_spinor_create(op->ctx, &X);
    if (!rc) rc = lym_solver_solve(s, op, phi, Y, nullptr, &infos[0]);
    if (!rc) rc = lym_solver_solve(s, op, Y, X, nullptr, &infos[1]);
    if (!rc) rc = lym_fermion_force(op, X, Y, force);
    lym_spinor_destroy(X);
    lym_spinor_destroy(Y);
    return rc;
}
