// measure.cpp -- point-source propagators and connected meson correlators
// (MesonCorrelator::execute, source/correlators/MesonCorrelator.cpp:22-200),
// on top of the operator and solver objects of the C ABI.
//
// The 12 spin-colour point sources (StochasticEstimator::generateSource,
// source/fermion_measurements/StochasticEstimator.cpp:149-161) are solved
// with the caller's solver and the operator's D (gamma5 off, as the reference
// sets it, MesonCorrelator.cpp:71); a mixed-precision CG solver takes all 12
// in one batched call (lym_solver_solve_batch), any other kind one at a time.
// The contractions run on the device (correlators.hip).
#include <hip/hip_runtime.h>

#include <cstring>
#include <vector>

#include "../../include/lym.h"
#include "lym_kernels.h"

using lym::zd;

namespace lym {
void launch_meson_contractions(const Geom& g, const zd* const* prop_dev, double* site_vals, double* res,
                               hipStream_t st);
void launch_set_one(zd* f, long idx, hipStream_t st);
}  // namespace lym

namespace lym_internal {
int fail(int code, const char* fmt, ...);
int solver_kind(const lym_solver* s);  // solvers.cpp
int solver_even_odd(const lym_solver* s);
}  // namespace lym_internal
using lym_internal::fail;

namespace {

struct Cleanup {
    std::vector<lym_spinor*> spinors;
    std::vector<void*> dev;
    lym_op* op = nullptr;
    int gamma5 = 1;
    ~Cleanup() {
        if (op) lym_op_set_gamma5(op, gamma5);
        for (lym_spinor* s : spinors) lym_spinor_destroy(s);
        for (void* p : dev) (void)hipFree(p);
    }
};

#define MCHECK(expr)                                                                               \
    do {                                                                                           \
        hipError_t e_ = (expr);                                                                    \
        if (e_ != hipSuccess) return fail(LYM_ERR_DEVICE, "%s: %s", #expr, hipGetErrorString(e_)); \
    } while (0)

}  // namespace

extern "C" int lym_meson_correlators(lym_solver* solver, lym_op* op, const int source_site[4], double* out,
                                     lym_solve_info* infos) {
    if (!solver || !op || !source_site || !out) return fail(LYM_ERR_INVALID, "null argument");
    lym_ctx* c = op->ctx;
    const lym::Geom& g = c->g;
    for (int m = 0; m < 4; ++m)
        if (source_site[m] < 0 || source_site[m] >= c->dims[m]) return fail(LYM_ERR_INVALID, "source site outside the lattice");
    if (op->power != 1) return fail(LYM_ERR_INVALID, "the propagator needs the operator to the first power");
    double kappa = 0;
    int kind = 0;
    lym_op_get_params(op, &kappa, nullptr, nullptr, &kind, nullptr);
    Cleanup cl;
    // D, not gamma5 D (restored on return)
    cl.gamma5 = op->gamma5;
    cl.op = op;
    lym_op_set_gamma5(op, 0);
    // the 12 sources and solutions
    std::vector<lym_spinor*> src(12), sol(12);
    for (int k = 0; k < 12; ++k) {
        int rc = lym_spinor_create(c, &src[k]);
        if (!rc) cl.spinors.push_back(src[k]);
        if (!rc) rc = lym_spinor_create(c, &sol[k]);
        if (!rc) cl.spinors.push_back(sol[k]);
        if (rc) return rc;
    }
    const int* s = source_site;
    const long lex = (((long)s[0] * c->dims[1] + s[1]) * c->dims[2] + s[2]) * c->dims[3] + s[3];
    const int par = (s[0] + s[1] + s[2] + s[3]) & g.tsh;
    const long eo = lex >> g.tsh;
    for (int al = 0; al < 4; ++al)
        for (int col = 0; col < 3; ++col) {
            // column c * 4 + alpha holds the source (alpha, c) (MesonCorrelator.cpp:94-98)
            lym_spinor* f = src[col * 4 + al];
            int rc = lym_spinor_set_zero(f);
            if (rc) return rc;
            lym::launch_set_one(reinterpret_cast<zd*>(f->d), ((long)par * 12 + 3 * al + col) * g.stride + eo, c->stream);
        }
    MCHECK(hipGetLastError());
    std::vector<lym_solve_info> local(12);
    lym_solve_info* inf = infos ? infos : local.data();
    // the solver kind decides between one batched call and 12 solves; it must
    // return D^-1 b: BiCGStab on D, or an even-odd CG driver (clover operator,
    // even extents)
    const int kindS = lym_internal::solver_kind(solver);
    const bool even = c->dims[0] % 2 == 0 && c->dims[1] % 2 == 0 && c->dims[2] % 2 == 0 && c->dims[3] % 2 == 0;
    const bool eo_driver = lym_internal::solver_even_odd(solver) && even &&
                           (kind == LYM_OP_CLOVER || kind == LYM_OP_EO_CLOVER);
    const bool ok = kindS == LYM_SOLVER_BICGSTAB || kindS == LYM_SOLVER_PRECOND_BICGSTAB ||
                    ((kindS == LYM_SOLVER_PRECOND_CG || kindS == LYM_SOLVER_PRECOND_MIXED_CG) && eo_driver);
    if (!ok) return fail(LYM_ERR_INVALID, "solver kind %d does not return D^-1 b for this operator", kindS);
    int rc = LYM_OK;
    if (kindS == LYM_SOLVER_MIXED_CG || kindS == LYM_SOLVER_PRECOND_MIXED_CG) {
        rc = lym_solver_solve_batch(solver, op, 12, src.data(), sol.data(), nullptr, inf);
    } else {
        for (int k = 0; !rc && k < 12; ++k) rc = lym_solver_solve(solver, op, src[k], sol[k], nullptr, &inf[k]);
    }
    if (rc) return rc;
    // contractions on the device
    zd** pdev = nullptr;
    double *site_vals = nullptr, *res = nullptr;
    MCHECK(hipMalloc((void**)&pdev, 12 * sizeof(zd*)));
    cl.dev.push_back(pdev);
    MCHECK(hipMalloc((void**)&site_vals, 3 * 2 * (size_t)g.stride * sizeof(double)));
    cl.dev.push_back(site_vals);
    MCHECK(hipMalloc((void**)&res, 3 * (size_t)g.L[3] * sizeof(double)));
    cl.dev.push_back(res);
    std::vector<zd*> ph(12);
    for (int k = 0; k < 12; ++k) ph[k] = reinterpret_cast<zd*>(sol[k]->d);
    MCHECK(hipMemcpyAsync(pdev, ph.data(), 12 * sizeof(zd*), hipMemcpyHostToDevice, c->stream));
    lym::launch_meson_contractions(g, pdev, site_vals, res, c->stream);
    MCHECK(hipGetLastError());
    const int T = g.L[3];
    std::vector<double> h(3 * (size_t)T);
    MCHECK(hipMemcpyAsync(h.data(), res, h.size() * sizeof(double), hipMemcpyDeviceToHost, c->stream));
    MCHECK(hipStreamSynchronize(c->stream));
    // outputs relative to the source timeslice, as the reference's source at
    // t = 0 of links shifted by t_source_origin (MesonCorrelator.cpp:54-66);
    // normalisation 4 kappa^2 (:162), PCAC ratio of the raw sums (:189-190)
    const double factor = 4.0 * kappa * kappa;
    const int t0 = s[3];
    for (int t = 0; t < T; ++t) {
        const int a = (t0 + t) % T;
        out[t] = factor * h[a];
        out[T + t] = factor * h[T + a];
        out[2 * T + t] = factor * h[2 * T + a];
    }
    for (int t = 0; t < T; ++t) {
        const int am = (t0 + (t == 0 ? T - 1 : t - 1)) % T, ap = (t0 + (t == T - 1 ? 0 : t + 1)) % T;
        out[3 * T + t] = (h[2 * T + am] - h[2 * T + ap]) / (4.0 * h[(t0 + t) % T]);
    }
    return LYM_OK;
}
