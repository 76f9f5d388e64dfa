// generate.cpp -- seeded synthetic inputs on the host.
//
// The reference seeds its generators from time(NULL)
// (source/utils/RandomSeed.cpp:17-28) and builds hot starts by Haar sampling
// (source/starters/HotStartGaugeConfiguration.cpp:18-65); the benchmark inputs
// here are instead U = exp(i eps H) with a counter-based generator, so every
// link and source component is a pure function of (seed, site, index) and the
// same on any machine and thread count.
#include <cmath>
#include <complex>
#include <cstdint>
#include <cstring>
#include <thread>
#include <vector>

#include "../../include/lym.h"

namespace lym_internal {
int fail(int code, const char* fmt, ...);
}
using lym_internal::fail;

namespace {

typedef std::complex<double> cd;

inline uint64_t mix(uint64_t x) {
    x += 0x9E3779B97F4A7C15ull;
    x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ull;
    x = (x ^ (x >> 27)) * 0x94D049BB133111EBull;
    return x ^ (x >> 31);
}

// uniform in (0, 1]
inline double uni(uint64_t seed, uint64_t ctr) { return ((mix(seed ^ mix(ctr)) >> 11) + 1) * 0x1.0p-53; }

// standard normal pair by Box-Muller from counters 2k, 2k+1
inline void gauss2(uint64_t seed, uint64_t k, double& a, double& b) {
    double u1 = uni(seed, 2 * k), u2 = uni(seed, 2 * k + 1);
    double r = std::sqrt(-2.0 * std::log(u1));
    a = r * std::cos(2 * M_PI * u2);
    b = r * std::sin(2 * M_PI * u2);
}

template <class F>
void parallel_for(long n, F f) {
    unsigned nt = std::thread::hardware_concurrency();
    if (nt == 0) nt = 1;
    if (nt > 32) nt = 32;
    if (n < 4096) nt = 1;
    std::vector<std::thread> th;
    long chunk = (n + nt - 1) / nt;
    for (unsigned t = 0; t < nt; ++t) {
        long lo = t * chunk, hi = std::min(n, lo + chunk);
        if (lo >= hi) break;
        th.emplace_back([=]() {
            for (long i = lo; i < hi; ++i) f(i);
        });
    }
    for (auto& t : th) t.join();
}

struct M3 {
    cd a[3][3];
};

M3 mul(const M3& A, const M3& B) {
    M3 C;
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) {
            cd s = 0;
            for (int k = 0; k < 3; ++k) s += A.a[i][k] * B.a[k][j];
            C.a[i][j] = s;
        }
    return C;
}

// exp(X) by scaling and squaring with a degree-16 Taylor polynomial
M3 expm(M3 X) {
    double nrm = 0;
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) nrm += std::norm(X.a[i][j]);
    nrm = std::sqrt(nrm);
    int s = 0;
    while (nrm > 0.25) {
        nrm *= 0.5;
        ++s;
    }
    double f = std::ldexp(1.0, -s);
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) X.a[i][j] *= f;
    M3 R, T;
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) R.a[i][j] = T.a[i][j] = (i == j) ? 1.0 : 0.0;
    for (int k = 1; k <= 16; ++k) {
        T = mul(T, X);
        for (int i = 0; i < 3; ++i)
            for (int j = 0; j < 3; ++j) {
                T.a[i][j] /= (double)k;
                R.a[i][j] += T.a[i][j];
            }
    }
    for (int k = 0; k < s; ++k) R = mul(R, R);
    return R;
}

// Gram-Schmidt on rows 0, 1 and row 2 = conj(row0 x row1): unitary, det 1
void reunitarise(M3& U) {
    cd* r0 = U.a[0];
    cd* r1 = U.a[1];
    double n0 = std::sqrt(std::norm(r0[0]) + std::norm(r0[1]) + std::norm(r0[2]));
    for (int j = 0; j < 3; ++j) r0[j] /= n0;
    cd d = std::conj(r0[0]) * r1[0] + std::conj(r0[1]) * r1[1] + std::conj(r0[2]) * r1[2];
    for (int j = 0; j < 3; ++j) r1[j] -= d * r0[j];
    double n1 = std::sqrt(std::norm(r1[0]) + std::norm(r1[1]) + std::norm(r1[2]));
    for (int j = 0; j < 3; ++j) r1[j] /= n1;
    U.a[2][0] = std::conj(r0[1] * r1[2] - r0[2] * r1[1]);
    U.a[2][1] = std::conj(r0[2] * r1[0] - r0[0] * r1[2]);
    U.a[2][2] = std::conj(r0[0] * r1[1] - r0[1] * r1[0]);
}

bool dims_ok(const int* d) {
    if (!d) return false;
    for (int i = 0; i < 4; ++i)
        if (d[i] < 1) return false;
    return true;
}

}  // namespace

extern "C" {

int lym_gen_su3_links(const int dims[4], double eps, uint64_t seed, double* links) {
    if (!dims_ok(dims) || !links) return fail(LYM_ERR_INVALID, "bad arguments");
    const long nl = 4L * dims[0] * dims[1] * dims[2] * dims[3];
    const double r3 = 1.0 / std::sqrt(3.0);
    cd* out = reinterpret_cast<cd*>(links);
    parallel_for(nl, [&](long l) {
        double w[8];
        for (int k = 0; k < 4; ++k) gauss2(seed, (uint64_t)l * 4 + k, w[2 * k], w[2 * k + 1]);
        // H = sum_a w_a lambda_a / 2 (Gell-Mann matrices)
        M3 H;
        H.a[0][0] = 0.5 * (w[2] + w[7] * r3);
        H.a[1][1] = 0.5 * (-w[2] + w[7] * r3);
        H.a[2][2] = 0.5 * (-2.0 * w[7] * r3);
        H.a[0][1] = 0.5 * cd(w[0], -w[1]);
        H.a[1][0] = 0.5 * cd(w[0], w[1]);
        H.a[0][2] = 0.5 * cd(w[3], -w[4]);
        H.a[2][0] = 0.5 * cd(w[3], w[4]);
        H.a[1][2] = 0.5 * cd(w[5], -w[6]);
        H.a[2][1] = 0.5 * cd(w[5], w[6]);
        M3 X;
        for (int i = 0; i < 3; ++i)
            for (int j = 0; j < 3; ++j) X.a[i][j] = cd(0, eps) * H.a[i][j];
        M3 U = expm(X);
        reunitarise(U);
        for (int i = 0; i < 3; ++i)
            for (int j = 0; j < 3; ++j) out[l * 9 + i + 3 * j] = U.a[i][j];
    });
    return LYM_OK;
}

int lym_gen_su2_links(const int dims[4], double eps, uint64_t seed, double* links) {
    if (!dims_ok(dims) || !links) return fail(LYM_ERR_INVALID, "bad arguments");
    const long nl = 4L * dims[0] * dims[1] * dims[2] * dims[3];
    cd* out = reinterpret_cast<cd*>(links);
    parallel_for(nl, [&](long l) {
        double a[4];
        gauss2(seed, (uint64_t)l * 2, a[0], a[1]);
        gauss2(seed, (uint64_t)l * 2 + 1, a[2], a[3]);
        double n = std::sqrt(a[0] * a[0] + a[1] * a[1] + a[2] * a[2]);
        double th = eps * n, c = std::cos(th), s = n > 0 ? std::sin(th) / n : 0.0;
        // cos th + i sin th (n . sigma), column-major 2x2
        out[l * 4 + 0] = cd(c, s * a[2]);
        out[l * 4 + 1] = cd(-s * a[1], s * a[0]);
        out[l * 4 + 2] = cd(s * a[1], s * a[0]);
        out[l * 4 + 3] = cd(c, -s * a[2]);
    });
    return LYM_OK;
}

int lym_su2_to_adjoint(long nlinks, const double* su2, double* adj) {
    if (nlinks < 0 || !su2 || !adj) return fail(LYM_ERR_INVALID, "bad arguments");
    const cd I(0, 1);
    const cd sig[3][2][2] = {{{0, 1}, {1, 0}}, {{0, -I}, {I, 0}}, {{1, 0}, {0, -1}}};
    const cd* in = reinterpret_cast<const cd*>(su2);
    parallel_for(nlinks, [&](long l) {
        cd U[2][2], Ud[2][2];
        for (int i = 0; i < 2; ++i)
            for (int j = 0; j < 2; ++j) U[i][j] = in[l * 4 + i + 2 * j];
        for (int i = 0; i < 2; ++i)
            for (int j = 0; j < 2; ++j) Ud[i][j] = std::conj(U[j][i]);
        // R_ab = 2 Re tr(T_a U T_b U^+), T = sigma / 2
        for (int a = 0; a < 3; ++a)
            for (int b = 0; b < 3; ++b) {
                cd tr = 0;
                for (int i = 0; i < 2; ++i)
                    for (int j = 0; j < 2; ++j)
                        for (int k = 0; k < 2; ++k)
                            for (int m = 0; m < 2; ++m) tr += sig[a][i][j] * U[j][k] * sig[b][k][m] * Ud[m][i];
                adj[l * 9 + a + 3 * b] = 0.5 * tr.real();
            }
    });
    return LYM_OK;
}

int lym_apply_bc(const int dims[4], int rep, int bc_t, double* links) {
    if (!dims_ok(dims) || !links) return fail(LYM_ERR_INVALID, "bad arguments");
    if (bc_t == 1) return LYM_OK;
    if (bc_t != -1) return fail(LYM_ERR_INVALID, "bc_t must be +1 or -1");
    const long V = (long)dims[0] * dims[1] * dims[2] * dims[3];
    const int nr = rep == LYM_REP_SU2_ADJ ? 9 : 18;
    for (long s = 0; s < V; ++s)
        if (s % dims[3] == dims[3] - 1)
            for (int k = 0; k < nr; ++k) links[(s * 4 + 3) * nr + k] = -links[(s * 4 + 3) * nr + k];
    return LYM_OK;
}

int lym_gen_gaussian_spinor(long nsites, uint64_t seed, double* out) {
    if (nsites < 0 || !out) return fail(LYM_ERR_INVALID, "bad arguments");
    const double sig = 1.0 / std::sqrt(2.0);
    parallel_for(nsites, [&](long s) {
        for (int k = 0; k < 12; ++k) {
            double a, b;
            gauss2(seed, (uint64_t)s * 12 + k, a, b);
            out[(s * 12 + k) * 2] = sig * a;
            out[(s * 12 + k) * 2 + 1] = sig * b;
        }
    });
    return LYM_OK;
}

int lym_point_source(const int dims[4], const int site[4], int spin, int colour, double* out) {
    if (!dims_ok(dims) || !site || !out) return fail(LYM_ERR_INVALID, "bad arguments");
    if (spin < 0 || spin > 3 || colour < 0 || colour > 2) return fail(LYM_ERR_INVALID, "spin/colour out of range");
    for (int i = 0; i < 4; ++i)
        if (site[i] < 0 || site[i] >= dims[i]) return fail(LYM_ERR_INVALID, "site outside the lattice");
    const long V = (long)dims[0] * dims[1] * dims[2] * dims[3];
    std::memset(out, 0, sizeof(double) * 24 * V);
    long s = ((long)(site[0] * dims[1] + site[1]) * dims[2] + site[2]) * dims[3] + site[3];
    out[(s * 12 + spin * 3 + colour) * 2] = 1.0;
    return LYM_OK;
}

}  // extern "C"
