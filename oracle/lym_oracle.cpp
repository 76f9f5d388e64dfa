// lym_oracle.cpp -- CPU FP64 restatement of LeonardYM's fermion-solve path.
//
// TEST INFRASTRUCTURE ONLY.  This file is the correctness oracle and the
// timed CPU baseline.  It is loaded only by tests/, by smoke() in
// __graft_entry__.py and by the cpu_baseline leg of bench.py, always as the
// checker, never as the thing measured or shipped.  The product path
// (leonardym_amd/) never links, loads or calls it.
//
// Data at the boundary (identical to the reference's raw vectors):
//   spinor  : V x [4 spin][3 colour] complex<double>, interleaved re/im
//   links   : V x [4 mu][3x3 column-major] complex<double>, BC folded in
//   F       : V x [6 planes][3x3 column-major] complex<double>
//   inverse : V x [12x12 column-major] complex<double>
// Site order is t-fastest lexicographic, idx = ((x*Ly + y)*Lz + z)*Lt + t
// (source/MPILattice/LocalLayout.h:149-155); mu = 0,1,2,3 <-> x,y,z,t.
//
// Global sums accumulate in long double, as the reference does
// (source/algebra_utils/AlgebraUtils.h:17-27,182-196).
//
// Three independent forms of the Wilson operator are kept so that each pins
// the others: the unrolled form (DiracWilsonOperator.cpp:16-391), the
// projector form (BasicDiracWilsonOperator.cpp:12-178) and a textbook form
// built from explicit 4x4 gamma matrices (Gamma.cpp:9-10).  The clover term
// likewise has the reference's chiral-block form
// (ImprovedDiracWilsonOperator.cpp:200-243) and the textbook
// Sheikholeslami-Wohlert form  -kappa*csw*sum_{mu<nu} sigma_{mu nu} Fhat_{mu nu}
// with Fhat = -i F.
//
// Parity pinning (all in tests/test_oracle.py): free-field plane-wave closed
// form, constant abelian field closed form, gauge covariance, gamma5
// hermiticity, form A == form B == form C, csw = 0 -> Wilson, abelian
// field-strength closed form, R * R^-1 = 1, Schur identity, solver solutions
// against a dense LU solve of the same operator.

#include <algorithm>
#include <cmath>
#include <complex>
#include <cstdint>
#include <cstring>
#include <map>
#include <mutex>
#include <vector>

#ifdef _OPENMP
#include <omp.h>
#endif

typedef std::complex<double> cd;
typedef long double ld;
typedef std::complex<long double> cld;

namespace {

// ---------------------------------------------------------------- geometry
// LocalLayout::initialize builds sup/down tables once
// (source/MPILattice/LocalLayout.h:34-110); cached here per lattice shape.
struct Geo {
    int L[4];
    long V;
    std::vector<int> up, dn;  // [V][4]
    std::vector<int> par;     // (x+y+z+t) % 2
};

std::mutex geo_mutex;
std::map<std::vector<int>, Geo*> geo_cache;

inline int wrap(int v, int m) { return v >= m ? v - m : (v < 0 ? v + m : v); }

const Geo& geo(const int* d) {
    std::vector<int> key(d, d + 4);
    std::lock_guard<std::mutex> lock(geo_mutex);
    auto it = geo_cache.find(key);
    if (it != geo_cache.end()) return *it->second;
    Geo* g = new Geo;
    for (int i = 0; i < 4; ++i) g->L[i] = d[i];
    g->V = (long)d[0] * d[1] * d[2] * d[3];
    g->up.resize(g->V * 4);
    g->dn.resize(g->V * 4);
    g->par.resize(g->V);
    for (int x = 0; x < d[0]; ++x)
        for (int y = 0; y < d[1]; ++y)
            for (int z = 0; z < d[2]; ++z)
                for (int t = 0; t < d[3]; ++t) {
                    int c[4] = {x, y, z, t};
                    long s = ((long)(x * d[1] + y) * d[2] + z) * d[3] + t;
                    for (int mu = 0; mu < 4; ++mu) {
                        int cu[4] = {x, y, z, t}, cdn[4] = {x, y, z, t};
                        cu[mu] = wrap(c[mu] + 1, d[mu]);
                        cdn[mu] = wrap(c[mu] - 1, d[mu]);
                        g->up[s * 4 + mu] = ((cu[0] * d[1] + cu[1]) * d[2] + cu[2]) * d[3] + cu[3];
                        g->dn[s * 4 + mu] = ((cdn[0] * d[1] + cdn[1]) * d[2] + cdn[2]) * d[3] + cdn[3];
                    }
                    g->par[s] = (x + y + z + t) & 1;
                }
    geo_cache[key] = g;
    return *g;
}

// ---------------------------------------------------------------- 3x3 algebra
// column-major: element (i,j) at i + 3*j, as Eigen stores FermionicGroup.
struct M3 {
    cd a[9];
    cd& operator()(int i, int j) { return a[i + 3 * j]; }
    const cd& operator()(int i, int j) const { return a[i + 3 * j]; }
};

inline M3 load3(const cd* p) { M3 m; for (int k = 0; k < 9; ++k) m.a[k] = p[k]; return m; }
inline M3 mul(const M3& A, const M3& B) {
    M3 C;
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) {
            cd s = 0;
            for (int k = 0; k < 3; ++k) s += A(i, k) * B(k, j);
            C(i, j) = s;
        }
    return C;
}
inline M3 dag(const M3& A) {
    M3 C;
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) C(i, j) = std::conj(A(j, i));
    return C;
}
inline M3 add(const M3& A, const M3& B) { M3 C; for (int k = 0; k < 9; ++k) C.a[k] = A.a[k] + B.a[k]; return C; }

// U * v and U^dagger * v for a colour vector
inline void mv(const cd* U, const cd* v, cd* o) {
    for (int i = 0; i < 3; ++i) o[i] = U[i] * v[0] + U[i + 3] * v[1] + U[i + 6] * v[2];
}
inline void mdv(const cd* U, const cd* v, cd* o) {
    for (int i = 0; i < 3; ++i)
        o[i] = std::conj(U[3 * i]) * v[0] + std::conj(U[3 * i + 1]) * v[1] + std::conj(U[3 * i + 2]) * v[2];
}

// ---------------------------------------------------------------- gamma tables
// Chiral basis of source/utils/Gamma.cpp:9-10 (rows listed).
const cd I(0., 1.);
cd GAMMA[4][4][4];
cd SIGMA[6][4][4];  // sigma_{mu nu} = (i/2)[gamma_mu, gamma_nu], planes 01,02,03,12,13,23
const int PLANE_MU[6] = {0, 0, 0, 1, 1, 2};
const int PLANE_NU[6] = {1, 2, 3, 2, 3, 3};

struct GammaInit {
    GammaInit() {
        const cd g[4][4][4] = {
            {{0, 0, 0, -I}, {0, 0, -I, 0}, {0, I, 0, 0}, {I, 0, 0, 0}},
            {{0, 0, 0, -1.}, {0, 0, 1., 0}, {0, 1., 0, 0}, {-1., 0, 0, 0}},
            {{0, 0, -I, 0}, {0, 0, 0, I}, {I, 0, 0, 0}, {0, -I, 0, 0}},
            {{0, 0, 1., 0}, {0, 0, 0, 1.}, {1., 0, 0, 0}, {0, 1., 0, 0}}};
        std::memcpy(GAMMA, g, sizeof(g));
        for (int k = 0; k < 6; ++k) {
            int m = PLANE_MU[k], n = PLANE_NU[k];
            for (int a = 0; a < 4; ++a)
                for (int b = 0; b < 4; ++b) {
                    cd s = 0;
                    for (int c = 0; c < 4; ++c) s += GAMMA[m][a][c] * GAMMA[n][c][b] - GAMMA[n][a][c] * GAMMA[m][c][b];
                    SIGMA[k][a][b] = 0.5 * I * s;
                }
        }
    }
} gamma_init;

inline const cd* site_spinor(const cd* v, long s) { return v + s * 12; }
inline cd* site_spinor(cd* v, long s) { return v + s * 12; }
inline const cd* link(const cd* U, long s, int mu) { return U + (s * 4 + mu) * 9; }

// ---------------------------------------------------------------- form A: unrolled
// Per direction mu the (1 -+ gamma_mu) projection keeps two spin components.
// h+ (from x+mu, multiplied by U_mu(x)) and h- (from x-mu, multiplied by
// U_mu(x-mu)^dagger) are formed for upper spins a = 0,1; the lower spin p
// partnered with a receives -g_pa (U h+ - U^dag h-), where g_pa is the
// gamma_mu entry.  Written out direction by direction as
// DiracWilsonOperator.cpp:21-192 does.
inline cd mi(cd z) { return cd(z.imag(), -z.real()); }  // -i z
inline cd pi(cd z) { return cd(-z.imag(), z.real()); }  //  i z

void hop_unrolled(const Geo& g, const cd* U, const cd* in, long s, cd H[12]) {
    cd hp[3], hm[3], tmm[3], tmp[3];
    for (int k = 0; k < 12; ++k) H[k] = 0;
    // mu = 0: h+_0 = psi0 + i psi3, h+_1 = psi1 + i psi2; lower: H3 += -i(...), H2 += -i(...)
    {
        long su = g.up[s * 4 + 0], sd = g.dn[s * 4 + 0];
        const cd *pu = site_spinor(in, su), *pd = site_spinor(in, sd);
        const cd *Uf = link(U, s, 0), *Ub = link(U, sd, 0);
        for (int c = 0; c < 3; ++c) { hp[c] = pu[c] + pi(pu[9 + c]); hm[c] = pd[c] - pi(pd[9 + c]); }
        mv(Uf, hp, tmm); mdv(Ub, hm, tmp);
        for (int c = 0; c < 3; ++c) { H[c] += tmm[c] + tmp[c]; H[9 + c] += mi(tmm[c] - tmp[c]); }
        for (int c = 0; c < 3; ++c) { hp[c] = pu[3 + c] + pi(pu[6 + c]); hm[c] = pd[3 + c] - pi(pd[6 + c]); }
        mv(Uf, hp, tmm); mdv(Ub, hm, tmp);
        for (int c = 0; c < 3; ++c) { H[3 + c] += tmm[c] + tmp[c]; H[6 + c] += mi(tmm[c] - tmp[c]); }
    }
    // mu = 1: h+_0 = psi0 + psi3 (H3 += +), h+_1 = psi1 - psi2 (H2 += -)
    {
        long su = g.up[s * 4 + 1], sd = g.dn[s * 4 + 1];
        const cd *pu = site_spinor(in, su), *pd = site_spinor(in, sd);
        const cd *Uf = link(U, s, 1), *Ub = link(U, sd, 1);
        for (int c = 0; c < 3; ++c) { hp[c] = pu[c] + pu[9 + c]; hm[c] = pd[c] - pd[9 + c]; }
        mv(Uf, hp, tmm); mdv(Ub, hm, tmp);
        for (int c = 0; c < 3; ++c) { H[c] += tmm[c] + tmp[c]; H[9 + c] += tmm[c] - tmp[c]; }
        for (int c = 0; c < 3; ++c) { hp[c] = pu[3 + c] - pu[6 + c]; hm[c] = pd[3 + c] + pd[6 + c]; }
        mv(Uf, hp, tmm); mdv(Ub, hm, tmp);
        for (int c = 0; c < 3; ++c) { H[3 + c] += tmm[c] + tmp[c]; H[6 + c] -= tmm[c] - tmp[c]; }
    }
    // mu = 2: h+_0 = psi0 + i psi2 (H2 += -i), h+_1 = psi1 - i psi3 (H3 += +i)
    {
        long su = g.up[s * 4 + 2], sd = g.dn[s * 4 + 2];
        const cd *pu = site_spinor(in, su), *pd = site_spinor(in, sd);
        const cd *Uf = link(U, s, 2), *Ub = link(U, sd, 2);
        for (int c = 0; c < 3; ++c) { hp[c] = pu[c] + pi(pu[6 + c]); hm[c] = pd[c] - pi(pd[6 + c]); }
        mv(Uf, hp, tmm); mdv(Ub, hm, tmp);
        for (int c = 0; c < 3; ++c) { H[c] += tmm[c] + tmp[c]; H[6 + c] += mi(tmm[c] - tmp[c]); }
        for (int c = 0; c < 3; ++c) { hp[c] = pu[3 + c] - pi(pu[9 + c]); hm[c] = pd[3 + c] + pi(pd[9 + c]); }
        mv(Uf, hp, tmm); mdv(Ub, hm, tmp);
        for (int c = 0; c < 3; ++c) { H[3 + c] += tmm[c] + tmp[c]; H[9 + c] += pi(tmm[c] - tmp[c]); }
    }
    // mu = 3: h+_0 = psi0 - psi2 (H2 += -), h+_1 = psi1 - psi3 (H3 += -)
    {
        long su = g.up[s * 4 + 3], sd = g.dn[s * 4 + 3];
        const cd *pu = site_spinor(in, su), *pd = site_spinor(in, sd);
        const cd *Uf = link(U, s, 3), *Ub = link(U, sd, 3);
        for (int c = 0; c < 3; ++c) { hp[c] = pu[c] - pu[6 + c]; hm[c] = pd[c] + pd[6 + c]; }
        mv(Uf, hp, tmm); mdv(Ub, hm, tmp);
        for (int c = 0; c < 3; ++c) { H[c] += tmm[c] + tmp[c]; H[6 + c] -= tmm[c] - tmp[c]; }
        for (int c = 0; c < 3; ++c) { hp[c] = pu[3 + c] - pu[9 + c]; hm[c] = pd[3 + c] + pd[9 + c]; }
        mv(Uf, hp, tmm); mdv(Ub, hm, tmp);
        for (int c = 0; c < 3; ++c) { H[3 + c] += tmm[c] + tmp[c]; H[9 + c] -= tmm[c] - tmp[c]; }
    }
}

// ---------------------------------------------------------------- form B: projector
// BasicDiracWilsonOperator.cpp:16-93: tmp_plus = U h+ + U^dag h-,
// tmp_minus = U h+ - U^dag h- for both upper spins and all mu, then the
// lower spins are assembled from tmp_minus in one final expression.
void hop_projector(const Geo& g, const cd* U, const cd* in, long s, cd H[12]) {
    cd tp[4][2][3], tm[4][2][3];
    cd h[2][3], a[3], b[3];
    for (int mu = 0; mu < 4; ++mu) {
        const cd* p = site_spinor(in, g.up[s * 4 + mu]);
        for (int c = 0; c < 3; ++c) {
            switch (mu) {
                case 0: h[0][c] = p[c] + pi(p[9 + c]); h[1][c] = p[3 + c] + pi(p[6 + c]); break;
                case 1: h[0][c] = p[c] + p[9 + c]; h[1][c] = p[3 + c] - p[6 + c]; break;
                case 2: h[0][c] = p[c] + pi(p[6 + c]); h[1][c] = p[3 + c] - pi(p[9 + c]); break;
                default: h[0][c] = p[c] - p[6 + c]; h[1][c] = p[3 + c] - p[9 + c]; break;
            }
        }
        for (int n = 0; n < 2; ++n) mv(link(U, s, mu), h[n], tp[mu][n]);
    }
    for (int mu = 0; mu < 4; ++mu) {
        long sd = g.dn[s * 4 + mu];
        const cd* p = site_spinor(in, sd);
        for (int c = 0; c < 3; ++c) {
            switch (mu) {
                case 0: h[0][c] = p[c] - pi(p[9 + c]); h[1][c] = p[3 + c] - pi(p[6 + c]); break;
                case 1: h[0][c] = p[c] - p[9 + c]; h[1][c] = p[3 + c] + p[6 + c]; break;
                case 2: h[0][c] = p[c] - pi(p[6 + c]); h[1][c] = p[3 + c] + pi(p[9 + c]); break;
                default: h[0][c] = p[c] + p[6 + c]; h[1][c] = p[3 + c] + p[9 + c]; break;
            }
        }
        for (int n = 0; n < 2; ++n) {
            mdv(link(U, sd, mu), h[n], a);
            for (int c = 0; c < 3; ++c) { b[c] = tp[mu][n][c]; tm[mu][n][c] = b[c] - a[c]; tp[mu][n][c] = b[c] + a[c]; }
        }
    }
    for (int c = 0; c < 3; ++c) {
        H[c] = tp[0][0][c] + tp[1][0][c] + tp[2][0][c] + tp[3][0][c];
        H[3 + c] = tp[0][1][c] + tp[1][1][c] + tp[2][1][c] + tp[3][1][c];
        // lower spins: D_2 = psi_2 + kappa (tm11 + tm30 + i tm01 + i tm20),
        //              D_3 = psi_3 + kappa (i tm00 - i tm21 - tm10 + tm31)  (:85-89)
        H[6 + c] = -tm[1][1][c] - tm[3][0][c] - pi(tm[0][1][c]) - pi(tm[2][0][c]);
        H[9 + c] = -pi(tm[0][0][c]) + pi(tm[2][1][c]) + tm[1][0][c] - tm[3][1][c];
    }
}

// ---------------------------------------------------------------- form C: textbook
// D psi(x) = psi(x) - kappa sum_mu [(1 - g_mu) U_mu(x) psi(x+mu) + (1 + g_mu) U_mu(x-mu)^dag psi(x-mu)]
// with full 4x4 gamma matrices, no projection.
void hop_textbook(const Geo& g, const cd* U, const cd* in, long s, cd H[12]) {
    for (int k = 0; k < 12; ++k) H[k] = 0;
    cd w[4][3];
    for (int mu = 0; mu < 4; ++mu) {
        long su = g.up[s * 4 + mu], sd = g.dn[s * 4 + mu];
        const cd *pu = site_spinor(in, su), *pd = site_spinor(in, sd);
        for (int b = 0; b < 4; ++b) mv(link(U, s, mu), pu + 3 * b, w[b]);
        for (int a = 0; a < 4; ++a)
            for (int c = 0; c < 3; ++c) {
                cd acc = w[a][c];
                for (int b = 0; b < 4; ++b) acc -= GAMMA[mu][a][b] * w[b][c];
                H[3 * a + c] += acc;
            }
        for (int b = 0; b < 4; ++b) mdv(link(U, sd, mu), pd + 3 * b, w[b]);
        for (int a = 0; a < 4; ++a)
            for (int c = 0; c < 3; ++c) {
                cd acc = w[a][c];
                for (int b = 0; b < 4; ++b) acc += GAMMA[mu][a][b] * w[b][c];
                H[3 * a + c] += acc;
            }
    }
}

// ---------------------------------------------------------------- clover
// Reference chiral-block combination, ImprovedDiracWilsonOperator.cpp:200-243:
// clover_a = sum_j B_aj(F) psi_j, added to Q as +kappa*csw*clover.
void clover_block(const cd* F, const cd* psi, cd out[12]) {
    const cd *F0 = F, *F1 = F + 9, *F2 = F + 18, *F3 = F + 27, *F4 = F + 36, *F5 = F + 45;
    for (int i = 0; i < 3; ++i) {
        cd c0 = 0, c1 = 0, c2 = 0, c3 = 0;
        for (int j = 0; j < 3; ++j) {
            int e = i + 3 * j;
            c0 += pi((-F0[e] + F5[e]) * psi[j]);
            c1 += pi((F0[e] - F5[e]) * psi[3 + j]);
            c2 += pi((F0[e] + F5[e]) * psi[6 + j]);
            c3 += pi((-F0[e] - F5[e]) * psi[9 + j]);
            c0 += (F1[e] + F4[e] + pi(F2[e] - F3[e])) * psi[3 + j];
            c1 += (-F1[e] - F4[e] + pi(F2[e] - F3[e])) * psi[j];
            c2 += (-F1[e] + F4[e] + pi(F2[e] + F3[e])) * psi[9 + j];
            c3 += (F1[e] - F4[e] + pi(F2[e] + F3[e])) * psi[6 + j];
        }
        out[i] = c0; out[3 + i] = c1; out[6 + i] = c2; out[9 + i] = c3;
    }
}

// Textbook SW term in D: -kappa*csw*sum_{mu<nu} sigma_{mu nu} Fhat_{mu nu} psi, Fhat = -i F.
void clover_sigma(const cd* F, const cd* psi, double kc, cd out[12]) {
    for (int k = 0; k < 12; ++k) out[k] = 0;
    cd w[4][3];
    for (int p = 0; p < 6; ++p) {
        const cd* Fp = F + 9 * p;
        for (int b = 0; b < 4; ++b)
            for (int i = 0; i < 3; ++i) {
                cd acc = 0;
                for (int j = 0; j < 3; ++j) acc += (-I * Fp[i + 3 * j]) * psi[3 * b + j];
                w[b][i] = acc;
            }
        for (int a = 0; a < 4; ++a)
            for (int i = 0; i < 3; ++i) {
                cd acc = 0;
                for (int b = 0; b < 4; ++b) acc += SIGMA[p][a][b] * w[b][i];
                out[3 * a + i] -= kc * acc;
            }
    }
}

// ---------------------------------------------------------------- operator kinds
enum { K_WILSON = 0, K_WILSON_BASIC = 1, K_CLOVER = 2, K_EO_CLOVER = 3, K_TEXTBOOK = 4 };

struct Op {
    const Geo* g;
    int kind, power, gamma5;
    double kappa, csw;
    const cd *U, *F, *inv;
};

// out = Op v1 + alpha v2 for the single-power operators (Q when gamma5, else D).
// The gamma5 = false fix-up of the reference (DiracWilsonOperator.cpp:383-388)
// reduces to exactly this.
void apply1(const Op& op, cd* out, const cd* v1, const cd* v2, cd alpha) {
    const Geo& g = *op.g;
    const long V = g.V;
#pragma omp parallel for schedule(static)
    for (long s = 0; s < V; ++s) {
        cd H[12], C[12];
        const cd* ps = site_spinor(v1, s);
        cd* o = site_spinor(out, s);
        if (op.kind == K_WILSON || op.kind == K_CLOVER) hop_unrolled(g, op.U, v1, s, H);
        else if (op.kind == K_WILSON_BASIC) hop_projector(g, op.U, v1, s, H);
        else hop_textbook(g, op.U, v1, s, H);
        // Q = gamma5 (psi - kappa H)
        for (int k = 0; k < 6; ++k) o[k] = ps[k] - op.kappa * H[k];
        for (int k = 6; k < 12; ++k) o[k] = -(ps[k] - op.kappa * H[k]);
        if (op.kind == K_CLOVER && op.csw != 0.) {
            clover_block(op.F + s * 54, ps, C);
            for (int k = 0; k < 12; ++k) o[k] += (op.kappa * op.csw) * C[k];
        } else if (op.kind == K_TEXTBOOK && op.csw != 0.) {
            clover_sigma(op.F + s * 54, ps, op.kappa * op.csw, C);  // term in D
            for (int k = 0; k < 6; ++k) o[k] += C[k];
            for (int k = 6; k < 12; ++k) o[k] -= C[k];
        }
        if (!op.gamma5) for (int k = 6; k < 12; ++k) o[k] = -o[k];
        if (v2) for (int k = 0; k < 12; ++k) o[k] += alpha * v2[s * 12 + k];
    }
}

// ---------------------------------------------------------------- even-odd
// EvenOddImprovedDiracWilsonOperator.cpp:204-281: on sites of parity `part`
// out = -kappa H in (the hopping part of D), elsewhere out = in.  Aliasing of
// out and in is allowed as in the reference (it reads the other parity only).
void eo_hop(const Op& op, cd* out, const cd* in, int part) {
    const Geo& g = *op.g;
    std::vector<cd> res(g.V * 12);
#pragma omp parallel for schedule(static)
    for (long s = 0; s < g.V; ++s) {
        if (g.par[s] == part) {
            cd H[12];
            hop_unrolled(g, op.U, in, s, H);
            for (int k = 0; k < 12; ++k) res[s * 12 + k] = -op.kappa * H[k];
        } else {
            for (int k = 0; k < 12; ++k) res[s * 12 + k] = in[s * 12 + k];
        }
    }
    std::memcpy(out, res.data(), sizeof(cd) * 12 * g.V);
}

// R = 1 + kappa*csw*gamma5*B on one site as a dense 12x12 (index 3*spin + colour),
// EvenOddImprovedDiracWilsonOperator.cpp:291-311.
void site_R(const cd* F, double kc, cd R[144]) {
    for (int k = 0; k < 144; ++k) R[k] = 0;
    cd col[12], unit[12];
    for (int b = 0; b < 12; ++b) {
        for (int k = 0; k < 12; ++k) unit[k] = (k == b) ? 1. : 0.;
        clover_block(F, unit, col);
        for (int a = 0; a < 12; ++a) {
            cd v = kc * col[a];
            if (a >= 6) v = -v;
            R[a + 12 * b] = v + (a == b ? 1. : 0.);
        }
    }
}

// Gauss-Jordan with partial pivoting, column-major n x n.
bool invert(cd* A, cd* Ainv, int n) {
    std::vector<cd> M(A, A + n * n);
    for (int i = 0; i < n * n; ++i) Ainv[i] = 0;
    for (int i = 0; i < n; ++i) Ainv[i + n * i] = 1;
    for (int c = 0; c < n; ++c) {
        int piv = c;
        double best = std::abs(M[c + n * c]);
        for (int r = c + 1; r < n; ++r)
            if (std::abs(M[r + n * c]) > best) { best = std::abs(M[r + n * c]); piv = r; }
        if (best == 0.) return false;
        if (piv != c)
            for (int j = 0; j < n; ++j) { std::swap(M[c + n * j], M[piv + n * j]); std::swap(Ainv[c + n * j], Ainv[piv + n * j]); }
        cd d = 1. / M[c + n * c];
        for (int j = 0; j < n; ++j) { M[c + n * j] *= d; Ainv[c + n * j] *= d; }
        for (int r = 0; r < n; ++r) {
            if (r == c) continue;
            cd f = M[r + n * c];
            if (f == 0.) continue;
            for (int j = 0; j < n; ++j) { M[r + n * j] -= f * M[c + n * j]; Ainv[r + n * j] -= f * Ainv[c + n * j]; }
        }
    }
    return true;
}

// R^-1 applied on the even sites in place (multiplyEvenEvenInverse, :61-89).
void eo_inv_even(const Op& op, cd* v) {
    const Geo& g = *op.g;
#pragma omp parallel for schedule(static)
    for (long s = 0; s < g.V; ++s) {
        if (g.par[s] != 0) continue;
        cd a[12], *o = v + s * 12;
        const cd* M = op.inv + s * 144;
        for (int k = 0; k < 12; ++k) a[k] = o[k];
        for (int r = 0; r < 12; ++r) {
            cd acc = 0;
            for (int c = 0; c < 12; ++c) acc += M[r + 12 * c] * a[c];
            o[r] = acc;
        }
    }
}

// multiplyOddOddMinusIdentity (:148-201): on parity `part`,
// out = R in - out  (R = 1 + kappa csw gamma5 B).
void eo_oddodd_minus(const Op& op, cd* out, const cd* in, int part) {
    const Geo& g = *op.g;
#pragma omp parallel for schedule(static)
    for (long s = 0; s < g.V; ++s) {
        if (g.par[s] != part) continue;
        cd C[12];
        const cd* ps = in + s * 12;
        cd* o = out + s * 12;
        clover_block(op.F + s * 54, ps, C);
        for (int k = 0; k < 6; ++k) o[k] = ps[k] + (op.kappa * op.csw) * C[k] - o[k];
        for (int k = 6; k < 12; ++k) o[k] = ps[k] - (op.kappa * op.csw) * C[k] - o[k];
    }
}

// EvenOdd multiply (:26-47): out_e = in_e; out_o = gamma5^[gamma5] (R_oo - D_oe R_ee^-1 D_eo) in_o.
void eo_apply(const Op& op, cd* out, const cd* in) {
    const Geo& g = *op.g;
    eo_hop(op, out, in, 0);
    eo_inv_even(op, out);
    eo_hop(op, out, out, 1);
    eo_oddodd_minus(op, out, in, 1);
#pragma omp parallel for schedule(static)
    for (long s = 0; s < g.V; ++s) {
        if (g.par[s] == 0) {
            for (int k = 0; k < 12; ++k) out[s * 12 + k] = in[s * 12 + k];
        } else if (op.gamma5) {
            for (int k = 6; k < 12; ++k) out[s * 12 + k] = -out[s * 12 + k];
        }
    }
}

void apply_op(const Op& op, cd* out, const cd* in, std::vector<cd>& tmp) {
    if (op.kind == K_EO_CLOVER) {
        if (op.power == 1) { eo_apply(op, out, in); return; }
        tmp.resize(op.g->V * 12);
        eo_apply(op, tmp.data(), in);
        eo_apply(op, out, tmp.data());
        return;
    }
    if (op.power == 1) { apply1(op, out, in, nullptr, 0.); return; }
    tmp.resize(op.g->V * 12);
    apply1(op, tmp.data(), in, nullptr, 0.);
    apply1(op, out, tmp.data(), nullptr, 0.);
}

// ---------------------------------------------------------------- BLAS-1 (long double sums)
ld norm2(long n, const cd* a) {
    ld r = 0;
#pragma omp parallel for reduction(+ : r) schedule(static)
    for (long i = 0; i < n; ++i) r += (ld)std::norm(a[i]);
    return r;
}
cld dotc(long n, const cd* a, const cd* b) {
    ld re = 0, im = 0;
#pragma omp parallel for reduction(+ : re, im) schedule(static)
    for (long i = 0; i < n; ++i) {
        cd p = std::conj(a[i]) * b[i];
        re += p.real();
        im += p.imag();
    }
    return cld(re, im);
}

Op make_op(const int* dims, int kind, int power, double kappa, double csw, int gamma5,
           const double* links, const double* F, const double* inv) {
    Op op;
    op.g = &geo(dims);
    op.kind = kind; op.power = power; op.gamma5 = gamma5;
    op.kappa = kappa; op.csw = csw;
    op.U = reinterpret_cast<const cd*>(links);
    op.F = reinterpret_cast<const cd*>(F);
    op.inv = reinterpret_cast<const cd*>(inv);
    return op;
}

// splitmix64, used only for the BiCGStab random start on a zero source
inline uint64_t splitmix(uint64_t x) {
    x += 0x9E3779B97F4A7C15ull;
    x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ull;
    x = (x ^ (x >> 27)) * 0x94D049BB133111EBull;
    return x ^ (x >> 31);
}

}  // namespace

extern "C" {

typedef struct {
    int dims[4];
    int kind;    // 0 DiracWilson, 1 BasicDiracWilson, 2 Improved (clover), 3 EvenOdd clover, 4 textbook
    int power;   // 1 or 2
    double kappa, csw;
    int gamma5;
    const double* links;
    const double* F;
    const double* inv;
} orc_op_desc;

int orc_version(void) { return 1; }

int orc_num_threads(void) {
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}

// neighbour tables (LocalLayout.h:71-104)
void orc_tables(const int* dims, int* sup, int* sdn, int* parity) {
    const Geo& g = geo(dims);
    std::memcpy(sup, g.up.data(), sizeof(int) * g.V * 4);
    std::memcpy(sdn, g.dn.data(), sizeof(int) * g.V * 4);
    if (parity) std::memcpy(parity, g.par.data(), sizeof(int) * g.V);
}

// switchAntiperiodicBc (LatticeBoundaryConditions.h:8-14): U_t(x) -> -U_t(x) at t = T-1
void orc_apply_bc(const int* dims, double* links) {
    const Geo& g = geo(dims);
    cd* U = reinterpret_cast<cd*>(links);
    for (long s = 0; s < g.V; ++s)
        if (s % g.L[3] == g.L[3] - 1)
            for (int k = 0; k < 9; ++k) U[(s * 4 + 3) * 9 + k] = -U[(s * 4 + 3) * 9 + k];
}

// out = Op v1 (+ alpha v2 if v2 != NULL), single power, kinds 0,1,2,4
void orc_multiply(const orc_op_desc* d, double* out, const double* v1, const double* v2, double are, double aim) {
    Op op = make_op(d->dims, d->kind, d->power, d->kappa, d->csw, d->gamma5, d->links, d->F, d->inv);
    cd* o = reinterpret_cast<cd*>(out);
    const cd* a = reinterpret_cast<const cd*>(v1);
    const cd* b = reinterpret_cast<const cd*>(v2);
    std::vector<cd> tmp;
    if (op.kind == K_EO_CLOVER || op.power == 2) {
        apply_op(op, o, a, tmp);
        if (b) {
            const long n = op.g->V * 12;
            cd al(are, aim);
            for (long i = 0; i < n; ++i) o[i] += al * b[i];
        }
        return;
    }
    apply1(op, o, a, b, cd(are, aim));
}

// Field strength, ImprovedDiracWilsonOperator.cpp:502-523.  Clover leaves of
// plane (mu,nu) at x, all starting and ending at x; F = (Q - Q^dag)/8.
void orc_field_strength(const int* dims, const double* links, double* Fout) {
    const Geo& g = geo(dims);
    const cd* U = reinterpret_cast<const cd*>(links);
    cd* F = reinterpret_cast<cd*>(Fout);
#pragma omp parallel for schedule(static)
    for (long s = 0; s < g.V; ++s) {
        for (int p = 0; p < 6; ++p) {
            int m = PLANE_MU[p], n = PLANE_NU[p];
            long xm = g.dn[s * 4 + m], xn = g.dn[s * 4 + n];
            long xmn = g.dn[xm * 4 + n];
            long xpm = g.up[s * 4 + m], xpn = g.up[s * 4 + n];
            long xnpm = g.up[xn * 4 + m];   // x - nu + mu
            long xmpn = g.up[xm * 4 + n];   // x - mu + nu
            M3 Um_x = load3(link(U, s, m)), Un_x = load3(link(U, s, n));
            M3 Um_xm = load3(link(U, xm, m)), Un_xm = load3(link(U, xm, n));
            M3 Um_xmn = load3(link(U, xmn, m)), Un_xmn = load3(link(U, xmn, n));
            M3 Un_xn = load3(link(U, xn, n)), Um_xn = load3(link(U, xn, m));
            M3 Un_xnpm = load3(link(U, xnpm, n));
            M3 Un_xpm = load3(link(U, xpm, n)), Um_xpn = load3(link(U, xpn, m));
            M3 Um_xmpn = load3(link(U, xmpn, m));
            M3 l1 = mul(mul(dag(Um_xm), dag(Un_xmn)), mul(Um_xmn, Un_xn));
            M3 l2 = mul(mul(dag(Un_xn), Um_xn), mul(Un_xnpm, dag(Um_x)));
            M3 l3 = mul(mul(Um_x, Un_xpm), mul(dag(Um_xpn), dag(Un_x)));
            M3 l4 = mul(mul(Un_x, dag(Um_xmpn)), mul(dag(Un_xm), Um_xm));
            M3 Q = add(add(l1, l2), add(l3, l4));
            M3 Qd = dag(Q);
            for (int k = 0; k < 9; ++k) F[(s * 6 + p) * 9 + k] = 0.125 * (Q.a[k] - Qd.a[k]);
        }
    }
}

// R_ee^-1 for every site (calculateInverseEvenEven, :285-317).  Returns the
// number of singular sites.
int orc_clover_inverse(const int* dims, const double* Fin, double kappa, double csw, double* inv_out) {
    const Geo& g = geo(dims);
    const cd* F = reinterpret_cast<const cd*>(Fin);
    cd* inv = reinterpret_cast<cd*>(inv_out);
    int bad = 0;
#pragma omp parallel for reduction(+ : bad) schedule(static)
    for (long s = 0; s < g.V; ++s) {
        cd R[144];
        site_R(F + s * 54, kappa * csw, R);
        if (!invert(R, inv + s * 144, 12)) ++bad;
    }
    return bad;
}

// dense site matrix R = 1 + kappa csw gamma5 B (for tests)
void orc_clover_site_matrix(const double* Fsite, double kappa, double csw, double* R) {
    site_R(reinterpret_cast<const cd*>(Fsite), kappa * csw, reinterpret_cast<cd*>(R));
}

// EO building blocks on full-V vectors
void orc_eo_hop(const orc_op_desc* d, double* out, const double* in, int part) {
    Op op = make_op(d->dims, K_EO_CLOVER, 1, d->kappa, d->csw, d->gamma5, d->links, d->F, d->inv);
    eo_hop(op, reinterpret_cast<cd*>(out), reinterpret_cast<const cd*>(in), part);
}
void orc_eo_inv_even(const orc_op_desc* d, double* v) {
    Op op = make_op(d->dims, K_EO_CLOVER, 1, d->kappa, d->csw, d->gamma5, d->links, d->F, d->inv);
    eo_inv_even(op, reinterpret_cast<cd*>(v));
}
void orc_eo_oddodd_minus(const orc_op_desc* d, double* out, const double* in, int part) {
    Op op = make_op(d->dims, K_EO_CLOVER, 1, d->kappa, d->csw, d->gamma5, d->links, d->F, d->inv);
    eo_oddodd_minus(op, reinterpret_cast<cd*>(out), reinterpret_cast<const cd*>(in), part);
}

// ------------------------------------------------ AlgebraUtils (AlgebraUtils.h)
double orc_squared_norm(long nsites, const double* a) {
    return (double)norm2(nsites * 12, reinterpret_cast<const cd*>(a));
}
void orc_dot(long nsites, const double* a, const double* b, double* re, double* im) {
    cld r = dotc(nsites * 12, reinterpret_cast<const cd*>(a), reinterpret_cast<const cd*>(b));
    *re = (double)r.real();
    *im = (double)r.imag();
}
void orc_gamma5dot(long nsites, const double* a, const double* b, double* re, double* im) {
    const cd* x = reinterpret_cast<const cd*>(a);
    const cd* y = reinterpret_cast<const cd*>(b);
    ld sr = 0, si = 0;
    for (long s = 0; s < nsites; ++s)
        for (int k = 0; k < 12; ++k) {
            cd p = std::conj(x[s * 12 + k]) * y[s * 12 + k];
            if (k < 6) { sr += p.real(); si += p.imag(); }
            else { sr -= p.real(); si -= p.imag(); }
        }
    *re = (double)sr;
    *im = (double)si;
}
double orc_difference_norm(long nsites, const double* a, const double* b) {
    const cd* x = reinterpret_cast<const cd*>(a);
    const cd* y = reinterpret_cast<const cd*>(b);
    ld r = 0;
    for (long i = 0; i < nsites * 12; ++i) r += (ld)std::norm(x[i] - y[i]);
    return (double)r;
}

// ------------------------------------------------ solvers
typedef struct {
    int converged;
    unsigned steps;     // lastSteps as the reference reports it
    double last_error;  // final |r|^2
    unsigned applies;   // operator applications
} orc_solve_info;

// ConjugateGradient::solve (ConjugateGradient.cpp:10-74), stop |r|^2 < eps.
int orc_cg(const orc_op_desc* d, const double* bsrc, double* xout, const double* x0, double eps,
           unsigned max_steps, orc_solve_info* info) {
    Op op = make_op(d->dims, d->kind, d->power, d->kappa, d->csw, d->gamma5, d->links, d->F, d->inv);
    const long n = op.g->V * 12;
    const cd* b = reinterpret_cast<const cd*>(bsrc);
    std::vector<cd> x(n), r(n), p(n), tmp(n), scratch;
    const cd* x0c = x0 ? reinterpret_cast<const cd*>(x0) : b;
    std::copy(x0c, x0c + n, x.begin());
    unsigned applies = 0;
    apply_op(op, tmp.data(), x.data(), scratch); ++applies;
    for (long i = 0; i < n; ++i) { r[i] = b[i] - tmp[i]; p[i] = r[i]; }
    ld norm = norm2(n, r.data()), norm_next = norm;
    int ok = 0;
    unsigned steps = max_steps;
    for (unsigned step = 0; step < max_steps; ++step) {
        apply_op(op, tmp.data(), p.data(), scratch); ++applies;
        norm = norm_next;
        cld pAp = dotc(n, p.data(), tmp.data());
        cd alpha = cd(cld(norm, 0) / pAp);
#pragma omp parallel for schedule(static)
        for (long i = 0; i < n; ++i) { x[i] += alpha * p[i]; r[i] -= alpha * tmp[i]; }
        norm_next = norm2(n, r.data());
        if (norm_next < eps) { steps = step; ok = 1; break; }
        double beta = (double)(norm_next / norm);
#pragma omp parallel for schedule(static)
        for (long i = 0; i < n; ++i) p[i] = r[i] + beta * p[i];
    }
    std::copy(x.begin(), x.end(), reinterpret_cast<cd*>(xout));
    if (info) { info->converged = ok; info->steps = steps; info->last_error = (double)norm_next; info->applies = applies; }
    return ok;
}

// BiConjugateGradient::solve, plain BiCGStab (BiConjugateGradient.cpp:321-507),
// and the shifted overload (:510-690), which is the same iteration on
// A + shift (every A v there is dirac->multiplyAdd(out, v, v, shift)).
static int bicgstab_impl(const orc_op_desc* d, const double* bsrc, double* xout, const double* x0,
                         double precision, unsigned max_steps, cd shift, orc_solve_info* info);

int orc_bicgstab(const orc_op_desc* d, const double* bsrc, double* xout, const double* x0, double precision,
                 unsigned max_steps, orc_solve_info* info) {
    return bicgstab_impl(d, bsrc, xout, x0, precision, max_steps, cd(0, 0), info);
}

int orc_bicgstab_shift(const orc_op_desc* d, const double* bsrc, double* xout, const double* x0, double precision,
                       unsigned max_steps, double shift_re, double shift_im, orc_solve_info* info) {
    return bicgstab_impl(d, bsrc, xout, x0, precision, max_steps, cd(shift_re, shift_im), info);
}

static int bicgstab_impl(const orc_op_desc* d, const double* bsrc, double* xout, const double* x0,
                         double precision, unsigned max_steps, cd shift, orc_solve_info* info) {
    Op op0 = make_op(d->dims, d->kind, d->power, d->kappa, d->csw, d->gamma5, d->links, d->F, d->inv);
    const long n = op0.g->V * 12;
    auto apply_op = [&](const Op& o, cd* out, const cd* in, std::vector<cd>& scr) {
        ::apply_op(o, out, in, scr);
        if (shift != cd(0, 0))
            for (long i = 0; i < n; ++i) out[i] += shift * in[i];
    };
    const Op& op = op0;
    const cd* b = reinterpret_cast<const cd*>(bsrc);
    std::vector<cd> x(n), r(n), rh(n), p(n), nu(n), s(n), t(n), scratch;
    if (x0) {
        std::copy(reinterpret_cast<const cd*>(x0), reinterpret_cast<const cd*>(x0) + n, x.begin());
    } else if ((double)norm2(n, b) > precision) {
        std::copy(b, b + n, x.begin());
    } else {
        // the reference draws a time-seeded random vector here; a fixed seed is used instead
        for (long i = 0; i < n; ++i)
            x[i] = cd((splitmix(2 * i) >> 11) * 0x1.0p-53, (splitmix(2 * i + 1) >> 11) * 0x1.0p-53);
    }
    unsigned applies = 0;
    apply_op(op, p.data(), x.data(), scratch); ++applies;
    for (long i = 0; i < n; ++i) { r[i] = b[i] - p[i]; rh[i] = b[i] + p[i]; p[i] = 0; nu[i] = 0; }
    cd alpha = 1., omega = 1.;
    cld rho = 1.;
    unsigned step = 0;
    ld norm = 0;
    int ok = 0;
    double last = 0;
    while (step < max_steps) {
        cld rho_next = dotc(n, rh.data(), r.data());
        if (std::norm(rho_next) == 0.) { ok = 0; last = (double)norm; goto done; }
        {
            cd beta = cd(rho_next / rho) * (alpha / omega);
#pragma omp parallel for schedule(static)
            for (long i = 0; i < n; ++i) p[i] = r[i] + beta * (p[i] - omega * nu[i]);
            apply_op(op, nu.data(), p.data(), scratch); ++applies;
            cld at = dotc(n, rh.data(), nu.data());
            alpha = cd(rho_next / at);
#pragma omp parallel for schedule(static)
            for (long i = 0; i < n; ++i) s[i] = r[i] - alpha * nu[i];
            apply_op(op, t.data(), s.data(), scratch); ++applies;
            cld t1 = dotc(n, t.data(), s.data());
            cld t2 = dotc(n, t.data(), t.data());
            omega = cd(t1 / t2);
            if (t2.real() == 0) {
                std::copy(b, b + n, x.begin());
                ok = 1;
                last = 0;
                goto done;
            }
#pragma omp parallel for schedule(static)
            for (long i = 0; i < n; ++i) { x[i] += alpha * p[i] + omega * s[i]; r[i] = s[i] - omega * t[i]; }
            norm = norm2(n, r.data());
            last = (double)norm;
            if (norm < precision) { ok = 1; goto done; }
            rho = rho_next;
            ++step;
        }
    }
    step = max_steps;
done:
    std::copy(x.begin(), x.end(), reinterpret_cast<cd*>(xout));
    if (info) { info->converged = ok; info->steps = step; info->last_error = last; info->applies = applies; }
    return ok;
}

// PreconditionedBiCGStab::solve (PreconditionedBiCGStab.cpp:33-100).  d must
// describe the clover operator (kind 2) with F and inv (built with the same
// kappa, csw).  With all four extents even, the even-odd branch: returns
// x = D^-1 b.  Otherwise plain BiCGStab on the operator as d describes it
// (the reference's fallback, :43).
int orc_precond_bicgstab(const orc_op_desc* d, const double* bsrc, double* xout, const double* x0,
                         double precision, unsigned max_steps, orc_solve_info* info) {
    if ((d->dims[0] | d->dims[1] | d->dims[2] | d->dims[3]) % 2)
        return orc_bicgstab(d, bsrc, xout, x0, precision, max_steps, info);
    Op eo = make_op(d->dims, K_EO_CLOVER, 1, d->kappa, d->csw, 1, d->links, d->F, d->inv);
    const Geo& g = *eo.g;
    const long n = g.V * 12;
    const cd* b = reinterpret_cast<const cd*>(bsrc);
    std::vector<cd> src(b, b + n), src_eo(n), xo(n), xe(n);
    eo_inv_even(eo, src.data());
    eo_hop(eo, src.data(), src.data(), 1);
    for (long s = 0; s < g.V; ++s) {
        for (int k = 0; k < 12; ++k) {
            cd v = (g.par[s] == 1) ? b[s * 12 + k] - src[s * 12 + k] : b[s * 12 + k];
            src[s * 12 + k] = (k >= 6) ? -v : v;
        }
    }
    eo_apply(eo, src_eo.data(), src.data());
    orc_op_desc sq = *d;
    sq.kind = K_EO_CLOVER;
    sq.power = 2;
    sq.gamma5 = 1;
    int ok = orc_bicgstab(&sq, reinterpret_cast<const double*>(src_eo.data()), reinterpret_cast<double*>(xo.data()),
                          x0, precision, max_steps, info);
    eo_hop(eo, xe.data(), xo.data(), 0);
    for (long s = 0; s < g.V; ++s)
        for (int k = 0; k < 12; ++k)
            xe[s * 12 + k] = (g.par[s] == 0) ? b[s * 12 + k] - xe[s * 12 + k] : xo[s * 12 + k];
    eo_inv_even(eo, xe.data());
    std::copy(xe.begin(), xe.end(), reinterpret_cast<cd*>(xout));
    return ok;
}

// MesonCorrelator::execute, connected part (MesonCorrelator.cpp:22-199).  The
// 12 unit sources at `site` (StochasticEstimator.cpp:149-161, there the
// origin) are solved with PreconditionedBiCGStab on D (gamma5 off, :71): the
// even-odd driver for the clover operator (kind 2) on a lattice with all
// extents even, plain BiCGStab otherwise (PreconditionedBiCGStab.cpp:43 falls
// back to it); the propagator column of
// source (alpha, c) is c * 4 + alpha (:94-98).  The pion (:104-118), scalar
// (:121-139) and pseudoscalar-axial (:142-160) contractions are summed per
// timeslice in long double, normalised by 4 kappa^2 (:162), with the PCAC
// ratio of the raw sums (:189-190).  Every output is indexed by the sink time
// relative to the source's, as the reference's source at t = 0 on links
// shifted by t_source_origin (:54-66).  out: 4 x Lt (pion, scalar, ps, pcac).
// Returns 1 when all 12 solves converged.
int orc_meson_correlators(const orc_op_desc* d, const int* site, double precision, unsigned max_steps, double* out) {
    const Geo& g = geo(d->dims);
    const long V = g.V;
    const int T = g.L[3];
    // gamma5 gamma_3 and gamma_3 (Gamma.cpp:8-9, rows listed)
    const cd G53[4][4] = {{0, 0, 1, 0}, {0, 0, 0, 1}, {-1, 0, 0, 0}, {0, -1, 0, 0}};
    const cd G3[4][4] = {{0, 0, 1, 0}, {0, 0, 0, 1}, {1, 0, 0, 0}, {0, 1, 0, 0}};
    std::vector<std::vector<cd>> P(12, std::vector<cd>(V * 12));
    const long idx = (((long)site[0] * g.L[1] + site[1]) * g.L[2] + site[2]) * T + site[3];
    orc_op_desc dd = *d;
    dd.power = 1;
    dd.gamma5 = 0;
    int all_ok = 1;
    for (int alpha = 0; alpha < 4; ++alpha)
        for (int c = 0; c < 3; ++c) {
            std::vector<cd> src(V * 12, cd(0, 0));
            src[idx * 12 + 3 * alpha + c] = 1;
            orc_solve_info info;
            double* x = reinterpret_cast<double*>(P[c * 4 + alpha].data());
            const double* b = reinterpret_cast<const double*>(src.data());
            int ok = dd.kind == K_CLOVER ? orc_precond_bicgstab(&dd, b, x, nullptr, precision, max_steps, &info)
                                         : orc_bicgstab(&dd, b, x, nullptr, precision, max_steps, &info);
            all_ok = all_ok && ok;
        }
    auto S = [&](int col, long s, int mu, int dcol) { return P[col][s * 12 + 3 * mu + dcol]; };
    std::vector<long double> pion(T, 0.0L), scalar(T, 0.0L), ps(T, 0.0L);
    for (int alpha = 0; alpha < 4; ++alpha)
        for (int c = 0; c < 3; ++c)
            for (long s = 0; s < V; ++s)
                for (int mu = 0; mu < 4; ++mu)
                    for (int dc = 0; dc < 3; ++dc) pion[s % T] += std::norm(S(c * 4 + alpha, s, mu, dc));
    for (int alpha = 0; alpha < 4; ++alpha)
        for (int beta = 0; beta < 4; ++beta)
            for (int c = 0; c < 3; ++c)
                for (long s = 0; s < V; ++s)
                    for (int mu = 0; mu < 4; ++mu)
                        for (int nu = 0; nu < 4; ++nu)
                            for (int dc = 0; dc < 3; ++dc)
                                scalar[s % T] += std::real(G53[mu][nu] * G53[alpha][beta] *
                                                           std::conj(S(c * 4 + alpha, s, mu, dc)) *
                                                           S(c * 4 + beta, s, nu, dc));
    for (int alpha = 0; alpha < 4; ++alpha)
        for (int beta = 0; beta < 4; ++beta)
            for (int c = 0; c < 3; ++c)
                for (long s = 0; s < V; ++s)
                    for (int mu = 0; mu < 4; ++mu)
                        for (int dc = 0; dc < 3; ++dc) {
                            cd v = -G3[alpha][beta] * std::conj(S(c * 4 + alpha, s, mu, dc)) * S(c * 4 + beta, s, mu, dc);
                            v += G3[alpha][beta] * std::conj(S(c * 4 + mu, s, alpha, dc)) * S(c * 4 + mu, s, beta, dc);
                            ps[s % T] += std::real(v);
                        }
    const long double factor = 4.0L * d->kappa * d->kappa;
    const int t0 = site[3];
    for (int t = 0; t < T; ++t) {
        const int a = (t0 + t) % T;
        const int am = (t0 + (t == 0 ? T - 1 : t - 1)) % T, ap = (t0 + (t == T - 1 ? 0 : t + 1)) % T;
        out[t] = (double)(factor * pion[a]);
        out[T + t] = (double)(factor * scalar[a]);
        out[2 * T + t] = (double)(factor * ps[a]);
        out[3 * T + t] = (double)((ps[am] - ps[ap]) / (4.0L * pion[a]));
    }
    return all_ok;
}

}  // extern "C"

// ---------------------------------------------------------------- stout smearing
// StoutSmearing::smearing (source/utils/StoutSmearing.cpp:12-20) with smearLink
// (:55-62), WilsonGaugeAction::staple (source/actions/WilsonGaugeAction.cpp:9-20)
// and ExponentialMap::exp (source/utils/ExpMap.h:100-140 for SU(3), :142-155 for
// SU(2)), on fundamental gauge links V x [4][N x N column-major], N = 2 or 3.
// The reference divides by zero when the exponent is exactly zero (cold links,
// rho = 0); the restatement returns the identity there, the limit.

namespace {
struct MN {
    int n;
    cd a[9];
    cd& operator()(int i, int j) { return a[i + n * j]; }
    const cd& operator()(int i, int j) const { return a[i + n * j]; }
};
MN mn_zero(int n) {
    MN m;
    m.n = n;
    for (int k = 0; k < 9; ++k) m.a[k] = 0;
    return m;
}
MN mn_load(int n, const cd* p) {
    MN m = mn_zero(n);
    for (int k = 0; k < n * n; ++k) m.a[k] = p[k];
    return m;
}
MN mn_mul(const MN& A, const MN& B) {
    MN C = mn_zero(A.n);
    for (int i = 0; i < A.n; ++i)
        for (int j = 0; j < A.n; ++j) {
            cd s = 0;
            for (int k = 0; k < A.n; ++k) s += A(i, k) * B(k, j);
            C(i, j) = s;
        }
    return C;
}
MN mn_dag(const MN& A) {
    MN C = mn_zero(A.n);
    for (int i = 0; i < A.n; ++i)
        for (int j = 0; j < A.n; ++j) C(i, j) = std::conj(A(j, i));
    return C;
}
MN mn_axpy(cd a, const MN& X, const MN& Y) {  // a X + Y
    MN C = mn_zero(X.n);
    for (int k = 0; k < X.n * X.n; ++k) C.a[k] = a * X.a[k] + Y.a[k];
    return C;
}

// ExponentialMap::exp of an anti-hermitian traceless X
MN expmap(const MN& X) {
    const int n = X.n;
    MN R = mn_zero(n);
    if (n == 2) {
        const double a = X(0, 0).imag(), b = X(0, 1).real(), c = X(0, 1).imag();
        const double rho = std::sqrt(a * a + b * b + c * c);
        if (rho == 0.0) {
            R(0, 0) = R(1, 1) = 1.0;
            return R;
        }
        const cd phi(std::cos(rho), std::sin(rho));
        R(0, 0) = (-(a - rho) * std::conj(phi) + (a + rho) * phi) / (2. * rho);
        R(0, 1) = (cd(-c, b) * std::conj(phi) + cd(c, -b) * phi) / (2. * rho);
        R(1, 0) = -std::conj(R(0, 1));
        R(1, 1) = std::conj(R(0, 0));
        return R;
    }
    // Q = -i X hermitian traceless; c0 = det Q, c1 = tr(Q^2)/2 (Cayley-Hamilton)
    MN Q = mn_zero(3);
    for (int k = 0; k < 9; ++k) Q.a[k] = cd(X.a[k].imag(), -X.a[k].real());
    const cd det = Q(0, 0) * (Q(1, 1) * Q(2, 2) - Q(1, 2) * Q(2, 1)) - Q(0, 1) * (Q(1, 0) * Q(2, 2) - Q(1, 2) * Q(2, 0)) +
                   Q(0, 2) * (Q(1, 0) * Q(2, 1) - Q(1, 1) * Q(2, 0));
    const MN X2 = mn_mul(X, X);
    double c0 = det.real();
    const double c1 = -0.5 * (X2(0, 0) + X2(1, 1) + X2(2, 2)).real();
    if (c1 == 0.0) {
        for (int i = 0; i < 3; ++i) R(i, i) = 1.0;
        return R;
    }
    bool inv = false;
    if (c0 < 0) {
        c0 = -c0;
        inv = true;
    }
    const double c0max = (2. * c1 / 3.) * std::sqrt(c1 / 3.);
    const double theta = std::acos(std::min(1.0, c0 / c0max));
    const double u = std::sqrt(c1 / 3.) * std::cos(theta / 3.);
    const double w = std::sqrt(c1) * std::sin(theta / 3.);
    const cd e2(std::cos(2. * u), std::sin(2. * u));
    const cd e1(std::cos(u), -std::sin(u));
    const double xi = std::fabs(w) < 0.05 ? 1. - (1. / 6.) * w * w * (1. - (1. / 20.) * w * w * (1. - (1. / 42.) * w * w))
                                          : std::sin(w) / w;
    const cd ii(0., 1.);
    const cd h0 = (u * u - w * w) * e2 + e1 * (8. * u * u * std::cos(w) + 2. * u * ii * (3. * u * u + w * w) * xi);
    const cd h1 = 2. * u * e2 - e1 * (2. * u * std::cos(w) - ii * (3. * u * u - w * w) * xi);
    const cd h2 = e2 - e1 * (std::cos(w) + 3. * u * ii * xi);
    const double den = 9. * u * u - w * w;
    cd f0 = h0 / den, f1 = h1 / den, f2 = h2 / den;
    if (inv) {
        f0 = std::conj(f0);
        f1 = -std::conj(f1);
        f2 = std::conj(f2);
    }
    // f0 + f1 Q + f2 Q^2 with Q = -i X, Q^2 = -X^2
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) R(i, j) = -f2 * X2(i, j) + f1 * Q(i, j) + (i == j ? f0 : cd(0, 0));
    return R;
}

MN gauge_link(const double* U, int n, long s, int mu) {
    return mn_load(n, reinterpret_cast<const cd*>(U) + (s * 4 + mu) * n * n);
}
}  // namespace

extern "C" {

void orc_expmap(int n, const double* X, double* out) {
    MN R = expmap(mn_load(n, reinterpret_cast<const cd*>(X)));
    std::memcpy(out, R.a, sizeof(cd) * n * n);
}

// one level of stout smearing; staple_out (or NULL) receives the staples
void orc_stout_smear(const int* dims, int n, double rho, const double* in, double* out, double* staple_out) {
    const Geo& g = geo(dims);
#pragma omp parallel for
    for (long s = 0; s < g.V; ++s)
        for (int mu = 0; mu < 4; ++mu) {
            MN st = mn_zero(n);
            const long smu = g.up[s * 4 + mu];
            for (int nu = 0; nu < 4; ++nu) {
                if (nu == mu) continue;
                const long snu = g.up[s * 4 + nu], sdn = g.dn[s * 4 + nu];
                const long smudn = g.up[sdn * 4 + mu];
                // U_nu(x+mu) U_mu(x+nu)^+ U_nu(x)^+
                MN a = mn_mul(mn_mul(gauge_link(in, n, smu, nu), mn_dag(gauge_link(in, n, snu, mu))),
                              mn_dag(gauge_link(in, n, s, nu)));
                // U_nu(x+mu-nu)^+ U_mu(x-nu)^+ U_nu(x-nu)
                MN b = mn_mul(mn_mul(mn_dag(gauge_link(in, n, smudn, nu)), mn_dag(gauge_link(in, n, sdn, mu))),
                              gauge_link(in, n, sdn, nu));
                st = mn_axpy(1.0, a, mn_axpy(1.0, b, st));
            }
            if (staple_out) std::memcpy(staple_out + (s * 4 + mu) * 2 * n * n, st.a, sizeof(cd) * n * n);
            const MN Umu = gauge_link(in, n, s, mu);
            // Omega = rho C U^+, C = staple^+; iStout = (Omega^+ - Omega)/2; X = -(iStout - tr/n)
            const MN Om = mn_mul(mn_dag(st), mn_dag(Umu));
            MN X = mn_zero(n);
            cd tr = 0;
            for (int i = 0; i < n; ++i)
                for (int j = 0; j < n; ++j) {
                    X(i, j) = -0.5 * rho * (std::conj(Om(j, i)) - Om(i, j));
                    if (i == j) tr += X(i, j);
                }
            for (int i = 0; i < n; ++i) X(i, i) -= tr / (double)n;
            const MN E = mn_mul(expmap(X), Umu);
            std::memcpy(out + (s * 4 + mu) * 2 * n * n, E.a, sizeof(cd) * n * n);
        }
}


// ---------------------------------------------------------------- fermion force
// TwoFlavorFermionAction::force (source/actions/TwoFlavorFermionAction.cpp:17-20)
// = -FermionForce::force (source/hmc_forces/FermionForce.cpp:33-58) of
// DiracWilsonFermionForce::derivative (source/hmc_forces/DiracWilsonFermionForce.cpp:10-54)
// for every link.  flinks: the operator's fermion links (boundary sign folded
// in, V x 4 x 3x3 complex column-major; real adjoint links given as complex);
// glinks: the gauge links V x 4 x N x N (no boundary sign); X, Y: spinors.
// rep 0: SU(3) fundamental fermions (N = 3); rep 1: SU(2) with adjoint
// fermions (N = 2; the adjoint link of the projection is toAdjoint of the
// gauge link, ConvertLattice.h:24-34).  Generators: lambda_a / 2 for SU(3),
// sigma_a / 2 for SU(2) and (T^adj_a)_{bc} = -i eps_{abc} for the adjoint
// (source/utils/LieGenerators.h).
void orc_wilson_fermion_force(const int* dims, int rep, double kappa, const double* flinks_, const double* glinks_,
                              const double* X_, const double* Y_, double* force_) {
    const Geo& g = geo(dims);
    const cd* W = reinterpret_cast<const cd*>(flinks_);
    const cd* G = reinterpret_cast<const cd*>(glinks_);
    const cd* X = reinterpret_cast<const cd*>(X_);
    const cd* Y = reinterpret_cast<const cd*>(Y_);
    cd* F = reinterpret_cast<cd*>(force_);
    const int N = rep == 0 ? 3 : 2;
    const int NG = N * N - 1;
    // gamma5 (1 -+ gamma_mu) from the chiral gamma matrices
    cd g5m[4][4][4], g5p[4][4][4];
    for (int mu = 0; mu < 4; ++mu)
        for (int a = 0; a < 4; ++a)
            for (int b = 0; b < 4; ++b) {
                const double s5 = a < 2 ? 1.0 : -1.0;
                const cd id = a == b ? cd(1, 0) : cd(0, 0);
                g5m[mu][a][b] = s5 * (id - GAMMA[mu][a][b]);
                g5p[mu][a][b] = s5 * (id + GAMMA[mu][a][b]);
            }
    // generators of the gauge group (fundamental) and of the fermion representation
    std::vector<MN> Tg(NG), Tf(NG);
    if (N == 3) {
        for (int a = 0; a < 8; ++a) Tg[a] = mn_zero(3);
        const double r3 = 1.0 / std::sqrt(3.0);
        Tg[0](0, 1) = Tg[0](1, 0) = 0.5;
        Tg[1](0, 1) = cd(0, -0.5); Tg[1](1, 0) = cd(0, 0.5);
        Tg[2](0, 0) = 0.5; Tg[2](1, 1) = -0.5;
        Tg[3](0, 2) = Tg[3](2, 0) = 0.5;
        Tg[4](0, 2) = cd(0, -0.5); Tg[4](2, 0) = cd(0, 0.5);
        Tg[5](1, 2) = Tg[5](2, 1) = 0.5;
        Tg[6](1, 2) = cd(0, -0.5); Tg[6](2, 1) = cd(0, 0.5);
        Tg[7](0, 0) = Tg[7](1, 1) = 0.5 * r3; Tg[7](2, 2) = -r3;
        Tf = Tg;
    } else {
        for (int a = 0; a < 3; ++a) {
            Tg[a] = mn_zero(2);
            Tf[a] = mn_zero(3);
        }
        Tg[0](0, 1) = Tg[0](1, 0) = 0.5;
        Tg[1](0, 1) = cd(0, -0.5); Tg[1](1, 0) = cd(0, 0.5);
        Tg[2](0, 0) = 0.5; Tg[2](1, 1) = -0.5;
        const int eps[3][3][3] = {{{0, 0, 0}, {0, 0, 1}, {0, -1, 0}}, {{0, 0, -1}, {0, 0, 0}, {1, 0, 0}},
                                  {{0, 1, 0}, {-1, 0, 0}, {0, 0, 0}}};
        for (int a = 0; a < 3; ++a)
            for (int b = 0; b < 3; ++b)
                for (int c = 0; c < 3; ++c) Tf[a](b, c) = cd(0, -(double)eps[a][b][c]);
    }
#pragma omp parallel for
    for (long s = 0; s < g.V; ++s)
        for (int mu = 0; mu < 4; ++mu) {
            const long sp = g.up[s * 4 + mu];
            const cd* Wl = W + (s * 4 + mu) * 9;
            MN D = mn_zero(3);
            // two orderings: (A, B) = (X, Y) and (Y, X)
            for (int ord = 0; ord < 2; ++ord) {
                const cd* A = ord == 0 ? X : Y;
                const cd* B = ord == 0 ? Y : X;
                for (int al = 0; al < 4; ++al) {
                    // from A(x)^+ g5 (1 - g_mu) W B(x+mu): tmp = -kappa sum_b g5m B(x+mu)_b
                    cd tmp[3] = {0, 0, 0}, tmq[3] = {0, 0, 0};
                    for (int be = 0; be < 4; ++be) {
                        for (int c = 0; c < 3; ++c) {
                            tmp[c] += -kappa * g5m[mu][al][be] * B[sp * 12 + 3 * be + c];
                            // from A(x+mu)^+ g5 (1 + g_mu) W^+ B(x): tmq = -kappa sum_b g5p W^+ B(x)_b
                            cd wb = 0;
                            for (int k = 0; k < 3; ++k) wb += std::conj(Wl[c * 3 + k]) * B[s * 12 + 3 * be + k];
                            tmq[c] += -kappa * g5p[mu][al][be] * wb;
                        }
                    }
                    // W A(x+mu)_alpha
                    cd wa[3];
                    for (int i = 0; i < 3; ++i) {
                        wa[i] = 0;
                        for (int k = 0; k < 3; ++k) wa[i] += Wl[i + 3 * k] * A[sp * 12 + 3 * al + k];
                    }
                    // tensor(x, y)(i, j) = y_i conj(x_j)
                    for (int i = 0; i < 3; ++i)
                        for (int j = 0; j < 3; ++j)
                            D(i, j) += tmp[i] * std::conj(A[s * 12 + 3 * al + j]) - tmq[i] * std::conj(wa[j]);
                }
            }
            // projection: result = -i sum_a Im tr(D T^f_a L) T^g_a with L the fermion-representation
            // link of the gauge link (no boundary sign); the action's force is minus it
            MN L = mn_zero(3);
            const cd* Gl = G + (s * 4 + mu) * N * N;
            if (N == 3) {
                for (int k = 0; k < 9; ++k) L.a[k] = Gl[k];
            } else {
                const MN U2 = mn_load(2, Gl);
                for (int a = 0; a < 3; ++a)
                    for (int b = 0; b < 3; ++b) {
                        const MN P = mn_mul(mn_mul(mn_mul(Tg[a], U2), Tg[b]), mn_dag(U2));
                        L(a, b) = 2.0 * (P(0, 0) + P(1, 1)).real();
                    }
            }
            MN R = mn_zero(N);
            for (int a = 0; a < NG; ++a) {
                const MN M = mn_mul(mn_mul(D, Tf[a]), L);
                cd tr = 0;
                for (int i = 0; i < 3; ++i) tr += M(i, i);
                R = mn_axpy(cd(0, tr.imag()), Tg[a], R);  // -(-i Im tr) T_a
            }
            std::memcpy(F + (s * 4 + mu) * N * N, R.a, sizeof(cd) * N * N);
        }
}

}  // extern "C"
