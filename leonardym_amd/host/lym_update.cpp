// lym_update.cpp -- implementation of the C++ drop-in (lym_update.h) on the C ABI.
#include "lym_update.h"

#include <cstring>
#include <memory>

namespace Update {

namespace {

void check(int rc, const char* what) {
    if (rc != LYM_OK) throw std::runtime_error(std::string(what) + ": " + lym_last_error());
}

inline int modulus(int v, int m) { return v >= m ? v - m : (v < 0 ? v + m : v); }

struct Tables {
    std::vector<int> up, dn, coord;
    lym_ctx* ctx = nullptr;
    int dims[4] = {0, 0, 0, 0};
    int rep = -1, bc = 0;
};
Tables& tables() {
    static Tables t;
    return t;
}

}  // namespace

// ------------------------------------------------------------------ geometry
int LocalLayout::glob_x = 0, LocalLayout::glob_y = 0, LocalLayout::glob_z = 0, LocalLayout::glob_t = 0;
int LocalLayout::localsize = 0, LocalLayout::completesize = 0, LocalLayout::globalVolume = 0;
int LocalLayout::representation = LYM_REP_SU3_FUND;
int LocalLayout::bc_t = -1;

void LocalLayout::initialize() {
    const int d[4] = {glob_x, glob_y, glob_z, glob_t};
    for (int i = 0; i < 4; ++i)
        if (d[i] < 1) throw std::runtime_error("LocalLayout: lattice extents not set");
    globalVolume = glob_x * glob_y * glob_z * glob_t;
    localsize = completesize = globalVolume;
    Tables& T = tables();
    T.up.assign((size_t)globalVolume * 4, 0);
    T.dn.assign((size_t)globalVolume * 4, 0);
    T.coord.assign((size_t)globalVolume * 4, 0);
    for (int x = 0; x < glob_x; ++x)
        for (int y = 0; y < glob_y; ++y)
            for (int z = 0; z < glob_z; ++z)
                for (int t = 0; t < glob_t; ++t) {
                    int s = getGlobalCoordinate(x, y, z, t);
                    int c[4] = {x, y, z, t};
                    for (int mu = 0; mu < 4; ++mu) {
                        int u[4] = {x, y, z, t}, w[4] = {x, y, z, t};
                        u[mu] = c[mu] + 1;
                        w[mu] = c[mu] - 1;
                        T.up[s * 4 + mu] = getGlobalCoordinate(u[0], u[1], u[2], u[3]);
                        T.dn[s * 4 + mu] = getGlobalCoordinate(w[0], w[1], w[2], w[3]);
                        T.coord[s * 4 + mu] = c[mu];
                    }
                }
}

int LocalLayout::getGlobalCoordinate(int x, int y, int z, int t) {
    return glob_t * (glob_z * (glob_y * modulus(x, glob_x) + modulus(y, glob_y)) + modulus(z, glob_z)) +
           modulus(t, glob_t);
}
int LocalLayout::sup(int site, int mu) { return tables().up[site * 4 + mu]; }
int LocalLayout::sdn(int site, int mu) { return tables().dn[site * 4 + mu]; }
int LocalLayout::globalIndexX(int site) { return tables().coord[site * 4]; }
int LocalLayout::globalIndexY(int site) { return tables().coord[site * 4 + 1]; }
int LocalLayout::globalIndexZ(int site) { return tables().coord[site * 4 + 2]; }
int LocalLayout::globalIndexT(int site) { return tables().coord[site * 4 + 3]; }

lym_ctx* LocalLayout::context() {
    Tables& T = tables();
    const int d[4] = {glob_x, glob_y, glob_z, glob_t};
    bool same = T.ctx && T.rep == representation && T.bc == bc_t;
    for (int i = 0; i < 4; ++i) same = same && T.dims[i] == d[i];
    if (!same) {
        if (T.ctx) lym_ctx_destroy(T.ctx);
        T.ctx = nullptr;
        check(lym_ctx_create(d, representation, bc_t, &T.ctx), "lym_ctx_create");
        for (int i = 0; i < 4; ++i) T.dims[i] = d[i];
        T.rep = representation;
        T.bc = bc_t;
    }
    return T.ctx;
}

void LocalLayout::finalize() {
    Tables& T = tables();
    if (T.ctx) lym_ctx_destroy(T.ctx);
    T.ctx = nullptr;
}

// ------------------------------------------------------------------ DiracVector
const int& DiracVector::localsize = LocalLayout::localsize;
const int& DiracVector::completesize = LocalLayout::completesize;
const int& FermionLattice::localsize = LocalLayout::localsize;
const int& FermionLattice::completesize = LocalLayout::completesize;

DiracVector::DiracVector()
    : host_((size_t)LocalLayout::completesize * 12), dev_(nullptr), host_valid_(true), dev_valid_(false) {}

DiracVector::DiracVector(const DiracVector& o) : dev_(nullptr), host_valid_(true), dev_valid_(false) {
    o.syncHost();
    host_ = o.host_;
}

DiracVector& DiracVector::operator=(const DiracVector& o) {
    if (this == &o) return *this;
    if (o.dev_valid_ && !o.host_valid_) {
        // device-to-device copy keeps the data on the GPU
        lym_spinor* d = deviceOut();
        check(lym_spinor_copy(d, o.dev_), "lym_spinor_copy");
        return *this;
    }
    host_ = o.host_;
    host_valid_ = true;
    dev_valid_ = false;
    return *this;
}

DiracVector::~DiracVector() {
    if (dev_) lym_spinor_destroy(dev_);
}

void DiracVector::syncHost() const {
    if (host_valid_) return;
    host_.resize((size_t)LocalLayout::completesize * 12);
    check(lym_spinor_download(dev_, reinterpret_cast<double*>(host_.data())), "lym_spinor_download");
    host_valid_ = true;
}

GaugeVector* DiracVector::operator[](int site) {
    syncHost();
    dev_valid_ = false;
    return reinterpret_cast<GaugeVector*>(host_.data() + (size_t)site * 12);
}
const GaugeVector* DiracVector::operator[](int site) const {
    syncHost();
    return reinterpret_cast<const GaugeVector*>(host_.data() + (size_t)site * 12);
}
complex* DiracVector::getRawData() {
    syncHost();
    dev_valid_ = false;
    return host_.data();
}
const complex* DiracVector::getRawData() const {
    syncHost();
    return host_.data();
}

lym_spinor* DiracVector::device() const {
    if (!dev_) check(lym_spinor_create(LocalLayout::context(), &dev_), "lym_spinor_create");
    if (!dev_valid_) {
        check(lym_spinor_upload(dev_, reinterpret_cast<const double*>(host_.data())), "lym_spinor_upload");
        dev_valid_ = true;
    }
    return dev_;
}

lym_spinor* DiracVector::deviceOut() {
    if (!dev_) check(lym_spinor_create(LocalLayout::context(), &dev_), "lym_spinor_create");
    dev_valid_ = true;
    host_valid_ = false;
    return dev_;
}

// ------------------------------------------------------------------ links
FermionLattice::FermionLattice() : links_((size_t)LocalLayout::completesize * 36), version_(0) {}

void switchAntiperiodicBc(FermionLattice& lattice) {
    for (int site = 0; site < LocalLayout::completesize; ++site)
        if (LocalLayout::globalIndexT(site) == LocalLayout::glob_t - 1)
            for (int i = 0; i < 3; ++i)
                for (int j = 0; j < 3; ++j) lattice(site, 3, i, j) = -lattice(site, 3, i, j);
}

// ------------------------------------------------------------------ AlgebraUtils
long_real_t AlgebraUtils::squaredNorm(const reduced_dirac_vector_t& v) {
    double r;
    check(lym_spinor_squared_norm(v.device(), &r), "squaredNorm");
    return r;
}
std::complex<long_real_t> AlgebraUtils::dot(const reduced_dirac_vector_t& a, const reduced_dirac_vector_t& b) {
    double re, im;
    check(lym_spinor_dot(a.device(), b.device(), &re, &im), "dot");
    return std::complex<long_real_t>(re, im);
}
std::complex<long_real_t> AlgebraUtils::gamma5dot(const reduced_dirac_vector_t& a, const reduced_dirac_vector_t& b) {
    double re, im;
    check(lym_spinor_gamma5dot(a.device(), b.device(), &re, &im), "gamma5dot");
    return std::complex<long_real_t>(re, im);
}
long_real_t AlgebraUtils::differenceNorm(const reduced_dirac_vector_t& a, const reduced_dirac_vector_t& b) {
    double r;
    check(lym_spinor_difference_norm(a.device(), b.device(), &r), "differenceNorm");
    return r;
}
void AlgebraUtils::setToZero(reduced_dirac_vector_t& v) { check(lym_spinor_set_zero(v.deviceOut()), "setToZero"); }
void AlgebraUtils::gamma5(reduced_dirac_vector_t& v) {
    lym_spinor* d = v.device();
    v.deviceOut();
    check(lym_spinor_gamma5(d), "gamma5");
}
void AlgebraUtils::generateRandomComplexGaussianVector(reduced_dirac_vector_t& v, unsigned long seed) {
    check(lym_gen_gaussian_spinor(LocalLayout::completesize, seed, reinterpret_cast<double*>(v.getRawData())),
          "generateRandomComplexGaussianVector");
}

// ------------------------------------------------------------------ DiracOperator
DiracOperator::DiracOperator() : DiracOperator(LYM_OP_WILSON, 1) {}

DiracOperator::DiracOperator(int kind, unsigned power)
    : kappa(0.), gamma5(true), kind_(kind), power_(power), csw_(0.), op_(nullptr), gauge_(nullptr),
      gauge_version_(0), have_lattice_(false) {}

DiracOperator::DiracOperator(const extended_fermion_lattice_t& l, double k, bool g) : DiracOperator(LYM_OP_WILSON, 1) {
    kappa = k;
    gamma5 = g;
    setLattice(l);
}

DiracOperator::~DiracOperator() {
    if (op_) lym_op_destroy(op_);
    if (gauge_) lym_gauge_destroy(gauge_);
}

DiracOperator* DiracOperator::getInstance(const std::string& name, unsigned int power, const StorageParameters& p,
                                          const std::string& basename) {
    DiracOperator* r = nullptr;
    if (power != 1 && power != 2) throw std::runtime_error("Power of the Dirac Wilson Operator not supported!");
    if (name == "DiracWilson") {
        r = power == 1 ? (DiracOperator*)new DiracWilsonOperator() : new SquareDiracWilsonOperator();
    } else if (name == "BasicDiracWilson") {
        r = power == 1 ? (DiracOperator*)new BasicDiracWilsonOperator() : new BasicSquareDiracWilsonOperator();
    } else if (name == "Improved") {
        r = power == 1 ? (DiracOperator*)new ImprovedDiracWilsonOperator() : new SquareImprovedDiracWilsonOperator();
        r->setCSWValue(p.get<double>(basename + "csw"));
    } else {
        throw std::runtime_error("Dirac Wilson Operator" + name + " not supported!");
    }
    r->setKappa(p.get<double>(basename + "kappa"));
    r->name = name;
    return r;
}

DiracOperator* DiracOperator::getSquare(DiracOperator* d) {
    if (d->power_ != 1) throw std::runtime_error("Power of the Dirac Wilson Operator not supported!");
    DiracOperator* r = new DiracOperator(d->kind_, 2);
    r->kappa = d->kappa;
    r->csw_ = d->csw_;
    r->name = d->name;
    r->gamma5 = d->gamma5;
    if (d->have_lattice_) r->setLattice(d->lattice);
    return r;
}

DiracOperator* DiracOperator::getSquareRoot(DiracOperator* d) {
    if (d->power_ != 2) throw std::runtime_error("Power of the Dirac Wilson Operator not supported!");
    DiracOperator* r = new DiracOperator(d->kind_, 1);
    r->kappa = d->kappa;
    r->csw_ = d->csw_;
    r->name = d->name;
    r->gamma5 = d->gamma5;
    if (d->have_lattice_) r->setLattice(d->lattice);
    return r;
}

void DiracOperator::registerParameters(std::map<std::string, Option>& desc, const std::string& basename) {
    desc[basename + "dirac_operator"] = Option(basename + "dirac_operator", "DiracWilson", "The name of the Dirac operator");
    desc[basename + "kappa"] = Option(basename + "kappa", 0.125, "set the value of the kappa for the Dirac Wilson operator");
    desc[basename + "csw"] = Option(basename + "csw", 1.0, "set the value of the clover term for the Dirac Wilson operator");
}

void DiracOperator::setKappa(real_t k) {
    kappa = k;
    if (op_) check(lym_op_set_kappa(op_, k), "setKappa");
}

void DiracOperator::setCSWValue(real_t csw) {
    csw_ = csw;
    if (op_) check(lym_op_set_csw(op_, csw), "setCSW");
}

void DiracOperator::setLattice(const extended_fermion_lattice_t& l) {
    lattice = l;
    have_lattice_ = true;
    gauge_version_ = ~0ul;  // force a fresh device copy
}

lym_op* DiracOperator::handle() {
    if (!have_lattice_) throw std::runtime_error(name + ": setLattice was not called");
    lym_ctx* ctx = LocalLayout::context();
    if (!gauge_ || gauge_version_ != lattice.version()) {
        if (gauge_) lym_gauge_destroy(gauge_);
        gauge_ = nullptr;
        if (LocalLayout::representation == LYM_REP_SU2_ADJ) {
            std::vector<double> re((size_t)LocalLayout::completesize * 36);
            const complex* c = lattice.getRawData();
            for (size_t i = 0; i < re.size(); ++i) re[i] = c[i].real();
            check(lym_gauge_create(ctx, re.data(), 1, LYM_LINKS_FULL, &gauge_), "lym_gauge_create");
        } else {
            check(lym_gauge_create(ctx, reinterpret_cast<const double*>(lattice.getRawData()), 1, LYM_LINKS_FULL,
                                   &gauge_),
                  "lym_gauge_create");
        }
        gauge_version_ = lattice.version();
        if (op_) check(lym_op_set_gauge(op_, gauge_), "setLattice");
    }
    if (!op_) {
        check(lym_op_create(ctx, kind_, power_, kappa, csw_, gamma5, &op_), "lym_op_create");
        check(lym_op_set_gauge(op_, gauge_), "setLattice");
    }
    check(lym_op_set_kappa(op_, kappa), "setKappa");
    if (kind_ == LYM_OP_CLOVER || kind_ == LYM_OP_EO_CLOVER) check(lym_op_set_csw(op_, csw_), "setCSW");
    check(lym_op_set_gamma5(op_, gamma5), "setGamma5");
    return op_;
}

void DiracOperator::multiply(reduced_dirac_vector_t& output, const reduced_dirac_vector_t& input) {
    lym_op* op = handle();
    lym_spinor* in = input.device();
    if (&output == &input) {
        reduced_dirac_vector_t tmp;
        check(lym_op_multiply(op, tmp.deviceOut(), in), "multiply");
        output = tmp;
        return;
    }
    check(lym_op_multiply(op, output.deviceOut(), in), "multiply");
}

void DiracOperator::multiplyAdd(reduced_dirac_vector_t& output, const reduced_dirac_vector_t& v1,
                                const reduced_dirac_vector_t& v2, const complex& alpha) {
    lym_op* op = handle();
    lym_spinor* a = v1.device();
    lym_spinor* b = v2.device();
    if (&output == &v1 || &output == &v2) {
        reduced_dirac_vector_t tmp;
        check(lym_op_multiply_add(op, tmp.deviceOut(), a, b, alpha.real(), alpha.imag()), "multiplyAdd");
        output = tmp;
        return;
    }
    check(lym_op_multiply_add(op, output.deviceOut(), a, b, alpha.real(), alpha.imag()), "multiplyAdd");
}

// ------------------------------------------------------------------ solvers
bool Solver::run(int kind, bool even_odd, DiracOperator* dirac, const reduced_dirac_vector_t& source,
                 reduced_dirac_vector_t& solution, reduced_dirac_vector_t const* initial_guess) {
    lym_op* op = dirac->handle();
    lym_ctx* ctx = LocalLayout::context();
    lym_solver* s = nullptr;
    check(lym_solver_create(ctx, kind, &s), "lym_solver_create");
    std::unique_ptr<lym_solver, void (*)(lym_solver*)> guard(s, lym_solver_destroy);
    check(lym_solver_set_precision(s, precision), "setPrecision");
    check(lym_solver_set_max_steps(s, maxSteps), "setMaximumSteps");
    check(lym_solver_set_even_odd(s, even_odd ? 1 : 0), "setUseEvenOddPreconditioning");
    lym_spinor* b = source.device();
    lym_spinor* g = initial_guess ? initial_guess->device() : nullptr;
    lym_solve_info info;
    if (&solution == &source || &solution == initial_guess) {
        reduced_dirac_vector_t tmp;
        check(lym_solver_solve(s, op, b, tmp.deviceOut(), g, &info), "solve");
        solution = tmp;
    } else {
        check(lym_solver_solve(s, op, b, solution.deviceOut(), g, &info), "solve");
    }
    lastSteps = info.steps;
    lastError = info.last_error;
    return info.converged != 0;
}

}  // namespace Update
