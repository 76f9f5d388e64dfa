// lym_update.h -- C++ drop-in for LeonardYM's fermion-solve interface, backed
// by the MI355X library through the C ABI of include/lym.h.
//
// Names, signatures and argument meanings follow the reference so that its
// callers (e.g. source/correlators/MesonCorrelator.cpp:25-99,
// source/actions/TwoFlavorFermionAction.cpp:21-50) compile against it:
//   program_options/Option.h:10-51, StorageParameters.h:17-73
//   MPILattice/LocalLayout.h:34-170 (static geometry, sup/sdn tables)
//   algebra_utils/AlgebraUtils.h:17-275
//   dirac_operators/DiracOperator.h:11-140 and the Wilson, clover, even-odd
//     and Square operators
//   inverters/Solver.h:8-67, ConjugateGradient.h:8-36,
//     BiConjugateGradient.h:8-30, PreconditionedBiCGStab.h:9-31
// Differences: configuration errors throw (the reference calls exit()),
// fields keep a device copy that is synchronised lazily, and getForce() returns
// nullptr (forces are not on this path).
#pragma once

#include <array>
#include <complex>
#include <map>
#include <sstream>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/lym.h"

namespace Update {

typedef double real_t;
typedef long double long_real_t;
typedef std::complex<real_t> complex;
const int diracVectorLength = 3;
typedef std::array<complex, diracVectorLength> GaugeVector;

// ------------------------------------------------------------------ options
class NotFoundOption : public std::runtime_error {
public:
    explicit NotFoundOption(const std::string& name) : std::runtime_error(name + ": option not found!") {}
};

class BadOption : public std::runtime_error {
public:
    explicit BadOption(const std::string& name)
        : std::runtime_error("Error, option " + name + " wrongly specified! (bad type?)") {}
};

class Option {
public:
    Option() {}
    explicit Option(const std::string& name) : name_(name) {}
    Option(const std::string& name, const std::string& help) : name_(name), help_(help) {}
    template <typename T>
    Option(const std::string& name, const T& value, const std::string& help) : name_(name), help_(help) {
        std::ostringstream os;
        os.precision(17);
        os << std::boolalpha << value;
        value_ = os.str();
    }
    Option(const std::string& name, const char* value, const std::string& help)
        : value_(value), name_(name), help_(help) {}
    Option& operator=(const std::string& v) {
        value_ = v;
        return *this;
    }
    template <typename T>
    T as() const;
    std::string getHelp() const { return name_ + ": " + help_; }
    std::string toString() const { return name_ + " = " + value_ + " # " + help_; }
    const std::string& raw() const { return value_; }

private:
    std::string value_, name_, help_;
};

template <typename T>
T Option::as() const {
    std::istringstream is(value_);
    T v;
    if (!(is >> std::boolalpha >> v)) throw BadOption(name_);
    return v;
}
template <>
inline std::string Option::as<std::string>() const {
    return value_;
}

class StorageParameters {
public:
    StorageParameters() {}
    explicit StorageParameters(const std::map<std::string, Option>& vm) : vm_(vm) {}
    template <typename T>
    T get(const std::string& name) const {
        auto it = vm_.find(name);
        if (it == vm_.end()) throw NotFoundOption(name);
        return it->second.as<T>();
    }

private:
    std::map<std::string, Option> vm_;
};

// ------------------------------------------------------------------ geometry
// Static single-process layout (LocalLayout.h): set glob_* then initialize().
struct LocalLayout {
    static int glob_x, glob_y, glob_z, glob_t;
    static int localsize, completesize, globalVolume;
    static int representation;  // LYM_REP_SU3_FUND or LYM_REP_SU2_ADJ
    static int bc_t;            // -1 "fermion_antiperiodic" (the reference's default), +1 periodic
    static void initialize();
    static int getGlobalCoordinate(int x, int y, int z, int t);
    static int sup(int site, int mu);
    static int sdn(int site, int mu);
    static int globalIndexX(int site);
    static int globalIndexY(int site);
    static int globalIndexZ(int site);
    static int globalIndexT(int site);
    static lym_ctx* context();  // device context of the current geometry
    static void finalize();
};

// ------------------------------------------------------------------ fields
// reduced_dirac_vector_t: V x [4 spin] GaugeVector, host storage in the
// reference's raw layout plus a lazily synchronised device copy.
class DiracVector {
public:
    typedef LocalLayout Layout;
    DiracVector();
    DiracVector(const DiracVector& other);
    DiracVector& operator=(const DiracVector& other);
    ~DiracVector();

    GaugeVector* operator[](int site);              // host access (invalidates the device copy)
    const GaugeVector* operator[](int site) const;  // host access
    complex* getRawData();
    const complex* getRawData() const;

    static int sup(int site, int mu) { return LocalLayout::sup(site, mu); }
    static int sdn(int site, int mu) { return LocalLayout::sdn(site, mu); }
    static const int& localsize;
    static const int& completesize;

    // library side
    lym_spinor* device() const;   // current device copy (uploads if the host side changed)
    lym_spinor* deviceOut();      // device copy to be overwritten (host side becomes stale)
    void updateHalo() {}          // single process: nothing to exchange

private:
    void syncHost() const;
    mutable std::vector<complex> host_;
    mutable lym_spinor* dev_;
    mutable bool host_valid_, dev_valid_;
};

// Fermion links, V x 4 x (3x3 column-major): complex for SU(3), real for the
// SU(2) adjoint representation (stored with zero imaginary part).
class FermionLattice {
public:
    typedef LocalLayout Layout;
    FermionLattice();
    complex& operator()(int site, int mu, int i, int j) { ++version_; return links_[((size_t)site * 4 + mu) * 9 + i + 3 * j]; }
    const complex& operator()(int site, int mu, int i, int j) const { return links_[((size_t)site * 4 + mu) * 9 + i + 3 * j]; }
    complex* getRawData() { ++version_; return links_.data(); }
    const complex* getRawData() const { return links_.data(); }
    unsigned long version() const { return version_; }
    static const int& localsize;
    static const int& completesize;

private:
    std::vector<complex> links_;
    unsigned long version_;
};

typedef DiracVector reduced_dirac_vector_t;
typedef DiracVector extended_dirac_vector_t;
typedef FermionLattice extended_fermion_lattice_t;
typedef FermionLattice reduced_fermion_lattice_t;

// switchAntiperiodicBc (LatticeBoundaryConditions.h:8-14)
void switchAntiperiodicBc(FermionLattice& lattice);

// ------------------------------------------------------------------ AlgebraUtils
struct AlgebraUtils {
    static long_real_t squaredNorm(const reduced_dirac_vector_t& v);
    static std::complex<long_real_t> dot(const reduced_dirac_vector_t& a, const reduced_dirac_vector_t& b);
    static std::complex<long_real_t> gamma5dot(const reduced_dirac_vector_t& a, const reduced_dirac_vector_t& b);
    static long_real_t differenceNorm(const reduced_dirac_vector_t& a, const reduced_dirac_vector_t& b);
    static void setToZero(reduced_dirac_vector_t& v);
    static void gamma5(reduced_dirac_vector_t& v);
    // seeded (the reference seeds from the clock)
    static void generateRandomComplexGaussianVector(reduced_dirac_vector_t& v, unsigned long seed = 1);
};

// ------------------------------------------------------------------ operators
class FermionForce;

class DiracOperator {
public:
    DiracOperator();
    DiracOperator(const extended_fermion_lattice_t& lattice, double kappa = 0., bool gamma5 = true);
    virtual ~DiracOperator();
    static DiracOperator* getInstance(const std::string& name, unsigned int power, const StorageParameters& parameters,
                                      const std::string& basename = "");
    static DiracOperator* getSquare(DiracOperator* dirac);
    static DiracOperator* getSquareRoot(DiracOperator* dirac);
    static void registerParameters(std::map<std::string, Option>& desc, const std::string& basename);

    virtual void multiply(reduced_dirac_vector_t& output, const reduced_dirac_vector_t& input);
    virtual void multiplyAdd(reduced_dirac_vector_t& output, const reduced_dirac_vector_t& vector1,
                             const reduced_dirac_vector_t& vector2, const complex& alpha);
    virtual FermionForce* getForce() const { return nullptr; }
    virtual void setKappa(real_t kappa);
    real_t getKappa() const { return kappa; }
    const reduced_fermion_lattice_t* getLattice() const { return &lattice; }
    virtual void setLattice(const extended_fermion_lattice_t& lattice);
    void setGamma5(bool g) { gamma5 = g; }
    bool getGamma5() const { return gamma5; }
    std::string getName() const { return name; }

    // library side
    lym_op* handle();
    int kind() const { return kind_; }
    unsigned power() const { return power_; }
    real_t cswValue() const { return csw_; }

protected:
    DiracOperator(int kind, unsigned power);
    void setCSWValue(real_t csw);
    reduced_fermion_lattice_t lattice;
    real_t kappa;
    bool gamma5;
    std::string name;

private:
    int kind_;
    unsigned power_;
    real_t csw_;
    lym_op* op_;
    lym_gauge* gauge_;
    unsigned long gauge_version_;
    bool have_lattice_;
};

class DiracWilsonOperator : public DiracOperator {
public:
    DiracWilsonOperator() : DiracOperator(LYM_OP_WILSON, 1) {}
};
class BasicDiracWilsonOperator : public DiracOperator {
public:
    BasicDiracWilsonOperator() : DiracOperator(LYM_OP_BASIC_WILSON, 1) {}
};
class SquareDiracWilsonOperator : public DiracOperator {
public:
    SquareDiracWilsonOperator() : DiracOperator(LYM_OP_WILSON, 2) {}
};
class BasicSquareDiracWilsonOperator : public DiracOperator {
public:
    BasicSquareDiracWilsonOperator() : DiracOperator(LYM_OP_BASIC_WILSON, 2) {}
};

class ImprovedDiracWilsonOperator : public DiracOperator {
public:
    ImprovedDiracWilsonOperator() : DiracOperator(LYM_OP_CLOVER, 1) {}
    real_t getCSW() const { return cswValue(); }
    void setCSW(real_t csw) { setCSWValue(csw); }

protected:
    ImprovedDiracWilsonOperator(int kind, unsigned power) : DiracOperator(kind, power) {}
};
class SquareImprovedDiracWilsonOperator : public DiracOperator {
public:
    SquareImprovedDiracWilsonOperator() : DiracOperator(LYM_OP_CLOVER, 2) {}
    real_t getCSW() const { return cswValue(); }
    void setCSW(real_t csw) { setCSWValue(csw); }
};
class EvenOddImprovedDiracWilsonOperator : public ImprovedDiracWilsonOperator {
public:
    EvenOddImprovedDiracWilsonOperator() : ImprovedDiracWilsonOperator(LYM_OP_EO_CLOVER, 1) {}
};
class SquareEvenOddImprovedDiracWilsonOperator : public DiracOperator {
public:
    SquareEvenOddImprovedDiracWilsonOperator() : DiracOperator(LYM_OP_EO_CLOVER, 2) {}
    real_t getCSW() const { return cswValue(); }
    void setCSW(real_t csw) { setCSWValue(csw); }
};

// ------------------------------------------------------------------ solvers
class Solver {
public:
    explicit Solver(const std::string& name = "") : name(name), precision(0.0000000001), maxSteps(1000), lastError(0), lastSteps(0) {}
    virtual ~Solver() {}
    virtual bool solve(DiracOperator*, const reduced_dirac_vector_t&, reduced_dirac_vector_t&,
                       reduced_dirac_vector_t const* = 0) {
        return false;
    }
    void setPrecision(real_t e) { precision = e; }
    real_t getPrecision() const { return precision; }
    void setMaximumSteps(unsigned int m) { maxSteps = m; }
    unsigned int getMaximumSteps() const { return maxSteps; }
    real_t getLastError() const { return lastError; }
    unsigned int getLastSteps() const { return lastSteps; }

protected:
    bool run(int kind, bool even_odd, DiracOperator* dirac, const reduced_dirac_vector_t& source,
             reduced_dirac_vector_t& solution, reduced_dirac_vector_t const* initial_guess);
    std::string name;
    real_t precision;
    unsigned int maxSteps;
    real_t lastError;
    unsigned int lastSteps;
};

class BiConjugateGradient : public Solver {
public:
    using Solver::solve;
    BiConjugateGradient() : Solver("BiConjugateGradient") {}
    bool solve(DiracOperator* dirac, const reduced_dirac_vector_t& source, reduced_dirac_vector_t& solution,
               reduced_dirac_vector_t const* initial_guess = 0) override {
        return run(LYM_SOLVER_BICGSTAB, false, dirac, source, solution, initial_guess);
    }
};

class PreconditionedBiCGStab : public Solver {
public:
    using Solver::solve;
    PreconditionedBiCGStab() : Solver("PreconditionedBiCGStab"), useEvenOdd(true) {}
    bool solve(DiracOperator* dirac, const reduced_dirac_vector_t& source, reduced_dirac_vector_t& solution,
               reduced_dirac_vector_t const* initial_guess = 0) override {
        return run(LYM_SOLVER_PRECOND_BICGSTAB, useEvenOdd, dirac, source, solution, initial_guess);
    }
    void setUseEvenOddPreconditioning(bool b) { useEvenOdd = b; }

private:
    bool useEvenOdd;
};

// ConjugateGradient is a standalone class in the reference (ConjugateGradient.h:8-36)
class ConjugateGradient : private Solver {
public:
    ConjugateGradient() : Solver("ConjugateGradient") {
        precision = 0.00000000001;
        maxSteps = 3000;
    }
    bool solve(DiracOperator* dirac, const reduced_dirac_vector_t& source, reduced_dirac_vector_t& solution,
               reduced_dirac_vector_t const* initial_guess = 0) {
        return run(LYM_SOLVER_CG, false, dirac, source, solution, initial_guess);
    }
    void setPrecision(double e) { precision = e; }
    double getPrecision() const { return precision; }
    double getLastError() const { return lastError; }
    unsigned int getLastSteps() const { return lastSteps; }
    void setMaximumSteps(unsigned int m) { maxSteps = m; }
    unsigned int getMaximumSteps() const { return maxSteps; }
};

}  // namespace Update
