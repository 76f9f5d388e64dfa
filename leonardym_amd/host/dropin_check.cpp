// dropin_check.cpp -- drives the C++ drop-in the way LeonardYM's callers do
// (source/correlators/MesonCorrelator.cpp:25-99: getInstance("Improved", 1, ...),
// setLattice, setGamma5(false), PreconditionedBiCGStab over the 12 spin-colour
// point sources; source/tests/TestLinearAlgebra.cpp:36-150 for multiply and
// multiplyAdd), and writes every result to a file for comparison.
//
// usage: dropin_check Lx Ly Lz Lt kappa csw links.bin out.bin
//   links.bin: V x 4 x 9 complex<double> (column-major, BC folded)
//   out.bin  : 12 solutions, then Q psi, D psi, Q psi + a psi2, Q^2 psi, CG solution,
//              getSquareRoot(Q^2) psi
//              (each V x 12 complex<double>), psi and psi2 seeded in the program
#include <cstdio>
#include <cstdlib>
#include <map>
#include <vector>

#include "lym_update.h"

using namespace Update;

static void write(FILE* f, const reduced_dirac_vector_t& v) {
    fwrite(v.getRawData(), sizeof(complex), (size_t)LocalLayout::completesize * 12, f);
}

int main(int argc, char** argv) {
    if (argc != 9) {
        fprintf(stderr, "usage: %s Lx Ly Lz Lt kappa csw links.bin out.bin\n", argv[0]);
        return 2;
    }
    LocalLayout::glob_x = atoi(argv[1]);
    LocalLayout::glob_y = atoi(argv[2]);
    LocalLayout::glob_z = atoi(argv[3]);
    LocalLayout::glob_t = atoi(argv[4]);
    LocalLayout::initialize();
    const int V = LocalLayout::completesize;

    std::map<std::string, Option> desc;
    DiracOperator::registerParameters(desc, "");
    desc["kappa"] = argv[5];
    desc["csw"] = argv[6];
    StorageParameters params(desc);

    extended_fermion_lattice_t lattice;
    FILE* f = fopen(argv[7], "rb");
    if (!f || fread(lattice.getRawData(), sizeof(complex), (size_t)V * 36, f) != (size_t)V * 36) {
        fprintf(stderr, "cannot read %s\n", argv[7]);
        return 1;
    }
    fclose(f);

    try {
        FILE* out = fopen(argv[8], "wb");
        // MesonCorrelator-style propagator: D^-1 on 12 point sources at the origin
        DiracOperator* dirac = DiracOperator::getInstance("Improved", 1, params);
        dirac->setLattice(lattice);
        dirac->setGamma5(false);
        PreconditionedBiCGStab* solver = new PreconditionedBiCGStab();
        solver->setPrecision(1e-24);
        solver->setMaximumSteps(5000);
        reduced_dirac_vector_t source, solution;
        for (int alpha = 0; alpha < 4; ++alpha)
            for (int c = 0; c < diracVectorLength; ++c) {
                for (int site = 0; site < V; ++site)
                    for (int mu = 0; mu < 4; ++mu) source[site][mu].fill(complex(0., 0.));
                source[LocalLayout::getGlobalCoordinate(0, 0, 0, 0)][alpha][c] = 1;
                if (!solver->solve(dirac, source, solution)) {
                    fprintf(stderr, "solve (%d,%d) did not converge\n", alpha, c);
                    return 1;
                }
                fprintf(stderr, "point source (%d,%d): %u steps\n", alpha, c, solver->getLastSteps());
                write(out, solution);
            }
        // operator applies on a seeded Gaussian vector
        reduced_dirac_vector_t psi, psi2, r;
        AlgebraUtils::generateRandomComplexGaussianVector(psi, 11);
        AlgebraUtils::generateRandomComplexGaussianVector(psi2, 12);
        dirac->setGamma5(true);
        dirac->multiply(r, psi);
        write(out, r);
        dirac->setGamma5(false);
        dirac->multiply(r, psi);
        write(out, r);
        dirac->setGamma5(true);
        dirac->multiplyAdd(r, psi, psi2, complex(0.25, -0.5));
        write(out, r);
        DiracOperator* sq = DiracOperator::getSquare(dirac);
        sq->multiply(r, psi);
        write(out, r);
        // CG on Q^2 (ConjugateGradient.cpp:10-74), source psi
        ConjugateGradient cg;
        cg.setPrecision(1e-22 * AlgebraUtils::squaredNorm(psi));
        cg.setMaximumSteps(5000);
        if (!cg.solve(sq, psi, r)) {
            fprintf(stderr, "CG did not converge\n");
            return 1;
        }
        fprintf(stderr, "CG: %u steps, last error %g\n", cg.getLastSteps(), cg.getLastError());
        write(out, r);
        // getSquareRoot of the squared operator: back to Q (DiracOperator.cpp:184-229)
        DiracOperator* root = DiracOperator::getSquareRoot(sq);
        root->multiply(r, psi);
        write(out, r);
        delete root;
        fclose(out);
        delete sq;
        delete solver;
        delete dirac;
        // an unknown operator is an exception, not exit(1)
        try {
            DiracOperator::getInstance("Overlap", 1, params);
            fprintf(stderr, "unknown operator accepted\n");
            return 1;
        } catch (const std::exception& e) {
            fprintf(stderr, "expected error: %s\n", e.what());
        }
    } catch (const std::exception& e) {
        fprintf(stderr, "error: %s\n", e.what());
        return 1;
    }
    LocalLayout::finalize();
    printf("dropin_check ok\n");
    return 0;
}
