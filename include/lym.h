/* lym.h -- C ABI of the MI355X (gfx950) LeonardYM fermion-solve library.
 *
 * Plain pointers, sizes and opaque handles; no C++ or torch types.  Every
 * function returns LYM_OK (0) or an error code and never exits the process;
 * lym_last_error() gives the message of the calling thread's last failure.
 * (The reference exits on configuration errors:
 *  source/dirac_operators/DiracOperator.cpp:84-87, EvenOddImprovedDiracWilsonOperator.cpp:49-52.)
 *
 * Host data layouts are the reference's raw vectors:
 *   spinor : V x [4 spin][3 colour] complex double (re, im interleaved)
 *            (reduced_dirac_vector_t, source/LatticeDefinitions.h:86-114)
 *   links  : V x [4 mu][3x3 column-major] complex double, or for the SU(2)
 *            adjoint representation V x [4][3x3 column-major] real double
 *            (FermionicGroup, source/MatrixTypedef.h:61-90)
 * with sites in t-fastest lexicographic order
 *   idx = ((x*Ly + y)*Lz + z)*Lt + t   (source/MPILattice/LocalLayout.h:149-155).
 * Any extents >= 2 (source/MPILattice/LocalLayout.h:34-110).  Even-odd
 * operators, the even-odd driver, low-precision applies, mixed-precision,
 * batched and served solves need all four extents even
 * (source/inverters/PreconditionedBiCGStab.cpp:43); on a lattice with an odd
 * extent the full-lattice FP64 operators, multiplyAdd, the FP64 CG and
 * BiCGStab solvers (the preconditioned driver falls back to plain BiCGStab),
 * the meson correlators and the fermion force run, and the rest returns
 * LYM_ERR_UNSUPPORTED.
 */
#ifndef LYM_H
#define LYM_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define LYM_OK 0
#define LYM_ERR_INVALID 1     /* bad argument or configuration */
#define LYM_ERR_DEVICE 2      /* HIP runtime failure or no device */
#define LYM_ERR_NOMEM 3       /* device allocation failed */
#define LYM_ERR_UNSUPPORTED 4 /* valid request the library does not implement */
#define LYM_ERR_SINGULAR 5    /* a clover block could not be inverted */

/* gauge-group representations of the fermion links */
#define LYM_REP_SU3_FUND 0 /* NUMCOLORS=3, fundamental */
#define LYM_REP_SU2_ADJ 1  /* NUMCOLORS=2, ADJOINT: real 3x3 links */

/* operator kinds (names of DiracOperator::getInstance, DiracOperator.cpp:20-165) */
#define LYM_OP_WILSON 0       /* "DiracWilson":      DiracWilsonOperator.cpp:16-391 */
#define LYM_OP_BASIC_WILSON 1 /* "BasicDiracWilson": BasicDiracWilsonOperator.cpp:12-178 (same operator) */
#define LYM_OP_CLOVER 2       /* "Improved":         ImprovedDiracWilsonOperator.cpp:22-487 */
#define LYM_OP_EO_CLOVER 3    /* EvenOddImprovedDiracWilsonOperator.cpp:26-47 */

/* solvers */
#define LYM_SOLVER_CG 0                 /* ConjugateGradient.cpp:10-74 */
#define LYM_SOLVER_BICGSTAB 1           /* BiConjugateGradient.cpp:321-507 */
#define LYM_SOLVER_PRECOND_BICGSTAB 2   /* PreconditionedBiCGStab.cpp:33-100 */
/* Forms beyond the reference (no counterpart in it; their parity is against
 * the FP64 oracle): */
#define LYM_SOLVER_MIXED_CG 3           /* CG on Q^2 with an FP32/FP16 sloppy operator and FP64 reliable updates */
#define LYM_SOLVER_PRECOND_CG 4         /* the even-odd driver of PreconditionedBiCGStab.cpp:43-91 with FP64 CG on Qhat^2 */
#define LYM_SOLVER_PRECOND_MIXED_CG 5   /* the even-odd driver with the mixed-precision CG on Qhat^2 */

/* storage precision of a sloppy operator */
#define LYM_PREC_DOUBLE 0
#define LYM_PREC_SINGLE 1  /* FP32 storage and arithmetic */
#define LYM_PREC_HALF 2    /* FP16 storage (per-site FP32 scale on spinors and clover), FP32 arithmetic */

/* link storage on the device */
#define LYM_LINKS_FULL 0          /* 18 reals (9 for adjoint) */
#define LYM_LINKS_RECONSTRUCT 1   /* 12 reals (6 for adjoint), third row rebuilt */

typedef struct lym_ctx lym_ctx;
typedef struct lym_gauge lym_gauge;
typedef struct lym_spinor lym_spinor;
typedef struct lym_op lym_op;
typedef struct lym_solver lym_solver;

typedef struct {
    int converged;      /* solve() return value of the reference */
    unsigned steps;     /* getLastSteps(): iteration index at convergence, or max steps */
    double last_error;  /* final |r|^2 of the Krylov recursion (getLastError()) */
    unsigned applies;   /* operator applications performed */
    double seconds;     /* wall time of the solve on the device stream */
    unsigned reliable_updates; /* FP64 true-residual recomputations (mixed-precision solvers) */
} lym_solve_info;

/* ------------------------------------------------------------------ errors / device */
const char* lym_last_error(void);
int lym_device_count(int* n);
/* Page-locked host memory (no counterpart in the reference): spinor and link
 * uploads/downloads from such buffers run as direct DMA at full link speed
 * instead of being staged through the driver's pinned pool. */
int lym_host_alloc(size_t bytes, void** out);
void lym_host_free(void* p);

/* ------------------------------------------------------------------ context
 * Lattice geometry + stream: replaces LocalLayout::initialize()
 * (source/MPILattice/LocalLayout.h:34-110) and the boundary-condition choice of
 * Environment::synchronize (source/Environment.cpp:58-78): bc_t = -1 is
 * "fermion_antiperiodic", +1 "fermion_periodic". */
int lym_ctx_create(const int dims[4], int rep, int bc_t, lym_ctx** out);
void lym_ctx_destroy(lym_ctx* ctx);
int lym_ctx_synchronize(lym_ctx* ctx);
/* hipStream_t the context launches on (for callers that order their own work) */
void* lym_ctx_stream(lym_ctx* ctx);
/* HIP events around work on the context stream; stop returns elapsed ms */
int lym_ctx_timer_start(lym_ctx* ctx);
int lym_ctx_timer_stop(lym_ctx* ctx, float* ms);

/* ------------------------------------------------------------------ gauge field
 * Device-resident fermion links: the copy DiracOperator::setLattice makes
 * (source/dirac_operators/DiracOperator.cpp:332-334), shared by any number of
 * operators.  links_have_bc_folded = 1 when the time-boundary sign is already
 * in the links, as Environment::synchronize leaves them. */
int lym_gauge_create(lym_ctx* ctx, const double* links, int links_have_bc_folded, int storage, lym_gauge** out);
void lym_gauge_destroy(lym_gauge* g);

/* ------------------------------------------------------------------ spinors
 * A full-lattice reduced_dirac_vector_t resident on the device. */
int lym_spinor_create(lym_ctx* ctx, lym_spinor** out);
void lym_spinor_destroy(lym_spinor* s);
int lym_spinor_upload(lym_spinor* s, const double* host);
int lym_spinor_download(const lym_spinor* s, double* host);
int lym_spinor_copy(lym_spinor* dst, const lym_spinor* src);

/* AlgebraUtils (source/algebra_utils/AlgebraUtils.h) */
int lym_spinor_set_zero(lym_spinor* s);                                                    /* setToZero :256-263 */
int lym_spinor_gamma5(lym_spinor* s);                                                      /* gamma5 :268-275 */
int lym_spinor_squared_norm(const lym_spinor* s, double* out);                             /* squaredNorm :17-27 */
int lym_spinor_dot(const lym_spinor* a, const lym_spinor* b, double* re, double* im);      /* dot :182-196 */
int lym_spinor_gamma5dot(const lym_spinor* a, const lym_spinor* b, double* re, double* im);/* gamma5dot :219-238 */
int lym_spinor_difference_norm(const lym_spinor* a, const lym_spinor* b, double* out);     /* differenceNorm :244-254 */
int lym_spinor_axpy(lym_spinor* y, double are, double aim, const lym_spinor* x);            /* y += a x */

/* ------------------------------------------------------------------ operators
 * DiracOperator (source/dirac_operators/DiracOperator.h:11-140).  power 2 gives
 * the Square* operators Q^2 (SquareDiracWilsonOperator.cpp:12-22,
 * SquareImprovedDiracWilsonOperator.cpp:13-23, SquareEvenOddImprovedDiracWilsonOperator.cpp:13-27).
 * gamma5 = 1: multiply computes Q = gamma5 D; 0: D (DiracWilsonOperator.cpp:194-199). */
int lym_op_create(lym_ctx* ctx, int kind, unsigned power, double kappa, double csw, int gamma5, lym_op** out);
/* by factory name "DiracWilson" | "BasicDiracWilson" | "Improved" | "EvenOddImproved"
 * (DiracOperator::getInstance, DiracOperator.cpp:20-165) */
int lym_op_create_by_name(lym_ctx* ctx, const char* name, unsigned power, double kappa, double csw, int gamma5,
                          lym_op** out);
void lym_op_destroy(lym_op* op);
/* setLattice (:125).  The gauge field may belong to another context of the
 * same geometry on the same device (one device-resident copy serving several
 * streams); its clover field-strength blocks are built on first use, which must
 * then not race between threads (prepare one operator on it first). */
int lym_op_set_gauge(lym_op* op, const lym_gauge* g);
int lym_op_set_kappa(lym_op* op, double kappa);         /* setKappa (:107); diagonal blocks rebuilt */
int lym_op_set_csw(lym_op* op, double csw);             /* ImprovedDiracWilsonOperator::setCSW */
int lym_op_set_gamma5(lym_op* op, int gamma5);          /* setGamma5 (:127) */
int lym_op_get_params(const lym_op* op, double* kappa, double* csw, int* gamma5, int* kind, unsigned* power);
/* multiply (DiracOperator.h:90) */
int lym_op_multiply(lym_op* op, lym_spinor* out, const lym_spinor* in);
/* The same operator applied in a low storage precision (no counterpart in the
 * reference; the sloppy operator of the mixed-precision solvers): the input is
 * rounded to prec (LYM_PREC_SINGLE, or LYM_PREC_HALF with a per-site scale),
 * links and clover blocks are the operator's low-precision copies, arithmetic
 * is FP32, and the result is returned in FP64.  For the even-odd operator the
 * even half is copied through as in lym_op_multiply. */
int lym_op_multiply_sloppy(lym_op* op, int prec, lym_spinor* out, const lym_spinor* in);
/* lym_op_multiply_sloppy for nrhs spinors at once in one multi-RHS field
 * (no counterpart in the reference; the sloppy operator of the batched
 * solver): nrhs one of 1, 2, 3, 4, 6, 8, 12, 16.  matrix_cores = 1 runs the
 * FP16 SU(3) stencil's link products on the MFMA units (nrhs 2, 4 or 8;
 * LYM_ERR_UNSUPPORTED otherwise). */
int lym_op_multiply_sloppy_batch(lym_op* op, int prec, int matrix_cores, unsigned nrhs, lym_spinor* const* outs,
                                 const lym_spinor* const* ins);
/* multiplyAdd (DiracOperator.h:99): out = Op v1 + alpha v2 */
int lym_op_multiply_add(lym_op* op, lym_spinor* out, const lym_spinor* v1, const lym_spinor* v2, double are,
                        double aim);

/* ------------------------------------------------------------------ solvers
 * Solver (source/inverters/Solver.h:8-67), ConjugateGradient
 * (ConjugateGradient.h:8-36), BiConjugateGradient, PreconditionedBiCGStab
 * (PreconditionedBiCGStab.h:9-31).  Stopping rule |r|^2 < precision
 * (absolute, as the reference).  Defaults: CG 1e-11 / 3000 steps
 * (ConjugateGradient.cpp:6), BiCGStab 1e-10 / 1000 (Solver.h:10). */
int lym_solver_create(lym_ctx* ctx, int kind, lym_solver** out);
void lym_solver_destroy(lym_solver* s);
int lym_solver_set_precision(lym_solver* s, double precision);
int lym_solver_set_max_steps(lym_solver* s, unsigned max_steps);
int lym_solver_set_even_odd(lym_solver* s, int use_even_odd);   /* setUseEvenOddPreconditioning */
/* mixed-precision solvers (no counterpart in the reference): sloppy storage
 * precision (LYM_PREC_SINGLE or LYM_PREC_HALF, default HALF) and the factor by
 * which the sloppy residual must fall before an FP64 reliable update (default 0.05) */
int lym_solver_set_sloppy_precision(lym_solver* s, int prec);
int lym_solver_set_reliable_delta(lym_solver* s, double delta);
/* BiConjugateGradient's shifted overload (BiConjugateGradient.cpp:510-690):
 * plain BiCGStab (LYM_SOLVER_BICGSTAB) on Op + shift, every apply being
 * multiplyAdd(out, v, v, shift).  Full-lattice operators only (the even-odd
 * operator has no multiplyAdd); 0 + 0i restores the plain solver. */
int lym_solver_set_shift(lym_solver* s, double shift_re, double shift_im);
/* solve(dirac, source, solution, initial_guess) -- x0 may be NULL */
int lym_solver_solve(lym_solver* s, lym_op* op, const lym_spinor* source, lym_spinor* solution,
                     const lym_spinor* initial_guess, lym_solve_info* info);
/* Many right-hand sides against one operator (no counterpart in the reference,
 * which solves them one call at a time, e.g. MesonCorrelator.cpp loops over 12
 * point sources).  Mixed-precision CG kinds only (LYM_SOLVER_MIXED_CG,
 * LYM_SOLVER_PRECOND_MIXED_CG).  RHS are solved in batches that share every
 * operator apply; each keeps its own recursion, tolerance (precisions[q], or
 * the solver's precision when precisions is NULL) and stopping point, with
 * results bit-identical to the same RHS solved in a batch of any other size.
 * Starting guess x0 = source.  info[q].seconds is the wall time of the call. */
int lym_solver_solve_batch(lym_solver* s, lym_op* op, unsigned nrhs, const lym_spinor* const* sources,
                           lym_spinor* const* solutions, const double* precisions, lym_solve_info* infos);
/* Batched solves whose right-hand sides have their own kappa (no counterpart
 * in the reference; the several-kappa propagators of MesonCorrelator.cpp:25-99
 * solved together): solves D(kappas[q]) x_q = b_q with the even-odd
 * mixed-precision CG driver of a clover operator `op` (power 1; its gauge
 * field and csw, kappa per RHS) to |r|^2 < precisions[q] on the even-odd
 * system (NULL: the solver's precision), as lym_solver_solve_batch does.  The
 * RHS are lanes of one batch (width: the RHS block; 0 = 12 for 12 or more
 * RHS, else 8), each with its own kappa, clover blocks and stopping point, in
 * one stencil launch per apply; an RHS that stops frees its lane for the next
 * one, and RHS enter in order of kappa so that lanes of one kappa share their
 * clover blocks.  A result does not depend on the other RHS, the order of
 * admission or the batch width; a kappa that is not finite is an
 * LYM_ERR_INVALID.  Solver kind LYM_SOLVER_PRECOND_MIXED_CG. */
int lym_solver_solve_batch_kappas(lym_solver* s, lym_op* op, unsigned nrhs, const double* kappas,
                                  const lym_spinor* const* sources, lym_spinor* const* solutions,
                                  const double* precisions, lym_solve_info* infos);
/* RHS per batch: 0 = autotune (timed per lattice shape, link storage and
 * precision on first use), else one of 1, 2, 3, 4, 6, 8, 12, 16 */
int lym_solver_set_rhs_block(lym_solver* s, unsigned nrhs_per_batch);
/* the batch size and workgroup size in use (threads 0: default) */
int lym_solver_get_rhs_block(lym_solver* s, unsigned* nrhs_per_batch, unsigned* threads);
/* The matrix-core form of the batched FP16 stencil (no counterpart in the
 * reference): with enable = 1, FP16 batches of SU(3) links may run the link
 * product of every hop on the MFMA units for batch widths 2, 4 and 8 (the
 * autotuner times it beside the vector form and keeps the faster; with a fixed
 * RHS block those widths use it).  Default 0.  Results differ from the vector
 * form by FP32 rounding, so an RHS gives the same bits in batches of those
 * widths, not across forms.  used: the form the last batched call chose. */
int lym_solver_set_matrix_cores(lym_solver* s, int enable);
int lym_solver_get_matrix_cores(const lym_solver* s, int* enabled, int* used);

/* ------------------------------------------------------------------ serving
 * A request loop over device-resident gauge fields (no counterpart in the
 * reference, whose callers -- MesonCorrelator.cpp:25-99,
 * ChiralCondensate.cpp:43-120 -- solve one source at a time in their own
 * loop).  Each request solves D x = b for the clover operator of its lattice
 * and kappa with the even-odd mixed-precision CG driver to
 * |b - D x| <= tolerance |b|.  Continuous batching: a worker thread (one HIP
 * stream each) runs a batched solve on one lattice whose lanes are requests
 * of any kappa and tolerance; between chunks of iterations it returns the
 * requests that have stopped and admits queued requests of that lattice into
 * free lanes (up to max_batch).  Device state is charged against an HBM
 * budget (0: 80% of the free HBM at creation); a worker releases its least
 * recently used lattice state to admit a new one.  The source must stay valid and the solution
 * writable until the request completes.  A result does not depend on the
 * other requests in flight. */
typedef struct lym_server lym_server;
#define LYM_REQ_PENDING 0
#define LYM_REQ_RUNNING 1
#define LYM_REQ_DONE 2
#define LYM_REQ_FAILED 3
typedef struct {
    int status;              /* LYM_REQ_* */
    int error;               /* LYM_ERR_* when failed */
    unsigned batch_size;     /* requests solved together with this one */
    int worker;
    double queue_seconds;    /* arrival -> start of service */
    double service_seconds;  /* start of service -> result in host memory */
    double latency_seconds;  /* arrival -> result in host memory */
    lym_solve_info solve;
} lym_request_info;
int lym_server_create(unsigned workers, size_t hbm_budget_bytes, unsigned max_batch, lym_server** out);
void lym_server_destroy(lym_server* s);   /* finishes queued requests first */
/* sloppy precision LYM_PREC_SINGLE | LYM_PREC_HALF (default HALF), max steps (0: keep 20000) */
int lym_server_set_solver(lym_server* s, int sloppy_prec, unsigned max_steps);
int lym_server_add_lattice(lym_server* s, const int dims[4], int rep, int bc_t, const double* links,
                           int links_have_bc_folded, int storage, double csw, int* lattice_id);
int lym_server_submit(lym_server* s, int lattice_id, double kappa, double tolerance, const double* source,
                      double* solution, uint64_t* request_id);
/* timeout < 0 waits for completion; otherwise info->status tells whether it completed */
int lym_server_wait(lym_server* s, uint64_t request_id, double timeout_seconds, lym_request_info* info);
int lym_server_release(lym_server* s, uint64_t request_id);   /* forget a completed request */
int lym_server_drain(lym_server* s);                           /* wait until nothing is queued or running */
int lym_server_memory(lym_server* s, size_t* budget_bytes, size_t* in_use_bytes);

/* ------------------------------------------------------------------ measurements
 * Connected meson correlators from point-source propagators:
 * MesonCorrelator::execute (source/correlators/MesonCorrelator.cpp:22-200).
 * The 12 spin-colour unit sources at source_site (StochasticEstimator::
 * generateSource, source/fermion_measurements/StochasticEstimator.cpp:149-161,
 * there at the origin) are solved for D x = e with the given solver and
 * operator (gamma5 is switched off for the solves, as MesonCorrelator.cpp:71
 * does, and restored); a mixed-precision CG solver solves all 12 in one
 * batched call.  out holds 4 x Lt doubles, each indexed by the sink time
 * relative to the source timeslice (the reference's source at t = 0 of links
 * shifted by t_source_origin, :54-66): pion_exact, scalar_exact, ps_exact
 * (normalised by 4 kappa^2, :162) and pcac_mass (:189-190).  infos (NULL or
 * 12) receives the solves, column c * 4 + alpha for source spin alpha, colour c.
 * The stochastic disconnected part (:201-) is not computed. */
int lym_meson_correlators(lym_solver* solver, lym_op* op, const int source_site[4], double* out,
                          lym_solve_info* infos);

/* ------------------------------------------------------------------ gauge-configuration files
 * Readers and writers of the reference's configuration formats on host arrays
 * of fundamental gauge links (gaugeLinkConfiguration raw data): V x [4 mu]
 * [N x N column-major] complex double, N = ncolors (2 or 3), sites in the
 * t-fastest lexicographic order above.  The fermion links of the operators
 * are made from them as Environment::synchronize does (source/Environment.cpp:
 * 21-79): lym_su2_to_adjoint for adjoint fermions, then the boundary sign.
 *   LYM_FORMAT_LEONARD : "leonard_format" (ReadStartGaugeConfiguration.cpp:51-130,
 *     OutputSweep.cpp:32-84): <dir><name>_<number>.descriptor.txt and one XDR
 *     file per process <dir><name>_<number>_<proc>.txt of (x, y, z, t, links
 *     as doubles); written as one process, read for any number of them
 *   LYM_FORMAT_MUENSTER : "muenster_format", XDR floats with a trailer;
 *     SU(2) <dir>0:<number>:<name> (ReadStartGaugeConfiguration.cpp:132-297,
 *     OutputSweep.cpp:87-283), SU(3) <dir><name>.config-<number>
 *     (ReadStartGaugeConfiguration.cpp:299-480, OutputSweep.cpp:284-440;
 *     read links are re-unitarised from their first two columns)
 * Errors (unreadable, short or malformed files, a lattice size different from
 * dims) return LYM_ERR_INVALID; the reference prints and returns false. */
#define LYM_FORMAT_LEONARD 0
#define LYM_FORMAT_MUENSTER 1
typedef struct {
    double plaquette;  /* temporal plaquette: of the links when writing, the file's value when reading */
    double beta;       /* muenster trailer (written as a float, as the reference does) */
    double kappa;      /* muenster SU(2) trailer (the last, positive kappa) */
    double c1;         /* muenster SU(3) trailer (-1/12 when written) */
    int sweep;         /* muenster SU(2) trailer: sweep number */
} lym_config_info;
/* info: beta, kappa, sweep are inputs; plaquette (and c1) are set on return */
int lym_config_write(int format, const char* directory, const char* name, int number, const int dims[4],
                     int ncolors, const double* links, lym_config_info* info);
/* info may be NULL */
int lym_config_read(int format, const char* directory, const char* name, int number, const int dims[4],
                    int ncolors, double* links, lym_config_info* info);
/* Plaquette::temporalPlaquette (source/wilson_loops/Plaquette.cpp:49-66):
 * sum over sites and mu < 3 of Re tr[U_mu(x) U_t(x+mu) U_mu(x+t)^+ U_t(x)^+] / (3 N V) */
int lym_temporal_plaquette(const int dims[4], int ncolors, const double* links, double* out);

/* ------------------------------------------------------------------ HMC fermion force
 * TwoFlavorFermionAction::force of every link (source/actions/TwoFlavorFermionAction.cpp:17-20):
 * minus FermionForce::force (source/hmc_forces/FermionForce.cpp:33-58) of
 * DiracWilsonFermionForce::derivative (source/hmc_forces/DiracWilsonFermionForce.cpp:10-54)
 * from the solutions X = Q^-2 phi and Y = Q^-1 phi, for a Wilson operator
 * ("DiracWilson" / "BasicDiracWilson"; its kappa and gauge field).  force:
 * host V x [4 mu][N x N column-major] complex double, N = 3 (SU(3)) or 2
 * (SU(2) with adjoint fermions), traceless anti-hermitian.  As in the
 * reference the derivative is taken with respect to the fermion link and
 * projected with the link without the boundary sign, so on the time links
 * that carry the antiperiodic sign it is the negative of the derivative with
 * respect to the gauge link. */
int lym_fermion_force(lym_op* op, const lym_spinor* X, const lym_spinor* Y, double* force);
/* lym_fermion_force without any device allocation: no counterpart in the
 * reference.  It is for callers that cannot allocate device memory (tight
 * HMC or serving loops, harnesses that forbid hipMalloc; it synchronises the
 * context stream, so it is not for use under graph capture).  Same arguments
 * and the same bits as lym_fermion_force; the
 * force passes through a fixed 16 MiB device array of the library (one per
 * GPU, held until the process ends) in chunks of whole sites, so any lattice
 * size is accepted.  Calls that overlap on one GPU share that array and race:
 * the caller serialises them. */
int lym_fermion_force_noalloc(lym_op* op, const lym_spinor* X, const lym_spinor* Y, double* force);
/* TwoFlavorFermionAction::updateForce (:21-43): Y = Q^-1 phi and X = Q^-1 Y
 * with the solver (LYM_SOLVER_BICGSTAB or LYM_SOLVER_PRECOND_BICGSTAB, its
 * precision and step limit; the reference uses a BiConjugateGradient), then
 * lym_fermion_force.  infos: NULL or 2 entries (the two solves). */
int lym_two_flavor_force(lym_solver* s, lym_op* op, const lym_spinor* phi, double* force, lym_solve_info* infos);

/* Stout smearing of host fundamental gauge links (layout as lym_config_*) on
 * the device, `levels` times with parameter rho: StoutSmearing::smearing
 * (source/utils/StoutSmearing.cpp:12-20, smearLink :55-62) as
 * Environment::synchronize applies it before building the fermion links
 * (source/Environment.cpp:22-47).  in and out may be the same array.  An
 * exponent that is exactly zero gives the identity (the reference divides
 * 0 by 0 there). */
int lym_stout_smear(const int dims[4], int ncolors, double rho, int levels, const double* in, double* out);

/* ------------------------------------------------------------------ diagnostics
 * The workgroup -> site-block mapping of one stencil launch (no counterpart in
 * the reference), evaluated on the host by the device's own mapping functions
 * (no device is touched).  parity 0 or 1: a half-lattice launch of the
 * even-odd operator; -1: a full-lattice apply.  nr right-hand sides per site,
 * threads per workgroup (0: the default), lowp = 1: a single-RHS FP32/FP16
 * apply (two threads per site where the lattice allows).  Returns the number
 * of workgroups, sites per workgroup, site blocks per parity and site blocks
 * per x-slice (0 when an x-slice is not a whole number of blocks).  visits
 * (2 * nbpp ints, or NULL) counts how often each (parity, block) is visited;
 * branches (3 ints, or NULL) counts the blocks mapped by each renumbering:
 * plain order, x-slice fastest, XCD pairs. */
int lym_launch_plan(const int dims[4], int parity, int nr, int threads, int lowp, int* grid, int* spb, int* nbpp,
                    int* xslice_blocks, int* visits, long visits_len, int* branches);

/* ------------------------------------------------------------------ host-side input preparation
 * Seeded synthetic inputs (the reference seeds from time(NULL),
 * source/utils/RandomSeed.cpp:17-28, so inputs are generated here instead). */
/* U = exp(i eps H), H = sum_a w_a lambda_a / 2, w_a ~ N(0,1); SU(3), column-major */
int lym_gen_su3_links(const int dims[4], double eps, uint64_t seed, double* links);
/* U = exp(i eps sigma.a), a ~ N(0,1)^3; SU(2) 2x2 column-major complex, V x 4 x 4 complex */
int lym_gen_su2_links(const int dims[4], double eps, uint64_t seed, double* links);
/* ConvertLattice::toAdjoint (source/utils/ConvertLattice.h:24-34): R_ab = 2 Re tr(T_a U T_b U^+), T = sigma/2 */
int lym_su2_to_adjoint(long nlinks, const double* su2, double* adj);
/* switchAntiperiodicBc (source/LatticeBoundaryConditions.h:8-14) on host links */
int lym_apply_bc(const int dims[4], int rep, int bc_t, double* links);
/* complex Gaussian, sigma = 1/sqrt(2) per real part (AlgebraUtils.h:92-127) */
int lym_gen_gaussian_spinor(long nsites, uint64_t seed, double* out);
/* unit point source at `site` (StochasticEstimator::generateSource, StochasticEstimator.cpp:149-161) */
int lym_point_source(const int dims[4], const int site[4], int spin, int colour, double* out);

#ifdef __cplusplus
}
#endif

#endif /* LYM_H */
