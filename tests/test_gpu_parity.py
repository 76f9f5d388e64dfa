"""GPU parity: every path through the C ABI against the CPU oracle on the same
seeded inputs.  Operator applies must agree per site to 1e-13 relative;
solves must reach true residual |b - D x|/|b| <= 1e-10 and agree with the
oracle's solution to 1e-9 (the tolerances the target states for the FP64
path)."""
import numpy as np
import pytest

from helpers import (GAMMA5, K_BASIC, K_CLOVER, K_EO, K_WILSON, OracleField, coords, fold_bc, rel_err_per_site)

import leonardym_amd as lym

pytestmark = pytest.mark.gpu

SITE_TOL = 1e-13


def g5(v):
    return np.einsum("ab,sbc->sac", GAMMA5, v)


class Case:
    """A lattice with the same seeded links on the GPU and in the oracle."""

    def __init__(self, dims, rep="su3", eps=0.5, kappa=0.12, csw=1.0, storage=lym.LINKS_FULL, seed=11, bc=-1):
        self.dims = dims
        self.V = int(np.prod(dims))
        if rep == "su3":
            U = lym.gen_su3_links(dims, eps, seed)
            self.ctx = lym.Context(dims, lym.REP_SU3_FUND, bc)
            Uf = fold_bc(dims, U, bc)
            self.gauge = lym.Gauge(self.ctx, Uf, folded=True, storage=storage)
            Uo = Uf
        else:
            R = lym.su2_to_adjoint(lym.gen_su2_links(dims, eps, seed))
            self.ctx = lym.Context(dims, lym.REP_SU2_ADJ, bc)
            Rf = fold_bc(dims, R, bc)
            self.gauge = lym.Gauge(self.ctx, Rf, folded=True, storage=storage)
            Uo = Rf.astype(complex)
        self.kappa, self.csw = kappa, csw
        self.orc = OracleField(dims, Uo, kappa=kappa, csw=csw, need_clover=True)
        self.rng = np.random.default_rng(seed + 100)

    def spinor(self, host=None):
        return lym.Spinor(self.ctx, host)

    def gauss(self):
        return lym.gaussian_spinor(self.V, int(self.rng.integers(1 << 40)))

    def op(self, kind, power=1, gamma5=True, kappa=None, csw=None):
        return lym.Operator(self.ctx, kind, power, self.kappa if kappa is None else kappa,
                            self.csw if csw is None else csw, gamma5, gauge=self.gauge)

    def apply(self, op, v):
        a, b = self.spinor(v), self.spinor()
        op.multiply(b, a)
        return b.download()


CASES = [
    dict(dims=(4, 4, 4, 4)),
    dict(dims=(8, 8, 8, 8)),
    dict(dims=(4, 6, 2, 8)),
    dict(dims=(4, 4, 4, 8), storage=lym.LINKS_RECONSTRUCT),
    dict(dims=(4, 4, 4, 4), rep="adj", kappa=0.17),
    dict(dims=(4, 4, 6, 4), rep="adj", kappa=0.17, storage=lym.LINKS_RECONSTRUCT),
    dict(dims=(4, 4, 4, 6), bc=1),
]


@pytest.fixture(scope="module", params=range(len(CASES)), ids=lambda i: "case%d" % i)
def case(request):
    if lym.device_count() < 1:
        pytest.fail("no HIP device visible")
    print("case", CASES[request.param], flush=True)
    return Case(**CASES[request.param])


def test_spinor_roundtrip_and_blas(case):
    a, b = case.gauss(), case.gauss()
    A, B = case.spinor(a), case.spinor(b)
    assert np.array_equal(A.download(), a)
    assert abs(A.squared_norm() - np.vdot(a, a).real) < 1e-12 * np.vdot(a, a).real
    d = A.dot(B)
    ref = np.vdot(a.ravel(), b.ravel())
    assert abs(d - ref) < 1e-12 * abs(ref) + 1e-12
    d5 = A.gamma5dot(B)
    ref5 = np.vdot(a.ravel(), g5(b).ravel())
    assert abs(d5 - ref5) < 1e-12 * np.linalg.norm(a) * np.linalg.norm(b)
    assert abs(A.difference_norm(B) - np.sum(np.abs(a - b) ** 2)) < 1e-12 * np.sum(np.abs(a - b) ** 2)
    A.axpy(0.5 - 2j, B)
    assert np.max(np.abs(A.download() - (a + (0.5 - 2j) * b))) < 1e-14
    A.gamma5()
    assert np.max(np.abs(A.download() - g5(a + (0.5 - 2j) * b))) < 1e-14
    A.set_zero()
    assert np.count_nonzero(A.download()) == 0


@pytest.mark.parametrize("kind", [lym.OP_WILSON, lym.OP_BASIC_WILSON, lym.OP_CLOVER])
@pytest.mark.parametrize("gamma5", [True, False])
def test_multiply_matches_oracle(case, kind, gamma5):
    psi = case.gauss()
    okind = {lym.OP_WILSON: K_WILSON, lym.OP_BASIC_WILSON: K_BASIC, lym.OP_CLOVER: K_CLOVER}[kind]
    got = case.apply(case.op(kind, gamma5=gamma5), psi)
    ref = case.orc.apply(okind, psi, gamma5=gamma5)
    assert rel_err_per_site(got, ref) < SITE_TOL


@pytest.mark.parametrize("kind", [lym.OP_WILSON, lym.OP_CLOVER])
@pytest.mark.parametrize("gamma5", [True, False])
def test_multiply_add_matches_oracle(case, kind, gamma5):
    v1, v2 = case.gauss(), case.gauss()
    op = case.op(kind, gamma5=gamma5)
    A, B, C = case.spinor(v1), case.spinor(v2), case.spinor()
    op.multiply_add(C, A, B, 0.3 - 0.8j)
    ref = case.orc.apply(K_CLOVER if kind == lym.OP_CLOVER else K_WILSON, v1, gamma5=gamma5, v2=v2, alpha=0.3 - 0.8j)
    assert rel_err_per_site(C.download(), ref) < SITE_TOL


@pytest.mark.parametrize("kind", [lym.OP_WILSON, lym.OP_CLOVER, lym.OP_EO_CLOVER])
def test_square_operators(case, kind):
    psi, v2 = case.gauss(), case.gauss()
    okind = {lym.OP_WILSON: K_WILSON, lym.OP_CLOVER: K_CLOVER, lym.OP_EO_CLOVER: K_EO}[kind]
    op = case.op(kind, power=2)
    got = case.apply(op, psi)
    ref = case.orc.apply(okind, psi, power=2)
    assert rel_err_per_site(got, ref) < 10 * SITE_TOL
    A, B, C = case.spinor(psi), case.spinor(v2), case.spinor()
    op.multiply_add(C, A, B, 1.5j)
    assert rel_err_per_site(C.download(), ref + 1.5j * v2) < 10 * SITE_TOL


@pytest.mark.parametrize("gamma5", [True, False])
def test_even_odd_operator(case, gamma5):
    psi = case.gauss()
    got = case.apply(case.op(lym.OP_EO_CLOVER, gamma5=gamma5), psi)
    ref = case.orc.apply(K_EO, psi, gamma5=gamma5)
    assert rel_err_per_site(got, ref) < SITE_TOL
    op = case.op(lym.OP_EO_CLOVER)
    with pytest.raises(lym.LymError, match="not implemented"):
        op.multiply_add(case.spinor(), case.spinor(psi), case.spinor(psi), 1.0)


@pytest.mark.parametrize("csw", [1.0, 6.0])
def test_clover_inverse_paths_agree(case, csw, monkeypatch):
    """R_ee^-1 from the unpivoted LDL^H kernel and from the pivoted
    Gauss-Jordan fallback give the same EO operator as the oracle."""
    psi = case.gauss()
    fld = OracleField(case.dims, case.orc.U, kappa=case.kappa, csw=csw, need_clover=True)
    ref = fld.apply(K_EO, psi)
    got_ldl = case.apply(case.op(lym.OP_EO_CLOVER, csw=csw), psi)
    monkeypatch.setenv("LYM_CLOVER_INV_PIVOTED", "1")
    got_piv = case.apply(case.op(lym.OP_EO_CLOVER, csw=csw), psi)
    monkeypatch.delenv("LYM_CLOVER_INV_PIVOTED")
    tol = SITE_TOL if csw == 1.0 else 1e-11
    assert rel_err_per_site(got_ldl, ref) < tol
    assert rel_err_per_site(got_piv, ref) < tol


def test_kappa_change_rebuilds_clover(case):
    psi = case.gauss()
    op = case.op(lym.OP_EO_CLOVER)
    case.apply(op, psi)
    op.set_kappa(0.1)
    got = case.apply(op, psi)
    ref = case.orc.apply(K_EO, psi)  # oracle field built at case.kappa: rebuild at 0.1
    fld = OracleField(case.dims, case.orc.U, kappa=0.1, csw=case.csw, need_clover=True)
    ref = fld.apply(K_EO, psi)
    assert rel_err_per_site(got, ref) < SITE_TOL


def true_residual(case, x, b):
    """|b - D x| / |b| with the oracle's D."""
    Dx = case.orc.apply(K_CLOVER, x, gamma5=False)
    return np.linalg.norm((b - Dx).ravel()) / np.linalg.norm(b.ravel())


def test_cg_on_square_clover(case):
    """CG on Q^2 with rhs Q(gamma5 b) solves D x = b (ConjugateGradient.cpp:10-74)."""
    b = case.gauss()
    rhs = case.orc.apply(K_CLOVER, g5(b))
    eps = (1e-11 * np.linalg.norm(rhs)) ** 2
    solver = lym.Solver(case.ctx, lym.SOLVER_CG, precision=eps, max_steps=5000)
    X = case.spinor()
    info = solver.solve(case.op(lym.OP_CLOVER, power=2), case.spinor(rhs), X)
    assert info.converged
    x = X.download()
    assert true_residual(case, x, b) <= 1e-10
    xo, oinfo, ok = case.orc.solve("cg", rhs, eps, kind=K_CLOVER, power=2, max_steps=5000)
    assert ok
    assert np.linalg.norm((x - xo).ravel()) / np.linalg.norm(xo.ravel()) < 1e-9
    assert abs(int(info.steps) - int(oinfo.steps)) <= 2


@pytest.mark.parametrize("gamma5", [False, True])
def test_bicgstab_plain(case, gamma5):
    b = case.gauss()
    rhs = b if not gamma5 else g5(b)
    eps = (1e-11 * np.linalg.norm(rhs)) ** 2
    solver = lym.Solver(case.ctx, lym.SOLVER_BICGSTAB, precision=eps, max_steps=3000)
    X = case.spinor()
    info = solver.solve(case.op(lym.OP_CLOVER, gamma5=gamma5), case.spinor(rhs), X)
    assert info.converged
    x = X.download()
    assert true_residual(case, x, b) <= 1e-10
    xo, oinfo, ok = case.orc.solve("bicgstab", rhs, eps, kind=K_CLOVER, gamma5=gamma5, max_steps=3000)
    assert ok
    assert np.linalg.norm((x - xo).ravel()) / np.linalg.norm(xo.ravel()) < 1e-9


def test_preconditioned_bicgstab(case):
    """Even-odd driver returns D^-1 b whatever gamma5 says (PreconditionedBiCGStab.cpp:43-91)."""
    b = case.gauss()
    eps = (1e-11 * np.linalg.norm(b)) ** 2
    solver = lym.Solver(case.ctx, lym.SOLVER_PRECOND_BICGSTAB, precision=eps, max_steps=3000)
    X = case.spinor()
    info = solver.solve(case.op(lym.OP_CLOVER, gamma5=True), case.spinor(b), X)
    assert info.converged
    x = X.download()
    assert true_residual(case, x, b) <= 1e-10
    xo, oinfo, ok = case.orc.solve("precond_bicgstab", b, eps, kind=K_CLOVER, max_steps=3000)
    assert ok
    assert np.linalg.norm((x - xo).ravel()) / np.linalg.norm(xo.ravel()) < 1e-9


def test_cg_on_even_odd_square(case):
    """CG on Qhat^2 (odd half), even half of the solution = even half of the source."""
    b = case.gauss()
    eps = (1e-11 * np.linalg.norm(b)) ** 2
    solver = lym.Solver(case.ctx, lym.SOLVER_CG, precision=eps, max_steps=5000)
    X = case.spinor()
    info = solver.solve(case.op(lym.OP_EO_CLOVER, power=2), case.spinor(b), X)
    assert info.converged
    x = X.download()
    xo, oinfo, ok = case.orc.solve("cg", b, eps, kind=K_EO, power=2, max_steps=5000)
    assert ok
    assert np.linalg.norm((x - xo).ravel()) / np.linalg.norm(xo.ravel()) < 1e-9
    par = coords(case.dims).sum(1) % 2
    assert np.array_equal(x[par == 0], b[par == 0])


def test_solver_edge_cases(case):
    b = case.gauss()
    # non-convergence: max steps reached, reported as the reference does
    solver = lym.Solver(case.ctx, lym.SOLVER_CG, precision=1e-300, max_steps=3)
    info = solver.solve(case.op(lym.OP_CLOVER, power=2), case.spinor(b), case.spinor())
    assert not info.converged and info.steps == 3
    xo, oinfo, ok = case.orc.solve("cg", b, 1e-300, kind=K_CLOVER, power=2, max_steps=3)
    # zero source: BiCGStab starts from a seeded random vector and still converges
    z = np.zeros_like(b)
    s2 = lym.Solver(case.ctx, lym.SOLVER_BICGSTAB, precision=1e-20, max_steps=3000)
    X = case.spinor()
    info = s2.solve(case.op(lym.OP_CLOVER, gamma5=False), case.spinor(z), X)
    assert info.converged and np.linalg.norm(X.download()) < 1e-8
    # initial guess equal to the solution converges at once
    rhs = b
    eps = (1e-11 * np.linalg.norm(rhs)) ** 2
    s3 = lym.Solver(case.ctx, lym.SOLVER_BICGSTAB, precision=eps, max_steps=3000)
    X = case.spinor()
    s3.solve(case.op(lym.OP_CLOVER, gamma5=False), case.spinor(rhs), X)
    X2 = case.spinor()
    info2 = s3.solve(case.op(lym.OP_CLOVER, gamma5=False), case.spinor(rhs), X2, guess=X)
    x0 = X.download()
    xo, oinfo, ok = case.orc.solve("bicgstab", rhs, eps, kind=K_CLOVER, gamma5=False, max_steps=3000, x0=x0)
    assert info2.converged and ok and info2.steps <= 3 and abs(int(info2.steps) - int(oinfo.steps)) <= 1


def test_errors_are_codes_not_exits(case):
    with pytest.raises(lym.LymError, match="not supported"):
        lym.Operator(case.ctx, 0, name="Overlap")
    with pytest.raises(lym.LymError):
        lym.Operator(case.ctx, lym.OP_WILSON, power=3)
    op = lym.Operator(case.ctx, lym.OP_WILSON)  # no gauge yet
    with pytest.raises(lym.LymError, match="gauge"):
        op.multiply(case.spinor(), case.spinor())
    by_name = lym.Operator(case.ctx, 0, name="Improved", kappa=case.kappa, csw=case.csw, gauge=case.gauge)
    psi = case.gauss()
    assert rel_err_per_site(case.apply(by_name, psi), case.orc.apply(K_CLOVER, psi)) < SITE_TOL


def test_fresh_buffers_are_ordered_before_first_use():
    """Buffers are zeroed on the context stream, so objects used right after
    creation never see a late zero fill (regression: the fill once ran on the
    null stream, unordered with the non-blocking context stream)."""
    dims = (8, 8, 8, 8)
    V = int(np.prod(dims))
    ctx = lym.Context(dims)
    for r in range(6):
        U = fold_bc(dims, lym.gen_su3_links(dims, 0.5, 500 + r))
        g = lym.Gauge(ctx, U, folded=True)
        op = lym.Operator(ctx, lym.OP_EO_CLOVER, 1, 0.13, 1.0, True, gauge=g)
        psi = lym.gaussian_spinor(V, 600 + r)
        a, b = lym.Spinor(ctx, psi), lym.Spinor(ctx)
        op.multiply(b, a)
        orc = OracleField(dims, U, kappa=0.13, csw=1.0, need_clover=True)
        assert rel_err_per_site(b.download(), orc.apply(K_EO, psi)) < SITE_TOL


@pytest.mark.parametrize("sloppy", [lym.PREC_SINGLE, lym.PREC_HALF])
def test_mixed_cg_on_square_clover(case, sloppy):
    """Mixed-precision CG (FP32 arithmetic, FP32/FP16 storage, FP64 reliable
    updates) reaches the FP64 stopping rule and the FP64 oracle solution."""
    b = case.gauss()
    rhs = case.orc.apply(K_CLOVER, g5(b))
    eps = (1e-11 * np.linalg.norm(rhs)) ** 2
    solver = lym.Solver(case.ctx, lym.SOLVER_MIXED_CG, precision=eps, max_steps=20000, sloppy=sloppy)
    X = case.spinor()
    info = solver.solve(case.op(lym.OP_CLOVER, power=2), case.spinor(rhs), X)
    assert info.converged and info.reliable_updates >= 1
    x = X.download()
    assert true_residual(case, x, b) <= 1e-10
    xo, oinfo, ok = case.orc.solve("cg", rhs, eps, kind=K_CLOVER, power=2, max_steps=5000)
    err = np.linalg.norm((x - xo).ravel()) / np.linalg.norm(xo.ravel())
    print("mixed CG prec %d: %d iterations (FP64 oracle %d), %d reliable updates, error vs oracle %.2e"
          % (sloppy, info.steps, oinfo.steps, info.reliable_updates, err))
    assert err < 1e-9


@pytest.mark.parametrize("kind,sloppy", [(lym.SOLVER_PRECOND_CG, None), (lym.SOLVER_PRECOND_MIXED_CG, lym.PREC_HALF),
                                         (lym.SOLVER_PRECOND_MIXED_CG, lym.PREC_SINGLE)])
def test_even_odd_cg_drivers(case, kind, sloppy):
    """The even-odd driver with FP64 CG or mixed CG on Qhat^2 returns D^-1 b."""
    b = case.gauss()
    eps = (1e-11 * np.linalg.norm(b)) ** 2
    solver = lym.Solver(case.ctx, kind, precision=eps, max_steps=20000, sloppy=sloppy)
    X = case.spinor()
    info = solver.solve(case.op(lym.OP_CLOVER), case.spinor(b), X)
    assert info.converged
    x = X.download()
    assert true_residual(case, x, b) <= 1e-10
    xo, oinfo, ok = case.orc.solve("precond_bicgstab", b, eps, kind=K_CLOVER, max_steps=3000)
    assert np.linalg.norm((x - xo).ravel()) / np.linalg.norm(xo.ravel()) < 1e-9


def test_mixed_cg_max_steps(case):
    b = case.gauss()
    solver = lym.Solver(case.ctx, lym.SOLVER_MIXED_CG, precision=1e-300, max_steps=4)
    X = case.spinor()
    info = solver.solve(case.op(lym.OP_CLOVER, power=2), case.spinor(b), X)
    assert not info.converged and info.steps == 4
    assert np.all(np.isfinite(X.download()))


@pytest.mark.parametrize("kind,op_kind,power", [(lym.SOLVER_CG, lym.OP_CLOVER, 2),
                                                (lym.SOLVER_BICGSTAB, lym.OP_CLOVER, 1),
                                                (lym.SOLVER_PRECOND_BICGSTAB, lym.OP_CLOVER, 1),
                                                (lym.SOLVER_MIXED_CG, lym.OP_CLOVER, 2),
                                                (lym.SOLVER_PRECOND_MIXED_CG, lym.OP_CLOVER, 1)])
def test_graph_replay_equals_eager(case, kind, op_kind, power, monkeypatch):
    """Solver iterations replayed from captured HIP graphs give the same bits
    as the same kernels queued one by one; two solves in a row reuse the
    graph (replays with host reads and other work in between)."""
    b1, b2 = case.gauss(), case.gauss()
    op = case.op(op_kind, power=power)
    eps = (1e-11 * np.linalg.norm(b1)) ** 2

    def run():
        s = lym.Solver(case.ctx, kind, precision=eps, max_steps=5000, sloppy=lym.PREC_HALF)
        out = []
        for b in (b1, b2):
            X = case.spinor()
            info = s.solve(op, case.spinor(b), X)
            out.append((X.download(), int(info.steps), bool(info.converged)))
        return out

    monkeypatch.setenv("LYM_GRAPHS", "0")
    eager = run()
    monkeypatch.delenv("LYM_GRAPHS")
    graph = run()
    for (xe, se, ce), (xg, sg, cg) in zip(eager, graph):
        assert ce and cg and se == sg
        assert np.array_equal(xe, xg)


def test_graph_reuse_across_kappa_change(case, monkeypatch):
    """kappa lives in device memory, so a captured iteration chunk is replayed
    after set_kappa; results equal the eager ones and the oracle's at the new kappa."""
    b = case.gauss()
    eps = (1e-11 * np.linalg.norm(b)) ** 2
    k2 = case.kappa * 0.9

    def run():
        op = case.op(lym.OP_CLOVER)
        s = lym.Solver(case.ctx, lym.SOLVER_PRECOND_MIXED_CG, precision=eps, max_steps=5000, sloppy=lym.PREC_HALF)
        X1, X2 = case.spinor(), case.spinor()
        s.solve(op, case.spinor(b), X1)
        op.set_kappa(k2)
        info = s.solve(op, case.spinor(b), X2)
        return X1.download(), X2.download(), info

    monkeypatch.setenv("LYM_GRAPHS", "0")
    e1, e2, ie = run()
    monkeypatch.delenv("LYM_GRAPHS")
    g1, g2, ig = run()
    assert np.array_equal(e1, g1) and np.array_equal(e2, g2) and ie.steps == ig.steps
    fld = OracleField(case.dims, case.orc.U, kappa=k2, csw=case.csw, need_clover=True)
    xo, oinfo, ok = fld.solve("precond_bicgstab", b, eps, kind=K_CLOVER, max_steps=3000)
    assert ok
    assert np.linalg.norm((g2 - xo).ravel()) / np.linalg.norm(xo.ravel()) < 1e-9


@pytest.mark.parametrize("kind,power,gamma5", [(lym.OP_CLOVER, 1, False), (lym.OP_CLOVER, 1, True),
                                               (lym.OP_WILSON, 2, True)])
def test_shifted_bicgstab(case, kind, power, gamma5):
    """BiConjugateGradient's shifted overload (BiConjugateGradient.cpp:510-690):
    plain BiCGStab on Op + shift, against the oracle's, with (Op + shift) x = b."""
    b = case.gauss()
    shift = 0.05 - 0.02j
    eps = (1e-11 * np.linalg.norm(b)) ** 2
    s = lym.Solver(case.ctx, lym.SOLVER_BICGSTAB, precision=eps, max_steps=3000, shift=shift)
    op = case.op(kind, power=power, gamma5=gamma5)
    X = case.spinor()
    info = s.solve(op, case.spinor(b), X)
    assert info.converged
    x = X.download()
    okind = K_CLOVER if kind == lym.OP_CLOVER else K_WILSON
    Ax = case.orc.apply(okind, x, power=power, gamma5=gamma5)
    assert np.linalg.norm((Ax + shift * x - b).ravel()) / np.linalg.norm(b.ravel()) <= 1e-10
    xo, oinfo, ok = case.orc.solve("bicgstab", b, eps, kind=okind, power=power, gamma5=gamma5, max_steps=3000,
                                   shift=shift)
    assert ok
    assert np.linalg.norm((x - xo).ravel()) / np.linalg.norm(xo.ravel()) < 1e-9
    # other solver kinds and the even-odd operator refuse a shift
    with pytest.raises(lym.LymError):
        lym.Solver(case.ctx, lym.SOLVER_CG, shift=0.1).solve(case.op(lym.OP_CLOVER, power=2), case.spinor(b),
                                                             case.spinor())
    with pytest.raises(lym.LymError):
        s.solve(case.op(lym.OP_EO_CLOVER), case.spinor(b), case.spinor())


@pytest.mark.parametrize("kind,op_kind,power", [(lym.SOLVER_MIXED_CG, lym.OP_CLOVER, 2),
                                                (lym.SOLVER_PRECOND_MIXED_CG, lym.OP_CLOVER, 1)])
@pytest.mark.parametrize("sloppy", [lym.PREC_SINGLE, lym.PREC_HALF])
def test_mixed_cg_zero_source(case, kind, op_kind, power, sloppy):
    """A zero source has the solution x = 0: the mixed-precision CG stops at
    step 0 with a finite, all-zero solution, as the batched form does (the
    first alpha would otherwise be 0/0)."""
    z = np.zeros((case.V, 4, 3), complex)
    solver = lym.Solver(case.ctx, kind, precision=1e-20, max_steps=200, sloppy=sloppy)
    X = case.spinor(np.full((case.V, 4, 3), 7.0 + 1j))
    info = solver.solve(case.op(op_kind, power=power), case.spinor(z), X)
    x = X.download()
    assert info.converged and info.steps == 0
    assert np.all(np.isfinite(x)) and np.count_nonzero(x) == 0
    # the solver is reusable afterwards on a real source
    b = case.gauss()
    eps = (1e-11 * np.linalg.norm(b)) ** 2
    solver = lym.Solver(case.ctx, kind, precision=eps, max_steps=20000, sloppy=sloppy)
    info = solver.solve(case.op(op_kind, power=power), case.spinor(b), X)
    assert info.converged and info.steps > 0


def test_graph_captures_follow_operator_lifetimes(case, monkeypatch):
    """One solver, a sequence of operators created and destroyed in turn (so a
    later one may reuse an earlier one's addresses): captured iterations are
    dropped with their operator and never replayed for another, and every
    solve equals the same solve with the kernels queued one by one."""
    b = case.gauss()
    eps = (1e-11 * np.linalg.norm(b)) ** 2
    kappas = [case.kappa * f for f in (1.0, 0.95, 0.9, 0.95, 1.0, 0.85)]

    def run():
        s = lym.Solver(case.ctx, lym.SOLVER_PRECOND_MIXED_CG, precision=eps, max_steps=5000, sloppy=lym.PREC_HALF)
        out = []
        for k in kappas:
            op = case.op(lym.OP_CLOVER, kappa=k)
            X = case.spinor()
            info = s.solve(op, case.spinor(b), X)
            out.append((X.download(), int(info.steps)))
            op.close()
        return out

    monkeypatch.setenv("LYM_GRAPHS", "0")
    eager = run()
    monkeypatch.delenv("LYM_GRAPHS")
    graph = run()
    for (xe, se), (xg, sg) in zip(eager, graph):
        assert se == sg and np.array_equal(xe, xg)
