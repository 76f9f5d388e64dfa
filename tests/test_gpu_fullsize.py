"""Full-size checks (the 32^3x64 SU(3) and 24^3x48 SU(2)-adjoint workloads)
through size-independent properties, with no oracle: gamma5-hermiticity of Q,
the square operator equal to two applies, D x = b for the even-odd driver's
solution (true residual from the full-lattice apply), and a batched solve
equal bit for bit to the same RHS solved alone."""
import numpy as np
import pytest

import leonardym_amd as lym

pytestmark = pytest.mark.gpu

FULL = [dict(dims=(32, 32, 32, 64), rep="su3", kappa=0.134, seed=1003),
        dict(dims=(24, 24, 24, 48), rep="adj", kappa=0.17, seed=1002)]


class Big:
    def __init__(self, dims, rep, kappa, seed):
        self.dims, self.V, self.kappa = dims, int(np.prod(dims)), kappa
        if rep == "su3":
            U = lym.gen_su3_links(dims, 0.5, seed)
            self.ctx = lym.Context(dims, lym.REP_SU3_FUND, -1)
        else:
            U = lym.su2_to_adjoint(lym.gen_su2_links(dims, 0.5, seed))
            self.ctx = lym.Context(dims, lym.REP_SU2_ADJ, -1)
        self.gauge = lym.Gauge(self.ctx, U, folded=False)

    def op(self, kind=lym.OP_CLOVER, power=1, gamma5=True):
        return lym.Operator(self.ctx, kind, power, self.kappa, 1.0, gamma5, gauge=self.gauge)

    def spinor(self, seed=None):
        s = lym.Spinor(self.ctx)
        if seed is not None:
            s.upload(lym.gaussian_spinor(self.V, seed))
        return s


@pytest.fixture(scope="module", params=range(len(FULL)), ids=["su3_32x64", "adj_24x48"])
def big(request):
    if lym.device_count() < 1:
        pytest.fail("no HIP device visible")
    print("full-size case", FULL[request.param], flush=True)
    return Big(**FULL[request.param])


def test_gamma5_hermiticity_and_square(big):
    a, b = big.spinor(11), big.spinor(12)
    Q = big.op()
    Qa, Qb = big.spinor(), big.spinor()
    Q.multiply(Qa, a)
    Q.multiply(Qb, b)
    lhs, rhs = b.dot(Qa), Qb.dot(a)
    scale = np.sqrt(a.squared_norm() * Qb.squared_norm())
    assert abs(lhs - rhs) < 1e-12 * scale
    QQa, Q2a = big.spinor(), big.spinor()
    Q.multiply(QQa, Qa)
    big.op(power=2).multiply(Q2a, a)
    assert np.sqrt(QQa.difference_norm(Q2a) / QQa.squared_norm()) < 1e-14


def test_even_odd_solve_true_residual(big):
    b = big.spinor(13)
    x, Dx = big.spinor(), big.spinor()
    s = lym.Solver(big.ctx, lym.SOLVER_PRECOND_MIXED_CG, precision=1e-22 * b.squared_norm(), max_steps=4000,
                   sloppy=lym.PREC_HALF)
    info = s.solve(big.op(), b, x)
    assert info.converged
    big.op(gamma5=False).multiply(Dx, x)
    assert np.sqrt(Dx.difference_norm(b) / b.squared_norm()) <= 1e-10


def test_batched_equals_alone_full_size(big):
    srcs = [lym.Spinor(big.ctx, lym.point_source(big.dims, (1, 2, 3, 4), sc, 0)) for sc in range(3)]
    op = big.op()
    alone = lym.Solver(big.ctx, lym.SOLVER_PRECOND_MIXED_CG, precision=1e-22, max_steps=4000, sloppy=lym.PREC_HALF)
    alone.set_rhs_block(1)
    batch = lym.Solver(big.ctx, lym.SOLVER_PRECOND_MIXED_CG, precision=1e-22, max_steps=4000, sloppy=lym.PREC_HALF)
    batch.set_rhs_block(4)
    xa = [big.spinor() for _ in srcs]
    xb = [big.spinor() for _ in srcs]
    ia = alone.solve_batch(op, srcs, xa)
    ib = batch.solve_batch(op, srcs, xb)
    for q in range(3):
        assert ia[q].converged and ib[q].converged and ia[q].steps == ib[q].steps
        assert np.array_equal(xa[q].download(), xb[q].download())
