"""The matrix-core (MFMA) form of the multi-RHS FP16 stencil.

Each hop's link product, spin projection and reconstruction run as one
v_mfma_f32_16x16x16_f16 per site over the 8 (or 4, 2) right-hand sides of
the site, with the FP16 site scales applied to the FP32 product.  Per site it
is held to the same bound as the vector form of the FP16 stencil against the
FP64 apply (the oracle-checked operator), and against the vector form itself
to a few FP16 output roundings; then batched even-odd solves with it reach
the FP64 stopping rule and the vector form's solutions, and an RHS gives the
same bits in batches of 2, 4 and 8 (the MFMA computes each column from that
column alone)."""
import numpy as np
import pytest

from helpers import K_CLOVER, coords, fold_bc
from test_gpu_launch_shapes import LOWP, neighbour_scale
from test_gpu_parity import Case, true_residual

import leonardym_amd as lym

pytestmark = pytest.mark.gpu

KAPPA, CSW = 0.13, 1.0

SHAPES = [(4, 4, 4, 4), (6, 4, 4, 8), (8, 8, 8, 8), (16, 16, 16, 32)]


def setup(dims, bc=-1, storage=lym.LINKS_FULL, seed=31):
    ctx = lym.Context(dims, lym.REP_SU3_FUND, bc)
    U = fold_bc(dims, lym.gen_su3_links(dims, 0.5, seed), bc)
    g = lym.Gauge(ctx, U, folded=True, storage=storage)
    return ctx, g


# 16^3x32 runs only at the batch width the solver uses there, 8 RHS
DIMS_NR = [(d, nr) for d in SHAPES for nr in (8, 4, 2) if d != (16, 16, 16, 32) or nr == 8]


@pytest.mark.parametrize("dims,nr", DIMS_NR, ids=lambda v: "x".join(map(str, v)) if isinstance(v, tuple) else str(v))
@pytest.mark.parametrize("kind,power,gamma5", [(lym.OP_CLOVER, 1, True), (lym.OP_EO_CLOVER, 1, True),
                                               (lym.OP_WILSON, 1, False), (lym.OP_EO_CLOVER, 2, True)])
def test_mfma_apply_per_site(dims, nr, kind, power, gamma5):
    print("mfma apply: dims %s nr %d kind %d power %d gamma5 %d" % (dims, nr, kind, power, gamma5), flush=True)
    V = int(np.prod(dims))
    ctx, g = setup(dims)
    op = lym.Operator(ctx, kind, power, KAPPA, CSW, gamma5, gauge=g)
    psis = [lym.gaussian_spinor(V, 500 + q) for q in range(nr)]
    ins = [lym.Spinor(ctx, p) for p in psis]
    nan = np.full((V, 4, 3), np.nan + 1j * np.nan)
    mf = [lym.Spinor(ctx, nan) for _ in range(nr)]
    vec = [lym.Spinor(ctx, nan) for _ in range(nr)]
    op.multiply_sloppy_batch(mf, ins, lym.PREC_HALF, matrix_cores=True)
    op.multiply_sloppy_batch(vec, ins, lym.PREC_HALF, matrix_cores=False)
    ref = lym.Spinor(ctx)
    par = coords(dims).sum(1) % 2
    worst_ref = worst_vec = 0.0
    for q in range(nr):
        op.multiply(ref, ins[q])
        r, x, v = ref.download(), mf[q].download(), vec[q].download()
        assert np.all(np.isfinite(x)), "sites left unwritten"
        n = np.linalg.norm(psis[q].reshape(V, -1), axis=1)
        s = neighbour_scale(dims, n)
        if kind == lym.OP_EO_CLOVER:
            s = 2 * neighbour_scale(dims, s)
            if power == 2:
                s = 4 * neighbour_scale(dims, neighbour_scale(dims, s))
            assert np.array_equal(x[par == 0], psis[q][par == 0])
        rn = np.linalg.norm(r.reshape(V, -1), axis=1)
        bound = LOWP[lym.PREC_HALF] * (s + rn) * (2 if power == 2 else 1)
        worst_ref = max(worst_ref, float(np.max(np.linalg.norm((x - r).reshape(V, -1), axis=1) / bound)))
        if kind != lym.OP_EO_CLOVER:
            # one launch: both forms round the same FP16 inputs and differ only in
            # the order of their FP32 sums, so the stored FP16 outputs differ by
            # at most a few roundings of the site's largest component
            vn = np.max(np.abs(np.concatenate([v.real, v.imag], axis=-1)).reshape(V, -1), axis=1)
            d = np.max(np.abs(x - v).reshape(V, -1), axis=1)
            worst_vec = max(worst_vec, float(np.max(d / (4 * 2.0 ** -11 * vn + 1e-300))))
        else:
            # two or four launches: the FP16 intermediates differ by a rounding
            # here and there and carry it through the next hops
            worst_vec = max(worst_vec, float(np.linalg.norm(x - v) / np.linalg.norm(v)) / 2e-3)
    print("mfma: worst site at %.3f of the FP16 bound against FP64, %.3f of the bound against the vector form"
          % (worst_ref, worst_vec))
    assert worst_ref <= 1.0
    assert worst_vec <= 1.0


def test_mfma_periodic_and_unsupported():
    dims = (4, 4, 4, 8)
    V = int(np.prod(dims))
    ctx, g = setup(dims, bc=1)
    op = lym.Operator(ctx, lym.OP_CLOVER, 1, KAPPA, CSW, True, gauge=g)
    ins = [lym.Spinor(ctx, lym.gaussian_spinor(V, 600 + q)) for q in range(4)]
    mf = [lym.Spinor(ctx) for _ in range(4)]
    ref = lym.Spinor(ctx)
    op.multiply_sloppy_batch(mf, ins, lym.PREC_HALF, matrix_cores=True)
    for q in range(4):
        op.multiply(ref, ins[q])
        r = ref.download()
        assert np.linalg.norm(mf[q].download() - r) / np.linalg.norm(r) < 2e-3
    # widths and precisions without a matrix-core form are refused with a code
    outs3 = [lym.Spinor(ctx) for _ in range(3)]
    with pytest.raises(lym.LymError):
        op.multiply_sloppy_batch(outs3, ins[:3], lym.PREC_HALF, matrix_cores=True)
    with pytest.raises(lym.LymError):
        op.multiply_sloppy_batch(mf, ins, lym.PREC_SINGLE, matrix_cores=True)
    ctx12, g12 = setup(dims, storage=lym.LINKS_RECONSTRUCT)
    op12 = lym.Operator(ctx12, lym.OP_CLOVER, 1, KAPPA, CSW, True, gauge=g12)
    a12 = [lym.Spinor(ctx12, lym.gaussian_spinor(V, 700 + q)) for q in range(2)]
    with pytest.raises(lym.LymError):
        op12.multiply_sloppy_batch([lym.Spinor(ctx12) for _ in range(2)], a12, lym.PREC_HALF, matrix_cores=True)


@pytest.mark.parametrize("dims", [(4, 4, 4, 4), (8, 8, 8, 8)], ids=lambda d: "x".join(map(str, d)))
def test_mfma_batched_solves(dims):
    """Eight RHS through the even-odd mixed CG driver with the matrix-core
    stencil: batches of 8, 4 and 2 give the same bits, every solution meets
    the FP64 stopping rule and agrees with the vector form's solution."""
    print("mfma batched solves on", dims, flush=True)
    case = Case(dims=dims)
    srcs = [case.gauss() for _ in range(8)]
    eps = [(1e-11 * np.linalg.norm(b)) ** 2 for b in srcs]
    op = case.op(lym.OP_CLOVER)
    sols = {}
    for block, mc in ((8, True), (4, True), (2, True), (8, False)):
        solver = lym.Solver(case.ctx, lym.SOLVER_PRECOND_MIXED_CG, max_steps=20000, sloppy=lym.PREC_HALF)
        solver.set_rhs_block(block)
        solver.set_matrix_cores(mc)
        B = [case.spinor(b) for b in srcs]
        X = [case.spinor() for _ in srcs]
        infos = solver.solve_batch(op, B, X, eps)
        assert all(i.converged for i in infos)
        assert solver.matrix_cores() == (mc, mc)
        sols[(block, mc)] = ([x.download() for x in X], [int(i.steps) for i in infos])
    x8, st8 = sols[(8, True)]
    for key in ((4, True), (2, True)):
        xs, st = sols[key]
        assert st == st8, key
        assert all(np.array_equal(a, b) for a, b in zip(xs, x8)), key
    xv, _ = sols[(8, False)]
    for q, b in enumerate(srcs):
        assert true_residual(case, x8[q], b) <= 1e-10
        assert np.linalg.norm((x8[q] - xv[q]).ravel()) / np.linalg.norm(xv[q].ravel()) < 1e-9


def test_mfma_autotuned_point_sources_16x16x16x32():
    """The 12 point sources of a meson measurement on 16^3x32 with the
    autotuner free to pick the matrix-core form: every RHS converges to the
    FP64 stopping rule."""
    dims = (16, 16, 16, 32)
    V = int(np.prod(dims))
    ctx, g = setup(dims, seed=41)
    op = lym.Operator(ctx, lym.OP_CLOVER, 1, KAPPA, CSW, True, gauge=g)
    D = lym.Operator(ctx, lym.OP_CLOVER, 1, KAPPA, CSW, False, gauge=g)
    srcs = [lym.Spinor(ctx, lym.point_source(dims, (3, 5, 7, 9), sc // 3, sc % 3)) for sc in range(12)]
    sols = [lym.Spinor(ctx) for _ in range(12)]
    solver = lym.Solver(ctx, lym.SOLVER_PRECOND_MIXED_CG, precision=1e-22, max_steps=4000, sloppy=lym.PREC_HALF)
    solver.set_matrix_cores(True)
    infos = solver.solve_batch(op, srcs, sols)
    print("autotuned: %s RHS/batch, matrix cores %s, iterations %s" % (
        solver.rhs_block(), solver.matrix_cores(), [int(i.steps) for i in infos]))
    Dx = lym.Spinor(ctx)
    for q in range(12):
        assert infos[q].converged
        D.multiply(Dx, sols[q])
        assert np.sqrt(Dx.difference_norm(srcs[q])) <= 1e-10
