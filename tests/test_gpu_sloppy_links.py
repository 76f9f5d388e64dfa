"""Low-precision link copies.  SU(3) and SO(3) links are stored in FP32/FP16
as rows 0 and 1 (row 2 rebuilt in the stencil); links whose row 2 is not the
rebuilt one keep all three rows.  Either way the mixed-precision CG reaches the
FP64 stopping rule, the oracle's solution, and about the FP64 CG's iteration
count (a sloppy operator that differs from the FP64 one would need many more
iterations or stall)."""
import numpy as np
import pytest

from helpers import K_CLOVER, OracleField, fold_bc

import leonardym_amd as lym

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("rep,scale,sloppy", [("su3", 1.0, lym.PREC_HALF), ("su3", 0.9, lym.PREC_HALF),
                                              ("su3", 0.9, lym.PREC_SINGLE), ("adj", 1.0, lym.PREC_HALF),
                                              ("adj", 0.9, lym.PREC_HALF), ("adj", 0.9, lym.PREC_SINGLE)])
def test_mixed_cg_with_reconstructed_and_full_low_precision_links(rep, scale, sloppy):
    if lym.device_count() < 1:
        pytest.fail("no HIP device visible")
    dims = (4, 4, 4, 8)
    V = int(np.prod(dims))
    print("links", rep, "scale", scale, "sloppy", sloppy, flush=True)
    if rep == "su3":
        U = lym.gen_su3_links(dims, 0.5, 71) * scale
        ctx = lym.Context(dims, lym.REP_SU3_FUND, -1)
        kappa = 0.12
    else:
        U = lym.su2_to_adjoint(lym.gen_su2_links(dims, 0.5, 71)) * scale
        ctx = lym.Context(dims, lym.REP_SU2_ADJ, -1)
        kappa = 0.17
    Uf = fold_bc(dims, U, -1)
    gauge = lym.Gauge(ctx, Uf, folded=True)
    orc = OracleField(dims, Uf.astype(complex), kappa=kappa, csw=1.0, need_clover=True)
    op = lym.Operator(ctx, lym.OP_CLOVER, 1, kappa, 1.0, True, gauge=gauge)
    b = lym.gaussian_spinor(V, 7171)
    eps = (1e-11 * np.linalg.norm(b)) ** 2
    mixed = lym.Solver(ctx, lym.SOLVER_PRECOND_MIXED_CG, precision=eps, max_steps=4000, sloppy=sloppy)
    fp64 = lym.Solver(ctx, lym.SOLVER_PRECOND_CG, precision=eps, max_steps=4000)
    xm, x64 = lym.Spinor(ctx), lym.Spinor(ctx)
    im = mixed.solve(op, lym.Spinor(ctx, b), xm)
    i64 = fp64.solve(op, lym.Spinor(ctx, b), x64)
    print("mixed %d it, FP64 %d it" % (im.steps, i64.steps))
    assert im.converged and i64.converged
    xm, x64 = xm.download(), x64.download()
    Dx = orc.apply(K_CLOVER, xm, gamma5=False)
    assert np.linalg.norm((b - Dx).ravel()) / np.linalg.norm(b.ravel()) <= 1e-10
    assert np.linalg.norm((xm - x64).ravel()) / np.linalg.norm(x64.ravel()) < 1e-9
    assert im.steps <= 1.2 * i64.steps + 5
