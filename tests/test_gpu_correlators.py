"""Point-source propagators and connected meson correlators on the device
(lym_meson_correlators; MesonCorrelator::execute,
source/correlators/MesonCorrelator.cpp:22-199) against the oracle's 12 solves
and contractions on the same seeded 8^4 fields.  The oracle's contractions are
pinned on the CPU (free-field closed form, trace forms, gauge invariance)."""
import numpy as np
import pytest

from helpers import K_CLOVER, K_WILSON, OracleField, fold_bc

import leonardym_amd as lym

pytestmark = pytest.mark.gpu

CHANNELS = ("pion", "scalar", "ps", "pcac")


def compare(got, want, tol):
    for k in CHANNELS:
        scale = np.max(np.abs(want[k]))
        err = np.max(np.abs(got[k] - want[k]))
        print("%s: max |diff| %.2e of scale %.2e" % (k, err, scale))
        assert err <= tol * scale, k


@pytest.mark.parametrize("kind,sloppy", [(lym.SOLVER_PRECOND_BICGSTAB, None),
                                         (lym.SOLVER_PRECOND_MIXED_CG, lym.PREC_HALF),
                                         (lym.SOLVER_PRECOND_CG, None)])
def test_clover_correlators_match_oracle(kind, sloppy):
    if lym.device_count() < 1:
        pytest.fail("no HIP device visible")
    dims, kappa, csw = (8, 8, 8, 8), 0.13, 1.0
    site = (1, 2, 3, 5)
    U = fold_bc(dims, lym.gen_su3_links(dims, 0.5, 1001))
    ctx = lym.Context(dims)
    g = lym.Gauge(ctx, U, folded=True)
    op = lym.Operator(ctx, lym.OP_CLOVER, 1, kappa, csw, True, gauge=g)
    eps = 1e-24  # |b| = 1
    s = lym.Solver(ctx, kind, precision=eps, max_steps=5000, sloppy=sloppy)
    got = lym.meson_correlators(s, op, site)
    assert all(i.converged for i in got["infos"])
    orc = OracleField(dims, U, kappa=kappa, csw=csw, need_clover=True)
    want = orc.meson_correlators(site, eps)
    assert want["ok"]
    compare(got, want, 1e-9)
    # gamma5 of the operator is restored: it still applies Q
    psi = lym.gaussian_spinor(int(np.prod(dims)), 3)
    out = lym.Spinor(ctx)
    op.multiply(out, lym.Spinor(ctx, psi))
    ref = orc.apply(K_CLOVER, psi, gamma5=True)
    assert np.max(np.abs(out.download() - ref)) < 1e-12 * np.max(np.abs(ref))


def test_wilson_correlators_match_oracle():
    """The Wilson operator at the plain-path workload (8^4, kappa 0.12, eps 0.4):
    BiCGStab on D for every column."""
    dims, kappa = (8, 8, 8, 8), 0.12
    site = (0, 0, 0, 0)
    U = fold_bc(dims, lym.gen_su3_links(dims, 0.4, 1000))
    ctx = lym.Context(dims)
    g = lym.Gauge(ctx, U, folded=True)
    op = lym.Operator(ctx, lym.OP_WILSON, 1, kappa, 0.0, True, gauge=g)
    s = lym.Solver(ctx, lym.SOLVER_BICGSTAB, precision=1e-24, max_steps=5000)
    got = lym.meson_correlators(s, op, site)
    want = OracleField(dims, U, kappa=kappa).meson_correlators(site, 1e-24, kind=K_WILSON)
    assert want["ok"]
    compare(got, want, 1e-9)
    # a solver that does not return D^-1 b is refused
    with pytest.raises(lym.LymError):
        lym.meson_correlators(lym.Solver(ctx, lym.SOLVER_CG), op, site)


def test_correlators_on_an_odd_extent():
    """A lattice with odd spatial extents and odd Lt (one site class on the
    device): the point source, the 12 FP64 BiCGStab solves and the timeslice
    contractions against the oracle."""
    dims, kappa, csw = (5, 4, 3, 7), 0.12, 1.0
    site = (4, 1, 2, 6)
    U = fold_bc(dims, lym.gen_su3_links(dims, 0.4, 1002))
    ctx = lym.Context(dims)
    g = lym.Gauge(ctx, U, folded=True)
    op = lym.Operator(ctx, lym.OP_CLOVER, 1, kappa, csw, True, gauge=g)
    s = lym.Solver(ctx, lym.SOLVER_BICGSTAB, precision=1e-24, max_steps=5000)
    got = lym.meson_correlators(s, op, site)
    assert all(i.converged for i in got["infos"])
    want = OracleField(dims, U, kappa=kappa, csw=csw, need_clover=True).meson_correlators(site, 1e-24)
    assert want["ok"]
    compare(got, want, 1e-9)
