"""Lattices with an odd extent (the reference's LocalLayout takes any extents
for the operators without even-odd preconditioning,
source/MPILattice/LocalLayout.h:34-110).  Such a lattice has no two-colouring
under periodic wrapping, so the device keeps all sites in one class in
lexicographic order and every hop stays in it.  The full-lattice FP64
operators (Wilson, basic Wilson, clover; Q and D, multiplyAdd, squares) must
match the oracle per site to 1e-13, and FP64 CG and BiCGStab must reach true
residual 1e-10 and the oracle's solution to 1e-9 -- on lattices with an odd
spatial extent (which the even-odd layout would wrap onto the wrong parity),
an odd Lt, all extents odd, 12-real links, the SU(2) adjoint representation,
periodic and antiperiodic time, and one lattice large enough for the
XCD-pair block mapping and the LDS-staged stencil.  Everything that needs the
even-odd layout is refused with an error code."""
import numpy as np
import pytest

from helpers import K_BASIC, K_CLOVER, K_WILSON, rel_err_per_site

import leonardym_amd as lym
from test_gpu_parity import SITE_TOL, Case, g5, true_residual

pytestmark = pytest.mark.gpu

ODD_CASES = [
    dict(dims=(3, 4, 4, 4)),
    dict(dims=(4, 4, 4, 5)),
    dict(dims=(5, 5, 5, 5), bc=1, storage=lym.LINKS_RECONSTRUCT),
    dict(dims=(3, 3, 5, 4), rep="adj", kappa=0.17),
    dict(dims=(8, 9, 16, 32)),
]


@pytest.fixture(scope="module", params=range(len(ODD_CASES)), ids=lambda i: "x".join(map(str, ODD_CASES[i]["dims"])))
def case(request):
    if lym.device_count() < 1:
        pytest.fail("no HIP device visible")
    print("odd-extent case", ODD_CASES[request.param], flush=True)
    return Case(**ODD_CASES[request.param])


def test_layout_plan():
    """Every lattice above runs one parity-0 launch over all its sites; the
    large one takes the XCD-pair renumbering."""
    for c in ODD_CASES:
        r = lym.launch_plan(c["dims"], 0)
        assert r["spb"] * r["nbpp"] >= int(np.prod(c["dims"])) > r["spb"] * (r["nbpp"] - 1)
        assert np.array_equal(r["visits"][0], np.ones(r["nbpp"], np.int32))
    assert lym.launch_plan((8, 9, 16, 32), 0)["branches"][2] > 0


def test_spinor_roundtrip_and_blas(case):
    a, b = case.gauss(), case.gauss()
    A, B = case.spinor(a), case.spinor(b)
    assert np.array_equal(A.download(), a)
    assert abs(A.squared_norm() - np.vdot(a, a).real) < 1e-12 * np.vdot(a, a).real
    ref = np.vdot(a.ravel(), b.ravel())
    assert abs(A.dot(B) - ref) < 1e-12 * abs(ref) + 1e-12
    A.axpy(0.5 - 2j, B)
    A.gamma5()
    assert np.max(np.abs(A.download() - g5(a + (0.5 - 2j) * b))) < 1e-14


@pytest.mark.parametrize("kind", [lym.OP_WILSON, lym.OP_BASIC_WILSON, lym.OP_CLOVER])
@pytest.mark.parametrize("gamma5", [True, False])
def test_multiply_per_site(case, kind, gamma5):
    psi = case.gauss()
    okind = {lym.OP_WILSON: K_WILSON, lym.OP_BASIC_WILSON: K_BASIC, lym.OP_CLOVER: K_CLOVER}[kind]
    A = case.spinor(psi)
    B = case.spinor(np.full_like(psi, np.nan))  # every site must be written
    case.op(kind, gamma5=gamma5).multiply(B, A)
    got = B.download()
    assert np.all(np.isfinite(got))
    assert rel_err_per_site(got, case.orc.apply(okind, psi, gamma5=gamma5)) < SITE_TOL


@pytest.mark.parametrize("kind", [lym.OP_WILSON, lym.OP_CLOVER])
def test_multiply_add_and_square(case, kind):
    v1, v2 = case.gauss(), case.gauss()
    okind = K_CLOVER if kind == lym.OP_CLOVER else K_WILSON
    A, B, C = case.spinor(v1), case.spinor(v2), case.spinor()
    case.op(kind, gamma5=False).multiply_add(C, A, B, 0.3 - 0.8j)
    ref = case.orc.apply(okind, v1, gamma5=False, v2=v2, alpha=0.3 - 0.8j)
    assert rel_err_per_site(C.download(), ref) < SITE_TOL
    sq = case.op(kind, power=2)
    sq.multiply(C, A)
    assert rel_err_per_site(C.download(), case.orc.apply(okind, v1, power=2)) < 10 * SITE_TOL


def test_cg_on_square_clover(case):
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


def test_bicgstab(case):
    b = case.gauss()
    eps = (1e-11 * np.linalg.norm(b)) ** 2
    solver = lym.Solver(case.ctx, lym.SOLVER_BICGSTAB, precision=eps, max_steps=3000)
    X = case.spinor()
    info = solver.solve(case.op(lym.OP_CLOVER, gamma5=False), case.spinor(b), X)
    assert info.converged
    x = X.download()
    assert true_residual(case, x, b) <= 1e-10
    xo, oinfo, ok = case.orc.solve("bicgstab", b, eps, kind=K_CLOVER, gamma5=False, max_steps=3000)
    assert ok
    assert np.linalg.norm((x - xo).ravel()) / np.linalg.norm(xo.ravel()) < 1e-9
    # zero source: the seeded random start fills the lattice's one class only
    z = case.spinor(np.zeros_like(b))
    X0 = case.spinor()
    info = lym.Solver(case.ctx, lym.SOLVER_BICGSTAB, precision=1e-20, max_steps=3000).solve(
        case.op(lym.OP_CLOVER, gamma5=False), z, X0)
    assert info.converged and np.linalg.norm(X0.download()) < 1e-8


def test_even_odd_forms_are_refused(case):
    b = case.spinor(case.gauss())
    with pytest.raises(lym.LymError):
        case.op(lym.OP_EO_CLOVER)
    op = case.op(lym.OP_CLOVER)
    with pytest.raises(lym.LymError):
        lym.Solver(case.ctx, lym.SOLVER_MIXED_CG, max_steps=100).solve(op, b, case.spinor())
    with pytest.raises(lym.LymError):
        lym.Solver(case.ctx, lym.SOLVER_PRECOND_MIXED_CG, max_steps=100).solve_batch(op, [b], [case.spinor()])
    with pytest.raises(lym.LymError):
        op.multiply_sloppy(case.spinor(), b, lym.PREC_HALF)
    # the operator still works after the refusals
    out = case.spinor()
    op.multiply(out, b)
    assert np.all(np.isfinite(out.download()))


def test_server_refuses_odd_lattice():
    dims = (3, 4, 4, 4)
    srv = lym.Server(workers=1)
    try:
        with pytest.raises(lym.LymError):
            srv.add_lattice(dims, lym.gen_su3_links(dims, 0.5, 3))
    finally:
        srv.close()
