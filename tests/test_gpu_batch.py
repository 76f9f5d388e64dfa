"""Batched multi-RHS solves: every RHS of a batch reaches the FP64 stopping
rule and the oracle's solution, and gives the same bits whatever batch it is
solved in (batches of 1, 4 or the autotuned size; ragged last batches)."""
import numpy as np
import pytest

from helpers import K_CLOVER, K_EO, coords

import leonardym_amd as lym
from test_gpu_parity import Case, g5, true_residual

pytestmark = pytest.mark.gpu

BATCH_CASES = [
    dict(dims=(4, 4, 4, 4)),
    dict(dims=(4, 6, 2, 8)),
    dict(dims=(4, 4, 4, 8), storage=lym.LINKS_RECONSTRUCT),
    dict(dims=(4, 4, 4, 4), rep="adj", kappa=0.17),
]


@pytest.fixture(scope="module", params=range(len(BATCH_CASES)), ids=lambda i: "bcase%d" % i)
def case(request):
    if lym.device_count() < 1:
        pytest.fail("no HIP device visible")
    print("batch case", BATCH_CASES[request.param], flush=True)
    return Case(**BATCH_CASES[request.param])


def solve_batch(case, op, kind, srcs, block, precisions=None, sloppy=lym.PREC_HALF, eps=None, max_steps=20000):
    solver = lym.Solver(case.ctx, kind, precision=eps, max_steps=max_steps, sloppy=sloppy)
    solver.set_rhs_block(block)
    B = [case.spinor(b) for b in srcs]
    X = [case.spinor() for _ in srcs]
    infos = solver.solve_batch(op, B, X, precisions)
    return [x.download() for x in X], infos, solver


@pytest.mark.parametrize("sloppy", [lym.PREC_HALF, lym.PREC_SINGLE])
def test_even_odd_batch_matches_alone_and_oracle(case, sloppy):
    """Seven RHS through the even-odd mixed CG driver: batches of 4 (+3 ragged),
    of 1 and autotuned give identical bits; each solution satisfies D x = b."""
    srcs = [case.gauss() for _ in range(7)]
    eps = [(1e-11 * np.linalg.norm(b)) ** 2 for b in srcs]
    op = case.op(lym.OP_CLOVER)
    xs4, inf4, _ = solve_batch(case, op, lym.SOLVER_PRECOND_MIXED_CG, srcs, 4, eps, sloppy)
    xs1, inf1, _ = solve_batch(case, op, lym.SOLVER_PRECOND_MIXED_CG, srcs, 1, eps, sloppy)
    xsa, infa, sa = solve_batch(case, op, lym.SOLVER_PRECOND_MIXED_CG, srcs, 0, eps, sloppy)
    nr, thr = sa.rhs_block()
    print("autotuned batch: %d RHS, %d threads" % (nr, thr))
    assert nr in (4, 6, 8, 12, 16)
    for q, b in enumerate(srcs):
        assert inf4[q].converged and inf1[q].converged and infa[q].converged
        assert inf4[q].steps == inf1[q].steps == infa[q].steps
        assert np.array_equal(xs4[q], xs1[q]) and np.array_equal(xsa[q], xs1[q])
        assert true_residual(case, xs4[q], b) <= 1e-10
    xo, oinfo, ok = case.orc.solve("precond_bicgstab", srcs[0], eps[0], kind=K_CLOVER, max_steps=3000)
    assert ok
    assert np.linalg.norm((xs4[0] - xo).ravel()) / np.linalg.norm(xo.ravel()) < 1e-9


def test_batch_agrees_with_single_solves(case):
    """A batch and the single-RHS mixed CG solve the same systems (Q^2 x = Q g5 b)."""
    bs = [case.gauss() for _ in range(3)]
    rhs = [case.orc.apply(K_CLOVER, g5(b)) for b in bs]
    eps = (1e-11 * max(np.linalg.norm(r) for r in rhs)) ** 2
    op = case.op(lym.OP_CLOVER, power=2)
    xs, infos, _ = solve_batch(case, op, lym.SOLVER_MIXED_CG, rhs, 4, None, eps=eps)
    single = lym.Solver(case.ctx, lym.SOLVER_MIXED_CG, precision=eps, max_steps=20000, sloppy=lym.PREC_HALF)
    for q in range(3):
        X = case.spinor()
        info = single.solve(op, case.spinor(rhs[q]), X)
        assert info.converged and infos[q].converged
        assert true_residual(case, xs[q], bs[q]) <= 1e-10
        x1 = X.download()
        assert np.linalg.norm((xs[q] - x1).ravel()) / np.linalg.norm(x1.ravel()) < 1e-9
        assert abs(int(infos[q].steps) - int(info.steps)) <= max(3, info.steps // 20)


def test_even_odd_operator_batch(case):
    """Batched mixed CG on Qhat^2 (odd half); even half of each solution = source."""
    srcs = [case.gauss() for _ in range(2)]
    eps = (1e-11 * np.linalg.norm(srcs[0])) ** 2
    xs, infos, _ = solve_batch(case, case.op(lym.OP_EO_CLOVER, power=2), lym.SOLVER_MIXED_CG, srcs, 2, eps=eps)
    par = coords(case.dims).sum(1) % 2
    xo, oinfo, ok = case.orc.solve("cg", srcs[0], eps, kind=K_EO, power=2, max_steps=5000)
    assert ok and infos[0].converged
    assert np.linalg.norm((xs[0] - xo).ravel()) / np.linalg.norm(xo.ravel()) < 1e-9
    for q in range(2):
        assert np.array_equal(xs[q][par == 0], srcs[q][par == 0])


def test_batch_edge_cases(case):
    """A zero source, a loose and a tight tolerance, and max steps, in one batch."""
    b = case.gauss()
    zero = np.zeros_like(b)
    srcs = [b, zero, b]
    base = (1e-11 * np.linalg.norm(b)) ** 2
    xs, infos, _ = solve_batch(case, case.op(lym.OP_CLOVER), lym.SOLVER_PRECOND_MIXED_CG, srcs, 3,
                               [base * 1e10, base, base])
    assert infos[1].converged and infos[1].steps == 0 and not np.any(xs[1])
    assert infos[0].converged and infos[2].converged
    assert infos[0].steps < infos[2].steps
    assert true_residual(case, xs[2], b) <= 1e-10
    # the solve stops at max_steps per RHS and reports it
    xs, infos, _ = solve_batch(case, case.op(lym.OP_CLOVER), lym.SOLVER_PRECOND_MIXED_CG, [b, b], 2,
                               [1e-300, 1e-300], max_steps=5)
    assert all(not i.converged and i.steps == 5 for i in infos)
    assert all(np.all(np.isfinite(x)) for x in xs)
    # kinds without a batched form are refused with a code, not a crash
    solver = lym.Solver(case.ctx, lym.SOLVER_BICGSTAB)
    with pytest.raises(lym.LymError):
        solver.solve_batch(case.op(lym.OP_CLOVER), [case.spinor(b)], [case.spinor()])
    with pytest.raises(lym.LymError):
        lym.Solver(case.ctx, lym.SOLVER_MIXED_CG).set_rhs_block(5)


def test_batch_graph_replay_equals_eager(case, monkeypatch):
    srcs = [case.gauss() for _ in range(3)]
    eps = [(1e-11 * np.linalg.norm(b)) ** 2 for b in srcs]
    op = case.op(lym.OP_CLOVER)
    monkeypatch.setenv("LYM_GRAPHS", "0")
    xe, ie, _ = solve_batch(case, op, lym.SOLVER_PRECOND_MIXED_CG, srcs, 4, eps)
    monkeypatch.delenv("LYM_GRAPHS")
    xg, ig, _ = solve_batch(case, op, lym.SOLVER_PRECOND_MIXED_CG, srcs, 4, eps)
    for q in range(3):
        assert ie[q].steps == ig[q].steps and ie[q].converged and ig[q].converged
        assert np.array_equal(xe[q], xg[q])


@pytest.mark.parametrize("sloppy", [lym.PREC_HALF, lym.PREC_SINGLE])
def test_fused_batch_steps_against_separate_launches(case, sloppy, monkeypatch):
    """The even-odd batch with its CG sums fused into the stencil launches
    (|Qhat p|^2 per site from the launch writing Qhat p; x_s, r_s and |r_s|^2
    from the last launch's unrounded output) against the separate dot and
    update launches: both reach the FP64 stopping rule, the solutions agree to
    the solve's tolerance, the iteration counts to about ten per cent, and the
    fused form gives the same bits in batches of 4 and 1."""
    srcs = [case.gauss() for _ in range(5)]
    eps = [(1e-11 * np.linalg.norm(b)) ** 2 for b in srcs]
    op = case.op(lym.OP_CLOVER)
    monkeypatch.setenv("LYM_FUSED_BATCH", "0")
    xs_sep, inf_sep, _ = solve_batch(case, op, lym.SOLVER_PRECOND_MIXED_CG, srcs, 4, eps, sloppy)
    monkeypatch.delenv("LYM_FUSED_BATCH")
    xs4, inf4, _ = solve_batch(case, op, lym.SOLVER_PRECOND_MIXED_CG, srcs, 4, eps, sloppy)
    xs1, inf1, _ = solve_batch(case, op, lym.SOLVER_PRECOND_MIXED_CG, srcs, 1, eps, sloppy)
    print("iterations separate %s fused %s" % ([int(i.steps) for i in inf_sep], [int(i.steps) for i in inf4]))
    for q, b in enumerate(srcs):
        assert inf_sep[q].converged and inf4[q].converged and inf1[q].converged
        assert inf4[q].steps == inf1[q].steps and np.array_equal(xs4[q], xs1[q])
        # the two forms round differently (unrounded FP32 sums against stored
        # FP16 values), so the counts differ by a few iterations
        assert abs(int(inf4[q].steps) - int(inf_sep[q].steps)) <= max(6, 0.1 * inf_sep[q].steps)
        assert true_residual(case, xs4[q], b) <= 1e-10
        assert np.linalg.norm((xs4[q] - xs_sep[q]).ravel()) / np.linalg.norm(xs_sep[q].ravel()) < 1e-9


def test_fused_batch_steps_on_padded_parities():
    """6^3x8 has 864 sites per parity in a stride of 896: the fused steps' sums
    must skip the padding slots, where a wider batch (4 RHS) left values that
    the ragged batch after it (3 RHS) would otherwise add in.  Every RHS gives
    the same bits in batches of 4 (+3) as alone and meets the FP64 stopping
    rule."""
    case = Case(dims=(6, 6, 6, 8))
    assert lym.launch_plan(case.dims, 0)["nbpp"] * 256 >= 864
    srcs = [case.gauss() for _ in range(7)]
    eps = [(1e-11 * np.linalg.norm(b)) ** 2 for b in srcs]
    op = case.op(lym.OP_CLOVER)
    xs4, inf4, _ = solve_batch(case, op, lym.SOLVER_PRECOND_MIXED_CG, srcs, 4, eps)
    xs1, inf1, _ = solve_batch(case, op, lym.SOLVER_PRECOND_MIXED_CG, srcs, 1, eps)
    for q, b in enumerate(srcs):
        assert inf4[q].converged and inf1[q].converged
        assert inf4[q].steps == inf1[q].steps and np.array_equal(xs4[q], xs1[q]), q
        assert true_residual(case, xs4[q], b) <= 1e-10


def test_kappa_lanes_on_padded_parities():
    """The serving engine's lanes on 6^3x8 (padded parities): a run of 7
    sources widens the batch as they are admitted (the fields are relaid at
    each width); every source gives the same bits alone and in the mix and
    meets the FP64 stopping rule at its kappa."""
    from helpers import OracleField, fold_bc
    dims = (6, 6, 6, 8)
    case = Case(dims=dims)
    U = fold_bc(dims, lym.gen_su3_links(dims, 0.5, 11), -1)
    kap = [0.120, 0.125, 0.130, 0.120, 0.125, 0.130, 0.1275]
    srcs = [case.gauss() for _ in kap]
    eps = [(1e-11 * np.linalg.norm(b)) ** 2 for b in srcs]
    op = case.op(lym.OP_CLOVER)

    def run(idx, block=0):
        solver = lym.Solver(case.ctx, lym.SOLVER_PRECOND_MIXED_CG, max_steps=20000, sloppy=lym.PREC_HALF)
        solver.set_rhs_block(block)
        B = [case.spinor(srcs[q]) for q in idx]
        X = [case.spinor() for _ in idx]
        infos = solver.solve_batch_kappas(op, [kap[q] for q in idx], B, X, [eps[q] for q in idx])
        return [x.download() for x in X], infos

    xs, infos = run(list(range(len(kap))))
    for q in (0, 4, 6):
        alone, ia = run([q])
        assert ia[0].steps == infos[q].steps and np.array_equal(alone[0], xs[q]), q
    for q, k in enumerate(kap):
        assert infos[q].converged, q
        orc = OracleField(dims, U, kappa=k, csw=1.0, need_clover=True, need_inverse=False)
        Dx = orc.apply(K_CLOVER, xs[q], gamma5=False)
        assert np.linalg.norm((srcs[q] - Dx).ravel()) / np.linalg.norm(srcs[q].ravel()) <= 1e-10


def test_point_sources_16x16x16x32():
    """Point sources (a localised source decays below the smallest normal
    float within a few dozen hops): mixed CG with FP16 storage, alone and
    batched, reaches true residual 1e-10 and the FP64 solution."""
    if lym.device_count() < 1:
        pytest.fail("no HIP device visible")
    case = Case((16, 16, 16, 32), kappa=0.13, seed=31)
    site = (3, 5, 7, 9)
    srcs = [lym.point_source(case.dims, site, sc // 3, sc % 3) for sc in (0, 4, 11)]
    op = case.op(lym.OP_CLOVER)
    eps = 1e-22
    s64 = lym.Solver(case.ctx, lym.SOLVER_PRECOND_CG, precision=eps, max_steps=4000)
    for sloppy in (lym.PREC_HALF, lym.PREC_SINGLE):
        xs, infos, _ = solve_batch(case, op, lym.SOLVER_PRECOND_MIXED_CG, srcs, 2, None, sloppy, eps=eps, max_steps=4000)
        single = lym.Solver(case.ctx, lym.SOLVER_PRECOND_MIXED_CG, precision=eps, max_steps=4000, sloppy=sloppy)
        for q, b in enumerate(srcs):
            X1, X64 = case.spinor(), case.spinor()
            i1 = single.solve(op, case.spinor(b), X1)
            i64 = s64.solve(op, case.spinor(b), X64)
            print("point source %d sloppy %d: batch %d it, alone %d it, FP64 %d it" % (
                q, sloppy, infos[q].steps, i1.steps, i64.steps))
            assert infos[q].converged and i1.converged and i64.converged
            x64 = X64.download()
            assert true_residual(case, xs[q], b) <= 1e-10
            assert np.linalg.norm((xs[q] - x64).ravel()) / np.linalg.norm(x64.ravel()) < 1e-9
            assert np.linalg.norm((X1.download() - x64).ravel()) / np.linalg.norm(x64.ravel()) < 1e-9


def test_batch_after_gauge_storage_change():
    """An operator that ran a batch on 18-real links and is then given the same
    field in 12-real storage solves exactly as an operator made on the 12-real
    field (the batch links follow the new field)."""
    if lym.device_count() < 1:
        pytest.fail("no HIP device visible")
    dims = (4, 4, 4, 8)
    full = Case(dims, seed=17)
    comp = Case(dims, seed=17, storage=lym.LINKS_RECONSTRUCT)
    srcs = [full.gauss() for _ in range(3)]
    eps = [(1e-11 * np.linalg.norm(b)) ** 2 for b in srcs]
    op = full.op(lym.OP_CLOVER)
    x18, i18, _ = solve_batch(full, op, lym.SOLVER_PRECOND_MIXED_CG, srcs, 2, eps, max_steps=2000)
    op.set_gauge(comp.gauge)
    xsw, isw, _ = solve_batch(full, op, lym.SOLVER_PRECOND_MIXED_CG, srcs, 2, eps, max_steps=2000)
    xfr, ifr, _ = solve_batch(comp, comp.op(lym.OP_CLOVER), lym.SOLVER_PRECOND_MIXED_CG, srcs, 2, eps,
                              max_steps=2000)
    for q, b in enumerate(srcs):
        assert i18[q].converged and isw[q].converged and ifr[q].converged
        assert isw[q].steps == ifr[q].steps
        assert np.array_equal(xsw[q], xfr[q])
        assert true_residual(full, xsw[q], b) <= 1e-10


def test_batch_over_kappas():
    """Right-hand sides with their own kappa as lanes of one batch
    (lym_solver_solve_batch_kappas): each meets the FP64 stopping rule at its
    kappa (the oracle's D), gives the same bits alone, in the mix and at
    another batch width, and agrees with the per-kappa batched solve."""
    from helpers import OracleField, fold_bc
    dims = (8, 8, 8, 8)
    case = Case(dims=dims)
    U = fold_bc(dims, lym.gen_su3_links(dims, 0.5, 11), -1)
    kap = [0.120, 0.125, 0.130, 0.120, 0.125, 0.130, 0.1275]
    srcs = [case.gauss() for _ in kap]
    eps = [(1e-11 * np.linalg.norm(b)) ** 2 for b in srcs]
    op = case.op(lym.OP_CLOVER)

    def run(idx, block=0):
        solver = lym.Solver(case.ctx, lym.SOLVER_PRECOND_MIXED_CG, max_steps=20000, sloppy=lym.PREC_HALF)
        solver.set_rhs_block(block)
        B = [case.spinor(srcs[q]) for q in idx]
        X = [case.spinor() for _ in idx]
        infos = solver.solve_batch_kappas(op, [kap[q] for q in idx], B, X, [eps[q] for q in idx])
        return [x.download() for x in X], infos

    xs, infos = run(list(range(len(kap))))
    for q, k in enumerate(kap):
        assert infos[q].converged, q
        orc = OracleField(dims, U, kappa=k, csw=1.0, need_clover=True, need_inverse=False)
        Dx = orc.apply(K_CLOVER, xs[q], gamma5=False)
        res = np.linalg.norm((srcs[q] - Dx).ravel()) / np.linalg.norm(srcs[q].ravel())
        print("kappa %.4f: %d iterations, residual %.2e" % (k, infos[q].steps, res))
        assert res <= 1e-10
    # alone and at another width: the same bits
    for q in (0, 5):
        alone, ia = run([q])
        assert ia[0].steps == infos[q].steps and np.array_equal(alone[0], xs[q]), q
    xs4, _ = run(list(range(len(kap))), block=4)
    assert all(np.array_equal(a, b) for a, b in zip(xs4, xs))
    # the per-kappa batched solve reaches the same solutions
    for k in (0.120, 0.130):
        idx = [q for q in range(len(kap)) if kap[q] == k]
        solver = lym.Solver(case.ctx, lym.SOLVER_PRECOND_MIXED_CG, max_steps=20000, sloppy=lym.PREC_HALF)
        X = [case.spinor() for _ in idx]
        inf = solver.solve_batch(case.op(lym.OP_CLOVER, kappa=k), [case.spinor(srcs[q]) for q in idx], X,
                                 [eps[q] for q in idx])
        for j, q in enumerate(idx):
            assert inf[j].converged
            x = X[j].download()
            assert np.linalg.norm((x - xs[q]).ravel()) / np.linalg.norm(x.ravel()) < 1e-9


def test_batch_over_kappas_refuses_what_it_cannot_run():
    case = Case(dims=(4, 4, 4, 4))
    b = [case.spinor(case.gauss())]
    x = [case.spinor()]
    s = lym.Solver(case.ctx, lym.SOLVER_MIXED_CG, sloppy=lym.PREC_HALF)
    with pytest.raises(lym.LymError):
        s.solve_batch_kappas(case.op(lym.OP_CLOVER), [0.12], b, x)
    s = lym.Solver(case.ctx, lym.SOLVER_PRECOND_MIXED_CG, sloppy=lym.PREC_HALF)
    with pytest.raises(lym.LymError):
        s.solve_batch_kappas(case.op(lym.OP_CLOVER, power=2), [0.12], b, x)
    for bad in (float("nan"), float("inf")):
        with pytest.raises(lym.LymError):
            s.solve_batch_kappas(case.op(lym.OP_CLOVER), [bad], b, x)


def test_batch_over_kappas_admits_in_kappa_order():
    """Twelve right-hand sides whose kappas alternate enter the (default,
    12-lane) batch sorted by kappa; every one keeps its own index, converges,
    and gives the same bits as in a batch that receives them already grouped
    by kappa and as in an 8-lane batch."""
    case = Case(dims=(6, 6, 6, 8))
    kset = (0.120, 0.127, 0.133)
    kap = [kset[q % 3] for q in range(12)]
    srcs = [case.gauss() for _ in kap]
    op = case.op(lym.OP_CLOVER)

    def run(idx, block=0):
        solver = lym.Solver(case.ctx, lym.SOLVER_PRECOND_MIXED_CG, max_steps=20000, sloppy=lym.PREC_HALF)
        solver.set_rhs_block(block)
        X = [case.spinor() for _ in idx]
        infos = solver.solve_batch_kappas(op, [kap[q] for q in idx], [case.spinor(srcs[q]) for q in idx], X)
        return [x.download() for x in X], infos

    xs, infos = run(list(range(12)))
    assert all(i.converged for i in infos)
    grouped = sorted(range(12), key=lambda q: kap[q])
    xg, ig = run(grouped)
    for j, q in enumerate(grouped):
        assert ig[j].steps == infos[q].steps and np.array_equal(xg[j], xs[q]), q
    x8, _ = run(list(range(12)), block=8)
    assert all(np.array_equal(a, b) for a, b in zip(x8, xs))
