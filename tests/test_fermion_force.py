"""Two-flavour Wilson fermion force (TwoFlavorFermionAction::force and
updateForce, source/actions/TwoFlavorFermionAction.cpp:17-43, with
DiracWilsonFermionForce::derivative, source/hmc_forces/DiracWilsonFermionForce.cpp:10-54,
and FermionForce::force, source/hmc_forces/FermionForce.cpp:33-58).

The oracle restatement is pinned on the CPU by finite differences of the
action it differentiates: S(U) = |Q^-1 phi|^2 with Q = gamma5 D_Wilson (the
pseudofermion energy, TwoFlavorFermionAction.cpp:46-55).  Moving the gauge
link U_mu(x) to exp(i eps T_a) U_mu(x) changes S at the rate
dS/deps_a = -2 i tr(F T_a), F the force matrix (tr T_a T_b = delta_ab / 2),
for SU(3) fundamental and SU(2) adjoint fermions.  With the time-antiperiodic
boundary the reference differentiates with respect to the fermion link, so
on the links that carry the boundary sign its force is the negative of the
gauge-link derivative: the test states that too.  The device kernel is then
compared with the oracle per link."""
import numpy as np
import pytest
import scipy.linalg

from helpers import K_WILSON, OracleField, fold_bc, oracle_fermion_force

import leonardym_amd as lym

DIMS = (4, 4, 4, 4)
V = int(np.prod(DIMS))
KAPPA = 0.12


def gell_mann_half():
    T = np.zeros((8, 3, 3), complex)
    T[0][0, 1] = T[0][1, 0] = 0.5
    T[1][0, 1], T[1][1, 0] = -0.5j, 0.5j
    T[2][0, 0], T[2][1, 1] = 0.5, -0.5
    T[3][0, 2] = T[3][2, 0] = 0.5
    T[4][0, 2], T[4][2, 0] = -0.5j, 0.5j
    T[5][1, 2] = T[5][2, 1] = 0.5
    T[6][1, 2], T[6][2, 1] = -0.5j, 0.5j
    T[7] = np.diag([1, 1, -2]) / (2 * np.sqrt(3))
    return T


def pauli_half():
    return np.array([[[0, 1], [1, 0]], [[0, -1j], [1j, 0]], [[1, 0], [0, -1]]], complex) / 2


class Action:
    """S(U) = |Q^-1 phi|^2 on the oracle, for gauge links U."""

    def __init__(self, rep, bc, seed):
        self.rep, self.bc = rep, bc
        self.U = lym.gen_su3_links(DIMS, 0.6, seed) if rep == 0 else lym.gen_su2_links(DIMS, 0.6, seed)
        rng = np.random.default_rng(seed)
        self.phi = (rng.standard_normal((V, 4, 3)) + 1j * rng.standard_normal((V, 4, 3))) / np.sqrt(2)

    def fermion_links(self, U):
        W = U if self.rep == 0 else lym.su2_to_adjoint(U).astype(complex)
        return fold_bc(DIMS, W, self.bc)

    def solves(self, U):
        orc = OracleField(DIMS, self.fermion_links(U), kappa=KAPPA)
        Y, _, ok1 = orc.solve("bicgstab", self.phi, 1e-28, kind=K_WILSON)
        X, _, ok2 = orc.solve("bicgstab", Y, 1e-28, kind=K_WILSON)
        assert ok1 and ok2
        return X, Y

    def S(self, U):
        return float(np.sum(np.abs(self.solves(U)[1]) ** 2))


@pytest.mark.parametrize("rep,bc", [(0, 1), (1, 1), (0, -1)], ids=["su3-periodic", "su2adj-periodic", "su3-antiperiodic"])
def test_oracle_force_is_the_action_derivative(rep, bc):
    act = Action(rep, bc, 71 + rep)
    X, Y = act.solves(act.U)
    F = oracle_fermion_force(DIMS, rep, KAPPA, act.fermion_links(act.U), act.U, X, Y)
    T = gell_mann_half() if rep == 0 else pauli_half()
    N = T.shape[-1]
    # traceless anti-hermitian
    assert np.max(np.abs(F + np.conj(np.swapaxes(F, -1, -2)))) < 1e-12
    assert np.max(np.abs(np.trace(F, axis1=-2, axis2=-1))) < 1e-12
    t_last = DIMS[3] - 1
    links = [(0, 0, 1), (37, 2, 0), (V - 1, 3, len(T) - 1), (122, 1, 2)]
    if bc == -1:
        # a time link on the last timeslice carries the boundary sign
        links.append(((3 * DIMS[1] * DIMS[2] + 2) * DIMS[3] + t_last, 3, 4))
    def central(s, mu, a, h):
        up, dn = act.U.copy(), act.U.copy()
        up[s, mu] = scipy.linalg.expm(1j * h * T[a]) @ act.U[s, mu]
        dn[s, mu] = scipy.linalg.expm(-1j * h * T[a]) @ act.U[s, mu]
        return (act.S(up) - act.S(dn)) / (2 * h)

    h = 2e-3
    for s, mu, a in links:
        # Richardson extrapolation of central differences: O(h^4), with the
        # solves' rounding noise divided by h instead of by a tiny step
        fd = (4 * central(s, mu, a, h / 2) - central(s, mu, a, h)) / 3
        got = (-2j * np.trace(F[s, mu] @ T[a])).real
        sign = -1.0 if (bc == -1 and mu == 3 and s % DIMS[3] == t_last) else 1.0
        assert got == pytest.approx(sign * fd, rel=2e-7, abs=1e-9), (s, mu, a)
    assert N == (3 if rep == 0 else 2)


@pytest.mark.gpu
@pytest.mark.parametrize("rep", [0, 1], ids=["su3", "su2adj"])
@pytest.mark.parametrize("storage", [lym.LINKS_FULL, lym.LINKS_RECONSTRUCT], ids=["links18", "links12"])
def test_device_force_matches_oracle(rep, storage):
    print("fermion force: rep %d storage %d on %s" % (rep, storage, DIMS), flush=True)
    act = Action(rep, -1, 91 + rep)
    W = act.fermion_links(act.U)
    ctx = lym.Context(DIMS, lym.REP_SU3_FUND if rep == 0 else lym.REP_SU2_ADJ, -1)
    gauge = lym.Gauge(ctx, W.real if rep == 1 else W, folded=True, storage=storage)
    op = lym.Operator(ctx, lym.OP_WILSON, 1, KAPPA, 0.0, True, gauge=gauge)
    X, Y = act.solves(act.U)
    got = op.fermion_force(lym.Spinor(ctx, X), lym.Spinor(ctx, Y))
    want = oracle_fermion_force(DIMS, rep, KAPPA, W, act.U, X, Y)
    assert np.max(np.abs(got - want)) < 1e-13 * max(1.0, np.max(np.abs(want)))


@pytest.mark.gpu
def test_device_two_flavor_force_end_to_end():
    # updateForce: the two solves on the device, then the force, against the oracle's
    print("two-flavour force: solves + force on 6x4x4x8", flush=True)
    dims = (6, 4, 4, 8)
    v = int(np.prod(dims))
    U = lym.gen_su3_links(dims, 0.5, 93)
    W = fold_bc(dims, U, -1)
    phi = lym.gaussian_spinor(v, 94)
    ctx = lym.Context(dims, lym.REP_SU3_FUND, -1)
    gauge = lym.Gauge(ctx, W, folded=True)
    op = lym.Operator(ctx, lym.OP_WILSON, 1, KAPPA, 0.0, True, gauge=gauge)
    solver = lym.Solver(ctx, lym.SOLVER_BICGSTAB, precision=1e-26, max_steps=2000)
    got, infos = solver.two_flavor_force(op, lym.Spinor(ctx, phi))
    assert infos[0].converged and infos[1].converged
    orc = OracleField(dims, W, kappa=KAPPA)
    Y, _, _ = orc.solve("bicgstab", phi, 1e-26, kind=K_WILSON)
    X, _, _ = orc.solve("bicgstab", Y, 1e-26, kind=K_WILSON)
    want = oracle_fermion_force(dims, 0, KAPPA, W, U, X, Y)
    assert np.max(np.abs(got - want)) < 1e-9 * np.max(np.abs(want))


@pytest.mark.gpu
@pytest.mark.parametrize("dims", [(3, 4, 4, 6), (4, 4, 4, 5)], ids=lambda d: "x".join(map(str, d)))
def test_device_two_flavor_force_odd_extent(dims):
    """A lattice with an odd extent (one site class on the device): the two
    FP64 solves and the force against the oracle's."""
    print("two-flavour force: solves + force on", dims, flush=True)
    v = int(np.prod(dims))
    U = lym.gen_su3_links(dims, 0.5, 95)
    W = fold_bc(dims, U, -1)
    phi = lym.gaussian_spinor(v, 96)
    ctx = lym.Context(dims, lym.REP_SU3_FUND, -1)
    gauge = lym.Gauge(ctx, W, folded=True)
    op = lym.Operator(ctx, lym.OP_WILSON, 1, KAPPA, 0.0, True, gauge=gauge)
    solver = lym.Solver(ctx, lym.SOLVER_BICGSTAB, precision=1e-26, max_steps=2000)
    got, infos = solver.two_flavor_force(op, lym.Spinor(ctx, phi))
    assert infos[0].converged and infos[1].converged
    orc = OracleField(dims, W, kappa=KAPPA)
    Y, _, _ = orc.solve("bicgstab", phi, 1e-26, kind=K_WILSON)
    X, _, _ = orc.solve("bicgstab", Y, 1e-26, kind=K_WILSON)
    want = oracle_fermion_force(dims, 0, KAPPA, W, U, X, Y)
    assert np.max(np.abs(got - want)) < 1e-9 * np.max(np.abs(want))
