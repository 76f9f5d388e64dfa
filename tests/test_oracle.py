"""Pins the CPU oracle (oracle/lym_oracle.cpp) against closed forms,
invariants and brute force, chosen so that a plausible transcription mistake
(a wrong gamma sign, U in place of U^dagger, a swapped plane index, a wrong
link site, a dropped term) fails at least one test.

The reference ships no golden vectors or assertions; its printed checks
(source/tests/TestLinearAlgebra.cpp:36-150, TestCommunication.cpp:168-270) are
restated here as assertions.
"""
import ctypes

import numpy as np
import pytest

from helpers import (GAMMA, GAMMA5, K_BASIC, K_CLOVER, K_EO, K_TEXTBOOK, K_WILSON, OracleField, coords,
                     dense_matrix, fold_bc, gaussian_spinor, oracle, random_links, rel_err_per_site, vdot, _p)

DIMS = (4, 4, 4, 8)


def plane_wave(dims, p, u):
    x = coords(dims)
    ph = np.exp(1j * (x @ np.asarray(p)))
    return ph[:, None, None] * u[None, :, :]


def free_D(p, kappa):
    """Closed form D(p) = (1 - 2 kappa sum cos p) + 2 i kappa sum gamma_mu sin p_mu."""
    M = (1 - 2 * kappa * np.sum(np.cos(p))) * np.eye(4, dtype=complex)
    for mu in range(4):
        M = M + 2j * kappa * np.sin(p[mu]) * GAMMA[mu]
    return M


def momenta(dims, n):
    p = [2 * np.pi * n[mu] / dims[mu] for mu in range(3)]
    p.append((2 * n[3] + 1) * np.pi / dims[3])  # antiperiodic in time
    return np.array(p)


def test_neighbour_tables():
    """sup/sdn give coordinate +-1 mod L (TestCommunication.cpp:168-270)."""
    dims = (4, 6, 2, 8)
    V = int(np.prod(dims))
    sup = np.zeros((V, 4), np.int32)
    sdn = np.zeros((V, 4), np.int32)
    par = np.zeros(V, np.int32)
    oracle().orc_tables(_p(np.array(dims, np.int32)), _p(sup), _p(sdn), _p(par))
    c = coords(dims)
    for mu in range(4):
        up = c.copy()
        up[:, mu] = (up[:, mu] + 1) % dims[mu]
        dn = c.copy()
        dn[:, mu] = (dn[:, mu] - 1) % dims[mu]
        assert np.array_equal(c[sup[:, mu]], up)
        assert np.array_equal(c[sdn[:, mu]], dn)
    assert np.array_equal(par, c.sum(1) % 2)


def test_bc_fold():
    rng = np.random.default_rng(1)
    U = random_links(DIMS, rng)
    Ucm = np.ascontiguousarray(np.swapaxes(U, -1, -2))
    oracle().orc_apply_bc(_p(np.array(DIMS, np.int32)), _p(Ucm))
    Ub = np.swapaxes(Ucm, -1, -2)
    t = coords(DIMS)[:, 3]
    ref = U.copy()
    ref[t == DIMS[3] - 1, 3] *= -1
    assert np.array_equal(Ub, ref)
    assert np.array_equal(fold_bc(DIMS, U), ref)


@pytest.mark.parametrize("kind", [K_WILSON, K_BASIC, K_TEXTBOOK])
@pytest.mark.parametrize("gamma5", [True, False])
def test_free_field_plane_wave(kind, gamma5):
    """U = 1 with antiperiodic time: D e^{ipx}u = D(p) u e^{ipx}."""
    rng = np.random.default_rng(2)
    V = int(np.prod(DIMS))
    U = np.broadcast_to(np.eye(3, dtype=complex), (V, 4, 3, 3)).copy()
    fld = OracleField(DIMS, fold_bc(DIMS, U), kappa=0.13)
    for n in [(0, 0, 0, 0), (1, 2, 3, 1), (3, 1, 0, 6)]:
        p = momenta(DIMS, n)
        u = rng.standard_normal((4, 3)) + 1j * rng.standard_normal((4, 3))
        psi = plane_wave(DIMS, p, u)
        M = free_D(p, 0.13)
        if gamma5:
            M = GAMMA5 @ M
        expect = plane_wave(DIMS, p, M @ u)
        got = fld.apply(kind, psi, gamma5=gamma5)
        assert rel_err_per_site(got, expect) < 1e-13


@pytest.mark.parametrize("kind", [K_WILSON, K_BASIC, K_TEXTBOOK])
def test_constant_abelian_field(kind):
    """U_mu(x) = e^{i theta_mu} shifts the momentum: D[U](p) = D_free(p + theta).
    Distinguishes U from U^dagger on the backward hop."""
    rng = np.random.default_rng(3)
    V = int(np.prod(DIMS))
    theta = np.array([0.3, -0.7, 1.1, 0.45])
    U = np.zeros((V, 4, 3, 3), complex)
    for mu in range(4):
        U[:, mu] = np.exp(1j * theta[mu]) * np.eye(3)
    fld = OracleField(DIMS, fold_bc(DIMS, U), kappa=0.11)
    p = momenta(DIMS, (1, 0, 2, 3))
    u = rng.standard_normal((4, 3)) + 1j * rng.standard_normal((4, 3))
    got = fld.apply(kind, plane_wave(DIMS, p, u), gamma5=False)
    expect = plane_wave(DIMS, p, free_D(p + theta, 0.11) @ u)
    assert rel_err_per_site(got, expect) < 1e-13


def gauge_transform(dims, U, g):
    """U^g_mu(x) = g(x) U_mu(x) g(x+mu)^dagger."""
    V = int(np.prod(dims))
    c = coords(dims)
    Ug = np.empty_like(U)
    strides = np.array([dims[1] * dims[2] * dims[3], dims[2] * dims[3], dims[3], 1])
    for mu in range(4):
        cu = c.copy()
        cu[:, mu] = (cu[:, mu] + 1) % dims[mu]
        up = cu @ strides
        Ug[:, mu] = g @ U[:, mu] @ np.conj(np.swapaxes(g[up], 1, 2))
    return Ug


@pytest.mark.parametrize("kind,csw", [(K_WILSON, 0.0), (K_BASIC, 0.0), (K_TEXTBOOK, 1.0), (K_CLOVER, 1.0)])
@pytest.mark.parametrize("group", ["su3", "adj"])
def test_gauge_covariance(kind, csw, group):
    rng = np.random.default_rng(4)
    V = int(np.prod(DIMS))
    U = fold_bc(DIMS, random_links(DIMS, rng, group))
    g = random_links(DIMS, rng, group)[:, 0]
    psi = gaussian_spinor(V, rng)
    f1 = OracleField(DIMS, U, kappa=0.12, csw=csw, need_clover=csw != 0)
    f2 = OracleField(DIMS, gauge_transform(DIMS, U, g), kappa=0.12, csw=csw, need_clover=csw != 0)
    a = f2.apply(kind, np.einsum("sij,saj->sai", g, psi))
    b = np.einsum("sij,saj->sai", g, f1.apply(kind, psi))
    assert rel_err_per_site(a, b) < 1e-12


@pytest.mark.parametrize("kind,csw", [(K_WILSON, 0.0), (K_CLOVER, 1.0), (K_TEXTBOOK, 1.0)])
def test_gamma5_hermiticity(kind, csw):
    """<b, Q a> = <Q b, a> (TestLinearAlgebra.cpp:36-57)."""
    rng = np.random.default_rng(5)
    V = int(np.prod(DIMS))
    fld = OracleField(DIMS, fold_bc(DIMS, random_links(DIMS, rng)), kappa=0.1, csw=csw, need_clover=csw != 0)
    a, b = gaussian_spinor(V, rng), gaussian_spinor(V, rng)
    lhs = vdot(b, fld.apply(kind, a))
    rhs = vdot(fld.apply(kind, b), a)
    assert abs(lhs - rhs) < 1e-12 * abs(lhs)
    # and D itself is not hermitian: the test can fail
    lhs2 = vdot(b, fld.apply(kind, a, gamma5=False))
    rhs2 = vdot(fld.apply(kind, b, gamma5=False), a)
    assert abs(lhs2 - rhs2) > 1e-3 * abs(lhs2)


@pytest.mark.parametrize("gamma5", [True, False])
@pytest.mark.parametrize("group", ["su3", "adj"])
def test_three_wilson_forms_agree(gamma5, group):
    """Unrolled == projector (TestLinearAlgebra.cpp:145-150) == textbook."""
    rng = np.random.default_rng(6)
    V = int(np.prod(DIMS))
    fld = OracleField(DIMS, fold_bc(DIMS, random_links(DIMS, rng, group)), kappa=0.13)
    psi = gaussian_spinor(V, rng)
    a = fld.apply(K_WILSON, psi, gamma5=gamma5)
    b = fld.apply(K_BASIC, psi, gamma5=gamma5)
    c = fld.apply(K_TEXTBOOK, psi, gamma5=gamma5)
    assert rel_err_per_site(a, c) < 1e-13
    assert rel_err_per_site(b, c) < 1e-13


@pytest.mark.parametrize("gamma5", [True, False])
@pytest.mark.parametrize("group", ["su3", "adj"])
def test_clover_block_equals_sigma_form(gamma5, group):
    """The reference's chiral-block clover combination equals the textbook
    Sheikholeslami-Wohlert term -kappa csw sum sigma_{mu nu} Fhat_{mu nu}."""
    rng = np.random.default_rng(7)
    V = int(np.prod(DIMS))
    fld = OracleField(DIMS, fold_bc(DIMS, random_links(DIMS, rng, group, eps=None)), kappa=0.13, csw=1.3,
                      need_clover=True)
    psi = gaussian_spinor(V, rng)
    a = fld.apply(K_CLOVER, psi, gamma5=gamma5)
    c = fld.apply(K_TEXTBOOK, psi, gamma5=gamma5)
    w = fld.apply(K_WILSON, psi, gamma5=gamma5)
    assert rel_err_per_site(a, c) < 1e-13
    assert rel_err_per_site(a, w) > 1e-3  # the clover term is really there


def test_csw_zero_is_wilson():
    """TestLinearAlgebra.cpp:122-144."""
    rng = np.random.default_rng(8)
    V = int(np.prod(DIMS))
    fld = OracleField(DIMS, fold_bc(DIMS, random_links(DIMS, rng)), kappa=0.12, csw=0.0, need_clover=True)
    psi = gaussian_spinor(V, rng)
    w = fld.apply(K_WILSON, psi)
    c = fld.apply(K_CLOVER, psi)
    assert np.array_equal(w, c) or rel_err_per_site(w, c) < 1e-15
    c0 = fld.apply(K_CLOVER, psi, v2=psi, alpha=0.0)
    assert rel_err_per_site(c0, c) < 1e-15


@pytest.mark.parametrize("kind", [K_WILSON, K_BASIC, K_CLOVER])
@pytest.mark.parametrize("gamma5", [True, False])
def test_multiply_add(kind, gamma5):
    """multiplyAdd(out, v1, v2, a) = Op v1 + a v2, including the gamma5 = false
    fix-up (DiracWilsonOperator.cpp:383-388)."""
    rng = np.random.default_rng(9)
    V = int(np.prod(DIMS))
    fld = OracleField(DIMS, fold_bc(DIMS, random_links(DIMS, rng)), kappa=0.12, csw=1.0, need_clover=True)
    v1, v2 = gaussian_spinor(V, rng), gaussian_spinor(V, rng)
    alpha = 0.3 - 1.7j
    got = fld.apply(kind, v1, gamma5=gamma5, v2=v2, alpha=alpha)
    expect = fld.apply(kind, v1, gamma5=gamma5) + alpha * v2
    assert rel_err_per_site(got, expect) < 1e-14


@pytest.mark.parametrize("mu,nu", [(0, 1), (0, 2), (0, 3), (1, 2), (1, 3), (2, 3)])
def test_field_strength_abelian(mu, nu):
    """U_mu(x) = diag(e^{i B q_c x_nu}), B = 2 pi / L_nu: every leaf of plane
    (mu, nu) is e^{-i B q}, so F_{mu nu} = -i diag(sin B q), other planes 0.
    Plane order 01,02,03,12,13,23 (ImprovedDiracWilsonOperator.cpp:508-513)."""
    dims = (4, 6, 4, 8)
    V = int(np.prod(dims))
    c = coords(dims)
    q = np.array([1, 2, -3])
    B = 2 * np.pi / dims[nu]
    U = np.broadcast_to(np.eye(3, dtype=complex), (V, 4, 3, 3)).copy()
    ph = np.exp(1j * B * np.outer(c[:, nu], q))
    U[:, mu] = ph[:, :, None] * np.eye(3)
    fld = OracleField(dims, fold_bc(dims, U), kappa=0.1, csw=1.0, need_clover=True)
    planes = [(0, 1), (0, 2), (0, 3), (1, 2), (1, 3), (2, 3)]
    k = planes.index((mu, nu))
    expect = -1j * np.diag(np.sin(B * q))
    for p in range(6):
        want = expect if p == k else np.zeros((3, 3))
        assert np.max(np.abs(fld.F[:, p] - want)) < 1e-14
    # closed form of the clover term in D: -kappa csw sigma Fhat with Fhat = -sin(Bq)
    rng = np.random.default_rng(10)
    psi = gaussian_spinor(V, rng)
    dq = fld.apply(K_CLOVER, psi, gamma5=False) - fld.apply(K_WILSON, psi, gamma5=False)
    sig = 0.5j * (GAMMA[mu] @ GAMMA[nu] - GAMMA[nu] @ GAMMA[mu])
    want = 0.1 * 1.0 * np.einsum("ab,c,sbc->sac", sig, np.sin(B * q), psi)
    assert rel_err_per_site(dq, want) < 1e-13


def test_field_strength_properties():
    rng = np.random.default_rng(11)
    V = int(np.prod(DIMS))
    U = random_links(DIMS, rng)
    f_plain = OracleField(DIMS, U, need_clover=True)
    f_bc = OracleField(DIMS, fold_bc(DIMS, U), need_clover=True)
    F = f_plain.F
    assert np.max(np.abs(F + np.conj(np.swapaxes(F, -1, -2)))) < 1e-15  # anti-hermitian
    assert np.max(np.abs(F - f_bc.F)) < 1e-14  # antiperiodic BC does not change F
    cold = OracleField(DIMS, fold_bc(DIMS, np.broadcast_to(np.eye(3, dtype=complex), (V, 4, 3, 3)).copy()),
                       need_clover=True)
    assert np.max(np.abs(cold.F)) == 0


def test_clover_inverse():
    rng = np.random.default_rng(12)
    fld = OracleField(DIMS, fold_bc(DIMS, random_links(DIMS, rng, eps=0.5)), kappa=0.13, csw=1.0, need_clover=True)
    R = np.zeros((12, 12), complex)
    for s in [0, 7, 100, 511]:
        oracle().orc_clover_site_matrix(_p(np.ascontiguousarray(fld.F[s])), 0.13, 1.0, _p(R))
        Rm = R.T  # column-major storage
        inv = fld.inv[s].T
        assert np.max(np.abs(Rm @ inv - np.eye(12))) < 1e-14
        # chiral block diagonal, each 6x6 block hermitian
        assert np.max(np.abs(Rm[:6, 6:])) == 0 and np.max(np.abs(Rm[6:, :6])) == 0
        assert np.max(np.abs(Rm - np.conj(Rm.T))) < 1e-15


def test_eo_schur_identity():
    """With psi_e = -R_ee^-1 D_eo psi_o, (D psi)_e = 0 and (D psi)_o = Shat psi_o,
    which the even-odd operator must reproduce (EvenOdd...cpp:26-47)."""
    rng = np.random.default_rng(13)
    V = int(np.prod(DIMS))
    par = coords(DIMS).sum(1) % 2
    fld = OracleField(DIMS, fold_bc(DIMS, random_links(DIMS, rng, eps=0.5)), kappa=0.13, csw=1.0, need_clover=True)
    psi = gaussian_spinor(V, rng)
    psi[par == 0] = 0
    d = fld.desc(K_EO, gamma5=False)
    tmp = np.zeros_like(psi)
    oracle().orc_eo_hop(ctypes.byref(d), _p(tmp), _p(psi), 0)
    tmp[par == 1] = 0
    oracle().orc_eo_inv_even(ctypes.byref(d), _p(tmp))
    full = psi - tmp
    Dfull = fld.apply(K_CLOVER, full, gamma5=False)
    assert np.max(np.abs(Dfull[par == 0])) < 1e-13
    eo = fld.apply(K_EO, psi, gamma5=False)
    assert rel_err_per_site(eo[par == 1], Dfull[par == 1]) < 1e-13
    assert np.array_equal(eo[par == 0], psi[par == 0])
    eo5 = fld.apply(K_EO, psi, gamma5=True)
    assert rel_err_per_site(eo5[par == 1], np.einsum("ab,sbc->sac", GAMMA5, Dfull[par == 1])) < 1e-13


def test_blas_exact():
    rng = np.random.default_rng(14)
    n = 37
    a = (rng.integers(-5, 6, (n, 4, 3)) + 1j * rng.integers(-5, 6, (n, 4, 3))).astype(complex)
    b = (rng.integers(-5, 6, (n, 4, 3)) + 1j * rng.integers(-5, 6, (n, 4, 3))).astype(complex)
    lib = oracle()
    assert lib.orc_squared_norm(n, _p(a)) == np.sum(np.abs(a) ** 2)
    re, im = ctypes.c_double(), ctypes.c_double()
    lib.orc_dot(n, _p(a), _p(b), ctypes.byref(re), ctypes.byref(im))
    ref = np.vdot(a.ravel(), b.ravel())  # conj-linear in the first argument
    assert re.value == ref.real and im.value == ref.imag
    lib.orc_gamma5dot(n, _p(a), _p(b), ctypes.byref(re), ctypes.byref(im))
    ref5 = np.vdot(a.ravel(), np.einsum("ab,sbc->sac", GAMMA5, b).ravel())
    assert re.value == ref5.real and im.value == ref5.imag
    assert lib.orc_difference_norm(n, _p(a), _p(b)) == np.sum(np.abs(a - b) ** 2)


@pytest.fixture(scope="module")
def dense_case():
    """4^4 clover lattice, its dense D (gamma5 = false) and a Gaussian source."""
    dims = (4, 4, 4, 4)
    V = int(np.prod(dims))
    rng = np.random.default_rng(15)
    fld = OracleField(dims, fold_bc(dims, random_links(dims, rng, eps=0.5)), kappa=0.12, csw=1.0,
                      need_clover=True)
    D = dense_matrix(lambda e: fld.apply(K_CLOVER, e, gamma5=False), V)
    b = gaussian_spinor(V, rng)
    x = np.linalg.solve(D, b.ravel()).reshape(V, 4, 3)
    return fld, D, b, x


def _relerr(a, b):
    return np.linalg.norm((a - b).ravel()) / np.linalg.norm(b.ravel())


def test_dense_operator_is_gamma5_hermitian(dense_case):
    fld, D, b, x = dense_case
    G5 = np.kron(np.eye(D.shape[0] // 12), np.kron(GAMMA5, np.eye(3)))
    Q = G5 @ D
    assert np.max(np.abs(Q - Q.conj().T)) < 1e-14


def test_cg_solution_matches_dense(dense_case):
    """CG on Q^2 = D^dag D with rhs Q(gamma5 b) (ConjugateGradient.cpp:10-74)."""
    fld, D, b, x = dense_case
    rhs = fld.apply(K_CLOVER, np.einsum("ab,sbc->sac", GAMMA5, b))
    eps = (1e-12 * np.linalg.norm(rhs)) ** 2
    sol, info, ok = fld.solve("cg", rhs, eps, kind=K_CLOVER, power=2)
    assert ok and info.converged
    assert _relerr(sol, x) < 1e-9
    assert np.linalg.norm(D @ sol.ravel() - b.ravel()) / np.linalg.norm(b) < 1e-10


@pytest.mark.parametrize("gamma5", [False, True])
def test_bicgstab_solution_matches_dense(dense_case, gamma5):
    """Plain BiCGStab on Op (BiConjugateGradient.cpp:321-507)."""
    fld, D, b, x = dense_case
    rhs = b if not gamma5 else np.einsum("ab,sbc->sac", GAMMA5, b)
    eps = (1e-12 * np.linalg.norm(rhs)) ** 2
    sol, info, ok = fld.solve("bicgstab", rhs, eps, kind=K_CLOVER, power=1, gamma5=gamma5)
    assert ok
    assert _relerr(sol, x) < 1e-9


def test_precond_bicgstab_solution_matches_dense(dense_case):
    """Even-odd driver returns D^-1 b (PreconditionedBiCGStab.cpp:43-91)."""
    fld, D, b, x = dense_case
    eps = (1e-12 * np.linalg.norm(b)) ** 2
    sol, info, ok = fld.solve("precond_bicgstab", b, eps, kind=K_CLOVER)
    assert ok
    assert _relerr(sol, x) < 1e-9


@pytest.mark.parametrize("gamma5", [False, True])
def test_precond_bicgstab_falls_back_on_odd_extents(gamma5):
    """A lattice with an odd extent has no even-odd split: the driver runs plain
    BiCGStab on the operator as configured (PreconditionedBiCGStab.cpp:43), so
    it returns Op^-1 b -- checked against the dense operator of a 3x2x3x3
    lattice (the even-odd branch there would not solve D x = b)."""
    dims = (3, 2, 3, 3)
    V = int(np.prod(dims))
    rng = np.random.default_rng(16)
    fld = OracleField(dims, fold_bc(dims, random_links(dims, rng, eps=0.5)), kappa=0.12, csw=1.0,
                      need_clover=True)
    Op = dense_matrix(lambda e: fld.apply(K_CLOVER, e, gamma5=gamma5), V)
    b = gaussian_spinor(V, rng)
    x = np.linalg.solve(Op, b.ravel()).reshape(V, 4, 3)
    sol, info, ok = fld.solve("precond_bicgstab", b, (1e-12 * np.linalg.norm(b)) ** 2, kind=K_CLOVER,
                              gamma5=gamma5)
    assert ok
    assert _relerr(sol, x) < 1e-9


def test_cg_max_steps_failure(dense_case):
    fld, D, b, x = dense_case
    sol, info, ok = fld.solve("cg", b, 1e-30, kind=K_CLOVER, power=2, max_steps=3)
    assert not ok and info.steps == 3


def _np_cg(A, b, x0, eps, max_steps):
    """Textbook CG (Hestenes-Stiefel) on a dense HPD matrix, x0 given."""
    x = x0.copy()
    r = b - A @ x
    p = r.copy()
    nn = np.vdot(r, r).real
    for step in range(max_steps):
        Ap = A @ p
        alpha = nn / np.vdot(p, Ap)
        x += alpha * p
        r -= alpha * Ap
        nn2 = np.vdot(r, r).real
        if nn2 < eps:
            return x, step
        p = r + (nn2 / nn) * p
        nn = nn2
    return x, max_steps


def _np_bicgstab(A, b, x0, eps, max_steps):
    """van der Vorst BiCGStab with the reference's shadow residual b + A x0."""
    x = x0.copy()
    Ax = A @ x
    r = b - Ax
    rh = b + Ax
    p = np.zeros_like(b)
    nu = np.zeros_like(b)
    rho, alpha, omega = 1.0, 1.0, 1.0
    for step in range(max_steps):
        rho_n = np.vdot(rh, r)
        beta = (rho_n / rho) * (alpha / omega)
        p = r + beta * (p - omega * nu)
        nu = A @ p
        alpha = rho_n / np.vdot(rh, nu)
        s = r - alpha * nu
        t = A @ s
        omega = np.vdot(t, s) / np.vdot(t, t)
        x = x + alpha * p + omega * s
        r = s - omega * t
        if np.vdot(r, r).real < eps:
            return x, step
        rho = rho_n
    return x, max_steps


@pytest.mark.parametrize("shift", [0.3 - 0.2j, -0.05])
def test_shifted_bicgstab_matches_dense(dense_case, shift):
    """Shifted BiCGStab (BiConjugateGradient.cpp:510-690) solves (D + shift) x = b;
    its iterates after fixed steps are the textbook recurrences on D + shift,
    and shift 0 is the plain overload (to the run-to-run rounding of the
    oracle's threaded sums)."""
    fld, D, b, x = dense_case
    A = D + shift * np.eye(D.shape[0])
    xs = np.linalg.solve(A, b.ravel())
    eps = (1e-12 * np.linalg.norm(b)) ** 2
    sol, info, ok = fld.solve("bicgstab", b, eps, kind=K_CLOVER, power=1, gamma5=False, shift=shift)
    assert ok and _relerr(sol.ravel(), xs) < 1e-9
    sol, info, ok = fld.solve("bicgstab", b, 1e-300, kind=K_CLOVER, power=1, gamma5=False, max_steps=3, shift=shift)
    xn, _ = _np_bicgstab(A, b.ravel(), b.ravel().copy(), 1e-300, 3)
    assert not ok and _relerr(sol.ravel(), xn) < 1e-12
    s0, i0, ok0 = fld.solve("bicgstab", b, eps, kind=K_CLOVER, power=1, gamma5=False, shift=0.0)
    s1, i1, ok1 = fld.solve("bicgstab", b, eps, kind=K_CLOVER, power=1, gamma5=False)
    assert _relerr(s0, s1) < 1e-13 and abs(int(i0.steps) - int(i1.steps)) <= 1


def test_cg_iteration_path_matches_textbook(dense_case):
    """Pins the CG recurrences themselves (alpha, beta, x0 = b), which the
    converged solution alone does not see."""
    fld, D, b, x = dense_case
    G5 = np.kron(np.eye(D.shape[0] // 12), np.kron(GAMMA5, np.eye(3)))
    Q = G5 @ D
    rhs = fld.apply(K_CLOVER, np.einsum("ab,sbc->sac", GAMMA5, b))
    eps = (1e-8 * np.linalg.norm(rhs)) ** 2
    sol, info, ok = fld.solve("cg", rhs, eps, kind=K_CLOVER, power=2)
    xn, steps = _np_cg(Q @ Q, rhs.ravel(), rhs.ravel().copy(), eps, 5000)
    assert ok and abs(int(info.steps) - steps) <= 1
    assert _relerr(sol.ravel(), xn) < 1e-6


def test_bicgstab_iteration_path_matches_textbook(dense_case):
    fld, D, b, x = dense_case
    eps = (1e-8 * np.linalg.norm(b)) ** 2
    sol, info, ok = fld.solve("bicgstab", b, eps, kind=K_CLOVER, power=1, gamma5=False)
    xn, steps = _np_bicgstab(D, b.ravel(), b.ravel().copy(), eps, 5000)
    assert ok and abs(int(info.steps) - steps) <= 1
    assert _relerr(sol.ravel(), xn) < 1e-6


@pytest.mark.parametrize("nsteps", [1, 2, 6])
def test_solver_iterates_after_fixed_steps(dense_case, nsteps):
    """After a fixed number of unconverged iterations the oracle's iterate equals
    the textbook recurrences' iterate (pins alpha, beta, omega, rhat = b + A x0)."""
    fld, D, b, x = dense_case
    G5 = np.kron(np.eye(D.shape[0] // 12), np.kron(GAMMA5, np.eye(3)))
    Q = G5 @ D
    sol, info, ok = fld.solve("bicgstab", b, 1e-300, kind=K_CLOVER, power=1, gamma5=False, max_steps=nsteps)
    xn, _ = _np_bicgstab(D, b.ravel(), b.ravel().copy(), 1e-300, nsteps)
    assert not ok and _relerr(sol.ravel(), xn) < 1e-12
    sol, info, ok = fld.solve("cg", b, 1e-300, kind=K_CLOVER, power=2, max_steps=nsteps)
    xn, _ = _np_cg(Q @ Q, b.ravel(), b.ravel().copy(), 1e-300, nsteps)
    assert not ok and _relerr(sol.ravel(), xn) < 1e-12


# ---------------------------------------------------------------- meson correlators
# MesonCorrelator::execute (source/correlators/MesonCorrelator.cpp:22-199),
# restated in the oracle with the reference's index loops.  Pinned here by the
# free-field propagator in closed form with the contractions written as spin
# traces, by gauge invariance, and by the reference's own time shift.
G53 = GAMMA5 @ GAMMA[3]


def free_propagator(dims, kappa):
    """S(x) = D^-1 delta_{x,0} for U = 1, time-antiperiodic: (1/V) sum_p e^{ipx}
    D(p)^-1 with p_t = (2n+1) pi / T, as [Lx, Ly, Lz, Lt, sink spin, source spin]."""
    n = [np.arange(L) for L in dims]
    p = np.meshgrid(*[2 * np.pi * n[m] / dims[m] for m in range(3)], (2 * n[3] + 1) * np.pi / dims[3],
                    indexing="ij")
    D = (1 - 2 * kappa * sum(np.cos(q) for q in p))[..., None, None] * np.eye(4)
    for mu in range(4):
        D = D + 2j * kappa * np.sin(p[mu])[..., None, None] * GAMMA[mu]
    Dinv = np.linalg.inv(D)
    S = np.fft.ifftn(Dinv, axes=(0, 1, 2, 3))  # (1/V) sum_n e^{2 pi i n x / L}
    t = np.arange(dims[3])
    return S * np.exp(1j * np.pi * t / dims[3])[None, None, None, :, None, None]


def trace_correlators(S, kappa):
    """Per timeslice, 4 kappa^2 x: pion = tr S^+ S, scalar = Re tr(G^T S^+ G S),
    ps = Re[tr(g S S^+) - tr(g^T S^+ S)] with G = gamma5 gamma_3 (x) 1,
    g = gamma_3 (x) 1, for the propagator as matrices S[..., (sink spin, colour),
    (source spin, colour)] over [Lx, Ly, Lz, Lt]; and the PCAC ratio."""
    G = np.kron(G53, np.eye(3))
    g = np.kron(GAMMA[3], np.eye(3))
    Sd = np.conj(np.swapaxes(S, -1, -2))
    pion = np.einsum("...ab,...ba->...", Sd, S).real
    scalar = np.einsum("ab,...ab->...", G, Sd @ G @ S).real  # tr(A^T B) = sum_ab A_ab B_ab
    ps = (np.einsum("ab,...ba->...", g, S @ Sd) - np.einsum("ab,...ab->...", g, Sd @ S)).real
    f = 4 * kappa ** 2
    raw = {k: v.sum(axis=(0, 1, 2)) for k, v in (("pion", pion), ("scalar", scalar), ("ps", ps))}
    out = {k: f * v for k, v in raw.items()}
    out["pcac"] = (np.roll(raw["ps"], 1) - np.roll(raw["ps"], -1)) / (4 * raw["pion"])
    return out


@pytest.mark.parametrize("kind", [K_WILSON, K_CLOVER])
def test_meson_correlators_free_field(kind):
    """U = 1: the oracle's 12 solves and contractions equal the closed-form
    propagator's spin traces (BiCGStab on D for Wilson, the even-odd driver
    for clover, whose term vanishes at U = 1)."""
    dims = (4, 4, 4, 8)
    kappa = 0.11
    V = int(np.prod(dims))
    U = np.broadcast_to(np.eye(3, dtype=complex), (V, 4, 3, 3)).copy()
    fld = OracleField(dims, fold_bc(dims, U), kappa=kappa, csw=1.0, need_clover=kind == K_CLOVER)
    got = fld.meson_correlators((0, 0, 0, 0), 1e-26, kind=kind)
    assert got["ok"]
    want = trace_correlators(np.kron(free_propagator(dims, kappa), np.eye(3)), kappa)
    for k in ("pion", "scalar"):
        scale = np.max(np.abs(want[k]))
        assert np.max(np.abs(got[k] - want[k])) <= 1e-10 * scale, (k, got[k], want[k])
    assert np.all(got["pion"] > 0)
    # for S = a + b.gamma the PS combination tr(gamma_3 [S, S^+]) vanishes identically
    assert np.max(np.abs(got["ps"])) <= 1e-12 * np.max(got["pion"])
    assert np.max(np.abs(want["ps"])) <= 1e-12 * np.max(want["pion"])


def test_meson_correlators_random_links_trace_form():
    """Random links, where every channel is nonzero: the oracle's contractions
    (the reference's index loops) equal the spin-colour trace forms applied to
    the propagator columns solved separately with the oracle's even-odd driver."""
    rng = np.random.default_rng(37)
    dims = (4, 4, 4, 6)
    kappa = 0.12
    V = int(np.prod(dims))
    U = fold_bc(dims, random_links(dims, rng, eps=0.5))
    fld = OracleField(dims, U, kappa=kappa, csw=1.0, need_clover=True)
    site = (1, 0, 2, 1)
    got = fld.meson_correlators(site, 1e-26)
    idx = ((site[0] * dims[1] + site[1]) * dims[2] + site[2]) * dims[3] + site[3]
    S = np.zeros((V, 12, 12), complex)
    for alpha in range(4):
        for c in range(3):
            b = np.zeros((V, 4, 3), complex)
            b[idx, alpha, c] = 1
            x, info, ok = fld.solve("precond_bicgstab", b, 1e-26)
            assert ok
            S[:, :, 3 * alpha + c] = x.reshape(V, 12)
    want = trace_correlators(S.reshape(dims + (12, 12)), kappa)
    for k in ("pion", "scalar", "ps", "pcac"):
        want[k] = np.roll(want[k], -site[3])  # relative to the source timeslice
        assert np.max(np.abs(got[k] - want[k])) <= 1e-9 * np.max(np.abs(want[k])), (k, got[k], want[k])
    assert np.max(np.abs(want["ps"])) > 1e-6 * np.max(want["pion"])  # far above rounding: the channel is not empty


def random_gauge_transform(dims, U, rng):
    """U'_mu(x) = g(x) U_mu(x) g(x + mu)^+ with Haar-random g."""
    from helpers import random_su3
    V = int(np.prod(dims))
    g = random_su3(V, rng)
    c = coords(dims)
    out = np.empty_like(U)
    for mu in range(4):
        cu = c.copy()
        cu[:, mu] = (cu[:, mu] + 1) % dims[mu]
        iu = ((cu[:, 0] * dims[1] + cu[:, 1]) * dims[2] + cu[:, 2]) * dims[3] + cu[:, 3]
        out[:, mu] = g @ U[:, mu] @ np.conj(np.swapaxes(g[iu], -1, -2))
    return out


def test_meson_correlators_gauge_invariant_and_time_shift():
    """On random links: (i) a gauge transformation leaves every correlator
    unchanged (the colour indices are traced correctly); (ii) the source at
    time t0 gives, relative to t0, what the source at t = 0 gives on links
    shifted by t0 in time (the reference's t_source_origin, MesonCorrelator.cpp:54-66)."""
    rng = np.random.default_rng(31)
    dims = (4, 4, 4, 6)
    kappa = 0.12
    U = random_links(dims, rng, eps=0.5)
    base = OracleField(dims, fold_bc(dims, U), kappa=kappa, csw=1.0, need_clover=True)
    ref = base.meson_correlators((1, 2, 3, 2), 1e-26)
    gt = OracleField(dims, fold_bc(dims, random_gauge_transform(dims, U, rng)), kappa=kappa, csw=1.0,
                     need_clover=True).meson_correlators((1, 2, 3, 2), 1e-26)
    for k in ("pion", "scalar", "ps", "pcac"):
        assert np.max(np.abs(gt[k] - ref[k])) <= 1e-9 * np.max(np.abs(ref[k])), k
    # the reference's shift: lattice'[x, t] = lattice[x, t + t0] on the links
    # with the boundary sign folded in, source at time 0
    t0 = 2
    Uf = fold_bc(dims, U).reshape(dims + (4, 3, 3))
    shifted = np.roll(Uf, -t0, axis=3).reshape(-1, 4, 3, 3)
    sh = OracleField(dims, shifted, kappa=kappa, csw=1.0, need_clover=True).meson_correlators((1, 2, 3, 0), 1e-26)
    for k in ("pion", "scalar", "ps", "pcac"):
        assert np.max(np.abs(sh[k] - ref[k])) <= 1e-9 * np.max(np.abs(ref[k])), k
