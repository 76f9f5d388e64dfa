"""CPU-only checks of the product library: it loads, exports every symbol the
C header declares, and its host-side input generators are right.  No GPU
compute is called here."""
import os
import re

import numpy as np
import pytest

from helpers import ROOT, coords

import leonardym_amd as lym


def declared_symbols():
    text = open(os.path.join(ROOT, "include", "lym.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(lym_[a-z0-9_]+)\s*\(", text)))


def test_library_exports_every_declared_symbol():
    L = lym.lib()
    syms = declared_symbols()
    assert len(syms) >= 40
    missing = [s for s in syms if not hasattr(L, s)]
    assert not missing, missing
    # and the Python binding covers exactly the declared ABI
    assert sorted(lym.SIGNATURES) == syms


def test_su3_generator_is_special_unitary():
    dims = (4, 4, 4, 4)
    U = lym.gen_su3_links(dims, 0.5, 1234)
    V = np.prod(dims)
    u = U.reshape(-1, 3, 3)
    eye = np.eye(3)
    assert np.max(np.abs(u @ np.conj(np.swapaxes(u, 1, 2)) - eye)) < 1e-14
    assert np.max(np.abs(np.linalg.det(u) - 1)) < 1e-14
    # exp(i eps H): distance from 1 grows with eps, and the generator is seeded
    U2 = lym.gen_su3_links(dims, 0.5, 1234)
    assert np.array_equal(U, U2)
    U3 = lym.gen_su3_links(dims, 0.5, 1235)
    assert not np.array_equal(U, U3)
    d05 = np.mean(np.abs(u - eye))
    d01 = np.mean(np.abs(lym.gen_su3_links(dims, 0.1, 1234).reshape(-1, 3, 3) - eye))
    assert d01 < d05 < 1.0
    assert V > 0


def test_su3_generator_matches_matrix_exponential():
    """Pins U = exp(i eps H): log(U)/(i eps) is hermitian, traceless, with
    Gell-Mann components of unit variance."""
    from scipy.linalg import logm
    dims = (2, 2, 2, 2)
    eps = 0.3
    u = lym.gen_su3_links(dims, eps, 99).reshape(-1, 3, 3)
    comps = []
    for m in u:
        H = logm(m) / (1j * eps)
        assert np.max(np.abs(H - H.conj().T)) < 1e-10
        assert abs(np.trace(H)) < 1e-10
        comps.append([2 * H[0, 1].real, -2 * H[0, 1].imag, H[0, 0].real - H[1, 1].real])
    comps = np.array(comps)
    assert 0.5 < np.std(comps) < 1.5


def test_su2_adjoint_map():
    """R_ab = 2 Re tr(T_a U T_b U^+) (ConvertLattice.h:101-105): an SO(3)
    homomorphism, the identity for U = 1, and a rotation by theta about axis 3
    for U = exp(i theta sigma_3 / 2)."""
    U = lym.gen_su2_links((2, 2, 2, 2), 0.7, 5).reshape(-1, 2, 2)
    R = lym.su2_to_adjoint(U)
    assert np.max(np.abs(R @ np.swapaxes(R, 1, 2) - np.eye(3))) < 1e-14
    assert np.max(np.abs(np.linalg.det(R) - 1)) < 1e-14
    RR = lym.su2_to_adjoint(U[:8] @ U[8:16])
    assert np.max(np.abs(RR - R[:8] @ R[8:16])) < 1e-14
    assert np.max(np.abs(lym.su2_to_adjoint(np.eye(2, dtype=complex)[None]) - np.eye(3))) < 1e-15
    th = 0.37
    u = np.diag([np.exp(1j * th / 2), np.exp(-1j * th / 2)])[None]
    r = lym.su2_to_adjoint(u)[0]
    c, s = np.cos(th), np.sin(th)
    assert np.max(np.abs(np.abs(r[:2, :2]) - np.abs(np.array([[c, s], [s, c]])))) < 1e-14
    assert abs(r[2, 2] - 1) < 1e-15
    # SU(2) links are unitary with det 1
    assert np.max(np.abs(U @ np.conj(np.swapaxes(U, 1, 2)) - np.eye(2))) < 1e-14


def test_su2_adjoint_closed_form():
    """ConvertLattice::toAdjoint (source/utils/ConvertLattice.h:24-34) with the
    generators T_a = sigma_a / 2 in the order sigma_x, sigma_y, sigma_z
    (source/utils/LieGenerators.h:18-25), checked entry by entry, signs
    included, against the quaternion form of the rotation: for
    U = u0 + i (u1 sigma_1 + u2 sigma_2 + u3 sigma_3),
        R = (u0^2 - |u|^2) 1 + 2 u u^T - 2 u0 [u]_x,
    with [u]_x the cross-product matrix.  A permutation of the generators or
    a flipped sign changes off-diagonal entries and fails here (the group
    properties alone do not see it).  The reference's element (0, 1) for
    U = diag(e^{i th/2}, e^{-i th/2}) is +sin th; the form above agrees."""
    U = lym.gen_su2_links((2, 2, 2, 4), 0.9, 17).reshape(-1, 2, 2)
    R = lym.su2_to_adjoint(U)
    u0, u3 = U[:, 0, 0].real, U[:, 0, 0].imag
    u2, u1 = U[:, 0, 1].real, U[:, 0, 1].imag
    # the parametrisation holds for these links (SU(2): U10 = -conj(U01), U11 = conj(U00))
    assert np.max(np.abs(U[:, 1, 0] - (1j * u1 - u2))) < 1e-14
    assert np.max(np.abs(U[:, 1, 1] - (u0 - 1j * u3))) < 1e-14
    u = np.stack([u1, u2, u3], axis=1)
    cross = np.zeros((len(U), 3, 3))
    cross[:, 0, 1], cross[:, 0, 2], cross[:, 1, 2] = -u[:, 2], u[:, 1], -u[:, 0]
    cross -= np.swapaxes(cross, 1, 2)
    ref = ((u0 ** 2 - np.sum(u * u, axis=1))[:, None, None] * np.eye(3) + 2 * u[:, :, None] * u[:, None, :]
           - 2 * u0[:, None, None] * cross)
    assert np.max(np.abs(R - ref)) < 1e-14
    # rotation by th about each axis k, with its sign: R_ij = cos th (i, j != k),
    # R_{k+1, k+2} = +sin th (indices mod 3)
    th = 0.61
    c, s = np.cos(th), np.sin(th)
    sig = [np.array([[0, 1], [1, 0]], complex), np.array([[0, -1j], [1j, 0]]), np.diag([1.0 + 0j, -1.0])]
    for k in range(3):
        uk = np.cos(th / 2) * np.eye(2) + 1j * np.sin(th / 2) * sig[k]
        r = lym.su2_to_adjoint(uk[None])[0]
        i, j = (k + 1) % 3, (k + 2) % 3
        want = np.zeros((3, 3))
        want[k, k] = 1
        want[i, i] = want[j, j] = c
        want[i, j], want[j, i] = s, -s
        assert np.max(np.abs(r - want)) < 1e-15, (k, r)


def test_bc_and_sources():
    dims = (2, 2, 4, 6)
    U = lym.gen_su3_links(dims, 0.5, 3)
    Ub = lym.apply_bc(dims, U, lym.REP_SU3_FUND, -1)
    t = coords(dims)[:, 3]
    ref = U.copy()
    ref[t == dims[3] - 1, 3] *= -1
    assert np.array_equal(Ub, ref)
    assert np.array_equal(lym.apply_bc(dims, U, lym.REP_SU3_FUND, 1), U)
    R = np.random.default_rng(0).standard_normal((int(np.prod(dims)), 4, 3, 3))
    Rb = lym.apply_bc(dims, R, lym.REP_SU2_ADJ, -1)
    refr = R.copy()
    refr[t == dims[3] - 1, 3] *= -1
    assert np.array_equal(Rb, refr)
    s = lym.point_source(dims, (1, 0, 3, 5), 2, 1)
    idx = ((1 * dims[1] + 0) * dims[2] + 3) * dims[3] + 5
    assert s[idx, 2, 1] == 1 and np.count_nonzero(s) == 1
    g = lym.gaussian_spinor(20000, 7)
    assert abs(np.std(g.real) - 1 / np.sqrt(2)) < 0.01 and abs(np.mean(g)) < 0.01
    assert np.array_equal(g, lym.gaussian_spinor(20000, 7))


def test_invalid_arguments_are_errors():
    with pytest.raises(lym.LymError):
        lym.point_source((2, 2, 2, 2), (0, 0, 0, 0), 4, 0)
    with pytest.raises(lym.LymError):
        lym.point_source((2, 2, 2, 2), (0, 0, 0, 2), 0, 0)


def test_argument_errors_are_codes_before_any_device_work():
    """Bad arguments come back as LYM_ERR_INVALID with a message, checked
    before the library touches a device (so they hold on a CPU-only host)."""
    import ctypes
    L = lym.lib()
    out = ctypes.c_void_p()
    # above the largest lattice (2^27 sites per site class: 32-bit buffer offsets
    # per field component; a lattice with an odd extent has one class)
    for dims, what in (((4, 4, 4, 1), b">= 2"), ((256, 256, 256, 64), b"2^27"), ((3, 512, 512, 256), b"2^27")):
        d = (ctypes.c_int * 4)(*dims)
        assert L.lym_ctx_create(d, lym.REP_SU3_FUND, -1, ctypes.byref(out)) == lym.ERR_INVALID
        assert what in L.lym_last_error()
    d = (ctypes.c_int * 4)(4, 4, 4, 4)
    assert L.lym_ctx_create(d, 7, -1, ctypes.byref(out)) == lym.ERR_INVALID
    assert L.lym_ctx_create(d, lym.REP_SU3_FUND, 0, ctypes.byref(out)) == lym.ERR_INVALID
    assert L.lym_ctx_create(None, lym.REP_SU3_FUND, -1, ctypes.byref(out)) == lym.ERR_INVALID
    assert L.lym_op_set_kappa(None, 0.1) == lym.ERR_INVALID
    assert L.lym_solver_set_rhs_block(None, 4) == lym.ERR_INVALID
    assert L.lym_solver_solve_batch(None, None, 0, None, None, None, None) == lym.ERR_INVALID
    assert L.lym_server_create(0, 0, 4, ctypes.byref(out)) == lym.ERR_INVALID
    assert L.lym_server_create(4, 0, 17, ctypes.byref(out)) == lym.ERR_INVALID
    assert L.lym_server_submit(None, 0, 0.1, 1e-8, None, None, None) == lym.ERR_INVALID
    assert L.lym_host_alloc(0, ctypes.byref(out)) == lym.OK and not out.value


def test_stencil_block_mapping_is_a_bijection():
    """Every launch shape of the stencil visits each (parity, site block) of
    its lattice exactly once: the workgroup -> block mapping (XCD ranges,
    sweeping grids, x-slice-fastest and XCD-pair renumberings) evaluated on the
    host by the library's own device functions.  Covers single-RHS and batched
    widths, workgroup sizes, full and half-lattice launches, and lattices that
    take each renumbering."""
    import itertools
    seen = np.zeros(3, int)
    lattices = [(4, 4, 4, 4), (6, 6, 6, 8), (8, 8, 8, 8), (4, 6, 2, 8), (12, 12, 12, 12), (16, 16, 16, 32),
                (16, 8, 12, 16), (18, 4, 4, 32), (10, 4, 4, 32), (8, 4, 8, 32), (16, 4, 4, 12), (24, 24, 24, 48),
                (24, 24, 24, 32), (32, 32, 32, 64), (5, 7, 3, 6)]
    lattices += [(5, 8, 8, 16), (3, 3, 3, 3), (4, 4, 4, 5)]
    for dims, parity, (nr, threads) in itertools.product(
            lattices, (-1, 0, 1), [(1, 0), (1, 64), (1, 128), (2, 0), (3, 192), (4, 128), (6, 0), (8, 0), (12, 0),
                                   (16, 256)]):
        # a lattice with an odd extent has one site class: launches are parity 0
        # over all V sites
        odd = any(d % 2 for d in dims)
        if odd and parity != 0:
            with pytest.raises(lym.LymError):
                lym.launch_plan(dims, parity, nr=nr, threads=threads)
            continue
        per_class = int(np.prod(dims)) // (1 if odd else 2)
        for lowp in ((False, True) if nr == 1 else (False,)):
            r = lym.launch_plan(dims, parity, nr=nr, threads=threads, lowp=lowp)
            v = r["visits"]
            want = np.zeros_like(v)
            want[[0, 1] if parity < 0 else [parity]] = 1
            assert np.array_equal(v, want), (dims, parity, nr, threads, lowp)
            assert r["spb"] * r["nbpp"] >= per_class > r["spb"] * (r["nbpp"] - 1)
            seen += r["branches"] > 0
    assert np.all(seen > 0)  # every renumbering ran in at least one shape
