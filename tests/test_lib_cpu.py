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
    # above the largest lattice (2^28 sites: 32-bit buffer offsets per field component)
    for dims, what in (((4, 4, 4, 3), b"even"), ((4, 4, 4, 1), b">= 2"), ((256, 256, 256, 64), b"2^28")):
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
