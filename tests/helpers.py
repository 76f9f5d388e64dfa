"""Shared test helpers: ctypes binding of the CPU oracle and numpy field
generators that are independent of the product library.

Array conventions (the reference's raw layouts, t-fastest lexicographic):
  spinor : complex128 [V, 4, 3]
  links  : complex128 [V, 4, 3, 3] with element (i, j) at [..., i, j]
           (the oracle and the product take them column-major, i.e. the
           memory order of U.transpose(-1,-2); see links_to_colmajor)
"""
import ctypes
import os
import subprocess

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE_DIR = os.path.join(ROOT, "oracle")
ORACLE_SO = os.path.join(ORACLE_DIR, "liblym_oracle.so")

# chiral-basis gamma matrices (source/utils/Gamma.cpp:9-10)
I = 1j
GAMMA = np.array([
    [[0, 0, 0, -I], [0, 0, -I, 0], [0, I, 0, 0], [I, 0, 0, 0]],
    [[0, 0, 0, -1], [0, 0, 1, 0], [0, 1, 0, 0], [-1, 0, 0, 0]],
    [[0, 0, -I, 0], [0, 0, 0, I], [I, 0, 0, 0], [0, -I, 0, 0]],
    [[0, 0, 1, 0], [0, 0, 0, 1], [1, 0, 0, 0], [0, 1, 0, 0]],
], dtype=complex)
GAMMA5 = np.diag([1, 1, -1, -1]).astype(complex)

K_WILSON, K_BASIC, K_CLOVER, K_EO, K_TEXTBOOK = 0, 1, 2, 3, 4


class OpDesc(ctypes.Structure):
    _fields_ = [("dims", ctypes.c_int * 4), ("kind", ctypes.c_int), ("power", ctypes.c_int),
                ("kappa", ctypes.c_double), ("csw", ctypes.c_double), ("gamma5", ctypes.c_int),
                ("links", ctypes.c_void_p), ("F", ctypes.c_void_p), ("inv", ctypes.c_void_p)]


class SolveInfo(ctypes.Structure):
    _fields_ = [("converged", ctypes.c_int), ("steps", ctypes.c_uint),
                ("last_error", ctypes.c_double), ("applies", ctypes.c_uint)]


_oracle = None


def oracle():
    """Load (building if needed) the CPU oracle library."""
    global _oracle
    if _oracle is not None:
        return _oracle
    if not os.path.exists(ORACLE_SO):
        subprocess.check_call(["make", "-s", "-C", ORACLE_DIR])
    lib = ctypes.CDLL(ORACLE_SO)
    P = ctypes.c_void_p
    lib.orc_multiply.argtypes = [ctypes.POINTER(OpDesc), P, P, P, ctypes.c_double, ctypes.c_double]
    lib.orc_field_strength.argtypes = [P, P, P]
    lib.orc_clover_inverse.argtypes = [P, P, ctypes.c_double, ctypes.c_double, P]
    lib.orc_clover_inverse.restype = ctypes.c_int
    lib.orc_clover_site_matrix.argtypes = [P, ctypes.c_double, ctypes.c_double, P]
    lib.orc_eo_hop.argtypes = [ctypes.POINTER(OpDesc), P, P, ctypes.c_int]
    lib.orc_eo_inv_even.argtypes = [ctypes.POINTER(OpDesc), P]
    lib.orc_eo_oddodd_minus.argtypes = [ctypes.POINTER(OpDesc), P, P, ctypes.c_int]
    lib.orc_tables.argtypes = [P, P, P, P]
    lib.orc_apply_bc.argtypes = [P, P]
    lib.orc_squared_norm.argtypes = [ctypes.c_long, P]
    lib.orc_squared_norm.restype = ctypes.c_double
    lib.orc_dot.argtypes = [ctypes.c_long, P, P, P, P]
    lib.orc_gamma5dot.argtypes = [ctypes.c_long, P, P, P, P]
    lib.orc_difference_norm.argtypes = [ctypes.c_long, P, P]
    lib.orc_difference_norm.restype = ctypes.c_double
    for name in ("orc_cg", "orc_bicgstab", "orc_precond_bicgstab"):
        f = getattr(lib, name)
        f.argtypes = [ctypes.POINTER(OpDesc), P, P, P, ctypes.c_double, ctypes.c_uint, ctypes.POINTER(SolveInfo)]
        f.restype = ctypes.c_int
    lib.orc_bicgstab_shift.argtypes = [ctypes.POINTER(OpDesc), P, P, P, ctypes.c_double, ctypes.c_uint,
                                       ctypes.c_double, ctypes.c_double, ctypes.POINTER(SolveInfo)]
    lib.orc_bicgstab_shift.restype = ctypes.c_int
    lib.orc_num_threads.restype = ctypes.c_int
    lib.orc_meson_correlators.argtypes = [ctypes.POINTER(OpDesc), P, ctypes.c_double, ctypes.c_uint, P]
    lib.orc_meson_correlators.restype = ctypes.c_int
    lib.orc_expmap.argtypes = [ctypes.c_int, P, P]
    lib.orc_stout_smear.argtypes = [P, ctypes.c_int, ctypes.c_double, P, P, P]
    lib.orc_wilson_fermion_force.argtypes = [P, ctypes.c_int, ctypes.c_double, P, P, P, P, P]
    _oracle = lib
    return lib


def _p(a):
    return None if a is None else a.ctypes.data


def _dims(dims):
    return (ctypes.c_int * 4)(*dims)


def oracle_expmap(X):
    """ExponentialMap::exp of one anti-hermitian traceless N x N matrix (oracle)."""
    n = X.shape[-1]
    cm = np.ascontiguousarray(X.T, np.complex128)
    out = np.empty_like(cm)
    oracle().orc_expmap(n, cm.ctypes.data, out.ctypes.data)
    return np.ascontiguousarray(out.T)


def oracle_stout(dims, U, rho, levels=1, staples=False):
    """Stout smearing of gauge links [V, 4, N, N] by the oracle; with
    staples=True also the first level's staple sums [V, 4, N, N]."""
    n = U.shape[-1]
    cur = np.ascontiguousarray(np.swapaxes(U, -1, -2), np.complex128)
    st = np.empty_like(cur) if staples else None
    for lev in range(levels):
        out = np.empty_like(cur)
        oracle().orc_stout_smear(_dims(dims), n, float(rho), cur.ctypes.data, out.ctypes.data,
                                 st.ctypes.data if (staples and lev == 0) else None)
        cur = out
    res = np.ascontiguousarray(np.swapaxes(cur, -1, -2))
    if staples:
        return res, np.ascontiguousarray(np.swapaxes(st, -1, -2))
    return res


def oracle_fermion_force(dims, rep, kappa, Wf, Ug, X, Y):
    """TwoFlavorFermionAction::force of every link by the oracle: fermion links
    Wf [V, 4, 3, 3] (boundary sign in), gauge links Ug [V, 4, N, N], spinors
    X, Y [V, 4, 3]; rep 0 SU(3), 1 SU(2) adjoint.  Returns [V, 4, N, N]."""
    N = Ug.shape[-1]
    wf = np.ascontiguousarray(np.swapaxes(np.asarray(Wf, complex), -1, -2))
    ug = np.ascontiguousarray(np.swapaxes(np.asarray(Ug, complex), -1, -2))
    out = np.empty((Ug.shape[0], 4, N, N), complex)
    oracle().orc_wilson_fermion_force(_dims(dims), rep, float(kappa), wf.ctypes.data, ug.ctypes.data,
                                      np.ascontiguousarray(X).ctypes.data, np.ascontiguousarray(Y).ctypes.data,
                                      out.ctypes.data)
    return np.ascontiguousarray(np.swapaxes(out, -1, -2))


def links_to_colmajor(U):
    """[V,4,3,3] (i,j) -> memory with (i,j) at i + 3 j."""
    return np.ascontiguousarray(np.swapaxes(U, -1, -2))


def colmajor_to_links(a):
    return np.ascontiguousarray(np.swapaxes(a, -1, -2))


class OracleField:
    """A gauge field prepared for the oracle (col-major links, F, R_ee^-1)."""

    def __init__(self, dims, U, kappa=0.125, csw=0.0, need_clover=False, need_inverse=True):
        self.dims = tuple(dims)
        self.V = int(np.prod(dims))
        self.U = U
        self.Ucm = links_to_colmajor(U)
        self.kappa = kappa
        self.csw = csw
        self.F = None
        self.inv = None
        if need_clover:
            self.build_clover(need_inverse)

    def build_clover(self, need_inverse=True):
        """F for the clover term; the per-site 12x12 inverses R^-1 (needed by
        the even-odd operator and driver only) unless need_inverse is False."""
        lib = oracle()
        self.F = np.zeros((self.V, 6, 3, 3), complex)
        lib.orc_field_strength(_p(np.array(self.dims, np.int32)), _p(self.Ucm), _p(self.F))
        if not need_inverse:
            return
        self.inv = np.zeros((self.V, 12, 12), complex)
        bad = lib.orc_clover_inverse(_p(np.array(self.dims, np.int32)), _p(self.F), self.kappa, self.csw,
                                     _p(self.inv))
        assert bad == 0

    def desc(self, kind, power=1, gamma5=True, kappa=None, csw=None):
        d = OpDesc()
        d.dims = _dims(self.dims)
        d.kind = kind
        d.power = power
        d.kappa = self.kappa if kappa is None else kappa
        d.csw = self.csw if csw is None else csw
        d.gamma5 = int(bool(gamma5))
        d.links = self.Ucm.ctypes.data
        d.F = None if self.F is None else self.F.ctypes.data
        d.inv = None if self.inv is None else self.inv.ctypes.data
        return d

    def apply(self, kind, psi, power=1, gamma5=True, v2=None, alpha=0.0, **kw):
        d = self.desc(kind, power, gamma5, **kw)
        out = np.zeros_like(psi)
        psi = np.ascontiguousarray(psi)
        v2c = None if v2 is None else np.ascontiguousarray(v2)
        oracle().orc_multiply(ctypes.byref(d), _p(out), _p(psi), _p(v2c), float(np.real(alpha)),
                              float(np.imag(alpha)))
        return out

    def meson_correlators(self, site, eps, max_steps=5000, kind=K_CLOVER):
        """MesonCorrelator::execute (connected part): dict of pion, scalar, ps,
        pcac over the sink time relative to the source's, and whether every
        solve converged."""
        d = self.desc(kind, 1, False)
        T = self.dims[3]
        out = np.zeros(4 * T)
        ok = oracle().orc_meson_correlators(ctypes.byref(d), _p(np.array(site, np.int32)), eps, max_steps, _p(out))
        return {"pion": out[:T], "scalar": out[T:2 * T], "ps": out[2 * T:3 * T], "pcac": out[3 * T:], "ok": bool(ok)}

    def solve(self, which, b, eps, max_steps=5000, kind=K_CLOVER, power=1, gamma5=True, x0=None, shift=None):
        d = self.desc(kind, power, gamma5)
        x = np.zeros_like(b)
        info = SolveInfo()
        if shift is not None:  # BiConjugateGradient.cpp:510-690, A + shift
            assert which == "bicgstab"
            ok = oracle().orc_bicgstab_shift(ctypes.byref(d), _p(np.ascontiguousarray(b)), _p(x), _p(x0), eps,
                                             max_steps, complex(shift).real, complex(shift).imag,
                                             ctypes.byref(info))
            return x, info, ok
        fn = getattr(oracle(), "orc_" + which)
        ok = fn(ctypes.byref(d), _p(np.ascontiguousarray(b)), _p(x), _p(x0), eps, max_steps, ctypes.byref(info))
        return x, info, ok


# ---------------------------------------------------------------- generators (numpy, test side)
def coords(dims):
    Lx, Ly, Lz, Lt = dims
    x, y, z, t = np.meshgrid(np.arange(Lx), np.arange(Ly), np.arange(Lz), np.arange(Lt), indexing="ij")
    return np.stack([x.ravel(), y.ravel(), z.ravel(), t.ravel()], axis=1)


def random_su3(n, rng, eps=None):
    """Random SU(3) matrices: Haar via QR when eps is None, else exp(i eps H)."""
    if eps is None:
        z = (rng.standard_normal((n, 3, 3)) + 1j * rng.standard_normal((n, 3, 3))) / np.sqrt(2)
        q, r = np.linalg.qr(z)
        d = np.diagonal(r, axis1=1, axis2=2)
        q = q * (d / np.abs(d))[:, None, :]
    else:
        h = rng.standard_normal((n, 3, 3)) + 1j * rng.standard_normal((n, 3, 3))
        h = 0.5 * (h + np.conj(np.swapaxes(h, 1, 2)))
        h -= np.trace(h, axis1=1, axis2=2)[:, None, None] / 3 * np.eye(3)
        w, v = np.linalg.eigh(h)
        q = np.einsum("nij,nj,nkj->nik", v, np.exp(1j * eps * w), np.conj(v))
    det = np.linalg.det(q)
    q = q / (det ** (1.0 / 3.0))[:, None, None]
    return q


def random_so3(n, rng):
    """Random real orthogonal 3x3 with det +1 (adjoint SU(2) links)."""
    z = rng.standard_normal((n, 3, 3))
    q, r = np.linalg.qr(z)
    q = q * np.sign(np.diagonal(r, axis1=1, axis2=2))[:, None, :]
    det = np.linalg.det(q)
    q[:, :, 0] *= det[:, None]
    return q.astype(complex)


def random_links(dims, rng, group="su3", eps=None):
    V = int(np.prod(dims))
    if group == "su3":
        U = random_su3(V * 4, rng, eps)
    else:
        U = random_so3(V * 4, rng)
    return U.reshape(V, 4, 3, 3)


def fold_bc(dims, U, bc_t=-1):
    """switchAntiperiodicBc (source/LatticeBoundaryConditions.h:8-14)."""
    U = U.copy()
    if bc_t == -1:
        t = coords(dims)[:, 3]
        U[t == dims[3] - 1, 3] *= -1
    return U


def gaussian_spinor(V, rng):
    return (rng.standard_normal((V, 4, 3)) + 1j * rng.standard_normal((V, 4, 3))) / np.sqrt(2)


def vdot(a, b):
    return np.vdot(a.ravel(), b.ravel())


def rel_err_per_site(a, b):
    """max over sites of |a - b| / |b| (site norms over spin x colour)."""
    da = np.linalg.norm((a - b).reshape(a.shape[0], -1), axis=1)
    nb = np.linalg.norm(b.reshape(b.shape[0], -1), axis=1)
    return float(np.max(da / np.maximum(nb, 1e-300)))


def dense_matrix(apply, V):
    """Dense matrix of a linear map on V x 12 complex by applying it to unit vectors."""
    n = V * 12
    M = np.zeros((n, n), complex)
    e = np.zeros((V, 4, 3), complex)
    flat = e.reshape(-1)
    for j in range(n):
        flat[:] = 0
        flat[j] = 1
        M[:, j] = apply(e).reshape(-1)
    return M
