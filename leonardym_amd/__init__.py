"""Python binding of the MI355X LeonardYM fermion-solve library (liblym.so).

This is plumbing for tests and benchmarks: every call goes through the C ABI
declared in include/lym.h, every result is computed by the HIP kernels, and
there is no CPU fallback -- importing without a built liblym.so raises.
"""
import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
# LYM_LIB selects another build of the library (e.g. a variant under test)
LIB_PATH = os.environ.get("LYM_LIB") or os.path.join(HERE, "liblym.so")

OK, ERR_INVALID, ERR_DEVICE, ERR_NOMEM, ERR_UNSUPPORTED, ERR_SINGULAR = 0, 1, 2, 3, 4, 5
REP_SU3_FUND, REP_SU2_ADJ = 0, 1
OP_WILSON, OP_BASIC_WILSON, OP_CLOVER, OP_EO_CLOVER = 0, 1, 2, 3
SOLVER_CG, SOLVER_BICGSTAB, SOLVER_PRECOND_BICGSTAB = 0, 1, 2
SOLVER_MIXED_CG, SOLVER_PRECOND_CG, SOLVER_PRECOND_MIXED_CG = 3, 4, 5
PREC_DOUBLE, PREC_SINGLE, PREC_HALF = 0, 1, 2
REQ_PENDING, REQ_RUNNING, REQ_DONE, REQ_FAILED = 0, 1, 2, 3
LINKS_FULL, LINKS_RECONSTRUCT = 0, 1
FORMAT_LEONARD, FORMAT_MUENSTER = 0, 1

# C ABI: name -> (restype, argtypes)
_P = ctypes.c_void_p
_I = ctypes.c_int
_D = ctypes.c_double
_U = ctypes.c_uint
_U64 = ctypes.c_uint64
_L = ctypes.c_long
_PD = ctypes.POINTER(ctypes.c_double)
_PI = ctypes.POINTER(ctypes.c_int)


class SolveInfo(ctypes.Structure):
    _fields_ = [("converged", ctypes.c_int), ("steps", ctypes.c_uint), ("last_error", ctypes.c_double),
                ("applies", ctypes.c_uint), ("seconds", ctypes.c_double), ("reliable_updates", ctypes.c_uint)]


class RequestInfo(ctypes.Structure):
    _fields_ = [("status", ctypes.c_int), ("error", ctypes.c_int), ("batch_size", ctypes.c_uint),
                ("worker", ctypes.c_int), ("queue_seconds", ctypes.c_double), ("service_seconds", ctypes.c_double),
                ("latency_seconds", ctypes.c_double), ("solve", SolveInfo)]


class ConfigInfo(ctypes.Structure):
    _fields_ = [("plaquette", ctypes.c_double), ("beta", ctypes.c_double), ("kappa", ctypes.c_double),
                ("c1", ctypes.c_double), ("sweep", ctypes.c_int)]


SIGNATURES = {
    "lym_last_error": (ctypes.c_char_p, []),
    "lym_device_count": (_I, [_PI]),
    "lym_host_alloc": (_I, [ctypes.c_size_t, ctypes.POINTER(_P)]),
    "lym_host_free": (None, [_P]),
    "lym_ctx_create": (_I, [_P, _I, _I, ctypes.POINTER(_P)]),
    "lym_ctx_destroy": (None, [_P]),
    "lym_ctx_synchronize": (_I, [_P]),
    "lym_ctx_stream": (_P, [_P]),
    "lym_ctx_timer_start": (_I, [_P]),
    "lym_ctx_timer_stop": (_I, [_P, ctypes.POINTER(ctypes.c_float)]),
    "lym_gauge_create": (_I, [_P, _P, _I, _I, ctypes.POINTER(_P)]),
    "lym_gauge_destroy": (None, [_P]),
    "lym_spinor_create": (_I, [_P, ctypes.POINTER(_P)]),
    "lym_spinor_destroy": (None, [_P]),
    "lym_spinor_upload": (_I, [_P, _P]),
    "lym_spinor_download": (_I, [_P, _P]),
    "lym_spinor_copy": (_I, [_P, _P]),
    "lym_spinor_set_zero": (_I, [_P]),
    "lym_spinor_gamma5": (_I, [_P]),
    "lym_spinor_squared_norm": (_I, [_P, _PD]),
    "lym_spinor_dot": (_I, [_P, _P, _PD, _PD]),
    "lym_spinor_gamma5dot": (_I, [_P, _P, _PD, _PD]),
    "lym_spinor_difference_norm": (_I, [_P, _P, _PD]),
    "lym_spinor_axpy": (_I, [_P, _D, _D, _P]),
    "lym_op_create": (_I, [_P, _I, _U, _D, _D, _I, ctypes.POINTER(_P)]),
    "lym_op_create_by_name": (_I, [_P, ctypes.c_char_p, _U, _D, _D, _I, ctypes.POINTER(_P)]),
    "lym_op_destroy": (None, [_P]),
    "lym_op_set_gauge": (_I, [_P, _P]),
    "lym_op_set_kappa": (_I, [_P, _D]),
    "lym_op_set_csw": (_I, [_P, _D]),
    "lym_op_set_gamma5": (_I, [_P, _I]),
    "lym_op_get_params": (_I, [_P, _PD, _PD, _PI, _PI, ctypes.POINTER(_U)]),
    "lym_op_multiply": (_I, [_P, _P, _P]),
    "lym_op_multiply_add": (_I, [_P, _P, _P, _P, _D, _D]),
    "lym_op_multiply_sloppy": (_I, [_P, _I, _P, _P]),
    "lym_solver_create": (_I, [_P, _I, ctypes.POINTER(_P)]),
    "lym_solver_destroy": (None, [_P]),
    "lym_solver_set_precision": (_I, [_P, _D]),
    "lym_solver_set_max_steps": (_I, [_P, _U]),
    "lym_solver_set_even_odd": (_I, [_P, _I]),
    "lym_solver_set_sloppy_precision": (_I, [_P, _I]),
    "lym_solver_set_reliable_delta": (_I, [_P, _D]),
    "lym_solver_set_shift": (_I, [_P, _D, _D]),
    "lym_solver_solve": (_I, [_P, _P, _P, _P, _P, ctypes.POINTER(SolveInfo)]),
    "lym_solver_solve_batch": (_I, [_P, _P, _U, ctypes.POINTER(_P), ctypes.POINTER(_P), ctypes.POINTER(_D),
                                    ctypes.POINTER(SolveInfo)]),
    "lym_solver_set_rhs_block": (_I, [_P, _U]),
    "lym_solver_set_matrix_cores": (_I, [_P, _I]),
    "lym_solver_solve_batch_kappas": (_I, [_P, _P, _U, _PD, _P, _P, _PD, ctypes.POINTER(SolveInfo)]),
    "lym_solver_get_matrix_cores": (_I, [_P, _PI, _PI]),
    "lym_op_multiply_sloppy_batch": (_I, [_P, _I, _I, _U, _P, _P]),
    "lym_server_create": (_I, [_U, ctypes.c_size_t, _U, ctypes.POINTER(_P)]),
    "lym_server_destroy": (None, [_P]),
    "lym_server_set_solver": (_I, [_P, _I, _U]),
    "lym_server_add_lattice": (_I, [_P, _PI, _I, _I, _PD, _I, _I, _D, _PI]),
    "lym_server_submit": (_I, [_P, _I, _D, _D, _PD, _PD, ctypes.POINTER(_U64)]),
    "lym_server_wait": (_I, [_P, _U64, _D, ctypes.POINTER(RequestInfo)]),
    "lym_server_release": (_I, [_P, _U64]),
    "lym_server_drain": (_I, [_P]),
    "lym_server_memory": (_I, [_P, ctypes.POINTER(ctypes.c_size_t), ctypes.POINTER(ctypes.c_size_t)]),
    "lym_solver_get_rhs_block": (_I, [_P, ctypes.POINTER(_U), ctypes.POINTER(_U)]),
    "lym_gen_su3_links": (_I, [_P, _D, _U64, _P]),
    "lym_gen_su2_links": (_I, [_P, _D, _U64, _P]),
    "lym_su2_to_adjoint": (_I, [_L, _P, _P]),
    "lym_apply_bc": (_I, [_P, _I, _I, _P]),
    "lym_gen_gaussian_spinor": (_I, [_L, _U64, _P]),
    "lym_point_source": (_I, [_P, _P, _I, _I, _P]),
    "lym_launch_plan": (_I, [_P, _I, _I, _I, _I, _PI, _PI, _PI, _PI, _PI, _L, _PI]),
    "lym_meson_correlators": (_I, [_P, _P, _P, _PD, ctypes.POINTER(SolveInfo)]),
    "lym_config_write": (_I, [_I, ctypes.c_char_p, ctypes.c_char_p, _I, _PI, _I, _P, ctypes.POINTER(ConfigInfo)]),
    "lym_config_read": (_I, [_I, ctypes.c_char_p, ctypes.c_char_p, _I, _PI, _I, _P, ctypes.POINTER(ConfigInfo)]),
    "lym_temporal_plaquette": (_I, [_PI, _I, _P, _PD]),
    "lym_stout_smear": (_I, [_PI, _I, _D, _I, _P, _P]),
    "lym_fermion_force": (_I, [_P, _P, _P, _P]),
    "lym_fermion_force_noalloc": (_I, [_P, _P, _P, _P]),
    "lym_two_flavor_force": (_I, [_P, _P, _P, _P, ctypes.POINTER(SolveInfo)]),
}

_lib = None


class LymError(RuntimeError):
    pass


def lib():
    """The loaded liblym.so; raises if it has not been built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise LymError("liblym.so is not built (run __graft_entry__.build() or make -C leonardym_amd)")
        L = ctypes.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def _check(rc):
    if rc != 0:
        raise LymError("lym error %d: %s" % (rc, lib().lym_last_error().decode()))


def _arr(a):
    return a.ctypes.data_as(ctypes.c_void_p)


def device_count():
    n = ctypes.c_int()
    rc = lib().lym_device_count(ctypes.byref(n))
    return n.value if rc == 0 else 0


def launch_plan(dims, parity=-1, nr=1, threads=0, lowp=False):
    """The stencil's workgroup -> site-block mapping for one launch, computed
    on the host by the library (no device needed): dict with grid, spb, nbpp,
    xslice_blocks, visits [2, nbpp] and branches (plain, x-slice, XCD pairs)."""
    grid, spb, nbpp, W = (ctypes.c_int() for _ in range(4))
    _check(lib().lym_launch_plan(_arr(np.array(dims, np.int32)), parity, nr, threads, int(lowp), ctypes.byref(grid),
                                 ctypes.byref(spb), ctypes.byref(nbpp), ctypes.byref(W), None, 0, None))
    visits = np.zeros(2 * nbpp.value, np.int32)
    branches = np.zeros(3, np.int32)
    _check(lib().lym_launch_plan(_arr(np.array(dims, np.int32)), parity, nr, threads, int(lowp), ctypes.byref(grid),
                                 ctypes.byref(spb), ctypes.byref(nbpp), ctypes.byref(W),
                                 visits.ctypes.data_as(_PI), visits.size, branches.ctypes.data_as(_PI)))
    return {"grid": grid.value, "spb": spb.value, "nbpp": nbpp.value, "xslice_blocks": W.value,
            "visits": visits.reshape(2, nbpp.value), "branches": branches}


# ---------------------------------------------------------------- host-side inputs
def _dims_arr(dims):
    return np.ascontiguousarray(np.array(dims, np.int32))


def config_write(fmt, directory, name, number, dims, U, beta=0.0, kappa=0.0, sweep=0):
    """Write gauge links U [V, 4, N, N] (element (i, j) at [..., i, j]) in the
    reference's leonard or muenster format; returns the ConfigInfo (plaquette)."""
    N = U.shape[-1]
    cm = np.ascontiguousarray(np.swapaxes(np.asarray(U, np.complex128), -1, -2))
    d = _dims_arr(dims)
    info = ConfigInfo(0.0, float(beta), float(kappa), 0.0, int(sweep))
    _check(lib().lym_config_write(fmt, directory.encode(), name.encode(), int(number),
                                  d.ctypes.data_as(_PI), N, _arr(cm), ctypes.byref(info)))
    return info


def config_read(fmt, directory, name, number, dims, ncolors):
    """Read gauge links [V, 4, N, N] and the ConfigInfo of a configuration file."""
    V = int(np.prod(dims))
    cm = np.empty((V, 4, ncolors, ncolors), np.complex128)
    d = _dims_arr(dims)
    info = ConfigInfo()
    _check(lib().lym_config_read(fmt, directory.encode(), name.encode(), int(number), d.ctypes.data_as(_PI),
                                 ncolors, _arr(cm), ctypes.byref(info)))
    return np.ascontiguousarray(np.swapaxes(cm, -1, -2)), info


def temporal_plaquette(dims, U):
    cm = np.ascontiguousarray(np.swapaxes(np.asarray(U, np.complex128), -1, -2))
    d = _dims_arr(dims)
    out = ctypes.c_double()
    _check(lib().lym_temporal_plaquette(d.ctypes.data_as(_PI), U.shape[-1], _arr(cm), ctypes.byref(out)))
    return out.value


def stout_smear(dims, U, rho, levels=1):
    """Stout-smeared copy of fundamental gauge links U [V, 4, N, N] (on the device)."""
    cm = np.ascontiguousarray(np.swapaxes(np.asarray(U, np.complex128), -1, -2))
    out = np.empty_like(cm)
    d = _dims_arr(dims)
    _check(lib().lym_stout_smear(d.ctypes.data_as(_PI), U.shape[-1], float(rho), int(levels), _arr(cm), _arr(out)))
    return np.ascontiguousarray(np.swapaxes(out, -1, -2))


def gen_su3_links(dims, eps, seed):
    V = int(np.prod(dims))
    out = np.empty((V, 4, 3, 3), np.complex128)  # memory: column-major per link
    _check(lib().lym_gen_su3_links(_arr(np.array(dims, np.int32)), eps, seed, _arr(out)))
    return np.ascontiguousarray(np.swapaxes(out, -1, -2))  # -> [..., i, j]


def gen_su2_links(dims, eps, seed):
    V = int(np.prod(dims))
    out = np.empty((V, 4, 2, 2), np.complex128)
    _check(lib().lym_gen_su2_links(_arr(np.array(dims, np.int32)), eps, seed, _arr(out)))
    return np.ascontiguousarray(np.swapaxes(out, -1, -2))


def su2_to_adjoint(U2):
    """[..., 2, 2] complex (i, j) -> [..., 3, 3] real (a, b)."""
    shp = U2.shape[:-2]
    cm = np.ascontiguousarray(np.swapaxes(U2, -1, -2)).reshape(-1, 2, 2)
    out = np.empty((cm.shape[0], 3, 3), np.float64)
    _check(lib().lym_su2_to_adjoint(cm.shape[0], _arr(cm), _arr(out)))
    return np.ascontiguousarray(np.swapaxes(out, -1, -2)).reshape(shp + (3, 3))


def apply_bc(dims, links, rep, bc_t=-1):
    """Fold the time-boundary sign into [V,4,3,3] links (returns a copy)."""
    cm = np.ascontiguousarray(np.swapaxes(links, -1, -2))
    _check(lib().lym_apply_bc(_arr(np.array(dims, np.int32)), rep, bc_t, _arr(cm)))
    return np.ascontiguousarray(np.swapaxes(cm, -1, -2))


def gaussian_spinor(nsites, seed):
    out = np.empty((nsites, 4, 3), np.complex128)
    _check(lib().lym_gen_gaussian_spinor(nsites, seed, _arr(out)))
    return out


class _Pinned:
    """Owner of a page-locked host block; freed when the last array viewing it goes."""

    def __init__(self, nbytes):
        p = ctypes.c_void_p()
        _check(lib().lym_host_alloc(nbytes, ctypes.byref(p)))
        self.p, self.nbytes = p, nbytes

    def __del__(self):
        try:
            if self.p:
                lib().lym_host_free(self.p)
                self.p = None
        except Exception:
            pass


def pinned_empty(shape, dtype=np.complex128):
    """numpy array in page-locked host memory (fast uploads and downloads)."""
    dt = np.dtype(dtype)
    n = int(np.prod(shape)) * dt.itemsize
    owner = _Pinned(max(n, 1))
    buf = (ctypes.c_char * max(n, 1)).from_address(owner.p.value)
    buf._owner = owner  # keeps the block alive as long as the array
    return np.frombuffer(buf, dtype=dt, count=int(np.prod(shape))).reshape(shape)


def point_source(dims, site, spin, colour):
    V = int(np.prod(dims))
    out = np.empty((V, 4, 3), np.complex128)
    _check(lib().lym_point_source(_arr(np.array(dims, np.int32)), _arr(np.array(site, np.int32)), spin, colour,
                                  _arr(out)))
    return out


# ---------------------------------------------------------------- device objects
class Context:
    def __init__(self, dims, rep=REP_SU3_FUND, bc_t=-1):
        self.dims = tuple(int(d) for d in dims)
        self.V = int(np.prod(self.dims))
        self.rep = rep
        h = ctypes.c_void_p()
        _check(lib().lym_ctx_create(_arr(np.array(self.dims, np.int32)), rep, bc_t, ctypes.byref(h)))
        self.h = h

    def synchronize(self):
        _check(lib().lym_ctx_synchronize(self.h))

    def timer_start(self):
        _check(lib().lym_ctx_timer_start(self.h))

    def timer_stop(self):
        ms = ctypes.c_float()
        _check(lib().lym_ctx_timer_stop(self.h, ctypes.byref(ms)))
        return ms.value

    def close(self):
        if self.h:
            lib().lym_ctx_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Gauge:
    def __init__(self, ctx, links, folded=True, storage=LINKS_FULL):
        """links: [V,4,3,3] complex (SU(3)) or real (SU(2) adjoint), element (i, j)."""
        self.ctx = ctx
        if ctx.rep == REP_SU2_ADJ:
            cm = np.ascontiguousarray(np.swapaxes(np.asarray(links, np.float64), -1, -2))
        else:
            cm = np.ascontiguousarray(np.swapaxes(np.asarray(links, np.complex128), -1, -2))
        h = ctypes.c_void_p()
        _check(lib().lym_gauge_create(ctx.h, _arr(cm), int(folded), storage, ctypes.byref(h)))
        self.h = h

    def close(self):
        if self.h:
            lib().lym_gauge_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Spinor:
    def __init__(self, ctx, host=None):
        self.ctx = ctx
        h = ctypes.c_void_p()
        _check(lib().lym_spinor_create(ctx.h, ctypes.byref(h)))
        self.h = h
        if host is not None:
            self.upload(host)

    def upload(self, host):
        a = np.ascontiguousarray(host, np.complex128)
        assert a.size == self.ctx.V * 12
        _check(lib().lym_spinor_upload(self.h, _arr(a)))

    def download(self):
        out = np.empty((self.ctx.V, 4, 3), np.complex128)
        _check(lib().lym_spinor_download(self.h, _arr(out)))
        return out

    def squared_norm(self):
        r = ctypes.c_double()
        _check(lib().lym_spinor_squared_norm(self.h, ctypes.byref(r)))
        return r.value

    def dot(self, other):
        re, im = ctypes.c_double(), ctypes.c_double()
        _check(lib().lym_spinor_dot(self.h, other.h, ctypes.byref(re), ctypes.byref(im)))
        return complex(re.value, im.value)

    def gamma5dot(self, other):
        re, im = ctypes.c_double(), ctypes.c_double()
        _check(lib().lym_spinor_gamma5dot(self.h, other.h, ctypes.byref(re), ctypes.byref(im)))
        return complex(re.value, im.value)

    def difference_norm(self, other):
        r = ctypes.c_double()
        _check(lib().lym_spinor_difference_norm(self.h, other.h, ctypes.byref(r)))
        return r.value

    def axpy(self, a, x):
        _check(lib().lym_spinor_axpy(self.h, complex(a).real, complex(a).imag, x.h))

    def gamma5(self):
        _check(lib().lym_spinor_gamma5(self.h))

    def set_zero(self):
        _check(lib().lym_spinor_set_zero(self.h))

    def copy_from(self, other):
        _check(lib().lym_spinor_copy(self.h, other.h))

    def close(self):
        if self.h:
            lib().lym_spinor_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Operator:
    def __init__(self, ctx, kind, power=1, kappa=0.125, csw=1.0, gamma5=True, gauge=None, name=None):
        self.ctx = ctx
        h = ctypes.c_void_p()
        if name is not None:
            _check(lib().lym_op_create_by_name(ctx.h, name.encode(), power, kappa, csw, int(gamma5), ctypes.byref(h)))
        else:
            _check(lib().lym_op_create(ctx.h, kind, power, kappa, csw, int(gamma5), ctypes.byref(h)))
        self.h = h
        if gauge is not None:
            self.set_gauge(gauge)

    def set_gauge(self, g):
        self.gauge = g
        _check(lib().lym_op_set_gauge(self.h, g.h))

    def set_kappa(self, k):
        _check(lib().lym_op_set_kappa(self.h, k))

    def set_csw(self, c):
        _check(lib().lym_op_set_csw(self.h, c))

    def set_gamma5(self, g):
        _check(lib().lym_op_set_gamma5(self.h, int(g)))

    def multiply(self, out, inp):
        _check(lib().lym_op_multiply(self.h, out.h, inp.h))

    def multiply_add(self, out, v1, v2, alpha):
        a = complex(alpha)
        _check(lib().lym_op_multiply_add(self.h, out.h, v1.h, v2.h, a.real, a.imag))

    def _force_array(self):
        N = 3 if self.ctx.rep == REP_SU3_FUND else 2
        return np.empty((self.ctx.V, 4, N, N), np.complex128)

    def fermion_force(self, X, Y):
        """TwoFlavorFermionAction::force of every link, [V, 4, N, N] (i, j) at [..., i, j]."""
        out = self._force_array()
        _check(lib().lym_fermion_force(self.h, X.h, Y.h, _arr(out)))
        return np.ascontiguousarray(np.swapaxes(out, -1, -2))

    def multiply_sloppy(self, out, inp, prec):
        """out = Op inp in storage precision prec (PREC_SINGLE or PREC_HALF)."""
        _check(lib().lym_op_multiply_sloppy(self.h, prec, out.h, inp.h))

    def multiply_sloppy_batch(self, outs, ins, prec, matrix_cores=False):
        """outs[q] = Op ins[q] for every q in one multi-RHS apply at storage precision prec."""
        n = len(ins)
        if len(outs) != n:
            raise ValueError("outs and ins differ in length")
        o = (_P * n)(*[x.h for x in outs])
        i = (_P * n)(*[x.h for x in ins])
        _check(lib().lym_op_multiply_sloppy_batch(self.h, prec, int(bool(matrix_cores)), n, o, i))

    def close(self):
        if self.h:
            lib().lym_op_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Solver:
    def __init__(self, ctx, kind, precision=None, max_steps=None, even_odd=True, sloppy=None, delta=None,
                 shift=None):
        self.ctx = ctx
        h = ctypes.c_void_p()
        _check(lib().lym_solver_create(ctx.h, kind, ctypes.byref(h)))
        self.h = h
        if sloppy is not None:
            _check(lib().lym_solver_set_sloppy_precision(self.h, sloppy))
        if delta is not None:
            _check(lib().lym_solver_set_reliable_delta(self.h, delta))
        if precision is not None:
            _check(lib().lym_solver_set_precision(self.h, precision))
        if max_steps is not None:
            _check(lib().lym_solver_set_max_steps(self.h, max_steps))
        _check(lib().lym_solver_set_even_odd(self.h, int(even_odd)))
        if shift is not None:
            _check(lib().lym_solver_set_shift(self.h, complex(shift).real, complex(shift).imag))

    def solve(self, op, source, solution, guess=None):
        info = SolveInfo()
        _check(lib().lym_solver_solve(self.h, op.h, source.h, solution.h, guess.h if guess is not None else None,
                                      ctypes.byref(info)))
        return info

    def two_flavor_force(self, op, phi):
        """TwoFlavorFermionAction::updateForce: the two solves and the force
        [V, 4, N, N]; returns (force, (info_Y, info_X))."""
        N = 3 if op.ctx.rep == REP_SU3_FUND else 2
        out = np.empty((op.ctx.V, 4, N, N), np.complex128)
        infos = (SolveInfo * 2)()
        _check(lib().lym_two_flavor_force(self.h, op.h, phi.h, _arr(out), infos))
        return np.ascontiguousarray(np.swapaxes(out, -1, -2)), (infos[0], infos[1])

    def solve_batch(self, op, sources, solutions, precisions=None):
        """Solve op x_q = b_q for every q, batched; returns one SolveInfo per RHS."""
        n = len(sources)
        if len(solutions) != n:
            raise ValueError("sources and solutions differ in length")
        srcs = (_P * max(n, 1))(*[s.h for s in sources])
        sols = (_P * max(n, 1))(*[s.h for s in solutions])
        infos = (SolveInfo * max(n, 1))()
        prec = None
        if precisions is not None:
            prec = (_D * max(n, 1))(*[float(p) for p in precisions])
        _check(lib().lym_solver_solve_batch(self.h, op.h, n, srcs, sols, prec, infos))
        return [infos[i] for i in range(n)]

    def set_rhs_block(self, n):
        _check(lib().lym_solver_set_rhs_block(self.h, n))

    def solve_batch_kappas(self, op, kappas, sources, solutions, precisions=None):
        """Solve D(kappa_q) x_q = b_q for every q as lanes of one batch (even-odd
        mixed CG of a clover operator); returns one SolveInfo per RHS."""
        n = len(sources)
        if len(solutions) != n or len(kappas) != n:
            raise ValueError("kappas, sources and solutions differ in length")
        k = np.ascontiguousarray(kappas, dtype=np.float64)
        srcs = (_P * max(n, 1))(*[s.h for s in sources])
        sols = (_P * max(n, 1))(*[s.h for s in solutions])
        infos = (SolveInfo * max(n, 1))()
        prec = None
        if precisions is not None:
            prec = np.ascontiguousarray(precisions, dtype=np.float64)
        _check(lib().lym_solver_solve_batch_kappas(self.h, op.h, n, k.ctypes.data_as(_PD), srcs, sols,
                                                   prec.ctypes.data_as(_PD) if prec is not None else None, infos))
        return [infos[i] for i in range(n)]

    def rhs_block(self):
        a, b = _U(), _U()
        _check(lib().lym_solver_get_rhs_block(self.h, ctypes.byref(a), ctypes.byref(b)))
        return a.value, b.value

    def set_matrix_cores(self, enable):
        _check(lib().lym_solver_set_matrix_cores(self.h, int(bool(enable))))

    def matrix_cores(self):
        """(enabled, used by the last batched call)"""
        a, b = ctypes.c_int(), ctypes.c_int()
        _check(lib().lym_solver_get_matrix_cores(self.h, ctypes.byref(a), ctypes.byref(b)))
        return bool(a.value), bool(b.value)

    def close(self):
        if self.h:
            lib().lym_solver_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def meson_correlators(solver, op, source_site):
    """MesonCorrelator::execute on the device: dict of pion, scalar, ps (4 kappa^2
    normalised) and pcac arrays over the sink time relative to the source
    timeslice, and the 12 SolveInfo of the propagator columns."""
    T = op.ctx.dims[3]
    out = np.zeros(4 * T)
    infos = (SolveInfo * 12)()
    _check(lib().lym_meson_correlators(solver.h, op.h, _arr(np.array(source_site, np.int32)),
                                       out.ctypes.data_as(_PD), infos))
    return {"pion": out[:T], "scalar": out[T:2 * T], "ps": out[2 * T:3 * T], "pcac": out[3 * T:],
            "infos": [infos[i] for i in range(12)]}


class Server:
    """Request loop over device-resident gauge fields (see lym_server_* in lym.h).

    submit() keeps references to the source and solution arrays until the
    request is waited for, as the library reads and writes them asynchronously."""

    def __init__(self, workers=4, hbm_budget=0, max_batch=16, sloppy=None, max_steps=0):
        h = ctypes.c_void_p()
        _check(lib().lym_server_create(workers, hbm_budget, max_batch, ctypes.byref(h)))
        self.h = h
        self._live = {}
        if sloppy is not None or max_steps:
            _check(lib().lym_server_set_solver(self.h, PREC_HALF if sloppy is None else sloppy, max_steps))

    def add_lattice(self, dims, links, rep=REP_SU3_FUND, bc=-1, folded=False, storage=LINKS_FULL, csw=1.0):
        """links: [V,4,3,3], element (i, j), as for Gauge."""
        dt = np.float64 if rep == REP_SU2_ADJ else np.complex128
        links = np.ascontiguousarray(np.swapaxes(np.asarray(links, dt), -1, -2))
        d = (ctypes.c_int * 4)(*dims)
        lid = ctypes.c_int()
        _check(lib().lym_server_add_lattice(self.h, d, rep, bc, links.ctypes.data_as(_PD), int(folded), storage,
                                            csw, ctypes.byref(lid)))
        return lid.value

    def submit(self, lattice, kappa, tol, source, out=None):
        source = np.ascontiguousarray(source, dtype=np.complex128)
        if out is None:
            out = np.empty_like(source)
        if out.dtype != np.complex128 or not out.flags.c_contiguous or out.size != source.size:
            raise ValueError("solution must be a C-contiguous complex128 array of the source's size")
        rid = _U64()
        _check(lib().lym_server_submit(self.h, lattice, kappa, tol, source.ctypes.data_as(_PD),
                                       out.ctypes.data_as(_PD), ctypes.byref(rid)))
        self._live[rid.value] = (source, out)
        return rid.value

    def wait(self, rid, timeout=-1.0):
        """(RequestInfo, solution array); the request is released once complete."""
        info = RequestInfo()
        rc = lib().lym_server_wait(self.h, rid, timeout, ctypes.byref(info))
        if info.status in (REQ_DONE, REQ_FAILED):
            src, out = self._live.pop(rid)
            lib().lym_server_release(self.h, rid)
            _check(rc)
            return info, out
        _check(rc)
        return info, None

    def drain(self):
        _check(lib().lym_server_drain(self.h))

    def memory(self):
        a, b = ctypes.c_size_t(), ctypes.c_size_t()
        _check(lib().lym_server_memory(self.h, ctypes.byref(a), ctypes.byref(b)))
        return a.value, b.value

    def close(self):
        if self.h:
            lib().lym_server_destroy(self.h)
            self.h = None
            self._live.clear()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
