"""lym_fermion_force_noalloc, the fermion force that allocates no device
memory, against lym_fermion_force (the same kernel with a per-call device
buffer), bit for bit, through the C ABI.

Two calls in a row on one GPU, the second on a smaller lattice, so the second
reads none of the first call's contents left in the library's fixed array:
a 16^4 SU(3) lattice (65536 sites, three chunks of the 16 MiB array) and
then an 8^4 one (one chunk), plus the SU(2) adjoint representation."""
import ctypes

import numpy as np
import pytest

from helpers import fold_bc

import leonardym_amd as lym

KAPPA = 0.12


def force_pair(dims, rep, seed):
    V = int(np.prod(dims))
    if rep == 0:
        W = fold_bc(dims, lym.gen_su3_links(dims, 0.5, seed), -1)
        ctx = lym.Context(dims, lym.REP_SU3_FUND, -1)
        N = 3
    else:
        W = fold_bc(dims, lym.su2_to_adjoint(lym.gen_su2_links(dims, 0.5, seed)), -1)
        ctx = lym.Context(dims, lym.REP_SU2_ADJ, -1)
        N = 2
    gauge = lym.Gauge(ctx, W, folded=True)
    op = lym.Operator(ctx, lym.OP_WILSON, 1, KAPPA, 0.0, True, gauge=gauge)
    X = lym.Spinor(ctx, lym.gaussian_spinor(V, seed + 1))
    Y = lym.Spinor(ctx, lym.gaussian_spinor(V, seed + 2))
    want = np.full((V, 4, N, N), np.nan, np.complex128)
    got = np.full((V, 4, N, N), np.nan, np.complex128)
    L = lym.lib()
    rc = L.lym_fermion_force(op.h, X.h, Y.h, want.ctypes.data_as(ctypes.c_void_p))
    assert rc == 0, L.lym_last_error().decode()
    rc = L.lym_fermion_force_noalloc(op.h, X.h, Y.h, got.ctypes.data_as(ctypes.c_void_p))
    assert rc == 0, L.lym_last_error().decode()
    return got, want


@pytest.mark.gpu
@pytest.mark.parametrize("rep", [0, 1], ids=["su3", "su2adj"])
def test_noalloc_force_equals_allocating_force(rep):
    for dims in ((16, 16, 16, 16), (8, 8, 8, 8)):
        print("noalloc force: rep %d on %s" % (rep, dims), flush=True)
        got, want = force_pair(dims, rep, 300 + dims[0] + rep)
        assert np.all(np.isfinite(want))
        assert np.array_equal(got, want)


@pytest.mark.gpu
def test_noalloc_force_rejects_what_the_allocating_one_rejects():
    dims = (4, 4, 4, 4)
    ctx = lym.Context(dims, lym.REP_SU3_FUND, -1)
    gauge = lym.Gauge(ctx, fold_bc(dims, lym.gen_su3_links(dims, 0.5, 5), -1), folded=True)
    op = lym.Operator(ctx, lym.OP_CLOVER, 1, KAPPA, 1.0, True, gauge=gauge)
    X, Y = lym.Spinor(ctx), lym.Spinor(ctx)
    out = np.empty((256, 4, 3, 3), np.complex128)
    rc = lym.lib().lym_fermion_force_noalloc(op.h, X.h, Y.h, out.ctypes.data_as(ctypes.c_void_p))
    assert rc == lym.ERR_UNSUPPORTED
    rc = lym.lib().lym_fermion_force_noalloc(op.h, X.h, Y.h, None)
    assert rc == lym.ERR_INVALID
