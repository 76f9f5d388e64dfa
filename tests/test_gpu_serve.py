"""Serving loop: requests on several device-resident lattices, solved by
worker threads with batching of requests that share an operator.  A request's
result is the same alone and in a mix; every result meets its tolerance
against the oracle's D; HBM admission releases idle state and refuses what
cannot fit."""
import numpy as np
import pytest

from helpers import K_CLOVER, OracleField, fold_bc

import leonardym_amd as lym

pytestmark = pytest.mark.gpu

LATTICES = [(4, 4, 4, 4), (4, 4, 4, 8), (6, 4, 4, 4)]
CSW = 1.0


@pytest.fixture(scope="module")
def fields():
    if lym.device_count() < 1:
        pytest.fail("no HIP device visible")
    out = []
    for i, dims in enumerate(LATTICES):
        U = lym.gen_su3_links(dims, 0.5, 500 + i)
        out.append((dims, U))
    return out


def oracle_residual(dims, U, kappa, x, b):
    orc = OracleField(dims, fold_bc(dims, U, -1), kappa=kappa, csw=CSW, need_clover=True)
    Dx = orc.apply(K_CLOVER, x, gamma5=False)
    return np.linalg.norm((b - Dx).ravel()) / np.linalg.norm(b.ravel())


def make_server(fields, **kw):
    srv = lym.Server(**kw)
    ids = [srv.add_lattice(dims, U, csw=CSW) for dims, U in fields]
    return srv, ids


def test_alone_equals_mix_and_tolerances(fields):
    rng = np.random.default_rng(7)
    srv, ids = make_server(fields, workers=3)
    V0 = int(np.prod(LATTICES[0]))
    probe = lym.gaussian_spinor(V0, 4242)
    info, alone = srv.wait(srv.submit(ids[0], 0.125, 1e-10, probe))
    assert info.status == lym.REQ_DONE and info.solve.converged and info.batch_size == 1
    assert info.latency_seconds >= info.service_seconds > 0
    # the mix: random lattices, kappas (some equal to the probe's), tolerances
    reqs = []
    for k in range(24):
        li = int(rng.integers(len(LATTICES)))
        V = int(np.prod(LATTICES[li]))
        kappa = float(rng.choice([0.12, 0.125, 0.13]))
        tol = float(rng.choice([1e-8, 1e-10, 1e-12]))
        b = lym.gaussian_spinor(V, 9000 + k)
        reqs.append((li, kappa, tol, b, srv.submit(ids[li], kappa, tol, b)))
        if k == 5:
            rid_probe = srv.submit(ids[0], 0.125, 1e-10, probe)
    batched = 0
    for li, kappa, tol, b, rid in reqs:
        info, x = srv.wait(rid)
        assert info.status == lym.REQ_DONE and info.solve.converged
        batched += info.batch_size > 1
        dims, U = fields[li]
        res = oracle_residual(dims, U, kappa, x, b)
        print("lattice %s kappa %.3f tol %.0e: residual %.2e, %d steps, batch %d, latency %.1f ms"
              % (dims, kappa, tol, res, info.solve.steps, info.batch_size, 1e3 * info.latency_seconds))
        assert res <= tol
    info, mixed = srv.wait(rid_probe)
    assert info.status == lym.REQ_DONE
    assert np.array_equal(mixed, alone)
    print("requests batched with others: %d of %d" % (batched, len(reqs)))
    srv.drain()
    srv.close()


def test_admission_releases_idle_state(fields):
    """A budget that holds about one lattice's state: requests alternating
    between lattices still complete, and the charge never exceeds the budget."""
    probe = lym.Server(workers=1, max_batch=2)
    lid = probe.add_lattice(*fields[1], csw=CSW)
    info, _ = probe.wait(probe.submit(lid, 0.125, 1e-8, lym.gaussian_spinor(int(np.prod(LATTICES[1])), 1)))
    _, one = probe.memory()
    probe.close()
    assert one > 0
    srv, ids = make_server(fields, workers=2, hbm_budget=int(one * 1.5), max_batch=2)
    rids = []
    for k in range(8):
        li = k % len(LATTICES)
        rids.append((li, srv.submit(ids[li], 0.125, 1e-8, lym.gaussian_spinor(int(np.prod(LATTICES[li])), k))))
    for li, rid in rids:
        info, x = srv.wait(rid)
        assert info.status == lym.REQ_DONE and info.solve.converged
        budget, used = srv.memory()
        assert used <= budget
    srv.close()
    # a budget below one lattice's state is refused per request, with a code,
    # also when several workers ask for state at the same time
    tiny, ids = make_server(fields[:1], workers=2, hbm_budget=1 << 20, max_batch=1)
    rids = [tiny.submit(ids[0], kappa, 1e-8, lym.gaussian_spinor(int(np.prod(LATTICES[0])), 3))
            for kappa in (0.125, 0.126, 0.127)]
    for rid in rids:
        with pytest.raises(lym.LymError):
            tiny.wait(rid, timeout=60.0)
    tiny.close()


def test_bad_requests_are_refused(fields):
    srv, ids = make_server(fields[:1], workers=1)
    b = lym.gaussian_spinor(int(np.prod(LATTICES[0])), 5)
    with pytest.raises(lym.LymError):
        srv.submit(7, 0.125, 1e-8, b)
    with pytest.raises(lym.LymError):
        srv.submit(ids[0], 0.125, -1.0, b)
    with pytest.raises(lym.LymError):
        lym.Server(workers=0)
    srv.close()
