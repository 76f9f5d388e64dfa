"""Serving loop: requests on several device-resident lattices, solved by
worker threads with batching of requests that share an operator.  A request's
result is the same alone and in a mix; every result meets its tolerance
against the oracle's D; HBM admission releases idle state and refuses what
cannot fit."""
import numpy as np
import pytest

from helpers import K_CLOVER, K_WILSON, OracleField, fold_bc

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


def test_wilson_lattice_with_two_workers():
    """A csw = 0 lattice served by two workers at once: the even-odd driver
    needs the (unit) diagonal blocks, built once when the lattice is added,
    never lazily by concurrent workers.  Every result is checked against the
    oracle's D (Wilson)."""
    dims = (4, 4, 4, 8)
    U = lym.gen_su3_links(dims, 0.5, 610)
    srv = lym.Server(workers=2, max_batch=2)
    lid = srv.add_lattice(dims, U, csw=0.0)
    V = int(np.prod(dims))
    reqs = []
    for k in range(8):
        b = lym.gaussian_spinor(V, 620 + k)
        kappa = 0.120 + 0.002 * k
        reqs.append((kappa, b, srv.submit(lid, kappa, 1e-10, b)))
    for kappa, b, rid in reqs:
        info, x = srv.wait(rid)
        assert info.status == lym.REQ_DONE and info.solve.converged
        orc = OracleField(dims, fold_bc(dims, U, -1), kappa=kappa)
        Dx = orc.apply(K_WILSON, x, gamma5=False)  # the clover operator at csw = 0
        assert np.linalg.norm((b - Dx).ravel()) / np.linalg.norm(b.ravel()) <= 1e-10
    srv.close()


# The serving workload at its own sizes: lattices {8^4, 12^4, 16^3x32,
# 24^3x48} of SU(3) clover (csw 1.0, links exp(i eps H) at eps 0.5,
# antiperiodic t), kappa ~ U[0.120, 0.134], tolerance in {1e-8, 1e-10,
# 1e-12}, Gaussian sources, Poisson arrivals at 50 requests/s so that
# requests overlap on the device.
C4_LATTICES = [(8, 8, 8, 8), (12, 12, 12, 12), (16, 16, 16, 32), (24, 24, 24, 48)]


def test_serving_workload_at_full_size():
    import time
    rng = np.random.default_rng(2024)
    links = [lym.gen_su3_links(d, 0.5, 4000 + i) for i, d in enumerate(C4_LATTICES)]
    srv = lym.Server(workers=4, max_batch=4)
    ids = [srv.add_lattice(d, U, csw=CSW) for d, U in zip(C4_LATTICES, links)]
    nreq = 32
    spec = [(int(rng.integers(len(C4_LATTICES))), float(rng.uniform(0.120, 0.134)),
             float(rng.choice([1e-8, 1e-10, 1e-12]))) for _ in range(nreq)]
    spec[0] = (3, spec[0][1], 1e-12)  # at least one request per lattice at the tightest tolerance
    spec[1] = (2, spec[1][1], 1e-12)
    arrivals = np.cumsum(rng.exponential(1 / 50.0, nreq))
    srcs = [lym.gaussian_spinor(int(np.prod(C4_LATTICES[li])), 7000 + q) for q, (li, _, _) in enumerate(spec)]
    t0 = time.perf_counter()
    rids = []
    for q, ((li, kappa, tol), ta) in enumerate(zip(spec, arrivals)):
        dt = t0 + ta - time.perf_counter()
        if dt > 0:
            time.sleep(dt)
        rids.append(srv.submit(ids[li], kappa, tol, srcs[q]))
    results = []
    for q, rid in enumerate(rids):
        info, x = srv.wait(rid)
        assert info.status == lym.REQ_DONE and info.solve.converged, q
        results.append((info, x))
    print("batch sizes:", [int(i.batch_size) for i, _ in results], flush=True)
    # every result meets its tolerance: the oracle's D on the three smaller
    # lattices, the FP64 device D (per-site checked against the oracle at this
    # launch shape elsewhere) on 24^3x48
    big = {}
    for q, ((li, kappa, tol), (info, x)) in enumerate(zip(spec, results)):
        dims, b = C4_LATTICES[li], srcs[q]
        if li < 3:
            orc = OracleField(dims, fold_bc(dims, links[li], -1), kappa=kappa, csw=CSW, need_clover=True,
                              need_inverse=False)
            Dx = orc.apply(K_CLOVER, x, gamma5=False)
        else:
            if not big:
                big["ctx"] = lym.Context(dims, lym.REP_SU3_FUND, -1)
                big["g"] = lym.Gauge(big["ctx"], fold_bc(dims, links[li], -1), folded=True)
            D = lym.Operator(big["ctx"], lym.OP_CLOVER, 1, kappa, CSW, False, gauge=big["g"])
            out = lym.Spinor(big["ctx"])
            D.multiply(out, lym.Spinor(big["ctx"], x))
            Dx = out.download()
        res = np.linalg.norm((b - Dx).ravel()) / np.linalg.norm(b.ravel())
        print("request %d: lattice %s kappa %.4f tol %.0e: residual %.2e, %d steps, batch %d, latency %.1f ms"
              % (q, dims, kappa, tol, res, info.solve.steps, info.batch_size, 1e3 * info.latency_seconds),
              flush=True)
        assert res <= tol
    # the same requests alone give the same bits
    srv.drain()
    for q in (0, 1, 2, 5):
        li, kappa, tol = spec[q]
        info, alone = srv.wait(srv.submit(ids[li], kappa, tol, srcs[q]))
        assert info.status == lym.REQ_DONE
        assert np.array_equal(alone, results[q][1]), q
    srv.close()
