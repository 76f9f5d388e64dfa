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


def _oracle_check(dims, U, kappa, x, b, tol):
    orc = OracleField(dims, fold_bc(dims, U, -1), kappa=kappa, csw=CSW, need_clover=True, need_inverse=False)
    Dx = orc.apply(K_CLOVER, x, gamma5=False)
    res = np.linalg.norm((b - Dx).ravel()) / np.linalg.norm(b.ravel())
    assert res <= tol, (kappa, res, tol)
    return res


def test_continuous_batching_mixed_kappa_and_refill():
    """One worker, lanes for 3 requests: 9 requests of different kappa and
    tolerance on one lattice, submitted at once.  They share batches (a lane
    freed by a request that stops early takes the next queued one, at any
    kappa), every result meets its tolerance against the oracle, and each
    equals the same request solved alone, bit for bit."""
    dims = (8, 8, 8, 8)
    U = lym.gen_su3_links(dims, 0.5, 650)
    V = int(np.prod(dims))
    srv = lym.Server(workers=1, max_batch=3)
    lid = srv.add_lattice(dims, U, csw=CSW)
    spec = [(0.120 + 0.0015 * k, [1e-6, 1e-12, 1e-9][k % 3], lym.gaussian_spinor(V, 660 + k)) for k in range(9)]
    rids = [srv.submit(lid, kappa, tol, b) for kappa, tol, b in spec]
    mixed = [srv.wait(r) for r in rids]
    sizes = [int(i.batch_size) for i, _ in mixed]
    print("batch sizes:", sizes, "steps:", [int(i.solve.steps) for i, _ in mixed])
    assert max(sizes) == 3 and min(sizes) >= 2
    for (kappa, tol, b), (info, x) in zip(spec, mixed):
        assert info.status == lym.REQ_DONE and info.solve.converged
        _oracle_check(dims, U, kappa, x, b, tol)
    for (kappa, tol, b), (info, x) in zip(spec, mixed):
        ia, alone = srv.wait(srv.submit(lid, kappa, tol, b))
        assert ia.batch_size == 1
        assert np.array_equal(alone, x)
    srv.close()


def test_request_joins_a_running_batch():
    """With one worker, a request that arrives while another one is being
    solved on the same lattice joins its batch at the next chunk boundary
    (it does not wait for it to finish) and both results are unchanged."""
    import time
    dims = (16, 16, 16, 32)
    U = lym.gen_su3_links(dims, 0.5, 670)
    V = int(np.prod(dims))
    srv = lym.Server(workers=1, max_batch=4)
    lid = srv.add_lattice(dims, U, csw=CSW)
    b1, b2 = lym.gaussian_spinor(V, 671), lym.gaussian_spinor(V, 672)
    srv.wait(srv.submit(lid, 0.125, 1e-8, b1))  # warm: state and captured chunks
    r1 = srv.submit(lid, 0.133, 1e-12, b1)
    time.sleep(0.004)
    t2 = time.perf_counter()
    r2 = srv.submit(lid, 0.121, 1e-10, b2)
    i1, x1 = srv.wait(r1)
    i2, x2 = srv.wait(r2)
    print("first: %d steps, %.1f ms, batch %d; second: queued %.2f ms, batch %d"
          % (i1.solve.steps, 1e3 * i1.latency_seconds, i1.batch_size, 1e3 * i2.queue_seconds, i2.batch_size))
    assert i1.solve.converged and i2.solve.converged
    # the second arrived while the first was in service: it joined, not queued behind it
    assert i2.batch_size == 2 and i1.batch_size == 2
    assert i2.queue_seconds < 0.5 * i1.service_seconds
    _oracle_check(dims, U, 0.133, x1, b1, 1e-12)
    _oracle_check(dims, U, 0.121, x2, b2, 1e-10)
    _, a1 = srv.wait(srv.submit(lid, 0.133, 1e-12, b1))
    _, a2 = srv.wait(srv.submit(lid, 0.121, 1e-10, b2))
    assert np.array_equal(a1, x1) and np.array_equal(a2, x2)
    assert t2 > 0
    srv.close()
