"""The continuous-batching server (filt_serve_*) on the GPU: requests that
arrive while others run are admitted and batched, and each result equals the
same request run alone, bit for bit, in every precision form."""
import importlib

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

pkg = importlib.import_module("shmem-cuda-filtration_amd")
serving = importlib.import_module(pkg.__name__ + ".serving")


def alone(hip, precision, req, nsteps):
    p, P1, S2, perm, poro = req
    e = hip.ensemble(precision)
    try:
        mid = e.add(p, P1, S2, perm, poro)
        e.step(nsteps)
        P, S = e.get(mid)
        return P, S, e.mass(mid)
    finally:
        e.close()


@pytest.mark.parametrize("precision", [0, 1, 2])
def test_results_same_alone_and_in_the_mix(hip, precision):
    stream = serving.request_stream(rate=40.0, duration=0.4, seed=7, edge_lo=16, edge_hi=40,
                               steps_lo=10, steps_hi=60)
    assert len(stream) >= 8
    stats, res = serving.run(hip, precision, stream, workers=4, keep=range(len(stream)),
                             max_round_steps=3)
    assert stats["requests"] == len(stream)
    assert stats["max_active"] >= 2  # requests overlapped on the device
    for q, (_, shape, nsteps, mseed) in enumerate(stream):
        req = serving.make_request(shape, mseed)
        P, S, m = alone(hip, precision, req, nsteps)
        Pm, Sm, mm = res[q]
        assert np.array_equal(P, Pm) and np.array_equal(S, Sm), q
        assert m == mm


def test_admission_respects_slots_and_queue_order(hip):
    """max_active = 2: later requests wait in the queue (FIFO) and still
    finish with their own results."""
    srv = serving.Server(hip, 1, max_active=2, max_round_steps=4)
    try:
        reqs = [serving.make_request((12 + q, 10, 11), 100 + q) for q in range(5)]
        ids = [srv.submit(*r, 5 + q) for q, r in enumerate(reqs)]
        a, _ = srv.round()
        assert a == 2 and srv.status()["queued"] == 3
        for _ in range(40):
            srv.round()
            if all(srv.is_done(i) for i in ids):
                break
        assert all(srv.is_done(i) for i in ids)
        for q, i in enumerate(ids):
            P, S, m = srv.take(i)
            P0, S0, m0 = alone(hip, 1, reqs[q], 5 + q)
            assert np.array_equal(P, P0) and np.array_equal(S, S0) and m == m0
    finally:
        srv.close()


def test_request_that_can_never_fit_is_refused_at_submit(hip):
    """A request larger than the device (with the reserve) would block the
    FIFO queue for good: submit refuses it with FILT_ERR_NOMEM.  A request
    the stability guard refuses is refused at submit too."""
    srv = serving.Server(hip, 1, reserve_bytes=1 << 50, max_active=2)
    try:
        req = serving.make_request((12, 10, 11), 7)
        with pytest.raises(pkg.abi.FiltError) as err:
            srv.submit(*req, 5)
        assert err.value.code == pkg.abi.FILT_ERR_NOMEM
    finally:
        srv.close()
    srv = serving.Server(hip, 1)
    try:
        p, P1, S2, perm, poro = serving.make_request((12, 10, 11), 8)
        p.dt *= 1e6  # far past the explicit scheme's limit
        with pytest.raises(pkg.abi.FiltError) as err:
            srv.submit(p, P1, S2, perm, poro, 5)
        assert err.value.code == pkg.abi.FILT_ERR_UNSTABLE
        assert srv.status()["queued"] == 0
    finally:
        srv.close()
