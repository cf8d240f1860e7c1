"""Host-side checks of the service and of input validation (no device work: every case is
rejected before anything is uploaded)."""
import numpy as np
import pytest

from registration_amd import engine, serving, synthetic


def _request(rid, kind='surface', faces=None, vertices=None):
    V, F = synthetic.icosphere(1)
    if faces is not None:
        F = faces
    if vertices is not None:
        V = vertices
    return serving.Request(rid=rid, kind=kind, template=(F, V), target=(F, V), sigmaKernel=0.3,
                           sigmaDist=0.2, sigmaError=0.1, T=10, maxIter=5)


def test_check_faces_rejects_out_of_range_and_bad_shapes():
    V, F = synthetic.icosphere(1)
    engine.check_faces(F, len(V))
    engine.check_faces(np.zeros((0, 3), np.int64), 0)
    with pytest.raises(ValueError, match=r"\[0, 42\)"):
        engine.check_faces(np.where(F == 5, -1, F), len(V))
    with pytest.raises(ValueError, match=r"\[0, 42\)"):
        engine.check_faces(np.where(F == 5, len(V), F), len(V))
    with pytest.raises(ValueError, match="integer"):
        engine.check_faces(F.astype(np.float64), len(V))
    with pytest.raises(ValueError, match="integer"):
        engine.check_faces(F[:, :1], len(V))
    with pytest.raises(ValueError, match="integer"):
        engine.check_faces(F.reshape(-1), len(V))


def test_server_rejects_invalid_requests_and_keeps_serving():
    t = [0.0]
    srv = serving.Server(clock=lambda: t[0])
    srv.t0 = 0.0
    bad_kind = _request(0, kind='tetmesh')
    srv.submit(bad_kind)
    srv._admit()
    r = srv.results[0]
    assert r.stop_reason.startswith("rejected: ValueError") and "tetmesh" in r.stop_reason
    assert np.isnan(r.energy) and r.iterations == 0
    assert not srv.active and not srv.queue and srv.used == 0.0
    # A request that arrives later waits in the queue; rejection does not drop it.
    later = _request(1, kind='tetmesh')
    later.arrival = 1.0
    srv.submit(later)
    srv._admit()
    assert 1 not in srv.results and len(srv.queue) == 1
    t[0] = 2.0
    srv._admit()
    assert srv.results[1].stop_reason.startswith("rejected") and srv.results[1].admitted == 2.0


def test_submit_rejects_requests_above_the_admission_cap():
    srv = serving.Server(hbm_cap_bytes=1e5)
    with pytest.raises(ValueError, match="admission cap"):
        srv.submit(_request(0))
