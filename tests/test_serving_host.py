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


class _FakeOpt:
    """A registration stand-in: its optimiser asks for one objective and stops."""

    def __init__(self, obj):
        self.obj = obj
        self.history = [obj]
        self.stop_reason = 'small gradient'
        self.at = None
        self.affineDim = 0

    def cg_steps(self):
        yield (serving.cg.OBJECTIVE,)


def _active(srv, obj, precision):
    a = serving._Active.__new__(serving._Active)
    a.req = _request(7)
    a.admitted = 0.0
    a.twin = None
    a.err32 = float('nan')
    a.start(_FakeOpt(obj), precision)
    srv.used += serving.footprint_bytes(a.req)
    return a


def test_precision_guard_rule(monkeypatch):
    """The guard's decisions without a device: within certify_tol the FP32-compensated result
    is kept with its error and the FP64 energy; beyond it the registration restarts in FP64
    (a new FP64 registration, its optimiser from the start); an FP64 service, or
    certify_tol=None, never certifies."""
    t = [0.0]
    srv = serving.Server(clock=lambda: t[0], precision='fp32c')
    srv.t0 = 0.0
    assert srv.certify_tol == serving.DEFAULT_CERTIFY_TOL == 5e-7
    built = []
    monkeypatch.setattr(serving, 'build_registration',
                        lambda req, prec: built.append(prec) or _FakeOpt(2.0))
    monkeypatch.setattr(serving, 'fp64_twin', lambda req, opt: 'twin')
    a = _active(srv, 1.0, 'fp32c')
    assert srv._stopped(a) is False and a.pending == (serving.CERTIFY,) and a.twin == 'twin'
    assert srv._certified(a, 1.0 + 4e-7) is True
    r = srv.results[7]
    assert r.precision == 'fp32c' and r.energy == 1.0 and r.energy_fp64 == 1.0 + 4e-7
    assert abs(r.fp32c_error - 4e-7 / (1 + 4e-7)) < 1e-15 and srv.promoted == 0 and built == []
    b = _active(srv, 1.0, 'fp32c')
    srv._stopped(b)
    assert srv._certified(b, 1.0 + 1e-5) is False
    assert srv.promoted == 1 and built == ['fp64'] and b.precision == 'fp64'
    assert b.pending == (serving.cg.OBJECTIVE,) and b.opt.obj == 2.0 and b.err32 > 5e-7
    assert srv._stopped(b) is True                       # an FP64 run finishes without a check
    r = srv.results[7]
    assert r.precision == 'fp64' and r.energy == r.energy_fp64 == 2.0 and r.fp32c_error > 5e-7
    assert serving.Server(precision='fp64').certify_tol is None
    off = serving.Server(clock=lambda: t[0], precision='fp32c', certify_tol=None)
    off.t0 = 0.0
    c = _active(off, 1.0, 'fp32c')
    assert off._stopped(c) is True and off.results[7].precision == 'fp32c'
