"""Continuous batching: every registration's result is the same alone (plain path) and inside
a served mix of uneven sizes and kinds, and the service admits late arrivals while others run."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def requests():
    from registration_amd import serving, synthetic
    reqs = []
    rng = np.random.default_rng(5)
    for i, s in enumerate((2, 3, 2, 3)):
        V, F = synthetic.icosphere(s)
        Vr = V @ synthetic.random_rotation(rng).T
        reqs.append(serving.Request(rid=i, kind='surface', template=(F, Vr),
                                    target=(F, synthetic.rbf_field(rng, Vr)),
                                    sigmaKernel=float(rng.uniform(0.2, 0.5)), sigmaDist=0.1,
                                    sigmaError=0.1, T=(5 if i % 2 else 4), maxIter=6))
    for i, n in enumerate((128, 300)):
        X, _ = synthetic.circle(n)
        Y = synthetic.star_curve(n, 3 + i, 0.15)
        E = synthetic.closed_curve_edges(n)
        reqs.append(serving.Request(rid=10 + i, kind='curve', template=(E, X), target=(E, Y),
                                    sigmaKernel=0.3, sigmaDist=0.1, sigmaError=0.1, T=5,
                                    maxIter=6, typeKernel='laplacian' if i else 'gauss',
                                    order=1 if i else 10))
    # the measure data term (the reference's default errorType) in the same mix
    V, F = synthetic.icosphere(2)
    reqs.append(serving.Request(rid=30, kind='surface', template=(F, V),
                                target=(F, synthetic.bump_target(V)), sigmaKernel=0.4,
                                sigmaDist=0.2, sigmaError=0.1, T=4, maxIter=6, errorType='measure'))
    reqs.append(serving.Request(rid=32, kind='surface', template=(F, 0.9 * V),
                                target=(F, synthetic.bump_target(V)), sigmaKernel=0.4,
                                sigmaDist=0.25, sigmaError=0.1, T=5, maxIter=6,
                                errorType='varifold'))
    # global affine parts (surface similitude, curve affine) in the same mix
    reqs.append(serving.Request(rid=33, kind='surface', template=(F, V),
                                target=(F, 1.2 * synthetic.bump_target(V) + 0.05), sigmaKernel=0.4,
                                sigmaDist=0.2, sigmaError=0.1, T=4, maxIter=6,
                                affine='similitude', affineWeight=0.5))
    Xa, _ = synthetic.circle(140)
    Ea = synthetic.closed_curve_edges(140)
    reqs.append(serving.Request(rid=34, kind='curve', template=(Ea, Xa),
                                target=(Ea, 1.1 * synthetic.star_curve(140, 3, 0.2) + 0.1),
                                sigmaKernel=0.3, sigmaDist=0.1, sigmaError=0.1, T=5, maxIter=6,
                                affine='affine'))
    X, _ = synthetic.circle(160)
    E = synthetic.closed_curve_edges(160)
    reqs.append(serving.Request(rid=31, kind='curve', template=(E, X),
                                target=(E, synthetic.star_curve(160, 4, 0.2)), sigmaKernel=0.3,
                                sigmaDist=0.1, sigmaError=0.1, T=5, maxIter=6, errorType='measure'))
    for i, n in enumerate((200, 513)):
        X = synthetic.ball_points(rng, n)
        a0 = 0.2 * rng.standard_normal(X.shape)
        Y = synthetic._gpu_shoot(X, a0, 'gauss', 0.3, 5)
        reqs.append(serving.Request(rid=20 + i, kind='landmark', template=(None, X),
                                    target=(None, Y), sigmaKernel=0.3 + 0.1 * i, sigmaDist=0.1,
                                    sigmaError=0.1, T=5 - i, maxIter=6,
                                    typeKernel='laplacian' if i else 'gauss', order=3 if i else 10))
    return reqs


@pytest.mark.parametrize("prec", ["fp32c", "fp64"])
def test_results_identical_alone_and_in_the_mix(prec):
    from registration_amd import serving
    reqs = requests()
    alone = {r.rid: serving.plain_run(r, prec) for r in reqs}
    srv = serving.Server(precision=prec)
    for k, r in enumerate(reqs):
        r.arrival = 0.0 if k < 3 else 0.01 * (k - 2)   # later arrivals join a running batch
        srv.submit(r)
    res = srv.run()
    assert set(res) == set(alone)
    for rid, a in alone.items():
        b = res[rid]
        assert a.history == b.history, (rid, a.history, b.history)
        assert a.energy == b.energy and a.stop_reason == b.stop_reason
    # the service really batched: several registrations answered per round on average
    assert srv.max_batch >= 4 and srv.answered / srv.rounds > 1.5, (srv.max_batch, srv.answered, srv.rounds)


def test_admission_cap_limits_concurrency():
    from registration_amd import serving
    reqs = requests()[:3]
    cap = max(serving.footprint_bytes(r) for r in reqs) * 1.01
    srv = serving.Server(hbm_cap_bytes=cap, precision='fp32c')
    for r in reqs:
        srv.submit(r)
    peak = 0
    srv.t0 = srv.clock()
    while srv.queue or srv.active:
        srv._admit()
        peak = max(peak, len(srv.active))
        if srv.active:
            srv.step()
    assert peak == 1 and len(srv.results) == 3
    with pytest.raises(ValueError):
        serving.Server(hbm_cap_bytes=1000).submit(reqs[0])


def test_invalid_requests_are_rejected_and_the_rest_served():
    """Out-of-range face indices and mismatched landmark shapes fail their own requests at
    admission (before any kernel reads them); the valid requests in the same stream finish
    with the results they have alone."""
    from registration_amd import serving
    reqs = {r.rid: r for r in requests()}
    good = [reqs[0], reqs[10], reqs[20]]
    F, V = reqs[1].template
    # index len(V): the host Surface's own NumPy indexing fails, as the reference's does;
    # index -1 passes host NumPy (it wraps) and is caught before the faces reach the device.
    bad_faces = [serving.Request(rid=rid, kind='surface', template=(np.where(F == 7, bad, F), V),
                                 target=reqs[1].target, sigmaKernel=0.3, sigmaDist=0.1,
                                 sigmaError=0.1, T=4, maxIter=6)
                 for rid, bad in ((90, len(V)), (92, -1))]
    X = reqs[20].template[1]
    bad_shape = serving.Request(rid=91, kind='landmark', template=(None, X),
                                target=(None, X[:-1]), sigmaKernel=0.3, sigmaDist=0.1,
                                sigmaError=0.1, T=4, maxIter=6)
    alone = {r.rid: serving.plain_run(r, 'fp32c') for r in good}
    srv = serving.Server(precision='fp32c')
    for r in (good[0], bad_faces[0], good[1], bad_shape, bad_faces[1], good[2]):
        srv.submit(r)
    res = srv.run()
    assert res[90].stop_reason.startswith("rejected: IndexError") and np.isnan(res[90].energy)
    assert res[91].stop_reason.startswith("rejected: ValueError") and np.isnan(res[91].energy)
    assert res[92].stop_reason.startswith("rejected: ValueError") and "[0, 642)" in res[92].stop_reason
    for rid, a in alone.items():
        assert res[rid].history == a.history and res[rid].energy == a.energy
    assert srv.used == 0.0
