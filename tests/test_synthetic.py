"""Workload generators (host NumPy): the shapes and parameter ranges the benchmark streams
are defined with."""
import collections

import numpy as np

from registration_amd import synthetic


def test_icosphere_counts_and_orientation():
    for s, (nv, nf) in {3: (642, 1280), 6: (40962, 81920)}.items():
        V, F = synthetic.icosphere(s)
        assert V.shape == (nv, 3) and F.shape == (nf, 3)
        np.testing.assert_allclose(np.linalg.norm(V, axis=1), 1.0, rtol=1e-12)
        c = V[F].mean(axis=1)
        n = np.cross(V[F[:, 1]] - V[F[:, 0]], V[F[:, 2]] - V[F[:, 0]])
        assert np.all((c * n).sum(axis=1) > 0)   # outward faces


def test_mixed_service_stream_composition():
    calls = []

    def shoot(x0, a0, kernel, sigma, T):
        calls.append((x0.shape, kernel, T))
        return x0 + a0

    reqs = synthetic.mixed_service_requests(2000, seed=5, shoot=shoot)
    kinds = collections.Counter(r['kind'] for r in reqs)
    assert abs(kinds['curve'] / 2000 - 0.4) < 0.04 and abs(kinds['surface'] / 2000 - 0.4) < 0.04
    assert abs(kinds['landmark'] / 2000 - 0.2) < 0.03 and len(calls) == kinds['landmark']
    arr = np.array([r['arrival'] for r in reqs])
    assert np.all(np.diff(arr) > 0)
    # Poisson 200/s with 20% of gaps shortened 5x: mean gap (0.8 + 0.2 / 5) / 200
    assert abs(arr[-1] / 2000 - 0.84 / 200) < 0.1 * 0.84 / 200
    for r in reqs:
        assert r['T'] in (10, 20, 40) and 0.1 <= r['sigmaKernel'] <= 0.5
        assert r['typeKernel'] in ('gauss', 'laplacian')
        F, X = r['template']
        if r['kind'] == 'curve':
            assert X.shape[1] == 2 and 256 <= X.shape[0] <= 8192 and F.shape == (X.shape[0], 2)
        elif r['kind'] == 'surface':
            assert X.shape[0] in (642, 2562, 10242, 40962, 163842)
        else:
            assert F is None and X.shape[1] == 3 and 1000 <= X.shape[0] <= 200000
            assert np.all(np.linalg.norm(X, axis=1) <= 1.0)
    sizes = {k: [r['size'] for r in reqs if r['kind'] == k] for k in kinds}
    assert max(sizes['landmark']) > 100000 and min(sizes['landmark']) < 2000
    assert 163842 in sizes['surface'] and 642 in sizes['surface']


def test_landmark_momenta_scale():
    assert abs(synthetic.landmark_momenta_scale(0, 0.3) - 0.1) < 1e-15
    s1 = synthetic.landmark_momenta_scale(200000, 0.5)
    assert abs(s1 - 0.1 / np.sqrt(1 + 200000 * 0.125)) < 1e-15
