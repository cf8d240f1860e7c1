"""The oracle's symmetric surface matching (surfaceMatching.py with symmetric=True), pinned on
the CPU: its gradient, taken in the driver's own dot product, is the derivative of its
objective along seeded directions that move the momenta, the template's points and the affine
coefficients (central finite differences), and the optimiser lowers the objective."""
import numpy as np
import pytest

from oracle import lddmm_oracle as orc
from registration_amd import synthetic


def problem(affine='none'):
    V, F = synthetic.icosphere(1)
    rng = np.random.default_rng(9)
    W = V * np.array([1.15, 0.9, 1.0]) + 0.02 * rng.standard_normal(V.shape)
    return orc.SymmetricSurfaceMatching(orc.Surface(FV=(F, V)), orc.Surface(FV=(F, W)),
                                        orc.Kernel('gauss', sigma=0.6), orc.Kernel('gauss', sigma=0.4),
                                        sigmaError=0.2, timeStep=0.25, affine=affine)


@pytest.mark.parametrize("affine", ["none", "euclidean"])
def test_symmetric_gradient_is_the_derivative_of_the_objective(affine):
    m = problem(affine)
    rng = np.random.default_rng(4)
    m.at = 0.05 * rng.standard_normal(m.at.shape)
    m.x0 = m.x0 + 0.01 * rng.standard_normal(m.x0.shape)
    if m.affineDim:
        m.Afft = 0.05 * rng.standard_normal(m.Afft.shape)
    m.obj = None
    m.objectiveFun()
    m.fvInit.updateVertices(m.x0)
    grd = m.getGradient(1.0)
    for part in ("diff", "initx", "aff"):
        if part == "aff" and not m.affineDim:
            continue
        d = orc.Direction()
        d.diff = np.zeros(m.at.shape)
        d.initx = np.zeros(m.x0.shape)
        d.aff = np.zeros(m.Afft.shape)
        setattr(d, part, rng.standard_normal(getattr(d, part).shape))
        h = 1e-5
        fd = (m.updateTry(d, -h) - m.updateTry(d, h)) / (2 * h)
        [an] = m.dotProduct(grd, [d])
        assert abs(fd - an) <= 1e-6 * abs(an), (affine, part, fd, an)


def test_symmetric_optimiser_moves_the_template_and_lowers_the_objective():
    m = problem()
    m.maxIter = 5
    hist = m.optimizeMatching()
    assert len(hist) >= 3 and hist[-1] < hist[0]
    assert np.abs(m.x0 - m.fv0.vertices).max() > 0
