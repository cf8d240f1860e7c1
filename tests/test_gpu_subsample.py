"""SurfaceMatching(subsampleTargetSize=n) (surfaceMatching.py:131-133): the registered
template is the decimated copy of the template (Surface.Simplify); the FP64 CG sequence of
the GPU driver equals the oracle's run on that decimated template to 1e-9.  The template is
given with inward-facing triangles (negative volume) so that neither the decimation's
orientation rule nor the target flip applies and both drivers see the same surfaces."""
import numpy as np
import pytest

from oracle import lddmm_oracle as orc

pytestmark = pytest.mark.gpu


def test_subsampled_template_matches_oracle():
    from registration_amd import kernelFunctions as kfun, surfaceMatching as sm, surfaces, synthetic
    V, F = synthetic.icosphere(4)
    F = F[:, [0, 2, 1]]
    Y = synthetic.bump_target(V)
    p = sm.SurfaceMatchingParam(timeStep=0.2, KparDiff=kfun.Kernel('gauss', sigma=0.5, precision='fp64'),
                                KparDist=kfun.Kernel('gauss', sigma=0.3, precision='fp64'),
                                sigmaError=0.1, errorType='current')
    g = sm.SurfaceMatching(outputDir=None, Template=surfaces.Surface(FV=(F, V)),
                           Target=surfaces.Surface(FV=(F, Y)), param=p, testGradient=False,
                           maxIter=6, subsampleTargetSize=642)
    assert g.npt == 642 and g.fvInit.vertices.shape[0] == 642
    assert g.fv0.vertices.shape[0] == V.shape[0]          # the full template is kept
    assert g.fv0ori == -1 and g.fv1.surfVolume() < 0
    Vd, Fd = g.fvInit.vertices.copy(), g.fvInit.faces.copy()
    o = orc.SurfaceMatching(orc.Surface(FV=(Fd, Vd)), orc.Surface(FV=(F, Y)),
                            orc.Kernel('gauss', sigma=0.5), orc.Kernel('gauss', sigma=0.3),
                            sigmaError=0.1, timeStep=0.2, maxIter=6)
    print("subsampled template: 2562 -> 642 vertices", flush=True)
    g.optimizeMatching()
    o.optimizeMatching()
    hg, ho = np.array(g.history), np.array(o.history)
    assert len(hg) == len(ho) and len(hg) >= 4, (hg, ho)
    assert np.all(np.abs(hg - ho) <= 1e-9 * np.abs(ho)), (hg, ho)
