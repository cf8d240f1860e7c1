"""The product's conjugate gradient (registration_amd.conjugateGradient, host-side, generator
form) against the oracle's restatement of conjugateGradient.cg, on CPU-only optimisable
objects: the decision path and iterates must be identical."""
import numpy as np
import pytest

from oracle import lddmm_oracle as orc
from registration_amd import conjugateGradient as cgm


class Quadratic:
    def __init__(self, n=8, seed=0, with_dot=True, noise=0.0):
        rng = np.random.default_rng(seed)
        M = rng.standard_normal((n, n))
        self.A = M @ M.T + n * np.eye(n)
        self.b = rng.standard_normal(n)
        self.v = np.zeros(n)
        self.noise = noise
        self.calls = []
        if not with_dot:
            self.dotProduct = None

    def getVariable(self):
        return self.v

    def f(self, v):
        # a mildly non-quadratic objective, so line searches take several branches
        return 0.5 * v @ self.A @ v - self.b @ v + self.noise * np.sum(np.cos(3 * v))

    def objectiveFun(self):
        return self.f(self.v)

    def updateTry(self, d, eps, objRef=None):
        vt = self.v - eps * d
        o = self.f(vt)
        self.calls.append(("try", eps))
        if objRef is None or o < objRef:
            self.vTry, self.objTry = vt, o
        return o

    def acceptVarTry(self):
        self.v = self.vTry

    def getGradient(self, coeff=1.0):
        return (self.A @ self.v - self.b - 3 * self.noise * np.sin(3 * self.v)) / coeff

    def dotProduct(self, g1, g2):
        return np.array([np.dot(g1, h) for h in g2])

    def copyDir(self, d):
        return np.copy(d)

    def addProd(self, a, b, beta):
        return a + beta * b


@pytest.mark.parametrize("noise,eps0", [(0.0, 0.01), (0.5, 1.0), (2.0, 10.0)])
def test_cg_matches_oracle_restatement(noise, eps0):
    a, b = Quadratic(noise=noise), Quadratic(noise=noise)
    a.gradEps = b.gradEps = 1e-8
    cgm.cg(a, verb=False, maxIter=60, epsInit=eps0)
    orc.cg(b, verb=False, maxIter=60, epsInit=eps0)
    assert a.calls == b.calls
    assert np.array_equal(a.v, b.v)
    assert a.stop_reason == b.stop_reason


def test_cg_generator_yields_requests_in_order():
    q = Quadratic()
    gen = cgm.cg_steps(q, maxIter=3, epsInit=0.1)
    req = next(gen)
    kinds = []
    try:
        while True:
            kinds.append(req[0])
            req = gen.send(getattr(q, req[0])(*req[1:]))
    except StopIteration:
        pass
    assert kinds[0] == cgm.OBJECTIVE and kinds[1] == cgm.GRADIENT and kinds[2] == cgm.DOT
    assert cgm.TRY in kinds and cgm.ACCEPT in kinds


def test_cg_stops_on_bad_direction():
    class Bad(Quadratic):
        def getGradient(self, coeff=1.0):  # points uphill
            return -super().getGradient(coeff)
    q = Bad()
    cgm.cg(q, verb=False, maxIter=10, epsInit=0.1)
    assert q.stop_reason == 'bad direction'
