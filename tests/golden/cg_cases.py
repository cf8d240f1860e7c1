"""Seeded CPU objectives that drive a conjugate-gradient optimiser through every branch of
py-lddmm's conjugateGradient.cg (py-lddmm/conjugateGradient.py:33-217), and the recorder that
turns one run into a call trace.

Used by tests/golden/make_cg_golden.py (which runs the reference's own cg on them and writes
tests/golden/cg_trace.npz) and by tests/test_cg_golden.py (which runs the oracle's and the
product's cg on the same objects and compares the traces exactly).

A trace is the sequence of calls the optimiser makes on the object, each as a row
(code, a, b, c) of float64:
  OBJ   objectiveFun()                       a = returned value
  GRAD  getGradient(coeff)                   a = coeff
  DOT   dotProduct(g1, g2)                   a = len(g2); the results follow as DOTV rows
  DOTV  one dotProduct result                a = value
  TRY   updateTry(dir, eps, objRef)          a = eps, b = objRef (NaN if None), c = result
  ACC   acceptVarTry()
  ENDIT endOfIteration()
  START startOptim()   SITER startOfIteration()   ENDOPT endOptim()   RDIR randomDir()
  ADD   addProd(d1, d2, beta)                a = beta
plus the log lines the optimiser wrote (logging.info on the root logger).
"""
from __future__ import annotations

import logging

import numpy as np

OBJ, GRAD, DOT, DOTV, TRY, ACC, ENDIT, START, SITER, ENDOPT, RDIR, ADD = range(12)


class Problem:
    """f(v) = 1/2 v'Av - b'v + noise * sum(cos(3v)), scaled by ``scale``; every protocol
    method is recorded.  ``lock`` lists getGradient call numbers (0-based) after which every
    line-search trial returns obj + 1 until the next getGradient call: an object whose
    search direction is uphill at every step length, which is how the optimiser's
    bad-direction branches are reached."""

    def __init__(self, n=6, seed=0, noise=0.0, scale=1.0, cond=None, lock=(), flip=(),
                 hooks=False):
        rng = np.random.default_rng(seed)
        if cond is None:
            M = rng.standard_normal((n, n))
            self.A = M @ M.T + n * np.eye(n)
        else:
            Q, _ = np.linalg.qr(rng.standard_normal((n, n)))
            self.A = Q @ np.diag(np.geomspace(1.0, cond, n)) @ Q.T
        self.b = rng.standard_normal(n)
        self.noise, self.scale = noise, scale
        self.v = np.zeros(n)
        self.vTry = self.v
        self.objTry = None
        self.lock, self.flip = set(lock), set(flip)
        self.ngrad = 0
        self.locked = False
        self.trace = []
        self.rng = np.random.default_rng(seed + 1000)
        if hooks:
            self.startOptim = lambda: self.trace.append((START, 0., 0., 0.))
            self.startOfIteration = lambda: self.trace.append((SITER, 0., 0., 0.))
            self.endOptim = lambda: self.trace.append((ENDOPT, 0., 0., 0.))

    def f(self, v):
        return self.scale * (0.5 * v @ self.A @ v - self.b @ v + self.noise * np.sum(np.cos(3 * v)))

    def getVariable(self):
        return self.v

    def objectiveFun(self):
        o = self.f(self.v)
        self.trace.append((OBJ, o, 0., 0.))
        return o

    def updateTry(self, d, eps, objRef=None):
        vt = self.v - eps * d
        o = self.f(self.v) + 1.0 if self.locked else self.f(vt)
        self.trace.append((TRY, eps, np.nan if objRef is None else objRef, o))
        if objRef is None or o < objRef:
            self.vTry, self.objTry = vt, o
        return o

    def acceptVarTry(self):
        self.trace.append((ACC, 0., 0., 0.))
        self.v = self.vTry

    def getGradient(self, coeff=1.0):
        self.trace.append((GRAD, coeff, 0., 0.))
        self.locked = self.ngrad in self.lock
        g = self.scale * (self.A @ self.v - self.b - 3 * self.noise * np.sin(3 * self.v)) / coeff
        if self.ngrad in self.flip:
            g = -g
        self.ngrad += 1
        return g

    def dotProduct(self, g1, g2):
        r = [float(np.dot(g1, h)) for h in g2]
        self.trace.append((DOT, len(g2), 0., 0.))
        self.trace.extend((DOTV, x, 0., 0.) for x in r)
        return r

    def copyDir(self, d):
        return np.copy(d)

    def addProd(self, a, b, beta):
        self.trace.append((ADD, beta, 0., 0.))
        return a + beta * b

    def endOfIteration(self):
        self.trace.append((ENDIT, 0., 0., 0.))

    def randomDir(self):
        self.trace.append((RDIR, 0., 0., 0.))
        return self.rng.standard_normal(self.v.shape)


class BareProblem(Problem):
    """The same objective without dotProduct / copyDir / addProd / endOfIteration / randomDir:
    the optimiser's plain-array fallbacks (np.multiply(...).sum(), np.copy, g + beta * d)."""

    dotProduct = copyDir = addProd = endOfIteration = randomDir = None

    def __getattribute__(self, name):
        if name in ("dotProduct", "copyDir", "addProd", "endOfIteration", "randomDir"):
            raise AttributeError(name)
        return object.__getattribute__(self, name)


# (name, factory, attributes set on the object, cg keyword arguments)
CASES = [
    # small-gradient stop; the first trial overshoots (epsInit 10), so the epsSmall probe and
    # the halving loop run; hooks present
    ("small_gradient", lambda: Problem(seed=1, hooks=True), dict(gradEps=0.05),
     dict(verb=True, maxIter=50, epsInit=10.)),
    # maxIter stop on a non-quadratic objective, PR+ restarts every 4 iterations and a
    # re-initialised direction at cgBurnIn
    ("max_iter_restart", lambda: Problem(seed=2, noise=0.7, scale=100.),
     dict(gradEps=1e-12, restartRate=4, cgBurnIn=6, gradCoeff=3.0),
     dict(verb=True, maxIter=10, epsInit=1.)),
    # gradient pointing uphill at the first iteration: bad-direction stop (skipCG == 1)
    ("bad_direction_stop", lambda: Problem(seed=3, flip=(0,)), dict(gradEps=1e-9),
     dict(verb=True, maxIter=20, epsInit=0.1)),
    # uphill trials at iteration 3 with CG active, verb=True: CG disabled, no update; then
    # iteration 4 recovers with steepest descent
    ("bad_direction_disable", lambda: Problem(seed=4, noise=0.3, lock=(3,)), dict(gradEps=1e-9),
     dict(verb=True, maxIter=9, epsInit=1.)),
    # the same with verb=False: the reference then accepts the increased objective
    ("bad_direction_verb_false", lambda: Problem(seed=4, noise=0.3, lock=(3,)), dict(gradEps=1e-9),
     dict(verb=False, maxIter=9, epsInit=1.)),
    # uphill at iterations 3 and 4: CG disabled at 3, then the bad-direction stop at 4
    ("bad_direction_twice", lambda: Problem(seed=4, noise=0.3, lock=(3, 4)), dict(gradEps=1e-9),
     dict(verb=True, maxIter=9, epsInit=1.)),
    # a shallow objective: variation below 1e-6 with CG active (CG disabled, eps = 1), then the
    # small-variation stop
    ("small_variation", lambda: Problem(seed=5, scale=1e-5, cond=30.0, n=4), dict(gradEps=1e-14),
     dict(verb=True, maxIter=40, epsInit=1.)),
    # the gradient test (randomDir) on every iteration
    ("test_gradient", lambda: Problem(seed=6, noise=0.2), dict(gradEps=1e-6),
     dict(verb=True, maxIter=6, TestGradient=True, epsInit=1.)),
    # plain arrays: no dotProduct / copyDir / addProd / endOfIteration
    ("bare_arrays", lambda: BareProblem(seed=7, noise=0.4), dict(gradEps=1e-6),
     dict(verb=True, maxIter=12, epsInit=1.)),
]


class _Capture(logging.Handler):
    def __init__(self):
        super().__init__(level=logging.INFO)
        self.lines = []

    def emit(self, record):
        self.lines.append(record.getMessage())


def run(cgfun, case):
    """Run ``cgfun`` on a fresh object of ``case``; return (trace (n, 4) float64, log lines,
    final variable, converged flag)."""
    name, factory, attrs, kwargs = case
    opt = factory()
    for k, v in attrs.items():
        setattr(opt, k, v)
    root = logging.getLogger()
    cap = _Capture()
    old = root.level
    root.addHandler(cap)
    root.setLevel(logging.INFO)
    try:
        cgfun(opt, **kwargs)
    finally:
        root.removeHandler(cap)
        root.setLevel(old)
    return (np.array(opt.trace, dtype=np.float64).reshape(-1, 4), cap.lines, np.copy(opt.v),
            bool(getattr(opt, "converged", False)))


def stop_reason(lines):
    for s in lines:
        if s.startswith("Stopping Gradient Descent: "):
            return s[len("Stopping Gradient Descent: "):]
    return "maxIter"
