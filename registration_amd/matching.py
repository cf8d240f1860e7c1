"""Device-resident matching state and the batched implementations of the optimiser requests.

``DeviceMatching`` holds one registration on the GPU: template vertices x0, momenta at (T,N,d),
the trajectory xt of at, the target's currents (c2, cr2), and trial buffers.  The requests
the CG optimiser issues (objective, line-search trial, gradient, RKHS products) are
implemented once, for a *list* of registrations, by ``batch_*`` functions that issue one
segmented launch per kernel-sum stage for the whole list.  A registration's own methods call
them with a list of one, which is the plain single-pair path; the serving layer calls them
with every registration waiting on the same request.

Reference: py-lddmm/surfaceMatching.py:187-484 and curveMatching.py:148-352
(affine='none', symmetric=False, errorType 'current', 'measure' or, for surfaces,
'varifold'), plus a landmark-set data
term sum |x_T - y|^2 / sigmaError^2 over corresponding points
(landmarkMatching.LandmarkMatching).
Data-term energies (currents, measure and varifold norms) are always evaluated in FP64 pairwise
arithmetic; the flow, the covector, the RKHS products and the data-term gradient use the
kernels' precision.  Differences that do not change any result: the trajectory of an accepted trial is kept instead of being recomputed in
endOfIteration / landmarkHamiltonianCovector (the recomputation is deterministic and
bit-identical), and the energy sum reuses the flow's K(x_t) a_t instead of a second applyK.
"""
from __future__ import annotations

import math
from typing import Sequence

import numpy as np
import torch

from . import _lib
from . import engine
from . import runtime as rt
from .affineBasis import AffineBasis


CURRENT = 'current'
MEASURE = 'measure'
VARIFOLD = 'varifold'
LANDMARK = 'landmark'
SHAPE_TERMS = (CURRENT, MEASURE, VARIFOLD)
_DATA_OP = {CURRENT: _lib.OP_CURRENTS, MEASURE: _lib.OP_MEASURE, VARIFOLD: _lib.OP_VARIFOLD}
# operations whose descriptors take a dispatch order (lddmm_item_order)
_ORDERED_OPS = (_lib.OP_FLOW_STEP, _lib.OP_COVECTOR_STEP, _lib.OP_CURRENTS, _lib.OP_MEASURE)


class Direction:
    """Optimisation direction (surfaceMatching.py:52-56): diff is a (T, N, d) device tensor."""

    def __init__(self):
        self.diff = []
        self.aff = []
        self.initx = []


class DeviceMatching:
    """Common device state of SurfaceMatching / CurveMatching / LandmarkMatching."""

    # subclasses set: coef (1/3 surfaces, 1/2 curves), _new_shape(vertices, with_weights) and
    # measure_norm0_eps (the offset inside the target weights of measureNorm0: 1e-10 for
    # surfaces, surfaces.py:1095; none for curves, curves.py:624)
    coef = 1.0 / 3.0
    measure_norm0_eps = 1e-10
    # affine scalings of the gradient and dot products: 'surface' (coeffAff, switched from 1
    # to 100 after affBurnIn iterations; surfaceMatching.py:315-322,367-369,388-389) or 'curve'
    # (affineWeight; curveMatching.py:246-252,288-296)
    aff_style = 'surface'

    def _init_device(self, x0_np, faces_np, target_centers, target_cr, KparDiff, KparDist,
                     sigmaError, Tsize, regweight, target_points=None, data_term=CURRENT,
                     affine='none', affineWeight=1.0, rotWeight=None, scaleWeight=None,
                     transWeight=None, template_geometry=None):
        """Device state.  With ``target_points`` (N, d) the data term is the landmark one,
        sum |x_T - y|^2 / sigmaError^2 over corresponding points, and the shape arguments
        are unused; otherwise it is the currents (data_term 'current'), measure ('measure') or
        varifold ('varifold', surfaces) norm against the target faces (target_centers,
        target_cr).  ``affine`` adds the reference's global affine part A_t x + b_t to the
        velocity, with coefficients Afft (T, affineDim) on an AffineBasis, penalised by
        sum_t dt <affineWeight, Afft_t^2>; Afft and the directions' affine parts are device
        tensors.  ``template_geometry`` (centers, surfels of the template as given) makes the
        registration symmetric: the template's points x0 become a variable and the data term
        adds |S(x0) - S0|^2 / sigmaError^2 (see the template-term methods below)."""
        dev = rt.device()
        self.npt, self.dim = x0_np.shape
        self.Tsize = Tsize
        self.regweight = float(regweight)
        self.sigmaError = float(sigmaError)
        self.Kdiff = engine.DevKernel.of(KparDiff)
        self.x0 = rt.as_dev(x0_np)
        # Processing order of the template's points for the flow and RKHS-product sums
        # (spatially compact tiles; the trajectory keeps the points' own order): host work,
        # chosen on a worker thread while the device state below is set up.
        order_job = _order_pool().submit(engine.order_choice, np.array(x0_np, dtype=np.float64),
                                         self.Kdiff)
        T, N, d = Tsize, self.npt, self.dim
        if target_points is not None:
            self.data_kind = LANDMARK
            self.y = rt.as_dev(target_points)
            if tuple(self.y.shape) != (N, d):
                raise ValueError("landmark target must have the template's shape (N, d)")
            self.Kdist = None
        else:
            if data_term not in SHAPE_TERMS:
                raise NotImplementedError(f"data term {data_term!r} is not part of the GPU path")
            self.data_kind = data_term
            self.Kdist = engine.DevKernel.of(KparDist)
            if self.Kdiff.prec != self.Kdist.prec:
                raise ValueError("deformation and data kernels must share one precision")
            self.faces_np = np.asarray(faces_np)
            self.c2 = rt.as_dev(target_centers)
            self.cr2 = rt.as_dev(target_cr)
            if data_term == VARIFOLD and self.dim != 3:
                raise NotImplementedError("the varifold data term is defined for surfaces")
            if data_term in (MEASURE, VARIFOLD):
                self.a2 = rt.as_dev(engine.measure_weights(target_cr))
                self.a2_norm0 = rt.as_dev(engine.measure_weights(target_cr,
                                                                 self.measure_norm0_eps))
        self.at = torch.zeros((T, N, d), dtype=torch.float64, device=dev)
        self.atTry = torch.zeros_like(self.at)
        # Slice 0 of every trajectory buffer is x0 for the whole optimisation.
        self.xt = self.x0.expand(T + 1, N, d).clone()
        self._xt_try = self.xt.clone()
        self._xt_scratch = self.xt.clone()
        if self.data_kind in SHAPE_TERMS:
            weights = self.data_kind != CURRENT
            self.fvDef_dev = self._new_shape(self.x0, weights)
            self._try_shape = self._new_shape(self.x0, weights)
        else:
            self.fvDef_dev = self._try_shape = None
        # Pointer-free descriptor parts of this registration's launches, built once; the
        # batched requests fill in the buffer addresses of each call with NumPy.
        # Kept as raw bytes (runtime.raw / runtime.stack).
        # Dispatch orders of the large launches (work items with the most unculled pairs
        # first) are estimated from the template and the target on a worker thread from the
        # registration's first line-search trial on, and attached when ready (_count_trial):
        # they only schedule, and a registration that launches a few sums only (the
        # precision guard's FP64 twin) never asks for them.
        self._order_np = order_job.result()
        self._order = (None if self._order_np is None
                       else torch.from_numpy(self._order_np).pin_memory().to(dev, non_blocking=True))
        self._item_orders = []
        self._trials = 0
        self._order_future = None
        self._order_pts = np.array(x0_np, dtype=np.float64)
        self._flow_base = rt.raw(engine.flow_record(N, T, self.Kdiff, perm=self._order))
        self._flow_key = (self.Kdiff.kernel, self.Kdiff.prec, d)
        self._slice_bytes = N * d * 8
        # Gradient buffers, reused by every gradient request: the covector (T+1, N, d) with
        # the end-point covector written straight into its slice T, and the currents
        # gradient outputs.
        self._pxt = torch.empty_like(self.xt)
        self._pxt_tail = self._pxt[1:]
        self._pxT = self._pxt.data_ptr() + T * self._slice_bytes
        self._cov_base = rt.raw(engine.covector_record(N, T, self.Kdiff, self.regweight,
                                                       perm=self._order))
        self._def_stale = False
        if self.data_kind in SHAPE_TERMS:
            Ke = self.Kdist.energy()
            op = _DATA_OP[self.data_kind]
            # the target's points, for the data term's dispatch order
            self._order_tgt = (target_centers.cpu().numpy() if torch.is_tensor(target_centers)
                               else np.array(target_centers, dtype=np.float64))
            self._cur_try = rt.raw(self._data_record(self._try_shape, Ke))
            self._cur_key = (Ke.kernel, Ke.prec, d, op)
            self._geom_try = rt.raw(engine.geometry_record(self._try_shape))
            sh = self.fvDef_dev
            F = sh.centers.shape[0]
            self._z1 = torch.empty((F, d), dtype=torch.float64, device=dev)
            self._dz = torch.empty((F, d), dtype=torch.float64, device=dev)
            self._gtot = torch.empty(1, dtype=torch.float64, device=dev)
            g = self._data_record(sh, self.Kdist, gradient=True)
            g["out0"] = self._z1.data_ptr()
            g["out1"] = self._dz.data_ptr()
            g["total"] = self._gtot.data_ptr()
            # (no dispatch order: a registration runs this launch once per gradient, a few
            # times in all, and its order would cost more host time than it saves)
            self._cur_grad = rt.raw(g)
            self._cur_grad_key = (self.Kdist.kernel, self.Kdist.prec, d, op)
            self._gather = rt.raw(engine.gather_record(sh, self._z1, self._dz,
                                                       -2.0 / (self.sigmaError ** 2), self._pxT))
            self._geom_def = rt.raw(engine.geometry_record(sh))
        self.obj = None
        self.objTry = None
        self.obj0 = None
        self.gradCoeff = self.npt
        self.affine = affine
        self.affB = AffineBasis(self.dim, affine)
        self.affineDim = self.affB.affineDim
        self.affineBasis = self.affB.basis
        self.affineWeight = affineWeight * np.ones([self.affineDim, 1])
        if self.affB.rotComp and rotWeight is not None:
            self.affineWeight[self.affB.rotComp] = rotWeight
        if self.affB.simComp and scaleWeight is not None:
            self.affineWeight[self.affB.simComp] = scaleWeight
        if self.affB.transComp and transWeight is not None:
            self.affineWeight[self.affB.transComp] = transWeight
        self.coeffAff1, self.coeffAff2 = 1., 100.
        self.coeffAff = self.coeffAff1
        self.affBurnIn = 20
        if self.affineDim > 0:
            self.Afft = torch.zeros((T, self.affineDim), dtype=torch.float64, device=dev)
            self._basis = rt.as_dev(self.affineBasis)                   # (d^2 + d, affineDim)
            self._aff_w = rt.as_dev(self.affineWeight.reshape(-1))      # (affineDim,)
        else:
            self.Afft = np.zeros([self.Tsize, 0])
        self.AfftTry = self.Afft
        self.nobj = 0
        self.ngrad = 0
        self.iter = 0
        self.sym = None
        if template_geometry is not None:
            self.sym = _TemplateTerm(self, *template_geometry)

    def _order_launches(self):
        """(descriptor bytes, op, prec) of the launches that take a dispatch order: the flow
        and covector steps and the data-term energy."""
        out = [(self._flow_base, _lib.OP_FLOW_STEP, self.Kdiff.prec),
               (self._cov_base, _lib.OP_COVECTOR_STEP, self.Kdiff.prec)]
        if self.data_kind in SHAPE_TERMS:
            out.append((self._cur_try, _DATA_OP[self.data_kind], self.Kdist.energy().prec))
        return [x for x in out if x[1] in _ORDERED_OPS]

    def _estimate_orders(self, launches):
        """Host estimates of the dispatch orders of ``launches`` [(descriptor copy, op, prec)]
        from the template's points (no device work: it runs on a worker thread)."""
        out = []
        for rec, op, prec in launches:
            if op in (_lib.OP_FLOW_STEP, _lib.OP_COVECTOR_STEP):
                out.append(engine.item_order_host(op, prec, self.dim, rec, self._order_pts,
                                                  perm=self._order_np))
            else:
                c1 = self._order_pts[self.faces_np].mean(axis=1)
                out.append(engine.item_order_host(op, prec, self.dim, rec, c1,
                                                  x1=self._order_tgt))
        return out

    def _count_trial(self):
        """Called for every line-search trial launch.  The first one hands the estimate of
        the dispatch orders (host work, ~10 ms at 81,920 faces) to a worker thread; the first
        trial after it is done attaches them to the cached descriptors.  Orders only steer
        the schedule, so results do not depend on when they arrive."""
        self._trials += 1
        if self._trials == 1:
            launches = [(base, op, prec) for base, op, prec in self._order_launches()
                        if engine.item_count(op, prec, self.dim, base.view(rt.SEG_DT))
                        >= engine.MIN_ORDERED_ITEMS]
            if launches:
                copies = [(np.array(base, copy=True).view(rt.SEG_DT), op, prec)
                          for base, op, prec in launches]
                self._order_future = _order_pool().submit(self._estimate_orders, copies)
                self._order_bases = [base for base, _, _ in launches]
            return
        f = self._order_future
        if f is None or not f.done():
            return
        self._order_future = None
        for base, order in zip(self._order_bases, f.result()):
            o = engine.upload_order(order)
            if o is not None:
                self._item_orders.append(o)
                engine.attach_item_order(base.view(rt.SEG_DT), o)

    def _data_record(self, sh, K, gradient=False, target=None):
        """Descriptor of the data term of device shape ``sh`` against the target (outputs
        unset); ``target`` = (centers, surfels, weights) replaces the registration's target."""
        c2, cr2, a2 = (self.c2, self.cr2, getattr(self, "a2", None)) if target is None else target
        if self.data_kind == CURRENT:
            return engine.currents_record(sh.centers, sh.surfel, c2, cr2, K, self.coef)
        if self.data_kind == VARIFOLD:
            return engine.varifold_record(sh.centers, sh.weights, sh.surfel, c2, a2, cr2, K,
                                          self.coef)
        return engine.measure_record(sh.centers, sh.weights, c2, a2, K, self.coef,
                                     n1=sh.surfel if gradient else None)

    # ----- global affine part -----------------------------------------------------------
    def _aff_transforms(self, Afft):
        """(A (T, d, d), b (T, d)) device tensors of the coefficients Afft (T, affineDim):
        [vec(A_t); b_t] = basis Afft_t (affineBasis.py:67-79)."""
        d = self.dim
        AB = Afft @ self._basis.T
        return AB[:, :d * d].reshape(-1, d, d).contiguous(), AB[:, d * d:].contiguous()

    def _aff_penalty(self, Afft):
        """(T,) device tensor of dt <affineWeight, Afft_t^2> (surfaceMatching.py:214-215)."""
        return (Afft * Afft * self._aff_w).sum(dim=1) / self.Tsize

    def _aff_dot_weight(self):
        """(affineDim,) weights of the affine part of dot products: coeffAff (surfaces) or
        affineWeight (curves)."""
        if self.aff_style == 'curve':
            return self._aff_w
        return torch.full_like(self._aff_w, self.coeffAff)

    # ----- reference protocol, single registration -----------------------------------
    def objectiveFun(self):
        if self.obj is None:
            batch_objective([self])
        return self.obj

    def getVariable(self):
        if self.sym is not None:
            return [self.at, self.Afft, self.x0]
        return [self.at, self.Afft]

    def updateTry(self, dir, eps, objRef=None):
        return batch_update_try([(self, dir, eps, objRef)])[0]

    def getGradient(self, coeff=1.0):
        return batch_gradient([(self, coeff)])[0]

    def dotProduct(self, g1, g2):
        return batch_dot([(self, g1, g2)])[0]

    def acceptVarTry(self):
        self.obj = self.objTry
        self.at = self.atTry
        self.Afft = self.AfftTry
        self.xt, self._xt_try = self._xt_try, self.xt
        # The deformed shape's geometry follows xt[-1]; it is refreshed (batched over
        # registrations) before the next request that reads it.
        if self.fvDef_dev is not None:
            self._def_stale = True
        if self.sym is not None:
            # the kept trial's template points are slice 0 of its trajectory
            self.x0 = self.xt[0].clone()
            self.sym.stale = True

    def endOfIteration(self):
        self.iter += 1
        if self.affineDim > 0 and self.aff_style == 'surface' and self.iter >= self.affBurnIn:
            self.coeffAff = self.coeffAff2

    def addProd(self, dir1, dir2, beta):
        d = Direction()
        d.diff = dir1.diff + beta * dir2.diff
        d.aff = dir1.aff + beta * dir2.aff
        if self.sym is not None:
            d.initx = dir1.initx + beta * dir2.initx
        return d

    def copyDir(self, dir0):
        d = Direction()
        d.diff = dir0.diff.clone()
        d.aff = dir0.aff.clone() if isinstance(dir0.aff, torch.Tensor) else np.copy(dir0.aff)
        if self.sym is not None:
            d.initx = dir0.initx.clone()
        return d

    def randomDir(self):
        d = Direction()
        d.diff = torch.randn((self.Tsize, self.npt, self.dim), dtype=torch.float64,
                             device=self.x0.device)
        d.aff = (torch.randn((self.Tsize, self.affineDim), dtype=torch.float64, device=self.x0.device)
                 if self.affineDim > 0 else np.zeros([self.Tsize, 0]))
        if self.sym is not None:
            d.initx = torch.randn((self.npt, self.dim), dtype=torch.float64, device=self.x0.device)
        return d

    # ----- helpers ---------------------------------------------------------------------
    def deformed_vertices(self) -> torch.Tensor:
        return self.xt[-1]



class _TemplateTerm:
    """Symmetric matching's second data term (surfaceMatching.py:181-185,232-289,297-327,
    353-383): the template's points x0 are a variable, and the objective adds
    |S(x0) - S0|^2 / sigmaError^2, the data-term distance (currents, measure or varifold) from
    the moved template to the template as given, whose norm |S0|^2 / sigmaError^2 joins obj0.
    The gradient in x0 is (d|S(x0) - S0|^2/dx0 / sigmaError^2 - p_0) / (coeffInitx coeff), p_0
    the flow's covector at t = 0, and coeffInitx also weighs the x0 part of the dot product.

    Device state: the moved template S(x0) (geometry refreshed lazily after an accepted
    trial, as the reference refreshes it in endOfIteration), a trial shape S(x0Try), the
    original template's geometry (c0, cr0 and the measure weights), the gradient buffers and
    the descriptors of the energy, gradient and gather launches.  A trial writes its x0Try
    into slice 0 of its trajectory buffer, so the flow starts there and an accepted trial
    carries its template points with it."""

    coeffInitx = 0.1

    def __init__(self, o, centers0, cr0):
        if o.data_kind not in SHAPE_TERMS:
            raise NotImplementedError("symmetric matching needs a shape data term")
        dev = rt.device()
        weights = o.data_kind != CURRENT
        self.c0 = rt.as_dev(centers0)
        self.cr0 = rt.as_dev(cr0)
        self.a0 = self.a0_norm0 = None
        if weights:
            self.a0 = rt.as_dev(engine.measure_weights(cr0))
            self.a0_norm0 = rt.as_dev(engine.measure_weights(cr0, o.measure_norm0_eps))
        target = (self.c0, self.cr0, self.a0)
        self.shape = o._new_shape(o.x0, weights)        # S(x0)
        self.try_shape = o._new_shape(o.x0, weights)    # S(x0Try) of the trial in flight
        self.stale = False
        Ke = o.Kdist.energy()
        self.try_rec = rt.raw(o._data_record(self.try_shape, Ke, target=target))
        self.try_geom = rt.raw(engine.geometry_record(self.try_shape))
        F = self.shape.centers.shape[0]
        N, d = o.npt, o.dim
        self.z1 = torch.empty((F, d), dtype=torch.float64, device=dev)
        self.dz = torch.empty((F, d), dtype=torch.float64, device=dev)
        self.gtot = torch.empty(1, dtype=torch.float64, device=dev)
        self.px0 = torch.empty((N, d), dtype=torch.float64, device=dev)
        g = o._data_record(self.shape, o.Kdist, gradient=True, target=target)
        g["out0"] = self.z1.data_ptr()
        g["out1"] = self.dz.data_ptr()
        g["total"] = self.gtot.data_ptr()
        self.grad_rec = rt.raw(g)
        # +2/sigma^2 times the reference's vertex scatter: d|S(x0) - S0|^2/dx0 / sigma^2
        self.gather = rt.raw(engine.gather_record(self.shape, self.z1, self.dz,
                                                  2.0 / (o.sigmaError ** 2), self.px0.data_ptr()))
        self.geom = rt.raw(engine.geometry_record(self.shape))

    def norm0_job(self, o, total):
        """The energy job of |S0|^2 into ``total`` (a 1-element device view)."""
        Ke = o.Kdist.energy()
        if o.data_kind == CURRENT:
            return engine.KNormJob(self.c0, self.cr0, Ke, total)
        if o.data_kind == MEASURE:
            return engine.MeasureJob(self.c0, self.a0_norm0, None, None, Ke, total)
        return engine.VarifoldJob(self.c0, self.a0, self.cr0, None, None, None, Ke, total)


_ORDER_POOL = None


def _order_pool():
    """Worker threads for a registration's host-side order work: the choice of its
    processing order during set-up and the estimate of its dispatch orders.  Both call host
    code of the kernel library, which releases the interpreter lock."""
    global _ORDER_POOL
    if _ORDER_POOL is None:
        import concurrent.futures
        _ORDER_POOL = concurrent.futures.ThreadPoolExecutor(max_workers=2,
                                                            thread_name_prefix="lddmm-orders")
    return _ORDER_POOL


def _template_energies(opts, totals) -> None:
    """|S(x_0) - S0|^2 of symmetric registrations at their current template points into
    ``totals`` (1-element device views); the moved templates' geometry is refreshed."""
    for o in opts:
        o.sym.shape.vertices = o.x0
        o.sym.stale = False
    engine.geometry_records(rt.stack([o.sym.geom for o in opts], engine.GEOM_DT),
                            [o.x0.data_ptr() for o in opts])
    for idx in _by_key([o._cur_key for o in opts]):
        g = [opts[i] for i in idx]
        recs = rt.stack([rt.raw(o._data_record(o.sym.shape, o.Kdist.energy(),
                                               target=(o.sym.c0, o.sym.cr0, o.sym.a0)))
                         for o in g])
        key = g[0]._cur_key
        engine.currents_records(recs, key[1], key[2], [totals[i].data_ptr() for i in idx],
                                op=key[3])


def _refresh_template(opts) -> None:
    """Face geometry of the moved templates whose points changed (one launch)."""
    st = [o for o in opts if o.sym.stale]
    if not st:
        return
    engine.geometry_records(rt.stack([o.sym.geom for o in st], engine.GEOM_DT),
                            [o.x0.data_ptr() for o in st])
    for o in st:
        o.sym.shape.vertices = o.x0
        o.sym.stale = False


# --------------------------------------------------------------------------------------
# Batched request implementations
# --------------------------------------------------------------------------------------
class Pending:
    """Answers of a batched request whose device work is in flight.  The launch's FP64
    scalars are copied asynchronously to pinned host memory and an event is recorded after
    the copy; ``result()`` waits for that event only, so work launched afterwards (another
    group of registrations) keeps the GPU busy while the host turns these scalars into
    answers."""

    __slots__ = ("_dev", "_host", "_event", "_finish", "_done", "_value")

    def __init__(self, dev_buf: torch.Tensor, finish):
        self._dev = dev_buf
        self._host = torch.empty(dev_buf.shape, dtype=dev_buf.dtype, pin_memory=True)
        self._host.copy_(dev_buf, non_blocking=True)
        self._event = torch.cuda.Event()
        self._event.record()
        self._finish = finish
        self._done = False
        self._value = None

    def result(self):
        if not self._done:
            self._event.synchronize()
            self._value = self._finish(self._host.numpy())
            self._done = True
            self._dev = self._host = self._event = self._finish = None
        return self._value


class Done:
    """Answers available at launch time (device tensors)."""

    __slots__ = ("_value",)

    def __init__(self, value):
        self._value = value

    def result(self):
        return self._value


def _data_energies(opts, vertices, shapes, totals) -> None:
    """FP64 data-term energies of deformed templates into ``totals`` (1-element device
    views): currents norm against the target for currents registrations (after updating
    their face geometry in ``shapes``), squared distance to the target points for landmark
    registrations.  One launch per stage for the whole list."""
    cur = [i for i, o in enumerate(opts) if o.data_kind in SHAPE_TERMS]
    lmk = [i for i, o in enumerate(opts) if o.data_kind == LANDMARK]
    if cur:
        engine.geometry_batch([shapes[i] for i in cur], [vertices[i] for i in cur])
        engine.currents([engine.CurrentsJob(shapes[i].centers, shapes[i].surfel, opts[i].c2,
                                            opts[i].cr2, opts[i].Kdist.energy(), totals[i],
                                            coef=opts[i].coef)
                         for i in cur if opts[i].data_kind == CURRENT])
        engine.measure([engine.MeasureJob(shapes[i].centers, shapes[i].weights, opts[i].c2,
                                          opts[i].a2, opts[i].Kdist.energy(), totals[i],
                                          coef=opts[i].coef)
                        for i in cur if opts[i].data_kind == MEASURE])
        engine.varifold([engine.VarifoldJob(shapes[i].centers, shapes[i].weights,
                                            shapes[i].surfel, opts[i].c2, opts[i].a2,
                                            opts[i].cr2, opts[i].Kdist.energy(), totals[i],
                                            coef=opts[i].coef)
                         for i in cur if opts[i].data_kind == VARIFOLD])
    if lmk:
        engine.landmark_batch([vertices[i] for i in lmk], [opts[i].y for i in lmk],
                              [totals[i] for i in lmk])


def _refresh_def(opts) -> None:
    """Face geometry of the deformed shapes whose trajectory changed (one launch)."""
    st = [o for o in opts if o._def_stale]
    if not st:
        return
    engine.geometry_records(rt.stack([o._geom_def for o in st], engine.GEOM_DT),
                            [o.xt.data_ptr() + o.Tsize * o._slice_bytes for o in st])
    for o in st:
        o.fvDef_dev.vertices = o.xt[-1]
        o._def_stale = False


def batch_objective(opts: Sequence[DeviceMatching]) -> list:
    """objectiveFun (surfaceMatching.py:233-250): obj0 once, then the energy of at plus the
    data term of the deformed template."""
    return launch_objective(opts).result()


def launch_objective(opts: Sequence[DeviceMatching]) -> Pending:
    """batch_objective with its answers pending."""
    dev = rt.device()
    # per registration: T energies, the data term, the target norm, T affine penalties, and
    # (symmetric) the template term and the template's norm
    lens = [o.Tsize + 2 + (o.Tsize if o.affineDim > 0 else 0) + (2 if o.sym is not None else 0)
            for o in opts]
    scal = torch.zeros(sum(lens), dtype=torch.float64, device=dev)
    flows, norms, offs = [], [], []
    off = 0
    for o, L in zip(opts, lens):
        T = o.Tsize
        offs.append(off)
        A = b = None
        if o.affineDim > 0:
            A, b = o._aff_transforms(o.Afft)
            scal[off + T + 2:off + 2 * T + 2] = o._aff_penalty(o.Afft)
        flows.append(engine.FlowJob(o.x0, o.at, o.Kdiff, o.xt, scal[off:off + T], init=False,
                                    A=A, b=b, perm=o._order))
        off += L
    engine.shoot(flows)
    _data_energies(opts, [o.xt[-1] for o in opts], [o.fvDef_dev for o in opts],
                   [scal[off + o.Tsize:off + o.Tsize + 1] for o, off in zip(opts, offs)])
    for o in opts:
        o._def_stale = False
    # Target self-term (currentNorm0 / measureNorm0 / varifoldNorm0); zero for the landmark
    # data term.
    mnorms, vnorms = [], []
    for o, off in zip(opts, offs):
        T = o.Tsize
        if o.data_kind == CURRENT:
            norms.append(engine.KNormJob(o.c2, o.cr2, o.Kdist.energy(), scal[off + T + 1:off + T + 2]))
        elif o.data_kind == MEASURE:
            mnorms.append(engine.MeasureJob(o.c2, o.a2_norm0, None, None, o.Kdist.energy(),
                                            scal[off + T + 1:off + T + 2]))
        elif o.data_kind == VARIFOLD:
            vnorms.append(engine.VarifoldJob(o.c2, o.a2, o.cr2, None, None, None,
                                             o.Kdist.energy(), scal[off + T + 1:off + T + 2]))
    sym = [(o, off + L - 2) for o, off, L in zip(opts, offs, lens) if o.sym is not None]
    for o, at in sym:
        job = o.sym.norm0_job(o, scal[at + 1:at + 2])
        {CURRENT: norms, MEASURE: mnorms, VARIFOLD: vnorms}[o.data_kind].append(job)
    engine.knorm(norms)
    engine.measure(mnorms)
    engine.varifold(vnorms)
    if sym:
        _template_energies([o for o, _ in sym], [scal[at:at + 1] for _, at in sym])

    def finish(vals):
        out = []
        for o, off, L in zip(opts, offs, lens):
            T = o.Tsize
            s2 = o.sigmaError ** 2
            o.obj0 = vals[off + T + 1] / s2
            data = vals[off + T] / s2
            if o.sym is not None:
                o.obj0 += vals[off + L - 1] / s2
                data += vals[off + L - 2] / s2
            obj = 0
            dt = 1.0 / T
            for t in range(T):
                obj = obj + o.regweight * dt * vals[off + t]
                if o.affineDim > 0:
                    obj += vals[off + T + 2 + t]
            o.obj = obj + (o.obj0 + data)
            o.nobj += 1
            out.append(o.obj)
        return out
    return Pending(scal, finish)


def _by_key(keys):
    """Index arrays of equal keys, in first-appearance order."""
    groups = {}
    for i, k in enumerate(keys):
        groups.setdefault(k, []).append(i)
    return [np.asarray(v) for v in groups.values()]


def batch_update_try(items) -> list:
    """updateTry (surfaceMatching.py:255-289) for [(opt, dir, eps, objRef), ...]."""
    return launch_update_try(items).result()


def launch_update_try(items) -> Pending:
    """batch_update_try with its answers pending (the kept trial is recorded when they are
    read)."""
    dev = rt.device()
    n = len(items)
    opts = [it[0] for it in items]
    for o in opts:
        o._count_trial()
    Ts = np.fromiter((o.Tsize for o in opts), np.int64, n)
    offs = np.zeros(n, np.int64)
    np.cumsum(Ts[:-1] + 1, out=offs[1:])
    aff = [i for i, o in enumerate(opts) if o.affineDim > 0]
    # affine penalties of the trials after every registration's (T energies, data term)
    pen_off = {}
    end = int(offs[-1] + Ts[-1] + 1)
    for i in aff:
        pen_off[i] = end
        end += int(Ts[i])
    # symmetric registrations: the template term of the trial's x0Try
    sym = [i for i, o in enumerate(opts) if o.sym is not None]
    sym_off = {}
    for i in sym:
        sym_off[i] = end
        end += 1
    scal = torch.zeros(end, dtype=torch.float64, device=dev)
    sbase = np.uint64(scal.data_ptr())
    # atTry = at - eps dir for every registration (one launch per 32 registrations)
    ats = [torch.empty_like(o.at) for o in opts]
    engine.axpy_batch(ats, [o.at for o in opts], [it[1].diff for it in items],
                      [-float(it[2]) for it in items])
    for i in sym:
        # x0Try = x0 - eps dir.initx into slice 0 of the trial's trajectory
        o, dir_, eps = items[i][0], items[i][1], items[i][2]
        torch.add(o.x0, dir_.initx, alpha=-float(eps), out=o._xt_scratch[0])
    xt = np.fromiter((o._xt_scratch.data_ptr() for o in opts), np.uint64, n)
    at = np.fromiter((a.data_ptr() for a in ats), np.uint64, n)
    slice_bytes = np.fromiter((o._slice_bytes for o in opts), np.uint64, n)
    energy = sbase + np.uint64(8) * offs.astype(np.uint64)
    data_total = energy + np.uint64(8) * Ts.astype(np.uint64)
    final = xt + Ts.astype(np.uint64) * slice_bytes          # address of x_T in each trajectory
    afft_try = {}
    Aptr = np.zeros(n, np.uint64)
    bptr = np.zeros(n, np.uint64)
    keep = []                                                # affine parts live until launched
    for i in aff:
        o, dir_, eps = items[i][0], items[i][1], items[i][2]
        afft_try[i] = o.Afft - float(eps) * dir_.aff
        A, b = o._aff_transforms(afft_try[i])
        keep.append((A, b))
        Aptr[i], bptr[i] = A.data_ptr(), b.data_ptr()
        scal[pen_off[i]:pen_off[i] + int(Ts[i])] = o._aff_penalty(afft_try[i])
    for idx in _by_key([o._flow_key for o in opts]):
        o0 = opts[idx[0]]
        has_aff = bool(Aptr[idx].any())
        engine.shoot_records(rt.stack([opts[i]._flow_base for i in idx]), o0.Kdiff.prec,
                             o0.dim, Ts[idx], xt[idx], at[idx], energy[idx],
                             A=Aptr[idx] if has_aff else None, b=bptr[idx] if has_aff else None)
    del keep
    cur = [i for i, o in enumerate(opts) if o.data_kind in SHAPE_TERMS]
    if cur:
        engine.geometry_records(rt.stack([opts[i]._geom_try for i in cur], engine.GEOM_DT),
                                final[cur])
        for idx in _by_key([opts[i]._cur_key for i in cur]):
            g = [cur[i] for i in idx]
            key = opts[g[0]]._cur_key
            engine.currents_records(rt.stack([opts[i]._cur_try for i in g]),
                                    key[1], key[2], data_total[g], op=key[3])
    lmk = [i for i, o in enumerate(opts) if o.data_kind == LANDMARK]
    if lmk:
        engine.landmark_batch([opts[i]._xt_scratch[-1] for i in lmk], [opts[i].y for i in lmk],
                              [int(data_total[i]) for i in lmk])
    if sym:
        engine.geometry_records(rt.stack([opts[i].sym.try_geom for i in sym], engine.GEOM_DT),
                                xt[sym])
        for idx in _by_key([opts[i]._cur_key for i in sym]):
            g = [sym[i] for i in idx]
            key = opts[g[0]]._cur_key
            engine.currents_records(rt.stack([opts[i].sym.try_rec for i in g]), key[1], key[2],
                                    [int(sbase) + 8 * sym_off[i] for i in g], op=key[3])

    def finish(host):
        vals = host.tolist()
        out = []
        for i, ((o, dir, eps, objRef), atTry, off) in enumerate(zip(items, ats, offs.tolist())):
            T = o.Tsize
            c = o.regweight * (1.0 / T)
            objTry = o.obj0
            objDef = 0
            pen = vals[pen_off[i]:pen_off[i] + T] if i in pen_off else None
            for t, e in enumerate(vals[off:off + T]):
                objDef = objDef + c * e
                if pen is not None:
                    objDef += pen[t]
            objTry += objDef
            if i in sym_off:
                objTry += vals[off + T] / (o.sigmaError ** 2) + vals[sym_off[i]] / (o.sigmaError ** 2)
            else:
                objTry += vals[off + T] / (o.sigmaError ** 2)
            o.nobj += 1
            if math.isnan(objTry):
                out.append(float('inf'))
                continue
            if (objRef is None) or (objTry < objRef):
                o.atTry = atTry
                o.objTry = objTry
                if i in afft_try:
                    o.AfftTry = afft_try[i]
                o._xt_try, o._xt_scratch = o._xt_scratch, o._xt_try
            out.append(objTry)
        return out
    return Pending(scal, finish)


def launch_gradient(items) -> Done:
    """batch_gradient; the gradients are device tensors, available at once."""
    return Done(batch_gradient(items))


def batch_gradient(items) -> list:
    """getGradient (surfaceMatching.py:301-327): px1 = -grad(data)/sigma^2 at the deformed
    template, covector along the kept trajectory, dat = 2 lambda a - p, scaled by
    1/(coeff T)."""
    n = len(items)
    opts = [it[0] for it in items]
    _refresh_def(opts)
    # End-point covector px1, written into slice T of each registration's covector buffer.
    cur = [i for i, o in enumerate(opts) if o.data_kind in SHAPE_TERMS]
    if cur:
        # data-term gradient (z1, dz), then the vertex gathers with scale -2/sigma^2
        for idx in _by_key([opts[i]._cur_grad_key for i in cur]):
            g = [cur[i] for i in idx]
            key = opts[g[0]]._cur_grad_key
            rt.launch(key[3], key[1], key[2], rt.stack([opts[i]._cur_grad for i in g]))
        engine.gather_records(rt.stack([opts[i]._gather for i in cur], engine.GATHER_DT),
                              [opts[i].xt.data_ptr() + opts[i].Tsize * opts[i]._slice_bytes
                               for i in cur])
    sym = [o for o in opts if o.sym is not None]
    if sym:
        # d|S(x0) - S0|^2/dx0 / sigmaError^2 into each template term's px0
        _refresh_template(sym)
        for idx in _by_key([o._cur_grad_key for o in sym]):
            g = [sym[i] for i in idx]
            key = g[0]._cur_grad_key
            rt.launch(key[3], key[1], key[2], rt.stack([o.sym.grad_rec for o in g]))
        engine.gather_records(rt.stack([o.sym.gather for o in sym], engine.GATHER_DT),
                              [o.x0.data_ptr() for o in sym])
    lmk = [o for o in opts if o.data_kind == LANDMARK]
    if lmk:
        # px1 = -2 (x_T - y) / sigmaError^2
        engine.landmark_batch([o.xt[-1] for o in lmk], [o.y for o in lmk], [None] * len(lmk),
                              [o._pxT for o in lmk], [-2.0 / (o.sigmaError ** 2) for o in lmk])
    Ts = np.fromiter((o.Tsize for o in opts), np.int64, n)
    xt = np.fromiter((o.xt.data_ptr() for o in opts), np.uint64, n)
    at = np.fromiter((o.at.data_ptr() for o in opts), np.uint64, n)
    px = np.fromiter((o._pxt.data_ptr() for o in opts), np.uint64, n)
    Aptr = np.zeros(n, np.uint64)
    keep = []                                                # affine parts live until launched
    for i, o in enumerate(opts):
        if o.affineDim > 0:
            A, _ = o._aff_transforms(o.Afft)
            keep.append(A)
            Aptr[i] = A.data_ptr()
    for idx in _by_key([o._flow_key for o in opts]):
        o0 = opts[idx[0]]
        has_aff = bool(Aptr[idx].any())
        engine.covector_records(rt.stack([opts[i]._cov_base for i in idx]), o0.Kdiff.prec,
                                o0.dim, Ts[idx], xt[idx], at[idx], px[idx],
                                A=Aptr[idx] if has_aff else None)
    del keep
    # dat = 2 lambda a - p_{t+1};  grd = dat / (coeff T)
    dats = torch._foreach_mul([o.at for o in opts], [2 * o.regweight for o in opts])
    torch._foreach_sub_(dats, [o._pxt_tail for o in opts])
    torch._foreach_div_(dats, [it[1] * it[0].Tsize for it in items])
    out = []
    for (o, coeff), dat in zip(items, dats):
        g = Direction()
        g.diff = dat
        g.aff = _affine_gradient(o, coeff) if o.affineDim > 0 else np.zeros(o.Afft.shape)
        if o.sym is not None:
            g.initx = (o.sym.px0 - o._pxt[0]) / (o.sym.coeffInitx * coeff)
        o.ngrad += 1
        out.append(g)
    return out


def _affine_gradient(o, coeff):
    """Affine part of getGradient from the covector (surfaceMatching.py:313-322,
    curveMatching.py:246-252): dA_t = p_{t+1}^T x_t, db_t = sum p_{t+1}, projected on the
    basis, with the drivers' own scalings."""
    T, d = o.Tsize, o.dim
    dA = torch.einsum('tni,tnj->tij', o._pxt[1:], o.xt[:-1])
    db = o._pxt[1:].sum(dim=1)
    dAff = torch.cat([dA.reshape(T, d * d), db], dim=1) @ o._basis     # (T, affineDim)
    if o.aff_style == 'curve':
        return (2 * o.Afft - dAff / o._aff_w) / (coeff * T)
    return (2 * (o._aff_w * o.Afft) - dAff) / (o.coeffAff * coeff * T)


def batch_dot(items) -> list:
    """dotProduct (surfaceMatching.py:353-377): sum_t <g2_t, K(x_t) g1_t> for 1..2 g2's."""
    return launch_dot(items).result()


def launch_dot(items) -> Pending:
    """batch_dot with its answers pending."""
    dev = rt.device()
    n = len(items)
    opts = [it[0] for it in items]
    Ts = np.fromiter((o.Tsize for o in opts), np.int64, n)
    first = np.zeros(n, np.int64)
    np.cumsum(Ts[:-1], out=first[1:])
    buf = torch.zeros((int(Ts.sum()), 2), dtype=torch.float64, device=dev)
    out_ptr = np.uint64(buf.data_ptr()) + np.uint64(16) * first.astype(np.uint64)
    xt = np.fromiter((o.xt.data_ptr() for o in opts), np.uint64, n)
    g1 = np.fromiter((it[1].diff.data_ptr() for it in items), np.uint64, n)
    na = np.fromiter((len(it[2]) for it in items), np.int32, n)
    g2 = np.zeros((n, 2), np.uint64)
    for i, it in enumerate(items):
        for q, g in enumerate(it[2]):
            g2[i, q] = g.diff.data_ptr()
    for idx in _by_key([o._flow_key for o in opts]):
        o0 = opts[idx[0]]
        engine.kdot_records(rt.stack([opts[i]._flow_base for i in idx]), o0.Kdiff.prec,
                            o0.dim, Ts[idx], xt[idx], g1[idx], g2[idx], na[idx], out_ptr[idx])
    aff = [i for i, o in enumerate(opts) if o.affineDim > 0]
    if aff:
        # affine products per time step in columns 2..3 (surfaceMatching.py:366-369,
        # curveMatching.py:288-296)
        abuf = torch.zeros_like(buf)
        for i in aff:
            o, a1 = opts[i], items[i][1].aff * opts[i]._aff_dot_weight()
            for q, g in enumerate(items[i][2]):
                abuf[first[i]:first[i] + Ts[i], q] = (a1 * g.aff).sum(dim=1)
        buf = torch.cat([buf, abuf], dim=1)
    sym = [i for i, o in enumerate(opts) if o.sym is not None]
    sym_row = {}
    if sym:
        # x0 parts <g1.initx, g2.initx> (surfaceMatching.py:373-376), one row per registration
        sbuf = torch.zeros((len(sym), buf.shape[1]), dtype=torch.float64, device=dev)
        for r, i in enumerate(sym):
            sym_row[i] = int(Ts.sum()) + r
            for q, g in enumerate(items[i][2]):
                sbuf[r, q] = (items[i][1].initx * g.initx).sum()
        buf = torch.cat([buf, sbuf], dim=0)

    def finish(allv):
        out = []
        for i in range(n):
            # sequential sum over t (numpy's accumulate), as the reference's loop
            v = allv[first[i]:first[i] + Ts[i], :na[i]]
            if opts[i].affineDim > 0:
                a = allv[first[i]:first[i] + Ts[i], 2:2 + na[i]]
                r = np.zeros(na[i])
                for t in range(Ts[i]):
                    r = r + v[t]
                    r += a[t]
                out.append(r)
            else:
                out.append(np.cumsum(v, axis=0)[-1].copy())
            if i in sym_row:
                out[-1] = out[-1] + allv[sym_row[i], :na[i]] * opts[i].sym.coeffInitx
        return out
    return Pending(buf, finish)
