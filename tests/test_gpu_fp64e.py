"""FP64E, the data-term energy form of FP32-compensated registrations (FP64 arithmetic, pairs
below 2^-40 culled, the FP32 exponential below 2^-16), against the FP64 form of the same
energies: the currents, measure and varifold data terms of an icosphere-5 surface (20,480
faces) against targets from far to nearly matching, where the data term is a tiny difference
of the sums and the culled terms of its sums must cancel.  The bar is 1e-9 of the data
term (measured 2.2e-10 on the 81,920-face pair at a data term 6.4e-7 of the sums)."""
import numpy as np
import pytest
import torch

from registration_amd import _lib, engine, runtime as rt, surfaces, synthetic

pytestmark = pytest.mark.gpu


def _energy(op, prec, S, T, sigma, weights):
    tot = torch.empty(1, dtype=torch.float64, device="cuda")
    nf = S.centers.shape[0]
    rt.Plan(op, prec, 3, [rt.seg(y=S.centers, ny=nf, x1=T.centers, nx1=nf, sigma=sigma,
                                 coef=1.0 / 3.0, kernel=_lib.KERNEL_GAUSS, order=10, total=tot,
                                 **weights)]).run()
    # the target's own norm: KNORM of its surfels (currents), or the term without a target
    norm = torch.empty(1, dtype=torch.float64, device="cuda")
    own_op, own = ((_lib.OP_KNORM, dict(w0=T.surfel)) if op == _lib.OP_CURRENTS
                   else (op, dict(w0=T.weights, w2=T.surfel)))
    rt.Plan(own_op, prec, 3, [rt.seg(y=T.centers, ny=nf, sigma=sigma, coef=1.0 / 3.0,
                                     kernel=_lib.KERNEL_GAUSS, order=10, total=norm,
                                     **own)]).run()
    torch.cuda.synchronize()
    return tot.item(), norm.item()


@pytest.mark.parametrize("term", ["currents", "measure", "varifold"])
def test_fp64e_data_term_within_1e9_of_fp64(term):
    V, F = synthetic.icosphere(5)
    Y = synthetic.rbf_field(np.random.default_rng(7), V)   # V + a smooth displacement
    S = surfaces.DeviceSurface(F, torch.as_tensor(V, device="cuda"), with_weights=True)
    op = {"currents": _lib.OP_CURRENTS, "measure": _lib.OP_MEASURE,
          "varifold": _lib.OP_VARIFOLD}[term]
    for eps in (1.0, 1e-2, 1e-3):
        T = surfaces.DeviceSurface(F, torch.as_tensor(V + eps * (Y - V), device="cuda"),
                                   with_weights=True)
        weights = {"currents": dict(w0=S.surfel, w1=T.surfel),
                   "measure": dict(w0=S.weights, w1=T.weights, w2=S.surfel),
                   "varifold": dict(w0=S.weights, w1=T.weights, w2=S.surfel,
                                    sw=[T.surfel])}[term]
        d64, n64 = _energy(op, _lib.PREC_FP64, S, T, 0.1, weights)
        d64e, n64e = _energy(op, _lib.PREC_FP64E, S, T, 0.1, weights)
        data64, data64e = d64 + n64, d64e + n64e
        assert data64 > 0
        assert abs(data64e - data64) <= 1e-9 * data64, (term, eps, data64, data64e)
        assert abs(d64e - d64) <= 1e-12 * abs(d64) and abs(n64e - n64) <= 1e-12 * abs(n64)


def test_fp32c_kernels_take_fp64e_energies():
    from registration_amd.kernelFunctions import Kernel
    assert engine.DevKernel.of(Kernel('gauss', sigma=0.1, precision='fp32c')).energy().prec == \
        _lib.PREC_FP64E
    assert engine.DevKernel.of(Kernel('gauss', sigma=0.1, precision='fp64')).energy().prec == \
        _lib.PREC_FP64
