"""Fits on the GPU likelihood: the variable-metric minimiser with the
analytic gradient converges, and reaches the same minimum as the plain
finite-difference mode from the same start."""
import numpy as np
import pytest

from pwa_amd import abi, models, toys
from pwa_amd._lib import Context, generate_phsp
from pwa_amd.fit import Objective, minimize

pytestmark = pytest.mark.gpu


def _setup(model, nd, nm, seed):
    mcols = generate_phsp(nm, seed)
    dcols = toys.accept_reject(model, nd, seed + 1)
    ctx = Context(model.desc)
    ctx.upload(abi.PWA_DATA, dcols, np.ones(nd))
    ctx.upload(abi.PWA_MC, mcols, np.ones(nm))
    return ctx


def test_fit_analytic_and_plain_agree():
    model = models.config_small(1)
    ctx = _setup(model, 4000, 20000, 61)
    fl = np.flatnonzero(model.float_mask)
    rng = np.random.default_rng(62)
    x0 = model.params[fl] * (1.0 + 0.03 * rng.standard_normal(len(fl)))
    oa = Objective(ctx, model.params, model.float_mask, "analytic")
    ra = minimize(oa, x0)
    of = Objective(ctx, model.params, model.float_mask, "fd")
    rf = minimize(of, x0)
    f_true = ctx.fcn(model.params)
    print(f"[fit] analytic: fcn {ra.fval:.6f} edm {ra.edm:.1e} iter {ra.n_iter} grads {ra.n_grad} "
          f"{ra.seconds:.2f}s | plain: fcn {rf.fval:.6f} edm {rf.edm:.1e} iter {rf.n_iter} fcn calls {rf.n_fcn} "
          f"{rf.seconds:.2f}s | fcn(truth) {f_true:.6f}", flush=True)
    ctx.close()
    assert ra.converged and rf.converged
    assert ra.fval <= f_true + 1e-6
    assert abs(ra.fval - rf.fval) < 1e-4
    dev = np.max(np.abs(ra.x - rf.x) / ra.errors)
    print(f"[fit] max |x_analytic - x_plain| / error = {dev:.2e}", flush=True)
    assert dev < 1e-2
