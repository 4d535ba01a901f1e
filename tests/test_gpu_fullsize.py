"""The nominal and K-matrix workloads at their full size (300k accept-reject
data, 3M phase-space MC; 216 and 246 floating parameters) checked through
properties that do not depend on the size: repeated and value-only
evaluations agree bit for bit, the NLL is invariant under a global complex
rescaling of the couplings and under a rescaling of the MC weights, and the
analytic gradient matches a central difference along a random direction."""
import os
import sys
from types import SimpleNamespace

import numpy as np
import pytest

from pwa_amd import abi, models
from pwa_amd._lib import Context

sys.path.insert(0, os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
import bench  # noqa: E402
from test_oracle_pins import _scale_all_couplings  # noqa: E402

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module", params=["nominal", "kmatrix"])
def workload(request):
    model, p, dcols, mcols, _ = bench.build_workload(300_000, 3_000_000, name=request.param)
    print(f"[fullsize] {request.param}: {dcols.shape[1]} data, {mcols.shape[1]} MC events", flush=True)
    return model, p, dcols, mcols


def _ctx(model, dcols, mcols, wscale=1.0):
    c = Context(model.desc)
    c.upload(abi.PWA_DATA, dcols, np.ones(dcols.shape[1]), models.window_schemes(dcols[1]))
    c.upload(abi.PWA_MC, mcols, wscale * np.ones(mcols.shape[1]), models.window_schemes(mcols[1]))
    return c


def test_fullsize_properties(workload):
    model, p, dcols, mcols = workload
    with _ctx(model, dcols, mcols) as c:
        v1, g1 = c.eval_grad(p)
        v2, g2 = c.eval_grad(p)
        v0, _ = c.eval(p)
        assert v1 == v2 and np.array_equal(g1, g2), "repeated evaluations differ"
        assert v0 == v1, f"value-only NLL {v0!r} != gradient-evaluation NLL {v1!r}"
        # global complex rescaling of every coupling: |A|^2 scales, the NLL does not
        ps = _scale_all_couplings(SimpleNamespace(params=p, desc=model.desc), 1.7 * np.exp(0.6j))
        vs, _ = c.eval(ps)
        print(f"[fullsize] nll {v1:.10f}, rescaled couplings {vs:.10f}", flush=True)
        assert abs(vs - v1) < 1e-11 * abs(v1)
        # directional derivative along a random direction of the floating parameters
        fl = np.flatnonzero(model.float_mask)
        rng = np.random.default_rng(5)
        u = np.zeros_like(p)
        u[fl] = rng.standard_normal(len(fl)) * np.maximum(np.abs(p[fl]), 1e-2)
        def cd(h):
            return (c.eval(p + h * u)[0] - c.eval(p - h * u)[0]) / (2 * h)

        # central differences at h and h/2, Richardson-extrapolated (the
        # O(h^2) truncation of a step in all 216 parameters at once is
        # ~1e-5 relative at h = 1e-5)
        d1, d2 = cd(2e-5), cd(1e-5)
        fd = (4.0 * d2 - d1) / 3.0
        an = float(g1 @ u)
        print(f"[fullsize] directional derivative: analytic {an:.10e}, central differences {d1:.10e} (h=2e-5) "
              f"{d2:.10e} (h=1e-5), extrapolated {fd:.10e}", flush=True)
        assert abs(an - fd) < 1e-6 * max(abs(fd), 1.0)
    with _ctx(model, dcols, mcols, wscale=3.0) as c:
        v3, _ = c.eval(p)
    print(f"[fullsize] MC weights x3: nll {v3:.10f}", flush=True)
    assert abs(v3 - v1) < 1e-11 * abs(v1)
