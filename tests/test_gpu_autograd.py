"""torch.autograd.Function over the C ABI: backward equals the adjoint-kernel
gradient and central finite differences of the forward."""
import numpy as np
import pytest
import torch

from pwa_amd import abi, models
from pwa_amd._lib import Context, generate_phsp
from pwa_amd.autograd import fcn, nll

pytestmark = pytest.mark.gpu


def test_autograd_backward_matches_adjoint_and_fd():
    m = models.config_small(1)
    d, mc = generate_phsp(700, 51), generate_phsp(1400, 52)
    with Context(m.desc) as c:
        c.upload(abi.PWA_DATA, d, np.ones(700))
        c.upload(abi.PWA_MC, mc, np.ones(1400))
        x = torch.tensor(m.params, dtype=torch.float64, requires_grad=True)
        y = nll(x, c)
        y.backward()
        v, g = c.eval_grad(m.params)
        assert y.item() == v
        assert torch.allclose(x.grad, torch.from_numpy(g))
        k = int(np.flatnonzero(m.float_mask)[3])
        h = 1e-5 * abs(m.params[k])
        pp, pm = m.params.copy(), m.params.copy()
        pp[k] += h
        pm[k] -= h
        fd = (c.eval(pp)[0] - c.eval(pm)[0]) / (2 * h)
        assert abs(x.grad[k].item() - fd) <= 1e-5 * abs(fd) + 1e-6
        # FCN through the same path, gradient scaled by alpha
        x2 = torch.tensor(m.params, dtype=torch.float64, requires_grad=True)
        f = fcn(x2, c)
        f.backward()
        assert torch.allclose(x2.grad, torch.from_numpy(g) * c.alpha(), rtol=1e-12, atol=1e-12)
