"""Data and MC contributions to the NLL from per-event weights: the
decomposition of python/calculate_nll_contributions.py:24-33, on the arrays
pwa_eval exports (fx_data = |A|^2 per data event, fx_mc = w |A|^2 per MC
event, the reference's amp2 and amp2_weight branches).

    alpha = sum(w_data) / sum(w_data^2)
    data  = alpha * sum(w_data * ln amp2)          (amp2 <= 0 contributes 0)
    mc    = alpha * sum(w_data) * ln(sum(amp2_weight) / sum(w_mc))
    total = mc - data

Where every amp2 > 0, total = alpha * (NLL - NLL_SHIFT) of pwa_eval (its
fx <= 0 penalty differs: the reference's script maps those events to 0).
"""
import numpy as np


def nll_contributions(amp2, w_data, amp2_weight, w_mc):
    amp2 = np.asarray(amp2, dtype=np.float64)
    w_data = np.asarray(w_data, dtype=np.float64)
    amp2_weight = np.asarray(amp2_weight, dtype=np.float64)
    w_mc = np.asarray(w_mc, dtype=np.float64)
    sw = float(np.sum(w_data))
    alpha = sw / float(np.sum(np.square(w_data)))
    pos = amp2 > 0
    data = alpha * float(np.sum(w_data * np.where(pos, np.log(np.where(pos, amp2, 1.0)), 0.0)))
    mc = alpha * sw * float(np.log(np.sum(amp2_weight) / np.sum(w_mc)))
    return {"alpha": alpha, "data": data, "mc": mc, "total": mc - data}


def nll_contributions_at(ctx, params, w_data, w_mc):
    """The decomposition at a parameter point from a Context whose data and MC
    samples carry the weights w_data, w_mc (caller order)."""
    _, _, fxd, fxm = ctx.eval(params, want_fx=True)
    return nll_contributions(fxd, w_data, fxm, w_mc)
