"""Event samples: phase-space MC and data drawn by accept-reject from the model.

Accept-reject follows the reference's toy procedure (per-event |A|^2 weights,
src/RooDalitzAmplitude.cxx:1293-1628, then hit-and-reject,
gen_toys/hit_reject_for_amplitude_eff/hit_and_reject.py): candidates are
uniform phase-space events, |A|^2 comes from the GPU (pwa_amp2, unsmeared),
and an event is kept with probability |A|^2 / (1.2 * max |A|^2 of the batch).
"""
import numpy as np

from . import abi
from ._lib import Context, generate_phsp


def accept_reject(model, n_target, seed, batch=None):
    rng = np.random.default_rng(seed)
    batch = batch or max(4 * n_target, 20000)
    out, got, it = [], 0, 0
    with Context(model.desc) as c:
        while got < n_target:
            cand = generate_phsp(batch, seed * 7919 + it + 1)
            c.upload(abi.PWA_MC, cand, np.ones(batch))
            a2 = c.amp2(abi.PWA_MC, model.params)
            keep = rng.uniform(0.0, 1.2 * a2.max(), size=batch) < a2
            out.append(cand[:, keep])
            got += int(keep.sum())
            it += 1
    return np.concatenate(out, axis=1)[:, :n_target]
