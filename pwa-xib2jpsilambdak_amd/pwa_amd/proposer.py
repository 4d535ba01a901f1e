"""Learned start-point proposer: an MLP (BF16 MFMA kernels, pwa_amd.mlp)
trained on (parameters, FCN, gradient) samples that the GPU likelihood
produces along fit trajectories.  Besides the FCN and its gradient at a point,
the network learns the FCN at the end of the fit the point lies on (every
point of a trajectory flows to that fit's minimum): `rank` orders random
starting points by that predicted outcome; `propose` instead descends them on
the learned surface first.  Every proposal is refit with the FP64 likelihood,
which decides.
"""
import numpy as np

from .mlp import MLP


class Proposer:
    def __init__(self, n_float, lo, hi, n_hidden=256, n_layers=3, seed=1, grad_weight=0.5):
        self.P = n_float
        self.lo = np.asarray(lo, dtype=np.float64)
        self.hi = np.asarray(hi, dtype=np.float64)
        self.net = MLP(n_float, n_float + 2, n_hidden, n_layers, seed=seed)
        self.grad_weight = grad_weight

    def close(self):
        self.net.close()

    # normalisation: inputs standardised per parameter, FCN standardised,
    # gradient expressed in the standardised coordinates
    def _norm_x(self, X):
        return (X - self.mu_x) / self.sd_x

    def train(self, X, F, G, F_end, epochs=300, batch=512, lr=1e-3, seed=0):
        """Fits the network to the fit outcome F_end, the FCN and its
        gradient.  FCN values are compressed, u = (F - F_lo) / s, y = asinh(u)
        (F_lo = 2nd percentile, s = 10-50 percentile spread of F), so the low
        region of the surface carries most of the weight; the gradient target
        is the derivative of y with respect to the standardised inputs.
        Outputs: [outcome, FCN, gradient (P)]."""
        X = np.asarray(X, dtype=np.float64)
        F = np.asarray(F, dtype=np.float64)
        G = np.asarray(G, dtype=np.float64)
        F_end = np.asarray(F_end, dtype=np.float64)
        self.mu_x = X.mean(axis=0)
        self.sd_x = np.maximum(X.std(axis=0), 1e-6)
        self.f_lo = np.percentile(F, 2)
        self.s_f = max(np.percentile(F, 50) - np.percentile(F, 10), 1e-6)
        self.e_lo = F_end.min()
        self.s_e = max(np.std(F_end), 1e-6)
        u = (F - self.f_lo) / self.s_f
        yf = np.arcsinh(u)
        ye = (F_end - self.e_lo) / self.s_e
        yg = G * (self.sd_x / self.s_f) / np.sqrt(1.0 + u * u)[:, None]
        y = np.concatenate([ye[:, None], yf[:, None], np.clip(yg, -10.0, 10.0)], axis=1)
        ow = np.concatenate([[1.0, 1.0], np.full(self.P, self.grad_weight / self.P)])
        self.net.set_data(self._norm_x(X), y, ow)
        return self.net.train(epochs, batch=min(batch, len(X)), lr=lr, seed=seed)

    def predict(self, X):
        """Predicted fit outcome and FCN at X."""
        out = self.net.predict(self._norm_x(np.asarray(X, dtype=np.float64))).astype(np.float64)
        return self.e_lo + self.s_e * out[:, 0], self.f_lo + self.s_f * np.sinh(np.clip(out[:, 1], -700.0, 700.0))

    def rank(self, starts, n_top):
        """The n_top starts with the best predicted fit outcome, best first."""
        starts = np.asarray(starts, dtype=np.float64)
        outcome, _ = self.predict(starts)
        order = np.argsort(outcome)[:n_top]
        return starts[order], outcome[order]

    def propose(self, starts, n_top, steps=200, eta=0.02, min_dist=0.5):
        """Descend every start on the learned surface (normalised
        coordinates, steps of eta along the normalised predicted gradient,
        kept within the limits) and return up to n_top end points with the
        lowest predicted FCN, best first, at least min_dist apart (normalised
        Euclidean distance / sqrt(P))."""
        z = self._norm_x(np.asarray(starts, dtype=np.float64))
        zlo, zhi = self._norm_x(self.lo), self._norm_x(self.hi)
        for _ in range(steps):
            out = self.net.predict(z).astype(np.float64)
            g = out[:, 2:]
            g /= np.maximum(np.sqrt(np.mean(g * g, axis=1, keepdims=True)), 1e-6)
            z = np.clip(z - eta * g, zlo, zhi)
        score = self.net.predict(z)[:, 1]
        picked = []
        for j in np.argsort(score):
            if all(np.linalg.norm(z[j] - z[k]) / np.sqrt(self.P) > min_dist for k in picked):
                picked.append(j)
            if len(picked) == n_top:
                break
        picked = np.array(picked, dtype=int)
        return self.mu_x + self.sd_x * z[picked], self.f_lo + self.s_f * np.sinh(score[picked])
