"""Learned start-point proposer: an MLP (BF16 MFMA kernels, pwa_amd.mlp)
trained on (parameters, FCN, gradient) samples that the GPU likelihood
produces along fit trajectories; it ranks random starting points after a
short descent on the learned surface.  Every proposal is refit with the FP64
likelihood, which decides.
"""
import numpy as np

from .mlp import MLP


class Proposer:
    def __init__(self, n_float, lo, hi, n_hidden=256, n_layers=3, seed=1, grad_weight=0.5):
        self.P = n_float
        self.lo = np.asarray(lo, dtype=np.float64)
        self.hi = np.asarray(hi, dtype=np.float64)
        self.net = MLP(n_float, n_float + 1, n_hidden, n_layers, seed=seed)
        self.grad_weight = grad_weight

    def close(self):
        self.net.close()

    # normalisation: inputs standardised per parameter, FCN standardised,
    # gradient expressed in the standardised coordinates
    def _norm_x(self, X):
        return (X - self.mu_x) / self.sd_x

    def train(self, X, F, G, epochs=200, batch=256, lr=1e-3, seed=0):
        X = np.asarray(X, dtype=np.float64)
        F = np.asarray(F, dtype=np.float64)
        G = np.asarray(G, dtype=np.float64)
        self.mu_x = X.mean(axis=0)
        self.sd_x = np.maximum(X.std(axis=0), 1e-6)
        # robust FCN scale: the samples span many orders of the FCN; train on
        # a clipped, standardised value so the low region dominates
        self.f_ref = np.percentile(F, 50)
        self.sd_f = max(np.percentile(F, 90) - np.percentile(F, 10), 1e-6)
        z = self._norm_x(X)
        yf = np.clip((F - self.f_ref) / self.sd_f, -5.0, 5.0)
        yg = np.clip(G * self.sd_x / self.sd_f, -5.0, 5.0)
        y = np.concatenate([yf[:, None], yg], axis=1)
        ow = np.concatenate([[1.0], np.full(self.P, self.grad_weight / self.P)])
        self.net.set_data(z, y, ow)
        return self.net.train(epochs, batch=min(batch, len(X)), lr=lr, seed=seed)

    def predict(self, X):
        out = self.net.predict(self._norm_x(np.asarray(X, dtype=np.float64))).astype(np.float64)
        return self.f_ref + self.sd_f * out[:, 0], out[:, 1:] * self.sd_f / self.sd_x

    def propose(self, starts, n_top, steps=100, eta=0.05):
        """Descend every start on the learned surface (normalised
        coordinates, steps of eta along the normalised predicted gradient,
        kept within the limits) and return the n_top end points with the
        lowest predicted FCN, best first."""
        z = self._norm_x(np.asarray(starts, dtype=np.float64))
        zlo, zhi = self._norm_x(self.lo), self._norm_x(self.hi)
        for _ in range(steps):
            out = self.net.predict(z).astype(np.float64)
            g = out[:, 1:]
            g /= np.maximum(np.sqrt(np.mean(g * g, axis=1, keepdims=True)), 1e-6)
            z = np.clip(z - eta * g, zlo, zhi)
        score = self.net.predict(z)[:, 0]
        order = np.argsort(score)[:n_top]
        return self.mu_x + self.sd_x * z[order], score[order]
