"""The learned surrogate: an MLP from the parameter vector to all predicted
log bin contents, trained on the GPU (BF16 MFMA kernels of mlp.hip) on
parameter/bin pairs that the exact path generates, and its likelihood.
"""
from __future__ import annotations

import ctypes as C
from typing import Optional

import numpy as np

from . import _native as N


class MLP:
    """ctypes wrapper of gcm_mlp (n_in -> hidden -> hidden -> n_out, GELU)."""

    def __init__(self, n_in: int, n_hidden: int, n_out: int, seed: int = 0,
                 device: int = 0):
        self._lib = N.lib()
        h = C.c_void_p()
        N.check(self._lib.gcm_mlp_create(device, n_in, n_hidden, n_out, seed,
                                         C.byref(h)))
        self._h = h
        self.n_in, self.n_hidden, self.n_out = n_in, n_hidden, n_out
        n = C.c_int64()
        N.check(self._lib.gcm_mlp_param_count(self._h, C.byref(n)))
        self.n_params = n.value

    def params(self) -> np.ndarray:
        p = np.empty(self.n_params, np.float32)
        N.check(self._lib.gcm_mlp_get_params(self._h, p.ctypes.data))
        return p

    def set_params(self, p: np.ndarray) -> None:
        p = np.ascontiguousarray(p, dtype=np.float32)
        assert p.size == self.n_params
        N.check(self._lib.gcm_mlp_set_params(self._h, p.ctypes.data))

    def split(self, p: np.ndarray):
        """(W1, b1, W2, b2, W3, b3) views of a flat parameter vector."""
        i, h, o = self.n_in, self.n_hidden, self.n_out
        sizes = [h * i, h, h * h, h, o * h, o]
        shapes = [(h, i), (h,), (h, h), (h,), (o, h), (o,)]
        out, off = [], 0
        for s, sh in zip(sizes, shapes):
            out.append(p[off:off + s].reshape(sh))
            off += s
        return out

    def predict(self, x: np.ndarray) -> np.ndarray:
        x = np.ascontiguousarray(np.atleast_2d(x), dtype=np.float32)
        y = np.empty((x.shape[0], self.n_out), np.float32)
        N.check(self._lib.gcm_mlp_predict(self._h, x.ctypes.data, x.shape[0],
                                          y.ctypes.data))
        return y

    def gradients(self, x: np.ndarray, t: np.ndarray):
        x = np.ascontiguousarray(x, dtype=np.float32)
        t = np.ascontiguousarray(t, dtype=np.float32)
        g = np.empty(self.n_params, np.float32)
        loss = C.c_double()
        N.check(self._lib.gcm_mlp_gradients(self._h, x.ctypes.data, t.ctypes.data,
                                            x.shape[0], g.ctypes.data, C.byref(loss)))
        return loss.value, g

    def set_data(self, x: np.ndarray, t: np.ndarray) -> None:
        x = np.ascontiguousarray(x, dtype=np.float32)
        t = np.ascontiguousarray(t, dtype=np.float32)
        N.check(self._lib.gcm_mlp_set_data(self._h, x.ctypes.data, t.ctypes.data,
                                           x.shape[0]))

    def train(self, epochs: int, batch: int = 256, lr: float = 1e-3,
              seed: int = 0) -> np.ndarray:
        losses = np.zeros(max(1, epochs))
        N.check(self._lib.gcm_mlp_train(self._h, epochs, batch, lr, seed,
                                        losses.ctypes.data))
        return losses[:epochs]

    def close(self):
        if getattr(self, "_h", None):
            self._lib.gcm_mlp_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Surrogate:
    """Parameter point -> predicted bin contents (sum of weights), via an MLP
    on standardised inputs and chi2-weighted log contents.

    Targets: for a bin with mean content cbar_b over the training set, the
    network learns t_b = (log c_b - m_b) * sqrt(cbar_b) / G, where m_b is the
    mean log content and G one global scale.  An error e in t_b is an error
    e G / sqrt(cbar_b) in log c_b, which moves the Poisson chi2 of the bin by
    about cbar_b (e G / sqrt(cbar_b))^2 = (e G)^2: the mean squared error is
    proportional to the chi2 error, so the network spends its capacity on the
    bins the likelihood is sensitive to.  Bins without events in the training
    set are predicted empty."""

    def __init__(self, n_par: int, n_bins: int, hidden: int = 512, seed: int = 0,
                 device: int = 0):
        self.mlp = MLP(n_par, hidden, n_bins, seed=seed, device=device)
        self.n_par, self.n_bins = n_par, n_bins
        self.x_mu = np.zeros(n_par)
        self.x_sd = np.ones(n_par)
        self.mean = np.zeros(n_bins)
        self.scale = np.ones(n_bins)
        self.head_ready = False

    def _x(self, theta):
        return ((np.atleast_2d(theta) - self.x_mu) / self.x_sd).astype(np.float32)

    def targets(self, sums: np.ndarray) -> np.ndarray:
        """Fix the output normalisation from a training set of bin contents
        and return the normalised targets."""
        sums = np.asarray(sums, np.float64)
        cbar = sums.mean(0)
        live = cbar > 0
        floor = np.where(live, 1e-6 * cbar, 1.0)
        logc = np.log(np.maximum(sums, floor))
        m = logc.mean(0)
        u = np.sqrt(np.where(live, cbar, 0.0))
        dev = (logc - m) * u
        G = float(np.sqrt(np.mean(dev[:, live] ** 2))) if live.any() else 1.0
        G = G if G > 0 else 1.0
        self.mean = np.where(live, m, -np.inf)
        self.scale = np.where(live, G / np.where(live, u, 1.0), 0.0)
        return (dev / G).astype(np.float32)

    def fit(self, theta: np.ndarray, sums: np.ndarray, epochs: int = 30,
            batch: int = 256, lr: float = 1e-3, lr_final: Optional[float] = None,
            seed: int = 0, log=None, log_every: int = 25) -> np.ndarray:
        """Adam epochs on (theta, bin contents) pairs; the learning rate decays
        geometrically from lr to lr_final (default lr / 30).  Returns the
        sampled training loss of each epoch; log(message), when given, gets a
        progress line every log_every epochs."""
        theta = np.atleast_2d(np.asarray(theta, np.float64))
        self.x_mu = theta.mean(0)
        sd = theta.std(0)
        self.x_sd = np.where(sd > 0, sd, 1.0)
        t = self.targets(sums)
        self.mlp.set_data(self._x(theta), t)
        lr_final = lr / 30.0 if lr_final is None else lr_final
        losses = np.empty(epochs)
        for e in range(epochs):
            le = lr * (lr_final / lr) ** (e / max(1, epochs - 1))
            losses[e] = self.mlp.train(1, batch, le, seed * 100003 + e)[0]
            if log is not None and (e + 1) % max(1, log_every) == 0:
                log(f"surrogate epoch {e + 1}/{epochs}: loss {losses[e]:.4f}")
        self.head_ready = False
        return losses

    def predict_sums(self, theta: np.ndarray) -> np.ndarray:
        y = self.mlp.predict(self._x(theta)).astype(np.float64)
        with np.errstate(invalid="ignore"):
            out = np.exp(self.mean + self.scale * y)
        return np.where(np.isfinite(self.mean), out, 0.0)

    def set_head(self, data: np.ndarray, var_ratio: np.ndarray) -> None:
        """Data histogram and per-bin (sum w^2 / sum w) ratio for chi2()."""
        f = lambda a: np.ascontiguousarray(a, dtype=np.float64)
        self._head = [f(self.mean), f(self.scale), f(data), f(var_ratio)]
        N.check(self.mlp._lib.gcm_mlp_set_head(self.mlp._h, *[a.ctypes.data for a in self._head]))
        self.head_ready = True

    def chi2(self, theta: np.ndarray) -> np.ndarray:
        """Barlow-Beeston chi2 of the predicted contents, on the device."""
        if not self.head_ready:
            raise RuntimeError("Surrogate.set_head after fit")
        x = self._x(theta)
        out = np.empty(x.shape[0])
        N.check(self.mlp._lib.gcm_mlp_predict_llh(self.mlp._h, x.ctypes.data, x.shape[0],
                                                 out.ctypes.data))
        return out


def barlow_llh(data: np.ndarray, pred: np.ndarray, mcvar: np.ndarray) -> np.ndarray:
    """Barlow-Beeston OA2021 per row (BarlowBeestonBanff2022.h:150-294, the
    default options), vectorised over parameter points: rows of pred/mcvar."""
    data = np.broadcast_to(data, pred.shape)
    with np.errstate(divide="ignore", invalid="ignore", over="ignore"):
        frac = np.minimum(np.sqrt(mcvar) / pred, 1e150)
        temp = pred * frac * frac - 1
        temp2 = temp * temp + 4 * data * frac * frac
        beta = (-temp + np.sqrt(temp2)) / 2.0
        has = mcvar > 0
        newmc = np.where(has, np.maximum(pred * beta, np.finfo(float).tiny), pred)
        pen = np.where(has, (beta - 1) ** 2 / (2 * frac * frac), 0.0)
        tiny = np.finfo(float).tiny
        eps = np.finfo(float).eps
        stat = np.where(data <= eps, newmc,
                        np.where(newmc > tiny,
                                 newmc - data + data * (np.log(np.maximum(data, tiny))
                                                        - np.log(np.maximum(newmc, tiny))),
                                 tiny - data + data * (np.log(np.maximum(data, tiny)) - np.log(tiny))))
        chisq = 2.0 * stat + 2.0 * pen
    return chisq.sum(-1)
