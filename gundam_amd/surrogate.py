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
    on standardised inputs ((theta - prior mean) / prior sigma) and
    standardised log contents."""

    def __init__(self, n_par: int, n_bins: int, prior_mean: np.ndarray,
                 prior_sigma: np.ndarray, hidden: int = 512, seed: int = 0,
                 floor: float = 1e-3, device: int = 0):
        self.mlp = MLP(n_par, hidden, n_bins, seed=seed, device=device)
        self.mu = np.asarray(prior_mean, np.float64)
        self.sig = np.where(np.asarray(prior_sigma) > 0, prior_sigma, 1.0)
        self.floor = floor
        self.y_mean = np.zeros(n_bins)
        self.y_scale = np.ones(n_bins)

    def _x(self, theta):
        return ((np.atleast_2d(theta) - self.mu) / self.sig).astype(np.float32)

    def fit(self, theta: np.ndarray, sums: np.ndarray, epochs: int = 30,
            batch: int = 256, lr: float = 1e-3, seed: int = 0) -> np.ndarray:
        logy = np.log(np.maximum(sums, self.floor))
        self.y_mean = logy.mean(0)
        self.y_scale = np.maximum(logy.std(0), 1e-3)
        t = ((logy - self.y_mean) / self.y_scale).astype(np.float32)
        self.mlp.set_data(self._x(theta), t)
        return self.mlp.train(epochs, batch, lr, seed)

    def predict_sums(self, theta: np.ndarray) -> np.ndarray:
        y = self.mlp.predict(self._x(theta)).astype(np.float64)
        return np.exp(y * self.y_scale + self.y_mean)


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
