"""The learned cost model: a 3-layer MLP (in -> 256 -> 256 -> 1) whose
forward, backward and Adam steps run in libgat's HIP kernels (BF16 matrix
cores, FP32 accumulation). It regresses log(runtime) and is used only to
rank candidate configs; the tuner re-measures every pick."""
import ctypes as C

import numpy as np

from .._lib import check, lib


class MLPCostModel:
    def __init__(self, in_dim, hidden=256, seed=0):
        self.in_dim, self.hidden = in_dim, hidden
        h = C.c_void_p()
        check(lib().gat_mlp_create(in_dim, hidden, seed, C.byref(h)))
        self._h = h
        n = C.c_int64()
        check(lib().gat_mlp_param_count(self._h, C.byref(n)))
        self.n_params = n.value
        self.mu, self.sd = 0.0, 1.0
        self.trained = False

    def __del__(self):
        h = getattr(self, "_h", None)
        if h and h.value:
            lib().gat_mlp_free(h)
            self._h = None

    @staticmethod
    def _f32(a):
        return np.ascontiguousarray(a, dtype=np.float32)

    def params(self):
        p = np.zeros(self.n_params, dtype=np.float32)
        check(lib().gat_mlp_get_params(self._h, p.ctypes.data, self.n_params))
        return p

    def set_params(self, p):
        p = self._f32(p)
        check(lib().gat_mlp_set_params(self._h, p.ctypes.data, self.n_params))

    def forward(self, X):
        X = self._f32(X)
        y = np.zeros(len(X), dtype=np.float32)
        check(lib().gat_mlp_forward(self._h, X.ctypes.data, len(X), y.ctypes.data))
        return y

    def gradients(self, X, t):
        X, t = self._f32(X), self._f32(t)
        g = np.zeros(self.n_params, dtype=np.float32)
        loss = C.c_float()
        check(lib().gat_mlp_gradients(self._h, X.ctypes.data, t.ctypes.data, len(X), g.ctypes.data, C.byref(loss)))
        return g, loss.value

    def train_raw(self, X, t, steps, lr):
        X, t = self._f32(X), self._f32(t)
        loss = C.c_float()
        check(lib().gat_mlp_train(self._h, X.ctypes.data, t.ctypes.data, len(X), steps, lr, C.byref(loss)))
        return loss.value

    def fit(self, X, seconds, steps=400, lr=2e-3):
        y = np.log(np.maximum(np.asarray(seconds, dtype=np.float64), 1e-9))
        self.mu, self.sd = float(y.mean()), float(y.std() + 1e-6)
        loss = self.train_raw(X, (y - self.mu) / self.sd, steps, lr)
        self.trained = True
        return loss

    def predict(self, X):
        """Predicted seconds."""
        return np.exp(self.forward(X).astype(np.float64) * self.sd + self.mu)


def reference_forward(p, X, in_dim, hidden):
    """FP32 CPU restatement with bf16-rounded matmul operands (checker)."""
    W1, b1, W2, b2, W3, b3 = unpack(p, in_dim, hidden)
    r = bf16_round
    z1 = r(X) @ r(W1) + b1
    a1 = np.maximum(z1, 0)
    z2 = r(a1) @ r(W2) + b2
    a2 = np.maximum(z2, 0)
    y = r(a2) @ r(W3) + b3
    return y, (z1, a1, z2, a2)


def reference_gradients(p, X, t, in_dim, hidden):
    W1, b1, W2, b2, W3, b3 = unpack(p, in_dim, hidden)
    r = bf16_round
    y, (z1, a1, z2, a2) = reference_forward(p, X, in_dim, hidden)
    B = len(X)
    dy = 2.0 * (y - t) / B
    gW3 = r(a2).T @ r(dy)
    gb3 = dy.sum(keepdims=True)
    d2 = np.outer(r(dy), r(W3)) * (z2 > 0)
    gW2 = r(a1).T @ r(d2)
    gb2 = d2.sum(0)
    d1 = (r(d2) @ r(W2).T) * (z1 > 0)
    gW1 = r(X).T @ r(d1)
    gb1 = d1.sum(0)
    loss = float(np.mean((y - t) ** 2))
    return np.concatenate([gW1.ravel(), gb1, gW2.ravel(), gb2, gW3.ravel(), gb3.ravel()]).astype(np.float32), loss


def unpack(p, in_dim, hidden):
    o = 0
    W1 = p[o:o + in_dim * hidden].reshape(in_dim, hidden); o += in_dim * hidden
    b1 = p[o:o + hidden]; o += hidden
    W2 = p[o:o + hidden * hidden].reshape(hidden, hidden); o += hidden * hidden
    b2 = p[o:o + hidden]; o += hidden
    W3 = p[o:o + hidden]; o += hidden
    b3 = p[o:o + 1]
    return W1.astype(np.float64), b1.astype(np.float64), W2.astype(np.float64), b2.astype(np.float64), \
        W3.astype(np.float64), b3.astype(np.float64)


def bf16_round(a):
    """Round float values to bfloat16 (nearest even), returned as float64."""
    f = np.asarray(a, dtype=np.float32)
    u = f.view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000
    return u.astype(np.uint32).view(np.float32).astype(np.float64)
