"""The warm-start network on the device (libmcx mcx_mlp_*): 8 features ->
4 x 256 SiLU -> log T, BF16 matrix-core layers with FP32 accumulation.

`Network` owns one device handle; parameters are the flat FP32 vector of
include/mcx.h.  Inputs and targets are normalised by the statistics stored
with the network (`x_mean`, `x_std`, `t_mean`, `t_std`), so the device sees
features and targets of order one.
"""
import ctypes as C

import numpy as np

from . import _lib

NIN, WIDTH, DEPTH, NPARAMS = _lib.MLP_NIN, _lib.MLP_WIDTH, _lib.MLP_DEPTH, _lib.MLP_NPARAMS


def init_params(seed):
    """Uniform weights with the fan-in bound sqrt(6/fan_in)/1.1 (He's rule
    for a SiLU that passes about half the variance), zero biases"""
    rng = np.random.default_rng(seed)
    p = np.zeros(NPARAMS, dtype=np.float32)
    o = 0
    blocks = [(NIN, WIDTH), (WIDTH,)] + [(WIDTH, WIDTH), (WIDTH,)] * (DEPTH - 1) + [(WIDTH,), (1,)]
    for k, s in enumerate(blocks):
        n = int(np.prod(s))
        if len(s) == 2 or k == len(blocks) - 2:
            bound = np.sqrt(6.0 / s[0]) / 1.1
            p[o:o + n] = rng.uniform(-bound, bound, size=n)
        o += n
    assert o == NPARAMS
    return p


class Network:
    def __init__(self, device=0, params=None, seed=1):
        self.lib = _lib.load()
        h = C.c_void_p()
        rc = self.lib.mcx_mlp_create(C.byref(h), device, NIN, WIDTH, DEPTH)
        if rc != 0:
            raise RuntimeError(f"mcx_mlp_create failed ({rc})")
        self.h = h
        self.x_mean = np.zeros(NIN)
        self.x_std = np.ones(NIN)
        self.t_mean, self.t_std = 0.0, 1.0
        self.set_params(init_params(seed) if params is None else params)

    def _chk(self, rc, what):
        if rc != 0:
            msg = self.lib.mcx_mlp_last_error(self.h).decode()
            raise RuntimeError(f"{what} failed ({rc}): {msg}")

    def close(self):
        if self.h:
            self.lib.mcx_mlp_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_params(self, p):
        p = np.ascontiguousarray(p, dtype=np.float32)
        assert p.size == NPARAMS
        self._chk(self.lib.mcx_mlp_set_params(self.h, _lib.ptr(p)), "mcx_mlp_set_params")

    def params(self):
        p = np.empty(NPARAMS, dtype=np.float32)
        self._chk(self.lib.mcx_mlp_get_params(self.h, _lib.ptr(p)), "mcx_mlp_get_params")
        return p

    # raw (normalised) interface -------------------------------------------
    def forward_raw(self, x):
        x = np.ascontiguousarray(x, dtype=np.float32)
        y = np.empty(x.shape[0], dtype=np.float32)
        self._chk(self.lib.mcx_mlp_forward(self.h, x.shape[0], _lib.ptr(x), _lib.ptr(y)), "mcx_mlp_forward")
        return y

    def gradient_raw(self, x, t):
        x = np.ascontiguousarray(x, dtype=np.float32)
        t = np.ascontiguousarray(t, dtype=np.float32)
        g = np.empty(NPARAMS, dtype=np.float32)
        loss = C.c_double()
        self._chk(self.lib.mcx_mlp_gradient(self.h, x.shape[0], _lib.ptr(x), _lib.ptr(t), C.byref(loss),
                                            _lib.ptr(g)), "mcx_mlp_gradient")
        return loss.value, g

    def set_data_raw(self, x, t):
        x = np.ascontiguousarray(x, dtype=np.float32)
        t = np.ascontiguousarray(t, dtype=np.float32)
        self._chk(self.lib.mcx_mlp_set_data(self.h, x.shape[0], _lib.ptr(x), _lib.ptr(t)), "mcx_mlp_set_data")

    def train(self, nsteps, batch=8192, lr=1e-3, beta1=0.9, beta2=0.999, eps=1e-8, seed=1):
        """nsteps Adam steps on the data of set_data; returns the per-step loss"""
        o = _lib.MlpTrainOpts(batch=batch, nsteps=nsteps, lr=lr, beta1=beta1, beta2=beta2, eps=eps, seed=seed)
        loss = np.empty(max(1, nsteps), dtype=np.float32)
        self._chk(self.lib.mcx_mlp_train(self.h, C.byref(o), _lib.ptr(loss)), "mcx_mlp_train")
        return loss[:nsteps]

    def timing(self):
        f, s = C.c_double(), C.c_double()
        n = C.c_int64()
        self._chk(self.lib.mcx_mlp_last_timing(self.h, C.byref(f), C.byref(n), C.byref(s)), "mcx_mlp_last_timing")
        return dict(forward_ms=f.value, forward_rows=n.value, step_ms=s.value)

    # physical interface ----------------------------------------------------
    def fit_normalisation(self, X, t):
        self.x_mean = X.mean(axis=0)
        self.x_std = np.maximum(X.std(axis=0), 1e-6)
        self.t_mean, self.t_std = float(t.mean()), float(max(t.std(), 1e-6))

    def norm_x(self, X):
        return ((np.asarray(X) - self.x_mean) / self.x_std).astype(np.float32)

    def predict(self, X):
        """log T for features X [n][8] (physical units)"""
        return self.forward_raw(self.norm_x(X)).astype(np.float64) * self.t_std + self.t_mean

    def set_data(self, X, t):
        self.set_data_raw(self.norm_x(X), ((np.asarray(t) - self.t_mean) / self.t_std).astype(np.float32))

    def save(self, path):
        np.savez(path, params=self.params(), x_mean=self.x_mean, x_std=self.x_std,
                 t_mean=np.float64(self.t_mean), t_std=np.float64(self.t_std))

    @classmethod
    def load(cls, path, device=0):
        d = np.load(path, allow_pickle=False)
        net = cls(device=device, params=d["params"])
        net.x_mean, net.x_std = d["x_mean"], d["x_std"]
        net.t_mean, net.t_std = float(d["t_mean"]), float(d["t_std"])
        return net
