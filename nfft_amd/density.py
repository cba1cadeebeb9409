"""Density-compensation weights for the weighted CGNR, through the C ABI of
``libnfft3_hip.so`` (declared in include/nfft3_hip.h; no counterpart in the
reference, whose MRI applications read weights from files).

* ``pipe_menon(plan, iters)``: the Pipe-Menon fixed point w <- w / (B B^T w)
  with the plan's own Kaiser-Bessel interpolation and spreading.
* ``features(plan)``: per-node density features (radius, log node counts
  of 5 x 5 coarse-grid neighbourhoods at three scales, and the kernel
  density estimate log(v / mean v) with v = B B^T 1), a (M, 80) float32
  device tensor.
* ``DcfModel``: the MLP that maps features to log w, trained and evaluated by
  the library's BF16 MFMA kernels.

Device buffers are torch CUDA tensors (torch is used for device memory only);
the plan must be a float-family ``nfft_amd.Plan`` with d = 2 for the features.
"""
import ctypes

import numpy as np

from . import lib

_P = ctypes.c_void_p


def _L():
    L = lib()
    if not getattr(L, "_dcf_ready", False):
        L.nfft_dcf_num_features.restype = ctypes.c_int
        L.nfft_dcf_num_params.restype = ctypes.c_int
        L.nfft_dcf_model_create.restype = _P
        L.nfft_dcf_model_create.argtypes = [ctypes.c_uint]
        L.nfft_dcf_model_destroy.argtypes = [_P]
        L.nfft_dcf_model_get_params.argtypes = [_P, _P]
        L.nfft_dcf_model_set_params.argtypes = [_P, _P]
        L.nfft_dcf_model_loss_grad.restype = ctypes.c_float
        L.nfft_dcf_model_loss_grad.argtypes = [_P, _P, _P, ctypes.c_int, _P]
        L.nfft_dcf_model_train_step.restype = ctypes.c_float
        L.nfft_dcf_model_train_step.argtypes = [_P, _P, _P, ctypes.c_int, ctypes.c_float]
        L.nfft_dcf_model_predict.argtypes = [_P, _P, ctypes.c_int, _P]
        L.nfftf_density_pipe_menon.argtypes = [_P, _P, ctypes.c_int]
        L.nfftf_density_pipe_menon_device.argtypes = [_P, _P, ctypes.c_int]
        L.nfftf_dcf_features_device.argtypes = [_P, _P]
        L.nfftf_density_learned_device.argtypes = [_P, _P, _P]
        L._dcf_ready = True
    return L


def num_features():
    return _L().nfft_dcf_num_features()


def num_params():
    return _L().nfft_dcf_num_params()


def _plan_ptr(plan):
    if plan.precision != "single":
        raise ValueError("density weights are computed by float-family plans")
    return ctypes.cast(plan.pp, _P)


def pipe_menon(plan, iters=20, w0=None):
    """Pipe-Menon weights of the plan's nodes (host float32 array)."""
    w = np.ones(plan.M, dtype=np.float32) if w0 is None else np.array(w0, dtype=np.float32)
    _L().nfftf_density_pipe_menon(_plan_ptr(plan), w.ctypes.data, int(iters))
    return w


def pipe_menon_device(plan, w_dev, iters=20):
    """In place on a float32 CUDA tensor of M values."""
    _L().nfftf_density_pipe_menon_device(_plan_ptr(plan), w_dev.data_ptr(), int(iters))
    return w_dev


def features(plan):
    import torch
    feat = torch.empty((plan.M, num_features()), dtype=torch.float32, device="cuda")
    _L().nfftf_dcf_features_device(_plan_ptr(plan), feat.data_ptr())
    return feat


class DcfModel:
    """The density MLP (80 -> 64 -> 64 -> 1) held by the library."""

    def __init__(self, seed=0, params=None):
        self.L = _L()
        self.h = self.L.nfft_dcf_model_create(int(seed))
        if params is not None:
            self.set_params(params)

    @property
    def params(self):
        p = np.empty(num_params(), dtype=np.float32)
        self.L.nfft_dcf_model_get_params(self.h, p.ctypes.data)
        return p

    def set_params(self, p):
        p = np.ascontiguousarray(p, dtype=np.float32)
        assert p.size == num_params()
        self.L.nfft_dcf_model_set_params(self.h, p.ctypes.data)

    def loss_grad(self, feat, target):
        g = np.empty(num_params(), dtype=np.float32)
        loss = self.L.nfft_dcf_model_loss_grad(self.h, feat.data_ptr(), target.data_ptr(),
                                               int(feat.shape[0]), g.ctypes.data)
        return float(loss), g

    def train_step(self, feat, target, lr):
        return float(self.L.nfft_dcf_model_train_step(self.h, feat.data_ptr(), target.data_ptr(),
                                                      int(feat.shape[0]), float(lr)))

    def predict(self, feat):
        import torch
        out = torch.empty(feat.shape[0], dtype=torch.float32, device=feat.device)
        self.L.nfft_dcf_model_predict(self.h, feat.data_ptr(), int(feat.shape[0]), out.data_ptr())
        return out

    def weights(self, plan):
        """exp(prediction) on the plan's nodes, a float32 CUDA tensor."""
        import torch
        w = torch.empty(plan.M, dtype=torch.float32, device="cuda")
        self.L.nfftf_density_learned_device(_plan_ptr(plan), self.h, w.data_ptr())
        return w

    def save(self, path):
        np.save(path, self.params)

    @classmethod
    def load(cls, path):
        return cls(params=np.load(path, allow_pickle=False))

    def close(self):
        if self.h:
            self.L.nfft_dcf_model_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
