"""ctypes binding of the learned moment denoiser (include/rh_denoise.h).

Device volumes are passed as torch CUDA tensors (float32, contiguous,
[P][Z][Y][X][C]); torch only provides the device memory, every kernel is the
library's own.  The denoiser has no counterpart in the reference: it
corrects the noisy deposited moments (n, J) of a low-ppc run towards those
of a high-ppc run before the field update.
"""
from __future__ import annotations

import ctypes as C
from typing import Optional, Sequence, Tuple

import numpy as np

from .gpu import load_library

N_WEIGHTS = 64800
LAYERS = ((7, 32), (32, 32), (32, 32), (32, 4))  # (cin, cout), 3x3x3 kernels, no biases
RH_DN_OK, RH_DN_ERR_ARG, RH_DN_ERR_HIP, RH_DN_ERR_CAPACITY = 0, -1, -2, -4
STATIC_MAX_VOXELS = 262144  # RH_DN_STATIC_MAX_VOXELS: the static scratch of rh_dn_forward_static


class DenoiserError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__("%s (code %d)" % (msg, code))
        self.code = code


_bound = False


def _lib():
    global _bound
    L = load_library()
    if not _bound:
        vp, i32, i64, f32 = C.c_void_p, C.c_int32, C.c_int64, C.c_float
        L.rh_dn_create.argtypes = [i64, C.POINTER(vp)]
        L.rh_dn_destroy.argtypes = [vp]
        L.rh_dn_destroy.restype = None
        L.rh_dn_last_error.argtypes = [vp]
        L.rh_dn_last_error.restype = C.c_char_p
        L.rh_dn_set_weights.argtypes = [vp, vp, i64]
        L.rh_dn_get_weights.argtypes = [vp, vp, i64]
        L.rh_dn_set_scales.argtypes = [vp, vp]
        L.rh_dn_forward.argtypes = [vp, vp, vp, i32, i32, i32, i32]
        L.rh_dn_forward_static.argtypes = [vp, vp, vp, i32, i32, i32, i32]
        L.rh_dn_train_step.argtypes = [vp, vp, vp, i32, i32, i32, i32, f32, C.POINTER(f32)]
        L.rh_dn_gradient.argtypes = [vp, vp, vp, i32, i32, i32, i32, vp, C.POINTER(f32)]
        L.rh_dn_synchronize.argtypes = [vp]
        L.rh_dn_set_stream.argtypes = [vp, vp]
        L.rh_dn_set_timing.argtypes = [vp, i32]
        L.rh_dn_kernel_time.argtypes = [vp, C.c_char_p, C.POINTER(C.c_double), C.POINTER(i64)]
        L.rh_dn_reset_timing.argtypes = [vp]
        _bound = True
    return L


def init_weights(seed: int = 0, gain: float = 1.0) -> np.ndarray:
    """He-style initial weights in the library layout (layers concatenated,
    each [cout][cin][3][3][3])."""
    rng = np.random.default_rng(seed)
    parts = []
    for cin, cout in LAYERS:
        std = gain * np.sqrt(2.0 / (cin * 27))
        parts.append((rng.standard_normal((cout, cin, 3, 3, 3)) * std).astype(np.float32).ravel())
    return np.concatenate(parts)


def split_weights(w: np.ndarray):
    """The four layer tensors [cout][cin][3][3][3] of a flat weight vector."""
    out, off = [], 0
    for cin, cout in LAYERS:
        n = cout * cin * 27
        out.append(w[off:off + n].reshape(cout, cin, 3, 3, 3))
        off += n
    return out


class Denoiser:
    def __init__(self, max_voxels: int, scales: Optional[Sequence[float]] = None,
                 weights: Optional[np.ndarray] = None):
        L = _lib()
        h = C.c_void_p()
        rc = L.rh_dn_create(int(max_voxels), C.byref(h))
        if rc != RH_DN_OK:
            raise DenoiserError(rc, "rh_dn_create failed")
        self._h = h
        if scales is not None:
            self.set_scales(scales)
        self.set_weights(init_weights() if weights is None else weights)

    def _check(self, rc: int) -> None:
        if rc != RH_DN_OK:
            raise DenoiserError(rc, _lib().rh_dn_last_error(self._h).decode())

    def close(self) -> None:
        if getattr(self, "_h", None):
            _lib().rh_dn_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_weights(self, w: np.ndarray) -> None:
        w = np.ascontiguousarray(w, dtype=np.float32)
        self._check(_lib().rh_dn_set_weights(self._h, w.ctypes.data, w.size))

    def weights(self) -> np.ndarray:
        w = np.empty(N_WEIGHTS, dtype=np.float32)
        self._check(_lib().rh_dn_get_weights(self._h, w.ctypes.data, w.size))
        return w

    def set_scales(self, scales: Sequence[float]) -> None:
        s = np.ascontiguousarray(scales, dtype=np.float32)
        if s.size != 7:
            raise ValueError("7 channel scales expected")
        self._check(_lib().rh_dn_set_scales(self._h, s.ctypes.data))

    @staticmethod
    def _dims(x, channels: int) -> Tuple[int, int, int, int]:
        if x.dim() != 5 or x.shape[-1] != channels or not x.is_contiguous() or not x.is_cuda:
            raise ValueError("expected a contiguous CUDA tensor [P][Z][Y][X][%d]" % channels)
        P, Z, Y, X, _ = x.shape
        return P, X, Y, Z

    def forward(self, x, y=None):
        import torch
        P, X, Y, Z = self._dims(x, 7)
        if y is None:
            y = torch.empty((P, Z, Y, X, 4), dtype=torch.float32, device=x.device)
        self._check(_lib().rh_dn_forward(self._h, x.data_ptr(), y.data_ptr(), P, X, Y, Z))
        return y

    def forward_static(self, x, y):
        """forward() through rh_dn_forward_static: scratch in the library's
        static device array (at most STATIC_MAX_VOXELS voxels), nothing
        allocated; the output tensor y must be given."""
        P, X, Y, Z = self._dims(x, 7)
        if tuple(y.shape) != (P, Z, Y, X, 4):
            raise ValueError("y must have shape [P][Z][Y][X][4]")
        self._dims(y, 4)
        self._check(_lib().rh_dn_forward_static(self._h, x.data_ptr(), y.data_ptr(), P, X, Y, Z))
        return y

    def gradient(self, x, target) -> Tuple[np.ndarray, float]:
        P, X, Y, Z = self._dims(x, 7)
        self._dims(target, 4)
        g = np.empty(N_WEIGHTS, dtype=np.float32)
        loss = C.c_float()
        self._check(_lib().rh_dn_gradient(self._h, x.data_ptr(), target.data_ptr(), P, X, Y, Z,
                                          g.ctypes.data, C.byref(loss)))
        return g, float(loss.value)

    def train_step(self, x, target, lr: float) -> float:
        P, X, Y, Z = self._dims(x, 7)
        self._dims(target, 4)
        loss = C.c_float()
        self._check(_lib().rh_dn_train_step(self._h, x.data_ptr(), target.data_ptr(), P, X, Y, Z,
                                            float(lr), C.byref(loss)))
        return float(loss.value)

    def synchronize(self) -> None:
        self._check(_lib().rh_dn_synchronize(self._h))

    def set_timing(self, on: bool) -> None:
        self._check(_lib().rh_dn_set_timing(self._h, int(on)))

    def reset_timing(self) -> None:
        self._check(_lib().rh_dn_reset_timing(self._h))

    def kernel_time(self, name: str) -> Tuple[float, int]:
        ms, n = C.c_double(), C.c_int64()
        self._check(_lib().rh_dn_kernel_time(self._h, name.encode(), C.byref(ms), C.byref(n)))
        return ms.value, n.value


def forward_flops(voxels: int) -> float:
    """Multiply-adds x2 of one forward pass (the 3x3x3 convolutions)."""
    return 2.0 * voxels * 27 * sum(ci * co for ci, co in LAYERS)
