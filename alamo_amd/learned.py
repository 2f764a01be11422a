"""The learned coarse-level solver: Python mirror of include/alamo_unet.h.

UNet wraps one network handle of libalamo_elastic.so: a 3D U-Net (four
levels, channels 16-32-64-128, bfloat16 weights and activations on the matrix
cores, float32 accumulation) whose forward pass, backward pass, loss and Adam
step are the library's HIP kernels.  Fields are float32 CUDA tensors shaped
(components, nz, ny, nx), node index i fastest; there is no CPU fallback.
"""
from __future__ import annotations

import ctypes

import numpy as np

from . import elastic as _el
from .elastic import AlamoError

NN_EXPORTS = [
    "alamo_nn_create", "alamo_nn_destroy", "alamo_nn_num_params", "alamo_nn_shape", "alamo_nn_init_params",
    "alamo_nn_set_params", "alamo_nn_get_params", "alamo_nn_forward", "alamo_nn_backward", "alamo_nn_mse",
    "alamo_nn_adam", "alamo_nn_train_step", "alamo_nn_tape",
]

_bound = False


def lib():
    global _bound
    L = _el.lib()
    if not _bound:
        vp, ci, cd = ctypes.c_void_p, ctypes.c_int, ctypes.c_double
        for name in NN_EXPORTS:
            getattr(L, name)
        L.alamo_nn_create.argtypes = [ctypes.POINTER(ci), ci, ci, ctypes.POINTER(vp)]
        L.alamo_nn_destroy.argtypes = [vp]
        L.alamo_nn_num_params.argtypes = [vp, ctypes.POINTER(ctypes.c_long)]
        L.alamo_nn_shape.argtypes = [vp, ctypes.POINTER(ci), ctypes.POINTER(ci), ctypes.POINTER(ci)]
        L.alamo_nn_init_params.argtypes = [vp, ctypes.c_ulonglong, vp]
        L.alamo_nn_set_params.argtypes = [vp, vp, vp]
        L.alamo_nn_get_params.argtypes = [vp, vp, vp]
        L.alamo_nn_forward.argtypes = [vp, vp, vp, vp]
        L.alamo_nn_backward.argtypes = [vp, vp, vp, vp, vp]
        L.alamo_nn_mse.argtypes = [vp, vp, vp, vp, ctypes.POINTER(cd), vp]
        L.alamo_nn_adam.argtypes = [vp, vp, ctypes.c_float, vp]
        L.alamo_nn_train_step.argtypes = [vp, vp, vp, ctypes.c_float, ctypes.POINTER(cd), vp]
        L.alamo_nn_tape.argtypes = [vp, ci, vp, vp, vp]
        _bound = True
    return L


def _f32(t, shape=None):
    torch = _el._torch()
    if not isinstance(t, torch.Tensor) or not t.is_cuda or t.dtype != torch.float32 or not t.is_contiguous():
        raise AlamoError("expected a contiguous float32 CUDA tensor")
    if shape is not None and tuple(t.shape) != tuple(shape):
        raise AlamoError(f"expected shape {tuple(shape)}, got {tuple(t.shape)}")
    return ctypes.c_void_p(t.data_ptr())


class UNet:
    """nodes = (nx, ny, nz) of the finest network level (each minus 1 a
    multiple of 8); n_in input and n_out output components."""

    def __init__(self, nodes, n_in: int, n_out: int = 3, seed: int | None = 0):
        _el._torch()
        nodes = np.broadcast_to(np.asarray(nodes, dtype=np.int32), (3,)).copy()
        h = ctypes.c_void_p()
        _el._check(lib().alamo_nn_create(nodes.ctypes.data_as(ctypes.POINTER(ctypes.c_int)), int(n_in), int(n_out),
                                         ctypes.byref(h)), "alamo_nn_create")
        self.h = h
        self.nodes = tuple(int(v) for v in nodes)
        self.n_in, self.n_out = int(n_in), int(n_out)
        n = ctypes.c_long()
        _el._check(lib().alamo_nn_num_params(h, ctypes.byref(n)), "alamo_nn_num_params")
        self.num_params = int(n.value)
        if seed is not None:
            _el._check(lib().alamo_nn_init_params(h, int(seed), _el._stream()), "alamo_nn_init_params")

    def __del__(self):
        h = getattr(self, "h", None)
        if h:
            try:
                lib().alamo_nn_destroy(h)
            except Exception:
                pass
            self.h = None

    @property
    def field_shape(self):
        return (self.nodes[2], self.nodes[1], self.nodes[0])

    CHANNELS = (16, 32, 64, 128)

    def conv_table(self):
        """(cin, cout, level) of the 15 convolutions, in parameter order:
        two per encoder level, two per decoder level (2, 1, 0), the output."""
        ch = self.CHANNELS
        t = []
        for lev in range(4):
            t += [(self.n_in if lev == 0 else ch[lev - 1], ch[lev], lev), (ch[lev], ch[lev], lev)]
        for lev in (2, 1, 0):
            t += [(ch[lev + 1] + ch[lev], ch[lev], lev), (ch[lev], ch[lev], lev)]
        return t + [(ch[0], self.n_out, 0)]

    def level_shape(self, lev):
        n = list(self.nodes)
        for _ in range(lev):
            n = [(v - 1) // 2 + 1 for v in n]
        return (n[2], n[1], n[0])

    def tape(self, conv):
        """After forward(): the input (cin, nz, ny, nx) of convolution `conv`
        and its output after the ReLU (cout, nz, ny, nx; None for the output
        layer), as the forward pass kept them (alamo_nn_tape)."""
        torch = _el._torch()
        cin, cout, lev = self.conv_table()[conv]
        shape = self.level_shape(lev)
        x = torch.empty((cin,) + shape, dtype=torch.float32, device="cuda")
        y = None if conv == len(self.conv_table()) - 1 else torch.empty((cout,) + shape, dtype=torch.float32,
                                                                            device="cuda")
        _el._check(lib().alamo_nn_tape(self.h, int(conv), _f32(x), _f32(y) if y is not None else None,
                                       _el._stream()), "alamo_nn_tape")
        return x, y

    def params(self):
        torch = _el._torch()
        p = torch.empty(self.num_params, dtype=torch.float32, device="cuda")
        _el._check(lib().alamo_nn_get_params(self.h, _f32(p), _el._stream()), "alamo_nn_get_params")
        return p

    def set_params(self, p):
        _el._check(lib().alamo_nn_set_params(self.h, _f32(p, (self.num_params,)), _el._stream()),
                   "alamo_nn_set_params")

    def forward(self, x, out=None):
        torch = _el._torch()
        if out is None:
            out = torch.empty((self.n_out,) + self.field_shape, dtype=torch.float32, device="cuda")
        _el._check(lib().alamo_nn_forward(self.h, _f32(x, (self.n_in,) + self.field_shape),
                                          _f32(out, (self.n_out,) + self.field_shape), _el._stream()),
                   "alamo_nn_forward")
        return out

    __call__ = forward

    def backward(self, dout, want_input_grad=False):
        """Gradients of <out, dout> after forward(): (dparams, dinput or None)."""
        torch = _el._torch()
        dp = torch.empty(self.num_params, dtype=torch.float32, device="cuda")
        din = torch.empty((self.n_in,) + self.field_shape, dtype=torch.float32, device="cuda") if want_input_grad else None
        _el._check(lib().alamo_nn_backward(self.h, _f32(dout, (self.n_out,) + self.field_shape), _f32(dp),
                                           _f32(din) if din is not None else None, _el._stream()),
                   "alamo_nn_backward")
        return dp, din

    def mse(self, out, target, want_grad=False):
        torch = _el._torch()
        d = torch.empty_like(out) if want_grad else None
        loss = ctypes.c_double()
        _el._check(lib().alamo_nn_mse(self.h, _f32(out), _f32(target, out.shape), _f32(d) if d is not None else None,
                                      ctypes.byref(loss), _el._stream()), "alamo_nn_mse")
        return loss.value, d

    def adam(self, dparams, lr):
        _el._check(lib().alamo_nn_adam(self.h, _f32(dparams, (self.num_params,)), float(lr), _el._stream()),
                   "alamo_nn_adam")

    def train_step(self, x, target, lr):
        loss = ctypes.c_double()
        _el._check(lib().alamo_nn_train_step(self.h, _f32(x, (self.n_in,) + self.field_shape),
                                             _f32(target, (self.n_out,) + self.field_shape), float(lr),
                                             ctypes.byref(loss), _el._stream()), "alamo_nn_train_step")
        return loss.value
