"""The learned preconditioner's 3-D U-Net (csrc/cad_unet.hip) from Python.

The network: 3x3x3 convolutions with ReLU, channels 16 / 32 / 64 on three
levels (2x2x2 average pooling down, nearest-neighbour upsampling up, skip
connections by concatenation), a 1x1x1 output layer; 8 input channels, one
output channel, 390,913 parameters.  Parameters are one FP32 vector: per 3x3x3
layer its weights W[tap][cin][cout] (tap = kx + 3 ky + 9 kz) then its biases,
in the order of LAYERS, then the output layer's 16 weights and bias.

Inputs are BF16 tensors of shape (nb, nz+2, ny+2, nx+2, 8): channels last, a
zero halo of one voxel around each box.  Device memory comes from torch; the
kernels are the library's own.
"""
import ctypes as C

import numpy as np

from . import capi

# (level, cin, cout) of the 3x3x3 layers, in parameter order
LAYERS = ((0, 8, 16), (0, 16, 16), (1, 16, 32), (1, 32, 32), (2, 32, 64), (2, 64, 64), (2, 64, 32),
          (1, 64, 32), (1, 32, 32), (1, 32, 16), (0, 32, 16), (0, 16, 16))
NAMES = ("e1a", "e1b", "e2a", "e2b", "b3a", "b3b", "u3", "d2a", "d2b", "u2", "d1a", "d1b")
IN_CHANNELS = 8


def _bind():
    L = capi.lib()
    if getattr(L, "_unet_bound", False):
        return L
    vp = C.c_void_p
    L.cad_unet_param_count.restype = C.c_int
    L.cad_unet_create.argtypes = [C.c_int] * 5 + [vp, C.POINTER(vp)]
    L.cad_unet_destroy.argtypes = [vp]
    L.cad_unet_destroy.restype = None
    L.cad_unet_set_params.argtypes = [vp, C.POINTER(C.c_float)]
    L.cad_unet_get_params.argtypes = [vp, C.POINTER(C.c_float)]
    L.cad_unet_init_params.argtypes = [C.c_int, C.POINTER(C.c_float)]
    L.cad_unet_set_input.argtypes = [vp, vp]
    L.cad_unet_get_output.argtypes = [vp, vp]
    L.cad_unet_forward.argtypes = [vp]
    L.cad_unet_train_step.argtypes = [vp, vp, vp, C.c_float, C.POINTER(C.c_float)]
    L.cad_unet_gradients.argtypes = [vp, vp, vp, C.POINTER(C.c_float), C.POINTER(C.c_float)]
    L.cad_unet_train_step_dy.argtypes = [vp, vp, C.c_float]
    L.cad_unet_gradients_dy.argtypes = [vp, vp, C.POINTER(C.c_float)]
    L._unet_bound = True
    return L


def param_count():
    return int(_bind().cad_unet_param_count())


def layer_slices():
    """name -> (weight slice, bias slice, (cin, cout)) in the parameter vector."""
    out, o = {}, 0
    for name, (_, ci, co) in zip(NAMES, LAYERS):
        w = slice(o, o + 27 * ci * co)
        o += 27 * ci * co
        b = slice(o, o + co)
        o += co
        out[name] = (w, b, (ci, co))
    out["out"] = (slice(o, o + 16), slice(o + 16, o + 17), (16, 1))
    return out


def init_params(seed=0):
    """He-normal weights, zero biases (the library's initialiser)."""
    p = np.zeros(param_count(), np.float32)
    capi._check(_bind().cad_unet_init_params(int(seed), p.ctypes.data_as(C.POINTER(C.c_float))))
    return p


def _fp(a):
    return a.ctypes.data_as(C.POINTER(C.c_float))


class UNet:
    """A network for batches of nb boxes of nx x ny x nz voxels (nx a multiple
    of 64, ny and nz of 4); train=True also holds gradients and Adam state."""

    def __init__(self, nx, ny, nz, nb=1, train=False, stream=None):
        L = _bind()
        h = C.c_void_p()
        capi._check(L.cad_unet_create(int(nx), int(ny), int(nz), int(nb), int(bool(train)),
                                      stream, C.byref(h)))
        self._h = h
        self.box = (int(nx), int(ny), int(nz))
        self.nb = int(nb)

    def close(self):
        if getattr(self, "_h", None):
            _bind().cad_unet_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def set_params(self, p):
        p = np.ascontiguousarray(p, np.float32)
        assert p.size == param_count()
        capi._check(_bind().cad_unet_set_params(self._h, _fp(p)))

    def get_params(self):
        p = np.zeros(param_count(), np.float32)
        capi._check(_bind().cad_unet_get_params(self._h, _fp(p)))
        return p

    def set_input(self, x):
        """x: torch BF16 CUDA tensor (nb, nz+2, ny+2, nx+2, 8), zero halos."""
        nx, ny, nz = self.box
        assert tuple(x.shape) == (self.nb, nz + 2, ny + 2, nx + 2, IN_CHANNELS) and x.is_contiguous()
        capi._check(_bind().cad_unet_set_input(self._h, C.c_void_p(x.data_ptr())))

    def forward(self, out):
        """Runs the network on the input; out: torch FP32 CUDA tensor of
        nb*nz*ny*nx values (interior voxels, x fastest) receives the result."""
        L = _bind()
        capi._check(L.cad_unet_forward(self._h))
        capi._check(L.cad_unet_get_output(self._h, C.c_void_p(out.data_ptr())))
        return out

    def train_step(self, target, mask, lr):
        """One Adam step on loss = mean over boxes of sum m (y-e)^2 / sum m e^2
        (target e, mask m: FP32 CUDA tensors in the output's order); returns
        the loss before the step."""
        loss = C.c_float()
        capi._check(_bind().cad_unet_train_step(self._h, C.c_void_p(target.data_ptr()),
                                                C.c_void_p(mask.data_ptr()), float(lr),
                                                C.byref(loss)))
        return loss.value

    def train_step_dy(self, dy, lr):
        """One Adam step from dy, the gradient (FP32 CUDA tensor in the
        output's order) of a loss with respect to the last forward's output."""
        assert dy.is_contiguous() and dy.numel() == self.nb * self.box[0] * self.box[1] * self.box[2]
        capi._check(_bind().cad_unet_train_step_dy(self._h, C.c_void_p(dy.data_ptr()), float(lr)))

    def gradients_dy(self, dy):
        """Every parameter's gradient from dy (as for train_step_dy), no step."""
        assert dy.is_contiguous() and dy.numel() == self.nb * self.box[0] * self.box[1] * self.box[2]
        g = np.zeros(param_count(), np.float32)
        capi._check(_bind().cad_unet_gradients_dy(self._h, C.c_void_p(dy.data_ptr()), _fp(g)))
        return g

    def gradients(self, target, mask):
        """The loss and its gradient with respect to every parameter."""
        g = np.zeros(param_count(), np.float32)
        loss = C.c_float()
        capi._check(_bind().cad_unet_gradients(self._h, C.c_void_p(target.data_ptr()),
                                               C.c_void_p(mask.data_ptr()), _fp(g), C.byref(loss)))
        return loss.value, g
