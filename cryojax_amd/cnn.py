"""Pose-proposal CNN on hand-written BF16 MFMA kernels (include/cryomi_cnn.h).

Network (no counterpart in the reference; cryoJAX has no learned pose
proposals): a 64x64 image through four 3x3 stride-2 convolutions (32, 64,
128, 128 channels, ReLU), then 2048 -> 512 (ReLU) -> 96 = eight rotation
hypotheses (unconstrained 3x3, orthonormalised at use), eight (x, y) offsets
and eight confidence logits. Activations are NHWC bfloat16; weights are kept
as float32 masters (Adam) with a bfloat16 copy that the GEMMs read; every
product accumulates in float32 on MFMA.

torch only allocates the device buffers and supplies the stream; every
operation is a libcryomi call.
"""

from __future__ import annotations

import ctypes
import math

import numpy as np
import torch

from . import _lib

CONV = [(1, 32, 64), (32, 64, 32), (64, 128, 16), (128, 128, 8)]  # (c_in, c_out, input side)
FC = [(2048, 512), (512, 96)]
N_HYP = 8
N_OUT = 96
IMAGE = 64


def kpad(c_in):
    return ((9 * c_in + 7) // 8) * 8


def layer_shapes():
    out = []
    for i, (ci, co, _) in enumerate(CONV):
        out.append((f"conv{i + 1}", (co, kpad(ci)), co))
    for i, (fi, fo) in enumerate(FC):
        out.append((f"fc{i + 1}", (fo, fi), fo))
    return out


class _Gemm(ctypes.Structure):
    _fields_ = [
        ("a", ctypes.c_void_p), ("lda_m", ctypes.c_longlong), ("lda_k", ctypes.c_longlong),
        ("b", ctypes.c_void_p), ("ldb_k", ctypes.c_longlong), ("ldb_n", ctypes.c_longlong),
        ("m", ctypes.c_int), ("n", ctypes.c_int), ("k", ctypes.c_int), ("alpha", ctypes.c_float),
        ("bias", ctypes.c_void_p), ("mask", ctypes.c_void_p), ("ld_mask", ctypes.c_longlong),
        ("c32", ctypes.c_void_p), ("ldc32", ctypes.c_longlong), ("c16", ctypes.c_void_p),
        ("ldc16", ctypes.c_longlong), ("flags", ctypes.c_uint), ("split_k", ctypes.c_int),
    ]


def _ptr(t):
    return None if t is None else t.data_ptr()


def _stream(dev):
    return torch.cuda.current_stream(dev).cuda_stream


def gemm(a, lda_m, lda_k, b, ldb_k, ldb_n, m, n, k, *, c32=None, ldc32=0, c16=None, ldc16=0, bias=None, mask=None,
         ld_mask=0, relu=False, accumulate=False, split_k=1, alpha=1.0):
    """cryomi_gemm_bf16 on uint16 (bf16) device tensors; see include/cryomi_cnn.h."""
    lib = _lib.load()
    flags = (_lib.GEMM_RELU if relu else 0) | (_lib.GEMM_ACCUMULATE if accumulate else 0)
    args = _Gemm(_ptr(a), lda_m, lda_k, _ptr(b), ldb_k, ldb_n, m, n, k, alpha, _ptr(bias), _ptr(mask), ld_mask,
                 _ptr(c32), ldc32, _ptr(c16), ldc16, flags, split_k)
    _lib.check(lib.cryomi_gemm_bf16(ctypes.byref(args), _stream(a.device)), "cryomi_gemm_bf16")


class _Conv(ctypes.Structure):
    _fields_ = [
        ("op", ctypes.c_int), ("n", ctypes.c_int), ("h", ctypes.c_int), ("w", ctypes.c_int), ("c", ctypes.c_int),
        ("co", ctypes.c_int), ("x", ctypes.c_void_p), ("weights", ctypes.c_void_p), ("dy", ctypes.c_void_p),
        ("bias", ctypes.c_void_p), ("mask", ctypes.c_void_p), ("c32", ctypes.c_void_p), ("c16", ctypes.c_void_p),
        ("flags", ctypes.c_uint), ("split_k", ctypes.c_int), ("bias_grad", ctypes.c_void_p),
        ("workspace", ctypes.c_void_p), ("workspace_bytes", ctypes.c_size_t),
    ]


def conv_workspace_floats(c, co, split_k, bias_grad=True):
    """Workspace (float32 elements) of a weight-gradient conv call."""
    return split_k * co * (9 * c + (1 if bias_grad else 0))


def conv(op, n, h, w, c, co, *, x=None, weights=None, dy=None, bias=None, mask=None, c32=None, c16=None, relu=False,
         split_k=1, bias_grad=None, workspace=None, accumulate=False):
    """cryomi_conv3x3_s2_bf16: a 3x3 / stride-2 layer as an implicit GEMM (see include/cryomi_cnn.h)."""
    lib = _lib.load()
    dev = next(t for t in (x, weights, dy) if t is not None).device
    flags = (_lib.GEMM_RELU if relu else 0) | (_lib.GEMM_ACCUMULATE if accumulate else 0)
    wbytes = 0 if workspace is None else workspace.numel() * workspace.element_size()
    args = _Conv(op, n, h, w, c, co, _ptr(x), _ptr(weights), _ptr(dy), _ptr(bias), _ptr(mask), _ptr(c32), _ptr(c16),
                 flags, split_k, _ptr(bias_grad), _ptr(workspace), wbytes)
    _lib.check(lib.cryomi_conv3x3_s2_bf16(ctypes.byref(args), _stream(dev)), "cryomi_conv3x3_s2_bf16")


def _split(tiles, k):
    """Split-k layers for a weight-gradient GEMM: ~256 workgroups (one per CU), >= 512 k each."""
    s = max(1, min(math.ceil(256 / max(tiles, 1)), k // 512))
    return s


def to_bf16(x):
    """float32 cuda tensor -> uint16 bf16 bit patterns (cryomi_to_bf16)."""
    out = torch.empty(x.shape, dtype=torch.int16, device=x.device)
    _lib.check(_lib.load().cryomi_to_bf16(x.data_ptr(), x.numel(), out.data_ptr(), _stream(x.device)),
               "cryomi_to_bf16")
    return out


def from_bf16(h):
    """int16 tensor of bf16 bit patterns -> float32 (a bit shift; for inspection and tests)."""
    return ((h.to(torch.int32) & 0xFFFF) << 16).view(torch.float32)


class PoseNet:
    """Parameters, optimiser state and the forward / backward passes."""

    def __init__(self, device="cuda:0", seed=0, params=None):
        self.device = torch.device(device)
        self.shapes = layer_shapes()
        self.offsets = {}
        o = 0
        for name, (co, k), nb in self.shapes:
            o = (o + 7) // 8 * 8  # 16-byte aligned bf16 weights
            self.offsets[name + ".w"] = (o, (co, k))
            o += co * k
            self.offsets[name + ".b"] = (o, (nb,))
            o += nb
        self.n_params = (o + 7) // 8 * 8
        host = np.zeros(self.n_params, dtype=np.float32)
        init = params if params is not None else self.init_params(seed)
        for key, (off, shape) in self.offsets.items():
            host[off:off + int(np.prod(shape))] = np.asarray(init[key], dtype=np.float32).ravel()
        self.p = torch.from_numpy(host).to(self.device)
        self.p16 = to_bf16(self.p)
        self.g = torch.zeros_like(self.p)
        self.m = torch.zeros_like(self.p)
        self.v = torch.zeros_like(self.p)
        self.t = 0
        self._bufs = {}

    @staticmethod
    def init_params(seed=0):
        rng = np.random.default_rng(seed)
        p = {}
        for name, (co, k), nb in layer_shapes():
            fan_in = 9 * CONV[int(name[-1]) - 1][0] if name.startswith("conv") else k
            w = np.zeros((co, k), dtype=np.float32)
            scale = math.sqrt(2.0 / fan_in) if name != "fc2" else 0.01
            w[:, :fan_in] = rng.standard_normal((co, fan_in)) * scale
            p[name + ".w"] = w
            p[name + ".b"] = np.zeros(nb, dtype=np.float32)
        return p

    def view(self, key, which="p"):
        off, shape = self.offsets[key]
        t = getattr(self, which)
        return t[off:off + int(np.prod(shape))].view(*shape)

    def state_dict(self):
        return {k: self.view(k).cpu().numpy().copy() for k in self.offsets}

    def load_state_dict(self, d):
        for k in self.offsets:
            self.view(k).copy_(torch.from_numpy(np.asarray(d[k], dtype=np.float32)))
        self.p16 = to_bf16(self.p)

    # -- buffers -------------------------------------------------------------
    def _buf(self, key, shape, dtype):
        t = self._bufs.get(key)
        if t is None or tuple(t.shape) != tuple(shape):
            t = torch.empty(shape, dtype=dtype, device=self.device)
            self._bufs[key] = t
        return t

    # -- forward ---------------------------------------------------------------
    def forward(self, x16, n):
        """x16: (n, 64, 64) bf16 (int16 tensor). Returns out (n, 96) float32."""
        lib = _lib.load()
        st = _stream(self.device)
        a = x16
        for i, (ci, co, s) in enumerate(CONV):
            kp = kpad(ci)
            npix = n * (s // 2) ** 2
            y = self._buf(f"a{i + 1}", (npix, co), torch.int16)
            w16 = self.view(f"conv{i + 1}.w", "p16")
            if ci % 8 == 0:  # patches gathered inside the GEMM
                conv(_lib.CONV_FORWARD, n, s, s, ci, co, x=a, weights=w16, bias=self.view(f"conv{i + 1}.b"), c16=y,
                     relu=True)
            else:  # one input channel: explicit patches (kp = 16 columns)
                col = self._buf(f"col{i + 1}", (npix, kp), torch.int16)
                _lib.check(lib.cryomi_im2col_s2(a.data_ptr(), n, s, s, ci, kp, col.data_ptr(), st),
                           "cryomi_im2col_s2")
                gemm(col, kp, 1, w16, 1, kp, npix, co, kp, c16=y, ldc16=co, bias=self.view(f"conv{i + 1}.b"),
                     relu=True)
            self._bufs[f"in{i + 1}"] = a
            a = y
        f = a  # (n * 16, 128) == (n, 2048) NHWC flatten
        h = self._buf("h", (n, 512), torch.int16)
        gemm(f, 2048, 1, self.view("fc1.w", "p16"), 1, 2048, n, 512, 2048, c16=h, ldc16=512,
             bias=self.view("fc1.b"), relu=True)
        out = self._buf("out", (n, N_OUT), torch.float32)
        gemm(h, 512, 1, self.view("fc2.w", "p16"), 1, 512, n, N_OUT, 512, c32=out, ldc32=N_OUT,
             bias=self.view("fc2.b"))
        self._x = x16
        self._n = n
        return out

    # -- backward ----------------------------------------------------------------
    def backward(self, dout16):
        """dout16: (n, 96) bf16 = dL/dout. Accumulates into self.g (zeroed here)."""
        lib = _lib.load()
        st = _stream(self.device)
        n = self._n
        if n % 8:
            raise ValueError("training batches must be a multiple of 8 images")
        _lib.check(lib.cryomi_fill_zero(self.g.data_ptr(), self.g.numel(), st), "cryomi_fill_zero")
        h = self._bufs["h"]
        f = self._bufs["a4"]
        # fc2 (k = the batch: short, no split)
        gemm(dout16, 1, N_OUT, h, 512, 1, N_OUT, 512, n, c32=self.view("fc2.w", "g"), ldc32=512)
        _lib.check(lib.cryomi_column_sums_bf16(dout16.data_ptr(), n, N_OUT, N_OUT, self.view("fc2.b", "g").data_ptr(),
                                               st), "cryomi_column_sums_bf16")
        dh = self._buf("dh", (n, 512), torch.int16)
        gemm(dout16, N_OUT, 1, self.view("fc2.w", "p16"), 512, 1, n, 512, N_OUT, c16=dh, ldc16=512, mask=h,
             ld_mask=512)
        # fc1
        gemm(dh, 1, 512, f, 2048, 1, 512, 2048, n, c32=self.view("fc1.w", "g"), ldc32=2048)
        _lib.check(lib.cryomi_column_sums_bf16(dh.data_ptr(), n, 512, 512, self.view("fc1.b", "g").data_ptr(), st),
                   "cryomi_column_sums_bf16")
        dy = self._buf("dy4", (n * 16, 128), torch.int16)
        gemm(dh, 512, 1, self.view("fc1.w", "p16"), 2048, 1, n, 2048, 512, c16=dy, ldc16=2048, mask=f, ld_mask=2048)
        for i in range(len(CONV), 0, -1):
            ci, co, s = CONV[i - 1]
            kp = kpad(ci)
            npix = n * (s // 2) ** 2
            tiles = ((co + 127) // 128) * ((kp + 1 + 127) // 128)
            if ci % 8 == 0:  # weight and bias gradients in one implicit GEMM
                sk = _split(tiles, npix)
                ws = self._buf(f"ws{i}", (conv_workspace_floats(ci, co, sk),), torch.float32)
                conv(_lib.CONV_WEIGHT_GRAD, n, s, s, ci, co, x=self._bufs[f"in{i}"], dy=dy,
                     c32=self.view(f"conv{i}.w", "g"), bias_grad=self.view(f"conv{i}.b", "g"), split_k=sk,
                     workspace=ws)
            else:
                col = self._bufs[f"col{i}"]
                gemm(dy, 1, co, col, kp, 1, co, kp, npix, c32=self.view(f"conv{i}.w", "g"), ldc32=kp,
                     accumulate=True, split_k=_split(tiles, npix))
                _lib.check(lib.cryomi_column_sums_bf16(dy.data_ptr(), npix, co, co,
                                                       self.view(f"conv{i}.b", "g").data_ptr(), st),
                           "cryomi_column_sums_bf16")
            if i > 1:
                a_in = self._bufs[f"a{i - 1}"]
                dprev = self._buf(f"dy{i - 1}", (n * s * s, ci), torch.int16)
                conv(_lib.CONV_DATA_GRAD, n, s, s, ci, co, weights=self.view(f"conv{i}.w", "p16"), dy=dy, mask=a_in,
                     c16=dprev)
                dy = dprev

    def adam_step(self, lr=3e-4, beta1=0.9, beta2=0.999, eps=1e-8):
        self.t += 1
        _lib.check(_lib.load().cryomi_adam(self.p.data_ptr(), self.m.data_ptr(), self.v.data_ptr(), self.g.data_ptr(),
                                           self.p16.data_ptr(), self.p.numel(), lr, beta1, beta2, eps,
                                           1 - beta1**self.t, 1 - beta2**self.t, _stream(self.device)), "cryomi_adam")

    def loss(self, out, rot, off, eps=0.05, off_weight=0.01):
        """cryomi_pose_loss; returns (dout bf16, terms (n, 3))."""
        n = out.shape[0]
        dout = self._buf("dout", (n, N_OUT), torch.int16)
        terms = self._buf("terms", (n, 3), torch.float32)
        _lib.check(_lib.load().cryomi_pose_loss(out.data_ptr(), rot.data_ptr(), off.data_ptr(), n, eps, off_weight,
                                                dout.data_ptr(), None, terms.data_ptr(), _stream(self.device)),
                   "cryomi_pose_loss")
        return dout, terms

