"""Learned artifact correction on top of the plain FDK (no counterpart in the
reference, which has no learned part).

A 3-level residual encoder-decoder of 3x3x3 convolutions with ReLU, 32 / 64 /
128 channels, 2x2x2 average pooling down and nearest upsampling up, additive
skips at levels 1 and 2 (added before the ReLU of the first decoder
convolution of the level).  The network predicts a correction c(x) of the
plain reconstruction x, and the output is x + c(x).  Inputs and targets are in
units of HU / 1000 (mu / mu_water - 1).

Every operation is a kernel of libarecon.so (include/arecon_cnn.h): BF16
activations and weights on the matrix cores with FP32 accumulation, FP32
master weights and Adam moments.  torch provides device memory only.

Fallback: ``correct()`` returns the plain reconstruction wherever the RMS of
the correction over a 64^3 block exceeds ``bound`` (HU / 1000), so the plain
FDK remains the answer where the network would move far from it.
"""
from __future__ import annotations

import ctypes as C
import math
from dataclasses import dataclass

import numpy as np
import torch

from . import _lib

MU_WATER = 0.02

# name, input channels, output channels, ReLU
LAYERS = [
    ("e1a", 8, 32, True), ("e1b", 32, 32, True),
    ("e2a", 32, 64, True), ("e2b", 64, 64, True),
    ("b1", 64, 128, True), ("b2", 128, 128, True),
    ("d2a", 128, 64, True), ("d2b", 64, 64, True),
    ("d1a", 64, 32, True), ("d1b", 32, 32, True),
    ("out", 32, 8, False),          # channel 0 is the correction
]
CIN0 = 8                            # the one input channel, padded to 8


def _kp(c: int) -> int:
    return (27 * c + 31) // 32 * 32


def _r16(c: int) -> int:
    return (c + 15) // 16 * 16


def _dims(B, D, H, W):
    return (C.c_int * 4)(B, D, H, W)


def _ptr(t: torch.Tensor | None):
    return C.c_void_p(t.data_ptr()) if t is not None else C.c_void_p(0)


def _call(name, *args):
    _lib.check(getattr(_lib.lib(), name)(*args), name)


class Layer:
    def __init__(self, name, cin, cout, relu, gen: torch.Generator, dev):
        self.name, self.cin, self.cout, self.relu = name, cin, cout, relu
        fan_in = 27 * (1 if name == "e1a" else cin)
        w = torch.randn((cout, 27, cin), generator=gen) * math.sqrt(2.0 / fan_in)
        if name == "e1a":
            w[:, :, 1:] = 0.0                  # padded input channels
        if name == "out":
            w.zero_()                          # start from the plain reconstruction
        self.w = w.to(dev).contiguous()
        self.b = torch.zeros(cout, device=dev)
        self.gw = torch.zeros_like(self.w)
        self.gb = torch.zeros_like(self.b)
        self.mw, self.vw = torch.zeros_like(self.w), torch.zeros_like(self.w)
        self.mb, self.vb = torch.zeros_like(self.b), torch.zeros_like(self.b)
        self.wf = torch.zeros((_r16(cout), _kp(cin)), dtype=torch.int16, device=dev)
        self.wd = torch.zeros((_r16(cin), _kp(cout)), dtype=torch.int16, device=dev)

    def pack(self, stream):
        _call("arecon_cnn_pack_weights", _ptr(self.wf), _ptr(self.wd), _ptr(self.w), self.cout,
              self.cin, stream)


@dataclass
class Shape3:
    B: int
    D: int
    H: int
    W: int

    def half(self):
        return Shape3(self.B, self.D // 2, self.H // 2, self.W // 2)

    @property
    def n(self):
        return self.B * self.D * self.H * self.W

    def c(self):
        return _dims(self.B, self.D, self.H, self.W)


class Net:
    def __init__(self, seed: int = 0, device: str = "cuda"):
        self.dev = torch.device(device)
        gen = torch.Generator().manual_seed(seed)
        self.layers = {name: Layer(name, ci, co, r, gen, self.dev) for name, ci, co, r in LAYERS}
        self.step_count = 0
        self._ws = None
        self._ws_shape = None

    # -- parameters ---------------------------------------------------------
    def stream(self):
        return C.c_void_p(torch.cuda.current_stream().cuda_stream)

    def pack(self):
        s = self.stream()
        for L in self.layers.values():
            L.pack(s)

    def state_dict(self):
        return {n: (L.w.cpu(), L.b.cpu()) for n, L in self.layers.items()}

    def load_state_dict(self, sd):
        for n, (w, b) in sd.items():
            self.layers[n].w.copy_(w)
            self.layers[n].b.copy_(b)
        self.pack()

    # -- workspace ----------------------------------------------------------
    def _workspace(self, s1: Shape3):
        if self._ws_shape == (s1.B, s1.D, s1.H, s1.W):
            return self._ws
        assert s1.D % 4 == 0 and s1.H % 4 == 0 and s1.W % 32 == 0, \
            "volume dims: D, H multiples of 4, W a multiple of 32"
        s2, s3 = s1.half(), s1.half().half()
        bf = dict(dtype=torch.int16, device=self.dev)
        ws = {"L1": s1, "L2": s2, "L3": s3}
        acts = {"xin": (s1, CIN0), "a1": (s1, 32), "s1": (s1, 32), "p1": (s2, 32),
                "a2": (s2, 64), "s2": (s2, 64), "p2": (s3, 64), "a3": (s3, 128), "s3": (s3, 128),
                "u2": (s2, 128), "c2": (s2, 64), "d2": (s2, 64), "u1": (s1, 64), "c1": (s1, 32),
                "d1": (s1, 32), "out": (s1, 8)}
        for k, (sh, c) in acts.items():
            ws[k] = torch.empty((sh.n, c), **bf)
        ws["shape_of"] = {k: v[0] for k, v in acts.items()}
        self._ws, self._ws_shape = ws, (s1.B, s1.D, s1.H, s1.W)
        return ws

    # -- forward ------------------------------------------------------------
    def _conv(self, name, y, x, sh: Shape3, skip=None):
        L = self.layers[name]
        _call("arecon_cnn_conv3", _ptr(y), _ptr(x), _ptr(L.wf), _ptr(L.b), _ptr(skip), sh.c(),
              L.cin, L.cout, int(L.relu), self.stream())

    def forward(self, x0: torch.Tensor) -> torch.Tensor:
        """x0: FP32 [B, D, H, W] (HU / 1000).  Returns the BF16 output tensor
        [nvox][8] (channel 0 = correction); activations stay for backward."""
        B, D, H, W = x0.shape
        ws = self._workspace(Shape3(B, D, H, W))
        s1, s2, s3 = ws["L1"], ws["L2"], ws["L3"]
        st = self.stream()
        self.x0 = x0
        _call("arecon_cnn_to_bf16", _ptr(ws["xin"]), _ptr(x0), s1.n, CIN0, st)
        self._conv("e1a", ws["a1"], ws["xin"], s1)
        self._conv("e1b", ws["s1"], ws["a1"], s1)
        _call("arecon_cnn_pool", _ptr(ws["p1"]), _ptr(ws["s1"]), s2.c(), 32, st)
        self._conv("e2a", ws["a2"], ws["p1"], s2)
        self._conv("e2b", ws["s2"], ws["a2"], s2)
        _call("arecon_cnn_pool", _ptr(ws["p2"]), _ptr(ws["s2"]), s3.c(), 64, st)
        self._conv("b1", ws["a3"], ws["p2"], s3)
        self._conv("b2", ws["s3"], ws["a3"], s3)
        _call("arecon_cnn_upsample", _ptr(ws["u2"]), _ptr(ws["s3"]), s2.c(), 128, st)
        self._conv("d2a", ws["c2"], ws["u2"], s2, skip=ws["s2"])
        self._conv("d2b", ws["d2"], ws["c2"], s2)
        _call("arecon_cnn_upsample", _ptr(ws["u1"]), _ptr(ws["d2"]), s1.c(), 64, st)
        self._conv("d1a", ws["c1"], ws["u1"], s1, skip=ws["s1"])
        self._conv("d1b", ws["d1"], ws["c1"], s1)
        self._conv("out", ws["out"], ws["d1"], s1)
        return ws["out"]

    # -- backward -----------------------------------------------------------
    def _layer_bwd(self, name, dy, y, x, sh: Shape3, need_dx=True):
        """dy: gradient w.r.t. the layer output y (post-activation); x: the
        layer input.  Accumulates gw, gb; returns (dx or None, dy masked)."""
        L = self.layers[name]
        st = self.stream()
        dyp = torch.empty((sh.n, L.cout), dtype=torch.int16, device=self.dev)
        dypT = torch.empty((L.cout, sh.n), dtype=torch.int16, device=self.dev)
        _call("arecon_cnn_relu_bwd", _ptr(dyp), _ptr(dypT), _ptr(L.gb), _ptr(dy), _ptr(y),
              sh.c(), L.cout, int(L.relu), st)
        xT3 = torch.empty((3, L.cin, sh.n), dtype=torch.int16, device=self.dev)
        _call("arecon_cnn_channel_major3", _ptr(xT3), _ptr(x), sh.c(), L.cin, st)
        _call("arecon_cnn_conv3_wgrad", _ptr(L.gw), _ptr(dypT), _ptr(xT3), sh.c(), L.cin, L.cout,
              st)
        dx = None
        if need_dx:
            dx = torch.empty((sh.n, L.cin), dtype=torch.int16, device=self.dev)
            _call("arecon_cnn_conv3", _ptr(dx), _ptr(dyp), _ptr(L.wd), _ptr(None), _ptr(None),
                  sh.c(), L.cout, L.cin, 0, st)
        return dx, dyp

    def backward(self, dout: torch.Tensor) -> None:
        ws = self._ws
        s1, s2, s3 = ws["L1"], ws["L2"], ws["L3"]
        st = self.stream()
        d_d1, _ = self._layer_bwd("out", dout, ws["out"], ws["d1"], s1)
        d_c1, _ = self._layer_bwd("d1b", d_d1, ws["d1"], ws["c1"], s1)
        d_u1, d_s1 = self._layer_bwd("d1a", d_c1, ws["c1"], ws["u1"], s1)   # skip grad = dyp
        d_s1 = d_s1.clone()
        d_d2 = torch.empty((s2.n, 64), dtype=torch.int16, device=self.dev)
        _call("arecon_cnn_upsample_bwd", _ptr(d_d2), _ptr(d_u1), s2.c(), 64, st)
        d_c2, _ = self._layer_bwd("d2b", d_d2, ws["d2"], ws["c2"], s2)
        d_u2, d_s2 = self._layer_bwd("d2a", d_c2, ws["c2"], ws["u2"], s2)
        d_s2 = d_s2.clone()
        d_s3 = torch.empty((s3.n, 128), dtype=torch.int16, device=self.dev)
        _call("arecon_cnn_upsample_bwd", _ptr(d_s3), _ptr(d_u2), s3.c(), 128, st)
        d_a3, _ = self._layer_bwd("b2", d_s3, ws["s3"], ws["a3"], s3)
        d_p2, _ = self._layer_bwd("b1", d_a3, ws["a3"], ws["p2"], s3)
        _call("arecon_cnn_pool_bwd", _ptr(d_s2), _ptr(d_p2), s2.c(), 64, 1, st)
        d_a2, _ = self._layer_bwd("e2b", d_s2, ws["s2"], ws["a2"], s2)
        d_p1, _ = self._layer_bwd("e2a", d_a2, ws["a2"], ws["p1"], s2)
        _call("arecon_cnn_pool_bwd", _ptr(d_s1), _ptr(d_p1), s1.c(), 32, 1, st)
        d_a1, _ = self._layer_bwd("e1b", d_s1, ws["s1"], ws["a1"], s1)
        self._layer_bwd("e1a", d_a1, ws["a1"], ws["xin"], s1, need_dx=False)

    # -- training -----------------------------------------------------------
    def zero_grad(self):
        for L in self.layers.values():
            L.gw.zero_()
            L.gb.zero_()

    def loss_and_grad(self, x0: torch.Tensor, target: torch.Tensor) -> torch.Tensor:
        """Forward + backward for FP32 [B, D, H, W] input and target (HU/1000);
        returns the mean squared error (a device scalar)."""
        out = self.forward(x0)
        dout = torch.empty_like(out)
        loss = torch.zeros(1, device=self.dev)
        _call("arecon_cnn_mse_residual", _ptr(dout), _ptr(loss), _ptr(out), 8, _ptr(x0),
              _ptr(target), x0.numel(), self.stream())
        self.backward(dout)
        return loss

    def adam_step(self, lr=1e-3, b1=0.9, b2=0.999, eps=1e-8):
        self.step_count += 1
        st = self.stream()
        for L in self.layers.values():
            _call("arecon_cnn_adam", _ptr(L.w), _ptr(L.mw), _ptr(L.vw), _ptr(L.gw), L.w.numel(),
                  lr, b1, b2, eps, self.step_count, st)
            _call("arecon_cnn_adam", _ptr(L.b), _ptr(L.mb), _ptr(L.vb), _ptr(L.gb), L.b.numel(),
                  lr, b1, b2, eps, self.step_count, st)
            L.pack(st)

    def train_step(self, x0, target, lr=1e-3) -> torch.Tensor:
        self.zero_grad()
        loss = self.loss_and_grad(x0, target)
        self.adam_step(lr)
        return loss

    # -- inference ----------------------------------------------------------
    def apply(self, x0: torch.Tensor) -> torch.Tensor:
        """x0 + c(x0), FP32, same shape as x0."""
        out = self.forward(x0)
        y = torch.empty_like(x0)
        _call("arecon_cnn_residual_out", _ptr(y), _ptr(out), 8, _ptr(x0), x0.numel(),
              self.stream())
        return y

    def apply_tiled(self, x0: torch.Tensor, tile: int = 64, stride: int = 32,
                    batch: int = 16) -> torch.Tensor:
        """x0 + c(x0) evaluated on overlapping tile^3 crops, as in training: each
        tile keeps its central stride^3 region (its outer part at the volume's
        faces), so every voxel is predicted with the context a training patch
        gives its centre.  x0 is [1][D][H][W] with D, H, W >= tile."""
        assert x0.shape[0] == 1 and min(x0.shape[1:]) >= tile
        m = (tile - stride) // 2

        def starts(n):
            k = [i * stride for i in range(max(1, -(-(n - tile) // stride) + 1))]
            return sorted({min(s, n - tile) for s in k})

        def keep(s, n):
            return (0 if s == 0 else s + m), (n if s + tile >= n else s + tile - m)

        D, H, W = x0.shape[1:]
        y = torch.empty_like(x0)
        boxes = [(a, b, c) for a in starts(D) for b in starts(H) for c in starts(W)]
        for i in range(0, len(boxes), batch):
            grp = boxes[i:i + batch]
            xt = torch.stack([x0[0, a:a + tile, b:b + tile, c:c + tile] for a, b, c in grp])
            yt = self.apply(xt.contiguous())
            for k, (a, b, c) in enumerate(grp):
                (a0, a1), (b0, b1), (c0, c1) = keep(a, D), keep(b, H), keep(c, W)
                y[0, a0:a1, b0:b1, c0:c1] = yt[k, a0 - a:a1 - a, b0 - b:b1 - b, c0 - c:c1 - c]
        return y

    def correct(self, x0: torch.Tensor, bound: float = 0.3, block: int = 64,
                tiled: bool = False):
        """Corrected volume with the plain-FDK fallback: blocks of block^3 voxels
        (the last ones partial) whose correction RMS exceeds ``bound`` keep x0.
        ``tiled`` evaluates the network on overlapping 64^3 tiles
        (``apply_tiled``).  Returns (volume, fraction of blocks that fell back)."""
        y = self.apply_tiled(x0) if tiled else self.apply(x0)
        c = y - x0
        B, D, H, W = x0.shape
        nb = [-(-n // block) for n in (D, H, W)]
        pad = (0, nb[2] * block - W, 0, nb[1] * block - H, 0, nb[0] * block - D)
        sq = torch.nn.functional.pad(c * c, pad)
        ones = torch.nn.functional.pad(torch.ones_like(c), pad)
        shape = (B, nb[0], block, nb[1], block, nb[2], block)
        rms = (sq.reshape(shape).sum(dim=(2, 4, 6)) / ones.reshape(shape).sum(dim=(2, 4, 6))).sqrt()
        bad = rms > bound
        if bad.any():
            mask = (bad.repeat_interleave(block, 1).repeat_interleave(block, 2)
                    .repeat_interleave(block, 3))[:, :D, :H, :W]
            y[mask] = x0[mask]
        return y, float(bad.float().mean())
