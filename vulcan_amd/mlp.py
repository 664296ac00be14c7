"""Fully connected network with SiLU hidden layers, trained on the device.

Every matrix product and elementwise step runs in the HIP library
(BF16 MFMA with FP32 accumulation for the GEMMs, FP32 elementwise, Adam on
FP32 master weights with BF16 copies for the GEMMs); torch provides only the
buffers.  Shapes follow the row-major convention X[batch][features],
W[in][out].
"""
from __future__ import annotations

import ctypes as C
from typing import List, Optional, Sequence

import numpy as np
import torch

from . import _lib
from .device import dptr, stream_handle

F32 = torch.float32
BF16_T = torch.int16      # raw 16-bit storage of BF16 values


def to_bf16(x: torch.Tensor) -> torch.Tensor:
    """Device FP32 -> BF16 (round to nearest even) through the library."""
    x = x.contiguous()
    out = torch.empty(x.shape, dtype=BF16_T, device=x.device)
    _lib.check(_lib.load().vk_mlp_to_bf16(x.numel(), dptr(x), dptr(out), stream_handle()), 'vk_mlp_to_bf16')
    return out


def bf16_to_f32(t: torch.Tensor) -> torch.Tensor:
    return (t.to(torch.int32) << 16).view(torch.float32) if t.dtype == BF16_T else t


class MLP:
    def __init__(self, sizes: Sequence[int], seed: int = 0, device: str = 'cuda'):
        self.sizes = list(sizes)
        self.lib = _lib.load()
        g = np.random.default_rng(seed)
        dev = torch.device(device)
        self.W, self.Wb, self.b = [], [], []
        for fi, fo in zip(self.sizes[:-1], self.sizes[1:]):
            w = g.standard_normal((fi, fo)).astype(np.float32) * np.float32(np.sqrt(2.0 / fi))
            W = torch.as_tensor(w, device=dev)
            self.W.append(W)
            self.Wb.append(to_bf16(W))
            self.b.append(torch.zeros(fo, dtype=F32, device=dev))
        self.params = self.W + self.b
        self.m = [torch.zeros_like(p) for p in self.params]
        self.v = [torch.zeros_like(p) for p in self.params]
        self.grads = [torch.zeros_like(p) for p in self.params]
        self.step_count = 0
        self._cache = None

    @property
    def nlayers(self) -> int:
        return len(self.W)

    def _gemm(self, M, N, K, A, sam, sak, B, sbk, sbn, Cout, ldc, bias=None, acc=0, Yact=None, ldy=0):
        _lib.check(self.lib.vk_mlp_gemm(M, N, K, dptr(A), sam, sak, dptr(B), sbk, sbn, dptr(Cout), ldc,
                                        dptr(bias), acc, dptr(Yact), ldy, stream_handle()), 'vk_mlp_gemm')

    def forward(self, X: torch.Tensor, keep: bool = False) -> torch.Tensor:
        """X: [B][in] FP32 on the device -> output [B][out] FP32."""
        Bn = X.shape[0]
        A = to_bf16(X)
        acts, Zs = [A], []
        for l in range(self.nlayers):
            fi, fo = self.sizes[l], self.sizes[l + 1]
            Z = torch.empty((Bn, fo), dtype=F32, device=X.device)
            last = l == self.nlayers - 1
            Y = None if last else torch.empty((Bn, fo), dtype=BF16_T, device=X.device)
            self._gemm(Bn, fo, fi, A, fi, 1, self.Wb[l], fo, 1, Z, fo, bias=self.b[l], Yact=Y, ldy=fo)
            Zs.append(Z)
            if not last:
                acts.append(Y)
                A = Y
        if keep:
            self._cache = (acts, Zs)
        return Zs[-1]

    def loss_and_backward(self, X, T, Wmask=None):
        """Masked mean-squared error and the gradients of every parameter."""
        Y = self.forward(X, keep=True)
        acts, Zs = self._cache
        Bn = X.shape[0]
        n = Y.numel()
        count = float(Wmask.sum().item()) if Wmask is not None else float(n)
        dY = torch.empty_like(Y)
        dYb = torch.empty(Y.shape, dtype=BF16_T, device=Y.device)
        parts = torch.empty((n + 255) // 256, dtype=F32, device=Y.device)
        _lib.check(self.lib.vk_mlp_mse(n, dptr(Y), dptr(T.contiguous()), dptr(Wmask), C.c_float(1.0 / count),
                                       dptr(dY), dptr(dYb), dptr(parts), stream_handle()), 'vk_mlp_mse')
        dZ, dZb = dY, dYb
        L = self.nlayers
        for l in range(L - 1, -1, -1):
            fi, fo = self.sizes[l], self.sizes[l + 1]
            gW, gb = self.grads[l], self.grads[L + l]
            # dW = A_{l-1}^T dZ
            self._gemm(fi, fo, Bn, acts[l], 1, fi, dZb, fo, 1, gW, fo)
            _lib.check(self.lib.vk_mlp_colsum(Bn, fo, dptr(dZ), dptr(gb), stream_handle()), 'vk_mlp_colsum')
            if l == 0:
                break
            # dA_{l-1} = dZ W^T, then dZ_{l-1} = dA * silu'(Z_{l-1})
            dA = torch.empty((Bn, fi), dtype=F32, device=X.device)
            self._gemm(Bn, fi, fo, dZb, fo, 1, self.Wb[l], 1, fo, dA, fi)
            dZn = torch.empty_like(dA)
            dZnb = torch.empty(dA.shape, dtype=BF16_T, device=X.device)
            _lib.check(self.lib.vk_mlp_silu_bwd(dA.numel(), dptr(Zs[l - 1]), dptr(dA), dptr(dZn), dptr(dZnb),
                                                stream_handle()), 'vk_mlp_silu_bwd')
            dZ, dZb = dZn, dZnb
        return float(parts.sum().item()), Y

    def adam_step(self, lr=1e-3, b1=0.9, b2=0.999, eps=1e-8, wd=0.0):
        self.step_count += 1
        L = self.nlayers
        for q, (p, g, m, v) in enumerate(zip(self.params, self.grads, self.m, self.v)):
            wb = self.Wb[q] if q < L else None
            _lib.check(self.lib.vk_mlp_adam(p.numel(), dptr(p), dptr(g), dptr(m), dptr(v), C.c_float(lr),
                                            C.c_float(b1), C.c_float(b2), C.c_float(eps), self.step_count,
                                            C.c_float(wd if q < L else 0.0), dptr(wb), stream_handle()),
                       'vk_mlp_adam')

    def state_dict(self):
        return {'sizes': self.sizes, 'W': [w.cpu().numpy() for w in self.W], 'b': [b.cpu().numpy() for b in self.b]}

    def load_state_dict(self, d):
        for l, (w, b) in enumerate(zip(d['W'], d['b'])):
            self.W[l].copy_(torch.as_tensor(w))
            self.b[l].copy_(torch.as_tensor(b))
            self.Wb[l] = to_bf16(self.W[l])
