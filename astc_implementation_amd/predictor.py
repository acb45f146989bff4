"""Learned candidate predictor for the encoder's search.

An MLP over one block's texels (R, G, B / 255 of every texel: 48 inputs at
4x4, 108 at 6x6) with two GELU hidden layers of 256 and five softmax heads:

* ``pc``    4 logits   partition count 1..4 of the search's chosen encoding
* ``seed2`` 1024       best 2-partition seed the search found for the block
* ``seed3`` 1024       best 3-partition seed
* ``seed4`` 1024       best 4-partition seed
* ``mode``  n modes    trial block mode of the chosen encoding

Labels come from the encoder's own exhaustive (plain) search
(``astcg_get_labels``).  Every matrix product and elementwise step runs in
the library's HIP kernels (``include/astcg_predictor.h``): BF16 operands on
the gfx950 matrix cores with FP32 accumulation, FP32 master weights updated
by Adam.  PyTorch only provides device memory and the stream.
"""
import ctypes

import numpy as np

from . import library, _check

HIDDEN = 256
HEADS = ("pc", "seed2", "seed3", "seed4", "mode")


def _sig(lib):
    vp, i32, i64, f32 = ctypes.c_void_p, ctypes.c_int, ctypes.c_longlong, ctypes.c_float
    sig = {
        "astcg_mlp_gemm": [i32, i32, i32, vp, i32, i32, vp, i32, i32, vp, i32, vp, i32, vp],
        "astcg_mlp_gelu": [vp, vp, i64, vp],
        "astcg_mlp_gelu_backward": [vp, vp, vp, vp, i64, vp],
        "astcg_mlp_to_bf16": [vp, vp, i64, vp],
        "astcg_mlp_colsum": [vp, i32, i32, vp, i32, vp],
        "astcg_mlp_softmax_xent": [vp, i32, i32, i32, vp, vp, vp, f32, vp, vp, vp, vp],
        "astcg_mlp_rank_xent": [vp, i32, i32, i32, i32, vp, vp, f32, vp, vp, vp, vp],
        "astcg_mlp_adam": [vp, vp, vp, vp, vp, i64, f32, f32, f32, f32, i32, f32, vp],
        "astcg_mlp_features": [vp, i32, i32, i32, i32, i32, i32, vp, i32, vp],
        "astcg_mlp_topk": [vp, i32, i32, i32, vp, vp, i32, vp, vp, vp],
        "astcg_mlp_infer": [i32, vp, i32, i32, vp, vp, vp, vp, vp, vp, i32, i32, i32, i32, i32, vp, vp, vp, vp],
        "astcg_mlp_seed_topk": [vp, vp, i32, vp, vp, vp, i32, vp, i32, i32, vp],
    }
    for name, args in sig.items():
        fn = getattr(lib, name)
        fn.restype = ctypes.c_int
        fn.argtypes = args
    return lib


def _p(t):
    return ctypes.c_void_p(t.data_ptr()) if t is not None else None


def trial_mode_count(bx, by):
    n = ctypes.c_uint32()
    _check(library().astcg_trial_modes(bx, by, ctypes.byref(n), None, 0))
    return n.value


class Predictor:
    """The MLP, its Adam state and its training / inference passes."""

    def __init__(self, block=(4, 4), hidden=HIDDEN, seed=0, device="cuda"):
        import torch
        self.torch = torch
        self.lib = _sig(library())
        self.bx, self.by = block
        self.nin = 3 * self.bx * self.by
        self.hidden = hidden
        self.nmodes = trial_mode_count(self.bx, self.by)
        self.head_len = [4, 1024, 1024, 1024, self.nmodes]
        self.head_off = list(np.cumsum([0] + self.head_len[:-1]))
        self.nout = int(sum(self.head_len))
        self.device = torch.device(device)
        self.d_head_off = torch.tensor(self.head_off, dtype=torch.int32, device=self.device)
        self.d_head_len = torch.tensor(self.head_len, dtype=torch.int32, device=self.device)
        g = torch.Generator().manual_seed(seed)
        shapes = [(self.nin, hidden), (hidden,), (hidden, hidden), (hidden,), (hidden, self.nout), (self.nout,)]
        self.params = []
        for s in shapes:
            if len(s) == 2:
                bound = (6.0 / s[0]) ** 0.5 * (0.5 if s[1] == self.nout else 1.0)
                w = (torch.rand(s, generator=g) * 2.0 - 1.0) * bound
            else:
                w = torch.zeros(s)
            self.params.append(w.to(self.device).contiguous())
        self.params16 = [w.to(torch.bfloat16) if w.dim() == 2 else None for w in self.params]
        self.m = [torch.zeros_like(w) for w in self.params]
        self.v = [torch.zeros_like(w) for w in self.params]
        self.step = 0

    # ---- kernels -------------------------------------------------------------
    def _stream(self):
        return ctypes.c_void_p(self.torch.cuda.current_stream(self.device).cuda_stream)

    def gemm(self, m, n, k, a, lda, ta, b, ldb, tb, c, ldc, bias=None, accumulate=False):
        _check(self.lib.astcg_mlp_gemm(m, n, k, _p(a), lda, int(ta), _p(b), ldb, int(tb), _p(c), ldc, _p(bias),
                                       int(accumulate), self._stream()))

    def refresh_bf16(self):
        for w, w16 in zip(self.params, self.params16):
            if w16 is not None:
                _check(self.lib.astcg_mlp_to_bf16(_p(w), _p(w16), w.numel(), self._stream()))

    # ---- passes --------------------------------------------------------------
    def forward(self, x16):
        """x16: (B, nin) bf16 -> logits (B, nout) fp32; keeps activations."""
        torch = self.torch
        B, H = x16.shape[0], self.hidden
        W1, b1, W2, b2, W3, b3 = self.params
        W1h, _, W2h, _, W3h, _ = self.params16
        f32 = dict(dtype=torch.float32, device=self.device)
        bf = dict(dtype=torch.bfloat16, device=self.device)
        z1 = torch.empty((B, H), **f32)
        self.gemm(B, H, self.nin, x16, self.nin, 0, W1h, H, 0, z1, H, b1)
        h1 = torch.empty((B, H), **bf)
        _check(self.lib.astcg_mlp_gelu(_p(z1), _p(h1), z1.numel(), self._stream()))
        z2 = torch.empty((B, H), **f32)
        self.gemm(B, H, H, h1, H, 0, W2h, H, 0, z2, H, b2)
        h2 = torch.empty((B, H), **bf)
        _check(self.lib.astcg_mlp_gelu(_p(z2), _p(h2), z2.numel(), self._stream()))
        logits = torch.empty((B, self.nout), **f32)
        self.gemm(B, self.nout, H, h2, H, 0, W3h, self.nout, 0, logits, self.nout, b3)
        self.act = (x16, z1, h1, z2, h2)
        return logits

    def backward(self, logits, labels, rank=None, rank_weight=1.0):
        """Softmax cross-entropy of every head (labels (B, 5) int32, -1 =
        none), plus, with rank = (cols (B, 12) int32, target (B, 3) int32),
        a softmax over the 4 compat positions' seed logits of each partition
        count (cols: absolute logit columns, target: the winning position);
        returns the per-row losses and the gradients (fp32, like params)."""
        torch = self.torch
        x16, z1, h1, z2, h2 = self.act
        B, H, NO = x16.shape[0], self.hidden, self.nout
        W1h, _, W2h, _, W3h, _ = self.params16
        f32 = dict(dtype=torch.float32, device=self.device)
        bf = dict(dtype=torch.bfloat16, device=self.device)
        dlog = torch.empty((B, NO), **bf)
        dlog32 = torch.empty((B, NO), **f32)
        loss = torch.empty((B,), **f32)
        _check(self.lib.astcg_mlp_softmax_xent(_p(logits), B, NO, len(HEADS), _p(self.d_head_off),
                                               _p(self.d_head_len), _p(labels), 1.0 / B, _p(dlog), _p(dlog32),
                                               _p(loss), self._stream()))
        if rank is not None:
            cols, target = rank
            _check(self.lib.astcg_mlp_rank_xent(_p(logits), B, NO, 3, 4, _p(cols), _p(target), rank_weight / B,
                                                _p(dlog), _p(dlog32), _p(loss), self._stream()))
        grads = [torch.empty_like(w) for w in self.params]
        gW1, gb1, gW2, gb2, gW3, gb3 = grads
        self.gemm(H, NO, B, h2, H, 1, dlog, NO, 0, gW3, NO)
        _check(self.lib.astcg_mlp_colsum(_p(dlog32), B, NO, _p(gb3), 0, self._stream()))
        dh2 = torch.empty((B, H), **f32)
        self.gemm(B, H, NO, dlog, NO, 0, W3h, NO, 1, dh2, H)
        dz2 = torch.empty((B, H), **bf)
        dz2_32 = torch.empty((B, H), **f32)
        _check(self.lib.astcg_mlp_gelu_backward(_p(dh2), _p(z2), _p(dz2), _p(dz2_32), dh2.numel(), self._stream()))
        self.gemm(H, H, B, h1, H, 1, dz2, H, 0, gW2, H)
        _check(self.lib.astcg_mlp_colsum(_p(dz2_32), B, H, _p(gb2), 0, self._stream()))
        dh1 = torch.empty((B, H), **f32)
        self.gemm(B, H, H, dz2, H, 0, W2h, H, 1, dh1, H)
        dz1 = torch.empty((B, H), **bf)
        dz1_32 = torch.empty((B, H), **f32)
        _check(self.lib.astcg_mlp_gelu_backward(_p(dh1), _p(z1), _p(dz1), _p(dz1_32), dh1.numel(), self._stream()))
        self.gemm(self.nin, H, B, x16, self.nin, 1, dz1, H, 0, gW1, H)
        _check(self.lib.astcg_mlp_colsum(_p(dz1_32), B, H, _p(gb1), 0, self._stream()))
        return loss, grads

    def adam(self, grads, lr=1e-3, betas=(0.9, 0.999), eps=1e-8, weight_decay=0.0):
        self.step += 1
        for w, g, m, v, w16 in zip(self.params, grads, self.m, self.v, self.params16):
            _check(self.lib.astcg_mlp_adam(_p(w), _p(g), _p(m), _p(v), _p(w16), w.numel(), lr, betas[0], betas[1],
                                           eps, self.step, weight_decay, self._stream()))

    def train_step(self, x16, labels, lr=1e-3, rank=None):
        logits = self.forward(x16)
        loss, grads = self.backward(logits, labels, rank)
        self.adam(grads, lr=lr)
        return loss

    def topk(self, logits, k):
        """Top-k columns (and probabilities) of every head: (B, 5, k) each."""
        torch = self.torch
        B = logits.shape[0]
        idx = torch.empty((B, len(HEADS), k), dtype=torch.int32, device=self.device)
        prob = torch.empty((B, len(HEADS), k), dtype=torch.float32, device=self.device)
        _check(self.lib.astcg_mlp_topk(_p(logits), B, self.nout, len(HEADS), _p(self.d_head_off),
                                       _p(self.d_head_len), k, _p(idx), _p(prob), self._stream()))
        return idx, prob

    def features(self, d_rgba, width, height, first_block, num_blocks):
        """bf16 (num_blocks, nin) inputs of image blocks (device RGBA8 tensor)."""
        x = self.torch.empty((num_blocks, self.nin), dtype=self.torch.bfloat16, device=self.device)
        _check(self.lib.astcg_mlp_features(_p(d_rgba), width, height, self.bx, self.by, first_block, num_blocks,
                                           _p(x), self.nin, self._stream()))
        return x

    # ---- persistence ---------------------------------------------------------
    def state(self):
        return {"block": (self.bx, self.by), "hidden": self.hidden,
                "params": [w.detach().cpu().numpy() for w in self.params]}

    def load(self, state):
        for w, a in zip(self.params, state["params"]):
            w.copy_(self.torch.from_numpy(np.asarray(a, dtype=np.float32)).to(self.device))
        self.refresh_bf16()
