"""A learned field predictor for the guided step: a small convolutional
network, trained on slices the plain step generates, whose output replaces
the extrapolated field guess of ``step(..., guess=...)`` (the predictor
pass of lcode.py:841-881).  The reference has no learned predictor.

Every operation runs in this package's own HIP kernels
(``lcode_pred_*`` in include/lcode3d_hip.h): 3 x 3 convolutions forward and
backward, MSE gradient, Adam, and the FP64 <-> FP32 plane conversions.
torch only holds the device buffers.

Network: 14 input planes -> conv3x3 -> ReLU -> conv3x3 -> ReLU -> conv3x3
-> 6 planes.  Inputs: the six fields of slices n-1 and n-2 (each divided by
a per-field scale), the beam of slice n and the plasma density of slice
n-1.  Output: the change of slice n's fields from their linear
extrapolation 2 F(n-1) - F(n-2), in units of per-field target scales; the
guess is 2 F(n-1) - F(n-2) + scale * output.
"""

import ctypes
import math

import numpy as np
import torch

from . import _lib
from .lcode import _stream

FIELDS = ("Ex", "Ey", "Ez", "Bx", "By", "Bz")
N_IN, N_OUT = 14, 6


def _ptr(t):
    return ctypes.c_void_p(t.data_ptr()) if t is not None else None


class ConvPredictor:
    """The network and its Adam state on the current HIP device."""

    def __init__(self, grid_steps, hidden=(16, 16), seed=0, lr=1e-3):
        self.N = int(grid_steps)
        self.dev = torch.device("cuda", torch.cuda.current_device())
        chans = (N_IN,) + tuple(hidden) + (N_OUT,)
        self.layers = []            # (cin, cout, w offset, b offset, relu)
        off = 0
        for k in range(len(chans) - 1):
            cin, cout = chans[k], chans[k + 1]
            if not (1 <= cin <= 16 and 1 <= cout <= 16):
                raise ValueError("layer widths must be 1..16")
            relu = k < len(chans) - 2
            self.layers.append((cin, cout, off, off + cout * cin * 9, relu))
            off += cout * cin * 9 + cout
        self.nparams = off
        rng = np.random.default_rng(seed)
        p = np.zeros(off, dtype=np.float32)
        for li, (cin, cout, wo, bo, relu) in enumerate(self.layers):
            std = math.sqrt(2.0 / (cin * 9))           # He initialisation
            if li == len(self.layers) - 1:
                std *= 0.1     # start close to the linear extrapolation
            p[wo:bo] = rng.standard_normal(cout * cin * 9) * std
        self.params = torch.as_tensor(p, device=self.dev)
        self.grad = torch.zeros_like(self.params)
        self.m = torch.zeros_like(self.params)
        self.v = torch.zeros_like(self.params)
        self.t = 0
        self.lr = lr
        self.scale = np.ones(N_IN, dtype=np.float32)
        self.tscale = np.ones(N_OUT, dtype=np.float32)
        n2 = self.N * self.N
        self._acts = [torch.empty((c, self.N, self.N), dtype=torch.float32,
                                  device=self.dev) for c in chans]
        self._gacts = [torch.empty_like(a) for a in self._acts]
        self._scratch = torch.empty(max(
            _lib.load().lcode_pred_backward_weight_scratch(ci, co, self.N, self.N)
            for ci, co, *_ in self.layers), dtype=torch.float32, device=self.dev)
        self._loss_part = torch.empty((n2 * N_OUT + 255) // 256, dtype=torch.float32,
                                      device=self.dev)
        self._loss = torch.zeros(1, dtype=torch.float32, device=self.dev)
        self._guess = [torch.empty((self.N, self.N), dtype=torch.float64,
                                   device=self.dev) for _ in FIELDS]

    # -- parameters as the oracle's layer list (tests) -----------------------
    def layer_arrays(self):
        p = self.params.cpu().numpy().astype(np.float64)
        out = []
        for cin, cout, wo, bo, relu in self.layers:
            out.append((p[wo:bo].reshape(cout, cin, 3, 3), p[bo:bo + cout], relu))
        return out

    # -- FP64 fields <-> planes ---------------------------------------------
    def _io(self, s1, s2, beam, planes=None, nxt=None, pred=None, guess=None):
        io = _lib.LcodePredIo()
        for f, name in enumerate(FIELDS):
            io.f1[f] = getattr(s1, name).data_ptr()
            io.f2[f] = getattr(s2, name).data_ptr()
            if nxt is not None:
                io.next[f] = getattr(nxt, name).data_ptr()
            if guess is not None:
                io.guess[f] = guess[f].data_ptr()
        if isinstance(beam, torch.Tensor):
            io.beam = beam.data_ptr()
        else:
            io.beam_scalar = float(beam)
        io.ro = s1.ro.data_ptr()
        for c in range(N_IN):
            io.scale[c] = float(self.scale[c])
        for c in range(N_OUT):
            io.tscale[c] = float(self.tscale[c])
        io.planes = planes.data_ptr() if planes is not None else None
        io.pred = pred.data_ptr() if pred is not None else None
        io.n = self.N * self.N
        return io

    def inputs(self, s1, s2, beam, out=None):
        """Input planes [14, N, N] from states n-1, n-2 and beam n."""
        out = out if out is not None else torch.empty((N_IN, self.N, self.N),
                                                      dtype=torch.float32, device=self.dev)
        _lib.check(_lib.load().lcode_pred_inputs(
            ctypes.byref(self._io(s1, s2, beam, planes=out)), _stream()), "lcode_pred_inputs")
        return out

    def targets(self, s1, s2, nxt, out=None):
        """Target planes [6, N, N]: (F(n) - 2 F(n-1) + F(n-2)) / tscale."""
        out = out if out is not None else torch.empty((N_OUT, self.N, self.N),
                                                      dtype=torch.float32, device=self.dev)
        _lib.check(_lib.load().lcode_pred_targets(
            ctypes.byref(self._io(s1, s2, 0.0, planes=out, nxt=nxt)), _stream()),
            "lcode_pred_targets")
        return out

    def fit_scales(self, states):
        """Per-field RMS of the fields and of their slice-to-slice change
        from linear extrapolation, over a sequence of states."""
        f = np.array([[torch.sqrt((getattr(s, n) ** 2).mean()).item() for n in FIELDS]
                      for s in states])
        d = np.array([[torch.sqrt(((getattr(states[k], n) - 2 * getattr(states[k - 1], n)
                                     + getattr(states[k - 2], n)) ** 2).mean()).item()
                       for n in FIELDS] for k in range(2, len(states))])
        fs = np.maximum(f.max(axis=0), 1e-30)
        ds = np.maximum(d.max(axis=0), 1e-30)
        ro = max(max(torch.abs(s.ro).max().item() for s in states), 1e-30)
        self.scale[:6] = fs
        self.scale[6:12] = fs
        self.scale[12] = 0.1
        self.scale[13] = ro
        self.tscale[:] = ds

    # -- network -------------------------------------------------------------
    def forward(self, x):
        lib = _lib.load()
        self._acts[0].copy_(x)
        for li, (cin, cout, wo, bo, relu) in enumerate(self.layers):
            _lib.check(lib.lcode_pred_conv3x3_forward(
                _ptr(self._acts[li]), cin, _ptr(self.params[wo:]), _ptr(self.params[bo:]),
                _ptr(self._acts[li + 1]), cout, self.N, self.N, 1 if relu else 0,
                _stream()), "lcode_pred_conv3x3_forward")
        return self._acts[-1]

    def backward(self, gy):
        """Parameter gradients into self.grad (after forward)."""
        lib = _lib.load()
        self._gacts[-1].copy_(gy)
        for li in range(len(self.layers) - 1, -1, -1):
            cin, cout, wo, bo, relu = self.layers[li]
            _lib.check(lib.lcode_pred_conv3x3_backward_weight(
                _ptr(self._acts[li]), cin, _ptr(self._gacts[li + 1]), cout,
                _ptr(self._scratch), _ptr(self.grad[wo:]), self.N, self.N, _stream()),
                "lcode_pred_conv3x3_backward_weight")
            if li > 0:
                mask = self._acts[li] if self.layers[li - 1][4] else None
                _lib.check(lib.lcode_pred_conv3x3_backward_input(
                    _ptr(self._gacts[li + 1]), cout, _ptr(self.params[wo:]),
                    _ptr(self._gacts[li]), cin, _ptr(mask), self.N, self.N, _stream()),
                    "lcode_pred_conv3x3_backward_input")
        return self.grad

    def input_grad(self, gy):
        """Gradient of sum(gy * y) with respect to the network input."""
        self.backward(gy)
        cin, cout, wo, bo, relu = self.layers[0]
        _lib.check(_lib.load().lcode_pred_conv3x3_backward_input(
            _ptr(self._gacts[1]), cout, _ptr(self.params[wo:]), _ptr(self._gacts[0]),
            cin, None, self.N, self.N, _stream()), "lcode_pred_conv3x3_backward_input")
        return self._gacts[0]

    def train_step(self, x, t):
        """One Adam step on one sample; returns the loss as a 0-d device
        tensor (mean squared error in target units), without synchronising."""
        lib = _lib.load()
        y = self.forward(x)
        n = y.numel()
        _lib.check(lib.lcode_pred_mse_grad(_ptr(y), _ptr(t), n, _ptr(self._gacts[-1]),
                                           _ptr(self._loss_part), _ptr(self._loss),
                                           _stream()), "lcode_pred_mse_grad")
        self.backward(self._gacts[-1])
        self.t += 1
        _lib.check(lib.lcode_pred_adam(_ptr(self.params), _ptr(self.grad), _ptr(self.m),
                                       _ptr(self.v), self.nparams, self.lr, 0.9, 0.999,
                                       1e-8, self.t, _stream()), "lcode_pred_adam")
        return self._loss[0] / n

    def fit(self, samples, epochs, seed=0):
        """Adam over (inputs, targets) pairs for `epochs` shuffled passes;
        returns the mean loss of each epoch."""
        rng = np.random.default_rng(seed)
        history = []
        for _ in range(epochs):
            losses = [self.train_step(samples[k][0], samples[k][1])
                      for k in rng.permutation(len(samples))]
            history.append(torch.stack(losses).mean().item())
        return history

    # -- inference -----------------------------------------------------------
    def guess(self, s1, s2, beam):
        """The six guessed fields of slice n (FP64 device grids), for
        ``step(..., guess=[(1.0, fields)])``."""
        y = self.forward(self.inputs(s1, s2, beam, out=self._acts[0]))
        _lib.check(_lib.load().lcode_pred_guess(
            ctypes.byref(self._io(s1, s2, beam, pred=y, guess=self._guess)), _stream()),
            "lcode_pred_guess")
        return tuple(self._guess)
