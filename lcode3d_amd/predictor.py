"""A learned field predictor for the guided step: a small convolutional
network, trained on slices the plain step generates, whose output replaces
the extrapolated field guess of ``step(..., guess=...)`` (the predictor
pass of lcode.py:841-881).  The reference has no learned predictor.

Every operation runs in this package's own HIP kernels
(``lcode_pred_*`` in include/lcode3d_hip.h): 3 x 3 convolutions forward and
backward, MSE gradient, Adam, and the FP64 <-> FP32 plane conversions.
torch only holds the device buffers.

Network: 14 input planes -> conv3x3 -> ReLU -> conv3x3 -> ReLU -> conv3x3
-> 6 planes.  With Q = 3 F(n-1) - 3 F(n-2) + F(n-3), the quadratic
extrapolation of slice n's six fields, and D = F(n-1) - F(n-2), the inputs
are Q and D (each divided by a per-field scale), the beam of slice n and
the plasma density of slice n-1; the output is the change of slice n's
fields from Q in units of per-field target scales, and the guess is
Q + scale * output.  A zero network reproduces the quadratic extrapolation.
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
    def _io(self, s1, s2, s3, beam, planes=None, nxt=None, pred=None, guess=None):
        io = _lib.LcodePredIo()
        for f, name in enumerate(FIELDS):
            io.f1[f] = getattr(s1, name).data_ptr()
            io.f2[f] = getattr(s2, name).data_ptr()
            io.f3[f] = getattr(s3, name).data_ptr()
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

    def inputs(self, s1, s2, s3, beam, out=None):
        """Input planes [14, N, N] from states n-1, n-2, n-3 and beam n."""
        out = out if out is not None else torch.empty((N_IN, self.N, self.N),
                                                      dtype=torch.float32, device=self.dev)
        _lib.check(_lib.load().lcode_pred_inputs(
            ctypes.byref(self._io(s1, s2, s3, beam, planes=out)), _stream()),
            "lcode_pred_inputs")
        return out

    def targets(self, s1, s2, s3, nxt, out=None):
        """Target planes [6, N, N]: (F(n) - Q) / tscale."""
        out = out if out is not None else torch.empty((N_OUT, self.N, self.N),
                                                      dtype=torch.float32, device=self.dev)
        _lib.check(_lib.load().lcode_pred_targets(
            ctypes.byref(self._io(s1, s2, s3, 0.0, planes=out, nxt=nxt)), _stream()),
            "lcode_pred_targets")
        return out

    def fit_scales(self, runs):
        """Per-field RMS over runs (sequences of consecutive states) of Q,
        of D and of the target F(n) - Q; the beam and density scales."""
        acc = self.scale_accumulator()
        for states in runs:
            for k in range(3, len(states)):
                acc.add(states[k], states[k - 1], states[k - 2], states[k - 3])
        acc.finish()

    def scale_accumulator(self):
        """Accumulates the statistics of fit_scales window by window
        (states n, n-1, n-2, n-3), so long runs need not be kept."""
        net = self

        class _Acc:
            def __init__(self):
                self.q2, self.d2, self.t2 = np.zeros(6), np.zeros(6), np.zeros(6)
                self.ro2, self.cnt = 0.0, 0

            def add(self, s0, s1, s2, s3):
                for f, n in enumerate(FIELDS):
                    F1, F2, F3 = getattr(s1, n), getattr(s2, n), getattr(s3, n)
                    q = 3 * F1 - 3 * F2 + F3
                    self.q2[f] += (q ** 2).mean().item()
                    self.d2[f] += ((F1 - F2) ** 2).mean().item()
                    self.t2[f] += ((getattr(s0, n) - q) ** 2).mean().item()
                self.ro2 += (s1.ro ** 2).mean().item()
                self.cnt += 1

            def finish(self):
                def rms(a):
                    return np.maximum(np.sqrt(np.asarray(a) / max(self.cnt, 1)), 1e-30)
                net.scale[:6] = rms(self.q2)
                net.scale[6:12] = rms(self.d2)
                net.scale[12] = 0.1
                net.scale[13] = float(rms([self.ro2])[0])
                net.tscale[:] = rms(self.t2)
        return _Acc()

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
        """Adam over (inputs, targets) pairs for `epochs` shuffled passes,
        the learning rate decaying linearly to a tenth; returns the mean loss
        of each epoch."""
        rng = np.random.default_rng(seed)
        lr0 = self.lr
        history = []
        for e in range(epochs):
            self.lr = lr0 * (1 - 0.9 * e / max(epochs - 1, 1))
            losses = [self.train_step(samples[k][0], samples[k][1])
                      for k in rng.permutation(len(samples))]
            history.append(torch.stack(losses).mean().item())
        self.lr = lr0
        return history

    # -- inference -----------------------------------------------------------
    def guess(self, s1, s2, s3, beam):
        """The six guessed fields of slice n (FP64 device grids) from the
        states of slices n-1, n-2, n-3, for ``step(..., guess=[(1.0, fields)])``."""
        y = self.forward(self.inputs(s1, s2, s3, beam, out=self._acts[0]))
        _lib.check(_lib.load().lcode_pred_guess(
            ctypes.byref(self._io(s1, s2, s3, beam, pred=y, guess=self._guess)), _stream()),
            "lcode_pred_guess")
        return tuple(self._guess)
