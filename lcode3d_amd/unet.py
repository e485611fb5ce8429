"""The learned field predictor: a 4-level convolutional encoder-decoder
with 32 channels per level, trained on slices the plain step generates; its
output is the field guess of ``step(..., guess=...)`` (the predictor pass
of lcode.py:841-881).  The reference has no learned predictor.

Every operation runs in this package's HIP kernels (``lcode_unet_*`` in
include/lcode3d_hip.h): BF16 implicit-GEMM convolutions on the matrix cores
(forward, input and weight gradients), pooling and upsampling with their
adjoints, the loss, Adam, and the plane conversions.  torch only holds the
device buffers.

What it learns.  The guided step reproduces the plain step exactly when its
guess equals the plain predictor pass's fields F1(n); the corrector passes
errors of the guess in the transverse magnetic field through almost
unattenuated.  The network therefore predicts T(n) = F1(n) - Q(n), the gap
between the predictor's fields and the quadratic extrapolation
Q(n) = 3 F(n-1) - 3 F(n-2) + F(n-3) of the last three slices' final fields,
from the last two slices' fields, the beam and the plasma density, at a
quarter of the grid resolution (the input planes are 4 x 4 means; the
output is applied as 4 x 4 blocks): guess = Q + t * y.
"""

import ctypes
import math

import numpy as np
import torch

from . import _lib
from .lcode import _stream, is_mixed

FIELDS = ("Ex", "Ey", "Ez", "Bx", "By", "Bz")
N_IN, N_OUT = 14, 6
#: (cin, cout) of the eight convolutions C0 .. C7
LAYERS = ((14, 32), (32, 32), (32, 32), (32, 32), (64, 32), (64, 32), (64, 32),
          (32, 6))


def level0_size(grid_steps):
    return (int(grid_steps) + 3) // 4


def _p(t):
    return t.data_ptr() if t is not None else None


class UNetPredictor:
    """The network, its Adam state and its buffers on the current device."""

    def __init__(self, grid_steps, seed=0, lr=1e-3):
        lib = _lib.load()
        self.N = int(grid_steps)
        self.S = level0_size(self.N)
        self.dev = torch.device("cuda", torch.cuda.current_device())
        sizes = (ctypes.c_int64 * 5)()
        _lib.check(lib.lcode_unet_sizes(self.S, sizes), "lcode_unet_sizes")
        self.nparams, npack, nact, npart, nloss = (int(v) for v in sizes)
        f32, dev = torch.float32, self.dev
        self.params = torch.zeros(self.nparams, dtype=f32, device=dev)
        self.grad = torch.zeros_like(self.params)
        self.m = torch.zeros_like(self.params)
        self.v = torch.zeros_like(self.params)
        self.wpack = torch.zeros(npack, dtype=torch.int16, device=dev)
        self.act = torch.zeros(nact, dtype=torch.int16, device=dev)
        self.y = torch.zeros(self.S * self.S * 16, dtype=f32, device=dev)
        self.partial = torch.zeros(npart, dtype=f32, device=dev)
        self.loss_part = torch.zeros(nloss, dtype=f32, device=dev)
        self.loss = torch.zeros(1, dtype=f32, device=dev)
        self.t = 0
        self.lr = float(lr)
        self.scale = np.ones(N_IN, dtype=np.float32)
        self.tscale = np.ones(N_OUT, dtype=np.float32)
        self._x = torch.zeros((self.S, self.S, 32), dtype=torch.int16, device=dev)
        self._guess = {}
        #: guess(): evaluate the network on every `refresh_every`-th call and
        #: apply its last output in between (the learned correction changes
        #: on the plasma-period scale, hundreds of slices)
        self.refresh_every = 1
        #: the extrapolation the network corrects: 'quadratic' 3 F(n-1) -
        #: 3 F(n-2) + F(n-3), or 'alternating' 2 F(n-1) - 2 F(n-3) + F(n-4)
        #: (Variant A's sign-alternating part carried on as the step does, so
        #: the learned correction is smooth from slice to slice and can be
        #: reused between evaluations); training and guesses must agree
        self.base = "quadratic"
        self._calls = 0
        self.set_params(self.init_params(seed))

    # -- parameters ----------------------------------------------------------
    @staticmethod
    def init_params(seed=0):
        """He-normal weights (the output layer scaled by 0.1), zero biases:
        [(w[cout][cin][3][3], b[cout])] as float64 NumPy."""
        rng = np.random.default_rng(seed)
        out = []
        for k, (cin, cout) in enumerate(LAYERS):
            std = math.sqrt(2.0 / (cin * 9)) * (0.1 if k == len(LAYERS) - 1 else 1.0)
            out.append((rng.standard_normal((cout, cin, 3, 3)) * std, np.zeros(cout)))
        return out

    def set_params(self, layers):
        flat = np.concatenate([np.concatenate([w.ravel(), b.ravel()]) for w, b in layers])
        assert flat.size == self.nparams
        self.params.copy_(torch.as_tensor(flat.astype(np.float32)))
        _lib.check(_lib.load().lcode_unet_pack(ctypes.byref(self._net()), _stream()),
                   "lcode_unet_pack")

    def layer_arrays(self, source=None):
        """params (or a tensor of the same layout) as [(w, b)] float64."""
        p = (self.params if source is None else source).detach().cpu().numpy().astype(np.float64)
        out, off = [], 0
        for cin, cout in LAYERS:
            n = cout * cin * 9
            out.append((p[off:off + n].reshape(cout, cin, 3, 3), p[off + n:off + n + cout]))
            off += n + cout
        return out

    def tensor(self, index, level):
        """Activation/gradient tensor `index` (enum order of unet.hip) as a
        [s, s, 32] int16 view of bfloat16 bits."""
        lib = _lib.load()
        off = lib.lcode_unet_tensor_offset(self.S, index)
        s = self.S
        for _ in range(level):
            s = (s + 1) // 2
        return self.act[off:off + s * s * 32].view(s, s, 32)

    # -- buffers as the C structs -------------------------------------------
    def _net(self):
        n = _lib.LcodeUnet()
        n.S = self.S
        for name in ("params", "grad", "adam_m", "adam_v", "wpack", "act", "y",
                     "partial", "loss_part", "loss"):
            t = {"adam_m": self.m, "adam_v": self.v}.get(name, getattr(self, name, None))
            setattr(n, name, t.data_ptr())
        return n

    def _io(self, s1, s2, s3=None, beam=0, nxt=None, planes=None, planes_f32=None,
            targets=None, guess=None, s4=None):
        io = _lib.LcodeUnetIo()
        f32 = s1.Ex.dtype == torch.float32
        for f, name in enumerate(FIELDS):
            io.f1[f] = getattr(s1, name).data_ptr()
            io.f2[f] = getattr(s2, name).data_ptr()
            if s3 is not None:
                io.f3[f] = getattr(s3, name).data_ptr()
            if s4 is not None:
                io.f4[f] = getattr(s4, name).data_ptr()
            if nxt is not None:
                io.next[f] = nxt[f].data_ptr()
            if guess is not None:
                io.guess[f] = guess[f].data_ptr()
        io.ro = s1.ro.data_ptr()
        if isinstance(beam, torch.Tensor):
            if beam.dtype != torch.float64 or not beam.is_contiguous():
                raise ValueError("the beam must be a contiguous float64 device array")
            io.beam = beam.data_ptr()
        else:
            io.beam_scalar = float(beam)
        io.f32 = 1 if f32 else 0
        io.grid_steps = self.N
        for c in range(N_IN):
            io.scale[c] = float(self.scale[c])
        for c in range(N_OUT):
            io.tscale[c] = float(self.tscale[c])
        io.planes = _p(planes)
        io.planes_f32 = _p(planes_f32)
        io.targets = _p(targets)
        return io

    def _run_io(self, io, which):
        _lib.check(_lib.load().lcode_unet_io(ctypes.byref(self._net()), ctypes.byref(io),
                                             which, _stream()), "lcode_unet_io")

    # -- data ------------------------------------------------------------------
    def raw_inputs(self, s1, s2, beam):
        """Unscaled FP32 input planes [S, S, 16] (14 used) of states n-1,
        n-2 and beam n."""
        out = torch.empty((self.S, self.S, 16), dtype=torch.float32, device=self.dev)
        scale, self.scale = self.scale, np.ones(N_IN, dtype=np.float32)
        try:
            self._run_io(self._io(s1, s2, beam=beam, planes_f32=out), 0)
        finally:
            self.scale = scale
        return out

    def _s4(self, s4):
        if self.base == "alternating":
            if s4 is None:
                raise ValueError("the alternating base needs slice n-4")
            return s4
        return None

    def raw_targets(self, s1, s2, s3, f1, s4=None):
        """Unscaled FP32 targets [S, S, 6]: 4 x 4 means of F1(n) - Q(n)
        (Q the base extrapolation; s4 = slice n-4 for the alternating one)."""
        out = torch.empty((self.S, self.S, 6), dtype=torch.float32, device=self.dev)
        tscale, self.tscale = self.tscale, np.ones(N_OUT, dtype=np.float32)
        try:
            self._run_io(self._io(s1, s2, s3, nxt=f1, targets=out, s4=self._s4(s4)), 1)
        finally:
            self.tscale = tscale
        return out

    def fit_scales(self, raw_x, raw_t):
        """Per-channel RMS of the raw inputs and targets (lists of tensors)."""
        xs = torch.stack(raw_x)[..., :N_IN].double()
        ts = torch.stack(raw_t).double()
        sx = xs.pow(2).mean(dim=(0, 1, 2)).sqrt().cpu().numpy()
        st = ts.pow(2).mean(dim=(0, 1, 2)).sqrt().cpu().numpy()
        self.scale = np.where(sx > 0, sx, 1.0).astype(np.float32)
        self.tscale = np.where(st > 0, st, 1.0).astype(np.float32)

    def encode(self, raw_x):
        """Scaled bfloat16 input planes [S, S, 32] (int16 bits) from raw ones."""
        x = torch.zeros((self.S, self.S, 32), dtype=torch.float32, device=self.dev)
        x[..., :N_IN] = raw_x[..., :N_IN] / torch.as_tensor(self.scale, device=self.dev)
        return x.to(torch.bfloat16).view(torch.int16)

    def scaled_targets(self, raw_t):
        return (raw_t / torch.as_tensor(self.tscale, device=self.dev)).contiguous()

    # -- passes ----------------------------------------------------------------
    def forward(self, x):
        """y = net(x); returns the [S, S, 6] view of the output."""
        _lib.check(_lib.load().lcode_unet_forward(ctypes.byref(self._net()),
                                                  x.data_ptr(), _stream()),
                   "lcode_unet_forward")
        return self.y.view(self.S, self.S, 16)[..., :N_OUT]

    def infer(self, x):
        """forward() without the copies the backward pass needs."""
        _lib.check(_lib.load().lcode_unet_infer(ctypes.byref(self._net()),
                                                x.data_ptr(), _stream()),
                   "lcode_unet_infer")
        return self.y.view(self.S, self.S, 16)[..., :N_OUT]

    def backward(self, x, t):
        _lib.check(_lib.load().lcode_unet_backward(ctypes.byref(self._net()), x.data_ptr(),
                                                   t.data_ptr(), _stream()),
                   "lcode_unet_backward")

    def train_step(self, x, t):
        """Forward, loss, backward and one Adam step; returns the loss as a
        device scalar (no synchronisation)."""
        self.forward(x)
        self.backward(x, t)
        self.t += 1
        _lib.check(_lib.load().lcode_unet_adam(ctypes.byref(self._net()), self.t,
                                               self.lr, _stream()), "lcode_unet_adam")
        return self.loss.clone()

    def fit(self, samples, epochs, seed=0):
        """Adam over (x, t) pairs in a seeded random order; the mean loss of
        each epoch."""
        rng = np.random.default_rng(seed)
        hist = []
        for _ in range(epochs):
            losses = [self.train_step(*samples[k]) for k in rng.permutation(len(samples))]
            hist.append(float(torch.stack(losses).mean().item()))
        return hist

    def guess(self, s1, s2, s3, beam, s4=None):
        """Field guess of slice n from states n-1, n-2, n-3 (and n-4 for the
        alternating base; without it the first slices use the quadratic
        one) and beam n: the six fields in the states' grid type (buffers
        reused per call).  The network runs on every refresh_every-th call;
        the other calls add its last output to this slice's extrapolation."""
        key = s1.Ex.dtype
        if key not in self._guess:
            self._guess[key] = tuple(torch.empty((self.N, self.N), dtype=key, device=self.dev)
                                     for _ in FIELDS)
        g = self._guess[key]
        if self._calls % max(int(self.refresh_every), 1) == 0:
            self._run_io(self._io(s1, s2, beam=beam, planes=self._x), 0)
            self.infer(self._x)
        self._calls += 1
        older = s4 if self.base == "alternating" else None
        self._run_io(self._io(s1, s2, s3, guess=g, s4=older), 2)
        return g


def collect_training_data(lc, net, config, runs, slices, every=1, first=3,
                          beam_fn=None, on_slice=None):
    """Samples of the plain FP64 method for `net`: for each (amplitude,
    sigma) of `runs`, one run of `slices` slices from init; every `every`
    slices from slice `first` on, the raw input planes of states n-1, n-2
    and beam n and the raw target F1(n) - Q(n), with F1(n) the plain
    predictor pass's fields (lcode.predictor_fields) and Q(n) the net's
    base extrapolation (four states back for the alternating base).
    Returns the list of (raw_x, raw_t) and the per-channel sums of
    squares."""
    if beam_fn is None:
        from .beams import cosine_gaussian as beam_fn
    samples = []
    sx = torch.zeros(N_IN, dtype=torch.float64, device=net.dev)
    st_ = torch.zeros(N_OUT, dtype=torch.float64, device=net.dev)
    for amp, sig in runs:
        _, _, const, vp, st = lc.init(config)
        hist = []
        depth = 4 if net.base == "alternating" else 3
        for i in range(slices):
            beam = beam_fn(config, i, amplitude=amp, sigma=sig)
            if len(hist) >= depth and i >= first and (i - first) % every == 0:
                f1 = lc.predictor_fields(config, const, vp, hist[0], beam)
                x = net.raw_inputs(hist[0], hist[1], beam)
                t = net.raw_targets(hist[0], hist[1], hist[2], f1, *hist[3:4])
                sx += x[..., :N_IN].double().pow(2).sum(dim=(0, 1))
                st_ += t.double().pow(2).sum(dim=(0, 1))
                samples.append((x, t))
            st = lc.step(config, const, vp, st, beam)
            hist.insert(0, st)
            del hist[depth:]
            if on_slice:
                on_slice(i)
    return samples, sx, st_


def train(net, samples, sums, epochs, seed=0):
    """Scales from the sums of squares, the samples encoded (bfloat16
    inputs, scaled targets) in place of the raw ones, then `epochs` of Adam;
    returns the mean loss of every epoch."""
    sx, st = sums
    n = max(len(samples), 1) * net.S * net.S
    rx = (sx / n).sqrt().cpu().numpy()
    rt = (st / n).sqrt().cpu().numpy()
    net.scale = np.where(rx > 0, rx, 1.0).astype(np.float32)
    net.tscale = np.where(rt > 0, rt, 1.0).astype(np.float32)
    for k, (x, t) in enumerate(samples):
        samples[k] = (net.encode(x), net.scaled_targets(t))
    return net.fit(samples, epochs, seed=seed)
