"""WSM6 column surrogate: data, training and inline use.

A 369 -> 512 -> 512 -> 287 ReLU network maps one microphysics column (41
levels x T, p, rho, qv, qc, qr, qi, qs, qg) to the tendencies the WSM6 scheme
(pywrf/wrf_physics.py:137-1214) gives that column over one step (41 levels x
T, qv, qc, qr, qi, qs, qg).  Everything that touches the network runs in the
library's own kernels: BF16 matrix-core GEMMs with FP32 accumulation for the
forward and backward passes, the loss, the bias gradients and Adam
(csrc/mlp.hip); the inline step packs the columns, runs the three layers,
applies the tendencies with the fallback checks and re-runs the plain scheme
on the columns that fail them (csrc/surrogate.hip).  torch provides the
device memory, the shuffling of the training set and the bookkeeping.

Training data are column pairs written by the plain WSM6 kernel inside the
step: the state of every tile column just before microphysics_driver and the
change it made, over seeded runs of the 40 x 40 x 41 case (TRAIN_RECIPE).

    from pywrf_amd import surrogate
    data = surrogate.collect_pairs(runs=50, steps=140)       # about 5.1 M pairs
    net = surrogate.Surrogate.fit(data, epochs=8)
    ops = surrogate.surrogate_ops(net)                       # the step's functions, WSM6 replaced
"""
import ctypes
import math
import time

import torch

from . import _lib
from ._lib import call, grid, ptr, stream

NLEV, NIN, NOUT = 41, 9, 7
FIN, FOUT = NLEV * NIN, NLEV * NOUT          # 369, 287
KIN, HID, NO = 384, 512, 288                 # padded widths
IN_VARS = ("T", "p", "rho", "qv", "qc", "qr", "qi", "qs", "qg")
OUT_VARS = ("T", "qv", "qc", "qr", "qi", "qs", "qg")

# seeded 40 x 40 x 41 runs of the reference's CPU case (one 500 m hill) with
# the seed choosing six warm bubbles of 1-3 K and cloud / precipitation blobs
# of all five condensate species, so that every process of the scheme occurs
# in the training set; train_recipe() also draws the flow per seed (uniform
# wind, an upper-level jet, u/v noise) to cover the flows of the 240 x 240
# case
TRAIN_RECIPE = dict(n=40, hills=[(0.5, 0.5, 500.0, 45e3)], jet=None, u0=10.0, noise=0.5, bubbles="random",
                    clouds=6)


def train_recipe(seed):
    """TRAIN_RECIPE with the seed's flow: u0 ~ U(0, 12) m/s, with probability
    1/2 a jet of U(5, 25) m/s at 10 km, noise ~ U(0.3, 2) m/s."""
    import numpy as np
    rng = np.random.default_rng(10_000 + seed)
    u0, jet_amp, noise = rng.uniform(0, 12), rng.uniform(5, 25), rng.uniform(0.3, 2.0)
    jet = (float(jet_amp), 10e3, 5e3) if rng.uniform() < 0.5 else None
    return dict(TRAIN_RECIPE, seed=seed, u0=float(u0), jet=jet, noise=float(noise))


def _vp(t):
    return ctypes.c_void_p(t.data_ptr()) if t is not None else ctypes.c_void_p(0)


def _check_bf16(t, rows, cols):
    if t.dtype != torch.bfloat16 or not t.is_cuda or not t.is_contiguous() or tuple(t.shape) != (rows, cols):
        raise TypeError("expected a contiguous bf16 CUDA tensor of shape %s, got %s %s %s"
                        % ((rows, cols), t.dtype, t.device, tuple(t.shape)))


def gemm(epi, A, B, C, bias=None, H=None, ksplit=1):
    """C = A B^T with the library's matrix-core kernel (see pywrf_mlp_gemm):
    A bf16 [M x K], B bf16 [N x K]; C f32 or bf16 [M x N]."""
    M, K = A.shape
    N = B.shape[0]
    rows, cols = (N, M) if epi == 4 else (M, N)          # epilogue 4 writes C^T
    if B.shape[1] != K or C.shape[0] != rows or C.shape[1] < cols:
        raise IndexError("gemm: shapes %s x %s^T -> %s" % (tuple(A.shape), tuple(B.shape), tuple(C.shape)))
    _check_bf16(A, M, K)
    _check_bf16(B, N, K)
    if bias is not None and (bias.numel() < N or bias.dtype != torch.float32):
        raise IndexError("gemm: bias too short")
    ldh = 0
    if H is not None:
        _check_bf16(H, M, H.shape[1])
        ldh = H.shape[1]
    call("pywrf_mlp_gemm", epi, M, N, K, _vp(A), K, _vp(B), K, _vp(C), C.shape[1], _vp(bias), _vp(H), ldh, ksplit,
         stream())


def transpose(x, out):
    R, C = x.shape
    if tuple(out.shape) != (C, R):
        raise IndexError("transpose: %s -> %s" % (tuple(x.shape), tuple(out.shape)))
    call("pywrf_mlp_transpose", R, C, _vp(x), C, _vp(out), R, stream())


def rowsum(x, out):
    R, C = x.shape
    call("pywrf_mlp_rowsum", R, C, _vp(x), C, _vp(out), stream())


def _to_bf16(x):
    return x.to(torch.bfloat16).contiguous()


class MLP:
    """Parameters of the network: FP32 masters, their BF16 copies (W and
    W^T) and the Adam moments.  Padded input columns (369..383) and the padded
    output row (287) are zero and stay zero (their gradients are zero)."""

    LAYERS = ((HID, KIN, FIN), (HID, HID, HID), (NO, HID, HID))

    def __init__(self, seed=0, device="cuda"):
        gen = torch.Generator().manual_seed(seed)
        self.W, self.b = [], []
        for (rows, cols, fan_in), valid in zip(self.LAYERS, (HID, HID, FOUT)):
            w = torch.randn(rows, cols, generator=gen) * math.sqrt(2.0 / fan_in)
            w[:, fan_in:] = 0.0
            w[valid:] = 0.0
            self.W.append(w.to(device))
            self.b.append(torch.zeros(rows, device=device))
        self.mW = [torch.zeros_like(w) for w in self.W]
        self.vW = [torch.zeros_like(w) for w in self.W]
        self.mb = [torch.zeros_like(b) for b in self.b]
        self.vb = [torch.zeros_like(b) for b in self.b]
        self.t = 0
        self.sync()

    def sync(self):
        """BF16 copies of the FP32 masters."""
        self.Wb = [_to_bf16(w) for w in self.W]
        self.Wt = [_to_bf16(w.t()) for w in self.W]

    def state_dict(self):
        return {"W%d" % n: w for n, w in enumerate(self.W)} | {"b%d" % n: b for n, b in enumerate(self.b)}

    def load_state_dict(self, d):
        for n in range(3):
            self.W[n].copy_(d["W%d" % n])
            self.b[n].copy_(d["b%d" % n])
        self.sync()


class Workspace:
    """Activation buffers of one batch of M columns (M a multiple of 32)."""

    def __init__(self, M, device="cuda"):
        if M % 32:
            raise ValueError("batch size must be a multiple of 32")
        bf = dict(dtype=torch.bfloat16, device=device)
        self.M = M
        self.H1 = torch.empty(M, HID, **bf)
        self.H2 = torch.empty(M, HID, **bf)
        self.Y = torch.empty(M, NO, dtype=torch.float32, device=device)
        self.dY = torch.empty(M, NO, **bf)
        self.dH2 = torch.empty(M, HID, **bf)
        self.dH1 = torch.empty(M, HID, **bf)
        self.Xt = torch.empty(KIN, M, **bf)
        self.H1t = torch.empty(HID, M, **bf)
        self.H2t = torch.empty(HID, M, **bf)
        self.dYt = torch.empty(NO, M, **bf)
        self.dH1t = torch.empty(HID, M, **bf)
        self.dH2t = torch.empty(HID, M, **bf)
        self.gW = [torch.zeros(r, c, device=device) for r, c, _ in MLP.LAYERS]
        self.gb = [torch.zeros(r, device=device) for r, _, _ in MLP.LAYERS]
        self.loss = torch.zeros(1, device=device)


def forward(net, X, ws):
    """Y (FP32 [M x 288], standardised tendencies) of the BF16 inputs X [M x 384]."""
    M = X.shape[0]
    gemm(1, X, net.Wb[0], ws.H1[:M], net.b[0])
    gemm(1, ws.H1[:M], net.Wb[1], ws.H2[:M], net.b[1])
    gemm(0, ws.H2[:M], net.Wb[2], ws.Y[:M], net.b[2])
    return ws.Y[:M]


def backward(net, X, T, w, ws, wsum, ksplit=None):
    """Loss and parameter gradients (ws.gW, ws.gb) of one batch: the weighted
    mean squared error sum_n w[n] (Y - T)^2 / (M wsum) of the forward pass
    just run on X (standardised targets T, FP32 [M x 288]; wsum = sum w)."""
    M = ws.M
    ks = ksplit or max(1, min(64, M // 256))
    ws.loss.zero_()
    call("pywrf_mlp_mse", M, FOUT, _vp(ws.Y), NO, _vp(T), NO, _vp(w), 1.0 / (M * wsum), _vp(ws.dY), NO,
         _vp(ws.loss), stream())
    # layer 3: dW3 = dY^T H2, db3 = sum dY, dH2 = (dY W3) [H2 > 0]
    transpose(ws.dY, ws.dYt)
    transpose(ws.H2, ws.H2t)
    for g in ws.gW:
        g.zero_()
    gemm(3, ws.dYt, ws.H2t, ws.gW[2], ksplit=ks)
    rowsum(ws.dYt, ws.gb[2])
    gemm(2, ws.dY, net.Wt[2], ws.dH2, H=ws.H2)
    # layer 2
    transpose(ws.dH2, ws.dH2t)
    transpose(ws.H1, ws.H1t)
    gemm(3, ws.dH2t, ws.H1t, ws.gW[1], ksplit=ks)
    rowsum(ws.dH2t, ws.gb[1])
    gemm(2, ws.dH2, net.Wt[1], ws.dH1, H=ws.H1)
    # layer 1
    transpose(ws.dH1, ws.dH1t)
    transpose(X, ws.Xt)
    gemm(3, ws.dH1t, ws.Xt, ws.gW[0], ksplit=ks)
    rowsum(ws.dH1t, ws.gb[0])
    return ws.loss


def loss_weights(device="cuda"):
    """Per-output weights of the loss: 1 on the 7 x 40 tendencies, 0 on the
    padding (level 40 of each field, output 287)."""
    w = torch.ones(NO, device=device)
    w[FOUT:] = 0.0
    for v in range(NOUT):
        w[v * NLEV + NLEV - 1] = 0.0
    return w


def adam_step(net, ws, lr, beta1=0.9, beta2=0.999, eps=1e-8):
    net.t += 1
    for n in range(3):
        W = net.W[n]
        call("pywrf_mlp_adam", W.shape[0], W.shape[1], _vp(W), _vp(ws.gW[n]), _vp(net.mW[n]), _vp(net.vW[n]),
             float(lr), beta1, beta2, eps, 1.0, net.t, _vp(net.Wb[n]), _vp(net.Wt[n]), stream())
        b = net.b[n]
        call("pywrf_mlp_adam", 1, b.shape[0], _vp(b), _vp(ws.gb[n]), _vp(net.mb[n]), _vp(net.vb[n]), float(lr),
             beta1, beta2, eps, 1.0, net.t, None, None, stream())


# ----------------------------------------------------------------- data
class Pairs:
    """Column pairs on the device: raw inputs (FP32 [N x 369], feature
    var * 41 + k of IN_VARS, level 40 zero) and WSM6 tendencies (FP32
    [N x 287], var * 41 + k of OUT_VARS, per second)."""

    def __init__(self, X, Y):
        self.X, self.Y = X, Y

    def __len__(self):
        return self.X.shape[0]


def _tile(th):
    """The scheme's tile slices of a (41, ny, nx) field (wsm6's columns: the
    physics index set, one point in from the step's tile on each side, and
    levels [0, 40))."""
    return slice(6, th.shape[1] - 7), slice(6, th.shape[2] - 7)


def _columns(field, js, is_, km=NLEV - 1):
    """(ncol, 41) view-copy of a field's tile columns, level 40 zero."""
    c = field[:km, js, is_].permute(1, 2, 0).reshape(-1, km)
    return torch.nn.functional.pad(c, (0, NLEV - km))


class _Recorder:
    """Wraps microphysics_driver: the inputs and the change of each call are
    appended to the lists."""

    def __init__(self, ops):
        self._ops = ops
        self.inputs, self.tends = [], []

    def __getattr__(self, name):
        return getattr(self._ops, name)

    def microphysics_driver(self, th, rho, pi_phy, p, ht, dz8w, p8w, dt, *rest):
        js, is_ = _tile(th)
        q = rest[2:8]
        before_T = th * pi_phy
        ins = [before_T, p, rho] + list(q)
        X = torch.cat([_columns(f, js, is_) for f in ins], dim=1)
        old = [before_T.clone()] + [x.clone() for x in q]
        out = self._ops.microphysics_driver(th, rho, pi_phy, p, ht, dz8w, p8w, dt, *rest)
        new = [th * pi_phy] + list(q)
        Y = torch.cat([_columns((a - b) / dt, js, is_) for a, b in zip(new, old)], dim=1)
        self.inputs.append(X)
        self.tends.append(Y)
        return out


def collect_pairs(runs=50, steps=140, seed0=100, log=None):
    """Column pairs of the plain WSM6 kernel over `runs` seeded runs of
    TRAIN_RECIPE, `steps` RK3 steps each (729 columns per step at 40 x 40)."""
    import pywrf_amd
    from .gen import synthetic_state
    from .step import rk3_step, RunConfig
    cfg = RunConfig()
    Xs, Ys = [], []
    t0 = time.perf_counter()
    for r in range(runs):
        s = synthetic_state(train_recipe(seed0 + r), device="cuda")
        rec = _Recorder(pywrf_amd.ops())
        for _ in range(steps):
            rk3_step(s, rec, cfg)
        X, Y = torch.cat(rec.inputs), torch.cat(rec.tends)
        ok = torch.isfinite(X).all(1) & torch.isfinite(Y).all(1)
        dropped = int((~ok).sum())
        if dropped:
            X, Y = X[ok], Y[ok]
        Xs.append(X)
        Ys.append(Y)
        if log:
            log("collect: run %d/%d, %d pairs (%d non-finite dropped), %.1f s"
                % (r + 1, runs, sum(x.shape[0] for x in Xs), dropped, time.perf_counter() - t0))
    return Pairs(torch.cat(Xs), torch.cat(Ys))


# ----------------------------------------------------------------- the trained network
class Surrogate:
    """The network with its standardisation and fallback tolerances."""

    QTOL = 1e-7      # kg/kg: smaller negative mixing ratios fail the column
    WTOL = 1e-5      # relative column water gain that fails the column
    ENVELOPE = 1.05  # widening of the training envelope (see pywrf_surrogate_t)
    DTMAX = 0.25     # K per call: larger heating or cooling at a level fails the column

    def __init__(self, net, in_mean, in_rstd, out_mean, out_std, in_lo=None, in_hi=None, out_absmax=None):
        self.net = net
        dev = net.W[0].device
        self.in_lo = torch.zeros(KIN, device=dev)
        self.in_hi = torch.zeros(KIN, device=dev)
        self.out_absmax = torch.zeros(NO, device=dev)
        if in_lo is not None:
            self.in_lo[:FIN], self.in_hi[:FIN], self.out_absmax[:FOUT] = in_lo, in_hi, out_absmax
        self.in_mean = torch.zeros(KIN, device=dev)
        self.in_rstd = torch.zeros(KIN, device=dev)
        self.in_mean[:FIN] = in_mean
        self.in_rstd[:FIN] = in_rstd
        self.out_mean = torch.zeros(NO, device=dev)
        self.out_std = torch.zeros(NO, device=dev)
        self.out_mean[:FOUT] = out_mean
        self.out_std[:FOUT] = out_std
        self.history = []

    # smallest standard deviations used to standardise, per field: a feature
    # that (nearly) never varies in the training set must not be scaled up
    IN_FLOOR = dict(T=1e-2, p=1.0, rho=1e-4, qv=1e-6, qc=1e-7, qr=1e-7, qi=1e-7, qs=1e-7, qg=1e-7)
    OUT_FLOOR = dict(T=1e-6, qv=1e-10, qc=1e-10, qr=1e-10, qi=1e-10, qs=1e-10, qg=1e-10)

    @classmethod
    def untrained(cls, data, seed=0):
        """The network at its seeded initialisation, with the standardisation
        of `data` (per-feature mean and standard deviation, the deviation
        floored per field by IN_FLOOR / OUT_FLOOR)."""
        dev = data.X.device
        fin = torch.tensor([cls.IN_FLOOR[v] for v in IN_VARS for _ in range(NLEV)], device=dev)
        fout = torch.tensor([cls.OUT_FLOOR[v] for v in OUT_VARS for _ in range(NLEV)], device=dev)
        X, Y = data.X.double(), data.Y.double()
        in_mean, in_std = X.mean(0).float(), torch.maximum(X.std(0).float(), fin)
        out_mean, out_std = Y.mean(0).float(), torch.maximum(Y.std(0).float(), fout)
        # the envelope: observed ranges, widened by the floors (a feature that never
        # varies still admits values within its floor of the one it showed)
        lo, hi = data.X.min(0).values - fin, data.X.max(0).values + fin
        amax = torch.maximum(data.Y.abs().max(0).values, 10.0 * fout)
        return cls(MLP(seed, device=dev), in_mean, 1.0 / in_std, out_mean, out_std, lo, hi, amax)

    def standardise(self, data):
        X = torch.zeros(len(data), KIN, dtype=torch.bfloat16, device=data.X.device)
        X[:, :FIN] = ((data.X - self.in_mean[:FIN]) * self.in_rstd[:FIN]).to(torch.bfloat16)
        T = torch.zeros(len(data), NO, device=data.X.device)
        T[:, :FOUT] = (data.Y - self.out_mean[:FOUT]) / self.out_std[:FOUT]
        return X, T

    @classmethod
    def fit(cls, data, epochs=8, batch=8192, lr=1e-3, seed=0, log=None, valid=None):
        """Train on the GPU with the library's kernels (Adam, cosine decay)."""
        self = cls.untrained(data, seed)
        X, T = self.standardise(data)
        w = loss_weights(X.device)
        wsum = float(w.sum())
        self.loss_weight = w
        ws = Workspace(batch, device=X.device)
        n = X.shape[0]
        nb = n // batch
        total = epochs * nb
        gen = torch.Generator(device=X.device).manual_seed(seed)
        step = 0
        t0 = time.perf_counter()
        for ep in range(epochs):
            perm = torch.randperm(n, device=X.device, generator=gen)
            acc = torch.zeros((), device=X.device)
            for bi in range(nb):
                idx = perm[bi * batch:(bi + 1) * batch]
                xb = X.index_select(0, idx)
                tb = T.index_select(0, idx)
                forward(self.net, xb, ws)
                acc += backward(self.net, xb, tb, w, ws, wsum)[0]
                lr_t = lr * 0.5 * (1.0 + math.cos(math.pi * step / total))
                adam_step(self.net, ws, lr_t)
                step += 1
            loss = float(acc) / max(nb, 1)
            rec = {"epoch": ep + 1, "train_loss": loss, "seconds": time.perf_counter() - t0}
            if valid is not None:
                rec["valid_loss"] = self.loss(valid, batch)
            self.history.append(rec)
            if log:
                log("fit: " + " ".join("%s %.4g" % kv for kv in rec.items()))
        self.train_seconds = time.perf_counter() - t0
        self.train_steps = step
        return self

    def loss(self, data, batch=8192):
        """Weighted mean squared error (standardised) over a set of pairs."""
        X, T = self.standardise(data)
        ws = Workspace(batch, device=X.device)
        w = loss_weights(X.device)
        wsum = float(w.sum())
        tot, cnt = 0.0, 0
        for b0 in range(0, X.shape[0] - batch + 1, batch):
            forward(self.net, X[b0:b0 + batch], ws)
            tot += float(backward(self.net, X[b0:b0 + batch], T[b0:b0 + batch], w, ws, wsum)[0]) * batch
            cnt += batch
        return tot / max(cnt, 1)

    def predict(self, data, batch=8192):
        """De-standardised tendencies (FP32 [N x 287]) of a set of pairs."""
        X, _ = self.standardise(data)
        ws = Workspace(batch, device=X.device)
        out = []
        for b0 in range(0, X.shape[0], batch):
            xb = X[b0:b0 + batch]
            Y = forward(self.net, xb, ws)
            out.append(Y[:, :FOUT] * self.out_std[:FOUT] + self.out_mean[:FOUT])
        return torch.cat(out)

    def struct(self):
        """pywrf_surrogate_t of the current weights (with their fragment-order
        copies, which this object keeps alive)."""
        s = _lib.Surrogate()
        net = self.net
        s.w1, s.w2, s.w3 = (t.data_ptr() for t in net.Wb)
        self._frag = []
        for W in net.Wb:
            n, k = W.shape
            f = torch.empty(((n + 31) // 32 * 32, k), dtype=torch.bfloat16, device=W.device)
            call("pywrf_mlp_frag_pack", n, k, W.data_ptr(), W.stride(0), f.data_ptr(), stream())
            self._frag.append(f)
        s.w1f, s.w2f, s.w3f = (f.data_ptr() for f in self._frag)
        s.b1, s.b2, s.b3 = (t.data_ptr() for t in net.b)
        s.in_mean, s.in_rstd = self.in_mean.data_ptr(), self.in_rstd.data_ptr()
        s.out_mean, s.out_std = self.out_mean.data_ptr(), self.out_std.data_ptr()
        s.in_lo, s.in_hi, s.out_absmax = self.in_lo.data_ptr(), self.in_hi.data_ptr(), self.out_absmax.data_ptr()
        s.qtol, s.wtol, s.envelope, s.dtmax = self.QTOL, self.WTOL, self.ENVELOPE, self.DTMAX
        return s

    def save(self, path):
        torch.save({"net": {k: v.cpu() for k, v in self.net.state_dict().items()},
                    "in_mean": self.in_mean.cpu(), "in_rstd": self.in_rstd.cpu(),
                    "out_mean": self.out_mean.cpu(), "out_std": self.out_std.cpu(), "in_lo": self.in_lo.cpu(),
                    "in_hi": self.in_hi.cpu(), "out_absmax": self.out_absmax.cpu()}, path)

    @classmethod
    def load(cls, path, device="cuda"):
        d = torch.load(path, map_location=device, weights_only=True)
        self = cls.__new__(cls)
        self.net = MLP(0, device=device)
        self.net.load_state_dict(d["net"])
        self.in_mean, self.in_rstd = d["in_mean"], d["in_rstd"]
        self.out_mean, self.out_std = d["out_mean"], d["out_std"]
        self.in_lo, self.in_hi, self.out_absmax = d["in_lo"], d["in_hi"], d["out_absmax"]
        self.history = []
        return self


# ----------------------------------------------------------------- inline use
class _SurrogateOps:
    def __init__(self, ops, sur, chunk, qtol=None, wtol=None, envelope=None, dtmax=None, noalloc=False):
        self._ops, self.sur, self.chunk, self.noalloc = ops, sur, chunk, noalloc
        self._ws = {}
        self.nfallback = torch.zeros(1, dtype=torch.int32, device="cuda")
        self._struct = sur.struct()
        if qtol is not None:
            self._struct.qtol = qtol
        if wtol is not None:
            self._struct.wtol = wtol
        if envelope is not None:
            self._struct.envelope = envelope
        if dtmax is not None:
            self._struct.dtmax = dtmax

    def __getattr__(self, name):
        return getattr(self._ops, name)

    def microphysics_driver(self, th, rho, pi_phy, p, ht,
                            dz8w, p8w, dt, dx, dy,
                            qv_curr, qc_curr, qr_curr, qi_curr, qs_curr, qg_curr,
                            rainnc, rainncv, snownc, snowncv, sr,
                            refl_10cm, diagflag, do_radar_ref,
                            graupelnc, graupelncv,
                            has_reqc, has_reqi, has_reqs,
                            re_cloud, re_ice, re_snow,
                            ids, ide, jds, jde, kds, kde,
                            ims, ime, jms, jme, kms, kme,
                            its, ite, jts, jte, kts, kte):
        """microphysics_driver's signature and results (wrf_physics.py:94-134),
        the columns computed by the surrogate with per-column WSM6 fallback."""
        g = grid(th.shape, (ids, ide, jds, jde, kds, kde, ims, ime, jms, jme, kms, kme, its, ite, jts, jte, kts, kte))
        if self.noalloc:
            call("pywrf_wsm6_surrogate_noalloc", g, ctypes.addressof(self._struct), ptr(th), ptr(rho), ptr(pi_phy),
                 ptr(p), ptr(dz8w), float(dt), ptr(qv_curr), ptr(qc_curr), ptr(qr_curr), ptr(qi_curr), ptr(qs_curr),
                 ptr(qg_curr), ptr(rainnc), ptr(rainncv), ptr(snownc), ptr(snowncv), ptr(sr), ptr(graupelnc),
                 ptr(graupelncv), _vp(self.nfallback), stream())
            return (th, qv_curr, qc_curr, qr_curr, qi_curr, qs_curr, qg_curr, rainnc, rainncv, snownc, snowncv,
                    graupelnc, graupelncv, sr, refl_10cm)
        key = tuple(th.shape) + (its, ite, jts, jte)
        ws = self._ws.get(key)
        if ws is None:
            nbytes = ctypes.c_longlong()
            call("pywrf_wsm6_surrogate_workspace", g, self.chunk, ctypes.byref(nbytes))
            ws = self._ws[key] = torch.empty(nbytes.value, dtype=torch.uint8, device="cuda")
        call("pywrf_wsm6_surrogate", g, ctypes.addressof(self._struct), ptr(th), ptr(rho), ptr(pi_phy), ptr(p), ptr(dz8w),
             float(dt), ptr(qv_curr), ptr(qc_curr), ptr(qr_curr), ptr(qi_curr), ptr(qs_curr), ptr(qg_curr),
             ptr(rainnc), ptr(rainncv), ptr(snownc), ptr(snowncv), ptr(sr), ptr(graupelnc), ptr(graupelncv),
             _vp(ws), ws.numel(), _vp(self.nfallback), stream())
        return (th, qv_curr, qc_curr, qr_curr, qi_curr, qs_curr, qg_curr, rainnc, rainncv, snownc, snowncv,
                graupelnc, graupelncv, sr, refl_10cm)


def surrogate_ops(sur, ops=None, chunk=65536, qtol=None, wtol=None, envelope=None, dtmax=None, noalloc=False):
    """The step's functions with microphysics_driver served by the surrogate
    (ops: the namespace to wrap, pywrf_amd.ops() by default; qtol / wtol /
    envelope / dtmax override the fallback tolerances; noalloc: use the
    library's allocation-free entry point, pywrf_wsm6_surrogate_noalloc)."""
    import pywrf_amd
    return _SurrogateOps(ops or pywrf_amd.ops(), sur, chunk, qtol, wtol, envelope, dtmax, noalloc)


# ----------------------------------------------------------------- evaluation
def compare_runs(sur, config="config2", nsteps=60, chunk=65536, log=None):
    """The surrogate step against the plain step on the same seeded state:
    steps/s of each (HIP graphs of 10 steps, the two runs' replays
    interleaved so that clock drift falls on both alike, the first replay of
    each untimed), the RMS difference of qv and T (theta) after nsteps
    relative to the RMS of the plain run's field, and the mean fraction of
    columns that fell back to WSM6."""
    import pywrf_amd
    from .gen import synthetic_state
    from .graph import StepGraph
    from .state import clone_state
    from .step import RunConfig
    cfg = RunConfig()
    base = synthetic_state(config)
    out = {}
    per = 10 if nsteps % 10 == 0 else 1
    reps = nsteps // per
    runs = {}
    for name in ("wsm6", "surrogate"):
        s = clone_state(base, device="cuda")
        ops = pywrf_amd.ops() if name == "wsm6" else surrogate_ops(sur, chunk=chunk)
        runs[name] = (s, ops, StepGraph(s, cfg, ops, steps=per))
    times = {name: [] for name in runs}
    fb = []
    for r in range(reps):
        for name, (s, ops, g) in runs.items():
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            g.replay()
            torch.cuda.synchronize()
            if r > 0 or reps == 1:
                times[name].append(time.perf_counter() - t0)
            if name == "surrogate":
                fb.append(ops.nfallback.clone())
    for name, t in times.items():
        out[name + "_steps_per_s"] = per * len(t) / sum(t)
        if log:
            log("compare: %s %.1f steps/s" % (name, out[name + "_steps_per_s"]))
    gr = runs["surrogate"][0].grid
    ncol = (gr.nx - 13) * (gr.ny - 13)
    out["fallback_fraction"] = float(torch.cat(fb).float().mean()) / ncol
    finals = {name: (s.moist[1].double().cpu(), s.t.double().cpu()) for name, (s, _, _) in runs.items()}
    rms = lambda x: float(torch.sqrt(torch.mean(x * x)))
    (qv0, t0_), (qv1, t1) = finals["wsm6"], finals["surrogate"]
    out["qv_rms_rel"] = rms(qv1 - qv0) / rms(qv0)
    out["theta_rms_rel"] = rms(t1 - t0_) / rms(t0_ + 300.0)
    out["theta_rms_K"] = rms(t1 - t0_)
    out["theta_maxabs_K"] = float((t1 - t0_).abs().max())
    out["qv_maxabs"] = float((qv1 - qv0).abs().max())
    out["finite"] = bool(torch.isfinite(t1).all()) and bool(torch.isfinite(qv1).all())
    out["speedup"] = out["surrogate_steps_per_s"] / out["wsm6_steps_per_s"]
    out["steps"] = nsteps
    out["timed_steps"] = per * len(times["wsm6"])
    return out


def main(argv=None):
    import argparse
    import json
    import sys
    ap = argparse.ArgumentParser(description="train the WSM6 column surrogate on the GPU")
    ap.add_argument("--runs", type=int, default=50)
    ap.add_argument("--steps", type=int, default=140)
    ap.add_argument("--epochs", type=int, default=8)
    ap.add_argument("--batch", type=int, default=8192)
    ap.add_argument("--lr", type=float, default=1e-3)
    ap.add_argument("--out", help="where to save the trained network")
    ap.add_argument("--load", help="a trained network: skip training, run only the comparison")
    ap.add_argument("--compare", type=int, default=60, help="steps of the 240x240x41 comparison (0: none)")
    a = ap.parse_args(argv)
    log = lambda *m: print(*m, file=sys.stderr, flush=True)
    if a.load:
        print(json.dumps({"compare_240": compare_runs(Surrogate.load(a.load), nsteps=a.compare, log=log)}), flush=True)
        return
    if not a.out:
        ap.error("--out is required when training")
    data = collect_pairs(a.runs, a.steps, log=log)
    nval = max(a.batch, len(data) // 50)
    perm = torch.randperm(len(data), device=data.X.device, generator=torch.Generator(device="cuda").manual_seed(1))
    valid = Pairs(data.X[perm[:nval]], data.Y[perm[:nval]])
    train = Pairs(data.X[perm[nval:]], data.Y[perm[nval:]])
    del data
    untrained = Surrogate.untrained(train)
    untrained.loss_weight = loss_weights()
    sur = Surrogate.fit(train, epochs=a.epochs, batch=a.batch, lr=a.lr, log=log, valid=valid)
    sur.save(a.out)
    rec = {"pairs": len(train) + len(valid), "train_pairs": len(train), "valid_pairs": len(valid),
           "epochs": a.epochs, "batch": a.batch, "train_seconds": round(sur.train_seconds, 2),
           "train_steps": sur.train_steps, "valid_loss": sur.history[-1]["valid_loss"],
           "valid_loss_untrained": untrained.loss(valid, a.batch), "history": sur.history}
    if a.compare:
        rec["compare_240"] = compare_runs(sur, nsteps=a.compare, log=log)
    print(json.dumps(rec), flush=True)


if __name__ == "__main__":
    main()
